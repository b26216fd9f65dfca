"""The finite-difference check of the adjoint gradients (test_gpu_adjoint)
evaluates L(theta + h d) - L(theta - h d) through a rearranged, cancellation-
free identity.  This CPU test confirms the identity itself on exact sparse-LU
solves."""
import numpy as np

from faspxx_amd import problems as P


def test_difference_identity_exact_solves():
    """The rearranged difference equals L(x+) - L(x-) on exact sparse-LU
    solves of a small system (CPU), so the identity carries no modelling
    error; checked at 64 x 64 where the losses themselves are small."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    pr = P.config(5, 64)
    rowp = pr.A.row_ptr.astype(np.int64)
    cols = pr.A.col_ind.astype(np.int64)
    n = pr.A.nrow
    A0 = sp.csr_matrix((pr.A.values, cols, rowp), shape=(n, n)).tocsc()
    rng = np.random.default_rng(1)
    dv = rng.standard_normal(pr.A.nnz) * 1e-3
    db = rng.standard_normal(n) * 1e-3
    D = sp.csr_matrix((dv, cols, rowp), shape=(n, n)).tocsc()
    h = 1e-2

    def L(x):
        return 0.5 * float(np.sum((x - pr.xstar) ** 2))

    x0 = spl.spsolve(A0, pr.b)
    xp = spl.spsolve(A0 + h * D, pr.b + h * db)
    xm = spl.spsolve(A0 - h * D, pr.b - h * db)
    r0 = pr.b - A0 @ x0
    g = db - D @ x0
    dp = spl.spsolve(A0 + h * D, r0 + h * g)
    dm = spl.spsolve(A0 - h * D, r0 - h * g)
    direct = L(xp) - L(xm)
    ident = 0.5 * float((dp - dm) @ (2 * (x0 - pr.xstar) + dp + dm))
    assert abs(direct - ident) <= 1e-10 * abs(direct), (direct, ident)
