"""SparseSolve (faspxx_amd.autograd): adjoint gradients for the variable-
coefficient diffusion system (-div(kappa grad u) = b, harmonic-mean faces,
loss 0.5 ||x - x_t||^2).  Checked against an FP64 dense reference gradient at
32 x 32 and against central finite differences (h = 1e-4 and 1e-6) along 16
random directions at 512 x 512, with the gradient's forward and adjoint solves
at rtol 1e-10.  No reference counterpart: parity is pinned by the mathematics
(dL/db = A^{-T} g, dL/dA = -lam x^T on the pattern).

The library context has its own stream here (fx.Context(0)), while torch
works on its default stream: SparseSolve orders the two itself."""
import numpy as np
import pytest

from faspxx_amd import problems as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tctx():
    import torch
    from faspxx_amd import fx
    torch.cuda.set_device(0)
    c = fx.Context(0)
    assert c.stream != torch.cuda.current_stream().cuda_stream
    yield c
    c.close()


def loss_and_grads(tctx, pr, rtol):
    import torch
    from faspxx_amd.autograd import SolveOptions, SparsePattern, sparse_solve
    pat = SparsePattern(tctx, pr.A, SolveOptions(relTol=rtol))
    vals = torch.tensor(pr.A.values, device="cuda", requires_grad=True)
    b = torch.tensor(pr.b, device="cuda", requires_grad=True)
    xt = torch.tensor(pr.xstar, device="cuda")
    x = sparse_solve(vals, b, pat)
    L = 0.5 * torch.sum((x - xt) ** 2)
    L.backward()
    return pat, L.item(), vals.grad.cpu().numpy(), b.grad.cpu().numpy(), x.detach().cpu().numpy()


def test_dense_reference_gradient_32(tctx):
    pr = P.config(5, 32)
    pat, L, gv, gb, x = loss_and_grads(tctx, pr, 1e-13)
    A = pr.A.to_scipy().toarray()
    xd = np.linalg.solve(A, pr.b)
    lam = np.linalg.solve(A.T, xd - pr.xstar)
    rows = np.repeat(np.arange(pr.A.nrow), np.diff(pr.A.row_ptr.astype(np.int64)))
    gv_ref = -lam[rows] * xd[pr.A.col_ind]
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-10
    assert np.linalg.norm(gb - lam) / np.linalg.norm(lam) < 1e-9
    assert np.linalg.norm(gv - gv_ref) / np.linalg.norm(gv_ref) < 1e-9


def test_finite_differences_512(tctx):
    """Central differences along 16 random directions d = (dv, db) (dv
    symmetric in the values, so A stays SPD; unit norm over (values, b)).

    The loss is about 3.5e7 here, so L(theta + h d) - L(theta - h d) is not
    formed by subtracting two losses (at h = 1e-6 that difference is 1e-6 of
    each, and the rounding of L alone would be ~1e-3 of it).  The same
    difference is evaluated exactly instead: with x0 a base solution,
    r0 = b - A x0 and g = db - D x0 (D = dv on A's pattern),
        x(theta +- h d) = x0 + delta_pm,   (A +- h D) delta_pm = r0 +- h g,
        L+ - L- = 0.5 (delta_+ - delta_-)^T (2 (x0 - x_t) + delta_+ + delta_-),
    an identity in exact arithmetic.  Each delta is solved to rtol 1e-10 of
    its own right-hand side, so its error is O(h) smaller than a full solve's.
    The 1e-6 bar holds at both step sizes."""
    import scipy.sparse as sp
    import torch
    from faspxx_amd.autograd import SolveOptions, SparsePattern, sparse_solve
    pr = P.config(5)
    pat, L0, gv, gb, _ = loss_and_grads(tctx, pr, 1e-10)
    fpat = SparsePattern(tctx, pr.A, SolveOptions(relTol=1e-10))
    rng = np.random.default_rng(4)
    rowp = pr.A.row_ptr.astype(np.int64)
    rows = np.repeat(np.arange(pr.A.nrow), np.diff(rowp))
    cols = pr.A.col_ind.astype(np.int64)
    A0 = sp.csr_matrix((pr.A.values, cols, rowp), shape=(pr.A.nrow, pr.A.nrow))

    def solve(vals, rhs):
        with torch.no_grad():
            return sparse_solve(torch.tensor(vals, device="cuda"),
                                torch.tensor(rhs, device="cuda"), fpat).cpu().numpy()

    x0 = solve(pr.A.values, pr.b)
    r0 = pr.b - A0 @ x0
    e0 = 2.0 * (x0 - pr.xstar)
    errs = {1e-4: [], 1e-6: []}
    for k in range(16):
        dv = rng.standard_normal(pr.A.nnz)
        D = sp.csr_matrix((dv, cols, rowp), shape=(pr.A.nrow, pr.A.nrow))
        Ds = (0.5 * (D + D.T)).tocsr()
        Ds.sort_indices()
        # values of the symmetrised direction on A's pattern, in A's order
        dvs = np.asarray(Ds[rows, cols]).ravel()
        db = rng.standard_normal(pr.A.nrow)
        scale = np.sqrt(np.sum(dvs ** 2) + np.sum(db ** 2))
        dvs, db = dvs / scale, db / scale
        an = float(gv @ dvs + gb @ db)
        Dd = sp.csr_matrix((dvs, cols, rowp), shape=(pr.A.nrow, pr.A.nrow))
        g = db - Dd @ x0
        for h in errs:
            dp = solve(pr.A.values + h * dvs, r0 + h * g)
            dm = solve(pr.A.values - h * dvs, r0 - h * g)
            fd = 0.5 * float((dp - dm) @ (e0 + dp + dm)) / (2 * h)
            errs[h].append(abs(fd - an) / max(abs(an), 1e-30))
    for h, e in errs.items():
        print(f"FD relative errors, h = {h:g}:", ["%.2e" % v for v in e])
    assert max(errs[1e-4]) <= 1e-6, errs[1e-4]
    assert max(errs[1e-6]) <= 1e-6, errs[1e-6]

