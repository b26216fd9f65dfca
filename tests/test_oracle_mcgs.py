"""The oracle's multicolor Gauss-Seidel (no reference counterpart) pinned by
the mathematics, CPU only:
  * the greedy coloring is a proper coloring of A + A^T, and on the 5- and
    7-point stencils it is the red-black parity (i + j (+ k)) mod 2;
  * with every row its own color the sweep is textbook SOR,
    x+ = (D/w + L)^{-1} (b - (U + (1 - 1/w) D) x), solved here with SciPy's
    sparse triangular solver (the symmetric sweep adds the backward half);
  * on 2D Poisson the red-black sweep equals a vectorized NumPy red-black
    update written from the stencil."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from csrfile import read_csr
from faspxx_amd import problems as P
from oracle import Matrix


def proper(A, color):
    S = (A + A.T).tocoo()
    off = S.row != S.col
    return not np.any(color[S.row[off]] == color[S.col[off]])


@pytest.mark.parametrize("cid,scale", [(1, 16), (2, 8), (3, 32), (4, 3000), (5, 16)])
def test_coloring_is_proper(oracle, problems, cid, scale):
    pr = problems(cid, scale)
    nc, color = oracle.mc_color(Matrix.of(pr.A))
    assert proper(pr.A.to_scipy(), color)
    assert nc == color.max() + 1
    if cid in (1, 3, 5):  # 5-point: red-black
        m = scale
        i = np.arange(pr.A.nrow)
        assert nc == 2 and np.array_equal(color, ((i % m) + (i // m)) % 2)
    if cid == 2:
        m = scale
        i = np.arange(pr.A.nrow)
        assert nc == 2 and np.array_equal(color, ((i % m) + (i // m) % m + i // (m * m)) % 2)


def test_coloring_fixture_matrices(oracle):
    for name in ("fdm_10X10.csr", "fem_small.csr"):
        n, rp, ci, va = read_csr(name)
        M = Matrix(n, n, rp, ci, va, oracle.form_diagptr(n, rp, ci))
        nc, color = oracle.mc_color(M)
        A = sp.csr_matrix((va, ci.astype(np.int64), rp.astype(np.int64)), shape=(n, n))
        assert proper(A, color) and nc >= 2


@pytest.mark.parametrize("w", [1.0, 1.3, 0.7])
def test_one_color_per_row_is_sor(oracle, problems, w):
    pr = problems(3, 16)  # nonsymmetric upwind convection-diffusion, n = 256
    M = Matrix.of(pr.A)
    A = pr.A.to_scipy().tocsr()
    n = A.shape[0]
    x0 = np.random.default_rng(1).uniform(-1, 1, n)
    ident = np.arange(n, dtype=np.uint32)
    D = sp.diags(A.diagonal())
    L, U = sp.tril(A, -1), sp.triu(A, 1)
    fwd = spl.spsolve_triangular((D / w + L).tocsr(), pr.b - (U + (1 - 1 / w) * D) @ x0, lower=True)
    x1 = oracle.mcgs(M, ident, n, w, pr.b, x0, sweeps=1)
    assert np.linalg.norm(x1 - fwd) <= 1e-12 * np.linalg.norm(fwd)
    bwd = spl.spsolve_triangular((D / w + U).tocsr(), pr.b - (L + (1 - 1 / w) * D) @ fwd,
                                 lower=False)
    x2 = oracle.mcgs(M, ident, n, w, pr.b, x0, sweeps=1, symmetric=True)
    assert np.linalg.norm(x2 - bwd) <= 1e-12 * np.linalg.norm(bwd)


def test_red_black_poisson_closed_form(oracle, problems):
    m = 24
    pr = problems(1, m)
    M = Matrix.of(pr.A)
    nc, color = oracle.mc_color(M)
    x0 = np.random.default_rng(2).uniform(-1, 1, m * m)
    b = pr.b
    x = oracle.mcgs(M, color, nc, 1.0, b, x0, sweeps=2)
    # vectorized red-black GS on the 5-point stencil (diagonal 4, neighbours -1)
    X = x0.reshape(m, m).copy()
    B = b.reshape(m, m)
    jj, ii = np.meshgrid(np.arange(m), np.arange(m))
    red = ((ii + jj) % 2) == 0
    for _ in range(2):
        for mask in (red, ~red):
            P_ = np.pad(X, 1)
            nb = P_[:-2, 1:-1] + P_[2:, 1:-1] + P_[1:-1, :-2] + P_[1:-1, 2:]
            X = np.where(mask, (B + nb) / 4.0, X)
    assert np.linalg.norm(x - X.ravel()) <= 1e-13 * np.linalg.norm(x)


def test_mcgs_smooths_and_converges(oracle, problems):
    """As a stationary iteration on an SPD M-matrix the symmetric sweep
    converges, and faster than weighted Jacobi."""
    pr = problems(5, 16)
    M = Matrix.of(pr.A)
    nc, color = oracle.mc_color(M)
    x = np.zeros(pr.A.nrow)
    x = oracle.mcgs(M, color, nc, 1.0, pr.b, x, sweeps=50, symmetric=True)
    r_gs = np.linalg.norm(pr.b - pr.A.matvec(x))
    r_j = np.linalg.norm(pr.b - pr.A.matvec(
        oracle.solve(11, M, pr.b, np.zeros(pr.A.nrow), weight=2.0 / 3.0, maxIter=100,
                     relTol=1e-30, absTol=0.0)["x"]))
    assert r_gs < r_j < np.linalg.norm(pr.b)


def test_mcgs_preconditioned_cg(oracle, problems):
    """Symmetric multicolor GS as the CG preconditioner (the oracle's pc kind
    2): a symmetric SPD preconditioner, so PCG converges, in fewer iterations
    than Jacobi-PCG on the variable-coefficient system."""
    pr = problems(5, 32)
    M = Matrix.of(pr.A)
    prm = dict(maxIter=2000, relTol=1e-10, absTol=0.0)
    gs = oracle.solve(1, M, pr.b, np.zeros(pr.A.nrow), pc="mcgs", pc_iters=1, **prm)
    jac = oracle.solve(1, M, pr.b, np.zeros(pr.A.nrow), pc="jacobi", **prm)
    assert gs["retcode"] == 0 and jac["retcode"] == 0
    assert gs["numIter"] < 0.75 * jac["numIter"], (gs["numIter"], jac["numIter"])
    x = gs["x"]
    assert np.linalg.norm(pr.b - pr.A.matvec(x)) < 1e-9 * np.linalg.norm(pr.b)
