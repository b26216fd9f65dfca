"""SparseSolve (faspxx_amd.autograd): adjoint gradients for the variable-
coefficient diffusion system (-div(kappa grad u) = b, harmonic-mean faces,
loss 0.5 ||x - x_t||^2).  Checked against an FP64 dense reference gradient at
32 x 32 and against central finite differences (h = 1e-4 and 1e-6) along 16
random directions at 512 x 512, with the gradient's forward and adjoint solves
at rtol 1e-10 and the finite-difference solves at rtol 1e-12.  No reference
counterpart: parity is pinned by the mathematics (dL/db = A^{-T} g,
dL/dA = -lam x^T on the pattern)."""
import numpy as np
import pytest

from faspxx_amd import problems as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tctx():
    import torch
    from faspxx_amd import fx
    torch.cuda.set_device(0)
    s = torch.cuda.Stream()
    torch.cuda.set_stream(s)
    c = fx.Context(0, stream=s.cuda_stream)
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
    """Central differences along 16 random directions (symmetric in the
    values, so A stays SPD; unit norm over (values, b)).  The loss is about
    3.5e7 here, so at h = 1e-6 the difference quotient itself carries
    rounding noise of up to ~3e-6 relative (the same directions with exact
    sparse-LU solves on the CPU give 2e-7 .. 3.2e-6); the 1e-6 bar is
    therefore applied at h = 1e-4 (LU-exact noise ~1e-8, truncation O(h^2)),
    and h = 1e-6 is held to its own noise floor, 1e-5."""
    import scipy.sparse as sp
    import torch
    from faspxx_amd.autograd import SolveOptions, SparsePattern, sparse_solve
    pr = P.config(5)
    pat, L0, gv, gb, _ = loss_and_grads(tctx, pr, 1e-10)
    fpat = SparsePattern(tctx, pr.A, SolveOptions(relTol=1e-12))
    xt = torch.tensor(pr.xstar, device="cuda")
    rng = np.random.default_rng(4)
    rows = np.repeat(np.arange(pr.A.nrow), np.diff(pr.A.row_ptr.astype(np.int64)))
    cols = pr.A.col_ind.astype(np.int64)

    def L(vals, b):
        with torch.no_grad():
            x = sparse_solve(torch.tensor(vals, device="cuda"), torch.tensor(b, device="cuda"),
                             fpat)
            return 0.5 * torch.sum((x - xt) ** 2).item()

    errs = {1e-4: [], 1e-6: []}
    for k in range(16):
        dv = rng.standard_normal(pr.A.nnz)
        D = sp.csr_matrix((dv, cols, pr.A.row_ptr.astype(np.int64)), shape=(pr.A.nrow, pr.A.nrow))
        Ds = (0.5 * (D + D.T)).tocsr()
        Ds.sort_indices()
        # values of the symmetrised direction on A's pattern, in A's order
        dvs = np.asarray(Ds[rows, cols]).ravel()
        db = rng.standard_normal(pr.A.nrow)
        scale = np.sqrt(np.sum(dvs ** 2) + np.sum(db ** 2))
        dvs, db = dvs / scale, db / scale
        an = float(gv @ dvs + gb @ db)
        for h in errs:
            fd = (L(pr.A.values + h * dvs, pr.b + h * db)
                  - L(pr.A.values - h * dvs, pr.b - h * db)) / (2 * h)
            errs[h].append(abs(fd - an) / max(abs(an), 1e-30))
    for h, e in errs.items():
        print(f"FD relative errors, h = {h:g}:", ["%.2e" % v for v in e])
    assert max(errs[1e-4]) <= 1e-6, errs[1e-4]
    assert max(errs[1e-6]) <= 1e-5, errs[1e-6]
