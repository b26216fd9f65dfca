"""GPU multicolor Gauss-Seidel (fx_mcgs_*, no reference counterpart) against
the oracle's definition (oracle/fx_oracle.cpp orc_mc_color / orc_mcgs,
itself pinned by tests/test_oracle_mcgs.py).  Bars: colors identical;
sweeps bit-exact where rows are summed by one thread (all stencil systems),
1e-12 relative otherwise; the symmetric sweep as a preconditioner gives CG
iteration counts within +-2 of the oracle's."""
import numpy as np
import pytest

from conftest import rel_err
from csrfile import read_csr
from faspxx_amd import fx
from oracle import CG, Matrix

pytestmark = pytest.mark.gpu


def dev(ctx, a):
    return fx.Vec.from_numpy(ctx, np.asarray(a, dtype=np.float64))


CASES = [(1, None), (2, 32), (3, 256), (5, 64), (4, 3000)]


@pytest.mark.parametrize("cid,scale", CASES)
def test_colors_and_sweeps(ctx, oracle, problems, cid, scale):
    pr = problems(cid, scale)
    M = Matrix.of(pr.A)
    A = fx.Csr.from_problem(ctx, pr.A)
    onc, ocol = oracle.mc_color(M)
    for w, sweeps, sym in [(1.0, 3, False), (1.2, 2, True)]:
        s = fx.MCGS(ctx, A, w)
        nc, col = s.colors()
        assert nc == onc and np.array_equal(col, ocol)
        x0 = np.random.default_rng(cid).uniform(-1, 1, pr.A.nrow)
        x = dev(ctx, x0)
        s.sweep(dev(ctx, pr.b), x, sweeps, sym)
        ref = oracle.mcgs(M, ocol, onc, w, pr.b, x0, sweeps, sym)
        if cid == 4:  # long rows are summed by lane groups
            assert rel_err(x.numpy(), ref) < 1e-12
        else:
            assert np.array_equal(x.numpy(), ref)
        s.close()


def test_fixture_matrices(ctx, oracle):
    for name in ("fdm_10X10.csr", "fem_small.csr"):
        n, rp, ci, va = read_csr(name)
        dp = oracle.form_diagptr(n, rp, ci)
        M = Matrix(n, n, rp, ci, va, dp)
        A = fx.Csr(ctx, n, n, rp, ci, va, None)
        s = fx.MCGS(ctx, A, 1.0)
        onc, ocol = oracle.mc_color(M)
        nc, col = s.colors()
        assert nc == onc and np.array_equal(col, ocol)
        b = np.random.default_rng(n).standard_normal(n)
        x = dev(ctx, np.zeros(n))
        s.sweep(dev(ctx, b), x, 4, True)
        assert np.array_equal(x.numpy(), oracle.mcgs(M, ocol, onc, 1.0, b, np.zeros(n), 4, True))


@pytest.mark.parametrize("cid,scale", [(5, 64), (1, 64)])
def test_mcgs_preconditioned_cg(ctx, oracle, problems, cid, scale):
    pr = problems(cid, scale)
    M = Matrix.of(pr.A)
    prm = dict(maxIter=3000, relTol=1e-10, absTol=0.0)
    o = oracle.solve(CG, M, pr.b, np.zeros(M.nrow), pc="mcgs", pc_iters=1, **prm)
    A = fx.Csr.from_problem(ctx, pr.A)
    pc = fx.Precond(ctx, "mcgs", A, 1.0, 1)
    s = fx.Solver(ctx, fx.CG)
    s.Setup(A, pc, fx.params(**prm))
    x = fx.Vec(ctx, M.nrow)
    res = s.Solve(dev(ctx, pr.b), x)
    assert res.retcode == o["retcode"] == 0
    assert abs(res.numIter - o["numIter"]) <= 2, (res.numIter, o["numIter"])
    assert rel_err(x.numpy(), o["x"]) < 1e-8
