"""fx_pc_apply_noalloc (the allocation-free form of fx_pc_apply, for callers
that cannot allocate) against fx_pc_apply through the C ABI: bit for bit, on
two calls in a row, the second with a smaller system; multi-sweep Jacobi
from a nonzero and from a zero start (the residual's shortcut), identity,
and the size limit's error."""
import numpy as np
import pytest

from faspxx_amd import fx

pytestmark = pytest.mark.gpu


def both(ctx, A, pc_kind, weight, iters, b, x0):
    P = fx.Precond(ctx, pc_kind, A, weight, iters)
    bv = fx.Vec.from_numpy(ctx, b)
    xa = fx.Vec.from_numpy(ctx, x0)
    xb = fx.Vec.from_numpy(ctx, x0)
    P.Solve(bv, xa)
    P.SolveNoAlloc(bv, xb)
    ctx.sync()
    out = xa.numpy(), xb.numpy()
    P.close()
    return out


@pytest.mark.parametrize("kind,weight,iters", [("jacobi", 1.0, 1), ("jacobi", 2.0 / 3.0, 3),
                                               ("identity", 1.0, 1)])
def test_noalloc_equals_pc_apply(ctx, problems, kind, weight, iters):
    rng = np.random.default_rng(7)
    # two calls in a row, the second on a smaller system: 3D Poisson 64^3
    # (n = 262,144), then the variable-coefficient 96^2 system (n = 9,216)
    for cid, side in ((2, 64), (5, 96)):
        pr = problems(cid, side)
        A = fx.Csr.from_problem(ctx, pr.A)
        n = pr.A.nrow
        for x0 in (rng.uniform(-1, 1, n), np.zeros(n)):
            a, b = both(ctx, A, kind, weight, iters, pr.b, x0)
            assert np.array_equal(a, b), (cid, side, kind, iters)
            assert np.any(a != x0)
        A.close()


def test_noalloc_size_limit(ctx):
    n = (1 << 22) + 1
    rp = np.arange(n + 1, dtype=np.uint32)
    ci = np.arange(n, dtype=np.uint32)
    A = fx.Csr(ctx, n, n, rp, ci, np.full(n, 2.0), None)
    P = fx.Precond(ctx, "jacobi", A, 1.0, 1)
    b = fx.Vec.from_numpy(ctx, np.ones(n))
    x = fx.Vec(ctx, n)
    with pytest.raises(Exception) as e:
        P.SolveNoAlloc(b, x)
    assert "exceed" in str(e.value)
    # the ordinary entry point still takes it
    P.Solve(b, x)
    assert np.all(x.numpy() == 0.5)
    P.close()
    A.close()
