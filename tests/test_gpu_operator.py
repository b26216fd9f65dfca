"""Callback operators (fx_op_*, fx_solver_setup_op): the LOP-with-own-Apply
pattern of example/TestMatFree.cxx on the device solvers.  An operator whose
apply is the matrix's own SpMV must give the per-operation matrix solve bit
for bit (same operations in the same order); a host-side 5-point stencil
(TestMatFree's MatFree::Apply) must match the oracle's CG on the assembled
matrix within +-2 iterations."""
import os

import numpy as np
import pytest

from faspxx_amd import fx
from oracle import BICGSTAB, CG, FGMRES, GMRES, Matrix

pytestmark = pytest.mark.gpu


def per_operation(fn):
    old = os.environ.get("FASPXX_FX_NO_FAST")
    os.environ["FASPXX_FX_NO_FAST"] = "1"
    try:
        return fn()
    finally:
        if old is None:
            del os.environ["FASPXX_FX_NO_FAST"]
        else:
            os.environ["FASPXX_FX_NO_FAST"] = old


@pytest.mark.parametrize("kind", [fx.CG, fx.BICGSTAB, fx.GMRES, fx.FGMRES])
@pytest.mark.parametrize("pc", ["identity", "jacobi"])
def test_operator_equals_matrix_solve(ctx, problems, kind, pc):
    pr = problems(3, 64) if kind != fx.CG else problems(5, 64)
    n = pr.A.nrow
    A = fx.Csr.from_problem(ctx, pr.A)
    P = fx.Precond(ctx, pc, A, 1.0, 1)
    prm = fx.params(maxIter=3000, relTol=1e-8, absTol=0.0)
    b = fx.Vec.from_numpy(ctx, pr.b)

    def solve(op):
        s = fx.Solver(ctx, kind)
        s.Setup(op, P, prm)
        x = fx.Vec(ctx, n)
        r = s.Solve(b, x)
        return r, x.numpy()

    r0, x0 = per_operation(lambda: solve(A))
    calls = []

    def apply(x, y):
        calls.append(1)
        A.Apply(x, y)

    r1, x1 = solve(fx.Operator(ctx, n, apply))
    assert r1.retcode == r0.retcode  # FGMRES with Jacobi may stop at maxIter (stale Z)
    assert r1.numIter == r0.numIter
    assert np.array_equal(x1, x0)
    assert len(calls) >= r1.numIter


OKINDS = {fx.CG: CG, fx.BICGSTAB: BICGSTAB, fx.GMRES: GMRES, fx.FGMRES: FGMRES}


@pytest.mark.parametrize("kind", [fx.CG, fx.BICGSTAB, fx.GMRES, fx.FGMRES])
@pytest.mark.parametrize("pc", ["identity", "jacobi"])
def test_device_only_operator_fused(ctx, oracle, problems, kind, pc):
    """An operator whose apply only enqueues device work (here the matrix's
    own SpMV, fx_op_set_device_only) runs the fused solvers: its work is
    captured in the replay graphs, so that once a solver's graphs are made
    (two solves on it) a CG or BiCGStab solve calls apply only to set up
    and to finish.  Against the per-operation operator solve and the
    oracle's solve on the matrix: count +-2, x to 1e-6 relative (FGMRES
    with Jacobi: the 200-iteration budget, its residual norm to 1e-6)."""
    pr = problems(3, 64) if kind != fx.CG else problems(5, 64)
    n = pr.A.nrow
    A = fx.Csr.from_problem(ctx, pr.A)
    P = fx.Precond(ctx, pc, A, 1.0, 1)
    flex_budget = kind == fx.FGMRES and pc == "jacobi"
    prm = dict(maxIter=200, relTol=1e-12, absTol=0.0) if flex_budget else \
        dict(maxIter=3000, relTol=1e-8, absTol=0.0)
    b = fx.Vec.from_numpy(ctx, pr.b)
    calls = []

    def apply(x, y):
        calls.append(1)
        A.Apply(x, y)

    def solve(repeat=1):
        s = fx.Solver(ctx, kind)
        s.Setup(fx.Operator(ctx, n, apply, device_only=True), P, fx.params(**prm))
        for _ in range(repeat):
            calls.clear()
            x = fx.Vec(ctx, n)
            r = s.Solve(b, x)
        return r, x.numpy(), s.last_path

    r1, x1, path = solve(3 if kind in (fx.CG, fx.BICGSTAB) else 1)
    assert path == "fused"
    if kind in (fx.CG, fx.BICGSTAB):
        assert len(calls) <= 4 < r1.numIter, (len(calls), r1.numIter)
    r0, x0, path0 = per_operation(solve)
    assert path0 == "ops"
    o = oracle.solve(OKINDS[kind], Matrix.of(pr.A), pr.b, np.zeros(n), pc=pc, **prm)
    for r, x in ((r1, x1), (r0, x0)):
        assert r.retcode == o["retcode"]
        assert abs(r.numIter - o["numIter"]) <= 2, (r.numIter, o["numIter"])
        if flex_budget:
            assert abs(r.norm2 - o["norm2"]) <= 1e-6 * o["norm2"]
        else:
            assert np.linalg.norm(x - o["x"]) <= 1e-6 * np.linalg.norm(o["x"])


def test_device_only_operator_error(ctx, problems):
    pr = problems(1, 16)
    n = pr.A.nrow

    def boom(x, y):
        raise RuntimeError("apply failed")

    op = fx.Operator(ctx, n, boom, device_only=True)
    s = fx.Solver(ctx, fx.CG)
    s.Setup(op, fx.Precond(ctx, "identity", None, 1.0, 1), fx.params(maxIter=10, relTol=1e-8))
    with pytest.raises(fx.FxError, match="ERROR_INPUT_PAR"):
        s.Solve(fx.Vec.from_numpy(ctx, pr.b), fx.Vec(ctx, n))
    assert isinstance(op.error, RuntimeError)


def stencil_apply(m):
    """MatFree::Apply of example/TestMatFree.cxx on an m x m grid, on host
    copies: y = 4 x_c - x_l - x_r - x_d - x_u (missing neighbours dropped)."""
    def apply(x, y):
        X = x.numpy().reshape(m, m)
        Y = 4.0 * X
        Y[1:, :] -= X[:-1, :]
        Y[:-1, :] -= X[1:, :]
        Y[:, 1:] -= X[:, :-1]
        Y[:, :-1] -= X[:, 1:]
        y.upload(Y.ravel())
    return apply


def test_matfree_stencil_against_oracle(ctx, oracle, problems):
    m = 63  # TestMatFree's dim = 64 mesh
    pr = problems(1, m)  # the same 5-point matrix, assembled
    n = m * m
    b = np.random.default_rng(3).uniform(-1, 1, n)
    x0 = np.full(n, 0.25)  # TestMatFree's initial guess
    prm = dict(maxIter=10000, relTol=1e-8, absTol=1e-12)
    o = oracle.solve(CG, Matrix.of(pr.A), b, x0, **prm)
    op = fx.Operator(ctx, n, stencil_apply(m))
    s = fx.Solver(ctx, fx.CG)
    s.Setup(op, fx.Precond(ctx, "identity", None, 1.0, 1), fx.params(**prm))
    x = fx.Vec.from_numpy(ctx, x0)
    r = s.Solve(fx.Vec.from_numpy(ctx, b), x)
    assert r.retcode == o["retcode"] == 0
    assert abs(r.numIter - o["numIter"]) <= 2, (r.numIter, o["numIter"])
    assert np.linalg.norm(x.numpy() - o["x"]) <= 1e-8 * np.linalg.norm(o["x"])


def test_operator_errors(ctx, problems):
    pr = problems(1, 16)
    n = pr.A.nrow

    def boom(x, y):
        raise RuntimeError("apply failed")

    op = fx.Operator(ctx, n, boom)
    s = fx.Solver(ctx, fx.CG)
    s.Setup(op, fx.Precond(ctx, "identity", None, 1.0, 1), fx.params(maxIter=10, relTol=1e-8))
    with pytest.raises(fx.FxError, match="ERROR_INPUT_PAR"):
        s.Solve(fx.Vec.from_numpy(ctx, pr.b), fx.Vec(ctx, n))
    assert isinstance(op.error, RuntimeError)
    sj = fx.Solver(ctx, fx.JACOBI)
    with pytest.raises(fx.FxError):
        sj.Setup(op, fx.Precond(ctx, "identity", None, 1.0, 1), fx.params())
