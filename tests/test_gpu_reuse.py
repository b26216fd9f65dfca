"""Solver objects reused across matrices, and contexts on the null stream.

* One solver object is set up again after its matrix and preconditioner
  handles were destroyed and made again (a different size, then the same
  size with other values).  A freshly made handle usually reuses the
  address of the one just freed, so the solver must tell them apart by
  more than their address: it keeps its fused-path workspace and replay
  graphs only for the very same handles.  Each solve is checked against
  the oracle (iteration count +-2, solution).
* A context on the legacy null stream (fx_ctx_create_on_stream(dev, NULL))
  cannot capture graphs; every fused loop, the refinement included, then
  launches its kernels directly.  Checked against the oracle.
"""
import dataclasses

import numpy as np
import pytest

from conftest import rel_err
from faspxx_amd import fx
from oracle import BICGSTAB, CG, FGMRES, GMRES, JACOBI, Matrix

pytestmark = pytest.mark.gpu

PRM = dict(maxIter=8000, relTol=1e-8, absTol=0.0, restart=20)
KINDS = [(fx.CG, CG), (fx.BICGSTAB, BICGSTAB), (fx.GMRES, GMRES), (fx.FGMRES, FGMRES)]
# Per solver, three systems on which the reference's method converges to
# 1e-8 (a different size, then the same size with other values): CG needs
# symmetric ones; the reference's BiCGStab (P A recurrence) diverges on the
# variable-coefficient system; its FGMRES with a Jacobi preconditioner
# (stale Z) converges slowly, so it gets smaller ones.
# (config, grid side, factor on A and b)
SEQ = {fx.CG: [(1, 40, 1.0), (1, 32, 1.0), (5, 32, 1.0)],
       fx.BICGSTAB: [(3, 40, 1.0), (3, 32, 1.0), (1, 32, 1.0)],
       fx.GMRES: [(3, 40, 1.0), (3, 32, 1.0), (1, 32, 1.0)],
       fx.FGMRES: [(3, 28, 1.0), (3, 20, 1.0), (3, 20, 2.5)]}


def system(problems, cid, side, factor):
    pr = problems(cid, side)
    if factor == 1.0:
        return pr
    A = dataclasses.replace(pr.A, values=pr.A.values * factor)
    return dataclasses.replace(pr, A=A, b=pr.b * factor)


def dev(ctx, a):
    return fx.Vec.from_numpy(ctx, np.asarray(a))


def check_solve(ctx, oracle, s, pr, okind, state=(None, 0)):
    """One solve against the oracle.  GMRES and FGMRES keep state in the
    solver object between Solve calls: the restart length the last solve
    ended with (src/GMRES.cxx:255-265; Setup resets it) and, for FGMRES,
    the basis Z that each Jacobi application starts from
    (src/FGMRES.cxx:158-159).  The oracle replays the solve from the state
    the device solver holds, `state` = (Z, restart), and returns the state
    the device holds after it: the oracle's, stopped at the device's
    iteration count (the carried state depends on where the solve stops,
    and the counts may differ by the +-2 of the parity bar)."""
    x = dev(ctx, np.zeros(pr.A.nrow))
    res = s.Solve(dev(ctx, pr.b), x)
    M = Matrix.of(pr.A)
    x0 = np.zeros(M.nrow)
    if okind in (GMRES, FGMRES):
        z, rst = state
        if okind != FGMRES or (z is not None and z.shape[1] != M.nrow):
            z = None
        kw = dict(pc="jacobi", z=z, restart_in=rst, **PRM)
        o = oracle.krylov_traced(okind, M, pr.b, x0, **kw)
    else:
        o = oracle.solve(okind, M, pr.b, x0, pc="jacobi", **PRM)
    assert res.retcode == o["retcode"] == 0, (res.retcode, o["retcode"])
    assert o["numIter"] < PRM["maxIter"]  # the reference's method converged here
    assert abs(res.numIter - o["numIter"]) <= 2, (res.numIter, o["numIter"])
    xr = x.numpy()
    assert np.linalg.norm(pr.b - pr.A.matvec(xr)) / np.linalg.norm(pr.b) < 1e-7
    assert rel_err(xr, o["x"]) < 1e-6
    if okind not in (GMRES, FGMRES):
        return (None, 0)
    if res.numIter != o["numIter"]:
        kw["maxIter"] = res.numIter
        o = oracle.krylov_traced(okind, M, pr.b, x0, **kw)
    # the device converged in its last cycle, so it kept that cycle's length
    return (o["Z"], o["planned"][-1])


@pytest.mark.parametrize("kind,okind", KINDS)
def test_setup_again_after_handles_remade(ctx, oracle, problems, kind, okind):
    seq = [system(problems, *c) for c in SEQ[kind]]
    s = fx.Solver(ctx, kind)
    prm = fx.params(**PRM)
    addrs = []
    state = (None, 0)
    for pr in seq:
        A = fx.Csr.from_problem(ctx, pr.A)
        P = fx.Precond(ctx, "jacobi", A, 1.0, 1)
        addrs.append((A.h.value, P.h.value))
        s.Setup(A, P, prm)
        state = (state[0], 0)  # Setup resets the restart length, not Z
        for _ in range(2):  # the second solve reuses the kept workspace
            if kind == fx.FGMRES:
                # FGMRES's Jacobi applications start from the Z the last
                # solve left (src/FGMRES.cxx:158-159); a solve from such a Z
                # amplifies rounding beyond the +-2 bar (its count moves by
                # tens of iterations with the summation order), so each
                # solve here starts from Clean()'s state: Z = 0
                s.Clean()
                state = (None, 0)
            state = check_solve(ctx, oracle, s, pr, okind, state)
        P.close()
        A.close()
    print("handle addresses (matrix, preconditioner):", addrs)
    s.close()


def test_jacobi_solver_after_handles_remade(ctx, oracle, problems):
    s = fx.Solver(ctx, fx.JACOBI)
    prm = fx.params(maxIter=40, minIter=0, relTol=1e-30, absTol=0.0)
    for pr in (problems(1, 40), problems(1, 32), problems(5, 32)):
        A = fx.Csr.from_problem(ctx, pr.A)
        P = fx.Precond(ctx, "jacobi", A, 2.0 / 3.0, 1)
        s.Setup(A, P, prm)
        x = dev(ctx, np.zeros(pr.A.nrow))
        res = s.Solve(dev(ctx, pr.b), x)
        M = Matrix.of(pr.A)
        o = oracle.solve(JACOBI, M, pr.b, np.zeros(M.nrow), weight=2.0 / 3.0, maxIter=40,
                         minIter=0, relTol=1e-30, absTol=0.0)
        assert res.numIter == o["numIter"] == 40
        assert np.array_equal(x.numpy(), o["x"])  # sweeps are bit-exact
        P.close()
        A.close()
    s.close()


@pytest.fixture(scope="module")
def nctx():
    c = fx.Context(0, stream=0)
    assert c.stream == 0
    yield c
    c.close()


@pytest.mark.parametrize("kind,okind", KINDS)
def test_null_stream_context_solvers(nctx, oracle, problems, kind, okind):
    pr = system(problems, *SEQ[kind][1])
    A = fx.Csr.from_problem(nctx, pr.A)
    P = fx.Precond(nctx, "jacobi", A, 1.0, 1)
    s = fx.Solver(nctx, kind)
    s.Setup(A, P, fx.params(**PRM))
    check_solve(nctx, oracle, s, pr, okind)
    s.close()


def test_null_stream_context_jacobi_sweeps(nctx, oracle, problems):
    pr = problems(2, 24)
    A = fx.Csr.from_problem(nctx, pr.A)
    P = fx.Precond(nctx, "jacobi", A, 2.0 / 3.0, 1)
    s = fx.Solver(nctx, fx.JACOBI)
    s.Setup(A, P, fx.params(maxIter=50, minIter=50, relTol=1e-30, absTol=0.0))
    x = dev(nctx, np.zeros(pr.A.nrow))
    s.Solve(dev(nctx, pr.b), x)
    M = Matrix.of(pr.A)
    o = oracle.solve(JACOBI, M, pr.b, np.zeros(M.nrow), weight=2.0 / 3.0, maxIter=50,
                     minIter=50, relTol=1e-30, absTol=0.0)
    assert np.array_equal(x.numpy(), o["x"])


@pytest.mark.parametrize("cid,scale", [(2, 32), (4, 16384)])
def test_null_stream_context_refinement(nctx, oracle, problems, cid, scale):
    """The refinement (scaled f32 inner CG, graph-replayed on a real stream)
    on a null-stream context: final FP64 relres <= 1e-10, solution against
    the oracle's FP64 Jacobi-PCG at rtol 1e-12."""
    pr = problems(cid, scale)
    A = fx.Csr.from_problem(nctx, pr.A)
    x = fx.Vec(nctx, pr.A.nrow)
    res, inner = fx.solve_ir(nctx, A, dev(nctx, pr.b), x, 1.0,
                             fx.params(maxIter=50, relTol=1e-10, absTol=0.0),
                             fx.params(maxIter=5000, relTol=1e-4, absTol=0.0))
    assert res.retcode == 0 and inner > 0
    xr = x.numpy()
    assert np.linalg.norm(pr.b - pr.A.matvec(xr)) / np.linalg.norm(pr.b) < 1e-10
    M = Matrix.of(pr.A)
    o = oracle.solve(CG, M, pr.b, np.zeros(M.nrow), pc="jacobi", maxIter=20000, relTol=1e-12,
                     absTol=0.0)
    assert rel_err(xr, o["x"]) < 1e-6


def test_null_stream_context_repeated(nctx, problems):
    """Second and third solves on a null-stream context, whose replays are
    launched directly: the refinement with its outer steps queued ahead of
    and behind its inner solves, CG and the Jacobi sweeps with their last
    replay cut to the predicted count.  Each repeat agrees bit for bit with
    the first solve."""
    pr = problems(2, 24)
    n = pr.A.nrow
    A = fx.Csr.from_problem(nctx, pr.A)
    ir = fx.IR(nctx, A, 1.0, fx.params(maxIter=50, relTol=1e-10, absTol=0.0),
               fx.params(maxIter=5000, relTol=1e-4, absTol=0.0))
    runs = []
    for _ in range(3):
        x = fx.Vec(nctx, n)
        res, inner = ir.Solve(dev(nctx, pr.b), x)
        assert res.retcode == 0
        runs.append((res.numIter, inner, x.numpy()))
    ir.close()
    for k in (1, 2):
        assert runs[k][:2] == runs[0][:2] and np.array_equal(runs[k][2], runs[0][2])
    for kind, pc_w, prm in ((fx.CG, 1.0, dict(maxIter=2000, relTol=1e-8, absTol=0.0)),
                            (fx.JACOBI, 2.0 / 3.0, dict(maxIter=40, relTol=1e-30, absTol=0.0))):
        P = fx.Precond(nctx, "jacobi", A, pc_w, 1)
        s = fx.Solver(nctx, kind)
        s.Setup(A, P, fx.params(**prm))
        outs = []
        for _ in range(3):
            x = dev(nctx, np.zeros(n))
            res = s.Solve(dev(nctx, pr.b), x)
            assert res.retcode == 0
            outs.append((res.numIter, x.numpy()))
        for k in (1, 2):
            assert outs[k][0] == outs[0][0] and np.array_equal(outs[k][1], outs[0][1])
        s.close()
        P.close()
