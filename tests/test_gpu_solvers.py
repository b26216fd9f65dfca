"""GPU solver parity against the oracle (CPU restatement of src/CG.cxx,
src/BiCGStab.cxx, src/GMRES.cxx, src/FGMRES.cxx, src/Iter.cxx), through the
C ABI.  Bars: iteration counts within +-2 at the same relTol; solutions and
final residual norms within the stated relative tolerances; weighted-Jacobi
sweeps bit-exact (their updates contain no reduction)."""
import dataclasses
import json
import os
import pathlib

import numpy as np
import pytest

from conftest import rel_err
from faspxx_amd import fx
from oracle import BICGSTAB, CG, FGMRES, GMRES, JACOBI, Matrix
from perturb import ulp_perturbed

pytestmark = pytest.mark.gpu


def dev(ctx, a, dt=fx.F64):
    return fx.Vec.from_numpy(ctx, np.asarray(a), dt)


def gpu_solve(ctx, pr, kind, pc="identity", weight=1.0, pc_iters=1, x0=None, force_ops=False,
              **prm):
    A = fx.Csr.from_problem(ctx, pr.A)
    P = fx.Precond(ctx, "jacobi" if (pc == "jacobi" or kind == JACOBI) else "identity", A,
                   weight, pc_iters)
    s = fx.Solver(ctx, kind)
    s.Setup(A, P, fx.params(**prm))
    x = dev(ctx, np.zeros(pr.A.nrow) if x0 is None else x0)
    old = os.environ.get("FASPXX_FX_NO_FAST")
    if force_ops:
        os.environ["FASPXX_FX_NO_FAST"] = "1"
    try:
        res = s.Solve(dev(ctx, pr.b), x)
    finally:
        if force_ops:
            if old is None:
                del os.environ["FASPXX_FX_NO_FAST"]
            else:
                os.environ["FASPXX_FX_NO_FAST"] = old
    return res, x.numpy(), s.history()


PRM = dict(maxIter=3000, relTol=1e-8, absTol=0.0)

CG_CASES = [(1, None, "identity", 1.0), (1, None, "jacobi", 0.8), (2, 32, "jacobi", 1.0),
            (5, 64, "jacobi", 1.0), (4, 16384, "jacobi", 1.0)]


@pytest.mark.parametrize("force_ops", [False, True])
@pytest.mark.parametrize("cid,scale,pc,w", CG_CASES)
def test_cg_parity(ctx, oracle, problems, cid, scale, pc, w, force_ops):
    pr = problems(cid, scale)
    M = Matrix.of(pr.A)
    o = oracle.solve(CG, M, pr.b, np.zeros(M.nrow), pc=pc, weight=w, hist_cap=4000, **PRM)
    res, x, hist = gpu_solve(ctx, pr, fx.CG, pc, w, force_ops=force_ops, **PRM)
    assert res.retcode == o["retcode"] == 0
    assert abs(res.numIter - o["numIter"]) <= 2, (res.numIter, o["numIter"])
    assert rel_err(x, o["x"]) < 1e-7
    true_res = np.linalg.norm(pr.b - pr.A.matvec(x)) / np.linalg.norm(pr.b)
    assert true_res < 1e-8
    assert abs(res.norm2 - o["norm2"]) <= 1e-3 * o["norm2"] + 1e-14
    k = min(len(hist), 30)
    assert k > 10 and rel_err(hist[:k], o["hist"][:k]) < 1e-10


@pytest.mark.parametrize("force_ops", [False, True])
@pytest.mark.parametrize("cid,scale,sweeps,w", [(2, 32, 2, 1.0), (1, 64, 3, 0.8),
                                                (4, 16384, 2, 1.0), (5, 64, 3, 1.0)])
def test_cg_multisweep_jacobi(ctx, oracle, problems, cid, scale, sweeps, w, force_ops):
    """Jacobi preconditioner of several sweeps z += D^-1 (r - A z) from z = 0
    (src/Iter.cxx:66-77 with pcdIter > 1, example/TestPCG.cxx:71-76): the
    fused CG forms z1 = D^-1 r inside its update and each further sweep as
    one SpMV with the next sweep's max-norm; same bars as the one-sweep CG."""
    pr = problems(cid, scale)
    M = Matrix.of(pr.A)
    o = oracle.solve(CG, M, pr.b, np.zeros(M.nrow), pc="jacobi", weight=w, pc_iters=sweeps,
                     hist_cap=4000, **PRM)
    A = fx.Csr.from_problem(ctx, pr.A)
    P = fx.Precond(ctx, "jacobi", A, w, sweeps)
    s = fx.Solver(ctx, fx.CG)
    s.Setup(A, P, fx.params(**PRM))
    x = dev(ctx, np.zeros(M.nrow))
    old = os.environ.get("FASPXX_FX_NO_FAST")
    if force_ops:
        os.environ["FASPXX_FX_NO_FAST"] = "1"
    try:
        res = s.Solve(dev(ctx, pr.b), x)
    finally:
        if force_ops:
            if old is None:
                del os.environ["FASPXX_FX_NO_FAST"]
            else:
                os.environ["FASPXX_FX_NO_FAST"] = old
    assert s.last_path == ("ops" if force_ops else "fused")
    x, hist = x.numpy(), s.history()
    assert res.retcode == o["retcode"] == 0
    assert abs(res.numIter - o["numIter"]) <= 2, (res.numIter, o["numIter"])
    assert rel_err(x, o["x"]) < 1e-7
    assert np.linalg.norm(pr.b - pr.A.matvec(x)) / np.linalg.norm(pr.b) < 1e-8
    assert abs(res.norm2 - o["norm2"]) <= 1e-3 * o["norm2"] + 1e-14
    k = min(len(hist), 30)
    assert k > 5 and rel_err(hist[:k], o["hist"][:k]) < 1e-10


def test_cg_x0_and_defaults(ctx, oracle, problems):
    pr = problems(1, 64)
    M = Matrix.of(pr.A)
    x0 = np.ones(M.nrow)
    for prm in [dict(), dict(maxIter=25), dict(maxIter=25, minIter=5, relTol=1e-3)]:
        o = oracle.solve(CG, M, pr.b, x0, pc="jacobi", **prm)
        res, x, _ = gpu_solve(ctx, pr, fx.CG, "jacobi", x0=x0, **prm)
        assert res.numIter == o["numIter"] and res.retcode == o["retcode"], prm
        assert rel_err(x, o["x"]) < 1e-10


def test_cg_zero_rhs_and_maxit(ctx, problems):
    pr = problems(1, 32)

    class Z:
        A = pr.A
        b = np.zeros(pr.A.nrow)
    res, x, _ = gpu_solve(ctx, Z, fx.CG, maxIter=100, relTol=1e-8, absTol=0.0)
    assert res.numIter == 0 and np.all(x == 0)
    res, x, _ = gpu_solve(ctx, pr, fx.CG, maxIter=7, relTol=1e-12, absTol=0.0)
    assert res.numIter == 7


def test_cg_tolsmall_path(ctx, oracle, problems):
    """relTol below attainable accuracy exercises the host-side check III."""
    pr = problems(1, 16)
    M = Matrix.of(pr.A)
    prm = dict(maxIter=2000, relTol=1e-19, absTol=0.0, restart=2)
    o = oracle.solve(CG, M, pr.b, np.zeros(M.nrow), **prm)
    res, x, _ = gpu_solve(ctx, pr, fx.CG, **prm)
    assert res.retcode == o["retcode"]
    assert rel_err(x, o["x"]) < 1e-10


@pytest.mark.parametrize("force_ops", [False, True])
def test_jacobi_sweeps_bitexact(ctx, oracle, problems, force_ops):
    pr = problems(2, 32)
    M = Matrix.of(pr.A)
    o = oracle.solve(JACOBI, M, pr.b, np.zeros(M.nrow), weight=2 / 3, maxIter=50, relTol=1e-30,
                     absTol=0.0, hist_cap=60)
    res, x, hist = gpu_solve(ctx, pr, fx.JACOBI, weight=2 / 3, maxIter=50, relTol=1e-30,
                             absTol=0.0, force_ops=force_ops)
    assert res.numIter == o["numIter"] == 50
    assert np.array_equal(x, o["x"])
    assert abs(res.norm2 - o["norm2"]) <= 1e-13 * o["norm2"]
    assert rel_err(hist, o["hist"][:50]) < 1e-13


@pytest.mark.parametrize("maxIter", [1, 4, 29, 30, 31, 59, 60, 61])
def test_jacobi_replay_lengths(ctx, oracle, problems, maxIter):
    """Sweep counts around the replay lengths (sweeps are replayed in
    multiples of 6, up to 30 at a time; a run of maxIter sweeps launches
    maxIter + 1 and its last replay is cut to cover just those): bit-exact
    against the oracle, and a second solve on the same solver (which
    predicts its sweeps from the first) gives the same bits."""
    pr = problems(2, 32)
    M = Matrix.of(pr.A)
    prm = dict(maxIter=maxIter, relTol=1e-30, absTol=0.0)
    o = oracle.solve(JACOBI, M, pr.b, np.zeros(M.nrow), weight=2 / 3, hist_cap=80, **prm)
    A = fx.Csr.from_problem(ctx, pr.A)
    s = fx.Solver(ctx, fx.JACOBI)
    s.Setup(A, fx.Precond(ctx, "jacobi", A, 2 / 3, 1), fx.params(**prm))
    for k in range(2):
        x = dev(ctx, np.zeros(M.nrow))
        res = s.Solve(dev(ctx, pr.b), x)
        assert s.last_path == "fused"
        assert res.numIter == o["numIter"] == maxIter, (k, res.numIter, o["numIter"])
        assert np.array_equal(x.numpy(), o["x"])


@pytest.mark.parametrize("relTol", [1e-3, 1e-4])
def test_jacobi_tolerance_repeat(ctx, oracle, problems, relTol):
    """Weighted Jacobi stopped by the tolerance (36 and 128 sweeps at 32^3):
    the count within 1 of the oracle's and x bit-exact when the counts
    agree; a second solve on the same solver, whose last replay is cut to
    the first solve's count, gives the same count and bits."""
    pr = problems(2, 32)
    M = Matrix.of(pr.A)
    prm = dict(maxIter=2000, relTol=relTol, absTol=0.0)
    o = oracle.solve(JACOBI, M, pr.b, np.zeros(M.nrow), weight=2 / 3, hist_cap=2100, **prm)
    A = fx.Csr.from_problem(ctx, pr.A)
    s = fx.Solver(ctx, fx.JACOBI)
    s.Setup(A, fx.Precond(ctx, "jacobi", A, 2 / 3, 1), fx.params(**prm))
    out = []
    for _ in range(2):
        x = dev(ctx, np.zeros(M.nrow))
        res = s.Solve(dev(ctx, pr.b), x)
        assert s.last_path == "fused" and res.retcode == 0
        out.append((res.numIter, x.numpy()))
    assert abs(out[0][0] - o["numIter"]) <= 1, (out[0][0], o["numIter"])
    if out[0][0] == o["numIter"]:
        assert np.array_equal(out[0][1], o["x"])
    assert out[1][0] == out[0][0] and np.array_equal(out[1][1], out[0][1])


@pytest.mark.parametrize("relTol", [1e-3, 1e-4])
def test_jacobi_tolerance_at_max_iterations(ctx, oracle, problems, relTol):
    """A tolerance stop after K sweeps with maxIter = K - 1, K and K + 1
    (src/Iter.cxx:69-99: at maxIter = K the loop ends by its count with the
    same iterate; at K + 1 the sweep that decides the last allowed one finds
    convergence, and only its deciding workgroup runs): the counts, x bit
    for bit and the final residual against the unlimited solve and the
    oracle's fixed-count sweeps."""
    pr = problems(2, 32)
    M = Matrix.of(pr.A)
    A = fx.Csr.from_problem(ctx, pr.A)
    P = fx.Precond(ctx, "jacobi", A, 2 / 3, 1)

    def run(max_it, tol):
        s = fx.Solver(ctx, fx.JACOBI)
        s.Setup(A, P, fx.params(maxIter=max_it, relTol=tol, absTol=0.0))
        x = dev(ctx, np.zeros(M.nrow))
        res = s.Solve(dev(ctx, pr.b), x)
        assert s.last_path == "fused"
        return res, x.numpy()

    full, x_full = run(2000, relTol)
    K = full.numIter
    assert 2 < K < 2000
    for max_it, want_it in [(K - 1, K - 1), (K, K), (K + 1, K)]:
        res, x = run(max_it, relTol)
        assert res.numIter == want_it, (max_it, res.numIter, K)
        o = oracle.solve(JACOBI, M, pr.b, np.zeros(M.nrow), weight=2 / 3, maxIter=want_it,
                         relTol=1e-30, absTol=0.0, hist_cap=2100)
        assert np.array_equal(x, o["x"]), max_it
        assert abs(res.norm2 - o["norm2"]) <= 1e-13 * o["norm2"], max_it
        if want_it == K:
            assert np.array_equal(x, x_full)


def test_jacobi_converges_unit(ctx):
    """UnitTestsJacobi 3x3 (utest/src/UnitTestsJacobi.cxx:85-157)."""
    class PR:
        pass
    from faspxx_amd.problems import CSR
    PR.A = CSR(3, 3, np.array([0, 3, 6, 9], np.uint32), np.array([0, 1, 2] * 3, np.uint32),
               np.array([4, 2, 1, 3, 8, 2, 2, 1, 9], np.float64), np.array([0, 4, 8], np.uint32))
    PR.b = np.ones(3)
    for w in (1.0, 0.5):
        res, x, _ = gpu_solve(ctx, PR, fx.JACOBI, weight=w, maxIter=100, relTol=1e-9)
        assert np.all(np.abs(x - [0.221719, 0.027149, 0.058824]) < 1e-5)


def oracle_band(oracle, kind, M, b, seeds=(1, 2, 3, 4), **kw):
    """The oracle on b and on one-ulp perturbations of b: its result on b,
    the lowest and highest iteration count over all of them, and the largest
    relative distance of a perturbed solution from the solution on b."""
    base = oracle.solve(kind, M, b, np.zeros(M.nrow), **kw)
    counts, dx = [base["numIter"]], 0.0
    for s in seeds:
        o = oracle.solve(kind, M, ulp_perturbed(b, s), np.zeros(M.nrow), **kw)
        counts.append(o["numIter"])
        dx = max(dx, rel_err(o["x"], base["x"]))
    return base, min(counts), max(counts), dx


@pytest.mark.parametrize("force_ops", [False, True])
@pytest.mark.parametrize("cid,scale,pc", [(3, 128, "jacobi"), (3, 128, "identity"),
                                          (3, 256, "jacobi"), (3, 256, "identity")])
def test_bicgstab_parity(ctx, oracle, problems, cid, scale, pc, force_ops):
    """Fused device loop (default) and per-operation path against the oracle,
    on the nonsymmetric system BiCGStab is meant for.  (On the symmetric
    Poisson and variable-coefficient systems the reference's P.A BiCGStab is
    erratic: on the 64 x 64 variable-coefficient system the oracle's own
    iterate grows to 1e12, so those make no parity case.)

    The count bar is +-2 around the oracle's counts on b and on one-ulp
    perturbations of b: at 256 x 256 the oracle itself stops after 375 or 378
    iterations depending on the last bit of a few entries of b, and a
    parallel dot product changes the rounding by at least as much.  Likewise
    the solution bar is 1e-6 or twice the largest distance between the
    oracle's own solutions on b and on the perturbed b, whichever is larger.
    The device solution's true relative residual is held to 2e-8."""
    pr = problems(cid, scale)
    M = Matrix.of(pr.A)
    prm = dict(pc=pc, **PRM)
    o, lo, hi, dx = oracle_band(oracle, BICGSTAB, M, pr.b, hist_cap=4000, **prm)
    res, x, hist = gpu_solve(ctx, pr, fx.BICGSTAB, pc, force_ops=force_ops, **PRM)
    assert res.retcode == o["retcode"] == 0
    assert lo - 2 <= res.numIter <= hi + 2, (res.numIter, o["numIter"], lo, hi)
    if scale == 128:
        assert lo == hi and abs(res.numIter - o["numIter"]) <= 2
    assert rel_err(x, o["x"]) < max(1e-6, 2 * dx), (rel_err(x, o["x"]), dx)
    assert np.linalg.norm(pr.b - pr.A.matvec(x)) < 2e-8 * np.linalg.norm(pr.b)
    k = min(len(hist), 20)
    assert k > 5 and rel_err(hist[:k], o["hist"][:k]) < 1e-8


def test_bicgstab_x0_maxit_and_zero_rhs(ctx, oracle, problems):
    pr = problems(3, 64)
    M = Matrix.of(pr.A)
    x0 = np.random.default_rng(5).uniform(-1, 1, M.nrow)
    for prm in [dict(maxIter=13, relTol=1e-12, absTol=0.0), dict(maxIter=40, minIter=3),
                dict(maxIter=1000, relTol=1e-6, absTol=0.0, savIter=5)]:
        o = oracle.solve(BICGSTAB, M, pr.b, x0, pc="jacobi", **prm)
        res, x, _ = gpu_solve(ctx, pr, fx.BICGSTAB, "jacobi", x0=x0, **prm)
        assert res.retcode == o["retcode"], prm
        assert abs(res.numIter - o["numIter"]) <= 2, (prm, res.numIter, o["numIter"])
        assert rel_err(x, o["x"]) < 1e-6, prm

    class Z:
        A = pr.A
        b = np.zeros(pr.A.nrow)
    res, x, _ = gpu_solve(ctx, Z, fx.BICGSTAB, "jacobi", maxIter=100, relTol=1e-8, absTol=0.0)
    assert res.numIter == 0 and np.all(x == 0)


def test_bicgstab_convdiff_full_size(ctx, problems):
    """The 1024 x 1024 upwind convection-diffusion system (n = 1,048,576),
    Jacobi-preconditioned, rtol 1e-8: both paths reach the true relative
    residual, in a count within +-2 of the band of oracle counts recorded in
    tests/golden/solver_iterations.json (the oracle on b and on four one-ulp
    perturbations of b; the oracle takes about 90 s per solve at this size,
    too long to run here)."""
    gold = json.loads((pathlib.Path(__file__).parent / "golden" / "solver_iterations.json")
                      .read_text())["bicgstab_convdiff_1024_jacobi_rtol1e-8"]
    counts = [gold["numIter"]] + gold["numIter_b_ulp_perturbed"]
    lo, hi = min(counts), max(counts)
    pr = problems(3, None)
    for force_ops in (False, True):
        res, x, _ = gpu_solve(ctx, pr, fx.BICGSTAB, "jacobi", force_ops=force_ops,
                              maxIter=5000, relTol=1e-8, absTol=0.0)
        assert res.retcode == gold["retcode"] == 0
        true_res = np.linalg.norm(pr.b - pr.A.matvec(x)) / np.linalg.norm(pr.b)
        assert true_res < 2e-8, true_res
        assert lo - 2 <= res.numIter <= hi + 2, (force_ops, res.numIter, lo, hi)


@pytest.mark.parametrize("force_ops", [False, True])
@pytest.mark.parametrize("left", [0, 1])
@pytest.mark.parametrize("pc", ["jacobi", "identity"])
def test_gmres_parity(ctx, oracle, problems, pc, left, force_ops):
    """Fused cycle graphs (default) and per-operation path against the oracle."""
    pr = problems(3, 64)
    M = Matrix.of(pr.A)
    prm = dict(maxIter=4000, relTol=1e-8, absTol=0.0, leftPrecond=left)
    o = oracle.solve(GMRES, M, pr.b, np.zeros(M.nrow), pc=pc, hist_cap=5000, **prm)
    res, x, hist = gpu_solve(ctx, pr, fx.GMRES, pc, force_ops=force_ops, **prm)
    assert res.retcode == o["retcode"] == 0
    assert abs(res.numIter - o["numIter"]) <= 2, (res.numIter, o["numIter"])
    assert rel_err(x, o["x"]) < 1e-6
    assert abs(res.norm2 - o["norm2"]) <= 1e-4 * o["norm2"]
    k = min(len(hist), 40)
    assert k > 30 and rel_err(hist[:k], o["hist"][:k]) < 1e-8


@pytest.mark.parametrize("prm", [dict(maxIter=17, relTol=1e-12), dict(maxIter=200, minIter=12),
                                 dict(maxIter=500, restart=7, maxRestart=12, minRestart=4)])
def test_gmres_parameters(ctx, oracle, problems, prm):
    """maxIter inside a cycle, minIter, and the variable-restart clamping."""
    pr = problems(3, 48)
    M = Matrix.of(pr.A)
    prm = dict(absTol=0.0, **prm)
    x0 = np.random.default_rng(9).uniform(-1, 1, M.nrow)
    o = oracle.solve(GMRES, M, pr.b, x0, pc="jacobi", **prm)
    res, x, _ = gpu_solve(ctx, pr, fx.GMRES, "jacobi", x0=x0, **prm)
    assert res.retcode == o["retcode"], prm
    assert abs(res.numIter - o["numIter"]) <= 2, (prm, res.numIter, o["numIter"])
    assert rel_err(x, o["x"]) < 1e-6, prm


def test_gmres_convdiff_full_size(ctx, problems):
    """GMRES(30), Jacobi right-preconditioned, on the 1024 x 1024 system to
    rtol 1e-8: the fused and per-operation paths agree within 2 iterations."""
    pr = problems(3, None)
    out = []
    for force_ops in (False, True):
        res, x, _ = gpu_solve(ctx, pr, fx.GMRES, "jacobi", force_ops=force_ops,
                              maxIter=20000, relTol=1e-8, absTol=0.0)
        assert res.retcode == 0
        true_res = np.linalg.norm(pr.b - pr.A.matvec(x)) / np.linalg.norm(pr.b)
        assert true_res < 2e-8, true_res
        out.append(res.numIter)
    assert abs(out[0] - out[1]) <= 2, out


@pytest.mark.parametrize("force_ops", [False, True])
def test_fgmres_parity(ctx, oracle, problems, force_ops):
    """Fused cycle graphs (default) and per-operation path against the oracle."""
    pr = problems(3, 48)
    M = Matrix.of(pr.A)
    o = oracle.solve(FGMRES, M, pr.b, np.zeros(M.nrow), pc="identity", hist_cap=4000, **PRM)
    res, x, hist = gpu_solve(ctx, pr, fx.FGMRES, "identity", force_ops=force_ops, **PRM)
    assert abs(res.numIter - o["numIter"]) <= 2
    assert rel_err(x, o["x"]) < 1e-6
    k = min(len(hist), 20)
    assert k > 5 and rel_err(hist[:k], o["hist"][:k]) < 1e-8
    # Jacobi with the reference's stale initial guess (src/FGMRES.cxx:158-159):
    # same trajectory for a fixed iteration budget
    prm = dict(maxIter=200, relTol=1e-12, absTol=0.0)
    o = oracle.solve(FGMRES, M, pr.b, np.zeros(M.nrow), pc="jacobi", **prm)
    res, x, _ = gpu_solve(ctx, pr, fx.FGMRES, "jacobi", force_ops=force_ops, **prm)
    assert res.numIter == o["numIter"] == 200
    assert abs(res.norm2 - o["norm2"]) <= 1e-6 * o["norm2"]


def test_fgmres_stale_z_across_solves(ctx, problems):
    """The reference keeps Z in the solver, so a second Solve starts its
    Jacobi applications from the first solve's Z: the fused path and the
    per-operation path carry it over alike."""
    pr = problems(3, 48)
    A = fx.Csr.from_problem(ctx, pr.A)
    P = fx.Precond(ctx, "jacobi", A, 1.0, 1)
    prm = fx.params(maxIter=150, relTol=1e-12, absTol=0.0)
    out = []
    for force_ops in (False, True):
        s = fx.Solver(ctx, fx.FGMRES)
        s.Setup(A, P, prm)
        norms = []
        for _ in range(2):
            x = dev(ctx, np.zeros(pr.A.nrow))
            old = os.environ.get("FASPXX_FX_NO_FAST")
            if force_ops:
                os.environ["FASPXX_FX_NO_FAST"] = "1"
            try:
                r = s.Solve(dev(ctx, pr.b), x)
            finally:
                if force_ops:
                    if old is None:
                        del os.environ["FASPXX_FX_NO_FAST"]
                    else:
                        os.environ["FASPXX_FX_NO_FAST"] = old
            norms.append(r.norm2)
        out.append(norms)
        s.close()
    assert out[0][0] != out[0][1]  # the second solve starts from a different Z
    for a, b in zip(out[0], out[1]):
        assert abs(a - b) <= 1e-6 * b


def test_krylov_front_end(ctx, problems):
    pr = problems(1, 32)
    A = fx.Csr.from_problem(ctx, pr.A)
    P = fx.Precond(ctx, "identity")
    for name, ok in [("CG", True), ("BiCGStab", True), ("gmres", True), ("FGMRES", True),
                     ("nonsense", True)]:
        x = fx.Vec(ctx, pr.A.nrow)
        res = fx.krylov(ctx, A, dev(ctx, pr.b), x, P, name, fx.params(maxIter=2000, relTol=1e-8,
                                                                      absTol=0.0))
        assert res.retcode == 0
        assert np.linalg.norm(pr.b - pr.A.matvec(x.numpy())) / np.linalg.norm(pr.b) < 1e-7


@pytest.mark.parametrize("cid,scale", [(2, 32), (5, 64), (4, 16384)])
def test_ir_mixed_precision(ctx, oracle, problems, cid, scale):
    pr = problems(cid, scale)
    A = fx.Csr.from_problem(ctx, pr.A)
    x = fx.Vec(ctx, pr.A.nrow)
    res, inner = fx.solve_ir(ctx, A, dev(ctx, pr.b), x, 1.0,
                             fx.params(maxIter=50, relTol=1e-10, absTol=0.0),
                             fx.params(maxIter=5000, relTol=1e-4, absTol=0.0))
    assert res.retcode == 0
    xr = x.numpy()
    true_res = np.linalg.norm(pr.b - pr.A.matvec(xr)) / np.linalg.norm(pr.b)
    assert true_res < 1e-10
    M = Matrix.of(pr.A)
    o = oracle.solve(CG, M, pr.b, np.zeros(M.nrow), pc="jacobi", maxIter=20000, relTol=1e-12,
                     absTol=0.0)
    assert rel_err(xr, o["x"]) < 1e-6


def _ir_solve(ctx, pr, outer, inner, monkeypatch, refcg):
    """One refinement solve; refcg forces the f32 Jacobi-PCG inner solver
    instead of the scaled single-reduction one (read when the solver is made)."""
    if refcg:
        monkeypatch.setenv("FASPXX_FX_IR_REFCG", "1")
    else:
        monkeypatch.delenv("FASPXX_FX_IR_REFCG", raising=False)
    A = fx.Csr.from_problem(ctx, pr.A)
    x = fx.Vec(ctx, pr.A.nrow)
    ir = fx.IR(ctx, A, 1.0, outer, inner)
    res, inner_its = ir.Solve(dev(ctx, pr.b), x)
    ir.close()
    monkeypatch.delenv("FASPXX_FX_IR_REFCG", raising=False)
    return res, inner_its, x.numpy()


@pytest.mark.parametrize("cid,scale", [(2, 32), (5, 64), (4, 16384), (1, 64)])
def test_ir_scaled_inner_matches_jacobi_inner(ctx, problems, monkeypatch, cid, scale):
    """The scaled inner solver (CG on S A S, S = diag(a_ii)^-1/2, one
    reduction phase per iteration) is Jacobi-PCG in exact arithmetic: with
    the same tolerances it reaches the FP64 outer tolerance in the same
    number of outer steps (+-1) and a similar number of inner iterations as
    the f32 Jacobi-PCG inner solver, and the two solutions agree to the
    accuracy the tolerance allows."""
    pr = problems(cid, scale)
    outer = fx.params(maxIter=50, relTol=1e-10, absTol=0.0)
    inner = fx.params(maxIter=5000, relTol=1e-4, absTol=0.0)
    rs, its_s, xs = _ir_solve(ctx, pr, outer, inner, monkeypatch, False)
    rj, its_j, xj = _ir_solve(ctx, pr, outer, inner, monkeypatch, True)
    assert rs.retcode == 0 and rj.retcode == 0
    nb = np.linalg.norm(pr.b)
    assert np.linalg.norm(pr.b - pr.A.matvec(xs)) < 1e-10 * nb
    assert abs(rs.numIter - rj.numIter) <= 1
    assert abs(its_s - its_j) <= max(4, 0.15 * its_j), (its_s, its_j)
    assert rel_err(xs, xj) < 1e-8


def test_ir_nonpositive_diagonal_uses_jacobi_inner(ctx, problems, monkeypatch):
    """A matrix with a non-positive diagonal (here -A of 2D Poisson) cannot
    be scaled by diag(a_ii)^-1/2: the refinement then runs the f32
    Jacobi-PCG inner solver, bit for bit the same solve as when that solver
    is forced, and still converges (the signs cancel in Jacobi-PCG)."""
    pr = problems(1, 32)
    negA = dataclasses.replace(pr.A, values=-pr.A.values)
    neg = dataclasses.replace(pr, A=negA, b=-pr.b)
    outer = fx.params(maxIter=50, relTol=1e-10, absTol=0.0)
    inner = fx.params(maxIter=5000, relTol=1e-4, absTol=0.0)
    rd, its_d, xd = _ir_solve(ctx, neg, outer, inner, monkeypatch, False)
    rf, its_f, xf = _ir_solve(ctx, neg, outer, inner, monkeypatch, True)
    assert rd.retcode == 0 and its_d == its_f and rd.numIter == rf.numIter
    assert np.array_equal(xd, xf)
    assert np.linalg.norm(neg.b - neg.A.matvec(xd)) < 1e-10 * np.linalg.norm(neg.b)


def test_ir_scaled_inner_iteration_cap(ctx, problems, monkeypatch):
    """An inner solve cut off by its maxIter still hands its iterate to the
    outer loop, which then converges in more outer steps."""
    pr = problems(1, 16)
    outer = fx.params(maxIter=400, relTol=1e-10, absTol=0.0)
    inner = fx.params(maxIter=10, relTol=1e-6, absTol=0.0)
    res, its, xs = _ir_solve(ctx, pr, outer, inner, monkeypatch, False)
    assert res.retcode == 0 and res.numIter > 2
    assert its <= 10 * res.numIter
    assert np.linalg.norm(pr.b - pr.A.matvec(xs)) < 1e-10 * np.linalg.norm(pr.b)


def _f32_fma(a, b, c):
    """fma in f32 from f64: the product of two f32 values is exact in f64;
    the f64 sum is then rounded to f32 (a double rounding, which can differ
    from one correctly rounded fma by an ulp in rare cases)."""
    return (np.float64(a) * np.asarray(b, np.float64)
            + np.asarray(c, np.float64)).astype(np.float32)


def _scaled_inner_cg(A, b, m):
    """m iterations of the refinement's inner solve from x0 = 0, restated in
    numpy: single-reduction CG on S A S (S = diag(a_ii)^-1/2, values rounded
    to f32), f32 vectors, f64 sums, y += alpha p in every iteration.  Returns
    the outer step's x = S y (f64) and the sums (r, r) at the start of
    iterations 0..m, which decide where a tolerance stops the solve."""
    M = A.to_scipy().tocsr()
    d = M.diagonal()
    s = 1.0 / np.sqrt(d)
    rows = np.repeat(np.arange(A.nrow), np.diff(A.row_ptr.astype(np.int64)))
    cols = A.col_ind.astype(np.int64)
    vs = ((A.values * s[rows]) * s[cols]).astype(np.float32)
    import scipy.sparse as sp
    Ah = sp.csr_matrix((vs.astype(np.float64), cols, A.row_ptr.astype(np.int64)),
                       shape=(A.nrow, A.nrow))
    r = (s * b).astype(np.float32)
    gamma0 = float(np.dot(r.astype(np.float64), r.astype(np.float64)))
    p = sp_ = y = None
    gamma_prev = alpha_prev = 0.0
    gamma = gamma0
    gammas = [gamma0]
    for it in range(m):
        w = (Ah @ r.astype(np.float64)).astype(np.float32)
        delta = float(np.dot(w.astype(np.float64), r.astype(np.float64)))
        if it == 0:
            alpha, beta = gamma / delta, 0.0
        else:
            beta = gamma / gamma_prev
            alpha = gamma / (delta - beta * gamma / alpha_prev)
        a, bb = np.float32(alpha), np.float32(beta)
        if it == 0:
            p, sp_, y = r.copy(), w.copy(), (a * r).astype(np.float32)
        else:
            p = _f32_fma(bb, p, r)
            sp_ = _f32_fma(bb, sp_, w)
            y = _f32_fma(a, p, y)
        r = _f32_fma(-a, sp_, r)
        gamma_prev, alpha_prev = gamma, alpha
        gamma = float(np.dot(r.astype(np.float64), r.astype(np.float64)))
        gammas.append(gamma)
    return s * y.astype(np.float64), gammas


@pytest.mark.parametrize("m", [1, 2, 3, 4, 5, 8, 33])
def test_ir_scaled_inner_increment(ctx, problems, m):
    """The increment one outer step adds, after exactly m inner iterations
    (odd and even m; 33 crosses a replay boundary), against the numpy
    restatement of the inner solve.  The device adds alpha p to y two
    iterations at a time and leaves the last term of an odd count to the
    outer step, so an error there would show here, where the refinement's
    final residual would hide it (the outer loop corrects any increment).
    The inner solve is stopped once by its maxIter and once by a tolerance
    chosen between the sums of iterations m - 1 and m; the latter is solved
    twice on one refinement object, and the second solve, predicted from
    the first, queues its outer step behind the inner solve's last replay
    (the increment is then added there)."""
    pr = problems(1, 32)
    want, gammas = _scaled_inner_cg(pr.A, pr.b, m)
    outer = fx.params(maxIter=1, relTol=1e-14, absTol=0.0)
    A = fx.Csr.from_problem(ctx, pr.A)
    # stopped by maxIter
    ir = fx.IR(ctx, A, 1.0, outer, fx.params(maxIter=m, relTol=1e-30, absTol=0.0))
    try:
        x = fx.Vec(ctx, pr.A.nrow)
        _, inner_its = ir.Solve(dev(ctx, pr.b), x)
        assert inner_its == m
        assert rel_err(x.numpy(), want) < 1e-5, (m, rel_err(x.numpy(), want))
    finally:
        ir.close()
    # stopped by the tolerance at iteration m: the first gamma below thr
    lo, hi = gammas[m], min(gammas[:m])
    assert lo < 0.95 * hi, (m, lo, hi)  # a margin far above the f32 rounding differences
    rel_tol = float(np.sqrt(np.sqrt(lo * hi) / gammas[0]))
    ir = fx.IR(ctx, A, 1.0, outer, fx.params(maxIter=10 * m + 10, relTol=rel_tol, absTol=0.0))
    try:
        for _ in range(2):
            x = fx.Vec(ctx, pr.A.nrow)
            _, inner_its = ir.Solve(dev(ctx, pr.b), x)
            assert inner_its == m
            assert rel_err(x.numpy(), want) < 1e-5, (m, rel_err(x.numpy(), want))
    finally:
        ir.close()


@pytest.mark.parametrize("cid,scale", [(3, 128), (1, 64)])
def test_ir_gmres_inner(ctx, oracle, problems, cid, scale):
    """FP32-inner GMRES (f64 Hessenberg / Givens) with FP64 refinement, on the
    nonsymmetric upwind system and on 2D Poisson: the FP64 relative residual
    meets the outer tolerance and the solution matches the oracle's FP64
    GMRES to the accuracy the tolerance allows."""
    pr = problems(cid, scale)
    A = fx.Csr.from_problem(ctx, pr.A)
    x = fx.Vec(ctx, pr.A.nrow)
    res, inner = fx.solve_ir(ctx, A, dev(ctx, pr.b), x, 1.0,
                             fx.params(maxIter=50, relTol=1e-10, absTol=0.0),
                             fx.params(maxIter=5000, relTol=1e-4, absTol=0.0), fx.GMRES)
    assert res.retcode == 0 and res.numIter >= 2 and inner > 0
    xr = x.numpy()
    assert np.linalg.norm(pr.b - pr.A.matvec(xr)) < 1e-10 * np.linalg.norm(pr.b)
    M = Matrix.of(pr.A)
    o = oracle.solve(GMRES, M, pr.b, np.zeros(M.nrow), pc="jacobi", maxIter=20000,
                     relTol=1e-12, absTol=0.0)
    assert rel_err(xr, o["x"]) < 1e-6


def test_ir_headline_configuration(ctx, problems):
    """The benchmark's solve: 3D Poisson 128^3, FP32-inner CG (Jacobi, inner
    rtol 1e-4) refined in FP64 to 1e-8: two outer steps, FP64 relative
    residual below 1e-8, about as many inner iterations as the FP64 CG
    takes, and an error against the exact solution no worse than twice the
    FP64 CG's."""
    pr = problems(2, None)
    A = fx.Csr.from_problem(ctx, pr.A)
    b = dev(ctx, pr.b)
    x = fx.Vec(ctx, pr.A.nrow)
    res, inner = fx.solve_ir(ctx, A, b, x, 1.0, fx.params(maxIter=50, relTol=1e-8, absTol=0.0),
                             fx.params(maxIter=20000, relTol=1e-4, absTol=0.0))
    assert res.retcode == 0 and res.numIter == 2
    xr = x.numpy()
    assert np.linalg.norm(pr.b - pr.A.matvec(xr)) < 1e-8 * np.linalg.norm(pr.b)
    r64, x64, _ = gpu_solve(ctx, pr, fx.CG, "jacobi", maxIter=20000, relTol=1e-8, absTol=0.0)
    assert r64.retcode == 0 and abs(inner - r64.numIter) <= 10
    # the error against x* (b = A x*) is the FP64 solve's, to within a factor 2
    e_ir, e_64 = rel_err(xr, pr.xstar), rel_err(x64, pr.xstar)
    assert e_ir < 2 * e_64 + 1e-12, (e_ir, e_64)


@pytest.mark.parametrize("kind,okind", [(fx.GMRES, GMRES), (fx.FGMRES, FGMRES)])
@pytest.mark.parametrize("savIter", [10, 5000])
def test_gmres_safeguard_fused(ctx, oracle, problems, kind, okind, savIter):
    """The safe-guard copy (savIter <= maxIter, src/GMRES.cxx:234,277,
    src/FGMRES.cxx:224,268) on the fused cycle graphs: the reference's test
    after each cycle never holds (resAbsOld = resAbs after every inner
    step), so a solve that runs past savIter returns its initial guess, and
    one that stops before it returns its solution.  Same count and x as the
    oracle and the per-operation path."""
    pr = problems(3, 32)
    prm = dict(maxIter=8000, relTol=1e-8, absTol=0.0, restart=20, savIter=savIter)
    M = Matrix.of(pr.A)
    x0 = np.random.default_rng(9).uniform(-1, 1, M.nrow)
    o = oracle.solve(okind, M, pr.b, x0, pc="jacobi", **prm)
    outs = []
    for force_ops in (False, True):
        A = fx.Csr.from_problem(ctx, pr.A)
        P = fx.Precond(ctx, "jacobi", A, 1.0, 1)
        s = fx.Solver(ctx, kind)
        s.Setup(A, P, fx.params(**prm))
        x = dev(ctx, x0)
        old = os.environ.get("FASPXX_FX_NO_FAST")
        if force_ops:
            os.environ["FASPXX_FX_NO_FAST"] = "1"
        try:
            res = s.Solve(dev(ctx, pr.b), x)
        finally:
            if force_ops:
                if old is None:
                    del os.environ["FASPXX_FX_NO_FAST"]
                else:
                    os.environ["FASPXX_FX_NO_FAST"] = old
        assert s.last_path == ("ops" if force_ops else "fused")
        outs.append((res, x.numpy()))
    for res, x in outs:
        assert abs(res.numIter - o["numIter"]) <= 2, (res.numIter, o["numIter"])
        if o["numIter"] > savIter:
            assert np.array_equal(x, x0) and np.array_equal(o["x"], x0)
        else:
            assert rel_err(x, o["x"]) < 1e-6


def _sine_mode(m):
    """An eigenvector of the m^3 7-point Laplacian (Dirichlet-zero boundary):
    with it as the right-hand side, Jacobi-PCG and the refinement's scaled
    CG stop after one iteration."""
    t = np.sin(np.pi * np.arange(1, m + 1) / (m + 1))
    return np.einsum("k,j,i->kji", t, t, t).reshape(-1).copy()


@pytest.mark.parametrize("kind", [fx.CG, fx.BICGSTAB])
def test_solve_after_short_solve(ctx, problems, kind):
    """A solve on a set-up solver whose previous solve stopped within one
    replay: the replay loop is then predicted to stop within its first
    replay, finds it running, and must go on waiting for the replays of
    this solve only (not a host slot that the previous solve left holding
    its final state).  Same count and bit-identical x as a fresh solver."""
    m = 32
    pr = problems(2, m)
    A = fx.Csr.from_problem(ctx, pr.A)
    P = fx.Precond(ctx, "jacobi", A, 1.0, 1)
    s = fx.Solver(ctx, kind)
    s.Setup(A, P, fx.params(**PRM))
    x = dev(ctx, np.zeros(pr.A.nrow))
    short = s.Solve(dev(ctx, pr.A.matvec(_sine_mode(m))), x)
    assert short.retcode == 0 and 1 <= short.numIter <= 2, short.numIter
    x = dev(ctx, np.zeros(pr.A.nrow))
    res = s.Solve(dev(ctx, pr.b), x)
    assert s.last_path == "fused"
    ref, xref, _ = gpu_solve(ctx, pr, kind, "jacobi", **PRM)
    assert res.retcode == ref.retcode == 0
    assert res.numIter == ref.numIter > 40, (res.numIter, ref.numIter)
    assert np.array_equal(x.numpy(), xref)


def test_ir_solve_after_short_solve(ctx, problems):
    """The refinement's inner solves are predicted from the same outer step
    of the previous solve; after a solve whose first inner solve took one
    iteration (the right-hand side is an eigenvector), a solve whose first
    inner solve needs many runs it to the end: same outer and inner counts
    and bit-identical x as a fresh refinement object."""
    m = 32
    pr = problems(2, m)
    outer = fx.params(maxIter=50, relTol=1e-10, absTol=0.0)
    inner = fx.params(maxIter=5000, relTol=1e-4, absTol=0.0)
    A = fx.Csr.from_problem(ctx, pr.A)
    ir = fx.IR(ctx, A, 1.0, outer, inner)
    x = fx.Vec(ctx, pr.A.nrow)
    r1, in1 = ir.Solve(dev(ctx, pr.A.matvec(_sine_mode(m))), x)
    assert r1.retcode == 0
    x.set(0.0)
    r2, in2 = ir.Solve(dev(ctx, pr.b), x)
    ir.close()
    A2 = fx.Csr.from_problem(ctx, pr.A)
    ir2 = fx.IR(ctx, A2, 1.0, outer, inner)
    x2 = fx.Vec(ctx, pr.A.nrow)
    r3, in3 = ir2.Solve(dev(ctx, pr.b), x2)
    ir2.close()
    assert r2.retcode == r3.retcode == 0
    assert (r2.numIter, in2) == (r3.numIter, in3) and in3 > 40, (r2.numIter, in2, r3.numIter, in3)
    assert np.array_equal(x.numpy(), x2.numpy())


@pytest.mark.parametrize("kind", [fx.CG])
def test_solve_after_values_set(ctx, kind):
    """A solver set up on a structure-only matrix (every stored value 1,
    here the identity's pattern) and solved once, then given values: its
    next solve uses them, since the matrix is then another operator to it
    (its replay graphs were made for a matrix without values)."""
    n = 20000
    idx = np.arange(n, dtype=np.uint32)
    A = fx.Csr(ctx, n, n, np.arange(n + 1, dtype=np.uint32), idx)
    P = fx.Precond(ctx, "identity", A, 1.0, 1)
    s = fx.Solver(ctx, kind)
    s.Setup(A, P, fx.params(maxIter=100, relTol=1e-12, absTol=0.0))
    b = np.random.default_rng(5).uniform(1.0, 2.0, n)
    x = dev(ctx, np.zeros(n))
    r = s.Solve(dev(ctx, b), x)
    assert s.last_path == "fused" and r.retcode == 0
    assert rel_err(x.numpy(), b) < 1e-12
    d = 2.0 + (np.arange(n) % 3)
    vals = dev(ctx, d)
    A.set_values_dev(vals.data_ptr)
    x = dev(ctx, np.zeros(n))
    r = s.Solve(dev(ctx, b), x)
    assert s.last_path == "fused" and r.retcode == 0
    assert rel_err(x.numpy(), b / d) < 1e-12


def test_ir_outer_steps_queued_ahead(ctx, problems):
    """The refinement queues an inner solve before the host has looked at
    the residual that precedes it when the last solve went on past that
    residual; the residual's finisher stops it on the device when the outer
    loop is in fact done.  A zero right-hand side (done at once, with x left
    at 0) after a solve of several outer steps takes that path; the solves
    of the real right-hand side before and after it, the last one queued
    ahead again, agree bit for bit."""
    pr = problems(2, 32)
    n = pr.A.nrow
    outer = fx.params(maxIter=50, relTol=1e-10, absTol=0.0)
    inner = fx.params(maxIter=5000, relTol=1e-4, absTol=0.0)
    A = fx.Csr.from_problem(ctx, pr.A)
    ir = fx.IR(ctx, A, 1.0, outer, inner)
    runs = []
    for rhs in (pr.b, np.zeros(n), pr.b, pr.b):
        x = fx.Vec(ctx, n)
        res, inner_its = ir.Solve(dev(ctx, rhs), x)
        assert res.retcode == 0
        runs.append((res.numIter, inner_its, x.numpy()))
    ir.close()
    assert runs[0][0] >= 2
    assert runs[1][0] == 0 and runs[1][1] == 0 and not np.any(runs[1][2])
    for k in (2, 3):
        assert runs[k][:2] == runs[0][:2], (runs[k][:2], runs[0][:2])
        assert np.array_equal(runs[k][2], runs[0][2])


def test_ir_sorted_rows(ctx, problems):
    """The refinement on a matrix stored with its rows sorted by length (the
    power-law system at n = 200000): its f64 residual reads b and the
    scaling in that stored order.  Two solves on one object (the second
    with its outer steps queued ahead and behind) reach the outer tolerance
    in true FP64 residual and agree bit for bit."""
    pr = problems(4, 200000)
    n = pr.A.nrow
    A = fx.Csr.from_problem(ctx, pr.A)
    assert A.layout == "sorted"
    ir = fx.IR(ctx, A, 1.0, fx.params(maxIter=50, relTol=1e-10, absTol=0.0),
               fx.params(maxIter=5000, relTol=1e-4, absTol=0.0))
    runs = []
    for _ in range(2):
        x = fx.Vec(ctx, n)
        res, inner = ir.Solve(dev(ctx, pr.b), x)
        assert res.retcode == 0
        runs.append((res.numIter, inner, x.numpy()))
    ir.close()
    xr = runs[0][2]
    assert np.linalg.norm(pr.b - pr.A.matvec(xr)) < 1e-10 * np.linalg.norm(pr.b)
    assert runs[1][:2] == runs[0][:2] and np.array_equal(runs[1][2], xr)
