"""Full-size runs of the named configurations on the GPU against the
oracle's results on the same inputs (tests/golden/solver_iterations.json,
written on the CPU by tests/golden/make_iterations.py; the oracle is pinned
against the compiled reference by test_oracle*.py).

* 3D Poisson 128^3: Jacobi-PCG to 1e-8 (example/TestPCG.cxx:71-88) -- count
  +-2, final norm, solution; 50 weighted-Jacobi sweeps, omega = 2/3
  (example/TestJacobi.cxx:69-81) -- bit for bit (SHA-256 of x); the FP32 /
  FP64 refinement's solution against the oracle's.
* 1024^2 upwind convection-diffusion: GMRES(30) + Jacobi to 1e-8 -- count
  +-2; FGMRES(30) + Jacobi after its 2000-iteration budget -- the residual
  reached; BiCGStab + Jacobi -- the solution's error against x* beside the
  oracle's (its count band is tested in test_gpu_solvers.py).
* Power-law system, n = 2^20: FP64 Jacobi-CG to 1e-10 -- count +-2 and the
  solution against the oracle's, run here; the refinement's solution
  against the same.
"""
import hashlib
import json
import pathlib

import numpy as np
import pytest

from conftest import rel_err
from faspxx_amd import fx
from oracle import CG, Matrix

pytestmark = pytest.mark.gpu

GOLD = json.loads((pathlib.Path(__file__).parent / "golden" / "solver_iterations.json").read_text())


def dev(ctx, a):
    return fx.Vec.from_numpy(ctx, np.asarray(a))


def solve(ctx, pr, kind, weight=1.0, **prm):
    A = fx.Csr.from_problem(ctx, pr.A)
    P = fx.Precond(ctx, "jacobi", A, weight, 1)
    s = fx.Solver(ctx, kind)
    s.Setup(A, P, fx.params(**prm))
    x = dev(ctx, np.zeros(pr.A.nrow))
    res = s.Solve(dev(ctx, pr.b), x)
    return res, x.numpy(), A


def relres(pr, x):
    return np.linalg.norm(pr.b - pr.A.matvec(x)) / np.linalg.norm(pr.b)


def samples(x):
    return x[np.linspace(0, x.shape[0] - 1, 16).astype(np.int64)]


def test_pcg_poisson3d_128(ctx, problems):
    g = GOLD["pcg_poisson3d_128_jacobi_rtol1e-8"]
    pr = problems(2, None)
    res, x, A = solve(ctx, pr, fx.CG, maxIter=20000, relTol=1e-8, absTol=0.0)
    assert res.retcode == g["retcode"] == 0
    assert abs(res.numIter - g["numIter"]) <= 2, (res.numIter, g["numIter"])
    assert relres(pr, x) < 1e-8
    if res.numIter == g["numIter"]:  # the same iteration: the same residual and x to rounding
        assert abs(res.norm2 - g["norm2"]) <= 1e-6 * g["norm2"]
        assert abs(np.linalg.norm(x) - g["x_norm2"]) <= 1e-10 * g["x_norm2"]
        assert np.max(np.abs(samples(x) - g["x_samples"])) <= 1e-9 * np.max(np.abs(x))


def test_jacobi_sweeps_poisson3d_128(ctx, problems):
    """The 50 sweeps have no reduction in their update, so x is the
    oracle's bit for bit."""
    g = GOLD["jacobi_poisson3d_128_w2_3_50sweeps"]
    pr = problems(2, None)
    res, x, _ = solve(ctx, pr, fx.JACOBI, weight=2.0 / 3.0, maxIter=50, minIter=0, relTol=1e-30,
                      absTol=0.0)
    assert res.numIter == g["numIter"] == 50
    assert hashlib.sha256(np.ascontiguousarray(x, "<f8").tobytes()).hexdigest() == g["x_sha256"]
    assert abs(res.norm2 - g["norm2"]) <= 1e-12 * g["norm2"]


def test_refinement_poisson3d_128_against_oracle(ctx, problems):
    """FP32-inner / FP64-refined CG to 1e-8: final FP64 relres <= 1e-8, and
    the solution within 1e-6 of the oracle's FP64 Jacobi-PCG solution at
    the same tolerance (both are 1e-8 solutions: their error against x* is
    about 1.3e-6)."""
    g = GOLD["pcg_poisson3d_128_jacobi_rtol1e-8"]
    pr = problems(2, None)
    A = fx.Csr.from_problem(ctx, pr.A)
    x = fx.Vec(ctx, pr.A.nrow)
    res, inner = fx.solve_ir(ctx, A, dev(ctx, pr.b), x, 1.0,
                             fx.params(maxIter=50, relTol=1e-8, absTol=0.0),
                             fx.params(maxIter=20000, relTol=1e-4, absTol=0.0))
    xr = x.numpy()
    assert res.retcode == 0 and relres(pr, xr) <= 1e-8
    err = np.linalg.norm(xr - pr.xstar) / np.linalg.norm(pr.xstar)
    print(f"refined: {res.numIter} outer, {inner} inner; error vs x* {err:.3e} "
          f"(oracle FP64 PCG {g['err_vs_xstar']:.3e})")
    assert err <= 2.0 * g["err_vs_xstar"]
    scale = np.max(np.abs(xr))
    assert np.max(np.abs(samples(xr) - g["x_samples"])) <= 1e-6 * scale
    assert abs(np.linalg.norm(xr) - g["x_norm2"]) <= 1e-6 * g["x_norm2"]


def test_gmres30_convdiff_1024(ctx, problems):
    g = GOLD["gmres30_convdiff_1024_jacobi_rtol1e-8"]
    pr = problems(3, None)
    res, x, _ = solve(ctx, pr, fx.GMRES, maxIter=20000, relTol=1e-8, absTol=0.0, restart=30)
    assert res.retcode == g["retcode"] == 0
    assert abs(res.numIter - g["numIter"]) <= 2, (res.numIter, g["numIter"])
    assert relres(pr, x) < 1e-8
    err = np.linalg.norm(x - pr.xstar) / np.linalg.norm(pr.xstar)
    assert 0.5 * g["err_vs_xstar"] <= err <= 2.0 * g["err_vs_xstar"], (err, g["err_vs_xstar"])


def test_fgmres30_convdiff_1024_budget(ctx, problems):
    """The reference's FGMRES with its stale-Z Jacobi preconditioner
    (src/FGMRES.cxx:158-159) does not reach 1e-8 on this system; after the
    2000-iteration budget the device's relative residual is the oracle's to
    within 1% (measured: 2e-5 relative)."""
    g = GOLD["fgmres30_convdiff_1024_jacobi_maxit2000"]
    pr = problems(3, None)
    res, x, _ = solve(ctx, pr, fx.FGMRES, maxIter=2000, relTol=1e-8, absTol=0.0, restart=30)
    assert res.numIter == g["numIter"] == 2000
    rr = relres(pr, x)
    assert abs(rr - g["relres"]) <= 1e-2 * g["relres"], (rr, g["relres"])


def test_bicgstab_convdiff_1024_error(ctx, problems):
    """The count is held to the oracle's one-ulp band elsewhere; here the
    solution's error against x* is of the oracle's size."""
    g = GOLD["bicgstab_convdiff_1024_jacobi_rtol1e-8"]
    pr = problems(3, None)
    res, x, _ = solve(ctx, pr, fx.BICGSTAB, maxIter=20000, relTol=1e-8, absTol=0.0)
    assert res.retcode == 0 and relres(pr, x) < 1e-8
    err = np.linalg.norm(x - pr.xstar) / np.linalg.norm(pr.xstar)
    assert 0.2 * g["err_vs_xstar"] <= err <= 5.0 * g["err_vs_xstar"], (err, g["err_vs_xstar"])


def test_cg_and_refinement_powerlaw_1048576(ctx, oracle, problems):
    """FP64 Jacobi-CG to 1e-10 (count +-2, solution within 1e-9 of the
    oracle's, which runs here: 20 iterations) and the FP32-inner / FP64
    refinement to 1e-10 (solution within 1e-9 of the same)."""
    g = GOLD["cg_powerlaw_1048576_jacobi_rtol1e-10"]
    pr = problems(4, None)
    res, x, A = solve(ctx, pr, fx.CG, maxIter=20000, relTol=1e-10, absTol=0.0)
    assert res.retcode == 0 and abs(res.numIter - g["numIter"]) <= 2
    assert relres(pr, x) < 1e-10
    o = oracle.solve(CG, Matrix.of(pr.A), pr.b, np.zeros(pr.A.nrow), pc="jacobi", maxIter=20000,
                     relTol=1e-10, absTol=0.0)
    assert o["numIter"] == g["numIter"]
    assert np.allclose(samples(o["x"]), g["x_samples"], rtol=0, atol=0)
    assert rel_err(x, o["x"]) < 1e-9
    xi = fx.Vec(ctx, pr.A.nrow)
    r2, inner = fx.solve_ir(ctx, A, dev(ctx, pr.b), xi, 1.0,
                            fx.params(maxIter=50, relTol=1e-10, absTol=0.0),
                            fx.params(maxIter=20000, relTol=1e-4, absTol=0.0))
    xr = xi.numpy()
    assert r2.retcode == 0 and relres(pr, xr) < 1e-10
    e = rel_err(xr, o["x"])
    print(f"power-law refinement: {r2.numIter} outer, {inner} inner, |x - x_oracle|/|x| = {e:.2e}")
    assert e < 1e-9
