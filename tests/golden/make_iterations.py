"""Writes tests/golden/solver_iterations.json: results of the oracle (the CPU
restatement of the reference's solvers, itself pinned against the compiled
reference) on the full-size systems, too large for the oracle to run inside
the GPU test suite.  CPU only.

    python tests/golden/make_iterations.py [entry ...]   # default: all
    entries: bicgstab pcg jacobi gmres fgmres powerlaw   (about 12 minutes in all)

Entries (each merged into the existing file):
* bicgstab: BiCGStab + Jacobi on the 1024^2 convection-diffusion system to
  rtol 1e-8, its error against x*, and the counts for the same system with
  b perturbed by one unit in the last place in a random eighth of its
  entries (seeds 1..4).  BiCGStab's count on this nonsymmetric system is not
  determined by the algorithm to better than that spread: any change of
  summation order perturbs the iteration by the same order of rounding.
* pcg: Jacobi-PCG on the 3D Poisson 128^3 system to rtol 1e-8
  (example/TestPCG.cxx:71-88): count, final norms, ||x||, x at 16 indices.
* jacobi: 50 weighted-Jacobi sweeps (omega = 2/3) on the same system from
  x = 0 (example/TestJacobi.cxx:69-81): SHA-256 of x's bytes (the sweeps
  are bit-exact on the device), final norm, x at 16 indices.
* gmres: GMRES(30) + Jacobi (right) on the 1024^2 system to rtol 1e-8:
  count, norms, the cycle lengths of the variable restart.
* fgmres: FGMRES(30) + Jacobi on the 1024^2 system with a budget of 2000
  iterations (the stale-Z preconditioner, src/FGMRES.cxx:158-159, does not
  reach 1e-8 within it): final norm and true residual, cycle lengths.
* powerlaw: Jacobi-CG on the power-law system (n = 2^20) to rtol 1e-10:
  count, norms, ||x||, x at 16 indices.
"""
import hashlib
import json
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))

from faspxx_amd import problems as P  # noqa: E402
from oracle import BICGSTAB, CG, FGMRES, GMRES, JACOBI, Matrix, Oracle  # noqa: E402
from perturb import ulp_perturbed  # noqa: E402

PATH = pathlib.Path(__file__).with_name("solver_iterations.json")


def sample_idx(n):
    return np.linspace(0, n - 1, 16).astype(np.int64)


def relres(pr, x):
    return float(np.linalg.norm(pr.b - pr.A.matvec(x)) / np.linalg.norm(pr.b))


def common(o, pr, cpu_s):
    x = o["x"]
    out = {"numIter": o["numIter"], "retcode": o["retcode"], "norm2": o["norm2"],
           "normInf": o["normInf"], "relres": relres(pr, x), "x_norm2": float(np.linalg.norm(x)),
           "x_samples": x[sample_idx(x.shape[0])].tolist(), "cpu_s": cpu_s}
    if pr.xstar is not None:
        out["err_vs_xstar"] = float(np.linalg.norm(x - pr.xstar) / np.linalg.norm(pr.xstar))
    return out


def run(o, kind, pr, **kw):
    M = Matrix.of(pr.A)
    t = time.time()
    r = o.solve(kind, M, pr.b, np.zeros(M.nrow), **kw)
    return r, round(time.time() - t, 1)


def traced(o, kind, pr, **kw):
    M = Matrix.of(pr.A)
    t = time.time()
    r = o.krylov_traced(kind, M, pr.b, np.zeros(M.nrow), **kw)
    return r, round(time.time() - t, 1)


def entry_bicgstab(o):
    pr = P.config(3)
    prm = dict(pc="jacobi", maxIter=20000, relTol=1e-8, absTol=0.0)
    r, cpu_s = run(o, BICGSTAB, pr, **prm)
    out = common(r, pr, cpu_s)
    M = Matrix.of(pr.A)
    spread = []
    for seed in (1, 2, 3, 4):
        rs = o.solve(BICGSTAB, M, ulp_perturbed(pr.b, seed), np.zeros(M.nrow), **prm)
        assert rs["retcode"] == 0
        spread.append(rs["numIter"])
        print("seed", seed, rs["numIter"], flush=True)
    out["numIter_b_ulp_perturbed"] = spread
    return "bicgstab_convdiff_1024_jacobi_rtol1e-8", out


def entry_pcg(o):
    pr = P.config(2)
    r, cpu_s = run(o, CG, pr, pc="jacobi", weight=1.0, maxIter=20000, relTol=1e-8, absTol=0.0)
    return "pcg_poisson3d_128_jacobi_rtol1e-8", common(r, pr, cpu_s)


def entry_jacobi(o):
    pr = P.config(2)
    r, cpu_s = run(o, JACOBI, pr, weight=2.0 / 3.0, maxIter=50, minIter=0, relTol=1e-30,
                   absTol=0.0)
    out = common(r, pr, cpu_s)
    out["x_sha256"] = hashlib.sha256(np.ascontiguousarray(r["x"], "<f8").tobytes()).hexdigest()
    return "jacobi_poisson3d_128_w2_3_50sweeps", out


def entry_gmres(o):
    pr = P.config(3)
    r, cpu_s = traced(o, GMRES, pr, pc="jacobi", maxIter=20000, relTol=1e-8, absTol=0.0,
                      restart=30)
    out = common(r, pr, cpu_s)
    out["cycle_lengths"] = [c[0] for c in r["cycles"]]
    return "gmres30_convdiff_1024_jacobi_rtol1e-8", out


def entry_fgmres(o):
    pr = P.config(3)
    r, cpu_s = traced(o, FGMRES, pr, pc="jacobi", maxIter=2000, relTol=1e-8, absTol=0.0,
                      restart=30)
    out = common(r, pr, cpu_s)
    out["cycle_lengths"] = [c[0] for c in r["cycles"]]
    out["cycle_rtrue"] = [c[2] for c in r["cycles"]]
    return "fgmres30_convdiff_1024_jacobi_maxit2000", out


def entry_powerlaw(o):
    pr = P.config(4)
    r, cpu_s = run(o, CG, pr, pc="jacobi", weight=1.0, maxIter=20000, relTol=1e-10, absTol=0.0)
    return "cg_powerlaw_1048576_jacobi_rtol1e-10", common(r, pr, cpu_s)


ENTRIES = {"bicgstab": entry_bicgstab, "pcg": entry_pcg, "jacobi": entry_jacobi,
           "gmres": entry_gmres, "fgmres": entry_fgmres, "powerlaw": entry_powerlaw}


def main(names):
    o = Oracle()
    data = json.loads(PATH.read_text()) if PATH.exists() else {}
    for name in names or list(ENTRIES):
        t = time.time()
        key, out = ENTRIES[name](o)
        data[key] = out
        print(f"{key}: numIter {out['numIter']}, relres {out['relres']:.3e} "
              f"({time.time() - t:.0f} s)", flush=True)
        PATH.write_text(json.dumps(data, indent=1) + "\n")


if __name__ == "__main__":
    main(sys.argv[1:])
