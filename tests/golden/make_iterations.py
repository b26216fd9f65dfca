"""Writes tests/golden/solver_iterations.json: iteration counts of the oracle
(the CPU restatement of the reference's solvers) on systems too large for the
oracle to run inside the test suite.  CPU only; about 8 minutes.

Besides the count for the system itself, it records the counts for the same
system with b perturbed by one unit in the last place in a random eighth of
its entries (seeds 1..4).  BiCGStab's iteration count on this nonsymmetric
system is not determined by the algorithm to better than that spread: any
change of summation order (a parallel dot product) perturbs the iteration
by the same order of rounding.

    python tests/golden/make_iterations.py
"""
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
from oracle import BICGSTAB, Matrix, Oracle  # noqa: E402
from perturb import ulp_perturbed  # noqa: E402


def main():
    o = Oracle()
    out = {}
    pr = P.config(3)  # 2D upwind convection-diffusion 1024 x 1024
    M = Matrix.of(pr.A)
    prm = dict(pc="jacobi", maxIter=20000, relTol=1e-8, absTol=0.0)
    t = time.time()
    r = o.solve(BICGSTAB, M, pr.b, np.zeros(M.nrow), **prm)
    cpu_s = round(time.time() - t, 1)
    spread = []
    for seed in (1, 2, 3, 4):
        rs = o.solve(BICGSTAB, M, ulp_perturbed(pr.b, seed), np.zeros(M.nrow), **prm)
        assert rs["retcode"] == 0
        spread.append(rs["numIter"])
        print("seed", seed, rs["numIter"], flush=True)
    out["bicgstab_convdiff_1024_jacobi_rtol1e-8"] = {
        "numIter": r["numIter"], "retcode": r["retcode"],
        "relres": float(np.linalg.norm(pr.b - pr.A.matvec(r["x"])) / np.linalg.norm(pr.b)),
        "cpu_s": cpu_s, "numIter_b_ulp_perturbed": spread}
    path = pathlib.Path(__file__).with_name("solver_iterations.json")
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(out)


if __name__ == "__main__":
    main()
