"""The reference's Krylov driver (example/TestKrylov.cxx) built against the
drop-in classes (faspxx_amd/examples/TestKrylov.cpp) and run on the GPU from
its command line: b = 0, x0 = 1, identity preconditioner, each algorithm by
name, on the reference's own data files.  The iteration count must be within
+-2 of the oracle's and the true residual of the result must meet the
tolerance."""
import pathlib
import re
import subprocess

import numpy as np
import pytest

from csrfile import read_csr
from oracle import BICGSTAB, CG, FGMRES, GMRES, Matrix

ROOT = pathlib.Path(__file__).resolve().parents[1]
EXE = ROOT / "faspxx_amd" / "examples" / "TestKrylov"
GOLD = ROOT / "tests" / "golden"

pytestmark = pytest.mark.gpu

KINDS = {"cg": CG, "bicgstab": BICGSTAB, "gmres": GMRES, "fgmres": FGMRES}


def run(*args):
    if not EXE.exists():
        pytest.fail(f"{EXE} not built: run __graft_entry__.build()")
    p = subprocess.run([str(EXE), *args], capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert p.returncode == 0, p.stdout + p.stderr
    it = int(re.search(r"Number of iterations : (\d+)", p.stdout).group(1))
    m = re.search(r"result (-?\d+) (\S+) (\S+)", p.stdout)
    return int(m.group(1)), it, float(m.group(2))


@pytest.mark.parametrize("alg", ["cg", "bicgstab", "gmres", "fgmres", "CG"])
def test_testkrylov_fdm(oracle, alg):
    n, rp, ci, va = read_csr("fdm_10X10.csr")
    M = Matrix(n, n, rp, ci, va, oracle.form_diagptr(n, rp, ci))
    rc, it, rn = run("-mat", str(GOLD / "fdm_10X10.csr"), "-algName", alg, "-maxIter", "500",
                     "-resRel", "1e-10", "-resAbs", "0", "-verbose", "2")
    o = oracle.solve(KINDS[alg.lower()], M, np.zeros(n), np.ones(n), maxIter=500, relTol=1e-10,
                     absTol=0.0)
    assert rc == o["retcode"] == 0
    assert abs(it - o["numIter"]) <= 2, (it, o["numIter"])
    r0 = np.linalg.norm(M_apply(M, np.ones(n)))
    assert rn <= 1e-10 * r0 * 1.5


def M_apply(M, x):
    y = np.zeros(M.nrow)
    for i in range(M.nrow):
        s, e = M.rp[i], M.rp[i + 1]
        y[i] = np.dot(M.va[s:e], x[M.ci[s:e]])
    return y


def csrx_of(name):
    """The CSRx arrays ReadMat builds from a data file (tests/cpp/readdata_tests)."""
    exe = ROOT / "tests" / "cpp" / "readdata_tests"
    out = subprocess.run([str(exe), "csrx", str(GOLD / name)], capture_output=True, text=True,
                         check=True, timeout=60).stdout.split()
    rc, row, col, nnz = (int(v) for v in out[:4])
    assert rc == 0
    rest = out[4:]
    rp = np.array(rest[:row + 1], dtype=np.uint32)
    ci = np.array(rest[row + 1:row + 1 + nnz], dtype=np.uint32)
    va = np.array(rest[row + 1 + nnz:row + 1 + 2 * nnz], dtype=np.float64)
    return row, rp, ci, va


@pytest.mark.parametrize("alg", ["gmres"])
def test_testkrylov_mtx_defaults(oracle, alg):
    """nos7.mtx (n = 729, SPD, condition number about 4e9) through ReadMat's
    .mtx path with the SOLParams defaults except maxIter (relTol 1e-6,
    absTol 1e-8).  (Unpreconditioned CG does not converge on it within 3000
    iterations in the oracle either, so it makes no count comparison.)"""
    n, rp, ci, va = csrx_of("nos7.mtx")
    M = Matrix(n, n, rp, ci, va, oracle.form_diagptr(n, rp, ci))
    rc, it, rn = run("-mat", str(GOLD / "nos7.mtx"), "-algName", alg, "-maxIter", "3000",
                     "-verbose", "2")
    o = oracle.solve(KINDS[alg], M, np.zeros(n), np.ones(n), maxIter=3000)
    assert rc == o["retcode"], (rc, o["retcode"])
    assert abs(it - o["numIter"]) <= 2, (it, o["numIter"])
