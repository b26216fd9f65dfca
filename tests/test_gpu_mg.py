"""The drop-in MG<MAT> (faspxx_amd/host/MG.hxx) on the GPU, driven as
example/TestMG.cxx drives the reference's (tests/cpp/mg_tests.cpp), against
the oracle's restatement (orc_mg, itself pinned by tests/test_oracle_mg.py):
the same CSRx arrays (ReadMat's normalisation, read back through
tests/cpp/readdata_tests), the same levels, smoothers and parameters.  Bars:
the iteration count within +-2 (equal in practice) and the solution within
1e-12 of the oracle's relative to the initial guess, 1e-10 relative
otherwise."""
import pathlib
import subprocess

import numpy as np
import pytest

from oracle import Matrix

ROOT = pathlib.Path(__file__).resolve().parents[1]
EXE = ROOT / "tests" / "cpp" / "mg_tests"
RD = ROOT / "tests" / "cpp" / "readdata_tests"
GOLD = ROOT / "tests" / "golden"

pytestmark = pytest.mark.gpu

ONE = dict(maxIter=1, relTol=0.0, absTol=0.0)


def csrx(path, oracle):
    out = subprocess.run([str(RD), "csrx", str(path)], capture_output=True, text=True,
                         check=True, timeout=60).stdout.split()
    rc, row, col, nnz = (int(v) for v in out[:4])
    assert rc == 0
    rest = out[4:]
    rp = np.array(rest[:row + 1], dtype=np.uint32)
    ci = np.array(rest[row + 1:row + 1 + nnz], dtype=np.uint32)
    va = np.array(rest[row + 1 + nnz:row + 1 + 2 * nnz], dtype=np.float64)
    return Matrix(row, col, rp, ci, va, oracle.form_diagptr(row, rp, ci))


def run_mg(A_file, levels, ncycles, smoother, maxIter, minIter, relTol, absTol, extra=()):
    if not EXE.exists():
        pytest.fail(f"{EXE} not built: run __graft_entry__.build()")
    args = [str(EXE), str(A_file), str(levels), str(ncycles), smoother, str(maxIter),
            str(minIter), repr(relTol), repr(absTol), *map(str, extra)]
    print(" ".join(args), flush=True)
    p = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr
    lines = p.stdout.split("\n")
    i = next(k for k, l in enumerate(lines) if len(l.split()) == 4)
    rc, it, n2, ninf = lines[i].split()
    x = np.array([float(v) for v in lines[i + 1:] if v.strip()])
    return int(rc), int(it), float(n2), x


def write_vec(path, v):
    path.write_text(f"{len(v)}\n" + "\n".join(repr(float(a)) for a in v) + "\n")


def test_testmg_sample(oracle):
    """TestMG's sample run: fdm_10X10, three IdentityMatrix levels, default
    Jacobi smoothers, -maxIter 200 -minIter 100, b = 0, x0 = 1."""
    A = csrx(GOLD / "fdm_10X10.csr", oracle)
    rc, it, n2, x = run_mg(GOLD / "fdm_10X10.csr", 3, 1, "default", 200, 100, 1e-6, 1e-8)
    o = oracle.mg(A, np.zeros(A.nrow), np.ones(A.nrow), [(None, None, None)] * 3, maxIter=200,
                  minIter=100)
    assert rc == o["retcode"] == 0
    assert it == o["numIter"] == 100
    assert np.max(np.abs(x - o["x"])) <= 1e-12
    assert abs(n2 - o["norm2"]) <= 1e-6 * o["norm2"]


@pytest.mark.parametrize("levels,ncycles", [(3, 1), (3, 2), (4, 1)])
def test_one_sweep_solver(oracle, tmp_path, levels, ncycles):
    """One-sweep smoothers and coarse solvers, a random right-hand side, MG
    iterated to relTol 1e-10 (fdm_10X10: a cycle is 2L + 1 Jacobi sweeps
    here, too few for fem_small's 3969 unknowns within 400 cycles)."""
    A = csrx(GOLD / "fdm_10X10.csr", oracle)
    rng = np.random.default_rng(levels * 10 + ncycles)
    b = rng.standard_normal(A.nrow)
    write_vec(tmp_path / "b.vec", b)
    write_vec(tmp_path / "x0.vec", np.zeros(A.nrow))
    rc, it, n2, x = run_mg(GOLD / "fdm_10X10.csr", levels, ncycles, "one", 400, 0, 1e-10, 0.0,
                           ["-", tmp_path / "b.vec", tmp_path / "x0.vec"])
    o = oracle.mg(A, b, np.zeros(A.nrow), [(None, ONE, ONE)] * levels, ncycles=ncycles,
                  maxIter=400, relTol=1e-10, absTol=0.0)
    assert rc == o["retcode"] == 0
    assert abs(it - o["numIter"]) <= 2, (it, o["numIter"])
    assert 0 < it < 400
    assert np.linalg.norm(x - o["x"]) <= 1e-10 * np.linalg.norm(o["x"])


def test_square_transfer_operator(oracle, tmp_path):
    """A random nonsymmetric square transfer operator read from a .mtx file,
    two levels, three cycles of MG from a random x0."""
    A = csrx(GOLD / "fdm_10X10.csr", oracle)
    n = A.nrow
    rng = np.random.default_rng(11)
    Td = np.eye(n) * 0.7 + (rng.random((n, n)) < 0.03) * rng.uniform(-0.2, 0.2, (n, n))
    r, c = np.nonzero(Td)
    body = "\n".join(f"{i + 1} {j + 1} {float(Td[i, j])!r}" for i, j in zip(r, c))
    (tmp_path / "T.mtx").write_text(
        f"%%MatrixMarket matrix coordinate real general\n{n} {n} {len(r)}\n{body}\n")
    T = csrx(tmp_path / "T.mtx", oracle)
    b = rng.standard_normal(n)
    x0 = rng.standard_normal(n)
    write_vec(tmp_path / "b.vec", b)
    write_vec(tmp_path / "x0.vec", x0)
    rc, it, n2, x = run_mg(GOLD / "fdm_10X10.csr", 2, 1, "one", 3, 0, 0.0, 0.0,
                           [tmp_path / "T.mtx", tmp_path / "b.vec", tmp_path / "x0.vec"])
    o = oracle.mg(A, b, x0, [(T, ONE, ONE)] * 2, maxIter=3, relTol=0.0, absTol=0.0)
    assert it == o["numIter"] == 3
    assert np.linalg.norm(x - o["x"]) <= 1e-12 * np.linalg.norm(o["x"])
