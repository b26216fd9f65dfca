"""File ingestion of the drop-in layer (faspxx_amd/host/faspxx_readdata.cpp:
ReadCSR / ReadMTX / ReadMat with SortCSRRow and CSRtoMAT's zero-diagonal
insertion) on the reference's data files, CPU only.  The expected CSRx arrays
come from the compiled reference's own ReadMat / ReadCSR (src/ReadData.cxx,
oracle/_ref) where it is built, and from independent normalisations:
tests/csrfile.py for .csr, and for .mtx a NumPy one written from the format
(rows bucketed in file order, a stable sort by column, zero diagonals
inserted).  Structure and values must match exactly; the oracle's CheckCSRx
must accept the result."""
import pathlib
import subprocess

import numpy as np
import pytest

from csrfile import read_csr
from oracle import Matrix

ROOT = pathlib.Path(__file__).resolve().parents[1]
BIN = ROOT / "tests" / "cpp" / "readdata_tests"
GOLD = ROOT / "tests" / "golden"


def run(mode, name):
    if not BIN.exists():
        pytest.skip("readdata_tests not built (run __graft_entry__.build())")
    out = subprocess.run([str(BIN), mode, str(GOLD / name)], capture_output=True, text=True,
                         check=True).stdout.split()
    rc, row, col, nnz = (int(v) for v in out[:4])
    rest = out[4:]
    nrp = row + 1
    rp = np.array(rest[:nrp], dtype=np.int64)
    ci = np.array(rest[nrp:nrp + nnz], dtype=np.int64)
    va = np.array(rest[nrp + nnz:nrp + 2 * nnz], dtype=np.float64)
    return rc, row, col, rp, ci, va


def mtx_expected(name):
    lines = [l for l in (GOLD / name).read_text().splitlines() if l.strip() and l[0] != "%"]
    n, m, nnz = (int(v) for v in lines[0].split())
    trip = [l.split() for l in lines[1:1 + nnz]]
    r = np.array([int(t[0]) - 1 for t in trip])
    c = np.array([int(t[1]) - 1 for t in trip])
    v = np.array([float(t[2]) for t in trip])
    rows_c, rows_v = [], []
    for i in range(n):
        sel = np.nonzero(r == i)[0]  # file order
        ci, vi = c[sel], v[sel]
        order = np.argsort(ci, kind="stable")
        ci, vi = ci[order], vi[order]
        if i < m and not np.any(ci == i):
            k = np.searchsorted(ci, i)
            ci, vi = np.insert(ci, k, i), np.insert(vi, k, 0.0)
        rows_c.append(ci)
        rows_v.append(vi)
    rp = np.zeros(n + 1, np.int64)
    np.cumsum([len(x) for x in rows_c], out=rp[1:])
    return n, m, rp, np.concatenate(rows_c), np.concatenate(rows_v)


@pytest.mark.parametrize("name", ["fdm_10X10.csr", "fem_small.csr"])
def test_csr_files(oracle, name):
    rc, row, col, rp, ci, va = run("csrx", name)
    n, erp, eci, eva = read_csr(name)
    assert rc == 0 and row == col == n
    assert np.array_equal(rp, erp) and np.array_equal(ci, eci) and np.array_equal(va, eva)
    dp = oracle.form_diagptr(n, rp, ci)
    assert oracle.check_csrx(Matrix(n, n, rp, ci, va, dp)) == 0


@pytest.mark.parametrize("name", ["nos7.mtx", "nos7-.mtx"])
def test_mtx_files(oracle, name):
    rc, row, col, rp, ci, va = run("csrx", name)
    n, m, erp, eci, eva = mtx_expected(name)
    assert rc == 0 and (row, col) == (n, m)
    assert np.array_equal(rp, erp) and np.array_equal(ci, eci) and np.array_equal(va, eva)
    dp = oracle.form_diagptr(n, rp, ci)
    assert oracle.check_csrx(Matrix(n, m, rp, ci, va, dp)) == 0


def test_raw_csr_is_the_file():
    rc, row, col, rp, ci, va = run("raw", "fdm_10X10.csr")
    tok = (GOLD / "fdm_10X10.csr").read_text().split()
    n = int(tok[0])
    frp = np.array(tok[1:n + 2], dtype=np.int64)
    shift = 1 if frp[0] == 1 else 0
    assert rc == 0 and row == n
    assert np.array_equal(rp, frp - shift)
    nnz = int(frp[-1] - frp[0])
    assert np.array_equal(ci, np.array(tok[n + 2:n + 2 + nnz], dtype=np.int64) - shift)
    assert np.array_equal(va, np.array(tok[n + 2 + nnz:n + 2 + 2 * nnz], dtype=np.float64))


@pytest.mark.parametrize("name", ["fdm_10X10.csr", "fem_small.csr", "nos7.mtx", "nos7-.mtx"])
def test_readmat_against_compiled_reference(reference, name):
    """The normalised CSRx equals the reference's ReadMat bit for bit."""
    rc, row, col, rp, ci, va = run("csrx", name)
    ref = reference.read_mat(GOLD / name)
    assert not isinstance(ref, int), ref
    assert rc == 0 and (row, col) == ref[:2]
    assert np.array_equal(rp, ref[2]) and np.array_equal(ci, ref[3])
    assert np.array_equal(va, ref[4])


@pytest.mark.parametrize("name", ["fdm_10X10.csr", "fem_small.csr"])
def test_readcsr_raw_against_compiled_reference(reference, name):
    rc, row, col, rp, ci, va = run("raw", name)
    ref = reference.read_csr(GOLD / name)
    assert not isinstance(ref, int), ref
    assert rc == 0 and (row, col) == ref[:2]
    assert np.array_equal(rp, ref[2]) and np.array_equal(ci, ref[3])
    assert np.array_equal(va, ref[4])


def test_missing_file_code_matches_reference(reference):
    rc = subprocess.run([str(BIN), "csrx", str(GOLD / "missing.csr")], capture_output=True,
                        text=True).stdout.split()
    assert int(rc[0]) == reference.read_mat(GOLD / "missing.csr") == -10  # ERROR_OPEN_FILE
