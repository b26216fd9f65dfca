"""The FASP++-interface C++ classes (faspxx_amd/host) on the GPU: the C++
port of the reference's unit tests (utest/src/UnitTests{VEC,MAT,Jacobi}.cxx)
and the TestPCG / TestKrylov call sequences, and ReadMat on the reference's
data files (tests/cpp/dropin_tests.cpp)."""
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
BIN = ROOT / "tests" / "cpp" / "dropin_tests"

pytestmark = pytest.mark.gpu


def test_dropin_unit_tests():
    if not BIN.exists():
        pytest.fail(f"{BIN} not built: run __graft_entry__.build()")
    p = subprocess.run([str(BIN), str(ROOT / "tests" / "golden")], capture_output=True, text=True,
                       timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "PASSED (0 failures)" in p.stdout
