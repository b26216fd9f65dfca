"""The drop-in Parameters class (faspxx_amd/host/faspxx_param.cpp) through
tests/cpp/param_tests.cpp, CPU only: the reference's own unit test
(utest/src/UnitTestsParam.cxx:17-93, with its parameter files
utest/data/{single,multiple}_sol.param copied to tests/golden), the
parameter-file rules it leaves open, and SetSOLParams' option names."""
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
BIN = ROOT / "tests" / "cpp" / "param_tests"


def test_parameters_unit_test():
    if not BIN.exists():
        pytest.skip("param_tests not built (run __graft_entry__.build())")
    p = subprocess.run([str(BIN), str(ROOT / "tests" / "golden")], capture_output=True,
                       text=True, timeout=60)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "0 failures" in p.stdout


def test_help_exits_with_print_help():
    """-h prints the option list and exits with PRINT_HELP (0), as
    src/Param.cxx:96-100."""
    exe = ROOT / "faspxx_amd" / "examples" / "TestKrylov"
    if not exe.exists():
        pytest.skip("TestKrylov not built (run __graft_entry__.build())")
    p = subprocess.run([str(exe), "-h"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0
    for opt in ("-mat", "-rhs", "-xin", "-par", "-algName", "-maxIter", "-resRel", "-verbose"):
        assert opt in p.stdout
    p = subprocess.run([str(exe), "-mat"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0 and "Check the argument: -mat" in p.stdout
