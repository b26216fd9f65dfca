"""The C-ABI library loads and exports every entry point include/faspxx_hip.h
declares; the Python binding declares a signature for each.  No GPU calls."""
import pathlib
import re

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared():
    text = (ROOT / "include" / "faspxx_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert len(names) > 40
    assert "fx_spmv" in names and "fx_solver_solve" in names and "fx_solve_ir" in names


def test_library_exports_all_declared():
    import ctypes
    from faspxx_amd import fx
    if not fx.LIB_PATH.exists():
        pytest.fail(f"{fx.LIB_PATH} missing: run __graft_entry__.build()")
    lib = ctypes.CDLL(str(fx.LIB_PATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from faspxx_amd import fx
    assert set(declared()) <= set(fx._SIGS), set(declared()) - set(fx._SIGS)
    fx.load()  # dlopen + argtypes; no device call


def test_params_default_matches_solparams():
    from faspxx_amd import fx
    p = fx.params()
    # SOLParams(), src/Param.hxx:68-79
    assert (p.maxIter, p.minIter, p.savIter, p.restart) == (1, 0, 100000, 30)
    assert (p.relTol, p.absTol) == (1e-6, 1e-8)
    assert (p.maxRestart, p.minRestart, p.leftPrecond) == (30, 10, 0)


def test_dropin_library_builds_and_links():
    """The C++ drop-in layer is built and links against the C ABI library."""
    import ctypes
    lib = ROOT / "faspxx_amd" / "libfaspxx_dropin.so"
    if not lib.exists():
        pytest.fail(f"{lib} missing: run __graft_entry__.build()")
    ctypes.CDLL(str(ROOT / "faspxx_amd" / "libfaspxx_hip.so"))
    ctypes.CDLL(str(lib))
    assert (ROOT / "tests" / "cpp" / "dropin_tests").exists()
