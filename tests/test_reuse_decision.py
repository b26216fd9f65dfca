"""The solver's workspace-reuse decision (faspxx_amd/csrc/fx_reuse.hpp) on
the CPU: a handle freed and made again at the same address must not be
taken for the old one.  Compiled with the host compiler, no GPU needed."""
import pathlib
import subprocess

ROOT = pathlib.Path(__file__).resolve().parents[1]


def test_reuse_decision(tmp_path):
    exe = tmp_path / "reuse_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", f"-I{ROOT / 'faspxx_amd' / 'csrc'}",
                    str(ROOT / "tests" / "cpp" / "reuse_check.cpp"), "-o", str(exe)], check=True)
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=30)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "PASSED (0 failures)" in p.stdout
