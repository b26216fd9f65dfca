"""Build libfaspxx_hip.so (gfx950) in-tree with hipcc.

    python -m faspxx_amd.build            # compile every HIP source, link

Objects go to faspxx_amd/_build/, the library to faspxx_amd/libfaspxx_hip.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "libfaspxx_hip.so"
BUILD = PKG / "_build"

ARCH = os.environ.get("FASPXX_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: every fused multiply-add is written explicitly with fma().
# Packed-FP32 instructions are disabled for all device code.
CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def sources() -> list[pathlib.Path]:
    return sorted(CSRC.glob("*.hip"))


def _compile(src: pathlib.Path) -> tuple[pathlib.Path, str]:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", str(src), "-o", str(obj)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{p.stdout}\n{p.stderr}")
    return obj, p.stderr


def build(verbose: bool = False) -> pathlib.Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        results = list(ex.map(_compile, srcs))
    if verbose:
        for obj, err in results:
            msg = "\n".join(l for l in err.splitlines() if "packed-fp32-ops" not in l)
            if msg.strip():
                print(f"[{obj.name}] {msg}", file=sys.stderr)
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *[str(o) for o, _ in results]]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stdout}\n{p.stderr}")
    os.replace(tmp, OUT)
    build_dropin()
    return OUT


DROPIN = PKG / "libfaspxx_dropin.so"


EXAMPLES = ["TestKrylov"]


def build_dropin() -> None:
    """The FASP++-interface C++ classes (host code over the C ABI), the C++
    ports of the reference's unit tests that exercise them, and the example
    drivers (faspxx_amd/examples)."""
    cxx = os.environ.get("CXX", "g++")
    common = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{ROOT / 'include'}",
              f"-I{PKG / 'host'}"]
    host = PKG / "host"
    cmd = [cxx, *common, "-shared", "-o", str(DROPIN), str(host / "faspxx_dropin.cpp"),
           str(host / "faspxx_readdata.cpp"), str(host / "faspxx_param.cpp"), f"-L{PKG}",
           "-lfaspxx_hip", "-Wl,-rpath,$ORIGIN"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"drop-in library build failed:\n{p.stderr}")
    progs = [(ROOT / "tests" / "cpp" / f"{t}.cpp", ROOT / "tests" / "cpp" / t, "$ORIGIN/../../faspxx_amd")
             for t in ("dropin_tests", "readdata_tests", "param_tests", "mg_tests")]
    progs += [(PKG / "examples" / f"{t}.cpp", PKG / "examples" / t, "$ORIGIN/..") for t in EXAMPLES]
    for src, out, rpath in progs:
        cmd = [cxx, *common, "-o", str(out), str(src), f"-L{PKG}", "-lfaspxx_dropin",
               "-lfaspxx_hip", f"-Wl,-rpath,{PKG}", f"-Wl,-rpath,{rpath}"]
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError(f"{src.name} build failed:\n{p.stderr}")


if __name__ == "__main__":
    print(build(verbose=True))
