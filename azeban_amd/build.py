"""Build the HIP library libazeban_hip.so in-tree for gfx950.

Two hipcc translation units: azeban_amd/csrc/azb_solver.hip (the C ABI declared
in include/azeban_hip.h, the column passes and the seams) and
azeban_amd/csrc/azb_fused.hip (the fused row kernels and the k-space kernel).
The second is compiled with the instruction-level-parallelism scheduler, which
measured faster for those kernels and slower for the column passes.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "libazeban_hip.so"
CSRC = PKG / "csrc"
# (source, extra device-compiler flags)
UNITS = [
    (CSRC / "azb_solver.hip", []),
    (CSRC / "azb_fused.hip", ["-mllvm", "--amdgpu-sched-strategy=max-ilp"]),
]
SOURCES = [u[0] for u in UNITS]
HEADERS = [CSRC / "azb_fft.hpp", CSRC / "azb_kernels.hpp", CSRC / "azb_aux_kernels.hpp", CSRC / "azb_launch.hpp",
           ROOT / "include" / "azeban_hip.h"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # multiply-adds are fused only within one source expression: the backend's
    # cross-statement fusion depends on which outputs a kernel instantiation
    # discards (e.g. the truncated rows of a forward column pass for a given tile
    # width), which would make results differ in the last bit between launch shapes
    "-ffp-contract=on",
    # no packed-FP32 arithmetic in the generated gfx950 code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list, verbose: bool) -> subprocess.Popen:
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)


def build(verbose: bool = False) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    objs = [PKG / "csrc" / (src.stem + ".o") for src, _ in UNITS]
    # the translation units compile in parallel, then link
    procs = [_run([hipcc(), *HIPCC_FLAGS, *extra, "-c", str(src), "-o", str(obj)], verbose)
             for (src, extra), obj in zip(UNITS, objs)]
    failed = False
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = True
    if failed:
        raise RuntimeError("hipcc failed compiling libazeban_hip.so")
    link = _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)], verbose)
    out, _ = link.communicate()
    if link.returncode != 0:
        sys.stderr.write(out)
        raise RuntimeError("hipcc failed linking libazeban_hip.so")
    for obj in objs:
        obj.unlink()
    os.replace(tmp, LIB)
    return LIB


CPP_TEST_SRC = ROOT / "tests" / "cpp" / "test_dropin.cpp"
CPP_TEST_BIN = ROOT / "tests" / "cpp" / "test_dropin"


def build_cpp_tests(verbose: bool = False) -> Path:
    """The C++ drop-in layer test program, host-only C++ linked against the library."""
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{ROOT / 'include'}", str(CPP_TEST_SRC),
           f"-L{PKG}", "-lazeban_hip", "-Wl,-rpath,$ORIGIN/../../azeban_amd", "-o", str(CPP_TEST_BIN)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building the C++ drop-in tests")
    return CPP_TEST_BIN


def is_stale() -> bool:
    if not LIB.exists():
        return True
    t = LIB.stat().st_mtime
    return any(p.stat().st_mtime > t for p in SOURCES + HEADERS)


if __name__ == "__main__":
    print(build(verbose=True))
