"""Builds librasterizer_amd.so in-tree (rasterizer_amd/_lib/).

Host code (tables, bake, the C++ drop-in classes) is compiled for x86-64 with
g++; the kernels and the C ABI with hipcc for gfx950.  All floating-point code
is compiled with -ffp-contract=off so that products and sums round one at a
time, as the reference's SSE/AVX intrinsics do.  Packed FP32 instructions are
disabled for the device code.

Run as ``python -m rasterizer_amd.build`` or call :func:`build`.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIBDIR = PKG / "_lib"
LIB = LIBDIR / "librasterizer_amd.so"
ARCH = os.environ.get("RZ_OFFLOAD_ARCH", "gfx950")

HOST_FLAGS = ["-O2", "-g", "-fPIC", "-std=c++17", "-mavx2", "-mfma", "-msse4.1", "-ffp-contract=off",
              "-fno-fast-math", "-Wall", "-Wextra", "-Wno-unused-parameter", f"-I{ROOT / 'include'}"]
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off", "-fno-fast-math",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", f"-I{ROOT / 'include'}", f"-I{CSRC}"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the rasterizer needs ROCm's hipcc to build")


def _run_all(cmds: list[list[str]]) -> None:
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)) for c in cmds]
    failed = []
    for c, p in procs:
        out, _ = p.communicate()
        # hipcc's host pass does not know the packed-fp32 feature; that warning is expected.
        lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if p.returncode != 0:
            failed.append((c, "\n".join(lines)))
        elif lines and os.environ.get("RZ_BUILD_VERBOSE"):
            print("\n".join(lines), file=sys.stderr)
    if failed:
        msg = "\n\n".join(" ".join(c) + "\n" + o for c, o in failed)
        raise RuntimeError("rasterizer_amd build failed:\n" + msg)


def build(verbose: bool = False) -> Path:
    """Compiles all sources and links the shared library; returns its path."""
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    objs = {
        "rz_host.o": ["g++", *HOST_FLAGS, "-c", str(CSRC / "rz_host.cpp")],
        "rz_io.o": ["g++", *HOST_FLAGS, "-c", str(CSRC / "rz_io.cpp")],
        "rz_quad.o": ["g++", *HOST_FLAGS, "-O3", "-c", str(CSRC / "rz_quad.cpp")],
        "dropin.o": ["g++", *HOST_FLAGS, "-c", str(CSRC / "dropin.cpp")],
        "rz_kernels.o": [hipcc, *HIP_FLAGS, "-c", str(CSRC / "rz_kernels.hip")],
        "rz_api.o": [hipcc, *HIP_FLAGS, "-c", str(CSRC / "rz_api.cpp")],
        "rz_views.o": [hipcc, *HIP_FLAGS, "-c", str(CSRC / "rz_views.cpp")],
        "rz_learn.o": [hipcc, *HIP_FLAGS, "-c", str(CSRC / "rz_learn.cpp")],
        "rz_mlp.o": [hipcc, *HIP_FLAGS, "-c", str(CSRC / "rz_mlp.hip")],
        "rz_bake.o": [hipcc, *HIP_FLAGS, "-c", str(CSRC / "rz_bake.hip")],
        "rz_sah.o": [hipcc, *HIP_FLAGS, "-c", str(CSRC / "rz_sah.hip")],
        "rz_quad_dev.o": [hipcc, *HIP_FLAGS, "-c", str(CSRC / "rz_quad.hip")],
        "rz_tune.o": [hipcc, *HIP_FLAGS, "-c", str(CSRC / "rz_tune.cpp")],
    }
    cmds = [cmd + ["-o", str(BUILD / name)] for name, cmd in objs.items()]
    if verbose:
        for c in cmds:
            print(" ".join(c))
    _run_all(cmds)
    tmp = LIB.with_suffix(".so.tmp")
    link = [hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp),
            *[str(BUILD / n) for n in objs], "-lpthread"]
    _run_all([link])
    os.replace(tmp, LIB)
    # Test programs of the C++ drop-in headers (run by tests/test_gpu_sah.py
    # and tests/test_gpu_dropin.py).
    bindir = ROOT / "tests" / "_bin"
    progs = []
    for name in ("sah_dropin_check", "main_sequence_check"):
        src = ROOT / "tests" / "cpp" / f"{name}.cpp"
        if src.exists():
            bindir.mkdir(exist_ok=True)
            progs.append(["g++", "-O2", "-std=c++17", "-msse4.1", f"-I{ROOT / 'include' / 'rasterizer_amd'}",
                          f"-I{ROOT / 'include'}", str(src), f"-L{LIBDIR}", "-lrasterizer_amd",
                          "-Wl,-rpath,$ORIGIN/../../rasterizer_amd/_lib", "-Wl,-rpath,/opt/rocm/lib",
                          "-o", str(bindir / name)])
    if progs:
        _run_all(progs)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
