"""Builds libgi.so (HIP kernels for gfx950 + the C ABI of include/gi.h) in-tree.

The shared library lands next to this file so that it travels with the source
tree; objects go to ``_build/``.  Every HIP translation unit is compiled with
``-ffp-contract=off`` (each ``a*b+c`` in the kernels is a rounded multiply and a
rounded add, the numerics the passes are specified with) and with packed-FP32
instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libgi.so"

HIP_SOURCES = ["raster.hip", "screen.hip", "diff.hip", "batch.hip", "noalloc.hip"]
CXX_SOURCES = [CSRC / "gi_math.cpp", PKG / "host" / "scene_loader.cpp"]

ARCH = os.environ.get("GI_OFFLOAD_ARCH", "gfx950")


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the MI355X library cannot be built")


def hip_flags() -> list[str]:
    return [
        "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
        f"-I{ROOT / 'include'}",
    ]


def cxx_flags() -> list[str]:
    return ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
            "-Wall", "-Wextra", f"-I{ROOT / 'include'}"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ... ({res.returncode})")
    if verbose and (res.stdout or res.stderr):
        sys.stdout.write(res.stdout + res.stderr)


def build(verbose: bool = False) -> Path:
    """Compiles every source and links ``libgi.so``; always rebuilds."""
    BUILD.mkdir(exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        obj = BUILD / (Path(src).stem + ".hip.o")
        objs.append(obj)
        jobs.append([hipcc, *hip_flags(), "-c", str(CSRC / src), "-o", str(obj)])
    for src in CXX_SOURCES:
        obj = BUILD / (Path(src).stem + ".cpp.o")
        objs.append(obj)
        jobs.append(["g++", *cxx_flags(), "-c", str(src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)], verbose)
    os.replace(tmp, LIB)
    build_host_test(hipcc, verbose)
    return LIB


HOST_TEST = ROOT / "tests" / "cpp" / "host_pipeline_test"


def build_host_test(hipcc: str, verbose: bool = False) -> Path:
    """The C++ host pass interface (host/shadow_passes.cpp) linked into the
    driver the GPU tests run; it calls libgi.so through the C ABI only."""
    tmp = HOST_TEST.with_suffix(".tmp")
    _run([hipcc, "-O2", "-std=c++17", "-ffp-contract=off", f"-I{ROOT / 'include'}",
          str(ROOT / "tests" / "cpp" / "host_pipeline_test.cpp"), str(PKG / "host" / "shadow_passes.cpp"),
          f"-L{PKG}", "-lgi", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/../../globalillumination_amd",
          "-o", str(tmp)], verbose)
    os.replace(tmp, HOST_TEST)
    return HOST_TEST


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
