"""Builds libpfem_gpu.so in place (hipcc for gfx950; one object per translation unit)."""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpfem_gpu.so")
SOURCES = ["pattern.hip", "ell.hip", "pspg.hip", "heat.hip", "fracstep.hip", "krylov.hip", "picard.hip", "wc.hip", "extract.hip", "amg.hip", "microbench.hip",
           "amg_hierarchy.cpp",
           "engine.cpp", "capi.cpp", "mesher.cpp"]
# host-only translation units (no device code): plain C++, no floating-point contraction (the
# mesher's predicates bound the rounding of every product and sum separately)
HOST_ONLY = {"mesher.cpp"}
ARCH = os.environ.get("PFEM_OFFLOAD_ARCH", "gfx950")

CXXFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay disabled (hipcc would otherwise form them from float code)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # FP64 atomic add as the hardware instruction (device memory from hipMalloc)
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    return os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if src in HOST_ONLY:
        flags = [f for f in CXXFLAGS if not f.startswith("--offload-arch") and f != "-munsafe-fp-atomics"]
        cmd = [_hipcc(), *flags, "-ffp-contract=off", "-x", "c++", "-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = [_hipcc(), *CXXFLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
