"""Build libpme_hip.so (the HIP kernels + C ABI) in-tree for gfx950.

    python -m gpu_recip_amd.build        # or gpu_recip_amd.build.build()

hipcc cross-compiles without a GPU.  Every source is compiled as HIP with
packed-FP32 instruction selection disabled (-packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libpme_hip.so")
OBJ = os.path.join(HERE, "_build")

ARCH = os.environ.get("PME_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(CSRC, f))
    return out


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, "-x", "hip", "-c", src, "-o", obj] + COMMON
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    """Compile every csrc/ source and link libpme_hip.so; returns its path."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-o", tmp] + objs + [f"--offload-arch={ARCH}", "-fPIC"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


DROPIN_SRC = os.path.join(ROOT, "tests", "cpp", "gpu_recip_dropin.cpp")
DROPIN_BIN = os.path.join(ROOT, "tests", "cpp", "gpu_recip_dropin")
DIRECT_DROPIN_SRC = os.path.join(ROOT, "tests", "cpp", "direct_dropin.cpp")
DIRECT_DROPIN_BIN = os.path.join(ROOT, "tests", "cpp", "direct_dropin")


def build_dropin_test() -> str:
    """Compile the C++ drop-in test programs (host code over include/pme_recip.hpp,
    linked against libpme_hip.so with an rpath to this package)."""
    for src, out in ((DROPIN_SRC, DROPIN_BIN), (DIRECT_DROPIN_SRC, DIRECT_DROPIN_BIN)):
        cmd = [HIPCC, "-std=c++17", "-O2", f"-I{INCLUDE}", src, "-o", out, f"-L{HERE}", "-lpme_hip",
               f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/../../gpu_recip_amd"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"drop-in test build failed ({os.path.basename(src)}):\n{r.stdout}\n{r.stderr}")
    return DROPIN_BIN


if __name__ == "__main__":
    build(verbose=True)
    build_dropin_test()
    sys.exit(0)
