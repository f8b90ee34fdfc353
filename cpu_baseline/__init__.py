"""ctypes binding of cpu_baseline/libpme_cpu.so: the FP32 CPU restatement of
the reference's PME reciprocal solve (same index formula, B-spline recursion
and stage order, FP32 grids, FP64 energy/virial sums), multithreaded with
OpenMP.

For the CPU comparison only: bench.py times it as the CPU baseline and
tests/ check it against the FP64 oracle.  The product package never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO_PATH = os.path.join(HERE, "libpme_cpu.so")

_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile libpme_cpu.so with cpu_baseline/Makefile (g++ -fopenmp)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(SO_PATH):
        build()
    L = C.CDLL(SO_PATH)
    L.cpu_index_f32.argtypes = [C.c_int, _fp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _ip, _fp]
    L.cpu_pme_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    L.cpu_pme_create.restype = C.c_void_p
    L.cpu_pme_destroy.argtypes = [C.c_void_p]
    L.cpu_pme_run.argtypes = [C.c_void_p, C.c_double, _dp, C.c_int, _fp, _fp, C.POINTER(C.c_double), _dp, C.c_int]
    L.cpu_pme_run.restype = C.c_int
    L.cpu_max_threads.restype = C.c_int
    _lib = L
    return L


def index_f32(xyzq, recip_diag, n):
    """Grid index + fraction of the FP32 formula (cpu_index_f32)."""
    x = np.ascontiguousarray(xyzq, dtype=np.float32)
    idx = np.zeros((len(x), 3), np.int32)
    fr = np.zeros((len(x), 3), np.float32)
    lib().cpu_index_f32(len(x), x, *map(float, recip_diag), *n, idx, fr)
    return idx, fr


def max_threads() -> int:
    return lib().cpu_max_threads()


class CpuPme:
    """One solver per grid and order (cpu_pme_create), reused across solves."""

    def __init__(self, nfft, order):
        self.nfft = tuple(int(v) for v in nfft)
        self.order = int(order)
        self.h = lib().cpu_pme_create(*self.nfft, self.order)
        if not self.h:
            raise ValueError(f"cpu_pme_create: unsupported grid {self.nfft} / order {order}")

    def __del__(self):
        if getattr(self, "h", None):
            lib().cpu_pme_destroy(self.h)
            self.h = None

    def solve(self, xyzq, box, kappa, nthreads=0, force=None):
        """Full solve.  Returns dict(force (N,3) float32 = +dE/dr, energy, virial (9,))."""
        x = np.ascontiguousarray(xyzq, dtype=np.float32)
        f = force if force is not None else np.zeros((len(x), 3), np.float32)
        e = C.c_double(0.0)
        v = np.zeros(9)
        rc = lib().cpu_pme_run(self.h, float(kappa), np.asarray(box, np.float64), len(x), x, f, C.byref(e), v,
                               int(nthreads))
        if rc != 0:
            raise RuntimeError(f"cpu_pme_run failed rc={rc}")
        return {"force": f, "energy": e.value, "virial": v}


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), or the machine type."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or platform.machine()
