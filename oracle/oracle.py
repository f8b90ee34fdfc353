"""ctypes binding of oracle/liboracle.so (FP64 restatement of CudaPMERecip).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
Each wrapper names the C function it calls; the C source cites the reference
file:line each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO_PATH = os.path.join(HERE, "liboracle.so")

INDEX_F64, INDEX_F32 = 0, 1
NYQUIST_REFERENCE, NYQUIST_EXACT = 0, 1
CCELEC = 332.0716

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class _Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("order", C.c_int),
        ("kappa", C.c_double), ("recip", C.c_double * 9),
        ("index_mode", C.c_int), ("nyquist_mode", C.c_int), ("nthreads", C.c_int),
    ]


@dataclass
class OracleParams:
    nx: int
    ny: int
    nz: int
    order: int
    kappa: float
    box: tuple  # orthorhombic edge lengths (A)
    index_mode: int = INDEX_F64
    nyquist_mode: int = NYQUIST_REFERENCE
    nthreads: int = 0
    recip: np.ndarray = field(default=None)

    def __post_init__(self):
        if self.recip is None:
            r = np.zeros(9)
            r[0], r[4], r[8] = 1.0 / self.box[0], 1.0 / self.box[1], 1.0 / self.box[2]
            self.recip = r

    def c(self) -> _Params:
        p = _Params()
        p.nx, p.ny, p.nz, p.order = self.nx, self.ny, self.nz, self.order
        p.kappa = self.kappa
        for i in range(9):
            p.recip[i] = float(self.recip[i])
        p.index_mode, p.nyquist_mode, p.nthreads = self.index_mode, self.nyquist_mode, self.nthreads
        return p


def build() -> None:
    """Compile liboracle.so with oracle/Makefile (gcc -fopenmp)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(SO_PATH):
        build()
    L = C.CDLL(SO_PATH)
    P = C.POINTER(_Params)
    L.orc_index_f32.argtypes = [C.c_int, _fp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _ip, _fp]
    L.orc_index_f64.argtypes = [C.c_int, _fp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _ip, _dp]
    L.orc_bspline.argtypes = [C.c_int, C.c_double, _dp, _dp]
    L.orc_bsp_mod.argtypes = [C.c_int, C.c_int, _dp]
    L.orc_spread.argtypes = [P, C.c_int, _fp, _dp]
    L.orc_fft1d.argtypes = [C.c_int, _dp, C.c_int]
    L.orc_fft3d_r2c.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int]
    L.orc_fft3d_c2r.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int]
    L.orc_scalar_sum.argtypes = [P, _dp, C.POINTER(C.c_double), _dp]
    L.orc_gather.argtypes = [P, C.c_int, _fp, _dp, _dp]
    L.orc_self_energy.argtypes = [C.c_int, _fp, C.c_double]
    L.orc_self_energy.restype = C.c_double
    L.orc_solve.argtypes = [P, C.c_int, _fp, _dp, C.POINTER(C.c_double), _dp, C.c_void_p, C.c_void_p]
    L.orc_solve.restype = C.c_int
    L.orc_ewald_recip.argtypes = [C.c_int, _fp, _dp, C.c_double, _ip, C.POINTER(C.c_double), C.c_void_p, _dp, C.c_int]
    L.orc_ewald_recip9.argtypes = [C.c_int, _fp, _dp, C.c_double, _ip, C.POINTER(C.c_double), C.c_void_p, _dp, C.c_int]
    L.orc_direct_ewald.argtypes = [C.c_int, _fp, _dp, C.c_double, C.c_double, C.c_int, C.c_void_p, _dp, _dp, C.c_int]
    L.orc_direct_ewald.restype = C.c_double
    L.orc_pair14.argtypes = [C.c_int, _fp, _dp, C.c_double, C.c_double, C.c_int, C.c_void_p, _dp, _dp]
    L.orc_pair14.restype = C.c_double
    L.orc_direct_tiles.argtypes = [C.c_int, _fp, _dp, C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_void_p,
                                   C.c_void_p, _dp, _dp, _dp, C.c_int]
    L.orc_direct_tiles.restype = C.c_double
    _lib = L
    return L


def _xyzq(x) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    assert x.ndim == 2 and x.shape[1] == 4
    return x


def index_f32(xyzq, recip_diag, n):
    """Reference FP32 grid index + fraction (orc_index_f32)."""
    x = _xyzq(xyzq)
    idx = np.zeros((len(x), 3), np.int32)
    fr = np.zeros((len(x), 3), np.float32)
    lib().orc_index_f32(len(x), x, *map(float, recip_diag), *n, idx, fr)
    return idx, fr


def index_f64(xyzq, recip_diag, n):
    x = _xyzq(xyzq)
    idx = np.zeros((len(x), 3), np.int32)
    fr = np.zeros((len(x), 3), np.float64)
    lib().orc_index_f64(len(x), x, *map(float, recip_diag), *n, idx, fr)
    return idx, fr


def bspline(order, w):
    th = np.zeros(order)
    dth = np.zeros(order)
    lib().orc_bspline(order, float(w), th, dth)
    return th, dth


def bsp_mod(order, n):
    out = np.zeros(n)
    lib().orc_bsp_mod(order, n, out)
    return out


def spread(p: OracleParams, xyzq):
    x = _xyzq(xyzq)
    g = np.zeros(p.nx * p.ny * p.nz)
    lib().orc_spread(C.byref(p.c()), len(x), x, g)
    return g.reshape(p.nz, p.ny, p.nx)


def fft1d(a, sign=-1):
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    v = a.view(np.float64)
    lib().orc_fft1d(len(a), v, sign)
    return v.view(np.complex128)


def fft3d_r2c(real_zyx, nthreads=0):
    nz, ny, nx = real_zyx.shape
    r = np.ascontiguousarray(real_zyx, dtype=np.float64)
    out = np.zeros((nz, ny, nx // 2 + 1), np.complex128)
    lib().orc_fft3d_r2c(nx, ny, nz, r.reshape(-1), out.view(np.float64).reshape(-1), nthreads)
    return out


def fft3d_c2r(cplx_zyx, nx, nthreads=0):
    nz, ny, hx = cplx_zyx.shape
    assert hx == nx // 2 + 1
    c = np.ascontiguousarray(cplx_zyx, dtype=np.complex128)
    out = np.zeros((nz, ny, nx), np.float64)
    lib().orc_fft3d_c2r(nx, ny, nz, c.view(np.float64).reshape(-1), out.reshape(-1), nthreads)
    return out


def scalar_sum(p: OracleParams, cplx_zyx):
    """Returns (solved spectrum copy, energy, virial[9])."""
    c = np.ascontiguousarray(cplx_zyx, dtype=np.complex128).copy()
    e = C.c_double(0.0)
    v = np.zeros(9)
    lib().orc_scalar_sum(C.byref(p.c()), c.view(np.float64).reshape(-1), C.byref(e), v)
    return c, e.value, v


def gather(p: OracleParams, xyzq, phi_zyx):
    x = _xyzq(xyzq)
    f = np.zeros((len(x), 3))
    lib().orc_gather(C.byref(p.c()), len(x), x, np.ascontiguousarray(phi_zyx, np.float64).reshape(-1), f)
    return f


def self_energy(xyzq, kappa):
    x = _xyzq(xyzq)
    return lib().orc_self_energy(len(x), x, float(kappa))


def solve(p: OracleParams, xyzq, want_grids=False):
    """Full solve.  Returns dict(force (N,3) = +dE/dr, energy, virial (9,) [, grid, phi])."""
    x = _xyzq(xyzq)
    f = np.zeros((len(x), 3))
    e = C.c_double(0.0)
    v = np.zeros(9)
    G = p.nx * p.ny * p.nz
    grid = np.zeros(G) if want_grids else None
    phi = np.zeros(G) if want_grids else None
    rc = lib().orc_solve(C.byref(p.c()), len(x), x, f, C.byref(e), v,
                         grid.ctypes.data if want_grids else None,
                         phi.ctypes.data if want_grids else None)
    if rc != 0:
        raise RuntimeError(f"orc_solve failed rc={rc}")
    out = {"force": f, "energy": e.value, "virial": v}
    if want_grids:
        out["grid"] = grid.reshape(p.nz, p.ny, p.nx)
        out["phi"] = phi.reshape(p.nz, p.ny, p.nx)
    return out


def ewald_recip(xyzq, box, kappa, kmax, want_force=True, nthreads=0):
    x = _xyzq(xyzq)
    e = C.c_double(0.0)
    v = np.zeros(9)
    f = np.zeros((len(x), 3)) if want_force else None
    lib().orc_ewald_recip(len(x), x, np.asarray(box, np.float64), float(kappa),
                          np.asarray(kmax, np.int32), C.byref(e),
                          f.ctypes.data if want_force else None, v, nthreads)
    return {"energy": e.value, "force": f, "virial": v}


def direct_ewald(xyzq, box, kappa, roff, excl=None, nthreads=0):
    """Direct-space Ewald by brute force over all pairs and periodic images
    within roff, excluded pairs skipped (orc_direct_ewald).  Returns
    dict(energy, force (N,3) = +dE/dr, virial (9,) = sum_pairs r_a dE/dr_b)."""
    x = _xyzq(xyzq)
    f = np.zeros((len(x), 3))
    v = np.zeros(9)
    ex = None if excl is None or len(excl) == 0 else np.ascontiguousarray(excl, np.int32).reshape(-1, 2)
    e = lib().orc_direct_ewald(len(x), x, np.asarray(box, np.float64), float(kappa), float(roff),
                               0 if ex is None else len(ex), None if ex is None else ex.ctypes.data, f, v, nthreads)
    return {"energy": e, "force": f, "virial": v}


def direct_tiles(xyzq_sorted, box, kappa, roff, ientry, tile_indj, tile_excl, nthreads=0):
    """Direct-space sum over a 32-atom tile list (orc_direct_tiles), in the
    list's atom order.  Returns dict(energy, force (npad,3) = +dE/dr, virial
    (9,), sforce (27,3))."""
    x = _xyzq(xyzq_sorted)
    f = np.zeros((len(x), 3))
    v = np.zeros(9)
    sf = np.zeros(81)
    ie = np.ascontiguousarray(ientry, np.int32).reshape(-1, 4)
    tj = np.ascontiguousarray(tile_indj, np.int32)
    te = np.ascontiguousarray(tile_excl, np.uint32).reshape(-1, 32)
    e = lib().orc_direct_tiles(len(x), x, np.asarray(box, np.float64), float(kappa), float(roff), len(ie),
                               ie.ctypes.data, tj.ctypes.data, te.ctypes.data, f, v, sf, nthreads)
    return {"energy": e, "force": f, "virial": v, "sforce": sf.reshape(27, 3)}


def pair14(xyzq, box, kappa, fac, pairs):
    """1-4 list energy (orc_pair14): pairs (M,3) of (i, j, ishift); fac = e14fac
    for interactions, 0 for exclusions.  Returns dict(energy, force, virial)."""
    x = _xyzq(xyzq)
    f = np.zeros((len(x), 3))
    v = np.zeros(9)
    pl = np.ascontiguousarray(pairs, np.int32).reshape(-1, 3)
    e = lib().orc_pair14(len(x), x, np.asarray(box, np.float64), float(kappa), float(fac), len(pl),
                         pl.ctypes.data if len(pl) else None, f, v)
    return {"energy": e, "force": f, "virial": v}


def ewald_recip9(xyzq, recip9, kappa, kmax, want_force=True, nthreads=0):
    """Brute-force Ewald reciprocal sum for any (triclinic) recip (orc_ewald_recip9)."""
    x = _xyzq(xyzq)
    e = C.c_double(0.0)
    v = np.zeros(9)
    f = np.zeros((len(x), 3)) if want_force else None
    lib().orc_ewald_recip9(len(x), x, np.ascontiguousarray(recip9, np.float64).reshape(9), float(kappa),
                           np.asarray(kmax, np.int32), C.byref(e), f.ctypes.data if want_force else None, v,
                           nthreads)
    return {"energy": e.value, "force": f, "virial": v}
