"""ctypes binding of libpme_hip.so (the C ABI declared in include/pme_recip.h).

The library is loaded from this package directory only.  There is no
fallback: if the HIP library is missing or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpme_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "pme_recip.h")

PME_OK = 0
STATUS = {
    0: "PME_OK", 1: "PME_ERR_INVALID_ARG", 2: "PME_ERR_UNSUPPORTED", 3: "PME_ERR_OOM",
    4: "PME_ERR_BUDGET", 5: "PME_ERR_DEVICE", 6: "PME_ERR_STATE", 7: "PME_ERR_NOT_FOUND",
    8: "PME_ERR_STOPPED",
}
PME_ERR_STOPPED = 8
FFT_COLUMN, FFT_SLAB, FFT_BOX = 0, 1, 2
ACCUM_FLOAT, ACCUM_INT32, ACCUM_INT64 = 0, 1, 2
FORCE_FLOAT, FORCE_FLOAT3, FORCE_INT64 = 0, 1, 2
GRID_REAL, GRID_SPECTRUM = 0, 1
PREC_FP32, PREC_BF16X3, PREC_FP16X3 = 0, 1, 2
CALC_EV, CALC_SELF = 1, 2


class PmeError(RuntimeError):
    def __init__(self, status: int, func: str, msg: str):
        self.status = status
        self.name = STATUS.get(status, str(status))
        super().__init__(f"{func}: {self.name}: {msg}")


_vp = C.c_void_p
_i = C.c_int
_d = C.c_double
_pd = C.POINTER(C.c_double)



class SystemDesc(C.Structure):
    """pme_system_desc (include/pme_recip.h)."""
    _fields_ = [
        ("nfftx", C.c_int), ("nffty", C.c_int), ("nfftz", C.c_int), ("order", C.c_int),
        ("kappa", C.c_double), ("recip", C.c_double * 9), ("ncoord", C.c_int),
        ("d_xyzq", C.c_void_p), ("d_force", C.c_void_p), ("force_type", C.c_int),
        ("stride", C.c_int), ("calc_energy_virial", C.c_int), ("precision", C.c_int),
    ]


class ServerStats(C.Structure):
    """pme_server_stats (include/pme_recip.h)."""
    _fields_ = [
        ("submitted", C.c_ulonglong), ("completed", C.c_ulonglong), ("failed", C.c_ulonglong),
        ("batches", C.c_ulonglong), ("mean_batch", C.c_double), ("max_batch", C.c_int),
        ("evictions", C.c_ulonglong), ("admitted_new", C.c_ulonglong), ("resident", C.c_int),
        ("arena_used", C.c_size_t), ("arena_budget", C.c_size_t),
        ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Nlist(C.Structure):
    """pme_nlist: the tile neighbour list of the direct-space term (device pointers)."""
    _fields_ = [("n_ientry", C.c_int), ("ientry", C.c_void_p), ("n_tile", C.c_int), ("tile_indj", C.c_void_p),
                ("tile_excl", C.c_void_p)]


class RecipHeader(C.Structure):
    """pme_recip_header: the direct->recip request header (DomdecRecipComm.h:11-15)."""
    _fields_ = [("inv_boxx", C.c_double), ("inv_boxy", C.c_double), ("inv_boxz", C.c_double),
                ("ncoord", C.c_int), ("calc_energy", C.c_ubyte), ("calc_virial", C.c_ubyte)]


class RequestResult(C.Structure):
    """pme_request_result: what one served request returns."""
    _fields_ = [("energy", C.c_double), ("energy_self", C.c_double), ("virial", C.c_double * 9),
                ("latency_ms", C.c_double)]


_pi = C.POINTER(C.c_int)
_pu64 = C.POINTER(C.c_ulonglong)
_sz = C.c_size_t

# name -> argtypes (restype is pme_status = int unless listed in _RESTYPE)
SIGNATURES = {
    "pme_last_error": [],
    "pme_status_string": [_i],
    "pme_ev_create": [C.POINTER(_vp)],
    "pme_ev_destroy": [_vp],
    "pme_ev_insert": [_vp, C.c_char_p],
    "pme_ev_clear": [_vp, _vp],
    "pme_ev_clear_energy": [_vp, _vp],
    "pme_ev_clear_virial": [_vp, _vp],
    "pme_ev_copy_to_host": [_vp, _vp],
    "pme_ev_get_energy": [_vp, C.c_char_p, _pd],
    "pme_ev_get_virial": [_vp, _pd],
    "pme_ev_get_sforce": [_vp, _pd],
    "pme_ev_energy_pointer": [_vp, C.c_char_p, C.POINTER(_vp)],
    "pme_ev_virial_pointer": [_vp, C.POINTER(_vp)],
    "pme_recip_create": [_i, _i, _i, _i, _i, _i, _i, _i, _vp, C.c_char_p, C.c_char_p, _vp, C.POINTER(_vp)],
    "pme_recip_destroy": [_vp],
    "pme_recip_set_stream": [_vp, _vp],
    "pme_recip_set_order": [_vp, _i],
    "pme_recip_set_precision": [_vp, _i],
    "pme_recip_set_graph": [_vp, _i],
    "pme_recip_get_info": [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)],
    "pme_recip_spread_charge": [_vp, _vp, _i, _pd],
    "pme_recip_r2c_fft": [_vp],
    "pme_recip_scalar_sum": [_vp, _pd, _d, _i, _i],
    "pme_recip_c2r_fft": [_vp],
    "pme_recip_gather_force": [_vp, _vp, _i, _pd, _i, _i, _vp],
    "pme_recip_calc_self_energy": [_vp, _vp, _i, _d],
    "pme_recip_solve": [_vp, _vp, _i, _pd, _d, _i, _i, _i, _i, _vp],
    "pme_recip_solve_noalloc": [_vp, _vp, _i, _pd, _d, _i, _i, _i, _i, _vp],
    "pme_recip_copy_grid": [_vp, _i, _vp, C.c_size_t],
    "pme_recip_set_timing": [_vp, _i],
    "pme_recip_get_timing": [_vp, C.POINTER(C.c_float)],
    "pme_fill_bspline": [_i, _i, _i, _i, _pd, _vp, _i] + [_vp] * 10 + [_vp],
    "pme_bsp_mod": [_i, _i, _pd],
    "pme_batch_create": [_sz, _vp, C.POINTER(_vp)],
    "pme_batch_destroy": [_vp],
    "pme_batch_add": [_vp, C.POINTER(SystemDesc), _pi],
    "pme_batch_update": [_vp, _i, C.POINTER(SystemDesc)],
    "pme_batch_remove": [_vp, _i],
    "pme_batch_solve": [_vp, _i, _pi],
    "pme_batch_results": [_vp, _i, _pi, _vp, _vp],
    "pme_batch_results_self": [_vp, _i, _pi, _vp],
    "pme_batch_set_graph": [_vp, _i],
    "pme_batch_set_tuning": [_vp, _i, _i, _i],
    "pme_batch_get_tuning": [_vp, _pi, _pi, _pi],
    "pme_batch_set_chains": [_vp, _i],
    "pme_batch_get_chains": [_vp, _pi],
    "pme_batch_set_half_bricks": [_vp, _i],
    "pme_batch_get_half_bricks": [_vp, _pi],
    "pme_batch_autotune": [_vp, C.c_char_p, C.POINTER(C.c_float)],
    "pme_batch_memory": [_vp, C.POINTER(_sz), C.POINTER(_sz)],
    "pme_batch_set_timing": [_vp, _i],
    "pme_server_create": [_sz, _i, C.POINTER(_vp)],
    "pme_server_destroy": [_vp],
    "pme_server_submit": [_vp, _i, C.POINTER(SystemDesc), _vp, _pu64],
    "pme_server_submit_header": [_vp, _i, C.POINTER(RecipHeader), _i, _i, _i, _i, _d, _vp, _vp, _pu64],
    "pme_server_wait": [_vp, C.c_ulonglong, _pd, _pd, _pd],
    "pme_server_wait_result": [_vp, C.c_ulonglong, C.POINTER(RequestResult)],
    "pme_server_poll": [_vp, C.c_ulonglong, _pi, _pd, _pd, _pd],
    "pme_server_drain": [_vp],
    "pme_server_release": [_vp, _i],
    "pme_server_stats_get": [_vp, C.POINTER(ServerStats)],
    "pme_server_stats_reset": [_vp],
    "pme_server_set_tuning": [_vp, _i, _i, _i],
    "pme_batch_get_timing": [_vp, C.POINTER(C.c_float)],
    "pme_direct_create": [_vp, C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(_vp)],
    "pme_direct_destroy": [_vp],
    "pme_direct_setup": [_vp, _d, _d, _d, _d, _d, _d, _d, _i, _i],
    "pme_direct_get_setup": [_vp] + [C.POINTER(C.c_float)] * 7 + [C.POINTER(_i)] * 2,
    "pme_direct_set_box_size": [_vp, C.c_float, C.c_float, C.c_float],
    "pme_direct_set_calc_vdw": [_vp, _i],
    "pme_direct_set_calc_elec": [_vp, _i],
    "pme_direct_set_14_list": [_vp, _i, _i, _vp, _vp, _vp],
    "pme_direct_calc_14_force": [_vp, _vp, _i, _i, _i, _vp, _vp],
    "pme_direct_calc_force": [_vp, _vp, C.POINTER(Nlist), _i, _i, _i, _vp, _vp],
}
_RESTYPE = {"pme_last_error": C.c_char_p, "pme_status_string": C.c_char_p}

_lib = None


def header_symbols() -> list[str]:
    """Every function the C header declares (pme_* names followed by '(')."""
    with open(HEADER) as f:
        txt = f.read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pme_[a-z0-9_]+)\s*\(", txt)))


def load() -> C.CDLL:
    """Load libpme_hip.so and bind every declared symbol.  Raises if missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} not built: run `python -m gpu_recip_amd.build` "
                           "(there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, argt in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argt
        fn.restype = _RESTYPE.get(name, C.c_int)
    _lib = lib
    return lib


def check(status: int, func: str) -> None:
    if status != PME_OK:
        msg = load().pme_last_error()
        raise PmeError(status, func, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


_dbl9_cache: dict = {}


def dbl9(recip) -> C.Array:
    """A C double[9] of the reciprocal box.  The C ABI copies the values during
    the call, so the arrays of recently used boxes are kept and reused (a
    solve passes the same box to three calls)."""
    key = tuple(recip)
    arr = _dbl9_cache.get(key)
    if arr is None:
        if len(_dbl9_cache) >= 64:
            _dbl9_cache.clear()
        arr = _dbl9_cache[key] = (C.c_double * 9)(*[float(v) for v in key])
    return arr
