"""Python mirror of the reference's CudaPMERecip / CudaEnergyVirial interface
(CudaPMERecip.h:166-203, CudaEnergyVirial.h:20-53) over the C ABI.

Device buffers are torch tensors on a ROCm device (plumbing only); every
compute call goes through libpme_hip.so.  Method names and argument meaning
follow the reference; errors raise PmeError instead of exit(1).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def _ptr(t) -> int:
    return 0 if t is None else int(t.data_ptr())


def _stream_ptr(stream) -> int:
    if stream is None:
        return 0
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)


class CudaEnergyVirial:
    """Named FP64 energy slots + 3x3 virial (CudaEnergyVirial.h:20-53)."""

    def __init__(self):
        h = C.c_void_p()
        L.call("pme_ev_create", C.byref(h))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) is not None and L._lib is not None:
            L._lib.pme_ev_destroy(self.h)
            self.h = None

    def insert(self, name: str) -> None:
        L.call("pme_ev_insert", self.h, name.encode())

    def clear(self, stream=None) -> None:
        L.call("pme_ev_clear", self.h, _stream_ptr(stream))

    def clearEnergy(self, stream=None) -> None:
        L.call("pme_ev_clear_energy", self.h, _stream_ptr(stream))

    def clearVirial(self, stream=None) -> None:
        L.call("pme_ev_clear_virial", self.h, _stream_ptr(stream))

    def copyToHost(self, stream=None) -> None:
        L.call("pme_ev_copy_to_host", self.h, _stream_ptr(stream))

    def getEnergy(self, name: str) -> float:
        out = C.c_double()
        L.call("pme_ev_get_energy", self.h, name.encode(), C.byref(out))
        return out.value

    def getVirial(self) -> list[float]:
        v = (C.c_double * 9)()
        L.call("pme_ev_get_virial", self.h, v)
        return list(v)

    def getSforce(self) -> list[float]:
        """sforce[27*3] = sforce_dp + sforce_fp / 2^30 of the host copy."""
        v = (C.c_double * 81)()
        L.call("pme_ev_get_sforce", self.h, v)
        return list(v)

    def getEnergyPointer(self, name: str) -> int:
        """Device address of energy term `name` (a double)."""
        p = C.c_void_p()
        L.call("pme_ev_energy_pointer", self.h, name.encode(), C.byref(p))
        return int(p.value or 0)

    def getVirialPointer(self) -> int:
        """Device address of the Virial_t blob: sforce_dp[27][3] doubles,
        sforce_fp[27][3] int64, virmat[9] doubles, then the energy terms."""
        p = C.c_void_p()
        L.call("pme_ev_virial_pointer", self.h, C.byref(p))
        return int(p.value or 0)


class CudaPMERecip:
    """The reciprocal-space solver (CudaPMERecip<AT,CT,CT2>).

    xyzq: device float32 tensor (ncoord, 4); recip: 9 floats (row-major,
    orthorhombic); forces: device tensors laid out per force_type.
    """

    def __init__(self, nfftx, nffty, nfftz, order, energyVirial: CudaEnergyVirial,
                 nameRecip="recip", nameSelf="self", fft_type=L.FFT_BOX, nnode=1, mynode=0,
                 accum_type=L.ACCUM_FLOAT, stream=None):
        h = C.c_void_p()
        L.call("pme_recip_create", nfftx, nffty, nfftz, order, fft_type, nnode, mynode, accum_type,
               energyVirial.h, nameRecip.encode(), nameSelf.encode(), _stream_ptr(stream), C.byref(h))
        self.h = h
        self.ev = energyVirial  # keep alive
        self.nfft = (nfftx, nffty, nfftz)

    def __del__(self):
        if getattr(self, "h", None) is not None and L._lib is not None:
            L._lib.pme_recip_destroy(self.h)
            self.h = None

    def set_stream(self, stream) -> None:
        L.call("pme_recip_set_stream", self.h, _stream_ptr(stream))

    def set_order(self, order: int) -> None:
        L.call("pme_recip_set_order", self.h, order)

    def set_precision(self, precision: int) -> None:
        """PREC_FP32, PREC_BF16X3 or PREC_FP16X3 for the y/z FFT passes."""
        L.call("pme_recip_set_precision", self.h, int(precision))

    def set_graph(self, on: bool) -> None:
        """solve() as a hipGraph replay on the handle's own stream (off by default)."""
        L.call("pme_recip_set_graph", self.h, int(bool(on)))

    def get_info(self):
        v = [C.c_int() for _ in range(4)]
        L.call("pme_recip_get_info", self.h, *[C.byref(x) for x in v])
        return tuple(x.value for x in v)

    def spread_charge(self, xyzq, ncoord, recip) -> None:
        L.call("pme_recip_spread_charge", self.h, _ptr(xyzq), int(ncoord), L.dbl9(recip))

    def r2c_fft(self) -> None:
        L.call("pme_recip_r2c_fft", self.h)

    def scalar_sum(self, recip, kappa, calc_energy, calc_virial) -> None:
        L.call("pme_recip_scalar_sum", self.h, L.dbl9(recip), float(kappa), int(bool(calc_energy)),
               int(bool(calc_virial)))

    def c2r_fft(self) -> None:
        L.call("pme_recip_c2r_fft", self.h)

    def gather_force(self, xyzq, ncoord, recip, stride, force, force_type=L.FORCE_FLOAT) -> None:
        L.call("pme_recip_gather_force", self.h, _ptr(xyzq), int(ncoord), L.dbl9(recip), int(stride),
               int(force_type), _ptr(force))

    def calc_self_energy(self, xyzq, ncoord, kappa) -> None:
        L.call("pme_recip_calc_self_energy", self.h, _ptr(xyzq), int(ncoord), float(kappa))

    def solve(self, xyzq, ncoord, recip, kappa, calc_energy, calc_virial, stride, force,
              force_type=L.FORCE_FLOAT) -> None:
        L.call("pme_recip_solve", self.h, _ptr(xyzq), int(ncoord), L.dbl9(recip), float(kappa),
               int(bool(calc_energy)), int(bool(calc_virial)), int(stride), int(force_type), _ptr(force))

    def solve_noalloc(self, xyzq, ncoord, recip, kappa, calc_energy, calc_virial, stride, force,
                      force_type=L.FORCE_FLOAT) -> None:
        """solve() with the work buffers in the library's static device
        scratch (no allocation; PME_ERR_BUDGET beyond its size)."""
        L.call("pme_recip_solve_noalloc", self.h, _ptr(xyzq), int(ncoord), L.dbl9(recip), float(kappa),
               int(bool(calc_energy)), int(bool(calc_virial)), int(stride), int(force_type), _ptr(force))

    def copy_grid(self, which, dst) -> None:
        L.call("pme_recip_copy_grid", self.h, int(which), _ptr(dst), int(dst.numel() * dst.element_size()))

    def set_timing(self, enable: bool) -> None:
        L.call("pme_recip_set_timing", self.h, int(bool(enable)))

    def get_timing(self):
        ms = (C.c_float * 8)()
        L.call("pme_recip_get_timing", self.h, ms)
        return list(ms)


def force_stride(ncoord: int, elem_size: int = 4) -> int:
    """Force<T>::realloc stride (Force.h:63-69): aligned to 256-byte boundaries."""
    return (((ncoord - 1 + 32) * elem_size - 1) // 256 + 1) * 256 // elem_size


def fill_bspline(order, nfft, recip, xyzq, ncoord, stream=None):
    """Bspline<float>::fill_bspline (Bspline.cu:203-231): device arrays
    (gix, giy, giz, charge, thetax, thetay, thetaz, dthetax, dthetay, dthetaz)."""
    import torch

    dev = xyzq.device
    gi = [torch.empty(ncoord, dtype=torch.int32, device=dev) for _ in range(3)]
    q = torch.empty(ncoord, dtype=torch.float32, device=dev)
    th = [torch.empty(ncoord * order, dtype=torch.float32, device=dev) for _ in range(6)]
    L.call("pme_fill_bspline", int(order), int(nfft[0]), int(nfft[1]), int(nfft[2]), L.dbl9(recip),
           _ptr(xyzq), int(ncoord), *[_ptr(t) for t in gi], _ptr(q), *[_ptr(t) for t in th],
           _stream_ptr(stream))
    return (*gi, q, *th)


class PmeBatch:
    """Many independent systems solved per launch (pme_batch_* in the C ABI).

    Systems are admitted with add() against the HBM budget and stay resident;
    solve() enqueues one solve of a subset (all by default) on the batch's
    stream, results() waits and returns (energy[n], virial[n, 9]) as float64 arrays.
    """

    def __init__(self, hbm_budget_bytes: int, stream=None):
        h = C.c_void_p()
        L.call("pme_batch_create", C.c_size_t(int(hbm_budget_bytes)), _stream_ptr(stream), C.byref(h))
        self.h = h
        self._keep = {}  # id -> tensors the descriptor points at
        self._id_arrays = {}  # tuple of ids -> (n, ctypes int array), reused by repeated solves

    def __del__(self):
        if getattr(self, "h", None) is not None and L._lib is not None:
            L._lib.pme_batch_destroy(self.h)
            self.h = None

    @staticmethod
    def desc(nfft, order, kappa, recip, xyzq, ncoord, force, force_type=L.FORCE_FLOAT, stride=None,
             calc_energy_virial=True, precision=L.PREC_FP32) -> "L.SystemDesc":
        d = L.SystemDesc()
        d.nfftx, d.nffty, d.nfftz = (int(v) for v in nfft)
        d.order = int(order)
        d.kappa = float(kappa)
        for i, v in enumerate(recip):
            d.recip[i] = float(v)
        d.ncoord = int(ncoord)
        d.d_xyzq = _ptr(xyzq)
        d.d_force = _ptr(force)
        d.force_type = int(force_type)
        d.stride = int(ncoord if stride is None else stride)
        # True/False, or the L.CALC_EV / L.CALC_SELF bits
        d.calc_energy_virial = int(calc_energy_virial)
        d.precision = int(precision)
        return d

    def add(self, desc, keep=()) -> int:
        i = C.c_int()
        L.call("pme_batch_add", self.h, C.byref(desc), C.byref(i))
        self._keep[i.value] = tuple(keep)
        return i.value

    def update(self, sid: int, desc, keep=()) -> None:
        L.call("pme_batch_update", self.h, int(sid), C.byref(desc))
        self._keep[sid] = tuple(keep)

    def remove(self, sid: int) -> None:
        L.call("pme_batch_remove", self.h, int(sid))
        self._keep.pop(sid, None)

    def _ids(self, ids):
        # a solve loop passes the same ids every step: keep their C array
        if ids is None:
            return 0, None
        key = tuple(ids)
        hit = self._id_arrays.get(key)
        if hit is None:
            if len(self._id_arrays) >= 16:
                self._id_arrays.clear()
            hit = self._id_arrays[key] = (len(key), (C.c_int * len(key))(*[int(i) for i in key]))
        return hit

    def solve(self, ids=None) -> None:
        n, arr = self._ids(ids)
        L.call("pme_batch_solve", self.h, n, arr)

    def results(self, ids):
        """(energies (n,), virials (n, 9)) of the listed systems after the
        last solve, as float64 arrays (waits for the batch's stream)."""
        if ids is None:
            raise ValueError("PmeBatch.results: list the system ids")
        n, arr = self._ids(ids)
        e = np.empty(n, np.float64)
        v = np.empty((n, 9), np.float64)
        L.call("pme_batch_results", self.h, n, arr, e.ctypes.data, v.ctypes.data)
        return e, v

    def results_self(self, ids):
        """Self energies (n,) of the last solve for systems described with
        L.CALC_SELF (0 for the others)."""
        if ids is None:
            raise ValueError("PmeBatch.results_self: list the system ids")
        n, arr = self._ids(ids)
        es = np.empty(n, np.float64)
        L.call("pme_batch_results_self", self.h, n, arr, es.ctypes.data)
        return es

    def set_graph(self, enable: bool) -> None:
        L.call("pme_batch_set_graph", self.h, int(bool(enable)))

    def set_tuning(self, brick_target: int, fft_tile: int, wg_per_cu: int) -> None:
        L.call("pme_batch_set_tuning", self.h, int(brick_target), int(fft_tile), int(wg_per_cu))

    def get_tuning(self):
        v = [C.c_int() for _ in range(3)]
        L.call("pme_batch_get_tuning", self.h, *[C.byref(x) for x in v])
        return tuple(x.value for x in v)

    def set_chains(self, nchains: int) -> None:
        """1 to 4 stage chains per solve (pme_batch_set_chains)."""
        L.call("pme_batch_set_chains", self.h, int(nchains))

    def get_chains(self) -> int:
        v = C.c_int()
        L.call("pme_batch_get_chains", self.h, C.byref(v))
        return v.value

    def set_half_bricks(self, on: bool) -> None:
        """Half-size spread/gather bricks, two workgroups per CU (pme_batch_set_half_bricks)."""
        L.call("pme_batch_set_half_bricks", self.h, 1 if on else 0)

    def get_half_bricks(self) -> bool:
        v = C.c_int()
        L.call("pme_batch_get_half_bricks", self.h, C.byref(v))
        return bool(v.value)

    def autotune(self, cache_path=None) -> float:
        ms = C.c_float()
        L.call("pme_batch_autotune", self.h, None if cache_path is None else str(cache_path).encode(), C.byref(ms))
        return ms.value

    def set_timing(self, enable: bool) -> None:
        L.call("pme_batch_set_timing", self.h, int(bool(enable)))

    def get_timing(self):
        ms = (C.c_float * 8)()
        L.call("pme_batch_get_timing", self.h, ms)
        return list(ms)

    def memory(self):
        u, b = C.c_size_t(), C.c_size_t()
        L.call("pme_batch_memory", self.h, C.byref(u), C.byref(b))
        return u.value, b.value


class PmeServer:
    """Continuous-batching serving loop (pme_server_* in the C ABI).

    submit() queues a solve of system `key` and returns a ticket at once;
    wait(ticket) blocks for (energy, virial[9], latency_ms).  The caller keeps
    the descriptor's device buffers alive until the request completes.
    """

    def __init__(self, hbm_budget_bytes: int, nlanes: int = 2):
        h = C.c_void_p()
        L.call("pme_server_create", C.c_size_t(int(hbm_budget_bytes)), int(nlanes), C.byref(h))
        self.h = h
        self._keep = {}  # ticket -> buffers of the request

    def close(self) -> None:
        if getattr(self, "h", None) is not None and L._lib is not None:
            L._lib.pme_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        self.close()

    def submit(self, key: int, desc, keep=(), ready_event=None) -> int:
        t = C.c_ulonglong()
        L.call("pme_server_submit", self.h, int(key), C.byref(desc), ready_event, C.byref(t))
        if keep:
            self._keep[t.value] = tuple(keep)
        return t.value

    def submit_header(self, key: int, header_bytes: bytes, nfft, order: int, kappa: float, xyzq, force3,
                      keep=()) -> int:
        """Serve one direct->recip request: header_bytes is the wire header
        (three doubles, an int, two bool bytes; DomdecRecipComm.h:11-15),
        xyzq a device float4 array, force3 a device float3 array.  Computes
        what CudaDomdecRecip::calc does: the energy and virial when either
        flag is set and the self energy when calc_energy is (wait_result).
        A STOP header (ncoord < 0) raises PmeError with status
        L.PME_ERR_STOPPED."""
        h = L.RecipHeader.from_buffer_copy(header_bytes.ljust(C.sizeof(L.RecipHeader), b"\0"))
        t = C.c_ulonglong()
        L.call("pme_server_submit_header", self.h, int(key), C.byref(h), int(nfft[0]), int(nfft[1]), int(nfft[2]),
               int(order), float(kappa), _ptr(xyzq), _ptr(force3), C.byref(t))
        self._keep[t.value] = tuple(keep) + (xyzq, force3)
        return t.value

    def wait(self, ticket: int):
        e, ms = C.c_double(), C.c_double()
        v = (C.c_double * 9)()
        try:
            L.call("pme_server_wait", self.h, C.c_ulonglong(ticket), C.byref(e), v, C.byref(ms))
        finally:
            self._keep.pop(ticket, None)
        return e.value, list(v), ms.value

    def wait_result(self, ticket: int) -> dict:
        """Block for the request: energy, energy_self, virial (9), latency_ms."""
        r = L.RequestResult()
        try:
            L.call("pme_server_wait_result", self.h, C.c_ulonglong(ticket), C.byref(r))
        finally:
            self._keep.pop(ticket, None)
        return {"energy": r.energy, "energy_self": r.energy_self, "virial": list(r.virial),
                "latency_ms": r.latency_ms}

    def poll(self, ticket: int):
        """None while in progress, else (energy, virial, latency_ms)."""
        d, e, ms = C.c_int(), C.c_double(), C.c_double()
        v = (C.c_double * 9)()
        L.call("pme_server_poll", self.h, C.c_ulonglong(ticket), C.byref(d), C.byref(e), v, C.byref(ms))
        if not d.value:
            return None
        self._keep.pop(ticket, None)
        return e.value, list(v), ms.value

    def drain(self) -> None:
        L.call("pme_server_drain", self.h)

    def release(self, key: int) -> None:
        L.call("pme_server_release", self.h, int(key))

    def stats(self) -> dict:
        st = L.ServerStats()
        L.call("pme_server_stats_get", self.h, C.byref(st))
        return st.as_dict()

    def reset_stats(self) -> None:
        L.call("pme_server_stats_reset", self.h)

    def set_tuning(self, brick_target: int, fft_tile: int, wg_per_cu: int) -> None:
        L.call("pme_server_set_tuning", self.h, int(brick_target), int(fft_tile), int(wg_per_cu))
