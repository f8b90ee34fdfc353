"""Python mirror of the reference's CudaPMEDirectForce<long long, float>
(CudaPMEDirectForce.h:82-204) over the C ABI: the direct-space (real-space)
Ewald term on the reference's tile neighbour list.

Device buffers are torch tensors (plumbing only); every compute call goes
through libpme_hip.so.  Forces are ADDED to an int64 tensor of 3 * stride
entries in 2^40 fixed point (x at [i], y at [i + stride], z at [i + 2 stride])
as the gradient +dE/dr; energies go to the three named energy terms.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .pme import _ptr, _stream_ptr

VDW_NONE, ELEC_NONE, ELEC_EWALD, ELEC_EWALD_LOOKUP = 0, 0, 101, 111
FORCE_SCALE = 2.0 ** 40


class DeviceTileList:
    """A nlist.TileList uploaded to the device (pme_nlist)."""

    def __init__(self, tl, device):
        import torch

        self.tl = tl
        self.ientry = torch.from_numpy(np.ascontiguousarray(tl.ientry, np.int32)).to(device)
        self.tile_indj = torch.from_numpy(np.ascontiguousarray(tl.tile_indj, np.int32)).to(device)
        self.tile_excl = torch.from_numpy(np.ascontiguousarray(tl.tile_excl, np.uint32).view(np.int32)).to(device)
        self.xyzq = torch.from_numpy(np.ascontiguousarray(tl.xyzq, np.float32)).to(device)
        self.c = L.Nlist()
        self.c.n_ientry = len(tl.ientry)
        self.c.ientry = _ptr(self.ientry)
        self.c.n_tile = len(tl.tile_indj)
        self.c.tile_indj = _ptr(self.tile_indj)
        self.c.tile_excl = _ptr(self.tile_excl)


class CudaPMEDirectForce:
    def __init__(self, energyVirial, nameVdw="vdw", nameElec="elec", nameExcl="excl"):
        h = C.c_void_p()
        L.call("pme_direct_create", energyVirial.h, nameVdw.encode(), nameElec.encode(), nameExcl.encode(),
               C.byref(h))
        self.h = h
        self.ev = energyVirial  # keep alive

    def __del__(self):
        if getattr(self, "h", None) is not None and L._lib is not None:
            L._lib.pme_direct_destroy(self.h)
            self.h = None

    def setup(self, boxx, boxy, boxz, kappa, roff, ron, e14fac, vdw_model=VDW_NONE, elec_model=ELEC_EWALD):
        L.call("pme_direct_setup", self.h, float(boxx), float(boxy), float(boxz), float(kappa), float(roff),
               float(ron), float(e14fac), int(vdw_model), int(elec_model))

    def get_setup(self):
        f = [C.c_float() for _ in range(7)]
        i = [C.c_int() for _ in range(2)]
        L.call("pme_direct_get_setup", self.h, *[C.byref(x) for x in f], *[C.byref(x) for x in i])
        return tuple(x.value for x in f) + tuple(x.value for x in i)

    def set_box_size(self, boxx, boxy, boxz):
        L.call("pme_direct_set_box_size", self.h, float(boxx), float(boxy), float(boxz))

    def set_calc_elec(self, on: bool):
        L.call("pme_direct_set_calc_elec", self.h, int(bool(on)))

    def set_calc_vdw(self, on: bool):
        L.call("pme_direct_set_calc_vdw", self.h, int(bool(on)))

    def set_14_list(self, in14, ex14, stream=None):
        """in14, ex14: (M, 3) int arrays of (i, j, ishift)."""
        a = np.ascontiguousarray(np.asarray(in14, np.int32).reshape(-1, 3))
        b = np.ascontiguousarray(np.asarray(ex14, np.int32).reshape(-1, 3))
        L.call("pme_direct_set_14_list", self.h, len(a), len(b), a.ctypes.data if len(a) else None,
               b.ctypes.data if len(b) else None, _stream_ptr(stream))

    def calc_14_force(self, xyzq, calc_energy, calc_virial, stride, force, stream=None):
        L.call("pme_direct_calc_14_force", self.h, _ptr(xyzq), int(bool(calc_energy)), int(bool(calc_virial)),
               int(stride), _ptr(force), _stream_ptr(stream))

    def calc_force(self, xyzq, nlist: DeviceTileList, calc_energy, calc_virial, stride, force, stream=None):
        L.call("pme_direct_calc_force", self.h, _ptr(xyzq), C.byref(nlist.c), int(bool(calc_energy)),
               int(bool(calc_virial)), int(stride), _ptr(force), _stream_ptr(stream))
