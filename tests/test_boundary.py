"""CPU tests of the C-ABI boundary: the library loads, exports every symbol the
header declares, and the host-only entry points behave (no GPU work here)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from gpu_recip_amd import _lib as L
from gpu_recip_amd import systems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(L.LIB_PATH):
        from gpu_recip_amd import build
        build.build()
    return L.load()


def test_exports_every_declared_symbol(lib):
    declared = L.header_symbols()
    assert len(declared) >= 30
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    # and the dynamic symbol table really exports them (extern "C", unmangled)
    nm = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in nm.splitlines() if line.strip()}
    assert not [s for s in declared if s not in exported]
    # the ctypes table binds exactly the declared set
    assert sorted(L.SIGNATURES) == declared


def test_no_oracle_in_product():
    # the product library and package never reference the oracle
    nm = subprocess.run(["nm", "-D", L.LIB_PATH], capture_output=True, text=True).stdout
    assert "orc_" not in nm
    pkg = os.path.join(ROOT, "gpu_recip_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                with open(os.path.join(dirpath, f)) as fh:
                    txt = fh.read()
                assert "import oracle" not in txt and "from oracle" not in txt and "liboracle" not in txt, f


def test_status_strings(lib):
    assert lib.pme_status_string(0) == b"ok"
    assert lib.pme_status_string(6) == b"stage called out of order"


@pytest.mark.parametrize("order,n", [(4, 48), (6, 100), (8, 45), (4, 64), (6, 108)])
def test_bsp_mod_host_matches_oracle(lib, order, n):
    out = (C.c_double * n)()
    assert lib.pme_bsp_mod(order, n, out) == 0
    assert np.allclose(np.array(out), O.bsp_mod(order, n), rtol=1e-13)


def _create(lib, nx=48, ny=48, nz=48, order=4, fft=2, nnode=1, accum=0, ev=1):
    h = C.c_void_p()
    evp = C.c_void_p(ev) if ev else None
    return lib.pme_recip_create(nx, ny, nz, order, fft, nnode, 0, accum, evp, b"recip", b"self", None, C.byref(h))


@pytest.mark.parametrize("kw,status", [
    (dict(order=5), 2), (dict(order=7), 2), (dict(order=3), 2),
    (dict(nnode=2), 2),
    (dict(nx=77), 2),            # 7*11: not 2/3/5-smooth
    (dict(ny=1024), 2),          # above the 512 edge limit
    (dict(nz=4, order=6), 2),    # grid edge below the order
    (dict(accum=3), 1),
    (dict(fft=7), 1),
    (dict(ev=0), 1),
])
def test_create_rejects_before_touching_the_device(lib, kw, status):
    # validation happens before any HIP call, so this runs without a GPU
    assert _create(lib, **kw) == status
    assert lib.pme_last_error()


def test_null_arguments_rejected(lib):
    assert lib.pme_recip_spread_charge(None, None, 0, None) == 1
    assert lib.pme_recip_r2c_fft(None) == 1
    assert lib.pme_recip_gather_force(None, None, 0, None, 0, 0, None) == 1
    assert lib.pme_ev_create(None) == 1
    assert lib.pme_bsp_mod(5, 1, None) == 1


def test_triplet_generator():
    x = systems.triplets(1000, (30.0, 40.0, 50.0), 1)
    assert x.shape == (3000, 4) and x.dtype == np.float32
    assert abs(float(x[:, 3].astype(np.float64).sum())) < 1e-3
    t = x.reshape(1000, 3, 4)[:, :, :3].astype(np.float64)
    d1 = np.linalg.norm(t[:, 1] - t[:, 0], axis=1)
    d2 = np.linalg.norm(t[:, 2] - t[:, 0], axis=1)
    assert np.allclose(d1, 0.9572, atol=1e-5) and np.allclose(d2, 0.9572, atol=1e-5)
    cosang = ((t[:, 1] - t[:, 0]) * (t[:, 2] - t[:, 0])).sum(1) / (d1 * d2)
    assert np.allclose(np.degrees(np.arccos(cosang)), 104.52, atol=1e-3)
    # same seed, same system
    assert np.array_equal(x, systems.triplets(1000, (30.0, 40.0, 50.0), 1))


def test_batched_config_grids():
    sys64 = systems.config_batched()
    assert len(sys64) == 64
    for s in sys64:
        assert 1998 <= s.natom <= 40002 and s.natom % 3 == 0
        for L_, n in zip(s.box, s.nfft):
            assert 30 <= L_ <= 75
            assert n >= L_ / 0.9 and systems.smooth235_at_least(L_ / 0.9) == n


def test_algorithmic_bytes():
    s = systems.config_production()
    assert s.natom == 96000
    assert s.algorithmic_bytes() == 44 * 96000 + 16 * 1036800 + 32 * 528000


def test_serving_schedule():
    from gpu_recip_amd import serving

    reqs = serving.schedule(10.0)
    assert 2700 < len(reqs) < 3300  # Poisson, 300/s
    t = np.array([r.t for r in reqs])
    assert np.all(np.diff(t) > 0) and t[-1] < 10.0
    fu = [r for r in reqs if r.followup]
    assert 0.15 < len(fu) / len(reqs) < 0.25
    for r in fu:
        b = reqs[r.base]
        assert not b.followup and b.key == r.key and b.t < r.t and (b.natom, b.nfft, b.order) == (r.natom, r.nfft, r.order)
    new = [r for r in reqs if not r.followup]
    assert len({r.key for r in new}) == len(new)
    n = np.array([r.natom for r in new])
    assert n.min() >= 4998 and n.max() <= 400002 and np.all(n % 3 == 0)
    assert 0.4 < np.mean([r.order == 4 for r in new]) < 0.6
    for r in new[:200]:
        assert abs(r.natom / r.box ** 3 - 0.1) < 1e-9 and r.nfft >= r.box and systems.smooth235_at_least(r.box) == r.nfft
    assert len(serving.schedule(1e9, max_requests=7)) == 7


def test_recip_header_layout():
    import struct
    # DomdecRecipComm.h:11-15: double x3, int, bool, bool (natural alignment: 32 bytes)
    assert C.sizeof(L.RecipHeader) == 32
    raw = struct.pack("<dddi??", 1 / 62.23, 1 / 62.0, 1 / 61.0, 23558, True, False)
    h = L.RecipHeader.from_buffer_copy(raw.ljust(32, b"\0"))
    assert (h.ncoord, h.calc_energy, h.calc_virial) == (23558, 1, 0)
    assert abs(h.inv_boxx - 1 / 62.23) < 1e-18
