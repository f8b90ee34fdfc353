"""GPU tests of the rest of the drop-in boundary: the COLUMN and SLAB handle
types, set_stream, clearEnergy / clearVirial, the device energy and virial
pointers and the getSforce layout (CudaPMERecip.h:166-203,
CudaEnergyVirial.h:13-52, CudaEnergyVirial.cu:172-278)."""
import ctypes as C

import numpy as np
import pytest

from gpu_recip_amd import _lib as L
from gpu_recip_amd import systems
from gpu_recip_amd.pme import CudaEnergyVirial, CudaPMERecip, force_stride

from gpu_util import say, to_dev, torch_dev

pytestmark = pytest.mark.gpu

# the device blob (Virial_t, CudaEnergyVirial.h:13-18)
OFF_SFORCE_FP = 81 * 8
OFF_VIRMAT = OFF_SFORCE_FP + 81 * 8
OFF_ENERGY = OFF_VIRMAT + 9 * 8

_hip = None


def _hip_lib():
    """The HIP runtime this process already runs on (the copy torch and
    libpme_hip.so loaded: dlopen of the same path returns that handle)."""
    global _hip
    if _hip is None:
        path = "libamdhip64.so"
        with open("/proc/self/maps") as m:
            for line in m:
                if "libamdhip64.so" in line:
                    path = line.split()[-1]
                    break
        _hip = C.CDLL(path)
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpy.restype = C.c_int
    return _hip


def dev_read(ptr, nbytes):
    """nbytes of device memory at ptr (hipMemcpy device to host)."""
    buf = C.create_string_buffer(nbytes)
    assert _hip_lib().hipMemcpy(buf, C.c_void_p(ptr), nbytes, 2) == 0
    return buf.raw


def dev_write(ptr, data: bytes):
    assert _hip_lib().hipMemcpy(C.c_void_p(ptr), data, len(data), 1) == 0


def _system():
    s = systems.config_plain_small()  # 8,001 atoms, 48^3, order 4
    return s, to_dev(s.xyzq)


def _solve(pme, ev, s, x, f, stream=None, calc=True):
    ev.clear(stream)
    pme.spread_charge(x, s.natom, s.recip)
    pme.r2c_fft()
    pme.scalar_sum(s.recip, s.kappa, calc, calc)
    pme.c2r_fft()
    pme.gather_force(x, s.natom, s.recip, force_stride(s.natom), f, L.FORCE_FLOAT)
    ev.copyToHost(stream)


def test_column_and_slab_equal_box():
    """Every FFTtype runs the same single-GPU pipeline: forces, energy and
    virial equal bit for bit (fft_type, CudaPMERecip.cu:1433-1470)."""
    torch, dev = torch_dev()
    s, x = _system()
    st = force_stride(s.natom)
    out = {}
    for ft, name in ((L.FFT_BOX, "box"), (L.FFT_COLUMN, "column"), (L.FFT_SLAB, "slab")):
        say(f"  fft_type {name}")
        ev = CudaEnergyVirial()
        pme = CudaPMERecip(*s.nfft, s.order, ev, fft_type=ft)
        f = torch.zeros(3 * st, dtype=torch.float32, device=dev)
        _solve(pme, ev, s, x, f)
        torch.cuda.synchronize()
        out[name] = (f.cpu().numpy(), ev.getEnergy("recip"), np.array(ev.getVirial()))
    for name in ("column", "slab"):
        assert np.array_equal(out[name][0], out["box"][0]), name
        assert out[name][1] == out["box"][1], name
        assert np.array_equal(out[name][2], out["box"][2]), name


def test_set_stream_switch():
    """Solve on a side stream, switch to another stream and back to the
    null stream: every solve gives the same results (set_stream,
    CudaPMERecip.cu:1618-1638)."""
    torch, dev = torch_dev()
    s, x = _system()
    st = force_stride(s.natom)
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(*s.nfft, s.order, ev)
    f0 = torch.zeros(3 * st, dtype=torch.float32, device=dev)
    _solve(pme, ev, s, x, f0)
    torch.cuda.synchronize()
    e0, v0 = ev.getEnergy("recip"), np.array(ev.getVirial())
    side = [torch.cuda.Stream(), torch.cuda.Stream()]
    for i, strm in enumerate(side + [None]):
        say(f"  solve on stream {i}")
        pme.set_stream(strm)
        f = torch.zeros(3 * st, dtype=torch.float32, device=dev)
        torch.cuda.synchronize()  # the zeroed buffer is ready before the side stream reads it
        _solve(pme, ev, s, x, f, stream=strm)
        if strm is not None:
            strm.synchronize()
        torch.cuda.synchronize()
        assert np.array_equal(f.cpu().numpy(), f0.cpu().numpy())
        assert ev.getEnergy("recip") == e0
        assert np.array_equal(np.array(ev.getVirial()), v0)


def test_clear_energy_and_clear_virial():
    """clearEnergy zeroes only the energy terms, clearVirial only the
    sforce and virmat part; energies and virial accumulate until cleared
    (CudaEnergyVirial.cu:172-198)."""
    torch, dev = torch_dev()
    s, x = _system()
    st = force_stride(s.natom)
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(*s.nfft, s.order, ev)
    f = torch.zeros(3 * st, dtype=torch.float32, device=dev)
    _solve(pme, ev, s, x, f)
    torch.cuda.synchronize()
    e1, v1 = ev.getEnergy("recip"), np.array(ev.getVirial())
    assert e1 != 0.0 and np.abs(v1).max() > 0.0
    ev.clearEnergy()
    ev.copyToHost()
    torch.cuda.synchronize()
    assert ev.getEnergy("recip") == 0.0 and ev.getEnergy("self") == 0.0
    assert np.array_equal(np.array(ev.getVirial()), v1)
    # a second solve without clear adds its energy to 0 and its virial to v1
    pme.spread_charge(x, s.natom, s.recip)
    pme.r2c_fft()
    pme.scalar_sum(s.recip, s.kappa, True, True)
    pme.c2r_fft()
    pme.gather_force(x, s.natom, s.recip, st, f, L.FORCE_FLOAT)
    ev.copyToHost()
    torch.cuda.synchronize()
    assert ev.getEnergy("recip") == e1
    assert np.allclose(np.array(ev.getVirial()), 2 * v1, rtol=1e-14, atol=0)
    ev.clearVirial()
    ev.copyToHost()
    torch.cuda.synchronize()
    assert ev.getEnergy("recip") == e1
    assert not np.any(np.array(ev.getVirial()))
    assert not np.any(np.array(ev.getSforce()))


def test_device_pointers_and_sforce_layout():
    """getEnergyPointer / getVirialPointer address the device blob the way
    the reference's do (the Virial_t start and the energy slot), and
    getSforce combines sforce_dp + sforce_fp / 2^30 (CudaEnergyVirial.cu:231-278)."""
    torch, dev = torch_dev()
    s, x = _system()
    st = force_stride(s.natom)
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(*s.nfft, s.order, ev)
    f = torch.zeros(3 * st, dtype=torch.float32, device=dev)
    pme.calc_self_energy(x, s.natom, s.kappa)
    _solve(pme, ev, s, x, f)
    pme.calc_self_energy(x, s.natom, s.kappa)
    ev.copyToHost()
    torch.cuda.synchronize()
    e_rec, e_self, vir = ev.getEnergy("recip"), ev.getEnergy("self"), np.array(ev.getVirial())
    vp = ev.getVirialPointer()
    ep_rec, ep_self = ev.getEnergyPointer("recip"), ev.getEnergyPointer("self")
    torch.cuda.synchronize()
    assert ep_rec == vp + OFF_ENERGY + 8 * 0  # "recip" was inserted first
    assert ep_self == vp + OFF_ENERGY + 8 * 1
    blob = dev_read(vp, OFF_ENERGY + 16)
    assert np.array_equal(np.frombuffer(blob[OFF_VIRMAT:OFF_ENERGY], np.float64), vir)
    assert np.frombuffer(dev_read(ep_rec, 8), np.float64)[0] == e_rec
    assert np.frombuffer(dev_read(ep_self, 8), np.float64)[0] == e_self
    assert e_self < 0.0
    # sforce: the reciprocal path leaves it at zero; values another force
    # term writes through the Virial_t pointer come back through getSforce
    assert not np.any(np.frombuffer(blob[:OFF_VIRMAT], np.uint8))
    dp = np.arange(81, dtype=np.float64) * 0.25 - 3.0
    fp = (np.arange(81, dtype=np.int64) - 40) * (1 << 28)
    dev_write(vp, dp.tobytes() + fp.tobytes())
    ev.copyToHost()
    torch.cuda.synchronize()
    assert np.array_equal(np.array(ev.getSforce()), dp + fp.astype(np.float64) / 2.0 ** 30)
    assert np.array_equal(np.array(ev.getVirial()), vir)
