"""GPU tests of the split-precision FFT passes: the radix-16 DFT stage of the
y and z passes on the matrix cores as BF16 or FP16 hi-lo products with FP32
accumulation.  Spectra against the FP64 oracle's FFT, full solves of the
250,002-atom 144^3 workload against the FP64 oracle (target: force
relative RMS <= 1e-4), and the batched path against the plain one."""
import numpy as np
import pytest

import oracle as O
from gpu_recip_amd import _lib as L
from gpu_recip_amd import systems
from gpu_recip_amd.pme import CudaEnergyVirial, CudaPMERecip, PmeBatch, force_stride
from gpu_util import rel_err, rel_fro, rel_rms, say, to_dev, torch_dev

pytestmark = pytest.mark.gpu

# spectrum tolerance per mode (relative RMS vs the FP64 FFT of the same grid)
SPEC_TOL = {L.PREC_FP32: 3e-6, L.PREC_BF16X3: 5e-5, L.PREC_FP16X3: 5e-6}


def _solve(pme, ev, s, x, f, stride):
    ev.clear()
    pme.spread_charge(x, s.natom, s.recip)
    pme.r2c_fft()
    pme.scalar_sum(s.recip, s.kappa, True, True)
    pme.c2r_fft()
    pme.gather_force(x, s.natom, s.recip, stride, f, L.FORCE_FLOAT)
    ev.copyToHost()


def _forces(f, n, stride):
    fh = f.cpu().numpy()
    return np.stack([fh[:n], fh[stride:stride + n], fh[2 * stride:2 * stride + n]], 1).astype(np.float64)


@pytest.mark.parametrize("prec", [L.PREC_BF16X3, L.PREC_FP16X3])
@pytest.mark.parametrize("nfft", [(48, 64, 80), (36, 144, 32), (40, 48, 16), (30, 30, 30)])
def test_split_spectrum_and_roundtrip(nfft, prec):
    torch, dev = torch_dev()
    box = tuple(0.9 * n for n in nfft)
    s = systems.PmeSystem("t", systems.triplets(1500, box, 17), box, nfft, 4, 0.34)
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(*nfft, 4, ev)
    pme.set_precision(prec)
    x = to_dev(s.xyzq)
    nx, ny, nz = nfft
    g = torch.empty(nx * ny * nz, dtype=torch.float32, device=dev)
    c = torch.empty((nx // 2 + 1) * ny * nz * 2, dtype=torch.float32, device=dev)
    pme.spread_charge(x, s.natom, s.recip)
    pme.copy_grid(L.GRID_REAL, g)
    pme.r2c_fft()
    pme.copy_grid(L.GRID_SPECTRUM, c)
    torch.cuda.synchronize()
    grid = g.cpu().numpy().astype(np.float64).reshape(nz, ny, nx)
    spec = c.cpu().numpy().astype(np.float64).view(np.complex128).reshape(nz, ny, nx // 2 + 1)
    ref = O.fft3d_r2c(grid)
    err = rel_rms(spec, ref)
    split_axes = sum(1 for n in (ny, nz) if n % 16 == 0)
    say(f"  {nfft} prec {prec}: r2c rel err {err:.2e} ({split_axes} axes on the matrix cores)")
    assert err < (SPEC_TOL[prec] if split_axes else SPEC_TOL[L.PREC_FP32])
    pme.c2r_fft()
    pme.copy_grid(L.GRID_REAL, g)
    torch.cuda.synchronize()
    back = g.cpu().numpy().astype(np.float64).reshape(nz, ny, nx)
    assert rel_rms(back, grid * (nx * ny * nz)) < 2 * (SPEC_TOL[prec] if split_axes else SPEC_TOL[L.PREC_FP32])


def test_split_precision_workload_vs_fp64():
    """250,002 atoms, 144^3, order 4, kappa 0.30: FP32 vs BF16x3 vs FP16x3."""
    torch, dev = torch_dev()
    s = systems.config_split_precision()
    n = s.natom
    o = O.solve(O.OracleParams(*s.nfft, s.order, s.kappa, s.box, index_mode=O.INDEX_F64), s.xyzq)
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(*s.nfft, s.order, ev)
    x = to_dev(s.xyzq)
    stride = force_stride(n)
    f = torch.zeros(3 * stride, dtype=torch.float32, device=dev)
    res = {}
    for prec in (L.PREC_FP32, L.PREC_BF16X3, L.PREC_FP16X3):
        pme.set_precision(prec)
        _solve(pme, ev, s, x, f, stride)
        torch.cuda.synchronize()
        fr = _forces(f, n, stride)
        res[prec] = (rel_rms(fr, o["force"]), rel_err(ev.getEnergy("recip"), o["energy"]),
                     rel_fro(ev.getVirial(), o["virial"]))
        say(f"  prec {prec}: force rel-RMS {res[prec][0]:.2e} energy rel {res[prec][1]:.2e} virial {res[prec][2]:.2e}")
    for prec in (L.PREC_BF16X3, L.PREC_FP16X3):
        assert res[prec][0] <= 1e-4 and res[prec][1] <= 1e-4 and res[prec][2] <= 1e-4
    assert res[L.PREC_FP16X3][0] < 2 * res[L.PREC_FP32][0] + 1e-6


def test_split_batched_matches_plain():
    torch, dev = torch_dev()
    nfft, box = (40, 48, 64), (36.0, 43.0, 57.0)
    s = systems.PmeSystem("t", systems.triplets(4000, box, 23), box, nfft, 6, 0.34)
    n = s.natom
    stride = force_stride(n)
    x = to_dev(s.xyzq)
    b = PmeBatch(2 << 30)
    fb = {}
    ids = []
    for prec in (L.PREC_BF16X3, L.PREC_FP16X3):
        fb[prec] = torch.zeros(3 * stride, dtype=torch.float32, device=dev)
        ids.append(b.add(PmeBatch.desc(nfft, 6, s.kappa, s.recip, x, n, fb[prec], L.FORCE_FLOAT, stride,
                                       precision=prec), keep=(x, fb[prec])))
    b.solve()
    e, _ = b.results(ids)
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(*nfft, 6, ev)
    f = torch.zeros(3 * stride, dtype=torch.float32, device=dev)
    for k, prec in enumerate((L.PREC_BF16X3, L.PREC_FP16X3)):
        pme.set_precision(prec)
        _solve(pme, ev, s, x, f, stride)
        torch.cuda.synchronize()
        assert np.array_equal(_forces(fb[prec], n, stride), _forces(f, n, stride)), prec
        assert rel_err(e[k], ev.getEnergy("recip")) < 1e-12


def test_set_precision_checks():
    torch, dev = torch_dev()
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(32, 32, 32, 4, ev)
    with pytest.raises(L.PmeError) as ei:
        pme.set_precision(3)
    assert ei.value.status == 1
    s = systems.PmeSystem("t", systems.triplets(100, (30.0,) * 3, 1), (30.0,) * 3, (32, 32, 32), 4, 0.34)
    x = to_dev(s.xyzq)
    pme.spread_charge(x, s.natom, s.recip)
    pme.r2c_fft()
    with pytest.raises(L.PmeError) as ei:
        pme.set_precision(L.PREC_BF16X3)  # the forward transform is in flight
    assert ei.value.status == 6
