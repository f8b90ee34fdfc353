"""GPU test of pme_recip_solve_noalloc, the solve whose work buffers live in
the library's static device scratch: its forces, energy and virial equal
pme_recip_solve's bit for bit on two calls in a row (the second with fewer
atoms), and a solve too large for the scratch is refused before any device
work."""
import numpy as np
import pytest

from gpu_recip_amd import _lib as L
from gpu_recip_amd import systems
from gpu_recip_amd.pme import CudaEnergyVirial, CudaPMERecip, force_stride
from gpu_util import say, to_dev, torch_dev

pytestmark = pytest.mark.gpu


def _run(pme, ev, fn, x, n, s, f):
    ev.clear()
    fn(x, n, s.recip, s.kappa, True, True, force_stride(n), f, L.FORCE_FLOAT)
    ev.copyToHost()
    import torch

    torch.cuda.synchronize()
    return f.cpu().numpy().copy(), ev.getEnergy("recip"), np.array(ev.getVirial())


def test_noalloc_equals_solve_twice_then_smaller():
    torch, dev = torch_dev()
    for s in (systems.config_production(), systems.config_plain_small()):
        say(f"  {s.name}: natom={s.natom} nfft={s.nfft} order={s.order}")
        ev_a, ev_b = CudaEnergyVirial(), CudaEnergyVirial()
        pa = CudaPMERecip(*s.nfft, s.order, ev_a)
        pb = CudaPMERecip(*s.nfft, s.order, ev_b)
        x = to_dev(s.xyzq)
        for n in (s.natom, s.natom // 3):  # the second call with a smaller input
            st = force_stride(n)
            fa = torch.zeros(3 * st, dtype=torch.float32, device=dev)
            fb = torch.full((3 * st,), 7.0, dtype=torch.float32, device=dev)
            ra = _run(pa, ev_a, pa.solve, x, n, s, fa)
            rb = _run(pb, ev_b, pb.solve_noalloc, x, n, s, fb)
            say(f"    n={n}: energy {ra[1]:.9f} / {rb[1]:.9f}")
            assert np.array_equal(ra[0][:n], rb[0][:n])
            assert np.array_equal(ra[0][st:st + n], rb[0][st:st + n])
            assert np.array_equal(ra[0][2 * st:2 * st + n], rb[0][2 * st:2 * st + n])
            assert ra[1] == rb[1]
            assert np.array_equal(ra[2], rb[2])
        # the handle's own five-call path still works after the static solves
        f5 = torch.zeros(3 * force_stride(s.natom), dtype=torch.float32, device=dev)
        r5 = _run(pb, ev_b, pb.solve, x, s.natom, s, f5)
        r0 = _run(pa, ev_a, pa.solve, x, s.natom, s, torch.zeros_like(f5))
        assert np.array_equal(r5[0], r0[0]) and r5[1] == r0[1]


def test_noalloc_refuses_beyond_scratch():
    torch, dev = torch_dev()
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(256, 256, 256, 4, ev)  # 64 MiB of real grid alone
    x = torch.zeros((1000, 4), dtype=torch.float32, device=dev)
    f = torch.zeros(3 * force_stride(1000), dtype=torch.float32, device=dev)
    recip = [1 / 200.0, 0, 0, 0, 1 / 200.0, 0, 0, 0, 1 / 200.0]
    with pytest.raises(L.PmeError) as e:
        pme.solve_noalloc(x, 1000, recip, 0.34, True, True, force_stride(1000), f)
    assert e.value.status == 4  # PME_ERR_BUDGET
    assert not torch.any(f).item()
