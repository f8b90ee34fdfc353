"""pme_recip_solve with pme_recip_set_graph(1) on a caller's stream runs as a
hipGraph: the first call with new arguments runs the stages, the second
captures them, later calls replay the graph.  Every replay must equal the five-call sequence bit for bit
(forces, energy, virial), also after the coordinates move in place, with
other work and synchronisation between replays, after the arguments change
and back, and with the allocation-free solve interleaved."""
import numpy as np
import pytest

from gpu_recip_amd import _lib as L
from gpu_recip_amd import systems
from gpu_recip_amd.pme import CudaEnergyVirial, CudaPMERecip, force_stride
from gpu_util import say, to_dev, torch_dev

pytestmark = pytest.mark.gpu


def _five_calls(x, n, nfft, order, kappa, recip, st):
    torch, dev = torch_dev()
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(*nfft, order, ev)
    f = torch.zeros(3 * st, dtype=torch.float32, device=dev)
    ev.clear()
    pme.spread_charge(x, n, recip)
    pme.r2c_fft()
    pme.scalar_sum(recip, kappa, True, True)
    pme.c2r_fft()
    pme.gather_force(x, n, recip, st, f, L.FORCE_FLOAT)
    ev.copyToHost()
    torch.cuda.synchronize()
    return f.cpu().numpy(), ev.getEnergy("recip"), np.array(ev.getVirial())


def test_graph_replays_equal_the_stages():
    torch, dev = torch_dev()
    s = systems.config_plain_small()
    n, nfft, order, kappa = s.natom, s.nfft, s.order, s.kappa
    st = force_stride(n)
    x = to_dev(s.xyzq)
    side = torch.cuda.Stream()
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(*nfft, order, ev)
    pme.set_stream(side)
    pme.set_graph(True)
    f = torch.zeros(3 * st, dtype=torch.float32, device=dev)

    def solve(recip, kap=kappa):
        ev.clear(side)
        pme.solve(x, n, recip, kap, True, True, st, f, L.FORCE_FLOAT)
        ev.copyToHost(side)
        side.synchronize()
        return f.cpu().numpy(), ev.getEnergy("recip"), np.array(ev.getVirial())

    def check(got, want, what):
        assert np.array_equal(got[0], want[0]), what
        assert got[1] == want[1] and np.array_equal(got[2], want[2]), what

    r0 = list(s.recip)
    want = _five_calls(x, n, nfft, order, kappa, r0, st)
    for k in range(5):  # stages, capture, then replays
        check(solve(r0), want, f"call {k}")
        torch.zeros(1000, device=dev).sum().item()  # unrelated eager work in between
    # coordinates moved in place: the replay reads them
    rng = np.random.default_rng(3)
    xm = s.xyzq.copy()
    xm[:, :3] += rng.uniform(-0.3, 0.3, (n, 3)).astype(np.float32)
    x.copy_(torch.from_numpy(xm).to(dev))
    torch.cuda.synchronize()
    want_m = _five_calls(x, n, nfft, order, kappa, r0, st)
    for k in range(3):
        check(solve(r0), want_m, f"moved, call {k}")
    # another box, then back
    r1 = list(r0)
    r1[0] *= 1.01
    want_1 = _five_calls(x, n, nfft, order, kappa, r1, st)
    for k in range(3):
        check(solve(r1), want_1, f"box 1, call {k}")
    for k in range(3):
        check(solve(r0), want_m, f"box 0 again, call {k}")
    # the allocation-free solve in between leaves the graph valid
    g = torch.zeros(3 * st, dtype=torch.float32, device=dev)
    ev.clear(side)
    pme.solve_noalloc(x, n, r0, kappa, True, True, st, g, L.FORCE_FLOAT)
    side.synchronize()
    assert np.array_equal(g.cpu().numpy(), want_m[0])
    for k in range(2):
        check(solve(r0), want_m, f"after noalloc, call {k}")
    # the stages can still be called one by one on the same handle
    ev.clear(side)
    pme.spread_charge(x, n, r0)
    pme.r2c_fft()
    pme.scalar_sum(r0, kappa, True, True)
    pme.c2r_fft()
    pme.gather_force(x, n, r0, st, f, L.FORCE_FLOAT)
    ev.copyToHost(side)
    side.synchronize()
    check((f.cpu().numpy(), ev.getEnergy("recip"), np.array(ev.getVirial())), want_m, "five calls on the handle")
    say(f"  {n} atoms, grid {nfft}: 19 solves through the graph path equal the five-call sequence")
