"""GPU tests of the serving loop (pme_server_*): results are the same alone
and in the mix (bit for bit), equal the plain single-system path and agree
with the FP64 oracle; admission under a small HBM arena evicts idle systems
and follow-ups stay right; an open-loop run completes with latency stats."""
import numpy as np
import pytest

import oracle as O
from gpu_recip_amd import _lib as L
from gpu_recip_amd import serving
from gpu_recip_amd.pme import PmeServer
from gpu_util import rel_err, rel_rms, run_solve, say, to_dev, torch_dev

pytestmark = pytest.mark.gpu


def _forces(r):
    fh = r.force.cpu().numpy()
    n, st = r.natom, r.stride
    return np.stack([fh[:n], fh[st:st + n], fh[2 * st:2 * st + n]], 1).astype(np.float64)


def _small_schedule(n=24, nmax=60000, seed=11):
    reqs = serving.schedule(1e9, rate=300.0, nmin=3000, nmax=nmax, seed=seed, max_requests=n)
    return reqs


def test_alone_vs_mix():
    torch, dev = torch_dev()
    reqs = _small_schedule()
    assert any(r.followup for r in reqs) and len({r.order for r in reqs}) == 2
    serving.materialise(reqs, dev)
    srv = PmeServer(4 << 30, nlanes=2)
    # the mix: everything submitted at once
    serving.run_burst(srv, reqs)
    mix = [(_forces(r), r.result[0], r.result[1]) for r in reqs]
    st = srv.stats()
    say(f"  mix: {st['batches']} batches, mean batch {st['mean_batch']:.1f}, max {st['max_batch']}")
    assert st["completed"] == len(reqs) and st["max_batch"] > 1
    srv.close()
    # alone: a fresh server, one request at a time
    srv = PmeServer(4 << 30, nlanes=2)
    for i, r in enumerate(reqs):
        r.force.zero_()
        t = srv.submit(r.key, serving.desc_of(r))
        e, v, _ = srv.wait(t)
        f = _forces(r)
        assert np.array_equal(f, mix[i][0]), i
        assert e == mix[i][1] and v == mix[i][2], i
    srv.close()


def test_empty_request_in_a_burst():
    """A request with no atoms (null pointers) served in a burst with others:
    zero energy and virial for it, the others unchanged against serving them
    without it."""
    torch, dev = torch_dev()
    reqs = _small_schedule(n=6, nmax=20000, seed=9)
    serving.materialise(reqs, dev)
    srv = PmeServer(4 << 30, nlanes=2)
    serving.run_burst(srv, reqs)
    ref = [(_forces(r), r.result[0], r.result[1]) for r in reqs]
    srv.close()
    srv = PmeServer(4 << 30, nlanes=2)
    r0 = reqs[0]
    empty = L.SystemDesc()
    empty.nfftx = empty.nffty = empty.nfftz = r0.nfft
    empty.order = r0.order
    empty.kappa = r0.kappa
    for i in range(9):
        empty.recip[i] = (1.0 / r0.box) if i in (0, 4, 8) else 0.0
    empty.ncoord = 0
    empty.force_type = L.FORCE_FLOAT
    empty.calc_energy_virial = 1
    for r in reqs:
        r.force.zero_()
    tickets = [srv.submit(10 ** 6, empty)] + [srv.submit(r.key, serving.desc_of(r)) for r in reqs]
    srv.drain()
    e, v, _ = srv.wait(tickets[0])
    assert e == 0.0 and all(x == 0.0 for x in v)
    for r, t, (f, e0, v0) in zip(reqs, tickets[1:], ref):
        e1, v1, _ = srv.wait(t)
        assert np.array_equal(_forces(r), f) and e1 == e0 and v1 == v0
    srv.close()


def test_served_equals_plain_and_oracle():
    torch, dev = torch_dev()
    reqs = _small_schedule(n=10, nmax=30000, seed=5)
    serving.materialise(reqs, dev)
    srv = PmeServer(4 << 30, nlanes=3)
    serving.run_burst(srv, reqs)
    for r in reqs:
        xyzq = r.xyzq.cpu().numpy()
        box = (r.box,) * 3
        one = run_solve((r.nfft,) * 3, r.order, r.kappa, box, xyzq)
        f = _forces(r)
        assert np.array_equal(f, one["force"])
        assert rel_err(r.result[0], one["energy"]) < 1e-12
        if r.natom <= 15000:
            o = O.solve(O.OracleParams(r.nfft, r.nfft, r.nfft, r.order, r.kappa, box), xyzq)
            frms = rel_rms(f, o["force"])
            say(f"  key {r.key} n={r.natom} grid {r.nfft} p{r.order}: force rel-RMS vs FP64 {frms:.2e}")
            assert frms < 1e-4 and rel_err(r.result[0], o["energy"]) < 1e-4
    srv.close()


def test_eviction_under_small_arena():
    torch, dev = torch_dev()
    reqs = _small_schedule(n=30, nmax=40000, seed=3)
    serving.materialise(reqs, dev)
    srv = PmeServer(20 << 20, nlanes=2)  # room for a few systems only (the largest needs ~9 MB)
    serving.run_burst(srv, reqs)
    st = srv.stats()
    say(f"  small arena: evictions {st['evictions']}, new slots {st['admitted_new']}, resident {st['resident']}")
    assert st["completed"] == len(reqs) and st["evictions"] > 0
    assert st["arena_used"] <= st["arena_budget"]
    for r in reqs[-8:]:
        one = run_solve((r.nfft,) * 3, r.order, r.kappa, (r.box,) * 3, r.xyzq.cpu().numpy())
        assert np.array_equal(_forces(r), one["force"])
    srv.close()


def test_burst_into_a_full_arena():
    """A second burst of new systems into an arena full of idle residents
    evicts the least recently used ones (never a system of the batch being
    formed): every request completes with the plain solver's forces, and the
    arena stays within its budget."""
    torch, dev = torch_dev()
    first = _small_schedule(n=24, nmax=20000, seed=21)
    second = _small_schedule(n=24, nmax=20000, seed=22)
    for r in second:
        r.key += 1000  # new systems only
    serving.materialise(first, dev)
    serving.materialise(second, dev)
    srv = PmeServer(24 << 20, nlanes=2)
    serving.run_burst(srv, first)
    st1 = srv.stats()
    srv.reset_stats()
    serving.run_burst(srv, second)
    st2 = srv.stats()
    say(f"  full arena: first burst resident {st1['resident']}, second burst evictions {st2['evictions']}")
    assert st2["completed"] == len(second) and st2["failed"] == 0 and st2["evictions"] > 0
    assert st2["arena_used"] <= st2["arena_budget"]
    for r in second[::3]:
        one = run_solve((r.nfft,) * 3, r.order, r.kappa, (r.box,) * 3, r.xyzq.cpu().numpy())
        assert np.array_equal(_forces(r), one["force"])
    srv.close()


def test_rejects_and_release():
    torch, dev = torch_dev()
    reqs = _small_schedule(n=2, nmax=6000, seed=9)
    serving.materialise(reqs, dev)
    srv = PmeServer(2 << 20, nlanes=1)
    big = serving.ServeRequest(0.0, 99, 90000, 97.0, 100, 4, 0.34, False)
    serving.materialise([big], dev)
    with pytest.raises(L.PmeError) as ei:
        srv.submit(big.key, serving.desc_of(big))
    assert ei.value.status == 4  # PME_ERR_BUDGET: never fits the arena
    srv.close()
    srv = PmeServer(1 << 30, nlanes=1)
    r = reqs[0]
    t = srv.submit(r.key, serving.desc_of(r))
    srv.wait(t)
    assert srv.stats()["resident"] == 1
    srv.release(r.key)
    t = srv.submit(r.key + 1000, serving.desc_of(r))  # a different key: new slot
    srv.wait(t)
    st = srv.stats()
    assert st["resident"] == 1 and st["admitted_new"] == 2
    with pytest.raises(L.PmeError):
        srv.wait(t)  # a result is taken once
    srv.reset_stats()
    assert srv.stats()["completed"] == 0
    srv.close()


def test_open_loop_short():
    torch, dev = torch_dev()
    reqs = serving.schedule(1.0, rate=300.0, nmin=5000, nmax=100000, seed=21)
    serving.materialise(reqs, dev)
    srv = PmeServer(16 << 30, nlanes=2)
    wall, lag = serving.run_open_loop(srv, reqs)
    st = srv.stats()
    say(f"  open loop: {len(reqs)} requests in {wall:.2f} s, p50 {st['p50_ms']:.3f} ms, p99 {st['p99_ms']:.3f} ms, "
        f"mean batch {st['mean_batch']:.2f}, submit lag max {lag:.2f} ms")
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert all(r.result is not None and np.isfinite(r.result[0]) for r in reqs)
    srv.close()


def test_direct_recip_header_requests():
    """Requests in the direct->recip wire format (header + float4 coordinates
    in, float3 forces out) give the plain solver's float3 forces."""
    import struct

    torch, dev = torch_dev()
    from gpu_recip_amd import systems

    srv = PmeServer(1 << 30, nlanes=2)
    s = systems.config_plain_small()
    x = to_dev(s.xyzq)
    f3 = torch.zeros(3 * s.natom, dtype=torch.float32, device=dev)
    hdr = struct.pack("<dddi??", 1 / s.box[0], 1 / s.box[1], 1 / s.box[2], s.natom, True, True)
    t = srv.submit_header(7, hdr, s.nfft, s.order, s.kappa, x, f3)
    e, v, _ = srv.wait(t)
    one = run_solve(s.nfft, s.order, s.kappa, s.box, s.xyzq, force_type=L.FORCE_FLOAT3)
    assert np.array_equal(f3.cpu().numpy().reshape(-1, 3).astype(np.float64), one["force"])
    assert rel_err(e, one["energy"]) < 1e-12
    # forces only: energy/virial flags off
    hdr = struct.pack("<dddi??", 1 / s.box[0], 1 / s.box[1], 1 / s.box[2], s.natom, False, False)
    e, v, _ = srv.wait(srv.submit_header(7, hdr, s.nfft, s.order, s.kappa, x, f3))
    assert e == 0.0 and v == [0.0] * 9
    # CudaDomdecRecip::calc adds the self energy when calc_energy is set
    one = run_solve(s.nfft, s.order, s.kappa, s.box, s.xyzq, force_type=L.FORCE_FLOAT3, self_energy=True)
    hdr = struct.pack("<dddi??", 1 / s.box[0], 1 / s.box[1], 1 / s.box[2], s.natom, True, False)
    r = srv.wait_result(srv.submit_header(7, hdr, s.nfft, s.order, s.kappa, x, f3))
    assert rel_err(r["energy"], one["energy"]) < 1e-12
    assert rel_err(r["energy_self"], one["self"]) < 1e-12
    assert rel_err(r["energy_self"], O.self_energy(s.xyzq, s.kappa)) < 1e-6
    hdr = struct.pack("<dddi??", 1 / s.box[0], 1 / s.box[1], 1 / s.box[2], s.natom, False, True)
    r = srv.wait_result(srv.submit_header(7, hdr, s.nfft, s.order, s.kappa, x, f3))
    assert r["energy_self"] == 0.0 and rel_err(r["virial"][0], one["virial"][0]) < 1e-12
    # the STOP message (DomdecRecipComm::send_stop: ncoord = -1) is refused
    hdr = struct.pack("<dddi??", 1.0, 1.0, 1.0, -1, False, False)
    with pytest.raises(L.PmeError) as ei:
        srv.submit_header(7, hdr, s.nfft, s.order, s.kappa, x, f3)
    assert ei.value.status == L.PME_ERR_STOPPED
    assert srv.stats()["submitted"] == 4
    srv.close()


def test_full_size_burst_with_largest_requests():
    """The serving workload's largest requests at full size: a 360,000-atom
    order-6 system (cubic 153 A box, 160^3 grid) and a follow-up step of it,
    served in one burst with smaller requests.  Each result equals the plain
    single-system path bit for bit, and the large one is checked against the
    FP64 oracle (relative-RMS force error reported)."""
    torch, dev = torch_dev()
    natom = 360000
    box = (natom / 0.1) ** (1.0 / 3.0)
    nfft = serving.systems.smooth235_at_least(box / 1.0)
    assert nfft >= 150
    big = serving.ServeRequest(0.0, 900, natom, box, nfft, 6, 0.34, False)
    follow = serving.ServeRequest(0.0, 900, natom, box, nfft, 6, 0.34, True, base=0)
    small = _small_schedule(n=6, nmax=60000, seed=21)
    for i, r in enumerate(small):
        r.key += 1000
        if r.followup:
            r.base += 2
    reqs = [big, follow] + small
    serving.materialise(reqs, dev)
    say(f"  burst: {len(reqs)} requests, largest {natom} atoms, grid {nfft}^3, order 6")
    srv = PmeServer(16 << 30, nlanes=2)
    serving.run_burst(srv, reqs)
    st = srv.stats()
    assert st["completed"] == len(reqs)
    srv.close()
    for r in reqs:
        xyzq = r.xyzq.cpu().numpy()
        one = run_solve((r.nfft,) * 3, r.order, r.kappa, (r.box,) * 3, xyzq)
        assert np.array_equal(_forces(r), one["force"]), r.key
        assert rel_err(r.result[0], one["energy"]) < 1e-12
    xyzq = big.xyzq.cpu().numpy()
    o = O.solve(O.OracleParams(nfft, nfft, nfft, 6, 0.34, (box,) * 3), xyzq)
    frms = rel_rms(_forces(big), o["force"])
    erel = rel_err(big.result[0], o["energy"])
    say(f"  {natom} atoms, {nfft}^3, order 6: force rel-RMS vs FP64 {frms:.2e}, energy rel {erel:.2e}")
    assert frms < 1e-4 and erel < 1e-4
