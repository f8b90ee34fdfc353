"""GPU tests of the batched path (pme_batch_*): every system of a batch gives
the single-system solver's result (same kernels: forces bit-identical, energy
and virial to FP64 summation order) and agrees with the FP64 oracle; hipGraph
replays match eager solves; admission honours the HBM budget; subsets,
updates, removals and the autotuner keep results unchanged."""
import json
import math

import numpy as np
import pytest

import oracle as O
from gpu_recip_amd import _lib as L
from gpu_recip_amd import systems
from gpu_recip_amd.pme import PmeBatch, force_stride
from gpu_util import rel_err, rel_fro, rel_rms, run_solve, say, to_dev, torch_dev

pytestmark = pytest.mark.gpu


def _mixed_systems():
    """Heterogeneous set: orders 4/6/8, odd and even grids, 1 to 30k atoms."""
    out = []
    specs = [
        (4, (48, 48, 48), (40.0, 40.0, 40.0), 2667),
        (6, (45, 50, 54), (42.0, 47.0, 50.0), 3000),
        (8, (36, 40, 32), (33.0, 37.0, 30.0), 1500),
        (4, (60, 64, 72), (55.0, 60.0, 66.0), 10000),
        (6, (25, 27, 30), (22.0, 24.0, 27.0), 1),
        (4, (80, 75, 81), (72.0, 68.0, 74.0), 5),
    ]
    for i, (order, nfft, box, ntrip) in enumerate(specs):
        x = systems.triplets(ntrip, box, systems.SEED_BASE + 500 + i)
        out.append(systems.PmeSystem(f"mix{i}", x, box, nfft, order, 0.34))
    return out


def _add(b, s, torch, dev, force_type=L.FORCE_FLOAT, calc_ev=True):
    n = s.natom
    x = to_dev(s.xyzq)
    stride = force_stride(n)
    if force_type == L.FORCE_INT64:
        f = torch.zeros(3 * stride, dtype=torch.int64, device=dev)
    elif force_type == L.FORCE_FLOAT3:
        f = torch.zeros(3 * n, dtype=torch.float32, device=dev)
    else:
        f = torch.zeros(3 * stride, dtype=torch.float32, device=dev)
    d = PmeBatch.desc(s.nfft, s.order, s.kappa, s.recip, x, n, f, force_type, stride, calc_ev)
    sid = b.add(d, keep=(x, f))
    return sid, x, f, stride


def _forces(f, n, stride, force_type=L.FORCE_FLOAT):
    fh = f.cpu().numpy()
    if force_type == L.FORCE_FLOAT3:
        return fh.reshape(n, 3).astype(np.float64)
    out = np.stack([fh[:n], fh[stride:stride + n], fh[2 * stride:2 * stride + n]], 1).astype(np.float64)
    return out / 2.0 ** 40 if force_type == L.FORCE_INT64 else out


@pytest.mark.parametrize("graph", [True, False])
def test_batch_matches_single_and_oracle(graph):
    torch, dev = torch_dev()
    sl = _mixed_systems()
    b = PmeBatch(8 << 30)
    b.set_graph(graph)
    reg = [_add(b, s, torch, dev) for s in sl]
    ids = [r[0] for r in reg]
    say(f"batch solve of {len(ids)} systems, graph={graph}")
    b.solve()
    e, v = b.results(ids)
    for s, (sid, x, f, stride), ek, vk in zip(sl, reg, e, v):
        fb = _forces(f, s.natom, stride)
        one = run_solve(s.nfft, s.order, s.kappa, s.box, s.xyzq)
        assert np.array_equal(fb, one["force"]), s.name
        assert rel_err(ek, one["energy"]) < 1e-12, s.name
        assert rel_fro(vk, one["virial"]) < 1e-12, s.name
        o = O.solve(O.OracleParams(*s.nfft, s.order, s.kappa, s.box, index_mode=O.INDEX_F32), s.xyzq)
        frms = rel_rms(fb, o["force"])
        say(f"  {s.name}: force rel-RMS vs oracle {frms:.2e} energy rel {rel_err(ek, o['energy']):.2e}")
        assert frms < 1e-5 and rel_err(ek, o["energy"]) < 1e-5 and rel_fro(vk, o["virial"]) < 1e-5


@pytest.mark.parametrize("reps", [4, 8])
def test_plane_groups_match_single(reps):
    """A batch with g z planes per item for at least 2 items per CU runs its
    xy-plane FFTs g planes per workgroup (g up to 4, as many as fit LDS; on a
    256-CU GPU 4 copies of the mixed set give g = 2, 8 copies g = 4): odd and
    even nx, nz not a multiple of g (a shorter last group), orders 4/6/8.
    Forces stay bit-identical to the single-system solver, which transforms
    one plane per workgroup."""
    torch, dev = torch_dev()
    base = _mixed_systems()
    sl = []
    for rep in range(reps):
        for i, s in enumerate(base):
            x = systems.triplets(max(1, s.natom // 3), s.box, systems.SEED_BASE + 700 + 10 * rep + i)
            sl.append(systems.PmeSystem(f"{s.name}r{rep}", x, s.box, s.nfft, s.order, 0.34))
    g = 2 if reps == 4 else 4
    assert 2 * 256 * g <= sum(s.nfft[2] for s in sl) < 2 * 256 * (g + 1)
    b = PmeBatch(8 << 30)
    reg = [_add(b, s, torch, dev) for s in sl]
    ids = [r[0] for r in reg]
    say(f"batch of {len(ids)} systems, {sum(s.nfft[2] for s in sl)} z planes")
    for graph in (True, False):
        b.set_graph(graph)
        b.solve()
        e, v = b.results(ids)
        for s, (sid, x, f, stride), ek, vk in zip(sl, reg, e, v):
            fb = _forces(f, s.natom, stride)
            one = run_solve(s.nfft, s.order, s.kappa, s.box, s.xyzq)
            assert np.array_equal(fb, one["force"]), (s.name, graph)
            assert rel_err(ek, one["energy"]) < 1e-12, s.name
            assert rel_fro(vk, one["virial"]) < 1e-12, s.name


def test_full_and_half_plane_workgroups():
    """A batch mixing planes that fit a half-size plane workgroup (two per CU)
    with planes that need a full-size one (120 x 100 and 100 x 96: 100 and 81 KB
    of LDS, above the 75 KB half budget; and an odd 81 x 80 plane, whose x
    lines pair two real rows) runs two plane launches per direction; every
    system's forces equal the single-system solver's bit for bit."""
    torch, dev = torch_dev()
    specs = [(4, (120, 100, 40), (110.0, 92.0, 36.0), 4000), (6, (100, 96, 36), (92.0, 88.0, 33.0), 3000),
             (4, (81, 80, 30), (74.0, 73.0, 27.0), 2000), (4, (48, 48, 48), (40.0, 40.0, 40.0), 2667),
             (8, (45, 36, 40), (41.0, 33.0, 37.0), 1500)]
    sl = [systems.PmeSystem(f"pl{i}", systems.triplets(nt, box, systems.SEED_BASE + 900 + i), box, nfft, order, 0.34)
          for i, (order, nfft, box, nt) in enumerate(specs)]
    b = PmeBatch(8 << 30)
    reg = [_add(b, s, torch, dev) for s in sl]
    ids = [r[0] for r in reg]
    for wpc in (0, 4):
        b.set_tuning(24, 32, wpc)
        for graph in (True, False):
            b.set_graph(graph)
            b.solve()
            e, v = b.results(ids)
            for s, (sid, x, f, stride), ek in zip(sl, reg, e):
                one = run_solve(s.nfft, s.order, s.kappa, s.box, s.xyzq)
                assert np.array_equal(_forces(f, s.natom, stride), one["force"]), (s.name, wpc, graph)
                assert rel_err(ek, one["energy"]) < 1e-12, (s.name, wpc, graph)


def test_batch_edge_systems():
    """A batch holding an empty system (no atoms, null pointers), atoms far
    outside the box (any periodic image) and a clustered system whose atoms
    all fall in one brick: the empty system gives zero energy and virial, the
    others the single-system solver's forces bit for bit and the oracle's
    within FP32 rounding; graph and eager solves alike."""
    torch, dev = torch_dev()
    rng = np.random.default_rng(11)
    nfft, box = (48, 40, 36), (40.0, 33.0, 30.0)
    far = np.zeros((300, 4), np.float32)
    far[:, :3] = rng.uniform(-500, 500, (300, 3))
    far[:, 3] = rng.uniform(-1, 1, 300)
    cluster = np.zeros((6000, 4), np.float32)
    cluster[:, :3] = rng.uniform(0, 1.5, (6000, 3))
    cluster[:, 3] = rng.uniform(-1, 1, 6000)
    sl = [systems.PmeSystem("empty", np.zeros((0, 4), np.float32), box, nfft, 4, 0.34),
          systems.PmeSystem("far", far, box, nfft, 6, 0.34),
          systems.PmeSystem("cluster", cluster, box, nfft, 4, 0.34)]
    b = PmeBatch(4 << 30)
    reg = []
    for s in sl:
        if s.natom == 0:
            d = PmeBatch.desc(s.nfft, s.order, s.kappa, s.recip, None, 0, None, L.FORCE_FLOAT, 0)
            reg.append((b.add(d), None, None, 0))
        else:
            reg.append(_add(b, s, torch, dev))
    ids = [r[0] for r in reg]
    for graph in (True, False):
        b.set_graph(graph)
        b.solve()
        e, v = b.results(ids)
        assert e[0] == 0.0 and np.all(v[0] == 0.0)
        for k in (1, 2):
            s = sl[k]
            _, x, f, stride = reg[k]
            fb = _forces(f, s.natom, stride)
            one = run_solve(s.nfft, s.order, s.kappa, s.box, s.xyzq)
            assert np.array_equal(fb, one["force"]), (s.name, graph)
            assert rel_err(e[k], one["energy"]) < 1e-12, s.name
            o = O.solve(O.OracleParams(*s.nfft, s.order, s.kappa, s.box, index_mode=O.INDEX_F32), s.xyzq)
            frms = rel_rms(fb, o["force"])
            say(f"  {s.name} graph={graph}: force rel-RMS vs oracle {frms:.2e}")
            assert frms < 1e-4 and rel_err(e[k], o["energy"]) < 1e-4


def test_batch_self_energy():
    """Systems described with CALC_SELF get the self energy of
    calc_self_energy (CudaPMERecip.cu:2053-2061) from the binning pass:
    equal to the single-system solver's within FP64 summation order and to
    the oracle; CALC_SELF alone leaves energy and virial zero and the forces
    unchanged; systems without it report 0.  Graph and eager solves alike."""
    torch, dev = torch_dev()
    sl = systems.config_batched(6)
    flags = [L.CALC_EV | L.CALC_SELF, L.CALC_SELF, L.CALC_EV, L.CALC_EV | L.CALC_SELF, 0, L.CALC_SELF]
    b = PmeBatch(4 << 30)
    reg = [_add(b, s, torch, dev, calc_ev=fl) for s, fl in zip(sl, flags)]
    ids = [r[0] for r in reg]
    for graph in (True, False):
        b.set_graph(graph)
        b.solve()
        e, v = b.results(ids)
        es = b.results_self(ids)
        for k, (s, fl) in enumerate(zip(sl, flags)):
            one = run_solve(s.nfft, s.order, s.kappa, s.box, s.xyzq, self_energy=True)
            if fl & L.CALC_SELF:
                assert rel_err(es[k], one["self"]) < 1e-12, (s.name, graph)
                assert rel_err(es[k], O.self_energy(s.xyzq, s.kappa)) < 1e-6, s.name
            else:
                assert es[k] == 0.0
            if fl & L.CALC_EV:
                assert rel_err(e[k], one["energy"]) < 1e-12
            else:
                assert e[k] == 0.0 and np.all(v[k] == 0.0)
            _, x, f, stride = reg[k]
            assert np.array_equal(_forces(f, s.natom, stride), one["force"]), (s.name, graph)
    say(f"  self energies {es[0]:.6f} .. {es[-1]:.6f}")


@pytest.mark.parametrize("wg_per_cu", [0, 4])
def test_two_chains_match_one(wg_per_cu):
    """A solve split into two, three or four stage chains on as many streams
    gives the forces, energies, virials and self energies of the one-chain
    solve bit for bit, graph replays and eager solves alike (the same kernels
    per system)."""
    torch, dev = torch_dev()
    sl = systems.config_batched(9)
    b = PmeBatch(8 << 30)
    reg = [_add(b, s, torch, dev, calc_ev=L.CALC_EV | L.CALC_SELF) for s in sl]
    ids = [r[0] for r in reg]
    b.set_tuning(24, 32, wg_per_cu)
    out = {}
    for chains in (1, 2, 3, 4):
        b.set_chains(chains)
        assert b.get_chains() == chains
        for graph in (True, False):
            b.set_graph(graph)
            for f in (r[2] for r in reg):
                f.zero_()
            for _ in range(2):
                b.solve()
            e, v = b.results(ids)
            es = b.results_self(ids)
            out[(chains, graph)] = ([_forces(r[2], s.natom, r[3]) for s, r in zip(sl, reg)], e, v, es)
    ref = out[(1, False)]
    for key, (fs, e, v, es) in out.items():
        assert all(np.array_equal(a, c) for a, c in zip(fs, ref[0])), key
        assert np.array_equal(e, ref[1]) and np.array_equal(v, ref[2]) and np.array_equal(es, ref[3]), key
    with pytest.raises(L.PmeError):
        b.set_chains(5)
    with pytest.raises(L.PmeError):
        b.set_chains(0)


@pytest.mark.parametrize("order", [4, 6, 8])
def test_half_bricks_match_full(order):
    """Half bricks (smaller bricks, half-size spread/gather workgroups, two
    per CU) give the forces, energies and virials of full bricks bit for bit:
    the charge grid is an integer sum whatever the bricks, and each atom's
    gather reads the same potential values in the same order."""
    torch, dev = torch_dev()
    sl = [systems.PmeSystem(s.name, s.xyzq, s.box, s.nfft, order, s.kappa) for s in systems.config_batched(6)]
    b = PmeBatch(8 << 30)
    reg = [_add(b, s, torch, dev) for s in sl]
    ids = [r[0] for r in reg]
    out = {}
    for half in (False, True):
        b.set_half_bricks(half)
        assert b.get_half_bricks() == half
        for graph in (True, False):
            b.set_graph(graph)
            for f in (r[2] for r in reg):
                f.zero_()
            b.solve()
            e, v = b.results(ids)
            out[(half, graph)] = ([_forces(r[2], s.natom, r[3]) for s, r in zip(sl, reg)], e, v)
    ref = out[(False, False)]
    for key, (fs, e, v) in out.items():
        assert all(np.array_equal(a, c) for a, c in zip(fs, ref[0])), key
        assert np.array_equal(e, ref[1]) and np.array_equal(v, ref[2]), key
    with pytest.raises(L.PmeError):
        L.call("pme_batch_set_half_bricks", b.h, 2)


def test_batched_config_vs_oracle():
    """The 64-system batched workload: every system against the single-system
    solver, eight of them against the FP64 oracle (exact coordinates)."""
    torch, dev = torch_dev()
    sl = systems.config_batched()
    b = PmeBatch(16 << 30)
    reg = [_add(b, s, torch, dev) for s in sl]
    ids = [r[0] for r in reg]
    for _ in range(3):  # graph replays
        b.solve()
    e, v = b.results(ids)
    worst = 0.0
    for k, (s, (sid, x, f, stride)) in enumerate(zip(sl, reg)):
        fb = _forces(f, s.natom, stride)
        one = run_solve(s.nfft, s.order, s.kappa, s.box, s.xyzq)
        assert np.array_equal(fb, one["force"]), s.name
        assert rel_err(e[k], one["energy"]) < 1e-12
        if k % 8 == 0:
            o = O.solve(O.OracleParams(*s.nfft, s.order, s.kappa, s.box), s.xyzq)
            frms = rel_rms(fb, o["force"])
            worst = max(worst, frms)
            assert frms < 1e-4, (s.name, frms)
            assert rel_err(e[k], o["energy"]) < 1e-4
    say(f"  worst force rel-RMS vs FP64 oracle (exact index): {worst:.2e}")


def test_graph_replays_match_eager():
    """Replays of the captured graph, with eager work and coordinate updates
    between them, each agree with an eager (uncaptured) solve."""
    torch, dev = torch_dev()
    sl = _mixed_systems()[:4]
    bg = PmeBatch(4 << 30)
    be = PmeBatch(4 << 30)
    be.set_graph(False)
    rg = [_add(bg, s, torch, dev) for s in sl]
    re_ = [_add(be, s, torch, dev) for s in sl]
    rng = np.random.default_rng(7)
    for rep in range(4):
        # move the atoms of system rep % 4 in place (the descriptor is unchanged)
        k = rep % len(sl)
        dx = torch.from_numpy(rng.normal(0, 0.3, (sl[k].natom, 4)).astype(np.float32)).to(dev)
        dx[:, 3] = 0
        rg[k][1].add_(dx)
        re_[k][1].add_(dx)
        _ = torch.ones(1024, device=dev).sum().item()  # eager work between replays
        bg.solve()
        be.solve()
        eg, vg = bg.results([r[0] for r in rg])
        ee, ve = be.results([r[0] for r in re_])
        for s, a, c in zip(sl, rg, re_):
            assert np.array_equal(_forces(a[2], s.natom, a[3]), _forces(c[2], s.natom, c[3])), (rep, s.name)
        assert np.array_equal(np.array(eg), np.array(ee)) and np.array_equal(np.array(vg), np.array(ve))
        say(f"  replay {rep}: graph == eager")


def test_subset_update_remove():
    torch, dev = torch_dev()
    sl = _mixed_systems()
    b = PmeBatch(8 << 30)
    reg = [_add(b, s, torch, dev) for s in sl]
    # solve a subset: the others' outputs stay untouched (zero)
    sub = [reg[0][0], reg[3][0]]
    b.solve(sub)
    e, _ = b.results(sub)
    for k, r in enumerate(reg):
        fb = _forces(r[2], sl[k].natom, r[3])
        if r[0] in sub:
            assert np.abs(fb).sum() > 0
        else:
            assert not fb.any()
    # grow system 1 beyond its admitted capacity: the slot is re-planned
    s1 = sl[1]
    big = systems.triplets(4000, s1.box, 99)
    x = to_dev(big)
    stride = force_stride(len(big))
    f = torch.zeros(3 * stride, dtype=torch.float32, device=dev)
    b.update(reg[1][0], PmeBatch.desc(s1.nfft, s1.order, s1.kappa, s1.recip, x, len(big), f, L.FORCE_FLOAT, stride),
             keep=(x, f))
    # remove system 2 and admit another in its place
    b.remove(reg[2][0])
    s_new = systems.PmeSystem("new", systems.triplets(700, (30.0, 31.0, 32.0), 5), (30.0, 31.0, 32.0), (36, 36, 36), 6, 0.3)
    nid, nx_, nf, nstride = _add(b, s_new, torch, dev)
    assert nid == reg[2][0]  # the freed slot is reused
    ids = [reg[1][0], nid, reg[4][0]]
    b.solve(ids)
    e, v = b.results(ids)
    for sysm, xx, ff, st, ek in [(systems.PmeSystem("g", big, s1.box, s1.nfft, s1.order, s1.kappa), x, f, stride, e[0]),
                                 (s_new, nx_, nf, nstride, e[1])]:
        one = run_solve(sysm.nfft, sysm.order, sysm.kappa, sysm.box, sysm.xyzq)
        assert np.array_equal(_forces(ff, sysm.natom, st), one["force"])
        assert rel_err(ek, one["energy"]) < 1e-12
    with pytest.raises(L.PmeError):
        b.solve([reg[2][0] + 100])


def test_budget_admission():
    torch, dev = torch_dev()
    s = _mixed_systems()[3]
    b = PmeBatch(1 << 20)  # 1 MiB: the 60x64x72 grid alone needs more
    used0, budget = b.memory()
    assert budget == 1 << 20
    with pytest.raises(L.PmeError) as ei:
        _add(b, s, torch, dev)
    assert ei.value.status == 4  # PME_ERR_BUDGET
    assert b.memory()[0] == used0
    # a tiny system fits
    small = systems.PmeSystem("tiny", systems.triplets(10, (20.0, 20.0, 20.0), 3), (20.0, 20.0, 20.0), (16, 16, 16), 4, 0.4)
    sid = _add(b, small, torch, dev)[0]
    assert b.memory()[0] > used0
    b.solve([sid])
    b.results([sid])


def test_force_types_and_accumulation():
    torch, dev = torch_dev()
    s = _mixed_systems()[0]
    b = PmeBatch(2 << 30)
    r_float = _add(b, s, torch, dev, L.FORCE_FLOAT)
    r_f3 = _add(b, s, torch, dev, L.FORCE_FLOAT3)
    r_i64 = _add(b, s, torch, dev, L.FORCE_INT64, calc_ev=False)
    b.solve()
    b.solve()  # INT64 accumulates: twice the force
    e, _ = b.results([r_float[0], r_i64[0]])
    assert e[1] == 0.0
    ff = _forces(r_float[2], s.natom, r_float[3])
    f3 = _forces(r_f3[2], s.natom, r_f3[3], L.FORCE_FLOAT3)
    fi = _forces(r_i64[2], s.natom, r_i64[3], L.FORCE_INT64)
    assert np.array_equal(ff, f3)
    assert rel_rms(fi, 2 * ff) < 1e-6


def test_autotune_cache(tmp_path):
    torch, dev = torch_dev()
    sl = systems.config_batched(8)
    b = PmeBatch(8 << 30)
    reg = [_add(b, s, torch, dev) for s in sl]
    ids = [r[0] for r in reg]
    b.solve()
    e0, _ = b.results(ids)
    f0 = [_forces(r[2], s.natom, r[3]) for s, r in zip(sl, reg)]
    cache = tmp_path / "tune.json"
    ms = b.autotune(str(cache))
    tune = b.get_tuning()
    say(f"  autotune: {ms:.3f} ms/solve with (brick, tile, wg/cu) = {tune}")
    assert ms > 0 and cache.exists()
    entry = list(json.loads(cache.read_text()).values())[0]
    assert tuple(int(v) for v in entry[:3]) == tune
    assert int(entry[3]) == b.get_chains() and bool(entry[4]) == b.get_half_bricks()
    # results with the tuned plan: forces to FP32 rounding (tiles change the
    # FFT's summation order), energies to 1e-6
    b.solve()
    e1, _ = b.results(ids)
    for s, r, fa, ea, eb in zip(sl, reg, f0, e0, e1):
        assert rel_rms(_forces(r[2], s.natom, r[3]), fa) < 1e-5
        assert rel_err(eb, ea) < 1e-6
    # a second batch with the same set reads the cache
    b2 = PmeBatch(8 << 30)
    for s in sl:
        _add(b2, s, torch, dev)
    ms2 = b2.autotune(str(cache))
    assert b2.get_tuning() == tune and math.isclose(ms2, ms, rel_tol=1e-5)
    assert b2.get_half_bricks() == b.get_half_bricks()


def test_stage_timing():
    torch, dev = torch_dev()
    sl = _mixed_systems()[:3]
    b = PmeBatch(4 << 30)
    reg = [_add(b, s, torch, dev) for s in sl]
    ids = [r[0] for r in reg]
    b.solve()
    e0, _ = b.results(ids)
    b.set_timing(True)
    b.solve()
    e1, _ = b.results(ids)
    ms = b.get_timing()
    say(f"  batch stage ms: {[round(v, 4) for v in ms]}")
    assert len(ms) == 8 and all(v >= 0 for v in ms) and ms[1] > 0 and ms[4] > 0 and ms[7] > 0
    assert np.array_equal(e0, e1)
    b.set_timing(False)
    with pytest.raises(L.PmeError):
        b.get_timing()
