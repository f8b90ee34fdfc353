"""Parity of the HIP path (through the C ABI) with the oracle and with the
reference's own fixtures.  GPU only.

Tolerances (per result, FP32 path):
  * grid indices / stencil placement: exact (bit-equal int) vs the FP32
    restatement of CudaPMERecip.cu:512-529;
  * vs the oracle with the same FP32 index mapping (isolates FP32 arithmetic):
    force rel-RMS <= 1e-5, energy rel <= 1e-5, virial rel-Frobenius <= 1e-5;
  * vs the FP64 oracle (FP64 index mapping): force rel-RMS <= 1e-4;
  * reference fixtures: the reference's own absolute tolerances
    (gpu_recip.cu:146,157,286 for order 4; :373,384,515 for order 6).
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from gpu_recip_amd import _lib as L
from gpu_recip_amd import systems
from gpu_recip_amd.pme import CudaEnergyVirial, CudaPMERecip, fill_bspline, force_stride

from gpu_util import rel_err, rel_fro, rel_rms, run_solve, say, to_dev, torch_dev

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def oracle_params(s, index_mode=O.INDEX_F32):
    return O.OracleParams(*s.nfft, s.order, s.kappa, s.box, index_mode=index_mode)


def spread_error(got, p, xyzq, ref):
    """RMS spread error relative to the RMS of the grid spread with |q|: FP32
    accumulation error scales with the magnitude of the summed terms, not with
    the (possibly cancelled) result."""
    absq = xyzq.copy()
    absq[:, 3] = np.abs(absq[:, 3])
    scale = O.spread(p, absq).ravel()
    return float(np.sqrt(((got - ref) ** 2).sum() / max((scale ** 2).sum(), 1e-300)))


def small_system(nfft, box, order, ntrip, seed, kappa=0.34):
    return systems.PmeSystem("t", systems.triplets(ntrip, box, seed), box, nfft, order, kappa)


# ------------------------------------------------------------ index parity --

def test_index_parity_fill_bspline():
    torch, dev = torch_dev()
    rng = np.random.default_rng(5)
    for nfft, box in (((64, 64, 64), (62.23,) * 3), ((108, 100, 96), (100.0, 96.0, 90.0)), ((45, 48, 50), (37.7, 40.0, 41.0))):
        xs = []
        from test_oracle import boundary_coords
        for a in range(3):
            xs.append(boundary_coords(nfft[a], box[a], per_cell=2, images=(-1, 0, 1)))
        m = max(len(v) for v in xs)
        x = np.zeros((m + 5000, 4), np.float32)
        for a in range(3):
            x[:m, a] = np.resize(xs[a], m)
            x[m:, a] = rng.uniform(-2 * box[a], 3 * box[a], 5000)
        x[:, 3] = rng.uniform(-1, 1, len(x))
        recip = [1 / box[0], 0, 0, 0, 1 / box[1], 0, 0, 0, 1 / box[2]]
        for order in (4, 6, 8):
            say(f"  fill_bspline nfft={nfft} order={order} n={len(x)}")
            out = fill_bspline(order, nfft, recip, to_dev(x), len(x))
            torch.cuda.synchronize()
            gi = np.stack([out[0].cpu().numpy(), out[1].cpu().numpy(), out[2].cpu().numpy()], 1)
            idx, fr = O.index_f32(x, np.array(box) ** -1.0, nfft)
            assert np.array_equal(gi, idx), f"index mismatch at {np.argwhere(gi != idx)[:5]}"
            assert np.array_equal(out[3].cpu().numpy(), x[:, 3])
            # weights: FP32 recursion vs FP64 recursion on the same FP32 fraction
            for ax in range(3):
                th = out[4 + ax].cpu().numpy().reshape(-1, order)
                dth = out[7 + ax].cpu().numpy().reshape(-1, order)
                sel = np.arange(0, len(x), 97)
                for a in sel:
                    t64, d64 = O.bspline(order, float(fr[a, ax]))
                    assert np.abs(th[a] - t64).max() < 3e-7
                    assert np.abs(dth[a] - d64).max() < 3e-7


# ------------------------------------------------------------------ spread --

@pytest.mark.parametrize("nfft,box,order,ntrip", [
    ((48, 48, 48), (40.0, 40.0, 40.0), 4, 2667),
    ((108, 100, 96), (100.0, 96.0, 90.0), 6, 32000),
    ((45, 27, 50), (40.0, 24.0, 44.0), 6, 3000),
    ((8, 8, 8), (10.0, 10.0, 10.0), 4, 40),        # one brick per axis: stencil wraps inside the brick
    ((20, 24, 16), (18.0, 22.0, 15.0), 8, 500),
])
def test_spread_grid(nfft, box, order, ntrip):
    torch, dev = torch_dev()
    s = small_system(nfft, box, order, ntrip, 11)
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(*nfft, order, ev)
    x = to_dev(s.xyzq)
    pme.spread_charge(x, s.natom, s.recip)
    g = torch.empty(nfft[2] * nfft[1] * nfft[0], dtype=torch.float32, device=dev)
    pme.copy_grid(L.GRID_REAL, g)
    torch.cuda.synchronize()
    ref = O.spread(oracle_params(s), s.xyzq).ravel()
    got = g.cpu().numpy().astype(np.float64)
    err = spread_error(got, oracle_params(s), s.xyzq, ref)
    say(f"  spread rel-RMS {rel_rms(got, ref):.2e}  err/|q|-grid {err:.2e}")
    assert err < 1e-6
    assert abs(got.sum() - s.xyzq[:, 3].astype(np.float64).sum()) < 1e-3


def test_spread_edge_inputs():
    """empty, single atom, atoms far outside the box, and a clustered system
    whose atoms overflow their bin (all in one brick)."""
    torch, dev = torch_dev()
    nfft, box, order = (48, 40, 36), (40.0, 33.0, 30.0), 6
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(*nfft, order, ev)
    recip = [1 / box[0], 0, 0, 0, 1 / box[1], 0, 0, 0, 1 / box[2]]
    g = torch.empty(nfft[0] * nfft[1] * nfft[2], dtype=torch.float32, device=dev)
    rng = np.random.default_rng(3)
    far = np.zeros((300, 4), np.float32)
    far[:, :3] = rng.uniform(-500, 500, (300, 3))
    far[:, 3] = rng.uniform(-1, 1, 300)
    cluster = np.zeros((20000, 4), np.float32)
    cluster[:, :3] = rng.uniform(0, 1.5, (20000, 3))
    cluster[:, 3] = rng.uniform(-1, 1, 20000)
    cases = {"empty": np.zeros((0, 4), np.float32), "one": far[:1], "far": far, "cluster": cluster}
    for name, x in cases.items():
        say(f"  spread edge case {name} n={len(x)}")
        xd = to_dev(x) if len(x) else torch.zeros((1, 4), dtype=torch.float32, device=dev)
        pme.spread_charge(xd, len(x), recip)
        pme.copy_grid(L.GRID_REAL, g)
        torch.cuda.synchronize()
        got = g.cpu().numpy().astype(np.float64)
        p = O.OracleParams(*nfft, order, 0.34, box, index_mode=O.INDEX_F32)
        ref = O.spread(p, x).ravel() if len(x) else np.zeros_like(got)
        if len(x) == 0:
            assert np.all(got == 0)
        else:
            err = spread_error(got, p, x, ref)
            say(f"    err/|q|-grid {err:.2e}")
            assert err < 1e-6, name


# --------------------------------------------------------------------- FFT --

# Lengths whose transform runs as a register four-step (n = a*b, codelets
# 2, 3, 4, 5, 6, 8, 9, 10, 12, 16) on every axis, in plane and separate-pass
# layouts, next to Stockham-only lengths (75, 125) and the mixed shapes above.
FOUR_STEP_SHAPES = [(32, 10, 12), (36, 18, 27), (40, 50, 54), (48, 60, 64), (81, 81, 45), (96, 100, 128),
                    (256, 144, 160), (20, 16, 192), (150, 125, 75)]


@pytest.mark.parametrize("nfft", [(48, 48, 48), (108, 100, 96), (45, 27, 50), (64, 64, 64), (25, 36, 75), (144, 40, 30)]
                         + FOUR_STEP_SHAPES)
def test_r2c_spectrum_and_roundtrip(nfft):
    torch, dev = torch_dev()
    box = tuple(0.9 * n for n in nfft)
    s = small_system(nfft, box, 4, 1000, 7)
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(*nfft, 4, ev)
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
    say(f"  r2c rel err {rel_rms(spec, ref):.2e}")
    assert rel_rms(spec, ref) < 3e-6
    # c2r of the unsolved spectrum gives G * grid (both unnormalised)
    pme.c2r_fft()
    pme.copy_grid(L.GRID_REAL, g)
    torch.cuda.synchronize()
    back = g.cpu().numpy().astype(np.float64).reshape(nz, ny, nx)
    assert rel_rms(back, grid * (nx * ny * nz)) < 3e-6


# ------------------------------------------------------------- full solves --

CASES = [
    ("plain-8k-48c-p4", systems.config_plain_small),
    ("plain-96k-p6", systems.config_production),
]


@pytest.mark.parametrize("name,make", CASES)
def test_full_solve_vs_oracle(name, make):
    s = make()
    r = run_solve(s.nfft, s.order, s.kappa, s.box, s.xyzq, self_energy=True)
    o32 = O.solve(oracle_params(s, O.INDEX_F32), s.xyzq)
    o64 = O.solve(oracle_params(s, O.INDEX_F64), s.xyzq)
    e32, f32, v32 = rel_err(r["energy"], o32["energy"]), rel_rms(r["force"], o32["force"]), rel_fro(r["virial"], o32["virial"])
    e64, f64, v64 = rel_err(r["energy"], o64["energy"]), rel_rms(r["force"], o64["force"]), rel_fro(r["virial"], o64["virial"])
    say(f"  {name}: vs oracle(FP32 index) E {e32:.2e} F {f32:.2e} V {v32:.2e} | vs FP64 oracle E {e64:.2e} F {f64:.2e} V {v64:.2e}")
    assert e32 < 1e-5 and f32 < 1e-5 and v32 < 1e-5
    assert f64 < 1e-4 and e64 < 1e-4 and v64 < 1e-4
    assert rel_err(r["self"], O.self_energy(s.xyzq, s.kappa)) < 1e-12


@pytest.mark.parametrize("nfft,box,order", [
    ((36, 40, 45), (33.0, 36.0, 41.0), 8),
    ((45, 27, 25), (40.0, 24.0, 22.0), 6),
    ((8, 8, 8), (12.0, 12.0, 12.0), 4),
    ((60, 16, 30), (55.0, 15.0, 27.0), 4),
    # fused z pass as a register four-step: 75 = 5*15, 81 = 9*9, 128 = 8*16, 160 = 10*16, 12 = 3*4
    ((40, 36, 75), (36.0, 33.0, 68.0), 4),
    ((30, 32, 81), (27.0, 29.0, 73.0), 6),
    ((24, 30, 128), (22.0, 27.0, 115.0), 4),
    ((20, 24, 160), (18.0, 22.0, 144.0), 8),
    ((40, 45, 12), (36.0, 41.0, 11.0), 4),
])
def test_full_solve_shapes(nfft, box, order):
    s = small_system(nfft, box, order, 1500, 19)
    for five in (True, False):
        r = run_solve(s.nfft, s.order, s.kappa, s.box, s.xyzq, five_calls=five)
        o = O.solve(oracle_params(s, O.INDEX_F32), s.xyzq)
        f = rel_rms(r["force"], o["force"])
        say(f"  shape {nfft} p{order} five_calls={five}: F {f:.2e} E {rel_err(r['energy'], o['energy']):.2e}")
        assert f < 1e-5
        assert rel_err(r["energy"], o["energy"]) < 1e-5
        assert rel_fro(r["virial"], o["virial"]) < 1e-5


@pytest.mark.parametrize("order", [4, 6])
def test_reference_fixture(order):
    """gpu_recip test4/test6 (gpu_recip.cu:68-293, 298-522) with its tolerances."""
    d = np.load(os.path.join(GOLD, "recip_system_23558.npz"))
    with open(os.path.join(GOLD, "recip_constants.json")) as f:
        c = json.load(f)[f"order{order}"]
    s = systems.reference_fixture_system(d["xyzq"], order)
    r = run_solve(s.nfft, s.order, s.kappa, s.box, s.xyzq)
    fref = d[f"force_recip_{order}"]
    de = abs(r["energy"] - c["energy"])
    dv = np.abs(r["virial"] - np.array(c["virial"])).max()
    df = np.abs(r["force"] - fref).max()
    say(f"  fixture order {order}: dE {de:.2e} (tol {c['tol_energy']}) dV {dv:.2e} (tol {c['tol_virial']}) "
        f"dF {df:.2e} (tol {c['tol_force']})")
    assert de < c["tol_energy"]
    assert dv < c["tol_virial"]
    assert df < c["tol_force"]


def test_force_types_and_accumulation():
    s = small_system((48, 48, 48), (40.0, 40.0, 40.0), 4, 2000, 23)
    base = run_solve(s.nfft, s.order, s.kappa, s.box, s.xyzq, force_type=L.FORCE_FLOAT)
    f3 = run_solve(s.nfft, s.order, s.kappa, s.box, s.xyzq, force_type=L.FORCE_FLOAT3)
    assert rel_rms(f3["force"], base["force"]) < 1e-6
    fi = run_solve(s.nfft, s.order, s.kappa, s.box, s.xyzq, force_type=L.FORCE_INT64)
    # separate solves: the FP32 LDS accumulation order differs run to run (ulp level)
    assert rel_rms(fi["force"], base["force"]) < 1e-6
    # fixed point ADDS (gather_force<long long>): a second gather doubles it
    torch, dev = torch_dev()
    pme, x = fi["pme"], fi["x"]
    n = s.natom
    stride = force_stride(n)
    f = torch.zeros(3 * stride, dtype=torch.int64, device=dev)
    pme.gather_force(x, n, s.recip, stride, f, L.FORCE_INT64)
    pme.gather_force(x, n, s.recip, stride, f, L.FORCE_INT64)
    torch.cuda.synchronize()
    fh = f.cpu().numpy()
    two = np.stack([fh[:n], fh[stride:stride + n], fh[2 * stride:2 * stride + n]], 1) / 2.0 ** 40
    assert rel_rms(two, 2 * base["force"]) < 1e-6


def test_stage_state_machine():
    torch, dev = torch_dev()
    s = small_system((48, 40, 36), (40.0, 33.0, 30.0), 6, 2000, 29)
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(*s.nfft, s.order, ev)
    x = to_dev(s.xyzq)
    n = s.natom
    stride = force_stride(n)
    f = torch.zeros(3 * stride, dtype=torch.float32, device=dev)
    # gather before any potential exists is an error, not garbage
    with pytest.raises(L.PmeError) as ei:
        pme.gather_force(x, n, s.recip, stride, f)
    assert ei.value.name == "PME_ERR_STATE"
    with pytest.raises(L.PmeError):
        pme.r2c_fft()
    # scalar_sum twice multiplies by eterm twice (as the reference, which mutates in place)
    nx, ny, nz = s.nfft
    c = torch.empty((nx // 2 + 1) * ny * nz * 2, dtype=torch.float32, device=dev)
    pme.spread_charge(x, n, s.recip)
    pme.r2c_fft()
    pme.copy_grid(L.GRID_SPECTRUM, c)
    pme.scalar_sum(s.recip, s.kappa, False, False)
    pme.scalar_sum(s.recip, s.kappa, False, False)
    c2 = torch.empty_like(c)
    pme.copy_grid(L.GRID_SPECTRUM, c2)
    torch.cuda.synchronize()
    spec = c.cpu().numpy().astype(np.float64).view(np.complex128).reshape(nz, ny, nx // 2 + 1)
    p = oracle_params(s)
    once, _, _ = O.scalar_sum(p, spec)
    twice, _, _ = O.scalar_sum(p, once)
    got = c2.cpu().numpy().astype(np.float64).view(np.complex128).reshape(nz, ny, nx // 2 + 1)
    assert rel_rms(got, twice) < 1e-5
    # energies accumulate across scalar_sum calls until clear()
    ev.clear()
    pme.spread_charge(x, n, s.recip)
    pme.r2c_fft()
    pme.scalar_sum(s.recip, s.kappa, True, True)
    pme.c2r_fft()
    pme.spread_charge(x, n, s.recip)
    pme.r2c_fft()
    pme.scalar_sum(s.recip, s.kappa, True, True)
    ev.copyToHost()
    torch.cuda.synchronize()
    o = O.solve(p, s.xyzq)
    assert rel_err(ev.getEnergy("recip"), 2 * o["energy"]) < 1e-5
    assert rel_fro(ev.getVirial(), 2 * o["virial"]) < 1e-5


def test_gather_with_moved_atoms():
    """gather after the atoms moved since the spread (bins stale): the result
    is the oracle gather of the same potential at the new positions."""
    torch, dev = torch_dev()
    s = small_system((48, 48, 48), (40.0, 40.0, 40.0), 4, 2000, 31)
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(*s.nfft, s.order, ev)
    x = to_dev(s.xyzq)
    n = s.natom
    pme.spread_charge(x, n, s.recip)
    pme.r2c_fft()
    pme.scalar_sum(s.recip, s.kappa, False, False)
    pme.c2r_fft()
    g = torch.empty(48 ** 3, dtype=torch.float32, device=dev)
    pme.copy_grid(L.GRID_REAL, g)
    moved = s.xyzq.copy()
    moved[:, :3] += np.random.default_rng(2).normal(0, 3.0, (n, 3)).astype(np.float32)
    x.copy_(to_dev(moved))  # same pointer, new content: bins are stale
    stride = force_stride(n)
    f = torch.zeros(3 * stride, dtype=torch.float32, device=dev)
    pme.gather_force(x, n, s.recip, stride, f)
    torch.cuda.synchronize()
    phi = g.cpu().numpy().astype(np.float64).reshape(48, 48, 48)
    ref = O.gather(oracle_params(s), moved, phi)
    fh = f.cpu().numpy()
    got = np.stack([fh[:n], fh[stride:stride + n], fh[2 * stride:2 * stride + n]], 1)
    assert rel_rms(got, ref) < 1e-5


def test_set_order_recomputes():
    s = small_system((48, 48, 48), (40.0, 40.0, 40.0), 4, 1500, 37)
    torch, dev = torch_dev()
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(*s.nfft, 4, ev)
    pme.set_order(6)
    assert pme.get_info() == (48, 48, 48, 6)
    x = to_dev(s.xyzq)
    n = s.natom
    stride = force_stride(n)
    f = torch.zeros(3 * stride, dtype=torch.float32, device=dev)
    ev.clear()
    pme.solve(x, n, s.recip, s.kappa, True, True, stride, f)
    ev.copyToHost()
    torch.cuda.synchronize()
    s6 = systems.PmeSystem("t", s.xyzq, s.box, s.nfft, 6, s.kappa)
    o = O.solve(oracle_params(s6), s.xyzq)
    assert rel_err(ev.getEnergy("recip"), o["energy"]) < 1e-5
