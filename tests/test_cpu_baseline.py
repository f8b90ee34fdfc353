"""The FP32 CPU restatement timed as the benchmark's CPU baseline
(cpu_baseline/): its grid indices are the oracle's FP32 indices bit for bit,
and its forces, energy and virial match the FP64 oracle within FP32
precision."""
import numpy as np
import pytest

import cpu_baseline as CB
import oracle as O
from gpu_recip_amd import systems


def _rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def test_index_bit_equal_to_oracle():
    rng = np.random.default_rng(11)
    box = np.array([31.0, 28.5, 40.25])
    n = (32, 27, 45)
    x = np.zeros((4000, 4), np.float32)
    x[:, :3] = rng.uniform(-2 * box, 3 * box, (4000, 3)).astype(np.float32)
    # coordinates where w + 1/2 lands on a cell boundary, and one ulp either side
    k = np.arange(200)
    edge = (k[:, None] * (box / np.array(n)) - 0.5 * box / np.array(n)).astype(np.float32)
    for d in (-1, 0, 1):
        e = np.nextafter(edge, np.float32(np.inf) if d > 0 else np.float32(-np.inf)) if d else edge
        x[200 * (d + 1):200 * (d + 2), :3] = e
    recip = 1.0 / box
    i_c, f_c = CB.index_f32(x, recip, n)
    i_o, f_o = O.index_f32(x, recip, n)
    assert np.array_equal(i_c, i_o)
    assert np.array_equal(f_c, f_o)


@pytest.mark.parametrize("nfft,box,order,ntrip,kappa", [
    ((32, 32, 32), (30.0, 30.0, 30.0), 4, 900, 0.34),
    ((30, 36, 40), (29.0, 33.0, 38.0), 6, 1100, 0.32),
    ((27, 25, 24), (26.0, 25.0, 23.0), 8, 600, 0.30),
])
def test_solve_vs_oracle(nfft, box, order, ntrip, kappa):
    x = systems.triplets(ntrip, box, seed=order)
    for nthreads in (1, 3):
        r = CB.CpuPme(nfft, order).solve(x, box, kappa, nthreads=nthreads)
        o = O.solve(O.OracleParams(*nfft, order, kappa, box, index_mode=O.INDEX_F32), x)
        assert _rel_rms(r["force"].astype(np.float64), o["force"]) < 1e-5
        assert abs(r["energy"] - o["energy"]) / abs(o["energy"]) < 1e-5
        assert np.linalg.norm(r["virial"] - o["virial"]) / np.linalg.norm(o["virial"]) < 1e-5


def test_solve_independent_of_threads():
    nfft, box = (36, 36, 36), (34.0, 34.0, 34.0)
    x = systems.triplets(1500, box, seed=5)
    cp = CB.CpuPme(nfft, 4)
    a = cp.solve(x, box, 0.34, nthreads=1)
    b = cp.solve(x, box, 0.34, nthreads=4)
    # the spread and the transforms do not depend on the thread count; only
    # the FP64 energy partials are summed per z plane in fixed order
    assert np.array_equal(a["force"], b["force"])
    assert a["energy"] == b["energy"]


def test_unsupported():
    with pytest.raises(ValueError):
        CB.CpuPme((7, 32, 32), 4)
    with pytest.raises(ValueError):
        CB.CpuPme((32, 32, 32), 10)
