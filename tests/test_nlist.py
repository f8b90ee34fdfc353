"""CPU tests of the tile neighbour-list builder (gpu_recip_amd/nlist.py): the
pairs the list hands to the direct-space kernel, evaluated in FP64 with the
kernel's rules (rotation i = (lane + t) mod 32, exclusion bit t of word
lane, cutoff), reproduce the oracle's brute-force direct sum over all pairs
and images with the same exclusions."""
import numpy as np
import pytest

import oracle as O
from gpu_recip_amd import nlist, systems


def eval_list(tl: nlist.TileList, box, kappa, roff):
    """FP64 energy and gradient of the pairs the list selects (original atom order)."""
    from math import erfc

    x = tl.xyzq[:, :3].astype(np.float64)
    q = tl.xyzq[:, 3].astype(np.float64)
    box = np.asarray(box, np.float64)
    f = np.zeros((tl.npad, 3))
    e = 0.0
    npair = 0
    t = np.arange(32)
    for i0, ish, ts, te in tl.ientry.tolist():
        sh = nlist.SHIFTS[ish] * box
        for k in range(ts, te + 1):
            j0 = int(tl.tile_indj[k])
            for jl in range(32):
                word = int(tl.tile_excl[k, jl])
                ii = i0 + (jl + t) % 32
                keep = ((word >> t) & 1) == 0
                d = x[ii] + sh - x[j0 + jl]
                r2 = (d * d).sum(1)
                keep &= r2 < roff * roff
                for tt in np.nonzero(keep)[0]:
                    i = ii[tt]
                    r = np.sqrt(r2[tt])
                    qq = O.CCELEC * q[i] * q[j0 + jl]
                    e += qq * erfc(kappa * r) / r
                    du = -qq * (2 / np.sqrt(np.pi)) * kappa * np.exp(-kappa * kappa * r * r) - qq * erfc(kappa * r) / r
                    g = du / (r * r) * d[tt]
                    f[i] += g
                    f[j0 + jl] -= g
                    npair += 1
    out = np.zeros((tl.natom, 3))
    m = tl.perm >= 0
    out[tl.perm[m]] = f[m]
    return e, out, npair


@pytest.mark.parametrize("ntrip,boxl,roff", [(60, 16.0, 7.0), (120, 19.0, 9.0)])
def test_list_reproduces_brute_force_pairs(ntrip, boxl, roff):
    box = (boxl, boxl + 1.0, boxl + 2.0)
    x = systems.triplets(ntrip, box, seed=5 + ntrip)
    kappa = 0.32
    excl = nlist.water_exclusions(len(x))
    tl = nlist.build(x, box, roff, excl_pairs=excl)
    assert tl.npad % 32 == 0 and tl.npad >= len(x)
    e, f, npair = eval_list(tl, box, kappa, roff)
    # the brute force on the same wrapped coordinates (original order)
    xw = x.copy()
    xw[:, :3] = np.mod(x[:, :3].astype(np.float64), box).astype(np.float32)
    o = O.direct_ewald(xw, box, kappa, roff, excl=excl)
    assert npair > 0
    assert abs(e - o["energy"]) <= 1e-9 * abs(o["energy"])
    assert np.abs(f - o["force"]).max() <= 1e-9 * np.abs(o["force"]).max()


def test_ex14_shifts_give_minimum_images():
    box = (10.0, 11.0, 12.0)
    x = np.zeros((4, 4), np.float32)
    x[0, :3] = (0.2, 5.0, 11.9)
    x[1, :3] = (9.9, 5.1, 0.1)   # across the x and z faces from atom 0
    x[2, :3] = (5.0, 5.0, 5.0)
    x[3, :3] = (5.5, 5.0, 5.0)
    l14 = nlist.ex14_list(x, box, [(0, 1), (2, 3)])
    for (i, j, s) in l14.tolist():
        d = x[i, :3].astype(np.float64) + nlist.SHIFTS[s] * np.asarray(box) - x[j, :3]
        assert np.linalg.norm(d) < 1.0
    assert l14[1, 2] == 13
