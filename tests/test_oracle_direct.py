"""CPU tests pinning the FP64 direct-space Ewald oracle (orc_direct_ewald,
orc_pair14) without retyping its formula:
- with the brute-force Ewald reciprocal sum and the self energy, the total
  Ewald energy and forces do not depend on kappa (the splitting is exact);
- at small kappa the direct sum tends to the plain Coulomb sum in the cutoff;
- forces are the finite-difference gradient of the energy;
- the pair virial equals the strain derivative of the energy;
- an excluded pair corrected with an ex14 entry (fac 0) removes exactly its
  full Coulomb interaction from the Ewald total."""
import numpy as np
import pytest

import oracle as O


def _charges(n, box, seed, spread=1.0):
    """n charges, multiples of 1/4 (exact squares in FP32), neutral, placed
    in the first `spread` fraction of the box."""
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 4), np.float32)
    x[:, :3] = rng.uniform(0.0, spread, (n, 3)) * np.asarray(box)
    q = rng.choice([-0.75, -0.5, -0.25, 0.25, 0.5, 0.75], n)
    q[-1] -= q.sum()
    x[:, 3] = q
    return x


@pytest.mark.parametrize("n,seed", [(2, 1), (5, 2), (8, 3)])
def test_total_ewald_is_kappa_independent(n, seed):
    box = (24.0, 26.0, 28.0)
    x = _charges(n, box, seed, spread=0.2)  # every pair well inside the cutoff
    tot = []
    for kappa in (0.5, 0.6, 0.7):
        d = O.direct_ewald(x, box, kappa, roff=11.9)  # erfc(kappa roff) < 1e-16
        r = O.ewald_recip(x, box, kappa, (44, 44, 44))  # exp(-pi^2 m^2 / kappa^2) < 1e-24
        s = O.self_energy(x, kappa)
        tot.append((d["energy"] + r["energy"] + s, d["force"] + r["force"]))
    for e, f in tot[1:]:
        assert abs(e - tot[0][0]) <= 1e-9 * max(1.0, abs(tot[0][0]))
        assert np.abs(f - tot[0][1]).max() <= 1e-8 * max(1.0, np.abs(tot[0][1]).max())
    # the direct part carries a real share of the total at these kappas
    assert abs(O.direct_ewald(x, box, 0.5, roff=11.9)["energy"]) > 1e-3


def test_small_kappa_tends_to_cutoff_coulomb():
    box = (20.0, 20.0, 20.0)
    x = _charges(6, box, 7)
    roff = 9.0
    # plain Coulomb sum over every pair and image inside the cutoff
    e = 0.0
    for i in range(6):
        for j in range(i + 1, 6):
            for s in range(27):
                sh = np.array([s % 3 - 1, (s // 3) % 3 - 1, s // 9 - 1]) * np.asarray(box)
                d = x[i, :3].astype(np.float64) - x[j, :3] + sh
                r = np.linalg.norm(d)
                if r < roff:
                    e += O.CCELEC * float(x[i, 3]) * float(x[j, 3]) / r
    d = O.direct_ewald(x, box, 1e-7, roff)
    assert abs(d["energy"] - e) <= 1e-6 * abs(e)


def test_forces_are_the_energy_gradient():
    box = (15.0, 16.0, 17.0)
    x = _charges(6, box, 11)
    kappa, roff = 0.4, 7.0
    excl = [(0, 1)]
    d = O.direct_ewald(x, box, kappa, roff, excl=excl)
    h = 1e-4
    xd = x.astype(np.float64)
    for i in range(6):
        for a in range(3):
            xp = xd.copy()
            xm = xd.copy()
            xp[i, a] += h
            xm[i, a] -= h
            # recentre in float64 through a float32 array with the same values
            ep = O.direct_ewald(xp.astype(np.float32), box, kappa, roff, excl=excl)["energy"]
            em = O.direct_ewald(xm.astype(np.float32), box, kappa, roff, excl=excl)["energy"]
            fd = (ep - em) / (np.float32(xd[i, a] + h) - np.float32(xd[i, a] - h))
            assert abs(fd - d["force"][i, a]) <= 2e-3 * max(1.0, abs(d["force"][i, a]))


def test_pair_virial_is_the_strain_derivative():
    box = np.array([15.0, 16.0, 17.0])
    x = _charges(6, box, 13)
    kappa, roff = 0.4, 7.0
    d = O.direct_ewald(x, box, kappa, roff)
    vir = d["virial"].reshape(3, 3)
    eps = 1e-3  # the strain steps stay far above the FP32 rounding of the coordinates
    for a in range(3):
        xs = x.astype(np.float64).copy()
        bp, bm = box.copy(), box.copy()
        bp[a] *= 1 + eps
        bm[a] *= 1 - eps
        xp, xm = xs.copy(), xs.copy()
        xp[:, a] *= 1 + eps
        xm[:, a] *= 1 - eps
        ep = O.direct_ewald(xp.astype(np.float32), bp, kappa, roff)["energy"]
        em = O.direct_ewald(xm.astype(np.float32), bm, kappa, roff)["energy"]
        # dE/d(eps_aa) = sum_pairs r_a dE/dr_a
        assert abs((ep - em) / (2 * eps) - vir[a, a]) <= 1e-3 * max(1.0, abs(vir[a, a]))


def test_exclusion_correction_removes_the_full_coulomb_pair():
    box = (14.0, 14.0, 14.0)
    x = _charges(4, box, 17)
    x[1, :3] = x[0, :3] + np.array([0.9572, 0.0, 0.0], np.float32)  # a bonded pair
    kappa, roff = 0.34, 6.5
    full = O.direct_ewald(x, box, kappa, roff)
    excl = O.direct_ewald(x, box, kappa, roff, excl=[(0, 1)])
    corr = O.pair14(x, box, kappa, 0.0, [(0, 1, 13)])  # ishift 13: no shift
    r = float(np.linalg.norm(x[0, :3].astype(np.float64) - x[1, :3]))
    coul = O.CCELEC * float(x[0, 3]) * float(x[1, 3]) / r
    # full - pair erfc term - pair erf term = full - pair Coulomb
    assert abs((excl["energy"] + corr["energy"]) - (full["energy"] - coul)) <= 1e-9 * abs(coul)
    # an in14 entry with e14fac = 1 restores the pair's erfc term exactly
    in14 = O.pair14(x, box, kappa, 1.0, [(0, 1, 13)])
    assert abs(excl["energy"] + in14["energy"] - full["energy"]) <= 1e-9 * abs(full["energy"])
    assert np.allclose(excl["force"] + in14["force"], full["force"], rtol=0, atol=1e-9)


def test_tile_list_sum_equals_the_brute_force_sum():
    """orc_direct_tiles (the reference's traversal of the 32-atom tile list:
    entries, shifted i tiles, j tiles, exclusion words) equals the
    brute-force pair sum over all images with the same exclusions, for a list
    built by gpu_recip_amd.nlist (complete by test_nlist.py)."""
    from gpu_recip_amd import nlist, systems

    box = (22.0, 23.0, 24.0)
    x = systems.triplets(300, box, seed=21)
    x[:, :3] = np.mod(x[:, :3].astype(np.float64), box).astype(np.float32)
    water = nlist.water_exclusions(len(x))
    extra = np.array([[0, 10], [33, 70], [5, 200]], np.int64)
    excl = np.concatenate([water, extra])
    kappa, roff = 0.34, 9.0
    tl = nlist.build(x, box, roff, excl_pairs=excl)
    t = O.direct_tiles(tl.xyzq, box, kappa, roff, tl.ientry, tl.tile_indj, tl.tile_excl)
    b = O.direct_ewald(x, box, kappa, roff, excl=excl)
    fb = np.zeros((tl.npad, 3))
    fb[tl.sorted_index()] = b["force"]
    assert abs(t["energy"] - b["energy"]) <= 1e-10 * abs(b["energy"])
    assert np.abs(t["force"] - fb).max() <= 1e-9 * np.abs(fb).max()
    assert np.allclose(t["virial"], b["virial"], rtol=1e-10, atol=1e-8)
    # the shift forces reproduce the virial: sum_i x_i (x) g_i + sum_s s (x) sforce_s
    xs = tl.xyzq[:, :3].astype(np.float64)
    w = xs.T @ t["force"] + (nlist.SHIFTS * np.asarray(box)).T @ t["sforce"]
    assert np.allclose(w.reshape(9), b["virial"], rtol=1e-9, atol=1e-7)
    # dropping one exclusion word's bits changes the sum (the words are read)
    te = tl.tile_excl.copy()
    k = int(np.argmax((te != 0xFFFFFFFF).any(1) & (te != 0).any(1)))
    te[k] = 0
    t2 = O.direct_tiles(tl.xyzq, box, kappa, roff, tl.ientry, tl.tile_indj, te)
    assert t2["energy"] != t["energy"]
