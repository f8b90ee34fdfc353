"""Pins the FP64 oracle (oracle/pme_oracle.c) before it is trusted as the checker.

Each check is independent of the oracle's own formulas: the reference's golden
data (gpu_recip.cu fixtures, test_fbfft.cu goldens), closed forms, brute-force
Ewald and finite differences.  CPU only.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _fixture():
    d = np.load(os.path.join(GOLD, "recip_system_23558.npz"))
    with open(os.path.join(GOLD, "recip_constants.json")) as f:
        c = json.load(f)
    return d, c


# --- end to end against the reference's gpu_recip test (gpu_recip.cu:68-293, 298-522)

@pytest.mark.parametrize("order", [4, 6])
@pytest.mark.parametrize("index_mode", [O.INDEX_F64, O.INDEX_F32])
def test_reference_fixture(order, index_mode):
    d, c = _fixture()
    k = c[f"order{order}"]
    p = O.OracleParams(*c["nfft"], order, c["kappa"], tuple(c["box"]), index_mode=index_mode)
    r = O.solve(p, d["xyzq"])
    # the reference's own absolute tolerances
    assert abs(r["energy"] - k["energy"]) < k["tol_energy"]
    assert np.abs(r["virial"] - np.array(k["virial"])).max() < k["tol_virial"]
    fref = d[f"force_recip_{order}"]
    assert np.abs(r["force"] - fref).max() < k["tol_force"]
    if index_mode == O.INDEX_F64:
        rel = np.sqrt(((r["force"] - fref) ** 2).sum() / (fref ** 2).sum())
        assert rel < 1e-4


# --- FFT: test_fbfft.cu goldens (forward, unnormalised, e^{-2 pi i kn/N})

def test_fft_goldens_1d():
    g = np.load(os.path.join(GOLD, "fft_goldens.npz"))
    assert np.abs(O.fft1d(g["in64"].astype(complex), -1) - g["out64"]).max() < 1e-5
    for b in range(2):
        seg = slice(64 * b, 64 * (b + 1))
        assert np.abs(O.fft1d(g["in64x2"][seg].astype(complex), -1) - g["out64x2"][seg]).max() < 1e-5


@pytest.mark.parametrize("n", [4, 64])
def test_fft_goldens_2d_r2c(n):
    g = np.load(os.path.join(GOLD, "fft_goldens.npz"))
    a = g[f"in{n}x{n}"].reshape(n, n)          # [y][x], x fastest
    ref = g[f"out{n}x{n}"].reshape(n, n)
    c = O.fft3d_r2c(a[None].copy())
    assert np.abs(c[0] - ref[:, : n // 2 + 1]).max() < 1e-4


@pytest.mark.parametrize("n", [2, 3, 5, 7, 12, 36, 45, 48, 75, 96, 100, 108, 144])
def test_fft_brute_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)
    assert np.abs(O.fft1d(x, -1) - W @ x).max() < 1e-10
    assert np.abs(O.fft1d(x, +1) - W.conj() @ x).max() < 1e-10


@pytest.mark.parametrize("shape", [(6, 5, 4), (9, 10, 15), (12, 8, 7)])
def test_fft3d_brute_and_roundtrip(shape):
    nz, ny, nx = shape
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal(shape)
    c = O.fft3d_r2c(a)
    # brute 3-D DFT of the real grid, half spectrum
    full = np.zeros(shape, complex)
    for kz in range(nz):
        for ky in range(ny):
            for kx in range(nx // 2 + 1):
                ph = np.exp(-2j * np.pi * (kx * np.arange(nx)[None, None, :] / nx
                                           + ky * np.arange(ny)[None, :, None] / ny
                                           + kz * np.arange(nz)[:, None, None] / nz))
                full[kz, ky, kx] = (a * ph).sum()
    assert np.abs(c - full[:, :, : nx // 2 + 1]).max() < 1e-9
    # c2r(r2c(a)) = G * a (both unnormalised)
    back = O.fft3d_c2r(c, nx)
    assert np.abs(back - a * a.size).max() < 1e-9 * a.size
    # Parseval on the half spectrum
    w = np.full(nx // 2 + 1, 2.0)
    w[0] = 1.0
    if nx % 2 == 0:
        w[-1] = 1.0
    assert np.isclose((np.abs(c) ** 2 * w).sum(), a.size * (a ** 2).sum(), rtol=1e-10)


# --- B-splines: closed forms of M_4 and M_6 (Essmann et al. 1995 cardinal B-splines)

def _M(p, u):
    """Cardinal B-spline M_p(u) by the explicit sum formula (Schoenberg), evaluated
    in exact rational arithmetic (the alternating sum cancels badly in floats)."""
    from fractions import Fraction
    from math import comb, factorial
    u = Fraction(u)
    s = sum((-1) ** k * comb(p, k) * max(u - k, Fraction(0)) ** (p - 1) for k in range(p + 1))
    return float(s / factorial(p - 1))


@pytest.mark.parametrize("order", [4, 6, 8])
@pytest.mark.parametrize("w", [0.0, 0.1, 0.37, 0.5, 0.999])
def test_bspline_closed_form(order, w):
    th, dth = O.bspline(order, w)
    for j in range(order):
        u = w + order - 1 - j
        assert abs(th[j] - _M(order, u)) < 1e-13
        # derivative: M_p'(u) = M_{p-1}(u) - M_{p-1}(u-1)
        assert abs(dth[j] - (_M(order - 1, u) - _M(order - 1, u - 1))) < 1e-13
    assert abs(th.sum() - 1.0) < 1e-14
    assert abs(dth.sum()) < 1e-14


@pytest.mark.parametrize("n", [48, 64, 100, 45])
def test_bsp_mod_closed_form(n):
    th = 2 * np.pi * np.arange(n) / n
    # |sum M_p(j) e^{i j th}| for symmetric node sets
    b4 = 9.0 / (2.0 + np.cos(th)) ** 2
    assert np.allclose(O.bsp_mod(4, n), b4, rtol=1e-12)
    b6 = (120.0 / (66.0 + 52.0 * np.cos(th) + 2.0 * np.cos(2 * th))) ** 2
    assert np.allclose(O.bsp_mod(6, n), b6, rtol=1e-12)


# --- index mapping: fr = n * frac(x*r + 0.5) (half-box shifted), CudaPMERecip.cu:512-529

def test_index_closed_form():
    rng = np.random.default_rng(1)
    L = np.array([40.0, 37.5, 52.0])
    n = (48, 45, 64)
    x = np.zeros((2000, 4), np.float32)
    x[:, :3] = rng.uniform(-2 * L, 3 * L, size=(2000, 3))
    idx, fr = O.index_f64(x, 1.0 / L, n)
    u = np.array(n) * np.mod(x[:, :3].astype(np.float64) / L + 0.5, 1.0)
    assert np.allclose(idx + fr, u, atol=1e-9)
    assert idx.min() >= 0 and (idx < np.array(n)).all()
    i32, f32 = O.index_f32(x, 1.0 / L, n)
    # FP32 agrees with FP64 except within FP32 rounding of the cell boundary
    tot32 = i32 + f32.astype(np.float64)
    d = np.abs(tot32 - u)
    d = np.minimum(d, np.abs(d - np.array(n)))
    assert d.max() < 1e-4


def boundary_coords(n, L, per_cell=3, images=(-2, -1, 0, 1, 2)):
    """Coordinates at and one FP32 ulp either side of every shifted cell boundary
    x = (k/n - 1/2) L + m L -- the inputs where an FP32 index can flip."""
    out = []
    for m in images:
        for k in range(n):
            x0 = np.float32((k / n - 0.5) * L + m * L)
            v = x0
            for _ in range(per_cell):
                v = np.nextafter(v, np.float32(-np.inf))
            for _ in range(2 * per_cell + 1):
                out.append(v)
                v = np.nextafter(v, np.float32(np.inf))
    return np.array(out, np.float32)


def test_index_f32_boundaries():
    # the FP32 index of CudaPMERecip.cu:512-529 stays in [0, n); the only rounding
    # artefact is a fraction a few ulp below zero when w + 0.5f rounds up to an
    # integer (then i = 0).  i = n would need n*(1 - ulp) to round to n, which
    # FP32 never does for this formula, so the wrap of :566-568 is a guard only.
    for n, L in ((64, 62.23), (48, 40.0), (45, 37.7), (108, 100.0)):
        xs = boundary_coords(n, L)
        x = np.zeros((len(xs), 4), np.float32)
        x[:, 0] = x[:, 1] = x[:, 2] = xs
        i, f = O.index_f32(x, (1 / L,) * 3, (n, n, n))
        assert i.min() >= 0 and i.max() < n
        assert f.min() > -1e-5 and f.max() < 1.0
        u = n * np.mod(xs.astype(np.float64) / L + 0.5, 1.0)
        d = np.abs(i[:, 0] + f[:, 0].astype(np.float64) - u)
        d = np.minimum(d, n - d)
        assert d.max() < 1e-3


# --- spread invariants

def test_spread_charge_and_centroid():
    rng = np.random.default_rng(7)
    L = 30.0
    n = 36
    x = np.zeros((50, 4), np.float32)
    x[:, :3] = rng.uniform(0, L, (50, 3))
    x[:, 3] = rng.uniform(-1, 1, 50)
    for order in (4, 6):
        p = O.OracleParams(n, n, n, order, 0.34, (L, L, L))
        for a in range(5):
            g = O.spread(p, x[a : a + 1])
            assert np.isclose(g.sum(), x[a, 3], rtol=1e-12)
            # the charge-weighted stencil centroid sits p/2 - 1 points above
            # u = n*frac(x/L + 0.5) (theta[j] = M_p(w+p-1-j) at point i+j)
            u = n * np.mod(x[a, :3].astype(np.float64) / L + 0.5, 1.0)
            # g is [z][y][x]; first moment of M_p sampled at integers is exact
            for dim, axes in ((0, (0, 1)), (1, (0, 2)), (2, (1, 2))):
                prof = g.sum(axis=axes)
                expect = u[dim] + order / 2.0 - 1.0
                k = np.arange(n, dtype=np.float64)
                k_unwrapped = expect + (k - expect + n / 2) % n - n / 2
                mean = (prof * k_unwrapped).sum() / prof.sum()
                assert abs(mean - expect) < 1e-9
                assert np.count_nonzero(np.abs(prof) > 1e-15) <= order


# --- against brute-force Ewald (no grid, no splines)

def _few_charges(m, L, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((m, 4), np.float32)
    x[:, :3] = rng.uniform(0, L, (m, 3))
    q = rng.uniform(-1, 1, m)
    x[:, 3] = q - q.mean()
    return x


@pytest.mark.parametrize("order", [6, 8])
def test_against_brute_ewald(order):
    L = (12.0, 10.0, 14.0)
    x = _few_charges(8, 10.0, 3)
    kappa = 0.5
    p = O.OracleParams(48, 40, 54, order, kappa, L, nyquist_mode=O.NYQUIST_EXACT)
    r = O.solve(p, x)
    e = O.ewald_recip(x, L, kappa, (14, 12, 16))
    assert abs(r["energy"] - e["energy"]) < 1e-6 * max(1.0, abs(e["energy"]))
    rel = np.linalg.norm(r["force"] - e["force"]) / np.linalg.norm(e["force"])
    assert rel < 1e-5
    assert np.abs(r["virial"] - e["virial"]).max() < 1e-5 * np.abs(e["virial"]).max()


def test_brute_ewald_force_is_gradient():
    L = (10.0, 10.0, 10.0)
    x = _few_charges(5, 10.0, 9)
    kappa, km = 0.4, (8, 8, 8)
    e0 = O.ewald_recip(x, L, kappa, km)
    h = 1e-3
    for a in range(2):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, d] += h
            xm[a, d] -= h
            fd = (O.ewald_recip(xp, L, kappa, km, False)["energy"] - O.ewald_recip(xm, L, kappa, km, False)["energy"]) / (
                float(xp[a, d]) - float(xm[a, d]))
            assert abs(fd - e0["force"][a, d]) < 1e-5 * max(1, abs(fd))


# --- gather = +dE/dr (sign and scale), CudaPMERecip.cu:1238-1241

@pytest.mark.parametrize("order", [4, 6])
def test_gather_is_energy_gradient(order):
    L = (20.0, 18.0, 22.0)
    x = _few_charges(12, 18.0, 11)
    p = O.OracleParams(24, 20, 25, order, 0.35, L, nyquist_mode=O.NYQUIST_EXACT)
    r = O.solve(p, x)
    h = 2e-3
    for a in (0, 5):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, d] += h
            xm[a, d] -= h
            fd = (O.solve(p, xp)["energy"] - O.solve(p, xm)["energy"]) / (float(xp[a, d]) - float(xm[a, d]))
            assert abs(fd - r["force"][a, d]) < 2e-5 * max(1.0, np.abs(r["force"]).max())


def test_virial_diagonal_is_strain_derivative():
    # W_aa = -dE/d eps_a under x_a -> x_a (1+eps), L_a -> L_a (1+eps)
    L = np.array([20.0, 18.0, 22.0])
    x = _few_charges(12, 18.0, 21)
    p0 = O.OracleParams(24, 20, 25, 6, 0.35, tuple(L), nyquist_mode=O.NYQUIST_EXACT)
    v = O.solve(p0, x)["virial"].reshape(3, 3)
    eps = 1e-5
    for a in range(3):
        es = []
        for s in (+1, -1):
            Ls = L.copy()
            Ls[a] *= 1 + s * eps
            es.append(_energy_with_box(p0, x, Ls))
        dE = (es[0] - es[1]) / (2 * eps)
        assert abs(v[a, a] + dE) < 1e-6 * max(1.0, abs(dE))


def _energy_with_box(p, x, Ls):
    """Energy at fixed fractional coordinates with box Ls: spread with the original
    recip (same grid charges), influence function with the scaled box."""
    g = O.spread(p, x)
    c = O.fft3d_r2c(g)
    q = O.OracleParams(p.nx, p.ny, p.nz, p.order, p.kappa, tuple(Ls), nyquist_mode=p.nyquist_mode)
    _, e, _ = O.scalar_sum(q, c)
    return e


def test_self_energy_closed_form():
    x = np.array([[1, 2, 3, 0.5], [4, 5, 6, -1.0]], np.float32)
    assert np.isclose(O.self_energy(x, 0.34), -0.34 * O.CCELEC / np.sqrt(np.pi) * (0.25 + 1.0), rtol=1e-14)


def test_nyquist_weight_difference_small():
    # the reference double-counts the k1 = nx/2 plane (CudaPMERecip.cu:708); the
    # difference to the exact weight is reported, and tiny at these kappa/grids
    d, c = _fixture()
    p = O.OracleParams(64, 64, 64, 4, 0.32, (62.23,) * 3)
    e_ref = O.solve(p, d["xyzq"])["energy"]
    p.nyquist_mode = O.NYQUIST_EXACT
    e_ex = O.solve(p, d["xyzq"])["energy"]
    assert e_ref >= e_ex
    assert abs(e_ref - e_ex) < 1e-8 * abs(e_ex)


# ---------------------------------------------------------------- triclinic --
def _tric_recip(H):
    """recip (row-major) of a box whose columns are the box vectors: H^-1."""
    return np.linalg.inv(np.asarray(H, np.float64)).reshape(9)


def test_brute_ewald_is_lattice_basis_independent():
    """A sheared basis of the same cubic lattice (c' = c + a) gives the same
    brute-force reciprocal energy: the general m-hat of orc_ewald_recip9."""
    rng = np.random.default_rng(21)
    x = np.zeros((6, 4), np.float32)
    x[:, :3] = rng.uniform(0, 15.0, (6, 3))
    x[:, 3] = [0.5, -0.5, 0.25, -0.25, 0.75, -0.75]
    cube = O.ewald_recip(x, (15.0, 15.0, 15.0), 0.35, (10, 10, 10))
    sheared = O.ewald_recip9(x, _tric_recip([[15.0, 0.0, 15.0], [0.0, 15.0, 0.0], [0.0, 0.0, 15.0]]), 0.35,
                             (20, 10, 10))
    assert abs(sheared["energy"] - cube["energy"]) <= 1e-10 * abs(cube["energy"])
    assert np.allclose(sheared["force"], cube["force"], rtol=0, atol=1e-9)
    assert np.allclose(sheared["virial"], cube["virial"], rtol=0, atol=1e-6)  # k-space truncations differ


def _tric_params(H, n, order, kappa, mode=O.INDEX_F64):
    p = O.OracleParams(n[0], n[1], n[2], order, kappa, (1.0, 1.0, 1.0), index_mode=mode,
                       nyquist_mode=O.NYQUIST_EXACT, recip=_tric_recip(H))
    return p


H_TRIC = [[24.0, 3.0, -2.0], [0.0, 22.0, 4.0], [0.0, 0.0, 20.0]]


def test_triclinic_pme_matches_brute_ewald():
    rng = np.random.default_rng(22)
    x = np.zeros((10, 4), np.float32)
    frac = rng.uniform(0, 1, (10, 3))
    x[:, :3] = frac @ np.asarray(H_TRIC).T
    q = rng.choice([-0.75, -0.5, 0.5, 0.75], 10)
    q[-1] -= q.sum()
    x[:, 3] = q
    kappa = 0.3
    o = O.solve(_tric_params(H_TRIC, (64, 64, 60), 8, kappa), x)
    b = O.ewald_recip9(x, _tric_recip(H_TRIC), kappa, (14, 14, 14))
    assert abs(o["energy"] - b["energy"]) <= 1e-6 * abs(b["energy"])
    assert np.sqrt(((o["force"] - b["force"]) ** 2).sum() / (b["force"] ** 2).sum()) < 1e-5
    assert np.linalg.norm(o["virial"] - b["virial"]) <= 1e-5 * np.linalg.norm(b["virial"])


def test_triclinic_pme_forces_are_its_energy_gradient():
    rng = np.random.default_rng(23)
    x = np.zeros((6, 4), np.float32)
    x[:, :3] = rng.uniform(0, 1, (6, 3)) @ np.asarray(H_TRIC).T
    x[:, 3] = [0.5, -0.5, 0.25, -0.25, 0.75, -0.75]
    p = _tric_params(H_TRIC, (32, 30, 30), 6, 0.35)
    o = O.solve(p, x)
    h = 1e-3
    for i in range(6):
        for a in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, a] += h
            xm[i, a] -= h
            dx = float(xp[i, a]) - float(xm[i, a])
            fd = (O.solve(p, xp)["energy"] - O.solve(p, xm)["energy"]) / dx
            assert abs(fd - o["force"][i, a]) <= 1e-3 * max(1.0, abs(o["force"][i, a]))


def test_orthorhombic_recip_takes_the_reference_path():
    """A diagonal recip given as a general matrix: the same indices and results."""
    rng = np.random.default_rng(24)
    x = np.zeros((50, 4), np.float32)
    x[:, :3] = rng.uniform(-5, 35, (50, 3))
    x[:, 3] = np.tile([0.5, -0.5], 25)
    p = O.OracleParams(32, 32, 32, 4, 0.34, (30.0, 31.0, 32.0), index_mode=O.INDEX_F32)
    o1 = O.solve(p, x)
    p2 = O.OracleParams(32, 32, 32, 4, 0.34, (1.0, 1.0, 1.0), index_mode=O.INDEX_F32,
                        recip=np.array([1 / 30.0, 0, 0, 0, 1 / 31.0, 0, 0, 0, 1 / 32.0]))
    o2 = O.solve(p2, x)
    assert np.array_equal(o1["force"], o2["force"]) and o1["energy"] == o2["energy"]
