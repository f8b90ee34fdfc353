"""GPU tests of triclinic boxes on the plain path: the full recip matrix in the
grid index, m-hat, the volume and the gather transform.  The reference exits
on non-orthorhombic boxes (CudaPMERecip.cu:2042-2045, 2240-2249) and holds no
triclinic fixture, so the oracle (pinned by the brute-force Ewald sum, lattice
invariance and finite differences) is the check: the FP32 index of the
oracle's restatement, forces within 1e-5 relative RMS, energy and virial within
1e-5; against the FP64-index oracle within 1e-4."""
import numpy as np
import pytest

import oracle as O
from gpu_recip_amd import _lib as L
from gpu_recip_amd import systems
from gpu_recip_amd.pme import CudaEnergyVirial, CudaPMERecip, force_stride
from gpu_util import rel_err, rel_fro, rel_rms, say, to_dev, torch_dev

pytestmark = pytest.mark.gpu

BOXES = {
    "monoclinic": [[40.0, 0.0, 8.0], [0.0, 38.0, 0.0], [0.0, 0.0, 36.0]],
    "triclinic": [[44.0, 6.0, -5.0], [0.0, 41.0, 7.0], [0.0, 0.0, 39.0]],
}


def _system(H, ntrip, seed):
    """Water-like triplets: oxygens uniform in the cell (fractional coordinates
    mapped through H), each molecule's own geometry kept in Angstrom."""
    H = np.asarray(H, np.float64)
    x = systems.triplets(ntrip, (1.0, 1.0, 1.0), seed=seed)
    xo = np.repeat(x[0::3, :3].astype(np.float64), 3, axis=0)   # the oxygen of each atom's molecule
    x[:, :3] = (xo @ H.T + (x[:, :3] - xo)).astype(np.float32)
    return x, np.linalg.inv(H).reshape(9)


def _gpu(nfft, order, kappa, recip, x, five_calls=True):
    torch, dev = torch_dev()
    n = len(x)
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(*nfft, order, ev)
    xd = to_dev(x)
    st = force_stride(n)
    f = torch.zeros(3 * st, dtype=torch.float32, device=dev)
    ev.clear()
    if five_calls:
        pme.spread_charge(xd, n, recip)
        pme.r2c_fft()
        pme.scalar_sum(recip, kappa, True, True)
        pme.c2r_fft()
        pme.gather_force(xd, n, recip, st, f, L.FORCE_FLOAT)
    else:
        pme.solve(xd, n, recip, kappa, True, True, st, f, L.FORCE_FLOAT)
    ev.copyToHost()
    torch.cuda.synchronize()
    fh = f.cpu().numpy()
    force = np.stack([fh[:n], fh[st:st + n], fh[2 * st:2 * st + n]], 1).astype(np.float64)
    return force, ev.getEnergy("recip"), np.array(ev.getVirial())


@pytest.mark.parametrize("name", sorted(BOXES))
@pytest.mark.parametrize("order,nfft", [(4, (48, 45, 40)), (6, (45, 40, 36))])
def test_triclinic_solve_against_oracle(name, order, nfft):
    x, recip = _system(BOXES[name], 2000, seed=31 + order)
    kappa = 0.34
    say(f"  {name} box, {len(x)} atoms, grid {nfft}, order {order}")
    f, e, v = _gpu(nfft, order, kappa, recip, x)
    p32 = O.OracleParams(*nfft, order, kappa, (1.0, 1.0, 1.0), index_mode=O.INDEX_F32, recip=recip)
    o = O.solve(p32, x)
    frms = rel_rms(f, o["force"])
    say(f"    vs FP32-index oracle: force rel-RMS {frms:.2e}, energy {rel_err(e, o['energy']):.2e}, "
        f"virial {rel_fro(v, o['virial']):.2e}")
    assert frms < 1e-5
    assert rel_err(e, o["energy"]) < 1e-5
    assert rel_fro(v, o["virial"]) < 1e-5
    o64 = O.solve(O.OracleParams(*nfft, order, kappa, (1.0, 1.0, 1.0), recip=recip), x)
    assert rel_rms(f, o64["force"]) < 1e-4
    # the whole-solve entry point gives the same result
    f2, e2, v2 = _gpu(nfft, order, kappa, recip, x, five_calls=False)
    assert np.array_equal(f, f2) and e == e2 and np.array_equal(v, v2)


def test_sheared_basis_of_an_orthorhombic_lattice():
    """The same periodic system given by an orthorhombic box and by a sheared
    basis of the same lattice (c' = c + a): energies and forces agree to the
    PME discretisation error."""
    L0 = 36.0
    x = systems.triplets(1500, (L0, L0, L0), seed=41)
    kappa = 0.34
    fo, eo, _ = _gpu((40, 40, 40), 6, kappa, [1 / L0, 0, 0, 0, 1 / L0, 0, 0, 0, 1 / L0], x)
    H = np.array([[L0, 0.0, L0], [0.0, L0, 0.0], [0.0, 0.0, L0]])
    ft, et, _ = _gpu((40, 40, 40), 6, kappa, np.linalg.inv(H).reshape(9), x)
    say(f"  sheared vs orthorhombic: energy {rel_err(et, eo):.2e}, force rel-RMS {rel_rms(ft, fo):.2e}")
    assert rel_err(et, eo) < 1e-3
    assert rel_rms(ft, fo) < 1e-2


def test_bad_box_rejected():
    torch, dev = torch_dev()
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(32, 32, 32, 4, ev)
    x = to_dev(np.zeros((10, 4), np.float32))
    with pytest.raises(L.PmeError):
        pme.spread_charge(x, 10, [1 / 30.0, 0, 0, 0, 0.0, 0, 0, 0, 1 / 30.0])  # singular
    with pytest.raises(L.PmeError):
        pme.spread_charge(x, 10, [-1 / 30.0, 0, 0, 0, 1 / 30.0, 0, 0, 0, 1 / 30.0])  # left-handed
