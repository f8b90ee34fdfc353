"""Helpers for the GPU tests: run the HIP path through the C ABI on torch device
buffers and compare with the oracle."""
import sys

import numpy as np

from gpu_recip_amd import _lib as L
from gpu_recip_amd.pme import CudaEnergyVirial, CudaPMERecip, force_stride


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def torch_dev():
    import torch

    assert torch.cuda.is_available(), "GPU test needs a ROCm device"
    return torch, torch.device("cuda:0")


def to_dev(x):
    torch, dev = torch_dev()
    return torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)


def run_solve(nfft, order, kappa, box, xyzq, calc_ev=True, force_type=L.FORCE_FLOAT, five_calls=True,
              self_energy=False):
    """One solve on the GPU.  Returns dict(force (N,3) float64, energy, virial (9,), self)."""
    torch, dev = torch_dev()
    n = len(xyzq)
    say(f"  gpu solve nfft={nfft} order={order} natom={n} ft={force_type}")
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(*nfft, order, ev)
    recip = [1 / box[0], 0, 0, 0, 1 / box[1], 0, 0, 0, 1 / box[2]]
    x = to_dev(xyzq)
    stride = force_stride(n)
    if force_type == L.FORCE_INT64:
        f = torch.zeros(3 * stride, dtype=torch.int64, device=dev)
    elif force_type == L.FORCE_FLOAT3:
        f = torch.zeros(3 * n, dtype=torch.float32, device=dev)
    else:
        f = torch.zeros(3 * stride, dtype=torch.float32, device=dev)
    ev.clear()
    if five_calls:
        pme.spread_charge(x, n, recip)
        pme.r2c_fft()
        pme.scalar_sum(recip, kappa, calc_ev, calc_ev)
        pme.c2r_fft()
        pme.gather_force(x, n, recip, stride, f, force_type)
    else:
        pme.solve(x, n, recip, kappa, calc_ev, calc_ev, stride, f, force_type)
    if self_energy:
        pme.calc_self_energy(x, n, kappa)
    ev.copyToHost()
    torch.cuda.synchronize()
    fh = f.cpu().numpy()
    if force_type == L.FORCE_FLOAT3:
        force = fh.reshape(n, 3).astype(np.float64)
    elif force_type == L.FORCE_INT64:
        force = np.stack([fh[:n], fh[stride:stride + n], fh[2 * stride:2 * stride + n]], 1).astype(np.float64) / 2.0 ** 40
    else:
        force = np.stack([fh[:n], fh[stride:stride + n], fh[2 * stride:2 * stride + n]], 1).astype(np.float64)
    return {"force": force, "energy": ev.getEnergy("recip"), "virial": np.array(ev.getVirial()),
            "self": ev.getEnergy("self"), "pme": pme, "ev": ev, "x": x}


def rel_rms(a, b):
    """||a - b|| / ||b|| (real or complex)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.sqrt((np.abs(a - b) ** 2).sum() / max((np.abs(b) ** 2).sum(), 1e-300)))


def rel_err(a, b):
    return float(abs(a - b) / max(abs(b), 1e-300))


def rel_fro(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))
