"""The direct-space kernels' erfc (erfc_from_exp in pme_direct.hip): t P(t)
e^{-x^2} with t = 1 / (1 + p x), coefficients read from the kernel source and
evaluated in FP32 as the kernel does, against SciPy's FP64 erfc.  Relative
error below 1e-6 on [0, 6] (kappa r of every pair inside a cutoff of up to
six Ewald widths); the FP32 evaluation itself is about 4e-7."""
import os
import re

import numpy as np
from scipy.special import erfc

SRC = os.path.join(os.path.dirname(__file__), "..", "gpu_recip_amd", "csrc", "pme_direct.hip")


def _coeffs():
    text = open(SRC).read()
    p = float(re.search(r"kErfcP\s*=\s*([-0-9.eE+]+)f", text).group(1))
    body = re.search(r"kErfcC\[\d+\]\s*=\s*\{([^}]*)\}", text).group(1)
    c = [float(v.strip().rstrip("f")) for v in body.split(",") if v.strip()]
    return np.float32(p), np.array(c, np.float32)


def _erfc32(x, p, c):
    x = np.asarray(x, np.float32)
    t = (np.float32(1) / (p * x + np.float32(1))).astype(np.float32)
    acc = np.full_like(x, c[-1])
    for k in range(len(c) - 2, -1, -1):
        acc = (acc * t + c[k]).astype(np.float32)
    e = np.exp(-(x.astype(np.float64) ** 2)).astype(np.float32)  # expf: correctly rounded here
    return (acc * t * e).astype(np.float32)


def test_erfc_relative_error():
    p, c = _coeffs()
    assert len(c) == 10
    x = np.linspace(0.0, 6.0, 400001).astype(np.float32)
    got = _erfc32(x, p, c).astype(np.float64)
    ref = erfc(x.astype(np.float64))
    rel = np.abs(got / ref - 1.0)
    assert rel.max() < 1e-6, rel.max()
    assert got[0] == np.float32(1.0) or abs(got[0] - 1.0) < 4e-7


def test_erfc_pair_energy_error_is_negligible():
    """The potential qq erfc(kappa r)/r at kappa = 0.34, r in [0.8, 12] A:
    absolute error per unit charge product below 2e-6 kcal/mol."""
    p, c = _coeffs()
    kappa = 0.34
    r = np.linspace(0.8, 12.0, 100001)
    got = 332.0716 * _erfc32((kappa * r).astype(np.float32), p, c).astype(np.float64) / r
    ref = 332.0716 * erfc(np.float32(kappa) * r.astype(np.float32).astype(np.float64)) / r
    assert np.abs(got - ref).max() < 2e-6 * 332.0716
