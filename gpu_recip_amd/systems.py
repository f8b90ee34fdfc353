"""Synthetic PME systems: seeded TIP3P-like water triplets in orthorhombic boxes.

Geometry: O-H 0.9572 A, H-O-H 104.52 deg, charges O -0.834 e, H +0.417 e
(net neutral per triplet), molecule centres uniform in the box, orientations
uniform on SO(3) (random unit quaternions).  Coordinates are left unwrapped;
the grid index formula is periodic.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

Q_O, Q_H = -0.834, 0.417
R_OH = 0.9572
THETA_HOH = math.radians(104.52)
CCELEC = 332.0716
SEED_BASE = 20261202


@dataclass
class PmeSystem:
    name: str
    xyzq: np.ndarray          # (N, 4) float32
    box: tuple                # (Lx, Ly, Lz), A
    nfft: tuple               # (nx, ny, nz)
    order: int
    kappa: float

    @property
    def natom(self) -> int:
        return int(self.xyzq.shape[0])

    @property
    def recip(self) -> list:
        r = [0.0] * 9
        r[0], r[4], r[8] = 1.0 / self.box[0], 1.0 / self.box[1], 1.0 / self.box[2]
        return r

    def algorithmic_bytes(self) -> int:
        """Compulsory FP32 I/O of one solve: 44 N + 16 G + 32 Gc bytes
        (each stage reads its inputs and writes its outputs once)."""
        nx, ny, nz = self.nfft
        G = nx * ny * nz
        Gc = (nx // 2 + 1) * ny * nz
        return 44 * self.natom + 16 * G + 32 * Gc


def random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """n rotation matrices uniform on SO(3) from uniform unit quaternions."""
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def triplets(ntrip: int, box, seed: int) -> np.ndarray:
    """(3*ntrip, 4) float32 water-like triplets, net neutral."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 1.0, (ntrip, 3)) * np.asarray(box, np.float64)
    R = random_rotations(rng, ntrip)
    s, c = math.sin(THETA_HOH / 2), math.cos(THETA_HOH / 2)
    local = np.array([[0.0, 0.0, 0.0], [R_OH * s, R_OH * c, 0.0], [-R_OH * s, R_OH * c, 0.0]])
    pos = centres[:, None, :] + np.einsum("nij,kj->nki", R, local)
    out = np.empty((ntrip, 3, 4), np.float32)
    out[:, :, :3] = pos
    out[:, 0, 3] = Q_O
    out[:, 1:, 3] = Q_H
    return out.reshape(-1, 4)


def smooth235_at_least(x: float) -> int:
    n = max(1, int(math.ceil(x - 1e-9)))
    while True:
        m = n
        for f in (2, 3, 5):
            while m % f == 0:
                m //= f
        if m == 1:
            return n
        n += 1


# ---- the benchmark configurations ------------------------------------------

def config_plain_small() -> PmeSystem:
    """Plain FP32 solve: 2,667 triplets (8,001 atoms), 40 A cubic, 48^3, order 4, kappa 0.34."""
    box = (40.0, 40.0, 40.0)
    return PmeSystem("plain-8k-48c-p4", triplets(2667, box, SEED_BASE + 0), box, (48, 48, 48), 4, 0.34)


def config_production() -> PmeSystem:
    """Production plain solve: 32,000 triplets (96,000 atoms), box 100x96x90 A,
    grid 108x100x96, order 6, kappa 0.32."""
    box = (100.0, 96.0, 90.0)
    return PmeSystem("plain-96k-108x100x96-p6", triplets(32000, box, SEED_BASE + 1), box, (108, 100, 96), 6, 0.32)


def config_batched(nsys: int = 64) -> list[PmeSystem]:
    """64 systems: atoms log-uniform in [2000, 40000] (triplets), edges U(30, 75) A,
    grid edge = smallest 2/3/5-smooth >= edge/0.9, order 4, kappa 0.34."""
    rng = np.random.default_rng(SEED_BASE + 2)
    out = []
    for i in range(nsys):
        natom = math.exp(rng.uniform(math.log(2000), math.log(40000)))
        ntrip = max(1, int(round(natom / 3)))
        box = tuple(float(v) for v in rng.uniform(30.0, 75.0, 3))
        nfft = tuple(smooth235_at_least(L / 0.9) for L in box)
        out.append(PmeSystem(f"batch{i}", triplets(ntrip, box, SEED_BASE + 1000 + i), box, nfft, 4, 0.34))
    return out


def config_split_precision() -> PmeSystem:
    """250,002 atoms (83,334 triplets), 135 A cubic, 144^3, order 4, kappa 0.30."""
    box = (135.0, 135.0, 135.0)
    return PmeSystem("split-250k-144c-p4", triplets(83334, box, SEED_BASE + 3), box, (144, 144, 144), 4, 0.30)


def reference_fixture_system(xyzq: np.ndarray, order: int) -> PmeSystem:
    """The reference's gpu_recip test system (gpu_recip.cu:71-80): 62.23 A, 64^3, kappa 0.32."""
    box = (62.23, 62.23, 62.23)
    return PmeSystem(f"ref23558-p{order}", np.ascontiguousarray(xyzq, np.float32), box, (64, 64, 64), order, 0.32)
