"""Serving workload: open-loop Poisson arrivals of PME solves of uneven sizes.

Each request is a water-like triplet system (the generator of systems.py,
here on the device so that thousands of inputs can be made quickly): atom
count log-uniform in [nmin, nmax] (rounded to triplets), cubic box of side
(N / density)^(1/3), grid edge = smallest 2/3/5-smooth >= side / spacing,
B-spline order 4 or 6 with equal probability, kappa 0.34.  A fraction of
the requests are follow-up steps on an earlier system: the same key, the
earlier coordinates displaced by N(0, sigma) in every coordinate.

The driver submits every request at its arrival time (the submission does
not wait for earlier results), and the server reports per-request latency
from submission to the result being host-visible.
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from . import systems
from .pme import PmeServer, force_stride

SERVE_SEED = systems.SEED_BASE + 4


@dataclass
class ServeRequest:
    t: float             # arrival time (s from the start)
    key: int             # system key
    natom: int
    box: float           # cubic edge, A
    nfft: int            # cubic grid edge
    order: int
    kappa: float
    followup: bool
    base: int = -1       # index of the request whose coordinates are displaced (follow-ups)
    xyzq: object = None  # device tensor (N, 4)
    force: object = None
    stride: int = 0
    ticket: int = 0
    result: tuple = field(default=None, repr=False)


def schedule(seconds: float, rate: float = 300.0, nmin: int = 5000, nmax: int = 400000, density: float = 0.1,
             spacing: float = 1.0, followup_frac: float = 0.2, kappa: float = 0.34, seed: int = SERVE_SEED,
             recent: int = 32, max_requests: int | None = None):
    """The request list (arrival times, sizes, keys) for `seconds` of arrivals
    (at most max_requests)."""
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    new_idx = []
    next_key = 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds or (max_requests is not None and len(out) >= max_requests):
            break
        if new_idx and rng.uniform() < followup_frac:
            b = new_idx[-recent:][int(rng.integers(0, min(recent, len(new_idx))))]
            rb = out[b]
            out.append(ServeRequest(t, rb.key, rb.natom, rb.box, rb.nfft, rb.order, rb.kappa, True, base=b))
            continue
        n = math.exp(rng.uniform(math.log(nmin), math.log(nmax)))
        ntrip = max(1, int(round(n / 3)))
        natom = 3 * ntrip
        box = (natom / density) ** (1.0 / 3.0)
        nfft = systems.smooth235_at_least(box / spacing)
        order = 4 if rng.uniform() < 0.5 else 6
        out.append(ServeRequest(t, next_key, natom, box, nfft, order, kappa, False))
        new_idx.append(len(out) - 1)
        next_key += 1
    return out


def device_triplets(ntrip: int, box, seed: int, device):
    """(3*ntrip, 4) float32 device tensor of water-like triplets (same geometry
    and charges as systems.triplets, drawn with torch's generator)."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    f64 = torch.float64
    centres = torch.rand((ntrip, 3), generator=g, device=device, dtype=f64) * torch.tensor(box, dtype=f64, device=device)
    q = torch.randn((ntrip, 4), generator=g, device=device, dtype=f64)
    q = q / q.norm(dim=1, keepdim=True)
    w, x, y, z = q.unbind(1)
    R = torch.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
        2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
        2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1).view(ntrip, 3, 3)
    s, c = math.sin(systems.THETA_HOH / 2), math.cos(systems.THETA_HOH / 2)
    local = torch.tensor([[0.0, 0.0, 0.0], [systems.R_OH * s, systems.R_OH * c, 0.0],
                          [-systems.R_OH * s, systems.R_OH * c, 0.0]], dtype=f64, device=device)
    out = torch.empty((ntrip, 3, 4), dtype=torch.float32, device=device)
    for k in range(3):
        out[:, k, :3] = (centres + (R * local[k][None, None, :]).sum(-1)).to(torch.float32)
    out[:, 0, 3] = systems.Q_O
    out[:, 1:, 3] = systems.Q_H
    return out.view(-1, 4)


def materialise(reqs, device, sigma: float = 0.01, seed: int = SERVE_SEED):
    """Device inputs and force buffers for every request (inputs resident in
    HBM before the run)."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed + 7)
    for i, r in enumerate(reqs):
        if r.followup:
            base = reqs[r.base].xyzq
            d = torch.randn((r.natom, 3), generator=g, device=device, dtype=torch.float32) * sigma
            x = base.clone()
            x[:, :3] += d
            r.xyzq = x
        else:
            r.xyzq = device_triplets(r.natom // 3, (r.box, r.box, r.box), seed + 1000 + i, device)
        r.stride = force_stride(r.natom)
        r.force = torch.empty(3 * r.stride, dtype=torch.float32, device=device)
    if torch.device(device).type == "cuda":
        torch.cuda.synchronize()


def desc_of(r: ServeRequest, force_type=L.FORCE_FLOAT):
    from .pme import PmeBatch

    recip = [1.0 / r.box, 0, 0, 0, 1.0 / r.box, 0, 0, 0, 1.0 / r.box]
    return PmeBatch.desc((r.nfft,) * 3, r.order, r.kappa, recip, r.xyzq, r.natom, r.force, force_type, r.stride)


def run_open_loop(server: PmeServer, reqs, time_scale: float = 1.0):
    """Submit each request at its arrival time; returns (wall seconds, max
    submission lag in ms).  Results are collected with wait() afterwards."""
    t0 = time.perf_counter()
    lag = 0.0
    for r in reqs:
        target = t0 + r.t * time_scale
        while True:
            now = time.perf_counter()
            if now >= target:
                break
            if target - now > 2e-3:
                time.sleep(target - now - 1e-3)
        lag = max(lag, time.perf_counter() - target)
        r.ticket = server.submit(r.key, desc_of(r))
    server.drain()
    wall = time.perf_counter() - t0
    for r in reqs:
        r.result = server.wait(r.ticket)
    return wall, lag * 1e3


def run_burst(server: PmeServer, reqs):
    """Submit every request at once (saturation); returns wall seconds from the
    first submission to the last completion."""
    t0 = time.perf_counter()
    for r in reqs:
        r.ticket = server.submit(r.key, desc_of(r))
    server.drain()
    wall = time.perf_counter() - t0
    for r in reqs:
        r.result = server.wait(r.ticket)
    return wall


def run_plain(reqs):
    """The same requests one at a time through the plain five-call path
    (one solver per (grid, order), reused), each followed by its readback and
    a synchronisation; returns wall seconds."""
    import torch

    from .pme import CudaEnergyVirial, CudaPMERecip

    solvers = {}
    t0 = time.perf_counter()
    for r in reqs:
        k = (r.nfft, r.order)
        if k not in solvers:
            ev = CudaEnergyVirial()
            solvers[k] = (ev, CudaPMERecip(r.nfft, r.nfft, r.nfft, r.order, ev))
        ev, pme = solvers[k]
        recip = [1.0 / r.box, 0, 0, 0, 1.0 / r.box, 0, 0, 0, 1.0 / r.box]
        ev.clear()
        pme.spread_charge(r.xyzq, r.natom, recip)
        pme.r2c_fft()
        pme.scalar_sum(recip, r.kappa, True, True)
        pme.c2r_fft()
        pme.gather_force(r.xyzq, r.natom, recip, r.stride, r.force, L.FORCE_FLOAT)
        ev.copyToHost()
        torch.cuda.synchronize()
    return time.perf_counter() - t0


class Submitter(threading.Thread):
    """Open-loop submission on a thread of its own (for drivers that do
    other work meanwhile)."""

    def __init__(self, server, reqs, time_scale=1.0):
        super().__init__(daemon=True)
        self.server, self.reqs, self.time_scale = server, reqs, time_scale
        self.wall = self.lag_ms = None
        self.error = None

    def run(self):
        try:
            self.wall, self.lag_ms = run_open_loop(self.server, self.reqs, self.time_scale)
        except Exception as e:  # surfaced by the caller
            self.error = e
