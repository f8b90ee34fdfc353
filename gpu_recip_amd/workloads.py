"""Run one benchmark workload for a fixed number of solves (for profilers).

    python -m gpu_recip_amd.workloads production [--solves 20] [--precision fp32|bf16x3|fp16x3]
                                                 [--side-stream] [--one-shot] [--graph]
    python -m gpu_recip_amd.workloads split      [--solves 20] [--precision ...]
    python -m gpu_recip_amd.workloads batched    [--solves 20] [--tuning brick,tile,wg_per_cu[,chains[,half]]]
    python -m gpu_recip_amd.workloads small      [--solves 20]
    python -m gpu_recip_amd.workloads served     [--solves 1000]
    python -m gpu_recip_amd.workloads direct     [--solves 20]

production: 96,000 atoms, 108x100x96, order 6 (plain five-call path)
split:      250,002 atoms, 144^3, order 4 (plain path, chosen FFT precision)
batched:    the 64-system set, one batched solve per step (autotuned plan, as
            in the benchmark)
small:      8,001 atoms, 48^3, order 4 (plain path)
served:     the first --solves requests of the serving schedule (5k-400k
            atoms, grids 40-160, order 4 or 6), submitted at once to a
            2-lane server (saturation), after a short warm-up burst
direct:     the direct-space Ewald term on the production system (roff 9 A,
            water exclusions as 1-4 exclusion corrections), calc_force +
            calc_14_force with energy and virial per solve
Inputs are uploaded once; every solve ends with the energy/virial readback.
The plain workloads run the five stage calls on the null stream; --side-stream
runs them on a stream of the handle's own, --one-shot through pme_recip_solve,
--graph through pme_recip_solve replayed as a hipGraph (implies both).
"""
from __future__ import annotations

import argparse
import time

PRECS = {"fp32": 0, "bf16x3": 1, "fp16x3": 2}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=["production", "split", "batched", "small", "served", "direct"])
    ap.add_argument("--solves", type=int, default=20)
    ap.add_argument("--precision", choices=sorted(PRECS), default="fp32")
    ap.add_argument("--side-stream", action="store_true", help="plain: run on a stream of the handle's own")
    ap.add_argument("--one-shot", action="store_true", help="plain: pme_recip_solve instead of the five calls")
    ap.add_argument("--graph", action="store_true", help="plain: pme_recip_solve replayed as a hipGraph")
    ap.add_argument("--tuning", default=None,
                    help="batched: brick,tile,wg_per_cu[,chains[,half bricks]] to use instead of autotuning "
                         "(e.g. 24,32,4,2,1); served: brick,tile,wg_per_cu of the server")
    args = ap.parse_args(argv)

    import torch

    from . import _lib as L
    from . import systems
    from .pme import CudaEnergyVirial, CudaPMERecip, PmeBatch, force_stride

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    if args.workload == "served":
        from . import serving
        from .pme import PmeServer

        reqs = serving.schedule(1e9, max_requests=args.solves)
        serving.materialise(reqs, dev)
        srv = PmeServer(64 << 30, nlanes=2)
        if args.tuning:
            srv.set_tuning(*[int(t) for t in args.tuning.split(",")][:3])
        warm = serving.schedule(0.2, seed=serving.SERVE_SEED + 99)
        serving.materialise(warm, dev)
        serving.run_burst(srv, warm)
        for r in warm:
            srv.release(r.key)
        wall = serving.run_burst(srv, reqs)
        st = srv.stats()
        print(f"served: {len(reqs)} requests in {wall:.3f} s ({len(reqs) / wall:.0f} solves/s), "
              f"mean batch {st['mean_batch']:.1f}", flush=True)
        srv.close()
        return
    if args.workload == "batched":
        b = PmeBatch(32 << 30)
        ids = []
        for s in systems.config_batched():
            x = torch.from_numpy(s.xyzq).to(dev)
            st = force_stride(s.natom)
            f = torch.zeros(3 * st, dtype=torch.float32, device=dev)
            ids.append(b.add(PmeBatch.desc(s.nfft, s.order, s.kappa, s.recip, x, s.natom, f, L.FORCE_FLOAT, st,
                                           precision=PRECS[args.precision]), keep=(x, f)))
        if args.tuning:
            v = [int(t) for t in args.tuning.split(",")]
            if len(v) > 4:
                b.set_half_bricks(bool(v[4]))
            b.set_tuning(*v[:3])
            if len(v) > 3:
                b.set_chains(v[3])
            print(f"batched: plan (brick, tile, wg/cu) = {b.get_tuning()}, {b.get_chains()} chain(s), "
                  f"half bricks {b.get_half_bricks()}", flush=True)
        else:
            ms = b.autotune(None)  # the plan the benchmark uses
            print(f"batched: autotuned (brick, tile, wg/cu) = {b.get_tuning()}, {b.get_chains()} chain(s), "
                  f"half bricks {b.get_half_bricks()}, {ms:.3f} ms per batched solve", flush=True)
        t0 = time.perf_counter()
        for _ in range(args.solves):
            b.solve()
            b.results(ids)
        print(f"batched: {args.solves} x 64 solves in {time.perf_counter() - t0:.3f} s", flush=True)
        return
    if args.workload == "direct":
        import numpy as np

        from . import nlist
        from .direct import CudaPMEDirectForce, DeviceTileList

        s = systems.config_production()
        tl, ex14 = nlist.water_box(s.xyzq, s.box, 9.0)
        ev = CudaEnergyVirial()
        dr = CudaPMEDirectForce(ev, "vdw", "elec", "excl")
        dr.setup(*s.box, s.kappa, 9.0, 8.0, 1.0)
        dr.set_14_list(np.zeros((0, 3), np.int32), ex14)
        dl = DeviceTileList(tl, dev)
        st = force_stride(tl.npad, 8)
        f = torch.zeros(3 * st, dtype=torch.int64, device=dev)
        t0 = time.perf_counter()
        for _ in range(args.solves):
            ev.clear()
            dr.calc_force(dl.xyzq, dl, True, True, st, f)
            dr.calc_14_force(dl.xyzq, True, True, st, f)
            ev.copyToHost()
            torch.cuda.synchronize()
        print(f"direct ({len(tl.ientry)} i entries, {len(tl.tile_indj)} j tiles): {args.solves} calls in "
              f"{time.perf_counter() - t0:.3f} s, elec {ev.getEnergy('elec'):.6f}", flush=True)
        return
    s = {"production": systems.config_production, "split": systems.config_split_precision,
         "small": systems.config_plain_small}[args.workload]()
    ev = CudaEnergyVirial()
    pme = CudaPMERecip(*s.nfft, s.order, ev)
    pme.set_precision(PRECS[args.precision])
    one_shot = args.one_shot or args.graph
    stream = torch.cuda.Stream() if (args.side_stream or args.graph) else None
    if stream is not None:
        pme.set_stream(stream)
    pme.set_graph(args.graph)
    x = torch.from_numpy(s.xyzq).to(dev)
    st = force_stride(s.natom)
    f = torch.zeros(3 * st, dtype=torch.float32, device=dev)

    def step():
        ev.clear(stream)
        if one_shot:
            pme.solve(x, s.natom, s.recip, s.kappa, True, True, st, f, L.FORCE_FLOAT)
        else:
            pme.spread_charge(x, s.natom, s.recip)
            pme.r2c_fft()
            pme.scalar_sum(s.recip, s.kappa, True, True)
            pme.c2r_fft()
            pme.gather_force(x, s.natom, s.recip, st, f, L.FORCE_FLOAT)
        ev.copyToHost(stream)
        if stream is not None:
            stream.synchronize()
        else:
            torch.cuda.synchronize()

    for _ in range(3):
        step()
    t0 = time.perf_counter()
    for _ in range(args.solves):
        step()
    dt = time.perf_counter() - t0
    how = ("graph" if args.graph else "one-shot" if one_shot else "five calls") + (", own stream" if stream else "")
    print(f"{s.name} ({args.precision}, {how}): {args.solves} solves in {dt:.3f} s, {dt / args.solves * 1e6:.1f} us "
          f"per solve, energy {ev.getEnergy('recip'):.6f}", flush=True)


if __name__ == "__main__":
    main()
