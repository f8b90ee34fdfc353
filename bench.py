"""PME reciprocal-solve benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Headline workload: the batched set -- 64 independent orthorhombic systems
(atoms log-uniform in 2,000-40,000 as seeded TIP3P-like triplets, box edges
U(30, 75) A, grid edge = smallest 2/3/5-smooth >= edge/0.9, order 4, kappa
0.34, FP32).  A step is one batched solve of all 64 systems (one hipGraph
replay: binning, spread, 3-D R2C FFT, influence function with energy and
virial, 3-D C2R FFT, gather) plus the energy/virial copy to the host, so
value = 64 x steps / seconds.  Beside it, the same 64 systems solved one at
a time through the plain five-call path (the comparison the batched path is
built to beat) and the production plain solve (96,000 atoms, 108x100x96,
order 6).  Inputs are resident in HBM before every timed region.

Prints ONE JSON line.  The roofline object is for the batched kernel with
the largest average duration, timed with HIP events recorded on the batch's
stream.  The CPU baseline is the FP32 restatement of the reference's solve
(cpu_baseline/, C++ with OpenMP) timed on this host's cores, 1 thread and all
threads, on the small, production and 144^3 systems and the batched mix; the
FP64 oracle is timed beside it as the accuracy reference.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
# timing slots of pme_recip_get_timing / pme_batch_get_timing; when a grid's
# xy plane fits one workgroup the x and y passes run as one plane kernel
# (slots 2 and 5), and slots 3 and 6 are empty
STAGES = ["bin", "spread", "fft_fwd_x_or_xy", "fft_fwd_y", "fft_z_solve", "fft_inv_y_or_xy", "fft_inv_x", "gather"]
METRIC = "PME recip solves/s + atom-steps/s, batched/served vs plain; p99 latency @1e-4 err"
PLANE_LDS_BUDGET = 150 * 1024


def plane_mode(nfft):
    """Whether the library runs the x and y passes of this grid as one xy
    plane kernel: the same LDS test as fft_plane_lds_bytes (pme_fft.hip)."""
    nx, ny = nfft[0], nfft[1]
    nf = nx // 2 if nx % 2 == 0 else nx
    odd = lambda n: (n + 1) | 1  # noqa: E731
    buf = max(nf * odd(ny), ny * odd(nx // 2 + 1))
    return 8 * (nx + ny + 2 * buf) <= PLANE_LDS_BUDGET


def stage_bytes(s):
    """Algorithmic HBM bytes per launch of each timing slot for one system
    (FP32 data): every kernel reads its inputs and writes its outputs once."""
    N = s.natom
    nx, ny, nz = s.nfft
    G = nx * ny * nz
    Gc = (nx // 2 + 1) * ny * nz
    plane = plane_mode(s.nfft)
    return {
        "bin": 16 * N + 4 * N,                           # xyzq in, atom index into its bin out
        "spread": 16 * N + 4 * N + 4 * G,                 # xyzq + bin index in, charge grid out
        "fft_fwd_x_or_xy": 4 * G + 8 * Gc,                # real grid in, half spectrum out (x rows or xy planes)
        "fft_fwd_y": 0 if plane else 16 * Gc,             # spectrum in and out
        "fft_z_solve": 16 * Gc,                           # spectrum in, solved + z-inverted spectrum out
        "fft_inv_y_or_xy": (8 * Gc + 4 * G) if plane else 16 * Gc,
        "fft_inv_x": 0 if plane else 8 * Gc + 4 * G,
        "gather": 16 * N + 4 * N + 4 * G + 12 * N,        # xyzq + bin index + potential in, forces out
    }


# kernels behind each timing slot of the batched solve of the 64-system set
# (order 4, full bricks, every xy plane fitting a half-size plane workgroup,
# every z length a four-step pair of register class 0)
BATCH_KERNELS = {"bin": ("pme::k_batch_bin",), "spread": ("void pme::k_batch_spread<4, 1024>",),
                 "fft_fwd_x_or_xy": ("void pme::k_batch_plane<0, false, 512>",),
                 "fft_z_solve": ("void pme::k_batch_zsolve<0>",),
                 "fft_inv_y_or_xy": ("void pme::k_batch_plane<1, false, 512>",),
                 "gather": ("void pme::k_batch_gather<4, 1024>",)}


def stage_traffic(kernels, workload="batched"):
    """HBM bytes of one timing slot: the sum over its kernels' launches, or
    None when one of them has no measurement."""
    vals = [measured_traffic(k, workload) for k in kernels]
    return None if not vals or any(v is None for v in vals) else sum(vals)


def measured_traffic(kernel, workload="batched"):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.md")
    if not os.path.exists(path):
        return None
    col = None  # the "HBM bytes (calibrated)" column, found from the table header
    for line in open(path):
        if not line.lstrip().startswith("|"):
            continue
        cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if col is None:
            if "HBM bytes (calibrated)" in cells:
                col = cells.index("HBM bytes (calibrated)")
            continue
        if len(cells) > col and cells[0] == kernel and cells[1] == workload:
            try:
                return int(float(cells[col].replace(",", "")))
            except ValueError:
                return None
    return None


def forces_of(f, n, stride):
    fh = f.cpu().numpy()
    return np.stack([fh[:n], fh[stride:stride + n], fh[2 * stride:2 * stride + n]], 1).astype(np.float64)


def _time_solves(fn, items, budget_s):
    """Run fn(item) over items in turn for about budget_s (at least one
    solve); returns (solves/s, atoms/s, solves)."""
    n = atoms = 0
    t0 = time.perf_counter()
    while n == 0 or time.perf_counter() - t0 < budget_s:
        s = items[n % len(items)]
        fn(s)
        n += 1
        atoms += s.natom
    dt = time.perf_counter() - t0
    return n / dt, atoms / dt, n


def cpu_baseline(sl, s1, s2, s4, budget_s=2.0):
    """The FP32 restatement of the reference's solve (cpu_baseline/, C++ with
    OpenMP) on this host's cores: 1 thread and every thread, for the small
    (C1), production (C2) and split-precision-size (C4) systems and the
    batched mix.  `value` is the batched mix on every thread (the headline's
    workload).  The FP64 oracle is timed beside it on the mix, as the
    accuracy reference it is."""
    import cpu_baseline as CB
    import oracle as O

    nproc = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or CB.max_threads() or nproc
    solvers = {}

    def run(s, nt):
        key = (s.nfft, s.order)
        if key not in solvers:
            solvers[key] = CB.CpuPme(s.nfft, s.order)
        solvers[key].solve(s.xyzq, s.box, s.kappa, nthreads=nt)

    per = {}
    for name, items, budget in (("small_8k_48c_p4", [s1], budget_s), ("production_96k_p6", [s2], budget_s),
                                ("split_size_250k_144c_p4", [s4], budget_s), ("batched_mix_64", sl, 2 * budget_s)):
        row = {}
        for nt in (1, threads):
            run(items[0], nt)  # warm (allocations, first touch)
            v, a, n = _time_solves(lambda s: run(s, nt), items, budget)
            row[f"threads_{nt}"] = {"solves_per_s": v, "atom_steps_per_s": a, "solves": n}
        per[name] = row
    solvers.clear()
    mix = per["batched_mix_64"][f"threads_{threads}"]
    ov, oa, on = _time_solves(
        lambda s: O.solve(O.OracleParams(*s.nfft, s.order, s.kappa, s.box, index_mode=O.INDEX_F32,
                                         nthreads=threads), s.xyzq), sl, 2 * budget_s)
    return {
        "value": mix["solves_per_s"], "unit": "solves/s", "cores": threads, "kind": "port",
        "port": "fp32-restatement: cpu_baseline/pme_cpu.cpp, the reference's index formula (fmaf), B-spline "
                "recursion and stage order with FP32 grids and FFTs and FP64 energy/virial sums, OpenMP",
        "sample": f"{mix['solves']} solves of the 64 batched systems taken in turn on {threads} threads "
                  f"(about {2 * budget_s:g} s); per_config: about {budget_s:g} s per config and thread count",
        "atom_steps_per_s": mix["atom_steps_per_s"],
        "cpu_model": CB.cpu_model(), "nproc": nproc, "threads": threads,
        "per_config": per,
        "oracle_fp64": {"solves_per_s": ov, "atom_steps_per_s": oa, "solves": on, "threads": threads,
                        "note": "the FP64 C oracle (oracle/) on the batched mix: the accuracy reference, "
                                "not the speed baseline"},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-split", action="store_true", help="skip the split-precision (144^3) leg")
    ap.add_argument("--no-direct", action="store_true", help="skip the direct-space Ewald leg")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="seconds of open-loop arrivals (300 requests/s) for the served leg; 0 skips it")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device copies of 256 MiB (scale for the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this PME engine runs on one GPU; whole systems shard across GPUs "
                                   "without any exchange, so --gpus N>1 is not a supported configuration"}))
        sys.exit(2)

    import torch

    from gpu_recip_amd import _lib as L
    from gpu_recip_amd import systems
    from gpu_recip_amd.pme import CudaEnergyVirial, CudaPMERecip, PmeBatch, force_stride

    dev = torch.device("cuda:0")
    if args.copy_calibration:
        a = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"bytes_read": a.numel() * 4, "bytes_written": b.numel() * 4,
                                               "launches": 5}}))
        return

    # ---------------- batched: 64 systems, one graph replay per step ----------
    sl = systems.config_batched()
    # the batch runs on a stream of the caller's (as an MD engine would pass
    # its own), so no ordering against the null stream is needed per solve;
    # the inputs below are complete before the first solve (synchronize)
    bstream = torch.cuda.Stream()
    batch = PmeBatch(32 << 30, stream=bstream)
    reg = []
    for s in sl:
        x = torch.from_numpy(s.xyzq).to(dev)
        stride = force_stride(s.natom)
        f = torch.zeros(3 * stride, dtype=torch.float32, device=dev)
        sid = batch.add(PmeBatch.desc(s.nfft, s.order, s.kappa, s.recip, x, s.natom, f, L.FORCE_FLOAT, stride),
                        keep=(x, f))
        reg.append((sid, x, f, stride))
    ids = [r[0] for r in reg]
    torch.cuda.synchronize()
    tune_ms = None if args.no_autotune else batch.autotune(None)
    tuning = batch.get_tuning()
    chains = batch.get_chains()
    half_bricks = batch.get_half_bricks()

    def bstep():
        batch.solve()
        return batch.results(ids)

    for _ in range(args.warmup):
        bstep()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e_b, v_b = bstep()
    torch.cuda.synchronize()
    dt_b = time.perf_counter() - t0
    nsolve = len(sl)
    atoms_b = sum(s.natom for s in sl)

    # per-stage kernel times of the batched solve (HIP events on its stream):
    # every stage as one launch over all 64 systems (one chain; one workgroup
    # per work item, spread and gather one per CU from a work queue), the tuned
    # brick and FFT tile
    batch.set_tuning(tuning[0], tuning[1], 0)
    batch.set_timing(True)
    stage_ms = np.zeros(len(STAGES))
    nt = max(10, args.steps // 4)
    for _ in range(nt):
        bstep()
        stage_ms += np.array(batch.get_timing())
    batch.set_timing(False)
    batch.set_tuning(*tuning)
    stage_ms /= nt
    sb = {k: sum(stage_bytes(s)[k] for s in sl) for k in STAGES}
    k = int(np.argmax(stage_ms))
    kname = STAGES[k]
    achieved = sb[kname] / (stage_ms[k] * 1e-3) / 1e9
    bytes_b = sum(s.algorithmic_bytes() for s in sl)
    ms_b = dt_b / args.steps * 1e3

    # ---------------- plain: the same 64 systems one at a time ----------------
    plain = []
    for s, (sid, x, f, stride) in zip(sl, reg):
        ev = CudaEnergyVirial()
        pme = CudaPMERecip(*s.nfft, s.order, ev)
        fp = torch.zeros_like(f)
        plain.append((s, ev, pme, x, fp, stride))

    def pstep():
        for s, ev, pme, x, fp, stride in plain:
            r = s.recip
            ev.clear()
            pme.spread_charge(x, s.natom, r)
            pme.r2c_fft()
            pme.scalar_sum(r, s.kappa, True, True)
            pme.c2r_fft()
            pme.gather_force(x, s.natom, r, stride, fp, L.FORCE_FLOAT)
            ev.copyToHost()
            torch.cuda.synchronize()

    psteps = max(5, args.steps // 10)
    for _ in range(max(2, args.warmup // 5)):
        pstep()
    t0 = time.perf_counter()
    for _ in range(psteps):
        pstep()
    dt_p = time.perf_counter() - t0
    plain_c3 = nsolve * psteps / dt_p

    # ---------------- plain production solve (96,000 atoms) -------------------
    s2 = systems.config_production()
    ev2 = CudaEnergyVirial()
    pme2 = CudaPMERecip(*s2.nfft, s2.order, ev2)
    x2 = torch.from_numpy(s2.xyzq).to(dev)
    st2 = force_stride(s2.natom)
    f2 = torch.zeros(3 * st2, dtype=torch.float32, device=dev)

    def p2step():
        r = s2.recip
        ev2.clear()
        pme2.spread_charge(x2, s2.natom, r)
        pme2.r2c_fft()
        pme2.scalar_sum(r, s2.kappa, True, True)
        pme2.c2r_fft()
        pme2.gather_force(x2, s2.natom, r, st2, f2, L.FORCE_FLOAT)
        ev2.copyToHost()
        torch.cuda.synchronize()

    for _ in range(args.warmup):
        p2step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        p2step()
    dt_2 = time.perf_counter() - t0
    pme2.set_timing(True)
    st_ms2 = np.zeros(len(STAGES))
    for _ in range(nt):
        p2step()
        st_ms2 += np.array(pme2.get_timing())
    pme2.set_timing(False)
    st_ms2 /= nt
    sb2 = stage_bytes(s2)
    stage_frac2 = {STAGES[i]: sb2[STAGES[i]] / (st_ms2[i] * 1e-3) / (HBM_PEAK_GBS * 1e9)
                   for i in range(len(STAGES)) if sb2[STAGES[i]] > 0 and st_ms2[i] > 0}
    # the same solve through pme_recip_solve on a stream of its own: replayed
    # as one hipGraph after the first two calls
    side2 = torch.cuda.Stream()
    ev2g = CudaEnergyVirial()
    pme2g = CudaPMERecip(*s2.nfft, s2.order, ev2g)
    pme2g.set_stream(side2)
    pme2g.set_graph(True)
    f2g = torch.zeros(3 * st2, dtype=torch.float32, device=dev)

    def p2gstep():
        ev2g.clear(side2)
        pme2g.solve(x2, s2.natom, s2.recip, s2.kappa, True, True, st2, f2g, L.FORCE_FLOAT)
        ev2g.copyToHost(side2)
        side2.synchronize()

    for _ in range(max(3, args.warmup)):
        p2gstep()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        p2gstep()
    dt_2g = time.perf_counter() - t0
    assert np.array_equal(f2g.cpu().numpy(), f2.cpu().numpy()), "graph solve differs from the five calls"

    # ---------------- split precision: 250,002 atoms, 144^3, order 4 ---------
    split = None
    if not args.no_split:
        s4 = systems.config_split_precision()
        ev4 = CudaEnergyVirial()
        pme4 = CudaPMERecip(*s4.nfft, s4.order, ev4)
        x4 = torch.from_numpy(s4.xyzq).to(dev)
        st4 = force_stride(s4.natom)
        split = {"workload": s4.name, "natom": s4.natom, "nfft": list(s4.nfft), "order": s4.order, "kappa": s4.kappa,
                 "fp64_steps": "bsp_mod tables, energy/virial partial sums and their reduction",
                 "matrix_core_steps": "radix-16 DFT stage of the y and z passes (144 = 16 x 9), forward and inverse",
                 "modes": {}}
        f4 = {}
        for prec, pname in ((L.PREC_FP32, "fp32"), (L.PREC_BF16X3, "bf16x3"), (L.PREC_FP16X3, "fp16x3")):
            pme4.set_precision(prec)
            f4[pname] = torch.zeros(3 * st4, dtype=torch.float32, device=dev)

            def p4step():
                r = s4.recip
                ev4.clear()
                pme4.spread_charge(x4, s4.natom, r)
                pme4.r2c_fft()
                pme4.scalar_sum(r, s4.kappa, True, True)
                pme4.c2r_fft()
                pme4.gather_force(x4, s4.natom, r, st4, f4[pname], L.FORCE_FLOAT)
                ev4.copyToHost()
                torch.cuda.synchronize()

            for _ in range(max(3, args.warmup // 2)):
                p4step()
            k4 = max(10, args.steps // 4)
            t0 = time.perf_counter()
            for _ in range(k4):
                p4step()
            dt4 = time.perf_counter() - t0
            pme4.set_timing(True)
            ms4 = np.zeros(len(STAGES))
            for _ in range(10):
                p4step()
                ms4 += np.array(pme4.get_timing())
            pme4.set_timing(False)
            ms4 /= 10
            split["modes"][pname] = {"solves_per_s": k4 / dt4, "ms_per_solve": dt4 / k4 * 1e3,
                                     "energy": ev4.getEnergy("recip"),
                                     "stage_ms": {STAGES[i]: float(ms4[i]) for i in range(len(STAGES))}}
            if prec != L.PREC_FP32:
                # matrix-core work issued: every radix-16 column of 16 complex
                # points is 12 v_mfma_f32_32x32x8 per 32 columns (3 hi-lo
                # products x 4 K steps), 2*32*32*8 flops each: 384 flops per
                # complex point per pass; the y passes run one pass, the fused
                # z pass two (forward and inverse)
                gc4 = (s4.nfft[0] // 2 + 1) * s4.nfft[1] * s4.nfft[2]
                passes = {"fft_fwd_y": 1, "fft_z_solve": 2, "fft_inv_y_or_xy": 1}
                mf = {}
                for st_name, npass in passes.items():
                    fl = 384.0 * gc4 * npass
                    t_s = float(ms4[STAGES.index(st_name)]) * 1e-3
                    mf[st_name] = {"mfma_flops": fl, "tflops": fl / t_s / 1e12,
                                   "utilisation_vs_2.5_pflops": fl / t_s / 2.5e15}
                split["modes"][pname]["mfma"] = mf
        split["_forces"] = f4
        split["_stride"] = st4
        split["_system"] = s4

    # ---------------- direct-space Ewald on the production system -------------
    direct = None
    if not args.no_direct:
        from gpu_recip_amd import nlist
        from gpu_recip_amd.direct import CudaPMEDirectForce, DeviceTileList

        roff = 9.0
        tl, ex14 = nlist.water_box(s2.xyzq, s2.box, roff)
        checks, inter = nlist.count_pairs(tl, s2.box, roff)
        evd = CudaEnergyVirial()
        dr = CudaPMEDirectForce(evd, "vdw", "elec", "excl")
        dr.setup(*s2.box, s2.kappa, roff, roff - 1.0, 1.0)
        dr.set_14_list(np.zeros((0, 3), np.int32), ex14)
        dl = DeviceTileList(tl, dev)
        std = force_stride(tl.npad, 8)
        fd = torch.zeros(3 * std, dtype=torch.int64, device=dev)

        def dstep():
            evd.clear()
            dr.calc_force(dl.xyzq, dl, True, True, std, fd)
            dr.calc_14_force(dl.xyzq, True, True, std, fd)
            evd.copyToHost()
            torch.cuda.synchronize()

        for _ in range(max(3, args.warmup)):
            dstep()
        kd = max(10, args.steps)
        t0 = time.perf_counter()
        for _ in range(kd):
            dstep()
        dt_d = (time.perf_counter() - t0) / kd
        # the pair kernel alone, HIP events on the stream it runs on (the null stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(kd):
            dr.calc_force(dl.xyzq, dl, True, True, std, fd)
        e1.record()
        torch.cuda.synchronize()
        ms_pair = e0.elapsed_time(e1) / kd
        # FP32 work: 6 operations per pair check (3 differences, 3 multiply-adds
        # for r^2), 28 per interacting pair (the Ewald term with erfc, exp and
        # rsqrt counted as one operation each, force components and sums)
        flops = 6.0 * checks + 28.0 * inter
        direct = {
            "workload": f"{s2.name}: direct-space Ewald, roff {roff} A, kappa {s2.kappa}, water exclusions "
                        f"as 1-4 exclusion corrections; reference-format 32-atom tile list built on the host",
            "natom": s2.natom, "tiles": tl.npad // 32, "i_entries": int(len(tl.ientry)),
            "j_tiles": int(len(tl.tile_indj)), "pair_checks": checks, "interacting_pairs": inter,
            "ms_per_call": dt_d * 1e3, "calls_per_s": 1.0 / dt_d,
            "step": "calc_force + calc_14_force with energy and virial, energy/virial readback, sync",
            "pair_kernel_ms": ms_pair, "interactions_per_s": inter / (ms_pair * 1e-3),
            "roofline": {"bound": "valu", "kernel": "k_direct_force<true, true>",
                         "achieved": flops / (ms_pair * 1e-3) / 1e12, "peak": 157.3, "unit": "TFLOP/s",
                         "frac": flops / (ms_pair * 1e-3) / 1e12 / 157.3,
                         "flops_per_launch": flops,
                         "counting": "6 per pair check + 28 per interacting pair"},
            "energy_elec": evd.getEnergy("elec"), "energy_excl": evd.getEnergy("excl"),
        }

    # ---------------- served: open-loop Poisson arrivals, continuous batching --
    served = None
    if args.serve_seconds > 0:
        from gpu_recip_amd import serving
        from gpu_recip_amd.pme import PmeServer

        reqs = serving.schedule(args.serve_seconds)
        serving.materialise(reqs, dev)
        srv = PmeServer(64 << 30, nlanes=2)
        warm = serving.schedule(0.2, seed=serving.SERVE_SEED + 99)
        serving.materialise(warm, dev)
        serving.run_burst(srv, warm)  # first-use table uploads, lane buffers
        for r in warm:
            srv.release(r.key)
        srv.reset_stats()
        wall_o, lag_ms = serving.run_open_loop(srv, reqs)
        so = srv.stats()
        atoms_o = sum(r.natom for r in reqs)
        # saturation: the first requests (up to 3,000) all at once, and the
        # same requests one at a time through the plain path
        sat = reqs[:3000]
        for r in sat:
            srv.release(r.key)
        srv.reset_stats()
        wall_s = serving.run_burst(srv, sat)
        ss = srv.stats()
        srv.close()
        wall_p = serving.run_plain(sat)
        atoms_s = sum(r.natom for r in sat)
        served = {
            "workload": f"open-loop Poisson 300 requests/s for {args.serve_seconds:g} s; atoms log-uniform 5k-400k, "
                        "cubic box at 0.1 atoms/A^3, grid spacing ~1 A (2/3/5-smooth), order 4 or 6, kappa 0.34, "
                        "20% follow-ups on resident systems displaced by N(0, 0.01 A)",
            "requests": len(reqs), "followups": sum(r.followup for r in reqs),
            "solves_per_s": so["completed"] / wall_o, "atom_steps_per_s": atoms_o / wall_o,
            "p50_ms": so["p50_ms"], "p99_ms": so["p99_ms"], "mean_ms": so["mean_ms"], "max_ms": so["max_ms"],
            "mean_batch": so["mean_batch"], "max_batch": so["max_batch"], "evictions": so["evictions"],
            "new_slots": so["admitted_new"], "failed": so["failed"], "max_submit_lag_ms": lag_ms,
            "saturation": {"requests": len(sat), "solves_per_s": len(sat) / wall_s, "atom_steps_per_s": atoms_s / wall_s,
                           "p50_ms": ss["p50_ms"], "p99_ms": ss["p99_ms"], "mean_batch": ss["mean_batch"]},
            "plain_same_requests": {"solves_per_s": len(sat) / wall_p, "atom_steps_per_s": atoms_s / wall_p},
            "speedup_served_vs_plain": wall_p / wall_s,
        }

    # ---------------- accuracy vs the FP64 oracle (the checker) ---------------
    acc = {}
    try:
        import oracle as O

        worst_f = worst_e = 0.0
        for i in range(0, nsolve, 8):
            s = sl[i]
            o = O.solve(O.OracleParams(*s.nfft, s.order, s.kappa, s.box, index_mode=O.INDEX_F64), s.xyzq)
            fb = forces_of(reg[i][2], s.natom, reg[i][3])
            worst_f = max(worst_f, float(np.sqrt(((fb - o["force"]) ** 2).sum() / (o["force"] ** 2).sum())))
            worst_e = max(worst_e, float(abs(e_b[i] - o["energy"]) / abs(o["energy"])))
        o2 = O.solve(O.OracleParams(*s2.nfft, s2.order, s2.kappa, s2.box, index_mode=O.INDEX_F64), s2.xyzq)
        fp2 = forces_of(f2, s2.natom, st2)
        acc = {
            "batched_worst_force_rel_rms_vs_fp64": worst_f,
            "batched_worst_energy_rel_vs_fp64": worst_e,
            "batched_systems_checked": list(range(0, nsolve, 8)),
            "production_force_rel_rms_vs_fp64": float(np.sqrt(((fp2 - o2["force"]) ** 2).sum() / (o2["force"] ** 2).sum())),
            "production_energy_rel_vs_fp64": float(abs(ev2.getEnergy("recip") - o2["energy"]) / abs(o2["energy"])),
        }
        if split is not None:
            s4 = split.pop("_system")
            st4 = split.pop("_stride")
            f4 = split.pop("_forces")
            o4 = O.solve(O.OracleParams(*s4.nfft, s4.order, s4.kappa, s4.box, index_mode=O.INDEX_F64), s4.xyzq)
            for pname, ft in f4.items():
                fr = forces_of(ft, s4.natom, st4)
                m = split["modes"][pname]
                m["force_rel_rms_vs_fp64"] = float(np.sqrt(((fr - o4["force"]) ** 2).sum() / (o4["force"] ** 2).sum()))
                m["energy_rel_vs_fp64"] = float(abs(m.pop("energy") - o4["energy"]) / abs(o4["energy"]))
        if served is not None:
            worst = 0.0
            small = [r for r in reqs if r.natom <= 30000 and not r.followup][:3]
            for r in small:
                o = O.solve(O.OracleParams(r.nfft, r.nfft, r.nfft, r.order, r.kappa, (r.box,) * 3), r.xyzq.cpu().numpy())
                fh = r.force.cpu().numpy()
                fs = np.stack([fh[:r.natom], fh[r.stride:r.stride + r.natom],
                               fh[2 * r.stride:2 * r.stride + r.natom]], 1).astype(np.float64)
                worst = max(worst, float(np.sqrt(((fs - o["force"]) ** 2).sum() / (o["force"] ** 2).sum())))
            acc["served_worst_force_rel_rms_vs_fp64"] = worst
            acc["served_requests_checked"] = [r.key for r in small]
    except Exception as e:  # the oracle is the checker; its absence must not hide the GPU number
        acc = {"error": str(e)}

    if split is not None:
        for k_ in ("_forces", "_stride", "_system"):
            split.pop(k_, None)
    out = {
        "metric": METRIC,
        "value": nsolve * args.steps / dt_b,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_b,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded TIP3P-like triplets)",
        "config": {"workload": "batched-64-systems-2k-40k-atoms-p4", "systems": nsolve, "atoms_total": atoms_b,
                   "grid_edges": [min(min(s.nfft) for s in sl), max(max(s.nfft) for s in sl)], "order": 4,
                   "kappa": 0.34, "step": "one batched solve of all 64 systems (hipGraph replay) + energy/virial "
                                          "readback", "tuning": {"brick_target": tuning[0], "fft_tile": tuning[1],
                                                                 "wg_per_cu": tuning[2], "chains": chains,
                                                                 "half_bricks": half_bricks,
                                                                 "autotune_ms": tune_ms}},
        "atom_steps_per_s": atoms_b * args.steps / dt_b,
        "plain_same_systems": {"solves_per_s": plain_c3, "atom_steps_per_s": atoms_b * psteps / dt_p,
                               "path": "the 64 systems one at a time: 5 stage calls + readback + sync each"},
        "speedup_batched_vs_plain": (nsolve * args.steps / dt_b) / plain_c3,
        "production_plain": {"workload": s2.name, "solves_per_s": args.steps / dt_2,
                             "atom_steps_per_s": s2.natom * args.steps / dt_2,
                             "ms_per_solve": dt_2 / args.steps * 1e3,
                             "stage_ms": {STAGES[i]: float(st_ms2[i]) for i in range(len(STAGES))},
                             "stage_roofline_frac": stage_frac2,
                             "stage_timing_note": "HIP events per stage on the handle's stream (each slot carries "
                                                  "a few us of event overhead); rocprofv3 kernel times in profiles/",
                             "solve_roofline_frac": s2.algorithmic_bytes() / (dt_2 / args.steps) / (HBM_PEAK_GBS * 1e9),
                             "graph": {"step": "pme_recip_solve on a stream of its own (hipGraph replay), "
                                               "energy/virial clear and readback, sync",
                                       "solves_per_s": args.steps / dt_2g, "ms_per_solve": dt_2g / args.steps * 1e3,
                                       "solve_roofline_frac": s2.algorithmic_bytes() / (dt_2g / args.steps)
                                       / (HBM_PEAK_GBS * 1e9)}},
        "stage_ms": {STAGES[i]: float(stage_ms[i]) for i in range(len(STAGES))},
        # each stage's algorithmic bytes over its mean time, against 8 TB/s (the
        # empty slots of the plane layout carry only event overhead)
        "stage_roofline_frac": {STAGES[i]: sb[STAGES[i]] / (stage_ms[i] * 1e-3) / (HBM_PEAK_GBS * 1e9)
                                for i in range(len(STAGES)) if sb[STAGES[i]] > 0},
        "solve_algorithmic_bytes": bytes_b,
        "solve_roofline_frac": bytes_b / (ms_b * 1e-3) / (HBM_PEAK_GBS * 1e9),
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": stage_traffic(BATCH_KERNELS.get(kname, ())),
                     "traffic_source": "profiles/hbm_traffic.md (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated)",
                     "algorithmic_bytes_per_launch": sb[kname], "avg_launch_ms": float(stage_ms[k])},
        "served": served,
        "split_precision": split,
        "direct_space": direct,
        "accuracy": acc,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sl, systems.config_plain_small(), s2, systems.config_split_precision())
    print(json.dumps(out))


if __name__ == "__main__":
    main()
