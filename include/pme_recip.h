/*
 * pme_recip.h -- C ABI of the MI355X (gfx950) PME reciprocal-space engine.
 *
 * Drop-in boundary for the CudaPMERecip<AT,CT,CT2> / CudaEnergyVirial call
 * surface of ap-hynninen/gpu_recip.  Each entry point names the reference
 * member it replaces (file:line in that repository).  Plain pointers and sizes
 * only: device pointers are HIP device addresses, streams are hipStream_t
 * passed as void*.
 *
 * Deliberate deviations from the reference (which prints and calls exit(1) on
 * every error): every function returns a pme_status and pme_last_error()
 * gives the message of the calling thread's last failure.
 *
 * Threading: one handle = one stream, not re-entrant (as the reference).
 */
#ifndef PME_RECIP_H
#define PME_RECIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pme_status {
  PME_OK = 0,
  PME_ERR_INVALID_ARG = 1,   /* null pointer, negative count, bad enum        */
  PME_ERR_UNSUPPORTED = 2,   /* order/grid/box the engine does not handle     */
  PME_ERR_OOM = 3,           /* device or host allocation failed              */
  PME_ERR_BUDGET = 4,        /* request exceeds the configured HBM budget     */
  PME_ERR_DEVICE = 5,        /* HIP runtime error                             */
  PME_ERR_STATE = 6,         /* stage called out of order (e.g. gather before c2r) */
  PME_ERR_NOT_FOUND = 7,     /* unknown energy term name                      */
  PME_ERR_STOPPED = 8        /* a STOP request header (ncoord < 0)            */
} pme_status;

typedef void* pme_stream_t;                       /* hipStream_t */
typedef struct pme_energy_virial_s* pme_ev_t;     /* CudaEnergyVirial */
typedef struct pme_recip_s* pme_recip_t;          /* CudaPMERecip<AT,CT,CT2> */

/* FFTtype, CudaPMERecip.h:30.  Single GPU: all three give the same results;
 * intermediate grid views are always returned in the BOX layout. */
enum { PME_FFT_COLUMN = 0, PME_FFT_SLAB = 1, PME_FFT_BOX = 2 };

/* AT template argument (accumulation type of the charge grid),
 * CudaPMERecip.cu:2584-2600.  The engine accumulates every charge grid as
 * 64-bit fixed point with scale 2^40 (the reference's `long long` mode,
 * CudaPMERecip.cu:51-58), which is order independent and at least as precise
 * as the reference's float and 2^31 int modes; all three values are accepted
 * and give the same grid. */
enum { PME_ACCUM_FLOAT = 0, PME_ACCUM_INT32 = 1, PME_ACCUM_INT64 = 2 };

/* FT template argument of gather_force, CudaPMERecip.cu:985-1023 */
enum {
  PME_FORCE_FLOAT = 0,   /* float*: strided SoA store  f[i], f[i+stride], f[i+2 stride] */
  PME_FORCE_FLOAT3 = 1,  /* float3*: AoS store (stride ignored)                          */
  PME_FORCE_INT64 = 2    /* long long*: fixed point 2^40, ADDED atomically, strided SoA */
};

/* grid views, CudaPMERecip.h:41-64 (charge_grid / fft_grid / solved_grid) */
enum { PME_GRID_REAL = 0, PME_GRID_SPECTRUM = 1 };

const char* pme_last_error(void);
const char* pme_status_string(pme_status s);

/* ---------------- energy / virial buffer: CudaEnergyVirial ----------------
 * Device blob laid out as the reference's Virial_t (CudaEnergyVirial.h:13-18):
 * double sforce_dp[27][3]; long long sforce_fp[27][3]; double virmat[9];
 * then one double per energy term in insertion order. */
pme_status pme_ev_create(pme_ev_t* out);                                  /* CudaEnergyVirial.cu:139-145 */
pme_status pme_ev_destroy(pme_ev_t ev);                                   /* :150-153 */
pme_status pme_ev_insert(pme_ev_t ev, const char* name);                  /* EnergyVirial.cpp:16-27 */
pme_status pme_ev_clear(pme_ev_t ev, pme_stream_t stream);                /* CudaEnergyVirial.cu:172-176 */
pme_status pme_ev_clear_energy(pme_ev_t ev, pme_stream_t stream);         /* :181-187 */
pme_status pme_ev_clear_virial(pme_ev_t ev, pme_stream_t stream);         /* :192-197 */
pme_status pme_ev_copy_to_host(pme_ev_t ev, pme_stream_t stream);         /* :223-226 (async) */
pme_status pme_ev_get_energy(pme_ev_t ev, const char* name, double* out); /* :245-254 (host copy) */
pme_status pme_ev_get_virial(pme_ev_t ev, double virmat[9]);              /* :261-265 (host copy) */
pme_status pme_ev_get_sforce(pme_ev_t ev, double sforce[81]);             /* :270-278 */
pme_status pme_ev_energy_pointer(pme_ev_t ev, const char* name, double** dptr); /* :231-234 */
/* *dptr = start of the device Virial_t blob (sforce_dp[27][3] first; virmat at
 * double 162), as getVirialPointer returns (Virial_t*)d_buffer, :239-242 */
pme_status pme_ev_virial_pointer(pme_ev_t ev, double** dptr);

/* ---------------- the reciprocal solver: CudaPMERecip ---------------- */

/* CudaPMERecip(nfftx, nffty, nfftz, order, fft_type, nnode, mynode,
 *              energyVirial, nameRecip, nameSelf, stream), CudaPMERecip.cu:1412-1493.
 * order in {4, 6, 8}; every grid edge >= order and 2/3/5-smooth, <= 512;
 * nnode must be 1 (one GPU). */
pme_status pme_recip_create(int nfftx, int nffty, int nfftz, int order, int fft_type,
                            int nnode, int mynode, int accum_type, pme_ev_t ev,
                            const char* name_recip, const char* name_self,
                            pme_stream_t stream, pme_recip_t* out);
pme_status pme_recip_destroy(pme_recip_t h);                              /* :1645-1707 */
pme_status pme_recip_set_stream(pme_recip_t h, pme_stream_t stream);      /* :1618-1638 */
pme_status pme_recip_set_order(pme_recip_t h, int order);                 /* :2476-2479 */
/* Precision of the y and z FFT passes (no counterpart in the reference,
 * which has FP32 cuFFT only).  PME_PREC_FP32 (default): every stage FP32.
 * PME_PREC_BF16X3 / PME_PREC_FP16X3: on an axis whose length is a multiple
 * of 16, the radix-16 DFT stage runs on the matrix cores as a hi-lo split
 * product (x = hi + lo in BF16 or FP16; Ah*Bh + Ah*Bl + Al*Bh accumulated in
 * FP32; FP16 columns scaled by a power of two); the remaining stages, the x
 * pass, the influence function and the spline work stay FP32, and bsp_mod and
 * the energy/virial sums stay FP64.  Takes effect from the next spread. */
enum { PME_PREC_FP32 = 0, PME_PREC_BF16X3 = 1, PME_PREC_FP16X3 = 2 };
pme_status pme_recip_set_precision(pme_recip_t h, int precision);
pme_status pme_recip_get_info(pme_recip_t h, int* nfftx, int* nffty, int* nfftz,
                              int* order);                                /* CudaPMERecip.h:199-202 */

/* spread_charge(const float4* xyzq, int ncoord, const double* recip),
 * CudaPMERecip.cu:1797-1923.  d_xyzq: device float4[ncoord]; recip: host
 * double[9], row-major, must be diagonal (orthorhombic). */
pme_status pme_recip_spread_charge(pme_recip_t h, const void* d_xyzq, int ncoord,
                                   const double recip[9]);
pme_status pme_recip_r2c_fft(pme_recip_t h);                              /* :2352-2427 */
/* scalar_sum(recip, kappa, calc_energy, calc_virial), :1929-2047.  Energy is
 * ADDED to the name_recip slot, the virial ADDED into virmat. */
pme_status pme_recip_scalar_sum(pme_recip_t h, const double recip[9], double kappa,
                                int calc_energy, int calc_virial);
pme_status pme_recip_c2r_fft(pme_recip_t h);                              /* :2433-2470 */
/* gather_force<FT>(xyzq, ncoord, recip, stride, force), :2136-2252.  Output is
 * the gradient +dE/dr.  force_type: PME_FORCE_*. */
pme_status pme_recip_gather_force(pme_recip_t h, const void* d_xyzq, int ncoord,
                                  const double recip[9], int stride, int force_type,
                                  void* d_force);
/* calc_self_energy(xyzq, ncoord, kappa), :2053-2061: ADDS -kappa*ccelec/sqrt(pi)*sum q^2
 * to the name_self slot. */
pme_status pme_recip_calc_self_energy(pme_recip_t h, const void* d_xyzq, int ncoord,
                                      double kappa);

/* Whole solve in one call: spread -> r2c -> scalar_sum -> c2r -> gather
 * (the sequence of CudaDomdecRecip::calc, CudaDomdecRecip.h:50-57, without
 * the self energy).  Same results as the five calls. */
pme_status pme_recip_solve(pme_recip_t h, const void* d_xyzq, int ncoord,
                           const double recip[9], double kappa, int calc_energy,
                           int calc_virial, int stride, int force_type, void* d_force);

/* pme_recip_solve as a hipGraph replay (on = 1; default 0).  No counterpart in
 * the reference.  On a handle with a stream of its own (not the null stream)
 * the first solve with new arguments runs the stages, the second captures
 * them and later ones replay the graph; results are identical.  On ROCm 7 the
 * replay measured slower than the six launches it replaces (117.5 against
 * 108.3 us per solve on a 96,000-atom, 108x100x96 system), hence off by default. */
pme_status pme_recip_set_graph(pme_recip_t h, int on);

/* pme_recip_solve that allocates nothing.  No counterpart in the reference:
 * for callers that cannot call hipMalloc (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  The work buffers of the
 * solve (grid, spectrum, binning records, energy/virial partials) come from a
 * static device array of 64 MiB per GPU instead of the handle; a solve whose
 * buffers need more returns PME_ERR_BUDGET before any device work.  The
 * handle supplies its plan, stream and energy/virial buffer (created
 * beforehand), the caller inputs and outputs.  Results equal
 * pme_recip_solve's.  Calls that overlap on one GPU (other streams or
 * threads) race on the shared array; calls ordered on one stream do not.
 * Afterwards the handle's stage state is reset (the next call starts with
 * spread_charge or a solve). */
pme_status pme_recip_solve_noalloc(pme_recip_t h, const void* d_xyzq, int ncoord,
                                   const double recip[9], double kappa, int calc_energy,
                                   int calc_virial, int stride, int force_type, void* d_force);

/* Copy a grid view (device to device, on the handle's stream).
 * PME_GRID_REAL: charge grid after spread_charge, potential after c2r_fft,
 *   float[nfftz][nffty][nfftx].
 * PME_GRID_SPECTRUM: half spectrum after r2c_fft / scalar_sum,
 *   float2[nfftz][nffty][nfftx/2+1]. */
pme_status pme_recip_copy_grid(pme_recip_t h, int which, void* d_dst, size_t bytes);

/* Per-stage device time of the last solve stages (ms, from HIP events) if
 * enabled with pme_recip_set_timing(h, 1).  Order: bin, spread, fft_x_r2c,
 * fft_y_fwd, fft_z_solve, fft_y_inv, fft_x_c2r, gather. */
pme_status pme_recip_set_timing(pme_recip_t h, int enable);
pme_status pme_recip_get_timing(pme_recip_t h, float ms[8]);

/* ---------------- B-spline precompute: Bspline<float> ----------------
 * fill_bspline(xyzq, ncoord), Bspline.cu:203-231 (the reference has order 4
 * only): grid indices and theta/dtheta per atom, with the exact index formula
 * the spread and gather kernels use (CudaPMERecip.cu:512-529).
 * Outputs are device arrays: gi{x,y,z}[ncoord], charge[ncoord],
 * theta{x,y,z}[ncoord*order], dtheta{x,y,z}[ncoord*order] (atom-major). */
pme_status pme_fill_bspline(int order, int nfftx, int nffty, int nfftz,
                            const double recip[9], const void* d_xyzq, int ncoord,
                            int* d_gix, int* d_giy, int* d_giz, float* d_charge,
                            float* d_thetax, float* d_thetay, float* d_thetaz,
                            float* d_dthetax, float* d_dthetay, float* d_dthetaz,
                            pme_stream_t stream);

/* ---------------- batched path (no counterpart in the reference) ----------------
 * Many independent orthorhombic systems solved together: every stage is one
 * launch over all systems (persistent workgroups over per-stage work lists),
 * and the whole sequence is captured once in a hipGraph and replayed.  Each
 * system keeps its grids, FFT plans and binning store resident in HBM
 * ("slots"); admission is checked against the HBM budget of the batch.
 * A solve of a system is exactly the five-call sequence of the solver above
 * (same kernels, same results), energy and virial assigned per system. */
typedef struct pme_batch_s* pme_batch_t;

/* pme_system_desc.calc_energy_virial bits (1 keeps its meaning: energy and virial) */
enum { PME_CALC_EV = 1, PME_CALC_SELF = 2 };

typedef struct pme_system_desc {
  int nfftx, nffty, nfftz;   /* grid (2/3/5-smooth, >= order, <= 512) */
  int order;                 /* 4, 6 or 8 */
  double kappa;              /* 1/A */
  double recip[9];           /* row-major reciprocal box, diagonal */
  int ncoord;                /* atoms */
  const void* d_xyzq;        /* device float4[ncoord] */
  void* d_force;             /* device output, laid out per force_type */
  int force_type;            /* PME_FORCE_* (INT64 adds, the others store) */
  int stride;                /* SoA stride for FLOAT / INT64 (>= ncoord) */
  int calc_energy_virial;    /* PME_CALC_* bits: PME_CALC_EV the reciprocal energy and
                                virial, PME_CALC_SELF the self energy (calc_self_energy) */
  int precision;             /* PME_PREC_* of the y/z FFT passes */
} pme_system_desc;

pme_status pme_batch_create(size_t hbm_budget_bytes, pme_stream_t stream, pme_batch_t* out);
pme_status pme_batch_destroy(pme_batch_t b);
/* Admit a system: plan it and allocate its resident state.  PME_ERR_BUDGET
 * if it does not fit the remaining HBM budget (nothing is allocated then). */
pme_status pme_batch_add(pme_batch_t b, const pme_system_desc* sys, int* id);
/* New coordinates/box/outputs for a resident system (ncoord may grow up to
 * the capacity it was admitted with; a larger ncoord re-plans the slot). */
pme_status pme_batch_update(pme_batch_t b, int id, const pme_system_desc* sys);
pme_status pme_batch_remove(pme_batch_t b, int id);
/* Enqueue one solve of the listed systems (n = 0: every resident system) on
 * the batch's stream.  Forces are written on the device; energies/virials
 * are read with pme_batch_results. */
pme_status pme_batch_solve(pme_batch_t b, int n, const int* ids);
/* Wait for the last solve and copy energy[n] and virial[n*9] of the listed
 * systems (either pointer may be NULL). */
pme_status pme_batch_results(pme_batch_t b, int n, const int* ids, double* energy, double* virial);
/* Self energies -kappa*ccelec/sqrt(pi)*sum(q*q) (CudaPMERecip.cu:2053-2061)
 * of the listed systems from the last solve, for systems described with
 * PME_CALC_SELF (0 for the others); waits like pme_batch_results. */
pme_status pme_batch_results_self(pme_batch_t b, int n, const int* ids, double* energy_self);
/* hipGraph capture/replay of the solve sequence (default on). */
pme_status pme_batch_set_graph(pme_batch_t b, int enable);
/* Tuning knobs: brick edge target for spread/gather (>= 8), y/z FFT tile
 * width target (power of two), persistent workgroups per CU (0: one
 * workgroup per work item).  Re-plans every resident system. */
pme_status pme_batch_set_tuning(pme_batch_t b, int brick_target, int fft_tile, int wg_per_cu);
pme_status pme_batch_get_tuning(pme_batch_t b, int* brick_target, int* fft_tile, int* wg_per_cu);
/* Stage chains of a solve, 1 to 4 (default 2): the systems are dealt in turn,
 * in decreasing cost, into that many parts whose stage sequences run on as
 * many streams side by side (joined before the energy/virial copy), so one
 * part's kernel tails and launch gaps are filled by the others' work; 1 runs
 * one sequence.  Results do not depend on it.  Solves with stage timing on
 * run one chain. */
pme_status pme_batch_set_chains(pme_batch_t b, int nchains);
pme_status pme_batch_get_chains(pme_batch_t b, int* nchains);
/* Half bricks (on = 1; default 0): bricks planned within half the LDS of a
 * spread/gather workgroup, run as half-size workgroups, two per CU, so one's
 * HBM phases overlap the other's LDS work.  Re-plans every resident system;
 * results do not depend on it. */
pme_status pme_batch_set_half_bricks(pme_batch_t b, int on);
pme_status pme_batch_get_half_bricks(pme_batch_t b, int* on);
/* Autotuner: time the current resident set over a grid of tuning choices
 * (half bricks, brick target, FFT tile, workgroups per CU, stage chains), each
 * solve timed alone (as a caller that waits for every solve's results), keep
 * the fastest, and persist it in a JSON cache keyed by the set's signature and
 * the GPU architecture (cache_path may be NULL).  Returns the best time per
 * solve (ms). */
pme_status pme_batch_autotune(pme_batch_t b, const char* cache_path, float* best_ms);
pme_status pme_batch_memory(pme_batch_t b, size_t* used_bytes, size_t* budget_bytes);
/* Per-stage kernel times of the last solve (HIP events on the batch's
 * stream; solves run eagerly while timing is on).  ms[8]: bin, spread,
 * x forward, y forward, z forward+solve+inverse, y inverse, x inverse,
 * gather -- the order of pme_recip_get_timing. */
pme_status pme_batch_set_timing(pme_batch_t b, int enable);
pme_status pme_batch_get_timing(pme_batch_t b, float* ms);

/* ---------------- serving loop (no counterpart in the reference) ----------------
 * Continuous batching of PME solves that arrive one by one.  A worker thread
 * admits submitted requests into an HBM arena of hbm_budget_bytes (one
 * allocation at creation), keeps each system's grids, plans and binning
 * store resident under its key for follow-up requests (idle systems are
 * evicted least-recently-used first when space is needed), and launches
 * every admitted request that is not already in flight as one batched
 * solve on one of `nlanes` streams, so batches overlap on the device.  The
 * results of a request do not depend on what it was batched with (the
 * same kernels; integer charge accumulation).
 * Inputs (d_xyzq) must be complete when the request is submitted, or
 * ready_event (a hipEvent_t recorded after they were written) is given; the
 * caller keeps d_xyzq and d_force alive until the request completes. */
typedef struct pme_server_s* pme_server_t;

typedef struct pme_server_stats {
  unsigned long long submitted, completed, failed;
  unsigned long long batches;      /* batched solves launched */
  double mean_batch;               /* requests per batched solve */
  int max_batch;
  unsigned long long evictions;    /* resident systems evicted for space */
  unsigned long long admitted_new; /* requests that needed a new slot */
  int resident;                    /* systems resident now */
  size_t arena_used, arena_budget;
  double p50_ms, p99_ms, mean_ms, max_ms;  /* submit -> result host-visible */
} pme_server_stats;

pme_status pme_server_create(size_t hbm_budget_bytes, int nlanes, pme_server_t* out);
/* Finishes every submitted request, then stops the worker and frees. */
pme_status pme_server_destroy(pme_server_t s);
/* key >= 0 names the system: a resident key reuses its state (re-planned
 * only if the grid or order changes or ncoord exceeds what it was admitted
 * with).  PME_ERR_BUDGET if the system could never fit the arena. */
pme_status pme_server_submit(pme_server_t s, int key, const pme_system_desc* sys, void* ready_event,
                             unsigned long long* ticket);
/* The direct->recip request header of DomdecRecipComm.h:11-15 (same
 * layout: three doubles, an int, two one-byte flags).  A recip node serves
 * each header with a float4 coordinate array in and a float3 force array
 * out (CudaDomdecRecipLooper.cu:3-37); pme_server_submit_header does that
 * for a server-side system of fixed grid, order and kappa under `key`, as
 * CudaDomdecRecip::calc (CudaDomdecRecip.h:70-78): the reciprocal energy
 * and virial when either flag is set, and the self energy when calc_energy
 * is set (read with pme_server_wait_result).  A header with ncoord < 0 is
 * the reference's STOP message (DomdecRecipComm.cpp:140-144): it is
 * answered with PME_ERR_STOPPED and queues nothing. */
typedef struct pme_recip_header {
  double inv_boxx, inv_boxy, inv_boxz;
  int ncoord;
  unsigned char calc_energy, calc_virial;
} pme_recip_header;
pme_status pme_server_submit_header(pme_server_t s, int key, const pme_recip_header* h, int nfftx, int nffty,
                                    int nfftz, int order, double kappa, const void* d_xyzq, void* d_force3,
                                    unsigned long long* ticket);
/* Block until the request completes; returns its status.  energy/virial
 * (9)/latency_ms may be NULL.  A ticket's result can be taken once. */
pme_status pme_server_wait(pme_server_t s, unsigned long long ticket, double* energy, double* virial,
                           double* latency_ms);
/* Everything a request returns: reciprocal energy, self energy (requests
 * with PME_CALC_SELF), virial and submit-to-visible latency. */
typedef struct pme_request_result {
  double energy, energy_self;
  double virial[9];
  double latency_ms;
} pme_request_result;
pme_status pme_server_wait_result(pme_server_t s, unsigned long long ticket, pme_request_result* out);
/* Non-blocking: *done = 0 while in progress (result left in place). */
pme_status pme_server_poll(pme_server_t s, unsigned long long ticket, int* done, double* energy, double* virial,
                           double* latency_ms);
/* Wait until every request submitted so far has completed. */
pme_status pme_server_drain(pme_server_t s);
/* Drop a resident system (no-op for an unknown key, or one that is in flight
 * or queued at the time the worker handles the release). */
pme_status pme_server_release(pme_server_t s, int key);
pme_status pme_server_stats_get(pme_server_t s, pme_server_stats* out);
pme_status pme_server_stats_reset(pme_server_t s);
/* Plan of the server's batches (as pme_batch_set_tuning); default brick 24,
 * FFT tile 32, one workgroup per work item. */
pme_status pme_server_set_tuning(pme_server_t s, int brick_target, int fft_tile, int wg_per_cu);

/* Host FP64 bsp_mod table (calc_prefac/dftmod, CudaPMERecip.cu:2481-2579). */
pme_status pme_bsp_mod(int order, int n, double* out);

/* ---------------- direct-space Ewald: CudaPMEDirectForce<long long, float> ----------------
 * The real-space half of the Ewald sum over the reference's tile neighbour
 * list (CudaPMEDirectForce.h:82-204; kernels CudaDirectForce_util.h:184-251,
 * CudaDirectForceKernels.cu:352-760).  The reference's one instantiation has
 * AT = long long: forces are ADDED to a device long long array as 2^40 fixed
 * point, x at [i], y at [i + stride], z at [i + 2 stride], as the gradient
 * +dE/dr.  Energies are ADDED to the three named terms of the energy/virial
 * buffer, the shift forces of the virial to its sforce_dp[27][3].  Only the
 * electrostatic models are built: vdw_model must be PME_VDW_NONE. */
typedef struct pme_direct_s* pme_direct_t;
/* models, CudaDirectForceTypes.h:62-65 (EWALD_LOOKUP is evaluated analytically, as EWALD) */
enum { PME_VDW_NONE = 0, PME_ELEC_NONE = 0, PME_ELEC_EWALD = 101, PME_ELEC_EWALD_LOOKUP = 111 };
/* ientry_t (CudaNeighborListBuild.h:24-29): the i tile of 32 atoms starting at
 * iatomStart, shifted by box image ish = (sx+1) + 3(sy+1) + 9(sz+1), against the
 * j tiles tile_indj[tileStart..tileEnd] (inclusive) */
typedef struct { int iatomStart, ish, tileStart, tileEnd; } pme_ientry;
/* tile_excl_t<32> (CudaNeighborListBuild.h:19-22): bit t of excl[l] excludes
 * the pair (i = (l + t) mod 32, j = l) of the tile */
typedef struct { unsigned int excl[32]; } pme_tile_excl;
/* xx14list_t (Bonded_struct.h:122-124): pair (i, j), shift ishift added to xi - xj */
typedef struct { int i, j, ishift; } pme_xx14;
/* the tile list of CudaNeighborListBuild<32> (get_n_ientry, get_ientry,
 * get_n_tile, get_tile_indj, get_tile_excl, :186-200), device pointers; every
 * tile holds 32 atoms of the coordinate array (the reference pads to 32) */
typedef struct {
  int n_ientry;
  const pme_ientry* ientry;
  int n_tile;
  const int* tile_indj;
  const pme_tile_excl* tile_excl;
} pme_nlist;

/* CudaPMEDirectForce(energyVirial, nameVdw, nameElec, nameExcl), CudaPMEDirectForce.cu:62-118 */
pme_status pme_direct_create(pme_ev_t ev, const char* name_vdw, const char* name_elec, const char* name_excl,
                             pme_direct_t* out);
pme_status pme_direct_destroy(pme_direct_t h);                              /* :123-134 */
/* setup(boxx, boxy, boxz, kappa, roff, ron, e14fac, vdw_model, elec_model), :151-258 */
pme_status pme_direct_setup(pme_direct_t h, double boxx, double boxy, double boxz, double kappa, double roff,
                            double ron, double e14fac, int vdw_model, int elec_model);
pme_status pme_direct_get_setup(pme_direct_t h, float* boxx, float* boxy, float* boxz, float* kappa, float* roff,
                                float* ron, float* e14fac, int* vdw_model, int* elec_model); /* :263-276 */
pme_status pme_direct_set_box_size(pme_direct_t h, float boxx, float boxy, float boxz); /* :293-298 */
pme_status pme_direct_set_calc_vdw(pme_direct_t h, int calc_vdw);          /* :304-306 */
pme_status pme_direct_set_calc_elec(pme_direct_t h, int calc_elec);        /* :312-314 */
/* set_14_list(nin14list, nex14list, h_in14list, h_ex14list, stream), :575-593
 * (host lists, copied before the call returns) */
pme_status pme_direct_set_14_list(pme_direct_t h, int nin14, int nex14, const pme_xx14* h_in14,
                                  const pme_xx14* h_ex14, pme_stream_t stream);
/* calc_14_force(xyzq, calc_energy, calc_virial, stride, force, stream), :671-716:
 * 1-4 interactions, qq (erfc + e14fac - 1)/r into the elec term, and 1-4
 * exclusions, qq (erfc - 1)/r into the excl term */
pme_status pme_direct_calc_14_force(pme_direct_t h, const void* d_xyzq, int calc_energy, int calc_virial,
                                    int stride, long long* d_force, pme_stream_t stream);
/* calc_force(xyzq, nlist, calc_energy, calc_virial, stride, force, stream),
 * :719-770: qq erfc(kappa r)/r over the list's pairs closer than roff into
 * the elec term */
pme_status pme_direct_calc_force(pme_direct_t h, const void* d_xyzq, const pme_nlist* nlist, int calc_energy,
                                 int calc_virial, int stride, long long* d_force, pme_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* PME_RECIP_H */
