// pme_host.cpp -- host side of the PME reciprocal engine: the C ABI of
// include/pme_recip.h, per-handle plans and buffers, and the stage state
// machine that lets the reference's five-call sequence run as fused kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pme_recip.h"
#include "pme_internal.hpp"
#include "pme_plan.hpp"

using namespace pme;

// ------------------------------------------------------------------ errors --
static thread_local std::string g_err;

pme_status pme::set_error(pme_status s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return s;
}

#define fail pme::set_error

#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(e_ == hipErrorOutOfMemory ? PME_ERR_OOM : PME_ERR_DEVICE, "%s: %s (%s:%d)", #expr, \
                  hipGetErrorString(e_), __FILE__, __LINE__);                             \
  } while (0)

#define PME_TRY(expr)             \
  do {                            \
    pme_status s_ = (expr);       \
    if (s_ != PME_OK) return s_;  \
  } while (0)

extern "C" const char* pme_last_error(void) { return g_err.c_str(); }

extern "C" const char* pme_status_string(pme_status s) {
  switch (s) {
    case PME_OK: return "ok";
    case PME_ERR_INVALID_ARG: return "invalid argument";
    case PME_ERR_UNSUPPORTED: return "unsupported";
    case PME_ERR_OOM: return "out of memory";
    case PME_ERR_BUDGET: return "over budget";
    case PME_ERR_DEVICE: return "device error";
    case PME_ERR_STATE: return "stage called out of order";
    case PME_ERR_NOT_FOUND: return "not found";
    case PME_ERR_STOPPED: return "stop request";
  }
  return "unknown";
}

struct pme_recip_s;

// ---------------------------------------------------------- energy/virial --
// Virial_t layout, CudaEnergyVirial.h:13-18
static const size_t kOffSforceDp = 0;
static const size_t kOffSforceFp = 81 * sizeof(double);
static const size_t kOffVirmat = kOffSforceFp + 81 * sizeof(long long);
static const size_t kOffEnergy = kOffVirmat + 9 * sizeof(double);
static const int kMaxTerms = 64;
static const size_t kEvBytes = kOffEnergy + kMaxTerms * sizeof(double);

struct pme_energy_virial_s {
  std::map<std::string, int> index;
  int n = 0;
  char* d_buf = nullptr;
  char* h_buf = nullptr;
  std::vector<pme_recip_t> solvers;  // solvers writing into this buffer
};

// A solver may hold per-workgroup energy/virial partials whose total is not
// yet in the device buffer (it is normally folded into the next kernel);
// every operation that reads or clears the buffer flushes them first.
static pme_status flush_pending(pme_recip_t h, hipStream_t target, bool order_target);
static pme_status ev_flush(pme_ev_t ev, hipStream_t target, bool order_target = true) {
  for (pme_recip_t h : ev->solvers) PME_TRY(flush_pending(h, target, order_target));
  return PME_OK;
}

extern "C" pme_status pme_ev_create(pme_ev_t* out) {
  if (!out) return fail(PME_ERR_INVALID_ARG, "pme_ev_create: null out");
  pme_ev_t ev = new pme_energy_virial_s();
  hipError_t e = hipMalloc(&ev->d_buf, kEvBytes);
  if (e != hipSuccess) {
    delete ev;
    return fail(PME_ERR_OOM, "pme_ev_create: hipMalloc: %s", hipGetErrorString(e));
  }
  e = hipHostMalloc(&ev->h_buf, kEvBytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipFree(ev->d_buf);
    delete ev;
    return fail(PME_ERR_OOM, "pme_ev_create: hipHostMalloc: %s", hipGetErrorString(e));
  }
  std::memset(ev->h_buf, 0, kEvBytes);
  e = hipMemset(ev->d_buf, 0, kEvBytes);
  if (e != hipSuccess) return fail(PME_ERR_DEVICE, "pme_ev_create: hipMemset: %s", hipGetErrorString(e));
  *out = ev;
  return PME_OK;
}

extern "C" pme_status pme_ev_destroy(pme_ev_t ev) {
  if (!ev) return PME_OK;
  (void)hipFree(ev->d_buf);
  (void)hipHostFree(ev->h_buf);
  delete ev;
  return PME_OK;
}

extern "C" pme_status pme_ev_insert(pme_ev_t ev, const char* name) {
  if (!ev || !name) return fail(PME_ERR_INVALID_ARG, "pme_ev_insert: null argument");
  if (ev->index.count(name)) return PME_OK;  // only added if new, EnergyVirial.cpp:16-21
  if (ev->n >= kMaxTerms) return fail(PME_ERR_UNSUPPORTED, "pme_ev_insert: more than %d energy terms", kMaxTerms);
  ev->index[name] = ev->n++;
  return PME_OK;
}

static pme_status ev_slot(pme_ev_t ev, const char* name, int* slot) {
  if (!ev || !name) return fail(PME_ERR_INVALID_ARG, "energy term: null argument");
  auto it = ev->index.find(name);
  if (it == ev->index.end()) return fail(PME_ERR_NOT_FOUND, "energy term %s not found", name);
  *slot = it->second;
  return PME_OK;
}

extern "C" pme_status pme_ev_clear(pme_ev_t ev, pme_stream_t stream) {
  if (!ev) return fail(PME_ERR_INVALID_ARG, "pme_ev_clear: null ev");
  PME_TRY(ev_flush(ev, (hipStream_t)stream));
  HIP_TRY(launch_fill_zero(ev->d_buf, kEvBytes, (hipStream_t)stream));
  std::memset(ev->h_buf, 0, kEvBytes);
  return PME_OK;
}

extern "C" pme_status pme_ev_clear_energy(pme_ev_t ev, pme_stream_t stream) {
  if (!ev) return fail(PME_ERR_INVALID_ARG, "pme_ev_clear_energy: null ev");
  PME_TRY(ev_flush(ev, (hipStream_t)stream));
  HIP_TRY(launch_fill_zero(ev->d_buf + kOffEnergy, kMaxTerms * sizeof(double), (hipStream_t)stream));
  std::memset(ev->h_buf + kOffEnergy, 0, kMaxTerms * sizeof(double));
  return PME_OK;
}

extern "C" pme_status pme_ev_clear_virial(pme_ev_t ev, pme_stream_t stream) {
  if (!ev) return fail(PME_ERR_INVALID_ARG, "pme_ev_clear_virial: null ev");
  PME_TRY(ev_flush(ev, (hipStream_t)stream));
  HIP_TRY(launch_fill_zero(ev->d_buf, kOffEnergy, (hipStream_t)stream));
  std::memset(ev->h_buf, 0, kOffEnergy);
  return PME_OK;
}

extern "C" pme_status pme_ev_copy_to_host(pme_ev_t ev, pme_stream_t stream) {
  if (!ev) return fail(PME_ERR_INVALID_ARG, "pme_ev_copy_to_host: null ev");
  PME_TRY(ev_flush(ev, (hipStream_t)stream));
  HIP_TRY(hipMemcpyAsync(ev->h_buf, ev->d_buf, kEvBytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  return PME_OK;
}

extern "C" pme_status pme_ev_get_energy(pme_ev_t ev, const char* name, double* out) {
  int slot;
  PME_TRY(ev_slot(ev, name, &slot));
  if (!out) return fail(PME_ERR_INVALID_ARG, "pme_ev_get_energy: null out");
  std::memcpy(out, ev->h_buf + kOffEnergy + slot * sizeof(double), sizeof(double));
  return PME_OK;
}

extern "C" pme_status pme_ev_get_virial(pme_ev_t ev, double virmat[9]) {
  if (!ev || !virmat) return fail(PME_ERR_INVALID_ARG, "pme_ev_get_virial: null argument");
  std::memcpy(virmat, ev->h_buf + kOffVirmat, 9 * sizeof(double));
  return PME_OK;
}

extern "C" pme_status pme_ev_get_sforce(pme_ev_t ev, double sforce[81]) {
  if (!ev || !sforce) return fail(PME_ERR_INVALID_ARG, "pme_ev_get_sforce: null argument");
  const double* dp = reinterpret_cast<const double*>(ev->h_buf + kOffSforceDp);
  const long long* fp = reinterpret_cast<const long long*>(ev->h_buf + kOffSforceFp);
  for (int i = 0; i < 81; i++) sforce[i] = dp[i] + (double)fp[i] / (double)(1ll << 30);
  return PME_OK;
}

extern "C" pme_status pme_ev_energy_pointer(pme_ev_t ev, const char* name, double** dptr) {
  int slot;
  PME_TRY(ev_slot(ev, name, &slot));
  if (!dptr) return fail(PME_ERR_INVALID_ARG, "pme_ev_energy_pointer: null out");
  PME_TRY(ev_flush(ev, nullptr, false));
  *dptr = reinterpret_cast<double*>(ev->d_buf + kOffEnergy) + slot;
  return PME_OK;
}

extern "C" pme_status pme_ev_virial_pointer(pme_ev_t ev, double** dptr) {
  if (!ev || !dptr) return fail(PME_ERR_INVALID_ARG, "pme_ev_virial_pointer: null argument");
  PME_TRY(ev_flush(ev, nullptr, false));
  // the start of the Virial_t blob (sforce_dp first, virmat at double 162),
  // as getVirialPointer returns it (CudaEnergyVirial.cu:239-242)
  *dptr = reinterpret_cast<double*>(ev->d_buf);
  return PME_OK;
}

// -------------------------------------------------------------- host math --
static const double kPi = 3.14159265358979323846;

extern "C" pme_status pme_bsp_mod(int order, int n, double* out) {
  if (order < 3 || order > 16 || n < 2 || !out) return fail(PME_ERR_INVALID_ARG, "pme_bsp_mod: bad arguments");
  bsp_mod_host(order, n, out);
  return PME_OK;
}

// ------------------------------------------------------------------ handle --
enum GridState {
  ST_EMPTY = 0,      // nothing computed yet
  ST_CHARGE,         // real grid holds the charge grid
  ST_XY,             // spectrum transformed along x and y only
  ST_SPECTRUM,       // full half spectrum (possibly solved)
  ST_SOLVED_ZINV,    // solved spectrum, z already inverse transformed
  ST_POTENTIAL       // real grid holds the potential
};


// Everything the launches of one pme_recip_solve depend on: the arguments,
// the handle's stream, plan and buffers, and the self-energy partials the
// inverse transform folds in.  Equal keys give the same launch sequence.
struct SolveKey {
  const void* xyzq;
  void* force;
  hipStream_t stream;
  const void* buf[6];
  double recip[9];
  double kappa, pend_kcs;
  int n, calc_energy, calc_virial, stride, force_type;
  int order, prec, pend_self, nbins, chunk, nchunk, nchunk_cap, bins_for, partial_cap;
  int nb[3], wmax[3];
};

struct pme_recip_s {
  int nx = 0, ny = 0, nz = 0, order = 0, fft_type = PME_FFT_BOX, accum = PME_ACCUM_FLOAT;
  hipStream_t stream = nullptr;
  pme_ev_t ev = nullptr;
  int slot_recip = 0, slot_self = 0;

  float* grid = nullptr;      // real grid nx*ny*nz
  float2* cgrid = nullptr;    // half spectrum (nx/2+1)*ny*nz
  float* prefac[3] = {nullptr, nullptr, nullptr};
  float2* tw[3] = {nullptr, nullptr, nullptr};
  RadixPlan rp[3];
  RadixPlan rp_split[3];      // radix-16 first stage (y, z) for the split-precision modes
  bool has_split[3] = {false, false, false};
  int prec = PME_PREC_FP32;
  bool plane = false;         // x and y passes as one xy plane kernel (the plane fits one workgroup's LDS)
  int NR = 16, TXy = 16, TXz = 16;

  // bricks and the chunked record store of the binning
  int nb[3] = {1, 1, 1}, wmax[3] = {0, 0, 0}, nbins = 0;
  int bins_for = -1;          // atom capacity of the record store
  int chunk = 0, nchunk = 0, nchunk_cap = 0;
  int* seg_off = nullptr;     // [nchunk_cap][seg_stride(nbins)]
  float4* rec_pos = nullptr;  // [nchunk_cap * 8 * chunk]
  int* rec_idx = nullptr;

  // energy/virial partials and the reduction still pending on them
  double* partial = nullptr;       // [partial_cap * 7] scalar-sum partials
  int partial_cap = 0;
  double* self_partial = nullptr;  // [1024]
  int pend_solve = 0, pend_self = 0;
  double pend_kcs = 0.0;
  hipEvent_t flush_ev = nullptr;

  int state = ST_EMPTY;
  const void* bin_key_ptr = nullptr;
  int bin_key_n = -1;
  float bin_key_r[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // (float) recip of the last binning

  bool timing = false;
  hipEvent_t tev[9] = {};

  // pme_recip_solve replayed as a graph (solve_graph), when use_graph is set
  bool use_graph = false;
  SolveKey gkey{};
  bool gkey_valid = false;
  hipGraphExec_t gexec = nullptr;
};

static void free_handle(pme_recip_t h) {
  if (!h) return;
  if (h->ev) {
    auto& v = h->ev->solvers;
    for (size_t i = 0; i < v.size(); i++)
      if (v[i] == h) {
        v.erase(v.begin() + i);
        break;
      }
  }
  (void)hipFree(h->grid);
  (void)hipFree(h->cgrid);
  for (int i = 0; i < 3; i++) {
    (void)hipFree(h->prefac[i]);
    (void)hipFree(h->tw[i]);
  }
  (void)hipFree(h->seg_off);
  (void)hipFree(h->rec_pos);
  (void)hipFree(h->rec_idx);
  (void)hipFree(h->partial);
  (void)hipFree(h->self_partial);
  if (h->flush_ev) (void)hipEventDestroy(h->flush_ev);
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  for (auto& e : h->tev)
    if (e) (void)hipEventDestroy(e);
  delete h;
}

static EvReduce pending_reduce(pme_recip_t h) {
  EvReduce r;
  r.partial = h->partial;
  r.nsolve = h->pend_solve;
  r.energy = reinterpret_cast<double*>(h->ev->d_buf + kOffEnergy) + h->slot_recip;
  r.virmat = reinterpret_cast<double*>(h->ev->d_buf + kOffVirmat);
  r.self_partial = h->self_partial;
  r.nself = h->pend_self;
  r.self_slot = reinterpret_cast<double*>(h->ev->d_buf + kOffEnergy) + h->slot_self;
  r.kcs = h->pend_kcs;
  return r;
}

static EvReduce no_reduce() {
  EvReduce r{};
  return r;
}

// Launch the pending energy/virial reduction on the solver's stream; order
// `target` after it when it is another stream.
static pme_status flush_pending(pme_recip_t h, hipStream_t target, bool order_target) {
  if (h->pend_solve == 0 && h->pend_self == 0) return PME_OK;
  HIP_TRY(launch_ev_reduce(pending_reduce(h), h->stream));
  h->pend_solve = h->pend_self = 0;
  if (order_target && target != h->stream) {
    HIP_TRY(hipEventRecord(h->flush_ev, h->stream));
    HIP_TRY(hipStreamWaitEvent(target, h->flush_ev, 0));
  }
  return PME_OK;
}

static pme_status upload_prefac(pme_recip_t h) {
  const int n[3] = {h->nx, h->ny, h->nz};
  for (int a = 0; a < 3; a++) {
    std::vector<double> m(n[a]);
    bsp_mod_host(h->order, n[a], m.data());
    std::vector<float> f(n[a]);
    for (int i = 0; i < n[a]; i++) f[i] = (float)m[i];
    HIP_TRY(hipMemcpy(h->prefac[a], f.data(), sizeof(float) * n[a], hipMemcpyHostToDevice));
  }
  return PME_OK;
}

static bool plan_bricks(pme_recip_t h) {
  const int n[3] = {h->nx, h->ny, h->nz};
  BrickPlan bp;
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      ncu <= 0)
    ncu = 256;
  if (!pme::plan_bricks_auto(n, h->order, ncu, &bp)) return false;
  for (int a = 0; a < 3; a++) {
    h->nb[a] = bp.nb[a];
    h->wmax[a] = bp.wmax[a];
  }
  h->nbins = bp.nbins;
  return true;
}

// Size the record store for n atoms: chunks of kMinChunk atoms (more per
// chunk beyond kMaxChunks chunks), 8 record slots per atom (a stencil touches
// at most 8 bricks).
static pme_status ensure_bins(pme_recip_t h, int n) {
  const ChunkPlan cp = plan_chunks(n);
  const int chunk = cp.chunk, nchunk = cp.nchunk;
  if (n <= h->bins_for && chunk == h->chunk && nchunk <= h->nchunk_cap) {
    h->nchunk = nchunk;
    return PME_OK;
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  (void)hipFree(h->seg_off);
  (void)hipFree(h->rec_pos);
  (void)hipFree(h->rec_idx);
  h->seg_off = nullptr;
  h->rec_pos = nullptr;
  h->rec_idx = nullptr;
  h->bins_for = -1;
  const size_t rs = 8 * (size_t)chunk;
  HIP_TRY(hipMalloc(&h->seg_off, sizeof(int) * (size_t)nchunk * seg_stride(h->nbins)));
  HIP_TRY(hipMalloc(&h->rec_pos, sizeof(float4) * rs * nchunk));
  HIP_TRY(hipMalloc(&h->rec_idx, sizeof(int) * rs * nchunk));
  h->chunk = chunk;
  h->nchunk = nchunk;
  h->nchunk_cap = nchunk;
  h->bins_for = std::max(n, 1);
  return PME_OK;
}

static bool is_ortho(const double* r) {
  return r[1] == 0.0 && r[2] == 0.0 && r[3] == 0.0 && r[5] == 0.0 && r[6] == 0.0 && r[7] == 0.0;
}

// det(recip) = 1/V; a usable box has a finite, positive one
static double recip_det(const double* r) {
  return r[0] * (r[4] * r[8] - r[5] * r[7]) - r[1] * (r[3] * r[8] - r[5] * r[6]) + r[2] * (r[3] * r[7] - r[4] * r[6]);
}
static bool box_ok(const double* r) {
  const double d = recip_det(r);
  return std::isfinite(d) && d > 0.0;
}

static BinArgs bin_args(pme_recip_t h, const double* recip) {
  BinArgs a;
  a.nx = h->nx;
  a.ny = h->ny;
  a.nz = h->nz;
  a.r1 = (float)recip[0];
  a.r2 = (float)recip[4];
  a.r3 = (float)recip[8];
  a.rxy = (float)recip[1];
  a.rxz = (float)recip[2];
  a.ryx = (float)recip[3];
  a.ryz = (float)recip[5];
  a.rzx = (float)recip[6];
  a.rzy = (float)recip[7];
  a.tric = is_ortho(recip) ? 0 : 1;
  a.nbx = h->nb[0];
  a.nby = h->nb[1];
  a.nbz = h->nb[2];
  a.order = h->order;
  a.chunk = h->chunk;
  a.nchunk = h->nchunk;
  a.rec_stride = 8 * h->chunk;
  a.seg_off = h->seg_off;
  a.rec_pos = h->rec_pos;
  a.rec_idx = h->rec_idx;
  return a;
}

static AxisArgs axis_args(pme_recip_t h, int a) {
  AxisArgs ax;
  const bool split = h->prec != PME_PREC_FP32 && h->has_split[a];
  ax.rp = split ? h->rp_split[a] : h->rp[a];
  ax.tw = h->tw[a];
  ax.ntw = a == 0 ? h->nx : (a == 1 ? h->ny : h->nz);
  ax.prec = split ? h->prec : PME_PREC_FP32;
  // the fused z pass runs as a register four-step FFT where the length has a
  // factor pair, the y passes where their radix plan's pair is in the table
  ax.f4 = (a == 2 && !split) ? four_step_code(h->nz) : -1;
  ax.f4col = (a == 1 && !split) ? four_step_code_for_plan(ax.rp) : -1;
  return ax;
}


static SolveArgs solve_args(pme_recip_t h, const double* recip, double kappa, bool calc_ev) {
  SolveArgs s;
  s.n1 = h->nx;
  s.n2 = h->ny;
  s.n3 = h->nz;
  s.nf1 = h->nx / 2 + (h->nx % 2);
  s.nf2 = h->ny / 2 + (h->ny % 2);
  s.nf3 = h->nz / 2 + (h->nz % 2);
  s.r1 = (float)recip[0];
  s.r2 = (float)recip[4];
  s.r3 = (float)recip[8];
  // CudaPMERecip.cu:2024-2026: inv_vol in FP64, then cast; det(recip) for a triclinic box
  s.tric = is_ortho(recip) ? 0 : 1;
  const double inv_vol = s.tric ? recip_det(recip) : recip[0] * recip[4] * recip[8];
  s.piv_inv = (float)(inv_vol / kPi);
  for (int i = 0; i < 9; i++) s.rt[i] = (float)recip[i];
  s.fac = (float)(kPi * kPi / (kappa * kappa));
  s.pf1 = h->prefac[0];
  s.pf2 = h->prefac[1];
  s.pf3 = h->prefac[2];
  s.calc_ev = calc_ev ? 1 : 0;
  s.partial = h->partial;
  return s;
}

static void pick_tiles(pme_recip_t h) {
  // the z tile is sized for the FP32 four-step z pass (one LDS buffer); a split
  // precision set later runs that tile in two buffers, within the 160 KB of LDS
  const FftTiles t = pick_fft_tiles(h->nx, h->ny, h->nz, h->rp[0].n, kDefaultFftTile, four_step_code(h->nz) >= 0);
  h->NR = t.NR;
  h->TXy = t.TXy;
  h->TXz = t.TXz;
}

extern "C" pme_status pme_recip_create(int nfftx, int nffty, int nfftz, int order, int fft_type, int nnode,
                                       int mynode, int accum_type, pme_ev_t ev, const char* name_recip,
                                       const char* name_self, pme_stream_t stream, pme_recip_t* out) {
  if (!out || !ev || !name_recip || !name_self) return fail(PME_ERR_INVALID_ARG, "pme_recip_create: null argument");
  *out = nullptr;
  if (fft_type != PME_FFT_BOX && fft_type != PME_FFT_SLAB && fft_type != PME_FFT_COLUMN)
    return fail(PME_ERR_INVALID_ARG, "pme_recip_create: fft_type %d", fft_type);
  if (nnode != 1 || mynode != 0)
    return fail(PME_ERR_UNSUPPORTED, "pme_recip_create: nnode=%d: one GPU per solver", nnode);
  if (accum_type != PME_ACCUM_FLOAT && accum_type != PME_ACCUM_INT32 && accum_type != PME_ACCUM_INT64)
    return fail(PME_ERR_INVALID_ARG, "pme_recip_create: accum_type %d", accum_type);
  if (order != 4 && order != 6 && order != 8)
    return fail(PME_ERR_UNSUPPORTED, "pme_recip_create: order %d (4, 6 or 8)", order);
  const int n[3] = {nfftx, nffty, nfftz};
  for (int a = 0; a < 3; a++) {
    if (n[a] < order || n[a] > 512 || !smooth235(n[a]))
      return fail(PME_ERR_UNSUPPORTED, "pme_recip_create: grid edge %d (need order <= n <= 512, 2/3/5-smooth)", n[a]);
  }
  pme_recip_t h = new pme_recip_s();
  h->nx = nfftx;
  h->ny = nffty;
  h->nz = nfftz;
  h->order = order;
  h->fft_type = fft_type;
  h->accum = accum_type;
  h->stream = (hipStream_t)stream;
  if (!plan_bricks(h)) {
    delete h;
    return fail(PME_ERR_UNSUPPORTED, "pme_recip_create: no brick plan fits the LDS budget");
  }
  // CudaPMERecip.cu:1426-1429
  pme_status st = pme_ev_insert(ev, name_recip);
  if (st == PME_OK) st = pme_ev_insert(ev, name_self);
  if (st == PME_OK) st = ev_slot(ev, name_recip, &h->slot_recip);
  if (st == PME_OK) st = ev_slot(ev, name_self, &h->slot_self);
  if (st != PME_OK) {
    delete h;
    return st;
  }
  auto init = [&]() -> pme_status {
    const size_t G = (size_t)h->nx * h->ny * h->nz;
    const size_t Gc = (size_t)(h->nx / 2 + 1) * h->ny * h->nz;
    HIP_TRY(hipMalloc(&h->grid, sizeof(float) * G));
    HIP_TRY(hipMalloc(&h->cgrid, sizeof(float2) * (size_t)spec_pitch(h->nx / 2 + 1, h->ny) * h->nz));
    for (int a = 0; a < 3; a++) {
      HIP_TRY(hipMalloc(&h->prefac[a], sizeof(float) * n[a]));
      HIP_TRY(hipMalloc(&h->tw[a], sizeof(float2) * n[a]));
      const std::vector<float2> t = twiddles(n[a]);
      HIP_TRY(hipMemcpy(h->tw[a], t.data(), sizeof(float2) * n[a], hipMemcpyHostToDevice));
    }
    const bool even = (h->nx % 2) == 0;
    h->has_split[1] = radix_plan_split(h->ny, 1, &h->rp_split[1]);
    h->has_split[2] = radix_plan_split(h->nz, 1, &h->rp_split[2]);
    if (!radix_plan(even ? h->nx / 2 : h->nx, even ? 2 : 1, &h->rp[0]) || !radix_plan(h->ny, 1, &h->rp[1]) ||
        !radix_plan(h->nz, 1, &h->rp[2]))
      return fail(PME_ERR_UNSUPPORTED, "pme_recip_create: no radix plan");
    pick_tiles(h);
    h->plane = fft_plane_lds_bytes(h->nx, h->ny, h->rp[0].n, 1) != 0;
    PME_TRY(upload_prefac(h));
    const int hx = h->nx / 2 + 1;
    h->partial_cap = std::max(cols_blocks(hx * h->ny, h->TXz, 1), scalar_sum_blocks((int)Gc));
    HIP_TRY(hipMalloc(&h->partial, sizeof(double) * 7 * (size_t)h->partial_cap));
    HIP_TRY(hipMalloc(&h->self_partial, sizeof(double) * 1024));
    HIP_TRY(hipEventCreateWithFlags(&h->flush_ev, hipEventDisableTiming));
    for (auto& e : h->tev) HIP_TRY(hipEventCreate(&e));
    return PME_OK;
  };
  st = init();
  if (st != PME_OK) {
    free_handle(h);
    return st;
  }
  h->ev = ev;
  ev->solvers.push_back(h);
  *out = h;
  return PME_OK;
}

extern "C" pme_status pme_recip_destroy(pme_recip_t h) {
  if (!h) return PME_OK;
  (void)flush_pending(h, h->stream, false);
  (void)hipStreamSynchronize(h->stream);
  free_handle(h);
  return PME_OK;
}

extern "C" pme_status pme_recip_set_stream(pme_recip_t h, pme_stream_t stream) {
  if (!h) return fail(PME_ERR_INVALID_ARG, "pme_recip_set_stream: null handle");
  // partials written on the old stream are reduced there first
  PME_TRY(flush_pending(h, (hipStream_t)stream, true));
  h->stream = (hipStream_t)stream;
  return PME_OK;
}

extern "C" pme_status pme_recip_set_precision(pme_recip_t h, int precision) {
  if (!h) return fail(PME_ERR_INVALID_ARG, "pme_recip_set_precision: null handle");
  if (precision < PME_PREC_FP32 || precision > PME_PREC_FP16X3)
    return fail(PME_ERR_INVALID_ARG, "pme_recip_set_precision: precision %d", precision);
  if (h->state != ST_EMPTY && h->state != ST_CHARGE && h->state != ST_POTENTIAL)
    return fail(PME_ERR_STATE, "pme_recip_set_precision: a transform is in progress");
  h->prec = precision;
  return PME_OK;
}

extern "C" pme_status pme_recip_set_order(pme_recip_t h, int order) {
  if (!h) return fail(PME_ERR_INVALID_ARG, "pme_recip_set_order: null handle");
  if (order != 4 && order != 6 && order != 8) return fail(PME_ERR_UNSUPPORTED, "pme_recip_set_order: order %d", order);
  if (h->nx < order || h->ny < order || h->nz < order)
    return fail(PME_ERR_UNSUPPORTED, "pme_recip_set_order: grid smaller than order %d", order);
  PME_TRY(flush_pending(h, h->stream, false));
  HIP_TRY(hipStreamSynchronize(h->stream));
  const int old_order = h->order, old_nbins = h->nbins;
  int old_nb[3] = {h->nb[0], h->nb[1], h->nb[2]}, old_w[3] = {h->wmax[0], h->wmax[1], h->wmax[2]};
  h->order = order;
  if (!plan_bricks(h)) {
    h->order = old_order;
    h->nbins = old_nbins;
    for (int a = 0; a < 3; a++) {
      h->nb[a] = old_nb[a];
      h->wmax[a] = old_w[a];
    }
    return fail(PME_ERR_UNSUPPORTED, "pme_recip_set_order: no brick plan fits the LDS budget");
  }
  PME_TRY(upload_prefac(h));
  if (h->nbins != old_nbins) {  // segment table is sized by the brick count
    (void)hipFree(h->seg_off);
    h->seg_off = nullptr;
    h->bins_for = -1;
    h->nchunk_cap = 0;
  }
  h->bin_key_n = -1;
  h->state = ST_EMPTY;
  return PME_OK;
}

extern "C" pme_status pme_recip_get_info(pme_recip_t h, int* nx, int* ny, int* nz, int* order) {
  if (!h) return fail(PME_ERR_INVALID_ARG, "pme_recip_get_info: null handle");
  if (nx) *nx = h->nx;
  if (ny) *ny = h->ny;
  if (nz) *nz = h->nz;
  if (order) *order = h->order;
  return PME_OK;
}

static void mark(pme_recip_t h, int i) {
  if (h->timing) (void)hipEventRecord(h->tev[i], h->stream);
}

static pme_status do_bin(pme_recip_t h, const void* xyzq, int n, const double* recip) {
  PME_TRY(ensure_bins(h, n));
  BinArgs a = bin_args(h, recip);
  HIP_TRY(launch_bin_atoms((const float4*)xyzq, n, a, h->stream));
  h->bin_key_ptr = xyzq;
  h->bin_key_n = n;
  for (int i = 0; i < 9; i++) h->bin_key_r[i] = (float)recip[i];
  return PME_OK;
}

extern "C" pme_status pme_recip_spread_charge(pme_recip_t h, const void* d_xyzq, int ncoord, const double recip[9]) {
  if (!h || !recip || (ncoord > 0 && !d_xyzq)) return fail(PME_ERR_INVALID_ARG, "pme_recip_spread_charge: null argument");
  if (ncoord < 0) return fail(PME_ERR_INVALID_ARG, "pme_recip_spread_charge: ncoord %d", ncoord);
  if (!box_ok(recip)) return fail(PME_ERR_INVALID_ARG, "pme_recip_spread_charge: recip has no positive determinant");
  mark(h, 0);
  PME_TRY(do_bin(h, d_xyzq, ncoord, recip));
  mark(h, 1);
  BinArgs a = bin_args(h, recip);
  HIP_TRY(launch_spread(h->order, a, h->wmax, h->grid, h->stream));
  h->state = ST_CHARGE;
  return PME_OK;
}

extern "C" pme_status pme_recip_r2c_fft(pme_recip_t h) {
  if (!h) return fail(PME_ERR_INVALID_ARG, "pme_recip_r2c_fft: null handle");
  if (h->state != ST_CHARGE && h->state != ST_POTENTIAL)
    return fail(PME_ERR_STATE, "pme_recip_r2c_fft: no real grid (call spread_charge first)");
  const int hx = h->nx / 2 + 1;
  SolveArgs dummy{};
  mark(h, 2);
  if (h->plane) {
    HIP_TRY(launch_fft_plane_fwd(h->grid, h->cgrid, h->nx, h->ny, h->nz, axis_args(h, 0), axis_args(h, 1), h->stream));
    mark(h, 3);
  } else {
    HIP_TRY(launch_fft_x_r2c(h->grid, h->cgrid, h->nx, h->ny, h->nz, h->NR, axis_args(h, 0), h->stream));
    mark(h, 3);
    HIP_TRY(launch_fft_cols(COL_FWD, h->cgrid, h->ny, hx, h->TXy, hx, h->nz, spec_pitch(hx, h->ny), axis_args(h, 1),
                            dummy, no_reduce(), hx, h->stream));
  }
  h->state = ST_XY;
  return PME_OK;
}

static pme_status z_pass(pme_recip_t h, int mode, const SolveArgs& sa) {
  const int hx = h->nx / 2 + 1;
  HIP_TRY(launch_fft_cols(mode, h->cgrid, h->nz, hx * h->ny, h->TXz, spec_pitch(hx, h->ny), 1, 0, axis_args(h, 2), sa,
                          no_reduce(), hx, h->stream));
  return PME_OK;
}

extern "C" pme_status pme_recip_scalar_sum(pme_recip_t h, const double recip[9], double kappa, int calc_energy,
                                           int calc_virial) {
  if (!h || !recip) return fail(PME_ERR_INVALID_ARG, "pme_recip_scalar_sum: null argument");
  if (!box_ok(recip)) return fail(PME_ERR_INVALID_ARG, "pme_recip_scalar_sum: recip has no positive determinant");
  if (!(kappa > 0.0)) return fail(PME_ERR_INVALID_ARG, "pme_recip_scalar_sum: kappa %g", kappa);
  if (h->state != ST_XY && h->state != ST_SPECTRUM && h->state != ST_SOLVED_ZINV)
    return fail(PME_ERR_STATE, "pme_recip_scalar_sum: no spectrum (call r2c_fft first)");
  // the partial buffer is about to be rewritten
  if (h->pend_solve) PME_TRY(flush_pending(h, h->stream, false));
  const bool ev = calc_energy || calc_virial;
  SolveArgs sa = solve_args(h, recip, kappa, ev);
  const int hx = h->nx / 2 + 1;
  mark(h, 4);
  if (h->state == ST_XY && !sa.tric) {
    PME_TRY(z_pass(h, COL_FWD_SOLVE_INV, sa));
    if (ev) h->pend_solve = cols_blocks(hx * h->ny, h->TXz, 1);
    h->state = ST_SOLVED_ZINV;
  } else {
    // the spectrum in full first (a triclinic box: the fused z pass holds
    // the orthorhombic influence function only)
    if (h->state == ST_SOLVED_ZINV || h->state == ST_XY) PME_TRY(z_pass(h, COL_FWD, sa));
    HIP_TRY(launch_scalar_sum(h->cgrid, hx, sa, h->stream));
    if (ev) h->pend_solve = scalar_sum_blocks(hx * h->ny * h->nz);
    h->state = ST_SPECTRUM;
  }
  return PME_OK;
}

extern "C" pme_status pme_recip_c2r_fft(pme_recip_t h) {
  if (!h) return fail(PME_ERR_INVALID_ARG, "pme_recip_c2r_fft: null handle");
  const int hx = h->nx / 2 + 1;
  SolveArgs dummy{};
  if (h->state == ST_XY) {
    PME_TRY(z_pass(h, COL_FWD_INV, dummy));
  } else if (h->state == ST_SPECTRUM) {
    PME_TRY(z_pass(h, COL_INV, dummy));
  } else if (h->state != ST_SOLVED_ZINV) {
    return fail(PME_ERR_STATE, "pme_recip_c2r_fft: no spectrum (call r2c_fft first)");
  }
  mark(h, 5);
  // the y-inverse (or xy plane inverse) pass also folds in the pending
  // energy/virial reduction
  if (h->plane) {
    HIP_TRY(launch_fft_plane_inv(h->cgrid, h->grid, h->nx, h->ny, h->nz, axis_args(h, 0), axis_args(h, 1),
                                 pending_reduce(h), h->stream));
    h->pend_solve = h->pend_self = 0;
    mark(h, 6);
  } else {
    HIP_TRY(launch_fft_cols(COL_INV, h->cgrid, h->ny, hx, h->TXy, hx, h->nz, spec_pitch(hx, h->ny), axis_args(h, 1),
                            dummy, pending_reduce(h), hx, h->stream));
    h->pend_solve = h->pend_self = 0;
    mark(h, 6);
    HIP_TRY(launch_fft_x_c2r(h->cgrid, h->grid, h->nx, h->ny, h->nz, h->NR, axis_args(h, 0), h->stream));
  }
  h->state = ST_POTENTIAL;
  return PME_OK;
}

extern "C" pme_status pme_recip_gather_force(pme_recip_t h, const void* d_xyzq, int ncoord, const double recip[9],
                                             int stride, int force_type, void* d_force) {
  if (!h || !recip || (ncoord > 0 && (!d_xyzq || !d_force)))
    return fail(PME_ERR_INVALID_ARG, "pme_recip_gather_force: null argument");
  if (ncoord < 0) return fail(PME_ERR_INVALID_ARG, "pme_recip_gather_force: ncoord %d", ncoord);
  if (force_type < PME_FORCE_FLOAT || force_type > PME_FORCE_INT64)
    return fail(PME_ERR_INVALID_ARG, "pme_recip_gather_force: force_type %d", force_type);
  if (force_type != PME_FORCE_FLOAT3 && stride < ncoord)
    return fail(PME_ERR_INVALID_ARG, "pme_recip_gather_force: stride %d < ncoord %d", stride, ncoord);
  if (!box_ok(recip)) return fail(PME_ERR_INVALID_ARG, "pme_recip_gather_force: recip has no positive determinant");
  if (h->state != ST_POTENTIAL) return fail(PME_ERR_STATE, "pme_recip_gather_force: no potential (call c2r_fft first)");
  if (ncoord == 0) return PME_OK;
  bool same_box = true;
  for (int i = 0; i < 9; i++) same_box = same_box && (float)recip[i] == h->bin_key_r[i];
  // the records of the last binning say which brick gathers which atom; the
  // gather itself reads the current coordinates, so stale records only cost
  // a detour through HBM.  Different atoms or box: bin again.
  if (d_xyzq != h->bin_key_ptr || ncoord != h->bin_key_n || !same_box) {
    PME_TRY(do_bin(h, d_xyzq, ncoord, recip));
  }
  mark(h, 7);
  BinArgs a = bin_args(h, recip);
  HIP_TRY(launch_gather(h->order, force_type, (const float4*)d_xyzq, a, h->wmax, h->grid, stride, d_force, h->stream));
  mark(h, 8);
  return PME_OK;
}

extern "C" pme_status pme_recip_calc_self_energy(pme_recip_t h, const void* d_xyzq, int ncoord, double kappa) {
  if (!h || (ncoord > 0 && !d_xyzq)) return fail(PME_ERR_INVALID_ARG, "pme_recip_calc_self_energy: null argument");
  if (ncoord < 0) return fail(PME_ERR_INVALID_ARG, "pme_recip_calc_self_energy: ncoord %d", ncoord);
  if (ncoord == 0) return PME_OK;
  if (h->pend_self) PME_TRY(flush_pending(h, h->stream, false));
  int nblk = 0;
  HIP_TRY(launch_self_energy((const float4*)d_xyzq, ncoord, h->self_partial, &nblk, h->stream));
  h->pend_self = nblk;
  h->pend_kcs = kappa * kCcelec / std::sqrt(kPi);
  return PME_OK;
}

static pme_status solve_stages(pme_recip_t h, const void* d_xyzq, int ncoord, const double recip[9], double kappa,
                               int calc_energy, int calc_virial, int stride, int force_type, void* d_force) {
  PME_TRY(pme_recip_spread_charge(h, d_xyzq, ncoord, recip));
  PME_TRY(pme_recip_r2c_fft(h));
  PME_TRY(pme_recip_scalar_sum(h, recip, kappa, calc_energy, calc_virial));
  PME_TRY(pme_recip_c2r_fft(h));
  PME_TRY(pme_recip_gather_force(h, d_xyzq, ncoord, recip, stride, force_type, d_force));
  return PME_OK;
}

static SolveKey solve_key(pme_recip_t h, const void* d_xyzq, int ncoord, const double recip[9], double kappa,
                          int calc_energy, int calc_virial, int stride, int force_type, void* d_force, int pend_self,
                          double pend_kcs) {
  SolveKey k;
  std::memset(&k, 0, sizeof k);  // padding too: keys compare bytewise
  k.xyzq = d_xyzq;
  k.force = d_force;
  k.stream = h->stream;
  const void* b[6] = {h->grid, h->cgrid, h->seg_off, h->rec_pos, h->rec_idx, h->partial};
  for (int i = 0; i < 6; i++) k.buf[i] = b[i];
  for (int i = 0; i < 9; i++) k.recip[i] = recip[i];
  k.kappa = kappa;
  k.pend_kcs = pend_kcs;
  k.n = ncoord;
  k.calc_energy = calc_energy != 0;
  k.calc_virial = calc_virial != 0;
  k.stride = stride;
  k.force_type = force_type;
  k.order = h->order;
  k.prec = h->prec;
  k.pend_self = pend_self;
  k.nbins = h->nbins;
  k.chunk = h->chunk;
  k.nchunk = h->nchunk;
  k.nchunk_cap = h->nchunk_cap;
  k.bins_for = h->bins_for;
  k.partial_cap = h->partial_cap;
  for (int a = 0; a < 3; a++) {
    k.nb[a] = h->nb[a];
    k.wmax[a] = h->wmax[a];
  }
  return k;
}

// The one-shot solve as a hipGraph on the handle's stream: the first call
// with a new key runs the stages eagerly (allocating what they need), the
// second captures them into a graph and launches it, later ones replay it.
// A replay leaves the handle in the state the stages leave it in.  The null
// stream cannot be captured, nor are timed solves; those run eagerly.
static pme_status solve_graph(pme_recip_t h, const void* d_xyzq, int ncoord, const double recip[9], double kappa,
                              int calc_energy, int calc_virial, int stride, int force_type, void* d_force) {
  const int pend_self = h->pend_self;
  const double pend_kcs = h->pend_kcs;
  if (h->gkey_valid && h->gexec) {
    const SolveKey k = solve_key(h, d_xyzq, ncoord, recip, kappa, calc_energy, calc_virial, stride, force_type,
                                 d_force, pend_self, pend_kcs);
    if (std::memcmp(&k, &h->gkey, sizeof k) == 0) {
      HIP_TRY(hipGraphLaunch(h->gexec, h->stream));
      h->state = ST_POTENTIAL;
      h->pend_solve = h->pend_self = 0;
      h->bin_key_ptr = d_xyzq;
      h->bin_key_n = ncoord;
      for (int i = 0; i < 9; i++) h->bin_key_r[i] = (float)recip[i];
      return PME_OK;
    }
  }
  if (h->gexec) {
    (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
  }
  const bool seen = h->gkey_valid;
  const SolveKey before = solve_key(h, d_xyzq, ncoord, recip, kappa, calc_energy, calc_virial, stride, force_type,
                                    d_force, pend_self, pend_kcs);
  if (seen && std::memcmp(&before, &h->gkey, sizeof before) == 0) {
    // the same solve as last time, with nothing left to allocate: capture it
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const pme_status st =
        solve_stages(h, d_xyzq, ncoord, recip, kappa, calc_energy, calc_virial, stride, force_type, d_force);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (st != PME_OK || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      h->gkey_valid = false;
      if (st != PME_OK) return st;
      HIP_TRY(e);
    }
    const hipError_t ei = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      h->gexec = nullptr;
      h->gkey_valid = false;
      HIP_TRY(ei);
    }
    HIP_TRY(hipGraphLaunch(h->gexec, h->stream));
    return PME_OK;
  }
  PME_TRY(solve_stages(h, d_xyzq, ncoord, recip, kappa, calc_energy, calc_virial, stride, force_type, d_force));
  // the key as the next identical call will see it (buffers as allocated now)
  h->gkey = solve_key(h, d_xyzq, ncoord, recip, kappa, calc_energy, calc_virial, stride, force_type, d_force,
                      pend_self, pend_kcs);
  h->gkey_valid = true;
  return PME_OK;
}

extern "C" pme_status pme_recip_set_graph(pme_recip_t h, int on) {
  if (!h) return fail(PME_ERR_INVALID_ARG, "pme_recip_set_graph: null handle");
  h->use_graph = on != 0;
  if (!h->use_graph && h->gexec) {
    (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
  }
  h->gkey_valid = false;
  return PME_OK;
}

extern "C" pme_status pme_recip_solve(pme_recip_t h, const void* d_xyzq, int ncoord, const double recip[9],
                                      double kappa, int calc_energy, int calc_virial, int stride, int force_type,
                                      void* d_force) {
  if (!h || !recip) return fail(PME_ERR_INVALID_ARG, "pme_recip_solve: null argument");
  if (!box_ok(recip)) return fail(PME_ERR_INVALID_ARG, "pme_recip_solve: recip has no positive determinant");
  // a scalar-sum reduction still pending would be flushed inside the sequence
  if (!h->use_graph || h->timing || h->stream == nullptr || h->pend_solve != 0)
    return solve_stages(h, d_xyzq, ncoord, recip, kappa, calc_energy, calc_virial, stride, force_type, d_force);
  return solve_graph(h, d_xyzq, ncoord, recip, kappa, calc_energy, calc_virial, stride, force_type, d_force);
}

// ------------------------------------------------- allocation-free solve --
// pme_recip_solve with every device work buffer of the solve (real grid, half
// spectrum, binning record store and segment table, energy/virial partials)
// carved out of one file-scope device array instead of the handle's buffers,
// for callers that cannot allocate (graph capture, serving loops).  Each GPU
// holds its own copy of the array until the process ends; its device address
// comes from hipGetSymbolAddress once per device (the array's name is only a
// host-side stand-in).  Every call on a GPU shares that GPU's copy: calls that
// overlap on one GPU (other streams, other threads) race, calls ordered on one
// stream do not.  Nothing in it needs clearing between calls: every kernel of
// the solve writes what it later reads.
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kScratchDevices] = {};

static pme_status scratch_address(void** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices)
    return fail(PME_ERR_UNSUPPORTED, "pme_recip_solve_noalloc: device %d (at most %d devices)", dev, kScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(PME_ERR_DEVICE, "pme_recip_solve_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return PME_OK;
}

// Offsets of the work buffers of an n-atom solve in the scratch array
// (256-byte aligned); returns the bytes used.
struct ScratchLayout {
  size_t grid, cgrid, seg_off, rec_pos, rec_idx, partial, total;
  int chunk, nchunk;
};
static ScratchLayout scratch_layout(pme_recip_t h, int n) {
  ScratchLayout L{};
  const ChunkPlan cp = plan_chunks(n);
  L.chunk = cp.chunk;
  L.nchunk = cp.nchunk;
  const size_t G = (size_t)h->nx * h->ny * h->nz;
  const size_t rs = 8 * (size_t)cp.chunk * cp.nchunk;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += (bytes + 255) & ~(size_t)255;
    return at;
  };
  L.grid = take(sizeof(float) * G);
  L.cgrid = take(sizeof(float2) * (size_t)spec_pitch(h->nx / 2 + 1, h->ny) * h->nz);
  L.seg_off = take(sizeof(int) * (size_t)cp.nchunk * seg_stride(h->nbins));
  L.rec_pos = take(sizeof(float4) * rs);
  L.rec_idx = take(sizeof(int) * rs);
  L.partial = take(sizeof(double) * 7 * (size_t)h->partial_cap);
  L.total = off;
  return L;
}

extern "C" pme_status pme_recip_solve_noalloc(pme_recip_t h, const void* d_xyzq, int ncoord, const double recip[9],
                                              double kappa, int calc_energy, int calc_virial, int stride,
                                              int force_type, void* d_force) {
  if (!h || !recip) return fail(PME_ERR_INVALID_ARG, "pme_recip_solve_noalloc: null argument");
  if (ncoord < 0) return fail(PME_ERR_INVALID_ARG, "pme_recip_solve_noalloc: ncoord %d", ncoord);
  if (!box_ok(recip))
    return fail(PME_ERR_INVALID_ARG, "pme_recip_solve_noalloc: recip has no positive determinant");
  const ScratchLayout L = scratch_layout(h, ncoord);
  if (L.total > kScratchBytes)
    return fail(PME_ERR_BUDGET,
                "pme_recip_solve_noalloc: %d atoms on a %dx%dx%d grid need %zu bytes of work buffers, "
                "more than the %zu-byte static scratch",
                ncoord, h->nx, h->ny, h->nz, L.total, kScratchBytes);
  void* vbase = nullptr;
  PME_TRY(scratch_address(&vbase));
  char* base = static_cast<char*>(vbase);
  // reductions still pending read the handle's own partials: fold them in first
  PME_TRY(flush_pending(h, h->stream, false));
  float* grid = h->grid;
  float2* cgrid = h->cgrid;
  int* seg_off = h->seg_off;
  float4* rec_pos = h->rec_pos;
  int* rec_idx = h->rec_idx;
  double* partial = h->partial;
  const int bins_for = h->bins_for, chunk = h->chunk, nchunk = h->nchunk, nchunk_cap = h->nchunk_cap;
  h->grid = reinterpret_cast<float*>(base + L.grid);
  h->cgrid = reinterpret_cast<float2*>(base + L.cgrid);
  h->seg_off = reinterpret_cast<int*>(base + L.seg_off);
  h->rec_pos = reinterpret_cast<float4*>(base + L.rec_pos);
  h->rec_idx = reinterpret_cast<int*>(base + L.rec_idx);
  h->partial = reinterpret_cast<double*>(base + L.partial);
  // the record store is sized for exactly this call, so the binning never reallocates
  h->chunk = L.chunk;
  h->nchunk = h->nchunk_cap = L.nchunk;
  h->bins_for = std::max(ncoord, 1);
  h->bin_key_n = -1;
  // the stages run eagerly: instantiating a graph may allocate
  pme_status st = solve_stages(h, d_xyzq, ncoord, recip, kappa, calc_energy, calc_virial, stride, force_type,
                               d_force);
  // a reduction still pending reads the scratch partials: launch it now
  if (st == PME_OK) st = flush_pending(h, h->stream, false);
  h->grid = grid;
  h->cgrid = cgrid;
  h->seg_off = seg_off;
  h->rec_pos = rec_pos;
  h->rec_idx = rec_idx;
  h->partial = partial;
  h->bins_for = bins_for;
  h->chunk = chunk;
  h->nchunk = nchunk;
  h->nchunk_cap = nchunk_cap;
  // the potential and the records are in the scratch, not in the handle
  h->state = ST_EMPTY;
  h->bin_key_n = -1;
  h->pend_solve = 0;
  return st;
}

extern "C" pme_status pme_recip_copy_grid(pme_recip_t h, int which, void* d_dst, size_t bytes) {
  if (!h || !d_dst) return fail(PME_ERR_INVALID_ARG, "pme_recip_copy_grid: null argument");
  const size_t G = (size_t)h->nx * h->ny * h->nz;
  const size_t Gc = (size_t)(h->nx / 2 + 1) * h->ny * h->nz;
  if (which == PME_GRID_REAL) {
    if (h->state != ST_CHARGE && h->state != ST_POTENTIAL)
      return fail(PME_ERR_STATE, "pme_recip_copy_grid: no real grid");
    if (bytes < G * sizeof(float)) return fail(PME_ERR_INVALID_ARG, "pme_recip_copy_grid: buffer too small");
    HIP_TRY(hipMemcpyAsync(d_dst, h->grid, G * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
  } else if (which == PME_GRID_SPECTRUM) {
    if (h->state == ST_XY || h->state == ST_SOLVED_ZINV) {
      SolveArgs dummy{};
      PME_TRY(z_pass(h, COL_FWD, dummy));
      h->state = ST_SPECTRUM;
    }
    if (h->state != ST_SPECTRUM) return fail(PME_ERR_STATE, "pme_recip_copy_grid: no spectrum");
    if (bytes < Gc * sizeof(float2)) return fail(PME_ERR_INVALID_ARG, "pme_recip_copy_grid: buffer too small");
    // the reference's BOX layout, without the plane padding of the pitch
    const size_t row = sizeof(float2) * (size_t)(h->nx / 2 + 1) * h->ny;
    HIP_TRY(hipMemcpy2DAsync(d_dst, row, h->cgrid, sizeof(float2) * (size_t)spec_pitch(h->nx / 2 + 1, h->ny), row,
                             h->nz, hipMemcpyDeviceToDevice, h->stream));
  } else {
    return fail(PME_ERR_INVALID_ARG, "pme_recip_copy_grid: which %d", which);
  }
  return PME_OK;
}

extern "C" pme_status pme_recip_set_timing(pme_recip_t h, int enable) {
  if (!h) return fail(PME_ERR_INVALID_ARG, "pme_recip_set_timing: null handle");
  h->timing = enable != 0;
  return PME_OK;
}

extern "C" pme_status pme_recip_get_timing(pme_recip_t h, float ms[8]) {
  if (!h || !ms) return fail(PME_ERR_INVALID_ARG, "pme_recip_get_timing: null argument");
  if (!h->timing) return fail(PME_ERR_STATE, "pme_recip_get_timing: timing not enabled");
  HIP_TRY(hipEventSynchronize(h->tev[8]));
  for (int i = 0; i < 8; i++) HIP_TRY(hipEventElapsedTime(&ms[i], h->tev[i], h->tev[i + 1]));
  return PME_OK;
}

extern "C" pme_status pme_fill_bspline(int order, int nfftx, int nffty, int nfftz, const double recip[9],
                                       const void* d_xyzq, int ncoord, int* d_gix, int* d_giy, int* d_giz,
                                       float* d_charge, float* d_thetax, float* d_thetay, float* d_thetaz,
                                       float* d_dthetax, float* d_dthetay, float* d_dthetaz, pme_stream_t stream) {
  if (!recip || ncoord < 0) return fail(PME_ERR_INVALID_ARG, "pme_fill_bspline: bad arguments");
  if (order != 4 && order != 6 && order != 8) return fail(PME_ERR_UNSUPPORTED, "pme_fill_bspline: order %d", order);
  if (nfftx < 1 || nffty < 1 || nfftz < 1) return fail(PME_ERR_INVALID_ARG, "pme_fill_bspline: grid");
  if (ncoord == 0) return PME_OK;
  if (!d_xyzq || !d_gix || !d_giy || !d_giz || !d_charge || !d_thetax || !d_thetay || !d_thetaz || !d_dthetax ||
      !d_dthetay || !d_dthetaz)
    return fail(PME_ERR_INVALID_ARG, "pme_fill_bspline: null device pointer");
  HIP_TRY(launch_fill_bspline(order, (const float4*)d_xyzq, ncoord, (float)recip[0], (float)recip[4], (float)recip[8],
                              nfftx, nffty, nfftz, d_gix, d_giy, d_giz, d_charge, d_thetax, d_thetay, d_thetaz,
                              d_dthetax, d_dthetay, d_dthetaz, (hipStream_t)stream));
  return PME_OK;
}
