// pme_batch.cpp -- batched PME solves of many independent systems.
//
// Each admitted system owns a slot: its grids, binning store, FFT tiles and
// energy/virial partials stay resident in HBM, admission is checked against
// the batch's HBM budget.  A solve of a subset of slots runs the same stage
// bodies as the single-system solver, but every stage is ONE launch over all
// systems (persistent workgroups walking a per-stage work list), and the
// whole 8-launch sequence is captured once in a hipGraph and replayed while
// the subset and the launch shapes stay the same.  Coordinates, box and
// outputs are read from the device descriptor table at run time, so updating
// them does not invalidate the graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pme_recip.h"
#include "pme_internal.hpp"
#include "pme_plan.hpp"

using namespace pme;

#define BHIP(expr)                                                                                  \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess)                                                                           \
      return set_error(e_ == hipErrorOutOfMemory ? PME_ERR_OOM : PME_ERR_DEVICE, "%s: %s", #expr, \
                       hipGetErrorString(e_));                                                      \
  } while (0)

#define BTRY(expr)                \
  do {                            \
    pme_status s_ = (expr);       \
    if (s_ != PME_OK) return s_;  \
  } while (0)

namespace {

const double kPiB = 3.14159265358979323846;

struct Slot {
  bool used = false;
  pme_system_desc in{};
  int cap_atoms = 0;
  BrickPlan bp{};
  ChunkPlan cp{};
  FftTiles ft{};
  RadixPlan rp[3]{};
  SysDesc d{};
  size_t bytes = 0;
  float* grid = nullptr;
  float2* cgrid = nullptr;
  int* seg_off = nullptr;
  float4* rec_pos = nullptr;
  int* rec_idx = nullptr;
  double* partial = nullptr;
};

struct Launch {
  int stage;     // BatchStage
  int order;     // spread/gather template order (0 otherwise)
  int offset;    // first item in the concatenated list
  int count;     // items
  size_t lds;    // dynamic LDS bytes (max over the launch's systems)
};

}  // namespace

struct pme_batch_s {
  hipStream_t stream = nullptr;
  // With no stream given, the batch runs on a stream of its own (captures
  // are not allowed on the null stream) and orders itself after / before
  // the null stream's work with the two events.
  bool own_stream = false;
  hipEvent_t ev_in = nullptr, ev_done = nullptr;
  size_t budget = 0, used = 0;
  int brick_target = kDefaultBrick, fft_tile = kDefaultFftTile, wg_per_cu = 0;
  bool use_graph = true;
  int ncu = 256;
  std::string arch;
  std::vector<Slot> slots;
  SysDesc* d_sys = nullptr;  // [slots.capacity]
  int d_sys_cap = 0;
  double* d_ev = nullptr;    // [d_sys_cap * 10]
  double* h_ev = nullptr;    // pinned mirror
  std::map<int, float2*> tw_cache;
  std::map<std::pair<int, int>, float*> pf_cache;
  // work lists of the current subset
  std::vector<int> cur_ids;
  bool lists_valid = false;
  std::vector<Launch> launches;
  WorkItem* d_items = nullptr;
  size_t items_cap = 0;
  // graph of the current subset
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool graph_valid = false;
  // per-stage timing (eager launches bracketed by events; graph bypassed)
  bool timing = false;
  hipEvent_t tev[9] = {};
  float stage_ms[8] = {};
};

static void drop_graph(pme_batch_t b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_valid = false;
}

static void invalidate(pme_batch_t b) {
  b->lists_valid = false;
  drop_graph(b);
}

static void free_slot_state(pme_batch_t b, Slot& s) {
  (void)hipFree(s.grid);
  (void)hipFree(s.cgrid);
  (void)hipFree(s.seg_off);
  (void)hipFree(s.rec_pos);
  (void)hipFree(s.rec_idx);
  (void)hipFree(s.partial);
  s.grid = nullptr;
  s.cgrid = nullptr;
  s.seg_off = nullptr;
  s.rec_pos = nullptr;
  s.rec_idx = nullptr;
  s.partial = nullptr;
  b->used -= s.bytes;
  s.bytes = 0;
}

static pme_status shared_tables(pme_batch_t b, int order, int n, const float2** tw, const float** pf) {
  auto it = b->tw_cache.find(n);
  if (it == b->tw_cache.end()) {
    float2* d = nullptr;
    BHIP(hipMalloc(&d, sizeof(float2) * n));
    const std::vector<float2> t = twiddles(n);
    BHIP(hipMemcpy(d, t.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
    b->used += sizeof(float2) * n;
    it = b->tw_cache.emplace(n, d).first;
  }
  *tw = it->second;
  auto jt = b->pf_cache.find({order, n});
  if (jt == b->pf_cache.end()) {
    std::vector<double> m(n);
    bsp_mod_host(order, n, m.data());
    std::vector<float> f(n);
    for (int i = 0; i < n; i++) f[i] = (float)m[i];
    float* d = nullptr;
    BHIP(hipMalloc(&d, sizeof(float) * n));
    BHIP(hipMemcpy(d, f.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    b->used += sizeof(float) * n;
    jt = b->pf_cache.emplace(std::make_pair(order, n), d).first;
  }
  *pf = jt->second;
  return PME_OK;
}

static pme_status validate(const pme_system_desc* s) {
  if (!s) return set_error(PME_ERR_INVALID_ARG, "system: null descriptor");
  if (s->order != 4 && s->order != 6 && s->order != 8) return set_error(PME_ERR_UNSUPPORTED, "system: order %d", s->order);
  const int n[3] = {s->nfftx, s->nffty, s->nfftz};
  for (int a = 0; a < 3; a++)
    if (n[a] < s->order || n[a] > 512 || !smooth235(n[a]))
      return set_error(PME_ERR_UNSUPPORTED, "system: grid edge %d", n[a]);
  const double* r = s->recip;
  if (r[1] != 0.0 || r[2] != 0.0 || r[3] != 0.0 || r[5] != 0.0 || r[6] != 0.0 || r[7] != 0.0)
    return set_error(PME_ERR_UNSUPPORTED, "system: only orthorhombic boxes are supported");
  if (!(r[0] > 0.0 && r[4] > 0.0 && r[8] > 0.0)) return set_error(PME_ERR_INVALID_ARG, "system: recip diagonal");
  if (!(s->kappa > 0.0)) return set_error(PME_ERR_INVALID_ARG, "system: kappa %g", s->kappa);
  if (s->ncoord < 0) return set_error(PME_ERR_INVALID_ARG, "system: ncoord %d", s->ncoord);
  if (s->ncoord > 0 && (!s->d_xyzq || !s->d_force)) return set_error(PME_ERR_INVALID_ARG, "system: null device pointer");
  if (s->force_type < PME_FORCE_FLOAT || s->force_type > PME_FORCE_INT64)
    return set_error(PME_ERR_INVALID_ARG, "system: force_type %d", s->force_type);
  if (s->force_type != PME_FORCE_FLOAT3 && s->stride < s->ncoord)
    return set_error(PME_ERR_INVALID_ARG, "system: stride %d < ncoord %d", s->stride, s->ncoord);
  return PME_OK;
}

// Values of the descriptor that change with coordinates, box and outputs.
static void fill_dynamic(Slot& s) {
  const pme_system_desc& in = s.in;
  SysDesc& d = s.d;
  d.bin.r1 = (float)in.recip[0];
  d.bin.r2 = (float)in.recip[4];
  d.bin.r3 = (float)in.recip[8];
  d.bin.nchunk = std::max(1, (in.ncoord + s.cp.chunk - 1) / s.cp.chunk);
  d.xyzq = (const float4*)in.d_xyzq;
  d.natom = in.ncoord;
  d.force = in.d_force;
  d.force_type = in.force_type;
  d.stride = in.stride;
  d.sa.r1 = d.bin.r1;
  d.sa.r2 = d.bin.r2;
  d.sa.r3 = d.bin.r3;
  const double inv_vol = in.recip[0] * in.recip[4] * in.recip[8];
  d.sa.piv_inv = (float)(inv_vol / kPiB);
  d.sa.fac = (float)(kPiB * kPiB / (in.kappa * in.kappa));
  d.sa.calc_ev = in.calc_energy_virial ? 1 : 0;
}

// Plan one system under the batch's tuning; returns the bytes its resident
// state needs.  Touches no device memory.
static pme_status plan_only(pme_batch_t b, Slot& s, size_t* bytes) {
  const pme_system_desc& in = s.in;
  const int n[3] = {in.nfftx, in.nffty, in.nfftz};
  if (!plan_bricks(n, in.order, b->brick_target, &s.bp))
    return set_error(PME_ERR_UNSUPPORTED, "system: no brick plan fits the LDS budget");
  s.cp = plan_chunks(s.cap_atoms);
  const bool even = (n[0] % 2) == 0;
  if (!radix_plan(even ? n[0] / 2 : n[0], even ? 2 : 1, &s.rp[0]) || !radix_plan(n[1], 1, &s.rp[1]) ||
      !radix_plan(n[2], 1, &s.rp[2]))
    return set_error(PME_ERR_UNSUPPORTED, "system: no radix plan");
  s.ft = pick_fft_tiles(n[0], n[1], n[2], s.rp[0].n, b->fft_tile);
  const int hx = n[0] / 2 + 1;
  const int nzblk = cols_blocks(hx, s.ft.TXz, n[1]);
  const size_t G = (size_t)n[0] * n[1] * n[2];
  const size_t Gc = (size_t)hx * n[1] * n[2];
  const size_t rs = 8 * (size_t)s.cp.chunk;
  *bytes = sizeof(float) * G + sizeof(float2) * Gc + sizeof(int) * (size_t)s.cp.nchunk * (s.bp.nbins + 1) +
           (sizeof(float4) + sizeof(int)) * rs * s.cp.nchunk + sizeof(double) * 7 * (size_t)nzblk;
  return PME_OK;
}

// Allocate the planned state of slot `slot_index` and fill its descriptor.
static pme_status alloc_slot(pme_batch_t b, Slot& s, int slot_index, size_t bytes) {
  const pme_system_desc& in = s.in;
  const int n[3] = {in.nfftx, in.nffty, in.nfftz};
  const int hx = n[0] / 2 + 1;
  const int nzblk = cols_blocks(hx, s.ft.TXz, n[1]);
  const size_t G = (size_t)n[0] * n[1] * n[2];
  const size_t Gc = (size_t)hx * n[1] * n[2];
  const size_t rs = 8 * (size_t)s.cp.chunk;
  s.bytes = bytes;
  b->used += bytes;
  BHIP(hipMalloc(&s.grid, sizeof(float) * G));
  BHIP(hipMalloc(&s.cgrid, sizeof(float2) * Gc));
  BHIP(hipMalloc(&s.seg_off, sizeof(int) * (size_t)s.cp.nchunk * (s.bp.nbins + 1)));
  BHIP(hipMalloc(&s.rec_pos, sizeof(float4) * rs * s.cp.nchunk));
  BHIP(hipMalloc(&s.rec_idx, sizeof(int) * rs * s.cp.nchunk));
  BHIP(hipMalloc(&s.partial, sizeof(double) * 7 * (size_t)nzblk));

  SysDesc& d = s.d;
  std::memset(&d, 0, sizeof(d));
  d.bin.nx = n[0];
  d.bin.ny = n[1];
  d.bin.nz = n[2];
  d.bin.nbx = s.bp.nb[0];
  d.bin.nby = s.bp.nb[1];
  d.bin.nbz = s.bp.nb[2];
  d.bin.order = in.order;
  d.bin.chunk = s.cp.chunk;
  d.bin.rec_stride = (int)rs;
  d.bin.seg_off = s.seg_off;
  d.bin.rec_pos = s.rec_pos;
  d.bin.rec_idx = s.rec_idx;
  for (int a = 0; a < 3; a++) d.wmax[a] = s.bp.wmax[a];
  d.grid = s.grid;
  d.cgrid = s.cgrid;
  for (int a = 0; a < 3; a++) {
    const float2* tw = nullptr;
    const float* pf = nullptr;
    BTRY(shared_tables(b, in.order, n[a], &tw, &pf));
    d.ax[a].rp = s.rp[a];
    d.ax[a].tw = tw;
    d.ax[a].ntw = n[a];
    if (a == 0) d.sa.pf1 = pf;
    if (a == 1) d.sa.pf2 = pf;
    if (a == 2) d.sa.pf3 = pf;
  }
  d.NR = s.ft.NR;
  d.TXy = s.ft.TXy;
  d.TXz = s.ft.TXz;
  d.sa.n1 = n[0];
  d.sa.n2 = n[1];
  d.sa.n3 = n[2];
  d.sa.nf1 = n[0] / 2 + (n[0] % 2);
  d.sa.nf2 = n[1] / 2 + (n[1] % 2);
  d.sa.nf3 = n[2] / 2 + (n[2] % 2);
  d.sa.partial = s.partial;
  d.nzblk = nzblk;
  d.ev_out = b->d_ev + 10 * (size_t)slot_index;
  fill_dynamic(s);
  return PME_OK;
}

// Plan, check the budget (nothing allocated when it does not fit), allocate.
static pme_status plan_slot(pme_batch_t b, Slot& s, int slot_index) {
  size_t bytes = 0;
  BTRY(plan_only(b, s, &bytes));
  if (b->used + bytes > b->budget)
    return set_error(PME_ERR_BUDGET, "system needs %zu bytes, %zu of %zu in use", bytes, b->used, b->budget);
  return alloc_slot(b, s, slot_index, bytes);
}

static pme_status upload_desc(pme_batch_t b, int id) {
  BHIP(hipStreamSynchronize(b->stream));
  BHIP(hipMemcpy(b->d_sys + id, &b->slots[id].d, sizeof(SysDesc), hipMemcpyHostToDevice));
  return PME_OK;
}

static pme_status ensure_desc_capacity(pme_batch_t b, int need) {
  if (need <= b->d_sys_cap) return PME_OK;
  int cap = std::max(64, b->d_sys_cap * 2);
  while (cap < need) cap *= 2;
  BHIP(hipStreamSynchronize(b->stream));
  SysDesc* nd = nullptr;
  double* nev = nullptr;
  double* nh = nullptr;
  BHIP(hipMalloc(&nd, sizeof(SysDesc) * cap));
  BHIP(hipMalloc(&nev, sizeof(double) * 10 * cap));
  BHIP(hipHostMalloc(&nh, sizeof(double) * 10 * cap, hipHostMallocDefault));
  BHIP(hipMemset(nev, 0, sizeof(double) * 10 * cap));
  std::memset(nh, 0, sizeof(double) * 10 * cap);
  (void)hipFree(b->d_sys);
  (void)hipFree(b->d_ev);
  (void)hipHostFree(b->h_ev);
  b->d_sys = nd;
  b->d_ev = nev;
  b->h_ev = nh;
  b->d_sys_cap = cap;
  // every resident descriptor points into d_ev: refresh and re-upload
  for (size_t i = 0; i < b->slots.size(); i++) {
    if (!b->slots[i].used) continue;
    b->slots[i].d.ev_out = b->d_ev + 10 * i;
    BHIP(hipMemcpy(b->d_sys + i, &b->slots[i].d, sizeof(SysDesc), hipMemcpyHostToDevice));
  }
  invalidate(b);
  return PME_OK;
}

extern "C" pme_status pme_batch_create(size_t hbm_budget_bytes, pme_stream_t stream, pme_batch_t* out) {
  if (!out) return set_error(PME_ERR_INVALID_ARG, "pme_batch_create: null out");
  *out = nullptr;
  pme_batch_t b = new pme_batch_s();
  b->stream = (hipStream_t)stream;
  b->budget = hbm_budget_bytes;
  if (!b->stream) {
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming) != hipSuccess) {
      pme_batch_destroy(b);
      return set_error(PME_ERR_DEVICE, "pme_batch_create: stream/event creation failed");
    }
    b->own_stream = true;
  }
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
    b->ncu = prop.multiProcessorCount;
    b->arch = prop.gcnArchName;
  }
  pme_status st = ensure_desc_capacity(b, 64);
  if (st != PME_OK) {
    pme_batch_destroy(b);
    return st;
  }
  *out = b;
  return PME_OK;
}

extern "C" pme_status pme_batch_destroy(pme_batch_t b) {
  if (!b) return PME_OK;
  (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  for (auto& s : b->slots)
    if (s.used) free_slot_state(b, s);
  for (auto& kv : b->tw_cache) (void)hipFree(kv.second);
  for (auto& kv : b->pf_cache) (void)hipFree(kv.second);
  (void)hipFree(b->d_sys);
  (void)hipFree(b->d_ev);
  (void)hipHostFree(b->h_ev);
  (void)hipFree(b->d_items);
  for (hipEvent_t e : b->tev)
    if (e) (void)hipEventDestroy(e);
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_done) (void)hipEventDestroy(b->ev_done);
  if (b->own_stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return PME_OK;
}

extern "C" pme_status pme_batch_add(pme_batch_t b, const pme_system_desc* sys, int* id) {
  if (!b || !id) return set_error(PME_ERR_INVALID_ARG, "pme_batch_add: null argument");
  BTRY(validate(sys));
  int k = -1;
  for (size_t i = 0; i < b->slots.size(); i++)
    if (!b->slots[i].used) {
      k = (int)i;
      break;
    }
  if (k < 0) {
    k = (int)b->slots.size();
    b->slots.emplace_back();
  }
  BTRY(ensure_desc_capacity(b, k + 1));
  Slot& s = b->slots[k];
  s = Slot();
  s.in = *sys;
  s.cap_atoms = std::max(sys->ncoord, 1);
  pme_status st = plan_slot(b, s, k);
  if (st != PME_OK) {
    free_slot_state(b, s);
    s.used = false;
    return st;
  }
  s.used = true;
  BTRY(upload_desc(b, k));
  invalidate(b);
  *id = k;
  return PME_OK;
}

extern "C" pme_status pme_batch_update(pme_batch_t b, int id, const pme_system_desc* sys) {
  if (!b || id < 0 || id >= (int)b->slots.size() || !b->slots[id].used)
    return set_error(PME_ERR_INVALID_ARG, "pme_batch_update: unknown id %d", id);
  BTRY(validate(sys));
  Slot& s = b->slots[id];
  const pme_system_desc& o = s.in;
  const bool same_plan = sys->nfftx == o.nfftx && sys->nffty == o.nffty && sys->nfftz == o.nfftz &&
                         sys->order == o.order && sys->ncoord <= s.cap_atoms;
  const int old_nchunk = s.d.bin.nchunk;
  BTRY(hipStreamSynchronize(b->stream) == hipSuccess ? PME_OK : set_error(PME_ERR_DEVICE, "sync"));
  if (same_plan) {
    s.in = *sys;
    fill_dynamic(s);
    if (s.d.bin.nchunk != old_nchunk) invalidate(b);
  } else {
    free_slot_state(b, s);
    s.in = *sys;
    s.cap_atoms = std::max(sys->ncoord, 1);
    pme_status st = plan_slot(b, s, id);
    if (st != PME_OK) {
      free_slot_state(b, s);
      s.used = false;
      invalidate(b);
      return st;
    }
    invalidate(b);
  }
  return upload_desc(b, id);
}

extern "C" pme_status pme_batch_remove(pme_batch_t b, int id) {
  if (!b || id < 0 || id >= (int)b->slots.size() || !b->slots[id].used)
    return set_error(PME_ERR_INVALID_ARG, "pme_batch_remove: unknown id %d", id);
  BHIP(hipStreamSynchronize(b->stream));
  free_slot_state(b, b->slots[id]);
  b->slots[id].used = false;
  invalidate(b);
  return PME_OK;
}

// Work lists of the subset: per stage (and per order for spread/gather) one
// contiguous run of items, systems in decreasing size so the largest tiles
// are dispatched first.
static pme_status build_lists(pme_batch_t b, const std::vector<int>& ids) {
  std::vector<int> order_ids = ids;
  std::sort(order_ids.begin(), order_ids.end(), [&](int x, int y) {
    const Slot& a = b->slots[x];
    const Slot& c = b->slots[y];
    const double ca = (double)a.in.ncoord * a.in.order * a.in.order * a.in.order + (double)a.in.nfftx * a.in.nffty * a.in.nfftz;
    const double cc = (double)c.in.ncoord * c.in.order * c.in.order * c.in.order + (double)c.in.nfftx * c.in.nffty * c.in.nfftz;
    return ca > cc || (ca == cc && x < y);
  });
  std::vector<WorkItem> items;
  b->launches.clear();
  auto add_stage = [&](int stage, int order) {
    Launch L{stage, order, (int)items.size(), 0, 0};
    for (int id : order_ids) {
      const Slot& s = b->slots[id];
      const SysDesc& d = s.d;
      if ((stage == BS_SPREAD || stage == BS_GATHER) && s.in.order != order) continue;
      int cnt = 0;
      size_t lds = 0;
      const int hx = d.bin.nx / 2 + 1;
      switch (stage) {
        case BS_BIN:
          cnt = d.bin.nchunk;
          lds = bin_lds_bytes(s.bp.nbins);
          break;
        case BS_SPREAD:
          cnt = s.bp.nbins;
          lds = spread_lds_bytes(s.in.order, s.bp.wmax, d.bin.nchunk);
          break;
        case BS_GATHER:
          cnt = s.bp.nbins;
          lds = gather_lds_bytes(s.in.order, s.bp.wmax, d.bin.nchunk);
          break;
        case BS_FFTX:
          cnt = (d.bin.ny * d.bin.nz + d.NR - 1) / d.NR;
          lds = fft_x_lds_bytes(d.ax[0].rp.n, d.NR, d.bin.nx);
          break;
        case BS_FFTY:
          cnt = cols_blocks(hx, d.TXy, d.bin.nz);
          lds = fft_cols_lds_bytes(d.bin.ny, d.TXy);
          break;
        case BS_FFTZ:
          cnt = cols_blocks(hx, d.TXz, d.bin.ny);
          lds = fft_cols_lds_bytes(d.bin.nz, d.TXz);
          break;
      }
      for (int i = 0; i < cnt; i++) items.push_back(WorkItem{id, i});
      L.count += cnt;
      L.lds = std::max(L.lds, lds);
    }
    if (L.count > 0) b->launches.push_back(L);
  };
  add_stage(BS_BIN, 0);
  for (int o : {4, 6, 8}) add_stage(BS_SPREAD, o);
  add_stage(BS_FFTX, 0);
  add_stage(BS_FFTY, 0);
  add_stage(BS_FFTZ, 0);
  for (int o : {4, 6, 8}) add_stage(BS_GATHER, o);
  BHIP(hipStreamSynchronize(b->stream));
  if (items.size() > b->items_cap) {
    (void)hipFree(b->d_items);
    b->d_items = nullptr;
    b->items_cap = std::max(items.size(), 2 * b->items_cap);
    BHIP(hipMalloc(&b->d_items, sizeof(WorkItem) * b->items_cap));
  }
  if (!items.empty()) BHIP(hipMemcpy(b->d_items, items.data(), sizeof(WorkItem) * items.size(), hipMemcpyHostToDevice));
  b->cur_ids = ids;
  b->lists_valid = true;
  drop_graph(b);
  return PME_OK;
}

static int grid_for(pme_batch_t b, int count, bool big_block) {
  if (b->wg_per_cu <= 0) return count;
  const int per_cu = big_block ? 1 : b->wg_per_cu;
  return std::max(1, std::min(count, b->ncu * per_cu));
}

// Stage order: bin, spread, x forward, y forward, z forward + influence
// function + z inverse, y inverse (with the per-system energy/virial
// reduction), x inverse, gather.  With timing on, event k is recorded before
// stage k and event 8 after the last.
static pme_status enqueue(pme_batch_t b) {
  auto mark = [&](int k) -> pme_status {
    if (b->timing) BHIP(hipEventRecord(b->tev[k], b->stream));
    return PME_OK;
  };
  auto run = [&](int stage, bool inverse) -> pme_status {
    for (const Launch& L : b->launches) {
      if (L.stage != stage) continue;
      const WorkItem* it = b->d_items + L.offset;
      const bool big = stage == BS_SPREAD || stage == BS_GATHER;
      const int grid = grid_for(b, L.count, big);
      switch (stage) {
        case BS_BIN: BHIP(launch_batch_bin(b->d_sys, it, L.count, grid, L.lds, b->stream)); break;
        case BS_SPREAD: BHIP(launch_batch_spread(L.order, b->d_sys, it, L.count, grid, L.lds, b->stream)); break;
        case BS_FFTX: BHIP(launch_batch_fft_x(inverse ? 1 : 0, b->d_sys, it, L.count, grid, L.lds, b->stream)); break;
        case BS_FFTY:
          BHIP(launch_batch_cols(1, inverse ? COL_INV : COL_FWD, inverse ? 1 : 0, b->d_sys, it, L.count, grid, L.lds,
                                 b->stream));
          break;
        case BS_FFTZ:
          BHIP(launch_batch_cols(2, COL_FWD_SOLVE_INV, 0, b->d_sys, it, L.count, grid, L.lds, b->stream));
          break;
        case BS_GATHER: BHIP(launch_batch_gather(L.order, b->d_sys, it, L.count, grid, L.lds, b->stream)); break;
        default: break;
      }
    }
    return PME_OK;
  };
  const int seq[8][2] = {{BS_BIN, 0}, {BS_SPREAD, 0}, {BS_FFTX, 0}, {BS_FFTY, 0},
                         {BS_FFTZ, 0}, {BS_FFTY, 1}, {BS_FFTX, 1}, {BS_GATHER, 0}};
  for (int k = 0; k < 8; k++) {
    BTRY(mark(k));
    BTRY(run(seq[k][0], seq[k][1] != 0));
  }
  return mark(8);
}

static pme_status run_solve(pme_batch_t b);

static std::vector<int> resolve_ids(pme_batch_t b, int n, const int* ids) {
  std::vector<int> v;
  if (n <= 0 || !ids) {
    for (size_t i = 0; i < b->slots.size(); i++)
      if (b->slots[i].used) v.push_back((int)i);
  } else {
    v.assign(ids, ids + n);
  }
  return v;
}

extern "C" pme_status pme_batch_solve(pme_batch_t b, int n, const int* ids) {
  if (!b) return set_error(PME_ERR_INVALID_ARG, "pme_batch_solve: null batch");
  const std::vector<int> v = resolve_ids(b, n, ids);
  for (int id : v)
    if (id < 0 || id >= (int)b->slots.size() || !b->slots[id].used)
      return set_error(PME_ERR_INVALID_ARG, "pme_batch_solve: unknown id %d", id);
  {
    std::vector<int> sorted = v;
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
      return set_error(PME_ERR_INVALID_ARG, "pme_batch_solve: id listed twice");
  }
  if (v.empty()) return PME_OK;
  if (!b->lists_valid || v != b->cur_ids) BTRY(build_lists(b, v));
  if (b->own_stream) {  // after the null stream's pending work (inputs written there)
    BHIP(hipEventRecord(b->ev_in, nullptr));
    BHIP(hipStreamWaitEvent(b->stream, b->ev_in, 0));
  }
  BTRY(run_solve(b));
  if (b->own_stream) {  // and before the null stream's later work (outputs read there)
    BHIP(hipEventRecord(b->ev_done, b->stream));
    BHIP(hipStreamWaitEvent(nullptr, b->ev_done, 0));
  }
  return PME_OK;
}

static pme_status run_solve(pme_batch_t b) {
  if (!b->use_graph || b->timing) return enqueue(b);
  if (!b->graph_valid) {
    BHIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    pme_status st = enqueue(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (st != PME_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (e != hipSuccess) return set_error(PME_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    b->graph = g;
    BHIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_valid = true;
  }
  BHIP(hipGraphLaunch(b->exec, b->stream));
  return PME_OK;
}

extern "C" pme_status pme_batch_results(pme_batch_t b, int n, const int* ids, double* energy, double* virial) {
  if (!b) return set_error(PME_ERR_INVALID_ARG, "pme_batch_results: null batch");
  const std::vector<int> v = resolve_ids(b, n, ids);
  BHIP(hipStreamSynchronize(b->stream));
  int hi = 0;
  for (int id : v) {
    if (id < 0 || id >= (int)b->slots.size() || !b->slots[id].used)
      return set_error(PME_ERR_INVALID_ARG, "pme_batch_results: unknown id %d", id);
    hi = std::max(hi, id + 1);
  }
  if (hi > 0) BHIP(hipMemcpy(b->h_ev, b->d_ev, sizeof(double) * 10 * hi, hipMemcpyDeviceToHost));
  if (b->timing && b->lists_valid)
    for (int k = 0; k < 8; k++) BHIP(hipEventElapsedTime(&b->stage_ms[k], b->tev[k], b->tev[k + 1]));
  for (size_t k = 0; k < v.size(); k++) {
    const double* e = b->h_ev + 10 * (size_t)v[k];
    const bool on = b->slots[v[k]].in.calc_energy_virial != 0;
    if (energy) energy[k] = on ? e[0] : 0.0;
    if (virial)
      for (int c = 0; c < 9; c++) virial[9 * k + c] = on ? e[1 + c] : 0.0;
  }
  return PME_OK;
}

extern "C" pme_status pme_batch_set_graph(pme_batch_t b, int enable) {
  if (!b) return set_error(PME_ERR_INVALID_ARG, "pme_batch_set_graph: null batch");
  b->use_graph = enable != 0;
  drop_graph(b);
  return PME_OK;
}

// Re-plan every resident system under the current tuning.  The new plans
// are checked against the budget before anything is freed; on failure the
// caller restores the previous tuning.
static pme_status replan_all(pme_batch_t b) {
  size_t total = b->used;
  for (Slot& s : b->slots) {
    if (!s.used) continue;
    Slot t = s;
    size_t bytes = 0;
    BTRY(plan_only(b, t, &bytes));
    total = total - s.bytes + bytes;
  }
  if (total > b->budget) return set_error(PME_ERR_BUDGET, "re-plan needs %zu of %zu bytes", total, b->budget);
  BHIP(hipStreamSynchronize(b->stream));
  invalidate(b);
  for (size_t i = 0; i < b->slots.size(); i++) {
    Slot& s = b->slots[i];
    if (!s.used) continue;
    free_slot_state(b, s);
    size_t bytes = 0;
    pme_status st = plan_only(b, s, &bytes);
    if (st == PME_OK) st = alloc_slot(b, s, (int)i, bytes);
    if (st != PME_OK) {
      free_slot_state(b, s);
      s.used = false;
      return st;
    }
    BTRY(upload_desc(b, (int)i));
  }
  return PME_OK;
}

extern "C" pme_status pme_batch_set_timing(pme_batch_t b, int enable) {
  if (!b) return set_error(PME_ERR_INVALID_ARG, "pme_batch_set_timing: null batch");
  BHIP(hipStreamSynchronize(b->stream));
  if (enable && !b->tev[0])
    for (int k = 0; k < 9; k++) BHIP(hipEventCreate(&b->tev[k]));
  b->timing = enable != 0;
  for (float& v : b->stage_ms) v = 0.f;
  return PME_OK;
}

extern "C" pme_status pme_batch_get_timing(pme_batch_t b, float* ms) {
  if (!b || !ms) return set_error(PME_ERR_INVALID_ARG, "pme_batch_get_timing: null argument");
  if (!b->timing) return set_error(PME_ERR_STATE, "pme_batch_get_timing: timing is off");
  BHIP(hipStreamSynchronize(b->stream));
  if (b->lists_valid)
    for (int k = 0; k < 8; k++) BHIP(hipEventElapsedTime(&b->stage_ms[k], b->tev[k], b->tev[k + 1]));
  for (int k = 0; k < 8; k++) ms[k] = b->stage_ms[k];
  return PME_OK;
}

extern "C" pme_status pme_batch_set_tuning(pme_batch_t b, int brick_target, int fft_tile, int wg_per_cu) {
  if (!b) return set_error(PME_ERR_INVALID_ARG, "pme_batch_set_tuning: null batch");
  if (brick_target < 8 || brick_target > 64 || fft_tile < 1 || fft_tile > 64 || (fft_tile & (fft_tile - 1)) ||
      wg_per_cu < 0 || wg_per_cu > 64)
    return set_error(PME_ERR_INVALID_ARG, "pme_batch_set_tuning: brick %d tile %d wg/cu %d", brick_target, fft_tile,
                     wg_per_cu);
  const bool replan = brick_target != b->brick_target || fft_tile != b->fft_tile;
  const int ob = b->brick_target, ot = b->fft_tile;
  b->brick_target = brick_target;
  b->fft_tile = fft_tile;
  b->wg_per_cu = wg_per_cu;
  drop_graph(b);
  if (replan) {
    pme_status st = replan_all(b);
    if (st == PME_ERR_BUDGET || st == PME_ERR_UNSUPPORTED) {  // nothing was freed
      b->brick_target = ob;
      b->fft_tile = ot;
    }
    return st;
  }
  return PME_OK;
}

extern "C" pme_status pme_batch_get_tuning(pme_batch_t b, int* brick_target, int* fft_tile, int* wg_per_cu) {
  if (!b) return set_error(PME_ERR_INVALID_ARG, "pme_batch_get_tuning: null batch");
  if (brick_target) *brick_target = b->brick_target;
  if (fft_tile) *fft_tile = b->fft_tile;
  if (wg_per_cu) *wg_per_cu = b->wg_per_cu;
  return PME_OK;
}

extern "C" pme_status pme_batch_memory(pme_batch_t b, size_t* used_bytes, size_t* budget_bytes) {
  if (!b) return set_error(PME_ERR_INVALID_ARG, "pme_batch_memory: null batch");
  if (used_bytes) *used_bytes = b->used;
  if (budget_bytes) *budget_bytes = b->budget;
  return PME_OK;
}

// ---- autotuner ---------------------------------------------------------------

// FNV-1a over the resident set: grid, order and atom-count class of every
// system (sorted), plus the GPU architecture.
static std::string signature(pme_batch_t b) {
  std::vector<std::string> parts;
  for (const Slot& s : b->slots) {
    if (!s.used) continue;
    char buf[96];
    std::snprintf(buf, sizeof(buf), "%dx%dx%d:p%d:n%d", s.in.nfftx, s.in.nffty, s.in.nfftz, s.in.order,
                  (int)std::ceil(std::log2((double)std::max(s.in.ncoord, 1))));
    parts.push_back(buf);
  }
  std::sort(parts.begin(), parts.end());
  unsigned long long h = 1469598103934665603ull;
  std::string all = b->arch;
  for (auto& p : parts) all += "|" + p;
  for (unsigned char c : all) {
    h ^= c;
    h *= 1099511628211ull;
  }
  char key[64];
  std::snprintf(key, sizeof(key), "%s-%016llx", b->arch.empty() ? "gpu" : b->arch.substr(0, b->arch.find(':')).c_str(), h);
  return key;
}

// cache file: one JSON object {"key": [brick, tile, wg_per_cu, ms], ...}
static std::map<std::string, std::vector<double>> read_cache(const char* path) {
  std::map<std::string, std::vector<double>> m;
  if (!path) return m;
  std::ifstream f(path);
  if (!f) return m;
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string t = ss.str();
  size_t i = 0;
  while ((i = t.find('"', i)) != std::string::npos) {
    const size_t j = t.find('"', i + 1);
    if (j == std::string::npos) break;
    const std::string key = t.substr(i + 1, j - i - 1);
    const size_t lb = t.find('[', j), rb = t.find(']', j);
    if (lb == std::string::npos || rb == std::string::npos) break;
    std::vector<double> vals;
    std::stringstream vs(t.substr(lb + 1, rb - lb - 1));
    std::string tok;
    while (std::getline(vs, tok, ',')) vals.push_back(std::atof(tok.c_str()));
    if (vals.size() == 4) m[key] = vals;
    i = rb + 1;
  }
  return m;
}

static void write_cache(const char* path, const std::map<std::string, std::vector<double>>& m) {
  if (!path) return;
  std::ofstream f(path);
  f.precision(9);
  f << "{\n";
  size_t k = 0;
  for (auto& kv : m) {
    f << "  \"" << kv.first << "\": [" << (int)kv.second[0] << ", " << (int)kv.second[1] << ", " << (int)kv.second[2]
      << ", " << kv.second[3] << "]" << (++k < m.size() ? "," : "") << "\n";
  }
  f << "}\n";
}

static pme_status time_solves(pme_batch_t b, int warm, int reps, float* ms) {
  hipEvent_t e0, e1;
  BHIP(hipEventCreate(&e0));
  BHIP(hipEventCreate(&e1));
  for (int i = 0; i < warm; i++) BTRY(pme_batch_solve(b, 0, nullptr));
  BHIP(hipEventRecord(e0, b->stream));
  for (int i = 0; i < reps; i++) BTRY(pme_batch_solve(b, 0, nullptr));
  BHIP(hipEventRecord(e1, b->stream));
  BHIP(hipEventSynchronize(e1));
  float t = 0.f;
  BHIP(hipEventElapsedTime(&t, e0, e1));
  *ms = t / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return PME_OK;
}

extern "C" pme_status pme_batch_autotune(pme_batch_t b, const char* cache_path, float* best_ms) {
  if (!b) return set_error(PME_ERR_INVALID_ARG, "pme_batch_autotune: null batch");
  const std::string key = signature(b);
  auto cache = read_cache(cache_path);
  auto hit = cache.find(key);
  if (hit != cache.end()) {
    BTRY(pme_batch_set_tuning(b, (int)hit->second[0], (int)hit->second[1], (int)hit->second[2]));
    if (best_ms) *best_ms = (float)hit->second[3];
    return PME_OK;
  }
  // the outputs of INT64 systems accumulate: tuning solves would add into them
  for (const Slot& s : b->slots)
    if (s.used && s.in.force_type == PME_FORCE_INT64)
      return set_error(PME_ERR_STATE, "pme_batch_autotune: resident INT64 (accumulating) outputs");
  float best = 1e30f;
  int bb = b->brick_target, bt = b->fft_tile, bw = b->wg_per_cu;
  for (int brick : {12, 16, 20, 24})
    for (int tile : {8, 16, 32})
      for (int wpc : {0, 4}) {
        if (pme_batch_set_tuning(b, brick, tile, wpc) != PME_OK) continue;
        float ms = 0.f;
        if (time_solves(b, 2, 5, &ms) != PME_OK) continue;
        if (ms < best) {
          best = ms;
          bb = brick;
          bt = tile;
          bw = wpc;
        }
      }
  BTRY(pme_batch_set_tuning(b, bb, bt, bw));
  cache[key] = {(double)bb, (double)bt, (double)bw, (double)best};
  write_cache(cache_path, cache);
  if (best_ms) *best_ms = best;
  return PME_OK;
}
