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
#include "pme_sched.hpp"

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

struct Slot {
  bool used = false;
  pme_system_desc in{};
  int cap_atoms = 0;
  SystemPlan plan{};
  SysDesc d{};
  void* region = nullptr;  // one allocation: grid, spectrum, record store, partials
  size_t bytes = 0;
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
  Tuning tune;
  bool use_graph = true;
  int ncu = 256;
  std::string arch;
  std::vector<Slot> slots;
  SysDesc* d_sys = nullptr;  // [slots.capacity]
  int d_sys_cap = 0;
  double* d_ev = nullptr;    // [d_sys_cap * kEvOut]
  double* h_ev = nullptr;    // pinned mirror
  int ev_copied = 0;         // slots the solve's own d_ev -> h_ev copy covers
  TableCache tables;
  // work lists of the current subset: with several chains the systems are
  // dealt into parts of about equal cost whose stage sequences run on
  // `stream` and the auxiliary streams side by side (one chain's kernel tails
  // and launch gaps are filled by the others' work)
  std::vector<int> cur_ids;
  bool lists_valid = false;
  int nchains = kBatchChains;
  std::vector<Launch> launches;                        // chain 0, on `stream`
  std::vector<Launch> launches_aux[kMaxBatchChains - 1];  // chain c on aux[c - 1] (empty: unused)
  hipStream_t aux[kMaxBatchChains - 1] = {};
  hipEvent_t ev_fork = nullptr, ev_join[kMaxBatchChains - 1] = {};
  int* d_ctr = nullptr;  // work-queue counters: 2 per launch, chain c at c * 2 * kMaxLaunches
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
  b->ev_copied = 0;
  drop_graph(b);
}

static void free_slot_state(pme_batch_t b, Slot& s) {
  (void)hipFree(s.region);
  s.region = nullptr;
  b->used -= s.bytes;
  s.bytes = 0;
}

// Allocate the planned region of slot `slot_index` and fill its descriptor.
static pme_status alloc_slot(pme_batch_t b, Slot& s, int slot_index) {
  BHIP(hipMalloc(&s.region, s.plan.total));
  s.bytes = s.plan.total;
  b->used += s.bytes;
  BTRY(fill_desc(s.d, s.in, s.plan, s.region, b->tables));
  s.d.ev_out = b->d_ev + kEvOut * (size_t)slot_index;
  return PME_OK;
}

// Plan, check the budget (nothing allocated when it does not fit), allocate.
static pme_status plan_slot(pme_batch_t b, Slot& s, int slot_index) {
  BTRY(plan_system(s.in, s.cap_atoms, b->tune, &s.plan));
  if (b->used + s.plan.total > b->budget)
    return set_error(PME_ERR_BUDGET, "system needs %zu bytes, %zu of %zu in use", s.plan.total, b->used, b->budget);
  return alloc_slot(b, s, slot_index);
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
  BHIP(hipMalloc(&nev, sizeof(double) * kEvOut * cap));
  BHIP(hipHostMalloc(&nh, sizeof(double) * kEvOut * cap, hipHostMallocDefault));
  BHIP(hipMemset(nev, 0, sizeof(double) * kEvOut * cap));
  std::memset(nh, 0, sizeof(double) * kEvOut * cap);
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
    b->slots[i].d.ev_out = b->d_ev + kEvOut * i;
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
  bool ok = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) == hipSuccess &&
            hipMalloc(&b->d_ctr, sizeof(int) * 2 * kMaxLaunches * kMaxBatchChains) == hipSuccess &&
            hipMemset(b->d_ctr, 0, sizeof(int) * 2 * kMaxLaunches * kMaxBatchChains) == hipSuccess;
  for (int c = 0; ok && c < kMaxBatchChains - 1; c++)
    ok = hipStreamCreateWithFlags(&b->aux[c], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&b->ev_join[c], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    pme_batch_destroy(b);
    return set_error(PME_ERR_DEVICE, "pme_batch_create: auxiliary streams, events or work-queue counters failed");
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
  for (hipStream_t a : b->aux)
    if (a) (void)hipStreamSynchronize(a);
  drop_graph(b);
  for (auto& s : b->slots)
    if (s.used) free_slot_state(b, s);
  b->tables.release();
  (void)hipFree(b->d_sys);
  (void)hipFree(b->d_ev);
  (void)hipHostFree(b->h_ev);
  (void)hipFree(b->d_items);
  (void)hipFree(b->d_ctr);
  for (hipEvent_t e : b->tev)
    if (e) (void)hipEventDestroy(e);
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_done) (void)hipEventDestroy(b->ev_done);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  for (hipEvent_t e : b->ev_join)
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t a : b->aux)
    if (a) (void)hipStreamDestroy(a);
  if (b->own_stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return PME_OK;
}

extern "C" pme_status pme_batch_add(pme_batch_t b, const pme_system_desc* sys, int* id) {
  if (!b || !id) return set_error(PME_ERR_INVALID_ARG, "pme_batch_add: null argument");
  BTRY(validate_system(sys));
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
  BTRY(validate_system(sys));
  Slot& s = b->slots[id];
  const pme_system_desc& o = s.in;
  const bool same_plan = sys->nfftx == o.nfftx && sys->nffty == o.nffty && sys->nfftz == o.nfftz &&
                         sys->order == o.order && sys->precision == o.precision && sys->ncoord <= s.cap_atoms;
  const int old_nchunk = s.d.bin.nchunk;
  BTRY(hipStreamSynchronize(b->stream) == hipSuccess ? PME_OK : set_error(PME_ERR_DEVICE, "sync"));
  if (same_plan) {
    s.in = *sys;
    fill_dynamic(s.d, s.in, s.plan);
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
// contiguous run of items, systems in decreasing cost so the largest tiles
// are dispatched first.  Items index the slot table.
static pme_status build_lists(pme_batch_t b, const std::vector<int>& ids) {
  std::vector<int> order_ids = ids;
  std::sort(order_ids.begin(), order_ids.end(), [&](int x, int y) {
    const double ca = system_cost(b->slots[x].in), cc = system_cost(b->slots[y].in);
    return ca > cc || (ca == cc && x < y);
  });
  // several chains (not while timing stages): systems dealt in turn in
  // decreasing cost, each chain planned for its share of the CUs
  const int nc = b->timing ? 1 : std::max(1, std::min(b->nchains, (int)order_ids.size()));
  std::vector<int> part[kMaxBatchChains];
  for (size_t k = 0; k < order_ids.size(); k++) part[k % nc].push_back(order_ids[k]);
  std::vector<WorkItem> items;
  for (auto& l : b->launches_aux) l.clear();
  for (int c = 0; c < nc; c++) {
    std::vector<const SysDesc*> descs;
    std::vector<const SystemPlan*> plans;
    std::vector<int> orders;
    for (int id : part[c]) {
      descs.push_back(&b->slots[id].d);
      plans.push_back(&b->slots[id].plan);
      orders.push_back(b->slots[id].in.order);
    }
    std::vector<WorkItem> it;
    std::vector<Launch>& L = c == 0 ? b->launches : b->launches_aux[c - 1];
    build_items(descs.data(), plans.data(), orders.data(), (int)part[c].size(), std::max(1, b->ncu / nc), it, L);
    for (Launch& l : L) l.offset += (int)items.size();
    for (WorkItem& w : it) w.sys = part[c][w.sys];
    items.insert(items.end(), it.begin(), it.end());
  }
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

// The stages, then the copy of every slot's energy/virial into the pinned
// mirror (part of the captured graph), so reading results needs no further
// device work.
static pme_status enqueue(pme_batch_t b) {
  if (!b->launches_aux[0].empty()) {  // fork the other chains onto aux, join before the copy
    BHIP(hipEventRecord(b->ev_fork, b->stream));
    for (int c = 1; c < kMaxBatchChains; c++)
      if (!b->launches_aux[c - 1].empty()) BHIP(hipStreamWaitEvent(b->aux[c - 1], b->ev_fork, 0));
    BTRY(enqueue_stages(b->launches, b->d_sys, b->d_items, b->d_ctr, b->tune, b->ncu, b->stream, nullptr));
    for (int c = 1; c < kMaxBatchChains; c++) {
      if (b->launches_aux[c - 1].empty()) continue;
      BTRY(enqueue_stages(b->launches_aux[c - 1], b->d_sys, b->d_items, b->d_ctr + 2 * kMaxLaunches * c, b->tune,
                          b->ncu, b->aux[c - 1], nullptr));
      BHIP(hipEventRecord(b->ev_join[c - 1], b->aux[c - 1]));
      BHIP(hipStreamWaitEvent(b->stream, b->ev_join[c - 1], 0));
    }
  } else {
    BTRY(enqueue_stages(b->launches, b->d_sys, b->d_items, b->d_ctr, b->tune, b->ncu, b->stream,
                        b->timing ? b->tev : nullptr));
  }
  const int n = (int)b->slots.size();
  if (n > 0)
    BHIP(hipMemcpyAsync(b->h_ev, b->d_ev, sizeof(double) * kEvOut * (size_t)n, hipMemcpyDeviceToHost, b->stream));
  b->ev_copied = n;
  return PME_OK;
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
  // after the null stream's pending work (inputs written there); an idle
  // null stream has nothing to order against, and skipping the cross-stream
  // wait then saves a dependency hop on the device per solve
  if (b->own_stream && hipStreamQuery(nullptr) != hipSuccess) {
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

// Wait for the stream and make h_ev current for the listed ids (v: resolved).
static pme_status fetch_ev(pme_batch_t b, int n, const int* ids, const char* who, std::vector<int>& v) {
  v = resolve_ids(b, n, ids);
  BHIP(hipStreamSynchronize(b->stream));
  int hi = 0;
  for (int id : v) {
    if (id < 0 || id >= (int)b->slots.size() || !b->slots[id].used)
      return set_error(PME_ERR_INVALID_ARG, "%s: unknown id %d", who, id);
    hi = std::max(hi, id + 1);
  }
  // the solves copy d_ev into h_ev themselves; slots added since then are
  // read here
  if (hi > b->ev_copied) BHIP(hipMemcpy(b->h_ev, b->d_ev, sizeof(double) * kEvOut * hi, hipMemcpyDeviceToHost));
  if (b->timing && b->lists_valid)
    for (int k = 0; k < 8; k++) BHIP(hipEventElapsedTime(&b->stage_ms[k], b->tev[k], b->tev[k + 1]));
  return PME_OK;
}

extern "C" pme_status pme_batch_results(pme_batch_t b, int n, const int* ids, double* energy, double* virial) {
  if (!b) return set_error(PME_ERR_INVALID_ARG, "pme_batch_results: null batch");
  std::vector<int> v;
  const pme_status st = fetch_ev(b, n, ids, "pme_batch_results", v);
  if (st != PME_OK) return st;
  for (size_t k = 0; k < v.size(); k++) {
    const double* e = b->h_ev + kEvOut * (size_t)v[k];
    const bool on = (b->slots[v[k]].in.calc_energy_virial & PME_CALC_EV) != 0;
    if (energy) energy[k] = on ? e[0] : 0.0;
    if (virial)
      for (int c = 0; c < 9; c++) virial[9 * k + c] = on ? e[1 + c] : 0.0;
  }
  return PME_OK;
}

extern "C" pme_status pme_batch_results_self(pme_batch_t b, int n, const int* ids, double* energy_self) {
  if (!b) return set_error(PME_ERR_INVALID_ARG, "pme_batch_results_self: null batch");
  std::vector<int> v;
  const pme_status st = fetch_ev(b, n, ids, "pme_batch_results_self", v);
  if (st != PME_OK) return st;
  for (size_t k = 0; k < v.size(); k++) {
    const bool on = (b->slots[v[k]].in.calc_energy_virial & PME_CALC_SELF) != 0;
    if (energy_self) energy_self[k] = on ? b->h_ev[kEvOut * (size_t)v[k] + 10] : 0.0;
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
    SystemPlan np;
    BTRY(plan_system(s.in, s.cap_atoms, b->tune, &np));
    total = total - s.bytes + np.total;
  }
  if (total > b->budget) return set_error(PME_ERR_BUDGET, "re-plan needs %zu of %zu bytes", total, b->budget);
  BHIP(hipStreamSynchronize(b->stream));
  invalidate(b);
  for (size_t i = 0; i < b->slots.size(); i++) {
    Slot& s = b->slots[i];
    if (!s.used) continue;
    free_slot_state(b, s);
    pme_status st = plan_system(s.in, s.cap_atoms, b->tune, &s.plan);
    if (st == PME_OK) st = alloc_slot(b, s, (int)i);
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
  if (b->timing != (enable != 0)) b->lists_valid = false;  // timed solves run as one chain
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
  const bool replan = brick_target != b->tune.brick_target || fft_tile != b->tune.fft_tile;
  const Tuning old = b->tune;
  b->tune.brick_target = brick_target;
  b->tune.fft_tile = fft_tile;
  b->tune.wg_per_cu = wg_per_cu;
  drop_graph(b);
  if (replan) {
    pme_status st = replan_all(b);
    if (st == PME_ERR_BUDGET || st == PME_ERR_UNSUPPORTED) b->tune = old;  // nothing was freed
    return st;
  }
  return PME_OK;
}

extern "C" pme_status pme_batch_set_half_bricks(pme_batch_t b, int on) {
  if (!b) return set_error(PME_ERR_INVALID_ARG, "pme_batch_set_half_bricks: null batch");
  if (on != 0 && on != 1) return set_error(PME_ERR_INVALID_ARG, "pme_batch_set_half_bricks: %d (0 or 1)", on);
  if (on == b->tune.half_bricks) return PME_OK;
  const Tuning old = b->tune;
  b->tune.half_bricks = on;
  drop_graph(b);
  pme_status st = replan_all(b);
  if (st == PME_ERR_BUDGET || st == PME_ERR_UNSUPPORTED) b->tune = old;  // nothing was freed
  return st;
}

extern "C" pme_status pme_batch_get_half_bricks(pme_batch_t b, int* on) {
  if (!b || !on) return set_error(PME_ERR_INVALID_ARG, "pme_batch_get_half_bricks: null argument");
  *on = b->tune.half_bricks;
  return PME_OK;
}

extern "C" pme_status pme_batch_set_chains(pme_batch_t b, int nchains) {
  if (!b) return set_error(PME_ERR_INVALID_ARG, "pme_batch_set_chains: null batch");
  if (nchains < 1 || nchains > kMaxBatchChains)
    return set_error(PME_ERR_INVALID_ARG, "pme_batch_set_chains: %d (1 to %d)", nchains, kMaxBatchChains);
  if (nchains != b->nchains) {
    BHIP(hipStreamSynchronize(b->stream));
    b->nchains = nchains;
    b->lists_valid = false;
    drop_graph(b);
  }
  return PME_OK;
}

extern "C" pme_status pme_batch_get_chains(pme_batch_t b, int* nchains) {
  if (!b || !nchains) return set_error(PME_ERR_INVALID_ARG, "pme_batch_get_chains: null argument");
  *nchains = b->nchains;
  return PME_OK;
}

extern "C" pme_status pme_batch_get_tuning(pme_batch_t b, int* brick_target, int* fft_tile, int* wg_per_cu) {
  if (!b) return set_error(PME_ERR_INVALID_ARG, "pme_batch_get_tuning: null batch");
  if (brick_target) *brick_target = b->tune.brick_target;
  if (fft_tile) *fft_tile = b->tune.fft_tile;
  if (wg_per_cu) *wg_per_cu = b->tune.wg_per_cu;
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

// cache file: one JSON object {"key": [brick, tile, wg_per_cu, chains, half_bricks, ms], ...}
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
    if (vals.size() == 6) m[key] = vals;  // brick, tile, wg/cu, chains, half bricks, ms
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
      << ", " << (int)kv.second[3] << ", " << (int)kv.second[4] << ", " << kv.second[5] << "]"
      << (++k < m.size() ? "," : "") << "\n";
  }
  f << "}\n";
}

// Mean device time of one solve, each solve timed alone (events around it,
// then a wait), as a caller that reads every solve's results before the next
// one runs it: back-to-back solves would overlap one solve's tail with the
// next one's start and favour plans that only pipeline well.
static pme_status time_solves(pme_batch_t b, int warm, int reps, float* ms) {
  hipEvent_t e0, e1;
  BHIP(hipEventCreate(&e0));
  BHIP(hipEventCreate(&e1));
  for (int i = 0; i < warm; i++) BTRY(pme_batch_solve(b, 0, nullptr));
  BHIP(hipStreamSynchronize(b->stream));
  float sum = 0.f;
  for (int i = 0; i < reps; i++) {
    BHIP(hipEventRecord(e0, b->stream));
    BTRY(pme_batch_solve(b, 0, nullptr));
    BHIP(hipEventRecord(e1, b->stream));
    BHIP(hipEventSynchronize(e1));
    float t = 0.f;
    BHIP(hipEventElapsedTime(&t, e0, e1));
    sum += t;
  }
  *ms = sum / reps;
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
    BTRY(pme_batch_set_half_bricks(b, (int)hit->second[4]));
    BTRY(pme_batch_set_tuning(b, (int)hit->second[0], (int)hit->second[1], (int)hit->second[2]));
    BTRY(pme_batch_set_chains(b, (int)hit->second[3]));
    if (best_ms) *best_ms = (float)hit->second[5];
    return PME_OK;
  }
  // the outputs of INT64 systems accumulate: tuning solves would add into them
  for (const Slot& s : b->slots)
    if (s.used && s.in.force_type == PME_FORCE_INT64)
      return set_error(PME_ERR_STATE, "pme_batch_autotune: resident INT64 (accumulating) outputs");
  // a sweep of 10 solves per plan, then the 4 fastest again with 40 each: the
  // minimum over 120 short timings alone favours plans that were lucky once
  struct Cand {
    float ms;
    int half, brick, tile, wpc, ch;
  };
  std::vector<Cand> cand;
  auto apply = [&](const Cand& c) {
    return pme_batch_set_half_bricks(b, c.half) == PME_OK && pme_batch_set_tuning(b, c.brick, c.tile, c.wpc) == PME_OK &&
           pme_batch_set_chains(b, c.ch) == PME_OK;
  };
  for (int half : {0, 1})
    for (int brick : {12, 16, 20, 24, 32})
      for (int tile : {8, 16, 32})
        for (int wpc : {0, 4})
          for (int ch : {1, 2}) {
            Cand c{0.f, half, brick, tile, wpc, ch};
            if (!apply(c) || time_solves(b, 3, 10, &c.ms) != PME_OK) continue;
            cand.push_back(c);
          }
  if (cand.empty()) return set_error(PME_ERR_UNSUPPORTED, "pme_batch_autotune: no plan could be timed");
  std::sort(cand.begin(), cand.end(), [](const Cand& x, const Cand& y) { return x.ms < y.ms; });
  float best = 1e30f;
  Cand win = cand[0];
  for (size_t i = 0; i < cand.size() && i < 4; i++) {
    Cand c = cand[i];
    if (!apply(c) || time_solves(b, 2, 40, &c.ms) != PME_OK) continue;
    if (c.ms < best) {
      best = c.ms;
      win = c;
    }
  }
  if (best >= 1e30f) best = win.ms;
  const int bb = win.brick, bt = win.tile, bw = win.wpc, bc = win.ch, bh = win.half;
  BTRY(pme_batch_set_half_bricks(b, bh));
  BTRY(pme_batch_set_tuning(b, bb, bt, bw));
  BTRY(pme_batch_set_chains(b, bc));
  cache[key] = {(double)bb, (double)bt, (double)bw, (double)bc, (double)bh, (double)best};
  write_cache(cache_path, cache);
  if (best_ms) *best_ms = best;
  return PME_OK;
}
