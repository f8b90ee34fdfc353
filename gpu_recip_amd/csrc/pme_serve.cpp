// pme_serve.cpp -- continuous-batching serving loop for PME solves.
//
// Requests (a system descriptor under a caller-chosen key) are queued by
// any thread.  One worker thread owns the device side:
//   retire  -- a lane whose completion event has fired publishes the
//              energies/virials of its requests (already copied to pinned
//              host memory on the lane's stream) and their latencies;
//   admit   -- queued requests, oldest first, get their resident state: a
//              resident key is reused (or re-planned if its plan no longer
//              fits), a new key is planned and placed in the HBM arena,
//              evicting idle resident systems least-recently-used first;
//   launch  -- every admitted request that is not in flight goes into ONE
//              batched solve on an idle lane: descriptors and work lists are
//              staged in the lane's pinned buffer and copied with the stage
//              kernels on the lane's stream, so nothing blocks the host.
// Several lanes let a new batch start while earlier ones still run; a key
// is never in two lanes at once (its grids are shared state).
// The arena is one hipMalloc at creation: admission and eviction never call
// the allocator (hipFree would synchronise the device).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/pme_recip.h"
#include "pme_arena.hpp"
#include "pme_sched.hpp"

using namespace pme;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Resident {
  pme_system_desc in{};
  int cap_atoms = 0;
  SystemPlan plan{};
  SysDesc d{};
  size_t off = 0, bytes = 0;
  bool busy = false;  // in a lane
  unsigned long long last_use = 0;
};

struct Request {
  unsigned long long ticket;
  int key;
  pme_system_desc in;
  hipEvent_t ready;
  double t_submit;
};

struct Result {
  bool done = false;
  pme_status status = PME_OK;
  std::string msg;
  double energy = 0.0;
  double energy_self = 0.0;
  double virial[9] = {};
  double latency_ms = 0.0;
};

struct Lane {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<Request> reqs;  // in flight, in lane order
  SysDesc* d_sys = nullptr;
  WorkItem* d_items = nullptr;
  int* d_ctr = nullptr;              // work-queue counters, 2 per launch (zeroed at creation)
  double* d_ev = nullptr;
  size_t cap_sys = 0, cap_items = 0;
  unsigned char* h_stage = nullptr;  // pinned: descriptors then items
  size_t stage_cap = 0;
  double* h_ev = nullptr;            // pinned, cap_sys * kEvOut
  std::vector<void*> retired;        // buffers replaced while growing, freed at the end
};

}  // namespace

struct pme_server_s {
  int device = 0;
  int ncu = 256;
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  std::deque<Request> incoming;                        // guarded by mu
  std::unordered_map<unsigned long long, Result> results;  // guarded by mu
  unsigned long long next_ticket = 1;
  unsigned long long outstanding = 0;                  // submitted, not completed (mu)
  Tuning tune{kServeBrick, kServeFftTile, 0, 0};       // guarded by mu
  std::vector<int> release_keys;                       // guarded by mu
  // stats (guarded by mu)
  pme_server_stats st{};
  std::vector<double> lat_ms;
  unsigned long long batched_reqs = 0;
  // worker-only
  void* arena_base = nullptr;
  Arena arena;
  TableCache tables;
  std::unordered_map<int, Resident> resident;
  std::set<std::pair<unsigned long long, int>> lru;  // (last_use, key) of every resident, oldest first
  std::deque<Request> pending;
  std::vector<Lane> lanes;
  unsigned long long use_clock = 0;
};

// Mark a resident as just used (its place in the LRU order).
static void lru_touch(pme_server_t s, int key, Resident& R) {
  if (R.last_use) s->lru.erase({R.last_use, key});
  R.last_use = ++s->use_clock;
  s->lru.insert({R.last_use, key});
}

// Drop a resident: its arena block, LRU entry and map entry.  Returns the
// free block its memory joined.
static std::pair<size_t, size_t> drop_resident(pme_server_t s, std::unordered_map<int, Resident>::iterator it) {
  const std::pair<size_t, size_t> blk = s->arena.release(it->second.off, it->second.bytes);
  s->lru.erase({it->second.last_use, it->first});
  s->resident.erase(it);
  return blk;
}

static void publish(pme_server_t s, const Request& r, pme_status status, const std::string& msg, const double* ev,
                    double t_done) {
  std::lock_guard<std::mutex> g(s->mu);
  Result& res = s->results[r.ticket];
  res.done = true;
  res.status = status;
  res.msg = msg;
  if (ev) {  // kEvOut doubles: energy, virial[9], self energy
    res.energy = ev[0];
    for (int c = 0; c < 9; c++) res.virial[c] = ev[1 + c];
    res.energy_self = ev[10];
  }
  res.latency_ms = (t_done - r.t_submit) * 1e3;
  if (status == PME_OK) {
    s->st.completed++;
    s->lat_ms.push_back(res.latency_ms);
  } else {
    s->st.failed++;
  }
  s->outstanding--;
}

// ---- admission ----------------------------------------------------------------

// Evict idle residents (LRU first) until `bytes` fit; never touches `keep`
// (the keys already selected for the batch being formed).  The walk from the
// oldest entry skips only residents in flight or selected, which were used
// last.  When the first-fit search fails, only the block an eviction frees
// (merged with its free neighbours) can newly fit, so that block alone is
// checked after each eviction.
static bool make_room(pme_server_t s, size_t bytes, const std::unordered_set<int>& keep, size_t* off) {
  if (s->arena.alloc(bytes, off)) return true;
  const size_t need = (bytes + 255) & ~(size_t)255;
  for (;;) {
    auto victim = s->resident.end();
    for (const auto& e : s->lru) {
      auto it = s->resident.find(e.second);
      if (it->second.busy || keep.count(e.second)) continue;
      victim = it;
      break;
    }
    if (victim == s->resident.end()) return false;
    const std::pair<size_t, size_t> blk = drop_resident(s, victim);
    {
      std::lock_guard<std::mutex> g(s->mu);
      s->st.evictions++;
    }
    if (blk.second >= need) {
      s->arena.take(blk.first, need, off);
      return true;
    }
  }
}

enum AdmitResult { ADMIT_OK, ADMIT_WAIT, ADMIT_FAIL };

static AdmitResult admit(pme_server_t s, const Request& r, const Tuning& tune, const std::unordered_set<int>& keep,
                         std::string* err) {
  auto it = s->resident.find(r.key);
  if (it != s->resident.end()) {
    Resident& R = it->second;
    const bool same = r.in.nfftx == R.in.nfftx && r.in.nffty == R.in.nffty && r.in.nfftz == R.in.nfftz &&
                      r.in.order == R.in.order && r.in.precision == R.in.precision && r.in.ncoord <= R.cap_atoms;
    if (same) {
      R.in = r.in;
      fill_dynamic(R.d, R.in, R.plan);
      lru_touch(s, r.key, R);
      return ADMIT_OK;
    }
    drop_resident(s, it);  // idle (keys in flight are not admitted)
  }
  Resident R;
  R.in = r.in;
  R.cap_atoms = std::max(r.in.ncoord, 1);
  if (plan_system(R.in, R.cap_atoms, tune, &R.plan) != PME_OK) {
    *err = pme_last_error();
    return ADMIT_FAIL;
  }
  R.bytes = R.plan.total;
  if (R.bytes > s->arena.size) {
    *err = "system needs more than the whole arena";
    return ADMIT_FAIL;
  }
  if (!make_room(s, R.bytes, keep, &R.off)) return ADMIT_WAIT;
  if (fill_desc(R.d, R.in, R.plan, static_cast<unsigned char*>(s->arena_base) + R.off, s->tables) != PME_OK) {
    s->arena.release(R.off, R.bytes);
    *err = pme_last_error();
    return ADMIT_FAIL;
  }
  Resident& slot = s->resident[r.key] = R;
  lru_touch(s, r.key, slot);
  std::lock_guard<std::mutex> g(s->mu);
  s->st.admitted_new++;
  return ADMIT_OK;
}

// ---- lanes ----------------------------------------------------------------------

static pme_status grow_lane(Lane& L, size_t nsys, size_t nitems) {
  if (nsys > L.cap_sys) {
    size_t cap = std::max(nsys, 2 * L.cap_sys);
    if (L.d_sys) L.retired.push_back(L.d_sys);
    if (L.d_ev) L.retired.push_back(L.d_ev);
    if (L.h_ev) (void)hipHostFree(L.h_ev);  // the lane is idle: no copy into it is pending
    L.d_sys = nullptr;
    L.d_ev = nullptr;
    L.h_ev = nullptr;
    if (hipMalloc(&L.d_sys, sizeof(SysDesc) * cap) != hipSuccess ||
        hipMalloc(&L.d_ev, sizeof(double) * kEvOut * cap) != hipSuccess ||
        hipHostMalloc(&L.h_ev, sizeof(double) * kEvOut * cap, hipHostMallocDefault) != hipSuccess)
      return set_error(PME_ERR_OOM, "serving lane buffers");
    L.cap_sys = cap;
  }
  if (nitems > L.cap_items) {
    size_t cap = std::max(nitems, 2 * L.cap_items);
    if (L.d_items) L.retired.push_back(L.d_items);
    L.d_items = nullptr;
    if (hipMalloc(&L.d_items, sizeof(WorkItem) * cap) != hipSuccess) return set_error(PME_ERR_OOM, "serving lane items");
    L.cap_items = cap;
  }
  const size_t need = sizeof(SysDesc) * L.cap_sys + sizeof(WorkItem) * L.cap_items;
  if (need > L.stage_cap) {
    if (L.h_stage) (void)hipHostFree(L.h_stage);
    L.h_stage = nullptr;
    if (hipHostMalloc(&L.h_stage, need, hipHostMallocDefault) != hipSuccess)
      return set_error(PME_ERR_OOM, "serving lane staging");
    L.stage_cap = need;
  }
  return PME_OK;
}

static pme_status launch_lane(pme_server_t s, Lane& L, std::vector<Request>& sel, const Tuning& tune) {
  // largest first, so the big tiles are dispatched first
  std::stable_sort(sel.begin(), sel.end(), [&](const Request& a, const Request& b) {
    return system_cost(a.in) > system_cost(b.in);
  });
  const int m = (int)sel.size();
  std::vector<SysDesc> descs(m);
  std::vector<const SysDesc*> dp(m);
  std::vector<const SystemPlan*> pp(m);
  std::vector<int> orders(m);
  for (int k = 0; k < m; k++) {
    Resident& R = s->resident[sel[k].key];
    descs[k] = R.d;
    dp[k] = &descs[k];
    pp[k] = &R.plan;
    orders[k] = R.in.order;
  }
  std::vector<WorkItem> items;
  std::vector<Launch> launches;
  build_items(dp.data(), pp.data(), orders.data(), m, s->ncu, items, launches);
  pme_status st = grow_lane(L, m, items.size());
  if (st != PME_OK) return st;
  for (int k = 0; k < m; k++) descs[k].ev_out = L.d_ev + kEvOut * (size_t)k;
  std::memcpy(L.h_stage, descs.data(), sizeof(SysDesc) * m);
  std::memcpy(L.h_stage + sizeof(SysDesc) * L.cap_sys, items.data(), sizeof(WorkItem) * items.size());
  for (const Request& r : sel)
    if (r.ready && hipStreamWaitEvent(L.stream, r.ready, 0) != hipSuccess)
      return set_error(PME_ERR_DEVICE, "hipStreamWaitEvent on a ready event");
  if (hipMemcpyAsync(L.d_sys, L.h_stage, sizeof(SysDesc) * m, hipMemcpyHostToDevice, L.stream) != hipSuccess ||
      hipMemcpyAsync(L.d_items, L.h_stage + sizeof(SysDesc) * L.cap_sys, sizeof(WorkItem) * items.size(),
                     hipMemcpyHostToDevice, L.stream) != hipSuccess)
    return set_error(PME_ERR_DEVICE, "serving lane upload");
  st = enqueue_stages(launches, L.d_sys, L.d_items, L.d_ctr, tune, s->ncu, L.stream, nullptr);
  if (st != PME_OK) return st;
  if (hipMemcpyAsync(L.h_ev, L.d_ev, sizeof(double) * kEvOut * m, hipMemcpyDeviceToHost, L.stream) != hipSuccess ||
      hipEventRecord(L.done, L.stream) != hipSuccess)
    return set_error(PME_ERR_DEVICE, "serving lane readback");
  for (const Request& r : sel) s->resident[r.key].busy = true;
  L.reqs = sel;
  L.busy = true;
  std::lock_guard<std::mutex> g(s->mu);
  s->st.batches++;
  s->batched_reqs += m;
  s->st.max_batch = std::max(s->st.max_batch, m);
  return PME_OK;
}

static bool retire(pme_server_t s) {
  bool any = false;
  for (Lane& L : s->lanes) {
    if (!L.busy) continue;
    const hipError_t q = hipEventQuery(L.done);
    if (q == hipErrorNotReady) continue;
    const double t = now_s();
    const std::string msg = q == hipSuccess ? "" : hipGetErrorString(q);
    for (size_t k = 0; k < L.reqs.size(); k++) {
      const Request& r = L.reqs[k];
      auto it = s->resident.find(r.key);
      if (it != s->resident.end()) {
        it->second.busy = false;
        lru_touch(s, r.key, it->second);
      }
      double ev[kEvOut] = {};
      const double* h = L.h_ev + kEvOut * k;
      if (r.in.calc_energy_virial & PME_CALC_EV) std::memcpy(ev, h, sizeof(double) * 10);
      if (r.in.calc_energy_virial & PME_CALC_SELF) ev[10] = h[10];
      publish(s, r, q == hipSuccess ? PME_OK : PME_ERR_DEVICE, msg, ev, t);
    }
    L.reqs.clear();
    L.busy = false;
    any = true;
  }
  if (any) s->cv_done.notify_all();
  return any;
}

static void worker_main(pme_server_t s) {
  (void)hipSetDevice(s->device);
  for (;;) {
    Tuning tune;
    std::vector<int> rel;
    {
      std::unique_lock<std::mutex> g(s->mu);
      while (!s->incoming.empty()) {
        s->pending.push_back(s->incoming.front());
        s->incoming.pop_front();
      }
      tune = s->tune;
      rel.swap(s->release_keys);
      bool lanes_busy = false;
      for (const Lane& L : s->lanes) lanes_busy = lanes_busy || L.busy;
      if (s->stop && s->pending.empty() && !lanes_busy) break;
    }
    if (!rel.empty()) {
      std::unordered_set<int> queued;
      for (const Request& r : s->pending) queued.insert(r.key);
      for (int key : rel) {
        auto it = s->resident.find(key);
        if (it != s->resident.end() && !it->second.busy && !queued.count(key)) drop_resident(s, it);
      }
    }
    bool progress = retire(s);
    Lane* idle = nullptr;
    for (Lane& L : s->lanes)
      if (!L.busy) {
        idle = &L;
        break;
      }
    if (idle && !s->pending.empty()) {
      std::vector<Request> sel;
      std::unordered_set<int> keys;
      std::deque<Request> rest;
      bool blocked = false;
      bool any_busy = false;
      for (const Lane& L : s->lanes) any_busy = any_busy || L.busy;
      while (!s->pending.empty()) {
        Request r = s->pending.front();
        s->pending.pop_front();
        auto it = s->resident.find(r.key);
        const bool in_flight = it != s->resident.end() && it->second.busy;
        if (blocked || in_flight || keys.count(r.key)) {
          rest.push_back(r);  // keeps its place for the next batch
          continue;
        }
        std::string err;
        const AdmitResult a = admit(s, r, tune, keys, &err);
        if (a == ADMIT_OK) {
          sel.push_back(r);
          keys.insert(r.key);
        } else if (a == ADMIT_WAIT && (any_busy || !sel.empty())) {
          blocked = true;  // oldest first: wait for space instead of overtaking
          rest.push_back(r);
        } else {
          publish(s, r, a == ADMIT_WAIT ? PME_ERR_BUDGET : PME_ERR_UNSUPPORTED,
                  a == ADMIT_WAIT ? std::string("does not fit the arena") : err, nullptr, now_s());
          s->cv_done.notify_all();
        }
      }
      s->pending.swap(rest);
      if (!sel.empty()) {
        const pme_status st = launch_lane(s, *idle, sel, tune);
        if (st != PME_OK) {
          const std::string msg = pme_last_error();
          for (const Request& r : sel) {
            auto it = s->resident.find(r.key);
            if (it != s->resident.end()) it->second.busy = false;
            publish(s, r, st, msg, nullptr, now_s());
          }
          idle->reqs.clear();
          idle->busy = false;
          s->cv_done.notify_all();
        }
        progress = true;
      }
    }
    bool lanes_busy = false;
    for (const Lane& L : s->lanes) lanes_busy = lanes_busy || L.busy;
    {
      std::unique_lock<std::mutex> g(s->mu);
      s->st.resident = (int)s->resident.size();
      s->st.arena_used = s->arena.used;
      // Idle: sleep until a submission.  Lanes in flight: poll their events
      // (a yield, not a timed sleep, so completion is seen within
      // microseconds).
      if (!progress && !lanes_busy && s->incoming.empty() && !s->stop)
        s->cv_work.wait_for(g, std::chrono::milliseconds(5));
    }
    if (!progress && lanes_busy) std::this_thread::yield();
  }
}

// ---- C ABI ---------------------------------------------------------------------

extern "C" pme_status pme_server_create(size_t hbm_budget_bytes, int nlanes, pme_server_t* out) {
  if (!out) return set_error(PME_ERR_INVALID_ARG, "pme_server_create: null out");
  *out = nullptr;
  if (nlanes < 1 || nlanes > 8) return set_error(PME_ERR_INVALID_ARG, "pme_server_create: nlanes %d", nlanes);
  if (hbm_budget_bytes < (1u << 20)) return set_error(PME_ERR_INVALID_ARG, "pme_server_create: budget below 1 MiB");
  pme_server_t s = new pme_server_s();
  hipDeviceProp_t prop;
  if (hipGetDevice(&s->device) != hipSuccess || hipGetDeviceProperties(&prop, s->device) != hipSuccess) {
    delete s;
    return set_error(PME_ERR_DEVICE, "pme_server_create: no device");
  }
  s->ncu = prop.multiProcessorCount;
  hipError_t e = hipMalloc(&s->arena_base, hbm_budget_bytes);
  if (e != hipSuccess) {
    delete s;
    return set_error(PME_ERR_OOM, "pme_server_create: arena of %zu bytes: %s", hbm_budget_bytes, hipGetErrorString(e));
  }
  s->arena.init(hbm_budget_bytes);
  s->lanes.resize(nlanes);
  for (Lane& L : s->lanes) {
    if (hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&L.done, hipEventDisableTiming) != hipSuccess ||
        hipMalloc(&L.d_ctr, sizeof(int) * 2 * kMaxLaunches) != hipSuccess ||
        hipMemset(L.d_ctr, 0, sizeof(int) * 2 * kMaxLaunches) != hipSuccess) {
      pme_server_destroy(s);
      return set_error(PME_ERR_DEVICE, "pme_server_create: lane stream, event or work-queue counters failed");
    }
  }
  s->st.arena_budget = s->arena.size;
  s->worker = std::thread(worker_main, s);
  *out = s;
  return PME_OK;
}

extern "C" pme_status pme_server_destroy(pme_server_t s) {
  if (!s) return PME_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (Lane& L : s->lanes) {
    if (L.stream) (void)hipStreamSynchronize(L.stream);
    (void)hipFree(L.d_sys);
    (void)hipFree(L.d_items);
    (void)hipFree(L.d_ctr);
    (void)hipFree(L.d_ev);
    for (void* p : L.retired) (void)hipFree(p);
    if (L.h_stage) (void)hipHostFree(L.h_stage);
    if (L.h_ev) (void)hipHostFree(L.h_ev);
    if (L.done) (void)hipEventDestroy(L.done);
    if (L.stream) (void)hipStreamDestroy(L.stream);
  }
  s->tables.release();
  (void)hipFree(s->arena_base);
  delete s;
  return PME_OK;
}

extern "C" pme_status pme_server_submit(pme_server_t s, int key, const pme_system_desc* sys, void* ready_event,
                                        unsigned long long* ticket) {
  if (!s || !ticket) return set_error(PME_ERR_INVALID_ARG, "pme_server_submit: null argument");
  if (key < 0) return set_error(PME_ERR_INVALID_ARG, "pme_server_submit: key %d", key);
  pme_status st = validate_system(sys);
  if (st != PME_OK) return st;
  Request r;
  r.key = key;
  r.in = *sys;
  r.ready = (hipEvent_t)ready_event;
  {
    std::lock_guard<std::mutex> g(s->mu);
    SystemPlan p;
    st = plan_system(*sys, std::max(sys->ncoord, 1), s->tune, &p);
    if (st != PME_OK) return st;
    if (p.total > s->arena.size)
      return set_error(PME_ERR_BUDGET, "pme_server_submit: system needs %zu bytes, arena is %zu", p.total, s->arena.size);
    if (s->stop) return set_error(PME_ERR_STATE, "pme_server_submit: server is stopping");
    r.ticket = s->next_ticket++;
    r.t_submit = now_s();
    s->incoming.push_back(r);
    s->results[r.ticket] = Result();
    s->outstanding++;
    s->st.submitted++;
  }
  s->cv_work.notify_one();
  *ticket = r.ticket;
  return PME_OK;
}

extern "C" pme_status pme_server_submit_header(pme_server_t s, int key, const pme_recip_header* h, int nfftx,
                                               int nffty, int nfftz, int order, double kappa, const void* d_xyzq,
                                               void* d_force3, unsigned long long* ticket) {
  if (!h) return set_error(PME_ERR_INVALID_ARG, "pme_server_submit_header: null header");
  // negative ncoord: the STOP message of DomdecRecipComm::send_stop
  if (h->ncoord < 0) return set_error(PME_ERR_STOPPED, "pme_server_submit_header: STOP header (ncoord %d)", h->ncoord);
  pme_system_desc d{};
  d.nfftx = nfftx;
  d.nffty = nffty;
  d.nfftz = nfftz;
  d.order = order;
  d.kappa = kappa;
  d.recip[0] = h->inv_boxx;
  d.recip[4] = h->inv_boxy;
  d.recip[8] = h->inv_boxz;
  d.ncoord = h->ncoord;
  d.d_xyzq = d_xyzq;
  d.d_force = d_force3;
  d.force_type = PME_FORCE_FLOAT3;
  d.stride = h->ncoord;
  // CudaDomdecRecip::calc (CudaDomdecRecip.h:70-78): scalar_sum with the
  // two flags, then calc_self_energy when calc_energy is set
  d.calc_energy_virial = ((h->calc_energy || h->calc_virial) ? PME_CALC_EV : 0) | (h->calc_energy ? PME_CALC_SELF : 0);
  d.precision = PME_PREC_FP32;
  return pme_server_submit(s, key, &d, nullptr, ticket);
}

static pme_status take_result(pme_server_t s, std::unordered_map<unsigned long long, Result>::iterator it,
                              double* energy, double* virial, double* latency_ms) {
  const Result r = it->second;
  s->results.erase(it);
  if (energy) *energy = r.energy;
  if (virial)
    for (int c = 0; c < 9; c++) virial[c] = r.virial[c];
  if (latency_ms) *latency_ms = r.latency_ms;
  if (r.status != PME_OK) return set_error(r.status, "request failed: %s", r.msg.c_str());
  return PME_OK;
}

extern "C" pme_status pme_server_wait(pme_server_t s, unsigned long long ticket, double* energy, double* virial,
                                      double* latency_ms) {
  if (!s) return set_error(PME_ERR_INVALID_ARG, "pme_server_wait: null server");
  std::unique_lock<std::mutex> g(s->mu);
  auto it = s->results.find(ticket);
  if (it == s->results.end()) return set_error(PME_ERR_NOT_FOUND, "pme_server_wait: unknown ticket %llu", ticket);
  s->cv_done.wait(g, [&] { return s->results[ticket].done; });
  return take_result(s, s->results.find(ticket), energy, virial, latency_ms);
}

extern "C" pme_status pme_server_wait_result(pme_server_t s, unsigned long long ticket, pme_request_result* out) {
  if (!s || !out) return set_error(PME_ERR_INVALID_ARG, "pme_server_wait_result: null argument");
  std::unique_lock<std::mutex> g(s->mu);
  auto it = s->results.find(ticket);
  if (it == s->results.end())
    return set_error(PME_ERR_NOT_FOUND, "pme_server_wait_result: unknown ticket %llu", ticket);
  s->cv_done.wait(g, [&] { return s->results[ticket].done; });
  it = s->results.find(ticket);
  out->energy_self = it->second.energy_self;
  return take_result(s, it, &out->energy, out->virial, &out->latency_ms);
}

extern "C" pme_status pme_server_poll(pme_server_t s, unsigned long long ticket, int* done, double* energy,
                                      double* virial, double* latency_ms) {
  if (!s || !done) return set_error(PME_ERR_INVALID_ARG, "pme_server_poll: null argument");
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->results.find(ticket);
  if (it == s->results.end()) return set_error(PME_ERR_NOT_FOUND, "pme_server_poll: unknown ticket %llu", ticket);
  *done = it->second.done ? 1 : 0;
  if (!*done) return PME_OK;
  return take_result(s, it, energy, virial, latency_ms);
}

extern "C" pme_status pme_server_drain(pme_server_t s) {
  if (!s) return set_error(PME_ERR_INVALID_ARG, "pme_server_drain: null server");
  std::unique_lock<std::mutex> g(s->mu);
  s->cv_done.wait(g, [&] { return s->outstanding == 0; });
  return PME_OK;
}

extern "C" pme_status pme_server_release(pme_server_t s, int key) {
  if (!s) return set_error(PME_ERR_INVALID_ARG, "pme_server_release: null server");
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->release_keys.push_back(key);
  }
  s->cv_work.notify_one();
  return PME_OK;
}

static double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t i = (size_t)pos;
  const double f = pos - (double)i;
  return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

extern "C" pme_status pme_server_stats_get(pme_server_t s, pme_server_stats* out) {
  if (!s || !out) return set_error(PME_ERR_INVALID_ARG, "pme_server_stats_get: null argument");
  std::lock_guard<std::mutex> g(s->mu);
  *out = s->st;
  out->mean_batch = s->st.batches ? (double)s->batched_reqs / (double)s->st.batches : 0.0;
  out->p50_ms = percentile(s->lat_ms, 0.50);
  out->p99_ms = percentile(s->lat_ms, 0.99);
  double sum = 0.0, mx = 0.0;
  for (double v : s->lat_ms) {
    sum += v;
    mx = std::max(mx, v);
  }
  out->mean_ms = s->lat_ms.empty() ? 0.0 : sum / (double)s->lat_ms.size();
  out->max_ms = mx;
  return PME_OK;
}

extern "C" pme_status pme_server_stats_reset(pme_server_t s) {
  if (!s) return set_error(PME_ERR_INVALID_ARG, "pme_server_stats_reset: null server");
  std::lock_guard<std::mutex> g(s->mu);
  const size_t budget = s->st.arena_budget;
  s->st = pme_server_stats{};
  s->st.arena_budget = budget;
  s->lat_ms.clear();
  s->batched_reqs = 0;
  return PME_OK;
}

extern "C" pme_status pme_server_set_tuning(pme_server_t s, int brick_target, int fft_tile, int wg_per_cu) {
  if (!s) return set_error(PME_ERR_INVALID_ARG, "pme_server_set_tuning: null server");
  if (brick_target < 8 || brick_target > 64 || fft_tile < 1 || fft_tile > 64 || (fft_tile & (fft_tile - 1)) ||
      wg_per_cu < 0 || wg_per_cu > 64)
    return set_error(PME_ERR_INVALID_ARG, "pme_server_set_tuning: brick %d tile %d wg/cu %d", brick_target, fft_tile,
                     wg_per_cu);
  std::lock_guard<std::mutex> g(s->mu);
  s->tune.brick_target = brick_target;
  s->tune.fft_tile = fft_tile;
  s->tune.wg_per_cu = wg_per_cu;
  return PME_OK;
}
