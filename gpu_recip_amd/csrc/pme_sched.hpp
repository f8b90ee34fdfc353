// pme_sched.hpp -- multi-system scheduling core shared by the batched path
// (pme_batch.cpp) and the serving loop (pme_serve.cpp): planning one
// system's resident state, filling its device descriptor, building the
// per-stage work lists of a set of systems and enqueueing the stage launches.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <map>
#include <utility>
#include <vector>

#include "../../include/pme_recip.h"
#include "pme_internal.hpp"
#include "pme_plan.hpp"

namespace pme {

// Resident state of one system under a tuning: plans and the byte size of
// every buffer (each piece 256-byte aligned inside one region).
struct SystemPlan {
  BrickPlan bp;
  ChunkPlan cp;
  RadixPlan rp[3];
  int prec[3];          // PREC_* per axis (the radix-16 stage on the matrix cores)
  FftTiles ft;
  int nzblk;            // z-pass tiles = energy/virial partials
  int plane;            // 1: x and y passes as one xy plane kernel
  size_t off[7];        // grid, cgrid, seg_off, rec_pos, rec_idx, partial, self_partial
  size_t total;         // bytes of the region
};

constexpr int kBatchChains = 2;     // default stage chains of a batched solve
constexpr int kMaxBatchChains = 4;  // most stage chains (streams) of a batched solve

// Default plan of the serving loop: its batches are many systems at once, so
// larger bricks (fewer halo records and per-brick phases) and wider FFT tiles
// keep the GPU full; measured on 3,000 requests of the serving schedule at
// saturation: 29,405 solves/s against 22,859 with the single-system defaults
// (brick 16, tile 16).
constexpr int kServeBrick = 24;
constexpr int kServeFftTile = 32;

struct Tuning {
  int brick_target = kDefaultBrick;
  int fft_tile = kDefaultFftTile;
  int wg_per_cu = 0;    // 0: one workgroup per work item
  int half_bricks = 0;  // 1: bricks within half the LDS budget, spread and gather in half-size
                        // workgroups, two per CU (one's HBM phases overlap the other's LDS work)
};

pme_status validate_system(const pme_system_desc* s);
// Plans for up to cap_atoms atoms; touches no device memory.
pme_status plan_system(const pme_system_desc& in, int cap_atoms, const Tuning& t, SystemPlan* p);

// Device twiddle and prefac tables shared by every system with the same
// edge (and order).  get() uploads on first use (synchronous copy).
struct TableCache {
  std::map<int, float2*> tw;
  std::map<std::pair<int, int>, float*> pf;
  size_t bytes = 0;
  pme_status get(int order, int n, const float2** tw_out, const float** pf_out);
  void release();
};

// Descriptor of a system whose region starts at `base`.  ev_out is left for
// the caller.
pme_status fill_desc(SysDesc& d, const pme_system_desc& in, const SystemPlan& p, void* base, TableCache& tables);
// The parts that change with coordinates, box and outputs.
void fill_dynamic(SysDesc& d, const pme_system_desc& in, const SystemPlan& p);

// One launch of a stage over the items [offset, offset + count).
struct Launch {
  int stage;   // BatchStage
  int order;   // spread/gather template order; largest plane group of a plane launch; z pass: 1 + the
               // four-step register class, 0 for the Stockham / split kernels (0 otherwise)
  int split;   // FFT launches: 1 for the kernels with the split-precision matrix-core stage
  int half;    // plane launches: 1 for half-size workgroups (two per CU)
  int offset;
  int count;
  size_t lds;  // dynamic LDS bytes: max over the launch's systems
};

// Work lists of systems descs[0..n): items refer to the index in descs.
// Systems are visited in the given order (callers sort by cost).
void build_items(const SysDesc* const* descs, const SystemPlan* const* plans, const int* orders, int n, int ncu,
                 std::vector<WorkItem>& items, std::vector<Launch>& launches);

// Enqueue the eight stages (bin, spread, x, y, z+solve, y inverse with the
// energy/virial reduction, x inverse, gather) on `stream`.  With tev != NULL,
// tev[k] is recorded before stage k and tev[8] after the last.  d_ctr: two
// zeroed ints per launch (the persistent launches' work-queue counters, left
// zeroed by each launch); launches that may run at the same time need
// distinct counters.
constexpr int kMaxLaunches = 64;  // launches of one stage sequence (d_ctr holds 2 ints each)
pme_status enqueue_stages(const std::vector<Launch>& launches, const SysDesc* d_sys, const WorkItem* d_items,
                          int* d_ctr, const Tuning& t, int ncu, hipStream_t stream, hipEvent_t* tev);

// Cost used to order systems (largest first).
double system_cost(const pme_system_desc& in);

}  // namespace pme
