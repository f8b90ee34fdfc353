// pme_plan.hpp -- host-side planning shared by the single-system solver and
// the batched path: FP64 bsp_mod, FFT radix plans and tiles, brick plans,
// record-store sizing.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "../../include/pme_recip.h"
#include "pme_internal.hpp"

namespace pme {

// Records a printf-style message for pme_last_error() (per thread) and
// returns s.
pme_status set_error(pme_status s, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

constexpr size_t kLdsBudget = 64 * 1024;        // FFT kernels (several workgroups per CU)
constexpr size_t kBrickLdsBudget = 150 * 1024;  // spread/gather: one 1024-thread workgroup per CU (160 KB LDS)
constexpr int kMaxBins = 8192;                  // LDS histogram of the binning kernel (two counters per brick)
constexpr int kDefaultBrick = 16;               // grid points per brick edge (target)
constexpr int kDefaultFftTile = 16;             // lines per y/z FFT tile (target)

// bsp_mod[k] = 1/|sum_{j=0}^{n-2} arr[j] e^{2 pi i jk/n}|^2, FP64
// (dftmod + calc_prefac, CudaPMERecip.cu:2481-2505, 2544-2579)
void bsp_mod_host(int order, int n, double* out);
bool smooth235(int n);
// Stockham radices for length n (8, 4, 2, 5, 3), table stride tstride
bool radix_plan(int n, int tstride, RadixPlan* rp);
// Plan whose first stage has radix 16 (run on the matrix cores in the
// split-precision modes), the rest as radix_plan(n/16).  False if 16 does
// not divide n.
bool radix_plan_split(int n, int tstride, RadixPlan* rp);
// e^{-2 pi i k/n}, k < n, computed in FP64
std::vector<float2> twiddles(int n);

struct BrickPlan {
  int nb[3];
  int wmax[3];
  int nbins;
};
// bricks of about `target` points per edge, each at least order-1 wide, LDS
// footprints of spread and gather within `budget` bytes
bool plan_bricks(const int n[3], int order, int target, BrickPlan* bp, size_t budget = kBrickLdsBudget);
// The brick plan of a single-system solve on ncu CUs: of the plans for brick
// targets 16 to 32, the one with the least modelled spread time per CU
bool plan_bricks_auto(const int n[3], int order, int ncu, BrickPlan* bp);

struct FftTiles {
  int NR, TXy, TXz;
};
// z_four_step: the fused z pass runs as a register four-step (one LDS buffer)
FftTiles pick_fft_tiles(int nx, int ny, int nz, int nf_x, int tile_target, bool z_four_step);

struct ChunkPlan {
  int chunk, nchunk;
};
ChunkPlan plan_chunks(int natom);

}  // namespace pme
