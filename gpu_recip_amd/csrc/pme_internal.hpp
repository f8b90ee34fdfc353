// pme_internal.hpp -- launch interface between the host side (pme_host.cpp)
// and the kernels (pme_*.hip).  Not part of the public C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include "pme_device.hpp"

namespace pme {

// Spatial bins of atoms: one bin per grid brick, holding the atoms whose base
// grid index (the lowest stencil point) lies in that brick.
struct BinArgs {
  int nx, ny, nz;          // grid
  float r1, r2, r3;        // (float) recip diagonal, as the reference casts it
  int nbx, nby, nbz;       // bricks per axis
  int cap;                 // atom slots per bin
  int* bin_count;          // [nbins] live counters, zero between binnings
  int* bin_saved;          // [nbins] counts of the last binning
  int* bin_idx;            // [nbins * cap] atom indices
  int* ovf;                // [0] live overflow count, [1] saved overflow count
  int* ovf_idx;            // [ovf_cap] atoms that did not fit their bin
  int ovf_cap;
  unsigned int* ticket;    // last-workgroup ticket of the spread kernel
};

struct AxisArgs {
  RadixPlan rp;            // transform length and radices
  const float2* tw;        // e^{-2 pi i k/ntw}, k < ntw (FP64-computed, stored FP32)
  int ntw;                 // table length (the grid edge)
};

// Influence function + energy/virial (scalar_sum), CudaPMERecip.cu:594-901
struct SolveArgs {
  int n1, n2, n3;          // nfftx, nffty, nfftz
  int nf1, nf2, nf3;       // n/2 + n%2
  float r1, r2, r3;        // recip diagonal (float)
  float fac;               // pi^2/kappa^2
  float piv_inv;           // 1/(pi V)
  const float* pf1;        // prefac (bsp_mod) tables
  const float* pf2;
  const float* pf3;
  int calc_ev;             // accumulate energy and virial
  double* partial;         // [nblocks * 7] per-workgroup partial sums
  unsigned int* ticket;
  double* energy;          // device energy slot (added to)
  double* virmat;          // device virmat[9] (added to)
};

enum ColMode { COL_FWD = 0, COL_INV = 1, COL_FWD_SOLVE_INV = 2, COL_FWD_INV = 3 };

// ---- launchers (return the HIP error of the launch) ----
hipError_t launch_bin_atoms(const float4* xyzq, int n, const BinArgs& a, hipStream_t s);
hipError_t launch_bins_finalize(const BinArgs& a, hipStream_t s);
hipError_t launch_spread(int order, const float4* xyzq, const BinArgs& a, const int wmax[3],
                         float* grid, hipStream_t s);
hipError_t launch_gather(int order, int force_type, const float4* xyzq, const BinArgs& a,
                         const int wmax[3], const float* phi, int stride, void* force,
                         hipStream_t s);

hipError_t launch_fft_x_r2c(const float* grid, float2* cg, int nx, int nrows, int NR,
                            const AxisArgs& ax, hipStream_t s);
hipError_t launch_fft_x_c2r(const float2* cg, float* grid, int nx, int nrows, int NR,
                            const AxisArgs& ax, hipStream_t s);
// lines along one axis of the half spectrum: tile of TX consecutive kx, grid
// (ceil(hx/TX), nouter); element (line l, point j) at base + j*pt_stride + l,
// base = kx0 + outer*outer_stride.  For the z pass outer is ky (needed by
// COL_FWD_SOLVE_INV).
hipError_t launch_fft_cols(int mode, float2* cg, int n, int hx, int TX, long long pt_stride,
                           int nouter, long long outer_stride, const AxisArgs& ax,
                           const SolveArgs& sa, hipStream_t s);
hipError_t launch_scalar_sum(float2* cg, int hx, const SolveArgs& sa, hipStream_t s);

hipError_t launch_self_energy(const float4* xyzq, int n, double kappa_ccelec_sqrtpi, double* partial,
                              unsigned int* ticket, double* slot, hipStream_t s);
hipError_t launch_fill_zero(void* p, size_t bytes, hipStream_t s);
hipError_t launch_fill_bspline(int order, const float4* xyzq, int n, float r1, float r2, float r3,
                               int nx, int ny, int nz, int* gix, int* giy, int* giz, float* q,
                               float* thx, float* thy, float* thz, float* dthx, float* dthy,
                               float* dthz, hipStream_t s);

// LDS sizes (bytes) the launchers request, exposed for plan validation
size_t spread_lds_bytes(int order, const int wmax[3]);
size_t gather_lds_bytes(int order, const int wmax[3]);
size_t fft_x_lds_bytes(int nf, int NR, int ntw);
size_t fft_cols_lds_bytes(int n, int TX);

}  // namespace pme
