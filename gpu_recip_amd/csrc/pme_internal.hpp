// pme_internal.hpp -- launch interface between the host side (pme_host.cpp)
// and the kernels (pme_*.hip).  Not part of the public C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/pme_recip.h"
#include "pme_device.hpp"

namespace pme {

// Atoms sorted into grid bricks by a chunked counting sort.  Binning
// workgroup w takes atoms [w*chunk, (w+1)*chunk) and writes one record per
// brick an atom's stencil touches (1 to 8) into its own region of rec_stride
// slots, grouped by brick, the brick of the atom's base grid index (its home)
// first.  Per chunk the segment table holds seg_stride(nbins) ints: seg[b] =
// start of brick b's records in the region (exclusive prefix, seg[nbins] =
// total), then hend[b] = end of its home records (seg[b] <= hend[b] <=
// seg[b+1]).  rec_pos holds every record's (x, y, z, q); rec_idx the atom
// index of the home records only.  No global atomics: counts and cursors live
// in LDS.
__host__ __device__ constexpr int seg_stride(int nbins) { return 2 * nbins + 1; }
struct BinArgs {
  int nx, ny, nz;          // grid
  float r1, r2, r3;        // (float) recip diagonal, as the reference casts it
  float rxy, rxz, ryx, ryz, rzx, rzy;  // (float) off-diagonal recip[0][1], [0][2], [1][0], [1][2], [2][0], [2][1]
  int tric;                // 1: some off-diagonal entry is non-zero (triclinic box)
  int nbx, nby, nbz;       // bricks per axis
  int order;               // B-spline order (stencil width)
  int chunk;               // atoms per binning workgroup
  int nchunk;              // binning workgroups
  int rec_stride;          // record slots per chunk region (8 * chunk)
  int* seg_off;            // [nchunk][seg_stride(nbins)]
  float4* rec_pos;         // [nchunk * rec_stride] (x, y, z, q) copies
  int* rec_idx;            // [nchunk * rec_stride] atom index (home records)
};

// Grid indices and fractions of one atom on the three axes
// (CudaPMERecip.cu:512-529 per axis).  The fractional coordinate of axis a is
// row a of recip times (x, y, z): a triclinic recip first fuses the two
// off-diagonal terms of the row into the +2 shift (the larger axis innermost)
// and adds the diagonal term last, so a diagonal recip gives the reference's
// single-FMA formula exactly.
__device__ __forceinline__ void atom_grid_index(const BinArgs& a, const float4& p, int& ix, int& iy, int& iz,
                                                float& fx, float& fy, float& fz) {
#pragma clang fp contract(off)
  float bx = 2.0f, by = 2.0f, bz = 2.0f;
  if (a.tric) {
    bx = __builtin_fmaf(p.y, a.rxy, __builtin_fmaf(p.z, a.rxz, 2.0f));
    by = __builtin_fmaf(p.x, a.ryx, __builtin_fmaf(p.z, a.ryz, 2.0f));
    bz = __builtin_fmaf(p.x, a.rzx, __builtin_fmaf(p.y, a.rzy, 2.0f));
  }
  grid_index_at(p.x, a.r1, bx, a.nx, ix, fx);
  grid_index_at(p.y, a.r2, by, a.ny, iy, fy);
  grid_index_at(p.z, a.r3, bz, a.nz, iz, fz);
}

// Deferred energy/virial reduction: per-workgroup partials written by the
// scalar-sum and self-energy kernels are summed in fixed order by one
// workgroup of a later kernel (no contended atomics).
struct EvReduce {
  const double* partial;   // [nsolve * 7] scalar-sum partials
  int nsolve;              // 0: none pending
  double* energy;          // recip slot
  double* virmat;          // virmat[9]
  const double* self_partial;  // [nself]
  int nself;               // 0: none pending
  double* self_slot;
  double kcs;              // kappa*ccelec/sqrt(pi)
};

struct AxisArgs {
  RadixPlan rp;            // transform length and radices
  const float2* tw;        // e^{-2 pi i k/ntw}, k < ntw (FP64-computed, stored FP32)
  int ntw;                 // table length (the grid edge)
  int prec;                // PREC_*: a radix-16 first stage runs on the matrix cores
  int f4;                  // z pass: four-step factor pair (four_step_code), -1: Stockham stages
  int f4col;               // y passes: four-step pair of the one-buffer column pass, -1: generic stages
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
  float rt[9];             // (float) recip, row-major, for a triclinic box
  int tric;                // 1: m-hat from the full recip (stand-alone scalar sum only)
};

enum ColMode { COL_FWD = 0, COL_INV = 1, COL_FWD_SOLVE_INV = 2, COL_FWD_INV = 3 };

// ---- batched path: many independent systems per launch ----------------------
// Everything the stage bodies need for one system, resident on the device.
struct SysDesc {
  BinArgs bin;             // grid, recip, bricks and record store
  int wmax[3];             // largest brick per axis
  const float4* xyzq;      // caller's coordinates (x, y, z, q)
  int natom;
  float* grid;             // real grid
  float2* cgrid;           // half spectrum
  AxisArgs ax[3];          // FFT plans and twiddles
  int NR, TXy, TXz;        // FFT tiles
  SolveArgs sa;            // influence function; sa.partial = this system's partials
  int nzblk;               // z-pass tiles (= partial count)
  int plane;               // 1: x and y passes as one xy plane kernel
  void* force;             // caller's force output
  int force_type, stride;
  double* ev_out;          // [kEvOut]: energy, virial[9], self energy of the last solve (assigned)
  double* self_partial;    // [bin.nchunk]: sum of q*q per binning chunk (calc_self)
  int calc_self;           // 1: self energy (CudaPMERecip.cu:2053-2061)
  double kcs;              // kappa*ccelec/sqrt(pi)
};
constexpr int kEvOut = 11;  // doubles per system in the batched energy/virial output

// One workgroup-sized task of a batched stage: system index and the tile of
// that system (chunk, brick, row tile or column tile).
struct WorkItem {
  int sys;
  int idx;
};

// Work distribution of the batched spread and gather kernels (one brick per
// item, one workgroup per CU at a time).  Workgroup b starts with item b.
// With fewer workgroups than items (persistent launches) a workgroup that
// finishes a brick takes the next one from the launch's counter ctr[0], so
// the bricks balance as under the hardware dispatcher.  The fetch comes after
// the brick's stores are issued and before the next brick's loads, since a
// wait for those loads would also wait for an earlier fetch.  The last
// workgroup to finish (counted in ctr[1]) puts both counters back to zero for
// the next launch that uses them (stream order makes that launch see the
// zeros).  One lane per workgroup touches the counters, with vector atomics.
// The other batched kernels stride statically: their items are many and
// short, and one counter shared by every XCD would serialize them.
__device__ __forceinline__ int next_item(int nitems, int* ctr) {
  if ((int)gridDim.x >= nitems) return nitems;
  __shared__ int s_next;
  __syncthreads();
  if (threadIdx.x == 0) s_next = (int)gridDim.x + atomicAdd(&ctr[0], 1);
  __syncthreads();
  return s_next;
}
__device__ __forceinline__ void items_done(int nitems, int* ctr) {
  if ((int)gridDim.x >= nitems) return;
  if (threadIdx.x == 0 && atomicAdd(&ctr[1], 1) == (int)gridDim.x - 1) {
    atomicExch(&ctr[0], 0);
    atomicExch(&ctr[1], 0);
  }
}

// Work items are the same for every lane of a workgroup: make that explicit
// so descriptor fields are read with scalar loads and branch uniformly.
__device__ __forceinline__ WorkItem uniform_item(WorkItem w) {
  w.sys = __builtin_amdgcn_readfirstlane(w.sys);
  w.idx = __builtin_amdgcn_readfirstlane(w.idx);
  return w;
}

// Batched stages; each has its own work-item list.
enum BatchStage { BS_BIN = 0, BS_SPREAD, BS_FFTX, BS_FFTY, BS_FFTZ, BS_GATHER, BS_PLANE, BS_COUNT };

// Segments of brick b in the chunked record store: s_base[w] = first record
// of b in chunk w's region, s_pre[w] = exclusive prefix of the counts over w
// (s_pre[nchunk] = total).  Returns the total.  NT == blockDim.x.
// The segment bounds of chunk w = threadIdx.x (every chunk has a thread:
// nchunk <= kMaxChunks <= NT), loaded into registers so that a caller can
// issue the loads first and overlap their latency with other work.  HOME:
// the home records of the brick only (its gather), else all its records.
struct SegBounds {
  int lo, hi;
};
template <bool HOME>
__device__ __forceinline__ SegBounds brick_seg_load(const BinArgs& a, int b) {
  SegBounds s{0, 0};
  const int w = threadIdx.x;
  if (w < a.nchunk) {
    const int nbins = a.nbx * a.nby * a.nbz;
    const int* seg = a.seg_off + (size_t)w * seg_stride(nbins);
    s.lo = seg[b];
    s.hi = HOME ? seg[nbins + 1 + b] : seg[b + 1];
  }
  return s;
}
template <int NT>
__device__ __forceinline__ int brick_segments(const BinArgs& a, const SegBounds& sb, int* s_pre, int* s_base,
                                              int* sh_scan) {
  if (a.nchunk <= 64) {
    // every chunk's bounds are in wave 0's registers: one wave-level scan and
    // one barrier
    if (threadIdx.x < 64) {
      const int lane = threadIdx.x;
      const int cnt = lane < a.nchunk ? sb.hi - sb.lo : 0;
      int v = cnt;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
      }
      if (lane < a.nchunk) {
        s_base[lane] = sb.lo;
        s_pre[lane] = v - cnt;
      }
      if (lane == 63) s_pre[a.nchunk] = v;  // lanes past nchunk add nothing
    }
    __syncthreads();
    return s_pre[a.nchunk];
  }
  if ((int)threadIdx.x < a.nchunk) {
    s_base[threadIdx.x] = sb.lo;
    s_pre[threadIdx.x] = sb.hi - sb.lo;
  }
  __syncthreads();
  const int total = block_exclusive_scan<NT>(s_pre, a.nchunk, sh_scan);
  if (threadIdx.x == 0) s_pre[a.nchunk] = total;
  __syncthreads();
  return total;
}

// record slot of the v-th record of the brick (0 <= v < total)
__device__ __forceinline__ size_t brick_record(const BinArgs& a, int v, const int* s_pre, const int* s_base) {
  int lo = 0, hi = a.nchunk;  // s_pre[lo] <= v < s_pre[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (s_pre[mid] <= v) lo = mid;
    else hi = mid;
  }
  return (size_t)lo * a.rec_stride + s_base[lo] + (v - s_pre[lo]);
}

// ---- launchers (return the HIP error of the launch) ----
hipError_t launch_bin_atoms(const float4* xyzq, int n, const BinArgs& a, hipStream_t s);
hipError_t launch_spread(int order, const BinArgs& a, const int wmax[3], float* grid, hipStream_t s);
hipError_t launch_gather(int order, int force_type, const float4* xyzq, const BinArgs& a,
                         const int wmax[3], const float* phi, int stride, void* force,
                         hipStream_t s);

// The half spectrum holds nz planes of pitch spec_pitch(hx, ny) complex
// values, each plane ny rows of hx = nx/2 + 1 (kx fastest).
hipError_t launch_fft_x_r2c(const float* grid, float2* cg, int nx, int ny, int nz, int NR,
                            const AxisArgs& ax, hipStream_t s);
hipError_t launch_fft_x_c2r(const float2* cg, float* grid, int nx, int ny, int nz, int NR,
                            const AxisArgs& ax, hipStream_t s);
// lines along one axis of the half spectrum: tile of TX consecutive lines,
// grid (ceil(hx/TX), nouter); element (line l, point j) at base + j*pt_stride
// + l, base = l0 + outer*outer_stride.  The y pass: hx = nx/2 + 1 lines per
// kz (outer).  The z pass: hx = (nx/2 + 1) ny lines, the flattened (ky, kx)
// of a plane, nouter = 1, hxd = nx/2 + 1 to recover kx and ky (needed by
// COL_FWD_SOLVE_INV, which writes per-workgroup energy/virial partials).
// Workgroup (0,0) also performs a pending energy/virial reduction if red
// has one.
hipError_t launch_fft_cols(int mode, float2* cg, int n, int hx, int TX, long long pt_stride,
                           int nouter, long long outer_stride, const AxisArgs& ax,
                           const SolveArgs& sa, const EvReduce& red, int hxd, hipStream_t s);
int cols_blocks(int hx, int TX, int nouter);
hipError_t launch_scalar_sum(float2* cg, int hx, const SolveArgs& sa, hipStream_t s);
int scalar_sum_blocks(int total);

hipError_t launch_self_energy(const float4* xyzq, int n, double* partial, int* nblocks, hipStream_t s);
hipError_t launch_ev_reduce(const EvReduce& red, hipStream_t s);
hipError_t launch_fill_zero(void* p, size_t bytes, hipStream_t s);
hipError_t launch_fill_bspline(int order, const float4* xyzq, int n, float r1, float r2, float r3,
                               int nx, int ny, int nz, int* gix, int* giy, int* giz, float* q,
                               float* thx, float* thy, float* thz, float* dthx, float* dthy,
                               float* dthz, hipStream_t s);

// Sum the pending partials in fixed order (one workgroup) and add them into
// the energy/virial buffer (CudaPMERecip.cu:871-886 for the scalar sum,
// :1272-1274 for the self energy).
template <int NT = kBlock>
__device__ __forceinline__ void ev_reduce_block(const EvReduce& r) {
  __shared__ double sh[7 * (NT / 64)];
  if (r.nsolve > 0) {
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < r.nsolve; i += NT)
#pragma unroll
      for (int c = 0; c < 7; c++) v[c] += r.partial[(size_t)i * 7 + c];
    block_sum<7, NT>(v, sh);
    if (threadIdx.x == 0) {
      const double hc = 0.5 * kCcelec;
      r.energy[0] += v[0] * hc;
      const double v0 = -hc * v[1], v1 = -hc * v[2], v2 = -hc * v[3];
      const double v3 = -hc * v[4], v4 = -hc * v[5], v5 = -hc * v[6];
      r.virmat[0] += v0; r.virmat[1] += v1; r.virmat[2] += v2;
      r.virmat[3] += v1; r.virmat[4] += v3; r.virmat[5] += v4;
      r.virmat[6] += v2; r.virmat[7] += v4; r.virmat[8] += v5;
    }
  }
  if (r.nself > 0) {
    double v[1] = {0.0};
    for (int i = threadIdx.x; i < r.nself; i += NT) v[0] += r.self_partial[i];
    block_sum<1, NT>(v, sh);
    if (threadIdx.x == 0) r.self_slot[0] += -v[0] * r.kcs;
  }
}

// xy plane kernels (x and y passes of one z plane in one workgroup):
// forward reads the real plane and writes its half-spectrum plane, inverse
// the reverse; a pending energy/virial reduction runs in one extra
// workgroup of the inverse launch.
hipError_t launch_fft_plane_fwd(const float* grid, float2* cg, int nx, int ny, int nz, const AxisArgs& ax,
                                const AxisArgs& ay, hipStream_t s);
hipError_t launch_fft_plane_inv(const float2* cg, float* grid, int nx, int ny, int nz, const AxisArgs& ax,
                                const AxisArgs& ay, const EvReduce& red, hipStream_t s);
// Factor pair code of the register four-step z pass for a z length n
// (-1: none, the Stockham stages are used)
int four_step_code(int n);
// the factor pair (n1, n2) of a four_step_code
bool four_step_pair(int code, int* n1, int* n2);
// four_step_code of a radix plan's four-step pair when the table holds that
// pair, else -1 (the y passes then run the generic stages)
int four_step_code_for_plan(const RadixPlan& rp);
// register class of a four-step pair (0: N2 <= 10 and 5 x 15, 1: the others), -1: none
int four_step_class(int code);

// LDS bytes of a plane kernel over np planes per workgroup (0 when they do
// not fit one workgroup)
size_t fft_plane_lds_bytes(int nx, int ny, int nf_x, int np);
// planes per workgroup of a system in a batched plane launch allowing up to
// maxnp (the largest group that fits the LDS budget of a full-size or, with
// half, a half-size workgroup)
int plane_group_host(int nx, int ny, int nf_x, int maxnp, bool half);
// one plane fits the LDS budget of a half-size plane workgroup
bool plane_fits_half(int nx, int ny, int nf_x);
// batched plane kernels: items are (system, first z plane), with
// plane_group_host(.., maxnp, half) planes each; with reduce != 0 the plane-0
// item of every system assigns that system's energy/virial.  split != 0: the
// kernel with the matrix-core y stage (systems whose y axis is not FP32);
// split == 0 requires FP32 y axes.  half != 0: kPlaneBlock / 2 threads per
// workgroup (two workgroups per CU), for systems with plane_fits_half only.
hipError_t launch_batch_plane(int inverse, int reduce, int maxnp, int split, int half, const SysDesc* sys,
                              const WorkItem* items, int nitems, int grid, size_t lds, hipStream_t s);

// batched launchers: `grid` persistent workgroups loop over the work items
hipError_t launch_batch_bin(const SysDesc* sys, const WorkItem* items, int nitems, int grid, size_t lds, hipStream_t s);
// spread / gather: half != 0 runs kSpreadBlock / 2 threads per workgroup
// (bricks planned within half the LDS budget, two workgroups per CU)
hipError_t launch_batch_spread(int order, int half, const SysDesc* sys, const WorkItem* items, int nitems, int* ctr,
                               int grid, size_t lds, hipStream_t s);
hipError_t launch_batch_gather(int order, int half, const SysDesc* sys, const WorkItem* items, int nitems, int* ctr,
                               int grid, size_t lds, hipStream_t s);
hipError_t launch_batch_fft_x(int inverse, const SysDesc* sys, const WorkItem* items, int nitems, int grid, size_t lds,
                              hipStream_t s);
// axis 1 (y) or 2 (z); mode COL_*; with reduce != 0 the first tile of every
// system sums that system's energy/virial partials into ev_out; split as
// for launch_batch_plane, for the transformed axis
// fused z pass of the systems whose z length has a four-step pair of class cls
hipError_t launch_batch_zsolve(int cls, const SysDesc* sys, const WorkItem* items, int nitems, int grid, size_t lds,
                               hipStream_t s);
hipError_t launch_batch_cols(int axis, int mode, int reduce, int split, const SysDesc* sys, const WorkItem* items,
                             int nitems, int grid, size_t lds, hipStream_t s);

// LDS sizes (bytes) the launchers request, exposed for plan validation
size_t bin_lds_bytes(int nbins);
size_t spread_lds_bytes(int order, const int wmax[3], int nchunk);
size_t gather_lds_bytes(int order, const int wmax[3], int nchunk);
size_t fft_x_lds_bytes(int nf, int NR, int ntw);
// column passes: twiddles and nbuf n*TX buffers (two for the Stockham stages,
// one for the fused z pass as a register four-step)
size_t fft_cols_lds_bytes(int n, int TX, int nbuf = 2);

constexpr int kSpreadBlock = 1024;  // threads per spread / gather workgroup (16 waves per CU)
constexpr int kRecPerThread = 2;    // gather: brick records read per thread per round
constexpr int kGatherStageU = 24;   // gather: halo loads in flight per thread (one round up to 24 K points)
constexpr int kBinBlock = 512;     // threads per binning workgroup (one atom each at the default chunk)
constexpr int kPlaneBlock = 1024;   // threads per xy plane FFT workgroup
constexpr int kPlaneGroupMax = 4;   // z planes per batched plane workgroup (at most)
constexpr size_t kPlaneLdsBudget = 150 * 1024;  // the plane kernels run one workgroup per CU at most
constexpr int kMaxChunks = 512;    // binning workgroups per solve (chunk grows with n)
static_assert(kMaxChunks <= kSpreadBlock, "brick_segments: one thread per chunk");
constexpr int kMinChunk = 512;    // atoms per binning workgroup


// ---- direct-space Ewald (pme_direct.hip) ----
// The settings of the Ewald pair term (DirectSettings_t fields it reads,
// CudaDirectForceTypes.h:23-67)
struct DirectParams {
  float boxx, boxy, boxz;
  float kappa, kappa2;
  float roff2;
  float e14fac;
};
constexpr int kDirWaves = 4;         // waves per non-bonded workgroup
constexpr int kDirChunk = 8;         // j tiles per chunk of the pair kernel (kDirChunk / 2 per half-wave)
constexpr int kDir14Block = 256;     // 1-4 entries per workgroup
constexpr int kDirPartial = 2 + 81;  // doubles per workgroup partial: elec, excl, sforce[27][3]
hipError_t launch_direct_force(const DirectParams& d, int n_ientry, int n_tile, int nwave,
                               const pme_ientry* ientry, int* tile_entry, const int* tile_indj,
                               const pme_tile_excl* tile_excl, const float4* xyzq, int stride,
                               unsigned long long* force, bool energy, bool virial, double* partial,
                               hipStream_t s);
hipError_t launch_direct_14(const DirectParams& d, int nin14, int nex14, const pme_xx14* in14, const pme_xx14* ex14,
                            const float4* xyzq, int stride, unsigned long long* force, bool energy, bool virial,
                            double* partial, hipStream_t s);
// adds the nblk partials (fixed order) to the elec / excl terms and sforce_dp[27][3]
hipError_t launch_direct_reduce(const double* partial, int nblk, bool energy, bool virial, double* e_elec,
                                double* e_excl, double* sforce_dp, hipStream_t s);
int direct_force_blocks(int nwave);
int direct_force_waves(int n_tile, int ncu);
int direct_14_blocks(int n);

}  // namespace pme
