// pme_gather.hip -- force gather (CudaPMERecip.cu:1031-1245 semantics) for CDNA4.
//
// One workgroup per grid brick.  The workgroup stages the potential of its
// brick plus a (p-1) halo on the upper side in LDS with coalesced loads.  It
// then takes the atoms the binning put first in the brick's record segments
// (its home atoms), loads their CURRENT coordinates (so the result is right
// even if the caller moved atoms after spread_charge) and gathers each atom
// in one thread: B-spline weights and derivatives in registers, the p x p
// stencil rows read from LDS.  An atom whose current base index is no longer
// in this brick reads the potential from HBM.
#include "pme_internal.hpp"

namespace pme {

__device__ __forceinline__ void store_force(int ft, void* force, int atom, int stride, float fx, float fy, float fz) {
  if (ft == 0) {  // float: strided SoA store, CudaPMERecip.cu:1005-1012
    float* f = reinterpret_cast<float*>(force);
    f[atom] = fx;
    f[(size_t)atom + stride] = fy;
    f[(size_t)atom + 2 * (size_t)stride] = fz;
  } else if (ft == 1) {  // float3: AoS store, :1016-1023
    float* f = reinterpret_cast<float*>(force) + 3 * (size_t)atom;
    f[0] = fx;
    f[1] = fy;
    f[2] = fz;
  } else {  // long long: fixed point 2^40 added atomically, :992-1001
    unsigned long long* f = reinterpret_cast<unsigned long long*>(force);
    const long long ix = (long long)rintf(kForceScale * fx);
    const long long iy = (long long)rintf(kForceScale * fy);
    const long long iz = (long long)rintf(kForceScale * fz);
    atomicAdd(&f[atom], (unsigned long long)ix);
    atomicAdd(&f[(size_t)atom + stride], (unsigned long long)iy);
    atomicAdd(&f[(size_t)atom + 2 * (size_t)stride], (unsigned long long)iz);
  }
}

// Gradient sums of one atom from the potential: with w(j) = theta, w'(j) =
// dtheta per axis,
//   f1 = sum dthx thy thz phi,  f2 = sum thx dthy thz phi,  f3 = sum thx thy dthz phi,
// evaluated row by row: a = sum_x thx phi, b = sum_x dthx phi per (y, z) row,
// then y sums per z plane.  at(jx, jy, jz) returns phi at the stencil point.
template <int P, typename At>
__device__ __forceinline__ void gather_sums(const float* tx, const float* dtx, const float* ty, const float* dty,
                                            const float* tz, const float* dtz, At at, float& f1, float& f2,
                                            float& f3) {
  f1 = f2 = f3 = 0.f;
#pragma unroll
  for (int jz = 0; jz < P; jz++) {
    float ay = 0.f, by = 0.f, cy = 0.f;
#pragma unroll
    for (int jy = 0; jy < P; jy++) {
      float ra = 0.f, rb = 0.f;
#pragma unroll
      for (int jx = 0; jx < P; jx++) {
        const float v = at(jx, jy, jz);
        ra = fmaf(tx[jx], v, ra);
        rb = fmaf(dtx[jx], v, rb);
      }
      ay = fmaf(ty[jy], ra, ay);
      by = fmaf(ty[jy], rb, by);
      cy = fmaf(dty[jy], ra, cy);
    }
    f1 = fmaf(tz[jz], by, f1);
    f2 = fmaf(tz[jz], cy, f2);
    f3 = fmaf(dtz[jz], ay, f3);
  }
}

// One atom's gather: B-spline weights and derivatives in registers, its
// p x p stencil rows read from the brick's LDS potential (pl, [hz][hy][hx]),
// or from HBM when the atom's current base index is no longer in the brick
// (coordinates changed since the binning).
template <int P>
__device__ __forceinline__ void gather_atom(const BinArgs& a, const float4& p, int at, int x0, int y0, int z0, int wx,
                                            int wy, int wz, int hx, int hy, const float* __restrict__ pl,
                                            const float* __restrict__ phi, float ccelec_f, int ft, int stride,
                                            void* force) {
  int ix, iy, iz;
  float fx, fy, fz;
  atom_grid_index(a, p, ix, iy, iz, fx, fy, fz);
  ix = safe_index(ix, a.nx);
  iy = safe_index(iy, a.ny);
  iz = safe_index(iz, a.nz);
  float tx[P], dtx[P], ty[P], dty[P], tz[P], dtz[P];
  bspline_theta_dtheta<P>(fx, tx, dtx);
  bspline_theta_dtheta<P>(fy, ty, dty);
  bspline_theta_dtheta<P>(fz, tz, dtz);
  int dx = ix - x0, dy = iy - y0, dz = iz - z0;
  if (dx >= wx) dx -= a.nx;
  if (dy >= wy) dy -= a.ny;
  if (dz >= wz) dz -= a.nz;
  float f1, f2, f3;
  if (dx >= 0 && dy >= 0 && dz >= 0) {
    const float* col = pl + mul24(mul24(dz, hy) + dy, hx) + dx;
    const int rs = hx, ps = mul24(hy, hx);
    gather_sums<P>(tx, dtx, ty, dty, tz, dtz, [&](int jx, int jy, int jz) { return col[jz * ps + jy * rs + jx]; }, f1,
                   f2, f3);
  } else {
    gather_sums<P>(tx, dtx, ty, dty, tz, dtz,
                   [&](int jx, int jy, int jz) {
                     int gx = ix + jx, gy = iy + jy, gz = iz + jz;
                     if (gx >= a.nx) gx -= a.nx;
                     if (gy >= a.ny) gy -= a.ny;
                     if (gz >= a.nz) gz -= a.nz;
                     return phi[(size_t)gx + (size_t)a.nx * ((size_t)gy + (size_t)a.ny * gz)];
                   },
                   f1, f2, f3);
  }
  // CudaPMERecip.cu:1231-1241: q*ccelec, then q*recip*f*nfft in FP32; a
  // triclinic box: dE/dr_b = sum_a (q f_a n_a) recip[a][b]
  const float q = p.w * ccelec_f;
  if (!a.tric) {
    store_force(ft, force, at, stride, q * a.r1 * f1 * (float)a.nx, q * a.r2 * f2 * (float)a.ny,
                q * a.r3 * f3 * (float)a.nz);
  } else {
    const float g1 = q * f1 * (float)a.nx, g2 = q * f2 * (float)a.ny, g3 = q * f3 * (float)a.nz;
    store_force(ft, force, at, stride, fmaf(g3, a.rzx, fmaf(g2, a.ryx, g1 * a.r1)),
                fmaf(g3, a.rzy, fmaf(g2, a.r2, g1 * a.rxy)), fmaf(g3, a.r3, fmaf(g2, a.ryz, g1 * a.rxz)));
  }
}

// Brick b of one grid (one NT-thread workgroup); smem: the
// workgroup's dynamic LDS (gather_lds_bytes).  ft: PME_FORCE_* store mode.
// The brick's potential plus a (p-1) upper halo is staged in LDS; the home
// records of the brick (the atoms whose base grid index it holds, first in
// each of its record segments) are then gathered one atom per thread.
// FROM_REC: the coordinates are the binning's copy in the record store
// (contiguous per brick) instead of xyzq[atom] (scattered); only for solves
// that bin and gather the same coordinates (the batched path: one solve reads
// the caller's array once).  Otherwise the current coordinates are read, so
// an atom moved since the binning is still right (gather_atom).
template <int P, bool FROM_REC, int NT>
__device__ __forceinline__ void gather_brick_body(const float4* __restrict__ xyzq, const BinArgs& a, int b, int wxm,
                                                  int wym, int wzm, const float* __restrict__ phi, float ccelec_f,
                                                  int ft, int stride, void* force, unsigned char* smem) {
  __shared__ int sh_scan[NT / 64];
  const int hmax = (wxm + P - 1) * (wym + P - 1) * (wzm + P - 1);
  float* pl = reinterpret_cast<float*>(smem);                 // halo brick, [hz][hy][hx]
  int* s_pre = reinterpret_cast<int*>(pl + ((hmax + 3) & ~3));  // [nchunk + 1]
  int* s_base = s_pre + a.nchunk + 1;                           // [nchunk]

  const int tid = threadIdx.x;
  const int bxi = b % a.nbx, byi = (b / a.nbx) % a.nby, bzi = b / (a.nbx * a.nby);
  const int x0 = brick_lo(bxi, a.nx, a.nbx), wx = brick_lo(bxi + 1, a.nx, a.nbx) - x0;
  const int y0 = brick_lo(byi, a.ny, a.nby), wy = brick_lo(byi + 1, a.ny, a.nby) - y0;
  const int z0 = brick_lo(bzi, a.nz, a.nbz), wz = brick_lo(bzi + 1, a.nz, a.nbz) - z0;

  // the home-record segment bounds are in flight while the halo is staged
  const SegBounds sb = brick_seg_load<true>(a, b);

  // stage brick + upper halo (periodic) in LDS in this brick's own [hz][hy][hx]
  // layout: thread c takes column (lx, ly) = (c mod hx, c / hx) of the halo
  // brick and walks it in z, kGatherStageU loads in flight (one division per
  // thread; consecutive threads read consecutive x, as rows of the grid)
  const int hx = wx + P - 1, hy = wy + P - 1, hz = wz + P - 1;
  const int ncol = hx * hy;
  const int nxny = mul24(a.nx, a.ny);  // grid offsets fit 32 bits (edges <= 512)
  for (int c = tid; c < ncol; c += NT) {
    const int ly = c / hx;
    const int lx = c - ly * hx;
    int gx = x0 + lx, gy = y0 + ly;
    if (gx >= a.nx) gx -= a.nx;
    if (gy >= a.ny) gy -= a.ny;
    const float* col = phi + (gx + mul24(a.nx, gy));
    for (int lz0 = 0; lz0 < hz; lz0 += kGatherStageU) {
      float v[kGatherStageU];
#pragma unroll
      for (int u = 0; u < kGatherStageU; u++) {
        // z0 + lz < nz + P - 1 < 2 nz for lz < hz: one wrap; 0 past the halo
        const int lz = lz0 + u;
        int gz = z0 + lz;
        if (gz >= a.nz) gz -= a.nz;
        v[u] = lz < hz ? col[mul24(nxny, lz < hz ? gz : 0)] : 0.f;
      }
#pragma unroll
      for (int u = 0; u < kGatherStageU; u++)
        if (lz0 + u < hz) pl[mul24(lz0 + u, ncol) + c] = v[u];
    }
  }
  const int total = brick_segments<NT>(a, sb, s_pre, s_base, sh_scan);  // ends with a barrier

  // up to kRecPerThread home records per thread per round, every load of the
  // round issued before the first is used
  for (int base = 0; base < total; base += kRecPerThread * NT) {
    int atom[kRecPerThread];
    float4 pos[kRecPerThread];
#pragma unroll
    for (int r = 0; r < kRecPerThread; r++) {
      const int v = base + r * NT + tid;
      atom[r] = -1;
      pos[r] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (v < total) {
        const size_t slot = brick_record(a, v, s_pre, s_base);
        atom[r] = a.rec_idx[slot];
        if (FROM_REC) pos[r] = a.rec_pos[slot];
      }
    }
    if (!FROM_REC) {
#pragma unroll
      for (int r = 0; r < kRecPerThread; r++)
        if (atom[r] >= 0) pos[r] = xyzq[atom[r]];
    }
#pragma unroll
    for (int r = 0; r < kRecPerThread; r++)
      if (atom[r] >= 0)
        gather_atom<P>(a, pos[r], atom[r], x0, y0, z0, wx, wy, wz, hx, hy, pl, phi, ccelec_f, ft, stride, force);
  }
}

template <int P>
__global__ __launch_bounds__(kSpreadBlock) void k_gather_brick(const float4* __restrict__ xyzq, BinArgs a, int wxm,
                                                               int wym, int wzm, const float* __restrict__ phi,
                                                               float ccelec_f, int ft, int stride, void* force) {
  extern __shared__ __align__(16) unsigned char smem[];
  gather_brick_body<P, false, kSpreadBlock>(xyzq, a, blockIdx.x, wxm, wym, wzm, phi, ccelec_f, ft, stride, force,
                                            smem);
}

size_t gather_lds_bytes(int order, const int wmax[3], int nchunk) {
  const size_t halo = (size_t)(wmax[0] + order - 1) * (wmax[1] + order - 1) * (wmax[2] + order - 1);
  return sizeof(float) * ((halo + 3) & ~(size_t)3) + sizeof(int) * (2 * (size_t)nchunk + 1);
}

template <int P>
static hipError_t launch_gather_p(int ft, const float4* xyzq, const BinArgs& a, const int wmax[3], const float* phi,
                                  int stride, void* force, hipStream_t s) {
  const int nbins = a.nbx * a.nby * a.nbz;
  const size_t lds = gather_lds_bytes(P, wmax, a.nchunk);
  hipLaunchKernelGGL(k_gather_brick<P>, dim3(nbins), dim3(kSpreadBlock), lds, s, xyzq, a, wmax[0], wmax[1], wmax[2], phi,
                     (float)kCcelec, ft, stride, force);
  return hipGetLastError();
}

hipError_t launch_gather(int order, int force_type, const float4* xyzq, const BinArgs& a, const int wmax[3],
                         const float* phi, int stride, void* force, hipStream_t s) {
  switch (order) {
    case 4: return launch_gather_p<4>(force_type, xyzq, a, wmax, phi, stride, force, s);
    case 6: return launch_gather_p<6>(force_type, xyzq, a, wmax, phi, stride, force, s);
    case 8: return launch_gather_p<8>(force_type, xyzq, a, wmax, phi, stride, force, s);
    default: return hipErrorInvalidValue;
  }
}

// ---- batched: persistent workgroups over the work items of many systems ----

// NT = kSpreadBlock / 2: half-size workgroups, two per CU, for bricks
// planned within half the LDS budget (four waves per SIMD: at most 128 VGPRs)
template <int P, int NT>
__global__ __launch_bounds__(NT, 4) void k_batch_gather(const SysDesc* __restrict__ sys,
                                                     const WorkItem* __restrict__ items, int nitems, int* ctr) {
  extern __shared__ __align__(16) unsigned char smem[];
  for (int it = blockIdx.x; it < nitems; it = next_item(nitems, ctr)) {
    const WorkItem w = uniform_item(items[it]);
    const SysDesc& d = sys[w.sys];
    gather_brick_body<P, true, NT>(d.xyzq, d.bin, w.idx, d.wmax[0], d.wmax[1], d.wmax[2], d.grid, (float)kCcelec, d.force_type,
                         d.stride, d.force, smem);
    __syncthreads();
  }
  items_done(nitems, ctr);
}

hipError_t launch_batch_gather(int order, int half, const SysDesc* sys, const WorkItem* items, int nitems, int* ctr,
                               int grid, size_t lds, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  constexpr int H = kSpreadBlock / 2;
#define PME_BGATHER(P)                                                                                           \
  if (half)                                                                                                      \
    hipLaunchKernelGGL((k_batch_gather<P, H>), dim3(grid), dim3(H), lds, s, sys, items, nitems, ctr);            \
  else                                                                                                           \
    hipLaunchKernelGGL((k_batch_gather<P, kSpreadBlock>), dim3(grid), dim3(kSpreadBlock), lds, s, sys, items,    \
                       nitems, ctr);
  switch (order) {
    case 4: PME_BGATHER(4) break;
    case 6: PME_BGATHER(6) break;
    case 8: PME_BGATHER(8) break;
    default: return hipErrorInvalidValue;
  }
#undef PME_BGATHER
  return hipGetLastError();
}

}  // namespace pme
