// pme_gather.hip -- force gather (CudaPMERecip.cu:1031-1245 semantics) for CDNA4.
//
// One workgroup per grid brick.  The workgroup stages the potential of its
// brick plus a (p-1) halo on the upper side in LDS with coalesced loads.  It
// then takes the atoms whose home copy the binning put in this brick,
// compacts them with a wave ballot, loads their CURRENT coordinates (so the
// result is right even if the caller moved atoms after spread_charge) and
// gathers with G lanes per atom (G = 4 for order 4, 8 for orders 6 and 8):
// lane g owns stencil column x = ix+g and walks the p*p (y, z) rows out of
// LDS; the G partial sums are combined with wave shuffles.  An atom whose
// current base index is no longer in this brick reads the potential from HBM.
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

// Brick b of one grid (one kSpreadBlock-thread workgroup); smem: the
// workgroup's dynamic LDS (gather_lds_bytes).  ft: PME_FORCE_* store mode.
template <int P>
__device__ __forceinline__ void gather_brick_body(const float4* __restrict__ xyzq, const BinArgs& a, int b, int wxm,
                                                  int wym, int wzm, const float* __restrict__ phi, float ccelec_f,
                                                  int ft, int stride, void* force, unsigned char* smem) {
  constexpr int G = (P <= 4) ? 4 : 8;
  constexpr int NT = 6 * P;  // theta x/y/z and dtheta x/y/z per atom
  __shared__ int sh_scan[kSpreadBlock / 64];
  __shared__ int s_cnt;
  const int hxs = wxm + P - 1, hys = wym + P - 1, hzs = wzm + P - 1;
  float* pl = reinterpret_cast<float*>(smem);                               // halo brick
  float* s_th = pl + ((hxs * hys * hzs + 3) & ~3);                          // [kGatherBatch][NT]
  int4* s_meta = reinterpret_cast<int4*>(s_th + kGatherBatch * NT);         // [kGatherBatch] (ix, iy, iz, atom)
  float* s_q = reinterpret_cast<float*>(s_meta + kGatherBatch);             // [kGatherBatch]
  int* s_pre = reinterpret_cast<int*>(s_q + kGatherBatch);                  // [nchunk + 1]
  int* s_base = s_pre + a.nchunk + 1;                                       // [nchunk]

  const int tid = threadIdx.x;
  const int bxi = b % a.nbx, byi = (b / a.nbx) % a.nby, bzi = b / (a.nbx * a.nby);
  const int x0 = brick_lo(bxi, a.nx, a.nbx), wx = brick_lo(bxi + 1, a.nx, a.nbx) - x0;
  const int y0 = brick_lo(byi, a.ny, a.nby), wy = brick_lo(byi + 1, a.ny, a.nby) - y0;
  const int z0 = brick_lo(bzi, a.nz, a.nbz), wz = brick_lo(bzi + 1, a.nz, a.nbz) - z0;

  // stage brick + upper halo (periodic) in LDS, kStageU loads in flight
  const int hx = wx + P - 1, hy = wy + P - 1, hz = wz + P - 1;
  const int nh = hx * hy * hz;
  for (int i0 = tid; i0 < nh; i0 += kStageU * kSpreadBlock) {
    float v[kStageU];
    int dst[kStageU];
#pragma unroll
    for (int u = 0; u < kStageU; u++) {
      const int i = i0 + u * kSpreadBlock;
      const int t = i / hx;
      const int lx = i - t * hx;
      const int lz = t / hy;
      const int ly = t - lz * hy;
      int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
      if (gx >= a.nx) gx -= a.nx;
      if (gy >= a.ny) gy -= a.ny;
      if (gz >= a.nz) gz -= a.nz;
      dst[u] = (lz * hys + ly) * hxs + lx;
      v[u] = (i < nh) ? phi[(size_t)gx + (size_t)a.nx * ((size_t)gy + (size_t)a.ny * gz)] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < kStageU; u++)
      if (i0 + u * kSpreadBlock < nh) pl[dst[u]] = v[u];
  }
  const int total = brick_segments<kSpreadBlock>(a, b, s_pre, s_base, sh_scan);  // ends with a barrier

  const int g = tid % G;
  for (int base = 0; base < total; base += kGatherBatch) {
    if (tid == 0) s_cnt = 0;
    __syncthreads();
    // 1. home atoms of this batch of records, compacted with a wave ballot
    const int v = base + tid;
    int atom = 0;
    bool home = false;
    if (tid < kGatherBatch && v < total) {
      const int r = a.rec_idx[brick_record(a, v, s_pre, s_base)];
      home = r < 0;
      atom = r & 0x7fffffff;
    }
    const unsigned long long m = __ballot(home);
    int wbase = 0;
    if ((tid & 63) == 0 && m != 0ull) wbase = atomicAdd(&s_cnt, __popcll(m));
    wbase = __shfl(wbase, 0, 64);
    // 2. current coordinates -> grid index and weights, once per atom
    if (home) {
      const int slot = wbase + __popcll(m & __lanemask_lt());
      const float4 p = xyzq[atom];
      int ix, iy, iz;
      float wxf, wyf, wzf;
      grid_index(p.x, a.r1, a.nx, ix, wxf);
      grid_index(p.y, a.r2, a.ny, iy, wyf);
      grid_index(p.z, a.r3, a.nz, iz, wzf);
      s_meta[slot] = make_int4(safe_index(ix, a.nx), safe_index(iy, a.ny), safe_index(iz, a.nz), atom);
      s_q[slot] = p.w;
      float* th = s_th + slot * NT;
      float t[P], d[P];
      bspline_theta_dtheta<P>(wxf, t, d);
#pragma unroll
      for (int j = 0; j < P; j++) {
        th[j] = t[j];
        th[3 * P + j] = d[j];
      }
      bspline_theta_dtheta<P>(wyf, t, d);
#pragma unroll
      for (int j = 0; j < P; j++) {
        th[P + j] = t[j];
        th[4 * P + j] = d[j];
      }
      bspline_theta_dtheta<P>(wzf, t, d);
#pragma unroll
      for (int j = 0; j < P; j++) {
        th[2 * P + j] = t[j];
        th[5 * P + j] = d[j];
      }
    }
    __syncthreads();
    // 3. G lanes per atom: lane g owns stencil column x = ix + g
    const int na = s_cnt;
    for (int a0 = 0; a0 < na; a0 += kSpreadBlock / G) {
      const int ai = a0 + tid / G;
      const bool active = ai < na;
      float f1 = 0.f, f2 = 0.f, f3 = 0.f;
      if (active && g < P) {
        const int4 mt = s_meta[ai];
        const float* th = s_th + ai * NT;
        const float tx = th[g], dtx = th[3 * P + g];
        float thy[P], thz[P], dthy[P], dthz[P];
#pragma unroll
        for (int j = 0; j < P; j++) {
          thy[j] = th[P + j];
          thz[j] = th[2 * P + j];
          dthy[j] = th[4 * P + j];
          dthz[j] = th[5 * P + j];
        }
        int dx = mt.x - x0, dy = mt.y - y0, dz = mt.z - z0;
        if (dx >= wx) dx -= a.nx;
        if (dy >= wy) dy -= a.ny;
        if (dz >= wz) dz -= a.nz;
        // f1 = dtx sum_z thz sum_y thy phi, f2 = tx sum_z thz sum_y dthy phi,
        // f3 = tx sum_z dthz sum_y thy phi: two y sums per z plane
        float g1 = 0.f, g2 = 0.f, g3 = 0.f;
        if (dx >= 0 && dy >= 0 && dz >= 0) {
          const float* col = pl + (dz * hys + dy) * hxs + dx + g;
#pragma unroll
          for (int jz = 0; jz < P; jz++) {
            float sy = 0.f, sdy = 0.f;
#pragma unroll
            for (int jy = 0; jy < P; jy++) {
              const float vv = col[(jz * hys + jy) * hxs];
              sy = fmaf(thy[jy], vv, sy);
              sdy = fmaf(dthy[jy], vv, sdy);
            }
            g1 = fmaf(thz[jz], sy, g1);
            g2 = fmaf(thz[jz], sdy, g2);
            g3 = fmaf(dthz[jz], sy, g3);
          }
        } else {
          // coordinates changed since the binning: potential from HBM
          int gx = mt.x + g;
          if (gx >= a.nx) gx -= a.nx;
#pragma unroll
          for (int jz = 0; jz < P; jz++) {
            int gz = mt.z + jz;
            if (gz >= a.nz) gz -= a.nz;
            float sy = 0.f, sdy = 0.f;
#pragma unroll
            for (int jy = 0; jy < P; jy++) {
              int gy = mt.y + jy;
              if (gy >= a.ny) gy -= a.ny;
              const float vv = phi[(size_t)gx + (size_t)a.nx * ((size_t)gy + (size_t)a.ny * gz)];
              sy = fmaf(thy[jy], vv, sy);
              sdy = fmaf(dthy[jy], vv, sdy);
            }
            g1 = fmaf(thz[jz], sy, g1);
            g2 = fmaf(thz[jz], sdy, g2);
            g3 = fmaf(dthz[jz], sy, g3);
          }
        }
        f1 = dtx * g1;
        f2 = tx * g2;
        f3 = tx * g3;
      }
#pragma unroll
      for (int off = 1; off < G; off <<= 1) {
        f1 += __shfl_xor(f1, off, 64);
        f2 += __shfl_xor(f2, off, 64);
        f3 += __shfl_xor(f3, off, 64);
      }
      if (active && g == 0) {
        // CudaPMERecip.cu:1231-1241: q*ccelec, then q*recip*f*nfft in FP32
        const float q = s_q[ai] * ccelec_f;
        const float fx = q * a.r1 * f1 * (float)a.nx;
        const float fy = q * a.r2 * f2 * (float)a.ny;
        const float fz = q * a.r3 * f3 * (float)a.nz;
        store_force(ft, force, s_meta[ai].w, stride, fx, fy, fz);
      }
    }
    __syncthreads();
  }
}

template <int P>
__global__ __launch_bounds__(kSpreadBlock) void k_gather_brick(const float4* __restrict__ xyzq, BinArgs a, int wxm,
                                                               int wym, int wzm, const float* __restrict__ phi,
                                                               float ccelec_f, int ft, int stride, void* force) {
  extern __shared__ __align__(16) unsigned char smem[];
  gather_brick_body<P>(xyzq, a, blockIdx.x, wxm, wym, wzm, phi, ccelec_f, ft, stride, force, smem);
}

size_t gather_lds_bytes(int order, const int wmax[3], int nchunk) {
  const size_t halo = (size_t)(wmax[0] + order - 1) * (wmax[1] + order - 1) * (wmax[2] + order - 1);
  return sizeof(float) * ((halo + 3) & ~(size_t)3) + sizeof(float) * kGatherBatch * 6 * order +
         sizeof(int4) * kGatherBatch + sizeof(float) * kGatherBatch + sizeof(int) * (2 * (size_t)nchunk + 1);
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

template <int P>
__global__ __launch_bounds__(kSpreadBlock) void k_batch_gather(const SysDesc* __restrict__ sys,
                                                               const WorkItem* __restrict__ items, int nitems) {
  extern __shared__ __align__(16) unsigned char smem[];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const WorkItem w = uniform_item(items[it]);
    const SysDesc& d = sys[w.sys];
    gather_brick_body<P>(d.xyzq, d.bin, w.idx, d.wmax[0], d.wmax[1], d.wmax[2], d.grid, (float)kCcelec, d.force_type,
                         d.stride, d.force, smem);
    __syncthreads();
  }
}

hipError_t launch_batch_gather(int order, const SysDesc* sys, const WorkItem* items, int nitems, int grid, size_t lds,
                               hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  switch (order) {
    case 4: hipLaunchKernelGGL(k_batch_gather<4>, dim3(grid), dim3(kSpreadBlock), lds, s, sys, items, nitems); break;
    case 6: hipLaunchKernelGGL(k_batch_gather<6>, dim3(grid), dim3(kSpreadBlock), lds, s, sys, items, nitems); break;
    case 8: hipLaunchKernelGGL(k_batch_gather<8>, dim3(grid), dim3(kSpreadBlock), lds, s, sys, items, nitems); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace pme
