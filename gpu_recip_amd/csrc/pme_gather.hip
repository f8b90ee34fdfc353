// pme_gather.hip -- force gather (CudaPMERecip.cu:1031-1245 semantics) for CDNA4.
//
// One workgroup per grid brick.  The workgroup stages the potential of its
// brick plus a (p-1) halo on the upper side in LDS with coalesced loads, then
// processes the atoms of its bin (binned by the spread): G lanes per atom
// (G = 4 for order 4, 8 for orders 6 and 8), lane g owns stencil column
// x = ix+g and walks the p*p (y, z) rows out of LDS; the G partial sums are
// combined with wave shuffles.  Atoms that are no longer in their bin's brick
// (coordinates changed since the binning) and bin-overflow atoms read the
// potential from HBM instead, so the result never depends on the binning.
#include "pme_internal.hpp"

namespace pme {

template <int FT>
__device__ __forceinline__ void store_force(void* force, int atom, int stride, float fx, float fy, float fz) {
  if (FT == 0) {  // float: strided SoA store, CudaPMERecip.cu:1005-1012
    float* f = reinterpret_cast<float*>(force);
    f[atom] = fx;
    f[(size_t)atom + stride] = fy;
    f[(size_t)atom + 2 * (size_t)stride] = fz;
  } else if (FT == 1) {  // float3: AoS store, :1016-1023
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

template <int P, int FT>
__global__ __launch_bounds__(kBlock) void k_gather_brick(const float4* __restrict__ xyzq, BinArgs a, int wxm,
                                                         int wym, int wzm, const float* __restrict__ phi,
                                                         float ccelec_f, int stride, void* force) {
  constexpr int G = (P <= 4) ? 4 : 8;
  extern __shared__ __align__(16) unsigned char smem[];
  float* pl = reinterpret_cast<float*>(smem);
  const int hxs = wxm + P - 1, hys = wym + P - 1;

  const int tid = threadIdx.x;
  const int b = blockIdx.x;
  const int nbins = a.nbx * a.nby * a.nbz;
  const int bxi = b % a.nbx, byi = (b / a.nbx) % a.nby, bzi = b / (a.nbx * a.nby);
  const int x0 = brick_lo(bxi, a.nx, a.nbx), wx = brick_lo(bxi + 1, a.nx, a.nbx) - x0;
  const int y0 = brick_lo(byi, a.ny, a.nby), wy = brick_lo(byi + 1, a.ny, a.nby) - y0;
  const int z0 = brick_lo(bzi, a.nz, a.nbz), wz = brick_lo(bzi + 1, a.nz, a.nbz) - z0;

  // stage brick + upper halo (periodic) in LDS
  const int hx = wx + P - 1, hy = wy + P - 1, hz = wz + P - 1;
  const int nh = hx * hy * hz;
  for (int i = tid; i < nh; i += kBlock) {
    const int lx = i % hx;
    const int t = i / hx;
    const int ly = t % hy;
    const int lz = t / hy;
    int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
    if (gx >= a.nx) gx -= a.nx;
    if (gy >= a.ny) gy -= a.ny;
    if (gz >= a.nz) gz -= a.nz;
    pl[(lz * hys + ly) * hxs + lx] = phi[(size_t)gx + (size_t)a.nx * ((size_t)gy + (size_t)a.ny * gz)];
  }
  __syncthreads();

  int cnt = a.bin_saved[b];
  if (cnt > a.cap) cnt = a.cap;
  int novf = a.ovf[1];
  if (novf > a.ovf_cap) novf = a.ovf_cap;
  const int chunk = (novf + nbins - 1) / nbins;
  int ovf_lo = b * chunk;
  int ovf_n = novf - ovf_lo;
  if (ovf_n > chunk) ovf_n = chunk;
  if (ovf_n < 0) ovf_n = 0;
  const int ntot = cnt + ovf_n;

  const int g = tid % G;
  for (int base = 0; base < ntot; base += kBlock / G) {
    const int ai = base + tid / G;
    const bool active = ai < ntot;
    int atom = 0;
    if (active) atom = (ai < cnt) ? a.bin_idx[(size_t)b * a.cap + ai] : a.ovf_idx[ovf_lo + ai - cnt];
    float f1 = 0.f, f2 = 0.f, f3 = 0.f;
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
    if (active) {
      p = xyzq[atom];
      int ix, iy, iz;
      float wxf, wyf, wzf;
      grid_index(p.x, a.r1, a.nx, ix, wxf);
      grid_index(p.y, a.r2, a.ny, iy, wyf);
      grid_index(p.z, a.r3, a.nz, iz, wzf);
      ix = safe_index(ix, a.nx);
      iy = safe_index(iy, a.ny);
      iz = safe_index(iz, a.nz);
      float thx[P], thy[P], thz[P], dthx[P], dthy[P], dthz[P];
      bspline_theta_dtheta<P>(wxf, thx, dthx);
      bspline_theta_dtheta<P>(wyf, thy, dthy);
      bspline_theta_dtheta<P>(wzf, thz, dthz);
      float tx = 0.f, dtx = 0.f;
#pragma unroll
      for (int j = 0; j < P; j++)
        if (g == j) {
          tx = thx[j];
          dtx = dthx[j];
        }
      if (g < P) {
        int dx = ix - x0, dy = iy - y0, dz = iz - z0;
        if (dx >= wx) dx -= a.nx;
        if (dy >= wy) dy -= a.ny;
        if (dz >= wz) dz -= a.nz;
        if (dx >= 0 && dy >= 0 && dz >= 0) {
          const float* col = pl + (dz * hys + dy) * hxs + dx + g;
#pragma unroll
          for (int jz = 0; jz < P; jz++) {
#pragma unroll
            for (int jy = 0; jy < P; jy++) {
              const float v = col[(jz * hys + jy) * hxs];
              const float tyz = thy[jy] * thz[jz];
              f1 += dtx * tyz * v;
              f2 += tx * dthy[jy] * thz[jz] * v;
              f3 += tx * thy[jy] * dthz[jz] * v;
            }
          }
        } else {
          int gx = ix + g;
          if (gx >= a.nx) gx -= a.nx;
#pragma unroll
          for (int jz = 0; jz < P; jz++) {
            int gz = iz + jz;
            if (gz >= a.nz) gz -= a.nz;
#pragma unroll
            for (int jy = 0; jy < P; jy++) {
              int gy = iy + jy;
              if (gy >= a.ny) gy -= a.ny;
              const float v = phi[(size_t)gx + (size_t)a.nx * ((size_t)gy + (size_t)a.ny * gz)];
              const float tyz = thy[jy] * thz[jz];
              f1 += dtx * tyz * v;
              f2 += tx * dthy[jy] * thz[jz] * v;
              f3 += tx * thy[jy] * dthz[jz] * v;
            }
          }
        }
      }
    }
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
      f1 += __shfl_xor(f1, off, 64);
      f2 += __shfl_xor(f2, off, 64);
      f3 += __shfl_xor(f3, off, 64);
    }
    if (active && g == 0) {
      // CudaPMERecip.cu:1231-1241: q*ccelec, then q*recip*f*nfft in FP32
      const float q = p.w * ccelec_f;
      const float fx = q * a.r1 * f1 * (float)a.nx;
      const float fy = q * a.r2 * f2 * (float)a.ny;
      const float fz = q * a.r3 * f3 * (float)a.nz;
      store_force<FT>(force, atom, stride, fx, fy, fz);
    }
  }
}

size_t gather_lds_bytes(int order, const int wmax[3]) {
  return sizeof(float) * (size_t)(wmax[0] + order - 1) * (wmax[1] + order - 1) * (wmax[2] + order - 1);
}

template <int P>
static hipError_t launch_gather_p(int ft, const float4* xyzq, const BinArgs& a, const int wmax[3], const float* phi,
                                  int stride, void* force, hipStream_t s) {
  const int nbins = a.nbx * a.nby * a.nbz;
  const size_t lds = gather_lds_bytes(P, wmax);
  const float cc = (float)kCcelec;
  switch (ft) {
    case 0:
      hipLaunchKernelGGL((k_gather_brick<P, 0>), dim3(nbins), dim3(kBlock), lds, s, xyzq, a, wmax[0], wmax[1], wmax[2], phi, cc, stride, force);
      break;
    case 1:
      hipLaunchKernelGGL((k_gather_brick<P, 1>), dim3(nbins), dim3(kBlock), lds, s, xyzq, a, wmax[0], wmax[1], wmax[2], phi, cc, stride, force);
      break;
    case 2:
      hipLaunchKernelGGL((k_gather_brick<P, 2>), dim3(nbins), dim3(kBlock), lds, s, xyzq, a, wmax[0], wmax[1], wmax[2], phi, cc, stride, force);
      break;
    default:
      return hipErrorInvalidValue;
  }
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

}  // namespace pme
