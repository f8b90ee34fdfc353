// pme_misc.hip -- small kernels: zero fill (used instead of hipMemsetAsync so
// that clears can sit inside captured graphs), self energy, B-spline fill.
#include "pme_internal.hpp"

namespace pme {

__global__ __launch_bounds__(kBlock) void k_fill_zero(unsigned long long* p, size_t nwords, unsigned char* tail,
                                                      int ntail) {
  const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  const size_t step = (size_t)gridDim.x * kBlock;
  for (size_t k = i; k < nwords; k += step) p[k] = 0ull;
  if (i < (size_t)ntail) tail[i] = 0;
}

hipError_t launch_fill_zero(void* p, size_t bytes, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  const size_t nwords = bytes / 8;
  const int ntail = (int)(bytes - nwords * 8);
  size_t nblk = (nwords + kBlock - 1) / kBlock;
  if (nblk < 1) nblk = 1;
  if (nblk > 2048) nblk = 2048;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)nblk), dim3(kBlock), 0, s, (unsigned long long*)p, nwords,
                     (unsigned char*)p + nwords * 8, ntail);
  return hipGetLastError();
}

// calc_self_energy_kernel, CudaPMERecip.cu:1252-1275: q*q in FP32, summed in
// FP64.  Grid-stride with at most 1024 workgroups writing partials; the sum
// and the -kappa*ccelec/sqrt(pi) factor are applied by ev_reduce_block.
__global__ __launch_bounds__(kBlock) void k_self_energy(const float4* __restrict__ xyzq, int n, double* partial) {
  __shared__ double sh[kBlock / 64];
  double v[1] = {0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const float q = xyzq[i].w;
    v[0] += (double)(q * q);
  }
  block_sum<1>(v, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = v[0];
}

hipError_t launch_self_energy(const float4* xyzq, int n, double* partial, int* nblocks, hipStream_t s) {
  int nblk = (n + kBlock - 1) / kBlock;
  if (nblk < 1) nblk = 1;
  if (nblk > 1024) nblk = 1024;
  *nblocks = nblk;
  hipLaunchKernelGGL(k_self_energy, dim3(nblk), dim3(kBlock), 0, s, xyzq, n, partial);
  return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_ev_reduce(EvReduce red) { ev_reduce_block<kBlock>(red); }

hipError_t launch_ev_reduce(const EvReduce& red, hipStream_t s) {
  hipLaunchKernelGGL(k_ev_reduce, dim3(1), dim3(kBlock), 0, s, red);
  return hipGetLastError();
}

// Bspline<float>::fill_bspline, Bspline.cu:9-145, 203-231, for orders 4/6/8
// with the spread/gather index formula.
template <int P>
__global__ __launch_bounds__(kBlock) void k_fill_bspline(const float4* __restrict__ xyzq, int n, float r1, float r2,
                                                         float r3, int nx, int ny, int nz, int* gix, int* giy,
                                                         int* giz, float* q, float* thx, float* thy, float* thz,
                                                         float* dthx, float* dthy, float* dthz) {
  const int a = blockIdx.x * kBlock + threadIdx.x;
  if (a >= n) return;
  const float4 p = xyzq[a];
  int ix, iy, iz;
  float wx, wy, wz;
  grid_index(p.x, r1, nx, ix, wx);
  grid_index(p.y, r2, ny, iy, wy);
  grid_index(p.z, r3, nz, iz, wz);
  gix[a] = ix;
  giy[a] = iy;
  giz[a] = iz;
  q[a] = p.w;
  float t[P], d[P];
  bspline_theta_dtheta<P>(wx, t, d);
#pragma unroll
  for (int j = 0; j < P; j++) {
    thx[(size_t)a * P + j] = t[j];
    dthx[(size_t)a * P + j] = d[j];
  }
  bspline_theta_dtheta<P>(wy, t, d);
#pragma unroll
  for (int j = 0; j < P; j++) {
    thy[(size_t)a * P + j] = t[j];
    dthy[(size_t)a * P + j] = d[j];
  }
  bspline_theta_dtheta<P>(wz, t, d);
#pragma unroll
  for (int j = 0; j < P; j++) {
    thz[(size_t)a * P + j] = t[j];
    dthz[(size_t)a * P + j] = d[j];
  }
}

hipError_t launch_fill_bspline(int order, const float4* xyzq, int n, float r1, float r2, float r3, int nx, int ny,
                               int nz, int* gix, int* giy, int* giz, float* q, float* thx, float* thy, float* thz,
                               float* dthx, float* dthy, float* dthz, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const dim3 g((n + kBlock - 1) / kBlock), b(kBlock);
  switch (order) {
    case 4:
      hipLaunchKernelGGL(k_fill_bspline<4>, g, b, 0, s, xyzq, n, r1, r2, r3, nx, ny, nz, gix, giy, giz, q, thx, thy, thz, dthx, dthy, dthz);
      break;
    case 6:
      hipLaunchKernelGGL(k_fill_bspline<6>, g, b, 0, s, xyzq, n, r1, r2, r3, nx, ny, nz, gix, giy, giz, q, thx, thy, thz, dthx, dthy, dthz);
      break;
    case 8:
      hipLaunchKernelGGL(k_fill_bspline<8>, g, b, 0, s, xyzq, n, r1, r2, r3, nx, ny, nz, gix, giy, giz, q, thx, thy, thz, dthx, dthy, dthz);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace pme
