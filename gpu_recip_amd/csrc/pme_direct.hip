// pme_direct.hip -- the direct-space (real-space) half of the Ewald sum for
// CDNA4: the non-bonded pair kernel over the reference's tile neighbour list
// and the 1-4 interaction / exclusion kernel, with the semantics of
// CudaPMEDirectForce<long long, float> (CudaDirectForce_util.h:184-251,
// CudaDirectForceKernels.cu:352-400, 506-716).
//
// Non-bonded pairs: one wave64 per i-tile entry (32 i atoms, one periodic
// shift); each half of the wave takes one of two consecutive j tiles, one j
// atom per lane.  The 32 shifted i atoms sit in LDS and are visited in the
// rotation i = (lane + t) mod 32, so the 32 lanes of a half touch 32
// different i atoms at every step: their force sums are LDS float adds with
// no two lanes on one word, in a fixed order (deterministic).  Pair forces
// are summed in FP32 registers (j) and LDS (i) and added to the caller's
// force array once per tile as 2^40 fixed point (integer adds: the result
// does not depend on the order of the tiles).  Energies and the shift forces
// of the virial are per-workgroup FP64 partials summed in fixed order by a
// one-workgroup reduction into the energy/virial buffer.
#include "pme_internal.hpp"

namespace pme {

// calc_box_shift<float>, gpu_utils.h:312-320: ish = (ix+1) + 3(iy+1) + 9(iz+1)
__device__ __forceinline__ void box_shift(int ish, const DirectParams& d, float& sx, float& sy, float& sz) {
  const int iz = ish / 9, r = ish - 9 * iz, iy = r / 3, ix = r - 3 * iy;
  sx = (float)(ix - 1) * d.boxx;
  sy = (float)(iy - 1) * d.boxy;
  sz = (float)(iz - 1) * d.boxz;
}

// The EWALD pair term (CudaDirectForceKernels.cu:352-390) for charges already
// multiplied by ccelec: pot = qq (erfc(kappa r) + fac_m1) / r with fac_m1 = 0
// for non-bonded pairs, e14fac - 1 for 1-4 interactions and -1 for 1-4
// exclusions.  Returns fij = (r dU/dr) / r^2, so that fij * (xi - xj) is the
// gradient on i.  erfcf and expf are the ROCm device library's.
__device__ __forceinline__ float ewald_fij(float r2, float rinv, float qq, float fac_m1, const DirectParams& d,
                                           float& pot) {
  const float r = r2 * rinv;
  const float erfc_val = erfcf(d.kappa * r);
  const float exp_val = expf(-d.kappa2 * r2);
  const float qq_efac_rinv = qq * (erfc_val + fac_m1) * rinv;
  pot = qq_efac_rinv;
  const float two_sqrtpi = 1.12837916709551f;  // 2/sqrt(pi)
  const float fij_elec = -qq * two_sqrtpi * d.kappa * exp_val - qq_efac_rinv;
  return fij_elec * rinv * rinv;
}

// f (+=) as 2^40 fixed point, write_force<long long> (gpu_utils.h:292-302)
__device__ __forceinline__ void add_fixed(unsigned long long* force, int i, int stride, float fx, float fy, float fz) {
  atomicAdd(&force[i], (unsigned long long)(long long)rintf(kForceScale * fx));
  atomicAdd(&force[(size_t)i + stride], (unsigned long long)(long long)rintf(kForceScale * fy));
  atomicAdd(&force[(size_t)i + 2 * (size_t)stride], (unsigned long long)(long long)rintf(kForceScale * fz));
}

__device__ __forceinline__ void lds_fadd(float* p, float v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The workgroup's partials: elec, excl, then sforce[27][3] (the shift forces
// of the virial, CudaEnergyVirial.h:14).  Each thread brings its energies and
// one shift-force contribution (ish < 0: none); the shift forces are summed
// per shift in thread order and the energies with block_sum, both in fixed
// order.
template <int NT>
__device__ __forceinline__ void write_direct_partial(double elec, double excl, int ish, double sfx, double sfy,
                                                     double sfz, bool energy, bool virial, double* out) {
  __shared__ double s_red[2 * (NT / 64)];
  __shared__ double s_sf[NT][3];
  __shared__ int s_ish[NT];
  const int tid = threadIdx.x;
  if (energy) {
    double v[2] = {elec, excl};
    block_sum<2, NT>(v, s_red);
    if (tid == 0) {
      out[0] = v[0];
      out[1] = v[1];
    }
  }
  if (virial) {
    s_ish[tid] = ish;
    s_sf[tid][0] = sfx;
    s_sf[tid][1] = sfy;
    s_sf[tid][2] = sfz;
    __syncthreads();
    for (int k = tid; k < 81; k += NT) {
      const int s = k / 3, c = k - 3 * s;
      double acc = 0.0;
      for (int t = 0; t < NT; t++)
        if (s_ish[t] == s) acc += s_sf[t][c];
      out[2 + k] = acc;
    }
  }
}

// One i-tile entry per wave; kDirWaves waves per workgroup.
template <bool ENERGY, bool VIRIAL>
__global__ __launch_bounds__(kDirWaves * 64) void k_direct_force(DirectParams d, int n_ientry,
                                                                 const pme_ientry* __restrict__ ientry,
                                                                 const int* __restrict__ tile_indj,
                                                                 const pme_tile_excl* __restrict__ tile_excl,
                                                                 const float4* __restrict__ xyzq, int stride,
                                                                 unsigned long long* __restrict__ force,
                                                                 double* __restrict__ partial) {
  __shared__ float4 s_xi[kDirWaves][32];     // shifted i atoms, q * ccelec
  __shared__ float s_fi[kDirWaves][2][3][32];  // i forces per half of the wave
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, h = lane >> 5, jl = lane & 31;
  const int e = blockIdx.x * kDirWaves + wv;
  const bool active = e < n_ientry;
  pme_ientry en{0, 13, 1, 0};
  if (active) en = ientry[e];
  float sx, sy, sz;
  box_shift(en.ish, d, sx, sy, sz);
  if (active && h == 0) {
    const float4 p = xyzq[en.iatomStart + jl];
    s_xi[wv][jl] = make_float4(p.x + sx, p.y + sy, p.z + sz, p.w * (float)kCcelec);
  }
  s_fi[wv][h][0][jl] = 0.f;
  s_fi[wv][h][1][jl] = 0.f;
  s_fi[wv][h][2][jl] = 0.f;
  __syncthreads();
  double pot = 0.0;
  for (int base = en.tileStart; base <= en.tileEnd; base += 2) {
    const int jt = base + h;
    const bool valid = jt <= en.tileEnd;
    unsigned excl = 0xffffffffu;
    float4 pj = make_float4(0.f, 0.f, 0.f, 0.f);
    int indj = 0;
    if (valid) {
      indj = tile_indj[jt];
      excl = tile_excl[jt].excl[jl];
      pj = xyzq[indj + jl];
    }
    float fjx = 0.f, fjy = 0.f, fjz = 0.f;
#pragma unroll 4
    for (int t = 0; t < 32; t++) {
      const int ii = (jl + t) & 31;
      const float4 pi = s_xi[wv][ii];
      const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      const float r2 = dx * dx + dy * dy + dz * dz;
      if (!(excl & 1u) && r2 < d.roff2) {
        float pe;
        const float fij = ewald_fij(r2, rsqrtf(r2), pi.w * pj.w, 0.0f, d, pe);
        if (ENERGY) pot += (double)pe;
        const float fx = fij * dx, fy = fij * dy, fz = fij * dz;
        fjx -= fx;
        fjy -= fy;
        fjz -= fz;
        lds_fadd(&s_fi[wv][h][0][ii], fx);
        lds_fadd(&s_fi[wv][h][1][ii], fy);
        lds_fadd(&s_fi[wv][h][2][ii], fz);
      }
      excl >>= 1;
    }
    if (valid && (fjx != 0.f || fjy != 0.f || fjz != 0.f)) add_fixed(force, indj + jl, stride, fjx, fjy, fjz);
  }
  __syncthreads();
  // i forces: the two halves' sums, one i atom per lane of the first half
  float fix = 0.f, fiy = 0.f, fiz = 0.f;
  if (active && h == 0) {
    fix = s_fi[wv][0][0][jl] + s_fi[wv][1][0][jl];
    fiy = s_fi[wv][0][1][jl] + s_fi[wv][1][1][jl];
    fiz = s_fi[wv][0][2][jl] + s_fi[wv][1][2][jl];
    if (fix != 0.f || fiy != 0.f || fiz != 0.f) add_fixed(force, en.iatomStart + jl, stride, fix, fiy, fiz);
  }
  if (ENERGY || VIRIAL) {
    // shift force of the entry (CudaDirectForce_util.h:321-347): the sum of
    // its i forces, except for the central image
    double sf[3] = {(double)fix, (double)fiy, (double)fiz};
#pragma unroll
    for (int c = 0; c < 3; c++) sf[c] = wave_sum(sf[c]);
    const int ish = (VIRIAL && active && lane == 0 && en.ish != 13) ? en.ish : -1;
    write_direct_partial<kDirWaves * 64>(pot, 0.0, ish, sf[0], sf[1], sf[2], ENERGY, VIRIAL,
                                         partial + (size_t)blockIdx.x * kDirPartial);
  }
}

// 1-4 interactions [0, nin14) and exclusions [nin14, nin14 + nex14), one
// entry per thread (calc_14_force_kernel, CudaDirectForceKernels.cu:664-760).
template <bool ENERGY, bool VIRIAL>
__global__ __launch_bounds__(kDir14Block) void k_direct_14(DirectParams d, int nin14, int nex14,
                                                           const pme_xx14* __restrict__ in14,
                                                           const pme_xx14* __restrict__ ex14,
                                                           const float4* __restrict__ xyzq, int stride,
                                                           unsigned long long* __restrict__ force,
                                                           double* __restrict__ partial) {
  const int k = blockIdx.x * kDir14Block + threadIdx.x;
  double elec = 0.0, excl = 0.0, sfx = 0.0, sfy = 0.0, sfz = 0.0;
  int ish_out = -1;
  if (k < nin14 + nex14) {
    const bool is_in = k < nin14;
    const pme_xx14 p = is_in ? in14[k] : ex14[k - nin14];
    float sx, sy, sz;
    box_shift(p.ishift, d, sx, sy, sz);
    const float4 xi = xyzq[p.i], xj = xyzq[p.j];
    const float dx = xi.x - xj.x + sx, dy = xi.y - xj.y + sy, dz = xi.z - xj.z + sz;
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float qq = (float)kCcelec * xi.w * xj.w;
    const float r = sqrtf(r2);
    float pe;
    const float fij = ewald_fij(r2, 1.0f / r, qq, is_in ? d.e14fac - 1.0f : -1.0f, d, pe);
    if (is_in) elec = (double)pe;
    else excl = (double)pe;
    const float fx = fij * dx, fy = fij * dy, fz = fij * dz;
    add_fixed(force, p.i, stride, fx, fy, fz);
    add_fixed(force, p.j, stride, -fx, -fy, -fz);
    if (p.ishift != 13) {
      ish_out = p.ishift;
      sfx = (double)fx;
      sfy = (double)fy;
      sfz = (double)fz;
    }
  }
  if (ENERGY || VIRIAL)
    write_direct_partial<kDir14Block>(elec, excl, VIRIAL ? ish_out : -1, sfx, sfy, sfz, ENERGY, VIRIAL,
                                      partial + (size_t)blockIdx.x * kDirPartial);
}

// Sum the partials of nblk workgroups in fixed order and add them into the
// energy/virial buffer: the elec and excl slots and sforce_dp[27][3].
__global__ __launch_bounds__(128) void k_direct_reduce(const double* __restrict__ partial, int nblk, int energy,
                                                       int virial, double* e_elec, double* e_excl,
                                                       double* sforce_dp) {
  const int c = threadIdx.x;
  if (c >= kDirPartial) return;
  if (c < 2 && !energy) return;
  if (c >= 2 && !virial) return;
  double acc = 0.0;
  for (int b = 0; b < nblk; b++) acc += partial[(size_t)b * kDirPartial + c];
  if (c == 0) e_elec[0] += acc;
  else if (c == 1) e_excl[0] += acc;
  else sforce_dp[c - 2] += acc;
}

hipError_t launch_direct_force(const DirectParams& d, int n_ientry, const pme_ientry* ientry, const int* tile_indj,
                               const pme_tile_excl* tile_excl, const float4* xyzq, int stride,
                               unsigned long long* force, bool energy, bool virial, double* partial,
                               hipStream_t s) {
  const int nblk = direct_force_blocks(n_ientry);
  if (nblk == 0) return hipSuccess;
  const dim3 g(nblk), b(kDirWaves * 64);
  if (energy && virial)
    hipLaunchKernelGGL((k_direct_force<true, true>), g, b, 0, s, d, n_ientry, ientry, tile_indj, tile_excl, xyzq, stride,
                       force, partial);
  else if (energy)
    hipLaunchKernelGGL((k_direct_force<true, false>), g, b, 0, s, d, n_ientry, ientry, tile_indj, tile_excl, xyzq,
                       stride, force, partial);
  else if (virial)
    hipLaunchKernelGGL((k_direct_force<false, true>), g, b, 0, s, d, n_ientry, ientry, tile_indj, tile_excl, xyzq,
                       stride, force, partial);
  else
    hipLaunchKernelGGL((k_direct_force<false, false>), g, b, 0, s, d, n_ientry, ientry, tile_indj, tile_excl, xyzq,
                       stride, force, partial);
  return hipGetLastError();
}

hipError_t launch_direct_14(const DirectParams& d, int nin14, int nex14, const pme_xx14* in14, const pme_xx14* ex14,
                            const float4* xyzq, int stride, unsigned long long* force, bool energy, bool virial,
                            double* partial, hipStream_t s) {
  const int nblk = direct_14_blocks(nin14 + nex14);
  if (nblk == 0) return hipSuccess;
  const dim3 g(nblk), b(kDir14Block);
  if (energy && virial)
    hipLaunchKernelGGL((k_direct_14<true, true>), g, b, 0, s, d, nin14, nex14, in14, ex14, xyzq, stride, force, partial);
  else if (energy)
    hipLaunchKernelGGL((k_direct_14<true, false>), g, b, 0, s, d, nin14, nex14, in14, ex14, xyzq, stride, force, partial);
  else if (virial)
    hipLaunchKernelGGL((k_direct_14<false, true>), g, b, 0, s, d, nin14, nex14, in14, ex14, xyzq, stride, force, partial);
  else
    hipLaunchKernelGGL((k_direct_14<false, false>), g, b, 0, s, d, nin14, nex14, in14, ex14, xyzq, stride, force,
                       partial);
  return hipGetLastError();
}

hipError_t launch_direct_reduce(const double* partial, int nblk, bool energy, bool virial, double* e_elec,
                                double* e_excl, double* sforce_dp, hipStream_t s) {
  if (nblk == 0 || !(energy || virial)) return hipSuccess;
  hipLaunchKernelGGL(k_direct_reduce, dim3(1), dim3(128), 0, s, partial, nblk, energy ? 1 : 0, virial ? 1 : 0, e_elec,
                     e_excl, sforce_dp);
  return hipGetLastError();
}

int direct_force_blocks(int n_ientry) { return (n_ientry + kDirWaves - 1) / kDirWaves; }
int direct_14_blocks(int n) { return (n + kDir14Block - 1) / kDir14Block; }

}  // namespace pme
