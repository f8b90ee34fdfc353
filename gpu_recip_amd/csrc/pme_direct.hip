// pme_direct.hip -- the direct-space (real-space) half of the Ewald sum for
// CDNA4: the non-bonded pair kernel over the reference's tile neighbour list
// and the 1-4 interaction / exclusion kernel, with the semantics of
// CudaPMEDirectForce<long long, float> (CudaDirectForce_util.h:184-251,
// CudaDirectForceKernels.cu:352-400, 506-716).
//
// Non-bonded pairs.  The j tiles of the list are split into equal
// contiguous ranges, one per wave, so the waves carry equal work whatever the
// spread of the entries' lengths (central-image entries hold tens of tiles,
// the others a few); a small kernel first maps every j tile to its entry.
// Each half of a wave walks its half of the range one tile per pass, one j
// atom per lane, holding the i tile of the tile's entry (32 atoms, moved by
// the entry's periodic shift on the j side) in LDS and flushing the i-force
// sums whenever the entry changes.  A pass has two phases.  The check phase
// tests the 32 x 64 pairs against the cutoff, each lane reading the i atoms
// at LDS addresses common to its half (broadcast reads) and building a 32-bit
// mask of its interacting i atoms; the tile's exclusion words, rotated into
// the same bit order, clear the excluded pairs.  The dense phase packs the
// wave's interacting pairs into an LDS queue (a wave prefix sum of the mask
// counts gives each lane its slots) and evaluates them 64 at a time, so the
// Ewald term runs only for pairs inside the cutoff (about one check in eight
// for a 32-atom tile list).  Pair forces are rounded once to 2^40 fixed
// point and summed with integer LDS atomics per i and per j atom; the sums
// go to the caller's force array as integer atomics, the j sums after every
// pass and the i sums when the entry changes, and the shift forces of the
// virial are summed the same way: none of these depends on the order of the
// pairs.  Energies are per-lane FP64 sums over a fixed set of pairs, summed
// per workgroup in wave order; a parallel fixed-order reduction adds the
// workgroups' partials to the energy/virial buffer.
#include "pme_internal.hpp"

namespace pme {

// calc_box_shift<float>, gpu_utils.h:312-320: ish = (ix+1) + 3(iy+1) + 9(iz+1)
__device__ __forceinline__ void box_shift(int ish, const DirectParams& d, float& sx, float& sy, float& sz) {
  const int iz = ish / 9, r = ish - 9 * iz, iy = r / 3, ix = r - 3 * iy;
  sx = (float)(ix - 1) * d.boxx;
  sy = (float)(iy - 1) * d.boxy;
  sz = (float)(iz - 1) * d.boxz;
}

// erfc(x) for x >= 0 from e^{-x^2}: erfc(x) e^{x^2} = t P(t) with
// t = 1 / (1 + p x), the form of Hastings' approximation (Abramowitz &
// Stegun 7.1.26) with ten terms whose coefficients were fitted here by
// minimax least squares on the relative error of erfc(x) e^{x^2} over
// [0, 12].  Evaluated in FP32 with the hardware reciprocal its relative error
// is below 4e-7 on [0, 6] (about three ulp, checked against an FP64 erfc).
// The pair term needs e^{-x^2} for the force anyway, so erfc costs a
// reciprocal and ten FMAs.
constexpr float kErfcP = 0.57f;
constexpr float kErfcC[10] = {0.3215866983f,  0.3216636181f, 0.2678903937f, 0.1791002899f,  -0.0476588681f,
                              0.2127946466f,  -0.7263579965f, 0.7648586631f, -0.3605679274f, 0.0666905344f};
__device__ __forceinline__ float erfc_from_exp(float x, float exp_neg_x2) {
  const float t = __builtin_amdgcn_rcpf(fmaf(kErfcP, x, 1.0f));
  float acc = kErfcC[9];
#pragma unroll
  for (int k = 8; k >= 0; k--) acc = fmaf(acc, t, kErfcC[k]);
  return acc * t * exp_neg_x2;
}

// The EWALD pair term (CudaDirectForceKernels.cu:352-390) for charges already
// multiplied by ccelec: pot = qq (erfc(kappa r) + fac_m1) / r with fac_m1 = 0
// for non-bonded pairs, e14fac - 1 for 1-4 interactions and -1 for 1-4
// exclusions.  Returns fij = (r dU/dr) / r^2, so that fij * (xi - xj) is the
// gradient on i.  expf is the ROCm device library's.
__device__ __forceinline__ float ewald_fij(float r2, float rinv, float qq, float fac_m1, const DirectParams& d,
                                           float& pot) {
  const float r = r2 * rinv;
  const float exp_val = expf(-d.kappa2 * r2);
  const float erfc_val = erfc_from_exp(d.kappa * r, exp_val);
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

// A pair-force component as 2^40 fixed point, rounded once to the nearest
// integer.  Below 2^11 in magnitude (every pair outside close contact) one
// FMA does it (to_fixed_product: f * 2^40 is exact in FP64 and below 2^51);
// larger f * 2^40 are integers already and convert exactly, held inside the
// format's range of +-2^23.
__device__ __forceinline__ unsigned long long pair_fixed(float f) {
  if (fabsf(f) < 2048.0f) return to_fixed_product((double)f, 1099511627776.0);
  const float c = fminf(fmaxf(f, -8388607.0f), 8388607.0f);
  return (unsigned long long)(long long)((double)c * 1099511627776.0);
}

// A pair energy as 2^32 fixed point (integer sums do not depend on which wave
// takes which pairs); |e| < 2^19 by one FMA, larger ones held inside +-2^30.
__device__ __forceinline__ long long energy_fixed(float e) {
  if (fabsf(e) < 524288.0f) return (long long)to_fixed_product((double)e, 4294967296.0);
  const float c = fminf(fmaxf(e, -1073741824.0f), 1073741824.0f);
  return (long long)((double)c * 4294967296.0);
}

// Inclusive prefix sum over the 64 lanes of a wave with DPP moves: shifts by
// 1, 2, 4 and 8 lanes inside each row of 16, then row 0's and row 2's last
// lanes into rows 1 and 3, and lane 31 into rows 2 and 3 (all VALU, no LDS).
__device__ __forceinline__ int wave_inclusive_scan(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

// Orders the LDS accesses of the lanes of one wave (all of them issue from
// the same instruction stream, so a compiler barrier and the LDS counter wait
// are enough; no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The workgroup's partials: elec, excl, then sforce[27][3] (the shift forces
// of the virial, CudaEnergyVirial.h:14).  Each thread brings its energies and
// one shift-force contribution (ish < 0: none); the energies are summed with
// block_sum in fixed order, the shift forces per shift as 2^40 fixed point
// with integer LDS atomics (exact, so the order does not matter).
template <int NT>
__device__ __forceinline__ void write_direct_partial(double elec, double excl, int ish, float sfx, float sfy,
                                                     float sfz, bool energy, bool virial, double* out) {
  __shared__ double s_red[2 * (NT / 64)];
  __shared__ unsigned long long s_sf[81];
  const int tid = threadIdx.x;
  if (virial) {
    for (int k = tid; k < 81; k += NT) s_sf[k] = 0ull;
    __syncthreads();
    if (ish >= 0) {
      atomicAdd(&s_sf[3 * ish + 0], pair_fixed(sfx));
      atomicAdd(&s_sf[3 * ish + 1], pair_fixed(sfy));
      atomicAdd(&s_sf[3 * ish + 2], pair_fixed(sfz));
    }
  }
  if (energy) {
    double v[2] = {elec, excl};
    block_sum<2, NT>(v, s_red);
    if (tid == 0) {
      out[0] = v[0];
      out[1] = v[1];
    }
  }
  if (virial) {
    __syncthreads();
    for (int k = tid; k < 81; k += NT) out[2 + k] = (double)(long long)s_sf[k] * (1.0 / 1099511627776.0);
  }
}

// tile_entry[t] = -1 for every tile (the map kernel then fills in the tiles
// of the entries).
__global__ __launch_bounds__(256) void k_direct_tile_clear(int n_tile, int* __restrict__ tile_entry) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t < n_tile) tile_entry[t] = -1;
}

// Every j tile's entry: tile_entry[t] = e for tileStart <= t <= tileEnd of
// entry e, one wave per entry; tiles of no entry stay -1.
__global__ __launch_bounds__(256) void k_direct_tile_map(int n_ientry, const pme_ientry* __restrict__ ientry,
                                                         int n_tile, int* __restrict__ tile_entry) {
  const int e = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (e >= n_ientry) return;
  const pme_ientry en = ientry[e];
  const int hi = min(en.tileEnd, n_tile - 1);
  for (int t = max(en.tileStart, 0) + lane; t <= hi; t += 64) tile_entry[t] = e;
}

// LDS of one wave of the pair kernel
struct DirWave {
  float4 xi[2][33];                 // i tile of each half's entry, q * ccelec (rows padded: other banks)
  float4 xj[64];                    // j atoms of the pass moved by minus their shift, q
  unsigned long long fi[2][3][32];  // i force sums of each half's entry, 2^40 fixed point
  unsigned long long fj[3][64];     // j force sums of the pass
  unsigned short pairs[32 * 64];    // interacting pairs of the pass: i | (j lane << 5)
};

// Adds the i-force sums of a half's entry to the force array and, for a
// shifted image, to the shift forces; clears them.  Lanes of the half only.
template <bool VIRIAL>
__device__ __forceinline__ void flush_entry(DirWave& W, int h, int jl, int iatom, int ish, int stride,
                                            unsigned long long* __restrict__ force, unsigned long long* s_sf) {
  unsigned long long g[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    g[c] = W.fi[h][c][jl];
    W.fi[h][c][jl] = 0ull;
  }
  if (g[0] | g[1] | g[2]) {
    atomicAdd(&force[iatom + jl], g[0]);
    atomicAdd(&force[(size_t)(iatom + jl) + stride], g[1]);
    atomicAdd(&force[(size_t)(iatom + jl) + 2 * (size_t)stride], g[2]);
    // the entry's shift force (CudaDirectForce_util.h:321-347): the sum of
    // its i forces, except for the central image
    if (VIRIAL && ish != 13) {
#pragma unroll
      for (int c = 0; c < 3; c++) atomicAdd(&s_sf[3 * ish + c], g[c]);
    }
  }
}

// The list's j tiles in chunks of kDirChunk, dealt to the waves in turn
// (chunk c to wave c mod nwave): every wave sees tiles from the whole list,
// whose pair density varies tenfold from tile to tile, so the waves' loads
// even out without a shared counter.  Within a chunk each half of the wave
// takes kDirChunk / 2 consecutive tiles, one per pass.  The j atoms of a
// chunk are loaded together at its start, and the tile records (entry, atom
// index, exclusions) of the wave's next chunk while the current one is
// processed: one memory round trip per chunk, not a chain of them per pass.
template <bool ENERGY, bool VIRIAL>
__global__ __launch_bounds__(kDirWaves * 64) void k_direct_force(DirectParams d, int n_ientry, int n_tile,
                                                                 const pme_ientry* __restrict__ ientry,
                                                                 const int* __restrict__ tile_entry,
                                                                 const int* __restrict__ tile_indj,
                                                                 const pme_tile_excl* __restrict__ tile_excl,
                                                                 const float4* __restrict__ xyzq, int stride,
                                                                 unsigned long long* __restrict__ force,
                                                                 double* __restrict__ partial) {
  __shared__ DirWave s_w[kDirWaves];
  __shared__ unsigned long long s_sf[81];  // shift forces, 2^40 fixed point
  __shared__ unsigned long long s_pot;     // energy, 2^32 fixed point
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, h = lane >> 5, jl = lane & 31;
  DirWave& W = s_w[wv];
  if (threadIdx.x < 81) s_sf[threadIdx.x] = 0ull;
  if (threadIdx.x == 0) s_pot = 0ull;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    W.fi[h][c][jl] = 0ull;
    W.fj[c][lane] = 0ull;
  }
  __syncthreads();
  const int nchunk = (n_tile + kDirChunk - 1) / kDirChunk;
  const int wtotal = (int)gridDim.x * kDirWaves;
  int chunk = (int)blockIdx.x * kDirWaves + wv;
  // the half's current entry
  int cur = -1, cur_lo = 0, cur_hi = -1, cur_ish = 13, cur_i0 = 0;
  float sx = 0.f, sy = 0.f, sz = 0.f;
  long long pot = 0;
  constexpr int NP = kDirChunk / 2;  // passes per chunk
  // tile records of the half's NP tiles of a chunk
  int ne[NP], nindj[NP];
  unsigned nexcl[NP];
  auto load_records = [&](int c) {
#pragma unroll
    for (int k = 0; k < NP; k++) {
      const int t = c * kDirChunk + h * NP + k;
      const bool in = c < nchunk && t < n_tile;
      ne[k] = in ? tile_entry[t] : -1;
      nindj[k] = in ? tile_indj[t] : 0;
      nexcl[k] = in ? tile_excl[t].excl[jl] : 0xffffffffu;
    }
  };
  load_records(chunk);
  while (chunk < nchunk) {
    const int tbase = chunk * kDirChunk + h * NP;
    int ce[NP], cindj[NP];
    unsigned cexcl[NP];
    float4 cpj[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) {
      ce[k] = ne[k];
      cindj[k] = nindj[k];
      cexcl[k] = nexcl[k];
    }
    // the j atoms of this chunk, then the records of the next one
#pragma unroll
    for (int k = 0; k < NP; k++) cpj[k] = ce[k] >= 0 ? xyzq[cindj[k] + jl] : make_float4(0.f, 0.f, 0.f, 0.f);
    load_records(chunk + wtotal);
#pragma unroll
    for (int k = 0; k < NP; k++) {
      const int t = tbase + k;
      const int e = ce[k];
      if (e != cur) {
        if (cur >= 0) flush_entry<VIRIAL>(W, h, jl, cur_i0, cur_ish, stride, force, s_sf);
        cur = -1;
        if (e >= 0 && e < n_ientry) {
          const pme_ientry en = ientry[e];
          if (t >= en.tileStart && t <= en.tileEnd) {
            cur = e;
            cur_lo = en.tileStart;
            cur_hi = en.tileEnd;
            cur_ish = en.ish;
            cur_i0 = en.iatomStart;
            box_shift(cur_ish, d, sx, sy, sz);
            const float4 p = xyzq[cur_i0 + jl];
            W.xi[h][jl] = make_float4(p.x, p.y, p.z, p.w * (float)kCcelec);
          }
        }
      }
      wave_lds_sync();
      const bool ok = cur >= 0 && t >= cur_lo && t <= cur_hi && t < n_tile;
      const unsigned excl = cexcl[k];
      const float4 pj = cpj[k];
      const int indj = cindj[k];
      // xi - (xj - shift) is the separation of i's image from j
      const float xjx = pj.x - sx, xjy = pj.y - sy, xjz = pj.z - sz;
      W.xj[lane] = make_float4(xjx, xjy, xjz, pj.w);
      // check phase: bit i of m = pair (i, j) inside the cutoff
      unsigned m = 0u;
#pragma unroll
      for (int g = 3; g >= 0; g--) {
        float4 pg[8];  // eight i atoms read together, then tested
#pragma unroll
        for (int u = 0; u < 8; u++) pg[u] = W.xi[h][8 * g + u];
#pragma unroll
        for (int u = 7; u >= 0; u--) {
          const float dx = pg[u].x - xjx, dy = pg[u].y - xjy, dz = pg[u].z - xjz;
          const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
          m = (m << 1) | (r2 < d.roff2 ? 1u : 0u);
        }
      }
      // bit t of excl[l] excludes i = (l + t) mod 32: rotated left by l it is bit i
      m &= ~__builtin_amdgcn_alignbit(excl, excl, (32 - jl) & 31);
      if (!ok) m = 0u;
      // dense phase: the wave's pairs packed in lane order
      const int cnt = __popc(m);
      const int incl = wave_inclusive_scan(cnt);
      const int total = __builtin_amdgcn_readlane(incl, 63);
      int slot = incl - cnt;
      while (m) {
        const int i = __ffs(m) - 1;
        m &= m - 1u;
        if (slot < 32 * 64) W.pairs[slot] = (unsigned short)(i | (lane << 5));
        slot++;
      }
      wave_lds_sync();
      for (int q = lane; q < total; q += 64) {
        const unsigned pr = W.pairs[q];
        const int i = pr & 31, l = pr >> 5, hl = l >> 5;
        const float4 pi = W.xi[hl][i], pjl = W.xj[l];
        const float dx = pi.x - pjl.x, dy = pi.y - pjl.y, dz = pi.z - pjl.z;
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        float pe;
        const float fij = ewald_fij(r2, rsqrtf(r2), pi.w * pjl.w, 0.0f, d, pe);
        if (ENERGY) pot += energy_fixed(pe);
        const float fx = fij * dx, fy = fij * dy, fz = fij * dz;
        atomicAdd(&W.fi[hl][0][i], pair_fixed(fx));
        atomicAdd(&W.fi[hl][1][i], pair_fixed(fy));
        atomicAdd(&W.fi[hl][2][i], pair_fixed(fz));
        atomicAdd(&W.fj[0][l], pair_fixed(-fx));
        atomicAdd(&W.fj[1][l], pair_fixed(-fy));
        atomicAdd(&W.fj[2][l], pair_fixed(-fz));
      }
      wave_lds_sync();
      unsigned long long g[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        g[c] = W.fj[c][lane];
        W.fj[c][lane] = 0ull;
      }
      if (ok && (g[0] | g[1] | g[2])) {
        atomicAdd(&force[indj + jl], g[0]);
        atomicAdd(&force[(size_t)(indj + jl) + stride], g[1]);
        atomicAdd(&force[(size_t)(indj + jl) + 2 * (size_t)stride], g[2]);
      }
      wave_lds_sync();
    }
    chunk += wtotal;
  }
  if (cur >= 0) flush_entry<VIRIAL>(W, h, jl, cur_i0, cur_ish, stride, force, s_sf);
  if (ENERGY || VIRIAL) {
    if (ENERGY && pot != 0) atomicAdd(&s_pot, (unsigned long long)pot);
    __syncthreads();
    double* out = partial + (size_t)blockIdx.x * kDirPartial;
    const int tid = threadIdx.x;
    if (ENERGY && tid == 0) {
      out[0] = (double)(long long)s_pot * (1.0 / 4294967296.0);
      out[1] = 0.0;
    }
    if (VIRIAL && tid < 81) out[2 + tid] = (double)(long long)s_sf[tid] * (1.0 / 1099511627776.0);
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
  double elec = 0.0, excl = 0.0;
  float sfx = 0.f, sfy = 0.f, sfz = 0.f;
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
      sfx = fx;
      sfy = fy;
      sfz = fz;
    }
  }
  if (ENERGY || VIRIAL)
    write_direct_partial<kDir14Block>(elec, excl, VIRIAL ? ish_out : -1, sfx, sfy, sfz, ENERGY, VIRIAL,
                                      partial + (size_t)blockIdx.x * kDirPartial);
}

// Sum the partials of nblk workgroups and add them into the energy/virial
// buffer: the elec and excl slots and sforce_dp[27][3].  One workgroup per
// slot; thread k sums workgroups k, k + 256, ... and block_sum adds the
// threads in fixed order (the same result on every run).
__global__ __launch_bounds__(256) void k_direct_reduce(const double* __restrict__ partial, int nblk, int energy,
                                                       int virial, double* e_elec, double* e_excl,
                                                       double* sforce_dp) {
  __shared__ double s_red[4];
  const int c = blockIdx.x;
  if ((c < 2 && !energy) || (c >= 2 && !virial)) return;
  double acc = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 256) acc += partial[(size_t)b * kDirPartial + c];
  block_sum<1, 256>(&acc, s_red);
  if (threadIdx.x == 0) {
    if (c == 0) e_elec[0] += acc;
    else if (c == 1) e_excl[0] += acc;
    else sforce_dp[c - 2] += acc;
  }
}

hipError_t launch_direct_force(const DirectParams& d, int n_ientry, int n_tile, int nwave,
                               const pme_ientry* ientry, int* tile_entry, const int* tile_indj,
                               const pme_tile_excl* tile_excl, const float4* xyzq, int stride,
                               unsigned long long* force, bool energy, bool virial, double* partial,
                               hipStream_t s) {
  if (n_ientry <= 0 || n_tile <= 0 || nwave <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_direct_tile_clear, dim3((n_tile + 255) / 256), dim3(256), 0, s, n_tile, tile_entry);
  hipLaunchKernelGGL(k_direct_tile_map, dim3((n_ientry + 3) / 4), dim3(256), 0, s, n_ientry, ientry, n_tile,
                     tile_entry);
  const dim3 g(direct_force_blocks(nwave)), b(kDirWaves * 64);
  if (energy && virial)
    hipLaunchKernelGGL((k_direct_force<true, true>), g, b, 0, s, d, n_ientry, n_tile, ientry, tile_entry, tile_indj,
                       tile_excl, xyzq, stride, force, partial);
  else if (energy)
    hipLaunchKernelGGL((k_direct_force<true, false>), g, b, 0, s, d, n_ientry, n_tile, ientry, tile_entry, tile_indj,
                       tile_excl, xyzq, stride, force, partial);
  else if (virial)
    hipLaunchKernelGGL((k_direct_force<false, true>), g, b, 0, s, d, n_ientry, n_tile, ientry, tile_entry, tile_indj,
                       tile_excl, xyzq, stride, force, partial);
  else
    hipLaunchKernelGGL((k_direct_force<false, false>), g, b, 0, s, d, n_ientry, n_tile, ientry, tile_entry, tile_indj,
                       tile_excl, xyzq, stride, force, partial);
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
  hipLaunchKernelGGL(k_direct_reduce, dim3(kDirPartial), dim3(256), 0, s, partial, nblk, energy ? 1 : 0,
                     virial ? 1 : 0, e_elec, e_excl, sforce_dp);
  return hipGetLastError();
}

int direct_force_blocks(int nwave) { return (nwave + kDirWaves - 1) / kDirWaves; }

int direct_force_waves(int n_tile, int ncu) {
  // 16 waves per CU (four workgroups of kDirWaves), at most one per chunk
  return std::max(1, std::min(ncu * 16, (n_tile + kDirChunk - 1) / kDirChunk));
}
int direct_14_blocks(int n) { return (n + kDir14Block - 1) / kDir14Block; }

}  // namespace pme
