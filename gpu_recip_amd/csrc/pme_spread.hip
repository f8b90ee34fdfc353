// pme_spread.hip -- charge spreading onto the PME grid (CudaPMERecip.cu:479-580
// semantics), re-designed for CDNA4:
//
//  1. k_bin_chunks: a chunked counting sort.  Workgroup w bins its chunk of
//     atoms by grid brick with LDS counters only (histogram, block scan,
//     cursors): an atom gets one record in every brick its p^3 stencil
//     touches (1 to 8), the copy in the brick of its base index is flagged
//     "home".  No global atomics, so nothing contends across the 8 XCDs.
//  2. k_spread_brick: one workgroup per brick.  It reads exactly the records
//     of its brick (all of them touch it), computes the B-spline weights into
//     LDS and accumulates q*thx*thy*thz in LDS as 64-bit fixed point with
//     scale 2^40 -- the reference's `long long` grid accumulation
//     (CudaPMERecip.cu:51-58, write_grid<long long>; reduce_force, reduce.cu:66-78).
//     Integer LDS atomics are an order of magnitude faster than float ones
//     on gfx950 and integer sums do not depend on the order of the adds, so
//     the charge grid is bitwise reproducible.  The brick is converted to
//     float and written to HBM once with coalesced stores (no memset, no
//     global atomics).
#include "pme_internal.hpp"

namespace pme {

// Divisors of the brick arithmetic of one grid, as reciprocals: grid edges
// are at most 512 and a brick count at most the edge, so every numerator
// below is under 2^19 and sdiv gives the exact integer quotient.
struct BrickDivs {
  SmallDiv n[3], nb[3];
};
__device__ __forceinline__ BrickDivs brick_divs(const BinArgs& a) {
  return BrickDivs{{small_div(a.nx), small_div(a.ny), small_div(a.nz)},
                   {small_div(a.nbx), small_div(a.nby), small_div(a.nbz)}};
}
// brick_lo(b, n, nb) and brick_of(i, n, nb) without integer division
__device__ __forceinline__ int brick_lo_d(int b, int n, const SmallDiv& dnb) { return sdiv(b * n, dnb); }
__device__ __forceinline__ int brick_of_d(int i, int nb, const SmallDiv& dn) { return sdiv((i + 1) * nb - 1, dn); }

// Does the stencil i..i+P-1 of an atom in brick b run past the brick's end
// (into brick b+1, periodic)?  An axis with one brick wraps onto itself.
__device__ __forceinline__ int spills(int i, int b, int n, int nb, int P, const SmallDiv& dnb) {
  if (nb == 1) return 0;
  return (i + P - 1 >= brick_lo_d(b + 1, n, dnb)) ? 1 : 0;
}

struct AtomBricks {
  int bx, by, bz;  // brick of the base index
  int sx, sy, sz;  // 1 if the stencil spills into the next brick
};

__device__ __forceinline__ AtomBricks atom_bricks(const float4& p, const BinArgs& a, const BrickDivs& dv) {
  int ix, iy, iz;
  float wx, wy, wz;
  atom_grid_index(a, p, ix, iy, iz, wx, wy, wz);
  ix = safe_index(ix, a.nx);
  iy = safe_index(iy, a.ny);
  iz = safe_index(iz, a.nz);
  AtomBricks r;
  r.bx = brick_of_d(ix, a.nbx, dv.n[0]);
  r.by = brick_of_d(iy, a.nby, dv.n[1]);
  r.bz = brick_of_d(iz, a.nbz, dv.n[2]);
  r.sx = spills(ix, r.bx, a.nx, a.nbx, a.order, dv.nb[0]);
  r.sy = spills(iy, r.by, a.ny, a.nby, a.order, dv.nb[1]);
  r.sz = spills(iz, r.bz, a.nz, a.nbz, a.order, dv.nb[2]);
  return r;
}

__device__ __forceinline__ int next_brick(int b, int d, int nb) {
  b += d;
  return b >= nb ? b - nb : b;
}

// Chunk w of the counting sort (one kBinBlock-thread workgroup).  hist:
// 2*nbins ints of LDS: per brick the count of home records, then the count of
// all records; after the scan, the cursors of the home records and of the
// other (spill) records.  Each brick's segment holds its home records first,
// so that the gather reads exactly the atoms it owns.  With the default chunk
// of kBinBlock atoms each thread bins one atom and keeps it and its bricks in
// registers between the counting and the placing pass.
__device__ __forceinline__ void bin_chunk_body(const float4* __restrict__ xyzq, int n, const BinArgs& a, int w,
                                               int* hist) {
  __shared__ int sh_scan[kBinBlock / 64];
  const int nbins = a.nbx * a.nby * a.nbz;
  int* hcnt = hist;          // [nbins] home counts, then home cursors
  int* tcnt = hist + nbins;  // [nbins] record counts, then spill cursors
  const int tid = threadIdx.x;
  const int a0 = w * a.chunk;
  const int a1 = min(a0 + a.chunk, n);
  const BrickDivs dv = brick_divs(a);
  for (int i = tid; i < 2 * nbins; i += kBinBlock) hist[i] = 0;
  // first atom of this thread, loaded while the histogram is cleared
  const int first = a0 + tid;
  float4 p0 = make_float4(0.f, 0.f, 0.f, 0.f);
  if (first < a1) p0 = xyzq[first];
  const AtomBricks k0 = atom_bricks(p0, a, dv);
  __syncthreads();
  for (int atom = first; atom < a1; atom += kBinBlock) {
    const AtomBricks k = atom == first ? k0 : atom_bricks(xyzq[atom], a, dv);
    atomicAdd(&hcnt[k.bx + a.nbx * (k.by + a.nby * k.bz)], 1);
    for (int dz = 0; dz <= k.sz; dz++)
      for (int dy = 0; dy <= k.sy; dy++)
        for (int dx = 0; dx <= k.sx; dx++) {
          const int b = next_brick(k.bx, dx, a.nbx) + a.nbx * (next_brick(k.by, dy, a.nby) + a.nby * next_brick(k.bz, dz, a.nbz));
          atomicAdd(&tcnt[b], 1);
        }
  }
  __syncthreads();
  const int total = block_exclusive_scan<kBinBlock>(tcnt, nbins, sh_scan);
  int* seg = a.seg_off + (size_t)w * seg_stride(nbins);
  int* hend = seg + nbins + 1;
  for (int i = tid; i < nbins; i += kBinBlock) {
    const int start = tcnt[i], nh = hcnt[i];
    seg[i] = start;
    hend[i] = start + nh;
    hcnt[i] = start;       // home cursor
    tcnt[i] = start + nh;  // spill cursor
  }
  if (tid == 0) seg[nbins] = total;
  __syncthreads();
  float4* rp = a.rec_pos + (size_t)w * a.rec_stride;
  int* ri = a.rec_idx + (size_t)w * a.rec_stride;
  for (int atom = first; atom < a1; atom += kBinBlock) {
    const float4 p = atom == first ? p0 : xyzq[atom];
    const AtomBricks k = atom == first ? k0 : atom_bricks(p, a, dv);
    for (int dz = 0; dz <= k.sz; dz++)
      for (int dy = 0; dy <= k.sy; dy++)
        for (int dx = 0; dx <= k.sx; dx++) {
          const int b = next_brick(k.bx, dx, a.nbx) + a.nbx * (next_brick(k.by, dy, a.nby) + a.nby * next_brick(k.bz, dz, a.nbz));
          if ((dx | dy | dz) == 0) {
            const int pos = atomicAdd(&hcnt[b], 1);
            rp[pos] = p;
            ri[pos] = atom;
          } else {
            rp[atomicAdd(&tcnt[b], 1)] = p;
          }
        }
  }
}

__global__ __launch_bounds__(kBinBlock) void k_bin_chunks(const float4* __restrict__ xyzq, int n, BinArgs a) {
  extern __shared__ int hist[];  // [nbins]: counts, then cursors
  bin_chunk_body(xyzq, n, a, blockIdx.x, hist);
}

size_t bin_lds_bytes(int nbins) { return 2 * sizeof(int) * (size_t)nbins; }

hipError_t launch_bin_atoms(const float4* xyzq, int n, const BinArgs& a, hipStream_t s) {
  const int nbins = a.nbx * a.nby * a.nbz;
  hipLaunchKernelGGL(k_bin_chunks, dim3(a.nchunk), dim3(kBinBlock), bin_lds_bytes(nbins), s, xyzq, n, a);
  return hipGetLastError();
}

// local offset of a grid index relative to a brick start x0 of width w:
// values >= w belong to the periodic image below the brick.
__device__ __forceinline__ int local_offset(int i, int x0, int w, int n) {
  int d = i - x0;
  if (d >= w) d -= n;
  return d;
}

// x stride of the LDS brick: P-1 margin points on either side, so a stencil
// row never needs a bounds test along x (margin sums are dropped, or folded
// back when one brick spans the whole periodic axis); odd in 64-bit words so
// rows of consecutive lanes fall in different banks.
__host__ __device__ constexpr int spread_xstride(int wxm, int P) { return (wxm + 2 * (P - 1)) | 1; }

// One record's p^3 stencil added into the brick accumulator acc (the x
// margin absorbs rows that leave the brick in x; rows outside it in y or z
// are skipped).
template <int P>
__device__ __forceinline__ void spread_record(const BinArgs& a, float4 p, int x0, int wx, int y0, int wy, int z0,
                                              int wz, bool wrapy, bool wrapz, int bxs, int plane,
                                              unsigned long long* __restrict__ acc) {
  int ix, iy, iz;
  float fx, fy, fz;
  atom_grid_index(a, p, ix, iy, iz, fx, fy, fz);
  float tx[P], ty[P], tz[P];
  bspline_theta<P>(fx, tx);
  bspline_theta<P>(fy, ty);
  bspline_theta<P>(fz, tz);
  // q * 2^40 (exact): the fixed-point scale rides on theta_x, widened to
  // FP64 once per record so each stencil point costs one FMA
  const float qs = p.w * 1099511627776.0f;
  double txd[P];
#pragma unroll
  for (int j = 0; j < P; j++) txd[j] = (double)(tx[j] * qs);
  // x offset into the margin-padded row: in [0, wx + 2(P-1)) always
  const int dx = local_offset(safe_index(ix, a.nx), x0, wx, a.nx) + (P - 1);
  const int dy = local_offset(safe_index(iy, a.ny), y0, wy, a.ny);
  const int dz = local_offset(safe_index(iz, a.nz), z0, wz, a.nz);
  // stencil rows inside the brick (bit j of ym / zm) and their LDS offsets,
  // once per record instead of per row
  int yo[P], zo[P];
  unsigned ym = 0u, zm = 0u;
#pragma unroll
  for (int j = 0; j < P; j++) {
    int ly = dy + j;
    ly -= (wrapy && ly >= wy) ? a.ny : 0;
    ym |= (unsigned)(ly >= 0 && ly < wy) << j;
    yo[j] = mul24(ly, bxs);
    int lz = dz + j;
    lz -= (wrapz && lz >= wz) ? a.nz : 0;
    zm |= (unsigned)(lz >= 0 && lz < wz) << j;
    zo[j] = mul24(lz, plane) + dx;
  }
#pragma unroll
  for (int jz = 0; jz < P; jz++) {
    if (!((zm >> jz) & 1u)) continue;
#pragma unroll
    for (int jy = 0; jy < P; jy++) {
      if (!((ym >> jy) & 1u)) continue;
      const double c = (double)(ty[jy] * tz[jz]);
      unsigned long long* row = acc + (zo[jz] + yo[jy]);
#pragma unroll
      for (int jx = 0; jx < P; jx++) atomicAdd(&row[jx], to_fixed_product(txd[jx], c));
    }
  }
}

// Store phase of a brick: pair i of x-neighbours (lx, lx + 1) of row (ly,
// lz) converted from fixed point and stored; the second can fall on the
// row's margin, inside the row's LDS stride, and is then not stored.  WRAP
// (one brick along x): the high margin is the periodic image of the row
// start and is added in.  Products of small non-negative integers in 24 bits
// (edges <= 512), grid offsets in 32 bits (nx*ny < 2^19, offsets < 2^27).
template <int P, bool WRAP, int NT>
__device__ __forceinline__ void store_brick_pairs(const unsigned long long* __restrict__ acc, float* __restrict__ grid,
                                                  int nx, int ny, int x0, int y0, int z0, int wx, int wy, int hw,
                                                  int npair, int bxs, int plane) {
  const int nxny = mul24(nx, ny);
  const SmallDiv dhw = small_div(hw), dwy = small_div(wy);
  for (int i = threadIdx.x; i < npair; i += NT) {
    const int t = sdiv(i, dhw);
    const int lx = 2 * (i - mul24(t, hw));
    const int lz = sdiv(t, dwy);
    const int ly = t - mul24(lz, wy);
    const unsigned long long* row = acc + (mul24(lz, plane) + mul24(ly, bxs) + (P - 1));
    unsigned long long v0 = row[lx], v1 = row[lx + 1];
    if (WRAP) {
      if (lx < P - 1) v0 += row[lx + nx];
      if (lx + 1 < P - 1) v1 += row[lx + 1 + nx];
    }
    float* g = grid + ((x0 + lx) + mul24(nx, y0 + ly) + mul24(nxny, z0 + lz));
    g[0] = from_fixed40(v0);
    if (lx + 1 < wx) g[1] = from_fixed40(v1);
  }
}

// Brick b of one grid (one NT-thread workgroup); smem: the workgroup's
// dynamic LDS (spread_lds_bytes).  Each thread takes one record
// per batch (an atom whose stencil touches the brick), keeps q*theta_x,
// theta_y, theta_z in registers and adds its p^3 stencil into the brick
// (spread_record).
template <int P, int NT>
__device__ __forceinline__ void spread_brick_body(const BinArgs& a, int b, int wxm, int wym, int wzm,
                                                  float* __restrict__ grid, unsigned char* smem) {
  __shared__ int sh_scan[NT / 64];

  const int bxs = spread_xstride(wxm, P);
  unsigned long long* acc = reinterpret_cast<unsigned long long*>(smem);
  const int acc_n = wzm * wym * bxs;
  int* s_pre = reinterpret_cast<int*>(acc + acc_n);  // [nchunk + 1]
  int* s_base = s_pre + a.nchunk + 1;                 // [nchunk]

  const int tid = threadIdx.x;
  const int bxi = b % a.nbx, byi = (b / a.nbx) % a.nby, bzi = b / (a.nbx * a.nby);
  const int x0 = brick_lo(bxi, a.nx, a.nbx), wx = brick_lo(bxi + 1, a.nx, a.nbx) - x0;
  const int y0 = brick_lo(byi, a.ny, a.nby), wy = brick_lo(byi + 1, a.ny, a.nby) - y0;
  const int z0 = brick_lo(bzi, a.nz, a.nbz), wz = brick_lo(bzi + 1, a.nz, a.nbz) - z0;
  const bool wrapx = (a.nbx == 1), wrapy = (a.nby == 1), wrapz = (a.nbz == 1);
  const int plane = mul24(wym, bxs);

  const SegBounds sb = brick_seg_load<false>(a, b);  // in flight while the brick is cleared
  for (int i = tid; i < acc_n; i += NT) acc[i] = 0ull;
  const int total = brick_segments<NT>(a, sb, s_pre, s_base, sh_scan);  // ends with a barrier

  for (int base = 0; base < total; base += NT) {
    // records of a batch in a strided order (stride 17, coprime with the
    // batch size): consecutive records are often atoms of one molecule with
    // the same stencil rows, and integer sums do not depend on the order
    const int nbat = min(NT, total - base);
    if (tid >= nbat) break;
    const int stride = (nbat % 17) ? 17 : ((nbat % 19) ? 19 : 23);
    const int lt = tid * stride;
    const int v = base + (lt - mul24(sdiv(lt, small_div(nbat)), nbat));
    spread_record<P>(a, a.rec_pos[brick_record(a, v, s_pre, s_base)], x0, wx, y0, wy, z0, wz, wrapy, wrapz, bxs,
                     plane, acc);
  }
  __syncthreads();

  // the brick is complete: one coalesced store per grid point, no memset
  // needed; two x-neighbours per thread
  const int hw = (wx + 1) >> 1;
  const int npair = hw * wy * wz;
  if (wrapx)
    store_brick_pairs<P, true, NT>(acc, grid, a.nx, a.ny, x0, y0, z0, wx, wy, hw, npair, bxs, plane);
  else
    store_brick_pairs<P, false, NT>(acc, grid, a.nx, a.ny, x0, y0, z0, wx, wy, hw, npair, bxs, plane);
}

template <int P>
__global__ __launch_bounds__(kSpreadBlock) void k_spread_brick(BinArgs a, int wxm, int wym, int wzm,
                                                               float* __restrict__ grid) {
  extern __shared__ __align__(16) unsigned char smem[];
  spread_brick_body<P, kSpreadBlock>(a, blockIdx.x, wxm, wym, wzm, grid, smem);
}

size_t spread_lds_bytes(int order, const int wmax[3], int nchunk) {
  return sizeof(unsigned long long) * (size_t)wmax[2] * wmax[1] * spread_xstride(wmax[0], order) +
         sizeof(int) * (2 * (size_t)nchunk + 1);
}

hipError_t launch_spread(int order, const BinArgs& a, const int wmax[3], float* grid, hipStream_t s) {
  const int nbins = a.nbx * a.nby * a.nbz;
  const size_t lds = spread_lds_bytes(order, wmax, a.nchunk);
  switch (order) {
    case 4:
      hipLaunchKernelGGL(k_spread_brick<4>, dim3(nbins), dim3(kSpreadBlock), lds, s, a, wmax[0], wmax[1], wmax[2], grid);
      break;
    case 6:
      hipLaunchKernelGGL(k_spread_brick<6>, dim3(nbins), dim3(kSpreadBlock), lds, s, a, wmax[0], wmax[1], wmax[2], grid);
      break;
    case 8:
      hipLaunchKernelGGL(k_spread_brick<8>, dim3(nbins), dim3(kSpreadBlock), lds, s, a, wmax[0], wmax[1], wmax[2], grid);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---- batched: persistent workgroups over the work items of many systems ----

__global__ __launch_bounds__(kBinBlock) void k_batch_bin(const SysDesc* __restrict__ sys,
                                                      const WorkItem* __restrict__ items, int nitems) {
  extern __shared__ int hist[];
  __shared__ double sh_self[kBinBlock / 64];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const WorkItem w = uniform_item(items[it]);
    const SysDesc& d = sys[w.sys];
    bin_chunk_body(d.xyzq, d.natom, d.bin, w.idx, hist);
    if (d.calc_self) {
      // this chunk's share of the self energy (calc_self_energy_kernel,
      // CudaPMERecip.cu:1252-1275): q*q in FP32, summed in FP64
      const int a0 = w.idx * d.bin.chunk, a1 = min(a0 + d.bin.chunk, d.natom);
      double v[1] = {0.0};
      for (int i = a0 + (int)threadIdx.x; i < a1; i += kBinBlock) {
        const float q = d.xyzq[i].w;
        v[0] += (double)(q * q);
      }
      block_sum<1, kBinBlock>(v, sh_self);
      if (threadIdx.x == 0) d.self_partial[w.idx] = v[0];
    }
    __syncthreads();
  }
}

// NT = kSpreadBlock / 2: half-size workgroups, two per CU, for bricks
// planned within half the LDS budget (four waves per SIMD: at most 128 VGPRs)
template <int P, int NT>
__global__ __launch_bounds__(NT, 4) void k_batch_spread(const SysDesc* __restrict__ sys,
                                                     const WorkItem* __restrict__ items, int nitems, int* ctr) {
  extern __shared__ __align__(16) unsigned char smem[];
  for (int it = blockIdx.x; it < nitems; it = next_item(nitems, ctr)) {
    const WorkItem w = uniform_item(items[it]);
    const SysDesc& d = sys[w.sys];
    spread_brick_body<P, NT>(d.bin, w.idx, d.wmax[0], d.wmax[1], d.wmax[2], d.grid, smem);
    __syncthreads();
  }
  items_done(nitems, ctr);
}

hipError_t launch_batch_bin(const SysDesc* sys, const WorkItem* items, int nitems, int grid, size_t lds,
                            hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_bin, dim3(grid), dim3(kBinBlock), lds, s, sys, items, nitems);
  return hipGetLastError();
}

hipError_t launch_batch_spread(int order, int half, const SysDesc* sys, const WorkItem* items, int nitems, int* ctr,
                               int grid, size_t lds, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  constexpr int H = kSpreadBlock / 2;
#define PME_BSPREAD(P)                                                                                           \
  if (half)                                                                                                      \
    hipLaunchKernelGGL((k_batch_spread<P, H>), dim3(grid), dim3(H), lds, s, sys, items, nitems, ctr);            \
  else                                                                                                           \
    hipLaunchKernelGGL((k_batch_spread<P, kSpreadBlock>), dim3(grid), dim3(kSpreadBlock), lds, s, sys, items,    \
                       nitems, ctr);
  switch (order) {
    case 4: PME_BSPREAD(4) break;
    case 6: PME_BSPREAD(6) break;
    case 8: PME_BSPREAD(8) break;
    default: return hipErrorInvalidValue;
  }
#undef PME_BSPREAD
  return hipGetLastError();
}

}  // namespace pme
