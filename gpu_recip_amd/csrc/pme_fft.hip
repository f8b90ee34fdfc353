// pme_fft.hip -- 3-D real-to-complex FFT of the PME grid and the reciprocal
// convolution (scalar sum), hand-written for CDNA4.
//
// Layouts (the reference's BOX layout, CudaPMERecip.cu:1383, 1399-1400):
//   real grid     float [nz][ny][nx]            (x fastest, packed)
//   half spectrum float2[nz][ny][nx/2+1]        (x fastest)
// Forward transforms use e^{-2 pi i}, inverse e^{+2 pi i}, both unnormalised.
//
// Passes: x (R2C / C2R over rows, even nx through a half-length complex FFT),
// y and z over strided columns, tiled as TX consecutive kx so every HBM access
// is a contiguous run.  Every 1-D transform runs in LDS as mixed-radix
// (2/3/4/5/8) Stockham stages.  The z pass can run forward FFT -> influence
// function (energy/virial in FP64) -> inverse FFT in one kernel, which is the
// reference's r2c z-stage + scalar_sum + c2r z-stage in one read and one write
// of the spectrum.
#include "pme_internal.hpp"

namespace pme {

// ---------------------------------------------------------------- x pass ----

// Row tile `tile` (NR rows) of one grid; sm: dynamic LDS (fft_x_lds_bytes).
// Offset of spectrum row r (= ky + ny kz) in the padded half spectrum.
__device__ __forceinline__ size_t spec_row(int r, int ny, int hx, long long pz) {
  const int kz = sdiv(r, small_div(ny));
  return (size_t)kz * (size_t)pz + (size_t)(r - kz * ny) * hx;
}

__device__ __forceinline__ void fft_x_r2c_body(const float* __restrict__ grid, float2* __restrict__ cg, int nx,
                                               int nrows, int ny, int NR, const AxisArgs& ax, int tile, float2* sm) {
  const int tid = threadIdx.x;
  const int nf = ax.rp.n;
  const int LS = NR + 1;
  float2* tw = sm;
  float2* b0 = sm + ax.ntw;
  float2* b1 = b0 + nf * LS;
  for (int i = tid; i < ax.ntw; i += kBlock) tw[i] = ax.tw[i];
  const int row0 = tile * NR;
  const int nr = min(NR, nrows - row0);
  const int hx = nx / 2 + 1;
  const bool even = (nx & 1) == 0;
  if (even) {
    // the tile's rows are contiguous in HBM: keep U loads in flight per thread
    const float2* src = reinterpret_cast<const float2*>(grid + (size_t)row0 * nx);
    const int ntile = NR * nf, nvalid = nr * nf;
    for (int e0 = tid; e0 < ntile; e0 += kStageU * kBlock) {
      float2 v[kStageU];
#pragma unroll
      for (int u = 0; u < kStageU; u++) {
        const int e = e0 + u * kBlock;
        v[u] = (e < nvalid) ? src[e] : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int u = 0; u < kStageU; u++) {
        const int e = e0 + u * kBlock;
        if (e < ntile) {
          const int row = sdiv(e, small_div(nf));
          b0[(e - row * nf) * LS + row] = v[u];
        }
      }
    }
  } else {
    const float* src = grid + (size_t)row0 * nx;
    const int ntile = NR * nx, nvalid = nr * nx;
    for (int e0 = tid; e0 < ntile; e0 += kStageU * kBlock) {
      float v[kStageU];
#pragma unroll
      for (int u = 0; u < kStageU; u++) {
        const int e = e0 + u * kBlock;
        v[u] = (e < nvalid) ? src[e] : 0.f;
      }
#pragma unroll
      for (int u = 0; u < kStageU; u++) {
        const int e = e0 + u * kBlock;
        if (e < ntile) {
          const int row = sdiv(e, small_div(nx));
          b0[(e - row * nx) * LS + row] = make_float2(v[u], 0.f);
        }
      }
    }
  }
  __syncthreads();
  float2* res = fft_lines<-1>(b0, b1, tw, ax.rp, NR, LS, tid, kBlock);
  if (even) {
    // X[k] = E[k] + w^k O[k], E = (Z[k] + Z*[h-k])/2, O = (Z[k] - Z*[h-k])/(2i)
    const int h = nf;
    for (int e = tid; e < nr * hx; e += kBlock) {
      const int row = sdiv(e, small_div(hx));
      const int k = e - row * hx;
      const float2 zk = res[(k == h ? 0 : k) * LS + row];
      const float2 zc = cconj(res[(k == 0 ? 0 : h - k) * LS + row]);
      const float2 E = cscale(cadd(zk, zc), 0.5f);
      const float2 D = csub(zk, zc);
      const float2 O = make_float2(0.5f * D.y, -0.5f * D.x);
      cg[spec_row(row0 + row, ny, hx, spec_pitch(hx, ny)) + k] = cadd(E, cmul(tw[k], O));
    }
  } else {
    for (int e = tid; e < nr * hx; e += kBlock) {
      const int row = sdiv(e, small_div(hx));
      const int k = e - row * hx;
      cg[spec_row(row0 + row, ny, hx, spec_pitch(hx, ny)) + k] = res[k * LS + row];
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_fft_x_r2c(const float* __restrict__ grid, float2* __restrict__ cg, int nx,
                                                      int nrows, int ny, int NR, AxisArgs ax) {
  extern __shared__ __align__(16) float2 sm[];
  fft_x_r2c_body(grid, cg, nx, nrows, ny, NR, ax, blockIdx.x, sm);
}

__device__ __forceinline__ void fft_x_c2r_body(const float2* __restrict__ cg, float* __restrict__ grid, int nx,
                                               int nrows, int ny, int NR, const AxisArgs& ax, int tile, float2* sm) {
  const int tid = threadIdx.x;
  const int nf = ax.rp.n;
  const int LS = NR + 1;
  float2* tw = sm;
  float2* b0 = sm + ax.ntw;
  float2* b1 = b0 + nf * LS;
  for (int i = tid; i < ax.ntw; i += kBlock) tw[i] = ax.tw[i];
  __syncthreads();
  const int row0 = tile * NR;
  const int nr = min(NR, nrows - row0);
  const int hx = nx / 2 + 1;
  const bool even = (nx & 1) == 0;
  if (even) {
    // Z[k] = (X[k] + X*[h-k]) + i w^{-k} (X[k] - X*[h-k]),  k < h
    const int h = nf;
    for (int e0 = tid; e0 < NR * h; e0 += kStageU * kBlock) {
      float2 xk[kStageU], xc[kStageU];
#pragma unroll
      for (int u = 0; u < kStageU; u++) {
        const int e = e0 + u * kBlock;
        const int row = sdiv(e, small_div(h));
        const int k = e - row * h;
        xk[u] = xc[u] = make_float2(0.f, 0.f);
        if (row < nr) {
          const float2* x = cg + spec_row(row0 + row, ny, hx, spec_pitch(hx, ny));
          xk[u] = x[k];
          xc[u] = cconj(x[h - k]);
        }
      }
#pragma unroll
      for (int u = 0; u < kStageU; u++) {
        const int e = e0 + u * kBlock;
        if (e < NR * h) {
          const int row = sdiv(e, small_div(h));
          const int k = e - row * h;
          const float2 A = cadd(xk[u], xc[u]);
          const float2 B = cmul(cconj(tw[k]), csub(xk[u], xc[u]));
          b0[k * LS + row] = make_float2(A.x - B.y, A.y + B.x);
        }
      }
    }
  } else {
    for (int e = tid; e < NR * nx; e += kBlock) {
      const int row = sdiv(e, small_div(nx));
      const int k = e - row * nx;
      float2 z = make_float2(0.f, 0.f);
      if (row < nr) {
        const float2* x = cg + spec_row(row0 + row, ny, hx, spec_pitch(hx, ny));
        z = (k < hx) ? x[k] : cconj(x[nx - k]);
      }
      b0[k * LS + row] = z;
    }
  }
  __syncthreads();
  float2* res = fft_lines<+1>(b0, b1, tw, ax.rp, NR, LS, tid, kBlock);
  if (even) {
    float2* dst = reinterpret_cast<float2*>(grid + (size_t)row0 * nx);
    for (int e = tid; e < nr * nf; e += kBlock) {
      const int row = sdiv(e, small_div(nf));
      const int m = e - row * nf;
      dst[(size_t)row * nf + m] = res[m * LS + row];
    }
  } else {
    for (int e = tid; e < nr * nx; e += kBlock) {
      const int row = sdiv(e, small_div(nx));
      const int m = e - row * nx;
      grid[(size_t)(row0 + row) * nx + m] = res[m * LS + row].x;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_fft_x_c2r(const float2* __restrict__ cg, float* __restrict__ grid, int nx,
                                                      int nrows, int ny, int NR, AxisArgs ax) {
  extern __shared__ __align__(16) float2 sm[];
  fft_x_c2r_body(cg, grid, nx, nrows, ny, NR, ax, blockIdx.x, sm);
}

// ------------------------------------------------------------- xy planes --
// One workgroup transforms np (1 or 2) consecutive z planes through x and y
// in LDS, so each plane is read once and written once with contiguous
// accesses:
//   forward: real rows --x R2C--> half-spectrum rows, transposed in LDS to
//            columns --y FFT--> the planes of the half spectrum;
//   inverse: the reverse (y inverse, C2R pre-processing, x inverse).
// The rows of the np planes are one batch of np*ny x lines (row r of plane p
// is line p*ny + r, the order they have in memory), and their columns one
// batch of np*hx y lines (column k of plane p is line p*hx + k), so two
// planes share every FFT stage and barrier.
// LDS: twiddles of x and y, two buffers of plane_buf() points.  x lines
// have point stride LSX, y lines point stride LSY, both odd so that 8-byte
// accesses of consecutive lines hit distinct bank pairs.
// Odd nx: rows 2j and 2j+1 of a plane are the real and imaginary parts of
// one x line (one complex FFT of length nx transforms two rows; an odd ny
// pairs the plane's last row with zeros), so a plane has (ny + 1) / 2 x
// lines; the pairs never cross planes, so a group of planes rounds exactly
// as its planes one at a time.

__host__ __device__ constexpr int odd_pad(int n) { return (n + 1) | 1; }

// x lines of one plane
__host__ __device__ constexpr int plane_xlines(int nx, int ny) { return (nx & 1) ? (ny + 1) >> 1 : ny; }

__host__ __device__ constexpr int plane_buf(int nx, int ny, int nf, int np) {
  return (nf * odd_pad(np * plane_xlines(nx, ny)) > ny * odd_pad(np * (nx / 2 + 1)))
             ? nf * odd_pad(np * plane_xlines(nx, ny))
             : ny * odd_pad(np * (nx / 2 + 1));
}

__host__ __device__ constexpr size_t plane_lds_need(int nx, int ny, int nf, int np) {
  return sizeof(float2) * ((size_t)nx + ny + 2 * (size_t)plane_buf(nx, ny, nf, np));
}

size_t fft_plane_lds_bytes(int nx, int ny, int nf_x, int np) {
  const size_t b = plane_lds_need(nx, ny, nf_x, np);
  return b <= kPlaneLdsBudget ? b : 0;
}

// LDS budget of one plane workgroup: kPlaneLdsBudget for the full-size
// (kPlaneBlock-thread) workgroup, alone on its CU; half of it for the
// half-size one, two of which share a CU
__host__ __device__ constexpr size_t plane_budget(bool half) { return half ? kPlaneLdsBudget / 2 : kPlaneLdsBudget; }

// planes per workgroup of a batched plane launch allowing up to maxnp: the
// largest group whose buffers fit (the same on host and device)
__host__ __device__ inline int plane_group(int nx, int ny, int nf, int maxnp, bool half) {
  int g = maxnp;
  while (g > 1 && plane_lds_need(nx, ny, nf, g) > plane_budget(half)) g--;
  return g;
}

int plane_group_host(int nx, int ny, int nf, int maxnp, bool half) { return plane_group(nx, ny, nf, maxnp, half); }

bool plane_fits_half(int nx, int ny, int nf) { return plane_lds_need(nx, ny, nf, 1) <= plane_budget(true); }

// planes z .. z+np-1 of the real grid -> the same planes of the half spectrum
// (SPLIT: the y pass may use the split-precision matrix-core stage)
// Thread map of a row-major block of rows of w elements (w <= NT) for loops
// without a division per element: thread t takes column t mod w of rows
// t / w, t / w + step, ... with step = NT / w (threads past step * w idle).
struct ColMap {
  int c, r0, step;
  bool on;
};
template <int NT>
__device__ __forceinline__ ColMap col_map(int w) {
  ColMap m;
  m.step = NT / w;
  m.c = (int)threadIdx.x % w;
  m.r0 = (int)threadIdx.x / w;
  m.on = m.r0 < m.step;
  return m;
}

template <bool SPLIT, int NT>
__device__ __forceinline__ void fft_plane_fwd_body(const float* __restrict__ grid, float2* __restrict__ cg, int nx,
                                                   int ny, const AxisArgs& axx, const AxisArgs& axy, int z, int np,
                                                   float2* sm) {
  const int tid = threadIdx.x;
  const int nf = axx.rp.n;  // x line length: nx/2 (even nx) or nx
  const int hx = nx / 2 + 1;
  const int nrow = np * ny;                   // rows
  const int hyl = plane_xlines(nx, ny);      // x lines per plane
  const int nxl = np * hyl;                   // x lines
  const int ncol = np * hx;                   // y lines
  const int LSX = odd_pad(nxl), LSY = odd_pad(ncol);
  const int bsz = plane_buf(nx, ny, nf, np);
  float2* twx = sm;
  float2* twy = twx + nx;
  float2* b0 = twy + ny;
  float2* b1 = b0 + bsz;
  for (int i = tid; i < nx; i += NT) twx[i] = axx.tw[i];
  for (int i = tid; i < ny; i += NT) twy[i] = axy.tw[i];
  const bool even = (nx & 1) == 0;
  const size_t plane = (size_t)z * nx * ny;
  if (even) {
    const float2* src = reinterpret_cast<const float2*>(grid + plane);
    const ColMap m = col_map<NT>(nf);
    if (m.on)
      for (int rb = m.r0; rb < nrow; rb += kStageU * m.step) {
        float2 v[kStageU];
#pragma unroll
        for (int u = 0; u < kStageU; u++) {
          const int r = rb + u * m.step;
          v[u] = r < nrow ? src[(size_t)r * nf + m.c] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < kStageU; u++) {
          const int r = rb + u * m.step;
          if (r < nrow) b0[m.c * LSX + r] = v[u];
        }
      }
  } else {
    // rows 2j, 2j+1 of plane pl -> real and imaginary parts of x line
    // pl*hyl + j (an odd ny pairs the last row of each plane with zeros)
    const float* src = grid + plane;
    const SmallDiv dny = small_div(ny);
    const ColMap m = col_map<NT>(nx);
    if (m.on)
      for (int rb = m.r0; rb < nrow; rb += kStageU * m.step) {
        float v[kStageU];
#pragma unroll
        for (int u = 0; u < kStageU; u++) {
          const int r = rb + u * m.step;
          v[u] = r < nrow ? src[(size_t)r * nx + m.c] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < kStageU; u++) {
          const int r = rb + u * m.step;
          const int pl = sdiv(r, dny);
          const int y = r - pl * ny;
          if (r < nrow) reinterpret_cast<float*>(b0 + (m.c * LSX + pl * hyl + (y >> 1)))[y & 1] = v[u];
        }
      }
    if ((ny & 1) && m.on)
      for (int pl = m.r0; pl < np; pl += m.step) b0[m.c * LSX + pl * hyl + (ny >> 1)].y = 0.f;
  }
  __syncthreads();
  float2* res = fft_lines<-1>(b0, b1, twx, axx.rp, nxl, LSX, tid, NT);
  float2* q = res == b0 ? b1 : b0;
  // R2C post-processing (even nx) and transpose to columns:
  // X[k] = E[k] + w^k O[k], E = (Z[k] + Z*[h-k])/2, O = (Z[k] - Z*[h-k])/(2i)
  const SmallDiv dny = small_div(ny);
  const ColMap mh = col_map<NT>(hx);
  for (int r = mh.r0; mh.on && r < nrow; r += mh.step) {  // x line r (row of the plane group), point k
    const int k = mh.c;
    const int pl = sdiv(r, dny);
    const int y = r - pl * ny;
    float2 X;
    if (even) {
      const int h = nf;
      const float2 zk = res[(k == h ? 0 : k) * LSX + r];
      const float2 zc = cconj(res[(k == 0 ? 0 : h - k) * LSX + r]);
      const float2 E = cscale(cadd(zk, zc), 0.5f);
      const float2 D = csub(zk, zc);
      const float2 O = make_float2(0.5f * D.y, -0.5f * D.x);
      X = cadd(E, cmul(twx[k], O));
    } else {
      // Z = FFT(a + i b) of the row pair: A[k] = (Z[k] + Z*[nx-k]) / 2,
      // B[k] = (Z[k] - Z*[nx-k]) / (2i); row y is a (y even) or b (y odd)
      const int i = pl * hyl + (y >> 1);
      const float2 zk = res[k * LSX + i];
      const float2 zc = cconj(res[(k == 0 ? 0 : nx - k) * LSX + i]);
      const float2 D = csub(zk, zc);
      X = (y & 1) ? make_float2(0.5f * D.y, -0.5f * D.x) : cscale(cadd(zk, zc), 0.5f);
    }
    q[y * LSY + pl * hx + k] = X;
  }
  __syncthreads();
  float2* r2 = fft_lines<-1, SPLIT>(q, res, twy, axy.rp, ncol, LSY, tid, NT, axy.prec);
  const long long pz = spec_pitch(hx, ny);
  float2* dst = cg + (size_t)z * (size_t)pz;
  for (int r = mh.r0; mh.on && r < nrow; r += mh.step) {
    const int pl = sdiv(r, dny);
    const int y = r - pl * ny;
    dst[(size_t)pl * (size_t)pz + y * hx + mh.c] = r2[y * LSY + pl * hx + mh.c];
  }
}

// planes z .. z+np-1 of the half spectrum -> the same planes of the real grid
template <bool SPLIT, int NT>
__device__ __forceinline__ void fft_plane_inv_body(const float2* __restrict__ cg, float* __restrict__ grid, int nx,
                                                   int ny, const AxisArgs& axx, const AxisArgs& axy, int z, int np,
                                                   float2* sm) {
  const int tid = threadIdx.x;
  const int nf = axx.rp.n;
  const int hx = nx / 2 + 1;
  const int nrow = np * ny;                   // rows
  const int hyl = plane_xlines(nx, ny);      // x lines per plane
  const int nxl = np * hyl;                   // x lines
  const int ncol = np * hx;                   // y lines
  const int LSX = odd_pad(nxl), LSY = odd_pad(ncol);
  const int bsz = plane_buf(nx, ny, nf, np);
  float2* twx = sm;
  float2* twy = twx + nx;
  float2* b0 = twy + ny;
  float2* b1 = b0 + bsz;
  for (int i = tid; i < nx; i += NT) twx[i] = axx.tw[i];
  for (int i = tid; i < ny; i += NT) twy[i] = axy.tw[i];
  const long long pz = spec_pitch(hx, ny);
  const float2* src = cg + (size_t)z * (size_t)pz;
  const SmallDiv dny = small_div(ny);
  const ColMap mh = col_map<NT>(hx);
  if (mh.on)
    for (int rb = mh.r0; rb < nrow; rb += kStageU * mh.step) {
      float2 v[kStageU];
      int at[kStageU];  // LDS column offset of row r: y * LSY + pl * hx
#pragma unroll
      for (int u = 0; u < kStageU; u++) {
        const int r = rb + u * mh.step;
        const int pl = sdiv(r, dny);
        const int y = r - pl * ny;
        at[u] = y * LSY + pl * hx;
        v[u] = r < nrow ? src[(size_t)pl * (size_t)pz + y * hx + mh.c] : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int u = 0; u < kStageU; u++)
        if (rb + u * mh.step < nrow) b0[at[u] + mh.c] = v[u];
    }
  __syncthreads();
  float2* r = fft_lines<+1, SPLIT>(b0, b1, twy, axy.rp, ncol, LSY, tid, NT, axy.prec);
  float2* p = r == b0 ? b1 : b0;
  const bool even = (nx & 1) == 0;
  if (even) {
    // Z[k] = (X[k] + X*[h-k]) + i w^{-k} (X[k] - X*[h-k]),  k < h
    const int h = nf;
    const ColMap m = col_map<NT>(h);
    for (int row = m.r0; m.on && row < nrow; row += m.step) {
      const int k = m.c;
      const int pl = sdiv(row, dny);
      const float2* col = r + (row - pl * ny) * LSY + pl * hx;
      const float2 xk = col[k];
      const float2 xc = cconj(col[h - k]);
      const float2 A = cadd(xk, xc);
      const float2 B = cmul(cconj(twx[k]), csub(xk, xc));
      p[k * LSX + row] = make_float2(A.x - B.y, A.y + B.x);
    }
  } else {
    // x line pl*hyl + j: Z = A + i B from the full Hermitian spectra A, B of
    // rows 2j and 2j+1 of plane pl (zero past its last row); its inverse FFT
    // has row 2j in its real and row 2j+1 in its imaginary part
    const SmallDiv dhyl = small_div(hyl);
    const ColMap m = col_map<NT>(nx);
    for (int i = m.r0; m.on && i < nxl; i += m.step) {
      const int k = m.c;
      const int kk = k < hx ? k : nx - k;
      const int pl = sdiv(i, dhyl);
      const int ya = 2 * (i - pl * hyl);
      const float2* col = r + pl * hx + kk;
      float2 A = col[ya * LSY];
      float2 B = make_float2(0.f, 0.f);
      if (ya + 1 < ny) B = col[(ya + 1) * LSY];
      if (k >= hx) {
        A = cconj(A);
        B = cconj(B);
      }
      p[k * LSX + i] = make_float2(A.x - B.y, A.y + B.x);
    }
  }
  __syncthreads();
  float2* s = fft_lines<+1>(p, r, twx, axx.rp, nxl, LSX, tid, NT);
  const size_t plane = (size_t)z * nx * ny;
  if (even) {
    float2* dst = reinterpret_cast<float2*>(grid + plane);
    const ColMap m = col_map<NT>(nf);
    for (int row = m.r0; m.on && row < nrow; row += m.step) dst[(size_t)row * nf + m.c] = s[m.c * LSX + row];
  } else {
    float* dst = grid + plane;
    const ColMap m = col_map<NT>(nx);
    for (int row = m.r0; m.on && row < nrow; row += m.step) {
      const int pl = sdiv(row, dny);
      const int y = row - pl * ny;
      dst[(size_t)row * nx + m.c] = reinterpret_cast<const float*>(s + (m.c * LSX + pl * hyl + (y >> 1)))[y & 1];
    }
  }
}

template <bool SPLIT>
__global__ __launch_bounds__(kPlaneBlock) void k_fft_plane_fwd(const float* __restrict__ grid, float2* __restrict__ cg,
                                                              int nx, int ny, AxisArgs axx, AxisArgs axy) {
  extern __shared__ __align__(16) float2 sm[];
  fft_plane_fwd_body<SPLIT, kPlaneBlock>(grid, cg, nx, ny, axx, axy, blockIdx.x, 1, sm);
}

// Workgroup z < nz transforms plane z; with a pending energy/virial
// reduction the launch has one more workgroup that performs it alone (on a
// CU the planes leave free), so no plane's workgroup waits for it.
template <bool SPLIT>
__global__ __launch_bounds__(kPlaneBlock) void k_fft_plane_inv(const float2* __restrict__ cg, float* __restrict__ grid,
                                                              int nx, int ny, int nz, AxisArgs axx, AxisArgs axy,
                                                              EvReduce red) {
  extern __shared__ __align__(16) float2 sm[];
  if ((int)blockIdx.x >= nz) {
    ev_reduce_block<kPlaneBlock>(red);
    return;
  }
  fft_plane_inv_body<SPLIT, kPlaneBlock>(cg, grid, nx, ny, axx, axy, blockIdx.x, 1, sm);
}

hipError_t launch_fft_plane_fwd(const float* grid, float2* cg, int nx, int ny, int nz, const AxisArgs& ax,
                                const AxisArgs& ay, hipStream_t s) {
  const size_t lds = fft_plane_lds_bytes(nx, ny, ax.rp.n, 1);
  if (!lds) return hipErrorInvalidValue;
  if (ay.prec != PREC_FP32)
    hipLaunchKernelGGL(k_fft_plane_fwd<true>, dim3(nz), dim3(kPlaneBlock), lds, s, grid, cg, nx, ny, ax, ay);
  else
    hipLaunchKernelGGL(k_fft_plane_fwd<false>, dim3(nz), dim3(kPlaneBlock), lds, s, grid, cg, nx, ny, ax, ay);
  return hipGetLastError();
}

hipError_t launch_fft_plane_inv(const float2* cg, float* grid, int nx, int ny, int nz, const AxisArgs& ax,
                                const AxisArgs& ay, const EvReduce& red, hipStream_t s) {
  const size_t lds = fft_plane_lds_bytes(nx, ny, ax.rp.n, 1);
  if (!lds) return hipErrorInvalidValue;
  const dim3 g(nz + ((red.nsolve > 0 || red.nself > 0) ? 1 : 0));
  if (ay.prec != PREC_FP32)
    hipLaunchKernelGGL(k_fft_plane_inv<true>, g, dim3(kPlaneBlock), lds, s, cg, grid, nx, ny, nz, ax, ay, red);
  else
    hipLaunchKernelGGL(k_fft_plane_inv<false>, g, dim3(kPlaneBlock), lds, s, cg, grid, nx, ny, nz, ax, ay, red);
  return hipGetLastError();
}

// ------------------------------------------------------- influence function --

// eterm at (k1, k2, k3) and the energy/virial terms of that point,
// scalar_sum_ortho_kernel CudaPMERecip.cu:671-704
__device__ __forceinline__ float eterm_at(const SolveArgs& s, int k1, int k2, int k3, float& mh1, float& mh2,
                                          float& mh3, float& msq_inv) {
  const int m1 = k1 >= s.nf1 ? k1 - s.n1 : k1;
  const int m2 = k2 >= s.nf2 ? k2 - s.n2 : k2;
  const int m3 = k3 >= s.nf3 ? k3 - s.n3 : k3;
  mh1 = s.r1 * (float)m1;
  mh2 = s.r2 * (float)m2;
  mh3 = s.r3 * (float)m3;
  // grouped as in solve_lines, which takes the (k1, k2) factors out of its
  // loop over k3, so both give the same value; the fused multiply-adds are
  // explicit so that every kernel rounds them alike
  const float s12 = fmaf(mh2, mh2, mh1 * mh1);
  const float msq = fmaf(mh3, mh3, s12);
  const float p12 = s.piv_inv * s.pf1[k1] * s.pf2[k2];
  // hardware reciprocal and exponential (1-2 ulp), as the reference's fast-math build
  msq_inv = __builtin_amdgcn_rcpf(msq);
  return __expf(-s.fac * msq) * p12 * s.pf3[k3] * msq_inv;
}

// Multiply one half-spectrum point by eterm; accumulate energy/virial with
// the reference's weights (mirror term for every k1 >= 1, :708-741).  The
// mirror point (n1-k1, -k2, -k3) has the same |m|^2 and, bsp_mod being
// symmetric, the same eterm; its m components are -m except on a Nyquist
// plane of an even axis (m = -n/2 maps to itself), so the diagonal terms
// double and each off-diagonal term m_a m_b doubles or cancels.
__device__ __forceinline__ float2 solve_point(const SolveArgs& s, float2 q, int k1, int k2, int k3, double* acc) {
  if ((k1 | k2 | k3) == 0) return make_float2(0.f, 0.f);
  float mh1, mh2, mh3, msq_inv;
  const float eterm = eterm_at(s, k1, k2, k3, mh1, mh2, mh3, msq_inv);
  if (s.calc_ev) {
    const float q2 = fmaf(q.y, q.y, q.x * q.x);
    const float tmp1 = eterm * q2;
    const float vterm = 2.0f * (s.fac + msq_inv);
    const float tmp2 = tmp1 * vterm;
    // weight of the point (1) plus its mirror (1 when k1 >= 1)
    const double w = k1 >= 1 ? 2.0 : 1.0;
    // mirror sign of each m component: +1 when it maps to itself
    const bool f1 = !(2 * k1 == s.n1), f2 = !(2 * k2 == s.n2 || k2 == 0), f3 = !(2 * k3 == s.n3 || k3 == 0);
    const double w12 = k1 >= 1 ? (f1 == f2 ? 2.0 : 0.0) : 1.0;
    const double w13 = k1 >= 1 ? (f1 == f3 ? 2.0 : 0.0) : 1.0;
    const double w23 = k1 >= 1 ? (f2 == f3 ? 2.0 : 0.0) : 1.0;
    acc[0] += w * (double)tmp1;
    acc[1] += w * (double)(tmp1 * (vterm * mh1 * mh1 - 1.0f));
    acc[2] += w12 * (double)(tmp2 * mh1 * mh2);
    acc[3] += w13 * (double)(tmp2 * mh1 * mh3);
    acc[4] += w * (double)(tmp1 * (vterm * mh2 * mh2 - 1.0f));
    acc[5] += w23 * (double)(tmp2 * mh2 * mh3);
    acc[6] += w * (double)(tmp1 * (vterm * mh3 * mh3 - 1.0f));
  }
  return cscale(q, eterm);
}

// eterm of a triclinic box at mapped wave numbers (m1, m2, m3), k3 = its z
// index: m-hat = m1 R[0][:] + m2 R[1][:] + m3 R[2][:] (rows of recip).
__device__ __forceinline__ float eterm_tric(const SolveArgs& s, int k1, int k2, int k3, int m1, int m2, int m3,
                                            float* mh, float& msq_inv) {
  const float f1 = (float)m1, f2 = (float)m2, f3 = (float)m3;
  mh[0] = fmaf(s.rt[6], f3, fmaf(s.rt[3], f2, s.rt[0] * f1));
  mh[1] = fmaf(s.rt[7], f3, fmaf(s.rt[4], f2, s.rt[1] * f1));
  mh[2] = fmaf(s.rt[8], f3, fmaf(s.rt[5], f2, s.rt[2] * f1));
  const float msq = fmaf(mh[2], mh[2], fmaf(mh[1], mh[1], mh[0] * mh[0]));
  msq_inv = __builtin_amdgcn_rcpf(msq);
  return __expf(-s.fac * msq) * (s.piv_inv * s.pf1[k1] * s.pf2[k2]) * s.pf3[k3] * msq_inv;
}

// energy and virial terms of one point (weight 1), the order of acc as in solve_point
__device__ __forceinline__ void ev_terms(const SolveArgs& s, float eterm, float q2, const float* mh, float msq_inv,
                                         double* acc) {
  const float tmp1 = eterm * q2;
  const float vterm = 2.0f * (s.fac + msq_inv);
  const float tmp2 = tmp1 * vterm;
  acc[0] += (double)tmp1;
  acc[1] += (double)(tmp1 * (vterm * mh[0] * mh[0] - 1.0f));
  acc[2] += (double)(tmp2 * mh[0] * mh[1]);
  acc[3] += (double)(tmp2 * mh[0] * mh[2]);
  acc[4] += (double)(tmp1 * (vterm * mh[1] * mh[1] - 1.0f));
  acc[5] += (double)(tmp2 * mh[1] * mh[2]);
  acc[6] += (double)(tmp1 * (vterm * mh[2] * mh[2] - 1.0f));
}

// solve_point for a triclinic box: the same weights (the point, plus its
// mirror (n1-k1, -k2, -k3) for every k1 >= 1), but the mirror's m-hat is not
// minus the point's on a Nyquist plane, so its eterm and virial terms are
// evaluated from its own wave numbers.
__device__ __forceinline__ float2 solve_point_tric(const SolveArgs& s, float2 q, int k1, int k2, int k3, double* acc) {
  if ((k1 | k2 | k3) == 0) return make_float2(0.f, 0.f);
  const int m1 = k1 >= s.nf1 ? k1 - s.n1 : k1;
  const int m2 = k2 >= s.nf2 ? k2 - s.n2 : k2;
  const int m3 = k3 >= s.nf3 ? k3 - s.n3 : k3;
  float mh[3], msq_inv;
  const float eterm = eterm_tric(s, k1, k2, k3, m1, m2, m3, mh, msq_inv);
  if (s.calc_ev) {
    const float q2 = fmaf(q.y, q.y, q.x * q.x);
    ev_terms(s, eterm, q2, mh, msq_inv, acc);
    if (k1 >= 1) {
      const int k1s = s.n1 - k1, k2s = k2 ? s.n2 - k2 : 0, k3s = k3 ? s.n3 - k3 : 0;
      const int m1s = k1s >= s.nf1 ? k1s - s.n1 : k1s;
      const int m2s = k2s >= s.nf2 ? k2s - s.n2 : k2s;
      const int m3s = k3s >= s.nf3 ? k3s - s.n3 : k3s;
      float ms[3], msq_inv_s;
      const float eterm_s = eterm_tric(s, k1s, k2s, k3s, m1s, m2s, m3s, ms, msq_inv_s);
      ev_terms(s, eterm_s, q2, ms, msq_inv_s, acc);
    }
  }
  return cscale(q, eterm);
}

// The influence function over the z lines of a column tile (in LDS, point j
// of line l at res[j * TX + l]; lines l < nl are kx = kx0 + l at ky = k2).
// Thread t owns line t mod TX and visits k3 = t / TX, t / TX + kBlock / TX,
// ..., so the (k1, k2) factors of eterm and of the energy/virial terms are
// formed once per thread: scalar_sum_ortho_kernel CudaPMERecip.cu:671-747
// per point, with the sums over k3 of one line taken before the line's
// constant factors are applied.  Per line, with the point weight w (2 when
// k1 >= 1: the point and its mirror), tmp1 = eterm |Q|^2, tmp2 = tmp1 vterm:
//   E   += w sum tmp1                  V11 += w (mh1^2 sum tmp2 - sum tmp1)
//   V22 += w (mh2^2 sum tmp2 - sum tmp1)  V33 += w (sum tmp2 mh3^2 - sum tmp1)
//   V12 += w12 mh1 mh2 sum tmp2      V13 += mh1 sum w13 tmp2 mh3
//   V23 += mh2 sum w23 tmp2 mh3
// (the off-diagonal mirror weights are those of solve_point; w13 and w23
// depend on k3 through the zero/Nyquist test of m3).
__device__ __forceinline__ void solve_lines(const SolveArgs& s, float2* __restrict__ res, int n, int TX, int nl,
                                            int c0, int hxd, double* acc) {
  const int tid = threadIdx.x;
  const SmallDiv dtx = small_div(TX);
  const int j0 = sdiv(tid, dtx);
  const int l = tid - j0 * TX;
  const int jstep = sdiv(kBlock, dtx);  // >= 4: column tiles are at most 64 wide
  static_assert(kBlock >= 4 * 64, "solve_lines: every line needs a thread");
  if (j0 >= jstep || l >= nl) return;
  // line c0 + l of the z pass: flattened (ky, kx) of the z plane, rows of hxd
  const int c = c0 + l;
  const int k2 = sdiv(c, small_div(hxd));
  const int k1 = c - k2 * hxd;
  const int m1 = k1 >= s.nf1 ? k1 - s.n1 : k1;
  const int m2 = k2 >= s.nf2 ? k2 - s.n2 : k2;
  const float mh1 = s.r1 * (float)m1;
  const float mh2 = s.r2 * (float)m2;
  const float s12 = fmaf(mh2, mh2, mh1 * mh1);
  const float p12 = s.piv_inv * s.pf1[k1] * s.pf2[k2];
  const bool f1 = !(2 * k1 == s.n1), f2 = !(2 * k2 == s.n2 || k2 == 0);
  const bool k1pos = k1 >= 1;
  double sA = 0.0, sB = 0.0, sC = 0.0, sD13 = 0.0, sD23 = 0.0;
  for (int j = j0; j < n; j += jstep) {
    float2* pt = res + mul24(j, TX) + l;
    if ((k1 | k2 | j) == 0) {
      *pt = make_float2(0.f, 0.f);
      continue;
    }
    const int m3 = j >= s.nf3 ? j - s.n3 : j;
    const float mh3 = s.r3 * (float)m3;
    const float msq = fmaf(mh3, mh3, s12);
    const float msq_inv = __builtin_amdgcn_rcpf(msq);
    const float eterm = __expf(-s.fac * msq) * p12 * s.pf3[j] * msq_inv;
    const float2 q = *pt;
    if (s.calc_ev) {
      const float q2 = fmaf(q.y, q.y, q.x * q.x);
      const float tmp1 = eterm * q2;
      const float vterm = 2.0f * (s.fac + msq_inv);
      const float tmp2 = tmp1 * vterm;
      const float t3 = tmp2 * mh3;
      const bool f3 = !(2 * j == s.n3 || j == 0);
      const float w13 = k1pos ? (f1 == f3 ? 2.f : 0.f) : 1.f;
      const float w23 = k1pos ? (f2 == f3 ? 2.f : 0.f) : 1.f;
      sA += (double)tmp1;
      sB += (double)tmp2;
      sC += (double)(t3 * mh3);
      sD13 += (double)(w13 * t3);
      sD23 += (double)(w23 * t3);
    }
    *pt = cscale(q, eterm);
  }
  if (s.calc_ev) {
    const double w = k1pos ? 2.0 : 1.0;
    const double w12 = k1pos ? (f1 == f2 ? 2.0 : 0.0) : 1.0;
    const double a1 = (double)mh1, a2 = (double)mh2;
    acc[0] += w * sA;
    acc[1] += w * (a1 * a1 * sB - sA);
    acc[2] += w12 * a1 * a2 * sB;
    acc[3] += a1 * sD13;
    acc[4] += w * (a2 * a2 * sB - sA);
    acc[5] += a2 * sD23;
    acc[6] += w * (sC - sA);
  }
}

// Per-workgroup energy/virial partial (fixed summation order inside the
// workgroup).  The total is formed later by ev_reduce_block in one workgroup.
__device__ __forceinline__ void ev_partial(const SolveArgs& s, double* acc, double* sh, int blk) {
  block_sum<7>(acc, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < 7; c++) s.partial[(size_t)blk * 7 + c] = acc[c];
  }
}

// ---- z pass as a register four-step FFT -------------------------------------
// n = N1 N2, j = j2 + N2 j1, k = k1 + N1 k2:
//   X[k1 + N1 k2] = sum_j2 w_N2^{j2 k2} w_n^{j2 k1} sum_j1 w_N1^{j1 k1} x[j2 + N2 j1].
// Work unit (line l, q): in the forward half q = j2 (an N1-point DFT of the
// column j2 + N2 j1 in registers, the twiddle w_n^{j2 k1}, one LDS
// exchange), then q = k1 (an N2-point DFT in registers).  That thread then
// holds the whole line's spectrum at k3 = k1 + N1 k2, applies the influence
// function there and runs the inverse the same way back (the conjugate
// steps in reverse order).  Two LDS exchanges per direction instead of one
// per radix stage, and no index arithmetic inside the DFTs.
#define PME_F4_PAIRS(X)                                                                             \
  X(0, 4, 4) X(1, 4, 5) X(2, 4, 6) X(3, 5, 5) X(4, 3, 9) X(5, 5, 6) X(6, 4, 8) X(7, 6, 6) X(8, 5, 8) \
  X(9, 5, 9) X(10, 6, 8) X(11, 5, 10) X(12, 6, 9) X(13, 6, 10) X(14, 8, 8) X(15, 8, 9) X(16, 5, 15)  \
  X(17, 8, 10) X(18, 9, 9) X(19, 9, 10) X(20, 8, 12) X(21, 10, 10) X(22, 9, 12) X(23, 10, 12)        \
  X(24, 8, 16) X(25, 9, 15) X(26, 12, 12) X(27, 10, 15) X(28, 10, 16) X(29, 12, 15) X(30, 12, 16)

int four_step_code(int n) {
#define PME_F4_CODE(c, a, b) \
  if (n == (a) * (b)) return c;
  PME_F4_PAIRS(PME_F4_CODE)
#undef PME_F4_CODE
  return -1;
}

// Register class of a factor pair.  A kernel that dispatches over several
// pairs is allocated the registers of their union, which grows well beyond
// the largest single pair (all 31 pairs and the Stockham stages: 157 VGPRs,
// three waves per SIMD; one pair: 72 to 118).  The batched z pass therefore
// has one kernel per class: class 0 holds the pairs with N2 <= 10 and 5 x 15
// (75, a common edge; held to 128 VGPRs, four waves per SIMD, with three
// spilled), class 1 the rest (125 VGPRs).
constexpr int f4_class_of(int n1, int n2) { return (n2 <= 10 || (n1 == 5 && n2 == 15)) ? 0 : 1; }
int four_step_code_for_plan(const RadixPlan& rp) {
  const int code = four_step_code(rp.n);
  int a = 0, b = 0;
  return (rp.f4a > 0 && code >= 0 && four_step_pair(code, &a, &b) && a == rp.f4a && b == rp.f4b) ? code : -1;
}

bool four_step_pair(int code, int* n1, int* n2) {
#define PME_F4_PAIR(c, a, b) \
  if (code == (c)) {         \
    *n1 = (a);               \
    *n2 = (b);               \
    return true;             \
  }
  PME_F4_PAIRS(PME_F4_PAIR)
#undef PME_F4_PAIR
  return false;
}

int four_step_class(int code) {
#define PME_F4_CLS(c, a, b) \
  if (code == (c)) return f4_class_of(a, b);
  PME_F4_PAIRS(PME_F4_CLS)
#undef PME_F4_CLS
  return -1;
}

// eterm of one point of a z line; (k1, k2) factors precomputed (as in
// solve_lines, so the value is the same)
struct LineFactors {
  float mh1, mh2, s12, p12;
  bool f1, f2, k1pos;
  int k1, k2;
};
__device__ __forceinline__ LineFactors line_factors(const SolveArgs& s, int k1, int k2) {
  LineFactors f;
  const int m1 = k1 >= s.nf1 ? k1 - s.n1 : k1;
  const int m2 = k2 >= s.nf2 ? k2 - s.n2 : k2;
  f.mh1 = s.r1 * (float)m1;
  f.mh2 = s.r2 * (float)m2;
  f.s12 = fmaf(f.mh2, f.mh2, f.mh1 * f.mh1);
  f.p12 = s.piv_inv * s.pf1[k1] * s.pf2[k2];
  f.f1 = !(2 * k1 == s.n1);
  f.f2 = !(2 * k2 == s.n2 || k2 == 0);
  f.k1pos = k1 >= 1;
  f.k1 = k1;
  f.k2 = k2;
  return f;
}

// Q <- eterm Q at k3 of the line; energy/virial sums as in solve_lines
__device__ __forceinline__ float2 solve_line_point(const SolveArgs& s, const LineFactors& f, int k3, float2 q,
                                                   double* sums) {
  if ((f.k1 | f.k2 | k3) == 0) return make_float2(0.f, 0.f);
  const int m3 = k3 >= s.nf3 ? k3 - s.n3 : k3;
  const float mh3 = s.r3 * (float)m3;
  const float msq = fmaf(mh3, mh3, f.s12);
  const float msq_inv = __builtin_amdgcn_rcpf(msq);
  const float eterm = __expf(-s.fac * msq) * f.p12 * s.pf3[k3] * msq_inv;
  if (s.calc_ev) {
    const float q2 = fmaf(q.y, q.y, q.x * q.x);
    const float tmp1 = eterm * q2;
    const float vterm = 2.0f * (s.fac + msq_inv);
    const float tmp2 = tmp1 * vterm;
    const float t3 = tmp2 * mh3;
    const bool f3 = !(2 * k3 == s.n3 || k3 == 0);
    const float w13 = f.k1pos ? (f.f1 == f3 ? 2.f : 0.f) : 1.f;
    const float w23 = f.k1pos ? (f.f2 == f3 ? 2.f : 0.f) : 1.f;
    sums[0] += (double)tmp1;
    sums[1] += (double)tmp2;
    sums[2] += (double)(t3 * mh3);
    sums[3] += (double)(w13 * t3);
    sums[4] += (double)(w23 * t3);
  }
  return cscale(q, eterm);
}

__device__ __forceinline__ void fold_line_sums(const LineFactors& f, const double* sums, double* acc) {
  const double w = f.k1pos ? 2.0 : 1.0;
  const double w12 = f.k1pos ? (f.f1 == f.f2 ? 2.0 : 0.0) : 1.0;
  const double a1 = (double)f.mh1, a2 = (double)f.mh2;
  acc[0] += w * sums[0];
  acc[1] += w * (a1 * a1 * sums[1] - sums[0]);
  acc[2] += w12 * a1 * a2 * sums[1];
  acc[3] += a1 * sums[3];
  acc[4] += w * (a2 * a2 * sums[1] - sums[0]);
  acc[5] += a2 * sums[4];
  acc[6] += w * (sums[2] - sums[0]);
}

// One column tile of the z pass (as fft_cols_body<COL_FWD_SOLVE_INV>) with
// n = N1 N2: TX consecutive lines c0 .. c0+TX-1 of the flattened (ky, kx)
// index of a z plane (rows of hxd), point j of line l at base + j pt_stride
// + l.  sm: twiddles (ntw) and ONE n*TX buffer b: the first half reads its
// columns straight from HBM into registers and writes b at (k1 N2 + j2); the
// middle pass reads and rewrites the same N2 points of b per unit (in place);
// the last half reads b and stores straight to HBM.  Lanes run along l, so
// every HBM access of a wave is two runs of TX consecutive float2.
template <int N1, int N2>
__device__ __forceinline__ void zsolve_four_step(float2* __restrict__ cg, int hx, int TX, long long pt_stride,
                                                 long long outer_stride, const AxisArgs& ax, const SolveArgs& sa,
                                                 int tx, int ty, int blk, float2* sm, int hxd) {
  __shared__ double sh_red[7 * (kBlock / 64)];
  const int tid = threadIdx.x;
  float2* tw = sm;
  float2* b = sm + ax.ntw;
  for (int i = tid; i < ax.ntw; i += kBlock) tw[i] = ax.tw[i];
  const int ts = ax.rp.tstride;
  const int kx0 = tx * TX;
  const int nl = min(TX, hx - kx0);
  // the tile's first point (uniform) and 32-bit point offsets from it: a z
  // line spans n * pt_stride < 2^27 points (edges <= 512)
  float2* __restrict__ tile = cg + ((size_t)kx0 + (size_t)ty * (size_t)outer_stride);
  const unsigned pst = (unsigned)pt_stride;
  const int tsh = __builtin_ctz(TX);  // TX is a power of two (pick_fft_tiles)
  // forward, first half: unit (l, j2), column j2 + N2 j1 loaded from HBM
  for (int u0 = 0; u0 < N2 * TX; u0 += kBlock) {
    const int u = u0 + tid;
    const int j2 = (u >> tsh);
    const int l = u - j2 * TX;
    const bool live = u < N2 * TX;
    float2 x[N1];
    const unsigned o = (unsigned)j2 * pst + (unsigned)l;
#pragma unroll
    for (int j1 = 0; j1 < N1; j1++)
      x[j1] = (live && l < nl) ? tile[o + (unsigned)(N2 * j1) * pst] : make_float2(0.f, 0.f);
    if (u0 == 0) __syncthreads();  // twiddles in LDS (uniform branch)
    if (live) {
      DftN<N1, -1>::run(x);
#pragma unroll
      for (int k1 = 0; k1 < N1; k1++) {
        const float2 y = k1 == 0 ? x[0] : cmul(x[k1], tw[mul24(j2 * k1, ts)]);
        b[(k1 * N2 + j2) * TX + l] = y;
      }
    }
  }
  __syncthreads();
  // forward second half, influence function, inverse first half: unit (l, k1),
  // in place on the N2 points (k1 N2 + j2) of b
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int u = tid; u < N1 * TX; u += kBlock) {
    const int k1 = (u >> tsh);
    const int l = u - k1 * TX;
    float2 X[N2];
#pragma unroll
    for (int j2 = 0; j2 < N2; j2++) X[j2] = b[(k1 * N2 + j2) * TX + l];
    DftN<N2, -1>::run(X);  // X[k2]: point k3 = k1 + N1 k2
    if (l < nl) {
      const int ky = sdiv(kx0 + l, small_div(hxd));
      const LineFactors f = line_factors(sa, kx0 + l - ky * hxd, ky);
      double sums[5] = {0, 0, 0, 0, 0};
#pragma unroll
      for (int k2 = 0; k2 < N2; k2++) X[k2] = solve_line_point(sa, f, k1 + N1 * k2, X[k2], sums);
      if (sa.calc_ev) fold_line_sums(f, sums, acc);
    }
    DftN<N2, +1>::run(X);  // X[j2]
#pragma unroll
    for (int j2 = 0; j2 < N2; j2++) {
      const float2 t = tw[mul24(j2 * k1, ts)];
      const float2 z = j2 == 0 ? X[0] : cmul(X[j2], make_float2(t.x, -t.y));
      b[(k1 * N2 + j2) * TX + l] = z;
    }
  }
  __syncthreads();
  // inverse second half: unit (l, j2), stored straight to HBM
  for (int u = tid; u < N2 * TX; u += kBlock) {
    const int j2 = (u >> tsh);
    const int l = u - j2 * TX;
    float2 z[N1];
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) z[k1] = b[(k1 * N2 + j2) * TX + l];
    DftN<N1, +1>::run(z);  // z[j1]: point j = j2 + N2 j1
    if (l < nl) {
      const unsigned o = (unsigned)j2 * pst + (unsigned)l;
#pragma unroll
      for (int j1 = 0; j1 < N1; j1++) tile[o + (unsigned)(N2 * j1) * pst] = z[j1];
    }
  }
  if (sa.calc_ev) ev_partial(sa, acc, sh_red, blk);
}

// the pairs of register class CLS only
template <int CLS>
__device__ __forceinline__ void zsolve_four_step_dispatch(int code, float2* __restrict__ cg, int hx, int TX,
                                                          long long pt_stride, long long outer_stride,
                                                          const AxisArgs& ax, const SolveArgs& sa, int tx, int ty,
                                                          int blk, float2* sm, int hxd) {
  switch (code) {
#define PME_F4_CASE(c, a, b)                                                                          \
  case c:                                                                                             \
    if constexpr (f4_class_of(a, b) == CLS)                                                              \
      zsolve_four_step<a, b>(cg, hx, TX, pt_stride, outer_stride, ax, sa, tx, ty, blk, sm, hxd);      \
    break;
    PME_F4_PAIRS(PME_F4_CASE)
#undef PME_F4_CASE
    default: break;
  }
}

// One tile of a y pass (S = -1 forward, +1 inverse) as a register four-step
// with n = N1 N2 in ONE LDS buffer: the first half loads its columns
// j2 + N2 j1 from HBM into registers, the second half reads its N2 points
// (k1 N2 + j2) back and stores X[k1 + N1 k2] straight to HBM (one LDS round
// trip instead of a staging copy and two).  Tile (tx, ty): lines
// kx0 = tx*TX ... of the hx lines at outer index ty.
template <int N1, int N2, int S>
__device__ __forceinline__ void cols_four_step(float2* __restrict__ cg, int hx, int TX, long long pt_stride,
                                               long long outer_stride, const AxisArgs& ax, int tx, int ty,
                                               float2* sm) {
  const int tid = threadIdx.x;
  float2* tw = sm;
  float2* b = sm + ax.ntw;
  for (int i = tid; i < ax.ntw; i += kBlock) tw[i] = ax.tw[i];
  const int ts = ax.rp.tstride;
  const int kx0 = tx * TX;
  const int nl = min(TX, hx - kx0);
  // the tile's first point (uniform) and 32-bit point offsets from it: a y
  // line spans n * pt_stride < 2^18 points (edges <= 512, pt_stride = hx)
  float2* __restrict__ tile = cg + ((size_t)kx0 + (size_t)ty * (size_t)outer_stride);
  const unsigned pst = (unsigned)pt_stride;
  const SmallDiv dtx = small_div(TX);
  for (int u0 = 0; u0 < N2 * TX; u0 += kBlock) {
    const int u = u0 + tid;
    const int j2 = sdiv(u, dtx);
    const int l = u - j2 * TX;
    const bool live = u < N2 * TX;
    float2 x[N1];
    const unsigned o = (unsigned)j2 * pst + (unsigned)l;
#pragma unroll
    for (int j1 = 0; j1 < N1; j1++)
      x[j1] = (live && l < nl) ? tile[o + (unsigned)(N2 * j1) * pst] : make_float2(0.f, 0.f);
    if (u0 == 0) __syncthreads();  // twiddles in LDS (uniform branch)
    if (live) {
      DftN<N1, S>::run(x);
      const int tj = mul24(j2, ts);
      b[j2 * TX + l] = x[0];
#pragma unroll
      for (int k1 = 1; k1 < N1; k1++) {
        float2 w = tw[k1 * tj];
        if (S > 0) w.y = -w.y;
        b[(k1 * N2 + j2) * TX + l] = cmul(x[k1], w);
      }
    }
  }
  __syncthreads();
  for (int u = tid; u < N1 * TX; u += kBlock) {
    const int k1 = sdiv(u, dtx);
    const int l = u - k1 * TX;
    float2 X[N2];
#pragma unroll
    for (int j2 = 0; j2 < N2; j2++) X[j2] = b[(k1 * N2 + j2) * TX + l];
    DftN<N2, S>::run(X);
    if (l < nl) {
      const unsigned o = (unsigned)k1 * pst + (unsigned)l;
#pragma unroll
      for (int k2 = 0; k2 < N2; k2++) tile[o + (unsigned)(N1 * k2) * pst] = X[k2];
    }
  }
}

// the y pass of a tile through cols_four_step for the pair of code
template <int S>
__device__ __forceinline__ void cols_four_step_dispatch(int code, float2* __restrict__ cg, int hx, int TX,
                                                        long long pt_stride, long long outer_stride,
                                                        const AxisArgs& ax, int tx, int ty, float2* sm) {
  switch (code) {
#define PME_YC_CASE(c, a, b) \
  case c: cols_four_step<a, b, S>(cg, hx, TX, pt_stride, outer_stride, ax, tx, ty, sm); break;
    PME_F4_PAIRS(PME_YC_CASE)
#undef PME_YC_CASE
    default: break;
  }
}

// The fused z pass of one grid as a register four-step with n = N1 N2 (one
// kernel per pair, so each gets the registers of its own pair only): column
// tiles of TX lines of the flattened (ky, kx) of a z plane.
template <int N1, int N2>
__global__ __launch_bounds__(kBlock) void k_zsolve(float2* __restrict__ cg, int nlines, int TX, long long pt_stride,
                                                   AxisArgs ax, SolveArgs sa, EvReduce red, int hxd) {
  extern __shared__ __align__(16) float2 sm[];
  if ((red.nsolve > 0 || red.nself > 0) && blockIdx.x == 0) ev_reduce_block<kBlock>(red);
  zsolve_four_step<N1, N2>(cg, nlines, TX, pt_stride, 0, ax, sa, blockIdx.x, 0, blockIdx.x, sm, hxd);
}


// ------------------------------------------------------------ y / z passes --

// Column tile (tx, ty) of one grid: lines kx0 = tx*TX ... of the hx lines at
// outer index ty.  The y pass: hx = nx/2 + 1 columns per (kz = ty).  The z
// pass: hx = the hxd * ny lines of a z plane, flattened (ky, kx) with rows of
// hxd = nx/2 + 1, and ty = 0.  blk indexes the energy/virial partial of the
// tile.
template <int MODE, bool SPLIT>
__device__ __forceinline__ void fft_cols_body(float2* __restrict__ cg, int n, int hx, int TX, long long pt_stride,
                                              long long outer_stride, const AxisArgs& ax, const SolveArgs& sa,
                                              int tx, int ty, int blk, float2* sm, int hxd) {
  // y passes with a four-step pair (ax.f4col, set for the y axis when the pair
  // is the radix plan's): one LDS buffer, columns straight from and to HBM
  if ((MODE == COL_FWD || MODE == COL_INV) && !SPLIT && ax.f4col >= 0) {
    cols_four_step_dispatch<MODE == COL_FWD ? -1 : 1>(ax.f4col, cg, hx, TX, pt_stride, outer_stride, ax, tx, ty, sm);
    return;
  }
  __shared__ double sh_red[7 * (kBlock / 64)];
  const int tid = threadIdx.x;
  float2* tw = sm;
  float2* b0 = sm + ax.ntw;
  float2* b1 = b0 + n * TX;
  for (int i = tid; i < ax.ntw; i += kBlock) tw[i] = ax.tw[i];
  const int kx0 = tx * TX;
  const int nl = min(TX, hx - kx0);
  const size_t base = (size_t)kx0 + (size_t)ty * (size_t)outer_stride;
  for (int e0 = tid; e0 < n * TX; e0 += kStageU * kBlock) {
    float2 v[kStageU];
#pragma unroll
    for (int u = 0; u < kStageU; u++) {
      const int e = e0 + u * kBlock;
      const int j = sdiv(e, small_div(TX));
      const int l = e - j * TX;
      v[u] = (e < n * TX && l < nl) ? cg[base + (size_t)j * pt_stride + l] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < kStageU; u++) {
      const int e = e0 + u * kBlock;
      if (e < n * TX) b0[e] = v[u];
    }
  }
  __syncthreads();
  float2* res;
  if (MODE == COL_INV) {
    res = fft_lines<+1, SPLIT>(b0, b1, tw, ax.rp, TX, TX, tid, kBlock, ax.prec);
  } else {
    res = fft_lines<-1, SPLIT>(b0, b1, tw, ax.rp, TX, TX, tid, kBlock, ax.prec);
  }
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  if (MODE == COL_FWD_SOLVE_INV) {
    solve_lines(sa, res, n, TX, nl, kx0, hxd, acc);
    __syncthreads();
  }
  if (MODE == COL_FWD_SOLVE_INV || MODE == COL_FWD_INV) {
    res = fft_lines<+1, SPLIT>(res, res == b0 ? b1 : b0, tw, ax.rp, TX, TX, tid, kBlock, ax.prec);
  }
  for (int e = tid; e < n * TX; e += kBlock) {
    const int j = sdiv(e, small_div(TX));
    const int l = e - j * TX;
    if (l < nl) cg[base + (size_t)j * pt_stride + l] = res[e];
  }
  if (MODE == COL_FWD_SOLVE_INV && sa.calc_ev) ev_partial(sa, acc, sh_red, blk);
}

template <int MODE, bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_fft_cols(float2* __restrict__ cg, int n, int hx, int TX,
                                                     long long pt_stride, long long outer_stride, AxisArgs ax,
                                                     SolveArgs sa, EvReduce red, int hxd) {
  extern __shared__ __align__(16) float2 sm[];
  if ((red.nsolve > 0 || red.nself > 0) && blockIdx.x == 0 && blockIdx.y == 0) ev_reduce_block<kBlock>(red);
  fft_cols_body<MODE, SPLIT>(cg, n, hx, TX, pt_stride, outer_stride, ax, sa, blockIdx.x, blockIdx.y,
                      blockIdx.x + gridDim.x * blockIdx.y, sm, hxd);
}

// Stand-alone scalar sum over the whole half spectrum (used when the spectrum
// has been materialised, e.g. scalar_sum called twice, and for triclinic
// boxes).
__global__ __launch_bounds__(kBlock) void k_scalar_sum(float2* __restrict__ cg, int hx, int total, SolveArgs sa) {
  __shared__ double sh_red[7 * (kBlock / 64)];
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx < total) {
    const int k1 = idx % hx;
    const int t = idx / hx;
    const int k2 = t % sa.n2;
    const int k3 = t / sa.n2;
    const size_t a = (size_t)k3 * (size_t)spec_pitch(hx, sa.n2) + (size_t)k2 * hx + k1;
    cg[a] = sa.tric ? solve_point_tric(sa, cg[a], k1, k2, k3, acc) : solve_point(sa, cg[a], k1, k2, k3, acc);
  }
  if (sa.calc_ev) ev_partial(sa, acc, sh_red, blockIdx.x);
}

// ----------------------------------------------------------------- launch --

size_t fft_x_lds_bytes(int nf, int NR, int ntw) { return sizeof(float2) * ((size_t)ntw + 2 * (size_t)nf * (NR + 1)); }
size_t fft_cols_lds_bytes(int n, int TX, int nbuf) { return sizeof(float2) * ((size_t)n + (size_t)nbuf * n * TX); }

hipError_t launch_fft_x_r2c(const float* grid, float2* cg, int nx, int ny, int nz, int NR, const AxisArgs& ax,
                            hipStream_t s) {
  const size_t lds = fft_x_lds_bytes(ax.rp.n, NR, ax.ntw);
  const int nrows = ny * nz;
  hipLaunchKernelGGL(k_fft_x_r2c, dim3((nrows + NR - 1) / NR), dim3(kBlock), lds, s, grid, cg, nx, nrows, ny, NR,
                     ax);
  return hipGetLastError();
}

hipError_t launch_fft_x_c2r(const float2* cg, float* grid, int nx, int ny, int nz, int NR, const AxisArgs& ax,
                            hipStream_t s) {
  const size_t lds = fft_x_lds_bytes(ax.rp.n, NR, ax.ntw);
  const int nrows = ny * nz;
  hipLaunchKernelGGL(k_fft_x_c2r, dim3((nrows + NR - 1) / NR), dim3(kBlock), lds, s, cg, grid, nx, nrows, ny, NR,
                     ax);
  return hipGetLastError();
}

int cols_blocks(int hx, int TX, int nouter) { return ((hx + TX - 1) / TX) * nouter; }

hipError_t launch_fft_cols(int mode, float2* cg, int n, int hx, int TX, long long pt_stride, int nouter,
                           long long outer_stride, const AxisArgs& ax, const SolveArgs& sa, const EvReduce& red,
                           int hxd, hipStream_t s) {
  const bool split = ax.prec != PREC_FP32;
  const bool four_step = mode == COL_FWD_SOLVE_INV && !split && ax.f4 >= 0;
  const size_t lds = fft_cols_lds_bytes(n, TX, four_step ? 1 : 2);
  const dim3 grid((hx + TX - 1) / TX, nouter);
  // the fused z pass indexes its tile with shifts
  if (mode == COL_FWD_SOLVE_INV && (TX <= 0 || (TX & (TX - 1)) != 0)) return hipErrorInvalidValue;
  if (four_step) {
    if (nouter != 1 || outer_stride != 0) return hipErrorInvalidValue;  // one flattened plane of lines
    switch (ax.f4) {
#define PME_ZS_CASE(c, a, b)                                                                                       \
  case c:                                                                                                          \
    hipLaunchKernelGGL((k_zsolve<a, b>), grid, dim3(kBlock), lds, s, cg, hx, TX, pt_stride, ax, sa, red, hxd); \
    break;
      PME_F4_PAIRS(PME_ZS_CASE)
#undef PME_ZS_CASE
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
#define PME_COLS(M)                                                                                         \
  if (split)                                                                                              \
    hipLaunchKernelGGL((k_fft_cols<M, true>), grid, dim3(kBlock), lds, s, cg, n, hx, TX, pt_stride, outer_stride, ax, sa, red, hxd); \
  else                                                                                                    \
    hipLaunchKernelGGL((k_fft_cols<M, false>), grid, dim3(kBlock), lds, s, cg, n, hx, TX, pt_stride, outer_stride, ax, sa, red, hxd);
  switch (mode) {
    case COL_FWD: PME_COLS(COL_FWD); break;
    case COL_INV: PME_COLS(COL_INV); break;
    case COL_FWD_SOLVE_INV: PME_COLS(COL_FWD_SOLVE_INV); break;
    case COL_FWD_INV: PME_COLS(COL_FWD_INV); break;
    default:
      return hipErrorInvalidValue;
  }
#undef PME_COLS
  return hipGetLastError();
}

int scalar_sum_blocks(int total) { return (total + kBlock - 1) / kBlock; }

hipError_t launch_scalar_sum(float2* cg, int hx, const SolveArgs& sa, hipStream_t s) {
  const int total = hx * sa.n2 * sa.n3;
  hipLaunchKernelGGL(k_scalar_sum, dim3((total + kBlock - 1) / kBlock), dim3(kBlock), 0, s, cg, hx, total, sa);
  return hipGetLastError();
}

// ---- batched: persistent workgroups over the work items of many systems ----

template <int INV>
__global__ __launch_bounds__(kBlock) void k_batch_fft_x(const SysDesc* __restrict__ sys,
                                                        const WorkItem* __restrict__ items, int nitems) {
  extern __shared__ __align__(16) float2 sm[];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const WorkItem w = uniform_item(items[it]);
    const SysDesc& d = sys[w.sys];
    const int nrows = d.bin.ny * d.bin.nz;
    if (INV)
      fft_x_c2r_body(d.cgrid, d.grid, d.bin.nx, nrows, d.bin.ny, d.NR, d.ax[0], w.idx, sm);
    else
      fft_x_r2c_body(d.grid, d.cgrid, d.bin.nx, nrows, d.bin.ny, d.NR, d.ax[0], w.idx, sm);
    __syncthreads();
  }
}

// energy/virial of one system from its z-pass partials and its self energy
// from the binning's per-chunk sums of q*q (fixed order, one workgroup),
// assigned to ev_out
template <int NT = kBlock>
__device__ __forceinline__ void batch_ev_reduce(const SysDesc& d) {
  __shared__ double sh[7 * (NT / 64)];
  if (d.sa.calc_ev) {
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < d.nzblk; i += NT)
#pragma unroll
      for (int c = 0; c < 7; c++) v[c] += d.sa.partial[(size_t)i * 7 + c];
    block_sum<7, NT>(v, sh);
    if (threadIdx.x == 0) {
      const double hc = 0.5 * kCcelec;
      double* o = d.ev_out;
      o[0] = v[0] * hc;
      const double v0 = -hc * v[1], v1 = -hc * v[2], v2 = -hc * v[3];
      const double v3 = -hc * v[4], v4 = -hc * v[5], v5 = -hc * v[6];
      o[1] = v0; o[2] = v1; o[3] = v2;
      o[4] = v1; o[5] = v3; o[6] = v4;
      o[7] = v2; o[8] = v4; o[9] = v5;
    }
  }
  if (d.calc_self) {
    double v[1] = {0.0};
    for (int i = threadIdx.x; i < d.bin.nchunk; i += NT) v[0] += d.self_partial[i];
    block_sum<1, NT>(v, sh);
    if (threadIdx.x == 0) d.ev_out[10] = -v[0] * d.kcs;
  }
}

template <int AXIS, int MODE, bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_batch_cols(const SysDesc* __restrict__ sys,
                                                       const WorkItem* __restrict__ items, int nitems, int reduce) {
  extern __shared__ __align__(16) float2 sm[];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const WorkItem w = uniform_item(items[it]);
    const SysDesc& d = sys[w.sys];
    if (reduce && w.idx == 0 && (d.sa.calc_ev || d.calc_self)) batch_ev_reduce<kBlock>(d);
    const int hx = d.bin.nx / 2 + 1;
    const long long pz = spec_pitch(hx, d.bin.ny);
    if (AXIS == 1) {  // y: tiles of kx per kz
      const int ntx = (hx + d.TXy - 1) / d.TXy;
      fft_cols_body<MODE, SPLIT>(d.cgrid, d.bin.ny, hx, d.TXy, hx, pz, d.ax[1], d.sa, w.idx % ntx, w.idx / ntx,
                                 w.idx, sm, hx);
    } else {  // z: tiles of the flattened (ky, kx) of a plane
      fft_cols_body<MODE, SPLIT>(d.cgrid, d.bin.nz, hx * d.bin.ny, d.TXz, pz, 0, d.ax[2], d.sa, w.idx, 0, w.idx, sm,
                                 hx);
    }
    __syncthreads();
  }
}

// The fused z pass of systems whose z length has a four-step pair of class
// CLS (their items only): tiles of the flattened (ky, kx) of a z plane.
template <int CLS>
__global__ __launch_bounds__(kBlock, 4) void k_batch_zsolve(const SysDesc* __restrict__ sys,
                                                         const WorkItem* __restrict__ items, int nitems) {
  extern __shared__ __align__(16) float2 sm[];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const WorkItem w = uniform_item(items[it]);
    const SysDesc& d = sys[w.sys];
    const int hx = d.bin.nx / 2 + 1;
    zsolve_four_step_dispatch<CLS>(d.ax[2].f4, d.cgrid, hx * d.bin.ny, d.TXz, spec_pitch(hx, d.bin.ny), 0, d.ax[2],
                                   d.sa, w.idx, 0, w.idx, sm, hx);
    __syncthreads();
  }
}

// items are (system, first plane z); each covers planes z .. z+g-1 with g =
// plane_group(.., maxnp, NT < kPlaneBlock) of its system (fewer at the top of
// the grid).  SPLIT: every system of the launch may use the matrix-core stage
// in y.  NT = kPlaneBlock / 2: half-size workgroups, two per CU, for the
// systems whose plane fits half the LDS budget: one workgroup's HBM loads and
// stores then overlap the other's transforms.
template <int INV, bool SPLIT, int NT>
__global__ __launch_bounds__(NT, 4) void k_batch_plane(const SysDesc* __restrict__ sys,
                                                    const WorkItem* __restrict__ items, int nitems, int reduce,
                                                    int maxnp) {
  extern __shared__ __align__(16) float2 sm[];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const WorkItem w = uniform_item(items[it]);
    const SysDesc& d = sys[w.sys];
    const int npl = min(plane_group(d.bin.nx, d.bin.ny, d.ax[0].rp.n, maxnp, NT < kPlaneBlock), d.bin.nz - w.idx);
    if (INV) {
      if (reduce && w.idx == 0 && (d.sa.calc_ev || d.calc_self)) batch_ev_reduce<NT>(d);
      fft_plane_inv_body<SPLIT, NT>(d.cgrid, d.grid, d.bin.nx, d.bin.ny, d.ax[0], d.ax[1], w.idx, npl, sm);
    } else {
      fft_plane_fwd_body<SPLIT, NT>(d.grid, d.cgrid, d.bin.nx, d.bin.ny, d.ax[0], d.ax[1], w.idx, npl, sm);
    }
    __syncthreads();
  }
}

hipError_t launch_batch_plane(int inverse, int reduce, int maxnp, int split, int half, const SysDesc* sys,
                              const WorkItem* items, int nitems, int grid, size_t lds, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  if (maxnp < 1 || maxnp > kPlaneGroupMax) return hipErrorInvalidValue;
  const dim3 g(grid);
#define PME_BPLANE(I, S)                                                                                        \
  if (half)                                                                                                     \
    hipLaunchKernelGGL((k_batch_plane<I, S, kPlaneBlock / 2>), g, dim3(kPlaneBlock / 2), lds, s, sys, items,   \
                       nitems, reduce, maxnp);                                                                  \
  else                                                                                                          \
    hipLaunchKernelGGL((k_batch_plane<I, S, kPlaneBlock>), g, dim3(kPlaneBlock), lds, s, sys, items, nitems,    \
                       reduce, maxnp);
  if (inverse && split) {
    PME_BPLANE(1, true)
  } else if (inverse) {
    PME_BPLANE(1, false)
  } else if (split) {
    PME_BPLANE(0, true)
  } else {
    PME_BPLANE(0, false)
  }
#undef PME_BPLANE
  return hipGetLastError();
}

hipError_t launch_batch_fft_x(int inverse, const SysDesc* sys, const WorkItem* items, int nitems, int grid, size_t lds,
                              hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  if (inverse)
    hipLaunchKernelGGL(k_batch_fft_x<1>, dim3(grid), dim3(kBlock), lds, s, sys, items, nitems);
  else
    hipLaunchKernelGGL(k_batch_fft_x<0>, dim3(grid), dim3(kBlock), lds, s, sys, items, nitems);
  return hipGetLastError();
}

hipError_t launch_batch_zsolve(int cls, const SysDesc* sys, const WorkItem* items, int nitems, int grid, size_t lds,
                               hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  if (cls == 0)
    hipLaunchKernelGGL(k_batch_zsolve<0>, dim3(grid), dim3(kBlock), lds, s, sys, items, nitems);
  else if (cls == 1)
    hipLaunchKernelGGL(k_batch_zsolve<1>, dim3(grid), dim3(kBlock), lds, s, sys, items, nitems);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_batch_cols(int axis, int mode, int reduce, int split, const SysDesc* sys, const WorkItem* items,
                             int nitems, int grid, size_t lds, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  const dim3 g(grid), b(kBlock);
#define PME_BCOLS(A, M)                                                                                 \
  if (split)                                                                                          \
    hipLaunchKernelGGL((k_batch_cols<A, M, true>), g, b, lds, s, sys, items, nitems, reduce);         \
  else                                                                                                \
    hipLaunchKernelGGL((k_batch_cols<A, M, false>), g, b, lds, s, sys, items, nitems, reduce);
  if (axis == 1 && mode == COL_FWD) { PME_BCOLS(1, COL_FWD); }
  else if (axis == 1 && mode == COL_INV) { PME_BCOLS(1, COL_INV); }
  else if (axis == 2 && mode == COL_FWD_SOLVE_INV) { PME_BCOLS(2, COL_FWD_SOLVE_INV); }
  else return hipErrorInvalidValue;
#undef PME_BCOLS
  return hipGetLastError();
}

}  // namespace pme
