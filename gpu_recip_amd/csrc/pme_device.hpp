// pme_device.hpp -- device helpers shared by the PME kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pme {

constexpr int kBlock = 256;          // threads per workgroup for every kernel here
constexpr int kStageU = 8;           // global loads kept in flight per thread when staging to LDS
constexpr double kCcelec = 332.0716; // CudaPMERecip.h:27
constexpr float kForceScale = 1099511627776.0f;  // 2^40, gpu_utils.h:9

// Plane pitch of the half spectrum in complex elements: the hx * ny values of
// a z plane (rows of hx = nx/2 + 1, ky outer) rounded up to a multiple of 32,
// so that every plane starts on a 256-byte boundary and the z pass, which
// reads aligned runs of each plane, touches whole cache lines only.
__host__ __device__ __forceinline__ long long spec_pitch(int hx, int ny) {
  return ((long long)hx * ny + 31) & ~31LL;
}

// Grid index of one coordinate, CudaPMERecip.cu:512-529.  The reference is
// built with -fmad=true, which fuses x*recip + 2.0f into one FMA; nothing else
// in the expression can fuse.  Contraction is switched off here so the
// compiler cannot fuse anything beyond the explicit fmaf.
// base: 2 for a diagonal recip; a triclinic one brings the off-diagonal terms
// of the row in it (atom_grid_index).
__device__ __forceinline__ void grid_index_at(float x, float r, float base, int n, int& i, float& w) {
#pragma clang fp contract(off)
  const float t = __builtin_fmaf(x, r, base);
  const float fr = (float)n * (t - (floorf(t + 0.5f) - 0.5f));
  i = (int)fr;
  w = fr - (float)i;
}
__device__ __forceinline__ void grid_index(float x, float r, int n, int& i, float& w) {
  grid_index_at(x, r, 2.0f, n, i, w);
}

// Grid index brought into [0, n) before it addresses memory.  For finite
// coordinates of any sane size the formula above already gives 0 <= i < n
// (n*(1-ulp) never rounds up to n in FP32); non-finite or astronomically large
// coordinates can give anything, and this keeps every access in bounds.
__device__ __forceinline__ int safe_index(int i, int n) {
  if ((unsigned)i < (unsigned)n) return i;
  i %= n;
  return i < 0 ? i + n : i;
}

// theta[j] = M_p(w + p-1-j): the cardinal B-spline recursion of
// CudaPMERecip.cu:181-208 (same operation order, FP32).
template <int P>
__device__ __forceinline__ void bspline_theta(float w, float* th) {
  th[P - 1] = 0.0f;
  th[1] = w;
  th[0] = 1.0f - w;
#pragma unroll
  for (int k = 3; k <= P - 1; k++) {
    const float div = 1.0f / (float)(k - 1);
    th[k - 1] = div * w * th[k - 2];
#pragma unroll
    for (int j = 1; j <= k - 2; j++)
      th[k - j - 1] = div * ((w + j) * th[k - j - 2] + (k - j - w) * th[k - j - 1]);
    th[0] = div * (1.0f - w) * th[0];
  }
  const float div = 1.0f / (float)(P - 1);
  th[P - 1] = div * w * th[P - 2];
#pragma unroll
  for (int j = 1; j <= P - 2; j++)
    th[P - j - 1] = div * ((w + j) * th[P - j - 2] + (P - j - w) * th[P - j - 1]);
  th[0] = div * (1.0f - w) * th[0];
}

// theta and dtheta = d theta/dw, CudaPMERecip.cu:120-176.
template <int P>
__device__ __forceinline__ void bspline_theta_dtheta(float w, float* th, float* dth) {
  th[P - 1] = 0.0f;
  th[1] = w;
  th[0] = 1.0f - w;
#pragma unroll
  for (int k = 3; k <= P - 1; k++) {
    const float div = 1.0f / (float)(k - 1);
    th[k - 1] = div * w * th[k - 2];
#pragma unroll
    for (int j = 1; j <= k - 2; j++)
      th[k - j - 1] = div * ((w + j) * th[k - j - 2] + (k - j - w) * th[k - j - 1]);
    th[0] = div * (1.0f - w) * th[0];
  }
  dth[0] = -th[0];
#pragma unroll
  for (int j = 2; j <= P; j++) dth[j - 1] = th[j - 2] - th[j - 1];
  const float div = 1.0f / (float)(P - 1);
  th[P - 1] = div * w * th[P - 2];
#pragma unroll
  for (int j = 1; j <= P - 2; j++)
    th[P - j - 1] = div * ((w + j) * th[P - j - 2] + (P - j - w) * th[P - j - 1]);
  th[0] = div * (1.0f - w) * th[0];
}

// 64-bit fixed point with scale 2^40 (gpu_utils.h:9): a contribution v is
// stored as round(v * 2^40) to the nearest integer (ties to even), as
// lliroundf(FORCE_SCALE*val) in write_grid<long long> (CudaPMERecip.cu:51-58).
// The caller folds 2^40 into an operand beforehand (a power-of-two factor,
// so the product is exactly the 2^40 multiple of the unscaled one) and
// passes the two FP32 factors a, b widened to FP64 (|a b| < 2^51).  The
// product of two FP32 values has at most 48 significant bits, so it is exact
// in FP64, and one FMA with 1.5*2^52 rounds it once, to the nearest integer
// (ties to even), into the low mantissa bits as 64-bit two's complement;
// subtracting the constant's bit pattern leaves it.  The constant's low word
// is zero, so the subtraction is one 32-bit add on the high word.
__device__ __forceinline__ unsigned long long to_fixed_product(double a, double b) {
  const double t = __fma_rn(a, b, 6755399441055744.0);
  return (unsigned long long)__double_as_longlong(t) - 0x4338000000000000ull;
}
// back to FP32: ((float)val)*2^-40 as reduce_force<long long, float> (reduce.cu:66-78)
__device__ __forceinline__ float from_fixed40(unsigned long long u) {
  return (float)((double)(long long)u * (1.0 / 1099511627776.0));
}

// Brick decomposition of one grid axis into nb nearly equal bricks:
// brick b covers [floor(b n/nb), floor((b+1) n/nb)).
__device__ __forceinline__ int brick_lo(int b, int n, int nb) { return (b * n) / nb; }
// brick holding grid point i (0 <= i < n): the unique b with b n < (i+1) nb <= (b+1) n
__device__ __forceinline__ int brick_of(int i, int n, int nb) { return ((i + 1) * nb - 1) / n; }

// ---- complex helpers -----------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by s*i (s = +1 or -1)
__device__ __forceinline__ float2 cmul_si(float2 a, float s) { return make_float2(-s * a.y, s * a.x); }

// Small DFTs y_k = sum_j a_j e^{S 2 pi i jk/R}, S = -1 forward, +1 inverse.
template <int R, int S> struct Dft;

template <int S> struct Dft<2, S> {
  static __device__ __forceinline__ void run(float2* a) {
    float2 t = a[1];
    a[1] = csub(a[0], t);
    a[0] = cadd(a[0], t);
  }
};

template <int S> struct Dft<3, S> {
  static __device__ __forceinline__ void run(float2* a) {
    const float c = -0.5f, s3 = 0.866025403784438647f;  // cos(2pi/3), sin(2pi/3)
    float2 t = cadd(a[1], a[2]);
    float2 d = csub(a[1], a[2]);
    float2 m = make_float2(a[0].x + c * t.x, a[0].y + c * t.y);
    float2 r = cmul_si(cscale(d, s3), (float)S);
    a[0] = cadd(a[0], t);
    a[1] = cadd(m, r);
    a[2] = csub(m, r);
  }
};

template <int S> struct Dft<4, S> {
  static __device__ __forceinline__ void run(float2* a) {
    float2 t0 = cadd(a[0], a[2]);
    float2 t1 = csub(a[0], a[2]);
    float2 t2 = cadd(a[1], a[3]);
    float2 t3 = cmul_si(csub(a[1], a[3]), (float)S);
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};

template <int S> struct Dft<5, S> {
  static __device__ __forceinline__ void run(float2* a) {
    const float c1 = 0.309016994374947424f;   // cos(2pi/5)
    const float c2 = -0.809016994374947424f;  // cos(4pi/5)
    const float s1 = 0.951056516295153572f;   // sin(2pi/5)
    const float s2 = 0.587785252292473129f;   // sin(4pi/5)
    float2 t1 = cadd(a[1], a[4]), d1 = csub(a[1], a[4]);
    float2 t2 = cadd(a[2], a[3]), d2 = csub(a[2], a[3]);
    float2 m1 = make_float2(a[0].x + c1 * t1.x + c2 * t2.x, a[0].y + c1 * t1.y + c2 * t2.y);
    float2 m2 = make_float2(a[0].x + c2 * t1.x + c1 * t2.x, a[0].y + c2 * t1.y + c1 * t2.y);
    float2 r1 = cmul_si(make_float2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y), (float)S);
    float2 r2 = cmul_si(make_float2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y), (float)S);
    a[0] = cadd(a[0], cadd(t1, t2));
    a[1] = cadd(m1, r1);
    a[4] = csub(m1, r1);
    a[2] = cadd(m2, r2);
    a[3] = csub(m2, r2);
  }
};

template <int S> struct Dft<8, S> {
  static __device__ __forceinline__ void run(float2* a) {
    // radix-2 split: evens and odds as DFT4, odds twiddled by e^{S 2 pi i k/8}
    float2 e[4] = {a[0], a[2], a[4], a[6]};
    float2 o[4] = {a[1], a[3], a[5], a[7]};
    Dft<4, S>::run(e);
    Dft<4, S>::run(o);
    const float h = 0.707106781186547524f;
    // k=1: (h, S h); k=2: S i; k=3: (-h, S h)
    float2 o1 = cmul(o[1], make_float2(h, S * h));
    float2 o2 = cmul_si(o[2], (float)S);
    float2 o3 = cmul(o[3], make_float2(-h, S * h));
    a[0] = cadd(e[0], o[0]);
    a[4] = csub(e[0], o[0]);
    a[1] = cadd(e[1], o1);
    a[5] = csub(e[1], o1);
    a[2] = cadd(e[2], o2);
    a[6] = csub(e[2], o2);
    a[3] = cadd(e[3], o3);
    a[7] = csub(e[3], o3);
  }
};

// ---- register DFTs of composite length (four-step FFT codelets) -----------
// cos and sin of 2 pi m / N in double at compile time: Taylor series after
// reduction to [-pi, pi]; the quarter-turn points are exact.
constexpr double ct_trig(int m, int N, bool sine) {
  m %= N;
  if (m < 0) m += N;
  if ((4 * m) % N == 0) {
    const int q = (4 * m) / N;  // quarter turns
    const double c[4] = {1.0, 0.0, -1.0, 0.0}, s[4] = {0.0, 1.0, 0.0, -1.0};
    return sine ? s[q] : c[q];
  }
  double x = 2.0 * 3.14159265358979323846 * (double)(2 * m > N ? m - N : m) / (double)N;
  double term = sine ? x : 1.0, sum = term;
  for (int i = 1; i < 24; i++) {
    term *= -x * x / (sine ? (double)(2 * i) * (2 * i + 1) : (double)(2 * i - 1) * (2 * i));
    sum += term;
  }
  return sum;
}
template <int N> struct CtTwiddles {
  float c[N], s[N];
  constexpr CtTwiddles() : c(), s() {
    for (int m = 0; m < N; m++) {
      c[m] = (float)ct_trig(m, N, false);
      s[m] = (float)ct_trig(m, N, true);
    }
  }
};

// y_k = sum_j a_j e^{S 2 pi i jk/N} in registers for N = A * B (A in {2,3,4,5}):
// DFT_A over j = b + B a for each b, twiddle e^{S 2 pi i b ka / N}, DFT_B over
// b for each ka, output k = ka + A kb.  Lengths 1, 2, 3, 4, 5, 8 use Dft.
template <int N, int S> struct DftN {
  static constexpr int A = N % 4 == 0 ? 4 : (N % 2 == 0 ? 2 : (N % 3 == 0 ? 3 : 5));
  static constexpr int B = N / A;
  static __device__ __forceinline__ void run(float2* x) {
    constexpr CtTwiddles<N> W{};
    float2 t[N];
#pragma unroll
    for (int b = 0; b < B; b++) {
      float2 u[A];
#pragma unroll
      for (int a = 0; a < A; a++) u[a] = x[b + B * a];
      Dft<A, S>::run(u);
#pragma unroll
      for (int ka = 0; ka < A; ka++)
        t[b * A + ka] = (b * ka) % N == 0 ? u[ka] : cmul(u[ka], make_float2(W.c[(b * ka) % N], (float)S * W.s[(b * ka) % N]));
    }
#pragma unroll
    for (int ka = 0; ka < A; ka++) {
      float2 v[B];
#pragma unroll
      for (int b = 0; b < B; b++) v[b] = t[b * A + ka];
      DftN<B, S>::run(v);
#pragma unroll
      for (int kb = 0; kb < B; kb++) x[ka + A * kb] = v[kb];
    }
  }
};
template <int S> struct DftN<1, S> {
  static __device__ __forceinline__ void run(float2*) {}
};
template <int S> struct DftN<2, S> : Dft<2, S> {};
template <int S> struct DftN<3, S> : Dft<3, S> {};
template <int S> struct DftN<4, S> : Dft<4, S> {};
template <int S> struct DftN<5, S> : Dft<5, S> {};
template <int S> struct DftN<8, S> : Dft<8, S> {};

// Radix plan of one transform length (Stockham autosort, mixed radix).
struct RadixPlan {
  int n;         // transform length
  int nrad;      // number of stages
  int rad[12];   // radix of each stage (2, 3, 4, 5, 8; 16 for a matrix-core first stage)
  int tstride;   // twiddle table stride: table length / n
  unsigned long long rcode;  // rad[] packed 5 bits per stage (read by the kernels: no indexed arrays)
  int f4a, f4b;  // register four-step factors n = f4a * f4b (codelet lengths), 0: Stockham stages
};

__host__ __device__ inline int plan_radix(const RadixPlan& rp, int s) { return (int)((rp.rcode >> (5 * s)) & 31); }

// Quotient of small non-negative ints (n + d < 2^21) from the FP32
// reciprocal (v_rcp_f32, 1 ulp): the product (n + 1/2) * (1/d) is within
// 1.5 * 2^-23 relative of n/d + 1/(2d), less than 1/(2d) from it, so the
// truncation is floor(n/d).  A few full-rate instructions instead of the
// integer division sequence.
struct SmallDiv {
  float inv;
  int d;
};
__device__ __forceinline__ SmallDiv small_div(int d) { return SmallDiv{__builtin_amdgcn_rcpf((float)d), d}; }
__device__ __forceinline__ int sdiv(int n, const SmallDiv& s) { return (int)(((float)n + 0.5f) * s.inv); }
// 24-bit multiply (full rate) for index arithmetic on small operands
__device__ __forceinline__ int mul24(int a, int b) { return __mul24(a, b); }

// One Stockham stage of radix R over nlines interleaved lines.
// Element j of line l lives at buf[j*ls + l].  Stage with p = product of the
// previous radices: butterfly i reads i + r T (T = n/R), twiddles input r by
// w^{r k n/(pR)} (k = i mod p) and writes (i/p) p R + k + r p.
// tscale = (n / (p R)) * (twiddle table stride), passed in so that no stage
// needs a run-time integer division.
template <int R, int S>
__device__ __forceinline__ void stockham_stage(const float2* __restrict__ src, float2* __restrict__ dst,
                                               const float2* __restrict__ tw, int n, int p, int tscale,
                                               int nlines, int ls, int tid, int nthr) {
  const int T = n / R;
  const int nbf = T * nlines;
  const SmallDiv dl = small_div(nlines), dp = small_div(p);
  const int Tls = mul24(T, ls), pls = mul24(p, ls);
  for (int bf = tid; bf < nbf; bf += nthr) {
    const int i = sdiv(bf, dl);
    const int l = bf - mul24(i, nlines);
    const int ip = sdiv(i, dp);
    const int k = i - mul24(ip, p);
    float2 u[R];
    const float2* s0 = src + mul24(i, ls) + l;
#pragma unroll
    for (int r = 0; r < R; r++) u[r] = s0[r * Tls];
    if (p > 1) {
      const int kt = mul24(k, tscale);
#pragma unroll
      for (int r = 1; r < R; r++) {
        float2 w = tw[r * kt];
        if (S > 0) w.y = -w.y;
        u[r] = cmul(u[r], w);
      }
    }
    Dft<R, S>::run(u);
    const int j = mul24(mul24(ip, p), R) + k;
    float2* d0 = dst + mul24(j, ls) + l;
#pragma unroll
    for (int r = 0; r < R; r++) d0[r * pls] = u[r];
  }
}

// ---- split-precision radix-16 stage on the matrix cores ---------------------
// Precision of the radix-16 first stage: FP32 (no MFMA stage), BF16 hi-lo
// with three products, FP16 hi-lo with three products and a per-column
// power-of-two scale (FP16's range).  Accumulation is FP32 in both.
enum { PREC_FP32 = 0, PREC_BF16X3 = 1, PREC_FP16X3 = 2 };

typedef short v4s16 __attribute__((ext_vector_type(4)));
typedef _Float16 v4f16 __attribute__((ext_vector_type(4)));
typedef float v16f32 __attribute__((ext_vector_type(16)));

// round-to-nearest-even FP32 -> BF16 bits (finite inputs)
__device__ __forceinline__ unsigned short bf16_bits(float x) {
  const unsigned u = __float_as_uint(x);
  return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_value(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

// x = hi + lo in the chosen 16-bit format (hi, lo as raw 16-bit patterns)
template <int PREC>
__device__ __forceinline__ void split16(float x, short& hi, short& lo) {
  if (PREC == PREC_BF16X3) {
    const unsigned short h = bf16_bits(x);
    hi = (short)h;
    lo = (short)bf16_bits(x - bf16_value(h));
  } else {
    const _Float16 h = (_Float16)x;
    const _Float16 l = (_Float16)(x - (float)h);
    hi = __builtin_bit_cast(short, h);
    lo = __builtin_bit_cast(short, l);
  }
}

// cos(2 pi e / 16) for 0 <= e < 16, from three constants by symmetry (selects,
// no table in memory)
__device__ __forceinline__ float cos16(int e) {
  const float c1 = 0.92387953251128674f, c2 = 0.70710678118654752f, c3 = 0.38268343236508977f;
  const int q = e & 7;
  const float v = q == 0 ? 1.0f : q == 1 ? c1 : q == 2 ? c2 : q == 3 ? c3 : q == 4 ? 0.0f : q == 5 ? -c3 : q == 6 ? -c2 : -c1;
  return (e & 8) ? -v : v;
}

template <int PREC>
__device__ __forceinline__ v16f32 mfma32x32x8(v4s16 a, v4s16 b, v16f32 c) {
  if (PREC == PREC_BF16X3) return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_32x32x8f16(__builtin_bit_cast(v4f16, a), __builtin_bit_cast(v4f16, b), c, 0, 0, 0);
}

// 16-point DFTs of ncols columns as matrix products on the matrix cores: for
// column c, y[r] = sum_m W16^(S r m) x[m], x[m] = load(ctx, m), each y[r]
// handed to store(ctx, r, y[r]), ctx = prep(c) (the column's decomposition,
// formed once).  In real arithmetic this is the 32x32 matrix
// [[Fr, -Fi], [Fi, Fr]] times the 32-row column [Re x; Im x], done per 32
// columns by one wave with v_mfma_f32_32x32x8 (K = 8, four K steps).  Operand
// layouts of the 32x32x8 instruction: lane L = 32h + j holds A[j][8q + 4h + t]
// and B[8q + 4h + t][j] (t < 4) and returns D[8(r/4) + 4h + r%4][j] in
// accumulator register r.  hi-lo: D = Ah Bh + Ah Bl + Al Bh.
template <int S, int PREC, typename Prep, typename Load, typename Store>
__device__ __forceinline__ void mfma_dft16_cols(int ncols, Prep prep, Load load, Store store, int tid, int nthr) {
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int nwave = nthr >> 6;
  const int h = lane >> 5;
  int j = lane & 31;
  // opaque to the optimiser: keeps the operand set-up below inside this
  // stage instead of hoisted (and held in registers) across callers' loops
  asm volatile("" : "+v"(j));
  // A operand (the DFT matrix in real form) for the four K steps
  v4s16 ah[4], al[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int k = 8 * q + 4 * h + t;
      const int r = j & 15, m = k & 15;
      const int e = (r * m) & 15;
      const float cr = cos16(e);
      const float si = (float)S * cos16((e + 12) & 15);  // S * sin(2 pi e / 16)
      float a;
      if (j < 16) a = (k < 16) ? cr : -si;
      else a = (k < 16) ? si : cr;
      short hi, lo;
      split16<PREC>(a, hi, lo);
      ah[q][t] = hi;
      al[q][t] = lo;
    }
  }
  const int nblk = (ncols + 31) >> 5;
  for (int blk = wave; blk < nblk; blk += nwave) {
    const int c = blk * 32 + j;
    const bool valid = c < ncols;
    const auto ctx = prep(valid ? c : 0);
    float2 x[8];  // x[4h + t] and x[8 + 4h + t]
#pragma unroll
    for (int t = 0; t < 4; t++) {
      x[t] = valid ? load(ctx, 4 * h + t) : make_float2(0.f, 0.f);
      x[4 + t] = valid ? load(ctx, 8 + 4 * h + t) : make_float2(0.f, 0.f);
    }
    float scale = 1.0f, unscale = 1.0f;
    if (PREC == PREC_FP16X3) {
      float mx = 0.0f;
#pragma unroll
      for (int t = 0; t < 8; t++) mx = fmaxf(mx, fmaxf(fabsf(x[t].x), fabsf(x[t].y)));
      mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
      if (mx > 0.0f) {
        const int ex = ilogbf(mx);
        scale = ldexpf(1.0f, 13 - ex);
        unscale = ldexpf(1.0f, ex - 13);
      }
    }
    v4s16 bh[4], bl[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      short hi, lo;
      split16<PREC>(x[t].x * scale, hi, lo);      // q = 0: Re x[4h + t]
      bh[0][t] = hi;
      bl[0][t] = lo;
      split16<PREC>(x[4 + t].x * scale, hi, lo);  // q = 1: Re x[8 + 4h + t]
      bh[1][t] = hi;
      bl[1][t] = lo;
      split16<PREC>(x[t].y * scale, hi, lo);      // q = 2: Im x[4h + t]
      bh[2][t] = hi;
      bl[2][t] = lo;
      split16<PREC>(x[4 + t].y * scale, hi, lo);  // q = 3: Im x[8 + 4h + t]
      bh[3][t] = hi;
      bl[3][t] = lo;
    }
    v16f32 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 4; q++) acc = mfma32x32x8<PREC>(ah[q], bl[q], acc);
#pragma unroll
    for (int q = 0; q < 4; q++) acc = mfma32x32x8<PREC>(al[q], bh[q], acc);
#pragma unroll
    for (int q = 0; q < 4; q++) acc = mfma32x32x8<PREC>(ah[q], bh[q], acc);
    if (valid) {
#pragma unroll
      for (int r = 0; r < 8; r++) {
        const int row = 8 * (r >> 2) + 4 * h + (r & 3);  // 0..15: output point
        store(ctx, row, make_float2(acc[r] * unscale, acc[r + 8] * unscale));
      }
    }
  }
}

// First Stockham stage with radix 16 (p = 1, no twiddles) on the matrix
// cores: column c = i*nlines + l (i < T = n/16) is x[m] = src[(i + m T) ls + l],
// y[r] -> dst[(16 i + r) ls + l].
template <int S, int PREC>
__device__ __forceinline__ void mfma_dft16_stage(const float2* __restrict__ src, float2* __restrict__ dst, int n,
                                                 int nlines, int ls, int tid, int nthr) {
  const int T = n >> 4;
  const SmallDiv dl = small_div(nlines);
  mfma_dft16_cols<S, PREC>(
      T * nlines,
      [&](int c) {
        const int i = sdiv(c, dl);
        return make_int2(i, c - i * nlines);
      },
      [&](int2 ctx, int m) { return src[(ctx.x + m * T) * ls + ctx.y]; },
      [&](int2 ctx, int r, float2 y) { dst[(16 * ctx.x + r) * ls + ctx.y] = y; }, tid, nthr);
}

// Pass A of a four-step transform n = 16 B on the matrix cores (the DFT_16 of
// four_step_pass_a<16>): unit (line l, j2) takes the 16 points j2 + B j1,
// multiplies output k1 by w_n^{S j2 k1} and writes it to (j2 + B k1).
template <int S, int PREC>
__device__ __forceinline__ void mfma_four_step_pass_a(const float2* __restrict__ src, float2* __restrict__ dst,
                                                      const float2* __restrict__ tw, int B, int tstride, int nlines,
                                                      int ls, int tid, int nthr) {
  const SmallDiv dl = small_div(nlines);
  const int Bls = mul24(B, ls);
  mfma_dft16_cols<S, PREC>(
      B * nlines,
      [&](int c) {
        const int j2 = sdiv(c, dl);
        return make_int2(j2, c - mul24(j2, nlines));
      },
      [&](int2 ctx, int m) { return src[mul24(ctx.x, ls) + ctx.y + m * Bls]; },
      [&](int2 ctx, int k1, float2 y) {
        float2 w = tw[mul24(mul24(ctx.x, k1), tstride)];
        if (S > 0) w.y = -w.y;
        dst[mul24(ctx.x, ls) + ctx.y + k1 * Bls] = k1 == 0 ? y : cmul(y, w);
      },
      tid, nthr);
}

// ---- four-step transform of lines in LDS ----------------------------------
// n = A B, j = j2 + B j1, k = k1 + A k2 (A, B codelet lengths).  Pass A: unit
// (line l, j2) loads the A points j2 + B j1 into registers, runs DFT_A and
// multiplies by w_n^{S j2 k1}, writing (k1, j2) to the other buffer; pass B:
// unit (l, k1) runs DFT_B over j2 and writes X[k1 + A k2] back in natural
// order.  Two LDS passes (two barriers) for the whole transform; consecutive
// units take consecutive lines, so the accesses of a wave are contiguous.
template <int A, int S>
__device__ __forceinline__ void four_step_pass_a(const float2* __restrict__ src, float2* __restrict__ dst,
                                                 const float2* __restrict__ tw, int B, int tstride, int nlines,
                                                 int ls, int tid, int nthr) {
  const SmallDiv dl = small_div(nlines);
  const int Bls = mul24(B, ls);
  for (int u = tid; u < B * nlines; u += nthr) {
    const int j2 = sdiv(u, dl);
    const int l = u - mul24(j2, nlines);
    float2 x[A];
    const float2* s0 = src + mul24(j2, ls) + l;
#pragma unroll
    for (int j1 = 0; j1 < A; j1++) x[j1] = s0[j1 * Bls];
    DftN<A, S>::run(x);
    float2* d0 = dst + mul24(j2, ls) + l;
    const int ts = mul24(j2, tstride);
    d0[0] = x[0];
#pragma unroll
    for (int k1 = 1; k1 < A; k1++) {
      float2 w = tw[k1 * ts];
      if (S > 0) w.y = -w.y;
      d0[k1 * Bls] = cmul(x[k1], w);
    }
  }
}
template <int B, int S>
__device__ __forceinline__ void four_step_pass_b(const float2* __restrict__ src, float2* __restrict__ dst, int A,
                                                 int nlines, int ls, int tid, int nthr) {
  const SmallDiv dl = small_div(nlines);
  const int Als = mul24(A, ls);
  for (int u = tid; u < A * nlines; u += nthr) {
    const int k1 = sdiv(u, dl);
    const int l = u - mul24(k1, nlines);
    float2 x[B];
    const float2* s0 = src + mul24(mul24(k1, B), ls) + l;
#pragma unroll
    for (int j2 = 0; j2 < B; j2++) x[j2] = s0[j2 * ls];
    DftN<B, S>::run(x);
    float2* d0 = dst + mul24(k1, ls) + l;
#pragma unroll
    for (int k2 = 0; k2 < B; k2++) d0[k2 * Als] = x[k2];
  }
}
// codelet lengths of the four-step passes
#define PME_F4_LENGTHS(X) X(2) X(3) X(4) X(5) X(6) X(8) X(9) X(10) X(12) X(16)
template <int S, bool SPLIT = false>
__device__ __forceinline__ float2* fft_lines_four_step(float2* buf0, float2* buf1, const float2* tw,
                                                       const RadixPlan& rp, int nlines, int ls, int tid, int nthr,
                                                       int prec = PREC_FP32) {
  if (SPLIT && prec != PREC_FP32 && rp.f4a == 16) {
    // the 16-point DFTs as hi-lo split matrix products
    if (prec == PREC_BF16X3)
      mfma_four_step_pass_a<S, PREC_BF16X3>(buf0, buf1, tw, rp.f4b, rp.tstride, nlines, ls, tid, nthr);
    else
      mfma_four_step_pass_a<S, PREC_FP16X3>(buf0, buf1, tw, rp.f4b, rp.tstride, nlines, ls, tid, nthr);
  } else {
    switch (rp.f4a) {
#define PME_F4A(L) \
  case L: four_step_pass_a<L, S>(buf0, buf1, tw, rp.f4b, rp.tstride, nlines, ls, tid, nthr); break;
      PME_F4_LENGTHS(PME_F4A)
#undef PME_F4A
      default: break;
    }
  }
  __syncthreads();
  switch (rp.f4b) {
#define PME_F4B(L) \
  case L: four_step_pass_b<L, S>(buf1, buf0, rp.f4a, nlines, ls, tid, nthr); break;
    PME_F4_LENGTHS(PME_F4B)
#undef PME_F4B
    default: break;
  }
  __syncthreads();
  return buf0;
}

// Full transform of nlines lines held in LDS (buf0 -> result).  Returns the
// buffer holding the result (buf0 or buf1).  Ends with a barrier.  With
// SPLIT, prec != PREC_FP32 and a plan whose first radix is 16, that stage runs
// on the matrix cores (mfma_dft16_stage).  Without SPLIT the matrix-core code
// is not compiled in, which keeps the register budget (and so the occupancy)
// of FP32-only kernels at what their Stockham stages need.  Lengths with a
// four-step factor pair (rp.f4a > 0) run fft_lines_four_step instead of
// the Stockham stages (except the matrix-core variant); FOUR_STEP = false
// keeps the Stockham stages.
template <int S, bool SPLIT = false, bool FOUR_STEP = true>
__device__ __forceinline__ float2* fft_lines(float2* buf0, float2* buf1, const float2* tw, const RadixPlan& rp,
                             int nlines, int ls, int tid, int nthr, int prec = PREC_FP32) {
  float2* src = buf0;
  float2* dst = buf1;
  if (FOUR_STEP && rp.f4a > 0 && (!(SPLIT && prec != PREC_FP32) || rp.f4a == 16))
    return fft_lines_four_step<S, SPLIT>(buf0, buf1, tw, rp, nlines, ls, tid, nthr, prec);
  int p = 1;
  int s0 = 0;
  if (SPLIT && prec != PREC_FP32 && rp.nrad > 0 && plan_radix(rp, 0) == 16) {
    if (prec == PREC_BF16X3)
      mfma_dft16_stage<S, PREC_BF16X3>(src, dst, rp.n, nlines, ls, tid, nthr);
    else
      mfma_dft16_stage<S, PREC_FP16X3>(src, dst, rp.n, nlines, ls, tid, nthr);
    __syncthreads();
    float2* t = src;
    src = dst;
    dst = t;
    p = 16;
    s0 = 1;
  }
  int q = p == 1 ? rp.n : rp.n / 16;  // n / p, then divided by each stage's (compile-time) radix
  for (int s = s0; s < rp.nrad; s++) {
    const int R = plan_radix(rp, s);
    switch (R) {
      case 2: q /= 2; stockham_stage<2, S>(src, dst, tw, rp.n, p, q * rp.tstride, nlines, ls, tid, nthr); break;
      case 3: q /= 3; stockham_stage<3, S>(src, dst, tw, rp.n, p, q * rp.tstride, nlines, ls, tid, nthr); break;
      case 4: q /= 4; stockham_stage<4, S>(src, dst, tw, rp.n, p, q * rp.tstride, nlines, ls, tid, nthr); break;
      case 5: q /= 5; stockham_stage<5, S>(src, dst, tw, rp.n, p, q * rp.tstride, nlines, ls, tid, nthr); break;
      default: q /= 8; stockham_stage<8, S>(src, dst, tw, rp.n, p, q * rp.tstride, nlines, ls, tid, nthr); break;
    }
    __syncthreads();
    float2* t = src;
    src = dst;
    dst = t;
    p *= R;
  }
  return src;
}

// ---- reductions ------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum of NV doubles over NT threads (NT == blockDim.x).  The result
// is valid in thread 0.  sh must hold NV * (NT/64) doubles.  Fixed order.
template <int NV, int NT = kBlock>
__device__ __forceinline__ void block_sum(double* v, double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < NV; c++) v[c] = wave_sum(v[c]);
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < NV; c++) sh[wid * NV + c] = v[c];
  }
  __syncthreads();
  // thread c sums component c over the waves in wave order (thread c alone
  // reads and then writes sh[c])
  if (threadIdx.x < NV) {
    const int c = threadIdx.x;
    double s = 0.0;
    for (int w = 0; w < NT / 64; w++) s += sh[w * NV + c];
    sh[c] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NV; c++) v[c] = sh[c];
  }
  __syncthreads();
}

// Block-wide exclusive scan of data[0, n) in place over NT threads (NT ==
// blockDim.x).  Returns the total in every thread.  sh: NT/64 ints of LDS.
template <int NT = kBlock>
__device__ __forceinline__ int block_exclusive_scan(int* data, int n, int* sh) {
  const int tid = threadIdx.x;
  const int per = (n + NT - 1) / NT;
  const int lo = min(tid * per, n), hi = min(lo + per, n);
  int s = 0;
  for (int i = lo; i < hi; i++) s += data[i];
  const int lane = tid & 63, wid = tid >> 6;
  int v = s;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  if (lane == 63) sh[wid] = v;
  __syncthreads();
  int wbase = 0, total = 0;
#pragma unroll
  for (int k = 0; k < NT / 64; k++) {
    if (k < wid) wbase += sh[k];
    total += sh[k];
  }
  int run = wbase + v - s;
  for (int i = lo; i < hi; i++) {
    const int c = data[i];
    data[i] = run;
    run += c;
  }
  __syncthreads();
  return total;
}

}  // namespace pme
