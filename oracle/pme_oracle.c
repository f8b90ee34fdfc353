/*
 * pme_oracle.c -- FP64 host restatement of the CudaPMERecip reciprocal solve.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
 *
 * Build: see oracle/Makefile.  Compiled with -ffp-contract=off so that the only
 * fused multiply-add is the explicit fmaf() in the FP32 index mapping.
 */
#include "pme_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double PI = 3.14159265358979323846;   /* CudaPMERecip.cu:27 */

typedef struct { double re, im; } cplx;

static int pick_threads(int n) {
#ifdef _OPENMP
  return n > 0 ? n : omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* a1: fractional coordinate -> grid index.  CudaPMERecip.cu:512-529          */
/* ------------------------------------------------------------------------ */
static void index_one_f32(float x, float r, int n, int* i_out, float* w_out) {
  /* w = x*recip + 2.0f, contracted to one FMA by -fmad=true (Makefile:174) */
  float w = fmaf(x, r, 2.0f);
  float t = floorf(w + 0.5f) - 0.5f;
  float fr = (float)n * (w - t);
  int i = (int)fr;
  *i_out = i;
  *w_out = fr - (float)i;
}

void orc_index_f32(int natom, const float* xyzq, double r1, double r2, double r3,
                   int nx, int ny, int nz, int* idx, float* frac) {
  /* recip is cast to float on the host before launch (CudaPMERecip.cu:1810-1812) */
  const float fr1 = (float)r1, fr2 = (float)r2, fr3 = (float)r3;
  for (int a = 0; a < natom; a++) {
    index_one_f32(xyzq[4 * a + 0], fr1, nx, &idx[3 * a + 0], &frac[3 * a + 0]);
    index_one_f32(xyzq[4 * a + 1], fr2, ny, &idx[3 * a + 1], &frac[3 * a + 1]);
    index_one_f32(xyzq[4 * a + 2], fr3, nz, &idx[3 * a + 2], &frac[3 * a + 2]);
  }
}

static void index_one_f64(double x, double r, int n, int* i_out, double* w_out) {
  double w = x * r + 2.0;
  double fr = (double)n * (w - (floor(w + 0.5) - 0.5));
  int i = (int)fr;
  *i_out = i;
  *w_out = fr - (double)i;
}

void orc_index_f64(int natom, const float* xyzq, double r1, double r2, double r3,
                   int nx, int ny, int nz, int* idx, double* frac) {
  for (int a = 0; a < natom; a++) {
    index_one_f64(xyzq[4 * a + 0], r1, nx, &idx[3 * a + 0], &frac[3 * a + 0]);
    index_one_f64(xyzq[4 * a + 1], r2, ny, &idx[3 * a + 1], &frac[3 * a + 1]);
    index_one_f64(xyzq[4 * a + 2], r3, nz, &idx[3 * a + 2], &frac[3 * a + 2]);
  }
}

/* Per-atom index + fractional part in the mode the params ask for.  The
 * fractional coordinate of axis a is row a of recip times (x, y, z), so a
 * triclinic box only adds the off-diagonal terms; for an orthorhombic recip
 * they are exact zeros and the result is the reference's formula bit for bit.
 * FP32 mode: the off-diagonal terms of the two other axes (in increasing axis
 * order, the larger one innermost) are fused into the +2 term first, the
 * diagonal term last: t = fmaf(x_a, R_aa, fmaf(x_b, R_ab, fmaf(x_c, R_ac, 2))). */
static int is_tric(const double* r) {
  return r[1] != 0.0 || r[2] != 0.0 || r[3] != 0.0 || r[5] != 0.0 || r[6] != 0.0 || r[7] != 0.0;
}

static void index_t_f32(float t, int n, int* i_out, float* w_out) {
  float c = floorf(t + 0.5f) - 0.5f;
  float fr = (float)n * (t - c);
  int i = (int)fr;
  *i_out = i;
  *w_out = fr - (float)i;
}

static void atom_index(const orc_params* p, const float* q4, int ix[3], double w[3]) {
  const int n[3] = {p->nx, p->ny, p->nz};
  if (!is_tric(p->recip)) {
    if (p->index_mode == ORC_INDEX_F32) {
      float f;
      for (int a = 0; a < 3; a++) {
        index_one_f32(q4[a], (float)p->recip[4 * a], n[a], &ix[a], &f);
        w[a] = f;
      }
    } else {
      for (int a = 0; a < 3; a++) index_one_f64(q4[a], p->recip[4 * a], n[a], &ix[a], &w[a]);
    }
    return;
  }
  for (int a = 0; a < 3; a++) {
    const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;  /* the other two axes, b < c */
    const double* R = &p->recip[3 * a];
    if (p->index_mode == ORC_INDEX_F32) {
      const float t = fmaf(q4[a], (float)R[a], fmaf(q4[b], (float)R[b], fmaf(q4[c], (float)R[c], 2.0f)));
      float f;
      index_t_f32(t, n[a], &ix[a], &f);
      w[a] = f;
    } else {
      const double t = ((double)q4[a] * R[a] + (double)q4[b] * R[b] + (double)q4[c] * R[c]) + 2.0;
      const double fr = (double)n[a] * (t - (floor(t + 0.5) - 0.5));
      ix[a] = (int)fr;
      w[a] = fr - (double)ix[a];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* a2: cardinal B-spline recursion.  fill_bspline_host, CudaPMERecip.cu:2507  */
/* ------------------------------------------------------------------------ */
void orc_bspline(int order, double w, double* th, double* dth) {
  th[order - 1] = 0.0;
  th[1] = w;
  th[0] = 1.0 - w;
  for (int k = 3; k <= order - 1; k++) {
    double div = 1.0 / (double)(k - 1);
    th[k - 1] = div * w * th[k - 2];
    for (int j = 1; j <= k - 2; j++)
      th[k - j - 1] = div * ((w + j) * th[k - j - 2] + (k - j - w) * th[k - j - 1]);
    th[0] = div * (1.0 - w) * th[0];
  }
  /* derivative from the order-(p-1) values */
  if (dth) {
    dth[0] = -th[0];
    for (int j = 2; j <= order; j++) dth[j - 1] = th[j - 2] - th[j - 1];
  }
  /* final recursion to order p */
  {
    int k = order;
    double div = 1.0 / (double)(k - 1);
    th[k - 1] = div * w * th[k - 2];
    for (int j = 1; j <= k - 2; j++)
      th[k - j - 1] = div * ((w + j) * th[k - j - 2] + (k - j - w) * th[k - j - 1]);
    th[0] = div * (1.0 - w) * th[0];
  }
}

/* ------------------------------------------------------------------------ */
/* a9: bsp_mod.  dftmod + calc_prefac, CudaPMERecip.cu:2481-2505, 2544-2579   */
/* ------------------------------------------------------------------------ */
void orc_bsp_mod(int order, int n, double* bsp_mod) {
  double* arr = (double*)calloc((size_t)(n > order + 1 ? n : order + 1), sizeof(double));
  double* th = (double*)malloc(sizeof(double) * order);
  double* dth = (double*)malloc(sizeof(double) * order);
  orc_bspline(order, 0.0, th, dth);
  /* bsp_arr[0] = 0, bsp_arr[1..p] = M_p at the integer nodes */
  for (int i = 2; i <= order + 1; i++) arr[i - 1] = th[i - 2];
  const double rsmall = 1.0e-10;
  for (int k = 0; k < n; k++) {
    double s1 = 0.0, s2 = 0.0;
    /* the reference's sum stops at j = n-2 (CudaPMERecip.cu:2490) */
    for (int j = 0; j < n - 1; j++) {
      long long jk = ((long long)j * k) % n;
      double arg = 2.0 * PI * (double)jk / (double)n;
      s1 += arr[j] * cos(arg);
      s2 += arr[j] * sin(arg);
    }
    bsp_mod[k] = s1 * s1 + s2 * s2;
  }
  for (int k = 0; k < n; k++)
    if (bsp_mod[k] < rsmall) {
      double lo = bsp_mod[k > 0 ? k - 1 : n - 1];
      double hi = bsp_mod[k < n - 1 ? k + 1 : 0];
      bsp_mod[k] = 0.5 * (lo + hi);
    }
  for (int k = 0; k < n; k++) bsp_mod[k] = 1.0 / bsp_mod[k];
  free(arr);
  free(th);
  free(dth);
}

/* ------------------------------------------------------------------------ */
/* a3: spread.  CudaPMERecip.cu:479-580 (scatter q*thx*thy*thz on p^3 points) */
/* Threads own z-planes, so the sum per point is in atom order and            */
/* independent of the thread count.                                          */
/* ------------------------------------------------------------------------ */
void orc_spread(const orc_params* p, int natom, const float* xyzq, double* grid) {
  const int nx = p->nx, ny = p->ny, nz = p->nz, order = p->order;
  const size_t G = (size_t)nx * ny * nz;
  memset(grid, 0, G * sizeof(double));
  int nt = pick_threads(p->nthreads);
  if (nt > nz) nt = nz;
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    int tid = omp_get_thread_num(), nth = omp_get_num_threads();
#else
    int tid = 0, nth = 1;
#endif
    int z_lo = (int)((long long)nz * tid / nth), z_hi = (int)((long long)nz * (tid + 1) / nth);
    double thx[16], thy[16], thz[16], d[16];
    for (int a = 0; a < natom; a++) {
      int ix[3];
      double w[3];
      atom_index(p, &xyzq[4 * a], ix, w);
      /* quick reject: does the z stencil touch [z_lo, z_hi)? */
      int touch = 0;
      for (int k = 0; k < order; k++) {
        int z = ix[2] + k;
        if (z >= nz) z -= nz;
        if (z >= z_lo && z < z_hi) { touch = 1; break; }
      }
      if (!touch) continue;
      double q = xyzq[4 * a + 3];
      orc_bspline(order, w[0], thx, d);
      orc_bspline(order, w[1], thy, d);
      orc_bspline(order, w[2], thz, d);
      for (int k = 0; k < order; k++) {
        int z = ix[2] + k;
        if (z >= nz) z -= nz;
        if (z < z_lo || z >= z_hi) continue;
        for (int j = 0; j < order; j++) {
          int y = ix[1] + j;
          if (y >= ny) y -= ny;
          double qyz = q * thy[j] * thz[k];
          for (int i = 0; i < order; i++) {
            int x = ix[0] + i;
            if (x >= nx) x -= nx;
            grid[x + (size_t)nx * (y + (size_t)ny * z)] += qyz * thx[i];
          }
        }
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* a4/a6: FFT.  Textbook mixed-radix decimation in time, any length          */
/* (smallest prime factor first; naive DFT for the prime butterflies).       */
/* ------------------------------------------------------------------------ */
static int smallest_factor(int n) {
  if (n % 2 == 0) return 2;
  for (int f = 3; f * f <= n; f += 2)
    if (n % f == 0) return f;
  return n;
}

/* out[k] = sum_j in[j*stride] w^{jk}, w = tw[step] of the length-N root table */
static void fft_rec(int n, const cplx* in, int stride, cplx* out, const cplx* tw, int N,
                    int step, cplx* tmp) {
  if (n == 1) { out[0] = in[0]; return; }
  int r = smallest_factor(n), m = n / r;
  for (int q = 0; q < r; q++) fft_rec(m, in + (size_t)q * stride, stride * r, out + (size_t)q * m, tw, N, step * r, tmp);
  /* combine: X[k + s*m] = sum_q out[q*m + k] * w_n^{q (k + s m)} */
  for (int k = 0; k < m; k++) {
    for (int s = 0; s < r; s++) {
      double re = 0.0, im = 0.0;
      for (int q = 0; q < r; q++) {
        long long e = ((long long)q * (k + (long long)s * m)) % n;
        cplx w = tw[(size_t)(e * step) % N];
        cplx v = out[(size_t)q * m + k];
        re += v.re * w.re - v.im * w.im;
        im += v.re * w.im + v.im * w.re;
      }
      tmp[s].re = re;
      tmp[s].im = im;
    }
    for (int s = 0; s < r; s++) out[(size_t)k + (size_t)s * m] = tmp[s];
  }
}

typedef struct { int n; cplx* tw; cplx* buf; cplx* out; cplx* tmp; } fft_plan;

static void plan_init(fft_plan* pl, int n, int sign) {
  pl->n = n;
  pl->tw = (cplx*)malloc(sizeof(cplx) * n);
  pl->buf = (cplx*)malloc(sizeof(cplx) * n);
  pl->out = (cplx*)malloc(sizeof(cplx) * n);
  pl->tmp = (cplx*)malloc(sizeof(cplx) * n);
  for (int k = 0; k < n; k++) {
    double a = 2.0 * PI * (double)k / (double)n;
    pl->tw[k].re = cos(a);
    pl->tw[k].im = sign * sin(a);
  }
}
static void plan_free(fft_plan* pl) { free(pl->tw); free(pl->buf); free(pl->out); free(pl->tmp); }
static void plan_exec(fft_plan* pl) {  /* buf -> out */
  fft_rec(pl->n, pl->buf, 1, pl->out, pl->tw, pl->n, 1, pl->tmp);
}

void orc_fft1d(int n, double* data, int sign) {
  fft_plan pl;
  plan_init(&pl, n, sign < 0 ? -1 : 1);
  memcpy(pl.buf, data, sizeof(cplx) * n);
  plan_exec(&pl);
  memcpy(data, pl.out, sizeof(cplx) * n);
  plan_free(&pl);
}

/* c2c along y and z of a half spectrum (hx, ny, nz) in place */
static void fft_yz(int hx, int ny, int nz, cplx* c, int sign, int nt) {
#pragma omp parallel num_threads(nt)
  {
    fft_plan pl;
    plan_init(&pl, ny, sign);
#pragma omp for schedule(static)
    for (int zk = 0; zk < nz * hx; zk++) {
      int z = zk / hx, k = zk % hx;
      for (int y = 0; y < ny; y++) pl.buf[y] = c[k + (size_t)hx * (y + (size_t)ny * z)];
      plan_exec(&pl);
      for (int y = 0; y < ny; y++) c[k + (size_t)hx * (y + (size_t)ny * z)] = pl.out[y];
    }
    plan_free(&pl);
    plan_init(&pl, nz, sign);
#pragma omp for schedule(static)
    for (int yk = 0; yk < ny * hx; yk++) {
      int y = yk / hx, k = yk % hx;
      for (int z = 0; z < nz; z++) pl.buf[z] = c[k + (size_t)hx * (y + (size_t)ny * z)];
      plan_exec(&pl);
      for (int z = 0; z < nz; z++) c[k + (size_t)hx * (y + (size_t)ny * z)] = pl.out[z];
    }
    plan_free(&pl);
  }
}

void orc_fft3d_r2c(int nx, int ny, int nz, const double* real, double* cplx_out, int nthreads) {
  const int hx = nx / 2 + 1;
  cplx* c = (cplx*)cplx_out;
  int nt = pick_threads(nthreads);
#pragma omp parallel num_threads(nt)
  {
    fft_plan pl;
    plan_init(&pl, nx, -1);
#pragma omp for schedule(static)
    for (int yz = 0; yz < ny * nz; yz++) {
      for (int x = 0; x < nx; x++) {
        pl.buf[x].re = real[x + (size_t)nx * yz];
        pl.buf[x].im = 0.0;
      }
      plan_exec(&pl);
      for (int k = 0; k < hx; k++) c[k + (size_t)hx * yz] = pl.out[k];
    }
    plan_free(&pl);
  }
  fft_yz(hx, ny, nz, c, -1, nt);
}

void orc_fft3d_c2r(int nx, int ny, int nz, const double* cplx_in, double* real, int nthreads) {
  const int hx = nx / 2 + 1;
  const size_t Gc = (size_t)hx * ny * nz;
  cplx* c = (cplx*)malloc(sizeof(cplx) * Gc);
  memcpy(c, cplx_in, sizeof(cplx) * Gc);
  int nt = pick_threads(nthreads);
  fft_yz(hx, ny, nz, c, +1, nt);
#pragma omp parallel num_threads(nt)
  {
    fft_plan pl;
    plan_init(&pl, nx, +1);
#pragma omp for schedule(static)
    for (int yz = 0; yz < ny * nz; yz++) {
      /* Hermitian extension of the row: X[n-k] = conj(X[k]) */
      for (int k = 0; k < nx; k++) {
        if (k < hx) {
          pl.buf[k] = c[k + (size_t)hx * yz];
        } else {
          cplx v = c[(nx - k) + (size_t)hx * yz];
          pl.buf[k].re = v.re;
          pl.buf[k].im = -v.im;
        }
      }
      plan_exec(&pl);
      for (int x = 0; x < nx; x++) real[x + (size_t)nx * yz] = pl.out[x].re;
    }
    plan_free(&pl);
  }
  free(c);
}

/* ------------------------------------------------------------------------ */
/* a5: scalar sum.  scalar_sum_ortho_kernel CudaPMERecip.cu:594-901, host     */
/* setup :1929-2047 (BOX layout: k1 over nx/2+1, k2 over ny, k3 over nz).    */
/* ------------------------------------------------------------------------ */
static void ev_term(double eterm, double q2, double fac, double m1, double m2, double m3,
                    double msq_inv, double acc[7]) {
  double tmp1 = eterm * q2;
  double vterm = 2.0 * (fac + msq_inv);
  double tmp2 = tmp1 * vterm;
  acc[0] += tmp1;
  acc[1] += tmp1 * (vterm * m1 * m1 - 1.0);
  acc[2] += tmp2 * m1 * m2;
  acc[3] += tmp2 * m1 * m3;
  acc[4] += tmp1 * (vterm * m2 * m2 - 1.0);
  acc[5] += tmp2 * m2 * m3;
  acc[6] += tmp1 * (vterm * m3 * m3 - 1.0);
}

void orc_scalar_sum(const orc_params* p, double* cplx_io, double* energy, double* virial9) {
  const int n1 = p->nx, n2 = p->ny, n3 = p->nz;
  const int s1 = n1 / 2 + 1;
  const int nf1 = n1 / 2 + (n1 % 2), nf2 = n2 / 2 + (n2 % 2), nf3 = n3 / 2 + (n3 % 2);
  const double* R = p->recip;  /* m-hat = m1 R[0][:] + m2 R[1][:] + m3 R[2][:] */
  const double fac = PI * PI / (p->kappa * p->kappa);
  /* 1/V = det(recip): r1 r2 r3 for an orthorhombic box (CudaPMERecip.cu:2024-2026) */
  const double det = R[0] * (R[4] * R[8] - R[5] * R[7]) - R[1] * (R[3] * R[8] - R[5] * R[6]) +
                     R[2] * (R[3] * R[7] - R[4] * R[6]);
  const double piv_inv = is_tric(R) ? det / PI : (R[0] * R[4] * R[8]) / PI;
  double* b1 = (double*)malloc(sizeof(double) * n1);
  double* b2 = (double*)malloc(sizeof(double) * n2);
  double* b3 = (double*)malloc(sizeof(double) * n3);
  orc_bsp_mod(p->order, n1, b1);
  orc_bsp_mod(p->order, n2, b2);
  orc_bsp_mod(p->order, n3, b3);
  cplx* c = (cplx*)cplx_io;
  double* part = (double*)calloc((size_t)7 * n3, sizeof(double));
  int nt = pick_threads(p->nthreads);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int k3 = 0; k3 < n3; k3++) {
    double* acc = &part[7 * k3];
    for (int k2 = 0; k2 < n2; k2++) {
      for (int k1 = 0; k1 < s1; k1++) {
        size_t pos = k1 + (size_t)s1 * (k2 + (size_t)n2 * k3);
        if (k1 == 0 && k2 == 0 && k3 == 0) { c[pos].re = 0.0; c[pos].im = 0.0; continue; }
        int m1 = k1 >= nf1 ? k1 - n1 : k1;
        int m2 = k2 >= nf2 ? k2 - n2 : k2;
        int m3 = k3 >= nf3 ? k3 - n3 : k3;
        double mh1 = R[0] * m1 + R[3] * m2 + R[6] * m3;
        double mh2 = R[1] * m1 + R[4] * m2 + R[7] * m3;
        double mh3 = R[2] * m1 + R[5] * m2 + R[8] * m3;
        double msq = mh1 * mh1 + mh2 * mh2 + mh3 * mh3;
        double msq_inv = 1.0 / msq;
        double eterm = exp(-fac * msq) * piv_inv * b1[k1] * b2[k2] * b3[k3] * msq_inv;
        double q2 = c[pos].re * c[pos].re + c[pos].im * c[pos].im;
        ev_term(eterm, q2, fac, mh1, mh2, mh3, msq_inv, acc);
        int mirror = (k1 >= 1);
        if (p->nyquist_mode == ORC_NYQUIST_EXACT && (n1 % 2) == 0 && k1 == n1 / 2) mirror = 0;
        if (mirror) {
          int k1s = n1 - k1;
          int k2s = (n2 - k2) % n2;
          int k3s = (n3 - k3) % n3;
          int m1s = k1s >= nf1 ? k1s - n1 : k1s;
          int m2s = k2s >= nf2 ? k2s - n2 : k2s;
          int m3s = k3s >= nf3 ? k3s - n3 : k3s;
          double ms1 = R[0] * m1s + R[3] * m2s + R[6] * m3s;
          double ms2 = R[1] * m1s + R[4] * m2s + R[7] * m3s;
          double ms3 = R[2] * m1s + R[5] * m2s + R[8] * m3s;
          double msqs = ms1 * ms1 + ms2 * ms2 + ms3 * ms3;
          double msqs_inv = 1.0 / msqs;
          double eterms = exp(-fac * msqs) * piv_inv * b1[k1s] * b2[k2s] * b3[k3s] * msqs_inv;
          ev_term(eterms, q2, fac, ms1, ms2, ms3, msqs_inv, acc);
        }
        c[pos].re *= eterm;
        c[pos].im *= eterm;
      }
    }
  }
  double tot[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int k3 = 0; k3 < n3; k3++)
    for (int i = 0; i < 7; i++) tot[i] += part[7 * k3 + i];
  const double half_ccelec = 0.5 * ORC_CCELEC;
  if (energy) *energy = tot[0] * half_ccelec;
  if (virial9) {
    double v0 = -half_ccelec * tot[1], v1 = -half_ccelec * tot[2], v2 = -half_ccelec * tot[3];
    double v3 = -half_ccelec * tot[4], v4 = -half_ccelec * tot[5], v5 = -half_ccelec * tot[6];
    virial9[0] = v0; virial9[1] = v1; virial9[2] = v2;
    virial9[3] = v1; virial9[4] = v3; virial9[5] = v4;
    virial9[6] = v2; virial9[7] = v4; virial9[8] = v5;
  }
  free(part);
  free(b1);
  free(b2);
  free(b3);
}

/* ------------------------------------------------------------------------ */
/* a7: gather.  gather_force_ortho CudaPMERecip.cu:1031-1245                  */
/* ------------------------------------------------------------------------ */
void orc_gather(const orc_params* p, int natom, const float* xyzq, const double* phi,
                double* force) {
  const int nx = p->nx, ny = p->ny, nz = p->nz, order = p->order;
  int nt = pick_threads(p->nthreads);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int a = 0; a < natom; a++) {
    int ix[3];
    double w[3];
    double thx[16], thy[16], thz[16], dthx[16], dthy[16], dthz[16];
    atom_index(p, &xyzq[4 * a], ix, w);
    orc_bspline(order, w[0], thx, dthx);
    orc_bspline(order, w[1], thy, dthy);
    orc_bspline(order, w[2], thz, dthz);
    double f1 = 0.0, f2 = 0.0, f3 = 0.0;
    for (int k = 0; k < order; k++) {
      int z = ix[2] + k;
      if (z >= nz) z -= nz;
      for (int j = 0; j < order; j++) {
        int y = ix[1] + j;
        if (y >= ny) y -= ny;
        for (int i = 0; i < order; i++) {
          int x = ix[0] + i;
          if (x >= nx) x -= nx;
          double v = phi[x + (size_t)nx * (y + (size_t)ny * z)];
          f1 += dthx[i] * thy[j] * thz[k] * v;
          f2 += thx[i] * dthy[j] * thz[k] * v;
          f3 += thx[i] * thy[j] * dthz[k] * v;
        }
      }
    }
    double q = (double)xyzq[4 * a + 3] * ORC_CCELEC;
    /* dE/dr_b = sum_a dE/ds_a R[a][b], dE/ds_a = q n_a f_a (CudaPMERecip.cu:1238-1241
     * for the diagonal recip) */
    const double* R = p->recip;
    const double g1 = q * f1 * nx, g2 = q * f2 * ny, g3 = q * f3 * nz;
    if (is_tric(R)) {
      force[3 * a + 0] = g1 * R[0] + g2 * R[3] + g3 * R[6];
      force[3 * a + 1] = g1 * R[1] + g2 * R[4] + g3 * R[7];
      force[3 * a + 2] = g1 * R[2] + g2 * R[5] + g3 * R[8];
    } else {
      force[3 * a + 0] = q * R[0] * f1 * nx;
      force[3 * a + 1] = q * R[4] * f2 * ny;
      force[3 * a + 2] = q * R[8] * f3 * nz;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* a8: self energy.  calc_self_energy_kernel CudaPMERecip.cu:1252-1275        */
/* ------------------------------------------------------------------------ */
double orc_self_energy(int natom, const float* xyzq, double kappa) {
  double s = 0.0;
  for (int a = 0; a < natom; a++) {
    float q = xyzq[4 * a + 3];
    s += (double)(q * q);   /* q*q formed in FP32, summed in FP64, as the reference */
  }
  return -s * kappa * ORC_CCELEC / sqrt(PI);
}

/* ------------------------------------------------------------------------ */
/* Whole solve in the reference's stage order (gpu_recip.cu:118-138).         */
/* ------------------------------------------------------------------------ */
int orc_solve(const orc_params* p, int natom, const float* xyzq, double* force,
              double* energy, double* virial9, double* grid_out, double* phi_out) {
  if (p->nx < 2 || p->ny < 2 || p->nz < 2 || p->order < 3 || p->order > 16) return -1;
  const size_t G = (size_t)p->nx * p->ny * p->nz;
  const size_t Gc = (size_t)(p->nx / 2 + 1) * p->ny * p->nz;
  double* grid = grid_out ? grid_out : (double*)malloc(sizeof(double) * G);
  double* phi = phi_out ? phi_out : (double*)malloc(sizeof(double) * G);
  double* spec = (double*)malloc(sizeof(double) * 2 * Gc);
  if (!grid || !phi || !spec) return -2;
  orc_spread(p, natom, xyzq, grid);
  orc_fft3d_r2c(p->nx, p->ny, p->nz, grid, spec, p->nthreads);
  double e = 0.0, v[9];
  orc_scalar_sum(p, spec, &e, v);
  orc_fft3d_c2r(p->nx, p->ny, p->nz, spec, phi, p->nthreads);
  if (force) orc_gather(p, natom, xyzq, phi, force);
  if (energy) *energy = e;
  if (virial9) memcpy(virial9, v, sizeof(v));
  free(spec);
  if (!grid_out) free(grid);
  if (!phi_out) free(phi);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Brute-force Ewald reciprocal sum (independent of grids and splines).       */
/* E = ccelec/(2 pi V) sum_{m != 0} exp(-pi^2 |m|^2/kappa^2)/|m|^2 |S(m)|^2   */
/* with m = (m1/Lx, m2/Ly, m3/Lz), S(m) = sum_j q_j exp(2 pi i m.r_j).        */
/* ------------------------------------------------------------------------ */
void orc_ewald_recip(int natom, const float* xyzq, const double box[3], double kappa,
                     const int kmax[3], double* energy, double* force, double* virial9,
                     int nthreads) {
  const double r9[9] = {1.0 / box[0], 0, 0, 0, 1.0 / box[1], 0, 0, 0, 1.0 / box[2]};
  orc_ewald_recip9(natom, xyzq, r9, kappa, kmax, energy, force, virial9, nthreads);
}

void orc_ewald_recip9(int natom, const float* xyzq, const double recip9[9], double kappa,
                      const int kmax[3], double* energy, double* force, double* virial9,
                      int nthreads) {
  const double* R = recip9;
  const double V = 1.0 / (R[0] * (R[4] * R[8] - R[5] * R[7]) - R[1] * (R[3] * R[8] - R[5] * R[6]) +
                          R[2] * (R[3] * R[7] - R[4] * R[6]));
  const double fac = PI * PI / (kappa * kappa);
  const int K1 = 2 * kmax[0] + 1, K2 = 2 * kmax[1] + 1, K3 = 2 * kmax[2] + 1;
  const long long nk = (long long)K1 * K2 * K3;
  int nt = pick_threads(nthreads);
  double e_tot = 0.0;
  double vir[6] = {0, 0, 0, 0, 0, 0};
  if (force) memset(force, 0, sizeof(double) * 3 * natom);
#pragma omp parallel num_threads(nt)
  {
    double e_loc = 0.0, v_loc[6] = {0, 0, 0, 0, 0, 0};
    double* f_loc = force ? (double*)calloc((size_t)3 * natom, sizeof(double)) : NULL;
    double* cs = (double*)malloc(sizeof(double) * 2 * natom);
#pragma omp for schedule(dynamic, 16)
    for (long long t = 0; t < nk; t++) {
      int m1 = (int)(t % K1) - kmax[0];
      int m2 = (int)((t / K1) % K2) - kmax[1];
      int m3 = (int)(t / ((long long)K1 * K2)) - kmax[2];
      if (m1 == 0 && m2 == 0 && m3 == 0) continue;
      double h1 = R[0] * m1 + R[3] * m2 + R[6] * m3;
      double h2 = R[1] * m1 + R[4] * m2 + R[7] * m3;
      double h3 = R[2] * m1 + R[5] * m2 + R[8] * m3;
      double msq = h1 * h1 + h2 * h2 + h3 * h3;
      double g = exp(-fac * msq) / msq;
      double sr = 0.0, si = 0.0;
      for (int j = 0; j < natom; j++) {
        double arg = 2.0 * PI * (h1 * xyzq[4 * j] + h2 * xyzq[4 * j + 1] + h3 * xyzq[4 * j + 2]);
        double c = cos(arg), s = sin(arg), q = xyzq[4 * j + 3];
        cs[2 * j] = c;
        cs[2 * j + 1] = s;
        sr += q * c;
        si += q * s;
      }
      double s2 = sr * sr + si * si;
      double ek = ORC_CCELEC / (2.0 * PI * V) * g * s2;
      e_loc += ek;
      double vt = 2.0 * (fac + 1.0 / msq);
      v_loc[0] += ek * (vt * h1 * h1 - 1.0);
      v_loc[1] += ek * vt * h1 * h2;
      v_loc[2] += ek * vt * h1 * h3;
      v_loc[3] += ek * (vt * h2 * h2 - 1.0);
      v_loc[4] += ek * vt * h2 * h3;
      v_loc[5] += ek * (vt * h3 * h3 - 1.0);
      if (f_loc) {
        /* dE/dr_j = ccelec/(2 pi V) g * 2 Re[S^* q_j 2 pi i m e^{i arg_j}]
         *         = -(2 ccelec q_j / V) g m (cos_j * si - sin_j * sr) * (-1) */
        double pre = ORC_CCELEC / (2.0 * PI * V) * g * 2.0 * 2.0 * PI;
        for (int j = 0; j < natom; j++) {
          double q = xyzq[4 * j + 3];
          /* Re[ (sr - i si) * i * (c + i s) ] = Re[ i (sr c + si s) + (sr (i*i) s ...) ]
           * = -(sr s - si c) */
          double re = -(sr * cs[2 * j + 1] - si * cs[2 * j]);
          double d = pre * q * re;
          f_loc[3 * j + 0] += d * h1;
          f_loc[3 * j + 1] += d * h2;
          f_loc[3 * j + 2] += d * h3;
        }
      }
    }
#pragma omp critical
    {
      e_tot += e_loc;
      for (int i = 0; i < 6; i++) vir[i] += v_loc[i];
      if (f_loc)
        for (int i = 0; i < 3 * natom; i++) force[i] += f_loc[i];
    }
    free(cs);
    free(f_loc);
  }
  if (energy) *energy = e_tot;
  if (virial9) {
    virial9[0] = -vir[0]; virial9[1] = -vir[1]; virial9[2] = -vir[2];
    virial9[3] = -vir[1]; virial9[4] = -vir[3]; virial9[5] = -vir[4];
    virial9[6] = -vir[2]; virial9[7] = -vir[4]; virial9[8] = -vir[5];
  }
}

/* ------------------------------------------------------------------------
 * Direct-space Ewald (the real-space half; CudaPMEDirectForce).
 * Pair term for the EWALD model, CudaDirectForceKernels.cu:352-390:
 *   U(r) = ccelec qi qj (erfc(kappa r) + fac - 1) / r,
 * fac = 1 for the non-bonded pairs (CudaDirectForce_util.h:184-251: only
 * pairs within roff that the exclusion masks leave), e14fac for the 1-4
 * interactions and 0 for the 1-4 exclusions (:506-660).  Gradient +dU/dr.
 * ------------------------------------------------------------------------ */
static void pair_ewald(double dx, double dy, double dz, double qq, double kappa, double fac, double* u,
                       double g[3]) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double r = sqrt(r2);
  const double e = (erfc(kappa * r) + fac - 1.0) / r;
  /* r dU/dr = -qq (2 kappa / sqrt(pi)) exp(-kappa^2 r^2) - qq e */
  const double rdu = -qq * (2.0 / sqrt(PI)) * kappa * exp(-kappa * kappa * r2) - qq * e;
  *u = qq * e;
  const double s = rdu / r2;
  g[0] = s * dx;
  g[1] = s * dy;
  g[2] = s * dz;
}

static int excluded(int i, int j, int nexcl, const int* excl) {
  for (int k = 0; k < nexcl; k++) {
    const int a = excl[2 * k], b = excl[2 * k + 1];
    if ((a == i && b == j) || (a == j && b == i)) return 1;
  }
  return 0;
}

double orc_direct_ewald(int natom, const float* xyzq, const double box[3], double kappa, double roff, int nexcl,
                        const int* excl, double* force, double* virial9, int nthreads) {
  const int nt = pick_threads(nthreads);
  const double roff2 = roff * roff;
  double e_tot = 0.0, vir[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma omp parallel num_threads(nt)
  {
    double e_loc = 0.0, v_loc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double* f_loc = force ? (double*)calloc((size_t)3 * natom, sizeof(double)) : NULL;
#pragma omp for schedule(dynamic, 8)
    for (int i = 0; i < natom; i++) {
      for (int j = i + 1; j < natom; j++) {
        if (nexcl && excluded(i, j, nexcl, excl)) continue;
        const double qq = ORC_CCELEC * (double)xyzq[4 * i + 3] * (double)xyzq[4 * j + 3];
        for (int s = 0; s < 27; s++) {
          const double sx = (s % 3 - 1) * box[0], sy = ((s / 3) % 3 - 1) * box[1], sz = (s / 9 - 1) * box[2];
          const double dx = (double)xyzq[4 * i] - (double)xyzq[4 * j] + sx;
          const double dy = (double)xyzq[4 * i + 1] - (double)xyzq[4 * j + 1] + sy;
          const double dz = (double)xyzq[4 * i + 2] - (double)xyzq[4 * j + 2] + sz;
          if (dx * dx + dy * dy + dz * dz >= roff2) continue;
          double u, g[3];
          pair_ewald(dx, dy, dz, qq, kappa, 1.0, &u, g);
          e_loc += u;
          const double d[3] = {dx, dy, dz};
          for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) v_loc[3 * a + b] += d[a] * g[b];
          if (f_loc)
            for (int a = 0; a < 3; a++) {
              f_loc[3 * i + a] += g[a];
              f_loc[3 * j + a] -= g[a];
            }
        }
      }
    }
#pragma omp critical
    {
      e_tot += e_loc;
      for (int k = 0; k < 9; k++) vir[k] += v_loc[k];
      if (f_loc)
        for (int k = 0; k < 3 * natom; k++) force[k] += f_loc[k];
    }
    free(f_loc);
  }
  if (virial9)
    for (int k = 0; k < 9; k++) virial9[k] += vir[k];
  return e_tot;
}

double orc_pair14(int natom, const float* xyzq, const double box[3], double kappa, double fac, int npair,
                  const int* list, double* force, double* virial9) {
  double e = 0.0;
  for (int k = 0; k < npair; k++) {
    const int i = list[3 * k], j = list[3 * k + 1], s = list[3 * k + 2];
    if (i < 0 || j < 0 || i >= natom || j >= natom) continue;
    /* calc_box_shift, gpu_utils.h:312-320 */
    const double sx = (s % 3 - 1) * box[0], sy = ((s / 3) % 3 - 1) * box[1], sz = (s / 9 - 1) * box[2];
    const double dx = (double)xyzq[4 * i] - (double)xyzq[4 * j] + sx;
    const double dy = (double)xyzq[4 * i + 1] - (double)xyzq[4 * j + 1] + sy;
    const double dz = (double)xyzq[4 * i + 2] - (double)xyzq[4 * j + 2] + sz;
    const double qq = ORC_CCELEC * (double)xyzq[4 * i + 3] * (double)xyzq[4 * j + 3];
    double u, g[3];
    pair_ewald(dx, dy, dz, qq, kappa, fac, &u, g);
    e += u;
    const double d[3] = {dx, dy, dz};
    if (virial9)
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) virial9[3 * a + b] += d[a] * g[b];
    if (force)
      for (int a = 0; a < 3; a++) {
        force[3 * i + a] += g[a];
        force[3 * j + a] -= g[a];
      }
  }
  return e;
}

double orc_direct_tiles(int npad, const float* xyzq, const double box[3], double kappa, double roff, int n_ientry,
                        const int* ientry, const int* tile_indj, const unsigned* tile_excl, double* force,
                        double* virial9, double* sforce, int nthreads) {
  const int nt = pick_threads(nthreads);
  const double roff2 = roff * roff;
  double e_tot = 0.0, vir[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, sf[81];
  for (int k = 0; k < 81; k++) sf[k] = 0.0;
#pragma omp parallel num_threads(nt)
  {
    double e_loc = 0.0, v_loc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, sf_loc[81];
    for (int k = 0; k < 81; k++) sf_loc[k] = 0.0;
    double* f_loc = force ? (double*)calloc((size_t)3 * npad, sizeof(double)) : NULL;
#pragma omp for schedule(dynamic, 4)
    for (int e = 0; e < n_ientry; e++) {
      const int i0 = ientry[4 * e], ish = ientry[4 * e + 1], t0 = ientry[4 * e + 2], t1 = ientry[4 * e + 3];
      /* calc_box_shift, gpu_utils.h:312-320 */
      const double sx = (ish % 3 - 1) * box[0], sy = ((ish / 3) % 3 - 1) * box[1], sz = (ish / 9 - 1) * box[2];
      double gi[3] = {0.0, 0.0, 0.0};
      for (int t = t0; t <= t1; t++) {
        const int j0 = tile_indj[t];
        for (int l = 0; l < 32; l++) {
          const unsigned w = tile_excl[32 * (size_t)t + l];
          const int j = j0 + l;
          for (int b = 0; b < 32; b++) {
            if ((w >> b) & 1u) continue;
            const int i = i0 + (l + b) % 32;
            if (i < 0 || j < 0 || i >= npad || j >= npad) continue;
            const double dx = (double)xyzq[4 * i] + sx - (double)xyzq[4 * j];
            const double dy = (double)xyzq[4 * i + 1] + sy - (double)xyzq[4 * j + 1];
            const double dz = (double)xyzq[4 * i + 2] + sz - (double)xyzq[4 * j + 2];
            if (dx * dx + dy * dy + dz * dz >= roff2) continue;
            const double qq = ORC_CCELEC * (double)xyzq[4 * i + 3] * (double)xyzq[4 * j + 3];
            double u, g[3];
            pair_ewald(dx, dy, dz, qq, kappa, 1.0, &u, g);
            e_loc += u;
            const double d[3] = {dx, dy, dz};
            for (int a = 0; a < 3; a++) {
              gi[a] += g[a];
              for (int c = 0; c < 3; c++) v_loc[3 * a + c] += d[a] * g[c];
            }
            if (f_loc)
              for (int a = 0; a < 3; a++) {
                f_loc[3 * i + a] += g[a];
                f_loc[3 * j + a] -= g[a];
              }
          }
        }
      }
      if (ish != 13)
        for (int a = 0; a < 3; a++) sf_loc[3 * ish + a] += gi[a];
    }
#pragma omp critical
    {
      e_tot += e_loc;
      for (int k = 0; k < 9; k++) vir[k] += v_loc[k];
      for (int k = 0; k < 81; k++) sf[k] += sf_loc[k];
      if (f_loc)
        for (int k = 0; k < 3 * npad; k++) force[k] += f_loc[k];
    }
    free(f_loc);
  }
  if (virial9)
    for (int k = 0; k < 9; k++) virial9[k] += vir[k];
  if (sforce)
    for (int k = 0; k < 81; k++) sforce[k] += sf[k];
  return e_tot;
}

