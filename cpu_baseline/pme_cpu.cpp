// pme_cpu.cpp -- FP32 CPU restatement of the reference's PME reciprocal solve,
// timed as the CPU baseline of the benchmark.  It is a baseline, not part of
// the product (gpu_recip_amd/ never loads it) and not the checker (that is the
// FP64 oracle in oracle/).
//
// The reference (ap-hynninen/gpu_recip) has no CPU PME path; this restates its
// GPU algorithm with the reference's arithmetic on host cores:
//   - grid index: w = fmaf(x, r, 2), fr = n (w - (floor(w + 0.5) - 0.5)),
//     i = (int)fr, in FP32 (CudaPMERecip.cu:512-529; the FMA is the -fmad=true
//     contraction of Makefile:174);
//   - B-splines: the FP32 recursion of calc_theta_dtheta (CudaPMERecip.cu:120-208);
//   - spread: q theta_x theta_y theta_z accumulated in an FP32 grid
//     (spread_charge_ortho with AT = float, :479-580), stencil point (i+j) mod n;
//   - FFT: 3-D R2C / C2R, e^{-2 pi i} forward, e^{+2 pi i} inverse, unnormalised
//     (the cuFFT calls of :2421 and :2464), as FP32 mixed-radix Stockham
//     transforms written here (no FFT library in the image);
//   - scalar sum: FP32 eterm per point, FP64 sums of energy and virial with the
//     reference's weights (:594-901, :1929-2047), bsp_mod in FP64 (:2481-2579);
//   - gather: FP32 sums, output q ccelec recip_ii n_i f_i = +dE/dr (:1031-1245).
// Multithreaded with OpenMP; the spread splits the grid into z slabs, one per
// thread, so its FP32 sums run in atom order whatever the thread count.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

namespace {

typedef std::complex<float> cf;
const double kPi = 3.14159265358979323846;
const double kCcelec = 332.0716;  // CudaPMERecip.h:27

// ------------------------------------------------------------------ index --
// The only fused multiply-add of the reference's index expression is the
// explicit one; contraction is off for this function.
__attribute__((optimize("fp-contract=off"))) void grid_index(float x, float r, int n, int& i, float& w) {
  const float t = fmaf(x, r, 2.0f);
  const float fr = (float)n * (t - (floorf(t + 0.5f) - 0.5f));
  i = (int)fr;
  w = fr - (float)i;
}

int wrap_index(int i, int n) {
  if (i >= 0 && i < n) return i;
  i %= n;
  return i < 0 ? i + n : i;
}

// theta[j] = M_p(w + p-1-j) and its derivative, CudaPMERecip.cu:120-176
void bspline(int P, float w, float* th, float* dth) {
  th[P - 1] = 0.0f;
  th[1] = w;
  th[0] = 1.0f - w;
  for (int k = 3; k <= P - 1; k++) {
    const float div = 1.0f / (float)(k - 1);
    th[k - 1] = div * w * th[k - 2];
    for (int j = 1; j <= k - 2; j++) th[k - j - 1] = div * ((w + j) * th[k - j - 2] + (k - j - w) * th[k - j - 1]);
    th[0] = div * (1.0f - w) * th[0];
  }
  if (dth) {
    dth[0] = -th[0];
    for (int j = 2; j <= P; j++) dth[j - 1] = th[j - 2] - th[j - 1];
  }
  const float div = 1.0f / (float)(P - 1);
  th[P - 1] = div * w * th[P - 2];
  for (int j = 1; j <= P - 2; j++) th[P - j - 1] = div * ((w + j) * th[P - j - 2] + (P - j - w) * th[P - j - 1]);
  th[0] = div * (1.0f - w) * th[0];
}

// bsp_mod in FP64: dftmod of the B-spline values at the integer nodes,
// CudaPMERecip.cu:2481-2537 (the sum stops at j = n-2; tiny values replaced by
// the neighbour average)
std::vector<float> bsp_mod(int P, int n) {
  std::vector<double> arr(std::max(n, P + 1), 0.0), th(P, 0.0);
  th[P - 1] = 0.0;
  th[1] = 0.0;
  th[0] = 1.0;
  for (int k = 3; k <= P; k++) {
    const double div = 1.0 / (double)(k - 1);
    th[k - 1] = 0.0;
    for (int j = 1; j <= k - 2; j++) th[k - j - 1] = div * ((double)j * th[k - j - 2] + (double)(k - j) * th[k - j - 1]);
    th[0] = div * th[0];
  }
  for (int i = 2; i <= P + 1; i++) arr[i - 1] = th[i - 2];
  std::vector<double> m(n);
  for (int k = 0; k < n; k++) {
    double s1 = 0.0, s2 = 0.0;
    for (int j = 0; j < n - 1; j++) {
      const double a = 2.0 * kPi * (double)(((long long)j * k) % n) / (double)n;
      s1 += arr[j] * std::cos(a);
      s2 += arr[j] * std::sin(a);
    }
    m[k] = s1 * s1 + s2 * s2;
  }
  for (int k = 0; k < n; k++)
    if (m[k] < 1.0e-10) m[k] = 0.5 * (m[(k + n - 1) % n] + m[(k + 1) % n]);
  std::vector<float> out(n);
  for (int k = 0; k < n; k++) out[k] = (float)(1.0 / m[k]);
  return out;
}

// -------------------------------------------------------------------- FFT --
// Batched mixed-radix Stockham FFT over B interleaved lines: element j of line
// b at buf[j * B + b], so the inner loops run over lines and vectorise.
struct FftPlan {
  int n = 0;
  std::vector<int> rad;
  std::vector<cf> tw;  // e^{-2 pi i k/n}
  explicit FftPlan(int n_ = 1) : n(n_) {
    int m = n;
    for (int r : {4, 2, 3, 5})
      while (m % r == 0) {
        rad.push_back(r);
        m /= r;
      }
    tw.resize(n);
    for (int k = 0; k < n; k++) tw[k] = cf((float)std::cos(2 * kPi * k / n), (float)-std::sin(2 * kPi * k / n));
  }
};

template <int R>
inline void dft_small(cf* u, float sign) {
  if (R == 2) {
    const cf t = u[1];
    u[1] = u[0] - t;
    u[0] = u[0] + t;
  } else if (R == 3) {
    const float c = -0.5f, s = 0.866025403784438647f * sign;
    const cf t = u[1] + u[2], d = u[1] - u[2];
    const cf m = u[0] + c * t;
    const cf r(-s * d.imag(), s * d.real());
    u[0] = u[0] + t;
    u[1] = m + r;
    u[2] = m - r;
  } else if (R == 4) {
    const cf t0 = u[0] + u[2], t1 = u[0] - u[2], t2 = u[1] + u[3];
    const cf d = u[1] - u[3];
    const cf t3(-sign * d.imag(), sign * d.real());
    u[0] = t0 + t2;
    u[2] = t0 - t2;
    u[1] = t1 + t3;
    u[3] = t1 - t3;
  } else {  // 5
    const float c1 = 0.309016994374947424f, c2 = -0.809016994374947424f;
    const float s1 = 0.951056516295153572f * sign, s2 = 0.587785252292473129f * sign;
    const cf t1 = u[1] + u[4], d1 = u[1] - u[4], t2 = u[2] + u[3], d2 = u[2] - u[3];
    const cf m1 = u[0] + c1 * t1 + c2 * t2, m2 = u[0] + c2 * t1 + c1 * t2;
    const cf e1 = s1 * d1 + s2 * d2, e2 = s2 * d1 - s1 * d2;
    const cf r1(-e1.imag(), e1.real()), r2(-e2.imag(), e2.real());
    u[0] = u[0] + t1 + t2;
    u[1] = m1 + r1;
    u[4] = m1 - r1;
    u[2] = m2 + r2;
    u[3] = m2 - r2;
  }
}

template <int R>
void stockham_stage(const cf* src, cf* dst, const FftPlan& pl, int p, int B, float sign) {
  const int n = pl.n, T = n / R, step = n / (p * R);
  for (int i = 0; i < T; i++) {
    const int k = i % p, ip = i / p;
    cf w[R];
    for (int r = 0; r < R; r++) {
      w[r] = pl.tw[(r * k * step) % n];
      if (sign > 0) w[r] = std::conj(w[r]);
    }
    const int j = ip * p * R + k;
    for (int b = 0; b < B; b++) {
      cf u[R];
      for (int r = 0; r < R; r++) u[r] = src[(size_t)(i + r * T) * B + b] * w[r];
      dft_small<R>(u, sign);
      for (int r = 0; r < R; r++) dst[(size_t)(j + r * p) * B + b] = u[r];
    }
  }
}

// Transform the B lines in buf (buf -> result, tmp as the other buffer);
// returns the buffer holding the result.
cf* fft_lines(const FftPlan& pl, cf* buf, cf* tmp, int B, float sign) {
  cf* src = buf;
  cf* dst = tmp;
  int p = 1;
  for (int R : pl.rad) {
    switch (R) {
      case 2: stockham_stage<2>(src, dst, pl, p, B, sign); break;
      case 3: stockham_stage<3>(src, dst, pl, p, B, sign); break;
      case 4: stockham_stage<4>(src, dst, pl, p, B, sign); break;
      default: stockham_stage<5>(src, dst, pl, p, B, sign); break;
    }
    std::swap(src, dst);
    p *= R;
  }
  return src;
}

const int kLineBatch = 16;

struct Solver {
  int nx, ny, nz, P, hx;
  float r1, r2, r3;
  double kappa;
  FftPlan px, py, pz;
  std::vector<float> pf1, pf2, pf3;
  std::vector<float> grid;  // nz * ny * nx
  std::vector<cf> spec;     // nz * ny * hx

  Solver(int nx_, int ny_, int nz_, int P_)
      : nx(nx_), ny(ny_), nz(nz_), P(P_), hx(nx_ / 2 + 1), px(nx_), py(ny_), pz(nz_),
        pf1(bsp_mod(P_, nx_)), pf2(bsp_mod(P_, ny_)), pf3(bsp_mod(P_, nz_)),
        grid((size_t)nx_ * ny_ * nz_), spec((size_t)(nx_ / 2 + 1) * ny_ * nz_) {}

  // spread_charge_ortho (AT = float): thread t owns z planes [z0, z1) and adds
  // every atom's stencil points that fall there, atoms in order
  void spread(int natom, const float* xyzq) {
    std::fill(grid.begin(), grid.end(), 0.0f);
#pragma omp parallel
    {
      const int nt = omp_get_num_threads(), t = omp_get_thread_num();
      const int z0 = (int)((long long)nz * t / nt), z1 = (int)((long long)nz * (t + 1) / nt);
      float tx[8], ty[8], tz[8];
      for (int a = 0; a < natom; a++) {
        const float* p = xyzq + 4 * (size_t)a;
        int ix, iy, iz;
        float fx, fy, fz;
        grid_index(p[2], r3, nz, iz, fz);
        iz = wrap_index(iz, nz);
        // does the stencil's z range [iz, iz + P) (periodic) meet [z0, z1)?
        bool hit = false;
        for (int j = 0; j < P && !hit; j++) {
          int z = iz + j;
          if (z >= nz) z -= nz;
          hit = z >= z0 && z < z1;
        }
        if (!hit) continue;
        grid_index(p[0], r1, nx, ix, fx);
        grid_index(p[1], r2, ny, iy, fy);
        ix = wrap_index(ix, nx);
        iy = wrap_index(iy, ny);
        bspline(P, fx, tx, nullptr);
        bspline(P, fy, ty, nullptr);
        bspline(P, fz, tz, nullptr);
        const float q = p[3];
        for (int jz = 0; jz < P; jz++) {
          int z = iz + jz;
          if (z >= nz) z -= nz;
          if (z < z0 || z >= z1) continue;
          for (int jy = 0; jy < P; jy++) {
            int y = iy + jy;
            if (y >= ny) y -= ny;
            const float qyz = q * ty[jy] * tz[jz];
            float* row = grid.data() + (size_t)nx * (y + (size_t)ny * z);
            for (int jx = 0; jx < P; jx++) {
              int x = ix + jx;
              if (x >= nx) x -= nx;
              row[x] += qyz * tx[jx];
            }
          }
        }
      }
    }
  }

  // x R2C over all rows, two real rows per complex transform; then y and z
  void r2c() {
    const int nrows = ny * nz;
#pragma omp parallel
    {
      std::vector<cf> b0((size_t)nx * kLineBatch), b1((size_t)nx * kLineBatch);
#pragma omp for schedule(static)
      for (int r0 = 0; r0 < nrows; r0 += 2 * kLineBatch) {
        const int npair = std::min(kLineBatch, (nrows - r0 + 1) / 2);
        for (int b = 0; b < npair; b++) {
          const int ra = r0 + 2 * b, rb = ra + 1;
          const float* A = grid.data() + (size_t)ra * nx;
          const float* Bq = rb < nrows ? grid.data() + (size_t)rb * nx : nullptr;
          for (int j = 0; j < nx; j++) b0[(size_t)j * npair + b] = cf(A[j], Bq ? Bq[j] : 0.0f);
        }
        const cf* Z = fft_lines(px, b0.data(), b1.data(), npair, -1.0f);
        for (int b = 0; b < npair; b++) {
          const int ra = r0 + 2 * b, rb = ra + 1;
          for (int k = 0; k < hx; k++) {
            const cf zk = Z[(size_t)k * npair + b], zc = std::conj(Z[(size_t)((nx - k) % nx) * npair + b]);
            spec[(size_t)ra * hx + k] = 0.5f * (zk + zc);
            if (rb < nrows) spec[(size_t)rb * hx + k] = cf(0.0f, -0.5f) * (zk - zc);
          }
        }
      }
    }
    cols(py, hx, (size_t)hx, ny, nz, (size_t)hx * ny, -1.0f);
    cols(pz, hx, (size_t)hx * ny, nz, ny, (size_t)hx, -1.0f);
  }

  // 1-D transforms of length n along an axis of the half spectrum: lines
  // (kx, outer) with element j at kx + j * pt + outer * ostride
  void cols(const FftPlan& pl, int nkx, size_t pt, int n, int nouter, size_t ostride, float sign) {
    const int nblk = (nkx + kLineBatch - 1) / kLineBatch;
#pragma omp parallel
    {
      std::vector<cf> b0((size_t)n * kLineBatch), b1((size_t)n * kLineBatch);
#pragma omp for schedule(static)
      for (int t = 0; t < nblk * nouter; t++) {
        const int o = t / nblk, kx0 = (t % nblk) * kLineBatch;
        const int B = std::min(kLineBatch, nkx - kx0);
        cf* base = spec.data() + (size_t)o * ostride + kx0;
        for (int j = 0; j < n; j++)
          for (int b = 0; b < B; b++) b0[(size_t)j * B + b] = base[j * pt + b];
        const cf* res = fft_lines(pl, b0.data(), b1.data(), B, sign);
        for (int j = 0; j < n; j++)
          for (int b = 0; b < B; b++) base[j * pt + b] = res[(size_t)j * B + b];
      }
    }
  }

  // scalar_sum_ortho_kernel: eterm in FP32, energy/virial sums in FP64 with
  // the reference's weights (2 for every k1 >= 1, Nyquist included)
  void scalar_sum(double* energy, double* vir9) {
    const int nf1 = nx / 2 + nx % 2, nf2 = ny / 2 + ny % 2, nf3 = nz / 2 + nz % 2;
    const float fac = (float)(kPi * kPi / (kappa * kappa));
    const float piv = piv_inv;
    std::vector<double> part((size_t)nz * 7, 0.0);
#pragma omp parallel for schedule(static)
    for (int k3 = 0; k3 < nz; k3++) {
      double acc[7] = {0, 0, 0, 0, 0, 0, 0};
      const int m3 = k3 >= nf3 ? k3 - nz : k3;
      const float mh3 = r3 * (float)m3;
      for (int k2 = 0; k2 < ny; k2++) {
        const int m2 = k2 >= nf2 ? k2 - ny : k2;
        const float mh2 = r2 * (float)m2;
        cf* row = spec.data() + (size_t)hx * (k2 + (size_t)ny * k3);
        for (int k1 = 0; k1 < hx; k1++) {
          if ((k1 | k2 | k3) == 0) {
            row[0] = cf(0.0f, 0.0f);
            continue;
          }
          const int m1 = k1 >= nf1 ? k1 - nx : k1;
          const float mh1 = r1 * (float)m1;
          const float msq = mh1 * mh1 + mh2 * mh2 + mh3 * mh3;
          const float eterm = expf(-fac * msq) * piv * pf1[k1] * pf2[k2] * pf3[k3] / msq;
          const cf q = row[k1];
          const float q2 = q.real() * q.real() + q.imag() * q.imag();
          const float tmp1 = eterm * q2;
          const float vterm = 2.0f * (fac + 1.0f / msq);
          const float tmp2 = tmp1 * vterm;
          const double w = k1 >= 1 ? 2.0 : 1.0;
          // mirror point (n1-k1, -k2, -k3): m components flip sign except on
          // a Nyquist plane or at zero, so off-diagonal terms double or cancel
          const bool f1 = !(2 * k1 == nx), f2 = !(2 * k2 == ny || k2 == 0), f3 = !(2 * k3 == nz || k3 == 0);
          const double w12 = k1 >= 1 ? (f1 == f2 ? 2.0 : 0.0) : 1.0;
          const double w13 = k1 >= 1 ? (f1 == f3 ? 2.0 : 0.0) : 1.0;
          const double w23 = k1 >= 1 ? (f2 == f3 ? 2.0 : 0.0) : 1.0;
          acc[0] += w * tmp1;
          acc[1] += w * (tmp1 * (vterm * mh1 * mh1 - 1.0f));
          acc[2] += w12 * (tmp2 * mh1 * mh2);
          acc[3] += w13 * (tmp2 * mh1 * mh3);
          acc[4] += w * (tmp1 * (vterm * mh2 * mh2 - 1.0f));
          acc[5] += w23 * (tmp2 * mh2 * mh3);
          acc[6] += w * (tmp1 * (vterm * mh3 * mh3 - 1.0f));
          row[k1] = q * eterm;
        }
      }
      for (int c = 0; c < 7; c++) part[(size_t)k3 * 7 + c] = acc[c];
    }
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int k3 = 0; k3 < nz; k3++)
      for (int c = 0; c < 7; c++) v[c] += part[(size_t)k3 * 7 + c];
    const double hc = 0.5 * kCcelec;
    *energy = hc * v[0];
    const double vx = -hc * v[1], vxy = -hc * v[2], vxz = -hc * v[3], vy = -hc * v[4], vyz = -hc * v[5],
                 vz = -hc * v[6];
    const double V[9] = {vx, vxy, vxz, vxy, vy, vyz, vxz, vyz, vz};
    std::memcpy(vir9, V, sizeof(V));
  }

  void c2r() {
    cols(pz, hx, (size_t)hx * ny, nz, ny, (size_t)hx, +1.0f);
    cols(py, hx, (size_t)hx, ny, nz, (size_t)hx * ny, +1.0f);
    const int nrows = ny * nz;
#pragma omp parallel
    {
      std::vector<cf> b0((size_t)nx * kLineBatch), b1((size_t)nx * kLineBatch);
#pragma omp for schedule(static)
      for (int r0 = 0; r0 < nrows; r0 += 2 * kLineBatch) {
        const int npair = std::min(kLineBatch, (nrows - r0 + 1) / 2);
        // Z = Xa + i Xb over the full (Hermitian) spectra of two rows
        for (int b = 0; b < npair; b++) {
          const int ra = r0 + 2 * b, rb = ra + 1;
          const cf* A = spec.data() + (size_t)ra * hx;
          const cf* Bq = rb < nrows ? spec.data() + (size_t)rb * hx : nullptr;
          for (int k = 0; k < nx; k++) {
            const cf xa = k < hx ? A[k] : std::conj(A[nx - k]);
            const cf xb = Bq ? (k < hx ? Bq[k] : std::conj(Bq[nx - k])) : cf(0.0f, 0.0f);
            b0[(size_t)k * npair + b] = xa + cf(-xb.imag(), xb.real());
          }
        }
        const cf* z = fft_lines(px, b0.data(), b1.data(), npair, +1.0f);
        for (int b = 0; b < npair; b++) {
          const int ra = r0 + 2 * b, rb = ra + 1;
          for (int j = 0; j < nx; j++) {
            grid[(size_t)ra * nx + j] = z[(size_t)j * npair + b].real();
            if (rb < nrows) grid[(size_t)rb * nx + j] = z[(size_t)j * npair + b].imag();
          }
        }
      }
    }
  }

  // gather_force_ortho: FP32 sums per atom, force AoS natom*3 (+dE/dr)
  void gather(int natom, const float* xyzq, float* force) {
    const float ccf = (float)kCcelec;
#pragma omp parallel for schedule(static)
    for (int a = 0; a < natom; a++) {
      const float* p = xyzq + 4 * (size_t)a;
      int ix, iy, iz;
      float fx, fy, fz;
      grid_index(p[0], r1, nx, ix, fx);
      grid_index(p[1], r2, ny, iy, fy);
      grid_index(p[2], r3, nz, iz, fz);
      ix = wrap_index(ix, nx);
      iy = wrap_index(iy, ny);
      iz = wrap_index(iz, nz);
      float tx[8], dtx[8], ty[8], dty[8], tz[8], dtz[8];
      bspline(P, fx, tx, dtx);
      bspline(P, fy, ty, dty);
      bspline(P, fz, tz, dtz);
      float f1 = 0.f, f2 = 0.f, f3 = 0.f;
      for (int jz = 0; jz < P; jz++) {
        int z = iz + jz;
        if (z >= nz) z -= nz;
        float ay = 0.f, by = 0.f, cy = 0.f;
        for (int jy = 0; jy < P; jy++) {
          int y = iy + jy;
          if (y >= ny) y -= ny;
          const float* row = grid.data() + (size_t)nx * (y + (size_t)ny * z);
          float ra = 0.f, rb = 0.f;
          for (int jx = 0; jx < P; jx++) {
            int x = ix + jx;
            if (x >= nx) x -= nx;
            ra += tx[jx] * row[x];
            rb += dtx[jx] * row[x];
          }
          ay += ty[jy] * ra;
          by += ty[jy] * rb;
          cy += dty[jy] * ra;
        }
        f1 += tz[jz] * by;
        f2 += tz[jz] * cy;
        f3 += dtz[jz] * ay;
      }
      const float q = p[3] * ccf;
      force[3 * (size_t)a + 0] = q * r1 * f1 * (float)nx;
      force[3 * (size_t)a + 1] = q * r2 * f2 * (float)ny;
      force[3 * (size_t)a + 2] = q * r3 * f3 * (float)nz;
    }
  }

  float piv_inv = 0.f;
};

}  // namespace

extern "C" {

// Grid indices and fractions of the FP32 formula (the same function the
// spread and gather use), idx/frac 3 per atom.
void cpu_index_f32(int natom, const float* xyzq, double r1, double r2, double r3, int nx, int ny, int nz, int* idx,
                   float* frac) {
  const float r[3] = {(float)r1, (float)r2, (float)r3};
  const int n[3] = {nx, ny, nz};
  for (int a = 0; a < natom; a++)
    for (int d = 0; d < 3; d++) grid_index(xyzq[4 * (size_t)a + d], r[d], n[d], idx[3 * a + d], frac[3 * a + d]);
}

static bool smooth(int n, int order) {
  int m = n;
  for (int f : {2, 3, 5})
    while (m % f == 0) m /= f;
  return m == 1 && n >= order;
}

// A solver for one grid and order (plans, bsp_mod tables, grid buffers), as
// the reference constructs CudaPMERecip once per system.  NULL for an
// unsupported order (> 8) or grid (not 2/3/5-smooth).
void* cpu_pme_create(int nx, int ny, int nz, int order) {
  if (order < 3 || order > 8 || !smooth(nx, order) || !smooth(ny, order) || !smooth(nz, order)) return nullptr;
  return new Solver(nx, ny, nz, order);
}

void cpu_pme_destroy(void* h) { delete static_cast<Solver*>(h); }

// One full solve on a solver: spread -> r2c -> scalar sum -> c2r -> gather.
// box: the three edges (orthorhombic); force: natom*3 AoS (+dE/dr); energy
// and virial[9] assigned.  nthreads <= 0: the OpenMP default.
int cpu_pme_run(void* h, double kappa, const double box[3], int natom, const float* xyzq, float* force,
                double* energy, double* virial9, int nthreads) {
  Solver& s = *static_cast<Solver*>(h);
  if (nthreads > 0) omp_set_num_threads(nthreads);
  // CudaPMERecip.cu:2144-2152 and :2024-2026: the diagonal cast to float, the
  // volume in FP64
  const double rd[3] = {1.0 / box[0], 1.0 / box[1], 1.0 / box[2]};
  s.r1 = (float)rd[0];
  s.r2 = (float)rd[1];
  s.r3 = (float)rd[2];
  s.kappa = kappa;
  s.piv_inv = (float)(rd[0] * rd[1] * rd[2] / kPi);
  s.spread(natom, xyzq);
  s.r2c();
  s.scalar_sum(energy, virial9);
  s.c2r();
  s.gather(natom, xyzq, force);
  return 0;
}

// Hardware threads OpenMP would use by default.
int cpu_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
