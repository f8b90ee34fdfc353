// pme_plan.cpp -- host-side planning (see pme_plan.hpp).
#include "pme_plan.hpp"

#include <algorithm>
#include <cmath>

namespace pme {

static const double kPi = 3.14159265358979323846;

// B-spline values at w (FP64), fill_bspline_host CudaPMERecip.cu:2507-2537
static void bspline_host(int order, double w, double* th) {
  th[order - 1] = 0.0;
  th[1] = w;
  th[0] = 1.0 - w;
  for (int k = 3; k <= order; k++) {
    const double div = 1.0 / (double)(k - 1);
    th[k - 1] = div * w * th[k - 2];
    for (int j = 1; j <= k - 2; j++) th[k - j - 1] = div * ((w + j) * th[k - j - 2] + (k - j - w) * th[k - j - 1]);
    th[0] = div * (1.0 - w) * th[0];
  }
}

void bsp_mod_host(int order, int n, double* out) {
  std::vector<double> arr(std::max(n, order + 1), 0.0), th(order);
  bspline_host(order, 0.0, th.data());
  for (int i = 2; i <= order + 1; i++) arr[i - 1] = th[i - 2];
  for (int k = 0; k < n; k++) {
    double s1 = 0.0, s2 = 0.0;
    for (int j = 0; j < n - 1; j++) {
      const long long jk = ((long long)j * k) % n;
      const double arg = 2.0 * kPi * (double)jk / (double)n;
      s1 += arr[j] * std::cos(arg);
      s2 += arr[j] * std::sin(arg);
    }
    out[k] = s1 * s1 + s2 * s2;
  }
  for (int k = 0; k < n; k++)
    if (out[k] < 1.0e-10) out[k] = 0.5 * (out[(k + n - 1) % n] + out[(k + 1) % n]);
  for (int k = 0; k < n; k++) out[k] = 1.0 / out[k];
}

bool smooth235(int n) {
  if (n < 1) return false;
  for (int f : {2, 3, 5})
    while (n % f == 0) n /= f;
  return n == 1;
}

static void pack_radices(RadixPlan* rp) {
  rp->rcode = 0;
  for (int s = 0; s < rp->nrad; s++) rp->rcode |= (unsigned long long)rp->rad[s] << (5 * s);
}

// Factor pair of the register four-step transform: n = a * b with both
// codelet lengths (2, 3, 4, 5, 6, 8, 9, 10, 12, 16), the most balanced such
// pair; (0, 0) when there is none.
static void four_step_factors(int n, int* fa, int* fb) {
  static const int L[] = {2, 3, 4, 5, 6, 8, 9, 10, 12, 16};
  *fa = *fb = 0;
  int best = 1 << 30;
  for (int a : L)
    for (int b : L)
      if (a * b == n && a <= b && b < best) {
        best = b;
        *fa = a;
        *fb = b;
      }
}

bool radix_plan(int n, int tstride, RadixPlan* rp) {
  rp->n = n;
  rp->tstride = tstride;
  four_step_factors(n, &rp->f4a, &rp->f4b);
  rp->nrad = 0;
  int m = n;
  while (m % 8 == 0 && m > 8) { rp->rad[rp->nrad++] = 8; m /= 8; }
  while (m % 4 == 0) { rp->rad[rp->nrad++] = 4; m /= 4; }
  while (m % 2 == 0) { rp->rad[rp->nrad++] = 2; m /= 2; }
  while (m % 8 == 0) { rp->rad[rp->nrad++] = 8; m /= 8; }
  while (m % 5 == 0) { rp->rad[rp->nrad++] = 5; m /= 5; }
  while (m % 3 == 0) { rp->rad[rp->nrad++] = 3; m /= 3; }
  if (rp->nrad > 12) return false;
  pack_radices(rp);
  return m == 1;
}

bool radix_plan_split(int n, int tstride, RadixPlan* rp) {
  if (n % 16 != 0) return false;
  RadixPlan rest;
  if (n > 16) {
    if (!radix_plan(n / 16, tstride, &rest) || rest.nrad + 1 > 12) return false;
  } else {
    rest.nrad = 0;
  }
  rp->n = n;
  rp->tstride = tstride;
  // a register four-step with the 16-point DFTs on the matrix cores when n/16
  // is a codelet length; otherwise the matrix-core stage opens the Stockham
  // stages
  rp->f4a = rp->f4b = 0;
  for (int b : {2, 3, 4, 5, 6, 8, 9, 10, 12, 16})
    if (n == 16 * b) {
      rp->f4a = 16;
      rp->f4b = b;
    }
  rp->nrad = rest.nrad + 1;
  rp->rad[0] = 16;
  for (int s = 0; s < rest.nrad; s++) rp->rad[s + 1] = rest.rad[s];
  pack_radices(rp);
  return true;
}

std::vector<float2> twiddles(int n) {
  std::vector<float2> t(n);
  for (int k = 0; k < n; k++) {
    const double ang = 2.0 * kPi * (double)k / (double)n;
    t[k] = make_float2((float)std::cos(ang), (float)(-std::sin(ang)));
  }
  return t;
}

bool plan_bricks(const int n[3], int order, int target, BrickPlan* bp, size_t budget) {
  for (int a = 0; a < 3; a++) {
    int nb = (n[a] + target / 2) / target;
    if (nb < 1) nb = 1;
    while (nb > 1 && n[a] / nb < order - 1) nb--;
    bp->nb[a] = nb;
    bp->wmax[a] = (n[a] + nb - 1) / nb;
  }
  while (spread_lds_bytes(order, bp->wmax, kMaxChunks) > budget || gather_lds_bytes(order, bp->wmax, kMaxChunks) > budget) {
    int best = -1;
    for (int a = 0; a < 3; a++)
      if (n[a] / (bp->nb[a] + 1) >= order - 1 && (best < 0 || bp->wmax[a] > bp->wmax[best])) best = a;
    if (best < 0) return false;
    bp->nb[best]++;
    bp->wmax[best] = (n[best] + bp->nb[best] - 1) / bp->nb[best];
  }
  bp->nbins = bp->nb[0] * bp->nb[1] * bp->nb[2];
  return bp->nbins <= kMaxBins && bin_lds_bytes(bp->nbins) <= kLdsBudget;
}

// Modelled spread time per CU of a brick plan (us): rounds of bricks over the
// CUs times a per-brick fixed cost plus, per grid point of a brick, the store
// and the stencil work of its records (halo factor prod(1 + (p-1)/w) records
// per atom, p^3 stencil points each).  Constants from timing variants of the
// spread at 96,000 atoms, 108x100x96, order 6, brick 16: 7.7 us fixed per brick
// (start, clear, segment bounds and scan), 8.3e-4 us per point stored and
// 1.59e-3 us per point and record of an order-6 stencil; the gather's phases
// scale alike.  It ranks the measured plans right: at 96k atoms brick 16
// (99.9 us per solve) ahead of 18/20/24 (105.8/116.9/124.1), at 250,002 atoms,
// 144^3, order 4 brick 24 (145.2) ahead of 18/20/16 (149.1/157.3/161.2).
static double brick_cost(const int n[3], int order, int ncu, const BrickPlan& bp) {
  const double g = (double)n[0] * n[1] * n[2];
  const double pts = g / bp.nbins;
  double halo = 1.0;
  for (int a = 0; a < 3; a++) halo *= 1.0 + (double)(order - 1) * bp.nb[a] / n[a];
  const double rounds = (double)((bp.nbins + ncu - 1) / ncu);
  const double p3 = (double)order * order * order / 216.0;
  return rounds * (7.7 + pts * (8.3e-4 + 1.59e-3 * halo * p3));
}

bool plan_bricks_auto(const int n[3], int order, int ncu, BrickPlan* bp) {
  bool any = false;
  double best = 0.0;
  for (int target : {16, 18, 20, 24, 28, 32}) {
    BrickPlan t;
    if (!plan_bricks(n, order, target, &t)) continue;
    const double c = brick_cost(n, order, std::max(ncu, 1), t);
    if (!any || c < best) {
      best = c;
      *bp = t;
      any = true;
    }
  }
  return any;
}

FftTiles pick_fft_tiles(int nx, int ny, int nz, int nf_x, int tile_target, bool z_four_step) {
  FftTiles t;
  const int hx = nx / 2 + 1;
  // x rows per workgroup: enough that the four-step units of the x pass fill
  // the workgroup (at 144^3: 32 rows of 72, 142.7 -> 139.0 us per solve against
  // 8 rows in Stockham stages)
  t.NR = nf_x <= 64 ? 64 : (nf_x <= 128 ? 32 : (nf_x <= 256 ? 16 : 8));
  while (t.NR > 1 && fft_x_lds_bytes(nf_x, t.NR, nx) > kLdsBudget) t.NR /= 2;
  // column tiles of equal width covering hx without padding: the fewest
  // tiles of at most tile_target columns whose LDS fits the budget
  auto tile = [&](int n) {
    int ntile = (hx + tile_target - 1) / tile_target;
    int tx = (hx + ntile - 1) / ntile;
    while (tx > 1 && fft_cols_lds_bytes(n, tx) > kLdsBudget) {
      ntile++;
      tx = (hx + ntile - 1) / ntile;
    }
    return tx;
  };
  t.TXy = tile(ny);
  // z: the lines are the flattened (ky, kx) of a plane, so the tile width
  // need not divide hx; a power of two keeps every tile on whole, aligned
  // cache lines of the plane
  int tz = 1;
  while (tz * 2 <= tile_target) tz *= 2;
  while (tz > 1 && fft_cols_lds_bytes(nz, tz, z_four_step ? 1 : 2) > kLdsBudget) tz /= 2;
  t.TXz = tz;
  return t;
}

ChunkPlan plan_chunks(int natom) {
  ChunkPlan c;
  c.chunk = kMinChunk;
  while ((long long)c.chunk * kMaxChunks < natom) c.chunk *= 2;
  c.nchunk = std::max(1, (natom + c.chunk - 1) / c.chunk);
  return c;
}

}  // namespace pme
