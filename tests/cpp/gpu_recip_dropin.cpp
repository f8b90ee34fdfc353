// gpu_recip_dropin.cpp -- the reference's gpu_recip test (gpu_recip.cu:68-293
// test4, :298-522 test6) written against the C++ drop-in header
// include/pme_recip.hpp: same classes, same five calls, same tolerances.
// Unlike the reference (which always returns 0) it exits non-zero on failure.
//
//   gpu_recip_dropin <xyzq.f32> <force_recip_4.f64> <force_recip_6.f64>
// xyzq.f32: N x 4 float32 (x, y, z, q); force files: N x 3 float64.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pme_recip.hpp"

#define HIPCHECK(x)                                                             \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));              \
      std::exit(2);                                                             \
    }                                                                           \
  } while (0)

template <typename T>
static std::vector<T> load(const char* path) {
  std::vector<T> v;
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    std::fprintf(stderr, "cannot open %s\n", path);
    std::exit(2);
  }
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f) / (long)sizeof(T);
  std::fseek(f, 0, SEEK_SET);
  v.resize(n);
  if (std::fread(v.data(), sizeof(T), n, f) != (size_t)n) std::exit(2);
  std::fclose(f);
  return v;
}

// Force<T>::realloc stride (Force.h:63-69): aligned to 256-byte boundaries
static int force_stride(int n) { return ((((n - 1 + 32) * (int)sizeof(float)) - 1) / 256 + 1) * 256 / (int)sizeof(float); }

// AT: the charge-grid accumulation type of the instantiation (float as in
// gpu_recip.cu; int as CudaDomdecRecip.h:12, the production one; long long);
// fft_type: BOX as every reference caller, or COLUMN / SLAB.
template <typename AT>
static bool run(int order, FFTtype fft_type, const char* what, const std::vector<float>& h_xyzq,
                const std::vector<double>& fref, double energy_comp, const double* virial_comp, double tol_e,
                double tol_v, double tol_f) {
  std::printf("CudaPMERecip<%s, float, float2>, fft type %d, order %d\n", what, (int)fft_type, order);
  // Settings for the data, gpu_recip.cu:71-80
  const double boxx = 62.23, boxy = 62.23, boxz = 62.23, kappa = 0.320;
  const int nfftx = 64, nffty = 64, nfftz = 64;
  const int ncoord = (int)(h_xyzq.size() / 4);
  double recip[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  recip[0] = 1.0 / boxx;
  recip[4] = 1.0 / boxy;
  recip[8] = 1.0 / boxz;

  float4* xyzq = nullptr;
  float* force = nullptr;
  const int stride = force_stride(ncoord);
  HIPCHECK(hipMalloc(&xyzq, sizeof(float4) * ncoord));
  HIPCHECK(hipMalloc(&force, sizeof(float) * 3 * stride));
  HIPCHECK(hipMemcpy(xyzq, h_xyzq.data(), sizeof(float4) * ncoord, hipMemcpyHostToDevice));

  bool ok = true;
  {
    CudaEnergyVirial energyVirial;
    CudaPMERecip<AT, float, float2> PMErecip(nfftx, nffty, nfftz, order, fft_type, 1, 0, energyVirial, "recip",
                                             "self");
    // warm up with energy and virial (gpu_recip.cu:118-138)
    energyVirial.clear();
    PMErecip.spread_charge(xyzq, ncoord, recip);
    PMErecip.r2c_fft();
    PMErecip.scalar_sum(recip, kappa, true, true);
    PMErecip.c2r_fft();
    PMErecip.gather_force(xyzq, ncoord, recip, stride, force);
    energyVirial.copyToHost();
    HIPCHECK(hipDeviceSynchronize());
    const double energy = energyVirial.getEnergy("recip");
    double virial[9];
    energyVirial.getVirial(virial);
    double max_diff = std::fabs(energy_comp - energy);
    if (std::isnan(energy) || max_diff > tol_e) {
      std::printf("order %d energy comparison FAILED energy_comp = %.10f energy = %.10f\n", order, energy_comp, energy);
      ok = false;
    } else {
      std::printf("order %d energy comparison OK (tolerance %g max difference %g)\n", order, tol_e, max_diff);
    }
    max_diff = 0.0;
    for (int i = 0; i < 9; i++) max_diff = std::fmax(max_diff, std::fabs(virial_comp[i] - virial[i]));
    if (max_diff > tol_v) {
      std::printf("order %d virial comparison FAILED (max difference %g)\n", order, max_diff);
      ok = false;
    } else {
      std::printf("order %d virial comparison OK (tolerance %g max difference %g)\n", order, tol_v, max_diff);
    }
    // second solve, forces only (gpu_recip.cu:179-291)
    PMErecip.spread_charge(xyzq, ncoord, recip);
    PMErecip.r2c_fft();
    PMErecip.scalar_sum(recip, kappa, false, false);
    PMErecip.c2r_fft();
    PMErecip.gather_force(xyzq, ncoord, recip, stride, force);
    std::vector<float> h_force(3 * (size_t)stride);
    HIPCHECK(hipMemcpy(h_force.data(), force, sizeof(float) * 3 * stride, hipMemcpyDeviceToHost));
    // Force<T>::compare (Force.cu:56-96): max abs component difference
    max_diff = 0.0;
    for (int i = 0; i < ncoord; i++) {
      for (int d = 0; d < 3; d++) {
        const double a = h_force[(size_t)i + (size_t)d * stride];
        if (std::isnan(a)) max_diff = INFINITY;
        max_diff = std::fmax(max_diff, std::fabs(a - fref[3 * (size_t)i + d]));
      }
    }
    if (max_diff > tol_f) {
      std::printf("order %d force comparison FAILED (tolerance %g max difference %g)\n", order, tol_f, max_diff);
      ok = false;
    } else {
      std::printf("order %d force comparison OK (tolerance %g max difference %g)\n", order, tol_f, max_diff);
    }
    // the fixed-point accumulation and the float3 store of the other FT variants
    long long* fl = nullptr;
    HIPCHECK(hipMalloc(&fl, sizeof(long long) * 3 * stride));
    HIPCHECK(hipMemset(fl, 0, sizeof(long long) * 3 * stride));
    PMErecip.gather_force(xyzq, ncoord, recip, stride, fl);
    std::vector<long long> h_fl(3 * (size_t)stride);
    HIPCHECK(hipMemcpy(h_fl.data(), fl, sizeof(long long) * 3 * stride, hipMemcpyDeviceToHost));
    double dfl = 0.0;
    for (int i = 0; i < 3 * stride; i++)
      dfl = std::fmax(dfl, std::fabs((double)h_fl[i] / 1099511627776.0 - (double)h_force[i]));
    std::printf("order %d gather_force<long long> vs <float> max difference %g %s\n", order, dfl,
                dfl < 1e-5 ? "OK" : "FAILED");
    ok = ok && dfl < 1e-5;
    HIPCHECK(hipFree(fl));
  }
  HIPCHECK(hipFree(xyzq));
  HIPCHECK(hipFree(force));
  return ok;
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s xyzq.f32 force_recip_4.f64 force_recip_6.f64\n", argv[0]);
    return 2;
  }
  const std::vector<float> xyzq = load<float>(argv[1]);
  const std::vector<double> f4 = load<double>(argv[2]);
  const std::vector<double> f6 = load<double>(argv[3]);
  // energy/virial constants and tolerances, gpu_recip.cu:81-84, :311-314
  const double v4[9] = {-0.5282341105E+03, 0.5596133994E+01, -0.6254453157E+01, 0.5596133994E+01, -0.5054581254E+03,
                        -0.7652312605E-01, -0.6254453157E+01, -0.7652312605E-01, -0.5575246246E+03};
  const double v6[9] = {-0.5314776507E+03, 0.5670623182E+01, -0.6175456643E+01, 0.5670623182E+01, -0.5086150173E+03,
                        -0.1026251692E+00, -0.6175456643E+01, -0.1026251692E+00, -0.5606459148E+03};
  bool ok = true;
  try {
    ok = run<float>(4, BOX, "float", xyzq, f4, 0.5939730270E+03, v4, 2.0e-3, 4.0e-3, 1.0e-3) && ok;
    ok = run<float>(6, BOX, "float", xyzq, f6, 0.5951144996E+03, v6, 1.3e-3, 2.0e-3, 3.6e-4) && ok;
    ok = run<int>(4, BOX, "int", xyzq, f4, 0.5939730270E+03, v4, 2.0e-3, 4.0e-3, 1.0e-3) && ok;
    ok = run<int>(6, COLUMN, "int", xyzq, f6, 0.5951144996E+03, v6, 1.3e-3, 2.0e-3, 3.6e-4) && ok;
    ok = run<long long>(4, SLAB, "long long", xyzq, f4, 0.5939730270E+03, v4, 2.0e-3, 4.0e-3, 1.0e-3) && ok;
    ok = run<long long>(6, BOX, "long long", xyzq, f6, 0.5951144996E+03, v6, 1.3e-3, 2.0e-3, 3.6e-4) && ok;
  } catch (const pme::Error& e) {
    std::printf("pme error: %s\n", e.what());
    return 1;
  }
  std::printf(ok ? "ALL OK\n" : "FAILED\n");
  return ok ? 0 : 1;
}
