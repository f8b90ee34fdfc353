// direct_dropin.cpp -- the direct-space Ewald term through the C++ drop-in
// class CudaPMEDirectForce<long long, float> of include/pme_recip.hpp, called
// as the reference's caller calls it (CudaPMEForcefield.cu:189 setup, :298
// set_14_list, :444-460 calc_force + calc_14_force into one long long force
// array; CudaPMEDirectForce.h:147-203).  Inputs and expected values come
// from files written by tests/test_gpu_dropin.py; exits non-zero on failure.
//
//   direct_dropin <dir>
// <dir>/params.f64   boxx boxy boxz kappa roff e14fac e_elec e_excl tol_e tol_f
// <dir>/xyzq.f32     npad x 4 (the tile list's sorted coordinates)
// <dir>/ientry.i32   n_ientry x 4; tile_indj.i32 n_tile; tile_excl.u32 n_tile x 32
// <dir>/in14.i32, ex14.i32   (i, j, ishift) rows
// <dir>/force.f64    npad x 3 expected gradient (sorted order)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
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
static std::vector<T> load(const std::string& path) {
  std::vector<T> v;
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    std::fprintf(stderr, "cannot open %s\n", path.c_str());
    std::exit(2);
  }
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f) / (long)sizeof(T);
  std::fseek(f, 0, SEEK_SET);
  v.resize(n);
  if (n && std::fread(v.data(), sizeof(T), n, f) != (size_t)n) std::exit(2);
  std::fclose(f);
  return v;
}

template <typename T>
static T* upload(const std::vector<T>& h) {
  T* d = nullptr;
  HIPCHECK(hipMalloc(&d, sizeof(T) * (h.empty() ? 1 : h.size())));
  if (!h.empty()) HIPCHECK(hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return d;
}

static bool report(const char* what, double err, double tol) {
  const bool ok = err <= tol;
  std::printf("  %-28s %.3e %s (tolerance %.1e)\n", what, err, ok ? "OK" : "FAILED", tol);
  return ok;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <dir>\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  const std::vector<double> prm = load<double>(dir + "/params.f64");
  if (prm.size() != 10) return 2;
  const std::vector<float> xyzq = load<float>(dir + "/xyzq.f32");
  const std::vector<int> ientry = load<int>(dir + "/ientry.i32");
  const std::vector<int> tile_indj = load<int>(dir + "/tile_indj.i32");
  const std::vector<unsigned int> tile_excl = load<unsigned int>(dir + "/tile_excl.u32");
  std::vector<int> in14 = load<int>(dir + "/in14.i32");
  std::vector<int> ex14 = load<int>(dir + "/ex14.i32");
  const std::vector<double> fref = load<double>(dir + "/force.f64");
  const int npad = (int)(xyzq.size() / 4);
  const int stride = (npad + 63) / 64 * 64;
  std::printf("CudaPMEDirectForce<long long, float>: %d atoms, %zu i entries, %zu j tiles\n", npad,
              ientry.size() / 4, tile_indj.size());

  float4* d_xyzq = reinterpret_cast<float4*>(upload(xyzq));
  pme_ientry* d_ientry = reinterpret_cast<pme_ientry*>(upload(ientry));
  int* d_indj = upload(tile_indj);
  pme_tile_excl* d_excl = reinterpret_cast<pme_tile_excl*>(upload(tile_excl));
  long long* d_force = nullptr;
  HIPCHECK(hipMalloc(&d_force, sizeof(long long) * 3 * stride));
  pme_nlist nlist = {(int)(ientry.size() / 4), d_ientry, (int)tile_indj.size(), d_indj, d_excl};

  bool ok = true;
  try {
    CudaEnergyVirial energyVirial;
    CudaPMEDirectForce<long long, float> direct(energyVirial, "vdw", "elec", "excl");
    direct.setup(prm[0], prm[1], prm[2], prm[3], prm[4], prm[4] - 1.0, prm[5], NONE, EWALD);
    float bx, by, bz;
    direct.get_box_size(bx, by, bz);
    ok = report("get_box_size", std::fabs(bx - prm[0]) + std::fabs(by - prm[1]) + std::fabs(bz - prm[2]), 1e-5) && ok;
    direct.set_14_list((int)(in14.size() / 3), (int)(ex14.size() / 3), reinterpret_cast<xx14list_t*>(in14.data()),
                       reinterpret_cast<xx14list_t*>(ex14.data()));

    // one force/energy evaluation, as CudaPMEForcefield::calc does
    energyVirial.clear();
    HIPCHECK(hipMemset(d_force, 0, sizeof(long long) * 3 * stride));
    direct.calc_force(d_xyzq, nlist, true, true, stride, d_force);
    direct.calc_14_force(d_xyzq, true, true, stride, d_force);
    energyVirial.copyToHost();
    HIPCHECK(hipDeviceSynchronize());
    const double e_elec = energyVirial.getEnergy("elec");
    const double e_excl = energyVirial.getEnergy("excl");
    ok = report("elec energy (relative)", std::fabs(e_elec - prm[6]) / std::fabs(prm[6]), prm[8]) && ok;
    ok = report("excl energy (relative)", std::fabs(e_excl - prm[7]) / std::fabs(prm[7]), prm[8]) && ok;
    std::vector<long long> hf(3 * stride);
    HIPCHECK(hipMemcpy(hf.data(), d_force, sizeof(long long) * hf.size(), hipMemcpyDeviceToHost));
    double num = 0.0, den = 0.0;
    for (int i = 0; i < npad; i++)
      for (int c = 0; c < 3; c++) {
        const double g = (double)hf[i + c * stride] * std::ldexp(1.0, -40);
        const double r = fref[3 * i + c];
        num += (g - r) * (g - r);
        den += r * r;
      }
    ok = report("force (relative RMS)", std::sqrt(num / den), prm[9]) && ok;

    // with the electrostatics off nothing is added (set_calc_elec, :312-314)
    direct.set_calc_elec(false);
    ok = report("get_calc_elec after false", direct.get_calc_elec() ? 1.0 : 0.0, 0.0) && ok;
    energyVirial.clear();
    direct.calc_force(d_xyzq, nlist, true, true, stride, d_force);
    energyVirial.copyToHost();
    HIPCHECK(hipDeviceSynchronize());
    std::vector<long long> hf2(3 * stride);
    HIPCHECK(hipMemcpy(hf2.data(), d_force, sizeof(long long) * hf2.size(), hipMemcpyDeviceToHost));
    double changed = 0.0;
    for (size_t k = 0; k < hf.size(); k++) changed += (hf[k] != hf2[k]) ? 1.0 : 0.0;
    ok = report("elec off: forces unchanged", changed, 0.0) && ok;
    ok = report("elec off: elec energy", std::fabs(energyVirial.getEnergy("elec")), 0.0) && ok;
  } catch (const pme::Error& e) {
    std::printf("pme error: %s\n", e.what());
    return 1;
  }
  HIPCHECK(hipFree(d_force));
  HIPCHECK(hipFree(d_excl));
  HIPCHECK(hipFree(d_indj));
  HIPCHECK(hipFree(d_ientry));
  HIPCHECK(hipFree(d_xyzq));
  std::printf(ok ? "ALL OK\n" : "FAILED\n");
  return ok ? 0 : 1;
}
