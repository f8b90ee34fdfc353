// pme_recip.hpp -- C++ drop-in for the reference's CudaPMERecip /
// CudaEnergyVirial / CudaPMEDirectForce classes (ap-hynninen/gpu_recip
// CudaPMERecip.h:37-205, CudaEnergyVirial.h:20-53, CudaPMEDirectForce.h:80-204),
// header-only over the C ABI of pme_recip.h.
//
// Same class names, member names and argument meaning:
//   CudaEnergyVirial ev;
//   CudaPMERecip<float, float, float2> pme(nfftx, nffty, nfftz, order, BOX, 1, 0, ev, "recip", "self");
//   ev.clear();
//   pme.spread_charge(d_xyzq, ncoord, recip);
//   pme.r2c_fft();
//   pme.scalar_sum(recip, kappa, true, true);
//   pme.c2r_fft();
//   pme.gather_force(d_xyzq, ncoord, recip, stride, d_force);   // FT = float, float3 or long long
//   ev.copyToHost(); hipDeviceSynchronize(); ev.getEnergy("recip"); ev.getVirial(vir);
//
// Error behaviour: the reference prints and calls exit(1); this header throws
// pme::Error (a std::runtime_error carrying the pme_status) instead.
#ifndef PME_RECIP_HPP
#define PME_RECIP_HPP

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <type_traits>

#include "pme_recip.h"

namespace pme {

class Error : public std::runtime_error {
 public:
  Error(pme_status s, const std::string& where)
      : std::runtime_error(where + ": " + pme_status_string(s) + ": " + pme_last_error()), status(s) {}
  pme_status status;
};

inline void check(pme_status s, const char* where) {
  if (s != PME_OK) throw Error(s, where);
}

}  // namespace pme

// FFTtype, CudaPMERecip.h:30
enum FFTtype { COLUMN = PME_FFT_COLUMN, SLAB = PME_FFT_SLAB, BOX = PME_FFT_BOX };

// Layout of the device energy/virial blob, CudaEnergyVirial.h:13-18 (only
// used to compute addresses); the energy terms follow virmat.
struct Virial_t {
  double sforce_dp[27][3];
  long long int sforce_fp[27][3];
  double virmat[9];
};

// CudaEnergyVirial: named FP64 energy terms + 3x3 virial in one device blob
class CudaEnergyVirial {
 public:
  CudaEnergyVirial() { pme::check(pme_ev_create(&h_), "CudaEnergyVirial"); }
  ~CudaEnergyVirial() { pme_ev_destroy(h_); }
  CudaEnergyVirial(const CudaEnergyVirial&) = delete;
  CudaEnergyVirial& operator=(const CudaEnergyVirial&) = delete;

  void insert(const char* name) { pme::check(pme_ev_insert(h_, name), "insert"); }
  void insert(const std::string& name) { insert(name.c_str()); }
  void clear(hipStream_t stream = 0) { pme::check(pme_ev_clear(h_, stream), "clear"); }
  void clearEnergy(hipStream_t stream = 0) { pme::check(pme_ev_clear_energy(h_, stream), "clearEnergy"); }
  void clearVirial(hipStream_t stream = 0) { pme::check(pme_ev_clear_virial(h_, stream), "clearVirial"); }
  void copyToHost(hipStream_t stream = 0) { pme::check(pme_ev_copy_to_host(h_, stream), "copyToHost"); }
  double* getEnergyPointer(const std::string& name) {
    double* p = nullptr;
    pme::check(pme_ev_energy_pointer(h_, name.c_str(), &p), "getEnergyPointer");
    return p;
  }
  Virial_t* getVirialPointer() {
    double* p = nullptr;
    pme::check(pme_ev_virial_pointer(h_, &p), "getVirialPointer");
    return reinterpret_cast<Virial_t*>(p);
  }
  double getEnergy(const char* name) {
    double e = 0.0;
    pme::check(pme_ev_get_energy(h_, name, &e), "getEnergy");
    return e;
  }
  double getEnergy(const std::string& name) { return getEnergy(name.c_str()); }
  void getVirial(double* virmat) { pme::check(pme_ev_get_virial(h_, virmat), "getVirial"); }
  void getSforce(double* sforce) { pme::check(pme_ev_get_sforce(h_, sforce), "getSforce"); }

  pme_ev_t handle() const { return h_; }

 private:
  pme_ev_t h_ = nullptr;
};

// CudaPMERecip<AT, CT, CT2>: AT = accumulation type of the charge grid
// (float, int or long long: every one accumulates as 2^40 fixed point, see
// PME_ACCUM_* in pme_recip.h), CT/CT2 = float/float2.
template <typename AT, typename CT, typename CT2>
class CudaPMERecip {
  static_assert(std::is_same<CT, float>::value && std::is_same<CT2, float2>::value,
                "the MI355X engine computes in FP32 (CT = float, CT2 = float2)");

 public:
  CudaPMERecip(int nfftx, int nffty, int nfftz, int order, FFTtype fft_type, int nnode, int mynode,
               CudaEnergyVirial& energyVirial, const char* nameRecip, const char* nameSelf,
               hipStream_t stream = 0)
      : nfftx_(nfftx), nffty_(nffty), nfftz_(nfftz), order_(order) {
    pme::check(pme_recip_create(nfftx, nffty, nfftz, order, (int)fft_type, nnode, mynode, accum_type(),
                                energyVirial.handle(), nameRecip, nameSelf, stream, &h_),
               "CudaPMERecip");
  }
  ~CudaPMERecip() { pme_recip_destroy(h_); }
  CudaPMERecip(const CudaPMERecip&) = delete;
  CudaPMERecip& operator=(const CudaPMERecip&) = delete;

  void set_stream(hipStream_t stream) { pme::check(pme_recip_set_stream(h_, stream), "set_stream"); }

  void spread_charge(const float4* xyzq, const int ncoord, const double* recip) {
    pme::check(pme_recip_spread_charge(h_, xyzq, ncoord, recip), "spread_charge");
  }
  void r2c_fft() { pme::check(pme_recip_r2c_fft(h_), "r2c_fft"); }
  void scalar_sum(const double* recip, const double kappa, const bool calc_energy, const bool calc_virial) {
    pme::check(pme_recip_scalar_sum(h_, recip, kappa, calc_energy, calc_virial), "scalar_sum");
  }
  void c2r_fft() { pme::check(pme_recip_c2r_fft(h_), "c2r_fft"); }
  void calc_self_energy(const float4* xyzq, const int ncoord, const double kappa) {
    pme::check(pme_recip_calc_self_energy(h_, xyzq, ncoord, kappa), "calc_self_energy");
  }
  // FT = float (strided SoA store), float3 (AoS store), long long (2^40 fixed-point add)
  template <typename FT>
  void gather_force(const float4* xyzq, const int ncoord, const double* recip, const int stride, FT* force) {
    pme::check(pme_recip_gather_force(h_, xyzq, ncoord, recip, stride, force_type<FT>(), force), "gather_force");
  }

  int get_nfftx() { return nfftx_; }
  int get_nffty() { return nffty_; }
  int get_nfftz() { return nfftz_; }
  int get_order() { return order_; }
  void set_order(int order) {
    pme::check(pme_recip_set_order(h_, order), "set_order");
    order_ = order;
  }

  pme_recip_t handle() const { return h_; }

 private:
  static int accum_type() {
    if (std::is_same<AT, float>::value) return PME_ACCUM_FLOAT;
    if (std::is_same<AT, long long>::value || std::is_same<AT, long long int>::value) return PME_ACCUM_INT64;
    return PME_ACCUM_INT32;
  }
  template <typename FT>
  static int force_type() {
    if (std::is_same<FT, float>::value) return PME_FORCE_FLOAT;
    if (std::is_same<FT, float3>::value) return PME_FORCE_FLOAT3;
    static_assert(std::is_same<FT, float>::value || std::is_same<FT, float3>::value ||
                      std::is_same<FT, long long>::value,
                  "gather_force<FT>: FT is float, float3 or long long");
    return PME_FORCE_INT64;
  }

  pme_recip_t h_ = nullptr;
  int nfftx_, nffty_, nfftz_, order_;
};

// Models of the direct-space term, CudaDirectForceTypes.h:62-65 (only the
// electrostatic ones are built)
enum { NONE = PME_VDW_NONE, EWALD = PME_ELEC_EWALD, EWALD_LOOKUP = PME_ELEC_EWALD_LOOKUP };
typedef pme_xx14 xx14list_t;  // Bonded_struct.h:122-124

// CudaPMEDirectForce<AT, CT> (CudaPMEDirectForce.h:80-204): the direct-space
// Ewald term.  The reference instantiates AT = long long, CT = float only.
// calc_force takes the tile list as a pme_nlist of device pointers in place of
// CudaNeighborListBuild<32> (whose get_ientry/get_tile_indj/get_tile_excl
// fill it).
template <typename AT, typename CT>
class CudaPMEDirectForce {
  static_assert(std::is_same<AT, long long>::value && std::is_same<CT, float>::value,
                "CudaPMEDirectForce<long long, float> is the one instantiation");

 public:
  CudaPMEDirectForce(CudaEnergyVirial& energyVirial, const char* nameVdw, const char* nameElec,
                     const char* nameExcl) {
    pme::check(pme_direct_create(energyVirial.handle(), nameVdw, nameElec, nameExcl, &h_), "CudaPMEDirectForce");
  }
  ~CudaPMEDirectForce() { pme_direct_destroy(h_); }
  CudaPMEDirectForce(const CudaPMEDirectForce&) = delete;
  CudaPMEDirectForce& operator=(const CudaPMEDirectForce&) = delete;

  void setup(double boxx, double boxy, double boxz, double kappa, double roff, double ron, double e14fac,
             int vdw_model, int elec_model) {
    pme::check(pme_direct_setup(h_, boxx, boxy, boxz, kappa, roff, ron, e14fac, vdw_model, elec_model), "setup");
    calc_vdw_ = calc_elec_ = true;  // setup turns both terms on, CudaPMEDirectForce.cu:255-256
  }
  void get_setup(float& boxx, float& boxy, float& boxz, float& kappa, float& roff, float& ron, float& e14fac,
                 int& vdw_model, int& elec_model) {
    pme::check(pme_direct_get_setup(h_, &boxx, &boxy, &boxz, &kappa, &roff, &ron, &e14fac, &vdw_model, &elec_model),
               "get_setup");
  }
  void get_box_size(CT& boxx, CT& boxy, CT& boxz) {
    float kappa, roff, ron, e14fac;
    int vdw, elec;
    get_setup(boxx, boxy, boxz, kappa, roff, ron, e14fac, vdw, elec);
  }
  void set_box_size(const CT boxx, const CT boxy, const CT boxz) {
    pme::check(pme_direct_set_box_size(h_, boxx, boxy, boxz), "set_box_size");
  }
  void set_calc_vdw(const bool calc_vdw) {
    pme::check(pme_direct_set_calc_vdw(h_, calc_vdw), "set_calc_vdw");
    calc_vdw_ = calc_vdw;
  }
  void set_calc_elec(const bool calc_elec) {
    pme::check(pme_direct_set_calc_elec(h_, calc_elec), "set_calc_elec");
    calc_elec_ = calc_elec;
  }
  bool get_calc_vdw() { return calc_vdw_; }
  bool get_calc_elec() { return calc_elec_; }

  void set_14_list(int nin14list, int nex14list, xx14list_t* h_in14list, xx14list_t* h_ex14list,
                   hipStream_t stream = 0) {
    pme::check(pme_direct_set_14_list(h_, nin14list, nex14list, h_in14list, h_ex14list, stream), "set_14_list");
  }
  void calc_14_force(const float4* xyzq, const bool calc_energy, const bool calc_virial, const int stride, AT* force,
                     hipStream_t stream = 0) {
    pme::check(pme_direct_calc_14_force(h_, xyzq, calc_energy, calc_virial, stride, force, stream), "calc_14_force");
  }
  void calc_force(const float4* xyzq, const pme_nlist& nlist, const bool calc_energy, const bool calc_virial,
                  const int stride, AT* force, hipStream_t stream = 0) {
    pme::check(pme_direct_calc_force(h_, xyzq, &nlist, calc_energy, calc_virial, stride, force, stream),
               "calc_force");
  }

  pme_direct_t handle() const { return h_; }

 private:
  pme_direct_t h_ = nullptr;
  bool calc_vdw_ = true, calc_elec_ = true;
};

#endif  // PME_RECIP_HPP
