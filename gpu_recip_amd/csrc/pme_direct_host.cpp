// pme_direct_host.cpp -- C ABI of the direct-space Ewald term
// (CudaPMEDirectForce<long long, float>, CudaPMEDirectForce.h:82-204): setup,
// the 1-4 lists and the two force calls; the kernels are in pme_direct.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/pme_recip.h"
#include "pme_internal.hpp"
#include "pme_plan.hpp"

using namespace pme;

#define fail pme::set_error

#define HIP_TRY(expr)                                                                                     \
  do {                                                                                                    \
    hipError_t e_ = (expr);                                                                               \
    if (e_ != hipSuccess)                                                                                 \
      return fail(e_ == hipErrorOutOfMemory ? PME_ERR_OOM : PME_ERR_DEVICE, "%s: %s (%s:%d)", #expr,      \
                  hipGetErrorString(e_), __FILE__, __LINE__);                                             \
  } while (0)

#define PME_TRY(expr)             \
  do {                            \
    pme_status s_ = (expr);       \
    if (s_ != PME_OK) return s_;  \
  } while (0)

struct pme_direct_s {
  pme_ev_t ev = nullptr;
  double* e_vdw = nullptr;     // device energy slots in the energy/virial buffer
  double* e_elec = nullptr;
  double* e_excl = nullptr;
  double* sforce_dp = nullptr;  // Virial_t::sforce_dp[27][3] of the buffer
  DirectParams d{};
  double roff = 0.0, ron = 0.0;
  int vdw_model = PME_VDW_NONE, elec_model = PME_ELEC_NONE;
  bool have_setup = false;
  bool calc_elec = true, calc_vdw = true;
  int nin14 = 0, nex14 = 0;
  pme_xx14* in14 = nullptr;     // device lists
  pme_xx14* ex14 = nullptr;
  int in14_cap = 0, ex14_cap = 0;
  double* partial = nullptr;    // [partial_cap * kDirPartial] per-workgroup partials
  int partial_cap = 0;
  int* tile_entry = nullptr;    // [tile_cap] entry of every j tile of the list
  int tile_cap = 0;
  int ncu = 0;                  // compute units of the device
};

extern "C" pme_status pme_direct_create(pme_ev_t ev, const char* name_vdw, const char* name_elec,
                                        const char* name_excl, pme_direct_t* out) {
  if (!out || !ev || !name_vdw || !name_elec || !name_excl)
    return fail(PME_ERR_INVALID_ARG, "pme_direct_create: null argument");
  *out = nullptr;
  // CudaPMEDirectForce.cu:62-79: the three energy terms
  PME_TRY(pme_ev_insert(ev, name_vdw));
  PME_TRY(pme_ev_insert(ev, name_elec));
  PME_TRY(pme_ev_insert(ev, name_excl));
  pme_direct_t h = new pme_direct_s();
  h->ev = ev;
  double* blob = nullptr;
  pme_status st = pme_ev_energy_pointer(ev, name_vdw, &h->e_vdw);
  if (st == PME_OK) st = pme_ev_energy_pointer(ev, name_elec, &h->e_elec);
  if (st == PME_OK) st = pme_ev_energy_pointer(ev, name_excl, &h->e_excl);
  if (st == PME_OK) st = pme_ev_virial_pointer(ev, &blob);
  if (st != PME_OK) {
    delete h;
    return st;
  }
  h->sforce_dp = blob;  // sforce_dp[27][3] opens the Virial_t blob
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&h->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || h->ncu <= 0) {
    delete h;
    return fail(PME_ERR_DEVICE, "pme_direct_create: cannot query the device");
  }
  *out = h;
  return PME_OK;
}

extern "C" pme_status pme_direct_destroy(pme_direct_t h) {
  if (!h) return PME_OK;
  (void)hipDeviceSynchronize();
  (void)hipFree(h->in14);
  (void)hipFree(h->ex14);
  (void)hipFree(h->partial);
  (void)hipFree(h->tile_entry);
  delete h;
  return PME_OK;
}

extern "C" pme_status pme_direct_setup(pme_direct_t h, double boxx, double boxy, double boxz, double kappa,
                                       double roff, double ron, double e14fac, int vdw_model, int elec_model) {
  if (!h) return fail(PME_ERR_INVALID_ARG, "pme_direct_setup: null handle");
  if (vdw_model != PME_VDW_NONE)
    return fail(PME_ERR_UNSUPPORTED, "pme_direct_setup: vdw_model %d (only the electrostatic Ewald term is built)",
                vdw_model);
  if (elec_model != PME_ELEC_NONE && elec_model != PME_ELEC_EWALD && elec_model != PME_ELEC_EWALD_LOOKUP)
    return fail(PME_ERR_UNSUPPORTED, "pme_direct_setup: elec_model %d (NONE, EWALD or EWALD_LOOKUP)", elec_model);
  if (!(boxx > 0.0 && boxy > 0.0 && boxz > 0.0) || !(roff > 0.0) || !(kappa > 0.0))
    return fail(PME_ERR_INVALID_ARG, "pme_direct_setup: box, roff and kappa must be positive");
  // DirectSettings_t fields the Ewald term reads (CudaPMEDirectForce.cu:151-257)
  h->d.boxx = (float)boxx;
  h->d.boxy = (float)boxy;
  h->d.boxz = (float)boxz;
  h->d.kappa = (float)kappa;
  h->d.kappa2 = (float)(kappa * kappa);
  h->d.roff2 = (float)(roff * roff);
  h->d.e14fac = (float)e14fac;
  h->roff = roff;
  h->ron = ron;
  h->vdw_model = vdw_model;
  h->elec_model = elec_model;
  h->calc_vdw = h->calc_elec = true;
  h->have_setup = true;
  return PME_OK;
}

extern "C" pme_status pme_direct_get_setup(pme_direct_t h, float* boxx, float* boxy, float* boxz, float* kappa,
                                           float* roff, float* ron, float* e14fac, int* vdw_model, int* elec_model) {
  if (!h) return fail(PME_ERR_INVALID_ARG, "pme_direct_get_setup: null handle");
  if (boxx) *boxx = h->d.boxx;
  if (boxy) *boxy = h->d.boxy;
  if (boxz) *boxz = h->d.boxz;
  if (kappa) *kappa = h->d.kappa;
  if (roff) *roff = (float)h->roff;
  if (ron) *ron = (float)h->ron;
  if (e14fac) *e14fac = h->d.e14fac;
  if (vdw_model) *vdw_model = h->vdw_model;
  if (elec_model) *elec_model = h->elec_model;
  return PME_OK;
}

extern "C" pme_status pme_direct_set_box_size(pme_direct_t h, float boxx, float boxy, float boxz) {
  if (!h) return fail(PME_ERR_INVALID_ARG, "pme_direct_set_box_size: null handle");
  if (!(boxx > 0.f && boxy > 0.f && boxz > 0.f)) return fail(PME_ERR_INVALID_ARG, "pme_direct_set_box_size: box");
  h->d.boxx = boxx;
  h->d.boxy = boxy;
  h->d.boxz = boxz;
  return PME_OK;
}

extern "C" pme_status pme_direct_set_calc_elec(pme_direct_t h, int calc_elec) {
  if (!h) return fail(PME_ERR_INVALID_ARG, "pme_direct_set_calc_elec: null handle");
  h->calc_elec = calc_elec != 0;
  return PME_OK;
}

extern "C" pme_status pme_direct_set_calc_vdw(pme_direct_t h, int calc_vdw) {
  if (!h) return fail(PME_ERR_INVALID_ARG, "pme_direct_set_calc_vdw: null handle");
  h->calc_vdw = calc_vdw != 0;
  return PME_OK;
}

static pme_status upload_list(const pme_xx14* src, int n, pme_xx14** dst, int* cap, hipStream_t s) {
  if (n <= 0) return PME_OK;
  if (n > *cap) {
    HIP_TRY(hipStreamSynchronize(s));
    (void)hipFree(*dst);
    *dst = nullptr;
    *cap = 0;
    HIP_TRY(hipMalloc(dst, sizeof(pme_xx14) * (size_t)n));
    *cap = n;
  }
  HIP_TRY(hipMemcpyAsync(*dst, src, sizeof(pme_xx14) * (size_t)n, hipMemcpyHostToDevice, s));
  return PME_OK;
}

extern "C" pme_status pme_direct_set_14_list(pme_direct_t h, int nin14, int nex14, const pme_xx14* h_in14,
                                             const pme_xx14* h_ex14, pme_stream_t stream) {
  if (!h) return fail(PME_ERR_INVALID_ARG, "pme_direct_set_14_list: null handle");
  if (nin14 < 0 || nex14 < 0 || (nin14 > 0 && !h_in14) || (nex14 > 0 && !h_ex14))
    return fail(PME_ERR_INVALID_ARG, "pme_direct_set_14_list: bad list");
  for (int k = 0; k < nin14; k++)
    if (h_in14[k].ishift < 0 || h_in14[k].ishift > 26 || h_in14[k].i < 0 || h_in14[k].j < 0)
      return fail(PME_ERR_INVALID_ARG, "pme_direct_set_14_list: in14 entry %d", k);
  for (int k = 0; k < nex14; k++)
    if (h_ex14[k].ishift < 0 || h_ex14[k].ishift > 26 || h_ex14[k].i < 0 || h_ex14[k].j < 0)
      return fail(PME_ERR_INVALID_ARG, "pme_direct_set_14_list: ex14 entry %d", k);
  hipStream_t s = (hipStream_t)stream;
  PME_TRY(upload_list(h_in14, nin14, &h->in14, &h->in14_cap, s));
  PME_TRY(upload_list(h_ex14, nex14, &h->ex14, &h->ex14_cap, s));
  // the host arrays may be freed on return (copy_HtoD on the stream, CudaPMEDirectForce.cu:586-591)
  HIP_TRY(hipStreamSynchronize(s));
  h->nin14 = nin14;
  h->nex14 = nex14;
  return PME_OK;
}

static pme_status ensure_partial(pme_direct_t h, int nblk, hipStream_t s) {
  if (nblk <= h->partial_cap) return PME_OK;
  HIP_TRY(hipStreamSynchronize(s));
  (void)hipFree(h->partial);
  h->partial = nullptr;
  h->partial_cap = 0;
  HIP_TRY(hipMalloc(&h->partial, sizeof(double) * kDirPartial * (size_t)nblk));
  h->partial_cap = nblk;
  return PME_OK;
}

static bool elec_on(pme_direct_t h) { return h->calc_elec && h->elec_model != PME_ELEC_NONE; }

extern "C" pme_status pme_direct_calc_14_force(pme_direct_t h, const void* d_xyzq, int calc_energy, int calc_virial,
                                               int stride, long long* d_force, pme_stream_t stream) {
  if (!h) return fail(PME_ERR_INVALID_ARG, "pme_direct_calc_14_force: null handle");
  if (!h->have_setup) return fail(PME_ERR_STATE, "pme_direct_calc_14_force: call setup first");
  // only the Ewald models compute the 1-4 terms here (vdW is not built)
  if (!elec_on(h) || h->nin14 + h->nex14 == 0) return PME_OK;
  if (!d_xyzq || !d_force) return fail(PME_ERR_INVALID_ARG, "pme_direct_calc_14_force: null device pointer");
  if (stride < 0) return fail(PME_ERR_INVALID_ARG, "pme_direct_calc_14_force: stride %d", stride);
  hipStream_t s = (hipStream_t)stream;
  const bool ce = calc_energy != 0, cv = calc_virial != 0;
  const int nblk = direct_14_blocks(h->nin14 + h->nex14);
  if (ce || cv) PME_TRY(ensure_partial(h, nblk, s));
  HIP_TRY(launch_direct_14(h->d, h->nin14, h->nex14, h->in14, h->ex14, (const float4*)d_xyzq, stride,
                           reinterpret_cast<unsigned long long*>(d_force), ce, cv, h->partial, s));
  HIP_TRY(launch_direct_reduce(h->partial, nblk, ce, cv, h->e_elec, h->e_excl, h->sforce_dp, s));
  return PME_OK;
}

extern "C" pme_status pme_direct_calc_force(pme_direct_t h, const void* d_xyzq, const pme_nlist* nlist,
                                            int calc_energy, int calc_virial, int stride, long long* d_force,
                                            pme_stream_t stream) {
  if (!h || !nlist) return fail(PME_ERR_INVALID_ARG, "pme_direct_calc_force: null argument");
  if (!h->have_setup) return fail(PME_ERR_STATE, "pme_direct_calc_force: call setup first");
  if (nlist->n_ientry < 0 || nlist->n_tile < 0)
    return fail(PME_ERR_INVALID_ARG, "pme_direct_calc_force: negative list size");
  if (nlist->n_ientry == 0 || !elec_on(h)) return PME_OK;  // CudaPMEDirectForce.cu:741-744
  if (!nlist->ientry || !nlist->tile_indj || !nlist->tile_excl || !d_xyzq || !d_force)
    return fail(PME_ERR_INVALID_ARG, "pme_direct_calc_force: null device pointer");
  if (stride < 0) return fail(PME_ERR_INVALID_ARG, "pme_direct_calc_force: stride %d", stride);
  if (nlist->n_tile == 0) return PME_OK;
  hipStream_t s = (hipStream_t)stream;
  const bool ce = calc_energy != 0, cv = calc_virial != 0;
  const int nwave = direct_force_waves(nlist->n_tile, h->ncu);
  const int nblk = direct_force_blocks(nwave);
  if (ce || cv) PME_TRY(ensure_partial(h, nblk, s));
  if (nlist->n_tile > h->tile_cap) {
    HIP_TRY(hipStreamSynchronize(s));
    (void)hipFree(h->tile_entry);
    h->tile_entry = nullptr;
    h->tile_cap = 0;
    HIP_TRY(hipMalloc(&h->tile_entry, sizeof(int) * (size_t)nlist->n_tile));
    h->tile_cap = nlist->n_tile;
  }
  HIP_TRY(launch_direct_force(h->d, nlist->n_ientry, nlist->n_tile, nwave, nlist->ientry, h->tile_entry,
                              nlist->tile_indj, nlist->tile_excl, (const float4*)d_xyzq, stride,
                              reinterpret_cast<unsigned long long*>(d_force), ce, cv, h->partial, s));
  HIP_TRY(launch_direct_reduce(h->partial, nblk, ce, cv, h->e_elec, h->e_excl, h->sforce_dp, s));
  return PME_OK;
}
