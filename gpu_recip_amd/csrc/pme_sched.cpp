// pme_sched.cpp -- multi-system scheduling core (see pme_sched.hpp).
#include "pme_sched.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace pme {

static const double kPiS = 3.14159265358979323846;

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

pme_status validate_system(const pme_system_desc* s) {
  if (!s) return set_error(PME_ERR_INVALID_ARG, "system: null descriptor");
  if (s->order != 4 && s->order != 6 && s->order != 8) return set_error(PME_ERR_UNSUPPORTED, "system: order %d", s->order);
  const int n[3] = {s->nfftx, s->nffty, s->nfftz};
  for (int a = 0; a < 3; a++)
    if (n[a] < s->order || n[a] > 512 || !smooth235(n[a]))
      return set_error(PME_ERR_UNSUPPORTED, "system: grid edge %d", n[a]);
  const double* r = s->recip;
  if (r[1] != 0.0 || r[2] != 0.0 || r[3] != 0.0 || r[5] != 0.0 || r[6] != 0.0 || r[7] != 0.0)
    return set_error(PME_ERR_UNSUPPORTED, "system: only orthorhombic boxes are supported");
  if (!(r[0] > 0.0 && r[4] > 0.0 && r[8] > 0.0)) return set_error(PME_ERR_INVALID_ARG, "system: recip diagonal");
  if (!(s->kappa > 0.0)) return set_error(PME_ERR_INVALID_ARG, "system: kappa %g", s->kappa);
  if (s->ncoord < 0) return set_error(PME_ERR_INVALID_ARG, "system: ncoord %d", s->ncoord);
  if (s->ncoord > 0 && (!s->d_xyzq || !s->d_force)) return set_error(PME_ERR_INVALID_ARG, "system: null device pointer");
  if (s->force_type < PME_FORCE_FLOAT || s->force_type > PME_FORCE_INT64)
    return set_error(PME_ERR_INVALID_ARG, "system: force_type %d", s->force_type);
  if (s->precision < PME_PREC_FP32 || s->precision > PME_PREC_FP16X3)
    return set_error(PME_ERR_INVALID_ARG, "system: precision %d", s->precision);
  if (s->calc_energy_virial < 0 || s->calc_energy_virial > (PME_CALC_EV | PME_CALC_SELF))
    return set_error(PME_ERR_INVALID_ARG, "system: calc_energy_virial %d", s->calc_energy_virial);
  if (s->force_type != PME_FORCE_FLOAT3 && s->stride < s->ncoord)
    return set_error(PME_ERR_INVALID_ARG, "system: stride %d < ncoord %d", s->stride, s->ncoord);
  return PME_OK;
}

pme_status plan_system(const pme_system_desc& in, int cap_atoms, const Tuning& t, SystemPlan* p) {
  const int n[3] = {in.nfftx, in.nffty, in.nfftz};
  if (!plan_bricks(n, in.order, t.brick_target, &p->bp, t.half_bricks ? kBrickLdsBudget / 2 : kBrickLdsBudget))
    return set_error(PME_ERR_UNSUPPORTED, "system: no brick plan fits the LDS budget");
  p->cp = plan_chunks(std::max(cap_atoms, 1));
  const bool even = (n[0] % 2) == 0;
  if (!radix_plan(even ? n[0] / 2 : n[0], even ? 2 : 1, &p->rp[0]) || !radix_plan(n[1], 1, &p->rp[1]) ||
      !radix_plan(n[2], 1, &p->rp[2]))
    return set_error(PME_ERR_UNSUPPORTED, "system: no radix plan");
  p->prec[0] = PME_PREC_FP32;
  for (int a = 1; a < 3; a++) {
    p->prec[a] = PME_PREC_FP32;
    RadixPlan sp;
    if (in.precision != PME_PREC_FP32 && radix_plan_split(n[a], 1, &sp)) {
      p->rp[a] = sp;
      p->prec[a] = in.precision;
    }
  }
  p->ft = pick_fft_tiles(n[0], n[1], n[2], p->rp[0].n, t.fft_tile,
                         p->prec[2] == PME_PREC_FP32 && four_step_code(n[2]) >= 0);
  p->plane = fft_plane_lds_bytes(n[0], n[1], p->rp[0].n, 1) != 0 ? 1 : 0;
  const int hx = n[0] / 2 + 1;
  if (p->ft.TXz <= 0 || (p->ft.TXz & (p->ft.TXz - 1)) != 0)  // the z pass indexes its tiles with shifts
    return set_error(PME_ERR_UNSUPPORTED, "system: z tile %d is not a power of two", p->ft.TXz);
  p->nzblk = cols_blocks(hx * n[1], p->ft.TXz, 1);
  const size_t G = (size_t)n[0] * n[1] * n[2];
  const size_t Gc = (size_t)spec_pitch(hx, n[1]) * n[2];
  const size_t rs = 8 * (size_t)p->cp.chunk * p->cp.nchunk;
  const size_t sz[7] = {sizeof(float) * G,
                        sizeof(float2) * Gc,
                        sizeof(int) * (size_t)p->cp.nchunk * seg_stride(p->bp.nbins),
                        sizeof(float4) * rs,
                        sizeof(int) * rs,
                        sizeof(double) * 7 * (size_t)p->nzblk,
                        sizeof(double) * (size_t)p->cp.nchunk};
  size_t o = 0;
  for (int i = 0; i < 7; i++) {
    p->off[i] = o;
    o += align256(sz[i]);
  }
  p->total = o;
  return PME_OK;
}

pme_status TableCache::get(int order, int n, const float2** tw_out, const float** pf_out) {
  auto it = tw.find(n);
  if (it == tw.end()) {
    float2* d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(float2) * n);
    if (e != hipSuccess) return set_error(PME_ERR_OOM, "twiddle table: %s", hipGetErrorString(e));
    const std::vector<float2> t = twiddles(n);
    e = hipMemcpy(d, t.data(), sizeof(float2) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return set_error(PME_ERR_DEVICE, "twiddle upload: %s", hipGetErrorString(e));
    bytes += sizeof(float2) * n;
    it = tw.emplace(n, d).first;
  }
  *tw_out = it->second;
  auto jt = pf.find({order, n});
  if (jt == pf.end()) {
    std::vector<double> m(n);
    bsp_mod_host(order, n, m.data());
    std::vector<float> f(n);
    for (int i = 0; i < n; i++) f[i] = (float)m[i];
    float* d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(float) * n);
    if (e != hipSuccess) return set_error(PME_ERR_OOM, "prefac table: %s", hipGetErrorString(e));
    e = hipMemcpy(d, f.data(), sizeof(float) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return set_error(PME_ERR_DEVICE, "prefac upload: %s", hipGetErrorString(e));
    bytes += sizeof(float) * n;
    jt = pf.emplace(std::make_pair(order, n), d).first;
  }
  *pf_out = jt->second;
  return PME_OK;
}

void TableCache::release() {
  for (auto& kv : tw) (void)hipFree(kv.second);
  for (auto& kv : pf) (void)hipFree(kv.second);
  tw.clear();
  pf.clear();
  bytes = 0;
}

void fill_dynamic(SysDesc& d, const pme_system_desc& in, const SystemPlan& p) {
  d.bin.r1 = (float)in.recip[0];
  d.bin.r2 = (float)in.recip[4];
  d.bin.r3 = (float)in.recip[8];
  d.bin.nchunk = std::max(1, (in.ncoord + p.cp.chunk - 1) / p.cp.chunk);
  d.xyzq = (const float4*)in.d_xyzq;
  d.natom = in.ncoord;
  d.force = in.d_force;
  d.force_type = in.force_type;
  d.stride = in.stride;
  d.sa.r1 = d.bin.r1;
  d.sa.r2 = d.bin.r2;
  d.sa.r3 = d.bin.r3;
  const double inv_vol = in.recip[0] * in.recip[4] * in.recip[8];
  d.sa.piv_inv = (float)(inv_vol / kPiS);
  d.sa.fac = (float)(kPiS * kPiS / (in.kappa * in.kappa));
  d.sa.calc_ev = (in.calc_energy_virial & PME_CALC_EV) ? 1 : 0;
  d.calc_self = (in.calc_energy_virial & PME_CALC_SELF) ? 1 : 0;
  d.kcs = in.kappa * kCcelec / std::sqrt(kPiS);
}

pme_status fill_desc(SysDesc& d, const pme_system_desc& in, const SystemPlan& p, void* base, TableCache& tables) {
  unsigned char* b = static_cast<unsigned char*>(base);
  const int n[3] = {in.nfftx, in.nffty, in.nfftz};
  std::memset(&d, 0, sizeof(d));
  d.bin.nx = n[0];
  d.bin.ny = n[1];
  d.bin.nz = n[2];
  d.bin.nbx = p.bp.nb[0];
  d.bin.nby = p.bp.nb[1];
  d.bin.nbz = p.bp.nb[2];
  d.bin.order = in.order;
  d.bin.chunk = p.cp.chunk;
  d.bin.rec_stride = 8 * p.cp.chunk;
  d.bin.seg_off = reinterpret_cast<int*>(b + p.off[2]);
  d.bin.rec_pos = reinterpret_cast<float4*>(b + p.off[3]);
  d.bin.rec_idx = reinterpret_cast<int*>(b + p.off[4]);
  for (int a = 0; a < 3; a++) d.wmax[a] = p.bp.wmax[a];
  d.grid = reinterpret_cast<float*>(b + p.off[0]);
  d.cgrid = reinterpret_cast<float2*>(b + p.off[1]);
  for (int a = 0; a < 3; a++) {
    const float2* tw = nullptr;
    const float* pf = nullptr;
    pme_status st = tables.get(in.order, n[a], &tw, &pf);
    if (st != PME_OK) return st;
    d.ax[a].rp = p.rp[a];
    d.ax[a].tw = tw;
    d.ax[a].ntw = n[a];
    d.ax[a].prec = p.prec[a];
    d.ax[a].f4 = (a == 2 && p.prec[a] == PME_PREC_FP32) ? four_step_code(n[a]) : -1;
    d.ax[a].f4col = (a == 1 && p.prec[a] == PME_PREC_FP32) ? four_step_code_for_plan(p.rp[1]) : -1;
    if (a == 0) d.sa.pf1 = pf;
    if (a == 1) d.sa.pf2 = pf;
    if (a == 2) d.sa.pf3 = pf;
  }
  d.NR = p.ft.NR;
  d.TXy = p.ft.TXy;
  d.TXz = p.ft.TXz;
  d.sa.n1 = n[0];
  d.sa.n2 = n[1];
  d.sa.n3 = n[2];
  d.sa.nf1 = n[0] / 2 + (n[0] % 2);
  d.sa.nf2 = n[1] / 2 + (n[1] % 2);
  d.sa.nf3 = n[2] / 2 + (n[2] % 2);
  d.sa.partial = reinterpret_cast<double*>(b + p.off[5]);
  d.self_partial = reinterpret_cast<double*>(b + p.off[6]);
  d.nzblk = p.nzblk;
  d.plane = p.plane;
  fill_dynamic(d, in, p);
  return PME_OK;
}

void build_items(const SysDesc* const* descs, const SystemPlan* const* plans, const int* orders, int n, int ncu,
                 std::vector<WorkItem>& items, std::vector<Launch>& launches) {
  items.clear();
  launches.clear();
  // A plane workgroup takes up to `group` z planes at once (as many as fit
  // its LDS): each FFT stage and barrier then covers that many planes'
  // butterflies.  The group is the largest that still leaves at least two
  // items per CU.
  int total_planes = 0;
  for (int k = 0; k < n; k++)
    if (descs[k]->plane) total_planes += descs[k]->bin.nz;
  int group = kPlaneGroupMax;
  while (group > 1 && total_planes < 2 * ncu * group) group--;
  // order: the spread/gather order, or the largest plane group of a plane launch
  // split: the FFT launches of systems whose transformed axis uses the
  // split-precision matrix-core stage go to kernels compiled with it, the
  // others to FP32-only kernels
  // half: plane launches of the systems whose plane fits a half-size
  // workgroup (half = 1) or does not (half = 0)
  auto add_stage = [&](int stage, int order, int split, int half) {
    Launch L{stage, order, split, half, (int)items.size(), 0, 0};
    for (int k = 0; k < n; k++) {
      const SysDesc& d = *descs[k];
      const SystemPlan& p = *plans[k];
      if ((stage == BS_SPREAD || stage == BS_GATHER) && orders[k] != order) continue;
      if (stage == BS_FFTY || stage == BS_PLANE || stage == BS_FFTZ) {
        const int a = stage == BS_FFTZ ? 2 : 1;
        if ((d.ax[a].prec != PME_PREC_FP32) != (split != 0)) continue;
      }
      // z: order 1 + c takes the systems with a four-step pair of register
      // class c, order 0 the others (Stockham stages, split precision)
      if (stage == BS_FFTZ && order != (d.ax[2].f4 >= 0 ? 1 + four_step_class(d.ax[2].f4) : 0)) continue;
      int cnt = 0;
      size_t lds = 0;
      const int hx = d.bin.nx / 2 + 1;
      switch (stage) {
        case BS_BIN:
          cnt = d.bin.nchunk;
          lds = bin_lds_bytes(p.bp.nbins);
          break;
        case BS_SPREAD:
          cnt = p.bp.nbins;
          lds = spread_lds_bytes(orders[k], p.bp.wmax, d.bin.nchunk);
          break;
        case BS_GATHER:
          cnt = p.bp.nbins;
          lds = gather_lds_bytes(orders[k], p.bp.wmax, d.bin.nchunk);
          break;
        case BS_FFTX:
          if (d.plane) break;
          cnt = (d.bin.ny * d.bin.nz + d.NR - 1) / d.NR;
          lds = fft_x_lds_bytes(d.ax[0].rp.n, d.NR, d.bin.nx);
          break;
        case BS_FFTY:
          if (d.plane) break;
          cnt = cols_blocks(hx, d.TXy, d.bin.nz);
          lds = fft_cols_lds_bytes(d.bin.ny, d.TXy);
          break;
        case BS_PLANE: {
          if (!d.plane || plane_fits_half(d.bin.nx, d.bin.ny, d.ax[0].rp.n) != (half != 0)) break;
          const int g = plane_group_host(d.bin.nx, d.bin.ny, d.ax[0].rp.n, order, half != 0);
          lds = fft_plane_lds_bytes(d.bin.nx, d.bin.ny, d.ax[0].rp.n, g);
          for (int z = 0; z < d.bin.nz; z += g) items.push_back(WorkItem{k, z});
          L.count += (d.bin.nz + g - 1) / g;
          break;
        }
        case BS_FFTZ:
          cnt = cols_blocks(hx * d.bin.ny, d.TXz, 1);
          lds = fft_cols_lds_bytes(d.bin.nz, d.TXz, d.ax[2].f4 >= 0 ? 1 : 2);
          break;
      }
      for (int i = 0; i < cnt; i++) items.push_back(WorkItem{k, i});
      L.count += cnt;
      L.lds = std::max(L.lds, lds);
    }
    if (L.count > 0) launches.push_back(L);
  };
  add_stage(BS_BIN, 0, 0, 0);
  for (int o : {4, 6, 8}) add_stage(BS_SPREAD, o, 0, 0);
  add_stage(BS_FFTX, 0, 0, 0);
  for (int sp : {0, 1}) add_stage(BS_FFTY, 0, sp, 0);
  for (int sp : {0, 1})
    for (int h : {1, 0}) add_stage(BS_PLANE, group, sp, h);
  for (int c : {1, 2}) add_stage(BS_FFTZ, c, 0, 0);
  for (int sp : {0, 1}) add_stage(BS_FFTZ, 0, sp, 0);
  for (int o : {4, 6, 8}) add_stage(BS_GATHER, o, 0, 0);
}

// Workgroups of a batched launch.  Spread and gather (queue != 0) always run
// one persistent workgroup per CU (two half-size ones with half_bricks)
// taking bricks from the launch's work queue (their LDS brick leaves room for
// no more, and a persistent workgroup saves the start of a new one per brick).  The others
// run one workgroup per item with wg_per_cu == 0, else wg_per_cu persistent
// workgroups per CU striding over the items; the plane kernels (cap > 0)
// cap per CU (one full-size or two half-size workgroups fill a CU).
static int grid_for(const Tuning& t, int ncu, int count, int cap, bool queue) {
  if (queue) return std::max(1, std::min(count, ncu * (t.half_bricks ? 2 : 1)));
  if (t.wg_per_cu <= 0) return count;
  const int per_cu = cap > 0 ? cap : t.wg_per_cu;
  return std::max(1, std::min(count, ncu * per_cu));
}

#define SCHED_HIP(expr)                                                                       \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return set_error(PME_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

pme_status enqueue_stages(const std::vector<Launch>& launches, const SysDesc* d_sys, const WorkItem* d_items,
                          int* d_ctr, const Tuning& t, int ncu, hipStream_t stream, hipEvent_t* tev) {
  if (launches.size() > (size_t)kMaxLaunches) return set_error(PME_ERR_UNSUPPORTED, "batched solve: too many launches");
  // timing slot k: the stages run between events k and k+1
  //   0 bin, 1 spread, 2 x forward and xy-plane forward, 3 y forward, 4 z (+ influence function),
  //   5 y inverse and xy-plane inverse (with the energy/virial reduction), 6 x inverse, 7 gather
  const int seq[8][3] = {{BS_BIN, 0, -1},  {BS_SPREAD, 0, -1}, {BS_FFTX, 0, BS_PLANE}, {BS_FFTY, 0, -1},
                         {BS_FFTZ, 0, -1}, {BS_FFTY, 1, BS_PLANE}, {BS_FFTX, 1, -1},   {BS_GATHER, 0, -1}};
  for (int k = 0; k < 8; k++) {
    if (tev) SCHED_HIP(hipEventRecord(tev[k], stream));
    const bool inv = seq[k][1] != 0;
    for (int pass = 0; pass < 2; pass++) {
      const int stage = pass == 0 ? seq[k][0] : seq[k][2];
      if (stage < 0) continue;
      for (size_t li = 0; li < launches.size(); li++) {
        const Launch& L = launches[li];
        if (L.stage != stage) continue;
        const WorkItem* it = d_items + L.offset;
        int* ctr = d_ctr + 2 * li;  // this launch's work-queue counters (spread, gather)
        const bool queue = stage == BS_SPREAD || stage == BS_GATHER;
        const int grid = grid_for(t, ncu, L.count, stage == BS_PLANE ? (L.half ? 2 : 1) : 0, queue);
        switch (stage) {
          case BS_BIN: SCHED_HIP(launch_batch_bin(d_sys, it, L.count, grid, L.lds, stream)); break;
          case BS_SPREAD:
            SCHED_HIP(launch_batch_spread(L.order, t.half_bricks, d_sys, it, L.count, ctr, grid, L.lds, stream));
            break;
          case BS_FFTX: SCHED_HIP(launch_batch_fft_x(inv ? 1 : 0, d_sys, it, L.count, grid, L.lds, stream)); break;
          case BS_FFTY:
            SCHED_HIP(launch_batch_cols(1, inv ? COL_INV : COL_FWD, inv ? 1 : 0, L.split, d_sys, it, L.count, grid, L.lds,
                                        stream));
            break;
          case BS_PLANE:
            SCHED_HIP(launch_batch_plane(inv ? 1 : 0, inv ? 1 : 0, L.order, L.split, L.half, d_sys, it, L.count, grid, L.lds,
                                         stream));
            break;
          case BS_FFTZ:
            if (L.order > 0)
              SCHED_HIP(launch_batch_zsolve(L.order - 1, d_sys, it, L.count, grid, L.lds, stream));
            else
              SCHED_HIP(launch_batch_cols(2, COL_FWD_SOLVE_INV, 0, L.split, d_sys, it, L.count, grid, L.lds, stream));
            break;
          case BS_GATHER:
            SCHED_HIP(launch_batch_gather(L.order, t.half_bricks, d_sys, it, L.count, ctr, grid, L.lds, stream));
            break;
          default: break;
        }
      }
    }
  }
  if (tev) SCHED_HIP(hipEventRecord(tev[8], stream));
  return PME_OK;
}

double system_cost(const pme_system_desc& in) {
  const double p = in.order;
  return (double)in.ncoord * p * p * p + 4.0 * (double)in.nfftx * in.nffty * in.nfftz;
}

}  // namespace pme
