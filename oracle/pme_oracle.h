/*
 * pme_oracle.h -- FP64 host restatement of the PME reciprocal-space solve of
 * CudaPMERecip (ap-hynninen/gpu_recip).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  Nothing in gpu_recip_amd/ links or calls it.
 *
 * Every function cites the reference file:line it restates.  Conventions that
 * are kept exactly (the reference's, not textbook SPME):
 *   - grid index: w = x*r + 2, fr = n*(w - (floor(w+0.5) - 0.5)), i = (int)fr
 *     (CudaPMERecip.cu:512-529); in ORC_INDEX_F32 mode the same formula runs in
 *     FP32 with the x*r+2 term fused (fmaf), as the reference's -fmad=true build
 *     does (Makefile:174);
 *   - stencil: grid point (i+j) mod n receives theta[j] = M_p(w + p-1-j)
 *     (CudaPMERecip.cu:549-577, 120-208);
 *   - forward FFT e^{-2 pi i}, inverse e^{+2 pi i}, both unnormalised;
 *   - k = 0 dropped, energy weight 2 for every k1 >= 1 (reference) or the
 *     exact weight 1 on the Nyquist plane (ORC_NYQUIST_EXACT);
 *   - gather returns the gradient +dE/dr (CudaPMERecip.cu:1238-1241).
 */
#ifndef PME_ORACLE_H
#define PME_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_INDEX_F64 = 0, ORC_INDEX_F32 = 1 };
enum { ORC_NYQUIST_REFERENCE = 0, ORC_NYQUIST_EXACT = 1 };

#define ORC_CCELEC 332.0716

typedef struct orc_params {
  int nx, ny, nz;       /* grid */
  int order;            /* B-spline order p (any p >= 3) */
  double kappa;         /* Ewald splitting parameter, 1/A */
  double recip[9];      /* row-major reciprocal box: fractional coordinate a = row a . (x, y, z);
                           off-diagonal entries make the box triclinic */
  int index_mode;       /* ORC_INDEX_F64 or ORC_INDEX_F32 */
  int nyquist_mode;     /* ORC_NYQUIST_REFERENCE or ORC_NYQUIST_EXACT */
  int nthreads;         /* OpenMP threads (<= 0: library default) */
} orc_params;

/* a1: reference FP32 index mapping (CudaPMERecip.cu:512-529), idx/frac are 3 per atom */
void orc_index_f32(int natom, const float* xyzq, double r1, double r2, double r3,
                   int nx, int ny, int nz, int* idx, float* frac);
/* a1 in FP64 (same formula, double arithmetic) */
void orc_index_f64(int natom, const float* xyzq, double r1, double r2, double r3,
                   int nx, int ny, int nz, int* idx, double* frac);

/* a2: theta[j] = M_p(w+p-1-j), dtheta[j] = d theta[j]/dw  (CudaPMERecip.cu:2507-2537) */
void orc_bspline(int order, double w, double* theta, double* dtheta);

/* a9: bsp_mod[k] = 1/|sum_{j<=n-2} bsp_arr[j] e^{2 pi i jk/n}|^2 (CudaPMERecip.cu:2481-2579) */
void orc_bsp_mod(int order, int n, double* bsp_mod);

/* a3: charge grid Q, packed x-fastest nx*ny*nz (CudaPMERecip.cu:479-580) */
void orc_spread(const orc_params* p, int natom, const float* xyzq, double* grid);

/* a4/a6: FFTs. cplx is interleaved (re,im), half spectrum (nx/2+1)*ny*nz, x fastest. */
void orc_fft1d(int n, double* data, int sign);   /* in place, unnormalised, sign=-1 forward */
void orc_fft3d_r2c(int nx, int ny, int nz, const double* real, double* cplx, int nthreads);
void orc_fft3d_c2r(int nx, int ny, int nz, const double* cplx, double* real, int nthreads);

/* a5: influence function in place; energy (kcal/mol) and virial[9] are ASSIGNED
 * (CudaPMERecip.cu:594-901, 1929-2047) */
void orc_scalar_sum(const orc_params* p, double* cplx, double* energy, double* virial9);

/* a7: gradient +dE/dr, AoS natom*3 (CudaPMERecip.cu:1031-1245) */
void orc_gather(const orc_params* p, int natom, const float* xyzq, const double* phi,
                double* force);

/* a8: -kappa*ccelec/sqrt(pi)*sum q^2 (CudaPMERecip.cu:1252-1275, 2053-2061) */
double orc_self_energy(int natom, const float* xyzq, double kappa);

/* Whole solve: spread -> r2c -> scalar_sum -> c2r -> gather.
 * Returns 0 on success. grid_out (nx*ny*nz) and phi_out may be NULL. */
int orc_solve(const orc_params* p, int natom, const float* xyzq, double* force,
              double* energy, double* virial9, double* grid_out, double* phi_out);

/* Independent check: brute-force Ewald reciprocal sum over |m_i| <= kmax[i]
 * for an orthorhombic box (no grid, no splines).  force may be NULL. */
void orc_ewald_recip(int natom, const float* xyzq, const double box[3], double kappa,
                     const int kmax[3], double* energy, double* force, double* virial9,
                     int nthreads);

/* The same for any (triclinic) recip: reciprocal vectors m1 R[0][:] + m2 R[1][:] + m3 R[2][:],
 * volume 1/det(recip). */
void orc_ewald_recip9(int natom, const float* xyzq, const double recip9[9], double kappa,
                      const int kmax[3], double* energy, double* force, double* virial9,
                      int nthreads);

/* f3: direct-space Ewald, FP64 brute force over every pair i < j and every
 * periodic image (27 shifts of an orthorhombic box) closer than roff, except
 * the nexcl pairs excl[2k], excl[2k+1] (skipped in every image):
 * sum ccelec qi qj erfc(kappa r)/r (CudaDirectForceKernels.cu:352-390,
 * CudaDirectForce_util.h:184-251).  Gradient +dE/dr ADDED to force (AoS
 * natom*3, may be NULL); pair virial sum_pairs r_a dE/dr_b ADDED to virial9
 * (may be NULL).  Returns the energy. */
double orc_direct_ewald(int natom, const float* xyzq, const double box[3], double kappa, double roff,
                        int nexcl, const int* excl, double* force, double* virial9, int nthreads);
/* 1-4 lists: entries list[3k..3k+2] = (i, j, ishift); the shift of ishift
 * (gpu_utils.h:312-320) is added to xi - xj; pair energy
 * ccelec qi qj (erfc(kappa r) + fac - 1)/r with fac = e14fac for the 1-4
 * interactions and 0 for the exclusions (CudaDirectForceKernels.cu:506-660).
 * Gradient and pair virial ADDED as above.  Returns the energy. */
double orc_pair14(int natom, const float* xyzq, const double box[3], double kappa, double fac, int npair,
                  const int* list, double* force, double* virial9);
/* The direct-space sum over a 32-atom tile list, the traversal of
 * CudaDirectForce_util.h:184-251: for every entry ientry[4e..4e+3] =
 * (iatomStart, ish, tileStart, tileEnd), every j tile t in [tileStart,
 * tileEnd] starting at atom tile_indj[t], every j = lane l of the tile and
 * every bit t' of tile_excl[32 t + l] that is clear, the pair
 * (i = iatomStart + (l + t') mod 32, j) with the shift of ish added to xi,
 * when closer than roff: ccelec qi qj erfc(kappa r)/r.  Gradient ADDED to
 * force (npad*3, may be NULL), pair virial to virial9 (may be NULL); sforce
 * (27*3, may be NULL) gets, per shift other than 13, the sum of the i-side
 * gradients of its entries (the shift forces of CudaDirectForce_util.h:321-347).
 * Returns the energy. */
double orc_direct_tiles(int npad, const float* xyzq, const double box[3], double kappa, double roff, int n_ientry,
                        const int* ientry, const int* tile_indj, const unsigned* tile_excl, double* force,
                        double* virial9, double* sforce, int nthreads);

#ifdef __cplusplus
}
#endif
#endif /* PME_ORACLE_H */
