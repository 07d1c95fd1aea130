/*
 * CPU oracle for the MG-PICOLA particle-mesh COLA step and 2LPT initial conditions.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library; the product path
 * (libmgpu / libmgpicola) never links or calls it.
 *
 * This is an FP64 restatement of the reference algorithm, function by function,
 * with explicit arguments in place of the reference's global state.  Meshes use
 * the FFTW in-place layout the reference uses:
 *   real    r[(ix*N + iy)*2*(N/2+1) + iz]          (src/auxPM.c:335)
 *   complex c[(ix*N + iy)*(N/2+1) + iz]  (re,im)   (src/auxPM.c:477)
 * with one extra x-slab appended for the ghost plane (src/2LPT.c:97-99).
 * Transforms are unnormalised in both directions, as FFTW's are.
 *
 * Particles are SoA: pos[0..2][np] (float), as the reference keeps float Pos
 * in MEMORY_MODE (src/vars.h:293-353).
 */
#ifndef MGPICOLA_ORACLE_H
#define MGPICOLA_ORACLE_H

/* Planted mistakes, for the tests that show each pin catches a plausible slip
 * (tests/test_oracle_planted.py builds one library per number): 0 = none (the
 * default build); otherwise one line of the restatement is wrong on purpose.
 *   1 force Nyquist index -N/2 instead of +N/2      7 IC: the skipped k = 0 mode draws nothing
 *   2 2LPT factor +3/7 instead of -3/7              8 Kick: COLA term with + sign
 *   3 P(k): kz > 0 modes weighted 1 instead of 2    9 Drift: 2LPT term growth from order 1
 *   4 deposit: ghost slab not folded onto x = 0    10 LPT field: growth index not mirrored
 *   5 f(R) mass term without a^2                   11 seed table: two shell loops swapped
 *   6 IC: delta ~ P instead of sqrt(P)                                                    */
#ifndef OR_PLANT
#define OR_PLANT 0
#endif

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

int or_num_threads(void); /* threads of the OpenMP loops (OMP_NUM_THREADS) */

/* model ids (compile-time MODEL= switch of the reference Makefile:33-110) */
enum { OR_LCDM = 0, OR_FOFR = 1, OR_DGP = 2, OR_BRANSDICKE = 3, OR_MBETA = 4 };
/* smoothing filters for density screening (src/mg.h:318-338) */
enum { OR_FILTER_NONE = 0, OR_FILTER_GAUSSIAN = 1, OR_FILTER_TOPHAT = 2, OR_FILTER_SHARPK = 3 };

typedef struct {
  int nmesh;            /* Nmesh */
  int nsample;          /* Nsample (particles per dimension) */
  double box;           /* Box [Mpc/h] */
  double omega;         /* Omega_m */
  int model;            /* OR_* */
  int mg_active;        /* modified_gravity_active */
  int include_screening;
  int filter;           /* OR_FILTER_* (DGP smoothing) */
} or_cfg;

/* Scalars of the model hooks at a = aexp_global (src/user_defined_functions.h:413-773). */
typedef struct {
  double a;
  double coupling;      /* coupling_function(a) */
  double mass2;         /* mass2_of_a(a)  = m^2(a)/H0^2 */
  double phi_crit;      /* f(R): 1.5 fR0 ((Om+4(1-Om))/(Om a^-3+4(1-Om)))^(n+1); mbeta: phi_of_a(a) */
  double dgp_fac;       /* 8/9 Om (rcH0/beta_DGP(a))^2 */
  double rsmooth;       /* Rsmooth [Mpc/h] */
  double geff;          /* GeffoverG(a,0) (JBD) */
} or_mg;

size_t or_real_len(int n);      /* doubles in one padded real mesh incl. ghost slab */

/* ---- FFT (own mixed radix; FFTW r2c/c2r semantics) ---- */
void or_fft_c2c_1d(int n, double *data, int sign);               /* in place, unnormalised */
void or_fft_r2c_3d(int n, double *data);                         /* in place padded, forward e^{-i} */
void or_fft_c2r_3d(int n, double *data);                         /* in place padded, backward e^{+i} */

/* ---- PM step ---- */
void or_pm_deposit(const or_cfg *c, size_t np, const float *const pos[3], double *density);
void or_pofk(const or_cfg *c, const double *dk, int nbins, int bintype, double kmin, double kmax,
             int subtract_shotnoise, double *pofk, double *kmean, double *nmodes);
void or_adjust_pofk(int nmesh, int *nbins, int *bintype, int *subtract, double *kmin, double *kmax);
/* PtoMesh_RSD (axis 1 = y, 2 = z) and bin_up_RSD_power_spectrum */
void or_pm_deposit_rsd(const or_cfg *c, size_t np, const float *const pos[3], const float *const vel[3],
                       const float *const l0[3], const float *const l1[3], int axis, double vnorm,
                       int scale_dependent, int use_cola, double s1, double s2, double *density);
void or_rsd_multipoles(const or_cfg *c, const double *dk, int nbins, int bintype, double kmin, double kmax,
                       int subtract_shotnoise, double *P0, double *P2, double *P4, double *k, double *n);
/* ComputeFifthForce: dk = delta(k) (P3D), dx = copy of delta(x) (mgarray_two before the r2c).
 * phik receives P3D_mgarray_two (to be added to P3D in Forces); for JBD dk is scaled in place. */
int  or_mg_fifth_force(const or_cfg *c, const or_mg *m, double *dk, const double *dx, double *phik);
void or_pm_forces(const or_cfg *c, double *dk, const double *phik, double *f0, double *f1, double *f2);
void or_pm_gather(const or_cfg *c, size_t np, const float *const pos[3],
                  const double *f0, const double *f1, const double *f2, float *const disp[3], double sumD[3]);
/* Kick / Drift.  d1,d2 are the particle LPT arrays (P.D, P.D2 for Kick; P.D, P.D2
 * or P.dDdy, P.dD2dy for Drift) or NULL.  Non-scale-dependent (sd = 0): the COLA
 * term is d1*s1 + d2*s2 with s = (ddDddy, ddD2ddy) or (deltaD, deltaD2);
 * scale-dependent (sd = 1): the stored fields already hold those products. */
void or_kick(size_t np, float *const vel[3], float *const disp[3], const float *const d1[3],
             const float *const d2[3], double s1, double s2, int sd, double omega, double dda, double A,
             int use_cola, const double sumD[3], double sumV[3]);
void or_drift(size_t np, float *const pos[3], const float *const vel[3], const float *const d1[3],
              const float *const d2[3], double s1, double s2, int sd, double box, double dyyy, int use_cola,
              const double sumV[3]);

/* ---- RNG: GSL ranlxd1 (Luescher RANLUX, double precision, luxury p=202) ---- */
typedef struct { double x[12]; double carry; unsigned ir, jr, ir_old, pr; } or_ranlxd;
void or_ranlxd_seed(or_ranlxd *s, unsigned long seed);
double or_ranlxd_uniform(or_ranlxd *s);
unsigned long or_ranlxd_get(or_ranlxd *s);
void or_seedtable(int n, unsigned long seed, unsigned int *table);

/* ---- 2LPT initial conditions (src/2LPT.c:185-1520) ----
 * pk_n2[n2] = PowerSpec(2pi/Box sqrt(n2)) [times the MG rescaling], n2 = 0..3*(N/2)^2.
 * Non-SD: za/lpt receive the lattice displacements (float, [3][Np^3]).
 * SD:     cdelta1/cdelta2 receive the stored delta(k) and -source(k) (complex padded). */
void or_ic_delta(const or_cfg *c, unsigned long seed, int sphere_mode, int amp_fixed, int inverted,
                 const double *pk_n2, double *cdelta);
void or_ic_displacements(const or_cfg *c, const double *cdelta, int scale_dependent,
                         float *const za[3], float *const lpt[3], double *cdelta2);
/* from_cdisp_store_to_ZA + assignment: field_a(q) for the stored delta with per-|n|^2 growth g_n2
 * (already including normfactor), then mean removal over the lattice (src/2LPT.c:1539-1735). */
void or_lpt_field(const or_cfg *c, const double *cdelta, const double *g_n2, float *const out[3]);

#ifdef __cplusplus
}
#endif
#endif
