/*
 * mgpicola.h — C host of the MI355X MG-PICOLA path.
 *
 * The host keeps what the reference computes on the CPU around the hot path:
 * the paramfile interface (src/read_param.c), the model hooks
 * (src/user_defined_functions.h), the growth-factor and power-spectrum tables
 * (src/cosmo.c, src/power.c, src/jbd.c), the COLA time loop (src/main.c) and
 * the output formats; every mesh and particle operation goes through libmgpu
 * (include/mgpu.h).
 */
#ifndef MGPICOLA_HOST_H
#define MGPICOLA_HOST_H

#include <stdio.h>
#include "../../include/mgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MGP_PI 3.14159265358979323846          /* src/vars.h:61 */
#define MGP_GRAVITY 6.672e-8                   /* src/vars.h:62 */
#define MGP_HUBBLE 3.2407789e-18               /* src/vars.h:64 */
#define MGP_INVERSE_H0_MPCH 2997.92458         /* src/vars.h:65 */
#define MGP_MAXOUT 512

/* The reference builds one executable per MODEL= (Makefile:33-110); the host
 * selects the same configurations at run time. */
enum {
  MGP_MODEL_FOFR = 0,       /* FOFRGRAVITY + SCALEDEPENDENT */
  MGP_MODEL_FOFR_LCDM = 1,  /* FOFRGRAVITY */
  MGP_MODEL_DGP = 2,        /* DGPGRAVITY + GAUSSIANFILTER */
  MGP_MODEL_BRANSDICKE = 3, /* BRANSDICKE */
  MGP_MODEL_MBETA = 4,      /* MBETAMODEL + SCALEDEPENDENT */
  MGP_MODEL_FOFRNU = 5      /* FOFRGRAVITY + MASSIVE_NEUTRINOS + SCALEDEPENDENT */
};

typedef struct {
  /* build configuration */
  int model;
  int scale_dependent;
  int compute_pofk;
  int gadget_style;
  /* modified gravity tags (src/user_defined_functions.h:227-406) */
  int modified_gravity_active, use_lcdm_growth_factors, input_pofk_is_for_lcdm, input_sigma8_is_for_lcdm;
  int inverted_initial_condition, amplitude_fixed_initial_condition, include_screening;
  int pofk_compute_every_step, pofk_compute_rsd_pofk, pofk_nbins, pofk_bintype, pofk_subtract_shotnoise;
  double pofk_kmin, pofk_kmax;
  double fofr0, nfofr;
  double range_symm, assb_symm, beta_symm;
  double Rsmooth, rcH0_DGP;
  double wBD, Omegah2, Omegarh2, Omegavh2;
  /* core tags (src/read_param.c:167-319) */
  char OutputDir[500], FileBase[500], OutputRedshiftFile[500];
  int NumFilesWrittenInParallel, UseCOLA;
  double Buffer;
  int Nmesh, Nsample;
  double Box, Init_Redshift;
  int Seed, SphereMode, WhichSpectrum, WhichTransfer;
  char FileWithInputSpectrum[500], FileWithInputTransfer[500];
  double Omega, OmegaBaryon, HubbleParam, Sigma8, PrimordialIndex;
  double UnitVelocity_in_cm_per_s, UnitLength_in_cm, UnitMass_in_g, InputSpectrum_UnitLength_in_cm;
  /* output list (src/read_param.c:32-89), sorted by decreasing redshift */
  int Noutputs;
  double out_z[MGP_MAXOUT];
  int out_nsteps[MGP_MAXOUT];
  /* massive neutrinos (FOFRNU; src/user_defined_functions.h:261-272, :50-78) */
  int nu_include_massive_neutrinos;
  double nu_SumMassNuEV;
  char nu_FilenameTransferInfofile[500];
  double OmegaNu, OmegaCDM; /* derived: SumMassNuEV/(93.14 h^2), Omega - OmegaBaryon - OmegaNu */
  /* host option (not a paramfile tag): build the scale-dependent growth tables on the
   * GPU (mgpu_growth_tables, default) or with the host's own ODE solver */
  int growth_on_gpu;
  /* host option (not a paramfile tag): on steps with no output between Kick and
   * Drift, run both in one particle pass (mgpu_kick_drift, default 1).  Its drift
   * subtracts the mean velocity from before the kick, which equals the mean after
   * the kick up to rounding (the kick term has zero mean); 0 runs the reference's
   * separate Kick and Drift with the mean of the kicked velocities (src/main.c:771). */
  int fuse_kick_drift;
  /* host option: write the particle snapshot at each output (default 1); 0 keeps
   * the outputs' P(k) and RSD files but skips the particle download and files
   * (a 1024^3 snapshot is 26 GB) */
  int write_particles;
  /* host option: the steps between outputs through mgpu_step, with the means kept
   * on the device and no wait for the device inside a step (1), each step also
   * captured and launched as one hipGraph (2: only for a run that has the process
   * to itself), or the separate calls (0, default).  Used where the step is one
   * fused sequence (see mgpu_step); other steps take the separate calls. */
  int device_step;
} mgp_params;

/* ---- paramfile (src/read_param.c) ---- */
int mgp_model_from_name(const char *name);
void mgp_params_defaults(mgp_params *p, int model);
/* strict reader: every tag of the configuration must appear exactly once and
 * nothing else; returns 0 or -1 with a message in err */
int mgp_read_paramfile(const char *fname, mgp_params *p, char *err, size_t errlen);
int mgp_read_outputs(mgp_params *p, char *err, size_t errlen);
/* number of tags the configuration expects, and the i-th name */
int mgp_param_tag_count(const mgp_params *p);
const char *mgp_param_tag_name(const mgp_params *p, int i);

/* ---- cosmology, growth factors, power spectrum ---- */
typedef struct mgp_cosmo mgp_cosmo;
mgp_cosmo *mgp_cosmo_create(mgp_params *p, char *err, size_t errlen);
void mgp_cosmo_free(mgp_cosmo *c);
double mgp_hubble(const mgp_cosmo *c, double a);
double mgp_dhubbleda(const mgp_cosmo *c, double a);
double mgp_growth(const mgp_cosmo *c, int which, double a);     /* which: MGP_G_* */
double mgp_growth_lcdm(const mgp_cosmo *c, int which, double a);
double mgp_growth_sd(const mgp_cosmo *c, int which, double k, double a);
double mgp_growth_lcdmfit(const mgp_cosmo *c, int which, double a);
double mgp_Sq(const mgp_cosmo *c, double ai, double af, double aRef);
double mgp_Sphi(const mgp_cosmo *c, double ai, double af, double aRef);
double mgp_power_spec(const mgp_cosmo *c, double k);
double mgp_sigma8_norm(const mgp_cosmo *c);
double mgp_mg_pofk_ratio(const mgp_cosmo *c, double k, double a);
double mgp_mg_sigma8_enhancement(const mgp_cosmo *c, double a);
/* hooks */
double mgp_GeffoverG(const mgp_cosmo *c, double a, double k);
double mgp_coupling(const mgp_cosmo *c, double a);
double mgp_mass2(const mgp_cosmo *c, double a);
double mgp_beta(const mgp_cosmo *c, double a);
double mgp_beta_dgp(const mgp_cosmo *c, double a);
double mgp_phi_crit(const mgp_cosmo *c, double a);
double mgp_screening_potential(const mgp_cosmo *c, double a, double phin);
double mgp_screening_density(const mgp_cosmo *c, double a, double dens);
void mgp_mg_coeffs(const mgp_cosmo *c, double a, mgpu_mg *out);
/* per-|n|^2 tables for libmgpu */
void mgp_ic_pk_table(const mgp_cosmo *c, double *pk_n2);
/* scale-dependent LPT field tables: g_o(|n|^2) including normfactor for
 * field = MGP_FIELD_*, orders 1 and 2 */
void mgp_lpt_tables(const mgp_cosmo *c, int field, double A, double AFF, double *g1_n2, double *g2_n2);
/* the same as ln k slices for mgpu_lpt_fields_slices; buf: 8 * nk doubles (nk = 1000) */
int mgp_lpt_kslices(const mgp_cosmo *c, int field, double A, double AFF, mgpu_kslice *o1, mgpu_kslice *o2,
                    double *buf);

enum { MGP_G_D = 0, MGP_G_dDdy = 1, MGP_G_ddDddy = 2, MGP_G_D2 = 3, MGP_G_dD2dy = 4, MGP_G_ddD2ddy = 5 };
enum { MGP_FIELD_D = 0, MGP_FIELD_dDdy = 1, MGP_FIELD_ddDddy = 2, MGP_FIELD_deltaD = 3 };

/* ---- the simulation (src/main.c) ---- */
typedef struct {
  double step_seconds;     /* wall time of the step loop */
  long steps;              /* force evaluations in the loop */
  double final_A;
} mgp_run_stats;
int mgp_run(mgp_params *p, int verbose, mgp_run_stats *stats, char *err, size_t errlen);
/* the same run step by step: create (tables + GPU context), initial
 * conditions, then one time-loop iteration per advance call */
typedef struct mgp_sim mgp_sim;
mgp_sim *mgp_sim_create(mgp_params *p, int verbose, char *err, size_t errlen);
/* count runs of one configuration (ps[r] may differ in Seed, the IC pairing flags
 * and the output names only) whose GPU contexts form one mgpu batch; advanced
 * in lockstep (each member's mgp_sim_advance in turn, member 0 first), the steps
 * that go through mgpu_step advance all members with one mgpu_batch_step.  Each
 * member's results are those of its solo run.  0 or -1 (err filled). */
int mgp_sim_create_batch(mgp_params **ps, int count, int verbose, mgp_sim **out, char *err, size_t errlen);
int mgp_sim_ic(mgp_sim *s);
int mgp_sim_advance(mgp_sim *s, int *counted);
int mgp_sim_finished(const mgp_sim *s);
void mgp_sim_free(mgp_sim *s);
mgpu_ctx *mgp_sim_gpu(mgp_sim *s);
mgp_cosmo *mgp_sim_cosmo(mgp_sim *s);
const char *mgp_sim_error(const mgp_sim *s);
double mgp_sim_a(const mgp_sim *s);
/* learned Green's-function multiplier: m (not owned) applied from the next force
 * evaluation on; track = 1 logs a, the density feature and the CDM P(k) of every
 * force evaluation (P(k) when pofk_compute_every_step) */
int mgp_sim_set_learned(mgp_sim *s, mgpu_mlp *m, int track);
int mgp_sim_log_count(const mgp_sim *s);
int mgp_sim_log_entry(const mgp_sim *s, int i, double *a, double *feat, double *pk /* 3 nbins */);
int mgp_sim_nbins(const mgp_sim *s);

/* ---- output (src/main.c:792-1149, src/compute_pofk.c:239-261) ---- */
/* ---- massive neutrinos (CAMB transfer functions, src/read_CAMB_data.h) ---- */
double mgp_T_nu(const mgp_cosmo *c, double k, double a);
double mgp_T_cdm(const mgp_cosmo *c, double k, double a);
double mgp_T_b(const mgp_cosmo *c, double k, double a);
double mgp_T_tot(const mgp_cosmo *c, double k, double a);
double mgp_T_cb(const mgp_cosmo *c, double k, double a);
double mgp_pk_total(const mgp_cosmo *c, double k);
double mgp_pk_cb(const mgp_cosmo *c, double k);
double mgp_munu_1lpt(const mgp_cosmo *c, double a, double k);
double mgp_munu_2lpt(const mgp_cosmo *c, double a, double k);
/* OmegaNu/Omega N^3 T_nu(k, a)/T_cb(k, 1) on the scale-dependent ln k grid (buf: 4 ny doubles);
 * -1 without neutrinos */
int mgp_nu_slice(const mgp_cosmo *c, double a, mgpu_kslice *o, double *buf);
int mgp_write_rsd_pofk(const mgp_params *p, double a, const double *const y[5], const double *const z[5], int nbins);
int mgp_write_pofk(const mgp_params *p, double a, const char *label, const double *pofk, const double *kmean,
                   const double *nmodes, int nbins, int bintype, double kmin, double kmax);
int mgp_write_gadget(const mgp_params *p, double A, double z_label, const float *pos, const float *vel, size_t np);
/* the ASCII snapshot of builds without GADGET_STYLE (src/main.c:999-1027); gadget_style = 0 */
int mgp_write_ascii(const mgp_params *p, double A, double z_label, const float *pos, const float *vel, size_t np);
int mgp_write_info(const mgp_params *p, double A, size_t np);

#ifdef __cplusplus
}
#endif
#endif
