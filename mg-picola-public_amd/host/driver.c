/* The COLA simulation driver: initial conditions, the kick-drift-kick time
 * loop over the output list and the outputs, restating src/main.c:33-787 with
 * every mesh/particle operation on the GPU through libmgpu. */
#include "internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

struct mgp_sim {
  mgp_params *p;
  mgp_cosmo *c;
  mgpu_ctx *g;
  int verbose;
  int stepDistr, StdDA;
  double sumD[3], sumV[3];
  /* P(k) binning after adjust_pofk_parameters (integer-k units) */
  int nbins, bintype, subtract;
  double kmin, kmax;
  double *pk, *pkk, *pkn;
  double *g1, *g2;   /* per-|n|^2 LPT growth tables */
  char *err;
  size_t errlen;
  char errbuf[4096];
  /* time-loop state (src/main.c:386-611) */
  double A, AI, AF, AFF, Di, Di2, da;
  int i, ts, nsteps, NoutputStart, finished;
  double loop_time;
  long steps_done;
  /* learned Green's-function multiplier (host option) and the per-step log of
   * (a, density feature, CDM P(k)) that its training reads */
  int vmean_valid; /* sumV holds the mean of the current velocities (after a kick) */
  int means_on_device; /* the last steps went through mgpu_step: sumD and sumV are on the device */
  mgpu_mlp *mlp;
  int learn_track;
  double learn_theta;
  int nlog, caplog;
  double *log_a, *log_feat, *log_pk; /* log_pk: 3 nbins per entry (P, k mean, modes) */
  /* a member of a batch (mgp_sim_create_batch): the record shared by the members and this member's index */
  struct batch_rec *brec;
  int batch_idx;
};
typedef struct mgp_sim sim;

/* Shared by the members of a batch: their contexts in order, and the time-loop
 * iteration in which member 0 last stepped all of them with mgpu_batch_step. */
struct batch_rec {
  mgpu_ctx **ctxs;
  struct mgp_sim **sims;
  int count, refs;
  long stepped_iter;
  int ic_done; /* member 0 made every member's initial conditions with one batched IC */
};

static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

#define GPU(call)                                                                      \
  do {                                                                                 \
    int rc_ = (call);                                                                  \
    if (rc_) {                                                                         \
      snprintf(s->err, s->errlen, "%s failed (%d): %s", #call, rc_, mgpu_last_error(s->g)); \
      return -1;                                                                       \
    }                                                                                  \
  } while (0)

/* adjust_pofk_parameters (src/compute_pofk.c:758-805) on integer-k units */
static void adjust_pofk(const mgp_params *p, int *nbins, int *bintype, int *sub, double *kmin, double *kmax) {
  if (!(*bintype == 0 || *bintype == 1)) *bintype = 0;
  if (!(*sub == 0 || *sub == 1)) *sub = 1;
  if (*nbins <= 0) *nbins = p->Nmesh;
  if (*kmax <= *kmin) {
    if (*bintype == 0) *kmin = 0.0;
    if (*bintype == 1) *kmin = 1.0;
    *kmax = (double)p->Nmesh;
  }
  if (*kmin < 0.0) *kmin = *bintype == 0 ? 0.0 : 1.0;
  if (*bintype == 1 && *kmin == 0.0) *kmin = 1.0;
  if (*kmax > sqrt(3.0) * (double)p->Nmesh) *kmax = (double)p->Nmesh;
}

/* assign_displacment_field_to_particles for both LPT orders at once
 * (src/2LPT.c:1758-2005): the growth goes to the GPU as ln k slices */
static int lpt_fields(sim *s, int slot, int field, double A, double AFF) {
  mgpu_kslice o1, o2;
  if (mgp_lpt_kslices(s->c, field, A, AFF, &o1, &o2, s->g1)) {
    snprintf(s->err, s->errlen, "no scale-dependent growth tables");
    return -1;
  }
  GPU(mgpu_lpt_fields_slices(s->g, slot, &o1, &o2));
  return 0;
}

/* both LPT slots at once: slot 0 gets field f0, slot 1 field f1 */
static int lpt_fields2(sim *s, int f0, int f1, double A, double AFF) {
  mgpu_kslice a1, a2, b1, b2;
  if (mgp_lpt_kslices(s->c, f0, A, AFF, &a1, &a2, s->g1) || mgp_lpt_kslices(s->c, f1, A, AFF, &b1, &b2, s->g2)) {
    snprintf(s->err, s->errlen, "no scale-dependent growth tables");
    return -1;
  }
  GPU(mgpu_lpt_fields_slices2(s->g, &a1, &a2, &b1, &b2));
  return 0;
}

/* Kick, src/main.c:688-742 */
static int kick(sim *s, double AI, double AF, double A) {
  double dda;
  if (s->StdDA == 0) dda = mgp_Sphi(s->c, AI, AF, A);
  else {
    extern double mgp_SphiStd(const mgp_cosmo *, double, double);
    dda = mgp_SphiStd(s->c, AI, AF);
  }
  double s1 = 0, s2 = 0;
  if (!s->p->scale_dependent) {
    s1 = mgp_growth(s->c, MGP_G_ddDddy, A);
    s2 = mgp_growth(s->c, MGP_G_ddD2ddy, A);
  }
  GPU(mgpu_kick(s->g, dda, A, s1, s2, s->sumD, s->sumV));
  return 0;
}

/* Drift, src/main.c:747-787 */
static int drift(sim *s, double A, double AFF, double AF, double Di, double Di2) {
  double dyyy;
  if (s->StdDA == 0) dyyy = mgp_Sq(s->c, A, AFF, AF);
  else {
    extern double mgp_SqStd(const mgp_cosmo *, double, double);
    dyyy = mgp_SqStd(s->c, A, AFF);
  }
  double s1 = 0, s2 = 0;
  if (!s->p->scale_dependent) {
    s1 = mgp_growth(s->c, MGP_G_D, AFF) - Di;
    s2 = mgp_growth(s->c, MGP_G_D2, AFF) - Di2;
  }
  GPU(mgpu_drift(s->g, dyyy, s1, s2, s->sumV));
  return 0;
}

/* Kick then Drift in one GPU pass (the steps with no output between them): the
 * same time factors as kick() and drift(); the drift's mean velocity is the one
 * before the kick (s->sumV, valid once a kick has run). */
/* the coefficients of mgpu_kick_drift / mgpu_step for one step */
static void kick_drift_coeffs(sim *s, double AI, double AF, double A, double AFF, double Di, double Di2, double c[7]) {
  double dda, dyyy;
  if (s->StdDA == 0) {
    dda = mgp_Sphi(s->c, AI, AF, A);
    dyyy = mgp_Sq(s->c, A, AFF, AF);
  } else {
    extern double mgp_SphiStd(const mgp_cosmo *, double, double);
    extern double mgp_SqStd(const mgp_cosmo *, double, double);
    dda = mgp_SphiStd(s->c, AI, AF);
    dyyy = mgp_SqStd(s->c, A, AFF);
  }
  double s1k = 0, s2k = 0, s1d = 0, s2d = 0;
  if (!s->p->scale_dependent) {
    s1k = mgp_growth(s->c, MGP_G_ddDddy, A);
    s2k = mgp_growth(s->c, MGP_G_ddD2ddy, A);
    s1d = mgp_growth(s->c, MGP_G_D, AFF) - Di;
    s2d = mgp_growth(s->c, MGP_G_D2, AFF) - Di2;
  }
  c[0] = dda, c[1] = s1k, c[2] = s2k, c[3] = dyyy, c[4] = s1d, c[5] = s2d, c[6] = A;
}

/* mgpu_step may apply: nothing of this step is read back on the host (mgpu_step
 * itself refuses, with MGPU_E_STATE, the models whose step is not one fused
 * sequence, and those steps take the separate calls) */
static int device_step_ok(const sim *s) {
  const mgp_params *p = s->p;
  return p->device_step && !p->scale_dependent && !(p->compute_pofk && p->pofk_compute_every_step) &&
         !(p->compute_pofk && p->pofk_compute_rsd_pofk == 1) && !s->mlp && !s->learn_track;
}

static int kick_drift(sim *s, double AI, double AF, double A, double AFF, double Di, double Di2) {
  double dda, dyyy;
  if (s->StdDA == 0) {
    dda = mgp_Sphi(s->c, AI, AF, A);
    dyyy = mgp_Sq(s->c, A, AFF, AF);
  } else {
    extern double mgp_SphiStd(const mgp_cosmo *, double, double);
    extern double mgp_SqStd(const mgp_cosmo *, double, double);
    dda = mgp_SphiStd(s->c, AI, AF);
    dyyy = mgp_SqStd(s->c, A, AFF);
  }
  double s1k = 0, s2k = 0, s1d = 0, s2d = 0;
  if (!s->p->scale_dependent) {
    s1k = mgp_growth(s->c, MGP_G_ddDddy, A);
    s2k = mgp_growth(s->c, MGP_G_ddD2ddy, A);
    s1d = mgp_growth(s->c, MGP_G_D, AFF) - Di;
    s2d = mgp_growth(s->c, MGP_G_D2, AFF) - Di2;
  }
  GPU(mgpu_kick_drift(s->g, dda, A, s1k, s2k, s->sumD, dyyy, s1d, s2d, s->sumV));
  return 0;
}

/* compute_RSD_powerspectrum (src/compute_pofk.c:403-500): P0, P2, P4 of the y and z lines
 * of sight, averaged, to pofk_RSD_*.  Scale-dependent COLA needs dD/dy in LPT slot 1. */
static int rsd_pofk(sim *s, double A) {
  const mgp_params *p = s->p;
  const double UnitTime_in_s = p->UnitLength_in_cm / p->UnitVelocity_in_cm_per_s;
  const double Hubble = MGP_HUBBLE * UnitTime_in_s;
  const double vnorm = (Hubble / A) / (100.0 * A * mgp_hubble(s->c, A)) * ((double)p->Nmesh / p->Box);
  const double s1 = p->scale_dependent ? 0.0 : mgp_growth(s->c, MGP_G_dDdy, A);
  const double s2 = p->scale_dependent ? 0.0 : mgp_growth(s->c, MGP_G_dD2dy, A);
  const int nb = s->nbins;
  double *buf = calloc(10 * (size_t)nb, sizeof(double));
  const double *y[5], *z[5];
  for (int l = 0; l < 5; l++) {
    y[l] = buf + l * nb;
    z[l] = buf + (5 + l) * nb;
  }
  mgpu_rsd_pofk ry = {buf, buf + nb, buf + 2 * nb, buf + 3 * nb, buf + 4 * nb};
  mgpu_rsd_pofk rz = {buf + 5 * nb, buf + 6 * nb, buf + 7 * nb, buf + 8 * nb, buf + 9 * nb};
  int rc = mgpu_rsd_power_spectrum(s->g, 1, vnorm, s1, s2, &ry);
  if (!rc) rc = mgpu_rsd_power_spectrum(s->g, 2, vnorm, s1, s2, &rz);
  if (rc) {
    snprintf(s->err, s->errlen, "RSD power spectrum: %s", mgpu_last_error(s->g));
    free(buf);
    return -1;
  }
  if (mgp_write_rsd_pofk(p, A, y, z, nb) && s->verbose)
    fprintf(stderr, "warning: could not write the RSD P(k) file in %s\n", p->OutputDir);
  free(buf);
  return 0;
}

/* Output, src/main.c:792-1062 */
static int output(sim *s, double A, double AF, double AFF, double z_label) {
  const mgp_params *p = s->p;
  const size_t np = (size_t)p->Nsample * p->Nsample * p->Nsample;
  if (p->scale_dependent) {
    if (lpt_fields(s, 1, MGP_FIELD_dDdy, A, AFF)) return -1;
  }
  /* RSD multipoles at the output, with the velocities synchronised (src/main.c:875-879) */
  const int first_step = fabs(p->Init_Redshift - z_label) / p->Init_Redshift <= 1.0E-6;
  if (p->compute_pofk && p->pofk_compute_rsd_pofk >= 1 && !first_step && rsd_pofk(s, A)) return -1;
  if (!p->write_particles) {
    if (p->scale_dependent && lpt_fields(s, 1, MGP_FIELD_deltaD, A, AFF)) return -1;
    return 0;
  }
  float *pos = malloc(sizeof(float) * 3 * np), *vel = malloc(sizeof(float) * 3 * np);
  float *l0 = malloc(sizeof(float) * 3 * np), *l1 = malloc(sizeof(float) * 3 * np);
  int rc = mgpu_get_particles(s->g, pos, vel) || mgpu_get_lpt(s->g, 0, l0) || mgpu_get_lpt(s->g, 1, l1);
  if (!rc) {
    const double UnitTime_in_s = p->UnitLength_in_cm / p->UnitVelocity_in_cm_per_s;
    const double Hubble = MGP_HUBBLE * UnitTime_in_s;
    const double fac = Hubble / pow(A, 1.5), velfac = p->UnitVelocity_in_cm_per_s / 1.0e5;
    const double dDdy = mgp_growth(s->c, MGP_G_dDdy, A), dD2dy = mgp_growth(s->c, MGP_G_dD2dy, A);
    for (int k = 0; k < 3; k++)
      for (size_t n = 0; n < np; n++) {
        size_t i = k * np + n;
        double lptv = p->scale_dependent ? (double)l1[i] : (double)l0[i] * dDdy + (double)l1[i] * dD2dy;
        vel[i] = (float)(velfac * fac * (vel[i] - s->sumV[k] + lptv * p->UseCOLA));
      }
    rc = p->gadget_style ? mgp_write_gadget(p, A, z_label, pos, vel, np) : mgp_write_ascii(p, A, z_label, pos, vel, np);
    if (!rc) rc = mgp_write_info(p, A, np);
  }
  free(pos);
  free(vel);
  free(l0);
  free(l1);
  if (rc) {
    snprintf(s->err, s->errlen, "writing output at z = %f failed", z_label);
    return -1;
  }
  if (p->scale_dependent && lpt_fields(s, 1, MGP_FIELD_deltaD, A, AFF)) return -1;
  return 0;
}

/* one log entry per force evaluation: a, the density feature and the CDM P(k) */
static int learn_log(sim *s, double A, int have_pk) {
  double f = 0.0;
  GPU(mgpu_density_feature(s->g, &f));
  if (s->nlog == s->caplog) {
    const int cap = s->caplog ? 2 * s->caplog : 64;
    double *la = realloc(s->log_a, sizeof(double) * cap), *lf = realloc(s->log_feat, sizeof(double) * cap);
    if (la) s->log_a = la;
    if (lf) s->log_feat = lf;
    double *lp = realloc(s->log_pk, sizeof(double) * 3 * s->nbins * cap);
    if (lp) s->log_pk = lp;
    if (!la || !lf || !lp) {
      snprintf(s->err, s->errlen, "out of host memory for the step log");
      return -1;
    }
    s->caplog = cap;
  }
  double *o = s->log_pk + (size_t)3 * s->nbins * s->nlog;
  for (int b = 0; b < s->nbins; b++) {
    o[b] = have_pk && s->pk ? s->pk[b] : 0.0;
    o[s->nbins + b] = have_pk && s->pkk ? s->pkk[b] : 0.0;
    o[2 * s->nbins + b] = have_pk && s->pkn ? s->pkn[b] : 0.0;
  }
  s->log_a[s->nlog] = A;
  s->log_feat[s->nlog] = f;
  s->nlog++;
  return 0;
}

static int get_displacements(sim *s, double A) {
  const mgp_params *p = s->p;
  mgpu_mg mg;
  mgp_mg_coeffs(s->c, A, &mg);
  const int want_pk = p->compute_pofk && p->pofk_compute_every_step;
  mgpu_pofk pk = {s->pk, s->pkk, s->pkn};
  const int nu_model = p->model == MGP_MODEL_FOFRNU;
  if (nu_model && p->nu_include_massive_neutrinos) {
    /* the neutrino term of PtoMesh at this a (src/auxPM.c:383-420) */
    mgpu_kslice f;
    if (mgp_nu_slice(s->c, A, &f, s->g1)) {
      snprintf(s->err, s->errlen, "no neutrino transfer functions");
      return -1;
    }
    GPU(mgpu_set_neutrinos(s->g, (p->Omega - p->OmegaNu) / p->Omega, &f));
  }
  if (s->mlp || s->learn_track) GPU(mgpu_set_learned_multiplier(s->g, s->mlp, A, s->learn_theta));
  GPU(mgpu_get_displacements(s->g, p->modified_gravity_active ? &mg : NULL, want_pk ? &pk : NULL, s->sumD));
  if (s->learn_track && learn_log(s, A, want_pk)) return -1;
  if (want_pk && s->pk &&
      mgp_write_pofk(p, A, "CDM", s->pk, s->pkk, s->pkn, s->nbins, s->bintype, s->kmin, s->kmax) && s->verbose)
    fprintf(stderr, "warning: could not write P(k) file in %s\n", p->OutputDir);
  if (nu_model && want_pk && s->pk) {
    /* the total matter P(k) after the neutrino term (src/auxPM.c:422-426) */
    GPU(mgpu_power_spectrum(s->g, &pk));
    if (mgp_write_pofk(p, A, "total", s->pk, s->pkk, s->pkn, s->nbins, s->bintype, s->kmin, s->kmax) && s->verbose)
      fprintf(stderr, "warning: could not write P(k) file in %s\n", p->OutputDir);
  }
  /* RSD multipoles every step (src/auxPM.c:373-380).  The reference skips the steps it
   * takes to be output steps with (timeStep == 0) && (i != NoutputStart), where i is
   * PtoMesh's slab loop counter (2 alloc_slice by then), so in effect every first step
   * of an output interval.  Scale-dependent: dD/dy at A goes to LPT slot 1, which the
   * step refills before Kick. */
  if (p->compute_pofk && p->pofk_compute_rsd_pofk == 1 && s->ts != 0) {
    if (p->scale_dependent && p->UseCOLA && lpt_fields(s, 1, MGP_FIELD_dDdy, A, A)) return -1;
    if (rsd_pofk(s, A)) return -1;
  }
  return 0;
}


/* Set-up: growth tables and the GPU context (src/main.c:68-162). */
/* the host side of a run up to its GPU context: tables, P(k) binning and the context's parameters */
static sim *sim_alloc(mgp_params *p, int verbose, mgpu_params *gp) {
  sim *s = calloc(1, sizeof(sim));
  s->p = p;
  s->verbose = verbose;
  s->err = s->errbuf;
  s->errlen = sizeof s->errbuf;
  mkdir(p->OutputDir, 0755);
  if (p->UseCOLA) {
    s->stepDistr = 0;
    s->StdDA = 0;
  } else {
    s->stepDistr = 1;
    s->StdDA = 2;
  }
  const double t_start = now_s();
  s->c = mgp_cosmo_create(p, s->err, s->errlen);
  if (!s->c) return s;
  if (s->verbose) printf("[mgpicola] growth tables ready (%.2f s)\n", now_s() - t_start);
  memset(gp, 0, sizeof *gp);
  gp->nmesh = p->Nmesh;
  gp->nsample = p->Nsample;
  gp->box = p->Box;
  gp->omega = p->Omega;
  switch (p->model) {
    case MGP_MODEL_FOFR:
    case MGP_MODEL_FOFRNU:
    case MGP_MODEL_FOFR_LCDM: gp->model = MGPU_FOFR; break;
    case MGP_MODEL_DGP: gp->model = MGPU_DGP; break;
    case MGP_MODEL_BRANSDICKE: gp->model = MGPU_BRANSDICKE; break;
    default: gp->model = MGPU_MBETA;
  }
  gp->mg_active = p->modified_gravity_active;
  gp->include_screening = p->include_screening;
  gp->filter = p->model == MGP_MODEL_DGP ? MGPU_FILTER_GAUSSIAN : MGPU_FILTER_NONE;
  gp->scale_dependent = p->scale_dependent;
  gp->use_cola = p->UseCOLA;
  s->nbins = p->pofk_nbins;
  s->bintype = p->pofk_bintype;
  s->subtract = p->pofk_subtract_shotnoise;
  s->kmin = p->pofk_kmin * p->Box / (2.0 * M_PI);
  s->kmax = p->pofk_kmax * p->Box / (2.0 * M_PI);
  adjust_pofk(p, &s->nbins, &s->bintype, &s->subtract, &s->kmin, &s->kmax);
  gp->pofk_nbins = s->nbins;
  gp->pofk_bintype = s->bintype;
  gp->pofk_subtract_shotnoise = s->subtract;
  gp->pofk_kmin = s->kmin;
  gp->pofk_kmax = s->kmax;
  return s;
}

/* the column passes' tile width measured on this device, before any data */
static int sim_autotune(sim *s) {
  double ms[4];
  if (mgpu_autotune_fft(s->g, ms)) {
    snprintf(s->err, s->errlen, "FFT autotune failed: %s", mgpu_last_error(s->g));
    return -1;
  }
  int t0 = 0;
  mgpu_get_fft_tile(s->g, 0, &t0);
  if (s->verbose)
    printf("[mgpicola] FFT tiles: %d columns (8: %.3f ms, 16: %.3f ms per y pass pair)\n", t0, ms[0], ms[1]);
  return 0;
}

static void sim_buffers(sim *s) {
  const mgp_params *p = s->p;
  s->pk = calloc(s->nbins, sizeof(double));
  s->pkk = calloc(s->nbins, sizeof(double));
  s->pkn = calloc(s->nbins, sizeof(double));
  /* g1 also holds the 8 growth slices of 1000 ln k nodes (mgp_lpt_kslices) */
  const size_t n2len = mgpu_n2_table_len(p->Nmesh);
  s->g1 = malloc(sizeof(double) * (n2len > 8000 ? n2len : 8000));
  s->g2 = malloc(sizeof(double) * (n2len > 8000 ? n2len : 8000));
}

mgp_sim *mgp_sim_create(mgp_params *p, int verbose, char *err, size_t errlen) {
  mgpu_params gp;
  sim *s = sim_alloc(p, verbose, &gp);
  if (!s->c) goto fail;
  if (mgpu_create(&gp, &s->g)) {
    snprintf(s->err, s->errlen, "mgpu_create failed: %s", mgpu_last_error(NULL));
    goto fail;
  }
  if (sim_autotune(s)) goto fail;
  sim_buffers(s);
  return s;
fail:
  snprintf(err, errlen, "%s", s->err);
  mgp_sim_free(s);
  return NULL;
}

int mgp_sim_create_batch(mgp_params **ps, int count, int verbose, mgp_sim **out, char *err, size_t errlen) {
  if (count < 1) {
    snprintf(err, errlen, "empty batch");
    return -1;
  }
  for (int r = 0; r < count; r++) out[r] = NULL;
  mgpu_params gp0, gp;
  struct batch_rec *b = calloc(1, sizeof *b);
  b->ctxs = calloc(count, sizeof(mgpu_ctx *));
  b->sims = calloc(count, sizeof(struct mgp_sim *));
  b->count = count;
  b->stepped_iter = -1;
  int rc = 0;
  for (int r = 0; r < count && !rc; r++) {
    out[r] = sim_alloc(ps[r], verbose, r ? &gp : &gp0);
    b->sims[r] = out[r];
    out[r]->brec = b;
    out[r]->batch_idx = r;
    b->refs++;
    if (!out[r]->c) {
      snprintf(err, errlen, "member %d: %s", r, out[r]->err);
      rc = -1;
    } else if (r && memcmp(&gp, &gp0, sizeof gp)) {
      snprintf(err, errlen, "member %d: the members of a batch differ in mesh, particles, box, model or P(k) bins", r);
      rc = -1;
    }
  }
  if (!rc && mgpu_batch_create(&gp0, count, b->ctxs)) {
    snprintf(err, errlen, "mgpu_batch_create failed: %s", mgpu_last_error(NULL));
    rc = -1;
  }
  if (!rc) {
    for (int r = 0; r < count; r++) out[r]->g = b->ctxs[r];
    /* one width for all: the members' kernels are launched together */
    int t0 = 0;
    if (sim_autotune(out[0])) {
      snprintf(err, errlen, "%s", out[0]->err);
      rc = -1;
    } else {
      mgpu_get_fft_tile(out[0]->g, 0, &t0);
      for (int r = 1; r < count && !rc; r++)
        if (mgpu_set_fft_tile(out[r]->g, 0, t0)) {
          snprintf(err, errlen, "member %d: %s", r, mgpu_last_error(out[r]->g));
          rc = -1;
        }
    }
  }
  if (rc) {
    for (int r = 0; r < count; r++) {
      if (out[r] && !out[r]->g && b->ctxs[r]) mgpu_destroy(b->ctxs[r]);
      mgp_sim_free(out[r]);
      out[r] = NULL;
    }
    return -1;
  }
  for (int r = 0; r < count; r++) sim_buffers(out[r]);
  return 0;
}

void mgp_sim_free(mgp_sim *s) {
  if (!s) return;
  if (s->g) mgpu_destroy(s->g);
  if (s->brec && --s->brec->refs == 0) {
    free(s->brec->sims);
    free(s->brec->ctxs);
    free(s->brec);
  }
  mgp_cosmo_free(s->c);
  free(s->pk);
  free(s->pkk);
  free(s->pkn);
  free(s->g1);
  free(s->g2);
  free(s->log_a);
  free(s->log_feat);
  free(s->log_pk);
  free(s);
}

/* The network's theta input: (log10 f_R0 + 6)/2 for active f(R), -1 otherwise. */
int mgp_sim_set_learned(mgp_sim *s, mgpu_mlp *m, int track) {
  const mgp_params *p = s->p;
  const int fofr = p->model == MGP_MODEL_FOFR || p->model == MGP_MODEL_FOFR_LCDM || p->model == MGP_MODEL_FOFRNU;
  s->mlp = m;
  s->learn_track = track;
  s->learn_theta = fofr && p->modified_gravity_active && p->fofr0 > 0 ? (log10(p->fofr0) + 6.0) / 2.0 : -1.0;
  if (!m && !track) GPU(mgpu_set_learned_multiplier(s->g, NULL, -1.0, 0.0));
  return 0;
}
int mgp_sim_log_count(const mgp_sim *s) { return s->nlog; }
int mgp_sim_log_entry(const mgp_sim *s, int i, double *a, double *feat, double *pk) {
  if (i < 0 || i >= s->nlog) return -1;
  *a = s->log_a[i];
  *feat = s->log_feat[i];
  memcpy(pk, s->log_pk + (size_t)3 * s->nbins * i, sizeof(double) * 3 * s->nbins);
  return 0;
}
int mgp_sim_nbins(const mgp_sim *s) { return s->nbins; }

mgpu_ctx *mgp_sim_gpu(mgp_sim *s) { return s->g; }
mgp_cosmo *mgp_sim_cosmo(mgp_sim *s) { return s->c; }
const char *mgp_sim_error(const mgp_sim *s) { return s->errbuf; }
double mgp_sim_a(const mgp_sim *s) { return s->A; }

/* Initial conditions and particle set-up (src/main.c:164-384). */
int mgp_sim_ic(mgp_sim *s) {
  const mgp_params *p = s->p;
  const size_t n2len = mgpu_n2_table_len(p->Nmesh);
  s->A = 1.0 / (1.0 + p->Init_Redshift);
  s->Di = mgp_growth(s->c, MGP_G_D, s->A);
  s->Di2 = mgp_growth(s->c, MGP_G_D2, s->A);
  s->AI = s->AF = s->AFF = s->A;
  double *pk = malloc(sizeof(double) * n2len);
  mgp_ic_pk_table(s->c, pk);
  int rc = 0;
  struct batch_rec *b = s->brec;
  if (b && b->ic_done && s->batch_idx > 0) {
    /* made by member 0's batched IC */
  } else if (b && s->batch_idx == 0 && !p->scale_dependent) {
    /* one batched IC for all members when they share the sphere mode and the amplitude fixing */
    int same = 1;
    for (int r = 1; r < b->count; r++) {
      const mgp_params *q = b->sims[r]->p;
      same &= q->SphereMode == p->SphereMode &&
              q->amplitude_fixed_initial_condition == p->amplitude_fixed_initial_condition;
    }
    if (same) {
      unsigned long *seeds = malloc(sizeof(unsigned long) * b->count);
      int *inv = malloc(sizeof(int) * b->count);
      for (int r = 0; r < b->count; r++) {
        seeds[r] = (unsigned long)b->sims[r]->p->Seed;
        inv[r] = b->sims[r]->p->inverted_initial_condition;
      }
      rc = mgpu_batch_ic_displacement_fields(b->ctxs, b->count, seeds, p->SphereMode,
                                             p->amplitude_fixed_initial_condition, inv, pk);
      free(seeds);
      free(inv);
      b->ic_done = rc == 0;
    } else {
      rc = mgpu_ic_displacement_fields(s->g, (unsigned long)p->Seed, p->SphereMode,
                                       p->amplitude_fixed_initial_condition, p->inverted_initial_condition, pk);
    }
  } else {
    rc = mgpu_ic_displacement_fields(s->g, (unsigned long)p->Seed, p->SphereMode,
                                     p->amplitude_fixed_initial_condition, p->inverted_initial_condition, pk);
  }
  free(pk);
  if (rc) {
    snprintf(s->err, s->errlen, "IC generation failed: %s", mgpu_last_error(s->g));
    return -1;
  }
  if (p->scale_dependent) {
    if (lpt_fields2(s, MGP_FIELD_D, MGP_FIELD_dDdy, s->A, s->AFF)) return -1;
  }
  GPU(mgpu_init_particles(s->g, s->Di, s->Di2));
  s->sumV[0] = s->sumV[1] = s->sumV[2] = 0.0;
  s->NoutputStart = 0;
  if ((p->Init_Redshift - p->out_z[0]) / p->Init_Redshift <= 1.0E-6 || p->Init_Redshift <= 1.0e-6) {
    if (output(s, s->A, s->AF, s->AFF, p->out_z[0])) return -1;
    if (p->Noutputs == 1) {
      s->finished = 1;
      return 0;
    }
    s->NoutputStart++;
  }
  s->i = s->NoutputStart;
  s->ts = 0;
  s->nsteps = -1;
  return 0;
}

/* One iteration of the time loop body (src/main.c:405-610).  *counted is 1
 * when the iteration is one of the timed steps (not the final sync half-kick). */
int mgp_sim_advance(mgp_sim *s, int *counted) {
  const mgp_params *p = s->p;
  *counted = 0;
  if (s->finished) return 0;
  if (s->nsteps < 0) { /* entering output interval i */
    s->da = 0.0;
    if (s->i == p->Noutputs) s->nsteps = 1;
    else {
      s->nsteps = p->out_nsteps[s->i];
      double ao = 1.0 / (1.0 + p->out_z[s->i]);
      if (s->stepDistr == 0) s->da = (ao - s->A) / (double)s->nsteps;
      else s->da = (log(ao) - log(s->A)) / (double)s->nsteps;
    }
    s->ts = 0;
  }
  const double t0 = now_s();
  const double da = s->da;
  double A = s->A;
  if (s->ts == 0 && s->i != s->NoutputStart) s->AF = A;
  else s->AF = s->stepDistr == 0 ? A + da * 0.5 : A * exp(da * 0.5);
  s->AFF = s->stepDistr == 0 ? A + da : A * exp(da);
  if (s->verbose) {
    if (s->i != p->Noutputs) printf("[mgpicola] step a = %f -> %f\n", A, s->AFF);
    else printf("[mgpicola] final half timestep\n");
  }
  const int output_now = s->ts == 0 && s->i != s->NoutputStart;
  int stepped = 0;
  if (s->brec && s->batch_idx > 0 && s->brec->stepped_iter == s->steps_done) {
    /* member 0 stepped the whole batch in this iteration */
    stepped = 1;
    s->means_on_device = 1;
  } else if (!output_now && s->vmean_valid && p->fuse_kick_drift && device_step_ok(s) &&
             (!s->brec || s->batch_idx == 0)) {
    /* the whole step on the device: the means stay there */
    mgpu_mg mg;
    mgp_mg_coeffs(s->c, A, &mg);
    double c[7];
    kick_drift_coeffs(s, s->AI, s->AF, A, s->AFF, s->Di, s->Di2, c);
    const mgpu_mg *pm = p->modified_gravity_active ? &mg : NULL;
    int rc = s->brec ? mgpu_batch_step(s->brec->ctxs, s->brec->count, pm, c[0], c[6], c[1], c[2], c[3], c[4], c[5],
                                       p->device_step == 2)
                     : mgpu_step(s->g, pm, c[0], c[6], c[1], c[2], c[3], c[4], c[5], p->device_step == 2);
    if (s->brec) s->brec->stepped_iter = rc == 0 ? s->steps_done : -1;
    if (rc == 0) {
      stepped = 1;
      s->means_on_device = 1;
    } else if (rc != MGPU_E_STATE) {
      snprintf(s->err, s->errlen, "mgpu_step failed (%d): %s", rc, mgpu_last_error(s->g));
      return -1;
    }
  }
  if (!stepped && s->means_on_device) {
    /* back to the separate calls: the last kick's velocity mean from the device */
    GPU(mgpu_get_means(s->g, NULL, s->sumV));
    s->means_on_device = 0;
  }
  if (!stepped) {
  if (get_displacements(s, A)) return -1;
  if (p->scale_dependent) {
    if (lpt_fields2(s, MGP_FIELD_ddDddy, MGP_FIELD_deltaD, A, s->AFF)) return -1;
  }
  if (!output_now && s->vmean_valid && p->fuse_kick_drift) {
    if (kick_drift(s, s->AI, s->AF, A, s->AFF, s->Di, s->Di2)) return -1;
  } else {
    if (kick(s, s->AI, s->AF, A)) return -1;
    s->vmean_valid = 1;
    if (output_now) {
      if (output(s, A, s->AF, s->AFF, p->out_z[s->i - 1])) return -1;
      if (s->i == p->Noutputs) {
        s->finished = 1;
        return 0;
      }
      s->AI = A;
      s->AF = s->stepDistr == 0 ? A + da * 0.5 : A * exp(da * 0.5);
      s->sumD[0] = s->sumD[1] = s->sumD[2] = 0.0;
      if (kick(s, s->AI, s->AF, A)) return -1;
    }
    if (drift(s, A, s->AFF, s->AF, s->Di, s->Di2)) return -1;
  }
  }
  s->A = s->AFF;
  s->AI = s->AF;
  s->Di = mgp_growth(s->c, MGP_G_D, s->A);
  s->Di2 = mgp_growth(s->c, MGP_G_D2, s->A);
  if (!stepped) GPU(mgpu_sync(s->g)); /* device steps are left in flight: the next call queues behind them */
  s->loop_time += now_s() - t0;
  s->steps_done++;
  *counted = 1;
  if (++s->ts == s->nsteps) {
    s->i++;
    s->nsteps = -1;
    if (s->i > p->Noutputs) s->finished = 1;
  }
  return 0;
}

int mgp_sim_finished(const mgp_sim *s) { return s->finished; }

int mgp_run(mgp_params *p, int verbose, mgp_run_stats *st, char *err, size_t errlen) {
  err[0] = 0;
  sim *s = mgp_sim_create(p, verbose, err, errlen);
  if (!s) return -1;
  int rc = mgp_sim_ic(s);
  while (!rc && !s->finished) {
    int counted;
    rc = mgp_sim_advance(s, &counted);
  }
  if (rc) snprintf(err, errlen, "%s", s->errbuf);
  if (st) {
    st->step_seconds = s->loop_time;
    st->steps = s->steps_done;
    st->final_A = s->A;
  }
  mgp_sim_free(s);
  return rc;
}
