/* Linear power spectrum of the initial conditions (src/power.c): Eisenstein &
 * Hu no-wiggle transfer function, tabulated P(k) or T(k) files, sigma_8
 * normalisation with a top-hat window. */
#include "internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* TransferFunc_EH (src/power.c:447-476) */
static double transfer_eh(const struct mgp_cosmo *c, double k) {
  const mgp_params *p = c->p;
  double hubble = p->HubbleParam, omegam = p->Omega;
  double ombh2 = p->OmegaBaryon * p->HubbleParam * p->HubbleParam;
  if (p->OmegaBaryon == 0) ombh2 = 0.04 * p->HubbleParam * p->HubbleParam;
  k *= (3.085678e24 / p->UnitLength_in_cm);
  double theta = 2.728 / 2.7;
  double ommh2 = omegam * hubble * hubble;
  double s = 44.5 * log(9.83 / ommh2) / sqrt(1. + 10. * exp(0.75 * log(ombh2))) * hubble;
  double a = 1. - 0.328 * log(431. * ommh2) * ombh2 / ommh2 + 0.380 * log(22.3 * ommh2) * (ombh2 / ommh2) * (ombh2 / ommh2);
  double gamma = a + (1. - a) / (1. + exp(4 * log(0.43 * k * s)));
  gamma *= omegam * hubble;
  double q = k * theta * theta / gamma;
  double L0 = log(2. * exp(1.) + 1.8 * q);
  double C0 = 14.2 + 731. / (1. + 62.5 * q);
  return L0 / (L0 + C0 * q * q);
}

/* log-log linear table lookup (PowerSpec_Tabulated / TransferFunc_Tabulated,
 * src/power.c:179-216, 399-436) */
static double table_lookup(const struct mgp_cosmo *c, double k) {
  const mgp_params *p = c->p;
  k *= (p->InputSpectrum_UnitLength_in_cm / p->UnitLength_in_cm);
  double logk = log10(k);
  if (logk < c->tab_logk[0] || logk > c->tab_logk[c->ntab - 1]) return 0;
  int lo = 0, hi = c->ntab - 1;
  while (hi - lo > 1) {
    int mid = (lo + hi) / 2;
    if (logk < c->tab_logk[mid]) hi = mid;
    else lo = mid;
  }
  double dl = c->tab_logk[hi] - c->tab_logk[lo];
  if (dl == 0) return 0;
  double u = (logk - c->tab_logk[lo]) / dl;
  return pow(10.0, (1 - u) * c->tab_logp[lo] + u * c->tab_logp[hi]);
}

/* PowerSpec (src/power.c:373-394) */
double mgp_power_spec(const mgp_cosmo *c, double k) {
  const mgp_params *p = c->p;
  /* with massive neutrinos the CDM + baryon spectrum of the CAMB tables (src/power.c:376-379) */
  if (c->nu) return c->Norm * mgp_pk_cb(c, k);
  switch (p->WhichSpectrum) {
    case 0: {
      double t = p->WhichTransfer == 1 ? table_lookup(c, k) : transfer_eh(c, k);
      return c->Norm * pow(k, p->PrimordialIndex) * t * t;
    }
    case 1:
      return c->Norm * table_lookup(c, k);
    default: {
      double t = transfer_eh(c, k);
      return c->Norm * pow(k, p->PrimordialIndex) * t * t;
    }
  }
}

double mgp_sigma8_norm(const mgp_cosmo *c) { return c->Norm; }

/* TopHatSigma2 integrand (src/power.c:500-516) */
typedef struct { const struct mgp_cosmo *c; double r; } th_par;
static double sigma2_int(double k, void *vp) {
  const th_par *t = (const th_par *)vp;
  double kr = t->r * k, kr2 = kr * kr, kr3 = kr2 * kr;
  if (kr < 1e-8) return 0;
  double w = 3 * (sin(kr) / kr3 - cos(kr) / kr2);
  return k * k * w * w * mgp_power_spec(t->c, k);
}

static int cmp2(const void *a, const void *b) {
  double x = ((const double *)a)[0], y = ((const double *)b)[0];
  return x < y ? -1 : x > y ? 1 : 0;
}

static int read_table(struct mgp_cosmo *c, const char *fname, int normalise_T, char *err, size_t errlen) {
  FILE *fd = fopen(fname, "r");
  if (!fd) {
    snprintf(err, errlen, "can't read input table '%s'", fname);
    return -1;
  }
  int cap = 1024, n = 0;
  double *pairs = malloc(sizeof(double) * 2 * cap), k, v;
  while (fscanf(fd, " %lg %lg ", &k, &v) == 2) {
    if (n == cap) {
      cap *= 2;
      pairs = realloc(pairs, sizeof(double) * 2 * cap);
    }
    pairs[2 * n] = log10(k);
    pairs[2 * n + 1] = log10(v);
    n++;
  }
  fclose(fd);
  if (n < 2) {
    free(pairs);
    snprintf(err, errlen, "input table '%s' has fewer than 2 rows", fname);
    return -1;
  }
  qsort(pairs, n, 2 * sizeof(double), cmp2);
  c->ntab = n;
  c->tab_logk = malloc(sizeof(double) * n);
  c->tab_logp = malloc(sizeof(double) * n);
  for (int i = 0; i < n; i++) {
    c->tab_logk[i] = pairs[2 * i];
    c->tab_logp[i] = pairs[2 * i + 1] - (normalise_T ? pairs[1] : 0.0);
  }
  free(pairs);
  return 0;
}

/* initialize_powerspectrum (src/power.c:226-257) */
int mgp_init_power(struct mgp_cosmo *c, char *err, size_t errlen) {
  const mgp_params *p = c->p;
  if (c->nu) {
    /* read_and_spline_transfer_functions replaces the P(k) input (src/power.c:232-240) */
    if (mgp_camb_read(c, p->nu_FilenameTransferInfofile, err, errlen)) return -1;
  } else if (p->WhichSpectrum == 1 && read_table(c, p->FileWithInputSpectrum, 0, err, errlen)) {
    return -1;
  }
  if (p->WhichSpectrum == 0 && p->WhichTransfer == 1 && read_table(c, p->FileWithInputTransfer, 1, err, errlen))
    return -1;
  c->R8 = 8 * (3.085678e24 / p->UnitLength_in_cm);
  c->Norm = 1.0;
  th_par t = {c, c->R8};
  double res = mgp_integrate(sigma2_int, &t, 0.0, 500.0 / c->R8, 1e-12, 1e-10) / (2.0 * MGP_PI * MGP_PI);
  c->Norm = p->Sigma8 * p->Sigma8 / res;
  return 0;
}
