/* Massive-neutrino inputs (src/read_CAMB_data.h, src/user_defined_functions.h:50-78, :655-680,
 * src/power.c:232-240, :376-381): CAMB transfer functions at several redshifts, splined in
 * (ln k, z), and the quantities built from them. */
#include "internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { NCOL = 13, COL_K = 0, COL_CDM = 1, COL_B = 2, COL_NU = 5, COL_TOT = 6, NMAX_ROWS = 5000 };

/* read_single_transfer_function_file (src/read_CAMB_data.h:238-275): one header line,
 * 13 columns; ln k, T_cdm, T_b, T_massive_nu, T_total */
static int read_transfer_file(const char *fn, double *lnk, double *cdm, double *b, double *nu, double *tot, int *npts) {
  FILE *fp = fopen(fn, "r");
  if (!fp) return -1;
  int ch;
  while ((ch = fgetc(fp)) != EOF && ch != '\n') {
  }
  int n = 0;
  double t[NCOL];
  for (;;) {
    if (fscanf(fp, "%lf", &t[0]) != 1) break;
    for (int i = 1; i < NCOL; i++)
      if (fscanf(fp, "%lf", &t[i]) != 1) {
        fclose(fp);
        return -1;
      }
    if (n >= NMAX_ROWS) {
      fclose(fp);
      return -1;
    }
    lnk[n] = log(t[COL_K]);
    cdm[n] = t[COL_CDM];
    b[n] = t[COL_B];
    nu[n] = t[COL_NU];
    tot[n] = t[COL_TOT];
    n++;
  }
  fclose(fp);
  *npts = n;
  return 0;
}

/* read_and_spline_transfer_functions (src/read_CAMB_data.h:41-160): the info file holds
 * "<directory> <nredshift>" and then "<file> <z>" per redshift */
int mgp_camb_read(struct mgp_cosmo *c, const char *infofile, char *err, size_t errlen) {
  FILE *fp = fopen(infofile, "r");
  if (!fp) {
    snprintf(err, errlen, "cannot read the transfer info file [%s]", infofile);
    return -1;
  }
  char dir[400];
  int nz = 0;
  if (fscanf(fp, "%399s %d", dir, &nz) != 2 || nz < 2) {
    fclose(fp);
    snprintf(err, errlen, "bad transfer info file [%s]", infofile);
    return -1;
  }
  double *tmp = malloc(sizeof(double) * 5 * NMAX_ROWS);
  double *lnk = NULL, *T[4] = {NULL, NULL, NULL, NULL}, *zs = malloc(sizeof(double) * nz);
  int nk = 0, rc = 0;
  for (int i = 0; i < nz && !rc; i++) {
    char name[400], path[820];
    if (fscanf(fp, "%399s %lf", name, &zs[i]) != 2) {
      snprintf(err, errlen, "bad transfer info file [%s]", infofile);
      rc = -1;
      break;
    }
    snprintf(path, sizeof path, "%s/%s", dir, name);
    int npts = 0;
    if (read_transfer_file(path, tmp, tmp + NMAX_ROWS, tmp + 2 * NMAX_ROWS, tmp + 3 * NMAX_ROWS,
                           tmp + 4 * NMAX_ROWS, &npts) || npts < 2) {
      snprintf(err, errlen, "cannot read the transfer file [%s]", path);
      rc = -1;
      break;
    }
    if (i == 0) {
      nk = npts;
      lnk = malloc(sizeof(double) * nk);
      for (int f = 0; f < 4; f++) T[f] = malloc(sizeof(double) * (size_t)nz * nk);
      memcpy(lnk, tmp, sizeof(double) * nk);
    } else if (npts != nk) {
      /* the reference assumes one k grid for all redshifts */
      snprintf(err, errlen, "the k grid of [%s] differs from the first file", path);
      rc = -1;
      break;
    }
    /* T[0] = nu, T[1] = cdm, T[2] = baryon, T[3] = total, row i = redshift */
    for (int j = 0; j < nk; j++) {
      T[0][(size_t)i * nk + j] = tmp[3 * NMAX_ROWS + j];
      T[1][(size_t)i * nk + j] = tmp[NMAX_ROWS + j];
      T[2][(size_t)i * nk + j] = tmp[2 * NMAX_ROWS + j];
      T[3][(size_t)i * nk + j] = tmp[4 * NMAX_ROWS + j];
    }
  }
  fclose(fp);
  if (!rc)
    for (int f = 0; f < 4; f++) rc |= mgp_spline2d_init(&c->camb[f], lnk, zs, T[f], nk, nz);
  if (!rc) c->have_camb = 1;
  free(tmp);
  free(zs);
  free(lnk);
  for (int f = 0; f < 4; f++) free(T[f]);
  return rc;
}

/* get_*_transfer_function (src/read_CAMB_data.h:165-211) at z = 1/a - 1 */
double mgp_T_nu(const mgp_cosmo *c, double k, double a) { return mgp_spline2d_eval(&c->camb[0], log(k), 1.0 / a - 1.0); }
double mgp_T_cdm(const mgp_cosmo *c, double k, double a) { return mgp_spline2d_eval(&c->camb[1], log(k), 1.0 / a - 1.0); }
double mgp_T_b(const mgp_cosmo *c, double k, double a) { return mgp_spline2d_eval(&c->camb[2], log(k), 1.0 / a - 1.0); }
double mgp_T_tot(const mgp_cosmo *c, double k, double a) { return mgp_spline2d_eval(&c->camb[3], log(k), 1.0 / a - 1.0); }
/* CDM + baryons weighted by their densities (the species the particles carry) */
double mgp_T_cb(const mgp_cosmo *c, double k, double a) {
  const mgp_params *p = c->p;
  return (p->OmegaBaryon * mgp_T_b(c, k, a) + p->OmegaCDM * mgp_T_cdm(c, k, a)) / (p->OmegaBaryon + p->OmegaCDM);
}

/* get_total_power_spectrum / get_cdm_baryon_power_spectrum (src/read_CAMB_data.h:217-233), P(k)/A_s */
double mgp_pk_total(const mgp_cosmo *c, double k) {
  const mgp_params *p = c->p;
  const double h = p->HubbleParam, T = mgp_T_tot(c, k, 1.0);
  return 2.0 * MGP_PI * MGP_PI * T * T * (k * h) * pow(k * h / 0.05, p->PrimordialIndex - 1.0) * pow(h, 3);
}
double mgp_pk_cb(const mgp_cosmo *c, double k) {
  const double r = mgp_T_cb(c, k, 1.0) / mgp_T_tot(c, k, 1.0);
  return mgp_pk_total(c, k) * r * r;
}

/* GeffoverG_neutrino_1LPT / _2LPT (src/user_defined_functions.h:662-680) */
double mgp_munu_1lpt(const mgp_cosmo *c, double a, double k) {
  const mgp_params *p = c->p;
  if (!c->nu || p->OmegaNu <= 1e-6) return 1.0;
  return (p->Omega - p->OmegaNu) / p->Omega + p->OmegaNu / p->Omega * mgp_T_nu(c, k, a) / mgp_T_cb(c, k, a);
}
double mgp_munu_2lpt(const mgp_cosmo *c, double a, double k) {
  const mgp_params *p = c->p;
  if (!c->nu || p->OmegaNu <= 1e-6) return 1.0;
  (void)a;
  (void)k;
  return (p->Omega - p->OmegaNu) / p->Omega;
}

/* The neutrino part of PtoMesh (src/auxPM.c:383-420) at a: the factor of the IC spectrum,
 * OmegaNu/Omega N^3 T_nu(k, a)/T_cb(k, 1), as a slice on the uniform ln k grid of the
 * scale-dependent tables (values and d/d ln k by central differences of the spline) */
int mgp_nu_slice(const mgp_cosmo *c, double a, mgpu_kslice *o, double *buf) {
  const mgp_params *p = c->p;
  if (!c->nu || !c->have_sd) return -1;
  const int ny = c->sd[0].ny;
  const double fac = p->OmegaNu / p->Omega * pow((double)p->Nmesh, 3);
  double *v = buf, *dv = buf + ny, *den = buf + 2 * ny, *dden = buf + 3 * ny;
  for (int j = 0; j < ny; j++) {
    const double y = c->sd_lnkmin + c->sd_dlnk * j, h = 1e-4 * c->sd_dlnk;
    const double f0 = fac * mgp_T_nu(c, exp(y), a) / mgp_T_cb(c, exp(y), 1.0);
    const double fp = fac * mgp_T_nu(c, exp(y + h), a) / mgp_T_cb(c, exp(y + h), 1.0);
    const double fm = fac * mgp_T_nu(c, exp(y - h), a) / mgp_T_cb(c, exp(y - h), 1.0);
    v[j] = f0;
    dv[j] = (fp - fm) / (2.0 * h);
    den[j] = 1.0;
    dden[j] = 0.0;
  }
  *o = (mgpu_kslice){c->sd_lnkmin, c->sd_dlnk, ny, v, dv, den, dden, 1.0};
  return 0;
}
