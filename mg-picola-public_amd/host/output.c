/* Output formats of the reference: the P(k) text file
 * (src/compute_pofk.c:239-261), the GADGET-1 snapshot (src/main.c:884-997,
 * header src/main.c:1106-1149 / src/vars.h:204-224) and the info file
 * (src/main.c:1067-1099). */
#include "internal.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* k_from_index, src/compute_pofk.c:50-64 */
static double k_from_index(const mgp_params *p, int index, double kmin, double kmax, int nbins, int bintype) {
  if (bintype == 0) return (kmin + (kmax - kmin) / (double)nbins * index) * 2.0 * M_PI / p->Box;
  return exp(log(kmin) + log(kmax / kmin) / (double)nbins * index) * 2.0 * M_PI / p->Box;
}

int mgp_write_pofk(const mgp_params *p, double a, const char *label, const double *pofk, const double *kmean,
                   const double *nmodes, int nbins, int bintype, double kmin, double kmax) {
  double znow = 1.0 / a - 1.0;
  int zint = (int)znow, zfrac = (int)((znow - zint) * 1000);
  char fn[1200];
  snprintf(fn, sizeof fn, "%s/pofk_%s_z%d.%03d_%s.txt", p->OutputDir, p->FileBase, zint, zfrac, label);
  FILE *fp = fopen(fn, "w");
  if (!fp) return -1;
  fprintf(fp, "#  k_bin (h/Mpc)        P(k) (Mpc/h)^3     k_mean_bin (h/Mpc)     Delta = k^3P(k)/2pi^2\n");
  for (int i = 1; i < nbins; i++) {
    if (nmodes[i] > 0) {
      double kb = k_from_index(p, i, kmin, kmax, nbins, bintype);
      double delta = pofk[i] * kb * kb * kb / 2.0 / M_PI / M_PI;
      fprintf(fp, "%10.5f   %10.5f   %10.5f   %10.5f\n", kb, pofk[i], kmean[i], delta);
    }
  }
  fclose(fp);
  return 0;
}

/* compute_RSD_powerspectrum's file (src/compute_pofk.c:457-485): the multipoles of the y
 * and z lines of sight averaged, |y - z|/sqrt(2) as the error, for bins with modes and k > 0.
 * Each of y and z is {P0, P2, P4, k, n}. */
int mgp_write_rsd_pofk(const mgp_params *p, double a, const double *const y[5], const double *const z[5], int nbins) {
  double znow = 1.0 / a - 1.0;
  int zint = (int)znow, zfrac = (int)((znow - zint) * 1000);
  char fn[1200];
  snprintf(fn, sizeof fn, "%s/pofk_RSD_%s_z%d.%03d.txt", p->OutputDir, p->FileBase, zint, zfrac);
  FILE *fp = fopen(fn, "w");
  if (!fp) return -1;
  fprintf(fp, "#  k (h/Mpc)      P0 (Mpc/h)^3      P2 (Mpc/h)^3      P4 (Mpc/h)^3     sigma0     sigma2     sigma4\n");
  for (int i = 0; i < nbins; i++) {
    if (y[4][i] > 0 && y[3][i] > 0.0) {
      double P[3], e[3];
      for (int l = 0; l < 3; l++) {
        P[l] = (y[l][i] + z[l][i]) / 2.0;
        e[l] = fabs(y[l][i] - z[l][i]) / sqrt(2.0);
      }
      fprintf(fp, "%10.5f   %10.5f   %10.5f   %10.5f   %10.5f   %10.5f   %10.5f\n", y[3][i], P[0], P[1], P[2], e[0],
              e[1], e[2]);
    }
  }
  fclose(fp);
  return 0;
}

#pragma pack(push, 4)
typedef struct {
  uint32_t npart[6];
  double mass[6];
  double time;
  double redshift;
  int32_t flag_sfr;
  int32_t flag_feedback;
  uint32_t npartTotal[6];
  int32_t flag_cooling;
  int32_t num_files;
  double BoxSize;
  double Omega0;
  double OmegaLambda;
  double HubbleParam;
  int32_t flag_stellarage;
  int32_t flag_metals;
  uint32_t npartTotalHighWord[6];
  int32_t flag_entropy_instead_u;
  char fill[60];
} gadget_header;
#pragma pack(pop)

static int fw(const void *b, size_t sz, size_t n, FILE *fp) { return fwrite(b, sz, n, fp) == n ? 0 : -1; }

/* Output (GADGET_STYLE branch): pos in Mpc/h, vel block already in km/s */
int mgp_write_gadget(const mgp_params *p, double A, double z_label, const float *pos, const float *vel, size_t np) {
  _Static_assert(sizeof(gadget_header) == 256, "GADGET header must be 256 bytes");
  double Z = z_label;
  int zint, zfrac;
  if (Z < 0.0) zint = zfrac = 0;
  else {
    zint = (int)floor(Z);
    zfrac = (int)((Z - zint) * 1000);
  }
  char fn[1200];
  snprintf(fn, sizeof fn, "%s/%s_z%dp%03d.%d", p->OutputDir, p->FileBase, zint, zfrac, 0);
  FILE *fp = fopen(fn, "w");
  if (!fp) return -1;
  const double UnitTime_in_s = p->UnitLength_in_cm / p->UnitVelocity_in_cm_per_s;
  const double G = MGP_GRAVITY / pow(p->UnitLength_in_cm, 3) * p->UnitMass_in_g * pow(UnitTime_in_s, 2);
  const double Hubble = MGP_HUBBLE * UnitTime_in_s;
  const double lengthfac = p->UnitLength_in_cm / 3.085678e24;
  const unsigned long long tot = (unsigned long long)p->Nsample * p->Nsample * p->Nsample;
  gadget_header h;
  memset(&h, 0, sizeof h);
  h.npart[1] = (uint32_t)np;
  h.npartTotal[1] = (uint32_t)tot;
  h.npartTotalHighWord[1] = (uint32_t)(tot >> 32);
  h.mass[1] = (3.0 * p->Omega * Hubble * Hubble * p->Box * p->Box * p->Box) / (8.0 * MGP_PI * G * tot);
  h.time = A;
  h.redshift = 1.0 / A - 1.0;
  h.num_files = 1;
  h.BoxSize = p->Box * lengthfac;
  h.Omega0 = p->Omega;
  h.OmegaLambda = 1.0 - p->Omega;
  h.HubbleParam = p->HubbleParam;
  int dummy = sizeof(h), rc = 0;
  rc |= fw(&dummy, sizeof dummy, 1, fp) | fw(&h, sizeof h, 1, fp) | fw(&dummy, sizeof dummy, 1, fp);
  float *blk = malloc(sizeof(float) * 3 * np);
  for (size_t n = 0; n < np; n++)
    for (int k = 0; k < 3; k++) blk[3 * n + k] = (float)(lengthfac * pos[k * np + n]);
  dummy = (int)(sizeof(float) * 3 * np);
  rc |= fw(&dummy, sizeof dummy, 1, fp) | fw(blk, sizeof(float), 3 * np, fp) | fw(&dummy, sizeof dummy, 1, fp);
  for (size_t n = 0; n < np; n++)
    for (int k = 0; k < 3; k++) blk[3 * n + k] = vel[k * np + n];
  rc |= fw(&dummy, sizeof dummy, 1, fp) | fw(blk, sizeof(float), 3 * np, fp) | fw(&dummy, sizeof dummy, 1, fp);
  free(blk);
  fclose(fp);
  return rc;
}

/* Output, ASCII branch (builds without GADGET_STYLE), src/main.c:999-1027: the
 * particle count on the first line, then x y z [Mpc/h] vx vy vz [km/s] per
 * particle; the same file name as the GADGET file; vel already in km/s. */
int mgp_write_ascii(const mgp_params *p, double A, double z_label, const float *pos, const float *vel, size_t np) {
  double Z = z_label;
  int zint, zfrac;
  if (Z < 0.0) zint = zfrac = 0;
  else {
    zint = (int)floor(Z);
    zfrac = (int)((Z - zint) * 1000);
  }
  char fn[1200];
  snprintf(fn, sizeof fn, "%s/%s_z%dp%03d.%d", p->OutputDir, p->FileBase, zint, zfrac, 0);
  FILE *fp = fopen(fn, "w");
  if (!fp) return -1;
  const double lengthfac = p->UnitLength_in_cm / 3.085678e24;
  int rc = fprintf(fp, "%u\n", (unsigned)np) < 0;
  for (size_t n = 0; n < np && !rc; n++)
    rc = fprintf(fp, "%12.6f %12.6f %12.6f %12.6f %12.6f %12.6f\n", (float)(lengthfac * pos[n]),
                 (float)(lengthfac * pos[np + n]), (float)(lengthfac * pos[2 * np + n]), vel[n], vel[np + n],
                 vel[2 * np + n]) < 0;
  if (fclose(fp)) rc = 1;
  return rc ? -1 : 0;
}

/* Output_Info, src/main.c:1067-1099 (single task) */
int mgp_write_info(const mgp_params *p, double A, size_t np) {
  double Z = 1.0 / A - 1.0;
  char fn[1200];
  snprintf(fn, sizeof fn, "%s/info_%s_z%dp%03d.txt", p->OutputDir, p->FileBase, (int)Z, (int)rint((Z - (int)Z) * 1000));
  FILE *fp = fopen(fn, "w");
  if (!fp) return -1;
  fprintf(fp, "#    FILENUM      XMIN         YMIN        ZMIN         XMAX         YMAX         ZMAX         NPART    \n");
  fprintf(fp, "%12d %12.6lf %12.6lf %12.6lf %12.6lf %12.6lf %12.6lf %12u\n", 0, 0.0, 0.0, 0.0, p->Box, p->Box, p->Box,
          (unsigned)np);
  fclose(fp);
  return 0;
}
