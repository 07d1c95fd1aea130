/* Internal state of the C host. */
#ifndef MGP_INTERNAL_H
#define MGP_INTERNAL_H

#include "mgpicola.h"
#include "numerics.h"

struct mgp_cosmo {
  mgp_params *p;
  /* model flags derived from the build configuration */
  int fofr, dgp, jbd, mbeta;
  /* growth factors on x = ln a (src/cosmo.c:409-643) */
  int ngrowth;
  mgp_spline g[6], glcdm[6];
  /* scale-dependent tables on (ln a, ln k) (src/cosmo.c:798-981) */
  mgp_spline2d sd[6];
  int have_sd;
  double sd_lnkmin, sd_dlnk; /* uniform ln k grid */
  /* JBD background (src/jbd.c) */
  mgp_spline jbd_H, jbd_dHda, jbd_Geff;
  /* m-beta screening threshold phi(a) (src/cosmo.c:118-194) */
  mgp_spline phi_of_a;
  /* massive neutrinos: CAMB transfer functions nu, cdm, baryon, total on (ln k, z) and the
   * scale-dependent LCDM growth table that then replaces the scale-independent one */
  int nu, have_camb, have_sd_lcdm;
  mgp_spline2d camb[4];
  mgp_spline2d sd_lcdm[6];
  /* power spectrum (src/power.c) */
  double Norm;
  int ntab;
  double *tab_logk, *tab_logp; /* tabulated P(k) or T(k) in log10 */
  double R8;
};

/* models.c */
double mgp_hub(const struct mgp_cosmo *c, double a);
double mgp_dhub(const struct mgp_cosmo *c, double a);
double mgp_Factor_2LPT(const struct mgp_cosmo *c, double a);
double mgp_second_order_kernel(const struct mgp_cosmo *c, double k, double k1, double k2, double cost, double a);
int mgp_jbd_solve(struct mgp_cosmo *c, double *h);
int mgp_compute_phi_of_a(struct mgp_cosmo *c);

/* cosmology.c */
int mgp_solve_growth(struct mgp_cosmo *c);
/* camb.c */
int mgp_camb_read(struct mgp_cosmo *c, const char *infofile, char *err, size_t errlen);
/* power.c */
int mgp_init_power(struct mgp_cosmo *c, char *err, size_t errlen);

#define MGP_REDSHIFT_START(p) ((p)->Init_Redshift > 200.0 ? (p)->Init_Redshift : 200.0)
#define MGP_REDSHIFT_END (-0.5)

#endif
