/* Model hooks of MG-PICOLA (src/user_defined_functions.h:413-904), selected at
 * run time from the build configuration instead of #if defined(...), plus the
 * JBD background solver (src/jbd.c) and the m-beta phi(a) integral
 * (src/cosmo.c:93-194). */
#include "internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

static inline double pow2(double x) { return x * x; }
static inline double pow3(double x) { return x * x * x; }

/* hubble / dhubbleda (src/user_defined_functions.h:413-447) */
double mgp_hub(const struct mgp_cosmo *c, double a) {
  if (c->jbd) return mgp_spline_eval(&c->jbd_H, log(a));
  const double Om = c->p->Omega;
  return sqrt(Om / (a * a * a) + 1.0 - Om);
}
double mgp_dhub(const struct mgp_cosmo *c, double a) {
  if (c->jbd) return mgp_spline_eval(&c->jbd_dHda, log(a));
  const double Om = c->p->Omega;
  return 1.0 / (2.0 * mgp_hub(c, a)) * (-3.0 * Om / (a * a * a * a));
}
double mgp_hubble(const mgp_cosmo *c, double a) { return mgp_hub(c, a); }
double mgp_dhubbleda(const mgp_cosmo *c, double a) { return mgp_dhub(c, a); }

/* beta_DGP (src/user_defined_functions.h:453-456) */
double mgp_beta_dgp(const mgp_cosmo *c, double a) {
  return 1.0 + 2.0 * c->p->rcH0_DGP * (mgp_hub(c, a) + a * mgp_dhub(c, a) / 3.0);
}

/* beta_of_a (src/user_defined_functions.h:462-484) */
double mgp_beta(const mgp_cosmo *c, double a) {
  if (c->fofr) return 1.0 / sqrt(6.0);
  if (c->mbeta) {
    if (a > c->p->assb_symm) return c->p->beta_symm * sqrt(1.0 - pow3(c->p->assb_symm / a));
    return 0.0;
  }
  return 0.0;
}

/* mass2_of_a = m^2(a)/H0^2 (src/user_defined_functions.h:490-520) */
double mgp_mass2(const mgp_cosmo *c, double a) {
  const mgp_params *p = c->p;
  if (c->fofr) {
    double a3 = a * a * a;
    double fac = p->Omega / a3 + 4.0 * (1.0 - p->Omega);
    double fac0 = p->Omega + 4.0 * (1.0 - p->Omega);
    return fac0 * pow(fac / fac0, p->nfofr + 2.0) / ((1.0 + p->nfofr) * p->fofr0);
  }
  if (c->mbeta) {
    if (a > p->assb_symm) return 0.5 * pow2(2998.0 / p->range_symm) * (1.0 - pow3(p->assb_symm / a));
    return 1.0;
  }
  return 0.0;
}

/* dmass2_of_ada (src/user_defined_functions.h:527-560) */
static double dmass2_da(const mgp_cosmo *c, double a) {
  const mgp_params *p = c->p;
  if (c->fofr) {
    double a3 = a * a * a;
    double fac = p->Omega / a3 + 4.0 * (1.0 - p->Omega);
    double fac0 = p->Omega + 4.0 * (1.0 - p->Omega);
    double m = fac0 * (p->nfofr + 2.0) * pow(fac / fac0, p->nfofr + 1.0) / ((1.0 + p->nfofr) * p->fofr0 * fac0);
    return m * (-3.0 * p->Omega / a3 / a);
  }
  if (c->mbeta) {
    if (a > p->assb_symm) return 0.5 * pow2(2998.0 / p->range_symm) * (+3.0 * pow3(p->assb_symm / a) / a);
    return 0.0;
  }
  return 0.0;
}

/* coupling_function (src/user_defined_functions.h:587-612) */
double mgp_coupling(const mgp_cosmo *c, double a) {
  if (c->fofr || c->mbeta) return 2.0 * mgp_beta(c, a) * mgp_beta(c, a);
  if (c->dgp) return 1.0 / (3.0 * mgp_beta_dgp(c, a));
  if (c->jbd) return mgp_spline_eval(&c->jbd_Geff, log(a)) - 1.0;
  return 0.0;
}

/* GeffoverG = mu(a,k) (src/user_defined_functions.h:618-652) */
double mgp_GeffoverG(const mgp_cosmo *c, double a, double k) {
  double mu = 1.0;
  if (!c->p->modified_gravity_active) return mu;
  if (c->p->use_lcdm_growth_factors) return mu;
  if (c->fofr || c->mbeta) {
    double mass2a2 = a * a * mgp_mass2(c, a);
    double k2 = pow2(k * MGP_INVERSE_H0_MPCH);
    if (k2 == 0) return mu;
    return mu * (1.0 + 2.0 * mgp_beta(c, a) * mgp_beta(c, a) * k2 / (k2 + mass2a2));
  }
  if (c->dgp) return mu * (1.0 + 1.0 / (3.0 * mgp_beta_dgp(c, a)));
  if (c->jbd) return mu * (1.0 + mgp_coupling(c, a));
  return mu;
}

/* phi_crit(a) of screening_factor_potential (src/user_defined_functions.h:725-752) */
double mgp_phi_crit(const mgp_cosmo *c, double a) {
  const mgp_params *p = c->p;
  if (c->fofr)
    return 1.5 * p->fofr0 *
           pow((p->Omega + 4.0 * (1.0 - p->Omega)) / (1.0 / (a * a * a) * p->Omega + 4.0 * (1.0 - p->Omega)),
               p->nfofr + 1.0);
  if (c->mbeta) return mgp_spline_eval(&c->phi_of_a, log(a));
  return 0.0;
}

double mgp_screening_potential(const mgp_cosmo *c, double a, double phin) {
  if (!c->p->include_screening) return 1.0;
  if (phin >= 0.0) return 1.0;
  if (c->fofr || c->mbeta) {
    double s = fabs(mgp_phi_crit(c, a) / phin);
    return s > 1.0 ? 1.0 : s;
  }
  return 1.0;
}

/* screening_factor_density (src/user_defined_functions.h:757-773) */
static double dgp_fac(const mgp_cosmo *c, double a) {
  return 8.0 / 9.0 * c->p->Omega * pow(c->p->rcH0_DGP / mgp_beta_dgp(c, a), 2);
}
double mgp_screening_density(const mgp_cosmo *c, double a, double dens) {
  if (!c->p->include_screening) return 1.0;
  if (!c->dgp) return 1.0;
  double fac = dgp_fac(c, a) * (1.0 + dens);
  if (fac < 1e-5) return 1.0;
  return 2.0 * (sqrt(1.0 + fac) - 1.0) / fac;
}

/* Factor_2LPT (src/user_defined_functions.h:814-845) */
double mgp_Factor_2LPT(const struct mgp_cosmo *c, double a) {
  if (!c->p->modified_gravity_active || c->p->use_lcdm_growth_factors) return 1.0;
  if (c->dgp) {
    double b = mgp_beta_dgp(c, a);
    return 1.0 - (2.0 * c->p->rcH0_DGP * c->p->rcH0_DGP * c->p->Omega) /
                     (9.0 * a * a * a * pow3(b) * (1 + 1.0 / (3.0 * b)));
  }
  return 1.0;
}

/* second_order_kernel gamma_2 (src/user_defined_functions.h:862-904) */
static double fofr_pi_factor(const mgp_cosmo *c, double k, double a) {
  const mgp_params *p = c->p;
  double a3 = a * a * a;
  double fac = (p->Omega / a3 + 4.0 * (1.0 - p->Omega));
  double fac0 = (p->Omega + 4.0 * (1.0 - p->Omega));
  return pow2(k * MGP_INVERSE_H0_MPCH / a) + fac0 * pow(fac / fac0, p->nfofr + 2.0) / (1.0 + p->nfofr) / p->fofr0;
}

double mgp_second_order_kernel(const struct mgp_cosmo *c, double k, double k1, double k2, double cost, double a) {
  const mgp_params *p = c->p;
  if (!p->modified_gravity_active || p->use_lcdm_growth_factors) return 0.0;
  double gamma2 = 0.0;
  if (c->fofr) {
    double a3 = a * a * a;
    double fac = p->Omega / a3 + 4.0 * (1.0 - p->Omega);
    double fac0 = p->Omega + 4.0 * (1.0 - p->Omega);
    gamma2 = -(3.0 * (p->nfofr + 2.0) / (12.0 * (1.0 + p->nfofr) * (1.0 + p->nfofr))) *
             pow2(k * MGP_INVERSE_H0_MPCH / (a * mgp_hub(c, a))) * pow2(p->Omega / a3) * fac0 *
             pow(fac / fac0, 2.0 * p->nfofr + 3.0) / pow2(p->fofr0);
    gamma2 /= fofr_pi_factor(c, k, a) * fofr_pi_factor(c, k1, a) * fofr_pi_factor(c, k2, a);
  } else if (c->dgp) {
    gamma2 = -1.0 / 6.0 / pow3(mgp_beta_dgp(c, a)) * pow2(p->rcH0_DGP / mgp_hub(c, a)) *
             pow2(p->Omega / (a * a * a)) * (1.0 - cost * cost);
  } else if (c->mbeta) {
    double m2 = mgp_mass2(c, a), dm2da = dmass2_da(c, a);
    double pk = pow2(k * MGP_INVERSE_H0_MPCH / a) + m2;
    double pk1 = pow2(k1 * MGP_INVERSE_H0_MPCH / a) + m2;
    double pk2 = pow2(k2 * MGP_INVERSE_H0_MPCH / a) + m2;
    gamma2 = (dm2da / m2) * pow2(mgp_beta(c, a)) * p->Omega / (2.0 * a * a) * (m2 / pk1) * (m2 / pk2) *
             (1.0 - m2 / pk) / pow2(mgp_hub(c, a));
  }
  if (k2 > 1e-3 && k1 > 1e-3) {
    gamma2 += 1.5 * p->Omega / (a * a * a * mgp_hub(c, a) * mgp_hub(c, a)) *
              ((mgp_GeffoverG(c, a, k) - mgp_GeffoverG(c, a, k1)) * (k1 / k2) +
               (mgp_GeffoverG(c, a, k) - mgp_GeffoverG(c, a, k2)) * (k2 / k1)) *
              cost / 2.0;
  }
  return gamma2;
}

void mgp_mg_coeffs(const mgp_cosmo *c, double a, mgpu_mg *m) {
  m->a = a;
  m->coupling = mgp_coupling(c, a);
  m->mass2 = mgp_mass2(c, a);
  m->phi_crit = mgp_phi_crit(c, a);
  m->dgp_fac = c->dgp ? dgp_fac(c, a) : 0.0;
  m->rsmooth = c->p->Rsmooth > 0 ? c->p->Rsmooth : 1.0;
  m->geff = mgp_GeffoverG(c, a, 0.0);
}

/* ---------------------------------------------------------------------------
 * JBD background (src/jbd.c): y = ln(phi/phi0) against x = ln a from z = 1e6,
 * phi_i found by bisection so that phi(a=1) = phi0.
 * ------------------------------------------------------------------------- */
typedef struct { double Om, Or, Ov, w; } jbd_par;

static double jbd_H(const jbd_par *g, double x, double y, double dy) {
  return sqrt(exp(-y) * (g->Or * exp(-4 * x) + g->Om * exp(-3 * x) + g->Ov) / (1.0 + dy - g->w / 6 * dy * dy));
}
static double jbd_dHdx(const jbd_par *g, double x, double y, double dy) {
  double H = jbd_H(g, x, y, dy);
  double om = g->Om * exp(-y - 3.0 * x) / (H * H), ov = g->Ov * exp(-y) / (H * H);
  double ora = g->Or * exp(-y - 4.0 * x) / (H * H);
  return H / 2.0 * (-3.0 - ora - g->w / 2.0 * dy * dy + dy - 3.0 * (om + 4.0 * ov) / (3.0 + 2.0 * g->w) + 3.0 * ov);
}
static double jbd_H_of_z(const jbd_par *g, double x, double y, double z) {
  return sqrt(exp(-y) * (g->Or * exp(-4 * x) + g->Om * exp(-3 * x) + g->Ov) +
              z * z / 12.0 * (2.0 * g->w + 3.0) * exp(-2 * y - 6 * x)) -
         z / 2.0 * exp(-y - 3 * x);
}
static int jbd_rhs(double x, const double *y, double *dydx, void *par) {
  const jbd_par *g = (const jbd_par *)par;
  double h = jbd_H_of_z(g, x, y[0], y[1]);
  dydx[0] = y[1] * exp(-y[0] - 3 * x) / h;
  dydx[1] = 3.0 * exp(3.0 * x) / (2 * g->w + 3) / h * (g->Om * exp(-3 * x) + 4.0 * g->Ov);
  return 0;
}

int mgp_jbd_solve(struct mgp_cosmo *c, double *hout) {
  mgp_params *p = c->p;
  jbd_par g = {p->Omegah2, p->Omegarh2, p->Omegavh2, p->wBD};
  const int npts = 10000;
  const double xini = log(1.0 / (1.0 + 1e6)), xend = 0.0, dx = (xend - xini) / (npts - 1);
  double *xa = malloc(sizeof(double) * npts), *ya = malloc(sizeof(double) * npts), *dya = malloc(sizeof(double) * npts);
  double lo = 0.0, hi = 1.0;
  for (int it = 0; it < 200; it++) {
    double phinow = 0.5 * (lo + hi), yi = log(phinow), dyi = 0.0;
    double y[2] = {yi, jbd_H(&g, xini, yi, dyi) * exp(3.0 * xini + yi) * dyi};
    double x = xini, h = 1e-8;
    xa[0] = xini;
    ya[0] = y[0];
    dya[0] = y[1];
    for (int i = 1; i < npts; i++) {
      double xn = xini + i * dx;
      if (mgp_ode_integrate(jbd_rhs, &g, 2, &x, xn, y, &h, 1e-10, 1e-10)) goto fail;
      xa[i] = xn;
      ya[i] = y[0];
      dya[i] = y[1] / jbd_H_of_z(&g, xn, y[0], y[1]) * exp(-3.0 * xn - y[0]);
    }
    double ratio = exp(ya[npts - 1]);
    if (fabs(ratio - 1.0) < 1e-6) break;
    if (ratio < 1.0) lo = phinow;
    else hi = phinow;
  }
  double hnow = jbd_H(&g, 0.0, ya[npts - 1], dya[npts - 1]);
  *hout = hnow;
  double *H = malloc(sizeof(double) * npts), *dH = malloc(sizeof(double) * npts), *G = malloc(sizeof(double) * npts);
  for (int i = 0; i < npts; i++) {
    G[i] = exp(-ya[i]);
    H[i] = jbd_H(&g, xa[i], ya[i], dya[i]) / hnow;
    dH[i] = jbd_dHdx(&g, xa[i], ya[i], dya[i]) / hnow * exp(-xa[i]);
  }
  mgp_spline_init(&c->jbd_Geff, xa, G, npts);
  mgp_spline_init(&c->jbd_H, xa, H, npts);
  mgp_spline_init(&c->jbd_dHda, xa, dH, npts);
  free(H);
  free(dH);
  free(G);
  free(xa);
  free(ya);
  free(dya);
  return 0;
fail:
  free(xa);
  free(ya);
  free(dya);
  return -1;
}

/* phi(a) = 9 Om / (2 beta(a)) int beta/(m^2 a^3) dln a (src/cosmo.c:93-194) */
static double phi_integrand(double x, void *par) {
  const struct mgp_cosmo *c = (const struct mgp_cosmo *)par;
  double a = exp(x);
  double beta = mgp_beta(c, a), m2 = mgp_mass2(c, a);
  if (beta == 0.0) return 0.0;
  return 9.0 * c->p->Omega * beta / (m2 * a * a * a);
}

int mgp_compute_phi_of_a(struct mgp_cosmo *c) {
  const int npts = 1000;
  const double amin = 1.0 / (1.0 + MGP_REDSHIFT_START(c->p)), amax = 1.0 / (1.0 + MGP_REDSHIFT_END);
  const double xmin = log(amin), xmax = log(amax), dx = (xmax - xmin) / (npts - 1);
  double *x = malloc(sizeof(double) * npts), *ph = malloc(sizeof(double) * npts);
  x[0] = xmin;
  ph[0] = 0.0;
  for (int i = 1; i < npts; i++) {
    x[i] = xmin + i * dx;
    ph[i] = ph[i - 1] + mgp_integrate(phi_integrand, c, xmin + (i - 1) * dx, x[i], 0.0, 1e-9);
  }
  for (int i = 0; i < npts; i++) {
    double b = mgp_beta(c, exp(x[i]));
    if (b != 0.0) ph[i] /= 2.0 * b;
  }
  mgp_spline_init(&c->phi_of_a, x, ph, npts);
  free(x);
  free(ph);
  return 0;
}
