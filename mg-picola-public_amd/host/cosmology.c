/* Growth factors, COLA time integrals and the cosmology object of the host.
 * Restates src/cosmo.c (growth ODEs 206-312, splines 318-403 and 409-643,
 * scale-dependent tables 798-1010, Sq/Sphi 1114-1160, LCDM fits 1291-1398). */
#include "internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline double pow2(double x) { return x * x; }

double mgp_Qfactor(const struct mgp_cosmo *c, double a) { return mgp_hub(c, a) * a * a * a; }

/* ode_growth_D / ode_growth_DLCDM (src/cosmo.c:206-261): y = D, dD/dx, D2, dD2/dx */
typedef struct {
  const struct mgp_cosmo *c;
  double k;
  int lcdm;
} growth_par;

static int growth_rhs(double x, const double *D, double *dD, void *vp) {
  const growth_par *gp = (const growth_par *)vp;
  const struct mgp_cosmo *c = gp->c;
  const double a = exp(x), k = gp->k, Om = c->p->Omega;
  const double H = mgp_hub(c, a), dH = mgp_dhub(c, a);
  const double alphafac = 2.0 + a * dH / H;
  double mu = 1.0, modfac = 1.0;
  if (!gp->lcdm) {
    mu = mgp_GeffoverG(c, a, k);
    if (c->p->scale_dependent)
      modfac = 1.0 + 2.0 * pow2(a * a * H) * mgp_second_order_kernel(c, k, k / sqrt(2), k / sqrt(2), 0.0, a) /
                         (1.5 * Om * a * mu);
    else
      modfac = mgp_Factor_2LPT(c, a);
  }
  const double betafac = 1.5 * Om * mu / (a * a * a * H * H);
  /* massive neutrinos: effective mu of the CDM+baryon growth (src/cosmo.c:224-236, :253-265) */
  const double nu1 = mgp_munu_1lpt(c, a, k), nu2 = mgp_munu_2lpt(c, a, k);
  dD[0] = D[1];
  dD[1] = -alphafac * D[1] + betafac * nu1 * D[0];
  dD[2] = D[3];
  dD[3] = -alphafac * D[3] + betafac * nu2 * (D[2] - modfac * D[0] * D[0]);
  return 0;
}

#define ODE_ABS 1e-11
#define ODE_REL 1e-11

static mgpu_spline2d spline_view(const mgp_spline2d *s) {
  mgpu_spline2d v = {s->nx, s->ny, s->x, s->y, s->z, s->zx, s->zy, s->zxy};
  return v;
}

/* The scale-dependent table of every k on the GPU (mgpu_growth_tables): the same ODEs,
 * integrator and tolerances as the host loop below; z[f] gets nk * npts values, k-major. */
static int growth_table_gpu(const struct mgp_cosmo *c, int lcdm, double xini, double dx, int npts,
                            const double *logk, int nk, double *z[6]) {
  const mgp_params *p = c->p;
  mgpu_growth_model m;
  memset(&m, 0, sizeof m);
  m.model = c->fofr ? MGPU_FOFR : c->mbeta ? MGPU_MBETA : MGPU_LCDM;
  m.mg_active = p->modified_gravity_active && !p->use_lcdm_growth_factors;
  m.lcdm = lcdm;
  m.omega = p->Omega;
  m.fofr0 = p->fofr0;
  m.nfofr = p->nfofr;
  m.assb_symm = p->assb_symm;
  m.beta_symm = p->beta_symm;
  m.range_symm = p->range_symm;
  if (c->nu) {
    m.omega_nu = p->OmegaNu;
    m.omega_b = p->OmegaBaryon;
    m.omega_cdm = p->OmegaCDM;
    m.t_nu = spline_view(&c->camb[0]);
    m.t_cdm = spline_view(&c->camb[1]);
    m.t_b = spline_view(&c->camb[2]);
  }
  const size_t plane = (size_t)nk * npts;
  double *out = malloc(sizeof(double) * 6 * plane);
  if (!out) return -1;
  int rc = mgpu_growth_tables(&m, xini, dx, npts, logk, nk, out);
  if (!rc)
    for (int f = 0; f < 6; f++) memcpy(z[f], out + f * plane, sizeof(double) * plane);
  free(out);
  return rc;
}

/* The same table with the host's ODE solver, one OpenMP task per k.  lcdm: the
 * ode_growth_DLCDM equations and source terms without mu and the neutrino factors. */
static int growth_table_host(const struct mgp_cosmo *c, int lcdm, double xini, double dx, int npts,
                             const double *logk, int nk, double *z[6]) {
  const mgp_params *p = c->p;
  int failed = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(| : failed)
  for (int j = 0; j < nk; j++) {
    const double know = exp(logk[j]);
    growth_par g = {c, know, lcdm};
    double y[4] = {1.0, 1.0, -3.0 / 7.0, -6.0 / 7.0};
    double xs = xini, h = 1e-4;
    for (int i = 0; i < npts; i++) {
      const double xn = xini + i * dx, a = exp(xn);
      if (i > 0 && mgp_ode_integrate(growth_rhs, &g, 4, &xs, xn, y, &h, ODE_ABS, ODE_REL)) failed = 1;
      const double Q = mgp_Qfactor(c, a), Om = p->Omega;
      double f1 = 1.0, f2 = 1.0;
      if (!lcdm) {
        const double mu = mgp_GeffoverG(c, a, know);
        f1 = mu * mgp_munu_1lpt(c, a, know);
        f2 = mu * mgp_munu_2lpt(c, a, know);
      }
      const size_t o = (size_t)j * npts + i;
      z[0][o] = y[0];
      z[1][o] = y[1] * Q / a;
      z[2][o] = 1.5 * f1 * Om * a * y[0];
      z[3][o] = y[2];
      z[4][o] = y[3] * Q / a;
      z[5][o] = 1.5 * f2 * Om * a * (y[2] - pow2(y[0]));
    }
  }
  return failed ? -1 : 0;
}

/* solve_for_growth_factors (src/cosmo.c:409-643) and, for scale-dependent
 * builds, calculate_scale_dependent_growth_factor (src/cosmo.c:798-981) */
int mgp_solve_growth(struct mgp_cosmo *c) {
  const mgp_params *p = c->p;
  const int npts = 1000;
  const double zini = MGP_REDSHIFT_START(p), zend = MGP_REDSHIFT_END;
  const double aini = 1.0 / (1.0 + zini), xini = log(aini), xend = log(1.0 / (1.0 + zend));
  const double dx = (xend - xini) / (npts - 1);
  double *x = malloc(sizeof(double) * npts);
  double *arr[12];
  for (int i = 0; i < 12; i++) arr[i] = malloc(sizeof(double) * npts);
  growth_par gp = {c, 1e-3, 0}, gl = {c, 1e-3, 1};
  double D[4] = {1.0, 1.0, -3.0 / 7.0, -6.0 / 7.0}, L[4] = {1.0, 1.0, -3.0 / 7.0, -6.0 / 7.0};
  double xd = xini, xl = xini, hd = 1e-4, hl = 1e-4;
  for (int i = 0; i < npts; i++) {
    const double xn = xini + i * dx, a = exp(xn);
    if (i > 0) {
      if (mgp_ode_integrate(growth_rhs, &gp, 4, &xd, xn, D, &hd, ODE_ABS, ODE_REL)) return -1;
      if (mgp_ode_integrate(growth_rhs, &gl, 4, &xl, xn, L, &hl, ODE_ABS, ODE_REL)) return -1;
    }
    const double mu = mgp_GeffoverG(c, a, 1e-3), Q = mgp_Qfactor(c, a), Om = p->Omega;
    x[i] = xn;
    arr[0][i] = D[0];
    arr[1][i] = D[1] * Q / a;
    arr[2][i] = 1.5 * mu * Om * a * D[0];
    arr[3][i] = D[2];
    arr[4][i] = D[3] * Q / a;
    arr[5][i] = 1.5 * mu * Om * a * (D[2] - pow2(D[0]) * mgp_Factor_2LPT(c, a));
    arr[6][i] = L[0];
    arr[7][i] = L[1] * Q / a;
    arr[8][i] = 1.5 * Om * a * L[0];
    arr[9][i] = L[2];
    arr[10][i] = L[3] * Q / a;
    arr[11][i] = 1.5 * Om * a * (L[2] - pow2(L[0]));
  }
  for (int i = 0; i < 6; i++) {
    mgp_spline_init(&c->g[i], x, arr[i], npts);
    mgp_spline_init(&c->glcdm[i], x, arr[6 + i], npts);
  }
  for (int i = 0; i < 12; i++) free(arr[i]);

  if (p->scale_dependent) {
    const int nk = 1000;
    const double kmin = 2.0 * MGP_PI / p->Box * 0.5;
    const double kmax = 2.0 * MGP_PI / p->Box * p->Nmesh * sqrt(3.0) * 2.0;
    double *logk = malloc(sizeof(double) * nk);
    for (int j = 0; j < nk; j++) logk[j] = log(kmin) + log(kmax / kmin) * j / (double)(nk - 1);
    c->sd_lnkmin = logk[0];
    c->sd_dlnk = log(kmax / kmin) / (double)(nk - 1);
    double *z[6];
    for (int f = 0; f < 6; f++) z[f] = malloc(sizeof(double) * (size_t)npts * nk);
    /* with massive neutrinos LCDM growth is scale-dependent too: its table first
     * (src/cosmo.c:829-889), then the model's */
    int rc = 0;
    for (int lcdm = c->nu ? 1 : 0; lcdm >= 0 && !rc; lcdm--) {
      rc = p->growth_on_gpu ? growth_table_gpu(c, lcdm, xini, dx, npts, logk, nk, z)
                            : growth_table_host(c, lcdm, xini, dx, npts, logk, nk, z);
      for (int f = 0; f < 6 && !rc; f++) mgp_spline2d_init(lcdm ? &c->sd_lcdm[f] : &c->sd[f], x, logk, z[f], npts, nk);
    }
    for (int f = 0; f < 6; f++) free(z[f]);
    free(logk);
    if (rc) {
      free(x);
      return -1;
    }
    c->have_sd_lcdm = c->nu;
    c->have_sd = 1;
  }
  free(x);
  return 0;
}

/* growth_D ... growth_ddD2ddy (src/cosmo.c:318-403): normalised by D or D2 at a = 1 */
static int norm_of(int which) { return which < 3 ? MGP_G_D : MGP_G_D2; }

double mgp_growth(const mgp_cosmo *c, int which, double a) {
  return mgp_spline_eval(&c->g[which], log(a)) / mgp_spline_eval(&c->g[norm_of(which)], 0.0);
}
double mgp_growth_lcdm(const mgp_cosmo *c, int which, double a) {
  return mgp_spline_eval(&c->glcdm[which], log(a)) / mgp_spline_eval(&c->glcdm[norm_of(which)], 0.0);
}
/* growth_X_scaledependent(k, a) (src/cosmo.c:987-1010) */
double mgp_growth_sd(const mgp_cosmo *c, int which, double k, double a) {
  return mgp_spline2d_eval(&c->sd[which], log(a), log(k)) / mgp_spline2d_eval(&c->sd[norm_of(which)], 0.0, log(k));
}

/* LCDM fitting functions of the original PICOLA (src/cosmo.c:1291-1398):
 * growthD (hypergeometric closed form), growthD2, growthD2v, DprimeQ */
static double hyp2f1_series(double a, double b, double cc, double z) {
  /* |z| < 1 power series */
  double term = 1.0, sum = 1.0;
  for (int n = 0; n < 2000; n++) {
    term *= (a + n) * (b + n) / ((cc + n) * (n + 1.0)) * z;
    sum += term;
    if (fabs(term) < 1e-17 * fabs(sum)) break;
  }
  return sum;
}
/* D_+(a) of LCDM: a * 2F1(1/3, 1, 11/6; -a^3 (1-Om)/Om), normalised at a = 1 */
static double dplus(double Om, double a) {
  double z = -a * a * a * (1.0 - Om) / Om;
  if (fabs(z) < 0.9) return a * hyp2f1_series(1.0 / 3.0, 1.0, 11.0 / 6.0, z);
  /* Pfaff transformation for |z| >= 0.9: 2F1(a,b;c;z) = (1-z)^-b 2F1(c-a, b; c; z/(z-1)) */
  return a * pow(1.0 - z, -1.0) * hyp2f1_series(11.0 / 6.0 - 1.0 / 3.0, 1.0, 11.0 / 6.0, z / (z - 1.0));
}
double mgp_growth_lcdmfit(const mgp_cosmo *c, int which, double a) {
  const double Om = c->p->Omega;
  const double D = dplus(Om, a) / dplus(Om, 1.0);
  const double Oma = Om / (Om + (1.0 - Om) * a * a * a);
  const double Oma1 = 1.0;  /* Om(a=1)/Om normalisation handled below */
  (void)Oma1;
  switch (which) {
    case MGP_G_D: return D;
    case MGP_G_D2: {
      double d2 = D * D * pow(Oma, -1.0 / 143.0);
      return d2 / (1.0 * pow(Om / (Om + (1.0 - Om)), -1.0 / 143.0));
    }
    case MGP_G_ddDddy: return 1.5 * Om * D * a;
    default: return NAN;
  }
}

/* COLA time integrals (src/cosmo.c:1114-1160) with nLPT = -2.5, fullT = 1 */
typedef struct { const struct mgp_cosmo *c; int kind; } tpar;
static double t_integrand(double a, void *vp) {
  const tpar *t = (const tpar *)vp;
  double Q = mgp_Qfactor(t->c, a);
  if (t->kind == 0) return pow(a, -2.5) / Q; /* fun(): gpQ(a)/Q(a) */
  if (t->kind == 1) return 1.0 / Q;          /* funSqStd */
  return a / Q;                              /* funSphiStd */
}
double mgp_Sq(const mgp_cosmo *c, double ai, double af, double aRef) {
  tpar t = {c, 0};
  return mgp_integrate(t_integrand, &t, ai, af, 0.0, 1e-12) / pow(aRef, -2.5);
}
double mgp_Sphi(const mgp_cosmo *c, double ai, double af, double aRef) {
  return (pow(af, -2.5) - pow(ai, -2.5)) * aRef / mgp_Qfactor(c, aRef) / (-2.5 * pow(aRef, -3.5));
}
double mgp_SqStd(const mgp_cosmo *c, double ai, double af) {
  tpar t = {c, 1};
  return mgp_integrate(t_integrand, &t, ai, af, 0.0, 1e-12);
}
double mgp_SphiStd(const mgp_cosmo *c, double ai, double af) {
  tpar t = {c, 2};
  return mgp_integrate(t_integrand, &t, ai, af, 0.0, 1e-12);
}

/* mg_pofk_ratio (src/cosmo.c:688-712): raw (unnormalised) growth ratio squared */
double mgp_mg_pofk_ratio(const mgp_cosmo *c, double k, double a) {
  if (!c->p->modified_gravity_active) return 1.0;
  if (c->have_sd_lcdm)
    return pow2(mgp_spline2d_eval(&c->sd[MGP_G_D], log(a), log(k)) /
                mgp_spline2d_eval(&c->sd_lcdm[MGP_G_D], log(a), log(k)));
  double dl = mgp_spline_eval(&c->glcdm[MGP_G_D], log(a));
  if (c->p->scale_dependent) return pow2(mgp_spline2d_eval(&c->sd[MGP_G_D], log(a), log(k)) / dl);
  return pow2(mgp_spline_eval(&c->g[MGP_G_D], log(a)) / dl);
}

/* mg_sigma8_enhancement (src/cosmo.c:722-756) */
double mgp_mg_sigma8_enhancement(const mgp_cosmo *c, double a) {
  if (!c->p->modified_gravity_active) return 1.0;
  if (!c->p->scale_dependent)
    return mgp_spline_eval(&c->g[MGP_G_D], log(a)) / mgp_spline_eval(&c->glcdm[MGP_G_D], log(a));
  const double kmin = 0.001, kmax = 100.0;
  const int npts = 1000;
  double in = 0.0, inl = 0.0;
  const double dlogk = log(kmax / kmin) / (double)(npts - 1);
  for (int i = 0; i < npts; i++) {
    double know = exp(log(kmin) + log(kmax / kmin) * i / (double)(npts - 1));
    double kR8 = know * 8.0;
    double w = 3.0 / (kR8 * kR8 * kR8) * (sin(kR8) - kR8 * cos(kR8));
    double D = mgp_spline2d_eval(&c->sd[MGP_G_D], log(a), log(know));
    double DL = c->have_sd_lcdm ? mgp_spline2d_eval(&c->sd_lcdm[MGP_G_D], log(a), log(know))
                                : mgp_spline_eval(&c->glcdm[MGP_G_D], log(a));
    double pw = mgp_power_spec(c, know);
    in += pw * w * w * D * D * know * know * know * dlogk / (2.0 * M_PI * M_PI);
    inl += pw * w * w * DL * DL * know * know * know * dlogk / (2.0 * M_PI * M_PI);
  }
  return sqrt(in) / sqrt(inl);
}

/* per-|n|^2 table of the IC power spectrum: PowerSpec(k) with the MG rescaling
 * of src/2LPT.c:303-306, 394-410 */
void mgp_ic_pk_table(const mgp_cosmo *c, double *pk) {
  const mgp_params *p = c->p;
  const int N = p->Nmesh;
  const size_t len = (size_t)3 * (N / 2) * (N / 2) + 1;
  double s8r2 = 1.0;
  if (p->modified_gravity_active && p->input_pofk_is_for_lcdm) s8r2 = pow2(mgp_mg_sigma8_enhancement(c, 1.0));
  const int rescale = p->modified_gravity_active && p->input_pofk_is_for_lcdm;
  pk[0] = 0.0;
#pragma omp parallel for schedule(static)
  for (size_t n2 = 1; n2 < len; n2++) {
    const double k = 2.0 * MGP_PI / p->Box * sqrt((double)n2);
    double v = mgp_power_spec(c, k);
    if (rescale) {
      /* mg_pofk_ratio(k, 1) (src/2LPT.c:395-396): against the scale-dependent LCDM
       * growth when massive neutrinos make it scale-dependent */
      v *= mgp_mg_pofk_ratio(c, k, 1.0);
      if (!p->input_sigma8_is_for_lcdm) v /= s8r2;
    }
    pk[n2] = v;
  }
}

/* per-|n|^2 LPT growth tables for libmgpu (from_cdisp_store_to_ZA,
 * src/2LPT.c:1539-1616): g_o = normfactor_o * growth_o(k, A) */
void mgp_lpt_tables(const mgp_cosmo *c, int field, double A, double AFF, double *g1, double *g2) {
  const mgp_params *p = c->p;
  const int N = p->Nmesh;
  const size_t len = (size_t)3 * (N / 2) * (N / 2) + 1;
  const int ny = c->sd[0].ny;
  const int num1 = field == MGP_FIELD_D || field == MGP_FIELD_deltaD ? MGP_G_D
                   : field == MGP_FIELD_dDdy                        ? MGP_G_dDdy
                                                                    : MGP_G_ddDddy;
  const int num2 = num1 + 3;
  double *buf = malloc(sizeof(double) * ny * 12);
  double *n1v = buf, *n1d = buf + ny, *n2v = buf + 2 * ny, *n2d = buf + 3 * ny;
  double *d1v = buf + 4 * ny, *d1d = buf + 5 * ny, *d2v = buf + 6 * ny, *d2d = buf + 7 * ny;
  double *f1v = buf + 8 * ny, *f1d = buf + 9 * ny, *f2v = buf + 10 * ny, *f2d = buf + 11 * ny;
  mgp_spline2d_slice(&c->sd[num1], log(A), n1v, n1d);
  mgp_spline2d_slice(&c->sd[num2], log(A), n2v, n2d);
  if (field == MGP_FIELD_deltaD) {
    mgp_spline2d_slice(&c->sd[num1], log(AFF), f1v, f1d);
    mgp_spline2d_slice(&c->sd[num2], log(AFF), f2v, f2d);
    for (int j = 0; j < ny; j++) {
      n1v[j] = f1v[j] - n1v[j];
      n1d[j] = f1d[j] - n1d[j];
      n2v[j] = f2v[j] - n2v[j];
      n2d[j] = f2d[j] - n2d[j];
    }
  }
  mgp_spline2d_slice(&c->sd[MGP_G_D], 0.0, d1v, d1d);
  mgp_spline2d_slice(&c->sd[MGP_G_D2], 0.0, d2v, d2d);
  const double norm2 = -3.0 / 7.0 / ((double)N * N * N);
  g1[0] = g2[0] = 0.0;
#pragma omp parallel for schedule(static)
  for (size_t n2 = 1; n2 < len; n2++) {
    const double lk = log(2.0 * MGP_PI / p->Box * sqrt((double)n2));
    g1[n2] = mgp_slice_eval(&c->sd[0], n1v, n1d, lk) / mgp_slice_eval(&c->sd[0], d1v, d1d, lk);
    g2[n2] = norm2 * (mgp_slice_eval(&c->sd[0], n2v, n2d, lk) / mgp_slice_eval(&c->sd[0], d2v, d2d, lk));
  }
  free(buf);
}

/* The same growth as cubic Hermite slices in ln k for evaluation on the GPU
 * (mgpu_lpt_fields_slices): buf holds 8 * ny doubles. */
int mgp_lpt_kslices(const mgp_cosmo *c, int field, double A, double AFF, mgpu_kslice *o1, mgpu_kslice *o2,
                    double *buf) {
  if (!c->have_sd) return -1;
  const int ny = c->sd[0].ny;
  const int num1 = field == MGP_FIELD_D || field == MGP_FIELD_deltaD ? MGP_G_D
                   : field == MGP_FIELD_dDdy                        ? MGP_G_dDdy
                                                                    : MGP_G_ddDddy;
  const int num2 = num1 + 3;
  double *n1v = buf, *n1d = buf + ny, *d1v = buf + 2 * ny, *d1d = buf + 3 * ny;
  double *n2v = buf + 4 * ny, *n2d = buf + 5 * ny, *d2v = buf + 6 * ny, *d2d = buf + 7 * ny;
  mgp_spline2d_slice(&c->sd[num1], log(A), n1v, n1d);
  mgp_spline2d_slice(&c->sd[num2], log(A), n2v, n2d);
  if (field == MGP_FIELD_deltaD) {
    /* D(AFF) - D(A): the slices are linear in the tabulated values */
    double *tv = malloc(sizeof(double) * 2 * ny), *td = tv + ny;
    mgp_spline2d_slice(&c->sd[num1], log(AFF), tv, td);
    for (int j = 0; j < ny; j++) {
      n1v[j] = tv[j] - n1v[j];
      n1d[j] = td[j] - n1d[j];
    }
    mgp_spline2d_slice(&c->sd[num2], log(AFF), tv, td);
    for (int j = 0; j < ny; j++) {
      n2v[j] = tv[j] - n2v[j];
      n2d[j] = td[j] - n2d[j];
    }
    free(tv);
  }
  mgp_spline2d_slice(&c->sd[MGP_G_D], 0.0, d1v, d1d);
  mgp_spline2d_slice(&c->sd[MGP_G_D2], 0.0, d2v, d2d);
  const int N = c->p->Nmesh;
  *o1 = (mgpu_kslice){c->sd_lnkmin, c->sd_dlnk, ny, n1v, n1d, d1v, d1d, 1.0};
  *o2 = (mgpu_kslice){c->sd_lnkmin, c->sd_dlnk, ny, n2v, n2d, d2v, d2d, -3.0 / 7.0 / ((double)N * N * N)};
  return 0;
}

/* ------------------------------------------------------------------ object */
mgp_cosmo *mgp_cosmo_create(mgp_params *p, char *err, size_t errlen) {
  struct mgp_cosmo *c = calloc(1, sizeof(*c));
  c->p = p;
  c->fofr = p->model == MGP_MODEL_FOFR || p->model == MGP_MODEL_FOFR_LCDM || p->model == MGP_MODEL_FOFRNU;
  /* massive neutrinos (init_modified_version, src/user_defined_functions.h:50-78) */
  c->nu = p->model == MGP_MODEL_FOFRNU;
  p->OmegaNu = 0.0;
  p->OmegaCDM = p->Omega - p->OmegaBaryon;
  if (c->nu) {
    p->OmegaNu = p->nu_SumMassNuEV / (93.14 * p->HubbleParam * p->HubbleParam);
    p->OmegaCDM = p->Omega - p->OmegaBaryon - p->OmegaNu;
    if (p->OmegaNu < 1e-6) {
      p->OmegaNu = 0.0;
      p->OmegaCDM = p->Omega - p->OmegaBaryon;
      p->nu_include_massive_neutrinos = 0;
    }
  }
  c->dgp = p->model == MGP_MODEL_DGP;
  c->jbd = p->model == MGP_MODEL_BRANSDICKE;
  c->mbeta = p->model == MGP_MODEL_MBETA;
  if (c->jbd) {
    /* init_modified_version (src/user_defined_functions.h:167-178): h and Omega from the background */
    double h;
    if (mgp_jbd_solve(c, &h)) {
      snprintf(err, errlen, "JBD background solver failed");
      mgp_cosmo_free(c);
      return NULL;
    }
    p->HubbleParam = h;
    p->Omega = p->Omegah2 / h / h;
  }
  if (c->mbeta && p->modified_gravity_active) mgp_compute_phi_of_a(c);
  if (mgp_init_power(c, err, errlen)) {
    mgp_cosmo_free(c);
    return NULL;
  }
  if (mgp_solve_growth(c)) {
    snprintf(err, errlen, "growth-factor ODE integration failed");
    mgp_cosmo_free(c);
    return NULL;
  }
  return c;
}

void mgp_cosmo_free(mgp_cosmo *c) {
  if (!c) return;
  for (int i = 0; i < 6; i++) {
    if (c->g[i].n) mgp_spline_free(&c->g[i]);
    if (c->glcdm[i].n) mgp_spline_free(&c->glcdm[i]);
    if (c->sd[i].nx) mgp_spline2d_free(&c->sd[i]);
    if (c->sd_lcdm[i].nx) mgp_spline2d_free(&c->sd_lcdm[i]);
    if (i < 4 && c->camb[i].nx) mgp_spline2d_free(&c->camb[i]);
  }
  if (c->jbd_H.n) mgp_spline_free(&c->jbd_H);
  if (c->jbd_dHda.n) mgp_spline_free(&c->jbd_dHda);
  if (c->jbd_Geff.n) mgp_spline_free(&c->jbd_Geff);
  if (c->phi_of_a.n) mgp_spline_free(&c->phi_of_a);
  free(c->tab_logk);
  free(c->tab_logp);
  free(c);
}
