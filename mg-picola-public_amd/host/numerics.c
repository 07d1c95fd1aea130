/* Numerical building blocks of the host; see numerics.h. */
#include "numerics.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------ cubic spline */
int mgp_spline_init(mgp_spline *s, const double *x, const double *y, int n) {
  s->n = n;
  s->x = (double *)malloc(sizeof(double) * n);
  s->y = (double *)malloc(sizeof(double) * n);
  s->y2 = (double *)calloc(n, sizeof(double));
  if (!s->x || !s->y || !s->y2) return -1;
  memcpy(s->x, x, sizeof(double) * n);
  memcpy(s->y, y, sizeof(double) * n);
  if (n < 3) return 0;
  /* natural end conditions: y2[0] = y2[n-1] = 0; Thomas algorithm */
  double *c = (double *)malloc(sizeof(double) * n), *d = (double *)malloc(sizeof(double) * n);
  c[0] = 0.0;
  d[0] = 0.0;
  for (int i = 1; i < n - 1; i++) {
    double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
    double a = h0, b = 2.0 * (h0 + h1), cc = h1;
    double r = 6.0 * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
    double m = b - a * c[i - 1];
    c[i] = cc / m;
    d[i] = (r - a * d[i - 1]) / m;
  }
  s->y2[n - 1] = 0.0;
  for (int i = n - 2; i >= 1; i--) s->y2[i] = d[i] - c[i] * s->y2[i + 1];
  s->y2[0] = 0.0;
  free(c);
  free(d);
  return 0;
}

void mgp_spline_free(mgp_spline *s) {
  free(s->x);
  free(s->y);
  free(s->y2);
  memset(s, 0, sizeof(*s));
}

static int find_interval(const double *x, int n, double v) {
  if (v <= x[0]) return 0;
  if (v >= x[n - 1]) return n - 2;
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    int mid = (lo + hi) / 2;
    if (v < x[mid]) hi = mid;
    else lo = mid;
  }
  return lo;
}

double mgp_spline_eval(const mgp_spline *s, double x) {
  if (x < s->x[0]) x = s->x[0];
  if (x > s->x[s->n - 1]) x = s->x[s->n - 1];
  int i = find_interval(s->x, s->n, x);
  double h = s->x[i + 1] - s->x[i];
  double a = (s->x[i + 1] - x) / h, b = (x - s->x[i]) / h;
  return a * s->y[i] + b * s->y[i + 1] + ((a * a * a - a) * s->y2[i] + (b * b * b - b) * s->y2[i + 1]) * h * h / 6.0;
}

double mgp_spline_deriv(const mgp_spline *s, double x) {
  if (x < s->x[0]) x = s->x[0];
  if (x > s->x[s->n - 1]) x = s->x[s->n - 1];
  int i = find_interval(s->x, s->n, x);
  double h = s->x[i + 1] - s->x[i];
  double a = (s->x[i + 1] - x) / h, b = (x - s->x[i]) / h;
  return (s->y[i + 1] - s->y[i]) / h - (3.0 * a * a - 1.0) / 6.0 * h * s->y2[i] +
         (3.0 * b * b - 1.0) / 6.0 * h * s->y2[i + 1];
}

/* ------------------------------------------------------------ bicubic */
int mgp_spline2d_init(mgp_spline2d *s, const double *x, const double *y, const double *z, int nx, int ny) {
  s->nx = nx;
  s->ny = ny;
  size_t nz = (size_t)nx * ny;
  s->x = (double *)malloc(sizeof(double) * nx);
  s->y = (double *)malloc(sizeof(double) * ny);
  s->z = (double *)malloc(sizeof(double) * nz);
  s->zx = (double *)malloc(sizeof(double) * nz);
  s->zy = (double *)malloc(sizeof(double) * nz);
  s->zxy = (double *)malloc(sizeof(double) * nz);
  if (!s->x || !s->y || !s->z || !s->zx || !s->zy || !s->zxy) return -1;
  memcpy(s->x, x, sizeof(double) * nx);
  memcpy(s->y, y, sizeof(double) * ny);
  memcpy(s->z, z, sizeof(double) * nz);
  int nmax = nx > ny ? nx : ny;
#pragma omp parallel
  {
    double *line = (double *)malloc(sizeof(double) * nmax);
    mgp_spline sp;
    /* zx: spline along x for every y */
#pragma omp for
    for (int j = 0; j < ny; j++) {
      for (int i = 0; i < nx; i++) line[i] = z[(size_t)j * nx + i];
      mgp_spline_init(&sp, x, line, nx);
      for (int i = 0; i < nx; i++) s->zx[(size_t)j * nx + i] = mgp_spline_deriv(&sp, x[i]);
      mgp_spline_free(&sp);
    }
    /* zy: spline along y for every x */
#pragma omp for
    for (int i = 0; i < nx; i++) {
      for (int j = 0; j < ny; j++) line[j] = z[(size_t)j * nx + i];
      mgp_spline_init(&sp, y, line, ny);
      for (int j = 0; j < ny; j++) s->zy[(size_t)j * nx + i] = mgp_spline_deriv(&sp, y[j]);
      mgp_spline_free(&sp);
    }
    /* zxy: spline along y of zx for every x (after zx is complete) */
#pragma omp barrier
#pragma omp for
    for (int i = 0; i < nx; i++) {
      for (int j = 0; j < ny; j++) line[j] = s->zx[(size_t)j * nx + i];
      mgp_spline_init(&sp, y, line, ny);
      for (int j = 0; j < ny; j++) s->zxy[(size_t)j * nx + i] = mgp_spline_deriv(&sp, y[j]);
      mgp_spline_free(&sp);
    }
    free(line);
  }
  return 0;
}

void mgp_spline2d_free(mgp_spline2d *s) {
  free(s->x);
  free(s->y);
  free(s->z);
  free(s->zx);
  free(s->zy);
  free(s->zxy);
  memset(s, 0, sizeof(*s));
}

/* cubic Hermite basis on t in [0,1] with interval length h */
static inline void hermite(double t, double h, double *h00, double *h10, double *h01, double *h11) {
  double t2 = t * t, t3 = t2 * t;
  *h00 = 2 * t3 - 3 * t2 + 1;
  *h10 = (t3 - 2 * t2 + t) * h;
  *h01 = -2 * t3 + 3 * t2;
  *h11 = (t3 - t2) * h;
}
static inline void hermite_d(double t, double h, double *d00, double *d10, double *d01, double *d11) {
  double t2 = t * t;
  *d00 = (6 * t2 - 6 * t) / h;
  *d10 = 3 * t2 - 4 * t + 1;
  *d01 = (-6 * t2 + 6 * t) / h;
  *d11 = 3 * t2 - 2 * t;
}

double mgp_spline2d_eval(const mgp_spline2d *s, double x, double y) {
  if (x < s->x[0]) x = s->x[0];
  if (x > s->x[s->nx - 1]) x = s->x[s->nx - 1];
  if (y < s->y[0]) y = s->y[0];
  if (y > s->y[s->ny - 1]) y = s->y[s->ny - 1];
  int i = find_interval(s->x, s->nx, x), j = find_interval(s->y, s->ny, y);
  double hx = s->x[i + 1] - s->x[i], hy = s->y[j + 1] - s->y[j];
  double tx = (x - s->x[i]) / hx, ty = (y - s->y[j]) / hy;
  double ax[4], ay[4];
  hermite(tx, hx, &ax[0], &ax[1], &ax[2], &ax[3]);
  hermite(ty, hy, &ay[0], &ay[1], &ay[2], &ay[3]);
  const int nx = s->nx;
#define Z(a, b) s->z[(size_t)(j + (b)) * nx + i + (a)]
#define ZX(a, b) s->zx[(size_t)(j + (b)) * nx + i + (a)]
#define ZY(a, b) s->zy[(size_t)(j + (b)) * nx + i + (a)]
#define ZXY(a, b) s->zxy[(size_t)(j + (b)) * nx + i + (a)]
  /* p(x,y) = sum_{a,b} [H0a(x) z + H1a(x) zx] ... in tensor form */
  double v[2], dv[2];
  for (int b = 0; b < 2; b++) {
    v[b] = ax[0] * Z(0, b) + ax[1] * ZX(0, b) + ax[2] * Z(1, b) + ax[3] * ZX(1, b);
    dv[b] = ax[0] * ZY(0, b) + ax[1] * ZXY(0, b) + ax[2] * ZY(1, b) + ax[3] * ZXY(1, b);
  }
#undef Z
#undef ZX
#undef ZY
#undef ZXY
  return ay[0] * v[0] + ay[1] * dv[0] + ay[2] * v[1] + ay[3] * dv[1];
}

void mgp_spline2d_slice(const mgp_spline2d *s, double x, double *val, double *dval) {
  if (x < s->x[0]) x = s->x[0];
  if (x > s->x[s->nx - 1]) x = s->x[s->nx - 1];
  int i = find_interval(s->x, s->nx, x);
  double hx = s->x[i + 1] - s->x[i];
  double tx = (x - s->x[i]) / hx;
  double ax[4];
  hermite(tx, hx, &ax[0], &ax[1], &ax[2], &ax[3]);
  const int nx = s->nx;
  for (int j = 0; j < s->ny; j++) {
    size_t o = (size_t)j * nx + i;
    val[j] = ax[0] * s->z[o] + ax[1] * s->zx[o] + ax[2] * s->z[o + 1] + ax[3] * s->zx[o + 1];
    dval[j] = ax[0] * s->zy[o] + ax[1] * s->zxy[o] + ax[2] * s->zy[o + 1] + ax[3] * s->zxy[o + 1];
  }
}

double mgp_slice_eval(const mgp_spline2d *s, const double *val, const double *dval, double y) {
  if (y < s->y[0]) y = s->y[0];
  if (y > s->y[s->ny - 1]) y = s->y[s->ny - 1];
  int j = find_interval(s->y, s->ny, y);
  double hy = s->y[j + 1] - s->y[j], ty = (y - s->y[j]) / hy;
  double a0, a1, a2, a3;
  hermite(ty, hy, &a0, &a1, &a2, &a3);
  return a0 * val[j] + a1 * dval[j] + a2 * val[j + 1] + a3 * dval[j + 1];
}

/* ------------------------------------------------------------ ODE (Dormand-Prince 5(4)) */
#define NMAX 8
int mgp_ode_integrate(mgp_ode_rhs f, void *par, int n, double *xp, double x1, double *y, double *hp, double abstol,
                      double reltol) {
  static const double c2 = 1. / 5, c3 = 3. / 10, c4 = 4. / 5, c5 = 8. / 9;
  static const double a21 = 1. / 5;
  static const double a31 = 3. / 40, a32 = 9. / 40;
  static const double a41 = 44. / 45, a42 = -56. / 15, a43 = 32. / 9;
  static const double a51 = 19372. / 6561, a52 = -25360. / 2187, a53 = 64448. / 6561, a54 = -212. / 729;
  static const double a61 = 9017. / 3168, a62 = -355. / 33, a63 = 46732. / 5247, a64 = 49. / 176,
                      a65 = -5103. / 18656;
  static const double b1 = 35. / 384, b3 = 500. / 1113, b4 = 125. / 192, b5 = -2187. / 6784, b6 = 11. / 84;
  static const double e1 = 71. / 57600, e3 = -71. / 16695, e4 = 71. / 1920, e5 = -17253. / 339200,
                      e6 = 22. / 525, e7 = -1. / 40;
  double x = *xp, h = *hp;
  double k1[NMAX], k2[NMAX], k3[NMAX], k4[NMAX], k5[NMAX], k6[NMAX], k7[NMAX], yt[NMAX], yn[NMAX];
  if (n > NMAX) return -1;
  if (x1 == x) return 0;
  double dir = x1 > x ? 1.0 : -1.0;
  if (h <= 0) h = 1e-4 * fabs(x1 - x);
  int steps = 0;
  if (f(x, y, k1, par)) return -1;
  while ((x1 - x) * dir > 0) {
    if (++steps > 1000000) return -1;
    double hh = h;
    if ((x + dir * hh - x1) * dir > 0) hh = fabs(x1 - x);
    double hs = dir * hh;
    for (int i = 0; i < n; i++) yt[i] = y[i] + hs * a21 * k1[i];
    f(x + c2 * hs, yt, k2, par);
    for (int i = 0; i < n; i++) yt[i] = y[i] + hs * (a31 * k1[i] + a32 * k2[i]);
    f(x + c3 * hs, yt, k3, par);
    for (int i = 0; i < n; i++) yt[i] = y[i] + hs * (a41 * k1[i] + a42 * k2[i] + a43 * k3[i]);
    f(x + c4 * hs, yt, k4, par);
    for (int i = 0; i < n; i++) yt[i] = y[i] + hs * (a51 * k1[i] + a52 * k2[i] + a53 * k3[i] + a54 * k4[i]);
    f(x + c5 * hs, yt, k5, par);
    for (int i = 0; i < n; i++)
      yt[i] = y[i] + hs * (a61 * k1[i] + a62 * k2[i] + a63 * k3[i] + a64 * k4[i] + a65 * k5[i]);
    f(x + hs, yt, k6, par);
    for (int i = 0; i < n; i++) yn[i] = y[i] + hs * (b1 * k1[i] + b3 * k3[i] + b4 * k4[i] + b5 * k5[i] + b6 * k6[i]);
    f(x + hs, yn, k7, par);
    double err = 0.0;
    for (int i = 0; i < n; i++) {
      double e = hs * (e1 * k1[i] + e3 * k3[i] + e4 * k4[i] + e5 * k5[i] + e6 * k6[i] + e7 * k7[i]);
      double sc = abstol + reltol * fmax(fabs(y[i]), fabs(yn[i]));
      double r = e / sc;
      err += r * r;
    }
    err = sqrt(err / n);
    if (err <= 1.0 || hh < 1e-14) {
      x += hs;
      memcpy(y, yn, sizeof(double) * n);
      memcpy(k1, k7, sizeof(double) * n);
      double fac = err > 0 ? 0.9 * pow(err, -0.2) : 5.0;
      if (fac > 5.0) fac = 5.0;
      if (fac < 0.2) fac = 0.2;
      if (hh == h) h *= fac;
    } else {
      double fac = 0.9 * pow(err, -0.25);
      if (fac < 0.1) fac = 0.1;
      h = hh * fac;
    }
  }
  *xp = x1;
  *hp = h;
  return 0;
}

/* ------------------------------------------------------------ Gauss-Kronrod 7-15 */
static const double gk_x[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                               0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                               0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                               0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
static const double gk_wk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                                0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                                0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                                0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
static const double gk_wg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                                0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

static double gk15(mgp_fun1 f, void *par, double a, double b, double *err) {
  double c = 0.5 * (a + b), h = 0.5 * (b - a);
  double fc = f(c, par);
  double rk = fc * gk_wk[7], rg = fc * gk_wg[3];
  for (int i = 0; i < 7; i++) {
    double dx = h * gk_x[i];
    double f1 = f(c - dx, par), f2 = f(c + dx, par);
    rk += gk_wk[i] * (f1 + f2);
    if (i % 2 == 1) rg += gk_wg[i / 2] * (f1 + f2);
  }
  *err = fabs((rk - rg) * h);
  return rk * h;
}

static double adapt(mgp_fun1 f, void *par, double a, double b, double whole, double err, double abstol,
                    double reltol, int depth) {
  if (err <= fmax(abstol, reltol * fabs(whole)) || depth > 40) return whole;
  double m = 0.5 * (a + b), e1, e2;
  double l = gk15(f, par, a, m, &e1), r = gk15(f, par, m, b, &e2);
  return adapt(f, par, a, m, l, e1, 0.5 * abstol, reltol, depth + 1) +
         adapt(f, par, m, b, r, e2, 0.5 * abstol, reltol, depth + 1);
}

double mgp_integrate(mgp_fun1 f, void *par, double a, double b, double abstol, double reltol) {
  double err;
  double whole = gk15(f, par, a, b, &err);
  return adapt(f, par, a, b, whole, err, abstol, reltol, 0);
}
