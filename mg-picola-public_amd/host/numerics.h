/* Numerical building blocks of the host (own implementations): natural cubic
 * splines, bicubic Hermite patches with spline derivatives, an adaptive
 * Dormand-Prince ODE integrator and adaptive Gauss-Kronrod quadrature. */
#ifndef MGP_NUMERICS_H
#define MGP_NUMERICS_H

#include <stddef.h>

/* natural cubic spline y(x) on increasing knots (the gsl_interp_cspline
 * interpolant the reference builds, src/wrappers.c:141-150); lookups clamp x
 * to the knot range as Lookup_GSL_Spline does (src/wrappers.c:162-167) */
typedef struct {
  int n;
  double *x, *y, *y2; /* y2: second derivatives at the knots */
} mgp_spline;
int mgp_spline_init(mgp_spline *s, const double *x, const double *y, int n);
void mgp_spline_free(mgp_spline *s);
double mgp_spline_eval(const mgp_spline *s, double x);
double mgp_spline_deriv(const mgp_spline *s, double x);

/* bicubic interpolation of z(x_i, y_j) (z[j*nx + i]) as gsl_interp2d_bicubic
 * builds it: Hermite patches whose zx, zy, zxy come from natural cubic
 * splines along each grid line (src/wrappers.c:102-137) */
typedef struct {
  int nx, ny;
  double *x, *y, *z, *zx, *zy, *zxy;
} mgp_spline2d;
int mgp_spline2d_init(mgp_spline2d *s, const double *x, const double *y, const double *z, int nx, int ny);
void mgp_spline2d_free(mgp_spline2d *s);
double mgp_spline2d_eval(const mgp_spline2d *s, double x, double y);
/* restriction to fixed x: value and d/dy at every y knot */
void mgp_spline2d_slice(const mgp_spline2d *s, double x, double *val, double *dval);
/* cubic Hermite evaluation of a slice at y */
double mgp_slice_eval(const mgp_spline2d *s, const double *val, const double *dval, double y);

/* adaptive Dormand-Prince 5(4) integration of dy/dx = f(x, y) from *x to x1 */
typedef int (*mgp_ode_rhs)(double x, const double *y, double *dydx, void *par);
int mgp_ode_integrate(mgp_ode_rhs f, void *par, int n, double *x, double x1, double *y, double *h, double abstol,
                      double reltol);

/* adaptive Gauss-Kronrod (7-15) quadrature of f on [a, b] */
typedef double (*mgp_fun1)(double x, void *par);
double mgp_integrate(mgp_fun1 f, void *par, double a, double b, double abstol, double reltol);

#endif
