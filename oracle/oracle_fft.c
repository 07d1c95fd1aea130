/*
 * Oracle FFT: own mixed-radix Stockham FFT in FP64 (test infrastructure only).
 *
 * The reference calls FFTW3-MPI through my_fftw_execute (src/wrappers.c:63-94):
 * a 3D r2c forward transform (e^{-i}, unnormalised) and a 3D c2r backward
 * transform (e^{+i}, unnormalised), in place on padded arrays.  For c2r the
 * imaginary parts of the kz = 0 and kz = N/2 entries of each z-row are dropped
 * after the x and y passes, which is what FFTW's rdft2 backward transform does.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_MAXF 40

typedef struct {
  int n, nf, fac[OR_MAXF];
  double *tw; /* tw[2m], tw[2m+1] = cos, sin(2 pi m / n) */
} plan1d;

static plan1d g_plans[64];
static int g_nplans = 0;

static const plan1d *get_plan(int n) {
  for (int i = 0; i < g_nplans; i++)
    if (g_plans[i].n == n) return &g_plans[i];
  plan1d *p = &g_plans[g_nplans++];
  p->n = n;
  p->nf = 0;
  int m = n;
  while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
  while (m % 3 == 0) { p->fac[p->nf++] = 3; m /= 3; }
  while (m % 5 == 0) { p->fac[p->nf++] = 5; m /= 5; }
  for (int f = 7; m > 1; f += 2)
    while (m % f == 0) { p->fac[p->nf++] = f; m /= f; }
  p->tw = (double *)malloc(sizeof(double) * 2 * n);
  for (int k = 0; k < n; k++) {
    long double ang = 2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
    p->tw[2 * k] = (double)cosl(ang);
    p->tw[2 * k + 1] = (double)sinl(ang);
  }
  return p;
}

/* Build every plan a transform of size n will use before entering threads. */
static void prepare(int n) { (void)get_plan(n); }

/* one complex DFT of length r on v (in place), sign = -1 forward, +1 backward */
static void small_dft(int r, double *v, int sign, const plan1d *p) {
  double t[2 * 16];
  if (r == 2) {
    double a0 = v[0], a1 = v[1], b0 = v[2], b1 = v[3];
    v[0] = a0 + b0; v[1] = a1 + b1; v[2] = a0 - b0; v[3] = a1 - b1;
    return;
  }
  if (r == 4) {
    double a0r = v[0] + v[4], a0i = v[1] + v[5], a1r = v[0] - v[4], a1i = v[1] - v[5];
    double b0r = v[2] + v[6], b0i = v[3] + v[7], b1r = v[2] - v[6], b1i = v[3] - v[7];
    /* multiply b1 by -i*sign... : forward uses -i */
    double jr = sign < 0 ? b1i : -b1i, ji = sign < 0 ? -b1r : b1r;
    v[0] = a0r + b0r; v[1] = a0i + b0i;
    v[4] = a0r - b0r; v[5] = a0i - b0i;
    v[2] = a1r + jr;  v[3] = a1i + ji;
    v[6] = a1r - jr;  v[7] = a1i - ji;
    return;
  }
  /* generic: direct DFT with the plan's root table (n is a multiple of r) */
  int stride = p->n / r;
  for (int k = 0; k < r; k++) {
    double sr = 0, si = 0;
    for (int j = 0; j < r; j++) {
      int m = (j * k) % r;
      double c = p->tw[2 * m * stride], s = sign * p->tw[2 * m * stride + 1];
      sr += v[2 * j] * c - v[2 * j + 1] * s;
      si += v[2 * j] * s + v[2 * j + 1] * c;
    }
    t[2 * k] = sr; t[2 * k + 1] = si;
  }
  memcpy(v, t, sizeof(double) * 2 * r);
}

static void exec1d(const plan1d *p, double *x, double *work, int sign) {
  int n = p->n;
  double *src = x, *dst = work;
  int ns = 1;
  double v[2 * 16];
  for (int f = 0; f < p->nf; f++) {
    int r = p->fac[f];
    int m = n / r;
    for (int j = 0; j < m; j++) {
      int k = j % ns;
      for (int q = 0; q < r; q++) {
        double ar = src[2 * (j + q * m)], ai = src[2 * (j + q * m) + 1];
        /* twiddle exp(sign 2 pi i q k / (ns r)) = tw[q k n/(ns r)] */
        int idx = (int)(((long)q * k * (n / (ns * r))) % n);
        double c = p->tw[2 * idx], s = sign * p->tw[2 * idx + 1];
        v[2 * q] = ar * c - ai * s;
        v[2 * q + 1] = ar * s + ai * c;
      }
      small_dft(r, v, sign, p);
      int d = (j / ns) * ns * r + k;
      for (int q = 0; q < r; q++) {
        dst[2 * (d + q * ns)] = v[2 * q];
        dst[2 * (d + q * ns) + 1] = v[2 * q + 1];
      }
    }
    double *t = src; src = dst; dst = t;
    ns *= r;
  }
  if (src != x) memcpy(x, src, sizeof(double) * 2 * n);
}

void or_fft_c2c_1d(int n, double *data, int sign) {
  const plan1d *p = get_plan(n);
  double *w = (double *)malloc(sizeof(double) * 2 * n);
  exec1d(p, data, w, sign);
  free(w);
}

size_t or_real_len(int n) { return (size_t)(n + 1) * n * 2 * (n / 2 + 1); }

/* strided complex passes along y (axis 1) and x (axis 0) of the half spectrum */
static void pass_yx(int n, double *c, int axis, int sign) {
  const int h = n / 2 + 1;
  const plan1d *p = get_plan(n);
#pragma omp parallel
  {
    double *buf = (double *)malloc(sizeof(double) * 4 * n);
    double *w = buf + 2 * n;
#pragma omp for schedule(static)
    for (int outer = 0; outer < n; outer++) {
      for (int kz = 0; kz < h; kz++) {
        size_t base, stride;
        if (axis == 1) { base = ((size_t)outer * n) * h + kz; stride = h; }
        else { base = (size_t)outer * h + kz; stride = (size_t)n * h; }
        for (int i = 0; i < n; i++) {
          buf[2 * i] = c[2 * (base + i * stride)];
          buf[2 * i + 1] = c[2 * (base + i * stride) + 1];
        }
        exec1d(p, buf, w, sign);
        for (int i = 0; i < n; i++) {
          c[2 * (base + i * stride)] = buf[2 * i];
          c[2 * (base + i * stride) + 1] = buf[2 * i + 1];
        }
      }
    }
    free(buf);
  }
}

void or_fft_r2c_3d(int n, double *data) {
  const int h = n / 2 + 1;
  const plan1d *p = get_plan(n);
  prepare(n);
#pragma omp parallel
  {
    double *buf = (double *)malloc(sizeof(double) * 4 * n);
    double *w = buf + 2 * n;
#pragma omp for schedule(static)
    for (long row = 0; row < (long)n * n; row++) {
      double *r = data + (size_t)row * 2 * h;
      for (int i = 0; i < n; i++) { buf[2 * i] = r[i]; buf[2 * i + 1] = 0.0; }
      exec1d(p, buf, w, -1);
      for (int k = 0; k < h; k++) { r[2 * k] = buf[2 * k]; r[2 * k + 1] = buf[2 * k + 1]; }
    }
    free(buf);
  }
  pass_yx(n, data, 1, -1);
  pass_yx(n, data, 0, -1);
}

void or_fft_c2r_3d(int n, double *data) {
  const int h = n / 2 + 1;
  const plan1d *p = get_plan(n);
  prepare(n);
  pass_yx(n, data, 0, +1);
  pass_yx(n, data, 1, +1);
#pragma omp parallel
  {
    double *buf = (double *)malloc(sizeof(double) * 4 * n);
    double *w = buf + 2 * n;
#pragma omp for schedule(static)
    for (long row = 0; row < (long)n * n; row++) {
      double *r = data + (size_t)row * 2 * h;
      for (int k = 0; k < h; k++) { buf[2 * k] = r[2 * k]; buf[2 * k + 1] = r[2 * k + 1]; }
      buf[1] = 0.0;                 /* Im of DC dropped */
      if (n % 2 == 0) buf[n + 1] = 0.0; /* Im of Nyquist dropped */
      for (int k = h; k < n; k++) { buf[2 * k] = buf[2 * (n - k)]; buf[2 * k + 1] = -buf[2 * (n - k) + 1]; }
      exec1d(p, buf, w, +1);
      for (int i = 0; i < n; i++) r[i] = buf[2 * i];
      r[n] = 0.0;
      if (2 * h > n + 1) r[n + 1] = 0.0;
    }
    free(buf);
  }
}
