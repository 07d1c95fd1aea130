// Scale-dependent growth tables on the GPU: calculate_scale_dependent_growth_factor
// (src/cosmo.c:798-981).  One thread per k integrates the first- and second-order
// growth ODEs (ode_growth_D / ode_growth_DLCDM, src/cosmo.c:206-261) over
// x = ln a with an adaptive Dormand-Prince 5(4) method (Dormand & Prince 1980,
// J. Comp. Appl. Math. 6, 19) and writes the six table fields at every node of
// the uniform x grid, in FP64.  The step control is the host solver's: an RMS
// error norm with absolute and relative tolerance 1e-11, steps clipped to land
// on each node, the step size carried from one interval to the next.
//
// The hooks are restated for the device: the LCDM background (hubble,
// dhubbleda), GeffoverG and second_order_kernel of f(R) Hu-Sawicki and of the
// symmetron-like m(a)/beta(a) model (src/user_defined_functions.h:462-652,
// :800-904), and GeffoverG_neutrino_1LPT/2LPT (:662-680) from the CAMB
// transfer functions as bicubic Hermite splines in (ln k, z).
#include "mgpu_internal.h"

namespace {

constexpr double kInvH0 = MGPU_INVERSE_H0_MPCH;
constexpr double kTol = 1e-11;
constexpr int kMaxSteps = 1000000;

struct Spl {
  int nx, ny;
  const double *x, *y, *z, *zx, *zy, *zxy;
};

struct Model {
  int model, mg, lcdm, nu;
  double om, fofr0, nfofr, assb, beta_s, range_s;
  double onu, ob, ocdm;
  Spl tnu, tcdm, tb;
};

__device__ inline double sq(double v) { return v * v; }
__device__ inline double cube(double v) { return v * v * v; }

__device__ int interval(const double *x, int n, double v) {
  if (v <= x[0]) return 0;
  if (v >= x[n - 1]) return n - 2;
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (v < x[mid]) hi = mid;
    else lo = mid;
  }
  return lo;
}

// bicubic Hermite patch from values and first/cross derivatives at the knots
__device__ double spl_eval(const Spl &s, double x, double y) {
  x = fmin(fmax(x, s.x[0]), s.x[s.nx - 1]);
  y = fmin(fmax(y, s.y[0]), s.y[s.ny - 1]);
  const int i = interval(s.x, s.nx, x), j = interval(s.y, s.ny, y);
  const double hx = s.x[i + 1] - s.x[i], hy = s.y[j + 1] - s.y[j];
  const double tx = (x - s.x[i]) / hx, ty = (y - s.y[j]) / hy;
  auto basis = [](double t, double h, double b[4]) {
    const double t2 = t * t, t3 = t2 * t;
    b[0] = 2 * t3 - 3 * t2 + 1;
    b[1] = (t3 - 2 * t2 + t) * h;
    b[2] = -2 * t3 + 3 * t2;
    b[3] = (t3 - t2) * h;
  };
  double ax[4], ay[4];
  basis(tx, hx, ax);
  basis(ty, hy, ay);
  double v[2], dv[2];
  for (int b = 0; b < 2; b++) {
    const size_t o = (size_t)(j + b) * s.nx + i;
    v[b] = ax[0] * s.z[o] + ax[1] * s.zx[o] + ax[2] * s.z[o + 1] + ax[3] * s.zx[o + 1];
    dv[b] = ax[0] * s.zy[o] + ax[1] * s.zxy[o] + ax[2] * s.zy[o + 1] + ax[3] * s.zxy[o + 1];
  }
  return ay[0] * v[0] + ay[1] * dv[0] + ay[2] * v[1] + ay[3] * dv[1];
}

__device__ inline double hub(const Model &m, double a) { return sqrt(m.om / (a * a * a) + 1.0 - m.om); }
__device__ inline double dhub(const Model &m, double a) {
  return 1.0 / (2.0 * hub(m, a)) * (-3.0 * m.om / (a * a * a * a));
}

// m(a)^2/H0^2 and its a-derivative (src/user_defined_functions.h:490-560)
__device__ double mass2(const Model &m, double a) {
  if (m.model == MGPU_FOFR) {
    const double fac = m.om / (a * a * a) + 4.0 * (1.0 - m.om), fac0 = m.om + 4.0 * (1.0 - m.om);
    return fac0 * pow(fac / fac0, m.nfofr + 2.0) / ((1.0 + m.nfofr) * m.fofr0);
  }
  if (a > m.assb) return 0.5 * sq(2998.0 / m.range_s) * (1.0 - cube(m.assb / a));
  return 1.0;
}
__device__ double dmass2(const Model &m, double a) {
  if (a > m.assb) return 0.5 * sq(2998.0 / m.range_s) * (+3.0 * cube(m.assb / a) / a);
  return 0.0;
}
__device__ double beta(const Model &m, double a) {
  if (m.model == MGPU_FOFR) return 1.0 / sqrt(6.0);
  if (a > m.assb) return m.beta_s * sqrt(1.0 - cube(m.assb / a));
  return 0.0;
}

// GeffoverG(a, k) (src/user_defined_functions.h:618-652)
__device__ double geff(const Model &m, double a, double k) {
  if (!m.mg || m.model == MGPU_LCDM) return 1.0;
  const double mass2a2 = a * a * mass2(m, a), k2 = sq(k * kInvH0);
  if (k2 == 0) return 1.0;
  const double b = beta(m, a);
  return 1.0 * (1.0 + 2.0 * b * b * k2 / (k2 + mass2a2));
}

// second_order_kernel(k, k/sqrt2, k/sqrt2, cos = 0, a) (src/user_defined_functions.h:800-904):
// the cos-dependent term vanishes
__device__ double gamma2(const Model &m, double a, double k) {
  if (!m.mg || m.model == MGPU_LCDM) return 0.0;
  const double k1 = k / sqrt(2.0);
  if (m.model == MGPU_FOFR) {
    const double a3 = a * a * a, fac = m.om / a3 + 4.0 * (1.0 - m.om), fac0 = m.om + 4.0 * (1.0 - m.om);
    const double H = hub(m, a);
    double g = -(3.0 * (m.nfofr + 2.0) / (12.0 * (1.0 + m.nfofr) * (1.0 + m.nfofr))) * sq(k * kInvH0 / (a * H)) *
               sq(m.om / a3) * fac0 * pow(fac / fac0, 2.0 * m.nfofr + 3.0) / sq(m.fofr0);
    const double mterm = fac0 * pow(fac / fac0, m.nfofr + 2.0) / (1.0 + m.nfofr) / m.fofr0;
    const double pik = sq(k * kInvH0 / a) + mterm, pik1 = sq(k1 * kInvH0 / a) + mterm;
    return g / (pik * pik1 * pik1);
  }
  const double m2 = mass2(m, a), dm2da = dmass2(m, a), b = beta(m, a), H = hub(m, a);
  const double pk = sq(k * kInvH0 / a) + m2, pk1 = sq(k1 * kInvH0 / a) + m2;
  return (dm2da / m2) * sq(b) * m.om / (2.0 * a * a) * (m2 / pk1) * (m2 / pk1) * (1.0 - m2 / pk) / sq(H);
}

// GeffoverG_neutrino_1LPT (src/user_defined_functions.h:662-672): T_nu/T_cb at (k, a)
__device__ double munu1(const Model &m, double a, double lnk) {
  if (!m.nu) return 1.0;
  const double z = 1.0 / a - 1.0;
  const double tnu = spl_eval(m.tnu, lnk, z), tcdm = spl_eval(m.tcdm, lnk, z), tb = spl_eval(m.tb, lnk, z);
  const double tcb = (m.ob * tb + m.ocdm * tcdm) / (m.ob + m.ocdm);
  return (m.om - m.onu) / m.om + m.onu / m.om * tnu / tcb;
}
__device__ double munu2(const Model &m) { return m.nu ? (m.om - m.onu) / m.om : 1.0; }

// y = (D1, dD1/dx, D2, dD2/dx)
__device__ void rhs(const Model &m, double k, double lnk, double x, const double y[4], double d[4]) {
  const double a = exp(x), H = hub(m, a), dH = dhub(m, a);
  const double alphafac = 2.0 + a * dH / H;
  double mu = 1.0, modfac = 1.0;
  if (!m.lcdm) {
    mu = geff(m, a, k);
    modfac = 1.0 + 2.0 * sq(a * a * H) * gamma2(m, a, k) / (1.5 * m.om * a * mu);
  }
  const double betafac = 1.5 * m.om * mu / (a * a * a * H * H);
  const double nu1 = munu1(m, a, lnk), nu2 = munu2(m);
  d[0] = y[1];
  d[1] = -alphafac * y[1] + betafac * nu1 * y[0];
  d[2] = y[3];
  d[3] = -alphafac * y[3] + betafac * nu2 * (y[2] - modfac * y[0] * y[0]);
}

// Out: 6 fields x nk x na (field, k, a), the stored forms of the host tables:
// D1, dD1/dy, the 1LPT source 1.5 mu mu_nu1 Om a D1, D2, dD2/dy, the 2LPT source.
__global__ __launch_bounds__(64) void k_growth(Model m, const double *__restrict__ lnkv, int nk, double x0,
                                               double dx, int na, double *__restrict__ out,
                                               int *__restrict__ fail) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nk) return;
  // k and ln k as the host evaluates them: k = exp(ln k_j), the splines at log(k)
  const double k = exp(lnkv[j]), lnk = log(k);
  // Dormand-Prince 5(4) tableau
  const double c2 = 1. / 5, c3 = 3. / 10, c4 = 4. / 5, c5 = 8. / 9;
  const double a21 = 1. / 5, a31 = 3. / 40, a32 = 9. / 40, a41 = 44. / 45, a42 = -56. / 15, a43 = 32. / 9;
  const double a51 = 19372. / 6561, a52 = -25360. / 2187, a53 = 64448. / 6561, a54 = -212. / 729;
  const double a61 = 9017. / 3168, a62 = -355. / 33, a63 = 46732. / 5247, a64 = 49. / 176, a65 = -5103. / 18656;
  const double b1 = 35. / 384, b3 = 500. / 1113, b4 = 125. / 192, b5 = -2187. / 6784, b6 = 11. / 84;
  const double e1 = 71. / 57600, e3 = -71. / 16695, e4 = 71. / 1920, e5 = -17253. / 339200, e6 = 22. / 525,
               e7 = -1. / 40;
  double y[4] = {1.0, 1.0, -3.0 / 7.0, -6.0 / 7.0};
  double k1[4], k2[4], k3[4], k4[4], k5[4], k6[4], k7[4], yt[4], yn[4];
  double x = x0, h = 1e-4;
  int steps = 0, bad = 0;  // steps: all intervals together, a bound every thread reaches
  const size_t plane = (size_t)nk * na;
  for (int i = 0; i < na; i++) {
    const double x1 = x0 + i * dx;
    if (i > 0 && !bad) {
      // one host-style integrate call from x to x1
      rhs(m, k, lnk, x, y, k1);
      while (x1 - x > 0) {
        if (++steps > kMaxSteps) {
          bad = 1;
          break;
        }
        double hh = h;
        if (x + hh - x1 > 0) hh = fabs(x1 - x);
        const double hs = hh;
#pragma unroll
        for (int q = 0; q < 4; q++) yt[q] = y[q] + hs * a21 * k1[q];
        rhs(m, k, lnk, x + c2 * hs, yt, k2);
#pragma unroll
        for (int q = 0; q < 4; q++) yt[q] = y[q] + hs * (a31 * k1[q] + a32 * k2[q]);
        rhs(m, k, lnk, x + c3 * hs, yt, k3);
#pragma unroll
        for (int q = 0; q < 4; q++) yt[q] = y[q] + hs * (a41 * k1[q] + a42 * k2[q] + a43 * k3[q]);
        rhs(m, k, lnk, x + c4 * hs, yt, k4);
#pragma unroll
        for (int q = 0; q < 4; q++) yt[q] = y[q] + hs * (a51 * k1[q] + a52 * k2[q] + a53 * k3[q] + a54 * k4[q]);
        rhs(m, k, lnk, x + c5 * hs, yt, k5);
#pragma unroll
        for (int q = 0; q < 4; q++)
          yt[q] = y[q] + hs * (a61 * k1[q] + a62 * k2[q] + a63 * k3[q] + a64 * k4[q] + a65 * k5[q]);
        rhs(m, k, lnk, x + hs, yt, k6);
#pragma unroll
        for (int q = 0; q < 4; q++)
          yn[q] = y[q] + hs * (b1 * k1[q] + b3 * k3[q] + b4 * k4[q] + b5 * k5[q] + b6 * k6[q]);
        rhs(m, k, lnk, x + hs, yn, k7);
        double err = 0.0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const double e = hs * (e1 * k1[q] + e3 * k3[q] + e4 * k4[q] + e5 * k5[q] + e6 * k6[q] + e7 * k7[q]);
          const double r = e / (kTol + kTol * fmax(fabs(y[q]), fabs(yn[q])));
          err += r * r;
        }
        err = sqrt(err / 4);
        if (!(err == err)) {  // NaN: the solution is lost
          bad = 1;
          break;
        }
        if (err <= 1.0 || hh < 1e-14) {
          x += hs;
#pragma unroll
          for (int q = 0; q < 4; q++) {
            y[q] = yn[q];
            k1[q] = k7[q];
          }
          double fac = err > 0 ? 0.9 * pow(err, -0.2) : 5.0;
          fac = fmin(fmax(fac, 0.2), 5.0);
          if (hh == h) h *= fac;
        } else {
          const double fac = fmax(0.9 * pow(err, -0.25), 0.1);
          h = hh * fac;
        }
      }
      x = x1;
    }
    const double a = exp(x1), H = hub(m, a), Q = H * a * a * a;
    double s1 = 1.5 * m.om * a, s2 = s1;
    if (!m.lcdm) {
      const double mu = geff(m, a, k);
      s1 *= mu * munu1(m, a, lnk);
      s2 *= mu * munu2(m);
    }
    const size_t o = (size_t)j * na + i;
    out[o] = y[0];
    out[plane + o] = y[1] * Q / a;
    out[2 * plane + o] = s1 * y[0];
    out[3 * plane + o] = y[2];
    out[4 * plane + o] = y[3] * Q / a;
    out[5 * plane + o] = s2 * (y[2] - y[0] * y[0]);
  }
  if (bad) atomicOr(fail, 1);
}

Spl to_spl(const mgpu_spline2d &s, std::vector<double *> &dev, int &rc) {
  Spl o{s.nx, s.ny, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  const double *src[6] = {s.x, s.y, s.z, s.zx, s.zy, s.zxy};
  const size_t len[6] = {(size_t)s.nx, (size_t)s.ny, (size_t)s.nx * s.ny, (size_t)s.nx * s.ny,
                         (size_t)s.nx * s.ny, (size_t)s.nx * s.ny};
  const double **dst[6] = {&o.x, &o.y, &o.z, &o.zx, &o.zy, &o.zxy};
  for (int f = 0; f < 6 && !rc; f++) {
    double *p = nullptr;
    if (hipMalloc(&p, len[f] * sizeof(double)) != hipSuccess) {
      rc = MGPU_E_NOMEM;
      break;
    }
    dev.push_back(p);
    if (hipMemcpy(p, src[f], len[f] * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = MGPU_E_HIP;
    *dst[f] = p;
  }
  return o;
}

}  // namespace

int growth_tables(const mgpu_growth_model *gm, double x0, double dx, int na, const double *lnk, int nk,
                  double *out) {
  Model m{};
  m.model = gm->model;
  m.mg = gm->mg_active;
  m.lcdm = gm->lcdm;
  m.om = gm->omega;
  m.fofr0 = gm->fofr0;
  m.nfofr = gm->nfofr;
  m.assb = gm->assb_symm;
  m.beta_s = gm->beta_symm;
  m.range_s = gm->range_symm;
  m.nu = gm->omega_nu > 1e-6;
  m.onu = gm->omega_nu;
  m.ob = gm->omega_b;
  m.ocdm = gm->omega_cdm;
  std::vector<double *> dev;
  int rc = MGPU_OK;
  if (m.nu) {
    m.tnu = to_spl(gm->t_nu, dev, rc);
    m.tcdm = to_spl(gm->t_cdm, dev, rc);
    m.tb = to_spl(gm->t_b, dev, rc);
  }
  double *d_lnk = nullptr, *d_out = nullptr;
  int *d_fail = nullptr;
  const size_t outlen = 6 * (size_t)nk * na;
  if (!rc && (hipMalloc(&d_lnk, nk * sizeof(double)) != hipSuccess ||
              hipMalloc(&d_out, outlen * sizeof(double)) != hipSuccess || hipMalloc(&d_fail, sizeof(int)) != hipSuccess))
    rc = MGPU_E_NOMEM;
  if (!rc && (hipMemcpy(d_lnk, lnk, nk * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
              hipMemset(d_fail, 0, sizeof(int)) != hipSuccess))
    rc = MGPU_E_HIP;
  if (!rc) {
    hipLaunchKernelGGL(k_growth, dim3((nk + 63) / 64), dim3(64), 0, 0, m, (const double *)d_lnk, nk, x0, dx, na, d_out,
                       d_fail);
    int fail = 0;
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(&fail, d_fail, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(out, d_out, outlen * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      rc = MGPU_E_HIP;
    else if (fail)
      rc = MGPU_E_INVALID;
  }
  for (double *p : dev) (void)hipFree(p);
  if (d_lnk) (void)hipFree(d_lnk);
  if (d_out) (void)hipFree(d_out);
  if (d_fail) (void)hipFree(d_fail);
  return rc;
}
