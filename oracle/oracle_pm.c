/*
 * Oracle: PM/COLA step of MG-PICOLA restated in FP64 C (test infrastructure only).
 *
 * Every function follows one reference routine; the file:line it restates is
 * given above it.  Global state of the reference becomes explicit arguments.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define OR_PI 3.14159265358979323846          /* src/vars.h:61 */
#define OR_INVERSE_H0_MPCH 2997.92458         /* src/vars.h:65 */

static inline size_t ridx(int n, int ix, int iy, int iz) { return ((size_t)ix * n + iy) * 2 * (n / 2 + 1) + iz; }
static inline size_t cidx(int n, int ix, int iy, int iz) { return ((size_t)ix * n + iy) * (n / 2 + 1) + iz; }

/* ------------------------------------------------------------------------
 * PtoMesh: CIC deposit, density initialised to -1, ghost slab added to
 * slab 0 with +1.  src/auxPM.c:280-356.
 * ---------------------------------------------------------------------- */
/* threads the oracle's OpenMP loops use */
int or_num_threads(void) { return omp_get_max_threads(); }

/* one particle's eight CIC updates, in the reference's order */
static inline void cic_one(int N, double scaleBox, double WPAR, double *density, float px, float py, float pz) {
  double X = px * scaleBox, Y = py * scaleBox, Z = pz * scaleBox;
  unsigned IX = (unsigned)X, IY = (unsigned)Y, IZ = (unsigned)Z;
  double DX = X - (double)IX, DY = Y - (double)IY, DZ = Z - (double)IZ;
  double TX = 1.0 - DX, TY = 1.0 - DY, TZ = 1.0 - DZ;
  DY *= WPAR;
  TY *= WPAR;
  if (IY >= (unsigned)N) IY = 0;
  if (IZ >= (unsigned)N) IZ = 0;
  unsigned IXn = IX + 1, IYn = IY + 1, IZn = IZ + 1;
  if (IYn >= (unsigned)N) IYn = 0;
  if (IZn >= (unsigned)N) IZn = 0;
  density[ridx(N, IX, IY, IZ)] += TX * TY * TZ;
  density[ridx(N, IX, IY, IZn)] += TX * TY * DZ;
  density[ridx(N, IX, IYn, IZ)] += TX * DY * TZ;
  density[ridx(N, IX, IYn, IZn)] += TX * DY * DZ;
  density[ridx(N, IXn, IY, IZ)] += DX * TY * TZ;
  density[ridx(N, IXn, IY, IZn)] += DX * TY * DZ;
  density[ridx(N, IXn, IYn, IZ)] += DX * DY * TZ;
  density[ridx(N, IXn, IYn, IZn)] += DX * DY * DZ;
}

void or_pm_deposit(const or_cfg *c, size_t np, const float *const pos[3], double *density) {
  const int N = c->nmesh;
  const double scaleBox = (double)N / c->box;
  const double WPAR = pow((double)N / (double)c->nsample, 3);
  const size_t total = or_real_len(N);
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < total; i++) density[i] = -1.0;
  const int nth = omp_get_max_threads();
  if (nth == 1 || np < 65536 || N < 8) {
    for (size_t i = 0; i < np; i++) cic_one(N, scaleBox, WPAR, density, pos[0][i], pos[1][i], pos[2][i]);
  } else {
    /* Threads own blocks of S >= 2 x-slabs; a particle in slab IX writes slabs IX and
     * IX + 1, so blocks two apart never touch the same cells: even blocks run
     * concurrently, then odd blocks.  Particles are visited in index order within
     * their slab, so the result does not depend on the thread count. */
    int S = N / (4 * nth);
    if (S < 2) S = 2;
    const int nb = (N + S - 1) / S;
    size_t *start = (size_t *)calloc((size_t)N + 1, sizeof(size_t));
    unsigned *slab = (unsigned *)malloc(np * sizeof(unsigned));
    size_t *order = (size_t *)malloc(np * sizeof(size_t));
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < np; i++) slab[i] = (unsigned)(pos[0][i] * scaleBox);
    for (size_t i = 0; i < np; i++) start[slab[i] + 1]++;
    for (int x = 0; x < N; x++) start[x + 1] += start[x];
    size_t *fill = (size_t *)malloc((size_t)N * sizeof(size_t));
    memcpy(fill, start, (size_t)N * sizeof(size_t));
    for (size_t i = 0; i < np; i++) order[fill[slab[i]]++] = i;
    for (int phase = 0; phase < 2; phase++) {
#pragma omp parallel for schedule(dynamic, 1)
      for (int b = phase; b < nb; b += 2) {
        const int x1 = (b + 1) * S < N ? (b + 1) * S : N;
        for (size_t q = start[b * S]; q < start[x1]; q++) {
          const size_t i = order[q];
          cic_one(N, scaleBox, WPAR, density, pos[0][i], pos[1][i], pos[2][i]);
        }
      }
    }
    free(start);
    free(slab);
    free(order);
    free(fill);
  }
  /* ghost slab (x = N) onto slab 0, with the +1 of src/auxPM.c:353-355 */
  const size_t slice = (size_t)N * 2 * (N / 2 + 1);
  if (np != 0)
    for (size_t i = 0; i < slice; i++) density[i] += (OR_PLANT == 4 ? 0.0 : density[(size_t)N * slice + i]) + 1.0;
}

/* ------------------------------------------------------------------------
 * adjust_pofk_parameters, src/compute_pofk.c:758-805 (kmin/kmax in integer-k units)
 * ---------------------------------------------------------------------- */
void or_adjust_pofk(int nmesh, int *nbins, int *bintype, int *subtract, double *kmin, double *kmax) {
  if (!(*bintype == 0 || *bintype == 1)) *bintype = 0;
  if (!(*subtract == 0 || *subtract == 1)) *subtract = 1;
  if (*nbins <= 0) *nbins = nmesh;
  if (*kmax <= *kmin) {
    if (*bintype == 0) *kmin = 0.0;
    if (*bintype == 1) *kmin = 1.0;
    *kmax = (double)nmesh;
  }
  if (*kmin < 0.0) {
    if (*bintype == 0) *kmin = 0.0;
    if (*bintype == 1) *kmin = 1.0;
  }
  if (*bintype == 1 && *kmin == 0.0) *kmin = 1.0;
  if (*kmax > sqrt(3.0) * (double)nmesh) *kmax = (double)nmesh;
}

/* pofk_bin_index, src/compute_pofk.c:30-45 */
static int bin_index(double kmag, double kmin, double kmax, int nbins, int bintype) {
  if (bintype == 0) return (int)((kmag - kmin) / (kmax - kmin) * nbins + 0.5);
  if (kmag <= 0.0) return -1;
  return (int)(log(kmag / kmin) / log(kmax / kmin) * nbins + 0.5);
}

/* ------------------------------------------------------------------------
 * compute_power_spectrum binning + normalisation, src/compute_pofk.c:71-236.
 * pofk[i] is normalised (Box^3, shot noise), kmean[i] in h/Mpc, nmodes[i] raw.
 * ---------------------------------------------------------------------- */
void or_pofk(const or_cfg *c, const double *dk, int nbins, int bintype, double kmin, double kmax,
             int subtract_shotnoise, double *pofk, double *kmean, double *nmodes) {
  const int N = c->nmesh;
  for (int i = 0; i < nbins; i++) pofk[i] = kmean[i] = nmodes[i] = 0.0;
  const double norm = 1.0 / pow((double)N, 6);
  /* each thread bins the x planes it owns into its own sums; the sums are added in thread order */
  const int nth = omp_get_max_threads();
  double *acc = (double *)calloc((size_t)nth * 3 * nbins, sizeof(double));
#pragma omp parallel
  {
    double *pofk_t = acc + (size_t)omp_get_thread_num() * 3 * nbins, *kmean_t = pofk_t + nbins,
           *nmodes_t = kmean_t + nbins;
#pragma omp for schedule(static)
    for (int i = 0; i < N; i++) {
      int d0 = i > N / 2 ? N - i : i;
      double gx = d0 == 0 ? 1.0 : sin((OR_PI * d0) / (double)N) / ((OR_PI * d0) / (double)N);
      for (int j = 0; j < N / 2 + 1; j++) {
        int d1 = j;
        double gy = d1 == 0 ? 1.0 : sin((OR_PI * d1) / (double)N) / ((OR_PI * d1) / (double)N);
        for (int pass = 0; pass < 2; pass++) {
          int d2 = pass == 0 ? 0 : N / 2;
          double gz = pass == 0 ? 1.0 : 2.0 / OR_PI;
          double kmag = sqrt((double)(d0 * d0 + d1 * d1 + d2 * d2));
          int nk = bin_index(kmag, kmin, kmax, nbins, bintype);
          if (nk >= 0 && nk < nbins) {
            double gc = 1.0 / pow(gx * gy * gz, 4.0) * norm;
            const double *v = dk + 2 * cidx(N, i, j, d2);
            pofk_t[nk] += (v[0] * v[0] + v[1] * v[1]) * gc;
            kmean_t[nk] += kmag;
            nmodes_t[nk] += 1.0;
            if (j != N / 2 && j != 0) {
              v = dk + 2 * cidx(N, i, N - j, d2);
              pofk_t[nk] += (v[0] * v[0] + v[1] * v[1]) * gc;
              kmean_t[nk] += kmag;
              nmodes_t[nk] += 1.0;
            }
          }
        }
        for (int k = 1; k < N / 2; k++) {
          double gz = sin((OR_PI * k) / (double)N) / ((OR_PI * k) / (double)N);
          double kmag = sqrt((double)(d0 * d0 + d1 * d1 + k * k));
          int nk = bin_index(kmag, kmin, kmax, nbins, bintype);
          if (nk >= 0 && nk < nbins) {
            double gc = 1.0 / pow(gx * gy * gz, 4.0) * norm;
            const double *v = dk + 2 * cidx(N, i, j, k);
            const double w = OR_PLANT == 3 ? 1.0 : 2.0;
            pofk_t[nk] += w * (v[0] * v[0] + v[1] * v[1]) * gc;
            kmean_t[nk] += w * kmag;
            nmodes_t[nk] += w;
            if (j != N / 2 && j != 0) {
              v = dk + 2 * cidx(N, i, N - j, k);
              pofk_t[nk] += w * (v[0] * v[0] + v[1] * v[1]) * gc;
              kmean_t[nk] += w * kmag;
              nmodes_t[nk] += w;
            }
          }
        }
      }
    }
  }
  for (int t = 0; t < nth; t++)
    for (int b = 0; b < nbins; b++) {
      pofk[b] += acc[(size_t)t * 3 * nbins + b];
      kmean[b] += acc[(size_t)t * 3 * nbins + nbins + b];
      nmodes[b] += acc[(size_t)t * 3 * nbins + 2 * nbins + b];
    }
  free(acc);
  for (int i = 0; i < nbins; i++) {
    if (nmodes[i] > 0) {
      pofk[i] = (pofk[i] / nmodes[i]) * pow(c->box, 3);
      if (subtract_shotnoise) pofk[i] -= pow(c->box / (double)c->nsample, 3);
      kmean[i] = (kmean[i] / nmodes[i]) * 2.0 * M_PI / c->box;
    }
  }
}

/* n^2 of a half-spectrum entry with the mg.h convention (src/mg.h:42-46) */
static inline double rk_mg(int N, int i, int j, int k) {
  int ii = i > N / 2 ? N - i : i;
  int jj = j > N / 2 ? N - j : j;
  return (double)(k * k + ii * ii + jj * jj);
}

static double smoothing_filter(int filter, double kR) {
  switch (filter) {                       /* src/mg.h:318-338 */
    case OR_FILTER_GAUSSIAN: return exp(-0.5 * kR * kR);
    case OR_FILTER_TOPHAT: if (kR < 1e-5) return 1.0; return 3.0 / (kR * kR * kR) * (sin(kR) - kR * cos(kR));
    case OR_FILTER_SHARPK: return kR < 1.0 ? 1.0 : 0.0;
    default: return 1.0;
  }
}

/* screening_factor_potential, src/user_defined_functions.h:725-752 */
static inline double screen_potential(const or_cfg *c, const or_mg *m, double phin) {
  if (!c->include_screening) return 1.0;
  if (phin >= 0.0) return 1.0;
  if (c->model == OR_FOFR || c->model == OR_MBETA) {
    double s = fabs(m->phi_crit / phin);
    return s > 1.0 ? 1.0 : s;
  }
  return 1.0;
}

/* screening_factor_density, src/user_defined_functions.h:757-773 */
static inline double screen_density(const or_cfg *c, const or_mg *m, double dens) {
  if (!c->include_screening) return 1.0;
  if (c->model != OR_DGP) return 1.0;
  double fac = m->dgp_fac * (1.0 + dens);
  if (fac < 1e-5) return 1.0;
  return 2.0 * (sqrt(1.0 + fac) - 1.0) / fac;
}

/* ------------------------------------------------------------------------
 * ComputeFifthForce dispatch (src/user_defined_functions.h:687-719) with
 * _PotentialScreening (src/mg.h:147-189), _DensityScreening (src/mg.h:197-260)
 * and _TimeDepGeffModels (src/mg.h:127-137).  Returns 1 when phik was written
 * (allocate_mg_arrays), 0 otherwise.
 * ---------------------------------------------------------------------- */
int or_mg_fifth_force(const or_cfg *c, const or_mg *m, double *dk, const double *dx, double *phik) {
  const int N = c->nmesh, H = N / 2 + 1;
  const size_t total = or_real_len(N);
  if (!c->mg_active) return 0;
  if (c->model == OR_BRANSDICKE) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < (size_t)N * N * H + (size_t)N * H; i++) { dk[2 * i] *= m->geff; dk[2 * i + 1] *= m->geff; }
    return 0;
  }
  double *one = (double *)malloc(sizeof(double) * total);
  if (c->model == OR_FOFR || c->model == OR_MBETA) {
    const double massterm2 = (OR_PLANT == 5 ? 1.0 : pow(m->a, 2)) * m->mass2 / pow((2.0 * OR_PI) * OR_INVERSE_H0_MPCH / c->box, 2);
    if (c->include_screening) {
      /* DivideByLaplacian, src/mg.h:21-64 */
      double nf = 1.0 / pow((double)N, 3);
      nf *= 1.5 * c->omega / m->a * pow(c->box / OR_INVERSE_H0_MPCH / (2.0 * OR_PI), 2);
#pragma omp parallel for schedule(static)
      for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
          for (int k = 0; k < H; k++) {
            size_t ind = cidx(N, i, j, k);
            if (i == 0 && j == 0 && k == 0) { one[0] = one[1] = 0.0; continue; }
            double KK = -1.0 / rk_mg(N, i, j, k);
            one[2 * ind] = nf * dk[2 * ind] * KK;
            one[2 * ind + 1] = nf * dk[2 * ind + 1] * KK;
          }
      or_fft_c2r_3d(N, one);
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < total; i++) phik[i] = dx[i] * screen_potential(c, m, one[i]);
    } else {
      memcpy(phik, dx, sizeof(double) * total);
    }
    or_fft_r2c_3d(N, phik);
    /* EffDensitykToPhiofk, src/mg.h:74-116 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
        for (int k = 0; k < H; k++) {
          size_t ind = cidx(N, i, j, k);
          if (i == 0 && j == 0 && k == 0) { phik[0] = phik[1] = 0.0; continue; }
          double RK = rk_mg(N, i, j, k);
          double KK = RK / (RK + massterm2);
          phik[2 * ind] = m->coupling * phik[2 * ind] * KK;
          phik[2 * ind + 1] = m->coupling * phik[2 * ind + 1] * KK;
        }
  } else if (c->model == OR_DGP) {
    if (!c->include_screening) {
      for (size_t i = 0; i < total; i++) phik[i] = dk[i] * m->coupling;
    } else {
      /* SmoothDensityField, src/mg.h:268-309 */
      const double nf = 1.0 / pow((double)N, 3);
#pragma omp parallel for schedule(static)
      for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
          for (int k = 0; k < H; k++) {
            size_t ind = cidx(N, i, j, k);
            double RK = rk_mg(N, i, j, k);
            double s = smoothing_filter(c->filter, sqrt(RK) * 2.0 * OR_PI / c->box * m->rsmooth) * nf;
            one[2 * ind] = dk[2 * ind] * s;
            one[2 * ind + 1] = dk[2 * ind + 1] * s;
          }
      or_fft_c2r_3d(N, one);
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < total; i++) phik[i] = dx[i] * (m->coupling * screen_density(c, m, one[i]));
      or_fft_r2c_3d(N, phik);
    }
  } else {
    free(one);
    return 0;
  }
  free(one);
  return 1;
}

/* ------------------------------------------------------------------------
 * Forces: Green's function -i n/n^2 / (N^3 2pi/Box), Nyquist = +N/2,
 * then 3 c2r and the ghost slab.  src/auxPM.c:437-555.
 * ---------------------------------------------------------------------- */
void or_pm_forces(const or_cfg *c, double *dk, const double *phik, double *f0, double *f1, double *f2) {
  const int N = c->nmesh, H = N / 2 + 1;
  const double Scale = 2. * M_PI / c->box;
  const size_t ncplx = (size_t)N * N * H;
  if (phik) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < ncplx + (size_t)N * H; i++) { dk[2 * i] += phik[2 * i]; dk[2 * i + 1] += phik[2 * i + 1]; }
  }
  double *F[3] = {f0, f1, f2};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < H; k++) {
        size_t ind = cidx(N, i, j, k);
        if (i == 0 && j == 0 && k == 0) {
          for (int a = 0; a < 3; a++) F[a][0] = F[a][1] = 0.0;
          continue;
        }
        int dd[3];
        dd[0] = (OR_PLANT == 1 ? i >= N / 2 : i > N / 2) ? i - N : i;
        dd[1] = (OR_PLANT == 1 ? j >= N / 2 : j > N / 2) ? j - N : j;
        dd[2] = k;
        double RK = (double)(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
        double KK = -1.0 / RK;
        double d0 = (dk[2 * ind] * KK * 1.0) / pow((double)N, 3);
        double d1 = (-1.0 * dk[2 * ind + 1] * KK * 1.0) / pow((double)N, 3);
        for (int a = 0; a < 3; a++) {
          F[a][2 * ind] = d1 * dd[a] / Scale;
          F[a][2 * ind + 1] = d0 * dd[a] / Scale;
        }
      }
  const size_t slice = (size_t)N * 2 * H;
  for (int a = 0; a < 3; a++) {
    or_fft_c2r_3d(N, F[a]);
    memcpy(F[a] + (size_t)N * slice, F[a], sizeof(double) * slice);
  }
}

/* ------------------------------------------------------------------------
 * MtoParticles: CIC gather (WPAR = 1) and the particle mean, src/auxPM.c:560-644.
 * ---------------------------------------------------------------------- */
void or_pm_gather(const or_cfg *c, size_t np, const float *const pos[3],
                  const double *f0, const double *f1, const double *f2, float *const disp[3], double sumD[3]) {
  const int N = c->nmesh;
  const double scaleBox = (double)N / c->box;
  const double *F[3] = {f0, f1, f2};
  double s0 = 0, s1 = 0, s2 = 0;
#pragma omp parallel for reduction(+ : s0, s1, s2) schedule(static)
  for (size_t i = 0; i < np; i++) {
    double X = pos[0][i] * scaleBox, Y = pos[1][i] * scaleBox, Z = pos[2][i] * scaleBox;
    unsigned IX = (unsigned)X, IY = (unsigned)Y, IZ = (unsigned)Z;
    double DX = X - (double)IX, DY = Y - (double)IY, DZ = Z - (double)IZ;
    double TX = 1.0 - DX, TY = 1.0 - DY, TZ = 1.0 - DZ;
    if (IY >= (unsigned)N) IY = 0;
    if (IZ >= (unsigned)N) IZ = 0;
    unsigned IXn = IX + 1, IYn = IY + 1, IZn = IZ + 1;
    if (IYn >= (unsigned)N) IYn = 0;
    if (IZn >= (unsigned)N) IZn = 0;
    double d[3];
    for (int a = 0; a < 3; a++) {
      const double *g = F[a];
      d[a] = g[ridx(N, IX, IY, IZ)] * TX * TY * TZ + g[ridx(N, IX, IY, IZn)] * TX * TY * DZ +
             g[ridx(N, IX, IYn, IZ)] * TX * DY * TZ + g[ridx(N, IX, IYn, IZn)] * TX * DY * DZ +
             g[ridx(N, IXn, IY, IZ)] * DX * TY * TZ + g[ridx(N, IXn, IY, IZn)] * DX * TY * DZ +
             g[ridx(N, IXn, IYn, IZ)] * DX * DY * TZ + g[ridx(N, IXn, IYn, IZn)] * DX * DY * DZ;
      disp[a][i] = (float)d[a];
    }
    /* the reference sums the stored (float) values */
    s0 += disp[0][i];
    s1 += disp[1][i];
    s2 += disp[2][i];
  }
  sumD[0] = s0 / (double)np;
  sumD[1] = s1 / (double)np;
  sumD[2] = s2 / (double)np;
}

/* COLA term of particle n, axis a: the reference forms D*g1 + D2*g2 in double
 * (src/main.c:728, src/main.c:781) or, scale-dependent, the float sum D + D2
 * of the stored fields (src/main.c:712, src/main.c:767). */
static inline double lpt_term(const float *const d1[3], const float *const d2[3], double s1, double s2,
                              int sd, int a, size_t n) {
  if (!d1) return 0.0;
  if (sd) return (double)(float)(d1[a][n] + d2[a][n]);
  return d1[a][n] * s1 + d2[a][n] * s2;
}

/* ------------------------------------------------------------------------
 * Kick, src/main.c:688-742.
 * ---------------------------------------------------------------------- */
void or_kick(size_t np, float *const vel[3], float *const disp[3], const float *const d1[3],
             const float *const d2[3], double s1, double s2, int sd, double omega, double dda, double A,
             int use_cola, const double sumD[3], double sumV[3]) {
  double s0 = 0, s1v = 0, s2v = 0;
#pragma omp parallel for reduction(+ : s0, s1v, s2v) schedule(static)
  for (size_t n = 0; n < np; n++)
    for (int a = 0; a < 3; a++) {
      disp[a][n] -= sumD[a];
      double force = -1.5 * omega * disp[a][n] + (OR_PLANT == 8 ? 1 : -1) * use_cola * lpt_term(d1, d2, s1, s2, sd, a, n) / A;
      vel[a][n] += force * dda;
      if (a == 0) s0 += vel[a][n];
      else if (a == 1) s1v += vel[a][n];
      else s2v += vel[a][n];
    }
  sumV[0] = s0 / (double)np;
  sumV[1] = s1v / (double)np;
  sumV[2] = s2v / (double)np;
}

/* periodic_wrap (float version, MEMORY_MODE), src/auxPM.c:649-655 */
static inline float wrapf(float x, double box) {
  while (x >= (float)box) x -= (float)box;
  while (x < 0) x += (float)box;
  if (x == (float)box) x = 0.0f;
  return x;
}

/* Drift, src/main.c:747-787 */
void or_drift(size_t np, float *const pos[3], const float *const vel[3], const float *const d1[3],
              const float *const d2[3], double s1, double s2, int sd, double box, double dyyy, int use_cola,
              const double sumV[3]) {
#pragma omp parallel for schedule(static)
  for (size_t n = 0; n < np; n++)
    for (int a = 0; a < 3; a++) {
      pos[a][n] += (vel[a][n] - sumV[a]) * dyyy;
      pos[a][n] = wrapf((float)(pos[a][n] + use_cola * lpt_term(d1, d2, s1, OR_PLANT == 9 ? s1 : s2, sd, a, n)), box);
    }
}

/* ------------------------------------------------------------------------
 * PtoMesh_RSD, src/compute_pofk.c:279-395: the CIC deposit with axis
 * [axis] (1 = y, 2 = z) moved to redshift space.  Mesh coordinates are
 * (Pos0, Pos1, Pos2) for axis z and (Pos0, Pos2, Pos1) for axis y; the third
 * gets + vnorm * V with one periodic wrap, vnorm = (Hubble/A)/(100 A hubble(A)) * N/Box.
 * V = Vel[axis] (+ LPT velocity with COLA): scale-dependent, l1 holds dD/dy summed
 * over both orders (P.dDdy + P.dD2dy, a float sum); otherwise D dDdy(A) + D2 dD2dy(A)
 * with s1 = dDdy(A), s2 = dD2dy(A), l0 = D, l1 = D2.
 * ---------------------------------------------------------------------- */
void or_pm_deposit_rsd(const or_cfg *c, size_t np, const float *const pos[3], const float *const vel[3],
                       const float *const l0[3], const float *const l1[3], int axis, double vnorm,
                       int scale_dependent, int use_cola, double s1, double s2, double *density) {
  const int N = c->nmesh;
  const double scaleBox = (double)N / c->box;
  const double WPAR = pow((double)N / (double)c->nsample, 3);
  if (axis != 1 && axis != 2) axis = 2;
  const int ya = axis == 1 ? 2 : 1;
  const size_t total = or_real_len(N);
  for (size_t i = 0; i < total; i++) density[i] = -1.0;
  for (size_t i = 0; i < np; i++) {
    double X = pos[0][i] * scaleBox, Y = pos[ya][i] * scaleBox, Z = pos[axis][i] * scaleBox;
    double V;
    if (!use_cola) V = vel[axis][i];
    else if (scale_dependent) V = (float)(vel[axis][i] + l1[axis][i]);
    else V = vel[axis][i] + (l0[axis][i] * s1 + l1[axis][i] * s2);
    V *= vnorm;
    Z += V;
    if (Z >= (double)N) Z -= (double)N;
    if (Z < 0) Z += (double)N;
    unsigned IX = (unsigned)X, IY = (unsigned)Y, IZ = (unsigned)Z;
    double DX = X - (double)IX, DY = Y - (double)IY, DZ = Z - (double)IZ;
    double TX = 1.0 - DX, TY = 1.0 - DY, TZ = 1.0 - DZ;
    DY *= WPAR;
    TY *= WPAR;
    if (IY >= (unsigned)N) IY = 0;
    if (IZ >= (unsigned)N) IZ = 0;
    unsigned IXn = IX + 1, IYn = IY + 1, IZn = IZ + 1;
    if (IYn >= (unsigned)N) IYn = 0;
    if (IZn >= (unsigned)N) IZn = 0;
    density[ridx(N, IX, IY, IZ)] += TX * TY * TZ;
    density[ridx(N, IX, IY, IZn)] += TX * TY * DZ;
    density[ridx(N, IX, IYn, IZ)] += TX * DY * TZ;
    density[ridx(N, IX, IYn, IZn)] += TX * DY * DZ;
    density[ridx(N, IXn, IY, IZ)] += DX * TY * TZ;
    density[ridx(N, IXn, IY, IZn)] += DX * TY * DZ;
    density[ridx(N, IXn, IYn, IZ)] += DX * DY * TZ;
    density[ridx(N, IXn, IYn, IZn)] += DX * DY * DZ;
  }
  const size_t slice = (size_t)N * 2 * (N / 2 + 1);
  if (np != 0)
    for (size_t i = 0; i < slice; i++) density[i] += density[(size_t)N * slice + i] + 1.0;
}

/* k_from_index, src/compute_pofk.c:51-64 (h/Mpc) */
static double k_from_index(int index, double kmin, double kmax, int nbins, int bintype, double box) {
  if (bintype == 0) return (kmin + (kmax - kmin) / (double)nbins * index) * 2.0 * OR_PI / box;
  if (bintype == 1) return exp(log(kmin) + log(kmax / kmin) / (double)nbins * index) * 2.0 * OR_PI / box;
  return 0.0;
}

/* one mode of bin_up_RSD_power_spectrum: weight w, mu^2 = d2^2/k^2 */
static void rsd_add(const double *v, double gc, double w, double mu2, double kmag, int nk, double *p0, double *p2,
                    double *p4, double *kb, double *nb) {
  double pofk = (v[0] * v[0] + v[1] * v[1]) * gc;
  p0[nk] += w * pofk;
  p2[nk] += w * pofk * mu2;
  p4[nk] += w * pofk * mu2 * mu2;
  kb[nk] += w * kmag;
  nb[nk] += w;
}

/* ------------------------------------------------------------------------
 * bin_up_RSD_power_spectrum, src/compute_pofk.c:496-753: sums of |delta|^2,
 * mu^2 |delta|^2, mu^4 |delta|^2 (line of sight = mesh z) with the P(k) mode
 * weights and deconvolution, normalised by the mode counts and Box^3, turned
 * into the Legendre multipoles P0, P2, P4 (in that order of assignment), shot
 * noise subtracted from all three; k = k_from_index.  kmin/kmax in integer
 * wavenumbers as after adjust_pofk_parameters.
 * ---------------------------------------------------------------------- */
void or_rsd_multipoles(const or_cfg *c, const double *dk, int nbins, int bintype, double kmin, double kmax,
                       int subtract_shotnoise, double *P0, double *P2, double *P4, double *k, double *n) {
  const int N = c->nmesh;
  double *kb = calloc(nbins, sizeof(double));
  for (int i = 0; i < nbins; i++) P0[i] = P2[i] = P4[i] = n[i] = 0.0;
  const double norm = 1.0 / pow((double)N, 6);
  for (int i = 0; i < N; i++) {
    int d0 = i > N / 2 ? N - i : i;
    double gx = d0 == 0 ? 1.0 : sin((OR_PI * d0) / (double)N) / ((OR_PI * d0) / (double)N);
    for (int j = 0; j < N / 2 + 1; j++) {
      int d1 = j;
      double gy = d1 == 0 ? 1.0 : sin((OR_PI * d1) / (double)N) / ((OR_PI * d1) / (double)N);
      for (int pass = 0; pass < 2; pass++) {
        int d2 = pass == 0 ? 0 : N / 2;
        double gz = pass == 0 ? 1.0 : 2.0 / OR_PI;
        double kmag = sqrt((double)(d0 * d0 + d1 * d1 + d2 * d2));
        double mu2 = d2 * d2 / (kmag * kmag);
        if (pass == 0 && i == 0 && j == 0) mu2 = 0.0;
        int nk = bin_index(kmag, kmin, kmax, nbins, bintype);
        if (nk >= 0 && nk < nbins) {
          double gc = 1.0 / pow(gx * gy * gz, 4.0) * norm;
          rsd_add(dk + 2 * cidx(N, i, j, d2), gc, 1.0, mu2, kmag, nk, P0, P2, P4, kb, n);
          if (j != N / 2 && j != 0) rsd_add(dk + 2 * cidx(N, i, N - j, d2), gc, 1.0, mu2, kmag, nk, P0, P2, P4, kb, n);
        }
      }
      for (int kz = 1; kz < N / 2; kz++) {
        double gz = sin((OR_PI * kz) / (double)N) / ((OR_PI * kz) / (double)N);
        double kmag = sqrt((double)(d0 * d0 + d1 * d1 + kz * kz));
        double mu2 = kz * kz / kmag / kmag;
        int nk = bin_index(kmag, kmin, kmax, nbins, bintype);
        if (nk >= 0 && nk < nbins) {
          double gc = 1.0 / pow(gx * gy * gz, 4.0) * norm;
          rsd_add(dk + 2 * cidx(N, i, j, kz), gc, 2.0, mu2, kmag, nk, P0, P2, P4, kb, n);
          if (j != N / 2 && j != 0) rsd_add(dk + 2 * cidx(N, i, N - j, kz), gc, 2.0, mu2, kmag, nk, P0, P2, P4, kb, n);
        }
      }
    }
  }
  const double shot = pow(c->box / (double)c->nsample, 3), box3 = pow(c->box, 3);
  for (int i = 0; i < nbins; i++) {
    k[i] = 0.0;
    if (n[i] > 0) {
      P0[i] = (P0[i] / n[i]) * box3;
      P2[i] = (P2[i] / n[i]) * box3;
      P4[i] = (P4[i] / n[i]) * box3;
      P4[i] = 9.0 * (35.0 * P4[i] - 30.0 * P2[i] + 3.0 * P0[i]) / 8.0;
      P2[i] = 5.0 * (3.0 * P2[i] - 1.0 * P0[i]) / 2.0;
      if (subtract_shotnoise) {
        P0[i] -= shot;
        P2[i] -= shot;
        P4[i] -= shot;
      }
      k[i] = k_from_index(i, kmin, kmax, nbins, bintype, c->box);
    }
  }
  free(kb);
}
