/*
 * Oracle: 2LPT initial conditions and scale-dependent LPT fields of MG-PICOLA,
 * restated in FP64 C (test infrastructure only).
 *
 * The random stream is GSL's ranlxd1, which the reference allocates at
 * src/2LPT.c:259.  GSL is not vendored in the reference; its generator is
 * Luescher's RANLUX in the double-precision form GSL ships as ranlxd (48-bit
 * subtract-with-borrow, lags 5 and 12, luxury p = 202 values generated per 12
 * returned, seeding by a 31-bit shift register).  It is restated here from that
 * published description (GSL >= 2.0, as the reference requires).
 */
#include "oracle.h"
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI 3.14159265358979323846

static inline size_t ridx(int n, int ix, int iy, int iz) { return ((size_t)ix * n + iy) * 2 * (n / 2 + 1) + iz; }
static inline size_t cidx(int n, int ix, int iy, int iz) { return ((size_t)ix * n + iy) * (n / 2 + 1) + iz; }

/* ------------------------------- RANLUX ------------------------------- */
static const double kOneBit = 1.0 / 281474976710656.0; /* 2^-48 */

void or_ranlxd_seed(or_ranlxd *s, unsigned long seed) {
  int xbit[31];
  if (seed == 0) seed = 1;
  long i = (long)(seed & 0xFFFFFFFFUL);
  for (int k = 0; k < 31; k++) { xbit[k] = (int)(i % 2); i /= 2; }
  int ibit = 0, jbit = 18;
  for (int k = 0; k < 12; k++) {
    double x = 0;
    for (int l = 1; l <= 48; l++) {
      double y = (double)((xbit[ibit] + 1) % 2);
      x += x + y;
      xbit[ibit] = (xbit[ibit] + xbit[jbit]) % 2;
      ibit = (ibit + 1) % 31;
      jbit = (jbit + 1) % 31;
    }
    s->x[k] = kOneBit * x;
  }
  s->carry = 0;
  s->ir = 11;
  s->jr = 7;
  s->ir_old = 0;
  s->pr = 202;
}

/* advance the lagged subtract-with-borrow x_n = x_{n-5} - x_{n-12} - c (mod 1) by pr steps */
static void ranlxd_advance(or_ranlxd *s) {
  unsigned ir = s->ir, jr = s->jr;
  double carry = s->carry;
  for (unsigned k = 0; k < s->pr; k++) {
    double y = s->x[jr] - s->x[ir] - carry;
    if (y < 0) { carry = kOneBit; y += 1; } else carry = 0;
    s->x[ir] = y;
    ir = ir == 11 ? 0 : ir + 1;
    jr = jr == 11 ? 0 : jr + 1;
  }
  s->ir = ir;
  s->ir_old = ir;
  s->jr = jr;
  s->carry = carry;
}

double or_ranlxd_uniform(or_ranlxd *s) {
  unsigned ir = s->ir;
  s->ir = ir == 11 ? 0 : ir + 1;
  if (s->ir == s->ir_old) ranlxd_advance(s);
  return s->x[s->ir];
}

unsigned long or_ranlxd_get(or_ranlxd *s) { return (unsigned long)(or_ranlxd_uniform(s) * 4294967296.0); }

/* seed table, src/2LPT.c:259-271: shells i < N/2 measured from the grid edges */
void or_seedtable(int N, unsigned long seed, unsigned int *t) {
  or_ranlxd r;
  or_ranlxd_seed(&r, seed);
#define DRAW ((unsigned int)(0x7fffffff * or_ranlxd_uniform(&r)))
  for (int i = 0; i < N / 2; i++) {
    for (int j = 0; j < i; j++) t[i * N + j] = DRAW;
    for (int j = 0; j < i + 1; j++) t[j * N + i] = DRAW;
    for (int j = 0; j < i; j++) t[(N - 1 - i) * N + j] = DRAW;
    for (int j = 0; j < i + 1; j++) t[(N - 1 - j) * N + i] = DRAW;
    if (OR_PLANT == 11) {
      for (int j = 0; j < i + 1; j++) t[j * N + (N - 1 - i)] = DRAW;
      for (int j = 0; j < i; j++) t[i * N + (N - 1 - j)] = DRAW;
    } else {
      for (int j = 0; j < i; j++) t[i * N + (N - 1 - j)] = DRAW;
      for (int j = 0; j < i + 1; j++) t[j * N + (N - 1 - i)] = DRAW;
    }
    for (int j = 0; j < i; j++) t[(N - 1 - i) * N + (N - 1 - j)] = DRAW;
    for (int j = 0; j < i + 1; j++) t[(N - 1 - j) * N + (N - 1 - i)] = DRAW;
  }
#undef DRAW
}

/* 2LPT.c wavenumber convention: index < N/2 -> index, else -(N-index) (src/2LPT.c:1229-1245) */
static inline double kcomp(int N, int i, double box) {
  return i < N / 2 ? i * 2 * OR_PI / box : -(N - i) * 2 * OR_PI / box;
}

/* ------------------------------------------------------------------------
 * Gaussian random field delta(k) at z = 0, src/2LPT.c:337-495 (stored as the
 * SCALEDEPENDENT build stores cdelta_cdm; the non-SD build forms i k/k^2 delta
 * directly from the same numbers).
 * ---------------------------------------------------------------------- */
void or_ic_delta(const or_cfg *c, unsigned long seed, int sphere_mode, int amp_fixed, int inverted,
                 const double *pk_n2, double *cd) {
  const int N = c->nmesh, Np = c->nsample;
  const double box = c->box;
  unsigned int *seedtable = (unsigned int *)malloc(sizeof(unsigned int) * (size_t)N * N);
  or_seedtable(N, seed, seedtable);
  memset(cd, 0, sizeof(double) * or_real_len(N));
  for (int i = 0; i < N; i++) {
    for (int j = 0; j < N; j++) {
      or_ranlxd r;
      or_ranlxd_seed(&r, seedtable[i * N + j]);
      for (int k = 0; k < N / 2; k++) {
        if (OR_PLANT == 7 && i == 0 && j == 0 && k == 0) continue;
        double phase = or_ranlxd_uniform(&r) * 2 * OR_PI, ampl = 0.0;
        if (!amp_fixed) {
          do { ampl = or_ranlxd_uniform(&r); } while (ampl == 0);
        }
        if (i == N / 2 || j == N / 2 || k == N / 2) continue;
        if (i == 0 && j == 0 && k == 0) continue;
        double kv[3] = {kcomp(N, i, box), kcomp(N, j, box), kcomp(N, k, box)};
        double kmag2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        double kmag = sqrt(kmag2);
        if (sphere_mode == 1) {
          if (kmag * box / (2 * OR_PI) > Np / 2) continue;
        } else {
          if (fabs(kv[0]) * box / (2 * OR_PI) > Np / 2) continue;
          if (fabs(kv[1]) * box / (2 * OR_PI) > Np / 2) continue;
          if (fabs(kv[2]) * box / (2 * OR_PI) > Np / 2) continue;
        }
        int ii = i < N / 2 ? i : N - i, jj = j < N / 2 ? j : N - j;
        double p_of_k = pk_n2[ii * ii + jj * jj + k * k];
        if (!amp_fixed) p_of_k *= -log(ampl);
        double delta = pow(box, -1.5) * (OR_PLANT == 6 ? p_of_k : sqrt(p_of_k));
        if (inverted) delta *= -1.0;
        double re = delta * cos(phase), im = delta * sin(phase);
        if (k > 0) {
          size_t co = cidx(N, i, j, k);
          cd[2 * co] = re; cd[2 * co + 1] = im;
        } else if (i == 0) {
          if (j >= N / 2) continue;
          size_t co = cidx(N, 0, j, 0), cc = cidx(N, 0, N - j, 0);
          cd[2 * co] = re; cd[2 * co + 1] = im;
          cd[2 * cc] = re; cd[2 * cc + 1] = -im;
        } else {
          if (i >= N / 2) continue;
          int iim = N - i, jjm = N - j;
          if (jjm == N) jjm = 0;
          size_t co = cidx(N, i, j, 0), cc = cidx(N, iim, jjm, 0);
          cd[2 * co] = re; cd[2 * co + 1] = im;
          cd[2 * cc] = re; cd[2 * cc + 1] = -im;
        }
      }
    }
  }
  free(seedtable);
}

/* trilinear lattice read-out of a padded real mesh with ghost slab, src/2LPT.c:1407-1476 */
static double lattice_read(int N, int Np, const double *m, int n, int mm, int p) {
  double u = (double)(n * N) / (double)Np, v = (double)(mm * N) / (double)Np, w = (double)(p * N) / (double)Np;
  int i = (int)u, j = (int)v, k = (int)w;
  if (i == N) i = N - 1;
  if (j == N) j = N - 1;
  if (k == N) k = N - 1;
  u -= i; v -= j; w -= k;
  int ii = i + 1, jj = j + 1, kk = k + 1;
  if (jj >= N) jj -= N;
  if (kk >= N) kk -= N;
  double f1 = (1 - u) * (1 - v) * (1 - w), f2 = (1 - u) * (1 - v) * (w), f3 = (1 - u) * (v) * (1 - w);
  double f4 = (1 - u) * (v) * (w), f5 = (u) * (1 - v) * (1 - w), f6 = (u) * (1 - v) * (w);
  double f7 = (u) * (v) * (1 - w), f8 = (u) * (v) * (w);
  return m[ridx(N, i, j, k)] * f1 + m[ridx(N, i, j, kk)] * f2 + m[ridx(N, i, jj, k)] * f3 +
         m[ridx(N, i, jj, kk)] * f4 + m[ridx(N, ii, j, k)] * f5 + m[ridx(N, ii, j, kk)] * f6 +
         m[ridx(N, ii, jj, k)] * f7 + m[ridx(N, ii, jj, kk)] * f8;
}

static void ghost_copy(int N, double *m) {
  const size_t slice = (size_t)N * 2 * (N / 2 + 1);
  memcpy(m + (size_t)N * slice, m, sizeof(double) * slice);
}

/* ------------------------------------------------------------------------
 * Displacement gradient, 2LPT source and 2nd-order displacement,
 * src/2LPT.c:1204-1520.  With scale_dependent the routine stores
 * cdelta2 = -source(k) and returns (src/2LPT.c:1336-1373).
 * ---------------------------------------------------------------------- */
void or_ic_displacements(const or_cfg *c, const double *cd, int scale_dependent,
                         float *const za[3], float *const lpt[3], double *cdelta2) {
  const int N = c->nmesh, Np = c->nsample, H = N / 2 + 1;
  const double box = c->box;
  const size_t len = or_real_len(N);
  double *cdisp[3], *dig[6];
  for (int a = 0; a < 3; a++) cdisp[a] = (double *)calloc(len, sizeof(double));
  for (int a = 0; a < 6; a++) dig[a] = (double *)calloc(len, sizeof(double));
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < H; k++) {
        size_t co = cidx(N, i, j, k);
        double kv[3] = {kcomp(N, i, box), kcomp(N, j, box), kcomp(N, k, box)};
        double kmag2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        double re = cd[2 * co], im = cd[2 * co + 1];
        if (re == 0.0 && im == 0.0) continue;
        for (int a = 0; a < 3; a++) {
          cdisp[a][2 * co] = -kv[a] / kmag2 * im;
          cdisp[a][2 * co + 1] = kv[a] / kmag2 * re;
        }
      }
  /* d(dis_i)/d(q_j) -> i k_j dis_i: pairs (0,0) (0,1) (0,2) (1,1) (1,2) (2,2) */
  static const int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {0, 1, 2, 1, 2, 2};
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < H; k++) {
        size_t co = cidx(N, i, j, k);
        double kv[3] = {kcomp(N, i, box), kcomp(N, j, box), kcomp(N, k, box)};
        for (int g = 0; g < 6; g++) {
          dig[g][2 * co] = -cdisp[pa[g]][2 * co + 1] * kv[pb[g]];
          dig[g][2 * co + 1] = cdisp[pa[g]][2 * co] * kv[pb[g]];
        }
      }
  for (int g = 0; g < 6; g++) or_fft_c2r_3d(N, dig[g]);
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < N; k++) {
        size_t co = ridx(N, i, j, k);
        dig[3][co] = dig[0][co] * (dig[3][co] + dig[5][co]) + dig[3][co] * dig[5][co] - dig[1][co] * dig[1][co] -
                     dig[2][co] * dig[2][co] - dig[4][co] * dig[4][co];
      }
  or_fft_r2c_3d(N, dig[3]);
  double *src = dig[3];
  if (scale_dependent) {
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
        for (int k = 0; k < H; k++) {
          size_t co = cidx(N, i, j, k);
          double kv[3] = {kcomp(N, i, box), kcomp(N, j, box), kcomp(N, k, box)};
          double kmag2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
          if (kmag2 > 0.0) { cdelta2[2 * co] = -src[2 * co]; cdelta2[2 * co + 1] = -src[2 * co + 1]; }
          else cdelta2[2 * co] = cdelta2[2 * co + 1] = 0.0;
        }
  } else {
    double *cdisp2[3] = {dig[0], dig[1], dig[2]};
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
        for (int k = 0; k < H; k++) {
          size_t co = cidx(N, i, j, k);
          double kv[3] = {kcomp(N, i, box), kcomp(N, j, box), kcomp(N, k, box)};
          double kmag2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
          for (int a = 0; a < 3; a++) {
            if (kmag2 > 0.0) {
              cdisp2[a][2 * co] = src[2 * co + 1] * kv[a] / kmag2;
              cdisp2[a][2 * co + 1] = -src[2 * co] * kv[a] / kmag2;
            } else cdisp2[a][2 * co] = cdisp2[a][2 * co + 1] = 0.0;
          }
        }
    for (int a = 0; a < 3; a++) {
      or_fft_c2r_3d(N, cdisp[a]);
      or_fft_c2r_3d(N, cdisp2[a]);
      ghost_copy(N, cdisp[a]);
      ghost_copy(N, cdisp2[a]);
    }
    const double nmesh3 = (double)N * N * N;
    const size_t np = (size_t)Np * Np * Np;
    double s1[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
    for (int n = 0; n < Np; n++)
      for (int m = 0; m < Np; m++)
        for (int p = 0; p < Np; p++) {
          size_t q = ((size_t)n * Np + m) * Np + p;
          for (int a = 0; a < 3; a++) {
            double dis = lattice_read(N, Np, cdisp[a], n, m, p);
            double dis2 = lattice_read(N, Np, cdisp2[a], n, m, p) / nmesh3;
            za[a][q] = (float)dis;
            const double f37 = OR_PLANT == 2 ? 3. / 7. : -3. / 7.;
            lpt[a][q] = (float)(f37 * dis2);
            s1[a] += dis;
            s2[a] += f37 * dis2;
          }
        }
    for (int a = 0; a < 3; a++) {
      s1[a] /= (double)np;
      s2[a] /= (double)np;
      for (size_t q = 0; q < np; q++) { za[a][q] -= s1[a]; lpt[a][q] -= s2[a]; }
    }
  }
  for (int a = 0; a < 3; a++) free(cdisp[a]);
  for (int g = 0; g < 6; g++) free(dig[g]);
}

/* ------------------------------------------------------------------------
 * from_cdisp_store_to_ZA (src/2LPT.c:1539-1735) + the single-rank assignment
 * of assign_displacment_field_to_particles (src/2LPT.c:1758-1829).
 * g_n2[n2] = normfactor * growth(k = 2pi/Box sqrt(n2)).
 * ---------------------------------------------------------------------- */
void or_lpt_field(const or_cfg *c, const double *cd, const double *g_n2, float *const out[3]) {
  const int N = c->nmesh, Np = c->nsample, H = N / 2 + 1;
  const double box = c->box;
  const size_t len = or_real_len(N);
  double *f[3];
  for (int a = 0; a < 3; a++) f[a] = (double *)calloc(len, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < H; k++) {
        size_t co = cidx(N, i, j, k);
        double kv[3] = {kcomp(N, i, box), kcomp(N, j, box), kcomp(N, k, box)};
        double kmag2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        int ii = i < N / 2 ? i : N - i, jj = j < N / 2 ? j : N - j, kk = k < N / 2 ? k : N - k;
        if (OR_PLANT == 10) ii = i % (N / 2 + 1);
        double g = g_n2[ii * ii + jj * jj + kk * kk];
        for (int a = 0; a < 3; a++) {
          if (kmag2 > 0.0) {
            f[a][2 * co] = -cd[2 * co + 1] * kv[a] / kmag2 * g;
            f[a][2 * co + 1] = cd[2 * co] * kv[a] / kmag2 * g;
          }
        }
      }
  const size_t np = (size_t)Np * Np * Np;
  double *tmp = (double *)malloc(sizeof(double) * np);
  for (int a = 0; a < 3; a++) {
    or_fft_c2r_3d(N, f[a]);
    ghost_copy(N, f[a]);
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int n = 0; n < Np; n++)
      for (int m = 0; m < Np; m++)
        for (int p = 0; p < Np; p++) {
          size_t q = ((size_t)n * Np + m) * Np + p;
          tmp[q] = lattice_read(N, Np, f[a], n, m, p);
          s += tmp[q];
        }
    s /= (double)np;
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < np; q++) out[a][q] = (float)(tmp[q] - s);
    free(f[a]);
  }
  free(tmp);
}
