// 2LPT initial conditions and scale-dependent LPT fields for gfx950.
//
//   k_ic_delta         displacement_fields' Gaussian mode loop (src/2LPT.c:337-495):
//                      one thread per (i, j) column re-seeds RANLUX from the seed
//                      table and draws (phase, amplitude) for kz < N/2 in order.
//   k_lpt_spectrum     i k_a/k^2 delta(k) g(|k|) and the displacement-gradient
//                      spectra (src/2LPT.c:1224-1277, 1311-1373, 1583-1629),
//                      2LPT.c wavenumbers (Nyquist = -N/2).
//   k_lattice_read     trilinear read-out at the Lagrangian lattice
//                      (src/2LPT.c:1407-1476, 1657-1702) + per-block sums for the
//                      mean removal (src/2LPT.c:1493-1508, 1713-1729).
#include "mgpu_internal.h"
#include "ranlux.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

// Seed table staging of the allocation-free IC entry point: one copy per GPU,
// held until the process ends.  Host code never passes `scratch` itself to a
// kernel or a copy (the name is only a host-side stand-in for the symbol); it
// uses the per-device address that ic_scratch_init obtains with
// hipGetSymbolAddress.
static __device__ __align__(256) unsigned scratch[MGPU_IC_SCRATCH_SEEDS];

namespace {

constexpr int kBlock = 256;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned *g_scratch_ptr[kMaxDevices] = {};

// one thread per column (i, j) of the N x N grid
__global__ __launch_bounds__(kBlock) void k_ic_delta(const unsigned *__restrict__ seedtable, int N, int Np, int HP,
                                                     double box, int sphere_mode, int amp_fixed, int inverted,
                                                     const double *__restrict__ pk_n2, float2 *__restrict__ cd,
                                                     size_t zs, size_t seed_stride, const int *__restrict__ inv_z) {
  // a batch: realization blockIdx.z's seed table and inversion flag
  cd = zbuf(cd, zs);
  seedtable += (size_t)blockIdx.z * seed_stride;
  if (inv_z) inverted = inv_z[blockIdx.z];
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= N * N) return;
  const int i = col / N, j = col % N;
  Ranlxd1 r;
  r.seed(seedtable[col]);
  const double bnorm = pow(box, -1.5);
  for (int k = 0; k < N / 2; k++) {
    const double phase = r.uniform() * 2 * MGPU_PI;
    double ampl = 0.0;
    if (!amp_fixed) {
      do {
        ampl = r.uniform();
      } while (ampl == 0);
    }
    if (i == N / 2 || j == N / 2 || k == N / 2) continue;
    if (i == 0 && j == 0 && k == 0) continue;
    // src/2LPT.c:364-386: k-vector with index < N/2 -> index, else -(N-index); cube or sphere cut
    // same double expressions as the reference so the cut rounds identically
    const int ni = i < N / 2 ? i : -(N - i), nj = j < N / 2 ? j : -(N - j), nk = k;
    const double kx = ni * 2 * MGPU_PI / box, ky = nj * 2 * MGPU_PI / box, kz = nk * 2 * MGPU_PI / box;
    if (sphere_mode == 1) {
      const double kmag = sqrt(kx * kx + ky * ky + kz * kz);
      if (kmag * box / (2 * MGPU_PI) > Np / 2) continue;
    } else {
      if (fabs(kx) * box / (2 * MGPU_PI) > Np / 2) continue;
      if (fabs(ky) * box / (2 * MGPU_PI) > Np / 2) continue;
      if (fabs(kz) * box / (2 * MGPU_PI) > Np / 2) continue;
    }
    double p_of_k = pk_n2[ni * ni + nj * nj + nk * nk];
    if (!amp_fixed) p_of_k *= -log(ampl);
    double delta = bnorm * sqrt(p_of_k);
    if (inverted) delta *= -1.0;
    const float2 v = make_float2((float)(delta * cos(phase)), (float)(delta * sin(phase)));
    const float2 vc = make_float2(v.x, -v.y);
    if (k > 0) {
      cd[((size_t)i * N + j) * HP + k] = v;
    } else if (i == 0) {
      if (j >= N / 2) continue;
      cd[((size_t)0 * N + j) * HP] = v;
      cd[((size_t)0 * N + (N - j)) * HP] = vc;
    } else {
      if (i >= N / 2) continue;
      const int jj = (N - j) == N ? 0 : N - j;
      cd[((size_t)i * N + j) * HP] = v;
      cd[((size_t)(N - i) * N + jj) * HP] = vc;
    }
  }
}

// 2LPT.c wavenumber component of index i (Nyquist -> -N/2), in h/Mpc
__device__ __forceinline__ double kc(int i, int N, double kf) { return (i < N / 2 ? i : -(N - i)) * kf; }
__device__ __forceinline__ int nabs(int i, int N) { return i < N / 2 ? i : N - i; }

enum { SPEC_DISP = 0, SPEC_GRAD = 1, SPEC_DISP2 = 2, SPEC_FIELD = 3, SPEC_NEG = 4 };

// Spectral kernels of the IC / LPT fields.  mode:
//   SPEC_DISP   out = i k_a/k^2 cd                           (cdisp, src/2LPT.c:417-490)
//   SPEC_GRAD   out = i k_b (i k_a/k^2 cd) = -k_a k_b/k^2 cd  (cdigrad, src/2LPT.c:1247-1271)
//   SPEC_DISP2  out = -i k_a/k^2 src                        (cdisp2, src/2LPT.c:1359-1370)
//   SPEC_FIELD  out = i k_a/k^2 (g1 cd + g2 cd2)            (from_cdisp_store_to_ZA, both orders)
//   SPEC_NEG    out = -src, 0 at k = 0                      (cdelta_cdm2, src/2LPT.c:1336-1345)
__global__ __launch_bounds__(kBlock) void k_lpt_spectrum(int mode, const float2 *__restrict__ cd,
                                                         const float2 *__restrict__ cd2, float2 *__restrict__ out,
                                                         int N, int HP, double kf, int a, int b,
                                                         const double *__restrict__ g1,
                                                         const double *__restrict__ g2, size_t zs) {
  cd = zbuf(cd, zs);
  cd2 = zbuf(cd2, zs);
  out = zbuf(out, zs);
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)N * N * HP) return;
  const int k = (int)(e % HP);
  const size_t row = e / HP;
  const int j = (int)(row % N), i = (int)(row / N);
  float2 r = make_float2(0.f, 0.f);
  if (k <= N / 2) {
    const double kv[3] = {kc(i, N, kf), kc(j, N, kf), kc(k, N, kf)};
    const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
    if (k2 > 0.0) {
      double re = 0, im = 0;
      if (mode == SPEC_DISP || mode == SPEC_GRAD) {
        const float2 d = cd[e];
        // cdisp = (-k_a/k2 * im, k_a/k2 * re)
        re = -kv[a] / k2 * d.y;
        im = kv[a] / k2 * d.x;
        if (mode == SPEC_GRAD) {  // times i k_b
          const double r2 = -im * kv[b], i2 = re * kv[b];
          re = r2;
          im = i2;
        }
      } else if (mode == SPEC_DISP2) {
        const float2 s = cd[e];
        re = s.y * kv[a] / k2;
        im = -s.x * kv[a] / k2;
      } else if (mode == SPEC_FIELD) {
        const int n2 = nabs(i, N) * nabs(i, N) + nabs(j, N) * nabs(j, N) + nabs(k, N) * nabs(k, N);
        const float2 d = cd[e];
        double dr = d.x * g1[n2], di = d.y * g1[n2];
        if (cd2) {
          const float2 d2 = cd2[e];
          dr += d2.x * g2[n2];
          di += d2.y * g2[n2];
        }
        re = -kv[a] / k2 * di;
        im = kv[a] / k2 * dr;
      } else {  // SPEC_NEG
        const float2 s = cd[e];
        re = -s.x;
        im = -s.y;
      }
      r = make_float2((float)re, (float)im);
    }
  }
  out[e] = r;
}

// 2LPT source on padded real meshes: src = d00 (d11 + d22) + d11 d22 (first
// call), src -= d*d (off-diagonal calls).  src/2LPT.c:1282-1290
// src may be d00 (each element is read before it is written)
__global__ __launch_bounds__(kBlock) void k_source_diag(const float *d00, const float *__restrict__ d11,
                                                        const float *__restrict__ d22, float *src, size_t n,
                                                        size_t zs) {
  d00 = zbuf(d00, zs);
  d11 = zbuf(d11, zs);
  d22 = zbuf(d22, zs);
  src = zbuf(src, zs);
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double a = d00[e], b = d11[e], c = d22[e];
  src[e] = (float)(a * (b + c) + b * c);
}
__global__ __launch_bounds__(kBlock) void k_source_off(const float *__restrict__ d, float *__restrict__ src,
                                                       size_t n, size_t zs) {
  d = zbuf(d, zs);
  src = zbuf(src, zs);
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double v = d[e];
  src[e] = (float)((double)src[e] - v * v);
}

// trilinear lattice read-out, src/2LPT.c:1407-1476.  out[q] = scale * value
__global__ __launch_bounds__(kBlock) void k_lattice_read(const float *__restrict__ m, int N, int Np, int HP2,
                                                         double scale, float *__restrict__ out,
                                                         double *__restrict__ partial, size_t zs) {
  m = zbuf(m, zs);
  out = zbuf(out, zs);
  partial = zbuf(partial, zs);
  __shared__ double red[kBlock];
  const size_t np = (size_t)Np * Np * Np;
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double val = 0.0;
  if (q < np) {
    const int p = (int)(q % Np), mm = (int)((q / Np) % Np), n = (int)(q / ((size_t)Np * Np));
    double u = (double)(n * N) / (double)Np, v = (double)(mm * N) / (double)Np, w = (double)(p * N) / (double)Np;
    int i = (int)u, j = (int)v, k = (int)w;
    if (i >= N) i = N - 1;
    if (j >= N) j = N - 1;
    if (k >= N) k = N - 1;
    u -= i;
    v -= j;
    w -= k;
    const int ii = i + 1 >= N ? 0 : i + 1, jj = j + 1 >= N ? 0 : j + 1, kk = k + 1 >= N ? 0 : k + 1;
    auto at = [&](int x, int y, int z) { return (double)m[((size_t)x * N + y) * HP2 + z]; };
    val = at(i, j, k) * ((1 - u) * (1 - v) * (1 - w)) + at(i, j, kk) * ((1 - u) * (1 - v) * w) +
          at(i, jj, k) * ((1 - u) * v * (1 - w)) + at(i, jj, kk) * ((1 - u) * v * w) +
          at(ii, j, k) * (u * (1 - v) * (1 - w)) + at(ii, j, kk) * (u * (1 - v) * w) +
          at(ii, jj, k) * (u * v * (1 - w)) + at(ii, jj, kk) * (u * v * w);
    val *= scale;
    out[q] = (float)val;
  }
  red[threadIdx.x] = val;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(size_t)blockIdx.x * 3] = red[0];
}

__global__ __launch_bounds__(kBlock) void k_sub_mean(float *__restrict__ v, size_t n, const double *__restrict__ sum,
                                                     double inv_n, size_t zs) {
  v = zbuf(v, zs);
  sum = zbuf(sum, zs);
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  v[q] = (float)((double)v[q] - sum[0] * inv_n);
}

// sum partial[b*3] over blocks in a fixed order -> out[0]
__global__ __launch_bounds__(1024) void k_reduce1(const double *__restrict__ partial, int nb, double *__restrict__ out,
                                                  size_t zs) {
  partial = zbuf(partial, zs);
  out = zbuf(out, zs);
  __shared__ double red[1024];
  double s = 0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += partial[(size_t)b * 3];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// particle init, src/main.c:257-309
__global__ __launch_bounds__(kBlock) void k_init_particles(float *__restrict__ pos, float *__restrict__ vel,
                                                           const float *__restrict__ l0,
                                                           const float *__restrict__ l1, int Np, double box,
                                                           int sd, double Di, double Di2, int use_cola,
                                                           double v1, double v2) {
  const size_t np = (size_t)Np * Np * Np;
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const int c[3] = {(int)(q / ((size_t)Np * Np)), (int)((q / Np) % Np), (int)(q % Np)};
  const float fb = (float)box;
  for (int a = 0; a < 3; a++) {
    const size_t idx = (size_t)a * np + q;
    double x = c[a] * (box / (double)Np);
    if (sd) x += (double)l0[idx];
    else x += (double)l0[idx] * Di + (double)l1[idx] * Di2;
    float f = (float)x;
    for (int it = 0; it < 4 && f >= fb; ++it) f -= fb;
    for (int it = 0; it < 4 && f < 0.f; ++it) f += fb;
    if (!(f >= 0.f && f < fb)) f = 0.f;
    pos[idx] = f;
    if (use_cola) vel[idx] = 0.f;
    else vel[idx] = sd ? l1[idx] : (float)((double)l0[idx] * v1 + (double)l1[idx] * v2);
  }
}

// ---------------------------------------------------------------- lattice fold
// LPT field components X_a = i k_a/k^2 (g1 cd1 + g2 cd2) at half-spectrum index
// (i, j, l), l in [0, N/2], 2LPT.c wavenumbers (src/2LPT.c:1583-1629), in FP32
// (the stored spectra and the output fields are FP32).  ikf = 1 / (2 pi/Box).
// Growth sources of the fold, for NS LPT slots at once: eval(n2, y, g) sets
// g[s][o] = growth of order o + 1 for slot s at |n|^2 = n2, y = ln k = ln kf +
// ln|n| (only the slice source uses y).
// From FP32 per-|n|^2 tables (one slot):
struct GTab {
  static constexpr bool kNeedsLnk = false;
  const float *t1, *t2;
  __device__ __forceinline__ void eval(int n2, float, float g[1][2]) const {
    g[0][0] = t1[n2];
    g[0][1] = t2[n2];
  }
};
// From cubic Hermite slices on one uniform ln k grid shared by all slices,
// held in LDS node-major with 1 + NS float4 per node: (den_1, dden_1, den_2,
// dden_2), shared by the slots, then per slot (num_1, dnum_1, num_2, dnum_2);
// g = norm_o num / den.  The Hermite basis is computed once per mode.
template <int NS>
struct GShared {
  static constexpr bool kNeedsLnk = true;
  const float4 *sl;
  float y0, dy, idy, ymax, norm1, norm2;
  int ny;
  __device__ __forceinline__ void eval(int, float y, float g[NS][2]) const {
    y = fminf(fmaxf(y, y0), ymax);
    int j = (int)((y - y0) * idy);
    j = min(max(j, 0), ny - 2);
    const float t = (y - (y0 + (float)j * dy)) * idy, t2 = t * t, t3 = t2 * t;
    const float h00 = 2.f * t3 - 3.f * t2 + 1.f, h10 = (t3 - 2.f * t2 + t) * dy, h01 = -2.f * t3 + 3.f * t2,
                h11 = (t3 - t2) * dy;
    const float4 *A = sl + j * (1 + NS), *B = A + (1 + NS);
    float4 a = A[0], b = B[0];
    const float r1 = norm1 * __builtin_amdgcn_rcpf(h00 * a.x + h10 * a.y + h01 * b.x + h11 * b.y);
    const float r2 = norm2 * __builtin_amdgcn_rcpf(h00 * a.z + h10 * a.w + h01 * b.z + h11 * b.w);
#pragma unroll
    for (int s = 0; s < NS; s++) {
      a = A[1 + s];
      b = B[1 + s];
      g[s][0] = r1 * (h00 * a.x + h10 * a.y + h01 * b.x + h11 * b.y);
      g[s][1] = r2 * (h00 * a.z + h10 * a.w + h01 * b.z + h11 * b.w);
    }
  }
};

// LPT field components X[s][a] = i k_a/k^2 (g1 cd1 + g2 cd2) of slot s at half-spectrum
// index (i, j, l), l in [0, N/2], 2LPT.c wavenumbers (src/2LPT.c:1583-1629), in FP32
// (the stored spectra and the output fields are FP32).  ikf = 1 / (2 pi/Box).
template <int NS, class G>
__device__ __forceinline__ void lpt_xf(int i, int j, int l, int N, int HP, int Np, float ikf, float lnkf,
                                       const float2 *__restrict__ cd1, const float2 *__restrict__ cd2, const G &gr,
                                       float X[NS][6]) {
  const int ni = i < N / 2 ? i : i - N, nj = j < N / 2 ? j : j - N, nl = l < N / 2 ? l : l - N;
  const int n2 = ni * ni + nj * nj + nl * nl;
#pragma unroll
  for (int s = 0; s < NS; s++)
#pragma unroll
    for (int q = 0; q < 6; q++) X[s][q] = 0.f;
  if (n2 == 0) return;
  const size_t e = ((size_t)i * N + j) * HP + l;
  // the 1LPT field is zero outside the initial-condition cube |n_a| <= Nsample/2
  const bool cube = abs(ni) <= Np / 2 && abs(nj) <= Np / 2 && abs(nl) <= Np / 2;
  const float2 d1 = cube ? cd1[e] : make_float2(0.f, 0.f);
  const float2 d2 = cd2 ? cd2[e] : make_float2(0.f, 0.f);
  float g[NS][2];
  gr.eval(n2, G::kNeedsLnk ? lnkf + 0.5f * __logf((float)n2) : 0.f, g);
  // k_a / k^2 = n_a / (kf n^2)
  const float sc = ikf * __builtin_amdgcn_rcpf((float)n2);
  const float nv[3] = {(float)ni * sc, (float)nj * sc, (float)nl * sc};
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const float dr = d1.x * g[s][0] + d2.x * g[s][1], di = d1.y * g[s][0] + d2.y * g[s][1];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      X[s][2 * q] = -nv[q] * di;
      X[s][2 * q + 1] = nv[q] * dr;
    }
  }
}

// Sampling the c2r of a field at every m-th mesh point (the Lagrangian lattice
// when N = m * Nsample, where the read-out of src/2LPT.c:1657-1702 has weight 1
// on one node) equals an Nsample^3 c2r of the aliased spectrum
//   F_f(k') = sum_{k = k' (mod Nsample)} F(k),
// F the Hermitian extension of the half spectrum with the kz = 0, N/2 planes
// symmetrised as the c2r's real-part rule implies.  One thread per output mode;
// consecutive threads take consecutive kz, so every alias read is a coalesced
// row segment.  The folded DC mode is zeroed, which removes the lattice mean
// exactly (src/2LPT.c:1719-1729).
// One folded mode (ip, jp, lp), lp <= Np/2: the sum over its m^3 aliases, for
// NS slots.  MF > 0 is the fold factor m = N / Nsample as a compile-time
// constant, so the alias loop unrolls and all its loads are in flight
// together; MF = 0 takes m at run time.  PLANES = false is for 0 < lp < Np/2:
// then no alias lies on the kz = 0 or N/2 plane, and whether alias c lies
// below or above N/2 does not depend on lp, so the lanes of a wave take one
// path.
template <int MF, int NS, class G, bool PLANES = true>
__device__ __forceinline__ void fold_mode(int ip, int jp, int lp, int N, int HP, int Np, int m_rt, float ikf,
                                          float lnkf, const float2 *__restrict__ cd1, const float2 *__restrict__ cd2,
                                          const G &gr, float acc[NS][6]) {
  const int m = MF > 0 ? MF : m_rt;
  float X[NS][6], Y[NS][6];
#pragma unroll
  for (int s = 0; s < NS; s++)
#pragma unroll
    for (int q = 0; q < 6; q++) acc[s][q] = 0.f;
  if (ip == 0 && jp == 0 && lp == 0) return;
#pragma unroll
  for (int a = 0; a < m; a++) {
    const int i = ip + a * Np, mi = i == 0 ? 0 : N - i;
#pragma unroll
    for (int b = 0; b < m; b++) {
      const int j = jp + b * Np, mj = j == 0 ? 0 : N - j;
#pragma unroll
      for (int c = 0; c < m; c++) {
        const int l = lp + c * Np;
        if (l > 0 && l < N / 2) {
          lpt_xf<NS>(i, j, l, N, HP, Np, ikf, lnkf, cd1, cd2, gr, X);
#pragma unroll
          for (int s = 0; s < NS; s++)
#pragma unroll
            for (int q = 0; q < 6; q++) acc[s][q] += X[s][q];
        } else if (l > N / 2) {
          lpt_xf<NS>(mi, mj, N - l, N, HP, Np, ikf, lnkf, cd1, cd2, gr, X);
#pragma unroll
          for (int s = 0; s < NS; s++)
#pragma unroll
            for (int q = 0; q < 3; q++) {
              acc[s][2 * q] += X[s][2 * q];
              acc[s][2 * q + 1] -= X[s][2 * q + 1];
            }
        } else if (PLANES) {  // kz = 0 or N/2 plane: (X(k) + conj X(-k)) / 2
          lpt_xf<NS>(i, j, l, N, HP, Np, ikf, lnkf, cd1, cd2, gr, X);
          lpt_xf<NS>(mi, mj, l, N, HP, Np, ikf, lnkf, cd1, cd2, gr, Y);
#pragma unroll
          for (int s = 0; s < NS; s++)
#pragma unroll
            for (int q = 0; q < 3; q++) {
              acc[s][2 * q] += 0.5f * (X[s][2 * q] + Y[s][2 * q]);
              acc[s][2 * q + 1] += 0.5f * (X[s][2 * q + 1] - Y[s][2 * q + 1]);
            }
        }
      }
    }
  }
}

// growth from per-|n|^2 tables, one slot; grid (Nsample^2 rows, kz blocks)
template <int MF>
__global__ __launch_bounds__(kBlock) void k_lpt_fold(const float2 *__restrict__ cd1, const float2 *__restrict__ cd2,
                                                     GTab gr, int N, int HP, int Np, int HPn, int m_rt, float ikf,
                                                     float2 *__restrict__ out, size_t nl) {
  const int lp = blockIdx.y * blockDim.x + threadIdx.x;
  if (lp > Np / 2) return;
  const unsigned row = blockIdx.x;
  const int jp = (int)(row % (unsigned)Np), ip = (int)(row / (unsigned)Np);
  float acc[1][6];
  fold_mode<MF, 1>(ip, jp, lp, N, HP, Np, m_rt, ikf, 0.f, cd1, cd2, gr, acc);
  const size_t e = (size_t)row * HPn + lp;
#pragma unroll
  for (int a = 0; a < 3; a++) out[a * nl + e] = make_float2(acc[0][2 * a], acc[0][2 * a + 1]);
}

// growth from the shared-grid ln k slices of NS slots (node-major float4s,
// copied once per block into LDS).  The blocks stride over the Nsample^2 rows;
// the interior kz of a row first, then the kz = 0 and Nsample/2 modes of all
// rows.  Output spectrum (slot s, component a) at out + (3 s + a) nl.
template <int MF, int NS>
__global__ __launch_bounds__(kBlock) void k_lpt_fold_gs(const float2 *__restrict__ cd1,
                                                        const float2 *__restrict__ cd2,
                                                        const float4 *__restrict__ nodes, GShared<NS> gs,
                                                        float lnkf, int N, int HP, int Np, int HPn, int m_rt,
                                                        float ikf, float2 *__restrict__ out, size_t nl) {
  extern __shared__ float4 sl[];
  const int nn = gs.ny * (1 + NS);
  for (int e = threadIdx.x; e < nn; e += blockDim.x) sl[e] = nodes[e];
  __syncthreads();
  gs.sl = sl;
  const unsigned rows = (unsigned)Np * (unsigned)Np;
  const int H = Np / 2;
  for (unsigned row = blockIdx.x; row < rows; row += gridDim.x) {
    const int jp = (int)(row % (unsigned)Np), ip = (int)(row / (unsigned)Np);
    for (int lp = 1 + (int)threadIdx.x; lp < H; lp += blockDim.x) {
      float acc[NS][6];
      fold_mode<MF, NS, GShared<NS>, false>(ip, jp, lp, N, HP, Np, m_rt, ikf, lnkf, cd1, cd2, gs, acc);
      const size_t o = (size_t)row * HPn + lp;
#pragma unroll
      for (int s = 0; s < NS; s++)
#pragma unroll
        for (int a = 0; a < 3; a++) out[(3 * s + a) * nl + o] = make_float2(acc[s][2 * a], acc[s][2 * a + 1]);
    }
  }
  for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < 2 * rows; q += gridDim.x * blockDim.x) {
    const unsigned row = q >> 1;
    const int lp = (q & 1) ? H : 0;
    const int jp = (int)(row % (unsigned)Np), ip = (int)(row / (unsigned)Np);
    float acc[NS][6];
    fold_mode<MF, NS, GShared<NS>, true>(ip, jp, lp, N, HP, Np, m_rt, ikf, lnkf, cd1, cd2, gs, acc);
    const size_t o = (size_t)row * HPn + lp;
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
      for (int a = 0; a < 3; a++) out[(3 * s + a) * nl + o] = make_float2(acc[s][2 * a], acc[s][2 * a + 1]);
  }
}

__global__ __launch_bounds__(kBlock) void k_d2f(size_t n, const double *__restrict__ a, float *__restrict__ b) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// g_o(|n|^2) = norm_o num_o(ln k) / den_o(ln k) from cubic Hermite slices on a
// uniform ln k grid; t = [num, dnum, den, dden] (4*ny doubles)
struct KSl {
  double y0, dy, norm;
  int ny;
};
__device__ __forceinline__ double herm(const double *__restrict__ v, const double *__restrict__ dv, const KSl &s,
                                       double y) {
  const double ymax = s.y0 + s.dy * (s.ny - 1);
  if (y < s.y0) y = s.y0;
  if (y > ymax) y = ymax;
  int j = (int)floor((y - s.y0) / s.dy);
  if (j > s.ny - 2) j = s.ny - 2;
  if (j < 0) j = 0;
  const double h = s.dy, t = (y - (s.y0 + j * s.dy)) / h, t2 = t * t, t3 = t2 * t;
  return (2 * t3 - 3 * t2 + 1) * v[j] + (t3 - 2 * t2 + t) * h * dv[j] + (-2 * t3 + 3 * t2) * v[j + 1] +
         (t3 - t2) * h * dv[j + 1];
}
__global__ __launch_bounds__(kBlock) void k_growth_n2(size_t n2len, double kf, KSl s1, const double *__restrict__ t1,
                                                      KSl s2, const double *__restrict__ t2, double *__restrict__ g1,
                                                      double *__restrict__ g2, float *__restrict__ g1f,
                                                      float *__restrict__ g2f) {
  const size_t n2 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n2 >= n2len) return;
  double a = 0.0, b = 0.0;
  if (n2 > 0) {
    const double lk = log(kf * sqrt((double)n2));
    a = s1.norm * (herm(t1, t1 + s1.ny, s1, lk) / herm(t1 + 2 * s1.ny, t1 + 3 * s1.ny, s1, lk));
    b = s2.norm * (herm(t2, t2 + s2.ny, s2, lk) / herm(t2 + 2 * s2.ny, t2 + 3 * s2.ny, s2, lk));
  }
  g1[n2] = a;
  g2[n2] = b;
  g1f[n2] = (float)a;
  g2f[n2] = (float)b;
}

// ---------------------------------------------------------------- massive neutrinos
// The factor of PtoMesh's neutrino term per |n|^2 from a ln k slice:
// fac = norm num/den at ln(kf |n|) (OmegaNu/Omega N^3 T_nu(k, a)/T_cb(k, 1)).
__global__ __launch_bounds__(kBlock) void k_nu_fac(size_t n2len, double kf, KSl s, const double *__restrict__ t,
                                                   double *__restrict__ fac) {
  const size_t n2 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n2 >= n2len) return;
  double v = 0.0;
  if (n2 > 0) {
    const double lk = log(kf * sqrt((double)n2));
    v = s.norm * (herm(t, t + s.ny, s, lk) / herm(t + 2 * s.ny, t + 3 * s.ny, s, lk));
  }
  fac[n2] = v;
}

// PtoMesh's neutrino term (src/auxPM.c:383-420) on every stored mode but k = 0:
// delta = cdmfac delta + fac(|n|^2) cdelta_IC, in FP64 as the reference.  Grid
// (N*N rows, kz blocks).
__global__ __launch_bounds__(kBlock) void k_nu_add(float2 *__restrict__ dk, const float2 *__restrict__ cd, int N,
                                                   int HP, const double *__restrict__ fac, double cdmfac) {
  const int k = blockIdx.y * blockDim.x + threadIdx.x;
  if (k > N / 2) return;
  const unsigned row = blockIdx.x;
  const int j = (int)(row % (unsigned)N), i = (int)(row / (unsigned)N);
  if (i == 0 && j == 0 && k == 0) return;
  const int d0 = i > N / 2 ? N - i : i, d1 = j > N / 2 ? N - j : j;
  const double f = fac[d0 * d0 + d1 * d1 + k * k];
  const size_t e = (size_t)row * HP + k;
  const float2 v = dk[e], c = cd[e];
  dk[e] = make_float2((float)(cdmfac * (double)v.x + f * (double)c.x), (float)(cdmfac * (double)v.y + f * (double)c.y));
}

// seed table on the host (a single serial stream), src/2LPT.c:259-271
void seedtable_host(int N, unsigned long seed, std::vector<unsigned> &t) {
  Ranlxd1 r;
  r.seed(seed);
  t.assign((size_t)N * N, 0u);
  auto draw = [&]() { return (unsigned)(0x7fffffff * r.uniform()); };
  for (int i = 0; i < N / 2; i++) {
    for (int j = 0; j < i; j++) t[i * N + j] = draw();
    for (int j = 0; j < i + 1; j++) t[j * N + i] = draw();
    for (int j = 0; j < i; j++) t[(N - 1 - i) * N + j] = draw();
    for (int j = 0; j < i + 1; j++) t[(N - 1 - j) * N + i] = draw();
    for (int j = 0; j < i; j++) t[i * N + (N - 1 - j)] = draw();
    for (int j = 0; j < i + 1; j++) t[j * N + (N - 1 - i)] = draw();
    for (int j = 0; j < i; j++) t[(N - 1 - i) * N + (N - 1 - j)] = draw();
    for (int j = 0; j < i + 1; j++) t[(N - 1 - j) * N + (N - 1 - i)] = draw();
  }
}

int spectrum(mgpu_ctx *ctx, int mode, const float2 *cd, const float2 *cd2, float2 *out, int a, int b,
             const double *g1 = nullptr, const double *g2 = nullptr) {
  const double kf = 2 * MGPU_PI / ctx->p.box;
  const double modes = (double)ctx->N * ctx->N * (ctx->N / 2 + 1);
  {
    KScope ks(ctx, MGPU_K_LPT, 8.0 * modes * (cd2 ? 3 : 2));
    hipLaunchKernelGGL(k_lpt_spectrum, dim3(cdiv(ctx->slot_elems, kBlock), 1, ctx->bz), dim3(kBlock), 0, ctx->stream,
                       mode, cd, cd2, out, ctx->N, ctx->HP, kf, a, b, g1, g2, ctx->zs);
  }
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}

// lattice read-out of a real mesh into dst ([np] floats) with mean removal
int readout(mgpu_ctx *ctx, const float2 *mesh, double scale, float *dst) {
  const size_t np = ctx->np;
  const unsigned nb = cdiv(np, kBlock);
  {
    // compulsory: one mesh sample read and one float written per lattice point
    KScope ks(ctx, MGPU_K_READOUT, 8.0 * np);
    hipLaunchKernelGGL(k_lattice_read, dim3(nb, 1, ctx->bz), dim3(kBlock), 0, ctx->stream, (const float *)mesh, ctx->N,
                       ctx->Np, 2 * ctx->HP, scale, dst, ctx->partial, ctx->zs);
  }
  MGPU_HIP(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_reduce1, dim3(1, 1, ctx->bz), dim3(1024), 0, ctx->stream, ctx->partial, (int)nb, ctx->dsum,
                     ctx->zs);
  MGPU_HIP(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_sub_mean, dim3(nb, 1, ctx->bz), dim3(kBlock), 0, ctx->stream, dst, np, ctx->dsum,
                     1.0 / (double)np, ctx->zs);
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}

int upload_n2(mgpu_ctx *ctx, int which, const double *h) {
  MGPU_HIP(ctx, hipMemcpyAsync(ctx->n2tab[which], h, ctx->n2len * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_d2f, dim3(cdiv(ctx->n2len, kBlock)), dim3(kBlock), 0, ctx->stream, ctx->n2len,
                     (const double *)ctx->n2tab[which], ctx->n2tabf[which]);
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}

}  // namespace

// The device address of this GPU's seed-table scratch, looked up once per
// device (with that device current) and kept for the process.
int ic_scratch_init(mgpu_ctx *ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return mgpu_fail(ctx, MGPU_E_INVALID, "device index beyond the seed-table scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[ctx->device]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return mgpu_fail(ctx, MGPU_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_scratch_ptr[ctx->device] = (unsigned *)p;
  }
  ctx->ic_scratch = g_scratch_ptr[ctx->device];
  return MGPU_OK;
}

// displacement_fields.  seed_dev == nullptr: the seed table goes to a buffer
// allocated for the call; otherwise to seed_dev (the per-device scratch), and
// nothing is allocated on the device.
int ic_displacement_fields(mgpu_ctx *ctx, unsigned long seed, int sphere_mode, int amp_fixed, int inverted,
                           const double *pk_n2, unsigned *seed_dev) {
  PhaseTimer pt(ctx, &ctx->t.ic);
  const int N = ctx->N;
  int rc;
  std::vector<unsigned> st;
  seedtable_host(N, seed, st);
  unsigned *dseed = seed_dev;
  if (!dseed) MGPU_HIP(ctx, hipMalloc(&dseed, st.size() * sizeof(unsigned)));
  MGPU_HIP(ctx, hipMemcpyAsync(dseed, st.data(), st.size() * sizeof(unsigned), hipMemcpyHostToDevice, ctx->stream));
  if ((rc = upload_n2(ctx, 0, pk_n2))) return rc;
  float2 *cd = ctx->p.scale_dependent ? ctx->cdelta[0] : ctx->slot[0];
  MGPU_HIP(ctx, hipMemsetAsync(cd, 0, ctx->slot_elems * sizeof(float2), ctx->stream));
  hipLaunchKernelGGL(k_ic_delta, dim3(cdiv((size_t)N * N, kBlock)), dim3(kBlock), 0, ctx->stream, dseed, N, ctx->Np,
                     ctx->HP, ctx->p.box, sphere_mode, amp_fixed, inverted, ctx->n2tab[0], cd, (size_t)0, (size_t)0,
                     (const int *)nullptr);
  MGPU_HIP(ctx, hipGetLastError());
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (!seed_dev) MGPU_HIP(ctx, hipFree(dseed));
  return ic_lpt_from_delta(ctx, cd);
}

// From delta(k) in cd on: the 2LPT source and the ZA and 2LPT lattice
// displacements (or, scale-dependent, the stored -source).  Every launch covers
// ctx->bz realizations (a batch's member 0 with bz = count).
int ic_lpt_from_delta(mgpu_ctx *ctx, float2 *cd) {
  int rc;
  const int N = ctx->N;
  // displacement gradient -> 2LPT source (real), built in slot 1 over the
  // psi_00 gradient (the diagonal gradients in slots 1-3)
  const size_t nr = 2 * ctx->slot_elems;
  const unsigned gr = cdiv(nr, kBlock);
  static const int ga[3] = {0, 1, 2};
  for (int g = 0; g < 3; g++) {
    if ((rc = spectrum(ctx, SPEC_GRAD, cd, nullptr, ctx->slot[1 + g], ga[g], ga[g]))) return rc;
    if ((rc = fft_c2r_3d(ctx, &ctx->fft, ctx->slot[1 + g]))) return rc;
  }
  hipLaunchKernelGGL(k_source_diag, dim3(gr, 1, ctx->bz), dim3(kBlock), 0, ctx->stream, (const float *)ctx->slot[1],
                     (const float *)ctx->slot[2], (const float *)ctx->slot[3], (float *)ctx->slot[1], nr, ctx->zs);
  MGPU_HIP(ctx, hipGetLastError());
  static const int oa[3] = {0, 0, 1}, ob[3] = {1, 2, 2};
  for (int g = 0; g < 3; g++) {
    if ((rc = spectrum(ctx, SPEC_GRAD, cd, nullptr, ctx->slot[3], oa[g], ob[g]))) return rc;
    if ((rc = fft_c2r_3d(ctx, &ctx->fft, ctx->slot[3]))) return rc;
    hipLaunchKernelGGL(k_source_off, dim3(gr, 1, ctx->bz), dim3(kBlock), 0, ctx->stream, (const float *)ctx->slot[3],
                       (float *)ctx->slot[1], nr, ctx->zs);
    MGPU_HIP(ctx, hipGetLastError());
  }
  if ((rc = fft_r2c_3d(ctx, &ctx->fft, ctx->slot[1]))) return rc;

  if (ctx->p.scale_dependent) {
    if ((rc = spectrum(ctx, SPEC_NEG, ctx->slot[1], nullptr, ctx->cdelta[1], 0, 0))) return rc;
    return MGPU_OK;
  }
  // ZA and 2LPT displacements at the lattice, mean removed
  const double inv_n3 = 1.0 / ((double)N * N * N);
  for (int a = 0; a < 3; a++) {
    if ((rc = spectrum(ctx, SPEC_DISP, cd, nullptr, ctx->slot[3], a, 0))) return rc;
    if ((rc = fft_c2r_3d(ctx, &ctx->fft, ctx->slot[3]))) return rc;
    if ((rc = readout(ctx, ctx->slot[3], 1.0, ctx->lpt[0] + (size_t)a * ctx->np))) return rc;
    if ((rc = spectrum(ctx, SPEC_DISP2, ctx->slot[1], nullptr, ctx->slot[2], a, 0))) return rc;
    if ((rc = fft_c2r_3d(ctx, &ctx->fft, ctx->slot[2]))) return rc;
    if ((rc = readout(ctx, ctx->slot[2], -3. / 7. * inv_n3, ctx->lpt[1] + (size_t)a * ctx->np))) return rc;
  }
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MGPU_OK;
}

int ic_displacement_fields_batch(mgpu_ctx *const *m, int count, const unsigned long *seeds, int sphere_mode,
                                 int amp_fixed, const int *inverted, const double *pk_n2) {
  mgpu_ctx *ctx = m[0];
  PhaseTimer pt(ctx, &ctx->t.ic);
  if (ctx->p.scale_dependent)
    return mgpu_fail(ctx, MGPU_E_STATE, "batched initial conditions: not for scale-dependent runs");
  const int N = ctx->N;
  const size_t nn = (size_t)N * N;
  // all seed tables, then the inversion flags, in one upload
  std::vector<unsigned> all(count * nn + count), st;
  for (int r = 0; r < count; r++) {
    seedtable_host(N, seeds[r], st);
    std::copy(st.begin(), st.end(), all.begin() + r * nn);
    all[count * nn + r] = inverted[r] ? 1u : 0u;
  }
  unsigned *dbuf = nullptr;
  MGPU_HIP(ctx, hipMalloc(&dbuf, all.size() * sizeof(unsigned)));
  struct Free {
    unsigned *p;
    ~Free() { (void)hipFree(p); }
  } free_dbuf{dbuf};
  MGPU_HIP(ctx, hipMemcpyAsync(dbuf, all.data(), all.size() * sizeof(unsigned), hipMemcpyHostToDevice, ctx->stream));
  int rc;
  for (int r = 0; r < count; r++) {
    // every member holds the table (later calls on a member alone read its own)
    if ((rc = upload_n2(m[r], 0, pk_n2))) return rc;
    MGPU_HIP(ctx, hipMemsetAsync(m[r]->slot[0], 0, ctx->slot_elems * sizeof(float2), ctx->stream));
  }
  struct Restore {
    mgpu_ctx *c;
    ~Restore() { c->bz = 1; }
  } restore{ctx};
  ctx->bz = (unsigned)count;
  float2 *cd = ctx->slot[0];
  hipLaunchKernelGGL(k_ic_delta, dim3(cdiv(nn, kBlock), 1, count), dim3(kBlock), 0, ctx->stream, dbuf, N, ctx->Np,
                     ctx->HP, ctx->p.box, sphere_mode, amp_fixed, 0, ctx->n2tab[0], cd, ctx->zs, nn,
                     (const int *)(dbuf + count * nn));
  MGPU_HIP(ctx, hipGetLastError());
  if ((rc = ic_lpt_from_delta(ctx, cd))) return rc;
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MGPU_OK;
}

int ic_init_particles(mgpu_ctx *ctx, double Di, double Di2) {
  PhaseTimer pt(ctx, &ctx->t.ic);
  if (ctx->np == 0) return MGPU_OK;
  hipLaunchKernelGGL(k_init_particles, dim3(cdiv(ctx->np, kBlock)), dim3(kBlock), 0, ctx->stream, ctx->pos, ctx->vel,
                     ctx->lpt[0], ctx->lpt[1], ctx->Np, ctx->p.box, ctx->p.scale_dependent, Di, Di2,
                     ctx->p.use_cola, 0.0, 0.0);
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}

// shared-grid slices as the fold kernel takes them (node float4s in ctx->slice_buf)
struct SliceFold {
  float y0, dy, idy, ymax, norm1, norm2;
  int ny;
};

template <int MF, int NS>
int launch_fold_gs(mgpu_ctx *ctx, const SliceFold &f, float lnkf, float ikf, float2 *out, size_t nl) {
  GShared<NS> gs{};
  gs.sl = nullptr;
  gs.y0 = f.y0;
  gs.dy = f.dy;
  gs.idy = f.idy;
  gs.ymax = f.ymax;
  gs.norm1 = f.norm1;
  gs.norm2 = f.norm2;
  gs.ny = f.ny;
  const int Np = ctx->Np;
  const size_t lds = sizeof(float4) * (size_t)f.ny * (1 + NS);
  const unsigned nb = (unsigned)std::min<size_t>((size_t)Np * Np, 2048);
  hipLaunchKernelGGL((k_lpt_fold_gs<MF, NS>), dim3(nb), dim3(kBlock), lds, ctx->stream, ctx->cdelta[0],
                     ctx->cdelta[1], (const float4 *)ctx->slice_buf, gs, lnkf, ctx->N, ctx->HP, Np,
                     ctx->fft_lattice.hp, ctx->N / Np, ikf, out, nl);
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}

// LPT fields of ns slots (slots[s]).  With sl (shared-grid growth slices of
// every slot, already in ctx->slice_buf) the lattice fold evaluates the growth
// in LDS; otherwise (ns = 1) from the n2 tables.
static int lattice_fields(mgpu_ctx *ctx, int ns, const int *slots, const SliceFold *sl) {
  int rc;
  if (ctx->lattice_slot) {
    // exact lattice fold: N = m * Nsample; spectrum (s, a) at lattice_slot + (3 s + a) nl
    const int Np = ctx->Np, HPn = ctx->fft_lattice.hp, m = ctx->N / Np;
    const size_t nl = (size_t)Np * Np * HPn;
    const float ikf = (float)(ctx->p.box / (2 * MGPU_PI)), lnkf = (float)std::log(2 * MGPU_PI / ctx->p.box);
    {
      // compulsory: both stored spectra read once (the 1LPT one inside the IC cube only), 3 ns folded outputs written
      const double modes = (double)ctx->N * ctx->N * (ctx->N / 2 + 1);
      KScope ks(ctx, MGPU_K_FOLD, 8.0 * modes * (1.0 + 1.0 / (double)(m * m * m)) + 3.0 * ns * 8.0 * nl);
      if (sl && ns == 2) {
        rc = m == 1 ? launch_fold_gs<1, 2>(ctx, *sl, lnkf, ikf, ctx->lattice_slot, nl)
             : m == 2 ? launch_fold_gs<2, 2>(ctx, *sl, lnkf, ikf, ctx->lattice_slot, nl)
             : m == 3 ? launch_fold_gs<3, 2>(ctx, *sl, lnkf, ikf, ctx->lattice_slot, nl)
                      : launch_fold_gs<0, 2>(ctx, *sl, lnkf, ikf, ctx->lattice_slot, nl);
      } else if (sl) {
        rc = m == 1 ? launch_fold_gs<1, 1>(ctx, *sl, lnkf, ikf, ctx->lattice_slot, nl)
             : m == 2 ? launch_fold_gs<2, 1>(ctx, *sl, lnkf, ikf, ctx->lattice_slot, nl)
             : m == 3 ? launch_fold_gs<3, 1>(ctx, *sl, lnkf, ikf, ctx->lattice_slot, nl)
                      : launch_fold_gs<0, 1>(ctx, *sl, lnkf, ikf, ctx->lattice_slot, nl);
      } else {
        auto kern = m == 1 ? k_lpt_fold<1> : m == 2 ? k_lpt_fold<2> : m == 3 ? k_lpt_fold<3> : k_lpt_fold<0>;
        hipLaunchKernelGGL(kern, dim3((unsigned)((size_t)Np * Np), cdiv(Np / 2 + 1, kBlock)), dim3(kBlock), 0,
                           ctx->stream, ctx->cdelta[0], ctx->cdelta[1], GTab{ctx->n2tabf[0], ctx->n2tabf[1]}, ctx->N,
                           ctx->HP, Np, HPn, m, ikf, ctx->lattice_slot, nl);
        rc = MGPU_OK;
      }
      if (rc) return rc;
    }
    MGPU_HIP(ctx, hipGetLastError());
    for (int s = 0; s < ns; s++)
      for (int a = 0; a < 3; a++)
        if ((rc = fft_c2r_3d(ctx, &ctx->fft_lattice, ctx->lattice_slot + (3 * s + a) * nl,
                             ctx->lpt[slots[s]] + (size_t)a * ctx->np)))
          return rc;
    return MGPU_OK;
  }
  if (ns != 1) return mgpu_fail(ctx, MGPU_E_INVALID, "several LPT slots at once need the lattice fold");
  for (int a = 0; a < 3; a++) {
    if ((rc = spectrum(ctx, SPEC_FIELD, ctx->cdelta[0], ctx->cdelta[1], ctx->slot[3], a, 0, ctx->n2tab[0],
                       ctx->n2tab[1])))
      return rc;
    if ((rc = fft_c2r_3d(ctx, &ctx->fft, ctx->slot[3]))) return rc;
    if ((rc = readout(ctx, ctx->slot[3], 1.0, ctx->lpt[slots[0]] + (size_t)a * ctx->np))) return rc;
  }
  return MGPU_OK;
}

// pinned host staging of n bytes; the previous upload from it must be complete before it is refilled
static int slice_staging(mgpu_ctx *ctx, size_t bytes, void **host) {
  if (ctx->slice_ev) MGPU_HIP(ctx, hipEventSynchronize(ctx->slice_ev));
  else MGPU_HIP(ctx, hipEventCreateWithFlags(&ctx->slice_ev, hipEventDisableTiming));
  const size_t nd = (bytes + sizeof(double) - 1) / sizeof(double);
  if (ctx->slice_host_len < nd) {
    if (ctx->slice_host) MGPU_HIP(ctx, hipHostFree(ctx->slice_host));
    ctx->slice_host = nullptr;
    MGPU_HIP(ctx, hipHostMalloc((void **)&ctx->slice_host, nd * sizeof(double), hipHostMallocDefault));
    ctx->slice_host_len = nd;
  }
  if (!ctx->slice_buf || ctx->slice_len < nd) {
    if (ctx->slice_buf) {
      MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
      MGPU_HIP(ctx, hipFree(ctx->slice_buf));
    }
    ctx->slice_buf = nullptr;
    MGPU_HIP(ctx, hipMalloc(&ctx->slice_buf, nd * sizeof(double)));
    ctx->slice_len = nd;
  }
  *host = ctx->slice_host;
  return MGPU_OK;
}
static int slice_upload(mgpu_ctx *ctx, size_t bytes) {
  MGPU_HIP(ctx, hipMemcpyAsync(ctx->slice_buf, ctx->slice_host, bytes, hipMemcpyHostToDevice, ctx->stream));
  MGPU_HIP(ctx, hipEventRecord(ctx->slice_ev, ctx->stream));
  return MGPU_OK;
}

// ns slots, slices q[2 s + o] for (slot s, order o + 1)
int ic_lpt_fields_slices_n(mgpu_ctx *ctx, int ns, const int *slots, const mgpu_kslice *const *q) {
  PhaseTimer pt(ctx, &ctx->t.lpt_fields);
  if (!ctx->p.scale_dependent) return mgpu_fail(ctx, MGPU_E_STATE, "LPT fields need a scale-dependent context");
  for (int e = 0; e < 2 * ns; e++)
    if (!q[e] || q[e]->nk < 2 || q[e]->nk > 8192 || q[e]->dlnk <= 0)
      return mgpu_fail(ctx, MGPU_E_INVALID, "bad growth slice");
  const int ny = q[0]->nk;
  // the shared-grid fold needs one ln k grid for every slice, and (two slots) the
  // same denominators and norms in both slots
  bool shared = ctx->lattice_slot != nullptr && 16 * (size_t)ny * (1 + ns) <= 65536;
  for (int e = 1; e < 2 * ns && shared; e++)
    shared = q[e]->nk == ny && q[e]->lnk0 == q[0]->lnk0 && q[e]->dlnk == q[0]->dlnk;
  for (int o = 0; o < 2 && shared && ns == 2; o++)
    shared = q[2 + o]->norm == q[o]->norm && std::equal(q[o]->den, q[o]->den + ny, q[2 + o]->den) &&
             std::equal(q[o]->dden, q[o]->dden + ny, q[2 + o]->dden);
  if (!shared && ns == 2) {
    int rc = ic_lpt_fields_slices_n(ctx, 1, slots, q);
    return rc ? rc : ic_lpt_fields_slices_n(ctx, 1, slots + 1, q + 2);
  }
  int rc;
  if (shared) {
    const size_t bytes = sizeof(float) * 4 * (size_t)ny * (1 + ns);
    void *h;
    if ((rc = slice_staging(ctx, bytes, &h))) return rc;
    float *f = (float *)h;
    for (int j = 0; j < ny; j++) {
      *f++ = (float)q[0]->den[j];
      *f++ = (float)q[0]->dden[j];
      *f++ = (float)q[1]->den[j];
      *f++ = (float)q[1]->dden[j];
      for (int s = 0; s < ns; s++) {
        *f++ = (float)q[2 * s]->num[j];
        *f++ = (float)q[2 * s]->dnum[j];
        *f++ = (float)q[2 * s + 1]->num[j];
        *f++ = (float)q[2 * s + 1]->dnum[j];
      }
    }
    if ((rc = slice_upload(ctx, bytes))) return rc;
    SliceFold sf;
    sf.y0 = (float)q[0]->lnk0;
    sf.dy = (float)q[0]->dlnk;
    sf.idy = (float)(1.0 / q[0]->dlnk);
    sf.ymax = (float)(q[0]->lnk0 + q[0]->dlnk * (ny - 1));
    sf.norm1 = (float)q[0]->norm;
    sf.norm2 = (float)q[1]->norm;
    sf.ny = ny;
    return lattice_fields(ctx, ns, slots, &sf);
  }
  // general slices: the per-|n|^2 growth tables first (FP64 Hermite), then the table path
  const size_t total = 4 * (size_t)(q[0]->nk + q[1]->nk);
  void *h;
  if ((rc = slice_staging(ctx, total * sizeof(double), &h))) return rc;
  double *p = (double *)h;
  for (int e = 0; e < 2; e++)
    for (const double *src : {q[e]->num, q[e]->dnum, q[e]->den, q[e]->dden}) {
      std::copy(src, src + q[e]->nk, p);
      p += q[e]->nk;
    }
  if ((rc = slice_upload(ctx, total * sizeof(double)))) return rc;
  const KSl k0{q[0]->lnk0, q[0]->dlnk, q[0]->norm, q[0]->nk}, k1{q[1]->lnk0, q[1]->dlnk, q[1]->norm, q[1]->nk};
  hipLaunchKernelGGL(k_growth_n2, dim3(cdiv(ctx->n2len, kBlock)), dim3(kBlock), 0, ctx->stream, ctx->n2len,
                     2 * MGPU_PI / ctx->p.box, k0, (const double *)ctx->slice_buf, k1,
                     (const double *)ctx->slice_buf + 4 * q[0]->nk, ctx->n2tab[0], ctx->n2tab[1], ctx->n2tabf[0],
                     ctx->n2tabf[1]);
  MGPU_HIP(ctx, hipGetLastError());
  return lattice_fields(ctx, 1, slots, nullptr);
}

int ic_lpt_fields_slices(mgpu_ctx *ctx, int slot, const mgpu_kslice *o1, const mgpu_kslice *o2) {
  const mgpu_kslice *q[2] = {o1, o2};
  return ic_lpt_fields_slices_n(ctx, 1, &slot, q);
}

int ic_lpt_fields(mgpu_ctx *ctx, int slot, const double *g1_n2, const double *g2_n2) {
  PhaseTimer pt(ctx, &ctx->t.lpt_fields);
  if (!ctx->p.scale_dependent) return mgpu_fail(ctx, MGPU_E_STATE, "LPT fields need a scale-dependent context");
  int rc;
  if ((rc = upload_n2(ctx, 0, g1_n2))) return rc;
  if ((rc = upload_n2(ctx, 1, g2_n2))) return rc;
  if ((rc = lattice_fields(ctx, 1, &slot, nullptr))) return rc;
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MGPU_OK;
}


// the neutrino term's factors for the next PtoMesh (f == NULL: off)
int ic_nu_set(mgpu_ctx *ctx, double cdmfac, const mgpu_kslice *f) {
  if (!f) {
    ctx->nu_on = false;
    return MGPU_OK;
  }
  if (!f->num || !f->dnum || !f->den || !f->dden || f->nk < 2 || f->nk > 8192 || f->dlnk <= 0)
    return mgpu_fail(ctx, MGPU_E_INVALID, "bad neutrino slice");
  if (!ctx->cdelta[0]) return mgpu_fail(ctx, MGPU_E_STATE, "the neutrino term needs the stored IC spectrum (scale-dependent context)");
  const size_t n = 4 * (size_t)f->nk;
  std::vector<double> h(n);
  std::copy(f->num, f->num + f->nk, h.begin());
  std::copy(f->dnum, f->dnum + f->nk, h.begin() + f->nk);
  std::copy(f->den, f->den + f->nk, h.begin() + 2 * f->nk);
  std::copy(f->dden, f->dden + f->nk, h.begin() + 3 * f->nk);
  if (ctx->nu_slice_len < n) {
    MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->nu_slice) MGPU_HIP(ctx, hipFree(ctx->nu_slice));
    ctx->nu_slice = nullptr;
    MGPU_HIP(ctx, hipMalloc((void **)&ctx->nu_slice, n * sizeof(double)));
    ctx->nu_slice_len = n;
  }
  if (!ctx->nu_fac) MGPU_HIP(ctx, hipMalloc((void **)&ctx->nu_fac, ctx->n2len * sizeof(double)));
  MGPU_HIP(ctx, hipMemcpyAsync(ctx->nu_slice, h.data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  const KSl ks{f->lnk0, f->dlnk, f->norm, f->nk};
  hipLaunchKernelGGL(k_nu_fac, dim3(cdiv(ctx->n2len, kBlock)), dim3(kBlock), 0, ctx->stream, ctx->n2len,
                     2 * MGPU_PI / ctx->p.box, ks, (const double *)ctx->nu_slice, ctx->nu_fac);
  MGPU_HIP(ctx, hipGetLastError());
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));  // h goes out of scope
  ctx->nu_cdmfac = cdmfac;
  ctx->nu_on = true;
  return MGPU_OK;
}

// the neutrino term on delta(k) in slot 0
int ic_nu_apply(mgpu_ctx *ctx) {
  if (!ctx->nu_on) return mgpu_fail(ctx, MGPU_E_STATE, "no neutrino factors: call mgpu_set_neutrinos first");
  const int N = ctx->N;
  KScope ks(ctx, MGPU_K_SPECTRAL, 24.0 * (double)N * N * (N / 2 + 1));
  hipLaunchKernelGGL(k_nu_add, dim3((unsigned)((size_t)N * N), cdiv((size_t)(N / 2 + 1), kBlock)), dim3(kBlock), 0,
                     ctx->stream, ctx->slot[0], ctx->cdelta[0], N, ctx->HP, (const double *)ctx->nu_fac,
                     ctx->nu_cdmfac);
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}
