// Spectral passes fused into FFT passes (gfx950).
//
// The reference applies every k-space multiplier and every real-space
// screening factor as a separate sweep over FP64 meshes between FFTW calls
// (src/mg.h:21-309, src/auxPM.c:450-535).  Here each multiplier rides on the
// FFT pass that touches the same data, so no mesh is read or written only to
// be scaled:
//   green x pass (inverse)   S = delta(k) [+ phi(k)] times the Green's function
//                            or smoothing window, then the x transform; the
//                            force variant writes two transforms, of i n_x G
//                            and of G (the y and z components share G);
//   dual y pass (inverse)    y transform of T and of i n_y T from one read;
//   z screen pass            c2r of the potential (or smoothed density) row,
//                            the real-space screening against the delta(x)
//                            row, and the r2c of the effective density row, in
//                            one kernel with the row kept in LDS;
//   epilogue x pass (fwd)    x transform followed by coupling n^2/(n^2 + m^2);
//   z c2r with i n_z         (fft.hip) the z force component.
// Integer wavenumbers follow the Forces convention n in (-N/2, N/2] (the
// Nyquist index counts as +N/2, src/auxPM.c:467-479); only n^2 enters the
// MG and smoothing multipliers, where the sign convention does not matter.
#include "fft_core.h"

#include <cmath>

namespace {

using namespace fftc;

__device__ __forceinline__ int nsgn(int i, int N) { return i > N / 2 ? i - N : i; }

enum { GREEN_FORCE = FUSED_GREEN_FORCE, GREEN_LAPLACE = FUSED_GREEN_LAPLACE, GREEN_SMOOTH = FUSED_GREEN_SMOOTH };
struct GreenArgs {
  float cnorm;  // GREEN_FORCE: 1 / (N^3 2 pi / Box)
  float nf;     // GREEN_LAPLACE: 1.5 Omega / a (Box / (2 pi c/H0))^2 / N^3; GREEN_SMOOTH: 1 / N^3
  float kfac;   // GREEN_SMOOTH: 2 pi R / Box
  int filter;   // GREEN_SMOOTH: MGPU_FILTER_*
  const float *mult = nullptr;  // GREEN_FORCE: learned multiplier per n^2, or none
};

__device__ __forceinline__ float window(int filter, float kR) {
  switch (filter) {
    case MGPU_FILTER_GAUSSIAN: return expf(-0.5f * kR * kR);
    case MGPU_FILTER_TOPHAT: return kR < 1e-5f ? 1.f : 3.f / (kR * kR * kR) * (sinf(kR) - kR * cosf(kR));
    case MGPU_FILTER_SHARPK: return kR < 1.f ? 1.f : 0.f;
    default: return 1.f;
  }
}

// The x pass's Green's / smoothing multiplier of S at (n_x, n^2), valid = the
// column is a stored mode (kz <= N/2): o1 = G (or i n_x G for GREEN_FORCE,
// o2 = G), G = f S.
template <int MODE>
__device__ __forceinline__ void green_mul(float2 S, int nx, int n2, bool valid, const GreenArgs &ga, float2 &o1,
                                          float2 &o2) {
  float f = 0.f;
  if (valid) {
    if (MODE == GREEN_FORCE) {
      f = n2 > 0 ? -ga.cnorm / (float)n2 : 0.f;
      if (ga.mult) f *= ga.mult[n2];
    }
    if (MODE == GREEN_LAPLACE) f = n2 > 0 ? -ga.nf / (float)n2 : 0.f;
    if (MODE == GREEN_SMOOTH) f = window(ga.filter, sqrtf((float)n2) * ga.kfac) * ga.nf;
  }
  const float2 G = cscale(S, f);
  if (MODE == GREEN_FORCE) {
    o1 = cscale(mul_i<+1>(G), (float)nx);
    o2 = G;
  } else {
    o1 = G;
  }
}

// x pass, inverse.  Tile (bx, by) of xcd_tile: columns kz = bx*T + c, outer index ky = by,
// line index kx.  out1 is written by GREEN_FORCE only.
template <int L, int T, int MODE>
__global__ __launch_bounds__(col_threads(L, T)) void k_col_green(const float2 *__restrict__ a,
                                                                   const float2 *__restrict__ b, float2 *out0,
                                                                   float2 *out1, size_t ls, size_t os,
                                                                   const float2 *__restrict__ tw, GreenArgs ga) {
  extern __shared__ float2 lds[];
  constexpr int NT = col_threads(L, T), E = col_ept(L, T);
  constexpr int R0 = pick_radix(L, E), NB0 = E / R0, M0 = L / R0;
  constexpr int N = L;
  const int tid = threadIdx.x;
  int bx, ky;
  xcd_tile(bx, ky);
  const int kz0 = bx * T;
  const size_t tile = (size_t)ky * os + kz0;
  float2 p1[E], p2[E];
#pragma unroll
  for (int q = 0; q < NB0; ++q) {
    const int g = tid + q * NT, c = g % T, j = g / T;
    const float2 *pa = a + tile + (size_t)j * ls + c;
    const float2 *pb = b ? b + tile + (size_t)j * ls + c : nullptr;
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      float2 v = pa[(size_t)(r * M0) * ls];
      if (pb) v = cadd(v, pb[(size_t)(r * M0) * ls]);
      p1[q * R0 + r] = v;
    }
  }
  const int ny = nsgn(ky, N);
#pragma unroll
  for (int q = 0; q < NB0; ++q) {
    const int g = tid + q * NT, c = g % T, j = g / T;
    const int kz = kz0 + c;
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int nx = nsgn(j + r * M0, N);
      const int n2 = nx * nx + ny * ny + kz * kz;
      green_mul<MODE>(p1[q * R0 + r], nx, n2, kz <= N / 2, ga, p1[q * R0 + r], p2[q * R0 + r]);
    }
  }
  auto ld = [&](int, int, int) { return make_float2(0.f, 0.f); };
  auto st0 = [&](int c, int j, int o, float2 v) { out0[tile + (size_t)j * ls + c + (size_t)o * ls] = v; };
  fft_stages<L, T, NT, +1, 1, T, true, false, 1, false, true>(lds, tw, 1, tid, ld, st0, p1);
  if (MODE == GREEN_FORCE) {
    __syncthreads();
    auto st1 = [&](int c, int j, int o, float2 v) { out1[tile + (size_t)j * ls + c + (size_t)o * ls] = v; };
    fft_stages<L, T, NT, +1, 1, T, true, false, 1, false, true>(lds, tw, 1, tid, ld, st1, p2);
  }
}

// y pass, inverse, two outputs from one read: buf <- Y[buf], oy <- Y[i n_y buf].
// Tile (bx, by) of xcd_tile: columns kz, outer index kx = by, line index ky.
template <int L, int T>
__global__ __launch_bounds__(col_threads(L, T)) void k_col_dual(float2 *buf, float2 *__restrict__ oy, size_t ls,
                                                                  size_t os, const float2 *__restrict__ tw, size_t zs) {
  buf = zbuf(buf, zs);
  oy = zbuf(oy, zs);
  extern __shared__ float2 lds[];
  constexpr int NT = col_threads(L, T), E = col_ept(L, T);
  constexpr int R0 = pick_radix(L, E), NB0 = E / R0, M0 = L / R0;
  const int tid = threadIdx.x;
  int bx, by;
  xcd_tile(bx, by);
  const size_t tile = (size_t)by * os + (size_t)bx * T;
  float2 p1[E], p2[E];
#pragma unroll
  for (int q = 0; q < NB0; ++q) {
    const int g = tid + q * NT, c = g % T, j = g / T;
    const float2 *pa = buf + tile + (size_t)j * ls + c;
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const float2 v = pa[(size_t)(r * M0) * ls];
      p2[q * R0 + r] = v;
      p1[q * R0 + r] = cscale(mul_i<+1>(v), (float)nsgn(j + r * M0, L));
    }
  }
  auto ld = [&](int, int, int) { return make_float2(0.f, 0.f); };
  auto st0 = [&](int c, int j, int o, float2 v) { oy[tile + (size_t)j * ls + c + (size_t)o * ls] = v; };
  fft_stages<L, T, NT, +1, 1, T, true, false, 1, false, true>(lds, tw, 1, tid, ld, st0, p1);
  __syncthreads();
  auto st1 = [&](int c, int j, int o, float2 v) { buf[tile + (size_t)j * ls + c + (size_t)o * ls] = v; };
  fft_stages<L, T, NT, +1, 1, T, true, false, 1, false, true>(lds, tw, 1, tid, ld, st1, p2);
}

// x pass, forward, then phi = coupling n^2/(n^2 + m2) X (0 at n = 0): EffDensitykToPhiofk
// (src/mg.h:74-116) on the transform's output.  Tile: columns kz, outer ky, line kx.
template <int L, int T>
__global__ __launch_bounds__(col_threads(L, T)) void k_col_epi(float2 *buf, size_t ls, size_t os,
                                                                 const float2 *__restrict__ tw, float coupling,
                                                                 float m2) {
  extern __shared__ float2 lds[];
  constexpr int NT = col_threads(L, T);
  constexpr int N = L;
  const int tid = threadIdx.x;
  int bx, ky;
  xcd_tile(bx, ky);
  const int kz0 = bx * T;
  const size_t tile = (size_t)ky * os + kz0;
  const int ny = nsgn(ky, N);
  auto ld = [&](int c, int j, int o) { return buf[tile + (size_t)j * ls + c + (size_t)o * ls]; };
  auto st = [&](int c, int j, int o, float2 v) {
    const int kz = kz0 + c, nx = nsgn(j + o, N);
    const int n2 = nx * nx + ny * ny + kz * kz;
    const float f = (kz <= N / 2 && n2 > 0) ? coupling * ((float)n2 / ((float)n2 + m2)) : 0.f;
    buf[tile + (size_t)j * ls + c + (size_t)o * ls] = cscale(v, f);
  };
  fft_stages<L, T, NT, -1, 1, T, true, false, 1>(lds, tw, 1, tid, ld, st);
}

// x pass, forward, accumulating compute_power_spectrum's binned power
// (src/compute_pofk.c:71-236) from the transform's output as it is stored:
// weight 1 (kz = 0, N/2) or 2, CIC deconvolution winv[d0] winv[d1] winv[d2]
// (1/N^6 folded in), |n| binned linearly (bintype 0) or logarithmically.
// Bin sums go to LDS doubles, then one global atomic per bin and block.
template <int L, int T>
__global__ __launch_bounds__(col_threads(L, T)) void k_col_pofk(float2 *buf, size_t ls, size_t os,
                                                                  const float2 *__restrict__ tw,
                                                                  const float *__restrict__ winv,
                                                                  const int16_t *__restrict__ pbin, int nbins,
                                                                  double *__restrict__ pk) {
  extern __shared__ float2 lds[];
  constexpr int NT = col_threads(L, T);
  constexpr int N = L;
  double *bins = (double *)(lds + (size_t)L * T);
  float *wv = (float *)(bins + nbins);
  const int tid = threadIdx.x;
  for (int b = tid; b < nbins; b += NT) bins[b] = 0.0;
  for (int d = tid; d <= N / 2; d += NT) wv[d] = winv[d];
  __syncthreads();
  int bx, ky;
  xcd_tile(bx, ky);
  const int kz0 = bx * T;
  const size_t tile = (size_t)ky * os + kz0;
  const int d1 = ky > N / 2 ? N - ky : ky;
  auto ld = [&](int c, int j, int o) { return buf[tile + (size_t)j * ls + c + (size_t)o * ls]; };
  auto st = [&](int c, int j, int o, float2 v) {
    buf[tile + (size_t)j * ls + c + (size_t)o * ls] = v;
    const int kz = kz0 + c, kx = j + o;
    if (kz > N / 2) return;
    const int d0 = kx > N / 2 ? N - kx : kx;
    // the bin of |n|^2 comes from the table the host fills with the FP64 formula
    const int nk = pbin[d0 * d0 + d1 * d1 + kz * kz];
    if (nk < 0) return;
    const float w = (kz == 0 || kz == N / 2) ? 1.f : 2.f;
    atomicAdd(&bins[nk], (double)((v.x * v.x + v.y * v.y) * (w * wv[d0] * wv[d1] * wv[kz])));
  };
  fft_stages<L, T, NT, -1, 1, T, true, false, 1>(lds, tw, 1, tid, ld, st);
  __syncthreads();
  for (int b = tid; b < nbins; b += NT)
    if (bins[b] != 0.0) atomicAdd(&pk[b], bins[b]);
}

// x pass pair: a forward x transform and, from its output in LDS, an inverse
// x transform with a Green's / smoothing multiplier, so the spectrum between
// them is never written and read back.
//   FWD_DENSITY  the density's forward x pass: delta(k) stored to kout (when
//                not null) and its power binned into pk (when not null) as
//                k_col_pofk does; S = delta(k).
//   FWD_EPI      forward x pass of the screened effective density, the
//                EffDensitykToPhiofk epilogue (as k_col_epi), S = delta(k) + phi(k)
//                with delta(k) read from dk (as k_col_green adds them).
//   FWD_PLAIN    the same without the epilogue (DGP: phi(k) is the transform).
// Then S times the MODE multiplier and the inverse x transform(s), as
// k_col_green.  The forward's last stage stores its outputs into LDS in the
// tile layout (after a barrier), from where each thread takes its first-stage
// operands of the inverse.
enum { FWD_DENSITY = 0, FWD_EPI = 1, FWD_PLAIN = 2 };
template <int L, int T, int FWD, int MODE>
__global__ __launch_bounds__(col_threads(L, T)) void k_col_fused(
    const float2 *in, float2 *kout, const float2 *__restrict__ dk, float2 *out0, float2 *out1, size_t ls, size_t os,
    const float2 *__restrict__ tw, GreenArgs ga, float coupling, float m2, const float *__restrict__ winv,
    const int16_t *__restrict__ pbin, int nbins, double *__restrict__ pk, size_t zs) {
  in = zbuf(in, zs);
  kout = zbuf(kout, zs);
  dk = zbuf(dk, zs);
  out0 = zbuf(out0, zs);
  out1 = zbuf(out1, zs);
  pk = zbuf(pk, zs);
  extern __shared__ float2 lds[];
  constexpr int NT = col_threads(L, T), E = col_ept(L, T);
  constexpr int R0 = pick_radix(L, E), NB0 = E / R0, M0 = L / R0;
  constexpr int N = L;
  const int tid = threadIdx.x;
  int bx, ky;
  xcd_tile(bx, ky);
  const int kz0 = bx * T;
  const size_t tile = (size_t)ky * os + kz0;
  const int ny = nsgn(ky, N), d1 = ky > N / 2 ? N - ky : ky;
  const bool bin = FWD == FWD_DENSITY && pk != nullptr;
  double *bins = (double *)(lds + (size_t)L * T);
  float *wv = (float *)(bins + nbins);
  if (bin) {
    for (int b = tid; b < nbins; b += NT) bins[b] = 0.0;
    for (int d = tid; d <= N / 2; d += NT) wv[d] = winv[d];
    __syncthreads();
  }
  // forward transform; its outputs (row kx of column c) go to lds[c + kx T]
  auto ld = [&](int c, int j, int o) { return in[tile + (size_t)j * ls + c + (size_t)o * ls]; };
  auto st = [&](int c, int j, int o, float2 v) {
    const int kx = j + o, kz = kz0 + c;
    if (FWD == FWD_DENSITY) {
      if (kout) kout[tile + (size_t)kx * ls + c] = v;
      if (bin && kz <= N / 2) {
        const int d0 = kx > N / 2 ? N - kx : kx;
        const int nk = pbin[d0 * d0 + d1 * d1 + kz * kz];
        if (nk >= 0) {
          const float w = (kz == 0 || kz == N / 2) ? 1.f : 2.f;
          atomicAdd(&bins[nk], (double)((v.x * v.x + v.y * v.y) * (w * wv[d0] * wv[d1] * wv[kz])));
        }
      }
    }
    if (FWD == FWD_EPI) {
      const int nx = nsgn(kx, N), n2 = nx * nx + ny * ny + kz * kz;
      const float f = (kz <= N / 2 && n2 > 0) ? coupling * ((float)n2 / ((float)n2 + m2)) : 0.f;
      v = cscale(v, f);
    }
    lds[c + kx * T] = v;
  };
  fft_stages<L, T, NT, -1, 1, T, true, false, 1, true>(lds, tw, 1, tid, ld, st);
  // delta(k) to add, in the inverse's first-stage layout (requested after the
  // forward's loads: the vector-memory counter is waited on in order, so loads
  // issued earlier would delay the forward's first stage)
  float2 p1[E], p2[E];
  if (FWD != FWD_DENSITY) {
#pragma unroll
    for (int q = 0; q < NB0; ++q) {
      const int g = tid + q * NT, c = g % T, j = g / T;
#pragma unroll
      for (int r = 0; r < R0; ++r) p2[q * R0 + r] = dk[tile + (size_t)(j + r * M0) * ls + c];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NB0; ++q) {
    const int g = tid + q * NT, c = g % T, j = g / T;
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const float2 v = lds[c + (j + r * M0) * T];
      p1[q * R0 + r] = FWD == FWD_DENSITY ? v : cadd(p2[q * R0 + r], v);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NB0; ++q) {
    const int g = tid + q * NT, c = g % T, j = g / T;
    const int kz = kz0 + c;
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int nx = nsgn(j + r * M0, N);
      const int n2 = nx * nx + ny * ny + kz * kz;
      green_mul<MODE>(p1[q * R0 + r], nx, n2, kz <= N / 2, ga, p1[q * R0 + r], p2[q * R0 + r]);
    }
  }
  auto ldz = [&](int, int, int) { return make_float2(0.f, 0.f); };
  auto st0 = [&](int c, int j, int o, float2 v) { out0[tile + (size_t)j * ls + c + (size_t)o * ls] = v; };
  fft_stages<L, T, NT, +1, 1, T, true, false, 1, false, true>(lds, tw, 1, tid, ldz, st0, p1);
  if (MODE == GREEN_FORCE) {
    __syncthreads();
    auto st1 = [&](int c, int j, int o, float2 v) { out1[tile + (size_t)j * ls + c + (size_t)o * ls] = v; };
    fft_stages<L, T, NT, +1, 1, T, true, false, 1, false, true>(lds, tw, 1, tid, ldz, st1, p2);
  }
  if (bin) {
    __syncthreads();
    for (int b = tid; b < nbins; b += NT)
      if (bins[b] != 0.0) atomicAdd(&pk[b], bins[b]);
  }
}

enum { SCREEN_POTENTIAL = 0, SCREEN_DENSITY = 1 };

// z pass: c2r of the row (potential Phi_N or smoothed density delta_R), the
// screening against the delta(x) row, r2c of the effective density; the row
// stays in LDS between the three steps.
//   SCREEN_POTENTIAL (f(R), m-beta; src/mg.h:169-171, src/user_defined_functions.h:725-752):
//       delta_eff = delta * min(1, |p0 / Phi|) where Phi < 0, else delta       (p0 = phi_crit)
//   SCREEN_DENSITY (DGP; src/mg.h:235-244, src/user_defined_functions.h:757-773):
//       delta_eff = delta * p0 * 2 (sqrt(1 + f) - 1) / f, f = p1 (1 + delta_R), 1 if f < 1e-5
//                                                                   (p0 = coupling, p1 = dgp_fac)
template <int N, int KIND>
__global__ __launch_bounds__(RowCfg<N>::NT) void k_z_screen(float2 *buf, const float2 *__restrict__ dx, int HP,
                                                             size_t nrows, const float2 *__restrict__ tw, float p0,
                                                             float p1, size_t zs) {
  buf = zbuf(buf, zs);
  dx = zbuf(dx, zs);
  using C = RowCfg<N>;
  constexpr int ITER = C::RB * C::M / C::NT;
  static_assert(C::RB * C::M % C::NT == 0, "row block not a multiple of the block size");
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x;
  const size_t row0 = (size_t)blockIdx.x * C::RB;
  // the delta(x) pairs this thread screens, loaded first so they arrive during the c2r
  float2 dv[ITER];
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    const int e = tid + it * C::NT, r = e / C::M, n = e % C::M;
    const size_t row = row0 + r < nrows ? row0 + r : nrows - 1;
    dv[it] = dx[row * HP + n];
  }
  z_merge<N, false>(lds, buf, HP, row0, nrows, tw, tid);
  __syncthreads();
  auto ldl = [&](int r, int j, int o) { return lds[r * C::LS + j + o]; };
  auto stl = [&](int r, int j, int o, float2 v) { lds[r * C::LS + j + o] = v; };
  fft_stages<C::M, C::RB, C::NT, +1, C::LS, 1, false, true, 1>(lds, tw, 2, tid, ldl, stl);
  __syncthreads();
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    const int e = tid + it * C::NT, r = e / C::M, n = e % C::M;
    const float2 f = lds[r * C::LS + n];
    const float2 d = dv[it];
    float s0, s1;
    if (KIND == SCREEN_POTENTIAL) {
      s0 = f.x < 0.f ? fminf(1.f, fabsf(p0 / f.x)) : 1.f;
      s1 = f.y < 0.f ? fminf(1.f, fabsf(p0 / f.y)) : 1.f;
    } else {
      const float a0 = p1 * (1.f + f.x), a1 = p1 * (1.f + f.y);
      s0 = p0 * (a0 < 1e-5f ? 1.f : 2.f * (sqrtf(1.f + a0) - 1.f) / a0);
      s1 = p0 * (a1 < 1e-5f ? 1.f : 2.f * (sqrtf(1.f + a1) - 1.f) / a1);
    }
    lds[r * C::LS + n] = make_float2(d.x * s0, d.y * s1);
  }
  __syncthreads();
  float2 w[ZSplit<N>::ITER];
  z_split_twiddles<N>(w, tw, tid);
  fft_stages<C::M, C::RB, C::NT, -1, C::LS, 1, false, true, 1>(lds, tw, 2, tid, ldl, stl);
  __syncthreads();
  z_split<N>(lds, w, buf, HP, row0, nrows, tid);
}

template <class K>
int set_lds(mgpu_ctx *ctx, K kern, size_t lds) {
  if (lds > 65536) MGPU_HIP(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  return MGPU_OK;
}

template <int N, int T>
int green_launch(mgpu_ctx *ctx, int mode, const float2 *a, const float2 *b, float2 *o0, float2 *o1, GreenArgs ga) {
  const size_t HP = ctx->fft.hp, ls = (size_t)N * HP, os = HP;
  const size_t lds = (size_t)N * T * sizeof(float2);
  dim3 grid(HP / T, N), block(col_threads(N, T));
  int rc;
  if (mode == GREEN_FORCE) {
    if ((rc = set_lds(ctx, k_col_green<N, T, GREEN_FORCE>, lds))) return rc;
    hipLaunchKernelGGL((k_col_green<N, T, GREEN_FORCE>), grid, block, lds, ctx->stream, a, b, o0, o1, ls, os,
                       ctx->fft.tw, ga);
  } else if (mode == GREEN_LAPLACE) {
    if ((rc = set_lds(ctx, k_col_green<N, T, GREEN_LAPLACE>, lds))) return rc;
    hipLaunchKernelGGL((k_col_green<N, T, GREEN_LAPLACE>), grid, block, lds, ctx->stream, a, b, o0, o1, ls, os,
                       ctx->fft.tw, ga);
  } else {
    if ((rc = set_lds(ctx, k_col_green<N, T, GREEN_SMOOTH>, lds))) return rc;
    hipLaunchKernelGGL((k_col_green<N, T, GREEN_SMOOTH>), grid, block, lds, ctx->stream, a, b, o0, o1, ls, os,
                       ctx->fft.tw, ga);
  }
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}

struct FusedXArgs {
  int fwd, mode;
  const float2 *in;
  float2 *kout;
  const float2 *dk;
  float2 *o0, *o1;
  GreenArgs ga;
  float coupling, m2;
  double *pk;
};

template <int N, int T, int FWD, int MODE>
int fused_x_launch_k(mgpu_ctx *ctx, const FusedXArgs &a) {
  const size_t HP = ctx->fft.hp, ls = (size_t)N * HP, os = HP;
  const int nbins = a.pk ? ctx->p.pofk_nbins : 0;
  const size_t lds = (size_t)N * T * sizeof(float2) + (a.pk ? sizeof(double) * nbins + sizeof(float) * (N / 2 + 1) : 0);
  if (lds > ctx->lds_max) return mgpu_fail(ctx, MGPU_E_INVALID, "P(k) bins do not fit the LDS of the x pass");
  if (a.pk && !ctx->pk_bin) return mgpu_fail(ctx, MGPU_E_INVALID, "no P(k) bin table");
  int rc;
  auto kern = k_col_fused<N, T, FWD, MODE>;
  if ((rc = set_lds(ctx, kern, lds))) return rc;
  hipLaunchKernelGGL(kern, dim3(HP / T, N, ctx->bz), dim3(col_threads(N, T)), lds, ctx->stream, a.in, a.kout, a.dk,
                     a.o0, a.o1, ls, os, ctx->fft.tw, a.ga, a.coupling, a.m2, ctx->winvf, ctx->pk_bin, nbins, a.pk,
                     ctx->zs);
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}

template <int N, int T>
int fused_x_launch(mgpu_ctx *ctx, const FusedXArgs &a) {
  if (a.fwd == FWD_DENSITY && a.mode == GREEN_LAPLACE) return fused_x_launch_k<N, T, FWD_DENSITY, GREEN_LAPLACE>(ctx, a);
  if (a.fwd == FWD_DENSITY && a.mode == GREEN_SMOOTH) return fused_x_launch_k<N, T, FWD_DENSITY, GREEN_SMOOTH>(ctx, a);
  // the force pair keeps G and i n_x G in registers: only with 8-column tiles
  if constexpr (T == 8) {
    if (a.fwd == FWD_DENSITY && a.mode == GREEN_FORCE) return fused_x_launch_k<N, T, FWD_DENSITY, GREEN_FORCE>(ctx, a);
    if (a.fwd == FWD_EPI && a.mode == GREEN_FORCE) return fused_x_launch_k<N, T, FWD_EPI, GREEN_FORCE>(ctx, a);
    if (a.fwd == FWD_PLAIN && a.mode == GREEN_FORCE) return fused_x_launch_k<N, T, FWD_PLAIN, GREEN_FORCE>(ctx, a);
  }
  return mgpu_fail(ctx, MGPU_E_INVALID, "unsupported fused x pass");
}

template <int N, int T>
int dual_launch(mgpu_ctx *ctx, float2 *buf, float2 *oy) {
  const size_t HP = ctx->fft.hp, ls = HP, os = (size_t)N * HP;
  const size_t lds = (size_t)N * T * sizeof(float2);
  int rc;
  if ((rc = set_lds(ctx, k_col_dual<N, T>, lds))) return rc;
  hipLaunchKernelGGL((k_col_dual<N, T>), dim3(HP / T, N, ctx->bz), dim3(col_threads(N, T)), lds, ctx->stream, buf, oy,
                     ls, os, ctx->fft.tw, ctx->zs);
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}

template <int N, int T>
int epi_launch(mgpu_ctx *ctx, float2 *buf, float coupling, float m2) {
  const size_t HP = ctx->fft.hp, ls = (size_t)N * HP, os = HP;
  const size_t lds = (size_t)N * T * sizeof(float2);
  int rc;
  if ((rc = set_lds(ctx, k_col_epi<N, T>, lds))) return rc;
  hipLaunchKernelGGL((k_col_epi<N, T>), dim3(HP / T, N), dim3(col_threads(N, T)), lds, ctx->stream, buf, ls, os,
                     ctx->fft.tw, coupling, m2);
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}

template <int N, int T>
int pofk_launch(mgpu_ctx *ctx, float2 *buf, double *pk) {
  const size_t HP = ctx->fft.hp, ls = (size_t)N * HP, os = HP;
  const int nbins = ctx->p.pofk_nbins;
  const size_t lds = (size_t)N * T * sizeof(float2) + sizeof(double) * nbins + sizeof(float) * (N / 2 + 1);
  if (lds > ctx->lds_max) return mgpu_fail(ctx, MGPU_E_INVALID, "P(k) bins do not fit the LDS of the x pass");
  if (!ctx->pk_bin) return mgpu_fail(ctx, MGPU_E_INVALID, "no P(k) bin table");
  int rc;
  if ((rc = set_lds(ctx, k_col_pofk<N, T>, lds))) return rc;
  hipLaunchKernelGGL((k_col_pofk<N, T>), dim3(HP / T, N), dim3(col_threads(N, T)), lds, ctx->stream, buf, ls, os,
                     ctx->fft.tw, ctx->winvf, ctx->pk_bin, nbins, pk);
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}

template <int N>
int screen_launch(mgpu_ctx *ctx, int kind, float2 *buf, const float2 *dx, float p0, float p1) {
  using C = RowCfg<N>;
  const size_t nrows = (size_t)N * N;
  const size_t lds = (size_t)C::RB * C::LS * sizeof(float2);
  dim3 grid(cdiv(nrows, C::RB), 1, ctx->bz), block(C::NT);
  int rc;
  if (kind == SCREEN_POTENTIAL) {
    if ((rc = set_lds(ctx, k_z_screen<N, SCREEN_POTENTIAL>, lds))) return rc;
    hipLaunchKernelGGL((k_z_screen<N, SCREEN_POTENTIAL>), grid, block, lds, ctx->stream, buf, dx, (int)ctx->fft.hp,
                       nrows, ctx->fft.tw, p0, p1, ctx->zs);
  } else {
    if ((rc = set_lds(ctx, k_z_screen<N, SCREEN_DENSITY>, lds))) return rc;
    hipLaunchKernelGGL((k_z_screen<N, SCREEN_DENSITY>), grid, block, lds, ctx->stream, buf, dx, (int)ctx->fft.hp,
                       nrows, ctx->fft.tw, p0, p1, ctx->zs);
  }
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}

// column kernels are instantiated for the tile width the plan chose
template <int N, template <int, int> class F, class... A>
int by_tile(mgpu_ctx *ctx, A... args) {
  if constexpr (col_threads(N, 16) <= 1024) {
    if (ctx->fft.tcols == 16) return F<N, 16>::run(ctx, args...);
  }
  return F<N, 8>::run(ctx, args...);
}
template <int N, int T>
struct GreenF {
  static int run(mgpu_ctx *c, int m, const float2 *a, const float2 *b, float2 *o0, float2 *o1, GreenArgs g) {
    return green_launch<N, T>(c, m, a, b, o0, o1, g);
  }
};
template <int N, int T>
struct DualF {
  static int run(mgpu_ctx *c, float2 *buf, float2 *oy) { return dual_launch<N, T>(c, buf, oy); }
};
template <int N, int T>
struct PofkF {
  static int run(mgpu_ctx *c, float2 *buf, double *pk) { return pofk_launch<N, T>(c, buf, pk); }
};
template <int N, int T>
struct FusedXF {
  static int run(mgpu_ctx *c, const FusedXArgs &a) { return fused_x_launch<N, T>(c, a); }
};
template <int N, int T>
struct EpiF {
  static int run(mgpu_ctx *c, float2 *buf, float cp, float m2) { return epi_launch<N, T>(c, buf, cp, m2); }
};


}  // namespace

// ---------------------------------------------------------------- host entry points

int fused_green_x(mgpu_ctx *ctx, int mode, const float2 *a, const float2 *b, float2 *o0, float2 *o1, double p0,
                  double p1, int filter) {
  GreenArgs ga{0.f, 0.f, 0.f, filter};
  if (mode == GREEN_FORCE) ga.cnorm = (float)p0, ga.mult = ctx->mlp ? ctx->green_mult : nullptr;
  else ga.nf = (float)p0;
  ga.kfac = (float)p1;
  const double modes = (double)ctx->N * ctx->N * (ctx->N / 2 + 1);
  KScope ks(ctx, mode == GREEN_FORCE ? MGPU_K_FFT_GREEN : MGPU_K_FFT_POT,
            8.0 * modes * ((b ? 2 : 1) + (mode == GREEN_FORCE ? 2 : 1)));
  switch (ctx->N) {
#define X(S) \
  case S:    \
    return by_tile<S, GreenF>(ctx, mode, a, b, o0, o1, ga);
    MGPU_FFT_SIZES(X)
#undef X
    default:
      return mgpu_fail(ctx, MGPU_E_INVALID, "unsupported FFT size");
  }
}

int fused_x_pair(mgpu_ctx *ctx, int fwd, int mode, const float2 *in, float2 *kout, const float2 *dk, float2 *o0,
                 float2 *o1, double p0, double p1, int filter, double coupling, double massterm2, double *pk) {
  FusedXArgs a;
  a.fwd = fwd;
  a.mode = mode;
  a.in = in;
  a.kout = kout;
  a.dk = dk;
  a.o0 = o0;
  a.o1 = o1;
  a.ga = GreenArgs{0.f, 0.f, 0.f, filter};
  if (mode == GREEN_FORCE) a.ga.cnorm = (float)p0, a.ga.mult = ctx->mlp ? ctx->green_mult : nullptr;
  else a.ga.nf = (float)p0;
  a.ga.kfac = (float)p1;
  a.coupling = (float)coupling;
  a.m2 = (float)massterm2;
  a.pk = pk;
  // compulsory: the input (and delta(k)) read, the outputs (and delta(k)) written
  const double modes = (double)ctx->N * ctx->N * (ctx->N / 2 + 1);
  const int nio = 1 + (dk ? 1 : 0) + (kout ? 1 : 0) + (mode == GREEN_FORCE ? 2 : 1);
  KScope ks(ctx, fwd == FWD_DENSITY ? MGPU_K_FFT_XDENS : MGPU_K_FFT_XFORCE, 8.0 * modes * nio);
  switch (ctx->N) {
#define X(S)                                                               \
  case S:                                                                  \
    /* the force pairs hold G through the first inverse: 8-column tiles */ \
    if (mode == GREEN_FORCE) return FusedXF<S, 8>::run(ctx, a);            \
    return by_tile<S, FusedXF>(ctx, a);
    MGPU_FFT_SIZES(X)
#undef X
    default:
      return mgpu_fail(ctx, MGPU_E_INVALID, "unsupported FFT size");
  }
}

int fused_dual_y(mgpu_ctx *ctx, float2 *buf, float2 *oy) {
  const double modes = (double)ctx->N * ctx->N * (ctx->N / 2 + 1);
  KScope ks(ctx, MGPU_K_FFT_DUAL, 8.0 * modes * 3);
  switch (ctx->N) {
#define X(S) \
  case S:    \
    return by_tile<S, DualF>(ctx, buf, oy);
    MGPU_FFT_SIZES(X)
#undef X
    default:
      return mgpu_fail(ctx, MGPU_E_INVALID, "unsupported FFT size");
  }
}

int fused_epi_phi_x(mgpu_ctx *ctx, float2 *buf, double coupling, double massterm2) {
  const double modes = (double)ctx->N * ctx->N * (ctx->N / 2 + 1);
  KScope ks(ctx, MGPU_K_FFT_EPI, 16.0 * modes);
  switch (ctx->N) {
#define X(S) \
  case S:    \
    return by_tile<S, EpiF>(ctx, buf, (float)coupling, (float)massterm2);
    MGPU_FFT_SIZES(X)
#undef X
    default:
      return mgpu_fail(ctx, MGPU_E_INVALID, "unsupported FFT size");
  }
}

int fused_z_screen(mgpu_ctx *ctx, int density_kind, float2 *buf, const float *dx, double p0, double p1) {
  const double rows = (double)ctx->N * ctx->N;
  KScope ks(ctx, MGPU_K_FFT_SCREEN, rows * (16.0 * (ctx->N / 2 + 1) + 4.0 * ctx->N));
  const int kind = density_kind ? SCREEN_DENSITY : SCREEN_POTENTIAL;
  switch (ctx->N) {
#define X(S) \
  case S:    \
    return screen_launch<S>(ctx, kind, buf, (const float2 *)dx, (float)p0, (float)p1);
    MGPU_FFT_SIZES(X)
#undef X
    default:
      return mgpu_fail(ctx, MGPU_E_INVALID, "unsupported FFT size");
  }
}

int fused_pofk_x(mgpu_ctx *ctx, float2 *buf, double *pk) {
  const double modes = (double)ctx->N * ctx->N * (ctx->N / 2 + 1);
  KScope ks(ctx, MGPU_K_FFT_PK, 16.0 * modes);
  switch (ctx->N) {
#define X(S) \
  case S:    \
    return by_tile<S, PofkF>(ctx, buf, pk);
    MGPU_FFT_SIZES(X)
#undef X
    default:
      return mgpu_fail(ctx, MGPU_E_INVALID, "unsupported FFT size");
  }
}
