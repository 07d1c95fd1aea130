// 3D real<->complex FFT for gfx950, written for this project (replaces the
// FFTW-MPI plans of src/wrappers.c:63-94 and src/auxPM.c:49,68-70).
//
// A transform is three HBM passes over a mesh slot:
//   z pass   contiguous rows: a length-N real row is an N/2-point complex FFT
//            plus the split (r2c) or merge (c2r) step;
//   y pass   strided columns of the half spectrum, tiles of T columns;
//   x pass   same, stride N*HP.
// Every pass is a register-resident Stockham FFT: the first radix stage reads
// its operands straight from HBM (coalesced: consecutive lanes read
// consecutive columns or consecutive row elements), the last stage writes its
// results straight back, and only the stages in between exchange data through
// LDS.  Radix 16 stages keep a 1024-point transform to three stages, two LDS
// exchanges.  Twiddles come from one table W_N^m computed on the host in
// double precision.
//
// Semantics follow FFTW's rdft2: forward e^{-i}, backward e^{+i}, both
// unnormalised; the c2r drops Im of the kz = 0 and kz = N/2 entries.
#include "mgpu_internal.h"

#include <vector>
#include <cmath>

namespace {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by SIGN * i
template <int SIGN>
__device__ __forceinline__ float2 mul_i(float2 a) {
  return SIGN > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <int R, int SIGN>
struct Bfly;

template <int SIGN>
struct Bfly<2, SIGN> {
  __device__ __forceinline__ static void run(float2 *v) {
    float2 t = v[0];
    v[0] = cadd(t, v[1]);
    v[1] = csub(t, v[1]);
  }
};

template <int SIGN>
struct Bfly<3, SIGN> {
  __device__ __forceinline__ static void run(float2 *v) {
    const float s3 = 0.86602540378443864676f;
    float2 t1 = cadd(v[1], v[2]);
    float2 t2 = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
    float2 u = cscale(mul_i<SIGN>(csub(v[1], v[2])), s3);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, u);
    v[2] = csub(t2, u);
  }
};

template <int SIGN>
struct Bfly<4, SIGN> {
  __device__ __forceinline__ static void run(float2 *v) {
    float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    float2 b0 = cadd(v[1], v[3]), b1 = mul_i<SIGN>(csub(v[1], v[3]));
    v[0] = cadd(a0, b0);
    v[1] = cadd(a1, b1);
    v[2] = csub(a0, b0);
    v[3] = csub(a1, b1);
  }
};

template <int SIGN>
struct Bfly<8, SIGN> {
  __device__ __forceinline__ static void run(float2 *v) {
    const float c = 0.70710678118654752440f;
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    Bfly<4, SIGN>::run(e);
    Bfly<4, SIGN>::run(o);
    // o1 *= c(1 + S i), o2 *= S i, o3 *= c(-1 + S i)
    float2 o1 = make_float2(c * (o[1].x - SIGN * o[1].y), c * (o[1].y + SIGN * o[1].x));
    float2 o2 = mul_i<SIGN>(o[2]);
    float2 o3 = make_float2(c * (-o[3].x - SIGN * o[3].y), c * (-o[3].y + SIGN * o[3].x));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

// 16 points as 4 x 4: n = n1 + 4 n2, k = k2 + 4 k1;
// A[n1][k2] = DFT4_n2(v[n1 + 4 n2]) * W16^(n1 k2), X[k2 + 4 k1] = DFT4_n1(A[n1][k2]).
template <int SIGN>
struct Bfly<16, SIGN> {
  __device__ __forceinline__ static void run(float2 *v) {
    const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f, h = 0.70710678118654752440f;
    float2 a[4][4];
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) {
      a[n1][0] = v[n1];
      a[n1][1] = v[n1 + 4];
      a[n1][2] = v[n1 + 8];
      a[n1][3] = v[n1 + 12];
      Bfly<4, SIGN>::run(a[n1]);
    }
    // twiddles W16^(n1 k2) = exp(SIGN 2 pi i n1 k2 / 16)
    const float sg = (float)SIGN;
    const float2 w1 = make_float2(c1, sg * s1), w2 = make_float2(h, sg * h), w3 = make_float2(s1, sg * c1);
    const float2 w6 = make_float2(-h, sg * h), w9 = make_float2(-c1, -sg * s1);
    a[1][1] = cmul(a[1][1], w1);
    a[1][2] = cmul(a[1][2], w2);
    a[1][3] = cmul(a[1][3], w3);
    a[2][1] = cmul(a[2][1], w2);
    a[2][2] = mul_i<SIGN>(a[2][2]);
    a[2][3] = cmul(a[2][3], w6);
    a[3][1] = cmul(a[3][1], w3);
    a[3][2] = cmul(a[3][2], w6);
    a[3][3] = cmul(a[3][3], w9);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
      float2 b[4] = {a[0][k2], a[1][k2], a[2][k2], a[3][k2]};
      Bfly<4, SIGN>::run(b);
      v[k2] = b[0];
      v[k2 + 4] = b[1];
      v[k2 + 8] = b[2];
      v[k2 + 12] = b[3];
    }
  }
};

constexpr int pick_radix(int rem, int e) {
  return (rem % 16 == 0 && e % 16 == 0)  ? 16
         : (rem % 8 == 0 && e % 8 == 0)  ? 8
         : (rem % 4 == 0 && e % 4 == 0)  ? 4
         : (rem % 2 == 0 && e % 2 == 0)  ? 2
         : (rem % 3 == 0 && e % 3 == 0)  ? 3
                                         : 0;
}
// elements held per thread for a length-L transform
constexpr int elems_per_thread(int L) { return (L % 3 == 0) ? (L < 24 ? L : 24) : (L < 16 ? L : 16); }

// Stockham DIT on B independent transforms of length L, E = L*B/NT elements per
// thread.  The first stage reads element n = j + r*M of transform b through
// ld(b, j, r*M); the last stage writes it through st(b, j, r*M, value); stages
// in between exchange through LDS, element
// (b, n) at lds[b*BS + n*ES].  BFAST maps consecutive lanes to consecutive b.
// With IN_LDS the first stage's input lives in the same LDS buffer, so a
// barrier separates those reads from the first exchange write.  With pre !=
// nullptr the first stage's operands are already in registers, pre[q*R + r]
// holding what ld(b, j, r*M) would return for butterfly q (see first_loads).
template <int L, int B, int NT, int SIGN, int BS, int ES, bool BFAST, bool IN_LDS, int Ns, class LD, class ST>
__device__ __forceinline__ void fft_stages(float2 *lds, const float2 *__restrict__ tw, int twstride, int tid,
                                           const LD &ld, const ST &st, const float2 *pre = nullptr) {
  constexpr int E = L * B / NT;
  constexpr int R = pick_radix(L / Ns, E);
  static_assert(R > 0, "size not factorable with the per-thread element count");
  constexpr int NB = E / R;
  constexpr int M = L / R;
  constexpr bool FIRST = (Ns == 1), LAST = (Ns * R == L);
  float2 v[NB][R];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int g = tid + q * NT;
    int b, j;
    if constexpr (BFAST) { b = g % B; j = g / B; } else { j = g % M; b = g / M; }
    const int k = j % Ns;
    // twiddles W^(r k) of this stage: for radix 8 and 16, r = r1 + 4 r2 and
    // W^(r k) = W^(r1 k) W^(4 r2 k), so only W^(k), W^(2k), W^(3k) and the
    // W^(4 r2 k) are loaded (one extra rounding for the products)
    constexpr int R2 = (R == 16 || R == 8) ? R / 4 : 1;
    constexpr int R1 = R / R2;
    float2 wa[R1], wb[R2];
    if constexpr (!FIRST) {
      const int step = k * (L / (Ns * R));
#pragma unroll
      for (int r1 = 1; r1 < R1; ++r1) {
        float2 w = tw[(r1 * step) * twstride];
        wa[r1] = SIGN > 0 ? cconj(w) : w;
      }
#pragma unroll
      for (int r2 = 1; r2 < R2; ++r2) {
        float2 w = tw[(R1 * r2 * step) * twstride];
        wb[r2] = SIGN > 0 ? cconj(w) : w;
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      float2 x;
      if constexpr (FIRST) {
        x = pre ? pre[q * R + r] : ld(b, j, r * M);
      } else {
        x = lds[b * BS + (j + r * M) * ES];
        const int r1 = r % R1, r2 = r / R1;
        if (r1 > 0 && r2 > 0) x = cmul(x, cmul(wa[r1], wb[r2]));
        else if (r1 > 0) x = cmul(x, wa[r1]);
        else if (r2 > 0) x = cmul(x, wb[r2]);
      }
      v[q][r] = x;
    }
  }
#pragma unroll
  for (int q = 0; q < NB; ++q) Bfly<R, SIGN>::run(v[q]);
  if constexpr (LAST) {
    // last stage: output index j + r*Ns with j < Ns
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int g = tid + q * NT;
      int b, j;
      if constexpr (BFAST) { b = g % B; j = g / B; } else { j = g % M; b = g / M; }
#pragma unroll
      for (int r = 0; r < R; ++r) st(b, j, r * M, v[q][r]);
    }
  } else {
    if constexpr (!FIRST || IN_LDS) __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int g = tid + q * NT;
      int b, j;
      if constexpr (BFAST) { b = g % B; j = g / B; } else { j = g % M; b = g / M; }
      const int k = j % Ns;
      const int d = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) lds[b * BS + (d + r * Ns) * ES] = v[q][r];
    }
    __syncthreads();
    fft_stages<L, B, NT, SIGN, BS, ES, BFAST, false, Ns * R>(lds, tw, twstride, tid, ld, st);
  }
}

// The first stage's operand loads of fft_stages, into registers.
template <int L, int B, int NT, bool BFAST, class LD>
__device__ __forceinline__ void first_loads(float2 *pre, int tid, const LD &ld) {
  constexpr int E = L * B / NT;
  constexpr int R = pick_radix(L, E);
  constexpr int NB = E / R;
  constexpr int M = L / R;
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int g = tid + q * NT;
    int b, j;
    if constexpr (BFAST) { b = g % B; j = g / B; } else { j = g % M; b = g / M; }
#pragma unroll
    for (int r = 0; r < R; ++r) pre[q * R + r] = ld(b, j, r * M);
  }
}

// ---------------------------------------------------------------- strided column pass
constexpr int col_ept(int L, int T) { return elems_per_thread(L); }
constexpr int col_threads(int L, int T) { return L * T / col_ept(L, T); }

// TPB tiles per block, at outer indices blockIdx.y*TPB + i: the first-stage
// loads of tile i+1 are issued before the stages of tile i run, so HBM reads
// overlap the LDS exchanges of the previous tile.
template <int L, int T, int SIGN, int TPB>
__global__ __launch_bounds__(col_threads(L, T)) void k_fft_col(float2 *__restrict__ buf, size_t line_stride,
                                                                 size_t outer_stride,
                                                                 const float2 *__restrict__ tw) {
  extern __shared__ float2 lds[];
  constexpr int NT = col_threads(L, T);
  constexpr int E = col_ept(L, T);
  static_assert(NT <= 1024, "column tile too large");
  const int tid = threadIdx.x;
  auto tile = [&](int i) { return buf + (size_t)(blockIdx.y * TPB + i) * outer_stride + (size_t)blockIdx.x * T; };
  if constexpr (TPB == 1) {
    float2 *base = tile(0);
    auto ld = [&](int c, int j, int o) { return (base + (size_t)j * line_stride + c)[(size_t)o * line_stride]; };
    auto st = [&](int c, int j, int o, float2 v) { (base + (size_t)j * line_stride + c)[(size_t)o * line_stride] = v; };
    fft_stages<L, T, NT, SIGN, 1, T, true, false, 1>(lds, tw, 1, tid, ld, st);
    return;
  }
  float2 pre[E];
  {
    const float2 *base = tile(0);
    first_loads<L, T, NT, true>(
        pre, tid, [&](int c, int j, int o) { return (base + (size_t)j * line_stride + c)[(size_t)o * line_stride]; });
  }
#pragma unroll
  for (int i = 0; i < TPB; ++i) {
    float2 cur[E];
#pragma unroll
    for (int e = 0; e < E; ++e) cur[e] = pre[e];
    if (i + 1 < TPB) {
      const float2 *nb = tile(i + 1);
      first_loads<L, T, NT, true>(
          pre, tid, [&](int c, int j, int o) { return (nb + (size_t)j * line_stride + c)[(size_t)o * line_stride]; });
    }
    float2 *base = tile(i);
    auto ld = [&](int c, int j, int o) { return (base + (size_t)j * line_stride + c)[(size_t)o * line_stride]; };
    auto st = [&](int c, int j, int o, float2 v) { (base + (size_t)j * line_stride + c)[(size_t)o * line_stride] = v; };
    fft_stages<L, T, NT, SIGN, 1, T, true, false, 1>(lds, tw, 1, tid, ld, st, cur);
    if (i + 1 < TPB) __syncthreads();  // the next tile's first exchange reuses the LDS
  }
}

// ---------------------------------------------------------------- contiguous z pass
template <int N>
struct RowCfg {
  static constexpr int M = N / 2;
  static constexpr int E = elems_per_thread(M);
  static constexpr int RB = 256 * E / M > 0 ? 256 * E / M : 1;  // rows per block
  static constexpr int NT = M * RB / E;
  static constexpr int LS = M + 1;  // LDS row stride
  static_assert(NT <= 1024, "row tile too large");
};

// r2c along z: real row (N floats at complex offset row*HP) -> X[0..N/2].
// Entries N/2 < k < HP of the row are padding and are left untouched.
template <int N>
__global__ __launch_bounds__(RowCfg<N>::NT) void k_fft_z_r2c(float2 *__restrict__ buf, int HP, size_t nrows,
                                                              const float2 *__restrict__ tw) {
  using C = RowCfg<N>;
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x;
  const size_t row0 = (size_t)blockIdx.x * C::RB;
  // rows past the end (only when N*N is not a multiple of RB) read the last
  // row; their results are never stored
  auto ld = [&](int r, int j, int o) {
    const size_t row = row0 + r < nrows ? row0 + r : nrows - 1;
    return buf[row * HP + j + o];
  };
  auto st = [&](int r, int j, int o, float2 v) { lds[r * C::LS + j + o] = v; };
  constexpr int K = C::M + 1;
  constexpr int ITER = (C::RB * K + C::NT - 1) / C::NT;
  float2 w[ITER];
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    const int e = tid + it * C::NT;
    w[it] = e < C::RB * K ? tw[e % K] : make_float2(0.f, 0.f);
  }
  fft_stages<C::M, C::RB, C::NT, -1, C::LS, 1, false, false, 1>(lds, tw, 2, tid, ld, st);
  __syncthreads();
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    const int e = tid + it * C::NT;
    if (e >= C::RB * K) break;
    const int r = e / K, k = e % K;
    const size_t row = row0 + r;
    if (row >= nrows) continue;
    const float2 zk = lds[r * C::LS + (k % C::M)];
    const float2 zm = cconj(lds[r * C::LS + ((C::M - k) % C::M)]);
    const float2 a = cadd(zk, zm), b = csub(zk, zm);
    const float2 ib = mul_i<+1>(cmul(w[it], b));
    buf[row * HP + k] = make_float2(0.5f * (a.x - ib.x), 0.5f * (a.y - ib.y));
  }
}

// c2r along z: X[0..N/2] (Im of ends dropped) -> real row.  The row is written
// to out with a row stride of orow float2 (out = buf, orow = HP in place; or a
// compact [N^2][N] real array, orow = N/2).
template <int N>
__global__ __launch_bounds__(RowCfg<N>::NT) void k_fft_z_c2r(const float2 *buf, int HP, size_t nrows,
                                                              const float2 *__restrict__ tw, float2 *out, int orow) {
  using C = RowCfg<N>;
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x;
  const size_t row0 = (size_t)blockIdx.x * C::RB;
  constexpr int HALF = C::M / 2 + 1;
  constexpr int ITER = (C::RB * HALF + C::NT - 1) / C::NT;
  // merge: Z[k] = (X_k + X*_{M-k}) + i w^k (X_k - X*_{M-k}), w = e^{+2 pi i/N};
  // every global load of the block's rows is issued before any is used
  float2 xk[ITER], xm[ITER], wk[ITER], wm[ITER];
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    // out-of-range slots load from a valid address and are discarded below
    const int e = tid + it * C::NT < C::RB * HALF ? tid + it * C::NT : C::RB * HALF - 1;
    const int r = e / HALF, k = e % HALF;
    const size_t row = row0 + r < nrows ? row0 + r : nrows - 1;
    const float2 *g = buf + row * HP;
    xk[it] = g[k];
    xm[it] = g[C::M - k];
    wk[it] = tw[k];
    wm[it] = tw[C::M - k];
  }
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    const int e = tid + it * C::NT;
    if (e >= C::RB * HALF) break;
    const int r = e / HALF, k = e % HALF;
    float2 *lrow = lds + r * C::LS;
    if (k == 0) {
      lrow[0] = make_float2(xk[it].x + xm[it].x, xk[it].x - xm[it].x);
    } else {
      const int kk = C::M - k;
      const float2 a = xk[it], b = xm[it];
      const float2 zk = cadd(cadd(a, cconj(b)), mul_i<+1>(cmul(cconj(wk[it]), csub(a, cconj(b)))));
      const float2 zkk = cadd(cadd(b, cconj(a)), mul_i<+1>(cmul(cconj(wm[it]), csub(b, cconj(a)))));
      lrow[k] = zk;
      if (kk != k) lrow[kk] = zkk;
    }
  }
  __syncthreads();
  auto ld = [&](int r, int j, int o) { return lds[r * C::LS + j + o]; };
  auto st = [&](int r, int j, int o, float2 v) {
    const size_t row = row0 + r;
    if (row < nrows) out[row * orow + j + o] = v;
  };
  fft_stages<C::M, C::RB, C::NT, +1, C::LS, 1, false, true, 1>(lds, tw, 2, tid, ld, st);
}

#define MGPU_FFT_SIZES(X) \
  X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1536) X(2048)

// tiles per block of the column passes: two (software-pipelined) for the
// large power-of-two sizes, where the registers for the prefetch are there
constexpr int col_tpb(int L) { return (L % 3 != 0 && L >= 256) ? 2 : 1; }

template <int L, int T, int SIGN>
int launch_col(mgpu_ctx *ctx, float2 *buf, size_t line_stride, size_t outer_stride, int nouter, int ntiles,
               const float2 *tw) {
  size_t lds = (size_t)L * T * sizeof(float2);
  constexpr int TPB = col_tpb(L);
  auto kern = k_fft_col<L, T, SIGN, TPB>;
  static_assert(L % TPB == 0, "outer count must be a multiple of the tiles per block");
  if (lds > 65536) MGPU_HIP(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  dim3 grid(ntiles, nouter / TPB);
  hipLaunchKernelGGL(kern, grid, dim3(col_threads(L, T)), lds, ctx->stream, buf, line_stride, outer_stride, tw);
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}

template <int N>
int launch_z(mgpu_ctx *ctx, float2 *buf, int HP, const float2 *tw, int forward, float2 *out = nullptr) {
  using C = RowCfg<N>;
  size_t nrows = (size_t)N * N;
  size_t lds = (size_t)C::RB * C::LS * sizeof(float2);
  dim3 grid(cdiv(nrows, C::RB));
  if (forward) {
    auto kern = k_fft_z_r2c<N>;
    if (lds > 65536) MGPU_HIP(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, grid, dim3(C::NT), lds, ctx->stream, buf, HP, nrows, tw);
  } else {
    auto kern = k_fft_z_c2r<N>;
    if (lds > 65536) MGPU_HIP(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, grid, dim3(C::NT), lds, ctx->stream, (const float2 *)buf, HP, nrows, tw,
                       out ? out : buf, out ? N / 2 : HP);
  }
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}

template <int N>
int dispatch_col(mgpu_ctx *ctx, const FFTPlan *plan, float2 *buf, size_t ls, size_t os, int forward) {
  const int nouter = N;
  if constexpr (col_threads(N, 16) <= 1024) {
    if (plan->tcols == 16) {
      const int ntiles = plan->hp / 16;
      return forward ? launch_col<N, 16, -1>(ctx, buf, ls, os, nouter, ntiles, plan->tw)
                     : launch_col<N, 16, +1>(ctx, buf, ls, os, nouter, ntiles, plan->tw);
    }
  }
  const int ntiles = plan->hp / 8;
  return forward ? launch_col<N, 8, -1>(ctx, buf, ls, os, nouter, ntiles, plan->tw)
                 : launch_col<N, 8, +1>(ctx, buf, ls, os, nouter, ntiles, plan->tw);
}

}  // namespace

bool fft_size_supported(int n) {
  switch (n) {
#define X(S) case S:
    MGPU_FFT_SIZES(X)
#undef X
    return true;
    default:
      return false;
  }
}

int fft_plan_init(mgpu_ctx *ctx, FFTPlan *plan, int n, int tcols) {
  if (!fft_size_supported(n)) return mgpu_fail(ctx, MGPU_E_INVALID, "unsupported FFT size " + std::to_string(n));
  plan->n = n;
  plan->hp = ((n / 2 + 1) + 15) / 16 * 16;
  // 16-column tiles only where requested and the tile fits one workgroup's LDS and thread limit
  size_t need16 = (size_t)n * 16 * sizeof(float2);
  plan->tcols = (tcols == 16 && need16 <= ctx->lds_max && n * 16 / (n % 3 == 0 ? 24 : 16) <= 1024) ? 16 : 8;
  std::vector<float2> h(n);
  for (int m = 0; m < n; m++) {
    long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)n;
    h[m] = make_float2((float)cosl(ang), (float)sinl(ang));
  }
  MGPU_HIP(ctx, hipMalloc(&plan->tw, sizeof(float2) * n));
  MGPU_HIP(ctx, hipMemcpy(plan->tw, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  return MGPU_OK;
}

void fft_plan_free(FFTPlan *plan) {
  if (plan->tw) (void)hipFree(plan->tw);
  plan->tw = nullptr;
}

int fft_pass_z(mgpu_ctx *ctx, const FFTPlan *plan, float2 *buf, int forward, float *out) {
  // compulsory traffic: N^2 rows of N reals one way and N/2+1 complex the other
  const double rows = (double)plan->n * plan->n;
  const int cat = plan->kcat >= 0 ? plan->kcat : (forward ? MGPU_K_FFT_R2C : MGPU_K_FFT_C2R);
  KScope ks(ctx, cat, rows * (plan->n * 4.0 + (plan->n / 2 + 1) * 8.0));
  switch (plan->n) {
#define X(S) \
  case S:    \
    return launch_z<S>(ctx, buf, plan->hp, plan->tw, forward, (float2 *)out);
    MGPU_FFT_SIZES(X)
#undef X
    default:
      return mgpu_fail(ctx, MGPU_E_INVALID, "unsupported FFT size");
  }
}

int fft_pass_col(mgpu_ctx *ctx, const FFTPlan *plan, float2 *buf, int axis, int forward) {
  const size_t N = plan->n, HP = plan->hp;
  // axis 1 (y): element stride HP, one tile row per x; axis 0 (x): stride N*HP, one per y
  const size_t ls = axis == 1 ? HP : N * HP;
  const size_t os = axis == 1 ? N * HP : HP;
  // compulsory traffic: read + write of the N*N*(N/2+1) half spectrum
  const int cat = plan->kcat >= 0 ? plan->kcat : (forward ? MGPU_K_FFT_COL_FWD : MGPU_K_FFT_COL_INV);
  KScope ks(ctx, cat, 2.0 * 8.0 * (double)N * N * (N / 2 + 1));
  switch (plan->n) {
#define X(S) \
  case S:    \
    return dispatch_col<S>(ctx, plan, buf, ls, os, forward);
    MGPU_FFT_SIZES(X)
#undef X
    default:
      return mgpu_fail(ctx, MGPU_E_INVALID, "unsupported FFT size");
  }
}

int fft_r2c_3d(mgpu_ctx *ctx, const FFTPlan *plan, float2 *buf) {
  PhaseTimer pt(ctx, &ctx->t.fft);
  int rc;
  if ((rc = fft_pass_z(ctx, plan, buf, 1, nullptr))) return rc;
  if ((rc = fft_pass_col(ctx, plan, buf, 1, 1))) return rc;
  return fft_pass_col(ctx, plan, buf, 0, 1);
}

int fft_c2r_3d(mgpu_ctx *ctx, const FFTPlan *plan, float2 *buf, float *out) {
  PhaseTimer pt(ctx, &ctx->t.fft);
  int rc;
  if ((rc = fft_pass_col(ctx, plan, buf, 0, 0))) return rc;
  if ((rc = fft_pass_col(ctx, plan, buf, 1, 0))) return rc;
  return fft_pass_z(ctx, plan, buf, 0, out);
}
