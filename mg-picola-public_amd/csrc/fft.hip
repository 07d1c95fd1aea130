// 3D real<->complex FFT for gfx950, written for this project (replaces the
// FFTW-MPI plans of src/wrappers.c:63-94 and src/auxPM.c:49,68-70).
//
// A transform is three HBM passes over a mesh slot:
//   z pass   contiguous rows: a length-N real row is an N/2-point complex FFT
//            plus the split step; in place, forward (r2c) or inverse (c2r);
//   y pass   strided columns of the half spectrum, tiles of T columns;
//   x pass   same, stride N*HP.
// Each pass stages a tile in LDS with coalesced 128-byte row segments, runs a
// mixed-radix (8/4/2/3) Stockham FFT on it in LDS, and writes it back.
// Twiddles come from one table W_N^m computed on the host in double.
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

constexpr int pick_radix(int rem, int e) {
  return (rem % 8 == 0 && e % 8 == 0)   ? 8
         : (rem % 4 == 0 && e % 4 == 0) ? 4
         : (rem % 2 == 0 && e % 2 == 0) ? 2
         : (rem % 3 == 0 && e % 3 == 0) ? 3
                                        : 0;
}
// elements held per thread for a length-L transform
constexpr int elems_per_thread(int L) { return (L % 3 == 0) ? (L < 24 ? L : 24) : (L < 16 ? L : 16); }

// Stockham DIT passes on B transforms of length L held in LDS; element (b, n)
// at lds[b*BS + n*ES].  BFAST maps consecutive butterflies to consecutive b.
template <int L, int B, int NT, int SIGN, int BS, int ES, bool BFAST, int Ns>
__device__ __forceinline__ void stockham(float2 *lds, const float2 *__restrict__ tw, int twstride, int tid) {
  if constexpr (Ns < L) {
    constexpr int E = L * B / NT;
    constexpr int R = pick_radix(L / Ns, E);
    static_assert(R > 0, "size not factorable with the per-thread element count");
    constexpr int NB = E / R;
    constexpr int M = L / R;
    float2 v[NB][R];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int g = tid + q * NT;
      int b, j;
      if constexpr (BFAST) { b = g % B; j = g / B; } else { j = g % M; b = g / M; }
      const int k = j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 x = lds[b * BS + (j + r * M) * ES];
        if (Ns > 1 && r > 0) {
          float2 w = tw[(r * k * (L / (Ns * R))) * twstride];
          if (SIGN > 0) w.y = -w.y;
          x = cmul(x, w);
        }
        v[q][r] = x;
      }
      Bfly<R, SIGN>::run(v[q]);
    }
    __syncthreads();
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
    stockham<L, B, NT, SIGN, BS, ES, BFAST, Ns * R>(lds, tw, twstride, tid);
  }
}

// ---------------------------------------------------------------- strided column pass
template <int L, int T>
struct ColCfg {
  static constexpr int E = elems_per_thread(L);
  static constexpr int NT = L * T / E;
  static_assert(NT <= 1024, "column tile too large");
};

constexpr int col_threads(int L, int T) { return L * T / elems_per_thread(L); }

template <int L, int T, int SIGN>
__global__ __launch_bounds__(col_threads(L, T)) void k_fft_col(float2 *__restrict__ buf, size_t line_stride,
                                                                 size_t outer_stride,
                                                                 const float2 *__restrict__ tw) {
  extern __shared__ float2 lds[];
  constexpr int NT = ColCfg<L, T>::NT;
  const int tid = threadIdx.x;
  float2 *base = buf + (size_t)blockIdx.y * outer_stride + (size_t)blockIdx.x * T;
  for (int e = tid; e < L * T; e += NT) {
    const int c = e % T, n = e / T;
    lds[n * T + c] = base[(size_t)n * line_stride + c];
  }
  __syncthreads();
  stockham<L, T, NT, SIGN, 1, T, true, 1>(lds, tw, 1, tid);
  for (int e = tid; e < L * T; e += NT) {
    const int c = e % T, n = e / T;
    base[(size_t)n * line_stride + c] = lds[n * T + c];
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

// r2c along z: real row (N floats at complex offset row*HP) -> X[0..N/2], zero padding.
template <int N>
__global__ __launch_bounds__(RowCfg<N>::NT) void k_fft_z_r2c(float2 *__restrict__ buf, int HP, size_t nrows,
                                                              const float2 *__restrict__ tw) {
  using C = RowCfg<N>;
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x;
  const size_t row0 = (size_t)blockIdx.x * C::RB;
  for (int e = tid; e < C::RB * C::M; e += C::NT) {
    const int r = e / C::M, n = e % C::M;
    const size_t row = row0 + r;
    lds[r * C::LS + n] = row < nrows ? buf[row * HP + n] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  stockham<C::M, C::RB, C::NT, -1, C::LS, 1, false, 1>(lds, tw, 2, tid);
  for (int e = tid; e < C::RB * HP; e += C::NT) {
    const int r = e / HP, k = e % HP;
    const size_t row = row0 + r;
    if (row >= nrows) continue;
    float2 X = make_float2(0.f, 0.f);
    if (k <= C::M) {
      const float2 zk = lds[r * C::LS + (k % C::M)];
      const float2 zm = cconj(lds[r * C::LS + ((C::M - k) % C::M)]);
      const float2 a = cadd(zk, zm), b = csub(zk, zm);
      const float2 ib = mul_i<+1>(cmul(tw[k], b));
      X = make_float2(0.5f * (a.x - ib.x), 0.5f * (a.y - ib.y));
    }
    buf[row * HP + k] = X;
  }
}

// c2r along z: X[0..N/2] (Im of ends dropped) -> real row.  The row is written
// to out with a row stride of orow float2 (out = buf, orow = HP in place, the
// rest of the row zero padding; or a compact [N^2][N] real array, orow = N/2).
template <int N>
__global__ __launch_bounds__(RowCfg<N>::NT) void k_fft_z_c2r(float2 *__restrict__ buf, int HP, size_t nrows,
                                                              const float2 *__restrict__ tw, float2 *__restrict__ out,
                                                              int orow) {
  using C = RowCfg<N>;
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x;
  const size_t row0 = (size_t)blockIdx.x * C::RB;
  for (int e = tid; e < C::RB * (C::M + 1); e += C::NT) {
    const int r = e / (C::M + 1), k = e % (C::M + 1);
    const size_t row = row0 + r;
    lds[r * C::LS + k] = row < nrows ? buf[row * HP + k] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  constexpr int HALF = C::M / 2 + 1;
  for (int e = tid; e < C::RB * HALF; e += C::NT) {
    const int r = e / HALF, k = e % HALF;
    float2 *row = lds + r * C::LS;
    if (k == 0) {
      const float x0 = row[0].x, xm = row[C::M].x;
      row[0] = make_float2(x0 + xm, x0 - xm);
    } else {
      const int kk = C::M - k;
      const float2 xk = row[k], xm = row[kk];
      // Z[k] = (X_k + X*_{M-k}) + i w^k (X_k - X*_{M-k}),  w = e^{+2 pi i/N}
      const float2 wk = cconj(tw[k]), wkk = cconj(tw[kk]);
      const float2 zk = cadd(cadd(xk, cconj(xm)), mul_i<+1>(cmul(wk, csub(xk, cconj(xm)))));
      const float2 zkk = cadd(cadd(xm, cconj(xk)), mul_i<+1>(cmul(wkk, csub(xm, cconj(xk)))));
      row[k] = zk;
      if (kk != k) row[kk] = zkk;
    }
  }
  __syncthreads();
  stockham<C::M, C::RB, C::NT, +1, C::LS, 1, false, 1>(lds, tw, 2, tid);
  for (int e = tid; e < C::RB * orow; e += C::NT) {
    const int r = e / orow, k = e % orow;
    const size_t row = row0 + r;
    if (row >= nrows) continue;
    out[row * orow + k] = k < C::M ? lds[r * C::LS + k] : make_float2(0.f, 0.f);
  }
}

#define MGPU_FFT_SIZES(X) \
  X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1536) X(2048)

template <int L, int T, int SIGN>
int launch_col(mgpu_ctx *ctx, float2 *buf, size_t line_stride, size_t outer_stride, int nouter, int ntiles,
               const float2 *tw) {
  size_t lds = (size_t)L * T * sizeof(float2);
  auto kern = k_fft_col<L, T, SIGN>;
  if (lds > 65536) MGPU_HIP(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  dim3 grid(ntiles, nouter);
  hipLaunchKernelGGL(kern, grid, dim3(ColCfg<L, T>::NT), lds, ctx->stream, buf, line_stride, outer_stride, tw);
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
    hipLaunchKernelGGL(kern, grid, dim3(C::NT), lds, ctx->stream, buf, HP, nrows, tw, out ? out : buf,
                       out ? N / 2 : HP);
  }
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}

template <int N>
int dispatch_col(mgpu_ctx *ctx, const FFTPlan *plan, float2 *buf, size_t ls, size_t os, int forward) {
  const int nouter = N;
  if constexpr (N * 16 / elems_per_thread(N) <= 1024) {
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

int fft_plan_init(mgpu_ctx *ctx, FFTPlan *plan, int n) {
  if (!fft_size_supported(n)) return mgpu_fail(ctx, MGPU_E_INVALID, "unsupported FFT size " + std::to_string(n));
  plan->n = n;
  plan->hp = ((n / 2 + 1) + 15) / 16 * 16;
  // 16-column tiles when the tile fits the LDS of one workgroup, else 8
  size_t need16 = (size_t)n * 16 * sizeof(float2);
  plan->tcols = (need16 <= ctx->lds_max && n * 16 / (n % 3 == 0 ? 24 : 16) <= 1024) ? 16 : 8;
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
  KScope ks(ctx, MGPU_K_FFT_Z, rows * (plan->n * 4.0 + (plan->n / 2 + 1) * 8.0));
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
  KScope ks(ctx, axis == 1 ? MGPU_K_FFT_Y : MGPU_K_FFT_X, 2.0 * 8.0 * (double)N * N * (N / 2 + 1));
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
