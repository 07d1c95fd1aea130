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
#include "fft_core.h"

#include <vector>
#include <cmath>

namespace {

using namespace fftc;


// TPB tiles per block, at outer indices blockIdx.y*TPB + i: the first-stage
// loads of tile i+1 are issued before the stages of tile i run, so HBM reads
// overlap the LDS exchanges of the previous tile.
template <int L, int T, int SIGN, int TPB>
__global__ __launch_bounds__(col_threads(L, T)) void k_fft_col(float2 *__restrict__ buf, size_t line_stride,
                                                                 size_t outer_stride,
                                                                 const float2 *__restrict__ tw, size_t zs) {
  buf = zbuf(buf, zs);
  extern __shared__ float2 lds[];
  constexpr int NT = col_threads(L, T);
  constexpr int E = col_ept(L, T);
  static_assert(NT <= 1024, "column tile too large");
  const int tid = threadIdx.x;
  int bx, by;
  xcd_tile(bx, by);
  auto tile = [&](int i) { return buf + (size_t)(by * TPB + i) * outer_stride + (size_t)bx * T; };
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
    fft_stages<L, T, NT, SIGN, 1, T, true, false, 1, false, true>(lds, tw, 1, tid, ld, st, cur);
    if (i + 1 < TPB) __syncthreads();  // the next tile's first exchange reuses the LDS
  }
}


// r2c along z: real row (N floats at complex offset row*HP) -> X[0..N/2].
// Entries N/2 < k < HP of the row are padding and are left untouched.
template <int N>
__global__ __launch_bounds__(RowCfg<N>::NT) void k_fft_z_r2c(const float2 *in, float2 *buf, int HP, size_t nrows,
                                                              const float2 *__restrict__ tw, size_t zs) {
  in = zbuf(in, zs);
  buf = zbuf(buf, zs);
  using C = RowCfg<N>;
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x;
  const size_t row0 = (size_t)blockIdx.x * C::RB;
  // rows past the end (only when N*N is not a multiple of RB) read the last
  // row; their results are never stored
  auto ld = [&](int r, int j, int o) {
    const size_t row = row0 + r < nrows ? row0 + r : nrows - 1;
    return in[row * HP + j + o];
  };
  auto st = [&](int r, int j, int o, float2 v) { lds[r * C::LS + j + o] = v; };
  float2 w[ZSplit<N>::ITER];
  z_split_twiddles<N>(w, tw, tid);
  fft_stages<C::M, C::RB, C::NT, -1, C::LS, 1, false, false, 1>(lds, tw, 2, tid, ld, st);
  __syncthreads();
  z_split<N>(lds, w, buf, HP, row0, nrows, tid);
}

// c2r along z: X[0..N/2] (Im of ends dropped) -> real row.  The row is written
// to out with a row stride of orow float2 (out = buf, orow = HP in place; or a
// compact [N^2][N] real array, orow = N/2).  IKZ multiplies X[k] by i k first
// (the z force component, k the kz index).
template <int N, bool IKZ>
__global__ __launch_bounds__(RowCfg<N>::NT) void k_fft_z_c2r(const float2 *buf, int HP, size_t nrows,
                                                              const float2 *__restrict__ tw, float2 *out, int orow,
                                                              size_t zs) {
  buf = zbuf(buf, zs);
  out = zbuf(out, zs);
  using C = RowCfg<N>;
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x;
  const size_t row0 = (size_t)blockIdx.x * C::RB;
  z_merge<N, IKZ>(lds, buf, HP, row0, nrows, tw, tid);
  __syncthreads();
  auto ld = [&](int r, int j, int o) { return lds[r * C::LS + j + o]; };
  auto st = [&](int r, int j, int o, float2 v) {
    const size_t row = row0 + r;
    if (row < nrows) out[row * orow + j + o] = v;
  };
  fft_stages<C::M, C::RB, C::NT, +1, C::LS, 1, false, true, 1>(lds, tw, 2, tid, ld, st);
}

// tiles per block of the column passes: two (software-pipelined) for the
// large power-of-two sizes, where the registers for the prefetch are there
#ifndef MGPU_COL_TPB
#define MGPU_COL_TPB 2
#endif
constexpr int col_tpb(int L) { return (L % 3 != 0 && L >= 256) ? MGPU_COL_TPB : 1; }

template <int L, int T, int SIGN>
int launch_col(mgpu_ctx *ctx, float2 *buf, size_t line_stride, size_t outer_stride, int nouter, int ntiles,
               const float2 *tw) {
  size_t lds = (size_t)L * T * sizeof(float2);
  constexpr int TPB = col_tpb(L);
  auto kern = k_fft_col<L, T, SIGN, TPB>;
  static_assert(L % TPB == 0, "outer count must be a multiple of the tiles per block");
  if (lds > 65536) MGPU_HIP(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  dim3 grid(ntiles, nouter / TPB, ctx->bz);
  hipLaunchKernelGGL(kern, grid, dim3(col_threads(L, T)), lds, ctx->stream, buf, line_stride, outer_stride, tw, ctx->zs);
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}

template <int N>
int launch_z(mgpu_ctx *ctx, float2 *buf, int HP, const float2 *tw, int forward, float2 *out = nullptr,
             const float2 *in = nullptr) {
  using C = RowCfg<N>;
  size_t nrows = (size_t)N * N;
  size_t lds = (size_t)C::RB * C::LS * sizeof(float2);
  dim3 grid(cdiv(nrows, C::RB), 1, ctx->bz);
  if (forward) {
    auto kern = k_fft_z_r2c<N>;
    if (lds > 65536) MGPU_HIP(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, grid, dim3(C::NT), lds, ctx->stream, in ? in : (const float2 *)buf, buf, HP, nrows, tw,
                       ctx->zs);
  } else {
    auto kern = k_fft_z_c2r<N, false>;
    if (lds > 65536) MGPU_HIP(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, grid, dim3(C::NT), lds, ctx->stream, (const float2 *)buf, HP, nrows, tw,
                       out ? out : buf, out ? N / 2 : HP, ctx->zs);
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

// 16-column tiles fit one workgroup's LDS and thread limit
bool fft_tile16_fits(const mgpu_ctx *ctx, int n) {
  const size_t need16 = (size_t)n * 16 * sizeof(float2);
  return need16 <= ctx->lds_max && n * 16 / elems_per_thread(n) <= 1024;
}

int fft_plan_init(mgpu_ctx *ctx, FFTPlan *plan, int n, int tcols) {
  if (!fft_size_supported(n)) return mgpu_fail(ctx, MGPU_E_INVALID, "unsupported FFT size " + std::to_string(n));
  plan->n = n;
  plan->hp = ((n / 2 + 1) + 15) / 16 * 16;
  plan->tcols = (tcols == 16 && fft_tile16_fits(ctx, n)) ? 16 : 8;
  std::vector<float2> h(n);
  for (int m = 0; m < n; m++) {
    long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)n;
    h[m] = make_float2((float)cosl(ang), (float)sinl(ang));
  }
  MGPU_HIP(ctx, hipMalloc(&plan->tw, sizeof(float2) * n));
  MGPU_HIP(ctx, hipMemcpy(plan->tw, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  return MGPU_OK;
}

// The column passes' tile width measured on this device: the forward and the
// inverse y pass over buf (a scratch mesh of the plan's size) with each width
// that fits, one warm-up and the best of three; the faster becomes
// plan->tcols.  The width changes no column's arithmetic, only the schedule.
int fft_autotune(mgpu_ctx *ctx, FFTPlan *plan, float2 *buf, double ms[2]) {
  const bool can16 = fft_tile16_fits(ctx, plan->n);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return mgpu_fail(ctx, MGPU_E_HIP, "event creation failed");
  }
  const int widths[2] = {8, 16};
  int rc = MGPU_OK;
  for (int w = 0; w < 2 && !rc; w++) {
    ms[w] = -1.0;
    if (widths[w] == 16 && !can16) continue;
    plan->tcols = widths[w];
    for (int rep = 0; rep < 4 && !rc; rep++) {
      if (hipEventRecord(e0, ctx->stream) != hipSuccess) rc = MGPU_E_HIP;
      if (!rc) rc = fft_pass_col(ctx, plan, buf, 1, 1);
      if (!rc) rc = fft_pass_col(ctx, plan, buf, 1, 0);
      float t = 0.f;
      if (!rc && (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                  hipEventElapsedTime(&t, e0, e1) != hipSuccess))
        rc = MGPU_E_HIP;
      if (!rc && rep > 0 && (ms[w] < 0 || t < ms[w])) ms[w] = t;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc == MGPU_E_HIP ? mgpu_fail(ctx, rc, "FFT autotune timing failed") : rc;
  plan->tcols = (ms[1] >= 0 && ms[1] <= ms[0]) ? 16 : 8;
  return MGPU_OK;
}

void fft_plan_free(FFTPlan *plan) {
  if (plan->tw) (void)hipFree(plan->tw);
  plan->tw = nullptr;
}

int fft_pass_z(mgpu_ctx *ctx, const FFTPlan *plan, float2 *buf, int forward, float *out, const float *in) {
  // compulsory traffic: N^2 rows of N reals one way and N/2+1 complex the other
  const double rows = (double)plan->n * plan->n;
  const int cat = plan->kcat >= 0 ? plan->kcat : (forward ? MGPU_K_FFT_R2C : MGPU_K_FFT_C2R);
  KScope ks(ctx, cat, rows * (plan->n * 4.0 + (plan->n / 2 + 1) * 8.0));
  switch (plan->n) {
#define X(S) \
  case S:    \
    return launch_z<S>(ctx, buf, plan->hp, plan->tw, forward, (float2 *)out, (const float2 *)in);
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

int fft_pass_z_ikz(mgpu_ctx *ctx, const FFTPlan *plan, float2 *buf) {
  const double rows = (double)plan->n * plan->n;
  KScope ks(ctx, MGPU_K_FFT_C2R, rows * (plan->n * 4.0 + (plan->n / 2 + 1) * 8.0));
  switch (plan->n) {
#define X(S)                                                                                                    \
  case S: {                                                                                                     \
    using C = RowCfg<S>;                                                                                        \
    const size_t nrows = (size_t)S * S, lds = (size_t)C::RB * C::LS * sizeof(float2);                           \
    auto kern = k_fft_z_c2r<S, true>;                                                                           \
    if (lds > 65536)                                                                                            \
      MGPU_HIP(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
    hipLaunchKernelGGL(kern, dim3(cdiv(nrows, C::RB), 1, ctx->bz), dim3(C::NT), lds, ctx->stream,              \
                       (const float2 *)buf, plan->hp, nrows, plan->tw, buf, plan->hp, ctx->zs);                 \
    MGPU_HIP(ctx, hipGetLastError());                                                                           \
    return MGPU_OK;                                                                                             \
  }
    MGPU_FFT_SIZES(X)
#undef X
    default:
      return mgpu_fail(ctx, MGPU_E_INVALID, "unsupported FFT size");
  }
}

int fft_r2c_3d(mgpu_ctx *ctx, const FFTPlan *plan, float2 *buf, const float *in) {
  PhaseTimer pt(ctx, &ctx->t.fft);
  int rc;
  if ((rc = fft_pass_z(ctx, plan, buf, 1, nullptr, in))) return rc;
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
