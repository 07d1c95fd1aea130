// FFT building blocks shared by the plain transforms (fft.hip) and the fused
// spectral passes (fused.hip): radix 2/3/4/8/16 butterflies and the
// register-resident Stockham stage recursion used by every pass kernel.
#pragma once

#include "mgpu_internal.h"

namespace fftc {

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
// (radix-3 lengths: 12 up to 768, so the z passes and the fused column passes of
// the 192-768 meshes run two or more waves per SIMD; 24 from 1536 on, where 12
// would need 1024-thread column blocks capped at 128 VGPRs)
constexpr int elems_per_thread(int L) {
  return (L % 3 == 0) ? (L < 12 ? L : L < 1536 ? 12 : 24) : (L < 16 ? L : 16);
}

// Stockham DIT on B independent transforms of length L, E = L*B/NT elements per
// thread.  The first stage reads element n = j + r*M of transform b through
// ld(b, j, r*M); the last stage writes it through st(b, j, r*M, value); stages
// in between exchange through LDS, element
// (b, n) at lds[b*BS + n*ES].  BFAST maps consecutive lanes to consecutive b.
// With IN_LDS the first stage's input lives in the same LDS buffer, so a
// barrier separates those reads from the first exchange write.  With PRE the
// first stage's operands are already in registers, pre[q*R + r] holding what
// ld(b, j, r*M) would return for butterfly q (see first_loads); a compile-time
// flag, not a null test on pre, so the register array is never addressed
// through a pointer at run time (which would place it in scratch memory).
// With SYNC_ST a barrier precedes the last stage's stores, so st may write the
// same LDS buffer (every thread has read its last-stage operands by then).
template <int L, int B, int NT, int SIGN, int BS, int ES, bool BFAST, bool IN_LDS, int Ns, bool SYNC_ST = false,
          bool PRE = false, class LD, class ST>
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
        if constexpr (PRE) x = pre[q * R + r];
        else x = ld(b, j, r * M);
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
    if constexpr (SYNC_ST) __syncthreads();
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
    fft_stages<L, B, NT, SIGN, BS, ES, BFAST, false, Ns * R, SYNC_ST>(lds, tw, twstride, tid, ld, st);
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

// Block -> tile order for 2D grids: workgroups are dealt to the 8 XCDs round
// robin, so give each XCD a contiguous range of tiles (x fastest); tiles that
// share cache lines or rows then meet in one L2.
__device__ __forceinline__ void xcd_tile(int &bx, int &by) {
  const unsigned nbx = gridDim.x, nblk = gridDim.x * gridDim.y, id = blockIdx.x + blockIdx.y * nbx;
  const unsigned lin = nblk % 8 == 0 ? (id % 8) * (nblk / 8) + id / 8 : id;
  bx = (int)(lin % nbx);
  by = (int)(lin / nbx);
}

// ---------------------------------------------------------------- strided column pass
constexpr int col_ept(int L, int T) { return elems_per_thread(L); }
constexpr int col_threads(int L, int T) { return L * T / col_ept(L, T); }

// ---------------------------------------------------------------- contiguous z pass
// threads per block of the z passes (rows per block = this / (N/2) x elements per thread):
// one wave per block measured fastest at 1024^3 (r2c 1.99 -> 1.79 ms, screening pass
// 4.32 -> 3.65 ms against 256 threads)
#ifndef MGPU_ZROW_THREADS
#define MGPU_ZROW_THREADS 64
#endif
template <int N>
struct RowCfg {
  static constexpr int M = N / 2;
  static constexpr int E = elems_per_thread(M);
  static constexpr int RB = MGPU_ZROW_THREADS * E / M > 0 ? MGPU_ZROW_THREADS * E / M : 1;  // rows per block
  static constexpr int NT = M * RB / E;
  static constexpr int LS = M + 1;  // LDS row stride
  static_assert(NT <= 1024, "row tile too large");
};

// ---- split (r2c) and merge (c2r) steps of the z pass on RB rows in LDS
template <int N>
struct ZSplit {
  static constexpr int K = RowCfg<N>::M + 1;
  static constexpr int ITER = (RowCfg<N>::RB * K + RowCfg<N>::NT - 1) / RowCfg<N>::NT;
};
// twiddles of the split step, loaded before the FFT stages so they are in flight early
template <int N>
__device__ __forceinline__ void z_split_twiddles(float2 *w, const float2 *__restrict__ tw, int tid) {
  using C = RowCfg<N>;
  constexpr int K = ZSplit<N>::K;
#pragma unroll
  for (int it = 0; it < ZSplit<N>::ITER; ++it) {
    const int e = tid + it * C::NT;
    w[it] = e < C::RB * K ? tw[e % K] : make_float2(0.f, 0.f);
  }
}
// X[k] = (Z_k + Z*_{M-k})/2 - i w^k (Z_k - Z*_{M-k})/2, w = e^{-2 pi i/N}, k = 0..M, to out rows
template <int N>
__device__ __forceinline__ void z_split(const float2 *lds, const float2 *w, float2 *out, int HP, size_t row0,
                                        size_t nrows, int tid) {
  using C = RowCfg<N>;
  constexpr int K = ZSplit<N>::K;
#pragma unroll
  for (int it = 0; it < ZSplit<N>::ITER; ++it) {
    const int e = tid + it * C::NT;
    if (e >= C::RB * K) break;
    const int r = e / K, k = e % K;
    const size_t row = row0 + r;
    if (row >= nrows) continue;
    const float2 zk = lds[r * C::LS + (k % C::M)];
    const float2 zm = cconj(lds[r * C::LS + ((C::M - k) % C::M)]);
    const float2 a = cadd(zk, zm), b = csub(zk, zm);
    const float2 ib = mul_i<+1>(cmul(w[it], b));
    out[row * HP + k] = make_float2(0.5f * (a.x - ib.x), 0.5f * (a.y - ib.y));
  }
}
// merge: Z[k] = (X_k + X*_{M-k}) + i w^k (X_k - X*_{M-k}), w = e^{+2 pi i/N}, from
// the rows of buf (Im of X_0 and X_M dropped) into LDS.  Every global load of
// the block's rows is issued before any is used.  IKZ: X_k -> i k X_k first.
template <int N, bool IKZ>
__device__ __forceinline__ void z_merge(float2 *lds, const float2 *buf, int HP, size_t row0, size_t nrows,
                                        const float2 *__restrict__ tw, int tid) {
  using C = RowCfg<N>;
  constexpr int HALF = C::M / 2 + 1;
  constexpr int ITER = (C::RB * HALF + C::NT - 1) / C::NT;
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
    const int kk = C::M - k;
    float2 a = xk[it], b = xm[it];
    if (IKZ) {
      a = cscale(mul_i<+1>(a), (float)k);
      b = cscale(mul_i<+1>(b), (float)kk);
    }
    float2 *lrow = lds + r * C::LS;
    if (k == 0) {
      lrow[0] = make_float2(a.x + b.x, a.x - b.x);
    } else {
      const float2 zk = cadd(cadd(a, cconj(b)), mul_i<+1>(cmul(cconj(wk[it]), csub(a, cconj(b)))));
      const float2 zkk = cadd(cadd(b, cconj(a)), mul_i<+1>(cmul(cconj(wm[it]), csub(b, cconj(a)))));
      lrow[k] = zk;
      if (kk != k) lrow[kk] = zkk;
    }
  }
}

}  // namespace fftc

#define MGPU_FFT_SIZES(X) \
  X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1536) X(2048)
