// Learned Green's-function multiplier (no counterpart in the reference; it
// fills the grid_corr slot that Forces leaves at 1, src/auxPM.c:488-498).
//
// A multilayer perceptron 4 -> 64 -> 64 -> 64 -> 1 with GELU activations maps
// (|k|/k_Nyquist, a, log10 f_R0, s) to ln M, where s is the rms of ln(1 + delta)
// over the mesh after the deposit (ln(1 + delta) floored at ln 0.01), a global
// measure of how nonlinear the density is.  The force pass multiplies the
// Green's function by M(|n|^2), tabulated once per step for every |n|^2 of the
// mesh (one network evaluation per distinct |n|^2); modes beyond the Nyquist
// radius (the corners of the cube, |k| up to sqrt(3) k_Nyquist) take the
// multiplier at |k| = k_Nyquist.
//
// Numerics: FP32 master parameters (Adam state beside them); the forward pass
// uses BF16-rounded weights; the two 64x64 layers run on MFMA
// (v_mfma_f32_16x16x16_bf16: BF16 operands, FP32 accumulation) with the hidden
// activations rounded to BF16 as their A operand; the input and output layers
// and all biases run in FP32.  The backward pass uses BF16 copies of the
// pre-activation gradients as MFMA operands for dX = dZ W and dW = dZ^T X, FP32
// everywhere else.  One wave owns 16 rows at a time; gradients are summed per
// wave in registers, written per wave and reduced in a fixed order, so a
// gradient is bit-reproducible.
#include "mgpu_internal.h"

#include <cmath>
#include <cstring>

namespace {

constexpr int H = 64;   // hidden width
constexpr int LD = 72;  // LDS row stride of a 16 x 64 BF16 tile (144 B, staggers banks)
constexpr int kWaves = 4, kThreads = 64 * kWaves;
// parameter offsets (floats): W1 [64][4], b1, W2 [64][64], b2, W3 [64][64], b3, W4 [64], b4
constexpr int oW1 = 0, oB1 = 256, oW2 = 320, oB2 = 4416, oW3 = 4480, oB3 = 8576, oW4 = 8640, oB4 = 8704;
constexpr int kParams = MGPU_MLP_NPARAMS;
static_assert(kParams == 8705, "parameter layout");
constexpr int kPart = kParams + 1;  // per-wave partial: gradient + loss

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short to_bf16(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float from_bf16(unsigned short b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ float bf(float f) { return from_bf16(to_bf16(f)); }

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float dgelu(float z) {
  return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

__device__ __forceinline__ f32x4 mma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x4 ld4(const unsigned short *p) { return *(const bf16x4 *)p; }
__device__ __forceinline__ bf16x4 gather4(const unsigned short *p, int stride) {
  bf16x4 r;
  r[0] = (short)p[0];
  r[1] = (short)p[stride];
  r[2] = (short)p[2 * stride];
  r[3] = (short)p[3 * stride];
  return r;
}

// Block LDS: weights shared by the waves, then one scratch area per wave.
struct alignas(16) Shared {
  unsigned short w2[H * LD], w3[H * LD];    // [out][in], BF16
  unsigned short w2t[H * LD], w3t[H * LD];  // [in][out], BF16 (backward)
  float w1[H * 4], b1[H], b2[H], b3[H], w4[H], b4;
};
struct alignas(16) WaveLds {
  float x[16][4];
  unsigned short h1[16 * LD], h2[16 * LD], dz[16 * LD];
  float z1[16][H], z2[16][H], z3[16][H];
  float y[16], dy[16];
};

enum { MODE_FWD = 0, MODE_TRAIN = 1, MODE_TABLE = 2 };

struct TableArgs {
  int N;           // mesh size: |k|/k_Nyq = sqrt(n2)/(N/2)
  float a, theta;  // scale factor, log10 f_R0
  const double *feature;
  float *table;    // out: M per n2
};

// D = H_in (16 x 64, BF16 rows in LDS) W^T + b, per lane 4 n-blocks x 4 rows
__device__ __forceinline__ void dense64(const unsigned short *hin, const unsigned short *w, f32x4 acc[4]) {
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; kb++)
      acc[nb] = mma(ld4(hin + r * LD + kb * 16 + 4 * q), ld4(w + (nb * 16 + r) * LD + kb * 16 + 4 * q), acc[nb]);
  }
}

template <int MODE>
__global__ __launch_bounds__(kThreads) void k_mlp(const float *__restrict__ P, const float *__restrict__ X,
                                                    const float *__restrict__ T, const float *__restrict__ Wt,
                                                    int n, float inv_wsum, float *__restrict__ Y,
                                                    float *__restrict__ part, TableArgs ta) {
  extern __shared__ __align__(16) unsigned char smem[];
  Shared &S = *(Shared *)smem;
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, r16 = l & 15, q = l >> 4;
  WaveLds &W = ((WaveLds *)(smem + sizeof(Shared)))[wave];
  // weights into LDS (BF16-rounded; the biases stay FP32)
  for (int i = threadIdx.x; i < H * H; i += kThreads) {
    const int o = i / H, c = i % H;
    const unsigned short v2 = to_bf16(P[oW2 + i]), v3 = to_bf16(P[oW3 + i]);
    S.w2[o * LD + c] = v2;
    S.w3[o * LD + c] = v3;
    if (MODE == MODE_TRAIN) {
      S.w2t[c * LD + o] = v2;
      S.w3t[c * LD + o] = v3;
    }
  }
  for (int i = threadIdx.x; i < H * 4; i += kThreads) S.w1[i] = bf(P[oW1 + i]);
  for (int i = threadIdx.x; i < H; i += kThreads) {
    S.b1[i] = P[oB1 + i];
    S.b2[i] = P[oB2 + i];
    S.b3[i] = P[oB3 + i];
    S.w4[i] = bf(P[oW4 + i]);
  }
  if (threadIdx.x == 0) S.b4 = P[oB4];
  __syncthreads();
  float s_feat = 0.f;
  if (MODE == MODE_TABLE) s_feat = (float)*ta.feature;

  // per-lane gradient accumulators (TRAIN)
  f32x4 g2[16], g3[16];
  float g1w[4][4], g1b[4], g2b[4], g3b[4], g4w[4], g4b = 0.f, loss = 0.f;
  if (MODE == MODE_TRAIN) {
#pragma unroll
    for (int t = 0; t < 16; t++) g2[t] = g3[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      g1b[nb] = g2b[nb] = g3b[nb] = g4w[nb] = 0.f;
#pragma unroll
      for (int i = 0; i < 4; i++) g1w[nb][i] = 0.f;
    }
  }

  const int tiles = (n + 15) / 16, wstride = gridDim.x * kWaves;
  for (int tile = blockIdx.x * kWaves + wave; tile < tiles; tile += wstride) {
    const int row0 = tile * 16;
    // inputs of the 16 rows (rows past n are zeros and carry no weight)
    {
      const int rr = l >> 2, c = l & 3, row = row0 + rr;
      float v = 0.f;
      if (row < n) {
        if (MODE == MODE_TABLE) {
          v = c == 0 ? fminf(sqrtf((float)row) / (0.5f * (float)ta.N), 1.f) : c == 1 ? ta.a : c == 2 ? ta.theta : s_feat;
        } else {
          v = X[(size_t)row * 4 + c];
        }
      }
      W.x[rr][c] = v;
    }
    __builtin_amdgcn_wave_barrier();
    // layer 1 (FP32): lane holds rows 4q + v, columns nb*16 + r16
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const int col = nb * 16 + r16;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = 4 * q + v;
        float z = S.b1[col];
#pragma unroll
        for (int i = 0; i < 4; i++) z += S.w1[col * 4 + i] * W.x[row][i];
        if (MODE == MODE_TRAIN) W.z1[row][col] = z;
        W.h1[row * LD + col] = to_bf16(gelu(z));
      }
    }
    __builtin_amdgcn_wave_barrier();
    // layer 2 (MFMA)
    f32x4 acc[4];
    dense64(W.h1, S.w2, acc);
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const int col = nb * 16 + r16;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = 4 * q + v;
        const float z = acc[nb][v] + S.b2[col];
        if (MODE == MODE_TRAIN) W.z2[row][col] = z;
        W.h2[row * LD + col] = to_bf16(gelu(z));
      }
    }
    __builtin_amdgcn_wave_barrier();
    // layer 3 (MFMA) and the output layer (FP32): y = W4 . gelu(z3) + b4
    dense64(W.h2, S.w3, acc);
    float yp[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const int col = nb * 16 + r16;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const float z = acc[nb][v] + S.b3[col];
        if (MODE == MODE_TRAIN) W.z3[4 * q + v][col] = z;
        yp[v] += S.w4[col] * gelu(z);
      }
    }
    // sum over the 16 lanes that share rows (fixed butterfly order)
#pragma unroll
    for (int v = 0; v < 4; v++) {
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) yp[v] += __shfl_xor(yp[v], off, 64);
    }
    if (r16 == 0) {
#pragma unroll
      for (int v = 0; v < 4; v++) W.y[4 * q + v] = yp[v] + S.b4;
    }
    __builtin_amdgcn_wave_barrier();
    if (MODE == MODE_FWD) {
      if (l < 16 && row0 + l < n) Y[row0 + l] = W.y[l];
      continue;
    }
    if (MODE == MODE_TABLE) {
      if (l < 16 && row0 + l < n) ta.table[row0 + l] = expf(W.y[l]);
      continue;
    }
    // ---- backward: dL/dy = 2 w (y - t) / sum w
    if (l < 16) {
      const int row = row0 + l;
      float d = 0.f;
      if (row < n) {
        const float e = W.y[l] - T[row], w = Wt[row];
        d = 2.f * w * e * inv_wsum;
        loss += w * e * e * inv_wsum;
      }
      W.dy[l] = d;
      g4b += d;
    }
    __builtin_amdgcn_wave_barrier();
    // dz3 = dy W4 gelu'(z3); dW4 += dy gelu(z3); db3 += dz3
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const int col = nb * 16 + r16;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = 4 * q + v;
        const float z = W.z3[row][col], dy = W.dy[row];
        g4w[nb] += dy * gelu(z);
        const float dz = dy * S.w4[col] * dgelu(z);
        g3b[nb] += dz;
        W.dz[row * LD + col] = to_bf16(dz);
      }
    }
    __builtin_amdgcn_wave_barrier();
    // dW3 += dz3^T h2: tile (jb, ib), A[j][r] = dz3[r][j], B[r][i] = h2[r][i]
    {
      bf16x4 A[4], B[4];
#pragma unroll
      for (int b = 0; b < 4; b++) {
        A[b] = gather4(W.dz + (4 * q) * LD + b * 16 + r16, LD);
        B[b] = gather4(W.h2 + (4 * q) * LD + b * 16 + r16, LD);
      }
#pragma unroll
      for (int jb = 0; jb < 4; jb++)
#pragma unroll
        for (int ib = 0; ib < 4; ib++) g3[jb * 4 + ib] = mma(A[jb], B[ib], g3[jb * 4 + ib]);
    }
    // dh2 = dz3 W3 (A = dz3 rows, B[j][i] = W3[j][i] = w3t[i][j]); dz2 = dh2 gelu'(z2)
    dense64(W.dz, S.w3t, acc);
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const int col = nb * 16 + r16;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = 4 * q + v;
        const float dz = acc[nb][v] * dgelu(W.z2[row][col]);
        g2b[nb] += dz;
        W.dz[row * LD + col] = to_bf16(dz);
      }
    }
    __builtin_amdgcn_wave_barrier();
    // dW2 += dz2^T h1
    {
      bf16x4 A[4], B[4];
#pragma unroll
      for (int b = 0; b < 4; b++) {
        A[b] = gather4(W.dz + (4 * q) * LD + b * 16 + r16, LD);
        B[b] = gather4(W.h1 + (4 * q) * LD + b * 16 + r16, LD);
      }
#pragma unroll
      for (int jb = 0; jb < 4; jb++)
#pragma unroll
        for (int ib = 0; ib < 4; ib++) g2[jb * 4 + ib] = mma(A[jb], B[ib], g2[jb * 4 + ib]);
    }
    // dh1 = dz2 W2; dz1 = dh1 gelu'(z1); dW1 += dz1 x; db1 += dz1
    dense64(W.dz, S.w2t, acc);
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const int col = nb * 16 + r16;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = 4 * q + v;
        const float dz = acc[nb][v] * dgelu(W.z1[row][col]);
        g1b[nb] += dz;
#pragma unroll
        for (int i = 0; i < 4; i++) g1w[nb][i] += dz * W.x[row][i];
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (MODE != MODE_TRAIN) return;
  // per-wave partial gradient: lanes that share a column (r16) hold partial sums
  // over rows 4q..4q+3; sum them over q in a fixed order, then one lane writes
  float *o = part + (size_t)(blockIdx.x * kWaves + wave) * kPart;
  auto qsum = [&](float v) {
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
  };
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    const int col = nb * 16 + r16;
    const float b1 = qsum(g1b[nb]), b2 = qsum(g2b[nb]), b3 = qsum(g3b[nb]), w4 = qsum(g4w[nb]);
    float w1[4];
#pragma unroll
    for (int i = 0; i < 4; i++) w1[i] = qsum(g1w[nb][i]);
    if (q == 0) {
      o[oB1 + col] = b1;
      o[oB2 + col] = b2;
      o[oB3 + col] = b3;
      o[oW4 + col] = w4;
#pragma unroll
      for (int i = 0; i < 4; i++) o[oW1 + col * 4 + i] = w1[i];
    }
  }
  // dW tiles: lane holds dW[jb*16 + 4q + v][ib*16 + r16]
#pragma unroll
  for (int jb = 0; jb < 4; jb++)
#pragma unroll
    for (int ib = 0; ib < 4; ib++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int j = jb * 16 + 4 * q + v, i = ib * 16 + r16;
        o[oW2 + j * H + i] = g2[jb * 4 + ib][v];
        o[oW3 + j * H + i] = g3[jb * 4 + ib][v];
      }
  // db4 and the loss: lanes 0..15 hold row partials
#pragma unroll
  for (int off = 1; off < 16; off <<= 1) {
    g4b += __shfl_xor(g4b, off, 64);
    loss += __shfl_xor(loss, off, 64);
  }
  if (l == 0) {
    o[oB4] = g4b;
    o[kParams] = loss;
  }
}

// gradient (and loss) = sum of the per-wave partials in wave order
__global__ void k_mlp_reduce(const float *__restrict__ part, int nparts, float *__restrict__ grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kPart) return;
  float s = 0.f;
  for (int p = 0; p < nparts; p++) s += part[(size_t)p * kPart + i];
  grad[i] = s;
}

__global__ void k_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, float lr, float b1t, float b2t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / (1.f - b1t)) / (sqrtf(vi / (1.f - b2t)) + eps);
}

// density feature: partial sums of ln^2(max(1 + delta, 0.01)) over rows of the real mesh
constexpr int kFeatBlocks = 1024;
__global__ __launch_bounds__(256) void k_feature(const float *__restrict__ d, int N, size_t rs,
                                                 double *__restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  const size_t rows = (size_t)N * N;
  for (size_t row = blockIdx.x; row < rows; row += gridDim.x) {
    const float *p = d + row * rs;
    for (int i = threadIdx.x; i < N; i += 256) {
      const float v = logf(fmaxf(1.f + p[i], 0.01f));
      s += (double)(v * v);
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void k_feature_final(const double *__restrict__ part, int n, double cells, double *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < n; i++) s += part[i];
  *out = sqrt(s / cells);
}

constexpr size_t kMlpLds = sizeof(Shared) + kWaves * sizeof(WaveLds);
static_assert(kMlpLds <= 160 * 1024, "MLP LDS");

template <int MODE>
int mlp_launch(hipStream_t st, int blocks, const float *P, const float *X, const float *T, const float *Wt, int n,
               float inv_wsum, float *Y, float *part, TableArgs ta) {
  const size_t lds = kMlpLds;
  if (hipFuncSetAttribute((const void *)k_mlp<MODE>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
      hipSuccess)
    return MGPU_E_HIP;
  hipLaunchKernelGGL(k_mlp<MODE>, dim3(blocks), dim3(kThreads), lds, st, P, X, T, Wt, n, inv_wsum, Y, part, ta);
  return hipGetLastError() == hipSuccess ? MGPU_OK : MGPU_E_HIP;
}

int tiles_to_blocks(int n, int cap) {
  const int tiles = (n + 15) / 16;
  int b = (tiles + kWaves - 1) / kWaves;
  return b < 1 ? 1 : (b > cap ? cap : b);
}

// deterministic initial parameters: splitmix64 + Box-Muller, W ~ N(0, 1/fan_in),
// biases 0, output layer 0 (M = 1 until trained)
void init_params(unsigned long long seed, std::vector<float> &p) {
  p.assign(kParams, 0.f);
  unsigned long long s = seed * 0x9E3779B97F4A7C15ull + 1;
  auto next = [&]() {
    unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
  };
  auto normal = [&]() {
    const double u1 = next() + 1e-300, u2 = next();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * MGPU_PI * u2);
  };
  for (int i = 0; i < 256; i++) p[oW1 + i] = (float)(normal() / std::sqrt(4.0));
  for (int i = 0; i < H * H; i++) p[oW2 + i] = (float)(normal() / std::sqrt((double)H));
  for (int i = 0; i < H * H; i++) p[oW3 + i] = (float)(normal() / std::sqrt((double)H));
}

}  // namespace

struct mgpu_mlp {
  float *p = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *part = nullptr;
  int part_blocks = 0;
  int t = 0;  // Adam steps taken
  hipStream_t stream = nullptr;
};

static int mlp_fail(int code, const char *msg) { return mgpu_fail(nullptr, code, msg); }

int mlp_create(unsigned long long seed, mgpu_mlp **out) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return mlp_fail(MGPU_E_NODEVICE, "no HIP device visible");
  }
  mgpu_mlp *m = new mgpu_mlp();
  m->part_blocks = 64;
  bool ok = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc(&m->p, kParams * sizeof(float)) == hipSuccess &&
            hipMalloc(&m->m, kParams * sizeof(float)) == hipSuccess &&
            hipMalloc(&m->v, kParams * sizeof(float)) == hipSuccess &&
            hipMalloc(&m->grad, kPart * sizeof(float)) == hipSuccess &&
            hipMalloc(&m->part, (size_t)m->part_blocks * kWaves * kPart * sizeof(float)) == hipSuccess;
  std::vector<float> p;
  init_params(seed, p);
  ok = ok && hipMemcpy(m->p, p.data(), kParams * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemset(m->m, 0, kParams * sizeof(float)) == hipSuccess &&
       hipMemset(m->v, 0, kParams * sizeof(float)) == hipSuccess;
  if (!ok) {
    mlp_destroy(m);
    return mlp_fail(MGPU_E_NOMEM, "MLP allocation failed");
  }
  *out = m;
  return MGPU_OK;
}

void mlp_destroy(mgpu_mlp *m) {
  if (!m) return;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (float *q : {m->p, m->m, m->v, m->grad, m->part})
    if (q) (void)hipFree(q);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

int mlp_params(mgpu_mlp *m, float *get, const float *set) {
  if (hipStreamSynchronize(m->stream) != hipSuccess) return mlp_fail(MGPU_E_HIP, "MLP stream failed");
  if (set) {
    if (hipMemcpy(m->p, set, kParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
      return mlp_fail(MGPU_E_HIP, "MLP parameter upload failed");
    m->t = 0;
    if (hipMemset(m->m, 0, kParams * sizeof(float)) != hipSuccess ||
        hipMemset(m->v, 0, kParams * sizeof(float)) != hipSuccess)
      return mlp_fail(MGPU_E_HIP, "MLP state reset failed");
  }
  if (get && hipMemcpy(get, m->p, kParams * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return mlp_fail(MGPU_E_HIP, "MLP parameter download failed");
  return MGPU_OK;
}

namespace {
// device copies of a sample set, freed on scope exit
struct Samples {
  float *x = nullptr, *t = nullptr, *w = nullptr, *y = nullptr;
  ~Samples() {
    for (float *q : {x, t, w, y})
      if (q) (void)hipFree(q);
  }
  int upload(const float *hx, const float *ht, const float *hw, int n, bool want_y) {
    if (hipMalloc(&x, (size_t)n * 4 * sizeof(float)) != hipSuccess) return MGPU_E_NOMEM;
    if (hipMemcpy(x, hx, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return MGPU_E_HIP;
    if (ht) {
      if (hipMalloc(&t, (size_t)n * sizeof(float)) != hipSuccess || hipMalloc(&w, (size_t)n * sizeof(float)) != hipSuccess)
        return MGPU_E_NOMEM;
      if (hipMemcpy(t, ht, (size_t)n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(w, hw, (size_t)n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return MGPU_E_HIP;
    }
    if (want_y && hipMalloc(&y, (size_t)n * sizeof(float)) != hipSuccess) return MGPU_E_NOMEM;
    return MGPU_OK;
  }
};

int grad_on_device(mgpu_mlp *m, const Samples &s, int n, float inv_wsum) {
  const int blocks = tiles_to_blocks(n, m->part_blocks);
  int rc = mlp_launch<MODE_TRAIN>(m->stream, blocks, m->p, s.x, s.t, s.w, n, inv_wsum, nullptr, m->part, TableArgs{});
  if (rc) return rc;
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kPart + 255) / 256), dim3(256), 0, m->stream, (const float *)m->part,
                     blocks * kWaves, m->grad);
  return hipGetLastError() == hipSuccess ? MGPU_OK : MGPU_E_HIP;
}
}  // namespace

int mlp_forward(mgpu_mlp *m, const float *x, int n, float *y) {
  Samples s;
  int rc = s.upload(x, nullptr, nullptr, n, true);
  if (!rc) rc = mlp_launch<MODE_FWD>(m->stream, tiles_to_blocks(n, 4096), m->p, s.x, nullptr, nullptr, n, 0.f, s.y,
                                     nullptr, TableArgs{});
  if (!rc && (hipStreamSynchronize(m->stream) != hipSuccess ||
              hipMemcpy(y, s.y, (size_t)n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess))
    rc = MGPU_E_HIP;
  return rc ? mlp_fail(rc, "MLP forward failed") : MGPU_OK;
}

static double wsum_inv(const float *w, int n) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += w[i];
  return s > 0 ? 1.0 / s : 0.0;
}

int mlp_gradient(mgpu_mlp *m, const float *x, const float *t, const float *w, int n, float *grad, double *loss) {
  Samples s;
  int rc = s.upload(x, t, w, n, false);
  if (!rc) rc = grad_on_device(m, s, n, (float)wsum_inv(w, n));
  std::vector<float> g(kPart);
  if (!rc && (hipStreamSynchronize(m->stream) != hipSuccess ||
              hipMemcpy(g.data(), m->grad, kPart * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess))
    rc = MGPU_E_HIP;
  if (rc) return mlp_fail(rc, "MLP gradient failed");
  if (grad) std::memcpy(grad, g.data(), kParams * sizeof(float));
  if (loss) *loss = g[kParams];
  return MGPU_OK;
}

int mlp_train(mgpu_mlp *m, const float *x, const float *t, const float *w, int n, int iters, float lr,
              double *loss) {
  Samples s;
  int rc = s.upload(x, t, w, n, false);
  const float iw = (float)wsum_inv(w, n);
  for (int it = 0; it < iters && !rc; it++) {
    rc = grad_on_device(m, s, n, iw);
    if (rc) break;
    m->t++;
    hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, m->stream, m->p, m->m, m->v,
                       (const float *)m->grad, lr, (float)std::pow(0.9, m->t), (float)std::pow(0.999, m->t));
    if (hipGetLastError() != hipSuccess) rc = MGPU_E_HIP;
  }
  float l = 0.f;
  if (!rc && (hipStreamSynchronize(m->stream) != hipSuccess ||
              hipMemcpy(&l, m->grad + kParams, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess))
    rc = MGPU_E_HIP;
  if (rc) return mlp_fail(rc, "MLP training failed");
  if (loss) *loss = l;  // of the last gradient evaluation (before its update)
  return MGPU_OK;
}

// After a deposit: the density feature of delta(x) and, with a network set, the
// multiplier table M(n^2) that the force pass reads (ctx->green_mult).
int learned_refresh(mgpu_ctx *ctx, const float *delta_x) {
  const int N = ctx->N;
  hipLaunchKernelGGL(k_feature, dim3(kFeatBlocks), dim3(256), 0, ctx->stream, delta_x, N, 2 * (size_t)ctx->HP,
                     ctx->feat_part);
  hipLaunchKernelGGL(k_feature_final, dim3(1), dim3(64), 0, ctx->stream, (const double *)ctx->feat_part, kFeatBlocks,
                     (double)N * N * N, ctx->feat);
  MGPU_HIP(ctx, hipGetLastError());
  if (!ctx->mlp) return MGPU_OK;
  TableArgs ta{N, (float)ctx->learn_a, (float)ctx->learn_theta, ctx->feat, ctx->green_mult};
  const int n = (int)ctx->n2len;
  if (mlp_launch<MODE_TABLE>(ctx->stream, tiles_to_blocks(n, 2048), ctx->mlp->p, nullptr, nullptr, nullptr, n, 0.f,
                             nullptr, nullptr, ta))
    return mgpu_fail(ctx, MGPU_E_HIP, "MLP table launch failed");
  return MGPU_OK;
}

size_t learned_feature_blocks() { return kFeatBlocks; }
