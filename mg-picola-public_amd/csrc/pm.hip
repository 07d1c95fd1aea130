// Particle-mesh kernels of the COLA step for gfx950: CIC deposit, P(k)
// binning, the modified-gravity spectral and real-space passes, the Green's
// function, the CIC gather, Kick and Drift.  Each kernel cites the reference
// routine (HAWinther/MG-PICOLA-PUBLIC) whose arithmetic it carries out.
#include "mgpu_internal.h"
#include "cic.h"

#include <cmath>
#include <vector>

namespace {

constexpr int kBlock = 256;

// ------------------------------------------------------------------ P(k)
// compute_power_spectrum (src/compute_pofk.c:71-236): every mode of the half
// spectrum with weight 1 (kz = 0, N/2) or 2, |n| binned linearly or
// logarithmically, CIC deconvolution 1/(W_x W_y W_z)^4 and 1/N^6.
// Only bins [b0, b0 + nb) are accumulated (in LDS), so a bin count beyond
// the LDS takes several launches over bin ranges.
__global__ __launch_bounds__(kBlock) void k_pofk(const float2 *__restrict__ dk, int N, int HP,
                                                 const double *__restrict__ winv, const int16_t *__restrict__ pbin,
                                                 int nbins, int bintype, double kmin, double kmax, int b0, int nb,
                                                 double *__restrict__ out) {
  extern __shared__ double sh[];  // 3*nb
  for (int b = threadIdx.x; b < 3 * nb; b += blockDim.x) sh[b] = 0.0;
  __syncthreads();
  const int H = N / 2 + 1;
  const size_t total = (size_t)N * N * H;
  const double lnrange = bintype == 1 ? log(kmax / kmin) : 1.0;
  for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < total; m += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(m % H);
    const size_t rowi = m / H;
    const int j = (int)(rowi % N), i = (int)(rowi / N);
    const int d0 = i > N / 2 ? N - i : i, d1 = j > N / 2 ? N - j : j, d2 = k;
    const int n2 = d0 * d0 + d1 * d1 + d2 * d2;
    const double kmag = sqrt((double)n2);
    int nk;
    if (pbin) nk = pbin[n2];  // the same formula, evaluated by the host
    else if (bintype == 0) nk = (int)((kmag - kmin) / (kmax - kmin) * nbins + 0.5);
    else nk = kmag <= 0.0 ? -1 : (int)(log(kmag / kmin) / lnrange * nbins + 0.5);
    if (nk < 0 || nk >= nbins || nk < b0 || nk >= b0 + nb) continue;
    const double w = (k == 0 || k == N / 2) ? 1.0 : 2.0;
    const float2 v = dk[rowi * HP + k];
    const double p = ((double)v.x * v.x + (double)v.y * v.y) * winv[d0] * winv[d1] * winv[d2];
    atomicAdd(&sh[nk - b0], w * p);
    atomicAdd(&sh[nb + nk - b0], w * kmag);
    atomicAdd(&sh[2 * nb + nk - b0], w);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < 3 * nb; b += blockDim.x)
    if (sh[b] != 0.0) atomicAdd(&out[(b / nb) * nbins + b0 + b % nb], sh[b]);
}

// ------------------------------------------------------------------ RSD multipoles
// PtoMesh_RSD's positions (src/compute_pofk.c:279-335): mesh coordinates
// (Pos0, Pos1, Pos2) for the z line of sight and (Pos0, Pos2, Pos1) for y; the
// line-of-sight one moved by vnorm V, V = Vel (+ the LPT velocity with COLA:
// the dD/dy field summed over orders as a float sum when scale-dependent, else
// D dDdy(A) + D2 dD2dy(A)), then one periodic wrap.  Record: x, y in box units,
// Z in mesh units as FP64 (low, high words), the layout the deposit reads.
__global__ __launch_bounds__(kBlock) void k_rsd_coords(const float *__restrict__ pos, const float *__restrict__ vel,
                                                       const float *__restrict__ l0, const float *__restrict__ l1,
                                                       size_t np, int axis, double vnorm, double scale, int sd,
                                                       int use_cola, double s1, double s2, int N,
                                                       float4 *__restrict__ crd) {
#pragma clang fp contract(off)  // the rounding sequence of the CPU reference
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int ya = axis == 1 ? 2 : 1;
  const size_t ia = (size_t)axis * np + i;
  const float va = vel[ia];
  double V;
  if (!use_cola) V = va;
  else if (sd) V = (double)(va + l1[ia]);
  else V = (double)va + ((double)l0[ia] * s1 + (double)l1[ia] * s2);
  V *= vnorm;
  double Z = (double)pos[ia] * scale;
  Z += V;
  if (Z >= (double)N) Z -= (double)N;
  if (Z < 0.0) Z += (double)N;
  crd[i] = make_float4(pos[i], pos[(size_t)ya * np + i], __int_as_float(__double2loint(Z)),
                       __int_as_float(__double2hiint(Z)));
}

// bin_up_RSD_power_spectrum's sums (src/compute_pofk.c:496-704): per bin the
// weighted |delta|^2 (deconvolved, 1/N^6), its mu^2 and mu^4 moments with
// mu^2 = n_z^2/n^2, and the mode count; weights 1 (kz = 0, N/2) or 2.
__global__ __launch_bounds__(kBlock) void k_rsd_bins(const float2 *__restrict__ dk, int N, int HP,
                                                     const double *__restrict__ winv,
                                                     const int16_t *__restrict__ pbin, int nbins,
                                                     double *__restrict__ out) {
  extern __shared__ double sh[];  // 4*nbins
  for (int b = threadIdx.x; b < 4 * nbins; b += blockDim.x) sh[b] = 0.0;
  __syncthreads();
  const int H = N / 2 + 1;
  const size_t total = (size_t)N * N * H;
  for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < total; m += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(m % H);
    const size_t rowi = m / H;
    const int j = (int)(rowi % N), i = (int)(rowi / N);
    const int d0 = i > N / 2 ? N - i : i, d1 = j > N / 2 ? N - j : j;
    const int n2 = d0 * d0 + d1 * d1 + k * k;
    const int nk = pbin[n2];
    if (nk < 0) continue;
    const double kmag = sqrt((double)n2);
    const double mu2 = k == 0 ? 0.0 : (double)(k * k) / kmag / kmag;
    const double w = (k == 0 || k == N / 2) ? 1.0 : 2.0;
    const float2 v = dk[rowi * HP + k];
    const double p = ((double)v.x * v.x + (double)v.y * v.y) * winv[d0] * winv[d1] * winv[k];
    atomicAdd(&sh[nk], w * p);
    atomicAdd(&sh[nbins + nk], w * p * mu2);
    atomicAdd(&sh[2 * nbins + nk], w * p * mu2 * mu2);
    atomicAdd(&sh[3 * nbins + nk], w);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < 4 * nbins; b += blockDim.x)
    if (sh[b] != 0.0) atomicAdd(&out[b], sh[b]);
}

// ------------------------------------------------------------------ spectral helpers
// |n|^2 of a half-spectrum entry (identical under the Forces, mg.h and 2LPT.c
// index conventions)
__device__ __forceinline__ int n2_of(int i, int j, int k, int N) {
  const int a = i > N / 2 ? N - i : i, b = j > N / 2 ? N - j : j;
  return a * a + b * b + k * k;
}

// EffDensitykToPhiofk (src/mg.h:74-116): out = coupling * n^2/(n^2 + massterm2) * in, out(0) = 0
__global__ __launch_bounds__(kBlock) void k_eff_to_phi(const float2 *__restrict__ in, float2 *__restrict__ out, int N,
                                                       int HP, double coupling, double massterm2) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)N * N * HP) return;
  const int k = (int)(e % HP);
  const size_t row = e / HP;
  const int j = (int)(row % N), i = (int)(row / N);
  float2 r = make_float2(0.f, 0.f);
  if (k <= N / 2) {
    const int n2 = n2_of(i, j, k, N);
    if (n2 > 0) {
      const double RK = (double)n2;
      const double f = coupling * (RK / (RK + massterm2));
      const float2 v = in[e];
      r = make_float2((float)(v.x * f), (float)(v.y * f));
    }
  }
  out[e] = r;
}

// out = in * s over the whole slot (DGP without screening, src/mg.h:212-217;
// JBD Geff(a), src/mg.h:127-137)
__global__ __launch_bounds__(kBlock) void k_scale(const float2 *__restrict__ in, float2 *__restrict__ out, size_t n,
                                                  double s) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float2 v = in[e];
  out[e] = make_float2((float)(v.x * s), (float)(v.y * s));
}

// ------------------------------------------------------------------ gather
// Sums of 3 per-particle values in a fixed order: each block takes kChunk
// consecutive particles, each thread kRed of them at stride kBlock (in order),
// then a shuffle tree per wave and the waves in order.  The partials
// [block][3] are reduced by k_reduce3, so the means do not depend on launch
// order or timing.
constexpr int kRed = 8, kChunk = kBlock * kRed;

__device__ __forceinline__ void block_sum3(double s[3], double *__restrict__ partial) {
  __shared__ double red[3][kBlock / 64];
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
#pragma unroll
  for (int a = 0; a < 3; a++) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s[a] += __shfl_down(s[a], o);
    if (lane == 0) red[a][w] = s[a];
  }
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int a = 0; a < 3; a++) {
      double t = red[a][0];
      for (int q = 1; q < kBlock / 64; q++) t += red[a][q];
      partial[(size_t)blockIdx.x * 3 + a] = t;
    }
}

// MtoParticles (src/auxPM.c:560-644): CIC interpolation of the three force
// meshes (WPAR = 1), Disp stored as float.  This direct form serves when the
// deposit's bins do not describe the current positions; otherwise the gather
// goes by tiles (gather_tiled).  Either way k_sum3 forms the sums for the mean.
__global__ __launch_bounds__(kBlock) void k_gather(const float *__restrict__ pos, size_t np, int N, int HP2,
                                                   double scale, const float *__restrict__ g0,
                                                   const float *__restrict__ g1, const float *__restrict__ g2,
                                                   float *__restrict__ disp) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  unsigned IX, IY, IZ;
  double DX, DY, DZ;
  cell_of((double)pos[i] * scale, N, IX, DX);
  cell_of((double)pos[np + i] * scale, N, IY, DY);
  cell_of((double)pos[2 * np + i] * scale, N, IZ, DZ);
  const double TX = 1.0 - DX, TY = 1.0 - DY, TZ = 1.0 - DZ;
  const unsigned IXn = IX + 1 >= (unsigned)N ? 0 : IX + 1;
  const unsigned IYn = IY + 1 >= (unsigned)N ? 0 : IY + 1;
  const unsigned IZn = IZ + 1 >= (unsigned)N ? 0 : IZ + 1;
  const size_t r00 = ((size_t)IX * N + IY) * HP2, r01 = ((size_t)IX * N + IYn) * HP2;
  const size_t r10 = ((size_t)IXn * N + IY) * HP2, r11 = ((size_t)IXn * N + IYn) * HP2;
  const double w000 = TX * TY * TZ, w001 = TX * TY * DZ, w010 = TX * DY * TZ, w011 = TX * DY * DZ;
  const double w100 = DX * TY * TZ, w101 = DX * TY * DZ, w110 = DX * DY * TZ, w111 = DX * DY * DZ;
  const float *g[3] = {g0, g1, g2};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float *m = g[a];
    const double d = m[r00 + IZ] * w000 + m[r00 + IZn] * w001 + m[r01 + IZ] * w010 + m[r01 + IZn] * w011 +
                     m[r10 + IZ] * w100 + m[r10 + IZn] * w101 + m[r11 + IZ] * w110 + m[r11 + IZn] * w111;
    disp[(size_t)a * np + i] = (float)d;
  }
}

// sums of Disp for the mean, in particle order (see block_sum3)
__global__ __launch_bounds__(kBlock) void k_sum3(const float *__restrict__ disp, size_t np,
                                                 double *__restrict__ partial, size_t zs) {
  disp = zbuf(disp, zs);
  partial = zbuf(partial, zs);
  const size_t base = (size_t)blockIdx.x * kChunk + threadIdx.x;
  double s[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float v[kRed];
#pragma unroll
    for (int r = 0; r < kRed; r++) {
      const size_t i = base + (size_t)r * kBlock;
      v[r] = i < np ? disp[(size_t)a * np + i] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < kRed; r++) s[a] += (double)v[r];
  }
  block_sum3(s, partial);
}

// fixed-order reduction of per-block partials [nb][3] -> out[3]
__global__ __launch_bounds__(1024) void k_reduce3(const double *__restrict__ partial, int nb, double *__restrict__ out,
                                                  size_t zs) {
  partial = zbuf(partial, zs);
  out = zbuf(out, zs);
  __shared__ double red[3][1024];
  double s[3] = {0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    for (int a = 0; a < 3; a++) s[a] += partial[(size_t)b * 3 + a];
  for (int a = 0; a < 3; a++) red[a][threadIdx.x] = s[a];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int a = 0; a < 3; a++) red[a][threadIdx.x] += red[a][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int a = 0; a < 3; a++) out[a] = red[a][0];
}

// ------------------------------------------------------------------ Kick / Drift
// COLA term (src/main.c:712, 728, 767, 781): D*s1 + D2*s2 in double, or the
// float sum of the stored scale-dependent fields.
__device__ __forceinline__ double lpt_term(const float *__restrict__ l0, const float *__restrict__ l1, double s1,
                                           double s2, int sd, size_t idx) {
#pragma clang fp contract(off)
  if (sd == 1) return (double)l0[idx];
  if (sd == 2) return (double)l1[idx];
  return (double)l0[idx] * s1 + (double)l1[idx] * s2;
}

// Kick (src/main.c:688-742)
__global__ __launch_bounds__(kBlock) void k_kick(float *__restrict__ vel, float *__restrict__ disp,
                                                 const float *__restrict__ l0, const float *__restrict__ l1,
                                                 size_t np, double s1, double s2, int sd, double omega, double dda,
                                                 double A, int use_cola, double sD0, double sD1, double sD2,
                                                 double *__restrict__ partial) {
#pragma clang fp contract(off)  // same rounding sequence as the CPU reference
  const size_t base = (size_t)blockIdx.x * kChunk + threadIdx.x;
  double s[3] = {0, 0, 0};
  const double sD[3] = {sD0, sD1, sD2};
#pragma unroll 2
  for (int r = 0; r < kRed; r++) {
    const size_t i = base + (size_t)r * kBlock;
    if (i >= np) break;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const size_t idx = (size_t)a * np + i;
      const float dv = (float)((double)disp[idx] - sD[a]);
      disp[idx] = dv;
      const double lp = use_cola ? lpt_term(l0, l1, s1, s2, sd, idx) : 0.0;
      const double force = -1.5 * omega * (double)dv - use_cola * lp / A;
      const float v = (float)((double)vel[idx] + force * dda);
      vel[idx] = v;
      s[a] += (double)v;
    }
  }
  block_sum3(s, partial);
}

// periodic_wrap on float (src/auxPM.c:649-655) with a bounded loop
__device__ __forceinline__ float wrapf(float x, float box) {
  for (int it = 0; it < 4 && x >= box; ++it) x -= box;
  for (int it = 0; it < 4 && x < 0.f; ++it) x += box;
  if (!(x >= 0.f && x < box)) {
    x = isfinite(x) ? fmodf(x, box) : 0.f;
    if (x < 0.f) x += box;
    if (!(x >= 0.f && x < box)) x = 0.f;
  }
  if (x == box) x = 0.f;
  return x;
}

// Drift (src/main.c:747-787)
__global__ __launch_bounds__(kBlock) void k_drift(float *__restrict__ pos, const float *__restrict__ vel,
                                                  const float *__restrict__ l0, const float *__restrict__ l1,
                                                  size_t np, double s1, double s2, int sd, float box, double dyyy,
                                                  int use_cola, double sV0, double sV1, double sV2) {
#pragma clang fp contract(off)  // same rounding sequence as the CPU reference
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const double sV[3] = {sV0, sV1, sV2};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const size_t idx = (size_t)a * np + i;
    float p = (float)((double)pos[idx] + ((double)vel[idx] - sV[a]) * dyyy);
    const double lp = use_cola ? lpt_term(l0, l1, s1, s2, sd, idx) : 0.0;
    pos[idx] = wrapf((float)((double)p + use_cola * lp), box);
  }
}

// Kick followed by Drift in one pass over the particles (src/main.c:688-787 on
// the steps with no output between them).  The drift's mean velocity is the
// mean before the kick (sV): the kick term has zero mean (Disp - sumD, and the
// LPT fields, whose k = 0 mode is zero), so it is the mean after the kick up to
// rounding.  Disp is not written back (only a second kick at an output step
// needs Disp - sumD, and that takes the separate passes).
__global__ __launch_bounds__(kBlock) void k_kick_drift(float *__restrict__ vel, float *__restrict__ pos,
                                                       const float *__restrict__ disp, const float *__restrict__ l0,
                                                       const float *__restrict__ l1, size_t np, int sdk, double s1k,
                                                       double s2k, int sdd, double s1d, double s2d, double omega,
                                                       double dda, double A, double dyyy, float box, int use_cola,
                                                       double sD0, double sD1, double sD2, double sV0, double sV1,
                                                       double sV2, const double *__restrict__ dsum,
                                                       double *__restrict__ partial, size_t zs) {
#pragma clang fp contract(off)  // same rounding sequence as the separate passes
  vel = zbuf(vel, zs);
  pos = zbuf(pos, zs);
  disp = zbuf(disp, zs);
  l0 = zbuf(l0, zs);
  l1 = zbuf(l1, zs);
  dsum = zbuf(dsum, zs);
  partial = zbuf(partial, zs);
  const size_t base = (size_t)blockIdx.x * kChunk + threadIdx.x;
  double s[3] = {0, 0, 0};
  double sD[3] = {sD0, sD1, sD2}, sV[3] = {sV0, sV1, sV2};
  if (dsum) {
    // the means from the device-resident sums (Disp in dsum[0..2], the last kick's velocities in
    // dsum[3..5]), divided as the host divides them
#pragma unroll
    for (int a = 0; a < 3; a++) {
      sD[a] = dsum[a] / (double)np;
      sV[a] = dsum[3 + a] / (double)np;
    }
  }
#pragma unroll 2
  for (int r = 0; r < kRed; r++) {
    const size_t i = base + (size_t)r * kBlock;
    if (i >= np) break;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const size_t idx = (size_t)a * np + i;
      const float dv = (float)((double)disp[idx] - sD[a]);
      const double lk = use_cola ? lpt_term(l0, l1, s1k, s2k, sdk, idx) : 0.0;
      const double force = -1.5 * omega * (double)dv - use_cola * lk / A;
      const float v = (float)((double)vel[idx] + force * dda);
      vel[idx] = v;
      s[a] += (double)v;
      const float p = (float)((double)pos[idx] + ((double)v - sV[a]) * dyyy);
      const double ld = use_cola ? lpt_term(l0, l1, s1d, s2d, sdd, idx) : 0.0;
      pos[idx] = wrapf((float)((double)p + use_cola * ld), box);
    }
  }
  block_sum3(s, partial);
}

}  // namespace

// =================================================================== host side

// the per-block partials reduced into dst (3 doubles on the device: ctx->dsum
// for Disp, ctx->dsum + 3 for the velocities), and copied to out unless out is
// nullptr (then nothing waits for the device)
int pm_reduce3(mgpu_ctx *ctx, int nblocks, double out[3], double *dst) {
  hipLaunchKernelGGL(k_reduce3, dim3(1, 1, ctx->bz), dim3(1024), 0, ctx->stream, ctx->partial, nblocks, dst, ctx->zs);
  MGPU_HIP(ctx, hipGetLastError());
  if (!out) return MGPU_OK;
  MGPU_HIP(ctx, hipMemcpyAsync(out, dst, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MGPU_OK;
}

int pm_deposit(mgpu_ctx *ctx) {
  PhaseTimer pt(ctx, &ctx->t.ptomesh);
  const bool need_dx = ctx->p.mg_active && ctx->p.model != MGPU_BRANSDICKE && ctx->p.model != MGPU_LCDM;
  // With MG screening delta(x) goes to slot 1, where it stays for the screening
  // pass, and the r2c writes delta(k) to slot 0.
  ctx->dx_slot = need_dx ? 1 : 0;
  int rc = deposit_tiled(ctx, (float *)ctx->slot[ctx->dx_slot]);
  if (!rc && ctx->learn_on) rc = learned_refresh(ctx, (const float *)ctx->slot[ctx->dx_slot]);
  return rc;
}

namespace {
// the bin sums are cleared; the per-axis deconvolution table is ctx->winv
int pofk_setup(mgpu_ctx *ctx, double **dwin) {
  const int nbins = ctx->p.pofk_nbins;
  if (nbins <= 0) return mgpu_fail(ctx, MGPU_E_INVALID, "pofk_nbins must be positive");
  *dwin = ctx->winv;
  MGPU_HIP(ctx, hipMemsetAsync(ctx->pk, 0, 3 * nbins * sizeof(double), ctx->stream));
  return MGPU_OK;
}

// normalisation of the bin sums (src/compute_pofk.c:200-236); with geometry the
// |n| sums and mode counts come from ctx->pk_geom and only the power from the device
int pofk_finish(mgpu_ctx *ctx, mgpu_pofk *out, bool geometry_from_ctx) {
  const int nbins = ctx->p.pofk_nbins;
  std::vector<double> h(3 * nbins);
  MGPU_HIP(ctx, hipMemcpyAsync(h.data(), ctx->pk, (geometry_from_ctx ? 1 : 3) * nbins * sizeof(double),
                               hipMemcpyDeviceToHost, ctx->stream));
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (geometry_from_ctx) std::copy(ctx->pk_geom.begin(), ctx->pk_geom.end(), h.begin() + nbins);
  else ctx->pk_geom.assign(h.begin() + nbins, h.end());
  const double box3 = std::pow(ctx->p.box, 3);
  const double shot = std::pow(ctx->p.box / (double)ctx->Np, 3);
  for (int b = 0; b < nbins; b++) {
    double p = h[b], k = h[nbins + b], n = h[2 * nbins + b];
    if (n > 0) {
      p = (p / n) * box3;
      if (ctx->p.pofk_subtract_shotnoise) p -= shot;
      k = (k / n) * 2.0 * M_PI / ctx->p.box;
    }
    if (out->pofk) out->pofk[b] = p;
    if (out->kmean) out->kmean[b] = k;
    if (out->nmodes) out->nmodes[b] = n;
  }
  return MGPU_OK;
}
}  // namespace

// compute_power_spectrum of the delta(k) in dk: power, |n| sums and mode counts
int pm_pofk(mgpu_ctx *ctx, const float2 *dk, mgpu_pofk *out) {
  PhaseTimer pt(ctx, &ctx->t.pofk);
  const int N = ctx->N, nbins = ctx->p.pofk_nbins;
  double *dwin;
  int rc;
  if ((rc = pofk_setup(ctx, &dwin))) return rc;
  // bins per launch: as many as the LDS holds (three doubles each)
  const int chunk = std::max(1, std::min(nbins, (int)(ctx->lds_max / (3 * sizeof(double)))));
  const size_t shmem = 3 * (size_t)chunk * sizeof(double);
  if (shmem > 65536) {
    MGPU_HIP(ctx, hipFuncSetAttribute((const void *)k_pofk, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  }
  const size_t modes = (size_t)N * N * (N / 2 + 1);
  unsigned grid = std::min<size_t>(cdiv(modes, kBlock), 4096);
  for (int b0 = 0; b0 < nbins; b0 += chunk) {
    const int nb = std::min(chunk, nbins - b0);
    KScope ks(ctx, MGPU_K_POFK, 8.0 * modes);
    hipLaunchKernelGGL(k_pofk, dim3(grid), dim3(kBlock), 3 * (size_t)nb * sizeof(double), ctx->stream, dk, N, ctx->HP,
                       dwin, ctx->pk_bin, nbins, ctx->p.pofk_bintype, ctx->p.pofk_kmin, ctx->p.pofk_kmax, b0, nb,
                       ctx->pk);
    MGPU_HIP(ctx, hipGetLastError());
  }
  return pofk_finish(ctx, out, false);
}

// The x pass that bins P(k) holds its tile of T columns and the bins in LDS.
static bool pofk_fused_fits(const mgpu_ctx *ctx, int T) {
  const size_t need = (size_t)ctx->N * T * sizeof(float2) + sizeof(double) * (size_t)ctx->p.pofk_nbins +
                      sizeof(float) * (size_t)(ctx->N / 2 + 1);
  return need <= ctx->lds_max;
}

// PtoMesh's r2c of delta(x) (slot dx_slot) into slot 0, with P(k) when out != NULL.
// Once the bin geometry is known the power is binned inside the last (x) pass,
// when the bins fit that pass's LDS; otherwise by k_pofk after the transform.
int pm_transform(mgpu_ctx *ctx, mgpu_pofk *out) {
  float2 *dk = ctx->slot[0];
  const float *dx = (const float *)ctx->slot[ctx->dx_slot];
  int rc;
  if (!out || !ctx->pk_bin || ctx->pk_geom.size() != 2 * (size_t)ctx->p.pofk_nbins ||
      !pofk_fused_fits(ctx, ctx->fft.tcols)) {
    {
      PhaseTimer pt(ctx, &ctx->t.ptomesh);
      if ((rc = fft_r2c_3d(ctx, &ctx->fft, dk, dx))) return rc;
    }
    return out ? pm_pofk(ctx, dk, out) : MGPU_OK;
  }
  double *dwin;
  if ((rc = pofk_setup(ctx, &dwin))) return rc;
  {
    PhaseTimer pt(ctx, &ctx->t.ptomesh);
    PhaseTimer pf(ctx, &ctx->t.fft);
    if ((rc = fft_pass_z(ctx, &ctx->fft, dk, 1, nullptr, dx))) return rc;
    if ((rc = fft_pass_col(ctx, &ctx->fft, dk, 1, 1))) return rc;
    if ((rc = fused_pofk_x(ctx, dk, ctx->pk))) return rc;
  }
  PhaseTimer pt(ctx, &ctx->t.pofk);
  return pofk_finish(ctx, out, true);
}

static double force_norm(const mgpu_ctx *ctx) {
  return 1.0 / (std::pow((double)ctx->N, 3) * (2. * M_PI / ctx->p.box));
}

// The f(R)/m-beta mass term and DivideByLaplacian normalisation (src/mg.h:21-116)
static double fofr_massterm2(const mgpu_ctx *ctx, const mgpu_mg *mg) {
  return mg->a * mg->a * mg->mass2 / std::pow(2.0 * MGPU_PI * MGPU_INVERSE_H0_MPCH / ctx->p.box, 2);
}
static double laplace_norm(const mgpu_ctx *ctx, const mgpu_mg *mg) {
  double nf = 1.0 / std::pow((double)ctx->N, 3);
  return nf * 1.5 * ctx->p.omega / mg->a * std::pow(ctx->p.box / MGPU_INVERSE_H0_MPCH / (2.0 * MGPU_PI), 2);
}

// ComputeFifthForce (src/user_defined_functions.h:687-719).  With screening the
// chain c2r -> real-space screening -> r2c runs as five passes over slot 2:
// Green's/smoothing x pass, y pass, z screening pass, y pass, x pass with the
// EffDensitykToPhiofk epilogue (f(R), m-beta) or plain (DGP).  phi(k) ends in slot 2.
int pm_fifth_force(mgpu_ctx *ctx, const mgpu_mg *mg) {
  PhaseTimer pt(ctx, &ctx->t.modified_gravity);
  ctx->have_phik = false;
  if (!ctx->p.mg_active || ctx->p.model == MGPU_LCDM) return MGPU_OK;
  if (!mg) return mgpu_fail(ctx, MGPU_E_INVALID, "modified gravity active but no mgpu_mg coefficients given");
  const int N = ctx->N, HP = ctx->HP;
  const size_t nc = ctx->slot_elems;
  const unsigned gc = cdiv(nc, kBlock);
  float2 *dk = ctx->slot[0], *phi = ctx->slot[2];
  const float *dx = (const float *)ctx->slot[1];
  ctx->phi_slot = 2;
  int rc;
  const double cbytes = 16.0 * (double)N * N * (N / 2 + 1);
  switch (ctx->p.model) {
    case MGPU_BRANSDICKE: {
      // _TimeDepGeffModels (src/mg.h:127-137): P3D *= Geff(a)
      KScope ks(ctx, MGPU_K_SPECTRAL, cbytes);
      hipLaunchKernelGGL(k_scale, dim3(gc), dim3(kBlock), 0, ctx->stream, dk, dk, nc, mg->geff);
      MGPU_HIP(ctx, hipGetLastError());
      return MGPU_OK;
    }
    case MGPU_FOFR:
    case MGPU_MBETA: {
      const double massterm2 = fofr_massterm2(ctx, mg);
      if (ctx->p.include_screening) {
        // _PotentialScreening (src/mg.h:147-189) with DivideByLaplacian (src/mg.h:21-64)
        if ((rc = fused_green_x(ctx, FUSED_GREEN_LAPLACE, dk, nullptr, phi, nullptr, laplace_norm(ctx, mg), 0.0, 0)))
          return rc;
        if ((rc = fft_pass_col(ctx, &ctx->fft, phi, 1, 0))) return rc;
        if ((rc = fused_z_screen(ctx, 0, phi, dx, mg->phi_crit, 0.0))) return rc;
        if ((rc = fft_pass_col(ctx, &ctx->fft, phi, 1, 1))) return rc;
        if ((rc = fused_epi_phi_x(ctx, phi, mg->coupling, massterm2))) return rc;
      } else {
        // the reference re-transforms its delta(x) copy (src/mg.h:153-156), which reproduces delta(k)
        KScope ks(ctx, MGPU_K_SPECTRAL, cbytes);
        hipLaunchKernelGGL(k_eff_to_phi, dim3(gc), dim3(kBlock), 0, ctx->stream, dk, phi, N, HP, mg->coupling,
                           massterm2);
        MGPU_HIP(ctx, hipGetLastError());
      }
      ctx->have_phik = true;
      return MGPU_OK;
    }
    case MGPU_DGP: {
      if (!ctx->p.include_screening) {
        // src/mg.h:212-217: phi = delta(k) / (3 beta)
        KScope ks(ctx, MGPU_K_SPECTRAL, cbytes);
        hipLaunchKernelGGL(k_scale, dim3(gc), dim3(kBlock), 0, ctx->stream, dk, phi, nc, mg->coupling);
        MGPU_HIP(ctx, hipGetLastError());
      } else {
        // _DensityScreening (src/mg.h:197-260) with SmoothDensityField (src/mg.h:268-309)
        const double nf = 1.0 / std::pow((double)N, 3);
        if ((rc = fused_green_x(ctx, FUSED_GREEN_SMOOTH, dk, nullptr, phi, nullptr, nf,
                                2.0 * MGPU_PI / ctx->p.box * mg->rsmooth, ctx->p.filter)))
          return rc;
        if ((rc = fft_pass_col(ctx, &ctx->fft, phi, 1, 0))) return rc;
        if ((rc = fused_z_screen(ctx, 1, phi, dx, mg->coupling, mg->dgp_fac))) return rc;
        if ((rc = fft_pass_col(ctx, &ctx->fft, phi, 1, 1))) return rc;
        if ((rc = fft_pass_col(ctx, &ctx->fft, phi, 0, 1))) return rc;
      }
      ctx->have_phik = true;
      return MGPU_OK;
    }
    default:
      return mgpu_fail(ctx, MGPU_E_INVALID, "unknown model");
  }
}

// The force tail shared by pm_forces and the fused step: slot 3 holds the x
// transform of i n_x G and slot 1 that of G (the delta(x) copy and phi(k) are
// used up by then); one dual y pass (i n_y G into slot 2), the y pass of slot 3
// and the three z c2r give F_x (slot 3), F_y (slot 2), F_z (slot 1).
static int forces_tail(mgpu_ctx *ctx) {
  int rc;
  ctx->have_phik = false;
  if ((rc = fused_dual_y(ctx, ctx->slot[1], ctx->slot[2]))) return rc;
  if ((rc = fft_pass_col(ctx, &ctx->fft, ctx->slot[3], 1, 0))) return rc;
  if ((rc = fft_pass_z(ctx, &ctx->fft, ctx->slot[3], 0, nullptr))) return rc;
  if ((rc = fft_pass_z(ctx, &ctx->fft, ctx->slot[2], 0, nullptr))) return rc;
  if ((rc = fft_pass_z_ikz(ctx, &ctx->fft, ctx->slot[1]))) return rc;
  ctx->force_slot[0] = 3;
  ctx->force_slot[1] = 2;
  ctx->force_slot[2] = 1;
  return MGPU_OK;
}

bool pm_step_fusable(const mgpu_ctx *ctx, const mgpu_mg *mg, const mgpu_pofk *out) {
  if (ctx->nu_on) return false;  // the neutrino term sits between the density's x pass and the next
  if (out && (!ctx->pk_bin || ctx->pk_geom.size() != 2 * (size_t)ctx->p.pofk_nbins)) return false;
  // the density pair bins P(k) in LDS: 8-column tiles with the Green's force (LCDM), else the plan's width
  const bool lcdm = !ctx->p.mg_active || ctx->p.model == MGPU_LCDM;
  if (out && !pofk_fused_fits(ctx, lcdm ? 8 : ctx->fft.tcols)) return false;
  if (lcdm) return true;
  if (!mg || !ctx->p.include_screening) return false;
  return ctx->p.model == MGPU_FOFR || ctx->p.model == MGPU_MBETA || ctx->p.model == MGPU_DGP;
}

// GetDisplacements (src/auxPM.c:37-103) with the x passes in pairs: the
// density's forward x pass (with P(k)) carries the inverse x pass of the
// Green's force (LCDM) or of the Laplacian / smoothing (screened MG), and the
// screened effective density's forward x pass carries the MG epilogue and the
// inverse x pass of the force.  Same arithmetic as pm_transform, pm_fifth_force
// and pm_forces; phi(k) is not stored.
int pm_step_fused(mgpu_ctx *ctx, const mgpu_mg *mg, mgpu_pofk *out, double sumD[3]) {
  int rc;
  if ((rc = pm_deposit(ctx))) return rc;
  float2 *dk = ctx->slot[0];
  double *pk = nullptr;
  if (out) {
    double *dwin;
    if ((rc = pofk_setup(ctx, &dwin))) return rc;
    pk = ctx->pk;
  }
  {
    PhaseTimer pt(ctx, &ctx->t.ptomesh);
    PhaseTimer pf(ctx, &ctx->t.fft);
    if ((rc = fft_pass_z(ctx, &ctx->fft, dk, 1, nullptr, (const float *)ctx->slot[ctx->dx_slot]))) return rc;
    if ((rc = fft_pass_col(ctx, &ctx->fft, dk, 1, 1))) return rc;
  }
  ctx->have_phik = false;
  const double cnorm = force_norm(ctx);
  if (!ctx->p.mg_active || ctx->p.model == MGPU_LCDM) {
    PhaseTimer pt(ctx, &ctx->t.forces);
    if ((rc = fused_x_pair(ctx, FUSED_FWD_DENSITY, FUSED_GREEN_FORCE, dk, dk, nullptr, ctx->slot[3], ctx->slot[1], cnorm,
                           0.0, 0, 0.0, 0.0, pk)))
      return rc;
  } else {
    PhaseTimer pt(ctx, &ctx->t.modified_gravity);
    float2 *phi = ctx->slot[2];
    const float *dx = (const float *)ctx->slot[1];
    const bool dgp = ctx->p.model == MGPU_DGP;
    if (dgp) {
      if ((rc = fused_x_pair(ctx, FUSED_FWD_DENSITY, FUSED_GREEN_SMOOTH, dk, dk, nullptr, phi, nullptr,
                             1.0 / std::pow((double)ctx->N, 3), 2.0 * MGPU_PI / ctx->p.box * mg->rsmooth,
                             ctx->p.filter, 0.0, 0.0, pk)))
        return rc;
    } else if ((rc = fused_x_pair(ctx, FUSED_FWD_DENSITY, FUSED_GREEN_LAPLACE, dk, dk, nullptr, phi, nullptr,
                                  laplace_norm(ctx, mg), 0.0, 0, 0.0, 0.0, pk))) {
      return rc;
    }
    if ((rc = fft_pass_col(ctx, &ctx->fft, phi, 1, 0))) return rc;
    if ((rc = fused_z_screen(ctx, dgp ? 1 : 0, phi, dx, dgp ? mg->coupling : mg->phi_crit, dgp ? mg->dgp_fac : 0.0)))
      return rc;
    if ((rc = fft_pass_col(ctx, &ctx->fft, phi, 1, 1))) return rc;
    if ((rc = fused_x_pair(ctx, dgp ? FUSED_FWD_PLAIN : FUSED_FWD_EPI, FUSED_GREEN_FORCE, phi, nullptr, dk,
                           ctx->slot[3], ctx->slot[1], cnorm, 0.0, 0, dgp ? 0.0 : mg->coupling,
                           dgp ? 0.0 : fofr_massterm2(ctx, mg), nullptr)))
      return rc;
  }
  {
    PhaseTimer pt(ctx, &ctx->t.forces);
    if ((rc = forces_tail(ctx))) return rc;
  }
  if (out) {
    PhaseTimer pt(ctx, &ctx->t.pofk);
    if ((rc = pofk_finish(ctx, out, true))) return rc;
  }
  return pm_gather(ctx, sumD);
}

// Forces (src/auxPM.c:437-555): F_a = c2r(i n_a (-1/n^2) P / (N^3 2pi/Box)), P = delta(k) (+ phi(k)).
// G = -P/(n^2 N^3 2pi/Box) is formed once in the x pass, which also writes the x
// transform of i n_x G; one y pass yields the y transforms of G and i n_y G, and
// the z component's i n_z is applied in its c2r.  Meshes: F_x slot 3, F_y slot 5, F_z slot 4.
int pm_forces(mgpu_ctx *ctx) {
  PhaseTimer pt(ctx, &ctx->t.forces);
  int rc;
  if ((rc = fused_green_x(ctx, FUSED_GREEN_FORCE, ctx->slot[0], ctx->have_phik ? ctx->slot[ctx->phi_slot] : nullptr,
                          ctx->slot[3], ctx->slot[1], force_norm(ctx), 0.0, 0)))
    return rc;
  return forces_tail(ctx);
}

int pm_gather(mgpu_ctx *ctx, double sumD[3]) {
  PhaseTimer pt(ctx, &ctx->t.mtoparticles);
  if (ctx->np == 0) {
    if (sumD) sumD[0] = sumD[1] = sumD[2] = 0;
    return MGPU_OK;
  }
  const int N = ctx->N;
  const unsigned nb = cdiv(ctx->np, kChunk);
  const float *f0 = (const float *)ctx->slot[ctx->force_slot[0]], *f1 = (const float *)ctx->slot[ctx->force_slot[1]],
              *f2 = (const float *)ctx->slot[ctx->force_slot[2]];
  // compulsory: positions and Disp (12 B each per particle) and the three force meshes once
  const double bytes = 24.0 * ctx->np + 12.0 * (double)N * N * N;
  if (ctx->sorted_valid) {
    // by the deposit's tiles, then the sums in particle order
    {
      KScope ks(ctx, MGPU_K_GATHER, bytes);
      int rc = gather_tiled(ctx, f0, f1, f2);
      if (rc) return rc;
    }
  } else {
    if (ctx->bz > 1) return mgpu_fail(ctx, MGPU_E_STATE, "a batched gather needs the bins of the last deposit");
    KScope ks(ctx, MGPU_K_GATHER, bytes);
    hipLaunchKernelGGL(k_gather, dim3(cdiv(ctx->np, kBlock)), dim3(kBlock), 0, ctx->stream, ctx->pos, ctx->np, N,
                       2 * ctx->HP, (double)N / ctx->p.box, f0, f1, f2, ctx->disp);
  }
  MGPU_HIP(ctx, hipGetLastError());
  {
    KScope ks(ctx, MGPU_K_MEAN, 12.0 * ctx->np);
    hipLaunchKernelGGL(k_sum3, dim3(nb, 1, ctx->bz), dim3(kBlock), 0, ctx->stream, ctx->disp, ctx->np, ctx->partial,
                       ctx->zs);
  }
  MGPU_HIP(ctx, hipGetLastError());
  double s[3];
  int rc = pm_reduce3(ctx, (int)nb, sumD ? s : nullptr, ctx->dsum);
  if (rc || !sumD) return rc;
  for (int a = 0; a < 3; a++) sumD[a] = s[a] / (double)ctx->np;
  return MGPU_OK;
}

int pm_kick(mgpu_ctx *ctx, double dda, double A, double s1, double s2, const double sumD[3], double sumV[3]) {
  PhaseTimer pt(ctx, &ctx->t.kick);
  if (ctx->np == 0) {
    sumV[0] = sumV[1] = sumV[2] = 0;
    return MGPU_OK;
  }
  const unsigned nb = cdiv(ctx->np, kChunk);
  const int sd = ctx->p.scale_dependent ? 1 : 0;
  {
    // vel, Disp read + written, one (scale-dependent) or two LPT vectors read
    KScope ks(ctx, MGPU_K_KICK, (48.0 + (sd ? 12.0 : 24.0)) * ctx->np);
    hipLaunchKernelGGL(k_kick, dim3(nb), dim3(kBlock), 0, ctx->stream, ctx->vel, ctx->disp, ctx->lpt[0], ctx->lpt[1],
                       ctx->np, s1, s2, sd, ctx->p.omega, dda, A, ctx->p.use_cola, sumD[0], sumD[1], sumD[2],
                       ctx->partial);
  }
  MGPU_HIP(ctx, hipGetLastError());
  double s[3];
  ctx->vsum_valid = true;
  int rc = pm_reduce3(ctx, (int)nb, s, ctx->dsum + 3);
  if (rc) return rc;
  for (int a = 0; a < 3; a++) sumV[a] = s[a] / (double)ctx->np;
  return MGPU_OK;
}

int pm_drift(mgpu_ctx *ctx, double dyyy, double s1, double s2, const double sumV[3]) {
  PhaseTimer pt(ctx, &ctx->t.drift);
  ctx->sorted_valid = false;
  if (ctx->np == 0) return MGPU_OK;
  const int sd = ctx->p.scale_dependent ? 2 : 0;
  {
    // pos read + written, vel read, one or two LPT vectors read
    KScope ks(ctx, MGPU_K_DRIFT, (36.0 + (sd ? 12.0 : 24.0)) * ctx->np);
    hipLaunchKernelGGL(k_drift, dim3(cdiv(ctx->np, kBlock)), dim3(kBlock), 0, ctx->stream, ctx->pos, ctx->vel,
                       ctx->lpt[0], ctx->lpt[1], ctx->np, s1, s2, sd, (float)ctx->p.box, dyyy, ctx->p.use_cola,
                       sumV[0], sumV[1], sumV[2]);
  }
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}

// sumD and sumV nullptr: the means are read from the device-resident sums and
// the new velocity sums stay on the device (nothing waits for the device)
int pm_kick_drift(mgpu_ctx *ctx, double dda, double A, double s1k, double s2k, const double sumD[3], double dyyy,
                  double s1d, double s2d, double sumV[3]) {
  PhaseTimer pt(ctx, &ctx->t.kick);
  ctx->sorted_valid = false;
  const bool dev = !sumD;
  if (ctx->np == 0) {
    if (sumV) sumV[0] = sumV[1] = sumV[2] = 0;
    return MGPU_OK;
  }
  const double z3[3] = {0.0, 0.0, 0.0};
  const double *hD = dev ? z3 : sumD, *hV = dev ? z3 : sumV;
  const unsigned nb = cdiv(ctx->np, kChunk);
  const bool sdep = ctx->p.scale_dependent;
  {
    // vel, pos read + written, Disp and the LPT vectors read
    KScope ks(ctx, MGPU_K_KICK, 84.0 * ctx->np);
    hipLaunchKernelGGL(k_kick_drift, dim3(nb, 1, ctx->bz), dim3(kBlock), 0, ctx->stream, ctx->vel, ctx->pos,
                       (const float *)ctx->disp, (const float *)ctx->lpt[0], (const float *)ctx->lpt[1], ctx->np,
                       sdep ? 1 : 0, s1k, s2k, sdep ? 2 : 0, s1d, s2d, ctx->p.omega, dda, A, dyyy,
                       (float)ctx->p.box, ctx->p.use_cola, hD[0], hD[1], hD[2], hV[0], hV[1], hV[2],
                       dev ? (const double *)ctx->dsum : nullptr, ctx->partial, ctx->zs);
  }
  MGPU_HIP(ctx, hipGetLastError());
  double s[3];
  ctx->vsum_valid = true;
  int rc = pm_reduce3(ctx, (int)nb, dev ? nullptr : s, ctx->dsum + 3);
  if (rc || dev) return rc;
  for (int a = 0; a < 3; a++) sumV[a] = s[a] / (double)ctx->np;
  return MGPU_OK;
}

// compute_RSD_powerspectrum for one line of sight (src/compute_pofk.c:403-455,
// :496-753): redshift-space deposit, r2c and the multipoles, normalised as
// the reference does (Legendre combination P4 before P2, shot noise off all
// three, k of the bin from k_from_index).  Uses slot 3 and the deposit scratch.
int pm_rsd_pofk(mgpu_ctx *ctx, int axis, double vnorm, double s1, double s2, mgpu_rsd_pofk *out) {
  PhaseTimer pt(ctx, &ctx->t.pofk);
  const int N = ctx->N, nbins = ctx->p.pofk_nbins;
  if (axis != 1 && axis != 2) return mgpu_fail(ctx, MGPU_E_INVALID, "RSD line of sight must be y (1) or z (2)");
  if (!ctx->pk_bin) return mgpu_fail(ctx, MGPU_E_INVALID, "RSD multipoles need 0 < pofk_nbins <= 32767");
  if (!ctx->rsd_crd) {
    MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    MGPU_HIP(ctx, hipMalloc((void **)&ctx->rsd_crd, sizeof(float4) * std::max<size_t>(ctx->np, 1)));
  }
  ctx->have_forces = false;
  float2 *dk = ctx->slot[3];
  if (ctx->np) {
    KScope ks(ctx, MGPU_K_POFK, 52.0 * ctx->np);
    hipLaunchKernelGGL(k_rsd_coords, dim3(cdiv(ctx->np, kBlock)), dim3(kBlock), 0, ctx->stream, ctx->pos, ctx->vel,
                       ctx->lpt[0], ctx->lpt[1], ctx->np, axis, vnorm, (double)N / ctx->p.box,
                       ctx->p.scale_dependent ? 1 : 0, ctx->p.use_cola, s1, s2, N, ctx->rsd_crd);
    MGPU_HIP(ctx, hipGetLastError());
  }
  int rc;
  if ((rc = deposit_tiled(ctx, (float *)dk, ctx->rsd_crd))) return rc;
  if ((rc = fft_r2c_3d(ctx, &ctx->fft, dk, (const float *)dk))) return rc;
  MGPU_HIP(ctx, hipMemsetAsync(ctx->pk, 0, 4 * (size_t)nbins * sizeof(double), ctx->stream));
  const size_t shmem = 4 * (size_t)nbins * sizeof(double);
  if (shmem > 65536)
    MGPU_HIP(ctx, hipFuncSetAttribute((const void *)k_rsd_bins, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  const size_t modes = (size_t)N * N * (N / 2 + 1);
  {
    KScope ks(ctx, MGPU_K_POFK, 8.0 * modes);
    hipLaunchKernelGGL(k_rsd_bins, dim3(std::min<size_t>(cdiv(modes, kBlock), 4096)), dim3(kBlock), shmem, ctx->stream,
                       dk, N, ctx->HP, ctx->winv, ctx->pk_bin, nbins, ctx->pk);
  }
  MGPU_HIP(ctx, hipGetLastError());
  std::vector<double> h(4 * (size_t)nbins);
  MGPU_HIP(ctx, hipMemcpyAsync(h.data(), ctx->pk, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const double box3 = std::pow(ctx->p.box, 3), shot = std::pow(ctx->p.box / (double)ctx->Np, 3);
  const double kmin = ctx->p.pofk_kmin, kmax = ctx->p.pofk_kmax;
  for (int b = 0; b < nbins; b++) {
    double p0 = h[b], p2 = h[nbins + b], p4 = h[2 * nbins + b], n = h[3 * nbins + b], k = 0.0;
    if (n > 0) {
      p0 = (p0 / n) * box3;
      p2 = (p2 / n) * box3;
      p4 = (p4 / n) * box3;
      p4 = 9.0 * (35.0 * p4 - 30.0 * p2 + 3.0 * p0) / 8.0;
      p2 = 5.0 * (3.0 * p2 - 1.0 * p0) / 2.0;
      if (ctx->p.pofk_subtract_shotnoise) {
        p0 -= shot;
        p2 -= shot;
        p4 -= shot;
      }
      // k_from_index (src/compute_pofk.c:51-64)
      if (ctx->p.pofk_bintype == 0) k = (kmin + (kmax - kmin) / (double)nbins * b) * 2.0 * M_PI / ctx->p.box;
      else k = std::exp(std::log(kmin) + std::log(kmax / kmin) / (double)nbins * b) * 2.0 * M_PI / ctx->p.box;
    }
    if (out->P0) out->P0[b] = p0;
    if (out->P2) out->P2[b] = p2;
    if (out->P4) out->P4[b] = p4;
    if (out->k) out->k[b] = k;
    if (out->n) out->n[b] = n;
  }
  return MGPU_OK;
}
