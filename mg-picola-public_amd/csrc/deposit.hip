// CIC deposit for gfx950 by tiles (PtoMesh's scatter, src/auxPM.c:295-356).
//
// Global 64-bit atomics on an 8 N^3-byte accumulator are bound by the L2's
// atomic rate (one per particle corner).  Here every particle is instead
// binned into its 16^3-cell tile (wave-aggregated counters, prefix sum,
// scatter of particle indices), each tile accumulates its particles in LDS
// with 64-bit integer atomics on a (16+1)^3 block, writes the cells it owns
// with plain stores and its +1 halo faces to a per-tile buffer, and a fix-up
// pass lets every tile add the halo faces of its lower neighbours to its own
// boundary cells.  The weights and their 2^-32 fixed-point rounding are those
// of the direct deposit, and integer sums do not depend on order, so the
// accumulator is bit-identical to a serial deposit of the same particles.
#include "mgpu_internal.h"

#include <cmath>

namespace {

constexpr double kFix = 4294967296.0;  // 2^32
constexpr int kTB = 256;

struct TileGeom {
  int N, tx, ty, tz;   // mesh size and tile dims
  int ntx, nty, ntz;   // tiles per dimension
  int hs;              // halo cells per tile
};

__device__ __forceinline__ void cell_of(double X, int N, unsigned &I, double &D) {
  if (!(X >= 0.0 && X < (double)N)) {
    if (!isfinite(X)) X = 0.0;
    X -= floor(X / N) * N;
    if (!(X >= 0.0 && X < (double)N)) X = 0.0;
  }
  I = (unsigned)X;
  D = X - (double)I;
  if (I >= (unsigned)N) I = 0;
}

__device__ __forceinline__ unsigned tile_of(const TileGeom &g, unsigned IX, unsigned IY, unsigned IZ) {
  return ((IX / g.tx) * g.nty + IY / g.ty) * g.ntz + IZ / g.tz;
}

// position of a halo cell (one local coordinate equal to the tile size) in the per-tile halo buffer
__device__ __forceinline__ int halo_index(const TileGeom &g, int lx, int ly, int lz) {
  const bool fx = lx == g.tx, fy = ly == g.ty, fz = lz == g.tz;
  const int F = g.ty * g.tz + g.tx * g.tz + g.tx * g.ty;
  if (fx && !fy && !fz) return ly * g.tz + lz;
  if (!fx && fy && !fz) return g.ty * g.tz + lx * g.tz + lz;
  if (!fx && !fy && fz) return g.ty * g.tz + g.tx * g.tz + lx * g.ty + ly;
  if (fx && fy && !fz) return F + lz;
  if (fx && !fy && fz) return F + g.tz + ly;
  if (!fx && fy && fz) return F + g.tz + g.ty + lx;
  return F + g.tz + g.ty + g.tx;
}

// 1. tile of every particle and its rank inside the tile (wave-aggregated atomics)
__global__ __launch_bounds__(kTB) void k_dep_bin(const float *__restrict__ pos, size_t np, TileGeom g, double scale,
                                                 unsigned *__restrict__ count, unsigned *__restrict__ tile,
                                                 unsigned *__restrict__ rank) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool active = i < np;
  unsigned t = 0;
  if (active) {
    unsigned IX, IY, IZ;
    double D;
    cell_of((double)pos[i] * scale, g.N, IX, D);
    cell_of((double)pos[np + i] * scale, g.N, IY, D);
    cell_of((double)pos[2 * np + i] * scale, g.N, IZ, D);
    t = tile_of(g, IX, IY, IZ);
  }
  const int lane = threadIdx.x % warpSize;
  unsigned r = 0;
  const bool valid = active;
  while (true) {
    const unsigned long long amask = __ballot(active);
    if (amask == 0) break;
    const int leader = __ffsll((long long)amask) - 1;
    const unsigned tl = __shfl(t, leader);
    const unsigned long long same = __ballot(active && t == tl);
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(&count[tl], (unsigned)__popcll(same));
    base = __shfl(base, leader);
    if (active && t == tl) {
      r = base + (unsigned)__popcll(same & ((1ull << lane) - 1ull));
      active = false;
    }
  }
  if (valid) {
    tile[i] = t;
    rank[i] = r;
  }
}

// 2. exclusive prefix sum of the tile counts (three launches)
__global__ __launch_bounds__(1024) void k_scan_block(const unsigned *__restrict__ in, size_t n,
                                                     unsigned *__restrict__ out, unsigned *__restrict__ sums) {
  __shared__ unsigned s[1024];
  const size_t i = (size_t)blockIdx.x * 1024 + threadIdx.x;
  const unsigned v = i < n ? in[i] : 0u;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    unsigned t = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0u;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = s[threadIdx.x] - v;
  if (threadIdx.x == 1023) sums[blockIdx.x] = s[1023];
}
__global__ __launch_bounds__(1024) void k_scan_sums(unsigned *__restrict__ sums, int nb) {
  // nb <= 1024 * 1024 handled by a serial carry over 1024-chunks
  __shared__ unsigned s[1024];
  unsigned carry = 0;
  for (int c0 = 0; c0 < nb; c0 += 1024) {
    const int i = c0 + threadIdx.x;
    const unsigned v = i < nb ? sums[i] : 0u;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      unsigned t = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0u;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) sums[i] = s[threadIdx.x] - v + carry;
    const unsigned tot = s[1023];
    __syncthreads();
    carry += tot;
  }
}
__global__ __launch_bounds__(1024) void k_scan_add(unsigned *__restrict__ out, size_t n,
                                                   const unsigned *__restrict__ sums, unsigned total_tail,
                                                   size_t ntail) {
  const size_t i = (size_t)blockIdx.x * 1024 + threadIdx.x;
  if (i < n) out[i] += sums[blockIdx.x];
  if (i == 0) out[ntail] = total_tail;
}

// 3. particle indices in tile order
__global__ __launch_bounds__(kTB) void k_dep_scatter(const unsigned *__restrict__ tile,
                                                     const unsigned *__restrict__ rank,
                                                     const unsigned *__restrict__ off, size_t np,
                                                     unsigned *__restrict__ idx) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  idx[off[tile[i]] + rank[i]] = (unsigned)i;
}

// 4. one workgroup per tile: LDS accumulation of the (t+1)^3 block
__global__ __launch_bounds__(kTB) void k_dep_tile(const float *__restrict__ pos, size_t np,
                                                  const unsigned *__restrict__ idx, const unsigned *__restrict__ off,
                                                  TileGeom g, double scale, double wpar,
                                                  unsigned long long *__restrict__ acc,
                                                  unsigned long long *__restrict__ halo) {
  extern __shared__ unsigned long long blk[];
  const int bx = g.tx + 1, by = g.ty + 1, bz = g.tz + 1, nblk = bx * by * bz;
  const unsigned t = blockIdx.x;
  const int ti = t / (g.nty * g.ntz), tj = (t / g.ntz) % g.nty, tk = t % g.ntz;
  const unsigned X0 = ti * g.tx, Y0 = tj * g.ty, Z0 = tk * g.tz;
  for (int e = threadIdx.x; e < nblk; e += blockDim.x) blk[e] = 0ull;
  __syncthreads();
  const unsigned p0 = off[t], p1 = off[t + 1];
  for (unsigned p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const size_t i = idx[p];
    unsigned IX, IY, IZ;
    double DX, DY, DZ;
    cell_of((double)pos[i] * scale, g.N, IX, DX);
    cell_of((double)pos[np + i] * scale, g.N, IY, DY);
    cell_of((double)pos[2 * np + i] * scale, g.N, IZ, DZ);
    const double TX = 1.0 - DX, TZ = 1.0 - DZ;
    const double TY = (1.0 - DY) * wpar;
    DY *= wpar;
    const int lx = IX - X0, ly = IY - Y0, lz = IZ - Z0;
    const int b00 = (lx * by + ly) * bz + lz, b01 = (lx * by + ly + 1) * bz + lz;
    const int b10 = ((lx + 1) * by + ly) * bz + lz, b11 = ((lx + 1) * by + ly + 1) * bz + lz;
#define DEP(b, w) atomicAdd(&blk[(b)], (unsigned long long)__double2ll_rn((w) * kFix))
    DEP(b00, TX * TY * TZ);
    DEP(b00 + 1, TX * TY * DZ);
    DEP(b01, TX * DY * TZ);
    DEP(b01 + 1, TX * DY * DZ);
    DEP(b10, DX * TY * TZ);
    DEP(b10 + 1, DX * TY * DZ);
    DEP(b11, DX * DY * TZ);
    DEP(b11 + 1, DX * DY * DZ);
#undef DEP
  }
  __syncthreads();
  const size_t N = g.N;
  for (int e = threadIdx.x; e < nblk; e += blockDim.x) {
    const int lz = e % bz, ly = (e / bz) % by, lx = e / (bz * by);
    const unsigned long long v = blk[e];
    if (lx < g.tx && ly < g.ty && lz < g.tz)
      acc[((size_t)(X0 + lx) * N + (Y0 + ly)) * N + (Z0 + lz)] = v;
    else
      halo[(size_t)t * g.hs + halo_index(g, lx, ly, lz)] = v;
  }
}

// 5. every tile adds the halo faces its lower neighbours produced for its cells
__global__ __launch_bounds__(kTB) void k_dep_fixup(unsigned long long *__restrict__ acc,
                                                   const unsigned long long *__restrict__ halo, TileGeom g) {
  const unsigned t = blockIdx.x;
  const int ti = t / (g.nty * g.ntz), tj = (t / g.ntz) % g.nty, tk = t % g.ntz;
  const size_t N = g.N;
  const int ncell = g.tx * g.ty * g.tz;
  for (int e = threadIdx.x; e < ncell; e += blockDim.x) {
    const int lz = e % g.tz, ly = (e / g.tz) % g.ty, lx = e / (g.tz * g.ty);
    if (lx != 0 && ly != 0 && lz != 0) continue;
    unsigned long long s = 0;
    for (int d = 1; d < 8; d++) {
      const int dx = d & 1, dy = (d >> 1) & 1, dz = (d >> 2) & 1;
      if ((dx && lx) || (dy && ly) || (dz && lz)) continue;
      const int ni = dx ? (ti == 0 ? g.ntx - 1 : ti - 1) : ti;
      const int nj = dy ? (tj == 0 ? g.nty - 1 : tj - 1) : tj;
      const int nk = dz ? (tk == 0 ? g.ntz - 1 : tk - 1) : tk;
      const size_t nb = ((size_t)ni * g.nty + nj) * g.ntz + nk;
      s += halo[nb * g.hs + halo_index(g, dx ? g.tx : lx, dy ? g.ty : ly, dz ? g.tz : lz)];
    }
    acc[((size_t)(ti * g.tx + lx) * N + (tj * g.ty + ly)) * N + (tk * g.tz + lz)] += s;
  }
}

TileGeom make_geom(int N) {
  TileGeom g;
  g.N = N;
  const int t = N % 16 == 0 ? 16 : N % 8 == 0 ? 8 : 4;
  g.tx = g.ty = g.tz = t;
  g.ntx = N / g.tx;
  g.nty = N / g.ty;
  g.ntz = N / g.tz;
  g.hs = g.ty * g.tz + g.tx * g.tz + g.tx * g.ty + g.tz + g.ty + g.tx + 1;
  return g;
}

}  // namespace

// scratch sizes for the tiled deposit: counters + offsets, tile ids, ranks, indices, halo buffer
size_t deposit_scratch_bytes(int N, size_t np) {
  TileGeom g = make_geom(N);
  const size_t nt = (size_t)g.ntx * g.nty * g.ntz;
  const size_t nsum = (nt + 1023) / 1024;
  return sizeof(unsigned) * (2 * (nt + 1) + nsum + 3 * np) + sizeof(unsigned long long) * nt * g.hs + 256 * 8;
}

int deposit_tiled(mgpu_ctx *ctx) {
  const int N = ctx->N;
  const size_t np = ctx->np;
  TileGeom g = make_geom(N);
  const size_t nt = (size_t)g.ntx * g.nty * g.ntz;
  const size_t nsum = (nt + 1023) / 1024;
  if (nsum > 1024 * 1024) return mgpu_fail(ctx, MGPU_E_INVALID, "too many deposit tiles");
  // carve the scratch buffer (256-byte aligned pieces)
  char *b = (char *)ctx->dep_scratch;
  auto take = [&](size_t bytes) {
    char *r = b;
    b += (bytes + 255) / 256 * 256;
    return r;
  };
  unsigned long long *halo = (unsigned long long *)take(sizeof(unsigned long long) * nt * g.hs);
  unsigned *count = (unsigned *)take(sizeof(unsigned) * (nt + 1));
  unsigned *off = (unsigned *)take(sizeof(unsigned) * (nt + 1));
  unsigned *sums = (unsigned *)take(sizeof(unsigned) * nsum);
  unsigned *tile = (unsigned *)take(sizeof(unsigned) * np);
  unsigned *rank = (unsigned *)take(sizeof(unsigned) * np);
  unsigned *idx = (unsigned *)take(sizeof(unsigned) * np);
  if ((size_t)(b - (char *)ctx->dep_scratch) > ctx->dep_scratch_bytes)
    return mgpu_fail(ctx, MGPU_E_INVALID, "deposit scratch too small");
  const double scale = (double)N / ctx->p.box;
  const double wpar = std::pow((double)N / (double)ctx->Np, 3);
  MGPU_HIP(ctx, hipMemsetAsync(count, 0, sizeof(unsigned) * (nt + 1), ctx->stream));
  if (np) {
    KScope ks(ctx, MGPU_K_DEPOSIT, 20.0 * np);
    hipLaunchKernelGGL(k_dep_bin, dim3(cdiv(np, kTB)), dim3(kTB), 0, ctx->stream, ctx->pos, np, g, scale, count, tile,
                       rank);
  }
  MGPU_HIP(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_scan_block, dim3(nsum), dim3(1024), 0, ctx->stream, count, nt, off, sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, ctx->stream, sums, (int)nsum);
  hipLaunchKernelGGL(k_scan_add, dim3(nsum), dim3(1024), 0, ctx->stream, off, nt, sums, (unsigned)np, nt);
  MGPU_HIP(ctx, hipGetLastError());
  if (np) {
    KScope ks(ctx, MGPU_K_DEPOSIT, 12.0 * np);
    hipLaunchKernelGGL(k_dep_scatter, dim3(cdiv(np, kTB)), dim3(kTB), 0, ctx->stream, tile, rank, off, np, idx);
  }
  MGPU_HIP(ctx, hipGetLastError());
  const size_t shmem = sizeof(unsigned long long) * (g.tx + 1) * (g.ty + 1) * (g.tz + 1);
  {
    // compulsory: indices + positions read, accumulator and halo written
    KScope ks(ctx, MGPU_K_DEPOSIT, 16.0 * np + 8.0 * ((double)N * N * N + (double)nt * g.hs));
    hipLaunchKernelGGL(k_dep_tile, dim3(nt), dim3(kTB), shmem, ctx->stream, ctx->pos, np, idx, off, g, scale, wpar,
                       ctx->acc, halo);
  }
  MGPU_HIP(ctx, hipGetLastError());
  {
    const double bnd = (double)nt * (g.tx * g.ty * g.tz - (g.tx - 1) * (g.ty - 1) * (g.tz - 1));
    KScope ks(ctx, MGPU_K_DEPOSIT, 8.0 * ((double)nt * g.hs + 2.0 * bnd));
    hipLaunchKernelGGL(k_dep_fixup, dim3(nt), dim3(kTB), 0, ctx->stream, ctx->acc, halo, g);
  }
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}
