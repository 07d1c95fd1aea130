// CIC deposit for gfx950 by tiles (PtoMesh's scatter, src/auxPM.c:295-356).
//
// Global 64-bit atomics on an 8 N^3-byte accumulator are bound by the L2's
// atomic rate (one per particle corner).  Here every particle is instead
// binned into its 16^3-cell tile (wave-aggregated counters, prefix sum,
// scatter of particle indices), and each tile accumulates its particles in
// LDS with 64-bit integer atomics on a (16+1)^3 block.  The tile stores the
// cells it owns into a 64-bit accumulator mesh (full 128-byte row segments)
// and its +1 halo layer into plane, edge and corner buffers laid out like the
// mesh cells they belong to (the x = 16(ti+1) plane, ...).  A streaming pass
// then reads the accumulator, adds to every low-face cell (a local coordinate
// 0) the parts its lower neighbours left in those buffers -- coalesced reads,
// no per-cell search -- and writes delta = acc 2^-32 - 1 in the real mesh
// layout.  The weights and their 2^-32 fixed-point rounding are those of the
// direct deposit, and integer sums do not depend on order, so delta is
// bit-identical to a serial deposit of the same particles.
#include "mgpu_internal.h"

#include <cmath>

namespace {

constexpr double kFix = 4294967296.0;  // 2^32
constexpr int kTB = 256;

constexpr double kInvFix = 1.0 / 4294967296.0;

struct TileGeom {
  int N, tx, ty, tz;   // mesh size and tile dims (cubic tiles: tx = ty = tz)
  int ntx, nty, ntz;   // tiles per dimension
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

// halo parts of the tiles, stored at the mesh cell they belong to.  With nt
// tiles per dimension and p, q, r the tile-plane indices (cell / tile size):
//   hx[p][y][z], hy[x][q][z], hz[r][x][y]       (faces: x, y or z local 0)
//   exy[p][q][z], exz[p][r][y], eyz[q][r][x]     (edges)
//   c[p][q][r]                                   (corner)
struct Halo {
  unsigned long long *hx, *hy, *hz, *exy, *exz, *eyz, *c;
};
size_t halo_cells(int N, int t) {
  const size_t n = N, nt = N / t;
  return 3 * nt * n * n + 3 * nt * nt * n + nt * nt * nt;
}
Halo halo_carve(unsigned long long *base, int N, int t) {
  const size_t n = N, nt = N / t;
  Halo h;
  h.hx = base;
  h.hy = h.hx + nt * n * n;
  h.hz = h.hy + nt * n * n;
  h.exy = h.hz + nt * n * n;
  h.exz = h.exy + nt * nt * n;
  h.eyz = h.exz + nt * nt * n;
  h.c = h.eyz + nt * nt * n;
  return h;
}

__device__ __forceinline__ float to_delta(unsigned long long v) { return (float)((double)v * kInvFix - 1.0); }

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
                                                  unsigned long long *__restrict__ acc, Halo h) {
  extern __shared__ unsigned long long blk[];
  const int bx = g.tx + 1, by = g.ty + 1, bz = g.tz + 1, nblk = bx * by * bz;
  const unsigned t = blockIdx.x;
  const int ti = t / (g.nty * g.ntz), tj = (t / g.ntz) % g.nty, tk = t % g.ntz;
  const unsigned X0 = ti * g.tx, Y0 = tj * g.ty, Z0 = tk * g.tz;
  for (int e = threadIdx.x; e < nblk; e += blockDim.x) blk[e] = 0ull;
  __syncthreads();
  const unsigned p0 = off[t], p1 = off[t + 1];
  // U particles per thread and round: all index loads, then all position loads,
  // are in flight before the first deposit
  constexpr int U = 4;
  for (unsigned base = p0; base < p1; base += U * blockDim.x) {
    unsigned ii[U];
    float px[U], py[U], pz[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const unsigned q = base + u * blockDim.x + threadIdx.x;
      ii[u] = idx[q < p1 ? q : p1 - 1];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      px[u] = pos[ii[u]];
      py[u] = pos[np + ii[u]];
      pz[u] = pos[2 * np + ii[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
    if (base + u * blockDim.x + threadIdx.x >= p1) break;
    unsigned IX, IY, IZ;
    double DX, DY, DZ;
    cell_of((double)px[u] * scale, g.N, IX, DX);
    cell_of((double)py[u] * scale, g.N, IY, DY);
    cell_of((double)pz[u] * scale, g.N, IZ, DZ);
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
  }
  __syncthreads();
  const size_t N = g.N, nt = g.ntx;
  const int T = g.tx;
  const size_t p = (ti + 1) % g.ntx, q = (tj + 1) % g.nty, r = (tk + 1) % g.ntz;
  for (int e = threadIdx.x; e < nblk; e += blockDim.x) {
    const int lz = e % bz, ly = (e / bz) % by, lx = e / (bz * by);
    const unsigned long long v = blk[e];
    const bool fx = lx == T, fy = ly == T, fz = lz == T;
    const size_t x = X0 + lx, y = Y0 + ly, z = Z0 + lz;
    if (!fx && !fy && !fz) acc[(x * N + y) * N + z] = v;
    else if (fx && !fy && !fz) h.hx[(p * N + y) * N + z] = v;
    else if (!fx && fy && !fz) h.hy[(x * nt + q) * N + z] = v;
    else if (!fx && !fy && fz) h.hz[(r * N + x) * N + y] = v;
    else if (fx && fy && !fz) h.exy[(p * nt + q) * N + z] = v;
    else if (fx && !fy && fz) h.exz[(p * nt + r) * N + y] = v;
    else if (!fx && fy && fz) h.eyz[(q * nt + r) * N + x] = v;
    else h.c[(p * nt + q) * nt + r] = v;
  }
}

// 5. streaming pass over the cells: accumulator + (low-face cells) the halo
// parts of the lower neighbours, converted to delta in the real layout.
// Grid (N*N rows, z blocks); tile sizes are powers of two (lt = log2 of the tile size).
__global__ __launch_bounds__(kTB) void k_dep_finish(const unsigned long long *__restrict__ acc, Halo h, int N,
                                                    int lt, int HP2, float *__restrict__ d0) {
  const int z = blockIdx.y * blockDim.x + threadIdx.x;
  if (z >= N) return;
  const unsigned row = blockIdx.x;
  const size_t n = N, nt = N >> lt;
  const size_t y = row % (unsigned)N, x = row / (unsigned)N;
  unsigned long long s = acc[(size_t)row * N + z];
  const int tm = (1 << lt) - 1;
  const bool ox = (x & tm) == 0, oy = (y & tm) == 0, oz = (z & tm) == 0;
  const size_t p = x >> lt, q = y >> lt, r = (size_t)z >> lt;
  if (ox) s += h.hx[(p * n + y) * n + z];
  if (oy) s += h.hy[(x * nt + q) * n + z];
  if (oz) s += h.hz[(r * n + x) * n + y];
  if (ox && oy) s += h.exy[(p * nt + q) * n + z];
  if (ox && oz) s += h.exz[(p * nt + r) * n + y];
  if (oy && oz) s += h.eyz[(q * nt + r) * n + x];
  if (ox && oy && oz) s += h.c[(p * nt + q) * nt + r];
  d0[(size_t)row * HP2 + z] = to_delta(s);
}

TileGeom make_geom(int N) {
  TileGeom g;
  g.N = N;
  const int t = N % 16 == 0 ? 16 : N % 8 == 0 ? 8 : 4;
  g.tx = g.ty = g.tz = t;
  g.ntx = N / g.tx;
  g.nty = N / g.ty;
  g.ntz = N / g.tz;
  return g;
}

}  // namespace

// scratch sizes for the tiled deposit: counters + offsets, tile ids, ranks, indices, halo buffer
size_t deposit_scratch_bytes(int N, size_t np) {
  TileGeom g = make_geom(N);
  const size_t nt = (size_t)g.ntx * g.nty * g.ntz;
  const size_t nsum = (nt + 1023) / 1024;
  return sizeof(unsigned) * (2 * (nt + 1) + nsum + 3 * np) + sizeof(unsigned long long) * halo_cells(N, g.tx) +
         256 * 8;
}

int deposit_tiled(mgpu_ctx *ctx, unsigned long long *acc, float *d) {
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
  const Halo halo = halo_carve((unsigned long long *)take(sizeof(unsigned long long) * halo_cells(N, g.tx)), N, g.tx);
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
  const int HP2 = 2 * ctx->HP;
  const double cells = (double)N * N * N, hcells = (double)halo_cells(N, g.tx);
  {
    // compulsory: indices + positions read, accumulator and halo parts written
    KScope ks(ctx, MGPU_K_DEPOSIT, 16.0 * np + 8.0 * (cells + hcells));
    hipLaunchKernelGGL(k_dep_tile, dim3(nt), dim3(kTB), shmem, ctx->stream, ctx->pos, np, idx, off, g, scale, wpar,
                       acc, halo);
  }
  MGPU_HIP(ctx, hipGetLastError());
  {
    // accumulator and halo parts read, delta written
    KScope ks(ctx, MGPU_K_DENSITY, 8.0 * (cells + hcells) + 4.0 * cells);
    const int lt = g.tx == 16 ? 4 : g.tx == 8 ? 3 : 2;
    hipLaunchKernelGGL(k_dep_finish, dim3((unsigned)((size_t)N * N), cdiv((size_t)N, kTB)), dim3(kTB), 0,
                       ctx->stream, acc, halo, N, lt, HP2, d);
  }
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}
