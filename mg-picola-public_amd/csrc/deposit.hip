// CIC deposit for gfx950 by tiles (PtoMesh's scatter, src/auxPM.c:295-356).
//
// Global 64-bit atomics on an 8 N^3-byte accumulator are bound by the L2's
// atomic rate (one per particle corner).  Here every particle is instead
// binned by its T^3-cell tile and by its face class -- which of its cell's
// local coordinates equal T - 1, the cells whose +1 corners reach into the
// upper neighbours -- with wave-aggregated counters, a prefix sum and a
// scatter of particle records.  One workgroup then owns one tile: it reads
// the tile's particles and, from its seven lower neighbours, only the
// particles of the face classes that reach it, adds every corner that lands
// in the tile with 64-bit integer atomics on a T^3 block in LDS, and writes
// delta = acc 2^-32 - 1 for its cells straight into the real mesh layout.
// No accumulator mesh, halo buffer or second pass is needed.  The weights and
// their 2^-32 fixed-point rounding are those of the direct deposit, and
// integer sums do not depend on order, so delta is bit-identical to a serial
// deposit of the same particles.
//
// The gather (MtoParticles, src/auxPM.c:560-644) reuses the bins while the
// positions are unchanged: one workgroup per tile stages the three force
// meshes on the tile's (T+1)^3 cells in LDS and interpolates its particles
// from there, so the scattered per-particle traffic is 3 position loads and 3
// Disp stores instead of 24 mesh loads.
#include "mgpu_internal.h"
#include "cic.h"

#include <cmath>

namespace {

constexpr double kFix = 4294967296.0;  // 2^32
constexpr int kTB = 256;
// threads per block of the tile deposit (a multiple of 64, at least 64) and of the tiled gather
#ifndef MGPU_DEPTILE_THREADS
#define MGPU_DEPTILE_THREADS 256
#endif
#ifndef MGPU_GATHER_THREADS
#define MGPU_GATHER_THREADS 512
#endif
constexpr int kDepTileNT = MGPU_DEPTILE_THREADS, kGatherNT = MGPU_GATHER_THREADS;
static_assert(kDepTileNT % 64 == 0 && kDepTileNT >= 64 && kDepTileNT <= 1024, "deposit block size");
static_assert(kGatherNT % 64 == 0 && kGatherNT >= 64 && kGatherNT <= 1024, "gather block size");

constexpr double kInvFix = 1.0 / 4294967296.0;

struct TileGeom {
  int N, tx, ty, tz;   // mesh size and tile dims (cubic tiles: tx = ty = tz)
  int ntx, nty, ntz;   // tiles per dimension
};

__device__ __forceinline__ unsigned tile_of(const TileGeom &g, unsigned IX, unsigned IY, unsigned IZ) {
  return ((IX / g.tx) * g.nty + IY / g.ty) * g.ntz + IZ / g.tz;
}

__device__ __forceinline__ float to_delta(unsigned long long v) { return (float)((double)v * kInvFix - 1.0); }

// Mesh coordinates of a particle.  Plain: from the SoA positions.  Redshift
// space (RSD): from a record (x, y in box units, Z in mesh units as FP64 in
// the last two slots), written by k_rsd_coords.
__device__ __forceinline__ double rec_z(const float4 &r) {
  return __hiloint2double(__float_as_int(r.w), __float_as_int(r.z));
}
template <bool RSD>
__device__ __forceinline__ void particle_xyz(const float *__restrict__ pos, const float4 *__restrict__ crd, size_t np,
                                             size_t i, double scale, double &X, double &Y, double &Z) {
  if (RSD) {
    const float4 r = crd[i];
    X = (double)r.x * scale;
    Y = (double)r.y * scale;
    Z = rec_z(r);
  } else {
    X = (double)pos[i] * scale;
    Y = (double)pos[np + i] * scale;
    Z = (double)pos[2 * np + i] * scale;
  }
}

// Bin key of a particle: 8 tile + face class.
template <bool RSD>
__device__ __forceinline__ unsigned bin_key(const float *__restrict__ pos, const float4 *__restrict__ crd, size_t np,
                                            size_t i, const TileGeom &g, double scale) {
  unsigned IX, IY, IZ;
  double D, X, Y, Z;
  particle_xyz<RSD>(pos, crd, np, i, scale, X, Y, Z);
  cell_of(X, g.N, IX, D);
  cell_of(Y, g.N, IY, D);
  cell_of(Z, g.N, IZ, D);
  const unsigned m = (unsigned)g.tx - 1;
  const unsigned cls = ((IX & m) == m ? 4u : 0u) | ((IY & m) == m ? 2u : 0u) | ((IZ & m) == m ? 1u : 0u);
  return 8u * tile_of(g, IX, IY, IZ) + cls;
}

// 1. bin key of every particle, 8 tile + face class (class bits: x, y, z local
// coordinate = T - 1), and its rank inside the bin.  Each thread takes kBinU
// particles (workgroup chunks of kBinU kTB, coalesced per u).  The workgroup's
// keys go into a hash table in LDS, where each particle takes a local rank from
// an LDS atomic; then one global atomic per distinct key of the workgroup adds
// its count and returns the base of its ranks.  Only the rank is stored: the
// scatter forms the key again from the position.
constexpr int kBinU = 4;

// LDS hash slots of a workgroup's keys (kBinU kTB particles: at most half full, so
// every probe sequence ends at a free or matching slot)
constexpr int kBinHBits = 11, kBinH = 1 << kBinHBits;
static_assert(kBinH >= 2 * kBinU * kTB, "hash table too small for a workgroup's keys");
template <bool RSD>
__global__ __launch_bounds__(kTB) void k_dep_bin(const float *__restrict__ pos, const float4 *__restrict__ crd,
                                                 size_t np, TileGeom g, double scale, unsigned *__restrict__ count,
                                                 unsigned *__restrict__ rank, size_t zs) {
  pos = zbuf(pos, zs);
  crd = zbuf(crd, zs);
  count = zbuf(count, zs);
  rank = zbuf(rank, zs);
  constexpr unsigned kEmpty = 0xffffffffu;  // bin keys are below 8 * tiles < 2^32 - 1
  __shared__ unsigned hkey[kBinH], hcnt[kBinH];
  for (int s = threadIdx.x; s < kBinH; s += kTB) {
    hkey[s] = kEmpty;
    hcnt[s] = 0u;
  }
  __syncthreads();
  const size_t i0 = (size_t)blockIdx.x * (kBinU * kTB) + threadIdx.x;
  unsigned t[kBinU], loc[kBinU];
  int slot[kBinU];
#pragma unroll
  for (int u = 0; u < kBinU; u++) {
    const size_t i = i0 + (size_t)u * kTB;
    t[u] = i < np ? bin_key<RSD>(pos, crd, np, i, g, scale) : 0u;
    slot[u] = -1;
    loc[u] = 0u;
  }
  // the workgroup's keys into the LDS table (linear probing), a local rank per key
#pragma unroll
  for (int u = 0; u < kBinU; u++) {
    if (i0 + (size_t)u * kTB < np) {
      unsigned h = (t[u] * 2654435761u) >> (32 - kBinHBits);
      for (int p = 0; p < kBinH; ++p) {
        const unsigned prev = atomicCAS(&hkey[h], kEmpty, t[u]);
        if (prev == kEmpty || prev == t[u]) {
          slot[u] = (int)h;
          break;
        }
        h = (h + 1) & (kBinH - 1);
      }
      if (slot[u] >= 0) loc[u] = atomicAdd(&hcnt[slot[u]], 1u);
    }
  }
  __syncthreads();
  // one global atomic per distinct key of the workgroup: its base replaces the count
  for (int s = threadIdx.x; s < kBinH; s += kTB) {
    const unsigned k = hkey[s];
    if (k != kEmpty) hcnt[s] = atomicAdd(&count[k], hcnt[s]);
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kBinU; u++) {
    const size_t i = i0 + (size_t)u * kTB;
    if (i < np) rank[i] = slot[u] >= 0 ? hcnt[slot[u]] + loc[u] : atomicAdd(&count[t[u]], 1u);
  }
}

// the bin counters cleared (grid-stride)
__global__ __launch_bounds__(256) void k_zero_u32(unsigned *__restrict__ a, size_t n, size_t zs) {
  a = zbuf(a, zs);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = 0u;
}

// 2. exclusive prefix sum of the tile counts (three launches)
__global__ __launch_bounds__(1024) void k_scan_block(const unsigned *__restrict__ in, size_t n,
                                                     unsigned *__restrict__ out, unsigned *__restrict__ sums,
                                                     size_t zs) {
  in = zbuf(in, zs);
  out = zbuf(out, zs);
  sums = zbuf(sums, zs);
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
__global__ __launch_bounds__(1024) void k_scan_sums(unsigned *__restrict__ sums, int nb, size_t zs) {
  sums = zbuf(sums, zs);
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
                                                   size_t ntail, size_t zs) {
  out = zbuf(out, zs);
  sums = zbuf(sums, zs);
  const size_t i = (size_t)blockIdx.x * 1024 + threadIdx.x;
  if (i < n) out[i] += sums[blockIdx.x];
  if (i == 0) out[ntail] = total_tail;
}

// 3. particle records in bin order: (x, y, z, index as bits), or the RSD
// record; the bin key is formed again from the position (the same function as
// the binning, so the same key)
template <bool RSD>
__global__ __launch_bounds__(kTB) void k_dep_scatter(const unsigned *__restrict__ rank,
                                                     const unsigned *__restrict__ off, const float *__restrict__ pos,
                                                     const float4 *__restrict__ crd, size_t np, TileGeom g,
                                                     double scale, float4 *__restrict__ rec, size_t zs) {
  rank = zbuf(rank, zs);
  off = zbuf(off, zs);
  pos = zbuf(pos, zs);
  crd = zbuf(crd, zs);
  rec = zbuf(rec, zs);
  const size_t i0 = (size_t)blockIdx.x * (kBinU * kTB) + threadIdx.x;
  unsigned o[kBinU];
#pragma unroll
  for (int u = 0; u < kBinU; u++) {
    const size_t i = i0 + (size_t)u * kTB;
    o[u] = i < np ? off[bin_key<RSD>(pos, crd, np, i, g, scale)] + rank[i] : 0u;
  }
#pragma unroll
  for (int u = 0; u < kBinU; u++) {
    const size_t i = i0 + (size_t)u * kTB;
    if (i < np)
      rec[o[u]] = RSD ? crd[i] : make_float4(pos[i], pos[np + i], pos[2 * np + i], __uint_as_float((unsigned)i));
  }
}

// 4. one workgroup per tile at a time, by a persistent grid: each workgroup
// walks its share of the tiles.  The particle lists of a tile are up to 64
// bin ranges: for each lower-neighbour shift (sx, sy, sz) in {0, 1}^3 the bins
// of the neighbour tile whose face class contains the shift's bits.  A
// particle read through shift s has local coordinate (I mod T) - s T, so its
// corners at local coordinate -1 (s = 1) or T (s = 0) fall outside the tile
// and are skipped; every corner is added by exactly one tile.  The next
// tile's first records and the bin offsets of the one after are loaded while
// the current tile's cells are read out and stored (a barrier waits for LDS
// operations only, not for loads or stores in flight).  Tiles of one XCD are
// consecutive, so the two z neighbours sharing a 128-byte row segment write
// it through one L2.
struct RangeRegs {
  unsigned a, e;
};
__device__ __forceinline__ RangeRegs tile_ranges_issue(const unsigned *__restrict__ off, const TileGeom &g, unsigned t) {
  // lane q < 64: range q = (shift q / 8, face class q % 8) of tile t
  RangeRegs r{0u, 0u};
  const int ti = t / (g.nty * g.ntz), tj = (t / g.ntz) % g.nty, tk = t % g.ntz;
  const int sh = threadIdx.x >> 3, cls = threadIdx.x & 7;
  if (threadIdx.x < 64 && (cls & sh) == sh) {
    const int ni = (ti - (sh >> 2) + g.ntx) % g.ntx, nj = (tj - ((sh >> 1) & 1) + g.nty) % g.nty,
              nk = (tk - (sh & 1) + g.ntz) % g.ntz;
    const unsigned b = 8u * (unsigned)((ni * g.nty + nj) * g.ntz + nk) + (unsigned)cls;
    r.a = off[b];
    r.e = off[b + 1];
  }
  return r;
}
__device__ __forceinline__ void tile_ranges_store(const RangeRegs &r, unsigned *rstart, unsigned *rpre) {
  if (threadIdx.x < 64) {
    rstart[threadIdx.x] = r.a;
    unsigned v = r.e - r.a;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const unsigned u = __shfl_up(v, o);
      if (threadIdx.x >= (unsigned)o) v += u;
    }
    rpre[threadIdx.x + 1] = v;
    if (threadIdx.x == 0) rpre[0] = 0;
  }
}

constexpr int kTileU = 4;  // records per thread and round

// round `base` of a tile's records (ranges in rstart/rpre) into registers
__device__ __forceinline__ void tile_records_issue(const float4 *__restrict__ rec, const unsigned *rstart,
                                                   const unsigned *rpre, unsigned base, unsigned total,
                                                   float4 (&pr)[kTileU], int (&sh)[kTileU]) {
#pragma unroll
  for (int u = 0; u < kTileU; ++u) {
    const unsigned v = base + u * blockDim.x + threadIdx.x;
    sh[u] = -1;
    if (v < total) {
      int q = 0;
#pragma unroll
      for (int stp = 32; stp > 0; stp >>= 1)
        if (rpre[q + stp] <= v) q += stp;
      sh[u] = q >> 3;
      pr[u] = rec[rstart[q] + (v - rpre[q])];
    }
  }
}

template <bool RSD>
__device__ __forceinline__ void tile_records_add(unsigned long long *blk, const TileGeom &g, double scale,
                                                 double wpar, const float4 (&pr)[kTileU], const int (&sh)[kTileU]) {
  const int T = g.tx, tm = T - 1;
#pragma unroll
  for (int u = 0; u < kTileU; ++u) {
    if (sh[u] < 0) continue;
    unsigned IX, IY, IZ;
    double DX, DY, DZ;
    cell_of((double)pr[u].x * scale, g.N, IX, DX);
    cell_of((double)pr[u].y * scale, g.N, IY, DY);
    cell_of(RSD ? rec_z(pr[u]) : (double)pr[u].z * scale, g.N, IZ, DZ);
    const double TX = 1.0 - DX, TZ = 1.0 - DZ;
    const double TY = (1.0 - DY) * wpar;
    DY *= wpar;
    const int lx = (int)(IX & tm) - ((sh[u] >> 2) ? T : 0), ly = (int)(IY & tm) - (((sh[u] >> 1) & 1) ? T : 0),
              lz = (int)(IZ & tm) - ((sh[u] & 1) ? T : 0);
    const bool x0 = lx >= 0, x1 = lx + 1 < T, y0 = ly >= 0, y1 = ly + 1 < T, z0 = lz >= 0, z1 = lz + 1 < T;
    const int b00 = (lx * T + ly) * T + lz, b01 = b00 + T, b10 = b00 + T * T, b11 = b10 + T;
#define DEP(c, b, w) \
  if (c) atomicAdd(&blk[(b)], (unsigned long long)__double2ll_rn((w) * kFix))
    DEP(x0 && y0 && z0, b00, TX * TY * TZ);
    DEP(x0 && y0 && z1, b00 + 1, TX * TY * DZ);
    DEP(x0 && y1 && z0, b01, TX * DY * TZ);
    DEP(x0 && y1 && z1, b01 + 1, TX * DY * DZ);
    DEP(x1 && y0 && z0, b10, DX * TY * TZ);
    DEP(x1 && y0 && z1, b10 + 1, DX * TY * DZ);
    DEP(x1 && y1 && z0, b11, DX * DY * TZ);
    DEP(x1 && y1 && z1, b11 + 1, DX * DY * DZ);
#undef DEP
  }
}

// k-th tile of this workgroup, or false past its share.  With nt and the grid
// multiples of 8, XCD x (blocks b = x mod 8) owns tiles [x nt/8, (x+1) nt/8)
// and walks them in order; else tile b + k grid.
__device__ __forceinline__ bool tile_of_block(unsigned nt, unsigned k, unsigned &t) {
  const unsigned G = gridDim.x, b = blockIdx.x;
  if (nt % 8 == 0 && G % 8 == 0) {
    const unsigned j = (b >> 3) + k * (G >> 3);
    t = (b & 7) * (nt >> 3) + j;
    return j < (nt >> 3);
  }
  t = b + k * G;
  return t < nt;
}

template <bool RSD, int T>
__global__ __launch_bounds__(kDepTileNT) void k_dep_tiles(const float4 *__restrict__ rec,
                                                          const unsigned *__restrict__ off, TileGeom g, double scale,
                                                          double wpar, int HP2, float *__restrict__ d, size_t zs) {
  rec = zbuf(rec, zs);
  off = zbuf(off, zs);
  d = zbuf(d, zs);
  // The loop is rotated so that each tile's first records are loaded and
  // consumed inside one iteration with a fixed sequence of memory operations
  // between (the next tile's offsets, then this readout's stores): the wait
  // before the records' use then leaves those stores in flight.
  constexpr int T3 = T * T * T, tm = T - 1, NST = (T3 / 4 + kDepTileNT - 1) / kDepTileNT;
  extern __shared__ unsigned long long blk[];
  __shared__ unsigned rstart[2][64], rpre[2][65];
  const unsigned nt = (unsigned)g.ntx * g.nty * g.ntz;
  const size_t N = g.N;
  unsigned t, tn = 0;
  if (!tile_of_block(nt, 0, t)) return;  // uniform over the workgroup
  // prologue: tile 0's ranges stored and its records added; tile 1's offsets in flight
  tile_ranges_store(tile_ranges_issue(off, g, t), rstart[0], rpre[0]);
  bool have_next = tile_of_block(nt, 1, tn);
  RangeRegs rn{0u, 0u};
  if (have_next) rn = tile_ranges_issue(off, g, tn);
  for (int e = threadIdx.x; e < T3; e += blockDim.x) blk[e] = 0ull;
  __syncthreads();
  unsigned total = rpre[0][64];
  for (unsigned base = 0; base < total; base += kTileU * blockDim.x) {
    float4 xr[kTileU];
    int xs[kTileU];
    tile_records_issue(rec, rstart[0], rpre[0], base, total, xr, xs);
    tile_records_add<RSD>(blk, g, scale, wpar, xr, xs);
  }
  for (unsigned k = 0;; ++k) {
    const int nb = (k & 1) ^ 1;
    if (have_next) tile_ranges_store(rn, rstart[nb], rpre[nb]);
    __syncthreads();  // tile k's corners added; tile k+1's ranges visible
    float4 pr[kTileU];
    int sh[kTileU];
    bool have_next2 = false;
    unsigned t2 = 0;
    if (have_next) {
      total = rpre[nb][64];
      tile_records_issue(rec, rstart[nb], rpre[nb], 0, total, pr, sh);
      have_next2 = tile_of_block(nt, k + 2, t2);
      if (have_next2) rn = tile_ranges_issue(off, g, t2);
    }
    {
      // tile k's cells, four of a row per thread: 16-byte stores (rows start at multiples of 4 floats)
      const int ti = t / (g.nty * g.ntz), tj = (t / g.ntz) % g.nty, tk = t % g.ntz;
      const size_t X0 = (size_t)ti * T, Y0 = (size_t)tj * T, Z0 = (size_t)tk * T;
#pragma unroll
      for (int j = 0; j < NST; j++) {
        const int e4 = j * kDepTileNT + threadIdx.x;
        if (NST * kDepTileNT == T3 / 4 || e4 < T3 / 4) {
          const int e = 4 * e4;
          const int lz = e & tm, ly = (e / T) & tm, lx = e / (T * T);
          const float4 v =
              make_float4(to_delta(blk[e]), to_delta(blk[e + 1]), to_delta(blk[e + 2]), to_delta(blk[e + 3]));
          *(float4 *)&d[((X0 + lx) * N + Y0 + ly) * HP2 + Z0 + lz] = v;
        }
      }
    }
    if (!have_next) break;
    __syncthreads();  // cells read out before the clear
    for (int e = threadIdx.x; e < T3; e += blockDim.x) blk[e] = 0ull;
    __syncthreads();
    const int cb = nb;
    tile_records_add<RSD>(blk, g, scale, wpar, pr, sh);
    for (unsigned base = kTileU * blockDim.x; base < total; base += kTileU * blockDim.x) {
      float4 xr[kTileU];
      int xs[kTileU];
      tile_records_issue(rec, rstart[cb], rpre[cb], base, total, xr, xs);
      tile_records_add<RSD>(blk, g, scale, wpar, xr, xs);
    }
    t = tn;
    tn = t2;
    have_next = have_next2;
  }
}

// 5. gather by tiles.  Consecutive tiles in z go to one XCD (workgroups are
// dealt to the 8 XCDs round robin), so the +1 halo a tile reads was mostly
// fetched into the same L2 by its neighbour.
template <int T>
__global__ __launch_bounds__(kGatherNT) void k_gather_tile(const float4 *__restrict__ rec, size_t np,
                                                     const unsigned *__restrict__ off, int N, int nt1, int HP2,
                                                     double scale, const float *__restrict__ g0,
                                                     const float *__restrict__ g1, const float *__restrict__ g2,
                                                     float *__restrict__ disp, size_t zs) {
  rec = zbuf(rec, zs);
  off = zbuf(off, zs);
  g0 = zbuf(g0, zs);
  g1 = zbuf(g1, zs);
  g2 = zbuf(g2, zs);
  disp = zbuf(disp, zs);
  constexpr int B = T + 1, B2 = B * B, B3 = B2 * B;
  __shared__ float f[3][B3];
  const unsigned nt = (unsigned)nt1 * nt1 * nt1;
  const unsigned t = nt % 8 == 0 ? (blockIdx.x & 7) * (nt >> 3) + (blockIdx.x >> 3) : blockIdx.x;
  const int ti = t / (nt1 * nt1), tj = (t / nt1) % nt1, tk = t % nt1;
  const int X0 = ti * T, Y0 = tj * T, Z0 = tk * T;
  // the thread's first particle record is requested before the mesh fill
  const unsigned p1 = off[8 * t + 8];
  unsigned q = off[8 * t] + threadIdx.x;
  float4 nxt = q < p1 ? rec[q] : make_float4(0.f, 0.f, 0.f, 0.f);
  // rows (lx, ly) of each mesh: T contiguous cells, then the z = T column.
  // Every load of the thread is issued before the first LDS store, so the
  // fill costs one memory latency, not one per row group.
  constexpr int NT = kGatherNT, KR = (B2 * T + NT - 1) / NT, KC = (B2 + NT - 1) / NT;
  const int gz = Z0 + T < N ? Z0 + T : 0;
  float v[3][KR + KC];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float *__restrict__ m = a == 0 ? g0 : a == 1 ? g1 : g2;
#pragma unroll
    for (int k = 0; k < KR; k++) {
      const int e = k * NT + (int)threadIdx.x;
      if (e < B2 * T) {
        const int lz = e % T, r = e / T, lx = r / B, ly = r % B;
        const int gx = X0 + lx < N ? X0 + lx : 0, gy = Y0 + ly < N ? Y0 + ly : 0;
        v[a][k] = m[((size_t)gx * N + gy) * HP2 + Z0 + lz];
      }
    }
#pragma unroll
    for (int k = 0; k < KC; k++) {
      const int r = k * NT + (int)threadIdx.x;
      if (r < B2) {
        const int lx = r / B, ly = r % B;
        const int gx = X0 + lx < N ? X0 + lx : 0, gy = Y0 + ly < N ? Y0 + ly : 0;
        v[a][KR + k] = m[((size_t)gx * N + gy) * HP2 + gz];
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 3; a++) {
#pragma unroll
    for (int k = 0; k < KR; k++) {
      const int e = k * NT + (int)threadIdx.x;
      if (e < B2 * T) f[a][(e / T) * B + e % T] = v[a][k];
    }
#pragma unroll
    for (int k = 0; k < KC; k++) {
      const int r = k * NT + (int)threadIdx.x;
      if (r < B2) f[a][r * B + T] = v[a][KR + k];
    }
  }
  __syncthreads();
  for (; q < p1; q += NT) {
    const float4 r = nxt;
    if (q + NT < p1) nxt = rec[q + NT];
    const size_t i = __float_as_uint(r.w);
    unsigned IX, IY, IZ;
    double DX, DY, DZ;
    cell_of((double)r.x * scale, N, IX, DX);
    cell_of((double)r.y * scale, N, IY, DY);
    cell_of((double)r.z * scale, N, IZ, DZ);
    const double TX = 1.0 - DX, TY = 1.0 - DY, TZ = 1.0 - DZ;
    // the particle was binned into this tile, so 0 <= local index < T
    const int b00 = ((int)(IX - X0) * B + (int)(IY - Y0)) * B + (int)(IZ - Z0), b01 = b00 + B, b10 = b00 + B2,
              b11 = b10 + B;
    const double w000 = TX * TY * TZ, w001 = TX * TY * DZ, w010 = TX * DY * TZ, w011 = TX * DY * DZ;
    const double w100 = DX * TY * TZ, w101 = DX * TY * DZ, w110 = DX * DY * TZ, w111 = DX * DY * DZ;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const float *m = f[a];
      const double d = m[b00] * w000 + m[b00 + 1] * w001 + m[b01] * w010 + m[b01 + 1] * w011 + m[b10] * w100 +
                       m[b10 + 1] * w101 + m[b11] * w110 + m[b11 + 1] * w111;
      disp[(size_t)a * np + i] = (float)d;
    }
  }
}

TileGeom make_geom(int N) {
  TileGeom g;
  g.N = N;
  const int t = cic_tile_size(N);
  g.tx = g.ty = g.tz = t;
  g.ntx = N / g.tx;
  g.nty = N / g.ty;
  g.ntz = N / g.tz;
  return g;
}

}  // namespace

// scratch sizes for the tiled deposit: bin counters + offsets, scan sums, ranks, records
size_t deposit_scratch_bytes(int N, size_t np) {
  TileGeom g = make_geom(N);
  const size_t nb = 8 * (size_t)g.ntx * g.nty * g.ntz;
  const size_t nsum = (nb + 1023) / 1024;
  return sizeof(unsigned) * (2 * (nb + 1) + nsum + np) + sizeof(float4) * np + 256 * 8;
}

int deposit_tiled(mgpu_ctx *ctx, float *d, const float4 *rsd) {
  const int N = ctx->N;
  const size_t np = ctx->np;
  TileGeom g = make_geom(N);
  const size_t nt = (size_t)g.ntx * g.nty * g.ntz, nb = 8 * nt;
  const size_t nsum = (nb + 1023) / 1024;
  if (nsum > 1024 * 1024 || nb > 0xffffffffull) return mgpu_fail(ctx, MGPU_E_INVALID, "too many deposit tiles");
  if (np > 0xffffffffull) return mgpu_fail(ctx, MGPU_E_INVALID, "more than 2^32 - 1 particles");
  // carve the scratch buffer (256-byte aligned pieces)
  char *b = (char *)ctx->dep_scratch;
  auto take = [&](size_t bytes) {
    char *r = b;
    b += (bytes + 255) / 256 * 256;
    return r;
  };
  unsigned *count = (unsigned *)take(sizeof(unsigned) * (nb + 1));
  unsigned *off = (unsigned *)take(sizeof(unsigned) * (nb + 1));
  unsigned *sums = (unsigned *)take(sizeof(unsigned) * nsum);
  unsigned *rank = (unsigned *)take(sizeof(unsigned) * np);
  float4 *rec = (float4 *)take(sizeof(float4) * np);
  if ((size_t)(b - (char *)ctx->dep_scratch) > ctx->dep_scratch_bytes)
    return mgpu_fail(ctx, MGPU_E_INVALID, "deposit scratch too small");
  // the records stay valid for the gather until the positions change (never for
  // redshift-space positions)
  ctx->dep_rec = rec;
  ctx->dep_off = off;
  ctx->sorted_valid = np > 0 && !rsd;
  const double scale = (double)N / ctx->p.box;
  const double wpar = std::pow((double)N / (double)ctx->Np, 3);
  // cleared by a kernel, not hipMemsetAsync, so that the step can be captured in a hipGraph
  const unsigned bz = ctx->bz;
  const size_t zs = ctx->zs;
  hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)std::min<size_t>(cdiv(nb + 1, 1024), 65535), 1, bz), dim3(256), 0,
                     ctx->stream, count, nb + 1, zs);
  if (np) {
    KScope ks(ctx, MGPU_K_DEP_SORT, 16.0 * np);
    hipLaunchKernelGGL(rsd ? k_dep_bin<true> : k_dep_bin<false>, dim3(cdiv(np, kBinU * kTB), 1, bz), dim3(kTB), 0,
                       ctx->stream, ctx->pos, rsd, np, g, scale, count, rank, zs);
  }
  MGPU_HIP(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_scan_block, dim3(nsum, 1, bz), dim3(1024), 0, ctx->stream, count, nb, off, sums, zs);
  hipLaunchKernelGGL(k_scan_sums, dim3(1, 1, bz), dim3(1024), 0, ctx->stream, sums, (int)nsum, zs);
  hipLaunchKernelGGL(k_scan_add, dim3(nsum, 1, bz), dim3(1024), 0, ctx->stream, off, nb, sums, (unsigned)np, nb, zs);
  MGPU_HIP(ctx, hipGetLastError());
  if (np) {
    KScope ks(ctx, MGPU_K_DEP_SORT, 32.0 * np);
    hipLaunchKernelGGL(rsd ? k_dep_scatter<true> : k_dep_scatter<false>, dim3(cdiv(np, kBinU * kTB), 1, bz),
                       dim3(kTB), 0, ctx->stream, rank, off, ctx->pos, rsd, np, g, scale, rec, zs);
  }
  MGPU_HIP(ctx, hipGetLastError());
  const size_t shmem = sizeof(unsigned long long) * g.tx * g.ty * g.tz;
  const double cells = (double)N * N * N;
  {
    // compulsory: the particle records read once, delta written
    KScope ks(ctx, MGPU_K_DEPOSIT, 16.0 * np + 4.0 * cells);
    if (g.tx != 16 && g.tx != 8 && g.tx != 4) return mgpu_fail(ctx, MGPU_E_INVALID, "deposit tile side not 4, 8 or 16");
    auto kern = rsd ? (g.tx == 16 ? k_dep_tiles<true, 16> : g.tx == 8 ? k_dep_tiles<true, 8> : k_dep_tiles<true, 4>)
                    : (g.tx == 16 ? k_dep_tiles<false, 16> : g.tx == 8 ? k_dep_tiles<false, 8> : k_dep_tiles<false, 4>);
    // persistent grid: the workgroups that fit the device at once, a multiple of 8 when there are enough tiles
    if (!ctx->dep_tiles_grid) {
      int per_cu = 0, cus = 0;
      MGPU_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kDepTileNT, shmem));
      MGPU_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
      ctx->dep_tiles_grid = (unsigned)std::max(1, per_cu * cus);
    }
    // a batched launch shares the resident workgroups among its realizations (the integer
    // tile sums do not depend on which workgroup adds a tile)
    unsigned grid = (unsigned)std::min<size_t>(std::max<unsigned>(8u, ctx->dep_tiles_grid / bz), nt);
    if (grid >= 8) grid -= grid % 8;
    hipLaunchKernelGGL(kern, dim3(grid, 1, bz), dim3(kDepTileNT), shmem, ctx->stream, rec, off, g, scale, wpar,
                       2 * ctx->HP, d, zs);
  }
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}

// Disp at the particles from the force meshes f0, f1, f2 (real layout), by
// the bins of the last deposit (the positions must not have changed since)
int gather_tiled(mgpu_ctx *ctx, const float *f0, const float *f1, const float *f2) {
  const int N = ctx->N;
  if (!ctx->sorted_valid) return mgpu_fail(ctx, MGPU_E_STATE, "gather by tiles needs the bins of the last deposit");
  const int T = cic_tile_size(N), nt1 = N / T;
  const unsigned nt = (unsigned)nt1 * nt1 * nt1;
  const double scale = (double)N / ctx->p.box;
  auto kern = T == 16 ? k_gather_tile<16> : T == 8 ? k_gather_tile<8> : k_gather_tile<4>;
  hipLaunchKernelGGL(kern, dim3(nt, 1, ctx->bz), dim3(kGatherNT), 0, ctx->stream, ctx->dep_rec, ctx->np, ctx->dep_off,
                     N, nt1, 2 * ctx->HP, scale, f0, f1, f2, ctx->disp, ctx->zs);
  MGPU_HIP(ctx, hipGetLastError());
  return MGPU_OK;
}
