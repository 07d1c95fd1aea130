// CIC cell index and tile helpers shared by the deposit and the gather.
#pragma once
#include <hip/hip_runtime.h>

// Cell index and fractional offset of a scaled coordinate X = pos * N / Box,
// as src/auxPM.c:298-322 forms them.  Positions outside [0, Box) (never
// produced by Drift's periodic wrap) are wrapped so every index stays in range.
// The deposit bins and the gather reads with this one function, so a particle
// is always found in the tile it was binned into.
static __device__ __forceinline__ void cell_of(double X, int N, unsigned &I, double &D) {
  if (!(X >= 0.0 && X < (double)N)) {
    if (!isfinite(X)) X = 0.0;
    X -= floor(X / N) * N;
    if (!(X >= 0.0 && X < (double)N)) X = 0.0;
  }
  I = (unsigned)X;
  D = X - (double)I;
  if (I >= (unsigned)N) I = 0;
}

// tile side of the deposit/gather tiles for mesh size N (Nmesh is 2^a or 3 2^a >= 8)
static inline int cic_tile_size(int N) { return N % 16 == 0 ? 16 : N % 8 == 0 ? 8 : 4; }
