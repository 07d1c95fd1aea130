// RANLUX in the double-precision form GSL ships as ranlxd1 (the generator the
// reference allocates at src/2LPT.c:259): 48-bit lagged subtract-with-borrow
// x_n = x_{n-5} - x_{n-12} - c (mod 1), 202 values generated per 12 returned,
// state seeded from a 31-bit shift register.  Written for this project from
// Luescher's published algorithm; usable on host and device.
//
// The state is kept in canonical order y[0] = oldest ... y[11] = newest so that
// every index is a compile-time constant after unrolling (no private-memory
// spills on the GPU).
#pragma once

#include <hip/hip_runtime.h>

struct Ranlxd1 {
  double y[12];
  double carry;
  int pos;  // next output index into y after an advance; 12 = exhausted

  __host__ __device__ static constexpr double one_bit() { return 1.0 / 281474976710656.0; }

  __host__ __device__ void seed(unsigned long s) {
    int xbit[31];
    if (s == 0) s = 1;
    long i = (long)(s & 0xFFFFFFFFUL);
    for (int k = 0; k < 31; k++) {
      xbit[k] = (int)(i % 2);
      i /= 2;
    }
    int ibit = 0, jbit = 18;
    for (int k = 0; k < 12; k++) {
      double x = 0;
      for (int l = 1; l <= 48; l++) {
        double b = (double)((xbit[ibit] + 1) % 2);
        x += x + b;
        xbit[ibit] = (xbit[ibit] + xbit[jbit]) % 2;
        ibit = (ibit + 1) % 31;
        jbit = (jbit + 1) % 31;
      }
      y[k] = one_bit() * x;
    }
    carry = 0;
    pos = 12;
  }

  // one subtract-with-borrow step writing ring slot s (s = write position)
  __host__ __device__ __forceinline__ static double swb(double a, double b, double &c) {
    double t = a - b - c;
    if (t < 0) {
      c = one_bit();
      t += 1;
    } else {
      c = 0;
    }
    return t;
  }

  // 202 steps; afterwards y is re-canonicalised (oldest first)
  __host__ __device__ void advance() {
    double c = carry;
    for (int round = 0; round < 16; ++round) {
#pragma unroll
      for (int s = 0; s < 12; ++s) y[s] = swb(y[(s + 7) % 12], y[s], c);
    }
#pragma unroll
    for (int s = 0; s < 10; ++s) y[s] = swb(y[(s + 7) % 12], y[s], c);
    // ring now starts at slot 10: rotate left by 10
    double t[12];
#pragma unroll
    for (int s = 0; s < 12; ++s) t[s] = y[(s + 10) % 12];
#pragma unroll
    for (int s = 0; s < 12; ++s) y[s] = t[s];
    carry = c;
    pos = 0;
  }

  __host__ __device__ __forceinline__ double pick(int u) const {
    double r = y[0];
#pragma unroll
    for (int s = 1; s < 12; ++s) r = (u == s) ? y[s] : r;
    return r;
  }

  // gsl_rng_uniform: [0, 1) with 48-bit resolution
  __host__ __device__ double uniform() {
    if (pos == 12) advance();
    return pick(pos++);
  }
};
