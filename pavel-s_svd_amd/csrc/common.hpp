// Shared device helpers for the gfx950 SVD kernels (wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pavelsvd.h"

namespace psvd {

constexpr int kWave = 64;

template <typename T> struct Eps;
template <> struct Eps<float> { static constexpr double value = 1.1920928955078125e-07; };
template <> struct Eps<double> { static constexpr double value = 2.220446049250313e-16; };

// Butterfly sum over aligned groups of `width` lanes (width a power of two <= 64).  Every lane of
// a group ends with the bitwise-identical sum: at each level both partners add the same two
// operands, and IEEE addition is commutative.
__device__ __forceinline__ double group_sum(double v, int width) {
  for (int o = width >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, width);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
  for (int o = kWave >> 1; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = kWave >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Block-wide reductions through a small LDS array with one slot per wave.  All threads of the
// block must call them; they return the result to every thread.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *red) {
  v = wave_sum(v);
  const int w = threadIdx.x / kWave;
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < NT / kWave; ++i) t += red[i];
  return t;
}

template <int NT>
__device__ __forceinline__ double block_max(double v, double *red) {
  v = wave_max(v);
  const int w = threadIdx.x / kWave;
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) red[w] = v;
  __syncthreads();
  double t = red[0];
#pragma unroll
  for (int i = 1; i < NT / kWave; ++i) t = fmax(t, red[i]);
  return t;
}

// Round-robin tournament (svdbench/standalone/primitives.py:309-313) in closed form: in round r
// of an even-sized tournament on np players, seat 0 holds player 0 and seat j >= 1 holds player
// 1 + ((j - 1 - r) mod (np - 1)); pair k faces seats k and np-1-k, ordered (min, max).
__device__ __forceinline__ void rr_pair(int r, int k, int np, int &p, int &q) {
  const int nm1 = np - 1;
  auto seat = [&](int j) -> int {
    if (j == 0) return 0;
    int t = (j - 1 - r) % nm1;
    if (t < 0) t += nm1;
    return 1 + t;
  };
  const int x = seat(k), y = seat(np - 1 - k);
  p = x < y ? x : y;
  q = x < y ? y : x;
}

// Jacobi rotation for the 2x2 Gram block [[app, apq], [apq, aqq]] (primitives.py:340-347):
// tau = (aqq - app) / (2|apq|), t = sign(tau) / (|tau| + hypot(tau, 1)) (t = 1 at tau = 0),
// c = 1 / sqrt(1 + t^2), s = t c sign(apq).  Computed in FP64.
__device__ __forceinline__ void jacobi_rotation(double app, double aqq, double apq, double &c, double &s) {
  const double a = fabs(apq);
  const double tau = (aqq - app) / (2.0 * a);
  double t;
  if (tau == 0.0) {
    t = 1.0;
  } else {
    t = copysign(1.0, tau) / (fabs(tau) + hypot(tau, 1.0));
  }
  c = 1.0 / sqrt(1.0 + t * t);
  s = t * c * (apq / a);
}

}  // namespace psvd
