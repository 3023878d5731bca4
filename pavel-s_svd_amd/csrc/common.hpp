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

// DPP lane permutation of a double (two 32-bit moves).  CTRL is a gfx9 DPP control word that
// gives every lane a source inside its own row (the quad permutations and the row mirrors used
// here), so no lane keeps an old value and the moves need no copy of their source.
template <int CTRL>
__device__ __forceinline__ double dpp_double(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, false),
                          __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, false));
}

__device__ __forceinline__ double readlane_double(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// Sum over aligned groups of L lanes: DPP within rows of 16 (quad_perm xor 1, xor 2,
// row_half_mirror, row_mirror), then v_readlane of the row sums for 32 and 64.  Each DPP level
// pairs every lane with a partner in the other half of its block, and the readlane levels hand
// one scalar to the whole group, so all lanes of a group end with the same bits.
template <int L>
__device__ __forceinline__ double group_sum_dpp(double v) {
  if constexpr (L >= 2) v += dpp_double<0xB1>(v);   // quad_perm [1,0,3,2]
  if constexpr (L >= 4) v += dpp_double<0x4E>(v);   // quad_perm [2,3,0,1]
  if constexpr (L >= 8) v += dpp_double<0x141>(v);  // row_half_mirror
  if constexpr (L >= 16) v += dpp_double<0x140>(v); // row_mirror
  if constexpr (L == 32) {
    const double r01 = readlane_double(v, 0) + readlane_double(v, 16);
    const double r23 = readlane_double(v, 32) + readlane_double(v, 48);
    v = (threadIdx.x & 32) ? r23 : r01;
  } else if constexpr (L >= 64) {
    v = (readlane_double(v, 0) + readlane_double(v, 16)) + (readlane_double(v, 32) + readlane_double(v, 48));
  }
  return v;
}

// group_sum_dpp of three values with their levels interleaved, so that the DPP moves of one
// value fill the other two's wait states; each result has the bits group_sum_dpp gives it.
template <int L>
__device__ __forceinline__ void group_sum3_dpp(double &a, double &b, double &c) {
  if constexpr (L > 16) {
    a = group_sum_dpp<L>(a);
    b = group_sum_dpp<L>(b);
    c = group_sum_dpp<L>(c);
  } else {
    if constexpr (L >= 2) { a += dpp_double<0xB1>(a); b += dpp_double<0xB1>(b); c += dpp_double<0xB1>(c); }
    if constexpr (L >= 4) { a += dpp_double<0x4E>(a); b += dpp_double<0x4E>(b); c += dpp_double<0x4E>(c); }
    if constexpr (L >= 8) { a += dpp_double<0x141>(a); b += dpp_double<0x141>(b); c += dpp_double<0x141>(c); }
    if constexpr (L >= 16) { a += dpp_double<0x140>(a); b += dpp_double<0x140>(b); c += dpp_double<0x140>(c); }
  }
}

// 16 bytes of T moved by one global load or store (global_load_dwordx4); the address must be
// 16-byte aligned.
template <typename T>
struct alignas(16) Vec16 {
  static constexpr int n = 16 / (int)sizeof(T);
  T v[16 / sizeof(T)];
};

// A value every lane of the wave holds (read back from one LDS word, say) moved to a scalar
// register: the compiler then treats what is derived from it as uniform, so branches on it are
// scalar branches and loop trip counts need no per-lane masks or selects.
__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Maximum over the whole wave (DPP within rows, readlane across rows); wave-uniform result.
__device__ __forceinline__ double wave_max_fast(double v) {
  v = fmax(v, dpp_double<0xB1>(v));
  v = fmax(v, dpp_double<0x4E>(v));
  v = fmax(v, dpp_double<0x141>(v));
  v = fmax(v, dpp_double<0x140>(v));
  return fmax(fmax(readlane_double(v, 0), readlane_double(v, 16)),
              fmax(readlane_double(v, 32), readlane_double(v, 48)));
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

// Complex LDS-resident one-sided Jacobi (csrc/jacobi_lds_complex.hip), reached through
// psvd_jacobi_lds with dtype PSVD_C64 / PSVD_C128.
size_t clds_smem_bytes(int dtype, int m, int n, bool vectors);
int run_clds(int dtype, int64_t batch, int m, int n, const void *a, void *u, void *s, void *vh, int32_t *info,
             double tol, int max_sweeps, int flags, int lanes, void *ws, hipStream_t stream);

// Complex single-launch one-sided Jacobi (csrc/jacobi_grid_complex.hip), reached through
// psvd_jacobi_grid with dtype PSVD_C64 / PSVD_C128.
size_t cgrid_workspace_bytes(int dtype, int m, int n);
int cjacobi_grid(int dtype, int m, int n, const void *a, void *u, void *s, void *vh, int32_t *info, double tol,
                 int max_sweeps, int flags, void *workspace, size_t workspace_bytes, hipStream_t stream);

// Item e of the upper triangle (k1 <= k2) of an np2 x np2 grid, row by row.
__device__ __forceinline__ void tri_decode(int64_t e, int np2, int &k1, int &k2) {
  int a = (int)((2.0 * np2 + 1.0 - sqrt((2.0 * np2 + 1.0) * (2.0 * np2 + 1.0) - 8.0 * (double)e)) / 2.0);
  if (a < 0) a = 0;
  if (a > np2 - 1) a = np2 - 1;
  while (a > 0 && (int64_t)a * np2 - (int64_t)a * (a - 1) / 2 > e) --a;
  while (a < np2 - 1 && (int64_t)(a + 1) * np2 - (int64_t)(a + 1) * a / 2 <= e) ++a;
  k1 = a;
  k2 = a + (int)(e - ((int64_t)a * np2 - (int64_t)a * (a - 1) / 2));
}

// Round-robin tournament (svdbench/standalone/primitives.py:309-313) in closed form: in round r
// of an even-sized tournament on np players, seat 0 holds player 0 and seat j >= 1 holds player
// 1 + ((j - 1 - r) mod (np - 1)); pair k faces seats k and np-1-k, ordered (min, max).
__device__ __forceinline__ void rr_pair(int r, int k, int np, int &p, int &q) {
  const int nm1 = np - 1;
  // 0 <= j - 1 <= np - 2 and 0 <= r <= np - 2, so one conditional add reduces mod (np - 1)
  auto seat = [&](int j) -> int {
    if (j == 0) return 0;
    int t = j - 1 - r;
    if (t < 0) t += nm1;
    return 1 + t;
  };
  const int x = seat(k), y = seat(np - 1 - k);
  p = x < y ? x : y;
  q = x < y ? y : x;
}

// Pair k of rotation round rr of a subproblem sweep over two blocks of KB indices (2 KB in all):
// the full tournament (rr < 2 KB - 1), or a cross sweep -- index k of the first block against index
// KB + (k + rr) mod KB of the second (rr < KB) -- which treats every pair between the blocks once and
// none inside them.  A block-Jacobi sweep whose first tournament round sweeps its block pairs fully
// and every later round only across them still treats every index pair once per sweep, as the
// scalar cyclic method does, in about half the rotation rounds.
__device__ __forceinline__ void sub_pair(bool cross, int rr, int k, int KB, int &a, int &b) {
  if (cross) {   // KB is a power of two and k + rr >= 0: the remainder is a mask
    a = k;
    b = KB + ((k + rr) & (KB - 1));
  } else {
    rr_pair(rr, k, 2 * KB, a, b);
  }
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

// The rotation of jacobi_rotation with c and s in FP64 (c^2 + s^2 = 1 to FP64 rounding, so the
// accumulated rotations stay orthogonal) and, for FP32 data, the angle t itself in FP32: tau from
// an FP64 reciprocal with one Newton step, t = 1 / (|tau| + sqrt(tau^2 + 1)) with the native FP32
// reciprocal and square root, c = 1 / sqrt(1 + t^2) by v_rsq_f64 and one Newton step.  An FP32
// angle leaves an off-diagonal of about 1e-7 of the pair's spread, below the FP32 data's 10 eps
// tolerance; FP64 data takes jacobi_rotation.
template <typename T>
__device__ __forceinline__ void jacobi_rotation_fast(double app, double aqq, double apq, double &c, double &s);

template <>
__device__ __forceinline__ void jacobi_rotation_fast<double>(double app, double aqq, double apq, double &c,
                                                             double &s) {
  jacobi_rotation(app, aqq, apq, c, s);
}

template <>
__device__ __forceinline__ void jacobi_rotation_fast<float>(double app, double aqq, double apq, double &c,
                                                            double &s) {
  const double a2 = 2.0 * fabs(apq);
  double r = __builtin_amdgcn_rcp(a2);
  r = r * fma(-a2, r, 2.0);
  const double tau = (aqq - app) * r;
  const double at = fabs(tau);
  double t;
  if (tau == 0.0) {
    t = 1.0;
  } else if (at < 1e18) {
    const float af = (float)at;
    t = (double)__builtin_amdgcn_rcpf(af + __builtin_amdgcn_sqrtf(fmaf(af, af, 1.0f)));
  } else {
    t = 0.5 / at;
  }
  t = copysign(t, tau);
  const double x = fma(t, t, 1.0);
  double y = __builtin_amdgcn_rsq(x);
  y = y * fma(-0.5 * x * y, y, 1.5);
  c = y;
  s = apq < 0.0 ? -t * y : t * y;
}

// The same rotation with the Jacobi angle evaluated in the storage precision: tau from the FP64
// sums (difference taken in FP64), t, c, s in T.  `tb` returns t*|apq| for the exact-arithmetic
// norm update app' = app - t|apq|, aqq' = aqq + t|apq|.
template <typename T>
__device__ __forceinline__ void jacobi_rotation_t(double app, double aqq, double apq, T &c, T &s, double &tb);

template <>
__device__ __forceinline__ void jacobi_rotation_t<float>(double app, double aqq, double apq, float &c, float &s,
                                                         double &tb) {
  // t in FP32 (the angle needs no more); c = 1/sqrt(1 + t^2) and s = t c in FP64 (v_rsq_f64 plus
  // one Newton step) and rounded once, so that c^2 + s^2 = 1 to an ulp and V stays orthogonal
  // over thousands of rotations.
  // native v_rcp_f32 / v_sqrt_f32 (about 1 ulp): t to a few ulps is all the angle needs
  // Branch-free: at tau = 0 (never -0: app, aqq are sums of squares) the general formula gives
  // rcp(0 + sqrt(1)) = 1 exactly, the reference's t = 1.
  // tau in FP32 when its FP64 operands are normal FP32 numbers; otherwise (unscaled data whose sums
  // leave the FP32 range, e.g. entries near 1e-25 or 1e20) the quotient is formed in FP64 and rounded,
  // so |apq| never becomes 0 or inf in FP32 and tau stays finite (or +-inf, which gives t = 0)
  const double a2 = 2.0 * fabs(apq), dd = aqq - app;
  float tau;
  if (a2 > 1e-30 && a2 < 1e30 && fabs(dd) < 1e37)
    tau = (float)dd * __builtin_amdgcn_rcpf((float)a2);
  else
    tau = (float)(dd / a2);
  const float at = fabsf(tau);
  float t = at > 1e18f ? 0.5f * __builtin_amdgcn_rcpf(at)
                       : __builtin_amdgcn_rcpf(at + __builtin_amdgcn_sqrtf(fmaf(at, at, 1.0f)));
  t = copysignf(t, tau);
  const double td = (double)t;
  const double x = fma(td, td, 1.0);
  double y = __builtin_amdgcn_rsq(x);
  y = y * fma(-0.5 * x * y, y, 1.5);
  c = (float)y;
  s = (float)(apq < 0.0 ? -td * y : td * y);
  tb = td * fabs(apq);
}

template <>
__device__ __forceinline__ void jacobi_rotation_t<double>(double app, double aqq, double apq, double &c, double &s,
                                                          double &tb) {
  jacobi_rotation(app, aqq, apq, c, s);
  const double a = fabs(apq);
  const double tau = (aqq - app) / (2.0 * a);
  const double t = tau == 0.0 ? 1.0 : copysign(1.0, tau) / (fabs(tau) + hypot(tau, 1.0));
  tb = t * a;
}

// Two-sided Jacobi update of one 2x2 block of a symmetric matrix G (ld ld): the block (rows p1, q1) x
// (columns p2, q2) becomes J1^T B J2 -- columns first, then rows, each rounded to T -- and, off the
// diagonal, its mirror is written too.  Returns the block's new max |entry|.
template <typename T>
__device__ __forceinline__ double block_update(T *G, int ld, int p1, int q1, int p2, int q2, double c1, double s1,
                                               double c2, double s2, bool act1, bool act2) {
  T *a11 = G + p1 * ld + p2, *a12 = G + p1 * ld + q2, *a21 = G + q1 * ld + p2, *a22 = G + q1 * ld + q2;
  double b11 = (double)*a11, b12 = (double)*a12, b21 = (double)*a21, b22 = (double)*a22;
  if (act2) {  // columns (p2, q2): col_p = c col_p - s col_q, col_q = s col_p + c col_q
    const double t11 = c2 * b11 - s2 * b12, t12 = s2 * b11 + c2 * b12;
    const double t21 = c2 * b21 - s2 * b22, t22 = s2 * b21 + c2 * b22;
    b11 = (double)(T)t11; b12 = (double)(T)t12; b21 = (double)(T)t21; b22 = (double)(T)t22;
  }
  if (act1) {  // rows (p1, q1): row_p = c row_p - s row_q, row_q = s row_p + c row_q
    const double t11 = c1 * b11 - s1 * b21, t21 = s1 * b11 + c1 * b21;
    const double t12 = c1 * b12 - s1 * b22, t22 = s1 * b12 + c1 * b22;
    b11 = t11; b12 = t12; b21 = t21; b22 = t22;
  }
  const T n11 = (T)b11, n12 = (T)b12, n21 = (T)b21, n22 = (T)b22;
  *a11 = n11; *a12 = n12; *a21 = n21; *a22 = n22;
  if (p1 != p2) {  // mirror block (pair k2 rows) x (pair k1 cols) is the transpose
    G[p2 * ld + p1] = n11; G[q2 * ld + p1] = n12; G[p2 * ld + q1] = n21; G[q2 * ld + q1] = n22;
  }
  return fmax(fmax(fabs((double)n11), fabs((double)n12)), fmax(fabs((double)n21), fabs((double)n22)));
}

// Rows (p_k, q_k) of the accumulated rotation Qt (ld `ld`) rotated by (c_k, s_k) for the pairs
// k = qk + j QSTEP (j < QIT, k < NP) of a round, at column qcol: one round trip to LDS for all of the
// thread's pairs -- every (c, s), (p, q) and Qt load is issued first, then the FP64 rotations, then
// the stores.  The pairs of a round are disjoint, so no row is both read and written by two items.
template <typename T, int NP, int QSTEP, int QIT>
__device__ __forceinline__ void rotate_rows_batched(T *Qt, int ld, const double2 *cs, const int2 *pq, int qk,
                                                    int qcol) {
  double2 r[QIT];
  int2 a[QIT];
  bool on[QIT];
#pragma unroll
  for (int j = 0; j < QIT; ++j) {
    const int k = qk + j * QSTEP;
    on[j] = k < NP;
    r[j] = on[j] ? cs[k] : make_double2(1.0, 0.0);
    a[j] = on[j] ? pq[k] : make_int2(0, 0);
    on[j] = on[j] && r[j].y != 0.0;
  }
  double x[QIT], y[QIT];
#pragma unroll
  for (int j = 0; j < QIT; ++j)
    if (on[j]) {
      x[j] = (double)Qt[a[j].x * ld + qcol];
      y[j] = (double)Qt[a[j].y * ld + qcol];
    }
#pragma unroll
  for (int j = 0; j < QIT; ++j)
    if (on[j]) {
      Qt[a[j].x * ld + qcol] = (T)(r[j].x * x[j] - r[j].y * y[j]);
      Qt[a[j].y * ld + qcol] = (T)(r[j].y * x[j] + r[j].x * y[j]);
    }
}

// ---- grid barrier of the package's single-launch kernels ---------------------------------------
// K2, K2c and the tall QR panel run every workgroup of an ordinary launch of at most one workgroup
// per CU (all of them resident on an otherwise idle GPU) and separate their phases with this
// barrier instead of a cooperative launch: workgroups arrive on one of kBarGroups counters 256 bytes
// apart (one counter when there are at most kBarFlat workgroups), the last arrival of a group bumps
// the top counter, and every workgroup polls the top counter with agent-scope loads.  Counters only
// grow during a launch (bar_reset_kernel clears them in-stream before it), so barrier e of the launch
// waits for e x (number of groups).  Memory: every wave waits for its own stores before the
// workgroup barrier; with FENCE the arriving thread then writes the XCD's L2 back (agent-scope
// release) and, after the wait, invalidates its caches (agent-scope acquire), so plain global data
// written before the barrier is read fresh after it; without FENCE the data that crosses the
// barrier must be written and read with agent-scope atomic accesses (st_agent / ld_agent).  A wait
// that exceeds kBarSpinLimit polls (workgroups that are not co-resident) raises the abort word, every
// waiter then returns false, and the kernel ends early with its failure reported.
constexpr int kBarGroups = 8, kBarFlat = 64;
constexpr unsigned kBarSpinLimit = 1u << 23;
constexpr int kBarWords = 64 * (2 + kBarGroups);   // top, abort, group counters: one 256-byte line each

__global__ void bar_reset_kernel(unsigned *ctr, int words);

template <typename T>
__device__ __forceinline__ void st_agent(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T ld_agent(const T *p) {
  return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct GridBarrier {
  unsigned *ctr;      // kBarWords counters of this barrier (zeroed before the launch)
  unsigned nwg, idx;  // workgroups taking part and this workgroup's index among them
  unsigned epoch;

  __device__ GridBarrier(unsigned *c, unsigned n, unsigned i) : ctr(c), nwg(n), idx(i), epoch(0) {}

  // Every thread of every participating workgroup calls sync() the same number of times.  Returns
  // false (in every thread) when the wait was abandoned.
  template <bool FENCE>
  __device__ bool sync() {
    __shared__ int s_ok;
    ++epoch;
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // this wave's stores have reached L2
    __syncthreads();
    if (threadIdx.x == 0) {
      if (FENCE) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      const bool flat = nwg <= (unsigned)kBarFlat;
      const unsigned ng = flat ? 1u : (unsigned)kBarGroups, g = flat ? 0u : idx % ng;
      const unsigned gsz = flat ? nwg : nwg / ng + (g < nwg % ng ? 1u : 0u);
      unsigned *top = ctr, *abort_ = ctr + 64;
      if (flat) {
        __hip_atomic_fetch_add(top, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        const unsigned got = __hip_atomic_fetch_add(ctr + 64 * (2 + g), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (got + 1 == epoch * gsz) __hip_atomic_fetch_add(top, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      const unsigned target = epoch * (flat ? nwg : ng);
      int ok = 1;
      for (unsigned spins = 0; ld_agent(top) < target; ++spins) {
        if ((spins & 63) == 63 && ld_agent(abort_) != 0) {
          ok = 0;
          break;
        }
        if (spins > kBarSpinLimit) {
          st_agent(abort_, 1u);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      if (FENCE) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
  }
};

// Workgroups the barrier kernels may launch: at most one per CU, and none when the kernel cannot
// hold one workgroup per CU (its occupancy is below one).
int barrier_grid_limit(const void *kernel, int threads, size_t lds);

// MFMA GEMM with the psvd_gemm contract plus an optional split-K workspace (gemm.hip).
int gemm_ws(int mode, int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha, const void *a, int64_t lda,
            int64_t stride_a, const void *b, int64_t ldb, int64_t stride_b, double beta, void *c, int64_t ldc,
            int64_t stride_c, int64_t batch, void *ws, size_t ws_bytes, hipStream_t st);
size_t gemm_splitk_bytes(int mode, int64_t m, int64_t n, int64_t batch);

}  // namespace psvd
