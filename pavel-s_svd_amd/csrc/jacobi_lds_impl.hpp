// K1 kernel templates: batched one-sided Jacobi SVD with each matrix resident in LDS (the kernels
// and their launch templates, shared by the FP32 and FP64 translation units).
//
// K1: batched one-sided Jacobi SVD with each matrix resident in LDS.
//
// One workgroup solves one matrix at a time and pulls the next matrix index from a device work
// queue (a persistent launch: grid = resident workgroups, not batch).  W (m x n, column-major,
// odd leading dimension) and V (n x n, column-major) live in LDS for the whole solve.  A round of
// the round-robin tournament has n/2 disjoint column pairs; each pair is owned by a group of
// `lanes` lanes that strides over the rows, so the three column dot products of a pair are a
// group butterfly reduction with wavefront shuffles.  Dot products are accumulated in FP64 (the
// products of FP32 values are exact in FP64), the rotation is computed in FP64 and applied in
// the storage type.  Sweep convergence is the reference test (primitives.py:332-352) evaluated
// on those FP64 sums.
//
// Follows svdbench/standalone/primitives.py:316-367 (jacobi_svd), :309-313 (schedule),
// :430-455 (complete_null_columns) and solvers.py:294-299 (max-abs scaling).
#pragma once

#include <float.h>

#include "common.hpp"

namespace psvd {

template <typename T>
struct LdsParams {
  const T *a;
  T *u, *s, *vh;
  int32_t *info;
  const int32_t *dims;
  const int64_t *offs;
  const int32_t *order;
  int32_t *queue;
  int64_t batch;
  int mu, nu, max_m, max_n;
  double tol;
  int max_sweeps, flags, lanes;
};

__host__ __device__ inline int pad_even(int n) { return n + (n & 1); }
// Column stride of W (m rows) and V (np rows) in LDS: the row count rounded up to a multiple of
// 16, plus one so that the stride is odd (columns of one round start in different banks).  Rows
// past the matrix's own are zero and stay zero under every rotation, so a group of L <= 16 lanes
// can stride over round_up(rows, L) rows with no bounds test.
__host__ __device__ inline int lds_ld(int rows) { return ((rows + 15) & ~15) | 1; }

template <typename T>
__host__ __device__ inline size_t lds_bytes(int m, int n, bool vectors) {
  const int np = pad_even(n), ldw = lds_ld(m), ldv = lds_ld(np);
  size_t b = sizeof(double) * (size_t)(3 * (np / 2) + 2 * np + m);
  b += sizeof(T) * ((size_t)np * ldw + (vectors ? (size_t)np * ldv : 0));
  b += sizeof(int) * (size_t)np;
  return (b + 15) & ~size_t(15);
}

template <int D>
__global__ void zero_queue_kernel(int32_t *q) {
  if (threadIdx.x == 0) q[0] = 0;
}

// Sweeps of one matrix with L lanes per column pair; the caller picks NT and L so that the
// round's pairs times L fill the workgroup.  One barrier per round: a group streams its pair's
// rows in unconditional chunks of four (row index clamped into the column, contributions
// masked), reduces the three FP64 dot products over its L lanes, decides, and applies the
// rotation to its W and V rows.  The round's rotation floor eps^2 * max_j |w_j|^2
// (primitives.py:337) is the maximum the previous round left behind: every wave publishes the
// maximum of the exact-arithmetic norm updates app - t|apq|, aqq + t|apq| of its pairs in its
// own slot of a double-buffered per-wave array (no atomics, no clearing) and every thread takes
// the maximum over the slots.  Convergence (no active pair in a sweep) is gathered the same way
// once per sweep.
template <int NW>
__device__ __forceinline__ double slots_max(const double *slot) {
  double v = slot[0];
#pragma unroll
  for (int w = 1; w < NW; ++w) v = fmax(v, slot[w]);
  return v;
}

constexpr int kChunk = 4;

template <typename T, int L>
__device__ __forceinline__ void pair_dots(const T *wp, const T *wq, int m, int lig, double &app, double &aqq,
                                          double &apq) {
  app = aqq = apq = 0.0;
  for (int base = 0; base < m; base += kChunk * L) {
    T xs[kChunk], ys[kChunk];
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      const int i = base + lig + j * L;
      const int ic = i < m ? i : m - 1;
      xs[j] = wp[ic];
      ys[j] = wq[ic];
    }
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      const bool live = base + lig + j * L < m;
      const double x = live ? (double)xs[j] : 0.0, y = live ? (double)ys[j] : 0.0;
      app = fma(x, x, app);
      aqq = fma(y, y, aqq);
      apq = fma(x, y, apq);
    }
  }
}

template <typename T, int L>
__device__ __forceinline__ void pair_rotate(T *xp, T *xq, int rows, int lig, T c, T s) {
  for (int base = 0; base < rows; base += kChunk * L) {
    T xs[kChunk], ys[kChunk];
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      const int i = base + lig + j * L;
      const int ic = i < rows ? i : rows - 1;
      xs[j] = xp[ic];
      ys[j] = xq[ic];
    }
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      const int i = base + lig + j * L;
      if (i < rows) {
        xp[i] = c * xs[j] - s * ys[j];
        xq[i] = s * xs[j] + c * ys[j];
      }
    }
  }
}

// Bank-spread order of a round's pairs.  A half-wave's LDS access is served by 32 banks at once;
// with L lanes per pair, its G = 32 / L groups read L consecutive words of their columns each.  The
// tournament gives consecutive pair indices columns p, p +- 1, ..., so with a column stride of 1 or 17
// (mod 32) consecutive groups would start one bank apart and their ranges overlap.  Each chunk of c <= 32
// pairs is therefore laid out column-major over h = ceil(c / G) rows: pair k = b h + a goes to slot
// a q + min(a, r) + b (q = c / h, r = c mod h), so the groups of one half-wave take pairs h apart, whose
// columns start about h banks apart (L apart in a full chunk).  Pairs of a round are disjoint, so the
// order in which the groups serve them is free: the results are bit-identical to the identity order.
template <int L>
__device__ __forceinline__ int pair_slot(int k, int npairs) {
  if constexpr (L > 16) {
    return k;
  } else {
    constexpr int G = 32 / L;
    const int c0 = k & ~31, kk = k - c0, c = npairs - c0 < 32 ? npairs - c0 : 32;
    const int h = (c + G - 1) / G, q = c / h, r = c - q * h;
    const int a = kk % h, b = kk / h;
    return c0 + a * q + (a < r ? a : r) + b;
  }
}

// inverse of pair_slot: the pair served by slot t
template <int L>
__device__ __forceinline__ int slot_pair(int t, int npairs) {
  if constexpr (L > 16) {
    return t;
  } else {
    constexpr int G = 32 / L;
    const int c0 = t & ~31, tt = t - c0, c = npairs - c0 < 32 ? npairs - c0 : 32;
    const int h = (c + G - 1) / G, q = c / h, r = c - q * h;
    int a, b;
    if (tt < r * (q + 1)) {
      a = tt / (q + 1);
      b = tt - a * (q + 1);
    } else {
      const int u = tt - r * (q + 1);
      a = r + u / q;
      b = u - (a - r) * q;
    }
    return c0 + b * h + a;
  }
}

// The pairs of round r: the reference's tournament (primitives.py:309-313) over column labels that
// are the columns themselves (the reference's ordering) or, with PSVD_ORDER_SORTED, the columns in
// descending order of their norms at the start of the sweep (perm[label] = column; de Rijk's
// ordering).  A pair is returned as (smaller column, larger column), as the tournament orders it.
__device__ __forceinline__ void round_pair(int r, int k, int np, const int *perm, int &p, int &q) {
  rr_pair(r, k, np, p, q);
  if (perm) {
    const int a = perm[p], b = perm[q];
    p = a < b ? a : b;
    q = a < b ? b : a;
  }
}

// perm[label] = column, labels in descending order of the columns' FP64 norms (ties by column), from
// the current W; cn: np doubles of scratch.  Called by the whole workgroup between rounds.
template <typename T, int NT>
__device__ void sort_labels(const T *W, int m, int np, int ldw, double *cn, int *perm) {
  constexpr int NW = NT / kWave;
  const int tid = threadIdx.x, wave = tid / kWave, lane = tid & (kWave - 1);
  for (int c = wave; c < np; c += NW) {
    const T *wc = W + c * ldw;
    double acc = 0.0;
    for (int i = lane; i < m; i += kWave) {
      const double x = (double)wc[i];
      acc = fma(x, x, acc);
    }
    acc = wave_sum(acc);
    if (lane == 0) cn[c] = acc;
  }
  __syncthreads();
  for (int j = tid; j < np; j += NT) {
    const double v = cn[j];
    int rank = 0;
    for (int i = 0; i < np; ++i) rank += cn[i] > v || (cn[i] == v && i < j);
    perm[rank] = j;
  }
  __syncthreads();
}

template <typename T> struct CachedRows { static constexpr int value = sizeof(T) == 4 ? 8 : 4; };

template <typename T, int NT, int L>
__device__ void jacobi_sweeps(T *W, T *V, int m, int np, int ldw, int ldv, bool vectors, double tol, double eps,
                              int max_sweeps, double *rslot, int *fslot, int flags, double *cn, int *perm,
                              int &sweeps_out, bool &conv_out) {
  constexpr int NG = NT / L, NW = NT / kWave;
  constexpr int RC = CachedRows<T>::value;       // rows per lane kept in registers between phases
  const int tid = threadIdx.x, wave = tid / kWave, lane = tid & (kWave - 1);
  const int grp = tid / L, lig = tid - grp * L;
  const int npairs = np / 2;
  const bool cached = m <= RC * L;
  // row slots a lane actually holds (uniform over the workgroup): slots past them are skipped, not
  // loaded and masked
  const int rc = (m + L - 1) / L, rcv = (np + L - 1) / L;
  const bool vcached = np <= RC * L;
  const double tol2 = tol * tol, eps2 = eps * eps;
  int sweeps = 0;
  bool converged = false;
  int g = 0;
  const bool spread = flags & PSVD_SLOT_SPREAD, sorted = flags & PSVD_ORDER_SORTED;
  for (int sw = 1; sw <= max_sweeps; ++sw) {
    sweeps = sw;
    bool wave_active = false;
    if (sorted) sort_labels<T, NT>(W, m, np, ldw, cn, perm);
    for (int r = 0; r < np - 1; ++r, ++g) {
      const double floor_ = eps2 * slots_max<NW>(rslot + (g & 1) * NW);
      double mymax = 0.0;
      for (int k = grp; k < npairs; k += NG) {
        int p, q;
        round_pair(r, spread ? slot_pair<L>(k, npairs) : k, np, sorted ? perm : nullptr, p, q);
        T *wp = W + p * ldw, *wq = W + q * ldw;
        double app = 0.0, aqq = 0.0, apq = 0.0;
        T xs[RC], ys[RC];
        if (cached) {
          // all rows of this lane in registers: unconditional loads (index clamped), masked sums
#pragma unroll
          for (int j = 0; j < RC; ++j) {
            if (j < rc) {
              const int i = lig + j * L;
              const int ic = i < m ? i : m - 1;
              xs[j] = wp[ic];
              ys[j] = wq[ic];
            }
          }
#pragma unroll
          for (int j = 0; j < RC; ++j) {
            if (j < rc) {
              const bool live = lig + j * L < m;
              const double x = live ? (double)xs[j] : 0.0, y = live ? (double)ys[j] : 0.0;
              app = fma(x, x, app);
              aqq = fma(y, y, aqq);
              apq = fma(x, y, apq);
            }
          }
        } else {
          pair_dots<T, L>(wp, wq, m, lig, app, aqq, apq);
        }
        app = group_sum_dpp<L>(app);
        aqq = group_sum_dpp<L>(aqq);
        apq = group_sum_dpp<L>(apq);
        double nap = app, naq = aqq;
        // active: |apq| / (|w_p| |w_q|) > tol and min(app, aqq) > floor (primitives.py:335-338)
        if (apq * apq > tol2 * app * aqq && fmin(app, aqq) > floor_) {
          T c, s;
          double tb;
          jacobi_rotation_t<T>(app, aqq, apq, c, s, tb);
          nap = app - tb;
          naq = aqq + tb;
          wave_active = true;
          if (cached) {
#pragma unroll
            for (int j = 0; j < RC; ++j) {
              const int i = lig + j * L;
              if (j < rc && i < m) {
                wp[i] = c * xs[j] - s * ys[j];
                wq[i] = s * xs[j] + c * ys[j];
              }
            }
          } else {
            pair_rotate<T, L>(wp, wq, m, lig, c, s);
          }
          if (vectors) {
            T *vp = V + p * ldv, *vq = V + q * ldv;
            if (vcached) {
              for (int j = 0; j < rcv; ++j) {   // uniform trip count: the rows this lane holds
                const int i = lig + j * L;
                if (i < np) {
                  const T x = vp[i], y = vq[i];
                  vp[i] = c * x - s * y;
                  vq[i] = s * x + c * y;
                }
              }
            } else {
              pair_rotate<T, L>(vp, vq, np, lig, c, s);
            }
          }
        }
        mymax = fmax(mymax, fmax(nap, naq));
      }
      mymax = wave_max_fast(mymax);
      if (lane == 0) rslot[((g + 1) & 1) * NW + wave] = mymax;
      __syncthreads();
    }
    // convergence: no wave had an active pair in this sweep
    const int wave_any = __any(wave_active) ? 1 : 0;
    if (lane == 0) fslot[(sw & 1) * NW + wave] = wave_any;
    __syncthreads();
    int any = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) any |= fslot[(sw & 1) * NW + w];
    if (!wave_uniform(any)) {
      converged = true;
      break;
    }
  }
  sweeps_out = sweeps;
  conv_out = converged;
}

// The same sweeps for L <= 16 lanes per pair and exactly S row slots per lane (S = ceil(m / L)):
// with the zero-padded columns every slot is a real or a zero row, so the loads, the FP64 sums,
// the rotation and the stores run unconditionally on registers (no clamped indices, no masks and
// no selects), and the V rotation strides over round_up(np, L) rows.  The sums are formed in the
// same order as in jacobi_sweeps, so both give the same bits.
template <typename T, int NT, int L, int S>
__device__ void jacobi_sweeps_fixed(T *W, T *V, int m, int np, int ldw, int ldv, bool vectors, double tol,
                                    double eps, int max_sweeps, double *rslot, int *fslot, int *ptab, int flags,
                                    double *cn, int *perm, int &sweeps_out, bool &conv_out) {
  static_assert(L <= 16, "the zero padding covers groups of at most 16 lanes");
  constexpr int NG = NT / L, NW = NT / kWave;
  const int tid = threadIdx.x, wave = tid / kWave, lane = tid & (kWave - 1);
  const int grp = tid / L, lig = tid - grp * L;
  const int npairs = np / 2;
  const int rcv = (np + L - 1) / L;
  const double tol2 = tol * tol, eps2 = eps * eps;
  // the round's pairs come from a double-buffered table, ptab[slot * npairs + k] = p | q << 16:
  // threads k < npairs write the next round's entries during a round and its closing barrier
  // publishes them, so a pass reads one word instead of evaluating the tournament closed form
  const bool spread = flags & PSVD_SLOT_SPREAD, sorted = flags & PSVD_ORDER_SORTED;
  const int slot0 = spread && tid < npairs ? pair_slot<L>(tid, npairs) : tid;   // table slot of k = tid
  if (!sorted) {
    for (int k = tid; k < npairs; k += NT) {
      int p, q;
      rr_pair(0, k, np, p, q);
      ptab[k == tid ? slot0 : (spread ? pair_slot<L>(k, npairs) : k)] = p | (q << 16);
    }
    __syncthreads();
  }
  int sweeps = 0;
  bool converged = false;
  int g = 0;
  for (int sw = 1; sw <= max_sweeps; ++sw) {
    sweeps = sw;
    bool wave_active = false;
    if (sorted) {   // the sweep's labels, then its first round's table
      sort_labels<T, NT>(W, m, np, ldw, cn, perm);
      for (int k = tid; k < npairs; k += NT) {
        int p, q;
        round_pair(0, k, np, perm, p, q);
        ptab[(g & 1) * npairs + (k == tid ? slot0 : (spread ? pair_slot<L>(k, npairs) : k))] = p | (q << 16);
      }
      __syncthreads();
    }
    for (int r = 0; r < np - 1; ++r, ++g) {
      const double floor_ = eps2 * slots_max<NW>(rslot + (g & 1) * NW);
      const int *tab = ptab + (g & 1) * npairs;
      double mymax = 0.0;
      for (int k = grp; k < npairs; k += NG) {
        const int pq = tab[k];
        const int p = pq & 0xffff, q = pq >> 16;
        // column offsets stay far below 2^24: 24-bit multiplies (full rate, unlike v_mul_lo_u32)
        T *wp = W + __umul24(p, ldw) + lig, *wq = W + __umul24(q, ldw) + lig;
        T xs[S], ys[S];
#pragma unroll
        for (int j = 0; j < S; ++j) {
          xs[j] = wp[j * L];
          ys[j] = wq[j * L];
        }
        double app = 0.0, aqq = 0.0, apq = 0.0;
#pragma unroll
        for (int j = 0; j < S; ++j) {
          const double x = (double)xs[j], y = (double)ys[j];
          app = fma(x, x, app);
          aqq = fma(y, y, aqq);
          apq = fma(x, y, apq);
        }
        group_sum3_dpp<L>(app, aqq, apq);
        double nap = app, naq = aqq;
        // active: |apq| / (|w_p| |w_q|) > tol and min(app, aqq) > floor (primitives.py:335-338)
        if (apq * apq > tol2 * app * aqq && fmin(app, aqq) > floor_) {
          T c, s;
          double tb;
          jacobi_rotation_t<T>(app, aqq, apq, c, s, tb);
          nap = app - tb;
          naq = aqq + tb;
          wave_active = true;
#pragma unroll
          for (int j = 0; j < S; ++j) {
            wp[j * L] = c * xs[j] - s * ys[j];
            wq[j * L] = s * xs[j] + c * ys[j];
          }
          if (vectors) {
            // four slots per step, loads before stores, so one LDS round trip serves four rows
            T *vp = V + __umul24(p, ldv) + lig, *vq = V + __umul24(q, ldv) + lig;
            int j = 0;
            for (; j + 4 <= rcv; j += 4, vp += 4 * L, vq += 4 * L) {
              T x[4], y[4];
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                x[u] = vp[u * L];
                y[u] = vq[u * L];
              }
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                vp[u * L] = c * x[u] - s * y[u];
                vq[u * L] = s * x[u] + c * y[u];
              }
            }
            for (; j < rcv; ++j, vp += L, vq += L) {
              const T x = vp[0], y = vq[0];
              vp[0] = c * x - s * y;
              vq[0] = s * x + c * y;
            }
          }
        }
        mymax = fmax(mymax, fmax(nap, naq));
      }
      if (r + 1 < np - 1 || !sorted) {  // the next round's pairs (round 0 of the next sweep after the
                                         // last, unless the next sweep relabels the columns first)
        const int rn = r + 1 < np - 1 ? r + 1 : 0;
        for (int k = tid; k < npairs; k += NT) {
          int p, q;
          round_pair(rn, k, np, sorted ? perm : nullptr, p, q);
          ptab[((g + 1) & 1) * npairs + (k == tid ? slot0 : (spread ? pair_slot<L>(k, npairs) : k))] =
              p | (q << 16);
        }
      }
      mymax = wave_max_fast(mymax);
      if (lane == 0) rslot[((g + 1) & 1) * NW + wave] = mymax;
      __syncthreads();
    }
    // convergence: no wave had an active pair in this sweep
    const int wave_any = __any(wave_active) ? 1 : 0;
    if (lane == 0) fslot[(sw & 1) * NW + wave] = wave_any;
    __syncthreads();
    int any = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) any |= fslot[(sw & 1) * NW + w];
    if (!wave_uniform(any)) {
      converged = true;
      break;
    }
  }
  sweeps_out = sweeps;
  conv_out = converged;
}

// Picks the fixed-slot sweeps when the lane count and the matrix's rows allow it, else the
// general ones.  m is uniform over the workgroup, so the switch is a scalar branch.
template <typename T, int NT, int L>
__device__ void run_sweeps(T *W, T *V, int m, int np, int ldw, int ldv, bool vectors, double tol, double eps,
                           int max_sweeps, double *rslot, int *fslot, int *ptab, int flags, double *cn, int *perm,
                           int &sweeps_out, bool &conv_out) {
  if constexpr (L <= 16) {
    constexpr int RC = sizeof(T) == 4 ? 12 : 4;   // row slots a lane can keep in registers
    const int rc = (m + L - 1) / L;
#define PSVD_FIXED_CASE(SV)                                                                                  \
  if constexpr (SV <= RC)                                                                                    \
    if (rc == SV)                                                                                            \
      return jacobi_sweeps_fixed<T, NT, L, SV>(W, V, m, np, ldw, ldv, vectors, tol, eps, max_sweeps, rslot,    \
                                               fslot, ptab, flags, cn, perm, sweeps_out, conv_out);
    PSVD_FIXED_CASE(1) PSVD_FIXED_CASE(2) PSVD_FIXED_CASE(3) PSVD_FIXED_CASE(4)
    PSVD_FIXED_CASE(5) PSVD_FIXED_CASE(6) PSVD_FIXED_CASE(7) PSVD_FIXED_CASE(8)
    PSVD_FIXED_CASE(9) PSVD_FIXED_CASE(10) PSVD_FIXED_CASE(11) PSVD_FIXED_CASE(12)
#undef PSVD_FIXED_CASE
  }
  jacobi_sweeps<T, NT, L>(W, V, m, np, ldw, ldv, vectors, tol, eps, max_sweeps, rslot, fslot, flags, cn, perm,
                          sweeps_out, conv_out);
}

// Waves per SIMD the register budget is sized for (128 VGPRs: four waves per SIMD).
template <int NT> struct WavesPerEU { static constexpr int value = 4; };

template <typename T, int NT, int L>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WavesPerEU<NT>::value)))
void jacobi_lds_kernel(LdsParams<T> P) {
  extern __shared__ double smem_d[];
  __shared__ int s_idx;
  __shared__ int s_fslot[2 * (NT / kWave)];
  __shared__ double s_rslot[2 * (NT / kWave)];
  __shared__ int s_nk;
  __shared__ double s_red[NT / kWave];
  __shared__ int s_redi[NT / kWave];

  constexpr int NW = NT / kWave;
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1), wave = tid / kWave;
  const bool vectors = !(P.flags & PSVD_NO_VECTORS);
  const double eps = Eps<T>::value;
  const double tol = P.tol > 0.0 ? P.tol : 10.0 * eps;

  for (;;) {
    if (tid == 0) s_idx = atomicAdd(P.queue, 1);
    __syncthreads();
    const int64_t idx = wave_uniform(s_idx);
    __syncthreads();
    if (idx >= P.batch) return;
    const int64_t mi = P.order ? (int64_t)P.order[idx] : idx;
    int m, n;
    int64_t aoff, soff, voff;
    if (P.dims) {
      m = P.dims[2 * mi];
      n = P.dims[2 * mi + 1];
      aoff = P.offs[3 * mi];
      soff = P.offs[3 * mi + 1];
      voff = P.offs[3 * mi + 2];
    } else {
      m = P.mu;
      n = P.nu;
      aoff = mi * (int64_t)m * n;
      soff = mi * (int64_t)n;
      voff = mi * (int64_t)n * n;
    }
    if (m < 1 || n < 1 || n > m || m > P.max_m || n > P.max_n) {
      if (tid == 0) {
        P.info[2 * mi] = 0;
        P.info[2 * mi + 1] = PSVD_INFO_BADSHAPE;
      }
      continue;  // uniform across the block: every thread read the same descriptor
    }
    const int np = pad_even(n), ldw = lds_ld(m), ldv = lds_ld(np), npairs = np / 2;
    int *ptab = reinterpret_cast<int *>(smem_d);   // 2 npairs pair words (room for 6 npairs)
    double *sig = smem_d + 3 * npairs;
    double *ytmp = sig + np;
    double *vt = ytmp + np;
    T *W = reinterpret_cast<T *>(vt + m);
    T *V = W + (size_t)np * ldw;
    int *ord = reinterpret_cast<int *>(V + (vectors ? (size_t)np * ldv : 0));

    // ---- load A (row-major) into W (column-major), check finiteness, optional max-abs scale
    const T *A = P.a + aoff;
    double amax = 0.0, bad = 0.0;
    for (int e = tid; e < m * n; e += NT) {
      const int i = e / n, j = e - i * n;
      const T x = A[e];
      W[j * ldw + i] = x;
      const double ax = fabs((double)x);
      if (ax <= DBL_MAX) amax = fmax(amax, ax); else bad = 1.0;
    }
    {  // zero rows past m in every column, and the padding column when n is odd
      const int padr = ldw - m;
      for (int e = tid; e < np * padr; e += NT) {
        const int j = e / padr;
        W[j * ldw + m + (e - j * padr)] = T(0);
      }
    }
    if (np > n)
      for (int i = tid; i < m; i += NT) W[n * ldw + i] = T(0);
    if (vectors)
      for (int e = tid; e < np * ldv; e += NT) {
        const int j = e / ldv;
        V[e] = (e - j * ldv == j) ? T(1) : T(0);
      }
    if (block_max<NT>(bad, s_red) > 0.0) {
      if (tid == 0) {
        P.info[2 * mi] = 0;
        P.info[2 * mi + 1] = PSVD_INFO_NONFINITE;
      }
      __syncthreads();
      continue;
    }
    double scale = 1.0;
    if (P.flags & PSVD_SCALE) {
      const double mx = block_max<NT>(amax, s_red);
      if (mx > 0.0) {
        scale = mx;
        const T st = (T)mx;  // exact: the maximum is one of the stored values
        for (int e = tid; e < m * n; e += NT) {
          const int j = e / m, i = e - j * m;
          W[j * ldw + i] = W[j * ldw + i] / st;
        }
      }
    }
    __syncthreads();
    // initial rotation floor: max_j |w_j|^2, one slot per wave
    {
      double mx = 0.0;
      for (int c = wave; c < n; c += NW) {
        const T *wc = W + c * ldw;
        double acc = 0.0;
        for (int i = lane; i < m; i += kWave) {
          const double x = (double)wc[i];
          acc = fma(x, x, acc);
        }
        mx = fmax(mx, wave_sum(acc));
      }
      if (lane == 0) s_rslot[wave] = mx;
    }
    __syncthreads();

    // ---- sweeps
    int sweeps = 0;
    bool converged = false;
    // sig and ord are free until the sweeps end: the sorted ordering's norms and labels live there
    run_sweeps<T, NT, L>(W, V, m, np, ldw, ldv, vectors, tol, eps, P.max_sweeps, s_rslot, s_fslot, ptab, P.flags,
                         sig, ord, sweeps,
                         converged);
    int status = converged ? PSVD_INFO_OK : PSVD_INFO_NOCONV;

    // ---- sigma in FP64, descending order, cut
    for (int c = wave; c < n; c += NW) {
      const T *wc = W + c * ldw;
      double acc = 0.0;
      for (int i = lane; i < m; i += kWave) {
        const double x = (double)wc[i];
        acc = fma(x, x, acc);
      }
      acc = wave_sum(acc);
      if (lane == 0) sig[c] = sqrt(acc);
    }
    __syncthreads();
    for (int j = tid; j < n; j += NT) {
      const double sj = sig[j];
      int rank = 0;
      for (int k = 0; k < n; ++k) {
        const double sk = sig[k];
        rank += (sk > sj) || (sk == sj && k < j);
      }
      ord[rank] = j;
    }
    __syncthreads();
    const double s1 = sig[ord[0]];
    const double cut = (P.flags & PSVD_CUT_FLOOR) ? eps * s1 : eps * (double)(m > np ? m : np) * s1;
    if (tid == 0) {
      int nk = 0;
      while (nk < n && sig[ord[nk]] > cut) ++nk;
      s_nk = nk;
    }
    __syncthreads();
    const int nk = wave_uniform(s_nk);

    // ---- U = W / sigma for kept columns; coordinate-vector completion for the rest
    if (vectors) {
      for (int e = tid; e < n * m; e += NT) {
        const int r = e / m, i = e - r * m;
        const int c = ord[r];
        T *wc = W + c * ldw;
        wc[i] = r < nk ? (T)((double)wc[i] / sig[c]) : T(0);
      }
      __syncthreads();
      for (int r = nk; r < n; ++r) {
        double best = -DBL_MAX;
        int bi = 0x7fffffff;
        for (int i = tid; i < m; i += NT) {
          double acc = 0.0;
          for (int c = 0; c < n; ++c) {
            const double x = (double)W[c * ldw + i];
            acc = fma(x, x, acc);
          }
          const double lev = 1.0 - acc;
          if (lev > best) {
            best = lev;
            bi = i;
          }
        }
        for (int o = kWave >> 1; o > 0; o >>= 1) {
          const double ob = __shfl_xor(best, o, kWave);
          const int oi = __shfl_xor(bi, o, kWave);
          if (ob > best || (ob == best && oi < bi)) {
            best = ob;
            bi = oi;
          }
        }
        __syncthreads();
        if (lane == 0) {
          s_red[wave] = best;
          s_redi[wave] = bi;
        }
        __syncthreads();
        best = s_red[0];
        bi = s_redi[0];
        for (int w = 1; w < NW; ++w)
          if (s_red[w] > best || (s_red[w] == best && s_redi[w] < bi)) {
            best = s_red[w];
            bi = s_redi[w];
          }
        const int ix = bi;
        // v = e_ix - U (U^T e_ix)
        for (int i = tid; i < m; i += NT) {
          double acc = (i == ix) ? 1.0 : 0.0;
          for (int c = 0; c < n; ++c) acc -= (double)W[c * ldw + i] * (double)W[c * ldw + ix];
          vt[i] = acc;
        }
        __syncthreads();
        // v = v - U (U^T v)
        for (int c = wave; c < n; c += NW) {
          double acc = 0.0;
          for (int i = lane; i < m; i += kWave) acc = fma((double)W[c * ldw + i], vt[i], acc);
          acc = wave_sum(acc);
          if (lane == 0) ytmp[c] = acc;
        }
        __syncthreads();
        double l2 = 0.0;
        for (int i = tid; i < m; i += NT) {
          double acc = vt[i];
          for (int c = 0; c < n; ++c) acc -= (double)W[c * ldw + i] * ytmp[c];
          vt[i] = acc;
          l2 = fma(acc, acc, l2);
        }
        const double len = sqrt(block_sum<NT>(l2, s_red));
        if (len < 10.0 * eps) status = PSVD_INFO_COMPLETION;
        const double inv = 1.0 / fmax(len, DBL_MIN);
        T *wc = W + ord[r] * ldw;
        for (int i = tid; i < m; i += NT) wc[i] = (T)(vt[i] * inv);
        __syncthreads();
      }
    }

    // ---- outputs
    T *S = P.s + soff;
    for (int r = tid; r < n; r += NT) S[r] = r < nk ? (T)(sig[ord[r]] * scale) : T(0);
    if (vectors) {
      T *U = P.u + aoff;
      for (int e = tid; e < m * n; e += NT) {
        const int i = e / n, r = e - i * n;
        U[e] = W[ord[r] * ldw + i];
      }
      T *VH = P.vh + voff;
      for (int e = tid; e < n * n; e += NT) {
        const int r = e / n, j = e - r * n;
        VH[e] = V[ord[r] * ldv + j];
      }
    }
    if (tid == 0) {
      P.info[2 * mi] = sweeps;
      P.info[2 * mi + 1] = status;
    }
    __syncthreads();
  }
}

constexpr size_t kLdsCap = 160 * 1024;

struct DeviceProps {
  int cus = 0;
};

static DeviceProps &props() {
  static DeviceProps p;
  if (p.cus == 0) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&p.cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (p.cus <= 0) p.cus = 256;
  }
  return p;
}

template <typename T, int NT, int L>
static int launch_lds(const LdsParams<T> &P, size_t smem, hipStream_t stream) {
  static bool attr = false;
  auto kern = jacobi_lds_kernel<T, NT, L>;
  if (!attr) {
    if (hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(kLdsCap - 4096)) != hipSuccess)
      return PSVD_ELAUNCH;
    attr = true;
  }
  // occupancy per dynamic-LDS size, queried once outside any stream capture and then reused
  static size_t last_smem = ~size_t(0);
  static int occ = 1;
  if (smem != last_smem) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, NT, smem) != hipSuccess || occ < 1) occ = 1;
    last_smem = smem;
  }
  int64_t grid = (int64_t)props().cus * occ;
  if (grid > P.batch) grid = P.batch;
  hipLaunchKernelGGL(zero_queue_kernel<0>, dim3(1), dim3(64), 0, stream, P.queue);
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(NT), smem, stream, P);
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

template <typename T>
static int run_lds(int64_t batch, int mu, int nu, const int32_t *dims, const int64_t *offs,
                   const int32_t *order, int max_m, int max_n, const void *a, void *u, void *s,
                   void *vh, int32_t *info, double tol, int max_sweeps, int flags, int lanes,
                   int threads, void *ws, size_t ws_bytes, hipStream_t stream) {
  const bool vectors = !(flags & PSVD_NO_VECTORS);
  const size_t smem = lds_bytes<T>(max_m, max_n, vectors);
  if (smem > kLdsCap - 4096) return PSVD_ETOOBIG;
  if (lanes > kWave || (lanes > 0 && (lanes & (lanes - 1))) || lanes < -kWave) return PSVD_EINVAL;
  if (lanes <= 0) {  // about -lanes rows per lane (4 when 0) for the largest matrix of the call
    const int rpl = lanes < 0 ? -lanes : 4;
    const int target = (max_m + rpl - 1) / rpl;
    lanes = 2;
    while (lanes < target && lanes < kWave) lanes <<= 1;
  }
  if (lanes == 1) lanes = 2;
  LdsParams<T> P;
  P.a = static_cast<const T *>(a);
  P.u = static_cast<T *>(u);
  P.s = static_cast<T *>(s);
  P.vh = static_cast<T *>(vh);
  P.info = info;
  P.dims = dims;
  P.offs = offs;
  P.order = order;
  P.queue = static_cast<int32_t *>(ws);
  P.batch = batch;
  P.mu = mu;
  P.nu = nu;
  P.max_m = max_m;
  P.max_n = max_n;
  P.tol = tol;
  P.max_sweeps = max_sweeps;
  P.flags = flags;
  P.lanes = lanes;
  (void)ws_bytes;
  if (threads == 0) threads = 256;
#define PSVD_LDS_CASE(NTV, LV) \
  if (threads == NTV && lanes == LV) return launch_lds<T, NTV, LV>(P, smem, stream);
#define PSVD_LDS_NT(NTV) \
  PSVD_LDS_CASE(NTV, 2) PSVD_LDS_CASE(NTV, 4) PSVD_LDS_CASE(NTV, 8) PSVD_LDS_CASE(NTV, 16) \
  PSVD_LDS_CASE(NTV, 32) PSVD_LDS_CASE(NTV, 64)
  PSVD_LDS_NT(64) PSVD_LDS_NT(128) PSVD_LDS_NT(256) PSVD_LDS_NT(512) PSVD_LDS_NT(1024)
#undef PSVD_LDS_NT
#undef PSVD_LDS_CASE
  return PSVD_EINVAL;
}

}  // namespace psvd
