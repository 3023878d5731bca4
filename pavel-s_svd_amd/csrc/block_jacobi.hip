// Block-pair one-sided Jacobi (svdbench/standalone/solvers.py:30-68) on gfx950.
//
// The n columns of W (m x n, row-major) form groups of `width` consecutive columns (the last one
// may be short; an odd group count gets an empty partner group), paired by the round-robin
// tournament of primitives.py:309-313.  The reference solves each pair with a full one-sided
// Jacobi SVD of its k <= 2 width columns and replaces them by W_pair V_pair, V_pair sorted by
// descending sigma (solvers.py:43-50); V gets the same update.  Here one workgroup takes one
// pair of one matrix per round launch, and everything after the load stays on chip:
//   1. G = W_pair^T W_pair in FP64 (FP32 products are exact in FP64), replicated row slices summed
//      in a fixed order;
//   2. the one-sided Jacobi of W_pair carried out on G: in round rr of the tournament over the k
//      columns (padded to even with a zero column, as primitives.py:320-323), pair (p, q) rotates
//      iff |g_pq| / (sqrt g_pp sqrt g_qq) > tol and min(g_pp, g_qq) > eps^2 max_c g_cc
//      (primitives.py:331-338); the rotation (primitives.py:340-347) is applied two-sidedly,
//      G <- J^T G J, and accumulated in Q <- Q J, all in FP64 in LDS; sweeps end when no active
//      pair exceeds tol (:352), failing after max_sweeps (:354-355);
//   3. the columns of Q are ordered by descending g_cc (ties by index; primitives.py:359-361);
//   4. W_pair <- W_pair Q and V_pair <- V_pair Q, staged through LDS in row chunks, FP64 sums.
// A rotation thus costs O(k) on the Gram matrix instead of O(m) on the columns, and W is read
// twice and written once per pair and round.
//
// The convergence test of a sweep (solvers.py:51-60) follows in two kernels: FP64 column norms
// (with the per-matrix maximum) and the maximum of |w_i^T w_j| / (|w_i| |w_j|) over live column
// pairs, live meaning |w_i|^2 > (eps max(m, n))^2 max_c |w_c|^2, from 16 x 16 column tiles with
// FP64 sums (the reference forms W^H W in the data precision).  Per-matrix maxima are combined
// with device atomics on the bit patterns of non-negative doubles after a fill kernel clears them.
#include <float.h>

#include "common.hpp"

namespace psvd {

constexpr int kBjrThreads = 256;
constexpr int kBjrMaxK = 64;      // columns of one pair: 2 * block_size
constexpr int kBjrChunk = 4096;   // doubles of the row-chunk staging buffer
constexpr int kBjrMaxEpt = (kBjrMaxK * (kBjrMaxK + 1) / 2 + kBjrThreads - 1) / kBjrThreads;   // Gram entries per thread
constexpr int kCorrTile = 16;
constexpr int kCorrRows = 64;
constexpr int kNormCols = 16;     // columns per norms workgroup (16 row groups of 16 columns)

__device__ __forceinline__ int bjr_group_cols(int g, int width, int n, int groups) {
  if (g >= groups) return 0;
  const int c = n - g * width;
  return c < width ? c : width;
}

// entry e of the upper triangle (p <= q), row by row: row p holds k - p entries
__device__ __forceinline__ void bjr_unrank(int e, int k, int &p, int &q) {
  p = 0;
  while (e >= k - p) {
    e -= k - p;
    ++p;
  }
  q = p + e;
}

static size_t bjr_lds_bytes(int kmax) {
  const size_t doubles = 2 * (size_t)kmax * (kmax + 1) + kBjrChunk + kBjrThreads + kBjrMaxK + 8;
  const size_t ints = 2 * (kBjrMaxK / 2) + kBjrMaxK + 2;
  return doubles * 8 + ints * 4;
}

// Stage rows [r0, r0 + nr) of the pair's k columns of M into X (row-major, k per row), as FP64.
template <typename T, typename COL>
__device__ __forceinline__ void bjr_stage(double *X, const T *M, int64_t ldm, int r0, int nr, int k, COL col) {
  for (int e = threadIdx.x; e < nr * k; e += kBjrThreads) {
    const int r = e / k, c = e - r * k;
    X[e] = (double)M[(int64_t)(r0 + r) * ldm + col(c)];
  }
}

// Ordering point for LDS traffic among the lanes of one wavefront (its instructions issue in
// order; the fences keep the compiler from moving LDS accesses across).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One-sided Jacobi of a pair's k columns carried out on their Gram matrix G (k x k, row stride
// ld, FP64, LDS), rotations accumulated into Q.  Round rr of the tournament over the k columns
// (padded to even; the padding column is zero and never active) rotates pair (p, q) iff
// |g_pq| > tol sqrt(g_pp) sqrt(g_qq) and min(g_pp, g_qq) > eps^2 max_c g_cc -- the reference's
// |a_pq| / (sqrt a_pp sqrt a_qq) > tol test without the division (primitives.py:331-338) -- and
// applies G <- J^T G J by disjoint 2 x 2 blocks (rows of pair i x columns of pair j) and Q <- Q J
// in one pass.  A sweep with no active pair ends the solve (the reference's off <= tol, :352).
// WAVE: run by wave 0 alone (np2^2 <= 64 blocks), else by the whole workgroup.  Leaves the sweep
// count in flag[1] and convergence in flag[0].
template <typename T, bool WAVE>
__device__ __forceinline__ void bjr_inner(double *G, double *Q, int k, int ld, double tol, int max_sweeps,
                                          double *rc, double *rs, int *rp, int *rq, double *act, int *flag) {
  constexpr int NTH = WAVE ? kWave : kBjrThreads;
  const int tid = threadIdx.x, wave = tid / kWave, lane = tid & (kWave - 1);
  const int kp = k + (k & 1), np2 = kp / 2;
  const double eps2 = Eps<T>::value * Eps<T>::value;
  const double tol2 = tol * tol;
  int sweeps = 0;
  bool converged = false;
  for (int sweep = 1; sweep <= max_sweeps; ++sweep) {
    bool any = false;   // per lane of wave 0
    for (int rr = 0; rr < kp - 1; ++rr) {
      if (wave == 0) {
        const double dmax = wave_max_fast(lane < k ? G[lane * ld + lane] : 0.0);
        if (lane < np2) {
          int p, q;
          rr_pair(rr, lane, kp, p, q);
          double c = 1.0, s = 0.0;
          bool active = false;
          if (q < k) {
            const double app = G[p * ld + p], aqq = G[q * ld + q], apq = G[p * ld + q];
            // squares for FP32 data (Gram entries far from FP64 overflow), square roots for FP64
            const bool big = sizeof(T) == 4 ? apq * apq > tol2 * (app * aqq)
                                            : fabs(apq) > tol * (sqrt(app) * sqrt(aqq));
            active = big && fmin(app, aqq) > eps2 * dmax;
            if (active) jacobi_rotation_fast<T>(app, aqq, apq, c, s);
          }
          any = any || active;
          rp[lane] = p;
          rq[lane] = q;
          rc[lane] = c;
          rs[lane] = s;
          act[lane] = active ? 1.0 : 0.0;
        }
      }
      if constexpr (WAVE) wave_sync(); else __syncthreads();
      for (int it = tid; it < np2 * np2; it += NTH) {
        const int i = it / np2, j = it - i * np2;
        const bool ai = act[i] != 0.0, aj = act[j] != 0.0;
        if (!ai && !aj) continue;
        const int p1 = rp[i], q1 = rq[i], p2 = rp[j], q2 = rq[j];
        const bool vq1 = q1 < k, vq2 = q2 < k;   // the padding index has no row or column
        double b11 = G[p1 * ld + p2];
        double b12 = vq2 ? G[p1 * ld + q2] : 0.0;
        double b21 = vq1 ? G[q1 * ld + p2] : 0.0;
        double b22 = vq1 && vq2 ? G[q1 * ld + q2] : 0.0;
        if (aj) {
          const double c = rc[j], s = rs[j];
          const double t11 = c * b11 - s * b12, t12 = s * b11 + c * b12;
          const double t21 = c * b21 - s * b22, t22 = s * b21 + c * b22;
          b11 = t11; b12 = t12; b21 = t21; b22 = t22;
        }
        if (ai) {
          const double c = rc[i], s = rs[i];
          const double t11 = c * b11 - s * b21, t21 = s * b11 + c * b21;
          const double t12 = c * b12 - s * b22, t22 = s * b12 + c * b22;
          b11 = t11; b12 = t12; b21 = t21; b22 = t22;
        }
        G[p1 * ld + p2] = b11;
        if (vq2) G[p1 * ld + q2] = b12;
        if (vq1) G[q1 * ld + p2] = b21;
        if (vq1 && vq2) G[q1 * ld + q2] = b22;
      }
      for (int it = tid; it < k * np2; it += NTH) {
        const int r = it / np2, j = it - r * np2;
        if (act[j] == 0.0) continue;
        const double c = rc[j], s = rs[j];
        const int p = rp[j], q = rq[j];
        const double qp = Q[r * ld + p], qq = Q[r * ld + q];
        Q[r * ld + p] = c * qp - s * qq;
        Q[r * ld + q] = s * qp + c * qq;
      }
      if constexpr (WAVE) wave_sync(); else __syncthreads();
    }
    sweeps = sweep;
    if (wave == 0) {
      const bool none = !__any(any);
      if (lane == 0) flag[0] = none;
    }
    if constexpr (WAVE) wave_sync(); else __syncthreads();
    converged = flag[0] != 0;
    if constexpr (WAVE) wave_sync(); else __syncthreads();
    if (converged) break;
  }
  if (tid == 0) flag[1] = sweeps;
}

template <typename T>
__global__ __launch_bounds__(kBjrThreads) void bjr_round_kernel(T *W, int64_t ldw, int64_t sw, T *V, int64_t ldv,
                                                                int64_t sv, int m, int n, int width, int groups,
                                                                int gp, int round, double tol, int max_sweeps,
                                                                int32_t *info) {
  extern __shared__ double lds[];
  constexpr int NT = kBjrThreads;
  const int npairs = gp / 2;
  const int64_t b = blockIdx.x / npairs;
  const int pair = (int)(blockIdx.x - b * npairs);
  int gi, gj;
  rr_pair(round, pair, gp, gi, gj);
  const int ki = bjr_group_cols(gi, width, n, groups), kj = bjr_group_cols(gj, width, n, groups);
  const int k = ki + kj;
  if (k < 2) return;   // a lone single column: nothing to rotate (solvers.py:46-47)
  const int ci0 = gi * width, cj0 = gj * width;
  auto col = [=](int c) { return c < ki ? ci0 + c : cj0 + (c - ki); };
  const int ld = k + 1;
  double *G = lds;
  double *Q = G + k * ld;
  double *X = Q + k * ld;
  double *part = X + kBjrChunk;
  double *rc = part + NT;
  double *rs = rc + kBjrMaxK / 2;
  int *rp = reinterpret_cast<int *>(rs + kBjrMaxK / 2 + 8);
  int *rq = rp + kBjrMaxK / 2;
  int *perm = rq + kBjrMaxK / 2;
  int *flag = perm + kBjrMaxK;
  const int tid = threadIdx.x, wave = tid / kWave, lane = tid & (kWave - 1);
  T *Wb = W + b * sw;
  T *Vb = V + b * sv;
  const int rows_per_chunk = kBjrChunk / k;

  // 1. Gram matrix of the pair's columns from LDS-staged row chunks.  Fewer entries than threads:
  // R replicas of each entry sum interleaved rows, combined in replica order; otherwise each
  // thread owns entries tid, tid + NT, ...
  const int E = k * (k + 1) / 2;
  const int R = E < NT ? NT / E : 1;
  const int e0 = E < NT ? tid % E : tid, rep = E < NT ? tid / E : 0;
  const bool own = E < NT ? tid < R * E : true;
  int ep[kBjrMaxEpt], eq[kBjrMaxEpt];
  double acc[kBjrMaxEpt];
#pragma unroll
  for (int j = 0; j < kBjrMaxEpt; ++j) {
    acc[j] = 0.0;
    ep[j] = eq[j] = 0;
    const int e = e0 + j * NT;
    if (own && e < E && (j == 0 || E >= NT)) bjr_unrank(e, k, ep[j], eq[j]);
  }
  for (int r0 = 0; r0 < m; r0 += rows_per_chunk) {
    const int nr = m - r0 < rows_per_chunk ? m - r0 : rows_per_chunk;
    bjr_stage(X, Wb, ldw, r0, nr, k, col);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kBjrMaxEpt; ++j) {
      const int e = e0 + j * NT;
      if (own && e < E && (j == 0 || E >= NT)) {
        double a = acc[j];
        for (int r = rep; r < nr; r += R) a = fma(X[r * k + ep[j]], X[r * k + eq[j]], a);
        acc[j] = a;
      }
    }
    __syncthreads();
  }
  if (E < NT) {
    part[tid] = own ? acc[0] : 0.0;
    __syncthreads();
    if (tid < E) {
      double s = 0.0;
      for (int x = 0; x < R; ++x) s += part[x * E + tid];
      int p, q;
      bjr_unrank(tid, k, p, q);
      G[p * ld + q] = s;
      G[q * ld + p] = s;
    }
  } else {
#pragma unroll
    for (int j = 0; j < kBjrMaxEpt; ++j) {
      if (tid + j * NT < E) {
        G[ep[j] * ld + eq[j]] = acc[j];
        G[eq[j] * ld + ep[j]] = acc[j];
      }
    }
  }
  for (int e = tid; e < k * k; e += NT) {
    const int r = e / k, c = e - r * k;
    Q[r * ld + c] = r == c ? 1.0 : 0.0;
  }
  __syncthreads();

  // 2. one-sided Jacobi of the pair on its Gram matrix: pairs of up to 16 columns in wave 0 alone
  // (no workgroup barriers), wider ones by the whole workgroup
  int sweeps;
  bool converged;
  if (k <= 16) {
    if (wave == 0) bjr_inner<T, true>(G, Q, k, ld, tol, max_sweeps, rc, rs, rp, rq, part, flag);
    __syncthreads();
  } else {
    bjr_inner<T, false>(G, Q, k, ld, tol, max_sweeps, rc, rs, rp, rq, part, flag);
    __syncthreads();
  }
  sweeps = flag[1];
  converged = flag[0] != 0;

  // 3. columns of Q by descending sigma^2 = g_cc (ties by index), written into G's storage
  if (tid < k) {
    const double x = G[tid * ld + tid];
    int rank = 0;
    for (int i = 0; i < k; ++i) {
      const double y = G[i * ld + i];
      rank += (y > x) || (y == x && i < tid);
    }
    perm[rank] = tid;
  }
  __syncthreads();
  for (int e = tid; e < k * k; e += NT) {
    const int c = e / k, j = e - c * k;
    G[c * ld + j] = Q[c * ld + perm[j]];
  }
  __syncthreads();

  // 4. W_pair <- W_pair Q and V_pair <- V_pair Q, row chunks staged in LDS
  for (int pass = 0; pass < 2; ++pass) {
    T *M = pass == 0 ? Wb : Vb;
    const int64_t ldm = pass == 0 ? ldw : ldv;
    const int rows = pass == 0 ? m : n;
    for (int r0 = 0; r0 < rows; r0 += rows_per_chunk) {
      const int nr = rows - r0 < rows_per_chunk ? rows - r0 : rows_per_chunk;
      bjr_stage(X, M, ldm, r0, nr, k, col);
      __syncthreads();
      for (int e = tid; e < nr * k; e += NT) {
        const int r = e / k, j = e - r * k;
        const double *xr = X + r * k;
        double a = 0.0;
        for (int c = 0; c < k; ++c) a = fma(xr[c], G[c * ld + j], a);
        M[(int64_t)(r0 + r) * ldm + col(j)] = (T)a;
      }
      __syncthreads();
    }
  }
  if (tid == 0) {
    atomicMax(&info[b * 2], sweeps);
    if (!converged) atomicOr(&info[b * 2 + 1], 1);
  }
}

__global__ void bjr_reset_kernel(int32_t *info, unsigned long long *corr, unsigned long long *dmax, int64_t batch) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < batch) {
    info[2 * i] = 0;
    info[2 * i + 1] = 0;
    corr[i] = 0ull;
    dmax[i] = 0ull;
  }
}

// column norms^2 (FP64), the norms and the per-matrix maximum; grid (column tiles of 16, batch).
// Thread (g, c) sums rows g, g + 16, ... of column c (16 contiguous columns per row segment); the 16
// row-group partials are added in group order.
template <typename T>
__global__ __launch_bounds__(kBjrThreads) void bjr_norms_kernel(const T *W, int64_t ldw, int64_t sw, int m, int n,
                                                                double *d, double *sigma, unsigned long long *dmax) {
  constexpr int NG = kBjrThreads / kNormCols;
  __shared__ double part[NG][kNormCols];
  __shared__ double red[kBjrThreads / kWave];
  const int64_t b = blockIdx.y;
  const int g = threadIdx.x / kNormCols, c = threadIdx.x % kNormCols;
  const int j = blockIdx.x * kNormCols + c;
  double acc = 0.0;
  if (j < n) {
    const T *cp = W + b * sw + j;
    int r = g;
    for (; r + 3 * NG < m; r += 4 * NG) {
      const double x0 = (double)cp[(int64_t)r * ldw], x1 = (double)cp[(int64_t)(r + NG) * ldw];
      const double x2 = (double)cp[(int64_t)(r + 2 * NG) * ldw], x3 = (double)cp[(int64_t)(r + 3 * NG) * ldw];
      acc = fma(x0, x0, acc);
      acc = fma(x1, x1, acc);
      acc = fma(x2, x2, acc);
      acc = fma(x3, x3, acc);
    }
    for (; r < m; r += NG) {
      const double x = (double)cp[(int64_t)r * ldw];
      acc = fma(x, x, acc);
    }
  }
  part[g][c] = acc;
  __syncthreads();
  double tot = 0.0;
  if (g == 0 && j < n) {
    for (int x = 0; x < NG; ++x) tot += part[x][c];
    d[b * n + j] = tot;
    sigma[b * n + j] = sqrt(tot);
  }
  const double mx = block_max<kBjrThreads>(tot, red);
  if (threadIdx.x == 0) atomicMax(&dmax[b], (unsigned long long)__double_as_longlong(mx));
}

// max over live pairs i != j of |w_i^T w_j| / (|w_i| |w_j|); grid (tiles, tiles, batch), upper half
template <typename T>
__global__ __launch_bounds__(256) void bjr_corr_kernel(const T *W, int64_t ldw, int64_t sw, int m, int n,
                                                       const double *d, const unsigned long long *dmax,
                                                       double live_factor, unsigned long long *corr) {
  const int ti = blockIdx.x, tj = blockIdx.y;
  const int64_t b = blockIdx.z;
  if (ti > tj) return;
  __shared__ double A[kCorrRows][kCorrTile + 1], B[kCorrRows][kCorrTile + 1];
  __shared__ double red[256 / kWave];
  const int tid = threadIdx.x, a = tid / kCorrTile, c = tid % kCorrTile;
  const int ci = ti * kCorrTile + a, cj = tj * kCorrTile + c;
  const T *Wb = W + b * sw;
  double acc = 0.0;
  for (int r0 = 0; r0 < m; r0 += kCorrRows) {
    for (int e = tid; e < kCorrRows * kCorrTile; e += 256) {
      const int r = e / kCorrTile, x = e % kCorrTile, row = r0 + r;
      const int c1 = ti * kCorrTile + x, c2 = tj * kCorrTile + x;
      A[r][x] = (row < m && c1 < n) ? (double)Wb[(int64_t)row * ldw + c1] : 0.0;
      B[r][x] = (row < m && c2 < n) ? (double)Wb[(int64_t)row * ldw + c2] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int r = 0; r < kCorrRows; ++r) acc = fma(A[r][a], B[r][c], acc);
    __syncthreads();
  }
  double ratio = 0.0;
  if (ci < n && cj < n && ci != cj) {
    const double thr = live_factor * __longlong_as_double((long long)dmax[b]);
    const double di = d[b * n + ci], dj = d[b * n + cj];
    if (di > thr && dj > thr) ratio = fabs(acc) / fmax(sqrt(di) * sqrt(dj), DBL_MIN);
  }
  const double mx = block_max<256>(ratio, red);
  if (tid == 0) atomicMax(&corr[b], (unsigned long long)__double_as_longlong(mx));
}

template <typename T>
static int run_block_jacobi_sweep(int64_t batch, int m, int n, int width, T *w, int64_t ldw, int64_t sw, T *v,
                                  int64_t ldv, int64_t sv, double tol, int max_sweeps, int32_t *info, double *corr,
                                  double *sigma, double *d, unsigned long long *dmax, hipStream_t st) {
  const int groups = (n + width - 1) / width;
  const int gp = groups + (groups & 1);
  const int kmax = 2 * width < n ? 2 * width : n;
  const size_t lds = bjr_lds_bytes(kmax);
  if (lds > 64 * 1024 &&
      hipFuncSetAttribute(reinterpret_cast<const void *>(bjr_round_kernel<T>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
    return PSVD_ELAUNCH;
  unsigned long long *cbits = reinterpret_cast<unsigned long long *>(corr);
  hipLaunchKernelGGL(bjr_reset_kernel, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, st, info, cbits, dmax,
                     batch);
  const int64_t blocks = batch * (gp / 2);
  for (int r = 0; r < gp - 1; ++r)
    hipLaunchKernelGGL(bjr_round_kernel<T>, dim3((unsigned)blocks), dim3(kBjrThreads), lds, st, w, ldw, sw, v, ldv, sv,
                       m, n, width, groups, gp, r, tol, max_sweeps, info);
  hipLaunchKernelGGL(bjr_norms_kernel<T>, dim3((unsigned)((n + kNormCols - 1) / kNormCols), (unsigned)batch),
                     dim3(kBjrThreads), 0, st, w, ldw, sw, m, n, d, sigma, dmax);
  const int tiles = (n + kCorrTile - 1) / kCorrTile;
  const double live = Eps<T>::value * (m > n ? m : n);
  hipLaunchKernelGGL(bjr_corr_kernel<T>, dim3((unsigned)tiles, (unsigned)tiles, (unsigned)batch), dim3(256), 0, st,
                     w, ldw, sw, m, n, d, dmax, live * live, cbits);
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

}  // namespace psvd

extern "C" {

size_t psvd_block_jacobi_workspace_bytes(int64_t batch, int n) {
  if (batch < 0 || n < 0) return 0;
  return (size_t)8 * (size_t)batch * ((size_t)n + 1) + 256;
}

int psvd_block_jacobi_sweep(int dtype, int64_t batch, int m, int n, int block_size, void *w, int64_t ldw,
                            int64_t stride_w, void *v, int64_t ldv, int64_t stride_v, double tol, int max_sweeps,
                            int32_t *info, double *corr, double *sigma, void *workspace, size_t workspace_bytes,
                            psvd_stream_t stream) {
  if (batch < 0 || batch > 65535 || m < 1 || n < 2 || block_size < 1 || max_sweeps < 1) return PSVD_EINVAL;
  const int groups = (n + block_size - 1) / block_size;
  if (groups < 2) return PSVD_EINVAL;   // a single group is a plain one-sided Jacobi SVD
  if (2 * (int64_t)block_size > psvd::kBjrMaxK) return PSVD_ETOOBIG;
  if (batch == 0) return PSVD_OK;
  if (!w || !v || !info || !corr || !sigma || !workspace || ldw < n || ldv < n ||
      (batch > 1 && (stride_w < (int64_t)m * ldw || stride_v < (int64_t)n * ldv)))
    return PSVD_EINVAL;
  if (workspace_bytes < psvd_block_jacobi_workspace_bytes(batch, n)) return PSVD_EINVAL;
  if (batch * ((groups + (groups & 1)) / 2) > 0x7fffffffLL) return PSVD_ETOOBIG;
  char *base = static_cast<char *>(workspace);
  base += (256 - (reinterpret_cast<uintptr_t>(base) & 255)) & 255;
  double *d = reinterpret_cast<double *>(base);
  unsigned long long *dmax = reinterpret_cast<unsigned long long *>(d + batch * n);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == PSVD_F32) {
    if (!(tol > 0.0)) tol = 10.0 * psvd::Eps<float>::value;
    return psvd::run_block_jacobi_sweep<float>(batch, m, n, block_size, static_cast<float *>(w), ldw, stride_w,
                                               static_cast<float *>(v), ldv, stride_v, tol, max_sweeps, info, corr,
                                               sigma, d, dmax, st);
  }
  if (dtype == PSVD_F64) {
    if (!(tol > 0.0)) tol = 10.0 * psvd::Eps<double>::value;
    return psvd::run_block_jacobi_sweep<double>(batch, m, n, block_size, static_cast<double *>(w), ldw, stride_w,
                                                static_cast<double *>(v), ldv, stride_v, tol, max_sweeps, info, corr,
                                                sigma, d, dmax, st);
  }
  return PSVD_EINVAL;
}

}  // extern "C"
