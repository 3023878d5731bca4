// Block-pair one-sided Jacobi (svdbench/standalone/solvers.py:30-68) on gfx950.
//
// The n columns of W (m x n, row-major) form groups of `width` consecutive columns (the last one
// may be short; an odd group count gets an empty partner group), paired by the round-robin
// tournament of primitives.py:309-313.  The reference solves each pair with a full one-sided
// Jacobi SVD of its k <= 2 width columns and replaces them by W_pair V_pair, V_pair sorted by
// descending sigma (solvers.py:48-55); V gets the same update.  Here one workgroup takes one
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
// The convergence test of a sweep (solvers.py:56-62) follows in two kernels: FP64 column norms
// (with the per-matrix maximum) and the maximum of |w_i^T w_j| / (|w_i| |w_j|) over live column
// pairs, live meaning |w_i|^2 > (eps max(m, n))^2 max_c |w_c|^2, from 16 x 16 column tiles with
// FP64 sums (the reference forms W^H W in the data precision).  Per-matrix maxima are combined
// with device atomics on the bit patterns of non-negative doubles after a fill kernel clears them.
#include <float.h>

#include "common.hpp"

namespace psvd {

constexpr int kBjrThreads = 256;
constexpr int kBjrMaxK = 64;      // columns of one pair: 2 * block_size
constexpr int kBjrChunk = 1024;   // doubles of the row-chunk staging buffer
constexpr int kCorrTile = 16;
constexpr int kCorrRows = 64;

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
  if (k < 2) return;   // a lone single column: nothing to rotate (solvers.py:50-51)
  const int ci0 = gi * width, cj0 = gj * width;
  auto col = [&](int c) { return c < ki ? ci0 + c : cj0 + (c - ki); };
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

  // 1. Gram matrix of the pair's columns
  const int E = k * (k + 1) / 2;
  if (E < NT) {
    const int R = NT / E;
    double acc = 0.0;
    if (tid < R * E) {
      const int e = tid % E, rep = tid / E;
      int p, q;
      bjr_unrank(e, k, p, q);
      const T *cp = Wb + col(p), *cq = Wb + col(q);
      for (int r = rep; r < m; r += R) acc = fma((double)cp[(int64_t)r * ldw], (double)cq[(int64_t)r * ldw], acc);
    }
    part[tid] = acc;
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
    for (int e = tid; e < E; e += NT) {
      int p, q;
      bjr_unrank(e, k, p, q);
      const T *cp = Wb + col(p), *cq = Wb + col(q);
      double acc = 0.0;
      for (int r = 0; r < m; ++r) acc = fma((double)cp[(int64_t)r * ldw], (double)cq[(int64_t)r * ldw], acc);
      G[p * ld + q] = acc;
      G[q * ld + p] = acc;
    }
  }
  for (int e = tid; e < k * k; e += NT) {
    const int r = e / k, c = e - r * k;
    Q[r * ld + c] = r == c ? 1.0 : 0.0;
  }
  __syncthreads();

  // 2. one-sided Jacobi of the pair on its Gram matrix
  const int kp = k + (k & 1), np2 = kp / 2;
  const double eps2 = Eps<T>::value * Eps<T>::value;
  int sweeps = 0;
  bool converged = false;
  for (int sweep = 1; sweep <= max_sweeps; ++sweep) {
    double off = 0.0;   // kept by thread 0
    for (int rr = 0; rr < kp - 1; ++rr) {
      if (wave == 0) {
        const double dmax = wave_max(lane < k ? G[lane * ld + lane] : 0.0);
        if (tid < np2) {
          int p, q;
          rr_pair(rr, tid, kp, p, q);
          double c = 1.0, s = 0.0, ratio = 0.0;
          if (q < k) {   // q == k is the zero padding column: never active
            const double app = G[p * ld + p], aqq = G[q * ld + q], apq = G[p * ld + q];
            const double den = sqrt(app) * sqrt(aqq);
            ratio = fabs(apq) / fmax(den, DBL_MIN);
            if (ratio > tol && fmin(app, aqq) > eps2 * dmax)
              jacobi_rotation(app, aqq, apq, c, s);
            else
              ratio = 0.0;
          }
          rp[tid] = p;
          rq[tid] = q;
          rc[tid] = c;
          rs[tid] = ratio > 0.0 ? s : 0.0;
          part[tid] = ratio;
        }
      }
      __syncthreads();
      if (tid == 0)
        for (int i = 0; i < np2; ++i) off = fmax(off, part[i]);
      // G <- G J and Q <- Q J: columns p, q of every active pair
      for (int it = tid; it < np2 * k; it += NT) {
        const int i = it / k, r = it - i * k;
        if (part[i] == 0.0) continue;
        const double c = rc[i], s = rs[i];
        const int p = rp[i], q = rq[i];
        const double gp = G[r * ld + p], gq = G[r * ld + q];
        G[r * ld + p] = c * gp - s * gq;
        G[r * ld + q] = s * gp + c * gq;
        const double qp = Q[r * ld + p], qq = Q[r * ld + q];
        Q[r * ld + p] = c * qp - s * qq;
        Q[r * ld + q] = s * qp + c * qq;
      }
      __syncthreads();
      // G <- J^T G: rows p, q of every active pair
      for (int it = tid; it < np2 * k; it += NT) {
        const int i = it / k, r = it - i * k;
        if (part[i] == 0.0) continue;
        const double c = rc[i], s = rs[i];
        const int p = rp[i], q = rq[i];
        const double gp = G[p * ld + r], gq = G[q * ld + r];
        G[p * ld + r] = c * gp - s * gq;
        G[q * ld + r] = s * gp + c * gq;
      }
      __syncthreads();
    }
    sweeps = sweep;
    if (tid == 0) flag[0] = off <= tol;
    __syncthreads();
    converged = flag[0] != 0;
    __syncthreads();
    if (converged) break;
  }

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
  const int rows_per_chunk = kBjrChunk / k;
  for (int pass = 0; pass < 2; ++pass) {
    T *M = pass == 0 ? Wb : Vb;
    const int64_t ldm = pass == 0 ? ldw : ldv;
    const int rows = pass == 0 ? m : n;
    for (int r0 = 0; r0 < rows; r0 += rows_per_chunk) {
      const int nr = rows - r0 < rows_per_chunk ? rows - r0 : rows_per_chunk;
      for (int e = tid; e < nr * k; e += NT) {
        const int r = e / k, c = e - r * k;
        X[e] = (double)M[(int64_t)(r0 + r) * ldm + col(c)];
      }
      __syncthreads();
      for (int e = tid; e < nr * k; e += NT) {
        const int r = e / k, j = e - r * k;
        const double *xr = X + r * k;
        double acc = 0.0;
        for (int c = 0; c < k; ++c) acc = fma(xr[c], G[c * ld + j], acc);
        M[(int64_t)(r0 + r) * ldm + col(j)] = (T)acc;
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

// column norms^2 (FP64), the norms and the per-matrix maximum; grid (column chunks, batch)
template <typename T>
__global__ __launch_bounds__(kBjrThreads) void bjr_norms_kernel(const T *W, int64_t ldw, int64_t sw, int m, int n,
                                                                double *d, double *sigma, unsigned long long *dmax) {
  __shared__ double red[kBjrThreads / kWave];
  const int64_t b = blockIdx.y;
  const int j = blockIdx.x * kBjrThreads + threadIdx.x;
  double acc = 0.0;
  if (j < n) {
    const T *c = W + b * sw + j;
    for (int r = 0; r < m; ++r) {
      const double x = (double)c[(int64_t)r * ldw];
      acc = fma(x, x, acc);
    }
    d[b * n + j] = acc;
    sigma[b * n + j] = sqrt(acc);
  }
  const double mx = block_max<kBjrThreads>(acc, red);
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
  hipLaunchKernelGGL(bjr_norms_kernel<T>, dim3((unsigned)((n + kBjrThreads - 1) / kBjrThreads), (unsigned)batch),
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
