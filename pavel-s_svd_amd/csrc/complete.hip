// Completion of missing columns of a batch of orthonormal bases (svdbench/standalone/
// primitives.py:430-467): U (m x n, row-major) has live columns [0, nk_b) for matrix b and the
// columns [nk_b, n) are replaced, one after the other, by the coordinate vector at the row of
// largest remaining leverage 1 - |U(i, :)|^2 (first such row), projected twice against all
// current columns (classical Gram-Schmidt, twice) and normalised; a residual below 10 eps(T)
// sets the failure status (the reference raises).  Columns still to be completed are zero while
// earlier ones are used, exactly as in the reference, so the products run over the r columns
// already filled.  FP64 accumulation throughout.
//
// One workgroup per matrix.  The leverage vector is formed once and, after column r is written,
// lowered by its squared entries (the reference recomputes 1 - |U(i, :)|^2 from scratch; the two
// agree to rounding).  U(:, :r) u-products run four rows per wave with 16 lanes across each row
// (64-byte segments) and the lanes' partials added by DPP moves within the 16-lane group; U^T v
// runs row slices x columns with the slices added in a fixed order, so every sum has a fixed
// order and the result is deterministic.
#include <float.h>

#include <algorithm>

#include "common.hpp"

namespace psvd {

constexpr int kCompleteThreads = 1024;
// the row products U(:, :r) x: each wave takes kRowGroups rows at a time, kRowLanes lanes per row,
// so one 4-level DPP sum serves four rows (a full-wave sum per row made these loops issue-bound)
constexpr int kRowLanes = 16, kRowGroups = kWave / kRowLanes;
static_assert(kRowLanes == 16, "the leaders' sum below reads lanes 0, 16, 32, 48");

// LDSV: v, the leverage vector and y live in LDS (8 (2m + n) bytes of dynamic LDS) instead of the
// workspace, so no per-row global store sits in front of the next row's loads.
template <typename T, bool LDSV>
__global__ __launch_bounds__(kCompleteThreads) void complete_kernel(T *U, int64_t ldu, int64_t su, int m, int n,
                                                                    const int32_t *nk, int32_t *status, double *vbuf,
                                                                    double *ybuf, double *levbuf) {
  constexpr int NT = kCompleteThreads, NW = NT / kWave;
  __shared__ double red[NW];
  __shared__ int redi[NW];
  __shared__ double ypart[NT];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int grp = lane / kRowLanes, gl = lane % kRowLanes;   // row group of the wave, lane in it
  const int64_t b = blockIdx.x;
  T *Ub = U + b * su;
  extern __shared__ double smv[];
  double *v = LDSV ? smv : vbuf + b * (int64_t)m;
  double *lev = LDSV ? smv + m : levbuf + b * (int64_t)m;
  double *y = LDSV ? smv + 2 * m : ybuf + b * (int64_t)n;
  const int k0 = nk[b] < 0 ? 0 : (nk[b] > n ? n : nk[b]);
  const double eps = Eps<T>::value;
  int st = 0;
  if (k0 == n) {
    if (tid == 0) status[b] = 0;
    return;
  }
  // missing columns start at zero
  for (int64_t e = tid; e < (int64_t)m * (n - k0); e += NT) {
    const int i = (int)(e / (n - k0)), c = k0 + (int)(e - (int64_t)i * (n - k0));
    Ub[(int64_t)i * ldu + c] = T(0);
  }
  // leverage of the live columns
  for (int i = wave; i < m; i += NW) {
    const T *row = Ub + (int64_t)i * ldu;
    double a = 0.0;
    for (int c = lane; c < k0; c += kWave) a = fma((double)row[c], (double)row[c], a);
    a = group_sum_dpp<kWave>(a);
    if (lane == 0) lev[i] = 1.0 - a;
  }
  __syncthreads();
  for (int r = k0; r < n; ++r) {
    double best = -DBL_MAX;
    int bi = 0x7fffffff;
    for (int i = tid; i < m; i += NT) {
      const double l = lev[i];
      if (l > best) {
        best = l;
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
    if (lane == 0) {
      red[wave] = best;
      redi[wave] = bi;
    }
    __syncthreads();
    best = red[0];
    bi = redi[0];
    for (int w = 1; w < NW; ++w)
      if (red[w] > best || (red[w] == best && redi[w] < bi)) {
        best = red[w];
        bi = redi[w];
      }
    const int ix = bi;
    const T *rix = Ub + (int64_t)ix * ldu;
    // v = e_ix - U (U^T e_ix) = e_ix - U U(ix, :)^T
    for (int i0 = wave * kRowGroups; i0 < m; i0 += NW * kRowGroups) {
      const int i = i0 + grp;
      double a = 0.0;
      if (i < m) {
        const T *row = Ub + (int64_t)i * ldu;
        for (int c = gl; c < r; c += kRowLanes) a = fma((double)row[c], (double)rix[c], a);
      }
      a = group_sum_dpp<kRowLanes>(a);
      if (gl == 0 && i < m) v[i] = (i == ix ? 1.0 : 0.0) - a;
    }
    __syncthreads();
    // y = U(:, :r)^T v: S row slices per column, slices added in order
    if (r > 0) {
      const int S = r >= NT ? 1 : NT / r;
      for (int c0 = 0; c0 < r; c0 += NT) {
        const int c = c0 + tid % (r < NT ? r : NT), sl = r < NT ? tid / r : 0;
        double a = 0.0;
        if (sl < S && c < r)
          for (int i = sl; i < m; i += S) a = fma((double)Ub[(int64_t)i * ldu + c], v[i], a);
        ypart[tid] = a;
        __syncthreads();
        if (tid < (r - c0 < NT ? r - c0 : NT)) {
          double t = 0.0;
          for (int x = 0; x < S; ++x) t += ypart[x * (r < NT ? r : NT) + tid];
          y[c0 + tid] = t;
        }
        __syncthreads();
      }
    }
    // v -= U(:, :r) y and |v|^2
    double l2 = 0.0;
    for (int i0 = wave * kRowGroups; i0 < m; i0 += NW * kRowGroups) {
      const int i = i0 + grp;
      double a = 0.0;
      if (i < m) {
        const T *row = Ub + (int64_t)i * ldu;
        for (int c = gl; c < r; c += kRowLanes) a = fma((double)row[c], y[c], a);
      }
      a = group_sum_dpp<kRowLanes>(a);
      if (gl == 0 && i < m) {
        const double vi = v[i] - a;
        v[i] = vi;
        l2 = fma(vi, vi, l2);
      }
    }
    // the four group leaders' sums, in lane order
    l2 = (readlane_double(l2, 0) + readlane_double(l2, 16)) + (readlane_double(l2, 32) + readlane_double(l2, 48));
    if (lane == 0) red[wave] = l2;
    __syncthreads();
    double len2 = 0.0;
    for (int w = 0; w < NW; ++w) len2 += red[w];
    const double len = sqrt(len2);
    if (len < 10.0 * eps) st = PSVD_INFO_COMPLETION;
    const double inv = 1.0 / fmax(len, DBL_MIN);
    for (int i = tid; i < m; i += NT) {
      const T x = (T)(v[i] * inv);
      Ub[(int64_t)i * ldu + r] = x;
      lev[i] -= (double)x * (double)x;
    }
    __syncthreads();
  }
  if (tid == 0) status[b] = st;
}


// ---- Wide path: few large matrices, every column spread over the whole GPU --------------------
// For batch <= kWideBatch the per-column work (three products over the r columns already filled)
// runs as three launches per column over all CUs instead of one workgroup per matrix:
//  cw_project: global argmax of the leverage (from per-workgroup partials, fixed order) -> ix;
//              v = e_ix - U(:, :r) U(ix, :r)^T for the workgroup's rows; |v|^2 partial.
//  cw_gram:    y = U(:, :r)^T v as (column block, row slice) partials.
//  cw_finish:  y and the partials summed in a fixed order; |v - U y|^2 = |v|^2 - |y|^2 (v is already
//              orthogonal to U to rounding after the first pass, so |y| << |v| and nothing cancels);
//              column r = (v - U y) / |v - U y|, leverage lowered by its squares, argmax partial.
// The two projections and the leverage rule are the reference's (primitives.py:430-455).  Every
// sum has a fixed order, so the result does not depend on scheduling.  A wave takes one row at a
// time, its 64 lanes across the row (256-byte FP32 segments), the lanes' partials added by DPP.
constexpr int kCw = 256, kCwWaves = kCw / kWave, kCwMaxParts = 256, kCwMaxSlices = 16;
constexpr int kWideBatch = 8;

template <typename T>
struct CwParams {
  T *U;
  int64_t ldu, su;
  int m, n, rb, npart;
  const int32_t *nk;
  int32_t *status;
  double *v, *lev, *v2, *bestv, *yp;
  int *besti;
};

__device__ __forceinline__ int cw_k0(const int32_t *nk, int64_t b, int n) {
  const int k = nk[b];
  return k < 0 ? 0 : (k > n ? n : k);
}

// (value, index) maximum with ties to the smaller index (the first maximum, as argmax)
__device__ __forceinline__ void cw_better(double &bv, int &bi, double ov, int oi) {
  if (ov > bv || (ov == bv && oi < bi)) {
    bv = ov;
    bi = oi;
  }
}

// the workgroup's (value, index) maximum from every thread's candidate, combined in wave order
__device__ __forceinline__ void cw_block_argmax(double &bv, int &bi, double *rv, int *ri) {
  for (int o = kWave >> 1; o > 0; o >>= 1) cw_better(bv, bi, __shfl_xor(bv, o, kWave), __shfl_xor(bi, o, kWave));
  const int wave = threadIdx.x / kWave;
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) {
    rv[wave] = bv;
    ri[wave] = bi;
  }
  __syncthreads();
  bv = rv[0];
  bi = ri[0];
  for (int w = 1; w < kCwWaves; ++w) cw_better(bv, bi, rv[w], ri[w]);
}

// a fixed-order block sum that every workgroup evaluates to the same bits
__device__ __forceinline__ double cw_block_sum(double a, double *red) {
  a = group_sum_dpp<kWave>(a);
  const int wave = threadIdx.x / kWave;
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) red[wave] = a;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < kCwWaves; ++w) t += red[w];
  return t;
}

template <typename T>
__device__ __forceinline__ double cw_row_dot(const T *row, const T *x, int r, int lane) {
  double a = 0.0;
#pragma unroll 4
  for (int c = lane; c < r; c += kWave) a = fma((double)row[c], (double)x[c], a);
  return group_sum_dpp<kWave>(a);
}

// leverage of the live columns, argmax partials, status cleared
template <typename T>
__global__ __launch_bounds__(kCw) void cw_init_kernel(CwParams<T> P) {
  __shared__ double rv[kCwWaves];
  __shared__ int ri[kCwWaves];
  const int64_t b = blockIdx.y;
  const int wg = blockIdx.x, lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int k0 = cw_k0(P.nk, b, P.n);
  if (wg == 0 && threadIdx.x == 0) P.status[b] = 0;
  if (k0 >= P.n) return;
  const T *Ub = P.U + b * P.su;
  double bv = -DBL_MAX;
  int bi = 0x7fffffff;
  for (int j = wave; j < P.rb; j += kCwWaves) {
    const int i = wg * P.rb + j;
    if (i >= P.m) break;
    const T *row = Ub + (int64_t)i * P.ldu;
    const double l = 1.0 - cw_row_dot(row, row, k0, lane);
    if (lane == 0) P.lev[b * P.m + i] = l;
    cw_better(bv, bi, l, i);
  }
  cw_block_argmax(bv, bi, rv, ri);
  if (threadIdx.x == 0) {
    P.bestv[b * kCwMaxParts + wg] = bv;
    P.besti[b * kCwMaxParts + wg] = bi;
  }
}

template <typename T>
__global__ __launch_bounds__(kCw) void cw_project_kernel(CwParams<T> P, int r) {
  __shared__ double rv[kCwWaves];
  __shared__ int ri[kCwWaves];
  const int64_t b = blockIdx.y;
  const int wg = blockIdx.x, lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (r < cw_k0(P.nk, b, P.n)) return;
  double bv = -DBL_MAX;
  int bi = 0x7fffffff;
  for (int k = threadIdx.x; k < P.npart; k += kCw)
    cw_better(bv, bi, P.bestv[b * kCwMaxParts + k], P.besti[b * kCwMaxParts + k]);
  cw_block_argmax(bv, bi, rv, ri);
  const int ix = bi;
  const T *Ub = P.U + b * P.su, *rix = Ub + (int64_t)ix * P.ldu;
  double acc = 0.0;
  for (int j = wave; j < P.rb; j += kCwWaves) {
    const int i = wg * P.rb + j;
    if (i >= P.m) break;
    const double vi = (i == ix ? 1.0 : 0.0) - cw_row_dot(Ub + (int64_t)i * P.ldu, rix, r, lane);
    if (lane == 0) P.v[b * P.m + i] = vi;
    acc = fma(vi, vi, acc);
  }
  __syncthreads();
  if (lane == 0) rv[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kCwWaves; ++w) t += rv[w];
    P.v2[b * kCwMaxParts + wg] = t;
  }
}

// y partials: blockIdx.x = 64-column block, blockIdx.y = row slice, blockIdx.z = matrix
template <typename T>
__global__ __launch_bounds__(kCw) void cw_gram_kernel(CwParams<T> P, int r, int S) {
  __shared__ double part[kCwWaves][kWave];
  const int64_t b = blockIdx.z;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (r < cw_k0(P.nk, b, P.n)) return;
  const int c = blockIdx.x * kWave + lane, s = blockIdx.y;
  const int rps = (P.m + S - 1) / S, i0 = s * rps, i1 = i0 + rps < P.m ? i0 + rps : P.m;
  const T *Ub = P.U + b * P.su;
  const double *v = P.v + b * P.m;
  double acc = 0.0;
  if (c < r) {
#pragma unroll 8
    for (int i = i0 + wave; i < i1; i += kCwWaves) acc = fma((double)Ub[(int64_t)i * P.ldu + c], v[i], acc);
  }
  part[wave][lane] = acc;
  __syncthreads();
  if (wave == 0 && c < r) {
    double t = 0.0;
    for (int w = 0; w < kCwWaves; ++w) t += part[w][lane];
    P.yp[(b * kCwMaxSlices + s) * (int64_t)P.n + c] = t;
  }
}

template <typename T>
__global__ __launch_bounds__(kCw) void cw_finish_kernel(CwParams<T> P, int r, int S) {
  extern __shared__ double ys[];
  __shared__ double red[kCwWaves];
  __shared__ int ri[kCwWaves];
  const int64_t b = blockIdx.y;
  const int wg = blockIdx.x, lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (r < cw_k0(P.nk, b, P.n)) return;
  double y2 = 0.0;
  for (int c = threadIdx.x; c < r; c += kCw) {
    double t = 0.0;
    for (int s = 0; s < S; ++s) t += P.yp[(b * kCwMaxSlices + s) * (int64_t)P.n + c];
    ys[c] = t;
    y2 = fma(t, t, y2);
  }
  y2 = cw_block_sum(y2, red);
  double v2 = 0.0;
  for (int k = threadIdx.x; k < P.npart; k += kCw) v2 += P.v2[b * kCwMaxParts + k];
  v2 = cw_block_sum(v2, red);
  const double len = sqrt(fmax(v2 - y2, 0.0));
  if (len < 10.0 * Eps<T>::value && wg == 0 && threadIdx.x == 0) P.status[b] = PSVD_INFO_COMPLETION;
  const double inv = 1.0 / fmax(len, DBL_MIN);
  T *Ub = P.U + b * P.su;
  double bv = -DBL_MAX;
  int bi = 0x7fffffff;
  for (int j = wave; j < P.rb; j += kCwWaves) {
    const int i = wg * P.rb + j;
    if (i >= P.m) break;
    T *row = Ub + (int64_t)i * P.ldu;
    double a = 0.0;
#pragma unroll 4
    for (int c = lane; c < r; c += kWave) a = fma((double)row[c], ys[c], a);
    a = group_sum_dpp<kWave>(a);
    const T x = (T)((P.v[b * P.m + i] - a) * inv);
    const double l = P.lev[b * P.m + i] - (double)x * (double)x;
    if (lane == 0) {
      row[r] = x;
      P.lev[b * P.m + i] = l;
    }
    cw_better(bv, bi, l, i);
  }
  cw_block_argmax(bv, bi, red, ri);
  if (threadIdx.x == 0) {
    P.bestv[b * kCwMaxParts + wg] = bv;
    P.besti[b * kCwMaxParts + wg] = bi;
  }
}

static bool cw_eligible(int64_t batch, int m, int n) { return batch <= kWideBatch && m >= 256 && n <= 8192; }

static size_t cw_bytes(int64_t batch, int m, int n) {
  return (size_t)8 * batch * (2 * (size_t)m + (size_t)kCwMaxSlices * n) + (size_t)20 * batch * kCwMaxParts + 2048;
}

template <typename T>
static int complete_wide(int64_t batch, int m, int n, T *u, int64_t ldu, int64_t su, const int32_t *nk,
                         int32_t *status, void *ws, hipStream_t st) {
  static bool attr = false;
  if (!attr) {   // y of up to 8192 columns in dynamic LDS
    if (hipFuncSetAttribute((const void *)cw_finish_kernel<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            8 * 8192) != hipSuccess)
      return PSVD_ELAUNCH;
    attr = true;
  }
  // the first column to complete: the smallest live count of the batch (one host read)
  int32_t hk[kWideBatch];
  if (hipMemcpyAsync(hk, nk, sizeof(int32_t) * batch, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return PSVD_ELAUNCH;
  int kmin = n;
  for (int64_t b = 0; b < batch; ++b) kmin = std::min(kmin, std::max(0, std::min<int>(hk[b], n)));
  CwParams<T> P;
  P.U = u;
  P.ldu = ldu;
  P.su = su;
  P.m = m;
  P.n = n;
  P.rb = kCwWaves * ((m + kCwWaves * kCwMaxParts - 1) / (kCwWaves * kCwMaxParts));
  P.npart = (m + P.rb - 1) / P.rb;
  P.nk = nk;
  P.status = status;
  char *cur = static_cast<char *>(ws);
  auto take = [&](size_t bytes) {
    char *q = cur;
    cur += (bytes + 255) & ~size_t(255);
    return q;
  };
  P.v = reinterpret_cast<double *>(take(8 * batch * (size_t)m));
  P.lev = reinterpret_cast<double *>(take(8 * batch * (size_t)m));
  P.v2 = reinterpret_cast<double *>(take(8 * batch * (size_t)kCwMaxParts));
  P.bestv = reinterpret_cast<double *>(take(8 * batch * (size_t)kCwMaxParts));
  P.besti = reinterpret_cast<int *>(take(4 * batch * (size_t)kCwMaxParts));
  P.yp = reinterpret_cast<double *>(take(8 * batch * (size_t)kCwMaxSlices * n));
  const dim3 rows((unsigned)P.npart, (unsigned)batch), block(kCw);
  hipLaunchKernelGGL(cw_init_kernel<T>, rows, block, 0, st, P);
  for (int r = kmin; r < n; ++r) {
    hipLaunchKernelGGL(cw_project_kernel<T>, rows, block, 0, st, P, r);
    const int cb = (r + kWave - 1) / kWave;
    int S = 1;
    if (cb > 0) {
      S = (int)(kCwMaxParts / (cb * batch));
      S = S < 1 ? 1 : (S > kCwMaxSlices ? kCwMaxSlices : S);
      S = S > (m + 127) / 128 ? (m + 127) / 128 : S;   // at least 128 rows per slice
      hipLaunchKernelGGL(cw_gram_kernel<T>, dim3((unsigned)cb, (unsigned)S, (unsigned)batch), block, 0, st, P, r, S);
    }
    hipLaunchKernelGGL(cw_finish_kernel<T>, rows, block, sizeof(double) * (size_t)(r > 0 ? r : 1), st, P, r, S);
  }
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

}  // namespace psvd

extern "C" {

size_t psvd_complete_workspace_bytes(int64_t batch, int m, int n) {
  if (batch < 0 || m < 0 || n < 0) return 0;
  const size_t one = (size_t)8 * batch * (2 * (size_t)m + n) + 256;
  return psvd::cw_eligible(batch, m, n) ? std::max(one, psvd::cw_bytes(batch, m, n)) : one;
}

int psvd_complete_columns(int dtype, int64_t batch, int m, int n, void *u, int64_t ldu, int64_t stride_u,
                          const int32_t *nk, int32_t *status, void *workspace, size_t workspace_bytes,
                          psvd_stream_t stream) {
  if (batch < 0 || m < 1 || n < 0 || n > m || ldu < n || !u || !nk || !status || !workspace || batch > 0x7fffffff)
    return PSVD_EINVAL;
  if (batch == 0 || n == 0) return PSVD_OK;
  if (workspace_bytes < psvd_complete_workspace_bytes(batch, m, n)) return PSVD_EINVAL;
  double *vb = static_cast<double *>(workspace);
  double *yb = vb + batch * (int64_t)m;
  double *lb = yb + batch * (int64_t)n;
  hipStream_t st = static_cast<hipStream_t>(stream);
  // few large matrices: every column over the whole GPU (one host read of nk; not inside a capture)
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (psvd::cw_eligible(batch, m, n) && hipStreamIsCapturing(st, &cap) == hipSuccess &&
      cap == hipStreamCaptureStatusNone) {
    if (dtype == PSVD_F32)
      return psvd::complete_wide<float>(batch, m, n, static_cast<float *>(u), ldu, stride_u, nk, status, workspace, st);
    if (dtype == PSVD_F64)
      return psvd::complete_wide<double>(batch, m, n, static_cast<double *>(u), ldu, stride_u, nk, status, workspace,
                                         st);
    return PSVD_EINVAL;
  }
  // v, leverage and y in LDS when they fit, with the kernel's static arrays, in the 64 KB a kernel
  // may use without raising its dynamic-LDS limit
  const size_t lds = sizeof(double) * (2 * (size_t)m + n);
  constexpr size_t kStatic = sizeof(double) * (psvd::kCompleteThreads + psvd::kCompleteThreads / psvd::kWave) +
                             sizeof(int) * (psvd::kCompleteThreads / psvd::kWave);
  const bool in_lds = lds + kStatic <= 65536;
  const dim3 grid((unsigned)batch), block(psvd::kCompleteThreads);
  if (dtype == PSVD_F32) {
    float *uf = static_cast<float *>(u);
    if (in_lds)
      hipLaunchKernelGGL((psvd::complete_kernel<float, true>), grid, block, lds, st, uf, ldu, stride_u, m, n, nk, status,
                         vb, yb, lb);
    else
      hipLaunchKernelGGL((psvd::complete_kernel<float, false>), grid, block, 0, st, uf, ldu, stride_u, m, n, nk, status,
                         vb, yb, lb);
  } else if (dtype == PSVD_F64) {
    double *ud = static_cast<double *>(u);
    if (in_lds)
      hipLaunchKernelGGL((psvd::complete_kernel<double, true>), grid, block, lds, st, ud, ldu, stride_u, m, n, nk,
                         status, vb, yb, lb);
    else
      hipLaunchKernelGGL((psvd::complete_kernel<double, false>), grid, block, 0, st, ud, ldu, stride_u, m, n, nk,
                         status, vb, yb, lb);
  } else {
    return PSVD_EINVAL;
  }
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

}  // extern "C"
