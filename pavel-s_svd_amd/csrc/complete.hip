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

}  // namespace psvd

extern "C" {

size_t psvd_complete_workspace_bytes(int64_t batch, int m, int n) {
  if (batch < 0 || m < 0 || n < 0) return 0;
  return (size_t)8 * batch * (2 * (size_t)m + n) + 256;
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
  // v, leverage and y in LDS when they fit the default 64 KB of dynamic LDS
  const size_t lds = sizeof(double) * (2 * (size_t)m + n);
  const bool in_lds = lds <= 65536;
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
