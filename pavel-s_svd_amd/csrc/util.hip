// Batched element-wise and reduction kernels used between the solver stages: per-matrix maximum
// of |a| and Frobenius norm (the two input normalisations of solvers.py:294-295 and :124-125),
// the QDWH convergence pair max|x' - x|, max|x'| (solvers.py:161-163), scaling by a per-matrix
// device scalar, symmetrisation (H + H^T)/2 (solvers.py:156, 169), axpby updates and a descending sort
// of FP64 keys with the matching column permutation.
//
// Reductions run over a 2-D grid (chunks of one matrix x batch) and combine chunk results with
// device atomics on the bit patterns of non-negative doubles (monotone as integers) or double
// atomicAdd, after a fill kernel clears the outputs: no hipMemset, so the sequence can be
// captured in a hipGraph.
#include <float.h>

#include "common.hpp"

namespace psvd {

constexpr int kUtilThreads = 256;

__global__ void fill_u64_kernel(unsigned long long *p, int64_t n, unsigned long long v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// clears a GridBarrier's counters in-stream before the launch that uses them
__global__ void bar_reset_kernel(unsigned *ctr, int words) {
  for (int i = threadIdx.x; i < words; i += blockDim.x) ctr[i] = 0u;
}

int barrier_grid_limit(const void *kernel, int threads, size_t lds) {
  int dev = 0, cus = 0, occ = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, threads, lds) != hipSuccess || occ < 1) return 0;
  return cus;
}

static int fill_u64(unsigned long long *p, int64_t n, unsigned long long v, hipStream_t st) {
  if (n <= 0) return PSVD_OK;
  hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

// op 0: max |a|, op 1: sum a^2 (square root taken by the finish kernel), op 2: max |x - y| and max |x|
template <typename T, int OP>
__global__ __launch_bounds__(kUtilThreads) void reduce_kernel(const T *x, int64_t ldx, int64_t sx, const T *y,
                                                              int64_t ldy, int64_t sy, int rows, int cols,
                                                              unsigned long long *out, int64_t batch) {
  const int64_t b = blockIdx.y;
  const int64_t total = (int64_t)rows * cols;
  const int64_t stride = (int64_t)gridDim.x * kUtilThreads;
  double v0 = 0.0, v1 = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * kUtilThreads + threadIdx.x; e < total; e += stride) {
    const int r = (int)(e / cols), c = (int)(e - (int64_t)r * cols);
    const double a = (double)x[b * sx + (int64_t)r * ldx + c];
    if (OP == 0) v0 = fmax(v0, fabs(a));
    if (OP == 3 && !(c & 1)) v0 = fmax(v0, hypot(a, (double)x[b * sx + (int64_t)r * ldx + c + 1]));
    if (OP == 4 && !(c & 1)) {
      const double ai = (double)x[b * sx + (int64_t)r * ldx + c + 1];
      const double dr = (double)y[b * sy + (int64_t)r * ldy + c], di = (double)y[b * sy + (int64_t)r * ldy + c + 1];
      v0 = fmax(v0, hypot(a - dr, ai - di));
      v1 = fmax(v1, hypot(a, ai));
    }
    if (OP == 1) v0 = fma(a, a, v0);
    if (OP == 2) {
      const double d = (double)y[b * sy + (int64_t)r * ldy + c];
      v0 = fmax(v0, fabs(a - d));
      v1 = fmax(v1, fabs(a));
    }
  }
  if (OP == 1) {
    v0 = group_sum_dpp<kWave>(v0);
  } else {
    v0 = wave_max_fast(v0);
    if (OP == 2 || OP == 4) v1 = wave_max_fast(v1);
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    if (OP == 1) {
      atomicAdd(reinterpret_cast<double *>(&out[b]), v0);
    } else {
      atomicMax(&out[b], (unsigned long long)__double_as_longlong(v0));
      if (OP == 2 || OP == 4) atomicMax(&out[batch + b], (unsigned long long)__double_as_longlong(v1));
    }
  }
}

__global__ void sqrt_kernel(double *p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = sqrt(p[i]);
}

// y_b = (s / d_b) x_b (invert) or y_b = s * d_b * x_b; d_b <= 0 counts as 1 (a zero matrix is
// left unscaled, solvers.py:295); d may be NULL (d_b = 1).
template <typename T>
__global__ __launch_bounds__(kUtilThreads) void scale_kernel(const T *x, int64_t ldx, int64_t sx, T *y, int64_t ldy,
                                                             int64_t sy, int rows, int cols, const double *d,
                                                             int invert, double s) {
  const int64_t b = blockIdx.y;
  const int64_t total = (int64_t)rows * cols;
  double f = d ? d[b] : 1.0;
  if (!(f > 0.0)) f = 1.0;
  const T fac = invert ? (T)f : (T)(s * f);
  for (int64_t e = (int64_t)blockIdx.x * kUtilThreads + threadIdx.x; e < total; e += (int64_t)gridDim.x * kUtilThreads) {
    const int r = (int)(e / cols), c = (int)(e - (int64_t)r * cols);
    const T v = x[b * sx + (int64_t)r * ldx + c];
    // division by the stored scale, as the reference's x / scale (rounding included)
    y[b * sy + (int64_t)r * ldy + c] = invert ? (T)(s * (double)(v / fac)) : v * fac;
  }
}

template <typename T>
__global__ __launch_bounds__(kUtilThreads) void axpby_kernel(T alpha, const T *x, int64_t ldx, int64_t sx, T beta, T *y,
                                                             int64_t ldy, int64_t sy, int rows, int cols) {
  const int64_t b = blockIdx.y;
  const int64_t total = (int64_t)rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * kUtilThreads + threadIdx.x; e < total; e += (int64_t)gridDim.x * kUtilThreads) {
    const int r = (int)(e / cols), c = (int)(e - (int64_t)r * cols);
    T *py = y + b * sy + (int64_t)r * ldy + c;
    *py = alpha * x[b * sx + (int64_t)r * ldx + c] + beta * *py;
  }
}

template <typename T>
__global__ __launch_bounds__(kUtilThreads) void symmetrize_kernel(T *a, int64_t lda, int64_t sa, int n) {
  const int64_t b = blockIdx.y;
  const int64_t total = (int64_t)n * n;
  for (int64_t e = (int64_t)blockIdx.x * kUtilThreads + threadIdx.x; e < total; e += (int64_t)gridDim.x * kUtilThreads) {
    const int r = (int)(e / n), c = (int)(e - (int64_t)r * n);
    if (c <= r) continue;
    T *p = a + b * sa + (int64_t)r * lda + c, *q = a + b * sa + (int64_t)c * lda + r;
    const T v = (*p + *q) / T(2);
    *p = v;
    *q = v;
  }
}

static unsigned grid_chunks(int64_t total) {
  int64_t c = (total + kUtilThreads * 8 - 1) / (kUtilThreads * 8);
  if (c < 1) c = 1;
  if (c > 1024) c = 1024;
  return (unsigned)c;
}

template <typename T>
static int run_reduce(int op, int64_t batch, int rows, int cols, const void *x, int64_t ldx, int64_t sx,
                      const void *y, int64_t ldy, int64_t sy, double *out, hipStream_t st) {
  unsigned long long *o = reinterpret_cast<unsigned long long *>(out);
  const int64_t nout = (op == 2 || op == 4) ? 2 * batch : batch;
  int e = fill_u64(o, nout, 0ull, st);  // the bit pattern of +0.0
  if (e != PSVD_OK) return e;
  const dim3 grid(grid_chunks((int64_t)rows * cols), (unsigned)batch);
  const T *xp = static_cast<const T *>(x), *yp = static_cast<const T *>(y);
  if (op == 0) hipLaunchKernelGGL((reduce_kernel<T, 0>), grid, dim3(kUtilThreads), 0, st, xp, ldx, sx, yp, ldy, sy, rows, cols, o, batch);
  else if (op == 1) hipLaunchKernelGGL((reduce_kernel<T, 1>), grid, dim3(kUtilThreads), 0, st, xp, ldx, sx, yp, ldy, sy, rows, cols, o, batch);
  else if (op == 2) hipLaunchKernelGGL((reduce_kernel<T, 2>), grid, dim3(kUtilThreads), 0, st, xp, ldx, sx, yp, ldy, sy, rows, cols, o, batch);
  else if (op == 3) hipLaunchKernelGGL((reduce_kernel<T, 3>), grid, dim3(kUtilThreads), 0, st, xp, ldx, sx, yp, ldy, sy, rows, cols, o, batch);
  else hipLaunchKernelGGL((reduce_kernel<T, 4>), grid, dim3(kUtilThreads), 0, st, xp, ldx, sx, yp, ldy, sy, rows, cols, o, batch);
  if (op == 1) hipLaunchKernelGGL(sqrt_kernel, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, st, out, batch);
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}


// Descending order of per-matrix FP64 keys (ties by index, the stable order of a flip of an
// ascending sort) by rank counting, then y[:, c] = x[:, perm[c]] and s[c] = max(keys[perm[c]], 0).
__global__ __launch_bounds__(kUtilThreads) void order_desc_kernel(const double *keys, int n, int32_t *perm) {
  const int64_t b = blockIdx.y;
  const int j = blockIdx.x * kUtilThreads + threadIdx.x;
  if (j >= n) return;
  const double *k = keys + b * n;
  const double x = k[j];
  int rank = 0;
  for (int i = 0; i < n; ++i) {
    const double y = k[i];
    rank += (y > x) || (y == x && i < j);
  }
  perm[b * n + rank] = j;
}

template <typename T>
__global__ __launch_bounds__(kUtilThreads) void gather_cols_kernel(const double *keys, const int32_t *perm, int rows,
                                                                   int n, const T *x, int64_t ldx, int64_t sx, T *y,
                                                                   int64_t ldy, int64_t sy, T *s) {
  const int64_t b = blockIdx.y;
  const int64_t e = (int64_t)blockIdx.x * kUtilThreads + threadIdx.x;
  if (e >= (int64_t)(rows > 0 ? rows : 1) * n) return;
  const int r = (int)(e / n), c = (int)(e - (int64_t)r * n);
  int src = perm[b * n + c];
  src = src < 0 ? 0 : (src >= n ? n - 1 : src);   // NaN keys leave rank gaps; never index outside x
  if (rows > 0) y[b * sy + (int64_t)r * ldy + c] = x[b * sx + (int64_t)r * ldx + src];
  if (r == 0) s[b * n + c] = (T)fmax(keys[b * n + src], 0.0);
}

// x[:, c] /= d[c] for d[c] > 0 (others unchanged): U = W / sigma of solvers.py:18-21 and :116
template <typename T>
__global__ __launch_bounds__(kUtilThreads) void scale_cols_kernel(T *x, int64_t ldx, int64_t sx, int rows, int cols,
                                                                  const T *d, int64_t sd) {
  const int64_t b = blockIdx.y;
  const int64_t total = (int64_t)rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * kUtilThreads + threadIdx.x; e < total; e += (int64_t)gridDim.x * kUtilThreads) {
    const int r = (int)(e / cols), c = (int)(e - (int64_t)r * cols);
    const T f = d[b * sd + c];
    if (f > T(0)) x[b * sx + (int64_t)r * ldx + c] /= f;
  }
}

}  // namespace psvd

extern "C" {

int psvd_reduce(int dtype, int op, int64_t batch, int rows, int cols, const void *x, int64_t ldx, int64_t sx,
                const void *y, int64_t ldy, int64_t sy, double *out, psvd_stream_t stream) {
  if (batch < 0 || rows < 0 || cols < 0 || op < 0 || op > 4 || !out || batch > 65535) return PSVD_EINVAL;
  if (op >= 3 && (cols & 1)) return PSVD_EINVAL;
  if (batch == 0) return PSVD_OK;
  if (!x || ((op == 2 || op == 4) && !y)) return PSVD_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == PSVD_F32) return psvd::run_reduce<float>(op, batch, rows, cols, x, ldx, sx, y, ldy, sy, out, st);
  if (dtype == PSVD_F64) return psvd::run_reduce<double>(op, batch, rows, cols, x, ldx, sx, y, ldy, sy, out, st);
  return PSVD_EINVAL;
}

int psvd_scale(int dtype, int64_t batch, int rows, int cols, const void *x, int64_t ldx, int64_t sx, void *y,
               int64_t ldy, int64_t sy, const double *d, int invert, double s, psvd_stream_t stream) {
  if (batch < 0 || rows < 0 || cols < 0 || !x || !y || batch > 65535) return PSVD_EINVAL;
  if (batch == 0 || rows == 0 || cols == 0) return PSVD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 grid(psvd::grid_chunks((int64_t)rows * cols), (unsigned)batch);
  if (dtype == PSVD_F32)
    hipLaunchKernelGGL(psvd::scale_kernel<float>, grid, dim3(psvd::kUtilThreads), 0, st, static_cast<const float *>(x),
                       ldx, sx, static_cast<float *>(y), ldy, sy, rows, cols, d, invert, s);
  else if (dtype == PSVD_F64)
    hipLaunchKernelGGL(psvd::scale_kernel<double>, grid, dim3(psvd::kUtilThreads), 0, st,
                       static_cast<const double *>(x), ldx, sx, static_cast<double *>(y), ldy, sy, rows, cols, d,
                       invert, s);
  else
    return PSVD_EINVAL;
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

int psvd_axpby(int dtype, int64_t batch, int rows, int cols, double alpha, const void *x, int64_t ldx, int64_t sx,
               double beta, void *y, int64_t ldy, int64_t sy, psvd_stream_t stream) {
  if (batch < 0 || rows < 0 || cols < 0 || !x || !y || batch > 65535) return PSVD_EINVAL;
  if (batch == 0 || rows == 0 || cols == 0) return PSVD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 grid(psvd::grid_chunks((int64_t)rows * cols), (unsigned)batch);
  if (dtype == PSVD_F32)
    hipLaunchKernelGGL(psvd::axpby_kernel<float>, grid, dim3(psvd::kUtilThreads), 0, st, (float)alpha,
                       static_cast<const float *>(x), ldx, sx, (float)beta, static_cast<float *>(y), ldy, sy, rows, cols);
  else if (dtype == PSVD_F64)
    hipLaunchKernelGGL(psvd::axpby_kernel<double>, grid, dim3(psvd::kUtilThreads), 0, st, alpha,
                       static_cast<const double *>(x), ldx, sx, beta, static_cast<double *>(y), ldy, sy, rows, cols);
  else
    return PSVD_EINVAL;
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

int psvd_symmetrize(int dtype, int64_t batch, int n, void *a, int64_t lda, int64_t sa, psvd_stream_t stream) {
  if (batch < 0 || n < 0 || !a || batch > 65535) return PSVD_EINVAL;
  if (batch == 0 || n == 0) return PSVD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 grid(psvd::grid_chunks((int64_t)n * n), (unsigned)batch);
  if (dtype == PSVD_F32)
    hipLaunchKernelGGL(psvd::symmetrize_kernel<float>, grid, dim3(psvd::kUtilThreads), 0, st, static_cast<float *>(a),
                       lda, sa, n);
  else if (dtype == PSVD_F64)
    hipLaunchKernelGGL(psvd::symmetrize_kernel<double>, grid, dim3(psvd::kUtilThreads), 0, st,
                       static_cast<double *>(a), lda, sa, n);
  else
    return PSVD_EINVAL;
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

int psvd_sort_desc(int dtype, int64_t batch, int rows, int n, const double *keys, void *s, const void *x, int64_t ldx,
                   int64_t sx, void *y, int64_t ldy, int64_t sy, int32_t *perm, psvd_stream_t stream) {
  if (batch < 0 || rows < 0 || n < 0 || !keys || !s || !perm || ldx < n || ldy < n || batch > 65535 ||
      (rows > 0 && (!x || !y)))
    return PSVD_EINVAL;
  if (batch == 0 || n == 0) return PSVD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(psvd::order_desc_kernel, dim3((unsigned)((n + psvd::kUtilThreads - 1) / psvd::kUtilThreads),
                                                    (unsigned)batch),
                     dim3(psvd::kUtilThreads), 0, st, keys, n, perm);
  const int r = rows > 0 ? rows : 1;
  const dim3 grid((unsigned)(((int64_t)r * n + psvd::kUtilThreads - 1) / psvd::kUtilThreads), (unsigned)batch);
  if (dtype == PSVD_F32)
    hipLaunchKernelGGL(psvd::gather_cols_kernel<float>, grid, dim3(psvd::kUtilThreads), 0, st, keys, perm, rows, n,
                       static_cast<const float *>(x), ldx, sx, static_cast<float *>(y), ldy, sy,
                       static_cast<float *>(s));
  else if (dtype == PSVD_F64)
    hipLaunchKernelGGL(psvd::gather_cols_kernel<double>, grid, dim3(psvd::kUtilThreads), 0, st, keys, perm, rows, n,
                       static_cast<const double *>(x), ldx, sx, static_cast<double *>(y), ldy, sy,
                       static_cast<double *>(s));
  else
    return PSVD_EINVAL;
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

int psvd_scale_columns(int dtype, int64_t batch, int rows, int cols, void *x, int64_t ldx, int64_t sx, const void *d,
                       int64_t sd, psvd_stream_t stream) {
  if (batch < 0 || rows < 0 || cols < 0 || batch > 65535 || ldx < cols || sd < cols) return PSVD_EINVAL;
  if (batch == 0 || rows == 0 || cols == 0) return PSVD_OK;
  if (!x || !d) return PSVD_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 grid(psvd::grid_chunks((int64_t)rows * cols), (unsigned)batch);
  if (dtype == PSVD_F32)
    hipLaunchKernelGGL(psvd::scale_cols_kernel<float>, grid, dim3(psvd::kUtilThreads), 0, st, static_cast<float *>(x),
                       ldx, sx, rows, cols, static_cast<const float *>(d), sd);
  else if (dtype == PSVD_F64)
    hipLaunchKernelGGL(psvd::scale_cols_kernel<double>, grid, dim3(psvd::kUtilThreads), 0, st,
                       static_cast<double *>(x), ldx, sx, rows, cols, static_cast<const double *>(d), sd);
  else
    return PSVD_EINVAL;
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

}  // extern "C"
