// K4c: Householder QR of a batch of small complex matrices, each resident in LDS, and the small
// batched complex product that forms U = Q U_R after the Jacobi solve of R.
//
// One workgroup per matrix.  A (m x n, m >= n; row-major with interleaved (re, im) pairs, as torch
// stores complex tensors) is held column-major in LDS as a real and an imaginary plane.  Column j
// takes the reference's complex xLARFG reflector (primitives.py:41-59):
//   beta = -sign(Re alpha) hypot(|alpha|, |x2|) (real; sign(0) taken as +1),
//   tau = (beta - alpha) / beta,  v = x / (alpha - beta),  v_0 = 1,
// and the trivial reflector (tau = 0, v = 0, beta = Re alpha) when x2 = 0 and alpha is real.  The
// trailing columns get A <- (I - conj(tau) v v^H) A (primitives.py:62-65, :79-83), one group of 16
// lanes per column, with the v^H A sums in FP64.  The reflectors stay below the diagonal.  Q (thin,
// m x n) is then accumulated backwards from [I; 0] with A <- (I - tau v v^H) A
// (primitives.py:84-88); R is the upper triangle with the real diagonal beta.  Matrices whose
// planes exceed LDS run the same kernel with the planes in a global workspace (one region per
// matrix, served by L2): the same arithmetic and order, so the same bits.
#include <float.h>

#include "common.hpp"

namespace psvd {

constexpr int kCqrThreads = 256;

__host__ __device__ inline int cqr_ld(int m) { return m | 1; }

template <typename R>
__host__ __device__ inline size_t cqr_lds_bytes(int m, int n) {
  return sizeof(double) * 2 * (size_t)n + sizeof(R) * 4 * (size_t)n * cqr_ld(m);
}

constexpr size_t kCqrLdsCap = 160 * 1024 - 4096;

// Columns are processed by groups of kCqrGroup lanes (one DPP row each), so a wave works on four
// columns at once and a column's sums are reduced by DPP within its row.
constexpr int kCqrGroup = 16;

// x <- x - v z for the rows i >= j of one column (planes xr, xi), v from the reflector planes
// (v_j = 1), z complex; lig is the lane's index in its group.
template <typename R>
__device__ __forceinline__ void cqr_axpy(R *xr, R *xi, const R *vr, const R *vi, int j, int m, int lig, double zr,
                                         double zi) {
  for (int i = j + lig; i < m; i += kCqrGroup) {
    const double a = i == j ? 1.0 : (double)vr[i], b = i == j ? 0.0 : (double)vi[i];
    xr[i] = (R)((double)xr[i] - (a * zr - b * zi));
    xi[i] = (R)((double)xi[i] - (a * zi + b * zr));
  }
}

// w = v^H x over the rows i >= j (v_j = 1), summed in FP64 over the group; every lane of the group
// gets it.
template <typename R>
__device__ __forceinline__ void cqr_dot(const R *xr, const R *xi, const R *vr, const R *vi, int j, int m, int lig,
                                        double &wr, double &wi) {
  double sr = 0.0, si = 0.0;
  for (int i = j + lig; i < m; i += kCqrGroup) {
    const double a = i == j ? 1.0 : (double)vr[i], b = i == j ? 0.0 : (double)vi[i];
    const double x = xr[i], y = xi[i];
    sr = fma(a, x, fma(b, y, sr));    // conj(v) x = (a - ib)(x + iy)
    si = fma(a, y, fma(-b, x, si));
  }
  wr = group_sum_dpp<kCqrGroup>(sr);
  wi = group_sum_dpp<kCqrGroup>(si);
}

template <typename R, bool GL>
__global__ __launch_bounds__(kCqrThreads) void cqr_lds_kernel(const R *a, R *q, R *r, int m, int n, double *gws,
                                                             int64_t gstride) {
  extern __shared__ double smem_cq[];
  __shared__ double red[kCqrThreads / kWave];
  const int ld = cqr_ld(m);
  double *base = GL ? gws + (int64_t)blockIdx.x * gstride : smem_cq;
  double *taur = base, *taui = taur + n;
  R *Ar = reinterpret_cast<R *>(taui + n), *Ai = Ar + (size_t)n * ld;
  R *Qr = Ai + (size_t)n * ld, *Qi = Qr + (size_t)n * ld;
  constexpr int NG = kCqrThreads / kCqrGroup;   // column groups in the workgroup
  const int tid = threadIdx.x, grp = tid / kCqrGroup, lig = tid - grp * kCqrGroup;
  const int64_t b = blockIdx.x;
  const R *A = a + b * 2 * (int64_t)m * n;
  for (int e = tid; e < m * n; e += kCqrThreads) {
    const int i = e / n, j = e - i * n;
    Ar[j * ld + i] = A[2 * e];
    Ai[j * ld + i] = A[2 * e + 1];
  }
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    R *vr = Ar + j * ld, *vi = Ai + j * ld;
    double part = 0.0;
    for (int i = j + 1 + tid; i < m; i += kCqrThreads) {
      const double x = vr[i], y = vi[i];
      part = fma(x, x, fma(y, y, part));
    }
    const double xnorm = sqrt(block_sum<kCqrThreads>(part, red));
    const double ar = vr[j], ai = vi[j];
    const bool trivial = xnorm == 0.0 && ai == 0.0;
    double beta = ar, tr = 0.0, ti = 0.0;
    if (!trivial) {
      beta = -(ar < 0.0 ? -1.0 : 1.0) * hypot(hypot(ar, ai), xnorm);
      tr = (beta - ar) / beta;
      ti = -ai / beta;
      // v = x / (alpha - beta)
      const double dr = ar - beta, di = ai, den = dr * dr + di * di;
      const double ir = dr / den, ii = -di / den;
      for (int i = j + 1 + tid; i < m; i += kCqrThreads) {
        const double x = vr[i], y = vi[i];
        vr[i] = (R)(x * ir - y * ii);
        vi[i] = (R)(x * ii + y * ir);
      }
    }
    __syncthreads();   // every thread has read alpha and written its part of v
    if (tid == 0) {
      vr[j] = (R)beta;
      vi[j] = R(0);
      taur[j] = tr;
      taui[j] = ti;
    }
    if (!trivial) {   // A(j:, c) <- A(j:, c) - conj(tau) v (v^H A(j:, c)), c > j, one group per column
      for (int c = j + 1 + grp; c < n; c += NG) {
        R *xr = Ar + c * ld, *xi = Ai + c * ld;
        double wr, wi;
        cqr_dot(xr, xi, vr, vi, j, m, lig, wr, wi);
        cqr_axpy(xr, xi, vr, vi, j, m, lig, tr * wr + ti * wi, tr * wi - ti * wr);   // conj(tau) w
      }
    }
    __syncthreads();
  }
  // Q = H_0 ... H_{n-1} [I; 0], accumulated backwards; rows >= j of columns < j are still zero
  for (int e = tid; e < m * n; e += kCqrThreads) {
    const int c = e / m, i = e - c * m;
    Qr[c * ld + i] = i == c ? R(1) : R(0);
    Qi[c * ld + i] = R(0);
  }
  __syncthreads();
  for (int j = n - 1; j >= 0; --j) {
    const double tr = taur[j], ti = taui[j];
    if (tr != 0.0 || ti != 0.0) {
      const R *vr = Ar + j * ld, *vi = Ai + j * ld;
      for (int c = j + grp; c < n; c += NG) {
        R *xr = Qr + c * ld, *xi = Qi + c * ld;
        double wr, wi;
        cqr_dot(xr, xi, vr, vi, j, m, lig, wr, wi);
        cqr_axpy(xr, xi, vr, vi, j, m, lig, tr * wr - ti * wi, tr * wi + ti * wr);   // tau w
      }
    }
    __syncthreads();
  }
  R *Qo = q + b * 2 * (int64_t)m * n, *Ro = r + b * 2 * (int64_t)n * n;
  for (int e = tid; e < m * n; e += kCqrThreads) {
    const int i = e / n, c = e - i * n;
    Qo[2 * e] = Qr[c * ld + i];
    Qo[2 * e + 1] = Qi[c * ld + i];
  }
  for (int e = tid; e < n * n; e += kCqrThreads) {
    const int i = e / n, c = e - i * n;
    Ro[2 * e] = i <= c ? Ar[c * ld + i] : R(0);
    Ro[2 * e + 1] = i <= c ? Ai[c * ld + i] : R(0);
  }
}

// C (m x n) = op(A) op(B) for each matrix of the batch, row-major interleaved complex, one thread
// per entry of C with the sum in FP64.  op(A) is A (stored m x k) or, with ta, A^H (A stored
// k x m); op(B) is B (stored k x n) or, with tb, B^H (B stored n x k).
template <typename R>
__global__ __launch_bounds__(256) void cmatmul_kernel(const R *a, const R *bm, R *c, int m, int k, int n, int ta,
                                                      int tb) {
  const int64_t bb = blockIdx.y;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= m * n) return;
  const int i = e / n, j = e - i * n;
  const R *A = a + bb * 2 * (int64_t)m * k;
  const R *B = bm + bb * 2 * (int64_t)k * n;
  // element t of row i of op(A) and of column j of op(B): (offset, stride) in complex entries
  const int64_t a0 = ta ? i : (int64_t)i * k, as = ta ? m : 1;
  const int64_t b0 = tb ? (int64_t)j * k : j, bs = tb ? 1 : n;
  const double ca = ta ? -1.0 : 1.0, cb = tb ? -1.0 : 1.0;   // conjugation signs of the imaginary parts
  double sr = 0.0, si = 0.0;
  for (int t = 0; t < k; ++t) {
    const double xr = A[2 * (a0 + t * as)], xi = ca * A[2 * (a0 + t * as) + 1];
    const double yr = B[2 * (b0 + t * bs)], yi = cb * B[2 * (b0 + t * bs) + 1];
    sr = fma(xr, yr, fma(-xi, yi, sr));
    si = fma(xr, yi, fma(xi, yr, si));
  }
  R *C = c + bb * 2 * (int64_t)m * n;
  C[2 * e] = (R)sr;
  C[2 * e + 1] = (R)si;
}

// a <- (a + a^H) / 2 for each n x n complex matrix of the batch (contiguous, interleaved), one
// thread per entry pair (i <= j); the diagonal becomes real.
template <typename R>
__global__ __launch_bounds__(256) void chermitize_kernel(R *a, int n) {
  const int64_t bb = blockIdx.y;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n * n) return;
  const int i = e / n, j = e - i * n;
  if (i > j) return;
  R *A = a + bb * 2 * (int64_t)n * n;
  R *x = A + 2 * ((int64_t)i * n + j), *y = A + 2 * ((int64_t)j * n + i);
  const R re = (x[0] + y[0]) / R(2), im = (x[1] - y[1]) / R(2);
  x[0] = re;
  x[1] = im;
  y[0] = re;
  y[1] = -im;
}

// Completion of missing complex basis columns (primitives.py:430-455): for each matrix, columns
// [nk, n) of U (m x n, row-major interleaved) are replaced in order by the coordinate vector e_ix
// at the row of largest remaining leverage 1 - sum_{c<r} |u_ic|^2 (first such row), projected
// twice as v <- v - U (U^H v) against the columns before it and normalised; a residual below
// 10 eps sets status PSVD_INFO_COMPLETION.  One workgroup per matrix, v in LDS, FP64.
constexpr int kCcThreads = 256;

template <typename R>
__global__ __launch_bounds__(kCcThreads) void ccomplete_kernel(R *u, int m, int n, const int32_t *nk,
                                                               int32_t *status, double eps) {
  extern __shared__ double smem_cc[];
  double *vr = smem_cc, *vi = vr + m, *yr = vi + m, *yi = yr + n;
  __shared__ double s_best[kCcThreads / kWave];
  __shared__ int s_bi[kCcThreads / kWave];
  __shared__ double s_red[kCcThreads / kWave];
  constexpr int NW = kCcThreads / kWave;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int64_t b = blockIdx.x;
  R *U = u + b * 2 * (int64_t)m * n;
  auto re = [&](int i, int c) -> R & { return U[2 * ((int64_t)i * n + c)]; };
  auto im = [&](int i, int c) -> R & { return U[2 * ((int64_t)i * n + c) + 1]; };
  int st = PSVD_INFO_OK;
  const int k0 = nk[b] < 0 ? 0 : (nk[b] > n ? n : nk[b]);
  for (int r = k0; r < n; ++r) {
    // row of largest leverage (ties: the first)
    double best = -DBL_MAX;
    int bi = 0x7fffffff;
    for (int i = tid; i < m; i += kCcThreads) {
      double acc = 0.0;
      for (int c = 0; c < r; ++c) {
        const double x = re(i, c), y = im(i, c);
        acc = fma(x, x, fma(y, y, acc));
      }
      const double lev = 1.0 - acc;
      if (lev > best || (lev == best && i < bi)) {
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
    if (lane == 0) {
      s_best[wave] = best;
      s_bi[wave] = bi;
    }
    __syncthreads();
    best = s_best[0];
    bi = s_bi[0];
    for (int w = 1; w < NW; ++w)
      if (s_best[w] > best || (s_best[w] == best && s_bi[w] < bi)) {
        best = s_best[w];
        bi = s_bi[w];
      }
    const int ix = wave_uniform(bi);
    for (int i = tid; i < m; i += kCcThreads) {
      vr[i] = i == ix ? 1.0 : 0.0;
      vi[i] = 0.0;
    }
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass) {
      // y = U^H v over the columns before r, one wave per column
      for (int c = wave; c < r; c += NW) {
        double ar = 0.0, ai = 0.0;
        for (int i = lane; i < m; i += kWave) {
          const double ur = re(i, c), ui = im(i, c);
          ar = fma(ur, vr[i], fma(ui, vi[i], ar));
          ai = fma(ur, vi[i], fma(-ui, vr[i], ai));
        }
        ar = group_sum_dpp<kWave>(ar);
        ai = group_sum_dpp<kWave>(ai);
        if (lane == 0) {
          yr[c] = ar;
          yi[c] = ai;
        }
      }
      __syncthreads();
      // v = v - U y
      for (int i = tid; i < m; i += kCcThreads) {
        double ar = vr[i], ai = vi[i];
        for (int c = 0; c < r; ++c) {
          const double ur = re(i, c), ui = im(i, c);
          ar -= ur * yr[c] - ui * yi[c];
          ai -= ur * yi[c] + ui * yr[c];
        }
        vr[i] = ar;
        vi[i] = ai;
      }
      __syncthreads();
    }
    double l2 = 0.0;
    for (int i = tid; i < m; i += kCcThreads) l2 = fma(vr[i], vr[i], fma(vi[i], vi[i], l2));
    const double len = sqrt(block_sum<kCcThreads>(l2, s_red));
    if (len < 10.0 * eps) st = PSVD_INFO_COMPLETION;
    const double inv = 1.0 / fmax(len, DBL_MIN);
    for (int i = tid; i < m; i += kCcThreads) {
      re(i, r) = (R)(vr[i] * inv);
      im(i, r) = (R)(vi[i] * inv);
    }
    __syncthreads();
  }
  if (tid == 0) status[b] = st;
}

// Per-matrix region of the global-workspace variant (doubles, 256-byte aligned).
template <typename R>
static int64_t cqr_gstride(int m, int n) {
  return (int64_t)((cqr_lds_bytes<R>(m, n) + 255) / 256 * 32);
}

template <typename R>
static int run_cqr(int64_t batch, int m, int n, const void *a, void *q, void *r, void *ws, size_t ws_bytes,
                   hipStream_t st) {
  const size_t smem = cqr_lds_bytes<R>(m, n);
  if (smem > kCqrLdsCap) {   // planes in the workspace
    const int64_t gs = cqr_gstride<R>(m, n);
    if (!ws || ws_bytes < (size_t)(batch * gs) * sizeof(double)) return ws ? PSVD_EINVAL : PSVD_ETOOBIG;
    hipLaunchKernelGGL((cqr_lds_kernel<R, true>), dim3((unsigned)batch), dim3(kCqrThreads), 0, st,
                       static_cast<const R *>(a), static_cast<R *>(q), static_cast<R *>(r), m, n,
                       static_cast<double *>(ws), gs);
    return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
  }
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute((const void *)cqr_lds_kernel<R, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)kCqrLdsCap) != hipSuccess)
      return PSVD_ELAUNCH;
    attr = true;
  }
  hipLaunchKernelGGL((cqr_lds_kernel<R, false>), dim3((unsigned)batch), dim3(kCqrThreads), smem, st,
                     static_cast<const R *>(a), static_cast<R *>(q), static_cast<R *>(r), m, n, nullptr, int64_t(0));
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

}  // namespace psvd

extern "C" {

int psvd_cqr_lds_smem_bytes(int dtype, int m, int n, size_t *bytes) {
  if (m < 1 || n < 1 || n > m || !bytes) return PSVD_EINVAL;
  size_t b;
  if (dtype == PSVD_C64) b = psvd::cqr_lds_bytes<float>(m, n);
  else if (dtype == PSVD_C128) b = psvd::cqr_lds_bytes<double>(m, n);
  else return PSVD_EINVAL;
  *bytes = b;
  return b > psvd::kCqrLdsCap ? PSVD_ETOOBIG : PSVD_OK;
}

int psvd_cqr_lds(int dtype, int64_t batch, int m, int n, const void *a, void *q, void *r, psvd_stream_t stream) {
  if (batch < 0 || batch > 0x7fffffff || m < 1 || n < 1 || n > m || !a || !q || !r) return PSVD_EINVAL;
  if (batch == 0) return PSVD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == PSVD_C64) return psvd::run_cqr<float>(batch, m, n, a, q, r, nullptr, 0, st);
  if (dtype == PSVD_C128) return psvd::run_cqr<double>(batch, m, n, a, q, r, nullptr, 0, st);
  return PSVD_EINVAL;
}

size_t psvd_cqr_workspace_bytes(int dtype, int64_t batch, int m, int n) {
  if (batch < 0 || m < 1 || n < 1 || n > m) return 0;
  if (dtype == PSVD_C64)
    return psvd::cqr_lds_bytes<float>(m, n) > psvd::kCqrLdsCap ? (size_t)(batch * psvd::cqr_gstride<float>(m, n)) * 8
                                                               : 0;
  if (dtype == PSVD_C128)
    return psvd::cqr_lds_bytes<double>(m, n) > psvd::kCqrLdsCap
               ? (size_t)(batch * psvd::cqr_gstride<double>(m, n)) * 8
               : 0;
  return 0;
}

int psvd_cqr(int dtype, int64_t batch, int m, int n, const void *a, void *q, void *r, void *workspace,
             size_t workspace_bytes, psvd_stream_t stream) {
  if (batch < 0 || batch > 0x7fffffff || m < 1 || n < 1 || n > m || !a || !q || !r) return PSVD_EINVAL;
  if (batch == 0) return PSVD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == PSVD_C64) return psvd::run_cqr<float>(batch, m, n, a, q, r, workspace, workspace_bytes, st);
  if (dtype == PSVD_C128) return psvd::run_cqr<double>(batch, m, n, a, q, r, workspace, workspace_bytes, st);
  return PSVD_EINVAL;
}

int psvd_cmatmul(int dtype, int transa, int transb, int64_t batch, int m, int k, int n, const void *a, const void *b,
                 void *c, psvd_stream_t stream) {
  if (batch < 0 || m < 1 || k < 1 || n < 1 || !a || !b || !c) return PSVD_EINVAL;
  if ((transa != 0 && transa != 1) || (transb != 0 && transb != 1)) return PSVD_EINVAL;
  if (dtype != PSVD_C64 && dtype != PSVD_C128) return PSVD_EINVAL;
  if ((int64_t)m * n > 0x7fffffff - 255) return PSVD_ETOOBIG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const size_t es = dtype == PSVD_C64 ? 2 * sizeof(float) : 2 * sizeof(double);
  // the grid's y dimension carries the matrix index: launches of at most 65535 matrices
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    const int64_t nb = batch - b0 < 65535 ? batch - b0 : 65535;
    const dim3 grid((unsigned)(((int64_t)m * n + 255) / 256), (unsigned)nb);
    const char *ap = static_cast<const char *>(a) + (size_t)b0 * m * k * es;
    const char *bp = static_cast<const char *>(b) + (size_t)b0 * k * n * es;
    char *cp = static_cast<char *>(c) + (size_t)b0 * m * n * es;
    if (dtype == PSVD_C64)
      hipLaunchKernelGGL(psvd::cmatmul_kernel<float>, grid, dim3(256), 0, st, reinterpret_cast<const float *>(ap),
                         reinterpret_cast<const float *>(bp), reinterpret_cast<float *>(cp), m, k, n, transa,
                         transb);
    else
      hipLaunchKernelGGL(psvd::cmatmul_kernel<double>, grid, dim3(256), 0, st, reinterpret_cast<const double *>(ap),
                         reinterpret_cast<const double *>(bp), reinterpret_cast<double *>(cp), m, k, n, transa,
                         transb);
    if (hipGetLastError() != hipSuccess) return PSVD_ELAUNCH;
  }
  return PSVD_OK;
}

int psvd_ccomplete_columns(int dtype, int64_t batch, int m, int n, void *u, const int32_t *nk, int32_t *status,
                           psvd_stream_t stream) {
  if (batch < 0 || batch > 0x7fffffff || m < 1 || n < 1 || n > m || !u || !nk || !status) return PSVD_EINVAL;
  if (dtype != PSVD_C64 && dtype != PSVD_C128) return PSVD_EINVAL;
  const size_t smem = sizeof(double) * 2 * ((size_t)m + n);
  if (smem > psvd::kCqrLdsCap) return PSVD_ETOOBIG;
  if (batch == 0) return PSVD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == PSVD_C64) {
    static bool attr = false;
    if (!attr) {
      if (hipFuncSetAttribute((const void *)psvd::ccomplete_kernel<float>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)psvd::kCqrLdsCap) != hipSuccess)
        return PSVD_ELAUNCH;
      attr = true;
    }
    hipLaunchKernelGGL(psvd::ccomplete_kernel<float>, dim3((unsigned)batch), dim3(psvd::kCcThreads), smem, st,
                       static_cast<float *>(u), m, n, nk, status, 1.1920928955078125e-07);
  } else {
    static bool attr = false;
    if (!attr) {
      if (hipFuncSetAttribute((const void *)psvd::ccomplete_kernel<double>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)psvd::kCqrLdsCap) != hipSuccess)
        return PSVD_ELAUNCH;
      attr = true;
    }
    hipLaunchKernelGGL(psvd::ccomplete_kernel<double>, dim3((unsigned)batch), dim3(psvd::kCcThreads), smem, st,
                       static_cast<double *>(u), m, n, nk, status, 2.220446049250313e-16);
  }
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

int psvd_chermitize(int dtype, int64_t batch, int n, void *a, psvd_stream_t stream) {
  if (batch < 0 || n < 1 || !a) return PSVD_EINVAL;
  if (dtype != PSVD_C64 && dtype != PSVD_C128) return PSVD_EINVAL;
  if ((int64_t)n * n > 0x7fffffff - 255) return PSVD_ETOOBIG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const size_t es = dtype == PSVD_C64 ? 2 * sizeof(float) : 2 * sizeof(double);
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    const int64_t nb = batch - b0 < 65535 ? batch - b0 : 65535;
    const dim3 grid((unsigned)(((int64_t)n * n + 255) / 256), (unsigned)nb);
    char *ap = static_cast<char *>(a) + (size_t)b0 * n * n * es;
    if (dtype == PSVD_C64)
      hipLaunchKernelGGL(psvd::chermitize_kernel<float>, grid, dim3(256), 0, st, reinterpret_cast<float *>(ap), n);
    else
      hipLaunchKernelGGL(psvd::chermitize_kernel<double>, grid, dim3(256), 0, st, reinterpret_cast<double *>(ap), n);
    if (hipGetLastError() != hipSuccess) return PSVD_ELAUNCH;
  }
  return PSVD_OK;
}

}  // extern "C"
