// K5: blocked Cholesky factorisation and right-sided triangular solves (batched).
//
// psvd_cholesky: right-looking, kChol-column blocks: the diagonal block is factored by one
// workgroup in LDS (FP64 accumulation, the reference's failure mask -- a pivot <= 0 sets the flag
// and is clamped to the smallest normal, primitives.py:150-158), the panel below it is solved row
// by row against the block (L21 = A21 L11^-T), and the trailing matrix is updated with one GEMM
// (A22 -= L21 L21^T).  The strict upper triangle is cleared at the end.
// psvd_trsm_right: B <- B L^-T (trans) or B <- B L^-1 with L lower triangular; block columns are
// solved row-parallel against the diagonal block and then removed from the unsolved columns with
// one GEMM (right-looking).
// QDWH's x z^-1 = x L^-T L^-1 (solvers.py:159) is two calls, with no transposed copies.
#include <float.h>

#include "common.hpp"

namespace psvd {

constexpr int kChol = 32;

template <typename T>
__global__ __launch_bounds__(64) void chol_diag_kernel(T *A, int64_t lda, int64_t sa, int jb, int nb, int32_t *fail) {
  __shared__ double L[kChol][kChol + 1];
  const int tid = threadIdx.x;
  T *Ab = A + (int64_t)blockIdx.x * sa + (int64_t)jb * lda + jb;
  for (int e = tid; e < nb * nb; e += 64) L[e / nb][e % nb] = (double)Ab[(int64_t)(e / nb) * lda + e % nb];
  __syncthreads();
  int bad = 0;
  for (int j = 0; j < nb; ++j) {
    // every thread computes the pivot (uniform), then the column below it in parallel
    double s = 0.0;
    for (int k = 0; k < j; ++k) s = fma(L[j][k], L[j][k], s);
    const double piv = L[j][j] - s;
    bad |= !(piv > 0.0);
    const double ljj = sqrt(fmax(piv, DBL_MIN));
    __syncthreads();
    if (tid == 0) L[j][j] = ljj;
    for (int i = j + 1 + tid; i < nb; i += 64) {
      double r = L[i][j];
      for (int k = 0; k < j; ++k) r -= L[i][k] * L[j][k];
      L[i][j] = r / ljj;
    }
    __syncthreads();
  }
  for (int e = tid; e < nb * nb; e += 64) {
    const int i = e / nb, k = e % nb;
    Ab[(int64_t)i * lda + k] = k <= i ? (T)L[i][k] : T(0);
  }
  if (tid == 0 && bad) atomicOr(&fail[blockIdx.x], 1);
}

// rows x nb block X (ld ldx) <- X L11^-T (trans: solve x L11^T = r, forward) or X L11^-1
// (solve x L11 = r, backward); L11 is the nb x nb lower-triangular block at `Lb` (ld ldl).
template <typename T>
__global__ __launch_bounds__(256) void rsolve_kernel(const T *L, int64_t ldl, int64_t sl, T *X, int64_t ldx,
                                                     int64_t sx, int rows, int nb, int trans) {
  __shared__ double Ls[kChol][kChol + 1];
  const int tid = threadIdx.x;
  const T *Lb = L + (int64_t)blockIdx.y * sl;
  for (int e = tid; e < nb * nb; e += 256) Ls[e / nb][e % nb] = (double)Lb[(int64_t)(e / nb) * ldl + e % nb];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * 256 + tid;
  if (i >= rows) return;
  T *row = X + (int64_t)blockIdx.y * sx + i * ldx;
  double x[kChol];
#pragma unroll
  for (int k = 0; k < kChol; ++k) x[k] = k < nb ? (double)row[k] : 0.0;
  if (trans) {
    // x L^T = r  <=>  L x^T = r^T: forward substitution
#pragma unroll
    for (int k = 0; k < kChol; ++k) {
      if (k < nb) {
        double s = x[k];
#pragma unroll
        for (int l = 0; l < kChol; ++l)
          if (l < k) s -= Ls[k][l] * x[l];
        x[k] = s / Ls[k][k];
      }
    }
  } else {
    // x L = r  <=>  L^T x^T = r^T: backward substitution
#pragma unroll
    for (int k = kChol - 1; k >= 0; --k) {
      if (k < nb) {
        double s = x[k];
#pragma unroll
        for (int l = 0; l < kChol; ++l)
          if (l > k && l < nb) s -= Ls[l][k] * x[l];
        x[k] = s / Ls[k][k];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kChol; ++k)
    if (k < nb) row[k] = (T)x[k];
}

template <typename T>
__global__ void tril_kernel(T *A, int64_t lda, int64_t sa, int n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  const int i = (int)(e / n), k = (int)(e - (int64_t)i * n);
  if (k > i) A[(int64_t)blockIdx.y * sa + (int64_t)i * lda + k] = T(0);
}

__global__ void zero_i32_kernel(int32_t *p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

#define PSVD_TRY(x)                 \
  do {                              \
    const int _st = (x);            \
    if (_st != PSVD_OK) return _st; \
  } while (0)

static int launched() { return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH; }

template <typename T>
static int run_rsolve(const T *L, int64_t ldl, int64_t sl, T *X, int64_t ldx, int64_t sx, int rows, int nb, int trans,
                      int64_t batch, hipStream_t st) {
  if (rows <= 0) return PSVD_OK;
  hipLaunchKernelGGL(rsolve_kernel<T>, dim3((unsigned)((rows + 255) / 256), (unsigned)batch), dim3(256), 0, st, L, ldl,
                     sl, X, ldx, sx, rows, nb, trans);
  return launched();
}

template <typename T>
static int run_cholesky(int gmode, int64_t batch, int n, T *a, int64_t lda, int64_t sa, int32_t *fail,
                        hipStream_t st) {
  hipLaunchKernelGGL(zero_i32_kernel, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, st, fail, batch);
  PSVD_TRY(launched());
  for (int jb = 0; jb < n; jb += kChol) {
    const int nb = n - jb < kChol ? n - jb : kChol, rest = n - jb - nb;
    hipLaunchKernelGGL(chol_diag_kernel<T>, dim3((unsigned)batch), dim3(64), 0, st, a, lda, sa, jb, nb, fail);
    PSVD_TRY(launched());
    if (rest > 0) {
      T *L11 = a + (int64_t)jb * lda + jb, *L21 = a + (int64_t)(jb + nb) * lda + jb;
      PSVD_TRY(run_rsolve<T>(L11, lda, sa, L21, lda, sa, rest, nb, 1, batch, st));
      T *A22 = a + (int64_t)(jb + nb) * lda + jb + nb;
      PSVD_TRY(psvd_gemm(gmode, 0, 1, rest, rest, nb, -1.0, L21, lda, sa, L21, lda, sa, 1.0, A22, lda, sa, batch, st));
    }
  }
  hipLaunchKernelGGL(tril_kernel<T>, dim3((unsigned)(((int64_t)n * n + 255) / 256), (unsigned)batch), dim3(256), 0, st,
                     a, lda, sa, n);
  return launched();
}

// Right-looking: each solved block column updates all the columns still to be solved with one
// GEMM of K = kChol (a wide output that fills the GPU), instead of each new block column gathering
// the solved ones (a skinny output with long K).
template <typename T>
static int run_trsm_right(int gmode, int64_t batch, int trans, int n, int rows, const T *L, int64_t ldl, int64_t sl,
                          T *B, int64_t ldb, int64_t sb, hipStream_t st) {
  if (trans) {  // B L^-T, left to right: solve B[:, J], then B[:, J+1:] -= B[:, J] L[J+1:, J]^T
    for (int jb = 0; jb < n; jb += kChol) {
      const int nb = n - jb < kChol ? n - jb : kChol, rest = n - jb - nb;
      PSVD_TRY(run_rsolve<T>(L + (int64_t)jb * ldl + jb, ldl, sl, B + jb, ldb, sb, rows, nb, 1, batch, st));
      if (rest > 0)
        PSVD_TRY(psvd_gemm(gmode, 0, 1, rows, rest, nb, -1.0, B + jb, ldb, sb, L + (int64_t)(jb + nb) * ldl + jb, ldl,
                           sl, 1.0, B + jb + nb, ldb, sb, batch, st));
    }
  } else {  // B L^-1, right to left: solve B[:, J], then B[:, :J] -= B[:, J] L[J, :J]
    const int nblk = (n + kChol - 1) / kChol;
    for (int bi = nblk - 1; bi >= 0; --bi) {
      const int jb = bi * kChol, nb = n - jb < kChol ? n - jb : kChol;
      PSVD_TRY(run_rsolve<T>(L + (int64_t)jb * ldl + jb, ldl, sl, B + jb, ldb, sb, rows, nb, 0, batch, st));
      if (jb > 0)
        PSVD_TRY(psvd_gemm(gmode, 0, 0, rows, jb, nb, -1.0, B + jb, ldb, sb, L + (int64_t)jb * ldl, ldl, sl, 1.0, B, ldb,
                           sb, batch, st));
    }
  }
  return PSVD_OK;
}

}  // namespace psvd

extern "C" {

int psvd_cholesky(int dtype, int gemm_mode, int64_t batch, int n, void *a, int64_t lda, int64_t stride_a,
                  int32_t *fail, psvd_stream_t stream) {
  if (batch < 0 || n < 1 || lda < n || !a || !fail || batch > 65535) return PSVD_EINVAL;
  if (batch == 0) return PSVD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == PSVD_F32)
    return psvd::run_cholesky<float>(gemm_mode == PSVD_GEMM_BF16X3 ? PSVD_GEMM_BF16X3 : PSVD_GEMM_F32, batch, n,
                                     static_cast<float *>(a), lda, stride_a, fail, st);
  if (dtype == PSVD_F64)
    return psvd::run_cholesky<double>(PSVD_GEMM_F64, batch, n, static_cast<double *>(a), lda, stride_a, fail, st);
  return PSVD_EINVAL;
}

int psvd_trsm_right(int dtype, int gemm_mode, int64_t batch, int trans, int n, int rows, const void *l, int64_t ldl,
                    int64_t stride_l, void *b, int64_t ldb, int64_t stride_b, psvd_stream_t stream) {
  if (batch < 0 || n < 1 || rows < 0 || ldl < n || ldb < n || !l || !b || batch > 65535) return PSVD_EINVAL;
  if (batch == 0 || rows == 0) return PSVD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == PSVD_F32)
    return psvd::run_trsm_right<float>(gemm_mode == PSVD_GEMM_BF16X3 ? PSVD_GEMM_BF16X3 : PSVD_GEMM_F32, batch, trans,
                                       n, rows, static_cast<const float *>(l), ldl, stride_l, static_cast<float *>(b),
                                       ldb, stride_b, st);
  if (dtype == PSVD_F64)
    return psvd::run_trsm_right<double>(PSVD_GEMM_F64, batch, trans, n, rows, static_cast<const double *>(l), ldl,
                                        stride_l, static_cast<double *>(b), ldb, stride_b, st);
  return PSVD_EINVAL;
}

}  // extern "C"
