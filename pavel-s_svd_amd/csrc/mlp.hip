// K8: the learned QDWH lower bound -- a 32 -> 64 -> 64 -> 1 tanh MLP that maps sketched
// features of A to log10(sigma_min / |A|_F) -- with hand-written inference and training
// kernels, and the column-norm statistics among its features.  No reference counterpart: the
// reference starts QDWH from l0 = eps (svdbench/standalone/solvers.py:128-130); the prediction only
// replaces that starting bound.
//
// Parameters (FP32, one flat vector of kMlpParams): W1 (64 x 32), b1 (64), W2 (64 x 64), b2 (64),
// w3 (64), b3 (1); row-major, W x convention.  Loss = (1/2B) sum_i (y_i - t_i)^2.
//
// mlp_kernel: one wavefront per block of samples, lane j = hidden unit j, weights in registers
// (row j of W1 and W2) and W2 also in LDS for the transposed product of the backward pass.  With
// targets it accumulates the gradient of its samples in registers and writes one partial gradient
// per workgroup; mlp_grad_reduce adds the partials in workgroup order (deterministic).
#include <float.h>

#include "common.hpp"

namespace psvd {

constexpr int kIn = 32, kHid = 64;
constexpr int kMlpParams = kHid * kIn + kHid + kHid * kHid + kHid + kHid + 1;
constexpr int kOffB1 = kHid * kIn, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid, kOffW3 = kOffB2 + kHid,
              kOffB3 = kOffW3 + kHid;
constexpr int kSamplesPerGroup = 16;

__global__ __launch_bounds__(64) void mlp_kernel(const float *prm, const float *x, const float *t, float *y,
                                                 float *gpart, double *lpart, int batch) {
  __shared__ float W2s[kHid][kHid + 1];
  __shared__ float xs[kIn];
  __shared__ float h1s[kHid];
  __shared__ float dz2s[kHid];
  const int j = threadIdx.x;
  float w1[kIn], w2[kHid];
#pragma unroll
  for (int k = 0; k < kIn; ++k) w1[k] = prm[j * kIn + k];
#pragma unroll
  for (int k = 0; k < kHid; ++k) {
    w2[k] = prm[kOffW2 + j * kHid + k];
    W2s[j][k] = w2[k];
  }
  const float b1 = prm[kOffB1 + j], b2 = prm[kOffB2 + j], w3 = prm[kOffW3 + j], b3 = prm[kOffB3];
  const bool train = t != nullptr;
  float g1[kIn], g2[kHid], gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f;
  double loss = 0.0;
  if (train) {
#pragma unroll
    for (int k = 0; k < kIn; ++k) g1[k] = 0.f;
#pragma unroll
    for (int k = 0; k < kHid; ++k) g2[k] = 0.f;
  }
  const int s0 = blockIdx.x * kSamplesPerGroup;
  for (int s = s0; s < s0 + kSamplesPerGroup && s < batch; ++s) {
    __syncthreads();
    if (j < kIn) xs[j] = x[(int64_t)s * kIn + j];
    __syncthreads();
    float z1 = b1;
#pragma unroll
    for (int k = 0; k < kIn; ++k) z1 = fmaf(w1[k], xs[k], z1);
    const float h1 = tanhf(z1);
    h1s[j] = h1;
    __syncthreads();
    float z2 = b2;
#pragma unroll
    for (int k = 0; k < kHid; ++k) z2 = fmaf(w2[k], h1s[k], z2);
    const float h2 = tanhf(z2);
    float out = w3 * h2;
    for (int o = 32; o > 0; o >>= 1) out += __shfl_xor(out, o, 64);
    out += b3;
    if (y && j == 0) y[s] = out;
    if (!train) continue;
    const float dy = (out - t[s]) / (float)batch;
    if (j == 0) loss += 0.5 * (double)(out - t[s]) * (double)(out - t[s]);
    gw3 = fmaf(dy, h2, gw3);
    gb3 += dy;
    const float dz2 = dy * w3 * (1.f - h2 * h2);
    gb2 += dz2;
#pragma unroll
    for (int k = 0; k < kHid; ++k) g2[k] = fmaf(dz2, h1s[k], g2[k]);
    dz2s[j] = dz2;
    __syncthreads();
    float dh1 = 0.f;
#pragma unroll
    for (int i = 0; i < kHid; ++i) dh1 = fmaf(W2s[i][j], dz2s[i], dh1);
    const float dz1 = dh1 * (1.f - h1 * h1);
    gb1 += dz1;
#pragma unroll
    for (int k = 0; k < kIn; ++k) g1[k] = fmaf(dz1, xs[k], g1[k]);
  }
  if (!train) return;
  float *g = gpart + (int64_t)blockIdx.x * kMlpParams;
#pragma unroll
  for (int k = 0; k < kIn; ++k) g[j * kIn + k] = g1[k];
  g[kOffB1 + j] = gb1;
#pragma unroll
  for (int k = 0; k < kHid; ++k) g[kOffW2 + j * kHid + k] = g2[k];
  g[kOffB2 + j] = gb2;
  g[kOffW3 + j] = gw3;
  if (j == 0) {
    g[kOffB3] = gb3;
    lpart[blockIdx.x] = loss;
  }
}

// grad[p] = sum over workgroups (in order) of gpart[w][p]; loss = sum of lpart / batch.
__global__ void mlp_grad_reduce_kernel(const float *gpart, const double *lpart, int groups, int batch, float *grad,
                                       double *loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kMlpParams) {
    float acc = 0.f;
    for (int w = 0; w < groups; ++w) acc += gpart[(int64_t)w * kMlpParams + p];
    grad[p] = acc;
  }
  if (p == 0 && loss) {
    double l = 0.0;
    for (int w = 0; w < groups; ++w) l += lpart[w];
    *loss = l / batch;
  }
}

// Adam (Kingma and Ba, 2015) with bias correction; step >= 1.
__global__ void adam_kernel(float *prm, const float *grad, float *m, float *v, int n, float lr, float b1, float b2,
                            float eps, int step) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float g = grad[p];
  const float mp = b1 * m[p] + (1.f - b1) * g, vp = b2 * v[p] + (1.f - b2) * g * g;
  m[p] = mp;
  v[p] = vp;
  const float mh = mp / (1.f - powf(b1, (float)step)), vh = vp / (1.f - powf(b2, (float)step));
  prm[p] -= lr * mh / (sqrtf(vh) + eps);
}

// Column-norm statistics of each n x n (row-major, ld lda) matrix: out[b] = {min, max, mean, std}
// of log10 |A(:, j)| over the columns (zero columns count as 1e-30) and log10 of max|a| (5 values).
template <typename T>
__global__ __launch_bounds__(256) void colstats_kernel(const T *a, int64_t lda, int64_t sa, int rows, int cols,
                                                       float *out) {
  __shared__ double red[8];
  const T *A = a + (int64_t)blockIdx.x * sa;
  double lmin = DBL_MAX, lmax = -DBL_MAX, lsum = 0.0, lsq = 0.0, amax = 0.0;
  for (int c = threadIdx.x; c < cols; c += 256) {
    double s = 0.0;
    for (int r = 0; r < rows; ++r) {
      const double v = (double)A[(int64_t)r * lda + c];
      s = fma(v, v, s);
      amax = fmax(amax, fabs(v));
    }
    const double l = 0.5 * log10(fmax(s, 1e-60));
    lmin = fmin(lmin, l);
    lmax = fmax(lmax, l);
    lsum += l;
    lsq += l * l;
  }
  lsum = block_sum<256>(lsum, red);
  lsq = block_sum<256>(lsq, red);
  amax = block_max<256>(amax, red);
  lmax = block_max<256>(lmax, red);
  lmin = -block_max<256>(-lmin, red);
  if (threadIdx.x == 0) {
    const double mean = lsum / cols, var = fmax(lsq / cols - mean * mean, 0.0);
    float *o = out + (int64_t)blockIdx.x * 5;
    o[0] = (float)lmin;
    o[1] = (float)lmax;
    o[2] = (float)mean;
    o[3] = (float)sqrt(var);
    o[4] = (float)log10(fmax(amax, 1e-30));
  }
}

}  // namespace psvd

extern "C" {

int psvd_mlp_param_count(void) { return psvd::kMlpParams; }

size_t psvd_mlp_workspace_bytes(int64_t batch) {
  if (batch < 0) return 0;
  const int64_t groups = (batch + psvd::kSamplesPerGroup - 1) / psvd::kSamplesPerGroup;
  return (size_t)groups * (sizeof(float) * psvd::kMlpParams + sizeof(double)) + 256;
}

int psvd_mlp_forward(const float *params, const float *x, int64_t batch, float *y, psvd_stream_t stream) {
  if (batch < 0 || !params || !x || !y || batch > 0x7fffffff) return PSVD_EINVAL;
  if (batch == 0) return PSVD_OK;
  const int64_t groups = (batch + psvd::kSamplesPerGroup - 1) / psvd::kSamplesPerGroup;
  hipLaunchKernelGGL(psvd::mlp_kernel, dim3((unsigned)groups), dim3(64), 0, static_cast<hipStream_t>(stream), params, x,
                     nullptr, y, nullptr, nullptr, (int)batch);
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

int psvd_mlp_grad(const float *params, const float *x, const float *target, int64_t batch, float *y, float *grad,
                  double *loss, void *workspace, size_t workspace_bytes, psvd_stream_t stream) {
  if (batch < 1 || !params || !x || !target || !grad || !workspace || batch > 0x7fffffff) return PSVD_EINVAL;
  if (workspace_bytes < psvd_mlp_workspace_bytes(batch)) return PSVD_EINVAL;
  const int64_t groups = (batch + psvd::kSamplesPerGroup - 1) / psvd::kSamplesPerGroup;
  float *gpart = static_cast<float *>(workspace);
  double *lpart = reinterpret_cast<double *>(
      (reinterpret_cast<uintptr_t>(gpart + groups * psvd::kMlpParams) + 7) & ~uintptr_t(7));
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(psvd::mlp_kernel, dim3((unsigned)groups), dim3(64), 0, st, params, x, target, y, gpart, lpart,
                     (int)batch);
  hipLaunchKernelGGL(psvd::mlp_grad_reduce_kernel, dim3((psvd::kMlpParams + 255) / 256), dim3(256), 0, st, gpart, lpart,
                     (int)groups, (int)batch, grad, loss);
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

int psvd_adam_step(float *params, const float *grad, float *m, float *v, int n, float lr, float beta1, float beta2,
                   float eps, int step, psvd_stream_t stream) {
  if (n < 0 || step < 1 || !params || !grad || !m || !v) return PSVD_EINVAL;
  if (n == 0) return PSVD_OK;
  hipLaunchKernelGGL(psvd::adam_kernel, dim3((n + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream), params,
                     grad, m, v, n, lr, beta1, beta2, eps, step);
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

int psvd_colstats(int dtype, int64_t batch, int rows, int cols, const void *a, int64_t lda, int64_t stride_a,
                  float *out, psvd_stream_t stream) {
  if (batch < 0 || rows < 1 || cols < 1 || lda < cols || !a || !out || batch > 0x7fffffff) return PSVD_EINVAL;
  if (batch == 0) return PSVD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == PSVD_F32)
    hipLaunchKernelGGL(psvd::colstats_kernel<float>, dim3((unsigned)batch), dim3(256), 0, st,
                       static_cast<const float *>(a), lda, stride_a, rows, cols, out);
  else if (dtype == PSVD_F64)
    hipLaunchKernelGGL(psvd::colstats_kernel<double>, dim3((unsigned)batch), dim3(256), 0, st,
                       static_cast<const double *>(a), lda, stride_a, rows, cols, out);
  else
    return PSVD_EINVAL;
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

}  // extern "C"
