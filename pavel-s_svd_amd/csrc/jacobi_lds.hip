// K1, batched one-sided Jacobi SVD with each matrix resident in LDS: the FP32 kernel set, the C ABI
// (psvd_jacobi_lds, psvd_jacobi_lds_noalloc) and the non-allocating entry point's device scratch.  The
// kernels are in jacobi_lds_impl.hpp; the FP64 set is instantiated in jacobi_lds_f64.hip.
//
// Follows svdbench/standalone/primitives.py:316-367 (jacobi_svd), :309-313 (schedule),
// :430-455 (complete_null_columns) and solvers.py:294-299 (max-abs scaling).
#include <mutex>

#include "jacobi_lds_impl.hpp"

namespace psvd {
// run_lds<double> (jacobi_lds_f64.hip, its own translation unit so the two kernel sets compile in parallel)
int run_lds_f64(int64_t batch, int mu, int nu, const int32_t *dims, const int64_t *offs, const int32_t *order,
                int max_m, int max_n, const void *a, void *u, void *s, void *vh, int32_t *info, double tol,
                int max_sweeps, int flags, int lanes, int threads, void *ws, size_t ws_bytes, hipStream_t stream);
}  // namespace psvd


// ---- psvd_jacobi_lds_noalloc: the same launch with its scratch in a file-scope device array --------
// The only scratch of psvd_jacobi_lds is the device work-queue counter (one int32 per launch, cleared
// by zero_queue_kernel before the launch).  For callers that cannot allocate (graph capture, serving
// loops) it lives in `scratch` below: 256 bytes per device, held until the process ends.  Host code
// never passes the array's name to a kernel (on the host it is only a stand-in for the symbol): the
// device address is taken once per device with hipGetSymbolAddress and kept in scratch_dev[].
static __device__ __align__(256) int32_t scratch[64];

namespace psvd {

constexpr int kScratchDevices = 64;
static std::mutex scratch_mu;
static void *scratch_dev[kScratchDevices];

// the device address of `scratch` on the current device (looked up on the first call per device)
static int scratch_address(void **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kScratchDevices) return PSVD_ELAUNCH;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_dev[dev] == nullptr) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return PSVD_ELAUNCH;
    scratch_dev[dev] = p;
  }
  *out = scratch_dev[dev];
  return PSVD_OK;
}

}  // namespace psvd

extern "C" {

const char *psvd_version(void) { return "pavelsvd 0.2.0 gfx950"; }

size_t psvd_jacobi_lds_workspace_bytes(int64_t batch) {
  (void)batch;
  return 256;
}

int psvd_jacobi_lds_smem_bytes(int dtype, int max_m, int max_n, int flags, size_t *bytes) {
  if (max_m < 1 || max_n < 1 || max_n > max_m || !bytes) return PSVD_EINVAL;
  const bool vectors = !(flags & PSVD_NO_VECTORS);
  size_t b;
  if (dtype == PSVD_F32) b = psvd::lds_bytes<float>(max_m, max_n, vectors);
  else if (dtype == PSVD_F64) b = psvd::lds_bytes<double>(max_m, max_n, vectors);
  else if (dtype == PSVD_C64 || dtype == PSVD_C128) b = psvd::clds_smem_bytes(dtype, max_m, max_n, vectors);
  else return PSVD_EINVAL;
  *bytes = b;
  return b > psvd::kLdsCap - 4096 ? PSVD_ETOOBIG : PSVD_OK;
}

int psvd_jacobi_lds(int dtype, int64_t batch, int m_uniform, int n_uniform, const int32_t *dims,
                    const int64_t *offs, const int32_t *order, int max_m, int max_n, const void *a,
                    void *u, void *s, void *vh, int32_t *info, double tol, int max_sweeps, int flags,
                    int lanes, int threads, void *workspace, size_t workspace_bytes,
                    psvd_stream_t stream) {
  if (batch < 0 || max_sweeps < 1 || !a || !s || !info || !workspace) return PSVD_EINVAL;
  if (workspace_bytes < psvd_jacobi_lds_workspace_bytes(batch)) return PSVD_EINVAL;
  if (batch > 0x7fffffff) return PSVD_ETOOBIG;   // the work-queue counter is an int32
  if (!(flags & PSVD_NO_VECTORS) && (!u || !vh)) return PSVD_EINVAL;
  if (dims == nullptr) {
    if (m_uniform < 1 || n_uniform < 1 || n_uniform > m_uniform) return PSVD_EINVAL;
    max_m = m_uniform;
    max_n = n_uniform;
  } else if (!offs || max_m < 1 || max_n < 1) {
    return PSVD_EINVAL;
  }
  if (batch == 0) return PSVD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == PSVD_C64 || dtype == PSVD_C128) {   // uniform batches only; threads fixed at 256
    if (dims != nullptr || batch > 0x7fffffff) return PSVD_EINVAL;
    return psvd::run_clds(dtype, batch, m_uniform, n_uniform, a, u, s, vh, info, tol, max_sweeps, flags, lanes,
                          workspace, st);
  }
  if (dtype == PSVD_F32)
    return psvd::run_lds<float>(batch, m_uniform, n_uniform, dims, offs, order, max_m, max_n, a, u, s,
                                vh, info, tol, max_sweeps, flags, lanes, threads, workspace,
                                workspace_bytes, st);
  if (dtype == PSVD_F64)
    return psvd::run_lds_f64(batch, m_uniform, n_uniform, dims, offs, order, max_m, max_n, a, u, s, vh, info, tol,
                             max_sweeps, flags, lanes, threads, workspace, workspace_bytes, st);
  return PSVD_EINVAL;
}

int psvd_jacobi_lds_noalloc(int dtype, int64_t batch, int m_uniform, int n_uniform, const int32_t *dims,
                            const int64_t *offs, const int32_t *order, int max_m, int max_n, const void *a, void *u,
                            void *s, void *vh, int32_t *info, double tol, int max_sweeps, int flags, int lanes,
                            int threads, psvd_stream_t stream) {
  if (batch < 0 || max_sweeps < 1 || !a || !s || !info) return PSVD_EINVAL;
  if (batch > 0x7fffffff) return PSVD_ETOOBIG;   // beyond what one int32 queue counter can hand out
  if (batch == 0) return PSVD_OK;
  void *queue = nullptr;
  const int rc = psvd::scratch_address(&queue);
  if (rc != PSVD_OK) return rc;
  return psvd_jacobi_lds(dtype, batch, m_uniform, n_uniform, dims, offs, order, max_m, max_n, a, u, s, vh, info, tol,
                         max_sweeps, flags, lanes, threads, queue, sizeof(scratch), stream);
}

}  // extern "C"
