// K1 for FP64 matrices: the run_lds<double> instantiations (kernels in jacobi_lds_impl.hpp), kept in a
// translation unit of their own so that the FP32 and FP64 kernel sets compile in parallel.
#include "jacobi_lds_impl.hpp"

namespace psvd {

int run_lds_f64(int64_t batch, int mu, int nu, const int32_t *dims, const int64_t *offs, const int32_t *order,
                int max_m, int max_n, const void *a, void *u, void *s, void *vh, int32_t *info, double tol,
                int max_sweeps, int flags, int lanes, int threads, void *ws, size_t ws_bytes, hipStream_t stream) {
  return run_lds<double>(batch, mu, nu, dims, offs, order, max_m, max_n, a, u, s, vh, info, tol, max_sweeps, flags,
                         lanes, threads, ws, ws_bytes, stream);
}

}  // namespace psvd
