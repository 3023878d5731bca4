// K2c: complex one-sided Jacobi SVD of one matrix too large for LDS, as a single launch (the complex
// counterpart of K2, csrc/jacobi_grid.hip).
//
// W (m x np) and V (np x np) live in the workspace as column-major real and imaginary planes.  Every
// workgroup of a resident grid (at most one per CU) owns a strided subset of each round's column
// pairs; a pair's sums |w_p|^2, |w_q|^2 and w_p^H w_q are workgroup reductions accumulated in FP64;
// the complex rotation of svdbench/standalone/primitives.py:340-351 (c real, s = t c a_pq / |a_pq|,
// w_p <- c w_p - conj(s) w_q, w_q <- s w_p + c w_q, the same for V) is applied in place, and each
// round ends with a grid barrier (GridBarrier with cache write-back and invalidation).  The test is the reference's (:332-339): |a_pq| / sqrt(a_pp a_qq) >
// tol and min(a_pp, a_qq) > eps^2 max_j |w_j|^2, evaluated on the FP64 sums.  The launch ends with
// FP64 sigma, the descending order, the cut (:362) and the outputs -- U = W / sigma for the kept
// columns (zero for the rest), Vh = V^H -- and the count of kept columns; the cut columns are then
// completed by psvd_ccomplete_columns (:430-455) in the same stream.
#include <float.h>

#include "common.hpp"

namespace psvd {

constexpr int kCGridThreads = 256;

template <typename R>
struct CGridParams {
  const R *a;          // (m, n) interleaved complex, row-major
  R *u, *s, *vh;
  int32_t *info;
  int m, n, np, ldw, ldv;
  R *Wr, *Wi, *Vr, *Vi;
  double *sig;         // np
  int *ord;            // np
  int32_t *nk;         // 1
  unsigned long long *rmax;   // 4: three rotating floors, max modulus of A
  int *flag;                  // 4: three rotating sweep flags, non-finite input
  unsigned *bar;              // kBarWords grid-barrier counters
  double tol;
  int max_sweeps, flags;
};

__host__ __device__ inline int cpad64(int x) { return (x + 63) & ~63; }

template <typename R>
__host__ __device__ inline size_t cgrid_ws_bytes(int m, int n) {
  const int np = n + (n & 1);
  size_t b = 0;
  auto take = [&](size_t bytes) { b += (bytes + 255) & ~size_t(255); };
  take(sizeof(R) * 2 * (size_t)cpad64(m) * np);
  take(sizeof(R) * 2 * (size_t)np * cpad64(np));
  take(sizeof(double) * np);
  take(sizeof(int) * np);
  take(sizeof(int32_t) * 2);
  take(sizeof(unsigned long long) * 4);
  take(sizeof(int) * 4);
  take(sizeof(unsigned) * kBarWords);
  return b;
}

// Sums of k doubles over the workgroup, in a fixed order; results in every thread.
template <int K>
__device__ __forceinline__ void cgrid_block_sum(double (&v)[K], double *red) {
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = group_sum_dpp<kWave>(v[k]);
  const int w = threadIdx.x / kWave;
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) red[K * w + k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < kCGridThreads / kWave; ++i) t += red[K * i + k];
    v[k] = t;
  }
}

template <typename R>
__global__ __launch_bounds__(kCGridThreads) void cjacobi_grid_kernel(CGridParams<R> P) {
  GridBarrier bar(P.bar, gridDim.x, blockIdx.x);
  // a barrier wait abandoned (workgroups not co-resident): report it and end the launch
#define PSVD_CGRID_SYNC()                      \
  do {                                         \
    if (!bar.sync<true>()) {                   \
      if (blockIdx.x == 0 && threadIdx.x == 0) { \
        P.info[0] = 0;                         \
        P.info[1] = PSVD_INFO_BARRIER;         \
        *P.nk = P.n;                           \
      }                                        \
      return;                                  \
    }                                          \
  } while (0)
  __shared__ double red[4 * (kCGridThreads / kWave)];
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const int G = gridDim.x, b = blockIdx.x;
  const int64_t gtid = (int64_t)b * kCGridThreads + tid, gsz = (int64_t)G * kCGridThreads;
  const int m = P.m, n = P.n, np = P.np, npairs = np / 2, ldw = P.ldw, ldv = P.ldv;
  R *Wr = P.Wr, *Wi = P.Wi, *Vr = P.Vr, *Vi = P.Vi;
  const bool vectors = !(P.flags & PSVD_NO_VECTORS);
  const double eps = Eps<R>::value;
  const double tol = P.tol > 0.0 ? P.tol : 10.0 * eps;
  unsigned long long *rmax = P.rmax;
  int *flag = P.flag;

  if (gtid == 0) {
    rmax[0] = rmax[1] = rmax[2] = rmax[3] = 0ull;
    flag[0] = flag[1] = flag[2] = flag[3] = 0;
  }
  PSVD_CGRID_SYNC();
  {  // load, finiteness, largest modulus; V = I
    double amax = 0.0;
    int bad = 0;
    for (int64_t e = gtid; e < (int64_t)m * n; e += gsz) {
      const int i = (int)(e / n), j = (int)(e - (int64_t)i * n);
      const R xr = P.a[2 * e], xi = P.a[2 * e + 1];
      Wr[(int64_t)j * ldw + i] = xr;
      Wi[(int64_t)j * ldw + i] = xi;
      const double ax = hypot((double)xr, (double)xi);
      if (ax <= DBL_MAX) amax = fmax(amax, ax); else bad = 1;
    }
    if (np > n)
      for (int64_t i = gtid; i < m; i += gsz) {
        Wr[(int64_t)n * ldw + i] = R(0);
        Wi[(int64_t)n * ldw + i] = R(0);
      }
    if (vectors)
      for (int64_t e = gtid; e < (int64_t)np * np; e += gsz) {
        const int c = (int)(e / np), i = (int)(e - (int64_t)c * np);
        Vr[(int64_t)c * ldv + i] = (c == i) ? R(1) : R(0);
        Vi[(int64_t)c * ldv + i] = R(0);
      }
    amax = wave_max(amax);
    if (lane == 0 && amax > 0.0) atomicMax(&rmax[3], (unsigned long long)__double_as_longlong(amax));
    if (bad) atomicOr(&flag[3], 1);
  }
  PSVD_CGRID_SYNC();
  if (__atomic_load_n(&flag[3], __ATOMIC_RELAXED)) {
    if (gtid == 0) {
      P.info[0] = 0;
      P.info[1] = PSVD_INFO_NONFINITE;
      *P.nk = n;
    }
    return;   // uniform: every thread read the same flag after the barrier
  }
  double scale = 1.0;
  if (P.flags & PSVD_SCALE) {   // a / a.abs().amax(), solvers.py:294
    const double mx = __longlong_as_double((long long)__atomic_load_n(&rmax[3], __ATOMIC_RELAXED));
    if (mx > 0.0) {
      scale = mx;
      const R st = (R)mx;
      for (int64_t e = gtid; e < (int64_t)m * n; e += gsz) {
        const int c = (int)(e / m), i = (int)(e - (int64_t)c * m);
        Wr[(int64_t)c * ldw + i] = Wr[(int64_t)c * ldw + i] / st;
        Wi[(int64_t)c * ldw + i] = Wi[(int64_t)c * ldw + i] / st;
      }
    }
    PSVD_CGRID_SYNC();
  }
  for (int c = b; c < n; c += G) {   // initial floor: max_j |w_j|^2
    double v[1] = {0.0};
    for (int i = tid; i < m; i += kCGridThreads) {
      const double xr = Wr[(int64_t)c * ldw + i], xi = Wi[(int64_t)c * ldw + i];
      v[0] = fma(xr, xr, fma(xi, xi, v[0]));
    }
    cgrid_block_sum<1>(v, red);
    if (tid == 0) atomicMax(&rmax[0], (unsigned long long)__double_as_longlong(v[0]));
  }
  PSVD_CGRID_SYNC();

  const double tol2 = tol * tol, eps2 = eps * eps;
  int sweeps = 0;
  bool converged = false;
  int g = 0;
  for (int sw = 1; sw <= P.max_sweeps; ++sw) {
    sweeps = sw;
    if (gtid == 0) atomicExch(&flag[(sw + 1) % 3], 0);
    for (int r = 0; r < np - 1; ++r, ++g) {
      if (gtid == 0) atomicExch(&rmax[(g + 2) % 3], 0ull);
      const double floor_ = eps2 * __longlong_as_double((long long)__atomic_load_n(&rmax[g % 3], __ATOMIC_RELAXED));
      for (int k = b; k < npairs; k += G) {
        int p, q;
        rr_pair(r, k, np, p, q);
        R *pr = Wr + (int64_t)p * ldw, *pi = Wi + (int64_t)p * ldw;
        R *qr = Wr + (int64_t)q * ldw, *qi = Wi + (int64_t)q * ldw;
        double v[4] = {0.0, 0.0, 0.0, 0.0};   // |w_p|^2, |w_q|^2, Re and Im of w_p^H w_q
        for (int i = tid; i < m; i += kCGridThreads) {
          const double xr = pr[i], xi = pi[i], yr = qr[i], yi = qi[i];
          v[0] = fma(xr, xr, fma(xi, xi, v[0]));
          v[1] = fma(yr, yr, fma(yi, yi, v[1]));
          v[2] = fma(xr, yr, fma(xi, yi, v[2]));
          v[3] = fma(xr, yi, fma(-xi, yr, v[3]));
        }
        cgrid_block_sum<4>(v, red);
        const double app = v[0], aqq = v[1], ar = v[2], ai = v[3];
        const double a2 = ar * ar + ai * ai;
        double nap = app, naq = aqq;
        if (a2 > tol2 * app * aqq && fmin(app, aqq) > floor_) {
          const double a = sqrt(a2);
          const double tau = (aqq - app) / (2.0 * a);
          const double t = tau == 0.0 ? 1.0 : copysign(1.0, tau) / (fabs(tau) + hypot(tau, 1.0));
          const double cd = 1.0 / sqrt(1.0 + t * t), f = t * cd / a;
          const R c = (R)cd, sr = (R)(f * ar), si = (R)(f * ai);
          nap = app - t * a;
          naq = aqq + t * a;
          if (tid == 0) atomicOr(&flag[sw % 3], 1);
          for (int i = tid; i < m; i += kCGridThreads) {
            const R xr = pr[i], xi = pi[i], yr = qr[i], yi = qi[i];
            pr[i] = c * xr - (sr * yr + si * yi);
            pi[i] = c * xi - (sr * yi - si * yr);
            qr[i] = (sr * xr - si * xi) + c * yr;
            qi[i] = (sr * xi + si * xr) + c * yi;
          }
          if (vectors) {
            R *vpr = Vr + (int64_t)p * ldv, *vpi = Vi + (int64_t)p * ldv;
            R *vqr = Vr + (int64_t)q * ldv, *vqi = Vi + (int64_t)q * ldv;
            for (int i = tid; i < np; i += kCGridThreads) {
              const R xr = vpr[i], xi = vpi[i], yr = vqr[i], yi = vqi[i];
              vpr[i] = c * xr - (sr * yr + si * yi);
              vpi[i] = c * xi - (sr * yi - si * yr);
              vqr[i] = (sr * xr - si * xi) + c * yr;
              vqi[i] = (sr * xi + si * xr) + c * yi;
            }
          }
        }
        if (tid == 0) atomicMax(&rmax[(g + 1) % 3], (unsigned long long)__double_as_longlong(fmax(nap, naq)));
      }
      PSVD_CGRID_SYNC();
    }
    if (__atomic_load_n(&flag[sw % 3], __ATOMIC_RELAXED) == 0) {
      converged = true;
      break;
    }
  }

  // sigma (FP64), descending order, cut
  for (int c = b; c < n; c += G) {
    double v[1] = {0.0};
    for (int i = tid; i < m; i += kCGridThreads) {
      const double xr = Wr[(int64_t)c * ldw + i], xi = Wi[(int64_t)c * ldw + i];
      v[0] = fma(xr, xr, fma(xi, xi, v[0]));
    }
    cgrid_block_sum<1>(v, red);
    if (tid == 0) P.sig[c] = sqrt(v[0]);
  }
  PSVD_CGRID_SYNC();
  for (int64_t j = gtid; j < n; j += gsz) {
    const double sj = P.sig[j];
    int rank = 0;
    for (int k = 0; k < n; ++k) {
      const double sk = P.sig[k];
      rank += (sk > sj) || (sk == sj && k < j);
    }
    P.ord[rank] = (int)j;
  }
  PSVD_CGRID_SYNC();
  const double s1 = P.sig[P.ord[0]];
  const double cut = (P.flags & PSVD_CUT_FLOOR) ? eps * s1 : eps * (double)(m > np ? m : np) * s1;
  int lo = 0, hi = n;   // the first column at or below the cut
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (P.sig[P.ord[mid]] > cut) lo = mid + 1; else hi = mid;
  }
  const int nk = lo;
  for (int64_t r = gtid; r < n; r += gsz) P.s[r] = r < nk ? (R)(P.sig[P.ord[r]] * scale) : R(0);
  if (vectors) {
    for (int64_t e = gtid; e < (int64_t)m * n; e += gsz) {
      const int i = (int)(e / n), r = (int)(e - (int64_t)i * n);
      const int c = P.ord[r];
      const double inv = r < nk ? 1.0 / P.sig[c] : 0.0;
      P.u[2 * e] = (R)((double)Wr[(int64_t)c * ldw + i] * inv);
      P.u[2 * e + 1] = (R)((double)Wi[(int64_t)c * ldw + i] * inv);
    }
    for (int64_t e = gtid; e < (int64_t)n * n; e += gsz) {   // Vh = V^H
      const int r = (int)(e / n), j = (int)(e - (int64_t)r * n);
      P.vh[2 * e] = Vr[(int64_t)P.ord[r] * ldv + j];
      P.vh[2 * e + 1] = -Vi[(int64_t)P.ord[r] * ldv + j];
    }
  }
  if (gtid == 0) {
    *P.nk = vectors ? nk : n;
    P.info[0] = sweeps;
    P.info[1] = converged ? PSVD_INFO_OK : PSVD_INFO_NOCONV;
  }
}
#undef PSVD_CGRID_SYNC

__global__ void cgrid_status_kernel(int32_t *info, const int32_t *cstatus) {
  if (threadIdx.x == 0 && info[1] == PSVD_INFO_OK && *cstatus != 0) info[1] = PSVD_INFO_COMPLETION;
}

static int cgrid_size(int n) {
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (cus <= 0) cus = 1;
  // at most one workgroup per CU, so every workgroup of the grid is resident
  const int np = n + (n & 1);
  int g = np / 2;
  if (g < 64) g = 64;
  return g < cus ? g : cus;
}

template <typename R>
static int run_cgrid(int dtype, int m, int n, const void *a, void *u, void *s, void *vh, int32_t *info, double tol,
                     int max_sweeps, int flags, void *ws, size_t ws_bytes, hipStream_t stream) {
  if (ws_bytes < cgrid_ws_bytes<R>(m, n)) return PSVD_EINVAL;
  const int G = cgrid_size(n);
  CGridParams<R> P;
  P.a = static_cast<const R *>(a);
  P.u = static_cast<R *>(u);
  P.s = static_cast<R *>(s);
  P.vh = static_cast<R *>(vh);
  P.info = info;
  P.m = m;
  P.n = n;
  P.np = n + (n & 1);
  P.ldw = cpad64(m);
  P.ldv = cpad64(P.np);
  char *w = static_cast<char *>(ws);
  auto take = [&](size_t bytes) {
    char *p = w;
    w += (bytes + 255) & ~size_t(255);
    return p;
  };
  P.Wr = reinterpret_cast<R *>(take(sizeof(R) * 2 * (size_t)P.ldw * P.np));
  P.Wi = P.Wr + (size_t)P.ldw * P.np;
  P.Vr = reinterpret_cast<R *>(take(sizeof(R) * 2 * (size_t)P.np * P.ldv));
  P.Vi = P.Vr + (size_t)P.np * P.ldv;
  P.sig = reinterpret_cast<double *>(take(sizeof(double) * P.np));
  P.ord = reinterpret_cast<int *>(take(sizeof(int) * P.np));
  P.nk = reinterpret_cast<int32_t *>(take(sizeof(int32_t) * 2));
  P.rmax = reinterpret_cast<unsigned long long *>(take(sizeof(unsigned long long) * 4));
  P.flag = reinterpret_cast<int *>(take(sizeof(int) * 4));
  P.tol = tol;
  P.max_sweeps = max_sweeps;
  P.flags = flags;
  P.bar = reinterpret_cast<unsigned *>(take(sizeof(unsigned) * kBarWords));
  static const int limit = barrier_grid_limit(reinterpret_cast<const void *>(cjacobi_grid_kernel<R>), kCGridThreads, 0);
  if (G > limit) return PSVD_ELAUNCH;
  hipLaunchKernelGGL(bar_reset_kernel, dim3(1), dim3(256), 0, stream, P.bar, kBarWords);
  hipLaunchKernelGGL(cjacobi_grid_kernel<R>, dim3(G), dim3(kCGridThreads), 0, stream, P);
  if (hipGetLastError() != hipSuccess) return PSVD_ELAUNCH;
  if (!(flags & PSVD_NO_VECTORS)) {
    int32_t *cstatus = P.nk + 1;
    const int rc = psvd_ccomplete_columns(dtype, 1, m, n, u, P.nk, cstatus, stream);
    if (rc != PSVD_OK) return rc;
    hipLaunchKernelGGL(cgrid_status_kernel, dim3(1), dim3(64), 0, stream, info, cstatus);
  }
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

size_t cgrid_workspace_bytes(int dtype, int m, int n) {
  if (m < 1 || n < 1 || n > m) return 0;
  if (dtype == PSVD_C64) return cgrid_ws_bytes<float>(m, n);
  if (dtype == PSVD_C128) return cgrid_ws_bytes<double>(m, n);
  return 0;
}

int cjacobi_grid(int dtype, int m, int n, const void *a, void *u, void *s, void *vh, int32_t *info, double tol,
                 int max_sweeps, int flags, void *workspace, size_t workspace_bytes, hipStream_t stream) {
  if (dtype == PSVD_C64)
    return run_cgrid<float>(dtype, m, n, a, u, s, vh, info, tol, max_sweeps, flags, workspace, workspace_bytes,
                            stream);
  if (dtype == PSVD_C128)
    return run_cgrid<double>(dtype, m, n, a, u, s, vh, info, tol, max_sweeps, flags, workspace, workspace_bytes,
                             stream);
  return PSVD_EINVAL;
}

}  // namespace psvd
