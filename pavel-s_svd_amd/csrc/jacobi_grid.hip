// K2: one-sided Jacobi SVD of one matrix too large for LDS, as a single launch.
//
// W (m x np, column-major) and V (np x np, column-major) live in a caller-provided workspace in
// HBM.  Every workgroup of a resident grid (at most one per CU) owns a strided subset of the round's
// column pairs; a pair's three dot products are a workgroup reduction (DPP within waves, LDS across
// waves) accumulated in FP64; the rotation is applied in place and the round ends with a grid barrier
// (GridBarrier, with the agent-scope cache write-back and invalidation W and V need).
// The whole solve -- load, max-abs scale, sweeps with the FP64 convergence test, FP64 sigma,
// sort, cut, coordinate-vector completion, and the row-major outputs -- is one launch with no
// host synchronisation.
//
// Follows svdbench/standalone/primitives.py:316-367 (jacobi_svd), :309-313 (schedule),
// :430-455 (complete_null_columns) and solvers.py:294-299 (max-abs scaling).
#include <float.h>

#include "common.hpp"

namespace psvd {

constexpr int kGridThreads = 256;

template <typename T>
struct GridParams {
  const T *a;
  T *u, *s, *vh;
  int32_t *info;
  int m, n, np, ldw, ldv;            // ldw, ldv: column strides of W and V (multiples of 64)
  T *W, *V;
  double *sig, *vt, *y, *part;       // np, m, np, 2*grid
  int *ord, *parti;                  // np, grid
  unsigned long long *rmax;          // 3 (+1 for amax)
  int *flag;                         // 3 (+1 for nonfinite)
  unsigned *bar;                     // kBarWords grid-barrier counters
  double tol;
  int max_sweeps, flags;
};

__host__ __device__ inline int pad64(int x) { return (x + 63) & ~63; }

template <typename T>
__host__ __device__ inline size_t grid_ws_bytes(int m, int n, int grid) {
  const int np = n + (n & 1);
  size_t b = 0;
  auto take = [&](size_t bytes) { b += (bytes + 255) & ~size_t(255); };
  take(sizeof(T) * (size_t)pad64(m) * np);
  take(sizeof(T) * (size_t)np * pad64(np));
  take(sizeof(double) * np);
  take(sizeof(double) * m);
  take(sizeof(double) * np);
  take(sizeof(double) * 2 * grid);
  take(sizeof(int) * np);
  take(sizeof(int) * grid);
  take(sizeof(unsigned long long) * 4);
  take(sizeof(int) * 4);
  take(sizeof(unsigned) * kBarWords);
  return b;
}

// Sum of three doubles over the workgroup; result in every thread.
__device__ __forceinline__ void block_sum3(double &a, double &b, double &c, double *red) {
  a = group_sum_dpp<64>(a);
  b = group_sum_dpp<64>(b);
  c = group_sum_dpp<64>(c);
  const int w = threadIdx.x / kWave;
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) {
    red[3 * w] = a;
    red[3 * w + 1] = b;
    red[3 * w + 2] = c;
  }
  __syncthreads();
  a = b = c = 0.0;
#pragma unroll
  for (int i = 0; i < kGridThreads / kWave; ++i) {
    a += red[3 * i];
    b += red[3 * i + 1];
    c += red[3 * i + 2];
  }
}

__device__ __forceinline__ double block_sum1(double a, double *red) {
  a = group_sum_dpp<64>(a);
  const int w = threadIdx.x / kWave;
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) red[w] = a;
  __syncthreads();
  a = 0.0;
#pragma unroll
  for (int i = 0; i < kGridThreads / kWave; ++i) a += red[i];
  return a;
}

template <typename T>
__global__ __launch_bounds__(kGridThreads) void jacobi_grid_kernel(GridParams<T> P) {
  GridBarrier bar(P.bar, gridDim.x, blockIdx.x);
  // a barrier wait abandoned (workgroups not co-resident): report it and end the launch
#define PSVD_GRID_SYNC()                       \
  do {                                         \
    if (!bar.sync<true>()) {                   \
      if (gtid == 0) {                         \
        P.info[0] = 0;                         \
        P.info[1] = PSVD_INFO_BARRIER;         \
      }                                        \
      return;                                  \
    }                                          \
  } while (0)
  __shared__ double red[3 * (kGridThreads / kWave)];
  __shared__ int redi[kGridThreads / kWave];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int G = gridDim.x, b = blockIdx.x;
  const int64_t gtid = (int64_t)b * kGridThreads + tid, gsz = (int64_t)G * kGridThreads;
  const int m = P.m, n = P.n, np = P.np, npairs = np / 2, ldw = P.ldw, ldv = P.ldv;
  T *W = P.W, *V = P.V;
  const bool vectors = !(P.flags & PSVD_NO_VECTORS);
  const double eps = Eps<T>::value;
  const double tol = P.tol > 0.0 ? P.tol : 10.0 * eps;
  unsigned long long *rmax = P.rmax;   // rmax[3] = max |a|
  int *flag = P.flag;                  // flag[3] = non-finite input

  if (gtid == 0) {
    rmax[0] = rmax[1] = rmax[2] = rmax[3] = 0ull;
    flag[0] = flag[1] = flag[2] = flag[3] = 0;
  }
  PSVD_GRID_SYNC();
  // ---- load (row-major A -> column-major W), finiteness, max |a|; V = I
  {
    double amax = 0.0;
    int bad = 0;
    for (int64_t e = gtid; e < (int64_t)m * n; e += gsz) {
      const int i = (int)(e / n), j = (int)(e - (int64_t)i * n);
      const T x = P.a[e];
      W[(int64_t)j * ldw + i] = x;
      const double ax = fabs((double)x);
      if (ax <= DBL_MAX) amax = fmax(amax, ax); else bad = 1;
    }
    if (np > n)
      for (int64_t i = gtid; i < m; i += gsz) W[(int64_t)n * ldw + i] = T(0);
    if (vectors)
      for (int64_t e = gtid; e < (int64_t)np * np; e += gsz) {
        const int c = (int)(e / np), i = (int)(e - (int64_t)c * np);
        V[(int64_t)c * ldv + i] = (c == i) ? T(1) : T(0);
      }
    amax = wave_max(amax);
    if (lane == 0 && amax > 0.0) atomicMax(&rmax[3], (unsigned long long)__double_as_longlong(amax));
    if (bad) atomicOr(&flag[3], 1);
  }
  PSVD_GRID_SYNC();
  if (__atomic_load_n(&flag[3], __ATOMIC_RELAXED)) {
    if (gtid == 0) {
      P.info[0] = 0;
      P.info[1] = PSVD_INFO_NONFINITE;
    }
    return;  // uniform: every thread read the same flag after the barrier
  }
  double scale = 1.0;
  if (P.flags & PSVD_SCALE) {
    const double mx = __longlong_as_double((long long)__atomic_load_n(&rmax[3], __ATOMIC_RELAXED));
    if (mx > 0.0) {
      scale = mx;
      const T st = (T)mx;
      for (int64_t e = gtid; e < (int64_t)m * n; e += gsz) {
        const int c = (int)(e / m), i = (int)(e - (int64_t)c * m);
        W[(int64_t)c * ldw + i] = W[(int64_t)c * ldw + i] / st;
      }
    }
    PSVD_GRID_SYNC();
  }
  // ---- initial floor: max_j |w_j|^2
  for (int c = b; c < n; c += G) {
    double acc = 0.0;
    for (int i = tid; i < m; i += kGridThreads) {
      const double x = (double)W[(int64_t)c * ldw + i];
      acc = fma(x, x, acc);
    }
    acc = block_sum1(acc, red);
    if (tid == 0) atomicMax(&rmax[0], (unsigned long long)__double_as_longlong(acc));
  }
  PSVD_GRID_SYNC();

  // ---- sweeps: one grid barrier per round
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
        T *wp = W + (int64_t)p * ldw, *wq = W + (int64_t)q * ldw;
        double app = 0.0, aqq = 0.0, apq = 0.0;
        for (int i = tid; i < m; i += kGridThreads) {
          const double x = (double)wp[i], y = (double)wq[i];
          app = fma(x, x, app);
          aqq = fma(y, y, aqq);
          apq = fma(x, y, apq);
        }
        block_sum3(app, aqq, apq, red);
        double nap = app, naq = aqq;
        if (apq * apq > tol2 * app * aqq && fmin(app, aqq) > floor_) {
          T c, s;
          double tb;
          jacobi_rotation_t<T>(app, aqq, apq, c, s, tb);
          nap = app - tb;
          naq = aqq + tb;
          if (tid == 0) atomicOr(&flag[sw % 3], 1);
          for (int i = tid; i < m; i += kGridThreads) {
            const T x = wp[i], y = wq[i];
            wp[i] = c * x - s * y;
            wq[i] = s * x + c * y;
          }
          if (vectors) {
            T *vp = V + (int64_t)p * ldv, *vq = V + (int64_t)q * ldv;
            for (int i = tid; i < np; i += kGridThreads) {
              const T x = vp[i], y = vq[i];
              vp[i] = c * x - s * y;
              vq[i] = s * x + c * y;
            }
          }
        }
        if (tid == 0) atomicMax(&rmax[(g + 1) % 3], (unsigned long long)__double_as_longlong(fmax(nap, naq)));
      }
      PSVD_GRID_SYNC();
    }
    if (__atomic_load_n(&flag[sw % 3], __ATOMIC_RELAXED) == 0) {
      converged = true;
      break;
    }
  }
  int status = converged ? PSVD_INFO_OK : PSVD_INFO_NOCONV;

  // ---- sigma (FP64), descending order, cut
  for (int c = b; c < n; c += G) {
    double acc = 0.0;
    for (int i = tid; i < m; i += kGridThreads) {
      const double x = (double)W[(int64_t)c * ldw + i];
      acc = fma(x, x, acc);
    }
    acc = block_sum1(acc, red);
    if (tid == 0) P.sig[c] = sqrt(acc);
  }
  PSVD_GRID_SYNC();
  for (int64_t j = gtid; j < n; j += gsz) {
    const double sj = P.sig[j];
    int rank = 0;
    for (int k = 0; k < n; ++k) {
      const double sk = P.sig[k];
      rank += (sk > sj) || (sk == sj && k < j);
    }
    P.ord[rank] = (int)j;
  }
  PSVD_GRID_SYNC();
  const double s1 = P.sig[P.ord[0]];
  const double cut = (P.flags & PSVD_CUT_FLOOR) ? eps * s1 : eps * (double)(m > np ? m : np) * s1;
  int nk;
  {  // sorted descending: binary search for the first column at or below the cut
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = (lo + hi) / 2;
      if (P.sig[P.ord[mid]] > cut) lo = mid + 1; else hi = mid;
    }
    nk = lo;
  }

  if (vectors) {
    for (int64_t e = gtid; e < (int64_t)n * m; e += gsz) {
      const int r = (int)(e / m), i = (int)(e - (int64_t)r * m);
      const int c = P.ord[r];
      T *wc = W + (int64_t)c * ldw;
      wc[i] = r < nk ? (T)((double)wc[i] / P.sig[c]) : T(0);
    }
    PSVD_GRID_SYNC();
    for (int r = nk; r < n; ++r) {
      // leverage 1 - |U(i,:)|^2 and its first maximiser
      double best = -DBL_MAX;
      int bi = 0x7fffffff;
      for (int64_t i = gtid; i < m; i += gsz) {
        double acc = 0.0;
        for (int c = 0; c < n; ++c) {
          const double x = (double)W[(int64_t)c * ldw + i];
          acc = fma(x, x, acc);
        }
        const double lev = 1.0 - acc;
        if (lev > best) {
          best = lev;
          bi = (int)i;
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
      __syncthreads();
      if (lane == 0) {
        red[wave] = best;
        redi[wave] = bi;
      }
      __syncthreads();
      if (tid == 0) {
        for (int w = 1; w < kGridThreads / kWave; ++w)
          if (red[w] > best || (red[w] == best && redi[w] < bi)) {
            best = red[w];
            bi = redi[w];
          }
        P.part[b] = best;
        P.parti[b] = bi;
      }
      PSVD_GRID_SYNC();
      best = P.part[0];
      bi = P.parti[0];
      for (int w = 1; w < G; ++w)
        if (P.part[w] > best || (P.part[w] == best && P.parti[w] < bi)) {
          best = P.part[w];
          bi = P.parti[w];
        }
      const int ix = bi;
      for (int64_t i = gtid; i < m; i += gsz) {
        double acc = (i == ix) ? 1.0 : 0.0;
        for (int c = 0; c < n; ++c) acc -= (double)W[(int64_t)c * ldw + i] * (double)W[(int64_t)c * ldw + ix];
        P.vt[i] = acc;
      }
      PSVD_GRID_SYNC();
      for (int c = b; c < n; c += G) {
        double acc = 0.0;
        for (int i = tid; i < m; i += kGridThreads) acc = fma((double)W[(int64_t)c * ldw + i], P.vt[i], acc);
        acc = block_sum1(acc, red);
        if (tid == 0) P.y[c] = acc;
      }
      PSVD_GRID_SYNC();
      double l2 = 0.0;
      for (int64_t i = gtid; i < m; i += gsz) {
        double acc = P.vt[i];
        for (int c = 0; c < n; ++c) acc -= (double)W[(int64_t)c * ldw + i] * P.y[c];
        P.vt[i] = acc;
        l2 = fma(acc, acc, l2);
      }
      l2 = block_sum1(l2, red);
      if (tid == 0) P.part[G + b] = l2;
      PSVD_GRID_SYNC();
      double len2 = 0.0;
      for (int w = 0; w < G; ++w) len2 += P.part[G + w];
      const double len = sqrt(len2);
      if (len < 10.0 * eps) status = PSVD_INFO_COMPLETION;
      const double inv = 1.0 / fmax(len, DBL_MIN);
      T *wc = W + (int64_t)P.ord[r] * ldw;
      for (int64_t i = gtid; i < m; i += gsz) wc[i] = (T)(P.vt[i] * inv);
      PSVD_GRID_SYNC();
    }
  }

  // ---- outputs
  for (int64_t r = gtid; r < n; r += gsz) P.s[r] = r < nk ? (T)(P.sig[P.ord[r]] * scale) : T(0);
  if (vectors) {
    for (int64_t e = gtid; e < (int64_t)m * n; e += gsz) {
      const int i = (int)(e / n), r = (int)(e - (int64_t)i * n);
      P.u[e] = W[(int64_t)P.ord[r] * ldw + i];
    }
    for (int64_t e = gtid; e < (int64_t)n * n; e += gsz) {
      const int r = (int)(e / n), j = (int)(e - (int64_t)r * n);
      P.vh[e] = V[(int64_t)P.ord[r] * ldv + j];
    }
  }
  if (gtid == 0) {
    P.info[0] = sweeps;
    P.info[1] = status;
  }
}
#undef PSVD_GRID_SYNC

template <typename T>
static int grid_size(int n) {
  static int occ = -1;
  if (occ < 0) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, jacobi_grid_kernel<T>, kGridThreads, 0) != hipSuccess ||
        occ < 1)
      occ = 1;
  }
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (cus <= 0) cus = 1;
  // at most one workgroup per CU, so that every workgroup of the grid is resident without relying on
  // the occupancy estimate
  (void)occ;
  const int np = n + (n & 1);
  int g = np / 2;
  if (g < 64) g = 64;
  return g < cus ? g : cus;
}

template <typename T>
static int run_grid(int m, int n, const void *a, void *u, void *s, void *vh, int32_t *info, double tol,
                    int max_sweeps, int flags, void *ws, size_t ws_bytes, hipStream_t stream) {
  const int G = grid_size<T>(n);
  if (ws_bytes < grid_ws_bytes<T>(m, n, G)) return PSVD_EINVAL;
  GridParams<T> P;
  P.a = static_cast<const T *>(a);
  P.u = static_cast<T *>(u);
  P.s = static_cast<T *>(s);
  P.vh = static_cast<T *>(vh);
  P.info = info;
  P.m = m;
  P.n = n;
  P.np = n + (n & 1);
  P.ldw = pad64(m);
  P.ldv = pad64(P.np);
  char *w = static_cast<char *>(ws);
  auto take = [&](size_t bytes) {
    char *p = w;
    w += (bytes + 255) & ~size_t(255);
    return p;
  };
  P.W = reinterpret_cast<T *>(take(sizeof(T) * (size_t)P.ldw * P.np));
  P.V = reinterpret_cast<T *>(take(sizeof(T) * (size_t)P.np * P.ldv));
  P.sig = reinterpret_cast<double *>(take(sizeof(double) * P.np));
  P.vt = reinterpret_cast<double *>(take(sizeof(double) * m));
  P.y = reinterpret_cast<double *>(take(sizeof(double) * P.np));
  P.part = reinterpret_cast<double *>(take(sizeof(double) * 2 * G));
  P.ord = reinterpret_cast<int *>(take(sizeof(int) * P.np));
  P.parti = reinterpret_cast<int *>(take(sizeof(int) * G));
  P.rmax = reinterpret_cast<unsigned long long *>(take(sizeof(unsigned long long) * 4));
  P.flag = reinterpret_cast<int *>(take(sizeof(int) * 4));
  P.tol = tol;
  P.max_sweeps = max_sweeps;
  P.flags = flags;
  P.bar = reinterpret_cast<unsigned *>(take(sizeof(unsigned) * kBarWords));
  static const int limit = barrier_grid_limit(reinterpret_cast<const void *>(jacobi_grid_kernel<T>), kGridThreads, 0);
  if (G > limit) return PSVD_ELAUNCH;
  hipLaunchKernelGGL(bar_reset_kernel, dim3(1), dim3(256), 0, stream, P.bar, kBarWords);
  hipLaunchKernelGGL(jacobi_grid_kernel<T>, dim3(G), dim3(kGridThreads), 0, stream, P);
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

}  // namespace psvd

extern "C" {

size_t psvd_jacobi_grid_workspace_bytes(int dtype, int m, int n) {
  if (m < 1 || n < 1 || n > m) return 0;
  if (dtype == PSVD_F32) return psvd::grid_ws_bytes<float>(m, n, psvd::grid_size<float>(n));
  if (dtype == PSVD_F64) return psvd::grid_ws_bytes<double>(m, n, psvd::grid_size<double>(n));
  if (dtype == PSVD_C64 || dtype == PSVD_C128) return psvd::cgrid_workspace_bytes(dtype, m, n);
  return 0;
}

int psvd_jacobi_grid(int dtype, int m, int n, const void *a, void *u, void *s, void *vh, int32_t *info,
                     double tol, int max_sweeps, int flags, void *workspace, size_t workspace_bytes,
                     psvd_stream_t stream) {
  if (m < 1 || n < 1 || n > m || max_sweeps < 1 || !a || !s || !info || !workspace) return PSVD_EINVAL;
  if (!(flags & PSVD_NO_VECTORS) && (!u || !vh)) return PSVD_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == PSVD_F32)
    return psvd::run_grid<float>(m, n, a, u, s, vh, info, tol, max_sweeps, flags, workspace, workspace_bytes, st);
  if (dtype == PSVD_F64)
    return psvd::run_grid<double>(m, n, a, u, s, vh, info, tol, max_sweeps, flags, workspace, workspace_bytes, st);
  if (dtype == PSVD_C64 || dtype == PSVD_C128)
    return psvd::cjacobi_grid(dtype, m, n, a, u, s, vh, info, tol, max_sweeps, flags, workspace, workspace_bytes, st);
  return PSVD_EINVAL;
}

}  // extern "C"
