// K4: blocked Householder QR of a batch of tall matrices, A = Q R with thin Q (m x n) and
// R (n x n), in the reference's reflector convention (LAPACK xLARFG: beta = -sign(alpha) *
// hypot(alpha, |x2|), tau = (beta - alpha) / beta, v = x / (alpha - beta), v0 = 1, and the
// trivial reflector tau = 0 when x2 = 0), so R matches primitives.qr up to rounding.
//
// Per panel of kPanel columns: one workgroup per matrix factors the panel (rows owned by threads,
// norms and the panel-internal w = v^T A products as FP64 workgroup reductions), a GEMM forms
// G = V^T V, one workgroup per matrix builds the compact-WY factor T (H_1...H_k = I - V T V^T,
// LAPACK larft, forward columnwise), and three GEMMs apply Q^T = I - V T^T V^T to the trailing
// columns.  Q is accumulated backwards from [I; 0] with the same three GEMMs per panel (orgqr).
// The host loop below only enqueues kernels: no allocation, no synchronisation.
//
// Follows svdbench/standalone/primitives.py:41-59 (householder_vec), :62-65 (_apply_left) and
// :74-89 (qr).
#include <float.h>

#include "common.hpp"

namespace psvd {

constexpr int kPanel = 32;
constexpr int kPanelThreads = 512;

// Sum of K <= kPanel doubles per thread over the workgroup; results broadcast through `out`.
template <int NT>
__device__ __forceinline__ void block_sum_vec(double (&v)[kPanel], int k_lo, int k_hi, double *red, double *out) {
  constexpr int NW = NT / kWave;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < kPanel; ++k)
    if (k >= k_lo && k < k_hi) {
      const double s = group_sum_dpp<kWave>(v[k]);
      if (lane == 0) red[wave * kPanel + k] = s;
    }
  __syncthreads();
  if ((int)threadIdx.x >= k_lo && (int)threadIdx.x < k_hi) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += red[w * kPanel + threadIdx.x];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

template <int NT>
__device__ __forceinline__ double block_sum_one(double v, double *red) {
  constexpr int NW = NT / kWave;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  v = group_sum_dpp<kWave>(v);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < NW; ++w) s += red[w];
  return s;
}

// Factor columns j0 .. j0+nbc-1 of F (rows j0 .. m-1) in place: R entries stay in F (zeros
// written below the diagonal), the explicit unit-lower reflectors go to V (same layout as F),
// tau[j] per column.
template <typename T>
__global__ __launch_bounds__(kPanelThreads) void qr_panel_kernel(T *F, int64_t ldf, int64_t sf, T *V, int64_t sv,
                                                                 T *tau, int64_t stau, int m, int j0, int nbc) {
  constexpr int NT = kPanelThreads;
  __shared__ double red[(NT / kWave) * kPanel];
  __shared__ double wsum[kPanel];
  const int tid = threadIdx.x;
  T *Fb = F + (int64_t)blockIdx.x * sf;
  T *Vb = V + (int64_t)blockIdx.x * sv;
  T *tb = tau + (int64_t)blockIdx.x * stau;
  for (int c = 0; c < nbc; ++c) {
    const int j = j0 + c;
    double part = 0.0;
    for (int i = tid; i < m; i += NT)
      if (i > j) {
        const double x = (double)Fb[(int64_t)i * ldf + j];
        part = fma(x, x, part);
      }
    const double xnorm = sqrt(block_sum_one<NT>(part, red));
    const double alpha = (double)Fb[(int64_t)j * ldf + j];
    double beta, tj, scal;
    const bool trivial = xnorm == 0.0;
    if (trivial) {
      beta = alpha;
      tj = 0.0;
      scal = 0.0;
    } else {
      beta = -copysign(1.0, alpha == 0.0 ? 1.0 : alpha) * hypot(alpha, xnorm);
      tj = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    // v (explicit, unit at row j; zero for the trivial reflector), R diagonal, zeros below it
    for (int i = tid; i < m; i += NT) {
      if (i > j) {
        const T vi = (T)((double)Fb[(int64_t)i * ldf + j] * scal);
        Vb[(int64_t)i * ldf + j] = vi;
        Fb[(int64_t)i * ldf + j] = T(0);
      } else if (i == j) {
        Vb[(int64_t)i * ldf + j] = trivial ? T(0) : T(1);
        Fb[(int64_t)i * ldf + j] = (T)beta;
      } else if (i >= j0) {
        Vb[(int64_t)i * ldf + j] = T(0);
      }
    }
    if (tid == 0) tb[j] = (T)tj;
    if (!trivial && c + 1 < nbc) {
      // w_k = v^T F(:, k) over rows >= j, then F(:, k) -= tau v w_k, for the panel's later columns
      double w[kPanel];
#pragma unroll
      for (int k = 0; k < kPanel; ++k) w[k] = 0.0;
      for (int i = tid; i < m; i += NT)
        if (i >= j) {
          const double vi = (double)Vb[(int64_t)i * ldf + j];
          const T *row = Fb + (int64_t)i * ldf + j0;
#pragma unroll
          for (int k = 0; k < kPanel; ++k)
            if (k > c && k < nbc) w[k] = fma(vi, (double)row[k], w[k]);
        }
      block_sum_vec<NT>(w, c + 1, nbc, red, wsum);
      for (int i = tid; i < m; i += NT)
        if (i >= j) {
          const double tv = tj * (double)Vb[(int64_t)i * ldf + j];
          T *row = Fb + (int64_t)i * ldf + j0;
#pragma unroll
          for (int k = 0; k < kPanel; ++k)
            if (k > c && k < nbc) row[k] = (T)((double)row[k] - tv * wsum[k]);
        }
    }
    __syncthreads();
  }
}

// ---- multi-workgroup panel for tall panels ---------------------------------------------------
// The same reflectors as qr_panel_kernel, with the panel's rows spread over nw workgroups per
// matrix (KR rows per thread, held in registers for the whole panel: one read, one write).  Each
// column needs two reductions over all rows -- |x2|^2, then w = v^T F(:, later columns) -- done as
// per-workgroup partial sums in FP64 that every workgroup then adds in the same fixed order after
// a barrier over the matrix's workgroups (deterministic sums, no atomic adds).  The partials cross
// workgroups through agent-scope stores and loads, so the barrier (GridBarrier without cache
// maintenance) only orders them.  Scratch buffers alternate with the column parity so that a fast
// workgroup never overwrites partials a slow one is still reading.
// per workgroup and column: kPanel w partials and the five partial sums that carry the next
// column's norm and alpha through the same grid barrier (kCoopW doubles)
constexpr int kCoopW = kPanel + 8;
struct CoopScr {
  int nw;
  __host__ __device__ static size_t doubles(int nw) { return 2 * ((size_t)nw + 1 + (size_t)nw * kCoopW); }
};

constexpr int kCoopLd = kPanel + 1;

template <typename T, int KR>
__host__ __device__ constexpr size_t coop_lds_bytes() {
  return sizeof(T) * (size_t)256 * KR * kCoopLd;
}

template <typename T, int KR>
__global__ __launch_bounds__(256) void qr_panel_coop_kernel(T *F, int64_t ldf, int64_t sf, T *V, int64_t sv, T *tau,
                                                            int64_t stau, int m, int j0, int nbc, int nw,
                                                            double *scr, unsigned *bar) {
  constexpr int NT = 256;
  extern __shared__ double smem_coop[];
  __shared__ double red[(NT / kWave) * kPanel];
  __shared__ double wsum[kPanel];
  __shared__ double seg[NT / kCoopW][kCoopW];
  const int tid = threadIdx.x;
  const int b = blockIdx.x / nw, wi = blockIdx.x - b * nw;
  GridBarrier gb(bar + (size_t)b * kBarWords, (unsigned)nw, (unsigned)wi);   // the matrix's workgroups
  bool aborted = false;
  T *Fb = F + (int64_t)b * sf;
  T *Vb = V + (int64_t)b * sv;
  T *tb = tau + (int64_t)b * stau;
  double *sb = scr + (size_t)b * CoopScr::doubles(nw);
  const size_t half = (size_t)nw + 1 + (size_t)nw * kCoopW;
  __shared__ double red5[(NT / kWave) * 8];
  __shared__ double ext[8];
  // the thread's KR panel rows, private to it (no barrier needed around them)
  T *xs = reinterpret_cast<T *>(smem_coop);
  auto X = [&](int q, int k) -> T & { return xs[(q * NT + tid) * kCoopLd + k]; };
  int row[KR];
#pragma unroll
  for (int q = 0; q < KR; ++q) {
    row[q] = j0 + (wi * KR + q) * NT + tid;
#pragma unroll
    for (int k = 0; k < kPanel; ++k) X(q, k) = (row[q] < m && k < nbc) ? Fb[(int64_t)row[q] * ldf + j0 + k] : T(0);
  }
  // Column c's norm below the diagonal and alpha normally come through the reduction of column
  // c - 1 (see below), saving one grid barrier per column; they are reduced directly for the first
  // column, after a trivial reflector, and when the carried value would have cancelled.
  bool carried = false;
  double cx2 = 0.0, cal = 0.0;
  for (int c = 0; c < nbc; ++c) {
    const int j = j0 + c;
    double *np_ = sb + (c & 1) * half;        // nw norm partials, then alpha
    double *al = np_ + nw;
    double *wp = al + 1;                      // nw x kCoopW w partials and carried sums
    double xnorm, alpha;
    if (carried) {
      xnorm = sqrt(cx2);
      alpha = cal;
    } else {
      double part = 0.0;
#pragma unroll
      for (int q = 0; q < KR; ++q)
        if (row[q] > j && row[q] < m) part = fma((double)X(q, c), (double)X(q, c), part);
      part = block_sum_one<NT>(part, red);
      if (tid == 0) st_agent(np_ + wi, part);
#pragma unroll
      for (int q = 0; q < KR; ++q)
        if (row[q] == j) st_agent(al, (double)X(q, c));
      if (!gb.sync<false>()) {
        aborted = true;
        break;
      }
      double tot = 0.0;
      for (int t = tid; t < nw; t += NT) tot += ld_agent(np_ + t);
      xnorm = sqrt(block_sum_one<NT>(tot, red));
      alpha = ld_agent(al);
    }
    carried = false;
    double beta, tj, scal;
    const bool trivial = xnorm == 0.0;
    if (trivial) {
      beta = alpha;
      tj = 0.0;
      scal = 0.0;
    } else {
      beta = -copysign(1.0, alpha == 0.0 ? 1.0 : alpha) * hypot(alpha, xnorm);
      tj = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    double v[KR];
#pragma unroll
    for (int q = 0; q < KR; ++q) {
      const int i = row[q];
      v[q] = 0.0;
      if (i >= m) continue;
      if (i > j) {
        const T vi = (T)((double)X(q, c) * scal);
        v[q] = (double)vi;
        Vb[(int64_t)i * ldf + j] = vi;
        X(q, c) = T(0);
      } else if (i == j) {
        v[q] = trivial ? 0.0 : 1.0;
        Vb[(int64_t)i * ldf + j] = trivial ? T(0) : T(1);
        X(q, c) = (T)beta;
      } else {
        Vb[(int64_t)i * ldf + j] = T(0);
      }
    }
    if (wi == 0 && tid == 0) tb[j] = (T)tj;
    if (!trivial && c + 1 < nbc) {
      double w[kPanel];
#pragma unroll
      for (int k = 0; k < kPanel; ++k) {
        w[k] = 0.0;
        if (k > c && k < nbc)
#pragma unroll
          for (int q = 0; q < KR; ++q) w[k] = fma(v[q], (double)X(q, k), w[k]);
      }
      // the next column x = X(:, c+1) over rows > j+1 after the update x - tau v w becomes
      // |x|^2 - 2 tau w x.v + tau^2 w^2 |v|^2: its three sums, and x, v at row j+1 (its alpha)
      double e[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int q = 0; q < KR; ++q) {
        const int i = row[q];
        if (i > j + 1 && i < m) {
          const double x = (double)X(q, c + 1);
          e[0] = fma(x, x, e[0]);
          e[1] = fma(x, v[q], e[1]);
          e[2] = fma(v[q], v[q], e[2]);
        } else if (i == j + 1) {
          e[3] = (double)X(q, c + 1);
          e[4] = v[q];
        }
      }
      {
        const int lane = tid & (kWave - 1), wave = tid / kWave;
#pragma unroll
        for (int r = 0; r < 5; ++r) {
          const double t = group_sum_dpp<kWave>(e[r]);
          if (lane == 0) red5[wave * 8 + r] = t;
        }
      }
      block_sum_vec<NT>(w, c + 1, nbc, red, wsum);   // its barriers also publish red5
      if (tid > c && tid < nbc) st_agent(wp + (size_t)wi * kCoopW + tid, wsum[tid]);
      if (tid < 5) {
        double t = 0.0;
#pragma unroll
        for (int x = 0; x < NT / kWave; ++x) t += red5[x * 8 + tid];
        st_agent(wp + (size_t)wi * kCoopW + kPanel + tid, t);
      }
      if (!gb.sync<false>()) {
        aborted = true;
        break;
      }
      {  // thread (k = tid % kCoopW, s = tid / kCoopW), s < NSEG, adds the partials of workgroups
         // s, s + NSEG, ... in that order, four loads in flight at a time; then the NSEG segment
         // sums in order.  Every workgroup forms the same sums in the same order.
        constexpr int NSEG = NT / kCoopW;
        const int k = tid % kCoopW, s = tid / kCoopW;
        double acc = 0.0;
        if (s < NSEG && ((k > c && k < nbc) || (k >= kPanel && k < kPanel + 5))) {
          const double *src = wp + k;
          int t = s;
          for (; t + 3 * NSEG < nw; t += 4 * NSEG) {
            const double a0 = ld_agent(src + (size_t)t * kCoopW), a1 = ld_agent(src + (size_t)(t + NSEG) * kCoopW);
            const double a2 = ld_agent(src + (size_t)(t + 2 * NSEG) * kCoopW);
            const double a3 = ld_agent(src + (size_t)(t + 3 * NSEG) * kCoopW);
            acc += a0;
            acc += a1;
            acc += a2;
            acc += a3;
          }
          for (; t < nw; t += NSEG) acc += ld_agent(src + (size_t)t * kCoopW);
        }
        if (s < NSEG) seg[s][k] = acc;
      }
      __syncthreads();
      if (tid < kPanel + 5) {
        double acc = 0.0;
#pragma unroll
        for (int s = 0; s < NT / kCoopW; ++s) acc += seg[s][tid];
        if (tid < kPanel) wsum[tid] = acc;
        else ext[tid - kPanel] = acc;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < KR; ++q) {
        const double tv = tj * v[q];
#pragma unroll
        for (int k = 0; k < kPanel; ++k)
          if (k > c && k < nbc) X(q, k) = (T)((double)X(q, k) - tv * wsum[k]);
      }
      {  // next column: norm^2 below the diagonal and alpha (rounded as the owner rounds X)
        const double w1 = wsum[c + 1], s1 = ext[0];
        const double n2 = s1 - 2.0 * tj * w1 * ext[1] + tj * tj * w1 * w1 * ext[2];
        cal = (double)(T)(ext[3] - (tj * ext[4]) * w1);
        cx2 = n2;
        carried = n2 >= 1e-4 * s1 && n2 <= DBL_MAX;   // no cancellation: else reduce it directly
      }
      __syncthreads();
    }
  }
  // an abandoned barrier wait leaves the panel NaN, so the solve ends in the non-finite-output error
  const T poison = T(__builtin_nan(""));
#pragma unroll
  for (int q = 0; q < KR; ++q)
    if (row[q] < m)
#pragma unroll
      for (int k = 0; k < kPanel; ++k)
        if (k < nbc) Fb[(int64_t)row[q] * ldf + j0 + k] = aborted ? poison : X(q, k);
  if (aborted && wi == 0 && tid < nbc) tb[j0 + tid] = poison;
}

static int device_cus() {
  static int cus = -1;
  if (cus < 0) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
  }
  return cus;
}

// Rows per thread for the multi-workgroup panel of a `batch` of matrices with `rows` panel rows, or 0
// when even the largest setting would need more workgroups than there are CUs (one workgroup per
// CU keeps the whole grid resident without relying on occupancy).
template <typename T>
static int coop_rows_per_thread(int64_t batch, int rows) {
  const int cus = device_cus();
  const int kmax = sizeof(T) == 4 ? 4 : 2;
  for (int kr = 1; kr <= kmax; kr *= 2) {
    const int64_t nw = (rows + 256LL * kr - 1) / (256LL * kr);
    if (nw * batch <= cus) return kr;
  }
  return 0;
}

template <typename T>
static size_t coop_scr_bytes(int64_t batch, int m) {
  const int64_t nw = (m + 255) / 256;   // the largest nw any panel of this matrix can use
  return sizeof(double) * (size_t)batch * CoopScr::doubles((int)nw) + sizeof(unsigned) * (size_t)batch * kBarWords;
}

template <typename T, int KR>
static int launch_coop(T *F, int64_t ldf, int64_t sf, T *V, int64_t sv, T *tau, int64_t stau, int m, int j0, int nbc,
                       int64_t batch, double *scr, hipStream_t st) {
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(qr_panel_coop_kernel<T, KR>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)coop_lds_bytes<T, KR>()) != hipSuccess)
      return PSVD_ELAUNCH;
    attr = true;
  }
  static const int limit = barrier_grid_limit(reinterpret_cast<const void *>(qr_panel_coop_kernel<T, KR>), 256,
                                              coop_lds_bytes<T, KR>());
  const int nw = (int)((m - j0 + 256LL * KR - 1) / (256LL * KR));
  if (nw * batch > limit) return PSVD_ELAUNCH;
  // barrier counters after the partials of the largest nw (coop_scr_bytes)
  unsigned *bar = reinterpret_cast<unsigned *>(scr + (size_t)batch * CoopScr::doubles((m + 255) / 256));
  hipLaunchKernelGGL(bar_reset_kernel, dim3(1), dim3(256), 0, st, bar, (int)(batch * kBarWords));
  const size_t lds = coop_lds_bytes<T, KR>();
  hipLaunchKernelGGL((qr_panel_coop_kernel<T, KR>), dim3((unsigned)(nw * batch)), dim3(256), lds, st, F, ldf, sf, V,
                     sv, tau, stau, m, j0, nbc, nw, scr, bar);
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

// Compact-WY factor of one panel: T (nbc x nbc upper triangular, stored kPanel x kPanel) from
// G = V^T V and tau (LAPACK larft, forward columnwise): T_ii = tau_i,
// T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i).
template <typename T>
__global__ __launch_bounds__(64) void qr_larft_kernel(const T *G, int64_t sg, const T *tau, int64_t stau, int j0,
                                                      T *Tf, int64_t stf, int nbc) {
  __shared__ double Ts[kPanel][kPanel + 1];
  __shared__ double Gs[kPanel][kPanel + 1];
  const int tid = threadIdx.x;
  const T *Gb = G + (int64_t)blockIdx.x * sg;
  const T *taub = tau + (int64_t)blockIdx.x * stau + j0;
  T *Tb = Tf + (int64_t)blockIdx.x * stf;
  for (int e = tid; e < kPanel * kPanel; e += 64) {
    const int r = e / kPanel, c = e % kPanel;
    Ts[r][c] = 0.0;
    Gs[r][c] = (r < nbc && c < nbc) ? (double)Gb[r * kPanel + c] : 0.0;
  }
  __syncthreads();
  for (int i = 0; i < nbc; ++i) {
    const double ti = (double)taub[i];
    if (tid < i) {
      double s = 0.0;
      for (int l = tid; l < i; ++l) s = fma(Ts[tid][l], Gs[l][i], s);
      Ts[tid][i] = -ti * s;
    }
    if (tid == i) Ts[i][i] = ti;
    __syncthreads();
  }
  for (int e = tid; e < kPanel * kPanel; e += 64) Tb[e] = (T)Ts[e / kPanel][e % kPanel];
}

// dst (rows x cols, ld_d) = src (ld_s) for each batch member; optional unit diagonal fill.
template <typename T>
__global__ void copy2d_kernel(const T *src, int64_t lds, int64_t ss, T *dst, int64_t ldd, int64_t sd, int rows,
                              int cols, int eye) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)rows * cols) return;
  const int r = (int)(e / cols), c = (int)(e - (int64_t)r * cols);
  const int64_t b = blockIdx.y;
  dst[b * sd + (int64_t)r * ldd + c] = eye ? (r == c ? T(1) : T(0)) : src[b * ss + (int64_t)r * lds + c];
}

template <typename T>
static int copy2d(const T *src, int64_t lds, int64_t ss, T *dst, int64_t ldd, int64_t sd, int rows, int cols,
                  int64_t batch, bool eye, hipStream_t st) {
  const int64_t e = (int64_t)rows * cols;
  if (e == 0 || batch == 0) return PSVD_OK;
  hipLaunchKernelGGL(copy2d_kernel<T>, dim3((unsigned)((e + 255) / 256), (unsigned)batch), dim3(256), 0, st, src,
                     lds, ss, dst, ldd, sd, rows, cols, eye ? 1 : 0);
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

template <typename T>
size_t qr_ws_bytes(int64_t batch, int m, int n) {
  const int64_t np = (n + kPanel - 1) / kPanel;
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  return al(sizeof(T) * batch * m * n) * 2 + al(sizeof(T) * batch * np * kPanel * kPanel) + al(sizeof(T) * batch * n) +
         al(sizeof(T) * batch * kPanel * kPanel) + al(sizeof(T) * batch * kPanel * n) * 2 + al(coop_scr_bytes<T>(batch, m)) +
         al(gemm_splitk_bytes(sizeof(T) == 8 ? PSVD_GEMM_F64 : PSVD_GEMM_F32, kPanel, n, batch));
}

#define PSVD_TRY(x)            \
  do {                         \
    const int _st = (x);       \
    if (_st != PSVD_OK) return _st; \
  } while (0)

template <typename T>
static int run_qr(int gmode, int64_t batch, int m, int n, const T *a, int64_t lda, int64_t sa, T *q, T *r,
                  void *ws, size_t ws_bytes, hipStream_t st) {
  if (ws_bytes < qr_ws_bytes<T>(batch, m, n)) return PSVD_EINVAL;
  const int64_t np = (n + kPanel - 1) / kPanel, mn = (int64_t)m * n;
  char *w = static_cast<char *>(ws);
  auto take = [&](size_t bytes) {
    char *p = w;
    w += (bytes + 255) & ~size_t(255);
    return reinterpret_cast<T *>(p);
  };
  T *F = take(sizeof(T) * batch * mn);
  T *V = take(sizeof(T) * batch * mn);
  T *Tf = take(sizeof(T) * batch * np * kPanel * kPanel);
  T *tau = take(sizeof(T) * batch * n);
  T *G = take(sizeof(T) * batch * kPanel * kPanel);
  T *W1 = take(sizeof(T) * batch * kPanel * n);
  T *W2 = take(sizeof(T) * batch * kPanel * n);
  double *scr = reinterpret_cast<double *>(take(coop_scr_bytes<T>(batch, m)));
  const size_t skb = gemm_splitk_bytes(sizeof(T) == 8 ? PSVD_GEMM_F64 : PSVD_GEMM_F32, kPanel, n, batch);
  void *sk = take(skb);
  const int64_t stf = np * kPanel * kPanel;
  PSVD_TRY(copy2d<T>(a, lda, sa, F, n, mn, m, n, batch, false, st));
  for (int64_t p = 0; p < np; ++p) {
    const int j0 = (int)(p * kPanel), nbc = n - j0 < kPanel ? n - j0 : kPanel, rows = m - j0, n2 = n - j0 - nbc;
    // tall panels: rows spread over several workgroups per matrix; short or very many: one each
    const int kr = rows > kPanelThreads ? coop_rows_per_thread<T>(batch, rows) : 0;
    if (kr == 1) {
      PSVD_TRY((launch_coop<T, 1>(F, n, mn, V, mn, tau, n, m, j0, nbc, batch, scr, st)));
    } else if (kr == 2) {
      PSVD_TRY((launch_coop<T, 2>(F, n, mn, V, mn, tau, n, m, j0, nbc, batch, scr, st)));
    } else if (kr == 4 && sizeof(T) == 4) {
      PSVD_TRY((launch_coop<T, sizeof(T) == 4 ? 4 : 2>(F, n, mn, V, mn, tau, n, m, j0, nbc, batch, scr, st)));
    } else {
      hipLaunchKernelGGL(qr_panel_kernel<T>, dim3((unsigned)batch), dim3(kPanelThreads), 0, st, F, (int64_t)n, mn, V,
                         mn, tau, (int64_t)n, m, j0, nbc);
      if (hipGetLastError() != hipSuccess) return PSVD_ELAUNCH;
    }
    T *Vp = V + (int64_t)j0 * n + j0;  // (rows x nbc), ld n
    PSVD_TRY(gemm_ws(gmode, 1, 0, nbc, nbc, rows, 1.0, Vp, n, mn, Vp, n, mn, 0.0, G, kPanel, kPanel * kPanel, batch,
                     sk, skb, st));
    hipLaunchKernelGGL(qr_larft_kernel<T>, dim3((unsigned)batch), dim3(64), 0, st, G, (int64_t)kPanel * kPanel, tau,
                       (int64_t)n, j0, Tf + p * kPanel * kPanel, stf, nbc);
    if (hipGetLastError() != hipSuccess) return PSVD_ELAUNCH;
    if (n2 > 0) {
      T *F2 = F + (int64_t)j0 * n + j0 + nbc;  // trailing block (rows x n2), ld n
      T *Tp = Tf + p * kPanel * kPanel;
      PSVD_TRY(gemm_ws(gmode, 1, 0, nbc, n2, rows, 1.0, Vp, n, mn, F2, n, mn, 0.0, W1, n, (int64_t)kPanel * n, batch,
                       sk, skb, st));
      PSVD_TRY(psvd_gemm(gmode, 1, 0, nbc, n2, nbc, 1.0, Tp, kPanel, stf, W1, n, (int64_t)kPanel * n, 0.0, W2, n,
                         (int64_t)kPanel * n, batch, st));
      PSVD_TRY(psvd_gemm(gmode, 0, 0, rows, n2, nbc, -1.0, Vp, n, mn, W2, n, (int64_t)kPanel * n, 1.0, F2, n, mn, batch,
                         st));
    }
  }
  if (r) PSVD_TRY(copy2d<T>(F, n, mn, r, n, (int64_t)n * n, n, n, batch, false, st));
  if (q) {
    PSVD_TRY(copy2d<T>(nullptr, 0, 0, q, n, mn, m, n, batch, true, st));
    for (int64_t p = np - 1; p >= 0; --p) {
      const int j0 = (int)(p * kPanel), nbc = n - j0 < kPanel ? n - j0 : kPanel, rows = m - j0, nq = n - j0;
      const T *Vp = V + (int64_t)j0 * n + j0;
      T *Qs = q + (int64_t)j0 * n + j0;
      T *Tp = Tf + p * kPanel * kPanel;
      PSVD_TRY(gemm_ws(gmode, 1, 0, nbc, nq, rows, 1.0, Vp, n, mn, Qs, n, mn, 0.0, W1, n, (int64_t)kPanel * n, batch,
                       sk, skb, st));
      PSVD_TRY(psvd_gemm(gmode, 0, 0, nbc, nq, nbc, 1.0, Tp, kPanel, stf, W1, n, (int64_t)kPanel * n, 0.0, W2, n,
                         (int64_t)kPanel * n, batch, st));
      PSVD_TRY(psvd_gemm(gmode, 0, 0, rows, nq, nbc, -1.0, Vp, n, mn, W2, n, (int64_t)kPanel * n, 1.0, Qs, n, mn, batch,
                         st));
    }
  }
  return PSVD_OK;
}

}  // namespace psvd

extern "C" {

size_t psvd_qr_workspace_bytes(int dtype, int64_t batch, int m, int n) {
  if (batch < 0 || m < 1 || n < 1 || n > m) return 0;
  if (dtype == PSVD_F32) return psvd::qr_ws_bytes<float>(batch, m, n);
  if (dtype == PSVD_F64) return psvd::qr_ws_bytes<double>(batch, m, n);
  return 0;
}

int psvd_qr(int dtype, int gemm_mode, int64_t batch, int m, int n, const void *a, int64_t lda, int64_t stride_a,
            void *q, void *r, void *workspace, size_t workspace_bytes, psvd_stream_t stream) {
  if (batch < 0 || m < 1 || n < 1 || n > m || lda < n || !a || !workspace) return PSVD_EINVAL;
  if (batch == 0) return PSVD_OK;
  if (batch > 65535) return PSVD_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == PSVD_F32) {
    if (gemm_mode != PSVD_GEMM_F32 && gemm_mode != PSVD_GEMM_BF16X3) return PSVD_EINVAL;
    return psvd::run_qr<float>(gemm_mode, batch, m, n, static_cast<const float *>(a), lda, stride_a,
                               static_cast<float *>(q), static_cast<float *>(r), workspace, workspace_bytes, st);
  }
  if (dtype == PSVD_F64)
    return psvd::run_qr<double>(PSVD_GEMM_F64, batch, m, n, static_cast<const double *>(a), lda, stride_a,
                                static_cast<double *>(q), static_cast<double *>(r), workspace, workspace_bytes, st);
  return PSVD_EINVAL;
}

}  // extern "C"
