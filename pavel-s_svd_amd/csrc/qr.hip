// K4: blocked Householder QR of a batch of tall matrices, A = Q R with thin Q (m x n) and
// R (n x n), in the reference's reflector convention (LAPACK xLARFG: beta = -sign(alpha) *
// hypot(alpha, |x2|), tau = (beta - alpha) / beta, v = x / (alpha - beta), v0 = 1, and the
// trivial reflector tau = 0 when x2 = 0), so R matches primitives.qr up to rounding.
//
// Per panel of kPanel columns: the reflectors, tau and the compact-WY factor T (H_1...H_k = I - V T V^T,
// LAPACK larft, forward columnwise) come from the CholeskyQR panel (qr_panel_cqr_kernel: CholeskyQR2 in
// FP64 over the panel's workgroups and Householder reconstruction, T included) or, where it does not
// apply, from a column-by-column panel (several workgroups per matrix, or one) followed by a GEMM for
// G = V^T V and qr_larft_kernel; three GEMMs then apply Q^T = I - V T^T V^T to the trailing columns.
// Q is accumulated backwards from [I; 0] with the same three GEMMs per panel (orgqr).
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

// Per matrix: the column-by-column partials (CoopScr), then the CholeskyQR panel's two nw x 32 x 32 Gram
// partial areas and nw column-mask words (qr_panel_cqr_kernel).
struct PanelScr {
  __host__ __device__ static size_t doubles(int nw) {
    return CoopScr::doubles(nw) + (size_t)2 * nw * kPanel * kPanel + (size_t)nw;
  }
};

// The column-by-column panel of matrix b (workgroup wi of its nw), the body of qr_panel_coop_kernel and
// the fallback of qr_panel_cqr_kernel; smem holds the thread's KR rows (coop_lds_bytes).
template <typename T, int KR>
__device__ void coop_panel(T *F, int64_t ldf, int64_t sf, T *V, int64_t sv, T *tau, int64_t stau, int m, int j0,
                           int nbc, int nw, int b, int wi, double *scr, GridBarrier &gb, double *smem_coop) {
  constexpr int NT = 256;
  __shared__ double red[(NT / kWave) * kPanel];
  __shared__ double wsum[kPanel];
  __shared__ double seg[NT / kCoopW][kCoopW];
  const int tid = threadIdx.x;
  bool aborted = false;
  T *Fb = F + (int64_t)b * sf;
  T *Vb = V + (int64_t)b * sv;
  T *tb = tau + (int64_t)b * stau;
  double *sb = scr + (size_t)b * PanelScr::doubles(nw);
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

template <typename T, int KR>
__global__ __launch_bounds__(256) void qr_panel_coop_kernel(T *F, int64_t ldf, int64_t sf, T *V, int64_t sv, T *tau,
                                                            int64_t stau, int m, int j0, int nbc, int nw,
                                                            double *scr, unsigned *bar) {
  extern __shared__ double smem_coop[];
  const int b = blockIdx.x / nw, wi = blockIdx.x - b * nw;
  GridBarrier gb(bar + (size_t)b * kBarWords, (unsigned)nw, (unsigned)wi);   // the matrix's workgroups
  coop_panel<T, KR>(F, ldf, sf, V, sv, tau, stau, m, j0, nbc, nw, b, wi, scr, gb, smem_coop);
}

// ---- tall panels by CholeskyQR2 and Householder reconstruction ----------------------------------------
// The same reflectors, tau and R as the column-by-column panel, from a factorisation with two barriers
// instead of one per column: the panel (rows of workgroup wi, one per thread) is orthogonalised by
// CholeskyQR2 in FP64 -- G = X^T X by FP64 MFMA per workgroup, the partials added in workgroup order
// by every workgroup, R1 = chol(G), Q1 = X R1^-1, the same again for Q1 (R2), Q = Q1 R2^-1, R = R2 R1 --
// and the Householder vectors are rebuilt from Q (Ballard, Demmel, Grigori, Jacquelin, Knight, Nguyen,
// "Reconstructing Householder vectors from tall-skinny QR", 2015): the LU without pivoting of the top
// n x n block of Q - S, the sign s_k = -sign(pivot) chosen at each step so |u_kk| = |pivot| + 1, gives
// Y = [L1; Q2 U^-1], tau_k = -u_kk s_k (the diagonal of T = -U S L1^-T) and R_HH = S R: LAPACK's
// xLARFG convention (tau in [1, 2], beta = -sign(alpha) |x|), the reference's primitives.qr.  Every
// workgroup forms the small factors itself from the same sums in the same order, so they agree bit for
// bit and need no broadcast.  Panels this cannot take exactly -- a Cholesky pivot that is not positive,
// G2 = Q1^T Q1 further than 0.5 from I (condition beyond about 1e8), a column with no nonzero below its
// diagonal, or a reconstruction pivot of exactly +-1 (where LAPACK takes the trivial reflector) -- run
// the column-by-column panel (coop_panel) instead; the decision is the same in every workgroup.
constexpr int kCqrLd = kPanel + 1, kCqrMaxNw = 64;
typedef double q4d __attribute__((ext_vector_type(4)));

template <typename T>
__host__ __device__ constexpr size_t cqr_lds_bytes() {
  // 256 rows, then seven 32 x 32 areas (G/R1, G2/R2, the top block of Q and its LU, R = R2 R1, R1^-1,
  // R2^-1, U^-1)
  return sizeof(double) * ((size_t)256 + 7 * kPanel) * kCqrLd;
}

// 32 x 32 Gram partial of the workgroup's 256 rows in FP64 MFMA (wave w: the 16 x 16 tile (w / 2, w % 2)),
// stored workgroup-major (entry e = 32 i + j of workgroup wi at part[1024 wi + e]); the barrier that
// follows writes them back from L2 and invalidates the readers' caches, and the sums read them with
// coalesced loads.
__device__ __forceinline__ void cqr_gram_partial(const double *Xd, double *part, int nw, int wi) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int lr = lane & 15, lk = lane >> 4, ti = wave >> 1, tj = wave & 1;
  q4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int k = 0; k < 256; k += 4) {
    const double a = Xd[(k + lk) * kCqrLd + 16 * ti + lr], bb = Xd[(k + lk) * kCqrLd + 16 * tj + lr];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb, acc, 0, 0, 0);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {   // FP64 MFMA D layout: register q of lane l is row 4 q + l / 16
    const int i = 16 * ti + 4 * q + lk, j = 16 * tj + lr;
    part[(size_t)wi * kPanel * kPanel + i * kPanel + j] = acc[q];
  }
}

// The Gram from the nw partials, added in workgroup order: thread t forms entries t + 256 r (r < 4), its
// loads coalesced across the wave (consecutive entries), eight workgroups' partials in flight.  With
// CHECK, *bad is set when an entry of the leading n x n block is further than 0.5 from the identity's and
// *emax receives the largest such distance.  Entries of padding columns (index >= n) are the identity's.
template <bool CHECK>
__device__ __forceinline__ void cqr_gram_sum(const double *part, int nw, int n, double *A, int *bad,
                                             unsigned long long *emax = nullptr) {
  const int t = threadIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  int w = 0;
  for (; w + 8 <= nw; w += 8) {
    double v[8][4];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int r = 0; r < 4; ++r) v[u][r] = part[(size_t)(w + u) * kPanel * kPanel + t + 256 * r];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] += v[u][r];
  }
  for (; w < nw; ++w)
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] += part[(size_t)w * kPanel * kPanel + t + 256 * r];
  double dev = 0.0;
  bool far = false;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int e = t + 256 * r, i = e / kPanel, j = e - i * kPanel;
    const double id = i == j ? 1.0 : 0.0;
    if (i >= n || j >= n) acc[r] = id;
    else if (CHECK) {
      const double d = fabs(acc[r] - id);
      far = far || !(d <= 0.5);
      dev = fmax(dev, d);
    }
    A[i * kCqrLd + j] = acc[r];
  }
  if (CHECK) {
    dev = wave_max_fast(dev);
    const bool anyfar = __any(far);
    if ((t & (kWave - 1)) == 0) {
      if (anyfar) atomicOr(bad, 1);
      atomicMax(emax, (unsigned long long)__double_as_longlong(dev));
    }
  }
}

// LDS writes of this wave become visible to its own later LDS reads (a wave's LDS operations run in
// order; the fences keep the compiler from moving them).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 1/sqrt(d) and 1/u from the hardware approximations with Newton steps (relative error near FP64's)
__device__ __forceinline__ double rsqrt_nr(double d) {
  double y = __builtin_amdgcn_rsq(d);
  y = y * fma(-0.5 * d * y, y, 1.5);
  return y * fma(-0.5 * d * y, y, 1.5);
}
__device__ __forceinline__ double rcp_nr(double u) {
  double y = __builtin_amdgcn_rcp(u);
  y = y * fma(-u, y, 2.0);
  return y * fma(-u, y, 2.0);
}

// The three small routines below run over all 32 indices (the panel's n <= 32 columns padded with zero
// columns, whose factors are the identity), so they need no bounds tests, and as rolled loops: register
// arrays keep compile-time indices by shifting one place per step (entry 0 is always the current index
// k), which keeps the kernel small enough for the instruction cache.  Reads past index 31 land in the
// next LDS row or area and only feed register slots that are never used.

// Broadcast of lane `src`'s double to the whole wave (two v_readlane with a scalar lane index).
__device__ __forceinline__ double lane_double(double v, int src) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), src),
                          __builtin_amdgcn_readlane(__double2loint(v), src));
}

// Upper Cholesky factor R (A = R^T R) of A (32 x 32, LDS) in place, by one wave: lane j holds column j
// (rows k.. at step k); at step k every lane writes its entry of row k of R to A and reads the row back
// (LDS round trips measured faster than 31 readlane broadcasts on this chain); rinv[k] = 1 / R[k][k].
// Returns false unless every pivot was positive and finite.  Wave 0 only.
__device__ bool cqr_chol_wave(double *A, double *rinv) {
  const int lane = threadIdx.x, col = lane < kPanel ? lane : kPanel - 1;
  double a[kPanel];
#pragma unroll
  for (int i = 0; i < kPanel; ++i) a[i] = A[i * kCqrLd + col];
  bool ok = true;
#pragma unroll 1
  for (int k = 0; k < kPanel; ++k) {
    const double d = lane_double(a[0], k);
    ok = ok && d > 0.0 && d <= DBL_MAX;
    const double ri = rsqrt_nr(d), rk = lane == k ? d * ri : a[0] * ri;   // R[k][lane]
    if (lane == 0) rinv[k] = ri;
    if (lane < kPanel) A[k * kCqrLd + lane] = lane >= k ? rk : 0.0;
    wave_lds_sync();
    const double *row = A + k * kCqrLd + k;   // R[k][k + i] at row[i]
    double r[kPanel];
#pragma unroll
    for (int i = 1; i < kPanel; ++i) r[i] = row[i];   // all of row k in flight at once
#pragma unroll
    for (int i = 1; i < kPanel; ++i) a[i - 1] = fma(-r[i], rk, a[i]);   // A[k+i][j] -= R[k][k+i] R[k][j]
    a[kPanel - 1] = 0.0;
  }
  return ok;
}

// io <- io R^-1 for a row io (the thread's LDS row) and the upper triangular R (LDS) with diagonal
// reciprocals rinv.
__device__ __forceinline__ void cqr_row_solve(double *io, const double *R, const double *rinv) {
  double x[kPanel];
#pragma unroll
  for (int k = 0; k < kPanel; ++k) x[k] = io[k];
#pragma unroll 1
  for (int k = 0; k < kPanel; ++k) {
    const double xk = x[0] * rinv[k];
    io[k] = xk;
    const double *row = R + k * kCqrLd + k;
    double r[kPanel];
#pragma unroll
    for (int i = 1; i < kPanel; ++i) r[i] = row[i];
#pragma unroll
    for (int i = 1; i < kPanel; ++i) x[i - 1] = fma(-xk, r[i], x[i]);
    x[kPanel - 1] = 0.0;
  }
}

// One 16 x 16 tile of X B (rows r0.., columns c0.. ; X and B in LDS with row stride kCqrLd, K = 32) on
// FP64 MFMA; every caller of a tile gets the same bits.
__device__ __forceinline__ q4d cqr_tile(const double *X, int r0, const double *B, int c0) {
  const int lane = threadIdx.x & (kWave - 1), lr = lane & 15, lk = lane >> 4;
  q4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int kk = 0; kk < kPanel; kk += 4)
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(X[(r0 + lr) * kCqrLd + kk + lk], B[(kk + lk) * kCqrLd + c0 + lr], acc,
                                               0, 0, 0);
  return acc;
}

// X <- X B for the rows r0 .. r0 + 16 RT - 1 of X (LDS, stride kCqrLd), one wave: every tile is formed
// before any is written back.
template <int RT>
__device__ __forceinline__ void cqr_rows_times(double *X, int r0, const double *B) {
  const int lane = threadIdx.x & (kWave - 1), lr = lane & 15, lk = lane >> 4;
  q4d acc[RT][2];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = cqr_tile(X, r0 + 16 * rt, B, 16 * ct);
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
      for (int q = 0; q < 4; ++q)   // FP64 MFMA D layout: register q of lane l is row 4 q + l / 16
        X[(r0 + 16 * rt + 4 * q + lk) * kCqrLd + 16 * ct + lr] = acc[rt][ct][q];
}

// Rows of the inverse of the upper triangular R (thread t < 32: row t of R^-1, in Ri).
__device__ __forceinline__ void cqr_tri_inverse(double *Ri, const double *R, const double *rinv) {
  const int t = threadIdx.x;
  if (t < kPanel) {
    for (int k = 0; k < kPanel; ++k) Ri[t * kCqrLd + k] = k == t ? 1.0 : 0.0;
    cqr_row_solve(Ri + t * kCqrLd, R, rinv);
  }
}

// LU without pivoting of M (32 x 32, LDS) with the reconstruction's signs (one wave, lane j holds
// column j, rows k.. at step k): the Schur pivot q becomes u_kk = q - s_k with s_k = -1 for q >= 0, +1
// for q < 0; lane k writes column k of L1 (scaled by 1/u_kk) as a contiguous row of the scratch LT
// (paired 16-byte stores), every lane writes its entry of row k of U to M and reads the L column back
// from LT.  M then holds U (with its diagonal) and L1 (strictly below, copied from LT at the end);
// sg[k] = s_k, uinv[k] = 1 / u_kk.  Returns false when a pivot k < n is exactly +-1 (LAPACK's trivial
// reflector) or not finite.
__device__ bool cqr_lu_wave(double *M, double *LT, int n, double *sg, double *uinv) {
  const int lane = threadIdx.x, col = lane < kPanel ? lane : kPanel - 1;
  double a[kPanel];
#pragma unroll
  for (int i = 0; i < kPanel; ++i) a[i] = M[i * kCqrLd + col];
  bool ok = true;
#pragma unroll 1
  for (int k = 0; k < kPanel; ++k) {
    const double q = lane_double(a[0], k);
    ok = ok && (k >= n || (fabs(q) != 1.0 && fabs(q) <= DBL_MAX));   // padding rows are zero (u = 1)
    const double s = q >= 0.0 ? -1.0 : 1.0, u = q - s, ui = rcp_nr(u);
    const double ukj = lane == k ? u : a[0];   // U[k][lane] for lanes >= k
    if (lane == 0) {
      sg[k] = s;
      uinv[k] = ui;
    }
    double *lrow = LT + k * kCqrLd;   // LT[k][i] = L[k + i][k]
    if (lane == k) {
#pragma unroll
      for (int i = 1; i < kPanel; ++i) lrow[i] = a[i] * ui;   // past index 31: unused slots of the row
    }
    if (lane < kPanel && lane >= k) M[k * kCqrLd + lane] = ukj;
    wave_lds_sync();
    double l[kPanel];
#pragma unroll
    for (int i = 1; i < kPanel; ++i) l[i] = lrow[i];
#pragma unroll
    for (int i = 1; i < kPanel; ++i) a[i - 1] = fma(-l[i], ukj, a[i]);
    a[kPanel - 1] = 0.0;
  }
  if (lane < kPanel)   // L1 into the strict lower part of M: M[i][lane] = LT[lane][i - lane]
    for (int i = lane + 1; i < kPanel; ++i) M[i * kCqrLd + lane] = LT[lane * kCqrLd + (i - lane)];
  return ok;
}

template <typename T>
__global__ __launch_bounds__(256) void qr_panel_cqr_kernel(T *F, int64_t ldf, int64_t sf, T *V, int64_t sv, T *tau,
                                                           int64_t stau, int m, int j0, int nbc, int nw, double *scr,
                                                           unsigned *bar, T *Tf, int64_t stf) {
  constexpr int NT = 256, LD = kCqrLd;
  extern __shared__ double smem_cqr[];
  double *Xd = smem_cqr;                 // the workgroup's rows, FP64 [256][LD]
  double *R1 = Xd + NT * LD;             // G, then its factor R1
  double *R2 = R1 + kPanel * LD;         // G2, then R2
  double *Mq = R2 + kPanel * LD;         // top n x n block of Q, then its LU
  double *Rr = Mq + kPanel * LD;         // R = R2 R1
  double *Ri1 = Rr + kPanel * LD;        // R1^-1
  double *Ri2 = Ri1 + kPanel * LD;       // R2^-1
  double *Ui = Ri2 + kPanel * LD;        // U^-1
  __shared__ double r1inv[kPanel], r2inv[kPanel], sg[kPanel], uinv[kPanel];
  __shared__ unsigned long long s_emax;
  __shared__ unsigned s_mask;
  __shared__ int s_bad;
  const int tid = threadIdx.x, n = nbc;
  const int b = blockIdx.x / nw, wi = blockIdx.x - b * nw;
  GridBarrier gb(bar + (size_t)b * kBarWords, (unsigned)nw, (unsigned)wi);
  T *Fb = F + (int64_t)b * sf;
  T *Vb = V + (int64_t)b * sv;
  T *tb = tau + (int64_t)b * stau;
  double *sb = scr + (size_t)b * PanelScr::doubles(nw);
  double *g1p = sb + CoopScr::doubles(nw), *g2p = g1p + (size_t)nw * kPanel * kPanel;
  unsigned *mkp = reinterpret_cast<unsigned *>(g2p + (size_t)nw * kPanel * kPanel);   // nw words
  const int pr = wi * NT + tid, row = j0 + pr;   // the thread's row, panel-relative and absolute
  const bool live = row < m;
  // ---- rows in FP64; the columns with a nonzero below their diagonal
  unsigned mask = 0;
  {
    T v[kPanel];
#pragma unroll
    for (int k = 0; k < kPanel; ++k) v[k] = (live && k < n) ? Fb[(int64_t)row * ldf + j0 + k] : T(0);
#pragma unroll
    for (int k = 0; k < kPanel; ++k) {
      if (v[k] != T(0) && pr > k) mask |= 1u << k;
      Xd[tid * LD + k] = (double)v[k];
    }
  }
  if (tid == 0) {
    s_mask = 0;
    s_bad = 0;
    s_emax = 0ull;
  }
  __syncthreads();
#pragma unroll
  for (int o = kWave / 2; o > 0; o /= 2) mask |= (unsigned)__shfl_xor((int)mask, o, kWave);
  if ((tid & (kWave - 1)) == 0 && mask) atomicOr(&s_mask, mask);   // one LDS atomic per wave
  cqr_gram_partial(Xd, g1p, nw, wi);
  __syncthreads();
  if (tid == 0) mkp[wi] = s_mask;
  bool fast = true;
  if (!gb.sync<true>()) {
    for (int k = 0; k < n; ++k) if (live) Fb[(int64_t)row * ldf + j0 + k] = T(__builtin_nan(""));
    if (wi == 0 && tid < n) tb[j0 + tid] = T(__builtin_nan(""));
    return;
  }
  // ---- pass 1: G, R1 = chol(G), Q1 = X R1^-1
  cqr_gram_sum<false>(g1p, nw, n, R1, &s_bad);
  if (tid == 0) {
    unsigned all = 0;
    for (int w = 0; w < nw; ++w) all |= mkp[w];
    const unsigned need = n >= 32 ? 0xffffffffu : ((1u << n) - 1u);
    if ((all & need) != need) s_bad = 1;
  }
  __syncthreads();
  if (tid < kWave) {
    const bool ok = cqr_chol_wave(R1, r1inv);
    if (tid == 0 && !ok) s_bad = 1;
  }
  __syncthreads();
  fast = s_bad == 0;
  const int wave = tid / kWave;
  if (fast) {
    cqr_tri_inverse(Ri1, R1, r1inv);
    __syncthreads();
    cqr_rows_times<4>(Xd, 64 * wave, Ri1);   // Q1 = X R1^-1
    __syncthreads();
    cqr_gram_partial(Xd, g2p, nw, wi);
    if (!gb.sync<true>()) {
      for (int k = 0; k < n; ++k) if (live) Fb[(int64_t)row * ldf + j0 + k] = T(__builtin_nan(""));
      if (wi == 0 && tid < n) tb[j0 + tid] = T(__builtin_nan(""));
      return;
    }
    // ---- pass 2: G2 = I + E, R2, Q = Q1 R2^-1; R = R2 R1.  When max|E| <= 1e-8 (condition of the panel
    // below about 1e4) R2 = I + strict_upper(E) + diag(E) / 2 and R2^-1 = I - (R2 - I) to O(E^2) = 1e-16;
    // else the Cholesky factor and its inverse.
    cqr_gram_sum<true>(g2p, nw, n, R2, &s_bad, &s_emax);
    __syncthreads();
    if (__longlong_as_double((long long)s_emax) <= 1e-8) {
      for (int e = tid; e < kPanel * kPanel; e += NT) {
        const int i = e / kPanel, j = e - i * kPanel;
        const double g = R2[i * LD + j];
        const double r = i < j ? g : (i == j ? 1.0 + 0.5 * (g - 1.0) : 0.0);
        Ri2[i * LD + j] = i < j ? -g : (i == j ? 1.0 - 0.5 * (g - 1.0) : 0.0);
        R2[i * LD + j] = r;   // entry e is read and written by this thread only
      }
    } else {
      if (tid < kWave) {
        const bool ok = cqr_chol_wave(R2, r2inv);
        if (tid == 0 && !ok) s_bad = 1;
      }
      __syncthreads();
      cqr_tri_inverse(Ri2, R2, r2inv);
    }
    __syncthreads();
    fast = s_bad == 0;
  }
  if (fast) {
    cqr_rows_times<4>(Xd, 64 * wave, Ri2);   // Q = Q1 R2^-1
    // the top n rows of Q, formed by every workgroup from the panel's first rows exactly as their owner
    // forms them, then their LU with the reconstruction's signs
    if (tid < kPanel)
#pragma unroll
      for (int k = 0; k < kPanel; ++k)
        Mq[tid * LD + k] = (tid < n && k < n) ? (double)Fb[(int64_t)(j0 + tid) * ldf + j0 + k] : 0.0;
    __syncthreads();
    if (wave == 0) {
      cqr_rows_times<2>(Mq, 0, Ri1);
      cqr_rows_times<2>(Mq, 0, Ri2);
    }
    if (tid >= n && tid < kPanel)   // padding rows of the top block: zero
      for (int k = 0; k < kPanel; ++k) Mq[tid * LD + k] = 0.0;
    __syncthreads();
    if (tid < kWave) {
      const bool ok = cqr_lu_wave(Mq, Rr, n, sg, uinv);   // Rr: scratch until R is formed
      if (tid == 0 && !ok) s_bad = 1;
    }
    __syncthreads();
    fast = s_bad == 0;
  }
  T *Tb = Tf + (int64_t)b * stf;   // the panel's compact-WY factor (kPanel x kPanel, row-major)
  if (!fast) {
    __syncthreads();   // every thread is past its reads of the LDS rows before they are reused
    coop_panel<T, 1>(F, ldf, sf, V, sv, tau, stau, m, j0, nbc, nw, b, wi, scr, gb, smem_cqr);
    // T from the reflectors (LAPACK larft, forward columnwise): G = V^T V over the workgroups' rows of V
    // (each workgroup rereads the rows it wrote), then T_ii = tau_i, T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i)
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPanel; ++k) Xd[tid * LD + k] = (live && k < n) ? (double)Vb[(int64_t)row * ldf + j0 + k] : 0.0;
    __syncthreads();
    cqr_gram_partial(Xd, g1p, nw, wi);
    if (!gb.sync<true>()) return;   // abandoned: coop_panel has left the panel NaN
    cqr_gram_sum<false>(g1p, nw, n, R1, &s_bad);
    __syncthreads();
    if (wi == 0) {
      double *Ts = R2;   // T, row-major in LDS
      for (int e = tid; e < kPanel * kPanel; e += NT) Ts[(e / kPanel) * LD + e % kPanel] = 0.0;
      __syncthreads();
      for (int i = 0; i < n; ++i) {
        const double ti = (double)tb[j0 + i];
        if (tid < i) {
          double acc = 0.0;
          for (int l = tid; l < i; ++l) acc = fma(Ts[tid * LD + l], R1[l * LD + i], acc);
          Ts[tid * LD + i] = -ti * acc;
        }
        if (tid == i) Ts[i * LD + i] = ti;
        __syncthreads();
      }
      for (int e = tid; e < kPanel * kPanel; e += NT) Tb[e] = (T)Ts[(e / kPanel) * LD + e % kPanel];
    }
    return;
  }
  // ---- outputs: Y = [L1; Q2 U^-1] to V, R_HH = S R2 R1 over zeros to the panel of F, tau_k = -u_kk s_k
  if (wi == 0) {
    for (int e = tid; e < kPanel * kPanel; e += NT) {   // R = R2 R1 (upper)
      const int i = e / kPanel, j = e - i * kPanel;
      double acc = 0.0;
      if (i <= j && j < n)
        for (int k = i; k <= j; ++k) acc = fma(R2[i * LD + k], R1[k * LD + j], acc);
      Rr[i * LD + j] = acc;
    }
    if (tid < n) tb[j0 + tid] = (T)(-Mq[tid * LD + tid] * sg[tid]);
    // T = -U S L1^-T: row t of T solves t L1^T = -(U S)_t (L1^T unit upper, in Ri1 -- free by now)
    for (int e = tid; e < kPanel * kPanel; e += NT) {
      const int i = e / kPanel, j = e - i * kPanel;
      Ri1[i * LD + j] = i == j ? 1.0 : (i < j && j < n ? Mq[j * LD + i] : 0.0);
      Ri2[i * LD + j] = (i <= j && j < n) ? -Mq[i * LD + j] * sg[j] : 0.0;
    }
    if (tid < kPanel) r1inv[tid] = 1.0;
  }
  __syncthreads();
  cqr_tri_inverse(Ui, Mq, uinv);   // U^-1 (U: the upper part of Mq), wave 0
  if (wi == 0 && tid >= kWave && tid < kWave + kPanel) {   // T, wave 1 of workgroup 0, beside it
    const int t = tid - kWave;
    cqr_row_solve(Ri2 + t * LD, Ri1, r1inv);
    for (int k = 0; k < kPanel; ++k) Tb[t * kPanel + k] = (t < n && k < n) ? (T)Ri2[t * LD + k] : T(0);
  }
  __syncthreads();
  cqr_rows_times<4>(Xd, 64 * wave, Ui);   // Y = Q U^-1 (kept for the rows below the top block)
  __syncthreads();
  if (!live) return;
  T *frow = Fb + (int64_t)row * ldf + j0, *vrow = Vb + (int64_t)row * ldf + j0;
  if (pr < n) {   // workgroup 0: the top rows
    for (int k = 0; k < n; ++k) {
      frow[k] = k >= pr ? (T)(sg[pr] * Rr[pr * LD + k]) : T(0);
      vrow[k] = k < pr ? (T)Mq[pr * LD + k] : (k == pr ? T(1) : T(0));
    }
  } else {
    for (int k = 0; k < n; ++k) {
      frow[k] = T(0);
      vrow[k] = (T)Xd[tid * LD + k];
    }
  }
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
  return sizeof(double) * (size_t)batch * PanelScr::doubles((int)nw) + sizeof(unsigned) * (size_t)batch * kBarWords;
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
  unsigned *bar = reinterpret_cast<unsigned *>(scr + (size_t)batch * PanelScr::doubles((m + 255) / 256));
  hipLaunchKernelGGL(bar_reset_kernel, dim3(1), dim3(256), 0, st, bar, (int)(batch * kBarWords));
  const size_t lds = coop_lds_bytes<T, KR>();
  hipLaunchKernelGGL((qr_panel_coop_kernel<T, KR>), dim3((unsigned)(nw * batch)), dim3(256), lds, st, F, ldf, sf, V,
                     sv, tau, stau, m, j0, nbc, nw, scr, bar);
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

template <typename T>
static int launch_cqr(T *F, int64_t ldf, int64_t sf, T *V, int64_t sv, T *tau, int64_t stau, int m, int j0, int nbc,
                      int64_t batch, double *scr, T *Tp, int64_t stf, hipStream_t st) {
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(qr_panel_cqr_kernel<T>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)cqr_lds_bytes<T>()) != hipSuccess)
      return PSVD_ELAUNCH;
    attr = true;
  }
  static const int limit = barrier_grid_limit(reinterpret_cast<const void *>(qr_panel_cqr_kernel<T>), 256,
                                              cqr_lds_bytes<T>());
  const int nw = (m - j0 + 255) / 256;
  if (nw * batch > limit) return PSVD_ELAUNCH;
  unsigned *bar = reinterpret_cast<unsigned *>(scr + (size_t)batch * PanelScr::doubles((m + 255) / 256));
  hipLaunchKernelGGL(bar_reset_kernel, dim3(1), dim3(256), 0, st, bar, (int)(batch * kBarWords));
  const size_t lds = cqr_lds_bytes<T>();
  hipLaunchKernelGGL(qr_panel_cqr_kernel<T>, dim3((unsigned)(nw * batch)), dim3(256), lds, st, F, ldf, sf, V, sv, tau,
                     stau, m, j0, nbc, nw, scr, bar, Tp, stf);
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
    // the CholeskyQR panel whenever the matrices' workgroups fit on the CUs (short panels included: one
    // or two workgroups per matrix); else tall panels column by column over several workgroups, and
    // short panels of large batches one workgroup per matrix
    const int kr0 = coop_rows_per_thread<T>(batch, rows);
    const int kr = rows > kPanelThreads ? kr0 : 0;
    const bool cqr = kr0 == 1 && (rows + 255) / 256 <= kCqrMaxNw;   // writes the panel's T itself
    if (cqr) {
      PSVD_TRY((launch_cqr<T>(F, n, mn, V, mn, tau, n, m, j0, nbc, batch, scr, Tf + p * kPanel * kPanel, stf, st)));
    } else if (kr == 1) {
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
    if (!cqr) {
      PSVD_TRY(gemm_ws(gmode, 1, 0, nbc, nbc, rows, 1.0, Vp, n, mn, Vp, n, mn, 0.0, G, kPanel, kPanel * kPanel, batch,
                       sk, skb, st));
      hipLaunchKernelGGL(qr_larft_kernel<T>, dim3((unsigned)batch), dim3(64), 0, st, G, (int64_t)kPanel * kPanel, tau,
                         (int64_t)n, j0, Tf + p * kPanel * kPanel, stf, nbc);
      if (hipGetLastError() != hipSuccess) return PSVD_ELAUNCH;
    }
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
