// K2b: blocked one-sided Jacobi SVD of one large tall matrix (m x n, m >= n).
//
// The columns of W (m x N, N = n padded to a multiple of 2 KB with zero columns that never rotate)
// are grouped in blocks of KB = 16 or 32; round r of a tournament over the nb blocks pairs block I
// with block J.  For each pair p the KS x KS (KS = 2 KB) Gram matrix S_p = [W_I W_J]^T [W_I W_J] is
// formed in FP64 on MFMA (row slices of W, per-slice partials added in a fixed order), and S_p is
// swept by `inner_sweeps` cyclic sweeps of two-sided plane rotations in LDS -- the rotation of pair
// (a, b) of S is exactly the one-sided rotation of columns a and b of W (app = s_aa, aqq = s_bb,
// apq = s_ab), with the reference's test apq^2 > tol^2 app aqq and min(app, aqq) > eps^2 max_j |w_j|^2
// (primitives.py:337-338; the column-norm maximum of the round's start, published by the previous
// round's pair sweeps).  The accumulated KS x KS rotation Q_p then updates W[:, I u J] and V[:, I u J]
// on MFMA.  A sweep without a rotation ends the iteration (primitives.py:352).  The tail -- FP64
// sigma, descending order, cut at eps max(m, n) sigma_1 (or eps sigma_1), U = W / sigma,
// coordinate-vector completion, row-major outputs -- follows primitives.py:356-366 and :430-455.
//
// Two launches per round g: bj_fused_kernel applies round g - 1's rotations to W (double-buffered)
// and forms round g's Gram partials from the rotated rows in the same workgroup; bj_inner_kernel runs
// the pair sweeps (one 1024-thread workgroup per pair; one barrier per rotation round in the cross
// rounds) beside round g - 1's V tiles.  The host queues sweep s + 1 before it reads sweep s's flag.
//
// Work per round: 2 m KS^2 FP64 MFMA flops per pair for the Gram, 2 (m + N) KS^2 for the updates;
// nb - 1 rounds per sweep, so the work per sweep does not depend on KB.  What KB sets is the
// parallelism and the serial chain: a round runs nb / 2 pair sweeps side by side, each KS - 1
// rotation rounds of O(KS^2) LDS traffic.  Blocks of 16 give twice the pairs per round and a pair
// sweep of 31 short rounds instead of 63 rounds four times as heavy.
#include <float.h>

#include "common.hpp"

namespace psvd {

constexpr int kJbThreads = 256;

// the rotation flag of sweep sw (1-based), by sw mod 3: [3], [4], [0] for remainders 1, 2, 0
__host__ __device__ inline int bj_flag_slot(int sw) {
  const int k = sw % 3;
  return k == 1 ? 3 : (k == 2 ? 4 : 0);
}

// Block width KB: subproblem size KS, LDS row stride, threads of the pair-sweep kernel (one
// upper-triangle 2 x 2 block of a rotation round per thread: 528 for KB = 32, 136 for KB = 16; for
// KB = 16 the 512 Qt entries of a round get threads of their own, 1024 in all).
template <int KB> struct BjShape {
  static constexpr int KS = 2 * KB, LD = KS + 1, INNER = KB == 32 ? 576 : 1024;
};

typedef float j4f __attribute__((ext_vector_type(4)));
typedef double j4d __attribute__((ext_vector_type(4)));

template <typename T>
struct BjParams {
  const T *a;
  T *u, *s, *vh;
  int32_t *info;
  int m, n, N, nb, splits, inner_sweeps, flags;
  int full_pairs;           // every round's pair sweeps over all pairs (else the first round's only)
  double tol;
  T *W, *V, *Q;             // W: m x N, V: N x N (row-major, ld N); Q: [2][npair][KS][KS] by round parity
  T *W2;                    // W's second buffer: round g's update reads one buffer and writes the other
  unsigned long long *fmx;  // [2]: max column norm^2 bits at the start of the round, by round parity
  double *Sp;               // npair x splits x 64 x 64 Gram partials
  int *act;                 // [2][npair] by round parity: the pair rotated in its sweep
  unsigned long long *slot;  // [0] max|a| bits, [1] max column norm^2 bits (this sweep)
  int *flag;                // [3] / [4] / [0] rotation in sweep sw with sw mod 3 = 1 / 2 / 0, [1] non-finite
                            // input, [2] completion status
  double *sig;              // N
  int *ord;                 // N
  int32_t *nk;              // 1
};

template <int KB>
__device__ __forceinline__ int bj_col(int I, int J, int i) { return i < KB ? I * KB + i : J * KB + (i - KB); }

template <typename T>
__global__ __launch_bounds__(kJbThreads) void bj_absmax_kernel(BjParams<T> P) {
  double mx = 0.0;
  int bad = 0;
  for (int64_t e = (int64_t)blockIdx.x * kJbThreads + threadIdx.x; e < (int64_t)P.m * P.n;
       e += (int64_t)gridDim.x * kJbThreads) {
    const double x = (double)P.a[e];
    bad |= !isfinite(x);
    mx = fmax(mx, fabs(x));
  }
  // the workgroup's maximum (one atomic per workgroup: the launch has at most 256 of them)
  __shared__ double wmx[kJbThreads / kWave];
  __shared__ int wbad[kJbThreads / kWave];
  mx = wave_max_fast(mx);
  const bool wave_bad = __any(bad);
  if ((threadIdx.x & (kWave - 1)) == 0) {
    wmx[threadIdx.x / kWave] = mx;
    wbad[threadIdx.x / kWave] = wave_bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m2 = wmx[0];
    int b2 = wbad[0];
    for (int w = 1; w < kJbThreads / kWave; ++w) {
      m2 = fmax(m2, wmx[w]);
      b2 |= wbad[w];
    }
    atomicMax(&P.slot[0], (unsigned long long)__double_as_longlong(m2));
    if (b2) atomicOr(&P.flag[1], 1);
  }
}

template <typename T>
__global__ __launch_bounds__(kJbThreads) void bj_load_kernel(BjParams<T> P) {
  const double mx = __longlong_as_double((long long)P.slot[0]);
  const double sc = (P.flags & PSVD_SCALE) && mx > 0.0 && isfinite(mx) ? mx : 1.0;
  const int64_t total = (int64_t)P.m * P.N + (int64_t)P.N * P.N;
  for (int64_t e = (int64_t)blockIdx.x * kJbThreads + threadIdx.x; e < total; e += (int64_t)gridDim.x * kJbThreads) {
    if (e < (int64_t)P.m * P.N) {
      const int i = (int)(e / P.N), j = (int)(e - (int64_t)i * P.N);
      // T division, as the reference's x / scale
      P.W[e] = j < P.n ? (T)(P.a[(int64_t)i * P.n + j] / (T)sc) : T(0);
    } else {
      const int64_t f = e - (int64_t)P.m * P.N;
      const int i = (int)(f / P.N), j = (int)(f - (int64_t)i * P.N);
      P.V[f] = i == j ? T(1) : T(0);
    }
  }
}

// max_j |w_j|^2 of the loaded columns: the first round's floor (fmx[0], cleared before the launch)
template <typename T>
__global__ __launch_bounds__(kJbThreads) void bj_colmax_kernel(BjParams<T> P) {
  // 16 columns per workgroup, 16 row groups; thread (g, c) sums rows g, g + 16, ... with four
  // loads in flight, the row-group partials are added in group order, then the maximum
  constexpr int NC = 16, NG = kJbThreads / NC;
  __shared__ double part[NG][NC];
  const int g = threadIdx.x / NC, cl = threadIdx.x % NC;
  const int c = blockIdx.x * NC + cl;
  double acc = 0.0;
  if (c < P.n) {
    const T *col = P.W + c;
    int i = g;
    for (; i + 3 * NG < P.m; i += 4 * NG) {
      const double x0 = (double)col[(int64_t)i * P.N], x1 = (double)col[(int64_t)(i + NG) * P.N];
      const double x2 = (double)col[(int64_t)(i + 2 * NG) * P.N], x3 = (double)col[(int64_t)(i + 3 * NG) * P.N];
      acc = fma(x0, x0, acc);
      acc = fma(x1, x1, acc);
      acc = fma(x2, x2, acc);
      acc = fma(x3, x3, acc);
    }
    for (; i < P.m; i += NG) {
      const double x = (double)col[(int64_t)i * P.N];
      acc = fma(x, x, acc);
    }
  }
  part[g][cl] = acc;
  __syncthreads();
  if (threadIdx.x < kWave) {
    double tot = 0.0;
    if (threadIdx.x < NC)
      for (int x = 0; x < NG; ++x) tot += part[x][threadIdx.x];
    tot = wave_max_fast(tot);
    if (threadIdx.x == 0) atomicMax(&P.fmx[0], (unsigned long long)__double_as_longlong(tot));
  }
}

constexpr int kGramRows = 64;   // rows of W staged per step of the Gram partial

template <typename T, int KB>
__host__ __device__ inline size_t bj_inner_lds() {
  constexpr int KS = BjShape<KB>::KS, LD = BjShape<KB>::LD;
  return sizeof(double) * (4 * KB + 3 * (size_t)KS * LD) + sizeof(int) * 2 * KB;
}

template <typename T, int KB, int NTH>
__device__ __forceinline__ void bj_update_body(const BjParams<T> &P, int p, int t0, int r, int par, double *smem);

// The pair rotation of the sweeps (primitives.py:340-347) with the shortest dependent chain: for FP32
// data tau and t in FP32 (as the LDS kernel forms them: native rcp/sqrt, FP64 quotient outside the FP32
// range), c = 1/sqrt(1 + t^2) and s = t c in FP64 (rsq plus one Newton step); FP64 data takes the
// FP64 rotation.
template <typename T>
__device__ __forceinline__ void bj_rotation(double app, double aqq, double apq, double &c, double &s) {
  if constexpr (sizeof(T) == 4) {
    const double a2 = 2.0 * fabs(apq), dd = aqq - app;
    float tau;
    if (a2 > 1e-30 && a2 < 1e30 && fabs(dd) < 1e37)
      tau = (float)dd * __builtin_amdgcn_rcpf((float)a2);
    else
      tau = (float)(dd / a2);
    const float at = fabsf(tau);
    float t = at > 1e18f ? 0.5f * __builtin_amdgcn_rcpf(at)
                         : __builtin_amdgcn_rcpf(at + __builtin_amdgcn_sqrtf(fmaf(at, at, 1.0f)));
    t = copysignf(t, tau);
    const double td = (double)t, x = fma(td, td, 1.0);
    double y = __builtin_amdgcn_rsq(x);
    y = y * fma(-0.5 * x * y, y, 1.5);
    c = y;
    s = apq < 0.0 ? -td * y : td * y;
  } else {
    jacobi_rotation_fast<T>(app, aqq, apq, c, s);
  }
}

// One 2 x 2 block of S under a rotation round: columns by pair 2's (c2, s2), then rows by pair 1's
// (c1, s1), with the fused multiply-adds spelled out.
__device__ __forceinline__ void bj_rot2x2(double c1, double s1, double c2, double s2, double b11, double b12,
                                          double b21, double b22, double &n11, double &n12, double &n21,
                                          double &n22) {
  const double t11 = fma(c2, b11, -(s2 * b12)), t12 = fma(s2, b11, c2 * b12);
  const double t21 = fma(c2, b21, -(s2 * b22)), t22 = fma(s2, b21, c2 * b22);
  n11 = fma(c1, t11, -(s1 * t21));
  n21 = fma(s1, t11, c1 * t21);
  n12 = fma(c1, t12, -(s1 * t22));
  n22 = fma(s1, t12, c1 * t22);
}

// Sweeps of S_p (FP64, in LDS) with the one-sided test; Q_p = accumulated rotation, written to the
// round's parity slot.  One thread per upper-triangle 2 x 2 block of a rotation round (KB (KB + 1) / 2
// of them), so no thread carries a second block on the round's critical path.  Blocks [npair, ...) of
// the launch (16-wide blocks only) apply round g - 1's rotations (tournament round vr) to V, four
// 64-row tiles per workgroup, nvq workgroups per pair, beside the pair sweeps.
template <typename T, int KB>
__global__ __launch_bounds__(BjShape<KB>::INNER) void bj_inner_kernel(BjParams<T> P, int r, int g, int fslot, int vr,
                                                                      int nvq, int done_slot) {
  constexpr int KS = BjShape<KB>::KS, LD = BjShape<KB>::LD, NTI = BjShape<KB>::INNER;
  extern __shared__ __align__(16) double smem_bj[];
  double2 *cs = reinterpret_cast<double2 *>(smem_bj);   // (c, s) of pair k of the round
  double *S = reinterpret_cast<double *>(cs + KB), *Qt = S + KS * LD;
  int2 *pq = reinterpret_cast<int2 *>(Qt + KS * LD);   // (p, q) of pair k
  double2 *cs2 = reinterpret_cast<double2 *>(pq + KB);   // cross rounds: the other (c, s) buffer
  double *S2 = reinterpret_cast<double *>(cs2 + KB);     // cross rounds: the other S buffer
  const int p = blockIdx.x, tid = threadIdx.x;
  if (done_slot >= 0 && P.flag[done_slot] == 0) return;   // the previous sweep rotated nowhere
  const int npair = P.nb / 2;
  if constexpr (NTI % 256 == 0) {
    if (p >= npair) {
      const int b = p - npair;
      bj_update_body<T, KB, NTI>(P, b / nvq, (b % nvq) * (NTI / 256), vr, (g - 1) & 1, smem_bj);
      return;
    }
  }
  const double eps = Eps<T>::value, tol = P.tol > 0.0 ? P.tol : 10.0 * eps;
  // the reference's floor eps^2 max_j |w_j|^2 at the start of this round (primitives.py:337): the largest
  // diagonal the previous round's pair sweeps left (their column norms after its rotations), or for the
  // first round the norms of the loaded columns
  const double floor_ = eps * eps * __longlong_as_double((long long)P.fmx[g & 1]);
  for (int e = tid; e < KS * KS; e += NTI) {
    const int i = e / KS, j = e - i * KS;
    const double *sp = P.Sp + (size_t)p * P.splits * KS * KS + e;
    // the slices' partials added in slice order, eight loads in flight at a time
    double acc = 0.0;
    for (int z0 = 0; z0 < P.splits; z0 += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = z0 + u < P.splits ? sp[(size_t)(z0 + u) * KS * KS] : 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (z0 + u < P.splits) acc = z0 + u == 0 ? v[u] : acc + v[u];
    }
    S[i * LD + j] = acc;
    Qt[i * LD + j] = i == j ? 1.0 : 0.0;
  }
  __syncthreads();
  // a subproblem in which no pair passes the test rotates nowhere in its sweep: skip it
  int live = 0;
  for (int e = tid; e < KS * KS; e += NTI) {
    const int i = e / KS, j = e - i * KS;
    if (i < j) {
      const double app = S[i * LD + i], aqq = S[j * LD + j], apq = S[i * LD + j];
      live |= apq * apq > tol * tol * app * aqq && fmin(app, aqq) > floor_;
    }
  }
  live = __syncthreads_or(live);
  int any = 0;
  // this thread's items of the round, fixed for the launch: upper-triangle blocks (k1 <= k2) of
  // the KB x KB pair grid and the Qt row pairs k = tid / KS + j QSTEP at column tid % KS
  constexpr int NTRI = KB * (KB + 1) / 2, NIT = (NTRI + NTI - 1) / NTI;
  int it1[NIT], it2[NIT];
#pragma unroll
  for (int j = 0; j < NIT; ++j) {
    const int e = tid + j * NTI;
    it1[j] = it2[j] = -1;
    if (e < NTRI) tri_decode(e, KB, it1[j], it2[j]);
  }
  constexpr int QSTEP = NTI / KS, QIT = (KB + QSTEP - 1) / QSTEP;   // Qt row pairs per column
  static_assert(NTI % KS == 0, "the Qt items tile the columns exactly");
  // blocks and Qt entries on separate threads when the workgroup holds both (one item per thread)
  constexpr bool kSplitQt = NTI >= NTRI + KB * KS;
  static_assert(!kSplitQt || NIT == 1, "one block per thread");
  const int qcol = tid % KS, qk = tid / KS;
  // the sweep's first tournament round treats every pair of the subproblem, the others only the
  // pairs across its two blocks (the pairs inside a block were treated in the first round)
  const bool cross = r > 0 && !P.full_pairs;
  const int nrr = cross ? KB : KS - 1;
  // Qt threads start at a wave.  In the cross rounds pair k is (k, KB + (k + rr) mod KB), so
  // row k of Qt stays with pair k and row KB + j passes from pair k to pair k - 1 each round: the
  // thread of (pair k, column col) -- 16 consecutive lanes per column -- keeps Qt[k][col] and the
  // Qt row its pair meets in registers and hands the latter to the next lower lane of its row with one
  // DPP rotate; after a whole cross sweep (KB rounds) lane k holds row KB + k again.
  constexpr int QB = (NTRI + kWave - 1) / kWave * kWave;
  static_assert(!kSplitQt || (KB == 16 && NTI >= QB + KB * KS), "register Qt needs 16-lane rows");
  const bool qthr = kSplitQt && tid >= QB && tid < QB + KB * KS;
  const int qe = tid - QB, tk = cross ? (qe & (KB - 1)) : qe / KS, tc = cross ? (qe >> 4) : qe - (qe / KS) * KS;
  double qp = tk == tc ? 1.0 : 0.0, qq = KB + tk == tc ? 1.0 : 0.0;
  bool swept = false;   // the cross rounds below ran the sweeps
  if constexpr (kSplitQt) {
    if (cross) {
      // Cross rounds, one barrier per rotation round.  While the blocks update (from S into the other S
      // buffer, every entry written) and Qt turns in registers, the rotation threads form the next
      // round's rotations: pair k of round rr + 1 meets the q column of pair k + 1 of round rr, so its
      // three inputs are entries of round rr's blocks (k, k), (k + 1, k + 1) and (k, k + 1) (for
      // k = KB - 1 the mirror entry of block (0, KB - 1)), formed from the same loads by the same
      // arithmetic as the block threads form them.  Threads DB .. DB + KB - 2 (one wave) serve pairs
      // 0 .. KB - 2 and the first thread of the next wave pair KB - 1, whose entry is the mirror one.
      constexpr int DB = (QB + KB * KS + kWave - 1) / kWave * kWave;
      static_assert(NTI >= DB + kWave + 1, "the rotation threads fit in the workgroup");
      const int rk_ = tid == DB + kWave ? KB - 1 : (tid >= DB && tid < DB + KB - 1 ? tid - DB : -1);
      int act_sw = 0, act_next = 0;
      if (live && tid < KB) {   // round 0's rotations
        const int a = tid, b = KB + tid, ra = __mul24(a, LD), rb = __mul24(b, LD);
        const double app = S[ra + a], aqq = S[rb + b], apq = S[ra + b];
        const bool act = apq * apq > tol * tol * app * aqq && fmin(app, aqq) > floor_;
        double c = 1.0, sn = 0.0;
        if (act) bj_rotation<T>(app, aqq, apq, c, sn);
        cs[tid] = make_double2(c, act ? sn : 0.0);
        act_sw = act;
      }
      __syncthreads();
      for (int sw = 0; live && sw < P.inner_sweeps; ++sw) {
        if (sw > 0) {
          act_sw = act_next;
          act_next = 0;
        }
        for (int rr = 0; rr < KB; ++rr) {
          const double2 *csc = (rr & 1) ? cs2 : cs;
          double2 *csn = (rr & 1) ? cs : cs2;
          const double *Sc = (rr & 1) ? S2 : S;
          double *Sn = (rr & 1) ? S : S2;
          // the rotation threads set the round's length (about 1280 cycles against 680 for a block item
          // and 410 for a Qt item): the block and Qt threads start 128 and 192 cycles later, so the
          // rotation threads' LDS loads are served first (512x256: 8.8 -> 8.5 ms)
          if (tid < NTRI) {
            __builtin_amdgcn_s_sleep(2);
            const int k1 = it1[0], k2 = it2[0];
            const int p1 = k1, q1 = KB + ((k1 + rr) & (KB - 1)), p2 = k2, q2 = KB + ((k2 + rr) & (KB - 1));
            const int rp1 = __mul24(p1, LD), rq1 = __mul24(q1, LD), rp2 = __mul24(p2, LD), rq2 = __mul24(q2, LD);
            const double2 r1 = csc[k1], r2 = csc[k2];
            const double b11 = Sc[rp1 + p2], b12 = Sc[rp1 + q2], b21 = Sc[rq1 + p2], b22 = Sc[rq1 + q2];
            double n11, n12, n21, n22;
            bj_rot2x2(r1.x, r1.y, r2.x, r2.y, b11, b12, b21, b22, n11, n12, n21, n22);
            Sn[rp1 + p2] = n11;
            Sn[rp1 + q2] = n12;
            Sn[rq1 + p2] = n21;
            Sn[rq1 + q2] = n22;
            if (p1 != p2) {
              Sn[rp2 + p1] = n11;
              Sn[rq2 + p1] = n12;
              Sn[rp2 + q1] = n21;
              Sn[rq2 + q1] = n22;
            }
          } else if (qthr) {
            __builtin_amdgcn_s_sleep(3);
            const double2 rq = csc[tk];
            if (rq.y != 0.0) {
              const double x = qp, y = qq;
              qp = rq.x * x - rq.y * y;
              qq = rq.y * x + rq.x * y;
            }
            qq = dpp_double<0x12F>(qq);   // row_ror:15 -- lane i takes lane i + 1 of its row
          } else if (rk_ >= 0) {
            const int k = rk_, kn = (k + 1) & (KB - 1);
            const int pk = k, qk = KB + ((k + rr) & (KB - 1)), pn = kn, qn = KB + ((kn + rr) & (KB - 1));
            const int rpk = __mul24(pk, LD), rqk = __mul24(qk, LD), rpn = __mul24(pn, LD), rqn = __mul24(qn, LD);
            const double2 rk = csc[k], rn = csc[kn];
            const double a11 = Sc[rpk + pk], a12 = Sc[rpk + qk], a21 = Sc[rqk + pk], a22 = Sc[rqk + qk];
            const double b11 = Sc[rpn + pn], b12 = Sc[rpn + qn], b21 = Sc[rqn + pn], b22 = Sc[rqn + qn];
            double app, aqq, apq, x12, x21, x22;
            bj_rot2x2(rk.x, rk.y, rk.x, rk.y, a11, a12, a21, a22, app, x12, x21, x22);
            bj_rot2x2(rn.x, rn.y, rn.x, rn.y, b11, b12, b21, b22, x12, x21, x22, aqq);
            if (k < KB - 1) {   // block (k, k + 1): its n12
              const double c11 = Sc[rpk + pn], c12 = Sc[rpk + qn], c21 = Sc[rqk + pn], c22 = Sc[rqk + qn];
              bj_rot2x2(rk.x, rk.y, rn.x, rn.y, c11, c12, c21, c22, x22, apq, x12, x21);
            } else {            // block (0, KB - 1): the n21 its mirror stores
              const double c11 = Sc[rpn + pk], c12 = Sc[rpn + qk], c21 = Sc[rqn + pk], c22 = Sc[rqn + qk];
              bj_rot2x2(rn.x, rn.y, rk.x, rk.y, c11, c12, c21, c22, x22, x12, apq, x21);
            }
            // the rotation is formed whether or not the pair passes the test (an inactive pair's
            // possibly non-finite values are discarded by the select), so the test and the rotation's
            // chain overlap
            const bool act = apq * apq > tol * tol * app * aqq && fmin(app, aqq) > floor_;
            double c, sn;
            bj_rotation<T>(app, aqq, apq, c, sn);
            csn[k] = act ? make_double2(c, sn) : make_double2(1.0, 0.0);
            if (rr + 1 < KB)
              act_sw |= act;
            else
              act_next |= act;   // round 0 of the next sweep
          }
          __syncthreads();
        }
        act_sw = __syncthreads_or(act_sw);
        any |= act_sw;
        if (!act_sw) break;
      }
      swept = true;
    }
  }
  for (int sw = 0; !swept && live && sw < P.inner_sweeps; ++sw) {
    int act_sw = 0;
    for (int rr = 0; rr < nrr; ++rr) {
      if (tid < KB) {
        int a, b;
        sub_pair(cross, rr, tid, KB, a, b);
        const int ra = __mul24(a, LD), rb = __mul24(b, LD);   // 24-bit products: full-rate multiplies
        const double app = S[ra + a], aqq = S[rb + b], apq = S[ra + b];
        const bool act = apq * apq > tol * tol * app * aqq && fmin(app, aqq) > floor_;
        double c = 1.0, s = 0.0;
        if (act) bj_rotation<T>(app, aqq, apq, c, s);
        cs[tid] = make_double2(c, act ? s : 0.0);
        if constexpr (!kSplitQt) pq[tid] = make_int2(a, b);
        act_sw |= act;
      }
      __syncthreads();
      if constexpr (kSplitQt) {
        // threads [0, NTRI): one 2 x 2 block each; threads [NTRI, NTRI + KB KS): one Qt entry pair
        // each.  The pairs come from the closed form, so each thread's (c, s) and S or Qt loads are
        // issued together: one LDS round trip before the arithmetic.
        if (tid < NTRI) {
          int p1, q1, p2, q2;
          sub_pair(cross, rr, it1[0], KB, p1, q1);
          sub_pair(cross, rr, it2[0], KB, p2, q2);
          // every load of the item goes out before the activity test, so the item costs one LDS round trip
          const double2 r1 = cs[it1[0]], r2 = cs[it2[0]];
          const int rp1 = __mul24(p1, LD), rq1 = __mul24(q1, LD), rp2 = __mul24(p2, LD), rq2 = __mul24(q2, LD);
          const double b11 = S[rp1 + p2], b12 = S[rp1 + q2], b21 = S[rq1 + p2], b22 = S[rq1 + q2];
          asm volatile("" ::"v"(b11), "v"(b12), "v"(b21), "v"(b22));   // keeps the loads ahead of the test
          double n11, n12, n21, n22;
          bj_rot2x2(r1.x, r1.y, r2.x, r2.y, b11, b12, b21, b22, n11, n12, n21, n22);
          if (r1.y != 0.0 || r2.y != 0.0) {
            S[rp1 + p2] = n11;
            S[rp1 + q2] = n12;
            S[rq1 + p2] = n21;
            S[rq1 + q2] = n22;
            if (p1 != p2) {
              S[rp2 + p1] = n11;
              S[rq2 + p1] = n12;
              S[rp2 + q1] = n21;
              S[rq2 + q1] = n22;
            }
          }
        } else if (qthr) {
          const double2 rq = cs[tk];
          if (cross) {
            if (rq.y != 0.0) {
              const double x = qp, y = qq;
              qp = rq.x * x - rq.y * y;
              qq = rq.y * x + rq.x * y;
            }
            qq = dpp_double<0x12F>(qq);   // row_ror:15 -- lane i takes lane i + 1 of its row
          } else {
            int pp, q2;
            sub_pair(false, rr, tk, KB, pp, q2);
            const double x = Qt[pp * LD + tc], y = Qt[q2 * LD + tc];
            if (rq.y != 0.0) {
              Qt[pp * LD + tc] = rq.x * x - rq.y * y;
              Qt[q2 * LD + tc] = rq.y * x + rq.x * y;
            }
          }
        }
      } else {
#pragma unroll
        for (int j = 0; j < NIT; ++j) {
          if (it1[j] < 0) continue;
          const double2 r1 = cs[it1[j]], r2 = cs[it2[j]];
          if (r1.y == 0.0 && r2.y == 0.0) continue;
          const int2 a1 = pq[it1[j]], a2 = pq[it2[j]];
          block_update(S, LD, a1.x, a1.y, a2.x, a2.y, r1.x, r1.y, r2.x, r2.y, r1.y != 0.0, r2.y != 0.0);
        }
        rotate_rows_batched<double, KB, QSTEP, QIT>(Qt, LD, cs, pq, qk, qcol);
      }
      __syncthreads();
    }
    act_sw = __syncthreads_or(act_sw);
    any |= act_sw;
    if (!act_sw) break;
  }
  if (cross && qthr && any) {   // the register rows back into Qt
    Qt[tk * LD + tc] = qp;
    Qt[(KB + tk) * LD + tc] = qq;
  }
  if (kSplitQt && cross) __syncthreads();
  if (any)
    for (int e = tid; e < KS * KS; e += NTI) {
      const int i = e / KS, j = e - i * KS;
      P.Q[((size_t)(g & 1) * npair + p) * KS * KS + e] = (T)Qt[j * LD + i];   // Q = Qt^T
    }
  __syncthreads();
  if (tid < kWave) {   // the next round's floor: the largest column norm^2 after this round's rotations
    const double d = wave_max_fast(tid < KS ? S[tid * LD + tid] : 0.0);
    if (tid == 0) atomicMax(&P.fmx[(g + 1) & 1], (unsigned long long)__double_as_longlong(d));
  }
  if (tid == 0) {
    P.act[(g & 1) * npair + p] = any;
    if (any) atomicOr(&P.flag[fslot], 1);
  }
}

template <typename T> struct BjMfma;
template <> struct BjMfma<float> {
  using V = j4f;
  static __device__ __forceinline__ V op(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ __forceinline__ int row(int q, int lk) { return 4 * lk + q; }
};
template <> struct BjMfma<double> {
  using V = j4d;
  static __device__ __forceinline__ V op(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  static __device__ __forceinline__ int row(int q, int lk) { return 4 * q + lk; }
};

// V[:, I u J] <- V[:, I u J] Q_p for NTH / 256 consecutive 64-row tiles t0, t0 + 1, ... of V (pair p of
// tournament round r, rotation from parity slot par).
constexpr int kTileRows = 64;

template <typename T, int KB, int NTH>
__host__ __device__ inline size_t bj_update_lds() {
  return sizeof(T) * (size_t)(kTileRows * (NTH / 256) + BjShape<KB>::KS) * BjShape<KB>::LD;
}

// The pair-sweep launch: its V-tile workgroups (16-wide blocks) need more LDS than the sweeps.
template <typename T, int KB>
__host__ __device__ inline size_t bj_inner_launch_lds() {
  constexpr int NTI = BjShape<KB>::INNER;
  const size_t a = bj_inner_lds<T, KB>();
  if constexpr (NTI % 256 == 0) {
    // more than half of a CU's 160 KiB: one workgroup per CU, so no V tiles share a CU with a pair sweep
    const size_t b = bj_update_lds<T, KB, NTI>(), c = 81 * 1024;
    return a > b ? (a > c ? a : c) : (b > c ? b : c);
  }
  return a;
}

// Each 64 x KS product on four waves of its 256-thread group: wave w owns rows 32 (w / 2) .. + 32 and
// columns KB (w % 2) .. + KB.  Tiles past the last are zero rows that are not stored.
template <typename T, int KB, int NTH>
__device__ __forceinline__ void bj_update_body(const BjParams<T> &P, int p, int t0, int r, int par, double *smem) {
  constexpr int KS = BjShape<KB>::KS, LD = BjShape<KB>::LD, TJ = KB / 16;
  static_assert(NTH % 256 == 0, "whole 256-thread groups");
  const int npair = P.nb / 2;
  if (!P.act[par * npair + p]) return;   // uniform over the workgroup
  const int tid = threadIdx.x, qd = tid >> 8, qt = tid & 255, lane = tid & 63, wave = qt >> 6;
  T (*Qs)[LD] = reinterpret_cast<T (*)[LD]>(smem);
  T (*Xs)[LD] = Qs + KS + qd * kTileRows;
  const int lr = lane & 15, lk = lane >> 4, wm = (wave >> 1) * 32, wn = (wave & 1) * KB;
  const int r0 = (t0 + qd) * kTileRows;
  int I, J;
  rr_pair(r, p, P.nb, I, J);
  using V16 = Vec16<T>;
  constexpr int VE = V16::n, RV = KS / VE;   // 16-byte vectors: a row's two KB-column blocks
  const V16 *Qp = reinterpret_cast<const V16 *>(P.Q + ((size_t)par * npair + p) * KS * KS);
  for (int e = qt; e < kTileRows * RV; e += 256) {
    const int i = e / RV, j = (e - i * RV) * VE, row = r0 + i;
    V16 x;
    if (row < P.N) {
      x = *reinterpret_cast<const V16 *>(P.V + (int64_t)row * P.N + bj_col<KB>(I, J, j));
    } else {
#pragma unroll
      for (int u = 0; u < VE; ++u) x.v[u] = T(0);
    }
#pragma unroll
    for (int u = 0; u < VE; ++u) Xs[i][j + u] = x.v[u];
  }
  for (int e = tid; e < KS * RV; e += NTH) {
    const int i = e / RV, j = (e - i * RV) * VE;
    const V16 y = Qp[e];
#pragma unroll
    for (int u = 0; u < VE; ++u) Qs[i][j + u] = y.v[u];
  }
  __syncthreads();
  typename BjMfma<T>::V acc[2][TJ];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[i][j][q] = 0;
#pragma unroll 4
  for (int k = 0; k < KS; k += 4) {
    T a[2], b[TJ];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = Xs[wm + 16 * i + lr][k + lk];
#pragma unroll
    for (int j = 0; j < TJ; ++j) b[j] = Qs[k + lk][wn + 16 * j + lr];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j) acc[i][j] = BjMfma<T>::op(a[i], b[j], acc[i][j]);
  }
  __syncthreads();   // every wave has read Xs
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) Xs[wm + 16 * i + BjMfma<T>::row(q, lk)][wn + 16 * j + lr] = (T)acc[i][j][q];
  __syncthreads();
  for (int e = qt; e < kTileRows * RV; e += 256) {
    const int i = e / RV, j = (e - i * RV) * VE, row = r0 + i;
    if (row < P.N) {
      V16 x;
#pragma unroll
      for (int u = 0; u < VE; ++u) x.v[u] = Xs[i][j + u];
      *reinterpret_cast<V16 *>(P.V + (int64_t)row * P.N + bj_col<KB>(I, J, j)) = x;
    }
  }
}

// Row slice z of the two blocks of pair p of round r, as round rprev left them: each block X is rebuilt
// from the row slice of its round-rprev pair (old buffer Wo) times that pair's rotation -- the half of Q
// that makes X (the identity when the pair was inactive) -- on MFMA, written to the new buffer Wn and
// kept in LDS in FP64, where the Gram partial of round r is formed from it (rprev < 0: no update, the
// rows come straight from Wo).  Every block of round r is made by exactly one workgroup per slice, and
// the old buffer is only read, so no workgroup reads a row another one writes.
template <typename T, int KB>
__device__ __forceinline__ void bj_gram_update_body(const BjParams<T> &P, int p, int z, int r, int rprev, int qpar,
                                                    const T *Wo, T *Wn, bool gram, double *smem) {
  constexpr int KR = kGramRows, KS = BjShape<KB>::KS, LD = BjShape<KB>::LD, TI = KB / 16;
  double (*Ws)[LD] = reinterpret_cast<double (*)[LD]>(smem);               // KR x KS, FP64 new rows
  T (*Xs)[KR][LD] = reinterpret_cast<T (*)[KR][LD]>(Ws + KR);              // [2][KR][KS] old pair rows per half
  T (*Qs)[KS][KB + 1] = reinterpret_cast<T (*)[KS][KB + 1]>(Xs + 2);       // [2][KS][KB] rotation halves
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  int I, J;
  rr_pair(r, p, P.nb, I, J);
  // per half h (block X = I, J): its round-rprev pair (blocks A < B) and whether X is A
  int pa[2], pb[2], pside[2], pact[2], pp[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int X = h == 0 ? I : J;
    if (rprev >= 0) {
      // seat of player X in round rprev (closed form of rr_pair), then its pair
      const int nm1 = P.nb - 1;
      int seat = 0;
      if (X > 0) {
        seat = X - 1 + rprev;
        if (seat >= nm1) seat -= nm1;
        seat += 1;
      }
      pp[h] = seat < P.nb - 1 - seat ? seat : P.nb - 1 - seat;
      rr_pair(rprev, pp[h], P.nb, pa[h], pb[h]);
      pside[h] = X == pa[h] ? 0 : 1;
      pact[h] = P.act[qpar * (P.nb / 2) + pp[h]];
    } else {
      pp[h] = 0;
      pa[h] = pb[h] = X;
      pside[h] = 0;
      pact[h] = 0;
    }
  }
  const int rows_per = (P.m + P.splits - 1) / P.splits;
  const int r0 = z * rows_per, r1 = r0 + rows_per < P.m ? r0 + rows_per : P.m;
  using V16 = Vec16<T>;
  constexpr int VE = V16::n, RV = KS / VE, NPF = 2 * KR * RV / kJbThreads;
  static_assert(2 * KR * RV % kJbThreads == 0, "the staged rows tile the workgroup");
  // the old rows of a chunk (Xs[h] = [W_A W_B] of half h's previous pair), fetched into registers one
  // chunk ahead so their latency overlaps the previous chunk's products
  V16 pf[NPF];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int j = 0; j < NPF; ++j) {
      const int e = tid + j * kJbThreads, h = e / (KR * RV), f = e - h * KR * RV, rr = f / RV, c = (f - rr * RV) * VE;
      const int row = k0 + rr;
      // read whether or not the previous pair rotated (its rotation half is then the identity): no load
      // waits for the activity flags
      if (row < r1) {
        pf[j] = *reinterpret_cast<const V16 *>(Wo + (int64_t)row * P.N + (c < KB ? pa[h] * KB + c : pb[h] * KB + c - KB));
      } else {
#pragma unroll
        for (int u = 0; u < VE; ++u) pf[j].v[u] = T(0);
      }
    }
  };
  fetch(r0);
  // the rotation halves (identity for an inactive or absent previous pair): every thread's vectors are
  // loaded unconditionally and together (an inactive pair's slot holds an older round's rotation, which
  // the select discards), then stored
  {
    constexpr int QV = KB / VE, NQ = 2 * KS * QV, NQT = (NQ + kJbThreads - 1) / kJbThreads;
    V16 qv[NQT];
#pragma unroll
    for (int j = 0; j < NQT; ++j) {
      const int e = tid + j * kJbThreads, h = e / (KS * QV), f = e - h * KS * QV, i = f / QV, jv = (f - i * QV) * VE;
      if (e < NQ && rprev >= 0)
        qv[j] = *reinterpret_cast<const V16 *>(P.Q + ((size_t)qpar * (P.nb / 2) + pp[h]) * KS * KS + (size_t)i * KS +
                                               pside[h] * KB + jv);
    }
#pragma unroll
    for (int j = 0; j < NQT; ++j) {
      const int e = tid + j * kJbThreads, h = e / (KS * QV), f = e - h * KS * QV, i = f / QV, jv = (f - i * QV) * VE;
      if (e < NQ) {
#pragma unroll
        for (int u = 0; u < VE; ++u)
          Qs[h][i][jv + u] = pact[h] ? qv[j].v[u] : (i == pside[h] * KB + jv + u ? T(1) : T(0));
      }
    }
  }
  j4d acc[TI][TI];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TI; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[i][j][q] = 0.0;
  for (int k0 = r0; k0 < r1; k0 += KR) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NPF; ++j) {
      const int e = tid + j * kJbThreads, h = e / (KR * RV), f = e - h * KR * RV, rr = f / RV, c = (f - rr * RV) * VE;
#pragma unroll
      for (int u = 0; u < VE; ++u) Xs[h][rr][c + u] = pf[j].v[u];
    }
    if (k0 + KR < r1) fetch(k0 + KR);
    __syncthreads();
    // new rows: half h = Xs[h] (KR x KS) times Qs[h] (KS x KB); wave w: rows 16 (w + 4 t) .. of both halves
    {
      typename BjMfma<T>::V nacc[2][TI];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int j = 0; j < TI; ++j)
#pragma unroll
          for (int q = 0; q < 4; ++q) nacc[h][j][q] = 0;
      const int rt = 16 * wave;   // each wave one 16-row tile of the 64-row chunk
#pragma unroll 4
      for (int k = 0; k < KS; k += 4) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const T a = Xs[h][rt + lr][k + lk];
#pragma unroll
          for (int j = 0; j < TI; ++j) nacc[h][j] = BjMfma<T>::op(a, Qs[h][k + lk][16 * j + lr], nacc[h][j]);
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int j = 0; j < TI; ++j)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int rr = rt + BjMfma<T>::row(q, lk), row = k0 + rr, c = 16 * j + lr;
            const T v = (T)nacc[h][j][q];
            Ws[rr][h * KB + c] = row < r1 ? (double)v : 0.0;
            if (row < r1 && rprev >= 0) Wn[(int64_t)row * P.N + (h == 0 ? I : J) * KB + c] = v;
          }
    }
    if (!gram) continue;   // the last round's update alone (the next loads wait at the loop's barrier)
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < KR; k += 4) {
      double a[TI], b[TI];
#pragma unroll
      for (int i = 0; i < TI; ++i) a[i] = Ws[k + lk][(wave >> 1) * KB + 16 * i + lr];
#pragma unroll
      for (int j = 0; j < TI; ++j) b[j] = Ws[k + lk][(wave & 1) * KB + 16 * j + lr];
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TI; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
  if (!gram) return;
  const int wm = (wave >> 1) * KB, wn = (wave & 1) * KB;
  double *out = P.Sp + ((size_t)p * P.splits + z) * KS * KS;
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TI; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q)   // FP64 MFMA D layout: register q of lane l is row 4 q + l / 16
        out[(wm + 16 * i + 4 * q + lk) * KS + wn + 16 * j + lr] = acc[i][j][q];
}

template <typename T, int KB>
__host__ __device__ inline size_t bj_fused_lds() {
  constexpr int KR = kGramRows, KS = BjShape<KB>::KS, LD = BjShape<KB>::LD;
  return sizeof(double) * KR * LD + sizeof(T) * (2 * (size_t)KR * LD + 2 * (size_t)KS * (KB + 1));
}

// One launch per round g (tournament round r): workgroups [0, ngram) rebuild their row slice of the two
// blocks of pair b % npair from round rprev's rotations (parity slot qpar; old buffer Wo -> new buffer
// Wn) and, with gram, form that slice's Gram partial of round r; the rest apply round vr's rotations
// (parity slot vpar) to the V tiles (pair, tile) -- V is read by no kernel of the sweep.  Workgroup 0
// clears the floor slot the round's pair sweeps will publish into (clear >= 0) and the rotation flag
// slot clr_flag (>= 0: the next sweep's, in a sweep's first round).  Nothing is done when the previous
// sweep's flag (done_slot >= 0) says it rotated nowhere.
template <typename T, int KB>
__global__ __launch_bounds__(kJbThreads) void bj_fused_kernel(BjParams<T> P, int r, int rprev, int qpar, const T *Wo,
                                                             T *Wn, int ngram, int gram, int vr, int vpar, int nt,
                                                             int clear, int done_slot, int clr_flag) {
  extern __shared__ double smem_bf[];
  const int b = (int)blockIdx.x, npair = P.nb / 2;
  if (done_slot >= 0 && P.flag[done_slot] == 0) return;
  if (b == 0 && threadIdx.x == 0) {
    if (clear >= 0) P.fmx[clear] = 0ull;
    if (clr_flag >= 0) P.flag[clr_flag] = 0;
  }
  if (b < ngram) {
    bj_gram_update_body<T, KB>(P, b % npair, b / npair, r, rprev, qpar, Wo, Wn, gram != 0, smem_bf);
  } else {
    const int t = b - ngram;
    bj_update_body<T, KB, kJbThreads>(P, t / nt, t % nt, vr, vpar, smem_bf);
  }
}

template <typename T, int KB>
__host__ __device__ inline size_t bj_fused_launch_lds() {
  const size_t a = bj_fused_lds<T, KB>(), b = bj_update_lds<T, KB, kJbThreads>();
  return a > b ? a : b;
}

template <typename T>
__global__ __launch_bounds__(kJbThreads) void bj_sigma_kernel(BjParams<T> P) {
  // column norms in FP64: 16 columns per workgroup, 16 row groups added in group order
  constexpr int NC = 16, NG = kJbThreads / NC;
  __shared__ double part[NG][NC];
  const int g = threadIdx.x / NC, cl = threadIdx.x % NC;
  const int c = blockIdx.x * NC + cl;
  double acc = 0.0;
  if (c < P.n) {
    const T *col = P.W + c;
    int i = g;
    for (; i + 3 * NG < P.m; i += 4 * NG) {
      const double x0 = (double)col[(int64_t)i * P.N], x1 = (double)col[(int64_t)(i + NG) * P.N];
      const double x2 = (double)col[(int64_t)(i + 2 * NG) * P.N], x3 = (double)col[(int64_t)(i + 3 * NG) * P.N];
      acc = fma(x0, x0, acc);
      acc = fma(x1, x1, acc);
      acc = fma(x2, x2, acc);
      acc = fma(x3, x3, acc);
    }
    for (; i < P.m; i += NG) {
      const double x = (double)col[(int64_t)i * P.N];
      acc = fma(x, x, acc);
    }
  }
  part[g][cl] = acc;
  __syncthreads();
  if (threadIdx.x < NC && c < P.n) {
    double tot = 0.0;
    for (int x = 0; x < NG; ++x) tot += part[x][threadIdx.x];
    P.sig[c] = sqrt(tot);
  }
}

template <typename T>
__global__ __launch_bounds__(kJbThreads) void bj_order_kernel(BjParams<T> P) {
  // rank of sigma_j among all n (ties by index), the sigma read in LDS chunks of kChunk
  constexpr int kChunk = 2048;
  __shared__ double sc[kChunk];
  const int j = blockIdx.x * kJbThreads + threadIdx.x;
  const double sj = j < P.n ? P.sig[j] : 0.0;
  int rank = 0;
  for (int k0 = 0; k0 < P.n; k0 += kChunk) {
    const int len = P.n - k0 < kChunk ? P.n - k0 : kChunk;
    __syncthreads();
    for (int e = threadIdx.x; e < len; e += kJbThreads) sc[e] = P.sig[k0 + e];
    __syncthreads();
    for (int k = 0; k < len; ++k) {
      const double sk = sc[k];
      rank += (sk > sj) || (sk == sj && k0 + k < j);
    }
  }
  if (j < P.n) P.ord[rank] = j;
}

// s, U (kept columns W / sigma, others zero for the completion), Vh, nk, info
template <typename T>
__global__ __launch_bounds__(kJbThreads) void bj_out_kernel(BjParams<T> P, int sweeps, int converged) {
  const int n = P.n, m = P.m, np = n + (n & 1);
  const double eps = Eps<T>::value;
  const double s1 = P.sig[P.ord[0]];
  const double cut = (P.flags & PSVD_CUT_FLOOR) ? eps * s1 : eps * (double)(m > np ? m : np) * s1;
  int lo = 0, hi = n;   // sorted descending: the first column at or below the cut
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (P.sig[P.ord[mid]] > cut) lo = mid + 1; else hi = mid;
  }
  const int nk = lo;
  const double mx = __longlong_as_double((long long)P.slot[0]);
  const double sc = (P.flags & PSVD_SCALE) && mx > 0.0 && isfinite(mx) ? mx : 1.0;
  const bool vectors = !(P.flags & PSVD_NO_VECTORS);
  const int64_t gtid = (int64_t)blockIdx.x * kJbThreads + threadIdx.x, gsz = (int64_t)gridDim.x * kJbThreads;
  for (int64_t r = gtid; r < n; r += gsz) P.s[r] = r < nk ? (T)(P.sig[P.ord[r]] * sc) : T(0);
  if (vectors) {
    for (int64_t e = gtid; e < (int64_t)m * n; e += gsz) {
      const int i = (int)(e / n), r = (int)(e - (int64_t)i * n);
      const int c = P.ord[r];
      P.u[e] = r < nk ? (T)((double)P.W[(int64_t)i * P.N + c] / P.sig[c]) : T(0);
    }
    for (int64_t e = gtid; e < (int64_t)n * n; e += gsz) {
      const int r = (int)(e / n), j = (int)(e - (int64_t)r * n);
      P.vh[e] = P.V[(int64_t)j * P.N + P.ord[r]];
    }
  }
  if (gtid == 0) {
    *P.nk = vectors ? nk : n;
    P.info[0] = sweeps;
    P.info[1] = P.flag[1] ? PSVD_INFO_NONFINITE : (converged ? PSVD_INFO_OK : PSVD_INFO_NOCONV);
  }
}

__global__ void bj_status_kernel(int32_t *info, const int *cstatus) {
  if (threadIdx.x == 0 && info[1] == PSVD_INFO_OK && *cstatus != 0) info[1] = PSVD_INFO_COMPLETION;
}

// the solve's start: every slot and flag cleared
__global__ void bj_reset_kernel(unsigned long long *slot, int *flag, unsigned long long *fmx) {
  if (threadIdx.x == 0) {
    slot[0] = slot[1] = 0ull;
    fmx[0] = fmx[1] = 0ull;
    for (int i = 0; i < 5; ++i) flag[i] = 0;
  }
}

// Row slices per pair: about 512 workgroups per round (fewer partials for the pair sweeps to add than
// with 1024, more overlap of chunks than with 256), each over at least one 64-row chunk.
static int bj_splits(int m, int npair) {
  int s = (512 + npair - 1) / npair;
  const int smax = (m + kGramRows - 1) / kGramRows;
  if (s > smax) s = smax;
  if (s > 64) s = 64;
  return s < 1 ? 1 : s;
}

template <typename T, int KB>
size_t bj_ws_bytes_kb(int m, int n) {
  constexpr int KS = BjShape<KB>::KS;
  const int N = (n + KS - 1) / KS * KS, npair = N / KS;
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  return al(sizeof(T) * (size_t)m * N) * 2 + al(sizeof(T) * (size_t)N * N) + al(sizeof(T) * 2 * (size_t)npair * KS * KS) +
         al(sizeof(double) * (size_t)npair * bj_splits(m, npair) * KS * KS) + al(sizeof(int) * 2 * npair) + al(8 * 2) +
         al(4 * 5) + al(8 * 2) + al(sizeof(double) * N) + al(sizeof(int) * N) + al(4) +
         al(psvd_complete_workspace_bytes(1, m, n));
}

// Either block width fits in the same workspace.
template <typename T>
size_t bj_ws_bytes(int m, int n) {
  const size_t a = bj_ws_bytes_kb<T, 16>(m, n), b = bj_ws_bytes_kb<T, 32>(m, n);
  return a > b ? a : b;
}

// Pinned flag words for the asynchronous per-sweep reads: one pair per host thread, allocated on its
// first call and kept for the process (null if the allocation fails).
static int *bj_pinned_words() {
  thread_local int *pinned = nullptr;
  if (!pinned && hipHostMalloc((void **)&pinned, 2 * sizeof(int), hipHostMallocDefault) != hipSuccess) pinned = nullptr;
  return pinned;
}

template <typename T, int KB>
static int run_bj(int m, int n, const T *a, T *u, T *s, T *vh, int32_t *info, double tol, int max_sweeps, int flags,
                  int inner_sweeps, void *ws, size_t ws_bytes, hipStream_t st) {
  constexpr int KS = BjShape<KB>::KS;
  if (ws_bytes < bj_ws_bytes_kb<T, KB>(m, n)) return PSVD_EINVAL;
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute((const void *)bj_inner_kernel<T, KB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)bj_inner_launch_lds<T, KB>()) != hipSuccess ||
        hipFuncSetAttribute((const void *)bj_fused_kernel<T, KB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)bj_fused_launch_lds<T, KB>()) != hipSuccess)
      return PSVD_ELAUNCH;
    attr = true;
  }
  BjParams<T> P{};
  P.a = a;
  P.u = u;
  P.s = s;
  P.vh = vh;
  P.info = info;
  P.m = m;
  P.n = n;
  P.N = (n + KS - 1) / KS * KS;
  P.nb = P.N / KB;
  const int npair = P.nb / 2;
  P.splits = bj_splits(m, npair);
  P.inner_sweeps = inner_sweeps;
  P.full_pairs = (flags & PSVD_BLOCK_FULL_PAIRS) != 0;
  P.flags = flags & ~PSVD_BLOCK_FULL_PAIRS;
  P.tol = tol;
  char *cur = static_cast<char *>(ws);
  auto take = [&](size_t bytes) {
    char *q = cur;
    cur += (bytes + 255) & ~size_t(255);
    return q;
  };
  P.W = reinterpret_cast<T *>(take(sizeof(T) * (size_t)m * P.N));
  P.W2 = reinterpret_cast<T *>(take(sizeof(T) * (size_t)m * P.N));
  P.V = reinterpret_cast<T *>(take(sizeof(T) * (size_t)P.N * P.N));
  P.Q = reinterpret_cast<T *>(take(sizeof(T) * 2 * (size_t)npair * KS * KS));
  P.Sp = reinterpret_cast<double *>(take(sizeof(double) * (size_t)npair * P.splits * KS * KS));
  P.act = reinterpret_cast<int *>(take(sizeof(int) * 2 * npair));
  P.slot = reinterpret_cast<unsigned long long *>(take(8 * 2));
  P.flag = reinterpret_cast<int *>(take(4 * 5));
  P.fmx = reinterpret_cast<unsigned long long *>(take(8 * 2));
  P.sig = reinterpret_cast<double *>(take(sizeof(double) * P.N));
  P.ord = reinterpret_cast<int *>(take(sizeof(int) * P.N));
  P.nk = reinterpret_cast<int32_t *>(take(4));
  void *cws = take(psvd_complete_workspace_bytes(1, m, n));
  const int nt = (flags & PSVD_NO_VECTORS) ? 0 : (P.N + kTileRows - 1) / kTileRows;   // 64-row tiles of V
  auto blocks = [](int64_t items) {
    const int64_t b = (items + kJbThreads - 1) / kJbThreads;
    return (unsigned)(b < 4096 ? (b > 0 ? b : 1) : 4096);
  };
  hipLaunchKernelGGL(bj_reset_kernel, dim3(1), dim3(64), 0, st, P.slot, P.flag, P.fmx);
  hipLaunchKernelGGL(bj_absmax_kernel<T>, dim3(blocks((int64_t)m * n) < 256 ? blocks((int64_t)m * n) : 256u),
                     dim3(kJbThreads), 0, st, P);
  hipLaunchKernelGGL(bj_load_kernel<T>, dim3(blocks((int64_t)m * P.N + (int64_t)P.N * P.N)), dim3(kJbThreads), 0, st,
                     P);
  // the first round's floor from the loaded columns; later rounds take it from the pair sweeps before them
  hipLaunchKernelGGL(bj_colmax_kernel<T>, dim3((n + 15) / 16), dim3(kJbThreads), 0, st, P);
  const int ngram = npair * P.splits, nvt = npair * nt;
  // V tiles: with 16-wide blocks, four per workgroup of the pair-sweep launch (nvq per pair); else in the
  // fused launch
  constexpr bool kVInner = BjShape<KB>::INNER % 256 == 0;
  constexpr int kTq = BjShape<KB>::INNER / 256;
  const int nvq = kVInner ? (nt + kTq - 1) / kTq : 0;
  T *buf[2] = {P.W, P.W2};   // round g reads buf[(g - 1) & 1] (round g - 1's input) and writes buf[g & 1]
  const size_t flds = bj_fused_launch_lds<T, KB>(), ilds = bj_inner_launch_lds<T, KB>();
  // Sweep sw + 1 is queued before the host reads whether sweep sw rotated, so the GPU does not idle
  // through the host's wait; when sweep sw rotated nowhere, sweep sw + 1's kernels find its flag clear
  // and return, and the solve continues from the state sweep sw left (its round count g).  The flags
  // rotate over three slots: sweep sw's first launch clears sweep sw + 1's slot, which sweep sw - 2
  // used and whose read by the host is queued before.
  int g = 0, vr = -1;
  auto enqueue_sweep = [&](int sw, int *host_flag, hipEvent_t ev) {
    const int done_slot = sw > 1 ? bj_flag_slot(sw - 1) : -1;
    for (int r = 0; r < P.nb - 1; ++r, ++g) {
      // round g - 1's rotations applied to W and this round's Gram partials, beside round g - 1's V tiles
      const int rprev = g == 0 ? -1 : (r > 0 ? r - 1 : P.nb - 2);
      const T *wo = g == 0 ? buf[0] : buf[(g - 1) & 1];
      const int par = (g - 1) & 1;   // round g - 1's rotations and activity flags
      hipLaunchKernelGGL((bj_fused_kernel<T, KB>), dim3((unsigned)(ngram + (!kVInner && vr >= 0 ? nvt : 0))),
                         dim3(kJbThreads), flds, st, P, r, rprev, par, wo, buf[g & 1], ngram, 1, vr, par, nt, (g + 1) & 1,
                         done_slot, r == 0 ? bj_flag_slot(sw + 1) : -1);
      hipLaunchKernelGGL((bj_inner_kernel<T, KB>), dim3((unsigned)(npair + (kVInner && vr >= 0 ? npair * nvq : 0))),
                         dim3(BjShape<KB>::INNER), ilds, st, P, r, g, bj_flag_slot(sw), vr, nvq, done_slot);
      if (nt > 0) vr = r;
    }
    if (hipMemcpyAsync(host_flag, &P.flag[bj_flag_slot(sw)], sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipEventRecord(ev, st) != hipSuccess)
      return false;
    return hipGetLastError() == hipSuccess;
  };
  int *pinned = bj_pinned_words();
  if (!pinned) return PSVD_ELAUNCH;
  // the two sweep events; on leaving, the host waits for the last queued flag read (the pinned words are
  // reused by this thread's next solve) and releases them
  struct SweepEvents {
    hipEvent_t e[2] = {nullptr, nullptr};
    int wait = -1;
    ~SweepEvents() {
      if (wait >= 0) (void)hipEventSynchronize(e[wait]);
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } sev;
  hipEvent_t *ev = sev.e;
  if (hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) != hipSuccess)
    return PSVD_ELAUNCH;
  int sweeps = 0, converged = 0, rc = PSVD_OK;
  int g_end[2] = {0, 0}, vr_end[2] = {-1, -1};   // round count and V-tile round after sweep sw, by sw & 1
  if (max_sweeps >= 1) {
    sev.wait = 1;
    if (!enqueue_sweep(1, &pinned[1], ev[1])) rc = PSVD_ELAUNCH;
    g_end[1] = g;
    vr_end[1] = vr;
  }
  for (int sw = 1; rc == PSVD_OK && sw <= max_sweeps; ++sw) {
    if (sw < max_sweeps) {
      sev.wait = (sw + 1) & 1;
      if (!enqueue_sweep(sw + 1, &pinned[(sw + 1) & 1], ev[(sw + 1) & 1])) {
        rc = PSVD_ELAUNCH;
        break;
      }
      g_end[(sw + 1) & 1] = g;
      vr_end[(sw + 1) & 1] = vr;
    }
    if (hipEventSynchronize(ev[sw & 1]) != hipSuccess) {
      rc = PSVD_ELAUNCH;
      break;
    }
    sweeps = sw;
    if (!pinned[sw & 1]) {
      converged = 1;
      g = g_end[sw & 1];
      vr = vr_end[sw & 1];
      break;
    }
  }
  if (rc != PSVD_OK) return rc;
  // the last round's rotations on W (no Gram) and V; the tail reads W from the buffer they go to
  if (g > 0) {
    const int par = (g - 1) & 1;
    hipLaunchKernelGGL((bj_fused_kernel<T, KB>), dim3((unsigned)(ngram + (vr >= 0 ? nvt : 0))), dim3(kJbThreads), flds,
                       st, P, 0, P.nb - 2, par, buf[par], buf[g & 1], ngram, 0, vr, par, nt, -1, -1, -1);
    P.W = buf[g & 1];
  }
  hipLaunchKernelGGL(bj_sigma_kernel<T>, dim3((n + 15) / 16), dim3(kJbThreads), 0, st, P);
  hipLaunchKernelGGL(bj_order_kernel<T>, dim3((n + kJbThreads - 1) / kJbThreads), dim3(kJbThreads), 0, st, P);
  hipLaunchKernelGGL(bj_out_kernel<T>, dim3(blocks((int64_t)m * n)), dim3(kJbThreads), 0, st, P, sweeps, converged);
  if (hipGetLastError() != hipSuccess) return PSVD_ELAUNCH;
  if (!(flags & PSVD_NO_VECTORS)) {
    const int rc = psvd_complete_columns(sizeof(T) == 4 ? PSVD_F32 : PSVD_F64, 1, m, n, u, n, (int64_t)m * n, P.nk,
                                         &P.flag[2], cws, psvd_complete_workspace_bytes(1, m, n), st);
    if (rc != PSVD_OK) return rc;
    hipLaunchKernelGGL(bj_status_kernel, dim3(1), dim3(64), 0, st, info, &P.flag[2]);
  }
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

// Block width: PSVD_BLOCK_NARROW / PSVD_BLOCK_WIDE choose it; by default blocks of 16.
template <typename T>
static int run_bj_any(int m, int n, const T *a, T *u, T *s, T *vh, int32_t *info, double tol, int max_sweeps,
                      int flags, int inner_sweeps, void *ws, size_t ws_bytes, hipStream_t st) {
  const int f = flags & ~(PSVD_BLOCK_NARROW | PSVD_BLOCK_WIDE);
  if (flags & PSVD_BLOCK_WIDE)
    return run_bj<T, 32>(m, n, a, u, s, vh, info, tol, max_sweeps, f, inner_sweeps, ws, ws_bytes, st);
  return run_bj<T, 16>(m, n, a, u, s, vh, info, tol, max_sweeps, f, inner_sweeps, ws, ws_bytes, st);
}

}  // namespace psvd

extern "C" {

size_t psvd_jacobi_blocked_workspace_bytes(int dtype, int m, int n) {
  if (m < 1 || n < 1 || n > m) return 0;
  if (dtype == PSVD_F32) return psvd::bj_ws_bytes<float>(m, n);
  if (dtype == PSVD_F64) return psvd::bj_ws_bytes<double>(m, n);
  return 0;
}

int psvd_jacobi_blocked(int dtype, int m, int n, const void *a, void *u, void *s, void *vh, int32_t *info, double tol,
                        int max_sweeps, int flags, int inner_sweeps, void *workspace, size_t workspace_bytes,
                        psvd_stream_t stream) {
  if (m < 1 || n < 1 || n > m || !a || !s || !info || !workspace || max_sweeps < 1 || inner_sweeps < 1 ||
      (flags & ~(PSVD_SCALE | PSVD_CUT_FLOOR | PSVD_NO_VECTORS | PSVD_BLOCK_NARROW | PSVD_BLOCK_WIDE |
                 PSVD_BLOCK_FULL_PAIRS)) ||
      ((flags & PSVD_BLOCK_NARROW) && (flags & PSVD_BLOCK_WIDE)) || n > 65536)
    return PSVD_EINVAL;
  if (!(flags & PSVD_NO_VECTORS) && (!u || !vh)) return PSVD_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == PSVD_F32)
    return psvd::run_bj_any<float>(m, n, static_cast<const float *>(a), static_cast<float *>(u), static_cast<float *>(s),
                               static_cast<float *>(vh), info, tol, max_sweeps, flags, inner_sweeps, workspace,
                               workspace_bytes, st);
  if (dtype == PSVD_F64)
    return psvd::run_bj_any<double>(m, n, static_cast<const double *>(a), static_cast<double *>(u),
                                static_cast<double *>(s), static_cast<double *>(vh), info, tol, max_sweeps, flags,
                                inner_sweeps, workspace, workspace_bytes, st);
  return PSVD_EINVAL;
}

}  // extern "C"
