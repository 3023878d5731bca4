// K1 for complex matrices: batched one-sided Jacobi SVD of complex64 / complex128 matrices, each
// resident in LDS (svdbench/standalone/primitives.py:316-367 with the complex arithmetic of
// :331-351; completion :430-455; max-abs scaling solvers.py:294-299).
//
// Same organisation as the real kernel: a persistent launch pulls matrix indices from a device
// work queue, one workgroup holds W (m x n) and V (n x n) in LDS for the whole solve, and each
// column pair of a tournament round is owned by a group of L lanes striding over the rows.  W
// and V are kept as separate real and imaginary planes (column-major, odd leading dimension), so
// every access is a plain R load or store.  Per pair the group forms in FP64
//     a_pp = sum |w_p|^2,  a_qq = sum |w_q|^2,  a_pq = w_p^H w_q  (real and imaginary part),
// tests |a_pq|^2 > tol^2 a_pp a_qq and min(a_pp, a_qq) > eps^2 max_j |w_j|^2 (:335-338), and
// rotates with tau = (a_qq - a_pp) / (2 |a_pq|), t = sign(tau) / (|tau| + hypot(tau, 1)) (t = 1 at
// tau = 0), c = 1 / sqrt(1 + t^2), s = t c a_pq / |a_pq| (:340-347):
//     w_p <- c w_p - conj(s) w_q,  w_q <- s w_p + c w_q  (and the same for V, :349-351).
// sigma_j = |w_j| in FP64, descending order, the reference's cut, U = W / sigma, complex
// coordinate-vector completion (v <- v - U (U^H v), twice), and outputs U, sigma (real) and
// Vh = V^H, interleaved (re, im) row-major like torch's complex tensors.
#include <float.h>

#include "common.hpp"

namespace psvd {

template <typename R>
struct CLdsParams {
  const R *a;   // batch x m x n complex, interleaved, row-major
  R *u, *s, *vh;
  int32_t *info;
  int32_t *queue;
  int64_t batch;
  int m, n;
  double tol;
  int max_sweeps, flags;
  double *gws;       // GL: one region of gstride doubles per workgroup in place of LDS
  int64_t gstride;
};

template <int NW>
__device__ __forceinline__ double slots_max_c(const double *slot) {
  double v = slot[0];
#pragma unroll
  for (int w = 1; w < NW; ++w) v = fmax(v, slot[w]);
  return v;
}

__global__ void zero_queue_kernel_c(int32_t *q) {
  if (threadIdx.x == 0) q[0] = 0;
}

__host__ __device__ inline int c_pad_even(int n) { return n + (n & 1); }
// Column stride of the W (m rows) and V (np rows) planes: the row count rounded up to a multiple
// of 16, plus one (odd); the extra rows are zero and stay zero under the rotations, so a group of
// L <= 16 lanes strides over ceil(rows / L) row slots with no bounds test.
__host__ __device__ inline int c_ld(int rows) { return ((rows + 15) & ~15) | 1; }

template <typename R>
__host__ __device__ inline size_t clds_bytes(int m, int n, bool vectors) {
  const int np = c_pad_even(n), ldw = c_ld(m), ldv = c_ld(np);
  size_t b = sizeof(double) * (size_t)(3 * np + 2 * m);   // sigma, y (re, im), v (re, im)
  b += sizeof(R) * 2 * ((size_t)np * ldw + (vectors ? (size_t)np * ldv : 0));
  b += sizeof(int) * (size_t)np;
  return (b + 15) & ~size_t(15);
}

// rotation of the 2 x 2 Gram block [[app, apq], [conj(apq), aqq]], |apq| = a > 0
__device__ __forceinline__ void cplx_rotation(double app, double aqq, double a, double ar, double ai, double &c,
                                              double &sr, double &si, double &tb) {
  const double tau = (aqq - app) / (2.0 * a);
  const double t = tau == 0.0 ? 1.0 : copysign(1.0, tau) / (fabs(tau) + hypot(tau, 1.0));
  c = 1.0 / sqrt(1.0 + t * t);
  const double f = t * c / a;
  sr = f * ar;
  si = f * ai;
  tb = t * a;
}

// x <- c x - conj(s) y, y <- s x + c y over rows lig, lig + L, ... of the two columns
template <typename R, int L>
__device__ __forceinline__ void cplx_rotate(R *xr, R *xi, R *yr, R *yi, int rows, int lig, R c, R sr, R si) {
  for (int i = lig; i < rows; i += L) {
    const R ar = xr[i], ai = xi[i], br = yr[i], bi = yi[i];
    xr[i] = c * ar - (sr * br + si * bi);
    xi[i] = c * ai - (sr * bi - si * br);
    yr[i] = (sr * ar - si * ai) + c * br;
    yi[i] = (sr * ai + si * ar) + c * bi;
  }
}

// S > 0: every lane holds exactly S = ceil(m / L) row slots of a pair (L <= 16), kept in registers
// from the sums to the rotation with no bounds tests; S = 0: rows streamed with a bounds test.
template <typename R, int NT, int L, int S, bool GL = false>
__global__ __launch_bounds__(NT) void jacobi_lds_cplx_kernel(CLdsParams<R> P) {
  extern __shared__ double smem_d[];
  __shared__ int s_idx;
  __shared__ int s_fslot[2 * (NT / kWave)];
  __shared__ double s_rslot[2 * (NT / kWave)];
  __shared__ int s_nk;
  __shared__ double s_red[NT / kWave];
  __shared__ int s_redi[NT / kWave];

  constexpr int NW = NT / kWave, NG = NT / L;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int grp = tid / L, lig = tid - grp * L;
  const bool vectors = !(P.flags & PSVD_NO_VECTORS);
  const double eps = Eps<R>::value;
  const double tol = P.tol > 0.0 ? P.tol : 10.0 * eps;
  const double tol2 = tol * tol, eps2 = eps * eps;
  const int m = P.m, n = P.n;
  const int np = c_pad_even(n), ldw = c_ld(m), ldv = c_ld(np), npairs = np / 2;
  const int rcv = (np + L - 1) / L;
  // GL: the planes of matrices beyond LDS in the workgroup's region of a global workspace (L2)
  double *sig = GL ? P.gws + (int64_t)blockIdx.x * P.gstride : smem_d;
  double *yr = sig + np, *yi = yr + np;
  double *vr = yi + np, *vi = vr + m;
  R *Wr = reinterpret_cast<R *>(vi + m);
  R *Wi = Wr + (size_t)np * ldw;
  R *Vr = Wi + (size_t)np * ldw;
  R *Vi = Vr + (vectors ? (size_t)np * ldv : 0);
  int *ord = reinterpret_cast<int *>(Vi + (vectors ? (size_t)np * ldv : 0));

  for (;;) {
    if (tid == 0) s_idx = atomicAdd(P.queue, 1);
    __syncthreads();
    const int64_t mi = s_idx;
    __syncthreads();
    if (mi >= P.batch) return;
    const int64_t aoff = mi * (int64_t)m * n, soff = mi * (int64_t)n, voff = mi * (int64_t)n * n;

    // ---- load A into the planes, check finiteness, optional max-abs scale
    const R *A = P.a + 2 * aoff;
    double amax = 0.0, bad = 0.0;
    for (int e = tid; e < m * n; e += NT) {
      const int i = e / n, j = e - i * n;
      const R xr = A[2 * e], xi = A[2 * e + 1];
      Wr[j * ldw + i] = xr;
      Wi[j * ldw + i] = xi;
      const double ax = hypot((double)xr, (double)xi);
      if (ax <= DBL_MAX) amax = fmax(amax, ax); else bad = 1.0;
    }
    {  // zero rows past m in every column, and the padding column when n is odd
      const int padr = ldw - m;
      for (int e = tid; e < np * padr; e += NT) {
        const int j = e / padr, i = m + (e - j * padr);
        Wr[j * ldw + i] = R(0);
        Wi[j * ldw + i] = R(0);
      }
    }
    if (np > n)
      for (int i = tid; i < m; i += NT) {
        Wr[n * ldw + i] = R(0);
        Wi[n * ldw + i] = R(0);
      }
    if (vectors)
      for (int e = tid; e < np * ldv; e += NT) {
        const int j = e / ldv;
        Vr[e] = (e - j * ldv == j) ? R(1) : R(0);
        Vi[e] = R(0);
      }
    if (block_max<NT>(bad, s_red) > 0.0) {
      if (tid == 0) {
        P.info[2 * mi] = 0;
        P.info[2 * mi + 1] = PSVD_INFO_NONFINITE;
      }
      __syncthreads();
      continue;
    }
    double scale = 1.0;
    if (P.flags & PSVD_SCALE) {
      const double mx = block_max<NT>(amax, s_red);
      if (mx > 0.0) {
        scale = mx;
        const R st = (R)mx;
        for (int e = tid; e < m * n; e += NT) {
          const int j = e / m, i = e - j * m;
          Wr[j * ldw + i] = Wr[j * ldw + i] / st;
          Wi[j * ldw + i] = Wi[j * ldw + i] / st;
        }
      }
    }
    __syncthreads();
    {  // initial rotation floor: max_j |w_j|^2, one slot per wave
      double mx = 0.0;
      for (int c = wave; c < n; c += NW) {
        double acc = 0.0;
        for (int i = lane; i < m; i += kWave) {
          const double xr = Wr[c * ldw + i], xi = Wi[c * ldw + i];
          acc = fma(xr, xr, fma(xi, xi, acc));
        }
        mx = fmax(mx, wave_sum(acc));
      }
      if (lane == 0) s_rslot[wave] = mx;
    }
    __syncthreads();

    // ---- sweeps (one barrier per round; the floor is the maximum of the previous round's
    // exact-arithmetic norm updates, as in the real kernel)
    int sweeps = 0;
    bool converged = false;
    int g = 0;
    for (int sw = 1; sw <= P.max_sweeps; ++sw) {
      sweeps = sw;
      bool wave_active = false;
      for (int r = 0; r < np - 1; ++r, ++g) {
        const double floor_ = eps2 * slots_max_c<NW>(s_rslot + (g & 1) * NW);
        double mymax = 0.0;
        for (int k = grp; k < npairs; k += NG) {
          int p, q;
          rr_pair(r, k, np, p, q);
          R *pr = Wr + p * ldw, *pi = Wi + p * ldw, *qr = Wr + q * ldw, *qi = Wi + q * ldw;
          double app = 0.0, aqq = 0.0, ar = 0.0, ai = 0.0;
          R xr_[S > 0 ? S : 1], xi_[S > 0 ? S : 1], br_[S > 0 ? S : 1], bi_[S > 0 ? S : 1];
          if constexpr (S > 0) {
#pragma unroll
            for (int j = 0; j < S; ++j) {
              xr_[j] = pr[lig + j * L];
              xi_[j] = pi[lig + j * L];
              br_[j] = qr[lig + j * L];
              bi_[j] = qi[lig + j * L];
            }
#pragma unroll
            for (int j = 0; j < S; ++j) {
              const double xr = xr_[j], xi = xi_[j], br = br_[j], bi = bi_[j];
              app = fma(xr, xr, fma(xi, xi, app));
              aqq = fma(br, br, fma(bi, bi, aqq));
              ar = fma(xr, br, fma(xi, bi, ar));
              ai = fma(xr, bi, fma(-xi, br, ai));
            }
          } else {
            for (int i = lig; i < m; i += L) {
              const double xr = pr[i], xi = pi[i], br = qr[i], bi = qi[i];
              app = fma(xr, xr, fma(xi, xi, app));
              aqq = fma(br, br, fma(bi, bi, aqq));
              ar = fma(xr, br, fma(xi, bi, ar));
              ai = fma(xr, bi, fma(-xi, br, ai));
            }
          }
          group_sum3_dpp<L>(app, aqq, ar);
          ai = group_sum_dpp<L>(ai);
          const double apq2 = ar * ar + ai * ai;
          double nap = app, naq = aqq;
          if (apq2 > tol2 * app * aqq && fmin(app, aqq) > floor_) {
            double c, sr, si, tb;
            cplx_rotation(app, aqq, sqrt(apq2), ar, ai, c, sr, si, tb);
            nap = app - tb;
            naq = aqq + tb;
            wave_active = true;
            const R cc = (R)c, rs = (R)sr, is = (R)si;
            if constexpr (S > 0) {
#pragma unroll
              for (int j = 0; j < S; ++j) {
                const R ar_ = xr_[j], ai_ = xi_[j], br = br_[j], bi = bi_[j];
                pr[lig + j * L] = cc * ar_ - (rs * br + is * bi);
                pi[lig + j * L] = cc * ai_ - (rs * bi - is * br);
                qr[lig + j * L] = (rs * ar_ - is * ai_) + cc * br;
                qi[lig + j * L] = (rs * ai_ + is * ar_) + cc * bi;
              }
              if (vectors) {   // rows past np are zero: rcv slots with no bounds test
                R *vpr = Vr + p * ldv + lig, *vpi = Vi + p * ldv + lig, *vqr = Vr + q * ldv + lig,
                  *vqi = Vi + q * ldv + lig;
                for (int j = 0; j < rcv; ++j) {
                  const R ar_ = vpr[j * L], ai_ = vpi[j * L], br = vqr[j * L], bi = vqi[j * L];
                  vpr[j * L] = cc * ar_ - (rs * br + is * bi);
                  vpi[j * L] = cc * ai_ - (rs * bi - is * br);
                  vqr[j * L] = (rs * ar_ - is * ai_) + cc * br;
                  vqi[j * L] = (rs * ai_ + is * ar_) + cc * bi;
                }
              }
            } else {
              cplx_rotate<R, L>(pr, pi, qr, qi, m, lig, cc, rs, is);
              if (vectors) cplx_rotate<R, L>(Vr + p * ldv, Vi + p * ldv, Vr + q * ldv, Vi + q * ldv, np, lig, cc, rs,
                                             is);
            }
          }
          mymax = fmax(mymax, fmax(nap, naq));
        }
        mymax = wave_max_fast(mymax);
        if (lane == 0) s_rslot[((g + 1) & 1) * NW + wave] = mymax;
        __syncthreads();
      }
      const int wave_any = __any(wave_active) ? 1 : 0;
      if (lane == 0) s_fslot[(sw & 1) * NW + wave] = wave_any;
      __syncthreads();
      int any = 0;
#pragma unroll
      for (int w = 0; w < NW; ++w) any |= s_fslot[(sw & 1) * NW + w];
      if (!any) {
        converged = true;
        break;
      }
    }
    int status = converged ? PSVD_INFO_OK : PSVD_INFO_NOCONV;

    // ---- sigma in FP64, descending order, cut
    for (int c = wave; c < n; c += NW) {
      double acc = 0.0;
      for (int i = lane; i < m; i += kWave) {
        const double xr = Wr[c * ldw + i], xi = Wi[c * ldw + i];
        acc = fma(xr, xr, fma(xi, xi, acc));
      }
      acc = wave_sum(acc);
      if (lane == 0) sig[c] = sqrt(acc);
    }
    __syncthreads();
    for (int j = tid; j < n; j += NT) {
      const double sj = sig[j];
      int rank = 0;
      for (int k = 0; k < n; ++k) {
        const double sk = sig[k];
        rank += (sk > sj) || (sk == sj && k < j);
      }
      ord[rank] = j;
    }
    __syncthreads();
    const double s1 = sig[ord[0]];
    const double cut = (P.flags & PSVD_CUT_FLOOR) ? eps * s1 : eps * (double)(m > np ? m : np) * s1;
    if (tid == 0) {
      int nk = 0;
      while (nk < n && sig[ord[nk]] > cut) ++nk;
      s_nk = nk;
    }
    __syncthreads();
    const int nk = s_nk;

    // ---- U = W / sigma for kept columns; complex coordinate-vector completion for the rest
    if (vectors) {
      for (int e = tid; e < n * m; e += NT) {
        const int r = e / m, i = e - r * m;
        const int c = ord[r];
        if (r < nk) {
          Wr[c * ldw + i] = (R)((double)Wr[c * ldw + i] / sig[c]);
          Wi[c * ldw + i] = (R)((double)Wi[c * ldw + i] / sig[c]);
        } else {
          Wr[c * ldw + i] = R(0);
          Wi[c * ldw + i] = R(0);
        }
      }
      __syncthreads();
      for (int r = nk; r < n; ++r) {
        double best = -DBL_MAX;
        int bi = 0x7fffffff;
        for (int i = tid; i < m; i += NT) {
          double acc = 0.0;
          for (int c = 0; c < n; ++c) {
            const double xr = Wr[c * ldw + i], xi = Wi[c * ldw + i];
            acc = fma(xr, xr, fma(xi, xi, acc));
          }
          const double lev = 1.0 - acc;
          if (lev > best) {
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
        __syncthreads();
        if (lane == 0) {
          s_red[wave] = best;
          s_redi[wave] = bi;
        }
        __syncthreads();
        best = s_red[0];
        bi = s_redi[0];
        for (int w = 1; w < NW; ++w)
          if (s_red[w] > best || (s_red[w] == best && s_redi[w] < bi)) {
            best = s_red[w];
            bi = s_redi[w];
          }
        const int ix = bi;
        // v = e_ix - U (U^H e_ix) = e_ix - U conj(U(ix, :))^T
        for (int i = tid; i < m; i += NT) {
          double ar = (i == ix) ? 1.0 : 0.0, ai = 0.0;
          for (int c = 0; c < n; ++c) {
            const double ur = Wr[c * ldw + i], ui = Wi[c * ldw + i];
            const double xr = Wr[c * ldw + ix], xi = Wi[c * ldw + ix];
            ar -= ur * xr + ui * xi;
            ai -= ui * xr - ur * xi;
          }
          vr[i] = ar;
          vi[i] = ai;
        }
        __syncthreads();
        // y = U^H v, then v = v - U y
        for (int c = wave; c < n; c += NW) {
          double ar = 0.0, ai = 0.0;
          for (int i = lane; i < m; i += kWave) {
            const double ur = Wr[c * ldw + i], ui = Wi[c * ldw + i];
            ar = fma(ur, vr[i], fma(ui, vi[i], ar));
            ai = fma(ur, vi[i], fma(-ui, vr[i], ai));
          }
          ar = wave_sum(ar);
          ai = wave_sum(ai);
          if (lane == 0) {
            yr[c] = ar;
            yi[c] = ai;
          }
        }
        __syncthreads();
        double l2 = 0.0;
        for (int i = tid; i < m; i += NT) {
          double ar = vr[i], ai = vi[i];
          for (int c = 0; c < n; ++c) {
            const double ur = Wr[c * ldw + i], ui = Wi[c * ldw + i];
            ar -= ur * yr[c] - ui * yi[c];
            ai -= ur * yi[c] + ui * yr[c];
          }
          vr[i] = ar;
          vi[i] = ai;
          l2 = fma(ar, ar, fma(ai, ai, l2));
        }
        const double len = sqrt(block_sum<NT>(l2, s_red));
        if (len < 10.0 * eps) status = PSVD_INFO_COMPLETION;
        const double inv = 1.0 / fmax(len, DBL_MIN);
        const int c = ord[r];
        for (int i = tid; i < m; i += NT) {
          Wr[c * ldw + i] = (R)(vr[i] * inv);
          Wi[c * ldw + i] = (R)(vi[i] * inv);
        }
        __syncthreads();
      }
    }

    // ---- outputs: U (m x n), sigma (real), Vh = V^H (n x n), interleaved complex
    R *So = P.s + soff;
    for (int r = tid; r < n; r += NT) So[r] = r < nk ? (R)(sig[ord[r]] * scale) : R(0);
    if (vectors) {
      R *U = P.u + 2 * aoff;
      for (int e = tid; e < m * n; e += NT) {
        const int i = e / n, r = e - i * n;
        U[2 * e] = Wr[ord[r] * ldw + i];
        U[2 * e + 1] = Wi[ord[r] * ldw + i];
      }
      R *VH = P.vh + 2 * voff;
      for (int e = tid; e < n * n; e += NT) {
        const int r = e / n, j = e - r * n;
        VH[2 * e] = Vr[ord[r] * ldv + j];
        VH[2 * e + 1] = -Vi[ord[r] * ldv + j];
      }
    }
    if (tid == 0) {
      P.info[2 * mi] = sweeps;
      P.info[2 * mi + 1] = status;
    }
    __syncthreads();
  }
}

constexpr size_t kCLdsCap = 160 * 1024 - 4096;

template <typename R, int L, int S>
static int launch_clds(const CLdsParams<R> &P, size_t smem, hipStream_t stream) {
  constexpr int NT = 256;
  auto kern = jacobi_lds_cplx_kernel<R, NT, L, S>;
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kCLdsCap) !=
        hipSuccess)
      return PSVD_ELAUNCH;
    attr = true;
  }
  int occ = 1, cus = 256, dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (cus <= 0) cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, NT, smem) != hipSuccess || occ < 1) occ = 1;
  int64_t grid = (int64_t)cus * occ;
  if (grid > P.batch) grid = P.batch;
  hipLaunchKernelGGL(zero_queue_kernel_c, dim3(1), dim3(64), 0, stream, P.queue);
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(NT), smem, stream, P);
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

template <typename R>
static int run_clds_t(int64_t batch, int m, int n, const void *a, void *u, void *s, void *vh, int32_t *info,
                      double tol, int max_sweeps, int flags, int lanes, void *ws, hipStream_t stream) {
  const bool vectors = !(flags & PSVD_NO_VECTORS);
  const size_t smem = clds_bytes<R>(m, n, vectors);
  if (smem > kCLdsCap) return PSVD_ETOOBIG;
  if (lanes > kWave || (lanes > 0 && (lanes & (lanes - 1))) || lanes < -kWave) return PSVD_EINVAL;
  if (lanes <= 0) {  // about four rows per lane
    const int rpl = lanes < 0 ? -lanes : 4;
    const int target = (m + rpl - 1) / rpl;
    lanes = 2;
    while (lanes < target && lanes < kWave) lanes <<= 1;
  }
  if (lanes == 1) lanes = 2;
  CLdsParams<R> P;
  P.a = static_cast<const R *>(a);
  P.u = static_cast<R *>(u);
  P.s = static_cast<R *>(s);
  P.vh = static_cast<R *>(vh);
  P.info = info;
  P.queue = static_cast<int32_t *>(ws);
  P.batch = batch;
  P.m = m;
  P.n = n;
  P.tol = tol;
  P.max_sweeps = max_sweeps;
  P.flags = flags;
  P.gws = nullptr;
  P.gstride = 0;
  // row slots per lane kept in registers: up to 8 (complex64) or 4 (complex128), L <= 16
  constexpr int SMAX = sizeof(R) == 4 ? 8 : 4;
  const int slots = (m + lanes - 1) / lanes;
  const int S = (lanes <= 16 && slots <= SMAX) ? slots : 0;
#define PSVD_CLDS_L(LV)                                                                                   \
  switch (S) {                                                                                            \
    case 1: return launch_clds<R, LV, 1>(P, smem, stream);                                                \
    case 2: return launch_clds<R, LV, 2>(P, smem, stream);                                                \
    case 3: return launch_clds<R, LV, 3>(P, smem, stream);                                                \
    case 4: return launch_clds<R, LV, 4>(P, smem, stream);                                                \
    case 5: return launch_clds<R, LV, (SMAX >= 5 ? 5 : 0)>(P, smem, stream);                              \
    case 6: return launch_clds<R, LV, (SMAX >= 6 ? 6 : 0)>(P, smem, stream);                              \
    case 7: return launch_clds<R, LV, (SMAX >= 7 ? 7 : 0)>(P, smem, stream);                              \
    case 8: return launch_clds<R, LV, (SMAX >= 8 ? 8 : 0)>(P, smem, stream);                              \
    default: return launch_clds<R, LV, 0>(P, smem, stream);                                               \
  }
  switch (lanes) {
    case 2: PSVD_CLDS_L(2)
    case 4: PSVD_CLDS_L(4)
    case 8: PSVD_CLDS_L(8)
    case 16: PSVD_CLDS_L(16)
    case 32: return launch_clds<R, 32, 0>(P, smem, stream);
    default: return launch_clds<R, 64, 0>(P, smem, stream);
  }
#undef PSVD_CLDS_L
}

// Matrices whose planes exceed LDS: the general (streamed-row) kernel with 64 lanes per pair on
// regions of the workspace, one per resident workgroup (the work queue hands each its matrices).
constexpr int kCglGrid = 512;   // workgroups: two per CU

static int64_t cgl_stride(size_t bytes) { return (int64_t)((bytes + 255) / 256 * 32); }

template <typename R>
static size_t cgl_ws_bytes(int64_t batch, int m, int n, bool vectors) {
  const int64_t grid = batch < kCglGrid ? batch : kCglGrid;
  return 256 + (size_t)(grid * cgl_stride(clds_bytes<R>(m, n, vectors))) * sizeof(double);
}

template <typename R>
static int run_cgl(int64_t batch, int m, int n, const void *a, void *u, void *s, void *vh, int32_t *info, double tol,
                   int max_sweeps, int flags, void *ws, size_t ws_bytes, hipStream_t stream) {
  const bool vectors = !(flags & PSVD_NO_VECTORS);
  if (ws_bytes < cgl_ws_bytes<R>(batch, m, n, vectors)) return PSVD_EINVAL;
  CLdsParams<R> P;
  P.a = static_cast<const R *>(a);
  P.u = static_cast<R *>(u);
  P.s = static_cast<R *>(s);
  P.vh = static_cast<R *>(vh);
  P.info = info;
  P.queue = static_cast<int32_t *>(ws);
  P.batch = batch;
  P.m = m;
  P.n = n;
  P.tol = tol;
  P.max_sweeps = max_sweeps;
  P.flags = flags;
  P.gws = reinterpret_cast<double *>(static_cast<char *>(ws) + 256);
  P.gstride = cgl_stride(clds_bytes<R>(m, n, vectors));
  const int64_t grid = batch < kCglGrid ? batch : kCglGrid;
  hipLaunchKernelGGL(zero_queue_kernel_c, dim3(1), dim3(64), 0, stream, P.queue);
  hipLaunchKernelGGL((jacobi_lds_cplx_kernel<R, 256, 64, 0, true>), dim3((unsigned)grid), dim3(256), 0, stream, P);
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

size_t clds_smem_bytes(int dtype, int m, int n, bool vectors) {
  return dtype == PSVD_C64 ? clds_bytes<float>(m, n, vectors) : clds_bytes<double>(m, n, vectors);
}

int run_clds(int dtype, int64_t batch, int m, int n, const void *a, void *u, void *s, void *vh, int32_t *info,
             double tol, int max_sweeps, int flags, int lanes, void *ws, hipStream_t stream) {
  if (dtype == PSVD_C64)
    return run_clds_t<float>(batch, m, n, a, u, s, vh, info, tol, max_sweeps, flags, lanes, ws, stream);
  if (dtype == PSVD_C128)
    return run_clds_t<double>(batch, m, n, a, u, s, vh, info, tol, max_sweeps, flags, lanes, ws, stream);
  return PSVD_EINVAL;
}

}  // namespace psvd

extern "C" {

size_t psvd_cjacobi_batched_workspace_bytes(int dtype, int64_t batch, int m, int n, int flags) {
  if (batch < 0 || m < 1 || n < 1 || n > m) return 0;
  const bool vectors = !(flags & PSVD_NO_VECTORS);
  if (dtype == PSVD_C64)
    return psvd::clds_bytes<float>(m, n, vectors) > psvd::kCLdsCap ? psvd::cgl_ws_bytes<float>(batch, m, n, vectors)
                                                                   : 256;
  if (dtype == PSVD_C128)
    return psvd::clds_bytes<double>(m, n, vectors) > psvd::kCLdsCap ? psvd::cgl_ws_bytes<double>(batch, m, n, vectors)
                                                                    : 256;
  return 0;
}

int psvd_cjacobi_batched(int dtype, int64_t batch, int m, int n, const void *a, void *u, void *s, void *vh,
                         int32_t *info, double tol, int max_sweeps, int flags, void *workspace, size_t workspace_bytes,
                         psvd_stream_t stream) {
  if (batch < 0 || batch > 0x7fffffff || m < 1 || n < 1 || n > m || max_sweeps < 1 || !a || !s || !info ||
      !workspace || workspace_bytes < 256)
    return PSVD_EINVAL;
  if (!(flags & PSVD_NO_VECTORS) && (!u || !vh)) return PSVD_EINVAL;
  if (batch == 0) return PSVD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool vectors = !(flags & PSVD_NO_VECTORS);
  if (dtype == PSVD_C64) {
    if (psvd::clds_bytes<float>(m, n, vectors) <= psvd::kCLdsCap)
      return psvd::run_clds(dtype, batch, m, n, a, u, s, vh, info, tol, max_sweeps, flags, 0, workspace, st);
    return psvd::run_cgl<float>(batch, m, n, a, u, s, vh, info, tol, max_sweeps, flags, workspace, workspace_bytes, st);
  }
  if (dtype == PSVD_C128) {
    if (psvd::clds_bytes<double>(m, n, vectors) <= psvd::kCLdsCap)
      return psvd::run_clds(dtype, batch, m, n, a, u, s, vh, info, tol, max_sweeps, flags, 0, workspace, st);
    return psvd::run_cgl<double>(batch, m, n, a, u, s, vh, info, tol, max_sweeps, flags, workspace, workspace_bytes,
                                 st);
  }
  return PSVD_EINVAL;
}

}  // extern "C"
