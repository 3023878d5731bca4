// Complex Hermitian kernels for matrices held in LDS (one workgroup per matrix): the two-sided
// Jacobi eigensolver, the Cholesky factorisation and the right-sided solve X <- X (L L^H)^-1 of the
// reference's complex QDWH Cholesky branch.
//
// ceigh_lds_kernel  svdbench/standalone/primitives.py:370-427 (jacobi_eigh, complex branch): per
//                   round of the tournament (:309-313), g_pp, g_qq real, a = |g_pq|, active when
//                   a > tol max|G| (the whole matrix at the start of the round), tau = (g_qq - g_pp)
//                   / 2a, t = sign(tau) / (|tau| + hypot(tau, 1)) (t = 1 at tau = 0),
//                   c = 1 / sqrt(1 + t^2), s = t c g_pq / a; G <- G J (columns: c g_p - conj(s) g_q,
//                   s g_p + c g_q), V <- V J, then G <- J^H G (rows: c r_p - s r_q, conj(s) r_p +
//                   c r_q); stop after a sweep without an active pair; eigenvalues ascending.
// cchol_lds_kernel  primitives.py:145-159 (cholesky): column by column, pivot Re(g_jj) - sum |l_jk|^2,
//                   failure when it is <= 0, l_jj = sqrt(max(pivot, tiny)), the column below the
//                   diagonal (g_ij - sum l_ik conj(l_jk)) / l_jj.
// csolve_kernel     X Z^-1 with Z = L L^H as the reference forms it (solvers.py:159-160:
//                   trsm_upper(L^H, trsm_lower(L, X^H))^H, primitives.py:162-179), one row of X per
//                   thread: t = x L^-H by forward substitution, then w = t L^-1 by back substitution.
//
// Complex values are interleaved (re, im) pairs in global memory, as torch stores them, and
// separate real and imaginary planes in LDS.  Sums are accumulated in FP64 and rounded once.
// Matrices whose planes exceed LDS run the same kernels (GL = true) with the planes in a global
// workspace, one region per matrix, served by L2: the same arithmetic in the same order.
#include <float.h>

#include "common.hpp"

namespace psvd {

constexpr int kCsmThreads = 256;

__host__ __device__ inline int ceigh_ld(int np) { return np + 1; }

template <typename R>
__host__ __device__ inline size_t ceigh_lds_bytes(int n) {
  const int np = n + (n & 1), ld = ceigh_ld(np);
  size_t b = sizeof(R) * 4 * (size_t)np * ld;                 // G and V^T, real and imaginary planes
  b += sizeof(double) * 3 * (size_t)(np / 2);                 // c, Re s, Im s
  b += sizeof(int) * 2 * (size_t)(np / 2);                    // p, q
  b += sizeof(double) * np + sizeof(int) * np;                // eigenvalues, order
  return (b + 15) & ~size_t(15);
}

template <typename R>
struct CeighParams {
  const R *h;        // (batch, n, n) interleaved complex, Hermitian
  R *w, *v;          // w (batch, n) real ascending; v (batch, n, n) interleaved complex, eigenvectors as columns
  int32_t *info;
  int n, max_sweeps;
  double tol;
  double *gws;       // GL: the per-matrix regions (gstride doubles each) that replace LDS
  int64_t gstride;
};

// One 2x2 block (rows p1, q1; columns p2, q2) of G <- J1^H G J2: columns with pair 2's rotation, then
// rows with pair 1's, each rounded to R (the reference rounds every elementwise step to its dtype);
// off the diagonal of the pair grid the mirror block gets the conjugate transpose.  Returns the new
// largest modulus of the block.
template <typename R>
__device__ __forceinline__ double cblock_update(R *Gr, R *Gi, int ld, int p1, int q1, int p2, int q2, double c1,
                                                double s1r, double s1i, double c2, double s2r, double s2i,
                                                bool act1, bool act2) {
  const int e11 = p1 * ld + p2, e12 = p1 * ld + q2, e21 = q1 * ld + p2, e22 = q1 * ld + q2;
  double b11r = Gr[e11], b11i = Gi[e11], b12r = Gr[e12], b12i = Gi[e12];
  double b21r = Gr[e21], b21i = Gi[e21], b22r = Gr[e22], b22i = Gi[e22];
  if (act2) {   // col_p <- c col_p - conj(s) col_q, col_q <- s col_p + c col_q
    const double t11r = c2 * b11r - (s2r * b12r + s2i * b12i), t11i = c2 * b11i - (s2r * b12i - s2i * b12r);
    const double t12r = (s2r * b11r - s2i * b11i) + c2 * b12r, t12i = (s2r * b11i + s2i * b11r) + c2 * b12i;
    const double t21r = c2 * b21r - (s2r * b22r + s2i * b22i), t21i = c2 * b21i - (s2r * b22i - s2i * b22r);
    const double t22r = (s2r * b21r - s2i * b21i) + c2 * b22r, t22i = (s2r * b21i + s2i * b21r) + c2 * b22i;
    b11r = (R)t11r; b11i = (R)t11i; b12r = (R)t12r; b12i = (R)t12i;
    b21r = (R)t21r; b21i = (R)t21i; b22r = (R)t22r; b22i = (R)t22i;
  }
  if (act1) {   // row_p <- c row_p - s row_q, row_q <- conj(s) row_p + c row_q
    const double t11r = c1 * b11r - (s1r * b21r - s1i * b21i), t11i = c1 * b11i - (s1r * b21i + s1i * b21r);
    const double t21r = (s1r * b11r + s1i * b11i) + c1 * b21r, t21i = (s1r * b11i - s1i * b11r) + c1 * b21i;
    const double t12r = c1 * b12r - (s1r * b22r - s1i * b22i), t12i = c1 * b12i - (s1r * b22i + s1i * b22r);
    const double t22r = (s1r * b12r + s1i * b12i) + c1 * b22r, t22i = (s1r * b12i - s1i * b12r) + c1 * b22i;
    b11r = t11r; b11i = t11i; b12r = t12r; b12i = t12i;
    b21r = t21r; b21i = t21i; b22r = t22r; b22i = t22i;
  }
  const R n11r = (R)b11r, n11i = (R)b11i, n12r = (R)b12r, n12i = (R)b12i;
  const R n21r = (R)b21r, n21i = (R)b21i, n22r = (R)b22r, n22i = (R)b22i;
  Gr[e11] = n11r; Gi[e11] = n11i; Gr[e12] = n12r; Gi[e12] = n12i;
  Gr[e21] = n21r; Gi[e21] = n21i; Gr[e22] = n22r; Gi[e22] = n22i;
  if (p1 != p2) {   // mirror block: G[p2][p1] = conj(G[p1][p2]), ...
    Gr[p2 * ld + p1] = n11r; Gi[p2 * ld + p1] = -n11i;
    Gr[q2 * ld + p1] = n12r; Gi[q2 * ld + p1] = -n12i;
    Gr[p2 * ld + q1] = n21r; Gi[p2 * ld + q1] = -n21i;
    Gr[q2 * ld + q1] = n22r; Gi[q2 * ld + q1] = -n22i;
  }
  auto mod = [](R re, R im) { return hypot((double)re, (double)im); };
  return fmax(fmax(mod(n11r, n11i), mod(n12r, n12i)), fmax(mod(n21r, n21i), mod(n22r, n22i)));
}

template <typename R, bool GL>
__global__ __launch_bounds__(kCsmThreads) void ceigh_lds_kernel(CeighParams<R> P) {
  constexpr int NT = kCsmThreads, NW = NT / kWave;
  extern __shared__ double smem_ce[];
  __shared__ double s_slot[2 * NW];
  __shared__ int s_flag[2 * NW];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int n = P.n, np = n + (n & 1), ld = ceigh_ld(np), npairs = np / 2;
  const int64_t b = blockIdx.x;
  double *cc = GL ? P.gws + b * P.gstride : smem_ce, *sr = cc + npairs, *si = sr + npairs, *dw = si + npairs;
  int *pp = reinterpret_cast<int *>(dw + np), *qq = pp + npairs, *ord = qq + npairs;
  R *Gr = reinterpret_cast<R *>(ord + np + (np & 1));
  R *Gi = Gr + np * ld, *Vr = Gi + np * ld, *Vi = Vr + np * ld;
  const R *H = P.h + 2 * b * (int64_t)n * n;
  const double eps = Eps<R>::value;
  const double tol = P.tol > 0.0 ? P.tol : 10.0 * eps;

  double mx = 0.0;
  for (int e = tid; e < np * np; e += NT) {
    const int i = e / np, j = e - i * np;
    const bool in = i < n && j < n;
    const R xr = in ? H[2 * ((int64_t)i * n + j)] : R(0), xi = in ? H[2 * ((int64_t)i * n + j) + 1] : R(0);
    Gr[i * ld + j] = xr;
    Gi[i * ld + j] = xi;
    Vr[i * ld + j] = i == j ? R(1) : R(0);
    Vi[i * ld + j] = R(0);
    mx = fmax(mx, hypot((double)xr, (double)xi));
  }
  mx = wave_max_fast(mx);
  if (lane == 0) s_slot[wave] = mx;
  __syncthreads();
  int sweeps = 0;
  bool converged = false;
  int g = 0;
  for (int sw = 1; sw <= P.max_sweeps; ++sw) {
    sweeps = sw;
    bool wave_active = false;
    for (int r = 0; r < np - 1; ++r, ++g) {
      double scale = s_slot[(g & 1) * NW];
      for (int w = 1; w < NW; ++w) scale = fmax(scale, s_slot[(g & 1) * NW + w]);
      for (int k = tid; k < npairs; k += NT) {
        int p, q;
        rr_pair(r, k, np, p, q);
        const double gpp = (double)Gr[p * ld + p], gqq = (double)Gr[q * ld + q];
        const double gr = (double)Gr[p * ld + q], gi = (double)Gi[p * ld + q];
        const double a = hypot(gr, gi);
        const bool act = a > tol * scale;
        double c = 1.0, s_r = 0.0, s_i = 0.0;
        if (act) {
          const double tau = (gqq - gpp) / (2.0 * a);
          const double t = tau == 0.0 ? 1.0 : copysign(1.0, tau) / (fabs(tau) + hypot(tau, 1.0));
          c = 1.0 / sqrt(1.0 + t * t);
          s_r = t * c * (gr / a);
          s_i = t * c * (gi / a);
        }
        cc[k] = c;
        sr[k] = s_r;
        si[k] = s_i;
        pp[k] = p;
        qq[k] = q;
        wave_active |= act;
      }
      __syncthreads();
      double mymax = 0.0;
      const int nblk = npairs * (npairs + 1) / 2;
      for (int e = tid; e < nblk; e += NT) {
        int k1, k2;
        tri_decode(e, npairs, k1, k2);
        const bool a1 = sr[k1] != 0.0 || si[k1] != 0.0, a2 = sr[k2] != 0.0 || si[k2] != 0.0;
        if (a1 || a2) {
          mymax = fmax(mymax, cblock_update(Gr, Gi, ld, pp[k1], qq[k1], pp[k2], qq[k2], cc[k1], sr[k1], si[k1], cc[k2],
                                            sr[k2], si[k2], a1, a2));
        } else {
          const int p1 = pp[k1], q1 = qq[k1], p2 = pp[k2], q2 = qq[k2];
          auto mod = [&](int i, int j) { return hypot((double)Gr[i * ld + j], (double)Gi[i * ld + j]); };
          mymax = fmax(mymax, fmax(fmax(mod(p1, p2), mod(p1, q2)), fmax(mod(q1, p2), mod(q1, q2))));
        }
      }
      // V <- V J: rows of V^T, Vt_p <- c Vt_p - conj(s) Vt_q, Vt_q <- s Vt_p + c Vt_q
      for (int e = tid; e < npairs * np; e += NT) {
        const int k = e / np, j = e - k * np;
        const double s_r = sr[k], s_i = si[k];
        if (s_r == 0.0 && s_i == 0.0) continue;
        const double c = cc[k];
        const int ip = pp[k] * ld + j, iq = qq[k] * ld + j;
        const double xr = Vr[ip], xi = Vi[ip], yr = Vr[iq], yi = Vi[iq];
        Vr[ip] = (R)(c * xr - (s_r * yr + s_i * yi));
        Vi[ip] = (R)(c * xi - (s_r * yi - s_i * yr));
        Vr[iq] = (R)((s_r * xr - s_i * xi) + c * yr);
        Vi[iq] = (R)((s_r * xi + s_i * xr) + c * yi);
      }
      mymax = wave_max_fast(mymax);
      if (lane == 0) s_slot[((g + 1) & 1) * NW + wave] = mymax;
      __syncthreads();
    }
    const int wa = __any(wave_active) ? 1 : 0;
    if (lane == 0) s_flag[(sw & 1) * NW + wave] = wa;
    __syncthreads();
    int any = 0;
    for (int w = 0; w < NW; ++w) any |= s_flag[(sw & 1) * NW + w];
    if (!any) {
      converged = true;
      break;
    }
  }
  for (int j = tid; j < n; j += NT) dw[j] = (double)Gr[j * ld + j];
  __syncthreads();
  for (int j = tid; j < n; j += NT) {
    const double x = dw[j];
    int rank = 0;
    for (int k = 0; k < n; ++k) rank += (dw[k] < x) || (dw[k] == x && k < j);
    ord[rank] = j;
  }
  __syncthreads();
  R *W = P.w + b * (int64_t)n;
  R *Vo = P.v + 2 * b * (int64_t)n * n;
  for (int r = tid; r < n; r += NT) W[r] = (R)dw[ord[r]];
  for (int e = tid; e < n * n; e += NT) {
    const int i = e / n, r = e - i * n;
    Vo[2 * e] = Vr[ord[r] * ld + i];
    Vo[2 * e + 1] = Vi[ord[r] * ld + i];
  }
  if (tid == 0) {
    P.info[2 * b] = sweeps;
    P.info[2 * b + 1] = converged ? PSVD_INFO_OK : PSVD_INFO_NOCONV;
  }
}

// ---- Cholesky ---------------------------------------------------------------------------------
template <typename R>
__host__ __device__ inline size_t cchol_lds_bytes(int n) {
  return sizeof(R) * 2 * (size_t)n * (n + 1) + 16;
}

template <typename R, bool GL>
__global__ __launch_bounds__(kCsmThreads) void cchol_lds_kernel(const R *G, R *L, int32_t *fail, int n, double *gws,
                                                               int64_t gstride) {
  constexpr int NT = kCsmThreads, NW = NT / kWave;
  extern __shared__ double smem_cc[];
  __shared__ double s_red[NW];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int ld = n + 1;
  const int64_t b = blockIdx.x;
  R *Lr = reinterpret_cast<R *>(GL ? gws + b * gstride : smem_cc), *Li = Lr + (size_t)n * ld;
  const R *Gb = G + 2 * b * (int64_t)n * n;
  for (int e = tid; e < n * n; e += NT) {   // the lower triangle of G; L's columns overwrite it in place
    const int i = e / n, j = e - i * n;
    Lr[i * ld + j] = i >= j ? Gb[2 * e] : R(0);
    Li[i * ld + j] = i >= j ? Gb[2 * e + 1] : R(0);
  }
  __syncthreads();
  const double tiny = sizeof(R) == 4 ? (double)FLT_MIN : DBL_MIN;
  int failed = 0;
  for (int j = 0; j < n; ++j) {
    // pivot: Re(g_jj) - sum_k |l_jk|^2
    double s = 0.0;
    for (int k = tid; k < j; k += NT) s += (double)Lr[j * ld + k] * Lr[j * ld + k] + (double)Li[j * ld + k] * Li[j * ld + k];
    s = group_sum_dpp<kWave>(s);
    if (lane == 0) s_red[wave] = s;
    __syncthreads();
    double tot = 0.0;
    for (int w = 0; w < NW; ++w) tot += s_red[w];
    const R piv = (R)((double)Lr[j * ld + j] - tot);
    failed |= piv <= R(0);
    const R ljj = (R)sqrt(fmax((double)piv, tiny));
    // column j below the diagonal: (g_ij - sum_k l_ik conj(l_jk)) / l_jj
    for (int i = j + 1 + tid; i < n; i += NT) {
      double ar = Lr[i * ld + j], ai = Li[i * ld + j];
      for (int k = 0; k < j; ++k) {
        const double xr = Lr[i * ld + k], xi = Li[i * ld + k], yr = Lr[j * ld + k], yi = Li[j * ld + k];
        ar -= xr * yr + xi * yi;
        ai -= xi * yr - xr * yi;
      }
      Lr[i * ld + j] = (R)(ar / (double)ljj);
      Li[i * ld + j] = (R)(ai / (double)ljj);
    }
    __syncthreads();
    if (tid == 0) {
      Lr[j * ld + j] = ljj;
      Li[j * ld + j] = R(0);
    }
    __syncthreads();
  }
  R *Lb = L + 2 * b * (int64_t)n * n;
  for (int e = tid; e < n * n; e += NT) {
    const int i = e / n, j = e - i * n;
    Lb[2 * e] = i >= j ? Lr[i * ld + j] : R(0);
    Lb[2 * e + 1] = i >= j ? Li[i * ld + j] : R(0);
  }
  if (tid == 0) fail[b] = failed;
}

// ---- X <- X (L L^H)^-1 ------------------------------------------------------------------------
template <typename R>
__host__ __device__ inline size_t csolve_lds_bytes(int m, int n) {
  return sizeof(R) * 2 * ((size_t)n * (n + 1) + (size_t)m * (n + 1)) + 16;
}

template <typename R, bool GL>
__global__ __launch_bounds__(kCsmThreads) void csolve_kernel(const R *L, R *X, int m, int n, double *gws,
                                                            int64_t gstride) {
  constexpr int NT = kCsmThreads;
  extern __shared__ double smem_cs[];
  const int tid = threadIdx.x, ld = n + 1;
  const int64_t b = blockIdx.x;
  R *Lr = reinterpret_cast<R *>(GL ? gws + b * gstride : smem_cs), *Li = Lr + (size_t)n * ld;
  R *Xr = Li + (size_t)n * ld, *Xi = Xr + (size_t)m * ld;
  const R *Lb = L + 2 * b * (int64_t)n * n;
  R *Xb = X + 2 * b * (int64_t)m * n;
  for (int e = tid; e < n * n; e += NT) {
    const int i = e / n, j = e - i * n;
    Lr[i * ld + j] = Lb[2 * e];
    Li[i * ld + j] = Lb[2 * e + 1];
  }
  for (int e = tid; e < m * n; e += NT) {
    const int i = e / n, j = e - i * n;
    Xr[i * ld + j] = Xb[2 * e];
    Xi[i * ld + j] = Xb[2 * e + 1];
  }
  __syncthreads();
  for (int i = tid; i < m; i += NT) {
    R *xr = Xr + i * ld, *xi = Xi + i * ld;
    // t = x L^-H:  t_j = (x_j - sum_{k<j} t_k conj(L_jk)) / L_jj  (L_jj real)
    for (int j = 0; j < n; ++j) {
      double ar = xr[j], ai = xi[j];
      for (int k = 0; k < j; ++k) {
        const double tr = xr[k], ti = xi[k], lr = Lr[j * ld + k], li = Li[j * ld + k];
        ar -= tr * lr + ti * li;
        ai -= ti * lr - tr * li;
      }
      const double d = Lr[j * ld + j];
      xr[j] = (R)(ar / d);
      xi[j] = (R)(ai / d);
    }
    // w = t L^-1:  w_j = (t_j - sum_{k>j} w_k L_kj) / L_jj
    for (int j = n - 1; j >= 0; --j) {
      double ar = xr[j], ai = xi[j];
      for (int k = j + 1; k < n; ++k) {
        const double wr = xr[k], wi = xi[k], lr = Lr[k * ld + j], li = Li[k * ld + j];
        ar -= wr * lr - wi * li;
        ai -= wr * li + wi * lr;
      }
      const double d = Lr[j * ld + j];
      xr[j] = (R)(ar / d);
      xi[j] = (R)(ai / d);
    }
  }
  __syncthreads();
  for (int e = tid; e < m * n; e += NT) {
    const int i = e / n, j = e - i * n;
    Xb[2 * e] = Xr[i * ld + j];
    Xb[2 * e + 1] = Xi[i * ld + j];
  }
}

constexpr size_t kCsmLdsCap = 160 * 1024 - 4096;

template <typename K>
static int csm_attr(K kern, size_t bytes) {
  if (bytes > kCsmLdsCap) return PSVD_ETOOBIG;
  if (bytes > 65536 && hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)kCsmLdsCap) != hipSuccess)
    return PSVD_ELAUNCH;
  return PSVD_OK;
}

// Per-matrix region of the global-workspace variants, in doubles (256-byte aligned).
static int64_t csm_gstride(size_t bytes) { return (int64_t)((bytes + 255) / 256 * 32); }

// Workspace the kernels need for a batch: none when a matrix's planes fit in LDS.
static size_t csm_ws(int64_t batch, size_t bytes) {
  return bytes > kCsmLdsCap ? (size_t)(batch * csm_gstride(bytes)) * sizeof(double) : 0;
}

// The launch of kernel `lds` with `bytes` of dynamic LDS, or of `gl` over the workspace when the
// planes exceed LDS (without a workspace that is PSVD_ETOOBIG).  `launch(kernel, smem, gws, gstride)`
// issues it.
template <typename KL, typename KG, typename F>
static int csm_launch(KL lds, KG gl, size_t bytes, int64_t batch, void *ws, size_t ws_bytes, F launch) {
  if (bytes > kCsmLdsCap) {
    if (!ws) return PSVD_ETOOBIG;
    if (ws_bytes < csm_ws(batch, bytes)) return PSVD_EINVAL;
    launch(gl, size_t(0), static_cast<double *>(ws), csm_gstride(bytes));
  } else {
    const int rc = csm_attr(lds, bytes);
    if (rc != PSVD_OK) return rc;
    launch(lds, bytes, static_cast<double *>(nullptr), int64_t(0));
  }
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

template <typename R>
static int run_ceigh(int64_t batch, int n, const void *h, void *w, void *v, int32_t *info, double tol, int max_sweeps,
                     void *ws, size_t ws_bytes, hipStream_t st) {
  CeighParams<R> P{static_cast<const R *>(h), static_cast<R *>(w), static_cast<R *>(v), info, n, max_sweeps, tol,
                   nullptr, 0};
  return csm_launch(ceigh_lds_kernel<R, false>, ceigh_lds_kernel<R, true>, ceigh_lds_bytes<R>(n), batch, ws, ws_bytes,
                    [&](auto kern, size_t smem, double *gws, int64_t gs) {
                      P.gws = gws;
                      P.gstride = gs;
                      hipLaunchKernelGGL(kern, dim3((unsigned)batch), dim3(kCsmThreads), smem, st, P);
                    });
}

template <typename R>
static int run_cchol(int64_t batch, int n, const void *g, void *l, int32_t *fail, void *ws, size_t ws_bytes,
                     hipStream_t st) {
  return csm_launch(cchol_lds_kernel<R, false>, cchol_lds_kernel<R, true>, cchol_lds_bytes<R>(n), batch, ws, ws_bytes,
                    [&](auto kern, size_t smem, double *gws, int64_t gs) {
                      hipLaunchKernelGGL(kern, dim3((unsigned)batch), dim3(kCsmThreads), smem, st,
                                         static_cast<const R *>(g), static_cast<R *>(l), fail, n, gws, gs);
                    });
}

template <typename R>
static int run_csolve(int64_t batch, int m, int n, const void *l, void *x, void *ws, size_t ws_bytes, hipStream_t st) {
  return csm_launch(csolve_kernel<R, false>, csolve_kernel<R, true>, csolve_lds_bytes<R>(m, n), batch, ws, ws_bytes,
                    [&](auto kern, size_t smem, double *gws, int64_t gs) {
                      hipLaunchKernelGGL(kern, dim3((unsigned)batch), dim3(kCsmThreads), smem, st,
                                         static_cast<const R *>(l), static_cast<R *>(x), m, n, gws, gs);
                    });
}

}  // namespace psvd

extern "C" {

int psvd_ceigh_lds_smem_bytes(int dtype, int n, size_t *bytes) {
  if (n < 1 || !bytes) return PSVD_EINVAL;
  if (dtype == PSVD_C64) *bytes = psvd::ceigh_lds_bytes<float>(n);
  else if (dtype == PSVD_C128) *bytes = psvd::ceigh_lds_bytes<double>(n);
  else return PSVD_EINVAL;
  return *bytes > psvd::kCsmLdsCap ? PSVD_ETOOBIG : PSVD_OK;
}

size_t psvd_ceigh_workspace_bytes(int dtype, int64_t batch, int n) {
  if (batch < 0 || n < 1) return 0;
  if (dtype == PSVD_C64) return psvd::csm_ws(batch, psvd::ceigh_lds_bytes<float>(n));
  if (dtype == PSVD_C128) return psvd::csm_ws(batch, psvd::ceigh_lds_bytes<double>(n));
  return 0;
}

int psvd_ceigh(int dtype, int64_t batch, int n, const void *h, void *w, void *v, int32_t *info, double tol,
               int max_sweeps, void *workspace, size_t workspace_bytes, psvd_stream_t stream) {
  if (batch < 0 || batch > 0x7fffffff || n < 1 || !h || !w || !v || !info || max_sweeps < 1) return PSVD_EINVAL;
  if (batch == 0) return PSVD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == PSVD_C64)
    return psvd::run_ceigh<float>(batch, n, h, w, v, info, tol, max_sweeps, workspace, workspace_bytes, st);
  if (dtype == PSVD_C128)
    return psvd::run_ceigh<double>(batch, n, h, w, v, info, tol, max_sweeps, workspace, workspace_bytes, st);
  return PSVD_EINVAL;
}

int psvd_ceigh_lds(int dtype, int64_t batch, int n, const void *h, void *w, void *v, int32_t *info, double tol,
                   int max_sweeps, psvd_stream_t stream) {
  return psvd_ceigh(dtype, batch, n, h, w, v, info, tol, max_sweeps, nullptr, 0, stream);
}

size_t psvd_ccholesky_workspace_bytes(int dtype, int64_t batch, int n) {
  if (batch < 0 || n < 1) return 0;
  if (dtype == PSVD_C64) return psvd::csm_ws(batch, psvd::cchol_lds_bytes<float>(n));
  if (dtype == PSVD_C128) return psvd::csm_ws(batch, psvd::cchol_lds_bytes<double>(n));
  return 0;
}

int psvd_ccholesky(int dtype, int64_t batch, int n, const void *g, void *l, int32_t *fail, void *workspace,
                   size_t workspace_bytes, psvd_stream_t stream) {
  if (batch < 0 || batch > 0x7fffffff || n < 1 || !g || !l || !fail) return PSVD_EINVAL;
  if (batch == 0) return PSVD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == PSVD_C64) return psvd::run_cchol<float>(batch, n, g, l, fail, workspace, workspace_bytes, st);
  if (dtype == PSVD_C128) return psvd::run_cchol<double>(batch, n, g, l, fail, workspace, workspace_bytes, st);
  return PSVD_EINVAL;
}

int psvd_ccholesky_lds(int dtype, int64_t batch, int n, const void *g, void *l, int32_t *fail, psvd_stream_t stream) {
  return psvd_ccholesky(dtype, batch, n, g, l, fail, nullptr, 0, stream);
}

int psvd_csolve_lds_smem_bytes(int dtype, int m, int n, size_t *bytes) {
  if (m < 1 || n < 1 || !bytes) return PSVD_EINVAL;
  if (dtype == PSVD_C64) *bytes = psvd::csolve_lds_bytes<float>(m, n);
  else if (dtype == PSVD_C128) *bytes = psvd::csolve_lds_bytes<double>(m, n);
  else return PSVD_EINVAL;
  return *bytes > psvd::kCsmLdsCap ? PSVD_ETOOBIG : PSVD_OK;
}

size_t psvd_csolve_chol_workspace_bytes(int dtype, int64_t batch, int m, int n) {
  if (batch < 0 || m < 1 || n < 1) return 0;
  if (dtype == PSVD_C64) return psvd::csm_ws(batch, psvd::csolve_lds_bytes<float>(m, n));
  if (dtype == PSVD_C128) return psvd::csm_ws(batch, psvd::csolve_lds_bytes<double>(m, n));
  return 0;
}

int psvd_csolve_chol(int dtype, int64_t batch, int m, int n, const void *l, void *x, void *workspace,
                     size_t workspace_bytes, psvd_stream_t stream) {
  if (batch < 0 || batch > 0x7fffffff || m < 1 || n < 1 || !l || !x) return PSVD_EINVAL;
  if (batch == 0) return PSVD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == PSVD_C64) return psvd::run_csolve<float>(batch, m, n, l, x, workspace, workspace_bytes, st);
  if (dtype == PSVD_C128) return psvd::run_csolve<double>(batch, m, n, l, x, workspace, workspace_bytes, st);
  return PSVD_EINVAL;
}

int psvd_csolve_chol_lds(int dtype, int64_t batch, int m, int n, const void *l, void *x, psvd_stream_t stream) {
  return psvd_csolve_chol(dtype, batch, m, n, l, x, nullptr, 0, stream);
}

}  // extern "C"
