// K6: two-sided cyclic Jacobi eigensolver of symmetric matrices, G <- J^T G J with one plane
// rotation per pair of a round-robin round, V accumulated; eigenvalues ascending.
//
// A round is two phases.  (1) One thread per pair reads g_pp, g_qq, g_pq, applies the reference's
// test |g_pq| > tol * max|G| (the maximum over the whole matrix at the start of the round,
// primitives.py:397-399) and forms c, s in FP64 (t = sign(tau) / (|tau| + hypot(tau, 1))).
// (2) Every 2x2 block of G with rows from pair k1 and columns from pair k2 is read once,
// transformed as J_k1^T B J_k2 -- columns first, then rows, the reference's order
// (primitives.py:408-417) -- and written back with its mirror block; the same pass tracks the
// new max|G| for the next round's test.  V is stored transposed so its column rotations are
// contiguous row rotations.
//
// eigh_lds_kernel: G and V^T in LDS, one workgroup per matrix, two barriers per round.
// Larger matrices: G and V^T in HBM, two kernels per round, one host read per sweep.
// Follows svdbench/standalone/primitives.py:370-427 (jacobi_eigh), :309-313 (schedule).
#include <float.h>

#include "common.hpp"

namespace psvd {

template <typename T>
struct EighParams {
  const T *h;
  T *w, *v;
  int32_t *info;
  int n;
  double tol;
  int max_sweeps;
  // HBM version only
  T *G, *Vt;
  T *cs;                      // 2 * npairs
  int *pq;                    // 2 * npairs
  unsigned long long *slot;   // 3 max slots
  int *flag;                  // 3 sweep flags
  double *dw;                 // np (eigenvalues, FP64) for sorting
  int *ord;                   // np
};

__host__ __device__ inline int eigh_ld(int np) { return np + 1; }

template <typename T>
__host__ __device__ inline size_t eigh_lds_bytes(int n) {
  const int np = n + (n & 1), ld = eigh_ld(np);
  size_t b = sizeof(T) * 2 * (size_t)np * ld;       // G, V^T
  b += sizeof(double) * 2 * (size_t)(np / 2);       // c, s
  b += sizeof(int) * 2 * (size_t)(np / 2);          // p, q
  b += sizeof(double) * np + sizeof(int) * np;      // eigenvalues, order
  return (b + 15) & ~size_t(15);
}

// Phase-2 work item: the 2x2 block (pair k1 rows) x (pair k2 cols), k1 <= k2, and its mirror.
template <typename T>
__device__ __forceinline__ double block_update(T *G, int ld, int p1, int q1, int p2, int q2, double c1, double s1,
                                               double c2, double s2, bool act1, bool act2) {
  T *a11 = G + p1 * ld + p2, *a12 = G + p1 * ld + q2, *a21 = G + q1 * ld + p2, *a22 = G + q1 * ld + q2;
  double b11 = (double)*a11, b12 = (double)*a12, b21 = (double)*a21, b22 = (double)*a22;
  if (act2) {  // columns (p2, q2): col_p = c col_p - s col_q, col_q = s col_p + c col_q
    const double t11 = c2 * b11 - s2 * b12, t12 = s2 * b11 + c2 * b12;
    const double t21 = c2 * b21 - s2 * b22, t22 = s2 * b21 + c2 * b22;
    b11 = (double)(T)t11; b12 = (double)(T)t12; b21 = (double)(T)t21; b22 = (double)(T)t22;
  }
  if (act1) {  // rows (p1, q1): row_p = c row_p - s row_q, row_q = s row_p + c row_q
    const double t11 = c1 * b11 - s1 * b21, t21 = s1 * b11 + c1 * b21;
    const double t12 = c1 * b12 - s1 * b22, t22 = s1 * b12 + c1 * b22;
    b11 = t11; b12 = t12; b21 = t21; b22 = t22;
  }
  const T n11 = (T)b11, n12 = (T)b12, n21 = (T)b21, n22 = (T)b22;
  *a11 = n11; *a12 = n12; *a21 = n21; *a22 = n22;
  if (p1 != p2) {  // mirror block (pair k2 rows) x (pair k1 cols) is the transpose
    G[p2 * ld + p1] = n11; G[q2 * ld + p1] = n12; G[p2 * ld + q1] = n21; G[q2 * ld + q1] = n22;
  }
  return fmax(fmax(fabs((double)n11), fabs((double)n12)), fmax(fabs((double)n21), fabs((double)n22)));
}

template <typename T>
__device__ __forceinline__ void pair_params(const T *G, int ld, int p, int q, double tol, double scale, double &c,
                                            double &s, bool &active, double &ratio) {
  const double gpp = (double)G[p * ld + p], gqq = (double)G[q * ld + q], gpq = (double)G[p * ld + q];
  const double a = fabs(gpq);
  ratio = a / fmax(scale, DBL_MIN);
  active = a > tol * scale;
  c = 1.0;
  s = 0.0;
  if (active) jacobi_rotation(gpp, gqq, gpq, c, s);
}

constexpr int kEighThreads = 256;

template <typename T>
__global__ __launch_bounds__(kEighThreads) void eigh_lds_kernel(EighParams<T> P) {
  constexpr int NT = kEighThreads, NW = NT / kWave;
  extern __shared__ double smem_e[];
  __shared__ double s_slot[2 * NW];
  __shared__ int s_flag[2 * NW];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int n = P.n, np = n + (n & 1), ld = eigh_ld(np), npairs = np / 2;
  const int64_t b = blockIdx.x;
  double *cc = smem_e, *ss = cc + npairs, *dw = ss + npairs;
  int *pp = reinterpret_cast<int *>(dw + np), *qq = pp + npairs, *ord = qq + npairs;
  T *G = reinterpret_cast<T *>(ord + np + (np & 1));
  T *Vt = G + np * ld;
  const T *H = P.h + b * (int64_t)n * n;
  const double eps = Eps<T>::value;
  const double tol = P.tol > 0.0 ? P.tol : 10.0 * eps;

  double mx = 0.0;
  for (int e = tid; e < np * np; e += NT) {
    const int i = e / np, j = e - i * np;
    const T x = (i < n && j < n) ? H[(int64_t)i * n + j] : T(0);
    G[i * ld + j] = x;
    Vt[i * ld + j] = i == j ? T(1) : T(0);
    mx = fmax(mx, fabs((double)x));
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
        double c, s, ratio;
        bool act;
        pair_params(G, ld, p, q, tol, scale, c, s, act, ratio);
        cc[k] = act ? c : 1.0;
        ss[k] = act ? s : 0.0;
        pp[k] = p;
        qq[k] = q;
        wave_active |= act;
      }
      __syncthreads();
      double mymax = 0.0;
      const int nblk = npairs * (npairs + 1) / 2;
      for (int e = tid; e < nblk; e += NT) {
        // e -> (k1, k2) with k1 <= k2, row-major over the upper triangle of the pair grid
        int k1 = (int)((2.0 * npairs + 1.0 - sqrt((2.0 * npairs + 1.0) * (2.0 * npairs + 1.0) - 8.0 * e)) / 2.0);
        if (k1 < 0) k1 = 0;
        if (k1 > npairs - 1) k1 = npairs - 1;
        while (k1 > 0 && k1 * npairs - k1 * (k1 - 1) / 2 > e) --k1;
        while (k1 < npairs - 1 && (k1 + 1) * npairs - (k1 + 1) * k1 / 2 <= e) ++k1;
        const int k2 = k1 + (e - (k1 * npairs - k1 * (k1 - 1) / 2));
        const double c1 = cc[k1], s1 = ss[k1], c2 = cc[k2], s2 = ss[k2];
        const bool a1 = s1 != 0.0, a2 = s2 != 0.0;
        if (a1 || a2) {
          mymax = fmax(mymax, block_update(G, ld, pp[k1], qq[k1], pp[k2], qq[k2], c1, s1, c2, s2, a1, a2));
        } else {
          const int p1 = pp[k1], q1 = qq[k1], p2 = pp[k2], q2 = qq[k2];
          mymax = fmax(mymax, fmax(fmax(fabs((double)G[p1 * ld + p2]), fabs((double)G[p1 * ld + q2])),
                                   fmax(fabs((double)G[q1 * ld + p2]), fabs((double)G[q1 * ld + q2]))));
        }
      }
      for (int e = tid; e < npairs * np; e += NT) {
        const int k = e / np, j = e - k * np;
        const double s = ss[k];
        if (s == 0.0) continue;
        const double c = cc[k];
        T *vp = Vt + pp[k] * ld + j, *vq = Vt + qq[k] * ld + j;
        const double x = (double)*vp, y = (double)*vq;
        *vp = (T)(c * x - s * y);
        *vq = (T)(s * x + c * y);
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
  // eigenvalues ascending (ties by index), V columns in the same order
  for (int j = tid; j < n; j += NT) dw[j] = (double)G[j * ld + j];
  __syncthreads();
  for (int j = tid; j < n; j += NT) {
    const double x = dw[j];
    int rank = 0;
    for (int k = 0; k < n; ++k) rank += (dw[k] < x) || (dw[k] == x && k < j);
    ord[rank] = j;
  }
  __syncthreads();
  T *W = P.w + b * (int64_t)n;
  T *Vo = P.v + b * (int64_t)n * n;
  for (int r = tid; r < n; r += NT) W[r] = (T)dw[ord[r]];
  for (int e = tid; e < n * n; e += NT) {
    const int i = e / n, r = e - i * n;
    Vo[e] = Vt[ord[r] * ld + i];
  }
  if (tid == 0) {
    P.info[2 * b] = sweeps;
    P.info[2 * b + 1] = converged ? PSVD_INFO_OK : PSVD_INFO_NOCONV;
  }
}

// ---- multi-launch version for one matrix beyond LDS -------------------------------------
// Each round is two kernels (the kernel boundary is the barrier): eigh_params_kernel forms the
// pair rotations from max|G| left in slot[g & 1] by the previous round and clears the other slot;
// eigh_update_kernel applies the 2x2 block updates and the V^T row rotations and gathers the new
// max|G|.  The host reads the sweep's activity flag once per sweep (the reference also
// synchronises once per sweep, primitives.py:418).
template <typename T>
__global__ __launch_bounds__(kEighThreads) void eigh_load_kernel(EighParams<T> P) {
  const int n = P.n, np = n + (n & 1), ld = (np + 63) & ~63;
  const int64_t gtid = (int64_t)blockIdx.x * kEighThreads + threadIdx.x, gsz = (int64_t)gridDim.x * kEighThreads;
  double mx = 0.0;
  for (int64_t e = gtid; e < (int64_t)np * np; e += gsz) {
    const int i = (int)(e / np), j = (int)(e - (int64_t)i * np);
    const T x = (i < n && j < n) ? P.h[(int64_t)i * n + j] : T(0);
    P.G[(int64_t)i * ld + j] = x;
    P.Vt[(int64_t)i * ld + j] = i == j ? T(1) : T(0);
    mx = fmax(mx, fabs((double)x));
  }
  mx = wave_max_fast(mx);
  if ((threadIdx.x & (kWave - 1)) == 0) atomicMax(&P.slot[0], (unsigned long long)__double_as_longlong(mx));
}

template <typename T>
__global__ __launch_bounds__(kEighThreads) void eigh_params_kernel(EighParams<T> P, int r, int g) {
  const int n = P.n, np = n + (n & 1), ld = (np + 63) & ~63, npairs = np / 2;
  const int64_t k = (int64_t)blockIdx.x * kEighThreads + threadIdx.x;
  const double eps = Eps<T>::value;
  const double tol = P.tol > 0.0 ? P.tol : 10.0 * eps;
  if (k == 0) P.slot[(g + 1) & 1] = 0ull;
  if (k >= npairs) return;
  const double scale = __longlong_as_double((long long)P.slot[g & 1]);
  int p, q;
  rr_pair(r, (int)k, np, p, q);
  double c, s, ratio;
  bool act;
  pair_params(P.G, ld, p, q, tol, scale, c, s, act, ratio);
  P.cs[k] = act ? (T)c : T(1);
  P.cs[npairs + k] = act ? (T)s : T(0);
  P.pq[k] = p;
  P.pq[npairs + k] = q;
  if (act) atomicOr(P.flag, 1);
}

template <typename T>
__global__ __launch_bounds__(kEighThreads) void eigh_update_kernel(EighParams<T> P, int g) {
  const int n = P.n, np = n + (n & 1), ld = (np + 63) & ~63, npairs = np / 2;
  const int64_t gtid = (int64_t)blockIdx.x * kEighThreads + threadIdx.x, gsz = (int64_t)gridDim.x * kEighThreads;
  const T *cc = P.cs, *ss = P.cs + npairs;
  const int *pp = P.pq, *qq = P.pq + npairs;
  T *G = P.G, *Vt = P.Vt;
  double mymax = 0.0;
  const int64_t nblk = (int64_t)npairs * (npairs + 1) / 2;
  for (int64_t e = gtid; e < nblk; e += gsz) {
    int k1 = (int)((2.0 * npairs + 1.0 - sqrt((2.0 * npairs + 1.0) * (2.0 * npairs + 1.0) - 8.0 * (double)e)) / 2.0);
    if (k1 < 0) k1 = 0;
    if (k1 > npairs - 1) k1 = npairs - 1;
    while (k1 > 0 && (int64_t)k1 * npairs - (int64_t)k1 * (k1 - 1) / 2 > e) --k1;
    while (k1 < npairs - 1 && (int64_t)(k1 + 1) * npairs - (int64_t)(k1 + 1) * k1 / 2 <= e) ++k1;
    const int k2 = k1 + (int)(e - ((int64_t)k1 * npairs - (int64_t)k1 * (k1 - 1) / 2));
    const double c1 = (double)cc[k1], s1 = (double)ss[k1], c2 = (double)cc[k2], s2 = (double)ss[k2];
    const bool a1 = s1 != 0.0, a2 = s2 != 0.0;
    const int p1 = pp[k1], q1 = qq[k1], p2 = pp[k2], q2 = qq[k2];
    if (a1 || a2) {
      mymax = fmax(mymax, block_update(G, ld, p1, q1, p2, q2, c1, s1, c2, s2, a1, a2));
    } else {
      mymax = fmax(mymax, fmax(fmax(fabs((double)G[(int64_t)p1 * ld + p2]), fabs((double)G[(int64_t)p1 * ld + q2])),
                               fmax(fabs((double)G[(int64_t)q1 * ld + p2]), fabs((double)G[(int64_t)q1 * ld + q2]))));
    }
  }
  for (int64_t e = gtid; e < (int64_t)npairs * np; e += gsz) {
    const int k = (int)(e / np), j = (int)(e - (int64_t)k * np);
    const double s = (double)ss[k];
    if (s == 0.0) continue;
    const double c = (double)cc[k];
    T *vp = Vt + (int64_t)pp[k] * ld + j, *vq = Vt + (int64_t)qq[k] * ld + j;
    const double x = (double)*vp, y = (double)*vq;
    *vp = (T)(c * x - s * y);
    *vq = (T)(s * x + c * y);
  }
  mymax = wave_max_fast(mymax);
  if ((threadIdx.x & (kWave - 1)) == 0) atomicMax(&P.slot[(g + 1) & 1], (unsigned long long)__double_as_longlong(mymax));
}

template <typename T>
__global__ __launch_bounds__(kEighThreads) void eigh_diag_kernel(EighParams<T> P) {
  const int n = P.n, np = n + (n & 1), ld = (np + 63) & ~63;
  const int64_t j = (int64_t)blockIdx.x * kEighThreads + threadIdx.x;
  if (j < n) P.dw[j] = (double)P.G[j * ld + j];
}

template <typename T>
__global__ __launch_bounds__(kEighThreads) void eigh_sort_kernel(EighParams<T> P) {
  const int n = P.n;
  const int64_t j = (int64_t)blockIdx.x * kEighThreads + threadIdx.x;
  if (j >= n) return;
  const double x = P.dw[j];
  int rank = 0;
  for (int k = 0; k < n; ++k) rank += (P.dw[k] < x) || (P.dw[k] == x && k < j);
  P.ord[rank] = (int)j;
}

template <typename T>
__global__ __launch_bounds__(kEighThreads) void eigh_out_kernel(EighParams<T> P, int sweeps) {
  const int n = P.n, np = n + (n & 1), ld = (np + 63) & ~63;
  const int64_t gtid = (int64_t)blockIdx.x * kEighThreads + threadIdx.x, gsz = (int64_t)gridDim.x * kEighThreads;
  for (int64_t r = gtid; r < n; r += gsz) P.w[r] = (T)P.dw[P.ord[r]];
  for (int64_t e = gtid; e < (int64_t)n * n; e += gsz) {
    const int i = (int)(e / n), r = (int)(e - (int64_t)i * n);
    P.v[e] = P.Vt[(int64_t)P.ord[r] * ld + i];
  }
  if (gtid == 0) {
    P.info[0] = sweeps;
    P.info[1] = *P.flag ? PSVD_INFO_NOCONV : PSVD_INFO_OK;
  }
}

__global__ void eigh_reset_kernel(unsigned long long *slot, int *flag, int what) {
  if (threadIdx.x == 0) {
    if (what & 1) {
      slot[0] = 0ull;
      slot[1] = 0ull;
    }
    if (what & 2) *flag = 0;
  }
}

constexpr size_t kEighLdsCap = 160 * 1024 - 4096;

template <typename T>
size_t eigh_grid_ws_bytes(int n) {
  const int np = n + (n & 1), ld = (np + 63) & ~63, npairs = np / 2;
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  return al(sizeof(T) * (size_t)np * ld) * 2 + al(sizeof(T) * 2 * npairs) + al(sizeof(int) * 2 * npairs) +
         al(8 * 3) + al(4 * 3) + al(sizeof(double) * np) + al(sizeof(int) * np);
}

template <typename T>
static int run_eigh(int64_t batch, int n, const T *h, T *w, T *v, int32_t *info, double tol, int max_sweeps, void *ws,
                    size_t ws_bytes, hipStream_t st) {
  EighParams<T> P{};
  P.h = h;
  P.w = w;
  P.v = v;
  P.info = info;
  P.n = n;
  P.tol = tol;
  P.max_sweeps = max_sweeps;
  const size_t smem = eigh_lds_bytes<T>(n);
  if (smem <= kEighLdsCap) {
    static bool attr = false;
    if (!attr) {
      if (hipFuncSetAttribute((const void *)eigh_lds_kernel<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kEighLdsCap) != hipSuccess)
        return PSVD_ELAUNCH;
      attr = true;
    }
    hipLaunchKernelGGL(eigh_lds_kernel<T>, dim3((unsigned)batch), dim3(kEighThreads), smem, st, P);
    return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
  }
  if (ws_bytes < eigh_grid_ws_bytes<T>(n)) return PSVD_EINVAL;
  const int np = n + (n & 1), ld = (np + 63) & ~63, npairs = np / 2;
  char *p = static_cast<char *>(ws);
  auto take = [&](size_t bytes) {
    char *q = p;
    p += (bytes + 255) & ~size_t(255);
    return q;
  };
  P.G = reinterpret_cast<T *>(take(sizeof(T) * (size_t)np * ld));
  P.Vt = reinterpret_cast<T *>(take(sizeof(T) * (size_t)np * ld));
  P.cs = reinterpret_cast<T *>(take(sizeof(T) * 2 * npairs));
  P.pq = reinterpret_cast<int *>(take(sizeof(int) * 2 * npairs));
  P.slot = reinterpret_cast<unsigned long long *>(take(8 * 3));
  P.flag = reinterpret_cast<int *>(take(4 * 3));
  P.dw = reinterpret_cast<double *>(take(sizeof(double) * np));
  P.ord = reinterpret_cast<int *>(take(sizeof(int) * np));
  auto blocks = [](int64_t items) { return (unsigned)((items + kEighThreads - 1) / kEighThreads); };
  const unsigned gupd = blocks((int64_t)npairs * (npairs + 1) / 2 > (int64_t)npairs * np ? (int64_t)npairs * (npairs + 1) / 2
                                                                                          : (int64_t)npairs * np);
  const unsigned gload = blocks((int64_t)np * np) < 4096 ? blocks((int64_t)np * np) : 4096;
  for (int64_t b = 0; b < batch; ++b) {
    EighParams<T> Q = P;
    Q.h = h + b * (int64_t)n * n;
    Q.w = w + b * (int64_t)n;
    Q.v = v + b * (int64_t)n * n;
    Q.info = info + 2 * b;
    hipLaunchKernelGGL(eigh_reset_kernel, dim3(1), dim3(64), 0, st, Q.slot, Q.flag, 3);
    hipLaunchKernelGGL(eigh_load_kernel<T>, dim3(gload), dim3(kEighThreads), 0, st, Q);
    int sweeps = 0, g = 0;
    for (int sw = 1; sw <= max_sweeps; ++sw) {
      sweeps = sw;
      hipLaunchKernelGGL(eigh_reset_kernel, dim3(1), dim3(64), 0, st, Q.slot, Q.flag, 2);
      for (int r = 0; r < np - 1; ++r, ++g) {
        hipLaunchKernelGGL(eigh_params_kernel<T>, dim3(blocks(npairs)), dim3(kEighThreads), 0, st, Q, r, g);
        hipLaunchKernelGGL(eigh_update_kernel<T>, dim3(gupd), dim3(kEighThreads), 0, st, Q, g);
      }
      if (hipGetLastError() != hipSuccess) return PSVD_ELAUNCH;
      int active = 0;
      if (hipMemcpyAsync(&active, Q.flag, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess) return PSVD_ELAUNCH;
      if (hipStreamSynchronize(st) != hipSuccess) return PSVD_ELAUNCH;
      if (!active) break;
    }
    hipLaunchKernelGGL(eigh_diag_kernel<T>, dim3(blocks(n)), dim3(kEighThreads), 0, st, Q);
    hipLaunchKernelGGL(eigh_sort_kernel<T>, dim3(blocks(n)), dim3(kEighThreads), 0, st, Q);
    hipLaunchKernelGGL(eigh_out_kernel<T>, dim3(blocks((int64_t)n * n) < 4096 ? blocks((int64_t)n * n) : 4096),
                       dim3(kEighThreads), 0, st, Q, sweeps);
    if (hipGetLastError() != hipSuccess) return PSVD_ELAUNCH;
  }
  return PSVD_OK;
}

}  // namespace psvd

extern "C" {

size_t psvd_eigh_workspace_bytes(int dtype, int n) {
  if (n < 1) return 0;
  if (dtype == PSVD_F32) return psvd::eigh_grid_ws_bytes<float>(n);
  if (dtype == PSVD_F64) return psvd::eigh_grid_ws_bytes<double>(n);
  return 0;
}

int psvd_eigh(int dtype, int64_t batch, int n, const void *h, void *w, void *v, int32_t *info, double tol,
              int max_sweeps, void *workspace, size_t workspace_bytes, psvd_stream_t stream) {
  if (batch < 0 || n < 1 || max_sweeps < 1 || !h || !w || !v || !info || batch > 0x7fffffff) return PSVD_EINVAL;
  if (batch == 0) return PSVD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == PSVD_F32)
    return psvd::run_eigh<float>(batch, n, static_cast<const float *>(h), static_cast<float *>(w),
                                 static_cast<float *>(v), info, tol, max_sweeps, workspace, workspace_bytes, st);
  if (dtype == PSVD_F64)
    return psvd::run_eigh<double>(batch, n, static_cast<const double *>(h), static_cast<double *>(w),
                                  static_cast<double *>(v), info, tol, max_sweeps, workspace, workspace_bytes, st);
  return PSVD_EINVAL;
}

}  // extern "C"
