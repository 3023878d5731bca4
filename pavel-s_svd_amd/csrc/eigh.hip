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
  unsigned long long *slot;   // max|G| slots of kSub words
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

template <typename T>
__device__ __forceinline__ void pair_params(const T *G, int ld, int p, int q, double tol, double scale, double &c,
                                            double &s, bool &active, double &ratio, bool fast = false) {
  const double gpp = (double)G[p * ld + p], gqq = (double)G[q * ld + q], gpq = (double)G[p * ld + q];
  const double a = fabs(gpq);
  ratio = a / fmax(scale, DBL_MIN);
  active = a > tol * scale;
  c = 1.0;
  s = 0.0;
  // fast: the FP32 angle of jacobi_rotation_fast (blocked solver; the kernels that keep the
  // reference's rotation order and sweep counts use the FP64 formula)
  if (active) {
    if (fast)
      jacobi_rotation_fast<T>(gpp, gqq, gpq, c, s);
    else
      jacobi_rotation(gpp, gqq, gpq, c, s);
  }
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

// Each rotating max|G| slot of the HBM solvers is spread over kSub words 256 bytes apart: the
// thousands of workgroups of a round add their maxima to kSub different L2 lines instead of
// queueing on one, and a slot's value is the largest of its words.
constexpr int kSub = 16, kSubLd = 32;

__device__ __forceinline__ unsigned long long *sub_slot(unsigned long long *s, int which, int k) {
  return s + ((size_t)which * kSub + k) * kSubLd;
}

__device__ __forceinline__ double read_slot(const unsigned long long *s, int which) {
  unsigned long long v = 0ull;
#pragma unroll
  for (int k = 0; k < kSub; ++k) v = v > s[((size_t)which * kSub + k) * kSubLd] ? v : s[((size_t)which * kSub + k) * kSubLd];
  return __longlong_as_double((long long)v);
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
  if ((threadIdx.x & (kWave - 1)) == 0)
    atomicMax(sub_slot(P.slot, 0, (int)((blockIdx.x * (kEighThreads / kWave) + threadIdx.x / kWave) % kSub)),
              (unsigned long long)__double_as_longlong(mx));
}

template <typename T>
__global__ __launch_bounds__(kEighThreads) void eigh_params_kernel(EighParams<T> P, int r, int g) {
  const int n = P.n, np = n + (n & 1), ld = (np + 63) & ~63, npairs = np / 2;
  const int64_t k = (int64_t)blockIdx.x * kEighThreads + threadIdx.x;
  const double eps = Eps<T>::value;
  const double tol = P.tol > 0.0 ? P.tol : 10.0 * eps;
  if (k < kSub) *sub_slot(P.slot, (g + 1) & 1, (int)k) = 0ull;
  bool act = false;
  if (k < npairs) {
    const double scale = read_slot(P.slot, g & 1);
    int p, q;
    rr_pair(r, (int)k, np, p, q);
    double c, s, ratio;
    pair_params(P.G, ld, p, q, tol, scale, c, s, act, ratio);
    P.cs[k] = act ? (T)c : T(1);
    P.cs[npairs + k] = act ? (T)s : T(0);
    P.pq[k] = p;
    P.pq[npairs + k] = q;
  }
  // one flag update per wave with a rotation
  if (__any(act) && (threadIdx.x & (kWave - 1)) == 0) atomicOr(P.flag, 1);
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
  if ((threadIdx.x & (kWave - 1)) == 0)
    atomicMax(sub_slot(P.slot, (g + 1) & 1, (int)((blockIdx.x * (kEighThreads / kWave) + threadIdx.x / kWave) % kSub)),
              (unsigned long long)__double_as_longlong(mymax));
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
  const int t = threadIdx.x;   // 64 threads: one per word of the 2 x kSub slot words
  if ((what & 1) && t < 2 * kSub) slot[(size_t)t * kSubLd] = 0ull;
  if ((what & 2) && t == 0) *flag = 0;
}

constexpr size_t kEighLdsCap = 160 * 1024 - 4096;

template <typename T>
size_t eigh_grid_ws_bytes(int n) {
  const int np = n + (n & 1), ld = (np + 63) & ~63, npairs = np / 2;
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  return al(sizeof(T) * (size_t)np * ld) * 2 + al(sizeof(T) * 2 * npairs) + al(sizeof(int) * 2 * npairs) +
         al(8 * 3 * kSub * kSubLd) + al(4 * 3) + al(sizeof(double) * np) + al(sizeof(int) * np);
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
  P.slot = reinterpret_cast<unsigned long long *>(take(8 * 3 * kSub * kSubLd));
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

// ---- blocked two-sided Jacobi for matrices beyond LDS --------------------------------------
// Indices are grouped in blocks of kEb = 32 (n padded to a multiple of 64 with decoupled zero
// rows and columns that never rotate); round r of the tournament over the nb blocks pairs block
// I with block J, and the 64 x 64 subproblem G[I u J, I u J] is swept in LDS by `inner_sweeps`
// cyclic sweeps of the scalar path's plane rotations and test (|g_pq| > tol * max|G|, max|G| of
// the whole matrix at the start of the round, primitives.py:397-399).  The subproblem's
// accumulated rotation Q_p is then applied to the rest of G -- block (p, q) <- Q_p^T B Q_q and
// its mirror -- and to the rows of V^T, as 64 x 64 x 64 MFMA products (the V^T products of a
// round run in the next round's first launch, beside its subproblem sweeps).  A sweep in which no
// subproblem rotates leaves G unchanged, so every off-diagonal |g_ij| <= tol * max|G| at the end:
// the reference's stopping rule (primitives.py:418).  max|G| is kept per 32 x 32 block so that
// blocks no rotation touches still contribute to the next round's threshold without being read.

constexpr int kEb = 32, kEs = 2 * kEb, kEsLd = kEs + 1;

typedef float e4f __attribute__((ext_vector_type(4)));
typedef double e4d __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<float> {
  using V = e4f;
  static __device__ __forceinline__ V op(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  // D register r of lane l holds row 4 * (l / 16) + r, column l % 16
  static __device__ __forceinline__ int row(int r, int lk) { return 4 * lk + r; }
};
template <> struct Mfma<double> {
  using V = e4d;
  static __device__ __forceinline__ V op(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  // D register r of lane l holds row 4 * r + l / 16, column l % 16
  static __device__ __forceinline__ int row(int r, int lk) { return 4 * r + lk; }
};

template <typename T>
struct BlkParams {
  const T *h;
  int n, N, nb, inner_sweeps;
  double tol;
  T *G, *Vt, *Q;               // G, V^T: N x N (ld N); Q: nb/2 matrices Q_p^T, 64 x 64 row-major
  int *act;                    // nb/2: subproblem p rotated in this round
  double *bm;                  // nb x nb: max |G| of each 32 x 32 block
  unsigned long long *slot;    // 3 rotating max|G| slots of kSub words (bit patterns of non-negative doubles)
  int *flag;                   // any rotation in the current sweep
};

__device__ __forceinline__ int blk_row(int I, int J, int i) { return i < kEb ? I * kEb + i : J * kEb + (i - kEb); }

__device__ __forceinline__ void slot_max(unsigned long long *s, double v) {
  atomicMax(s, (unsigned long long)__double_as_longlong(v));
}


// upper-triangle index e of an np2 x np2 grid (k1 <= k2, row-major) -> (k1, k2)

template <typename T>
__global__ __launch_bounds__(256) void blk_load_kernel(BlkParams<T> P) {
  __shared__ double red[4];
  const int I = blockIdx.y, K = blockIdx.x, tid = threadIdx.x;
  double mx = 0.0;
  for (int e = tid; e < kEb * kEb; e += 256) {
    const int i = I * kEb + e / kEb, k = K * kEb + (e & (kEb - 1));
    const T x = (i < P.n && k < P.n) ? P.h[(int64_t)i * P.n + k] : T(0);
    P.G[(int64_t)i * P.N + k] = x;
    P.Vt[(int64_t)i * P.N + k] = i == k ? T(1) : T(0);
    mx = fmax(mx, fabs((double)x));
  }
  mx = block_max<256>(mx, red);
  if (tid == 0) {
    P.bm[I * P.nb + K] = mx;
    slot_max(sub_slot(P.slot, 0, (int)((blockIdx.y * gridDim.x + blockIdx.x) % kSub)), mx);
  }
}

__global__ void blk_reset_kernel(unsigned long long *slot, int *flag, int what) {
  const int t = threadIdx.x;   // 64 threads: one per word of the 3 x kSub slot words
  if ((what & 1) && t < 3 * kSub) slot[(size_t)t * kSubLd] = 0ull;
  if ((what & 2) && t == 0) *flag = 0;
}

template <typename T>
__host__ __device__ inline size_t blk_inner_lds() {
  return sizeof(double) * 2 * kEb + sizeof(int) * 2 * kEb + sizeof(T) * 2 * (size_t)kEs * kEsLd;
}

// 576 threads: the 528 upper-triangle 2 x 2 blocks of a round are one item per thread, so no
// thread carries a second block on the round's critical path (512 threads left 16 with two)
constexpr int kBlkInner = 576;

// Rotations Q_p^T and activity flags of round g live in buffer g & 1: the launch that sweeps the
// subproblems of round g also applies round g - 1's rotations to V^T (see blk_round_kernel).
template <typename T>
__device__ __forceinline__ T *blk_qbuf(const BlkParams<T> &P, int g) {
  return P.Q + (size_t)(g & 1) * (P.nb / 2) * kEs * kEs;
}

template <typename T>
__device__ __forceinline__ int *blk_actbuf(const BlkParams<T> &P, int g) {
  return P.act + (g & 1) * (P.nb / 2);
}

template <typename T>
__device__ __forceinline__ void blk_inner_body(const BlkParams<T> &P, int p, int r, int g, double *smem_bi) {
  double2 *cs = reinterpret_cast<double2 *>(smem_bi);   // (c, s) of pair k of the round
  int2 *pq = reinterpret_cast<int2 *>(cs + kEb);          // (p, q) of pair k
  T *S = reinterpret_cast<T *>(pq + kEb);
  T *Qt = S + kEs * kEsLd;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  int I, J;
  rr_pair(r, p, P.nb, I, J);
  if (p == 0 && tid < kSub) *sub_slot(P.slot, (g + 2) % 3, tid) = 0ull;
  const double scale = read_slot(P.slot, g % 3);
  const double tol = P.tol > 0.0 ? P.tol : 10.0 * Eps<T>::value;
  double offmax = 0.0;
  for (int e = tid; e < kEs * kEs; e += kBlkInner) {
    const int i = e / kEs, j = e - i * kEs;
    const T x = P.G[(int64_t)blk_row(I, J, i) * P.N + blk_row(I, J, j)];
    S[i * kEsLd + j] = x;
    Qt[i * kEsLd + j] = i == j ? T(1) : T(0);
    if (i != j) offmax = fmax(offmax, fabs((double)x));
  }
  // a subproblem whose off-diagonal is already below the threshold rotates nowhere in its sweep
  // (every pair test fails and nothing changes): skip it
  const int live = __syncthreads_or(offmax > tol * scale);
  int any = 0;
  // this thread's items of the round, fixed for the launch: upper-triangle blocks (k1 <= k2) of
  // the 32 x 32 pair grid (528: one or two per thread) and four Qt rows pairs at column tid % 64
  constexpr int NTRI = kEb * (kEb + 1) / 2, NIT = (NTRI + kBlkInner - 1) / kBlkInner;
  int it1[NIT], it2[NIT];
#pragma unroll
  for (int j = 0; j < NIT; ++j) {
    const int e = tid + j * kBlkInner;
    it1[j] = it2[j] = -1;
    if (e < NTRI) tri_decode(e, kEb, it1[j], it2[j]);
  }
  constexpr int QSTEP = kBlkInner / kEs, QIT = (kEb + QSTEP - 1) / QSTEP;   // Qt row pairs per column
  const int qcol = tid % kEs, qk = tid / kEs;
  // the sweep's first tournament round treats every pair of the subproblem, the others only the
  // pairs across its two blocks (the pairs inside a block were treated in the first round)
  const bool cross = r > 0;
  const int nrr = cross ? kEb : kEs - 1;
  for (int sw = 0; live && sw < P.inner_sweeps; ++sw) {
    int act_sw = 0;
    for (int rr = 0; rr < nrr; ++rr) {
      if (tid < kEb) {
        int a, b;
        sub_pair(cross, rr, tid, kEb, a, b);
        double c, s, ratio;
        bool act;
        pair_params(S, kEsLd, a, b, tol, scale, c, s, act, ratio, true);
        cs[tid] = make_double2(act ? c : 1.0, act ? s : 0.0);
        pq[tid] = make_int2(a, b);
        act_sw |= act;
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < NIT; ++j) {
        if (it1[j] < 0) continue;
        const double2 r1 = cs[it1[j]], r2 = cs[it2[j]];
        if (r1.y == 0.0 && r2.y == 0.0) continue;
        const int2 a1 = pq[it1[j]], a2 = pq[it2[j]];
        block_update(S, kEsLd, a1.x, a1.y, a2.x, a2.y, r1.x, r1.y, r2.x, r2.y, r1.y != 0.0, r2.y != 0.0);
      }
      rotate_rows_batched<T, kEb, QSTEP, QIT>(Qt, kEsLd, cs, pq, qk, qcol);
      __syncthreads();
    }
    act_sw = __syncthreads_or(act_sw);
    any |= act_sw;
    if (!act_sw) break;
  }
  if (any) {
    T *Qg = blk_qbuf(P, g) + (int64_t)p * kEs * kEs;
    for (int e = tid; e < kEs * kEs; e += kBlkInner) {
      const int i = e / kEs, j = e - i * kEs;
      P.G[(int64_t)blk_row(I, J, i) * P.N + blk_row(I, J, j)] = S[i * kEsLd + j];
      Qg[e] = Qt[i * kEsLd + j];
    }
  }
  // maxima of the four 32 x 32 blocks, one per wave of the first four
  if (wave < 4) {
    const int bi = wave >> 1, bj = wave & 1;
    double mx = 0.0;
    for (int e = lane; e < kEb * kEb; e += kWave) {
      const int i = e / kEb, j = e & (kEb - 1);
      mx = fmax(mx, fabs((double)S[(bi * kEb + i) * kEsLd + bj * kEb + j]));
    }
    mx = wave_max_fast(mx);
    if (lane == 0) {
      P.bm[(bi ? J : I) * P.nb + (bj ? J : I)] = mx;
      slot_max(sub_slot(P.slot, (g + 1) % 3, (p * 4 + wave) % kSub), mx);
    }
  }
  if (tid == 0) {
    blk_actbuf(P, g)[p] = any;
    if (any) atomicOr(P.flag, 1);
  }
}

// acc = A B on a 64 x 64 x 64 tile from LDS (ld kEsLd); BT: B is given transposed (B[k][j] at
// Bs[j][k]).  Wave w computes the 32 x 32 quadrant (w / 2, w % 2) as 2 x 2 MFMA tiles.
template <typename T, bool BT>
__device__ __forceinline__ void mm64(const T *As, const T *Bs, typename Mfma<T>::V (&acc)[2][2], int wm, int wn,
                                     int lr, int lk) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = 0;
#pragma unroll 4
  for (int k = 0; k < kEs; k += 4) {
    T a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = As[(wm + 16 * i + lr) * kEsLd + k + lk];
#pragma unroll
    for (int j = 0; j < 2; ++j)
      b[j] = BT ? Bs[(wn + 16 * j + lr) * kEsLd + k + lk] : Bs[(k + lk) * kEsLd + wn + 16 * j + lr];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = Mfma<T>::op(a[i], b[j], acc[i][j]);
  }
}

template <typename T>
__device__ __forceinline__ void store64(T *Ds, const typename Mfma<T>::V (&acc)[2][2], int wm, int wn, int lr, int lk) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) Ds[(wm + 16 * i + Mfma<T>::row(r, lk)) * kEsLd + wn + 16 * j + lr] = (T)acc[i][j][r];
}

template <typename T>
__host__ __device__ inline size_t blk_update_lds() {
  return sizeof(T) * 2 * (size_t)kEs * kEsLd;   // the tile and one rotation at a time
}

// Block e: G block (p, q), p < q (the e-th of the strictly upper pair grid), and its mirror.
// Global loads and stores move 16 bytes per lane (every tile row is made of 32-element segments that start on a 128-byte boundary: N and
// the block offsets are multiples of 32 elements, the buffers 256-byte aligned); LDS tiles keep
// the odd stride kEsLd of the MFMA products.
template <typename T>
__global__ __launch_bounds__(256) void blk_update_kernel(BlkParams<T> P, int r, int g) {
  extern __shared__ double smem_bu[];
  T *Bs = reinterpret_cast<T *>(smem_bu);
  T *Qa = Bs + kEs * kEsLd;   // Q_p, then Q_q: two tiles of LDS (twice the workgroups per CU in FP64)
  using V16 = Vec16<T>;
  constexpr int VE = V16::n, RV = kEs / VE;   // elements per vector, vectors per tile row
  const int npair = P.nb / 2;
  const int64_t nG = (int64_t)npair * (npair - 1) / 2;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave, lr = lane & 15, lk = lane >> 4;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int64_t bid = blockIdx.x;
  typename Mfma<T>::V acc[2][2];
  if (bid < nG) {
    // strictly upper triangle of the pair grid: e -> (p, q), p < q
    int p, q;
    tri_decode(bid, npair - 1, p, q);
    q += 1;
    int Ip, Jp, Iq, Jq;
    rr_pair(r, p, P.nb, Ip, Jp);
    rr_pair(r, q, P.nb, Iq, Jq);
    const int *act = blk_actbuf(P, g);
    const int ap = act[p], aq = act[q];
    if (!ap && !aq) {
      if (tid < 4) {
        const int X = (tid >> 1) ? Jp : Ip, Y = (tid & 1) ? Jq : Iq;
        slot_max(sub_slot(P.slot, (g + 1) % 3, (int)((bid * 4 + tid) % kSub)), P.bm[X * P.nb + Y]);
      }
      return;
    }
    const V16 *Qp = reinterpret_cast<const V16 *>(blk_qbuf(P, g) + (int64_t)p * kEs * kEs);
    const V16 *Qq = reinterpret_cast<const V16 *>(blk_qbuf(P, g) + (int64_t)q * kEs * kEs);
    // the tile with Q_p (or, when only q rotated, Q_q) beside it
    const V16 *Q1 = ap ? Qp : Qq;
    for (int e = tid; e < kEs * RV; e += 256) {
      const int i = e / RV, j = (e - i * RV) * VE;
      const V16 x = *reinterpret_cast<const V16 *>(P.G + (int64_t)blk_row(Ip, Jp, i) * P.N + blk_row(Iq, Jq, j));
      const V16 y = Q1[e];
#pragma unroll
      for (int u = 0; u < VE; ++u) {
        Bs[i * kEsLd + j + u] = x.v[u];
        Qa[i * kEsLd + j + u] = y.v[u];
      }
    }
    __syncthreads();
    if (ap) {  // B <- Q_p^T B
      mm64<T, false>(Qa, Bs, acc, wm, wn, lr, lk);
      __syncthreads();   // every wave has read Q_p and B
      store64<T>(Bs, acc, wm, wn, lr, lk);
      if (aq)
        for (int e = tid; e < kEs * RV; e += 256) {   // Q_q into Q_p's place
          const int i = e / RV, j = (e - i * RV) * VE;
          const V16 y = Qq[e];
#pragma unroll
          for (int u = 0; u < VE; ++u) Qa[i * kEsLd + j + u] = y.v[u];
        }
      __syncthreads();
    }
    if (aq) {  // B <- B Q_q
      mm64<T, true>(Bs, Qa, acc, wm, wn, lr, lk);
      __syncthreads();
      store64<T>(Bs, acc, wm, wn, lr, lk);
      __syncthreads();
    }
    for (int e = tid; e < kEs * RV; e += 256) {
      const int i = e / RV, j = (e - i * RV) * VE;
      V16 x, y;
#pragma unroll
      for (int u = 0; u < VE; ++u) {
        x.v[u] = Bs[i * kEsLd + j + u];
        y.v[u] = Bs[(j + u) * kEsLd + i];
      }
      *reinterpret_cast<V16 *>(P.G + (int64_t)blk_row(Ip, Jp, i) * P.N + blk_row(Iq, Jq, j)) = x;
      *reinterpret_cast<V16 *>(P.G + (int64_t)blk_row(Iq, Jq, i) * P.N + blk_row(Ip, Jp, j)) = y;
    }
    double mx = 0.0;
    for (int e = lane; e < 32 * 32; e += kWave) mx = fmax(mx, fabs((double)Bs[(wm + (e >> 5)) * kEsLd + wn + (e & 31)]));
    mx = wave_max_fast(mx);
    if (lane == 0) {
      const int X = wm ? Jp : Ip, Y = wn ? Jq : Iq;
      P.bm[X * P.nb + Y] = mx;
      P.bm[Y * P.nb + X] = mx;
      slot_max(sub_slot(P.slot, (g + 1) % 3, (int)((bid * 4 + wave) % kSub)), mx);
    }
  }
}

// V^T tile t (rows of pair p of round r, 64-column chunk c) <- Q_p^T tile, for the rotations of
// round g (buffer g & 1), on a workgroup of kBlkInner threads: all threads move the tile, the
// first four waves form the product.
template <typename T>
__device__ __forceinline__ void blk_vt_body(const BlkParams<T> &P, int64_t t, int r, int g, double *smem) {
  T *Bs = reinterpret_cast<T *>(smem);
  T *Qa = Bs + kEs * kEsLd;
  using V16 = Vec16<T>;
  constexpr int VE = V16::n, RV = kEs / VE;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave, lr = lane & 15, lk = lane >> 4;
  const int wm = ((wave & 3) >> 1) * 32, wn = (wave & 1) * 32;
  const int64_t chunks = P.N / kEs;
  const int p = (int)(t / chunks), c0 = (int)(t - (int64_t)p * chunks) * kEs;
  if (!blk_actbuf(P, g)[p]) return;
  int Ip, Jp;
  rr_pair(r, p, P.nb, Ip, Jp);
  const V16 *Qp = reinterpret_cast<const V16 *>(blk_qbuf(P, g) + (int64_t)p * kEs * kEs);
  for (int e = tid; e < kEs * RV; e += kBlkInner) {
    const int i = e / RV, j = (e - i * RV) * VE;
    const V16 x = *reinterpret_cast<const V16 *>(P.Vt + (int64_t)blk_row(Ip, Jp, i) * P.N + c0 + j);
    const V16 y = Qp[e];
#pragma unroll
    for (int u = 0; u < VE; ++u) {
      Bs[i * kEsLd + j + u] = x.v[u];
      Qa[i * kEsLd + j + u] = y.v[u];
    }
  }
  __syncthreads();
  typename Mfma<T>::V acc[2][2];
  if (wave < 4) mm64<T, false>(Qa, Bs, acc, wm, wn, lr, lk);
  __syncthreads();
  if (wave < 4) store64<T>(Bs, acc, wm, wn, lr, lk);
  __syncthreads();
  for (int e = tid; e < kEs * RV; e += kBlkInner) {
    const int i = e / RV, j = (e - i * RV) * VE;
    V16 x;
#pragma unroll
    for (int u = 0; u < VE; ++u) x.v[u] = Bs[i * kEsLd + j + u];
    *reinterpret_cast<V16 *>(P.Vt + (int64_t)blk_row(Ip, Jp, i) * P.N + c0 + j) = x;
  }
}

// One tournament round's first launch: workgroups [0, ninner) sweep the 64 x 64 subproblems of
// round (r, g); the next npair * N / 64 workgroups (when vg >= 0) apply round (vr, vg)'s rotations
// to V^T.  V^T is read by no other kernel of the sweep, so its update for round g - 1 runs beside
// the subproblem sweeps of round g, which occupy only nb / 2 workgroups.
template <typename T>
__global__ __launch_bounds__(kBlkInner) void blk_round_kernel(BlkParams<T> P, int r, int g, int ninner, int vr,
                                                              int vg) {
  extern __shared__ __align__(16) double smem_br[];
  if ((int)blockIdx.x < ninner)
    blk_inner_body<T>(P, (int)blockIdx.x, r, g, smem_br);
  else
    blk_vt_body<T>(P, (int64_t)blockIdx.x - ninner, vr, vg, smem_br);
}

template <typename T>
size_t blk_ws_bytes(int n) {
  const int N = (n + kEs - 1) / kEs * kEs, nb = N / kEb, npair = nb / 2;
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  return 2 * al(sizeof(T) * (size_t)N * N) + al(2 * sizeof(T) * (size_t)npair * kEs * kEs) + al(2 * sizeof(int) * npair) +
         al(sizeof(double) * (size_t)nb * nb) + al(8 * 3 * kSub * kSubLd) + al(4 * 3) + al(sizeof(double) * N) + al(sizeof(int) * N);
}

template <typename T>
static int run_eigh_blocked(int64_t batch, int n, const T *h, T *w, T *v, int32_t *info, double tol, int max_sweeps,
                            int inner_sweeps, void *ws, size_t ws_bytes, hipStream_t st) {
  if (ws_bytes < blk_ws_bytes<T>(n)) return PSVD_EINVAL;
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute((const void *)blk_round_kernel<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)blk_inner_lds<T>()) != hipSuccess ||
        hipFuncSetAttribute((const void *)blk_update_kernel<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)blk_update_lds<T>()) != hipSuccess)
      return PSVD_ELAUNCH;
    attr = true;
  }
  BlkParams<T> P{};
  P.n = n;
  P.N = (n + kEs - 1) / kEs * kEs;
  P.nb = P.N / kEb;
  P.inner_sweeps = inner_sweeps;
  P.tol = tol;
  const int npair = P.nb / 2;
  char *cur = static_cast<char *>(ws);
  auto take = [&](size_t bytes) {
    char *q = cur;
    cur += (bytes + 255) & ~size_t(255);
    return q;
  };
  P.G = reinterpret_cast<T *>(take(sizeof(T) * (size_t)P.N * P.N));
  P.Vt = reinterpret_cast<T *>(take(sizeof(T) * (size_t)P.N * P.N));
  P.Q = reinterpret_cast<T *>(take(2 * sizeof(T) * (size_t)npair * kEs * kEs));
  P.act = reinterpret_cast<int *>(take(2 * sizeof(int) * npair));
  P.bm = reinterpret_cast<double *>(take(sizeof(double) * (size_t)P.nb * P.nb));
  P.slot = reinterpret_cast<unsigned long long *>(take(8 * 3 * kSub * kSubLd));
  P.flag = reinterpret_cast<int *>(take(4 * 3));
  double *dw = reinterpret_cast<double *>(take(sizeof(double) * P.N));
  int *ord = reinterpret_cast<int *>(take(sizeof(int) * P.N));
  const int64_t nG = (int64_t)npair * (npair - 1) / 2;
  const unsigned nvt = (unsigned)((int64_t)npair * (P.N / kEs));
  auto blocks = [](int64_t items) { return (unsigned)((items + kEighThreads - 1) / kEighThreads); };
  for (int64_t b = 0; b < batch; ++b) {
    P.h = h + b * (int64_t)n * n;
    hipLaunchKernelGGL(blk_reset_kernel, dim3(1), dim3(64), 0, st, P.slot, P.flag, 3);
    hipLaunchKernelGGL(blk_load_kernel<T>, dim3(P.nb, P.nb), dim3(256), 0, st, P);
    int sweeps = 0, g = 0, vr = -1;
    for (int sw = 1; sw <= max_sweeps; ++sw) {
      sweeps = sw;
      hipLaunchKernelGGL(blk_reset_kernel, dim3(1), dim3(64), 0, st, P.slot, P.flag, 2);
      for (int r = 0; r < P.nb - 1; ++r, ++g) {
        // subproblems of round g beside the V^T update of round g - 1 (none before the first)
        hipLaunchKernelGGL(blk_round_kernel<T>, dim3(npair + (g > 0 ? nvt : 0u)), dim3(kBlkInner), blk_inner_lds<T>(),
                           st, P, r, g, npair, vr, g - 1);
        if (nG > 0)
          hipLaunchKernelGGL(blk_update_kernel<T>, dim3((unsigned)nG), dim3(256), blk_update_lds<T>(), st, P, r, g);
        vr = r;
      }
      if (hipGetLastError() != hipSuccess) return PSVD_ELAUNCH;
      int active = 0;
      if (hipMemcpyAsync(&active, P.flag, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess) return PSVD_ELAUNCH;
      if (hipStreamSynchronize(st) != hipSuccess) return PSVD_ELAUNCH;
      if (!active) break;
    }
    // the last round's rotations on V^T
    if (g > 0)
      hipLaunchKernelGGL(blk_round_kernel<T>, dim3(nvt), dim3(kBlkInner), blk_inner_lds<T>(), st, P, 0, g, 0, vr, g - 1);
    EighParams<T> E{};
    E.n = n;
    E.G = P.G;
    E.Vt = P.Vt;
    E.flag = P.flag;
    E.dw = dw;
    E.ord = ord;
    E.w = w + b * (int64_t)n;
    E.v = v + b * (int64_t)n * n;
    E.info = info + 2 * b;
    hipLaunchKernelGGL(eigh_diag_kernel<T>, dim3(blocks(n)), dim3(kEighThreads), 0, st, E);
    hipLaunchKernelGGL(eigh_sort_kernel<T>, dim3(blocks(n)), dim3(kEighThreads), 0, st, E);
    hipLaunchKernelGGL(eigh_out_kernel<T>, dim3(blocks((int64_t)n * n) < 4096 ? blocks((int64_t)n * n) : 4096),
                       dim3(kEighThreads), 0, st, E, sweeps);
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

size_t psvd_eigh_blocked_workspace_bytes(int dtype, int n) {
  if (n < 1) return 0;
  if (dtype == PSVD_F32) return psvd::blk_ws_bytes<float>(n);
  if (dtype == PSVD_F64) return psvd::blk_ws_bytes<double>(n);
  return 0;
}

int psvd_eigh_blocked(int dtype, int64_t batch, int n, const void *h, void *w, void *v, int32_t *info, double tol,
                      int max_sweeps, int inner_sweeps, void *workspace, size_t workspace_bytes,
                      psvd_stream_t stream) {
  if (batch < 0 || n < 1 || max_sweeps < 1 || inner_sweeps < 1 || !h || !w || !v || !info || !workspace ||
      n > 65536)
    return PSVD_EINVAL;
  if (batch == 0) return PSVD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == PSVD_F32)
    return psvd::run_eigh_blocked<float>(batch, n, static_cast<const float *>(h), static_cast<float *>(w),
                                         static_cast<float *>(v), info, tol, max_sweeps, inner_sweeps, workspace,
                                         workspace_bytes, st);
  if (dtype == PSVD_F64)
    return psvd::run_eigh_blocked<double>(batch, n, static_cast<const double *>(h), static_cast<double *>(w),
                                          static_cast<double *>(v), info, tol, max_sweeps, inner_sweeps, workspace,
                                          workspace_bytes, st);
  return PSVD_EINVAL;
}

}  // extern "C"
