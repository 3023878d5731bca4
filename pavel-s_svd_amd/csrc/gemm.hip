// K3: batched strided GEMM on MFMA, C = alpha * op(A) * op(B) + beta * C, row-major operands.
//
// Three arithmetic modes share one tiling:
//   PSVD_GEMM_F32    FP32 operands on v_mfma_f32_16x16x4_f32 (FP32 products, FP32 accumulation)
//   PSVD_GEMM_BF16X3 FP32 operands split on the way into LDS as x = hi + lo with hi = bf16(x),
//                    lo = bf16(x - hi); the product is hi*hi + hi*lo + lo*hi on
//                    v_mfma_f32_16x16x32_bf16 with FP32 accumulation (the three-term scheme of
//                    Henry, Tang and Heinecke, "Leveraging the bfloat16 artificial intelligence
//                    datatype for higher-precision computations", ARITH 2019)
//   PSVD_GEMM_F64    FP64 operands on v_mfma_f64_16x16x4_f64
// Tiles: 128x128x32 (FP32/BF16x3) or 64x64x16 (FP64) per 256-thread workgroup, four waves in a
// 2x2 arrangement, operands staged in LDS with k contiguous so that every MFMA operand is one
// LDS read per lane.  Edges are zero-filled on load and masked on store.
#include <type_traits>

#include "common.hpp"

namespace psvd {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

struct GemmParams {
  int64_t M, N, K, batch;
  const void *A, *B;
  void *C;
  int64_t lda, ldb, ldc, sa, sb, sc;
  int ta, tb;
  double alpha, beta;
  int64_t kchunk;   // split-K: z-slice z covers k in [(z % ksplit) * kchunk, ... + kchunk)
  int ksplit;       // 1: C = alpha op(A) op(B) + beta C; > 1: raw partial sums into C + z * M * N
};

constexpr int kGemmThreads = 256;

template <int MODE> struct GemmCfg;
template <> struct GemmCfg<PSVD_GEMM_F32> {
  using T = float;
  static constexpr int BM = 128, BN = 128, BK = 32, PAD = 1;
};
template <> struct GemmCfg<PSVD_GEMM_BF16X3> {
  using T = float;
  static constexpr int BM = 128, BN = 128, BK = 32, PAD = 8;
};
template <> struct GemmCfg<PSVD_GEMM_F64> {
  using T = double;
  static constexpr int BM = 64, BN = 64, BK = 16, PAD = 1;
};

__device__ __forceinline__ void split_bf16(float x, __bf16 &hi, __bf16 &lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

// One (rows x BK) operand tile held in registers between its global load and its LDS store, so
// that the next tile's loads are in flight while the MFMAs consume the current one.  `trans` says
// the operand is stored k-major (element (row, k) at k*ld + row) rather than row-major
// (row*ld + k).  Thread t owns the elements t, t + 256, ...: a fixed column (row-major) or row
// (k-major) and rows (columns) a constant stride apart, so the addresses advance by one multiply.
template <typename T, int ROWS, int BK>
struct TileRegs {
  static_assert(kGemmThreads % BK == 0 && kGemmThreads % ROWS == 0, "tile shape");
  static constexpr int E = ROWS * BK / kGemmThreads;
  static constexpr int RSTEP = kGemmThreads / BK, KSTEP = kGemmThreads / ROWS;
  T v[E];

  __device__ __forceinline__ void load(const T *src, int64_t ld, bool trans, int64_t row0, int64_t rows, int64_t k0,
                                       int64_t kend) {
    const int tid = threadIdx.x;
    if (!trans) {
      const int r = tid / BK, kk = tid - (tid / BK) * BK;
      const int64_t gk = k0 + kk;
      const T *p = src + (row0 + r) * ld + gk;
#pragma unroll
      for (int i = 0; i < E; ++i)
        v[i] = (row0 + r + i * RSTEP < rows && gk < kend) ? p[(int64_t)i * RSTEP * ld] : T(0);
    } else {
      const int kk = tid / ROWS, r = tid - (tid / ROWS) * ROWS;
      const int64_t gr = row0 + r;
      const T *p = src + (k0 + kk) * ld + gr;
#pragma unroll
      for (int i = 0; i < E; ++i)
        v[i] = (gr < rows && k0 + kk + i * KSTEP < kend) ? p[(int64_t)i * KSTEP * ld] : T(0);
    }
  }

  // LDS layout S[row][k] (ld LDS_LD); SPLIT writes hi = bf16(x) and lo = bf16(x - hi)
  template <typename S, int LDS_LD, bool SPLIT>
  __device__ __forceinline__ void store(S *dst, S *dst_lo, bool trans) const {
    const int tid = threadIdx.x;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      int r, kk;
      if (!trans) {
        r = tid / BK + i * RSTEP;
        kk = tid - (tid / BK) * BK;
      } else {
        kk = tid / ROWS + i * KSTEP;
        r = tid - (tid / ROWS) * ROWS;
      }
      if constexpr (SPLIT) {
        split_bf16((float)v[i], dst[r * LDS_LD + kk], dst_lo[r * LDS_LD + kk]);
      } else {
        dst[r * LDS_LD + kk] = v[i];
      }
    }
  }
};

// Stage one (rows x BK) operand tile straight from global memory into LDS (element loop, few
// registers: the variant for products whose K spans one or two tiles).  `trans` as for TileRegs.
template <typename T, typename S, int ROWS, int BK, int LDS_LD, bool SPLIT>
__device__ __forceinline__ void stage_tile(S *dst, S *dst_lo, const T *src, int64_t ld, bool trans, int64_t row0,
                                           int64_t rows, int64_t k0, int64_t K) {
  const int tid = threadIdx.x;
  constexpr int E = ROWS * BK;
  if (!trans) {
    for (int idx = tid; idx < E; idx += kGemmThreads) {
      const int r = idx / BK, kk = idx - r * BK;
      const int64_t gr = row0 + r, gk = k0 + kk;
      const T v = (gr < rows && gk < K) ? src[gr * ld + gk] : T(0);
      if constexpr (SPLIT) {
        split_bf16((float)v, dst[r * LDS_LD + kk], dst_lo[r * LDS_LD + kk]);
      } else {
        dst[r * LDS_LD + kk] = v;
      }
    }
  } else {
    for (int idx = tid; idx < E; idx += kGemmThreads) {
      const int kk = idx / ROWS, r = idx - kk * ROWS;
      const int64_t gr = row0 + r, gk = k0 + kk;
      const T v = (gr < rows && gk < K) ? src[gk * ld + gr] : T(0);
      if constexpr (SPLIT) {
        split_bf16((float)v, dst[r * LDS_LD + kk], dst_lo[r * LDS_LD + kk]);
      } else {
        dst[r * LDS_LD + kk] = v;
      }
    }
  }
}

// PF: the next k-tile's global loads are issued before the current tile's MFMAs (more registers,
// one workgroup per CU); without it each tile is loaded and stored in turn (higher occupancy,
// better when K spans one or two tiles, as in the K = 32 panel updates).
template <int MODE, bool PF>
__global__ __launch_bounds__(kGemmThreads) void gemm_kernel(GemmParams P) {
  using Cfg = GemmCfg<MODE>;
  using T = typename Cfg::T;
  constexpr int BM = Cfg::BM, BN = Cfg::BN, BK = Cfg::BK, LD = BK + Cfg::PAD;
  constexpr bool SPLIT = MODE == PSVD_GEMM_BF16X3;
  using S = typename std::conditional<SPLIT, __bf16, T>::type;
  constexpr int WM = BM / 2, WN = BN / 2, TI = WM / 16, TJ = WN / 16;
  __shared__ __align__(16) S As[BM * LD];
  __shared__ __align__(16) S Bs[BN * LD];
  __shared__ __align__(16) S Alo[SPLIT ? BM * LD : 1];
  __shared__ __align__(16) S Blo[SPLIT ? BN * LD : 1];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t bz = blockIdx.z / P.ksplit, kz = blockIdx.z - bz * P.ksplit;
  const T *A = static_cast<const T *>(P.A) + bz * P.sa;
  const T *B = static_cast<const T *>(P.B) + bz * P.sb;
  const bool partial = P.ksplit > 1;
  T *C = static_cast<T *>(P.C) + (partial ? (int64_t)blockIdx.z * P.M * P.N : bz * P.sc);
  const int64_t kbeg = kz * P.kchunk, kend = kbeg + P.kchunk < P.K ? kbeg + P.kchunk : P.K;
  const int64_t m0 = (int64_t)blockIdx.y * BM, n0 = (int64_t)blockIdx.x * BN;
  const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;
  const int lr = lane & 15, lk = lane >> 4;

  using Acc = typename std::conditional<MODE == PSVD_GEMM_F64, v4d, v4f>::type;
  Acc acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = 0;

  TileRegs<T, PF ? BM : kGemmThreads / BK, BK> ra;   // (a one-element tile when unused)
  TileRegs<T, PF ? BN : kGemmThreads / BK, BK> rb;
  if (PF && kbeg < kend) {
    ra.load(A, P.lda, P.ta, m0, P.M, kbeg, kend);
    rb.load(B, P.ldb, !P.tb, n0, P.N, kbeg, kend);
  }
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
    __syncthreads();
    if constexpr (PF) {
      ra.template store<S, LD, SPLIT>(As, Alo, P.ta);
      rb.template store<S, LD, SPLIT>(Bs, Blo, !P.tb);
    } else {
      stage_tile<T, S, BM, BK, LD, SPLIT>(As, Alo, A, P.lda, P.ta, m0, P.M, k0, kend);
      stage_tile<T, S, BN, BK, LD, SPLIT>(Bs, Blo, B, P.ldb, !P.tb, n0, P.N, k0, kend);
    }
    __syncthreads();
    if (PF && k0 + BK < kend) {   // next tile's loads overlap this tile's MFMAs
      ra.load(A, P.lda, P.ta, m0, P.M, k0 + BK, kend);
      rb.load(B, P.ldb, !P.tb, n0, P.N, k0 + BK, kend);
    }
    if constexpr (MODE == PSVD_GEMM_BF16X3) {
      v8bf ah[TI], al[TI], bh[TJ], bl[TJ];
#pragma unroll
      for (int i = 0; i < TI; ++i) {
        ah[i] = *reinterpret_cast<const v8bf *>(&As[(wm + 16 * i + lr) * LD + 8 * lk]);
        al[i] = *reinterpret_cast<const v8bf *>(&Alo[(wm + 16 * i + lr) * LD + 8 * lk]);
      }
#pragma unroll
      for (int j = 0; j < TJ; ++j) {
        bh[j] = *reinterpret_cast<const v8bf *>(&Bs[(wn + 16 * j + lr) * LD + 8 * lk]);
        bl[j] = *reinterpret_cast<const v8bf *>(&Blo[(wn + 16 * j + lr) * LD + 8 * lk]);
      }
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
    } else {
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        T a[TI], b[TJ];
#pragma unroll
        for (int i = 0; i < TI; ++i) a[i] = As[(wm + 16 * i + lr) * LD + kk + lk];
#pragma unroll
        for (int j = 0; j < TJ; ++j) b[j] = Bs[(wn + 16 * j + lr) * LD + kk + lk];
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
          for (int j = 0; j < TJ; ++j) {
            if constexpr (MODE == PSVD_GEMM_F64)
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
            else
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
          }
      }
    }
  }

  // D layout of the 16x16 MFMA tiles, column lane % 16 in both: register r of a lane holds row
  // 4*(lane/16) + r for the FP32-accumulating instructions and row 4*r + lane/16 for the FP64 one
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int drow = MODE == PSVD_GEMM_F64 ? 4 * r + lk : 4 * lk + r;
        const int64_t row = m0 + wm + 16 * i + drow, col = n0 + wn + 16 * j + lr;
        if (row < P.M && col < P.N) {
          if (partial) {
            C[row * P.N + col] = (T)acc[i][j][r];
          } else {
            T *c = C + row * P.ldc + col;
            const T v = (T)P.alpha * (T)acc[i][j][r];
            *c = P.beta == 0.0 ? v : v + (T)P.beta * *c;
          }
        }
      }
}

template <int MODE>
static int launch_gemm(const GemmParams &P, hipStream_t stream) {
  using Cfg = GemmCfg<MODE>;
  const dim3 grid((unsigned)((P.N + Cfg::BN - 1) / Cfg::BN), (unsigned)((P.M + Cfg::BM - 1) / Cfg::BM),
                  (unsigned)(P.batch * P.ksplit));
  const int64_t ktiles = (P.kchunk + Cfg::BK - 1) / Cfg::BK;
  if (ktiles > 2)
    hipLaunchKernelGGL((gemm_kernel<MODE, true>), grid, dim3(kGemmThreads), 0, stream, P);
  else
    hipLaunchKernelGGL((gemm_kernel<MODE, false>), grid, dim3(kGemmThreads), 0, stream, P);
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

// C_b = alpha * sum_z part[b, z] + beta * C_b, the z partials added in order (deterministic).
template <typename T>
__global__ void splitk_reduce_kernel(const T *part, int ksplit, int64_t M, int64_t N, T *C, int64_t ldc, int64_t sc,
                                     double alpha, double beta) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M * N) return;
  const int64_t b = blockIdx.y;
  const T *p = part + b * ksplit * M * N + e;
  T acc = p[0];
  for (int z = 1; z < ksplit; ++z) acc += p[(int64_t)z * M * N];
  T *c = C + b * sc + (e / N) * ldc + (e % N);
  const T v = (T)alpha * acc;
  *c = beta == 0.0 ? v : v + (T)beta * *c;
}

template <int MODE>
static int run_gemm(GemmParams &P, void *ws, size_t ws_bytes, hipStream_t st) {
  using Cfg = GemmCfg<MODE>;
  using T = typename Cfg::T;
  P.ksplit = 1;
  P.kchunk = P.K;
  const int64_t tiles = ((P.N + Cfg::BN - 1) / Cfg::BN) * ((P.M + Cfg::BM - 1) / Cfg::BM) * P.batch;
  int64_t want = tiles >= 256 ? 1 : (512 + tiles - 1) / tiles;
  const int64_t kmax = P.K / (8 * Cfg::BK);                   // at least eight k-steps per slice
  if (want > kmax) want = kmax;
  if (want > 32) want = 32;
  while (want > 1 && (size_t)want * P.batch * P.M * P.N * sizeof(T) > ws_bytes) --want;
  if (!ws || want <= 1 || P.batch * want > 65535) return launch_gemm<MODE>(P, st);
  P.ksplit = (int)want;
  P.kchunk = ((P.K + want - 1) / want + Cfg::BK - 1) / Cfg::BK * Cfg::BK;
  P.ksplit = (int)((P.K + P.kchunk - 1) / P.kchunk);
  T *C = static_cast<T *>(P.C);
  const int64_t ldc = P.ldc, sc = P.sc;
  P.C = ws;
  const int rc = launch_gemm<MODE>(P, st);
  if (rc != PSVD_OK) return rc;
  const int64_t e = P.M * P.N;
  hipLaunchKernelGGL(splitk_reduce_kernel<T>, dim3((unsigned)((e + 255) / 256), (unsigned)P.batch), dim3(256), 0, st,
                     static_cast<const T *>(ws), P.ksplit, P.M, P.N, C, ldc, sc, P.alpha, P.beta);
  return hipGetLastError() == hipSuccess ? PSVD_OK : PSVD_ELAUNCH;
}

// Workspace that lets psvd::gemm_ws split K for an M x N output (batch of them): 32 slices.
size_t gemm_splitk_bytes(int mode, int64_t m, int64_t n, int64_t batch) {
  return (size_t)32 * batch * m * n * (mode == PSVD_GEMM_F64 ? 8 : 4);
}

}  // namespace psvd

extern "C" int psvd_gemm(int mode, int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha,
                         const void *a, int64_t lda, int64_t stride_a, const void *b, int64_t ldb,
                         int64_t stride_b, double beta, void *c, int64_t ldc, int64_t stride_c, int64_t batch,
                         psvd_stream_t stream) {
  return psvd::gemm_ws(mode, transa, transb, m, n, k, alpha, a, lda, stride_a, b, ldb, stride_b, beta, c, ldc,
                       stride_c, batch, nullptr, 0, static_cast<hipStream_t>(stream));
}

namespace psvd {

// Same contract as psvd_gemm; with a workspace of gemm_splitk_bytes(), outputs too small to fill
// the GPU are computed as K slices and reduced in a fixed order.
int gemm_ws(int mode, int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha, const void *a, int64_t lda,
            int64_t stride_a, const void *b, int64_t ldb, int64_t stride_b, double beta, void *c, int64_t ldc,
            int64_t stride_c, int64_t batch, void *ws, size_t ws_bytes, hipStream_t st) {
  if (m < 0 || n < 0 || k < 0 || batch < 0 || !c) return PSVD_EINVAL;
  if (m == 0 || n == 0 || batch == 0) return PSVD_OK;
  if (k > 0 && (!a || !b)) return PSVD_EINVAL;
  if (ldc < n || lda < (transa ? m : k) || ldb < (transb ? k : n)) return PSVD_EINVAL;
  if (batch > 65535 || (n + 127) / 128 > 0x7fffffff || (m + 63) / 64 > 65535) return PSVD_EINVAL;
  GemmParams P;
  P.M = m;
  P.N = n;
  P.K = k;
  P.batch = batch;
  P.A = a;
  P.B = b;
  P.C = c;
  P.lda = lda;
  P.ldb = ldb;
  P.ldc = ldc;
  P.sa = stride_a;
  P.sb = stride_b;
  P.sc = stride_c;
  P.ta = transa;
  P.tb = transb;
  P.alpha = alpha;
  P.beta = beta;
  if (mode == PSVD_GEMM_F32) return run_gemm<PSVD_GEMM_F32>(P, ws, ws_bytes, st);
  if (mode == PSVD_GEMM_BF16X3) return run_gemm<PSVD_GEMM_BF16X3>(P, ws, ws_bytes, st);
  if (mode == PSVD_GEMM_F64) return run_gemm<PSVD_GEMM_F64>(P, ws, ws_bytes, st);
  return PSVD_EINVAL;
}

}  // namespace psvd
