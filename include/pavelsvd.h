/*
 * pavelsvd.h -- C ABI of the MI355X (gfx950) standalone SVD hot path.
 *
 * The reference (svdbench/standalone, pure Python + PyTorch) has no native interface; these
 * entry points are the device-side halves of its Python primitives, with the contract of
 * svdbench/standalone/solvers.py:264-311 and primitives.py: plain device pointers and sizes,
 * caller-allocated outputs and workspace, a workspace-size query, an int status return and
 * per-matrix device-side info.  No entry point allocates device memory or synchronises the host,
 * so every call can be captured in a hipGraph (except where noted).
 *
 * Layouts: matrices are row-major (the PyTorch default of the reference), element offsets are
 * counted in elements of the stated dtype.  `stream` is a hipStream_t (NULL = default stream).
 */
#ifndef PAVELSVD_H
#define PAVELSVD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *psvd_stream_t;

/* return status */
#define PSVD_OK 0
#define PSVD_EINVAL 1     /* bad argument (shape, dtype, flags, workspace too small) */
#define PSVD_ELAUNCH 2    /* the HIP runtime refused a launch */
#define PSVD_ETOOBIG 3    /* the problem does not fit this kernel's on-chip budget */

/* dtype codes */
#define PSVD_F32 0
#define PSVD_F64 1
#define PSVD_C64 2        /* complex64: interleaved (re, im) float pairs; psvd_jacobi_lds only */
#define PSVD_C128 3       /* complex128: interleaved (re, im) double pairs; psvd_jacobi_lds only */

/* per-matrix device info: info[2*i] = sweeps (or iterations), info[2*i+1] = status */
#define PSVD_INFO_OK 0
#define PSVD_INFO_NOCONV 1      /* maps to RuntimeError('... did not converge ...') */
#define PSVD_INFO_COMPLETION 2  /* maps to RuntimeError('null-space completion failed ...') */
#define PSVD_INFO_NONFINITE 3   /* maps to ValueError('input contains NaN or infinity') */
#define PSVD_INFO_BADSHAPE 4    /* a descriptor violates 1 <= n <= m, m <= max_m, n <= max_n */
#define PSVD_INFO_BARRIER 5     /* a single-launch kernel (psvd_jacobi_grid) abandoned a grid-barrier wait:
                                   its workgroups were not all resident (the GPU was shared with other
                                   launches that held its CUs); maps to RuntimeError */

/* flags for the Jacobi entry points */
#define PSVD_SCALE 1        /* divide by max|a| first and multiply sigma back (solvers.py:294-299) */
#define PSVD_CUT_FLOOR 2    /* zero sigma only below eps*sigma_1 (the Jacobi rotation floor)
                               instead of eps*max(m,n)*sigma_1 (primitives.py:362) */
#define PSVD_NO_VECTORS 4   /* sigma only: no V accumulation, U and Vh are not written */
#define PSVD_SLOT_SPREAD 8  /* psvd_jacobi_lds: a round's pairs served in bank-spread slot order (a
                               launch-shape choice; the results are bit-identical) */
#define PSVD_ORDER_SORTED 16 /* psvd_jacobi_lds: pair ordering -- at the start of every sweep the
                               tournament's labels are the columns in descending norm order (de Rijk)
                               instead of the reference's identity labels (primitives.py:309-313) */
#define PSVD_BLOCK_NARROW 32 /* psvd_jacobi_blocked: column blocks of 16 (32 x 32 subproblems; the
                               default when neither block flag is given) */
#define PSVD_BLOCK_WIDE 64   /* psvd_jacobi_blocked: column blocks of 32 (64 x 64 subproblems) */
#define PSVD_BLOCK_FULL_PAIRS 128 /* psvd_jacobi_blocked: every tournament round sweeps all pairs of its
                               subproblems (by default only a sweep's first round does; the later
                               rounds sweep the pairs across the two blocks) */

/* GEMM arithmetic modes */
#define PSVD_GEMM_F32 0     /* FP32 operands, FP32 MFMA, FP32 accumulation */
#define PSVD_GEMM_BF16X3 1  /* FP32 operands split hi+lo in bf16, three bf16 MFMA products, FP32 acc */
#define PSVD_GEMM_F64 2     /* FP64 operands, FP64 MFMA */

/* Library identification: returns a static string ("pavelsvd <version> gfx950"). */
const char *psvd_version(void);

/* ---------------------------------------------------------------------------------------
 * Batched one-sided Jacobi SVD, LDS resident: one workgroup per matrix drawn from a device
 * work queue (persistent launch), W and V held in LDS, column dot products reduced with
 * wavefront shuffles and accumulated in FP64, FP64 convergence test, FP64 sigma.
 * Replaces primitives.jacobi_svd (svdbench/standalone/primitives.py:316-367) including the
 * round-robin schedule (:309-313), the sort / cut / completion tail (:356-366,
 * complete_null_columns :430-455) and the max-abs scaling of solvers.svd (:294-299).
 *
 * Matrices are tall (m >= n; the caller transposes wide inputs as solvers.py:226-228 does).
 * Uniform mode (dims == NULL): batch matrices of m_uniform x n_uniform packed contiguously.
 * Uneven mode: dims[2i] = m_i, dims[2i+1] = n_i; offs[3i] = element offset of A_i and U_i,
 * offs[3i+1] = offset of s_i, offs[3i+2] = offset of Vh_i (n_i x n_i).  order (optional) is a
 * permutation of [0, batch) giving the processing order (longest first balances the queue).
 * max_m / max_n bound the shapes (they size the LDS).  tol <= 0 selects 10*eps(dtype).
 * lanes = lanes per column pair: 2..64 (power of two; 1 is raised to 2), or -k (k in 1..64) for
 * about k rows per lane of the largest matrix (0 means -4).  threads = 64, 128, 256, 512 or 1024 (0 selects 256).
 * One launch uses one lane count; callers with very uneven batches split them into size
 * classes, one call each (the LDS budget is also sized per call from max_m, max_n).
 * Complex dtypes (PSVD_C64 / PSVD_C128, uniform mode only, threads fixed at 256): a, u and vh
 * are interleaved (re, im) pairs, s is real (float / double); the complex rotation of
 * primitives.py:340-351 (c real, s = t c a_pq / |a_pq|, w_p <- c w_p - conj(s) w_q) and
 * vh = V^H.
 * ------------------------------------------------------------------------------------- */
size_t psvd_jacobi_lds_workspace_bytes(int64_t batch);
int psvd_jacobi_lds_smem_bytes(int dtype, int max_m, int max_n, int flags, size_t *bytes);
int psvd_jacobi_lds(int dtype, int64_t batch, int m_uniform, int n_uniform,
                    const int32_t *dims, const int64_t *offs, const int32_t *order,
                    int max_m, int max_n,
                    const void *a, void *u, void *s, void *vh, int32_t *info,
                    double tol, int max_sweeps, int flags, int lanes, int threads,
                    void *workspace, size_t workspace_bytes, psvd_stream_t stream);

/* psvd_jacobi_lds_noalloc: psvd_jacobi_lds without a workspace argument and without any device
 * allocation.  No counterpart in the reference: it is for callers that cannot allocate (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).  The launch's only scratch,
 * the device work-queue counter, lives in a library-owned 256-byte device array per GPU, held until
 * the process ends; it is cleared by a fill kernel at the start of every call.  Every call on a GPU
 * shares that GPU's array, so calls that overlap on one GPU (two streams, two host threads) race and
 * must be serialised by the caller.  Maximum: batch <= 2^31 - 1 (one int32 counter), else
 * PSVD_ETOOBIG.  Same arguments, outputs and statuses as psvd_jacobi_lds otherwise. */
int psvd_jacobi_lds_noalloc(int dtype, int64_t batch, int m_uniform, int n_uniform,
                            const int32_t *dims, const int64_t *offs, const int32_t *order,
                            int max_m, int max_n,
                            const void *a, void *u, void *s, void *vh, int32_t *info,
                            double tol, int max_sweeps, int flags, int lanes, int threads,
                            psvd_stream_t stream);

/* ---------------------------------------------------------------------------------------
 * Batched strided GEMM on MFMA: C_i = alpha * op(A_i) * op(B_i) + beta * C_i for i < batch,
 * row-major: op(A) is m x k (A is m x k, or k x m when transa), op(B) is k x n (B is k x n, or
 * n x k when transb); X_i = X + i * stride_x.  Replaces the `@` products of the reference's
 * QDWH, TSQR and QR paths (solvers.py:134, 151-153, 159-160, 168, 173; :93-94, :248;
 * primitives.py:64-65, 449) with one kernel whose arithmetic is chosen by `mode`.
 * ------------------------------------------------------------------------------------- */
int psvd_gemm(int mode, int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha,
              const void *a, int64_t lda, int64_t stride_a, const void *b, int64_t ldb, int64_t stride_b,
              double beta, void *c, int64_t ldc, int64_t stride_c, int64_t batch, psvd_stream_t stream);

/* ---------------------------------------------------------------------------------------
 * Batched utilities between solver stages (graph-capturable; outputs cleared by a kernel).
 *  psvd_reduce: op 0 out[b] = max|x_b|; op 1 out[b] = ||x_b||_F; op 2 out[b] = max|x_b - y_b| and
 *               out[batch + b] = max|x_b| (out: device doubles); op 3 out[b] = the largest modulus
 *               of the (re, im) pairs of each row (cols even: a complex matrix seen as real, the
 *               reference's x.abs().amax() for complex x); op 4 the same two outputs as op 2 with
 *               moduli of (re, im) pairs (cols even: (x - y).abs().amax() and x.abs().amax() of
 *               complex x, y).  The max-abs and Frobenius
 *               normalisations of solvers.py:294-295 and :124-125, the QDWH test :161-163.
 *  psvd_scale:  y_b = x_b / d_b (invert, T division as the reference's x / scale) or
 *               y_b = (s * d_b) x_b; d_b <= 0 counts as 1; d may be NULL.
 *  psvd_symmetrize: a_b = (a_b + a_b^T) / 2 in place (solvers.py:156, 169).
 *  psvd_axpby: y_b = alpha x_b + beta y_b (the QDWH update, solvers.py:153, 160).
 * ------------------------------------------------------------------------------------- */
int psvd_reduce(int dtype, int op, int64_t batch, int rows, int cols, const void *x, int64_t ldx, int64_t sx,
                const void *y, int64_t ldy, int64_t sy, double *out, psvd_stream_t stream);
int psvd_scale(int dtype, int64_t batch, int rows, int cols, const void *x, int64_t ldx, int64_t sx, void *y,
               int64_t ldy, int64_t sy, const double *d, int invert, double s, psvd_stream_t stream);
int psvd_symmetrize(int dtype, int64_t batch, int n, void *a, int64_t lda, int64_t sa, psvd_stream_t stream);
int psvd_axpby(int dtype, int64_t batch, int rows, int cols, double alpha, const void *x, int64_t ldx, int64_t sx,
               double beta, void *y, int64_t ldy, int64_t sy, psvd_stream_t stream);
/* psvd_sort_desc: perm_b = descending order of keys_b (n FP64 values, ties by index), then
 * s_b[c] = max(keys_b[perm_b[c]], 0) and y_b[:, c] = x_b[:, perm_b[c]] (rows x n; rows may be 0;
 * y must not alias x).  Orders the singular values of the FP64 polishing step and their vectors
 * (no reference counterpart; the reference flips an ascending eigensolve, solvers.py:171);
 * perm: batch * n int32 workspace. */
int psvd_sort_desc(int dtype, int64_t batch, int rows, int n, const double *keys, void *s, const void *x, int64_t ldx,
                   int64_t sx, void *y, int64_t ldy, int64_t sy, int32_t *perm, psvd_stream_t stream);
/* psvd_scale_columns: x_b[:, c] /= d_b[c] wherever d_b[c] > 0, in place (d in x's dtype, row stride
 * sd per matrix): U = W / sigma of solvers.py:18-21 (_finish_columns) and U = A V / sigma of
 * solvers.py:116 (_gram), before the completion of the columns whose sigma is cut. */
int psvd_scale_columns(int dtype, int64_t batch, int rows, int cols, void *x, int64_t ldx, int64_t sx, const void *d,
                       int64_t sd, psvd_stream_t stream);

/* ---------------------------------------------------------------------------------------
 * Completion of missing basis columns (primitives.py:430-467): for each b, columns [nk[b], n) of
 * U_b (m x n, row-major) are replaced in order by the projected coordinate vector at the row of
 * largest remaining leverage (two Gram-Schmidt passes, FP64); status[b] = PSVD_INFO_COMPLETION
 * when a residual falls below 10 eps.  Serves complete_null_columns (sorted sigma: the missing
 * columns are the tail) and complete_basis (full_matrices).
 * ------------------------------------------------------------------------------------- */
size_t psvd_complete_workspace_bytes(int64_t batch, int m, int n);
int psvd_complete_columns(int dtype, int64_t batch, int m, int n, void *u, int64_t ldu, int64_t stride_u,
                          const int32_t *nk, int32_t *status, void *workspace, size_t workspace_bytes,
                          psvd_stream_t stream);

/* ---------------------------------------------------------------------------------------
 * Blocked Householder QR of a batch of tall matrices: A_i = Q_i R_i, Q_i thin (m x n, row-major,
 * contiguous), R_i (n x n, upper triangular).  Reflectors in the reference's xLARFG convention,
 * so R matches primitives.qr (svdbench/standalone/primitives.py:41-89) up to rounding.  A is
 * read only (row stride lda, batch stride stride_a).  q or r may be NULL to skip that output.
 * gemm_mode: PSVD_GEMM_F32 or PSVD_GEMM_BF16X3 for dtype PSVD_F32 (ignored for PSVD_F64).
 * Panels taller than 512 rows are factored by several workgroups per matrix in one launch of at
 * most one workgroup per CU, its phases separated by the library's grid barrier (an ordinary
 * launch, capturable in a hipGraph; a barrier wait abandoned because the GPU was shared with other
 * launches holding its CUs leaves the panel NaN, so the caller's finiteness check reports it).
 * ------------------------------------------------------------------------------------- */
size_t psvd_qr_workspace_bytes(int dtype, int64_t batch, int m, int n);
int psvd_qr(int dtype, int gemm_mode, int64_t batch, int m, int n, const void *a, int64_t lda, int64_t stride_a,
            void *q, void *r, void *workspace, size_t workspace_bytes, psvd_stream_t stream);

/* ---------------------------------------------------------------------------------------
 * Householder QR of a batch of small complex matrices held in LDS (one workgroup per matrix):
 * replaces primitives.qr for complex inputs (svdbench/standalone/primitives.py:41-89, the complex
 * branch of householder_vec: beta real, tau = (beta - alpha) / beta, factorisation with conj(tau),
 * Q accumulated backwards with tau).  dtype PSVD_C64 or PSVD_C128; a (batch, m, n), q (batch, m, n)
 * and r (batch, n, n, upper triangular with a real diagonal) are contiguous, row-major, with
 * interleaved (re, im) pairs; m >= n.  psvd_cqr_lds_smem_bytes gives the LDS the shape needs and
 * returns PSVD_ETOOBIG when it does not fit (16 m n bytes for complex64, 32 m n for complex128).
 * psvd_cmatmul: c = op(a) op(b) for each matrix of the batch (op(a): m x k, a stored m x k, or
 * k x m with transa = 1 for a^H; op(b): k x n, b stored k x n, or n x k with transb = 1 for b^H;
 * c: m x n; contiguous, interleaved complex, FP64 sums): the complex products of the reference's
 * qr solver (q @ u, solvers.py:245-248) and QDWH iteration (top @ bot^H, x^H @ a, x @ v,
 * solvers.py:151-153, 168-173).
 * psvd_chermitize: a <- (a + a^H) / 2 for each n x n complex matrix (solvers.py:169).
 * psvd_ccomplete_columns: complex counterpart of psvd_complete_columns (primitives.py:430-467):
 * for each matrix of u (batch, m, n, contiguous interleaved), columns [nk[b], n) are replaced in
 * order by the projected coordinate vector at the row of largest remaining leverage (two
 * Gram-Schmidt passes, FP64); status[b] = PSVD_INFO_COMPLETION when a residual falls below 10 eps.
 * ------------------------------------------------------------------------------------- */
/* Batched complex one-sided Jacobi of a uniform batch at any size: psvd_jacobi_lds's complex path
 * (same arithmetic, outputs and info) when a matrix fits LDS; beyond it the same kernel runs with
 * each resident workgroup's planes in a region of the workspace (psvd_cjacobi_batched_workspace_bytes:
 * 256 bytes for the work queue, plus the regions when the shape exceeds LDS). */
size_t psvd_cjacobi_batched_workspace_bytes(int dtype, int64_t batch, int m, int n, int flags);
int psvd_cjacobi_batched(int dtype, int64_t batch, int m, int n, const void *a, void *u, void *s, void *vh,
                         int32_t *info, double tol, int max_sweeps, int flags, void *workspace, size_t workspace_bytes,
                         psvd_stream_t stream);

int psvd_cqr_lds_smem_bytes(int dtype, int m, int n, size_t *bytes);
int psvd_cqr_lds(int dtype, int64_t batch, int m, int n, const void *a, void *q, void *r, psvd_stream_t stream);
/* The same QR for any m >= n: shapes beyond LDS run the same kernel with the planes in the workspace
 * (one region per matrix; psvd_cqr_workspace_bytes, 0 when the shape fits LDS). */
size_t psvd_cqr_workspace_bytes(int dtype, int64_t batch, int m, int n);
int psvd_cqr(int dtype, int64_t batch, int m, int n, const void *a, void *q, void *r, void *workspace,
             size_t workspace_bytes, psvd_stream_t stream);
int psvd_cmatmul(int dtype, int transa, int transb, int64_t batch, int m, int k, int n, const void *a, const void *b,
                 void *c, psvd_stream_t stream);
int psvd_chermitize(int dtype, int64_t batch, int n, void *a, psvd_stream_t stream);
int psvd_ccomplete_columns(int dtype, int64_t batch, int m, int n, void *u, const int32_t *nk, int32_t *status,
                           psvd_stream_t stream);

/* ---------------------------------------------------------------------------------------
 * Complex Hermitian kernels for matrices held in LDS (one workgroup per matrix; dtype PSVD_C64 or
 * PSVD_C128; interleaved (re, im) pairs, contiguous (batch, rows, cols) row-major; FP64 sums):
 *  psvd_ceigh_lds: two-sided cyclic Jacobi of Hermitian h -- the complex branch of jacobi_eigh
 *      (primitives.py:370-427): c real, s = t c g_pq / |g_pq|, G <- G J then J^H G, V <- V J, the
 *      test |g_pq| > tol max|G| per round, stop after a sweep without rotations (tol <= 0 selects
 *      10 eps); w (batch, n) real ascending, v (batch, n, n) eigenvectors as columns, info = [sweeps,
 *      status].  psvd_ceigh_lds_smem_bytes gives the LDS the size needs (PSVD_ETOOBIG beyond).
 *  psvd_ccholesky_lds: L (lower, strict upper zero) with G = L L^H from the lower triangle of the
 *      Hermitian g, the reference's failure mask (a pivot <= 0 is clamped to tiny and flagged, fail[b] = 1;
 *      primitives.py:145-159).
 *  psvd_csolve_chol_lds: x <- x (L L^H)^-1 for x (batch, m, n), the combined solve
 *      trsm_upper(L^H, trsm_lower(L, x^H))^H of the complex QDWH Cholesky branch (solvers.py:159-160,
 *      primitives.py:162-179).  psvd_csolve_lds_smem_bytes gives its LDS.
 * ------------------------------------------------------------------------------------- */
int psvd_ceigh_lds_smem_bytes(int dtype, int n, size_t *bytes);
int psvd_ceigh_lds(int dtype, int64_t batch, int n, const void *h, void *w, void *v, int32_t *info, double tol,
                   int max_sweeps, psvd_stream_t stream);
int psvd_ccholesky_lds(int dtype, int64_t batch, int n, const void *g, void *l, int32_t *fail, psvd_stream_t stream);
int psvd_csolve_lds_smem_bytes(int dtype, int m, int n, size_t *bytes);
int psvd_csolve_chol_lds(int dtype, int64_t batch, int m, int n, const void *l, void *x, psvd_stream_t stream);
/* The same three for any size: shapes whose planes exceed LDS run the same kernels with the planes in
 * the workspace (one region per matrix; the *_workspace_bytes queries give 0 when the shape fits
 * LDS, where the workspace may be null). */
size_t psvd_ceigh_workspace_bytes(int dtype, int64_t batch, int n);
int psvd_ceigh(int dtype, int64_t batch, int n, const void *h, void *w, void *v, int32_t *info, double tol,
               int max_sweeps, void *workspace, size_t workspace_bytes, psvd_stream_t stream);
size_t psvd_ccholesky_workspace_bytes(int dtype, int64_t batch, int n);
int psvd_ccholesky(int dtype, int64_t batch, int n, const void *g, void *l, int32_t *fail, void *workspace,
                   size_t workspace_bytes, psvd_stream_t stream);
size_t psvd_csolve_chol_workspace_bytes(int dtype, int64_t batch, int m, int n);
int psvd_csolve_chol(int dtype, int64_t batch, int m, int n, const void *l, void *x, void *workspace,
                     size_t workspace_bytes, psvd_stream_t stream);

/* ---------------------------------------------------------------------------------------
 * Two-sided cyclic Jacobi eigensolver of a batch of symmetric n x n matrices (primitives.py:
 * 370-427): w (batch, n) ascending, v (batch, n, n) row-major eigenvectors as columns,
 * info = [sweeps, status].  Uses an LDS-resident kernel (one workgroup per matrix) when the
 * matrix fits, otherwise two kernels per round on G and V^T in the workspace of
 * psvd_eigh_workspace_bytes, with one host read of the activity flag per sweep (not capturable).
 * The rotation order is the reference's, so sweep counts track it.  tol <= 0 selects 10*eps(dtype).
 * ------------------------------------------------------------------------------------- */
size_t psvd_eigh_workspace_bytes(int dtype, int n);
int psvd_eigh(int dtype, int64_t batch, int n, const void *h, void *w, void *v, int32_t *info, double tol,
              int max_sweeps, void *workspace, size_t workspace_bytes, psvd_stream_t stream);

/* ---------------------------------------------------------------------------------------
 * Blocked two-sided Jacobi eigensolver for large symmetric matrices; same outputs, tolerance
 * test and stopping rule as psvd_eigh (primitives.py:370-427: |g_pq| > tol * max|G|, stop after
 * a sweep without rotations), with the rotations grouped: n is split into 32-index blocks, each
 * round of a tournament over the blocks sweeps the 64 x 64 subproblem of every block pair in LDS
 * (`inner_sweeps` cyclic sweeps, >= 1) and applies the accumulated 64 x 64 rotations to the rest
 * of G and to V with MFMA products.  info[0] counts block sweeps.  One host read per sweep.
 * ------------------------------------------------------------------------------------- */
size_t psvd_eigh_blocked_workspace_bytes(int dtype, int n);
int psvd_eigh_blocked(int dtype, int64_t batch, int n, const void *h, void *w, void *v, int32_t *info, double tol,
                      int max_sweeps, int inner_sweeps, void *workspace, size_t workspace_bytes,
                      psvd_stream_t stream);

/* ---------------------------------------------------------------------------------------
 * Blocked Cholesky A = L L^T in place (lower; strict upper cleared) with the reference's failure
 * mask fail[b] (a pivot <= 0 is clamped and flagged, primitives.py:145-159), and right-sided
 * triangular solves B <- B L^-T (trans = 1) or B <- B L^-1 (trans = 0) for rows x n blocks B
 * (primitives.py:162-179 in the transposed form QDWH uses, solvers.py:159).
 * ------------------------------------------------------------------------------------- */
int psvd_cholesky(int dtype, int gemm_mode, int64_t batch, int n, void *a, int64_t lda, int64_t stride_a,
                  int32_t *fail, psvd_stream_t stream);
int psvd_trsm_right(int dtype, int gemm_mode, int64_t batch, int trans, int n, int rows, const void *l, int64_t ldl,
                    int64_t stride_l, void *b, int64_t ldb, int64_t stride_b, psvd_stream_t stream);

/* ---------------------------------------------------------------------------------------
 * One-sided Jacobi SVD of ONE tall matrix that does not fit LDS (e.g. 512x256, 2048x32, the
 * n x n R of TSQR), as a single launch of at most one workgroup per CU: W and V in the workspace
 * (HBM), each workgroup owns a strided share of every round's column pairs, one grid barrier per
 * round (the library's own barrier on counters in the workspace; status PSVD_INFO_BARRIER if a wait
 * is abandoned because other launches held the GPU's CUs),
 * FP64 dot products / convergence test / sigma, sort, cut and completion on the device.
 * Same algorithm and flags as psvd_jacobi_lds (primitives.py:316-367); info = [sweeps, status].
 * Complex dtypes (PSVD_C64 / PSVD_C128): interleaved a, u, vh, real s, the complex rotation of
 * primitives.py:340-351, vh = V^H; the cut columns are completed by psvd_ccomplete_columns in the
 * same stream (m + n up to about 9700).
 * Ordinary launches only: capturable in a hipGraph.
 * ------------------------------------------------------------------------------------- */
size_t psvd_jacobi_grid_workspace_bytes(int dtype, int m, int n);
int psvd_jacobi_grid(int dtype, int m, int n, const void *a, void *u, void *s, void *vh, int32_t *info,
                     double tol, int max_sweeps, int flags, void *workspace, size_t workspace_bytes,
                     psvd_stream_t stream);

/* ---------------------------------------------------------------------------------------
 * Blocked one-sided Jacobi SVD of ONE large tall matrix (primitives.py:316-367 with the rotations
 * grouped): columns in blocks of 16 (or 32 with PSVD_BLOCK_WIDE), a tournament over the blocks, for
 * each block pair the 32 x 32 (64 x 64) Gram matrix in FP64 on MFMA swept by `inner_sweeps` (>= 1)
 * cyclic sweeps of the reference's one-sided rotations and test, the accumulated rotation applied
 * to W and V on MFMA; stops after a sweep without rotations.  Same flags as psvd_jacobi_grid plus
 * the block-width flags, same outputs and info (info[0] counts block sweeps).  One host read per
 * sweep, with the next sweep already queued (when a sweep rotated nowhere, the queued one finds a
 * done flag and does nothing); the flag words are 8 bytes of pinned host memory allocated on a host
 * thread's first call and kept for the process.  Not capturable.
 * ------------------------------------------------------------------------------------- */
size_t psvd_jacobi_blocked_workspace_bytes(int dtype, int m, int n);
int psvd_jacobi_blocked(int dtype, int m, int n, const void *a, void *u, void *s, void *vh, int32_t *info, double tol,
                        int max_sweeps, int flags, int inner_sweeps, void *workspace, size_t workspace_bytes,
                        psvd_stream_t stream);

/* ---------------------------------------------------------------------------------------
 * Block-pair one-sided Jacobi (solvers.py:30-68, `block_jacobi`): ONE sweep over the tournament of
 * column groups of `block_size` columns (the last group may be short; an odd count gets an empty
 * partner) on W (batch x m x n) and V (batch x n x n), both updated in place.  Each round runs one
 * workgroup per group pair and matrix: the pair's k <= 2 block_size (<= 64, else PSVD_ETOOBIG)
 * columns are orthogonalised by the reference's one-sided Jacobi (tolerance tol, 0 = 10 eps; at
 * most max_sweeps sweeps; primitives.py:316-367) carried out on their FP64 Gram matrix, and
 * W_pair, V_pair are multiplied by the accumulated rotation with its columns ordered by
 * descending norm (solvers.py:43-50).  Then, for the sweep's convergence test (solvers.py:51-60):
 * corr[b] = max over live columns i != j of |w_i^T w_j| / (|w_i| |w_j|) (live: |w_i|^2 >
 * (eps max(m, n))^2 max_c |w_c|^2), sigma[b * n + c] = |w_c| (FP64 sums throughout).
 * info[b] = {largest pair sweep count, 1 if a pair did not converge}.  The caller stops when
 * corr <= its tolerance (40 eps by default).  A single group (n <= block_size) is PSVD_EINVAL: that
 * case is the plain one-sided Jacobi SVD.  Workspace: psvd_block_jacobi_workspace_bytes(batch, n).
 * ------------------------------------------------------------------------------------- */
size_t psvd_block_jacobi_workspace_bytes(int64_t batch, int n);
int psvd_block_jacobi_sweep(int dtype, int64_t batch, int m, int n, int block_size, void *w, int64_t ldw,
                            int64_t stride_w, void *v, int64_t ldv, int64_t stride_v, double tol, int max_sweeps,
                            int32_t *info, double *corr, double *sigma, void *workspace, size_t workspace_bytes,
                            psvd_stream_t stream);

/* ---------------------------------------------------------------------------------------
 * Learned QDWH lower bound (no reference counterpart: the reference starts QDWH from l0 = eps,
 * solvers.py:128-130).  A 32 -> 64 -> 64 -> 1 tanh MLP, FP32, parameters in one flat vector of
 * psvd_mlp_param_count() floats: W1 (64 x 32), b1, W2 (64 x 64), b2, w3 (64), b3.
 *  psvd_mlp_forward: y[i] = MLP(x[i, :]) for x (batch x 32).
 *  psvd_mlp_grad: y as above and grad = d/dparams of (1 / 2 batch) sum (y_i - target_i)^2, loss
 *                 (device double, may be NULL); per-workgroup partial gradients are added in a
 *                 fixed order (deterministic).  Workspace of psvd_mlp_workspace_bytes(batch).
 *  psvd_adam_step: Adam update of n parameters with bias correction (step >= 1).
 *  psvd_colstats: per matrix, {min, max, mean, std} of log10 of the column norms and log10 max|a|
 *                 (out: batch x 5 floats), the column-norm part of the sketch features.
 * ------------------------------------------------------------------------------------- */
int psvd_mlp_param_count(void);
size_t psvd_mlp_workspace_bytes(int64_t batch);
int psvd_mlp_forward(const float *params, const float *x, int64_t batch, float *y, psvd_stream_t stream);
int psvd_mlp_grad(const float *params, const float *x, const float *target, int64_t batch, float *y, float *grad,
                  double *loss, void *workspace, size_t workspace_bytes, psvd_stream_t stream);
int psvd_adam_step(float *params, const float *grad, float *m, float *v, int n, float lr, float beta1, float beta2,
                   float eps, int step, psvd_stream_t stream);
int psvd_colstats(int dtype, int64_t batch, int rows, int cols, const void *a, int64_t lda, int64_t stride_a,
                  float *out, psvd_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* PAVELSVD_H */
