"""Device primitives of the standalone hot path, each one gfx950 kernel launch (or a fixed
sequence of them) through the C ABI of ``include/pavelsvd.h``.

Mirrors the names and return conventions of svdbench/standalone/primitives.py for the hot path:

    jacobi_svd   primitives.py:316-367  -> K1 (LDS-resident batched one-sided Jacobi) or,
                                           for matrices beyond LDS, K2 (single launch)
    qr           primitives.py:41-89    -> K4 (blocked Householder QR, xLARFG reflectors)
    jacobi_eigh  primitives.py:370-427  -> K6 (two-sided Jacobi, LDS-resident or in HBM)
    cholesky     primitives.py:145-159  -> K5 (blocked Cholesky with the failure mask)
    trsm_lower / trsm_upper  :162-179   -> K5 (right-sided blocked solves, transposed form)

Tensors must live on a ROCm device; there is no CPU path.
"""
import torch

from .. import native as N


def lds_fits(dtype, m, n, vectors=True):
    """Whether an m x n (m >= n) matrix fits the LDS-resident Jacobi kernel."""
    st, _ = N._hip.jacobi_lds_smem_bytes(N.dtype_code(dtype), m, n, 0 if vectors else N.NO_VECTORS)
    return st == N.OK


def jacobi_lds_uniform(x, *, tol=None, max_sweeps=60, scale=False, cut='reference', vectors=True,
                       lanes=None, threads=None):
    """One-sided Jacobi SVD of every (m, n) matrix of the contiguous tall batch ``x`` (b, m, n).

    Returns U (b, m, n), s (b, n) descending, Vh (b, n, n) (None, s, None without vectors) and
    the per-matrix info tensor (b, 2) = [sweeps, status] on the device; no host synchronisation.
    Complex matrices beyond LDS run the same complex kernel on regions of a workspace.
    """
    b, m, n = x.shape
    if m < n:
        raise ValueError('jacobi_lds_uniform takes tall matrices (m >= n)')
    x = x.contiguous()
    code = N.dtype_code(x.dtype)
    flags = (N.SCALE if scale else 0) | (N.CUT_FLOOR if cut == 'floor' else 0) | (0 if vectors else N.NO_VECTORS)
    dev = x.device
    u = torch.empty_like(x) if vectors else None
    s = torch.empty(b, n, dtype=x.real.dtype if x.is_complex() else x.dtype, device=dev)
    vh = torch.empty(b, n, n, dtype=x.dtype, device=dev) if vectors else None
    info = torch.empty(b, 2, dtype=torch.int32, device=dev)
    if x.is_complex() and not lds_fits(x.dtype, m, n, vectors):
        wsb = N._hip.cjacobi_batched_workspace_bytes(code, b, m, n, flags)
        ws = N.workspace(wsb, dev)
        st = N._hip.cjacobi_batched(code, b, m, n, N.ptr(x), N.ptr(u), N.ptr(s), N.ptr(vh), N.ptr(info),
                                    float(tol or 0.0), int(max_sweeps), flags, N.ptr(ws), wsb, N.stream_of(x))
        N.check(st, 'psvd_cjacobi_batched')
        return u, s, vh, info
    wsb = N._hip.jacobi_lds_workspace_bytes(b)
    ws = N.workspace(wsb, dev)
    if x.is_complex():   # the complex kernel runs 256 threads, lanes from the row count
        threads, lanes = 256, 0 if lanes is None else lanes
    if lanes is None or threads is None:
        from .batched import DEFAULT_POLICY, launch_shape
        t_auto, l_auto = launch_shape(code, m, n, vectors, DEFAULT_POLICY)
        threads = t_auto if threads is None else threads
        lanes = l_auto if lanes is None else lanes
    st = N._hip.jacobi_lds(code, b, m, n, 0, 0, 0, m, n, N.ptr(x), N.ptr(u), N.ptr(s), N.ptr(vh),
                           N.ptr(info), float(tol or 0.0), int(max_sweeps), flags, int(lanes), int(threads),
                           N.ptr(ws), wsb, N.stream_of(x))
    N.check(st, 'psvd_jacobi_lds')
    return u, s, vh, info


def jacobi_grid(x, *, tol=None, max_sweeps=60, scale=False, cut='reference', vectors=True):
    """One-sided Jacobi SVD of each tall matrix of ``x`` (b, m, n) too large for LDS: one
    single launch per matrix (grid barrier per round).  Same outputs as ``jacobi_lds_uniform``."""
    b, m, n = x.shape
    if m < n:
        raise ValueError('jacobi_grid takes tall matrices (m >= n)')
    x = x.contiguous()
    code = N.dtype_code(x.dtype)
    flags = (N.SCALE if scale else 0) | (N.CUT_FLOOR if cut == 'floor' else 0) | (0 if vectors else N.NO_VECTORS)
    dev = x.device
    u = torch.empty_like(x) if vectors else None
    s = torch.empty(b, n, dtype=x.real.dtype if x.is_complex() else x.dtype, device=dev)
    vh = torch.empty(b, n, n, dtype=x.dtype, device=dev) if vectors else None
    info = torch.empty(b, 2, dtype=torch.int32, device=dev)
    wsb = N._hip.jacobi_grid_workspace_bytes(code, m, n)
    ws = N.workspace(wsb, dev)
    stream = N.stream_of(x)
    for i in range(b):
        st = N._hip.jacobi_grid(code, m, n, N.ptr(x[i]), N.ptr(u[i]) if vectors else 0, N.ptr(s[i]),
                                N.ptr(vh[i]) if vectors else 0, N.ptr(info[i]), float(tol or 0.0),
                                int(max_sweeps), flags, N.ptr(ws), wsb, stream)
        N.check(st, 'psvd_jacobi_grid')
    return u, s, vh, info


def jacobi_blocked(x, *, tol=None, max_sweeps=60, scale=False, cut='reference', vectors=True, inner_sweeps=1,
                   block_width=16, full_pairs=False):
    """Blocked one-sided Jacobi SVD of each tall matrix of ``x`` (b, m, n): column blocks of
    ``block_width`` (16 or 32), FP64 Gram subproblems of twice that size and MFMA column updates,
    one call per matrix.  A sweep's first tournament round sweeps all pairs of its subproblems, the
    later rounds the pairs across the two blocks (``full_pairs``: all pairs in every round).  Same
    outputs as ``jacobi_grid``."""
    b, m, n = x.shape
    if m < n:
        raise ValueError('jacobi_blocked takes tall matrices (m >= n)')
    if block_width not in (16, 32):
        raise ValueError(f'block_width must be 16 or 32, not {block_width!r}')
    x = x.contiguous()
    code = N.dtype_code(x.dtype)
    flags = (N.SCALE if scale else 0) | (N.CUT_FLOOR if cut == 'floor' else 0) | (0 if vectors else N.NO_VECTORS)
    flags |= N.BLOCK_WIDE if block_width == 32 else N.BLOCK_NARROW
    flags |= N.BLOCK_FULL_PAIRS if full_pairs else 0
    dev = x.device
    u = torch.empty_like(x) if vectors else None
    s = torch.empty(b, n, dtype=x.dtype, device=dev)
    vh = torch.empty(b, n, n, dtype=x.dtype, device=dev) if vectors else None
    info = torch.empty(b, 2, dtype=torch.int32, device=dev)
    wsb = N._hip.jacobi_blocked_workspace_bytes(code, m, n)
    ws = N.workspace(wsb, dev)
    stream = N.stream_of(x)
    for i in range(b):
        st = N._hip.jacobi_blocked(code, m, n, N.ptr(x[i]), N.ptr(u[i]) if vectors else 0, N.ptr(s[i]),
                                   N.ptr(vh[i]) if vectors else 0, N.ptr(info[i]), float(tol or 0.0),
                                   int(max_sweeps), flags, int(inner_sweeps), N.ptr(ws), wsb, stream)
        N.check(st, 'psvd_jacobi_blocked')
    return u, s, vh, info


def jacobi_any(x, method='auto', **kw):
    """K1 when the shape fits LDS; otherwise the blocked kernel (n >= 64) or the single-launch one.
    Complex matrices beyond LDS: one matrix takes the complex single-launch kernel, a batch the batched
    complex kernel on workspace regions (both the reference's rotation order).  method 'grid' or
    'blocked' forces a large-matrix kernel."""
    b, m, n = x.shape
    block_width = kw.pop('block_width', None) or 16
    if method == 'auto' and lds_fits(x.dtype, m, n, kw.get('vectors', True)):
        return jacobi_lds_uniform(x, **kw)
    kw.pop('lanes', None)
    kw.pop('threads', None)
    if x.is_complex():   # the reference's rotation order at any size
        if method not in ('auto', 'grid'):
            raise ValueError(f'complex matrices beyond LDS take the grid kernel, not {method!r}')
        if b > 1 and method == 'auto':   # a batch: the batched kernel on workspace regions
            return jacobi_lds_uniform(x, **kw)
        return jacobi_grid(x, **kw)       # one matrix: the single-launch kernel over the whole GPU
    if method == 'blocked' or (method == 'auto' and n >= 64):
        return jacobi_blocked(x, block_width=block_width, **kw)
    if method not in ('auto', 'grid'):
        raise ValueError(f'unknown Jacobi method {method!r}')
    return jacobi_grid(x, **kw)


def jacobi_svd(A, tol=None, max_sweeps=30, *, cut='reference'):
    """One-sided Jacobi, round-robin ordering, batched (primitives.py:316-367).

    Returns U (b, m, n), s (b, n) descending, V (b, n, n), sweeps -- the reference's convention
    (V, not V^H).  Raises RuntimeError on non-convergence or failed completion."""
    b = A.shape[0]
    u, s, vh, info = jacobi_any(A, tol=tol, max_sweeps=max_sweeps, cut=cut)
    info = info.cpu()
    N.raise_for_info(info, f'one-sided Jacobi did not converge in {max_sweeps} sweeps')
    return u, s, vh.mH, int(info[:, 0].max()) if b else 0


def cqr_fits(dtype, m, n):
    """True when the complex LDS-resident QR takes an m x n matrix of this dtype."""
    st, _ = N._hip.cqr_lds_smem_bytes(N.dtype_code(dtype), m, n)
    return st == N.OK


def cqr(A):
    """Householder QR of the tall complex batch A (b, m, n): Q (b, m, n), R (b, n, n)
    (primitives.py:41-89, complex branch), each matrix in LDS or, beyond it, in a workspace region."""
    b, m, n = A.shape
    if m < n:
        raise ValueError('qr takes tall matrices (m >= n)')
    A = A.contiguous()
    code = N.dtype_code(A.dtype)
    q = torch.empty(b, m, n, dtype=A.dtype, device=A.device)
    r = torch.empty(b, n, n, dtype=A.dtype, device=A.device)
    wsb = N._hip.cqr_workspace_bytes(code, b, m, n)
    ws = N.workspace(wsb, A.device) if wsb else None
    N.check(N._hip.cqr(code, b, m, n, N.ptr(A), N.ptr(q), N.ptr(r), N.ptr(ws) if wsb else 0, wsb, N.stream_of(A)),
            'psvd_cqr')
    return q, r


def cmatmul(a, b, *, ha=False, hb=False):
    """Batched complex product op(a) @ op(b) on the device (FP64 sums); op(x) is x, or x^H when ha /
    hb is set (a is then stored (B, k, m), b (B, n, k))."""
    a, b = a.contiguous(), b.contiguous()
    B = a.shape[0]
    m, k = (a.shape[2], a.shape[1]) if ha else (a.shape[1], a.shape[2])
    n = b.shape[1] if hb else b.shape[2]
    c = torch.empty(B, m, n, dtype=a.dtype, device=a.device)
    N.check(N._hip.cmatmul(N.dtype_code(a.dtype), int(ha), int(hb), B, m, k, n, N.ptr(a), N.ptr(b), N.ptr(c),
                           N.stream_of(a)), 'psvd_cmatmul')
    return c


def hermitize_(a):
    """a <- (a + a^H) / 2 in place for a contiguous complex (b, n, n) device tensor."""
    N.check(N._hip.chermitize(N.dtype_code(a.dtype), a.shape[0], a.shape[-1], N.ptr(a), N.stream_of(a)),
            'psvd_chermitize')
    return a


def qr(A, *, gemm='f32', want_q=True):
    """Householder QR of the tall batch A (b, m, n): Q (b, m, n), R (b, n, n) (primitives.py:74-89).

    gemm: 'f32' or 'bf16x3' arithmetic for the FP32 panel updates (FP64 inputs use FP64).
    want_q=False skips forming Q (returns None for it).  Complex inputs take the LDS-resident
    complex kernel (``cqr``), for the shapes it holds."""
    b, m, n = A.shape
    if m < n:
        raise ValueError('qr takes tall matrices (m >= n)')
    if A.is_complex():
        q, r = cqr(A)
        return (q if want_q else None), r
    code = N.dtype_code(A.dtype)
    if A.stride(-1) != 1:
        A = A.contiguous()
    q = torch.empty(b, m, n, dtype=A.dtype, device=A.device) if want_q else None
    r = torch.empty(b, n, n, dtype=A.dtype, device=A.device)
    wsb = N._hip.qr_workspace_bytes(code, b, m, n)
    ws = N.workspace(wsb, A.device)
    gmode = {'f32': 0, 'bf16x3': 1}[gemm]
    st = N._hip.qr(code, gmode, b, m, n, N.ptr(A), A.stride(-2), A.stride(0) if b > 1 else 0, N.ptr(q), N.ptr(r),
                   N.ptr(ws), wsb, N.stream_of(A))
    N.check(st, 'psvd_qr')
    return q, r


# --- per-matrix reductions and scalings (batched utility kernels) -----------------------------
def _mat3(x):
    x3 = x if x.dim() == 3 else x.unsqueeze(0)
    if x3.stride(-1) != 1:
        x3 = x3.contiguous()
    return x3, x3.stride(-2), (x3.stride(0) if x3.shape[0] > 1 else 0)


def reduce(x, op, y=None):
    """Per-matrix reduction of (b, r, c) device tensors into FP64 device scalars:
    'maxabs' -> (b,), 'fro' -> (b,), 'maxdiff' -> (b, 2) = [max|x-y|, max|x|]; complex inputs take
    moduli for 'maxabs' and 'maxdiff' ('fro' is the same sum of squares over the real view)."""
    def real_view(t):
        tr = torch.view_as_real(t.contiguous())
        return tr.reshape(*tr.shape[:-2], tr.shape[-2] * 2)
    if x.is_complex():
        op = {'maxabs': 'maxmod', 'maxdiff': 'maxdiffmod', 'fro': 'fro'}[op]
        x = real_view(x)
        y = None if y is None else real_view(y)
    x3, ldx, sx = _mat3(x)
    b, rows, cols = x3.shape
    code = {'maxabs': 0, 'fro': 1, 'maxdiff': 2, 'maxmod': 3, 'maxdiffmod': 4}[op]
    pair = code in (2, 4)
    out = torch.empty(2 * b if pair else b, dtype=torch.float64, device=x.device)
    if pair:
        y3, ldy, sy = _mat3(y)
    else:
        y3, ldy, sy = None, 0, 0
    st = N._hip.reduce(N.dtype_code(x.dtype), code, b, rows, cols, N.ptr(x3), ldx, sx, N.ptr(y3), ldy, sy,
                       N.ptr(out), N.stream_of(x))
    N.check(st, 'psvd_reduce')
    return out.view(2, b).mT if pair else out


def scale(x, d=None, *, invert=False, s=1.0, out=None):
    """out_b = x_b / d_b (invert; x / scale as the reference rounds it) or (s * d_b) * x_b.  A complex
    x is scaled through its real view (both parts by the real d_b)."""
    if x.is_complex():
        if out is not None:
            raise ValueError('complex scaling returns a new tensor (out is not taken)')
        xr = torch.view_as_real(x.contiguous())
        flat = xr.reshape(*xr.shape[:-2], xr.shape[-2] * 2)
        res = scale(flat, d, invert=invert, s=s)
        return torch.view_as_complex(res.reshape(xr.shape))
    x3, ldx, sx = _mat3(x)
    b, rows, cols = x3.shape
    if out is None:
        out = torch.empty_like(x3)
    o3, ldo, so = _mat3(out)
    st = N._hip.scale(N.dtype_code(x.dtype), b, rows, cols, N.ptr(x3), ldx, sx, N.ptr(o3), ldo, so,
                      N.ptr(d), int(invert), float(s), N.stream_of(x))
    N.check(st, 'psvd_scale')
    return out if out.dim() == x.dim() else out.view_as(x)


def symmetrize_(a):
    """a <- (a + a^T) / 2 in place for a (b, n, n) or (n, n) device tensor."""
    a3, lda, sa = _mat3(a)
    st = N._hip.symmetrize(N.dtype_code(a.dtype), a3.shape[0], a3.shape[-1], N.ptr(a3), lda, sa, N.stream_of(a))
    N.check(st, 'psvd_symmetrize')
    return a


EIGH_LDS_CAP = 160 * 1024 - 4096


def eigh_lds_fits(dtype, n):
    """Whether the LDS-resident eigensolver holds an n x n matrix (G, V^T and the round tables)."""
    size = 4 if dtype == torch.float32 else 8
    np_ = n + (n & 1)
    b = size * 2 * np_ * (np_ + 1) + 16 * (np_ // 2) + 8 * (np_ // 2) + 12 * np_
    return ((b + 15) & ~15) <= EIGH_LDS_CAP


def ceigh_fits(dtype, n):
    """Whether the complex two-sided Jacobi kernel holds an n x n Hermitian matrix in LDS."""
    st, _ = N._hip.ceigh_lds_smem_bytes(N.dtype_code(dtype), n)
    return st == N.OK


def jacobi_eigh_raw(H, tol=None, max_sweeps=30, *, method='auto', inner_sweeps=1):
    """Two-sided Jacobi of the symmetric (or, complex, Hermitian) batch H (b, n, n): w ascending,
    V, info (device).

    method 'cyclic' keeps the reference's rotation order (LDS kernel, or the per-round HBM path
    for large n); 'blocked' groups rotations into 64 x 64 subproblems applied with MFMA; 'auto'
    takes the LDS kernel when the matrix fits and the blocked path otherwise.  Complex H takes the
    complex two-sided kernel (the reference's complex rotation, primitives.py:398-417), in LDS or,
    beyond it, in a workspace region."""
    b, n, _ = H.shape
    H = H.contiguous()
    code = N.dtype_code(H.dtype)
    if H.is_complex():
        w = torch.empty(b, n, dtype=H.real.dtype, device=H.device)
        v = torch.empty(b, n, n, dtype=H.dtype, device=H.device)
        info = torch.empty(b, 2, dtype=torch.int32, device=H.device)
        wsb = N._hip.ceigh_workspace_bytes(code, b, n)
        ws = N.workspace(wsb, H.device) if wsb else None
        st = N._hip.ceigh(code, b, n, N.ptr(H), N.ptr(w), N.ptr(v), N.ptr(info), float(tol or 0.0),
                          int(max_sweeps), N.ptr(ws) if wsb else 0, wsb, N.stream_of(H))
        N.check(st, 'psvd_ceigh')
        return w, v, info
    if method == 'auto':
        method = 'cyclic' if eigh_lds_fits(H.dtype, n) else 'blocked'
    if method not in ('cyclic', 'blocked'):
        raise ValueError(f'unknown eigensolver method {method!r}')
    w = torch.empty(b, n, dtype=H.dtype, device=H.device)
    v = torch.empty(b, n, n, dtype=H.dtype, device=H.device)
    info = torch.empty(b, 2, dtype=torch.int32, device=H.device)
    if method == 'cyclic':
        wsb = N._hip.eigh_workspace_bytes(code, n)
        ws = N.workspace(wsb, H.device)
        st = N._hip.eigh(code, b, n, N.ptr(H), N.ptr(w), N.ptr(v), N.ptr(info), float(tol or 0.0), int(max_sweeps),
                         N.ptr(ws), wsb, N.stream_of(H))
        N.check(st, 'psvd_eigh')
    else:
        wsb = N._hip.eigh_blocked_workspace_bytes(code, n)
        ws = N.workspace(wsb, H.device)
        st = N._hip.eigh_blocked(code, b, n, N.ptr(H), N.ptr(w), N.ptr(v), N.ptr(info), float(tol or 0.0),
                                 int(max_sweeps), int(inner_sweeps), N.ptr(ws), wsb, N.stream_of(H))
        N.check(st, 'psvd_eigh_blocked')
    return w, v, info


def jacobi_eigh(H, tol=None, max_sweeps=30, *, method='auto', inner_sweeps=1):
    """Hermitian (real symmetric) H (b, n, n) -> w (b, n) ascending, V (b, n, n), sweeps
    (primitives.py:370-427).  Raises RuntimeError on non-convergence."""
    w, v, info = jacobi_eigh_raw(H, tol, max_sweeps, method=method, inner_sweeps=inner_sweeps)
    info = info.cpu()
    N.raise_for_info(info, f'Hermitian Jacobi did not converge in {max_sweeps} sweeps')
    return w, v, int(info[:, 0].max()) if H.shape[0] else 0


def cholesky(G, *, gemm='f32'):
    """G (b, n, n) symmetric (complex: Hermitian) -> L lower (G = L L^H), fail (b,) bool
    (primitives.py:145-159).  Complex G is factored by one workgroup per matrix, in LDS or, beyond
    it, in a workspace region."""
    b, n, _ = G.shape
    if G.is_complex():
        G = G.contiguous()
        L = torch.empty_like(G)
        fail = torch.empty(b, dtype=torch.int32, device=G.device)
        code = N.dtype_code(G.dtype)
        wsb = N._hip.ccholesky_workspace_bytes(code, b, n)
        ws = N.workspace(wsb, G.device) if wsb else None
        N.check(N._hip.ccholesky(code, b, n, N.ptr(G), N.ptr(L), N.ptr(fail), N.ptr(ws) if wsb else 0, wsb,
                                 N.stream_of(G)), 'psvd_ccholesky')
        return L, fail != 0
    L = G.contiguous().clone()
    fail = torch.empty(b, dtype=torch.int32, device=G.device)
    st = N._hip.cholesky(N.dtype_code(G.dtype), {'f32': 0, 'bf16x3': 1}.get(gemm, 0), b, n, N.ptr(L), n, n * n,
                         N.ptr(fail), N.stream_of(G))
    N.check(st, 'psvd_cholesky')
    return L, fail != 0


def trsm_right_(B, L, *, trans, gemm='f32'):
    """In place: B <- B L^-T (trans) or B <- B L^-1 for B (b, rows, n), L (b, n, n) lower."""
    b, rows, n = B.shape
    if (B.stride(-1) != 1 and n > 1) or (L.stride(-1) != 1 and n > 1):
        raise ValueError('trsm_right_ takes row-contiguous operands')
    ldb = B.stride(-2) if rows > 1 else n
    ldl = L.stride(-2) if n > 1 else 1
    st = N._hip.trsm_right(N.dtype_code(B.dtype), {'f32': 0, 'bf16x3': 1}.get(gemm, 0), b, int(trans), n, rows,
                           N.ptr(L), ldl, L.stride(0) if b > 1 else 0, N.ptr(B), max(ldb, n),
                           B.stride(0) if b > 1 else 0, N.stream_of(B))
    N.check(st, 'psvd_trsm_right')
    return B


def csolve_fits(dtype, m, n):
    """Whether psvd_csolve_chol_lds holds an m x n block and its n x n factor in LDS."""
    st, _ = N._hip.csolve_lds_smem_bytes(N.dtype_code(dtype), m, n)
    return st == N.OK


def csolve_chol_(X, L):
    """In place for complex X (b, m, n) and lower L (b, n, n): X <- X (L L^H)^-1, the reference's
    trsm_upper(L^H, trsm_lower(L, X^H))^H (solvers.py:159-160)."""
    b, m, n = X.shape
    if not (X.is_contiguous() and L.is_contiguous()):
        raise ValueError('csolve_chol_ takes contiguous operands')
    code = N.dtype_code(X.dtype)
    wsb = N._hip.csolve_chol_workspace_bytes(code, b, m, n)
    ws = N.workspace(wsb, X.device) if wsb else None
    N.check(N._hip.csolve_chol(code, b, m, n, N.ptr(L), N.ptr(X), N.ptr(ws) if wsb else 0, wsb, N.stream_of(X)),
            'psvd_csolve_chol')
    return X


def trsm_lower(L, B):
    """Solve L X = B (primitives.py:162-169) as X^T = B^T L^-T."""
    Xt = B.mT.contiguous()
    trsm_right_(Xt, L.contiguous(), trans=True)
    return Xt.mT


def trsm_upper(U, B):
    """Solve U X = B with U upper (primitives.py:172-179): U = L^T, X^T = B^T L^-1."""
    Xt = B.mT.contiguous()
    trsm_right_(Xt, U.mT.contiguous(), trans=False)
    return Xt.mT


def axpby_(alpha, x, beta, y):
    """y <- alpha * x + beta * y in place (batched)."""
    x3, ldx, sx = _mat3(x)
    y3, ldy, sy = _mat3(y)
    b, rows, cols = y3.shape
    st = N._hip.axpby(N.dtype_code(y.dtype), b, rows, cols, float(alpha), N.ptr(x3), ldx, sx, float(beta), N.ptr(y3),
                      ldy, sy, N.stream_of(y))
    N.check(st, 'psvd_axpby')
    return y


def sort_desc(keys, x):
    """keys (b, n) float64, x (b, r, n) -> (s (b, n) in x's dtype = keys sorted descending and
    clamped at 0, y = x with its columns in the same order)."""
    x3, ldx, sx = _mat3(x)
    b, rows, n = x3.shape
    keys = keys.contiguous()
    if tuple(keys.shape) != (b, n) or keys.dtype != torch.float64:
        raise ValueError('sort_desc keys must be (batch, n) float64')
    s = torch.empty(b, n, dtype=x.dtype, device=x.device)
    y = torch.empty(b, rows, n, dtype=x.dtype, device=x.device)
    perm = torch.empty(b, n, dtype=torch.int32, device=x.device)
    st = N._hip.sort_desc(N.dtype_code(x.dtype), b, rows, n, N.ptr(keys), N.ptr(s), N.ptr(x3), ldx, sx, N.ptr(y),
                          y.stride(1), y.stride(0), N.ptr(perm), N.stream_of(x))
    N.check(st, 'psvd_sort_desc')
    return s, y


def scale_columns_(x, d):
    """In place: x[b, :, c] /= d[b, c] wherever d[b, c] > 0 (d (b, cols) in x's dtype)."""
    x3, ldx, sx = _mat3(x)
    b, rows, cols = x3.shape
    d = d.contiguous()
    if tuple(d.shape) != (b, cols) or d.dtype != x.dtype:
        raise ValueError('scale_columns_ takes d of shape (batch, cols) in the dtype of x')
    st = N._hip.scale_columns(N.dtype_code(x.dtype), b, rows, cols, N.ptr(x3), ldx, sx, N.ptr(d), cols,
                              N.stream_of(x))
    N.check(st, 'psvd_scale_columns')
    return x


class BlockJacobiSweeper:
    """Sweeps of the block-pair one-sided Jacobi (solvers.py:30-68) on W (b, m, n) and V (b, n, n),
    both row-contiguous and updated in place; device outputs per sweep: corr (b,) FP64 (the Gram
    test value), sigma (b, n) FP64 column norms, info (b, 2) int32 (largest pair sweep count,
    non-convergence flag)."""

    def __init__(self, w, v, block_size):
        b, m, n = w.shape
        if w.stride(-1) != 1 or v.stride(-1) != 1 or w.stride(-2) != n or v.stride(-2) != n:
            raise ValueError('block Jacobi takes row-contiguous W and V')
        self.w, self.v, self.block_size = w, v, int(block_size)
        dev = w.device
        self.info = torch.empty(b, 2, dtype=torch.int32, device=dev)
        self.corr = torch.empty(b, dtype=torch.float64, device=dev)
        self.sigma = torch.empty(b, n, dtype=torch.float64, device=dev)
        wsb = N._hip.block_jacobi_workspace_bytes(b, n)
        self.ws = N.workspace(wsb, dev)

    def sweep(self, tol, max_sweeps):
        w, v = self.w, self.v
        b, m, n = w.shape
        st = N._hip.block_jacobi_sweep(N.dtype_code(w.dtype), b, m, n, self.block_size, N.ptr(w), n, m * n,
                                       N.ptr(v), n, n * n, float(tol or 0.0), int(max_sweeps), N.ptr(self.info),
                                       N.ptr(self.corr), N.ptr(self.sigma), N.ptr(self.ws), self.ws.numel(),
                                       N.stream_of(w))
        N.check(st, 'psvd_block_jacobi_sweep')


def complete_columns_(u, nk):
    """Complete columns [nk_b, n) of u (b, m, n) in place (primitives.py:430-455); nk is a (b,)
    int32 device tensor.  Returns the (b,) status tensor (device, no synchronisation)."""
    b, m, n = u.shape
    status = torch.empty(b, dtype=torch.int32, device=u.device)
    if u.is_complex():
        if not u.is_contiguous():
            raise ValueError('complex completion takes a contiguous (b, m, n) tensor')
        N.check(N._hip.ccomplete_columns(N.dtype_code(u.dtype), b, m, n, N.ptr(u), N.ptr(nk), N.ptr(status),
                                         N.stream_of(u)), 'psvd_ccomplete_columns')
        return status
    wsb = N._hip.complete_workspace_bytes(b, m, n)
    ws = N.workspace(wsb, u.device)
    st = N._hip.complete_columns(N.dtype_code(u.dtype), b, m, n, N.ptr(u), u.stride(-2), u.stride(0) if b > 1 else 0,
                                 N.ptr(nk), N.ptr(status), N.ptr(ws), wsb, N.stream_of(u))
    N.check(st, 'psvd_complete_columns')
    return status


def complete_null_columns(u, keep):
    """Replace the columns of u (b, m, n) where keep (b, n) is False by completed basis vectors
    and keep the rest (primitives.py:430-455): the kept columns are moved to the front (stable),
    the missing ones completed in their original order by the completion kernel -- each against
    the kept and the already completed columns, as the reference's loop does -- and the columns
    moved back.  Raises RuntimeError if a residual vanishes (:451-452)."""
    if bool(keep.all()):
        return u
    b, m, n = u.shape
    keep = keep.to(device=u.device, dtype=torch.bool).reshape(b, n)
    idx = torch.arange(n, device=u.device).expand(b, n)
    perm = torch.argsort((~keep).to(torch.int8) * n + idx, dim=1)        # kept first, each part in order
    w = torch.gather(u, 2, perm[:, None, :].expand(b, m, n)).contiguous()
    status = complete_columns_(w, keep.sum(1).to(torch.int32))
    if bool((status != 0).any()):
        raise RuntimeError('null-space completion failed: supplied live columns are not independent')
    out = torch.empty_like(w)
    out.scatter_(2, perm[:, None, :].expand(b, m, n), w)
    return out


def _eps(dtype):
    """Machine epsilon of the (real part of the) dtype, as primitives.py's helper."""
    return torch.finfo(dtype.to_real() if dtype.is_complex else dtype).eps


def _real_dtype(dtype):
    return dtype.to_real() if dtype.is_complex else dtype


def _round_robin(n):
    """Rounds of the round-robin tournament on n (even) players (primitives.py:309-313): n - 1
    lists of n/2 disjoint pairs (min, max); seat 0 stays, the others rotate by one seat per round."""
    seats = list(range(n))
    for _ in range(n - 1):
        yield [(min(seats[i], seats[n - 1 - i]), max(seats[i], seats[n - 1 - i])) for i in range(n // 2)]
        seats = [seats[0], seats[-1]] + seats[1:-1]


def complete_basis(q, size):
    """Extend orthonormal columns q (b, m, k) to `size` columns (primitives.py:458-467)."""
    b, m, k = q.shape
    if size == k:
        return q
    if not k <= size <= m:
        raise ValueError('invalid basis size')
    out = torch.zeros(b, m, size, dtype=q.dtype, device=q.device)
    out[:, :, :k] = q
    nk = torch.full((b,), k, dtype=torch.int32, device=q.device)
    status = complete_columns_(out, nk)
    if bool((status != 0).any()):
        raise RuntimeError('null-space completion failed: supplied live columns are not independent')
    return out
