"""numpy restatement of the reference standalone hot path (test infrastructure only).

Every function follows the reference operation by operation, in the input dtype, vectorised over
the leading batch dimension, so that iteration counts and rounding behaviour track the reference:

    round_robin            svdbench/standalone/primitives.py:309-313
    householder_vec        svdbench/standalone/primitives.py:41-59
    apply_left             svdbench/standalone/primitives.py:62-65
    qr                     svdbench/standalone/primitives.py:74-89
    cholesky               svdbench/standalone/primitives.py:145-159
    trsm_lower/trsm_upper  svdbench/standalone/primitives.py:162-179
    jacobi_svd             svdbench/standalone/primitives.py:316-367
    jacobi_eigh            svdbench/standalone/primitives.py:370-427
    complete_null_columns  svdbench/standalone/primitives.py:430-455
    complete_basis         svdbench/standalone/primitives.py:458-467
    qdwh_weights           svdbench/standalone/solvers.py:143-149
    tsqr                   svdbench/standalone/solvers.py:71-96
    polar_qdwh             svdbench/standalone/solvers.py:122-177
    raw / svd              svdbench/standalone/solvers.py:224-311
    finish_columns         svdbench/standalone/solvers.py:18-21
    block_jacobi           svdbench/standalone/solvers.py:30-68
    gram_jacobi            svdbench/standalone/solvers.py:99-119 (the Jacobi branch)
    sketch / partial       svdbench/standalone/solvers.py:180-221 (randomized, subspace, block_krylov)

Shapes follow the reference: A is (b, m, n).  Real and complex dtypes are supported.
"""
import math

import numpy as np

ALGORITHMS = (
    'bidiag_qr', 'bidiag_dc', 'jacobi', 'block_jacobi', 'qr', 'tsqr',
    'gram_jacobi', 'gram_dc', 'polar_qdwh', 'polar_newton_schulz',
    'polar_halley', 'polar_polynomial', 'randomized', 'subspace', 'block_krylov', 'lanczos',
)
HOT_PATH = ('jacobi', 'tsqr', 'polar_qdwh')
CONSUMERS = ('qr', 'block_jacobi', 'gram_jacobi', 'randomized', 'subspace', 'block_krylov')
PARTIAL = ('randomized', 'subspace', 'block_krylov')


def real_dtype(dt):
    dt = np.dtype(dt)
    return {np.dtype(np.complex64): np.dtype(np.float32),
            np.dtype(np.complex128): np.dtype(np.float64)}.get(dt, dt)


def eps(dt):
    return float(np.finfo(real_dtype(dt)).eps)


def tiny(dt):
    return np.finfo(real_dtype(dt)).tiny


def mh(x):
    """Conjugate transpose of the last two axes (torch ``.mH``)."""
    return np.conj(np.swapaxes(x, -1, -2)) if np.iscomplexobj(x) else np.swapaxes(x, -1, -2)


def _sign(x):
    return np.sign(x)


# ---------------------------------------------------------------------------------------------
# Tournament schedule
# ---------------------------------------------------------------------------------------------
def round_robin(n):
    """n-1 rounds of n/2 disjoint (min, max) pairs (primitives.py:309-313)."""
    players = list(range(n))
    for _ in range(n - 1):
        yield [(min(players[i], players[n - 1 - i]), max(players[i], players[n - 1 - i]))
               for i in range(n // 2)]
        players = [players[0]] + [players[-1]] + players[1:-1]


# ---------------------------------------------------------------------------------------------
# Householder QR
# ---------------------------------------------------------------------------------------------
def householder_vec(x):
    """x (b, k) -> v (b, k) with v[:,0]=1, tau (b,), beta (b,) real, LAPACK xLARFG convention
    (primitives.py:41-59): (I - tau v v^H)^H x = beta e1."""
    b, k = x.shape
    rd = real_dtype(x.dtype)
    alpha = x[:, 0]
    if k > 1:
        xnorm = np.sqrt((np.abs(x[:, 1:]) ** 2).sum(-1))
    else:
        xnorm = np.zeros(b, dtype=rd)
    cplx = np.iscomplexobj(x)
    ar = alpha.real if cplx else alpha
    ai = alpha.imag if cplx else np.zeros_like(ar)
    beta = -_sign(ar + (ar == 0)) * np.hypot(np.abs(alpha), xnorm)
    trivial = (xnorm == 0) & (ai == 0)
    beta = np.where(trivial, ar, beta).astype(rd)
    bt = beta.astype(x.dtype)
    denom = alpha - bt
    denom = np.where(trivial, np.ones_like(denom), denom)
    tau = (bt - alpha) / np.where(beta != 0, bt, np.ones_like(bt))
    tau = np.where(trivial, np.zeros_like(tau), tau)
    v = x / denom[:, None]
    v[:, 0] = 1.0
    v = np.where(trivial[:, None], np.zeros_like(v), v)
    return v, tau, beta


def apply_left(A, v, tau):
    """A (b, k, c) <- (I - tau v v^H) A (primitives.py:62-65)."""
    w = (np.conj(v)[:, :, None] * A).sum(-2)
    return A - tau[:, None, None] * v[:, :, None] * w[:, None, :]


def qr(A):
    """Unblocked Householder QR with backward accumulation of the thin Q (primitives.py:74-89).
    Returns Q (b, m, n), R (b, n, n)."""
    A = np.array(A, copy=True)
    b, m, n = A.shape
    vs, taus = [], []
    for j in range(n):
        v, tau, beta = householder_vec(A[:, j:, j])
        A[:, j:, j + 1:] = apply_left(A[:, j:, j + 1:], v, np.conj(tau))
        A[:, j, j] = beta.astype(A.dtype)
        A[:, j + 1:, j] = 0
        vs.append(v)
        taus.append(tau)
    R = np.triu(A[:, :n, :])
    Q = np.zeros((b, m, n), dtype=A.dtype)
    Q[:, :n, :] = np.eye(n, dtype=A.dtype)
    for j in range(n - 1, -1, -1):
        Q[:, j:, :] = apply_left(Q[:, j:, :], vs[j], taus[j])
    return Q, R


# ---------------------------------------------------------------------------------------------
# Cholesky and triangular solves
# ---------------------------------------------------------------------------------------------
def cholesky(G):
    """Unblocked column Cholesky G = L L^H with a per-matrix failure mask (primitives.py:145-159)."""
    b, n, _ = G.shape
    L = np.zeros_like(G)
    fail = np.zeros(b, dtype=bool)
    cplx = np.iscomplexobj(G)
    for j in range(n):
        s = (np.abs(L[:, j, :j]) ** 2).sum(-1)
        piv = (G[:, j, j].real if cplx else G[:, j, j]) - s
        fail |= piv <= 0
        ljj = np.sqrt(np.maximum(piv, tiny(G.dtype)))
        L[:, j, j] = ljj.astype(G.dtype)
        if j + 1 < n:
            r = G[:, j + 1:, j] - (L[:, j + 1:, :j] * np.conj(L[:, j, :j])[:, None, :]).sum(-1)
            L[:, j + 1:, j] = r / ljj[:, None].astype(G.dtype)
    return L, fail


def trsm_lower(L, B):
    """Solve L X = B by forward substitution, row by row (primitives.py:162-169)."""
    n = L.shape[-1]
    X = np.zeros_like(B)
    for i in range(n):
        acc = (L[:, i, :i][:, :, None] * X[:, :i, :]).sum(-2) if i > 0 else 0
        X[:, i, :] = (B[:, i, :] - acc) / L[:, i, i][:, None]
    return X


def trsm_upper(U, B):
    """Solve U X = B by back substitution, row by row (primitives.py:172-179)."""
    n = U.shape[-1]
    X = np.zeros_like(B)
    for i in range(n - 1, -1, -1):
        acc = (U[:, i, i + 1:][:, :, None] * X[:, i + 1:, :]).sum(-2) if i + 1 < n else 0
        X[:, i, :] = (B[:, i, :] - acc) / U[:, i, i][:, None]
    return X


# ---------------------------------------------------------------------------------------------
# Jacobi
# ---------------------------------------------------------------------------------------------
def jacobi_svd(A, tol=None, max_sweeps=30):
    """One-sided Jacobi, round-robin ordering, batched (primitives.py:316-367).
    Returns U (b,m,n), s (b,n) descending, V (b,n,n), sweeps."""
    b, m, n = A.shape
    padded = n % 2 == 1
    W = np.array(A, copy=True)
    if padded:
        W = np.concatenate([W, np.zeros((b, m, 1), dtype=A.dtype)], -1)
        n += 1
    V = np.broadcast_to(np.eye(n, dtype=W.dtype), (b, n, n)).copy()
    tol = tol or 10 * eps(A.dtype)
    rd = real_dtype(A.dtype)
    e = eps(A.dtype)
    sched = [(np.array([p for p, _ in prs]), np.array([q for _, q in prs])) for prs in round_robin(n)]
    sweeps = 0
    off = rd.type(0)
    for sweeps in range(1, max_sweeps + 1):
        off = rd.type(0)
        for P, Q in sched:
            wp, wq = W[:, :, P], W[:, :, Q]
            app = (np.abs(wp) ** 2).sum(1)
            aqq = (np.abs(wq) ** 2).sum(1)
            apq = (np.conj(wp) * wq).sum(1)
            den = np.sqrt(app) * np.sqrt(aqq)
            ratio = np.abs(apq) / np.maximum(den, tiny(rd))
            floor = e ** 2 * np.maximum(app, aqq).max(-1, keepdims=True)
            active = (ratio > tol) & (np.minimum(app, aqq) > floor)
            off = np.maximum(off, np.where(active, ratio, np.zeros_like(ratio)).max())
            a = np.where(active, np.abs(apq), np.ones_like(den))
            tau = (aqq - app) / (2 * a)
            t = _sign(tau) / (np.abs(tau) + np.hypot(tau, np.ones_like(tau)))
            t = np.where(tau == 0, np.ones_like(t), t)
            c = 1 / np.sqrt(1 + t * t)
            s = (t * c).astype(A.dtype) * (apq / a.astype(A.dtype))
            c = np.where(active, c, np.ones_like(c)).astype(A.dtype)
            s = np.where(active, s, np.zeros_like(s))
            c = c[:, None, :]
            s = s[:, None, :]
            W[:, :, P], W[:, :, Q] = c * wp - np.conj(s) * wq, s * wp + c * wq
            vp, vq = V[:, :, P], V[:, :, Q]
            V[:, :, P], V[:, :, Q] = c * vp - np.conj(s) * vq, s * vp + c * vq
        if off <= tol:
            break
    if off > tol:
        raise RuntimeError(f'one-sided Jacobi did not converge in {max_sweeps} sweeps')
    s_vals = np.sqrt((np.abs(W) ** 2).sum(1))
    if padded:
        W, V, s_vals = W[:, :, :-1], V[:, :-1, :-1], s_vals[:, :-1]
    order = np.argsort(-s_vals, axis=1, kind='stable')
    s_vals = np.take_along_axis(s_vals, order, 1)
    W = np.take_along_axis(W, order[:, None, :], 2)
    V = np.take_along_axis(V, order[:, None, :], 2)
    cut = e * max(m, n) * s_vals[:, :1]
    keep = s_vals > cut
    U = W / np.where(keep, s_vals, np.ones_like(s_vals))[:, None, :]
    U = complete_null_columns(U, keep)
    s_vals = np.where(keep, s_vals, np.zeros_like(s_vals))
    return U, s_vals, V, sweeps


def jacobi_eigh(H, tol=None, max_sweeps=30):
    """Two-sided cyclic Hermitian Jacobi, round-robin ordering (primitives.py:370-427).
    Returns w (b,n) ascending, V (b,n,n), sweeps."""
    b, n, _ = H.shape
    padded = n % 2 == 1
    G = np.array(H, copy=True)
    if padded:
        Z = np.zeros((b, n + 1, n + 1), dtype=H.dtype)
        Z[:, :n, :n] = G
        G = Z
        n += 1
    V = np.broadcast_to(np.eye(n, dtype=G.dtype), (b, n, n)).copy()
    tol = tol or 10 * eps(H.dtype)
    rd = real_dtype(H.dtype)
    cplx = np.iscomplexobj(H)
    bidx = np.arange(b)[:, None]
    sched = [(np.array([p for p, _ in prs]), np.array([q for _, q in prs])) for prs in round_robin(n)]
    sweeps = 0
    off = rd.type(0)
    for sweeps in range(1, max_sweeps + 1):
        off = rd.type(0)
        for P, Q in sched:
            gpp = G[bidx, P, P]
            gqq = G[bidx, Q, Q]
            gpq = G[bidx, P, Q]
            if cplx:
                gpp, gqq = gpp.real, gqq.real
            a = np.abs(gpq)
            scale = np.broadcast_to(np.abs(G).max((-2, -1))[:, None], gpp.shape)
            active = a > tol * scale
            off = np.maximum(off, (a / np.maximum(scale, tiny(rd))).max())
            a_safe = np.where(active, a, np.ones_like(a))
            tau = (gqq - gpp) / (2 * a_safe)
            t = _sign(tau) / (np.abs(tau) + np.hypot(tau, np.ones_like(tau)))
            t = np.where(tau == 0, np.ones_like(t), t)
            c = 1 / np.sqrt(1 + t * t)
            s = (t * c).astype(G.dtype) * (gpq / a_safe.astype(G.dtype))
            c = np.where(active, c, np.ones_like(c)).astype(G.dtype)
            s = np.where(active, s, np.zeros_like(s))
            cc, ss = c[:, None, :], s[:, None, :]
            gp, gq = G[:, :, P], G[:, :, Q]
            G[:, :, P], G[:, :, Q] = cc * gp - np.conj(ss) * gq, ss * gp + cc * gq
            vp, vq = V[:, :, P], V[:, :, Q]
            V[:, :, P], V[:, :, Q] = cc * vp - np.conj(ss) * vq, ss * vp + cc * vq
            cc, ss = c[:, :, None], s[:, :, None]
            rp, rq = G[:, P, :], G[:, Q, :]
            G[:, P, :], G[:, Q, :] = cc * rp - ss * rq, np.conj(ss) * rp + cc * rq
        if off <= tol:
            break
    if off > tol:
        raise RuntimeError(f'Hermitian Jacobi did not converge in {max_sweeps} sweeps')
    w = np.diagonal(G, axis1=-2, axis2=-1)
    w = w.real.copy() if cplx else w.copy()
    if padded:
        w, V = w[:, :-1], V[:, :-1, :-1]
    order = np.argsort(w, axis=-1, kind='stable')
    return np.take_along_axis(w, order, 1), np.take_along_axis(V, order[:, None, :], 2), sweeps


def complete_null_columns(u, keep):
    """Complete missing columns with projected coordinate vectors (primitives.py:430-455)."""
    if keep.all():
        return u
    b, m, n = u.shape
    u = np.where(keep[:, None, :], u, np.zeros_like(u))
    batch = np.arange(b)
    for j in range(n):
        missing = ~keep[:, j]
        if not missing.any():
            continue
        leverage = 1 - (np.abs(u) ** 2).sum(-1)
        idx = leverage.argmax(-1)
        v = np.zeros((b, m, 1), dtype=u.dtype)
        v[batch, idx, 0] = 1
        for _ in range(2):
            v = v - u @ (mh(u) @ v)
        length = np.sqrt((np.abs(v) ** 2).sum(-2, keepdims=True))
        if (length[missing] < 10 * eps(u.dtype)).any():
            raise RuntimeError('null-space completion failed: supplied live columns are not independent')
        v = v / np.maximum(length, tiny(u.dtype))
        u[:, :, j] = np.where(missing[:, None], v[:, :, 0], u[:, :, j])
    return u


def complete_basis(q, size):
    """Extend an orthonormal basis to ``size`` columns (primitives.py:458-467)."""
    if size == q.shape[-1]:
        return q
    b, m, k = q.shape
    if not k <= size <= m:
        raise ValueError('invalid basis size')
    z = np.zeros((b, m, size - k), dtype=q.dtype)
    keep = np.broadcast_to(np.arange(size)[None, :] < k, (b, size))
    return complete_null_columns(np.concatenate([q, z], -1), keep)


# ---------------------------------------------------------------------------------------------
# Solvers
# ---------------------------------------------------------------------------------------------
def _eye(a, n):
    return np.broadcast_to(np.eye(n, dtype=a.dtype), (a.shape[0], n, n)).copy()


def _jacobi(x, stats, opts):
    u, s, v, sweeps = jacobi_svd(x, tol=opts.get('tol'), max_sweeps=opts.get('max_sweeps', 60))
    stats['jacobi_sweeps'] = max(stats.get('jacobi_sweeps', 0), sweeps)
    return u, s, v


def tsqr_leaf_bounds(m, n, rows):
    """Row ranges of the TSQR leaves (solvers.py:73-83): ``None`` when a single QR is used."""
    rows = max(rows, n)
    if m <= rows:
        return None
    leaves = []
    for start in range(0, m, rows):
        end = min(start + rows, m)
        if end - start < n:
            start = leaves.pop()[0]
        leaves.append((start, end))
    return leaves


def tsqr(a, rows, stats):
    """Flat leaves then a binary merge tree with explicit Q back-transformation (solvers.py:71-96)."""
    b, m, n = a.shape
    bounds = tsqr_leaf_bounds(m, n, rows)
    if bounds is None:
        return qr(a)
    leaves = [qr(a[:, s:e]) for s, e in bounds]
    stats['tsqr_leaves'] = len(leaves)
    nodes = leaves
    while len(nodes) > 1:
        nxt = []
        for i in range(0, len(nodes), 2):
            if i + 1 == len(nodes):
                nxt.append(nodes[i])
                continue
            q1, r1 = nodes[i]
            q2, r2 = nodes[i + 1]
            q, r = qr(np.concatenate([r1, r2], -2))
            nxt.append((np.concatenate([q1 @ q[:, :n], q2 @ q[:, n:]], -2), r))
        nodes = nxt
    return nodes[0]


def qdwh_weights(lower):
    """Dynamically weighted Halley coefficients (a, b, c) and the next lower bound for the
    current bound ``lower`` on sigma_min/sigma_max (solvers.py:143-149; Nakatsukasa, Bai and
    Gygi, SIAM J. Matrix Anal. Appl. 31(5), 2010)."""
    l2 = lower * lower
    dd = (4 * (1 - l2) / (l2 * l2)) ** (1 / 3)
    sd = math.sqrt(1 + dd)
    aa = sd + 0.5 * math.sqrt(max(8 - 4 * dd + 8 * (2 - l2) / (l2 * sd), 0))
    bb = (aa - 1) ** 2 / 4
    cc = aa + bb - 1
    nxt = min(1., lower * (aa + bb * l2) / (1 + cc * l2))
    return aa, bb, cc, nxt


def polar_qdwh(a, stats, opts):
    """QDWH polar decomposition followed by the Hermitian Jacobi eigensolve of H
    (solvers.py:122-177)."""
    n = a.shape[-1]
    norm = np.maximum(np.sqrt((np.abs(a) ** 2).sum((-2, -1), keepdims=True)),
                      tiny(a.dtype)).astype(real_dtype(a.dtype))
    x = a / norm
    ident = _eye(a, n)
    tol = opts.get('tol') or 10 * eps(a.dtype)
    lower = opts.get('lower_bound', eps(a.dtype))
    if not 0 < lower <= 1:
        raise ValueError('lower_bound must be in (0,1]')
    max_iter = opts.get('max_iter', 40)
    converged = False
    it = 0
    for it in range(max_iter):
        g = mh(x) @ x
        aa, bb, cc, lower = qdwh_weights(lower)
        if cc > 100:
            q, _ = qr(np.concatenate([math.sqrt(cc) * x, ident], -2))
            top, bot = q[:, :a.shape[-2]], q[:, a.shape[-2]:]
            xn = bb / cc * x + (aa - bb / cc) / math.sqrt(cc) * (top @ mh(bot))
        else:
            z = ident + cc * g
            chol, fail = cholesky((z + mh(z)) / 2)
            if fail.any():
                raise RuntimeError('polar Cholesky lost positive definiteness')
            invprod = mh(trsm_upper(mh(chol), trsm_lower(chol, mh(x))))
            xn = bb / cc * x + (aa - bb / cc) * invprod
        diff = np.abs(xn - x).max()
        x = xn
        if diff <= tol * max(np.abs(x).max(), 1):
            converged = True
            break
    if not converged:
        raise RuntimeError(f'polar_qdwh did not converge in {max_iter} iterations')
    h = mh(x) @ a
    h = (h + mh(h)) / 2
    w, v, sweeps = jacobi_eigh(h, max_sweeps=opts.get('max_sweeps', 60))
    s, v = np.maximum(w[:, ::-1], 0), v[:, :, ::-1]
    keep = s > eps(a.dtype) * max(a.shape[-2:]) * s[:, :1]
    u = complete_null_columns(x @ v, keep)
    s = np.where(keep, s, np.zeros_like(s))
    stats.update(polar_iterations=it + 1, eigen_sweeps=sweeps,
                 polar_fixed_point_error=float(np.abs(x @ (mh(x) @ x) - x).max()))
    stats['_polar_factor'] = x
    return u, s, v


def finish_columns(w, s):
    """U = W / sigma for sigma above eps * max(m, n) * sigma_1, completed elsewhere; sigma zeroed
    there (solvers.py:18-21)."""
    keep = s > eps(w.dtype) * max(w.shape[-2:]) * s[:, :1]
    u = w / np.where(keep, s, np.ones_like(s))[:, None, :]
    return complete_null_columns(u, keep), np.where(keep, s, np.zeros_like(s))


def block_jacobi(a, stats, opts):
    """Block-pair one-sided Jacobi (solvers.py:30-68): column groups of block_size, paired by the
    round-robin tournament; each pair's columns are replaced by W_pair V_pair from a full one-sided
    Jacobi SVD of W_pair (V_pair sorted by descending sigma); a sweep ends with the Gram test
    max |g_ij| / sqrt(g_ii g_jj) over live columns <= tol (40 eps by default)."""
    b, m, n = a.shape
    width = opts.get('block_size', 4)
    if width < 1:
        raise ValueError('block_size must be positive')
    groups = [list(range(j, min(j + width, n))) for j in range(0, n, width)]
    if len(groups) == 1:
        return _jacobi(a, stats, opts)
    if len(groups) % 2:
        groups.append([])
    w, v = a.copy(), _eye(a, n)
    tol = opts.get('tol') or 40 * eps(a.dtype)
    rdt = real_dtype(a.dtype)
    converged = False
    sweep = 0
    for sweep in range(opts.get('max_sweeps', 30)):
        for pairs in round_robin(len(groups)):
            for i, j in pairs:
                cols = groups[i] + groups[j]
                if len(cols) < 2:
                    continue
                _, _, vv = _jacobi(w[:, :, cols], stats, opts)
                w[:, :, cols] = w[:, :, cols] @ vv
                v[:, :, cols] = v[:, :, cols] @ vv
        g = mh(w) @ w
        d = np.maximum(np.diagonal(g, axis1=-2, axis2=-1).real, 0).astype(rdt)
        denom = np.sqrt(d)[:, :, None] * np.sqrt(d)[:, None, :]
        g[:, np.arange(n), np.arange(n)] = 0
        live = d > (eps(a.dtype) * max(m, n)) ** 2 * d.max(-1, keepdims=True)
        corr = np.where(live[:, :, None] & live[:, None, :], np.abs(g) / np.maximum(denom, tiny(rdt)), 0)
        if corr.max() <= tol:
            converged = True
            break
    if not converged:
        raise RuntimeError('block Jacobi did not converge')
    s = np.sqrt((np.abs(w) ** 2).sum(1))
    order = np.argsort(-s, axis=-1, kind='stable')
    s = np.take_along_axis(s, order, 1)
    w = np.take_along_axis(w, order[:, None, :], 2)
    v = np.take_along_axis(v, order[:, None, :], 2)
    u, s = finish_columns(w, s)
    stats['block_sweeps'] = sweep + 1
    return u, s, v


def gram_jacobi(a, stats, opts):
    """Eigen-decomposition of the Gram matrix (solvers.py:99-119, Jacobi branch): G = sym(A^H A),
    Hermitian Jacobi, eigenvalues descending; sigma = sqrt(w) above eps * n * w_1 (zero below,
    counted in gram_discarded_values), U = A V / sigma completed."""
    g = mh(a) @ a
    g = (g + mh(g)) / 2
    w, v, it = jacobi_eigh(g, max_sweeps=opts.get('max_sweeps', 60))
    stats['eigen_sweeps'] = it
    w, v = w[:, ::-1], v[:, :, ::-1]
    cut = eps(a.dtype) * a.shape[-1] * np.maximum(w[:, :1], 0)
    keep = w > cut
    s = np.where(keep, np.sqrt(np.maximum(w, 0)), 0).astype(w.dtype)
    u = (a @ v) / np.where(keep, s, np.ones_like(s))[:, None, :]
    u = complete_null_columns(u, keep)
    stats['gram_discarded_values'] = int((~keep).sum())
    return u, s, v


def sketch(a, columns, seed):
    """The Gaussian test matrix (b, n, columns) of solvers.py:180-182, drawn by torch's CPU generator
    (the reference's call on a CPU tensor) and returned as numpy."""
    import torch
    gen = torch.Generator().manual_seed(seed)
    tdt = {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64,
           np.dtype(np.complex64): torch.complex64, np.dtype(np.complex128): torch.complex128}[np.dtype(a.dtype)]
    return torch.randn(a.shape[0], a.shape[-1], columns, dtype=tdt, generator=gen).numpy()


def partial(a, stats, opts, method):
    """Randomized range finder with power (subspace) iterations or a block Krylov basis, then the
    Jacobi SVD of Q^H A (solvers.py:185-221)."""
    k = opts.get('rank', min(a.shape[-2:]))
    ell = min(k + opts.get('oversample', 8), min(a.shape[-2:]))
    omega = opts.get('initial_basis')
    if omega is None:
        omega = sketch(a, ell, opts.get('seed', 20261011))
    else:
        omega = np.asarray(omega)
        if omega.ndim == 2:
            omega = omega[None]
        if omega.shape != (a.shape[0], a.shape[-1], ell):
            raise ValueError('initial_basis shape must match batch, n, rank+oversample')
        omega = omega.astype(a.dtype)
        stats['warm_started'] = True
    q, _ = qr(a @ omega)
    bases = [q]
    iterations = opts.get('iterations', 2)
    for _ in range(iterations):
        z, _ = qr(mh(a) @ q)
        y = a @ z
        if method == 'block_krylov':
            joined = np.concatenate(bases, -1)
            available = a.shape[-2] - joined.shape[-1]
            if available <= 0:
                break
            for _ in range(2):
                y = y - joined @ (mh(joined) @ y)
            q, _ = qr(y[:, :, :min(ell, available)])
            bases.append(q)
        else:
            q, _ = qr(y)
    if method == 'block_krylov':
        q, _ = qr(np.concatenate(bases, -1))
    small = mh(q) @ a
    u, s, v = raw(small, 'jacobi', stats, opts)
    stats.update(sketch_columns=ell, range_columns=q.shape[-1], power_iterations=iterations)
    return (q @ u)[:, :, :k], s[:, :k], v[:, :, :k]


def raw(a, method, stats, opts):
    """Wide inputs are solved transposed, except by the partial methods; dispatch by name
    (solvers.py:224-261)."""
    if a.shape[-2] < a.shape[-1] and method not in PARTIAL:
        u, s, v = raw(mh(a), method, stats, opts)
        return v, s, u
    if method == 'jacobi':
        return _jacobi(a, stats, opts)
    if method == 'block_jacobi':
        return block_jacobi(a, stats, opts)
    if method in ('qr', 'tsqr'):
        q, r = qr(a) if method == 'qr' else tsqr(a, opts.get('tile_rows', 4 * a.shape[-1]), stats)
        u, s, v = _jacobi(r, stats, opts)
        return q @ u, s, v
    if method == 'gram_jacobi':
        return gram_jacobi(a, stats, opts)
    if method == 'polar_qdwh':
        return polar_qdwh(a, stats, opts)
    if method in PARTIAL:
        return partial(a, stats, opts, method)
    raise ValueError(f'the oracle restates only {HOT_PATH + CONSUMERS}, not {method!r}')


def svd(a, algorithm='jacobi', *, full_matrices=False, values_only=False, return_info=False, **opts):
    """Argument validation, max-abs scaling, dispatch, rescale and truncate or complete
    (solvers.py:264-311)."""
    a = np.asarray(a)
    if algorithm not in ALGORITHMS:
        raise ValueError(f'algorithm must be one of {ALGORITHMS}')
    if a.ndim < 2 or min(a.shape[-2:]) < 1:
        raise ValueError('nonempty matrix dimensions required')
    if a.dtype not in (np.float32, np.float64, np.complex64, np.complex128):
        raise TypeError('supported dtypes: float32/64, complex64/128')
    rank = opts.get('rank', min(a.shape[-2:]))
    if not 1 <= rank <= min(a.shape[-2:]):
        raise ValueError('rank outside [1,min(m,n)]')
    if full_matrices and rank != min(a.shape[-2:]):
        raise ValueError('full_matrices and truncated rank cannot be combined')
    for key in ('max_iter', 'max_sweeps', 'leaf_size'):
        if key in opts and opts[key] < 1:
            raise ValueError(f'{key} must be positive')
    for key in ('oversample', 'iterations'):
        if key in opts and opts[key] < 0:
            raise ValueError(f'{key} must be nonnegative')
    shape = a.shape[:-2]
    x = a.reshape(-1, *a.shape[-2:])
    if not np.isfinite(x).all():
        raise ValueError('input contains NaN or infinity')
    scale = np.abs(x).max((-2, -1), keepdims=True)
    scale = np.where(scale > 0, scale, np.ones_like(scale))
    stats = {'algorithm': algorithm, 'vendor_factorizations': 0, 'converged': True}
    u, s, v = raw(x / scale, algorithm, stats, opts)
    s = s * scale[:, 0, 0][:, None]
    if full_matrices:
        u = complete_basis(u, a.shape[-2])
        v = complete_basis(v, a.shape[-1])
    else:
        u, s, v = u[:, :, :rank], s[:, :rank], v[:, :, :rank]
    if not all(np.isfinite(z).all() for z in (u, s, v)):
        raise RuntimeError('standalone solver produced nonfinite output')
    stats.pop('_polar_factor', None)
    result = (None if values_only else u.reshape(*shape, *u.shape[-2:]),
              s.reshape(*shape, s.shape[-1]),
              None if values_only else mh(v).reshape(*shape, v.shape[-1], v.shape[-2]))
    return (*result, stats) if return_info else result
