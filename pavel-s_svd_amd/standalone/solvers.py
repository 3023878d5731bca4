"""Drop-in solver entry points of the standalone SVD hot path on MI355X.

Same names, signatures, options, stats keys and exceptions as svdbench/standalone/solvers.py:
``ALGORITHMS`` (:7-11), ``svd`` (:264-311), ``factory`` (:313-324).  The hot-path families
``jacobi``, ``qr``, ``tsqr`` and ``polar_qdwh`` run on hand-written gfx950 kernels, and so do their
consumers ``block_jacobi`` (its own fused pair kernel), ``gram_jacobi`` and the range finders
``randomized``, ``subspace``, ``block_krylov`` (GEMM, QR and Jacobi kernels of the hot path); the
other names (bidiagonal and tridiagonal D&C families, the other polar iterations, Lanczos) stay in
``ALGORITHMS`` (so argument validation is unchanged) and raise ``Unsupported`` -- a RuntimeError
the harness records as an unsupported row -- instead of falling back to anything.

Extension options (no reference counterpart):
    cut='reference' | 'floor'   sigma below eps*max(m,n)*sigma_1 is zeroed and completed (the
                                reference contract, default) or only below eps*sigma_1, the floor
                                under which the Jacobi test stops enforcing orthogonality
    lanes, threads              Jacobi launch shape (see ``include/pavelsvd.h``)
    jacobi_method='auto'        matrices beyond LDS: 'blocked' (Gram subproblems of two column
                                blocks on MFMA, the default for n >= 64) or 'grid' (one rotation per
                                pair and round, the reference's order)
    jacobi_block_width=16       'blocked': columns per block, 16 (32 x 32 subproblems) or 32
    gemm='f32' | 'bf16x3'       arithmetic of the FP32 QR / TSQR products and of the QR-branch QDWH
                                iterations on MFMA (QDWH's Cholesky-branch steps stay FP32)
    polish=False                qr / tsqr, FP32 data: one FP64 Newton-Schulz step on U and on V.
                                QDWH, FP32 data: the iteration stops once |X^T X - I|_F <= orth_stop
                                (default 1e-2), then FP64 Newton-Schulz steps on X (polish_steps,
                                default 3) and on V, H = sym(X^T A) in FP64, G = V^T H V finished by
                                FP64 Jacobi sweeps (stat polish_sweeps), sigma = eig(G), U = X V W
                                in FP64 (finish threshold polish_tol, default 1e-10 max|G|);
                                eigh_tol loosens the FP32 eigensolve's threshold then;
                                eigh_inner_sweeps: sweeps per 64 x 64 subproblem (blocked solver)
    l0_model=None               QDWH: start from the learned lower bound (an L0Model, a JSON path,
                                or True for the packaged model); stats l0, l0_source
    block_size=4                block_jacobi: columns per group (the reference's option; <= 32 here)
    eigh_method='auto'          QDWH and gram_jacobi eigensolver: 'cyclic' (the reference's rotation order),
                                'blocked' (64 x 64 subproblems + MFMA) or 'auto' (LDS kernel when
                                H fits, blocked otherwise)
"""
import math

import torch

from . import primitives as P
from .blas import gemm

try:  # the harness's own exception type, so its `except Unsupported` matches
    from svdbench.algorithms import Unsupported
except ImportError:  # pragma: no cover - depends on the caller's environment
    class Unsupported(RuntimeError):
        """This build does not provide the requested algorithm, dtype or device."""

ALGORITHMS = (
    'bidiag_qr', 'bidiag_dc', 'jacobi', 'block_jacobi', 'qr', 'tsqr',
    'gram_jacobi', 'gram_dc', 'polar_qdwh', 'polar_newton_schulz',
    'polar_halley', 'polar_polynomial', 'randomized', 'subspace', 'block_krylov', 'lanczos',
)
MAX_LAUNCH_BATCH = 65535   # matrices per launch of the kernels indexed by the grid's y dimension
SUPPORTED = ('jacobi', 'qr', 'tsqr', 'polar_qdwh', 'block_jacobi', 'gram_jacobi', 'randomized', 'subspace',
             'block_krylov')
PARTIAL = ('randomized', 'subspace', 'block_krylov')
EPS = {torch.float32: 1.1920928955078125e-07, torch.float64: 2.220446049250313e-16}


def _jacobi(x, stats, opts, scale):
    """Tall batch (b, m, n) -> U, s, Vh with sigma rescaled (solvers.py:24-27 plus :294-299)."""
    vectors = not opts.get('_values_only', False)
    max_sweeps = opts.get('max_sweeps', 60)
    u, s, vh, info = P.jacobi_any(x, method=opts.get('jacobi_method', 'auto'), tol=opts.get('tol'),
                                  max_sweeps=max_sweeps, scale=scale, cut=opts.get('cut', 'reference'),
                                  vectors=vectors, lanes=opts.get('lanes'), threads=opts.get('threads'),
                                  block_width=opts.get('jacobi_block_width'))
    info = info.cpu()
    from .. import native as N
    N.raise_for_info(info, f'one-sided Jacobi did not converge in {max_sweeps} sweeps')
    stats['jacobi_sweeps'] = max(stats.get('jacobi_sweeps', 0), int(info[:, 0].max()))
    return u, s, vh


def _finish_columns(w, s):
    """U = W / sigma for sigma above eps * max(m, n) * sigma_1, completed below the cut, where sigma is
    zeroed (solvers.py:18-21).  W (b, m, n) is overwritten; s is sorted descending."""
    b, m, n = w.shape
    keep = s > EPS[w.dtype] * max(m, n) * s[:, :1]
    P.scale_columns_(w, s)
    if not bool(keep.all()):
        status = P.complete_columns_(w, keep.sum(-1).to(torch.int32))
        if bool((status != 0).any()):
            raise RuntimeError('null-space completion failed: supplied live columns are not independent')
    return w, torch.where(keep, s, torch.zeros_like(s))


def _block_jacobi(x, stats, opts):
    """Block-pair one-sided Jacobi (solvers.py:30-68): column groups of block_size (default 4) paired
    by the round-robin tournament, each pair orthogonalised by a full one-sided Jacobi solve and
    replaced by W_pair V_pair (sorted), a sweep ending with the Gram test max |g_ij| / sqrt(g_ii
    g_jj) <= tol (40 eps by default) over live columns.  One fused kernel per round on the GPU
    (block_jacobi.hip); one host read per sweep, as the reference's bool(corr.max() <= tol)."""
    b, m, n = x.shape
    width = opts.get('block_size', 4)
    if width < 1:
        raise ValueError('block_size must be positive')
    if -(-n // width) == 1:
        return _jacobi(x, stats, opts, scale=False)
    if x.is_complex():
        return _block_jacobi_complex(x, stats, opts, width)
    if 2 * width > 64:
        raise Unsupported(f'block_size {width} > 32: the pair Gram matrices of this build are held on chip')
    from .. import native as N
    w = x.clone()
    v = _eye(b, n, x.dtype, x.device)
    sweeper = P.BlockJacobiSweeper(w, v, width)
    tol = opts.get('tol') or 40 * EPS[x.dtype]
    inner_sweeps = opts.get('max_sweeps', 60)
    jac, converged, sweep = 0, False, 0
    for sweep in range(opts.get('max_sweeps', 30)):
        sweeper.sweep(opts.get('tol'), inner_sweeps)
        info = sweeper.info.cpu()
        N.raise_for_info(info, f'one-sided Jacobi did not converge in {inner_sweeps} sweeps')
        jac = max(jac, int(info[:, 0].max()))
        if float(sweeper.corr.max()) <= tol:
            converged = True
            break
    stats['jacobi_sweeps'] = max(stats.get('jacobi_sweeps', 0), jac)
    if not converged:
        raise RuntimeError('block Jacobi did not converge')
    s, w = P.sort_desc(sweeper.sigma, w)
    _, v = P.sort_desc(sweeper.sigma, v)
    u, s = _finish_columns(w, s)
    stats['block_sweeps'] = sweep + 1
    return u, s, v.mT


def _block_jacobi_complex(x, stats, opts, width):
    """Complex block_jacobi (solvers.py:30-68) for group pairs whose m x 2 block_size columns fit in
    LDS: each tournament round gathers its group pairs into one batch, solves it with the complex
    one-sided Jacobi kernel (the reference's _jacobi per pair, V's columns by descending sigma) and
    replaces W_pair, V_pair by W_pair V, V_pair V through the complex product; a sweep ends with the
    reference's Gram test max |g_ij| / sqrt(g_ii g_jj) over live columns (one host read per sweep)."""
    from .. import native as N
    b, m, n = x.shape
    real = P._real_dtype(x.dtype)
    groups = [list(range(j, min(j + width, n))) for j in range(0, n, width)]
    if len(groups) % 2:
        groups.append([])
    # the rounds' column sets, grouped by width (full pairs together, the short group apart)
    rounds = []
    for pairs in P._round_robin(len(groups)):
        by_width = {}
        for i, j in pairs:
            cols = groups[i] + groups[j]
            if len(cols) >= 2:
                by_width.setdefault(len(cols), []).append(cols)
        rounds.append([(k, torch.tensor(c, device=x.device)) for k, c in by_width.items()])
    w = x.clone()
    v = _eye(b, n, x.dtype, x.device)
    tol = opts.get('tol') or 40 * EPS[real]
    inner_sweeps = opts.get('max_sweeps', 60)
    jac, converged, sweep = 0, False, 0
    for sweep in range(opts.get('max_sweeps', 30)):
        infos = []
        for sets in rounds:
            for k, idx in sets:
                p = idx.shape[0]
                wp = w[:, :, idx].permute(0, 2, 1, 3).reshape(b * p, m, k).contiguous()
                vp = v[:, :, idx].permute(0, 2, 1, 3).reshape(b * p, n, k).contiguous()
                _, _, vh, info = P.jacobi_lds_uniform(wp, tol=opts.get('tol'), max_sweeps=inner_sweeps)
                infos.append(info)
                w[:, :, idx] = P.cmatmul(wp, vh, hb=True).reshape(b, p, m, k).permute(0, 2, 1, 3)
                v[:, :, idx] = P.cmatmul(vp, vh, hb=True).reshape(b, p, n, k).permute(0, 2, 1, 3)
        info = torch.cat(infos).cpu()
        N.raise_for_info(info, f'one-sided Jacobi did not converge in {inner_sweeps} sweeps')
        jac = max(jac, int(info[:, 0].max()))
        g = P.cmatmul(w, w, ha=True)
        d = g.diagonal(dim1=-2, dim2=-1).real.clamp_min(0)
        sq = d.sqrt()
        live = d > (EPS[real] * max(m, n)) ** 2 * d.amax(-1, keepdim=True)
        ga = g.abs()
        ga.diagonal(dim1=-2, dim2=-1).zero_()
        corr = torch.where(live[:, :, None] & live[:, None, :],
                           ga / (sq[:, :, None] * sq[:, None, :]).clamp_min(torch.finfo(real).tiny), 0)
        if float(corr.max()) <= tol:
            converged = True
            break
    stats['jacobi_sweeps'] = max(stats.get('jacobi_sweeps', 0), jac)
    if not converged:
        raise RuntimeError('block Jacobi did not converge')
    s = w.abs().square().sum(1).sqrt()
    order = s.argsort(-1, descending=True)
    s = s.gather(1, order)
    w = w.gather(2, order[:, None].expand_as(w)).contiguous()
    v = v.gather(2, order[:, None].expand_as(v))
    keep = s > EPS[real] * max(m, n) * s[:, :1]
    P.scale_columns_(torch.view_as_real(w).reshape(b, m, 2 * n), s.repeat_interleave(2, dim=1).contiguous())
    if not bool(keep.all()):
        status = P.complete_columns_(w, keep.sum(-1).to(torch.int32))
        if bool((status != 0).any()):
            raise RuntimeError('null-space completion failed: supplied live columns are not independent')
    stats['block_sweeps'] = sweep + 1
    return w, torch.where(keep, s, torch.zeros_like(s)), v.mH.contiguous()


def _gram(x, stats, opts):
    """Gram matrix and Hermitian Jacobi (solvers.py:99-119, the gram_jacobi branch): G = sym(A^T A)
    on MFMA, eigenvalues descending, sigma = sqrt(w) above eps * n * w_1 (zero below, counted in
    gram_discarded_values), U = A V / sigma on MFMA with the cut columns completed."""
    b, m, n = x.shape
    if x.is_complex():
        return _gram_complex(x, stats, opts)
    mode = _gmode(x, opts)
    g = gemm(x.mT, x, mode=mode)
    P.symmetrize_(g)
    w, v, it = P.jacobi_eigh(g, max_sweeps=opts.get('max_sweeps', 60), method=opts.get('eigh_method', 'auto'))
    stats['eigen_sweeps'] = it
    w = w.flip(-1)
    v = v.flip(-1).contiguous()
    keep = w > EPS[x.dtype] * n * w[:, :1].clamp_min(0)
    s = torch.where(keep, w.clamp_min(0).sqrt(), torch.zeros_like(w))
    stats['gram_discarded_values'] = int((~keep).sum())
    if opts.get('_values_only', False):
        return None, s, None
    u = gemm(x, v, mode=mode)
    P.scale_columns_(u, s)
    if not bool(keep.all()):
        status = P.complete_columns_(u, keep.sum(-1).to(torch.int32))
        if bool((status != 0).any()):
            raise RuntimeError('null-space completion failed: supplied live columns are not independent')
    return u, s, v.mT


def _gram_complex(x, stats, opts):
    """Complex gram_jacobi (solvers.py:99-119): G = sym(A^H A), its eigenpairs from the complex
    two-sided Jacobi (the reference's jacobi_eigh, flipped to descending), sigma = sqrt(w) above
    eps * n * w_1, U = A V / sigma with the cut columns completed."""
    b, m, n = x.shape
    real = torch.float32 if x.dtype == torch.complex64 else torch.float64
    g = P.cmatmul(x, x, ha=True)
    P.hermitize_(g)
    from .. import native as N
    max_sweeps = opts.get('max_sweeps', 60)
    w, v, info = P.jacobi_eigh_raw(g, max_sweeps=max_sweeps)   # solvers.py:107, eigenvalues flipped
    w = w.flip(-1)
    vh = v.flip(-1).mH.contiguous()
    info = info.cpu()
    N.raise_for_info(info, f"Hermitian Jacobi did not converge in {max_sweeps} sweeps")
    stats['eigen_sweeps'] = int(info[:, 0].max())
    keep = w > EPS[real] * n * w[:, :1]
    s = torch.where(keep, w.clamp_min(0).sqrt(), torch.zeros_like(w))
    stats['gram_discarded_values'] = int((~keep).sum())
    if opts.get('_values_only', False):
        return None, s, None
    u = P.cmatmul(x, vh, hb=True)
    ur = torch.view_as_real(u)
    P.scale_columns_(ur.reshape(b, m, 2 * n), s.repeat_interleave(2, dim=1).contiguous())
    if not bool(keep.all()):
        status = P.complete_columns_(u, keep.sum(-1).to(torch.int32))
        if bool((status != 0).any()):
            raise RuntimeError('null-space completion failed: supplied live columns are not independent')
    return u, s, vh


def _sketch(x, columns, seed):
    """Gaussian test matrix (b, n, columns) from torch's generator on the input's device, seeded as
    the reference's (solvers.py:180-182)."""
    gen = torch.Generator(device=x.device).manual_seed(seed)
    return torch.randn(x.shape[0], x.shape[-1], columns, dtype=x.dtype, device=x.device, generator=gen)


def _partial(x, stats, opts, method):
    """Randomized range finder (solvers.py:185-221): Q = orth(A Omega), `iterations` power steps
    Q <- orth(A orth(A^T Q)) ('randomized', 'subspace'), or a block Krylov basis of the same steps
    re-orthogonalised twice against the earlier blocks ('block_krylov'); then the Jacobi SVD of
    Q^T A and U = Q U_small, truncated to `rank`.  Products on MFMA, QRs by the panel kernels."""
    b, m, n = x.shape
    k = opts.get('rank', min(m, n))
    ell = min(k + opts.get('oversample', 8), min(m, n))
    omega = opts.get('initial_basis')
    if omega is None:
        omega = _sketch(x, ell, opts.get('seed', 20261011))
    else:
        if omega.ndim == 2:
            omega = omega[None]
        if tuple(omega.shape) != (b, n, ell):
            raise ValueError('initial_basis shape must match batch, n, rank+oversample')
        omega = omega.to(device=x.device, dtype=x.dtype)
        stats['warm_started'] = True
    mode = _gmode(x, opts)
    qmode = 'bf16x3' if mode == 'bf16x3' else 'f32'
    cplx = x.is_complex()

    def orth(y):
        return P.qr(y.contiguous(), gemm=qmode)[0]

    def mm(a_, b_, ha=False):
        """op(a_) @ b_ with op(a_) = a_^H (a_^T for real data) when ha."""
        if cplx:
            return P.cmatmul(a_, b_, ha=ha)
        return gemm(a_.mT if ha else a_, b_, mode=mode)

    q = orth(mm(x, omega.contiguous()))
    bases = [q]
    iterations = opts.get('iterations', 2)
    for _ in range(iterations):
        z = orth(mm(x, q, ha=True))
        y = mm(x, z)
        if method == 'block_krylov':
            joined = torch.cat(bases, -1)
            available = m - joined.shape[-1]
            if available <= 0:
                break
            for _ in range(2):
                if cplx:
                    y = _axpby_complex(-1.0, P.cmatmul(joined, P.cmatmul(joined, y, ha=True)), 1.0, y.contiguous())
                else:
                    y = gemm(joined, gemm(joined.mT, y, mode=mode), alpha=-1.0, beta=1.0, out=y, mode=mode)
            q = orth(y[:, :, :min(ell, available)])
            bases.append(q)
        else:
            q = orth(y)
    if method == 'block_krylov':
        q = orth(torch.cat(bases, -1))
    small = mm(q, x, ha=True)
    u, s, vh = _raw(small, 'jacobi', stats, opts)
    stats.update(sketch_columns=ell, range_columns=q.shape[-1], power_iterations=iterations)
    if u is None:
        return None, s[:, :k], None
    return mm(q, u)[:, :, :k], s[:, :k], vh[:, :k, :]


def _gmode(x, opts):
    """GEMM arithmetic for FP32 data ('f32' or 'bf16x3' from opts['gemm']); FP64 data uses FP64."""
    return opts.get('gemm', 'f32') if x.dtype == torch.float32 else 'f64'


def tsqr_leaf_bounds(m, n, rows):
    """Row ranges of the TSQR leaves, or None for a single QR: leaves are `max(rows, n)` rows
    tall and a short last block joins the previous leaf (solvers.py:73-83)."""
    rows = max(rows, n)
    if m <= rows:
        return None
    bounds = []
    for start in range(0, m, rows):
        end = min(start + rows, m)
        if end - start < n:
            start = bounds.pop()[0]
        bounds.append((start, end))
    return bounds


def _tsqr(x, rows, stats, opts):
    """TSQR: leaf QRs, then a binary tree of [R1; R2] QRs with the explicit Q back-transformation
    Q <- [Q1 q_top; Q2 q_bot] at every level (solvers.py:71-96).  Equal-height leaves of a matrix
    are one batched QR launch; each tree level is one batched QR plus two GEMMs per merge."""
    b, m, n = x.shape
    mode = _gmode(x, opts)
    qmode = mode if mode != 'f64' else 'f32'
    bounds = tsqr_leaf_bounds(m, n, rows)
    if bounds is None:
        return P.qr(x, gemm=qmode)
    stats['tsqr_leaves'] = len(bounds)
    qs, rs = [], []
    for j in range(b):
        xj = x[j]
        h = bounds[0][1] - bounds[0][0]
        k = sum(1 for s0, e0 in bounds if e0 - s0 == h)
        nodes = []
        ql, rl = P.qr(xj[:k * h].reshape(k, h, n), gemm=qmode)
        nodes += [(ql[i], rl[i]) for i in range(k)]
        for s0, e0 in bounds[k:]:
            q1, r1 = P.qr(xj[s0:e0].unsqueeze(0), gemm=qmode)
            nodes.append((q1[0], r1[0]))
        while len(nodes) > 1:
            pairs = [(nodes[i], nodes[i + 1]) for i in range(0, len(nodes) - 1, 2)]
            stacked = torch.stack([torch.cat([r1, r2], 0) for (_, r1), (_, r2) in pairs])
            qm, rm = P.qr(stacked, gemm=qmode)
            nxt = []
            for i, ((q1, _), (q2, _)) in enumerate(pairs):
                if x.is_complex():
                    out = torch.cat([P.cmatmul(q1[None], qm[i:i + 1, :n])[0],
                                     P.cmatmul(q2[None], qm[i:i + 1, n:])[0]], 0)
                    nxt.append((out, rm[i]))
                    continue
                out = torch.empty(q1.shape[0] + q2.shape[0], n, dtype=x.dtype, device=x.device)
                gemm(q1, qm[i, :n], out=out[:q1.shape[0]], mode=mode)
                gemm(q2, qm[i, n:], out=out[q1.shape[0]:], mode=mode)
                nxt.append((out, rm[i]))
            if len(nodes) % 2:
                nxt.append(nodes[-1])
            nodes = nxt
        qs.append(nodes[0][0])
        rs.append(nodes[0][1])
    return torch.stack(qs), torch.stack(rs)


def qdwh_weights(lower):
    """Dynamically weighted Halley coefficients (a, b, c) for the current lower bound l on
    sigma_min / sigma_max and the bound after the step (solvers.py:143-149; Nakatsukasa, Bai and
    Gygi, SIAM J. Matrix Anal. Appl. 31(5), 2010): d = (4 (1 - l^2) / l^4)^(1/3),
    a = sqrt(1 + d) + sqrt(8 - 4 d + 8 (2 - l^2) / (l^2 sqrt(1 + d))) / 2, b = (a - 1)^2 / 4,
    c = a + b - 1, l' = min(1, l (a + b l^2) / (1 + c l^2))."""
    l2 = lower * lower
    d = (4.0 * (1.0 - l2) / (l2 * l2)) ** (1.0 / 3.0)
    sd = math.sqrt(1.0 + d)
    a = sd + 0.5 * math.sqrt(max(8.0 - 4.0 * d + 8.0 * (2.0 - l2) / (l2 * sd), 0.0))
    b = (a - 1.0) ** 2 / 4.0
    c = a + b - 1.0
    return a, b, c, min(1.0, lower * (a + b * l2) / (1.0 + c * l2))


def _newton_schulz(x):
    """One Newton-Schulz step toward the orthogonal polar factor, x <- x (3 I - x^T x) / 2, as two
    FP64 MFMA products (x: (b, m, n) float64, m >= n)."""
    x = x.contiguous()
    g = gemm(x.mT, x, mode='f64')
    return gemm(x, g, alpha=-0.5, beta=1.5, out=x.clone(), mode='f64')


def _eye(b, n, dtype, device):
    return torch.eye(n, dtype=dtype, device=device).expand(b, n, n).contiguous()


def _qdwh_iterate(x, lower, tol, max_iter, mode, qmode, orth_stop=None):
    """QDWH iterations from x = A / |A|_F and the lower bound `lower` (solvers.py:131-167);
    returns (X, iterations, QR-branch iterations).  x is not modified.

    orth_stop (for the FP64-polished form only): a Cholesky-branch step first forms X^T X; when
    |X^T X - I|_F <= orth_stop every singular value of X is within the quadratic basin of the FP64
    Newton-Schulz steps that follow, and the iteration ends there instead of waiting for the
    elementwise test, which FP32 rounding delays by several steps at large n."""
    b, m, n = x.shape
    dt, dev = x.dtype, x.device
    nqr = 0
    for it in range(max_iter):
        aa, bb, cc, lower = qdwh_weights(lower)
        if cc > 100:
            nqr += 1
            stacked = torch.empty(b, m + n, n, dtype=dt, device=dev)
            P.scale(x, None, s=math.sqrt(cc), out=stacked[:, :m])
            stacked[:, m:] = torch.eye(n, dtype=dt, device=dev)
            q, _ = P.qr(stacked, gemm=qmode)
            xn = x.clone()
            gemm(q[:, :m], q[:, m:].mT, alpha=(aa - bb / cc) / math.sqrt(cc), beta=bb / cc, out=xn, mode=mode)
        else:
            # the Cholesky-branch steps end the iteration: their products stay in full FP32 (FP64
            # for FP64 data) so that the fixed-point test at tol = 10 eps can be met -- BF16x3
            # carries about 2^-16 relative precision
            cmode = 'f32' if mode == 'bf16x3' else mode
            if orth_stop is not None:
                g = gemm(x.mT, x, mode=cmode)
                eye = _eye(b, n, dt, dev)
                dev_ = P.axpby_(1.0, g, -1.0, eye.clone())                 # X^T X - I
                if float(P.reduce(dev_, 'fro').max()) <= orth_stop:
                    return x, it, nqr
                z = P.axpby_(cc, g, 1.0, eye)
            else:
                z = _eye(b, n, dt, dev)
                gemm(x.mT, x, alpha=cc, beta=1.0, out=z, mode=cmode)
            P.symmetrize_(z)
            chol, fail = P.cholesky(z, gemm=cmode if cmode != 'f64' else 'f32')
            if bool(fail.any()):
                raise RuntimeError('polar Cholesky lost positive definiteness')
            w = x.clone()
            P.trsm_right_(w, chol, trans=True, gemm=cmode if cmode != 'f64' else 'f32')
            P.trsm_right_(w, chol, trans=False, gemm=cmode if cmode != 'f64' else 'f32')
            xn = P.axpby_(aa - bb / cc, w, bb / cc, x.clone())
        test = P.reduce(xn, 'maxdiff', x).amax(0).cpu()   # [max|x' - x|, max|x'|] over the batch
        x = xn
        if float(test[0]) <= tol * max(float(test[1]), 1.0):
            return x, it + 1, nqr
    raise RuntimeError(f'polar_qdwh did not converge in {max_iter} iterations')


def _fixed_point_error(x, mode):
    """max |X (X^T X) - X| over the batch (solvers.py:175-176)."""
    fp = gemm(x, gemm(x.mT, x, mode=mode), mode=mode)
    return float(P.reduce(fp, 'maxdiff', x)[:, 0].max())


def _polar(a, stats, opts):
    """QDWH polar decomposition, then the Jacobi eigensolve of H = sym(X^T A)
    (solvers.py:122-177) on the GPU kernels: the QR branch stacks [sqrt(c) X; I] for the blocked
    Householder QR and folds X' = (b/c) X + (a - b/c)/sqrt(c) Q_top Q_bot^T into one GEMM; the
    Cholesky branch forms Z = I + c X^T X with one GEMM, factors it and applies X Z^-1 as two
    right-sided triangular solves.  One host synchronisation per iteration reads the convergence
    test max|X' - X| <= tol * max(max|X'|, 1) (over the whole batch, as the reference).

    With ``l0_model`` (an L0Model, a path, or True for the packaged model) and no explicit
    lower_bound, the iteration starts from the learned bound; a run that raises or ends with a
    fixed-point error above 100 tol is discarded and repeated from the reference's l0 = eps."""
    b, m, n = a.shape
    dt, dev = a.dtype, a.device
    eps = EPS[dt]
    mode = _gmode(a, opts)
    qmode = mode if mode != 'f64' else 'f32'
    x0 = P.scale(a, P.reduce(a, 'fro'), invert=True)
    tol = opts.get('tol') or 10 * eps
    lower = opts.get('lower_bound', eps)
    if not 0 < lower <= 1:
        raise ValueError('lower_bound must be in (0,1]')
    max_iter = opts.get('max_iter', 40)
    polish = bool(opts.get('polish', False)) and dt == torch.float32
    orth_stop = opts.get('orth_stop', 1e-2) if polish else None
    x = None
    l0_model = opts.get('l0_model')
    if l0_model is not None and l0_model is not False and 'lower_bound' not in opts:
        from ..learned_l0 import L0Model, DEFAULT_MODEL
        model = l0_model if isinstance(l0_model, L0Model) else \
            L0Model.load(DEFAULT_MODEL if l0_model is True else l0_model, dev)
        learned = model.lower_bound(a, eps) if min(m, n) >= 16 else None
        stats['l0_source'] = 'fallback'
        if learned is not None:
            try:
                x, it, nqr = _qdwh_iterate(x0, learned, tol, max_iter, mode, qmode, orth_stop)
                if _fixed_point_error(x, mode) > (2 * orth_stop if orth_stop is not None else 100 * tol):
                    x = None
            except RuntimeError:
                x = None
            stats['l0_learned'] = learned
            if x is not None:
                stats['l0_source'] = 'learned'
                lower = learned
    if x is None:
        x, it, nqr = _qdwh_iterate(x0, lower, tol, max_iter, mode, qmode, orth_stop)
    stats['polar_qr_iterations'] = nqr
    if l0_model is not None and l0_model is not False:
        stats['l0'] = lower
    if polish:
        # FP64 tail: Newton-Schulz steps make X orthogonal to FP64 rounding (three from the
        # |X^T X - I|_F <= 1e-2 stop: 1e-2 -> 1e-4 -> 1e-8 -> 1e-16), H = sym(X^T A) is formed in
        # FP64 and rounded once for the FP32 eigensolve
        x64 = x.double()
        for _ in range(int(opts.get('polish_steps', 3))):
            x64 = _newton_schulz(x64)
        h64 = gemm(x64.mT, a.double(), mode='f64')
        P.symmetrize_(h64)
        h = h64.float()
    else:
        h = gemm(x.mT, a, mode=mode)
        P.symmetrize_(h)
    # with the FP64 finish the FP32 eigensolve only has to bring V^T H V close to diagonal
    w, v, info = P.jacobi_eigh_raw(h, tol=opts.get('eigh_tol') if polish else None,
                                   max_sweeps=opts.get('max_sweeps', 60), method=opts.get('eigh_method', 'auto'),
                                   inner_sweeps=opts.get('eigh_inner_sweeps', 1))
    info_h = info.cpu()
    from .. import native as N
    N.raise_for_info(info_h, f"Hermitian Jacobi did not converge in {opts.get('max_sweeps', 60)} sweeps")
    polish_sweeps = 0
    if polish:
        # V orthogonal to FP64 rounding, then G = V^T H V (nearly diagonal: its off-diagonal is the
        # FP32 eigensolve's backward error) finished by FP64 Jacobi sweeps; sigma = eig(G)
        # descending, V <- V W
        v64 = _newton_schulz(v.double())
        g64 = gemm(v64.mT, gemm(h64, v64, mode='f64'), mode='f64')
        P.symmetrize_(g64)
        # the finish stops at |g_pq| <= polish_tol * max|G| (default 1e-10): the eigenvalues then move
        # by at most the remaining off-diagonal norm, far below FP32 resolution
        w2, z, info2 = P.jacobi_eigh_raw(g64, tol=opts.get('polish_tol', 1e-10), max_sweeps=opts.get('max_sweeps', 60),
                                         inner_sweeps=opts.get('eigh_inner_sweeps', 1))
        info2 = info2.cpu()
        N.raise_for_info(info2, f"Hermitian Jacobi did not converge in {opts.get('max_sweeps', 60)} sweeps")
        polish_sweeps = int(info2[:, 0].max())
        s64, v64 = P.sort_desc(w2, gemm(v64, z, mode='f64'))
        s = s64.to(dt)
        u = gemm(x64, v64, mode='f64').to(dt)
        v = v64.to(dt)
    else:
        s = w.flip(-1).clamp_min(0)
        v = v.flip(-1).contiguous()
        u = gemm(x, v, mode=mode)
    cut_factor = eps if opts.get('cut', 'reference') == 'floor' else eps * max(m, n)
    keep = s > cut_factor * s[:, :1]
    nk = keep.sum(-1).to(torch.int32)
    if not bool(keep.all()):
        status = P.complete_columns_(u, nk)
        if bool((status != 0).any()):
            raise RuntimeError('null-space completion failed: supplied live columns are not independent')
    s = torch.where(keep, s, torch.zeros_like(s))
    if polish:
        stats['polish_sweeps'] = polish_sweeps
    stats.update(polar_iterations=it, eigen_sweeps=int(info_h[:, 0].max()),
                 polar_fixed_point_error=_fixed_point_error(x64, 'f64') if polish else _fixed_point_error(x, mode))
    return u, s, v


def _polar_complex(a, stats, opts):
    """QDWH polar SVD of a complex tall batch (solvers.py:122-178 with the complex arithmetic of
    primitives.py): X = A / ||A||_F and dynamically weighted Halley steps with the
    reference's branch rule -- while c > 100, X <- (b/c) X + (a - b/c) / sqrt(c) Q1 Q2^H from the QR
    of [sqrt(c) X; I]; afterwards Z = I + c X^H X, L = chol(sym Z) (RuntimeError when it loses positive
    definiteness) and X <- (b/c) X + (a - b/c) X Z^-1 by the two triangular solves -- stopping on
    max|X_new - X| <= tol max(max|X|, 1) over the batch; then H = sym(X^H A) and its eigenpairs from
    the complex two-sided Jacobi (primitives.py:370-427), sigma the flipped eigenvalues clamped at 0,
    U = X V with the cut columns completed.  The QR, Cholesky, solve and eigensolver kernels hold each
    matrix in LDS or, beyond it, in a workspace region."""
    b, m, n = a.shape
    real = torch.float32 if a.dtype == torch.complex64 else torch.float64
    x = P.scale(a, P.reduce(a, 'fro'), invert=True)
    tol = opts.get('tol') or 10 * EPS[real]
    lower = opts.get('lower_bound', EPS[real])
    if not 0 < lower <= 1:
        raise ValueError('lower_bound must be in (0,1]')
    max_iter = opts.get('max_iter', 40)
    ident = torch.zeros(b, n, n, dtype=a.dtype, device=a.device)
    ident.diagonal(dim1=-2, dim2=-1).fill_(1)
    converged, nqr = False, 0
    for it in range(max_iter):
        aa, bb, cc, lower = qdwh_weights(lower)
        if cc > 100:
            nqr += 1
            q, _ = P.cqr(torch.cat([P.scale(x, None, s=math.sqrt(cc)), ident], -2))
            xn = P.cmatmul(q[:, :m], q[:, m:], hb=True)
            _axpby_complex(bb / cc, x, (aa - bb / cc) / math.sqrt(cc), xn)
        else:
            z = _axpby_complex(cc, P.cmatmul(x, x, ha=True), 1.0, ident.clone())    # I + c X^H X
            P.hermitize_(z)
            chol, fail = P.cholesky(z)
            if bool(fail.any()):
                raise RuntimeError('polar Cholesky lost positive definiteness')
            xn = P.csolve_chol_(x.clone(), chol)                                     # X Z^-1
            _axpby_complex(bb / cc, x, aa - bb / cc, xn)
        d = P.reduce(xn, 'maxdiff', x).max(0).values.cpu()
        x = xn
        if float(d[0]) <= tol * max(float(d[1]), 1.0):
            converged = True
            break
    if not converged:
        raise RuntimeError(f'polar_qdwh did not converge in {max_iter} iterations')
    h = P.cmatmul(x, a, ha=True)
    P.hermitize_(h)
    from .. import native as N
    max_sweeps = opts.get('max_sweeps', 60)
    w, v, info = P.jacobi_eigh_raw(h, max_sweeps=max_sweeps)
    info = info.cpu()
    N.raise_for_info(info, f'Hermitian Jacobi did not converge in {max_sweeps} sweeps')
    w = w.flip(-1).clamp_min(0)
    vh = v.flip(-1).mH.contiguous()
    keep = w > EPS[real] * max(m, n) * w[:, :1]
    u = P.complete_null_columns(P.cmatmul(x, vh, hb=True), keep)
    w = torch.where(keep, w, torch.zeros_like(w))
    fixed = P.reduce(P.cmatmul(x, P.cmatmul(x, x, ha=True)), 'maxdiff', x)[:, 0].max()
    stats.update(polar_iterations=it + 1, polar_qr_iterations=nqr, eigen_sweeps=int(info[:, 0].max()),
                 polar_fixed_point_error=float(fixed))
    return u, w, vh


def _axpby_complex(alpha, x, beta, y):
    """y <- alpha x + beta y for complex x, y with real alpha, beta (through the real views)."""
    xr = torch.view_as_real(x.contiguous())
    yr = torch.view_as_real(y)
    P.axpby_(alpha, xr.reshape(*xr.shape[:-2], -1), beta, yr.reshape(*yr.shape[:-2], -1))
    return y


def _raw(x, method, stats, opts):
    """Wide inputs are solved transposed (solvers.py:226-228) except by the partial methods; dispatch
    by name (solvers.py:229-261)."""
    if x.shape[-2] < x.shape[-1] and method not in PARTIAL:
        u, s, vh = _raw(x.mT.contiguous(), method, stats, opts)
        return (None if vh is None else vh.mT), s, (None if u is None else u.mT)
    if method == 'jacobi':
        return _jacobi(x, stats, opts, scale=not opts.get('_prescaled', False))
    if method == 'block_jacobi':
        return _block_jacobi(x, stats, opts)
    if method == 'gram_jacobi':
        return _gram(x, stats, opts)
    if method in PARTIAL:
        return _partial(x, stats, opts, method)
    if method in ('qr', 'tsqr'):
        n = x.shape[-1]
        values = opts.get('_values_only', False)
        mode = _gmode(x, opts)
        if method == 'qr':
            q, r = P.qr(x, gemm=mode if mode != 'f64' else 'f32', want_q=not values)
        else:
            q, r = _tsqr(x, opts.get('tile_rows', 4 * n), stats, opts)
        u, s, vh = _jacobi(r, stats, opts, scale=False)
        if values:
            return None, s, vh
        if x.is_complex():
            return P.cmatmul(q, u), s, vh
        u = gemm(q, u, mode=mode)
        if opts.get('polish', False) and x.dtype == torch.float32:
            # Q's orthogonality is that of its products (about 2^-16 with BF16x3) and V's that of
            # the FP32 sweeps: one FP64 Newton-Schulz step on each restores it
            u = _newton_schulz(u.double()).float()
            vh = _newton_schulz(vh.mT.double()).mT.float().contiguous()
        return u, s, vh
    if method == 'polar_qdwh':
        if x.is_complex():
            return _polar_complex(x, stats, opts)
        u, s, v = _polar(x, stats, opts)
        return u, s, v.mT
    raise Unsupported(f'{method!r} is not built for MI355X yet')


def svd(a, algorithm='jacobi', *, full_matrices=False, values_only=False, return_info=False, **opts):
    """Return U, s, Vh (and optionally diagnostics) for batched real tensors on a ROCm device.

    Argument checks and their order follow solvers.py:272-292; the input is never overwritten.
    """
    if algorithm not in ALGORITHMS:
        raise ValueError(f'algorithm must be one of {ALGORITHMS}')
    if a.ndim < 2 or min(a.shape[-2:]) < 1:
        raise ValueError('nonempty matrix dimensions required')
    if a.dtype not in (torch.float32, torch.float64, torch.complex64, torch.complex128):
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
    if opts.get('cut', 'reference') not in ('reference', 'floor'):
        raise ValueError("cut must be 'reference' or 'floor'")
    shape = a.shape[:-2]
    x = a.reshape(-1, *a.shape[-2:])
    if not bool(torch.isfinite(x).all()):
        raise ValueError('input contains NaN or infinity')
    if algorithm not in SUPPORTED:
        raise Unsupported(f'{algorithm!r} is outside the MI355X hot path (built: {SUPPORTED})')
    if a.device.type != 'cuda':
        raise Unsupported('the MI355X solvers take ROCm device tensors; there is no CPU path')
    if algorithm != 'jacobi' and x.shape[0] > MAX_LAUNCH_BATCH:
        # the batched kernels of the other solvers index matrices by the grid's y dimension
        # (at most 65 535): solve in slices of that many matrices and join the results
        outs = [svd(x[i:i + MAX_LAUNCH_BATCH], algorithm, full_matrices=full_matrices, values_only=values_only,
                    return_info=True, **opts) for i in range(0, x.shape[0], MAX_LAUNCH_BATCH)]
        s = torch.cat([o[1] for o in outs])
        result = (None if values_only else torch.cat([o[0] for o in outs]).reshape(*shape, *outs[0][0].shape[-2:]),
                  s.reshape(*shape, s.shape[-1]),
                  None if values_only else torch.cat([o[2] for o in outs]).reshape(*shape, *outs[0][2].shape[-2:]))
        stats = dict(outs[0][3])
        for o in outs[1:]:
            for k, v in o[3].items():
                if isinstance(v, (int, float)) and not isinstance(v, bool) and isinstance(stats.get(k), (int, float)):
                    stats[k] = max(stats[k], v)
        return (*result, stats) if return_info else result
    stats = {'algorithm': algorithm, 'vendor_factorizations': 0, 'converged': True}
    m, n = x.shape[-2:]
    with torch.no_grad():
        x = x.contiguous()
        inner = dict(opts, _values_only=values_only)
        scale = None
        if algorithm != 'jacobi':
            # normalise before any sums of squares (solvers.py:293-295); the Jacobi kernel fuses it
            scale = P.reduce(x, 'maxabs')
            x = P.scale(x, scale, invert=True)
            inner['_prescaled'] = True
        u, s, vh = _raw(x, algorithm, stats, inner)
        if scale is not None:
            s = P.scale(s.unsqueeze(1), scale).squeeze(1)
        if full_matrices and not values_only:
            u = P.complete_basis(u.contiguous(), a.shape[-2])
            vh = P.complete_basis(vh.mT.contiguous(), a.shape[-1]).mT
        elif not values_only:
            u, vh = u[:, :, :rank], vh[:, :rank, :]
        s = s[:, :rank]
        outs = (s,) if values_only else (u, s, vh)
        finite = torch.isfinite(outs[0]).all()
        for z in outs[1:]:
            finite = finite & torch.isfinite(z).all()
        if not bool(finite):   # one host read for all outputs
            raise RuntimeError('standalone solver produced nonfinite output')
        result = (None if values_only else u.reshape(*shape, *u.shape[-2:]),
                  s.reshape(*shape, s.shape[-1]),
                  None if values_only else vh.reshape(*shape, *vh.shape[-2:]))
    return (*result, stats) if return_info else result


def factory(a, spec):
    """Adapter for the harness plugin hook ("module:factory"); algorithm is spec['method']
    (solvers.py:313-324)."""
    opts = {k: v for k, v in spec.items() if k not in ('name', 'method', 'mode', 'seed_offset')}

    from .. import native as N
    method = spec['method']

    def run():
        N.workspace_reset()
        u, s, vh, info = svd(a, method, full_matrices=spec.get('mode') == 'full',
                             values_only=spec.get('mode') == 'values', return_info=True, **opts)
        run.diagnostics = info
        # the library allocates no device memory: all of its workspace comes from torch's allocator
        # (inside the harness's allocator peaks); the last solve's requests are reported beside it
        run.native_workspace_bytes = 0
        run.library_workspace_bytes = N.workspace_stats()
        return u, s, vh
    control = {'jacobi': 'device-side sweeps and convergence; one host read of the status per call',
               'block_jacobi': 'one fused kernel per tournament round; one host read of the Gram test per sweep',
               'polar_qdwh': 'one host read of the convergence pair per QDWH iteration, one per sweep of the '
                             'eigensolver beyond LDS'}
    run.metadata = {'implementation': 'hand-written gfx950 HIP kernels', 'vendor_factorizations': False,
                    'values_only_forms_vectors': method in ('block_jacobi', 'polar_qdwh'),
                    'rng_in_solve': method in PARTIAL and 'initial_basis' not in spec,
                    'workspace': 'caller-allocated through the PyTorch allocator; no native allocations',
                    'gpu_control': control.get(method, 'device-side sweeps and convergence; host reads of '
                                                       'per-sweep or per-call status only')}
    return run
