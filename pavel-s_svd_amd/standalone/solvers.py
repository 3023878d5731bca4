"""Drop-in solver entry points of the standalone SVD hot path on MI355X.

Same names, signatures, options, stats keys and exceptions as svdbench/standalone/solvers.py:
``ALGORITHMS`` (:7-11), ``svd`` (:264-311), ``factory`` (:313-324).  The hot-path families
``jacobi``, ``tsqr`` and ``polar_qdwh`` run on hand-written gfx950 kernels; the other names stay
in ``ALGORITHMS`` (so argument validation is unchanged) and raise ``Unsupported`` -- a
RuntimeError the harness records as an unsupported row -- instead of falling back to anything.

Extension options (no reference counterpart):
    cut='reference' | 'floor'   sigma below eps*max(m,n)*sigma_1 is zeroed and completed (the
                                reference contract, default) or only below eps*sigma_1, the floor
                                under which the Jacobi test stops enforcing orthogonality
    lanes, threads              Jacobi launch shape (see ``include/pavelsvd.h``)
"""
import torch

from . import primitives as P

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
SUPPORTED = ('jacobi',)


def _jacobi(x, stats, opts, scale):
    """Tall batch (b, m, n) -> U, s, Vh with sigma rescaled (solvers.py:24-27 plus :294-299)."""
    vectors = not opts.get('_values_only', False)
    max_sweeps = opts.get('max_sweeps', 60)
    u, s, vh, info = P.jacobi_any(x, tol=opts.get('tol'), max_sweeps=max_sweeps, scale=scale,
                                  cut=opts.get('cut', 'reference'), vectors=vectors,
                                  lanes=opts.get('lanes'), threads=opts.get('threads'))
    info = info.cpu()
    from .. import native as N
    N.raise_for_info(info, f'one-sided Jacobi did not converge in {max_sweeps} sweeps')
    stats['jacobi_sweeps'] = max(stats.get('jacobi_sweeps', 0), int(info[:, 0].max()))
    return u, s, vh


def _raw(x, method, stats, opts):
    """Wide inputs are solved transposed (solvers.py:226-228); dispatch by name."""
    if x.shape[-2] < x.shape[-1]:
        u, s, vh = _raw(x.mT.contiguous(), method, stats, opts)
        return (None if vh is None else vh.mT), s, (None if u is None else u.mT)
    if method == 'jacobi':
        return _jacobi(x, stats, opts, scale=True)
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
    if a.is_complex():
        raise Unsupported('complex dtypes are not built for MI355X yet')
    if a.device.type != 'cuda':
        raise Unsupported('the MI355X solvers take ROCm device tensors; there is no CPU path')
    if full_matrices:
        raise Unsupported('full_matrices completion is not built for MI355X yet')
    stats = {'algorithm': algorithm, 'vendor_factorizations': 0, 'converged': True}
    m, n = x.shape[-2:]
    with torch.no_grad():
        u, s, vh = _raw(x.contiguous(), algorithm, stats, dict(opts, _values_only=values_only))
        if not values_only:
            u, vh = u[:, :, :rank], vh[:, :rank, :]
        s = s[:, :rank]
        outs = (s,) if values_only else (u, s, vh)
        if not all(bool(torch.isfinite(z).all()) for z in outs):
            raise RuntimeError('standalone solver produced nonfinite output')
        result = (None if values_only else u.reshape(*shape, *u.shape[-2:]),
                  s.reshape(*shape, s.shape[-1]),
                  None if values_only else vh.reshape(*shape, *vh.shape[-2:]))
    return (*result, stats) if return_info else result


def factory(a, spec):
    """Adapter for the harness plugin hook ("module:factory"); algorithm is spec['method']
    (solvers.py:313-324)."""
    opts = {k: v for k, v in spec.items() if k not in ('name', 'method', 'mode', 'seed_offset')}

    def run():
        u, s, vh, info = svd(a, spec['method'], full_matrices=spec.get('mode') == 'full',
                             values_only=spec.get('mode') == 'values', return_info=True, **opts)
        run.diagnostics = info
        return u, s, vh
    run.metadata = {'implementation': 'hand-written gfx950 HIP kernels', 'vendor_factorizations': False,
                    'values_only_forms_vectors': False, 'rng_in_solve': False,
                    'gpu_control': 'device-side sweeps and convergence; one host sync per call for status'}
    return run
