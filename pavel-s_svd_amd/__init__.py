"""MI355X-native (gfx950) standalone SVD hot path: one-sided Jacobi, QDWH polar and TSQR, and
the solvers built on them (block-pair Jacobi, Gram + Jacobi, randomized range finders).

Drop-in for ``svdbench.standalone``: ``svd(a, algorithm, ...)``, ``ALGORITHMS`` and
``factory(a, spec)`` keep the reference's names, options, stats and exceptions, and run on
hand-written HIP kernels behind the C ABI of ``include/pavelsvd.h``.  The directory name is not a
Python identifier, so load it with ``importlib.import_module('pavel-s_svd_amd')`` or through the
harness hook ``{"name": "pavel-s_svd_amd:factory", "method": "jacobi"}``.
"""
from .standalone import ALGORITHMS, SUPPORTED, Unsupported, factory, svd

__all__ = ['ALGORITHMS', 'SUPPORTED', 'Unsupported', 'svd', 'factory']
