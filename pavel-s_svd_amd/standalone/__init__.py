"""MI355X standalone SVD solvers: the drop-in for svdbench.standalone (svd, ALGORITHMS, factory)."""
from .solvers import ALGORITHMS, SUPPORTED, Unsupported, factory, svd

__all__ = ['ALGORITHMS', 'SUPPORTED', 'Unsupported', 'svd', 'factory']
