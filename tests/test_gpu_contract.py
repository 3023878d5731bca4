"""The reference's contract test on the GPU (tests/test_standalone.py:146-156): every algorithm name
runs without any torch.linalg factorization (each one patched to raise) and never writes its
input; the names this build does not provide raise Unsupported instead of running elsewhere."""
from contextlib import ExitStack
from unittest.mock import patch

import numpy as np
import pytest
import torch

from conftest import load_package

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'
FORBIDDEN = ['svd', 'svdvals', 'qr', 'eigh', 'eig', 'eigvalsh', 'cholesky', 'solve', 'inv']


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32, torch.complex64])
def test_no_vendor_factorizations_and_immutable_input(dtype):
    pkg = load_package()
    rng = np.random.default_rng(9)
    a = torch.from_numpy(rng.normal(size=(9, 6))).to(DEV, dtype)
    if dtype.is_complex:
        a = a + 1j * torch.from_numpy(rng.normal(size=(9, 6))).to(DEV, dtype)
    saved = a.clone()
    ran = []
    with ExitStack() as stack:
        for name in FORBIDDEN:
            stack.enter_context(patch.object(torch.linalg, name, side_effect=AssertionError('vendor factorization')))
        for name in pkg.ALGORITHMS:
            try:
                u, s, vh = pkg.svd(a, name)
            except pkg.Unsupported:
                assert name not in pkg.SUPPORTED
                continue
            ran.append(name)
            k = s.shape[-1]
            rec = (u[:, :k] * s.to(u.dtype)) @ vh[:k]
            bar = 1e-10 if dtype in (torch.float64,) else 1e-4
            assert float((rec - a).abs().max()) <= bar * float(a.abs().max()), name
    assert torch.equal(a, saved)
    assert sorted(ran) == sorted(pkg.SUPPORTED)
