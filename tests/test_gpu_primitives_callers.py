"""The primitives module as the reference's other callers use it by name (D&C leaves and merge
refinement, divide_conquer.py:79-82, 114-118, 127, 146; laboratory, algorithm_laboratory/
algorithms.py:13-24): tiny FP64 one-sided Jacobi SVDs of bidiagonal blocks with max_sweeps=80 and
basis completion, Jacobi eigensolves of tridiagonal leaves, masked null-column completion.

Bars: LAPACK to 1e-13 (FP64); complete_null_columns equal to the oracle's restatement (same
algorithm and order) to 1e-12, kept columns untouched, result orthonormal to 1e-13.
"""
from importlib import import_module

import numpy as np
import pytest
import scipy.linalg
import torch

from conftest import load_package
from oracle import restatement as R

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'


@pytest.fixture(scope='module')
def P():
    load_package()
    return import_module('pavel-s_svd_amd.standalone.primitives')


def _bidiag(n, m, rng):
    b = np.zeros((n, m))
    b[np.arange(n), np.arange(n)] = rng.standard_normal(n)
    b[np.arange(min(n, m - 1)), np.arange(1, min(n, m - 1) + 1)] = rng.standard_normal(min(n, m - 1))
    return b


@pytest.mark.parametrize('n,m', [(1, 1), (1, 2), (3, 3), (3, 4), (4, 4), (4, 5)])
def test_dc_leaf_svd_and_completion(P, n, m):
    rng = np.random.default_rng(n * 10 + m)
    B = torch.from_numpy(_bidiag(n, m, rng)).to(DEV)
    if m > n:   # divide_conquer.py:79-81
        v, s, u, sweeps = P.jacobi_svd(B.T[None], max_sweeps=80)
        v = P.complete_basis(v, m)
        u, s, v = u[0], s[0], v[0]
    else:       # :82
        u, s, v, sweeps = P.jacobi_svd(B[None], max_sweeps=80)
        u, s, v = u[0], s[0], v[0]
    b = B.cpu().numpy()
    un, sn, vn = u.cpu().numpy(), s.cpu().numpy(), v.cpu().numpy()
    assert np.allclose(sn, scipy.linalg.svdvals(b), atol=1e-13)
    assert np.allclose((un * sn) @ vn[:, :n].T, b, atol=1e-13)
    assert np.allclose(vn.T @ vn, np.eye(m), atol=1e-13) and np.allclose(un.T @ un, np.eye(n), atol=1e-13)


@pytest.mark.parametrize('n', [1, 2, 3, 4])
def test_dc_leaf_eigh(P, n):
    rng = np.random.default_rng(n)
    d, e = rng.standard_normal(n), rng.standard_normal(n - 1)
    h = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    w, v, sweeps = P.jacobi_eigh(torch.from_numpy(h[None]).to(DEV), max_sweeps=80)   # :127
    assert np.allclose(w[0].cpu().numpy(), scipy.linalg.eigvalsh(h), atol=1e-13)
    vv = v[0].cpu().numpy()
    assert np.allclose((vv * w[0].cpu().numpy()) @ vv.T, h, atol=1e-13)


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_complete_null_columns_matches_oracle(P, dtype):
    rng = np.random.default_rng(5)
    q = scipy.linalg.qr(rng.standard_normal((12, 4)), mode='economic')[0]
    u = np.zeros((2, 12, 7))
    keep = np.array([[True, False, True, False, False, True, True], [False, True, True, True, False, False, True]])
    for i in range(2):
        u[i][:, keep[i]] = q
    ref = R.complete_null_columns(u.copy(), keep)
    out = P.complete_null_columns(torch.from_numpy(u).to(DEV, dtype), torch.from_numpy(keep).to(DEV)).cpu().numpy()
    tol = 1e-12 if dtype == torch.float64 else 2e-6
    assert np.max(np.abs(out - ref)) < tol
    for i in range(2):
        assert np.array_equal(out[i][:, keep[i]], u[i][:, keep[i]].astype(out.dtype))
        assert np.allclose(out[i].T @ out[i], np.eye(7), atol=tol * 10)
    full = torch.eye(3, 4, dtype=torch.float64, device=DEV)[None]
    with pytest.raises(RuntimeError):
        P.complete_null_columns(full, torch.tensor([[True, True, True, False]], device=DEV))


@pytest.mark.parametrize('m', [3556, 3557])
def test_complete_columns_lds_and_global_variants(P, m):
    """The completion kernel keeps v, the leverages and y in LDS while 8 (2m + n) bytes plus its
    8384 static bytes fit in 64 KiB: at n = 32 that is m <= 3556, so the two sizes run the LDS and the
    global-memory variant of the same kernel (9 matrices: more than the whole-GPU path takes)."""
    rng = np.random.default_rng(m)
    b, n, k = 9, 32, 8
    u = np.zeros((b, m, n))
    keep = np.zeros((b, n), dtype=bool)
    keep[:, :k] = True
    for i in range(b):
        u[i][:, :k] = scipy.linalg.qr(rng.standard_normal((m, k)), mode='economic')[0]
    ref = R.complete_null_columns(u.copy(), keep)
    out = P.complete_null_columns(torch.from_numpy(u).to(DEV), torch.from_numpy(keep).to(DEV)).cpu().numpy()
    assert np.max(np.abs(out - ref)) < 1e-12
    for i in range(b):
        assert np.allclose(out[i].T @ out[i], np.eye(n), atol=1e-13)
