"""GPU parity of the blocked Householder QR against the oracle (the reference's unblocked
Householder QR, same reflector convention) and LAPACK.

Bars: R equals the oracle's R (signs included) to 1e-12 (FP64) / 2e-5 (FP32) relative to |A|;
|QR - A| / |A| and |Q^T Q - I| to the same bars.
"""
from importlib import import_module

import numpy as np
import pytest
import scipy.linalg
import torch

from oracle import restatement as R

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'


@pytest.fixture(scope='module')
def prims():
    return import_module('pavel-s_svd_amd.standalone.primitives')


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('b,m,n', [(1, 5, 3), (2, 40, 40), (3, 100, 33), (1, 300, 70), (1, 1000, 64), (2, 257, 129),
                                   (2, 3000, 70), (1, 9000, 40), (4, 20000, 33), (8, 20000, 8)])
def test_qr_matches_oracle(prims, dtype, b, m, n):
    rng = np.random.default_rng(m + n)
    a = rng.standard_normal((b, m, n)).astype(dtype)
    q, r = prims.qr(torch.from_numpy(a).to(DEV))
    q, r = q.cpu().numpy().astype(np.float64), r.cpu().numpy().astype(np.float64)
    bar = 1e-12 if dtype == np.float64 else 2e-5
    scale = np.abs(a).max()
    if m * n <= 40 * 40 * 2:
        qo, ro = R.qr(a)
    else:
        ro = np.stack([scipy.linalg.qr(x.astype(np.float64), mode='economic')[1] for x in a])
    assert np.max(np.abs(r - ro)) / scale < bar * np.sqrt(m)
    assert np.allclose(np.triu(r), r)
    assert np.max(np.abs(q @ r - a)) / scale < bar * np.sqrt(m)
    eye = np.eye(n)
    assert np.max(np.abs(np.swapaxes(q, 1, 2) @ q - eye)) < bar * np.sqrt(m)


def test_qr_rank_deficient_and_zero_columns(prims):
    a = np.zeros((50, 6))
    a[:, 1] = np.arange(50)
    a[:, 3] = a[:, 1] * 2
    q, r = prims.qr(torch.from_numpy(a[None]).to(DEV))
    qo, ro = R.qr(a[None])
    assert np.allclose(r.cpu().numpy(), ro, atol=1e-10)
    assert np.allclose(q.cpu().numpy() @ r.cpu().numpy(), a, atol=1e-10)


def test_qr_tall_rank_deficient_multi_workgroup_panel(prims):
    # panels taller than one workgroup: trivial reflectors (zero columns) and a dependent column
    a = np.zeros((3000, 6))
    a[:, 1] = np.sin(np.arange(3000))
    a[:, 3] = a[:, 1] * 2
    a[:, 4] = np.cos(np.arange(3000))
    for dt in (np.float64, np.float32):
        x = a.astype(dt)
        q, r = prims.qr(torch.from_numpy(x[None]).to(DEV))
        ro = scipy.linalg.qr(x.astype(np.float64), mode='economic')[1]
        bar = 1e-12 if dt == np.float64 else 2e-5
        q, r = q.cpu().numpy()[0].astype(np.float64), r.cpu().numpy()[0].astype(np.float64)
        assert np.max(np.abs(q @ r - x)) < bar * 60
        assert np.max(np.abs(q.T @ q - np.eye(6))) < bar * 60
        # R is unique up to the dependent column 3; past it the reflector of a rounding-noise
        # column (|x2| ~ 1e-14, not exactly zero) points in a noise direction in any implementation
        assert np.max(np.abs(r[:3, :4] - ro[:3, :4])) < bar * 60 * np.abs(ro).max()


def test_qr_bf16x3_panel_updates(prims):
    rng = np.random.default_rng(3)
    a = rng.standard_normal((1, 512, 96)).astype(np.float32)
    q, r = prims.qr(torch.from_numpy(a).to(DEV), gemm='bf16x3')
    q, r = q.cpu().numpy().astype(np.float64), r.cpu().numpy().astype(np.float64)
    assert np.max(np.abs(q @ r - a)) / np.abs(a).max() < 3e-4
    assert np.max(np.abs(q[0].T @ q[0] - np.eye(96))) < 3e-4


@pytest.mark.parametrize('method', ['qr', 'polar_qdwh'])
def test_solvers_take_batches_beyond_one_launch(method):
    """70 000 float32 6 x 3 matrices: batches past the 65 535 matrices one launch of the batched
    utility and QR kernels indexes are solved in slices; every matrix's sigma against LAPACK."""
    from conftest import load_package
    pkg = load_package()
    rng = np.random.default_rng(29)
    a = rng.standard_normal((70000, 6, 3)).astype(np.float32)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), method, return_info=True)
    ref = np.linalg.svd(a.astype(np.float64), compute_uv=False)
    assert tuple(s.shape) == (70000, 3) and tuple(u.shape) == (70000, 6, 3) and tuple(vh.shape) == (70000, 3, 3)
    assert np.abs(s.cpu().numpy() - ref).max() <= 2e-5 * ref.max()
    assert info['algorithm'] == method


def _graded(m, n, cond, seed, dtype):
    """A = U diag(sigma) V^T, sigma geometric from 1 to 1/cond, Haar U and V (seeded)."""
    rng = np.random.default_rng(seed)
    u = np.linalg.qr(rng.standard_normal((m, n)))[0]
    v = np.linalg.qr(rng.standard_normal((n, n)))[0]
    return ((u * np.geomspace(1.0, 1.0 / cond, n)) @ v.T).astype(dtype)


@pytest.mark.parametrize('dtype,cond', [(np.float64, 1e2), (np.float64, 1e6), (np.float64, 1e11),
                                        (np.float32, 1e2), (np.float32, 1e5)])
def test_qr_panel_paths_by_condition(prims, dtype, cond):
    """Tall panels across the CholeskyQR panel's branches: well conditioned (second pass to first order),
    moderately conditioned (second Cholesky factor) and, in FP64, beyond its reach (G2 = Q1^T Q1 far from
    I: the column-by-column panel).  Every branch must give LAPACK's factorisation: Q R = A and
    Q^T Q = I to the dtype's bar, R's diagonal signs equal LAPACK's (xLARFG convention), and R equal to
    LAPACK's R within the condition number's amplification of the bar."""
    m, n = 2000, 64
    a = _graded(m, n, cond, 7, dtype)
    q, r = prims.qr(torch.from_numpy(a[None]).to(DEV))
    q, r = q.cpu().numpy()[0].astype(np.float64), r.cpu().numpy()[0].astype(np.float64)
    ro = scipy.linalg.qr(a.astype(np.float64), mode='economic')[1]
    bar = 1e-12 if dtype == np.float64 else 2e-5
    scale = np.abs(a).max()
    assert np.max(np.abs(q @ r - a)) / scale < bar * np.sqrt(m)
    assert np.max(np.abs(q.T @ q - np.eye(n))) < bar * np.sqrt(m)
    assert np.array_equal(np.sign(np.diag(r)), np.sign(np.diag(ro)))
    assert np.max(np.abs(r - ro)) / np.abs(ro).max() < bar * np.sqrt(m) * cond
