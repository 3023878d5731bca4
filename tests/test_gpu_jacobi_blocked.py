"""GPU parity of the blocked one-sided Jacobi kernel (column blocks of 16 or 32: 32 x 32 or 64 x 64
FP64 Gram subproblems, MFMA column updates) against FP64 LAPACK, the oracle and the single-launch
kernel that keeps the reference's rotation order.

Bars: sigma error over sigma_1, orthogonality of U and V, reconstruction: 1e-5 (FP32) / 1e-12 (FP64);
sigma equal to the single-launch kernel's within the same bars; the reference cut zeroes the same
singular values as the oracle (+-2 at the cut boundary).
"""
from importlib import import_module

import numpy as np
import pytest
import torch

from conftest import load_package
from oracle import matrices as X
from oracle import metrics as M
from oracle import restatement as R

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'
KEYS = ('spectrum_error_over_sigma1', 'orthogonality_u_fro_over_sqrt_cols', 'orthogonality_v_fro_over_sqrt_rows',
        'reconstruction_relative_fro')


@pytest.fixture(scope='module')
def P():
    load_package()
    return import_module('pavel-s_svd_amd.standalone.primitives')


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize('width', [16, 32])
@pytest.mark.parametrize('dtype', [np.float32, np.float64])
@pytest.mark.parametrize('m,n', [(64, 64), (65, 64), (200, 100), (300, 256), (1000, 300), (700, 129)])
def test_blocked_matches_lapack(P, dtype, m, n, width):
    rng = np.random.default_rng(m * 7 + n)
    a = rng.standard_normal((1, m, n)).astype(dtype)
    u, s, vh, info = P.jacobi_blocked(torch.from_numpy(a).to(DEV), scale=True, block_width=width)
    assert int(info[0, 1]) == 0 and 1 <= int(info[0, 0]) < 60
    met = M.validate(a[0], (_np(u[0]), _np(s[0]), _np(vh[0])), M.reference(a[0]))['worst']
    bar = 1e-5 if dtype == np.float32 else 1e-12
    for k in KEYS:
        assert met[k] <= bar, (k, met[k])
    assert np.all(np.diff(_np(s[0])) <= 0)


@pytest.mark.parametrize('width', [16, 32])
def test_blocked_equals_single_launch_kernel_and_cut(P, width):
    a = X.workload_single_jacobi()
    x = torch.from_numpy(a[None]).to(DEV)
    _, s1, _, i1 = P.jacobi_blocked(x, tol=1e-7, scale=True, cut='floor', block_width=width)
    _, s2, _, i2 = P.jacobi_grid(x, tol=1e-7, scale=True, cut='floor')
    assert np.max(np.abs(_np(s1) - _np(s2))) <= 2e-6
    # reference cut: the same sigma are zeroed as by the oracle
    _, s3, _, _ = P.jacobi_blocked(x, scale=True, block_width=width)
    _, rs, _ = R.svd(a, 'jacobi', max_sweeps=60)
    assert abs(int((_np(s3) == 0).sum()) - int((rs == 0).sum())) <= 2


def test_blocked_rank_deficient_zero_and_values_only(P):
    rng = np.random.default_rng(3)
    a = (rng.standard_normal((300, 20)) @ rng.standard_normal((20, 150))).astype(np.float64)
    u, s, vh, info = P.jacobi_blocked(torch.from_numpy(a[None]).to(DEV), scale=True)
    assert int(info[0, 1]) == 0
    sv = _np(s[0])
    assert np.count_nonzero(sv) == 20
    uu = _np(u[0])
    assert np.max(np.abs(uu.T @ uu - np.eye(150))) < 1e-12          # completed null columns
    assert np.max(np.abs((uu * sv) @ _np(vh[0]) - a)) < 1e-11 * np.abs(a).max()
    z = torch.zeros(1, 130, 70, dtype=torch.float32, device=DEV)
    u, s, vh, info = P.jacobi_blocked(z, scale=True)
    assert int(info[0, 0]) == 1 and not s.abs().any()
    assert np.max(np.abs(_np(u[0]).T @ _np(u[0]) - np.eye(70))) < 1e-6
    b = rng.standard_normal((1, 400, 200)).astype(np.float32)
    _, s, _, _ = P.jacobi_blocked(torch.from_numpy(b).to(DEV), vectors=False, scale=True)
    ref = np.linalg.svd(b[0].astype(np.float64), compute_uv=False)
    assert np.max(np.abs(_np(s[0]) - ref)) / ref[0] < 1e-5


def test_svd_dispatches_large_matrices_to_blocked(P):
    pkg = load_package()
    a = np.random.default_rng(5).standard_normal((600, 200)).astype(np.float32)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'jacobi', return_info=True)
    u2, s2, vh2, info2 = pkg.svd(torch.from_numpy(a).to(DEV), 'jacobi', jacobi_method='grid', return_info=True)
    assert np.max(np.abs(_np(s) - _np(s2))) / float(s[0]) < 2e-6
    met = M.validate(a, (_np(u), _np(s), _np(vh)), M.reference(a))['worst']
    assert all(met[k] <= 1e-5 for k in KEYS), met


def test_blocked_sweep_limits_queued_sweep_and_inner_sweeps(P):
    """The next sweep is queued before the host reads the convergence flag: a solve capped at its own
    sweep count runs the same rounds as the uncapped one (whose extra sweep turns into no-ops), so the
    outputs are bit-identical; a solve stopped early still has A V = W exactly (U diag(s) V^T = A
    to rounding, V orthogonal); two inner sweeps per pair and the full pair sweeps stay at LAPACK's
    accuracy."""
    rng = np.random.default_rng(11)
    a = (rng.standard_normal((300, 128)) * np.logspace(0, -3, 128)).astype(np.float64)
    x = torch.from_numpy(a[None]).to(DEV)
    u, s, vh, info = P.jacobi_blocked(x, scale=True)
    sw = int(info[0, 0])
    assert int(info[0, 1]) == 0 and 2 < sw < 60
    u2, s2, vh2, info2 = P.jacobi_blocked(x, scale=True, max_sweeps=sw)
    assert _np(info2).tolist() == _np(info).tolist()
    for t1, t2 in ((u, u2), (s, s2), (vh, vh2)):
        assert torch.equal(t1, t2)
    for cap in (1, 2):
        u3, s3, vh3, info3 = P.jacobi_blocked(x, scale=True, max_sweeps=cap)
        assert _np(info3[0]).tolist() == [cap, 1]   # PSVD_INFO_NOCONV
        v3 = _np(vh3[0]).T
        assert np.max(np.abs(v3.T @ v3 - np.eye(128))) < 1e-12
        rec = (_np(u3[0]) * _np(s3[0])) @ _np(vh3[0])
        assert np.linalg.norm(rec - a) / np.linalg.norm(a) < 1e-13
    ref = M.reference(a)
    for kw in (dict(inner_sweeps=2), dict(full_pairs=True), dict(inner_sweeps=2, full_pairs=True)):
        u4, s4, vh4, info4 = P.jacobi_blocked(x, scale=True, **kw)
        assert int(info4[0, 1]) == 0, kw
        met = M.validate(a, (_np(u4[0]), _np(s4[0]), _np(vh4[0])), ref)['worst']
        assert all(met[k] <= 1e-12 for k in KEYS), (kw, met)


def test_blocked_wide_rank_sort_chunks(P):
    """n above the 2048-value LDS chunk of the sigma rank sort (two chunks) and m * n above what the
    max|a| launch covers in one pass per thread: sigma against LAPACK, sorted descending."""
    rng = np.random.default_rng(21)
    a = rng.standard_normal((1, 2200, 2100)).astype(np.float32)
    _, s, _, info = P.jacobi_blocked(torch.from_numpy(a).to(DEV), vectors=False, scale=True)
    assert int(info[0, 1]) == 0
    sv = _np(s[0])
    ref = np.linalg.svd(a[0].astype(np.float64), compute_uv=False)
    assert np.all(np.diff(sv) <= 0)
    assert np.max(np.abs(sv - ref)) / ref[0] < 1e-5


@pytest.mark.parametrize('width', [16, 32])
def test_blocked_unscaled_fp32_extremes(P, width):
    """Unscaled FP32 data near 1e-25 and 1e20 (jacobi_blocked does not scale by default): the FP64
    Gram entries leave the FP32 range, the rotation angle then takes its FP64 quotient, and U, sigma
    and V stay finite and equal to the unit-scale problem's."""
    rng = np.random.default_rng(33)
    a = rng.standard_normal((1, 300, 144)).astype(np.float32)
    ref = np.linalg.svd(a[0].astype(np.float64), compute_uv=False)
    for factor in (1e-25, 1e20):
        x = torch.from_numpy(a * np.float32(factor)).to(DEV)
        u, s, vh, info = P.jacobi_blocked(x, block_width=width)
        assert int(info[0, 1]) == 0, factor
        un, sv, vn = _np(u[0]).astype(np.float64), _np(s[0]).astype(np.float64), _np(vh[0]).astype(np.float64)
        assert np.isfinite(un).all() and np.isfinite(sv).all() and np.isfinite(vn).all(), factor
        assert np.max(np.abs(sv / factor - ref)) / ref[0] < 1e-5, factor
        assert np.linalg.norm(un.T @ un - np.eye(144)) / np.sqrt(144) < 1e-5, factor
        assert np.linalg.norm(vn @ vn.T - np.eye(144)) / np.sqrt(144) < 1e-5, factor
