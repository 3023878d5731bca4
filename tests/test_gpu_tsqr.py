"""GPU parity of the TSQR and QR families (leaf/merge tree, Jacobi on R, U = Q U_R) against the
oracle and FP64 LAPACK, on the reference's corpus and on edge shapes of the leaf rule.

Bars: sigma error over sigma_1, U/V orthogonality and reconstruction <= 1e-5 (FP32) / 1e-12
(FP64); leaf counts equal the reference's; sweep counts within 2 of the reference's.
"""
import numpy as np
import pytest
import torch

from conftest import corpus, load_package
from oracle import matrices as X
from oracle import metrics as M
from oracle import restatement as R

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'
KEYS = ('spectrum_error_over_sigma1', 'orthogonality_u_fro_over_sqrt_cols', 'orthogonality_v_fro_over_sqrt_rows',
        'reconstruction_relative_fro')


@pytest.fixture(scope='module')
def pkg():
    return load_package()


def _metrics(a, out):
    return M.validate(a, tuple(None if t is None else t.cpu().numpy() for t in out), M.reference(a))['worst']


@pytest.mark.parametrize('method', ['tsqr', 'qr'])
@pytest.mark.parametrize('case', ['square32', 'batch8x16', 'tall256x16', 'tall2048x32', 'condition1e8',
                                  'condition1e12', 'rank_deficient', 'zero'])
def test_corpus(pkg, expected, method, case):
    a = corpus(case)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), method, return_info=True)
    met = _metrics(a, (u, s, vh))
    bar = 1e-5 if a.dtype == np.float32 else 1e-12
    for k in KEYS:
        if met[k] is not None:
            assert met[k] <= bar, (k, met[k])
    runs = [r for r in expected[case]['runs'] if r['method'] == 'tsqr']
    if method == 'tsqr' and runs and 'tsqr_leaves' in runs[0]['diagnostics']:
        assert info['tsqr_leaves'] == runs[0]['diagnostics']['tsqr_leaves']
    if method == 'tsqr' and runs:
        assert abs(info['jacobi_sweeps'] - runs[0]['diagnostics']['jacobi_sweeps']) <= 2


@pytest.mark.parametrize('m,n,rows', [(100, 6, 24), (130, 7, 7), (20, 8, 4), (64, 64, 8), (1000, 40, 100)])
def test_leaf_rule_edges_match_oracle(pkg, m, n, rows):
    a = X.make_matrix(dict(m=m, n=n, family='gaussian', dtype='float64', seed=m + n))
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'tsqr', tile_rows=rows, return_info=True)
    _, rs, _, rinfo = R.svd(a, 'tsqr', tile_rows=rows, return_info=True)
    assert info.get('tsqr_leaves') == rinfo.get('tsqr_leaves')
    assert np.max(np.abs(s.cpu().numpy() - rs)) / rs[0] < 1e-12
    met = _metrics(a, (u, s, vh))
    assert all(met[k] <= 1e-12 for k in KEYS), met


def test_wide_and_batched_and_values_only(pkg):
    a = X.make_matrix(dict(m=12, n=40, batch=3, family='gaussian', dtype='float32', seed=9))
    t = torch.from_numpy(a).to(DEV)
    u, s, vh = pkg.svd(t, 'tsqr', tile_rows=8)
    assert tuple(u.shape) == (3, 12, 12) and tuple(vh.shape) == (3, 12, 40)
    met = _metrics(a, (u, s, vh))
    assert all(met[k] <= 1e-5 for k in KEYS), met
    none, s2, none2 = pkg.svd(t, 'qr', values_only=True)
    assert none is None and none2 is None
    assert torch.allclose(s2, s, rtol=0, atol=1e-5 * float(s.max()))


def test_tall_lowrank_workload_small(pkg):
    """The tall low-rank workload's structure at reduced size: 2048 x 256, rank 16, noise 1e-4,
    256-row tiles (clamped to n by the leaf rule)."""
    a = X.workload_tall_lowrank(m=2048, n=256, rank=16)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'tsqr', tile_rows=256, cut='floor', return_info=True)
    assert info['tsqr_leaves'] == 8
    met = _metrics(a, (u, s, vh))
    assert all(met[k] <= 1e-5 for k in KEYS), met


@pytest.mark.parametrize('gmode', ['bf16x3', 'f32'])
def test_tsqr_polish_restores_orthogonality(pkg, gmode):
    """polish=True: one FP64 Newton-Schulz step on U and on V brings both to FP32 rounding level."""
    a = np.random.default_rng(9).standard_normal((3000, 300)).astype(np.float32)
    u, s, vh = pkg.svd(torch.from_numpy(a).to(DEV), 'tsqr', tile_rows=600, gemm=gmode, polish=True)
    met = _metrics(a, (u, s, vh))
    assert met['orthogonality_u_fro_over_sqrt_cols'] <= 1e-6, met
    assert met['orthogonality_v_fro_over_sqrt_rows'] <= 1e-6, met
    assert met['spectrum_error_over_sigma1'] <= 1e-5 and met['reconstruction_relative_fro'] <= 3e-5, met
