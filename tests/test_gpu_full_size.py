"""Parity at the benchmark configurations' full sizes, checked on the host in FP64.

* QDWH polar SVD of the 4096 x 4096 float32 matrix U diag(sigma) V^T (Haar U, V from numpy's QR with
  the R-diagonal sign fix, sigma log-uniform on [1e-6, 1] with 10 % in a width-1e-4 cluster at 0.5,
  FP64 product rounded once), generated on the host by the oracle's workload generator:
  - the MI355X form (gemm='f32', polish=True, cut='floor') and the same with BF16x3 products:
    sigma against the generating spectrum within 1e-5 sigma_1 (rounding A to FP32 moves sigma by at
    most |A - fl(A)|_2 <= 2^-24 |A|_F, about 1e-6 here), |U^T U - I|_F / sqrt(n) and the same for V
    below 1e-5, reconstruction |A - U S V^T|_F / |A|_F below 1e-4 (the bar the BF16x3 products'
    2^-16 backward error allows);
  - the default options (the reference's cut at eps max(m, n) sigma_1, FP32 products, no polish):
    the kept sigma within the reference's FP32 envelope 3e-4 sigma_1, the cut ones exactly 0 and
    below the cut in the generating spectrum, U (with its completed columns) and V orthonormal to
    3e-4, reconstruction within the dropped part's norm plus 3e-4.
* TSQR of the 8192 x 2048 float32 rank-64 + N(0, 1e-8) matrix with tile_rows=256 (2048-row leaves):
  sigma against FP64 LAPACK within 1e-5 sigma_1, orthogonality below 1e-5 (the polished forms, FP32
  or BF16x3 products) / 1e-4 (the default FP32 form), reconstruction below 1e-4.
* Learned l0 at 1024 x 1024: the packaged model on 16 matrices of its test split; QDWH from the
  learned bound takes fewer iterations on average than from the reference's eps, at equal accuracy
  against FP64 LAPACK (each gate within twice the eps run's, or 2e-5).
"""
import importlib
import json

import numpy as np
import pytest
import scipy.linalg
import torch

from conftest import load_package
from oracle import matrices as X
from oracle import metrics as M

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'


@pytest.fixture(scope='module')
def pkg():
    return load_package()


def _orth(q):
    q = np.asarray(q, np.float64)
    g = q.T @ q
    g[np.diag_indices_from(g)] -= 1.0
    return float(np.linalg.norm(g) / np.sqrt(q.shape[1]))


def _recon(a64, u, s, vh):
    return float(np.linalg.norm(a64 - (u.astype(np.float64) * s.astype(np.float64)) @ vh.astype(np.float64))
                 / np.linalg.norm(a64))


@pytest.fixture(scope='module')
def qdwh_case():
    a, sig = X.workload_qdwh(4096)
    return a, sig


@pytest.mark.parametrize('form,gemm', [('mi355x', 'f32'), ('bf16x3_polish', 'bf16x3')])
def test_qdwh_4096_polished_forms(pkg, qdwh_case, form, gemm):
    a, sig = qdwh_case
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'polar_qdwh', gemm=gemm, polish=True, cut='floor',
                             return_info=True)
    u, s, vh = u.cpu().numpy(), s.cpu().numpy().astype(np.float64), vh.cpu().numpy()
    err = float(np.max(np.abs(s - sig)) / sig[0])
    ou, ov, rec = _orth(u), _orth(vh.T), _recon(a.astype(np.float64), u, s, vh)
    print(f"QDWH 4096 {form}: sigma {err:.2e} orth U {ou:.2e} V {ov:.2e} recon {rec:.2e} "
          f"iterations {info['polar_iterations']} eigen sweeps {info['eigen_sweeps']}")
    assert err <= 1e-5 and ou <= 1e-5 and ov <= 1e-5 and rec <= 1e-4


def test_qdwh_4096_default_options(pkg, qdwh_case):
    a, sig = qdwh_case
    n = a.shape[0]
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'polar_qdwh', return_info=True)
    u, s, vh = u.cpu().numpy(), s.cpu().numpy().astype(np.float64), vh.cpu().numpy()
    cut = np.finfo(np.float32).eps * n * s[0]
    kept = s > 0
    assert kept.sum() >= 1 and np.all(s[~kept] == 0)
    kept_err = float(np.max(np.abs(s[kept] - sig[kept])) / sig[0])
    assert kept_err <= 3e-4, kept_err
    assert np.all(sig[~kept] <= cut + 3e-4 * sig[0])
    ou, ov = _orth(u), _orth(vh.T)
    dropped = float(np.linalg.norm(sig[~kept]) / np.linalg.norm(sig))
    rec = _recon(a.astype(np.float64), u, s, vh)
    print(f"QDWH 4096 default: kept {int(kept.sum())} sigma {kept_err:.2e} orth U {ou:.2e} V {ov:.2e} "
          f"recon {rec:.2e} (dropped part {dropped:.2e})")
    assert ou <= 3e-4 and ov <= 3e-4
    assert rec <= dropped + 3e-4


@pytest.fixture(scope='module')
def tsqr_case():
    a = X.workload_tall_lowrank()
    return a, scipy.linalg.svdvals(a.astype(np.float64))


@pytest.mark.parametrize('form,opts,orth_bar', [('mi355x', dict(gemm='f32', polish=True), 1e-5),
                                                ('bf16x3_polish', dict(gemm='bf16x3', polish=True), 1e-5),
                                                ('default', dict(), 1e-4)])
def test_tsqr_8192x2048(pkg, tsqr_case, form, opts, orth_bar):
    a, ref = tsqr_case
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'tsqr', tile_rows=256, cut='floor', return_info=True,
                             **opts)
    assert info['tsqr_leaves'] == 4          # rows = max(tile_rows, n) = 2048 (solvers.py:73)
    u, s, vh = u.cpu().numpy(), s.cpu().numpy().astype(np.float64), vh.cpu().numpy()
    err = float(np.max(np.abs(s - ref)) / ref[0])
    ou, ov, rec = _orth(u), _orth(vh.T), _recon(a.astype(np.float64), u, s, vh)
    print(f"TSQR 8192x2048 {form}: sigma {err:.2e} orth U {ou:.2e} V {ov:.2e} recon {rec:.2e} "
          f"sweeps {info['jacobi_sweeps']}")
    assert err <= 1e-5 and ou <= orth_bar and ov <= orth_bar and rec <= 1e-4


def test_learned_l0_1024_test_split(pkg):
    T = importlib.import_module('pavel-s_svd_amd.train_l0')
    L = importlib.import_module('pavel-s_svd_amd.learned_l0')
    with open(L.DEFAULT_MODEL) as f:
        ev = json.load(f)['evaluation']
    start, n = ev['train'] + ev['holdout'], ev['n']
    a = T.generate(start, start + 16, n)
    it_eps, it_learned, sources = [], [], []
    for j in range(a.shape[0]):
        x = a[j]
        u0, s0, vh0, i0 = pkg.svd(x, 'polar_qdwh', return_info=True)
        u1, s1, vh1, i1 = pkg.svd(x, 'polar_qdwh', return_info=True, l0_model=True)
        it_eps.append(i0['polar_iterations'])
        it_learned.append(i1['polar_iterations'])
        sources.append(i1['l0_source'])
        host = x.cpu().numpy()
        ref = M.reference(host)
        m0 = M.validate(host, (u0.cpu().numpy(), s0.cpu().numpy(), vh0.cpu().numpy()), ref)['worst']
        m1 = M.validate(host, (u1.cpu().numpy(), s1.cpu().numpy(), vh1.cpu().numpy()), ref)['worst']
        for k in ('spectrum_error_over_sigma1', 'orthogonality_u_fro_over_sqrt_cols', 'reconstruction_relative_fro'):
            assert m1[k] <= max(2 * m0[k], 2e-5), (j, k, m0[k], m1[k])
    print('iterations from eps', it_eps, 'from the learned bound', it_learned, 'sources', sources)
    assert np.mean(it_learned) < np.mean(it_eps)
    assert sum(l > e for l, e in zip(it_learned, it_eps)) <= 1
