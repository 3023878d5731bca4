"""Pins the oracle's restatement of the hot path's consumers -- block-pair Jacobi, Gram + Jacobi
eigensolver, randomized / subspace / block Krylov range finders -- against the reference's golden
rows (counts and accuracy recorded for the corpus inputs), LAPACK and closed forms.

Golden counts: block_sweeps, jacobi_sweeps, eigen_sweeps, gram_discarded_values must equal the
recorded ones for every device row of block_jacobi and gram_jacobi.  The partial methods draw their
sketch from torch's generator: CPU rows are reproduced with torch's CPU generator (sketch and range
sizes equal, jacobi_sweeps within one sweep: the reference's CPU products ran on MKL, the oracle's
on OpenBLAS); CUDA rows used the CUDA generator and are not reproducible here.
"""
import hashlib

import numpy as np
import pytest
import scipy.linalg

from conftest import corpus
from oracle import metrics as M
from oracle import restatement as R

RNG = np.random.default_rng(4321)
CONSUMERS = ('block_jacobi', 'gram_jacobi', 'randomized', 'subspace', 'block_krylov')
RUNS = [(c, m) for c in ('square32', 'square64', 'batch8x16', 'tall256x16', 'tall2048x32', 'condition1e8',
                         'condition1e12', 'rank_deficient', 'zero', 'clustered_complex', 'wide_complex', 'lowrank512')
        for m in CONSUMERS]


def _golden_runs(expected, case, method):
    runs = [r for r in expected[case]['runs'] if r['method'] == method]
    if method in R.PARTIAL:
        runs = [r for r in runs if r['device'] == 'cpu']
    return runs


@pytest.mark.parametrize('case,method', RUNS)
def test_consumer_counts_and_metrics_match_reference_golden(expected, case, method):
    runs = _golden_runs(expected, case, method)
    if not runs:
        pytest.skip('no reproducible golden row for this case')
    a = corpus(case)
    assert hashlib.sha256(a.tobytes(order='C')).hexdigest() == expected[case]['sha256']
    refs = M.reference(a)
    for run in runs:
        u, s, vh, info = R.svd(a, method, return_info=True, **run['options'])
        for key, want in run['diagnostics'].items():
            if key == 'jacobi_sweeps' and method in R.PARTIAL:
                assert abs(info[key] - want) <= 1, (key, info[key], want, run['source'])
            else:
                assert info[key] == want, (key, info[key], want, run['source'])
        met = M.validate(a, (u, s, vh), refs)['worst']
        floor = 64 * np.finfo(a.real.dtype).eps * max(a.shape[-2:]) ** 0.5
        for key in ('spectrum_error_over_sigma1', 'reconstruction_relative_fro',
                    'orthogonality_u_fro_over_sqrt_cols', 'orthogonality_v_fro_over_sqrt_rows'):
            assert met[key] <= max(3 * run['metrics'][key], floor), (key, met[key], run['metrics'][key], run['source'])


# --- block Jacobi -----------------------------------------------------------------------------
@pytest.mark.parametrize('width', [1, 2, 3, 5])
@pytest.mark.parametrize('shape', [(30, 13), (16, 16), (40, 9)])
def test_block_jacobi_matches_lapack(width, shape):
    a = RNG.standard_normal((2,) + shape)
    u, s, vh, info = R.svd(a, 'block_jacobi', block_size=width, return_info=True)
    for i in range(2):
        ref = scipy.linalg.svdvals(a[i])
        assert np.allclose(s[i], ref, atol=1e-13 * ref[0])
        assert np.allclose((u[i] * s[i]) @ vh[i], a[i], atol=1e-13)
        assert np.allclose(u[i].T @ u[i], np.eye(shape[1]), atol=1e-13)
    assert info['block_sweeps'] >= 1 and info['jacobi_sweeps'] >= 1


def test_block_jacobi_single_group_is_plain_jacobi_and_bad_width_raises():
    a = RNG.standard_normal((1, 12, 7)).astype(np.float32)
    s1 = R.svd(a, 'block_jacobi', block_size=7, return_info=True)
    s2 = R.svd(a, 'jacobi', return_info=True)
    assert np.array_equal(s1[1], s2[1]) and 'block_sweeps' not in s1[3]
    with pytest.raises(ValueError):
        R.svd(a, 'block_jacobi', block_size=0)
    with pytest.raises(RuntimeError):
        R.svd(RNG.standard_normal((30, 20)), 'block_jacobi', block_size=2, max_sweeps=1)


def test_block_jacobi_pair_solves_keep_descending_order_inside_pairs():
    # after a converged run every group pair is a sorted Jacobi SVD: within a group the column
    # norms of W = U diag(s) decrease and the final sort only permutes whole columns
    a = RNG.standard_normal((1, 24, 12))
    stats = {}
    u, s, v = R.block_jacobi(a, stats, {'block_size': 3})
    assert np.all(np.diff(s[0]) <= 0)
    assert np.allclose(a[0] @ v[0], u[0] * s[0], atol=1e-13)


# --- Gram + Jacobi ----------------------------------------------------------------------------
def test_gram_jacobi_matches_lapack_and_discards_the_null_space():
    a = RNG.standard_normal((3, 20, 8))
    u, s, vh, info = R.svd(a, 'gram_jacobi', return_info=True)
    for i in range(3):
        ref = scipy.linalg.svdvals(a[i])
        assert np.allclose(s[i], ref, rtol=1e-10)
        assert np.allclose((u[i] * s[i]) @ vh[i], a[i], atol=1e-12)
    assert info['gram_discarded_values'] == 0
    low = (RNG.standard_normal((1, 20, 3)) @ RNG.standard_normal((1, 3, 8)))
    u, s, vh, info = R.svd(low, 'gram_jacobi', return_info=True)
    assert info['gram_discarded_values'] == 5 and np.all(s[0, 3:] == 0)
    assert np.allclose(u[0].T @ u[0], np.eye(8), atol=1e-12)
    assert np.allclose(s[0, :3], scipy.linalg.svdvals(low[0])[:3], rtol=1e-10)


def test_gram_cut_is_eps_n_times_the_top_eigenvalue():
    # sigma^2 = w exactly on a diagonal matrix: sqrt(w) survives iff w > eps * n * w_1
    e = np.finfo(np.float64).eps
    d = np.sqrt(np.array([1.0, 0.5, 6 * e, 3 * e]))        # n = 4: cut at 4 eps
    u, s, vh, info = R.svd(np.diag(d)[None], 'gram_jacobi', return_info=True)
    assert info['gram_discarded_values'] == 1
    assert np.allclose(s[0, :3], d[:3], rtol=1e-12) and s[0, 3] == 0


# --- randomized / subspace / block Krylov -----------------------------------------------------
def test_partial_recovers_an_exact_low_rank_matrix():
    a = RNG.standard_normal((2, 40, 5)) @ RNG.standard_normal((2, 5, 30))
    for method in R.PARTIAL:
        u, s, vh, info = R.svd(a, method, rank=5, oversample=3, iterations=1, return_info=True)
        for i in range(2):
            assert np.allclose(s[i], scipy.linalg.svdvals(a[i])[:5], rtol=1e-11)
            assert np.allclose((u[i] * s[i]) @ vh[i], a[i], atol=1e-11)
        assert info['sketch_columns'] == 8 and info['power_iterations'] == 1


def test_partial_counts_and_warm_start():
    a = RNG.standard_normal((1, 50, 20))
    _, _, _, info = R.svd(a, 'block_krylov', rank=4, oversample=4, iterations=3, return_info=True)
    assert info['range_columns'] == 8 * 4 and info['sketch_columns'] == 8
    _, _, _, info = R.svd(a[:, :12], 'block_krylov', rank=4, oversample=4, iterations=3, return_info=True)
    assert info['range_columns'] == 12                  # the basis stops at the row count
    om = RNG.standard_normal((20, 8))
    r1 = R.svd(a, 'randomized', rank=4, oversample=4, initial_basis=om, return_info=True)
    r2 = R.svd(a, 'subspace', rank=4, oversample=4, initial_basis=om[None], return_info=True)
    assert r1[3]['warm_started'] and np.array_equal(r1[1], r2[1])
    with pytest.raises(ValueError):
        R.svd(a, 'randomized', rank=4, oversample=4, initial_basis=om[:, :5])


def test_partial_power_iterations_converge_to_the_top_subspace():
    # sigma_j = 2^-j: each power iteration multiplies the error of the rank-4 part by (s5/s4)^2
    q1 = scipy.linalg.qr(RNG.standard_normal((60, 30)), mode='economic')[0]
    q2 = scipy.linalg.qr(RNG.standard_normal((30, 30)))[0]
    sig = 2.0 ** -np.arange(30)
    a = ((q1 * sig) @ q2.T)[None]
    errs = [np.max(np.abs(R.svd(a, 'subspace', rank=4, oversample=2, iterations=it)[1][0] - sig[:4]))
            for it in (0, 2, 4)]
    assert errs[0] > errs[1] > errs[2] and errs[2] < 1e-9
