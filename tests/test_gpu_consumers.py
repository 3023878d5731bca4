"""GPU parity of the hot path's consumers -- block_jacobi (fused block-pair kernel), gram_jacobi
(MFMA Gram + Jacobi eigensolver) and the range finders randomized / subspace / block_krylov (MFMA
products, panel QR, Jacobi) -- against the oracle on the same inputs, the reference's golden rows
and FP64 LAPACK.

Bars: the reference's accuracy metrics within 3x of the recorded ones on the corpus (as for the hot
path); against LAPACK sigma error over sigma_1, U/V orthogonality and reconstruction <= 1e-5
(FP32) / 1e-12 (FP64), gram_jacobi's sigma to the Gram method's own bound eps * sigma_1^2 / sigma_i;
counts: block_sweeps, eigen_sweeps, gram_discarded_values, sketch/range columns equal to the
reference's, jacobi_sweeps within 1 (the pair solves run their test in FP64 on the Gram matrix, the
reference in the data precision).  The range finders take the oracle's sketch (initial_basis) so
that both sides start from the same test matrix.
"""
import numpy as np
import pytest
import scipy.linalg
import torch

from conftest import corpus, load_package
from oracle import metrics as M
from oracle import restatement as R

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'
KEYS = ('spectrum_error_over_sigma1', 'orthogonality_u_fro_over_sqrt_cols', 'orthogonality_v_fro_over_sqrt_rows',
        'reconstruction_relative_fro')
REAL = ('square32', 'square64', 'batch8x16', 'tall256x16', 'tall2048x32', 'condition1e8', 'condition1e12',
        'rank_deficient', 'zero')


@pytest.fixture(scope='module')
def pkg():
    return load_package()


def _np(t):
    return None if t is None else t.detach().cpu().numpy()


def _metrics(a, out):
    return M.validate(a, tuple(_np(t) for t in out), M.reference(a))['worst']


@pytest.mark.parametrize('method', ['block_jacobi', 'gram_jacobi'])
@pytest.mark.parametrize('case', REAL)
def test_corpus_counts_and_metrics_match_reference_golden(pkg, expected, method, case):
    a = corpus(case)
    print(f'case {case} method {method}', flush=True)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), method, return_info=True)
    met = _metrics(a, (u, s, vh))
    runs = [r for r in expected[case]['runs'] if r['method'] == method]
    assert runs
    for run in runs:
        for key, want in run['diagnostics'].items():
            slack = 1 if key == 'jacobi_sweeps' else 0
            assert abs(info[key] - want) <= slack, (key, info[key], want, run['source'])
        floor = 64 * np.finfo(a.dtype).eps * max(a.shape[-2:]) ** 0.5
        for key in KEYS:
            assert met[key] <= max(3 * run['metrics'][key], floor), (key, met[key], run['metrics'][key], run['source'])


@pytest.mark.parametrize('method', ['randomized', 'subspace', 'block_krylov'])
@pytest.mark.parametrize('case,opts', [('square32', {}), ('batch8x16', {}), ('tall2048x32', {}),
                                       ('lowrank512', {'rank': 8, 'oversample': 8, 'iterations': 2, 'seed': 31}),
                                       ('lowrank512', {'rank': 8, 'oversample': 8, 'iterations': 2, 'seed': 113})])
def test_range_finders_match_oracle_on_the_same_sketch(pkg, expected, method, case, opts):
    a = corpus(case)
    print(f'case {case} method {method} opts {opts}', flush=True)
    a3 = a[None] if a.ndim == 2 else a
    k = opts.get('rank', min(a.shape[-2:]))
    ell = min(k + opts.get('oversample', 8), min(a.shape[-2:]))
    omega = R.sketch(a3, ell, opts.get('seed', 20261011))
    ru, rs, rvh, rinfo = R.svd(a, method, return_info=True, **opts)       # the oracle draws the same sketch
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), method, return_info=True,
                             initial_basis=torch.from_numpy(omega[0] if a.ndim == 2 else omega), **opts)
    for key in ('sketch_columns', 'range_columns', 'power_iterations'):
        assert info[key] == rinfo[key], key
    assert info['warm_started'] is True
    assert abs(info['jacobi_sweeps'] - rinfo['jacobi_sweeps']) <= 1
    # both sides carry the FP32 range finder's own error (about 2e-6 sigma_1 on square32 in the
    # reference's recorded rows), in different rounding directions
    top = max(float(np.max(rs)), 1e-30)
    assert np.max(np.abs(_np(s) - rs)) / top <= 1e-5
    met = _metrics(a, (u, s, vh))
    ref = _metrics(a, (torch.from_numpy(ru), torch.from_numpy(rs), torch.from_numpy(rvh)))
    for key in KEYS:
        assert met[key] <= max(3 * ref[key], 1e-6), (key, met[key], ref[key])
    runs = [r for r in expected[case]['runs'] if r['method'] == method and r['device'] == 'cpu'
            and r['options'] == opts]
    for run in runs:   # the reference's own CPU row drew the same sketch
        for key in ('sketch_columns', 'range_columns', 'power_iterations'):
            assert info[key] == run['diagnostics'][key]
        assert abs(info['jacobi_sweeps'] - run['diagnostics']['jacobi_sweeps']) <= 1


@pytest.mark.parametrize('dtype', [np.float32, np.float64])
@pytest.mark.parametrize('width', [1, 2, 3, 4, 5, 8, 16, 32])
@pytest.mark.parametrize('shape', [(40, 23), (64, 64), (130, 70)])
def test_block_jacobi_matches_oracle_and_lapack(pkg, dtype, width, shape):
    rng = np.random.default_rng(width * 131 + shape[1])
    a = rng.standard_normal((2,) + shape).astype(dtype)
    print(f'block_jacobi width {width} shape {shape} dtype {np.dtype(dtype).name}', flush=True)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'block_jacobi', block_size=width, return_info=True)
    met = _metrics(a, (u, s, vh))
    bar = 1e-5 if dtype == np.float32 else 1e-12
    for key in KEYS:
        assert met[key] <= bar, (key, met[key])
    if shape[0] <= 64 and width <= 8:
        _, rs, _, rinfo = R.svd(a, 'block_jacobi', block_size=width, return_info=True)
        # the oracle runs in the data precision: its own distance to LAPACK widens the bar
        lap = np.stack([scipy.linalg.svdvals(x.astype(np.float64)) for x in a])
        oracle_err = np.max(np.abs(rs - lap)) / lap.max()
        assert np.max(np.abs(_np(s) - rs)) / rs.max() <= bar + oracle_err
        if -(-shape[1] // width) > 1:
            assert abs(info['block_sweeps'] - rinfo['block_sweeps']) <= 1, (info, rinfo)


def test_block_jacobi_at_the_single_matrix_shape(pkg):
    # 512 x 256 FP32, sigma geometric on [1e-6, 1] (the single-matrix workload's spectrum)
    rng = np.random.default_rng(7)
    q1 = scipy.linalg.qr(rng.standard_normal((512, 256)), mode='economic')[0]
    q2 = scipy.linalg.qr(rng.standard_normal((256, 256)))[0]
    sig = np.geomspace(1, 1e-6, 256)
    a = ((q1 * sig) @ q2.T).astype(np.float32)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'block_jacobi', return_info=True)
    met = _metrics(a, (u, s, vh))
    # sigma below eps * max(m, n) * sigma_1 = 6.1e-5 is zeroed by the reference's contract
    ref = scipy.linalg.svdvals(a.astype(np.float64))
    sv = _np(s)
    cut = np.finfo(np.float32).eps * 512 * ref[0]
    kept = sv > 0
    assert np.max(np.abs(sv[kept] - ref[kept])) <= 1e-6 * ref[0]
    assert np.all(ref[~kept] <= 1.01 * cut) and np.all(ref[kept] >= 0.99 * cut)
    assert met['orthogonality_u_fro_over_sqrt_cols'] <= 1e-5
    assert met['orthogonality_v_fro_over_sqrt_rows'] <= 1e-5
    tail = np.sqrt(np.sum(ref[~kept] ** 2)) / np.linalg.norm(ref)      # the zeroed sigma's share
    assert met['reconstruction_relative_fro'] <= tail + 1e-5
    assert 1 <= info['block_sweeps'] <= 30


def test_block_jacobi_special_inputs_and_errors(pkg):
    z = torch.zeros(1, 20, 9, device=DEV)
    u, s, vh, info = pkg.svd(z, 'block_jacobi', block_size=2, return_info=True)
    assert not s.any() and info['block_sweeps'] == 1
    assert np.allclose(_np(u[0]).T @ _np(u[0]), np.eye(9), atol=1e-6)
    rng = np.random.default_rng(11)
    low = (rng.standard_normal((50, 4)) @ rng.standard_normal((4, 30))).astype(np.float64)
    u, s, vh = pkg.svd(torch.from_numpy(low).to(DEV), 'block_jacobi', block_size=3)
    assert np.count_nonzero(_np(s)) == 4
    assert np.max(np.abs(_np(u).T @ _np(u) - np.eye(30))) < 1e-12
    assert np.max(np.abs((_np(u) * _np(s)) @ _np(vh) - low)) < 1e-12 * np.abs(low).max()
    wide = rng.standard_normal((12, 40)).astype(np.float32)
    u, s, vh = pkg.svd(torch.from_numpy(wide).to(DEV), 'block_jacobi', block_size=4)
    assert np.max(np.abs(_np(s) - scipy.linalg.svdvals(wide.astype(np.float64)))) < 1e-5 * _np(s)[0]
    x = torch.from_numpy(rng.standard_normal((30, 20)).astype(np.float32)).to(DEV)
    with pytest.raises(ValueError):
        pkg.svd(x, 'block_jacobi', block_size=0)
    with pytest.raises(pkg.Unsupported):
        pkg.svd(torch.zeros(1, 80, 70, device=DEV) + 1, 'block_jacobi', block_size=33)
    with pytest.raises(RuntimeError):
        pkg.svd(x.double(), 'block_jacobi', block_size=2, max_sweeps=1)


@pytest.mark.parametrize('dtype', [np.float32, np.float64])
@pytest.mark.parametrize('shape', [(40, 23), (300, 100), (1000, 200)])
def test_gram_jacobi_matches_oracle_and_lapack(pkg, dtype, shape):
    rng = np.random.default_rng(shape[1])
    a = rng.standard_normal((2,) + shape).astype(dtype)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'gram_jacobi', return_info=True)
    eps = np.finfo(dtype).eps
    for i in range(2):
        ref = scipy.linalg.svdvals(a[i].astype(np.float64))
        # the Gram method resolves sigma_i to about eps * sigma_1^2 / sigma_i
        assert np.all(np.abs(_np(s[i]) - ref) <= 50 * shape[1] * eps * ref[0] ** 2 / ref), i
    met = _metrics(a, (u, s, vh))
    assert met['orthogonality_v_fro_over_sqrt_rows'] <= (1e-5 if dtype == np.float32 else 1e-12)
    assert met['reconstruction_relative_fro'] <= (1e-5 if dtype == np.float32 else 1e-12)
    if shape[0] <= 300:
        _, rs, _, rinfo = R.svd(a, 'gram_jacobi', return_info=True)
        assert np.max(np.abs(_np(s) - rs)) / rs.max() <= (1e-5 if dtype == np.float32 else 1e-12)
        assert info['gram_discarded_values'] == rinfo['gram_discarded_values']
        # sweep counts in the reference's rotation order ('auto' takes the blocked eigensolver
        # once G no longer fits in LDS, whose sweeps are block sweeps)
        _, s2, _, info2 = pkg.svd(torch.from_numpy(a).to(DEV), 'gram_jacobi', eigh_method='cyclic', return_info=True)
        assert np.max(np.abs(_np(s2) - rs)) / rs.max() <= (1e-5 if dtype == np.float32 else 1e-12)
        assert abs(info2['eigen_sweeps'] - rinfo['eigen_sweeps']) <= 1


def test_gram_jacobi_discards_the_null_space(pkg):
    # integer factors: A and its Gram matrix are exact in FP64, so the 19 null eigenvalues are
    # rounding noise of the eigensolver alone, far below the eps * n * w_1 cut
    rng = np.random.default_rng(2)
    low = (rng.integers(-3, 4, (3, 60, 5)) @ rng.integers(-3, 4, (3, 5, 24))).astype(np.float64)
    u, s, vh, info = pkg.svd(torch.from_numpy(low).to(DEV), 'gram_jacobi', return_info=True)
    assert info['gram_discarded_values'] == 3 * 19
    for i in range(3):
        assert np.all(_np(s[i, 5:]) == 0)
        assert np.max(np.abs(_np(u[i]).T @ _np(u[i]) - np.eye(24))) < 1e-12
        assert np.max(np.abs((_np(u[i]) * _np(s[i])) @ _np(vh[i]) - low[i])) < 1e-11 * np.abs(low[i]).max()


def test_range_finders_own_sketch_and_errors(pkg):
    rng = np.random.default_rng(9)
    a = (rng.standard_normal((1, 400, 6)) @ rng.standard_normal((1, 6, 300))).astype(np.float32)
    x = torch.from_numpy(a).to(DEV)
    for method in ('randomized', 'subspace', 'block_krylov'):
        u, s, vh, info = pkg.svd(x, method, rank=6, return_info=True)   # sketch from the device generator
        ref = scipy.linalg.svdvals(a[0].astype(np.float64))[:6]
        assert np.max(np.abs(_np(s[0]) - ref)) / ref[0] < 1e-5
        assert 'warm_started' not in info and info['sketch_columns'] == 14
    with pytest.raises(ValueError):
        pkg.svd(x, 'randomized', rank=6, initial_basis=torch.zeros(300, 5))
    wide = torch.from_numpy(a[0].T.copy()).to(DEV)   # wide inputs are not transposed by these methods
    _, s, _ = pkg.svd(wide, 'subspace', rank=6)
    assert np.max(np.abs(_np(s) - scipy.linalg.svdvals(a[0].astype(np.float64))[:6])) < 1e-5 * float(s[0])
