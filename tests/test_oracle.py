"""Pins the CPU oracle against things other than itself: the reference's golden data (corpus inputs
and the counts/metrics recorded for them), LAPACK through SciPy, closed forms and brute force.

A planted mistake in any oracle function (dropped term, wrong sign or index, transposed operand)
fails at least one of these tests.
"""
import hashlib
import math

import numpy as np
import pytest
import scipy.linalg

from conftest import corpus
from oracle import matrices as X
from oracle import metrics as M
from oracle import restatement as R

RNG = np.random.default_rng(1234)


# --- tournament schedule ----------------------------------------------------------------------
@pytest.mark.parametrize('n', [2, 4, 6, 8, 16, 30])
def test_round_robin_is_a_one_factorisation(n):
    rounds = list(R.round_robin(n))
    assert len(rounds) == n - 1
    seen = set()
    for prs in rounds:
        cols = [c for p in prs for c in p]
        assert sorted(cols) == list(range(n))          # perfect matching
        assert all(p < q for p, q in prs)             # (min, max)
        seen.update(prs)
    assert seen == {(p, q) for p in range(n) for q in range(p + 1, n)}


# --- Householder QR ---------------------------------------------------------------------------
@pytest.mark.parametrize('dtype', [np.float64, np.complex128])
def test_householder_vec_annihilates(dtype):
    x = RNG.standard_normal((5, 7)).astype(dtype)
    if np.iscomplexobj(x):
        x = x + 1j * RNG.standard_normal((5, 7))
    v, tau, beta = R.householder_vec(x)
    h = np.eye(7)[None] - tau[:, None, None] * v[:, :, None] * v.conj()[:, None, :]
    y = np.einsum('bji,bj->bi', h.conj(), x)       # H^H x
    assert np.allclose(y[:, 1:], 0, atol=1e-13)
    assert np.allclose(y[:, 0], beta, atol=1e-13)
    assert np.all(np.sign(beta) == -np.sign(x[:, 0].real))


def test_householder_vec_trivial_column():
    x = np.array([[3.0, 0.0, 0.0], [-2.0, 0.0, 0.0]])
    v, tau, beta = R.householder_vec(x)
    assert np.all(tau == 0) and np.all(v == 0) and np.all(beta == x[:, 0])


@pytest.mark.parametrize('dtype', [np.float64, np.complex128])
@pytest.mark.parametrize('shape', [(9, 5), (6, 6), (40, 12)])
def test_qr_matches_lapack_geqrf(dtype, shape):
    a = RNG.standard_normal((2,) + shape).astype(dtype)
    if np.iscomplexobj(a):
        a = a + 1j * RNG.standard_normal((2,) + shape)
    q, r = R.qr(a)
    for i in range(2):
        qs, rs = scipy.linalg.qr(a[i], mode='economic')   # same xLARFG sign convention
        assert np.allclose(r[i], rs, atol=1e-12)
        assert np.allclose(q[i], qs, atol=1e-12)
    assert np.allclose(q @ r, a, atol=1e-12)


# --- Cholesky and triangular solves -----------------------------------------------------------
def test_cholesky_matches_lapack_and_flags_indefinite():
    g = RNG.standard_normal((3, 8, 8))
    g = g @ g.transpose(0, 2, 1) + 8 * np.eye(8)
    L, fail = R.cholesky(g)
    assert not fail.any()
    for i in range(3):
        assert np.allclose(L[i], scipy.linalg.cholesky(g[i], lower=True), atol=1e-12)
    bad = g.copy()
    bad[1] -= 100 * np.eye(8)
    _, fail = R.cholesky(bad)
    assert fail.tolist() == [False, True, False]


def test_trsm_matches_lapack():
    L = np.tril(RNG.standard_normal((2, 7, 7))) + 5 * np.eye(7)
    B = RNG.standard_normal((2, 7, 3))
    X1 = R.trsm_lower(L, B)
    U = L.transpose(0, 2, 1).copy()
    X2 = R.trsm_upper(U, B)
    for i in range(2):
        assert np.allclose(X1[i], scipy.linalg.solve_triangular(L[i], B[i], lower=True), atol=1e-12)
        assert np.allclose(X2[i], scipy.linalg.solve_triangular(U[i], B[i], lower=False), atol=1e-12)


# --- Jacobi SVD -------------------------------------------------------------------------------
def test_jacobi_2x2_closed_form():
    # sigma of [[a, b], [0, d]] from the closed form of the 2x2 SVD.
    a, b, d = 3.0, 4.0, 1.5
    u, s, v, sweeps = R.jacobi_svd(np.array([[[a, b], [0.0, d]]]))
    fro2, det = a * a + b * b + d * d, abs(a * d)
    s1 = math.sqrt((fro2 + math.sqrt(fro2 ** 2 - 4 * det * det)) / 2)
    assert np.allclose(s[0], [s1, det / s1], rtol=1e-14)
    assert np.allclose((u[0] * s[0]) @ v[0].T, [[a, b], [0, d]], atol=1e-14)


GOLDEN_RUNS = [
    ('square32', 'jacobi'), ('square64', 'jacobi'), ('batch8x16', 'jacobi'), ('tall256x16', 'jacobi'),
    ('tall2048x32', 'jacobi'), ('condition1e8', 'jacobi'), ('condition1e12', 'jacobi'),
    ('rank_deficient', 'jacobi'), ('zero', 'jacobi'), ('clustered_complex', 'jacobi'),
    ('wide_complex', 'jacobi'),
    ('square32', 'tsqr'), ('batch8x16', 'tsqr'), ('tall256x16', 'tsqr'), ('tall2048x32', 'tsqr'),
    ('condition1e8', 'tsqr'), ('condition1e12', 'tsqr'), ('rank_deficient', 'tsqr'), ('zero', 'tsqr'),
    ('square32', 'polar_qdwh'), ('square64', 'polar_qdwh'), ('batch8x16', 'polar_qdwh'),
    ('tall256x16', 'polar_qdwh'), ('tall2048x32', 'polar_qdwh'), ('condition1e8', 'polar_qdwh'),
    ('condition1e12', 'polar_qdwh'), ('rank_deficient', 'polar_qdwh'), ('zero', 'polar_qdwh'),
    ('clustered_complex', 'polar_qdwh'), ('wide_complex', 'polar_qdwh'),
]


@pytest.mark.parametrize('case,method', GOLDEN_RUNS)
def test_counts_and_metrics_match_reference_golden(expected, case, method):
    a = corpus(case)
    assert hashlib.sha256(a.tobytes(order='C')).hexdigest() == expected[case]['sha256']
    u, s, vh, info = R.svd(a, method, return_info=True)
    met = M.validate(a, (u, s, vh), M.reference(a))['worst']
    runs = [r for r in expected[case]['runs'] if r['method'] == method]
    assert runs
    for run in runs:
        for key in ('jacobi_sweeps', 'tsqr_leaves', 'polar_iterations', 'eigen_sweeps'):
            if key in run['diagnostics']:
                assert info[key] == run['diagnostics'][key], (key, run['source'])
        # accuracy within a small factor of what the reference recorded for the same input
        for key in ('spectrum_error_over_sigma1', 'reconstruction_relative_fro',
                    'orthogonality_u_fro_over_sqrt_cols', 'orthogonality_v_fro_over_sqrt_rows'):
            ref = run['metrics'][key]
            floor = 64 * np.finfo(a.real.dtype).eps * max(a.shape[-2:]) ** 0.5
            assert met[key] <= max(3 * ref, floor), (key, met[key], ref, run['source'])


def test_sigma_matches_lapack_on_uneven_gaussians():
    flat, ms, ns, offs = X.workload_uneven_batch(count=12, seed=5)
    for a in X.unpack(flat, ms, ns, offs):
        _, s, _ = R.svd(a, 'jacobi')
        sr = scipy.linalg.svdvals(a.astype(np.float64))
        assert np.max(np.abs(s - sr)) / sr[0] < 2e-5   # reference FP32 envelope is 3e-4


def test_jacobi_nonconvergence_raises():
    a = RNG.standard_normal((12, 9))
    with pytest.raises(RuntimeError):
        R.svd(a, 'jacobi', max_sweeps=1)


# --- Hermitian Jacobi -------------------------------------------------------------------------
@pytest.mark.parametrize('n', [1, 2, 7, 16])
def test_jacobi_eigh_matches_lapack(n):
    h = RNG.standard_normal((2, n, n))
    h = (h + h.transpose(0, 2, 1)) / 2
    w, v, sweeps = R.jacobi_eigh(h)
    for i in range(2):
        assert np.allclose(w[i], scipy.linalg.eigh(h[i], eigvals_only=True), atol=1e-13)
        assert np.allclose(v[i].T @ v[i], np.eye(n), atol=1e-13)
        assert np.allclose((v[i] * w[i]) @ v[i].T, h[i], atol=1e-13)


# --- QDWH -------------------------------------------------------------------------------------
def test_qdwh_weights_reduce_to_halley_at_one():
    aa, bb, cc, nxt = R.qdwh_weights(1.0)
    assert (round(aa, 12), round(bb, 12), round(cc, 12), nxt) == (3.0, 1.0, 3.0, 1.0)


@pytest.mark.parametrize('lower', [1.1920929e-07, 1e-3, 0.3, 0.9])
def test_qdwh_weights_are_the_optimal_rational_map(lower):
    aa, bb, cc, nxt = R.qdwh_weights(lower)
    f = lambda x: x * (aa + bb * x * x) / (1 + cc * x * x)
    xs = np.geomspace(lower, 1.0, 200001)
    fx = f(xs)
    assert abs(f(1.0) - 1.0) < 1e-12                  # fixes sigma = 1
    assert abs(fx.max() - 1.0) < 1e-9                 # maps [l, 1] into [l', 1]: max is 1 ...
    assert abs(fx.min() - nxt) < 1e-9 * max(nxt, 1e-3)  # ... and the minimum is the new bound
    assert np.all(np.diff(fx[: len(xs) // 4]) > 0)     # increasing near the lower end
    x_star = xs[np.argmax(fx[: -10])]                  # equioscillation: interior maximum is 1 too
    assert lower < x_star < 1 and abs(f(x_star) - 1) < 1e-8


def test_qdwh_bound_reaches_one_within_six_steps_from_eps32():
    lower, steps = 1.1920929e-07, 0
    while lower < 1 - 1e-15 and steps < 10:
        lower = R.qdwh_weights(lower)[3]
        steps += 1
    assert steps <= 6


def test_polar_factor_equals_u_vh():
    a = X.make_matrix(dict(m=20, n=9, family='geometric', condition=1e5, dtype='float64'))
    scale = np.abs(a).max()
    stats = {}
    R.polar_qdwh((a / scale)[None], stats, {})
    ur, _, vhr = scipy.linalg.svd(a)
    assert np.allclose(stats['_polar_factor'][0], ur[:, :9] @ vhr, atol=1e-12)


# --- TSQR -------------------------------------------------------------------------------------
def test_tsqr_leaves_and_factor():
    a = RNG.standard_normal((1, 100, 6))
    assert R.tsqr_leaf_bounds(100, 6, 24) == [(0, 24), (24, 48), (48, 72), (72, 100)]
    assert R.tsqr_leaf_bounds(100, 6, 200) is None
    assert R.tsqr_leaf_bounds(20, 8, 4) == [(0, 8), (8, 20)]
    stats = {}
    q, r = R.tsqr(a, 24, stats)
    assert stats['tsqr_leaves'] == 4
    assert np.allclose(q @ r, a, atol=1e-13)
    assert np.allclose(q[0].T @ q[0], np.eye(6), atol=1e-13)
    rs = scipy.linalg.qr(a[0], mode='r')[0][:6]
    assert np.allclose(np.abs(r[0]), np.abs(rs), atol=1e-12)   # unique up to row signs


# --- completion -------------------------------------------------------------------------------
def test_complete_null_columns_orthonormal_and_keeps_live():
    q = scipy.linalg.qr(RNG.standard_normal((10, 3)), mode='economic')[0]
    u = np.concatenate([q, np.zeros((10, 4))], 1)[None]
    keep = np.array([[True, False, True, False, True, False, False]])
    u[0, :, 2], u[0, :, 4] = q[:, 1], q[:, 2]
    u[0, :, 1] = 0
    out = R.complete_null_columns(u.copy(), keep)
    assert np.allclose(out[0].T @ out[0], np.eye(7), atol=1e-14)
    assert np.array_equal(out[0][:, keep[0]], u[0][:, keep[0]])
    full = np.concatenate([np.eye(3), np.zeros((3, 1))], 1)[None]   # live columns span R^3
    with pytest.raises(RuntimeError):
        R.complete_null_columns(full, np.array([[True, True, True, False]]))


def test_zero_matrix_gives_zero_sigma_and_orthonormal_u():
    u, s, vh = R.svd(np.zeros((9, 5)), 'jacobi')
    assert np.all(s == 0)
    assert np.allclose(u.T @ u, np.eye(5), atol=1e-14)


# --- svd wrapper contract ---------------------------------------------------------------------
def test_svd_wrapper_contract():
    a = RNG.standard_normal((12, 9))
    for kw in ({'rank': 0}, {'iterations': -1}, {'oversample': -1}, {'leaf_size': 0}):
        with pytest.raises(ValueError):
            R.svd(a, **kw)
    with pytest.raises(ValueError):
        R.svd(a * np.nan)
    with pytest.raises(TypeError):
        R.svd(a.astype(np.int32))
    with pytest.raises(ValueError):
        R.svd(a, 'nope')
    u, s, vh = R.svd(a.T.copy(), 'jacobi', full_matrices=True)
    assert u.shape == (9, 9) and vh.shape == (12, 12)
    assert np.allclose((u * s) @ vh[:9], a.T, atol=1e-13)
    none, ss, none2 = R.svd(a, 'jacobi', values_only=True)
    assert none is None and none2 is None
    assert np.allclose(ss, scipy.linalg.svdvals(a), atol=1e-13)


# --- metrics (svdbench/metrics.py) ------------------------------------------------------------
def test_metrics_gate_semantics():
    a = np.diag(np.array([4., 3., 2., 1.]))
    w = M.validate(a, (np.eye(4)[:, :2], np.array([4., 3.]), np.eye(4)[:2]), M.reference(a))['worst']
    assert abs(w['residual_over_optimal_tail'] - 1) < 1e-12
    assert abs(w['right_subspace_leakage_fro_over_sqrt_k']) < 1e-12
    a = np.diag([3., 3., 1.])
    q = np.array([[0., -1.], [1., 0.]])
    u = np.eye(3)[:, :2] @ q
    w = M.validate(a, (u, np.array([3., 3.]), u.T), M.reference(a))['worst']
    assert abs(w['reconstruction_relative_fro'] - 1 / np.sqrt(19)) < 1e-12
    w = M.validate(a, (np.eye(3)[:, :1], np.array([3.]), np.eye(3)[:1]), M.reference(a))['worst']
    assert w['right_subspace_leakage_fro_over_sqrt_k'] is None
    a = X.make_matrix(dict(m=24, n=16, family='geometric', condition=1e6, dtype='float32'))
    out = tuple(M.reference(a)[0])
    assert M.validate(a, out, M.reference(a))['worst']['spectrum_error_over_sigma1'] == 0


# --- generators -------------------------------------------------------------------------------
@pytest.mark.parametrize('case,spec', [
    ('square32', dict(m=32, n=32)), ('square64', dict(m=64, n=64)),
    ('batch8x16', dict(m=16, n=16, batch=8)), ('tall256x16', dict(m=256, n=16)),
    ('tall2048x32', dict(m=2048, n=32)), ('zero', dict(m=12, n=8, family='zero')),
])
def test_make_matrix_regenerates_reference_corpus(expected, case, spec):
    a = X.make_matrix({'dtype': 'float32', 'family': 'gaussian', **spec})
    assert hashlib.sha256(a.tobytes(order='C')).hexdigest() == expected[case]['sha256']


def test_workload_generators_shapes_and_spectra():
    ms, ns = X.uneven_shapes(4096)
    assert ms.min() >= 16 and ms.max() <= 96 and np.all(ns >= 8) and np.all(ns <= ms)
    a = X.workload_single_jacobi(m=64, n=32)
    s = scipy.linalg.svdvals(a.astype(np.float64))
    assert np.allclose(s, np.geomspace(1, 1e-6, 32), rtol=1e-3, atol=1e-7)
    a, s_true = X.workload_qdwh(n=64)
    assert np.allclose(scipy.linalg.svdvals(a.astype(np.float64)), s_true, atol=1e-6)
    assert np.sum(np.abs(s_true - 0.5) < 1e-4) >= 6
