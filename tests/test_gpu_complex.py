"""GPU parity of the complex one-sided Jacobi kernel and the complex Householder QR (complex64 /
complex128, LDS resident) against the reference's golden rows for its complex corpus inputs, the
oracle (which restates the complex arithmetic of primitives.py:41-89 and :331-351) and complex
LAPACK.

Bars: sigma error over sigma_1, U/V orthogonality and reconstruction <= 1e-5 (complex64) / 1e-12
(complex128); sweep counts within 1 of the reference's records and of the oracle's (FP64 sums);
metrics within 3x of the reference's recorded ones on the corpus.
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


@pytest.fixture(scope='module')
def pkg():
    return load_package()


def _np(t):
    return None if t is None else t.detach().cpu().numpy()


def _metrics(a, out):
    return M.validate(a, tuple(_np(t) for t in out), M.reference(a))['worst']


@pytest.mark.parametrize('case', ['clustered_complex', 'wide_complex'])
def test_complex_corpus_matches_reference_golden(pkg, expected, case):
    a = corpus(case)
    print(f'complex corpus {case} {a.shape} {a.dtype}', flush=True)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'jacobi', return_info=True)
    assert u.dtype == torch.from_numpy(a).dtype and s.dtype == torch.from_numpy(a.real.copy()).dtype
    met = _metrics(a, (u, s, vh))
    runs = [r for r in expected[case]['runs'] if r['method'] == 'jacobi']
    assert runs
    for run in runs:
        assert abs(info['jacobi_sweeps'] - run['diagnostics']['jacobi_sweeps']) <= 1, run['source']
        floor = 64 * np.finfo(a.real.dtype).eps * max(a.shape[-2:]) ** 0.5
        for key in KEYS:
            assert met[key] <= max(3 * run['metrics'][key], floor), (key, met[key], run['metrics'][key])


@pytest.mark.parametrize('dtype', [np.complex64, np.complex128])
@pytest.mark.parametrize('shape', [(3, 8, 8), (2, 40, 17), (1, 64, 64), (1, 96, 96), (4, 33, 1), (2, 70, 50)])
def test_complex_random_matches_lapack_and_oracle(pkg, dtype, shape):
    rng = np.random.default_rng(shape[1] * 7 + shape[2])
    a = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)
    print(f'complex random {shape} {np.dtype(dtype).name}', flush=True)
    # 96 x 96 complex128 (300 KB) is beyond LDS: the complex single-launch kernel solves it
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'jacobi', return_info=True)
    met = _metrics(a, (u, s, vh))
    bar = 1e-5 if dtype == np.complex64 else 1e-12
    for key in KEYS:
        assert met[key] <= bar, (key, met[key])
    if shape[1] <= 40:
        _, rs, _, rinfo = R.svd(a, 'jacobi', return_info=True)
        lap = np.stack([np.linalg.svd(x.astype(np.complex128), compute_uv=False) for x in a])
        oracle_err = np.max(np.abs(rs - lap)) / lap.max()
        assert np.max(np.abs(_np(s) - rs)) / rs.max() <= bar + oracle_err
        assert abs(info['jacobi_sweeps'] - rinfo['jacobi_sweeps']) <= 1, (info, rinfo)


def test_complex_rank_deficient_zero_wide_and_values_only(pkg):
    rng = np.random.default_rng(21)
    f1 = rng.standard_normal((30, 3)) + 1j * rng.standard_normal((30, 3))
    f2 = rng.standard_normal((3, 12)) + 1j * rng.standard_normal((3, 12))
    low = (f1 @ f2).astype(np.complex128)
    u, s, vh = pkg.svd(torch.from_numpy(low).to(DEV), 'jacobi')
    sv, uu = _np(s), _np(u)
    assert np.count_nonzero(sv) == 3
    assert np.max(np.abs(uu.conj().T @ uu - np.eye(12))) < 1e-12          # completed null columns
    assert np.max(np.abs((uu * sv) @ _np(vh) - low)) < 1e-12 * np.abs(low).max()
    z = torch.zeros(2, 10, 6, dtype=torch.complex64, device=DEV)
    u, s, vh, info = pkg.svd(z, 'jacobi', return_info=True)
    assert not s.any() and info['jacobi_sweeps'] == 1
    uu = _np(u[0])
    assert np.max(np.abs(uu.conj().T @ uu - np.eye(6))) < 1e-6
    w = (rng.standard_normal((9, 25)) + 1j * rng.standard_normal((9, 25))).astype(np.complex64)
    u, s, vh = pkg.svd(torch.from_numpy(w).to(DEV), 'jacobi')
    assert np.max(np.abs((_np(u) * _np(s)) @ _np(vh) - w)) < 1e-5 * np.abs(w).max()
    _, sv, _ = pkg.svd(torch.from_numpy(w).to(DEV), 'jacobi', values_only=True)
    assert np.max(np.abs(_np(sv) - np.linalg.svd(w.astype(np.complex128), compute_uv=False))) < 1e-5 * float(sv[0])


@pytest.mark.parametrize('dtype', [np.complex64, np.complex128])
@pytest.mark.parametrize('shape', [(3, 8, 8), (2, 40, 17), (1, 64, 64), (4, 33, 1), (2, 70, 50), (1, 96, 48)])
def test_complex_qr_factors_match_oracle_and_lapack(dtype, shape):
    """Q and R of the LDS complex QR against the oracle's restatement of primitives.qr (LAPACK's
    zgeqrf convention, so also scipy's economic QR): entrywise to 2e-5 (complex64) / 1e-12
    (complex128) relative to max |A|, Q^H Q = I and QR = A to the same bars."""
    import scipy.linalg
    from importlib import import_module
    P = import_module('pavel-s_svd_amd.standalone.primitives')
    rng = np.random.default_rng(shape[1] * 11 + shape[2])
    a = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)
    print(f'complex qr {shape} {np.dtype(dtype).name}', flush=True)
    q, r = P.qr(torch.from_numpy(a).to(DEV))
    q, r = _np(q), _np(r)
    tol = 2e-5 if dtype == np.complex64 else 1e-12
    qo, ro = R.qr(a.astype(np.complex128))
    scale = np.abs(a).max()
    assert np.abs(r - ro).max() <= tol * scale
    assert np.abs(q - qo).max() <= tol
    for i in range(shape[0]):
        qs, rs = scipy.linalg.qr(a[i].astype(np.complex128), mode='economic')
        assert np.abs(r[i] - rs).max() <= tol * scale
        assert np.abs(np.diag(r[i]).imag).max() == 0.0     # beta is real
        assert np.abs(q[i].conj().T @ q[i] - np.eye(shape[2])).max() <= tol
        assert np.abs(q[i] @ r[i] - a[i]).max() <= tol * scale


def test_complex_qr_trivial_reflector():
    """A column that is already reduced (real alpha, zeros below) takes the trivial reflector:
    tau = 0, beta = alpha, so R keeps alpha's sign and Q keeps e_j (primitives.py:49-57)."""
    from importlib import import_module
    P = import_module('pavel-s_svd_amd.standalone.primitives')
    a = np.zeros((1, 5, 3), dtype=np.complex128)
    a[0, 0, 0] = -2.0
    a[0, :, 1] = [1 + 1j, 3.0, 0.0, 1j, 2.0]
    a[0, :, 2] = [0.5, -1j, 2.0, 0.0, 1.0]
    q, r = P.qr(torch.from_numpy(a).to(DEV))
    qo, ro = R.qr(a)
    assert _np(r)[0, 0, 0] == -2.0 and np.allclose(_np(q)[0, :, 0], [1, 0, 0, 0, 0])
    assert np.abs(_np(r) - ro).max() < 1e-13 and np.abs(_np(q) - qo).max() < 1e-13


@pytest.mark.parametrize('case', ['clustered_complex', 'wide_complex'])
def test_complex_qr_solver_matches_oracle_and_lapack(pkg, case):
    """svd(a, 'qr') on the complex corpus inputs: QR of the (transposed, for wide) input, one-sided
    Jacobi of R, U = Q U_R (solvers.py:245-248).  sigma against the oracle's qr solver and LAPACK,
    metrics to the complex bars."""
    a = corpus(case)
    print(f'complex qr solver {case} {a.shape} {a.dtype}', flush=True)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'qr', return_info=True)
    met = _metrics(a, (u, s, vh))
    bar = 1e-5 if a.dtype == np.complex64 else 1e-12
    for key in KEYS:
        assert met[key] <= bar, (key, met[key])
    _, so, _, oinfo = R.svd(a, 'qr', return_info=True)
    assert np.abs(_np(s) - so).max() <= bar * so.max()
    assert abs(info['jacobi_sweeps'] - oinfo['jacobi_sweeps']) <= 1


def test_complex_qr_solver_zero_rank_deficient_and_scaled(pkg):
    """Zero input (every reflector trivial, sigma = 0, completed basis), a rank-2 input (columns
    past the rank cut and completed) and a 1e30-scaled input (max-modulus scaling before the sums)
    through svd(a, 'qr'), against LAPACK."""
    rng = np.random.default_rng(17)
    z = np.zeros((1, 12, 5), dtype=np.complex128)
    u, s, vh = pkg.svd(torch.from_numpy(z).to(DEV), 'qr')
    assert float(_np(s).max()) == 0.0
    uu = _np(u)[0]
    assert np.abs(uu.conj().T @ uu - np.eye(5)).max() < 1e-12
    f = rng.standard_normal((30, 2)) + 1j * rng.standard_normal((30, 2))
    g = rng.standard_normal((2, 9)) + 1j * rng.standard_normal((2, 9))
    lr = (f @ g)[None].astype(np.complex64)
    u, s, vh = pkg.svd(torch.from_numpy(lr).to(DEV), 'qr')
    ref = np.linalg.svd(lr[0].astype(np.complex128), compute_uv=False)
    assert np.abs(_np(s)[0] - ref).max() <= 1e-5 * ref[0]
    assert np.abs(_np(u)[0].conj().T @ _np(u)[0] - np.eye(9)).max() < 1e-5
    big = (1e30 * (rng.standard_normal((1, 20, 7)) + 1j * rng.standard_normal((1, 20, 7)))).astype(np.complex128)
    u, s, vh = pkg.svd(torch.from_numpy(big).to(DEV), 'qr')
    ref = np.linalg.svd(big[0], compute_uv=False)
    assert np.abs(_np(s)[0] - ref).max() <= 1e-12 * ref[0]


def test_complex_qr_batch_beyond_one_launch_of_the_product(pkg):
    """70 000 complex64 6 x 3 matrices through svd(a, 'qr'): the product U = Q U_R is launched in
    chunks of at most 65 535 matrices; every matrix's sigma against LAPACK."""
    rng = np.random.default_rng(23)
    a = (rng.standard_normal((70000, 6, 3)) + 1j * rng.standard_normal((70000, 6, 3))).astype(np.complex64)
    u, s, vh = pkg.svd(torch.from_numpy(a).to(DEV), 'qr')
    ref = np.linalg.svd(a.astype(np.complex128), compute_uv=False)
    assert np.abs(_np(s) - ref).max() <= 1e-5 * ref.max()
    rec = _np(u) @ (_np(s)[:, :, None] * _np(vh))
    assert np.abs(rec - a).max() <= 1e-5 * np.abs(a).max()


@pytest.mark.parametrize('dtype', [np.complex64, np.complex128])
def test_complex_qr_solver_random_batch(pkg, dtype):
    rng = np.random.default_rng(5)
    a = (rng.standard_normal((3, 60, 24)) + 1j * rng.standard_normal((3, 60, 24))).astype(dtype)
    u, s, vh = pkg.svd(torch.from_numpy(a).to(DEV), 'qr')
    uj, sj, vhj = pkg.svd(torch.from_numpy(a).to(DEV), 'jacobi')
    bar = 1e-5 if dtype == np.complex64 else 1e-12
    for i in range(3):
        met = _metrics(a[i], (u[i], s[i], vh[i]))
        for key in KEYS:
            assert met[key] <= bar, (key, met[key])
    assert np.abs(_np(s) - _np(sj)).max() <= bar * float(_np(sj).max())


@pytest.mark.parametrize('dtype,shape,rows', [(np.complex64, (2, 200, 16), 48), (np.complex128, (1, 130, 24), 40),
                                               (np.complex64, (1, 97, 8), 32)])
def test_complex_tsqr_matches_oracle(pkg, dtype, shape, rows):
    """svd(a, 'tsqr', tile_rows=rows) on complex inputs with several leaves (a short last block
    joins the previous leaf) and a binary tree of [R1; R2] merges (solvers.py:71-96): leaf count,
    sweeps and sigma against the oracle's TSQR, metrics to the complex bars."""
    rng = np.random.default_rng(shape[1] + rows)
    a = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)
    print(f'complex tsqr {shape} rows {rows} {np.dtype(dtype).name}', flush=True)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'tsqr', tile_rows=rows, return_info=True)
    _, so, _, oinfo = R.svd(a, 'tsqr', tile_rows=rows, return_info=True)
    assert info['tsqr_leaves'] == oinfo['tsqr_leaves'] > 1
    assert abs(info['jacobi_sweeps'] - oinfo['jacobi_sweeps']) <= 1
    bar = 1e-5 if dtype == np.complex64 else 1e-12
    assert np.abs(_np(s) - so).max() <= bar * so.max()
    for i in range(shape[0]):
        met = _metrics(a[i], (u[i], s[i], vh[i]))
        for key in KEYS:
            assert met[key] <= bar, (key, met[key])


@pytest.mark.parametrize('case', ['clustered_complex', 'wide_complex'])
def test_complex_qdwh_matches_reference_golden(pkg, expected, case):
    """svd(a, 'polar_qdwh') on the complex corpus: the reference's branch rule (QR form while c > 100,
    then the Cholesky form) gives its recorded QDWH iteration count exactly and the two-sided
    eigensolver of H its eigen sweeps within one; sigma against the oracle's solver, metrics within 3x
    of the reference's recorded ones (or the data type's floor)."""
    a = corpus(case)
    print(f'complex qdwh {case} {a.shape} {a.dtype}', flush=True)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'polar_qdwh', return_info=True)
    met = _metrics(a, (u, s, vh))
    runs = [r for r in expected[case]['runs'] if r['method'] == 'polar_qdwh']
    assert runs
    for run in runs:
        assert info['polar_iterations'] == run['diagnostics']['polar_iterations'], run['source']
        assert abs(info['eigen_sweeps'] - run['diagnostics']['eigen_sweeps']) <= 1, run['source']
        floor = 64 * np.finfo(a.real.dtype).eps * max(a.shape[-2:]) ** 0.5
        for key in KEYS:
            assert met[key] <= max(3 * run['metrics'][key], floor), (key, met[key], run['metrics'][key])
    _, so, _, oinfo = R.svd(a, 'polar_qdwh', return_info=True)
    bar = 1e-5 if a.dtype == np.complex64 else 1e-12
    assert np.abs(_np(s) - so).max() <= bar * so.max()
    assert info['polar_iterations'] == oinfo['polar_iterations']
    assert abs(info['eigen_sweeps'] - oinfo['eigen_sweeps']) <= 1
    assert 1 <= info['polar_qr_iterations'] < info['polar_iterations']


@pytest.mark.parametrize('dtype,shape', [(np.complex64, (3, 40, 17)), (np.complex128, (2, 48, 48)),
                                         (np.complex64, (1, 30, 62)), (np.complex128, (1, 64, 8))])
def test_complex_qdwh_random_matches_lapack(pkg, dtype, shape):
    rng = np.random.default_rng(shape[1] * 3 + shape[2])
    a = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)
    print(f'complex qdwh random {shape} {np.dtype(dtype).name}', flush=True)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'polar_qdwh', return_info=True)
    bar = 1e-5 if dtype == np.complex64 else 1e-12
    for i in range(shape[0]):
        met = _metrics(a[i], (u[i], s[i], vh[i]))
        for key in KEYS:
            assert met[key] <= bar, (key, met[key])
    real = np.float32 if dtype == np.complex64 else np.float64
    assert info['polar_fixed_point_error'] <= 100 * np.finfo(real).eps


@pytest.mark.parametrize('case', ['clustered_complex', 'wide_complex'])
def test_complex_gram_matches_reference_golden(pkg, expected, case):
    """svd(a, 'gram_jacobi') on the complex corpus: the number of discarded values equals the
    reference's record; the kept sigma against the oracle's Gram solver; metrics within 3x of the
    reference's recorded ones (or the data type's floor)."""
    a = corpus(case)
    print(f'complex gram {case} {a.shape} {a.dtype}', flush=True)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'gram_jacobi', return_info=True)
    met = _metrics(a, (u, s, vh))
    runs = [r for r in expected[case]['runs'] if r['method'] == 'gram_jacobi']
    assert runs
    for run in runs:
        assert info['gram_discarded_values'] == run['diagnostics']['gram_discarded_values'], run['source']
        assert abs(info['eigen_sweeps'] - run['diagnostics']['eigen_sweeps']) <= 1, run['source']
        floor = 64 * np.finfo(a.real.dtype).eps * max(a.shape[-2:]) ** 0.5
        for key in ('orthogonality_u_fro_over_sqrt_cols', 'orthogonality_v_fro_over_sqrt_rows',
                    'reconstruction_relative_fro', 'spectrum_error_over_sigma1'):
            assert met[key] <= max(3 * run['metrics'][key], floor), (key, met[key], run['metrics'][key])
    _, so, _, oinfo = R.svd(a, 'gram_jacobi', return_info=True)
    kept = so > 0
    bar = 1e-5 if a.dtype == np.complex64 else 1e-12
    assert np.abs(_np(s)[..., :so.shape[-1]][kept] - so[kept]).max() <= 10 * bar * so.max()


@pytest.mark.parametrize('case', ['clustered_complex', 'wide_complex'])
def test_complex_block_jacobi_matches_reference_golden(pkg, expected, case):
    """svd(a, 'block_jacobi') on the complex corpus: block sweeps within one of the reference's
    record and of the oracle's; sigma against the oracle's; metrics within 3x of the reference's
    recorded ones (or the data type's floor)."""
    a = corpus(case)
    print(f'complex block_jacobi {case} {a.shape} {a.dtype}', flush=True)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'block_jacobi', return_info=True)
    met = _metrics(a, (u, s, vh))
    runs = [r for r in expected[case]['runs'] if r['method'] == 'block_jacobi']
    assert runs
    for run in runs:
        assert abs(info['block_sweeps'] - run['diagnostics']['block_sweeps']) <= 1, run['source']
        floor = 64 * np.finfo(a.real.dtype).eps * max(a.shape[-2:]) ** 0.5
        for key in KEYS:
            assert met[key] <= max(3 * run['metrics'][key], floor), (key, met[key], run['metrics'][key])
    _, so, _, oinfo = R.svd(a, 'block_jacobi', return_info=True)
    assert abs(info['block_sweeps'] - oinfo['block_sweeps']) <= 1
    bar = 1e-5 if a.dtype == np.complex64 else 1e-12
    assert np.abs(_np(s) - so).max() <= 10 * bar * so.max()


@pytest.mark.parametrize('dtype', [np.complex64, np.complex128])
@pytest.mark.parametrize('shape,width', [((3, 40, 24), 4), ((2, 33, 21), 4), ((1, 64, 48), 8), ((2, 20, 20), 3),
                                         ((5, 16, 9), 2)])
def test_complex_block_jacobi_random_matches_lapack(pkg, dtype, shape, width):
    """Complex block_jacobi over even and odd group counts (the empty group of solvers.py:37-38),
    short last groups and several matrices per launch: sigma against LAPACK, U and V unitary,
    U diag(s) Vh = A."""
    rng = np.random.default_rng(sum(shape) + width)
    a = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)
    print(f'complex block_jacobi {shape} {width} {np.dtype(dtype).name}', flush=True)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'block_jacobi', return_info=True, block_size=width)
    u, s, vh = _np(u), _np(s), _np(vh)
    bar = 1e-5 if dtype == np.complex64 else 1e-12
    ref = np.linalg.svd(a.astype(np.complex128), compute_uv=False)
    assert np.abs(s - ref).max() <= 10 * bar * ref.max()
    n = shape[-1]
    assert np.abs(np.conj(np.swapaxes(u, -1, -2)) @ u - np.eye(n)).max() <= 10 * bar
    assert np.abs(vh @ np.conj(np.swapaxes(vh, -1, -2)) - np.eye(n)).max() <= 10 * bar
    assert np.abs((u * s[:, None, :]) @ vh - a).max() <= 10 * bar * np.abs(a).max()
    assert 1 <= info['block_sweeps'] <= 30


@pytest.mark.parametrize('method', ['randomized', 'subspace', 'block_krylov'])
@pytest.mark.parametrize('case,opts', [('clustered_complex', {}), ('wide_complex', {}),
                                       ('clustered_complex', {'rank': 6, 'oversample': 4, 'iterations': 2, 'seed': 7})])
def test_complex_range_finders_match_oracle_on_the_same_sketch(pkg, expected, method, case, opts):
    """Complex randomized / subspace / block_krylov (solvers.py:185-221) with the oracle's complex
    sketch passed as initial_basis: column counts equal the oracle's and the reference's CPU rows,
    sigma against the oracle's, metrics within 3x of the oracle's."""
    a = corpus(case)
    print(f'complex {method} {case} {opts}', flush=True)
    a3 = a[None] if a.ndim == 2 else a
    k = opts.get('rank', min(a.shape[-2:]))
    ell = min(k + opts.get('oversample', 8), min(a.shape[-2:]))
    omega = R.sketch(a3, ell, opts.get('seed', 20261011))
    ru, rs, rvh, rinfo = R.svd(a, method, return_info=True, **opts)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), method, return_info=True,
                             initial_basis=torch.from_numpy(omega[0] if a.ndim == 2 else omega), **opts)
    for key in ('sketch_columns', 'range_columns', 'power_iterations'):
        assert info[key] == rinfo[key], key
    assert abs(info['jacobi_sweeps'] - rinfo['jacobi_sweeps']) <= 1
    bar = 1e-5 if a.dtype == np.complex64 else 1e-12
    top = max(float(np.max(rs)), 1e-30)
    assert np.max(np.abs(_np(s) - rs)) / top <= bar
    met = _metrics(a, (u, s, vh))
    ref = _metrics(a, (torch.from_numpy(ru), torch.from_numpy(rs), torch.from_numpy(rvh)))
    for key in KEYS:
        assert met[key] <= max(3 * ref[key], 10 * bar), (key, met[key], ref[key])
    runs = [r for r in expected[case]['runs'] if r['method'] == method and r['device'] == 'cpu'
            and r['options'] == opts]
    for run in runs:
        for key in ('sketch_columns', 'range_columns', 'power_iterations'):
            assert info[key] == run['diagnostics'][key]


def test_complex_unsupported_paths(pkg):
    x = torch.zeros(1, 20, 10, dtype=torch.complex64, device=DEV) + 1
    with pytest.raises(pkg.Unsupported):
        pkg.svd(x, 'lanczos')


@pytest.mark.parametrize('dtype', [np.complex64, np.complex128])
def test_complex_batched_jacobi_beyond_lds(pkg, dtype):
    """A batch of complex matrices beyond LDS (the batched complex kernel on workspace regions, the
    reference's rotation order): sigma, U, V against complex LAPACK, sweep counts within one of the
    oracle's restatement; and complex block_jacobi whose 1200 x 16 group pairs exceed LDS against
    LAPACK."""
    from importlib import import_module
    P = import_module('pavel-s_svd_amd.standalone.primitives')
    tdt = torch.complex64 if dtype == np.complex64 else torch.complex128
    rng = np.random.default_rng(9)
    a = (rng.standard_normal((3, 260, 120)) + 1j * rng.standard_normal((3, 260, 120))).astype(dtype)
    assert not P.lds_fits(tdt, 260, 120, True)
    print(f'complex batched jacobi beyond LDS 3x260x120 {np.dtype(dtype).name}', flush=True)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'jacobi', return_info=True)
    bar = 1e-5 if dtype == np.complex64 else 1e-12
    for i in range(3):
        met = _metrics(a[i], (u[i], s[i], vh[i]))
        for key in KEYS:
            assert met[key] <= bar, (i, key, met[key])
    _, _, _, oinfo = R.svd(a[0], 'jacobi', return_info=True)
    assert abs(info['jacobi_sweeps'] - oinfo['jacobi_sweeps']) <= 1
    b = (rng.standard_normal((1, 1200, 40)) + 1j * rng.standard_normal((1, 1200, 40))).astype(dtype)
    assert not P.lds_fits(tdt, 1200, 16, True)
    print(f'complex block_jacobi with group pairs beyond LDS 1200x40 bs 8 {np.dtype(dtype).name}', flush=True)
    u, s, vh = pkg.svd(torch.from_numpy(b).to(DEV), 'block_jacobi', block_size=8)
    met = _metrics(b[0], (u[0], s[0], vh[0]))
    for key in KEYS:
        assert met[key] <= bar * 10, (key, met[key])


@pytest.mark.parametrize('method', ['jacobi', 'qr', 'tsqr', 'polar_qdwh', 'block_jacobi'])
@pytest.mark.parametrize('dtype', [np.complex64, np.complex128])
def test_complex_full_matrices_and_rank_deficient(pkg, method, dtype):
    """Rectangular full_matrices (complete_basis on complex bases) and a rank-2 input (columns past
    the cut completed, primitives.py:430-467) through every complex solver: U and V unitary, sigma
    against LAPACK, U diag(s) Vh = A."""
    rng = np.random.default_rng(41)
    f = rng.standard_normal((1, 24, 2)) + 1j * rng.standard_normal((1, 24, 2))
    g = rng.standard_normal((1, 2, 9)) + 1j * rng.standard_normal((1, 2, 9))
    a = (f @ g).astype(dtype)
    bar = 1e-5 if dtype == np.complex64 else 1e-12
    u, s, vh = pkg.svd(torch.from_numpy(a).to(DEV), method, full_matrices=True)
    u, s, vh = _np(u)[0], _np(s)[0], _np(vh)[0]
    assert u.shape == (24, 24) and vh.shape == (9, 9)
    ref = np.linalg.svd(a[0].astype(np.complex128), compute_uv=False)
    assert np.abs(s - ref).max() <= bar * ref[0] * 10
    assert np.abs(u.conj().T @ u - np.eye(24)).max() <= 10 * bar
    assert np.abs(vh @ vh.conj().T - np.eye(9)).max() <= 10 * bar
    assert np.abs((u[:, :9] * s) @ vh - a[0]).max() <= 10 * bar * np.abs(a).max()


@pytest.mark.parametrize('dtype', [np.complex64, np.complex128])
@pytest.mark.parametrize('b,n', [(3, 1), (2, 2), (4, 7), (2, 16), (1, 33), (2, 64), (2, 120)])
def test_complex_eigh_matches_lapack_and_oracle(dtype, b, n):
    """The complex two-sided Jacobi (primitives.py:370-427, complex branch): eigenvalues against
    LAPACK's zheevd, V unitary and V diag(w) V^H = H, sweep counts within one of the oracle's (the
    reference's rotation order; FP64 sums change at most the last sweep).  n = 120 is beyond the LDS
    budget (both dtypes): the same kernel on a workspace region."""
    from importlib import import_module
    P = import_module('pavel-s_svd_amd.standalone.primitives')
    if n == 120:
        assert not P.ceigh_fits(torch.complex64 if dtype == np.complex64 else torch.complex128, n)
    rng = np.random.default_rng(7 * n + b)
    h = rng.standard_normal((b, n, n)) + 1j * rng.standard_normal((b, n, n))
    h = ((h + np.conj(np.swapaxes(h, 1, 2))) / 2).astype(dtype)
    print(f'complex eigh {b}x{n} {np.dtype(dtype).name}', flush=True)
    w, v, sweeps = P.jacobi_eigh(torch.from_numpy(h).to(DEV), max_sweeps=60)
    w, v = w.cpu().numpy().astype(np.float64), v.cpu().numpy().astype(np.complex128)
    bar = 2e-6 if dtype == np.complex64 else 1e-12
    for i in range(b):
        ref = scipy.linalg.eigh(h[i].astype(np.complex128), eigvals_only=True)
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.all(np.diff(w[i]) >= 0)
        assert np.max(np.abs(w[i] - ref)) / scale < bar * np.sqrt(n)
        assert np.max(np.abs(np.conj(v[i].T) @ v[i] - np.eye(n))) < bar * n
        assert np.max(np.abs((v[i] * w[i]) @ np.conj(v[i].T) - h[i])) / scale < bar * n
    _, _, rsw = R.jacobi_eigh(h)
    assert abs(sweeps - rsw) <= 1


@pytest.mark.parametrize('dtype', [np.complex64, np.complex128])
@pytest.mark.parametrize('b,m,n', [(3, 1, 1), (2, 20, 9), (2, 64, 48), (1, 96, 64), (2, 300, 130)])
def test_complex_cholesky_and_solve(dtype, b, m, n):
    """Complex Cholesky against LAPACK's zpotrf (lower), the reference's failure mask on an
    indefinite matrix, and X (L L^H)^-1 against a direct solve (solvers.py:159-160); 300 x 130 is
    beyond the LDS budget of both kernels (workspace regions)."""
    from importlib import import_module
    P = import_module('pavel-s_svd_amd.standalone.primitives')
    if m == 300:
        assert not P.csolve_fits(torch.complex64 if dtype == np.complex64 else torch.complex128, m, n)
    rng = np.random.default_rng(m + n)
    x = (rng.standard_normal((b, n, n)) + 1j * rng.standard_normal((b, n, n)))
    z = (np.eye(n) + 0.3 * np.conj(np.swapaxes(x, 1, 2)) @ x).astype(dtype)
    L, fail = P.cholesky(torch.from_numpy(z).to(DEV))
    assert not bool(fail.any())
    L = L.cpu().numpy().astype(np.complex128)
    bar = 3e-5 if dtype == np.complex64 else 1e-12
    for i in range(b):
        ref = scipy.linalg.cholesky(z[i].astype(np.complex128), lower=True)
        assert np.max(np.abs(L[i] - ref)) <= bar * np.abs(ref).max()
    rhs = (rng.standard_normal((b, m, n)) + 1j * rng.standard_normal((b, m, n))).astype(dtype)
    xt = torch.from_numpy(rhs).to(DEV)
    P.csolve_chol_(xt, torch.from_numpy(L.astype(dtype)).to(DEV))
    got = xt.cpu().numpy().astype(np.complex128)
    for i in range(b):
        ref = np.linalg.solve(z[i].astype(np.complex128).T, rhs[i].astype(np.complex128).T).T   # X Z^-1
        assert np.max(np.abs(got[i] - ref)) <= 50 * bar * np.abs(ref).max()
    bad = z.copy()
    bad[0, n - 1, n - 1] = -1.0
    _, fail = P.cholesky(torch.from_numpy(bad).to(DEV))
    assert bool(fail[0]) and not bool(fail[1:].any())


@pytest.mark.parametrize('dtype,shape,rank', [(np.complex64, (300, 130), None), (np.complex128, (200, 100), None),
                                              (np.complex64, (100, 300), None), (np.complex128, (180, 120), 40)])
def test_complex_jacobi_beyond_lds(pkg, dtype, shape, rank):
    """Complex one-sided Jacobi on matrices larger than LDS (the complex single-launch kernel, the
    reference's rotation order): sigma, U, V against complex LAPACK; sigma against the oracle's
    restatement of the reference's solver within the oracle's own error (its sums run in the input
    dtype) and sweep counts within one of it for full-rank inputs (for a rank-deficient one the
    columns in the null space stop rotating at a floor taken one round apart, so only the results are
    compared); rank-deficient inputs completed."""
    rng = np.random.default_rng(shape[0] + shape[1])
    m, n = shape
    if rank is None:
        a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    else:
        a = (rng.standard_normal((m, rank)) + 1j * rng.standard_normal((m, rank))) @ \
            (rng.standard_normal((rank, n)) + 1j * rng.standard_normal((rank, n)))
    a = a.astype(dtype)
    from importlib import import_module
    P = import_module('pavel-s_svd_amd.standalone.primitives')
    tdt = torch.complex64 if dtype == np.complex64 else torch.complex128
    assert not P.lds_fits(tdt, max(shape), min(shape), True)
    print(f'complex jacobi beyond LDS {shape} {np.dtype(dtype).name} rank {rank}', flush=True)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'jacobi', return_info=True)
    met = _metrics(a, (u, s, vh))
    bar = 1e-5 if dtype == np.complex64 else 1e-12
    for key in KEYS:
        assert met[key] <= bar, (key, met[key])
    _, so, _, oinfo = R.svd(a, 'jacobi', return_info=True)
    lap = np.linalg.svd(a.astype(np.complex128), compute_uv=False)
    oracle_err = np.max(np.abs(so - lap)) / lap.max()
    assert np.abs(_np(s) - so).max() <= (bar + oracle_err) * so.max()
    if rank is None:
        assert abs(info['jacobi_sweeps'] - oinfo['jacobi_sweeps']) <= 1


@pytest.mark.parametrize('method,opts', [('qr', {}), ('tsqr', dict(tile_rows=150)), ('polar_qdwh', {}),
                                         ('gram_jacobi', {})])
@pytest.mark.parametrize('dtype', [np.complex64, np.complex128])
def test_complex_solvers_beyond_lds(pkg, method, opts, dtype):
    """The complex solvers on a 400 x 140 matrix, whose QRs, QDWH stack, Cholesky factor, Gram and
    eigen problems exceed LDS (the kernels run on workspace regions, the Jacobi of R on the complex
    single-launch kernel): sigma, U, V against complex LAPACK at the complex bars (gram_jacobi: the kept
    values above sqrt(eps n) sigma_1 at the Gram method's accuracy); QDWH iteration counts equal to the
    oracle's restatement of the reference's solver on the same input."""
    rng = np.random.default_rng(140)
    m, n = 400, 140
    a = (rng.standard_normal((1, m, n)) + 1j * rng.standard_normal((1, m, n))).astype(dtype)
    from importlib import import_module
    P = import_module('pavel-s_svd_amd.standalone.primitives')
    tdt = torch.complex64 if dtype == np.complex64 else torch.complex128
    assert not P.cqr_fits(tdt, m, n) and not P.ceigh_fits(tdt, n)
    print(f'complex {method} beyond LDS {m}x{n} {np.dtype(dtype).name}', flush=True)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), method, return_info=True, **opts)
    lap = np.linalg.svd(a[0].astype(np.complex128), compute_uv=False)
    bar = 1e-5 if dtype == np.complex64 else 1e-12
    if method == 'gram_jacobi':
        keep = _np(s[0]) > 0
        gbar = 2e-4 if dtype == np.complex64 else 1e-8
        assert np.abs(_np(s[0])[keep] - lap[keep]).max() <= gbar * lap[0]
        return
    met = _metrics(a[0], (u[0], s[0], vh[0]))
    for key in KEYS:
        assert met[key] <= bar, (key, met[key])
    if method == 'polar_qdwh':
        _, _, _, oinfo = R.svd(a[0], 'polar_qdwh', return_info=True)
        assert info['polar_iterations'] == oinfo['polar_iterations']
