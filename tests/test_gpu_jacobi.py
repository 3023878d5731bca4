"""GPU parity of the LDS-resident batched Jacobi kernel against the CPU oracle and FP64 LAPACK.

Bars (stated per test): sigma error over sigma_1 and U/V orthogonality against the FP64 gesvd
oracle, sweep counts within the reference's recorded counts (the FP64 convergence test may stop a
sweep earlier or later than the reference's FP32 test), and the reference's exceptions.
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


@pytest.fixture(scope='module')
def pkg():
    return load_package()


def _np(t):
    return None if t is None else t.detach().cpu().numpy()


def _metrics(a, out):
    return M.validate(a, tuple(_np(t) for t in out), M.reference(a))['worst']


REAL_CASES = ['square32', 'square64', 'batch8x16', 'tall256x16', 'condition1e8', 'condition1e12',
              'rank_deficient', 'zero']


@pytest.mark.parametrize('case', REAL_CASES)
def test_corpus_against_oracle_and_lapack(pkg, expected, case):
    a = corpus(case)
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'jacobi', return_info=True)
    ref_u, ref_s, ref_vh, ref_info = R.svd(a, 'jacobi', return_info=True)
    met = _metrics(a, (u, s, vh))
    eps = np.finfo(a.dtype).eps
    # sigma: within the oracle's own FP64-LAPACK distance plus a few eps (tolerance 8 eps * sqrt n)
    top = max(float(np.abs(ref_s).max()), np.finfo(float).tiny)
    assert np.max(np.abs(_np(s) - ref_s)) / top <= max(8 * eps * np.sqrt(a.shape[-1]), 1e-300) + \
        M.validate(a, (ref_u, ref_s, ref_vh), M.reference(a))['worst']['spectrum_error_over_sigma1']
    bar = 1e-5 if a.dtype == np.float32 else 1e-13
    for key in ('spectrum_error_over_sigma1', 'orthogonality_u_fro_over_sqrt_cols',
                'orthogonality_v_fro_over_sqrt_rows', 'reconstruction_relative_fro'):
        if met[key] is not None:
            assert met[key] <= bar, (key, met[key])
    golden = [r['diagnostics']['jacobi_sweeps'] for r in expected[case]['runs'] if r['method'] == 'jacobi']
    assert abs(info['jacobi_sweeps'] - golden[0]) <= 2, (info['jacobi_sweeps'], golden)
    assert info['jacobi_sweeps'] >= 1


def test_wide_matrix_is_solved_transposed(pkg):
    a = X.make_matrix(dict(m=9, n=23, family='gaussian', dtype='float64', seed=3))
    u, s, vh = pkg.svd(torch.from_numpy(a).to(DEV), 'jacobi')
    assert tuple(u.shape) == (9, 9) and tuple(vh.shape) == (9, 23)
    met = _metrics(a, (u, s, vh))
    assert met['reconstruction_relative_fro'] < 1e-14 and met['orthogonality_v_fro_over_sqrt_rows'] < 1e-14


@pytest.mark.parametrize('m,n', [(1, 1), (8, 1), (7, 7), (33, 17), (96, 96), (120, 40)])
def test_shapes_odd_and_degenerate(pkg, m, n):
    a = X.make_matrix(dict(m=m, n=n, family='gaussian', dtype='float32', seed=m * 100 + n))
    u, s, vh = pkg.svd(torch.from_numpy(a).to(DEV), 'jacobi')
    met = _metrics(a, (u, s, vh))
    assert met['spectrum_error_over_sigma1'] < 1e-5
    assert met['orthogonality_u_fro_over_sqrt_cols'] < 1e-5
    assert met['orthogonality_v_fro_over_sqrt_rows'] < 1e-5


# Launch shapes across the fixed-slot loop (L <= 16 lanes and ceil(m / L) <= 12 row slots per
# lane for FP32, <= 4 for FP64) and the general loop (L = 32, or more slots): row counts at, one
# below and one above a multiple of 16 (the padded column stride), n odd and n = m odd (np > m).
@pytest.mark.parametrize('dtype', ['float32', 'float64'])
@pytest.mark.parametrize('m,n', [(16, 16), (17, 9), (31, 31), (48, 47), (63, 20), (95, 95), (96, 64), (97, 33)])
def test_launch_shapes_fixed_and_general_loops(m, n, dtype):
    import importlib
    P = importlib.import_module('pavel-s_svd_amd.standalone.primitives')
    a = np.stack([X.make_matrix(dict(m=m, n=n, family='gaussian', dtype=dtype, seed=7 * m + n + k))
                  for k in range(3)])
    x = torch.from_numpy(a).to(DEV)
    tol = 1e-5 if dtype == 'float32' else 1e-12
    ref = None
    for threads, lanes in ((64, 2), (128, 4), (256, 8), (256, 16), (512, 16), (512, 32)):
        u, s, vh, info = P.jacobi_lds_uniform(x, lanes=lanes, threads=threads)
        assert (info[:, 1] == 0).all(), f'status {info[:, 1].tolist()} at {threads}x{lanes}'
        for b in range(a.shape[0]):
            met = _metrics(a[b], (u[b], s[b], vh[b]))
            assert met['spectrum_error_over_sigma1'] < tol, (threads, lanes, met)
            assert met['orthogonality_u_fro_over_sqrt_cols'] < tol, (threads, lanes, met)
            assert met['orthogonality_v_fro_over_sqrt_rows'] < tol, (threads, lanes, met)
        sv = _np(s)
        if ref is None:
            ref = sv
        else:   # every launch shape gives the same spectrum to the accuracy of the data type
            assert np.max(np.abs(sv - ref)) <= tol * np.max(ref), (threads, lanes)


def test_rank_deficient_completion_and_zero(pkg):
    for a in (np.zeros((12, 8), np.float32), np.diag([0., 2., 0., 3., 1.]),
              X.make_matrix(dict(m=15, n=7, dtype='float64', family='low_rank', rank=3, tail=0))):
        u, s, vh = pkg.svd(torch.from_numpy(np.ascontiguousarray(a)).to(DEV), 'jacobi')
        _, rs, _ = R.svd(a, 'jacobi')
        assert np.array_equal(_np(s) == 0, rs == 0)          # same cut decisions as the oracle
        uu = _np(u).astype(np.float64)
        assert np.allclose(uu.T @ uu, np.eye(uu.shape[1]), atol=1e-5)


def test_values_only_matches_vectors_bitwise(pkg):
    a = torch.from_numpy(X.make_matrix(dict(m=40, n=24, batch=3, dtype='float32', family='gaussian'))).to(DEV)
    _, s1, _ = pkg.svd(a, 'jacobi')
    none, s2, none2 = pkg.svd(a, 'jacobi', values_only=True)
    assert none is None and none2 is None
    assert torch.equal(s1, s2)


def test_errors_map_to_reference_exceptions(pkg):
    a = torch.from_numpy(np.random.default_rng(0).normal(size=(12, 9))).to(DEV)
    with pytest.raises(RuntimeError):
        pkg.svd(a, 'jacobi', max_sweeps=1)
    with pytest.raises(ValueError):
        pkg.svd(a * float('nan'), 'jacobi')
    saved = a.clone()
    pkg.svd(a, 'jacobi')
    assert torch.equal(a, saved)                                  # input never overwritten


def test_uneven_batch_against_lapack_and_oracle(pkg):
    from importlib import import_module
    B = import_module('pavel-s_svd_amd.standalone.batched')
    flat, ms, ns, offs = X.workload_uneven_batch(count=300, seed=11)
    batch = B.UnevenBatch(torch.from_numpy(flat).to(DEV), ms, ns)
    batch.configure(cut='floor').solve()
    stats = batch.check()
    mats = X.unpack(flat, ms, ns, offs)
    worst = {}
    for i, ((u, s, vh), a) in enumerate(zip(batch.results(), mats)):
        met = _metrics(a, (u, s, vh))
        for k in ('spectrum_error_over_sigma1', 'orthogonality_u_fro_over_sqrt_cols',
                  'orthogonality_v_fro_over_sqrt_rows', 'reconstruction_relative_fro'):
            worst[k] = max(worst.get(k, 0.0), met[k])
        if i < 20:                                                   # oracle (reference FP32 path)
            _, rs, _ = R.svd(a, 'jacobi')
            assert np.max(np.abs(_np(s) - rs)) / rs[0] < 2e-5
    assert all(v <= 1e-5 for v in worst.values()), worst
    assert 1 <= stats['jacobi_sweeps'] <= 20


def test_uneven_graph_replay_matches_eager(pkg):
    from importlib import import_module
    B = import_module('pavel-s_svd_amd.standalone.batched')
    flat, ms, ns, _ = X.workload_uneven_batch(count=64, seed=12)
    batch = B.UnevenBatch(torch.from_numpy(flat).to(DEV), ms, ns)
    batch.solve()
    torch.cuda.synchronize()
    eager = (batch.u.clone(), batch.s.clone(), batch.vh.clone())
    batch.capture()
    for _ in range(3):
        batch.u.zero_(); batch.s.zero_(); batch.vh.zero_()
        torch.cuda.synchronize()
        batch.replay()
        torch.cuda.synchronize()
        assert torch.equal(batch.u, eager[0]) and torch.equal(batch.s, eager[1]) and torch.equal(batch.vh, eager[2])
        batch.check()


def test_autotuner_and_class_workgroup_overrides(pkg, monkeypatch):
    """The autotuner (two policy candidates, then the per-class stage) leaves the batch configured
    with valid launch shapes, and an explicit per-class workgroup size changes only the launch: sigma
    against LAPACK, and against the default shape to 1e-5 of sigma_1."""
    from importlib import import_module
    B = import_module('pavel-s_svd_amd.standalone.batched')
    T = import_module('pavel-s_svd_amd.autotune')
    monkeypatch.delenv('PAVELSVD_TUNE_CACHE', raising=False)
    monkeypatch.setattr(T, '_CACHE', {})
    flat, ms, ns, _ = X.workload_uneven_batch(count=300, seed=31)
    batch = B.UnevenBatch(torch.from_numpy(flat).to(DEV), ms, ns)
    best = T.tune_uneven(batch, candidates=T.CANDIDATES[:2], finalists=2, rounds=1)
    assert set(best['class_threads']) <= {c['name'] for c in batch.classes}
    assert all(c['threads'] in (64, 128, 256, 512, 1024) for c in batch.classes)
    batch.configure(class_threads={})
    batch.solve()
    s_default = batch.s.cpu().numpy().copy()
    big = max(batch.classes, key=lambda c: c['cost'])
    other = big['threads'] * 2 if big['threads'] < 1024 else big['threads'] // 2
    batch.configure(class_threads={big['name']: other})
    assert [c['threads'] for c in batch.classes if c['name'] == big['name']] == [other]
    batch.solve()
    batch.check()
    s_over = batch.s.cpu().numpy()
    for i, (m, n) in enumerate(zip(ms, ns)):
        so, sd = s_over[batch.s_off[i]:batch.s_off[i] + n], s_default[batch.s_off[i]:batch.s_off[i] + n]
        ai = flat[batch.a_off[i]:batch.a_off[i] + m * n].reshape(m, n).astype(np.float64)
        ref = np.linalg.svd(ai, compute_uv=False)
        assert np.abs(so - ref).max() <= 1e-5 * ref[0]
        assert np.abs(so - sd).max() <= 1e-5 * ref[0]
    with pytest.raises(ValueError):
        batch.configure(class_threads={big['name']: 96})


def test_factory_plugin_hook(pkg):
    a = torch.from_numpy(corpus('square32')).to(DEV)
    run = pkg.factory(a, {'name': 'pavel-s_svd_amd:factory', 'method': 'jacobi'})
    u, s, vh = run()
    assert run.diagnostics['jacobi_sweeps'] >= 1 and run.metadata['vendor_factorizations'] is False


# --- K2: single-launch one-sided Jacobi for matrices beyond LDS ---------------------------------
def _prims():
    from importlib import import_module
    return import_module('pavel-s_svd_amd.standalone.primitives')


def test_grid_kernel_matches_lds_kernel():
    P = _prims()
    a = torch.from_numpy(X.make_matrix(dict(m=80, n=37, batch=2, dtype='float64', family='gaussian', seed=4))).to(DEV)
    u1, s1, vh1, i1 = P.jacobi_lds_uniform(a, scale=True)
    u2, s2, vh2, i2 = P.jacobi_grid(a, scale=True)
    torch.cuda.synchronize()
    assert (i1[:, 1] == 0).all() and (i2[:, 1] == 0).all()
    assert torch.allclose(s1, s2, rtol=0, atol=1e-13)
    for u, s, vh in ((u2, s2, vh2),):
        rec = (u * s[:, None, :]) @ vh
        assert torch.allclose(rec, a, atol=1e-13)


def test_tall2048x32_corpus_via_grid_kernel(pkg, expected):
    a = corpus('tall2048x32')
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'jacobi', return_info=True)
    met = _metrics(a, (u, s, vh))
    assert met['spectrum_error_over_sigma1'] < 1e-5 and met['orthogonality_u_fro_over_sqrt_cols'] < 1e-5
    assert met['orthogonality_v_fro_over_sqrt_rows'] < 1e-5 and met['reconstruction_relative_fro'] < 1e-5
    assert abs(info['jacobi_sweeps'] - 7) <= 2
    _, rs, _ = R.svd(a, 'jacobi')
    assert np.max(np.abs(_np(s) - rs)) / rs[0] < 1e-5


def test_single_512x256_geometric_tol_1e7():
    """Single-matrix workload: 512x256 float32, sigma geometric 1..1e-6, Haar U and V, tol 1e-7
    with the FP64 convergence test; gates against FP64 LAPACK."""
    pkg = load_package()
    a = X.workload_single_jacobi()
    u, s, vh, info = pkg.svd(torch.from_numpy(a).to(DEV), 'jacobi', tol=1e-7, cut='floor', return_info=True)
    met = _metrics(a, (u, s, vh))
    assert met['spectrum_error_over_sigma1'] <= 1e-5, met
    assert met['orthogonality_u_fro_over_sqrt_cols'] <= 1e-5, met
    assert met['orthogonality_v_fro_over_sqrt_rows'] <= 1e-5, met
    assert info['jacobi_sweeps'] <= 30
    # reference cut semantics: the same sigma are zeroed as by the oracle
    u2, s2 = pkg.svd(torch.from_numpy(a).to(DEV), 'jacobi')[:2]
    _, rs, _ = R.svd(a, 'jacobi', max_sweeps=60)
    assert abs(int((_np(s2) == 0).sum()) - int((rs == 0).sum())) <= 2


@pytest.mark.parametrize('dtype', [np.float32, np.float64])
def test_pair_schedules(pkg, dtype):
    """The tuner's schedule axis on an uneven batch (every size class, fixed-slot and general loops):
    slots='spread' serves the same pairs per round in another order, so U, s, Vh and the sweep counts
    are bit-identical to the identity order; ordering='sorted' (the tournament over columns relabelled
    by descending norm each sweep) is a different Jacobi ordering whose results meet the same bars
    against LAPACK and the oracle's sigma (the reference's FP32 path)."""
    from importlib import import_module
    B = import_module('pavel-s_svd_amd.standalone.batched')
    flat, ms, ns, offs = X.workload_uneven_batch(count=240, seed=41)
    flat = flat.astype(dtype)
    batch = B.UnevenBatch(torch.from_numpy(flat).to(DEV), ms, ns)
    outs = {}
    for ordering in ('tournament', 'sorted'):
        for slots in ('identity', 'spread'):
            batch.configure(ordering=ordering, slots=slots)
            batch.solve()
            st = batch.check()
            outs[ordering, slots] = (batch.u.clone(), batch.s.clone(), batch.vh.clone(), st['sweeps_per_matrix'])
    for ordering in ('tournament', 'sorted'):
        a, b = outs[ordering, 'identity'], outs[ordering, 'spread']
        assert all(torch.equal(x, y) for x, y in zip(a[:3], b[:3])) and a[3] == b[3]
    mats = X.unpack(flat, ms, ns, offs)
    bar = 1e-5 if dtype == np.float32 else 1e-12
    for key, (u, s, vh, sweeps) in outs.items():
        worst = {}
        for i, a in enumerate(mats):
            so, ao = batch.s_off[i], batch.a_off[i]
            m, n = a.shape
            met = _metrics(a, (u[ao:ao + m * n].view(m, n), s[so:so + n],
                               vh[batch.v_off[i]:batch.v_off[i] + n * n].view(n, n)))
            for k in ('spectrum_error_over_sigma1', 'orthogonality_u_fro_over_sqrt_cols',
                      'orthogonality_v_fro_over_sqrt_rows', 'reconstruction_relative_fro'):
                worst[k] = max(worst.get(k, 0.0), met[k])
            if i < 12 and dtype == np.float32:
                _, rs, _ = R.svd(a, 'jacobi')
                assert np.max(np.abs(_np(s[so:so + n]) - rs)) / rs[0] < 2e-5
        assert all(v <= bar for v in worst.values()), (key, worst)
        assert 1 <= max(sweeps) <= 20


@pytest.mark.parametrize('dtype', [np.float32, np.float64])
def test_slot_spread_is_bit_identical(pkg, dtype):
    """Same launch shape, identity vs spread slot order: bit-identical outputs and sweep counts."""
    from importlib import import_module
    B = import_module('pavel-s_svd_amd.standalone.batched')
    flat, ms, ns, _ = X.workload_uneven_batch(count=300, seed=43)
    batch = B.UnevenBatch(torch.from_numpy(flat.astype(dtype)).to(DEV), ms, ns)
    res = []
    for slots in ('identity', 'spread'):
        batch.configure(slots=slots)
        batch.solve()
        st = batch.check()
        res.append((batch.u.clone(), batch.s.clone(), batch.vh.clone(), st['sweeps_per_matrix']))
    assert torch.equal(res[0][0], res[1][0]) and torch.equal(res[0][1], res[1][1])
    assert torch.equal(res[0][2], res[1][2]) and res[0][3] == res[1][3]
