"""Benchmark: batched uneven one-sided Jacobi SVD on one MI355X (the headline line), plus the
single-matrix, QDWH, TSQR, learned-l0 and consumer-solver (block_jacobi, gram_jacobi, range
finders) workloads under "configs".

Workload: 4096 float32 matrices, m ~ U{16..96}, n ~ U{8..m}, entries i.i.d. N(0,1) (seed
20261011), packed row-major in HBM.  One step = one solve of the whole batch: a single
hipGraph replay of the persistent LDS-resident Jacobi launch (device work queue, FP64 dot
products and convergence test, FP64 sigma).  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-check] [--no-plain] [--only c2]
"""
import argparse
import importlib
import json
import math
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 (vector and matrix peaks are equal on gfx950)
PEAK_HBM_GBS = 8000.0
COUNT = 4096
COPY_BYTES = 256 << 20
SEED = 20261011


def uneven_workload():
    rng = np.random.default_rng(SEED)
    ms = rng.integers(16, 97, size=COUNT)
    ns = np.array([rng.integers(8, m + 1) for m in ms])
    sizes = ms * ns
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    flat = np.random.default_rng(SEED + 1).standard_normal(int(sizes.sum()), dtype=np.float32)
    return flat, ms.astype(np.int64), ns.astype(np.int64), offs.astype(np.int64)


def accuracy(batch, flat, ms, ns, offs):
    """Worst FP64-LAPACK gates over the whole batch (sigma error over sigma_1, orthogonality)."""
    import scipy.linalg
    res = batch.results()
    u_all, s_all, vh_all = batch.u.cpu().numpy(), batch.s.cpu().numpy(), batch.vh.cpu().numpy()
    worst = dict(spectrum_error_over_sigma1=0.0, orthogonality_u=0.0, orthogonality_v=0.0,
                 reconstruction_relative_fro=0.0)
    for i, (m, n, o) in enumerate(zip(ms, ns, offs)):
        a = flat[o:o + m * n].reshape(m, n).astype(np.float64)
        so, vo = batch.s_off[i], batch.v_off[i]
        s = s_all[so:so + n].astype(np.float64)
        u = u_all[o:o + m * n].reshape(m, n).astype(np.float64)
        vh = vh_all[vo:vo + n * n].reshape(n, n).astype(np.float64)
        sr = scipy.linalg.svdvals(a)
        worst['spectrum_error_over_sigma1'] = max(worst['spectrum_error_over_sigma1'], float(np.max(np.abs(s - sr)) / sr[0]))
        worst['orthogonality_u'] = max(worst['orthogonality_u'], float(np.linalg.norm(u.T @ u - np.eye(n)) / np.sqrt(n)))
        worst['orthogonality_v'] = max(worst['orthogonality_v'], float(np.linalg.norm(vh @ vh.T - np.eye(n)) / np.sqrt(n)))
        worst['reconstruction_relative_fro'] = max(worst['reconstruction_relative_fro'],
                                                   float(np.linalg.norm(a - (u * s) @ vh) / np.linalg.norm(a)))
    del res
    return worst


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo) and the cores this process may run on."""
    name = 'unknown'
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    name = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    return name, len(os.sched_getaffinity(0))


_POOL_BATCH = None   # (flat, ms, ns, offs) handed to forked workers


def _cpu_worker(args):
    """Worker w of N: the oracle on matrices w, w + N, ... (one host core, one BLAS thread) until
    `budget_s` has passed; returns (solved, seconds)."""
    w, nw, budget_s = args
    from threadpoolctl import threadpool_limits
    from oracle import restatement as R
    flat, ms, ns, offs = _POOL_BATCH
    done, t0 = 0, time.perf_counter()
    with threadpool_limits(1):
        for i in range(w, len(ms), nw):
            m, n, o = ms[i], ns[i], offs[i]
            R.svd(flat[o:o + m * n].reshape(m, n), 'jacobi')
            done += 1
            if time.perf_counter() - t0 > budget_s:
                break
    return done, time.perf_counter() - t0


def cpu_baseline(flat, ms, ns, offs, budget_s=10.0):
    """The oracle (numpy restatement of the reference's standalone Jacobi path) on the host, one
    svd(a_i, 'jacobi') call per matrix as the reference's uniform-shape API requires, for about
    ``budget_s`` seconds: with all cores of this process's affinity (one worker process per core,
    at most 16, each on an interleaved share of the batch) and with 4 workers, as the reference's
    published CPU runs used 4 threads.  Runs before this process touches the GPU (forked workers)."""
    import multiprocessing as mp
    global _POOL_BATCH
    _POOL_BATCH = (flat, ms, ns, offs)
    model, cores = cpu_model()
    rows = {}
    for workers in sorted({min(cores, 16), min(cores, 4)}, reverse=True):
        t0 = time.perf_counter()
        with mp.get_context('fork').Pool(workers) as pool:
            res = pool.map(_cpu_worker, [(w, workers, budget_s) for w in range(workers)])
        wall = time.perf_counter() - t0
        done = sum(d for d, _ in res)
        rows[workers] = {'value': done / max(t for _, t in res), 'matrices': done, 'wall_s': wall}
    best = max(rows)
    out = {'value': rows[best]['value'], 'unit': 'SVDs/s', 'cores': best, 'kind': 'port', 'cpu': model,
           'affinity_cores': cores,
           'sample': f'{rows[best]["matrices"]} of the {COUNT} matrices (interleaved shares, {budget_s:.0f} s per '
                     f'worker), one svd(..., "jacobi") call each, {best} worker processes x 1 thread (numpy, FP32 '
                     f'arithmetic as the reference)'}
    if 4 in rows and best != 4:
        out['cores_4'] = {'value': rows[4]['value'], 'matrices': rows[4]['matrices']}
    return out


def _scaled_phase(fn, small, full):
    """Seconds of fn() at size `small`, scaled by (full / small)^3 (the O(n^3) phases of the oracle)."""
    t0 = time.perf_counter()
    try:
        fn()
    except RuntimeError:       # a capped sweep count raises, as the reference does; the time is what counts
        pass
    return (time.perf_counter() - t0) * (full / small) ** 3


def cpu_single(a):
    """The oracle's one-sided Jacobi on the 512 x 256 matrix at tol 1e-7 (whether the reference's
    FP32 arithmetic converges there) and at the reference's default tol (10 eps32), one core."""
    from threadpoolctl import threadpool_limits
    from oracle import restatement as R
    out = {'unit': 's/solve', 'cores': 1, 'kind': 'port', 'cpu': cpu_model()[0]}
    with threadpool_limits(1):
        for key, tol in (('tol_1e-7', 1e-7), ('tol_default', None)):
            t0 = time.perf_counter()
            try:
                _, _, _, sweeps = R.jacobi_svd(a[None], tol=tol, max_sweeps=60)
                row = {'converged': True, 'jacobi_sweeps': int(sweeps)}
            except RuntimeError:
                row = {'converged': False, 'jacobi_sweeps': 60}
            row['seconds'] = time.perf_counter() - t0
            out[key] = row
    out['value'] = out['tol_1e-7']['seconds']
    out['sample'] = ('the whole solve (numpy, FP32 arithmetic as the reference; numpy elementwise work runs on one '
                     'core whatever the thread count)')
    return out


def cpu_qdwh(n, qr_iterations, chol_iterations, eigen_sweeps, small=768, seed=SEED):
    """Extrapolated oracle time of one QDWH solve at n: each phase of the restatement (the QR of
    [sqrt(c) X; I], the Cholesky and two triangular solves, one sweep of the two-sided Jacobi
    eigensolver) timed once at n = `small` on the same workload family, scaled by (n / small)^3 and
    multiplied by the counts the GPU run executed.  In-cache small problems make this a lower bound."""
    from threadpoolctl import threadpool_limits
    from oracle import matrices as X
    from oracle import restatement as R
    a, _ = X.workload_qdwh(small, seed)
    x = (a / np.linalg.norm(a)).astype(np.float32)
    g = (np.eye(small, dtype=np.float32) + 3.0 * x.T @ x)[None]
    with threadpool_limits(1):
        stacked = np.concatenate([x, np.eye(small, dtype=np.float32)])[None]
        t_qr = _scaled_phase(lambda: R.qr(stacked), small, n)
        t_chol = _scaled_phase(lambda: R.cholesky(g), small, n)
        low = np.tril(g)
        t_trsm = _scaled_phase(lambda: R.trsm_lower(low, x.T[None].copy()), small, n)
        h = (x.T @ a)[None]
        h = ((h + np.swapaxes(h, 1, 2)) / 2).astype(np.float32)
        t_sweep = _scaled_phase(lambda: R.jacobi_eigh(h, max_sweeps=1), small, n)
    total = qr_iterations * t_qr + chol_iterations * (t_chol + 2 * t_trsm) + eigen_sweeps * t_sweep
    return {'value': total, 'unit': 's/solve', 'cores': 1, 'kind': 'port', 'cpu': cpu_model()[0],
            'extrapolated': True,
            'phases_s': {'qr_iteration': t_qr, 'cholesky_iteration': t_chol + 2 * t_trsm, 'eigen_sweep': t_sweep},
            'counts': {'qr_iterations': qr_iterations, 'cholesky_iterations': chol_iterations,
                       'eigen_sweeps': eigen_sweeps},
            'sample': f'each phase of the oracle timed once at n={small} (same workload family), scaled by '
                      f'(n/{small})^3, times the counts of the GPU run (its eigen sweeps are the blocked solver\'s)'}


def cpu_tsqr(m, n, leaves, sweeps, small=512, seed=SEED):
    """Extrapolated oracle time of TSQR + Jacobi on R at m x n with `leaves` square n x n leaves:
    the leaf QR (n x n), the merge QR (2n x n) and one sweep of the one-sided Jacobi on R, each timed at
    n = `small` and scaled by (n / small)^3, the back-transformation products by their flops; times the
    GPU run's counts (leaves, leaves - 1 merges, sweeps)."""
    from threadpoolctl import threadpool_limits
    from oracle import restatement as R
    rng = np.random.default_rng(seed)
    leaf = rng.standard_normal((1, small, small)).astype(np.float32)
    merge = rng.standard_normal((1, 2 * small, small)).astype(np.float32)
    with threadpool_limits(1):
        t_leaf = _scaled_phase(lambda: R.qr(leaf), small, n)
        t_merge = _scaled_phase(lambda: R.qr(merge), small, n)
        r = np.triu(rng.standard_normal((1, small, small))).astype(np.float32)
        t_sweep = _scaled_phase(lambda: R.jacobi_svd(r, max_sweeps=1), small, n)
        q = rng.standard_normal((small, small)).astype(np.float32)
        t_gemm = _scaled_phase(lambda: q @ q, small, n)      # one n x n x n product
    merges = leaves - 1
    # each tree level multiplies all m rows of Q by n x n blocks (Q <- [Q1 q_top; Q2 q_bot]), then U = Q U_R
    products = (math.ceil(math.log2(leaves)) + 1) * m / n
    total = leaves * t_leaf + merges * t_merge + products * t_gemm + sweeps * t_sweep
    return {'value': total, 'unit': 's/solve', 'cores': 1, 'kind': 'port', 'cpu': cpu_model()[0],
            'extrapolated': True,
            'phases_s': {'leaf_qr': t_leaf, 'merge_qr': t_merge, 'jacobi_sweep_on_r': t_sweep, 'n_cubed_product': t_gemm},
            'counts': {'leaves': leaves, 'merges': merges, 'jacobi_sweeps': sweeps},
            'sample': f'each phase of the oracle timed once at n={small}, scaled by (n/{small})^3, times the counts '
                      f'of the GPU run (its sweeps are the blocked solver\'s)'}


def plain_gpu_form(pkg, flat, ms, ns, offs, dev, count=512):
    """The reference API on the GPU, one svd(a_i, 'jacobi') call per matrix (the API takes
    uniform shapes only): the same kernels, tolerance, scaling and cut as the batched launch, without
    batching, the persistent queue or graph replay; the first `count` matrices of the batch."""
    mats = [torch.from_numpy(flat[o:o + m * n].reshape(m, n)).to(dev) for m, n, o in zip(ms[:count], ns[:count],
                                                                                           offs[:count])]
    for a in mats[:8]:
        pkg.svd(a, 'jacobi', cut='floor')
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a in mats:
        pkg.svd(a, 'jacobi', cut='floor')
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {'value': len(mats) / dt, 'unit': 'SVDs/s', 'sample': f'first {len(mats)} of the {COUNT} matrices, one '
            f'svd(..., "jacobi", cut="floor") call each on cuda:0, {dt:.2f} s'}


def _haar64(n, k, gen, dev):
    """n x k Haar columns: Householder QR (this package's FP64 kernel) of seeded N(0, 1), R-diagonal
    sign fix."""
    P = importlib.import_module('pavel-s_svd_amd.standalone.primitives')
    g = torch.randn(1, n, k, generator=gen, device=dev, dtype=torch.float64)
    q, r = P.qr(g)
    return (q * torch.sign(torch.diagonal(r, dim1=1, dim2=2))[:, None, :])[0]


def _product(u, s, v):
    gemm = importlib.import_module('pavel-s_svd_amd.standalone.blas').gemm
    return gemm((u * s[None, :]).contiguous(), v.mT, mode='f64')


def _orth(u):
    """|U^T U - I|_F / sqrt(k) in FP64 on the GPU."""
    gemm = importlib.import_module('pavel-s_svd_amd.standalone.blas').gemm
    u = u.double()
    g = gemm(u.mT, u, mode='f64')
    g.diagonal().sub_(1.0)
    return float(g.norm() / math.sqrt(u.shape[1]))


def _timed(fn, reps):
    """(median seconds over `reps` timed calls after one warm-up call, the last result)."""
    torch.cuda.synchronize()
    out = fn()                                          # warm-up (also returns the result)
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), out


def _recon(a, u, s, vh):
    """|A - U diag(s) Vh|_F / |A|_F in FP64 on the GPU."""
    gemm = importlib.import_module('pavel-s_svd_amd.standalone.blas').gemm
    a64 = a.double()
    resid = gemm((u.double() * s.double()[None, :]).contiguous(), vh.double(), mode='f64') - a64
    return float(resid.norm() / a64.norm())


def _roofline(flops, seconds, what):
    achieved = flops / seconds / 1e12
    return {'bound': 'valu+mfma', 'achieved': achieved, 'peak': PEAK_FP32_TFLOPS, 'unit': 'TFLOP/s',
            'frac': achieved / PEAK_FP32_TFLOPS, 'algorithmic_flops': flops, 'work': what}


def jacobi_flops(m, n, sweeps):
    """The reference's one-sided Jacobi work: n(n-1)/2 (12 m + 6 n) flops per sweep plus 3 m n."""
    return sweeps * n * (n - 1) / 2 * (12 * m + 6 * n) + 3 * m * n


def config_single(pkg, dev, cpu=True):
    """512 x 256 float32, Haar U and V, sigma geometric 1 .. 1e-6; one-sided Jacobi at tol 1e-7."""
    gen = torch.Generator(device=dev)
    gen.manual_seed(SEED)
    s = torch.from_numpy(np.geomspace(1.0, 1e-6, 256)).to(dev)
    a = _product(_haar64(512, 256, gen, dev), s, _haar64(256, 256, gen, dev)).float()
    res = {}
    for cut in ('reference', 'floor'):
        t, (u, sv, vh, info) = _timed(lambda: pkg.svd(a, 'jacobi', tol=1e-7, cut=cut, return_info=True), 5)
        res[cut] = {'seconds': t, 'jacobi_sweeps': info['jacobi_sweeps'],
                    'spectrum_error_over_sigma1': float((sv.double() - s).abs().max()),
                    'orthogonality_u': _orth(u), 'orthogonality_v': _orth(vh.mT),
                    'reconstruction_relative_fro': _recon(a, u, sv, vh)}
    out = {'workload': '512x256 float32, Haar U/V, sigma geomspace(1, 1e-6); jacobi tol 1e-7 (the cut '
                       "'reference' zeroes sigma < eps*max(m,n)*sigma_1 = 6.1e-5 by the reference's contract)",
           'unit': 's/solve', 'results': res,
           'roofline': _roofline(jacobi_flops(512, 256, res['floor']['jacobi_sweeps']), res['floor']['seconds'],
                                 "reference one-sided Jacobi work x executed sweeps (cut 'floor'), whole solve")}
    if cpu:
        out['cpu_baseline'] = cpu_single(a.cpu().numpy())
    return out


def qdwh_flops(n, polar_iterations, qr_iterations, eigen_sweeps):
    """QDWH work as counted for the reference's path: (26/3) n^3 per QR-branch iteration, (16/3) n^3 per
    Cholesky-branch iteration (the reference also forms X^T X: 2 n^3), 4 n^3 for H and U, 9 n^3 per
    eigensolver sweep."""
    chol = polar_iterations - qr_iterations
    return n ** 3 * (26 / 3 * qr_iterations + 16 / 3 * chol + 4 + 9 * eigen_sweeps)


def config_qdwh(pkg, dev, n=4096, cpu=True):
    """n x n float32, Haar U and V, sigma log-uniform on [1e-6, 1] with 10% in a width-1e-4 cluster
    at 0.5: the MI355X form (FP32 products, FP64 polish), the same with BF16x3 QR-branch products,
    the plain FP32 form and the reference's defaults (cut at eps max(m, n) sigma_1 with the
    completion of the cut columns, FP32 products, no polish)."""
    gen = torch.Generator(device=dev)
    gen.manual_seed(SEED)
    rng = np.random.default_rng(SEED)
    k = int(round(0.1 * n))
    sig = np.sort(np.concatenate([np.exp(rng.uniform(np.log(1e-6), 0.0, n - k)),
                                  0.5 + 1e-4 * (rng.random(k) - 0.5)]))[::-1].copy()
    sig[0] = 1.0
    s = torch.from_numpy(sig).to(dev)
    a = _product(_haar64(n, n, gen, dev), s, _haar64(n, n, gen, dev)).float()
    forms = {'mi355x': dict(gemm='f32', polish=True, cut='floor'),
             'bf16x3_polish': dict(gemm='bf16x3', polish=True, cut='floor'),
             'plain_fp32': dict(gemm='f32', cut='floor'),
             'reference_defaults': dict()}
    res = {}
    for name, opts in forms.items():
        t, (u, sv, vh, info) = _timed(lambda: pkg.svd(a, 'polar_qdwh', return_info=True, **opts), 3)
        kept = sv > 0
        res[name] = {'seconds': t, 'options': opts, 'polar_iterations': info['polar_iterations'],
                     'polar_qr_iterations': info.get('polar_qr_iterations'), 'eigen_sweeps': info['eigen_sweeps'],
                     'polish_sweeps': info.get('polish_sweeps'),
                     'kept_sigma': int(kept.sum()),
                     'spectrum_error_over_sigma1': float((sv.double() - s)[kept].abs().max()),
                     'orthogonality_u': _orth(u), 'orthogonality_v': _orth(vh.mT),
                     'reconstruction_relative_fro': _recon(a, u, sv, vh)}
        res[name]['roofline'] = _roofline(qdwh_flops(n, info['polar_iterations'], info.get('polar_qr_iterations', 0),
                                                     info['eigen_sweeps']), t,
                                          'QDWH iterations + H, U + eigensolver sweeps as the reference counts them '
                                          '(FP64 polish work not counted), whole solve')
        del u, sv, vh
    out = {'workload': f'QDWH polar SVD, {n}x{n} float32, Haar U/V, sigma log-uniform [1e-6, 1] with 10% in a '
                       'width-1e-4 cluster at 0.5; sigma error over the kept values against the generating spectrum '
                       '(the reference defaults zero sigma < eps*max(m,n)*sigma_1 = 4.9e-4 and complete those U '
                       'columns)',
           'unit': 's/solve', 'results': res,
           'bf16x3_speedup': res['mi355x']['seconds'] / res['bf16x3_polish']['seconds'],
           'speedup_vs_plain_fp32': res['plain_fp32']['seconds'] / res['mi355x']['seconds']}
    if cpu:
        r = res['reference_defaults']
        out['cpu_baseline'] = cpu_qdwh(n, r['polar_qr_iterations'], r['polar_iterations'] - r['polar_qr_iterations'],
                                       r['eigen_sweeps'])
    return out


def config_tsqr(pkg, dev, cpu=True):
    """8192 x 2048 float32: rank-64 G1 diag(sigma) G2^T (orthonormal factors, sigma geometric 1 ..
    1e-3) plus N(0, 1e-8) noise; TSQR with 256-row panels, then Jacobi on R: the MI355X form (FP32
    products, FP64 polish of U and V), the same with BF16x3 products, and the plain FP32 form."""
    gen = torch.Generator(device=dev)
    gen.manual_seed(SEED)
    s = torch.from_numpy(np.geomspace(1.0, 1e-3, 64)).to(dev)
    a = _product(_haar64(8192, 64, gen, dev), s, _haar64(2048, 64, gen, dev))
    a = (a + 1e-4 * torch.randn(8192, 2048, generator=gen, device=dev, dtype=torch.float64)).float()
    res = {}
    m, n = a.shape
    for name, opts in (('mi355x', dict(gemm='f32', polish=True)), ('bf16x3_polish', dict(gemm='bf16x3', polish=True)),
                       ('plain_fp32', dict(gemm='f32'))):
        t, (u, sv, vh, info) = _timed(lambda: pkg.svd(a, 'tsqr', tile_rows=256, return_info=True, **opts), 3)
        res[name] = {'seconds': t, 'options': opts, 'tsqr_leaves': info.get('tsqr_leaves'),
                     'jacobi_sweeps': info['jacobi_sweeps'],
                     'reconstruction_relative_fro': _recon(a, u, sv, vh),
                     'orthogonality_u': _orth(u), 'orthogonality_v': _orth(vh.mT)}
        flops = 2 * (2 * m * n * n - 2 * n ** 3 / 3) + 2 * m * n * n + jacobi_flops(n, n, info['jacobi_sweeps'])
        res[name]['roofline'] = _roofline(flops, t, 'geqrf + orgqr + Q U_R + reference Jacobi work on R x executed '
                                                    'sweeps, whole solve')
        del u, sv, vh
    out = {'workload': 'TSQR, 8192x2048 float32, rank-64 (sigma 1 .. 1e-3) plus N(0, 1e-8) noise (its |E|_2 ~ '
                       '0.014 sets the spectrum, so accuracy is reconstruction and orthogonality in FP64), 256-row '
                       'panels (2048-row leaves: rows = max(tile_rows, n)), Jacobi on R', 'unit': 's/solve',
           'results': res, 'bf16x3_speedup': res['mi355x']['seconds'] / res['bf16x3_polish']['seconds']}
    if cpu:
        r = res['plain_fp32']
        out['cpu_baseline'] = cpu_tsqr(m, n, r['tsqr_leaves'] or 1, r['jacobi_sweeps'])
    return out


def config_consumers(pkg, dev):
    """Solvers built on the hot path's kernels: block_jacobi (fused block-pair rounds) and gram_jacobi
    on the 512 x 256 single-matrix workload; complex jacobi on 1024 complex64 48 x 48 matrices,
    complex qr on 1024 complex64 96 x 48, complex polar_qdwh and complex block_jacobi on 256 complex64
    48 x 48;
    randomized / subspace / block_krylov (rank 64,
    oversample 8, two power steps) on the 8192 x 2048 rank-64-plus-noise TSQR workload."""
    gen = torch.Generator(device=dev)
    gen.manual_seed(SEED)
    s1 = torch.from_numpy(np.geomspace(1.0, 1e-6, 256)).to(dev)
    a1 = _product(_haar64(512, 256, gen, dev), s1, _haar64(256, 256, gen, dev)).float()
    eps32 = float(np.finfo(np.float32).eps)
    res = {}
    for name, opts in (('block_jacobi', {}), ('block_jacobi_16', {'block_size': 16}), ('gram_jacobi', {})):
        method = name.split('_16')[0]
        t, (u, sv, vh, info) = _timed(lambda: pkg.svd(a1, method, return_info=True, **opts), 3)
        cut = eps32 * 512 * float(sv[0]) if method == 'block_jacobi' else math.sqrt(eps32 * 256) * float(sv[0])
        kept = s1 > cut
        res[name] = {'seconds': t, 'options': opts,
                     **{k: info[k] for k in ('block_sweeps', 'jacobi_sweeps', 'eigen_sweeps', 'gram_discarded_values')
                        if k in info},
                     'kept_sigma': int((sv > 0).sum()),
                     'kept_spectrum_error_over_sigma1': float((sv.double() - s1)[kept].abs().max()),
                     'orthogonality_v': _orth(vh.mT)}
    gen.manual_seed(SEED)
    s4 = torch.from_numpy(np.geomspace(1.0, 1e-3, 64)).to(dev)
    a4 = _product(_haar64(8192, 64, gen, dev), s4, _haar64(2048, 64, gen, dev))
    a4 = (a4 + 1e-4 * torch.randn(8192, 2048, generator=gen, device=dev, dtype=torch.float64)).float()
    ref = None
    for method in ('randomized', 'subspace', 'block_krylov'):
        t, (u, sv, vh, info) = _timed(lambda: pkg.svd(a4, method, rank=64, oversample=8, iterations=2, seed=31,
                                                      return_info=True), 3)
        if ref is None:     # FP64 Jacobi of the same matrix: the top-64 reference values
            ref = pkg.svd(a4.double(), 'tsqr', tile_rows=256, values_only=True)[1][:64]
        res[method] = {'seconds': t, **{k: info[k] for k in ('jacobi_sweeps', 'sketch_columns', 'range_columns')},
                       'top64_sigma_error_over_sigma1': float((sv.double() - ref).abs().max() / ref[0]),
                       'orthogonality_u': _orth(u), 'orthogonality_v': _orth(vh.mT)}
    rng = np.random.default_rng(SEED)
    ac = (rng.standard_normal((1024, 48, 48)) + 1j * rng.standard_normal((1024, 48, 48))).astype(np.complex64)
    xc = torch.from_numpy(ac).to(dev)
    t, (u, sv, vh, info) = _timed(lambda: pkg.svd(xc, 'jacobi', return_info=True), 3)
    ref = np.linalg.svd(ac[:64].astype(np.complex128), compute_uv=False)
    err = float(np.max(np.abs(sv[:64].double().cpu().numpy() - ref) / ref[:, :1]))
    uu = u[:64].to(torch.complex128)
    orth = float((uu.mH @ uu - torch.eye(48, dtype=torch.complex128, device=dev)).abs().max())
    res['jacobi_complex64_1024x48x48'] = {'seconds': t, 'svds_per_second': 1024 / t,
                                          'jacobi_sweeps': info['jacobi_sweeps'],
                                          'spectrum_error_over_sigma1_first64': err, 'orthogonality_u_max_first64': orth}
    aq = (rng.standard_normal((1024, 96, 48)) + 1j * rng.standard_normal((1024, 96, 48))).astype(np.complex64)
    xq = torch.from_numpy(aq).to(dev)
    t, (u, sv, vh, info) = _timed(lambda: pkg.svd(xq, 'qr', return_info=True), 3)
    ref = np.linalg.svd(aq[:64].astype(np.complex128), compute_uv=False)
    err = float(np.max(np.abs(sv[:64].double().cpu().numpy() - ref) / ref[:, :1]))
    uu = u[:64].to(torch.complex128)
    orth = float((uu.mH @ uu - torch.eye(48, dtype=torch.complex128, device=dev)).abs().max())
    res['qr_complex64_1024x96x48'] = {'seconds': t, 'svds_per_second': 1024 / t, 'jacobi_sweeps': info['jacobi_sweeps'],
                                      'spectrum_error_over_sigma1_first64': err, 'orthogonality_u_max_first64': orth}
    ap = (rng.standard_normal((256, 48, 48)) + 1j * rng.standard_normal((256, 48, 48))).astype(np.complex64)
    xp = torch.from_numpy(ap).to(dev)
    t, (u, sv, vh, info) = _timed(lambda: pkg.svd(xp, 'polar_qdwh', return_info=True), 3)
    ref = np.linalg.svd(ap[:64].astype(np.complex128), compute_uv=False)
    err = float(np.max(np.abs(sv[:64].double().cpu().numpy() - ref) / ref[:, :1]))
    uu = u[:64].to(torch.complex128)
    orth = float((uu.mH @ uu - torch.eye(48, dtype=torch.complex128, device=dev)).abs().max())
    res['polar_qdwh_complex64_256x48x48'] = {'seconds': t, 'svds_per_second': 256 / t,
                                             'polar_iterations': info['polar_iterations'],
                                             'eigen_sweeps': info['eigen_sweeps'],
                                             'spectrum_error_over_sigma1_first64': err,
                                             'orthogonality_u_max_first64': orth}
    ab = (rng.standard_normal((256, 48, 48)) + 1j * rng.standard_normal((256, 48, 48))).astype(np.complex64)
    xb = torch.from_numpy(ab).to(dev)
    t, (u, sv, vh, info) = _timed(lambda: pkg.svd(xb, 'block_jacobi', return_info=True), 3)
    ref = np.linalg.svd(ab[:64].astype(np.complex128), compute_uv=False)
    err = float(np.max(np.abs(sv[:64].double().cpu().numpy() - ref) / ref[:, :1]))
    uu = u[:64].to(torch.complex128)
    orth = float((uu.mH @ uu - torch.eye(48, dtype=torch.complex128, device=dev)).abs().max())
    res['block_jacobi_complex64_256x48x48'] = {'seconds': t, 'svds_per_second': 256 / t,
                                               'block_sweeps': info['block_sweeps'],
                                               'jacobi_sweeps': info['jacobi_sweeps'],
                                               'spectrum_error_over_sigma1_first64': err,
                                               'orthogonality_u_max_first64': orth}
    return {'workload': 'block_jacobi (block_size 4 and 16) and gram_jacobi on 512x256 float32, Haar U/V, sigma '
                        'geomspace(1, 1e-6) (errors over the sigma each method keeps: block_jacobi cuts at '
                        'eps*max(m,n)*sigma_1, gram_jacobi at sqrt(eps*n)*sigma_1); randomized / subspace / '
                        'block_krylov, rank 64 + 8, two power steps, seed 31, on 8192x2048 float32 rank-64 '
                        '(sigma 1 .. 1e-3) plus N(0, 1e-8) noise, top-64 sigma against FP64 TSQR + Jacobi; '
                        'complex jacobi on 1024 complex64 48x48 Gaussian matrices (one batched call, LDS kernel) '
                        'and complex qr on 1024 complex64 96x48 (LDS QR, Jacobi of R, Q U_R) and complex polar_qdwh '
                        'on 256 complex64 48x48 (QR-form steps) and complex block_jacobi (block_size 4) on 256 '
                        'complex64 48x48 (each round one batched Jacobi launch over its group pairs), sigma of the '
                        'first 64 against complex128 LAPACK',
            'unit': 's/solve', 'results': res}


def config_l0(pkg, dev, count=32, cpu=True):
    """The learned QDWH lower bound on `count` matrices of its test split (1024 x 1024)."""
    T = importlib.import_module('pavel-s_svd_amd.train_l0')
    L = importlib.import_module('pavel-s_svd_amd.learned_l0')
    if not os.path.exists(L.DEFAULT_MODEL):
        return {'skipped': 'no trained model (pavel-s_svd_amd/l0_model.json)'}
    with open(L.DEFAULT_MODEL) as f:
        ev = json.load(f).get('evaluation', {})
    model = L.L0Model.load()
    n, start = ev.get('n', 1024), ev.get('train', 1280) + ev.get('holdout', 256)
    rows = []
    for i0 in range(start, start + count, T.CHUNK):
        a = T.generate(i0, min(i0 + T.CHUNK, start + count), n)
        for j in range(a.shape[0]):
            T.polar_time(a[j:j + 1], None)                  # warm
            rows.append([T.polar_time(a[j:j + 1], mdl, stop) for stop in (None, 1e-2) for mdl in (None, model)])
    arr = np.array([[v for run in r for v in run] for r in rows])
    out = {'workload': f'QDWH polar phase on {count} test matrices {n}x{n} of the learned-l0 workload (log10 '
                       'cond ~ U[1, 12], geometric or clustered spectra); learned time includes features and '
                       'inference; orth_stop = the polished form (stop at |X^T X - I|_F <= 1e-2, three FP64 '
                       'Newton-Schulz steps)', 'unit': 's/solve'}
    for k, name in enumerate(('reference_stop_eps', 'reference_stop_learned', 'orth_stop_eps', 'orth_stop_learned')):
        c = 4 * k
        out[name] = {'seconds': float(arr[:, c].mean()), 'iterations': float(arr[:, c + 1].mean()),
                     'qr_iterations': float(arr[:, c + 2].mean())}
    out['model_evaluation'] = ev
    if cpu:
        e = out['reference_stop_eps']
        qr_it = int(round(e['qr_iterations']))
        base = cpu_qdwh(n, qr_it, int(round(e["iterations"])) - qr_it, 0, small=512)
        base['sample'] = ('the polar phase only (as the GPU figures above), from l0 = eps: ' + base['sample']
                          + '; mean counts of the eps runs rounded')
        out['cpu_baseline'] = base
    return out


def config_gemm(pkg, dev):
    """K3 (psvd_gemm) per arithmetic mode at the shapes the QDWH and TSQR solves issue: the square
    products of QDWH 4096 (X^T X, Q_top Q_bot^T, H = X^T A) and TSQR's 2048^3 back-transformation are
    compute-bound; the two products of a QR panel's trailing update on the stacked [sqrt(c) X; I]
    (W = V^T A: 32 x 4064 x 8192, then A -= V W': 8192 x 4064 x 32) read or rewrite the whole trailing
    matrix for 2 x 32 flops per element and are HBM-bound.  TF/s against the mode's MFMA peak and GB/s
    of the minimum traffic (C read and written once, A and B read once) against 8 TB/s.  The public
    psvd_gemm has no workspace, so the 32 x 4064 output runs unsplit on 32 workgroups; the QR solver
    issues that product through the split-K path (32 K-slices and an ordered reduction)."""
    gemm = importlib.import_module('pavel-s_svd_amd.standalone.blas').gemm
    shapes = {'square_4096': (4096, 4096, 4096, False, False), 'qdwh_qtop_qbotT_4096': (4096, 4096, 4096, False, True),
              'tsqr_backtransform_2048': (2048, 2048, 2048, False, False),
              'qr_panel_vtA_32x4064x8192': (32, 4064, 8192, True, False),
              'qr_trailing_update_8192x4064x32': (8192, 4064, 32, False, False)}
    peaks = {'f32': PEAK_FP32_TFLOPS, 'bf16x3': 2500.0 / 3, 'f64': 78.6}
    gen = torch.Generator(device=dev)
    gen.manual_seed(SEED)
    res = {}
    for name, (m, n, k, ta, tb) in shapes.items():
        for mode in ('f32', 'bf16x3', 'f64'):
            dt = torch.float64 if mode == 'f64' else torch.float32
            a = torch.randn(*((k, m) if ta else (m, k)), generator=gen, device=dev, dtype=dt)
            b = torch.randn(*((n, k) if tb else (k, n)), generator=gen, device=dev, dtype=dt)
            c = torch.zeros(m, n, device=dev, dtype=dt)
            aa, bb = (a.mT if ta else a), (b.mT if tb else b)
            t, _ = _timed(lambda: gemm(aa, bb, alpha=-1.0, beta=1.0, out=c, mode=mode), 5)
            size = 8 if mode == 'f64' else 4
            tf = 2.0 * m * n * k / t / 1e12
            gbs = size * (2 * m * n + m * k + k * n) / t / 1e9
            res[f'{name}:{mode}'] = {'ms': t * 1e3, 'tflops': tf, 'frac_of_mode_peak': tf / peaks[mode],
                                     'min_traffic_gbs': gbs, 'frac_of_hbm': gbs / PEAK_HBM_GBS}
            del a, b, c
    return {'workload': 'psvd_gemm C = A B - C at the QDWH / TSQR shapes, three arithmetic modes, median of 5 calls'
                            ' (public entry point, no split-K workspace)',
            'unit': 'ms', 'peaks_tflops': peaks, 'results': res}


def load_traffic():
    path = os.path.join(ROOT, 'profiles', 'jacobi_lds_uneven4096_traffic.json')
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get('hbm_bytes_per_launch')
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-check', action='store_true')
    ap.add_argument('--no-tune', action='store_true', help='use the default launch-shape policy')
    ap.add_argument('--no-plain', action='store_true', help='skip the plain GPU form (per-matrix calls), e.g. when '
                    'profiling the batched launch alone')
    ap.add_argument('--only', default='all', choices=['all', 'c2', 'single', 'qdwh', 'tsqr', 'l0', 'consumers', 'gemm'],
                    help="'c2': the batched-Jacobi line alone; single / qdwh / tsqr / l0 / consumers: that leg alone "
                         "(printed as {\"configs\": ...}, for profiling)")
    ap.add_argument('--copy-calibration', action='store_true',
                    help='only run a plain 256 MiB device-to-device copy 5 times (the reference kernel for '
                         'scaling FETCH_SIZE / WRITE_SIZE counter readings to bytes) and exit')
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py measures one GPU: every workload here runs on a single MI355X '
              '(multi-GPU sharding is a design note, not a measured path).', file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        x = torch.randn(COPY_BYTES // 4, device='cuda:0')
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({'copy_bytes_read': COPY_BYTES, 'copy_bytes_written': COPY_BYTES}), flush=True)
        return
    flat, ms, ns, offs = uneven_workload()
    cpu = None
    if not args.no_cpu and args.only in ('all', 'c2'):
        # host-only work first: its worker processes are forked before this process touches the GPU
        cpu = cpu_baseline(flat, ms, ns, offs)
        print(f'cpu baseline done: {cpu["value"]:.1f} SVDs/s on {cpu["cores"]} cores', file=sys.stderr, flush=True)
    pkg = importlib.import_module('pavel-s_svd_amd')
    legs = {'single': ('single_512x256', config_single), 'qdwh': ('qdwh_4096', config_qdwh),
            'tsqr': ('tsqr_8192x2048', config_tsqr), 'l0': ('learned_l0_1024', config_l0),
            'consumers': ('consumers', config_consumers), 'gemm': ('gemm', config_gemm)}
    if args.only in legs:
        name, fn = legs[args.only]
        kw = {} if name in ('consumers', 'gemm') else {'cpu': not args.no_cpu}
        print(json.dumps({'configs': {name: fn(pkg, torch.device('cuda:0'), **kw)}}), flush=True)
        return
    B = importlib.import_module('pavel-s_svd_amd.standalone.batched')
    T = importlib.import_module('pavel-s_svd_amd.autotune')
    dev = torch.device('cuda:0')
    batch = B.UnevenBatch(torch.from_numpy(flat).to(dev), ms, ns)
    params = {} if args.no_tune else T.tune_uneven(batch)
    batch.configure(cut='floor', **params)
    batch.capture()
    for _ in range(args.warmup):
        batch.replay()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for e0, e1 in ev:
        e0.record(stream)
        batch.replay()
        e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    stats = batch.check()
    kernel_ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in ev]))
    ms_per_step = wall * 1e3 / args.steps
    sweeps = np.array(stats['sweeps_per_matrix'])
    flops = float(np.sum(ns * (ns - 1) // 2 * (12 * ms + 6 * ns) * sweeps))
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    out = {
        'metric': 'batched Jacobi SVDs/s (uneven 4096-matrix batch, FP32, <=1e-5 sigma error over sigma_1)',
        'value': COUNT * args.steps / wall, 'unit': 'SVDs/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': ms_per_step, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic (seeded N(0,1) entries)',
        'config': {'workload': 'uneven batch: 4096 float32 matrices, m~U{16..96}, n~U{8..m}, one-sided Jacobi '
                               '(FP64 dot products and convergence test, tol 10*eps32), persistent LDS launch in a hipGraph',
                   'count': COUNT, 'seed': SEED, 'params': batch.params,
                   'classes': [{k: c[k] for k in ('threads', 'lanes', 'count', 'max_m', 'max_n', 'stream')} for c in batch.classes],
                   'sweeps_max': int(sweeps.max()), 'sweeps_mean': float(sweeps.mean()),
                   'schedule_tuning': (T.report(batch) or {}).get('schedules')},
        'roofline': {'bound': 'valu', 'achieved': achieved, 'peak': PEAK_FP32_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': achieved / PEAK_FP32_TFLOPS, 'traffic': load_traffic(),
                     'kernel': 'jacobi_lds_kernel<float,*>', 'kernel_ms': kernel_ms,
                     'algorithmic_flops_per_launch': flops},
    }
    if not args.no_check:
        out['accuracy_vs_fp64_lapack'] = accuracy(batch, flat, ms, ns, offs)
    if not args.no_plain:
        plain = plain_gpu_form(pkg, flat, ms, ns, offs, dev)
        plain['batched_speedup'] = out['value'] / plain['value']
        out['plain_gpu_form'] = plain
    if cpu is not None:
        out['cpu_baseline'] = cpu
    if args.only != 'c2':
        del batch
        torch.cuda.empty_cache()
        out['configs'] = {}
        for name, fn in legs.values():
            t0 = time.perf_counter()
            out['configs'][name] = fn(pkg, dev) if name in ('consumers', 'gemm') else fn(pkg, dev, cpu=not args.no_cpu)
            print(f'{name} done in {time.perf_counter() - t0:.1f} s', file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
