"""Benchmark: batched uneven one-sided Jacobi SVD on one MI355X.

Workload: 4096 float32 matrices, m ~ U{16..96}, n ~ U{8..m}, entries i.i.d. N(0,1) (seed
20261011), packed row-major in HBM.  One step = one solve of the whole batch: a single
hipGraph replay of the persistent LDS-resident Jacobi launch (device work queue, FP64 dot
products and convergence test, FP64 sigma).  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-check]
"""
import argparse
import importlib
import json
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


def cpu_baseline(flat, ms, ns, offs, budget_s=15.0):
    """The oracle (numpy restatement of the reference's standalone Jacobi path) on one host core,
    looping over the batch as the reference API requires, for about ``budget_s`` seconds."""
    from threadpoolctl import threadpool_limits
    from oracle import restatement as R
    done, t0 = 0, time.perf_counter()
    with threadpool_limits(1):
        for m, n, o in zip(ms, ns, offs):
            R.svd(flat[o:o + m * n].reshape(m, n), 'jacobi')
            done += 1
            if time.perf_counter() - t0 > budget_s and done >= 16:
                break
    dt = time.perf_counter() - t0
    return {'value': done / dt, 'unit': 'SVDs/s', 'cores': 1, 'kind': 'port',
            'sample': f'first {done} of the {COUNT} matrices, one svd(..., "jacobi") call each, '
                      f'{dt:.1f} s on 1 host core (numpy, FP32 arithmetic as the reference)'}


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
    pkg = importlib.import_module('pavel-s_svd_amd')
    B = importlib.import_module('pavel-s_svd_amd.standalone.batched')
    T = importlib.import_module('pavel-s_svd_amd.autotune')
    dev = torch.device('cuda:0')
    flat, ms, ns, offs = uneven_workload()
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
                   'sweeps_max': int(sweeps.max()), 'sweeps_mean': float(sweeps.mean())},
        'roofline': {'bound': 'mfma', 'achieved': achieved, 'peak': PEAK_FP32_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': achieved / PEAK_FP32_TFLOPS, 'traffic': load_traffic(),
                     'kernel': 'jacobi_lds_kernel<float,*>', 'kernel_ms': kernel_ms,
                     'algorithmic_flops_per_launch': flops},
    }
    if not args.no_check:
        out['accuracy_vs_fp64_lapack'] = accuracy(batch, flat, ms, ns, offs)
    if not args.no_cpu:
        out['cpu_baseline'] = cpu_baseline(flat, ms, ns, offs)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
