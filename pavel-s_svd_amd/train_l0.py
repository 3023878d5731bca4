"""Train and evaluate the learned QDWH lower bound on its benchmark workload.

Workload: `count` float32 n x n matrices A = U diag(sigma) V^T, U and V Haar (Householder QR of
seeded N(0, 1) matrices with the R-diagonal sign fix, FP64 kernels), the product formed in FP64
and rounded once; log10 condition ~ U[1, 12]; spectra alternate between geometric (1 down to
10^-cond) and clustered (log-uniform on that range with a tenth of the values in a cluster of
relative width 1e-4).  Matrix i depends only on (seed, i), so any slice can be regenerated.

Targets: log10(sigma_min / |A|_F) of the FP32 matrix, from the plain one-sided Jacobi solver run in
FP64 on it -- svd(A.double(), 'jacobi', values_only=True, cut='floor') -- floored at 1e-16: an FP32
solver resolves sigma_min only down to its own backward error (about 1e-7 |A|), while the bound QDWH
needs for a numerically singular FP32 matrix lies below that.

    python -m pavel-s_svd_amd.train_l0 data --out DIR [--n 1024 --start 0 --count 2048]
    python -m pavel-s_svd_amd.train_l0 fit --data DIR [--train 1536 --holdout 256] [--model PATH]

`data` writes DIR/l0_data_<start>_<stop>.npz (features, targets, conditions of that slice);
`fit` joins the slices in DIR, trains on the first
train - holdout matrices with the GPU kernels, takes the margin from the holdout, evaluates the
remaining (test) matrices -- QDWH from eps against QDWH from the learned bound, with the reference's
elementwise stop and with the polished form's orthogonality stop: iterations, QR-branch iterations,
polar-phase time including the feature and inference cost -- and saves the model with the
evaluation in its metadata.
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

pkg = importlib.import_module(__package__ or 'pavel-s_svd_amd')
P = importlib.import_module(pkg.__name__ + '.standalone.primitives')
S = importlib.import_module(pkg.__name__ + '.standalone.solvers')
L = importlib.import_module(pkg.__name__ + '.learned_l0')
blas = importlib.import_module(pkg.__name__ + '.standalone.blas')

SEED = 20261012
CHUNK = 16


def spectrum(n, rng, cond, family):
    lo = 10.0 ** -cond
    if family == 'geometric':
        return np.geomspace(1.0, lo, n)
    k = max(1, n // 10)
    s = np.exp(rng.uniform(np.log(lo), 0.0, n - k))
    c = np.exp(rng.uniform(np.log(lo), 0.0)) * (1.0 + 1e-4 * (rng.random(k) - 0.5))
    s = np.sort(np.concatenate([s, c]))[::-1]
    s[0], s[-1] = 1.0, lo
    return s


def meta(i, n, seed=SEED):
    rng = np.random.default_rng([seed, i])
    cond = rng.uniform(1.0, 12.0)
    family = 'geometric' if i % 2 == 0 else 'clustered'
    return cond, family, spectrum(n, rng, cond, family)


def generate(i0, i1, n, seed=SEED, device='cuda:0'):
    """Matrices i0 .. i1-1 of the workload, (i1 - i0, n, n) float32 on the device."""
    dev = torch.device(device)
    gens, sig = [], []
    for i in range(i0, i1):
        g = torch.Generator(device=dev)
        g.manual_seed(seed * 1000003 + i)
        gens.append(torch.randn(2, n, n, generator=g, device=dev, dtype=torch.float64))
        sig.append(meta(i, n, seed)[2])
    gm = torch.stack(gens)                                   # (b, 2, n, n)
    b = gm.shape[0]
    q, r = P.qr(gm.reshape(2 * b, n, n))
    q = q * torch.sign(torch.diagonal(r, dim1=1, dim2=2))[:, None, :]
    q = q.reshape(b, 2, n, n)
    s = torch.from_numpy(np.stack(sig)).to(dev)
    us = q[:, 0] * s[:, None, :]
    a = blas.gemm(us, q[:, 1].mT, mode='f64')
    return a.float().contiguous()


def targets_of(a):
    _, s, _ = pkg.svd(a.double(), 'jacobi', values_only=True, cut='floor')
    return torch.log10((s[:, -1] / s.square().sum(1).sqrt()).clamp_min(1e-16)).float()


def cmd_data(args):
    os.makedirs(args.out, exist_ok=True)
    feats, targ, conds = [], [], []
    t0 = time.time()
    stop = args.start + args.count
    for i0 in range(args.start, stop, CHUNK):
        i1 = min(i0 + CHUNK, stop)
        a = generate(i0, i1, args.n)
        feats.append(L.features(a).cpu().numpy())
        for j in range(i1 - i0):           # one matrix per solve: the batched iteration's count is the max
            targ.append(float(targets_of(a[j:j + 1])[0]))
        conds += [meta(i, args.n)[0] for i in range(i0, i1)]
        print(f'data {i1 - args.start}/{args.count} {time.time() - t0:.0f}s', flush=True)
    np.savez(os.path.join(args.out, f'l0_data_{args.start:05d}_{stop:05d}.npz'), features=np.concatenate(feats),
             targets=np.array(targ), conds=np.array(conds), n=args.n)


def load_data(path):
    """Join the l0_data_<start>_<stop>.npz slices of a directory (contiguous from 0)."""
    parts = sorted(f for f in os.listdir(path) if f.startswith('l0_data_') and f.endswith('.npz'))
    out, nxt, n = {'features': [], 'targets': [], 'conds': []}, 0, None
    for f in parts:
        start, stop = (int(x) for x in f[8:-4].split('_'))
        if start != nxt:
            raise ValueError(f'{path}: slices are not contiguous at {start}')
        d = np.load(os.path.join(path, f))
        for k in out:
            out[k].append(d[k])
        n, nxt = int(d['n']), stop
    return {k: np.concatenate(v) for k, v in out.items()}, n


def polar_time(a, model, orth_stop=None):
    """Polar phase of QDWH on one matrix (FP32 products): (seconds, iterations, QR-branch
    iterations, fixed-point error); with a model the time includes features and inference.
    orth_stop: the polished form's stop (|X^T X - I|_F <= orth_stop, then FP64 Newton-Schulz,
    whose three steps are included in the time)."""
    x0 = P.scale(a, P.reduce(a, 'fro'), invert=True)
    eps = S.EPS[torch.float32]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lower = eps if model is None else model.lower_bound(a, eps)
    x, it, nqr = S._qdwh_iterate(x0, lower, 10 * eps, 40, 'f32', 'f32', orth_stop)
    if orth_stop is not None:
        x = x.double()
        for _ in range(3):
            x = S._newton_schulz(x)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return dt, it, nqr, S._fixed_point_error(x, 'f64' if orth_stop is not None else 'f32')


def cmd_fit(args):
    d, n = load_data(args.data)
    f, t = torch.from_numpy(d['features']), torch.from_numpy(d['targets']).float()
    ntr = args.train - args.holdout
    model = L.L0Model(L.glorot(0))
    hist = model.fit(f[:ntr], t[:ntr], steps=args.steps, lr=args.lr, val=(f[ntr:args.train], t[ntr:args.train]))
    print('loss', hist, 'margin', model.margin, flush=True)
    pred = model.predict(f[args.train:]).cpu().numpy()
    err = pred - t[args.train:].numpy()
    res = {'n': n, 'train': ntr, 'holdout': args.holdout, 'test': int(len(f) - args.train), 'steps': args.steps,
           'loss_first_last': [hist[0][1], hist[-1][1]], 'margin': model.margin,
           'test_mean_abs_error_decades': float(np.mean(np.abs(err))),
           'test_fraction_bound_above_target': float(np.mean(err - model.margin > 0))}
    test = list(range(args.train, len(f)))[:args.eval]
    rows = []
    for i0 in range(0, len(test), CHUNK):
        idx = test[i0:i0 + CHUNK]
        a = generate(idx[0], idx[-1] + 1, n)
        for j in range(len(idx)):
            rows.append(tuple(polar_time(a[j:j + 1], mdl, stop) for stop in (None, 1e-2) for mdl in (None, model)))
        print(f'eval {len(rows)}/{len(test)}', flush=True)
    arr = np.array([[v for run in r for v in run] for r in rows])      # 4 runs x (s, it, qr, fpe)
    for k, name in enumerate(('reference_stop_eps', 'reference_stop_learned', 'orth_stop_eps', 'orth_stop_learned')):
        c = 4 * k
        res[name] = {'seconds': float(arr[:, c].mean()), 'iterations': float(arr[:, c + 1].mean()),
                     'qr_iterations': float(arr[:, c + 2].mean()), 'fixed_point_error_max': float(arr[:, c + 3].max())}
    res.update(eval_matrices=len(rows),
               learned_more_iterations_reference_stop=int(np.sum(arr[:, 5] > arr[:, 1])),
               learned_more_iterations_orth_stop=int(np.sum(arr[:, 13] > arr[:, 9])))
    print(json.dumps(res), flush=True)
    model.save(args.model, evaluation=res)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest='cmd', required=True)
    a = sub.add_parser('data')
    a.add_argument('--out', required=True)
    a.add_argument('--n', type=int, default=1024)
    a.add_argument('--start', type=int, default=0)
    a.add_argument('--count', type=int, default=2048)
    b = sub.add_parser('fit')
    b.add_argument('--data', required=True)
    b.add_argument('--train', type=int, default=1536)
    b.add_argument('--holdout', type=int, default=256)
    b.add_argument('--steps', type=int, default=4000)
    b.add_argument('--lr', type=float, default=3e-3)
    b.add_argument('--eval', type=int, default=512)
    b.add_argument('--model', default=L.DEFAULT_MODEL)
    args = ap.parse_args(argv)
    if args.cmd == 'data':
        cmd_data(args)
    else:
        cmd_fit(args)


if __name__ == '__main__':
    sys.exit(main())
