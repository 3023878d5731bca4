"""Autotuner for the batched Jacobi launch shape.

Tunes, on the actual batch, the launch-shape policy -- the waves per CU the resident matrices
should supply (which sets each matrix's workgroup size from its LDS footprint) and the minimum
rows per lane that bounds the lanes per column pair; together they fix the size classes the
batch is split into -- by timing each candidate solve with HIP events on the current stream.  Results are cached in-process per (dtype, max_m, max_n, batch
size bucket, vectors) and, when ``PAVELSVD_TUNE_CACHE`` names a JSON file, across processes.
"""
import json
import os
import sys

import torch

CANDIDATES = [dict(target_waves=w, min_rows=r, passes=p, streams=4) for w in (8, 16, 32) for r in (2, 4, 6)
              for p in (1, 2)]
_CACHE = {}


def _key(batch):
    bucket = 1 << max(batch.batch - 1, 0).bit_length()
    return f'{batch.code}:{batch.max_m}:{batch.max_n}:{bucket}:{int(batch.vectors)}'


def _load():
    path = os.environ.get('PAVELSVD_TUNE_CACHE')
    if path and os.path.exists(path):
        with open(path) as f:
            _CACHE.update(json.load(f))


def _store():
    path = os.environ.get('PAVELSVD_TUNE_CACHE')
    if path:
        with open(path, 'w') as f:
            json.dump(_CACHE, f, indent=1, sort_keys=True)


def time_launch(batch, reps=3):
    stream = torch.cuda.current_stream(batch.a.device)
    batch.solve()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        batch.solve()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_uneven(batch, candidates=None, reps=3, verbose=None):
    """Return the fastest launch parameters for ``batch`` (an UnevenBatch) and leave the batch
    configured with them."""
    if verbose is None:
        verbose = bool(os.environ.get('PAVELSVD_TUNE_VERBOSE'))
    if not _CACHE:
        _load()
    key = _key(batch)
    if key in _CACHE:
        best = _CACHE[key]['params']
        batch.configure(**best)
        return dict(best)
    saved = dict(batch.params)
    timings = []
    for cand in candidates or CANDIDATES:
        batch.configure(**cand)
        ms = time_launch(batch, reps)
        timings.append((ms, cand))
        if verbose:
            print(f'autotune {key} {cand}: {ms:.3f} ms', file=sys.stderr, flush=True)
    ms, best = min(timings, key=lambda t: t[0])
    batch.configure(**{k: v for k, v in saved.items() if k not in best}, **best)
    _CACHE[key] = {'params': best, 'ms': ms, 'timings': [[t, c] for t, c in timings]}
    _store()
    return dict(best)
