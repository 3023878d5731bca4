"""One measurement row in the reference harness's JSONL format for this package's solvers.

The reference harness (svdbench/cli.py) already runs these solvers through the plugin hook
``{"name": "pavel-s_svd_amd:factory", ...}``; this module produces the same row without it, for a
caller that drives the MI355X solvers directly: the timing contract of cli.py:86-105
(synchronise, wall clock and device events per sample, median / p10 / p90) and the row of
cli.py:108-183 (input copy, prepare, warm-ups, resident-solver timing, allocator memory, adapter
metadata, native diagnostics, output copy, metrics, end-to-end timing, warnings, status with
``unsupported`` for an ``Unsupported`` solver).  ``memory.native_workspace_bytes`` is filled from
the adapter -- 0, since the library takes all workspace from the caller through torch's allocator
-- and ``memory.library_workspace_bytes`` records what the last solve requested.  The metrics are
the caller's (``validate(a_cpu, host_result)``), as the harness computes them against its own FP64
reference; without one the row carries ``metrics: None``.
"""
import gc
import hashlib
import time
import warnings

import numpy as np
import torch

from .standalone.solvers import Unsupported, factory


def synchronize(device):
    if device.type == 'cuda':
        torch.cuda.synchronize(device)


def to_host(result):
    return tuple(None if x is None else x.detach().cpu().resolve_conj().resolve_neg().numpy().copy() for x in result)


def summary(samples):
    return {'samples_ms': samples, 'median_ms': float(np.median(samples)),
            'p10_ms': float(np.percentile(samples, 10)), 'p90_ms': float(np.percentile(samples, 90))}


def measure(fn, device, repeat, events=False):
    """Wall-clock (and device-event) samples of `repeat` synchronised calls (cli.py:86-105)."""
    wall, dev = [], []
    for _ in range(repeat):
        start = end = None
        if events and device.type == 'cuda':
            start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        synchronize(device)
        t = time.perf_counter()
        if start:
            start.record()
        result = fn()
        if end:
            end.record()
        synchronize(device)
        wall.append((time.perf_counter() - t) * 1e3)
        if start:
            dev.append(start.elapsed_time(end))
        del result
    return {'wall': summary(wall), 'device_event': summary(dev) if dev else None}


def prepare(a, spec):
    """The harness's checks before a plugin factory (algorithms.py:20-27: mode, rank), then this
    package's factory with the spec unchanged, as the plugin hook passes it (algorithms.py:90-93)."""
    mode = spec.get('mode', 'reduced')
    if mode not in ('reduced', 'full', 'values'):
        raise ValueError('mode must be reduced, full or values')
    rank = spec.get('rank', min(a.shape[-2:]))
    if not 1 <= rank <= min(a.shape[-2:]):
        raise ValueError('rank must be between 1 and min(m,n)')
    return factory(a, spec)


def measurement_row(a_cpu, case, spec, device, *, warmup=2, repeat=5, end_to_end=False, validate=None):
    """The harness row for solving the numpy matrix `a_cpu` with `spec` on `device` (cli.py:108-183)."""
    device = torch.device(device)
    row = {'type': 'measurement', 'case': case, 'algorithm': spec, 'device': str(device), 'status': 'error'}
    a = fn = result = None
    try:
        with warnings.catch_warnings(record=True) as caught, torch.inference_mode():
            warnings.simplefilter('always')
            synchronize(device)
            t = time.perf_counter()
            a = torch.from_numpy(a_cpu).to(device=device, copy=True)
            synchronize(device)
            row['initial_input_copy_ms'] = (time.perf_counter() - t) * 1e3
            t = time.perf_counter()
            fn = prepare(a, spec)
            row['adapter'] = fn.metadata
            synchronize(device)
            row['prepare_ms'] = (time.perf_counter() - t) * 1e3
            for _ in range(warmup):
                result = fn()
                synchronize(device)
                del result
                result = None
            gc.collect()
            if device.type == 'cuda':
                torch.cuda.reset_peak_memory_stats(device)
                baseline = torch.cuda.memory_allocated(device)
                reserved_baseline = torch.cuda.memory_reserved(device)
            row['resident_solver'] = measure(fn, device, repeat, events=True)
            native = getattr(fn, 'native_workspace_bytes', None)
            if device.type == 'cuda':
                row['memory'] = {
                    'input_bytes': a_cpu.nbytes,
                    'allocator_baseline_bytes': baseline,
                    'allocator_peak_allocated_bytes': torch.cuda.max_memory_allocated(device),
                    'allocator_peak_increment_bytes': torch.cuda.max_memory_allocated(device) - baseline,
                    'allocator_reserved_baseline_bytes': reserved_baseline,
                    'allocator_peak_reserved_bytes': torch.cuda.max_memory_reserved(device),
                    'native_workspace_bytes': native,
                    'library_workspace_bytes': getattr(fn, 'library_workspace_bytes', None),
                    'note': 'PyTorch allocator; the library allocates no device memory of its own, its workspace '
                            'is inside the allocator figures. Warm caches retained.',
                }
            else:
                row['memory'] = {'input_bytes': a_cpu.nbytes, 'native_workspace_bytes': native}
            result = fn()
            synchronize(device)
            row['native_diagnostics'] = fn.diagnostics
            t = time.perf_counter()
            host_result = to_host(result)
            synchronize(device)
            row['output_copy_ms'] = (time.perf_counter() - t) * 1e3
            row['memory']['output_bytes'] = sum(x.nbytes for x in host_result if x is not None)
            row['metrics'] = validate(a_cpu, host_result) if validate is not None else None
            del result
            result = None
            if end_to_end:
                def e2e():
                    x = torch.from_numpy(a_cpu).to(device=device, copy=True)
                    return to_host(prepare(x, spec)())
                row['end_to_end'] = measure(e2e, device, repeat)
            row['warnings'] = sorted({str(w.message) for w in caught})
            row['status'] = 'ok'
    except Unsupported as e:
        row.update(status='unsupported', error=str(e))
    except Exception as e:
        row.update(status='error', error=f'{type(e).__name__}: {e}')
    finally:
        del result, fn, a
        gc.collect()
    row['input_sha256'] = hashlib.sha256(np.ascontiguousarray(a_cpu).tobytes()).hexdigest()
    return row
