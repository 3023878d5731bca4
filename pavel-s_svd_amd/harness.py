"""Measurement rows in the reference harness's JSONL wire format, written by this package.

The reference's own CLI (svdbench/cli.py) drives these solvers unchanged through its plugin hook
``{"name": "pavel-s_svd_amd:factory", ...}``.  This module is for a caller that drives the MI355X
solvers directly and still wants rows that the reference's report tools read: the field names and
their meaning follow the reference's rows (cli.py:108-183, pinned by the row schema extracted from
its committed results); the code is organised around this package's own needs.

Design: a row is produced by a fixed pipeline of stages (upload, prepare, warm-up, resident timing,
checked solve, download, optional end-to-end timing).  Each stage adds its fields to the row; the
first exception ends the pipeline and sets the row's status (``unsupported`` for ``Unsupported``,
``error`` otherwise), so a partially measured row still carries the stages that completed.  Timing
uses one clock object that drains the device before and after every sample, reads the host wall
clock, and brackets the call with HIP events on torch's current stream.  The memory block reports
the caching allocator's figures and the library's own device allocation, which is always 0: every
workspace of ``libpavelsvd.so`` is a caller tensor from the allocator
(``memory.native_workspace_bytes``; the reference hardcodes None there, cli.py:145), with the last
solve's requests beside it (``memory.library_workspace_bytes``).
"""
import gc
import hashlib
import time
import warnings

import numpy as np
import torch

from .standalone.solvers import Unsupported, factory

_QUANTILES = (0.5, 0.1, 0.9)   # median, p10, p90 with numpy's linear interpolation


def _stats(samples):
    """{'samples_ms', 'median_ms', 'p10_ms', 'p90_ms'} of a list of millisecond samples."""
    med, lo, hi = (float(q) for q in np.quantile(np.asarray(samples, dtype=np.float64), _QUANTILES))
    return {'samples_ms': list(samples), 'median_ms': med, 'p10_ms': lo, 'p90_ms': hi}


class _Clock:
    """Synchronised samples on one device: host wall time and, on a GPU, HIP event time."""

    def __init__(self, device):
        self.device = device
        self.gpu = device.type == 'cuda'

    def drain(self):
        if self.gpu:
            torch.cuda.synchronize(self.device)

    def sample(self, fn, events):
        """One call of fn: (wall ms, event ms or None).  The result is dropped before returning, so no
        two samples' outputs are alive together (the allocator peak is one solve's)."""
        marks = [torch.cuda.Event(enable_timing=True) for _ in range(2)] if events and self.gpu else None
        self.drain()
        t0 = time.perf_counter()
        if marks:
            marks[0].record()
        out = fn()
        if marks:
            marks[1].record()
        self.drain()
        wall = (time.perf_counter() - t0) * 1e3
        del out
        return wall, (marks[0].elapsed_time(marks[1]) if marks else None)

    def series(self, fn, repeat, events=False):
        walls, evs = [], []
        for _ in range(repeat):
            w, e = self.sample(fn, events)
            walls.append(w)
            if e is not None:
                evs.append(e)
        return {'wall': _stats(walls), 'device_event': _stats(evs) if evs else None}

    def stopwatch(self, fn):
        """(result, ms) of one drained call whose result is kept."""
        self.drain()
        t0 = time.perf_counter()
        out = fn()
        self.drain()
        return out, (time.perf_counter() - t0) * 1e3


def _host_copy(outputs):
    """numpy copies of the solver's (U, s, Vh); conjugate and negative views are materialised."""
    return tuple(None if t is None else t.detach().cpu().resolve_conj().resolve_neg().numpy().copy()
                 for t in outputs)


def prepare(a, spec):
    """The harness's checks before a plugin factory (svdbench/algorithms.py:20-27: mode, rank), then
    this package's factory with the spec unchanged, as the plugin hook passes it (algorithms.py:90-93)."""
    mode = spec.get('mode', 'reduced')
    if mode not in ('reduced', 'full', 'values'):
        raise ValueError('mode must be reduced, full or values')
    rank = spec.get('rank', min(a.shape[-2:]))
    if not 1 <= rank <= min(a.shape[-2:]):
        raise ValueError('rank must be between 1 and min(m,n)')
    return factory(a, spec)


class _RowPipeline:
    """State shared by the stages of one row."""

    def __init__(self, a_cpu, spec, device, warmup, repeat, end_to_end, validate):
        self.a_cpu, self.spec, self.device = a_cpu, spec, device
        self.warmup, self.repeat, self.want_e2e, self.validate = warmup, repeat, end_to_end, validate
        self.clock = _Clock(device)
        self.a = self.solver = None

    # -- stages: each adds its fields to the row ------------------------------------------------------
    def upload(self, row):
        self.a, row['initial_input_copy_ms'] = self.clock.stopwatch(
            lambda: torch.from_numpy(self.a_cpu).to(device=self.device, copy=True))

    def build(self, row):
        self.solver, ms = self.clock.stopwatch(lambda: prepare(self.a, self.spec))
        row['adapter'] = self.solver.metadata
        row['prepare_ms'] = ms

    def warm(self, row):
        for _ in range(self.warmup):
            self.clock.sample(self.solver, events=False)
        gc.collect()

    def resident(self, row):
        before = self._allocator() if self.clock.gpu else None
        row['resident_solver'] = self.clock.series(self.solver, self.repeat, events=True)
        row['memory'] = self._memory(before)

    def checked_solve(self, row):
        outputs, _ = self.clock.stopwatch(self.solver)
        row['native_diagnostics'] = self.solver.diagnostics
        host, row['output_copy_ms'] = self.clock.stopwatch(lambda: _host_copy(outputs))
        del outputs
        row['memory']['output_bytes'] = sum(x.nbytes for x in host if x is not None)
        row['metrics'] = None if self.validate is None else self.validate(self.a_cpu, host)

    def end_to_end(self, row):
        if not self.want_e2e:
            return

        def full_call():
            x = torch.from_numpy(self.a_cpu).to(device=self.device, copy=True)
            return _host_copy(prepare(x, self.spec)())
        row['end_to_end'] = self.clock.series(full_call, self.repeat)

    STAGES = ('upload', 'build', 'warm', 'resident', 'checked_solve', 'end_to_end')

    # -- memory block ---------------------------------------------------------------------------------
    def _allocator(self):
        torch.cuda.reset_peak_memory_stats(self.device)
        return torch.cuda.memory_allocated(self.device), torch.cuda.memory_reserved(self.device)

    def _memory(self, before):
        mem = {'input_bytes': self.a_cpu.nbytes,
               'native_workspace_bytes': getattr(self.solver, 'native_workspace_bytes', None)}
        if before is None:
            return mem
        allocated, reserved = before
        peak = torch.cuda.max_memory_allocated(self.device)
        mem.update(allocator_baseline_bytes=allocated, allocator_peak_allocated_bytes=peak,
                   allocator_peak_increment_bytes=peak - allocated, allocator_reserved_baseline_bytes=reserved,
                   allocator_peak_reserved_bytes=torch.cuda.max_memory_reserved(self.device),
                   library_workspace_bytes=getattr(self.solver, 'library_workspace_bytes', None),
                   note='PyTorch caching allocator; libpavelsvd.so allocates no device memory, its workspace is '
                        'requested from the allocator and included above. Warm caches retained.')
        return mem

    def release(self):
        self.a = self.solver = None
        gc.collect()


def measurement_row(a_cpu, case, spec, device, *, warmup=2, repeat=5, end_to_end=False, validate=None):
    """One harness row for solving the numpy matrix `a_cpu` with `spec` on `device`.

    Fields as the reference's rows (cli.py:108-183): input copy, prepare, warm-ups, resident-solver
    timing (wall and device events, median / p10 / p90), allocator memory, adapter metadata, native
    diagnostics, output copy, metrics from ``validate(a_cpu, host_outputs)`` (None without one),
    optional end-to-end timing, warnings, status, input digest."""
    device = torch.device(device)
    row = {'type': 'measurement', 'case': case, 'algorithm': spec, 'device': str(device), 'status': 'error'}
    job = _RowPipeline(a_cpu, spec, device, warmup, repeat, end_to_end, validate)
    try:
        with warnings.catch_warnings(record=True) as caught, torch.inference_mode():
            warnings.simplefilter('always')
            for stage in job.STAGES:
                getattr(job, stage)(row)
            row['warnings'] = sorted({str(w.message) for w in caught})
        row['status'] = 'ok'
    except Unsupported as exc:
        row['status'], row['error'] = 'unsupported', str(exc)
    except Exception as exc:   # the row records the failure instead of raising, as the harness does
        row['status'], row['error'] = 'error', f'{type(exc).__name__}: {exc}'
    finally:
        job.release()
    row['input_sha256'] = hashlib.sha256(np.ascontiguousarray(a_cpu).tobytes()).hexdigest()
    return row
