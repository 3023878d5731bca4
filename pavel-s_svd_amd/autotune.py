"""Autotuner for the batched Jacobi launch shape.

Tunes, on the actual batch, the launch-shape policy -- the waves per CU the resident matrices
should supply (which sets each matrix's workgroup size from its LDS footprint) and the minimum
rows per lane that bounds the lanes per column pair; together they fix the size classes the
batch is split into -- and then, class by class, the workgroup size of each class.  Three
stages, all on captured-graph replays (the path a caller replays) timed with HIP events on the
current stream: every policy candidate over a few replays; the fastest few again in interleaved
rounds, ranked by their median replay time (single short timings differ by a few percent from
run to run, as much as the candidates differ); then each class, largest first, at twice and half
its workgroup size, a change kept when it beats the current setting, timed in alternation with
it, by more than 3 %; last, the pair schedule -- the pair ordering (the reference's tournament, or
the tournament over columns relabelled in descending norm order every sweep) and the order in which a
round's pairs are served (identity, or spread over the LDS banks) -- four combinations timed in
alternation with their sweep counts recorded, one kept when it beats the reference's schedule by
more than 2 %.
Results are cached in-process per (dtype, max_m, max_n, batch size bucket, vectors) and, when
``PAVELSVD_TUNE_CACHE`` names a JSON file, across processes.
"""
import json
import os
import sys

import torch

CANDIDATES = [dict(target_waves=w, min_rows=r, passes=p, streams=4) for w in (8, 16, 32) for r in (2, 4, 6, 8)
              for p in (1, 2)]
_CACHE = {}


def _key(batch):
    bucket = 1 << max(batch.batch - 1, 0).bit_length()
    return f'{batch.code}:{batch.max_m}:{batch.max_n}:{bucket}:{int(batch.vectors)}:v2'


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


def time_replays(batch, reps=8):
    """Per-replay times (ms) of the batch's captured graph, after one warm replay."""
    stream = torch.cuda.current_stream(batch.a.device)
    batch.capture()
    batch.replay()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for e0, e1 in ev:
        e0.record(stream)
        batch.replay()
        e1.record(stream)
    ev[-1][1].synchronize()
    return [e0.elapsed_time(e1) for e0, e1 in ev]


def _median_pair(batch, a, b, rounds):
    """Replay medians of configurations a and b, timed in alternation."""
    sa, sb = [], []
    for _ in range(rounds):
        batch.configure(**a)
        sa += time_replays(batch)
        batch.configure(**b)
        sb += time_replays(batch)
    return sorted(sa)[len(sa) // 2], sorted(sb)[len(sb) // 2]


def _tune_classes(batch, policy, rounds, verbose, key, gain=0.97):
    """Second stage: with the policy fixed, try each size class (largest cost first) at twice and
    half its workgroup size; keep a change when its replay median beats the current one by more
    than 1 - gain (the two timed in alternation).  Returns the class_threads map and its median."""
    batch.configure(**policy, class_threads={})
    cur = {}
    base = None
    for cl in sorted(batch.classes, key=lambda c: -c['cost']):
        for nt in (cl['threads'] * 2, cl['threads'] // 2):
            if nt < 64 or nt > 1024:
                continue
            trial = dict(cur, **{cl['name']: nt})
            m_cur, m_trial = _median_pair(batch, dict(policy, class_threads=cur), dict(policy, class_threads=trial),
                                          rounds)
            if verbose:
                print(f'autotune {key} class {cl["name"]} -> {nt}: {m_trial:.3f} vs {m_cur:.3f} ms', file=sys.stderr,
                      flush=True)
            base = m_cur if base is None else base
            if m_trial < gain * m_cur:
                cur, base = trial, m_trial
                break
    return cur, base


SCHEDULES = [dict(ordering=o, slots=sl) for o in ('tournament', 'sorted') for sl in ('identity', 'spread')]


def _tune_schedule(batch, params, rounds, verbose, key, gain=0.98, slot_gain=0.995):
    """Third stage: with the launch shape fixed, the pair ordering (the reference's tournament or
    the norm-sorted one) and the slot order of a round's pairs, each timed in alternation with the
    current best over `rounds` rounds, with the batch's sweep counts recorded.  Another ordering
    replaces the reference's tournament only when its median beats it by more than 1 - gain; within
    the chosen ordering the slot order is taken at a margin of 1 - slot_gain."""
    rows, medians = [], {}
    samples = {i: [] for i in range(len(SCHEDULES))}
    for _ in range(rounds):
        for i, sched in enumerate(SCHEDULES):
            batch.configure(**params, **sched)
            samples[i] += time_replays(batch)
    for i, sched in enumerate(SCHEDULES):
        batch.configure(**params, **sched)
        batch.solve()
        sweeps = batch.info[:, 0].float()
        medians[i] = sorted(samples[i])[len(samples[i]) // 2]
        rows.append(dict(sched, median_ms=medians[i], sweeps_mean=float(sweeps.mean()), sweeps_max=int(sweeps.max())))
        if verbose:
            print(f'autotune {key} schedule {sched}: {medians[i]:.3f} ms, sweeps {rows[-1]["sweeps_mean"]:.2f}',
                  file=sys.stderr, flush=True)
    # the ordering changes the rotations (and the sweep counts): it replaces the reference's only by
    # more than 1 - gain; the slot order only changes which lanes serve a pair (bit-identical
    # results), so within the chosen ordering the faster slot order is taken above a 0.5 % margin
    best = 0
    for i in range(1, len(SCHEDULES)):
        if SCHEDULES[i]['ordering'] != SCHEDULES[best]['ordering'] and medians[i] < gain * medians[best]:
            best = i
    for i in range(len(SCHEDULES)):
        if SCHEDULES[i]['ordering'] == SCHEDULES[best]['ordering'] and medians[i] < slot_gain * medians[best]:
            best = i
    return dict(SCHEDULES[best]), medians[best], rows


def tune_uneven(batch, candidates=None, reps=3, verbose=None, finalists=4, rounds=2, class_stage=True,
                schedule_stage=True):
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
        t = sorted(time_replays(batch, reps + 1))
        ms = t[len(t) // 2]
        timings.append((ms, cand))
        if verbose:
            print(f'autotune {key} {cand}: {ms:.3f} ms', file=sys.stderr, flush=True)
    timings.sort(key=lambda t: t[0])
    top = [c for _, c in timings[:max(1, finalists)]]
    samples = [[] for _ in top]
    for _ in range(rounds):
        for i, cand in enumerate(top):
            batch.configure(**cand)
            samples[i] += time_replays(batch)
    medians = [sorted(x)[len(x) // 2] for x in samples]
    i_best = min(range(len(top)), key=lambda i: medians[i])
    best, ms = dict(top[i_best]), medians[i_best]
    if verbose:
        for cand, med in zip(top, medians):
            print(f'autotune {key} replay {cand}: median {med:.3f} ms', file=sys.stderr, flush=True)
    if class_stage:
        best['class_threads'], ms = _tune_classes(batch, best, rounds, verbose, key)
    schedules = None
    if schedule_stage:
        sched, ms, schedules = _tune_schedule(batch, best, rounds, verbose, key)
        best.update(sched)
    batch.configure(**{k: v for k, v in saved.items() if k not in best}, **best)
    _CACHE[key] = {'params': best, 'ms': ms, 'timings': [[t, c] for t, c in timings],
                   'replay_medians': [[med, c] for med, c in zip(medians, top)], 'schedules': schedules}
    _store()
    return dict(best)


def report(batch):
    """The cached tuning record of ``batch``'s key (timings of every stage), or None."""
    return _CACHE.get(_key(batch))
