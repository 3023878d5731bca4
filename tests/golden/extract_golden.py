"""Write the golden fixtures for the standalone hot path from the reference's own data files.

Reads (as data only) the committed corpus inputs ``corpus/standalone/*.npy`` and the measurement
rows of ``results/standalone-{cpu-hetzner,hetzner,scale-hetzner}.jsonl`` of the reference tree and
writes:

* ``tests/golden/corpus/<case>.npy`` -- the exact input bytes (numpy, ``allow_pickle=False``);
* ``tests/golden/row_schema.json``   -- the field names every successful GPU measurement row of
  ``results/standalone-hetzner.jsonl`` carries (row, ``memory``, timing summary);
* ``tests/golden/expected.json``      -- per case: shape, dtype, sha256 of the input, and for each
  of ``jacobi``, ``tsqr``, ``polar_qdwh`` (the hot path) and ``block_jacobi``, ``gram_jacobi``,
  ``randomized``, ``subspace``, ``block_krylov`` (its consumers) the solver options of the row, the
  recorded counts (``jacobi_sweeps``, ``tsqr_leaves``, ``polar_iterations``, ``eigen_sweeps``,
  ``polar_fixed_point_error``, ``block_sweeps``, ``gram_discarded_values``, ``sketch_columns``,
  ``range_columns``, ``power_iterations``) and the worst-case accuracy metrics, for every device the
  reference recorded (``cpu`` / ``cuda:0``).

Run once in a container where the reference tree is mounted::

    python tests/golden/extract_golden.py /root/reference
"""
import hashlib
import json
import shutil
import sys
from pathlib import Path

import numpy as np

HOT = ('jacobi', 'tsqr', 'polar_qdwh', 'block_jacobi', 'gram_jacobi', 'randomized', 'subspace', 'block_krylov')
KEYS = ('jacobi_sweeps', 'tsqr_leaves', 'polar_iterations', 'eigen_sweeps', 'polar_fixed_point_error',
        'block_sweeps', 'gram_discarded_values', 'sketch_columns', 'range_columns', 'power_iterations')
FILES = ('standalone-cpu-hetzner.jsonl', 'standalone-hetzner.jsonl', 'standalone-scale-hetzner.jsonl')
CASES = ('square32', 'square64', 'batch8x16', 'tall256x16', 'tall2048x32', 'condition1e8',
         'condition1e12', 'rank_deficient', 'zero', 'clustered_complex', 'wide_complex', 'lowrank512')


def main(ref):
    ref = Path(ref)
    here = Path(__file__).resolve().parent
    (here / 'corpus').mkdir(exist_ok=True)
    out = {}
    for case in CASES:
        src = ref / 'corpus' / 'standalone' / f'{case}.npy'
        a = np.load(src, allow_pickle=False)
        shutil.copyfile(src, here / 'corpus' / f'{case}.npy')
        out[case] = {'shape': list(a.shape), 'dtype': str(a.dtype),
                     'sha256': hashlib.sha256(a.tobytes(order='C')).hexdigest(), 'runs': []}
    for name in FILES:
        for line_no, line in enumerate((ref / 'results' / name).read_text().splitlines(), 1):
            row = json.loads(line)
            if row.get('type') != 'measurement' or row.get('status') != 'ok':
                continue
            method = row['algorithm'].get('method')
            case = row['case']['name']
            if method not in HOT or case not in out:
                continue
            if row['input_sha256'] != out[case]['sha256']:
                raise SystemExit(f'{name}:{line_no}: input hash does not match the corpus file')
            diag = {k: v for k, v in row['native_diagnostics'].items() if k in KEYS}
            options = {k: v for k, v in row['algorithm'].items() if k not in ('name', 'method', 'mode')}
            out[case]['runs'].append({'source': f'results/{name}:{line_no}', 'method': method,
                                      'device': row['device'], 'case': row['case'], 'options': options,
                                      'diagnostics': diag, 'metrics': row['metrics']['worst']})
    (here / 'expected.json').write_text(json.dumps(out, indent=1, sort_keys=True) + '\n')
    rows = [json.loads(line) for line in (ref / 'results' / FILES[1]).read_text().splitlines()]
    ok = [r for r in rows if r.get('type') == 'measurement' and r.get('status') == 'ok']
    schema = {'source': f'results/{FILES[1]}', 'rows': len(ok),
              'row_keys': sorted(set.intersection(*(set(r) for r in ok))),
              'memory_keys': sorted(set.intersection(*(set(r['memory']) for r in ok))),
              'timing_keys': sorted(set.intersection(*(set(r['resident_solver']) for r in ok))),
              'summary_keys': sorted(set.intersection(*(set(r['resident_solver']['wall']) for r in ok)))}
    (here / 'row_schema.json').write_text(json.dumps(schema, indent=1) + '\n')


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
