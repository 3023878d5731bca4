"""Harness rows on the GPU: the row of svdbench/cli.py:108-183 for this package's solvers carries every
field the reference's own GPU rows carry (tests/golden/row_schema.json, extracted from
results/standalone-hetzner.jsonl), reports native_workspace_bytes = 0 (the library allocates no device
memory; its workspace is inside the allocator figures) with the requested workspace beside it, and
its diagnostics and metrics agree with the reference's recorded row for the same corpus input."""
import json
from importlib import import_module

import numpy as np
import pytest

from conftest import GOLDEN, corpus, load_package
from oracle import metrics as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def H():
    load_package()
    return import_module('pavel-s_svd_amd.harness')


def _validate(a, out):
    return M.validate(a, out, M.reference(a))


@pytest.mark.parametrize('case,spec', [
    ('square32', {'method': 'jacobi'}),
    ('tall256x16', {'method': 'tsqr'}),
    ('square32', {'method': 'block_jacobi'}),
    ('rank_deficient', {'method': 'gram_jacobi'}),
    ('lowrank512', {'method': 'block_krylov', 'rank': 8, 'oversample': 8, 'iterations': 2, 'seed': 31}),
])
def test_rows_carry_the_reference_fields(H, expected, case, spec):
    schema = json.loads((GOLDEN / 'row_schema.json').read_text())
    a = corpus(case)
    spec = dict(spec, name='pavel-s_svd_amd:factory')
    row = H.measurement_row(a, {'name': case}, spec, 'cuda:0', warmup=1, repeat=3, end_to_end=True,
                            validate=_validate)
    assert row['status'] == 'ok', row.get('error')
    json.dumps(row, allow_nan=False)
    assert set(schema['row_keys']) <= set(row)
    assert set(schema['memory_keys']) <= set(row['memory'])
    assert set(schema['timing_keys']) <= set(row['resident_solver'])
    assert set(schema['summary_keys']) <= set(row['resident_solver']['wall'])
    assert row['memory']['native_workspace_bytes'] == 0
    assert row['memory']['library_workspace_bytes']['requests'] >= 1
    assert row['input_sha256'] == expected[case]['sha256']
    assert row['adapter']['rng_in_solve'] is False or 'initial_basis' not in spec
    runs = [r for r in expected[case]['runs'] if r['method'] == spec['method']
            and r['options'] == {k: v for k, v in spec.items() if k not in ('name', 'method')}]
    assert runs
    for run in runs:
        for key, want in run['diagnostics'].items():
            if key in ('jacobi_sweeps', 'tsqr_leaves', 'block_sweeps', 'eigen_sweeps', 'gram_discarded_values',
                       'sketch_columns', 'range_columns', 'power_iterations'):
                slack = 1 if key == 'jacobi_sweeps' else 0
                if spec['method'] == 'block_krylov' and key == 'jacobi_sweeps':
                    continue   # the sketch here comes from the ROCm device's generator, not the row's
                assert abs(row['native_diagnostics'][key] - want) <= slack, (key, run['source'])
        floor = 64 * np.finfo(a.dtype).eps * max(a.shape[-2:]) ** 0.5
        worst = row['metrics']['worst']
        for key in ('spectrum_error_over_sigma1', 'orthogonality_u_fro_over_sqrt_cols',
                    'reconstruction_relative_fro'):
            assert worst[key] <= max(3 * run['metrics'][key], floor), (key, worst[key], run['source'])


def test_unsupported_rows(H):
    a = corpus('square32')
    row = H.measurement_row(a, {'name': 'square32'}, {'name': 'pavel-s_svd_amd:factory', 'method': 'lanczos'},
                            'cuda:0', warmup=0, repeat=1)
    assert row['status'] == 'unsupported'
    c = corpus('clustered_complex')
    row = H.measurement_row(c, {'name': 'clustered_complex'}, {'name': 'pavel-s_svd_amd:factory',
                                                               'method': 'bidiag_dc'}, 'cuda:0', warmup=0, repeat=1)
    assert row['status'] == 'unsupported'
    for method in ('jacobi', 'qr', 'tsqr', 'polar_qdwh', 'block_jacobi'):
        row = H.measurement_row(c, {'name': 'clustered_complex'}, {'name': 'pavel-s_svd_amd:factory',
                                                                   'method': method}, 'cuda:0', warmup=0, repeat=1)
        assert row['status'] == 'ok', (method, row.get('error'))
    json.dumps(row, allow_nan=False)
