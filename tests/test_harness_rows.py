"""The harness row writer on the CPU: the reference's checks before a plugin factory
(svdbench/algorithms.py:20-27), an unsupported device giving an ``unsupported`` row rather than a CPU
solve (cli.py:176-177), the input digest the reference records (cli.py:247-248), JSON without NaN."""
import hashlib
import json
from importlib import import_module

import numpy as np
import pytest
import torch

from conftest import corpus, load_package


@pytest.fixture(scope='module')
def H():
    load_package()
    return import_module('pavel-s_svd_amd.harness')


def test_cpu_device_row_is_unsupported_with_the_reference_digest(H, expected):
    a = corpus('square32')
    row = H.measurement_row(a, {'name': 'square32', 'm': 32, 'n': 32}, {'name': 'pavel-s_svd_amd:factory',
                                                                       'method': 'jacobi'}, 'cpu')
    assert row['status'] == 'unsupported' and 'ROCm' in row['error']
    assert row['input_sha256'] == expected['square32']['sha256']
    assert row['type'] == 'measurement' and row['device'] == 'cpu'
    json.dumps(row, allow_nan=False)


def test_prepare_checks_mode_and_rank(H):
    a = torch.zeros(1, 6, 4)
    with pytest.raises(ValueError):
        H.prepare(a, {'name': 'x:factory', 'method': 'jacobi', 'mode': 'economy'})
    with pytest.raises(ValueError):
        H.prepare(a, {'name': 'x:factory', 'method': 'jacobi', 'rank': 5})
    run = H.prepare(a, {'name': 'x:factory', 'method': 'randomized', 'rank': 2})
    assert run.metadata['rng_in_solve'] is True
    assert H.prepare(a, {'method': 'randomized', 'initial_basis': None})   # plain dict specs pass through
    row = H.measurement_row(np.zeros((6, 4), np.float32), {'name': 'z'}, {'method': 'jacobi', 'mode': 'bad'}, 'cpu')
    assert row['status'] == 'error' and row['error'].startswith('ValueError')
    assert row['input_sha256'] == hashlib.sha256(np.zeros((6, 4), np.float32).tobytes()).hexdigest()
