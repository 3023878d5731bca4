"""The host-side helpers of the primitives module that the reference's callers import by name
(svdbench/standalone/primitives.py: _round_robin :309-313, _eps), checked on the CPU."""
from importlib import import_module

import pytest
import torch

from conftest import load_package
from oracle import restatement as R


@pytest.fixture(scope='module')
def P():
    load_package()
    return import_module('pavel-s_svd_amd.standalone.primitives')


@pytest.mark.parametrize('n', [2, 4, 8, 14])
def test_round_robin_equals_the_pinned_oracle_schedule(P, n):
    assert list(P._round_robin(n)) == list(R.round_robin(n))


def test_eps_and_real_dtype(P):
    assert P._eps(torch.float32) == 2.0 ** -23 and P._eps(torch.complex128) == 2.0 ** -52
    assert P._real_dtype(torch.complex64) == torch.float32 and P._real_dtype(torch.float64) == torch.float64
