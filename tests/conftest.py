import importlib
import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / 'tests' / 'golden'
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (gfx950) GPU and the built HIP library')


def load_package():
    """The product package (its directory name is not a Python identifier)."""
    return importlib.import_module('pavel-s_svd_amd')


@pytest.fixture(scope='session')
def expected():
    return json.loads((GOLDEN / 'expected.json').read_text())


def corpus(case):
    return np.load(GOLDEN / 'corpus' / f'{case}.npy', allow_pickle=False)
