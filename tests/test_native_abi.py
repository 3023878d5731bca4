"""CPU checks of the C ABI: the library loads, exports every function declared in
include/pavelsvd.h, and the binding's argument plumbing reaches the C entry points (exercised with
arguments the library rejects before any GPU call)."""
import ast
import ctypes
import re
from pathlib import Path

import pytest

from conftest import ROOT, load_package

HEADER = ROOT / 'include' / 'pavelsvd.h'
LIB = ROOT / 'pavel-s_svd_amd' / 'libpavelsvd.so'


def declared_functions():
    text = re.sub(r'/\*.*?\*/', '', HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r'\b(psvd_[a-z0-9_]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    names = declared_functions()
    assert 'psvd_jacobi_lds' in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_queries_without_gpu():
    lib = ctypes.CDLL(str(LIB))
    lib.psvd_version.restype = ctypes.c_char_p
    assert lib.psvd_version().startswith(b'pavelsvd')
    N = __import__('importlib').import_module('pavel-s_svd_amd.native')
    st, b = N._hip.jacobi_lds_smem_bytes(N.F32, 96, 96, 0)
    assert st == N.OK and 0 < b < 160 * 1024
    assert N._hip.jacobi_lds_smem_bytes(N.F32, 2048, 32, 0)[0] == N.ETOOBIG
    assert N._hip.jacobi_lds_smem_bytes(N.F32, 4, 8, 0)[0] == N.EINVAL     # wide is rejected


def test_binding_argument_plumbing_reaches_c():
    N = __import__('importlib').import_module('pavel-s_svd_amd.native')
    # batch < 0 and a NULL workspace are rejected by the C entry point itself
    assert N._hip.jacobi_lds(N.F32, -1, 4, 4, 0, 0, 0, 4, 4, 0, 0, 0, 0, 0, 0.0, 60, 0, 0, 256, 0, 0, 0) == N.EINVAL
    assert N._hip.jacobi_lds(N.F32, 3, 4, 4, 0, 0, 0, 4, 4, 8, 8, 8, 8, 8, 0.0, 60, 0, 0, 256, 0, 0, 0) == N.EINVAL


def test_product_package_has_no_cpu_fallback_and_no_oracle_import():
    pkg = load_package()
    import torch
    with pytest.raises(pkg.Unsupported):
        pkg.svd(torch.randn(6, 4), 'jacobi')
    for path in (ROOT / 'pavel-s_svd_amd').rglob('*.py'):
        tree = ast.parse(path.read_text())
        for node in ast.walk(tree):
            if isinstance(node, (ast.Import, ast.ImportFrom)):
                names = [a.name for a in node.names] + [getattr(node, 'module', '') or '']
                assert not any(n.split('.')[0] == 'oracle' for n in names), path


def test_standalone_package_audit_like_reference():
    """The reference's AST audit (tests/test_standalone.py:157-166): no numpy, scipy or ctypes
    imports and no `.linalg` attribute in the standalone solver package."""
    for path in (ROOT / 'pavel-s_svd_amd' / 'standalone').glob('*.py'):
        tree = ast.parse(path.read_text())
        for node in ast.walk(tree):
            if isinstance(node, ast.Import):
                assert all(n.name not in ('numpy', 'scipy', 'ctypes') for n in node.names), path
            if isinstance(node, ast.Attribute):
                assert node.attr != 'linalg', path


def test_argument_validation_order_matches_reference():
    pkg = load_package()
    import torch
    a = torch.randn(12, 9, dtype=torch.float64)
    for kw in ({'rank': 0}, {'iterations': -1}, {'oversample': -1}, {'leaf_size': 0}, {'max_sweeps': 0}):
        with pytest.raises(ValueError):
            pkg.svd(a, **kw)
    with pytest.raises(ValueError):
        pkg.svd(a * float('nan'))
    with pytest.raises(TypeError):
        pkg.svd(a.to(torch.int32))
    with pytest.raises(ValueError):
        pkg.svd(a, 'not_an_algorithm')
    with pytest.raises(ValueError):
        pkg.svd(torch.empty(0, 3))
