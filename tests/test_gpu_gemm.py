"""GPU parity of the MFMA GEMM (FP32, BF16x3, FP64 modes) against FP64 numpy products.

Bars, on max |C - C_ref| / sqrt(k) over all entries for N(0,1) operands: FP32 mode
2e-6 * (1 + sqrt(k) / 8) (FP32 accumulation of k products, worst entry of up to 2e4), BF16x3
6e-5 (the three-term bf16 split keeps about 16 significant bits per operand), FP64 1e-13.
"""
from importlib import import_module

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'


@pytest.fixture(scope='module')
def blas():
    return import_module('pavel-s_svd_amd.standalone.blas')


def _ref(a, b, ta, tb):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    a = np.swapaxes(a, -1, -2) if ta else a
    b = np.swapaxes(b, -1, -2) if tb else b
    return a @ b


def tol(mode, k):
    return {'f32': 2e-6 * (1 + np.sqrt(k) / 8), 'bf16x3': 6e-5, 'f64': 1e-13}[mode]


@pytest.mark.parametrize('mode', ['f32', 'bf16x3', 'f64'])
@pytest.mark.parametrize('m,n,k', [(1, 1, 1), (17, 33, 5), (128, 128, 32), (130, 257, 70), (300, 64, 513)])
@pytest.mark.parametrize('ta,tb', [(False, False), (True, False), (False, True), (True, True)])
def test_gemm_matches_fp64(blas, mode, m, n, k, ta, tb):
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    dt = np.float64 if mode == 'f64' else np.float32
    a = rng.standard_normal((k, m) if ta else (m, k)).astype(dt)
    b = rng.standard_normal((n, k) if tb else (k, n)).astype(dt)
    at = torch.from_numpy(a).to(DEV)
    bt = torch.from_numpy(b).to(DEV)
    out = blas.gemm(at.mT if ta else at, bt.mT if tb else bt, mode=mode)
    ref = _ref(a, b, ta, tb)
    scale = np.sqrt(k) * 1.0
    err = np.max(np.abs(out.cpu().numpy() - ref)) / scale
    assert err <= tol(mode, k), err


@pytest.mark.parametrize('mode', ['f32', 'bf16x3', 'f64'])
def test_batched_alpha_beta(blas, mode):
    rng = np.random.default_rng(5)
    dt = np.float64 if mode == 'f64' else np.float32
    a = rng.standard_normal((3, 40, 24)).astype(dt)
    b = rng.standard_normal((3, 24, 50)).astype(dt)
    c = rng.standard_normal((3, 40, 50)).astype(dt)
    ct = torch.from_numpy(c).to(DEV)
    blas.gemm(torch.from_numpy(a).to(DEV), torch.from_numpy(b).to(DEV), alpha=-0.5, beta=2.0, out=ct, mode=mode)
    ref = -0.5 * (a.astype(np.float64) @ b) + 2.0 * c
    assert np.max(np.abs(ct.cpu().numpy() - ref)) / np.sqrt(24) <= 2 * tol(mode, 24)


def test_broadcast_operand(blas):
    rng = np.random.default_rng(6)
    a = rng.standard_normal((4, 30, 20)).astype(np.float32)
    b = rng.standard_normal((20, 10)).astype(np.float32)
    out = blas.gemm(torch.from_numpy(a).to(DEV), torch.from_numpy(b).to(DEV))
    assert np.max(np.abs(out.cpu().numpy() - a.astype(np.float64) @ b)) < 1e-4
