"""The learned-l0 oracle (oracle/mlp.py) pinned by mathematics, not by itself: backpropagation
against central finite differences, the forward pass on closed-form parameter sets, and the sketch
features on diagonal matrices (column norms known exactly; Ritz values interlace the spectrum)."""
import numpy as np

from oracle import mlp as O
from oracle import matrices as X


def _params(seed):
    rng = np.random.default_rng(seed)
    return rng.normal(0, 0.3, O.NPARAM)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(2)
    x = rng.normal(size=(9, O.IN))
    t = rng.normal(size=9)
    _, g = O.loss_grad(p, x, t)
    idx = rng.choice(O.NPARAM, 80, replace=False)
    idx = np.concatenate([idx, [O.NPARAM - 1, 0, O.HID * O.IN, O.NPARAM - 2]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (O.loss_grad(pp, x, t)[0] - O.loss_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_forward_closed_forms():
    p = np.zeros(O.NPARAM)
    p[-1] = 0.7                                    # only b3: y = b3
    x = np.random.default_rng(0).normal(size=(5, O.IN))
    assert np.allclose(O.forward(p, x)[0], 0.7)
    w1, b1, w2, b2, w3, b3 = O.unpack(np.arange(O.NPARAM, dtype=float))
    assert w1.shape == (64, 32) and w2.shape == (64, 64) and w3.shape == (64,) and b3 == O.NPARAM - 1
    p = np.zeros(O.NPARAM)                         # y = sum_j w3_j tanh(tanh(x_0)) with W1[:, 0] = 1, W2 = I
    p[:O.HID * O.IN].reshape(O.HID, O.IN)[:, 0] = 1.0
    off = O.HID * O.IN + O.HID
    p[off:off + O.HID * O.HID].reshape(O.HID, O.HID)[:] = np.eye(O.HID)
    p[off + O.HID * O.HID + O.HID:off + O.HID * O.HID + 2 * O.HID] = 0.5
    assert np.allclose(O.forward(p, x)[0], 32.0 * np.tanh(np.tanh(x[:, 0])))


def test_features_on_diagonal_matrices():
    s = np.geomspace(1.0, 1e-3, 64)
    a = np.diag(s)
    f = O.features(a)
    assert f.shape == (32,)
    assert np.isclose(f[29], 3.0) and np.isclose(f[31], np.log10(64))
    assert np.isclose(f[30], np.log10(s).std())
    r = O.ritz(a)
    assert np.all(r <= s[:8] * (1 + 1e-12))            # Ritz values interlace the spectrum
    assert np.all(np.diff(r[:, 0]) >= -1e-12)          # and the top one only grows with the steps
    cs = O.colstats(a)
    assert np.allclose(cs[:4], [-3.0, 0.0, np.log10(s).mean(), np.log10(s).std()])
    assert np.isclose(f[0], np.log10(np.linalg.norm(s) / (r[-1, 0] * 8.0)))


def test_l0_workload_spectra():
    a, conds, fams = X.workload_l0(6, 24, seed=5, cond_lo=1, cond_hi=5)
    for ai, c in zip(a, conds):
        sv = np.linalg.svd(ai.astype(np.float64), compute_uv=False)
        assert abs(np.log10(sv[0] / sv[-1]) - c) < 1e-3
    assert fams[:2] == ['geometric', 'clustered']
