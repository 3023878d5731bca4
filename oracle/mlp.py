"""numpy restatement of the learned QDWH lower bound (test infrastructure only).

No reference counterpart: the reference starts QDWH from l0 = eps (svdbench/standalone/solvers.py:
128-130).  This module restates, in FP64, what the GPU kernels compute, so that the kernels are
checked against something other than themselves:

    unpack / forward / loss_grad   the 32 -> 64 -> 64 -> 1 tanh MLP, its loss
                                   (1/2B) sum (y - t)^2 and the gradient by backpropagation
                                   (pinned by central finite differences in tests/test_oracle.py)
    features                       the 32 sketch features of a tall matrix (column-norm statistics
                                   and Ritz values of four steps of subspace iteration on A^T A
                                   from a fixed 8-column Gaussian start; pinned by closed forms on
                                   diagonal matrices in tests/test_oracle.py)
"""
import numpy as np

IN, HID = 32, 64
NPARAM = HID * IN + HID + HID * HID + HID + HID + 1
OMEGA_SEED = 20261011
SKETCH = 8


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    w1 = p[o:o + HID * IN].reshape(HID, IN); o += HID * IN
    b1 = p[o:o + HID]; o += HID
    w2 = p[o:o + HID * HID].reshape(HID, HID); o += HID * HID
    b2 = p[o:o + HID]; o += HID
    w3 = p[o:o + HID]; o += HID
    b3 = p[o]
    return w1, b1, w2, b2, w3, b3


def forward(p, x):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    h1 = np.tanh(x @ w1.T + b1)
    h2 = np.tanh(h1 @ w2.T + b2)
    return h2 @ w3 + b3, (h1, h2)


def loss_grad(p, x, t):
    """Loss (1/2B) sum (y - t)^2 and its gradient in the flat parameter layout."""
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    w1, b1, w2, b2, w3, b3 = unpack(p)
    y, (h1, h2) = forward(p, x)
    bsz = x.shape[0]
    dy = (y - t) / bsz
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    dz2 = np.outer(dy, w3) * (1 - h2 ** 2)
    gw2 = dz2.T @ h1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ w2) * (1 - h1 ** 2)
    gw1 = dz1.T @ x
    gb1 = dz1.sum(0)
    grad = np.concatenate([gw1.ravel(), gb1, gw2.ravel(), gb2, gw3, [gb3]])
    return 0.5 * np.mean((y - t) ** 2), grad


def omega(n):
    """Fixed orthonormal 8-column start of the subspace iteration (QR of seeded N(0, 1))."""
    q, _ = np.linalg.qr(np.random.default_rng(OMEGA_SEED).standard_normal((n, SKETCH)))
    return q


def colstats(a):
    """{min, max, mean, std} of log10 of the column norms, log10 max|a| (zero columns: 1e-30)."""
    a = np.asarray(a, dtype=np.float64)
    l = 0.5 * np.log10(np.maximum((a * a).sum(0), 1e-60))
    return np.array([l.min(), l.max(), l.mean(), l.std(), np.log10(max(np.abs(a).max(), 1e-30))])


def ritz(a, steps=4):
    """Ritz values (descending) of A^T A's top-8 subspace after each of `steps` iterations, as
    singular-value estimates: Y = A X, sqrt(eig(Y^T Y)), X <- orth(A^T Y)."""
    a = np.asarray(a, dtype=np.float64)
    x = omega(a.shape[1])
    out = []
    for t in range(steps):
        y = a @ x
        w = np.linalg.eigvalsh(y.T @ y)[::-1]
        out.append(np.sqrt(np.maximum(w, 0.0)))
        if t + 1 < steps:
            x, _ = np.linalg.qr(a.T @ y)
    return np.array(out)


def features(a):
    """The 32 features of one tall matrix (m >= n)."""
    a = np.asarray(a, dtype=np.float64)
    m, n = a.shape
    lg = lambda v: np.log10(np.maximum(v, 1e-30))
    fro = np.sqrt((a * a).sum())
    r = ritz(a)
    s1 = max(r[-1, 0], 1e-30)
    cs = colstats(a)
    f = [lg(fro / (s1 * np.sqrt(n)))]
    f += [lg(r[t, k] / s1) for t in range(4) for k in range(1, 8)]
    f += [cs[1] - cs[0], cs[3], np.log10(n)]
    return np.clip(np.array(f), -30.0, 30.0)
