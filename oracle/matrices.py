"""Seeded test inputs (test infrastructure only).

``make_matrix`` restates ``svdbench/matrices.py:8-62`` (families gaussian, sparse, zero,
geometric, clustered, low_rank) so that the reference's own corpus bytes regenerate exactly for the
QR-free families.  The ``workload_*`` functions generate the five benchmark workloads:

* ``workload_single_jacobi``  512x256 float32, sigma geometric 1..1e-6, Haar U and V
* ``workload_uneven_batch``   4096 float32 matrices, m ~ U{16..96}, n ~ U{8..m}, N(0,1) entries
* ``workload_qdwh``           n x n float32, log-uniform sigma in [1e-6, 1] with a 10% cluster
* ``workload_tall_lowrank``   8192x2048 float32 rank-64 product plus N(0, 1e-8) noise
"""
import numpy as np

SEED = 20261011


def make_matrix(case):
    """Restatement of ``make_matrix`` (matrices.py:8-62)."""
    m, n = case['m'], case['n']
    b = case.get('batch', 1)
    rng = np.random.default_rng(case.get('seed', SEED))
    dtype = np.dtype(case.get('dtype', 'float32'))
    if dtype not in [np.dtype(x) for x in ('float32', 'float64', 'complex64', 'complex128')]:
        raise ValueError('Use float32/64 or complex64/128')

    def random(shape):
        a = rng.standard_normal(shape)
        if np.issubdtype(dtype, np.complexfloating):
            a = (a + 1j * rng.standard_normal(shape)) / np.sqrt(2)
        return a

    matrices = []
    family = case.get('family', 'gaussian')
    r = min(m, n)
    for _ in range(b):
        if family == 'gaussian':
            a = random((m, n)) / np.sqrt(max(m, n))
        elif family == 'sparse':
            density = case.get('density', 0.05)
            a = random((m, n)) / np.sqrt(max(m, n))
            a *= rng.random((m, n)) < density
        elif family == 'zero':
            a = np.zeros((m, n))
        else:
            q, rq = np.linalg.qr(random((m, r)), mode='reduced')
            v, rv = np.linalg.qr(random((n, r)), mode='reduced')
            if case.get('haar'):
                # Sign fix so that Q is Haar distributed (Mezzadri, Notices AMS 54(5), 2007).
                q = q * np.sign(np.diagonal(rq))
                v = v * np.sign(np.diagonal(rv))
            if family == 'geometric':
                s = np.geomspace(1, 1 / case.get('condition', 1e6), r)
            elif family == 'clustered':
                cluster = case.get('cluster_size', 4)
                spread = case.get('cluster_spread', 0.0)
                condition = case.get('condition', 1e3)
                centers = np.geomspace(1, 1 / condition, (r + cluster - 1) // cluster)
                s = np.repeat(centers, cluster)[:r]
                if spread:
                    s *= np.tile(np.linspace(1 + spread / 2, 1 - spread / 2, cluster), len(centers))[:r]
                    s = np.sort(s)[::-1]
            elif family == 'low_rank':
                k = case.get('rank', max(1, r // 8))
                s = np.zeros(r)
                s[:k] = np.geomspace(1, 0.1, k)
                s[k:] = case.get('tail', 1e-5)
            else:
                raise ValueError(f'Unknown matrix family: {family}')
            a = (q * s) @ v.conj().T
        matrices.append(a.astype(dtype))
    a = np.stack(matrices) if b > 1 else matrices[0]
    return np.ascontiguousarray(a * case.get('scale', 1.0))


def workload_single_jacobi(seed=SEED, m=512, n=256, condition=1e6):
    """512x256 float32, A = U diag(sigma) V^T with Haar U, V and sigma geometric 1 .. 1/condition."""
    return make_matrix(dict(m=m, n=n, family='geometric', condition=condition, dtype='float32',
                            seed=seed, haar=True))


def uneven_shapes(count=4096, seed=SEED, m_lo=16, m_hi=96, n_lo=8):
    """Shapes of the uneven batch: m ~ U{m_lo..m_hi}, n ~ U{n_lo..m}."""
    rng = np.random.default_rng(seed)
    ms = rng.integers(m_lo, m_hi + 1, size=count)
    ns = np.array([rng.integers(n_lo, m + 1) for m in ms])
    return ms.astype(np.int64), ns.astype(np.int64)


def workload_uneven_batch(count=4096, seed=SEED, m_lo=16, m_hi=96, n_lo=8):
    """Uneven batch of row-major float32 matrices with i.i.d. N(0,1) entries.

    Returns (flat, ms, ns, offsets): matrix i occupies flat[offsets[i]:offsets[i]+ms[i]*ns[i]]."""
    ms, ns = uneven_shapes(count, seed, m_lo, m_hi, n_lo)
    sizes = ms * ns
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    rng = np.random.default_rng(seed + 1)
    flat = rng.standard_normal(int(sizes.sum()), dtype=np.float32)
    return flat, ms, ns, offsets


def unpack(flat, ms, ns, offsets):
    return [flat[o:o + m * n].reshape(m, n) for o, m, n in zip(offsets, ms, ns)]


def haar(n, rng, k=None):
    """n x k Haar-distributed orthonormal columns (QR of N(0,1) with the R-diagonal sign fix)."""
    q, r = np.linalg.qr(rng.standard_normal((n, k or n)))
    return q * np.sign(np.diagonal(r))


def qdwh_spectrum(n, rng, lo=1e-6, cluster_frac=0.1, center=0.5, width=1e-4):
    """log-uniform sigma in [lo, 1] with ``cluster_frac`` of the values in a cluster of the given
    width around ``center``; sorted descending, largest forced to 1."""
    k = int(round(cluster_frac * n))
    s = np.exp(rng.uniform(np.log(lo), 0.0, n - k))
    c = center + width * (rng.random(k) - 0.5)
    s = np.sort(np.concatenate([s, c]))[::-1]
    s[0] = 1.0
    return s


def workload_qdwh(n=4096, seed=SEED):
    """n x n float32 A = U diag(sigma) V^T (FP64 product, rounded once), Haar U and V."""
    rng = np.random.default_rng(seed)
    s = qdwh_spectrum(n, rng)
    u = haar(n, rng)
    v = haar(n, rng)
    return ((u * s) @ v.T).astype(np.float32), s


def workload_tall_lowrank(m=8192, n=2048, rank=64, seed=SEED):
    """Rank-``rank`` product G1 G2^T with sigma geometric 1 .. 1e-3 plus N(0, 1e-8) noise."""
    rng = np.random.default_rng(seed)
    g1 = haar(m, rng, rank)
    g2 = haar(n, rng, rank)
    s = np.geomspace(1, 1e-3, rank)
    a = (g1 * s) @ g2.T + rng.normal(0.0, 1e-4, size=(m, n))
    return a.astype(np.float32)


def l0_spectrum(n, rng, log10_cond, family):
    """Singular values for the learned-l0 workload: 'geometric' from 1 to 10^-log10_cond, or
    'clustered' -- log-uniform on that range with a tenth of the values in a cluster of relative
    width 1e-4 at a random position; sorted descending with the extremes pinned."""
    lo = 10.0 ** -log10_cond
    if family == 'geometric':
        return np.geomspace(1.0, lo, n)
    k = max(1, n // 10)
    s = np.exp(rng.uniform(np.log(lo), 0.0, n - k))
    c = np.exp(rng.uniform(np.log(lo), 0.0)) * (1.0 + 1e-4 * (rng.random(k) - 0.5))
    s = np.sort(np.concatenate([s, c]))[::-1]
    s[0], s[-1] = 1.0, lo
    return s


def workload_l0(count, n, seed=SEED, cond_lo=1.0, cond_hi=12.0):
    """`count` float32 n x n matrices U diag(sigma) V^T (Haar U, V; FP64 product rounded once) with
    log10 condition ~ U[cond_lo, cond_hi] and geometric or clustered spectra (alternating).
    Returns (a (count, n, n) float32, log10 condition numbers, families)."""
    rng = np.random.default_rng(seed)
    conds = rng.uniform(cond_lo, cond_hi, count)
    fams = ['geometric' if i % 2 == 0 else 'clustered' for i in range(count)]
    out = np.empty((count, n, n), dtype=np.float32)
    for i in range(count):
        s = l0_spectrum(n, rng, conds[i], fams[i])
        out[i] = ((haar(n, rng) * s) @ haar(n, rng).T).astype(np.float32)
    return out, conds, fams
