"""Uneven batches: many tall matrices of different shapes solved by one persistent launch.

The reference API takes one (..., m, n) tensor (solvers.py:289-290), so a batch of mixed shapes
is a Python loop of ``svd`` calls there.  ``UnevenBatch`` packs the matrices row-major into one
device buffer, builds the descriptor arrays once (shapes, element offsets, and a longest-first
processing order by the Jacobi cost n^2 (12 m + 6 n) that balances the device work queue), and
solves the whole batch with one launch of the LDS-resident Jacobi kernel.  ``capture`` records
that launch in a hipGraph (through torch.cuda.CUDAGraph), after which ``replay`` re-solves the
current contents of ``a`` with a single graph launch.
"""
import torch

from .. import native as N


class UnevenBatch:
    """A packed batch of tall matrices on one ROCm device.

    ``flat``: 1-D device tensor holding matrix i row-major at ``offsets[i]`` (default: packed
    back to back); ``ms``, ``ns``: host sequences of the shapes (m_i >= n_i >= 1).
    """

    def __init__(self, flat, ms, ns, offsets=None, *, vectors=True):
        ms = [int(v) for v in ms]
        ns = [int(v) for v in ns]
        if len(ms) != len(ns):
            raise ValueError('ms and ns differ in length')
        if any(n < 1 or n > m for m, n in zip(ms, ns)):
            raise ValueError('every matrix must be tall and nonempty (1 <= n <= m)')
        if flat.dim() != 1 or flat.device.type != 'cuda':
            raise ValueError('flat must be a 1-D ROCm device tensor')
        self.code = N.dtype_code(flat.dtype)
        self.batch = len(ms)
        self.ms, self.ns = ms, ns
        self.max_m, self.max_n = max(ms, default=1), max(ns, default=1)
        self.vectors = vectors
        self.a = flat
        dev = flat.device
        a_off, s_off, v_off = [], [], []
        pa = ps = pv = 0
        for i, (m, n) in enumerate(zip(ms, ns)):
            a_off.append(pa if offsets is None else int(offsets[i]))
            s_off.append(ps)
            v_off.append(pv)
            pa += m * n
            ps += n
            pv += n * n
        if offsets is None and pa != flat.numel():
            raise ValueError(f'flat holds {flat.numel()} elements, the shapes need {pa}')
        if offsets is not None and any(o < 0 or o + m * n > flat.numel() for o, m, n in zip(a_off, ms, ns)):
            raise ValueError('an offset places a matrix outside flat')
        self.a_off, self.s_off, self.v_off = a_off, s_off, v_off
        cost = [n * n * (12 * m + 6 * n) for m, n in zip(ms, ns)]
        order = sorted(range(self.batch), key=lambda i: -cost[i])
        host_dims = torch.tensor([[m, n] for m, n in zip(ms, ns)], dtype=torch.int32).reshape(-1)
        host_offs = torch.tensor(list(zip(a_off, s_off, v_off)), dtype=torch.int64).reshape(-1)
        self.dims = host_dims.to(dev)
        self.offs = host_offs.to(dev)
        self.order = torch.tensor(order, dtype=torch.int32).to(dev)
        self.u = torch.empty(flat.numel(), dtype=flat.dtype, device=dev) if vectors else None
        self.s = torch.empty(ps, dtype=flat.dtype, device=dev)
        self.vh = torch.empty(pv, dtype=flat.dtype, device=dev) if vectors else None
        self.info = torch.empty(self.batch, 2, dtype=torch.int32, device=dev)
        self.ws_bytes = N._hip.jacobi_lds_workspace_bytes(self.batch)
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=dev)
        st, self.smem = N._hip.jacobi_lds_smem_bytes(self.code, self.max_m, self.max_n,
                                                      0 if vectors else N.NO_VECTORS)
        N.check(st, f'uneven batch with max shape {self.max_m}x{self.max_n}')
        self.graph = None
        self.params = dict(tol=None, max_sweeps=60, scale=True, cut='reference', lanes=0, threads=256)

    @property
    def flops_per_sweep(self):
        """Algorithmic flops of one sweep over the batch: n(n-1)/2 pairs x (6m dot products +
        6(m+n) rotation) per matrix, the count used for the roofline."""
        return sum(n * (n - 1) // 2 * (12 * m + 6 * n) for m, n in zip(self.ms, self.ns))

    def configure(self, **params):
        unknown = set(params) - set(self.params)
        if unknown:
            raise ValueError(f'unknown parameters {sorted(unknown)}')
        self.params.update(params)
        self.graph = None
        return self

    def solve(self):
        """Launch the solve on the current stream; no host synchronisation."""
        p = self.params
        flags = (N.SCALE if p['scale'] else 0) | (N.CUT_FLOOR if p['cut'] == 'floor' else 0) | \
                (0 if self.vectors else N.NO_VECTORS)
        st = N._hip.jacobi_lds(self.code, self.batch, 0, 0, N.ptr(self.dims), N.ptr(self.offs),
                               N.ptr(self.order), self.max_m, self.max_n, N.ptr(self.a), N.ptr(self.u),
                               N.ptr(self.s), N.ptr(self.vh), N.ptr(self.info), float(p['tol'] or 0.0),
                               int(p['max_sweeps']), flags, int(p['lanes']), int(p['threads']),
                               N.ptr(self.ws), self.ws_bytes, N.stream_of(self.a))
        N.check(st, 'psvd_jacobi_lds (uneven)')

    def capture(self):
        """Record one solve in a hipGraph; ``replay`` then re-runs it with one graph launch."""
        self.solve()                       # first launch outside capture (sets kernel attributes)
        torch.cuda.synchronize(self.a.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.solve()
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            self.capture()
        self.graph.replay()

    def check(self):
        """Synchronise, raise the reference's exceptions for failed matrices, return the stats."""
        info = self.info.cpu()
        N.raise_for_info(info, f"one-sided Jacobi did not converge in {self.params['max_sweeps']} sweeps")
        return {'algorithm': 'jacobi', 'vendor_factorizations': 0, 'converged': True,
                'jacobi_sweeps': int(info[:, 0].max()) if self.batch else 0,
                'sweeps_per_matrix': info[:, 0].tolist()}

    def results(self):
        """Per-matrix (U, s, Vh) views into the packed outputs."""
        out = []
        for m, n, ao, so, vo in zip(self.ms, self.ns, self.a_off, self.s_off, self.v_off):
            s = self.s[so:so + n]
            u = self.u[ao:ao + m * n].view(m, n) if self.vectors else None
            vh = self.vh[vo:vo + n * n].view(n, n) if self.vectors else None
            out.append((u, s, vh))
        return out


def svd_uneven(matrices, **params):
    """Convenience: SVD of a list of tall 2-D device tensors in one launch; returns a list of
    (U, s, Vh) and the stats dict."""
    if not matrices:
        return [], {'algorithm': 'jacobi', 'jacobi_sweeps': 0}
    flat = torch.cat([m.reshape(-1) for m in matrices])
    batch = UnevenBatch(flat, [m.shape[0] for m in matrices], [m.shape[1] for m in matrices])
    batch.configure(**params).solve()
    stats = batch.check()
    return batch.results(), stats
