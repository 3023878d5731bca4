"""Uneven batches: many tall matrices of different shapes solved by one persistent launch.

The reference API takes one (..., m, n) tensor (solvers.py:289-290), so a batch of mixed shapes
is a Python loop of ``svd`` calls there.  ``UnevenBatch`` packs the matrices row-major into one
device buffer and builds the descriptor arrays once: shapes, element offsets, and a processing
order.  The order groups the matrices into size classes by the lanes-per-column-pair count their
row count calls for (about ``rows_per_lane`` rows per lane), and within a class puts the largest
Jacobi cost n^2 (12 m + 6 n) first so the device work queue balances.  Each class is one
persistent launch of the LDS-resident Jacobi kernel sized for that class (its own LDS budget,
hence its own occupancy); ``capture`` records the launches of all classes in one hipGraph
(through torch.cuda.CUDAGraph), after which ``replay`` re-solves the current contents of ``a``
with a single graph launch.
"""
import torch

from .. import native as N

# Launch-shape policy.  LDS decides how many matrices can be resident on a CU (160 KiB each);
# the waves per matrix are chosen so that the resident matrices together supply about
# `target_waves` waves per CU (one to sixteen waves per matrix), and the lanes per column pair
# fill the workgroup with the matrix's n/2 pairs (at most `passes` pairs per group per round)
# while keeping at least `min_rows` rows per lane.
DEFAULT_POLICY = dict(target_waves=16, min_rows=4, passes=2)
LDS_PER_CU = 160 * 1024


def launch_shape(code, m, n, vectors, policy):
    """(threads, lanes per column pair) for an m x n matrix under ``policy``."""
    st, lds = N._hip.jacobi_lds_smem_bytes(code, m, n, 0 if vectors else N.NO_VECTORS)
    resident = max(1, LDS_PER_CU // max(lds, 1))
    waves = 1
    while waves * resident < policy['target_waves'] and waves < 16:
        waves *= 2
    threads = 64 * waves
    npairs = (n + 1) // 2
    lanes = 2
    while lanes * 2 * npairs <= threads * policy['passes'] and lanes * 2 * policy['min_rows'] <= m and lanes < 64:
        lanes *= 2
    return threads, lanes


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
        host_dims = torch.tensor([[m, n] for m, n in zip(ms, ns)], dtype=torch.int32).reshape(-1)
        host_offs = torch.tensor(list(zip(a_off, s_off, v_off)), dtype=torch.int64).reshape(-1)
        self.dims = host_dims.to(dev)
        self.offs = host_offs.to(dev)
        self.classes = []
        self.order = None
        self.u = torch.empty(flat.numel(), dtype=flat.dtype, device=dev) if vectors else None
        self.s = torch.empty(ps, dtype=flat.dtype, device=dev)
        self.vh = torch.empty(pv, dtype=flat.dtype, device=dev) if vectors else None
        self.info = torch.empty(self.batch, 2, dtype=torch.int32, device=dev)
        self.ws_bytes = N._hip.jacobi_lds_workspace_bytes(self.batch)
        self.ws = None
        st, self.smem = N._hip.jacobi_lds_smem_bytes(self.code, self.max_m, self.max_n,
                                                      0 if vectors else N.NO_VECTORS)
        N.check(st, f'uneven batch with max shape {self.max_m}x{self.max_n}')
        self.graph = None
        # class_threads: per size class ("threads,lanes" of the policy) a workgroup size replacing the
        # policy's, as the autotuner's second stage picks it
        # ordering: 'tournament' (the reference's round-robin labels) or 'sorted' (labels in descending
        # column-norm order each sweep); slots: 'identity' or 'spread' (the order in which a round's
        # pairs are served -- a launch-shape choice, bit-identical results)
        self.params = dict(tol=None, max_sweeps=60, scale=True, cut='reference', streams=4, class_threads={},
                           ordering='tournament', slots='identity', **DEFAULT_POLICY)
        self.side = []
        self._plan()

    def _plan(self):
        """Size classes (one launch shape each) and the processing order (longest first inside
        each class)."""
        cost = [n * n * (12 * m + 6 * n) for m, n in zip(self.ms, self.ns)]
        groups = {}
        for i, (m, n) in enumerate(zip(self.ms, self.ns)):
            groups.setdefault(launch_shape(self.code, m, n, self.vectors, self.params), []).append(i)
        order, classes = [], []
        overrides = self.params.get('class_threads') or {}
        for (threads, lanes) in sorted(groups, reverse=True):
            idx = sorted(groups[(threads, lanes)], key=lambda i: -cost[i])
            mm = max(self.ms[i] for i in idx)
            nn = max(self.ns[i] for i in idx)
            st, _ = N._hip.jacobi_lds_smem_bytes(self.code, mm, nn, 0 if self.vectors else N.NO_VECTORS)
            N.check(st, f'size class with max shape {mm}x{nn}')
            name = f'{threads},{lanes}'
            nt = int(overrides.get(name, threads))
            if nt not in (64, 128, 256, 512, 1024):
                raise ValueError(f'class_threads[{name!r}] must be 64, 128, 256, 512 or 1024')
            classes.append(dict(threads=nt, lanes=lanes, name=name, start=len(order), count=len(idx), max_m=mm,
                                max_n=nn, cost=sum(cost[i] for i in idx)))
            order.extend(idx)
        # classes run concurrently on side streams, greedily balanced by total cost
        nstreams = max(1, min(int(self.params['streams']), len(classes)))
        load = [0] * nstreams
        for cl in sorted(classes, key=lambda c: -c['cost']):
            k = load.index(min(load))
            cl['stream'] = k
            load[k] += cl['cost']
        self.classes = classes
        self.order = torch.tensor(order, dtype=torch.int32).to(self.a.device)
        dev = self.a.device
        self.ws = N.workspace(self.ws_bytes * max(1, len(classes)), dev)
        self.side = [torch.cuda.Stream(device=dev) for _ in range(nstreams)] if len(classes) > 1 else []

    @property
    def flops_per_sweep(self):
        """Algorithmic flops of one sweep over the batch: n(n-1)/2 pairs x (6m dot products +
        6(m+n) rotation) per matrix, the count used for the roofline."""
        return sum(n * (n - 1) // 2 * (12 * m + 6 * n) for m, n in zip(self.ms, self.ns))

    def configure(self, **params):
        unknown = set(params) - set(self.params)
        if unknown:
            raise ValueError(f'unknown parameters {sorted(unknown)}')
        if params.get('ordering', 'tournament') not in ('tournament', 'sorted'):
            raise ValueError("ordering must be 'tournament' or 'sorted'")
        if params.get('slots', 'identity') not in ('identity', 'spread'):
            raise ValueError("slots must be 'identity' or 'spread'")
        replan = any((k in DEFAULT_POLICY or k in ('streams', 'class_threads')) and v != self.params[k]
                     for k, v in params.items())
        self.params.update(params)
        if replan:
            self._plan()
        self.graph = None
        return self

    def solve(self):
        """Launch the solve on the current stream; no host synchronisation."""
        p = self.params
        flags = (N.SCALE if p['scale'] else 0) | (N.CUT_FLOOR if p['cut'] == 'floor' else 0) | \
                (0 if self.vectors else N.NO_VECTORS) | (N.ORDER_SORTED if p['ordering'] == 'sorted' else 0) | \
                (N.SLOT_SPREAD if p['slots'] == 'spread' else 0)
        main = torch.cuda.current_stream(self.a.device)
        if self.side:
            fork = torch.cuda.Event()
            fork.record(main)
            for st in self.side:
                st.wait_event(fork)
        for ci, cl in enumerate(self.classes):
            stream = self.side[cl['stream']].cuda_stream if self.side else main.cuda_stream
            order_ptr = N.ptr(self.order) + 4 * cl['start']
            st = N._hip.jacobi_lds(self.code, cl['count'], 0, 0, N.ptr(self.dims), N.ptr(self.offs), order_ptr,
                                   cl['max_m'], cl['max_n'], N.ptr(self.a), N.ptr(self.u), N.ptr(self.s),
                                   N.ptr(self.vh), N.ptr(self.info), float(p['tol'] or 0.0),
                                   int(p['max_sweeps']), flags, cl['lanes'], cl['threads'],
                                   N.ptr(self.ws) + ci * self.ws_bytes, self.ws_bytes, stream)
            N.check(st, f"psvd_jacobi_lds (uneven, {cl['threads']}x{cl['lanes']} class)")
        for st in self.side:
            join = torch.cuda.Event()
            join.record(st)
            main.wait_event(join)

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
