"""psvd_jacobi_lds_noalloc (the batched Jacobi entry point whose work-queue counter lives in a
library-owned device array instead of a workspace) against psvd_jacobi_lds through the C ABI: two
calls in a row, the second on a smaller batch, outputs bit-identical to the allocating entry point on
the same inputs (uniform FP32, uneven FP32, uniform FP64, uniform complex64); batches beyond one
int32 counter are refused before any launch."""
import numpy as np
import pytest
import torch

from conftest import load_package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def N():
    load_package()
    return __import__('importlib').import_module('pavel-s_svd_amd.native')


def _outputs(x, vectors=True):
    b, m, n = x.shape
    real = x.real.dtype if x.is_complex() else x.dtype
    return (torch.full_like(x, float('nan')), torch.full((b, n), float('nan'), dtype=real, device=x.device),
            torch.full((b, n, n), float('nan'), dtype=x.dtype, device=x.device),
            torch.full((b, 2), -7, dtype=torch.int32, device=x.device))


def _uniform_pair(N, x, lanes, threads):
    """(allocating outputs, non-allocating outputs) for a uniform batch."""
    b, m, n = x.shape
    code = N.dtype_code(x.dtype)
    flags = N.SCALE
    res = []
    for noalloc in (False, True):
        u, s, vh, info = _outputs(x)
        if noalloc:
            st = N._hip.jacobi_lds_noalloc(code, b, m, n, 0, 0, 0, m, n, N.ptr(x), N.ptr(u), N.ptr(s), N.ptr(vh),
                                           N.ptr(info), 0.0, 60, flags, lanes, threads, N.stream_of(x))
        else:
            wsb = N._hip.jacobi_lds_workspace_bytes(b)
            ws = torch.empty(wsb, dtype=torch.uint8, device=x.device)
            st = N._hip.jacobi_lds(code, b, m, n, 0, 0, 0, m, n, N.ptr(x), N.ptr(u), N.ptr(s), N.ptr(vh),
                                   N.ptr(info), 0.0, 60, flags, lanes, threads, N.ptr(ws), wsb, N.stream_of(x))
        assert st == N.OK, (noalloc, st)
        torch.cuda.synchronize()
        res.append((u, s, vh, info))
    return res


def _same(ref, got):
    for r, g in zip(ref, got):
        assert torch.equal(r, g), (r.dtype, r.shape)
    assert int(got[3][:, 1].abs().max()) == 0 and int(got[3][:, 0].min()) >= 1


@pytest.mark.parametrize('dtype,shapes,lanes,threads', [
    (torch.float32, [(300, 40, 24), (37, 17, 9)], 8, 256),
    (torch.float64, [(64, 48, 48), (5, 20, 20)], 16, 512),
    (torch.complex64, [(96, 30, 18), (11, 12, 7)], 0, 256),
])
def test_noalloc_matches_allocating_entry_point_uniform(N, dtype, shapes, lanes, threads):
    g = torch.Generator().manual_seed(11)
    for b, m, n in shapes:   # the second call is smaller and reuses the same device array
        x = torch.randn(b, m, n, generator=g, dtype=dtype).to('cuda:0')
        ref, got = _uniform_pair(N, x, lanes, threads)
        _same(ref, got)


def test_noalloc_matches_allocating_entry_point_uneven(N):
    rng = np.random.default_rng(5)
    for count in (500, 61):
        ms = rng.integers(16, 97, count)
        ns = np.array([rng.integers(8, m + 1) for m in ms])
        dims = torch.tensor(np.stack([ms, ns], 1).reshape(-1), dtype=torch.int32, device='cuda:0')
        a_off = np.concatenate([[0], np.cumsum(ms * ns)[:-1]])
        s_off = np.concatenate([[0], np.cumsum(ns)[:-1]])
        v_off = np.concatenate([[0], np.cumsum(ns * ns)[:-1]])
        offs = torch.tensor(np.stack([a_off, s_off, v_off], 1).reshape(-1), dtype=torch.int64, device='cuda:0')
        order = torch.tensor(np.argsort(-(ns.astype(np.int64) ** 2 * ms), kind='stable'), dtype=torch.int32,
                             device='cuda:0')
        a = torch.randn(int((ms * ns).sum()), dtype=torch.float32).to('cuda:0')
        outs = []
        for noalloc in (False, True):
            u = torch.full_like(a, float('nan'))
            s = torch.full((int(ns.sum()),), float('nan'), device='cuda:0')
            vh = torch.full((int((ns * ns).sum()),), float('nan'), device='cuda:0')
            info = torch.full((count, 2), -7, dtype=torch.int32, device='cuda:0')
            args = (N.F32, count, 0, 0, N.ptr(dims), N.ptr(offs), N.ptr(order), 96, 96, N.ptr(a), N.ptr(u), N.ptr(s),
                    N.ptr(vh), N.ptr(info), 0.0, 60, N.SCALE | N.CUT_FLOOR, 16, 512)
            if noalloc:
                st = N._hip.jacobi_lds_noalloc(*args, N.stream_of(a))
            else:
                wsb = N._hip.jacobi_lds_workspace_bytes(count)
                ws = torch.empty(wsb, dtype=torch.uint8, device='cuda:0')
                st = N._hip.jacobi_lds(*args, N.ptr(ws), wsb, N.stream_of(a))
            assert st == N.OK
            torch.cuda.synchronize()
            outs.append((u, s, vh, info))
        _same(*outs)


def test_noalloc_refuses_batches_beyond_one_counter(N):
    x = torch.zeros(1, 4, 4, device='cuda:0')
    st = N._hip.jacobi_lds_noalloc(N.F32, 2 ** 31, 4, 4, 0, 0, 0, 4, 4, N.ptr(x), N.ptr(x), N.ptr(x), N.ptr(x),
                                   N.ptr(x), 0.0, 60, 0, 4, 256, N.stream_of(x))
    assert st == N.ETOOBIG
