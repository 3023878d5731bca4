"""Binding to the gfx950 kernel library (``libpavelsvd.so`` through the ``_hip`` module).

There is no fallback: importing this module without the built library raises ImportError, and
every entry point refuses tensors that are not on a ROCm device.
"""
import torch

try:
    from . import _hip
except ImportError as exc:  # pragma: no cover - exercised only on an unbuilt tree
    raise ImportError('the gfx950 kernel library is not built: run `python pavel-s_svd_amd/build.py` '
                      '(or __graft_entry__.build()) first') from exc

F32, F64, C64, C128 = 0, 1, 2, 3
DTYPE_CODE = {torch.float32: F32, torch.float64: F64, torch.complex64: C64, torch.complex128: C128}

OK, EINVAL, ELAUNCH, ETOOBIG = 0, 1, 2, 3
INFO_OK, INFO_NOCONV, INFO_COMPLETION, INFO_NONFINITE, INFO_BADSHAPE, INFO_BARRIER = 0, 1, 2, 3, 4, 5

SCALE, CUT_FLOOR, NO_VECTORS, SLOT_SPREAD, ORDER_SORTED, BLOCK_NARROW, BLOCK_WIDE = 1, 2, 4, 8, 16, 32, 64
BLOCK_FULL_PAIRS = 128

_STATUS = {EINVAL: 'invalid argument', ELAUNCH: 'kernel launch failed', ETOOBIG: 'problem exceeds the kernel budget'}


class NativeError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, what, status):
        super().__init__(f'{what}: {_STATUS.get(status, status)}')
        self.status = status


def check(status, what):
    if status != OK:
        raise NativeError(what, status)


def version():
    return _hip.version()


def stream_of(t):
    """hipStream_t handle of torch's current stream on the tensor's device."""
    return torch.cuda.current_stream(t.device).cuda_stream


def ptr(t):
    return 0 if t is None else t.data_ptr()


def dtype_code(dtype):
    try:
        return DTYPE_CODE[dtype]
    except KeyError:
        raise TypeError(f'the gfx950 kernels take float32/float64 (complex64/128: LDS Jacobi), not {dtype}') from None


def raise_for_info(info, algorithm_error):
    """Map per-matrix device status codes to the reference's exceptions
    (primitives.py:355, :451-452; solvers.py:291-292)."""
    codes = info[..., 1]
    if bool((codes == INFO_NONFINITE).any()):
        raise ValueError('input contains NaN or infinity')
    if bool((codes == INFO_BADSHAPE).any()):
        raise ValueError('matrix descriptor outside 1 <= n <= m <= max_m, n <= max_n')
    if bool((codes == INFO_BARRIER).any()):
        raise RuntimeError('grid barrier wait abandoned: the launch could not keep all its workgroups resident '
                           '(the GPU was shared with other work)')
    if bool((codes == INFO_NOCONV).any()):
        raise RuntimeError(algorithm_error)
    if bool((codes == INFO_COMPLETION).any()):
        raise RuntimeError('null-space completion failed: supplied live columns are not independent')


# --- workspace accounting -----------------------------------------------------------------------
# Every entry point takes caller-allocated workspace; this package allocates it from torch's caching
# allocator (so the harness's allocator peaks include it) and the library itself allocates no
# device memory.  The counters let an adapter report what one solve requested.
_WS = {'largest': 0, 'total': 0, 'requests': 0}


def workspace(nbytes, device):
    """A uint8 device tensor of `nbytes` for one entry point's workspace, counted."""
    nbytes = int(nbytes)
    _WS['largest'] = max(_WS['largest'], nbytes)
    _WS['total'] += nbytes
    _WS['requests'] += 1
    return torch.empty(nbytes, dtype=torch.uint8, device=device)


def workspace_reset():
    _WS.update(largest=0, total=0, requests=0)


def workspace_stats():
    """{'largest', 'total', 'requests'} since the last reset."""
    return dict(_WS)

