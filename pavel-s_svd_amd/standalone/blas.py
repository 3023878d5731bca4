"""Matrix products on the MFMA GEMM kernel (K3) for the standalone hot path.

``gemm`` accepts 2-D or batched 3-D device tensors whose last two dimensions are row- or
column-contiguous (a transposed view is passed as the transposed flag of the kernel, never
copied), and writes into a fresh or caller-supplied output.
"""
import torch

from .. import native as N

MODES = {'f32': 0, 'bf16x3': 1, 'f64': 2}


def _operand(t):
    """(pointer, leading dimension, transposed, batch stride) of a 2-D/3-D tensor view."""
    if t.stride(-1) == 1 and t.stride(-2) >= max(t.shape[-1], 1):
        return t, t.stride(-2), False
    if t.stride(-2) == 1 and t.stride(-1) >= max(t.shape[-2], 1):
        return t, t.stride(-1), True
    t = t.contiguous()
    return t, t.stride(-2), False


def gemm(a, b, *, alpha=1.0, beta=0.0, out=None, mode=None):
    """out = alpha * a @ b + beta * out for (b?, m, k) @ (b?, k, n) device tensors.

    mode: 'f32' (FP32 MFMA), 'bf16x3' (split FP32 on bf16 MFMA) or 'f64'; default from dtype."""
    if mode is None:
        mode = 'f64' if a.dtype == torch.float64 else 'f32'
    want = torch.float64 if mode == 'f64' else torch.float32
    if a.dtype != want or b.dtype != want:
        raise TypeError(f'gemm mode {mode} takes {want} operands')
    batched = a.dim() == 3 or b.dim() == 3
    a3 = a if a.dim() == 3 else a.unsqueeze(0)
    b3 = b if b.dim() == 3 else b.unsqueeze(0)
    batch = max(a3.shape[0], b3.shape[0])
    m, k = a3.shape[-2:]
    k2, n = b3.shape[-2:]
    if k != k2:
        raise ValueError(f'inner dimensions differ: {tuple(a.shape)} @ {tuple(b.shape)}')
    if out is None:
        out = torch.empty((batch, m, n) if batched else (m, n), dtype=want, device=a.device)
        if beta != 0.0:
            raise ValueError('beta != 0 needs an output tensor')
    o3 = out if out.dim() == 3 else out.unsqueeze(0)
    if not (o3.stride(-1) == 1 and o3.is_contiguous() or o3.stride(-1) == 1):
        raise ValueError('out must be row-contiguous')
    a3, lda, ta = _operand(a3)
    b3, ldb, tb = _operand(b3)
    sa = a3.stride(0) if a3.shape[0] > 1 else 0
    sb = b3.stride(0) if b3.shape[0] > 1 else 0
    sc = o3.stride(0) if o3.shape[0] > 1 else 0
    st = N._hip.gemm(MODES[mode], int(ta), int(tb), m, n, k, float(alpha), N.ptr(a3), lda, sa, N.ptr(b3), ldb, sb,
                     float(beta), N.ptr(o3), o3.stride(-2), sc, batch, N.stream_of(a))
    N.check(st, 'psvd_gemm')
    return out
