"""PyTorch autograd over the permanent engine.

``permanent(A)`` takes a complex128 tensor of shape (..., n, n) on the GPU
and returns perm(A) for every matrix of the batch.  When A requires a
gradient, the forward pass runs the adjoint sweep (pqb_perm_grad_device),
which yields perm(A) and G = d perm / d A from the same Gray walk; G is kept
for the backward pass, which only scales it: for the holomorphic perm,
PyTorch's convention for complex inputs gives

    grad_A = grad_out * conj(G).

Without a gradient the forward pass is the batched plain sweep on the same
device pointers.  There is no CPU path: a tensor that is not on the GPU is
rejected.
"""

from __future__ import annotations

import torch

from . import _capi as C


def _check_input(A: torch.Tensor) -> int:
    if not isinstance(A, torch.Tensor):
        raise TypeError("permanent expects a torch.Tensor")
    if A.dtype != torch.complex128:
        raise TypeError("permanent expects complex128 matrices")
    if not A.is_cuda:
        raise ValueError("permanent runs on the GPU: move the tensor to a CUDA (HIP) device first")
    if A.dim() < 2 or A.shape[-1] != A.shape[-2]:
        raise ValueError("permanent expects square matrices in the last two dimensions")
    n = A.shape[-1]
    if n < 1 or n > C.MAX_COLS:
        raise ValueError(f"matrix order must be in [1, {C.MAX_COLS}]")
    return n


def _grad_sweep(Ab: torch.Tensor, n: int):
    """Ab: (count, n, n) contiguous complex128 on the GPU -> (perm, G) on the GPU."""
    count = Ab.shape[0]
    perm = torch.empty(count, dtype=torch.complex128, device=Ab.device)
    G = torch.empty_like(Ab)
    wb = C.grad_workspace_bytes(count, n)
    work = torch.empty(wb, dtype=torch.uint8, device=Ab.device)
    stream = torch.cuda.current_stream(Ab.device).cuda_stream
    C.permanent_grad_device(Ab.data_ptr(), count, n, perm.data_ptr(), G.data_ptr(), work.data_ptr(), wb, stream)
    return perm, G


def _forward_only(Ab: torch.Tensor, n: int) -> torch.Tensor:
    """Batched plain sweep on device pointers (pqb_perm_batch_device)."""
    count = Ab.shape[0]
    out = torch.empty((count, 6), dtype=torch.float64, device=Ab.device)
    wb = C.batch_device_workspace_bytes(count, n, n)
    work = torch.zeros(wb, dtype=torch.uint8, device=Ab.device)  # per-matrix counters start at zero
    stream = torch.cuda.current_stream(Ab.device).cuda_stream
    C.permanent_batch_device(Ab.data_ptr(), count, n, n, out.data_ptr(), work.data_ptr(), wb, stream)
    return torch.complex(out[:, 0], out[:, 1])


class Permanent(torch.autograd.Function):
    """perm over the last two dimensions; the backward pass uses the adjoint
    sweep's gradient saved by the forward pass."""

    @staticmethod
    def forward(ctx, A):
        n = _check_input(A)
        batch = A.shape[:-2]
        Ab = A.detach().reshape(-1, n, n).contiguous()
        if ctx.needs_input_grad[0]:
            perm, G = _grad_sweep(Ab, n)
            ctx.save_for_backward(G)
            ctx.in_shape = A.shape
        else:
            perm = _forward_only(Ab, n)
        return perm.reshape(batch)

    @staticmethod
    def backward(ctx, grad_out):
        (G,) = ctx.saved_tensors
        g = grad_out.reshape(-1, 1, 1).to(torch.complex128)
        return (g * G.conj()).reshape(ctx.in_shape)


def permanent(A: torch.Tensor) -> torch.Tensor:
    """Differentiable permanent of each n x n matrix in A (complex128, GPU)."""
    return Permanent.apply(A)
