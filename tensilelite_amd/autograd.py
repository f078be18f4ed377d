"""torch.autograd.Function for C = A @ B on the gfx950 kernels.

Row-major torch matrices; the library is column-major, so (see gemm.mm)
  forward  C  = A @ B      runs the NN kernel,
  backward dA = dC @ B^T   runs the TN adjoint kernel,
           dB = A^T @ dC   runs the NT adjoint kernel,
all through the C ABI (ctypes).  float64 inputs run the FP64 MFMA path, which
is what the finite-difference check (torch.autograd.gradcheck) uses.
"""
from __future__ import annotations

import torch

from .gemm import mm


class GemmFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
        a = a.contiguous()
        b = b.contiguous()
        ctx.save_for_backward(a, b)
        return mm(a, b)

    @staticmethod
    def backward(ctx, dc: torch.Tensor):
        a, b = ctx.saved_tensors
        dc = dc.contiguous().to(a.dtype)
        da = mm(dc, b, trans_b=True) if ctx.needs_input_grad[0] else None
        db = mm(a, dc, trans_a=True) if ctx.needs_input_grad[1] else None
        return da, db


def matmul(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """Differentiable ``a @ b`` on the library's kernels."""
    return GemmFunction.apply(a, b)
