"""FP8 e4m3 (OCP) per-tensor quantisation for the fp8 GEMM form.

scale = 448 / max|x| is computed on the device in fp32; q = rne_sat(x*scale)
is written K-contiguous, the layout the fp8 GEMM (TN) reads.  The GEMM folds
1/(scaleA*scaleB) into alpha.
"""
from __future__ import annotations

import torch

from . import _capi as C
from .gemm import _stream


def quantize(x: torch.Tensor, rows: int, k: int, ldx: int, trans_x: bool, ldq: int | None = None,
             stream=None):
    """Quantise a column-major rows x k operand view of fp32 ``x``.

    trans_x=True: element (r, kk) at x[kk + r*ldx] (K-contiguous);
    trans_x=False: at x[r + kk*ldx].  Returns (q uint8 [rows*ldq], scale fp32[1]).
    """
    assert x.dtype == torch.float32 and x.is_cuda
    ldq = ldq or k
    q = torch.empty(rows * ldq, dtype=torch.uint8, device=x.device)
    scale = torch.empty(1, dtype=torch.float32, device=x.device)
    work = torch.empty(C.lib().ffgemm_fp8_workspace_size(), dtype=torch.uint8, device=x.device)
    C.check(C.lib().ffgemm_fp8_quantize(x.data_ptr(), rows, k, ldx, int(trans_x), q.data_ptr(), ldq,
                                        scale.data_ptr(), work.data_ptr(), _stream(stream)))
    return q, scale
