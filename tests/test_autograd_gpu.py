"""Differentiation form: dA = dC B^T (TN adjoint kernel) and dB = A^T dC (NT)
behind a torch.autograd.Function, checked by finite differences in FP64 and
against torch's own gradient."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from tensilelite_amd.autograd import GemmFunction, matmul
from tests._util import elem_ok

pytestmark = pytest.mark.gpu


def test_gradcheck_fp64_finite_differences(cuda):
    """Central differences in FP64 at M=64, N=48, K=32, U[-1,1], seed 4."""
    print("case gradcheck f64", flush=True)
    g = torch.Generator().manual_seed(4)
    a = (torch.rand(64, 32, generator=g, dtype=torch.float64) * 2 - 1).to(cuda).requires_grad_()
    b = (torch.rand(32, 48, generator=g, dtype=torch.float64) * 2 - 1).to(cuda).requires_grad_()
    assert torch.autograd.gradcheck(GemmFunction.apply, (a, b), eps=1e-6, atol=1e-8, rtol=1e-7)


def test_fp64_forward_backward_vs_oracle(cuda):
    g = torch.Generator().manual_seed(4)
    a = (torch.rand(64, 32, generator=g, dtype=torch.float64) * 2 - 1)
    b = (torch.rand(32, 48, generator=g, dtype=torch.float64) * 2 - 1)
    dc = torch.randn(64, 48, generator=g, dtype=torch.float64)
    ad, bd = a.to(cuda).requires_grad_(), b.to(cuda).requires_grad_()
    c = matmul(ad, bd)
    c.backward(dc.to(cuda))
    # row-major c = a b  <=>  column-major c^T = b^T a^T: NN with A' = b (48x32, ld 48), B' = a (32x64, ld 32)
    ref_c = oracle.gemm_f64("NN", 48, 64, 32, 1.0, b.numpy().ravel(), 48, a.numpy().ravel(), 32).reshape(64, 48)
    np.testing.assert_allclose(c.detach().cpu().numpy(), ref_c, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ad.grad.cpu().numpy(), (dc @ b.T).numpy(), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(bd.grad.cpu().numpy(), (a.T @ dc).numpy(), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("shape", [(256, 192, 128), (8192, 4096, 4096)])
def test_bf16_grads_vs_torch_fp32(cuda, shape):
    """bf16 forward/backward vs torch's gradient computed in fp32 from the same
    bf16 values (per-element 1e-2 bound, rms-scaled)."""
    M, N, K = shape
    print(f"case bf16 autograd {shape}", flush=True)
    g = torch.Generator().manual_seed(4)
    a = torch.randn(M, K, generator=g).to(torch.bfloat16).to(cuda)
    b = torch.randn(K, N, generator=g).to(torch.bfloat16).to(cuda)
    dc = torch.randn(M, N, generator=g).to(torch.bfloat16).to(cuda)
    ad, bd = a.clone().requires_grad_(), b.clone().requires_grad_()
    c = matmul(ad, bd)
    c.backward(dc)
    a32, b32, dc32 = a.float(), b.float(), dc.float()
    rows = slice(0, min(M, 512))
    ref_c = (a32[rows] @ b32)
    ref_da = (dc32[rows] @ b32.T.contiguous())
    ref_db = (a32.T.contiguous() @ dc32)
    for got, ref in ((c.detach().float()[rows], ref_c), (ad.grad.float()[rows], ref_da), (bd.grad.float(), ref_db)):
        bad, worst = elem_ok(got.cpu().numpy(), ref.cpu().numpy())
        assert bad == 0, worst
