"""The 8192 members of the reference's phantom shape set
(test/phantom_test.sh:5-53: TN, one free dimension at 480 or 512, the other
two 8192) at full size, for fp32, bf16 and fp16, through the library's own
plan with a workspace (tuned tile, split-K where the tuning table says so).
Sampled output columns against the FP64 oracle on the same rounded inputs.

Tolerances: fp32 rel-Frobenius <= 1e-5; 16-bit per element
|d - r| <= 1e-2 (|r| + rms(r)).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

import oracle
import tensilelite_amd as tl
from tensilelite_amd import phantom
from tests._util import elem_ok, rel_fro

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("family", ["480xN", "512xN", "Nx480", "Nx512"])
def test_phantom_8192_sampled_columns(cuda, name, family):
    M, N, K = phantom.shape(family, 8192)
    code, tdt, _ = phantom.DTYPES[name]
    g = torch.Generator(device="cpu").manual_seed(11)
    A = (torch.randn(M * K, generator=g) / np.sqrt(K)).to(tdt)    # op(A) rows K-contiguous: (k, m) at k + m*K
    B = (torch.randn(N * K, generator=g) / np.sqrt(K)).to(tdt)    # (k, n) at k + n*K
    p = tl.make_problem("TN", code, M, N, K, K, K, M, M)
    plan = tl.query_tile_ws(p)
    _say(f"case phantom {name} {M}x{N}x{K} tile {plan.mt_m}x{plan.mt_n} split {plan.split_k}")
    n = tl.workspace_size(p)
    ws = torch.empty(n, dtype=torch.uint8, device=cuda) if n else None
    D = torch.empty(M * N, dtype=tdt, device=cuda)
    tl.run(p, A.to(cuda), B.to(cuda), None, D, workspace=ws)
    torch.cuda.synchronize()
    cols = np.random.default_rng(M + N).choice(N, 8, replace=False)
    A64 = A.double().numpy()
    Bc = B.double().numpy().reshape(N, K)[cols]
    ref = oracle.gemm_f64("TN", M, len(cols), K, 1.0, A64, K, Bc.ravel(), K).reshape(len(cols), M)
    got = D.float().cpu().numpy().reshape(N, M)[cols]
    if name == "f32":
        err = rel_fro(got, ref)
        _say(f"rel_fro={err:.3e}")
        assert err <= 1e-5
    else:
        bad, worst = elem_ok(got, ref)
        assert bad == 0, worst
