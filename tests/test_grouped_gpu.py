"""Grouped GEMM (one persistent launch) and hipGraph replay, vs the oracle."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
import tensilelite_amd as tl
from tensilelite_amd import grouped
from tests._util import elem_ok

pytestmark = pytest.mark.gpu


def _check(probs, As, Ds):
    for p, A, B, D in zip(probs, As[0], As[1], Ds):
        M, N, K, lda, ldd = p.M, p.N, p.K, p.lda, p.ldd
        ref = oracle.gemm_f64("NN", M, N, K, 1.0, A.double().cpu().numpy(), lda, B.double().cpu().numpy(), K,
                              ldd=ldd)
        got = D.float().cpu().numpy()
        r = ref.reshape(N, ldd)[:, :M]
        g = got.reshape(N, ldd)[:, :M]
        bad, worst = elem_ok(g, r)
        assert bad == 0, (M, N, K, worst)


@pytest.mark.parametrize("tile", [(256, 128), (128, 128), (64, 64)])
def test_grouped_uneven_shapes(cuda, tile):
    print(f"case grouped tile={tile}", flush=True)
    shapes = [(72, 256, 512), (1000, 256, 512), (64, 1024, 2048), (300, 256, 2048), (8, 256, 512)]
    g = torch.Generator().manual_seed(3)
    probs, As, Bs, Ds = [], [], [], []
    for (M, N, K) in shapes:
        lda = (M + 7) // 8 * 8
        A = (torch.randn(lda * K, generator=g) / K ** 0.5).to(torch.bfloat16)
        B = (torch.randn(K * N, generator=g) / K ** 0.5).to(torch.bfloat16)
        As.append(A.to(cuda)); Bs.append(B.to(cuda))
        Ds.append(torch.zeros(lda * N, dtype=torch.bfloat16, device=cuda))
        probs.append(tl.make_problem("NN", torch.bfloat16, M, N, K, lda, K, lda, lda))
    plan = grouped.GroupedPlan(probs, As, Bs, None, Ds, tile=tile)
    plan.run()
    torch.cuda.synchronize()
    _check(probs, (As, Bs), Ds)
    plan.close()


def test_grouped_config_graph_replays_match_eager(cuda):
    """The batching configuration (64 problems, seed 3) captured in a hipGraph:
    every replay, with eager work in between, equals the eager launch."""
    print("case grouped config + graph", flush=True)
    probs, As, Bs, Ds = grouped.make_config_problems(cuda, count=64)
    plan = grouped.GroupedPlan(probs, As, Bs, None, Ds, tile=(128, 128))
    plan.run()
    torch.cuda.synchronize()
    eager = [d.clone() for d in Ds]
    # oracle check on a subset (the whole set is checked through the eager copy)
    sub = list(range(0, 64, 9))
    _check([probs[i] for i in sub], ([As[i] for i in sub], [Bs[i] for i in sub]), [eager[i] for i in sub])
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        plan.run()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        plan.run()
    for rep in range(3):
        for d in Ds:
            d.fill_(7.0)      # eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for p, d, e in zip(probs, Ds, eager):
            # the kernel writes rows [0, M) of each column; rows [M, ldd) keep the fill
            got = d.view(p.N, p.ldd)[:, : p.M]
            want = e.view(p.N, p.ldd)[:, : p.M]
            assert torch.equal(got, want), f"replay {rep} differs"
            assert torch.all(d.view(p.N, p.ldd)[:, p.M:] == 7.0)
    plan.close()
