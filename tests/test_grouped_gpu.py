"""Grouped GEMM (one persistent launch) and hipGraph replay, vs the oracle."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
import tensilelite_amd as tl
from tensilelite_amd import grouped
from tests._util import elem_ok, extract

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


# Uneven problems with ragged K (not a multiple of 64) whose tiles outnumber the
# workgroups of the persistent grid at every tile below (256 x 256: 265 tiles
# on 256 workgroups; 128 x 256: 523 on 256 or 512, by layout; 64 x 64: 4066 on
# 1024), so workgroups run several tiles of different problems one after the other.
_LAYOUT_SHAPES = [(4096, 2048, 64), (3000, 1000, 200), (2048, 2048, 96), (72, 256, 512), (1000, 304, 64),
                  (8, 40, 24), (520, 1032, 328)]


@pytest.mark.parametrize("layout", ["NN", "NT", "TN", "TT"])
@pytest.mark.parametrize("tile,dtype", [((256, 256), torch.bfloat16), ((128, 256), torch.bfloat16),
                                        ((256, 256), torch.float16), ((64, 64), torch.bfloat16)])
def test_grouped_layouts_persistent(cuda, layout, tile, dtype):
    print(f"case grouped {layout} tile={tile} {dtype}", flush=True)
    ta, tb = layout[0] == "T", layout[1] == "T"
    g = torch.Generator().manual_seed(11)
    probs, As, Bs, Ds, lds = [], [], [], [], []
    for (M, N, K) in _LAYOUT_SHAPES:
        lda, ldb = (K if ta else M), (N if tb else K)
        A = (torch.randn(M * K, generator=g) / K ** 0.5).to(dtype)
        B = (torch.randn(K * N, generator=g) / K ** 0.5).to(dtype)
        As.append(A.to(cuda)); Bs.append(B.to(cuda))
        Ds.append(torch.full((M * N,), 7.0, dtype=dtype, device=cuda))
        probs.append(tl.make_problem(layout, dtype, M, N, K, lda, ldb, M, M))
        lds.append((lda, ldb))
    plan = grouped.GroupedPlan(probs, As, Bs, None, Ds, tile=tile)
    info = plan.info()
    assert info["tile"] == tile and info["total_tiles"] > info["grid"], info
    plan.run()
    torch.cuda.synchronize()
    for (M, N, K), (lda, ldb), A, B, D in zip(_LAYOUT_SHAPES, lds, As, Bs, Ds):
        ref = oracle.gemm_f64(layout, M, N, K, 1.0, A.double().cpu().numpy(), lda, B.double().cpu().numpy(), ldb)
        bad, worst = elem_ok(D.float().cpu().numpy(), ref)
        assert bad == 0, (layout, M, N, K, worst)
    plan.close()


@pytest.mark.parametrize("tile", [(128, 128), (256, 256)])
def test_grouped_config_graph_replays_match_eager(cuda, tile):
    """The batching configuration (64 problems, seed 3) captured in a hipGraph:
    every replay, with eager work in between, equals the eager launch."""
    print(f"case grouped config + graph tile={tile}", flush=True)
    probs, As, Bs, Ds = grouped.make_config_problems(cuda, count=64)
    plan = grouped.GroupedPlan(probs, As, Bs, None, Ds, tile=tile)
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


# 256 x 256 tiles with C: odd K-tile counts (K = 152 runs three K-tiles of 64,
# padded to four in the NN streaming engine), a single K-tile (K = 40), K = 0
# (D = beta * C), and more tiles than workgroups so tiles follow one another.
_BETA_SHAPES = [(3000, 2048, 152), (1000, 520, 40), (520, 264, 0), (2048, 2048, 136), (296, 1032, 328),
                (4096, 2048, 64)]


@pytest.mark.parametrize("layout", ["NN", "TN"])
def test_grouped_256_alpha_beta_odd_ktiles(cuda, layout):
    print(f"case grouped {layout} 256x256 alpha/beta", flush=True)
    ta, tb = layout[0] == "T", layout[1] == "T"
    alpha, beta = 0.5, -1.25
    g = torch.Generator().manual_seed(17)
    probs, As, Bs, Cs, Ds, lds = [], [], [], [], [], []
    for (M, N, K) in _BETA_SHAPES:
        # K = 0: operands of 8 K-rows (never read, as every loader masks k >= K)
        Ka = max(K, 8)
        lda, ldb = (Ka if ta else M), (N if tb else Ka)
        A = (torch.randn(M * Ka, generator=g) / Ka ** 0.5).to(torch.bfloat16)
        B = (torch.randn(Ka * N, generator=g) / Ka ** 0.5).to(torch.bfloat16)
        Cm = torch.randn(M * N, generator=g).to(torch.bfloat16)
        As.append(A.to(cuda)); Bs.append(B.to(cuda)); Cs.append(Cm.to(cuda))
        Ds.append(torch.full((M * N,), 7.0, dtype=torch.bfloat16, device=cuda))
        probs.append(tl.make_problem(layout, torch.bfloat16, M, N, K, lda, ldb, M, M, alpha=alpha, beta=beta))
        lds.append((lda, ldb))
    plan = grouped.GroupedPlan(probs, As, Bs, Cs, Ds, tile=(256, 256))
    info = plan.info()
    assert info["tile"] == (256, 256) and info["total_tiles"] > info["grid"], info
    plan.run()
    torch.cuda.synchronize()
    for (M, N, K), (lda, ldb), A, B, Cm, D in zip(_BETA_SHAPES, lds, As, Bs, Cs, Ds):
        ref = oracle.gemm_f64(layout, M, N, K, alpha, A.double().cpu().numpy(), lda, B.double().cpu().numpy(), ldb,
                              beta=beta, C=Cm.double().cpu().numpy())
        bad, worst = elem_ok(D.float().cpu().numpy(), ref)
        assert bad == 0, (layout, M, N, K, worst)
    plan.close()


# NN at 256 x 256 (the streaming split-ring engine) with ragged last rows of
# tiles: M mod 256 at 1..8, 128 and 136, M <= 128, ragged N, K = 0, alpha and
# beta with C, two runs of one plan.
_HALF_SHAPES = [(8, 300, 64), (128, 512, 200), (136, 256, 96), (384, 776, 512), (392, 256, 72), (640, 1000, 0),
                (264, 64, 1024), (2148, 1536, 256), (1, 40, 16), (3000, 2048, 152)]


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
def test_grouped_256_ragged_rows(cuda, dtype):
    print(f"case grouped NN 256x256 ragged rows {dtype}", flush=True)
    alpha, beta = 1.5, 0.5
    g = torch.Generator().manual_seed(23)
    probs, As, Bs, Cs, Ds, lds = [], [], [], [], [], []
    for (M, N, K) in _HALF_SHAPES:
        Ka, ld = max(K, 8), (M + 7) // 8 * 8
        A = (torch.randn(ld * Ka, generator=g) / Ka ** 0.5).to(dtype)
        B = (torch.randn(Ka * N, generator=g) / Ka ** 0.5).to(dtype)
        Cm = torch.randn(ld * N, generator=g).to(dtype)
        As.append(A.to(cuda)); Bs.append(B.to(cuda)); Cs.append(Cm.to(cuda))
        Ds.append(torch.full((ld * N,), 7.0, dtype=dtype, device=cuda))
        probs.append(tl.make_problem("NN", dtype, M, N, K, ld, Ka, ld, ld, alpha=alpha, beta=beta))
        lds.append((ld, Ka))
    plan = grouped.GroupedPlan(probs, As, Bs, Cs, Ds, tile=(256, 256))
    assert plan.info()["total_tiles"] == sum(-(-M // 256) * -(-N // 256) for (M, N, K) in _HALF_SHAPES)
    for rep in range(2):  # a second run over the same schedule gives the same result
        plan.run()
        torch.cuda.synchronize()
        for (M, N, K), (ld, Ka), A, B, Cm, D in zip(_HALF_SHAPES, lds, As, Bs, Cs, Ds):
            ref = oracle.gemm_f64("NN", M, N, K, alpha, A.double().cpu().numpy(), ld, B.double().cpu().numpy(), Ka,
                                  beta, Cm.double().cpu().numpy(), ld, ld)
            got = D.float().cpu().numpy()
            bad, worst = elem_ok(extract(got, M, N, ld), extract(ref, M, N, ld))
            assert bad == 0, (rep, M, N, K, worst)
            assert np.all(got.reshape(N, ld)[:, M:] == 7.0), (M, N, K, "padding rows written")
    plan.close()


def test_grouped_rejects_16x16_mfma(cuda):
    """The grouped launch runs 32 x 32 MFMAs only; a record asking for 16 x 16 is refused."""
    from tensilelite_amd import _capi as C
    A = torch.zeros(64 * 64, dtype=torch.bfloat16, device=cuda)
    B = torch.zeros(64 * 64, dtype=torch.bfloat16, device=cuda)
    D = torch.zeros(64 * 64, dtype=torch.bfloat16, device=cuda)
    p = tl.make_problem("NN", torch.bfloat16, 64, 64, 64, 64, 64, 64, 64)
    with pytest.raises(C.FFGemmError, match="RTN_ERR"):
        grouped.GroupedPlan([p], [A], [B], None, [D], tile=(64, 64, 0, 0, 0, 16))
