"""hipGraph capture of the single-GEMM entry points.

ffgemm_run / ffgemm_run_ws / ffgemm_fp8_quantize allocate nothing and never
synchronise, so a caller can capture them.  Each test captures one call
sequence, replays the graph several times with eager work (a fill of the
outputs) between replays, and checks every replay bit-exactly against the
eager result (the kernels are deterministic: no atomics in the GEMMs, and the
fp8 amax reduction is a maximum).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
import tensilelite_amd as tl
from tensilelite_amd import _capi as C
from tests._util import rel_fro

pytestmark = pytest.mark.gpu


def _capture(fn):
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fn()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        fn()
    return graph


def _replays_match(graph, outs, eager, reps=3):
    for rep in range(reps):
        for o in outs:
            o.fill_(7)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for o, e in zip(outs, eager):
            assert torch.equal(o, e), f"replay {rep} differs from the eager result"


@pytest.mark.parametrize("layout", ["TN", "NN", "NT"])
def test_gemm_graph_replays(cuda, layout):
    M, N, K = 1024, 768, 512
    print(f"case graph bf16 {layout}", flush=True)
    g = torch.Generator().manual_seed(5)
    ta, tb = layout[0] == "T", layout[1] == "T"
    A = (torch.randn(M * K, generator=g) / 16).to(torch.bfloat16).to(cuda)
    B = (torch.randn(K * N, generator=g) / 16).to(torch.bfloat16).to(cuda)
    Cm = torch.randn(M * N, generator=g).to(torch.bfloat16).to(cuda)
    D = torch.empty(M * N, dtype=torch.bfloat16, device=cuda)
    fn = lambda: tl.gemm(layout, M, N, K, A, B, Cm, D, alpha=0.5, beta=1.0, tile=(256, 256))
    fn()
    torch.cuda.synchronize()
    eager = D.clone()
    ref = oracle.gemm_f64(layout, M, N, K, 0.5, A.float().cpu().numpy().astype(np.float64), K if ta else M,
                          B.float().cpu().numpy().astype(np.float64), N if tb else K, 1.0,
                          Cm.float().cpu().numpy().astype(np.float64), M, M)
    assert rel_fro(eager.float().cpu().numpy(), ref) < 1e-2
    _replays_match(_capture(fn), [D], [eager])


def test_bf16x3_workspace_graph_replays(cuda):
    M, N, K = 520, 264, 200
    print("case graph bf16x3 workspace", flush=True)
    g = torch.Generator().manual_seed(6)
    A = torch.randn(M * K, generator=g).to(cuda)
    B = torch.randn(K * N, generator=g).to(cuda)
    D = torch.empty(M * N, device=cuda)
    p = tl.make_problem("NN", C.F32, M, N, K, M, K, M, M, compute=C.COMPUTE_BF16X3)
    ws = torch.empty(tl.workspace_size(p), dtype=torch.uint8, device=cuda)
    fn = lambda: tl.run(p, A, B, None, D, workspace=ws)
    fn()
    torch.cuda.synchronize()
    eager = D.clone()
    ref = oracle.gemm_f64("NN", M, N, K, 1.0, A.cpu().numpy().astype(np.float64), M,
                          B.cpu().numpy().astype(np.float64), K)
    assert rel_fro(eager.cpu().numpy(), ref) <= 1e-5
    _replays_match(_capture(fn), [D], [eager])


@pytest.mark.parametrize("layout", ["NN", "NT"])
def test_repack_graph_replays(cuda, layout):
    """The K-major repack plan (copy launch, then the TN problem on the rest of
    the workspace, here split in two K slices) captured and replayed."""
    M, N, K = 520, 392, 1024
    print(f"case graph repack {layout}", flush=True)
    g = torch.Generator().manual_seed(9)
    A = (torch.randn(M * K, generator=g) / K ** 0.5).to(torch.bfloat16).to(cuda)   # M-contiguous
    B = (torch.randn(K * N, generator=g) / K ** 0.5).to(torch.bfloat16).to(cuda)
    Cm = torch.randn(M * N, generator=g).to(torch.bfloat16).to(cuda)
    D = torch.empty(M * N, dtype=torch.bfloat16, device=cuda)
    p = tl.make_problem(layout, C.BF16, M, N, K, M, K if layout[1] == "N" else N, M, M, 0.5, 1.0)
    rec = (128, 128, 0, 0, 0, 0, 2, 2)
    ws = torch.empty(tl.workspace_size(p, rec), dtype=torch.uint8, device=cuda)
    fn = lambda: tl.run(p, A, B, Cm, D, tile=rec, workspace=ws)
    fn()
    torch.cuda.synchronize()
    eager = D.clone()
    ref = oracle.gemm_f64(layout, M, N, K, 0.5, A.float().cpu().numpy().astype(np.float64), M,
                          B.float().cpu().numpy().astype(np.float64), K if layout[1] == "N" else N, 1.0,
                          Cm.float().cpu().numpy().astype(np.float64), M, M)
    assert rel_fro(eager.float().cpu().numpy(), ref) < 1e-2
    _replays_match(_capture(fn), [D], [eager])


def test_fp8_quantize_and_gemm_graph_replays(cuda):
    M, N, K = 512, 384, 256
    print("case graph fp8 quantize + gemm", flush=True)
    g = torch.Generator().manual_seed(7)
    A = torch.randn(M * K, generator=g).to(cuda)   # K-contiguous
    B = torch.randn(N * K, generator=g).to(cuda)
    qa = torch.empty(M * K, dtype=torch.uint8, device=cuda)
    qb = torch.empty(N * K, dtype=torch.uint8, device=cuda)
    sa = torch.empty(1, device=cuda)
    sb = torch.empty(1, device=cuda)
    wa = torch.empty(C.lib().ffgemm_fp8_workspace_size(), dtype=torch.uint8, device=cuda)
    wb = torch.empty_like(wa)
    D = torch.empty(M * N, device=cuda)
    s = lambda: torch.cuda.current_stream().cuda_stream

    def fn():
        C.check(C.lib().ffgemm_fp8_quantize(A.data_ptr(), M, K, K, 1, qa.data_ptr(), K, sa.data_ptr(),
                                            wa.data_ptr(), s()))
        C.check(C.lib().ffgemm_fp8_quantize(B.data_ptr(), N, K, K, 1, qb.data_ptr(), K, sb.data_ptr(),
                                            wb.data_ptr(), s()))
        tl.gemm("TN", M, N, K, qa, qb, D=D, lda=K, ldb=K, scaleA=sa, scaleB=sb, dtype_out=torch.float32)

    fn()
    torch.cuda.synchronize()
    eager = [D.clone(), qa.clone(), qb.clone(), sa.clone(), sb.clone()]
    _replays_match(_capture(fn), [D, qa, qb, sa, sb], eager)


def test_two_streams_concurrent(cuda):
    """Different GEMMs in flight on two streams at once equal the same GEMMs run
    one after the other (the library keeps no per-call device state)."""
    print("case two streams", flush=True)
    g = torch.Generator().manual_seed(8)
    jobs = []
    for (lay, M, N, K) in [("TN", 768, 1024, 512), ("NN", 640, 512, 384)]:
        ta, tb = lay[0] == "T", lay[1] == "T"
        A = (torch.randn(M * K, generator=g) / 16).to(torch.bfloat16).to(cuda)
        B = (torch.randn(K * N, generator=g) / 16).to(torch.bfloat16).to(cuda)
        jobs.append((lay, M, N, K, A, B))
    seq = [tl.gemm(l, M, N, K, A, B) for (l, M, N, K, A, B) in jobs]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = [torch.empty_like(x) for x in seq]
    for rep in range(4):
        for (l, M, N, K, A, B), st, o in zip(jobs, streams, outs):
            st.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(st):
                tl.gemm(l, M, N, K, A, B, D=o)
        torch.cuda.synchronize()
        for o, e in zip(outs, seq):
            assert torch.equal(o, e), f"round {rep}"
