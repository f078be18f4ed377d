"""Split-K on the GPU (ffgemm_run_ws with a workspace): K cut into S slices
whose fp32 partial products are summed by the reduction pass.  Every dtype
and layout the slices can run, ragged M and N, K not a multiple of the slice
(a shorter last slice), padded leading dimensions, alpha / beta with C,
against the FP64 oracle on the same (rounded) inputs.

Tolerances: 16-bit per element |d - r| <= 1e-2 (|r| + rms(r)); fp32
rel-Frobenius <= 1e-5; fp8 rel-Frobenius <= 5e-5 against the FP64 product of
the dequantised operands (the GEMM's own error; see test_gemm_gpu.py).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

import oracle
import tensilelite_amd as tl
from tensilelite_amd import _capi as C
from tests._util import bits_to_f32, colmajor, elem_ok, extract, op_dims, quantize_bits, rel_fro

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev(x, name, device):
    if name == "f32":
        return torch.from_numpy(x).to(device)
    bits = quantize_bits(x, name)
    return torch.from_numpy(bits.view(np.int16).copy()).to(device).view(torch.bfloat16 if name == "bf16" else torch.float16)


def _vals(x, name):
    return x if name == "f32" else bits_to_f32(quantize_bits(x, name), name)


@pytest.mark.parametrize("name", ["bf16", "f16", "f32"])
@pytest.mark.parametrize("layout", ["NN", "NT", "TN", "TT"])
@pytest.mark.parametrize("split", [2, 3, 5])
def test_split_k_layouts(cuda, name, layout, split):
    M, N, K = 200, 136, 1160          # slices of 256 (or 384 / 256) K, the last shorter
    tile = (128, 128) if name != "f32" else (64, 64)
    _say(f"case split-K {name} {layout} S={split}")
    rng = np.random.default_rng(split * 7 + len(layout))
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    lda, ldb, ldc = ra + 8, rb + 8, M + 8
    A = colmajor(rng, ra, ca, lda, kind="normal", scale=1 / np.sqrt(K))
    B = colmajor(rng, rb, cb, ldb, kind="normal", scale=1 / np.sqrt(K))
    Cm = colmajor(rng, M, N, ldc, kind="normal")
    p = tl.make_problem(layout, {"bf16": C.BF16, "f16": C.F16, "f32": C.F32}[name], M, N, K, lda, ldb, ldc, ldc,
                        0.75, -0.5)
    rec = tile + (0, 0, 0, 0, split)
    t = tl.query_tile_ws(p, rec)
    assert t.split_k == split
    ws = torch.empty((split * ((M + 3) // 4 * 4) * N * 4 + 255) // 256 * 256, dtype=torch.uint8, device=cuda)
    D = torch.full((ldc * N,), float("nan"), dtype=torch.float32 if name == "f32" else
                   (torch.bfloat16 if name == "bf16" else torch.float16), device=cuda)
    tl.run(p, _dev(A, name, cuda), _dev(B, name, cuda), _dev(Cm, name, cuda), D, tile=rec, workspace=ws)
    torch.cuda.synchronize()
    ref = oracle.gemm_f64(layout, M, N, K, 0.75, _vals(A, name).astype(np.float64), lda,
                          _vals(B, name).astype(np.float64), ldb, -0.5, _vals(Cm, name).astype(np.float64), ldc, ldc)
    got = D.float().cpu().numpy()
    g, r = extract(got, M, N, ldc), extract(ref, M, N, ldc)
    if name == "f32":
        assert rel_fro(g, r) <= 1e-5
    else:
        bad, worst = elem_ok(g, r)
        assert bad == 0, worst
    assert np.isnan(got.reshape(N, ldc)[:, M:]).all()      # padding rows untouched


def test_split_k_fp8(cuda):
    M, N, K = 264, 200, 2304
    _say("case split-K fp8")
    rng = np.random.default_rng(8)
    A = rng.standard_normal(M * K).astype(np.float32)   # K-contiguous
    B = rng.standard_normal(N * K).astype(np.float32)
    sa, sb = np.float32(oracle.fp8_scale(np.abs(A).max())), np.float32(oracle.fp8_scale(np.abs(B).max()))
    qa, qb = oracle.f32_to_e4m3_bits(A, sa), oracle.f32_to_e4m3_bits(B, sb)
    # the problem holds the scales' device pointers: keep the tensors alive
    dsa, dsb = torch.tensor([sa], device=cuda), torch.tensor([sb], device=cuda)
    p = tl.make_problem("TN", C.F8E4M3, M, N, K, K, K, M, M, scaleA=dsa, scaleB=dsb, dtype_out=C.F32)
    rec = (128, 128, 0, 0, 0, 0, 3)
    ws = torch.empty(3 * M * N * 4 + 256, dtype=torch.uint8, device=cuda)
    D = torch.empty(M * N, device=cuda)
    tl.run(p, torch.from_numpy(qa).to(cuda), torch.from_numpy(qb).to(cuda), None, D, tile=rec, workspace=ws)
    torch.cuda.synchronize()
    ref = oracle.gemm_f64("TN", M, N, K, 1.0, oracle.e4m3_bits_to_f32(qa).astype(np.float64) / sa, K,
                          oracle.e4m3_bits_to_f32(qb).astype(np.float64) / sb, K)
    assert rel_fro(D.cpu().numpy(), ref) <= 5e-5


def test_split_k_bf16_truncate_and_beta_zero_nan_c(cuda):
    """The reduction's bf16 output honours truncation, and beta = 0 never reads C."""
    M, N, K = 96, 80, 1024
    rng = np.random.default_rng(9)
    A = colmajor(rng, K, M, kind="int")       # TN, small integers: exact sums
    B = colmajor(rng, K, N, kind="int")
    p = tl.make_problem("TN", C.BF16, M, N, K, K, K, M, M, 0.3, 0.0, round=C.ROUND_TRUNCATE)
    Cn = torch.full((M * N,), float("nan"), dtype=torch.bfloat16, device=cuda)
    D = torch.empty(M * N, dtype=torch.bfloat16, device=cuda)
    rec = (64, 64, 0, 0, 0, 0, 4)
    ws = torch.empty(4 * M * N * 4, dtype=torch.uint8, device=cuda)
    tl.run(p, _dev(A, "bf16", cuda), _dev(B, "bf16", cuda), Cn, D, tile=rec, workspace=ws)
    torch.cuda.synchronize()
    sums = oracle.gemm_f64("TN", M, N, K, 1.0, A.astype(np.float64), K, B.astype(np.float64), K)
    want = oracle.f32_to_bf16_bits((np.float32(0.3) * sums.astype(np.float32)).astype(np.float32), truncate=True)
    got = D.view(torch.int16).cpu().numpy().view(np.uint16)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("name,shape", [("f32", (480, 1024, 1024)), ("f32", (512, 480, 512)),
                                        ("bf16", (256, 256, 8192)), ("bf16", (64, 64, 65536))])
def test_split_k_library_choice(cuda, name, shape):
    """Problems the library itself splits (no tile given, workspace="auto"):
    the plan has split_k > 1, and the result meets the dtype's bar."""
    M, N, K = shape
    _say(f"case auto split {name} {shape}")
    code = C.F32 if name == "f32" else C.BF16
    p = tl.make_problem("TN", code, M, N, K, K, K, M, M)
    assert tl.query_tile_ws(p).split_k > 1
    rng = np.random.default_rng(K)
    A = rng.standard_normal(M * K).astype(np.float32)
    B = rng.standard_normal(N * K).astype(np.float32)
    D = tl.gemm("TN", M, N, K, _dev(A, name, cuda), _dev(B, name, cuda))
    torch.cuda.synchronize()
    ref = oracle.gemm_f64("TN", M, N, K, 1.0, _vals(A, name).astype(np.float64), K, _vals(B, name).astype(np.float64), K)
    if name == "f32":
        assert rel_fro(D.cpu().numpy(), ref) <= 1e-5
    else:
        bad, worst = elem_ok(D.float().cpu().numpy(), ref)
        assert bad == 0, worst


def test_split_k_graph_replay(cuda):
    """A split plan (two launches and the reduction) captured into a hipGraph
    replays to the eager result, with eager work between replays."""
    M, N, K = 480, 512, 4096
    p = tl.make_problem("TN", C.F32, M, N, K, K, K, M, M)
    rec = (64, 64, 0, 0, 0, 0, 6)
    g = torch.Generator(device="cpu").manual_seed(3)
    A = torch.randn(M * K, generator=g).to(cuda)
    B = torch.randn(N * K, generator=g).to(cuda)
    ws = torch.empty(6 * M * N * 4 + 256, dtype=torch.uint8, device=cuda)
    D0 = torch.empty(M * N, device=cuda)
    tl.run(p, A, B, None, D0, tile=rec, workspace=ws)
    D = torch.zeros(M * N, device=cuda)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        tl.run(p, A, B, None, D, tile=rec, workspace=ws, stream=side)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        tl.run(p, A, B, None, D, tile=rec, workspace=ws)
    for i in range(3):
        D.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(D, D0), i
        (A * 2).sum().item()   # eager work between replays


@pytest.mark.parametrize("name,shape,beta", [("bf16", (4608, 4608, 4608), 0.5), ("f16", (4608, 4608, 4608), 0.0),
                                             ("bf16", (69120, 256, 2048), 1.0)])
def test_tail_split(cuda, name, shape, beta):
    """Problems whose 256 x 256 tiles leave a short last pass run, on the
    workspace path with no tile given, as a data-parallel part plus a split-K
    remainder (cut between tile columns for the square shape, between tile rows
    for the tall one).  Sampled output columns (both parts; every row) against
    the FP64 oracle, alpha / beta with C."""
    M, N, K = shape
    code = C.BF16 if name == "bf16" else C.F16
    tdt = torch.bfloat16 if name == "bf16" else torch.float16
    p = tl.make_problem("TN", code, M, N, K, K, K, M, M, 1.25, beta)
    assert tl.workspace_size(p) > 0 and tl.query_tile_ws(p).split_k == 1
    _say(f"case tail split {name} {shape} beta={beta}")
    g = torch.Generator(device="cpu").manual_seed(M + K)
    A = (torch.randn(M * K, generator=g) / np.sqrt(K)).to(tdt)
    B = (torch.randn(N * K, generator=g) / np.sqrt(K)).to(tdt)
    Cm = torch.randn(M * N, generator=g).to(tdt)
    D = torch.full((M * N,), float("nan"), dtype=tdt, device=cuda)
    tl.run(p, A.to(cuda), B.to(cuda), Cm.to(cuda), D)
    torch.cuda.synchronize()
    rng = np.random.default_rng(N)
    cols = np.unique(np.concatenate([rng.choice(N, 3, replace=False), [N - 1, 0]]))
    Bc = B.double().numpy().reshape(N, K)[cols]
    ref = oracle.gemm_f64("TN", M, len(cols), K, 1.25, A.double().numpy(), K, Bc.ravel(), K).reshape(len(cols), M)
    ref += beta * Cm.double().numpy().reshape(N, M)[cols]
    got = D.float().cpu().numpy().reshape(N, M)[cols]
    bad, worst = elem_ok(got, ref)
    assert bad == 0, worst


def test_tail_split_graph_replay(cuda):
    """A tail-split plan (the data-parallel launch, the split slices and the
    reduction) captured into a hipGraph replays to the eager result, with
    eager work between replays."""
    M = N = K = 4608
    p = tl.make_problem("TN", C.BF16, M, N, K, K, K, M, M)
    n = tl.workspace_size(p)
    assert n > 0
    g = torch.Generator(device="cpu").manual_seed(7)
    A = (torch.randn(M * K, generator=g) / np.sqrt(K)).to(torch.bfloat16).to(cuda)
    B = (torch.randn(N * K, generator=g) / np.sqrt(K)).to(torch.bfloat16).to(cuda)
    ws = torch.empty(n, dtype=torch.uint8, device=cuda)
    D0 = torch.empty(M * N, dtype=torch.bfloat16, device=cuda)
    tl.run(p, A, B, None, D0, workspace=ws)
    D = torch.zeros(M * N, dtype=torch.bfloat16, device=cuda)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        tl.run(p, A, B, None, D, workspace=ws, stream=side)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        tl.run(p, A, B, None, D, workspace=ws)
    for i in range(3):
        D.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(D, D0), i
        (A.float() * 2).sum().item()   # eager work between replays
