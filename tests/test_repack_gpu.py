"""K-major repack on the workspace path: bf16/fp16 problems in layouts NN, NT
and TT copy their M/N-contiguous operands K-contiguous into the workspace
(kpack16_kernel) and run as TN.  Each case is checked against the FP64 oracle
on the same rounded inputs, per element |d - r| <= 1e-2 (|r| + rms(r)).

The forced cases (tile record repack 2) cover ragged rows (a short last
8-row block in the copy), K not a multiple of the copy's 256-wide workgroup,
batches with padded strides, beta with C, a caller's macro tile, and a
split-K plan of the repacked problem; the 8192^3 cases run the plan the
library picks by itself (NN, TT) or forced (NT), with sampled columns.
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

CODES = {"bf16": C.BF16, "f16": C.F16}
TORCH = {"bf16": torch.bfloat16, "f16": torch.float16}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev(bits, name, device):
    return torch.from_numpy(bits.view(np.int16).copy()).to(device).view(TORCH[name])


# (M, N, K, batch, tile): tile None = the library's tile for the repacked problem
CASES = [
    (997, 515, 264, 1, None),              # ragged rows, K past one 256-wide copy block
    (256, 256, 64, 1, None),               # whole tiles, one K-tile
    (300, 200, 136, 2, None),              # batch with padded strides
    (64, 1040, 512, 1, (128, 128)),        # a caller's tile
    (520, 480, 1024, 1, (128, 128, 0, 0, 0, 0, 2)),  # the repacked problem split in two K slices
]


@pytest.mark.parametrize("name", ["bf16", "f16"])
@pytest.mark.parametrize("layout", ["NN", "NT", "TT"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[:4])) + ("" if c[4] is None else "_tile"))
def test_repack_forced(cuda, name, layout, case):
    M, N, K, batch, tile = case
    _say(f"case repack {name} {layout} M={M} N={N} K={K} batch={batch} tile={tile}")
    rng = np.random.default_rng(M + 3 * N + K)
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    pad = 8 if batch > 1 else 0
    lda, ldb, ldc = (ra + 7) // 8 * 8 + pad, (rb + 7) // 8 * 8 + pad, (M + 7) // 8 * 8 + pad
    bsA, bsB, bsC = lda * ca + 8 * pad, ldb * cb + 8 * pad, ldc * N + 8 * pad
    A = np.concatenate([np.concatenate([colmajor(rng, ra, ca, lda, kind="normal", scale=1 / np.sqrt(K)),
                                        np.zeros(bsA - lda * ca, np.float32)]) for _ in range(batch)])
    B = np.concatenate([np.concatenate([colmajor(rng, rb, cb, ldb, kind="normal", scale=1 / np.sqrt(K)),
                                        np.zeros(bsB - ldb * cb, np.float32)]) for _ in range(batch)])
    Cm = rng.standard_normal(bsC * batch).astype(np.float32)
    qa, qb, qc = quantize_bits(A, name), quantize_bits(B, name), quantize_bits(Cm, name)
    p = tl.make_problem(layout, CODES[name], M, N, K, lda, ldb, ldc, ldc, 0.75, -0.5, batch=batch,
                        bsA=bsA, bsB=bsB, bsC=bsC, bsD=bsC)
    rec = tuple(tile or (0, 0)) + (0,) * (7 - len(tile or (0, 0))) + (2,)
    plan = tl.query_tile_ws(p, tl.make_tile(rec))
    assert plan.repack == 2
    if tile and len(tile) > 6:
        assert plan.split_k == tile[6]
    rows_a = M if layout[0] == "N" else 0
    rows_b = N if layout[1] == "T" else 0
    need = tl.workspace_size(p, rec)
    assert need >= (rows_a * K * 2 * batch + 255) // 256 * 256 + (rows_b * K * 2 * batch + 255) // 256 * 256
    D = torch.full((bsC * batch,), float("nan"), dtype=TORCH[name], device=cuda)
    tl.run(p, _dev(qa, name, cuda), _dev(qb, name, cuda), _dev(qc, name, cuda), D, tile=rec)
    torch.cuda.synchronize()
    got = D.float().cpu().numpy()
    Av, Bv, Cv = bits_to_f32(qa, name), bits_to_f32(qb, name), bits_to_f32(qc, name)
    for b in range(batch):
        ref = oracle.gemm_f64(layout, M, N, K, 0.75, Av[b * bsA:].astype(np.float64), lda,
                              Bv[b * bsB:].astype(np.float64), ldb, -0.5,
                              Cv[b * bsC:(b + 1) * bsC].astype(np.float64), ldc, ldc)
        gb = got[b * bsC:(b + 1) * bsC]
        bad, worst = elem_ok(extract(gb, M, N, ldc), extract(ref, M, N, ldc))
        assert bad == 0, (b, worst)
        if ldc > M:
            assert np.isnan(gb[:ldc * N].reshape(N, ldc)[:, M:]).all()  # padding rows untouched


@pytest.mark.parametrize("layout,forced", [("NN", False), ("TT", False), ("NT", True)])
def test_repack_8192_sampled_columns(cuda, layout, forced):
    """bf16 at 8192^3: NN and TT repack by the library's own choice, NT forced
    (its tuning-table entry keeps the native engine).  24 sampled output
    columns against the FP64 oracle, every column sum against the FP64 sums."""
    M = N = K = 8192
    _say(f"case repack 8192 bf16 {layout} forced={forced}")
    p = tl.make_problem(layout, C.BF16, M, N, K, M if layout[0] == "N" else K, K if layout[1] == "N" else N, M, M)
    rec = (0, 0, 0, 0, 0, 0, 0, 2) if forced else None
    plan = tl.query_tile_ws(p, tl.make_tile(rec)) if forced else tl.query_tile_ws(p)
    assert plan.repack == 2, "the plan repacks"
    g = torch.Generator(device="cpu").manual_seed(2)
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    A16 = (torch.randn(ra * ca, generator=g) / np.sqrt(K)).to(torch.bfloat16)
    B16 = (torch.randn(rb * cb, generator=g) / np.sqrt(K)).to(torch.bfloat16)
    D = torch.empty(M * N, dtype=torch.bfloat16, device=cuda)
    tl.run(p, A16.to(cuda), B16.to(cuda), None, D, tile=rec)
    torch.cuda.synchronize()
    Dh = D.float().cpu().numpy().reshape(N, M)
    A64 = A16.double().numpy()
    B64 = B16.double().numpy()
    opA = A64.reshape(ca, ra).T if layout[0] == "N" else A64.reshape(ca, ra)      # (M, K)
    opB = B64.reshape(cb, rb).T if layout[1] == "N" else B64.reshape(cb, rb)      # (K, N)
    cols = np.random.default_rng(1).choice(N, 24, replace=False)
    sub = np.ascontiguousarray(opB[:, cols])
    ref_cols = oracle.gemm_f64("TN", M, len(cols), K, 1.0, np.ascontiguousarray(opA).ravel(), K,
                               sub.T.ravel(), K).reshape(len(cols), M).T
    bad, worst = elem_ok(Dh[cols].T, ref_cols)
    assert bad == 0, worst
    colsum_ref = opA.sum(axis=0) @ opB
    assert rel_fro(Dh.astype(np.float64).sum(axis=1), colsum_ref) < 2e-3
