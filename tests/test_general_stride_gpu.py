"""Problems off the 16-byte grid on the GPU: leading dimensions of any length
(the reference's `-p` padding gives ld = rows + pad, src/gemmMfma.cpp:210-217),
K not a multiple of the 16-byte vector for a K-contiguous operand, batch
strides of any length, and matrices that start one element past a 16-byte
boundary.  These run the general-stride kernel (element-wise loads and
stores); each is checked against the FP64 oracle on the same rounded inputs.

Tolerances: 16-bit per element |d - r| <= 1e-2 (|r| + rms(r)); fp32 and
bf16x3 rel-Frobenius <= 1e-5; fp8 rel-Frobenius <= 5e-5 against the FP64
product of the dequantised operands.
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

CODES = {"bf16": C.BF16, "f16": C.F16, "f32": C.F32, "x3": C.F32}
TORCH = {"bf16": torch.bfloat16, "f16": torch.float16, "f32": torch.float32, "x3": torch.float32}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _vals(x, name):
    return x if name in ("f32", "x3") else bits_to_f32(quantize_bits(x, name), name)


def _dev(x, name, device, shift):
    """x (float32, column-major with its ld) on the device, starting `shift`
    elements into its allocation (shift 1: off the 16-byte grid)."""
    assert x.dtype == np.float32
    if name in ("f32", "x3"):
        t = torch.from_numpy(np.concatenate([np.zeros(shift, np.float32), x])).to(device)
    else:
        bits = np.concatenate([np.zeros(shift, np.uint16), quantize_bits(x, name)])
        t = torch.from_numpy(bits.view(np.int16).copy()).to(device).view(TORCH[name])
    return t[shift:]


def _check(name, got, ref):
    if name in ("f32", "x3"):
        assert rel_fro(got, ref) <= 1e-5
    else:
        bad, worst = elem_ok(got, ref)
        assert bad == 0, worst


CASES = [
    # (M, N, K, pad, shift): pad elements past the stored rows of every matrix
    (67, 45, 37, 1, 0),       # odd ld, K off the vector
    (130, 70, 200, 3, 1),     # odd ld, every matrix one element off the grid
    (64, 64, 64, 1, 0),       # whole tiles, odd ld only
    (300, 260, 129, 0, 1),    # aligned ld, shifted pointers, K = 129
]


@pytest.mark.parametrize("name", ["bf16", "f16", "f32", "x3"])
@pytest.mark.parametrize("layout", ["NN", "NT", "TN", "TT"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c)))
def test_general_stride(cuda, name, layout, case):
    M, N, K, pad, shift = case
    _say(f"case general-stride {name} {layout} {case}")
    rng = np.random.default_rng(M * 7 + K)
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    lda, ldb, ldc = ra + pad, rb + pad, M + pad
    A = colmajor(rng, ra, ca, lda, kind="normal", scale=1 / np.sqrt(K))
    B = colmajor(rng, rb, cb, ldb, kind="normal", scale=1 / np.sqrt(K))
    Cm = colmajor(rng, M, N, ldc, kind="normal")
    comp = C.COMPUTE_BF16X3 if name == "x3" else C.COMPUTE_DEFAULT
    p = tl.make_problem(layout, CODES[name], M, N, K, lda, ldb, ldc, ldc, 0.75, -0.5, compute=comp)
    D = torch.full((ldc * N + shift,), float("nan"), dtype=TORCH[name], device=cuda)[shift:]
    tl.run(p, _dev(A, name, cuda, shift), _dev(B, name, cuda, shift), _dev(Cm, name, cuda, shift), D,
           workspace=None)
    torch.cuda.synchronize()
    ref = oracle.gemm_f64(layout, M, N, K, 0.75, _vals(A, name).astype(np.float64), lda,
                          _vals(B, name).astype(np.float64), ldb, -0.5, _vals(Cm, name).astype(np.float64), ldc, ldc)
    got = D.float().cpu().numpy()
    _check(name, extract(got, M, N, ldc), extract(ref, M, N, ldc))
    if pad:
        assert np.isnan(got.reshape(N, ldc)[:, M:]).all()      # padding rows untouched


@pytest.mark.parametrize("name", ["bf16", "f32"])
def test_general_stride_batched_and_workspace(cuda, name):
    """Batch 3 with odd batch strides, through ffgemm_run_ws (the bf16x3-free
    workspace path: no split for off-grid problems) and an explicit 128 x 128
    tile record."""
    M, N, K, batch = 150, 90, 77, 3
    layout = "TN"
    rng = np.random.default_rng(3)
    lda, ldb, ldc = K + 1, K + 3, M + 1
    bsA, bsB, bsC = lda * M + 5, ldb * N + 1, ldc * N + 3
    A = (rng.standard_normal(bsA * batch) / np.sqrt(K)).astype(np.float32)
    B = (rng.standard_normal(bsB * batch) / np.sqrt(K)).astype(np.float32)
    Cm = rng.standard_normal(bsC * batch).astype(np.float32)
    p = tl.make_problem(layout, CODES[name], M, N, K, lda, ldb, ldc, ldc, 1.25, 0.5, batch=batch,
                        bsA=bsA, bsB=bsB, bsC=bsC, bsD=bsC)
    plan = tl.query_tile_ws(p)
    assert plan.split_k == 1 and (plan.waves_m, plan.waves_n, plan.lds_buffers) == (2, 2, 2)
    D = torch.full((bsC * batch,), float("nan"), dtype=TORCH[name], device=cuda)
    ws = torch.empty(1 << 20, dtype=torch.uint8, device=cuda)
    tl.run(p, _dev(A, name, cuda, 0), _dev(B, name, cuda, 0), _dev(Cm, name, cuda, 0), D, tile=(128, 128),
           workspace=ws)
    torch.cuda.synchronize()
    got = D.float().cpu().numpy()
    Av, Bv, Cv = _vals(A, name), _vals(B, name), _vals(Cm, name)
    for b in range(batch):
        ref = oracle.gemm_f64(layout, M, N, K, 1.25, Av[b * bsA:].astype(np.float64), lda,
                              Bv[b * bsB:].astype(np.float64), ldb, 0.5, Cv[b * bsC:(b + 1) * bsC].astype(np.float64),
                              ldc, ldc)
        _check(name, extract(got[b * bsC:(b + 1) * bsC], M, N, ldc), extract(ref, M, N, ldc))


def test_general_stride_bf16x3_planes(cuda):
    """bf16x3 with a workspace splits into aligned planes; an odd ldd makes the
    plane GEMM itself run the general-stride kernel."""
    M, N, K = 200, 120, 300
    rng = np.random.default_rng(5)
    lda, ldb, ldc = M + 1, K + 1, M + 1
    A = colmajor(rng, M, K, lda, kind="normal")
    B = colmajor(rng, K, N, ldb, kind="normal")
    p = tl.make_problem("NN", C.F32, M, N, K, lda, ldb, ldc, ldc, compute=C.COMPUTE_BF16X3)
    assert tl.workspace_size(p) > 0
    D = torch.full((ldc * N,), float("nan"), device=cuda)
    tl.run(p, torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda), None, D)
    torch.cuda.synchronize()
    ref = oracle.gemm_f64("NN", M, N, K, 1.0, A.astype(np.float64), lda, B.astype(np.float64), ldb, 0.0, None, ldc,
                          ldc)
    assert rel_fro(extract(D.cpu().numpy(), M, N, ldc), extract(ref, M, N, ldc)) <= 1e-5


@pytest.mark.parametrize("K,pad", [(40, 0), (100, 3), (257, 1)])
def test_general_stride_fp8(cuda, K, pad):
    """fp8 e4m3 TN with K off the 16-element vector and odd leading dimensions."""
    M, N = 136, 72
    _say(f"case general-stride fp8 K={K} pad={pad}")
    rng = np.random.default_rng(K)
    lda, ldb, ldd = K + pad, K + pad, M + pad
    A = rng.standard_normal(M * lda).astype(np.float32)
    B = rng.standard_normal(N * ldb).astype(np.float32)
    sa, sb = np.float32(oracle.fp8_scale(np.abs(A).max())), np.float32(oracle.fp8_scale(np.abs(B).max()))
    qa, qb = oracle.f32_to_e4m3_bits(A, sa), oracle.f32_to_e4m3_bits(B, sb)
    dsa, dsb = torch.tensor([sa], device=cuda), torch.tensor([sb], device=cuda)
    p = tl.make_problem("TN", C.F8E4M3, M, N, K, lda, ldb, ldd, ldd, scaleA=dsa, scaleB=dsb, dtype_out=C.F32)
    assert tl.query_tile(p).mt_m == 128
    D = torch.full((ldd * N,), float("nan"), device=cuda)
    tl.run(p, torch.from_numpy(qa).to(cuda), torch.from_numpy(qb).to(cuda), None, D, workspace=None)
    torch.cuda.synchronize()
    ref = oracle.gemm_f64("TN", M, N, K, 1.0, oracle.e4m3_bits_to_f32(qa).astype(np.float64) / sa, lda,
                          oracle.e4m3_bits_to_f32(qb).astype(np.float64) / sb, ldb, 0.0, None, ldd, ldd)
    assert rel_fro(extract(D.cpu().numpy(), M, N, ldd), extract(ref, M, N, ldd)) <= 5e-5
