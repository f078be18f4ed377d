"""GPU parity of the gemmMfma path (through the C ABI) against the oracle.

Tolerances:
  fp16/bf16 inputs: per element |d - r| <= 1e-2 * (|r| + rms(r)) vs the FP64
      oracle on the same (rounded) inputs;
  fp32 (fp32 MFMA) and bf16x3: relative Frobenius error <= 1e-5 vs FP64;
  fp8 e4m3: quantised bytes bit-exact vs the oracle's e4m3 encoder, GEMM
      relative Frobenius <= 5e-5 vs FP64 on the dequantised operands;
  fp64: relative Frobenius <= 1e-12 vs numpy float64.
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

TILES = [(256, 256), (256, 128), (128, 256), (128, 128), (64, 64)]
LAYOUTS = ["NN", "NT", "TN", "TT"]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev16(bits, name, device):
    t = torch.from_numpy(bits.view(np.int16).copy()).to(device)
    return t.view(torch.bfloat16 if name == "bf16" else torch.float16)


def _f64_ref(layout, M, N, K, A, lda, B, ldb, alpha=1.0, beta=0.0, Cm=None, ldc=None, ldd=None,
             batch=1, bsA=0, bsB=0, bsC=0, bsD=0):
    return oracle.gemm_f64(layout, M, N, K, alpha, A.astype(np.float64), lda, B.astype(np.float64), ldb,
                           beta, None if Cm is None else Cm.astype(np.float64), ldc, ldd, batch, bsA, bsB,
                           bsC, bsD)


@pytest.mark.parametrize("name", ["bf16", "f16"])
@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("tile", TILES)
def test_16bit_layouts_tiles_edges(cuda, name, layout, tile):
    # M, N not multiples of any tile; K = 200 leaves a partial K-tile of 8.
    M, N, K = 328, 272, 200
    _say(f"case {name} {layout} tile={tile} M={M} N={N} K={K}")
    rng = np.random.default_rng(11)
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    A = colmajor(rng, ra, ca, kind="normal", scale=1 / np.sqrt(K))
    B = colmajor(rng, rb, cb, kind="normal", scale=1 / np.sqrt(K))
    Ab, Bb = quantize_bits(A, name), quantize_bits(B, name)
    Ar, Br = bits_to_f32(Ab, name), bits_to_f32(Bb, name)
    D = tl.gemm(layout, M, N, K, _dev16(Ab, name, cuda), _dev16(Bb, name, cuda), tile=tile)
    torch.cuda.synchronize()
    got = D.float().cpu().numpy()
    ref = _f64_ref(layout, M, N, K, Ar, ra, Br, rb)
    bad, worst = elem_ok(got, ref)
    assert bad == 0, f"{bad} elements off (worst scaled error {worst:.3g})"


@pytest.mark.parametrize("name", ["bf16", "f16"])
@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("tile", [(256, 256), (128, 128), (64, 64)])
def test_16bit_ragged_m_n_padded_ld(cuda, name, layout, tile):
    """M and N that are not multiples of the 8-element vector, with padded
    leading dimensions: the last vector of an M/N-contiguous operand ends past
    the matrix and must still be read (regression: it read as zero)."""
    M, N, K = 65, 37, 136
    _say(f"case ragged {name} {layout} tile={tile}")
    rng = np.random.default_rng(12)
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    lda, ldb, ldd = (ra + 7) // 8 * 8, (rb + 7) // 8 * 8, (M + 7) // 8 * 8
    A = colmajor(rng, ra, ca, lda, kind="normal")
    B = colmajor(rng, rb, cb, ldb, kind="normal")
    A, B = A[: lda * (ca - 1) + ra], B[: ldb * (cb - 1) + rb]   # no slack after the last element
    Ab, Bb = quantize_bits(A, name), quantize_bits(B, name)
    D = tl.gemm(layout, M, N, K, _dev16(Ab, name, cuda), _dev16(Bb, name, cuda), lda=lda, ldb=ldb, ldd=ldd,
                tile=tile)
    torch.cuda.synchronize()
    ref = _f64_ref(layout, M, N, K, bits_to_f32(Ab, name), lda, bits_to_f32(Bb, name), ldb, ldd=ldd)
    bad, worst = elem_ok(extract(D.float().cpu().numpy(), M, N, ldd), extract(ref, M, N, ldd))
    assert bad == 0, worst


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("tile", [(256, 256), (128, 128), (64, 64)])
@pytest.mark.parametrize("compute", ["f32", "bf16x3", "bf16x3_regsplit"])
def test_fp32_accuracy_forms(cuda, layout, tile, compute):
    """fp32 MFMA and both bf16x3 forms (hi/lo planes in a workspace, or the
    split in registers) within 1e-5 of FP64."""
    M, N, K = 200, 264, 132
    _say(f"case {compute} {layout} tile={tile}")
    rng = np.random.default_rng(2)
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    A = colmajor(rng, ra, ca, kind="normal")
    B = colmajor(rng, rb, cb, kind="normal")
    comp = C.COMPUTE_DEFAULT if compute == "f32" else C.COMPUTE_BF16X3
    D = tl.gemm(layout, M, N, K, torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda),
                compute=comp, tile=tile, workspace=None if compute == "bf16x3_regsplit" else "auto")
    torch.cuda.synchronize()
    err = rel_fro(D.cpu().numpy(), _f64_ref(layout, M, N, K, A, ra, B, rb))
    _say(f"rel_fro={err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("tile", [(128, 256), (256, 128)])
def test_bf16x3_regsplit_rect_tiles(cuda, layout, tile):
    """The in-register bf16x3 split on the rectangular tiles, with M and N
    past the first 128 rows / columns of a tile (both halves of a 256 side
    valid, ragged ends), padded leading dimensions and alpha / beta with C."""
    M, N, K = 300, 520, 200
    _say(f"case bf16x3_regsplit {layout} tile={tile}")
    rng = np.random.default_rng(13)
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    lda, ldb, ldc = ra + 4, rb + 8, M + 4
    A = colmajor(rng, ra, ca, lda, kind="normal")
    B = colmajor(rng, rb, cb, ldb, kind="normal")
    Cm = colmajor(rng, M, N, ldc, kind="normal")
    D = tl.gemm(layout, M, N, K, torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda),
                torch.from_numpy(Cm).to(cuda), lda=lda, ldb=ldb, ldc=ldc, ldd=ldc, alpha=0.75, beta=-0.5,
                compute=C.COMPUTE_BF16X3, tile=tile, workspace=None)
    torch.cuda.synchronize()
    ref = _f64_ref(layout, M, N, K, A, lda, B, ldb, 0.75, -0.5, Cm, ldc, ldc)
    got = D.cpu().numpy()
    err = rel_fro(extract(got, M, N, ldc), extract(ref, M, N, ldc))
    _say(f"rel_fro={err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("layout", ["NN", "TN"])
def test_bf16x3_regsplit_default_tile_4096(cuda, layout):
    """bf16x3 with neither a workspace nor a tile at 4096^3, where the size
    pick says 256 x 256: the library runs the in-register split on 128 x 256
    instead (that kernel spills at 256 x 256), which ffgemm_query_tile reports;
    8 sampled output columns within 1e-5 of FP64 (normal(0, 1), seed 2)."""
    S = 4096
    _say(f"case bf16x3 regsplit default tile {layout} 4096")
    p = tl.make_problem(layout, C.F32, S, S, S, S, S, S, S, compute=C.COMPUTE_BF16X3)
    t = tl.query_tile(p)
    assert (t.mt_m, t.mt_n) == (128, 256)
    g = torch.Generator(device="cpu").manual_seed(2)
    A = torch.randn(S * S, generator=g)
    B = torch.randn(S * S, generator=g)
    D = torch.empty(S * S, device=cuda)
    tl.run(p, A.to(cuda), B.to(cuda), None, D, workspace=None)
    torch.cuda.synchronize()
    cols = np.random.default_rng(3).choice(S, 8, replace=False)
    Bc = B.numpy().reshape(S, S)[cols].astype(np.float64)                 # column n of B (K-contiguous)
    Am = A.numpy().reshape(S, S)
    At = (Am.T if layout == "NN" else Am).astype(np.float64).ravel()      # op(A) with K contiguous
    ref = oracle.gemm_f64("TN", S, len(cols), S, 1.0, At, S, Bc.ravel(), S).reshape(len(cols), S)
    got = D.cpu().numpy().reshape(S, S)[cols]
    err = rel_fro(got, ref)
    _say(f"rel_fro={err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("name", ["bf16", "f16", "f32"])
def test_alpha_beta_padding(cuda, name):
    M, N, K = 256, 192, 128
    lda, ldb, ldc, ldd = K + 32, K + 64, M + 16, M + 8  # TN with padded rows, as the reference
    _say(f"case alpha/beta {name}")
    rng = np.random.default_rng(5)
    A = colmajor(rng, K, M, lda)
    B = colmajor(rng, K, N, ldb)
    Cm = colmajor(rng, M, N, ldc)
    Dinit = np.full(ldd * N, 7.0, dtype=np.float32)
    if name == "f32":
        dev = lambda x: torch.from_numpy(x).to(cuda)
        Ar, Br, Cr = A, B, Cm
    else:
        bits = [quantize_bits(x, name) for x in (A, B, Cm, Dinit)]
        dev = None
        Ar, Br, Cr = (bits_to_f32(b, name) for b in bits[:3])
    alpha, beta = 0.314, 0.2718
    if name == "f32":
        Ad, Bd, Cd, Dd = dev(A), dev(B), dev(Cm), dev(Dinit)
    else:
        Ad, Bd, Cd, Dd = (_dev16(b, name, cuda) for b in bits)
    tl.gemm("TN", M, N, K, Ad, Bd, Cd, Dd, lda=lda, ldb=ldb, ldc=ldc, ldd=ldd, alpha=alpha, beta=beta)
    torch.cuda.synchronize()
    got = Dd.float().cpu().numpy()
    ref = _f64_ref("TN", M, N, K, Ar, lda, Br, ldb, alpha, beta, Cr, ldc, ldd)
    g, r = extract(got, M, N, ldd), extract(ref, M, N, ldd)
    if name in ("f32", "x3"):
        assert rel_fro(g, r) <= 1e-5
    else:
        bad, worst = elem_ok(g, r)
        assert bad == 0, worst
    # padding rows of D are untouched
    pad = got.reshape(N, ldd)[:, M:]
    assert np.all(pad == 7.0)


def test_beta_zero_never_reads_c(cuda):
    M, N, K = 128, 128, 64
    _say("case beta0 NaN C")
    rng = np.random.default_rng(6)
    A = colmajor(rng, M, K)
    B = colmajor(rng, K, N)
    Cm = torch.full((M * N,), float("nan"), device=cuda)
    D = tl.gemm("NN", M, N, K, torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda), Cm, beta=0.0)
    torch.cuda.synchronize()
    assert torch.isfinite(D).all()


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("tile", [(256, 256), (128, 128), (64, 64)])
def test_16bit_beta_zero_never_reads_c(cuda, layout, tile):
    """beta = 0 on every 16-bit engine: a NaN-filled C must not reach D."""
    M, N, K = 264, 200, 96
    _say(f"case bf16 beta0 NaN C {layout} tile={tile}")
    rng = np.random.default_rng(12)
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    Ab = quantize_bits(colmajor(rng, ra, ca, kind="normal", scale=1 / np.sqrt(K)), "bf16")
    Bb = quantize_bits(colmajor(rng, rb, cb, kind="normal", scale=1 / np.sqrt(K)), "bf16")
    Cm = torch.full((M * N,), float("nan"), dtype=torch.bfloat16, device=cuda)
    D = tl.gemm(layout, M, N, K, _dev16(Ab, "bf16", cuda), _dev16(Bb, "bf16", cuda), Cm, beta=0.0, tile=tile)
    torch.cuda.synchronize()
    got = D.float().cpu().numpy()
    assert np.isfinite(got).all()
    ref = _f64_ref(layout, M, N, K, bits_to_f32(Ab, "bf16"), ra, bits_to_f32(Bb, "bf16"), rb)
    bad, worst = elem_ok(got, ref)
    assert bad == 0, worst


def test_tn_operand_past_2gib(cuda):
    """bf16 TN with A of 140000 rows x K = 8192 (2.29 GB): rows whose bytes lie
    past 2 GiB (row 131072 on) are computed from the rebased per-tile buffer;
    sampled rows on both sides of the boundary are checked against the oracle."""
    M, N, K = 140000, 256, 8192
    _say(f"case TN operand past 2 GiB M={M} N={N} K={K}")
    g = torch.Generator(device=cuda).manual_seed(13)
    A = torch.empty(M * K, dtype=torch.bfloat16, device=cuda)
    for r0 in range(0, M, 20000):  # fill in slices to bound the fp32 temporary
        r1 = min(M, r0 + 20000)
        A[r0 * K:r1 * K] = (torch.randn((r1 - r0) * K, generator=g, device=cuda) / K ** 0.5).to(torch.bfloat16)
    B = (torch.randn(N * K, generator=g, device=cuda) / K ** 0.5).to(torch.bfloat16)
    D = tl.gemm("TN", M, N, K, A, B)
    torch.cuda.synchronize()
    rows = [0, 1, 255, 256, 131071, 131072, 131073, 131327, 131328, 139743, 139999]
    rows += torch.randint(0, M, (21,), generator=torch.Generator().manual_seed(14)).tolist()
    rows = sorted(set(rows))
    Ar = A.view(M, K)[rows].float().cpu().numpy().astype(np.float64).ravel()
    Bh = B.float().cpu().numpy().astype(np.float64)
    ref = oracle.gemm_f64("TN", len(rows), N, K, 1.0, Ar, K, Bh, K).reshape(N, len(rows))
    got = D.view(N, M)[:, rows].float().cpu().numpy()
    bad, worst = elem_ok(got, ref)
    assert bad == 0, worst
    del A, B, D
    torch.cuda.empty_cache()


def test_nn_m_contiguous_operand_past_2gib(cuda):
    """bf16 NN with an M-contiguous A of K = 8192 rows x lda = 140000 (2.29 GB):
    the library runs it in two K chunks accumulating into D (beta * C applied
    once); sampled rows on both sides of the 2 GiB mark are checked."""
    M, N, K = 140000, 256, 8192
    _say(f"case NN M-contiguous operand past 2 GiB M={M} N={N} K={K}")
    g = torch.Generator(device=cuda).manual_seed(17)
    A = torch.empty(M * K, dtype=torch.bfloat16, device=cuda)
    for k0 in range(0, K, 1024):  # fill in slices to bound the fp32 temporary
        A[k0 * M:(k0 + 1024) * M] = (torch.randn(1024 * M, generator=g, device=cuda) / K ** 0.5).to(torch.bfloat16)
    B = (torch.randn(K * N, generator=g, device=cuda) / K ** 0.5).to(torch.bfloat16)
    Cm = torch.randn(M * N, generator=g, device=cuda).to(torch.bfloat16)
    D = tl.gemm("NN", M, N, K, A, B, Cm, alpha=1.0, beta=0.5)
    torch.cuda.synchronize()
    rows = [0, 1, 255, 256, 65535, 65536, 76695, 76696, 139743, 139999]
    rows += torch.randint(0, M, (22,), generator=torch.Generator().manual_seed(18)).tolist()
    rows = sorted(set(rows))
    Ar = A.view(K, M)[:, rows].float().cpu().numpy().astype(np.float64).ravel()   # R-contiguous, K columns
    Bh = B.float().cpu().numpy().astype(np.float64)
    Cr = Cm.view(N, M)[:, rows].float().cpu().numpy().astype(np.float64).ravel()
    ref = oracle.gemm_f64("NN", len(rows), N, K, 1.0, Ar, len(rows), Bh, K, 0.5, Cr, len(rows)).reshape(N, len(rows))
    got = D.view(N, M)[:, rows].float().cpu().numpy()
    bad, worst = elem_ok(got, ref)
    assert bad == 0, worst
    del A, B, Cm, D
    torch.cuda.empty_cache()


def test_k_zero_and_alpha_zero(cuda):
    M, N = 96, 80
    _say("case K=0 / alpha=0")
    rng = np.random.default_rng(7)
    Cm = colmajor(rng, M, N)
    Cd = torch.from_numpy(Cm).to(cuda)
    A = torch.zeros(M * 8, device=cuda)
    B = torch.zeros(8 * N, device=cuda)
    D = tl.gemm("NN", M, N, 0, A, B, Cd, lda=M, ldb=8, alpha=1.0, beta=0.5)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(D.cpu().numpy(), (0.5 * Cm).astype(np.float32))
    A = torch.from_numpy(colmajor(rng, M, 64)).to(cuda)
    B = torch.from_numpy(colmajor(rng, 64, N)).to(cuda)
    D = tl.gemm("NN", M, N, 64, A, B, Cd, alpha=0.0, beta=1.0)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(D.cpu().numpy(), Cm)


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("tile", [(256, 256), (128, 128), (64, 64)])
def test_16bit_k_zero_gives_beta_c(cuda, layout, tile):
    """K = 0 on every 16-bit engine: no K-tile runs and D = beta * C exactly."""
    M, N = 264, 200
    _say(f"case bf16 K=0 {layout} tile={tile}")
    rng = np.random.default_rng(15)
    Cb = quantize_bits(colmajor(rng, M, N), "bf16")
    ta, tb = layout[0] == "T", layout[1] == "T"
    A = torch.zeros(8 * M, dtype=torch.bfloat16, device=cuda)
    B = torch.zeros(8 * N, dtype=torch.bfloat16, device=cuda)
    D = tl.gemm(layout, M, N, 0, A, B, _dev16(Cb, "bf16", cuda), lda=8 if ta else M, ldb=N if tb else 8,
                alpha=1.0, beta=0.5, tile=tile)
    torch.cuda.synchronize()
    want = (0.5 * bits_to_f32(Cb, "bf16")).astype(np.float32)  # a power-of-two scale of bf16 is exact
    np.testing.assert_array_equal(D.float().cpu().numpy(), want)


def test_zero_sized_is_noop(cuda):
    D = torch.full((16,), 3.0, device=cuda)
    tl.gemm("NN", 0, 4, 8, torch.zeros(64, device=cuda), torch.zeros(32, device=cuda), D=D, lda=8, ldd=4)
    torch.cuda.synchronize()
    assert torch.all(D == 3.0)


@pytest.mark.parametrize("name", ["bf16", "f32", "x3"])
def test_batched(cuda, name):
    M, N, K, batch = 136, 120, 96, 3
    _say(f"case batch {name}")
    rng = np.random.default_rng(8)
    lda, ldb, ldd = K, K, M
    bsA, bsB, bsD = lda * M + 64, ldb * N + 128, ldd * N + 32
    A = np.concatenate([np.pad(colmajor(rng, K, M, lda), (0, 64)) for _ in range(batch)])
    B = np.concatenate([np.pad(colmajor(rng, K, N, ldb), (0, 128)) for _ in range(batch)])
    if name in ("f32", "x3"):
        Ad, Bd, Ar, Br = torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda), A, B
    else:
        Ab, Bb = quantize_bits(A, name), quantize_bits(B, name)
        Ad, Bd, Ar, Br = _dev16(Ab, name, cuda), _dev16(Bb, name, cuda), bits_to_f32(Ab, name), bits_to_f32(Bb, name)
    D = tl.gemm("TN", M, N, K, Ad, Bd, batch=batch, bsA=bsA, bsB=bsB, bsD=bsD, lda=lda, ldb=ldb, ldd=ldd,
                D=torch.zeros(bsD * batch, dtype=Ad.dtype, device=cuda),
                compute=C.COMPUTE_BF16X3 if name == "x3" else C.COMPUTE_DEFAULT)
    torch.cuda.synchronize()
    ref = _f64_ref("TN", M, N, K, Ar, lda, Br, ldb, ldd=ldd, batch=batch, bsA=bsA, bsB=bsB, bsD=bsD)
    g = extract(D.float().cpu().numpy(), M, N, ldd, batch, bsD)
    r = extract(ref, M, N, ldd, batch, bsD)
    if name in ("f32", "x3"):
        assert rel_fro(g, r) <= 1e-5
    else:
        assert elem_ok(g, r)[0] == 0


@pytest.mark.parametrize("dtype", [1, 2, 3])
def test_reference_kat_tn_kernarg(cuda, dtype):
    """The reference's own run: InitRealData patterns (A = idx%7, B = -(idx%11),
    C = idx%7), alpha 0.314, beta 0.2718, padding 32, TN, through the
    72-byte-kernarg-shaped entry point (src/gemmMfma.cpp:106-119, 210-235)."""
    M, N, K, pad = 480, 512, 512, 32
    _say(f"case KAT dtype={dtype}")
    lda = ldb = K + pad
    ldc = ldd = M + pad
    A = oracle.init_real_data(2.4, lda, M)
    B = oracle.init_real_data(-2.4, ldb, N)
    Cm = oracle.init_real_data(2.4, ldc, N)
    alpha, beta = 0.314, 0.2718
    ref = oracle.gemm_f32("TN", M, N, K, alpha, A, lda, B, ldb, beta, Cm, ldc, ldd)
    if dtype == 1:
        dev = [torch.from_numpy(x).to(cuda) for x in (A, B, Cm)]
        D = torch.zeros(ldd * N, device=cuda)
    else:
        name = "f16" if dtype == 2 else "bf16"
        # the reference converts bf16 by truncation (inc/ff_utils.h:119-122); integers are exact either way
        dev = [_dev16(quantize_bits(x, name, truncate=True), name, cuda) for x in (A, B, Cm)]
        D = torch.zeros(ldd * N, dtype=dev[0].dtype, device=cuda)
    st = C.lib().ffgemm_run_tn_kernarg(D.data_ptr(), dev[2].data_ptr(), dev[0].data_ptr(), dev[1].data_ptr(),
                                       alpha, beta, ldd, ldd * N, ldc, ldc * N, lda, lda * M, ldb, ldb * N,
                                       M, N, K, 1, dtype, None, None)
    C.check(st)
    torch.cuda.synchronize()
    g = extract(D.float().cpu().numpy(), M, N, ldd)
    r = extract(ref, M, N, ldd)
    if dtype == 1:
        # the reference's own rule: |gpu - cpu| <= 0.1 (src/gemmMfma.cpp:401)
        assert np.max(np.abs(g - r)) <= 0.1
    else:
        bad, worst = elem_ok(g, r)
        assert bad == 0, worst


@pytest.mark.parametrize("name", ["bf16", "f16", "f32"])
def test_tensile_args_entry(cuda, name):
    """A Tensile-style caller's 148-byte argument buffer (BBH / HBH / SB), batch 2
    with padded batch strides; alpha / beta as those kernels take them."""
    from tests.test_capi import tensile_args
    M, N, K, batch = 200, 136, 96, 2
    _say(f"case Tensile argument buffer {name}")
    rng = np.random.default_rng(16)
    lda = ldb = K + 8
    ldc = ldd = M + 8
    bsA, bsB, bsC = lda * M + 64, ldb * N + 32, ldc * N + 16
    A, B, Cm = (rng.uniform(-1.0, 1.0, n).astype(np.float32) for n in (bsA * batch, bsB * batch, bsC * batch))
    alpha, beta = 0.75, -0.5
    if name == "f32":
        Ad, Bd, Cd = (torch.from_numpy(np.ascontiguousarray(x)).to(cuda) for x in (A, B, Cm))
        Ar, Br, Cr = A, B, Cm
        ab = [int(np.float32(v).view(np.uint32)) for v in (alpha, beta)]
        dt = C.F32
    else:
        bits = [quantize_bits(np.ascontiguousarray(x), name) for x in (A, B, Cm)]
        Ad, Bd, Cd = (_dev16(b_, name, cuda) for b_ in bits)
        Ar, Br, Cr = (bits_to_f32(b_, name) for b_ in bits)
        if name == "f16":   # HBH: fp16 bits in the low half of each slot
            ab = [int(np.float16(v).view(np.uint16)) for v in (alpha, beta)]
        else:
            ab = [int(np.float32(v).view(np.uint32)) for v in (alpha, beta)]
        dt = C.F16 if name == "f16" else C.BF16
    D = torch.full((bsC * batch,), 7.0, dtype=Cd.dtype, device=cuda)
    buf = tensile_args(D.data_ptr(), Cd.data_ptr(), Ad.data_ptr(), Bd.data_ptr(), ab[0], ab[1], M, N, K, batch,
                       ldd, bsC, ldc, bsC, lda, bsA, ldb, bsB, sizes=(bsC * batch, bsA * batch, bsB * batch))
    C.check(C.lib().ffgemm_run_tensile_args(buf, len(buf), dt, None, None))
    torch.cuda.synchronize()
    ref = _f64_ref("TN", M, N, K, Ar, lda, Br, ldb, alpha, beta, Cr, ldc, ldd, batch, bsA, bsB, bsC, bsC)
    got = D.float().cpu().numpy()
    for b in range(batch):
        r = ref[b * bsC: b * bsC + ldd * N].reshape(N, ldd)[:, :M]
        g = got[b * bsC: b * bsC + ldd * N].reshape(N, ldd)[:, :M]
        if name == "f32":
            assert rel_fro(g, r) <= 1e-5
        else:
            bad, worst = elem_ok(g, r)
            assert bad == 0, (b, worst)


def test_bf16_truncate_matches_reference_store(cuda):
    """bf16 D rounded by truncation, as the reference's store
    (src/gemmMfmaKernelWriter.h:1661-1664).  Integer data keeps every fp32 sum
    exact, so D must equal trunc(fma(alpha, sum, beta*c)) to the bit."""
    M, N, K = 128, 128, 128
    _say("case bf16 truncate")
    rng = np.random.default_rng(9)
    A = colmajor(rng, K, M, kind="int")
    B = colmajor(rng, K, N, kind="int")
    Cm = colmajor(rng, M, N, kind="int")
    name = "bf16"
    Ad, Bd, Cd = (_dev16(quantize_bits(x, name), name, cuda) for x in (A, B, Cm))
    alpha, beta = 0.5, 0.25  # powers of two keep alpha*sum + beta*c exact
    D = tl.gemm("TN", M, N, K, Ad, Bd, Cd, alpha=alpha, beta=beta, round=C.ROUND_TRUNCATE)
    torch.cuda.synchronize()
    exact = oracle.gemm_f64("TN", M, N, K, alpha, A.astype(np.float64), K, B.astype(np.float64), K, beta,
                            Cm.astype(np.float64), M, M).astype(np.float32)
    want = oracle.f32_to_bf16_bits(exact, truncate=True)
    got = D.view(torch.int16).cpu().numpy().view(np.uint16)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("layout", ["TN", "NT"])
@pytest.mark.parametrize("name,rounding", [("bf16", "rne"), ("bf16", "truncate"), ("f16", "rne")])
def test_staged_beta_epilogue_bit_exact(cuda, layout, name, rounding):
    """beta != 0 on the eight-wave engine: whole 256 x 256 tiles leave through
    LDS (fp32 staging, C read as 8-byte column pieces), the edge tiles of the
    same launch through the per-fragment epilogue.  C and D have different
    padded leading dimensions.  Integer data and power-of-two alpha / beta keep
    every fp32 result exact, so D must equal the oracle's conversion of the
    exact result to the bit, under either rounding."""
    M, N, K = 640, 512, 192          # M: two whole tile rows and one of 128 rows
    ldc, ldd = M + 24, M + 8
    _say(f"case staged beta {layout} {name} {rounding}")
    rng = np.random.default_rng(17)
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    A = colmajor(rng, ra, ca, kind="int")
    B = colmajor(rng, rb, cb, kind="int")
    Cm = colmajor(rng, M, N, ldc, kind="int")
    Ad, Bd, Cd = (_dev16(quantize_bits(x, name), name, cuda) for x in (A, B, Cm))
    Dd = torch.full((ldd * N,), float("nan"), dtype=Ad.dtype, device=cuda)
    alpha, beta = 0.5, -0.25
    rnd = C.ROUND_TRUNCATE if rounding == "truncate" else C.ROUND_RNE
    # the eight-wave engine at 256 x 256, no repack
    tl.gemm(layout, M, N, K, Ad, Bd, Cd, Dd, ldc=ldc, ldd=ldd, alpha=alpha, beta=beta, round=rnd,
            tile=(256, 256, 0, 2, 4, 0, 0, 1))
    torch.cuda.synchronize()
    exact = oracle.gemm_f64(layout, M, N, K, alpha, A.astype(np.float64), ra, B.astype(np.float64), rb, beta,
                            Cm.astype(np.float64), ldc, ldd).astype(np.float32)
    got = Dd.view(torch.int16).cpu().numpy().view(np.uint16).reshape(N, ldd)
    want = quantize_bits(np.ascontiguousarray(extract(exact, M, N, ldd)[0]).ravel(), name,
                         truncate=rounding == "truncate").reshape(N, M)
    np.testing.assert_array_equal(got[:, :M], want)
    assert np.isnan(Dd.float().cpu().numpy().reshape(N, ldd)[:, M:]).all()   # padding rows untouched


def test_fp8_quantize_bit_exact_and_gemm(cuda):
    M, N, K = 192, 160, 256
    _say("case fp8")
    rng = np.random.default_rng(2)
    A = colmajor(rng, M, K, kind="normal")   # NN-stored inputs: A M-contiguous, B K-contiguous
    B = colmajor(rng, K, N, kind="normal")
    Ad, Bd = torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda)
    qa, sa = tl.fp8.quantize(Ad, rows=M, k=K, ldx=M, trans_x=False)
    qb, sb = tl.fp8.quantize(Bd, rows=N, k=K, ldx=K, trans_x=True)
    torch.cuda.synchronize()
    sa_ref = oracle.fp8_scale(np.abs(A).max())
    sb_ref = oracle.fp8_scale(np.abs(B).max())
    assert sa.item() == np.float32(sa_ref) and sb.item() == np.float32(sb_ref)
    # oracle encoding of the same operands in the K-contiguous layout
    At = A.reshape(K, M).T.copy().ravel()   # (m, k) at k + m*K
    qa_ref = oracle.f32_to_e4m3_bits(At, np.float32(sa_ref))
    qb_ref = oracle.f32_to_e4m3_bits(B, np.float32(sb_ref))
    np.testing.assert_array_equal(qa.cpu().numpy(), qa_ref)
    np.testing.assert_array_equal(qb.cpu().numpy(), qb_ref)
    D = tl.gemm("TN", M, N, K, qa, qb, lda=K, ldb=K, scaleA=sa, scaleB=sb, dtype_out=torch.float32)
    torch.cuda.synchronize()
    deq_a = oracle.e4m3_bits_to_f32(qa_ref).astype(np.float64) / sa_ref
    deq_b = oracle.e4m3_bits_to_f32(qb_ref).astype(np.float64) / sb_ref
    ref = oracle.gemm_f64("TN", M, N, K, 1.0, deq_a, K, deq_b, K)
    err = rel_fro(D.cpu().numpy(), ref)
    err_vs_fp32_inputs = rel_fro(D.cpu().numpy(), _f64_ref("NN", M, N, K, A, M, B, K))
    _say(f"fp8 rel_fro vs dequantised {err:.3e}; vs fp32 inputs {err_vs_fp32_inputs:.3e}")
    # the products of e4m3 values are exact; the bound covers the f8f6f4 MFMA's
    # internal accumulation (measured 1.5e-5 here) and the fp32 dequant multiply
    assert err <= 5e-5
    # the form's total error is the quantisation's own: the oracle's dequantised
    # FP64 product has the same error against the fp32 inputs (e4m3: a few %)
    err_q = rel_fro(ref, _f64_ref("NN", M, N, K, A, M, B, K))
    assert 0.01 < err_q < 0.1
    assert abs(err_vs_fp32_inputs - err_q) <= 1e-4


@pytest.mark.parametrize("tile", [(256, 256), (128, 128), (64, 64)])
@pytest.mark.parametrize("out", ["f32", "bf16"])
def test_fp8_tiles_edges(cuda, tile, out):
    """FP8 GEMM on every engine: M, N not multiples of any tile and K = 208
    (a partial K-tile of 80), operands quantised by the oracle's encoder."""
    M, N, K = 328, 272, 208
    _say(f"case fp8 tile={tile} out={out}")
    rng = np.random.default_rng(21)
    A = rng.standard_normal(M * K).astype(np.float32)   # K-contiguous: (m, k) at k + m*K
    B = rng.standard_normal(N * K).astype(np.float32)
    sa, sb = np.float32(oracle.fp8_scale(np.abs(A).max())), np.float32(oracle.fp8_scale(np.abs(B).max()))
    qa, qb = oracle.f32_to_e4m3_bits(A, sa), oracle.f32_to_e4m3_bits(B, sb)
    dt = torch.float32 if out == "f32" else torch.bfloat16
    D = tl.gemm("TN", M, N, K, torch.from_numpy(qa).to(cuda), torch.from_numpy(qb).to(cuda), lda=K, ldb=K,
                scaleA=torch.tensor([sa], device=cuda), scaleB=torch.tensor([sb], device=cuda), dtype_out=dt,
                tile=tile)
    torch.cuda.synchronize()
    deq_a = oracle.e4m3_bits_to_f32(qa).astype(np.float64) / sa
    deq_b = oracle.e4m3_bits_to_f32(qb).astype(np.float64) / sb
    ref = oracle.gemm_f64("TN", M, N, K, 1.0, deq_a, K, deq_b, K)
    got = D.float().cpu().numpy()
    if out == "f32":
        assert rel_fro(got, ref) <= 5e-5
    else:
        bad, worst = elem_ok(got, ref)
        assert bad == 0, worst


@pytest.mark.parametrize("layout", LAYOUTS)
def test_fp64_oracle_product(cuda, layout):
    M, N, K = 130, 70, 90
    _say(f"case f64 {layout}")
    rng = np.random.default_rng(4)
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    A = rng.uniform(-1, 1, ra * ca)
    B = rng.uniform(-1, 1, rb * cb)
    Cm = rng.uniform(-1, 1, M * N)
    D = tl.gemm(layout, M, N, K, torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda),
                torch.from_numpy(Cm).to(cuda), alpha=0.75, beta=-0.5)
    torch.cuda.synchronize()
    ref = oracle.gemm_f64(layout, M, N, K, 0.75, A, ra, B, rb, -0.5, Cm, M, M)
    assert rel_fro(D.cpu().numpy(), ref) <= 1e-12


@pytest.mark.parametrize("name", ["bf16", "f16"])
@pytest.mark.parametrize("layout", ["NN", "NT", "TN"])
@pytest.mark.parametrize("beta", [0.0, 1.0])
def test_4096_sampled_columns(cuda, name, layout, beta):
    """Config 2 at full size, bf16 and fp16 (output in the input type): 24
    random output columns against the FP64 oracle and every column sum against
    (1^T op(A)) op(B) + beta * (1^T C) in FP64."""
    M = N = K = 4096
    _say(f"case 4096 {name} {layout} beta={beta}")
    tdt = torch.bfloat16 if name == "bf16" else torch.float16
    g = torch.Generator(device="cpu").manual_seed(1)
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    A32 = (torch.randn(ra * ca, generator=g) / np.sqrt(K)).to(tdt)
    B32 = (torch.randn(rb * cb, generator=g) / np.sqrt(K)).to(tdt)
    C32 = (torch.rand(M * N, generator=g) * 2 - 1).to(tdt)
    D = tl.gemm(layout, M, N, K, A32.to(cuda), B32.to(cuda), C32.to(cuda), beta=beta)
    torch.cuda.synchronize()
    Dh = D.float().cpu().numpy().reshape(N, M)
    A64 = A32.double().numpy()
    B64 = B32.double().numpy()
    C64 = C32.double().numpy().reshape(N, M)
    rng = np.random.default_rng(0)
    cols = rng.choice(N, 24, replace=False)
    opA = A64.reshape(ca, ra).T if layout[0] == "N" else A64.reshape(ca, ra)      # (M, K)
    opB = B64.reshape(cb, rb).T if layout[1] == "N" else B64.reshape(cb, rb)      # (K, N)
    sub = np.ascontiguousarray(opB[:, cols])                                     # (K, 24)
    ref_cols = oracle.gemm_f64("TN", M, len(cols), K, 1.0, np.ascontiguousarray(opA).ravel(), K,
                               sub.T.ravel(), K).reshape(len(cols), M).T + beta * C64[cols].T
    bad, worst = elem_ok(Dh[cols].T, ref_cols)
    assert bad == 0, worst
    colsum_ref = (opA.sum(axis=0) @ opB) + beta * C64.sum(axis=1)
    colsum = Dh.astype(np.float64).sum(axis=1)
    assert rel_fro(colsum, colsum_ref) < 2e-3


def test_config1_fp32_nn_480x512x512_vs_reference_loop(cuda):
    """The reference's own CPU case (fp32 NN 480 x 512 x 512, alpha 1, beta 0,
    uniform[-1, 1] from seed 0) on the GPU, against the oracle's restatement of
    the reference CPU GEMM (src/gemmMfma.cpp:325-346)."""
    M, N, K = 480, 512, 512
    _say("case config1 fp32 NN 480x512x512")
    rng = np.random.default_rng(0)
    A = rng.uniform(-1, 1, M * K).astype(np.float32)
    B = rng.uniform(-1, 1, K * N).astype(np.float32)
    D = tl.gemm("NN", M, N, K, torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda))
    torch.cuda.synchronize()
    ref = oracle.gemm_f32("NN", M, N, K, 1.0, A, M, B, K)
    assert rel_fro(D.cpu().numpy(), ref) <= 1e-5
    assert rel_fro(D.cpu().numpy(), _f64_ref("NN", M, N, K, A, M, B, K)) <= 1e-5


def test_fp8_4096_sampled_columns(cuda):
    """fp8 e4m3 TN at 4096^3 (device quantisation, fp32 out): 16 sampled output
    columns against the FP64 product of the dequantised operands."""
    S = 4096
    _say("case fp8 4096 sampled")
    g = torch.Generator(device="cpu").manual_seed(2)
    A = torch.randn(S * S, generator=g).to(cuda)   # K-contiguous (k + m*K)
    B = torch.randn(S * S, generator=g).to(cuda)
    qa, sa = tl.fp8.quantize(A, rows=S, k=S, ldx=S, trans_x=True)
    qb, sb = tl.fp8.quantize(B, rows=S, k=S, ldx=S, trans_x=True)
    D = tl.gemm("TN", S, S, S, qa, qb, lda=S, ldb=S, scaleA=sa, scaleB=sb, dtype_out=torch.float32)
    torch.cuda.synchronize()
    cols = np.random.default_rng(1).choice(S, 16, replace=False)
    deq_a = oracle.e4m3_bits_to_f32(qa.cpu().numpy()).astype(np.float64) / sa.item()
    qb_h = qb.cpu().numpy().reshape(S, S)[cols]
    deq_b = oracle.e4m3_bits_to_f32(qb_h.ravel()).astype(np.float64) / sb.item()
    ref = oracle.gemm_f64("TN", S, len(cols), S, 1.0, deq_a, S, deq_b, S).reshape(len(cols), S)
    got = D.cpu().numpy().reshape(S, S)[cols]
    assert rel_fro(got, ref) <= 5e-5


def test_fp8_8192_sampled_columns(cuda):
    """Config 3's FP8 form at full size: A, B ~ normal(0, 1) from seed 2 (fp32,
    M-contiguous A and K-contiguous B as the NN precision problem stores them),
    quantised on the device with per-tensor amax scales, fp32 out.  16 sampled
    output columns against the FP64 product of the dequantised operands
    (rel-Frobenius <= 5e-5: the GEMM's own error), and the whole form's error
    against the FP64 product of the unquantised inputs, which is the e4m3
    quantisation's (a few %)."""
    S = 8192
    _say("case fp8 8192 sampled")
    g = torch.Generator(device="cpu").manual_seed(2)
    A = torch.randn(S * S, generator=g).to(cuda)   # M-contiguous: (m, k) at m + k*S
    B = torch.randn(S * S, generator=g).to(cuda)   # K-contiguous: (k, n) at k + n*S
    qa, sa = tl.fp8.quantize(A, rows=S, k=S, ldx=S, trans_x=False)
    qb, sb = tl.fp8.quantize(B, rows=S, k=S, ldx=S, trans_x=True)
    D = tl.gemm("TN", S, S, S, qa, qb, lda=S, ldb=S, scaleA=sa, scaleB=sb, dtype_out=torch.float32)
    torch.cuda.synchronize()
    cols = np.random.default_rng(5).choice(S, 16, replace=False)
    # the quantised bytes are the oracle encoder's (the quantiser is bit-exact;
    # checked here on the sampled operand rows too)
    a_h = A.cpu().numpy()
    qa_h = qa.cpu().numpy()
    sa_h, sb_h = sa.item(), sb.item()
    assert sa_h == np.float32(oracle.fp8_scale(np.abs(a_h).max()))
    At = a_h.reshape(S, S).T          # (m, k), K-contiguous rows
    del a_h
    rows = np.random.default_rng(6).choice(S, 64, replace=False)
    assert np.array_equal(qa_h.reshape(S, S)[rows], oracle.f32_to_e4m3_bits(np.ascontiguousarray(At[rows]).ravel(),
                                                                          np.float32(sa_h)).reshape(64, S))
    deq_a = oracle.e4m3_bits_to_f32(qa_h).astype(np.float64) / sa_h
    del qa_h
    qb_h = qb.cpu().numpy().reshape(S, S)[cols]
    deq_b = oracle.e4m3_bits_to_f32(qb_h.ravel()).astype(np.float64) / sb_h
    ref = oracle.gemm_f64("TN", S, len(cols), S, 1.0, deq_a, S, deq_b, S).reshape(len(cols), S)
    del deq_a
    got = D.cpu().numpy().reshape(S, S)[cols]
    err = rel_fro(got, ref)
    _say(f"rel_fro vs dequantised FP64 = {err:.3e}")
    assert err <= 5e-5
    b_cols = B.cpu().numpy().reshape(S, S)[cols].astype(np.float64)
    ref_x = oracle.gemm_f64("TN", S, len(cols), S, 1.0, np.ascontiguousarray(At).astype(np.float64).ravel(), S,
                            b_cols.ravel(), S).reshape(len(cols), S)
    err_x = rel_fro(got, ref_x)
    _say(f"rel_fro vs unquantised FP64 = {err_x:.3e}")
    assert 0.01 < err_x < 0.1


@pytest.mark.parametrize("compute", ["f32", "bf16x3"])
def test_fp32_forms_8192_sampled_columns(cuda, compute):
    """The precision configuration at full size (NN, 8192^3, normal(0, 1) from
    seed 2): 8 sampled output columns within 1e-5 of FP64."""
    S = 8192
    _say(f"case {compute} 8192 sampled")
    g = torch.Generator(device="cpu").manual_seed(2)
    A = torch.randn(S * S, generator=g)
    B = torch.randn(S * S, generator=g)
    comp = C.COMPUTE_BF16X3 if compute == "bf16x3" else C.COMPUTE_DEFAULT
    D = tl.gemm("NN", S, S, S, A.to(cuda), B.to(cuda), compute=comp)
    torch.cuda.synchronize()
    cols = np.random.default_rng(3).choice(S, 8, replace=False)
    Bc = B.numpy().reshape(S, S)[cols]          # column n of B (K-contiguous) = row n of this view
    At = A.numpy().reshape(S, S).T.astype(np.float64).ravel()   # op(A) K-contiguous: the oracle's TN form
    ref = oracle.gemm_f64("TN", S, len(cols), S, 1.0, At, S, Bc.astype(np.float64).ravel(), S).reshape(len(cols), S)
    del At
    got = D.cpu().numpy().reshape(S, S)[cols]
    err = rel_fro(got, ref)
    _say(f"rel_fro={err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("K", [8, 40, 64, 72, 136])
def test_256_tile_short_k(cuda, layout, K):
    """One or two K-tiles (partial ones included) on every 256 x 256 engine:
    the pipelines' prologue, tail and barrier balance with almost no main loop."""
    M, N = 264, 520
    _say(f"case short K={K} {layout}")
    rng = np.random.default_rng(K)
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    A = colmajor(rng, ra, ca, kind="normal")
    B = colmajor(rng, rb, cb, kind="normal")
    Ab, Bb = quantize_bits(A, "bf16"), quantize_bits(B, "bf16")
    D = tl.gemm(layout, M, N, K, _dev16(Ab, "bf16", cuda), _dev16(Bb, "bf16", cuda), tile=(256, 256))
    torch.cuda.synchronize()
    ref = _f64_ref(layout, M, N, K, bits_to_f32(Ab, "bf16"), ra, bits_to_f32(Bb, "bf16"), rb)
    bad, worst = elem_ok(D.float().cpu().numpy(), ref)
    assert bad == 0, worst


@pytest.mark.parametrize("K", [16, 112, 144])
def test_fp8_256_tile_short_k(cuda, K):
    M, N = 264, 300
    _say(f"case fp8 short K={K}")
    rng = np.random.default_rng(K)
    A = rng.standard_normal(M * K).astype(np.float32)
    B = rng.standard_normal(N * K).astype(np.float32)
    sa, sb = np.float32(oracle.fp8_scale(np.abs(A).max())), np.float32(oracle.fp8_scale(np.abs(B).max()))
    qa, qb = oracle.f32_to_e4m3_bits(A, sa), oracle.f32_to_e4m3_bits(B, sb)
    D = tl.gemm("TN", M, N, K, torch.from_numpy(qa).to(cuda), torch.from_numpy(qb).to(cuda), lda=K, ldb=K,
                scaleA=torch.tensor([sa], device=cuda), scaleB=torch.tensor([sb], device=cuda),
                dtype_out=torch.float32, tile=(256, 256))
    torch.cuda.synchronize()
    deq_a = oracle.e4m3_bits_to_f32(qa).astype(np.float64) / sa
    deq_b = oracle.e4m3_bits_to_f32(qb).astype(np.float64) / sb
    ref = oracle.gemm_f64("TN", M, N, K, 1.0, deq_a, K, deq_b, K)
    assert rel_fro(D.cpu().numpy(), ref) <= 5e-5


@pytest.mark.parametrize("layout", LAYOUTS)
def test_batched_256_tile(cuda, layout):
    """Batch strides with padding between batches on every 256 x 256 engine."""
    M, N, K, batch = 300, 264, 136, 3
    _say(f"case batched 256 {layout}")
    rng = np.random.default_rng(9)
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    lda, ldb, ldd = (ra + 7) // 8 * 8, (rb + 7) // 8 * 8, (M + 7) // 8 * 8 + 8
    bsA, bsB, bsD = lda * ca + 64, ldb * cb + 128, ldd * N + 32
    A = np.concatenate([np.pad(colmajor(rng, ra, ca, lda, kind="normal"), (0, 64)) for _ in range(batch)])
    B = np.concatenate([np.pad(colmajor(rng, rb, cb, ldb, kind="normal"), (0, 128)) for _ in range(batch)])
    Ab, Bb = quantize_bits(A, "bf16"), quantize_bits(B, "bf16")
    D = tl.gemm(layout, M, N, K, _dev16(Ab, "bf16", cuda), _dev16(Bb, "bf16", cuda), batch=batch, bsA=bsA,
                bsB=bsB, bsD=bsD, lda=lda, ldb=ldb, ldd=ldd,
                D=torch.zeros(bsD * batch, dtype=torch.bfloat16, device=cuda), tile=(256, 256))
    torch.cuda.synchronize()
    ref = _f64_ref(layout, M, N, K, bits_to_f32(Ab, "bf16"), lda, bits_to_f32(Bb, "bf16"), ldb, ldd=ldd,
                   batch=batch, bsA=bsA, bsB=bsB, bsD=bsD)
    g = extract(D.float().cpu().numpy(), M, N, ldd, batch, bsD)
    r = extract(ref, M, N, ldd, batch, bsD)
    assert elem_ok(g, r)[0] == 0


@pytest.mark.parametrize("name", ["bf16", "f16"])
@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("waves", [(2, 4), (2, 2)])
def test_256_wave_tilings(cuda, name, layout, waves):
    """Both wave tilings of the 256 x 256 tile on every layout (the eight-wave
    engine with transposing reads for NN and TT, the four-wave engines for TN
    and NT included), ragged in M, N and K."""
    M, N, K = 328, 272, 200
    _say(f"case waves {waves} {name} {layout}")
    rng = np.random.default_rng(13)
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    A = colmajor(rng, ra, ca, kind="normal", scale=1 / np.sqrt(K))
    B = colmajor(rng, rb, cb, kind="normal", scale=1 / np.sqrt(K))
    Ab, Bb = quantize_bits(A, name), quantize_bits(B, name)
    D = tl.gemm(layout, M, N, K, _dev16(Ab, name, cuda), _dev16(Bb, name, cuda), tile=(256, 256, 0) + waves)
    torch.cuda.synchronize()
    ref = _f64_ref(layout, M, N, K, bits_to_f32(Ab, name), ra, bits_to_f32(Bb, name), rb)
    bad, worst = elem_ok(D.float().cpu().numpy(), ref)
    assert bad == 0, worst


@pytest.mark.parametrize("name", ["bf16", "f16"])
@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("mt", [64, 128])
def test_16x16_mfma_form(cuda, name, layout, mt):
    """The tile record's MFMA knob (mfma_mn = 16: 16 x 16 x 32 MFMAs) against
    the FP64 oracle: a partial K-tile and edges in M and N, then ragged M, N
    (rows past M inside a lane's four) with padded leading dimensions and
    alpha / beta with C; D's padding rows stay untouched."""
    tile = (mt, mt, 0, 0, 0, 16)
    _say(f"case 16x16 MFMA {name} {layout} tile={mt}x{mt}")
    rng = np.random.default_rng(21)
    M, N, K = 328, 272, 200
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    Ab = quantize_bits(colmajor(rng, ra, ca, kind="normal", scale=1 / np.sqrt(K)), name)
    Bb = quantize_bits(colmajor(rng, rb, cb, kind="normal", scale=1 / np.sqrt(K)), name)
    D = tl.gemm(layout, M, N, K, _dev16(Ab, name, cuda), _dev16(Bb, name, cuda), tile=tile)
    torch.cuda.synchronize()
    ref = _f64_ref(layout, M, N, K, bits_to_f32(Ab, name), ra, bits_to_f32(Bb, name), rb)
    bad, worst = elem_ok(D.float().cpu().numpy(), ref)
    assert bad == 0, f"{bad} elements off (worst scaled error {worst:.3g})"

    M, N, K = 65, 37, 136
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    lda, ldb, ldc, ldd = (ra + 7) // 8 * 8, (rb + 7) // 8 * 8, M + 15, (M + 7) // 8 * 8 + 8
    ldc = (ldc + 7) // 8 * 8
    A = colmajor(rng, ra, ca, lda, kind="normal")
    B = colmajor(rng, rb, cb, ldb, kind="normal")
    Cm = colmajor(rng, M, N, ldc)
    bits = [quantize_bits(x, name) for x in (A, B, Cm, np.full(ldd * N, 7.0, dtype=np.float32))]
    Ad, Bd, Cd, Dd = (_dev16(b, name, cuda) for b in bits)
    alpha, beta = 0.5, -1.5
    tl.gemm(layout, M, N, K, Ad, Bd, Cd, Dd, lda=lda, ldb=ldb, ldc=ldc, ldd=ldd, alpha=alpha, beta=beta, tile=tile)
    torch.cuda.synchronize()
    got = Dd.float().cpu().numpy()
    ref = _f64_ref(layout, M, N, K, bits_to_f32(bits[0], name), lda, bits_to_f32(bits[1], name), ldb, alpha, beta,
                   bits_to_f32(bits[2], name), ldc, ldd)
    bad, worst = elem_ok(extract(got, M, N, ldd), extract(ref, M, N, ldd))
    assert bad == 0, worst
    assert np.all(got.reshape(N, ldd)[:, M:] == 7.0)


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("mt", [64, 128])
def test_16x16_mfma_form_fp32(cuda, layout, mt):
    """fp32 operands on 16 x 16 x 4 MFMAs (mfma_mn = 16): a partial K-tile,
    ragged M and N, padded leading dimensions, alpha / beta with C, against
    the FP64 oracle (rel-Frobenius <= 1e-5); D's padding rows untouched."""
    M, N, K = 330, 270, 212
    _say(f"case 16x16 MFMA fp32 {layout} tile={mt}x{mt}")
    rng = np.random.default_rng(23)
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    lda, ldb, ldd = (ra + 3) // 4 * 4 + 4, (rb + 3) // 4 * 4, (M + 3) // 4 * 4 + 8
    A = colmajor(rng, ra, ca, lda, kind="normal")
    B = colmajor(rng, rb, cb, ldb, kind="normal")
    Cm = colmajor(rng, M, N, ldd)
    Dd = torch.full((ldd * N,), 7.0, device=cuda)
    p = tl.make_problem(layout, C.F32, M, N, K, lda, ldb, ldd, ldd, 0.5, -1.5)
    t = tl.query_tile(p, (mt, mt, 0, 0, 0, 16))
    assert (t.mt_m, t.mfma_mn, t.depth_u) == (mt, 16, 32)
    tl.run(p, torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda), torch.from_numpy(Cm).to(cuda), Dd,
           tile=(mt, mt, 0, 0, 0, 16), workspace=None)
    torch.cuda.synchronize()
    got = Dd.cpu().numpy()
    ref = _f64_ref(layout, M, N, K, A, lda, B, ldb, 0.5, -1.5, Cm, ldd, ldd)
    assert rel_fro(extract(got, M, N, ldd), extract(ref, M, N, ldd)) <= 1e-5
    assert np.all(got.reshape(N, ldd)[:, M:] == 7.0)


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("mt", [64, 128])
def test_16x16_mfma_form_f32_out(cuda, layout, mt):
    """bf16 inputs, fp32 D and C through the 16 x 16 MFMA form."""
    M, N, K = 144, 104, 264
    _say(f"case 16x16 MFMA bf16->f32 {layout} tile={mt}x{mt}")
    rng = np.random.default_rng(22)
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    Ab = quantize_bits(colmajor(rng, ra, ca, kind="normal", scale=1 / np.sqrt(K)), "bf16")
    Bb = quantize_bits(colmajor(rng, rb, cb, kind="normal", scale=1 / np.sqrt(K)), "bf16")
    Cm = colmajor(rng, M, N)
    Cd = torch.from_numpy(Cm).to(cuda)
    D = tl.gemm(layout, M, N, K, _dev16(Ab, "bf16", cuda), _dev16(Bb, "bf16", cuda), Cd, alpha=2.0, beta=0.25,
                dtype_out=torch.float32, tile=(mt, mt, 0, 0, 0, 16))
    torch.cuda.synchronize()
    ref = _f64_ref(layout, M, N, K, bits_to_f32(Ab, "bf16"), ra, bits_to_f32(Bb, "bf16"), rb, 2.0, 0.25, Cm)
    assert rel_fro(D.cpu().numpy(), ref) <= 1e-5


# (dtype, layout, M, N, K, tile record or None)
INPLACE_CASES = [
    ("bf16", "TN", 512, 512, 128, (256, 256, 0, 2, 4, 0, 0, 1)),   # eight-wave engine, staged beta epilogue
    ("bf16", "TN", 520, 264, 128, (256, 256, 0, 2, 4, 0, 0, 1)),   # eight-wave engine, edge tiles
    ("f16", "NN", 520, 264, 96, (256, 256, 0, 2, 2, 0, 0, 1)),     # split-ring engine, C requested first
    ("f16", "TN", 300, 200, 136, None),                            # the library's tile
    ("bf16", "NT", 264, 392, 512, (128, 128, 0, 0, 0, 0, 2)),      # split-K into the workspace, then the reduce
    ("bf16", "NN", 384, 256, 200, (0, 0, 0, 0, 0, 0, 0, 2)),       # K-major repack, then TN
    ("f32", "NN", 200, 136, 64, None),                             # fp32
]


@pytest.mark.parametrize("case", INPLACE_CASES, ids=lambda c: "_".join(map(str, c[:5])))
def test_in_place_beta_update(cuda, case):
    """D = alpha*op(A)*op(B) + beta*C with D the same buffer as C (ldc = ldd):
    every epilogue reads each C element in the thread that later writes that
    D element, so the in-place result equals the out-of-place one bit for bit."""
    name, layout, M, N, K, tile = case
    _say(f"case in-place {name} {layout} {M}x{N}x{K} tile={tile}")
    rng = np.random.default_rng(M + N + K)
    (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
    ld = (M + 7) // 8 * 8 + 8
    A = colmajor(rng, ra, ca, kind="normal", scale=1 / np.sqrt(K))
    B = colmajor(rng, rb, cb, kind="normal", scale=1 / np.sqrt(K))
    Cm = colmajor(rng, M, N, ld, kind="normal")
    if name == "f32":
        Ad, Bd, Cd = (torch.from_numpy(x).to(cuda) for x in (A, B, Cm))
        code = C.F32
    else:
        Ad, Bd, Cd = (_dev16(quantize_bits(x, name), name, cuda) for x in (A, B, Cm))
        code = C.BF16 if name == "bf16" else C.F16
    p = tl.make_problem(layout, code, M, N, K, ra, rb, ld, ld, 0.625, -1.25)
    D = torch.full_like(Cd, float("nan"))
    tl.run(p, Ad, Bd, Cd, D, tile=tile)
    inplace = Cd.clone()
    tl.run(p, Ad, Bd, inplace, inplace, tile=tile)
    torch.cuda.synchronize()
    got = inplace.cpu().view(torch.int32 if name == "f32" else torch.int16).numpy().reshape(N, ld)
    want = D.cpu().view(torch.int32 if name == "f32" else torch.int16).numpy().reshape(N, ld)
    np.testing.assert_array_equal(got[:, :M], want[:, :M])
    # the padding rows keep C's values
    cref = Cd.cpu().view(torch.int32 if name == "f32" else torch.int16).numpy().reshape(N, ld)
    np.testing.assert_array_equal(got[:, M:], cref[:, M:])
