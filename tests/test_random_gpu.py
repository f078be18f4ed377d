"""Seeded random problems against the FP64 oracle: every layout, 16-bit dtype,
tile record (macro tile, wave tiling, MFMA instruction), ragged M / N / K,
padded leading dimensions and alpha / beta with C, single and grouped."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
import tensilelite_amd as tl
from tensilelite_amd import _capi as C
from tensilelite_amd import grouped
from tests._util import bits_to_f32, colmajor, elem_ok, extract, op_dims, quantize_bits

pytestmark = pytest.mark.gpu

# (mt_m, mt_n, group_m, waves_m, waves_n, mfma_mn); None is the library's choice
_TILES = [None, (256, 256, 0, 2, 4, 0), (256, 256, 0, 2, 2, 0), (256, 128), (128, 256), (128, 128), (64, 64),
          (128, 128, 0, 0, 0, 16), (64, 64, 0, 0, 0, 16), (256, 256, 3, 0, 0, 0)]


def _dev16(bits, name, device):
    t = torch.from_numpy(bits.view(np.int16).copy()).to(device)
    return t.view(torch.bfloat16 if name == "bf16" else torch.float16)


def _round8(x):
    return (x + 7) // 8 * 8


def _problem(rng):
    layout = ["NN", "NT", "TN", "TT"][rng.integers(4)]
    name = ["bf16", "f16"][rng.integers(2)]
    M, N = int(rng.integers(1, 600)), int(rng.integers(1, 600))
    K = int(rng.integers(0, 70)) * 8          # K-contiguous operands need K % 8 == 0
    beta = [0.0, 0.0, 1.0, -0.75][rng.integers(4)]
    alpha = [1.0, 0.5, -2.0][rng.integers(3)]
    return layout, name, M, N, K, alpha, beta


@pytest.mark.parametrize("seed", range(12))
def test_random_single_problems(cuda, seed):
    rng = np.random.default_rng(1000 + seed)
    for case in range(8):
        layout, name, M, N, K, alpha, beta = _problem(rng)
        tile = _TILES[rng.integers(len(_TILES))]
        print(f"case seed={seed}.{case} {name} {layout} M={M} N={N} K={K} a={alpha} b={beta} tile={tile}",
              flush=True)
        (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
        lda = _round8(max(ra, 1)) + 8 * int(rng.integers(3))
        ldb = _round8(max(rb, 1)) + 8 * int(rng.integers(3))
        ldc = _round8(M) + 8 * int(rng.integers(2))
        ldd = _round8(M) + 8 * int(rng.integers(2))
        A = colmajor(rng, max(ra, 1), max(ca, 1), lda, kind="normal", scale=1 / np.sqrt(max(K, 1)))
        B = colmajor(rng, max(rb, 1), max(cb, 1), ldb, kind="normal", scale=1 / np.sqrt(max(K, 1)))
        Cm = colmajor(rng, M, N, ldc)
        bits = [quantize_bits(x, name) for x in (A, B, Cm, np.full(ldd * N, 7.0, dtype=np.float32))]
        Ad, Bd, Cd, Dd = (_dev16(b, name, cuda) for b in bits)
        tl.gemm(layout, M, N, K, Ad, Bd, Cd if beta else None, Dd, lda=lda, ldb=ldb, ldc=ldc if beta else None,
                ldd=ldd, alpha=alpha, beta=beta, tile=tile)
        torch.cuda.synchronize()
        got = Dd.float().cpu().numpy()
        ref = oracle.gemm_f64(layout, M, N, K, alpha, bits_to_f32(bits[0], name).astype(np.float64), lda,
                              bits_to_f32(bits[1], name).astype(np.float64), ldb, beta,
                              bits_to_f32(bits[2], name).astype(np.float64) if beta else None,
                              ldc if beta else None, ldd)
        bad, worst = elem_ok(extract(got, M, N, ldd), extract(ref, M, N, ldd))
        assert bad == 0, (seed, case, layout, name, M, N, K, tile, worst)
        assert np.all(got.reshape(N, ldd)[:, M:] == 7.0), (seed, case, "padding rows written")


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("tile", [(256, 256), (128, 256), (64, 64)])
def test_random_grouped_plans(cuda, seed, tile):
    rng = np.random.default_rng(2000 + seed)
    layout = ["NN", "NT", "TN", "TT"][rng.integers(4)]
    name = ["bf16", "f16"][rng.integers(2)]
    count = int(rng.integers(5, 24))
    print(f"case grouped seed={seed} {name} {layout} tile={tile} problems={count}", flush=True)
    probs, As, Bs, Cs, Ds, host = [], [], [], [], [], []
    for _ in range(count):
        M, N = int(rng.integers(1, 600)), int(rng.integers(1, 600))
        K = int(rng.integers(0, 50)) * 8
        alpha, beta = [1.0, 0.5][rng.integers(2)], [0.0, 1.0, -0.5][rng.integers(3)]
        (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
        lda, ldb, ldd = _round8(max(ra, 1)), _round8(max(rb, 1)), _round8(M)
        A = colmajor(rng, max(ra, 1), max(ca, 1), lda, kind="normal", scale=1 / np.sqrt(max(K, 1)))
        B = colmajor(rng, max(rb, 1), max(cb, 1), ldb, kind="normal", scale=1 / np.sqrt(max(K, 1)))
        Cm = colmajor(rng, M, N, ldd)
        bits = [quantize_bits(x, name) for x in (A, B, Cm)]
        As.append(_dev16(bits[0], name, cuda))
        Bs.append(_dev16(bits[1], name, cuda))
        Cs.append(_dev16(bits[2], name, cuda))
        Ds.append(torch.full((ldd * N,), 7.0, dtype=As[-1].dtype, device=cuda))
        probs.append(tl.make_problem(layout, As[-1].dtype, M, N, K, lda, ldb, ldd, ldd, alpha=alpha, beta=beta))
        host.append((M, N, K, lda, ldb, ldd, alpha, beta, bits))
    plan = grouped.GroupedPlan(probs, As, Bs, Cs, Ds, tile=tile)
    plan.run()
    torch.cuda.synchronize()
    plan.close()
    for (M, N, K, lda, ldb, ldd, alpha, beta, bits), D in zip(host, Ds):
        f = [bits_to_f32(b, name).astype(np.float64) for b in bits]
        ref = oracle.gemm_f64(layout, M, N, K, alpha, f[0], lda, f[1], ldb, beta, f[2], ldd, ldd)
        got = D.float().cpu().numpy()
        bad, worst = elem_ok(extract(got, M, N, ldd), extract(ref, M, N, ldd))
        assert bad == 0, (seed, layout, name, M, N, K, tile, worst)
        assert np.all(got.reshape(N, ldd)[:, M:] == 7.0)


# fp32 tiles run the first engine (two K-tiles in flight up to 128 x 128)
_F32_TILES = [None, (256, 256), (256, 128), (128, 256), (128, 128), (64, 64)]


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("form", ["f32", "bf16x3_planes", "bf16x3_regsplit"])
def test_random_fp32_forms(cuda, seed, form):
    """fp32 MFMA and both bf16x3 paths (workspace planes, in-register split):
    rel-Frobenius <= 1e-5 against FP64 of the fp32 operands."""
    rng = np.random.default_rng(3000 + seed)
    for case in range(4):
        layout = ["NN", "NT", "TN", "TT"][rng.integers(4)]
        M, N = int(rng.integers(1, 500)), int(rng.integers(1, 500))
        K = int(rng.integers(1, 120)) * 4         # K-contiguous fp32 operands need K % 4 == 0
        beta = [0.0, 1.0, -0.5][rng.integers(3)]
        # fp32 MFMA also on its 16 x 16 x 4 form (mfma_mn = 16)
        tiles = _F32_TILES + ([(64, 64, 0, 0, 0, 16), (128, 128, 0, 0, 0, 16)] if form == "f32" else [])
        tile = tiles[rng.integers(len(tiles))]
        print(f"case {form} seed={seed}.{case} {layout} M={M} N={N} K={K} b={beta} tile={tile}", flush=True)
        (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
        lda, ldb, ldd = (ra + 3) // 4 * 4, (rb + 3) // 4 * 4, (M + 3) // 4 * 4
        A = colmajor(rng, ra, ca, lda, kind="normal")
        B = colmajor(rng, rb, cb, ldb, kind="normal")
        Cm = colmajor(rng, M, N, ldd)
        dev = lambda x: torch.from_numpy(x).to(cuda)
        compute = C.COMPUTE_DEFAULT if form == "f32" else C.COMPUTE_BF16X3
        D = tl.gemm(layout, M, N, K, dev(A), dev(B), dev(Cm) if beta else None, lda=lda, ldb=ldb,
                    ldc=ldd if beta else None, ldd=ldd, beta=beta, compute=compute, tile=tile,
                    workspace=None if form == "bf16x3_regsplit" else "auto")
        torch.cuda.synchronize()
        ref = oracle.gemm_f64(layout, M, N, K, 1.0, A.astype(np.float64), lda, B.astype(np.float64), ldb, beta,
                              Cm.astype(np.float64) if beta else None, ldd if beta else None, ldd)
        g, r = extract(D.cpu().numpy(), M, N, ldd), extract(ref, M, N, ldd)
        err = float(np.linalg.norm(g - r) / max(np.linalg.norm(r), 1e-300))
        assert err <= 1e-5, (form, seed, case, layout, M, N, K, tile, err)


@pytest.mark.parametrize("seed", range(6))
def test_random_fp8(cuda, seed):
    """FP8 e4m3 TN on every tile: rel-Frobenius <= 5e-5 against FP64 of the
    quantised operands (fp32 out), per element 1e-2 (bf16 out)."""
    rng = np.random.default_rng(4000 + seed)
    for case in range(4):
        M, N = int(rng.integers(1, 500)), int(rng.integers(1, 500))
        K = int(rng.integers(1, 40)) * 16           # 16-byte rows: K % 16 == 0 for e4m3
        tile = [None, (256, 256), (128, 128), (64, 64)][rng.integers(4)]
        out = ["f32", "bf16"][rng.integers(2)]
        print(f"case fp8 seed={seed}.{case} M={M} N={N} K={K} tile={tile} out={out}", flush=True)
        A = rng.standard_normal(M * K).astype(np.float32)
        B = rng.standard_normal(N * K).astype(np.float32)
        sa, sb = np.float32(oracle.fp8_scale(np.abs(A).max())), np.float32(oracle.fp8_scale(np.abs(B).max()))
        qa, qb = oracle.f32_to_e4m3_bits(A, sa), oracle.f32_to_e4m3_bits(B, sb)
        dt = torch.float32 if out == "f32" else torch.bfloat16
        ldd = _round8(M)
        D = tl.gemm("TN", M, N, K, torch.from_numpy(qa).to(cuda), torch.from_numpy(qb).to(cuda), lda=K, ldb=K,
                    ldd=ldd, scaleA=torch.tensor([sa], device=cuda), scaleB=torch.tensor([sb], device=cuda),
                    dtype_out=dt, tile=tile)
        torch.cuda.synchronize()
        deq_a = oracle.e4m3_bits_to_f32(qa).astype(np.float64) / sa
        deq_b = oracle.e4m3_bits_to_f32(qb).astype(np.float64) / sb
        ref = extract(oracle.gemm_f64("TN", M, N, K, 1.0, deq_a, K, deq_b, K, ldd=ldd), M, N, ldd)
        got = extract(D.float().cpu().numpy(), M, N, ldd)
        if out == "f32":
            err = float(np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300))
            assert err <= 5e-5, (seed, case, M, N, K, tile, err)
        else:
            bad, worst = elem_ok(got, ref)
            assert bad == 0, (seed, case, M, N, K, tile, worst)
