"""Seeded random problems against the FP64 oracle: every layout, 16-bit dtype,
tile record (macro tile, wave tiling, MFMA instruction), ragged M / N / K,
padded leading dimensions and alpha / beta with C, single and grouped."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
import tensilelite_amd as tl
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
