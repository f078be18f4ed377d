"""Pins the CPU oracle (no GPU): golden KATs, the reference's own converters
(compiled from its header into oracle/_ref when available), closed forms,
library routines (numpy/torch) and brute force on tiny inputs."""
from __future__ import annotations

import glob
import os

import numpy as np
import pytest
import torch

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build(with_ref=True)


# ---------------------------------------------------------------- InitRealData
def test_init_real_data_patterns_hand_values():
    a = oracle.init_real_data(2.4, 5, 3)
    assert a.tolist() == [0, 1, 2, 3, 4, 5, 6, 0, 1, 2, 3, 4, 5, 6, 0]
    b = oracle.init_real_data(-2.4, 4, 4)
    assert b.tolist() == [0, -1, -2, -3, -4, -5, -6, -7, -8, -9, -10, 0, -1, -2, -3, -4]
    assert np.all(oracle.init_real_data(0.5, 8, 2) == 0)
    assert np.all(oracle.init_real_data(-55.55, 3, 2) == np.float32(-55.55))
    c = oracle.init_real_data(2.2, 3, 2, 2)   # (i+1)*0.01 + ch + 100*bt
    np.testing.assert_allclose(c[[0, 4, 11]], [0.01, 1.02, 101.03], rtol=1e-6)


# ---------------------------------------------------------------- converters
def _sample_f32(n=1 << 18, seed=0):
    rng = np.random.default_rng(seed)
    mag = rng.standard_normal(n).astype(np.float32) * np.float32(2.0) ** rng.integers(-30, 18, n).astype(np.float32)
    raw = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32).view(np.float32)
    edge = np.array([0.0, -0.0, 65504, 65519.99, 65520, -65520, 6.1035156e-05, 5.96e-8, 2.98e-8, 2.9802322e-08,
                     np.inf, -np.inf, np.nan, 1 / 3, 0.1, 1e-8, 1e30], dtype=np.float32)
    return np.concatenate([mag, raw, edge])


def test_f32_to_f16_matches_ieee_rne():
    x = _sample_f32()
    got = oracle.f32_to_f16_bits(x)
    want = x.astype(np.float16).view(np.uint16)
    nan = np.isnan(x)
    np.testing.assert_array_equal(got[~nan], want[~nan])
    # the reference maps every NaN to sign|0x7e00 (inc/ff_utils.h:107-108)
    assert np.all((got[nan] & 0x7fff) == 0x7e00)


def test_f32_to_f16_known_probe_values():
    # values probed from the reference's converter
    assert oracle.lib().orc_f32_to_f16(0.1) == 0x2E66
    assert oracle.lib().orc_f32_to_f16(1 / 3) == 0x3555


def test_converters_bit_exact_vs_reference_header():
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    x = _sample_f32(1 << 17, seed=1)
    want = np.empty(x.size, dtype=np.uint16)
    ref.ref_cvt_f32_to_f16_n(x.ctypes.data, want.ctypes.data, x.size)
    np.testing.assert_array_equal(oracle.f32_to_f16_bits(x), want)
    for v in x[:2000]:
        assert oracle.lib().orc_f32_to_bf16_trunc(float(v)) == (ref.ref_cvt_f32_to_bf16(float(v)) & 0xFFFF)
    for h in range(65536):
        a = np.float32(oracle.lib().orc_f16_to_f32_quirk(h))
        b = np.float32(ref.ref_cvt_f16_to_f32(h))
        assert a.view(np.uint32) == b.view(np.uint32) or (np.isnan(a) and np.isnan(b)), hex(h)
        assert oracle.lib().orc_bf16_to_f32(h) == ref.ref_cvt_bf16_to_f32(h) or np.isnan(ref.ref_cvt_bf16_to_f32(h))


def test_f16_to_f32_exhaustive_ieee():
    h = np.arange(65536, dtype=np.uint16)
    got = oracle.f16_bits_to_f32(h)
    want = h.view(np.float16).astype(np.float32)
    nan = np.isnan(want)
    np.testing.assert_array_equal(got[~nan].view(np.uint32), want[~nan].view(np.uint32))
    assert np.all(np.isnan(got[nan]))


def test_f16_to_f32_quirk_closed_form():
    q = oracle.lib().orc_f16_to_f32_quirk
    assert abs(q(0x00A8) - 3.5524368e-05) < 1e-11   # subnormal: 2^-15 * (1 + 168/1024)
    assert q(0x7C00) == 65536.0                        # inf decodes as 2^16
    assert q(0x3C00) == 1.0 and q(0xBC00) == -1.0      # normals are right
    assert q(0) == 0.0


def test_bf16_rne_vs_torch_and_trunc_closed_form():
    x = _sample_f32(1 << 16, seed=3)
    fin = np.isfinite(x)
    got = oracle.f32_to_bf16_bits(x[fin])
    want = torch.from_numpy(x[fin]).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    np.testing.assert_array_equal(got, want)
    tr = oracle.f32_to_bf16_bits(x, truncate=True)
    np.testing.assert_array_equal(tr, (x.view(np.uint32) >> 16).astype(np.uint16))
    assert oracle.lib().orc_f32_to_bf16_trunc(0.1) == 0x3DCC        # probe value
    assert oracle.lib().orc_f32_to_bf16_trunc(65504.0) == 0x477F


def test_split_bf16x3_properties():
    x = _sample_f32(1 << 14, seed=4)
    x = x[np.isfinite(x) & (np.abs(x) < 1e30) & (np.abs(x) > 1e-30)]
    hi, lo = oracle.split_bf16x3(x)
    h = oracle.bf16_bits_to_f32(hi).astype(np.float64)
    l = oracle.bf16_bits_to_f32(lo).astype(np.float64)
    np.testing.assert_array_equal(hi, oracle.f32_to_bf16_bits(x))
    # |x - hi - lo| <= 2^-16 |x|: two 8-bit significands cover 16+ bits
    assert np.all(np.abs(x - h - l) <= np.abs(x) * 2.0 ** -16)
    # lo is x - hi rounded to nearest even in bf16: x - hi is exact in fp32
    # (Sterbenz), so torch's RNE bf16 cast of that fp32 difference pins every
    # bit of lo.  A truncating lo meets the error bound above but fails here.
    diff = (x.astype(np.float64) - h).astype(np.float32)
    assert np.array_equal(diff.astype(np.float64), x.astype(np.float64) - h)
    want_lo = torch.from_numpy(diff).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    np.testing.assert_array_equal(lo, want_lo)
    trunc_lo = (diff.view(np.uint32) >> 16).astype(np.uint16)
    assert np.count_nonzero(trunc_lo != want_lo) > x.size // 4   # the pin separates the two roundings


# ---------------------------------------------------------------- e4m3
def test_e4m3_decode_all_codes_closed_form_and_torch():
    codes = np.arange(256, dtype=np.uint8)
    got = oracle.e4m3_bits_to_f32(codes)
    for c in range(256):
        e, m, s = (c >> 3) & 15, c & 7, -1.0 if c & 0x80 else 1.0
        if e == 15 and m == 7:
            assert np.isnan(got[c])
        elif e == 0:
            assert got[c] == s * m * 2.0 ** -9
        else:
            assert got[c] == s * (1 + m / 8) * 2.0 ** (e - 7)
    if hasattr(torch, "float8_e4m3fn"):
        t = torch.from_numpy(codes).view(torch.float8_e4m3fn).float().numpy()
        ok = ~np.isnan(t)
        np.testing.assert_array_equal(got[ok], t[ok])


def test_e4m3_encode_roundtrip_rne_saturation():
    codes = np.array([c for c in range(256) if (c & 0x7F) != 0x7F], dtype=np.uint8)
    vals = oracle.e4m3_bits_to_f32(codes)
    back = oracle.f32_to_e4m3_bits(vals)
    np.testing.assert_array_equal(back[vals != 0], codes[vals != 0])
    enc = oracle.lib().orc_f32_to_e4m3
    assert enc(1000.0) == 0x7E and enc(-1e9) == 0xFE and enc(float("inf")) == 0x7E   # saturate to 448
    assert enc(1.0625) == 0x38          # midpoint of 1.0 and 1.125 -> even (1.0)
    assert enc(1.1875) == 0x3A          # midpoint of 1.125 and 1.25 -> even (1.25)
    assert enc(2.0 ** -10) == 0x00 and enc(3 * 2.0 ** -10) == 0x02   # subnormal ties to even
    if hasattr(torch, "float8_e4m3fn"):
        x = _sample_f32(1 << 15, seed=5)
        x = x[np.isfinite(x)]
        x = np.clip(x, -448, 448)
        want = torch.from_numpy(x).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
        np.testing.assert_array_equal(oracle.f32_to_e4m3_bits(x), want)


# ---------------------------------------------------------------- GEMM oracle
@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "kat_*.npz"))))
def test_gemm_f32_golden_kat(path):
    d = np.load(path)   # allow_pickle defaults to False
    M, N, K = int(d["M"]), int(d["N"]), int(d["K"])
    ldd = int(d["ldd"])
    D = oracle.gemm_f32("TN", M, N, K, float(d["alpha"]), d["A"], int(d["lda"]), d["B"], int(d["ldb"]),
                        float(d["beta"]), d["C"], int(d["ldc"]), ldd)
    got = D.reshape(N, ldd)[:, :M]
    np.testing.assert_array_equal(got, d["D"])
    # beta = 0, alpha = 1 makes D the exact integer sums
    if float(d["beta"]) == 0.0 and float(d["alpha"]) == 1.0:
        np.testing.assert_array_equal(got.astype(np.int64), d["sums"])


@pytest.mark.parametrize("layout", ["NN", "NT", "TN", "TT"])
def test_gemm_oracles_vs_torch_float64(layout):
    rng = np.random.default_rng(7)
    M, N, K = 37, 29, 23
    ta, tb = layout[0] == "T", layout[1] == "T"
    lda, ldb, ldc, ldd = (K if ta else M) + 3, (N if tb else K) + 5, M + 2, M + 4
    A = rng.uniform(-1, 1, lda * (M if ta else K)).astype(np.float32)
    B = rng.uniform(-1, 1, ldb * (K if tb else N)).astype(np.float32)
    Cm = rng.uniform(-1, 1, ldc * N).astype(np.float32)
    At = torch.from_numpy(A.astype(np.float64)).reshape(-1, lda)[:, : (K if ta else M)]
    Bt = torch.from_numpy(B.astype(np.float64)).reshape(-1, ldb)[:, : (N if tb else K)]
    opA = At if ta else At.T    # (M, K)
    opB = Bt if tb else Bt.T    # (K, N): N-layout rows are columns n of length K
    ref = (0.5 * opA @ opB + 0.25 * torch.from_numpy(Cm.astype(np.float64)).reshape(N, ldc)[:, :M].T).numpy()
    d64 = oracle.gemm_f64(layout, M, N, K, 0.5, A, lda, B, ldb, 0.25, Cm, ldc, ldd)
    d32 = oracle.gemm_f32(layout, M, N, K, 0.5, A, lda, B, ldb, 0.25, Cm, ldc, ldd)
    np.testing.assert_allclose(d64.reshape(N, ldd)[:, :M].T, ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(d32.reshape(N, ldd)[:, :M].T, ref, rtol=1e-5, atol=1e-5)


def test_gemm_f32_brute_force_tiny_and_mt():
    rng = np.random.default_rng(8)
    M, N, K = 3, 2, 4
    A = rng.integers(-3, 4, M * K).astype(np.float32)     # NN: A[m + k*M]
    B = rng.integers(-3, 4, K * N).astype(np.float32)     # B[k + n*K]
    D = oracle.gemm_f32("NN", M, N, K, 1.0, A, M, B, K)
    for m in range(M):
        for n in range(N):
            s = sum(float(A[m + k * M]) * float(B[k + n * K]) for k in range(K))
            assert D[m + n * M] == s
    # the threaded variant runs the identical per-element loop
    M, N, K = 64, 57, 33
    A = rng.standard_normal(M * K).astype(np.float32)
    B = rng.standard_normal(K * N).astype(np.float32)
    d1 = oracle.gemm_f32("NN", M, N, K, 1.0, A, M, B, K)
    d4 = oracle.gemm_f32("NN", M, N, K, 1.0, A, M, B, K, nthreads=4)
    np.testing.assert_array_equal(d1, d4)


def test_gemm_planted_mistakes_are_caught():
    """A transposed operand, a dropped beta term or a wrong leading dimension
    must fail the checks above: plant each in the oracle call and confirm the
    torch float64 reference disagrees with it (while the correct call agrees)."""
    rng = np.random.default_rng(9)
    M = N = K = 8
    lda = M + 4
    A = rng.uniform(-1, 1, lda * K).astype(np.float32)
    B = rng.uniform(-1, 1, K * N).astype(np.float32)
    Cm = rng.uniform(-1, 1, M * N).astype(np.float32)
    opA = torch.from_numpy(A.astype(np.float64)).reshape(K, lda)[:, :M].T    # (M, K)
    opB = torch.from_numpy(B.astype(np.float64)).reshape(N, K).T             # (K, N)
    ref = (opA @ opB + torch.from_numpy(Cm.astype(np.float64)).reshape(N, M).T).numpy()

    def dev(d):
        return np.max(np.abs(d.reshape(N, M).T - ref))

    good = oracle.gemm_f64("NN", M, N, K, 1.0, A, lda, B, K, 1.0, Cm, M, M)
    assert dev(good) < 1e-12
    wrong_layout = oracle.gemm_f64("TN", M, N, K, 1.0, A, lda, B, K, 1.0, Cm, M, M)
    no_beta = oracle.gemm_f64("NN", M, N, K, 1.0, A, lda, B, K, 0.0, Cm, M, M)
    wrong_ld = oracle.gemm_f64("NN", M, N, K, 1.0, A, M, B, K, 1.0, Cm, M, M)
    assert dev(wrong_layout) > 1e-3
    assert dev(no_beta) > 1e-3
    assert dev(wrong_ld) > 1e-3
