"""Writes the golden known-answer fixtures for the oracle tests.

The reference commits no numeric golden output; its own check is a
known-answer test in disguise: InitRealData fills A with idx % 7 and B with
-(idx % 11) (ff_gpu_src/ff_gpu.cpp:134-136, src/gemmMfma.cpp:231-234), so
every product and partial sum is an integer and the fp32 CPU loop
(src/gemmMfma.cpp:325-346) computes sum exactly.  This script computes the
expected D = alpha*sum + beta*c with exact integer sums (numpy int64) and one
final float64 -> float32 rounding -- independently of the oracle -- and
stores inputs and outputs as .npz (no pickles).

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def pattern(init_type: float, dim0: int, dim1: int) -> np.ndarray:
    idx = np.arange(dim0 * dim1, dtype=np.int64)
    if init_type == 2.4:
        return (idx % 7).astype(np.float32)
    if init_type == -2.4:
        return (-(idx % 11)).astype(np.float32)
    raise ValueError(init_type)


def kat(M: int, N: int, K: int, pad: int, alpha: float, beta: float):
    lda = ldb = K + pad
    ldc = ldd = M + pad
    A = pattern(2.4, lda, M)
    B = pattern(-2.4, ldb, N)
    C = pattern(2.4, ldc, N)
    Ai = A.reshape(M, lda)[:, :K].astype(np.int64)       # (m, k) at lda*m + k
    Bi = B.reshape(N, ldb)[:, :K].astype(np.int64)       # (n, k) at ldb*n + k
    S = Bi @ Ai.T                                          # (n, m) exact integer sums
    Cv = C.reshape(N, ldc)[:, :M].astype(np.float64)
    a32, b32 = np.float32(alpha), np.float32(beta)
    # reference order: t = alpha*t (fp32), bc = beta*c (fp32), t + bc (fp32)
    t = (np.float64(a32) * S.astype(np.float64)).astype(np.float32)
    bc = (np.float64(b32) * Cv).astype(np.float32)
    D = (t.astype(np.float64) + bc.astype(np.float64)).astype(np.float32)
    return dict(M=M, N=N, K=K, pad=pad, alpha=a32, beta=b32, lda=lda, ldb=ldb, ldc=ldc, ldd=ldd,
                A=A, B=B, C=C, sums=S.astype(np.int64), D=D)


def main():
    cases = {
        "kat_tn_64x48x40_p32": kat(64, 48, 40, 32, 0.314, 0.2718),
        "kat_tn_96x80x128_p0": kat(96, 80, 128, 0, 0.314, 0.2718),
        "kat_tn_32x16x520_p8_beta0": kat(32, 16, 520, 8, 1.0, 0.0),
    }
    for name, d in cases.items():
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
        print("wrote", name)


if __name__ == "__main__":
    main()
