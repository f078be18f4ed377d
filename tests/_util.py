"""Helpers shared by the tests: column-major test data and comparisons.

Matrices are flat float32 numpy arrays in column-major order with an explicit
leading dimension, exactly as the reference's DataMem buffers
(src/gemmMfma.cpp:210-235).
"""
from __future__ import annotations

import numpy as np

import oracle


def colmajor(rng, rows, cols, ld=None, kind="uniform", scale=1.0):
    ld = ld or rows
    x = np.zeros(ld * cols, dtype=np.float32)
    if kind == "uniform":
        v = rng.uniform(-1.0, 1.0, size=(cols, rows))
    elif kind == "normal":
        v = rng.standard_normal(size=(cols, rows))
    elif kind == "int":
        v = rng.integers(-4, 5, size=(cols, rows))
    else:
        raise ValueError(kind)
    x.reshape(cols, ld)[:, :rows] = (v * scale).astype(np.float32)
    return x


def op_dims(layout, M, N, K):
    """(rows, cols) of the stored A and B for a column-major layout."""
    ta, tb = layout[0] == "T", layout[1] == "T"
    return ((K, M) if ta else (M, K)), ((N, K) if tb else (K, N))


def quantize_bits(x32, dtype_name, truncate=False):
    """fp32 -> 16-bit storage bits via the oracle converters (RNE unless truncate)."""
    if dtype_name == "bf16":
        return oracle.f32_to_bf16_bits(x32, truncate=truncate)
    return oracle.f32_to_f16_bits(x32)


def bits_to_f32(bits, dtype_name):
    if dtype_name == "bf16":
        return oracle.bf16_bits_to_f32(bits)
    return oracle.f16_bits_to_f32(bits)


def extract(D, M, N, ld, batch=1, bs=0):
    """Column-major flat buffer -> (batch, N, M) array of the valid region."""
    out = []
    for b in range(batch):
        base = b * bs
        out.append(np.stack([D[base + n * ld: base + n * ld + M] for n in range(N)]))
    return np.stack(out)


def rel_fro(x, ref):
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300))


def elem_ok(x, ref, rel=1e-2):
    """Per-element check |x - ref| <= rel * (|ref| + rms(ref)).

    The rms term covers outputs that cancel to near zero, where an fp32
    accumulation of O(1) products cannot hold a pure relative bound."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    rms = float(np.sqrt(np.mean(ref * ref))) if ref.size else 0.0
    bad = np.abs(x - ref) > rel * (np.abs(ref) + rms)
    return int(bad.sum()), float(np.max(np.abs(x - ref) / (np.abs(ref) + rms + 1e-300))) if ref.size else 0.0
