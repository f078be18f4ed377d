"""CPU oracle for the gemmMfma path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this package, and only to check results (or as
the timed CPU baseline).  The product package ``tensilelite_amd`` never imports
it.  The C restatement lives in ``oracle.c`` (each function cites the
TensileLite file:line it follows); this module loads it with ctypes and adds
numpy conveniences.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref_cvt.so")

NN, NT, TN, TT = 0, 1, 2, 3
LAYOUTS = {"NN": NN, "NT": NT, "TN": TN, "TT": TT}

_lib = None
_ref = None


def build(with_ref: bool = False) -> None:
    """Compile the oracle (and oracle/_ref when the reference tree exists)."""
    targets = ["all"]
    if with_ref and os.path.isfile("/root/reference/inc/ff_utils.h"):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", _HERE] + targets, check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.isfile(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        i64, f32, f64, vp, ci = ctypes.c_int64, ctypes.c_float, ctypes.c_double, ctypes.c_void_p, ctypes.c_int
        L.orc_init_real_data.argtypes = [vp, f32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.orc_f32_to_f16.argtypes = [f32]; L.orc_f32_to_f16.restype = ctypes.c_uint16
        L.orc_f16_to_f32.argtypes = [ctypes.c_uint16]; L.orc_f16_to_f32.restype = f32
        L.orc_f16_to_f32_quirk.argtypes = [ctypes.c_uint16]; L.orc_f16_to_f32_quirk.restype = f32
        L.orc_f32_to_bf16_trunc.argtypes = [f32]; L.orc_f32_to_bf16_trunc.restype = ctypes.c_uint16
        L.orc_f32_to_bf16_rne.argtypes = [f32]; L.orc_f32_to_bf16_rne.restype = ctypes.c_uint16
        L.orc_bf16_to_f32.argtypes = [ctypes.c_uint16]; L.orc_bf16_to_f32.restype = f32
        L.orc_f32_to_e4m3.argtypes = [f32]; L.orc_f32_to_e4m3.restype = ctypes.c_uint8
        L.orc_e4m3_to_f32.argtypes = [ctypes.c_uint8]; L.orc_e4m3_to_f32.restype = f32
        L.orc_cvt_f32_to_f16_n.argtypes = [vp, vp, i64]
        L.orc_cvt_f32_to_bf16_n.argtypes = [vp, vp, i64, ci]
        L.orc_cvt_f16_to_f32_n.argtypes = [vp, vp, i64]
        L.orc_cvt_bf16_to_f32_n.argtypes = [vp, vp, i64]
        L.orc_cvt_f32_to_e4m3_n.argtypes = [vp, vp, i64, f32]
        L.orc_cvt_e4m3_to_f32_n.argtypes = [vp, vp, i64]
        L.orc_split_bf16x3_n.argtypes = [vp, vp, vp, i64]
        g32 = [ci, i64, i64, i64, i64, f32, vp, i64, i64, vp, i64, i64, f32, vp, i64, i64, vp, i64, i64]
        L.orc_gemm_f32.argtypes = g32
        L.orc_gemm_f32_mt.argtypes = [ci] + g32
        L.orc_gemm_f64.argtypes = [ci, i64, i64, i64, i64, f64, vp, i64, i64, vp, i64, i64, f64, vp,
                                   i64, i64, vp, i64, i64]
        _lib = L
    return _lib


def ref_lib():
    """The reference's own converters compiled from its header, or None."""
    global _ref
    if _ref is None and os.path.isfile(_REF_PATH):
        R = ctypes.CDLL(_REF_PATH)
        R.ref_cvt_f32_to_f16.argtypes = [ctypes.c_float]; R.ref_cvt_f32_to_f16.restype = ctypes.c_int16
        R.ref_cvt_f16_to_f32.argtypes = [ctypes.c_uint16]; R.ref_cvt_f16_to_f32.restype = ctypes.c_float
        R.ref_cvt_f32_to_bf16.argtypes = [ctypes.c_float]; R.ref_cvt_f32_to_bf16.restype = ctypes.c_int16
        R.ref_cvt_bf16_to_f32.argtypes = [ctypes.c_uint16]; R.ref_cvt_bf16_to_f32.restype = ctypes.c_float
        R.ref_cvt_f32_to_f16_n.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        _ref = R
    return _ref


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def init_real_data(init_type: float, dim0: int, dim1: int = 1, dim2: int = 1) -> np.ndarray:
    x = np.zeros(dim0 * dim1 * dim2, dtype=np.float32)
    lib().orc_init_real_data(_p(x), ctypes.c_float(init_type), dim0, dim1, dim2)
    return x


def f32_to_f16_bits(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, dtype=np.uint16)
    lib().orc_cvt_f32_to_f16_n(_p(x), _p(out), x.size)
    return out


def f32_to_bf16_bits(x: np.ndarray, truncate: bool = False) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, dtype=np.uint16)
    lib().orc_cvt_f32_to_bf16_n(_p(x), _p(out), x.size, int(truncate))
    return out


def f16_bits_to_f32(h: np.ndarray) -> np.ndarray:
    h = np.ascontiguousarray(h, dtype=np.uint16)
    out = np.empty(h.shape, dtype=np.float32)
    lib().orc_cvt_f16_to_f32_n(_p(h), _p(out), h.size)
    return out


def bf16_bits_to_f32(b: np.ndarray) -> np.ndarray:
    b = np.ascontiguousarray(b, dtype=np.uint16)
    out = np.empty(b.shape, dtype=np.float32)
    lib().orc_cvt_bf16_to_f32_n(_p(b), _p(out), b.size)
    return out


def f32_to_e4m3_bits(x: np.ndarray, scale: float = 1.0) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, dtype=np.uint8)
    lib().orc_cvt_f32_to_e4m3_n(_p(x), _p(out), x.size, ctypes.c_float(scale))
    return out


def e4m3_bits_to_f32(q: np.ndarray) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.uint8)
    out = np.empty(q.shape, dtype=np.float32)
    lib().orc_cvt_e4m3_to_f32_n(_p(q), _p(out), q.size)
    return out


def split_bf16x3(x: np.ndarray):
    x = np.ascontiguousarray(x, dtype=np.float32)
    hi = np.empty(x.shape, dtype=np.uint16)
    lo = np.empty(x.shape, dtype=np.uint16)
    lib().orc_split_bf16x3_n(_p(x), _p(hi), _p(lo), x.size)
    return hi, lo


def fp8_scale(amax: float) -> float:
    """Per-tensor scale 448/amax computed in fp32, as the device does."""
    a = np.float32(amax)
    return float(np.float32(448.0) / a) if a > 0 else 1.0


def gemm_f32(layout, M, N, K, alpha, A, lda, B, ldb, beta=0.0, C=None, ldc=None, ldd=None,
             batch=1, bsA=0, bsB=0, bsC=0, bsD=0, nthreads=1) -> np.ndarray:
    """fp32 reference GEMM (gemmMfma.cpp:325-346). Arrays are flat column-major buffers."""
    layout = LAYOUTS.get(layout, layout)
    A = np.ascontiguousarray(A, dtype=np.float32).ravel()
    B = np.ascontiguousarray(B, dtype=np.float32).ravel()
    ldd = ldd or M
    ldc = ldc or ldd
    bsD = bsD or ldd * N
    D = np.zeros(max(bsD * batch, ldd * N), dtype=np.float32)
    if C is None:
        C = np.zeros(1, dtype=np.float32)
    C = np.ascontiguousarray(C, dtype=np.float32).ravel()
    args = (layout, M, N, K, batch, ctypes.c_float(alpha), _p(A), lda, bsA, _p(B), ldb, bsB,
            ctypes.c_float(beta), _p(C), ldc, bsC, _p(D), ldd, bsD)
    if nthreads > 1:
        lib().orc_gemm_f32_mt(nthreads, *args)
    else:
        lib().orc_gemm_f32(*args)
    return D


def gemm_f64(layout, M, N, K, alpha, A, lda, B, ldb, beta=0.0, C=None, ldc=None, ldd=None,
             batch=1, bsA=0, bsB=0, bsC=0, bsD=0) -> np.ndarray:
    layout = LAYOUTS.get(layout, layout)
    A = np.ascontiguousarray(A, dtype=np.float64).ravel()
    B = np.ascontiguousarray(B, dtype=np.float64).ravel()
    ldd = ldd or M
    ldc = ldc or ldd
    bsD = bsD or ldd * N
    D = np.zeros(max(bsD * batch, ldd * N), dtype=np.float64)
    if C is None:
        C = np.zeros(1, dtype=np.float64)
    C = np.ascontiguousarray(C, dtype=np.float64).ravel()
    lib().orc_gemm_f64(layout, M, N, K, batch, ctypes.c_double(alpha), _p(A), lda, bsA, _p(B), ldb,
                       bsB, ctypes.c_double(beta), _p(C), ldc, bsC, _p(D), ldd, bsD)
    return D
