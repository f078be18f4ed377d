"""ctypes binding of libffgemm.so (include/ffgemm.h).

The library is built in-tree by ``python -m tensilelite_amd.build``; loading
fails loudly when it is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FFGEMM_LIB") or os.path.join(HERE, "lib", "libffgemm.so")

SUCCESS, WARN, ERR, FATAL = 0, 1, 2, 3
NN, NT, TN, TT = 0, 1, 2, 3
LAYOUTS = {"NN": NN, "NT": NT, "TN": TN, "TT": TT}
F32, F16, BF16, F8E4M3, F64 = 1, 2, 3, 4, 5
COMPUTE_DEFAULT, COMPUTE_BF16X3 = 0, 1
ROUND_RNE, ROUND_TRUNCATE = 0, 1


class Problem(ctypes.Structure):
    """ffgemm_problem."""
    _fields_ = [
        ("layout", ctypes.c_int32), ("dtype", ctypes.c_int32), ("dtype_out", ctypes.c_int32),
        ("compute", ctypes.c_int32), ("round", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("M", ctypes.c_int64), ("N", ctypes.c_int64), ("K", ctypes.c_int64), ("batch", ctypes.c_int64),
        ("lda", ctypes.c_int64), ("ldb", ctypes.c_int64), ("ldc", ctypes.c_int64), ("ldd", ctypes.c_int64),
        ("bsA", ctypes.c_int64), ("bsB", ctypes.c_int64), ("bsC", ctypes.c_int64), ("bsD", ctypes.c_int64),
        ("alpha", ctypes.c_float), ("beta", ctypes.c_float),
        ("scaleA", ctypes.c_void_p), ("scaleB", ctypes.c_void_p),
    ]


class Tile(ctypes.Structure):
    """ffgemm_tile."""
    _fields_ = [
        ("mt_m", ctypes.c_int32), ("mt_n", ctypes.c_int32), ("depth_u", ctypes.c_int32),
        ("waves_m", ctypes.c_int32), ("waves_n", ctypes.c_int32), ("mfma_mn", ctypes.c_int32),
        ("group_m", ctypes.c_int32), ("lds_buffers", ctypes.c_int32), ("split_k", ctypes.c_int32),
        ("repack", ctypes.c_int32),
    ]

    def as_tuple(self):
        return (self.mt_m, self.mt_n)


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.isfile(LIB_PATH):
        raise RuntimeError(f"libffgemm.so not built ({LIB_PATH}); run python -m tensilelite_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, i32, f32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_float
    PP, PT = ctypes.POINTER(Problem), ctypes.POINTER(Tile)
    L.ffgemm_init.restype = i32
    L.ffgemm_shutdown.restype = None
    L.ffgemm_status_string.argtypes = [i32]; L.ffgemm_status_string.restype = ctypes.c_char_p
    L.ffgemm_last_error.restype = ctypes.c_char_p
    L.ffgemm_version.restype = ctypes.c_char_p
    L.ffgemm_device_cu_count.restype = i32
    L.ffgemm_query_tile.argtypes = [PP, PT]; L.ffgemm_query_tile.restype = i32
    L.ffgemm_query_tile_ws.argtypes = [PP, PT]; L.ffgemm_query_tile_ws.restype = i32
    L.ffgemm_tile_from_reference.argtypes = [PP] + [i32] * 7 + [PT]
    L.ffgemm_tile_from_reference.restype = i32
    L.ffgemm_tune_db_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(i32)]; L.ffgemm_tune_db_load.restype = i32
    L.ffgemm_run.argtypes = [PP, vp, vp, vp, vp, PT, vp]; L.ffgemm_run.restype = i32
    L.ffgemm_workspace_size.argtypes = [PP, ctypes.POINTER(ctypes.c_size_t)]
    L.ffgemm_workspace_size.restype = i32
    L.ffgemm_workspace_size_ex.argtypes = [PP, PT, ctypes.POINTER(ctypes.c_size_t)]
    L.ffgemm_workspace_size_ex.restype = i32
    L.ffgemm_run_ws.argtypes = [PP, vp, vp, vp, vp, PT, vp, ctypes.c_size_t, vp]; L.ffgemm_run_ws.restype = i32
    L.ffgemm_run_x3_noalloc.argtypes = [PP, vp, vp, vp, vp, PT, vp]; L.ffgemm_run_x3_noalloc.restype = i32
    u32 = ctypes.c_uint32
    L.ffgemm_run_tn_kernarg.argtypes = [vp, vp, vp, vp, f32, f32] + [u32] * 12 + [i32, PT, vp]
    L.ffgemm_run_tensile_args.argtypes = [vp, ctypes.c_size_t, i32, PT, vp]; L.ffgemm_run_tensile_args.restype = i32
    L.ffgemm_run_tn_kernarg.restype = i32
    L.ffgemm_fp8_workspace_size.restype = ctypes.c_size_t
    L.ffgemm_fp8_quantize.argtypes = [vp, i64, i64, i64, i32, vp, i64, vp, vp, vp]
    L.ffgemm_fp8_quantize.restype = i32
    L.ffgemm_convert.argtypes = [vp, vp, i64, i32, vp]; L.ffgemm_convert.restype = i32
    L.ffgemm_grouped_plan_create.argtypes = [ctypes.POINTER(vp), i32, PP, ctypes.POINTER(vp),
                                             ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), PT]
    L.ffgemm_grouped_plan_create.restype = i32
    L.ffgemm_grouped_plan_run.argtypes = [vp, vp]; L.ffgemm_grouped_plan_run.restype = i32
    L.ffgemm_grouped_plan_info.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(ctypes.c_int32), PT]
    L.ffgemm_grouped_plan_info.restype = i32
    L.ffgemm_grouped_plan_destroy.argtypes = [vp]; L.ffgemm_grouped_plan_destroy.restype = i32
    _lib = L
    return L


class FFGemmError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{lib().ffgemm_status_string(status).decode()}: {msg}")
        self.status = status


def check(status: int) -> None:
    if status != SUCCESS:
        raise FFGemmError(status, lib().ffgemm_last_error().decode())
