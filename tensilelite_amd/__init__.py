"""tensilelite_amd -- MI355X (gfx950) GEMM for the TensileLite gemmMfma path.

The compute path is libffgemm.so (hand-written HIP kernels for gfx950 behind
the C ABI in include/ffgemm.h).  This package binds it with ctypes and adds
the torch-facing pieces: row-major ``mm``, FP8 quantisation, grouped-GEMM
plans, the autograd Function and the autotuner.
"""
from . import _capi, fp8
from ._capi import FFGemmError, Problem, Tile, check, lib
from .gemm import (convert, gemm, make_problem, make_tile, mm, query_tile, query_tile_ws, run, tile_from_reference,
                   workspace_size)

__all__ = ["fp8", "FFGemmError", "Problem", "Tile", "check", "lib", "convert", "gemm", "make_problem",
           "make_tile", "mm", "query_tile", "query_tile_ws", "run", "tile_from_reference", "workspace_size", "_capi"]
