"""Torch-facing wrappers over the C ABI.

Two views of the same kernels:

* ``gemm(...)`` -- the reference's column-major problem (src/gemmMfma.cpp:
  210-217): flat device buffers, explicit M/N/K, leading dimensions, layout,
  alpha/beta.  This is what the parity tests drive.
* ``mm(a, b, trans_a, trans_b)`` -- row-major torch matrices,
  ``c = op(a) @ op(b)``.  A row-major r x c tensor is the column-major c x r
  matrix, so ``c^T = op(b)^T op(a)^T`` runs as a column-major problem with the
  operands swapped.
"""
from __future__ import annotations

import ctypes

import torch

from . import _capi as C

_DT = {torch.float32: C.F32, torch.float16: C.F16, torch.bfloat16: C.BF16, torch.float64: C.F64,
       torch.uint8: C.F8E4M3}
if hasattr(torch, "float8_e4m3fn"):
    _DT[torch.float8_e4m3fn] = C.F8E4M3
_TORCH = {C.F32: torch.float32, C.F16: torch.float16, C.BF16: torch.bfloat16, C.F64: torch.float64}


def _ptr(t):
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def _stream(stream):
    if stream is None:
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, torch.cuda.Stream):
        return ctypes.c_void_p(stream.cuda_stream)
    return ctypes.c_void_p(stream)


def make_problem(layout, dtype, M, N, K, lda, ldb, ldc, ldd, alpha=1.0, beta=0.0, batch=1,
                 bsA=0, bsB=0, bsC=0, bsD=0, dtype_out=0, compute=C.COMPUTE_DEFAULT,
                 round=C.ROUND_RNE, scaleA=None, scaleB=None) -> C.Problem:
    p = C.Problem()
    p.layout = C.LAYOUTS.get(layout, layout)
    p.dtype = _DT.get(dtype, dtype)
    p.dtype_out = _DT.get(dtype_out, dtype_out) if dtype_out else 0
    p.compute = compute
    p.round = round
    p.M, p.N, p.K, p.batch = M, N, K, batch
    p.lda, p.ldb, p.ldc, p.ldd = lda, ldb, ldc, ldd
    p.bsA, p.bsB, p.bsC, p.bsD = bsA, bsB, bsC, bsD
    p.alpha, p.beta = alpha, beta
    p.scaleA = _ptr(scaleA)
    p.scaleB = _ptr(scaleB)
    return p


def make_tile(tile) -> C.Tile | None:
    """Tile record from (mt_m, mt_n[, group_m[, waves_m, waves_n[, mfma_mn[, split_k[, repack]]]]])
    or a Tile.

    waves_m x waves_n picks the wave tiling: 2 x 4 the eight-wave engine,
    2 x 2 a four-wave one, 0 x 0 (or absent) the tile's default.  mfma_mn 16
    picks 16 x 16 MFMAs at 64 x 64 and 128 x 128 (16 x 16 x 32 for bf16 /
    fp16, 16 x 16 x 4 for fp32); 0 or 32 the 32 x 32 ones.  split_k >= 2 cuts K into that many slices (with a
    workspace only).  repack 2 copies M/N-contiguous 16-bit operands K-contiguous into the workspace and runs
    the problem as TN, 1 never does, 0 leaves it to the library (mt_m 0 keeps the library's tile)."""
    if tile is None:
        return None
    if isinstance(tile, C.Tile):
        return tile
    t = C.Tile()
    t.mt_m, t.mt_n = int(tile[0]), int(tile[1])
    if len(tile) > 2:
        t.group_m = int(tile[2])
    if len(tile) > 4:
        t.waves_m, t.waves_n = int(tile[3]), int(tile[4])
    if len(tile) > 5:
        t.mfma_mn = int(tile[5])
    if len(tile) > 6:
        t.split_k = int(tile[6])
    if len(tile) > 7:
        t.repack = int(tile[7])
    return t


def workspace_size(problem: C.Problem, tile=None) -> int:
    """Device bytes ffgemm_run_ws can use for this problem under `tile` (None:
    the library's choice); 0: no workspace path."""
    n = ctypes.c_size_t()
    t = make_tile(tile)
    C.check(C.lib().ffgemm_workspace_size_ex(ctypes.byref(problem), ctypes.byref(t) if t is not None else None,
                                             ctypes.byref(n)))
    return n.value


def _tuned(problem: C.Problem):
    from .autotune import best_tile
    return best_tile(problem.dtype, ["NN", "NT", "TN", "TT"][problem.layout], problem.M, problem.N, problem.K)


def run(problem: C.Problem, A, B, Cm, D, tile=None, stream=None, workspace="auto") -> None:
    """ffgemm_run on device buffers (tensors or raw pointers).

    tile: None (the library's choice), a tile record / tuple (see make_tile),
    or "tuned" (the autotuner database's best for this problem, else None).
    workspace: "auto" allocates what ffgemm_workspace_size asks for (from
    torch's caching allocator) and runs ffgemm_run_ws; None runs ffgemm_run;
    a uint8 device tensor is used as given."""
    if isinstance(tile, str):
        if tile != "tuned":
            raise ValueError(f"tile={tile!r}: expected None, a tile, or 'tuned'")
        tile = _tuned(problem)
    t = make_tile(tile)
    tp = ctypes.byref(t) if t is not None else None
    if isinstance(workspace, str):
        n = workspace_size(problem, t)
        workspace = torch.empty(n, dtype=torch.uint8, device=torch.device("cuda", torch.cuda.current_device())) if n else None
    if workspace is None:
        C.check(C.lib().ffgemm_run(ctypes.byref(problem), _ptr(A), _ptr(B), _ptr(Cm), _ptr(D), tp, _stream(stream)))
    else:
        C.check(C.lib().ffgemm_run_ws(ctypes.byref(problem), _ptr(A), _ptr(B), _ptr(Cm), _ptr(D), tp,
                                      _ptr(workspace), workspace.numel(), _stream(stream)))


def query_tile(problem: C.Problem, tile=None) -> C.Tile:
    t = make_tile(tile) or C.Tile()
    C.check(C.lib().ffgemm_query_tile(ctypes.byref(problem), ctypes.byref(t)))
    return t


def query_tile_ws(problem: C.Problem, tile=None) -> C.Tile:
    """The record ffgemm_run_ws follows with a workspace (split_k included)."""
    t = make_tile(tile) or C.Tile()
    C.check(C.lib().ffgemm_query_tile_ws(ctypes.byref(problem), ctypes.byref(t)))
    return t


def tile_from_reference(problem: C.Problem, r: int, s: int, x: int, y: int, z: int = 32, depth_u: int = 0,
                        lds_buffers: int = 0):
    """The tile record that runs the reference's tuning record (-r -s -x -y -z
    -u -f) on ``problem``: (status, Tile, message).  status is SUCCESS when a
    kernel has that exact record, WARN when it runs as the nearest supported
    one; RTN_ERR raises."""
    t = C.Tile()
    st = C.lib().ffgemm_tile_from_reference(ctypes.byref(problem), r, s, x, y, z, depth_u, lds_buffers,
                                            ctypes.byref(t))
    if st not in (C.SUCCESS, C.WARN):
        C.check(st)
    msg = C.lib().ffgemm_last_error().decode() if st == C.WARN else ""
    return st, t, msg


def gemm(layout, M, N, K, A, B, Cm=None, D=None, *, lda=None, ldb=None, ldc=None, ldd=None,
         alpha=1.0, beta=0.0, batch=1, bsA=0, bsB=0, bsC=0, bsD=0, compute=C.COMPUTE_DEFAULT,
         round=C.ROUND_RNE, dtype_out=None, scaleA=None, scaleB=None, tile=None, stream=None,
         workspace="auto"):
    """Column-major D = alpha*op(A)*op(B) + beta*C on flat device tensors.

    A, B, C, D are 1-D (or any contiguous) device tensors holding the
    column-major matrices; D is allocated when not given.  ``workspace`` as in
    :func:`run`.
    """
    ta, tb = layout[0] == "T", layout[1] == "T"
    lda = lda or (K if ta else M)
    ldb = ldb or (N if tb else K)
    ldd = ldd or M
    ldc = ldc or ldd
    in_code = _DT[A.dtype]
    out_dtype = dtype_out or (torch.float32 if in_code == C.F8E4M3 else A.dtype)
    if D is None:
        n = max(bsD * batch, ldd * N) if batch > 1 else ldd * N
        D = torch.empty(n, dtype=out_dtype, device=A.device)
    p = make_problem(layout, in_code, M, N, K, lda, ldb, ldc, ldd, alpha, beta, batch, bsA, bsB, bsC,
                     bsD, _DT[out_dtype], compute, round, scaleA, scaleB)
    run(p, A, B, Cm, D, tile, stream, workspace)
    return D


def mm(a: torch.Tensor, b: torch.Tensor, trans_a: bool = False, trans_b: bool = False, *,
       out_dtype=None, alpha: float = 1.0, beta: float = 0.0, c: torch.Tensor | None = None,
       compute=C.COMPUTE_DEFAULT, tile=None, out: torch.Tensor | None = None) -> torch.Tensor:
    """Row-major ``alpha * op(a) @ op(b) + beta * c`` with contiguous 2-D tensors."""
    assert a.is_contiguous() and b.is_contiguous(), "mm expects contiguous row-major tensors"
    Mr = a.shape[1] if trans_a else a.shape[0]
    Ka = a.shape[0] if trans_a else a.shape[1]
    Kb = b.shape[1] if trans_b else b.shape[0]
    Nr = b.shape[0] if trans_b else b.shape[1]
    assert Ka == Kb, f"inner dimensions differ: {Ka} vs {Kb}"
    layout = ("T" if trans_b else "N") + ("T" if trans_a else "N")
    out_dtype = out_dtype or (torch.float32 if _DT[a.dtype] == C.F8E4M3 else a.dtype)
    if out is None:
        out = torch.empty((Mr, Nr), dtype=out_dtype, device=a.device)
    if beta != 0.0:
        assert c is not None and c.shape == (Mr, Nr) and c.is_contiguous()
    p = make_problem(layout, _DT[a.dtype], Nr, Mr, Ka, b.shape[1], a.shape[1], Nr, Nr, alpha, beta,
                     dtype_out=_DT[out_dtype], compute=compute)
    run(p, b, a, c, out, tile)
    return out


def convert(src: torch.Tensor, dtype) -> torch.Tensor:
    """Device fp32 -> bf16/fp16 conversion (RNE) through the library."""
    out = torch.empty(src.shape, dtype=dtype, device=src.device)
    C.check(C.lib().ffgemm_convert(src.data_ptr(), out.data_ptr(), src.numel(), _DT[dtype], _stream(None)))
    return out
