"""ffgemm_run_x3_noalloc: bf16x3 with its hi/lo planes in the library's
static device array instead of a caller workspace.  Through the C ABI, two
calls in a row (the second smaller, so the array holds the first call's
planes when it starts), each bit-identical to ffgemm_run_ws with a caller
workspace on the same problem and tile (same kernels, same planes) and within
the bf16x3 bar, rel-Frobenius 1e-5, of the FP64 oracle.  A problem whose
planes exceed the array is refused with RTN_ERR before anything runs (no GPU
needed for that check)."""
from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest

import oracle
import tensilelite_amd as tl
from tensilelite_amd import _capi as C
from tests._util import colmajor, extract, op_dims, rel_fro

SCRATCH_BYTES = 64 << 20


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_noalloc_refuses_planes_beyond_the_scratch():
    S = 4096   # (4096 + 4096) * 3 * 4096 * 2 bytes = 192 MiB of planes
    p = tl.make_problem("NN", C.F32, S, S, S, S, S, S, S, compute=C.COMPUTE_BF16X3)
    need = tl.workspace_size(p)
    assert need > SCRATCH_BYTES
    fake = ctypes.c_void_p(0x100000)
    st = C.lib().ffgemm_run_x3_noalloc(ctypes.byref(p), fake, fake, None, fake, None, None)
    assert st == C.ERR
    assert "no-allocation scratch" in C.lib().ffgemm_last_error().decode()
    # a non-bf16x3 problem is refused too
    q = tl.make_problem("NN", C.F32, 64, 64, 64, 64, 64, 64, 64)
    assert C.lib().ffgemm_run_x3_noalloc(ctypes.byref(q), fake, fake, None, fake, None, None) == C.ERR


@pytest.mark.gpu
def test_noalloc_matches_workspace_entry_two_calls(cuda):
    import torch

    cases = [("NN", 1000, 1200, 1000, (256, 256)), ("TN", 300, 200, 136, None)]   # second call smaller
    for layout, M, N, K, tile in cases:
        (ra, ca), (rb, cb) = op_dims(layout, M, N, K)
        p = tl.make_problem(layout, C.F32, M, N, K, ra, rb, M, M, 0.75, -0.5, compute=C.COMPUTE_BF16X3)
        need = tl.workspace_size(p)
        _say(f"case noalloc {layout} M={M} N={N} K={K} tile={tile} planes={need} bytes")
        assert 0 < need <= SCRATCH_BYTES
        rng = np.random.default_rng(M)
        A = colmajor(rng, ra, ca, kind="normal")
        B = colmajor(rng, rb, cb, kind="normal")
        Cm = colmajor(rng, M, N, kind="normal")
        dA, dB, dC = (torch.from_numpy(x).to(cuda) for x in (A, B, Cm))
        t = tl.make_tile(tile) if tile else None
        tp = ctypes.byref(t) if t is not None else None
        ws = torch.empty(need, dtype=torch.uint8, device=cuda)
        D_ws = torch.full((M * N,), float("nan"), device=cuda)
        D_na = torch.full((M * N,), float("nan"), device=cuda)
        C.check(C.lib().ffgemm_run_ws(ctypes.byref(p), dA.data_ptr(), dB.data_ptr(), dC.data_ptr(),
                                      D_ws.data_ptr(), tp, ws.data_ptr(), need, None))
        C.check(C.lib().ffgemm_run_x3_noalloc(ctypes.byref(p), dA.data_ptr(), dB.data_ptr(), dC.data_ptr(),
                                              D_na.data_ptr(), tp, None))
        torch.cuda.synchronize()
        assert torch.equal(D_na.view(torch.int32), D_ws.view(torch.int32))
        ref = oracle.gemm_f64(layout, M, N, K, 0.75, A.astype(np.float64), ra, B.astype(np.float64), rb, -0.5,
                              Cm.astype(np.float64), M, M)
        err = rel_fro(extract(D_na.cpu().numpy(), M, N, M), extract(ref, M, N, M))
        _say(f"rel_fro={err:.3e}")
        assert err <= 1e-5
