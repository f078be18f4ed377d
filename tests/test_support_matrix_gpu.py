"""The reference's function-support matrix on the GPU (test/gemm_test.sh:5-73,
verdicts in test/result data/fuction_test.xlsx, committed as
tests/golden/support_matrix.json): 32 tuning records per data type on
M=768, N=640, K=256, run (i) through the ffgemm_bench harness with the
reference's own flags and (ii) through the reference-kernarg entry
ffgemm_run_tn_kernarg with the record ffgemm_tile_from_reference maps them
to.  Every record must run and verify, including those the reference
refuses (its "N" rows): the supported set is a superset.

Tolerances: the harness's own verification (fp32 rel-Frobenius <= 1e-5,
16-bit per element 1e-2 * (|ref| + rms), against its fp32 CPU loop, which
restates src/gemmMfma.cpp:325-346); the kernarg runs use the InitRealData
integer patterns (ff_gpu_src/ff_gpu.cpp:134-136), whose sums are exact, and
compare with the oracle's restatement of that loop: fp32 exactly up to the
alpha/beta roundings (rel-Frobenius 1e-6), 16-bit per element as above.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys

import numpy as np
import pytest
import torch

import oracle
import tensilelite_amd as tl
from tensilelite_amd import _capi as C
from tests._util import elem_ok, rel_fro

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tensilelite_amd", "bin", "ffgemm_bench")
MATRIX = json.load(open(os.path.join(ROOT, "tests", "golden", "support_matrix.json")))["dtypes"]
RECORDS = [(c["r"], c["s"], c["x"], c["y"], c["u"]) for c in MATRIX["1"]["configs"]]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("dtype", [1, 2, 3])
def test_support_matrix_through_harness(cuda, dtype):
    """ffgemm_bench -d D -m 768 -n 640 -k 256 -r r -s s -x x -y y -u u: every
    record exits 0 with 'Verify Success.' (the harness's exit code is 1 when
    verification fails)."""
    if not os.path.isfile(BIN):
        pytest.skip("ffgemm_bench not built")
    prob = MATRIX["1"]["problem"]
    for r, s, x, y, u in RECORDS:
        args = [BIN, "-d", str(dtype), "-m", str(prob["M"]), "-n", str(prob["N"]), "-k", str(prob["K"]),
                "-r", str(r), "-s", str(s), "-x", str(x), "-y", str(y), "-u", str(u), "-l", "2"]
        _say("case", " ".join(args[1:]))
        out = subprocess.run(args, capture_output=True, text=True, timeout=60)
        assert out.returncode == 0, (out.returncode, out.stdout[-1500:], out.stderr[-500:])
        assert "Verify Success." in out.stdout


@pytest.mark.parametrize("dtype", [1, 2, 3])
def test_support_matrix_through_kernarg_entry(cuda, dtype):
    """The reference's kernarg list (TN, 768 x 640 x 256, ld = dim + 32,
    alpha 0.314, beta 0.2718 as its problem hard-codes) with each record's
    mapped tile."""
    prob = MATRIX["1"]["problem"]
    M, N, K = prob["M"], prob["N"], prob["K"]
    pad = 32
    lda = ldb = K + pad
    ldc = ldd = M + pad
    alpha, beta = 0.314, 0.2718
    A = oracle.init_real_data(2.4, lda, M)          # idx % 7
    B = oracle.init_real_data(-2.4, ldb, N)         # -(idx % 11)
    Cm = oracle.init_real_data(2.4, ldc, N)
    name = {1: "f32", 2: "f16", 3: "bf16"}[dtype]
    tdt = {1: torch.float32, 2: torch.float16, 3: torch.bfloat16}[dtype]
    # the patterns are small integers: exact in every dtype
    dA, dB, dC = (torch.from_numpy(v).to(tdt).to(cuda) for v in (A, B, Cm))
    ref = oracle.gemm_f32("TN", M, N, K, alpha, A, lda, B, ldb, beta, Cm, ldc, ldd)
    ref = ref.reshape(N, ldd)[:, :M]
    p = tl.make_problem("TN", dtype, M, N, K, lda, ldb, ldc, ldd, alpha, beta)
    seen = set()
    for r, s, x, y, u in RECORDS:
        st, t, msg = tl.tile_from_reference(p, r, s, x, y, 32, u, 0)
        assert st in (C.SUCCESS, C.WARN)
        _say(f"case {name} r={r} s={s} x={x} y={y} u={u} -> MT{t.mt_m}x{t.mt_n} waves {t.waves_m}x{t.waves_n}")
        seen.add((t.mt_m, t.mt_n))
        D = torch.full((ldd * N,), float("nan"), dtype=tdt, device=cuda)
        C.check(C.lib().ffgemm_run_tn_kernarg(D.data_ptr(), dC.data_ptr(), dA.data_ptr(), dB.data_ptr(),
                                              alpha, beta, ldd, ldd * N, ldc, ldc * N, lda, lda * M, ldb,
                                              ldb * N, M, N, K, 1, dtype, t, None))
        torch.cuda.synchronize()
        got = D.float().cpu().numpy().reshape(N, ldd)[:, :M]
        if dtype == 1:
            assert rel_fro(got, ref) <= 1e-6
        else:
            bad, worst = elem_ok(got, ref)
            assert bad == 0, worst
        # padding rows of D are never written
        assert torch.isnan(D.view(N, ldd)[:, M:].float()).all()
    # the records' macro tiles (32..128 per side) run on the 64 x 64 and 128 x 128 kernels
    assert seen == {(64, 64), (128, 128)}
