"""The Tensile-named kernels (Cijk_Alik_Bljk_{SB,HBH,BBH}_MT..._SE_K1) loaded
from their code object with hipModuleLoad and launched with the reference's
148-byte Tensile argument buffer (src/gemmMfma.cpp:70-103), as a Tensile
client runs a replacement kernel; results against the FP64 oracle.

Tolerances: 16-bit per element |d - r| <= 1e-2 (|r| + rms(r)); fp32
rel-Frobenius <= 1e-5.  The CPU test checks the code object's metadata: each
kernel's name and a 148-byte argument segment with the reference's offsets.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys

import numpy as np
import pytest

import oracle
from tensilelite_amd import tensile
from tests._util import bits_to_f32, colmajor, elem_ok, quantize_bits, rel_fro


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_code_object_metadata():
    if not os.path.isfile(tensile.CO_PATH):
        pytest.skip("code object not built")
    readelf = shutil.which("llvm-readelf") or "/opt/rocm/llvm/bin/llvm-readelf"
    if not os.path.isfile(readelf):
        pytest.skip("llvm-readelf absent")
    notes = subprocess.run([readelf, "--notes", tensile.CO_PATH], capture_output=True, text=True).stdout
    names = re.findall(r"^\s+\.name:\s+(Cijk_\S+)", notes, re.M)
    assert sorted(names) == sorted(tensile.KERNELS)
    sizes = re.findall(r"\.kernarg_segment_size:\s+(\d+)", notes)
    assert sizes and all(int(s) == tensile.KERNARG_BYTES for s in sizes)
    # the 30 arguments at the reference's offsets: sizeC/A/B at 0/8/16, D C A B at 24..48, alpha 56, beta 60,
    # then 21 32-bit values 64..144
    want = [0, 8, 16, 24, 32, 40, 48, 56, 60] + list(range(64, 148, 4))
    first = notes.split(".name:")[0]
    offs = [int(x) for x in re.findall(r"\.offset:\s+(\d+)", first)]
    assert offs[:30] == want


def test_pack_kernarg_layout():
    buf = tensile.pack_kernarg("HBH", 0x1000, 0x2000, 0x3000, 0x4000, 1.5, -0.25, 300, 200, 64, 72, 72, 304, 304,
                               batch=2, bsA=72 * 300, bsB=72 * 200, bsC=304 * 200, bsD=304 * 200, sizes=(1, 2, 3))
    assert len(buf) == 148
    import struct
    assert struct.unpack_from("<QQQ", buf, 0) == (1, 2, 3)
    assert struct.unpack_from("<QQQQ", buf, 24) == (0x1000, 0x2000, 0x3000, 0x4000)
    assert struct.unpack_from("<e", buf, 56)[0] == 1.5 and struct.unpack_from("<e", buf, 60)[0] == -0.25
    assert struct.unpack_from("<8I", buf, 64) == (304, 304 * 200, 304, 304 * 200, 72, 72 * 300, 72, 72 * 200)
    assert struct.unpack_from("<4I", buf, 96) == (300, 200, 2, 64)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(tensile.KERNELS))
def test_tensile_kernel_through_module_load(cuda, name):
    import torch

    mt0, mt1, block = tensile.KERNELS[name]
    kind = name.split("_")[3]
    M, N, K, batch = 2 * mt0 + 72, mt1 + 40, 200, 2        # ragged edges, partial K-tile, two batches
    lda, ldb, ldc = K + 8, K + 16, M + 8
    bsA, bsB, bsC = lda * M + 64, ldb * N + 32, ldc * N + 16
    alpha, beta = 0.75, -0.5
    _say(f"case {name} M={M} N={N} K={K} batch={batch}")
    rng = np.random.default_rng(31)
    A = rng.standard_normal(bsA * batch).astype(np.float32) / np.float32(np.sqrt(K))
    B = rng.standard_normal(bsB * batch).astype(np.float32) / np.float32(np.sqrt(K))
    Cm = rng.uniform(-1, 1, bsC * batch).astype(np.float32)
    if kind == "SB":
        Ad, Bd, Cd = A, B, Cm
        tdt = torch.float32
        dA, dB, dC = (torch.from_numpy(x).to(cuda) for x in (A, B, Cm))
    else:
        nm = "bf16" if kind == "BBH" else "f16"
        tdt = torch.bfloat16 if kind == "BBH" else torch.float16
        bits = [quantize_bits(x, nm) for x in (A, B, Cm)]
        Ad, Bd, Cd = (bits_to_f32(b, nm) for b in bits)
        dA, dB, dC = (torch.from_numpy(b.view(np.int16).copy()).to(cuda).view(tdt) for b in bits)
    D = torch.full((bsC * batch,), float("nan"), dtype=tdt, device=cuda)
    buf = tensile.pack_kernarg(kind, D.data_ptr(), dC.data_ptr(), dA.data_ptr(), dB.data_ptr(), alpha, beta,
                               M, N, K, lda, ldb, ldc, ldc, batch, bsA, bsB, bsC, bsC,
                               sizes=(bsC * batch, bsA * batch, bsB * batch))
    mod = tensile.TensileModule()
    try:
        grid = ((M + mt0 - 1) // mt0, (N + mt1 - 1) // mt1, batch)
        mod.launch(name, buf, grid, block, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        got = D.float().cpu().numpy()
        for b in range(batch):
            ref = oracle.gemm_f64("TN", M, N, K, alpha, Ad[b * bsA:].astype(np.float64), lda,
                                  Bd[b * bsB:].astype(np.float64), ldb, beta, Cd[b * bsC:].astype(np.float64),
                                  ldc, ldc)
            g = got[b * bsC: b * bsC + ldc * N].reshape(N, ldc)
            r = ref.reshape(N, ldc)
            if kind == "SB":
                assert rel_fro(g[:, :M], r[:, :M]) <= 1e-5
            else:
                bad, worst = elem_ok(g[:, :M], r[:, :M])
                assert bad == 0, worst
            assert np.isnan(g[:, M:]).all()            # padding rows untouched
        # a grid wider than the tiles: the extra workgroups return without writing
        D2 = torch.full_like(D, float("nan"))
        buf2 = tensile.pack_kernarg(kind, D2.data_ptr(), dC.data_ptr(), dA.data_ptr(), dB.data_ptr(), alpha, beta,
                                    M, N, K, lda, ldb, ldc, ldc, batch, bsA, bsB, bsC, bsC)
        mod.launch(name, buf2, (grid[0] + 1, grid[1] + 2, batch + 1), block, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert torch.equal(D2.view(torch.int16 if kind != "SB" else torch.int32),
                           D.view(torch.int16 if kind != "SB" else torch.int32))
    finally:
        mod.close()
