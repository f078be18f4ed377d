"""The reference's "phantom" shape set: TN GEMMs with one free dimension at
480 or 512 and the other two equal, 512 ... 8192 (test/phantom_test.sh:5-53),
with the fraction of its chip's MFMA peak the reference published for each
(doc/mfma gemm.pptx slides 10-12, gfx908 at 1289 MHz; context only, not a
target on another chip).  ``bench`` times each shape with the library's own
choice (split-K included: it runs through the workspace path) and reports
TFLOP/s and the fraction of MI355X's peak for the dtype.
"""
from __future__ import annotations

import torch

from . import _capi as C
from .gemm import make_problem, query_tile_ws, run, workspace_size

SIZES = (512, 1024, 2048, 4096, 8192)
# (dtype name, family) -> published fractions of peak for SIZES, families
# "480xN" (M = 480, N = K = size), "512xN", "Nx480" (N = 480, M = K = size), "Nx512"
REFERENCE_EFF = {
    ("f32", "480xN"): (31.21, 50.66, 58.29, 57.28, 62.53),
    ("f32", "512xN"): (35.51, 45.05, 58.18, 70.86, 77.10),
    ("f32", "Nx480"): (32.96, 53.36, 54.88, 52.60, 43.54),
    ("f32", "Nx512"): (35.53, 44.95, 58.44, 65.94, 65.66),
    ("bf16", "480xN"): (22.94, 45.57, 59.85, 62.43, 62.48),
    ("bf16", "512xN"): (27.93, 42.14, 55.37, 66.22, 80.80),
    ("bf16", "Nx480"): (25.68, 51.23, 61.71, 59.73, 57.40),
    ("bf16", "Nx512"): (27.48, 42.07, 55.63, 66.33, 78.74),
    ("f16", "480xN"): (13.38, 26.11, 32.86, 35.84, 36.65),
    ("f16", "512xN"): (17.03, 33.25, 44.32, 48.61, 64.81),
    ("f16", "Nx480"): (15.03, 28.27, 33.24, 35.12, 33.85),
    ("f16", "Nx512"): (17.21, 33.28, 44.19, 48.08, 64.15),
}
DTYPES = {"f32": (C.F32, torch.float32, 157.3), "bf16": (C.BF16, torch.bfloat16, 2500.0),
          "f16": (C.F16, torch.float16, 2500.0)}


def shape(family: str, s: int):
    """(M, N, K) of a family member."""
    return {"480xN": (480, s, s), "512xN": (512, s, s), "Nx480": (s, 480, s), "Nx512": (s, 512, s)}[family]


def problems(dtypes=("f32", "bf16", "f16")):
    """[(dtype name, family, M, N, K, published %)] for every member."""
    out = []
    for dt in dtypes:
        for fam in ("480xN", "512xN", "Nx480", "Nx512"):
            for s, eff in zip(SIZES, REFERENCE_EFF[(dt, fam)]):
                out.append((dt, fam) + shape(fam, s) + (eff,))
    return out


def bench(device, dtypes=("f32", "bf16", "f16"), min_ms=20.0):
    """Times every shape (TN, alpha 1, beta 0, normal(0, 1/sqrt(K)) inputs) in
    blocks of at least min_ms; returns rows of TFLOP/s, % of MI355X's peak for
    the dtype, the plan (tile, split) and the reference's published %."""
    rows = []
    g = torch.Generator(device="cpu").manual_seed(0)
    for dt, fam, M, N, K, ref in problems(dtypes):
        code, tdt, peak = DTYPES[dt]
        A = (torch.randn(M * K, generator=g) / K ** 0.5).to(tdt).to(device)
        B = (torch.randn(N * K, generator=g) / K ** 0.5).to(tdt).to(device)
        D = torch.empty(M * N, dtype=tdt, device=device)
        p = make_problem("TN", code, M, N, K, K, K, M, M)
        n = workspace_size(p)
        ws = torch.empty(n, dtype=torch.uint8, device=device) if n else None
        fn = lambda: run(p, A, B, None, D, workspace=ws)
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        s = torch.cuda.current_stream()
        steps = 10
        while True:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for _ in range(steps):
                fn()
            e1.record(s)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1)
            if ms >= min_ms or steps >= 100000:
                break
            steps = int(steps * max(2.0, 1.2 * min_ms / max(ms, 1e-3)))
        ms /= steps
        tf = 2.0 * M * N * K / (ms * 1e-3) / 1e12
        t = query_tile_ws(p)
        rows.append({"dtype": dt, "M": M, "N": N, "K": K, "us": round(ms * 1e3, 2), "tflops": round(tf, 2),
                     "pct_peak": round(100 * tf / peak, 2), "tile": [t.mt_m, t.mt_n], "split_k": max(1, t.split_k),
                     "reference_pct_peak_gfx908": ref})
        del A, B, D, ws
    return rows
