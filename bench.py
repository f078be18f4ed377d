#!/usr/bin/env python3
"""Benchmark of the gemmMfma path on one MI355X.

Headline workload (one "step" = one GEMM launch): bf16 GEMM with fp32
accumulate in the reference's layout (TN: A and B K-contiguous, column-major
D), M = N = K = 4096, alpha = 1, beta = 0, A and B i.i.d. normal(0, 1/sqrt(K))
from seed 1, inputs resident in HBM.  NN and NT run under "extras".  The JSON
line reports TFLOP/s (2*M*N*K per launch) and the fraction of the bf16 MFMA
peak, the roofline of the dominant kernel measured with HIP events on the
launch stream, the CPU reference GEMM timed on the host cores beside it, and
the other forms of the metric (fp16, NT/TN, beta = 1, fp8, bf16x3 vs fp32,
grouped) under "extras".

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extras] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

PEAK_TFLOPS = {"bf16": 2500.0, "f16": 2500.0, "f8": 5000.0, "f32": 157.3, "f64": 78.6}
HBM_GBPS = 8000.0
MAX_SCLK_MHZ = 2400.0   # the clock the published peaks assume


def _args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", default="", help="run only these extras, comma-separated (profiling)")
    ap.add_argument("--preload-s", type=float, default=1.5,
                    help="untimed seconds of the headline kernel before its warm-up (clock settling)")
    return ap.parse_args()


def time_launches(fn, steps, warmup, torch, min_ms=0.0):
    """Average duration of fn() over back-to-back launches, HIP events on the
    stream the kernels are launched on (torch's current stream).  `steps`
    launches, or, when min_ms is set, as many as fill at least min_ms (a block
    of `steps` first measures one launch): a short block falls inside the
    power manager's clock ramp and reads low."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()

    def block(n):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(n):
            fn()
        e1.record(s)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    ms = block(steps)
    if min_ms and ms < min_ms:
        n = max(steps, int(math.ceil(steps * min_ms / max(ms, 1e-3))))
        return block(n) / n
    return ms / steps  # ms


def preload(fn, seconds, torch):
    """Untimed launches of fn() for `seconds` of wall time: the clock settles
    at the power limit's steady state for this load before the timed region
    (the driver's few warm-up launches fall inside the ramp).  Returns the
    launch count."""
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(50):
            fn()
        n += 50
        torch.cuda.synchronize()
    return n


class ClockSampler:
    """Samples the GPU's graphics clock (amdsmi GPU metrics) from a background
    thread; mean() is the average of the samples inside a window.  Every
    method is a no-op when amdsmi is unavailable."""

    def __init__(self, period_s=0.002):
        import threading
        self.samples = []
        self.ok = False
        self._stop = threading.Event()
        try:
            import amdsmi
            amdsmi.amdsmi_init()
            self._smi = amdsmi
            self._h = amdsmi.amdsmi_get_processor_handles()[0]
            self._read()
            self.ok = True
        except Exception:
            return
        self._period = period_s
        self._t = threading.Thread(target=self._run, daemon=True)
        self._t.start()

    def _read(self):
        m = self._smi.amdsmi_get_gpu_metrics_info(self._h)
        v = m.get("current_gfxclks")
        if isinstance(v, (list, tuple)):
            v = [x for x in v if isinstance(x, (int, float)) and 0 < x < 10000]
            return sum(v) / len(v) if v else None
        v = m.get("current_gfxclk", m.get("average_gfxclk_frequency"))
        return float(v) if isinstance(v, (int, float)) and 0 < v < 10000 else None

    def _run(self):
        while not self._stop.is_set():
            try:
                v = self._read()
                if v:
                    self.samples.append((time.perf_counter(), v))
            except Exception:
                pass
            time.sleep(self._period)

    def mean(self, t0, t1):
        v = [c for t, c in self.samples if t0 <= t <= t1]
        return round(sum(v) / len(v), 1) if v else None

    def stop(self):
        if self.ok:
            self._stop.set()
            self._t.join(timeout=1)
            try:
                self._smi.amdsmi_shut_down()
            except Exception:
                pass


def main():
    a = _args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the GEMM path runs on a "
              "single GPU; across GPUs it would shard N, each GPU computing a block of D's columns)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import tensilelite_amd as tl
    from tensilelite_amd import _capi as C

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if a.only == "copy":
        # calibration for HBM byte counters: a plain device copy of 256 MiB
        x = torch.ones(64 << 20, device=dev)
        y = torch.empty_like(x)
        ms = time_launches(lambda: y.copy_(x), a.steps, a.warmup, torch)
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": y.numel() * 4,
                          "ms": ms, "GBps": 2 * x.numel() * 4 / (ms * 1e-3) / 1e9}))
        return
    M = N = K = 4096
    flop = 2.0 * M * N * K
    g = torch.Generator(device="cpu").manual_seed(1)
    A = (torch.randn(M * K, generator=g) / math.sqrt(K)).to(torch.bfloat16).to(dev)
    B = (torch.randn(K * N, generator=g) / math.sqrt(K)).to(torch.bfloat16).to(dev)
    D = torch.empty(M * N, dtype=torch.bfloat16, device=dev)
    prob = tl.make_problem("TN", C.BF16, M, N, K, K, K, M, M)
    tile = tl.query_tile(prob)
    fn = lambda: tl.run(prob, A, B, None, D)

    # clock settling: untimed launches until the power manager's clock is at
    # its steady state for this load, then the W warm-up launches and the K
    # timed ones (barrier + sync on both sides; one rank)
    clk = ClockSampler()
    tp = time.perf_counter()
    n_pre = preload(fn, a.preload_s, torch)
    tp1 = time.perf_counter()
    for _ in range(a.warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(a.steps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    ms = e0.elapsed_time(e1) / a.steps
    tflops = flop / (ms * 1e-3) / 1e12
    time.sleep(0.01)
    # the clock over the last half second of the preload and the timed region
    sclk = clk.mean(max(tp, tp1 - 0.5), time.perf_counter())
    clk.stop()

    out = {
        "metric": "TFLOP/s & % MFMA peak, fp16/bf16/fp8 4096^3; bf16x3 vs fp32 at equal error",
        "value": round(tflops, 2),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic: A,B ~ normal(0,1/sqrt(K)) seed 1 (torch CPU generator), resident in HBM",
        "config": {"workload": "gemm bf16 TN M=N=K=4096 alpha=1 beta=0, fp32 accumulate, bf16 out",
                   "M": M, "N": N, "K": K, "layout": "TN", "tile": [tile.mt_m, tile.mt_n, tile.depth_u],
                   "parallelism": "single GPU"},
        "pct_of_peak": round(100.0 * tflops / PEAK_TFLOPS["bf16"], 2),
        "roofline": {"bound": "mfma", "achieved": round(tflops, 2), "peak": PEAK_TFLOPS["bf16"],
                     "unit": "TFLOP/s", "frac": round(tflops / PEAK_TFLOPS["bf16"], 4),
                     "traffic": profiled_traffic(), "traffic_source": TRAFFIC_PROFILE,
                     "kernel": "ffgemm::v4::gemm_kernel<bf16,TN,%dx%d>" % (tile.mt_m, tile.mt_n),
                     "algorithmic_flop_per_launch": flop, "wall_s": round(wall, 4),
                     "sclk_mhz": sclk,
                     "peak_at_measured_clock": round(PEAK_TFLOPS["bf16"] * sclk / MAX_SCLK_MHZ, 1) if sclk else None,
                     "frac_at_measured_clock": round(tflops / (PEAK_TFLOPS["bf16"] * sclk / MAX_SCLK_MHZ), 4)
                     if sclk else None},
        "preload": {"seconds": a.preload_s, "launches": n_pre,
                    "note": "untimed launches of the headline kernel before the warm-up, so the timed "
                            "launches run at the steady-state clock of this load"},
    }

    extras = {}
    if not a.no_extras:
        extras = run_extras(a, torch, np, tl, C, dev)
    out["extras"] = extras

    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(np, ms)
    print(json.dumps(out))


TRAFFIC_PROFILE = "profiles/gemm_bf16_tn_4096_pmc.md"


def profiled_traffic():
    """HBM-side bytes per launch of the headline kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (each scaled by a plain copy's
    reading of the same counter); None when the summary is absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), TRAFFIC_PROFILE)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def run_extras(a, torch, np, tl, C, dev):
    ex = {}
    steps = max(10, a.steps // 4)
    M = N = K = 4096
    flop = 2.0 * M * N * K
    g = torch.Generator(device="cpu").manual_seed(1)
    only = set(x for x in a.only.split(",") if x)

    def want(name):
        if only and name not in only:
            return False
        # progress on stderr (stdout carries only the JSON line)
        print(f"[bench] {name}", file=sys.stderr, flush=True)
        return True

    # fp16 / bf16 across layouts and beta.  The forms are timed in three
    # interleaved rounds and each reports its median block, so no form pays
    # alone for the clock state the one before it left (the power limit sets
    # the clock from recent load).
    forms, bufs = [], {}
    for dt, tdt in (("bf16", torch.bfloat16), ("f16", torch.float16)):
        A = (torch.randn(M * K, generator=g) / math.sqrt(K)).to(tdt).to(dev)
        B = (torch.randn(K * N, generator=g) / math.sqrt(K)).to(tdt).to(dev)
        C0 = (torch.rand(M * N, generator=g) * 2 - 1).to(tdt).to(dev)
        D = torch.empty(M * N, dtype=tdt, device=dev)
        bufs[dt] = (A, B, C0, D)
        for layout in ("NN", "NT", "TN", "TT"):
            for beta in (0.0, 1.0):
                name = f"{dt}_{layout}_beta{int(beta)}_4096"
                if want(name):
                    forms.append((name, dt, tl.make_problem(layout, tdt, M, N, K, M if layout[0] == "N" else K,
                                                            K if layout[1] == "N" else N, M, M, 1.0, beta)))
    times = {name: [] for name, _, _ in forms}
    for _ in range(3):
        for name, dt, p in forms:
            A, B, C0, D = bufs[dt]
            times[name].append(time_launches(lambda: tl.run(p, A, B, C0, D), max(10, steps // 3), 3, torch, min_ms=60))
    for name, _, _ in forms:
        ms = sorted(times[name])[1]
        tf = flop / (ms * 1e-3) / 1e12
        ex[name] = {"tflops": round(tf, 1), "ms": round(ms, 4), "pct_peak": round(100 * tf / 2500.0, 1)}
    del bufs

    # bf16 at 8192^3 per layout: the library's own plan (ffgemm_run_ws; NN, NT
    # and TT repack their M/N-contiguous operands K-contiguous and run as TN),
    # the native engine of the layout (repack never, no workspace) and torch's
    # matmul (hipBLASLt) on the same column-major buffers, for context; three
    # interleaved rounds, best of each
    if want("layouts_8192"):
        S8 = 8192
        fl8 = 2.0 * S8 ** 3
        X8 = (torch.randn(S8 * S8, generator=g) / math.sqrt(S8)).to(torch.bfloat16).to(dev)
        Y8 = (torch.randn(S8 * S8, generator=g) / math.sqrt(S8)).to(torch.bfloat16).to(dev)
        D8 = torch.empty(S8 * S8, dtype=torch.bfloat16, device=dev)
        runs8, plans8, wss = {}, {}, []
        for layout in ("NN", "NT", "TT", "TN"):
            p8 = tl.make_problem(layout, C.BF16, S8, S8, S8, S8, S8, S8, S8)
            nws = tl.workspace_size(p8)
            ws8 = torch.empty(nws, dtype=torch.uint8, device=dev) if nws else None
            wss.append(ws8)
            plans8[layout] = "repack" if tl.query_tile_ws(p8).repack == 2 else "native"
            # D^T (N x M row-major) = op(B)^T op(A)^T over the same buffers
            at = X8.view(S8, S8) if layout[0] == "N" else X8.view(S8, S8).t()
            bt = Y8.view(S8, S8) if layout[1] == "N" else Y8.view(S8, S8).t()
            runs8[layout] = {
                "library": (lambda p8=p8, ws8=ws8: tl.run(p8, X8, Y8, None, D8, workspace=ws8)),
                "native": (lambda p8=p8: tl.run(p8, X8, Y8, None, D8, tile=(0, 0, 0, 0, 0, 0, 0, 1), workspace=None)),
                "hipblaslt": (lambda at=at, bt=bt: torch.matmul(bt, at, out=D8.view(S8, S8))),
            }
        best8 = {(l, k): float("inf") for l in runs8 for k in runs8[l]}
        for _ in range(3):
            for l in runs8:
                for k, fn in runs8[l].items():
                    best8[(l, k)] = min(best8[(l, k)], time_launches(fn, 5, 2, torch, min_ms=100))
        ex["bf16_layouts_8192"] = {
            l: {"plan": plans8[l], **{f"{k}_tflops": round(fl8 / (best8[(l, k)] * 1e-3) / 1e12, 1) for k in runs8[l]},
                "library_ms": round(best8[(l, "library")], 4)} for l in runs8}
        del X8, Y8, D8, wss, runs8

    # torch (hipBLASLt) bf16 at the same size, for context only: the headline
    # problem itself (A and B K-contiguous: D^T = B A^T in torch's row-major
    # terms) and a plain row-major torch.mm, in interleaved rounds with this
    # library's kernel (best round of each)
    if want("torch_bf16_4096"):
        a_km = torch.randn(M, K, device=dev).to(torch.bfloat16)     # row m = column m of op(A)
        b_kn = torch.randn(N, K, device=dev).to(torch.bfloat16)     # row n = column n of op(B)
        d_nm = torch.empty(N, M, dtype=torch.bfloat16, device=dev)
        x = torch.randn(M, K, dtype=torch.bfloat16, device=dev)
        y = torch.randn(K, N, dtype=torch.bfloat16, device=dev)
        p = tl.make_problem("TN", C.BF16, M, N, K, K, K, M, M)
        runs = {"ffgemm_TN": lambda: tl.run(p, a_km, b_kn, None, d_nm, workspace=None),
                "hipblaslt_TN_same_problem": lambda: torch.matmul(b_kn, a_km.t(), out=d_nm),
                "hipblaslt_row_major_mm": lambda: torch.mm(x, y)}
        best = {k: float("inf") for k in runs}
        for _ in range(3):
            for k, fn in runs.items():
                best[k] = min(best[k], time_launches(fn, steps, 5, torch, min_ms=200))
        ex["torch_bf16_4096_hipblaslt"] = {k: {"tflops": round(flop / (ms * 1e-3) / 1e12, 1), "ms": round(ms, 4)}
                                           for k, ms in best.items()}
        del a_km, b_kn, d_nm, x, y

    # fp8 e4m3 with per-tensor scales, 4096^3 and 8192^3
    for S in (4096, 8192):
        name = f"f8_TN_{S}"
        if not want(name):
            continue
        x = torch.randn(S * S, device=dev)
        y = torch.randn(S * S, device=dev)
        qa, sa = tl.fp8.quantize(x, rows=S, k=S, ldx=S, trans_x=False)
        qb, sb = tl.fp8.quantize(y, rows=S, k=S, ldx=S, trans_x=True)
        D = torch.empty(S * S, dtype=torch.float32, device=dev)
        p = tl.make_problem("TN", C.F8E4M3, S, S, S, S, S, S, S, scaleA=sa, scaleB=sb, dtype_out=C.F32)
        ms = time_launches(lambda: tl.run(p, qa, qb, None, D), max(5, steps // 4), 3, torch, min_ms=300)
        tf = 2.0 * S ** 3 / (ms * 1e-3) / 1e12
        # error of the whole form (quantisation included) against the FP64 GPU
        # product of the unquantised fp32 inputs, on 256 output columns: D[m, n]
        # = sum_k x[m + k*S] * y[k + n*S], an NN product of x and y
        cols = 256
        Dref = torch.empty(S * cols, dtype=torch.float64, device=dev)
        p64 = tl.make_problem("NN", C.F64, S, cols, S, S, S, S, S)
        tl.run(p64, x.double(), y[: S * cols].double(), None, Dref)
        torch.cuda.synchronize()
        err = (torch.linalg.norm(D[: S * cols].double() - Dref) / torch.linalg.norm(Dref)).item()
        ex[name] = {"tflops": round(tf, 1), "ms": round(ms, 4), "pct_peak": round(100 * tf / 5000.0, 1),
                    "rel_fro_err_vs_fp64": err}
        del x, y, qa, qb, D, Dref

    # precision: fp32 MFMA vs bf16x3 at 8192^3 (A, B ~ normal(0,1) seed 2)
    S = 8192
    if want("precision_8192"):
        g2 = torch.Generator(device="cpu").manual_seed(2)
        A = torch.randn(S * S, generator=g2).to(dev)
        B = torch.randn(S * S, generator=g2).to(dev)
        D = torch.empty(S * S, device=dev)
        res = {}
        for comp, name in ((C.COMPUTE_DEFAULT, "f32_mfma"), (C.COMPUTE_BF16X3, "bf16x3"),
                           (C.COMPUTE_BF16X3, "bf16x3_regsplit")):
            p = tl.make_problem("NN", C.F32, S, S, S, S, S, S, S, compute=comp)
            # the bf16x3 planes go to a workspace allocated once, as a caller would keep it;
            # the register-split form (no workspace) splits inside the GEMM
            nws = 0 if name == "bf16x3_regsplit" else tl.workspace_size(p)
            ws = torch.empty(nws, dtype=torch.uint8, device=dev) if nws else None
            # a few untimed launches first: each form runs at its own clock under the power limit
            ms = time_launches(lambda: tl.run(p, A, B, None, D, workspace=ws), 5, 3, torch, min_ms=300)
            del ws
            tf = 2.0 * S ** 3 / (ms * 1e-3) / 1e12
            # error vs the FP64 GPU oracle product on 256 sampled columns
            cols = 256
            Dref = torch.empty(S * cols, dtype=torch.float64, device=dev)
            p64 = tl.make_problem("NN", C.F64, S, cols, S, S, S, S, S)
            tl.run(p64, A.double(), B[: S * cols].double(), None, Dref)
            torch.cuda.synchronize()
            err = (torch.linalg.norm(D[: S * cols].double() - Dref) / torch.linalg.norm(Dref)).item()
            res[name] = {"tflops": round(tf, 1), "ms": round(ms, 3), "rel_fro_err_vs_fp64": err}
            if name == "bf16x3_regsplit":
                t = tl.query_tile(p)
                res[name]["tile"] = [t.mt_m, t.mt_n]
        res["bf16x3_over_f32"] = round(res["bf16x3"]["tflops"] / res["f32_mfma"]["tflops"], 2)
        ex["precision_8192"] = res
        del A, B, D

    # differentiation: bf16 C = A @ B, M=8192 N=4096 K=4096, forward NN and
    # backward TN (dA = dC B^T) + NT (dB = A^T dC) through the autograd Function
    if want("autograd_bf16"):
        from tensilelite_amd.autograd import matmul
        Mg, Ng, Kg = 8192, 4096, 4096
        ga = torch.Generator(device="cpu").manual_seed(4)
        a_ = torch.randn(Mg, Kg, generator=ga).to(torch.bfloat16).to(dev).requires_grad_()
        b_ = torch.randn(Kg, Ng, generator=ga).to(torch.bfloat16).to(dev).requires_grad_()
        dc = torch.randn(Mg, Ng, generator=ga).to(torch.bfloat16).to(dev)

        def step():
            c_ = matmul(a_, b_)
            a_.grad = None
            b_.grad = None
            c_.backward(dc)

        ms_fb = time_launches(step, max(5, steps // 4), 3, torch, min_ms=300)
        ms_f = time_launches(lambda: matmul(a_.detach(), b_.detach()), max(5, steps // 4), 3, torch, min_ms=300)
        fl = 2.0 * Mg * Ng * Kg
        ex["autograd_bf16_8192x4096x4096"] = {"fwd_tflops": round(fl / (ms_f * 1e-3) / 1e12, 1),
                                              "fwd_bwd_tflops": round(3 * fl / (ms_fb * 1e-3) / 1e12, 1),
                                              "fwd_ms": round(ms_f, 4), "fwd_bwd_ms": round(ms_fb, 4)}
        del a_, b_, dc

    # small shapes, where launch and latency dominate: the reference's analysed
    # kernel shape (fp16 TN 480x512x512) and its CPU case on the GPU (fp32 NN
    # 480x512x512).  Timed eagerly (each launch goes through Python and ctypes)
    # and as a hipGraph of 100 launches replayed back to back.
    if want("small_480"):
        res = {}
        m, n, k = 480, 512, 512
        for name, layout, tdt in (("f16_TN", "TN", torch.float16), ("f32_NN", "NN", torch.float32)):
            A = (torch.rand(m * k, generator=g) * 2 - 1).to(tdt).to(dev)
            B = (torch.rand(k * n, generator=g) * 2 - 1).to(tdt).to(dev)
            D = torch.empty(m * n, dtype=tdt, device=dev)
            p = tl.make_problem(layout, tdt, m, n, k, k if layout[0] == "T" else m, k, m, m)
            fn = lambda: tl.run(p, A, B, None, D)
            ms_eager = time_launches(fn, 200, 20, torch)
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                fn()
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                for _ in range(100):
                    fn()
            ms_graph = time_launches(graph.replay, 10, 2, torch) / 100
            fl = 2.0 * m * n * k
            roof_us = max(fl / (2500e12 if tdt != torch.float32 else 157.3e12),
                          (m * k + k * n + m * n) * A.element_size() / 8e12) * 1e6
            res[name] = {"us_eager": round(ms_eager * 1e3, 2), "us_graph": round(ms_graph * 1e3, 2),
                         "tflops_graph": round(fl / (ms_graph * 1e-3) / 1e12, 2), "roofline_us": round(roof_us, 3)}
            del A, B, D, graph
        ex["small_480x512x512"] = res

    # the reference's thin "phantom" shapes (one free dimension 480 / 512):
    # fp32 / bf16 / fp16 TN with the library's own plan (split-K through the
    # workspace path included), beside the fraction of peak the reference
    # published on its own chip; rows [dtype, M, N, K, TFLOP/s, % of MI355X
    # peak, tile, split_k, reference % of gfx908 peak]
    if want("phantom"):
        from tensilelite_amd import phantom
        rows = phantom.bench(dev)
        ex["phantom"] = {"columns": ["dtype", "M", "N", "K", "tflops", "pct_peak", "tile", "split_k",
                                     "reference_pct_peak_gfx908"],
                         "rows": [[r["dtype"], r["M"], r["N"], r["K"], r["tflops"], r["pct_peak"],
                                   "%dx%d" % tuple(r["tile"]), r["split_k"], r["reference_pct_peak_gfx908"]]
                                  for r in rows]}

    # grouped: 64 problems, one persistent launch, hipGraph replay
    if want("grouped_64"):
        from tensilelite_amd import grouped
        # blocks of at least 100 launches (about 40 ms each), so no block sits in the clock ramp
        ex["grouped_64"] = grouped.bench_config(dev, steps=max(100, steps // 4))
    return ex


def cpu_baseline(np, gpu_ms):
    """The oracle's restatement of the reference CPU GEMM (src/gemmMfma.cpp:
    325-346: fp32, k innermost, compiled -O2 without fast-math), timed on this
    host's cores:
      * value: a bounded sample of the headline workload on all cores, 1024 of
        the 4096 output columns of the 4096^3 TN GEMM (about 15 core-seconds);
      * the full 4096^3 GEMM on one core, extrapolated from 128 of its output
        columns (each column is an independent M x K x 1 product with the same
        loop, so time scales with the column count);
      * the reference's own CPU case, fp32 NN 480x512x512 (U[-1, 1], seed 0),
        on one core (median of 3) and on all cores."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle

    cores = len(os.sched_getaffinity(0))
    M = K = 4096
    ncols = 1024
    rng = np.random.default_rng(1)
    A = (rng.standard_normal(M * K) / 64.0).astype(np.float32)
    B = (rng.standard_normal(K * ncols) / 64.0).astype(np.float32)
    t0 = time.perf_counter()
    oracle.gemm_f32("TN", M, ncols, K, 1.0, A, K, B, K, nthreads=cores)
    dt = time.perf_counter() - t0
    tf = 2.0 * M * ncols * K / dt / 1e12
    # 4096^3 on one core: 128 columns, scaled to 4096
    n1 = 128
    t0 = time.perf_counter()
    oracle.gemm_f32("TN", M, n1, K, 1.0, A, K, B[: K * n1], K)
    dt1 = time.perf_counter() - t0
    full_1core_s = dt1 * (4096 / n1)
    # reference CPU case: fp32 NN 480x512x512, U[-1,1] seed 0
    m, n, k = 480, 512, 512
    r0 = np.random.default_rng(0)
    a = r0.uniform(-1, 1, m * k).astype(np.float32)
    b = r0.uniform(-1, 1, k * n).astype(np.float32)

    def med3(threads):
        ts = []
        for _ in range(3):
            t1 = time.perf_counter()
            oracle.gemm_f32("NN", m, n, k, 1.0, a, m, b, k, nthreads=threads)
            ts.append(time.perf_counter() - t1)
        return sorted(ts)[1]

    c1 = med3(1)
    call = med3(cores)
    fl1 = 2.0 * m * n * k
    return {"value": round(tf, 6), "unit": "TFLOP/s", "cores": cores, "kind": "port",
            "sample": f"4096x{ncols}x4096 fp32 TN ({ncols} of 4096 columns of the headline GEMM), "
                      f"{dt:.2f} s on {cores} threads",
            "gpu_speedup": round((2.0 * 4096 ** 3 / (gpu_ms * 1e-3) / 1e12) / tf, 1),
            "gemm_4096_1core": {"sample": f"{n1} of 4096 columns, {dt1:.2f} s", "extrapolated_s": round(full_1core_s, 1),
                                "gflops": round(2.0 * M * n1 * K / dt1 / 1e9, 3)},
            "config1_480x512x512": {"1core_s": round(c1, 4), "1core_gflops": round(fl1 / c1 / 1e9, 3),
                                    "allcores_s": round(call, 4), "allcores_gflops": round(fl1 / call / 1e9, 3),
                                    "allcores": cores},
            "compiler": "gcc -O2 -ffp-contract=off -fno-fast-math (oracle/Makefile)"}


if __name__ == "__main__":
    main()
