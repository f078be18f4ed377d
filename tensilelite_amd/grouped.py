"""Grouped GEMM: many independent column-major problems in one persistent
launch (ffgemm_grouped_plan_*), capturable into a hipGraph."""
from __future__ import annotations

import ctypes
import math

import torch

from . import _capi as C
from .gemm import _DT, _stream, make_problem, make_tile


class GroupedPlan:
    """Descriptors uploaded once; ``run()`` only launches the persistent kernel."""

    def __init__(self, problems, As, Bs, Cs, Ds, tile=None):
        n = len(problems)
        arr = (C.Problem * n)(*problems)
        vp = ctypes.c_void_p
        pa = (vp * n)(*[t.data_ptr() for t in As])
        pb = (vp * n)(*[t.data_ptr() for t in Bs])
        pc = (vp * n)(*[(t.data_ptr() if t is not None else None) for t in Cs]) if Cs is not None else None
        pd = (vp * n)(*[t.data_ptr() for t in Ds])
        t = make_tile(tile)
        self._h = vp()
        C.check(C.lib().ffgemm_grouped_plan_create(ctypes.byref(self._h), n, arr, pa, pb, pc, pd,
                                                   ctypes.byref(t) if t is not None else None))
        self._keep = (As, Bs, Cs, Ds)

    def run(self, stream=None):
        C.check(C.lib().ffgemm_grouped_plan_run(self._h, _stream(stream)))

    def info(self):
        tot = ctypes.c_int64()
        grid = ctypes.c_int32()
        t = C.Tile()
        C.check(C.lib().ffgemm_grouped_plan_info(self._h, ctypes.byref(tot), ctypes.byref(grid), ctypes.byref(t)))
        return {"total_tiles": tot.value, "grid": grid.value, "tile": (t.mt_m, t.mt_n)}

    def close(self):
        if self._h:
            C.lib().ffgemm_grouped_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def config_shapes(seed=3, count=64):
    """The batching configuration: M log-uniform in [64, 4096], N in
    {256, 1024, 4096}, K in {512, 2048}; deterministic from `seed`."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    shapes = []
    for _ in range(count):
        u = torch.rand(1, generator=g).item()
        M = max(1, int(round(math.exp(math.log(64) + u * (math.log(4096) - math.log(64))))))
        N = [256, 1024, 4096][int(torch.randint(0, 3, (1,), generator=g).item())]
        K = [512, 2048][int(torch.randint(0, 2, (1,), generator=g).item())]
        shapes.append((M, N, K))
    return shapes


def make_config_problems(device, seed=3, count=64, dtype=torch.bfloat16):
    """bf16 NN problems of the batching configuration with normal(0, 1/sqrt(K)) data."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    probs, As, Bs, Ds = [], [], [], []
    for (M, N, K) in config_shapes(seed, count):
        lda = (M + 7) // 8 * 8
        A = torch.zeros(lda * K, dtype=dtype)
        A.view(K, lda)[:, :M] = (torch.randn(K, M, generator=g) / math.sqrt(K)).to(dtype)
        B = (torch.randn(N * K, generator=g) / math.sqrt(K)).to(dtype)
        As.append(A.to(device))
        Bs.append(B.to(device))
        Ds.append(torch.zeros(lda * N, dtype=dtype, device=device))
        probs.append(make_problem("NN", _DT[dtype], M, N, K, lda, K, lda, lda))
    return probs, As, Bs, Ds


def bench_config(device, steps=20, tile=None):
    """Times the grouped batching configuration eagerly and as a hipGraph replay."""
    from . import autotune

    probs, As, Bs, Ds = make_config_problems(device)
    flop = sum(2.0 * p.M * p.N * p.K for p in probs)
    byts = sum(2.0 * (p.M * p.K + p.K * p.N + p.M * p.N) for p in probs)
    if tile is None:
        tile = autotune.tune_grouped(probs, As, Bs, Ds)["best"]
    plan = GroupedPlan(probs, As, Bs, None, Ds, tile=tile)
    for _ in range(3):
        plan.run()
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        plan.run()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    with torch.cuda.graph(graph):
        plan.run()
    graph.replay()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn):
        e0.record(s)
        for _ in range(steps):
            fn()
        e1.record(s)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    # eager and graph blocks interleaved (the clock drifts with load), median of each
    eager, graphed = [], []
    for _ in range(3):
        eager.append(timed(plan.run))
        graphed.append(timed(graph.replay))
    ms_eager = sorted(eager)[1]
    ms_graph = sorted(graphed)[1]
    # per-problem roofline: max(MFMA time, HBM time)
    t_roof = sum(max(2.0 * p.M * p.N * p.K / 2500e12, 2.0 * (p.M * p.K + p.K * p.N + p.M * p.N) / 8e12)
                 for p in probs)
    info = plan.info()
    plan.close()
    return {"problems": len(probs), "tile": list(info["tile"]), "total_tiles": info["total_tiles"],
            "grid": info["grid"], "tflops_eager": round(flop / (ms_eager * 1e-3) / 1e12, 1),
            "tflops_graph": round(flop / (ms_graph * 1e-3) / 1e12, 1), "ms_eager": round(ms_eager, 4),
            "ms_graph": round(ms_graph, 4), "roofline_ms": round(t_roof * 1e3, 4),
            "roofline_frac": round(t_roof * 1e3 / ms_graph, 3), "gflop": round(flop / 1e9, 2),
            "mbytes": round(byts / 1e6, 1)}
