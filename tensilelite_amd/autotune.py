"""The project's autotuner: successor of the reference's perf_test.py sweep
(test/perf_test.py:58-167) for the gfx950 knobs.

The reference sweeps padding, LDS buffers, MFMA size, MFMA tiles per wave
(r, s), waves per group (x, y) and DepthU, keeps the best TFLOP/s and prints
``Performance = X TFLOPs``.  Here the knobs that exist in the gfx950 kernels
are the macro tile (MT_M x MT_N), the wave tiling (for 256 x 256: eight waves
of 128 x 64, or four of 128 x 128 -- which also fixes the engine, its LDS
ring depth and DepthU) and the tile-raster group height.  Results are kept in
a JSON database keyed by (dtype, layout, M, N, K) and looked up by
``best_tile``.
"""
from __future__ import annotations

import json
import os

import torch

from . import _capi as C
from .gemm import make_problem, make_tile, run

# (mt_m, mt_n, waves_m, waves_n); waves (0, 0) is the tile's default engine
TILES = [(256, 256, 2, 4), (256, 256, 2, 2), (256, 128, 0, 0), (128, 256, 0, 0), (128, 128, 0, 0), (64, 64, 0, 0)]
DB_PATH = os.environ.get("FFGEMM_TUNE_DB", os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune_db.json"))


def _time(fn, steps=10, warmup=3):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(steps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def load_db(path=None):
    path = path or DB_PATH
    if os.path.isfile(path):
        with open(path) as f:
            return json.load(f)
    return {}


def save_db(db, path=None):
    path = path or DB_PATH
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(db, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def key(dtype_code, layout, M, N, K):
    return f"{dtype_code}:{layout}:{M}x{N}x{K}"


def tune_gemm(problem: C.Problem, A, B, Cm, D, tiles=TILES, group_ms=(0,), steps=10, verbose=False):
    """Times every candidate tile for one problem; returns {'best': tile, 'results': [...]}."""
    results = []
    flop = 2.0 * problem.M * problem.N * problem.K * max(1, problem.batch)
    for t in tiles:
        wm, wn = (t[2], t[3]) if len(t) > 3 else (0, 0)
        for gm in group_ms:
            tile = make_tile((t[0], t[1], gm, wm, wn))
            try:
                ms = _time(lambda: run(problem, A, B, Cm, D, tile))
            except C.FFGemmError:
                continue  # no kernel for this tile / wave tiling with this dtype and layout
            tf = flop / (ms * 1e-3) / 1e12
            results.append({"tile": [t[0], t[1], gm, wm, wn], "ms": ms, "tflops": tf})
            if verbose:
                print(f"MT{t[0]}x{t[1]} waves={wm}x{wn} gm={gm}: Performance = {tf:.2f} TFLOPs")
    results.sort(key=lambda r: r["ms"])
    return {"best": tuple(results[0]["tile"]) if results else None, "results": results}


def tune_and_record(problem, A, B, Cm, D, db=None, path=None, **kw):
    db = load_db(path) if db is None else db
    r = tune_gemm(problem, A, B, Cm, D, **kw)
    if r["best"] is not None:
        lay = ["NN", "NT", "TN", "TT"][problem.layout]
        db[key(problem.dtype, lay, problem.M, problem.N, problem.K)] = {"tile": list(r["best"]),
                                                                         "tflops": r["results"][0]["tflops"]}
        save_db(db, path)
    return r


def best_tile(dtype_code, layout, M, N, K, path=None):
    e = load_db(path).get(key(dtype_code, layout, M, N, K))
    return tuple(e["tile"]) if e else None


def tune_grouped(problems, As, Bs, Ds, tiles=((256, 256), (256, 128), (128, 256), (128, 128), (64, 64)), steps=10,
                 rounds=2):
    """Picks the grouped-GEMM macro tile by timing the whole persistent launch.

    The candidates are timed in `rounds` interleaved rounds and each keeps its
    best time, so the first candidate is not the one that pays for the clocks
    ramping up."""
    from .grouped import GroupedPlan

    plans = {}
    for t in tiles:
        try:
            plans[tuple(t)] = GroupedPlan(problems, As, Bs, None, Ds, tile=t)
        except C.FFGemmError:
            continue
    best = {}
    for _ in range(rounds):
        for t, plan in plans.items():
            best[t] = min(best.get(t, float("inf")), _time(plan.run, steps=steps))
    for plan in plans.values():
        plan.close()
    results = sorted(({"tile": list(t), "ms": ms} for t, ms in best.items()), key=lambda r: r["ms"])
    return {"best": tuple(results[0]["tile"]), "results": results}
