"""The project's autotuner: successor of the reference's perf_test.py sweep
(test/perf_test.py:58-167) for the gfx950 knobs.

The reference sweeps padding, LDS buffers, MFMA size, MFMA tiles per wave
(r, s), waves per group (x, y) and DepthU, keeps the best TFLOP/s and prints
``Performance = X TFLOPs``.  Here the knobs that exist in the gfx950 kernels
are the macro tile (MT_M x MT_N), the wave tiling (for 256 x 256: eight waves
of 128 x 64, or four of 128 x 128 -- which also fixes the engine, its LDS
ring depth and DepthU), the MFMA instruction (32 x 32 x 16, or 16 x 16 x 32
at 64 x 64 and 128 x 128), the tile-raster group height and, for problems
with few output tiles, split-K (K cut into S slices whose fp32 partials a
reduction pass sums; it runs with a workspace).  Results are kept in a CSV
table keyed by (dtype, layout, M, N, K), one line
``dtype,layout,M,N,K,mt_m,mt_n,group_m,waves_m,waves_n,mfma_mn,split_k,tflops``
per problem.
The C library reads the same table (ffgemm_tune_db_load): a problem run with
no tile takes its entry.  ``best_tile`` looks an entry up from Python.
"""
from __future__ import annotations

import csv
import ctypes
import os

import torch

from . import _capi as C
from .gemm import make_problem, make_tile, query_tile_ws, run

# (mt_m, mt_n, waves_m, waves_n, mfma_mn); waves (0, 0) is the tile's default
# engine, mfma 0 its 32 x 32 MFMAs
TILES = [(256, 256, 2, 4, 0), (256, 256, 2, 2, 0), (256, 128, 0, 0, 0), (128, 256, 0, 0, 0), (128, 128, 0, 0, 0),
         (128, 128, 0, 0, 16), (64, 64, 0, 0, 0), (64, 64, 0, 0, 16)]
DB_PATH = os.environ.get("FFGEMM_TUNE_DB", os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune_db.csv"))
FIELDS = ["dtype", "layout", "M", "N", "K", "mt_m", "mt_n", "group_m", "waves_m", "waves_n", "mfma_mn", "split_k",
          "tflops"]
# split-K candidates for problems whose best unsplit tile leaves CUs idle
SPLITS = (1, 2, 3, 4, 6, 8, 12, 16)
# an entry is recorded only when it beats the library's own choice by this
# much: run-to-run differences on one box reach ~3 %
MARGIN = 0.05


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
    db = {}
    if os.path.isfile(path):
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                db[key(int(r["dtype"]), r["layout"], int(r["M"]), int(r["N"]), int(r["K"]))] = {
                    "tile": [int(r[k]) for k in ("mt_m", "mt_n", "group_m", "waves_m", "waves_n")]
                    + [int(r.get("mfma_mn") or 0), int(r.get("split_k") or 0)],
                    "tflops": float(r["tflops"])}
    return db


def save_db(db, path=None):
    """Writes the table and has the C library use it from now on."""
    path = path or DB_PATH
    tmp = path + ".tmp"
    with open(tmp, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(FIELDS)
        for k in sorted(db):
            dt, lay, mnk = k.split(":")
            M, N, K = mnk.split("x")
            e = db[k]
            tile = list(e["tile"]) + [0] * (7 - len(e["tile"]))
            w.writerow([dt, lay, M, N, K] + tile + [round(e["tflops"], 1)])
    os.replace(tmp, path)
    C.check(C.lib().ffgemm_tune_db_load(path.encode(), None))


def key(dtype_code, layout, M, N, K):
    return f"{dtype_code}:{layout}:{M}x{N}x{K}"


def split_workspace_bytes(problem: C.Problem, S: int) -> int:
    """Workspace of a split into S slices: S fp32 planes of roundup(M, 4) x N
    (the C library's layout), rounded up to 256 bytes."""
    return (S * ((problem.M + 3) // 4 * 4) * problem.N * 4 + 255) // 256 * 256


def tune_gemm(problem: C.Problem, A, B, Cm, D, tiles=TILES, group_ms=(0,), splits=(1,), steps=10, rounds=2,
              verbose=False):
    """Times every candidate tile (x split-K count) for one problem in `rounds`
    interleaved rounds (each keeps its best time, so no candidate pays alone
    for the clock ramp); returns {'best': tile, 'results': [...]} fastest
    first.  Every candidate runs through ffgemm_run_ws with one workspace
    large enough for the largest split."""
    flop = 2.0 * problem.M * problem.N * problem.K * max(1, problem.batch)
    nws = max(split_workspace_bytes(problem, S) for S in splits) if max(splits) > 1 else 0
    ws = torch.empty(max(nws, 256), dtype=torch.uint8, device=torch.device("cuda", torch.cuda.current_device()))
    cands = []
    for t in tiles:
        wm, wn = (t[2], t[3]) if len(t) > 3 else (0, 0)
        mf = t[4] if len(t) > 4 else 0
        for gm in group_ms:
            for S in splits:
                rec = (t[0], t[1], gm, wm, wn, mf, S)
                try:
                    run(problem, A, B, Cm, D, make_tile(rec), workspace=ws)
                    # the split the library runs (slices are at least 128
                    # elements of K, so a large S can come to fewer slices)
                    eff = query_tile_ws(problem, make_tile(rec)).split_k if S > 1 else 1
                except C.FFGemmError:
                    continue  # no kernel for this tile / wave tiling / split with this dtype and layout
                rec = rec[:6] + (max(1, eff),)
                if rec not in cands:
                    cands.append(rec)
    best = {}
    for _ in range(rounds):
        for rec in cands:
            tile = make_tile(rec)
            best[rec] = min(best.get(rec, float("inf")),
                            _time(lambda: run(problem, A, B, Cm, D, tile, workspace=ws), steps=steps))
    results = sorted(({"tile": list(rec), "ms": ms, "tflops": flop / (ms * 1e-3) / 1e12} for rec, ms in best.items()),
                     key=lambda r: r["ms"])
    if verbose:
        for r in results:
            t = r["tile"]
            print(f"MT{t[0]}x{t[1]} waves={t[3]}x{t[4]} mfma={t[5] or 32} gm={t[2]} split={t[6]}: "
                  f"Performance = {r['tflops']:.2f} TFLOPs")
    return {"best": tuple(results[0]["tile"]) if results else None, "results": results}


def library_default(problem: C.Problem):
    """The candidate record of the tile the library picks with an empty table
    and a workspace (at 256 x 256 with its wave tiling, which the candidates
    name there; its split-K count)."""
    C.check(C.lib().ffgemm_tune_db_load(b"", None))
    try:
        t = C.Tile()
        C.check(C.lib().ffgemm_query_tile_ws(ctypes.byref(problem), ctypes.byref(t)))
    finally:
        C.check(C.lib().ffgemm_tune_db_load(None, None))
    S = max(1, t.split_k)
    if (t.mt_m, t.mt_n) == (256, 256):
        return (256, 256, 0, t.waves_m, t.waves_n, 0, S)
    return (t.mt_m, t.mt_n, 0, 0, 0, 0, S)


def tune_and_record(problem, A, B, Cm, D, db=None, path=None, **kw):
    """Tunes one problem and records its best tile when that beats the
    library's own choice by MARGIN; an entry that no longer does is dropped."""
    db = load_db(path) if db is None else db
    r = tune_gemm(problem, A, B, Cm, D, **kw)
    if r["best"] is None:
        return r
    lay = ["NN", "NT", "TN", "TT"][problem.layout]
    k = key(problem.dtype, lay, problem.M, problem.N, problem.K)
    default = library_default(problem)
    ms_default = next((x["ms"] for x in r["results"] if tuple(x["tile"]) == default), None)
    r["default"] = default
    if ms_default is None or r["results"][0]["ms"] < (1.0 - MARGIN) * ms_default:
        db[k] = {"tile": list(r["best"]), "tflops": r["results"][0]["tflops"]}
    else:
        db.pop(k, None)
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


# The problems the benchmark runs, then common shapes away from the square
# case: (dtype, layout, M, N, K, output dtype or None)
TUNE_PROBLEMS = (
    [(dt, lay, 4096, 4096, 4096, None) for dt in (C.BF16, C.F16) for lay in ("NN", "NT", "TN", "TT")]
    + [(C.BF16, "NN", 4096, 8192, 4096, None), (C.BF16, "TN", 4096, 8192, 4096, None),   # autograd fwd, dA
       (C.BF16, "NT", 4096, 4096, 8192, None)]                                            # autograd dB
    + [(C.F8E4M3, "TN", 4096, 4096, 4096, C.F32), (C.F8E4M3, "TN", 8192, 8192, 8192, C.F32),
       (C.F32, "NN", 8192, 8192, 8192, None),
       (C.F16, "TN", 480, 512, 512, None), (C.F32, "NN", 480, 512, 512, None)]
    + [(C.BF16, "TN", M, N, K, None) for (M, N, K) in
       ((2048, 2048, 8192), (8192, 1024, 8192), (1024, 8192, 8192), (16384, 16384, 1024), (3072, 3072, 3072),
        (6144, 6144, 6144), (2048, 8192, 2048), (8192, 2048, 2048), (1024, 1024, 16384), (12288, 4096, 4096))]
    + [(C.BF16, "NN", M, N, K, None) for (M, N, K) in ((2048, 4096, 4096), (8192, 8192, 2048), (3072, 3072, 3072))]
    + [(dt, "NT", 8192, 8192, 8192, None) for dt in (C.BF16, C.F16)])


def _operands(dt, lay, M, N, K, out, dev):
    ta, tb = lay[0] == "T", lay[1] == "T"
    tdt = {C.BF16: torch.bfloat16, C.F16: torch.float16, C.F32: torch.float32}.get(dt)
    if dt == C.F8E4M3:  # e4m3 bytes below the NaN encodings; speed does not depend on values
        A = torch.randint(0, 0x70, (M * K,), dtype=torch.uint8, device=dev)
        B = torch.randint(0, 0x70, (K * N,), dtype=torch.uint8, device=dev)
    else:
        A = (torch.randn(M * K, device=dev) / K ** 0.5).to(tdt)
        B = (torch.randn(K * N, device=dev) / K ** 0.5).to(tdt)
    D = torch.empty(M * N, dtype=torch.float32 if (out == C.F32 or dt in (C.F32, C.F8E4M3)) else tdt, device=dev)
    p = make_problem(lay, dt, M, N, K, K if ta else M, N if tb else K, M, M, dtype_out=out or 0)
    return p, A, B, D


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="Tune the benchmark's problems and record the table.")
    ap.add_argument("--db", default=None, help="table to update (default: the package's tune_db.csv)")
    ap.add_argument("--rounds", type=int, default=3)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--group-ms", default="0", help="tile-raster group heights to try, e.g. 0,2,8 (0: default)")
    ap.add_argument("--set", default="bench", choices=("bench", "phantom", "all"),
                    help="bench: the benchmark's problems and common shapes (split-K candidates where 256 x 256 "
                         "tiles leave CUs idle); phantom: the reference's thin shapes (test/phantom_test.sh) with "
                         "split-K candidates")
    a = ap.parse_args(argv)
    gms = tuple(int(x) for x in a.group_ms.split(","))
    dev = torch.device("cuda", torch.cuda.current_device())
    db = load_db(a.db)
    todo = []
    if a.set in ("bench", "all"):
        # split-K candidates where 256 x 256 tiles leave CUs idle
        ncu = torch.cuda.get_device_properties(dev).multi_processor_count
        todo += [(pr, SPLITS if -(-pr[2] // 256) * -(-pr[3] // 256) < ncu else (1,)) for pr in TUNE_PROBLEMS]
    if a.set in ("phantom", "all"):
        from . import phantom
        code = {"f32": C.F32, "bf16": C.BF16, "f16": C.F16}
        seen = set()
        for dt, _, M, N, K, _ in phantom.problems():
            if (dt, M, N, K) not in seen:
                seen.add((dt, M, N, K))
                todo.append(((code[dt], "TN", M, N, K, None), SPLITS))
    for (dt, lay, M, N, K, out), splits in todo:
        p, A, B, D = _operands(dt, lay, M, N, K, out, dev)
        r = tune_and_record(p, A, B, None, D, db=db, path=a.db, rounds=a.rounds, steps=a.steps, group_ms=gms,
                            splits=splits)
        best = r["results"][0]
        ms_def = next((x["ms"] for x in r["results"] if tuple(x["tile"]) == r["default"]), float("nan"))
        rec = "recorded" if key(dt, lay, M, N, K) in db else "library default kept"
        m16 = next((x for x in r["results"] if x["tile"][5] == 16), None)
        m16s = f", best 16x16 MFMA {tuple(m16['tile'])} {m16['ms'] * 1e3:.1f} us" if m16 else ""
        print(f"dtype {dt} {lay} {M}x{N}x{K}: best {tuple(best['tile'])} {best['ms'] * 1e3:.1f} us, "
              f"default {r['default']} {ms_def * 1e3:.1f} us{m16s}: {rec}", flush=True)
        del A, B, D


if __name__ == "__main__":
    main()
