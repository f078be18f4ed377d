"""C ABI checks that need no GPU: the library loads, exports every entry point
include/ffgemm.h declares, and rejects what checkKernelParameters would
(src/gemmMfmaKernelWriter.h:392-399) with RTN_ERR instead of exiting."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

import tensilelite_amd as tl
from tensilelite_amd import _capi as C

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "ffgemm.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ffgemm_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = C.lib()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings_and_version():
    lib = C.lib()
    assert [lib.ffgemm_status_string(i).decode() for i in range(4)] == ["SUCCESS", "RTN_WARN", "RTN_ERR", "RTN_FATAL"]
    assert b"gfx950" in lib.ffgemm_version()


def _p(layout="NN", dtype=C.BF16, M=512, N=512, K=512, **kw):
    ta, tb = layout[0] == "T", layout[1] == "T"
    d = dict(lda=K if ta else M, ldb=N if tb else K, ldc=M, ldd=M)
    d.update(kw)
    return tl.make_problem(layout, dtype, M, N, K, d["lda"], d["ldb"], d["ldc"], d["ldd"],
                           alpha=kw.get("alpha", 1.0), beta=kw.get("beta", 0.0), compute=kw.get("compute", 0),
                           dtype_out=kw.get("dtype_out", 0))


@pytest.mark.parametrize("layout", ["NN", "NT", "TN", "TT"])
@pytest.mark.parametrize("dtype", [C.BF16, C.F16, C.F32])
def test_query_tile_accepts_supported(layout, dtype):
    t = tl.query_tile(_p(layout, dtype, 4096, 4096, 4096))
    assert (t.mt_m, t.mt_n) == (256, 256)
    assert t.mfma_mn == 32


@pytest.mark.parametrize("shape,tile", [
    ((4096, 4096, 4096), (256, 256)),
    ((3072, 3072, 3072), (256, 256)),    # 144 tiles in one pass beat 288 in two (256 x 128)
    ((2048, 2048, 8192), (128, 128)),
    ((8192, 1024, 8192), (256, 128)),
    ((1024, 1024, 16384), (64, 64)),
    ((6144, 6144, 6144), (256, 256)),
])
def test_default_tile_is_the_measured_best(shape, tile):
    """With no tile and an empty table the library estimates passes x tile
    time; these shapes' tiles are the autotuner's measured winners (bf16 TN)."""
    t = tl.query_tile(_p("TN", C.BF16, *shape))
    assert (t.mt_m, t.mt_n) == tile


def test_query_tile_small_shapes_pick_small_tiles():
    t = tl.query_tile(_p("NN", C.BF16, 480, 512, 512))
    assert (t.mt_m, t.mt_n) == (64, 64)


@pytest.mark.parametrize("bad, msg", [
    (dict(lda=100), "lda"),                                # lda < M
    (dict(layout="NN", dtype=C.F8E4M3), "TN"),             # fp8 is TN only
    (dict(compute=C.COMPUTE_BF16X3), "fp32 operands"),     # bf16x3 needs fp32
    (dict(dtype=9), "dtype"),
])
def test_query_tile_rejects_with_rtn_err(bad, msg):
    kw = dict(bad)
    layout = kw.pop("layout", "NN")
    dtype = kw.pop("dtype", C.BF16)
    K = kw.pop("K", 512)
    p = _p(layout, dtype, 512, 512, K, **kw)
    with pytest.raises(tl.FFGemmError) as e:
        tl.query_tile(p)
    assert e.value.status == C.ERR
    assert msg in str(e.value)


def test_operand_size_limits():
    """Buffer offsets are 31-bit: a K-contiguous operand is addressed from each
    tile's first row, so it may exceed 2 GiB; a larger M/N-contiguous one runs
    in K chunks; only an ld too long for 320 rows in 2 GiB is refused."""
    M, N, K = 140000, 256, 8192                   # A: 140000 rows of 8192 bf16 = 2.29 GB
    t = tl.query_tile(_p("TN", C.BF16, M, N, K))  # A and B K-contiguous
    assert t.mt_m > 0
    t = tl.query_tile(_p("NN", C.BF16, M, N, K))  # NN: A M-contiguous, 8192 rows of lda = M: K chunks
    assert t.mt_m > 0
    Kw = 4194304                                  # 257 rows of 2^22 bf16 exceed 2 GiB
    with pytest.raises(tl.FFGemmError) as e:
        tl.query_tile(_p("TN", C.BF16, 512, 512, Kw))
    assert e.value.status == C.ERR and "2 GiB" in str(e.value)
    with pytest.raises(tl.FFGemmError):           # M-contiguous: 320 rows of lda = 2^22 exceed 2 GiB
        tl.query_tile(_p("NN", C.BF16, Kw, 64, 1024))
    # bf16x3 planes are K-contiguous too: a workspace is offered past 2 GiB of planes
    x3 = tl.make_problem("TN", C.F32, M, N, K, K, K, M, M, compute=C.COMPUTE_BF16X3)
    assert tl.workspace_size(x3) == ((M * 3 * K * 2 + 255) // 256 + (N * 3 * K * 2 + 255) // 256) * 256


@pytest.mark.parametrize("kw, mt", [
    (dict(K=500, layout="TN", lda=504, ldb=504), 64),      # K-contiguous operand, K % 8
    (dict(ldd=513), 64),                                   # D rows off the 16-byte grid
    (dict(lda=4097, M=4096, N=4096), 128),                 # enough 128 x 128 tiles for the CUs
    (dict(dtype=C.F32, ldb=513), 64),
    (dict(dtype=C.F32, compute=C.COMPUTE_BF16X3, lda=515), 64),
    (dict(dtype=C.F8E4M3, layout="TN", K=200, lda=200, ldb=200, dtype_out=C.F32), 128),
])
def test_general_stride_plan(kw, mt):
    """Leading dimensions, K or batch strides off the 16-byte grid run the
    general-stride kernel (element-wise loads and stores): a square 64 x 64 or
    128 x 128 tile (fp8: 128), 2 x 2 waves, two LDS stages, never split."""
    kw = dict(kw)
    layout = kw.pop("layout", "NN")
    dtype = kw.pop("dtype", C.BF16)
    M, N, K = kw.pop("M", 512), kw.pop("N", 512), kw.pop("K", 512)
    p = _p(layout, dtype, M, N, K, **kw)
    t = tl.query_tile(p)
    assert (t.mt_m, t.mt_n, t.waves_m, t.waves_n, t.lds_buffers, t.mfma_mn) == (mt, mt, 2, 2, 2, 32)
    assert tl.query_tile_ws(p).split_k == 1
    # a record of the form runs as given; any other is refused with the reason
    assert tl.query_tile(p, (64, 64) if dtype != C.F8E4M3 else (128, 128)).mt_m in (64, 128)
    with pytest.raises(tl.FFGemmError) as e:
        tl.query_tile(p, (256, 256))
    assert e.value.status == C.ERR and "general-stride" in str(e.value)
    with pytest.raises(tl.FFGemmError):
        tl.query_tile(p, (128, 128, 0, 0, 0, 0, 2))    # no split
    with pytest.raises(tl.FFGemmError):
        tl.query_tile(p, (64, 64, 0, 0, 0, 16))        # no 16 x 16 MFMA form


def test_tile_from_reference_off_grid_problem():
    """A reference record whose tile the general-stride kernel lacks runs as
    the library's choice for the problem, with RTN_WARN saying so."""
    p = tl.make_problem("TN", C.BF16, 512, 512, 512, 513, 513, 513, 513)
    st, t, msg = tl.tile_from_reference(p, 4, 2, 2, 4, 32)      # MT 256 x 256, 2 x 4 waves
    assert st == C.WARN and (t.mt_m, t.mt_n, t.waves_m, t.waves_n) == (64, 64, 2, 2)
    assert "runs as MT64x64" in msg
    st, t, _ = tl.tile_from_reference(p, 2, 2, 2, 2, 32)        # MT 128 x 128, 2 x 2 waves: exact
    assert st == C.SUCCESS and (t.mt_m, t.mt_n) == (128, 128)


def test_general_stride_pointer_alignment():
    """The general-stride form needs pointers aligned to their element size: a
    bf16 matrix at an odd byte address is refused before any launch."""
    p = _p("NN", C.BF16, 64, 64, 64, lda=65)
    fake = ctypes.c_void_p(0x100000)
    odd = ctypes.c_void_p(0x100001)
    st = C.lib().ffgemm_run(ctypes.byref(p), odd, fake, None, fake, None, None)
    assert st == C.ERR and "element size" in C.lib().ffgemm_last_error().decode()


def test_grouped_plan_rejects_off_grid_problems():
    """Grouped plans run the aligned engines only: a member off the 16-byte
    grid is refused at create, before anything is allocated."""
    p = _p("NN", C.BF16, 256, 256, 256, lda=257)
    probs = (C.Problem * 2)(p, p)
    fake = ctypes.c_void_p(0x100000)
    arr = (ctypes.c_void_p * 2)(fake, fake)
    plan = ctypes.c_void_p(0)
    t = C.Tile()
    t.mt_m = t.mt_n = 64
    st = C.lib().ffgemm_grouped_plan_create(ctypes.byref(plan), 2, probs, arr, arr, None, arr, ctypes.byref(t))
    assert st == C.ERR and plan.value is None
    assert "16-byte grid" in C.lib().ffgemm_last_error().decode()


def test_bad_tile_request_is_rejected():
    with pytest.raises(tl.FFGemmError) as e:
        tl.query_tile(_p(), tile=(96, 96))
    assert "unsupported macro tile" in str(e.value)


def test_kernarg_entry_rejects_bad_dtype_without_launch():
    st = C.lib().ffgemm_run_tn_kernarg(None, None, None, None, 1.0, 0.0, 64, 0, 64, 0, 64, 0, 64, 0,
                                       64, 64, 64, 1, 7, None, None)
    assert st == C.ERR
    assert b"data type" in C.lib().ffgemm_last_error()


def tensile_args(D, C_, A, B, alpha_bits, beta_bits, M, N, K, batch, ldd, bsD, ldc, bsC, lda, bsA, ldb, bsB,
                 sizes=(0, 0, 0)):
    """The reference's 148-byte Tensile-layout argument buffer (src/gemmMfma.cpp:72-102)."""
    import struct
    buf = struct.pack("<3Q4Q2I8I4I9I", sizes[0], sizes[1], sizes[2], D, C_, A, B, alpha_bits, beta_bits,
                      ldd, bsD, ldc, bsC, lda, bsA, ldb, bsB, M, N, batch, K, *([0] * 9))
    assert len(buf) == 148
    return buf


def test_tensile_args_entry_rejects_without_launch():
    lib = C.lib()
    buf = tensile_args(16, 16, 16, 16, 0x3F800000, 0, 64, 64, 64, 1, 64, 4096, 64, 4096, 64, 4096, 64, 4096)
    assert lib.ffgemm_run_tensile_args(buf, 100, C.BF16, None, None) == C.ERR          # short buffer
    assert b"148" in lib.ffgemm_last_error()
    assert lib.ffgemm_run_tensile_args(buf, 148, 7, None, None) == C.ERR               # data type
    small = tensile_args(16, 16, 16, 16, 0x3F800000, 0, 64, 64, 64, 1, 64, 4096, 64, 4096, 64, 4096, 64, 4096,
                         sizes=(4096, 64 * 63, 4096))                                 # sizeA one row short
    assert lib.ffgemm_run_tensile_args(small, 148, C.BF16, None, None) == C.ERR
    assert b"sizeA" in lib.ffgemm_last_error()


def test_grouped_plan_rejects_empty_group():
    h = ctypes.c_void_p()
    st = C.lib().ffgemm_grouped_plan_create(ctypes.byref(h), 0, None, None, None, None, None, None)
    assert st == C.ERR


def test_workspace_size_bf16x3_planes():
    """bf16x3 asks for hi/lo planes of (M + N) rows x 3*roundup(K, 64) bf16 per
    batch (each operand rounded to 256 bytes); other forms need none."""
    M, N, K = 200, 264, 132
    p = tl.make_problem("NN", C.F32, M, N, K, M, K, M, M, compute=C.COMPUTE_BF16X3)
    kp = 192
    assert tl.workspace_size(p) == ((M * 3 * kp * 2 + 255) // 256 + (N * 3 * kp * 2 + 255) // 256) * 256
    p.batch = 3
    assert tl.workspace_size(p) == ((3 * M * 3 * kp * 2 + 255) // 256 + (3 * N * 3 * kp * 2 + 255) // 256) * 256
    q = tl.make_problem("NN", C.F32, M, N, K, M, K, M, M)
    assert tl.workspace_size(q) == 0
    r = tl.make_problem("TN", C.BF16, M, N, K, K, K, M, M)
    assert tl.workspace_size(r) == 0


@pytest.mark.parametrize("layout,dtype,geo", [
    ("TN", C.BF16, (256, 256, 2, 4, 64, 3)),    # eight-wave engine, K-contiguous rows
    ("NT", C.F16, (256, 256, 2, 4, 64, 3)),     # eight-wave engine, transposing reads
    ("NN", C.BF16, (256, 256, 2, 2, 64, 3)),    # split-ring engine
    ("TT", C.BF16, (256, 256, 2, 2, 32, 4)),    # pipelined 4-wave engine
    ("TN", C.F8E4M3, (256, 256, 2, 4, 128, 3)),
    ("NN", C.F32, (256, 256, 2, 2, 32, 2)),
])
def test_query_tile_reports_engine_geometry(layout, dtype, geo):
    """The tile record describes the engine that will run the 4096^3 problem."""
    p = _p(layout, dtype, 4096, 4096, 4096)
    if dtype == C.F8E4M3:
        p.dtype_out = C.F32
    t = tl.query_tile(p)
    assert (t.mt_m, t.mt_n, t.waves_m, t.waves_n, t.depth_u, t.lds_buffers) == geo


@pytest.mark.parametrize("layout", ["NN", "NT", "TN", "TT"])
@pytest.mark.parametrize("dtype", [C.BF16, C.F32])
@pytest.mark.parametrize("size", [4096, 1000, 296])
def test_query_tile_record_round_trips(layout, dtype, size):
    """A tile record reported by ffgemm_query_tile is accepted when passed back."""
    p = _p(layout, dtype, size, size, size)
    t = tl.query_tile(p)
    t2 = tl.query_tile(p, t)
    assert (t2.mt_m, t2.mt_n, t2.waves_m, t2.waves_n, t2.depth_u, t2.lds_buffers) == \
        (t.mt_m, t.mt_n, t.waves_m, t.waves_n, t.depth_u, t.lds_buffers)


@pytest.mark.parametrize("layout,dtype,waves,geo", [
    ("TN", C.BF16, (2, 2), (2, 2, 32, 4)),     # four-wave pipelined engine
    ("NN", C.BF16, (2, 4), (2, 4, 64, 3)),     # eight-wave engine, transposing reads for A
    ("NN", C.BF16, (2, 2), (2, 2, 64, 3)),     # split-ring engine (the default)
    ("TT", C.F16, (2, 4), (2, 4, 64, 3)),
    ("TN", C.F8E4M3, (2, 2), (2, 2, 128, 2)),  # fp8 on the first engine
])
def test_wave_tiling_knob(layout, dtype, waves, geo):
    p = _p(layout, dtype, 4096, 4096, 4096)
    if dtype == C.F8E4M3:
        p.dtype_out = C.F32
    t = C.Tile()
    t.mt_m, t.mt_n, t.waves_m, t.waves_n = 256, 256, waves[0], waves[1]
    r = tl.query_tile(p, t)
    assert (r.waves_m, r.waves_n, r.depth_u, r.lds_buffers) == geo


@pytest.mark.parametrize("layout,dtype,tile,waves", [
    ("NN", C.F32, (256, 256), (2, 4)),   # no eight-wave fp32 kernel
    ("TN", C.BF16, (128, 128), (2, 4)),  # eight waves only at 256 x 256
    ("TN", C.BF16, (256, 256), (4, 2)),
    ("TN", C.BF16, (256, 256), (2, 3)),
])
def test_wave_tiling_knob_rejects(layout, dtype, tile, waves):
    p = _p(layout, dtype, 4096, 4096, 4096)
    t = C.Tile()
    t.mt_m, t.mt_n, t.waves_m, t.waves_n = tile[0], tile[1], waves[0], waves[1]
    with pytest.raises(C.FFGemmError, match="RTN_ERR"):
        tl.query_tile(p, t)


@pytest.mark.parametrize("layout", ["NN", "NT", "TN", "TT"])
@pytest.mark.parametrize("dtype", [C.BF16, C.F16, C.F32])
@pytest.mark.parametrize("mt", [64, 128])
def test_mfma_knob_16x16(layout, dtype, mt):
    """mfma_mn = 16 selects the 16 x 16 MFMA form of the first engine (16 x 16
    x 32 for 16-bit operands, 16 x 16 x 4 for fp32; four waves of 2 x 2, 128
    bytes of K per stage, two LDS stages) and is reported back."""
    t = C.Tile()
    t.mt_m, t.mt_n, t.mfma_mn = mt, mt, 16
    r = tl.query_tile(_p(layout, dtype, 1000, 520, 328), t)
    du = 32 if dtype == C.F32 else 64
    assert (r.mt_m, r.mt_n, r.mfma_mn, r.waves_m, r.waves_n, r.depth_u, r.lds_buffers) == (mt, mt, 16, 2, 2, du, 2)
    t32 = C.Tile()
    t32.mt_m, t32.mt_n, t32.mfma_mn = mt, mt, 32
    assert tl.query_tile(_p(layout, dtype, 1000, 520, 328), t32).mfma_mn == 32


@pytest.mark.parametrize("dtype,tile,mfma,waves", [
    (C.BF16, (256, 256), 16, (0, 0)),  # 16 x 16 only at 64 x 64 and 128 x 128
    (C.BF16, (128, 256), 16, (0, 0)),
    (C.BF16, (64, 64), 16, (2, 4)),    # four waves only
    (C.BF16, (64, 64), 8, (0, 0)),     # no such instruction width
])
def test_mfma_knob_rejects(dtype, tile, mfma, waves):
    t = C.Tile()
    t.mt_m, t.mt_n, t.mfma_mn, t.waves_m, t.waves_n = tile[0], tile[1], mfma, waves[0], waves[1]
    with pytest.raises(C.FFGemmError, match="RTN_ERR"):
        tl.query_tile(_p("NN", dtype, 1000, 520, 328), t)


def test_mfma_knob_16x16_not_for_bf16x3_or_fp8():
    """The 16 x 16 form takes bf16, fp16 and fp32 operands; bf16x3 and fp8 run
    32 x 32 MFMAs only."""
    for p in (_p("NN", C.F32, 1000, 520, 328, compute=C.COMPUTE_BF16X3),
              _p("TN", C.F8E4M3, 1000, 520, 320, dtype_out=C.F32)):
        with pytest.raises(C.FFGemmError, match="16x16"):
            tl.query_tile(p, (64, 64, 0, 0, 0, 16))


def test_tuned_tile_lookup(tmp_path, monkeypatch):
    """tile='tuned' reads the autotuner table (a missing entry means the library's choice)."""
    import importlib
    from tensilelite_amd import autotune
    G = importlib.import_module("tensilelite_amd.gemm")   # the module (the package re-exports gemm())
    monkeypatch.setattr(autotune, "DB_PATH", str(tmp_path / "db.csv"))
    try:
        autotune.save_db({autotune.key(C.BF16, "TN", 4096, 4096, 4096): {"tile": [256, 256, 0, 2, 2], "tflops": 1.0},
                          autotune.key(C.F16, "NN", 480, 512, 512): {"tile": [64, 64, 0, 0, 0, 16], "tflops": 1.0}})
        p = _p("TN", C.BF16, 4096, 4096, 4096)
        assert list(G._tuned(p)) == [256, 256, 0, 2, 2, 0, 0]
        t = tl.query_tile(p, G.make_tile(G._tuned(p)))
        assert (t.waves_m, t.waves_n, t.mfma_mn) == (2, 2, 32)
        p16 = _p("NN", C.F16, 480, 512, 512)
        assert list(G._tuned(p16)) == [64, 64, 0, 0, 0, 16, 0]
        assert tl.query_tile(p16).mfma_mn == 16        # the C library reads the same table
        assert G._tuned(_p("TN", C.BF16, 1024, 1024, 1024)) is None
    finally:
        C.check(C.lib().ffgemm_tune_db_load(b"", None))


def _load_table(path):
    n = ctypes.c_int32(-1)
    st = C.lib().ffgemm_tune_db_load(str(path).encode(), ctypes.byref(n))
    return st, n.value


def test_tune_table_consulted_by_the_library(tmp_path):
    """Problems given no tile take the table's entry; an entry that does not
    fit its problem is passed over; a malformed table is refused."""
    f = tmp_path / "t.csv"
    f.write_text("dtype,layout,M,N,K,mt_m,mt_n,group_m,waves_m,waves_n,tflops\n"
                 f"{C.BF16},TN,4096,4096,4096,256,256,0,2,2,1000.0\n"
                 f"{C.F32},NN,4096,4096,4096,256,256,0,2,4,100.0\n"      # no eight-wave fp32 kernel
                 f"{C.BF16},NN,480,512,512,128,128,4,0,0,50.0\n"
                 f"{C.F16},TT,480,512,512,64,64,0,0,0,16,40.0\n")      # with the mfma_mn field
    try:
        assert _load_table(f) == (C.SUCCESS, 4)
        t = tl.query_tile(_p("TN", C.BF16, 4096, 4096, 4096))          # the table's four-wave engine
        assert (t.mt_m, t.mt_n, t.waves_m, t.waves_n) == (256, 256, 2, 2)
        t = tl.query_tile(_p("NN", C.F32, 4096, 4096, 4096))           # passed over: the default
        assert (t.mt_m, t.mt_n) == (256, 256) and t.waves_n == 2
        t = tl.query_tile(_p("NN", C.BF16, 480, 512, 512))
        assert (t.mt_m, t.mt_n, t.group_m, t.mfma_mn) == (128, 128, 4, 32)
        t = tl.query_tile(_p("TT", C.F16, 480, 512, 512))
        assert (t.mt_m, t.mt_n, t.mfma_mn) == (64, 64, 16)
        t = tl.query_tile(_p("TN", C.BF16, 4096, 4096, 1024))          # not in the table
        assert (t.mt_m, t.mt_n, t.waves_n) == (256, 256, 4)
        bad = tmp_path / "bad.csv"
        bad.write_text("1,TN,64,64\n")
        assert _load_table(bad)[0] == C.ERR
        assert _load_table(tmp_path / "missing.csv") == (C.SUCCESS, 0)
    finally:
        C.check(C.lib().ffgemm_tune_db_load(b"", None))
    t = tl.query_tile(_p("NN", C.BF16, 480, 512, 512))
    assert (t.mt_m, t.mt_n) == (64, 64)


def test_committed_tune_table_is_valid():
    """Every entry of the committed table is accepted as is for its problem."""
    from tensilelite_amd import autotune
    path = os.path.join(os.path.dirname(os.path.abspath(autotune.__file__)), "tune_db.csv")
    db = autotune.load_db(path)
    assert os.path.isfile(path)
    try:
        st, n = _load_table(path)
        assert st == C.SUCCESS and n == len(db)
        for k, e in db.items():
            dt, lay, mnk = k.split(":")
            M, N, K = (int(x) for x in mnk.split("x"))
            p = _p(lay, int(dt), M, N, K)
            mt_m, mt_n, gm, wm, wn, mf, sk = e["tile"]
            # an entry with split-K is followed on the workspace path
            t = tl.query_tile_ws(p) if sk > 1 else tl.query_tile(p)
            assert (t.mt_m, t.mt_n, t.mfma_mn) == (mt_m, mt_n, mf or 32), k
            if sk > 1:
                assert t.split_k == sk, k
            if wm:
                assert (t.waves_m, t.waves_n) == (wm, wn), k
            if gm:
                assert t.group_m == gm, k
    finally:
        C.check(C.lib().ffgemm_tune_db_load(b"", None))


def _support_matrix():
    import json
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "support_matrix.json")
    return json.load(open(path))


def test_support_matrix_fixture_matches_reference_script():
    """The committed support table holds the 32 records of test/gemm_test.sh:
    5-73 (r, s, x, y in {1, 2}, u in {16, 32}) on M=768, N=640, K=256, with the
    reference's Y/N verdicts: every record for fp16, 22 of 32 for fp32."""
    d = _support_matrix()["dtypes"]
    want = {(r, s, x, y, u) for r in (1, 2) for s in (1, 2) for x in (1, 2) for y in (1, 2) for u in (16, 32)}
    for dt, n_yes in (("1", 22), ("2", 32)):
        assert d[dt]["problem"] == {"M": 768, "N": 640, "K": 256}
        cfgs = d[dt]["configs"]
        assert {(c["r"], c["s"], c["x"], c["y"], c["u"]) for c in cfgs} == want
        assert sum(c["support"] == "Y" for c in cfgs) == n_yes
        assert all(c["reason"] for c in cfgs if c["support"] == "N")


@pytest.mark.parametrize("dtype", [C.F32, C.F16, C.BF16])
def test_tile_from_reference_runs_every_support_matrix_record(dtype):
    """Every record of the reference's support matrix maps to a tile record
    that the library accepts for the matrix's problem (a superset of the
    reference's Y set: its N records run too), as SUCCESS when exact and as
    RTN_WARN naming what changed otherwise."""
    d = _support_matrix()["dtypes"]["1" if dtype == C.F32 else "2"]
    M, N, K = d["problem"]["M"], d["problem"]["N"], d["problem"]["K"]
    p = _p("TN", dtype, M, N, K)
    for c in d["configs"]:
        st, t, msg = tl.tile_from_reference(p, c["r"], c["s"], c["x"], c["y"], 32, c["u"], 0)
        assert st in (C.SUCCESS, C.WARN), c
        assert (st == C.WARN) == bool(msg)
        mt0, mt1 = 32 * c["r"] * c["x"], 32 * c["s"] * c["y"]
        assert t.mt_m >= mt0 and t.mt_n >= mt1, (c, t.mt_m, t.mt_n)
        q = tl.query_tile(p, t)
        assert (q.mt_m, q.mt_n, q.depth_u) == (t.mt_m, t.mt_n, t.depth_u)


def test_tile_from_reference_exact_and_nearest():
    p = _p("TN", C.BF16, 4096, 4096, 4096)
    # 256 x 256 with 2 x 4 waves: the eight-wave engine, exactly
    st, t, _ = tl.tile_from_reference(p, 4, 2, 2, 4, 32)
    assert st == C.SUCCESS and (t.mt_m, t.mt_n, t.waves_m, t.waves_n) == (256, 256, 2, 4)
    # 128 x 128 with 2 x 2 waves and DepthU 64 (the pipelined engine's 128-byte K rows, both operands
    # K-contiguous): exact
    st, t, _ = tl.tile_from_reference(p, 2, 2, 2, 2, 32, 64)
    assert st == C.SUCCESS and (t.mt_m, t.mt_n, t.depth_u) == (128, 128, 64)
    # ... with DepthU 32 or 16: the same tile, flagged
    for du in (32, 16):
        st, t, msg = tl.tile_from_reference(p, 2, 2, 2, 2, 32, du)
        assert st == C.WARN and (t.mt_m, t.mt_n, t.depth_u) == (128, 128, 64) and f"depth_u {du}" in msg
    # 128 x 64 has no kernel: the smallest covering tile, 128 x 128
    st, t, msg = tl.tile_from_reference(p, 2, 1, 2, 2, 32)
    assert st == C.WARN and (t.mt_m, t.mt_n) == (128, 128) and "MT128x64" in msg
    # 32 x 32 -> 64 x 64; 16 x 16 MFMAs at 64 x 64 (z=16, r=s=2, x=y=2): exact
    st, t, _ = tl.tile_from_reference(p, 1, 1, 1, 1, 32)
    assert st == C.WARN and (t.mt_m, t.mt_n) == (64, 64)
    st, t, _ = tl.tile_from_reference(p, 2, 2, 2, 2, 16)
    assert st == C.SUCCESS and (t.mt_m, t.mt_n, t.mfma_mn) == (64, 64, 16)
    # larger than any tile: 256 x 256
    st, t, _ = tl.tile_from_reference(p, 8, 8, 2, 2, 32)
    assert st == C.WARN and (t.mt_m, t.mt_n) == (256, 256)
    with pytest.raises(tl.FFGemmError):
        tl.tile_from_reference(p, 0, 1, 1, 1, 32)
    with pytest.raises(tl.FFGemmError):
        tl.tile_from_reference(p, 1, 1, 1, 1, 8)


@pytest.mark.parametrize("kind", ["f32", "bf16x3", "fp8", "bf16_to_f32", "f64"])
def test_grouped_plan_create_rejects_non_16bit_groups(kind):
    """The grouped kernels run bf16 -> bf16 and fp16 -> fp16 only: any other
    group is refused by ffgemm_grouped_plan_create itself (before it
    allocates), not later at ffgemm_grouped_plan_run, which a caller may be
    capturing into a graph.  No device memory is touched: the pointers are
    placeholders that are never dereferenced."""
    M = N = K = 256
    dt = {"f32": C.F32, "bf16x3": C.F32, "fp8": C.F8E4M3, "bf16_to_f32": C.BF16, "f64": C.F64}[kind]
    layout = "TN" if kind == "fp8" else "NN"
    p = _p(layout, dt, M, N, K, compute=C.COMPUTE_BF16X3 if kind == "bf16x3" else 0,
           dtype_out=C.F32 if kind == "bf16_to_f32" else 0)
    probs = (C.Problem * 2)(p, p)
    fake = [ctypes.c_void_p(0x100000 * (i + 1)) for i in range(3)]
    arr = lambda v: (ctypes.c_void_p * 2)(v, v)
    plan = ctypes.c_void_p(0)
    st = C.lib().ffgemm_grouped_plan_create(ctypes.byref(plan), 2, probs, arr(fake[0]), arr(fake[1]), None,
                                            arr(fake[2]), None)
    assert st == C.ERR
    assert plan.value is None
    msg = C.lib().ffgemm_last_error().decode()
    assert ("bf16 or fp16" in msg) or ("dtype" in msg), msg


def test_split_k_planning():
    """Split-K is planned only for the workspace path: ffgemm_query_tile (the
    ffgemm_run plan) never splits, ffgemm_query_tile_ws does for a problem with
    few output tiles, and the workspace it needs is S fp32 planes of
    roundup(M, 4) x N."""
    p = _p("TN", C.F32, 480, 1024, 1024)
    assert tl.query_tile(p).split_k == 1
    t = tl.query_tile_ws(p)
    assert t.split_k >= 2
    ldw = (480 + 3) // 4 * 4
    assert tl.workspace_size(p) == (t.split_k * ldw * 1024 * 4 + 255) // 256 * 256
    # plenty of tiles: no split, no workspace
    q = _p("TN", C.BF16, 4096, 4096, 4096)
    assert tl.query_tile_ws(q).split_k == 1 and tl.workspace_size(q) == 0
    # a caller's record with split_k: followed on the workspace path
    t = tl.query_tile_ws(q, (128, 128, 0, 0, 0, 0, 3))
    assert (t.mt_m, t.mt_n, t.split_k) == (128, 128, 3)


def test_tail_split_planning():
    """With no tile on the workspace path, a problem whose 256 x 256 tiles fill
    whole passes over the CUs plus a short last pass asks for the remainder's
    split-K planes (the data-parallel part needs none); one that fills whole
    passes, or whose last pass is long, asks for nothing.  ffgemm_query_tile_ws
    reports the single-launch record, and a caller's tile (or a tuning-table
    entry) runs as given."""
    cus = C.lib().ffgemm_device_cu_count()
    if cus != 256:
        pytest.skip(f"planning figures assume 256 CUs ({cus} here)")
    for M, N, K, want in [(6144, 6144, 6144, True), (4608, 4608, 4608, True), (7168, 7168, 7168, True),
                          (69120, 256, 2048, True), (4096, 4096, 4096, False), (8192, 8192, 8192, False),
                          (5120, 5120, 5120, False), (6144, 6144, 2048, False)]:
        p = _p("TN", C.BF16, M, N, K)
        assert (tl.workspace_size(p) > 0) == want, (M, N, K)
        assert tl.query_tile_ws(p).split_k == 1
    # 4608^3: 18 x 18 tiles, 252 in one pass, the remaining 4 tile columns (72 tiles) in 3 slices
    p = _p("TN", C.BF16, 4608, 4608, 4608)
    assert tl.workspace_size(p) == 3 * 4608 * 1024 * 4


@pytest.mark.parametrize("case", ["run", "batch", "x3", "short_k", "negative"])
def test_split_k_rejections(case):
    """A split the problem cannot take is refused with RTN_ERR: through
    ffgemm_run (no workspace), with batch > 1, for bf16x3, with K below two
    minimal slices, or a negative count."""
    fake = ctypes.c_void_p(0x100000)
    t = C.Tile()
    t.mt_m, t.mt_n, t.split_k = 128, 128, (-1 if case == "negative" else 4)
    if case == "batch":
        p = tl.make_problem("TN", C.BF16, 512, 512, 4096, 4096, 4096, 512, 512, batch=2, bsA=512 * 4096,
                            bsB=512 * 4096, bsD=512 * 512)
    elif case == "x3":
        p = _p("NN", C.F32, 512, 512, 4096, compute=C.COMPUTE_BF16X3)
    elif case == "short_k":
        p = _p("TN", C.BF16, 512, 512, 256)
    else:
        p = _p("TN", C.BF16, 512, 512, 4096)
    with pytest.raises(C.FFGemmError, match="RTN_ERR"):
        tl.query_tile_ws(p, t) if case != "run" else tl.query_tile(p, t)
    if case == "run":
        st = C.lib().ffgemm_run(ctypes.byref(p), fake, fake, None, fake, ctypes.byref(t), None)
        assert st == C.ERR and "workspace" in C.lib().ffgemm_last_error().decode()


def test_repack_planning():
    """The workspace path repacks a bf16/fp16 NN/NT/TT problem's M/N-contiguous
    operands K-contiguous (rows x K bf16 per operand at the front of the
    workspace) and runs it as TN when the cost model puts that ahead: problems
    of whole 256 x 256 passes (4096^3, 8192^3), not small ones, never TN; the
    no-workspace plan (ffgemm_query_tile) never repacks; a caller's tile, or
    repack 1, keeps the native layout; repack 2 forces it at any size."""
    cus = C.lib().ffgemm_device_cu_count()
    if cus != 256:
        pytest.skip(f"planning figures assume 256 CUs ({cus} here)")
    S = 8192
    for layout, nbytes in [("NN", S * S * 2), ("TT", S * S * 2), ("NT", 2 * S * S * 2), ("TN", 0)]:
        p = _p(layout, C.BF16, S, S, S)
        assert tl.workspace_size(p) == nbytes, layout
        assert tl.query_tile_ws(p).repack == (2 if nbytes else 0), layout
        assert tl.query_tile(p).repack == 0
        assert tl.workspace_size(p, (256, 256)) == 0
        assert tl.workspace_size(p, (0, 0, 0, 0, 0, 0, 0, 1)) == 0
    for layout in ("NN", "NT", "TT"):
        n4 = 4096 * 4096 * 2 * (2 if layout == "NT" else 1)
        assert tl.workspace_size(_p(layout, C.BF16, 4096, 4096, 4096)) == n4, layout
        assert tl.workspace_size(_p(layout, C.F16, 1024, 1024, 1024)) == 0, layout
        # 256 x 256 tiles that leave CUs idle: every layout repacks, fp16 too
        assert tl.query_tile_ws(_p(layout, C.F16, 3072, 3072, 3072)).repack == 2, layout
    # 128 x 128 tiles: NT repacks (its native engine is slowest there); TT at
    # 1024 x 4096 x 4096 and NN / TT at 2048^3 stay native
    assert tl.query_tile_ws(_p("NT", C.BF16, 2048, 2048, 8192)).repack == 2
    assert tl.query_tile_ws(_p("TT", C.BF16, 1024, 4096, 4096)).repack == 0
    assert tl.query_tile_ws(_p("NN", C.BF16, 2048, 2048, 2048)).repack == 0
    assert tl.query_tile_ws(_p("TT", C.BF16, 2048, 2048, 2048)).repack == 0
    # forced: each repacked operand's bytes rounded to 256, then the TN problem's own plan
    p = _p("NT", C.F16, 1000, 520, 256)
    t = tl.query_tile_ws(p, tl.make_tile((0, 0, 0, 0, 0, 0, 0, 2)))
    tn = tl.query_tile_ws(_p("TN", C.F16, 1000, 520, 256))
    assert t.repack == 2 and (t.mt_m, t.mt_n, t.waves_m, t.waves_n) == (tn.mt_m, tn.mt_n, tn.waves_m, tn.waves_n)
    assert tl.workspace_size(p, (0, 0, 0, 0, 0, 0, 0, 2)) == (1000 * 256 * 2 + 255) // 256 * 256 + 520 * 256 * 2
    # a forced repack whose TN problem splits: the partial planes follow the copies
    p = _p("NN", C.BF16, 520, 480, 1024)
    rec = (128, 128, 0, 0, 0, 0, 2, 2)
    t = tl.query_tile_ws(p, tl.make_tile(rec))
    assert (t.mt_m, t.mt_n, t.split_k, t.repack) == (128, 128, 2, 2)
    assert tl.workspace_size(p, rec) == 520 * 1024 * 2 + 2 * 520 * 480 * 4


@pytest.mark.parametrize("case", ["f32", "x3", "k_odd", "off_grid", "value", "no_ws"])
def test_repack_rejections(case):
    """A forced repack the problem cannot take is refused with RTN_ERR, before
    anything runs: fp32 and bf16x3 operands, K not a multiple of 8, strides off
    the 16-byte grid, a repack value outside 0-2, and ffgemm_run_ws without
    the workspace it needs."""
    fake = ctypes.c_void_p(0x100000)
    rec = (0, 0, 0, 0, 0, 0, 0, 3 if case == "value" else 2)
    if case == "f32":
        p = _p("NN", C.F32, 512, 512, 512)
    elif case == "x3":
        p = _p("NN", C.F32, 512, 512, 512, compute=C.COMPUTE_BF16X3)
    elif case == "k_odd":
        p = _p("NT", C.BF16, 512, 512, 516)
    elif case == "off_grid":
        p = _p("NN", C.BF16, 512, 512, 512, lda=515)
    else:
        p = _p("NN", C.BF16, 512, 512, 512)
    if case == "x3":
        # bf16x3 planes take the workspace path of their own: the record's
        # repack is not consulted there, and its plain run refuses nothing
        assert tl.workspace_size(p, rec) > 0
        return
    if case == "no_ws":
        t = tl.make_tile(rec)
        st = C.lib().ffgemm_run_ws(ctypes.byref(p), fake, fake, None, fake, ctypes.byref(t), fake, 256, None)
        assert st == C.ERR and "workspace" in C.lib().ffgemm_last_error().decode()
        return
    with pytest.raises(C.FFGemmError, match="RTN_ERR"):
        tl.workspace_size(p, rec)
    with pytest.raises(C.FFGemmError, match="RTN_ERR"):
        tl.query_tile_ws(p, tl.make_tile(rec))
