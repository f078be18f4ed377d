// C ABI of libffgemm.so (include/ffgemm.h): problem validation, tile choice
// and kernel launch.  Replaces GemmMfmaAsmSolution::generateKernel
// (src/gemmMfma.cpp:24-185), GemmMfmaKernelWriter::checkKernelParameters
// (src/gemmMfmaKernelWriter.h:321-405) and StreamHip::Launch
// (ff_gpu_src/ff_hip_runtime.cpp:278-320) for the gemmMfma path.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <tuple>
#include <queue>
#include <string>
#include <vector>

#include "../../include/ffgemm.h"
#include "aux_kernels.hpp"
#include "grouped.hpp"

using namespace ffgemm;

namespace {

thread_local std::string g_last_error;

int fail(int status, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return status;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(FFGEMM_FATAL, "%s: %s", what, hipGetErrorString(e));
}

std::mutex g_dev_mu;
int g_cu_count[64] = {0};

int cu_count() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (g_cu_count[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        g_cu_count[dev] = n;
    }
    return g_cu_count[dev];
}

struct Tile {
    int m, n;
};
const Tile kTiles[] = {{256, 256}, {256, 128}, {128, 256}, {128, 128}, {64, 64}};

bool tile_supported(int m, int n) {
    for (const Tile& t : kTiles)
        if (t.m == m && t.n == n) return true;
    return false;
}

// Resident workgroups per CU for a 16-bit tile (VGPR/AGPR and LDS limited),
// as the pipelined engine's configuration declares them.
int tile_occupancy(int m, int n, int layout) {
    if (m == 64 && n == 64) return 4;
    const bool ak = layout >= FFGEMM_TN, bk = (layout & 1) == 0;
#define FFGEMM_OCC(TM_, TN_)                                                               \
    if (m == TM_ && n == TN_) {                                                            \
        if (ak && bk) return v2::Cfg<TM_, TN_, true, true>::OCC;                           \
        if (ak) return v2::Cfg<TM_, TN_, true, false>::OCC;                                \
        if (bk) return v2::Cfg<TM_, TN_, false, true>::OCC;                                \
        return v2::Cfg<TM_, TN_, false, false>::OCC;                                       \
    }
    FFGEMM_OCC(256, 128) FFGEMM_OCC(128, 256) FFGEMM_OCC(128, 128)
#undef FFGEMM_OCC
    return 1;
}

int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Largest tile whose grid still fills every CU; else the one with most tiles.
// The tile with the least estimated time: passes of tiles over the CUs times
// one tile's time, which is its area over the tile's relative MFMA efficiency
// (16-bit 4096^3 sweep on MI355X: 256x256 1.0, 256x128 and 128x256 0.82,
// 128x128 0.70, 64x64 0.51).  The last pass counts whole, so a shape with a
// few tiles more than the CUs takes the next larger tile rather than a
// second, nearly empty pass.  Ties go to the larger tile.
Tile pick_tile(int64_t M, int64_t N, int64_t batch) {
    const int64_t cus = cu_count();
    static const double eff[] = {1.0, 0.82, 0.82, 0.70, 0.51};  // kTiles order
    Tile best = kTiles[0];
    double best_t = 0;
    for (int i = 0; i < 5; ++i) {
        const Tile& t = kTiles[i];
        const int64_t passes = cdiv(cdiv(M, t.m) * cdiv(N, t.n) * batch, cus);
        const double est = (double)passes * t.m * t.n / eff[i];
        if (i == 0 || est < best_t) { best = t; best_t = est; }
    }
    return best;
}

// ---------------------------------------------------------------------------
// Split-K planning.  A problem with too few output tiles for the CUs can cut
// K into S slices: one batched GEMM writes S fp32 partial planes to a
// workspace and splitk_reduce_kernel sums them into D.
// ---------------------------------------------------------------------------
// Shortest slice: four 128-byte K-tiles.
int64_t split_min_k(int mode) { return 4 * 128 / mode_elem_bytes(mode); }
// Slice length: ceil(K / S) rounded up to 128 elements, which keeps every
// slice's operand offset 16-byte aligned and whole K-tiles in all but the last.
int64_t split_slice_k(int64_t K, int S) { return cdiv(cdiv(K, S), 128) * 128; }
int64_t split_ldw(int64_t M) { return cdiv(M, 4) * 4; }

// Modelled time (µs) of a launch plan on MI355X: passes of workgroups over the
// CUs times one workgroup's time, i.e. its flop at 70 % of the mode's MFMA rate
// at 1.9 GHz divided by the tile's relative efficiency (as pick_tile), plus
// 4 µs of prologue and epilogue; for S > 1 also the reduction (S fp32 planes
// read, D written, C read, at 4 TB/s) and 4 µs for its launch.
double plan_cost(int64_t M, int64_t N, int64_t K, int mode, int osz, bool use_c, int ti, int S) {
    // relative efficiency per tile (kTiles order): 16-bit / fp8 from the
    // 4096^3 sweep; fp32 nearly flat (the phantom-shape sweep on MI355X put
    // its 64 x 64 and 128 x 128 tiles level with or ahead of 256 x 256)
    static const double eff16[] = {1.0, 0.82, 0.82, 0.70, 0.51};
    static const double eff32[] = {0.85, 0.9, 0.9, 0.95, 1.0};
    const double* eff = mode == kF32 ? eff32 : eff16;
    const double rate = (mode == kF8 ? 8192.0 : mode == kF32 ? 256.0 : 4096.0) * 1.9e3 * 0.7;  // flop/µs/CU
    const Tile& t = kTiles[ti];
    const int64_t ks = S > 1 ? split_slice_k(K, S) : K;
    const int64_t slices = S > 1 ? cdiv(K, ks) : 1;
    const int64_t passes = cdiv(cdiv(M, t.m) * cdiv(N, t.n) * slices, cu_count());
    double c = (double)passes * (2.0 * t.m * t.n * (double)ks / (rate * eff[ti]) + 4.0);
    if (slices > 1) c += ((double)slices * M * N * 4 + (double)M * N * osz * (use_c ? 2 : 1)) / 4e6 + 4.0;
    return c;
}

struct SplitChoice {
    Tile t;
    int S;
};
// The split with the least modelled time against the no-split plan `base`
// (the tile pick_tile or the caller chose); a split is taken only when the
// model puts it at least 15 % ahead.  With `fixed` the tile stays the caller's.
SplitChoice pick_split(int64_t M, int64_t N, int64_t K, int mode, int osz, bool use_c, const Tile* fixed, Tile base) {
    int bi = 0;
    for (int i = 0; i < 5; ++i)
        if (kTiles[i].m == base.m && kTiles[i].n == base.n) bi = i;
    const double none = plan_cost(M, N, K, mode, osz, use_c, bi, 1);
    SplitChoice best{base, 1};
    double best_c = none * 0.85;
    static const int cand[] = {2, 3, 4, 6, 8, 12, 16};
    for (int i = 0; i < 5; ++i) {
        if (fixed && i != bi) continue;
        for (int S : cand) {
            if (K / S < split_min_k(mode)) break;
            const double c = plan_cost(M, N, K, mode, osz, use_c, i, S);
            if (c < best_c) { best_c = c; best = SplitChoice{kTiles[i], S}; }
        }
    }
    return best;
}

int default_group_m(int64_t tiles_m, int64_t tiles_n) {
    int64_t per_xcd = std::max<int64_t>(1, tiles_m * tiles_n / 8);
    int g = 1;
    while ((int64_t)(g * 2) * (g * 2) <= per_xcd && g * 2 <= tiles_m) g *= 2;
    return g;
}

int out_kind_of(int dtype, int dtype_out) {
    int o = dtype_out ? dtype_out : (dtype == FFGEMM_F8E4M3 ? FFGEMM_F32 : dtype);
    if (o == FFGEMM_F32) return kOutF32;
    if (o == FFGEMM_F16) return kOutF16;
    if (o == FFGEMM_BF16) return kOutBF16;
    return -1;
}

int elem_bytes(int dtype) {
    switch (dtype) {
        case FFGEMM_F32: return 4;
        case FFGEMM_F16: case FFGEMM_BF16: return 2;
        case FFGEMM_F8E4M3: return 1;
        case FFGEMM_F64: return 8;
    }
    return 0;
}

struct Launch {
    int mode;
    int out;
    int layout;
    Tile tile;
    int waves;  // 0: the default engine for the tile; 4 or 8: that wave count
    int mfma;   // 0: 32 x 32 MFMAs; 16: the 16 x 16 x 32 form
    TileArgs args;
    int64_t grid;
    int split;   // K slices (1: none)
    int64_t ks;  // K per slice
    bool any;    // the general-stride form (gemm_kernel_any)
};

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// ---------------------------------------------------------------------------
// Tuning table: the autotuner's best tile per (dtype, layout, M, N, K), kept
// as CSV lines `dtype,layout,M,N,K,mt_m,mt_n,group_m,waves_m,waves_n,mfma_mn,tflops`
// (layout as NN/NT/TN/TT; a line of 11 fields has no mfma_mn: 32 x 32).  The successor of the best-of selection in
// test/perf_test.py:58-167: problems run with no tile take the table's tile.
// ---------------------------------------------------------------------------
using TuneKey = std::tuple<int, int, int64_t, int64_t, int64_t>;
std::mutex g_tune_mu;
std::map<TuneKey, ffgemm_tile> g_tune;
bool g_tune_loaded = false;

// FFGEMM_TUNE_DB, else tune_db.csv in the directory above the library's own
// (the package directory when the library sits in <package>/lib).
std::string default_tune_path() {
    if (const char* e = getenv("FFGEMM_TUNE_DB")) return e;
    Dl_info info;
    if (dladdr((void*)&default_tune_path, &info) && info.dli_fname) {
        std::string so = info.dli_fname;
        const size_t a = so.find_last_of('/');
        if (a != std::string::npos) {
            const std::string dir = so.substr(0, a);
            const size_t b = dir.find_last_of('/');
            if (b != std::string::npos) return dir.substr(0, b) + "/tune_db.csv";
        }
    }
    return "";
}

// Parses the table at `path` into `out`; returns the number of entries, or -1
// when the file exists but a line is malformed.  A missing file is an empty table.
int read_tune_table(const std::string& path, std::map<TuneKey, ffgemm_tile>& out) {
    out.clear();
    if (path.empty()) return 0;
    std::ifstream f(path);
    if (!f) return 0;
    std::string line;
    int n = 0;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#' || line.compare(0, 5, "dtype") == 0) continue;
        std::vector<std::string> v;
        std::stringstream ss(line);
        std::string cell;
        while (std::getline(ss, cell, ',')) v.push_back(cell);
        if (v.size() < 10) return -1;
        int layout = -1;
        const char* names[4] = {"NN", "NT", "TN", "TT"};
        for (int i = 0; i < 4; ++i)
            if (v[1] == names[i]) layout = i;
        char* end = nullptr;
        if (layout < 0) layout = (int)strtol(v[1].c_str(), &end, 10);
        long long x[10];
        for (int i = 0; i < 10; ++i) {
            if (i == 1) continue;
            x[i] = strtoll(v[i].c_str(), &end, 10);
            if (end == v[i].c_str()) return -1;
        }
        if (layout < 0 || layout > 3) return -1;
        ffgemm_tile t;
        memset(&t, 0, sizeof(t));
        t.mt_m = (int)x[5]; t.mt_n = (int)x[6]; t.group_m = (int)x[7];
        t.waves_m = (int)x[8]; t.waves_n = (int)x[9];
        if (v.size() >= 12) {
            t.mfma_mn = (int)strtol(v[10].c_str(), &end, 10);
            if (end == v[10].c_str()) return -1;
        }
        if (v.size() >= 13) {  // split_k before tflops
            t.split_k = (int)strtol(v[11].c_str(), &end, 10);
            if (end == v[11].c_str() || t.split_k < 0) return -1;
        }
        out[TuneKey((int)x[0], layout, x[2], x[3], x[4])] = t;
        ++n;
    }
    return n;
}

bool tuned_tile(const ffgemm_problem* p, ffgemm_tile* t) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!g_tune_loaded) {
        if (read_tune_table(default_tune_path(), g_tune) < 0) g_tune.clear();
        g_tune_loaded = true;
    }
    if (g_tune.empty()) return false;
    auto it = g_tune.find(TuneKey(p->dtype, p->layout, p->M, p->N, p->K));
    if (it == g_tune.end()) return false;
    *t = it->second;
    return true;
}

// Checks the problem against what the kernels support and fills `L`.
// allow_split: the caller has a workspace for split-K partials (L.split may
// then be > 1).
int prepare_with(const ffgemm_problem* p, const void* A, const void* B, const void* C, void* D,
                 const ffgemm_tile* tile, Launch& L, bool allow_split = false) {
    if (!p) return fail(FFGEMM_ERR, "null problem");
    if (p->layout < 0 || p->layout > 3) return fail(FFGEMM_ERR, "bad layout %d", p->layout);
    const int64_t M = p->M, N = p->N, K = p->K;
    const int64_t batch = p->batch <= 0 ? 1 : p->batch;
    if (M < 0 || N < 0 || K < 0) return fail(FFGEMM_ERR, "negative size");
    const int64_t lim = (int64_t)1 << 31;
    if (M >= lim || N >= lim || K >= lim || batch >= lim) return fail(FFGEMM_ERR, "size exceeds 2^31");
    const bool ta = p->layout >= 2, tb = (p->layout & 1) != 0;
    int mode;
    switch (p->dtype) {
        case FFGEMM_BF16: mode = kBF16; break;
        case FFGEMM_F16: mode = kF16; break;
        case FFGEMM_F32: mode = (p->compute == FFGEMM_COMPUTE_BF16X3) ? kX3 : kF32; break;
        case FFGEMM_F8E4M3: mode = kF8; break;
        default: return fail(FFGEMM_ERR, "unsupported dtype %d", p->dtype);
    }
    if (p->compute == FFGEMM_COMPUTE_BF16X3 && p->dtype != FFGEMM_F32)
        return fail(FFGEMM_ERR, "bf16x3 compute needs fp32 operands");
    const int out = out_kind_of(p->dtype, p->dtype_out);
    if (out < 0) return fail(FFGEMM_ERR, "unsupported output dtype %d", p->dtype_out);
    if ((mode == kBF16 && out == kOutF16) || (mode == kF16 && out == kOutBF16) ||
        ((mode == kF32 || mode == kX3) && out != kOutF32) || (mode == kF8 && out == kOutF16))
        return fail(FFGEMM_ERR, "unsupported input/output dtype pair %d/%d", p->dtype, p->dtype_out);
    if (mode == kF8 && p->layout != FFGEMM_TN) return fail(FFGEMM_ERR, "fp8 supports the TN layout only");

    // leading dimensions (column-major)
    const int64_t rowsA = ta ? K : M, rowsB = tb ? N : K;
    if (p->lda < std::max<int64_t>(1, rowsA)) return fail(FFGEMM_ERR, "lda %lld < %lld", (long long)p->lda, (long long)rowsA);
    if (p->ldb < std::max<int64_t>(1, rowsB)) return fail(FFGEMM_ERR, "ldb %lld < %lld", (long long)p->ldb, (long long)rowsB);
    if (p->ldd < std::max<int64_t>(1, M)) return fail(FFGEMM_ERR, "ldd %lld < M", (long long)p->ldd);
    const bool use_c = p->beta != 0.0f;
    if (use_c && p->ldc < std::max<int64_t>(1, M)) return fail(FFGEMM_ERR, "ldc %lld < M", (long long)p->ldc);
    if (p->lda >= lim || p->ldb >= lim || p->ldc >= lim || p->ldd >= lim) return fail(FFGEMM_ERR, "ld exceeds 2^31");

    if (M == 0 || N == 0) {  // nothing to compute or store
        L.mode = mode; L.out = out; L.layout = p->layout; L.tile = kTiles[4]; L.grid = 0; L.waves = 0;
        L.split = 1; L.ks = K; L.any = false;
        memset(&L.args, 0, sizeof(L.args));
        return FFGEMM_SUCCESS;
    }
    const int es = mode_elem_bytes(mode);
    const int ve = 16 / es;          // elements per 16-byte vector
    const int os = out == kOutF32 ? 4 : 2;
    if (M && N && (!A || !B || !D)) return fail(FFGEMM_ERR, "null matrix pointer");
    if (use_c && M && N && !C) return fail(FFGEMM_ERR, "beta != 0 but C is null");
    // The engines move 16-byte vectors: leading dimensions, batch strides and
    // pointers at 16-byte multiples, K a multiple of the vector for a
    // K-contiguous operand.  Any other problem runs the general-stride form
    // (element-wise loads and stores), which needs only element alignment.
    const bool vec_ok = (p->lda * es) % 16 == 0 && (p->ldb * es) % 16 == 0 && (p->ldd * os) % 16 == 0 &&
                        (!use_c || (p->ldc * os) % 16 == 0) && (!(ta || !tb) || K % ve == 0) &&
                        aligned16(A) && aligned16(B) && aligned16(D) && (!use_c || aligned16(C)) &&
                        (batch <= 1 || ((p->bsA * es) % 16 == 0 && (p->bsB * es) % 16 == 0 &&
                                        (p->bsD * os) % 16 == 0 && (!use_c || (p->bsC * os) % 16 == 0)));
    const bool any = !vec_ok;
    if (any && (((uintptr_t)A % es) || ((uintptr_t)B % es) || ((uintptr_t)D % os) || (use_c && ((uintptr_t)C % os))))
        return fail(FFGEMM_ERR, "matrix pointers must be aligned to their element size");
    // Buffer offsets are 31-bit.  A K-contiguous operand is addressed from each
    // tile's first row (at most 256 rows, and a margin row), so any number of
    // rows fits; an M/N-contiguous operand is addressed whole, all K rows of it.
    const int64_t exA = (ta ? 257 : K + 256) * p->lda * es;
    const int64_t exB = (tb ? K + 256 : 257) * p->ldb * es;
    if (exA >= lim || exB >= lim)
        return fail(FFGEMM_ERR, "operand exceeds 2 GiB per batch (K-contiguous operands: 257 rows of ld)");

    Tile t = pick_tile(M, N, batch);
    // the in-register bf16x3 split does not fit 256 x 256 without spilling
    // (8192^3 on MI355X: 165-203 TFLOP/s there, 284-289 at 128 x 256)
    if (mode == kX3 && t.m == 256 && t.n == 256) t = Tile{128, 256};
    // the general-stride form: 128 x 128 when that still gives a tile per CU
    // (fp8 always: its 64 x 64 form spills)
    if (any) t = (mode == kF8 || cdiv(M, 128) * cdiv(N, 128) * batch >= cu_count()) ? Tile{128, 128} : Tile{64, 64};
    int group_m = 0;
    int waves = 0;
    int mfma = 0;
    if (tile && tile->mt_m && any) {
        // one engine, two square tiles, 32 x 32 MFMAs, 2 x 2 waves, two stages
        const int du = mode_depth_u(mode);
        if (!(((tile->mt_m == 64 && mode != kF8) || tile->mt_m == 128) && tile->mt_n == tile->mt_m) ||
            (tile->waves_m && (tile->waves_m != 2 || tile->waves_n != 2)) ||
            (tile->mfma_mn && tile->mfma_mn != 32) || (tile->lds_buffers && tile->lds_buffers != 2) ||
            (tile->depth_u && tile->depth_u != du))
            return fail(FFGEMM_ERR,
                        "leading dimensions, batch strides, K or pointers off the 16-byte grid run the "
                        "general-stride kernel: MT64x64 (not fp8) or MT128x128, 2x2 waves, 32x32 MFMAs, 2 LDS stages, "
                        "depth_u %d (MT%dx%d asked)", du, tile->mt_m, tile->mt_n);
        t = Tile{tile->mt_m, tile->mt_m};
        group_m = tile->group_m;
    } else if (tile && tile->mt_m) {
        if (!tile_supported(tile->mt_m, tile->mt_n))
            return fail(FFGEMM_ERR, "unsupported macro tile %dx%d", tile->mt_m, tile->mt_n);
        // waves_m x waves_n picks the wave tiling (0: the tile's default engine;
        // 2 x 4: the eight-wave engine; 2 x 2: a four-wave engine); the other
        // knobs follow from that engine: zero, or the engine's own value (what
        // ffgemm_query_tile reports)
        if (tile->waves_m || tile->waves_n) {
            waves = tile->waves_m * tile->waves_n;
            if (tile->waves_m != 2 || (waves != 4 && waves != 8) ||
                engine_for(mode, ta, !tb, tile->mt_m, tile->mt_n, waves) == kEngNone)
                return fail(FFGEMM_ERR, "no %dx%d-wave kernel for MT%dx%d with this dtype and layout",
                            tile->waves_m, tile->waves_n, tile->mt_m, tile->mt_n);
        }
        // mfma_mn: 0 or 32 (32 x 32 MFMAs), or 16 (16 x 16 x 32 for bf16 / fp16,
        // 16 x 16 x 4 for fp32, at 64 x 64 and 128 x 128)
        if (tile->mfma_mn == 16) {
            if (engine_for(mode, ta, !tb, tile->mt_m, tile->mt_n, waves, 16) == kEngNone)
                return fail(FFGEMM_ERR, "no 16x16 MFMA kernel for MT%dx%d with this dtype (bf16/fp16/fp32 at 64x64, "
                            "128x128)", tile->mt_m, tile->mt_n);
            mfma = 16;
        } else if (tile->mfma_mn && tile->mfma_mn != 32) {
            return fail(FFGEMM_ERR, "unsupported mfma_mn %d (16 or 32)", tile->mfma_mn);
        }
        const EngineGeometry geo = engine_geometry(mode, ta, !tb, tile->mt_m, tile->mt_n, waves, mfma);
        if ((tile->lds_buffers && tile->lds_buffers != geo.lds_buffers) ||
            (tile->depth_u && tile->depth_u != geo.depth_u))
            return fail(FFGEMM_ERR,
                        "unsupported tile knobs for MT%dx%d (waves %dx%d, %d LDS stages, depth_u %d)",
                        tile->mt_m, tile->mt_n, geo.waves_m, geo.waves_n, geo.lds_buffers, geo.depth_u);
        t = Tile{tile->mt_m, tile->mt_n};
        group_m = tile->group_m;
    }
    // split-K: only with a workspace for the fp32 partial planes, for single
    // problems with an fp32 partial form (not bf16x3), each slice at least
    // four K-tiles long
    const bool split_ok = allow_split && batch == 1 && mode != kX3 && !any && K >= 2 * split_min_k(mode);
    int split = 1;
    if (tile && tile->split_k < 0) return fail(FFGEMM_ERR, "split_k %d < 0", tile->split_k);
    if (tile && tile->mt_m) {
        // a caller's tile splits only when it says so
        if (tile->split_k > 0) split = tile->split_k;
    } else if (split_ok) {
        const SplitChoice c = pick_split(M, N, K, mode, out == kOutF32 ? 4 : 2, use_c, nullptr, t);
        if (c.S > 1) { t = c.t; split = c.S; }
    }
    int64_t ks = K;
    if (split > 1) {
        if (!split_ok)
            return fail(FFGEMM_ERR, "split_k %d needs ffgemm_run_ws with a workspace, batch 1, not bf16x3, "
                        "16-byte-aligned strides and pointers, and K >= %lld", split, (long long)(2 * split_min_k(mode)));
        ks = split_slice_k(K, split);
        split = (int)cdiv(K, ks);
        if (split < 2) ks = K;
    }
    L.split = split;
    L.ks = ks;
    const int64_t tiles_m = cdiv(M, t.m), tiles_n = cdiv(N, t.n);
    const int64_t grid = tiles_m * tiles_n * batch;
    if (grid >= lim) return fail(FFGEMM_ERR, "too many tiles");

    TileArgs& a = L.args;
    memset(&a, 0, sizeof(a));
    a.A = A; a.B = B; a.C = use_c ? C : nullptr; a.D = D;
    a.bsA = p->bsA; a.bsB = p->bsB; a.bsC = p->bsC; a.bsD = p->bsD;
    a.M = (int32_t)M; a.N = (int32_t)N; a.K = (int32_t)K;
    a.lda = (int32_t)p->lda; a.ldb = (int32_t)p->ldb; a.ldc = (int32_t)(use_c ? p->ldc : p->ldd); a.ldd = (int32_t)p->ldd;
    a.alpha = p->alpha; a.beta = p->beta;
    a.scaleA = p->scaleA; a.scaleB = p->scaleB;
    a.tiles_m = (int32_t)tiles_m; a.tiles_n = (int32_t)tiles_n; a.batch = (int32_t)batch;
    a.group_m = group_m > 0 ? group_m : default_group_m(tiles_m, tiles_n);
    a.flags = (use_c ? kFlagBeta : 0u) | (p->round == FFGEMM_ROUND_TRUNCATE ? kFlagTrunc : 0u) |
              (grid <= cu_count() ? kFlagOneRound : 0u);
    L.mode = mode; L.out = out; L.layout = p->layout; L.tile = t; L.grid = grid; L.waves = waves; L.mfma = mfma;
    L.any = any;
    return FFGEMM_SUCCESS;
}

// An M/N-contiguous operand is addressed whole, all K rows of its ld, with
// 31-bit buffer offsets.  When that exceeds 2 GiB the problem runs in K chunks
// (a multiple of 64 that fits), accumulating into D; 0 when no chunking is
// needed (or possible: then prepare reports the size).
int64_t mn_k_chunk(const ffgemm_problem* p) {
    if (!p || p->K <= 0 || p->M <= 0 || p->N <= 0) return 0;
    const bool ta = p->layout >= FFGEMM_TN, tb = (p->layout & 1) != 0;
    const int es = p->dtype == FFGEMM_F32 ? 4 : p->dtype == FFGEMM_F8E4M3 ? 1 : 2;
    int64_t ld = 0;
    if (!ta) ld = std::max<int64_t>(ld, p->lda);  // A M-contiguous: K rows of lda
    if (tb) ld = std::max<int64_t>(ld, p->ldb);   // B N-contiguous: K rows of ldb
    const int64_t lim = (int64_t)1 << 31;
    if (ld <= 0 || (p->K + 256) * ld * es < lim) return 0;
    const int64_t kc = ((lim - 1) / (ld * es) - 256) / 64 * 64;
    return kc >= 64 ? kc : 0;
}

// prepare_with, where a problem given no tile first tries the tuning table's
// tile for it (single problems in the default compute and output type); an
// entry that does not fit the problem is passed over.
int prepare(const ffgemm_problem* p, const void* A, const void* B, const void* C, void* D,
            const ffgemm_tile* tile, Launch& L, bool allow_split = false) {
    ffgemm_tile tuned;
    if (p && (!tile || !tile->mt_m) && p->batch <= 1 && p->compute == FFGEMM_COMPUTE_DEFAULT &&
        (p->dtype_out == 0 || p->dtype_out == p->dtype || p->dtype == FFGEMM_F8E4M3) && tuned_tile(p, &tuned)) {
        // an entry with split-K runs as recorded with a workspace, else unsplit
        if (!allow_split && tuned.split_k > 1) tuned.split_k = 1;
        if (prepare_with(p, A, B, C, D, &tuned, L, allow_split) == FFGEMM_SUCCESS) return FFGEMM_SUCCESS;
    }
    return prepare_with(p, A, B, C, D, tile, L, allow_split);
}

// The engine geometry of a plan, as ffgemm_query_tile reports it.
EngineGeometry launch_geometry(const Launch& L) {
    if (L.any) return {2, 2, mode_depth_u(L.mode), 2};
    return engine_geometry(L.mode, L.layout >= FFGEMM_TN, (L.layout & 1) == 0, L.tile.m, L.tile.n, L.waves, L.mfma);
}

int launch(const Launch& L, hipStream_t s) {
    int r = -1;
    if (L.any) {
        r = (L.mode == kBF16 || L.mode == kF16) ? launch_any16(L.mode, L.layout, L.out, L.tile.m, L.args, (int)L.grid, s)
                                                : launch_any32(L.mode, L.layout, L.out, L.tile.m, L.args, (int)L.grid, s);
    } else switch (L.mode) {
        case kBF16: r = launch_bf16(L.layout, L.out, L.tile.m, L.tile.n, L.waves, L.mfma, L.args, (int)L.grid, s); break;
        case kF16: r = launch_f16(L.layout, L.out, L.tile.m, L.tile.n, L.waves, L.mfma, L.args, (int)L.grid, s); break;
        case kF32: r = launch_f32(L.layout, L.out, L.tile.m, L.tile.n, L.waves, L.mfma, L.args, (int)L.grid, s); break;
        case kX3: r = launch_x3(L.layout, L.out, L.tile.m, L.tile.n, L.waves, L.mfma, L.args, (int)L.grid, s); break;
        case kF8: r = launch_f8(L.layout, L.out, L.tile.m, L.tile.n, L.waves, L.mfma, L.args, (int)L.grid, s); break;
    }
    if (r != 0) return fail(FFGEMM_ERR, "no kernel for mode %d out %d layout %d tile %dx%d", L.mode, L.out,
                            L.layout, L.tile.m, L.tile.n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return FFGEMM_SUCCESS;
}

int run_f64(const ffgemm_problem* p, const void* A, const void* B, const void* C, void* D, hipStream_t s) {
    const int64_t batch = p->batch <= 0 ? 1 : p->batch;
    if (p->M < 0 || p->N < 0 || p->K < 0) return fail(FFGEMM_ERR, "negative size");
    if (p->layout < 0 || p->layout > 3) return fail(FFGEMM_ERR, "bad layout %d", p->layout);
    const bool ta = p->layout >= 2, tb = (p->layout & 1) != 0;
    if (p->lda < (ta ? p->K : p->M) || p->ldb < (tb ? p->N : p->K) || p->ldd < p->M ||
        (p->beta != 0.0f && p->ldc < p->M))
        return fail(FFGEMM_ERR, "leading dimension too small");
    if (p->M == 0 || p->N == 0) return FFGEMM_SUCCESS;
    if (batch > 65535) return fail(FFGEMM_ERR, "batch too large for the fp64 kernel");
    F64Args a;
    a.A = (const double*)A; a.B = (const double*)B; a.C = p->beta != 0.0f ? (const double*)C : nullptr;
    a.D = (double*)D;
    a.M = p->M; a.N = p->N; a.K = p->K; a.lda = p->lda; a.ldb = p->ldb; a.ldc = p->ldc; a.ldd = p->ldd;
    a.bsA = p->bsA; a.bsB = p->bsB; a.bsC = p->bsC; a.bsD = p->bsD;
    a.alpha = p->alpha; a.beta = p->beta; a.transA = ta; a.transB = tb;
    dim3 grid((unsigned)cdiv(p->M, 64), (unsigned)cdiv(p->N, 64), (unsigned)batch);
    hipLaunchKernelGGL(gemm_f64_kernel, grid, dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FFGEMM_SUCCESS : hip_fail(e, "fp64 kernel launch");
}

}  // namespace

// Scratch of ffgemm_run_x3_noalloc: the bf16x3 hi/lo planes of one problem,
// in a device array every GPU holds from module load until the process ends.
// Host code never passes `scratch` itself to a kernel (on the host it names a
// stand-in, not device memory): it uses the per-device address from
// hipGetSymbolAddress below.
constexpr size_t kNoAllocScratchBytes = (size_t)FFGEMM_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};

// Device address of `scratch` on the current device, looked up once per device.
int scratch_ptr(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= 64) return fail(FFGEMM_FATAL, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(FFGEMM_FATAL, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return FFGEMM_SUCCESS;
}
}  // namespace

extern "C" {

int ffgemm_init(void) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    cu_count();
    void* p = nullptr;
    return scratch_ptr(&p);
}

void ffgemm_shutdown(void) {}

const char* ffgemm_status_string(int status) {
    switch (status) {
        case FFGEMM_SUCCESS: return "SUCCESS";
        case FFGEMM_WARN: return "RTN_WARN";
        case FFGEMM_ERR: return "RTN_ERR";
        case FFGEMM_FATAL: return "RTN_FATAL";
    }
    return "UNKNOWN";
}

const char* ffgemm_last_error(void) { return g_last_error.c_str(); }

int ffgemm_tune_db_load(const char* path, int* entries) {
    std::map<TuneKey, ffgemm_tile> t;
    const std::string where = path ? std::string(path) : default_tune_path();
    const int n = read_tune_table(where, t);
    if (n < 0) return fail(FFGEMM_ERR, "malformed tuning table %s", where.c_str());
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.swap(t);
    g_tune_loaded = true;
    if (entries) *entries = n;
    return FFGEMM_SUCCESS;
}

const char* ffgemm_version(void) { return "ffgemm 0.1 gfx950"; }

int ffgemm_device_cu_count(void) { return cu_count(); }

int ffgemm_query_tile(const ffgemm_problem* p, ffgemm_tile* out) {
    if (p && p->dtype == FFGEMM_F64) {
        if (out) { memset(out, 0, sizeof(*out)); out->mt_m = 64; out->mt_n = 64; out->depth_u = 16; }
        return FFGEMM_SUCCESS;
    }
    // validation with dummy aligned pointers; nothing is dereferenced.  A
    // problem that runs in K chunks is checked through its first chunk.
    static const int64_t dummy[4] = {0, 0, 0, 0};
    ffgemm_problem q;
    if (const int64_t kc = mn_k_chunk(p)) {
        q = *p;
        q.K = kc;
        p = &q;
    }
    Launch L;
    int r = prepare(p, dummy, dummy, dummy, (void*)dummy, out && out->mt_m ? out : nullptr, L);
    if (r != FFGEMM_SUCCESS) return r;
    if (out) {
        memset(out, 0, sizeof(*out));
        const EngineGeometry geo = launch_geometry(L);
        out->mt_m = L.tile.m; out->mt_n = L.tile.n; out->depth_u = geo.depth_u;
        out->waves_m = geo.waves_m; out->waves_n = geo.waves_n; out->mfma_mn = L.mfma ? L.mfma : 32;
        out->group_m = L.args.group_m;
        out->lds_buffers = geo.lds_buffers;
        out->split_k = 1;
    }
    return FFGEMM_SUCCESS;
}

int ffgemm_tile_from_reference(const ffgemm_problem* p, int r, int s, int x, int y, int z, int depth_u,
                               int lds_buffers, ffgemm_tile* out) {
    if (!p || !out) return fail(FFGEMM_ERR, "null pointer");
    if (r <= 0 || s <= 0 || x <= 0 || y <= 0 || (z != 16 && z != 32))
        return fail(FFGEMM_ERR, "tuning record r=%d s=%d x=%d y=%d z=%d out of range", r, s, x, y, z);
    // MT0 = z*r*x, MT1 = z*s*y (src/gemmMfma.cpp:53); 64-bit so large factors
    // cannot overflow before the comparison below
    const int64_t mt0 = (int64_t)z * r * x, mt1 = (int64_t)z * s * y;
    const int64_t waves = (int64_t)x * y;
    // the exact record, when a kernel has it
    ffgemm_tile t;
    memset(&t, 0, sizeof(t));
    bool exact = false;
    if (tile_supported((int)std::min<int64_t>(mt0, 1 << 20), (int)std::min<int64_t>(mt1, 1 << 20))) {
        t.mt_m = (int)mt0; t.mt_n = (int)mt1;
        t.mfma_mn = z;
        if (x == 2 && (y == 2 || y == 4)) { t.waves_m = 2; t.waves_n = y; }
        ffgemm_tile q = t;
        exact = ffgemm_query_tile(p, &q) == FFGEMM_SUCCESS;
        if (!exact && t.waves_m) {  // the shape exists, the wave tiling not: the tile's own engine
            t.waves_m = t.waves_n = 0;
            q = t;
            exact = ffgemm_query_tile(p, &q) == FFGEMM_SUCCESS;
        }
        if (exact) t = q;
    }
    if (!exact) {
        // the smallest macro tile covering the record's (largest when none covers
        // it), 32 x 32 MFMAs, its default engine
        int best = -1;
        for (int i = 0; i < 5; ++i) {
            if (kTiles[i].m < mt0 || kTiles[i].n < mt1) continue;
            if (best < 0 || kTiles[i].m * kTiles[i].n < kTiles[best].m * kTiles[best].n) best = i;
        }
        if (best < 0) best = 0;
        memset(&t, 0, sizeof(t));
        t.mt_m = kTiles[best].m; t.mt_n = kTiles[best].n;
        ffgemm_tile q = t;
        int rc = ffgemm_query_tile(p, &q);
        if (rc != FFGEMM_SUCCESS) {
            // no kernel has that tile for this problem (strides off the 16-byte
            // grid run two square tiles only): the library's own choice
            memset(&q, 0, sizeof(q));
            rc = ffgemm_query_tile(p, &q);
            if (rc != FFGEMM_SUCCESS) return rc;
        }
        t = q;
    }
    *out = t;
    const bool same = exact && (depth_u <= 0 || depth_u == t.depth_u) &&
                      (lds_buffers <= 0 || lds_buffers == t.lds_buffers) &&
                      (waves == (int64_t)t.waves_m * t.waves_n);
    if (same) return FFGEMM_SUCCESS;
    return fail(FFGEMM_WARN, "record MT%lldx%lld (z=%d, %lld waves, depth_u %d, %d LDS buffers) runs as MT%dx%d "
                "(z=%d, %dx%d waves, depth_u %d, %d LDS stages)", (long long)mt0, (long long)mt1, z,
                (long long)waves, depth_u, lds_buffers, t.mt_m, t.mt_n, t.mfma_mn, t.waves_m, t.waves_n,
                t.depth_u, t.lds_buffers);
}

int ffgemm_run(const ffgemm_problem* p, const void* A, const void* B, const void* C, void* D,
               const ffgemm_tile* tile, void* stream) {
    if (p && p->dtype == FFGEMM_F64) return run_f64(p, A, B, C, D, (hipStream_t)stream);
    if (const int64_t kc = mn_k_chunk(p)) {
        // K chunks: the first applies the caller's beta * C, the rest add into D
        // (C = D, beta = 1; each element is read and written by one thread)
        const bool ta = p->layout >= FFGEMM_TN, tb = (p->layout & 1) != 0;
        const int es = p->dtype == FFGEMM_F32 ? 4 : 2;
        for (int64_t k0 = 0; k0 < p->K; k0 += kc) {
            ffgemm_problem q = *p;
            q.K = std::min<int64_t>(kc, p->K - k0);
            const void* Aq = (const char*)A + (ta ? k0 : k0 * p->lda) * es;
            const void* Bq = (const char*)B + (tb ? k0 * p->ldb : k0) * es;
            const void* Cq = C;
            if (k0 > 0) { q.beta = 1.0f; Cq = D; q.ldc = p->ldd; q.bsC = p->bsD; }
            Launch L;
            int r = prepare(&q, Aq, Bq, Cq, D, tile, L);
            if (r == FFGEMM_SUCCESS) r = launch(L, (hipStream_t)stream);
            if (r != FFGEMM_SUCCESS) return r;
        }
        return FFGEMM_SUCCESS;
    }
    Launch L;
    int r = prepare(p, A, B, C, D, tile, L);
    if (r != FFGEMM_SUCCESS) return r;
    if (p->M == 0 || p->N == 0) return FFGEMM_SUCCESS;
    return launch(L, (hipStream_t)stream);
}

}  // extern "C"

// bf16x3 through bf16 hi/lo planes in a workspace: K padded to the 64-element
// K-tile; A planes (M rows x 3kp: hi, hi, lo) then B planes (N rows: hi, lo, hi).
namespace {
int64_t x3_kp(int64_t K) { return cdiv(K, 64) * 64; }
int64_t round256(int64_t b) { return (b + 255) / 256 * 256; }
bool x3_planes_apply(const ffgemm_problem* p) {
    if (p->dtype != FFGEMM_F32 || p->compute != FFGEMM_COMPUTE_BF16X3) return false;
    if (p->M <= 0 || p->N <= 0 || p->K <= 0) return false;
    const int64_t batch = p->batch > 0 ? p->batch : 1;
    const int64_t row = 3 * x3_kp(p->K) * 2;  // bytes per plane row
    // the planes are K-contiguous operands of the bf16 GEMM, addressed from each
    // tile's first row (its 31-bit offset check); split grid limits
    return 257 * row < (1ll << 31) &&
           cdiv(std::max(p->M, p->N), 64) <= 65535 &&
           batch <= 65535;
}
int64_t x3_ws_bytes(const ffgemm_problem* p) {
    const int64_t batch = p->batch > 0 ? p->batch : 1;
    const int64_t row = 3 * x3_kp(p->K) * 2;
    return round256(p->M * row * batch) + round256(p->N * row * batch);
}

// Split-K workspace: one fp32 plane of ldw x N per slice.
int64_t split_ws_bytes(const Launch& L) { return round256((int64_t)L.split * split_ldw(L.args.M) * L.args.N * 4); }

// The plan ffgemm_run_ws follows for a problem without bf16x3 planes: split-K
// allowed (problems that run in K chunks do not split).
int split_plan(const ffgemm_problem* p, const void* A, const void* B, const void* C, void* D,
               const ffgemm_tile* tile, Launch& L) {
    return prepare(p, A, B, C, D, tile, L, mn_k_chunk(p) == 0);
}

// Runs a split plan: slices of ks K as one batched GEMM of fp32 partials into
// the workspace (alpha 1, beta 0; a shorter last slice as a second launch),
// then splitk_reduce_kernel forms D = alpha * sum + beta * C.
int run_split(const ffgemm_problem* p, const void* A, const void* B, const void* C, void* D, const Launch& L,
              float* ws, hipStream_t s) {
    const int S = L.split;
    const int64_t ks = L.ks, M = p->M, N = p->N, K = p->K;
    const bool ta = p->layout >= FFGEMM_TN, tb = (p->layout & 1) != 0;
    const int es = mode_elem_bytes(L.mode);
    const int64_t ldw = split_ldw(M);
    ffgemm_tile t;
    memset(&t, 0, sizeof(t));
    t.mt_m = L.tile.m; t.mt_n = L.tile.n; t.mfma_mn = L.mfma; t.split_k = 1;
    if (L.waves) { t.waves_m = 2; t.waves_n = L.waves / 2; }
    ffgemm_problem q = *p;
    q.dtype_out = FFGEMM_F32; q.round = FFGEMM_ROUND_RNE; q.alpha = 1.0f; q.beta = 0.0f;
    q.ldc = q.ldd = ldw; q.bsC = q.bsD = ldw * N;
    q.K = ks; q.bsA = ta ? ks : ks * p->lda; q.bsB = tb ? ks * p->ldb : ks;
    const int64_t last = K - (int64_t)(S - 1) * ks;
    const int nb = last == ks ? S : S - 1;
    q.batch = nb;
    int r = ffgemm_run(&q, A, B, nullptr, ws, &t, s);
    if (r != FFGEMM_SUCCESS) return r;
    if (nb < S) {
        q.batch = 1;
        q.K = last;
        const int64_t k0 = (int64_t)(S - 1) * ks;
        const void* Al = (const char*)A + (ta ? k0 : k0 * p->lda) * es;
        const void* Bl = (const char*)B + (tb ? k0 * p->ldb : k0) * es;
        r = ffgemm_run(&q, Al, Bl, nullptr, ws + (int64_t)(S - 1) * ldw * N, &t, s);
        if (r != FFGEMM_SUCCESS) return r;
    }
    SplitReduceArgs ra;
    ra.W = ws; ra.slice = ldw * N; ra.S = S; ra.M = (int32_t)M; ra.N = (int32_t)N; ra.ldw = (int32_t)ldw;
    ra.C = p->beta != 0.0f ? C : nullptr; ra.D = D; ra.ldc = (int32_t)p->ldc; ra.ldd = (int32_t)p->ldd;
    ra.alpha = p->alpha; ra.beta = p->beta; ra.out = L.out;
    ra.trunc = (L.out == kOutBF16 && p->round == FFGEMM_ROUND_TRUNCATE) ? 1 : 0;
    const int64_t work = cdiv(M, 4) * N;
    const int blocks = (int)std::min<int64_t>(cdiv(work, 256), 8 * 65536);
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(blocks), dim3(256), 0, s, ra);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FFGEMM_SUCCESS : hip_fail(e, "split-K reduce launch");
}

// Tail split: a problem whose 256 x 256 tiles fill r >= 1 whole passes over
// the CUs plus a short last pass runs as two sub-problems on the workspace
// path, the data-parallel part (whole 256 x 256 tile columns, or rows, that
// fill the r passes) and the remainder with split-K, so the last pass runs S
// times shorter slices on S times more CUs instead of leaving most CUs idle.
// Taken when the cost model puts it 15 % ahead of the single launch the
// library would otherwise make (its own tile choice, no split).
struct TailPlan {
    bool along_n;  // cut between tile columns (else tile rows)
    int64_t cut;   // first column (row) of the remainder
    int S;         // K slices of the remainder
};

// The sub-problem of rows [m0, m1) and columns [n0, n1) of D, with the
// operand, C and D pointers offset to it.
void sub_problem(const ffgemm_problem* p, const void*& A, const void*& B, const void*& C, void*& D, int64_t m0,
                 int64_t m1, int64_t n0, int64_t n1, ffgemm_problem& q) {
    const bool ta = p->layout >= FFGEMM_TN, tb = (p->layout & 1) != 0;
    const int es = p->dtype == FFGEMM_F32 ? 4 : p->dtype == FFGEMM_F8E4M3 ? 1 : 2;
    const int out = out_kind_of(p->dtype, p->dtype_out);
    const int os = out == kOutF32 ? 4 : 2;
    q = *p;
    q.M = m1 - m0;
    q.N = n1 - n0;
    A = (const char*)A + (ta ? m0 * p->lda : m0) * es;   // op(A) row m0
    B = (const char*)B + (tb ? n0 : n0 * p->ldb) * es;   // op(B) column n0
    if (C) C = (const char*)C + (n0 * p->ldc + m0) * os;
    D = (char*)D + (n0 * p->ldd + m0) * os;
}

bool tail_plan(const ffgemm_problem* p, const Launch& L, TailPlan& tp) {
    if (L.split != 1 || L.any || L.grid <= 0 || p->batch > 1 || L.waves || L.mfma) return false;
    if (L.mode != kBF16 && L.mode != kF16 && L.mode != kF8) return false;
    if (mn_k_chunk(p)) return false;
    const int64_t M = p->M, N = p->N, K = p->K, G = cu_count();
    const int64_t tm = cdiv(M, 256), tn = cdiv(N, 256), tiles = tm * tn, rounds = tiles / G;
    if (rounds < 1 || tiles % G == 0) return false;
    const int osz = L.out == kOutF32 ? 4 : 2;
    const bool use_c = p->beta != 0.0f;
    int ti = 0;  // the single launch's tile, as planned
    for (int i = 0; i < 5; ++i)
        if (kTiles[i].m == L.tile.m && kTiles[i].n == L.tile.n) ti = i;
    const double full = plan_cost(M, N, K, L.mode, osz, use_c, ti, 1);
    double best = full * 0.85;
    bool found = false;
    static const int cand[] = {2, 3, 4, 6, 8};
    for (int along_n = 1; along_n >= 0; --along_n) {
        const int64_t whole = along_n ? std::min(tn - 1, rounds * G / tm) : std::min(tm - 1, rounds * G / tn);
        if (whole < 1) continue;
        const int64_t cut = whole * 256;
        const double dp = along_n ? plan_cost(M, cut, K, L.mode, osz, use_c, 0, 1)
                                  : plan_cost(cut, N, K, L.mode, osz, use_c, 0, 1);
        for (int S : cand) {
            if (K / S < split_min_k(L.mode)) break;
            const double c = dp + (along_n ? plan_cost(M, N - cut, K, L.mode, osz, use_c, 0, S)
                                           : plan_cost(M - cut, N, K, L.mode, osz, use_c, 0, S));
            if (c < best) { best = c; tp = TailPlan{along_n != 0, cut, S}; found = true; }
        }
    }
    return found;
}

// The tail plan ffgemm_run_ws follows for p (no caller tile, no tuning-table
// entry), with the remainder's launch plan; false when it runs as one plan.
bool tail_plan_for(const ffgemm_problem* p, const ffgemm_tile* tile, const Launch& L, TailPlan& tp, Launch& LT) {
    ffgemm_tile tuned;
    if ((tile && tile->mt_m) || tuned_tile(p, &tuned) || !tail_plan(p, L, tp)) return false;
    static const int64_t dummy[4] = {0, 0, 0, 0};
    const void *A = dummy, *B = dummy, *C = dummy;
    void* D = (void*)dummy;
    ffgemm_problem q;
    if (tp.along_n) sub_problem(p, A, B, C, D, 0, p->M, tp.cut, p->N, q);
    else sub_problem(p, A, B, C, D, tp.cut, p->M, 0, p->N, q);
    ffgemm_tile t;
    memset(&t, 0, sizeof(t));
    t.mt_m = t.mt_n = 256;
    t.split_k = tp.S;
    return prepare_with(&q, dummy, dummy, dummy, (void*)dummy, &t, LT, true) == FFGEMM_SUCCESS && LT.split > 1;
}

// Runs a tail plan: the data-parallel part, then the remainder with split-K.
int run_tail(const ffgemm_problem* p, const void* A, const void* B, const void* C, void* D, const TailPlan& tp,
             float* ws, hipStream_t s) {
    ffgemm_problem q0, q1;
    const void *A0 = A, *B0 = B, *C0 = C, *A1 = A, *B1 = B, *C1 = C;
    void *D0 = D, *D1 = D;
    if (tp.along_n) {
        sub_problem(p, A0, B0, C0, D0, 0, p->M, 0, tp.cut, q0);
        sub_problem(p, A1, B1, C1, D1, 0, p->M, tp.cut, p->N, q1);
    } else {
        sub_problem(p, A0, B0, C0, D0, 0, tp.cut, 0, p->N, q0);
        sub_problem(p, A1, B1, C1, D1, tp.cut, p->M, 0, p->N, q1);
    }
    ffgemm_tile t;
    memset(&t, 0, sizeof(t));
    t.mt_m = t.mt_n = 256;
    t.split_k = 1;
    int r = ffgemm_run(&q0, A0, B0, C0, D0, &t, s);
    if (r != FFGEMM_SUCCESS) return r;
    Launch LT;
    t.split_k = tp.S;
    r = prepare_with(&q1, A1, B1, C1, D1, &t, LT, true);
    if (r != FFGEMM_SUCCESS) return r;
    return run_split(&q1, A1, B1, C1, D1, LT, ws, s);
}

// K-major repack (workspace path, bf16/fp16).  The engines run fastest with
// both operands K-contiguous (TN); at 256 x 256 the other layouts reach
// 0.84-0.86 (NN), 0.76-0.84 (TT) and 0.71-0.78 (NT) of its rate on MI355X
// (bf16, 4096^3 and 8192^3), their M/N-contiguous operands going through
// register transposes or transposing LDS reads.  Such a problem can instead
// copy those operands K-contiguous into the workspace (kpack16_kernel,
// HBM-bound) and run as TN, through this same workspace path.  The library
// takes it when the model below puts it 3 % ahead.  Measured on MI355X
// against the native engines (bf16): 4096^3 NN +3-4 %, NT +2-7 %, TT +6-7 %;
// 8192^3 NN +7 %, NT +12 %, TT +19 %; the autograd forward shape (NN 4096 x
// 8192 x 4096) +11 %.
struct RepackPlan {
    bool a, b;        // repack A (stored M-contiguous), B (stored N-contiguous)
    int64_t bytes_a;  // workspace bytes of A's copy (256-byte rounded); B's follows
    int64_t bytes_b;
};

// Copy rate of kpack16_kernel for the model: bytes read plus written per µs
// (5.7-6.5 TB/s measured on MI355X, 4096^2 and 8192^2 bf16).
constexpr double kRepackBytesPerUs = 5.5e6;

// The problem the repacked operands form: TN, packed K-contiguous copies
// (leading dimension K, batch stride rows * K).
ffgemm_problem repacked_problem(const ffgemm_problem* p, const RepackPlan& rp) {
    ffgemm_problem q = *p;
    q.layout = FFGEMM_TN;
    if (rp.a) { q.lda = p->K; q.bsA = p->M * p->K; }
    if (rp.b) { q.ldb = p->K; q.bsB = p->N * p->K; }
    return q;
}

// Whether p (native plan L) repacks under `tile` (its repack field: 0 the
// model's choice, 1 never, 2 always).  A forced repack the problem cannot take
// is an error; TN, K = 0 and empty problems have nothing to repack.
int plan_repack(const ffgemm_problem* p, const ffgemm_tile* tile, const Launch& L, RepackPlan& rp, bool& take) {
    take = false;
    const int force = tile ? tile->repack : 0;
    if (force < 0 || force > 2) return fail(FFGEMM_ERR, "repack %d (0, 1 or 2)", force);
    if (force == 1 || p->layout == FFGEMM_TN || L.grid <= 0 || p->K <= 0) return FFGEMM_SUCCESS;
    const int64_t batch = p->batch > 0 ? p->batch : 1;
    const char* need = (L.mode != kBF16 && L.mode != kF16) ? "bf16/fp16 operands"
                       : L.any ? "16-byte aligned leading dimensions, batch strides and pointers"
                       : (p->K % 8) ? "K a multiple of 8"
                       : batch > 65535 ? "batch <= 65535"
                       : (cdiv(p->M, 64) + cdiv(p->N, 64)) * cdiv(p->K, 64) * batch >= ((int64_t)1 << 31)
                           ? "fewer than 2^31 blocks of 64 x 64"
                           : nullptr;
    if (need) return force == 2 ? fail(FFGEMM_ERR, "repack needs %s", need) : FFGEMM_SUCCESS;
    const bool ta = p->layout >= FFGEMM_TN, tb = (p->layout & 1) != 0;
    rp.a = !ta;
    rp.b = tb;
    rp.bytes_a = rp.a ? round256(p->M * p->K * 2 * batch) : 0;
    rp.bytes_b = rp.b ? round256(p->N * p->K * 2 * batch) : 0;
    if (force == 2) {
        take = true;
        return FFGEMM_SUCCESS;
    }
    // the library's own choice: no caller tile, no tuning-table entry, a
    // single launch of 256 x 256 or 128 x 128 tiles (split-K plans keep their
    // layout; the repacked problem takes the same tile)
    ffgemm_tile tuned;
    const bool big = L.tile.m == 256 && L.tile.n == 256, mid = L.tile.m == 128 && L.tile.n == 128;
    if ((tile && tile->mt_m) || L.split > 1 || !(big || mid) || L.waves || L.mfma || tuned_tile(p, &tuned))
        return FFGEMM_SUCCESS;
    // model (µs): passes of tiles at 70 % of the MFMA rate at 1.9 GHz times the
    // tile's relative efficiency (as plan_cost), plus 4 µs each, over the
    // layout's rate relative to TN; against the same TN time plus the copies
    // at kRepackBytesPerUs and 2 µs for their launch.  The layouts' relative
    // rates (NN, NT, TN, TT), measured on MI355X: at 256 x 256, bf16 4096^3
    // and 8192^3; fp16 at 4096^3, whose TN runs at a lower clock when every CU
    // is busy, so its other layouts lose less there (a grid that leaves CUs
    // idle, 3072^3 fp16, lost as much as bf16); at 128 x 128 (2048 x 2048 x
    // 8192, 1024 x 4096 x 4096), where NT's two register rings cost most.
    static const double eff_big[4] = {0.85, 0.76, 1.0, 0.83}, eff_big_f16[4] = {0.88, 0.82, 1.0, 0.85};
    static const double eff_mid[4] = {0.85, 0.60, 1.0, 0.80};
    const int64_t tiles = cdiv(p->M, L.tile.m) * cdiv(p->N, L.tile.n) * batch;
    const bool full = tiles >= cu_count();
    const double* lay_eff = mid ? eff_mid : (L.mode == kF16 && full) ? eff_big_f16 : eff_big;
    const double passes = (double)cdiv(tiles, cu_count());
    const double tile_eff = big ? 1.0 : 0.70;
    const double tn = passes * (2.0 * L.tile.m * L.tile.n * (double)p->K / (4096.0 * 1.9e3 * 0.7 * tile_eff) + 4.0);
    const double moved = 2.0 * (double)(rp.bytes_a + rp.bytes_b);
    const double cost = tn + moved / kRepackBytesPerUs + 2.0;  // one copy launch
    take = cost < 0.97 * tn / lay_eff[p->layout];
    return FFGEMM_SUCCESS;
}

// Runs a repack plan: the copies into the front of the workspace, then the TN
// problem on the workspace path with the rest of it.
int run_repack(const ffgemm_problem* p, const void* A, const void* B, const void* C, void* D,
               const ffgemm_tile* tile, const RepackPlan& rp, void* ws, size_t ws_bytes, hipStream_t s) {
    const int64_t batch = p->batch > 0 ? p->batch : 1;
    uint16_t* wa = (uint16_t*)ws;
    uint16_t* wb = (uint16_t*)((char*)ws + rp.bytes_a);
    // one launch for both operands: A's blocks (if repacked) first
    KPackOperand op[2];
    int n = 0;
    auto add = [&](const void* x, uint16_t* y, int64_t rows, int64_t ldx, int64_t bsx) {
        KPackOperand& o = op[n++];
        o.x = (const uint16_t*)x; o.y = y; o.rows = rows; o.ldx = ldx; o.bsx = bsx; o.bsy = rows * p->K;
        o.tiles_r = (uint32_t)cdiv(rows, 64);
        o.blocks = (uint32_t)(cdiv(rows, 64) * cdiv(p->K, 64));
    };
    if (rp.a) add(A, wa, p->M, p->lda, p->bsA);
    if (rp.b) add(B, wb, p->N, p->ldb, p->bsB);
    if (n == 1) { op[1] = op[0]; op[1].blocks = 0; }
    const int64_t total = ((int64_t)op[0].blocks + op[1].blocks) * batch;
    // four waves a workgroup, one 64 x 64 block each per pass; at most two
    // workgroups a CU, so every wave but those of the smallest copies takes
    // several blocks and overlaps their loads and stores (4096^2 bf16 on
    // MI355X: 11.5 us against 12.4 with four workgroups a CU; 8192^2 level)
    const int64_t grid = std::min<int64_t>(cdiv(total, 4), 2 * (int64_t)cu_count());
    hipLaunchKernelGGL(kpack16_kernel, dim3((unsigned)grid), dim3(256), 0, s, op[0], op[1], p->K, (uint32_t)batch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "repack launch");
    const ffgemm_problem q = repacked_problem(p, rp);
    ffgemm_tile t;
    if (tile) t = *tile;
    else memset(&t, 0, sizeof(t));
    t.repack = 1;
    const size_t used = (size_t)(rp.bytes_a + rp.bytes_b);
    void* rest = ws_bytes > used ? (void*)((char*)ws + used) : nullptr;
    return ffgemm_run_ws(&q, rp.a ? (const void*)wa : A, rp.b ? (const void*)wb : B, C, D, &t, rest,
                         rest ? ws_bytes - used : 0, s);
}
}  // namespace

extern "C" {

int ffgemm_workspace_size_ex(const ffgemm_problem* p, const ffgemm_tile* tile, size_t* bytes) {
    if (!p || !bytes) return fail(FFGEMM_ERR, "null pointer");
    *bytes = 0;
    if (x3_planes_apply(p)) {
        *bytes = (size_t)x3_ws_bytes(p);
        return FFGEMM_SUCCESS;
    }
    // the repacked operands, then the split-K partial planes of the plan
    // ffgemm_run_ws would follow (validation with dummy aligned pointers;
    // nothing is dereferenced)
    static const int64_t dummy[4] = {0, 0, 0, 0};
    Launch L;
    if (p->dtype != FFGEMM_F64 && split_plan(p, dummy, dummy, dummy, (void*)dummy, tile, L) == FFGEMM_SUCCESS) {
        RepackPlan rp;
        bool take = false;
        const int r = plan_repack(p, tile, L, rp, take);
        if (r != FFGEMM_SUCCESS) return r;
        if (take) {
            const ffgemm_problem q = repacked_problem(p, rp);
            ffgemm_tile t;
            if (tile) t = *tile;
            else memset(&t, 0, sizeof(t));
            t.repack = 1;
            size_t rest = 0;
            const int rq = ffgemm_workspace_size_ex(&q, &t, &rest);
            if (rq != FFGEMM_SUCCESS) return rq;
            *bytes = (size_t)(rp.bytes_a + rp.bytes_b) + rest;
            return FFGEMM_SUCCESS;
        }
        TailPlan tp;
        Launch LT;
        if (L.split > 1) *bytes = (size_t)split_ws_bytes(L);
        else if (tail_plan_for(p, tile, L, tp, LT)) *bytes = (size_t)split_ws_bytes(LT);  // the remainder's planes
    }
    return FFGEMM_SUCCESS;
}

int ffgemm_workspace_size(const ffgemm_problem* p, size_t* bytes) { return ffgemm_workspace_size_ex(p, nullptr, bytes); }

int ffgemm_query_tile_ws(const ffgemm_problem* p, ffgemm_tile* out) {
    if (!p || !out) return fail(FFGEMM_ERR, "null pointer");
    if (x3_planes_apply(p) || p->dtype == FFGEMM_F64) return ffgemm_query_tile(p, out);
    static const int64_t dummy[4] = {0, 0, 0, 0};
    Launch L;
    const ffgemm_tile* given = (out->mt_m || out->repack || out->split_k) ? out : nullptr;
    const int r = split_plan(p, dummy, dummy, dummy, (void*)dummy, given, L);
    if (r != FFGEMM_SUCCESS) return r;
    {
        // a repacking plan reports the TN problem's record, with repack 2
        RepackPlan rp;
        bool take = false;
        const int rr = plan_repack(p, given, L, rp, take);
        if (rr != FFGEMM_SUCCESS) return rr;
        if (take) {
            const ffgemm_problem q = repacked_problem(p, rp);
            ffgemm_tile t = *out;
            if (!given) memset(&t, 0, sizeof(t));
            t.repack = 1;
            const int rq = ffgemm_query_tile_ws(&q, &t);
            if (rq != FFGEMM_SUCCESS) return rq;
            *out = t;
            out->repack = 2;
            return FFGEMM_SUCCESS;
        }
    }
    if (L.split <= 1) return ffgemm_query_tile(p, out);
    const EngineGeometry geo = launch_geometry(L);
    memset(out, 0, sizeof(*out));
    out->mt_m = L.tile.m; out->mt_n = L.tile.n; out->depth_u = geo.depth_u;
    out->waves_m = geo.waves_m; out->waves_n = geo.waves_n; out->mfma_mn = L.mfma ? L.mfma : 32;
    out->group_m = L.args.group_m; out->lds_buffers = geo.lds_buffers; out->split_k = L.split;
    return FFGEMM_SUCCESS;
}

int ffgemm_run_ws(const ffgemm_problem* p, const void* A, const void* B, const void* C, void* D,
                  const ffgemm_tile* tile, void* workspace, size_t workspace_bytes, void* stream) {
    if (!p) return fail(FFGEMM_ERR, "null problem");
    if (!x3_planes_apply(p)) {
        if (p->dtype == FFGEMM_F64) return ffgemm_run(p, A, B, C, D, tile, stream);
        Launch L;
        const int r = split_plan(p, A, B, C, D, tile, L);
        if (r != FFGEMM_SUCCESS) return r;
        RepackPlan rp;
        bool take = false;
        const int rr = plan_repack(p, tile, L, rp, take);
        if (rr != FFGEMM_SUCCESS) return rr;
        if (take) {
            size_t need = 0;
            const int rs = ffgemm_workspace_size_ex(p, tile, &need);
            if (rs != FFGEMM_SUCCESS) return rs;
            if (workspace && workspace_bytes >= need) {
                if ((uintptr_t)workspace % 256) return fail(FFGEMM_ERR, "workspace must be 256-byte aligned");
                return run_repack(p, A, B, C, D, tile, rp, workspace, workspace_bytes, (hipStream_t)stream);
            }
            if (tile && tile->repack == 2)
                return fail(FFGEMM_ERR, "repack needs a workspace of %zu bytes (%zu given)", need,
                            workspace ? workspace_bytes : (size_t)0);
        }
        if (L.split > 1) {
            const int64_t need = split_ws_bytes(L);
            if (workspace && workspace_bytes >= (size_t)need) {
                if ((uintptr_t)workspace % 256) return fail(FFGEMM_ERR, "workspace must be 256-byte aligned");
                return run_split(p, A, B, C, D, L, (float*)workspace, (hipStream_t)stream);
            }
            if (tile && tile->mt_m && tile->split_k > 1)
                return fail(FFGEMM_ERR, "split_k %d needs a workspace of %lld bytes (%zu given)", tile->split_k,
                            (long long)need, workspace ? workspace_bytes : (size_t)0);
        }
        if (tile && tile->split_k > 1) {  // a requested split that came to one slice
            ffgemm_tile t = *tile;
            t.split_k = 1;
            return ffgemm_run(p, A, B, C, D, &t, stream);
        }
        TailPlan tp;
        Launch LT;
        if (L.split <= 1 && tail_plan_for(p, tile, L, tp, LT) && workspace &&
            workspace_bytes >= (size_t)split_ws_bytes(LT)) {
            if ((uintptr_t)workspace % 256) return fail(FFGEMM_ERR, "workspace must be 256-byte aligned");
            return run_tail(p, A, B, C, D, tp, (float*)workspace, (hipStream_t)stream);
        }
        return ffgemm_run(p, A, B, C, D, tile, stream);
    }
    if (!workspace || workspace_bytes < (size_t)x3_ws_bytes(p) || mn_k_chunk(p))
        return ffgemm_run(p, A, B, C, D, tile, stream);
    if ((uintptr_t)workspace % 256) return fail(FFGEMM_ERR, "workspace must be 256-byte aligned");
    Launch L;
    int r = prepare(p, A, B, C, D, nullptr, L);  // the fp32 problem's own checks
    if (r != FFGEMM_SUCCESS) return r;
    hipStream_t s = (hipStream_t)stream;
    const int64_t batch = p->batch > 0 ? p->batch : 1;
    const int64_t kp = x3_kp(p->K);
    uint16_t* wa = (uint16_t*)workspace;
    uint16_t* wb = (uint16_t*)((char*)workspace + round256(p->M * 3 * kp * 2 * batch));
    const bool aK = p->layout >= FFGEMM_TN, bK = (p->layout & 1) == 0;
    hipLaunchKernelGGL(x3_split_kernel, dim3((unsigned)(kp / 64), (unsigned)cdiv(p->M, 64), (unsigned)batch),
                       dim3(256), 0, s, (const float*)A, p->M, p->K, p->lda, p->bsA, (int)aK, wa, kp, 4);
    hipLaunchKernelGGL(x3_split_kernel, dim3((unsigned)(kp / 64), (unsigned)cdiv(p->N, 64), (unsigned)batch),
                       dim3(256), 0, s, (const float*)B, p->N, p->K, p->ldb, p->bsB, (int)bK, wb, kp, 2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "bf16x3 split launch");
    ffgemm_problem q = *p;
    q.layout = FFGEMM_TN;
    q.dtype = FFGEMM_BF16;
    q.dtype_out = FFGEMM_F32;
    q.compute = FFGEMM_COMPUTE_DEFAULT;
    q.K = 3 * kp;
    q.lda = q.ldb = 3 * kp;
    q.bsA = p->M * 3 * kp;
    q.bsB = p->N * 3 * kp;
    return ffgemm_run(&q, wa, wb, C, D, tile, stream);
}

int ffgemm_run_x3_noalloc(const ffgemm_problem* p, const void* A, const void* B, const void* C, void* D,
                          const ffgemm_tile* tile, void* stream) {
    if (!p) return fail(FFGEMM_ERR, "null problem");
    if (p->dtype != FFGEMM_F32 || p->compute != FFGEMM_COMPUTE_BF16X3)
        return fail(FFGEMM_ERR, "ffgemm_run_x3_noalloc runs bf16x3 problems (fp32 operands, compute BF16X3)");
    // problems without a planes path split in registers: no scratch either
    if (!x3_planes_apply(p) || mn_k_chunk(p)) return ffgemm_run(p, A, B, C, D, tile, stream);
    const int64_t need = x3_ws_bytes(p);
    if (need > (int64_t)kNoAllocScratchBytes)
        return fail(FFGEMM_ERR, "bf16x3 planes of this problem need %lld bytes; the no-allocation scratch holds %zu "
                    "((M + N) * 3 * roundup(K, 64) * 2 * batch must fit; use ffgemm_run_ws with a workspace)",
                    (long long)need, kNoAllocScratchBytes);
    void* ws = nullptr;
    const int r = scratch_ptr(&ws);
    if (r != FFGEMM_SUCCESS) return r;
    // the split kernels write every element of both planes (K padding as
    // zeros), so nothing left by an earlier call is read
    return ffgemm_run_ws(p, A, B, C, D, tile, ws, kNoAllocScratchBytes, stream);
}

int ffgemm_run_tn_kernarg(void* D, const void* C, const void* A, const void* B, float alpha, float beta,
                          uint32_t ldd, uint32_t bsD, uint32_t ldc, uint32_t bsC, uint32_t lda, uint32_t bsA,
                          uint32_t ldb, uint32_t bsB, uint32_t M, uint32_t N, uint32_t K, uint32_t batch,
                          int dtype, const ffgemm_tile* tile, void* stream) {
    if (dtype < FFGEMM_F32 || dtype > FFGEMM_BF16) return fail(FFGEMM_ERR, "data type %d not in {1,2,3}", dtype);
    ffgemm_problem p;
    memset(&p, 0, sizeof(p));
    p.layout = FFGEMM_TN; p.dtype = dtype; p.dtype_out = dtype;
    p.M = M; p.N = N; p.K = K; p.batch = batch ? batch : 1;
    p.lda = lda; p.ldb = ldb; p.ldc = ldc; p.ldd = ldd;
    p.bsA = bsA; p.bsB = bsB; p.bsC = bsC; p.bsD = bsD;
    p.alpha = alpha; p.beta = beta;
    return ffgemm_run(&p, A, B, C, D, tile, stream);
}

int ffgemm_run_tensile_args(const void* kernarg, size_t kernarg_bytes, int dtype, const ffgemm_tile* tile,
                            void* stream) {
    // byte offsets of the reference's Tensile-layout argument buffer
    // (src/gemmMfma.cpp:72-102; its kernels' metadata lists the same)
    enum : size_t {
        kSizeC = 0, kSizeA = 8, kSizeB = 16, kD = 24, kC = 32, kA = 40, kB = 48, kAlpha = 56, kBeta = 60,
        kStrideD0 = 64, kStrideD1 = 68, kStrideC0 = 72, kStrideC1 = 76, kStrideA0 = 80, kStrideA1 = 84,
        kStrideB0 = 88, kStrideB1 = 92, kFree0 = 96, kFree1 = 100, kFree2 = 104, kSum0 = 108, kBytes = 148
    };
    if (!kernarg || kernarg_bytes < kBytes)
        return fail(FFGEMM_ERR, "Tensile argument buffer of %zu bytes; %d expected", kernarg_bytes, (int)kBytes);
    if (dtype < FFGEMM_F32 || dtype > FFGEMM_BF16) return fail(FFGEMM_ERR, "data type %d not in {1,2,3}", dtype);
    const char* k = (const char*)kernarg;
    auto u32 = [&](size_t off) { uint32_t v; memcpy(&v, k + off, 4); return v; };
    auto u64 = [&](size_t off) { uint64_t v; memcpy(&v, k + off, 8); return v; };
    auto ptr = [&](size_t off) { void* v; memcpy(&v, k + off, sizeof(v)); return v; };
    // alpha / beta: 32-bit slots; the half-precision (HBH) kernels take fp16 in
    // the low half (src/gemmMfmaKernelWriter.h:1845-1854), the others fp32
    auto scalar = [&](size_t off) {
        const uint32_t bits = u32(off);
        if (dtype == FFGEMM_F16) {
            const uint16_t h = (uint16_t)(bits & 0xffffu);
            _Float16 x;
            memcpy(&x, &h, 2);
            return (float)x;
        }
        float f;
        memcpy(&f, &bits, 4);
        return f;
    };
    ffgemm_problem p;
    memset(&p, 0, sizeof(p));
    p.layout = FFGEMM_TN; p.dtype = dtype; p.dtype_out = dtype;
    p.M = u32(kFree0); p.N = u32(kFree1); p.K = u32(kSum0);
    p.batch = u32(kFree2) ? u32(kFree2) : 1;
    p.ldd = u32(kStrideD0); p.bsD = u32(kStrideD1);
    p.ldc = u32(kStrideC0); p.bsC = u32(kStrideC1);
    p.lda = u32(kStrideA0); p.bsA = u32(kStrideA1);
    p.ldb = u32(kStrideB0); p.bsB = u32(kStrideB1);
    p.alpha = scalar(kAlpha); p.beta = scalar(kBeta);
    // sizeA/B/C (elements; 0: not given) must cover what the problem addresses;
    // sizeC is the C/D tensor: D always, C when beta != 0
    if (p.M && p.N) {
        const int64_t b = p.batch - 1;
        const int64_t needA = p.K ? b * p.bsA + (p.M - 1) * p.lda + p.K : 0;
        const int64_t needB = p.K ? b * p.bsB + (p.N - 1) * p.ldb + p.K : 0;
        int64_t needC = b * p.bsD + (p.N - 1) * p.ldd + p.M;
        if (p.beta != 0.0f) needC = std::max<int64_t>(needC, b * p.bsC + (p.N - 1) * p.ldc + p.M);
        const int64_t sA = (int64_t)u64(kSizeA), sB = (int64_t)u64(kSizeB), sC = (int64_t)u64(kSizeC);
        if ((sA && sA < needA) || (sB && sB < needB) || (sC && sC < needC))
            return fail(FFGEMM_ERR, "sizeA/sizeB/sizeC smaller than the problem addresses");
    }
    return ffgemm_run(&p, ptr(kA), ptr(kB), ptr(kC), ptr(kD), tile, stream);
}

size_t ffgemm_fp8_workspace_size(void) { return 256; }

int ffgemm_fp8_quantize(const float* x, int64_t rows, int64_t k, int64_t ldx, int trans_x, uint8_t* q,
                        int64_t ldq, float* scale_out, void* work, void* stream) {
    if (rows < 0 || k < 0) return fail(FFGEMM_ERR, "negative size");
    if (!x || !q || !scale_out || !work) return fail(FFGEMM_ERR, "null pointer");
    if (ldx < (trans_x ? k : rows) || ldq < k) return fail(FFGEMM_ERR, "leading dimension too small");
    if (rows >= (1ll << 31) / 32 || k >= (1ll << 31) / 32) return fail(FFGEMM_ERR, "matrix too large");
    hipStream_t s = (hipStream_t)stream;
    uint32_t* amax = (uint32_t*)work;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, amax, 1);
    const int64_t total = rows * k;
    int blocks = (int)std::min<int64_t>(std::max<int64_t>(1, cdiv(total, 256 * 8)), 2048);
    if (total > 0) hipLaunchKernelGGL(amax_kernel, dim3(blocks), dim3(256), 0, s, x, rows, k, ldx, trans_x, amax);
    hipLaunchKernelGGL(fp8_scale_kernel, dim3(1), dim3(64), 0, s, (const uint32_t*)amax, scale_out);
    if (total > 0) {
        dim3 grid((unsigned)cdiv(k, 32), (unsigned)cdiv(rows, 32));
        hipLaunchKernelGGL(fp8_quant_kernel, grid, dim3(256), 0, s, x, rows, k, ldx, trans_x, q, ldq,
                           (const float*)scale_out);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FFGEMM_SUCCESS : hip_fail(e, "fp8 quantize launch");
}

int ffgemm_convert(const float* src, void* dst, int64_t n, int dtype_dst, void* stream) {
    if (dtype_dst != FFGEMM_BF16 && dtype_dst != FFGEMM_F16) return fail(FFGEMM_ERR, "convert to bf16/f16 only");
    if (n < 0 || (n && (!src || !dst))) return fail(FFGEMM_ERR, "bad arguments");
    if (n == 0) return FFGEMM_SUCCESS;
    int blocks = (int)std::min<int64_t>(cdiv(n, 256 * 4), 8192);
    hipLaunchKernelGGL(convert_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, src, dst, n, dtype_dst);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FFGEMM_SUCCESS : hip_fail(e, "convert launch");
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Grouped GEMM plans
// ---------------------------------------------------------------------------
struct ffgemm_grouped_plan_s {
    void* d_mem = nullptr;  // descriptors, then the schedule's first[grid + 1], then the tiles
    const GroupDesc* d_descs = nullptr;
    const int32_t* d_first = nullptr;
    const GroupTile* d_sched = nullptr;
    int count = 0;
    int64_t total = 0;
    int grid = 0;
    int mode = 0, out = 0, layout = 0;
    Tile tile{0, 0};
    int device = 0;
};

extern "C" {

int ffgemm_grouped_plan_create(ffgemm_grouped_plan* plan, int count, const ffgemm_problem* probs,
                               const void* const* A, const void* const* B, const void* const* C,
                               void* const* D, const ffgemm_tile* tile) {
    if (!plan) return fail(FFGEMM_ERR, "null plan pointer");
    *plan = nullptr;
    if (count <= 0 || !probs || !A || !B || !D) return fail(FFGEMM_ERR, "empty or null group");
    // one tile for the whole group: the caller's, else the largest that fills the CUs
    Tile t{0, 0};
    if (tile && tile->mt_m) {
        if (!tile_supported(tile->mt_m, tile->mt_n))
            return fail(FFGEMM_ERR, "unsupported macro tile %dx%d", tile->mt_m, tile->mt_n);
        if (tile->mfma_mn && tile->mfma_mn != 32)
            return fail(FFGEMM_ERR, "the grouped launch runs 32x32 MFMAs (mfma_mn %d)", tile->mfma_mn);
        t = Tile{tile->mt_m, tile->mt_n};
    } else {
        for (const Tile& c : kTiles) {
            int64_t n = 0;
            for (int i = 0; i < count; ++i) n += cdiv(probs[i].M, c.m) * cdiv(probs[i].N, c.n);
            t = c;
            if (n >= 2 * cu_count()) break;
        }
    }
    ffgemm_tile tt;
    memset(&tt, 0, sizeof(tt));
    tt.mt_m = t.m; tt.mt_n = t.n;
    std::vector<GroupDesc> descs(count);
    int mode = -1, out = -1, layout = -1;
    for (int i = 0; i < count; ++i) {
        const ffgemm_problem& p = probs[i];
        if ((p.batch > 1)) return fail(FFGEMM_ERR, "grouped problems must have batch 1");
        Launch L;
        int r = prepare(&p, A[i], B[i], C ? C[i] : nullptr, D[i], &tt, L);
        if (r != FFGEMM_SUCCESS) return r;
        if (L.any)
            return fail(FFGEMM_ERR, "grouped problems need leading dimensions and pointers on the 16-byte grid and K "
                        "a multiple of 8 for a K-contiguous operand (problem %d)", i);
        if (i == 0) { mode = L.mode; out = L.out; layout = L.layout; }
        else if (L.mode != mode || L.out != out || L.layout != layout)
            return fail(FFGEMM_ERR, "grouped problems must share dtype, compute and layout");
        // the persistent kernels exist for bf16 -> bf16 and fp16 -> fp16 only;
        // refuse anything else here, not at ffgemm_grouped_plan_run (which a
        // caller may be capturing into a graph)
        if (!((L.mode == kBF16 && L.out == kOutBF16) || (L.mode == kF16 && L.out == kOutF16)))
            return fail(FFGEMM_ERR, "grouped GEMM runs bf16 or fp16 problems with the output in the input type "
                        "(problem %d: dtype %d, output %d, compute %d)", i, p.dtype, p.dtype_out, p.compute);
        descs[i].a = L.args;  // group height as the single GEMM's default (prepare)
    }
    // Tile schedule, longest-processing-time first: tiles in K-descending
    // order, each dealt to the workgroup with the least work so far (cost of a
    // tile: its K-steps of 64 plus two for the prologue and epilogue), so the
    // short tiles fill the tail of every workgroup's list.
    std::vector<int> order(count);
    for (int i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return probs[x].K > probs[y].K; });
    int64_t total = 0;
    for (int i = 0; i < count; ++i) total += (int64_t)descs[i].a.tiles_m * descs[i].a.tiles_n;
    if (total > INT32_MAX) return fail(FFGEMM_ERR, "grouped GEMM of %lld tiles exceeds 2^31", (long long)total);
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(total, (int64_t)cu_count() * tile_occupancy(t.m, t.n, layout)));
    std::vector<std::vector<GroupTile>> lists(grid);
    {
        // Workgroups are dispatched to the 8 XCDs round-robin (workgroup b on
        // XCD b % 8), each XCD with its own L2.  Ties go to the lowest rank,
        // and ranks run XCD by XCD, so consecutive tiles, which share A and B
        // panels, start on one XCD.
        const int xcds = (grid % 8 == 0) ? 8 : 1;
        auto rank = [&](int b) { return (b % xcds) * (grid / xcds) + b / xcds; };
        using Slot = std::pair<int64_t, int>;  // (work so far, rank)
        std::priority_queue<Slot, std::vector<Slot>, std::greater<Slot>> heap;
        std::vector<int> of_rank(grid);
        for (int b = 0; b < grid; ++b) {
            of_rank[rank(b)] = b;
            heap.push({0, rank(b)});
        }
        for (int i : order) {
            const int nt = descs[i].a.tiles_m * descs[i].a.tiles_n;
            const int64_t cost = cdiv(probs[i].K, 64) + 2;
            for (int k = 0; k < nt; ++k) {
                Slot w = heap.top();
                heap.pop();
                lists[of_rank[w.second]].push_back(GroupTile{i, k});
                heap.push({w.first + cost, w.second});
            }
        }
    }
    std::vector<int32_t> first(grid + 1, 0);
    std::vector<GroupTile> sched;
    sched.reserve((size_t)total);
    for (int b = 0; b < grid; ++b) {
        first[b] = (int32_t)sched.size();
        sched.insert(sched.end(), lists[b].begin(), lists[b].end());
    }
    first[grid] = (int32_t)sched.size();

    ffgemm_grouped_plan_s* P = new ffgemm_grouped_plan_s();
    P->count = count;
    P->total = total;
    P->grid = grid;
    P->mode = mode; P->out = out; P->layout = layout; P->tile = t;
    if (hipError_t e = hipGetDevice(&P->device); e != hipSuccess) {
        delete P;
        return hip_fail(e, "hipGetDevice");
    }
    if (total > 0) {
        const size_t off_first = round256(sizeof(GroupDesc) * count);
        const size_t off_sched = off_first + round256(sizeof(int32_t) * (grid + 1));
        const size_t bytes = off_sched + sizeof(GroupTile) * sched.size();
        std::vector<char> host(bytes, 0);
        memcpy(host.data(), descs.data(), sizeof(GroupDesc) * count);
        memcpy(host.data() + off_first, first.data(), sizeof(int32_t) * (grid + 1));
        memcpy(host.data() + off_sched, sched.data(), sizeof(GroupTile) * sched.size());
        hipError_t e = hipMalloc(&P->d_mem, bytes);
        if (e != hipSuccess) { delete P; return hip_fail(e, "hipMalloc(grouped plan)"); }
        e = hipMemcpy(P->d_mem, host.data(), bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) { (void)hipFree(P->d_mem); delete P; return hip_fail(e, "hipMemcpy(grouped plan)"); }
        P->d_descs = (const GroupDesc*)P->d_mem;
        P->d_first = (const int32_t*)((const char*)P->d_mem + off_first);
        P->d_sched = (const GroupTile*)((const char*)P->d_mem + off_sched);
    }
    *plan = P;
    return FFGEMM_SUCCESS;
}

int ffgemm_grouped_plan_run(ffgemm_grouped_plan plan, void* stream) {
    if (!plan) return fail(FFGEMM_ERR, "null plan");
    if (plan->total == 0) return FFGEMM_SUCCESS;
    int r = launch_grouped(plan->mode, plan->out, plan->layout, plan->tile.m, plan->tile.n, plan->d_descs,
                           plan->d_first, plan->d_sched, plan->grid, (hipStream_t)stream);
    if (r != 0) return fail(FFGEMM_ERR, "no grouped kernel for mode %d out %d layout %d", plan->mode, plan->out,
                            plan->layout);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FFGEMM_SUCCESS : hip_fail(e, "grouped launch");
}

int ffgemm_grouped_plan_info(ffgemm_grouped_plan plan, int64_t* total_tiles, int32_t* grid, ffgemm_tile* tile) {
    if (!plan) return fail(FFGEMM_ERR, "null plan");
    if (total_tiles) *total_tiles = plan->total;
    if (grid) *grid = plan->grid;
    if (tile) {
        memset(tile, 0, sizeof(*tile));
        const EngineGeometry geo = engine_geometry(plan->mode, plan->layout >= FFGEMM_TN, (plan->layout & 1) == 0,
                                                   plan->tile.m, plan->tile.n);
        tile->mt_m = plan->tile.m; tile->mt_n = plan->tile.n; tile->waves_m = geo.waves_m;
        tile->waves_n = geo.waves_n; tile->mfma_mn = 32; tile->lds_buffers = geo.lds_buffers;
        tile->depth_u = geo.depth_u; tile->group_m = 1;
    }
    return FFGEMM_SUCCESS;
}

int ffgemm_grouped_plan_destroy(ffgemm_grouped_plan plan) {
    if (!plan) return FFGEMM_SUCCESS;
    const hipError_t e = plan->d_mem ? hipFree(plan->d_mem) : hipSuccess;
    delete plan;
    return e == hipSuccess ? FFGEMM_SUCCESS : hip_fail(e, "hipFree(grouped plan)");
}

}  // extern "C"
