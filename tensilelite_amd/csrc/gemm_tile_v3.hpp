// Tile engine for 256 x 256 tiles of 16-bit operands (bf16 / fp16).
//
// What limits this tile on MI355X is the panel stream into LDS (256 + 256
// rows x K per tile, one tile per CU at 4096^3).  Two measured facts shape the
// engine: 128-byte row segments per request stream about 1.7x faster than
// 64-byte ones, and the stream only reaches that rate with roughly two K-tiles
// (~96 KiB per CU) continuously in flight.  Three full 64 KiB stages do not
// fit in 160 KiB of LDS, so the two operands are staged differently:
//  * B: three LDS stages (96 KiB); K-contiguous B goes HBM -> LDS directly
//    (buffer_load_dwordx4 ... lds) two K-tiles ahead; M/N-contiguous B goes
//    through a 2-slot register ring and is transposed into LDS;
//  * A: two LDS stages (64 KiB) fed from a 2-slot register ring loaded two
//    K-tiles ahead (16-byte loads of whole 128-byte row segments, or the
//    transposing loader when A is M-contiguous);
//  * K = 64 per stage in 128-byte rows, XOR-swizzled 16-byte chunks;
//  * one barrier per K-tile; the MFMAs of each read step run while the
//    fragments of the next step load, and the last step of K-tile kt-1 runs
//    after the barrier.
#pragma once
#include "gemm_tile_v2.hpp"

namespace ffgemm {
namespace v3 {

using v2::DirectLoader;
using v2::lds_addr;
using v2::static_for;
using v2::TransLoader;
using v2::wait_vmcnt;

constexpr int MT = 256;
constexpr int RB = 128;          // bytes of K per LDS row per stage
constexpr int DU = 64;           // K per stage
constexpr int RS = 4;            // MFMA read steps per stage
constexpr int NSA = 2, NSB = 3;  // LDS stages of A and B
constexpr int STAGE = MT * RB;   // 32 KiB per operand stage
constexpr int LDS_BYTES = (NSA + NSB) * STAGE;

// K-contiguous operand through registers: thread t loads the 16-byte chunk
// t%8 of rows t/8 + 32i (whole 128-byte row segments per 8 lanes).
template <bool PERM, int PF>
struct RegLoader {
    static constexpr bool kDirect = false;
    static constexpr int NLD = MT * RB / (16 * kThreads);  // 8
    static constexpr int NVMEM = NLD;
    static constexpr int NSTORE = NLD;
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t voff[NLD];
    int32_t klim;
    int tid;  // thread index in the workgroup
    i32x4_t v[PF][NLD];

    __device__ inline void init(const void* base, uint32_t extent, int32_t ld, int32_t rows, int32_t K,
                                int32_t row0, int t) {
        rsrc = make_rsrc(base, extent);
        tid = t;
        const int c = t & 7;
        klim = K - c * 8;
#pragma unroll
        for (int i = 0; i < NLD; ++i) {
            const int r = row0 + (t >> 3) + 32 * i;
            voff[i] = (r < rows) ? (uint32_t)(((int64_t)r * ld + c * 8) * 2) : kOOB;
        }
    }
    template <int S>
    __device__ inline void issue(int kt, char*) {
        const bool ok = kt * DU < klim;
#pragma unroll
        for (int i = 0; i < NLD; ++i)
            v[S][i] = buf_load16(rsrc, (ok && voff[i] != kOOB) ? voff[i] + (uint32_t)kt * RB : kOOB);
    }
    template <int S>
    __device__ inline void store(char* region) const {
        const int t = tid, c = t & 7;
#pragma unroll
        for (int i = 0; i < NLD; ++i) {
            const int r = (t >> 3) + 32 * i;
            *(i32x4_t*)(region + lds_addr<RB>(PERM ? perm_a_row(r) : r, c * 16)) = v[S][i];
        }
    }
};

template <int NM, int NR, int SYNC, int MASK>
__device__ inline void sched_block(int nx) { v2::sched_block<NM, NR, SYNC, MASK>(nx); }

template <int NM, int NR, int NX, int SYNC>
__device__ inline void sched_mix() { v2::sched_mix<NM, NR, NX, SYNC>(); }

// Epilogue of a whole tile with beta = 0 (after the K loop, LDS free): the
// four waves' 128 x 128 blocks go to LDS as a column-major tile (column stride
// 528 bytes, free of bank conflicts), then every instruction stores whole
// 512-byte column pieces (8 columns per pass of 256 threads; fp32 D in two
// rounds of 128 rows), non-temporal.  The per-fragment stores write 16-byte
// pieces of 32 columns per instruction.  Dt: the tile's first element.
template <int OUT>
__device__ inline void store_tile_staged(const TileArgs& a, char* Dt, const f32x16_t (&acc)[4][4], char* smem,
                                         int tid) {
    constexpr int OS = (OUT == kOutF32) ? 4 : 2;
    constexpr int CS = 528;           // LDS bytes per staged column (512 + 16)
    constexpr int ROWS = 512 / OS;    // rows per staging round: 256 (16-bit), 128 (fp32)
    constexpr int NPG = MT / ROWS;    // staging rounds: 1 or 2
    constexpr int NT = 2;             // buffer store: non-temporal
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int g = lane >> 5;
    const float alpha = a.alpha;
    const bool trunc = OUT == kOutBF16 && (a.flags & kFlagTrunc) != 0;
    const __amdgpu_buffer_rsrc_t rd = make_rsrc(Dt, (uint32_t)((int64_t)(MT - 1) * a.ldd * OS + MT * OS));
#pragma unroll
    for (int pg = 0; pg < NPG; ++pg) {
        if (NPG == 1 || wm == pg) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const f32x16_t& x = acc[i][j];
                    const int ml = (NPG == 1 ? 128 * wm : 0) + 32 * i + 16 * g;
                    const int nl = 128 * wn + 32 * j + (lane & 31);
                    char* p = smem + nl * CS + ml * OS;
                    if constexpr (OUT == kOutF32) {
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            *(f32x4_t*)(p + 16 * q) = f32x4_t{alpha * x[4 * q], alpha * x[4 * q + 1],
                                                              alpha * x[4 * q + 2], alpha * x[4 * q + 3]};
                    } else {
#pragma unroll
                        for (int q = 0; q < 2; ++q) {
                            i32x4_t v;
#pragma unroll
                            for (int e = 0; e < 4; ++e) {
                                const float lo = alpha * x[8 * q + 2 * e], hi = alpha * x[8 * q + 2 * e + 1];
                                v[e] = (int)(OUT == kOutF16 ? pack_f16x2(lo, hi)
                                                            : (trunc ? pack_bf16x2_trunc(lo, hi) : pack_bf16x2_rne(lo, hi)));
                            }
                            *(i32x4_t*)(p + 16 * q) = v;
                        }
                    }
                }
        }
        lds_barrier();
#pragma unroll
        for (int pass = 0; pass < 32; ++pass) {
            const int col = 8 * pass + (tid >> 5), ch = tid & 31;
            const i32x4_t v = *(const i32x4_t*)(smem + col * CS + ch * 16);
            const uint32_t off = (uint32_t)(((int64_t)col * a.ldd + ROWS * pg) * OS + ch * 16);
            __builtin_amdgcn_raw_buffer_store_b128(v, rd, off, 0, NT);
        }
        if (pg + 1 < NPG) lds_barrier();
    }
}

// CPRE: preload C in the epilogue.  The launcher takes it only with beta != 0
// (the beta = 0 kernel keeps the register allocation it has without it).
template <int MODE, bool AK, bool BK, int OUT, bool CPRE = true>
__device__ inline void gemm_tile(const TileArgs& a, int tm, int tn, int b, char* smem, int tid) {
    static_assert(MODE == kBF16 || MODE == kF16, "16-bit engine");
    constexpr int WT = MT / 2, T = WT / 32;  // 128 x 128 per wave, 4 x 4 MFMA tiles
    // A: register ring of 2 slots (two K-tiles ahead) -> 2 LDS stages
    using LA = std::conditional_t<AK, RegLoader<true, 2>, TransLoader<MT, true, 2, RB>>;
    // B: direct into 3 LDS stages, or register ring -> 3 LDS stages
    using LB = std::conditional_t<BK, DirectLoader<MT, false, 2, RB>, TransLoader<MT, false, 2, RB>>;
    constexpr int NV = LA::NVMEM + LB::NVMEM;
    constexpr int NW = LA::NSTORE + LB::NSTORE;

    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int m0 = tm * MT, n0 = tn * MT;
    const int g = lane >> 5;

    const char* Ab = (const char*)a.A + (int64_t)b * a.bsA * 2;
    const char* Bb = (const char*)a.B + (int64_t)b * a.bsB * 2;
    LA la;
    LB lb;
    // K-contiguous operands: the buffer starts at the tile's first row
    if constexpr (AK) la.init(Ab + (int64_t)m0 * a.lda * 2, extent_bytes(a.M - m0, a.lda, a.K, 2), a.lda, a.M - m0, a.K, 0, tid);
    else la.init(Ab, extent_bytes(a.K, a.lda, a.M, 2), a.lda, a.M, a.K, m0, tid);
    if constexpr (BK) lb.init(Bb + (int64_t)n0 * a.ldb * 2, extent_bytes(a.N - n0, a.ldb, a.K, 2), a.ldb, a.N - n0, a.K, 0, tid);
    else lb.init(Bb, extent_bytes(a.K, a.ldb, a.N, 2), a.ldb, a.N, a.K, n0, tid);

    f32x16_t acc[T][T];
#pragma unroll
    for (int i = 0; i < T; ++i)
#pragma unroll
        for (int j = 0; j < T; ++j)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;
    i32x4_t fa[2][T], fb[2][T];
#pragma unroll
    for (int i = 0; i < T; ++i) {
        fa[1][i] = i32x4_t{0, 0, 0, 0};
        fb[1][i] = i32x4_t{0, 0, 0, 0};
    }

    const int arow = wm * WT + (lane & 31);
    const int brow = wn * WT + (lane & 31);
    const int KT = (a.K + DU - 1) / DU;

    char* const sA = smem;                  // NSA stages of A
    char* const sB = smem + NSA * STAGE;    // NSB stages of B
    auto stA = [&](int kt) { return sA + (kt & 1) * STAGE; };
    auto stB = [&](int kt) { return sB + (kt % NSB) * STAGE; };

    // prologue: K-tiles 0 and 1 in flight; K-tile 0 written to its stages
    la.template issue<0>(0, nullptr);
    lb.template issue<0>(0, stB(0));
    la.template issue<1>(1, nullptr);
    lb.template issue<1>(1, stB(1));
    la.template store<0>(stA(0));
    lb.template store<0>(stB(0));

    auto iteration = [&](auto slot, int kt) {
        constexpr int S = decltype(slot)::value;  // kt % 2
        const char* sa = stA(kt);
        const char* sb = stB(kt);
        // B's direct loads of K-tile kt were issued two iterations ago: the
        // loads of the last iteration and the A loads issued after them may
        // still be in flight
        if constexpr (BK) wait_vmcnt<NV + LA::NVMEM>();
        lds_barrier();  // K-tile kt complete in LDS; K-tile kt-1 fully consumed
        // K-tile kt+2: B into stage (kt+2)%3 (K-tile kt-1 is done) during the
        // first read step, A into register slot S (its K-tile kt is in LDS
        // now) during the second
        static_for<RS>([&](auto step) {
            constexpr int R = decltype(step)::value;
            constexpr int CUR = R & 1, PREV = (R + 1) & 1;
#pragma unroll
            for (int i = 0; i < T; ++i)
                fa[CUR][i] = *(const i32x4_t*)(sa + lds_addr<RB>(arow + 32 * i, (2 * R + g) * 16));
#pragma unroll
            for (int j = 0; j < T; ++j)
                fb[CUR][j] = *(const i32x4_t*)(sb + lds_addr<RB>(brow + 32 * j, (2 * R + g) * 16));
#pragma unroll
            for (int i = 0; i < T; ++i)
#pragma unroll
                for (int j = 0; j < T; ++j) v2::mfma16<MODE>(fa[PREV][i], fb[PREV][j], acc[i][j]);
            if constexpr (R == RS - 1) {
                // K-tile kt+1 from registers into its stages (A always; B when
                // it is transposed)
                la.template store<S ^ 1>(stA(kt + 1));
                lb.template store<S ^ 1>(stB(kt + 1));
                sched_block<T * T, 2 * T, R + 1, 0x200>(NW);
            } else if constexpr (R == 0) {
                lb.template issue<S>(kt + 2, stB(kt + 2));
                sched_mix<T * T, 2 * T, LB::NVMEM, R + 1>();
            } else if constexpr (R == 1) {
                la.template issue<S>(kt + 2, nullptr);
                sched_mix<T * T, 2 * T, LA::NVMEM, R + 1>();
            } else {
                sched_block<T * T, 2 * T, R + 1, 0x020>(0);
            }
        });
    };

    int kt = 0;
    for (; kt + 2 <= KT; kt += 2) {
        iteration(std::integral_constant<int, 0>{}, kt);
        iteration(std::integral_constant<int, 1>{}, kt + 1);
    }
    if (kt < KT) iteration(std::integral_constant<int, 0>{}, kt);
#pragma unroll
    for (int i = 0; i < T; ++i)
#pragma unroll
        for (int j = 0; j < T; ++j) v2::mfma16<MODE>(fa[1][i], fb[1][j], acc[i][j]);
    // drain the loads issued past the end before the LDS is reused
    wait_vmcnt<0>();
    lds_barrier();

    const char* Cb = (const char*)a.C + (int64_t)b * a.bsC * ((OUT == kOutF32) ? 4 : 2);
    char* Db = (char*)a.D + (int64_t)b * a.bsD * ((OUT == kOutF32) ? 4 : 2);
    // a whole tile with beta = 0 leaves through LDS as whole columns
    if constexpr (!CPRE) {
        constexpr int OS = (OUT == kOutF32) ? 4 : 2;
        if (m0 + MT <= a.M && n0 + MT <= a.N && (a.flags & kFlagBeta) == 0 &&
            (int64_t)(MT - 1) * a.ldd * OS + MT * OS < 0x7ffffff0ll) {
            store_tile_staged<OUT>(a, Db + ((int64_t)n0 * a.ldd + m0) * OS, acc, smem, tid);
            return;
        }
    }
    // 16-bit C with beta != 0: every fragment's C slice is requested before the
    // first is used (the loader and fragment registers are free by now); a
    // fragment not wholly inside D reads C's first rows instead and stores
    // element by element
    const bool pre = CPRE && OUT != kOutF32 && (a.flags & kFlagBeta) != 0 && a.M >= 16;
    i32x4_t cpre[T][T][2];
    if constexpr (CPRE && OUT != kOutF32) {
        if (pre) {
#pragma unroll
            for (int i = 0; i < T; ++i)
#pragma unroll
                for (int j = 0; j < T; ++j) {
                    const int m = m0 + wm * WT + 32 * i + 16 * g, n = n0 + wn * WT + 32 * j + (lane & 31);
                    const bool full = n < a.N && m + 16 <= a.M;
                    const char* cp = full ? Cb + ((int64_t)n * a.ldc + m) * 2 : Cb;
                    cpre[i][j][0] = *(const i32x4_t*)cp;
                    cpre[i][j][1] = *(const i32x4_t*)(cp + 16);
                }
        }
    }
#pragma unroll
    for (int i = 0; i < T; ++i) {
        const int m = m0 + wm * WT + 32 * i + 16 * g;
#pragma unroll
        for (int j = 0; j < T; ++j) {
            const int n = n0 + wn * WT + 32 * j + (lane & 31);
            if (n < a.N && m < a.M) {
                float v[16];
#pragma unroll
                for (int e = 0; e < 16; ++e) v[e] = acc[i][j][e];
                if constexpr (CPRE && OUT != kOutF32) {
                    if (pre && m + 16 <= a.M) {
                        store16_pre<OUT>(a, Db, m, n, v, cpre[i][j][0], cpre[i][j][1], a.alpha, a.beta);
                        continue;
                    }
                }
                store16<OUT>(a, Cb, Db, m, n, v, a.alpha, a.beta);
            }
        }
    }
}

// Persistent form for the grouped launch: runs tiles begin .. end-1 (tile_at(i,
// a, tm, tn) names each; batch 0) as one stream of K-tiles.  The ring keeps
// two K-tiles in flight across tile boundaries: the last two K-tiles of a tile
// issue the loads of the next tile's first two, so the memory latency at a
// tile start overlaps the previous tile's last K-tiles instead of following
// its epilogue.  Each tile runs an even number of K-tiles, at least two (a
// padding K-tile loads only zeros), so every tile starts on register slot 0
// and A stage 0.  C is read fragment by fragment.
template <int MODE, bool AK, bool BK, int OUT, class TileAt>
__device__ inline void gemm_tiles_stream(TileAt&& tile_at, int begin, int end, char* smem, int tid) {
    static_assert(MODE == kBF16 || MODE == kF16, "16-bit engine");
    constexpr int WT = MT / 2, T = WT / 32;
    using LA = std::conditional_t<AK, RegLoader<true, 2>, TransLoader<MT, true, 2, RB>>;
    using LB = std::conditional_t<BK, DirectLoader<MT, false, 2, RB>, TransLoader<MT, false, 2, RB>>;
    constexpr int NV = LA::NVMEM + LB::NVMEM;
    constexpr int NW = LA::NSTORE + LB::NSTORE;
    if (begin >= end) return;

    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int g = lane >> 5;
    const int arow = wm * WT + (lane & 31);
    const int brow = wn * WT + (lane & 31);
    char* const sA = smem;
    char* const sB = smem + NSA * STAGE;
    // B's stage by position s in the stream of K-tiles
    auto stB = [&](int s) { return sB + (s % NSB) * STAGE; };
    auto ktiles = [](const TileArgs* a) { return max(2, ((a->K + 2 * DU - 1) / (2 * DU)) * 2); };

    // the loaders address the tile `ia` (the compute tile, or the next one
    // during the compute tile's last two K-tiles)
    LA la;
    LB lb;
    int itm, itn;
    const TileArgs* ia;
    auto init_loaders = [&]() {
        const TileArgs& a = *ia;
        const int t = tid;
        const int m0 = itm * MT, n0 = itn * MT;
        const char* Ab = (const char*)a.A;
        const char* Bb = (const char*)a.B;
        if constexpr (AK) la.init(Ab + (int64_t)m0 * a.lda * 2, extent_bytes(a.M - m0, a.lda, a.K, 2), a.lda, a.M - m0, a.K, 0, t);
        else la.init(Ab, extent_bytes(a.K, a.lda, a.M, 2), a.lda, a.M, a.K, m0, t);
        if constexpr (BK) lb.init(Bb + (int64_t)n0 * a.ldb * 2, extent_bytes(a.N - n0, a.ldb, a.K, 2), a.ldb, a.N - n0, a.K, 0, t);
        else lb.init(Bb, extent_bytes(a.K, a.ldb, a.N, 2), a.ldb, a.N, a.K, n0, t);
    };
    int ci = begin;
    tile_at(ci, ia, itm, itn);
    init_loaders();
    const TileArgs* ca = ia;
    int ctm = itm, ctn = itn, cKT = ktiles(ca);

    f32x16_t acc[T][T];
    i32x4_t fa[2][T], fb[2][T];
    auto clear = [&]() {
#pragma unroll
        for (int i = 0; i < T; ++i)
#pragma unroll
            for (int j = 0; j < T; ++j)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;
#pragma unroll
        for (int i = 0; i < T; ++i) {
            fa[1][i] = i32x4_t{0, 0, 0, 0};
            fb[1][i] = i32x4_t{0, 0, 0, 0};
        }
    };
    clear();

    int s = 0;
    la.template issue<0>(0, nullptr);
    lb.template issue<0>(0, stB(0));
    la.template issue<1>(1, nullptr);
    lb.template issue<1>(1, stB(1));
    la.template store<0>(sA);
    lb.template store<0>(stB(0));

    // the K-tile at stream position s (register slot and A stage S = s & 1),
    // as one iteration of gemm_tile's loop; the loaders fetch their tile's
    // K-tile kl (masked when kl is past K)
    auto step = [&](auto slot, int kl) {
        constexpr int S = decltype(slot)::value;
        const char* sa = sA + S * STAGE;
        const char* sb = stB(s);
        if constexpr (BK) wait_vmcnt<NV + LA::NVMEM>();
        lds_barrier();
        static_for<RS>([&](auto rstep) {
            constexpr int R = decltype(rstep)::value;
            constexpr int CUR = R & 1, PREV = (R + 1) & 1;
#pragma unroll
            for (int i = 0; i < T; ++i)
                fa[CUR][i] = *(const i32x4_t*)(sa + lds_addr<RB>(arow + 32 * i, (2 * R + g) * 16));
#pragma unroll
            for (int j = 0; j < T; ++j)
                fb[CUR][j] = *(const i32x4_t*)(sb + lds_addr<RB>(brow + 32 * j, (2 * R + g) * 16));
#pragma unroll
            for (int i = 0; i < T; ++i)
#pragma unroll
                for (int j = 0; j < T; ++j) v2::mfma16<MODE>(fa[PREV][i], fb[PREV][j], acc[i][j]);
            if constexpr (R == RS - 1) {
                // position s+1 (the next tile's first K-tile after a tile's
                // last) from registers into its stages
                la.template store<S ^ 1>(sA + (S ^ 1) * STAGE);
                lb.template store<S ^ 1>(stB(s + 1));
                sched_block<T * T, 2 * T, R + 1, 0x200>(NW);
            } else if constexpr (R == 0) {
                lb.template issue<S>(kl, stB(s + 2));
                sched_mix<T * T, 2 * T, LB::NVMEM, R + 1>();
            } else if constexpr (R == 1) {
                la.template issue<S>(kl, nullptr);
                sched_mix<T * T, 2 * T, LA::NVMEM, R + 1>();
            } else {
                sched_block<T * T, 2 * T, R + 1, 0x020>(0);
            }
        });
        ++s;
    };

    while (true) {
        const bool more = ci + 1 < end;
        for (int k = 0; k < cKT; k += 2) {
            // the last two K-tiles load the next tile's first two (past the
            // last tile, K-tiles beyond K: masked)
            int kl = k + 2;
            if (kl == cKT && more) {
                tile_at(ci + 1, ia, itm, itn);
                init_loaders();
                kl = 0;
            }
            step(std::integral_constant<int, 0>{}, kl);
            step(std::integral_constant<int, 1>{}, kl + 1);
        }
#pragma unroll
        for (int i = 0; i < T; ++i)
#pragma unroll
            for (int j = 0; j < T; ++j) v2::mfma16<MODE>(fa[1][i], fb[1][j], acc[i][j]);
        {
            const TileArgs& a = *ca;
            const char* Cb = (const char*)a.C;
            char* Db = (char*)a.D;
#pragma unroll
            for (int i = 0; i < T; ++i) {
                const int m = ctm * MT + wm * WT + 32 * i + 16 * g;
#pragma unroll
                for (int j = 0; j < T; ++j) {
                    const int n = ctn * MT + wn * WT + 32 * j + (lane & 31);
                    if (n < a.N && m < a.M) {
                        float v[16];
#pragma unroll
                        for (int e = 0; e < 16; ++e) v[e] = acc[i][j][e];
                        store16<OUT>(a, Cb, Db, m, n, v, a.alpha, a.beta);
                    }
                }
            }
        }
        if (!more) break;
        // With stores and loads both pending, the compiler takes the memory
        // counter as out of order and drains it to zero at the next wait it
        // needs, which would fall inside the next tile's K loop; draining it
        // here instead costs only the remaining latency of the next tile's
        // second K-tile, issued one K-tile before the epilogue.
        wait_vmcnt<0>();
        ++ci;
        ca = ia;
        ctm = itm;
        ctn = itn;
        cKT = ktiles(ca);
        clear();
    }
    // no LDS-bound load may be in flight when the workgroup ends
    wait_vmcnt<0>();
}

template <int MODE, bool AK, bool BK, int OUT, bool CPRE>
__global__ void __launch_bounds__(kThreads, 1) gemm_kernel(TileArgs a) {
    __shared__ __attribute__((aligned(16))) char smem[LDS_BYTES];
    int tb, tm, tn;
    map_tile(blockIdx.x, gridDim.x, a, tb, tm, tn);
    v3::gemm_tile<MODE, AK, BK, OUT, CPRE>(a, tm, tn, tb, smem, threadIdx.x);
}

}  // namespace v3
}  // namespace ffgemm
