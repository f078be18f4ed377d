// Pipelined tile engine for 16-bit operands (bf16 / fp16).
//
// Measured on MI355X (4096^3, 256 x 256 tiles, one workgroup per CU): the
// panel stream HBM/MALL -> L2 -> LDS is the bottleneck, and its throughput
// depends on the request shape: a CU fetching whole 128-byte row segments
// sustains 25 TB/s aggregate (LDS-direct), 64-byte segments only 14.7 TB/s.
// Design:
//  * LDS rows of RB bytes of K per stage (64 here; the 256 x 256 engine in
//    gemm_tile_v3.hpp uses 128-byte rows), chunk index XOR-swizzled per row
//    so 16-byte MFMA operand reads and 16-byte writes are conflict-free;
//  * K-contiguous operands go straight from HBM to LDS
//    (buffer_load_dwordx4 ... lds), PF K-tiles ahead; each lane fetches the
//    16-byte chunk whose swizzled LDS slot is lane*16, so the swizzle is free;
//  * M/N-contiguous operands are loaded PF K-tiles ahead into a register ring
//    and transposed into LDS (v_perm with per-lane selectors; row order
//    rotated by strip parity so the stores are conflict-free) one K-tile ahead;
//  * one barrier per K-tile; every read step's fragments are loaded while the
//    MFMAs of the previous step run, and the last step of K-tile kt-1 runs
//    after the barrier, while the first fragments of K-tile kt load.
// Loads beyond K, M or N read zero (buffer range checks), so the loop body
// has no data-dependent branches.
#pragma once
#include <type_traits>

#include "gemm_tile.hpp"

namespace ffgemm {
namespace v2 {

// Per-tile pipeline configuration: 64-byte rows (K = 32 per stage), 3-4
// stages; two workgroups per CU where the registers allow, so one group's
// MFMAs cover the other's waits.  Tiles with both operands K-contiguous
// (TN) take 128-byte rows (K = 64 per stage): twice the K in flight and half
// the barriers.  Measured on MI355X: 64 x 64 fp16 / bf16 4096 x 512 x 4096
// 34.9 -> 28.3 us, the one-round shapes of 512 x 512 x 512 to 512 x 2048 x
// 2048 within 2 %; 128 x 128, now one workgroup per CU (128 KiB of LDS), 1-4 %
// faster from 2048^3 to 4096^3 and on 480 x 8192 x 8192; 256 x 128 and 128 x
// 256, now one workgroup per CU with three stages (144 KiB), 2-6 % faster
// (fp16 8192 x 1024 x 8192 143.6 -> 135.7 us).  RBX > 0 sets the row bytes
// instead.
template <int MT_M, int MT_N, bool AK, bool BK, int RBX = 0>
struct Cfg {
    static constexpr bool kBig = (MT_M == 256 && MT_N == 256);
    static constexpr bool kSmall = (MT_M == 128 && MT_N == 128);
    // (256 x 256 with a K-contiguous B runs the split-ring engine instead;
    // with an M/N-contiguous B, 64-byte rows and four stages measured best)
    static constexpr int RB = RBX ? RBX : (!kBig && AK && BK ? 128 : 64);  // LDS bytes per row per stage
    static constexpr int DU = RB / 2;                  // K per stage (16-bit)
    static constexpr int RS = DU / 16;                 // MFMA read steps per stage
    // two workgroups per CU where the registers allow it without spills: the
    // 128 x 128 tiles, both-K-contiguous tiles, and the two shapes whose one
    // M/N-contiguous operand is the 128-row side
    static constexpr bool kTwoFit = kSmall || (AK && BK) || (MT_M == 128 && MT_N == 256 && !AK && BK) ||
                                    (MT_M == 256 && MT_N == 128 && AK && !BK);
    // (one, with 128-byte rows above 64 x 64: their stages fill most of the LDS)
    static constexpr int OCC = kBig ? 1 : ((kTwoFit && !(MT_M + MT_N > 128 && RB == 128)) ? 2 : 1);
    // four stages (three K-tiles in flight), except the two-per-CU 128 x 256 /
    // 256 x 128 shapes and those shapes with 128-byte rows (four would exceed
    // the LDS); 64 x 64, whose small problems are a latency chain of K-tiles,
    // measured 7-21 % faster with four than with three
    static constexpr int NS = kBig ? 4 : (((OCC == 2 && !kSmall && MT_M > 64) || (RB == 128 && MT_M + MT_N > 256)) ? 3 : 4);
    static constexpr int PF = NS - 1;                  // K-tiles in flight
};

// 16-byte chunk swizzle of a stage row.  For 64-byte rows the fragment reads
// (32 consecutive rows at one chunk per half-wave) are conflict-free with
// bits 2-4 of the row (SQ_LDS_BANK_CONFLICT 0 on gfx950); the earlier
// ((r >> 1) ^ (r >> 3)) & 3 cost 4 conflict cycles per ds_read_b128.
template <int RB>
__host__ __device__ constexpr int swz(int r) {
    return RB == 128 ? ((r ^ (r >> 3)) & 7) : (((r >> 2) ^ (r >> 3)) & 3);
}
template <int RB>
__host__ __device__ constexpr int lds_addr(int row, int off) {
    return row * RB + ((((off >> 4) ^ swz<RB>(row)) & (RB / 16 - 1)) << 4) + (off & 15);
}
// Inverse of perm_a_row: LDS row -> row of the A tile it holds.
__host__ __device__ constexpr int perm_a_inv(int lr) {
    int i = lr & 31;
    int g = (i >> 3) & 3, h = (i >> 2) & 1, q = i & 3;
    return (lr & ~31) | (h << 4) | (g << 2) | q;
}

typedef __attribute__((address_space(3))) void* lds_ptr_t;

// K-contiguous operand: HBM -> LDS directly, 16 bytes per lane.  One
// instruction fills 1 KiB of LDS = 1024/RB rows.
template <int MT, bool PERM, int PF, int RB>
struct DirectLoader {
    static constexpr bool kDirect = true;
    static constexpr int DU = RB / 2;
    static constexpr int LPR = RB / 16;              // lanes per row
    static constexpr int NI = MT * RB / 4096;        // instructions per wave per K-tile
    static constexpr int NVMEM = NI;
    static constexpr int NSTORE = 0;
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t voff[NI];
    int32_t klim[NI];  // K - (first K of this lane's chunk); masked rows: 0
    int wave;          // wave-uniform (SGPR)

    __device__ inline void init(const void* base, uint32_t extent, int32_t ld, int32_t rows, int32_t K,
                                int32_t row0, int tid) {
        rsrc = make_rsrc(base, extent);
        wave = __builtin_amdgcn_readfirstlane(tid >> 6);
        const int lane = tid & 63;
#pragma unroll
        for (int j = 0; j < NI; ++j) {
            const int lr = (wave * NI + j) * (64 / LPR) + lane / LPR;
            const int c = (lane % LPR) ^ swz<RB>(lr);
            const int r = row0 + (PERM ? perm_a_inv(lr) : lr);
            klim[j] = (r < rows) ? K - c * 8 : 0;
            voff[j] = (uint32_t)(((int64_t)r * ld + c * 8) * 2);
        }
    }
    template <int S>
    __device__ inline void issue(int kt, char* region) {
        char* dst = region + wave * (NI * 1024);
#pragma unroll
        for (int j = 0; j < NI; ++j) {
            const uint32_t off = (kt * DU < klim[j]) ? voff[j] + (uint32_t)kt * RB : kOOB;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_ptr_t)(dst + j * 1024), 16, off, 0, 0, 0);
        }
    }
    template <int S>
    __device__ inline void store(char*) const {}
};

// M/N-contiguous operand: 16-byte loads of 8 consecutive M (N) values for KB
// consecutive K, held in a PF-slot ring, transposed on the way to LDS.
template <int MT, bool PERM, int PF, int RB>
struct TransLoader {
    static constexpr bool kDirect = false;
    static constexpr int DU = RB / 2;
    static constexpr int KB = DU * MT / 2048;        // K rows per thread per K-tile
    static constexpr int NKG = DU / KB;              // K groups across consecutive lanes
    static constexpr int NVMEM = KB;
    static constexpr int NSTORE = 8;
    static_assert(KB == 2 || KB == 4 || KB == 8, "transposing loader geometry");
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t voff[KB];
    int32_t klim[KB];
    uint32_t kstep;
    int mc, kg;
    uint32_t sel0, sel1;
    i32x4_t v[PF][KB];

    __device__ inline void init(const void* base, uint32_t extent, int32_t ld, int32_t rows, int32_t K,
                                int32_t row0, int t) {
        rsrc = make_rsrc(base, extent);
        kg = t % NKG;
        mc = t / NKG;
        const int r = row0 + mc * 8;
#pragma unroll
        for (int j = 0; j < KB; ++j) {
            const int k = kg * KB + j;
            klim[j] = (r < rows) ? K - k : 0;
            voff[j] = (uint32_t)(((int64_t)k * ld + r) * 2);
        }
        kstep = (uint32_t)(DU * ld * 2);
        // v_perm selectors: low halves (0x05040100) or high halves (0x07060302)
        // of two dwords; odd strips swap them because their rows are rotated.
        sel0 = (mc & 1) ? 0x07060302u : 0x05040100u;
        sel1 = (mc & 1) ? 0x05040100u : 0x07060302u;
    }
    template <int S>
    __device__ inline void issue(int kt, char*) {
#pragma unroll
        for (int j = 0; j < KB; ++j)
            v[S][j] = buf_load16(rsrc, (kt * DU < klim[j]) ? voff[j] + (uint32_t)kt * kstep : kOOB);
    }
    // Row mc*8 + (i ^ (mc&1)) of the tile receives element i ^ (mc&1) of each
    // of the KB loaded vectors.
    template <int S>
    __device__ inline void store(char* region) const {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int row = mc * 8 + (i ^ (mc & 1));
            const int lr = PERM ? perm_a_row(row) : row;
            const uint32_t sel = (i & 1) ? sel1 : sel0;
            uint32_t w[KB / 2];
#pragma unroll
            for (int p = 0; p < KB / 2; ++p)
                w[p] = __builtin_amdgcn_perm((uint32_t)v[S][2 * p + 1][i >> 1], (uint32_t)v[S][2 * p][i >> 1], sel);
            char* dst = region + lds_addr<RB>(lr, kg * KB * 2);
            if constexpr (KB == 8) *(i32x4_t*)dst = i32x4_t{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
            else if constexpr (KB == 4) *(i32x2_t*)dst = i32x2_t{(int)w[0], (int)w[1]};
            else *(int*)dst = (int)w[0];
        }
    }
};

// Compile-time loop: f(integral_constant<int, 0>) ... f(integral_constant<int, N-1>).
template <int I, int N, class F>
__device__ inline void static_for_impl(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for_impl<I + 1, N>(f);
    }
}
template <int N, class F>
__device__ inline void static_for(F&& f) { static_for_impl<0, N>(f); }

// s_waitcnt vmcnt(N) with lgkmcnt/expcnt left free (gfx9 encoding).
template <int N>
__device__ inline void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

// Interleave for one block of NM MFMAs: the NR LDS reads one after each of
// the first MFMAs, then `nx` instructions of class MASK (VMEM loads 0x020 or
// LDS writes 0x200) one after each of the following MFMAs.
template <int NM, int NR, int SYNC, int MASK>
__device__ inline void sched_block(int nx) {
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, SYNC);
        if (i < NR) __builtin_amdgcn_sched_group_barrier(0x100, 1, SYNC);
        else if (i - NR < nx) __builtin_amdgcn_sched_group_barrier(MASK, 1, SYNC);
    }
    __builtin_amdgcn_sched_barrier(0);
}

// One block of NM MFMAs with the NR LDS reads after the even MFMAs and NX VMEM
// loads after the odd ones.  Issuing a load stalls its wave while the other
// waves' loads are queued at the texture addresser, so the split-ring engine
// spreads a K-tile's loads over two such blocks instead of packing them after
// one block's reads.
template <int NM, int NR, int NX, int SYNC>
__device__ inline void sched_mix() {
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, SYNC);
        if (i % 2 == 0 && i / 2 < NR) __builtin_amdgcn_sched_group_barrier(0x100, 1, SYNC);
        else if (i % 2 == 1 && i / 2 < NX) __builtin_amdgcn_sched_group_barrier(0x020, 1, SYNC);
    }
    __builtin_amdgcn_sched_barrier(0);
}

template <int MODE>
__device__ inline void mfma16(const i32x4_t& a, const i32x4_t& b, f32x16_t& acc) {
    if constexpr (MODE == kBF16)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b),
                                                      acc, 0, 0, 0);
    else
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8_t, a), __builtin_bit_cast(f16x8_t, b),
                                                     acc, 0, 0, 0);
}

template <int MODE, bool AK, bool BK, int MT_M, int MT_N, int OUT, int RBX = 0>
__device__ inline void gemm_tile(const TileArgs& a, int tm, int tn, int b, char* smem, int tid) {
    static_assert(MODE == kBF16 || MODE == kF16, "16-bit engine");
    using C = Cfg<MT_M, MT_N, AK, BK, RBX>;
    constexpr int RB = C::RB, DU = C::DU, RS = C::RS, NS = C::NS, PF = C::PF;
    constexpr int WT_M = MT_M / 2, WT_N = MT_N / 2;
    constexpr int TM = WT_M / 32, TN = WT_N / 32;
    constexpr int STAGE = (MT_M + MT_N) * RB;
    constexpr int BOFF = MT_M * RB;
    using LA = std::conditional_t<AK, DirectLoader<MT_M, true, PF, RB>, TransLoader<MT_M, true, PF, RB>>;
    using LB = std::conditional_t<BK, DirectLoader<MT_N, false, PF, RB>, TransLoader<MT_N, false, PF, RB>>;
    constexpr int NV = LA::NVMEM + LB::NVMEM;        // VMEM instructions per K-tile
    constexpr int NW = LA::NSTORE + LB::NSTORE;      // transposed LDS stores per K-tile
    static_assert(RS % 2 == 0, "read steps pair up in the fragment double buffer");

    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int m0 = tm * MT_M, n0 = tn * MT_N;
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

    f32x16_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;
    i32x4_t fa[2][TM], fb[2][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) fa[1][i] = i32x4_t{0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < TN; ++j) fb[1][j] = i32x4_t{0, 0, 0, 0};

    const int arow = wm * WT_M + (lane & 31);
    const int brow = wn * WT_N + (lane & 31);
    const int KT = (a.K + DU - 1) / DU;

    auto stage = [&](int kt) { return smem + (kt % NS) * STAGE; };

    // prologue: K-tiles 0..PF-1 in flight, K-tile 0 transposed into stage 0
    static_for<PF>([&](auto j) {
        constexpr int J = decltype(j)::value;
        la.template issue<J>(J, stage(J));
        lb.template issue<J>(J, stage(J) + BOFF);
    });
    la.template store<0>(stage(0));
    lb.template store<0>(stage(0) + BOFF);

    auto iteration = [&](auto slot, int kt) {
        constexpr int S = decltype(slot)::value;
        const char* sa = stage(kt);
        const char* sb = sa + BOFF;
        wait_vmcnt<(PF - 1) * NV>();  // K-tile kt's direct loads have landed (this wave)
        lds_barrier();                // ... and every wave's; stage kt-1 is free again
        la.template issue<S>(kt + PF, stage(kt + PF));
        lb.template issue<S>(kt + PF, stage(kt + PF) + BOFF);
        static_for<RS>([&](auto step) {
            constexpr int R = decltype(step)::value;
            constexpr int CUR = R & 1, PREV = (R + 1) & 1;
            // fragments of read step R: lane group g takes chunk 2R + g
#pragma unroll
            for (int i = 0; i < TM; ++i)
                fa[CUR][i] = *(const i32x4_t*)(sa + lds_addr<RB>(arow + 32 * i, (2 * R + g) * 16));
#pragma unroll
            for (int j = 0; j < TN; ++j)
                fb[CUR][j] = *(const i32x4_t*)(sb + lds_addr<RB>(brow + 32 * j, (2 * R + g) * 16));
            // MFMAs of the previous step (for R = 0: the last step of K-tile
            // kt-1, zeros on the first iteration)
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) mfma16<MODE>(fa[PREV][i], fb[PREV][j], acc[i][j]);
            if constexpr (R == RS - 1) {
                // K-tile kt+1 (register path) into its stage
                la.template store<(S + 1) % PF>(stage(kt + 1));
                lb.template store<(S + 1) % PF>(stage(kt + 1) + BOFF);
                sched_block<TM * TN, TM + TN, R + 1, 0x200>(NW);
            } else if constexpr (R == 0) {
                // (sched_mix here measured 0.3-0.7 % slower for TT at 256 x 256)
                sched_block<TM * TN, TM + TN, R + 1, 0x020>(NV);
            } else {
                sched_block<TM * TN, TM + TN, R + 1, 0x020>(0);
            }
        });
    };

    int kt = 0;
    for (; kt + PF <= KT; kt += PF)
        static_for<PF>([&](auto j) { iteration(j, kt + decltype(j)::value); });
    static_for<PF - 1>([&](auto j) {
        if (kt < KT) iteration(j, kt++);
    });
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) mfma16<MODE>(fa[(RS - 1) & 1][i], fb[(RS - 1) & 1][j], acc[i][j]);
    // drain the loads issued for K-tiles past the end before the LDS is reused
    wait_vmcnt<0>();
    lds_barrier();

    const char* Cb = (const char*)a.C + (int64_t)b * a.bsC * ((OUT == kOutF32) ? 4 : 2);
    char* Db = (char*)a.D + (int64_t)b * a.bsD * ((OUT == kOutF32) ? 4 : 2);
#pragma unroll
    for (int i = 0; i < TM; ++i) {
        const int m = m0 + wm * WT_M + 32 * i + 16 * g;
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int n = n0 + wn * WT_N + 32 * j + (lane & 31);
            if (n < a.N && m < a.M) {
                float v[16];
#pragma unroll
                for (int e = 0; e < 16; ++e) v[e] = acc[i][j][e];
                store16<OUT>(a, Cb, Db, m, n, v, a.alpha, a.beta);
            }
        }
    }
}

// LDS bytes the engine needs for a tile.
template <int MT_M, int MT_N, bool AK, bool BK, int RBX = 0>
__host__ __device__ constexpr int smem_bytes() {
    using C = Cfg<MT_M, MT_N, AK, BK, RBX>;
    return C::NS * (MT_M + MT_N) * C::RB;
}

template <int MODE, bool AK, bool BK, int MT_M, int MT_N, int OUT>
__global__ void __launch_bounds__(kThreads, (Cfg<MT_M, MT_N, AK, BK>::OCC)) gemm_kernel(TileArgs a) {
    __shared__ __attribute__((aligned(16))) char smem[smem_bytes<MT_M, MT_N, AK, BK>()];
    int tb, tm, tn;
    map_tile(blockIdx.x, gridDim.x, a, tb, tm, tn);
    v2::gemm_tile<MODE, AK, BK, MT_M, MT_N, OUT>(a, tm, tn, tb, smem, threadIdx.x);
}

}  // namespace v2
}  // namespace ffgemm
