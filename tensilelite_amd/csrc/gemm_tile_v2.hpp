// Deep-pipelined tile engine for 16-bit operands (bf16 / fp16).
//
// Why a second engine: with one workgroup of 4 waves per CU (one wave per
// SIMD, 256 accumulator registers each), nothing hides global-load latency
// except the software pipeline itself.  Measured on the first engine (one
// K-tile of register prefetch): MFMA busy 43%, LDS active 30%, HBM far from
// its roofline -- the waves were waiting on loads.  This engine keeps three
// K-tiles in flight:
//  * K = 32 per iteration, LDS ring of 4 stages of (MT_M + MT_N) rows x 64 B
//    (128 KiB for 256 x 256), chunk index XOR-swizzled per row;
//  * K-contiguous operands go straight from HBM to LDS
//    (buffer_load_dwordx4 ... lds), issued three K-tiles ahead; each lane
//    fetches the 16-byte chunk whose swizzled LDS slot is lane*16, so the
//    swizzle costs nothing;
//  * M/N-contiguous operands are loaded three K-tiles ahead into a 3-slot
//    register ring and transposed into LDS (v_perm with per-lane selectors,
//    row order rotated by strip parity so the stores are conflict-free) one
//    K-tile ahead;
//  * one barrier per K-tile; the MFMAs of the second half of K-tile kt-1 are
//    issued after the barrier, while the fragments of K-tile kt load.
// Loads beyond K, M or N read zero (buffer range checks), so the loop body
// has no data-dependent branches.
#pragma once
#include <type_traits>

#include "gemm_tile.hpp"

namespace ffgemm {
namespace v2 {

constexpr int RB = 64;  // LDS bytes per row per stage
constexpr int DU = 32;  // K per iteration (16-bit elements)

// Pipeline depth and residency per macro tile.  NS LDS stages keep PF = NS-1
// K-tiles in flight.  256 x 256 fills a CU alone (256 accumulators per lane,
// 128 KiB LDS); the smaller tiles run two workgroups per CU so one group's
// MFMAs cover the other's barrier and LDS waits.
// A 256 x 128 tile with a transposing operand needs its register ring and
// stays at one workgroup per CU (four stages); with two direct operands it
// fits two per CU (three stages, 2 x 72 KiB).
template <int MT_M, int MT_N, bool AK, bool BK>
struct Cfg {
    static constexpr bool kBig = (MT_M == 256 && MT_N == 256);
    static constexpr bool kSmall = (MT_M == 128 && MT_N == 128);
    static constexpr int OCC = kBig ? 1 : ((kSmall || (AK && BK)) ? 2 : 1);
    static constexpr int NS = (OCC == 2 && !kSmall) ? 3 : 4;
    static constexpr int PF = NS - 1;
};

__host__ __device__ constexpr int swz4(int r) { return ((r >> 1) ^ (r >> 3)) & 3; }
__host__ __device__ constexpr int lds_addr(int row, int off) {
    return row * RB + ((((off >> 4) ^ swz4(row)) & 3) << 4) + (off & 15);
}
// Inverse of perm_a_row: LDS row -> row of the A tile it holds.
__host__ __device__ constexpr int perm_a_inv(int lr) {
    int i = lr & 31;
    int g = (i >> 3) & 3, h = (i >> 2) & 1, q = i & 3;
    return (lr & ~31) | (h << 4) | (g << 2) | q;
}

typedef __attribute__((address_space(3))) void* lds_ptr_t;

// K-contiguous operand: HBM -> LDS directly, 16 bytes per lane.
template <int MT, bool PERM, int PF>
struct DirectLoader {
    static constexpr bool kDirect = true;
    static constexpr int NI = MT / 64;  // buffer_load ... lds per wave per K-tile
    static constexpr int NVMEM = NI;
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t voff[NI];
    int32_t klim[NI];  // K - (first K of this lane's chunk); masked rows: 0
    int wave;          // wave-uniform (SGPR)

    __device__ inline void init(const void* base, uint32_t extent, int32_t ld, int32_t rows, int32_t K,
                                int32_t row0) {
        rsrc = make_rsrc(base, extent);
        wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int j = 0; j < NI; ++j) {
            const int lr = (wave * NI + j) * 16 + (lane >> 2);
            const int c = (lane & 3) ^ swz4(lr);
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
            const uint32_t off = (kt * DU < klim[j]) ? voff[j] + (uint32_t)kt * (DU * 2) : kOOB;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_ptr_t)(dst + j * 1024), 16, off, 0, 0, 0);
        }
    }
    template <int S>
    __device__ inline void store(char*) const {}
};

// M/N-contiguous operand: 16-byte loads of 8 consecutive M (N) values for
// KB consecutive K, held in a 3-slot ring, transposed on the way to LDS.
template <int MT, bool PERM, int PF>
struct TransLoader {
    static constexpr bool kDirect = false;
    static constexpr int KB = MT / 64;      // K rows per thread per K-tile
    static constexpr int NKG = DU / KB;     // K groups across consecutive lanes
    static constexpr int NVMEM = KB;
    static_assert(KB == 2 || KB == 4, "transposing loader needs MT of 128 or 256");
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t voff[KB];
    int32_t klim[KB];
    uint32_t kstep;
    int mc, kg;
    uint32_t sel0, sel1;
    i32x4_t v[PF][KB];

    __device__ inline void init(const void* base, uint32_t extent, int32_t ld, int32_t rows, int32_t K,
                                int32_t row0) {
        rsrc = make_rsrc(base, extent);
        const int t = threadIdx.x;
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
    // Row mc*8 + (i ^ (mc&1)) of the tile receives element i ^ (mc&1) of
    // each of the KB loaded vectors.
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
            char* dst = region + lds_addr(lr, kg * KB * 2);
            if constexpr (KB == 4) *(i32x2_t*)dst = i32x2_t{(int)w[0], (int)w[1]};
            else *(int*)dst = (int)w[0];
        }
    }
};

// s_waitcnt vmcnt(N) with lgkmcnt/expcnt left free (gfx9 encoding).
template <int N>
__device__ inline void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

// Interleave for one block of NM MFMAs: first the NR LDS reads one after each
// MFMA, then `nx` instructions of class `mask` (VMEM loads or LDS writes).
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

template <int MODE>
__device__ inline void mfma16(const i32x4_t& a, const i32x4_t& b, f32x16_t& acc) {
    if constexpr (MODE == kBF16)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b),
                                                      acc, 0, 0, 0);
    else
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8_t, a), __builtin_bit_cast(f16x8_t, b),
                                                     acc, 0, 0, 0);
}

template <int MODE, bool AK, bool BK, int MT_M, int MT_N, int OUT>
__device__ inline void gemm_tile(const TileArgs& a, int tm, int tn, int b, char* smem) {
    static_assert(MODE == kBF16 || MODE == kF16, "16-bit engine");
    constexpr int WT_M = MT_M / 2, WT_N = MT_N / 2;
    constexpr int TM = WT_M / 32, TN = WT_N / 32;
    constexpr int STAGE = (MT_M + MT_N) * RB;
    constexpr int BOFF = MT_M * RB;
    constexpr int NS = Cfg<MT_M, MT_N, AK, BK>::NS, PF = Cfg<MT_M, MT_N, AK, BK>::PF;
    using LA = std::conditional_t<AK, DirectLoader<MT_M, true, PF>, TransLoader<MT_M, true, PF>>;
    using LB = std::conditional_t<BK, DirectLoader<MT_N, false, PF>, TransLoader<MT_N, false, PF>>;
    constexpr int NV = LA::NVMEM + LB::NVMEM;  // VMEM instructions per K-tile

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int m0 = tm * MT_M, n0 = tn * MT_N;
    const int g = lane >> 5;

    const char* Ab = (const char*)a.A + (int64_t)b * a.bsA * 2;
    const char* Bb = (const char*)a.B + (int64_t)b * a.bsB * 2;
    LA la;
    LB lb;
    la.init(Ab, AK ? extent_bytes(a.M, a.lda, a.K, 2) : extent_bytes(a.K, a.lda, a.M, 2), a.lda, a.M, a.K, m0);
    lb.init(Bb, BK ? extent_bytes(a.N, a.ldb, a.K, 2) : extent_bytes(a.K, a.ldb, a.N, 2), a.ldb, a.N, a.K, n0);

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
    la.template issue<0>(0, stage(0));
    lb.template issue<0>(0, stage(0) + BOFF);
    la.template issue<1>(1, stage(1));
    lb.template issue<1>(1, stage(1) + BOFF);
    if constexpr (PF > 2) {
        la.template issue<2 % PF>(2, stage(2));
        lb.template issue<2 % PF>(2, stage(2) + BOFF);
    }
    la.template store<0>(stage(0));
    lb.template store<0>(stage(0) + BOFF);

    auto iteration = [&](auto slot, int kt) {
        constexpr int S = decltype(slot)::value;
        const char* sa = stage(kt);
        const char* sb = sa + BOFF;
        wait_vmcnt<(PF - 1) * NV>();  // K-tile kt's direct loads have landed (this wave)
        lds_barrier();         // ... and every wave's; stage kt-1 is free again
        la.template issue<S>(kt + PF, stage(kt + PF));
        lb.template issue<S>(kt + PF, stage(kt + PF) + BOFF);
#pragma unroll
        for (int i = 0; i < TM; ++i) fa[0][i] = *(const i32x4_t*)(sa + lds_addr(arow + 32 * i, g * 16));
#pragma unroll
        for (int j = 0; j < TN; ++j) fb[0][j] = *(const i32x4_t*)(sb + lds_addr(brow + 32 * j, g * 16));
        // second half of K-tile kt-1 (zeros on the first iteration)
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) mfma16<MODE>(fa[1][i], fb[1][j], acc[i][j]);
        sched_block<TM * TN, TM + TN, 1, 0x020>(NV);
#pragma unroll
        for (int i = 0; i < TM; ++i) fa[1][i] = *(const i32x4_t*)(sa + lds_addr(arow + 32 * i, (2 + g) * 16));
#pragma unroll
        for (int j = 0; j < TN; ++j) fb[1][j] = *(const i32x4_t*)(sb + lds_addr(brow + 32 * j, (2 + g) * 16));
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) mfma16<MODE>(fa[0][i], fb[0][j], acc[i][j]);
        // K-tile kt+1 (register path) into its stage
        la.template store<(S + 1) % PF>(stage(kt + 1));
        lb.template store<(S + 1) % PF>(stage(kt + 1) + BOFF);
        sched_block<TM * TN, TM + TN, 2, 0x200>((LA::kDirect ? 0 : 8) + (LB::kDirect ? 0 : 8));
    };

    int kt = 0;
    for (; kt + PF <= KT; kt += PF) {
        iteration(std::integral_constant<int, 0>{}, kt);
        iteration(std::integral_constant<int, 1>{}, kt + 1);
        if constexpr (PF > 2) iteration(std::integral_constant<int, 2 % PF>{}, kt + 2);
    }
    if (kt < KT) iteration(std::integral_constant<int, 0>{}, kt++);
    if constexpr (PF > 2)
        if (kt < KT) iteration(std::integral_constant<int, 1>{}, kt++);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) mfma16<MODE>(fa[1][i], fb[1][j], acc[i][j]);
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
template <int MT_M, int MT_N, bool AK, bool BK>
__host__ __device__ constexpr int smem_bytes() { return Cfg<MT_M, MT_N, AK, BK>::NS * (MT_M + MT_N) * RB; }

template <int MODE, bool AK, bool BK, int MT_M, int MT_N, int OUT>
__global__ void __launch_bounds__(kThreads, (Cfg<MT_M, MT_N, AK, BK>::OCC)) gemm_kernel(TileArgs a) {
    __shared__ __attribute__((aligned(16))) char smem[smem_bytes<MT_M, MT_N, AK, BK>()];
    int tb, tm, tn;
    map_tile(blockIdx.x, gridDim.x, a, tb, tm, tn);
    v2::gemm_tile<MODE, AK, BK, MT_M, MT_N, OUT>(a, tm, tn, tb, smem);
}

}  // namespace v2
}  // namespace ffgemm
