// Eight-wave ping-pong engine: 256 x 256 tiles of bf16 / fp16 / fp8 (OCP
// e4m3) operands.  K-contiguous operands (TN, the reference's layout) load
// into 128-byte K rows; M/N-contiguous 16-bit operands (NT) load in their own
// orientation and reach the MFMAs through transposing LDS reads
// (ds_read_b64_tr_b16), so every load is a direct HBM -> LDS copy.
//
// Eight waves, two per SIMD.  The four waves of group 0 (waves 0-3) own
// output rows 0-127 of the tile, group 1 (waves 4-7) rows 128-255; wave s of
// a group owns 64 columns, a 128 x 64 block of 4 x 2 MFMA 32x32x16 tiles.
// The groups run one "window" apart: while the wave of one group runs the 32
// MFMAs of a K-tile (its compute window), the other wave on the same SIMD
// loads its fragments for the next K-tile from LDS and issues the HBM -> LDS
// loads (its memory window), so the MFMA pipe of every SIMD always has a wave
// with ready operands.  One workgroup barrier closes every window.  A window
// covers 64 bytes of K: two v_mfma_f32_32x32x16 K-steps per tile (16-bit) or
// one v_mfma_scale_f32_32x32x64_f8f6f4 (fp8), 512 MFMA cycles either way.
//
// LDS (160 KiB): 128 bytes of K per stage and row, with XOR-swizzled 16-byte
// chunks (conflict-free fragment reads), B in three stages and A in two.
// Every load goes HBM -> LDS directly (buffer_load_dwordx4 ... lds).  With
// both operands K-contiguous (TN, fp8) only memory windows issue loads:
// issuing a load costs its wave tens of cycles, which a compute window would
// take from the MFMA pipe.  Each memory window issues four loads per wave, two
// before the fragment reads of each K-step:
//   * the first memory window of K-tile kt: the group's half g (128 rows) of
//     A for K-tile kt+1 (only group g reads that half; stage (kt+1) % 2 was
//     last read in K-tile kt-1), landing four windows later;
//   * the second: half g of B for K-tile kt+2 (stage (kt+2) % 3 was last read
//     by both groups in K-tile kt-1), landing five windows later.
// So the load traffic is spread evenly over the four windows of a K-tile.
// NT keeps B(kt+2) at the start of the first memory window and A(kt+2) in the
// second compute window, one load before each group of four MFMAs (its memory
// windows already carry 24 transposing reads; splitting B's loads around them
// measured 15-25 % slower).
#pragma once
#include "gemm_tile_v2.hpp"

namespace ffgemm {
namespace v4 {

using v2::lds_addr;
using v2::lds_ptr_t;
using v2::wait_vmcnt;

constexpr int kThreads8 = 512;
constexpr int MT = 256;
constexpr int RB = 128;  // bytes of K per LDS row per stage
constexpr int STAGE = MT * RB;
constexpr int NSA = 2, NSB = 3;
constexpr int LDS_BYTES = (NSA + NSB) * STAGE;
constexpr int TI = 4, TJ = 2;  // MFMA tiles per wave along M and N
constexpr int NLD = 4;         // load instructions per wave per operand per K-tile

// Half (128 rows) of a 256-row K-contiguous operand panel for one K-tile,
// loaded by the four waves of one group: wave s fills LDS rows
// 128h + 32s + 8q + lane/8 (q = 0..3), one 1 KiB instruction per q.  Lane
// p of a row takes the logical chunk whose swizzled slot is p.  PERM applies
// the A-row permutation (LDS row -> tile row perm_a_inv), which is 4 rows
// per q inside a 32-row block instead of 8.
template <bool PERM, int ES = 2>
struct HalfLoader {
    static constexpr int QROWS = PERM ? 4 : 8;  // tile rows between q and q+1
    static constexpr int DU = RB / ES;          // K per stage
    static constexpr int CE = 16 / ES;          // K per 16-byte chunk
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t base;    // byte offset of this lane's q = 0 row
    int32_t rlim;     // rows - (this lane's q = 0 row)
    int32_t x;        // logical chunk of q = 0 (chunk of q is x ^ q)
    uint32_t qstride; // bytes between the rows of q and q + 1
    int32_t K;
    int32_t lds_off;  // this wave's first LDS row, in bytes

    __device__ inline void init(const void* ptr, uint32_t extent, int32_t ld, int32_t rows, int32_t K_,
                                int32_t row0, int h, int s, int tid) {
        rsrc = make_rsrc(ptr, extent);
        const int lane = tid & 63;
        const int j = lane >> 3;
        const int lr0 = 128 * h + 32 * s + j;  // LDS row of q = 0
        const int r = row0 + (PERM ? v2::perm_a_inv(lr0) : lr0);
        rlim = rows - r;
        x = ((lane & 7) ^ j ^ (4 * (s & 1))) & 7;
        base = (uint32_t)((int64_t)r * ld * ES);
        qstride = (uint32_t)(QROWS * ld * ES);
        K = K_;
        lds_off = (128 * h + 32 * s) * RB;
    }
    // loads q = Q0 .. Q0 + NQ - 1 of K-tile kt (rows and K past the matrix read zero)
    template <int Q0 = 0, int NQ = NLD>
    __device__ inline void issue(int kt, char* stage) const {
        char* dst = stage + lds_off;
#pragma unroll
        for (int q = Q0; q < Q0 + NQ; ++q) {
            const int c = x ^ q;
            const bool ok = (QROWS * q < rlim) && (kt * DU + c * CE < K);
            const uint32_t off = ok ? base + q * qstride + (uint32_t)kt * RB + (uint32_t)c * 16 : kOOB;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_ptr_t)(dst + q * 1024), 16, off, 0, 0, 0);
        }
    }
    // Per-q lane offsets for issue_full: rows past the matrix at kOOB.
    __device__ inline void full_offsets(uint32_t (&v)[NLD]) const {
#pragma unroll
        for (int q = 0; q < NLD; ++q)
            v[q] = (QROWS * q < rlim) ? base + q * qstride + (uint32_t)(x ^ q) * 16 : kOOB;
    }
    // issue() for a K-tile that lies wholly inside K (kt * DU + DU <= K): the
    // K-tile's byte offset goes in the instruction's scalar offset, so the loads
    // need no per-lane arithmetic.  Rows past the matrix read zero through
    // their kOOB lane offset (out of range with or without the scalar part).
    template <int Q0 = 0, int NQ = NLD>
    __device__ inline void issue_full(int kt, char* stage, const uint32_t (&v)[NLD]) const {
        char* dst = stage + lds_off;
        const int soff = __builtin_amdgcn_readfirstlane(kt * RB);
#pragma unroll
        for (int q = Q0; q < Q0 + NQ; ++q)
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_ptr_t)(dst + q * 1024), 16, v[q], soff, 0, 0);
    }
};

// Half (128 rows) of a 256-row M/N-contiguous 16-bit operand for one K-tile,
// loaded straight into LDS in its own orientation: region [64 k][128 rows],
// 256 bytes per k, 16-byte chunk c of k stored in slot c ^ 4*(k % 4), which
// makes the transposing fragment reads below conflict-free.  Wave s of the
// group fills k = 16s + 4q + lane/16 (q = 0..3), one 1 KiB instruction per q:
// lane p of a k takes the 8 rows of chunk p ^ 4*(k % 4), coalesced 256 bytes
// per k.  k beyond K reads zero; rows beyond the matrix only reach output rows
// that are never stored.
struct MNHalfLoader {
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t base;     // this lane's byte offset for kt = 0, q = 0
    int32_t klim;      // K - (this lane's k for kt = 0, q = 0)
    uint32_t qstride;  // 4 k
    uint32_t kstride;  // 64 k
    int32_t lds_off;   // this wave's first k, in bytes of the region

    __device__ inline void init(const void* ptr, uint32_t extent, int32_t ld, int32_t K, int32_t row0, int s,
                                int tid) {
        rsrc = make_rsrc(ptr, extent);
        const int lane = tid & 63;
        const int kr = 16 * s + (lane >> 4);
        const int c = (lane & 15) ^ (4 * (lane >> 4));
        base = (uint32_t)(((int64_t)kr * ld + row0 + 8 * c) * 2);
        klim = K - kr;
        qstride = (uint32_t)(4 * ld * 2);
        kstride = (uint32_t)(64 * ld * 2);
        lds_off = 16 * s * 256;
    }
    template <int Q0 = 0, int NQ = NLD>
    __device__ inline void issue(int kt, char* region) const {
        char* dst = region + lds_off;
#pragma unroll
        for (int q = Q0; q < Q0 + NQ; ++q) {
            const bool ok = kt * 64 + 4 * q < klim;
            const uint32_t off = ok ? base + (uint32_t)kt * kstride + q * qstride : kOOB;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_ptr_t)(dst + q * 1024), 16, off, 0, 0, 0);
        }
    }
    // (no scalar-offset form: rows past the matrix are masked only by the
    // range check of the per-lane offset)
    __device__ inline void full_offsets(uint32_t (&)[NLD]) const {}
    template <int Q0 = 0, int NQ = NLD>
    __device__ inline void issue_full(int kt, char* region, const uint32_t (&)[NLD]) const {
        issue<Q0, NQ>(kt, region);
    }
};

typedef short i16x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) i16x4_t lds_i16x4_t;

// ds_read_b64_tr_b16: within each 16-lane group, source lane s reads 4
// consecutive rows at one k, and lane i receives element j of source lane
// 4j + i/4 (position i % 4), i.e. 4 consecutive k of one row.
__device__ inline i16x4_t read_tr16(const char* p) {
    return __builtin_amdgcn_ds_read_tr16_b64_v4i16(
        (lds_i16x4_t*)((__attribute__((address_space(3))) char*)(p)));
}

// Per-lane offset, inside an [64 k][128 rows] region, of the transposing read
// that gives lane l the 4 k from k0 = 8*(l/32) (+ kb, a multiple of 4, added by
// the caller as kb*256) of row `row` of the MFMA operand, where row = r0 +
// (PERM ? perm_a_inv(l % 32) : l % 32): source lane s reads rows cs..cs+3 at
// k = k0 + s/4.
template <bool PERM>
__device__ inline int tr_offset(int lane, int r0) {
    const int G = lane >> 4, s = lane & 15, q = s & 3, gp = G & 1;
    const int cs = PERM ? (16 * (q & 1) + 8 * gp + 4 * (q >> 1)) : (16 * gp + 4 * q);
    const int k = 8 * (G >> 1) + (s >> 2);
    const int b = (r0 + cs) * 2;  // byte of the row inside the 256-byte k line
    return k * 256 + ((((b >> 4) ^ (4 * (k & 3))) & 15) << 4) + (b & 15);
}

__device__ inline void barrier() { asm volatile("s_barrier" ::: "memory"); }

// OCP e4m3 x e4m3, 32x32x64, unit block scales (e8m0 127 = 1.0).
__device__ inline void mfma_f8(const i32x4_t& a0, const i32x4_t& a1, const i32x4_t& b0, const i32x4_t& b1,
                               f32x16_t& acc) {
    const i32x8_t av = {a0[0], a0[1], a0[2], a0[3], a1[0], a1[1], a1[2], a1[3]};
    const i32x8_t bv = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
    acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, acc, 0, 0, 0, 127, 0, 127);
}

// Epilogue of a whole tile with beta = 0 (called right after the K loop, every
// wave's LDS-direct loads drained).  Each wave's accumulators go to LDS as a
// column-major tile (column stride 528 bytes: the fragment writes and the column
// reads are free of bank conflicts), then every instruction stores whole
// 512-byte column pieces: 16 passes of 16 columns (bf16 / fp16), or two rounds
// of 128 rows for fp32 D (the 256 KiB tile exceeds LDS).  The per-fragment
// stores write 16-byte pieces of 32 columns per instruction, so every 128-byte
// line of D was assembled from several partial writes; whole-line stores
// measured fp8 4096^3 63.3 -> 55.3 us, 8192^3 412.8 -> 367.7 us, bf16 4096^3
// 103.1 -> 100.9 us.  Stores are non-temporal, or system-coherent (sc0 sc1)
// for fp32 D when all tiles run in one round: fp8 4096^3 then measured 2.5-4 %
// faster (and 2 % slower at 8192^3, four rounds; bf16 D 0.4 % slower).  Dt:
// the tile's first element.
template <int OUT>
__device__ inline void store_tile_staged(const TileArgs& a, char* Dt, const f32x16_t (&acc)[TI][TJ], float alpha,
                                         char* smem, int tid) {
    constexpr int OS = (OUT == kOutF32) ? 4 : 2;
    constexpr int CS = 528;              // LDS bytes per staged column (512 + 16)
    constexpr int ROWS = 512 / OS;       // rows per staging round: 256 (16-bit), 128 (fp32)
    constexpr int NPG = MT / ROWS;       // staging rounds: 1 or 2
    constexpr int NT = 2, SC01 = 17;     // buffer store cache policies: nt, sc0 sc1
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int grp = wave >> 2, s = wave & 3;
    const int h = lane >> 5;
    const bool trunc = OUT == kOutBF16 && (a.flags & kFlagTrunc) != 0;
    const bool one_round = OUT == kOutF32 && (a.flags & kFlagOneRound) != 0;
    const __amdgpu_buffer_rsrc_t rd = make_rsrc(Dt, (uint32_t)((int64_t)(MT - 1) * a.ldd * OS + MT * OS));
    // the groups realign (group 1 ran one barrier ahead), then every wave's
    // loads have landed and the stages are free
    if (grp == 0) barrier();
    barrier();
#pragma unroll
    for (int pg = 0; pg < NPG; ++pg) {
        if (NPG == 1 || grp == pg) {
#pragma unroll
            for (int i = 0; i < TI; ++i)
#pragma unroll
                for (int j = 0; j < TJ; ++j) {
                    const f32x16_t& x = acc[i][j];
                    const int ml = (NPG == 1 ? 128 * grp : 0) + 32 * i + 16 * h;
                    const int nl = 64 * s + 32 * j + (lane & 31);
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
        for (int pass = 0; pass < 16; ++pass) {
            const int col = 16 * pass + (tid >> 5), ch = tid & 31;
            const i32x4_t v = *(const i32x4_t*)(smem + col * CS + ch * 16);
            const uint32_t off = (uint32_t)(((int64_t)col * a.ldd + ROWS * pg) * OS + ch * 16);
            if (one_round) __builtin_amdgcn_raw_buffer_store_b128(v, rd, off, 0, SC01);
            else __builtin_amdgcn_raw_buffer_store_b128(v, rd, off, 0, NT);
        }
        if (pg + 1 < NPG) lds_barrier();
    }
}

// beta != 0 with 16-bit D and C: the unscaled accumulators are staged in fp32,
// two rounds of 128 rows (group pg's half in round pg, as for fp32 D).  Each
// round's C pieces (4 consecutive rows of one column per thread, 8 bytes) are
// loaded before the round's staging barrier, so their latency overlaps the
// staging; the column pass then forms fma(alpha, acc, beta * C), the
// per-fragment epilogue's arithmetic, and stores whole 256-byte column pieces
// of D.  Dt, Ct: the tile's first element of D and of C.
template <int OUT>
__device__ inline void store_tile_staged_beta(const TileArgs& a, char* Dt, const char* Ct,
                                              const f32x16_t (&acc)[TI][TJ], float alpha, char* smem, int tid) {
    static_assert(OUT == kOutBF16 || OUT == kOutF16, "16-bit D and C");
    typedef int v2i_t __attribute__((ext_vector_type(2)));
    constexpr int CS = 528;   // LDS bytes per staged column (512 + 16)
    constexpr int ROWS = 128; // rows per staging round
    constexpr int NT = 2;     // buffer store cache policy: nt
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int grp = wave >> 2, s = wave & 3;
    const int h = lane >> 5, ch = tid & 31;
    const bool trunc = OUT == kOutBF16 && (a.flags & kFlagTrunc) != 0;
    const float beta = a.beta;
    const __amdgpu_buffer_rsrc_t rd = make_rsrc(Dt, (uint32_t)((int64_t)(MT - 1) * a.ldd * 2 + MT * 2));
    const __amdgpu_buffer_rsrc_t rc = make_rsrc(Ct, (uint32_t)((int64_t)(MT - 1) * a.ldc * 2 + MT * 2));
    // the groups realign (group 1 ran one barrier ahead), then every wave's
    // loads have landed and the stages are free
    if (grp == 0) barrier();
    barrier();
#pragma unroll
    for (int pg = 0; pg < 2; ++pg) {
        // pass p of this round: column 16 p + tid / 32, rows 128 pg + 4 ch .. + 3
        v2i_t c[16];
#pragma unroll
        for (int pass = 0; pass < 16; ++pass) {
            const int col = 16 * pass + (tid >> 5);
            c[pass] = __builtin_amdgcn_raw_buffer_load_b64(
                rc, (uint32_t)(((int64_t)col * a.ldc + ROWS * pg + 4 * ch) * 2), 0, 0);
        }
        if (grp == pg) {
#pragma unroll
            for (int i = 0; i < TI; ++i)
#pragma unroll
                for (int j = 0; j < TJ; ++j) {
                    const f32x16_t& x = acc[i][j];
                    char* p = smem + (64 * s + 32 * j + (lane & 31)) * CS + (32 * i + 16 * h) * 4;
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        *(f32x4_t*)(p + 16 * q) = f32x4_t{x[4 * q], x[4 * q + 1], x[4 * q + 2], x[4 * q + 3]};
                }
        }
        lds_barrier();
#pragma unroll
        for (int pass = 0; pass < 16; ++pass) {
            const int col = 16 * pass + (tid >> 5);
            const f32x4_t v = *(const f32x4_t*)(smem + col * CS + ch * 16);
            const uint32_t u0 = (uint32_t)c[pass][0], u1 = (uint32_t)c[pass][1];
            const float cv[4] = {OUT == kOutBF16 ? bf16lo(u0) : f16lo(u0), OUT == kOutBF16 ? bf16hi(u0) : f16hi(u0),
                                 OUT == kOutBF16 ? bf16lo(u1) : f16lo(u1), OUT == kOutBF16 ? bf16hi(u1) : f16hi(u1)};
            float r[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) r[e] = __builtin_fmaf(alpha, v[e], beta * cv[e]);
            uint32_t w[2];
#pragma unroll
            for (int e = 0; e < 2; ++e)
                w[e] = OUT == kOutF16 ? pack_f16x2(r[2 * e], r[2 * e + 1])
                                      : (trunc ? pack_bf16x2_trunc(r[2 * e], r[2 * e + 1])
                                               : pack_bf16x2_rne(r[2 * e], r[2 * e + 1]));
            __builtin_amdgcn_raw_buffer_store_b64(v2i_t{(int)w[0], (int)w[1]}, rd,
                                                  (uint32_t)(((int64_t)col * a.ldd + ROWS * pg + 4 * ch) * 2), 0, NT);
        }
        if (pg == 0) lds_barrier();
    }
}

template <int MODE, bool AK, bool BK, int OUT>
__device__ inline void gemm_tile(const TileArgs& a, int tm, int tn, int b, char* smem, int tid) {
    static_assert(MODE == kBF16 || MODE == kF16 || MODE == kF8, "16-bit or fp8 engine");
    static_assert((AK && BK) || MODE != kF8, "fp8 operands are K-contiguous");
    constexpr int ES = (MODE == kF8) ? 1 : 2;
    constexpr int DU = RB / ES;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int grp = wave >> 2, s = wave & 3;
    const int h = lane >> 5;
    const int m0 = tm * MT, n0 = tn * MT;

    const char* Ab = (const char*)a.A + (int64_t)b * a.bsA * ES;
    const char* Bb = (const char*)a.B + (int64_t)b * a.bsB * ES;
    // K-contiguous operands: rows of 128 bytes of K; M/N-contiguous ones: the
    // group halves of a stage are [64 k][128 rows] regions (MNHalfLoader)
    using LA = std::conditional_t<AK, HalfLoader<true, ES>, MNHalfLoader>;
    using LB = std::conditional_t<BK, HalfLoader<false, ES>, MNHalfLoader>;
    LA la;
    LB lb;
    // K-contiguous operands: the buffer starts at the tile's first row, so the
    // 32-bit offsets stay within the tile's 256 rows however large the matrix
    if constexpr (AK)
        la.init(Ab + (int64_t)m0 * a.lda * ES, extent_bytes(a.M - m0, a.lda, a.K, ES), a.lda, a.M - m0, a.K, 0, grp, s, tid);
    else la.init(Ab, extent_bytes(a.K, a.lda, a.M, 2), a.lda, a.K, m0 + 128 * grp, s, tid);
    if constexpr (BK)
        lb.init(Bb + (int64_t)n0 * a.ldb * ES, extent_bytes(a.N - n0, a.ldb, a.K, ES), a.ldb, a.N - n0, a.K, 0, grp, s, tid);
    else lb.init(Bb, extent_bytes(a.K, a.ldb, a.N, 2), a.ldb, a.K, n0 + 128 * grp, s, tid);
    // where this group's loads go inside a stage (K-contiguous loaders add it themselves)
    const int lhalf_a = AK ? 0 : 16384 * grp, lhalf_b = BK ? 0 : 16384 * grp;

    char* const sA = smem;
    char* const sB = smem + NSA * STAGE;
    auto stA = [&](int kt) { return sA + (kt & 1) * STAGE; };
    auto stB = [&](int kt) { return sB + (kt % NSB) * STAGE; };

    // Fragment addresses.  A row 128*grp + 32i + lane%32 and B row
    // 64s + 32j + lane%32 both have swizzle (lane&7) ^ ((lane>>3)&3) ^ 4*(i|j & 1),
    // so chunk 2ks + h sits in slot y ^ c with c = 2ks ^ 4*(i|j & 1): four
    // per-lane offsets serve every fragment; the rest is uniform.  (fp8: the
    // 32-byte operand of lane group h for K-step s is chunks 4s + h and
    // 4s + 2 + h, the chunks of 16-bit K-steps 2s and 2s + 1, in A and in B.)
    const int y = (h ^ (lane & 7) ^ ((lane >> 3) & 3)) & 7;
    int P[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) P[c] = (lane & 31) * RB + ((y ^ (2 * c)) << 4);
    const int aoff = 128 * grp * RB, boff = 64 * s * RB;
    // transposed fragment reads: per-tile offsets inside this wave's region
    int TA[TI], TB[TJ];
#pragma unroll
    for (int i = 0; i < TI; ++i) TA[i] = tr_offset<true>(lane, 32 * i);
#pragma unroll
    for (int j = 0; j < TJ; ++j) TB[j] = tr_offset<false>(lane, 64 * (s & 1) + 32 * j);
    const int aregion = 16384 * grp, bregion = 16384 * (s >> 1);
    auto tr_frag = [&](const char* region, int off, int ks) {
        const i16x4_t lo = read_tr16(region + off + 16 * ks * 256);
        const i16x4_t hi = read_tr16(region + off + (16 * ks + 4) * 256);
        const i32x2_t l2 = __builtin_bit_cast(i32x2_t, lo), h2 = __builtin_bit_cast(i32x2_t, hi);
        return i32x4_t{l2[0], l2[1], h2[0], h2[1]};
    };

    f32x16_t acc[TI][TJ];
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;
    i32x4_t fa[2][TI], fb[2][TJ];

    const int KT = (a.K + DU - 1) / DU;

    // fp8: the per-tensor quantisation scales, read before the K loop so their
    // latency is not added to the epilogue
    float alpha = a.alpha;
    if constexpr (MODE == kF8) {
        const float sa = a.scaleA ? *a.scaleA : 1.0f;
        const float sb = a.scaleB ? *a.scaleB : 1.0f;
        alpha = alpha / (sa * sb);
    }

    // Load schedule.  BAL (both operands K-contiguous: TN, fp8): loads only in
    // memory windows, four per wave each (see the top of this file).  With
    // transposing reads (NT) a memory window already carries 24 LDS reads; its
    // loads then measured faster where they were: B(kt+2) at the start of the
    // first memory window and A(kt+2) at the start of the second compute window
    // (NT bf16 4096^3: 1029 against 633 TFLOP/s with the balanced schedule).
    constexpr bool BAL = AK && BK;

    // prologue.  BAL: B(0), A(0) and B(1) in flight, B(0) and A(0) complete
    // before the barrier (A(1) is issued by the first memory window).
    // Otherwise: K-tiles 0 and 1 in flight, K-tile 0 complete.
    lb.issue(0, stB(0) + lhalf_b);
    la.issue(0, stA(0) + lhalf_a);
    lb.issue(1, stB(1) + lhalf_b);
    if constexpr (BAL) {
        wait_vmcnt<NLD>();
    } else {
        la.issue(1, stA(1) + lhalf_a);
        wait_vmcnt<2 * NLD>();
    }
    lds_barrier();
    if (grp == 1) barrier();  // group 1 runs one window behind group 0

    // K-contiguous operands: K-tiles wholly inside K load with fixed per-lane
    // offsets and the K position in the instruction's scalar offset (no
    // per-lane arithmetic in the windows); the tail and the K-tiles past the
    // end take the masked path, which reads zeros beyond K
    uint32_t va[NLD], vb[NLD];
    if constexpr (AK) la.full_offsets(va);
    if constexpr (BK) lb.full_offsets(vb);
    const int KF = a.K / DU;
    // loads q = Q0 .. Q0 + NQ - 1 of a K-tile of this group's half
    auto loadB = [&](int kt, auto full, auto q0, auto nq) {
        constexpr int Q0 = decltype(q0)::value, NQ = decltype(nq)::value;
        if constexpr (BK && decltype(full)::value) lb.template issue_full<Q0, NQ>(kt, stB(kt) + lhalf_b, vb);
        else lb.template issue<Q0, NQ>(kt, stB(kt) + lhalf_b);
    };
    auto loadA = [&](int kt, char* stage, auto full, auto q0, auto nq) {
        constexpr int Q0 = decltype(q0)::value, NQ = decltype(nq)::value;
        if constexpr (AK && decltype(full)::value) la.template issue_full<Q0, NQ>(kt, stage + lhalf_a, va);
        else la.template issue<Q0, NQ>(kt, stage + lhalf_a);
    };
    using q0_t = std::integral_constant<int, 0>;
    using q2_t = std::integral_constant<int, 2>;
    using n2_t = std::integral_constant<int, 2>;
    using n4_t = std::integral_constant<int, NLD>;

    // Windows per K-tile: memory(half 0), compute(half 0), memory(half 1),
    // compute(half 1); a half is K = 32 (two MFMA K-steps).
    auto ktile = [&](int kt, auto full) {
        const char* sa = stA(kt) + aoff;
        const char* sb = stB(kt) + boff;
        const char* ra = stA(kt) + aregion;
        const char* rb = stB(kt) + bregion;
        auto frags = [&](int half, int k2) {
            const int ks = 2 * half + k2;
#pragma unroll
            for (int i = 0; i < TI; ++i) {
                if constexpr (AK) fa[k2][i] = *(const i32x4_t*)(sa + P[ks ^ (2 * (i & 1))] + 4096 * i);
                else fa[k2][i] = tr_frag(ra, TA[i], ks);
            }
#pragma unroll
            for (int j = 0; j < TJ; ++j) {
                if constexpr (BK) fb[k2][j] = *(const i32x4_t*)(sb + P[ks ^ (2 * j)] + 4096 * j);
                else fb[k2][j] = tr_frag(rb, TB[j], ks);
            }
        };
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            // ---- memory window
            if constexpr (BAL) {
                // each K-step's fragment reads behind two loads, of A(kt+1)
                // (half 0) or B(kt+2) (half 1)
                if (half == 0) loadA(kt + 1, stA(kt + 1), full, q0_t{}, n2_t{});
                else loadB(kt + 2, full, q0_t{}, n2_t{});
                frags(half, 0);
                if (half == 0) loadA(kt + 1, stA(kt + 1), full, q2_t{}, n2_t{});
                else loadB(kt + 2, full, q2_t{}, n2_t{});
                frags(half, 1);
            } else {
                if (half == 0) loadB(kt + 2, full, q0_t{}, n4_t{});
                frags(half, 0);
                frags(half, 1);
            }
            // BAL: B(kt+1), issued by this wave two memory windows ago, lands
            // before the barrier that precedes its first reader
            if (half == 1) wait_vmcnt<2 * NLD>();
            lds_barrier();
            __builtin_amdgcn_sched_barrier(0);
            // ---- compute window
            __builtin_amdgcn_s_setprio(1);
            if constexpr (MODE == kF8) {
#pragma unroll
                for (int i = 0; i < TI; ++i)
#pragma unroll
                    for (int j = 0; j < TJ; ++j) mfma_f8(fa[0][i], fa[1][i], fb[0][j], fb[1][j], acc[i][j]);
            } else if constexpr (!BAL) {
                // this group's half of A(kt) is consumed: A(kt+2)'s four loads
                // go one before each group of four MFMAs (half 1)
#pragma unroll
                for (int k2 = 0; k2 < 2; ++k2)
#pragma unroll
                    for (int i = 0; i < TI; ++i) {
                        if (half == 1 && (i & 1) == 0) {
                            if (k2 == 0 && i == 0) loadA(kt + 2, stA(kt), full, std::integral_constant<int, 0>{}, std::integral_constant<int, 1>{});
                            if (k2 == 0 && i == 2) loadA(kt + 2, stA(kt), full, std::integral_constant<int, 1>{}, std::integral_constant<int, 1>{});
                            if (k2 == 1 && i == 0) loadA(kt + 2, stA(kt), full, std::integral_constant<int, 2>{}, std::integral_constant<int, 1>{});
                            if (k2 == 1 && i == 2) loadA(kt + 2, stA(kt), full, std::integral_constant<int, 3>{}, std::integral_constant<int, 1>{});
                        }
#pragma unroll
                        for (int j = 0; j < TJ; ++j) v2::mfma16<MODE>(fa[k2][i], fb[k2][j], acc[i][j]);
                        if ((i & 1) == 1) __builtin_amdgcn_sched_barrier(0);
                    }
            } else {
#pragma unroll
                for (int k2 = 0; k2 < 2; ++k2)
#pragma unroll
                    for (int i = 0; i < TI; ++i)
#pragma unroll
                        for (int j = 0; j < TJ; ++j) v2::mfma16<MODE>(fa[k2][i], fb[k2][j], acc[i][j]);
            }
            __builtin_amdgcn_s_setprio(0);
            __builtin_amdgcn_sched_barrier(0);
            // BAL: A(kt+1) of this group lands before its first reader's
            // window; otherwise K-tile kt+1 lands before the next window
            if (half == 1) {
                if constexpr (BAL) wait_vmcnt<NLD>();
                else wait_vmcnt<2 * NLD>();
            }
            barrier();
        }
    };
    // K-tiles whose loads lie wholly inside K, then the rest
    int kt = 0;
    for (; kt < KF - 2; ++kt) ktile(kt, std::true_type{});
    for (; kt < KT; ++kt) ktile(kt, std::false_type{});
    // loads issued past the end write zeros into free stages; drain them
    // before the workgroup's LDS is released
    wait_vmcnt<0>();

    const char* Cb = (const char*)a.C + (int64_t)b * a.bsC * ((OUT == kOutF32) ? 4 : 2);
    char* Db = (char*)a.D + (int64_t)b * a.bsD * ((OUT == kOutF32) ? 4 : 2);
    // A whole tile leaves through LDS as whole columns (store_tile_staged; with
    // beta != 0 and 16-bit D, store_tile_staged_beta); edge tiles, fp32 D with
    // beta != 0 and C or D extents past the 31-bit buffer range store fragment
    // by fragment.
    constexpr int OS = (OUT == kOutF32) ? 4 : 2;
    if (m0 + MT <= a.M && n0 + MT <= a.N && (int64_t)(MT - 1) * a.ldd * OS + MT * OS < 0x7ffffff0ll) {
        if ((a.flags & kFlagBeta) == 0) {
            store_tile_staged<OUT>(a, Db + ((int64_t)n0 * a.ldd + m0) * OS, acc, alpha, smem, tid);
            return;
        }
        if constexpr (OUT != kOutF32) {
            if ((int64_t)(MT - 1) * a.ldc * OS + MT * OS < 0x7ffffff0ll) {
                store_tile_staged_beta<OUT>(a, Db + ((int64_t)n0 * a.ldd + m0) * OS, Cb + ((int64_t)n0 * a.ldc + m0) * OS,
                                            acc, alpha, smem, tid);
                return;
            }
        }
    }
    int fm[TI * TJ], fn[TI * TJ];
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) {
            fm[i * TJ + j] = m0 + 128 * grp + 32 * i + 16 * h;
            fn[i * TJ + j] = n0 + 64 * s + 32 * j + (lane & 31);
        }
    store_frags<OUT, TI * TJ, (OUT == kOutF32) ? 4 : 8>(a, Cb, Db, fm, fn, &acc[0][0], alpha, a.beta);
    if (grp == 0) barrier();  // matches group 1's extra barrier
}

template <int MODE, bool AK, bool BK, int OUT>
__global__ void __launch_bounds__(kThreads8, 1) gemm_kernel(TileArgs a) {
    __shared__ __attribute__((aligned(16))) char smem[LDS_BYTES];
    int tb, tm, tn;
    map_tile(blockIdx.x, gridDim.x, a, tb, tm, tn);
    v4::gemm_tile<MODE, AK, BK, OUT>(a, tm, tn, tb, smem, threadIdx.x);
}

}  // namespace v4
}  // namespace ffgemm
