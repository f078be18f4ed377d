// The tile engine: one MT_M x MT_N output tile of D = alpha*op(A)*op(B) + beta*C.
//
// Replaces the fetch/math wave pair of src/gemmMfmaKernelWriter.h
// (fetch_loop :1226-1316, math_loop :1403-1447, fetch_c :1822-1906,
// mfma_mfma_pang_with_store :1572-1768) with a single-role design: every wave
// both loads (16-byte buffer loads into VGPRs, one K-tile ahead, two for
// tiles up to 128 x 128) and computes
// (MFMA from a two-stage LDS ring), which on gfx950 keeps the MFMA pipe busy
// without spending half the waves on address generation.
#pragma once
#include "gemm_common.hpp"

namespace ffgemm {

// ---------------------------------------------------------------------------
// Operand loader: one MT x (128 bytes of K) tile of A (rows = M) or B (rows = N)
// ---------------------------------------------------------------------------
// KC: operand is K-contiguous in HBM (A of TN/TT, B of NN/TN).
// Otherwise it is M/N-contiguous and each thread loads KB consecutive K-rows
// of an 8 (16-bit) or 4 (32-bit) element column strip and transposes it in
// registers.
// PF: register slots, i.e. K-tiles whose loads can be in flight at once.
// ANY: the general-stride form.  Each 16-byte register vector is gathered by
// element loads (1, 2 or 4 bytes), so leading dimensions, batch strides and K
// need no 16-byte multiple and pointers only their element alignment; elements
// past K (and, M/N-contiguous, past the last row) read as zero.
template <int MODE, bool KC, int MT, bool PERM, int PF = 1, bool ANY = false>
struct Loader {
    static constexpr int ES = mode_elem_bytes(MODE);
    static constexpr int DU = mode_depth_u(MODE);
    static constexpr bool X3 = (MODE == kX3);
    // loads of 16 bytes per thread
    static constexpr int NLD = (MT * kRowBytes) / (16 * kThreads);
    static_assert(NLD >= 1, "tile too small");
    // M/N-contiguous: elements per 16-byte load along M/N, K-rows per thread
    static constexpr int VE = 16 / ES;
    static constexpr int KB = KC ? 0 : (DU * MT) / (VE * kThreads);
    static_assert(KC || (KB == NLD && KB >= 2), "bad transposing geometry");
    static_assert(KC || MODE != kF8, "fp8 operands must be K-contiguous");
    // LDS write instructions per K-tile
    static constexpr int NSTORE = KC ? (MODE == kX3 ? 2 * NLD : NLD)
                                     : (ES == 2 ? 8 : (MODE == kX3 ? 8 : (KB == 2 ? 4 : 4 * (KB / 4))));

    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t voff[NLD];     // byte offset of each load for K-tile 0 (kOOB if masked)
    int32_t kidx[NLD];      // K index of each load for K-tile 0
    uint32_t kstep;         // bytes to advance per K-tile
    int32_t K;
    int32_t rlim;           // M/N-contiguous, ANY: rows of the vector inside the matrix
    int tid;                // thread index in the workgroup
    i32x4_t v[PF][NLD];

    // ld: leading dimension (elements); rows: M or N; row0: first row of tile
    __device__ inline void init(const void* base, uint32_t extent_bytes, int32_t ld, int32_t rows,
                                int32_t K_, int32_t row0, int t) {
        rsrc = make_rsrc(base, extent_bytes);
        K = K_;
        tid = t;
        if constexpr (KC) {
            const int c = t & 7;
#pragma unroll
            for (int i = 0; i < NLD; ++i) {
                int r = row0 + (t >> 3) + 32 * i;
                int k = c * VE;
                kidx[i] = k;
                voff[i] = (r < rows) ? (uint32_t)(((int64_t)r * ld + k) * ES) : kOOB;
            }
            kstep = kRowBytes;
        } else {
            constexpr int NC = MT / VE;  // column strips across the tile
            const int mc = t % NC, kg = t / NC;
            const int r = row0 + mc * VE;
            rlim = rows - r;
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                int k = kg * KB + j;
                kidx[j] = k;
                voff[j] = (r < rows) ? (uint32_t)(((int64_t)k * ld + r) * ES) : kOOB;
            }
            kstep = (uint32_t)(DU * ld * ES);
        }
    }

    // Issue the 16-byte loads of K-tile kt into register slot S.
    template <int S = 0>
    __device__ inline void load(int kt) {
        const int k0 = kt * DU;
#pragma unroll
        for (int i = 0; i < NLD; ++i) {
            if constexpr (ANY) {
                // valid elements of the vector: K-contiguous, VE consecutive K
                // of one row; M/N-contiguous, VE consecutive rows at one K
                int nv;
                if constexpr (KC) nv = voff[i] == kOOB ? 0 : min(VE, K - (kidx[i] + k0));
                else nv = (voff[i] == kOOB || kidx[i] + k0 >= K) ? 0 : min(VE, rlim);
                v[S][i] = gather(voff[i] + (uint32_t)kt * kstep, nv);
            } else {
                bool ok = (kidx[i] + k0) < K && voff[i] != kOOB;
                uint32_t off = ok ? voff[i] + (uint32_t)kt * kstep : kOOB;
                v[S][i] = buf_load16(rsrc, off);
            }
        }
    }

    // The VE elements at byte offset `base`, the first nv of them loaded, the
    // rest zero (masked loads read at kOOB, which returns zero).
    __device__ inline i32x4_t gather(uint32_t base, int nv) const {
        i32x4_t x;
        if constexpr (ES == 4) {
#pragma unroll
            for (int e = 0; e < 4; ++e)
                x[e] = (int)__builtin_amdgcn_raw_buffer_load_b32(rsrc, e < nv ? base + 4 * e : kOOB, 0, 0);
        } else if constexpr (ES == 2) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint32_t lo = __builtin_amdgcn_raw_buffer_load_b16(rsrc, 2 * e < nv ? base + 4 * e : kOOB, 0, 0);
                const uint32_t hi = __builtin_amdgcn_raw_buffer_load_b16(rsrc, 2 * e + 1 < nv ? base + 4 * e + 2 : kOOB,
                                                                         0, 0);
                x[e] = (int)(lo | (hi << 16));
            }
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                uint32_t w = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    w |= (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(rsrc, 4 * e + q < nv ? base + 4 * e + q : kOOB,
                                                                       0, 0) << (8 * q);
                x[e] = (int)w;
            }
        }
        return x;
    }

    // Write the K-tile in register slot S into one LDS stage (K-contiguous
    // 128-byte rows).
    template <int S = 0>
    __device__ inline void store(char* stage) const {
        const int t = tid;
        if constexpr (KC) {
            const int c = t & 7;
#pragma unroll
            for (int i = 0; i < NLD; ++i) {
                int r = (t >> 3) + 32 * i;
                int lr = PERM ? perm_a_row(r) : r;
                if constexpr (!X3) {
                    *(i32x4_t*)(stage + lds_addr(lr, c * 16)) = v[S][i];
                } else {
                    // 4 fp32 of K [4c, 4c+4) -> hi bf16 at bytes 8c, lo at 64 + 8c
                    i32x2_t hi, lo;
                    split4(v[S][i], hi, lo);
                    *(i32x2_t*)(stage + lds_addr(lr, c * 8)) = hi;
                    *(i32x2_t*)(stage + lds_addr(lr, 64 + c * 8)) = lo;
                }
            }
        } else {
            constexpr int NC = MT / VE;
            const int mc = t % NC, kg = t / NC;
            if constexpr (ES == 2) {
                // KB rows of 8 x 16-bit -> 8 rows of KB x 16-bit
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    int r = mc * 8 + i;
                    int lr = PERM ? perm_a_row(r) : r;
                    uint32_t w[KB / 2];
#pragma unroll
                    for (int p = 0; p < KB / 2; ++p) {
                        uint32_t a = (uint32_t)v[S][2 * p][i >> 1];
                        uint32_t b = (uint32_t)v[S][2 * p + 1][i >> 1];
                        w[p] = (i & 1) ? ((a >> 16) | (b & 0xffff0000u)) : ((a & 0xffffu) | (b << 16));
                    }
                    char* dst = stage + lds_addr(lr, kg * KB * 2);
                    if constexpr (KB == 8) *(i32x4_t*)dst = i32x4_t{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
                    else if constexpr (KB == 4) *(i32x2_t*)dst = i32x2_t{(int)w[0], (int)w[1]};
                    else *(int*)dst = (int)w[0];
                }
            } else {
                // fp32: KB rows of 4 floats -> 4 rows of KB floats
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    int r = mc * 4 + i;
                    int lr = PERM ? perm_a_row(r) : r;
                    if constexpr (!X3) {
#pragma unroll
                        for (int h = 0; h < KB / 4; ++h) {
                            i32x4_t o = {v[S][4 * h][i], v[S][4 * h + 1][i], v[S][4 * h + 2][i], v[S][4 * h + 3][i]};
                            *(i32x4_t*)(stage + lds_addr(lr, kg * KB * 4 + 16 * h)) = o;
                        }
                        if constexpr (KB == 2) {
                            *(i32x2_t*)(stage + lds_addr(lr, kg * 8)) = i32x2_t{v[S][0][i], v[S][1][i]};
                        }
                    } else {
                        uint32_t hw[KB / 2], lw[KB / 2];
#pragma unroll
                        for (int p = 0; p < KB / 2; ++p) {
                            float a = as_float(v[S][2 * p][i]);
                            float b = as_float(v[S][2 * p + 1][i]);
                            split2(a, b, hw[p], lw[p]);
                        }
                        char* dh = stage + lds_addr(lr, kg * KB * 2);
                        char* dl = stage + lds_addr(lr, 64 + kg * KB * 2);
                        if constexpr (KB == 8) {
                            *(i32x4_t*)dh = i32x4_t{(int)hw[0], (int)hw[1], (int)hw[2], (int)hw[3]};
                            *(i32x4_t*)dl = i32x4_t{(int)lw[0], (int)lw[1], (int)lw[2], (int)lw[3]};
                        } else if constexpr (KB == 4) {
                            *(i32x2_t*)dh = i32x2_t{(int)hw[0], (int)hw[1]};
                            *(i32x2_t*)dl = i32x2_t{(int)lw[0], (int)lw[1]};
                        } else {
                            *(int*)dh = (int)hw[0];
                            *(int*)dl = (int)lw[0];
                        }
                    }
                }
            }
        }
    }

    // bf16x3 split (Henry, Tang & Heinecke 2019): hi = rne(x), lo = rne(x - hi)
    __device__ static inline void split2(float a, float b, uint32_t& hi, uint32_t& lo) {
        __bf16 ha = (__bf16)a, hb = (__bf16)b;
        __bf16 la = (__bf16)(a - (float)ha), lb = (__bf16)(b - (float)hb);
        hi = (uint32_t)__builtin_bit_cast(uint16_t, ha) | ((uint32_t)__builtin_bit_cast(uint16_t, hb) << 16);
        lo = (uint32_t)__builtin_bit_cast(uint16_t, la) | ((uint32_t)__builtin_bit_cast(uint16_t, lb) << 16);
    }
    __device__ static inline void split4(i32x4_t x, i32x2_t& hi, i32x2_t& lo) {
        uint32_t h0, l0, h1, l1;
        split2(as_float(x[0]), as_float(x[1]), h0, l0);
        split2(as_float(x[2]), as_float(x[3]), h1, l1);
        hi = i32x2_t{(int)h0, (int)h1};
        lo = i32x2_t{(int)l0, (int)l1};
    }
};

// ---------------------------------------------------------------------------
// MFMA operand fragments and the per-read-step MFMA issue
// ---------------------------------------------------------------------------
template <int MODE>
struct Frag;

template <> struct Frag<kBF16> { i32x4_t x; };
template <> struct Frag<kF16> { i32x4_t x; };
template <> struct Frag<kF32> { i32x4_t x; };
template <> struct Frag<kX3> { i32x4_t hi, lo; };
template <> struct Frag<kF8> { i32x4_t x0, x1; };

// read steps per K-tile
template <int MODE>
__host__ __device__ constexpr int read_steps() { return (MODE == kX3 || MODE == kF8) ? 2 : 4; }

// Lane (row r of the operand stage, lane group g = lane/32) reads the chunks
// that carry its K slice for read step rs.  A and B use the same assignment,
// so every MFMA pairs matching K indices.
template <int MODE>
__device__ inline void read_frag(const char* stage, int row, int g, int rs, Frag<MODE>& f) {
    if constexpr (MODE == kX3) {
        f.hi = *(const i32x4_t*)(stage + lds_addr(row, (2 * rs + g) * 16));
        f.lo = *(const i32x4_t*)(stage + lds_addr(row, (4 + 2 * rs + g) * 16));
    } else if constexpr (MODE == kF8) {
        f.x0 = *(const i32x4_t*)(stage + lds_addr(row, (4 * rs + g) * 16));
        f.x1 = *(const i32x4_t*)(stage + lds_addr(row, (4 * rs + 2 + g) * 16));
    } else {
        f.x = *(const i32x4_t*)(stage + lds_addr(row, (2 * rs + g) * 16));
    }
}

template <int MODE>
__device__ inline void mfma(const Frag<MODE>& a, const Frag<MODE>& b, f32x16_t& acc) {
    if constexpr (MODE == kBF16) {
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a.x),
                                                      __builtin_bit_cast(bf16x8_t, b.x), acc, 0, 0, 0);
    } else if constexpr (MODE == kF16) {
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8_t, a.x),
                                                     __builtin_bit_cast(f16x8_t, b.x), acc, 0, 0, 0);
    } else if constexpr (MODE == kF32) {
#pragma unroll
        for (int e = 0; e < 4; ++e)
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(as_float(a.x[e]), as_float(b.x[e]), acc, 0, 0, 0);
    } else if constexpr (MODE == kX3) {
        // small cross terms first, then the dominant hi*hi product
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a.lo),
                                                      __builtin_bit_cast(bf16x8_t, b.hi), acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a.hi),
                                                      __builtin_bit_cast(bf16x8_t, b.lo), acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a.hi),
                                                      __builtin_bit_cast(bf16x8_t, b.hi), acc, 0, 0, 0);
    } else {
        i32x8_t av = {a.x0[0], a.x0[1], a.x0[2], a.x0[3], a.x1[0], a.x1[1], a.x1[2], a.x1[3]};
        i32x8_t bv = {b.x0[0], b.x0[1], b.x0[2], b.x0[3], b.x1[0], b.x1[1], b.x1[2], b.x1[3]};
        // format 0 = OCP e4m3 for both operands; e8m0 scale 127 = 1.0
        acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, acc, 0, 0, 0, 127, 0, 127);
    }
}

// ---------------------------------------------------------------------------
// Instruction interleaving for one read step: the MFMAs of the step with the
// LDS fragment reads of the next step spread between them, plus `nv` global
// loads and `nw` LDS writes (sched_group_barrier masks: 0x008 MFMA,
// 0x020 VMEM read, 0x100 DS read, 0x200 DS write).
// ---------------------------------------------------------------------------
template <int MODE, int TM, int TN, int SYNC = 0>
__device__ inline void sched_step(int nv, int nw) {
    constexpr int MPS = (MODE == kF32) ? 4 : (MODE == kX3 ? 3 : 1);
    constexpr int RPF = (MODE == kX3 || MODE == kF8) ? 2 : 1;
    constexpr int NM = TM * TN * MPS;
    constexpr int NR = (TM + TN) * RPF;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, SYNC);
        if (i < nv) __builtin_amdgcn_sched_group_barrier(0x020, 1, SYNC);
        if ((i * NR) / NM != ((i + 1) * NR) / NM) __builtin_amdgcn_sched_group_barrier(0x100, 1, SYNC);
        if (i >= NM / 2 && (i - NM / 2) < nw) __builtin_amdgcn_sched_group_barrier(0x200, 1, SYNC);
    }
}

// ---------------------------------------------------------------------------
// Epilogue: D = alpha*acc + beta*C (fp32 FMA), converted to the output type.
// Lane l of MFMA tile (ti, tj) holds the 16 consecutive rows
// 16*(l/32) + 0..15 of column (l%32).
// ---------------------------------------------------------------------------
// ANY: element by element always (the general-stride form: ldc, ldd and the
// pointers need only their element alignment).
template <int OUT, bool ANY = false>
__device__ inline void store16(const TileArgs& a, const char* Cb, char* Db, int m, int n, const float* acc,
                               float alpha, float beta) {
    constexpr int OS = (OUT == kOutF32) ? 4 : 2;
    const bool use_c = (a.flags & kFlagBeta) != 0;
    char* dp = Db + ((int64_t)n * a.ldd + m) * OS;
    const char* cp = Cb + ((int64_t)n * a.ldc + m) * OS;
    if (!ANY && m + 16 <= a.M) {
        float r[16];
        if (use_c) {
            float c[16];
            if constexpr (OUT == kOutF32) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    f32x4_t cv = *(const f32x4_t*)(cp + 16 * q);
                    c[4 * q] = cv[0]; c[4 * q + 1] = cv[1]; c[4 * q + 2] = cv[2]; c[4 * q + 3] = cv[3];
                }
            } else {
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    i32x4_t cv = *(const i32x4_t*)(cp + 16 * q);
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        uint32_t u = (uint32_t)cv[e];
                        c[8 * q + 2 * e] = (OUT == kOutBF16) ? bf16lo(u) : f16lo(u);
                        c[8 * q + 2 * e + 1] = (OUT == kOutBF16) ? bf16hi(u) : f16hi(u);
                    }
                }
            }
#pragma unroll
            for (int e = 0; e < 16; ++e) r[e] = __builtin_fmaf(alpha, acc[e], beta * c[e]);
        } else {
#pragma unroll
            for (int e = 0; e < 16; ++e) r[e] = alpha * acc[e];
        }
        if constexpr (OUT == kOutF32) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                *(f32x4_t*)(dp + 16 * q) = f32x4_t{r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]};
        } else {
            const bool trunc = (a.flags & kFlagTrunc) != 0;
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                uint32_t w[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    float x = r[8 * q + 2 * e], y = r[8 * q + 2 * e + 1];
                    if constexpr (OUT == kOutBF16) w[e] = trunc ? pack_bf16x2_trunc(x, y) : pack_bf16x2_rne(x, y);
                    else w[e] = pack_f16x2(x, y);
                }
                *(i32x4_t*)(dp + 16 * q) = i32x4_t{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
            }
        }
    } else {
        for (int e = 0; e < 16 && m + e < a.M; ++e) {
            float c = 0.0f;
            if (use_c) {
                if constexpr (OUT == kOutF32) c = ((const float*)cp)[e];
                else {
                    uint16_t u = ((const uint16_t*)cp)[e];
                    c = (OUT == kOutBF16) ? bf16lo(u) : f16lo(u);
                }
            }
            float r = use_c ? __builtin_fmaf(alpha, acc[e], beta * c) : alpha * acc[e];
            if constexpr (OUT == kOutF32) ((float*)dp)[e] = r;
            else if constexpr (OUT == kOutBF16) {
                uint32_t w = (a.flags & kFlagTrunc) ? pack_bf16x2_trunc(r, 0.f) : pack_bf16x2_rne(r, 0.f);
                ((uint16_t*)dp)[e] = (uint16_t)(w & 0xffffu);
            } else {
                ((uint16_t*)dp)[e] = (uint16_t)(pack_f16x2(r, 0.f) & 0xffffu);
            }
        }
    }
}

// Epilogue of one lane of a 16 x 16 MFMA block: rows m .. m+3 of column n.
template <int OUT>
__device__ inline void store4(const TileArgs& a, const char* Cb, char* Db, int m, int n, const f32x4_t& acc,
                              float alpha, float beta) {
    constexpr int OS = (OUT == kOutF32) ? 4 : 2;
    const bool use_c = (a.flags & kFlagBeta) != 0;
    const bool trunc = (a.flags & kFlagTrunc) != 0;
    char* dp = Db + ((int64_t)n * a.ldd + m) * OS;
    const char* cp = Cb + ((int64_t)n * a.ldc + m) * OS;
    if (m + 4 <= a.M) {
        float c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (use_c) {
            if constexpr (OUT == kOutF32) {
                const f32x4_t cv = *(const f32x4_t*)cp;
                c[0] = cv[0]; c[1] = cv[1]; c[2] = cv[2]; c[3] = cv[3];
            } else {
                const i32x2_t cv = *(const i32x2_t*)cp;
                const uint32_t u0 = (uint32_t)cv[0], u1 = (uint32_t)cv[1];
                c[0] = (OUT == kOutBF16) ? bf16lo(u0) : f16lo(u0);
                c[1] = (OUT == kOutBF16) ? bf16hi(u0) : f16hi(u0);
                c[2] = (OUT == kOutBF16) ? bf16lo(u1) : f16lo(u1);
                c[3] = (OUT == kOutBF16) ? bf16hi(u1) : f16hi(u1);
            }
        }
        float r[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) r[e] = use_c ? __builtin_fmaf(alpha, acc[e], beta * c[e]) : alpha * acc[e];
        if constexpr (OUT == kOutF32) {
            *(f32x4_t*)dp = f32x4_t{r[0], r[1], r[2], r[3]};
        } else if constexpr (OUT == kOutBF16) {
            const uint32_t w0 = trunc ? pack_bf16x2_trunc(r[0], r[1]) : pack_bf16x2_rne(r[0], r[1]);
            const uint32_t w1 = trunc ? pack_bf16x2_trunc(r[2], r[3]) : pack_bf16x2_rne(r[2], r[3]);
            *(i32x2_t*)dp = i32x2_t{(int)w0, (int)w1};
        } else {
            *(i32x2_t*)dp = i32x2_t{(int)pack_f16x2(r[0], r[1]), (int)pack_f16x2(r[2], r[3])};
        }
    } else {
        for (int e = 0; e < 4 && m + e < a.M; ++e) {
            float c = 0.0f;
            if (use_c) {
                if constexpr (OUT == kOutF32) c = ((const float*)cp)[e];
                else {
                    const uint16_t u = ((const uint16_t*)cp)[e];
                    c = (OUT == kOutBF16) ? bf16lo(u) : f16lo(u);
                }
            }
            const float r = use_c ? __builtin_fmaf(alpha, acc[e], beta * c) : alpha * acc[e];
            if constexpr (OUT == kOutF32) ((float*)dp)[e] = r;
            else if constexpr (OUT == kOutBF16)
                ((uint16_t*)dp)[e] = (uint16_t)((trunc ? pack_bf16x2_trunc(r, 0.f) : pack_bf16x2_rne(r, 0.f)) & 0xffffu);
            else ((uint16_t*)dp)[e] = (uint16_t)(pack_f16x2(r, 0.f) & 0xffffu);
        }
    }
}

// store16 for a fragment wholly inside D whose 16 C values (16-bit) were
// loaded ahead into c0, c1; beta != 0.
template <int OUT>
__device__ inline void store16_pre(const TileArgs& a, char* Db, int m, int n, const float* acc, i32x4_t c0,
                                   i32x4_t c1, float alpha, float beta) {
    static_assert(OUT != kOutF32, "16-bit C");
    float r[16];
    const i32x4_t cv[2] = {c0, c1};
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t u = (uint32_t)cv[q][e];
            const float lo = (OUT == kOutBF16) ? bf16lo(u) : f16lo(u), hi = (OUT == kOutBF16) ? bf16hi(u) : f16hi(u);
            r[8 * q + 2 * e] = __builtin_fmaf(alpha, acc[8 * q + 2 * e], beta * lo);
            r[8 * q + 2 * e + 1] = __builtin_fmaf(alpha, acc[8 * q + 2 * e + 1], beta * hi);
        }
    char* dp = Db + ((int64_t)n * a.ldd + m) * 2;
    const bool trunc = (a.flags & kFlagTrunc) != 0;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        uint32_t w[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float lo = r[8 * q + 2 * e], hi = r[8 * q + 2 * e + 1];
            if constexpr (OUT == kOutBF16) w[e] = trunc ? pack_bf16x2_trunc(lo, hi) : pack_bf16x2_rne(lo, hi);
            else w[e] = pack_f16x2(lo, hi);
        }
        *(i32x4_t*)(dp + 16 * q) = i32x4_t{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
    }
}

// Epilogue of a wave's NF accumulator fragments held in VGPRs (the eight-wave
// engine's 128 registers); fragment f covers rows m[f] .. m[f]+15 of column
// n[f] (lane layout as in store16).  With beta != 0 the C slices of up to NB
// fragments are requested before the first one is used, so their memory
// latencies overlap instead of adding up; a fragment not wholly inside D reads
// C's first 16 rows instead (valid when M >= 16) and takes store16's
// per-element path.  The beta and rounding choices are template parameters
// behind one uniform branch each.  (The four-wave engines keep 256
// accumulator registers in AGPRs and store fragment by fragment with store16.)
template <int OUT, int NF, int NB, bool USE_C, bool TRUNC>
__device__ inline void store_frags_as(const TileArgs& a, const char* Cb, char* Db, const int* m, const int* n,
                                      const f32x16_t* acc, float alpha, float beta) {
    constexpr int OS = (OUT == kOutF32) ? 4 : 2;
    constexpr int CV = OS;  // 16-byte C vectors per fragment: 16 values of OS bytes
    static_assert(NF % NB == 0 || NF < NB, "fragment batches");
    constexpr int B = NF < NB ? NF : NB;
#pragma unroll
    for (int f0 = 0; f0 < NF; f0 += B) {
        i32x4_t cv[B][CV];
        if constexpr (USE_C) {
            if (a.M >= 16) {
#pragma unroll
                for (int f = 0; f < B; ++f) {
                    const bool full = n[f0 + f] < a.N && m[f0 + f] + 16 <= a.M;
                    const char* cp = full ? Cb + ((int64_t)n[f0 + f] * a.ldc + m[f0 + f]) * OS : Cb;
#pragma unroll
                    for (int q = 0; q < CV; ++q) cv[f][q] = *(const i32x4_t*)(cp + 16 * q);
                }
            }
        }
#pragma unroll
        for (int f = 0; f < B; ++f) {
            const int mm = m[f0 + f], nn = n[f0 + f];
            if (nn >= a.N || mm >= a.M) continue;
            const f32x16_t& x = acc[f0 + f];
            if (mm + 16 > a.M) {  // rows past M: element by element
                float v[16];
#pragma unroll
                for (int e = 0; e < 16; ++e) v[e] = x[e];
                store16<OUT>(a, Cb, Db, mm, nn, v, alpha, beta);
                continue;
            }
            float r[16];
            if constexpr (USE_C) {
                float c[16];
#pragma unroll
                for (int q = 0; q < CV; ++q) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const uint32_t u = (uint32_t)cv[f][q][e];
                        if constexpr (OUT == kOutF32) {
                            c[4 * q + e] = __builtin_bit_cast(float, u);
                        } else {
                            c[8 * q + 2 * e] = (OUT == kOutBF16) ? bf16lo(u) : f16lo(u);
                            c[8 * q + 2 * e + 1] = (OUT == kOutBF16) ? bf16hi(u) : f16hi(u);
                        }
                    }
                }
#pragma unroll
                for (int e = 0; e < 16; ++e) r[e] = __builtin_fmaf(alpha, x[e], beta * c[e]);
            } else {
#pragma unroll
                for (int e = 0; e < 16; ++e) r[e] = alpha * x[e];
            }
            char* dp = Db + ((int64_t)nn * a.ldd + mm) * OS;
            if constexpr (OUT == kOutF32) {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    *(f32x4_t*)(dp + 16 * q) = f32x4_t{r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]};
            } else {
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    uint32_t w[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const float lo = r[8 * q + 2 * e], hi = r[8 * q + 2 * e + 1];
                        if constexpr (OUT == kOutBF16) w[e] = TRUNC ? pack_bf16x2_trunc(lo, hi) : pack_bf16x2_rne(lo, hi);
                        else w[e] = pack_f16x2(lo, hi);
                    }
                    *(i32x4_t*)(dp + 16 * q) = i32x4_t{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
                }
            }
        }
    }
}

// NB: fragments whose C is in flight together; each holds 2 (16-bit C) or 4
// (fp32 C) 16-byte vectors, so engines size it to their free registers.
template <int OUT, int NF, int NB>
__device__ inline void store_frags(const TileArgs& a, const char* Cb, char* Db, const int* m, const int* n,
                                   const f32x16_t* acc, float alpha, float beta) {
    const bool trunc = OUT == kOutBF16 && (a.flags & kFlagTrunc) != 0;
    if (a.flags & kFlagBeta) {
        if (trunc) store_frags_as<OUT, NF, NB, true, true>(a, Cb, Db, m, n, acc, alpha, beta);
        else store_frags_as<OUT, NF, NB, true, false>(a, Cb, Db, m, n, acc, alpha, beta);
    } else {
        if (trunc) store_frags_as<OUT, NF, NB, false, true>(a, Cb, Db, m, n, acc, alpha, beta);
        else store_frags_as<OUT, NF, NB, false, false>(a, Cb, Db, m, n, acc, alpha, beta);
    }
}

// Extent in bytes of one batch's operand as a buffer: the last row ends at
// `tail` elements, rounded up to the 16-byte vector that holds it (a partly
// out-of-range vector would read as all zeros; an aligned 16-byte vector never
// crosses a page, so the rounded extent stays inside mapped memory).  Clamped
// to 2 GiB - 16; the host rejects problems that do not fit.
__device__ inline uint32_t extent_bytes(int64_t rows, int64_t cols_ld, int64_t tail, int es) {
    const int64_t ve = 16 / es;
    int64_t e = ((rows - 1) * cols_ld + (tail + ve - 1) / ve * ve) * es;
    if (e < 0) e = 0;
    if (e > 0x7ffffff0ll) e = 0x7ffffff0ll;
    return (uint32_t)e;
}
// The same to the last element exactly (the general-stride form's element
// loads never read past it).
__device__ inline uint32_t extent_bytes_exact(int64_t rows, int64_t cols_ld, int64_t tail, int es) {
    int64_t e = ((rows - 1) * cols_ld + tail) * es;
    if (e < 0) e = 0;
    if (e > 0x7ffffff0ll) e = 0x7ffffff0ll;
    return (uint32_t)e;
}

// ---------------------------------------------------------------------------
// One output tile.  smem: 2 stages x (MT_M + MT_N) x 128 bytes.
// ---------------------------------------------------------------------------
// ANY: the general-stride form (element-wise loads and stores, see Loader).
template <int MODE, bool AK, bool BK, int MT_M, int MT_N, int OUT, bool ANY = false>
__device__ inline void gemm_tile(const TileArgs& a, int tm, int tn, int b, char* smem, int tid) {
    constexpr int ES = mode_elem_bytes(MODE);
    constexpr int DU = mode_depth_u(MODE);
    constexpr int WT_M = MT_M / 2, WT_N = MT_N / 2;
    constexpr int TM = WT_M / 32, TN = WT_N / 32;
    constexpr int RS = read_steps<MODE>();
    constexpr int STAGE = (MT_M + MT_N) * kRowBytes;
    static_assert(TM >= 1 && TN >= 1, "wave tile below one MFMA tile");

    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int m0 = tm * MT_M, n0 = tn * MT_N;

    const char* Ab = (const char*)a.A + (int64_t)b * a.bsA * ES;
    const char* Bb = (const char*)a.B + (int64_t)b * a.bsB * ES;

    // Tiles up to 128 x 128 keep two K-tiles of loads in flight (a register
    // slot each): each K-tile otherwise waits on the memory latency with only
    // the next K-tile's loads in flight.  The larger tiles lack the registers.
    constexpr int PF = (MT_M * MT_N <= 128 * 128) ? 2 : 1;
    Loader<MODE, AK, MT_M, true, PF, ANY> la;
    Loader<MODE, BK, MT_N, false, PF, ANY> lb;
    // A: K-contiguous = rows M of length lda; M-contiguous = rows K of length lda
    // A K-contiguous operand's buffer starts at the tile's first row, so its
    // offsets stay within the tile's rows however large the matrix is.
    auto ext = [](int64_t rows, int64_t ld, int64_t tail) {
        return ANY ? extent_bytes_exact(rows, ld, tail, ES) : extent_bytes(rows, ld, tail, ES);
    };
    if constexpr (AK) la.init(Ab + (int64_t)m0 * a.lda * ES, ext(a.M - m0, a.lda, a.K), a.lda, a.M - m0, a.K, 0, tid);
    else la.init(Ab, ext(a.K, a.lda, a.M), a.lda, a.M, a.K, m0, tid);
    if constexpr (BK) lb.init(Bb + (int64_t)n0 * a.ldb * ES, ext(a.N - n0, a.ldb, a.K), a.ldb, a.N - n0, a.K, 0, tid);
    else lb.init(Bb, ext(a.K, a.ldb, a.N), a.ldb, a.N, a.K, n0, tid);

    f32x16_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;

    const int KT = (a.K + DU - 1) / DU;
    const int g = lane >> 5;
    const int arow = wm * WT_M + (lane & 31);
    const int brow = wn * WT_N + (lane & 31);

    // Software pipeline (one LDS barrier per K-tile):
    //   step 0      : loads of K-tile kt+1 -> VGPRs, interleaved with MFMAs
    //   step rs     : LDS reads of step rs+1 interleaved with the MFMAs of step rs
    //   step RS-2   : VGPRs of K-tile kt+1 -> the other LDS stage
    //   barrier     : then the MFMAs of the last step run while the first
    //                 fragments of K-tile kt+1 are read, so the MFMA pipe does
    //                 not drain across the barrier.
    // Loads past K are masked to zero, so the loop has no branches.
    // With PF = 2, K-tile kt sits in register slot kt & 1 until it is
    // written to LDS; iteration kt refills that slot with K-tile kt + 2.
    Frag<MODE> fa[2][TM], fb[2][TN];
    la.template load<0>(0);
    lb.template load<0>(0);
    if constexpr (PF == 2) {
        la.template load<1>(1);
        lb.template load<1>(1);
    }
    la.template store<0>(smem);
    lb.template store<0>(smem + MT_M * kRowBytes);
    lds_barrier();
#pragma unroll
    for (int i = 0; i < TM; ++i) read_frag<MODE>(smem, arow + 32 * i, g, 0, fa[0][i]);
#pragma unroll
    for (int j = 0; j < TN; ++j) read_frag<MODE>(smem + MT_M * kRowBytes, brow + 32 * j, g, 0, fb[0][j]);

    auto iteration = [&](auto slot, int kt) {
        constexpr int S = decltype(slot)::value;     // slot of K-tile kt (0 when PF = 1)
        constexpr int SN = (PF == 2) ? (S ^ 1) : 0;  // slot of K-tile kt + 1
        const char* sa = smem + (kt & 1) * STAGE;
        const char* sb = sa + MT_M * kRowBytes;
        char* na = smem + ((kt + 1) & 1) * STAGE;
        la.template load<S>(kt + PF);
        lb.template load<S>(kt + PF);
#pragma unroll
        for (int rs = 0; rs + 1 < RS; ++rs) {
#pragma unroll
            for (int i = 0; i < TM; ++i) read_frag<MODE>(sa, arow + 32 * i, g, rs + 1, fa[(rs + 1) & 1][i]);
#pragma unroll
            for (int j = 0; j < TN; ++j) read_frag<MODE>(sb, brow + 32 * j, g, rs + 1, fb[(rs + 1) & 1][j]);
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) mfma<MODE>(fa[rs & 1][i], fb[rs & 1][j], acc[i][j]);
            if (rs + 2 == RS) {
                la.template store<SN>(na);
                lb.template store<SN>(na + MT_M * kRowBytes);
            }
            sched_step<MODE, TM, TN>(rs == 0 && !ANY ? la.NLD + lb.NLD : 0,
                                     rs + 2 == RS ? la.NSTORE + lb.NSTORE : 0);
        }
        lds_barrier();
#pragma unroll
        for (int i = 0; i < TM; ++i) read_frag<MODE>(na, arow + 32 * i, g, 0, fa[0][i]);
#pragma unroll
        for (int j = 0; j < TN; ++j) read_frag<MODE>(na + MT_M * kRowBytes, brow + 32 * j, g, 0, fb[0][j]);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) mfma<MODE>(fa[(RS - 1) & 1][i], fb[(RS - 1) & 1][j], acc[i][j]);
        sched_step<MODE, TM, TN>(0, 0);
    };
    if constexpr (PF == 1) {
        for (int kt = 0; kt < KT; ++kt) iteration(std::integral_constant<int, 0>{}, kt);
    } else {
        int kt = 0;
        for (; kt + 2 <= KT; kt += 2) {
            iteration(std::integral_constant<int, 0>{}, kt);
            iteration(std::integral_constant<int, 1>{}, kt + 1);
        }
        if (kt < KT) iteration(std::integral_constant<int, 0>{}, kt);
    }

    // epilogue
    float alpha = a.alpha;
    if constexpr (MODE == kF8) {
        float sa = a.scaleA ? *a.scaleA : 1.0f;
        float sb = a.scaleB ? *a.scaleB : 1.0f;
        alpha = alpha / (sa * sb);
    }
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
                store16<OUT, ANY>(a, Cb, Db, m, n, v, alpha, a.beta);
            }
        }
    }
}

// Tile index -> (batch, tile_m, tile_n).  Workgroups are dispatched to the 8
// XCDs round-robin (block b runs on XCD b % 8); the remap gives each XCD a
// contiguous range of the grouped raster so its L2 sees a compact block of
// A rows and B columns.
__device__ inline void map_tile(int bid, int nblocks, const TileArgs& a, int& tb, int& tm, int& tn) {
    int per = nblocks >> 3;
    int v = (bid < (per << 3)) ? ((bid & 7) * per + (bid >> 3)) : bid;
    const int tiles = a.tiles_m * a.tiles_n;
    tb = v / tiles;
    int t = v - tb * tiles;
    const int gm = a.group_m > 0 ? a.group_m : 1;
    const int gsize = gm * a.tiles_n;
    const int grp = t / gsize;
    const int first = grp * gm;
    const int h = min(a.tiles_m - first, gm);
    const int r = t - grp * gsize;
    tm = first + r % h;
    tn = r / h;
}

template <int MODE, bool AK, bool BK, int MT_M, int MT_N, int OUT>
__global__ void __launch_bounds__(kThreads, 1) gemm_kernel(TileArgs a) {
    __shared__ __attribute__((aligned(16))) char smem[2 * (MT_M + MT_N) * kRowBytes];
    int tb, tm, tn;
    map_tile(blockIdx.x, gridDim.x, a, tb, tm, tn);
    gemm_tile<MODE, AK, BK, MT_M, MT_N, OUT>(a, tm, tn, tb, smem, threadIdx.x);
}

// The general-stride form: any leading dimension, batch stride and K, pointers
// aligned to their element size; MT x MT tiles.
template <int MODE, bool AK, bool BK, int MT, int OUT>
__global__ void __launch_bounds__(kThreads, 1) gemm_kernel_any(TileArgs a) {
    __shared__ __attribute__((aligned(16))) char smem[2 * (MT + MT) * kRowBytes];
    int tb, tm, tn;
    map_tile(blockIdx.x, gridDim.x, a, tb, tm, tn);
    gemm_tile<MODE, AK, BK, MT, MT, OUT, true>(a, tm, tn, tb, smem, threadIdx.x);
}

// ---------------------------------------------------------------------------
// The same tile with 16 x 16 MFMAs (tile record mfma_mn = 16): 16 x 16 x 32 for
// bf16 and fp16, 16 x 16 x 4 for fp32.  Lane l supplies row (column) l % 16 of
// its operand block from 16-byte chunk l / 16 of the read step: 8 consecutive
// K of 16-bit data (one MFMA covers 32 K), or 4 fp32 (four MFMAs, element e of
// every lane's chunk in MFMA e: 16 K).  Either way a 128-byte K-tile takes two
// read steps, and A and B lanes pair the same K.  The result block gives lane l
// rows 4 * (l / 16) .. +3 of column l % 16.  Those rows are consecutive in the
// natural row order, so A is staged unpermuted.  The pipeline is the 32 x 32
// tile's with two read steps and two K-tiles of loads in flight.
// ---------------------------------------------------------------------------
template <int MODE>
__device__ inline void mfma16x16(const i32x4_t& a, const i32x4_t& b, f32x4_t& acc) {
    if constexpr (MODE == kBF16) {
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b),
                                                      acc, 0, 0, 0);
    } else if constexpr (MODE == kF16) {
        acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8_t, a), __builtin_bit_cast(f16x8_t, b),
                                                     acc, 0, 0, 0);
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(as_float(a[e]), as_float(b[e]), acc, 0, 0, 0);
    }
}

template <int MODE, bool AK, bool BK, int MT_M, int MT_N, int OUT>
__device__ inline void gemm_tile_m16(const TileArgs& a, int tm, int tn, int b, char* smem, int tid) {
    static_assert(MODE == kBF16 || MODE == kF16 || MODE == kF32, "16 x 16 MFMA: 16-bit or fp32 operands");
    constexpr int ES = mode_elem_bytes(MODE);
    constexpr int DU = mode_depth_u(MODE);
    constexpr int WT_M = MT_M / 2, WT_N = MT_N / 2;
    constexpr int TM = WT_M / 16, TN = WT_N / 16;
    constexpr int STAGE = (MT_M + MT_N) * kRowBytes;
    static_assert(kRowBytes == 128, "two read steps of 64 bytes of K per K-tile");

    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int m0 = tm * MT_M, n0 = tn * MT_N;

    const char* Ab = (const char*)a.A + (int64_t)b * a.bsA * ES;
    const char* Bb = (const char*)a.B + (int64_t)b * a.bsB * ES;
    // two K-tiles of loads in flight (register slot kt & 1 holds K-tile kt)
    Loader<MODE, AK, MT_M, false, 2> la;
    Loader<MODE, BK, MT_N, false, 2> lb;
    if constexpr (AK) la.init(Ab + (int64_t)m0 * a.lda * ES, extent_bytes(a.M - m0, a.lda, a.K, ES), a.lda, a.M - m0, a.K, 0, tid);
    else la.init(Ab, extent_bytes(a.K, a.lda, a.M, ES), a.lda, a.M, a.K, m0, tid);
    if constexpr (BK) lb.init(Bb + (int64_t)n0 * a.ldb * ES, extent_bytes(a.N - n0, a.ldb, a.K, ES), a.ldb, a.N - n0, a.K, 0, tid);
    else lb.init(Bb, extent_bytes(a.K, a.ldb, a.N, ES), a.ldb, a.N, a.K, n0, tid);

    f32x4_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};

    const int KT = (a.K + DU - 1) / DU;
    const int q = lane >> 4;
    const int arow = wm * WT_M + (lane & 15);
    const int brow = wn * WT_N + (lane & 15);
    // fragment of read step rs (K 32*rs .. 32*rs+31 of the K-tile)
    auto rd = [&](const char* st, int row, int rs) { return *(const i32x4_t*)(st + lds_addr(row, (4 * rs + q) * 16)); };

    i32x4_t fa[2][TM], fb[2][TN];
    la.template load<0>(0);
    lb.template load<0>(0);
    la.template load<1>(1);
    lb.template load<1>(1);
    la.template store<0>(smem);
    lb.template store<0>(smem + MT_M * kRowBytes);
    lds_barrier();
#pragma unroll
    for (int i = 0; i < TM; ++i) fa[0][i] = rd(smem, arow + 16 * i, 0);
#pragma unroll
    for (int j = 0; j < TN; ++j) fb[0][j] = rd(smem + MT_M * kRowBytes, brow + 16 * j, 0);

    auto iteration = [&](auto slot, int kt) {
        constexpr int S = decltype(slot)::value;
        const char* sa = smem + (kt & 1) * STAGE;
        const char* sb = sa + MT_M * kRowBytes;
        char* na = smem + ((kt + 1) & 1) * STAGE;
        la.template load<S>(kt + 2);
        lb.template load<S>(kt + 2);
        // read step 0: fragments of step 1 load while the MFMAs of step 0 run,
        // then K-tile kt+1 goes from registers to the other stage
#pragma unroll
        for (int i = 0; i < TM; ++i) fa[1][i] = rd(sa, arow + 16 * i, 1);
#pragma unroll
        for (int j = 0; j < TN; ++j) fb[1][j] = rd(sb, brow + 16 * j, 1);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) mfma16x16<MODE>(fa[0][i], fb[0][j], acc[i][j]);
        la.template store<S ^ 1>(na);
        lb.template store<S ^ 1>(na + MT_M * kRowBytes);
        sched_step<MODE, TM, TN>(la.NLD + lb.NLD, la.NSTORE + lb.NSTORE);
        lds_barrier();
        // read step 1's MFMAs run while the first fragments of K-tile kt+1 load
#pragma unroll
        for (int i = 0; i < TM; ++i) fa[0][i] = rd(na, arow + 16 * i, 0);
#pragma unroll
        for (int j = 0; j < TN; ++j) fb[0][j] = rd(na + MT_M * kRowBytes, brow + 16 * j, 0);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) mfma16x16<MODE>(fa[1][i], fb[1][j], acc[i][j]);
        sched_step<MODE, TM, TN>(0, 0);
    };
    int kt = 0;
    for (; kt + 2 <= KT; kt += 2) {
        iteration(std::integral_constant<int, 0>{}, kt);
        iteration(std::integral_constant<int, 1>{}, kt + 1);
    }
    if (kt < KT) iteration(std::integral_constant<int, 0>{}, kt);

    const char* Cb = (const char*)a.C + (int64_t)b * a.bsC * ((OUT == kOutF32) ? 4 : 2);
    char* Db = (char*)a.D + (int64_t)b * a.bsD * ((OUT == kOutF32) ? 4 : 2);
#pragma unroll
    for (int i = 0; i < TM; ++i) {
        const int m = m0 + wm * WT_M + 16 * i + 4 * q;
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int n = n0 + wn * WT_N + 16 * j + (lane & 15);
            if (n < a.N && m < a.M) store4<OUT>(a, Cb, Db, m, n, acc[i][j], a.alpha, a.beta);
        }
    }
}

template <int MODE, bool AK, bool BK, int MT_M, int MT_N, int OUT>
__global__ void __launch_bounds__(kThreads, 1) gemm_kernel_m16(TileArgs a) {
    __shared__ __attribute__((aligned(16))) char smem[2 * (MT_M + MT_N) * kRowBytes];
    int tb, tm, tn;
    map_tile(blockIdx.x, gridDim.x, a, tb, tm, tn);
    gemm_tile_m16<MODE, AK, BK, MT_M, MT_N, OUT>(a, tm, tn, tb, smem, threadIdx.x);
}

}  // namespace ffgemm
