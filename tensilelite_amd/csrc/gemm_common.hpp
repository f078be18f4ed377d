// Shared device-side definitions for the gfx950 GEMM kernels.
//
// Design (replaces the generated GCN kernel of src/gemmMfmaKernelWriter.h):
//  * 256 threads = 4 waves of 64 in a 2 x 2 grid; each wave owns a
//    (MT_M/2) x (MT_N/2) sub-tile made of 32x32 MFMA tiles.
//  * Every K-iteration moves 128 bytes of K per tile row ("DepthU" of 64
//    bf16/f16, 32 f32, 128 fp8).  Global -> VGPR uses 16-byte buffer loads
//    (out-of-range lanes read zero); VGPR -> LDS writes rows of 128 bytes in
//    8 chunks of 16 bytes whose chunk index is XOR-swizzled by the row, so
//    both the 16-byte LDS writes and the 16-byte MFMA operand reads are free
//    of bank conflicts.  Operands stored M/N-contiguous in HBM are transposed
//    in registers on their way to LDS, so LDS always holds K-contiguous rows.
//  * Two LDS stages: while the waves run MFMAs on stage s, the next K-tile
//    sits in VGPRs and is written to stage s^1 before a single barrier.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ffgemm {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8_t;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;
typedef __attribute__((__vector_size__(4 * sizeof(double)))) double f64x4_t;
typedef __attribute__((__vector_size__(4 * sizeof(int)))) int i32x4_t;
typedef __attribute__((__vector_size__(2 * sizeof(int)))) int i32x2_t;
typedef __attribute__((__vector_size__(8 * sizeof(int)))) int i32x8_t;

// Operand/compute modes.
enum Mode : int {
    kBF16 = 0,  // bf16 in, v_mfma_f32_32x32x16_bf16
    kF16 = 1,   // fp16 in, v_mfma_f32_32x32x16_f16
    kF32 = 2,   // fp32 in, v_mfma_f32_32x32x2_f32
    kX3 = 3,    // fp32 in, split hi/lo bf16, 3 x v_mfma_f32_32x32x16_bf16
    kF8 = 4     // OCP e4m3 in, v_mfma_scale_f32_32x32x64_f8f6f4
};

// Output element kinds.
enum OutKind : int { kOutBF16 = 0, kOutF16 = 1, kOutF32 = 2 };

constexpr int kThreads = 256;
constexpr int kRowBytes = 128;  // bytes of K per LDS row per stage
constexpr uint32_t kOOB = 0x80000000u;  // buffer offset that always reads 0

__host__ __device__ constexpr int mode_elem_bytes(int mode) {
    return (mode == kBF16 || mode == kF16) ? 2 : (mode == kF8 ? 1 : 4);
}
// K elements per main-loop iteration.
__host__ __device__ constexpr int mode_depth_u(int mode) { return kRowBytes / mode_elem_bytes(mode); }

// XOR swizzle of the 16-byte chunk index inside a 128-byte LDS row.  Rows
// r..r+7 map one logical chunk to 8 different physical chunks, and so do the
// rows r, r+8, ..., r+56 that the transposing writer touches together.
__host__ __device__ constexpr int swz(int row) { return (row ^ (row >> 3)) & 7; }

// Row permutation inside each 32-row block of the A tile.  Physical row
// m = 16h + 4g + q is stored at LDS row 8g + 4h + q, so the 32x32 MFMA
// accumulator of lane l holds the 16 consecutive rows 16*(l/32) + 0..15 of its
// column, which the epilogue writes as whole 16-byte vectors.
__host__ __device__ constexpr int perm_a_row(int r) {
    int m = r & 31;
    int h = (m >> 4) & 1, g = (m >> 2) & 3, q = m & 3;
    return (r & ~31) | (g << 3) | (h << 2) | q;
}

// Byte address of (row, logical byte offset) inside one operand stage.
__host__ __device__ constexpr int lds_addr(int row, int byte_off) {
    return row * kRowBytes + ((((byte_off >> 4) ^ swz(row)) & 7) << 4) + (byte_off & 15);
}

__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
    // 0x00020000: raw buffer, 32-bit data format (CDNA buffer resource word 3).
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                             0x00020000);
}

__device__ inline i32x4_t buf_load16(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
}

// LDS barrier: wait for this wave's LDS traffic, then the workgroup barrier.
// The asm memory clobber keeps the compiler from moving LDS accesses across it.
__device__ inline void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

__device__ inline uint32_t pack_bf16x2_rne(float a, float b) {
    __bf16 x = (__bf16)a, y = (__bf16)b;
    return (uint32_t)__builtin_bit_cast(uint16_t, x) | ((uint32_t)__builtin_bit_cast(uint16_t, y) << 16);
}
__device__ inline uint32_t pack_bf16x2_trunc(float a, float b) {
    return (__builtin_bit_cast(uint32_t, a) >> 16) | (__builtin_bit_cast(uint32_t, b) & 0xffff0000u);
}
__device__ inline uint32_t pack_f16x2(float a, float b) {
    _Float16 x = (_Float16)a, y = (_Float16)b;
    return (uint32_t)__builtin_bit_cast(uint16_t, x) | ((uint32_t)__builtin_bit_cast(uint16_t, y) << 16);
}
// Reinterpret an int taken by value.  (A bit_cast applied directly to a
// vector-element lvalue such as v[i] reads element 0 with this compiler.)
__device__ inline float as_float(int v) { return __builtin_bit_cast(float, v); }

__device__ inline float bf16lo(uint32_t v) { return __builtin_bit_cast(float, v << 16); }
__device__ inline float bf16hi(uint32_t v) { return __builtin_bit_cast(float, v & 0xffff0000u); }
__device__ inline float f16lo(uint32_t v) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(v & 0xffffu)); }
__device__ inline float f16hi(uint32_t v) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(v >> 16)); }

// Problem as the kernels see it (one batch-strided GEMM).
struct TileArgs {
    const void* A;
    const void* B;
    const void* C;
    void* D;
    int64_t bsA, bsB, bsC, bsD;  // elements
    int32_t M, N, K;
    int32_t lda, ldb, ldc, ldd;
    float alpha, beta;
    const float* scaleA;  // fp8: q = x * scale; D *= 1/(scaleA*scaleB)
    const float* scaleB;
    int32_t tiles_m, tiles_n, batch, group_m;
    uint32_t flags;  // kFlagBeta | kFlagTrunc
};
constexpr uint32_t kFlagBeta = 1u;   // beta != 0: C is read
constexpr uint32_t kFlagTrunc = 2u;  // bf16 output by truncation
constexpr uint32_t kFlagOneRound = 4u;  // every tile's workgroup is resident at once (grid <= CUs)

}  // namespace ffgemm
