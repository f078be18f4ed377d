// Kernels under the names and argument list of the reference's Tensile-layout
// kernels (its -t 1 mode, src/gemmMfma.cpp:70-103, named as
// src/gemmMfmaKernelWriter.h:374-387), so a Tensile client can load them from
// a code object as replacement kernels (test/work/feifei_rpk/run_feifei_rpk.py
// swaps such a kernel in for the one Tensile generates under the same name).
//
// Name: Cijk_Alik_Bljk_{SB,HBH,BBH}_MT{MT0}x{MT1}x{DU}_SE_K1 -- TN, batched
// (C[i,j,k], A[l,i,k], B[l,j,k]; l is the summation index).
// Arguments (148 bytes, natural alignment): sizeC, sizeA, sizeB (u64), D, C,
// A, B (pointers), alpha, beta (32-bit slots: fp16 in the low half for HBH,
// fp32 otherwise, as the reference's HBH kernels read them,
// src/gemmMfmaKernelWriter.h:1845-1854), strideD0, strideD1, StrideC0,
// StrideC1, StrideA0, StrideA1, StrideB0, StrideB1 (elements: leading
// dimension, then batch stride), SizesFree0 (M), SizesFree1 (N), SizesFree2
// (batch), SizesSum0 (K), then OrigStaggerUIter, NumWorkGroups0,
// NumWorkGroups1, MagicNumberProblemNumGroupTiles0, GridNumWorkGroups0,
// NumFullBlocks, WgmRemainder1, MagicNumberWgmRemainder1 and a pad, which
// these kernels ignore (they map tiles themselves).
// Launch: grid (ceil(M/MT0), ceil(N/MT1), batch) as Tensile computes it for
// the solution; workgroup size 512 for the 256 x 256 kernels (eight-wave
// engine), 256 for the others.  Workgroups beyond the tile grid return.
//
// Each kernel runs the library's engine for that tile and layout (the same
// code the C ABI launches), with the linear workgroup index remapped so that
// each XCD works on a compact block of tiles.  The reference's interface
// writes bf16 D by truncation; these kernels round to nearest even, as the
// library does by default.
#include "gemm_tile_v4.hpp"

namespace ffgemm {
namespace tensile {

__device__ inline float slot_f32(uint32_t v) { return __builtin_bit_cast(float, v); }
__device__ inline float slot_f16(uint32_t v) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(v & 0xffffu)); }

// Tile-raster group height: the library's default for this grid (each XCD's
// share of tiles as a near-square block), as the host picks it.
__device__ inline int group_height(int tiles_m, int tiles_n) {
    const int per_xcd = max(1, tiles_m * tiles_n / 8);
    int g = 1;
    while ((g * 2) * (g * 2) <= per_xcd && g * 2 <= tiles_m) g *= 2;
    return g;
}

// Fills TileArgs from the Tensile argument list; false when this workgroup has
// no tile.  (tb, tm, tn): batch and tile of this workgroup.
template <int MODE, int MT_M, int MT_N>
__device__ inline bool decode(TileArgs& a, int& tb, int& tm, int& tn, void* D, const void* C, const void* A,
                              const void* B, uint32_t alpha, uint32_t beta, uint32_t sd0, uint32_t sd1, uint32_t sc0,
                              uint32_t sc1, uint32_t sa0, uint32_t sa1, uint32_t sb0, uint32_t sb1, uint32_t m,
                              uint32_t n, uint32_t batch, uint32_t k) {
    a.A = A; a.B = B; a.D = D;
    a.M = (int32_t)m; a.N = (int32_t)n; a.K = (int32_t)k;
    a.lda = (int32_t)sa0; a.ldb = (int32_t)sb0; a.ldd = (int32_t)sd0;
    a.bsA = sa1; a.bsB = sb1; a.bsC = sc1; a.bsD = sd1;
    a.alpha = MODE == kF16 ? slot_f16(alpha) : slot_f32(alpha);
    a.beta = MODE == kF16 ? slot_f16(beta) : slot_f32(beta);
    const bool use_c = a.beta != 0.0f;
    a.C = use_c ? C : nullptr;
    a.ldc = (int32_t)(use_c ? sc0 : sd0);
    a.scaleA = nullptr; a.scaleB = nullptr;
    a.batch = (int32_t)(batch ? batch : 1);
    a.tiles_m = (a.M + MT_M - 1) / MT_M;
    a.tiles_n = (a.N + MT_N - 1) / MT_N;
    a.group_m = group_height(a.tiles_m, a.tiles_n);
    a.flags = use_c ? kFlagBeta : 0u;
    if (a.M <= 0 || a.N <= 0) return false;
    if ((int)blockIdx.x >= a.tiles_m || (int)blockIdx.y >= a.tiles_n || (int)blockIdx.z >= a.batch) return false;
    const int total = a.tiles_m * a.tiles_n * a.batch;
    const int bid = (int)blockIdx.x + a.tiles_m * ((int)blockIdx.y + a.tiles_n * (int)blockIdx.z);
    map_tile(bid, total, a, tb, tm, tn);
    return true;
}

}  // namespace tensile
}  // namespace ffgemm

using namespace ffgemm;

#define FFGEMM_TENSILE_PARAMS                                                                                     \
    uint64_t sizeC, uint64_t sizeA, uint64_t sizeB, void *D, const void *C, const void *A, const void *B,          \
        uint32_t alpha, uint32_t beta, uint32_t strideD0, uint32_t strideD1, uint32_t strideC0, uint32_t strideC1,  \
        uint32_t strideA0, uint32_t strideA1, uint32_t strideB0, uint32_t strideB1, uint32_t sizesFree0,            \
        uint32_t sizesFree1, uint32_t sizesFree2, uint32_t sizesSum0, uint32_t origStaggerUIter,                    \
        uint32_t numWorkGroups0, uint32_t numWorkGroups1, uint32_t magicNumberProblemNumGroupTiles0,                 \
        uint32_t gridNumWorkGroups0, uint32_t numFullBlocks, uint32_t wgmRemainder1,                               \
        uint32_t magicNumberWgmRemainder1, uint32_t padding
#define FFGEMM_TENSILE_DECODE(MODE, MT_M, MT_N)                                                                  \
    TileArgs a;                                                                                                   \
    int tb, tm, tn;                                                                                               \
    if (!tensile::decode<MODE, MT_M, MT_N>(a, tb, tm, tn, D, C, A, B, alpha, beta, strideD0, strideD1, strideC0,   \
                                           strideC1, strideA0, strideA1, strideB0, strideB1, sizesFree0,          \
                                           sizesFree1, sizesFree2, sizesSum0))                                    \
        return;

// 256 x 256, DepthU 64: the eight-wave ping-pong engine (512 threads).
#define FFGEMM_TENSILE_V4(NAME, MODE, OUT)                                                                        \
    extern "C" __global__ void __launch_bounds__(v4::kThreads8, 1) NAME(FFGEMM_TENSILE_PARAMS) {                   \
        __shared__ __attribute__((aligned(16))) char smem[v4::LDS_BYTES];                                         \
        FFGEMM_TENSILE_DECODE(MODE, 256, 256)                                                                     \
        v4::gemm_tile<MODE, true, true, OUT>(a, tm, tn, tb, smem, threadIdx.x);                                   \
    }
// 128 x 128 and 64 x 64 16-bit, DepthU 32 (64-byte K rows, as the names
// say): the pipelined four-wave engine.
#define FFGEMM_TENSILE_V2(NAME, MODE, OUT, MT)                                                                    \
    extern "C" __global__ void __launch_bounds__(kThreads, (v2::Cfg<MT, MT, true, true, 64>::OCC)) NAME(          \
        FFGEMM_TENSILE_PARAMS) {                                                                                  \
        __shared__ __attribute__((aligned(16))) char smem[v2::smem_bytes<MT, MT, true, true, 64>()];              \
        FFGEMM_TENSILE_DECODE(MODE, MT, MT)                                                                       \
        v2::gemm_tile<MODE, true, true, MT, MT, OUT, 64>(a, tm, tn, tb, smem, threadIdx.x);                       \
    }
// fp32, DepthU 32: the register-staged engine.
#define FFGEMM_TENSILE_V1(NAME, MT)                                                                               \
    extern "C" __global__ void __launch_bounds__(kThreads, 1) NAME(FFGEMM_TENSILE_PARAMS) {                       \
        __shared__ __attribute__((aligned(16))) char smem[2 * (MT + MT) * kRowBytes];                            \
        FFGEMM_TENSILE_DECODE(kF32, MT, MT)                                                                       \
        gemm_tile<kF32, true, true, MT, MT, kOutF32>(a, tm, tn, tb, smem, threadIdx.x);                           \
    }

FFGEMM_TENSILE_V4(Cijk_Alik_Bljk_BBH_MT256x256x64_SE_K1, kBF16, kOutBF16)
FFGEMM_TENSILE_V4(Cijk_Alik_Bljk_HBH_MT256x256x64_SE_K1, kF16, kOutF16)
FFGEMM_TENSILE_V2(Cijk_Alik_Bljk_BBH_MT128x128x32_SE_K1, kBF16, kOutBF16, 128)
FFGEMM_TENSILE_V2(Cijk_Alik_Bljk_HBH_MT128x128x32_SE_K1, kF16, kOutF16, 128)
FFGEMM_TENSILE_V2(Cijk_Alik_Bljk_BBH_MT64x64x32_SE_K1, kBF16, kOutBF16, 64)
FFGEMM_TENSILE_V2(Cijk_Alik_Bljk_HBH_MT64x64x32_SE_K1, kF16, kOutF16, 64)
FFGEMM_TENSILE_V1(Cijk_Alik_Bljk_SB_MT128x128x32_SE_K1, 128)
FFGEMM_TENSILE_V1(Cijk_Alik_Bljk_SB_MT64x64x32_SE_K1, 64)
