// Kernel instantiation tables.  Each gemm_<mode>.hip includes this header and
// instantiates the (layout x tile) table for its mode, so the heavy template
// compiles in parallel translation units.
#pragma once
#include "gemm_tile_v3.hpp"
#include "gemm_tile_v4.hpp"

namespace ffgemm {

// Supported macro tiles (MT_M x MT_N), 4 waves each.
#define FFGEMM_TILES(X) X(256, 256) X(256, 128) X(128, 256) X(128, 128) X(64, 64)

template <int MODE, bool AK, bool BK, int OUT>
inline int launch_layout(int mt_m, int mt_n, const TileArgs& a, int grid, hipStream_t s) {
    // 256 x 256, by what measured fastest on MI355X at 4096^3: the eight-wave
    // ping-pong engine (v4::) for TN and NT (M/N-contiguous operands through
    // transposing LDS reads) and fp8; the split-ring engine (v3::) for NN; the
    // pipelined 4-wave engine (v2::) for TT and for the other tiles of 128 and
    // up; the first engine for the rest (and fp32 / bf16x3 / fp8 below 256).
    constexpr bool k16 = (MODE == kBF16 || MODE == kF16);
#define FFGEMM_CASE(TM_, TN_)                                                                  \
    if (mt_m == TM_ && mt_n == TN_) {                                                          \
        if constexpr (TM_ == 256 && TN_ == 256 && ((k16 && AK == BK) || (MODE == kF8 && AK && BK))) \
            hipLaunchKernelGGL((v4::gemm_kernel<MODE, AK, BK, OUT>), dim3(grid), dim3(v4::kThreads8), 0, s, a); \
        else if constexpr (k16 && TM_ == 256 && TN_ == 256 && !AK && BK)                       \
            hipLaunchKernelGGL((v3::gemm_kernel<MODE, AK, BK, OUT>), dim3(grid), dim3(kThreads), 0, s, a); \
        else if constexpr (k16 && TM_ >= 128 && TN_ >= 128)                                    \
            hipLaunchKernelGGL((v2::gemm_kernel<MODE, AK, BK, TM_, TN_, OUT>), dim3(grid), dim3(kThreads), 0, s, a); \
        else                                                                                   \
            hipLaunchKernelGGL((gemm_kernel<MODE, AK, BK, TM_, TN_, OUT>), dim3(grid), dim3(kThreads), 0, s, a); \
        return 0;                                                                              \
    }
    FFGEMM_TILES(FFGEMM_CASE)
#undef FFGEMM_CASE
    return -1;
}

// layout: 0 NN, 1 NT, 2 TN, 3 TT.  A is K-contiguous for TN/TT, B for NN/TN.
template <int MODE, int OUT>
inline int launch_mode(int layout, int mt_m, int mt_n, const TileArgs& a, int grid, hipStream_t s) {
    switch (layout) {
        case 0: return launch_layout<MODE, false, true, OUT>(mt_m, mt_n, a, grid, s);
        case 1: return launch_layout<MODE, false, false, OUT>(mt_m, mt_n, a, grid, s);
        case 2: return launch_layout<MODE, true, true, OUT>(mt_m, mt_n, a, grid, s);
        case 3: return launch_layout<MODE, true, false, OUT>(mt_m, mt_n, a, grid, s);
    }
    return -1;
}

// Per-mode entry points (defined in gemm_<mode>.hip).
int launch_bf16(int layout, int out, int mt_m, int mt_n, const TileArgs& a, int grid, hipStream_t s);
int launch_f16(int layout, int out, int mt_m, int mt_n, const TileArgs& a, int grid, hipStream_t s);
int launch_f32(int layout, int out, int mt_m, int mt_n, const TileArgs& a, int grid, hipStream_t s);
int launch_x3(int layout, int out, int mt_m, int mt_n, const TileArgs& a, int grid, hipStream_t s);
int launch_f8(int layout, int out, int mt_m, int mt_n, const TileArgs& a, int grid, hipStream_t s);

}  // namespace ffgemm
