// Kernel instantiation tables.  Each gemm_<mode>.hip includes this header and
// instantiates the (layout x tile) table for its mode, so the heavy template
// compiles in parallel translation units.
#pragma once
#include "gemm_tile_v3.hpp"
#include "gemm_tile_v4.hpp"

namespace ffgemm {

// Supported macro tiles (MT_M x MT_N).
#define FFGEMM_TILES(X) X(256, 256) X(256, 128) X(128, 256) X(128, 128) X(64, 64)

// Which engine runs a (mode, operand contiguity, tile, waves).  waves = 0 takes
// the default, what measured fastest on MI355X at 4096^3: the eight-wave
// ping-pong engine (v4::) for 256 x 256 TN and NT (M/N-contiguous operands
// through transposing LDS reads) and fp8 TN; the split-ring engine (v3::) for
// 256 x 256 NN; the pipelined 4-wave engine (v2::) for 256 x 256 TT, the
// other 16-bit tiles of 128 and up, and 16-bit 64 x 64 TN (its four-stage
// direct-load ring shortens the latency chain of small problems); the first
// engine (gemm_tile.hpp) for the rest and for fp32 / bf16x3.  waves = 8 or 4 asks for the eight-wave or a
// four-wave engine (the wave-tiling knob of the tile record, which the
// autotuner sweeps); kEngNone when the combination has no kernel.  mfma = 16
// asks for 16 x 16 MFMAs (the MFMA-instruction knob): the first engine's
// 16 x 16 form (gemm_tile_m16) at 64 x 64 and 128 x 128 for bf16 / fp16
// (16 x 16 x 32) and fp32 (16 x 16 x 4).
enum Engine : int { kEngNone = 0, kEngV1 = 1, kEngV2 = 2, kEngV3 = 3, kEngV4 = 4, kEngV1M16 = 5 };
__host__ __device__ constexpr Engine engine_for(int mode, bool ak, bool bk, int tm, int tn, int waves = 0,
                                                int mfma = 0) {
    const bool k16 = (mode == kBF16 || mode == kF16);
    if (mfma == 16)
        return ((k16 || mode == kF32) && (waves == 0 || waves == 4) && tm == tn && (tm == 64 || tm == 128))
                   ? kEngV1M16 : kEngNone;
    if (mfma != 0 && mfma != 32) return kEngNone;
    const bool big = (tm == 256 && tn == 256);
    const bool v4ok = big && (k16 || (mode == kF8 && ak && bk));
    const bool v2ok = k16 && ((tm >= 128 && tn >= 128) || (tm == 64 && tn == 64 && ak && bk));
    const Engine four = (big && k16 && !ak && bk) ? kEngV3 : (v2ok ? kEngV2 : kEngV1);
    if (waves == 8) return v4ok ? kEngV4 : kEngNone;
    if (waves == 4) return four;
    if (waves != 0) return kEngNone;
    if (big && ((k16 && ak == bk) || (mode == kF8 && ak && bk))) return kEngV4;
    return four;
}

// Geometry of an engine, as ffgemm_query_tile reports it: waves along M and
// N, K per main-loop stage, LDS stages of the deeper operand ring.
struct EngineGeometry {
    int waves_m, waves_n, depth_u, lds_buffers;
};
inline EngineGeometry engine_geometry(int mode, bool ak, bool bk, int tm, int tn, int waves = 0, int mfma = 0) {
    const int es = mode_elem_bytes(mode);
    switch (engine_for(mode, ak, bk, tm, tn, waves, mfma)) {
        case kEngV4: return {2, 4, 128 / es, v4::NSB};
        case kEngV3: return {2, 2, v3::DU, v3::NSB};
        case kEngV2: {
            int ns = 0, du = 0;
#define FFGEMM_V2G(TM_, TN_, AK_, BK_) \
    if (tm == TM_ && tn == TN_ && ak == AK_ && bk == BK_) { ns = v2::Cfg<TM_, TN_, AK_, BK_>::NS; du = v2::Cfg<TM_, TN_, AK_, BK_>::DU; }
#define FFGEMM_V2L(TM_, TN_) FFGEMM_V2G(TM_, TN_, false, false) FFGEMM_V2G(TM_, TN_, false, true) \
    FFGEMM_V2G(TM_, TN_, true, false) FFGEMM_V2G(TM_, TN_, true, true)
            FFGEMM_V2L(256, 256) FFGEMM_V2L(256, 128) FFGEMM_V2L(128, 256) FFGEMM_V2L(128, 128) FFGEMM_V2L(64, 64)
#undef FFGEMM_V2L
#undef FFGEMM_V2G
            return {2, 2, du, ns};
        }
        case kEngV1:
        case kEngV1M16: return {2, 2, mode_depth_u(mode), 2};
        default: return {0, 0, 0, 0};
    }
}

template <Engine E, int MODE, bool AK, bool BK, int TM, int TN, int OUT>
inline void launch_engine(const TileArgs& a, int grid, hipStream_t s) {
    if constexpr (E == kEngV4)
        hipLaunchKernelGGL((v4::gemm_kernel<MODE, AK, BK, OUT>), dim3(grid), dim3(v4::kThreads8), 0, s, a);
    else if constexpr (E == kEngV3) {
        if (a.flags & kFlagBeta)
            hipLaunchKernelGGL((v3::gemm_kernel<MODE, AK, BK, OUT, true>), dim3(grid), dim3(kThreads), 0, s, a);
        else
            hipLaunchKernelGGL((v3::gemm_kernel<MODE, AK, BK, OUT, false>), dim3(grid), dim3(kThreads), 0, s, a);
    }
    else if constexpr (E == kEngV2)
        hipLaunchKernelGGL((v2::gemm_kernel<MODE, AK, BK, TM, TN, OUT>), dim3(grid), dim3(kThreads), 0, s, a);
    else if constexpr (E == kEngV1)
        hipLaunchKernelGGL((gemm_kernel<MODE, AK, BK, TM, TN, OUT>), dim3(grid), dim3(kThreads), 0, s, a);
    else if constexpr (E == kEngV1M16)
        hipLaunchKernelGGL((gemm_kernel_m16<MODE, AK, BK, TM, TN, OUT>), dim3(grid), dim3(kThreads), 0, s, a);
}

// Launches the engine engine_for picks; only the engines reachable for this
// (mode, layout, tile) are instantiated.
template <int MODE, bool AK, bool BK, int OUT>
inline int launch_layout(int mt_m, int mt_n, int waves, int mfma, const TileArgs& a, int grid, hipStream_t s) {
#define FFGEMM_CASE(TM_, TN_)                                                                  \
    if (mt_m == TM_ && mt_n == TN_) {                                                          \
        constexpr Engine E16 = engine_for(MODE, AK, BK, TM_, TN_, 0, 16);                      \
        if (mfma == 16) {                                                                      \
            if constexpr (E16 != kEngNone) {                                                   \
                launch_engine<E16, MODE, AK, BK, TM_, TN_, OUT>(a, grid, s);                   \
                return 0;                                                                      \
            }                                                                                  \
            return -1;                                                                         \
        }                                                                                      \
        constexpr Engine E0 = engine_for(MODE, AK, BK, TM_, TN_, 0);                           \
        constexpr Engine E4 = engine_for(MODE, AK, BK, TM_, TN_, 4);                           \
        constexpr Engine E8 = engine_for(MODE, AK, BK, TM_, TN_, 8);                           \
        const Engine e = engine_for(MODE, AK, BK, TM_, TN_, waves);                            \
        if (e == E0) { launch_engine<E0, MODE, AK, BK, TM_, TN_, OUT>(a, grid, s); return 0; } \
        if (e == E4) { launch_engine<E4, MODE, AK, BK, TM_, TN_, OUT>(a, grid, s); return 0; } \
        if (e == E8 && E8 != kEngNone) { launch_engine<E8, MODE, AK, BK, TM_, TN_, OUT>(a, grid, s); return 0; } \
        return -1;                                                                             \
    }
    FFGEMM_TILES(FFGEMM_CASE)
#undef FFGEMM_CASE
    return -1;
}

// layout: 0 NN, 1 NT, 2 TN, 3 TT.  A is K-contiguous for TN/TT, B for NN/TN.
template <int MODE, int OUT>
inline int launch_mode(int layout, int mt_m, int mt_n, int waves, int mfma, const TileArgs& a, int grid,
                       hipStream_t s) {
    switch (layout) {
        case 0: return launch_layout<MODE, false, true, OUT>(mt_m, mt_n, waves, mfma, a, grid, s);
        case 1: return launch_layout<MODE, false, false, OUT>(mt_m, mt_n, waves, mfma, a, grid, s);
        case 2: return launch_layout<MODE, true, true, OUT>(mt_m, mt_n, waves, mfma, a, grid, s);
        case 3: return launch_layout<MODE, true, false, OUT>(mt_m, mt_n, waves, mfma, a, grid, s);
    }
    return -1;
}

// The general-stride form (gemm_kernel_any: element-wise loads and stores) at
// MT x MT, MT = 64 or 128 (fp8: 128).  fp8 operands are K-contiguous (TN) only.
template <int MODE, int OUT>
inline int launch_any_mode(int layout, int mt, const TileArgs& a, int grid, hipStream_t s) {
    auto go = [&](auto ak, auto bk) -> int {
        constexpr bool AK = decltype(ak)::value, BK = decltype(bk)::value;
        if constexpr (MODE == kF8 && !(AK && BK)) {
            return -1;
        } else {
            if (mt == 64 && MODE != kF8)  // fp8 at 64 x 64 spills: 128 x 128 only
                hipLaunchKernelGGL((gemm_kernel_any<MODE, AK, BK, (MODE == kF8 ? 128 : 64), OUT>), dim3(grid),
                                   dim3(kThreads), 0, s, a);
            else if (mt == 128)
                hipLaunchKernelGGL((gemm_kernel_any<MODE, AK, BK, 128, OUT>), dim3(grid), dim3(kThreads), 0, s, a);
            else
                return -1;
            return 0;
        }
    };
    using F = std::false_type;
    using T = std::true_type;
    switch (layout) {
        case 0: return go(F{}, T{});
        case 1: return go(F{}, F{});
        case 2: return go(T{}, T{});
        case 3: return go(T{}, F{});
    }
    return -1;
}
// bf16 / fp16 (gemm_any16.hip); fp32, bf16x3 and fp8 (gemm_any32.hip)
int launch_any16(int mode, int layout, int out, int mt, const TileArgs& a, int grid, hipStream_t s);
int launch_any32(int mode, int layout, int out, int mt, const TileArgs& a, int grid, hipStream_t s);

// Per-mode entry points (defined in gemm_<mode>.hip).
// waves: 0 (default engine), 4 or 8; mfma: 0 / 32, or 16 (see engine_for)
int launch_bf16(int layout, int out, int mt_m, int mt_n, int waves, int mfma, const TileArgs& a, int grid, hipStream_t s);
int launch_f16(int layout, int out, int mt_m, int mt_n, int waves, int mfma, const TileArgs& a, int grid, hipStream_t s);
int launch_f32(int layout, int out, int mt_m, int mt_n, int waves, int mfma, const TileArgs& a, int grid, hipStream_t s);
int launch_x3(int layout, int out, int mt_m, int mt_n, int waves, int mfma, const TileArgs& a, int grid, hipStream_t s);
int launch_f8(int layout, int out, int mt_m, int mt_n, int waves, int mfma, const TileArgs& a, int grid, hipStream_t s);

}  // namespace ffgemm
