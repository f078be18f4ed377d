// fp32 inputs split into hi/lo bf16 pairs; hi*hi + hi*lo + lo*hi on the bf16
// MFMA with fp32 accumulate (bf16x3); fp32 D.
#include "gemm_dispatch.hpp"
namespace ffgemm {
int launch_x3(int layout, int out, int mt_m, int mt_n, int waves, int mfma, const TileArgs& a, int grid, hipStream_t s) {
    if (out == kOutF32) return launch_mode<kX3, kOutF32>(layout, mt_m, mt_n, waves, mfma, a, grid, s);
    return -1;
}
}  // namespace ffgemm
