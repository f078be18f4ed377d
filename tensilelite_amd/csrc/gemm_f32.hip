// fp32 inputs on the fp32 MFMA (v_mfma_f32_32x32x2_f32); fp32 D.
#include "gemm_dispatch.hpp"
namespace ffgemm {
int launch_f32(int layout, int out, int mt_m, int mt_n, int waves, int mfma, const TileArgs& a, int grid, hipStream_t s) {
    if (out == kOutF32) return launch_mode<kF32, kOutF32>(layout, mt_m, mt_n, waves, mfma, a, grid, s);
    return -1;
}
}  // namespace ffgemm
