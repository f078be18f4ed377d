// bf16 inputs, fp32 accumulate (v_mfma_f32_32x32x16_bf16); bf16 or fp32 D.
#include "gemm_dispatch.hpp"
namespace ffgemm {
int launch_bf16(int layout, int out, int mt_m, int mt_n, int waves, int mfma, const TileArgs& a, int grid, hipStream_t s) {
    if (out == kOutBF16) return launch_mode<kBF16, kOutBF16>(layout, mt_m, mt_n, waves, mfma, a, grid, s);
    if (out == kOutF32) return launch_mode<kBF16, kOutF32>(layout, mt_m, mt_n, waves, mfma, a, grid, s);
    return -1;
}
}  // namespace ffgemm
