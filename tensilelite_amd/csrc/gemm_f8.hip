// OCP e4m3 inputs (K-contiguous A and B, i.e. TN), per-tensor scales folded
// into alpha, fp32 accumulate on v_mfma_scale_f32_32x32x64_f8f6f4; fp32 or bf16 D.
#include "gemm_dispatch.hpp"
namespace ffgemm {
int launch_f8(int layout, int out, int mt_m, int mt_n, int waves, int mfma, const TileArgs& a, int grid, hipStream_t s) {
    if (layout != 2) return -1;
    if (out == kOutF32) return launch_layout<kF8, true, true, kOutF32>(mt_m, mt_n, waves, mfma, a, grid, s);
    if (out == kOutBF16) return launch_layout<kF8, true, true, kOutBF16>(mt_m, mt_n, waves, mfma, a, grid, s);
    return -1;
}
}  // namespace ffgemm
