// The general-stride form for fp32 (fp32 MFMA or bf16x3) and fp8 e4m3 inputs.
#include "gemm_dispatch.hpp"
namespace ffgemm {
int launch_any32(int mode, int layout, int out, int mt, const TileArgs& a, int grid, hipStream_t s) {
    if (mode == kF32 && out == kOutF32) return launch_any_mode<kF32, kOutF32>(layout, mt, a, grid, s);
    if (mode == kX3 && out == kOutF32) return launch_any_mode<kX3, kOutF32>(layout, mt, a, grid, s);
    if (mode == kF8 && out == kOutF32) return launch_any_mode<kF8, kOutF32>(layout, mt, a, grid, s);
    if (mode == kF8 && out == kOutBF16) return launch_any_mode<kF8, kOutBF16>(layout, mt, a, grid, s);
    return -1;
}
}  // namespace ffgemm
