// The general-stride form for 16-bit inputs (bf16, fp16), bf16 / fp16 / fp32 D.
#include "gemm_dispatch.hpp"
namespace ffgemm {
int launch_any16(int mode, int layout, int out, int mt, const TileArgs& a, int grid, hipStream_t s) {
    if (mode == kBF16 && out == kOutBF16) return launch_any_mode<kBF16, kOutBF16>(layout, mt, a, grid, s);
    if (mode == kBF16 && out == kOutF32) return launch_any_mode<kBF16, kOutF32>(layout, mt, a, grid, s);
    if (mode == kF16 && out == kOutF16) return launch_any_mode<kF16, kOutF16>(layout, mt, a, grid, s);
    if (mode == kF16 && out == kOutF32) return launch_any_mode<kF16, kOutF32>(layout, mt, a, grid, s);
    return -1;
}
}  // namespace ffgemm
