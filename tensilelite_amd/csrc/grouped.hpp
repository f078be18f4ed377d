// Grouped-GEMM descriptors and launcher (grouped.hip).
#pragma once
#include "gemm_dispatch.hpp"

namespace ffgemm {

struct GroupDesc {
    TileArgs a;          // tiles_m/tiles_n set, batch = 1
    int64_t tile_start;  // first global tile index of this problem
};

int launch_grouped(int mode, int out, int layout, int mt_m, int mt_n, const GroupDesc* descs, int count,
                   int64_t total, int grid, hipStream_t s);

}  // namespace ffgemm
