// Grouped-GEMM descriptors, tile schedule and launcher (grouped.hip).
#pragma once
#include "gemm_dispatch.hpp"

namespace ffgemm {

struct GroupDesc {
    TileArgs a;  // tiles_m/tiles_n set, batch = 1
};

// One scheduled output tile: the problem (descriptor index) and the tile's
// index inside that problem's row-major tile grid.
struct GroupTile {
    int32_t problem;
    int32_t tile;
};

// Workgroup b of the persistent grid runs sched[first[b]] .. sched[first[b + 1] - 1].
int launch_grouped(int mode, int out, int layout, int mt_m, int mt_n, const GroupDesc* descs, const int32_t* first,
                   const GroupTile* sched, int grid, hipStream_t s);

}  // namespace ffgemm
