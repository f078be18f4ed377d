// Grouped GEMM: many independent problems in one persistent launch.
//
// The plan (ffgemm_grouped_plan_create) stores one descriptor per problem and
// a tile schedule: every output tile of every problem, dealt to the
// workgroups of a persistent grid longest-first (K descending), each tile to
// the workgroup with the least work so far.  Workgroup b walks its own
// contiguous slice of the schedule and runs each tile with the same tile
// engine as the single GEMM; the split-ring engine runs the whole slice as one
// stream of K-tiles, so a tile's first loads overlap the previous tile's last
// K-tiles.  Nothing is allocated or copied at launch time, so the launch can
// be captured into a hipGraph and replayed.
#include "grouped.hpp"

namespace ffgemm {

#define FFGEMM_GROUPED_TILES(X) X(256, 256) X(256, 128) X(128, 256) X(128, 128) X(64, 64)

// Each tile runs the engine the single GEMM takes by default for it
// (engine_for): at 256 x 256 the eight-wave engine (512 threads) for TN / NT,
// the split-ring engine for NN and the pipelined engine for TT; the pipelined
// engine for the other 16-bit tiles of 128 and up; the first engine for 64 x 64.
template <int MODE, bool AK, bool BK, int MT_M, int MT_N>
struct GroupedEngine {
    static constexpr Engine E = engine_for(MODE, AK, BK, MT_M, MT_N, 0);
    static constexpr int THREADS = (E == kEngV4) ? v4::kThreads8 : kThreads;
    static constexpr int OCC = (E == kEngV2) ? v2::Cfg<MT_M, MT_N, AK, BK>::OCC : 1;
    static constexpr int LDS = (E == kEngV4)   ? v4::LDS_BYTES
                               : (E == kEngV3) ? v3::LDS_BYTES
                               : (E == kEngV2) ? v2::smem_bytes<MT_M, MT_N, AK, BK>()
                                               : 2 * (MT_M + MT_N) * kRowBytes;
};

template <int MODE, bool AK, bool BK, int MT_M, int MT_N, int OUT>
__global__ void __launch_bounds__((GroupedEngine<MODE, AK, BK, MT_M, MT_N>::THREADS),
                                  (GroupedEngine<MODE, AK, BK, MT_M, MT_N>::OCC))
    grouped_kernel(const GroupDesc* __restrict__ descs, const int32_t* __restrict__ first,
                   const GroupTile* __restrict__ sched) {
    using E = GroupedEngine<MODE, AK, BK, MT_M, MT_N>;
    __shared__ __attribute__((aligned(16))) char smem[E::LDS];
    const int end = first[blockIdx.x + 1];
    if constexpr (E::E == kEngV3) {
        // the split-ring engine runs the slice as one stream of K-tiles
        auto tile_at = [&](int i, const TileArgs*& a, int& tm, int& tn) {
            const GroupTile e = sched[i];
            a = &descs[e.problem].a;
            const int gm = a->group_m > 0 ? a->group_m : 1;
            const int gsize = gm * a->tiles_n;
            const int grp = e.tile / gsize, top = grp * gm;
            const int h = min(a->tiles_m - top, gm);
            const int r = e.tile - grp * gsize;
            tm = top + r % h;
            tn = r / h;
        };
        v3::gemm_tiles_stream<MODE, AK, BK, OUT>(tile_at, first[blockIdx.x], end, smem, threadIdx.x);
        return;
    }
    for (int i = first[blockIdx.x]; i < end; ++i) {
        // The thread index goes through an empty asm per tile, so the engine
        // derives its per-thread offsets inside the loop instead of having
        // them hoisted above it and held in registers across every tile.
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const GroupTile e = sched[i];
        const GroupDesc& d = descs[e.problem];
        const int local = e.tile;
        const int gm = d.a.group_m > 0 ? d.a.group_m : 1;
        const int gsize = gm * d.a.tiles_n;
        const int grp = local / gsize, first = grp * gm;
        const int h = min(d.a.tiles_m - first, gm);
        const int r = local - grp * gsize;
        const int tm = first + r % h, tn = r / h;
        if constexpr (E::E == kEngV4) v4::gemm_tile<MODE, AK, BK, OUT>(d.a, tm, tn, 0, smem, tid);
        else if constexpr (E::E == kEngV2) v2::gemm_tile<MODE, AK, BK, MT_M, MT_N, OUT>(d.a, tm, tn, 0, smem, tid);
        else gemm_tile<MODE, AK, BK, MT_M, MT_N, OUT>(d.a, tm, tn, 0, smem, tid);
        // the eight-wave engine leaves every wave past its last LDS read and
        // with no LDS-bound load in flight; one workgroup barrier keeps the
        // next tile's prologue behind the slowest wave
        if constexpr (E::E == kEngV4) __syncthreads();
    }
}

template <int MODE, bool AK, bool BK, int OUT>
static int grouped_layout(int mt_m, int mt_n, const GroupDesc* descs, const int32_t* first,
                          const GroupTile* sched, int grid, hipStream_t s) {
#define FFGEMM_GCASE(TM_, TN_)                                                                      \
    if (mt_m == TM_ && mt_n == TN_) {                                                               \
        hipLaunchKernelGGL((grouped_kernel<MODE, AK, BK, TM_, TN_, OUT>), dim3(grid),                \
                           dim3(GroupedEngine<MODE, AK, BK, TM_, TN_>::THREADS), 0, s,               \
                           descs, first, sched);                                                    \
        return 0;                                                                                   \
    }
    FFGEMM_GROUPED_TILES(FFGEMM_GCASE)
#undef FFGEMM_GCASE
    return -1;
}

template <int MODE, int OUT>
static int grouped_mode(int layout, int mt_m, int mt_n, const GroupDesc* descs, const int32_t* first,
                        const GroupTile* sched, int grid, hipStream_t s) {
    switch (layout) {
        case 0: return grouped_layout<MODE, false, true, OUT>(mt_m, mt_n, descs, first, sched, grid, s);
        case 1: return grouped_layout<MODE, false, false, OUT>(mt_m, mt_n, descs, first, sched, grid, s);
        case 2: return grouped_layout<MODE, true, true, OUT>(mt_m, mt_n, descs, first, sched, grid, s);
        case 3: return grouped_layout<MODE, true, false, OUT>(mt_m, mt_n, descs, first, sched, grid, s);
    }
    return -1;
}

int launch_grouped(int mode, int out, int layout, int mt_m, int mt_n, const GroupDesc* descs, const int32_t* first,
                   const GroupTile* sched, int grid, hipStream_t s) {
    if (mode == kBF16 && out == kOutBF16)
        return grouped_mode<kBF16, kOutBF16>(layout, mt_m, mt_n, descs, first, sched, grid, s);
    if (mode == kF16 && out == kOutF16)
        return grouped_mode<kF16, kOutF16>(layout, mt_m, mt_n, descs, first, sched, grid, s);
    return -1;
}

}  // namespace ffgemm
