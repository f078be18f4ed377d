// Grouped GEMM: many independent problems in one persistent launch.
//
// The host sorts problems by K (longest tiles first) and stores one
// descriptor per problem with the prefix sum of its tile count.  Each
// persistent workgroup walks the global tile list with a grid stride,
// binary-searches the owning problem, and runs the same tile engine as the
// single GEMM.  Nothing is allocated or copied at launch time, so the launch
// can be captured into a hipGraph and replayed.
#include "grouped.hpp"

namespace ffgemm {

// 256 x 256 is left out: with the descriptor walk it exceeds the register file.
#define FFGEMM_GROUPED_TILES(X) X(256, 128) X(128, 256) X(128, 128) X(64, 64)

// bf16/fp16 tiles of 128 and up use the pipelined engine (v2::), the rest
// the first engine.
template <int MODE, bool AK, bool BK, int MT_M, int MT_N>
struct GroupedEngine {
    static constexpr bool kV2 = (MODE == kBF16 || MODE == kF16) && MT_M >= 128 && MT_N >= 128;
    static constexpr int OCC = kV2 ? v2::Cfg<MT_M, MT_N, AK, BK>::OCC : 1;
    static constexpr int LDS = kV2 ? v2::smem_bytes<MT_M, MT_N, AK, BK>() : 2 * (MT_M + MT_N) * kRowBytes;
};

template <int MODE, bool AK, bool BK, int MT_M, int MT_N, int OUT>
__global__ void __launch_bounds__(kThreads, (GroupedEngine<MODE, AK, BK, MT_M, MT_N>::OCC))
    grouped_kernel(const GroupDesc* __restrict__ descs, int count, int64_t total_tiles) {
    using E = GroupedEngine<MODE, AK, BK, MT_M, MT_N>;
    __shared__ __attribute__((aligned(16))) char smem[E::LDS];
    for (int64_t t = blockIdx.x; t < total_tiles; t += gridDim.x) {
        int lo = 0, hi = count - 1;
        while (lo < hi) {  // last descriptor with tile_start <= t
            int mid = (lo + hi + 1) >> 1;
            if (descs[mid].tile_start <= t) lo = mid; else hi = mid - 1;
        }
        const GroupDesc& d = descs[lo];
        const int local = (int)(t - d.tile_start);
        const int gm = d.a.group_m > 0 ? d.a.group_m : 1;
        const int gsize = gm * d.a.tiles_n;
        const int grp = local / gsize, first = grp * gm;
        const int h = min(d.a.tiles_m - first, gm);
        const int r = local - grp * gsize;
        if constexpr (E::kV2) v2::gemm_tile<MODE, AK, BK, MT_M, MT_N, OUT>(d.a, first + r % h, r / h, 0, smem);
        else gemm_tile<MODE, AK, BK, MT_M, MT_N, OUT>(d.a, first + r % h, r / h, 0, smem);
    }
}

template <int MODE, bool AK, bool BK, int OUT>
static int grouped_layout(int mt_m, int mt_n, const GroupDesc* descs, int count, int64_t total, int grid,
                          hipStream_t s) {
#define FFGEMM_GCASE(TM_, TN_)                                                                      \
    if (mt_m == TM_ && mt_n == TN_) {                                                               \
        hipLaunchKernelGGL((grouped_kernel<MODE, AK, BK, TM_, TN_, OUT>), dim3(grid), dim3(kThreads), 0, s, \
                           descs, count, total);                                                    \
        return 0;                                                                                   \
    }
    FFGEMM_GROUPED_TILES(FFGEMM_GCASE)
#undef FFGEMM_GCASE
    return -1;
}

template <int MODE, int OUT>
static int grouped_mode(int layout, int mt_m, int mt_n, const GroupDesc* descs, int count, int64_t total,
                        int grid, hipStream_t s) {
    switch (layout) {
        case 0: return grouped_layout<MODE, false, true, OUT>(mt_m, mt_n, descs, count, total, grid, s);
        case 1: return grouped_layout<MODE, false, false, OUT>(mt_m, mt_n, descs, count, total, grid, s);
        case 2: return grouped_layout<MODE, true, true, OUT>(mt_m, mt_n, descs, count, total, grid, s);
        case 3: return grouped_layout<MODE, true, false, OUT>(mt_m, mt_n, descs, count, total, grid, s);
    }
    return -1;
}

int launch_grouped(int mode, int out, int layout, int mt_m, int mt_n, const GroupDesc* descs, int count,
                   int64_t total, int grid, hipStream_t s) {
    if (mode == kBF16 && out == kOutBF16) return grouped_mode<kBF16, kOutBF16>(layout, mt_m, mt_n, descs, count, total, grid, s);
    if (mode == kF16 && out == kOutF16) return grouped_mode<kF16, kOutF16>(layout, mt_m, mt_n, descs, count, total, grid, s);
    return -1;
}

}  // namespace ffgemm
