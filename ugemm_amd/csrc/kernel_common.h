// Device-side helpers shared by the GEMM kernel families.
#pragma once

#include <hip/hip_runtime.h>

namespace ugemm {

// Maps a launch's linear block id to a (tile_m, tile_n) pair of C tiles.
// 1) XCD de-interleave: hardware hands consecutive block ids round-robin to
//    the 8 XCDs of gfx950; renumber so that each XCD (one L2) owns a
//    contiguous range of tile ids.
// 2) Grouped order: GROUP_M tile-rows are walked column by column so that the
//    blocks resident at one time share A row-panels and B column-panels.
// The mapping is a bijection of [0, tiles_m * tiles_n).
__device__ __forceinline__ void tile_coords(int bid, int tiles_m, int tiles_n, int &tm, int &tn) {
    const int total = tiles_m * tiles_n;
    int pid = bid;
    if ((total & 7) == 0) pid = (bid & 7) * (total >> 3) + (bid >> 3);
    constexpr int GROUP_M = 8;
    const int per_group = GROUP_M * tiles_n;
    const int group = pid / per_group;
    const int first_m = group * GROUP_M;
    const int gsize = min(tiles_m - first_m, GROUP_M);
    const int in_group = pid - group * per_group;
    tm = first_m + in_group % gsize;
    tn = in_group / gsize;
}

}  // namespace ugemm
