// Workgroup -> output tile order shared by the GEMM kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace tns {

// blockIdx -> (tile_m, tile_n). Workgroups are dispatched round-robin over
// the eight XCDs, each with its own L2: the first 8*floor(grid/8) ids are
// regrouped so that every XCD gets one contiguous band of tiles (the tail
// keeps its order), and the band is walked in groups of eight tile-rows so
// that neighbouring workgroups share B panels. A bijection on [0, grid).
__device__ __forceinline__ void tile_of(int bid, int grid, int tiles_m, int tiles_n, int& tm, int& tn)
{
    const int per = grid >> 3;
    int id = bid;
    if (bid < (per << 3)) id = (bid & 7) * per + (bid >> 3);
    constexpr int GROUP = 8;
    const int group_size = GROUP * tiles_n;
    const int group = id / group_size;
    const int first = group * GROUP;
    const int gm = tiles_m - first < GROUP ? tiles_m - first : GROUP;
    const int in = id - group * group_size;
    tm = first + in % gm;
    tn = in / gm;
}

}  // namespace tns
