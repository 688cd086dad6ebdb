// pps_xcd.h -- workgroup numbering for the 8 XCDs of an MI355X.
#pragma once

#include <hip/hip_runtime.h>

namespace pps {

// Workgroups go to the 8 XCDs round-robin by id (each XCD has its own L2).
// Renumbers ids so that each XCD's workgroups are consecutive: neighbouring
// tiles, which read each other's edge rows, then run on one XCD and find
// those rows in its L2 (k_dq_tiled fetches 680 MB instead of 877 MB per
// 256^3 launch).
__device__ __forceinline__ int xcd_order(int b, int G) {
    const int q = G >> 3, r = G & 7, x = b & 7, s = b >> 3;
    return x < r ? x * (q + 1) + s : r * (q + 1) + (x - r) * q + s;
}

}  // namespace pps
