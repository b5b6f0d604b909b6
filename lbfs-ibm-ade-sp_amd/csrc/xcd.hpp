// Workgroup id remap for the 8 XCDs of the chip: the hardware hands
// workgroups to the XCDs round-robin, so workgroup b runs on XCD b % 8; the
// remap gives each XCD one contiguous range of logical ids, so neighbouring
// tiles (or nodes) share that XCD's L2.
#pragma once

#include <hip/hip_runtime.h>

namespace lbfs {

__device__ __forceinline__ int xcd_remap(int bid, int nblocks)
{
    const int per = (nblocks + 7) >> 3;
    const int xcd = bid & 7, idx = bid >> 3;
    const int full = nblocks - (per - 1) * 8;  // XCDs that get `per` blocks
    int base = xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1);
    return base + idx;
}

}  // namespace lbfs
