// sym_items.h -- how a wave of the batched symmetric kernels (sym.hip,
// sym_vjp.hip) takes its next work item (device code only).
#pragma once

#include <hip/hip_runtime.h>

#include "internal.h"

namespace utcnb {

// The next item of wave gw, or -1.  Static schedule (g.counter null): the
// items [start[gw], start[gw+1]) in turn.  Dynamic schedule: lane 0 takes the
// next index of the longest-first list with one atomic add; every wave stops
// at g.nitems, so all waves end once the list is used up.
__device__ __forceinline__ int batch_next_item(const BatchSymGeom& g, int gw, int& pos, int lane)
{
    if (!g.counter) {
        const int it = g.start[gw] + pos++;
        return it < g.start[gw + 1] ? it : -1;
    }
    int it = 0;
    if (lane == 0)
        it = atomicAdd(g.counter, 1);
    it = __shfl(it, 0, 64);
    return it < g.nitems ? it : -1;
}

}  // namespace utcnb
