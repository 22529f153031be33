// multi.h -- block-to-box map of the multi-box launches (ensembles of
// independent boxes): one launch runs one stage for every box, box k owning
// the blocks [start[k], start[k] + count[k]).  The starts are multiples of
// 8, so that a box's blocks meet the eight XCDs in the order of a launch of
// its own; the blocks between one box's last block and the next start
// return at once.  With ids, entry k of the map is box ids[k] (a launch over
// some of the boxes); without, entry k is box k.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct MultiMap {
    const uint32_t *start;  // nbox block starts (ascending), then nbox counts
    uint32_t nbox;
    const uint32_t *ids = nullptr;  // nbox box indices, or none
};

// The box of this block, the block's index within the box and the box's
// block count; false for the padding blocks after a box's last block.
__device__ __forceinline__ bool multi_block(const MultiMap &m, uint32_t &box, uint32_t &bid,
                                            uint32_t &nblk)
{
    const uint32_t b = blockIdx.x;
    uint32_t lo = 0, hi = m.nbox;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (m.start[mid] <= b) lo = mid;
        else hi = mid;
    }
    box = m.ids ? m.ids[lo] : lo;
    bid = b - m.start[lo];
    nblk = m.start[m.nbox + lo];
    return bid < nblk;
}
