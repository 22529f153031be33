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

// Tile index of block b of nb such that each of the 8 XCDs, which receive
// the blocks of a launch round-robin (b mod 8), works through one contiguous
// range of tiles: neighbouring particles, cells or voxels then share that
// XCD's L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb)
{
    const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u;
    return x * q + (x < r ? x : r) + (b >> 3);
}

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
