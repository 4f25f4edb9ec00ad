// Physical re-sort of a member's parcels into cell order.  The step reads
// parcels through the cell-sort permutation, so its gathers coalesce only as
// long as the physical order follows the cells; as parcels drift, every few
// steps the attribute arrays that outlive a step are gathered into the
// sorted order (dst[s] = src[perm[s]], through one scratch array) and
// copied back.  Parcel order is not observable: epic_get_parcels restores
// the caller's order through orig, which is permuted with the rest.
#include "common.h"
#include "launch.h"

#include <algorithm>

namespace epic {

__global__ __launch_bounds__(256) void k_permute(const PermArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const PermArgs &a = tab[ib.item];
#pragma unroll
    for (int r = 0; r < PER_BLOCK / 256; ++r) {
        const int64_t s = (int64_t)ib.blk * PER_BLOCK + r * 256 + threadIdx.x;
        if (s >= a.n) break;
        const int64_t q = a.perm ? gl(a.perm)[s] : s;   // no permutation: a copy
        if (a.wide) gl((double *)a.dst)[s] = gl((const double *)a.src)[q];
        else gl((float *)a.dst)[s] = gl((const float *)a.src)[q];   // 4-byte words (float or int)
    }
}

void add_permute(Plan &P, int phase, const std::vector<PermArgs> &items)
{
    std::vector<int> nb;
    for (const PermArgs &a : items) nb.push_back((int)std::max<int64_t>(1, (a.n + PER_BLOCK - 1) / PER_BLOCK));
    P.add((const void *)k_permute, 256, 0, phase, items, nb);
}

}  // namespace epic
