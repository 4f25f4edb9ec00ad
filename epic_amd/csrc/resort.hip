// Physical re-sort of a member's parcels into cell order.  The step reads
// parcels through the cell-sort permutation, so its gathers coalesce only as
// long as the physical order follows the cells; as parcels drift, every few
// steps the attribute arrays that outlive a step are gathered into the
// sorted order (dst[s] = src[perm[s]], into scratch arrays) and copied
// back.  Parcel order is not observable: epic_get_parcels restores
// the caller's order through orig, which is permuted with the rest.
#include "common.h"
#include "launch.h"

#include <algorithm>

namespace epic {

// Every array of a member in one pass: a thread gathers its slot's values
// (NW of 8 bytes, RESORT_F of 4) into the scratch arrays, reading the
// permutation once and keeping all the loads in flight; a second launch
// copies them back.
template <int NW>
__global__ __launch_bounds__(256) void k_resort(const ResortArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const ResortArgs &a = tab[ib.item];
#pragma unroll
    for (int r = 0; r < PER_BLOCK / 256; ++r) {
        const int64_t s = (int64_t)ib.blk * PER_BLOCK + r * 256 + threadIdx.x;
        if (s >= a.n) break;
        double w[NW > 0 ? NW : 1];
        float f[RESORT_F];
        if (a.gather) {
            const int64_t q = gl(a.perm)[s];
#pragma unroll
            for (int c = 0; c < NW; ++c) w[c] = gl(a.w[c])[q];
#pragma unroll
            for (int c = 0; c < RESORT_F; ++c) f[c] = gl(a.f[c])[q];
#pragma unroll
            for (int c = 0; c < NW; ++c) gl(a.tw[c])[s] = w[c];
#pragma unroll
            for (int c = 0; c < RESORT_F; ++c) gl(a.tf[c])[s] = f[c];
        } else {
#pragma unroll
            for (int c = 0; c < NW; ++c) w[c] = gl(a.tw[c])[s];
#pragma unroll
            for (int c = 0; c < RESORT_F; ++c) f[c] = gl(a.tf[c])[s];
#pragma unroll
            for (int c = 0; c < NW; ++c) gl(a.w[c])[s] = w[c];
#pragma unroll
            for (int c = 0; c < RESORT_F; ++c) gl(a.f[c])[s] = f[c];
        }
    }
}

// a per-slot array in caller-index order: out[orig[s]] = in[s]
__global__ __launch_bounds__(256) void k_to_caller(const ToCallerArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const ToCallerArgs &a = tab[ib.item];
    const int64_t s = (int64_t)ib.blk * 256 + threadIdx.x;
    if (s >= a.n) return;
    const int o = gl(a.orig)[s];
    if (o >= 0 && o < a.n) gl(a.out)[o] = gl(a.in)[s];
}

void add_resort(Plan &P, int phase, const std::vector<ResortArgs> &items, bool with_doubles)
{
    std::vector<int> nb;
    for (const ResortArgs &a : items) nb.push_back((int)std::max<int64_t>(1, (a.n + PER_BLOCK - 1) / PER_BLOCK));
    if (with_doubles)
        P.add((const void *)k_resort<RESORT_W>, 256, 0, phase, items, nb);
    else
        P.add((const void *)k_resort<0>, 256, 0, phase, items, nb);
}

void add_to_caller(Plan &P, int phase, const std::vector<ToCallerArgs> &items)
{
    std::vector<int> nb;
    for (const ToCallerArgs &a : items) nb.push_back((int)std::max<int64_t>(1, (a.n + 255) / 256));
    P.add((const void *)k_to_caller, 256, 0, phase, items, nb);
}

}  // namespace epic
