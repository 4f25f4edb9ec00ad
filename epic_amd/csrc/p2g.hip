// par2grid on gfx950: deterministic, atomic-free float accumulation.
//
// Pipeline (one member):
//   1. k_cell_keys     FP64 cell of every parcel centre (get_index,
//                      fields.f90:165-172), integer histogram of parcels and
//                      small parcels per cell (nparg / nsparg)
//   2. scan            exclusive prefix sum of the histogram -> cell_start
//   3. k_place         slot = cell_start + integer ticket (order not fixed yet)
//   4. k_cell_order    order every cell's slots by caller parcel index
//                      (epic_cell_sort only: the reference's cell order)
//   5. k_p2g_tiles     one work-group per (8x8 column tile, z chunk of tz
//                      layers), one thread per cell: support points in FP64,
//                      each of the 4 x 8 x 7 contributions of a parcel scaled
//                      by a per-member power of two, rounded and added to a
//                      64-bit integer LDS accumulator of its node; the box is
//                      scaled back exactly and written to a per-tile FP64
//                      staging box
//   6. k_p2g_finalize  per node, sums the staging boxes that cover it in a
//                      fixed tile order (the periodic halo accumulation of
//                      field_mpi.f90), applies the free-slip folds,
//                      normalises by volume, the zeta wall condition and the
//                      z extrapolation (parcel_interpl.f90:235-319) in FP64;
//                      each field is rounded to FP32 once, on its store
// Integer node sums are exact and the staging boxes are added in a fixed
// tile order, so the result depends on the parcel set alone, never on
// scheduling or parcel order: results are bitwise reproducible.
#include "common.h"
#include "ellipsoid.h"
#include "launch.h"

#include <algorithm>

namespace epic {

// ---------------------------------------------------------------------------
// 1. cell keys + histogram
// ---------------------------------------------------------------------------
__device__ inline int cell_key(const Grid &g, double x, double y, double z, int &err)
{
    const int i = (int)floor((x - g.lower[0]) * g.dxi[0]);
    const int j = (int)floor((y - g.lower[1]) * g.dxi[1]);
    const int k = (int)floor((z - g.lower[2]) * g.dxi[2]);
    if (k < -1 || k > g.nz) {
        err |= ERR_PARCEL_COUNT;
        return -1;
    }
    return ((k + 1) * g.ny + wrap(j, g.ny)) * g.nx + wrap(i, g.nx);
}

__global__ __launch_bounds__(256) void k_cell_keys(const KeysArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const KeysArgs &a = tab[ib.item];
    const Grid &g = a.g;
    const int64_t n = parcel_count(a.p);
    if ((int64_t)ib.blk * PER_BLOCK >= n) return;   // uniform over the block
    int e = 0;
    // every lane runs every round (the histogram adds are wave-aggregated
    // over runs of equal keys: consecutive parcels mostly share a cell)
#pragma unroll
    for (int r = 0; r < PER_BLOCK / 256; ++r) {
        const int64_t q = (int64_t)ib.blk * PER_BLOCK + r * 256 + threadIdx.x;
        int key = -1;
        bool small = false;
        if (q < n) {
            key = cell_key(g, gl(a.p.x)[q], gl(a.p.y)[q], gl(a.p.z)[q], e);
            gl(a.keys)[q] = key;
            small = key >= 0 && gl(a.p.vol)[q] <= g.vmin;
        }
        wave_run_add(key, gl(a.count));
        wave_run_add(small ? key : -1, gl(a.scount));
    }
    if (e) atomicOr(a.err, e);
}

// ---------------------------------------------------------------------------
// 2. exclusive scan (three passes, 1024 elements per block)
// ---------------------------------------------------------------------------
constexpr int SCAN_T = 256;
constexpr int SCAN_PER = 4;
constexpr int SCAN_BLOCK = SCAN_T * SCAN_PER;

__device__ inline int block_excl_scan(int v, int *lds, int &total)
{
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {
        int a = (t >= off) ? lds[t - off] : 0;
        __syncthreads();
        lds[t] += a;
        __syncthreads();
    }
    total = lds[SCAN_T - 1];
    const int incl = lds[t];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_reduce(const ScanArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    __shared__ int lds[SCAN_T];
    const ItemBlock ib = item_block(bmap);
    const ScanArgs &a = tab[ib.item];
    const int64_t base = ib.blk * (int64_t)SCAN_BLOCK + threadIdx.x * SCAN_PER;
    int s = 0;
#pragma unroll
    for (int u = 0; u < SCAN_PER; ++u)
        if (base + u < a.n) s += gl(a.in)[base + u];
    int total;
    block_excl_scan(s, lds, total);
    if (threadIdx.x == 0) gl(a.bsum)[ib.blk] = total;
}

// one block per item: exclusive scan of the item's block sums in place
__global__ __launch_bounds__(SCAN_T) void k_scan_blocks(const ScanArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    __shared__ int lds[SCAN_T];
    const ItemBlock ib = item_block(bmap);
    const ScanArgs &a = tab[ib.item];
    const int nb = (int)((a.n + SCAN_BLOCK - 1) / SCAN_BLOCK);
    int carry = 0;
    for (int base = 0; base < nb; base += SCAN_T * SCAN_PER) {
        const int i0 = base + threadIdx.x * SCAN_PER;
        int v[SCAN_PER], s = 0;
#pragma unroll
        for (int u = 0; u < SCAN_PER; ++u) {
            v[u] = (i0 + u < nb) ? gl(a.bsum)[i0 + u] : 0;
            s += v[u];
        }
        int total;
        int ex = block_excl_scan(s, lds, total);
        int run = carry + ex;
#pragma unroll
        for (int u = 0; u < SCAN_PER; ++u) {
            if (i0 + u < nb) gl(a.bsum)[i0 + u] = run;
            run += v[u];
        }
        carry += total;
        __syncthreads();
    }
}

// out[c] = exclusive prefix, out[n] = total; clears gl(a.clear)[c] (the
// placement cursors) on the way
__global__ __launch_bounds__(SCAN_T) void k_scan_apply(const ScanArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    __shared__ int lds[SCAN_T];
    const ItemBlock ib = item_block(bmap);
    const ScanArgs &a = tab[ib.item];
    const int64_t n = a.n;
    const int64_t base = ib.blk * (int64_t)SCAN_BLOCK + threadIdx.x * SCAN_PER;
    int v[SCAN_PER], s = 0;
#pragma unroll
    for (int u = 0; u < SCAN_PER; ++u) {
        v[u] = (base + u < n) ? gl(a.in)[base + u] : 0;
        s += v[u];
    }
    int total;
    int run = block_excl_scan(s, lds, total) + gl(a.bsum)[ib.blk];
#pragma unroll
    for (int u = 0; u < SCAN_PER; ++u) {
        if (base + u < n) {
            gl(a.out)[base + u] = run;
            if (a.clear) gl(a.clear)[base + u] = 0;
        }
        run += v[u];
    }
    // the thread owning element n-1 also writes the grand total out[n]
    if (base <= n - 1 && n - 1 < base + SCAN_PER) gl(a.out)[n] = run;
}

void add_exclusive_scan(Plan &P, int phase, const std::vector<ScanArgs> &items)
{
    std::vector<int> nb, one;
    for (const ScanArgs &a : items) {
        nb.push_back((int)((a.n + SCAN_BLOCK - 1) / SCAN_BLOCK));
        one.push_back(1);
    }
    P.add((const void *)k_scan_reduce, SCAN_T, 0, phase, items, nb);
    P.add((const void *)k_scan_blocks, SCAN_T, 0, phase, items, one);
    P.add((const void *)k_scan_apply, SCAN_T, 0, phase, items, nb);
}

int64_t scan_scratch_ints(int64_t n) { return (n + SCAN_BLOCK - 1) / SCAN_BLOCK + 1; }

// ---------------------------------------------------------------------------
// 3./4. placement and per-cell ordering
// ---------------------------------------------------------------------------
// Runs of equal keys in consecutive lanes (the parcels are nearly in cell
// order) take one ticket per run.
__global__ __launch_bounds__(256) void k_place(const PlaceArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const PlaceArgs &a = tab[ib.item];
    // the rounds' keys, cell starts and tickets are each requested together,
    // so a wave waits three memory round trips, not three per round
    constexpr int NR = PER_BLOCK / 256;   // every lane of a wave runs the same rounds
    const int lane = threadIdx.x & 63;
    const int64_t n = a.n_dev ? *a.n_dev : a.n;
    int key[NR], start[NR], head_lane[NR], got[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int64_t q = (int64_t)ib.blk * PER_BLOCK + r * 256 + threadIdx.x;
        key[r] = (q < n) ? gl(a.keys)[q] : -1;
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) start[r] = (key[r] >= 0) ? gl(a.cell_start)[key[r]] : 0;
#pragma unroll
    for (int r = 0; r < NR; ++r) {   // wave_run_ticket, its atomic left in flight
        const int prev = __shfl_up(key[r], 1, 64);
        const bool head = (lane == 0) || (prev != key[r]);
        const unsigned long long heads = __ballot(head);
        const unsigned long long below = heads & ((lane == 63) ? ~0ull : ((2ull << lane) - 1));
        head_lane[r] = 63 - __clzll(below);
        const unsigned long long above = heads & ~((2ull << lane) - 1 | (lane == 63 ? ~0ull : 0ull));
        const int next = above ? __ffsll((long long)above) - 1 : 64;
        got[r] = (head && key[r] >= 0) ? atomicAdd(&gl(a.cursor)[key[r]], next - lane) : 0;
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int base = __shfl(got[r], head_lane[r], 64) + (lane - head_lane[r]);
        if (key[r] >= 0) gl(a.perm)[start[r] + base] = (int)((int64_t)ib.blk * PER_BLOCK + r * 256 + threadIdx.x);
    }
}

__global__ void k_cell_order(int64_t nc, const int *__restrict__ cell_start, const int *__restrict__ perm_tmp,
                             const int *__restrict__ orig, int *__restrict__ perm)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
        const int s0 = cell_start[c], s1 = cell_start[c + 1];
        const int m = s1 - s0;
        if (m == 0) continue;
        if (m <= 32) {
            int pv[32], ov[32];
            for (int a = 0; a < m; ++a) {
                pv[a] = perm_tmp[s0 + a];
                ov[a] = orig[pv[a]];
            }
            for (int a = 1; a < m; ++a) {   // insertion sort by caller index
                const int kp = pv[a], ko = ov[a];
                int b = a - 1;
                while (b >= 0 && ov[b] > ko) {
                    pv[b + 1] = pv[b];
                    ov[b + 1] = ov[b];
                    --b;
                }
                pv[b + 1] = kp;
                ov[b + 1] = ko;
            }
            for (int a = 0; a < m; ++a) perm[s0 + a] = pv[a];
        } else {
            // crowded cell: selection by rank (stable: caller indices are unique)
            for (int a = 0; a < m; ++a) {
                const int pa = perm_tmp[s0 + a];
                const int oa = orig[pa];
                int rank = 0;
                for (int b = 0; b < m; ++b) rank += (orig[perm_tmp[s0 + b]] < oa);
                perm[s0 + rank] = pa;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// 5. tile scatter into fixed-point LDS accumulators
// ---------------------------------------------------------------------------
// One work-group per (8 x 8 column tile, tz cell layers), P2G_TPC threads per
// cell taking every P2G_TPC-th parcel of it, in neighbouring lanes: a wave
// covers 64 / P2G_TPC cells, so its loads share lines, and the points of
// different cells' parcels rarely add to one node in the same instruction.
// Every contribution w * (vol/4) * value is scaled by a power of two chosen
// per member and field from k_p2g_max's maxima (so that no single point's
// contribution exceeds 2^27),
// rounded to an integer and added to a 64-bit LDS accumulator of its node;
// the points of one parcel that fall in the same cell are combined first.
// Most points stay in their parcel's own cell: there the two threads of the
// cell combine their (rounded) contributions by a lane exchange and each
// adds the pair's sum to four of the cell's eight corners, so these atomics
// are half as many and no two lanes of an instruction hit one node.  Integer
// sums do not depend on the order of the additions, so the node sums are
// bitwise reproducible however the waves are scheduled, with no ordering
// pass and no barriers between parcels; the only rounding is that of each
// contribution (2^-28 of the member's largest) and of the final conversion.
// Node box of a tile: x, y in [-1, 9] around the tile, node planes
// [k0 - 1, k0 + tz + 1]; rows padded to lpitch 64-bit words (12 spreads the
// 8 x 8 lanes of a wave evenly over the 32 LDS banks).

__host__ __device__ constexpr int box_planes(int tz) { return tz + 3; }
// LDS row pitch in 64-bit words; at tz = 8 the unpadded 11 keeps two
// work-groups per CU within the LDS
__host__ __device__ constexpr int p2g_lpitch(int tz) { return tz >= 8 ? P2G_BOX : 12; }

// power-of-two scale s with max_abs * s < 2^27 (0 when max_abs is 0): the
// summed weights of up to four points of one parcel stay below 4, so every
// rounded contribution is below 2^29 and a node's sum over its parcels stays
// far below 2^47
__device__ inline float fixed_scale(float max_abs, int &e)
{
    if (!(max_abs > 0.0f)) {
        e = 0;
        return 0.0f;
    }
    frexpf(max_abs, &e);   // max_abs < 2^e
    return ldexpf(1.0f, 27 - e);
}

// the value of the other lane of a lane pair (2k, 2k+1): DPP quad
// permutation [1, 0, 3, 2]; both lanes must be active
__device__ inline int pair_swap(int v) { return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false); }
__device__ inline unsigned long long pair_swap64(unsigned long long v)
{
    const unsigned lo = (unsigned)pair_swap((int)(unsigned)v), hi = (unsigned)pair_swap((int)(unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

// One fixed-point contribution as a 64-bit accumulator term, in one FP64
// FMA: w * c is exact in FP64 (two FP32 factors), and adding 1.5 * 2^52
// rounds it to the nearest integer q (ties to even) with q held in the low
// bits of the result, whose pattern is F_MAGIC + q for |q| < 2^51.  F_MAGIC
// is zero below bit 48, so a sum of such terms equals the sum of the q
// modulo 2^48 in its low 48 bits, whatever the number of terms: the node
// sums (|sum| < 2^47) are the low 48 bits sign-extended (fixed_sum).
constexpr double F_MAGIC_D = 6755399441055744.0;   // 1.5 * 2^52
__device__ inline unsigned long long fixed_term(double w, double c)
{
    return (unsigned long long)__double_as_longlong(fma(w, c, F_MAGIC_D));
}
__device__ inline long long fixed_sum(unsigned long long acc) { return (long long)(acc << 16) >> 16; }

// the gridded quantities of parcel pi, read from its arrays (o4 = 4 pi)
__device__ inline void p2g_values(const Grid &g, const ParcelsDev &p, unsigned o4, double z, float v[P2G_NF])
{
    const float vor[3] = {at(p.vor[0], o4), at(p.vor[1], o4), at(p.vor[2], o4)};
    p2g_values_of(g, at(p.vol, o4), vor, at(p.buoy, o4), at(p.hum, o4), z, v);
}

// The largest |contribution| of every field over a member's parcels, and
// the largest |total buoyancy - buoyancy| one: one coalesced pass in
// storage order by at most P2G_MAX_BLOCKS work-groups, whose block maxima
// go to out[] by integer atomicMax on the bits of the non-negative floats.
// out[] is zero when this runs: it starts zeroed, and the finalize kernel
// of the same par2grid clears it once the tile kernel has taken its scales
// from it (a value left by an interrupted sequence would only be larger, a
// coarser but still valid scale).  Inside an RK step only the first stage
// runs this pass: each later stage's maxima come from the previous push,
// which has the parcels' new values at hand and bounds the moist
// contributions by the humidity (exact without humidity).
__global__ __launch_bounds__(256) void k_p2g_max(const P2GMaxArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const P2GMaxArgs &a = tab[ib.item];
    float mx[P2G_NF + 1];
#pragma unroll
    for (int f = 0; f <= P2G_NF; ++f) mx[f] = 0.0f;
    const int64_t n = parcel_count(a.p);
    for (int64_t q = (int64_t)ib.blk * 256 + threadIdx.x; q < n; q += (int64_t)a.nblk * 256) {
        float v[P2G_NF];
        p2g_values(a.g, a.p, (unsigned)q * 4u, gl(a.p.z)[q], v);
#pragma unroll
        for (int f = 0; f < P2G_NF; ++f) mx[f] = fmaxf(mx[f], fabsf(v[f]));
        mx[P2G_NF] = fmaxf(mx[P2G_NF], fabsf(v[4] - v[5]));
    }
    __shared__ float wm[4][P2G_NF + 1];
#pragma unroll
    for (int f = 0; f <= P2G_NF; ++f) {
        for (int o = 32; o > 0; o >>= 1) mx[f] = fmaxf(mx[f], __shfl_xor(mx[f], o, 64));
        if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6][f] = mx[f];
    }
    __syncthreads();
    if (threadIdx.x <= P2G_NF) {
        const int f = threadIdx.x;
        const float m = fmaxf(fmaxf(wm[0][f], wm[1][f]), fmaxf(wm[2][f], wm[3][f]));
        if (m > 0.0f) atomicMax((unsigned int *)a.out + f, __float_as_uint(m));
    }
}

// TZ (= a.tz) is a template parameter, so the LDS strides are constants and
// the 8 x 7 atomics of a cell are its base word plus immediate offsets
template <int TZ>
__global__ __launch_bounds__(64 * TZ * P2G_TPC, 8 / TZ) void k_p2g_tiles(const P2GArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const P2GArgs &a = tab[ib.item];
    const Grid &g = a.g;
    // stage 1 of a step after the first takes the cached axes (l_reuse)
    const float *const *axes = (a.use_alt && *a.use_alt) ? a.axes_alt : a.axes;
    constexpr int tz = TZ, lp = p2g_lpitch(TZ);
    const int tx = ib.blk % a.ntx;
    const int ty = (ib.blk / a.ntx) % a.nty;
    const int cz = ib.blk / (a.ntx * a.nty);
    const int i0 = tx * P2G_TX, j0 = ty * P2G_TY;
    const int k0 = -1 + cz * tz;
    const int tid = threadIdx.x;
    constexpr int nbox = box_planes(tz);
    constexpr int lplane = P2G_BOX * lp;          // words per LDS node plane
    constexpr int lfield = nbox * lplane;         // words per field
    extern __shared__ unsigned long long acc[];   // [field][plane][y][x]
    __shared__ int s_err, s_far;
    __shared__ unsigned short s_items[2 * TZ][64];   // per wave: the pass's items (source lane, first point)

    for (int u = tid; u < P2G_NF * lfield; u += blockDim.x) acc[u] = 0ull;
    if (tid == 0) s_err = s_far = 0;

    // this thread's cell and its share of the cell's parcels: the P2G_TPC
    // threads of a cell are neighbouring lanes, so one load instruction
    // covers 64 / P2G_TPC cells, whose parcels share fewer lines
    const int ct = tid / P2G_TPC, sub = tid % P2G_TPC;
    const int cx = ct & 7, cy = (ct >> 3) & 7, lz = ct >> 6;
    const int k = k0 + lz;
    int s0 = 0, s1 = 0;
    if (k <= g.nz && i0 + cx < g.nx && j0 + cy < g.ny) {   // cells of a partial tile beyond the grid: none
        const int c = ((k + 1) * g.ny + j0 + cy) * g.nx + i0 + cx;
        s0 = gl(a.cell_start)[c];
        s1 = gl(a.cell_start)[c + 1];
    }
    // fixed-point scales from the member's largest contributions; with no
    // condensation anywhere the total buoyancy equals the buoyancy (field 5
    // is a copy of field 4) and with no humidity field 6 is zero: their
    // atomics are then skipped and the results are the same bits
    float mx[P2G_NF + 1];
#pragma unroll
    for (int f = 0; f <= P2G_NF; ++f) mx[f] = gl(a.tmax)[f];
    float sc[P2G_NF];
#pragma unroll
    for (int f = 0; f < P2G_NF; ++f) {
        int e;
        sc[f] = fixed_scale(mx[f], e);
    }
    const bool own5 = mx[P2G_NF] > 0.0f;   // tbuoy differs from buoy somewhere
    const bool own6 = mx[6] > 0.0f;        // humidity somewhere
    __syncthreads();   // the accumulators are cleared

    // ---- pass 2: support points, trilinear weights, integer accumulation ----
    // the two threads of a cell take its parcels in turns and run the same
    // number of rounds (the odd thread idles in the last round of an odd
    // count), so that they can exchange their own-cell contributions
    static_assert(P2G_TPC == 2, "lane pairs");
    const bool odd = sub != 0;
    const int cown = (lz + 1) * lplane + (cy + 1) * lp + cx + 1;   // own cell's lower corner
    const int nown = cown + (odd ? lplane : 0);   // this lane's four corners: the lower (even) or upper (odd) plane
    // every lane of a wave runs the wave's largest number of rounds (the
    // non-own cells below are shared out over the whole wave), the lanes of
    // cells with fewer parcels idle
    int rounds = (s1 - s0 + 1) >> 1;
    for (int o = 32; o > 0; o >>= 1) rounds = max(rounds, __shfl_xor(rounds, o, 64));
    const int lane = tid & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned short *items = s_items[tid >> 6];
    int err = 0;
    for (int rd = 0; rd < rounds; ++rd) {
        const int s = s0 + 2 * rd + sub;
        const bool have = s < s1;
        const int pi = have ? gl(a.perm)[s] : 0;   // an idle lane reads parcel 0 and adds nothing
        const unsigned o4 = (unsigned)pi * 4u;   // byte offset (pi < 2^30): scalar base + 32-bit offset addressing
        const double pos[3] = {gl(a.p.x)[pi], gl(a.p.y)[pi], gl(a.p.z)[pi]};
        const double ve[3] = {at(axes[0], o4), at(axes[1], o4), at(axes[2], o4)};
        const double vt[3] = {at(axes[3], o4), at(axes[4], o4), at(axes[5], o4)};
        float v[P2G_NF];
        p2g_values(g, a.p, o4, pos[2], v);
        double csd[P2G_NF];
#pragma unroll
        for (int f = 0; f < P2G_NF; ++f) csd[f] = have ? (double)(v[f] * sc[f]) : 0.0;   // exact: sc is a power of two
        int cc[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) cc[d] = (int)floor((pos[d] - g.lower[d]) * g.dxi[d]);
        // the four support points: target cell (box corner index, -1 when
        // outside the neighbour cells) and cell fractions
        // tgt: >= 0 box corner; -1 none; < -1 a point beyond the neighbour
        // cells, -2 - its cell offsets (7 bits each, biased by 64)
        int tgt[4];
        float fr[4][3];
        unsigned farm = 0;   // bit jp: point jp lies beyond the neighbour cells
#pragma unroll
        for (int jp = 0; jp < 4; ++jp) {
            const double csn = support_cos(jp), snn = support_sin(jp);
            int off[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double pt = pos[d] + ve[d] * csn + vt[d] * snn;
                const double xyz = (pt - g.lower[d]) * g.dxi[d];
                const double fl = floor(xyz);
                off[d] = (int)fl - cc[d];
                fr[jp][d] = (float)(xyz - fl);
            }
            tgt[jp] = -1;
            if (cc[2] + off[2] < -1 || cc[2] + off[2] > g.nz) {   // beyond the halo layers (the reference fails too)
                if (have) err |= ERR_OUT_OF_GRID;
            } else if (off[0] < -1 || off[0] > 1 || off[1] < -1 || off[1] > 1 || off[2] < -1 || off[2] > 1) {
                if (have && off[0] >= -64 && off[0] < 64 && off[1] >= -64 && off[1] < 64 && off[2] >= -64 && off[2] < 64) {
                    farm |= 1u << jp;   // beyond the neighbour cells: global accumulator, below
                    tgt[jp] = -2 - (((off[2] + 64) << 14) | ((off[1] + 64) << 7) | (off[0] + 64));
                } else if (have) {
                    err |= ERR_OUT_OF_GRID;   // more than 64 cells away
                }
            } else {   // lower corner of the point's cell in the box
                tgt[jp] = (lz + off[2] + 1) * lplane + (cy + off[1] + 1) * lp + cx + off[0] + 1;
            }
        }
        // support points beyond the neighbour cells (strongly stretched
        // parcels, rare): each point's 8 x 7 fixed-point terms go to the
        // member's global accumulator by 64-bit atomics (exact integer sums,
        // so still independent of the order), which the finalize adds
        if (farm) {
            s_far = 1;
            const int64_t FS = field_stride(g);
#pragma unroll
            for (int jp = 0; jp < 4; ++jp) {
                if (!((farm >> jp) & 1u)) continue;
                const int code = -2 - tgt[jp];
                const int gi[3] = {cc[0] + (code & 127) - 64, cc[1] + ((code >> 7) & 127) - 64, cc[2] + (code >> 14) - 64};
                const float *f3 = fr[jp];
                const float wxy[4] = {(1.0f - f3[1]) * (1.0f - f3[0]), (1.0f - f3[1]) * f3[0], f3[1] * (1.0f - f3[0]),
                                      f3[1] * f3[0]};
#pragma unroll 1
                for (int c = 0; c < 8; ++c) {
                    const int ii = wrap(gi[0] + (c & 1), g.nx), jj = wrap(gi[1] + ((c >> 1) & 1), g.ny);
                    const int kk = gi[2] + (c >> 2);   // in [-1, nz + 1]
                    const int64_t node = ((int64_t)(kk + 1) * g.ny + jj) * g.nx + ii;
                    const double wd = (double)(((c & 4) ? f3[2] : 1.0f - f3[2]) * wxy[c & 3]);
#pragma unroll
                    for (int f = 0; f < P2G_NF; ++f) atomicAdd(a.far_acc + f * FS + node, fixed_term(wd, csd[f]));
                }
            }
        }
        // own cell: the summed weights (in point order) of the points in it,
        // on this lane's four corners and on the other lane's four
        {
            float wn[4] = {0.0f, 0.0f, 0.0f, 0.0f}, wf[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            bool any = false;
#pragma unroll
            for (int jp = 0; jp < 4; ++jp) {
                if (tgt[jp] != cown) continue;
                const float *f3 = fr[jp];
                const float wxy[4] = {(1.0f - f3[1]) * (1.0f - f3[0]), (1.0f - f3[1]) * f3[0], f3[1] * (1.0f - f3[0]),
                                      f3[1] * f3[0]};
                const float zl = 1.0f - f3[2], zu = f3[2];
                const float zn = odd ? zu : zl, zf = odd ? zl : zu;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    wn[c] = any ? wn[c] + zn * wxy[c] : zn * wxy[c];
                    wf[c] = any ? wf[c] + zf * wxy[c] : zf * wxy[c];
                }
                any = true;
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int node = nown + ((c >> 1) & 1) * lp + (c & 1);
                const double wnd = (double)wn[c], wfd = (double)wf[c];
#pragma unroll
                for (int f = 0; f < P2G_NF; ++f) {
                    if ((f == 5 && !own5) || (f == 6 && !own6)) continue;   // uniform over the work-group
                    // this lane's parcel on its plane plus the other lane's parcel on it
                    const unsigned long long t = fixed_term(wnd, csd[f]) + pair_swap64(fixed_term(wfd, csd[f]));
                    atomicAdd(&acc[f * lfield + node], t);
                }
            }
        }
        // every other cell the points reach: its 8 corners take the summed
        // weights of its points (in point order), 8 x 7 atomics per cell.
        // These (lane, cell) items, about one for every second parcel, are
        // shared out over the wave: item P[k] + (rank of the lane among the
        // lanes with more than k items) is a lane's k-th; the lanes take 64
        // items per pass, find their source lane through a per-wave list and
        // read its fractions, targets and values by ds_bpermute, so the
        // atomics run with nearly every lane active (one pass per round as a
        // rule, where a loop over the four points ran up to four sparse ones).
        // Each item's weights and terms are the same arithmetic as in its
        // source lane.
        unsigned leadm = 0;   // bit jp: point jp opens a cell other than its own and earlier points'
#pragma unroll
        for (int jp = 0; jp < 4; ++jp) {
            bool lead = tgt[jp] >= 0 && tgt[jp] != cown && have;
#pragma unroll
            for (int jq = 0; jq < jp; ++jq) lead = lead && (tgt[jq] != tgt[jp]);
            leadm |= (lead ? 1u : 0u) << jp;
        }
        const int ni = __popc(leadm);
        unsigned long long lb[4];
        int pk[5];
        pk[0] = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            lb[k] = __ballot(ni > k);
            pk[k + 1] = pk[k] + __popcll(lb[k]);
        }
        const int total = pk[4];   // uniform over the wave
        if (total > 0) {
            float csf[P2G_NF];
#pragma unroll
            for (int f = 0; f < P2G_NF; ++f) csf[f] = have ? v[f] * sc[f] : 0.0f;
            // box corners in 16 bits each; no target (also a far point) as -1
            int tb[4];
#pragma unroll
            for (int jq = 0; jq < 4; ++jq) tb[jq] = tgt[jq] >= 0 ? tgt[jq] : -1;
            const int t01 = (tb[0] & 0xffff) | (tb[1] << 16), t23 = (tb[2] & 0xffff) | (tb[3] << 16);
            for (int r = 0; r < total; r += 64) {
                unsigned lm = leadm;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int jp = __ffs(lm) - 1;   // the lane's k-th item: point jp
                    lm &= lm - 1u;
                    const int idx = pk[k] + __popcll(lb[k] & below) - r;
                    if (k < ni && idx >= 0 && idx < 64) items[idx] = (unsigned short)(lane | (jp << 8));
                }
                __builtin_amdgcn_wave_barrier();
                const bool mine = r + lane < total;
                const unsigned e = mine ? (unsigned)items[lane] : 0u;
                const int sa = (int)(e & 63u) << 2, jp0 = (int)(e >> 8) & 3;
                float g[4][3];
#pragma unroll
                for (int jq = 0; jq < 4; ++jq)
#pragma unroll
                    for (int d = 0; d < 3; ++d) g[jq][d] = __int_as_float(__builtin_amdgcn_ds_bpermute(sa, __float_as_int(fr[jq][d])));
                float cf[P2G_NF];
#pragma unroll
                for (int f = 0; f < P2G_NF; ++f)
                    cf[f] = ((f == 5 && !own5) || (f == 6 && !own6)) ? 0.0f
                            : __int_as_float(__builtin_amdgcn_ds_bpermute(sa, __float_as_int(csf[f])));
                const int u01 = __builtin_amdgcn_ds_bpermute(sa, t01), u23 = __builtin_amdgcn_ds_bpermute(sa, t23);
                __builtin_amdgcn_wave_barrier();   // the list is read before the next pass writes it
                if (mine) {
                    const int tg[4] = {(int)(short)(u01 & 0xffff), u01 >> 16, (int)(short)(u23 & 0xffff), u23 >> 16};
                    const int t = jp0 == 0 ? tg[0] : jp0 == 1 ? tg[1] : jp0 == 2 ? tg[2] : tg[3];
                    float w[8];
#pragma unroll
                    for (int c = 0; c < 8; ++c) w[c] = 0.0f;
#pragma unroll
                    for (int jq = 0; jq < 4; ++jq) {
                        if (jq < jp0 || (jq > jp0 && tg[jq] != t)) continue;
                        const float *f3 = g[jq];
                        const float wxy[4] = {(1.0f - f3[1]) * (1.0f - f3[0]), (1.0f - f3[1]) * f3[0],
                                              f3[1] * (1.0f - f3[0]), f3[1] * f3[0]};
#pragma unroll
                        for (int c = 0; c < 8; ++c) {
                            const float wc = ((c & 4) ? f3[2] : 1.0f - f3[2]) * wxy[c & 3];
                            w[c] = (jq == jp0) ? wc : w[c] + wc;
                        }
                    }
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        const int node = t + (c >> 2) * lplane + ((c >> 1) & 1) * lp + (c & 1);
                        const double wd = (double)w[c];
#pragma unroll
                        for (int f = 0; f < P2G_NF; ++f) {
                            if ((f == 5 && !own5) || (f == 6 && !own6)) continue;   // uniform over the work-group
                            atomicAdd(&acc[f * lfield + node], fixed_term(wd, (double)cf[f]));
                        }
                    }
                }
            }
        }
    }
    if (err) atomicOr(&s_err, err);
    __syncthreads();

    // ---- the box to staging (FP64, node-major, fields fastest) ----
    // q / scale is exact in FP64 (|q| < 2^53, the scale a power of two), so
    // the staged partial sums keep every bit of the integer sums
    double inv[P2G_NF];
#pragma unroll
    for (int f = 0; f < P2G_NF; ++f) inv[f] = (sc[f] > 0.0f) ? 1.0 / (double)sc[f] : 0.0;
    constexpr int PLANE = P2G_BOX * P2G_BOX * P2G_NF;
    EPIC_GLOBAL double *stage = gl(a.staging) + (int64_t)ib.blk * nbox * PLANE;
    for (int u = tid; u < nbox * PLANE; u += blockDim.x) {
        const int f = u % P2G_NF, nd = u / P2G_NF;
        const int pz = nd / (P2G_BOX * P2G_BOX), yx = nd - pz * (P2G_BOX * P2G_BOX);
        const int y = yx / P2G_BOX, x = yx - y * P2G_BOX;
        const int fs = (f == 5 && !own5) ? 4 : f;   // a skipped field 5 is field 4
        const long long q = (f == 6 && !own6) ? 0ll : fixed_sum(acc[fs * lfield + pz * lplane + y * lp + x]);
        stage[u] = (double)q * inv[fs];
    }
    if (tid == 0 && s_err) atomicOr(a.err, s_err);
    if (tid == 0 && s_far) atomicOr(a.far_flag, 1);
    if (ib.blk == 0 && tid < P2G_NF) a.far_scale[tid] = inv[tid];   // the same in every work-group
}

// ---------------------------------------------------------------------------
// 6. staging sum + folds + normalisation (one thread per (i, j) column)
// ---------------------------------------------------------------------------
// adds the staged sums of node (i, j, k) to out[], and its far-point sums
// when there are any (far)
__device__ inline void stage_add(const FinalizeArgs &a, int i, int j, int k, double out[P2G_NF], bool far)
{
    constexpr int PLANE = P2G_BOX * P2G_BOX * P2G_NF;
    const Grid &g = a.g;
    if (far) {
        const int64_t FS = field_stride(g), o = ((int64_t)(k + 1) * g.ny + j) * g.nx + i;
#pragma unroll
        for (int f = 0; f < P2G_NF; ++f) out[f] += (double)fixed_sum(a.far_acc[f * FS + o]) * a.far_scale[f];
    }
    const int nbox = a.tz + 3;
    const int cza = (k + 1) / a.tz;
    const int txa = i / P2G_TX, tya = j / P2G_TY;
    for (int dcz = -1; dcz <= 1; ++dcz) {
        const int cz = cza + dcz;
        if (cz < 0 || cz >= a.nchz) continue;
        const int lp = k - (cz * a.tz - 2);   // box plane index
        if (lp < 0 || lp >= nbox) continue;
            for (int dty = -1; dty <= 1; ++dty) {
            // distinct candidate tiles only (grids of one or two tiles wrap)
            if ((a.nty == 1 && dty != 0) || (a.nty == 2 && dty == 1)) continue;
            const int ty = wrap(tya + dty, a.nty);
            for (int ly = wrap(j - (ty * P2G_TY - 1), g.ny); ly < P2G_BOX; ly += g.ny) {
                for (int dtx = -1; dtx <= 1; ++dtx) {
                    if ((a.ntx == 1 && dtx != 0) || (a.ntx == 2 && dtx == 1)) continue;
                    const int tx = wrap(txa + dtx, a.ntx);
                    for (int lx = wrap(i - (tx * P2G_TX - 1), g.nx); lx < P2G_BOX; lx += g.nx) {
                        const int wg = (cz * a.nty + ty) * a.ntx + tx;
                        const EPIC_GLOBAL double *s = gl(a.staging) + ((int64_t)wg * nbox + lp) * PLANE + (ly * P2G_BOX + lx) * P2G_NF;
#pragma unroll
                        for (int f = 0; f < P2G_NF; ++f) out[f] += s[f];
                    }
                }
            }
        }
    }
}

// one thread per (column, node plane k = -1..nz+1), consecutive threads on
// consecutive columns.  Planes 0, 1, nz-1, nz carry the free-slip folds, the
// halo planes -1 and nz+1 the raw sums and the z extrapolation; every thread
// recomputes the raw planes it needs, so the planes are independent.
__global__ __launch_bounds__(256) void k_p2g_finalize(const FinalizeArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const FinalizeArgs &a = tab[ib.item];
    const Grid &g = a.g;
    const int nz = g.nz;
    const int64_t ncol = (int64_t)g.nx * g.ny;
    const int64_t tq = ib.blk * (int64_t)blockDim.x + threadIdx.x;
    if (a.clear_tmax && tq <= P2G_NF) gl(a.clear_tmax)[tq] = 0.0f;   // the tile kernel has read them
    if (tq >= ncol * (nz + 3)) return;
    const bool far = *a.far_flag != 0;   // k_far_clear resets it after this kernel
    const int k = (int)(tq / ncol) - 1;
    const int64_t col = tq - (int64_t)(k + 1) * ncol;
    const int i = (int)(col % g.nx), j = (int)(col / g.nx);
    const int64_t PS = plane_stride(g), FS = field_stride(g);
    const int64_t base = (int64_t)j * g.nx + i;
    auto F = [&](int slot, int kk) -> EPIC_GLOBAL float & { return gl(a.fld)[slot * FS + (int64_t)(kk + 1) * PS + base]; };

    // the node value of plane kk (folds of parcel_interpl.f90:258-287); the
    // staged sums add exactly, the folds and the normalisation run in FP64
    // and each field is rounded to FP32 once, when it is stored
    auto folded = [&](int kk, double r[P2G_NF]) {
#pragma unroll
        for (int f = 0; f < P2G_NF; ++f) r[f] = 0.0;
        stage_add(a, i, j, kk, r, far);
        if (kk == 0 || kk == nz) {
#pragma unroll
            for (int f = 0; f < P2G_NF; ++f) r[f] *= 2.0;
        } else if (kk == 1 || kk == nz - 1) {
            stage_add(a, i, j, kk == 1 ? -1 : nz + 1, r, far);
        }
    };
    auto zero_zeta = [&](int kk) { return (kk == 0 && g.zeta_zero[0]) || (kk == nz && g.zeta_zero[1]); };

    if (k >= 0 && k <= nz) {
        double r[P2G_NF];
        folded(k, r);
        const double v = r[0];
        F(F_VOLG, k) = (float)v;
        F(F_VORTG + 0, k) = (float)(r[1] / v);
        F(F_VORTG + 1, k) = (float)(r[2] / v);
        F(F_VORTG + 2, k) = zero_zeta(k) ? 0.0f : (float)(r[3] / v);
        F(F_TBUOYG, k) = (float)(r[4] / v);
        F(F_DBUOYG, k) = (float)(r[5] / v);
        F(F_HUMG, k) = (float)(r[6] / v);
    } else {
        // halo planes keep the raw sums for vol/dbuoy/hum; vortg and tbuoyg
        // are extrapolated from the normalised boundary planes (:306-318)
        double r[P2G_NF];
#pragma unroll
        for (int f = 0; f < P2G_NF; ++f) r[f] = 0.0;
        stage_add(a, i, j, k, r, far);
        F(F_VOLG, k) = (float)r[0];
        F(F_DBUOYG, k) = (float)r[5];
        F(F_HUMG, k) = (float)r[6];
        const int kb = (k < 0) ? 0 : nz, kn = (k < 0) ? 1 : nz - 1;
        double e2[4];   // twice the boundary plane's normalised values
        folded(kb, r);
#pragma unroll
        for (int c = 0; c < 3; ++c) e2[c] = (c == 2 && zero_zeta(kb)) ? 0.0 : 2.0 * (r[1 + c] / r[0]);
        e2[3] = 2.0 * (r[4] / r[0]);
        folded(kn, r);
#pragma unroll
        for (int c = 0; c < 3; ++c) F(F_VORTG + c, k) = (float)(e2[c] - r[1 + c] / r[0]);
        F(F_TBUOYG, k) = (float)(e2[3] - r[4] / r[0]);
    }
    // parcel counts (:201-205, folds :239-243)
    auto cnt = [&](const int *c, int kk) -> int {
        return (kk >= -1 && kk <= nz) ? gl(c)[((int64_t)(kk + 1) * g.ny + j) * g.nx + i] : 0;
    };
    int np = cnt(a.count, k), ns = cnt(a.scount, k);
    if (k == 0) {
        np += cnt(a.count, -1);
        ns += cnt(a.scount, -1);
    }
    if (k == nz - 1) {
        np += cnt(a.count, nz);
        ns += cnt(a.scount, nz);
    }
    gl(a.nparg)[(int64_t)(k + 1) * PS + base] = np;
    gl(a.nsparg)[(int64_t)(k + 1) * PS + base] = ns;
    // clear the histogram the next push fills (never the one read here)
    if (a.clear_count && k <= nz) {
        const int64_t c = ((int64_t)(k + 1) * g.ny + j) * g.nx + i;
        gl(a.clear_count)[c] = 0;
        gl(a.clear_scount)[c] = 0;
    }
}

// ---------------------------------------------------------------------------
// support-point half axes, once per parcel per stage (FP64 eigen-solve)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_axes(const AxesArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const AxesArgs &a = tab[ib.item];
    const ParcelsDev &p = a.p;
    const int64_t n = parcel_count(p);
    int e = 0;
#pragma unroll
    for (int r = 0; r < PER_BLOCK / 256; ++r) {
        const int64_t q = (int64_t)ib.blk * PER_BLOCK + r * 256 + threadIdx.x;
        if (q >= n) break;
        const double B[5] = {gl(p.B[0])[q], gl(p.B[1])[q], gl(p.B[2])[q], gl(p.B[3])[q], gl(p.B[4])[q]};
        double ve[3], vt[3];
        if (!support_axes(B, (double)gl(p.vol)[q], ve, vt)) {
            e |= ERR_B33;
            ve[0] = ve[1] = ve[2] = vt[0] = vt[1] = vt[2] = 0.0;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            gl(a.ax[d])[q] = (float)ve[d];
            gl(a.ax[3 + d])[q] = (float)vt[d];
        }
    }
    if (e) atomicOr(a.err, e);
}

// ---------------------------------------------------------------------------
// plan builders
// ---------------------------------------------------------------------------
static std::vector<int> parcel_blocks_of(const std::vector<int64_t> &n)
{
    std::vector<int> b;
    for (int64_t v : n) b.push_back((int)((v + PER_BLOCK - 1) / PER_BLOCK));
    return b;
}

// parcels a launch's grid covers: the allocated slots when known, so that a
// plan stays the same launch sequence when the parcel count changes
static int64_t grid_parcels(const ParcelsDev &p) { return std::max(p.n, p.cap); }

void add_cell_keys(Plan &P, int phase, const std::vector<KeysArgs> &items)
{
    std::vector<int64_t> n;
    for (const KeysArgs &a : items) n.push_back(grid_parcels(a.p));
    P.add((const void *)k_cell_keys, 256, 0, phase, items, parcel_blocks_of(n));
}

void add_place(Plan &P, int phase, const std::vector<PlaceArgs> &items)
{
    std::vector<int64_t> n;
    for (const PlaceArgs &a : items) n.push_back(std::max(a.n, a.cap));
    P.add((const void *)k_place, 256, 0, phase, items, parcel_blocks_of(n));
}

void add_p2g_max(Plan &P, int phase, const std::vector<P2GMaxArgs> &items)
{
    std::vector<int> nb;
    for (const P2GMaxArgs &a : items) nb.push_back(a.nblk);
    P.add((const void *)k_p2g_max, 256, 0, phase, items, nb);
}

int p2g_max_blocks(int64_t n) { return (int)std::min<int64_t>((n + 1023) / 1024, P2G_MAX_BLOCKS); }

void add_axes(Plan &P, int phase, const std::vector<AxesArgs> &items)
{
    std::vector<int64_t> n;
    for (const AxesArgs &a : items) n.push_back(grid_parcels(a.p));
    P.add((const void *)k_axes, 256, 0, phase, items, parcel_blocks_of(n));
}

void launch_cell_order(int64_t nc, const int *cell_start, const int *perm_tmp, const int *orig, int *perm, hipStream_t s)
{
    const int nb = (int)std::min<int64_t>((nc + 255) / 256, 8192);
    hipLaunchKernelGGL(k_cell_order, dim3(nb), dim3(256), 0, s, nc, cell_start, perm_tmp, orig, perm);
}

P2GLayout p2g_layout(const Grid &g)
{
    P2GLayout L;
    L.tz = 2;
    // the last tile of a row or column is partial when nx or ny is not a
    // multiple of 8: its cells beyond the grid hold no parcels
    L.ntx = (g.nx + P2G_TX - 1) / P2G_TX;
    L.nty = (g.ny + P2G_TY - 1) / P2G_TY;
    L.nchz = (g.nz + 2 + L.tz - 1) / L.tz;
    L.staging_elems = (int64_t)L.ntx * L.nty * L.nchz * (L.tz + 3) * P2G_BOX * P2G_BOX * P2G_NF;
    return L;
}


size_t p2g_lds_bytes(int tz) { return sizeof(unsigned long long) * P2G_NF * box_planes(tz) * P2G_BOX * p2g_lpitch(tz); }

// after a finalize: the far-point accumulator back to zero (only when the
// tile kernel put terms there), then the flag
constexpr int FAR_CLEAR_BLOCKS = 1024;
__global__ __launch_bounds__(256) void k_far_clear(const FarClearArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const FarClearArgs &a = tab[ib.item];
    if (*a.far_flag == 0) return;   // uniform over the launch
    for (int64_t q = (int64_t)ib.blk * 256 + threadIdx.x; q < a.n; q += (int64_t)FAR_CLEAR_BLOCKS * 256)
        gl(a.far_acc)[q] = 0ull;
}

__global__ __launch_bounds__(64) void k_far_reset(const FarClearArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    if (threadIdx.x == 0) gl(tab[ib.item].far_flag)[0] = 0;
}

constexpr size_t P2G_LDS_FLOOR = 160 * 1024 / 4 + 256;   // admits three work-groups per CU, not four

void add_p2g(Plan &P, int phase, const std::vector<P2GArgs> &tiles, const std::vector<FinalizeArgs> &fin,
             const std::vector<FarClearArgs> &far)
{
    // one grouped launch per z chunk (the work-group size is 64 tz P2G_TPC),
    // the kernel specialised on it (epic_set_tiling and the autotuner admit
    // 2, 4 and 8 only)
    std::vector<int> tzs;
    for (const P2GArgs &a : tiles)
        if (std::find(tzs.begin(), tzs.end(), a.tz) == tzs.end()) tzs.push_back(a.tz);
    for (int tz : tzs) {
        std::vector<P2GArgs> items;
        std::vector<int> nb;
        for (P2GArgs a : tiles) {
            if (a.tz != tz) continue;
            items.push_back(a);
            nb.push_back(a.ntx * a.nty * a.nchz);
        }
        const void *k = tz == 2 ? (const void *)k_p2g_tiles<2> : tz == 4 ? (const void *)k_p2g_tiles<4> : (const void *)k_p2g_tiles<8>;
        // at least a third of the CU's LDS a work-group: with four work-groups
        // of tz = 2 on a CU, the rows of parcels their waves touch leave the
        // L2 between a wave's rounds (the tile scatter 1.6 % slower than with
        // three); larger tz need more than that anyway
        const size_t lds = std::max<size_t>(p2g_lds_bytes(tz), P2G_LDS_FLOOR);
        P.add(k, 64 * tz * P2G_TPC, lds, phase == PH_P2G ? (int)PH_P2G_TILES : phase, items, nb);
    }
    std::vector<int> nf;
    for (const FinalizeArgs &f : fin) nf.push_back((int)(((int64_t)f.g.nx * f.g.ny * (f.g.nz + 3) + 255) / 256));
    P.add((const void *)k_p2g_finalize, 256, 0, phase, fin, nf);
    std::vector<int> nc(far.size(), FAR_CLEAR_BLOCKS), nr(far.size(), 1);
    P.add((const void *)k_far_clear, 256, 0, phase, far, nc);
    P.add((const void *)k_far_reset, 64, 0, phase, far, nr);
}

}  // namespace epic
