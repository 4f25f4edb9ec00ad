// par2grid on gfx950: deterministic, atomic-free float accumulation.
//
// Pipeline (one member):
//   1. k_cell_keys     FP64 cell of every parcel centre (get_index,
//                      fields.f90:165-172), integer histogram of parcels and
//                      small parcels per cell (nparg / nsparg)
//   2. scan            exclusive prefix sum of the histogram -> cell_start
//   3. k_place         slot = cell_start + integer ticket (order not fixed yet)
//   4. k_cell_order    order every cell's slots by caller parcel index, so the
//                      sorted order is a function of the parcel set alone
//   5. k_p2g_tiles     one work-group per (8x8 column tile, z chunk): marches
//                      up the chunk one cell layer at a time; the layer's
//                      parcels are staged in LDS (support points computed in
//                      FP64); every extended cell gathers the points that
//                      fall in it from its 3x3 neighbours in a fixed order,
//                      accumulates its 8 corners x 7 fields in registers and
//                      adds them to an LDS ring of node planes in 8
//                      parity phases (cells of equal parity share no node);
//                      finished planes are written to a per-tile staging box
//   6. k_p2g_finalize  per node, sums the staging boxes that cover it in a
//                      fixed tile order (the periodic halo accumulation of
//                      field_mpi.f90), applies the free-slip folds,
//                      normalises by volume, the zeta wall condition and the
//                      z extrapolation (parcel_interpl.f90:235-319)
// The sum at every node is formed in an order that depends only on the
// parcel set, never on scheduling: results are bitwise reproducible.
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
    int e = 0;
#pragma unroll
    for (int r = 0; r < PER_BLOCK / 256; ++r) {
        const int64_t q = (int64_t)ib.blk * PER_BLOCK + r * 256 + threadIdx.x;
        if (q >= a.p.n) break;
        const int key = cell_key(g, a.p.x[q], a.p.y[q], a.p.z[q], e);
        a.keys[q] = key;
        if (key >= 0) {
            atomicAdd(&a.count[key], 1);
            if (a.p.vol[q] <= g.vmin) atomicAdd(&a.scount[key], 1);
        }
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
        if (base + u < a.n) s += a.in[base + u];
    int total;
    block_excl_scan(s, lds, total);
    if (threadIdx.x == 0) a.bsum[ib.blk] = total;
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
            v[u] = (i0 + u < nb) ? a.bsum[i0 + u] : 0;
            s += v[u];
        }
        int total;
        int ex = block_excl_scan(s, lds, total);
        int run = carry + ex;
#pragma unroll
        for (int u = 0; u < SCAN_PER; ++u) {
            if (i0 + u < nb) a.bsum[i0 + u] = run;
            run += v[u];
        }
        carry += total;
        __syncthreads();
    }
}

// out[c] = exclusive prefix, out[n] = total; clears a.clear[c] (the
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
        v[u] = (base + u < n) ? a.in[base + u] : 0;
        s += v[u];
    }
    int total;
    int run = block_excl_scan(s, lds, total) + a.bsum[ib.blk];
#pragma unroll
    for (int u = 0; u < SCAN_PER; ++u) {
        if (base + u < n) {
            a.out[base + u] = run;
            if (a.clear) a.clear[base + u] = 0;
        }
        run += v[u];
    }
    // the thread owning element n-1 also writes the grand total out[n]
    if (base <= n - 1 && n - 1 < base + SCAN_PER) a.out[n] = run;
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
#pragma unroll
    for (int r = 0; r < PER_BLOCK / 256; ++r) {   // every lane of a wave runs the same rounds
        const int64_t q = (int64_t)ib.blk * PER_BLOCK + r * 256 + threadIdx.x;
        const int key = (q < a.n) ? a.keys[q] : -1;
        const int base = wave_run_ticket(key, a.cursor);
        if (key >= 0) a.perm[a.cell_start[key] + base] = (int)q;
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
// 5. tile scatter
// ---------------------------------------------------------------------------
constexpr int NEXT = 300;   // extended cells of a layer: 10 x 10 x 3

__device__ inline int box_planes(int tz) { return tz + 3; }

__global__ __launch_bounds__(P2G_THREADS, 3) void k_p2g_tiles(const P2GArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const P2GArgs &a = tab[ib.item];
    const Grid &g = a.g;
    // stage 1 of a step after the first takes the cached axes (l_reuse)
    const float *const *axes = (a.use_alt && *a.use_alt) ? a.axes_alt : a.axes;
    const int tx = ib.blk % a.ntx;
    const int ty = (ib.blk / a.ntx) % a.nty;
    const int cz = ib.blk / (a.ntx * a.nty);
    const int i0 = tx * P2G_TX, j0 = ty * P2G_TY;
    const int k0 = -1 + cz * a.tz;
    const int k1 = min(k0 + a.tz, g.nz + 1);   // layers [k0, k1)
    const int tid = threadIdx.x;

    constexpr int PLANE = P2G_BOX * P2G_BOX * P2G_NF;
    constexpr int NCELL = P2G_TX * P2G_TY;
    __shared__ float ring[4 * PLANE];
    __shared__ float frac[P2G_MAXP * 12];
    __shared__ float vals[P2G_MAXP * 7];
    __shared__ __attribute__((aligned(8))) unsigned short extc[P2G_MAXP * 4];   // extended cell of every support point (511: none)
    __shared__ unsigned short ord[P2G_MAXP];       // batch entries of each cell in caller order
    __shared__ unsigned char cell_of[P2G_MAXP];    // own cell of each batch entry
    __shared__ __attribute__((aligned(16))) unsigned short items[P2G_MAXP * 4];   // support points binned by extended cell
    int *const origk = (int *)items;               // caller index of each batch entry (until the binning)
    __shared__ int bin_start[NEXT];
    __shared__ int wave_tot[P2G_THREADS / 64];
    __shared__ int cell_off[NCELL + 1];            // concatenated offsets of the layer's cells
    __shared__ int cell_cnt[NCELL];
    __shared__ int row_slot[P2G_TY];
    __shared__ int row_off[P2G_TY + 1];
    __shared__ int batch_cell[NCELL + 1];          // cell-aligned batch boundaries
    __shared__ int s_nbatch, s_err;

    for (int u = tid; u < 4 * PLANE; u += P2G_THREADS) ring[u] = 0.0f;
    if (tid == 0) s_err = 0;

    const int nbox = box_planes(a.tz);
    float *stage = a.staging + (int64_t)ib.blk * nbox * PLANE;
    // extended cell owned by this thread in the accumulation (300 used)
    const bool ecell = tid < NEXT;
    const int ez = tid / 100 - 1;
    const int ey = (tid % 100) / 10 - 1;
    const int ex = tid % 10 - 1;

    auto flush_plane = [&](int P) {   // node plane P -> staging, then clear
        float *dst = stage + (int64_t)(P - (k0 - 1)) * PLANE;
        float *src = ring + ((P + 8) & 3) * PLANE;
        for (int u = tid; u < PLANE; u += P2G_THREADS) {
            dst[u] = src[u];
            src[u] = 0.0f;
        }
    };

    for (int k = k0; k < k1; ++k) {
        __syncthreads();
        // ---- the layer's cells: rows of 8 consecutive cells in the sort ----
        if (tid < P2G_TY) {
            const int c0 = ((k + 1) * g.ny + (j0 + tid)) * g.nx + i0;
            row_slot[tid] = a.cell_start[c0];
            row_off[tid + 1] = a.cell_start[c0 + P2G_TX] - a.cell_start[c0];
        }
        __syncthreads();
        if (tid == 0) {
            row_off[0] = 0;
            for (int r = 0; r < P2G_TY; ++r) row_off[r + 1] += row_off[r];
        }
        __syncthreads();
        if (tid < NCELL) {   // one lane per cell of the layer
            const int r = tid / P2G_TX, cx = tid % P2G_TX;
            const int base = ((k + 1) * g.ny + (j0 + r)) * g.nx + i0 + cx;
            const int s0 = a.cell_start[base];
            cell_cnt[tid] = a.cell_start[base + 1] - s0;
            cell_off[tid] = row_off[r] + (s0 - row_slot[r]);
            if (tid == 0) cell_off[NCELL] = row_off[P2G_TY];
        }
        __syncthreads();
        if (tid == 0) {
            // cell-aligned batches of at most P2G_MAXP parcels (one batch unless crowded)
            int nb = 0;
            batch_cell[0] = 0;
            if (row_off[P2G_TY] > P2G_MAXP) {
                int acc = 0;
                for (int c = 0; c < NCELL; ++c) {
                    const int m = cell_cnt[c];
                    if (m > P2G_MAXP) s_err |= ERR_OUT_OF_GRID;   // one cell beyond a batch: not supported
                    if (acc + m > P2G_MAXP) {
                        batch_cell[++nb] = c;
                        acc = 0;
                    }
                    acc += m;
                }
            }
            batch_cell[++nb] = NCELL;
            s_nbatch = nb;
        }
        __syncthreads();
        const int nbatch = s_nbatch;

        for (int bt = 0; bt < nbatch; ++bt) {
            const int cb0 = batch_cell[bt], cb1 = batch_cell[bt + 1];
            const int q0 = cell_off[cb0];
            const int nb = cell_off[cb1] - q0;
            if (nb > P2G_MAXP) continue;   // flagged above
            // owner cell of every entry
            for (int c = cb0 + tid; c < cb1; c += P2G_THREADS)
                for (int q = cell_off[c]; q < cell_off[c + 1]; ++q) cell_of[q - q0] = (unsigned char)c;
            __syncthreads();
            // ---- load: support points (FP64 positions, cached FP32 half axes) ----
            for (int q = tid; q < nb; q += P2G_THREADS) {
                const int gq = q0 + q;
                const int c = cell_of[q];
                const int r = c / P2G_TX, cx = c % P2G_TX;
                const int slot = row_slot[r] + (gq - row_off[r]);
                const int pi = a.perm[slot];
                origk[q] = a.p.orig[pi];
                const double pos[3] = {a.p.x[pi], a.p.y[pi], a.p.z[pi]};
                const double ve[3] = {axes[0][pi], axes[1][pi], axes[2][pi]};
                const double vt[3] = {axes[3][pi], axes[4][pi], axes[5][pi]};
                int cc[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) cc[d] = (int)floor((pos[d] - g.lower[d]) * g.dxi[d]);
#pragma unroll
                for (int jp = 0; jp < 4; ++jp) {
                    const double cs = support_cos(jp), sn = support_sin(jp);
                    int off[3];
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const double pt = pos[d] + ve[d] * cs + vt[d] * sn;
                        const double xyz = (pt - g.lower[d]) * g.dxi[d];
                        const double fl = floor(xyz);
                        off[d] = (int)fl - cc[d];
                        frac[q * 12 + jp * 3 + d] = (float)(xyz - fl);
                    }
                    int e = 511;
                    if (off[0] >= -1 && off[0] <= 1 && off[1] >= -1 && off[1] <= 1 && off[2] >= -1 && off[2] <= 1 &&
                        cc[2] + off[2] >= -1 && cc[2] + off[2] <= g.nz)
                        e = (off[2] + 1) * 100 + (r + off[1] + 1) * 10 + (cx + off[0] + 1);
                    else
                        atomicOr(&s_err, ERR_OUT_OF_GRID);
                    extc[q * 4 + jp] = (unsigned short)e;
                }
                const double qc = fmax(0.0, (double)a.p.hum[pi] - g.q0 * exp(g.lambda_c * (g.lower[2] - pos[2])));
                const double btot = (double)a.p.buoy[pi] + g.glat * qc;
                float *v = vals + q * 7;
                v[0] = (float)(0.25 * (double)a.p.vol[pi]);
                v[1] = a.p.vor[0][pi];
                v[2] = a.p.vor[1][pi];
                v[3] = a.p.vor[2][pi];
                v[4] = (float)btot;
                v[5] = a.p.buoy[pi];
                v[6] = a.p.hum[pi];
            }
            __syncthreads();
            // ---- order inside each cell by caller index (placement order is arbitrary) ----
            for (int c = cb0 + tid; c < cb1; c += P2G_THREADS) {
                const int lo = cell_off[c] - q0, hi = cell_off[c + 1] - q0;
                for (int q = lo; q < hi; ++q) {
                    const int oq = origk[q];
                    int rank = 0;
                    for (int q2 = lo; q2 < hi; ++q2) rank += (origk[q2] < oq);
                    ord[lo + rank] = (unsigned short)q;
                }
            }
            __syncthreads();
            // ---- every extended cell gathers its points (registers): the 3x3
            // cells of the layer around it in row-major order, their parcels in
            // caller order, the points of a parcel in order: a fixed summation
            // order that depends on the parcel set alone ----
            // scan of the 3x3 neighbourhood in that order, twice: count, then
            // fill this extended cell's bin (cheap passes); the accumulation
            // then runs over the bin only
            auto scan_nbh = [&](auto &&visit) {
                for (int r = ey - 1; r <= ey + 1; ++r) {
                    if (r < 0 || r >= P2G_TY) continue;
                    for (int cx = ex - 1; cx <= ex + 1; ++cx) {
                        if (cx < 0 || cx >= P2G_TX) continue;
                        const int c = r * P2G_TX + cx;
                        if (c < cb0 || c >= cb1) continue;
                        const int lo = cell_off[c] - q0, hi = cell_off[c + 1] - q0;
                        for (int x = lo; x < hi; ++x) {
                            const int q = ord[x];
                            const ushort4 e4 = *(const ushort4 *)(extc + q * 4);
                            if (e4.x == tid) visit(q * 4 + 0);
                            if (e4.y == tid) visit(q * 4 + 1);
                            if (e4.z == tid) visit(q * 4 + 2);
                            if (e4.w == tid) visit(q * 4 + 3);
                        }
                    }
                }
            };
            int nbin = 0;
            if (ecell) scan_nbh([&](int) { ++nbin; });
            {   // exclusive prefix of the bin sizes over the work-group
                const int lane = tid & 63, w = tid >> 6;
                int incl = nbin;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const int t = __shfl_up(incl, o, 64);
                    if (lane >= o) incl += t;
                }
                if (lane == 63) wave_tot[w] = incl;
                __syncthreads();
                int base = 0;
                for (int u = 0; u < w; ++u) base += wave_tot[u];
                if (ecell) bin_start[tid] = base + incl - nbin;
            }
            __syncthreads();   // origk is no longer read: items may be written
            if (ecell) {
                int k2 = bin_start[tid];
                scan_nbh([&](int it) { items[k2++] = (unsigned short)it; });
            }
            float acc[8][P2G_NF];
#pragma unroll
            for (int c = 0; c < 8; ++c)
#pragma unroll
                for (int f = 0; f < P2G_NF; ++f) acc[c][f] = 0.0f;
            const bool any = nbin > 0;
            if (ecell) {
                const int b0 = bin_start[tid];
                for (int x = b0; x < b0 + nbin; ++x) {
                    const int it = items[x];
                    const int q = it >> 2, jp = it & 3;
                    const float fx = frac[q * 12 + jp * 3 + 0];
                    const float fy = frac[q * 12 + jp * 3 + 1];
                    const float fz = frac[q * 12 + jp * 3 + 2];
                    const float *v = vals + q * 7;
                    const float wxy[4] = {(1.0f - fy) * (1.0f - fx), (1.0f - fy) * fx, fy * (1.0f - fx), fy * fx};
#pragma unroll
                    for (int cc = 0; cc < 8; ++cc) {
                        const float wz = (cc & 4) ? fz : 1.0f - fz;
                        const float wt = v[0] * (wz * wxy[cc & 3]);
                        acc[cc][0] += wt;
#pragma unroll
                        for (int f = 1; f < P2G_NF; ++f) acc[cc][f] += wt * v[f];
                    }
                }
            }
            // ---- parity-phased add into the node ring (equal parity: disjoint nodes) ----
            const int kz = k + ez;
            const int par = ((ex + 2) & 1) | (((ey + 2) & 1) << 1) | (((kz + 4) & 1) << 2);
            for (int ph = 0; ph < 8; ++ph) {
                __syncthreads();
                if (ecell && any && par == ph) {
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        const int nxl = ex + 1 + (c & 1);
                        const int nyl = ey + 1 + ((c >> 1) & 1);
                        const int P = kz + ((c >> 2) & 1);
                        float *dst = ring + ((P + 8) & 3) * PLANE + (nyl * P2G_BOX + nxl) * P2G_NF;
#pragma unroll
                        for (int f = 0; f < P2G_NF; ++f) dst[f] += acc[c][f];
                    }
                }
            }
            __syncthreads();
        }
        __syncthreads();
        // node plane k-1 is complete
        flush_plane(k - 1);
    }
    __syncthreads();
    flush_plane(k1 - 1);
    __syncthreads();
    flush_plane(k1);
    __syncthreads();
    flush_plane(k1 + 1);
    if (tid == 0 && s_err) atomicOr(a.err, s_err);
}

// ---------------------------------------------------------------------------
// 6. staging sum + folds + normalisation (one thread per (i, j) column)
// ---------------------------------------------------------------------------
__device__ inline void stage_sum(const FinalizeArgs &a, int i, int j, int k, float out[P2G_NF])
{
    constexpr int PLANE = P2G_BOX * P2G_BOX * P2G_NF;
    const Grid &g = a.g;
#pragma unroll
    for (int f = 0; f < P2G_NF; ++f) out[f] = 0.0f;
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
                        const float *s = a.staging + ((int64_t)wg * nbox + lp) * PLANE + (ly * P2G_BOX + lx) * P2G_NF;
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
    if (tq >= ncol * (nz + 3)) return;
    const int k = (int)(tq / ncol) - 1;
    const int64_t col = tq - (int64_t)(k + 1) * ncol;
    const int i = (int)(col % g.nx), j = (int)(col / g.nx);
    const int64_t PS = plane_stride(g), FS = field_stride(g);
    const int64_t base = (int64_t)j * g.nx + i;
    auto F = [&](int slot, int kk) -> float & { return a.fld[slot * FS + (int64_t)(kk + 1) * PS + base]; };

    // the node value of plane kk (folds of parcel_interpl.f90:258-287)
    auto folded = [&](int kk, float r[P2G_NF]) {
        float x[P2G_NF], y[P2G_NF];
        if (kk == 0 || kk == nz) {
            stage_sum(a, i, j, kk, x);
#pragma unroll
            for (int f = 0; f < P2G_NF; ++f) r[f] = 2.0f * x[f];
        } else if (kk == 1 || kk == nz - 1) {
            stage_sum(a, i, j, kk, x);
            stage_sum(a, i, j, kk == 1 ? -1 : nz + 1, y);
#pragma unroll
            for (int f = 0; f < P2G_NF; ++f) r[f] = x[f] + y[f];
        } else {
            stage_sum(a, i, j, kk, r);
        }
    };
    auto zero_zeta = [&](int kk) { return (kk == 0 && g.zeta_zero[0]) || (kk == nz && g.zeta_zero[1]); };

    if (k >= 0 && k <= nz) {
        float r[P2G_NF];
        folded(k, r);
        const float v = r[0];
        F(F_VOLG, k) = v;
        F(F_VORTG + 0, k) = r[1] / v;
        F(F_VORTG + 1, k) = r[2] / v;
        F(F_VORTG + 2, k) = zero_zeta(k) ? 0.0f : r[3] / v;
        F(F_TBUOYG, k) = r[4] / v;
        F(F_DBUOYG, k) = r[5] / v;
        F(F_HUMG, k) = r[6] / v;
    } else {
        // halo planes keep the raw sums for vol/dbuoy/hum; vortg and tbuoyg
        // are extrapolated from the normalised boundary planes (:306-318)
        float raw[P2G_NF], rb[P2G_NF], rn[P2G_NF];
        stage_sum(a, i, j, k, raw);
        const int kb = (k < 0) ? 0 : nz, kn = (k < 0) ? 1 : nz - 1;
        folded(kb, rb);
        folded(kn, rn);
        F(F_VOLG, k) = raw[0];
        F(F_DBUOYG, k) = raw[5];
        F(F_HUMG, k) = raw[6];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float vb = (c == 2 && zero_zeta(kb)) ? 0.0f : rb[1 + c] / rb[0];
            const float vn = rn[1 + c] / rn[0];
            F(F_VORTG + c, k) = 2.0f * vb - vn;
        }
        F(F_TBUOYG, k) = 2.0f * (rb[4] / rb[0]) - rn[4] / rn[0];
    }
    // parcel counts (:201-205, folds :239-243)
    auto cnt = [&](const int *c, int kk) -> int {
        return (kk >= -1 && kk <= nz) ? c[((int64_t)(kk + 1) * g.ny + j) * g.nx + i] : 0;
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
    a.nparg[(int64_t)(k + 1) * PS + base] = np;
    a.nsparg[(int64_t)(k + 1) * PS + base] = ns;
    // clear the histogram the next push fills (never the one read here)
    if (a.clear_count && k <= nz) {
        const int64_t c = ((int64_t)(k + 1) * g.ny + j) * g.nx + i;
        a.clear_count[c] = 0;
        a.clear_scount[c] = 0;
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
    int e = 0;
#pragma unroll
    for (int r = 0; r < PER_BLOCK / 256; ++r) {
        const int64_t q = (int64_t)ib.blk * PER_BLOCK + r * 256 + threadIdx.x;
        if (q >= p.n) break;
        const double B[5] = {p.B[0][q], p.B[1][q], p.B[2][q], p.B[3][q], p.B[4][q]};
        double ve[3], vt[3];
        if (!support_axes(B, (double)p.vol[q], ve, vt)) {
            e |= ERR_B33;
            ve[0] = ve[1] = ve[2] = vt[0] = vt[1] = vt[2] = 0.0;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            a.ax[d][q] = (float)ve[d];
            a.ax[3 + d][q] = (float)vt[d];
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

void add_cell_keys(Plan &P, int phase, const std::vector<KeysArgs> &items)
{
    std::vector<int64_t> n;
    for (const KeysArgs &a : items) n.push_back(a.p.n);
    P.add((const void *)k_cell_keys, 256, 0, phase, items, parcel_blocks_of(n));
}

void add_place(Plan &P, int phase, const std::vector<PlaceArgs> &items)
{
    std::vector<int64_t> n;
    for (const PlaceArgs &a : items) n.push_back(a.n);
    P.add((const void *)k_place, 256, 0, phase, items, parcel_blocks_of(n));
}

void add_axes(Plan &P, int phase, const std::vector<AxesArgs> &items)
{
    std::vector<int64_t> n;
    for (const AxesArgs &a : items) n.push_back(a.p.n);
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
    L.tz = 16;
    L.ntx = g.nx / P2G_TX;
    L.nty = g.ny / P2G_TY;
    L.nchz = (g.nz + 2 + L.tz - 1) / L.tz;
    L.staging_floats = (int64_t)L.ntx * L.nty * L.nchz * (L.tz + 3) * P2G_BOX * P2G_BOX * P2G_NF;
    return L;
}

void add_p2g(Plan &P, int phase, const std::vector<P2GArgs> &tiles, const std::vector<FinalizeArgs> &fin)
{
    std::vector<int> nb, nf;
    for (const P2GArgs &a : tiles) nb.push_back(a.ntx * a.nty * a.nchz);
    for (const FinalizeArgs &f : fin) nf.push_back((int)(((int64_t)f.g.nx * f.g.ny * (f.g.nz + 3) + 255) / 256));
    P.add((const void *)k_p2g_tiles, P2G_THREADS, 0, phase, tiles, nb);
    P.add((const void *)k_p2g_finalize, 256, 0, phase, fin, nf);
}

}  // namespace epic
