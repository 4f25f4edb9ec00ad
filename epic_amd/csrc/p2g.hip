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

__global__ void k_cell_keys(Grid g, ParcelsDev p, int *__restrict__ keys, int *__restrict__ count,
                            int *__restrict__ scount, int *__restrict__ err)
{
    int e = 0;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < p.n; q += (int64_t)gridDim.x * blockDim.x) {
        const int key = cell_key(g, p.x[q], p.y[q], p.z[q], e);
        keys[q] = key;
        if (key >= 0) {
            atomicAdd(&count[key], 1);
            if (p.vol[q] <= g.vmin) atomicAdd(&scount[key], 1);
        }
    }
    if (e) atomicOr(err, e);
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

__global__ void k_scan_reduce(const int *__restrict__ in, int64_t n, int *__restrict__ bsum)
{
    __shared__ int lds[SCAN_T];
    const int64_t base = blockIdx.x * (int64_t)SCAN_BLOCK + threadIdx.x * SCAN_PER;
    int s = 0;
#pragma unroll
    for (int u = 0; u < SCAN_PER; ++u)
        if (base + u < n) s += in[base + u];
    int total;
    block_excl_scan(s, lds, total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// single block: exclusive scan of nb block sums in place
__global__ void k_scan_blocks(int *__restrict__ bsum, int nb)
{
    __shared__ int lds[SCAN_T];
    int carry = 0;
    for (int base = 0; base < nb; base += SCAN_T * SCAN_PER) {
        const int i0 = base + threadIdx.x * SCAN_PER;
        int v[SCAN_PER], s = 0;
#pragma unroll
        for (int u = 0; u < SCAN_PER; ++u) {
            v[u] = (i0 + u < nb) ? bsum[i0 + u] : 0;
            s += v[u];
        }
        int total;
        int ex = block_excl_scan(s, lds, total);
        int run = carry + ex;
#pragma unroll
        for (int u = 0; u < SCAN_PER; ++u) {
            if (i0 + u < nb) bsum[i0 + u] = run;
            run += v[u];
        }
        carry += total;
        __syncthreads();
    }
}

__global__ void k_scan_apply(const int *__restrict__ in, int64_t n, const int *__restrict__ bsum, int *__restrict__ out)
{
    __shared__ int lds[SCAN_T];
    const int64_t base = blockIdx.x * (int64_t)SCAN_BLOCK + threadIdx.x * SCAN_PER;
    int v[SCAN_PER], s = 0;
#pragma unroll
    for (int u = 0; u < SCAN_PER; ++u) {
        v[u] = (base + u < n) ? in[base + u] : 0;
        s += v[u];
    }
    int total;
    int run = block_excl_scan(s, lds, total) + bsum[blockIdx.x];
#pragma unroll
    for (int u = 0; u < SCAN_PER; ++u) {
        if (base + u < n) out[base + u] = run;
        run += v[u];
    }
    if (base + SCAN_PER >= n && base < n + SCAN_PER) {
        // the thread owning element n-1 also writes the grand total out[n]
        if (base <= n - 1 && n - 1 < base + SCAN_PER) out[n] = run;
    }
}

void launch_exclusive_scan(const int *in, int64_t n, int *out, int *bsum, hipStream_t s)
{
    const int nb = (int)((n + SCAN_BLOCK - 1) / SCAN_BLOCK);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_T), 0, s, in, n, bsum);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(SCAN_T), 0, s, bsum, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_T), 0, s, in, n, bsum, out);
}

int64_t scan_scratch_ints(int64_t n) { return (n + SCAN_BLOCK - 1) / SCAN_BLOCK + 1; }

// ---------------------------------------------------------------------------
// 3./4. placement and per-cell ordering
// ---------------------------------------------------------------------------
// Runs of equal keys in consecutive lanes (the parcels are nearly in cell
// order) take one ticket per run.
__global__ void k_place(int64_t n, const int *__restrict__ keys, const int *__restrict__ cell_start,
                        int *__restrict__ cursor, int *__restrict__ perm_tmp)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t nround = (n + stride - 1) / stride;
    for (int64_t rnd = 0; rnd < nround; ++rnd) {   // every lane of a wave runs the same rounds
        const int64_t q = rnd * stride + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
        const int key = (q < n) ? keys[q] : -1;
        const int base = wave_run_ticket(key, cursor);
        if (key >= 0) perm_tmp[cell_start[key] + base] = (int)q;
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
struct P2GArgs {
    Grid g;
    ParcelsDev p;
    const float *axes[6];   // Veta (3), Vtau (3) per parcel, FP32
    const int *perm;
    const int *cell_start;
    float *staging;
    int *err;
    int tz;                 // cell layers per chunk
    int ntx, nty, nchz;
};

constexpr int NEXT = 300;   // extended cells of a layer: 10 x 10 x 3

__device__ inline int box_planes(int tz) { return tz + 3; }

__global__ __launch_bounds__(P2G_THREADS) void k_p2g_tiles(P2GArgs a)
{
    const Grid &g = a.g;
    const int tx = blockIdx.x % a.ntx;
    const int ty = (blockIdx.x / a.ntx) % a.nty;
    const int cz = blockIdx.x / (a.ntx * a.nty);
    const int i0 = tx * P2G_TX, j0 = ty * P2G_TY;
    const int k0 = -1 + cz * a.tz;
    const int k1 = min(k0 + a.tz, g.nz + 1);   // layers [k0, k1)
    const int tid = threadIdx.x;

    constexpr int PLANE = P2G_BOX * P2G_BOX * P2G_NF;
    constexpr int NCELL = P2G_TX * P2G_TY;
    __shared__ float ring[4 * PLANE];
    __shared__ float frac[P2G_MAXP * 12];
    __shared__ float vals[P2G_MAXP * 7];
    __shared__ unsigned short extc[P2G_MAXP * 4];  // extended cell of every support point (511: none)
    __shared__ unsigned short spos[P2G_MAXP];      // position of the parcel in caller order within the batch
    __shared__ unsigned char cell_of[P2G_MAXP];    // own cell of each batch entry
    __shared__ __attribute__((aligned(16))) unsigned short items[P2G_MAXP * 4];   // points binned by extended cell
    int *const origk = (int *)items;                // caller indices, only until the binning
    __shared__ int bin_cnt[NEXT], bin_start[NEXT + 1], bin_fill[NEXT];
    __shared__ int cell_off[NCELL + 1];            // concatenated offsets of the layer's cells
    __shared__ int row_slot[P2G_TY];
    __shared__ int row_off[P2G_TY + 1];
    __shared__ int batch_cell[NCELL + 1];          // cell-aligned batch boundaries
    __shared__ int s_nbatch, s_err;

    for (int u = tid; u < 4 * PLANE; u += P2G_THREADS) ring[u] = 0.0f;
    for (int u = tid; u < NEXT; u += P2G_THREADS) bin_cnt[u] = bin_fill[u] = 0;
    if (tid == 0) s_err = 0;

    const int nbox = box_planes(a.tz);
    float *stage = a.staging + (int64_t)blockIdx.x * nbox * PLANE;
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
            // cell-aligned batches of at most P2G_MAXP parcels
            int nb = 0, acc = 0;
            batch_cell[0] = 0;
            for (int c = 0; c < NCELL; ++c) {
                const int r = c / P2G_TX, cx = c % P2G_TX;
                const int base = ((k + 1) * g.ny + (j0 + r)) * g.nx + i0 + cx;
                const int m = a.cell_start[base + 1] - a.cell_start[base];
                cell_off[c] = row_off[r] + (a.cell_start[base] - row_slot[r]);
                if (m > P2G_MAXP) s_err |= ERR_OUT_OF_GRID;   // one cell beyond a batch: not supported
                if (acc + m > P2G_MAXP) {
                    batch_cell[++nb] = c;
                    acc = 0;
                }
                acc += m;
            }
            cell_off[NCELL] = row_off[P2G_TY];
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
                const double ve[3] = {a.axes[0][pi], a.axes[1][pi], a.axes[2][pi]};
                const double vt[3] = {a.axes[3][pi], a.axes[4][pi], a.axes[5][pi]};
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
                    spos[q] = (unsigned short)(lo + rank);
                }
            }
            __syncthreads();
            // ---- bin the support points by extended cell ----
            for (int u = tid; u < nb * 4; u += P2G_THREADS) {
                const int e = extc[u];
                if (e < NEXT) atomicAdd(&bin_cnt[e], 1);
            }
            __syncthreads();
            if (tid < 64) {   // exclusive scan of 300 counts by one wavefront
                int run = 0;
                for (int base = 0; base < NEXT; base += 64) {
                    const int e = base + tid;
                    const int v = (e < NEXT) ? bin_cnt[e] : 0;
                    int incl = v;
#pragma unroll
                    for (int o = 1; o < 64; o <<= 1) {
                        const int t = __shfl_up(incl, o, 64);
                        if (tid >= o) incl += t;
                    }
                    if (e < NEXT) bin_start[e] = run + incl - v;
                    run += __shfl(incl, 63, 64);
                }
                if (tid == 0) bin_start[NEXT] = run;
            }
            __syncthreads();
            for (int u = tid; u < nb * 4; u += P2G_THREADS) {
                const int e = extc[u];
                if (e < NEXT) items[bin_start[e] + atomicAdd(&bin_fill[e], 1)] = (unsigned short)u;
            }
            __syncthreads();
            // sort every bin by (caller order, point): the accumulation order is a
            // function of the parcel set alone
            if (ecell) {
                const int lo = bin_start[tid], hi = bin_start[tid + 1];
                for (int x = lo + 1; x < hi; ++x) {
                    const unsigned short it = items[x];
                    const int key = spos[it >> 2] * 4 + (it & 3);
                    int y = x - 1;
                    while (y >= lo && (spos[items[y] >> 2] * 4 + (items[y] & 3)) > key) {
                        items[y + 1] = items[y];
                        --y;
                    }
                    items[y + 1] = it;
                }
            }
            __syncthreads();
            // ---- every extended cell accumulates its points (registers) ----
            float acc[8][P2G_NF];
#pragma unroll
            for (int c = 0; c < 8; ++c)
#pragma unroll
                for (int f = 0; f < P2G_NF; ++f) acc[c][f] = 0.0f;
            bool any = false;
            if (ecell) {
                const int lo = bin_start[tid], hi = bin_start[tid + 1];
                any = hi > lo;
                for (int x = lo; x < hi; ++x) {
                    const int it = items[x];
                    const int q = it >> 2, jp = it & 3;
                    const float fx = frac[q * 12 + jp * 3 + 0];
                    const float fy = frac[q * 12 + jp * 3 + 1];
                    const float fz = frac[q * 12 + jp * 3 + 2];
                    const float *v = vals + q * 7;
                    const float wxy[4] = {(1.0f - fy) * (1.0f - fx), (1.0f - fy) * fx, fy * (1.0f - fx), fy * fx};
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        const float wz = (c & 4) ? fz : 1.0f - fz;
                        const float wt = v[0] * (wz * wxy[c & 3]);
                        acc[c][0] += wt;
#pragma unroll
                        for (int f = 1; f < P2G_NF; ++f) acc[c][f] += wt * v[f];
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
            for (int u = tid; u < NEXT; u += P2G_THREADS) bin_cnt[u] = bin_fill[u] = 0;
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
struct FinalizeArgs {
    Grid g;
    const float *staging;
    const int *count, *scount;
    float *fld;     // F_VOLG .. F_HUMG
    int *nparg, *nsparg;
    int tz, ntx, nty, nchz;
};

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

__global__ void k_p2g_finalize(FinalizeArgs a)
{
    const Grid &g = a.g;
    const int64_t col = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (col >= (int64_t)g.nx * g.ny) return;
    const int i = (int)(col % g.nx), j = (int)(col / g.nx);
    const int nz = g.nz;
    const int64_t PS = plane_stride(g), FS = field_stride(g);
    const int64_t base = (int64_t)j * g.nx + i;
    auto F = [&](int slot, int k) -> float & { return a.fld[slot * FS + (int64_t)(k + 1) * PS + base]; };

    float lo[3][P2G_NF], hi[3][P2G_NF];   // raw planes -1,0,1 and nz-1,nz,nz+1
    for (int k = -1; k <= nz + 1; ++k) {
        float r[P2G_NF];
        stage_sum(a, i, j, k, r);
        if (k <= 1) {
#pragma unroll
            for (int f = 0; f < P2G_NF; ++f) lo[k + 1][f] = r[f];
        } else if (k >= nz - 1) {
#pragma unroll
            for (int f = 0; f < P2G_NF; ++f) hi[k - (nz - 1)][f] = r[f];
        } else {
            const float v = r[0];
            F(F_VOLG, k) = v;
            F(F_VORTG + 0, k) = r[1] / v;
            F(F_VORTG + 1, k) = r[2] / v;
            F(F_VORTG + 2, k) = r[3] / v;
            F(F_TBUOYG, k) = r[4] / v;
            F(F_DBUOYG, k) = r[5] / v;
            F(F_HUMG, k) = r[6] / v;
        }
    }
    // free-slip folds (parcel_interpl.f90:258-287)
    float p0[P2G_NF], p1[P2G_NF], pn[P2G_NF], pn1[P2G_NF];
#pragma unroll
    for (int f = 0; f < P2G_NF; ++f) {
        p0[f] = 2.0f * lo[1][f];
        p1[f] = lo[2][f] + lo[0][f];
        pn[f] = 2.0f * hi[1][f];
        pn1[f] = hi[0][f] + hi[2][f];
    }
    auto store_norm = [&](int k, const float *r, bool zero_zeta) {
        const float v = r[0];
        F(F_VOLG, k) = v;
        F(F_VORTG + 0, k) = r[1] / v;
        F(F_VORTG + 1, k) = r[2] / v;
        F(F_VORTG + 2, k) = zero_zeta ? 0.0f : r[3] / v;
        F(F_TBUOYG, k) = r[4] / v;
        F(F_DBUOYG, k) = r[5] / v;
        F(F_HUMG, k) = r[6] / v;
    };
    store_norm(0, p0, g.zeta_zero[0] != 0);
    store_norm(1, p1, false);
    store_norm(nz - 1, pn1, false);
    store_norm(nz, pn, g.zeta_zero[1] != 0);
    // halo planes keep the raw sums for vol/dbuoy/hum; vortg and tbuoyg are
    // extrapolated (:306-318)
    const int halo[2] = {-1, nz + 1};
    const float *raw[2] = {lo[0], hi[2]};
    for (int h = 0; h < 2; ++h) {
        const int k = halo[h];
        F(F_VOLG, k) = raw[h][0];
        F(F_DBUOYG, k) = raw[h][5];
        F(F_HUMG, k) = raw[h][6];
        const int kb = (h == 0) ? 0 : nz, kn = (h == 0) ? 1 : nz - 1;
        for (int c = 0; c < 3; ++c) F(F_VORTG + c, k) = 2.0f * F(F_VORTG + c, kb) - F(F_VORTG + c, kn);
        F(F_TBUOYG, k) = 2.0f * F(F_TBUOYG, kb) - F(F_TBUOYG, kn);
    }
    // parcel counts (:201-205, folds :239-243)
    auto cnt = [&](const int *c, int k) -> int {
        return (k >= -1 && k <= nz) ? c[((int64_t)(k + 1) * g.ny + j) * g.nx + i] : 0;
    };
    for (int k = -1; k <= nz + 1; ++k) {
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
    }
}

// ---------------------------------------------------------------------------
// support-point half axes, once per parcel per stage (FP64 eigen-solve)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_axes(Grid g, ParcelsDev p, float *ax0, float *ax1, float *ax2, float *ax3,
                                              float *ax4, float *ax5, int *err)
{
    int e = 0;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < p.n; q += (int64_t)gridDim.x * blockDim.x) {
        const double B[5] = {p.B[0][q], p.B[1][q], p.B[2][q], p.B[3][q], p.B[4][q]};
        double ve[3], vt[3];
        if (!support_axes(B, (double)p.vol[q], ve, vt)) {
            e |= ERR_B33;
            ve[0] = ve[1] = ve[2] = vt[0] = vt[1] = vt[2] = 0.0;
        }
        ax0[q] = (float)ve[0];
        ax1[q] = (float)ve[1];
        ax2[q] = (float)ve[2];
        ax3[q] = (float)vt[0];
        ax4[q] = (float)vt[1];
        ax5[q] = (float)vt[2];
    }
    if (e) atomicOr(err, e);
}

void launch_axes(const Grid &g, const ParcelsDev &p, float *const axes[6], int *err, hipStream_t s)
{
    if (p.n == 0) return;
    const int nb = (int)std::min<int64_t>((p.n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_axes, dim3(nb), dim3(256), 0, s, g, p, axes[0], axes[1], axes[2], axes[3], axes[4], axes[5],
                       err);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
void launch_cell_keys(const Grid &g, const ParcelsDev &p, int *keys, int *count, int *scount, int *err, hipStream_t s)
{
    if (p.n == 0) return;
    const int nb = (int)std::min<int64_t>((p.n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_cell_keys, dim3(nb), dim3(256), 0, s, g, p, keys, count, scount, err);
}

void launch_place(int64_t n, const int *keys, const int *cell_start, int *cursor, int *perm_tmp, hipStream_t s)
{
    if (n == 0) return;
    const int nb = (int)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_place, dim3(nb), dim3(256), 0, s, n, keys, cell_start, cursor, perm_tmp);
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

void launch_p2g(const Grid &g, const ParcelsDev &p, const float *const axes[6], const int *perm,
                const int *cell_start, const int *count, const int *scount, float *staging, float *fld, int *nparg,
                int *nsparg, int *err, hipStream_t s)
{
    const P2GLayout L = p2g_layout(g);
    P2GArgs a;
    a.g = g;
    a.p = p;
    for (int c = 0; c < 6; ++c) a.axes[c] = axes[c];
    a.perm = perm;
    a.cell_start = cell_start;
    a.staging = staging;
    a.err = err;
    a.tz = L.tz;
    a.ntx = L.ntx;
    a.nty = L.nty;
    a.nchz = L.nchz;
    hipLaunchKernelGGL(k_p2g_tiles, dim3(L.ntx * L.nty * L.nchz), dim3(P2G_THREADS), 0, s, a);
    FinalizeArgs f;
    f.g = g;
    f.staging = staging;
    f.count = count;
    f.scount = scount;
    f.fld = fld;
    f.nparg = nparg;
    f.nsparg = nsparg;
    f.tz = L.tz;
    f.ntx = L.ntx;
    f.nty = L.nty;
    f.nchz = L.nchz;
    const int64_t cols = (int64_t)g.nx * g.ny;
    hipLaunchKernelGGL(k_p2g_finalize, dim3((unsigned)((cols + 127) / 128)), dim3(128), 0, s, f);
}

}  // namespace epic
