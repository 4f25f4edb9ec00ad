// Parcel merging (parcel_merge.f90, parcel_nearest.f90) on the device, with
// the semantics of the reference on one MPI rank.
//
// The reference's index of a parcel is its caller index (orig), whatever its
// physical slot (a re-sort puts the slots in cell order): every order below
// is the caller-index order, and the delete relabels caller indices as the
// reference moves parcels, while the slots are compacted on their own.
//
// 1. small parcels: volume < vmin (is_small, parcel_ellipsoid.f90:452-459),
//    compacted in caller-index order
// 2. periodic images: a small parcel whose nearest grid index is 0 or n in x
//    or y is sent to the rank itself (locate_parcel_in_boundary_cell,
//    parcel_nearest.f90:1384-1460; apply_periodic_shift :564-598); images are
//    numbered in the receive order of the directions (mpi_tags.f90), each
//    direction in small-parcel order
// 3. halo-cell binning of parcels and images (:196-257)
// 4. nearest other parcel of every small parcel and image over the 8 cells
//    around its nearest grid point (:608-662), the first minimum in the
//    reference's visiting order; image results return to their source
//    when strictly closer (:746-932)
// 5. tree resolution (:937-1377): the leaf iterations run in parallel (a
//    merging leaf is never a target, so each iteration is order free); the
//    few dual-link candidates left are decided by one thread in small-parcel
//    order, as the reference's sequential loop does
// 6. geometric merge (parcel_merge.f90:129-352) per target in FP64, summed
//    in small-parcel order
// 7. delete (parcel_container.f90:413-452): holes filled from the end, in
//    caller indices; the slots are compacted the same way
// These kernels are launched directly (the merge is not part of the step).
#include "common.h"
#include "ellipsoid.h"
#include "launch.h"

#include <algorithm>
#include <vector>

namespace epic {

namespace {

__device__ inline double fnint(double v) { return round(v); }   // Fortran nint: half away from zero

__device__ inline void edge_shift(const Grid &g, double &x, double &y)
{   // handle_periodic_edge_parcels (parcel_nearest.f90:172-188), one rank
    const int g0 = (int)floor((x - g.lower[0]) * g.dxi[0]);
    const int g1 = (int)floor((y - g.lower[1]) * g.dxi[1]);
    if (g0 < -1) x += g.extent[0];
    if (g0 > g.nx) x -= g.extent[0];
    if (g1 < -1) y += g.extent[1];
    if (g1 > g.ny) y -= g.extent[1];
}

__device__ inline int64_t gtid() { return blockIdx.x * (int64_t)blockDim.x + threadIdx.x; }

// receive order of the image directions: boundary list (x: 1 ix == nx,
// 0 ix == 0, -1 any; y likewise) and shift
__constant__ int c_lst_x[8] = {-1, -1, 1, 0, 1, 0, 1, 0};
__constant__ int c_lst_y[8] = {0, 1, -1, -1, 0, 0, 1, 1};
__constant__ int c_sh_x[8] = {0, 0, -1, 1, -1, 1, -1, 1};
__constant__ int c_sh_y[8] = {1, -1, 0, 0, 1, 1, -1, -1};
// order in which the image results return to their source
__constant__ int c_back[8] = {1, 0, 3, 2, 7, 6, 5, 4};

__global__ void k_small(Grid g, ParcelsDev p, double vmin, int *flag)
{
    const int64_t i = gtid();
    if (i >= p.n) return;
    double x = p.x[i], y = p.y[i];
    edge_shift(g, x, y);
    p.x[i] = x;
    p.y[i] = y;
    flag[i] = ((double)p.vol[i] < vmin) ? 1 : 0;
}

__global__ void k_compact(int64_t n, const int *flag, const int *rank, int *out)
{
    const int64_t i = gtid();
    if (i < n && flag[i]) out[rank[i]] = (int)i;
}

// out[orig[i]] = in[i]: a per-slot array in caller-index order
__global__ void k_to_caller(int64_t n, const int *orig, const int *in, int *out)
{
    const int64_t i = gtid();
    if (i < n) out[orig[i]] = in[i];
}

// the flagged slots listed in caller-index order: out[crank[orig[i]]] = i
__global__ void k_compact_caller(int64_t n, const int *flag, const int *orig, const int *crank, int *out)
{
    const int64_t i = gtid();
    if (i < n && flag[i]) out[crank[orig[i]]] = (int)i;
}

__global__ void k_image_flags(Grid g, ParcelsDev p, const int *sml, int ns, int *imgflag)
{
    const int64_t t = gtid();
    if (t >= 8 * (int64_t)ns) return;
    const int dir = (int)(t / ns), m = (int)(t % ns);
    const int i = sml[m];
    const int ix = (int)fnint(g.dxi[0] * (p.x[i] - g.lower[0]));
    const int iy = (int)fnint(g.dxi[1] * (p.y[i] - g.lower[1]));
    const int lx = c_lst_x[dir], ly = c_lst_y[dir];
    const bool inx = (lx < 0) || (lx == 0 && ix == 0) || (lx == 1 && ix == g.nx);
    const bool iny = (ly < 0) || (ly == 0 && iy == 0) || (ly == 1 && iy == g.ny);
    imgflag[t] = (inx && iny) ? 1 : 0;
}

__global__ void k_images(Grid g, ParcelsDev p, const int *sml, int ns, const int *imgflag, const int *imgrank,
                         double *ix_, double *iy_, double *iz_, int *ipid, int *imid, int *img_of)
{
    const int64_t t = gtid();
    if (t >= 8 * (int64_t)ns) return;
    if (!imgflag[t]) {
        img_of[t] = -1;
        return;
    }
    const int dir = (int)(t / ns), m = (int)(t % ns);
    const int i = sml[m];
    const int k = imgrank[t];
    const double x = p.x[i], y = p.y[i];
    const int ix = (int)fnint(g.dxi[0] * (x - g.lower[0]));
    const int iy = (int)fnint(g.dxi[1] * (y - g.lower[1]));
    const bool ex = (ix <= 0) || (ix >= g.nx), ey = (iy <= 0) || (iy >= g.ny);
    double qx = x + ((c_sh_x[dir] != 0 && ex) ? c_sh_x[dir] * g.extent[0] : 0.0);
    double qy = y + ((c_sh_y[dir] != 0 && ey) ? c_sh_y[dir] * g.extent[1] : 0.0);
    edge_shift(g, qx, qy);
    ix_[k] = qx;
    iy_[k] = qy;
    iz_[k] = p.z[i];
    ipid[k] = i;
    imid[k] = m;
    img_of[t] = k;
}

struct Pts {   // parcels then images
    const double *x, *y, *z;
    const double *ix, *iy, *iz;
    int64_t n;
    __device__ void get(int64_t e, double &a, double &b, double &c) const
    {
        if (e < n) {
            a = x[e];
            b = y[e];
            c = z[e];
        } else {
            a = ix[e - n];
            b = iy[e - n];
            c = iz[e - n];
        }
    }
};

__device__ inline int64_t halo_cell(const Grid &g, double x, double y, double z)
{
    const int jx = (int)(g.dxi[0] * (x - (g.lower[0] - g.dx[0])));
    const int jy = (int)(g.dxi[1] * (y - (g.lower[1] - g.dx[1])));
    int jz = (int)(g.dxi[2] * (z - g.lower[2]));
    if (jz > g.nz - 1) jz = g.nz - 1;
    return (int64_t)jx + (int64_t)(g.nx + 3) * jy + (int64_t)(g.nx + 3) * (g.ny + 3) * jz;
}

__global__ void k_bin_keys(Grid g, Pts P, int64_t ntot, int64_t ncell, int *key, int *cnt, int *err)
{
    const int64_t e = gtid();
    int k = -1;   // lanes past the end take part in the wave's histogram add with no key
    if (e < ntot) {
        double x, y, z;
        P.get(e, x, y, z);
        const int64_t c = halo_cell(g, x, y, z);
        if (c < 0 || c >= ncell) atomicOr(err, ERR_OUT_OF_GRID);
        else k = (int)c;
        key[e] = k;
    }
    wave_run_add(k, cnt);   // one add per run of equal keys (parcels mostly in cell order)
}

// nearest other parcel; first minimum in the reference's visiting order
// (cells iz, iy, ix ascending; inside a cell ascending index: the caller
// index of a parcel, images after every parcel in image order).  Eight
// lanes per small parcel or image, lane cs = 4 dz + 2 dy + dx taking the
// cell (iz0-1+dz, iy0-1+dy, ix0-1+dx), which numbers the cells in that
// visiting order: each lane finds its cell's minimum, and the eight combine
// to the smallest distance, the earlier cell on a tie, as the reference's
// strict comparison over cells keeps the first.
constexpr int SEARCH_LANES = 8;
__global__ __launch_bounds__(256) void k_search(Grid g, Pts P, const int *orig, const int *sml, int ns, int nimg,
                                                const int *start, const int *node, const int *ipid, int *iclo,
                                                double *dclo)
{
    const int64_t t = gtid();
    const int64_t m = t / SEARCH_LANES;   // the 8 lanes of one search are neighbours in a wave
    const int cs = (int)(t % SEARCH_LANES);
    const bool valid = m < (int64_t)ns + nimg;
    double bd = HUGE_VAL;
    int64_t bj = -1;
    int64_t is = -1;
    if (valid) {
        is = (m < ns) ? sml[m] : P.n + (m - ns);
        double xs, ys, zs;
        P.get(is, xs, ys, zs);
        const int nx = g.nx, ny = g.ny, nz = g.nz;
        const int ix0 = (int)fnint(g.dxi[0] * (xs - (g.lower[0] - g.dx[0])));
        const int iy0 = (int)fnint(g.dxi[1] * (ys - (g.lower[1] - g.dx[1])));
        const int iz0 = (int)fnint(g.dxi[2] * (zs - g.lower[2]));
        const int iz = iz0 - 1 + (cs >> 2), iy = iy0 - 1 + ((cs >> 1) & 1), ix = ix0 - 1 + (cs & 1);
        if (iz >= 0 && iz <= nz - 1 && iy >= 0 && iy <= ny + 1 && ix >= 0 && ix <= nx + 1) {
            const int64_t c = (int64_t)ix + (int64_t)(nx + 3) * iy + (int64_t)(nx + 3) * (ny + 3) * iz;
            // the cell's minimum under (distance, reference index), which does
            // not depend on the order the entries are taken in: four entries'
            // loads are issued together, and the reference index (a further
            // dependent load) is read only on a tie of distances
            auto ref = [&](int64_t j) { return (j < P.n) ? (int64_t)orig[j] : j; };
            auto take = [&](int64_t j, double xj, double yj, double zj) {
                if (j == is) return;
                const double delz = zj - zs;
                const double delx = xj - xs, dely = yj - ys;
                const double dsq = delx * delx + dely * dely + delz * delz;
                if (bj < 0 || dsq < bd || (dsq == bd && ref(j) < ref(bj))) {
                    bd = dsq;
                    bj = j;
                }
            };
            int k = start[c];
            const int kend = start[c + 1];
            for (; k + 4 <= kend; k += 4) {
                int64_t j4[4];
                double x4[4], y4[4], z4[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) j4[u] = node[k + u];
#pragma unroll
                for (int u = 0; u < 4; ++u) P.get(j4[u], x4[u], y4[u], z4[u]);
#pragma unroll
                for (int u = 0; u < 4; ++u) take(j4[u], x4[u], y4[u], z4[u]);
            }
            for (; k < kend; ++k) {
                const int64_t j = node[k];
                double xj, yj, zj;
                P.get(j, xj, yj, zj);
                take(j, xj, yj, zj);
            }
        }
    }
    // the eight lanes' minima: the smallest distance, then the earlier cell
    int bc = cs;
#pragma unroll
    for (int o = 1; o < SEARCH_LANES; o <<= 1) {
        const double od = __shfl_xor(bd, o, 64);
        const int64_t oj = __shfl_xor(bj, o, 64);
        const int oc = __shfl_xor(bc, o, 64);
        if (od < bd || (od == bd && oc < bc)) {
            bd = od;
            bj = oj;
            bc = oc;
        }
    }
    if (!valid || cs != 0) return;
    double dmin = g.extent[0] * g.extent[0] + g.extent[1] * g.extent[1] + g.extent[2] * g.extent[2];
    int64_t ic = -1;
    if (bj >= 0 && bd < dmin) {
        dmin = bd;
        ic = bj;
    }
    // update_remote_indices (parcel_nearest.f90:746-788)
    if (is >= P.n && ic >= P.n) dmin = HUGE_VAL;
    else if (ic >= P.n) ic = ipid[ic - P.n];
    iclo[m] = (int)ic;
    dclo[m] = dmin;
}

__global__ void k_reconcile(int ns, const int *img_of, int *iclo, double *dclo, int *err)
{
    const int64_t m = gtid();
    if (m >= ns) return;
    double d = dclo[m];
    int c = iclo[m];
    for (int o = 0; o < 8; ++o) {
        const int k = img_of[(int64_t)c_back[o] * ns + m];
        if (k < 0) continue;
        const double dk = dclo[ns + k];
        if (d > dk) {
            d = dk;
            c = iclo[ns + k];
        }
    }
    dclo[m] = d;
    iclo[m] = c;
    if (c < 0) atomicOr(err, ERR_MERGE);
}

// ---- tree resolution, first (iterative) stage ----
__global__ void k_tree1(const int *sml, int ns, const int *iclo, const unsigned char *merged, unsigned char *leaf,
                        unsigned char *avail)
{
    const int64_t m = gtid();
    if (m >= ns || merged[sml[m]]) return;
    leaf[sml[m]] = 1;
    avail[iclo[m]] = 1;
}
__global__ void k_tree2(const int *sml, int ns, const int *iclo, const unsigned char *merged, unsigned char *leaf)
{
    const int64_t m = gtid();
    if (m >= ns || merged[sml[m]]) return;
    leaf[iclo[m]] = 0;
}
__global__ void k_tree3(const int *sml, int ns, const int *iclo, const unsigned char *merged,
                        const unsigned char *leaf, unsigned char *avail)
{
    const int64_t m = gtid();
    if (m >= ns) return;
    const int is = sml[m];
    if (!merged[is] && !leaf[is]) avail[iclo[m]] = 0;
}
__global__ void k_tree4(const int *sml, int ns, const int *iclo, unsigned char *merged, const unsigned char *leaf,
                        const unsigned char *avail, int *cont)
{
    const int64_t m = gtid();
    if (m >= ns) return;
    const int is = sml[m], ic = iclo[m];
    if (!merged[is] && leaf[is] && avail[ic]) {
        merged[is] = 1;
        merged[ic] = 1;
        *cont = 1;
    }
}

// ---- second stage ----
__global__ void k_tree5(const int *sml, int ns, const int *iclo, const unsigned char *merged,
                        const unsigned char *leaf, unsigned char *avail)
{
    const int64_t m = gtid();
    if (m >= ns) return;
    const int is = sml[m];
    if (!merged[is] && leaf[is]) avail[iclo[m]] = 1;
}
// do-merge of the order-free cases; candidates (unmerged, no leaf, not
// available) are left for the sequential pass
__global__ void k_tree6(const int *sml, int ns, const unsigned char *merged, const unsigned char *leaf,
                        const unsigned char *avail, int *dom, int *cand)
{
    const int64_t m = gtid();
    if (m >= ns) return;
    const int is = sml[m];
    dom[m] = (leaf[is] != 0) ? 1 : 0;   // merged & leaf, or unmerged leaf
    cand[m] = (!merged[is] && !leaf[is] && !avail[is]) ? 1 : 0;
}
// The reference's sequential loop over the candidates in small-parcel order
// (parcel_nearest.f90, second stage of resolve_tree): candidate m merges if
// its target iclo[m] is available, else it is an isolated dual link and
// makes its own parcel available.  The only availability the loop changes is
// a candidate's own parcel, so m's decision depends on the candidate p that
// owns iclo[m], and only when p comes first: then iclo[m] is available
// exactly when p was isolated.  Each candidate therefore follows that chain
// (whose positions strictly decrease) to its first link to a parcel that is
// not an earlier candidate, and flips the answer there once per step: the
// decisions of the sequential loop, in parallel.  cpos[slot] is the
// candidate position of a candidate's own parcel, -1 for every other slot.
__global__ void k_cand_pos(int ncand, const int *clist, const int *sml, int *cpos)
{
    const int64_t u = gtid();
    if (u < ncand) cpos[sml[clist[u]]] = (int)u;
}
__global__ void k_tree7(const int *sml, const int *clist, int ncand, const int *iclo, const int *cpos,
                        unsigned char *avail, int *dom, int *iso)
{
    const int64_t u0 = gtid();
    if (u0 >= ncand) return;
    int64_t u = u0;
    bool flip = false, val;
    for (;;) {
        const int ic = iclo[clist[u]];
        const int q = cpos[ic];
        if (q < 0) {         // not a candidate's parcel: its availability is fixed during the loop
            val = avail[ic] != 0;
            break;
        }
        if (q >= u) {        // a later candidate's parcel: not yet made available
            val = false;
            break;
        }
        flip = !flip;        // available exactly when that earlier candidate was isolated
        u = q;
    }
    const bool merges = flip ? !val : val;
    const int m = clist[u0];
    if (merges) dom[m] = 1;
    else {
        iso[m] = 1;
        avail[sml[m]] = 1;   // a candidate's own parcel: no other thread reads it above
    }
}
__global__ void k_tree8(int ns, const int *iclo, const unsigned char *avail, const int *iso, int *dom)
{
    const int64_t m = gtid();
    if (m >= ns) return;
    if (!dom[m] && iso[m] && avail[iclo[m]]) dom[m] = 1;
}

__global__ void k_pairs(int ns, const int *dom, const int *rank, const int *sml, const int *iclo, int *pis, int *pic,
                        int *gcnt)
{
    const int64_t m = gtid();
    if (m >= ns || !dom[m]) return;
    const int k = rank[m];
    pis[k] = sml[m];
    pic[k] = iclo[m];
    atomicAdd(&gcnt[iclo[m]], 1);
}

__global__ void k_group_place(int nm, const int *pic, const int *gstart, int *gcur, int *gent)
{
    const int64_t k = gtid();
    if (k >= nm) return;
    const int ic = pic[k];
    gent[gstart[ic] + atomicAdd(&gcur[ic], 1)] = (int)k;
}

// collect_merge_stats (parcel_merge.f90:365-393), before the volumes change:
// cnt[0] counts the pairs whose target is big, cnt[j] the targets that
// min(7, their number of pairs) = j
__global__ void k_merge_stats_big(int nm, const int *pic, ParcelsDev p, double vmin, int *cnt)
{
    const int64_t k = gtid();
    if (k < nm && (double)p.vol[pic[k]] > vmin) atomicAdd(&cnt[0], 1);
}
// The histogram's bins are few and hot: each wave counts its lanes per bin
// with ballots over a grid-stride loop, the block sums its waves in LDS, and
// one atomic per block and bin reaches global memory (an atomic per target
// on 7 addresses serialises millions of updates).
constexpr int STATS_WAY_BLOCKS = 1024;
__global__ __launch_bounds__(256) void k_merge_stats_way(int64_t n, const int *gcnt, int *cnt)
{
    __shared__ int h[8];
    if (threadIdx.x < 8) h[threadIdx.x] = 0;
    __syncthreads();
    int w[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // wave-uniform counts
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = base + threadIdx.x;
        const int c = (i < n) ? gcnt[i] : 0;
        const int b = c < 7 ? c : 7;
#pragma unroll
        for (int j = 1; j < 8; ++j) w[j] += __popcll(__ballot(b == j));
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int j = 1; j < 8; ++j)
            if (w[j]) atomicAdd(&h[j], w[j]);
    }
    __syncthreads();
    if (threadIdx.x >= 1 && threadIdx.x < 8 && h[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], h[threadIdx.x]);
}

__device__ inline double pdelta(double d, double L) { return d - L * round(d * (1.0 / L)); }

// one thread per target parcel with mergers: do_group_merge + geometric_merge
__global__ void k_group_merge(Grid g, ParcelsDev p, const int *gstart, int *gent, const int *pis, int *err)
{
    const int64_t ic = gtid();
    if (ic >= p.n) return;
    const int s0 = gstart[ic], s1 = gstart[ic + 1];
    if (s1 == s0) return;
    // the group's pairs in small-parcel order
    for (int a = s0 + 1; a < s1; ++a) {
        const int v = gent[a];
        int b = a - 1;
        while (b >= s0 && gent[b] > v) {
            gent[b + 1] = gent[b];
            --b;
        }
        gent[b + 1] = v;
    }
    const double vic = p.vol[ic];
    double vm = vic;
    const double x0 = p.x[ic], y0 = p.y[ic];
    double px = 0.0, py = 0.0, pz = vic * p.z[ic];
    double bm = vic * (double)p.buoy[ic], hm = vic * (double)p.hum[ic];
    double wm[3] = {vic * (double)p.vor[0][ic], vic * (double)p.vor[1][ic], vic * (double)p.vor[2][ic]};
    for (int a = s0; a < s1; ++a) {
        const int is = pis[gent[a]];
        double xs = p.x[is], ys = p.y[is];
        xs = xs - g.extent[0] * trunc((xs - g.center[0]) * g.hli[0]);   // apply_periodic_bc on the small parcel
        ys = ys - g.extent[1] * trunc((ys - g.center[1]) * g.hli[1]);
        p.x[is] = xs;
        p.y[is] = ys;
        const double vs = p.vol[is];
        vm = vm + vs;
        px = px + vs * pdelta(xs - x0, g.extent[0]);
        py = py + vs * pdelta(ys - y0, g.extent[1]);
        pz = pz + vs * p.z[is];
        bm = bm + vs * (double)p.buoy[is];
        hm = hm + vs * (double)p.hum[is];
        for (int c = 0; c < 3; ++c) wm[c] = wm[c] + vs * (double)p.vor[c][is];
    }
    const double vi = 1.0 / vm;
    double cx = vi * px, cy = vi * py;
    cx = x0 + cx;
    cy = y0 + cy;
    const double cz = vi * pz;
    cx = cx - g.extent[0] * trunc((cx - g.center[0]) * g.hli[0]);
    cy = cy - g.extent[1] * trunc((cy - g.center[1]) * g.hli[1]);
    bm = vi * bm;
    hm = vi * hm;
    for (int c = 0; c < 3; ++c) wm[c] = vi * wm[c];
    // second moments about the merged centre
    double Bm[6];
    {
        const double B[5] = {p.B[0][ic], p.B[1][ic], p.B[2][ic], p.B[3][ic], p.B[4][ic]};
        double b33;
        if (!b33_of(B, vic, b33)) {
            atomicOr(err, ERR_B33);
            return;
        }
        const double dx = pdelta(x0 - cx, g.extent[0]), dy = pdelta(y0 - cy, g.extent[1]), dz = p.z[ic] - cz;
        const double mu = vic * vi;
        Bm[0] = mu * (5.0 * (dx * dx) + B[0]);
        Bm[1] = mu * (5.0 * dx * dy + B[1]);
        Bm[2] = mu * (5.0 * dx * dz + B[2]);
        Bm[3] = mu * (5.0 * (dy * dy) + B[3]);
        Bm[4] = mu * (5.0 * dy * dz + B[4]);
        Bm[5] = mu * (5.0 * (dz * dz) + b33);
    }
    for (int a = s0; a < s1; ++a) {
        const int is = pis[gent[a]];
        const double vs = p.vol[is];
        const double B[5] = {p.B[0][is], p.B[1][is], p.B[2][is], p.B[3][is], p.B[4][is]};
        double b33;
        if (!b33_of(B, vs, b33)) {
            atomicOr(err, ERR_B33);
            return;
        }
        const double dx = pdelta(p.x[is] - cx, g.extent[0]), dy = pdelta(p.y[is] - cy, g.extent[1]);
        const double dz = p.z[is] - cz;
        const double mu = vi * vs;
        Bm[0] = Bm[0] + mu * (5.0 * (dx * dx) + B[0]);
        Bm[1] = Bm[1] + mu * (5.0 * dx * dy + B[1]);
        Bm[2] = Bm[2] + mu * (5.0 * dx * dz + B[2]);
        Bm[3] = Bm[3] + mu * (5.0 * (dy * dy) + B[3]);
        Bm[4] = Bm[4] + mu * (5.0 * dy * dz + B[4]);
        Bm[5] = Bm[5] + mu * (5.0 * (dz * dz) + b33);
    }
    const double detB = Bm[0] * (Bm[3] * Bm[5] - Bm[4] * Bm[4]) - Bm[1] * (Bm[1] * Bm[5] - Bm[2] * Bm[4]) +
                        Bm[2] * (Bm[1] * Bm[4] - Bm[2] * Bm[3]);
    const double abc = 0.75 * vm * (1.0 / 3.14159265358979323846);
    const double factor = pow(abc * abc / detB, 1.0 / 3.0);
    for (int c = 0; c < 5; ++c) p.B[c][ic] = (float)(Bm[c] * factor);
    p.vol[ic] = (float)vm;
    p.x[ic] = cx;
    p.y[ic] = cy;
    p.z[ic] = cz;
    p.buoy[ic] = (float)bm;
    p.hum[ic] = (float)hm;
    for (int c = 0; c < 3; ++c) p.vor[c][ic] = (float)wm[c];
}

// delete (parcel_container.f90:413-452), in caller indices: hole k (the
// k-th smallest merged caller index below n - nm; the merged parcels pis[]
// are in caller order) takes the k-th valid caller index counted from the
// end.  The same rule on slots compacts the arrays: the k-th smallest
// merged slot below n - nm takes the k-th valid slot from the end.
__global__ void k_mark_holes(int nm, const int *pis, const int *orig, int *hole_c, int *hole_s)
{
    const int64_t k = gtid();
    if (k >= nm) return;
    hole_c[orig[pis[k]]] = 1;
    hole_s[pis[k]] = 1;
}
__global__ void k_tail_flags(int64_t n, int nm, const int *hole, int *tflag)
{   // tail positions from the top: t = n - 1 - u
    const int64_t u = gtid();
    if (u >= nm) return;
    tflag[u] = hole[n - 1 - u] ? 0 : 1;
}
// slot of every tail caller index t >= n - nm: tail_slot[t - (n - nm)]
__global__ void k_tail_slots(int64_t n, int nm, const int *orig, int *tail_slot)
{
    const int64_t i = gtid();
    if (i >= n) return;
    const int64_t t = orig[i] - (n - nm);
    if (t >= 0) tail_slot[t] = (int)i;
}
// the caller relabel: the k-th valid tail caller index from the top takes the
// k-th merged caller index (its parcel stays in its slot for now)
__global__ void k_relabel(int64_t n, int nm, const int *pis, const int *tflag_c, const int *trank_c,
                          const int *tail_slot, int *orig)
{
    const int64_t u = gtid();
    if (u >= nm || !tflag_c[u]) return;
    const int k = trank_c[u];
    orig[tail_slot[nm - 1 - u]] = orig[pis[k]];   // disjoint: valid parcels written, merged ones read
}
// slot compaction: hole k of hole_list (merged slots below n - nm, ascending)
// takes the k-th valid slot from the top, with all its arrays and caller index
__global__ void k_fill_holes(int64_t n, int nm, const int *hole_list, const int *tflag, const int *trank, ParcelsDev p,
                             float *attr, int64_t cap, int nattr)
{
    const int64_t u = gtid();
    if (u >= nm || !tflag[u]) return;
    const int k = trank[u];          // k-th valid parcel from the top
    const int64_t src = n - 1 - u, dst = hole_list[k];
    p.x[dst] = p.x[src];
    p.y[dst] = p.y[src];
    p.z[dst] = p.z[src];
    for (int a = 0; a < nattr; ++a) attr[a * cap + dst] = attr[a * cap + src];
    p.orig[dst] = p.orig[src];
}
__global__ void k_wrap_all(Grid g, ParcelsDev p)
{
    const int64_t i = gtid();
    if (i >= p.n) return;
    p.x[i] = p.x[i] - g.extent[0] * trunc((p.x[i] - g.center[0]) * g.hli[0]);
    p.y[i] = p.y[i] - g.extent[1] * trunc((p.y[i] - g.center[1]) * g.hli[1]);
}

inline unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

}  // namespace

int run_merge(const MergeCtx &c, int64_t *n_merge)
{
    const Grid &g = c.g;
    ParcelsDev p = c.p;
    const int64_t n = p.n;
    hipStream_t s = c.s;
    *n_merge = 0;
    if (n == 0) return 0;
    // scratch from the member's arena (hipFree would wait for the whole
    // device, other members' steps included); an arena too small for this
    // call is grown first to the largest need seen so far, at least 16
    // bytes per parcel slot (the first phase's flags and ranks)
    ScratchArena &sa = *c.scratch;
    {
        const size_t need = std::max(sa.want, (size_t)16 * (size_t)c.cap + 4096);
        if (sa.cap < need) {
            if (sa.base) {
                (void)hipStreamSynchronize(s);
                (void)hipFree(sa.base);
            }
            sa.base = nullptr;
            sa.cap = 0;
            const size_t grow = need + need / 4;
            if (hipMalloc((void **)&sa.base, grow) != hipSuccess) return -1;
            sa.cap = grow;
        }
    }
    size_t used = 0;
    std::vector<void *> owned;
    auto alloc = [&](void **ptr, size_t bytes) -> hipError_t {
        const size_t b = (std::max<size_t>(bytes, 4) + 255) & ~size_t(255);
        if (used + b <= sa.cap) {
            *ptr = sa.base + used;
            used += b;
            return hipSuccess;
        }
        used += b;   // counted for the next call's arena
        hipError_t e = hipMalloc(ptr, b);
        if (e == hipSuccess) owned.push_back(*ptr);
        return e;
    };
    auto done = [&](int rc) {
        (void)hipStreamSynchronize(s);
        for (void *q : owned) (void)hipFree(q);
        sa.want = std::max(sa.want, used);
        return rc;
    };
    auto host_int = [&](const int *dptr, int &v) -> hipError_t {
        hipError_t e = hipMemcpyAsync(&v, dptr, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        return e;
    };
    auto scan = [&](const int *in, int64_t len, int *out, int *bsum) {
        Plan P;
        add_exclusive_scan(P, PH_OTHER, {ScanArgs{in, len, out, bsum, nullptr}});
        return P.run(s);
    };
#define MOK(expr)                                         \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return done(-1);            \
    } while (0)

    // small parcels in caller-index order: flags by slot, moved to caller
    // indices, ranked there
    int *flag, *cflag, *rank, *bsum;
    MOK(alloc((void **)&flag, sizeof(int) * n));
    MOK(alloc((void **)&cflag, sizeof(int) * n));
    MOK(alloc((void **)&rank, sizeof(int) * (n + 1)));
    MOK(alloc((void **)&bsum, sizeof(int) * scan_scratch_ints(std::max<int64_t>(8 * n, c.ncell) + 1)));
    hipLaunchKernelGGL(k_small, dim3(nblk(n)), dim3(256), 0, s, g, p, c.vmin, flag);
    hipLaunchKernelGGL(k_to_caller, dim3(nblk(n)), dim3(256), 0, s, n, p.orig, flag, cflag);
    MOK(scan(cflag, n, rank, bsum));
    int ns = 0;
    MOK(host_int(rank + n, ns));
    if (ns == 0) return done(0);
    int *sml, *imgflag, *imgrank, *img_of;
    MOK(alloc((void **)&sml, sizeof(int) * ns));
    MOK(alloc((void **)&imgflag, sizeof(int) * 8 * (int64_t)ns));
    MOK(alloc((void **)&imgrank, sizeof(int) * (8 * (int64_t)ns + 1)));
    MOK(alloc((void **)&img_of, sizeof(int) * 8 * (int64_t)ns));
    hipLaunchKernelGGL(k_compact_caller, dim3(nblk(n)), dim3(256), 0, s, n, flag, p.orig, rank, sml);
    hipLaunchKernelGGL(k_image_flags, dim3(nblk(8 * (int64_t)ns)), dim3(256), 0, s, g, p, sml, ns, imgflag);
    MOK(scan(imgflag, 8 * (int64_t)ns, imgrank, bsum));
    int nimg = 0;
    MOK(host_int(imgrank + 8 * (int64_t)ns, nimg));
    double *ipx, *ipy, *ipz;
    int *ipid, *imid;
    MOK(alloc((void **)&ipx, sizeof(double) * nimg));
    MOK(alloc((void **)&ipy, sizeof(double) * nimg));
    MOK(alloc((void **)&ipz, sizeof(double) * nimg));
    MOK(alloc((void **)&ipid, sizeof(int) * nimg));
    MOK(alloc((void **)&imid, sizeof(int) * nimg));
    hipLaunchKernelGGL(k_images, dim3(nblk(8 * (int64_t)ns)), dim3(256), 0, s, g, p, sml, ns, imgflag, imgrank, ipx,
                       ipy, ipz, ipid, imid, img_of);
    // binning of parcels and images
    const int64_t ntot = n + nimg;
    Pts P{p.x, p.y, p.z, ipx, ipy, ipz, n};
    int *key, *cnt, *start, *cursor, *node;
    MOK(alloc((void **)&key, sizeof(int) * ntot));
    MOK(alloc((void **)&cnt, sizeof(int) * c.ncell));
    MOK(alloc((void **)&start, sizeof(int) * (c.ncell + 1)));
    MOK(alloc((void **)&cursor, sizeof(int) * c.ncell));
    MOK(alloc((void **)&node, sizeof(int) * ntot));
    MOK(hipMemsetAsync(cnt, 0, sizeof(int) * c.ncell, s));
    hipLaunchKernelGGL(k_bin_keys, dim3(nblk(ntot)), dim3(256), 0, s, g, P, ntot, c.ncell, key, cnt, c.err);
    {
        Plan Q;
        add_exclusive_scan(Q, PH_OTHER, {ScanArgs{cnt, c.ncell, start, bsum, cursor}});
        add_place(Q, PH_OTHER, {PlaceArgs{ntot, key, start, cursor, node}});
        MOK(Q.run(s));
    }
    // nearest search and the images' results
    const int64_t nsm = (int64_t)ns + nimg;
    int *iclo;
    double *dclo;
    MOK(alloc((void **)&iclo, sizeof(int) * nsm));
    MOK(alloc((void **)&dclo, sizeof(double) * nsm));
    hipLaunchKernelGGL(k_search, dim3(nblk((int64_t)SEARCH_LANES * nsm)), dim3(256), 0, s, g, P, p.orig, sml, ns, nimg,
                       start, node, ipid, iclo, dclo);
    hipLaunchKernelGGL(k_reconcile, dim3(nblk(ns)), dim3(256), 0, s, ns, img_of, iclo, dclo, c.err);
    int e = 0;
    MOK(host_int(c.err, e));
    if (e) return done(1);   // the caller reads the error word
    // tree resolution
    unsigned char *leaf, *avail, *merged;
    int *cont, *dom, *cand, *iso;
    MOK(alloc((void **)&leaf, n));
    MOK(alloc((void **)&avail, n));
    MOK(alloc((void **)&merged, n));
    MOK(alloc((void **)&cont, 2 * sizeof(int)));
    MOK(alloc((void **)&dom, sizeof(int) * ns));
    MOK(alloc((void **)&cand, sizeof(int) * ns));
    MOK(alloc((void **)&iso, sizeof(int) * ns));
    MOK(hipMemsetAsync(leaf, 0, n, s));
    MOK(hipMemsetAsync(avail, 0, n, s));
    MOK(hipMemsetAsync(merged, 0, n, s));
    MOK(hipMemsetAsync(iso, 0, sizeof(int) * ns, s));
    // the leaf iterations, two per read-back of the "merged any" flag: an
    // iteration that merges nothing leaves merged, leaf and avail as they
    // were (each kernel sets or clears the same bits again), so one past
    // convergence changes nothing, and the second flag of a pair says
    // whether the pair ended converged
    for (int it = 0;; it += 2) {
        MOK(hipMemsetAsync(cont, 0, 2 * sizeof(int), s));
        for (int h = 0; h < 2; ++h) {
            hipLaunchKernelGGL(k_tree1, dim3(nblk(ns)), dim3(256), 0, s, sml, ns, iclo, merged, leaf, avail);
            hipLaunchKernelGGL(k_tree2, dim3(nblk(ns)), dim3(256), 0, s, sml, ns, iclo, merged, leaf);
            hipLaunchKernelGGL(k_tree3, dim3(nblk(ns)), dim3(256), 0, s, sml, ns, iclo, merged, leaf, avail);
            hipLaunchKernelGGL(k_tree4, dim3(nblk(ns)), dim3(256), 0, s, sml, ns, iclo, merged, leaf, avail, cont + h);
        }
        int more = 0;
        MOK(host_int(cont + 1, more));
        if (!more) break;
        if (it > ns) return done(-1);   // every iteration that continues merges at least one parcel
    }
    hipLaunchKernelGGL(k_tree5, dim3(nblk(ns)), dim3(256), 0, s, sml, ns, iclo, merged, leaf, avail);
    hipLaunchKernelGGL(k_tree6, dim3(nblk(ns)), dim3(256), 0, s, sml, ns, merged, leaf, avail, dom, cand);
    {
        int *crank, *clist;
        MOK(alloc((void **)&crank, sizeof(int) * (ns + 1)));
        MOK(scan(cand, ns, crank, bsum));
        int ncand = 0;
        MOK(host_int(crank + ns, ncand));
        MOK(alloc((void **)&clist, sizeof(int) * std::max(ncand, 1)));
        hipLaunchKernelGGL(k_compact, dim3(nblk(ns)), dim3(256), 0, s, (int64_t)ns, cand, crank, clist);
        if (ncand > 0) {
            int *cpos;
            MOK(alloc((void **)&cpos, sizeof(int) * n));
            MOK(hipMemsetAsync(cpos, 0xff, sizeof(int) * n, s));
            hipLaunchKernelGGL(k_cand_pos, dim3(nblk(ncand)), dim3(256), 0, s, ncand, clist, sml, cpos);
            hipLaunchKernelGGL(k_tree7, dim3(nblk(ncand)), dim3(256), 0, s, sml, clist, ncand, iclo, cpos, avail, dom,
                               iso);
        }
    }
    hipLaunchKernelGGL(k_tree8, dim3(nblk(ns)), dim3(256), 0, s, ns, iclo, avail, iso, dom);
    // pairs in small-parcel order, grouped by target
    int *prank, *pis, *pic, *gcnt, *gstart, *gcur, *gent;
    MOK(alloc((void **)&prank, sizeof(int) * (ns + 1)));
    MOK(scan(dom, ns, prank, bsum));
    int nm = 0;
    MOK(host_int(prank + ns, nm));
    if (nm == 0) return done(0);
    MOK(alloc((void **)&pis, sizeof(int) * nm));
    MOK(alloc((void **)&pic, sizeof(int) * nm));
    MOK(alloc((void **)&gcnt, sizeof(int) * n));
    MOK(alloc((void **)&gstart, sizeof(int) * (n + 1)));
    MOK(alloc((void **)&gcur, sizeof(int) * n));
    MOK(alloc((void **)&gent, sizeof(int) * nm));
    MOK(hipMemsetAsync(gcnt, 0, sizeof(int) * n, s));
    MOK(hipMemsetAsync(gcur, 0, sizeof(int) * n, s));
    hipLaunchKernelGGL(k_pairs, dim3(nblk(ns)), dim3(256), 0, s, ns, dom, prank, sml, iclo, pis, pic, gcnt);
    MOK(scan(gcnt, n, gstart, bsum));
    hipLaunchKernelGGL(k_group_place, dim3(nblk(nm)), dim3(256), 0, s, nm, pic, gstart, gcur, gent);
    int *mstat;
    MOK(alloc((void **)&mstat, sizeof(int) * 8));
    MOK(hipMemsetAsync(mstat, 0, sizeof(int) * 8, s));
    hipLaunchKernelGGL(k_merge_stats_big, dim3(nblk(nm)), dim3(256), 0, s, nm, pic, p, c.vmin, mstat);
    hipLaunchKernelGGL(k_merge_stats_way, dim3((unsigned)std::min<int64_t>(nblk(n), STATS_WAY_BLOCKS)), dim3(256), 0,
                       s, n, gcnt, mstat);
    int mstat_h[8] = {};
    MOK(hipMemcpyAsync(mstat_h, mstat, sizeof(int) * 8, hipMemcpyDeviceToHost, s));
    hipLaunchKernelGGL(k_group_merge, dim3(nblk(n)), dim3(256), 0, s, g, p, gstart, gent, pis, c.err);
    MOK(host_int(c.err, e));
    if (e) return done(1);
    // delete (holes filled from the end): caller indices relabelled, then
    // the slots compacted; and the periodic wrap of every parcel
    int *hole_c, *hole_s, *tflag, *trank, *tail_slot, *hrank, *hole_list;
    MOK(alloc((void **)&hole_c, sizeof(int) * n));
    MOK(alloc((void **)&hole_s, sizeof(int) * n));
    MOK(alloc((void **)&tflag, sizeof(int) * nm));
    MOK(alloc((void **)&trank, sizeof(int) * (nm + 1)));
    MOK(alloc((void **)&tail_slot, sizeof(int) * nm));
    MOK(alloc((void **)&hrank, sizeof(int) * (n - nm + 1)));
    MOK(alloc((void **)&hole_list, sizeof(int) * nm));
    MOK(hipMemsetAsync(hole_c, 0, sizeof(int) * n, s));
    MOK(hipMemsetAsync(hole_s, 0, sizeof(int) * n, s));
    hipLaunchKernelGGL(k_mark_holes, dim3(nblk(nm)), dim3(256), 0, s, nm, pis, p.orig, hole_c, hole_s);
    hipLaunchKernelGGL(k_tail_flags, dim3(nblk(nm)), dim3(256), 0, s, n, nm, hole_c, tflag);
    MOK(scan(tflag, nm, trank, bsum));
    hipLaunchKernelGGL(k_tail_slots, dim3(nblk(n)), dim3(256), 0, s, n, nm, p.orig, tail_slot);
    hipLaunchKernelGGL(k_relabel, dim3(nblk(nm)), dim3(256), 0, s, n, nm, pis, tflag, trank, tail_slot, p.orig);
    // slots: merged slots below n - nm in ascending order, then the fill
    MOK(scan(hole_s, n - nm, hrank, bsum));
    hipLaunchKernelGGL(k_compact, dim3(nblk(n - nm)), dim3(256), 0, s, n - nm, hole_s, hrank, hole_list);
    hipLaunchKernelGGL(k_tail_flags, dim3(nblk(nm)), dim3(256), 0, s, n, nm, hole_s, tflag);
    MOK(scan(tflag, nm, trank, bsum));
    hipLaunchKernelGGL(k_fill_holes, dim3(nblk(nm)), dim3(256), 0, s, n, nm, hole_list, tflag, trank, p, c.attr, c.cap,
                       c.nattr);
    ParcelsDev q = p;
    q.n = n - nm;
    hipLaunchKernelGGL(k_wrap_all, dim3(nblk(q.n)), dim3(256), 0, s, g, q);
    MOK(hipGetLastError());
    *n_merge = nm;
#undef MOK
    const int rc = done(0);   // synchronises: mstat_h holds the counts
    if (c.stats)
        for (int j = 0; j < 8; ++j) c.stats[j] += mstat_h[j];
    return rc;
}

}  // namespace epic
