// grid2par (parcel_interpl.f90:405-470), and grid2par fused with dB/dt and
// the low-storage RK push (parcel_ellipsoid.f90:466-507, ls_rk.f90:142-217,
// parcel_bc.f90:17-40) for ls_rk_step.
//
// Per parcel: four support points from the half axes; each point gathers
// velocity (3), velocity gradient (8) and vorticity tendency (3) from the
// eight corner nodes with trilinear weights (FP64 cell fractions); the
// quarter-weighted sums update delta_pos / delta_vor.  In the fused form the
// gradient gives dB/dt (B33 from det B in FP64), the stage pushes position
// (FP64), vorticity and B by cb dt delta, applies the periodic wrap (int()
// truncation), scales the deltas by ca for the next stage and computes the
// next stage's cell key and cell histogram; on every stage but the last it
// also stores the half axes of the B it wrote (FP64 eigen-solve) and the
// next par2grid's fixed-point maxima.
#include "common.h"
#include "ellipsoid.h"
#include "launch.h"

#include <algorithm>

namespace epic {

// Stand-alone grid2par(add) (parcel_interpl.f90:405-470): support points
// from the current B (stored as the cache of the next par2grid, :442),
// quarter-weighted trilinear gathers of velocity, velocity gradient and
// vorticity tendency; delta_pos / delta_vor are set or accumulated, strain is
// set.  One parcel per thread.
__global__ __launch_bounds__(256) void k_grid2par(const G2PArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const G2PArgs &a = tab[ib.item];
    const Grid &g = a.g;
    const int64_t PS = plane_stride(g), FS = field_stride(g);
    int e = 0;
    const int64_t q = ib.blk * (int64_t)blockDim.x + threadIdx.x;
    if (q < a.p.n) {
        const double pos[3] = {gl(a.p.x)[q], gl(a.p.y)[q], gl(a.p.z)[q]};
        double B[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) B[c] = (double)gl(a.p.B[c])[q];
        const double vol = (double)gl(a.p.vol)[q];
        double ve[3], vt[3];
        if (!support_axes(B, vol, ve, vt)) {
            e |= ERR_B33;
            ve[0] = ve[1] = ve[2] = vt[0] = vt[1] = vt[2] = 0.0;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            gl(a.p.veta[d])[q] = (float)ve[d];
            gl(a.p.vtau[d])[q] = (float)vt[d];
        }
        float dp[3], dv[3], st[8];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            dp[c] = a.add ? gl(a.p.dpos[c])[q] : 0.0f;
            dv[c] = a.add ? gl(a.p.dvor[c])[q] : 0.0f;
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) st[c] = 0.0f;
        for (int jp = 0; jp < 4; ++jp) {
            const double cs = support_cos(jp), sn = support_sin(jp);
            int ii[3];
            float w[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double pt = pos[d] + ve[d] * cs + vt[d] * sn;
                const double xyz = (pt - g.lower[d]) * g.dxi[d];
                const double fl = floor(xyz);
                ii[d] = (int)fl;
                w[d] = (float)(xyz - fl);
            }
            if (ii[2] < -1 || ii[2] > g.nz) {
                e |= ERR_OUT_OF_GRID;
                continue;
            }
            const int x0 = wrap(ii[0], g.nx), x1 = wrap(ii[0] + 1, g.nx);
            const int y0 = wrap(ii[1], g.ny), y1 = wrap(ii[1] + 1, g.ny);
            const int64_t zp0 = (int64_t)(ii[2] + 1) * PS, zp1 = zp0 + PS;
            // corners in Fortran element order of ww(0:1,0:1,0:1): k fastest
            const int64_t off[8] = {zp0 + (int64_t)y0 * g.nx + x0, zp1 + (int64_t)y0 * g.nx + x0,
                                    zp0 + (int64_t)y1 * g.nx + x0, zp1 + (int64_t)y1 * g.nx + x0,
                                    zp0 + (int64_t)y0 * g.nx + x1, zp1 + (int64_t)y0 * g.nx + x1,
                                    zp0 + (int64_t)y1 * g.nx + x1, zp1 + (int64_t)y1 * g.nx + x1};
            const float pxc = 1.0f - w[0], pyc = 1.0f - w[1], pzc = 1.0f - w[2];
            const float w00 = pyc * pxc, w10 = pyc * w[0], w01 = w[1] * pxc, w11 = w[1] * w[0];
            const float wt[8] = {pzc * w00, w[2] * w00, pzc * w01, w[2] * w01,
                                 pzc * w10, w[2] * w10, pzc * w11, w[2] * w11};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const EPIC_GLOBAL float *f = gl(a.fld) + (F_VELOG + c) * FS;
                float s = 0.0f;
#pragma unroll
                for (int m = 0; m < 8; ++m) s += wt[m] * f[off[m]];
                dp[c] += 0.25f * s;
            }
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const EPIC_GLOBAL float *f = gl(a.fld) + (F_VELGRADG + c) * FS;
                float s = 0.0f;
#pragma unroll
                for (int m = 0; m < 8; ++m) s += wt[m] * f[off[m]];
                st[c] += 0.25f * s;
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const EPIC_GLOBAL float *f = gl(a.fld) + (F_VTEND + c) * FS;
                float s = 0.0f;
#pragma unroll
                for (int m = 0; m < 8; ++m) s += wt[m] * f[off[m]];
                dv[c] += 0.25f * s;
            }
        }

#pragma unroll
        for (int c = 0; c < 3; ++c) {
            gl(a.p.dpos[c])[q] = dp[c];
            gl(a.p.dvor[c])[q] = dv[c];
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) gl(a.p.strain[c])[q] = st[c];
    }
    if (e) atomicOr(a.err, e);
}

// ---------------------------------------------------------------------------
// Tiled form used by ls_rk_step: one work-group per (8x8 column tile, z chunk
// of tz cell layers).  The 14 gathered fields of the chunk's node box (node
// planes k0-1 .. k0+tz+1, x and y from -1 to 9 around the tile) are staged in
// LDS once as four planes of field groups (fields 0-3, 4-7, 8-11 as float4,
// 12-13 as float2), so every corner is three 16-byte and one 8-byte LDS read
// and neighbouring nodes sit in neighbouring banks; then every thread takes parcels of the
// chunk through the cell sort (the rows of 8 cells are contiguous slot
// ranges) with no barrier between parcels.  Support points come from the
// half axes cached for this stage.  The permutation entry of a thread's next
// parcel is loaded one round ahead, and parcel arrays are addressed as a
// table pointer plus a 32-bit byte offset.
// ---------------------------------------------------------------------------
constexpr int G2P_THREADS = 256;
size_t g2p_lds_bytes(int tz) { return (3 * sizeof(float4) + sizeof(float2)) * (tz + 3) * P2G_BOX * P2G_BOX; }

__global__ __launch_bounds__(G2P_THREADS, 4) void k_g2p_tiles(const G2PTileArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const G2PTileArgs &a = tab[ib.item];
    const Grid &g = a.g;
    const int tz = a.tz;
    const int tx = ib.blk % a.ntx;
    const int ty = (ib.blk / a.ntx) % a.nty;
    const int cz = ib.blk / (a.ntx * a.nty);
    const int i0 = tx * P2G_TX, j0 = ty * P2G_TY;
    const int k0 = -1 + cz * tz;
    const int nl = min(tz, g.nz + 1 - k0);   // layers k0 .. k0 + nl - 1
    const int tid = threadIdx.x;
    const int64_t PS = plane_stride(g), FS = field_stride(g);
    constexpr int NN = P2G_BOX * P2G_BOX;
    const int nbox = tz + 3;

    extern __shared__ float4 box[];                 // [group][plane][y][x]: 3 float4 groups, then a float2 group
    const int NB = nbox * NN;                       // nodes in the box
    float2 *box2 = (float2 *)(box + 3 * NB);
    __shared__ int row_slot[64], row_off[65];       // the chunk's rows of 8 cells
    __shared__ unsigned int s_pmax[P2G_NF + 1];     // the next par2grid's maxima (float bits)
    int e = 0;
    if (tid <= P2G_NF) s_pmax[tid] = 0u;   // ordered by the barrier after the box is staged

    // ---- stage the node box (plane P = k0 - 1 + lp) ----
    // thread = (node of the 11 x 11 box column, half of the 14 fields): the
    // wrapped column offset is found once, then the thread walks the planes
    // with its 7 loads in flight at a time
    static_assert(NN <= 128 && G2P_THREADS == 256, "two threads per box node column");
    {
        const int node = tid & 127, half = tid >> 7;
        if (node < NN) {
            const int ly = node / P2G_BOX, lx = node - ly * P2G_BOX;
            int gi = i0 - 1 + lx, gj = j0 - 1 + ly;   // in [-1, n + 9]
            gi = (gi < 0) ? gi + g.nx : (gi >= g.nx ? gi - g.nx : gi);
            gj = (gj < 0) ? gj + g.ny : (gj >= g.ny ? gj - g.ny : gj);
            const EPIC_GLOBAL float *col = gl(a.fld) + (int64_t)(F_VELOG + half) * FS + (int64_t)gj * g.nx + gi;
            for (int lp = 0; lp < nbox; ++lp) {
                const int P = k0 - 1 + lp;
                const bool in = (P >= -1 && P <= g.nz + 1);
                const int64_t po = (int64_t)(P + 1) * PS;
                float v[7];
#pragma unroll
                for (int t = 0; t < 7; ++t) v[t] = in ? col[(int64_t)(2 * t) * FS + po] : 0.0f;
                const int nd = lp * NN + node;
#pragma unroll
                for (int t = 0; t < 7; ++t) {
                    const int f = 2 * t + half;
                    if (f < 12) ((float *)box)[((f >> 2) * NB + nd) * 4 + (f & 3)] = v[t];
                    else ((float *)box2)[nd * 2 + (f - 12)] = v[t];
                }
            }
        }
    }
    const int nrow = nl * P2G_TY;
    if (tid < nrow) {   // a partial tile's rows are shorter (wx cells), or beyond the grid (empty)
        const int k = k0 + tid / P2G_TY, r = tid % P2G_TY;
        const int wx = min(P2G_TX, g.nx - i0);
        if (j0 + r < g.ny) {
            const int c0 = ((k + 1) * g.ny + (j0 + r)) * g.nx + i0;
            row_slot[tid] = gl(a.cell_start)[c0];
            row_off[tid + 1] = gl(a.cell_start)[c0 + wx] - gl(a.cell_start)[c0];
        } else {
            row_slot[tid] = 0;
            row_off[tid + 1] = 0;
        }
    }
    __syncthreads();
    if (tid < 64) {   // inclusive scan of the row sizes in one wave
        int v = (tid < nrow) ? row_off[tid + 1] : 0;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(v, o, 64);
            if (tid >= o) v += t;
        }
        if (tid < nrow) row_off[tid + 1] = v;
        if (tid == 0) row_off[0] = 0;
    }
    __syncthreads();
    const int total = row_off[nrow];

    // the row holding entry q: largest r with row_off[r] <= q
    auto row_of = [&](int q) {
        int r = 0;
        for (int step = 32; step > 0; step >>= 1)
            if (r + step < nrow && row_off[r + step] <= q) r += step;
        return r;
    };
    // The next round's entry is loaded one round ahead, and the position and
    // half axes of the parcel it names are requested in the middle of the
    // current round (after the gather), so the support points of a round
    // start without waiting for memory.
    int r_nx = 0, pi_nx = 0;
    double pn[3] = {0.0, 0.0, 0.0};
    float an[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    auto prefetch = [&](int pj) {
        const unsigned oj = (unsigned)pj * 4u;
        pn[0] = gl(a.p.x)[pj];
        pn[1] = gl(a.p.y)[pj];
        pn[2] = gl(a.p.z)[pj];
#pragma unroll
        for (int c = 0; c < 6; ++c) an[c] = at(a.axes[c], oj);
    };
    if (tid < total) {
        r_nx = row_of(tid);
        pi_nx = gl(a.perm)[row_slot[r_nx] + (tid - row_off[r_nx])];
        prefetch(pi_nx);
    }
    for (int base = 0; base < total; base += G2P_THREADS) {   // uniform rounds: the ballot needs every lane
        const int q = base + tid;
        const int r = r_nx, pi = pi_nx;
        const unsigned o4 = (unsigned)pi * 4u;   // byte offset (pi < 2^30): scalar base + 32-bit offset addressing
        const bool more = q + G2P_THREADS < total;
        if (more) {
            r_nx = row_of(q + G2P_THREADS);
            pi_nx = gl(a.perm)[row_slot[r_nx] + (q + G2P_THREADS - row_off[r_nx])];
        }
        double pos[3] = {pn[0], pn[1], pn[2]};
        const double ve[3] = {an[0], an[1], an[2]};
        const double vt[3] = {an[3], an[4], an[5]};
        int key = -1;
        bool small = false;
        auto process = [&]() {
            if (q >= total) return;
            const int k = k0 + r / P2G_TY;
            double B[5];
#pragma unroll
            for (int c = 0; c < 5; ++c) B[c] = (double)at(a.p.B[c], o4);
            const double vol = (double)at(a.p.vol, o4);
            int cc[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) cc[d] = (int)floor((pos[d] - g.lower[d]) * g.dxi[d]);
            const int cx = wrap(cc[0], g.nx) - i0, cy = wrap(cc[1], g.ny) - j0;
            if (cx < 0 || cx >= P2G_TX || cy < 0 || cy >= P2G_TY || cc[2] != k) {
                e |= ERR_CELL_MISMATCH;
                if (more) prefetch(pi_nx);
                return;
            }
            const bool acc_prev = a.stage > 1;
            float dp[3], dv[3], st[8];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                dp[c] = acc_prev ? at(a.p.dpos[c], o4) : 0.0f;
                dv[c] = acc_prev ? at(a.p.dvor[c], o4) : 0.0f;
            }
#pragma unroll
            for (int c = 0; c < 8; ++c) st[c] = 0.0f;
#pragma unroll
            for (int jp = 0; jp < 4; ++jp) {
                const double cs = support_cos(jp), sn = support_sin(jp);
                int off[3];
                float w[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double pt = pos[d] + ve[d] * cs + vt[d] * sn;
                    const double xyz = (pt - g.lower[d]) * g.dxi[d];
                    const double fl = floor(xyz);
                    off[d] = (int)fl - cc[d];
                    w[d] = (float)(xyz - fl);
                }
                if (k + off[2] < -1 || k + off[2] > g.nz) {   // beyond the halo layers (the reference fails too)
                    e |= ERR_OUT_OF_GRID;
                    continue;
                }
                // a point beyond the neighbour cells (strongly stretched
                // parcels, rare) gathers its corners from the global fields,
                // in the same order, so the sums are the same bits
                const bool far = off[0] < -1 || off[0] > 1 || off[1] < -1 || off[1] > 1 || off[2] < -1 || off[2] > 1;
                // box coordinates of the point's lower corner
                const int lx = cx + off[0] + 1, ly = cy + off[1] + 1, lp = k + off[2] - (k0 - 1);
                const float pxc = 1.0f - w[0], pyc = 1.0f - w[1], pzc = 1.0f - w[2];
                const float w00 = pyc * pxc, w10 = pyc * w[0], w01 = w[1] * pxc, w11 = w[1] * w[0];
                // corners in Fortran element order of ww(0:1,0:1,0:1): k fastest
                const float wt[8] = {pzc * w00, w[2] * w00, pzc * w01, w[2] * w01,
                                     pzc * w10, w[2] * w10, pzc * w11, w[2] * w11};
                float s[14];
#pragma unroll
                for (int f = 0; f < 14; ++f) s[f] = 0.0f;
                if (!far) {
#pragma unroll
                    for (int m = 0; m < 8; ++m) {
                        const int dz = m & 1, dy = (m >> 1) & 1, dx = (m >> 2) & 1;
                        const int nd = (lp + dz) * NN + (ly + dy) * P2G_BOX + (lx + dx);
                        const float4 q0 = box[nd], q1 = box[NB + nd], q2 = box[2 * NB + nd];
                        const float2 q3 = box2[nd];
                        const float vv[14] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w,
                                              q2.x, q2.y, q2.z, q2.w, q3.x, q3.y};
#pragma unroll
                        for (int f = 0; f < 14; ++f) s[f] += wt[m] * vv[f];
                    }
                } else {
#pragma unroll 1
                    for (int m = 0; m < 8; ++m) {
                        const int dz = m & 1, dy = (m >> 1) & 1, dx = (m >> 2) & 1;
                        const int ii = wrap(cc[0] + off[0] + dx, g.nx), jj = wrap(cc[1] + off[1] + dy, g.ny);
                        const int64_t o = (int64_t)(k + off[2] + dz + 1) * PS + (int64_t)jj * g.nx + ii;
                        const EPIC_GLOBAL float *f0 = gl(a.fld) + (int64_t)F_VELOG * FS + o;
#pragma unroll
                        for (int f = 0; f < 14; ++f) s[f] += wt[m] * f0[(int64_t)f * FS];
                    }
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) dp[c] += 0.25f * s[c];
#pragma unroll
                for (int c = 0; c < 8; ++c) st[c] += 0.25f * s[3 + c];
#pragma unroll
                for (int c = 0; c < 3; ++c) dv[c] += 0.25f * s[11 + c];
            }
            if (more) prefetch(pi_nx);   // the next round's position and axes, in flight during the rest
            // dB/dt (FP64, B33 from the volume constraint)
            double S[8], dB[5];
#pragma unroll
            for (int c = 0; c < 8; ++c) S[c] = (double)st[c];
            if (!dBdt(B, S, vol, dB)) {
                e |= ERR_B33;
                dB[0] = dB[1] = dB[2] = dB[3] = dB[4] = 0.0;
            }
            float db[5];
#pragma unroll
            for (int c = 0; c < 5; ++c) db[c] = acc_prev ? at(a.p.dB[c], o4) + (float)dB[c] : (float)dB[c];
            // push (ls_rk.f90:183-191), periodic wrap (parcel_bc.f90:17-21)
            const double cbdt = a.cb * (*a.dt);
#pragma unroll
            for (int c = 0; c < 3; ++c) pos[c] = pos[c] + cbdt * (double)dp[c];
#pragma unroll
            for (int c = 0; c < 3; ++c) at(a.p.vor[c], o4) = (float)((double)at(a.p.vor[c], o4) + cbdt * (double)dv[c]);
#pragma unroll
            for (int c = 0; c < 5; ++c) B[c] = B[c] + cbdt * (double)db[c];
            pos[0] = pos[0] - g.extent[0] * trunc((pos[0] - g.center[0]) * g.hli[0]);
            pos[1] = pos[1] - g.extent[1] * trunc((pos[1] - g.center[1]) * g.hli[1]);
            if (a.reflect) {   // parcel_bc.f90:26-40
                if (pos[2] > g.upper[2]) {
                    pos[2] = 2.0 * g.upper[2] - pos[2];
                    B[2] = -B[2];
                    B[4] = -B[4];
                } else if (pos[2] < g.lower[2]) {
                    pos[2] = 2.0 * g.lower[2] - pos[2];
                    B[2] = -B[2];
                    B[4] = -B[4];
                }
            }
            gl(a.p.x)[pi] = pos[0];
            gl(a.p.y)[pi] = pos[1];
            gl(a.p.z)[pi] = pos[2];
#pragma unroll
            for (int c = 0; c < 5; ++c) at(a.p.B[c], o4) = (float)B[c];
            if (a.stage < a.nstages) {
                const float ca = (float)a.ca;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    at(a.p.dpos[c], o4) = ca * dp[c];
                    at(a.p.dvor[c], o4) = ca * dv[c];
                }
#pragma unroll
                for (int c = 0; c < 5; ++c) at(a.p.dB[c], o4) = ca * db[c];
            }
            // cell of the new position for the next stage's sort
            const int ni = (int)floor((pos[0] - g.lower[0]) * g.dxi[0]);
            const int nj = (int)floor((pos[1] - g.lower[1]) * g.dxi[1]);
            const int nk = (int)floor((pos[2] - g.lower[2]) * g.dxi[2]);
            if (nk < -1 || nk > g.nz) e |= ERR_PARCEL_COUNT;
            else key = ((nk + 1) * g.ny + wrap(nj, g.ny)) * g.nx + wrap(ni, g.nx);
            at(a.keys, o4) = key;
            small = (vol <= g.vmin);
            if (a.axes_next[0]) {   // the next stage's support axes, from B as stored
                double Bs[5], ve2[3], vt2[3];
#pragma unroll
                for (int c = 0; c < 5; ++c) Bs[c] = (double)(float)B[c];
                if (!support_axes(Bs, vol, ve2, vt2)) {
                    e |= ERR_B33;
                    ve2[0] = ve2[1] = ve2[2] = vt2[0] = vt2[1] = vt2[2] = 0.0;
                }
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    at(a.axes_next[d], o4) = (float)ve2[d];
                    at(a.axes_next[3 + d], o4) = (float)vt2[d];
                }
            }
            if (a.tmax_next) {
                // bounds of the next par2grid's |contributions| of this
                // parcel, from what it stored: the condensate is at most the
                // humidity, so |tbuoy| <= |buoy| + glat hum and |tbuoy - buoy|
                // <= glat hum (exact when there is no humidity); a scale from
                // a bound is at most one power of two coarser, and the
                // fixed-point sums keep a factor 2 of headroom for rounding
                const float v0 = (float)(0.25 * vol), h = at(a.p.hum, o4);
                const float hb = fabsf(v0 * (float)g.glat * h);
                const float b5 = fabsf(v0 * at(a.p.buoy, o4));
                const float m[P2G_NF + 1] = {fabsf(v0), fabsf(v0 * at(a.p.vor[0], o4)), fabsf(v0 * at(a.p.vor[1], o4)),
                                             fabsf(v0 * at(a.p.vor[2], o4)), b5 + hb, b5, fabsf(v0 * h), hb};
                // into the work-group's maxima (bits of non-negative floats);
                // a lane below the running maximum skips the atomic
#pragma unroll
                for (int f = 0; f <= P2G_NF; ++f)
                    if (m[f] > __uint_as_float(s_pmax[f])) atomicMax(&s_pmax[f], __float_as_uint(m[f]));
            }
        };
        process();
        // next stage's cell histogram, one atomic per run of equal keys
        wave_run_add(key, gl(a.count));
        wave_run_add(small ? key : -1, gl(a.scount));
    }
    if (a.tmax_next) {
        __syncthreads();
        if (tid <= P2G_NF && s_pmax[tid]) atomicMax((unsigned int *)a.tmax_next + tid, s_pmax[tid]);
    }
    if (e) atomicOr(a.err, e);
}

void add_rk_stage_tiles(Plan &P, int phase, const std::vector<G2PTileArgs> &items)
{
    std::vector<int> nb;
    size_t lds = 0;
    for (const G2PTileArgs &a : items) {
        nb.push_back(a.ntx * a.nty * a.nchz);   // (empty rows: no work)
        lds = std::max(lds, g2p_lds_bytes(a.tz));
    }
    P.add((const void *)k_g2p_tiles, G2P_THREADS, lds, phase, items, nb);
}

void add_grid2par(Plan &P, int phase, const std::vector<G2PArgs> &items)
{
    std::vector<int> nb;
    for (const G2PArgs &a : items) nb.push_back((int)((a.p.n + 255) / 256));
    P.add((const void *)k_grid2par, 256, 0, phase, items, nb);
}

}  // namespace epic
