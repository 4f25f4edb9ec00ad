// Parcel splitting (parcel_split.f90:37-197) on the device.
//
// 1. k_split_flag   FP64 eigenvalues of every shape matrix (Scherzinger,
//                   B33 from the volume constraint); flag = aspect ratio
//                   sqrt(D1/D3) >= lambda_max or D1 >= amax^2
// 2. scan           exclusive prefix of the flags: the slot of each child
// 3. k_split_apply  every flagged parcel n: B -= 3/4 D1 v1 v1^T, volume / 2;
//                   its child (slot n0 + rank) copies B, vorticity, volume,
//                   buoyancy and humidity; positions x -/+ sqrt(3/5)/2
//                   sqrt(D1) v1; reflective then periodic BC on both
//                   (parcel_bc.f90:17-57, via parcel_communicate)
// Children land after the existing parcels in ascending parent order, the
// order of the reference run on one thread.
#include "common.h"
#include "ellipsoid.h"
#include "launch.h"

namespace epic {

__device__ inline bool full_shape(const ParcelsDev &p, int64_t q, double A[9])
{
    const double B[5] = {p.B[0][q], p.B[1][q], p.B[2][q], p.B[3][q], p.B[4][q]};
    double b33;
    if (!b33_of(B, (double)p.vol[q], b33)) return false;
    A[0] = B[0]; A[1] = B[1]; A[2] = B[2];
    A[3] = B[1]; A[4] = B[3]; A[5] = B[4];
    A[6] = B[2]; A[7] = B[4]; A[8] = b33;
    return true;
}

__global__ __launch_bounds__(256) void k_split_flag(const SplitArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const SplitArgs &a = tab[ib.item];
    int e = 0;
#pragma unroll
    for (int r = 0; r < PER_BLOCK / 256; ++r) {
        const int64_t q = (int64_t)ib.blk * PER_BLOCK + r * 256 + threadIdx.x;
        if (q >= a.p.n) break;
        double A[9], D[3], V[3][3];
        int f = 0;
        if (!full_shape(a.p, q, A)) {
            e |= ERR_B33;
        } else {
            scherzinger(A, D, V);
            const double lam = sqrt(D[0] / D[2]);   // get_aspect_ratio (parcel_ellipsoid.f90:423-429)
            f = (lam < a.lambda_max && D[0] < a.amax * a.amax) ? 0 : 1;
        }
        a.flag[q] = f;
    }
    if (e) atomicOr(a.err, e);
}

__global__ __launch_bounds__(256) void k_split_apply(const SplitArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const SplitArgs &a = tab[ib.item];
    const Grid &g = a.g;
    const ParcelsDev &p = a.p;
    const double dh = 0.5 * sqrt(3.0 / 5.0);
#pragma unroll
    for (int r = 0; r < PER_BLOCK / 256; ++r) {
        const int64_t q = (int64_t)ib.blk * PER_BLOCK + r * 256 + threadIdx.x;
        if (q >= p.n) break;
        if (!a.flag[q]) continue;
        const int64_t c = p.n + a.rank[q];
        double A[9], D[3], V[3][3];
        if (!full_shape(p, q, A)) continue;   // flagged only with a valid B33
        scherzinger(A, D, V);
        double B[5] = {A[0], A[1], A[2], A[4], A[5]};
        const double k = 0.75 * D[0];
        B[0] = B[0] - k * (V[0][0] * V[0][0]);
        B[1] = B[1] - k * V[0][0] * V[0][1];
        B[2] = B[2] - k * V[0][0] * V[0][2];
        B[3] = B[3] - k * (V[0][1] * V[0][1]);
        B[4] = B[4] - k * V[0][1] * V[0][2];
        const float vol = (float)(0.5 * (double)p.vol[q]);
        const double sd = sqrt(D[0]);
        double xc[3], xp[3];
        const double x0[3] = {p.x[q], p.y[q], p.z[q]};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double v = V[0][d] * dh * sd;
            xc[d] = x0[d] - v;
            xp[d] = x0[d] + v;
        }
        double Bc[5], Bp[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) Bc[i] = Bp[i] = B[i];
        auto reflect = [&](double x[3], double Bm[5]) {   // parcel_bc.f90:26-57
            if (x[2] > g.upper[2]) {
                x[2] = 2.0 * g.upper[2] - x[2];
                Bm[2] = -Bm[2];
                Bm[4] = -Bm[4];
            } else if (x[2] < g.lower[2]) {
                x[2] = 2.0 * g.lower[2] - x[2];
                Bm[2] = -Bm[2];
                Bm[4] = -Bm[4];
            }
            x[0] = x[0] - g.extent[0] * trunc((x[0] - g.center[0]) * g.hli[0]);   // :17-21
            x[1] = x[1] - g.extent[1] * trunc((x[1] - g.center[1]) * g.hli[1]);
        };
        reflect(xc, Bc);
        reflect(xp, Bp);
        p.x[q] = xp[0];
        p.y[q] = xp[1];
        p.z[q] = xp[2];
        p.x[c] = xc[0];
        p.y[c] = xc[1];
        p.z[c] = xc[2];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            p.B[i][q] = (float)Bp[i];
            p.B[i][c] = (float)Bc[i];
            p.dB[i][c] = 0.0f;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            p.vor[i][c] = p.vor[i][q];
            p.dpos[i][c] = 0.0f;
            p.dvor[i][c] = 0.0f;
        }
        p.vol[q] = vol;
        p.vol[c] = vol;
        p.buoy[c] = p.buoy[q];
        p.hum[c] = p.hum[q];
        p.orig[c] = (int)(p.n + a.crank[p.orig[q]]);   // children in their parents' caller order
    }
}

static std::vector<int> parcel_blocks(const std::vector<SplitArgs> &items)
{
    std::vector<int> b;
    for (const SplitArgs &a : items) b.push_back((int)((a.p.n + PER_BLOCK - 1) / PER_BLOCK));
    return b;
}

void add_split_flag(Plan &P, int phase, const std::vector<SplitArgs> &items)
{
    P.add((const void *)k_split_flag, 256, 0, phase, items, parcel_blocks(items));
}

void add_split_apply(Plan &P, int phase, const std::vector<SplitArgs> &items)
{
    P.add((const void *)k_split_apply, 256, 0, phase, items, parcel_blocks(items));
}

}  // namespace epic
