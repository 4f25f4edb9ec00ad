// grid2par, dB/dt and the low-storage RK push fused into one pass over the
// parcels (parcel_interpl.f90:405-470, parcel_ellipsoid.f90:466-507,
// ls_rk.f90:142-217, parcel_bc.f90:17-40).
//
// Per parcel: FP64 eigen-decomposition of B -> four support points; each point
// gathers velocity (3), velocity gradient (8) and vorticity tendency (3) from
// the eight corner nodes with trilinear weights (FP64 cell fractions); the
// quarter-weighted sums update delta_pos / delta_vor, the gradient gives
// dB/dt (B33 from det B in FP64), and the stage pushes position (FP64),
// vorticity and B by cb dt delta, applies the periodic wrap (int()
// truncation), scales the deltas by ca for the next stage and computes the
// next stage's cell key and cell histogram.
#include "common.h"
#include "ellipsoid.h"
#include "launch.h"

namespace epic {

struct G2PArgs {
    Grid g;
    ParcelsDev p;
    const float *fld;      // node fields (F_VELOG, F_VELGRADG, F_VTEND used)
    int mode;              // 0: stand-alone grid2par, 1: fused RK stage
    int add;               // grid2par: accumulate into delta_pos/delta_vor
    int stage;             // RK stage 1..nstages
    int nstages;
    int reflect;           // apply the reflective z boundary after the push
    int store_axes;        // cache Veta/Vtau for the next par2grid(l_reuse)
    double cbdt, ca;
    int *keys, *count, *scount;   // next-stage keys/histogram (may be null)
    int *err;
};

__global__ __launch_bounds__(256) void k_g2p(G2PArgs a)
{
    const Grid &g = a.g;
    const int64_t PS = plane_stride(g), FS = field_stride(g);
    int e = 0;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < a.p.n; q += (int64_t)gridDim.x * blockDim.x) {
        double pos[3] = {a.p.x[q], a.p.y[q], a.p.z[q]};
        double B[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) B[c] = (double)a.p.B[c][q];
        const double vol = (double)a.p.vol[q];
        double ve[3], vt[3];
        if (!support_axes(B, vol, ve, vt)) {
            e |= ERR_B33;
            ve[0] = ve[1] = ve[2] = vt[0] = vt[1] = vt[2] = 0.0;
        }
        if (a.store_axes) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                a.p.veta[d][q] = (float)ve[d];
                a.p.vtau[d][q] = (float)vt[d];
            }
        }
        float dp[3], dv[3], st[8];
        const bool accumulate = (a.mode == 0) ? (a.add != 0) : (a.stage > 1);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            dp[c] = accumulate ? a.p.dpos[c][q] : 0.0f;
            dv[c] = accumulate ? a.p.dvor[c][q] : 0.0f;
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
                const float *f = a.fld + (F_VELOG + c) * FS;
                float s = 0.0f;
#pragma unroll
                for (int m = 0; m < 8; ++m) s += wt[m] * f[off[m]];
                dp[c] += 0.25f * s;
            }
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const float *f = a.fld + (F_VELGRADG + c) * FS;
                float s = 0.0f;
#pragma unroll
                for (int m = 0; m < 8; ++m) s += wt[m] * f[off[m]];
                st[c] += 0.25f * s;
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float *f = a.fld + (F_VTEND + c) * FS;
                float s = 0.0f;
#pragma unroll
                for (int m = 0; m < 8; ++m) s += wt[m] * f[off[m]];
                dv[c] += 0.25f * s;
            }
        }

        if (a.mode == 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                a.p.dpos[c][q] = dp[c];
                a.p.dvor[c][q] = dv[c];
            }
#pragma unroll
            for (int c = 0; c < 8; ++c) a.p.strain[c][q] = st[c];
            continue;
        }

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
        for (int c = 0; c < 5; ++c) db[c] = (a.stage > 1) ? a.p.dB[c][q] + (float)dB[c] : (float)dB[c];

        // push (ls_rk.f90:183-191)
#pragma unroll
        for (int c = 0; c < 3; ++c) pos[c] = pos[c] + a.cbdt * (double)dp[c];
#pragma unroll
        for (int c = 0; c < 3; ++c) a.p.vor[c][q] = (float)((double)a.p.vor[c][q] + a.cbdt * (double)dv[c]);
#pragma unroll
        for (int c = 0; c < 5; ++c) B[c] = B[c] + a.cbdt * (double)db[c];
        // periodic wrap with Fortran int() truncation (parcel_bc.f90:17-21)
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
        a.p.x[q] = pos[0];
        a.p.y[q] = pos[1];
        a.p.z[q] = pos[2];
#pragma unroll
        for (int c = 0; c < 5; ++c) a.p.B[c][q] = (float)B[c];
        if (a.stage < a.nstages) {
            const float ca = (float)a.ca;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                a.p.dpos[c][q] = ca * dp[c];
                a.p.dvor[c][q] = ca * dv[c];
            }
#pragma unroll
            for (int c = 0; c < 5; ++c) a.p.dB[c][q] = ca * db[c];
        }
        if (a.keys) {
            const int i = (int)floor((pos[0] - g.lower[0]) * g.dxi[0]);
            const int j = (int)floor((pos[1] - g.lower[1]) * g.dxi[1]);
            const int k = (int)floor((pos[2] - g.lower[2]) * g.dxi[2]);
            int key = -1;
            if (k < -1 || k > g.nz) e |= ERR_PARCEL_COUNT;
            else key = ((k + 1) * g.ny + wrap(j, g.ny)) * g.nx + wrap(i, g.nx);
            a.keys[q] = key;
            if (key >= 0) {
                atomicAdd(&a.count[key], 1);
                if (vol <= g.vmin) atomicAdd(&a.scount[key], 1);
            }
        }
    }
    if (e) atomicOr(a.err, e);
}

static int g2p_blocks(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 16384); }

void launch_grid2par(const Grid &g, const ParcelsDev &p, const float *fld, int add, int *err, hipStream_t s)
{
    if (p.n == 0) return;
    G2PArgs a = {};
    a.g = g;
    a.p = p;
    a.fld = fld;
    a.mode = 0;
    a.add = add;
    a.store_axes = 1;
    a.err = err;
    hipLaunchKernelGGL(k_g2p, dim3(g2p_blocks(p.n)), dim3(256), 0, s, a);
}

void launch_rk_stage(const Grid &g, const ParcelsDev &p, const float *fld, int stage, int nstages, double cbdt,
                     double ca, int reflect, int store_axes, int *keys, int *count, int *scount, int *err,
                     hipStream_t s)
{
    if (p.n == 0) return;
    G2PArgs a = {};
    a.g = g;
    a.p = p;
    a.fld = fld;
    a.mode = 1;
    a.stage = stage;
    a.nstages = nstages;
    a.reflect = reflect;
    a.store_axes = store_axes;
    a.cbdt = cbdt;
    a.ca = ca;
    a.keys = keys;
    a.count = count;
    a.scount = scount;
    a.err = err;
    hipLaunchKernelGGL(k_g2p, dim3(g2p_blocks(p.n)), dim3(256), 0, s, a);
}

}  // namespace epic
