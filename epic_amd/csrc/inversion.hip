// Spectral vorticity inversion on gfx950 (inversion.f90:25-366,
// inversion_utils.f90:79-514).
//
// After the horizontal FFTs every horizontal operator of vor2vel is diagonal
// in (kx, ky), so the whole vertical part of the inversion is independent per
// spectral column.  k_vor2vel_columns gives each column to one wavefront and
// runs it in FP64 in LDS: decomposition into boundary values + sine series,
// Hou-Li filter, solenoidal projection, vertical Green's function, the
// hyperbolic boundary corrections theta/phi, the zero-mode mean flow and the
// velocity and velocity-gradient spectra.  The z transforms (DST-I / DCT-I
// with the sqrt(2/nz) normalisation of stafft.f90:410-555) are evaluated as
// one complex FFT of length 2 nz of the odd / even extension of the column.
// Tables (filter, Green, theta, phi, gamma) are evaluated on the fly in FP64
// from the same formulas as init_inversion; nothing is stored per column.
#include "common.h"
#include "fft.h"
#include "launch.h"

#include <algorithm>

namespace epic {

#define EPIC_PI 3.14159265358979323846

struct ColCtx {
    int nz;
    double2 *fa, *fb;
    const FftPlan *plan;
    const double2 *tw;   // e^{-2 pi i t / (2 nz)}, t < nz (the member's table; cached reads)
};

// DST-I of col[1..nz-1] in place (col[0], col[nz] untouched)
__device__ void zdst(const ColCtx &c, double2 *col)
{
    const int nz = c.nz, n2 = 2 * nz;
    for (int t = threadIdx.x; t < n2; t += blockDim.x) {
        double2 v = {0.0, 0.0};
        if (t > 0 && t < nz) v = col[t];
        else if (t > nz) v = double2{-col[n2 - t].x, -col[n2 - t].y};
        c.fa[t] = v;
    }
    __syncthreads();
    double2 *R = lds_fft_tw(c.fa, c.fb, 1, *c.plan, TwHalf<double2>{c.tw, nz}, false);
    const double s = 0.5 * sqrt(2.0 / (double)nz);
    for (int t = threadIdx.x + 1; t < nz; t += blockDim.x) col[t] = double2{-s * R[t].y, s * R[t].x};   // s i R
    __syncthreads();
}

// DCT-I of col[0..nz] in place
__device__ void zdct(const ColCtx &c, double2 *col)
{
    const int nz = c.nz, n2 = 2 * nz;
    for (int t = threadIdx.x; t < n2; t += blockDim.x) c.fa[t] = (t <= nz) ? col[t] : col[n2 - t];
    __syncthreads();
    double2 *R = lds_fft_tw(c.fa, c.fb, 1, *c.plan, TwHalf<double2>{c.tw, nz}, false);
    const double s = 0.5 * sqrt(2.0 / (double)nz);
    for (int t = threadIdx.x; t <= nz; t += blockDim.x) col[t] = double2{s * R[t].x, s * R[t].y};
    __syncthreads();
}

// field_decompose_semi_spectral (inversion_utils.f90:372-401)
__device__ void decompose(const ColCtx &c, double2 *col, const double *phim, const double *phip)
{
    const int nz = c.nz;
    const double2 b0 = col[0], bn = col[nz];
    __syncthreads();
    for (int z = threadIdx.x + 1; z < nz; z += blockDim.x)
        col[z] = double2{col[z].x - (b0.x * phim[z] + bn.x * phip[z]), col[z].y - (b0.y * phim[z] + bn.y * phip[z])};
    __syncthreads();
    zdst(c, col);
}

// field_combine_semi_spectral (inversion_utils.f90:426-456)
__device__ void combine(const ColCtx &c, double2 *col, const double *phim, const double *phip)
{
    const int nz = c.nz;
    zdst(c, col);
    const double2 b0 = col[0], bn = col[nz];
    for (int z = threadIdx.x + 1; z < nz; z += blockDim.x)
        col[z] = double2{col[z].x + b0.x * phim[z] + bn.x * phip[z], col[z].y + b0.y * phim[z] + bn.y * phip[z]};
    __syncthreads();
}

__device__ inline double2 muli(double2 a, double s) { return double2{-s * a.y, s * a.x}; }   // i s a

__global__ __launch_bounds__(64) void k_vor2vel_columns(const ColArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const ColArgs &a = tab[ib.item];
    const Grid &g = a.g;
    const int nz = g.nz, nz1 = nz + 1;
    const int col = ib.blk;
    const int kx = col % g.nxh, ky = col / g.nxh;
    const bool zero_mode = (kx == 0 && ky == 0);

    extern __shared__ double2 sm[];
    double2 *SX = sm, *SY = SX + nz1, *C = SY + nz1, *E = C + nz1, *DS = E + nz1, *AS = DS + nz1;
    double2 *fa = AS + nz1, *fb = fa + 2 * nz;
    double *phim = (double *)(fb + 2 * nz), *phip = phim + nz1;
    __shared__ double2 s_ds0, s_dsn, s_y0, s_yn, s_x0, s_xn;
    ColCtx cc{nz, fa, fb, &a.zplan, a.ztw};   // twiddles from global memory (cached)

    // wavenumbers (sta3dfft.f90:88-107, sta2dfft.f90:59-64): r = (pi/L) 2k
    const int kys = (2 * ky <= g.ny) ? ky : ky - g.ny;
    const double scx = EPIC_PI / g.extent[0], scy = EPIC_PI / g.extent[1], scz = EPIC_PI / g.extent[2];
    const double rkx = scx * (double)(2 * kx), rky = scy * (double)(2 * (kys < 0 ? -kys : kys));
    const double ikx = (2 * kx == g.nx) ? 0.0 : rkx;                       // diffx factor
    const double iky = (2 * ky == g.ny) ? 0.0 : (kys < 0 ? -rky : rky);   // diffy factor
    const double k2l2 = rkx * rkx + rky * rky;
    const double k2l2i = zero_mode ? 0.0 : 1.0 / k2l2;
    // Hou & Li filter (inversion_utils.f90:130-150)
    const double kxmaxi = 1.0 / (scx * (double)g.nx), kymaxi = 1.0 / (scy * (double)g.ny);
    const double kzmaxi = 1.0 / (scz * (double)nz);
    const double fh = exp(-36.0 * pow(kxmaxi * rkx, 36.0) + -36.0 * pow(kymaxi * rky, 36.0));
    // hyperbolic functions (inversion_utils.f90:305-351)
    const double Lz = g.extent[2];
    const double kl = sqrt(k2l2);
    const double ef = exp(-kl * Lz);
    const double dv = 1.0 / (1.0 - ef * ef);
    const double Q = dv * (1.0 + ef * ef), Rr = dv * 2.0 * ef, k2ifac = 0.5 * k2l2i;

    auto zcoord = [&](int iz, double &zm, double &zp) {
        const double z = g.lower[2] + g.dx[2] * (double)iz;
        zm = g.upper[2] - z;
        zp = z - g.lower[2];
    };

    // ---- load the three vorticity spectra and the phi tables ----
    const int64_t base = ((int64_t)ky * g.nxh + kx) * g.nzp;
    for (int z = threadIdx.x; z < nz1; z += blockDim.x) {
        const float2 x = gl(a.in[0])[base + z], y = gl(a.in[1])[base + z], w = gl(a.in[2])[base + z];
        SX[z] = double2{x.x, x.y};
        SY[z] = double2{y.x, y.y};
        C[z] = double2{w.x, w.y};
        double zm, zp;
        zcoord(z, zm, zp);
        if (zero_mode) {
            phim[z] = zm / Lz;
            phip[z] = zp / Lz;
        } else {
            const double ep = exp(-(kl * zp)), em = exp(-(kl * zm));
            phim[z] = dv * (ep - ef * em);
            phip[z] = dv * (em - ef * ep);
        }
    }
    __syncthreads();

    // ---- decompose + filter (inversion.f90:40-43) ----
    decompose(cc, SX, phim, phip);
    decompose(cc, SY, phim, phip);
    decompose(cc, C, phim, phip);
    for (int z = threadIdx.x; z < nz1; z += blockDim.x) {
        double f = fh;
        if (z > 0 && z < nz) f = fh * exp(-36.0 * pow(kzmaxi * (scz * (double)z), 36.0));
        if (zero_mode) f = 1.0;
        SX[z] = double2{f * SX[z].x, f * SX[z].y};
        SY[z] = double2{f * SY[z].x, f * SY[z].y};
        C[z] = double2{f * C[z].x, f * C[z].y};
    }
    __syncthreads();

    // ---- solenoidal projection (inversion.f90:50-87) ----
    combine(cc, C, phim, phip);   // zeta, semi-spectral
    {
        const double dzi = g.dxi[2], hdzi = 0.5 * g.dxi[2];
        for (int z = threadIdx.x; z < nz1; z += blockDim.x) {
            double2 d;
            if (z == 0) d = double2{dzi * (C[1].x - C[0].x), dzi * (C[1].y - C[0].y)};
            else if (z == nz) d = double2{dzi * (C[nz].x - C[nz - 1].x), dzi * (C[nz].y - C[nz - 1].y)};
            else d = double2{(C[z + 1].x - C[z - 1].x) * hdzi, (C[z + 1].y - C[z - 1].y) * hdzi};
            E[z] = d;
        }
        __syncthreads();
    }
    decompose(cc, E, phim, phip);
    if (!zero_mode) {
        for (int z = threadIdx.x; z < nz1; z += blockDim.x) {
            // D = B_x - A_y;  A = k2l2i (E_x + D_y), B = k2l2i (E_y - D_x)
            const double2 d = double2{muli(SY[z], ikx).x - muli(SX[z], iky).x, muli(SY[z], ikx).y - muli(SX[z], iky).y};
            const double2 ex = muli(E[z], ikx), ey = muli(E[z], iky), dx_ = muli(d, ikx), dy_ = muli(d, iky);
            SX[z] = double2{k2l2i * (ex.x + dy_.x), k2l2i * (ex.y + dy_.y)};
            SY[z] = double2{k2l2i * (ey.x - dx_.x), k2l2i * (ey.y - dx_.y)};
        }
    }
    __syncthreads();

    // ---- filtered solenoidal vorticity (inversion.f90:91-93) ----
    for (int z = threadIdx.x; z < nz1; z += blockDim.x) {
        DS[z] = SX[z];
        AS[z] = SY[z];
    }
    __syncthreads();
    combine(cc, DS, phim, phip);
    combine(cc, AS, phim, phip);
    for (int z = threadIdx.x; z < nz1; z += blockDim.x) {
        gl(a.out[0])[base + z] = float2{(float)DS[z].x, (float)DS[z].y};
        gl(a.out[1])[base + z] = float2{(float)AS[z].x, (float)AS[z].y};
        gl(a.out[2])[base + z] = float2{(float)C[z].x, (float)C[z].y};
    }
    __syncthreads();

    // ---- vertical velocity (inversion.f90:97-150) ----
    for (int z = threadIdx.x; z < nz1; z += blockDim.x) {
        const double2 dy_ = muli(SX[z], iky), dx_ = muli(SY[z], ikx);
        DS[z] = double2{dy_.x - dx_.x, dy_.y - dx_.y};
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        s_ds0 = DS[0];
        s_dsn = DS[nz];
    }
    __syncthreads();
    const double2 ds0 = s_ds0, dsn = s_dsn;
    for (int z = threadIdx.x; z < nz1; z += blockDim.x) {
        // es = ds(0) dtheta_-(z) + ds(nz) dtheta_+(z)
        double dtm = 0.0, dtp = 0.0;
        if (!zero_mode) {
            double zm, zp;
            zcoord(z, zm, zp);
            const double Lm = kl * zm, Lp = kl * zp;
            const double ep = exp(-Lp), em = exp(-Lm);
            const double pm = dv * (ep - ef * em), pp = dv * (em - ef * ep);
            const double dpm = -kl * dv * (ep + ef * em), dpp = kl * dv * (em + ef * ep);
            (void)pm;
            (void)pp;
            dtm = -k2ifac * ((Q * Lp - 1.0) * dpm - Rr * Lm * dpp);
            dtp = -k2ifac * ((Q * Lm - 1.0) * dpp - Rr * Lp * dpm);
        }
        E[z] = double2{ds0.x * dtm + dsn.x * dtp, ds0.y * dtm + dsn.y * dtp};
        if (z > 0 && z < nz) {
            const double rz = scz * (double)z;
            const double green = -1.0 / (k2l2 + rz * rz);
            DS[z] = double2{green * DS[z].x, green * DS[z].y};
            AS[z] = double2{rz * DS[z].x, rz * DS[z].y};
        } else {
            AS[z] = double2{0.0, 0.0};
        }
    }
    __syncthreads();
    zdct(cc, AS);
    zdst(cc, DS);
    for (int z = threadIdx.x; z < nz1; z += blockDim.x) {
        if (z == 0 || z == nz) {
            DS[z] = double2{0.0, 0.0};
        } else {
            double tm = 0.0, tp = 0.0;   // bs = ds(0) theta_-(z) + ds(nz) theta_+(z)
            if (!zero_mode) {
                double zm, zp;
                zcoord(z, zm, zp);
                const double Lm = kl * zm, Lp = kl * zp;
                const double ep = exp(-Lp), em = exp(-Lm);
                const double pm = dv * (ep - ef * em), pp = dv * (em - ef * ep);
                tm = k2ifac * (Rr * Lm * pp - Q * Lp * pm);
                tp = k2ifac * (Rr * Lp * pm - Q * Lm * pp);
            }
            DS[z] = double2{DS[z].x + (ds0.x * tm + dsn.x * tp), DS[z].y + (ds0.y * tm + dsn.y * tp)};
        }
        E[z] = double2{E[z].x + AS[z].x, E[z].y + AS[z].y};   // dw/dz
    }
    __syncthreads();

    // ---- horizontally averaged flow (inversion.f90:161-176) ----
    if (zero_mode) {
        if (threadIdx.x == 0) {
            s_y0 = SY[0];
            s_yn = SY[nz];
            s_x0 = SX[0];
            s_xn = SX[nz];
        }
        __syncthreads();
        for (int z = threadIdx.x; z < nz1; z += blockDim.x) {
            if (z == 0 || z == nz) {
                AS[z] = double2{0.0, 0.0};
                SX[z] = double2{0.0, 0.0};
            } else {
                const double rzi = 1.0 / (scz * (double)z);
                AS[z] = double2{-rzi * SY[z].x, -rzi * SY[z].y};   // ubar
                SX[z] = double2{rzi * SX[z].x, rzi * SX[z].y};     // vbar
            }
        }
        __syncthreads();
        zdct(cc, AS);
        zdct(cc, SX);
        for (int z = threadIdx.x; z < nz1; z += blockDim.x) {
            double zm, zp;
            zcoord(z, zm, zp);
            const double pt = zp / Lz;
            double zm2, zp2;
            zcoord(nz - z, zm2, zp2);
            const double pb = zp2 / Lz;
            const double gt = 0.5 * Lz * (pt * pt - 1.0 / 3.0);
            const double gb = 0.5 * Lz * (pb * pb - 1.0 / 3.0);
            AS[z] = double2{AS[z].x + s_yn.x * gt - s_y0.x * gb, AS[z].y + s_yn.y * gt - s_y0.y * gb};
            SX[z] = double2{SX[z].x - s_xn.x * gt + s_x0.x * gb, SX[z].y - s_xn.y * gt + s_x0.y * gb};
        }
        __syncthreads();
    }

    // ---- u, v, w and the velocity-gradient spectra (inversion.f90:180-233, 259-317) ----
    for (int z = threadIdx.x; z < nz1; z += blockDim.x) {
        double2 u, v;
        if (zero_mode) {
            u = AS[z];
            v = SX[z];
        } else {
            const double2 ex = muli(E[z], ikx), ey = muli(E[z], iky), cx = muli(C[z], ikx), cy = muli(C[z], iky);
            u = double2{k2l2i * (ex.x + cy.x), k2l2i * (ex.y + cy.y)};
            v = double2{k2l2i * (ey.x - cx.x), k2l2i * (ey.y - cx.y)};
        }
        const double2 w = DS[z];
        auto st = [&](int o, double2 val) { gl(a.out[o])[base + z] = float2{(float)val.x, (float)val.y}; };
        st(3, u);
        st(4, v);
        st(5, w);
        st(6, muli(u, ikx));    // du/dx
        st(7, muli(u, iky));    // du/dy
        st(8, muli(v, iky));    // dv/dy
        st(9, muli(w, ikx));    // dw/dx
        st(10, muli(w, iky));   // dw/dy
    }
}

// ---------------------------------------------------------------------------
// physical-space completion
// ---------------------------------------------------------------------------
// vor2vel: z halo of velocity (inversion.f90:241-248) and the derived
// gradient components (:278-313); one thread per (column, plane k =
// -1..nz+1).  A halo-plane thread forms its own halo values (linear
// extrapolation, or antisymmetry for w and its horizontal derivatives) from
// the interior planes before the derived components.
__global__ __launch_bounds__(256) void k_vor2vel_fixup(const FixArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const Grid &g = tab[ib.item].g;
    float *const fld = tab[ib.item].fld;
    const int nz = g.nz;
    const int64_t ncol = (int64_t)g.nx * g.ny;
    const int64_t tq = ib.blk * (int64_t)blockDim.x + threadIdx.x;
    if (tq >= ncol * (nz + 3)) return;
    const int k = (int)(tq / ncol) - 1;
    const int64_t col = tq - (int64_t)(k + 1) * ncol;
    const int64_t FS = field_stride(g), PS = plane_stride(g);
    auto F = [&](int slot, int kk) -> EPIC_GLOBAL float & { return gl(fld)[slot * FS + (int64_t)(kk + 1) * PS + col]; };
    float dwdx, dudy, dwdy;
    if (k == -1 || k == nz + 1) {
        const int kb = (k < 0) ? 0 : nz, kn = (k < 0) ? 1 : nz - 1;
        auto lin = [&](int slot) { return F(slot, k) = 2.0f * F(slot, kb) - F(slot, kn); };
        auto anti = [&](int slot) { return F(slot, k) = -F(slot, kn); };
        lin(F_VELOG + 0);
        lin(F_VELOG + 1);
        anti(F_VELOG + 2);
        lin(F_VELGRADG + DUDX);
        dudy = lin(F_VELGRADG + DUDY);
        dwdx = anti(F_VELGRADG + DWDX);
        lin(F_VELGRADG + DVDY);
        dwdy = anti(F_VELGRADG + DWDY);
    } else {
        dwdx = F(F_VELGRADG + DWDX, k);
        dudy = F(F_VELGRADG + DUDY, k);
        dwdy = F(F_VELGRADG + DWDY, k);
    }
    F(F_VELGRADG + DUDZ, k) = dwdx + F(F_VORTG + 1, k);
    F(F_VELGRADG + DVDX, k) = dudy + F(F_VORTG + 2, k);
    F(F_VELGRADG + DVDZ, k) = dwdy - F(F_VORTG + 0, k);
}

// vorticity_tendency: zeta walls and z extrapolation (inversion.f90:345-360)
__global__ __launch_bounds__(256) void k_vtend_fixup(const FixArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const Grid &g = tab[ib.item].g;
    float *const fld = tab[ib.item].fld;
    const int64_t col = ib.blk * (int64_t)blockDim.x + threadIdx.x;
    if (col >= (int64_t)g.nx * g.ny) return;
    const int64_t FS = field_stride(g), PS = plane_stride(g);
    const int nz = g.nz;
    auto F = [&](int slot, int k) -> EPIC_GLOBAL float & { return gl(fld)[slot * FS + (int64_t)(k + 1) * PS + col]; };
    // the transforms wrote the halo planes as the extrapolations 2 v(0) -
    // v(1) and 2 v(nz) - v(nz-1) of the unmodified tendency; where dzeta/dt
    // is held at zero on a wall, that extrapolation is -v(1) (-v(nz-1))
    if (g.zeta_zero[0]) {
        F(F_VTEND + 2, 0) = 0.0f;
        F(F_VTEND + 2, -1) = -F(F_VTEND + 2, 1);
    }
    if (g.zeta_zero[1]) {
        F(F_VTEND + 2, nz) = 0.0f;
        F(F_VTEND + 2, nz + 1) = -F(F_VTEND + 2, nz - 1);
    }
}

size_t col_lds_bytes(const Grid &g)
{
    const int nz1 = g.nz + 1;
    return (size_t)(6 * nz1 + 4 * g.nz) * sizeof(double2) + 2 * (size_t)nz1 * sizeof(double);
}

void add_vor2vel_columns(Plan &P, int phase, const std::vector<ColArgs> &items)
{
    std::vector<int> nb;
    size_t lds = 0;
    for (const ColArgs &a : items) {
        nb.push_back(a.g.ny * a.g.nxh);
        lds = std::max(lds, col_lds_bytes(a.g));
    }
    P.add((const void *)k_vor2vel_columns, 64, lds, phase, items, nb);
}

static std::vector<int> column_blocks(const std::vector<FixArgs> &items)
{
    std::vector<int> nb;
    for (const FixArgs &a : items) nb.push_back((int)(((int64_t)a.g.nx * a.g.ny + 255) / 256));
    return nb;
}

void add_vor2vel_fixup(Plan &P, int phase, const std::vector<FixArgs> &items)
{
    std::vector<int> nb;
    for (const FixArgs &a : items) nb.push_back((int)(((int64_t)a.g.nx * a.g.ny * (a.g.nz + 3) + 255) / 256));
    P.add((const void *)k_vor2vel_fixup, 256, 0, phase, items, nb);
}

void add_vtend_fixup(Plan &P, int phase, const std::vector<FixArgs> &items)
{
    P.add((const void *)k_vtend_fixup, 256, 0, phase, items, column_blocks(items));
}

}  // namespace epic
