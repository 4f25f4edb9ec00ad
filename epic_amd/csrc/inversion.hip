// Spectral vorticity inversion on gfx950 (inversion.f90:25-366,
// inversion_utils.f90:79-514).
//
// After the horizontal FFTs every horizontal operator of vor2vel is diagonal
// in (kx, ky), so the whole vertical part of the inversion is independent per
// spectral column.  k_vor2vel_columns gives each column to one wavefront and
// runs it in FP64, the column in registers (nz <= 191) or in LDS:
// decomposition into boundary values + sine series,
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

// columns of up to 64 COL_REG_SLOTS values (nz <= 191) keep their arrays in
// registers
constexpr int COL_REG_SLOTS = 3;

struct ColCtx {
    int nz;
    double2 *fa, *fb;
    const FftPlan *plan;
    const double2 *tw;   // e^{-2 pi i t / (2 nz)}, t < nz (the member's table; cached reads)
};

// One column of nz + 1 values per wave.  With NS > 0 the value at z = lane +
// 64 i sits in register slot i of the lane (NS slots, nz + 1 <= 64 NS): the
// column arrays then take no LDS, only the two FFT buffers do, and more
// columns share a CU.  NS == 0 keeps every column array in LDS (any nz).
template <int NS, typename T> struct ZCol {
    T r[NS > 0 ? NS : 1];
    T *l = nullptr;
    __device__ T &operator()(int i, int z)
    {
        if constexpr (NS > 0) return r[i];
        else return l[z];
    }
};

__device__ inline double read_lane(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
// the column's value at z (z uniform over the wave; with LDS storage, after
// a barrier that follows its last store)
template <int NS> __device__ __forceinline__ double2 zval(ZCol<NS, double2> &c, int z)
{
    if constexpr (NS > 0) {
        // every slot read from lane z mod 64, then the slot z / 64 chosen
        // among the scalar results (a select among the register slots
        // themselves would become an indexed access, which keeps the slots
        // in scratch memory)
        const int k = z >> 6, l = z & 63;
        double2 s = double2{0.0, 0.0};
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            const double2 v = double2{read_lane(c.r[i].x, l), read_lane(c.r[i].y, l)};
            if (k == i) s = v;
        }
        return s;
    } else {
        return c.l[z];
    }
}

// slots of this lane: z = lane + 64 i for i < nsl (z < nz + 1 checked by the caller)
#define ZSLOTS(i, z) \
    _Pragma("unroll") for (int i = 0; i < nsl; ++i) if (const int z = (int)threadIdx.x + 64 * i; true)

// DST-I of col[1..nz-1] in place (col[0], col[nz] untouched)
template <int NS> __device__ __forceinline__ void zdst(const ColCtx &c, ZCol<NS, double2> &col, int nsl)
{
    const int nz = c.nz, n2 = 2 * nz;
    ZSLOTS(i, z)
    {
        if (z > 0 && z < nz) {
            const double2 v = col(i, z);
            c.fa[z] = v;
            c.fa[n2 - z] = double2{-v.x, -v.y};   // the odd extension
        } else if (z == 0 || z == nz) {
            c.fa[z] = double2{0.0, 0.0};
        }
    }
    __syncthreads();
    double2 *R = lds_fft_tw(c.fa, c.fb, 1, *c.plan, TwHalf<double2>{c.tw, nz}, false);
    const double s = 0.5 * sqrt(2.0 / (double)nz);
    ZSLOTS(i, z)
    {
        if (z > 0 && z < nz) col(i, z) = double2{-s * R[z].y, s * R[z].x};   // s i R
    }
    __syncthreads();
}

// DCT-I of col[0..nz] in place
template <int NS> __device__ __forceinline__ void zdct(const ColCtx &c, ZCol<NS, double2> &col, int nsl)
{
    const int nz = c.nz, n2 = 2 * nz;
    ZSLOTS(i, z)
    {
        if (z <= nz) {
            const double2 v = col(i, z);
            c.fa[z] = v;
            if (z > 0 && z < nz) c.fa[n2 - z] = v;   // the even extension
        }
    }
    __syncthreads();
    double2 *R = lds_fft_tw(c.fa, c.fb, 1, *c.plan, TwHalf<double2>{c.tw, nz}, false);
    const double s = 0.5 * sqrt(2.0 / (double)nz);
    ZSLOTS(i, z)
    {
        if (z <= nz) col(i, z) = double2{s * R[z].x, s * R[z].y};
    }
    __syncthreads();
}

// field_decompose_semi_spectral (inversion_utils.f90:372-401)
template <int NS>
__device__ __forceinline__ void decompose(const ColCtx &c, ZCol<NS, double2> &col, ZCol<NS, double> &phim, ZCol<NS, double> &phip,
                          int nsl)
{
    const int nz = c.nz;
    const double2 b0 = zval(col, 0), bn = zval(col, nz);
    __syncthreads();
    ZSLOTS(i, z)
    {
        if (z > 0 && z < nz) {
            const double2 v = col(i, z);
            const double pm = phim(i, z), pp = phip(i, z);
            col(i, z) = double2{v.x - (b0.x * pm + bn.x * pp), v.y - (b0.y * pm + bn.y * pp)};
        }
    }
    __syncthreads();
    zdst(c, col, nsl);
}

// field_combine_semi_spectral (inversion_utils.f90:426-456)
template <int NS>
__device__ __forceinline__ void combine(const ColCtx &c, ZCol<NS, double2> &col, ZCol<NS, double> &phim, ZCol<NS, double> &phip,
                        int nsl)
{
    const int nz = c.nz;
    zdst(c, col, nsl);
    const double2 b0 = zval(col, 0), bn = zval(col, nz);
    ZSLOTS(i, z)
    {
        if (z > 0 && z < nz) {
            const double2 v = col(i, z);
            const double pm = phim(i, z), pp = phip(i, z);
            col(i, z) = double2{v.x + b0.x * pm + bn.x * pp, v.y + b0.y * pm + bn.y * pp};
        }
    }
    __syncthreads();
}

__device__ inline double2 muli(double2 a, double s) { return double2{-s * a.y, s * a.x}; }   // i s a

// (the slot loops of the LDS form, NS == 0, have a run-time trip count: no
// full unrolling there)
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wpass-failed"
template <int NS>
__global__ __launch_bounds__(64) void k_vor2vel_columns(const ColArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const ColArgs &a = tab[ib.item];
    const Grid &g = a.g;
    const int nz = g.nz, nz1 = nz + 1;
    const int nsl = NS > 0 ? NS : (nz1 + 63) / 64;
    const int col = ib.blk;
    const int kx = col % g.nxh, ky = col / g.nxh;
    const bool zero_mode = (kx == 0 && ky == 0);

    extern __shared__ double2 sm[];
    ZCol<NS, double2> SX, SY, C, E, DS, AS;
    ZCol<NS, double> phim, phip;
    double2 *fa, *fb;
    if constexpr (NS > 0) {
        fa = sm;
        fb = fa + 2 * nz;
    } else {
        SX.l = sm;
        SY.l = SX.l + nz1;
        C.l = SY.l + nz1;
        E.l = C.l + nz1;
        DS.l = E.l + nz1;
        AS.l = DS.l + nz1;
        fa = AS.l + nz1;
        fb = fa + 2 * nz;
        phim.l = (double *)(fb + 2 * nz);
        phip.l = phim.l + nz1;
    }
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
    ZSLOTS(i, z)
    {
        if (z < nz1) {
            const float2 x = gl(a.in[0])[base + z], y = gl(a.in[1])[base + z], w = gl(a.in[2])[base + z];
            SX(i, z) = double2{x.x, x.y};
            SY(i, z) = double2{y.x, y.y};
            C(i, z) = double2{w.x, w.y};
            double zm, zp;
            zcoord(z, zm, zp);
            if (zero_mode) {
                phim(i, z) = zm / Lz;
                phip(i, z) = zp / Lz;
            } else {
                const double ep = exp(-(kl * zp)), em = exp(-(kl * zm));
                phim(i, z) = dv * (ep - ef * em);
                phip(i, z) = dv * (em - ef * ep);
            }
        }
    }
    __syncthreads();

    // ---- decompose + filter (inversion.f90:40-43) ----
    decompose(cc, SX, phim, phip, nsl);
    decompose(cc, SY, phim, phip, nsl);
    decompose(cc, C, phim, phip, nsl);
    ZSLOTS(i, z)
    {
        if (z < nz1) {
            double f = fh;
            if (z > 0 && z < nz) f = fh * exp(-36.0 * pow(kzmaxi * (scz * (double)z), 36.0));
            if (zero_mode) f = 1.0;
            const double2 sx = SX(i, z), sy = SY(i, z), cz = C(i, z);
            SX(i, z) = double2{f * sx.x, f * sx.y};
            SY(i, z) = double2{f * sy.x, f * sy.y};
            C(i, z) = double2{f * cz.x, f * cz.y};
        }
    }
    __syncthreads();

    // ---- solenoidal projection (inversion.f90:50-87) ----
    combine(cc, C, phim, phip, nsl);   // zeta, semi-spectral
    {
        // neighbours through the FFT buffer
        ZSLOTS(i, z)
        {
            if (z < nz1) fa[z] = C(i, z);
        }
        __syncthreads();
        const double dzi = g.dxi[2], hdzi = 0.5 * g.dxi[2];
        ZSLOTS(i, z)
        {
            if (z < nz1) {
                double2 d;
                if (z == 0) d = double2{dzi * (fa[1].x - fa[0].x), dzi * (fa[1].y - fa[0].y)};
                else if (z == nz) d = double2{dzi * (fa[nz].x - fa[nz - 1].x), dzi * (fa[nz].y - fa[nz - 1].y)};
                else d = double2{(fa[z + 1].x - fa[z - 1].x) * hdzi, (fa[z + 1].y - fa[z - 1].y) * hdzi};
                E(i, z) = d;
            }
        }
        __syncthreads();
    }
    decompose(cc, E, phim, phip, nsl);
    if (!zero_mode) {
        ZSLOTS(i, z)
        {
            if (z < nz1) {
                // D = B_x - A_y;  A = k2l2i (E_x + D_y), B = k2l2i (E_y - D_x)
                const double2 sx = SX(i, z), sy = SY(i, z), e = E(i, z);
                const double2 d = double2{muli(sy, ikx).x - muli(sx, iky).x, muli(sy, ikx).y - muli(sx, iky).y};
                const double2 ex = muli(e, ikx), ey = muli(e, iky), dx_ = muli(d, ikx), dy_ = muli(d, iky);
                SX(i, z) = double2{k2l2i * (ex.x + dy_.x), k2l2i * (ex.y + dy_.y)};
                SY(i, z) = double2{k2l2i * (ey.x - dx_.x), k2l2i * (ey.y - dx_.y)};
            }
        }
    }
    __syncthreads();

    // ---- filtered solenoidal vorticity (inversion.f90:91-93) ----
    ZSLOTS(i, z)
    {
        if (z < nz1) {
            DS(i, z) = SX(i, z);
            AS(i, z) = SY(i, z);
        }
    }
    __syncthreads();
    combine(cc, DS, phim, phip, nsl);
    combine(cc, AS, phim, phip, nsl);
    ZSLOTS(i, z)
    {
        if (z < nz1) {
            const double2 ds = DS(i, z), as = AS(i, z), cz = C(i, z);
            gl(a.out[0])[base + z] = float2{(float)ds.x, (float)ds.y};
            gl(a.out[1])[base + z] = float2{(float)as.x, (float)as.y};
            gl(a.out[2])[base + z] = float2{(float)cz.x, (float)cz.y};
        }
    }
    __syncthreads();

    // ---- vertical velocity (inversion.f90:97-150) ----
    ZSLOTS(i, z)
    {
        if (z < nz1) {
            const double2 dy_ = muli(SX(i, z), iky), dx_ = muli(SY(i, z), ikx);
            DS(i, z) = double2{dy_.x - dx_.x, dy_.y - dx_.y};
        }
    }
    __syncthreads();
    const double2 ds0 = zval(DS, 0), dsn = zval(DS, nz);
    __syncthreads();
    ZSLOTS(i, z)
    {
        if (z < nz1) {
            // es = ds(0) dtheta_-(z) + ds(nz) dtheta_+(z)
            double dtm = 0.0, dtp = 0.0;
            if (!zero_mode) {
                double zm, zp;
                zcoord(z, zm, zp);
                const double Lm = kl * zm, Lp = kl * zp;
                const double ep = exp(-Lp), em = exp(-Lm);
                const double dpm = -kl * dv * (ep + ef * em), dpp = kl * dv * (em + ef * ep);
                dtm = -k2ifac * ((Q * Lp - 1.0) * dpm - Rr * Lm * dpp);
                dtp = -k2ifac * ((Q * Lm - 1.0) * dpp - Rr * Lp * dpm);
            }
            E(i, z) = double2{ds0.x * dtm + dsn.x * dtp, ds0.y * dtm + dsn.y * dtp};
            if (z > 0 && z < nz) {
                const double rz = scz * (double)z;
                const double green = -1.0 / (k2l2 + rz * rz);
                const double2 ds = DS(i, z);
                const double2 gd = double2{green * ds.x, green * ds.y};
                DS(i, z) = gd;
                AS(i, z) = double2{rz * gd.x, rz * gd.y};
            } else {
                AS(i, z) = double2{0.0, 0.0};
            }
        }
    }
    __syncthreads();
    zdct(cc, AS, nsl);
    zdst(cc, DS, nsl);
    ZSLOTS(i, z)
    {
        if (z < nz1) {
            if (z == 0 || z == nz) {
                DS(i, z) = double2{0.0, 0.0};
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
                const double2 ds = DS(i, z);
                DS(i, z) = double2{ds.x + (ds0.x * tm + dsn.x * tp), ds.y + (ds0.y * tm + dsn.y * tp)};
            }
            const double2 e = E(i, z), as = AS(i, z);
            E(i, z) = double2{e.x + as.x, e.y + as.y};   // dw/dz
        }
    }
    __syncthreads();

    // ---- horizontally averaged flow (inversion.f90:161-176) ----
    if (zero_mode) {
        const double2 y0 = zval(SY, 0), yn = zval(SY, nz), x0 = zval(SX, 0), xn = zval(SX, nz);
        __syncthreads();
        ZSLOTS(i, z)
        {
            if (z < nz1) {
                if (z == 0 || z == nz) {
                    AS(i, z) = double2{0.0, 0.0};
                    SX(i, z) = double2{0.0, 0.0};
                } else {
                    const double rzi = 1.0 / (scz * (double)z);
                    const double2 sy = SY(i, z), sx = SX(i, z);
                    AS(i, z) = double2{-rzi * sy.x, -rzi * sy.y};   // ubar
                    SX(i, z) = double2{rzi * sx.x, rzi * sx.y};     // vbar
                }
            }
        }
        __syncthreads();
        zdct(cc, AS, nsl);
        zdct(cc, SX, nsl);
        ZSLOTS(i, z)
        {
            if (z < nz1) {
                double zm, zp;
                zcoord(z, zm, zp);
                const double pt = zp / Lz;
                double zm2, zp2;
                zcoord(nz - z, zm2, zp2);
                const double pb = zp2 / Lz;
                const double gt = 0.5 * Lz * (pt * pt - 1.0 / 3.0);
                const double gb = 0.5 * Lz * (pb * pb - 1.0 / 3.0);
                const double2 as = AS(i, z), sx = SX(i, z);
                AS(i, z) = double2{as.x + yn.x * gt - y0.x * gb, as.y + yn.y * gt - y0.y * gb};
                SX(i, z) = double2{sx.x - xn.x * gt + x0.x * gb, sx.y - xn.y * gt + x0.y * gb};
            }
        }
        __syncthreads();
    }

    // ---- u, v, w and the velocity-gradient spectra (inversion.f90:180-233, 259-317) ----
    ZSLOTS(i, z)
    {
        if (z < nz1) {
            double2 u, v;
            if (zero_mode) {
                u = AS(i, z);
                v = SX(i, z);
            } else {
                const double2 e = E(i, z), cz = C(i, z);
                const double2 ex = muli(e, ikx), ey = muli(e, iky), cx = muli(cz, ikx), cy = muli(cz, iky);
                u = double2{k2l2i * (ex.x + cy.x), k2l2i * (ex.y + cy.y)};
                v = double2{k2l2i * (ey.x - cx.x), k2l2i * (ey.y - cx.y)};
            }
            const double2 w = DS(i, z);
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
}
#pragma clang diagnostic pop
#undef ZSLOTS

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

// register slots per lane for a column of nz + 1 values (0: the LDS form)
static int col_slots(const Grid &g)
{
    const int ns = (g.nz + 1 + 63) / 64;
    return ns <= COL_REG_SLOTS ? ns : 0;
}

size_t col_lds_bytes(const Grid &g)
{
    const int nz1 = g.nz + 1;
    const size_t fft = (size_t)4 * g.nz * sizeof(double2);   // the two FFT buffers of 2 nz points
    if (col_slots(g) > 0) return fft;
    return (size_t)6 * nz1 * sizeof(double2) + fft + 2 * (size_t)nz1 * sizeof(double);
}

// one launch per column form present (members of one batch may differ in nz)
void add_vor2vel_columns(Plan &P, int phase, const std::vector<ColArgs> &items)
{
    static const void *const kern[COL_REG_SLOTS + 1] = {(const void *)k_vor2vel_columns<0>,
                                                        (const void *)k_vor2vel_columns<1>,
                                                        (const void *)k_vor2vel_columns<2>,
                                                        (const void *)k_vor2vel_columns<3>};
    for (int ns = 0; ns <= COL_REG_SLOTS; ++ns) {
        std::vector<ColArgs> sel;
        std::vector<int> nb;
        size_t lds = 0;
        for (const ColArgs &a : items) {
            if (col_slots(a.g) != ns) continue;
            sel.push_back(a);
            nb.push_back(a.g.ny * a.g.nxh);
            lds = std::max(lds, col_lds_bytes(a.g));
        }
        if (!sel.empty()) P.add(kern[ns], 64, lds, phase, sel, nb);
    }
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
