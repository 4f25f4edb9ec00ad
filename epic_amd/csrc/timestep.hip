// get_time_step (rk_utils.f90:73-183) on the device.  For an adaptive step
// every node of planes 0..nz takes the largest eigenvalue magnitude of its
// symmetrised velocity gradient (get_strain, :30-66, FP64 Scherzinger
// eigenvalues) and the squared central-difference gradient of tbuoyg (x and y
// periodic, z over planes -1..nz+1); block maxima go to the member's
// StepTime by 64-bit atomic max on the bit patterns (non-negative doubles
// order like their bits, and a maximum does not depend on the order).  One
// thread then forms dt = min(alpha/gmax, alpha/bmax), gmax and bmax floored
// at epsilon, bmax the fourth root of the squared maximum, cut at the time
// limit with precise_stop; dt <= 0 raises the error of :179-182.
#include "common.h"
#include "ellipsoid.h"
#include "launch.h"

namespace epic {

__global__ __launch_bounds__(256) void k_ts_max(const TsArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const TsArgs &a = tab[ib.item];
    if (!a.tm->adaptive) return;   // uniform over the item's blocks
    const Grid &g = a.g;
    const int64_t nn = (int64_t)g.nx * g.ny * (g.nz + 1);
    const int64_t q = (int64_t)ib.blk * 256 + threadIdx.x;
    double gm = 0.0, b2 = 0.0;
    if (q < nn) {
        const int i = (int)(q % g.nx);
        const int j = (int)((q / g.nx) % g.ny);
        const int k = (int)(q / ((int64_t)g.nx * g.ny));
        const int64_t FS = field_stride(g), PS = plane_stride(g);
        const int64_t o = (int64_t)(k + 1) * PS + (int64_t)j * g.nx + i;
        double v[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) v[c] = (double)gl(a.fld)[(F_VELGRADG + c) * FS + o];
        const double A[9] = {v[DUDX], 0.5 * (v[DUDY] + v[DVDX]), 0.5 * (v[DUDZ] + v[DWDX]),
                             0.5 * (v[DUDY] + v[DVDX]), v[DVDY], 0.5 * (v[DVDZ] + v[DWDY]),
                             0.5 * (v[DUDZ] + v[DWDX]), 0.5 * (v[DVDZ] + v[DWDY]), -(v[DUDX] + v[DVDY])};
        double D[3], V[3][3];
        scherzinger(A, D, V);
        gm = fmax(fabs(D[0]), fmax(fabs(D[1]), fabs(D[2])));
        const EPIC_GLOBAL float *b = gl(a.fld) + F_TBUOYG * FS;
        const int ip = (i + 1 == g.nx) ? 0 : i + 1, im = (i == 0) ? g.nx - 1 : i - 1;
        const int jp = (j + 1 == g.ny) ? 0 : j + 1, jm = (j == 0) ? g.ny - 1 : j - 1;
        const int64_t ro = (int64_t)(k + 1) * PS;
        const double gx = 0.5 * g.dxi[0] * ((double)b[ro + (int64_t)j * g.nx + ip] - (double)b[ro + (int64_t)j * g.nx + im]);
        const double gy = 0.5 * g.dxi[1] * ((double)b[ro + (int64_t)jp * g.nx + i] - (double)b[ro + (int64_t)jm * g.nx + i]);
        const double gz = 0.5 * g.dxi[2] * ((double)b[o + PS] - (double)b[o - PS]);
        b2 = gx * gx + gy * gy + gz * gz;
    }
    // block maxima: waves by shuffles, then one atomic per wave
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        gm = fmax(gm, __shfl_xor(gm, off, 64));
        b2 = fmax(b2, __shfl_xor(b2, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&a.tm->gmax_bits, (unsigned long long)__double_as_longlong(gm));
        atomicMax(&a.tm->b2max_bits, (unsigned long long)__double_as_longlong(b2));
    }
}

__global__ void k_ts_final(const TsArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const TsArgs &a = tab[ib.item];
    if (threadIdx.x != 0 || !a.tm->adaptive) return;
    StepTime *tm = a.tm;
    const double eps = 2.220446049250313e-16;   // epsilon(1d0)
    const double gmax = fmax(eps, __longlong_as_double((long long)tm->gmax_bits));
    const double bmax = fmax(eps, sqrt(sqrt(__longlong_as_double((long long)tm->b2max_bits))));
    double dt = fmin(tm->alpha / gmax, tm->alpha / bmax);
    if (tm->precise_stop && (tm->t + dt > tm->limit)) dt = tm->limit - tm->t;
    if (!(dt > 0.0)) {
        atomicOr(a.err, ERR_DT);
        dt = 0.0;
    }
    tm->dt = dt;
    tm->gmax = gmax;
    tm->bmax = bmax;
    tm->gmax_bits = 0ull;   // ready for the next step
    tm->b2max_bits = 0ull;
}

void add_time_step(Plan &P, int phase, const std::vector<TsArgs> &items)
{
    std::vector<int> nb, one;
    for (const TsArgs &a : items) {
        nb.push_back((int)(((int64_t)a.g.nx * a.g.ny * (a.g.nz + 1) + 255) / 256));
        one.push_back(1);
    }
    P.add((const void *)k_ts_max, 256, 0, phase, items, nb);
    P.add((const void *)k_ts_final, 64, 0, phase, items, one);
}

}  // namespace epic
