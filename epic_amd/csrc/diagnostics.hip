// Parcel and field diagnostics (parcel_diagnostics.f90:51-159,
// field_diagnostics.f90:40-154) on the device, one rank, ape_calculation =
// 'none'.  Every quantity is a sum, minimum or maximum; sums are formed in
// FP64 per block (tree in LDS) and then over the blocks in block order by
// one block, so the results are reproducible.
#include "common.h"
#include "ellipsoid.h"
#include "launch.h"

#include <algorithm>

namespace epic {

namespace {

constexpr int DT = 256;
enum Op { SUM, MAX, MIN };

template <int NV>
__device__ inline void block_reduce(double v[NV], const int (&op)[NV], double (*s)[DT])
{
#pragma unroll
    for (int c = 0; c < NV; ++c) s[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int o = DT / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
#pragma unroll
            for (int c = 0; c < NV; ++c) {
                const double a = s[c][threadIdx.x], b = s[c][threadIdx.x + o];
                s[c][threadIdx.x] = (op[c] == SUM) ? a + b : (op[c] == MAX) ? fmax(a, b) : fmin(a, b);
            }
        }
        __syncthreads();
    }
}

__device__ inline double identity(int op)
{
    return op == SUM ? 0.0 : (op == MAX ? -1.0e300 : 1.0e300);
}

// parcels: ke, enstrophy, sum lam, sum lam^2, sum vol, sum vol^2, n_small,
// vol xi^2, vol eta^2, vol zeta^2 (sums); buoyancy max, min
constexpr int NP = 12;
__constant__ int c_pop[NP] = {SUM, SUM, SUM, SUM, SUM, SUM, SUM, SUM, SUM, SUM, MAX, MIN};

// src (may be null): the slot holding each parcel's tendencies (after a
// re-sort that left them in place); a grid-stride loop, so any number of
// blocks covers the parcels (one parcel per thread when nblocks(n) blocks)
__global__ __launch_bounds__(DT) void k_pdiag_partial(Grid g, ParcelsDev p, const int *src, double *part, int *err)
{
    __shared__ double s[NP][DT];
    int op[NP];
#pragma unroll
    for (int c = 0; c < NP; ++c) op[c] = c_pop[c];
    double v[NP];
#pragma unroll
    for (int c = 0; c < NP; ++c) v[c] = identity(op[c]);
    for (int64_t n = (int64_t)blockIdx.x * DT + threadIdx.x; n < p.n; n += (int64_t)gridDim.x * DT) {
        const int64_t t = src ? (int64_t)src[n] : n;
        const double vol = p.vol[n];
        const double u0 = p.dpos[0][t], u1 = p.dpos[1][t], u2 = p.dpos[2][t];
        const double w0 = p.vor[0][n], w1 = p.vor[1][n], w2 = p.vor[2][n];
        const double B[5] = {p.B[0][n], p.B[1][n], p.B[2][n], p.B[3][n], p.B[4][n]};
        double b33;
        if (!b33_of(B, vol, b33)) {
            atomicOr(err, ERR_B33);
            b33 = 0.0;
        }
        const double A[9] = {B[0], B[1], B[2], B[1], B[3], B[4], B[2], B[4], b33};
        double D[3], V[3][3];
        scherzinger(A, D, V);
        const double lam = sqrt(D[0] / D[2]);   // get_aspect_ratio (parcel_ellipsoid.f90:423-429)
        const double x[NP] = {(u0 * u0 + u1 * u1 + u2 * u2) * vol,
                              (w0 * w0 + w1 * w1 + w2 * w2) * vol,
                              lam,
                              lam * lam,
                              vol,
                              vol * vol,
                              (vol <= g.vmin) ? 1.0 : 0.0,
                              vol * w0 * w0,
                              vol * w1 * w1,
                              vol * w2 * w2,
                              (double)p.buoy[n],
                              (double)p.buoy[n]};
#pragma unroll
        for (int c = 0; c < NP; ++c) v[c] = (op[c] == SUM) ? v[c] + x[c] : (op[c] == MAX) ? fmax(v[c], x[c]) : fmin(v[c], x[c]);
    }
    block_reduce<NP>(v, op, s);
    if (threadIdx.x < NP) part[(int64_t)blockIdx.x * NP + threadIdx.x] = s[threadIdx.x][0];
}

// nodes (planes 0..nz): (volg - vcell)^2, nparg, nsparg (planes < nz),
// weighted volg |u|^2, weighted volg |w|^2 (sums); |volg - vcell|, nparg,
// tbuoyg (max); nparg, tbuoyg (min)
constexpr int NF = 10;
__constant__ int c_fop[NF] = {SUM, SUM, SUM, SUM, SUM, MAX, MAX, MAX, MIN, MIN};

__global__ __launch_bounds__(DT) void k_fdiag_partial(Grid g, const float *fld, const int *nparg, const int *nsparg,
                                                      double *part)
{
    __shared__ double s[NF][DT];
    int op[NF];
#pragma unroll
    for (int c = 0; c < NF; ++c) op[c] = c_fop[c];
    double v[NF];
#pragma unroll
    for (int c = 0; c < NF; ++c) v[c] = identity(op[c]);
    const int64_t nn = (int64_t)g.nx * g.ny * (g.nz + 1);
    const int64_t q = (int64_t)blockIdx.x * DT + threadIdx.x;
    if (q < nn) {
        const int64_t PS = plane_stride(g), FS = field_stride(g);
        const int k = (int)(q / PS);
        const int64_t o = q + PS;   // plane k + 1 of the stored layout
        const double vol = fld[F_VOLG * FS + o];
        const double dv = vol - g.vcell;
        double u2 = 0.0, w2 = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double u = fld[(F_VELOG + c) * FS + o], w = fld[(F_VORTG + c) * FS + o];
            u2 += u * u;
            w2 += w * w;
        }
        const double wt = (k == 0 || k == g.nz) ? 0.25 : 0.5;   // half weights on the boundary planes
        const double b = fld[F_TBUOYG * FS + o];
        v[0] = dv * dv;
        v[3] = wt * vol * u2;
        v[4] = wt * vol * w2;
        v[5] = fabs(dv);
        v[7] = b;
        v[9] = b;
        if (k < g.nz) {
            v[1] = nparg[o];
            v[2] = nsparg[o];
            v[6] = nparg[o];
            v[8] = nparg[o];
        }
    }
    block_reduce<NF>(v, op, s);
    if (threadIdx.x < NF) part[(int64_t)blockIdx.x * NF + threadIdx.x] = s[threadIdx.x][0];
}

// one block: the partials of nb blocks in block order per thread, then the
// tree; the operations are those of the parcel (NV == NP) or field table
template <int NV>
__global__ __launch_bounds__(DT) void k_diag_total(const double *part, int nb, double *out)
{
    __shared__ double s[NV][DT];
    int op[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) op[c] = (NV == NP) ? c_pop[c] : c_fop[c];
    double v[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) v[c] = identity(op[c]);
    for (int b = threadIdx.x; b < nb; b += DT)
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            const double x = part[(int64_t)b * NV + c];
            v[c] = (op[c] == SUM) ? v[c] + x : (op[c] == MAX) ? fmax(v[c], x) : fmin(v[c], x);
        }
    block_reduce<NV>(v, op, s);
    if (threadIdx.x < NV) out[threadIdx.x] = s[threadIdx.x][0];
}

inline int nblocks(int64_t n) { return (int)std::max<int64_t>(1, (n + DT - 1) / DT); }

}  // namespace

int64_t diag_scratch_doubles(const Grid &g, int64_t cap)
{
    const int64_t nn = (int64_t)g.nx * g.ny * (g.nz + 1);
    return std::max<int64_t>((int64_t)nblocks(cap) * NP, (int64_t)nblocks(nn) * NF) + NP + NF;
}

void launch_parcel_diagnostics(const Grid &g, const ParcelsDev &p, const int *src, double *scratch, double *out, int *err,
                               hipStream_t s, int max_blocks)
{
    const int nb = max_blocks > 0 ? std::min(nblocks(p.n), max_blocks) : nblocks(p.n);
    hipLaunchKernelGGL(k_pdiag_partial, dim3(nb), dim3(DT), 0, s, g, p, src, scratch, err);
    hipLaunchKernelGGL(k_diag_total<NP>, dim3(1), dim3(DT), 0, s, scratch, nb, out);
}

void launch_field_diagnostics(const Grid &g, const float *fld, const int *nparg, const int *nsparg, double *scratch,
                              double *out, hipStream_t s)
{
    const int nb = nblocks((int64_t)g.nx * g.ny * (g.nz + 1));
    hipLaunchKernelGGL(k_fdiag_partial, dim3(nb), dim3(DT), 0, s, g, fld, nparg, nsparg, scratch);
    hipLaunchKernelGGL(k_diag_total<NF>, dim3(1), dim3(DT), 0, s, scratch, nb, out);
}

}  // namespace epic
