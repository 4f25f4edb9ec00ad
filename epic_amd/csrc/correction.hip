// Parcel correction (parcel_correction.f90) on the device.
//
// apply_laplace (:137-172): phi = volg / vcell - 1 on the nodes, its
// horizontal FFT (the step's FP32 transforms), then per spectral column in
// FP64 the tridiagonal inversion of the semi-spectral Laplacian (lapinv1,
// inversion_utils.f90:565-589, tables :240-271) and the three gradient
// spectra (diverge, inversion.f90:410-455: i kx, i ky, 2nd-order z
// differences zero at the walls); after the inverse FFTs every parcel moves
// by grad phi interpolated at its centre.
// apply_gradient (:175-258): the limited gradient step from phi on the eight
// nodes around each parcel centre, in FP64.
// apply_vortcor (:91-133): the volume-weighted mean vorticity, reduced in a
// fixed order in FP64, reset to a target.
#include "common.h"
#include "launch.h"

#include <algorithm>

namespace epic {

namespace {
__device__ inline int64_t gtid() { return blockIdx.x * (int64_t)blockDim.x + threadIdx.x; }
inline unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }
}  // namespace

// one thread per (kx, ky) column; in: phi spectrum; fr, fi, et: FP64
// scratch of (nz+1) ncol entries each, plane-major so that neighbouring
// columns touch neighbouring words; out: the x, y, z gradient spectra
__global__ void k_lapinv_columns(Grid g, const float2 *in, double *fr, double *fi, double *et, float2 *ox, float2 *oy,
                                 float2 *oz)
{
    const int64_t col = gtid();
    const int ncol = g.ny * g.nxh;
    if (col >= ncol) return;
    const int kx = (int)(col % g.nxh), ky = (int)(col / g.nxh);
    const int nz = g.nz;
    const int64_t base = col * (int64_t)g.nzp;
    if (kx == 0 && ky == 0) {
        for (int z = 0; z <= nz; ++z) ox[base + z] = oy[base + z] = oz[base + z] = float2{0.0f, 0.0f};
        return;
    }
    const int kys = (2 * ky <= g.ny) ? ky : ky - g.ny;
    const double rkx = 2.0 * 3.14159265358979323846 * kx / g.extent[0];
    const double rky = 2.0 * 3.14159265358979323846 * kys / g.extent[1];
    const double ikx = (2 * kx == g.nx) ? 0.0 : rkx;     // diffx: zero at the x Nyquist
    const double iky = (2 * ky == g.ny) ? 0.0 : rky;     // diffy: zero at the y Nyquist
    const double k2l2 = rkx * rkx + rky * rky;
    const double dzisq = g.dxi[2] * g.dxi[2], ap = dzisq;
    const double a0 = -2.0 * dzisq - k2l2;
    fr += col;
    fi += col;
    et += col;
    const int64_t S = ncol;   // plane stride of the scratch
    // forward elimination (htdv / etdv on the fly)
    double h = 1.0 / a0;
    double e = -2.0 * ap * h;
    float2 r = in[base];
    fr[0] = (double)r.x * h;
    fi[0] = (double)r.y * h;
    et[0] = e;
    for (int z = 1; z <= nz - 1; ++z) {
        h = 1.0 / (a0 + ap * e);
        r = in[base + z];
        fr[z * S] = ((double)r.x - ap * fr[(z - 1) * S]) * h;
        fi[z * S] = ((double)r.y - ap * fi[(z - 1) * S]) * h;
        e = -ap * h;
        et[z * S] = e;
    }
    h = 1.0 / (a0 + 2.0 * ap * e);
    r = in[base + nz];
    fr[nz * S] = ((double)r.x - 2.0 * ap * fr[(nz - 1) * S]) * h;
    fi[nz * S] = ((double)r.y - 2.0 * ap * fi[(nz - 1) * S]) * h;
    // back substitution
    for (int z = nz - 1; z >= 0; --z) {
        fr[z * S] = et[z * S] * fr[(z + 1) * S] + fr[z * S];
        fi[z * S] = et[z * S] * fi[(z + 1) * S] + fi[z * S];
    }
    // gradient spectra: i kx f, i ky f, central z differences (0 at the walls)
    const double hdzi = 0.5 * g.dxi[2];
    for (int z = 0; z <= nz; ++z) {
        ox[base + z] = float2{(float)(-ikx * fi[z * S]), (float)(ikx * fr[z * S])};
        oy[base + z] = float2{(float)(-iky * fi[z * S]), (float)(iky * fr[z * S])};
        if (z == 0 || z == nz) oz[base + z] = float2{0.0f, 0.0f};
        else
            oz[base + z] = float2{(float)((fr[(z + 1) * S] - fr[(z - 1) * S]) * hdzi),
                                  (float)((fi[(z + 1) * S] - fi[(z - 1) * S]) * hdzi)};
    }
}

// z halo of the gradient fields: linear extrapolation
__global__ void k_grad_halo(Grid g, float *fld)
{
    const int64_t col = gtid();
    if (col >= (int64_t)g.nx * g.ny) return;
    const int64_t FS = field_stride(g), PS = plane_stride(g);
    const int nz = g.nz;
    for (int c = 0; c < 3; ++c) {
        float *f = fld + (F_GRAD + c) * FS + col;
        f[0] = 2.0f * f[PS] - f[2 * PS];
        f[(int64_t)(nz + 2) * PS] = 2.0f * f[(int64_t)(nz + 1) * PS] - f[(int64_t)nz * PS];
    }
}

__device__ inline void centre_weights(const Grid &g, const double x[3], int ijk[3], double w[8])
{   // trilinear (interpl.f90:17-56): weights w[kk*4 + jj*2 + ii]
    double p[3], pc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double xyz = (x[a] - g.lower[a]) * g.dxi[a];
        const double fl = floor(xyz);
        ijk[a] = (int)fl;
        p[a] = xyz - fl;
        pc[a] = 1.0 - p[a];
    }
#pragma unroll
    for (int kk = 0; kk < 2; ++kk)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii)
                w[kk * 4 + jj * 2 + ii] = (kk ? p[2] : pc[2]) * (jj ? p[1] : pc[1]) * (ii ? p[0] : pc[0]);
}

__device__ inline int64_t node_off(const Grid &g, int k, int j, int i)
{
    return (int64_t)(k + 1) * plane_stride(g) + (int64_t)wrap(j, g.ny) * g.nx + wrap(i, g.nx);
}

__global__ void k_laplace_push(Grid g, ParcelsDev p, const float *fld, int *err)
{
    const int64_t n = gtid();
    if (n >= p.n) return;
    double x[3] = {p.x[n], p.y[n], p.z[n]};
    int ijk[3];
    double w[8];
    centre_weights(g, x, ijk, w);
    if (ijk[2] < -1 || ijk[2] > g.nz) {
        atomicOr(err, ERR_OUT_OF_GRID);
        return;
    }
    const int64_t FS = field_stride(g);
    // sum over the 8 corners in Fortran element order (k fastest)
    static constexpr int ord[8] = {0, 4, 2, 6, 1, 5, 3, 7};
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        const float *f = fld + (F_GRAD + l) * FS;
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int c = ord[m];
            s += w[c] * (double)f[node_off(g, ijk[2] + (c >> 2), ijk[1] + ((c >> 1) & 1), ijk[0] + (c & 1))];
        }
        x[l] = x[l] + s;
    }
    x[0] = x[0] - g.extent[0] * trunc((x[0] - g.center[0]) * g.hli[0]);
    x[1] = x[1] - g.extent[1] * trunc((x[1] - g.center[1]) * g.hli[1]);
    p.x[n] = x[0];
    p.y[n] = x[1];
    p.z[n] = x[2];
}

__global__ void k_gradient_push(Grid g, ParcelsDev p, const float *fld, double prefactor, double max_compression,
                                int *err)
{
    const int64_t n = gtid();
    if (n >= p.n) return;
    double x[3] = {p.x[n], p.y[n], p.z[n]};
    int ijk[3];
    double w[8];
    centre_weights(g, x, ijk, w);
    const int is = ijk[0], js = ijk[1], ks = ijk[2];
    if (ks < -1 || ks > g.nz) {   // the node fields hold planes -1..nz+1
        atomicOr(err, ERR_OUT_OF_GRID);
        return;
    }
    const float *volg = fld + F_VOLG * field_stride(g);
    const double vci = 1.0 / g.vcell;
    auto PHI = [&](int k, int j, int i) { return (double)volg[node_off(g, k, j, i)] * vci - 1.0; };
    const double xf = w[1] + w[3] + w[5] + w[7];
    const double yf = w[2] + w[3] + w[6] + w[7];
    const double zf = w[4] + w[5] + w[6] + w[7];
    const double xfc = 1.0 - xf, yfc = 1.0 - yf, zfc = 1.0 - zf;
    double lim_x = g.dx[0] * xf * xfc;
    double xs = -prefactor * lim_x *
                (zfc * (yfc * (PHI(ks, js, is + 1) - PHI(ks, js, is)) + yf * (PHI(ks, js + 1, is + 1) - PHI(ks, js + 1, is))) +
                 zf * (yfc * (PHI(ks + 1, js, is + 1) - PHI(ks + 1, js, is)) +
                       yf * (PHI(ks + 1, js + 1, is + 1) - PHI(ks + 1, js + 1, is))));
    lim_x = lim_x * max_compression;
    xs = fmax(-lim_x, fmin(xs, lim_x));
    double lim_y = g.dx[1] * yf * yfc;
    double ys = -prefactor * lim_y *
                (zfc * (xfc * (PHI(ks, js + 1, is) - PHI(ks, js, is)) + xf * (PHI(ks, js + 1, is + 1) - PHI(ks, js, is + 1))) +
                 zf * (xfc * (PHI(ks + 1, js + 1, is) - PHI(ks + 1, js, is)) +
                       xf * (PHI(ks + 1, js + 1, is + 1) - PHI(ks + 1, js, is + 1))));
    lim_y = lim_y * max_compression;
    ys = fmax(-lim_y, fmin(ys, lim_y));
    double lim_z = g.dx[2] * zf * zfc;
    double zs = -prefactor * lim_z *
                (xfc * (yfc * (PHI(ks + 1, js, is) - PHI(ks, js, is)) + yf * (PHI(ks + 1, js + 1, is) - PHI(ks, js + 1, is))) +
                 xf * (yfc * (PHI(ks + 1, js, is + 1) - PHI(ks, js, is + 1)) +
                       yf * (PHI(ks + 1, js + 1, is + 1) - PHI(ks, js + 1, is + 1))));
    lim_z = lim_z * max_compression;
    zs = fmax(-lim_z, fmin(zs, lim_z));
    x[0] = x[0] + xs;
    x[1] = x[1] + ys;
    x[2] = x[2] + zs;
    x[0] = x[0] - g.extent[0] * trunc((x[0] - g.center[0]) * g.hli[0]);
    x[1] = x[1] - g.extent[1] * trunc((x[1] - g.center[1]) * g.hli[1]);
    p.x[n] = x[0];
    p.y[n] = x[1];
    p.z[n] = x[2];
}

// volume-weighted vorticity: per-block FP64 partials in a fixed order, then
// one block adds the partials in block order
constexpr int VW_T = 256;
__global__ __launch_bounds__(VW_T) void k_vw_partial(ParcelsDev p, double *part)
{
    __shared__ double s[4][VW_T];
    const int64_t n = gtid();
    double v = 0.0, w0 = 0.0, w1 = 0.0, w2 = 0.0;
    if (n < p.n) {
        v = p.vol[n];
        w0 = (double)p.vor[0][n] * v;
        w1 = (double)p.vor[1][n] * v;
        w2 = (double)p.vor[2][n] * v;
    }
    s[0][threadIdx.x] = v;
    s[1][threadIdx.x] = w0;
    s[2][threadIdx.x] = w1;
    s[3][threadIdx.x] = w2;
    __syncthreads();
    for (int o = VW_T / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
#pragma unroll
            for (int c = 0; c < 4; ++c) s[c][threadIdx.x] += s[c][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 4) part[(int64_t)blockIdx.x * 4 + threadIdx.x] = s[threadIdx.x][0];
}
__global__ __launch_bounds__(VW_T) void k_vw_total(const double *part, int nb, double *out)
{
    __shared__ double s[4][VW_T];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += VW_T)
#pragma unroll
        for (int c = 0; c < 4; ++c) a[c] += part[(int64_t)b * 4 + c];
#pragma unroll
    for (int c = 0; c < 4; ++c) s[c][threadIdx.x] = a[c];
    __syncthreads();
    for (int o = VW_T / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
#pragma unroll
            for (int c = 0; c < 4; ++c) s[c][threadIdx.x] += s[c][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 4) out[threadIdx.x] = s[threadIdx.x][0];
}
__global__ void k_vortcor(ParcelsDev p, double d0, double d1, double d2)
{
    const int64_t n = gtid();
    if (n >= p.n) return;
    p.vor[0][n] = (float)((double)p.vor[0][n] - d0);
    p.vor[1][n] = (float)((double)p.vor[1][n] - d1);
    p.vor[2][n] = (float)((double)p.vor[2][n] - d2);
}

void launch_lapinv_columns(const Grid &g, const float2 *in, double *fr, double *fi, double *et, float2 *ox, float2 *oy,
                           float2 *oz, hipStream_t s)
{
    hipLaunchKernelGGL(k_lapinv_columns, dim3(nblk((int64_t)g.ny * g.nxh)), dim3(256), 0, s, g, in, fr, fi, et, ox, oy,
                       oz);
}
void launch_grad_halo(const Grid &g, float *fld, hipStream_t s)
{
    hipLaunchKernelGGL(k_grad_halo, dim3(nblk((int64_t)g.nx * g.ny)), dim3(256), 0, s, g, fld);
}
void launch_laplace_push(const Grid &g, const ParcelsDev &p, const float *fld, int *err, hipStream_t s)
{
    hipLaunchKernelGGL(k_laplace_push, dim3(nblk(p.n)), dim3(256), 0, s, g, p, fld, err);
}
void launch_gradient_push(const Grid &g, const ParcelsDev &p, const float *fld, double prefactor, double max_compression,
                          int *err, hipStream_t s)
{
    hipLaunchKernelGGL(k_gradient_push, dim3(nblk(p.n)), dim3(256), 0, s, g, p, fld, prefactor, max_compression, err);
}
int64_t vw_partials(int64_t n) { return 4 * (int64_t)nblk(n) + 4; }
void launch_volume_weighted_vorticity(const ParcelsDev &p, double *part, double *out, hipStream_t s)
{
    const unsigned nb = nblk(p.n);
    hipLaunchKernelGGL(k_vw_partial, dim3(nb), dim3(VW_T), 0, s, p, part);
    hipLaunchKernelGGL(k_vw_total, dim3(1), dim3(VW_T), 0, s, part, (int)nb, out);
}
void launch_vortcor(const ParcelsDev &p, const double d[3], hipStream_t s)
{
    hipLaunchKernelGGL(k_vortcor, dim3(nblk(p.n)), dim3(256), 0, s, p, d[0], d[1], d[2]);
}

}  // namespace epic
