// Horizontal 2D real FFTs of the node planes z = 0..nz (FP32), the transform
// pair of fftxyp2s / fftxys2p (sta3dfft.f90:134-257) in complex half-spectrum
// form: S(kx, ky) = sum_{i,j} f(i, j) e^{-2 pi i (kx i/nx + ky j/ny)},
// kx in [0, nx/2].  The inverse carries the 1/(nx ny) normalisation.
//
// Two passes per direction; the intermediate T[j][kx][z] and the spectra
// S[ky][kx][z] keep ZB consecutive z planes contiguous, so every pass moves
// 128-byte segments and the vertical (z) column kernels read whole columns.
//
// Loaders fuse the physical products of the vorticity tendency into the
// forward pass; the inverse y pass can form i kx F1 + i ky F2 on the fly and
// the inverse x pass can add the second-order z derivative of a product.
#include "common.h"
#include "fft.h"
#include "launch.h"

#include <algorithm>

namespace epic {

// ---------------------------------------------------------------------------
// forward x pass: real rows -> half spectra
// ---------------------------------------------------------------------------
// The global -> LDS staging loops below request LOAD_BATCH elements per
// thread before the first LDS store, with clamped (always valid) addresses
// and no branches between the loads, so a thread waits one memory round
// trip per batch rather than one per element.
constexpr int LOAD_BATCH = 8;

// forward x input value at node offset o (mode fixed per launch item)
enum XfKind { XF_FIELD, XF_FLUX, XF_FLUX_B, XF_PHI };
template <int KIND> __device__ inline float phys_value(const FftXArgs &a, int64_t o, int64_t FS)
{
    if constexpr (KIND == XF_FIELD) return gl(a.fld)[a.slot * FS + o];
    // parcel-correction source: volg / vcell - 1
    if constexpr (KIND == XF_PHI) return gl(a.fld)[F_VOLG * FS + o] * (float)(1.0 / a.g.vcell) - 1.0f;
    // vorticity-tendency flux: (omega_a + f_a) * vel_b (+ sgn * tbuoy)
    float v = (gl(a.fld)[(F_VORTG + a.comp_a) * FS + o] + (float)a.g.f_cor[a.comp_a]) * gl(a.fld)[(F_VELOG + a.comp_b) * FS + o];
    if constexpr (KIND == XF_FLUX_B) v += (float)a.sgn * gl(a.fld)[F_TBUOYG * FS + o];
    return v;
}

// ZB rows of nx reals (planes zp0.., row j) packed as m = nx/2 complex
// values per row into buf[r m + q]; planes above nz are zero
template <int KIND> __device__ inline void stage_rows(const FftXArgs &a, float2 *buf, int j, int zp0, int m)
{
    const Grid &g = a.g;
    const int64_t FS = field_stride(g), PS = plane_stride(g);
    const int total = ZB * m;
    for (int t0 = threadIdx.x; t0 < total; t0 += LOAD_BATCH * blockDim.x) {
        float2 z[LOAD_BATCH];
#pragma unroll
        for (int u = 0; u < LOAD_BATCH; ++u) {
            const int t = min(t0 + u * (int)blockDim.x, total - 1);
            const int r = t / m, q = t - r * m;
            const int64_t o = (int64_t)(min(zp0 + r, g.nz) + 1) * PS + (int64_t)j * g.nx + 2 * q;
            z[u] = float2{phys_value<KIND>(a, o, FS), phys_value<KIND>(a, o + 1, FS)};
        }
#pragma unroll
        for (int u = 0; u < LOAD_BATCH; ++u) {
            const int t = t0 + u * blockDim.x;
            if (t < total) buf[t] = (zp0 + t / m <= g.nz) ? z[u] : float2{0.0f, 0.0f};
        }
    }
}

// ZB z-contiguous spectra columns: buf[r * len + c] = in[(c * stride + col) * nzp + zp0 + r]
// for c < len (MODE2: i kx in + i ky in2 of the y pass, kx factor kxw)
template <bool MODE2>
__device__ inline void stage_columns(const float2 *in, const float2 *in2, float2 *buf, int len, int64_t stride, int col,
                                     const Grid &g, int zp0, float kxw)
{
    const int total = ZB * len;
    for (int t0 = threadIdx.x; t0 < total; t0 += LOAD_BATCH * blockDim.x) {
        float2 v[LOAD_BATCH];
#pragma unroll
        for (int u = 0; u < LOAD_BATCH; ++u) {
            const int t = min(t0 + u * (int)blockDim.x, total - 1);
            const int c = t / ZB, r = t - c * ZB;
            const int64_t idx = ((int64_t)c * stride + col) * g.nzp + zp0 + r;
            if constexpr (MODE2) {
                const int kys = (c <= len / 2) ? c : c - len;
                const float kyw = (2 * c == len) ? 0.0f : (float)(2.0 * 3.14159265358979323846 * kys / g.extent[1]);
                const float2 f1 = gl(in)[idx], f2 = gl(in2)[idx];
                v[u] = float2{-(kxw * f1.y + kyw * f2.y), kxw * f1.x + kyw * f2.x};   // i kx f1 + i ky f2
            } else {
                v[u] = gl(in)[idx];
            }
        }
#pragma unroll
        for (int u = 0; u < LOAD_BATCH; ++u) {
            const int t = t0 + u * blockDim.x, c = t / ZB, r = t - c * ZB;
            if (t < total) buf[r * len + c] = v[u];
        }
    }
}

__global__ __launch_bounds__(256) void k_fft_xf(const FftXArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const FftXArgs &a = tab[ib.item];
    const Grid &g = a.g;
    const int j = ib.blk % g.ny;
    const int zp0 = (ib.blk / g.ny) * ZB;
    const int m = g.nx / 2;
    extern __shared__ float2 lds[];
    float2 *bufa = lds, *bufb = lds + ZB * (m + 1);
    float2 *tw = bufb + ZB * (m + 1);   // twiddles staged in LDS
    for (int t = threadIdx.x; t < g.nx; t += blockDim.x) tw[t] = gl(a.tw)[t];
    if (a.mode == 0) stage_rows<XF_FIELD>(a, bufa, j, zp0, m);
    else if (a.mode == 3) stage_rows<XF_PHI>(a, bufa, j, zp0, m);
    else if (a.sgn != 0) stage_rows<XF_FLUX_B>(a, bufa, j, zp0, m);
    else stage_rows<XF_FLUX>(a, bufa, j, zp0, m);
    __syncthreads();
    float2 *Z = lds_fft(bufa, bufb, ZB, a.plan, tw, 2, false);
    float2 *other = (Z == bufa) ? bufb : bufa;
    // untangle the packed real transform: X_k, k = 0..m
    for (int t = threadIdx.x; t < ZB * (m + 1); t += blockDim.x) {
        const int kx = t / ZB, r = t - kx * ZB;
        const float2 zk = Z[r * m + (kx % m)];
        const float2 zc = cconj(Z[r * m + ((m - kx) % m)]);
        const float2 w = tw[kx];   // e^{-2 pi i kx / nx}
        const float2 s = cadd(zk, zc), d = csub(zk, zc);
        const float2 wd = cmul(w, d);
        // X = (s - i w d) / 2
        other[r * (m + 1) + kx] = float2{0.5f * (s.x + wd.y), 0.5f * (s.y - wd.x)};
    }
    __syncthreads();
    const int nxh = m + 1;
    for (int t = threadIdx.x; t < ZB * nxh; t += blockDim.x) {
        const int kx = t / ZB, r = t - kx * ZB;
        gl(a.out)[((int64_t)j * nxh + kx) * g.nzp + zp0 + r] = other[r * (m + 1) + kx];
    }
}

// ---------------------------------------------------------------------------
// y pass (c2c), forward or inverse
// ---------------------------------------------------------------------------
// in place is allowed (out == in): a block reads and writes only its own
// (kx, z block) slab
__global__ __launch_bounds__(256) void k_fft_y(const FftYArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const FftYArgs &a = tab[ib.item];
    const Grid &g = a.g;
    const int kx = ib.blk % g.nxh;
    const int zp0 = (ib.blk / g.nxh) * ZB;
    const int ny = g.ny, nxh = g.nxh;
    extern __shared__ float2 lds[];
    float2 *bufa = lds, *bufb = lds + ZB * ny;
    float2 *tw = bufb + ZB * ny;   // twiddles staged in LDS
    for (int t = threadIdx.x; t < ny; t += blockDim.x) tw[t] = gl(a.tw)[t];
    // inverse combine factors: i kx (0 at the x Nyquist) and i ky (0 at the y Nyquist)
    const float kxw = (2 * kx == g.nx) ? 0.0f : (float)(2.0 * 3.14159265358979323846 * kx / g.extent[0]);
    if (a.mode == 2) stage_columns<true>(a.in, a.in2, bufa, ny, nxh, kx, g, zp0, kxw);
    else stage_columns<false>(a.in, nullptr, bufa, ny, nxh, kx, g, zp0, kxw);
    __syncthreads();
    float2 *R = lds_fft(bufa, bufb, ZB, a.plan, tw, 1, a.mode != 0);
    for (int t = threadIdx.x; t < ZB * ny; t += blockDim.x) {
        const int jy = t / ZB, r = t - jy * ZB;
        gl(a.out)[((int64_t)jy * nxh + kx) * g.nzp + zp0 + r] = R[r * ny + jy];
    }
}

// ---------------------------------------------------------------------------
// inverse x pass: half spectra -> real rows, with the output stage
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fft_xi(const FftXArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const FftXArgs &a = tab[ib.item];
    const Grid &g = a.g;
    const int j = ib.blk % g.ny;
    const int zp0 = (ib.blk / g.ny) * ZB;
    const int m = g.nx / 2, nxh = m + 1;
    extern __shared__ float2 lds[];
    float2 *bufa = lds, *bufb = lds + ZB * (m + 1);
    float2 *tw = bufb + ZB * (m + 1);   // twiddles staged in LDS
    for (int t = threadIdx.x; t < g.nx; t += blockDim.x) tw[t] = gl(a.tw)[t];
    stage_columns<false>(a.in, nullptr, bufb, nxh, 1, j * nxh, g, zp0, 0.0f);
    __syncthreads();
    for (int t = threadIdx.x; t < ZB * m; t += blockDim.x) {
        const int r = t / m, k = t - r * m;
        const float2 xk = bufb[r * nxh + k];
        const float2 xc = cconj(bufb[r * nxh + (m - k)]);
        const float2 w = cconj(tw[k]);   // e^{+2 pi i k / nx}
        const float2 s = cadd(xk, xc), d = csub(xk, xc);
        const float2 wd = cmul(w, d);
        bufa[r * m + k] = float2{s.x - wd.y, s.y + wd.x};   // s + i w d
    }
    __syncthreads();
    float2 *Z = lds_fft(bufa, bufb, ZB, a.plan, tw, 2, true);
    const float scale = 1.0f / ((float)g.nx * (float)g.ny);
    const int64_t FS = field_stride(g), PS = plane_stride(g);
    for (int t = threadIdx.x; t < ZB * g.nx; t += blockDim.x) {
        const int r = t / g.nx, i = t - r * g.nx;
        const int k = zp0 + r;
        if (k > g.nz) continue;
        const float2 z = Z[r * m + (i >> 1)];
        float v = ((i & 1) ? z.y : z.x) * scale;
        const int64_t o = (int64_t)(k + 1) * PS + (int64_t)j * g.nx + i;
        if (a.mode == 2) {
            // + d/dz of f3 = (zeta + f_z) vel_c, second order (central_diffz,
            // inversion_utils.f90:488-514)
            const float fz = (float)g.f_cor[2];
            auto f3 = [&](int kk) {
                const int64_t oo = (int64_t)(kk + 1) * PS + (int64_t)j * g.nx + i;
                return (gl(a.fld)[(F_VORTG + 2) * FS + oo] + fz) * gl(a.fld)[(F_VELOG + a.comp_b) * FS + oo];
            };
            const float dzi = (float)g.dxi[2];
            float dd;
            if (k == 0) dd = dzi * (f3(1) - f3(0));
            else if (k == g.nz) dd = dzi * (f3(g.nz) - f3(g.nz - 1));
            else dd = (f3(k + 1) - f3(k - 1)) * (0.5f * dzi);
            v = v + dd;
        }
        gl(a.fld_out)[a.slot * FS + o] = v;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static size_t x_lds(const Grid &g) { return (2 * ZB * (size_t)(g.nx / 2 + 1) + g.nx) * sizeof(float2); }
static size_t y_lds(const Grid &g) { return (2 * ZB * (size_t)g.ny + g.ny) * sizeof(float2); }

void add_fft_xf(Plan &P, int phase, const std::vector<FftXArgs> &items)
{
    std::vector<int> nb;
    size_t lds = 0;
    for (const FftXArgs &a : items) {
        nb.push_back(a.g.ny * (a.g.nzp / ZB));
        lds = std::max(lds, x_lds(a.g));
    }
    P.add((const void *)k_fft_xf, 256, lds, phase, items, nb);
}

void add_fft_y(Plan &P, int phase, const std::vector<FftYArgs> &items)
{
    std::vector<int> nb;
    size_t lds = 0;
    for (const FftYArgs &a : items) {
        nb.push_back(a.g.nxh * (a.g.nzp / ZB));
        lds = std::max(lds, y_lds(a.g));
    }
    P.add((const void *)k_fft_y, 256, lds, phase, items, nb);
}

void add_fft_xi(Plan &P, int phase, const std::vector<FftXArgs> &items)
{
    std::vector<int> nb;
    size_t lds = 0;
    for (const FftXArgs &a : items) {
        nb.push_back(a.g.ny * (a.g.nzp / ZB));
        lds = std::max(lds, x_lds(a.g));
    }
    P.add((const void *)k_fft_xi, 256, lds, phase, items, nb);
}

}  // namespace epic
