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
// The kernels are templated on the complex type of the spectra and of the
// arithmetic: float2 (the FP32 transforms) or double2 (the vorticity
// tendency's transforms, FP64 end to end, from the FP32 node fields).
//
// The global -> LDS staging loops below request LOAD_BATCH elements per
// thread before the first LDS store, with clamped (always valid) addresses
// and no branches between the loads, so a thread waits one memory round
// trip per batch rather than one per element.
constexpr int LOAD_BATCH = 8;

template <typename T2> __device__ inline const T2 *twiddles(const FftXArgs &a);
template <> __device__ inline const float2 *twiddles<float2>(const FftXArgs &a) { return a.tw; }
template <> __device__ inline const double2 *twiddles<double2>(const FftXArgs &a) { return a.tw64; }
template <typename T2> __device__ inline const T2 *twiddles(const FftYArgs &a);
template <> __device__ inline const float2 *twiddles<float2>(const FftYArgs &a) { return a.tw; }
template <> __device__ inline const double2 *twiddles<double2>(const FftYArgs &a) { return a.tw64; }

// forward x input value at node offset o (mode fixed per launch item)
enum XfKind { XF_FIELD, XF_FLUX, XF_FLUX_B, XF_PHI };
template <int KIND, typename R> __device__ inline R phys_value(const FftXArgs &a, int64_t o, int64_t FS)
{
    if constexpr (KIND == XF_FIELD) return (R)gl(a.fld)[a.slot * FS + o];
    // parcel-correction source: volg / vcell - 1
    if constexpr (KIND == XF_PHI) return (R)(gl(a.fld)[F_VOLG * FS + o] * (float)(1.0 / a.g.vcell) - 1.0f);
    // vorticity-tendency flux: (omega_a + f_a) * vel_b (+ sgn * tbuoy)
    R v = ((R)gl(a.fld)[(F_VORTG + a.comp_a) * FS + o] + (R)a.g.f_cor[a.comp_a]) *
          (R)gl(a.fld)[(F_VELOG + a.comp_b) * FS + o];
    if constexpr (KIND == XF_FLUX_B) v += (R)a.sgn * (R)gl(a.fld)[F_TBUOYG * FS + o];
    return v;
}

// ZB rows of nx reals (planes zp0.., row j) packed as m = nx/2 complex
// values per row into buf[r m + q]; planes above nz are zero, except with
// HALO, where z index nz + 1 carries node plane -1 and nz + 2 node plane
// nz + 1 as the linear extrapolations 2 f(0) - f(1) and 2 f(nz) - f(nz-1)
// of the input (so that the transform of a halo plane is the extrapolation
// of the transformed planes, without the error of the extrapolation)
template <int KIND, bool HALO, typename T2>
__device__ inline void stage_rows(const FftXArgs &a, T2 *buf, int j, int zp0, int m)
{
    using R = typename Cx<T2>::R;
    const Grid &g = a.g;
    const int nz = g.nz;
    const int64_t FS = field_stride(g), PS = plane_stride(g);
    const int total = ZB * m;
    for (int t0 = threadIdx.x; t0 < total; t0 += LOAD_BATCH * blockDim.x) {
        T2 z[LOAD_BATCH];
#pragma unroll
        for (int u = 0; u < LOAD_BATCH; ++u) {
            const int t = min(t0 + u * (int)blockDim.x, total - 1);
            const int r = fdiv(t, a.plan.n1), q = t - r * m;   // the x plan's length is m
            const int zp = zp0 + r;
            const int64_t row = (int64_t)j * g.nx + 2 * q;
            if constexpr (HALO) {
                // plane p0 with weight c0 plus plane p1 with weight c1
                int p0 = min(zp, nz), p1 = p0;
                R c0 = 1, c1 = 0;
                if (zp == nz + 1) { p0 = 0; p1 = 1; c0 = 2; c1 = -1; }
                else if (zp == nz + 2) { p0 = nz; p1 = nz - 1; c0 = 2; c1 = -1; }
                const int64_t o0 = (int64_t)(p0 + 1) * PS + row, o1 = (int64_t)(p1 + 1) * PS + row;
                z[u] = T2{c0 * phys_value<KIND, R>(a, o0, FS) + c1 * phys_value<KIND, R>(a, o1, FS),
                          c0 * phys_value<KIND, R>(a, o0 + 1, FS) + c1 * phys_value<KIND, R>(a, o1 + 1, FS)};
            } else {
                const int64_t o = (int64_t)(min(zp, nz) + 1) * PS + row;
                z[u] = T2{phys_value<KIND, R>(a, o, FS), phys_value<KIND, R>(a, o + 1, FS)};
            }
        }
#pragma unroll
        for (int u = 0; u < LOAD_BATCH; ++u) {
            const int t = t0 + u * blockDim.x;
            const int zmax = HALO ? nz + 2 : nz;
            if (t < total) buf[t] = (zp0 + fdiv(t, a.plan.n1) <= zmax) ? z[u] : T2{0, 0};
        }
    }
}

// ZB z-contiguous spectra columns: buf[r * len + c] = in[(c * stride + col) * nzp + zp0 + r]
// for c < len (MODE2: i kx in + i ky in2 of the y pass, kx factor kxw)
template <bool MODE2, typename T2>
__device__ inline void stage_columns(const T2 *in, const T2 *in2, T2 *buf, int len, int64_t stride, int col,
                                     const Grid &g, int zp0, typename Cx<T2>::R kxw)
{
    using R = typename Cx<T2>::R;
    const int total = ZB * len;
    for (int t0 = threadIdx.x; t0 < total; t0 += LOAD_BATCH * blockDim.x) {
        T2 v[LOAD_BATCH];
#pragma unroll
        for (int u = 0; u < LOAD_BATCH; ++u) {
            const int t = min(t0 + u * (int)blockDim.x, total - 1);
            const int c = t / ZB, r = t - c * ZB;
            const int64_t idx = ((int64_t)c * stride + col) * g.nzp + zp0 + r;
            if constexpr (MODE2) {
                const int kys = (c <= len / 2) ? c : c - len;
                const R kyw = (2 * c == len) ? (R)0 : (R)(2.0 * 3.14159265358979323846 * kys / g.extent[1]);
                const T2 f1 = gl(in)[idx], f2 = gl(in2)[idx];
                v[u] = T2{-(kxw * f1.y + kyw * f2.y), kxw * f1.x + kyw * f2.x};   // i kx f1 + i ky f2
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

template <typename T2> __global__ __launch_bounds__(256) void k_fft_xf(const FftXArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    using R = typename Cx<T2>::R;
    const ItemBlock ib = item_block(bmap);
    const FftXArgs &a = tab[ib.item];
    const Grid &g = a.g;
    const int j = ib.blk % g.ny;
    const int zp0 = (ib.blk / g.ny) * ZB;
    const int m = g.nx / 2;
    extern __shared__ char lds_raw[];
    T2 *lds = (T2 *)lds_raw;
    T2 *bufa = lds, *bufb = lds + ZB * (m + 1);
    T2 *tw = bufb + ZB * (m + 1);   // twiddles staged in LDS
    const T2 *twg = twiddles<T2>(a);
    for (int t = threadIdx.x; t < g.nx; t += blockDim.x) tw[t] = gl(twg)[t];
    if constexpr (sizeof(T2) == sizeof(double2)) {   // the tendency fluxes, with the halo planes
        if (a.sgn != 0) stage_rows<XF_FLUX_B, true>(a, bufa, j, zp0, m);
        else stage_rows<XF_FLUX, true>(a, bufa, j, zp0, m);
    } else {
        if (a.mode == 0) stage_rows<XF_FIELD, false>(a, bufa, j, zp0, m);
        else if (a.mode == 3) stage_rows<XF_PHI, false>(a, bufa, j, zp0, m);
        else if (a.sgn != 0) stage_rows<XF_FLUX_B, false>(a, bufa, j, zp0, m);
        else stage_rows<XF_FLUX, false>(a, bufa, j, zp0, m);
    }
    __syncthreads();
    T2 *Z = lds_fft(bufa, bufb, ZB, a.plan, tw, 2, false);
    T2 *other = (Z == bufa) ? bufb : bufa;
    // untangle the packed real transform: X_k, k = 0..m
    for (int t = threadIdx.x; t < ZB * (m + 1); t += blockDim.x) {
        const int kx = t / ZB, r = t - kx * ZB;
        const T2 zk = Z[r * m + (kx == m ? 0 : kx)];              // kx mod m, kx in [0, m]
        const T2 zc = cconj(Z[r * m + (kx == 0 ? 0 : m - kx)]);   // (m - kx) mod m
        const T2 w = tw[kx];   // e^{-2 pi i kx / nx}
        const T2 s = cadd(zk, zc), d = csub(zk, zc);
        const T2 wd = cmul(w, d);
        // X = (s - i w d) / 2
        other[r * (m + 1) + kx] = T2{(R)0.5 * (s.x + wd.y), (R)0.5 * (s.y - wd.x)};
    }
    __syncthreads();
    const int nxh = m + 1;
    T2 *out = (T2 *)a.out;
    for (int t = threadIdx.x; t < ZB * nxh; t += blockDim.x) {
        const int kx = t / ZB, r = t - kx * ZB;
        gl(out)[((int64_t)j * nxh + kx) * g.nzp + zp0 + r] = other[r * (m + 1) + kx];
    }
}

// ---------------------------------------------------------------------------
// y pass (c2c), forward or inverse
// ---------------------------------------------------------------------------
// in place is allowed (out == in): a block reads and writes only its own
// (kx, z block) slab
template <typename T2> __global__ __launch_bounds__(256) void k_fft_y(const FftYArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    using R = typename Cx<T2>::R;
    const ItemBlock ib = item_block(bmap);
    const FftYArgs &a = tab[ib.item];
    const Grid &g = a.g;
    const int kx = ib.blk % g.nxh;
    const int zp0 = (ib.blk / g.nxh) * ZB;
    const int ny = g.ny, nxh = g.nxh;
    extern __shared__ char lds_raw[];
    T2 *lds = (T2 *)lds_raw;
    T2 *bufa = lds, *bufb = lds + ZB * ny;
    T2 *tw = bufb + ZB * ny;   // twiddles staged in LDS
    const T2 *twg = twiddles<T2>(a);
    for (int t = threadIdx.x; t < ny; t += blockDim.x) tw[t] = gl(twg)[t];
    // inverse combine factors: i kx (0 at the x Nyquist) and i ky (0 at the y Nyquist)
    const R kxw = (2 * kx == g.nx) ? (R)0 : (R)(2.0 * 3.14159265358979323846 * kx / g.extent[0]);
    const T2 *in = (const T2 *)a.in, *in2 = (const T2 *)a.in2;
    if (a.mode == 2) stage_columns<true>(in, in2, bufa, ny, nxh, kx, g, zp0, kxw);
    else stage_columns<false>(in, (const T2 *)nullptr, bufa, ny, nxh, kx, g, zp0, kxw);
    __syncthreads();
    T2 *Rb = lds_fft(bufa, bufb, ZB, a.plan, tw, 1, a.mode != 0);
    T2 *out = (T2 *)a.out;
    for (int t = threadIdx.x; t < ZB * ny; t += blockDim.x) {
        const int jy = t / ZB, r = t - jy * ZB;
        gl(out)[((int64_t)jy * nxh + kx) * g.nzp + zp0 + r] = Rb[r * ny + jy];
    }
}

// ---------------------------------------------------------------------------
// inverse x pass: half spectra -> real rows, with the output stage
// ---------------------------------------------------------------------------
template <typename T2> __global__ __launch_bounds__(256) void k_fft_xi(const FftXArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    using R = typename Cx<T2>::R;
    constexpr bool HALO = sizeof(T2) == sizeof(double2);   // the tendency transforms carry the halo planes
    const ItemBlock ib = item_block(bmap);
    const FftXArgs &a = tab[ib.item];
    const Grid &g = a.g;
    const int j = ib.blk % g.ny;
    const int zp0 = (ib.blk / g.ny) * ZB;
    const int m = g.nx / 2, nxh = m + 1;
    extern __shared__ char lds_raw[];
    T2 *lds = (T2 *)lds_raw;
    T2 *bufa = lds, *bufb = lds + ZB * (m + 1);
    T2 *tw = bufb + ZB * (m + 1);   // twiddles staged in LDS
    const T2 *twg = twiddles<T2>(a);
    for (int t = threadIdx.x; t < g.nx; t += blockDim.x) tw[t] = gl(twg)[t];
    stage_columns<false>((const T2 *)a.in, (const T2 *)nullptr, bufb, nxh, 1, j * nxh, g, zp0, (R)0);
    __syncthreads();
    for (int t = threadIdx.x; t < ZB * m; t += blockDim.x) {
        const int r = fdiv(t, a.plan.n1), k = t - r * m;
        const T2 xk = bufb[r * nxh + k];
        const T2 xc = cconj(bufb[r * nxh + (m - k)]);
        const T2 w = cconj(tw[k]);   // e^{+2 pi i k / nx}
        const T2 s = cadd(xk, xc), d = csub(xk, xc);
        const T2 wd = cmul(w, d);
        bufa[r * m + k] = T2{s.x - wd.y, s.y + wd.x};   // s + i w d
    }
    __syncthreads();
    T2 *Z = lds_fft(bufa, bufb, ZB, a.plan, tw, 2, true);
    const R scale = (R)1 / ((R)g.nx * (R)g.ny);
    const int64_t FS = field_stride(g), PS = plane_stride(g);
    const int nz = g.nz;
    for (int t = threadIdx.x; t < ZB * g.nx; t += blockDim.x) {
        const int r = fdiv(t, a.plan.n2), i = t - r * g.nx;   // 2 m = nx
        const int zi = zp0 + r;
        // node plane written: 0..nz, and with HALO the z indices nz + 1 and
        // nz + 2 are node planes -1 and nz + 1
        int k = zi;
        if (HALO && zi == nz + 1) k = -1;
        else if (HALO && zi == nz + 2) k = nz + 1;
        else if (zi > nz) continue;
        const T2 z = Z[r * m + (i >> 1)];
        R v = ((i & 1) ? z.y : z.x) * scale;
        const int64_t o = (int64_t)(k + 1) * PS + (int64_t)j * g.nx + i;
        if (a.mode == 2) {
            // + d/dz of f3 = (zeta + f_z) vel_c, second order (central_diffz,
            // inversion_utils.f90:488-514); the halo planes take its linear
            // extrapolation, as the reference extrapolates the tendency
            const R fz = (R)g.f_cor[2];
            auto f3 = [&](int kk) {
                const int64_t oo = (int64_t)(kk + 1) * PS + (int64_t)j * g.nx + i;
                return ((R)gl(a.fld)[(F_VORTG + 2) * FS + oo] + fz) * (R)gl(a.fld)[(F_VELOG + a.comp_b) * FS + oo];
            };
            const R dzi = (R)g.dxi[2];
            auto dz = [&](int kk) -> R {
                if (kk == 0) return dzi * (f3(1) - f3(0));
                if (kk == nz) return dzi * (f3(nz) - f3(nz - 1));
                return (f3(kk + 1) - f3(kk - 1)) * ((R)0.5 * dzi);
            };
            R dd;
            if (k == -1) dd = (R)2 * dz(0) - dz(1);
            else if (k == nz + 1) dd = (R)2 * dz(nz) - dz(nz - 1);
            else dd = dz(k);
            v = v + dd;
        }
        gl(a.fld_out)[a.slot * FS + o] = (float)v;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <typename T2> static size_t x_lds(const Grid &g) { return (2 * ZB * (size_t)(g.nx / 2 + 1) + g.nx) * sizeof(T2); }
template <typename T2> static size_t y_lds(const Grid &g) { return (2 * ZB * (size_t)g.ny + g.ny) * sizeof(T2); }

size_t transform_lds_bytes(const Grid &g) { return std::max({x_lds<double2>(g), y_lds<double2>(g), col_lds_bytes(g)}); }

// every item of one launch has the same precision
template <class A> static bool fp64_items(const std::vector<A> &items) { return !items.empty() && items[0].fp64; }

void add_fft_xf(Plan &P, int phase, const std::vector<FftXArgs> &items)
{
    const bool d = fp64_items(items);
    std::vector<int> nb;
    size_t lds = 0;
    for (const FftXArgs &a : items) {
        nb.push_back(a.g.ny * (a.g.nzp / ZB));
        lds = std::max(lds, d ? x_lds<double2>(a.g) : x_lds<float2>(a.g));
    }
    P.add(d ? (const void *)k_fft_xf<double2> : (const void *)k_fft_xf<float2>, 256, lds, phase, items, nb);
}

void add_fft_y(Plan &P, int phase, const std::vector<FftYArgs> &items)
{
    const bool d = fp64_items(items);
    std::vector<int> nb;
    size_t lds = 0;
    for (const FftYArgs &a : items) {
        nb.push_back(a.g.nxh * (a.g.nzp / ZB));
        lds = std::max(lds, d ? y_lds<double2>(a.g) : y_lds<float2>(a.g));
    }
    P.add(d ? (const void *)k_fft_y<double2> : (const void *)k_fft_y<float2>, 256, lds, phase, items, nb);
}

void add_fft_xi(Plan &P, int phase, const std::vector<FftXArgs> &items)
{
    const bool d = fp64_items(items);
    std::vector<int> nb;
    size_t lds = 0;
    for (const FftXArgs &a : items) {
        nb.push_back(a.g.ny * (a.g.nzp / ZB));
        lds = std::max(lds, d ? x_lds<double2>(a.g) : x_lds<float2>(a.g));
    }
    P.add(d ? (const void *)k_fft_xi<double2> : (const void *)k_fft_xi<float2>, 256, lds, phase, items, nb);
}

}  // namespace epic
