// Batched complex FFT over LDS (Stockham autosort, radix 4/2/3/5) shared by
// the horizontal FP32 transforms and the FP64 vertical transforms.
#pragma once

#include "common.h"

namespace epic {

template <typename T2> struct Cx;
template <> struct Cx<float2> { using R = float; };
template <> struct Cx<double2> { using R = double; };

template <typename T2> __device__ inline T2 cadd(T2 a, T2 b) { return T2{a.x + b.x, a.y + b.y}; }
template <typename T2> __device__ inline T2 csub(T2 a, T2 b) { return T2{a.x - b.x, a.y - b.y}; }
template <typename T2> __device__ inline T2 cmul(T2 a, T2 b) { return T2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <typename T2> __device__ inline T2 cconj(T2 a) { return T2{a.x, -a.y}; }
// multiply by -i (forward) or +i (inverse)
template <typename T2> __device__ inline T2 cmuli(T2 a, bool inv) { return inv ? T2{-a.y, a.x} : T2{a.y, -a.x}; }
template <typename T2, typename R> __device__ inline T2 cscale(T2 a, R s) { return T2{a.x * s, a.y * s}; }

// In-place R-point DFT, R in {2, 3, 4, 5}; forward uses e^{-2 pi i/R}.
template <int R, typename T2> __device__ inline void butterfly(T2 *v, bool inv)
{
    using Rl = typename Cx<T2>::R;
    if constexpr (R == 2) {
        const T2 a = cadd(v[0], v[1]), b = csub(v[0], v[1]);
        v[0] = a;
        v[1] = b;
    } else if constexpr (R == 4) {
        const T2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        const T2 t2 = cadd(v[1], v[3]), t3 = cmuli(csub(v[1], v[3]), inv);
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    } else if constexpr (R == 3) {
        const Rl h = (Rl)0.5, s3 = (Rl)0.86602540378443864676;
        const T2 s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
        const T2 m = T2{v[0].x - h * s.x, v[0].y - h * s.y};
        const T2 rd = cmuli(cscale(d, s3), inv);   // -i s3 d (forward)
        v[0] = cadd(v[0], s);
        v[1] = cadd(m, rd);
        v[2] = csub(m, rd);
    } else {   // R == 5: pairs (1,4) and (2,3) share cosines; the sines
               // enter with -i (forward) on X1, X2 and +i on X4, X3
        static_assert(R == 5, "radix 2, 3, 4 or 5");
        const Rl c1 = (Rl)0.30901699437494742410, c2 = (Rl)-0.80901699437494742410;   // cos(2 pi/5), cos(4 pi/5)
        const Rl s1 = (Rl)0.95105651629515357212, s2 = (Rl)0.58778525229247312917;    // sin(2 pi/5), sin(4 pi/5)
        const T2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        const T2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
        const T2 a1 = T2{v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y};
        const T2 a2 = T2{v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y};
        const T2 b1 = cmuli(T2{s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y}, inv);
        const T2 b2 = cmuli(T2{s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y}, inv);
        v[0] = cadd(v[0], cadd(t1, t2));
        v[1] = cadd(a1, b1);
        v[4] = csub(a1, b1);
        v[2] = cadd(a2, b2);
        v[3] = csub(a2, b2);
    }
}

// twiddle accessors: w(t) = e^{-2 pi i t / n}, t < n
template <typename T2> struct TwStrided {   // table tw[t * stride]
    const T2 *__restrict__ tw;
    int stride;
    __device__ T2 operator()(int t) const { return tw[t * stride]; }
};
template <typename T2> struct TwHalf {   // first half of the table only (n even): w(t + n/2) = -w(t)
    const T2 *tw;
    int half;
    __device__ T2 operator()(int t) const
    {
        if (t < half) return tw[t];
        const T2 w = tw[t - half];
        return T2{-w.x, -w.y};
    }
};

// one Stockham stage of radix R over L lines of n points (registers only:
// the radix is a compile-time constant)
// (nb and ns: the plan's quotients by n / R and by Ns)
template <int R, typename T2, typename TW>
__device__ inline void fft_stage(const T2 *in, T2 *out, int L, int n, const FastDiv &nbd, const FastDiv &nsd,
                                 int tstep, const TW &tw, bool inv)
{
    const int nb = nbd.d, Ns = nsd.d;
    for (int t = threadIdx.x; t < L * nb; t += blockDim.x) {
        const int l = fdiv(t, nbd), jj = t - l * nb;
        const T2 *src = in + l * n;
        T2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src[jj + r * nb];
        const int q = fdiv(jj, nsd), kk = jj - q * Ns;
#pragma unroll
        for (int r = 1; r < R; ++r) {
            T2 w = tw(r * kk * tstep);
            if (inv) w = cconj(w);
            v[r] = cmul(v[r], w);
        }
        butterfly<R>(v, inv);
        T2 *dst = out + l * n + q * Ns * R + kk;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[r * Ns] = v[r];
    }
}

// L lines of n points each, line l at a[l*n]; every thread of the block
// calls this.  Returns the buffer holding the result (a or b).  Ends with a
// barrier.
template <typename T2, typename TW>
__device__ T2 *lds_fft_tw(T2 *a, T2 *b, int L, const FftPlan &pl, const TW &tw, bool inv)
{
    const int n = pl.n;
    T2 *in = a, *out = b;
    for (int st = 0; st < pl.nst; ++st) {
        const int R = pl.radix[st];
        if (R == 4) fft_stage<4>(in, out, L, n, pl.nb[st], pl.ns[st], pl.tstep[st], tw, inv);
        else if (R == 2) fft_stage<2>(in, out, L, n, pl.nb[st], pl.ns[st], pl.tstep[st], tw, inv);
        else if (R == 3) fft_stage<3>(in, out, L, n, pl.nb[st], pl.ns[st], pl.tstep[st], tw, inv);
        else fft_stage<5>(in, out, L, n, pl.nb[st], pl.ns[st], pl.tstep[st], tw, inv);
        __syncthreads();
        T2 *tmp = in;
        in = out;
        out = tmp;
    }
    return in;
}

// tw[t*tws] = e^{-2 pi i t / n}
template <typename T2>
__device__ T2 *lds_fft(T2 *a, T2 *b, int L, const FftPlan &pl, const T2 *__restrict__ tw, int tws, bool inv)
{
    return lds_fft_tw(a, b, L, pl, TwStrided<T2>{tw, tws}, inv);
}

inline FftPlan make_plan(int n)
{
    FftPlan p{};
    p.n = n;
    int m = n;
    while (m % 4 == 0) { p.radix[p.nst++] = 4; m /= 4; }
    while (m % 2 == 0) { p.radix[p.nst++] = 2; m /= 2; }
    while (m % 3 == 0) { p.radix[p.nst++] = 3; m /= 3; }
    while (m % 5 == 0) { p.radix[p.nst++] = 5; m /= 5; }
    if (m != 1 || n > FFT_MAX_N) {   // unsupported length
        p.nst = -1;
        return p;
    }
    int ns = 1;
    for (int st = 0; st < p.nst; ++st) {
        p.nb[st] = make_fastdiv(n / p.radix[st]);
        p.ns[st] = make_fastdiv(ns);
        p.tstep[st] = n / (ns * p.radix[st]);
        ns *= p.radix[st];
    }
    p.n1 = make_fastdiv(n);
    p.n2 = make_fastdiv(2 * n);
    return p;
}

}  // namespace epic
