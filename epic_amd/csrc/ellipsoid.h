// FP64 ellipsoid geometry on the device.
//
//  * B33 from the volume constraint det B = (3V/4pi)^2
//    (parcel_ellipsoid.f90:383-402, get_abc :409-415)
//  * eigen-decomposition of the 3x3 shape matrix by the Scherzinger-Dohrmann
//    algorithm with the reference's branch structure (scherzinger.f90:31-243):
//    diagonal shortcut, maxloc tie-breaking (first maximum) and the
//    three-swap descending sort, so near-degenerate shapes pick the same basis
//    (column lengths are compared squared: two whose lengths round to one
//    FP64 value but whose squares differ take the longer, an equally valid
//    pivot, where the first would be taken)
//  * half axes Veta = rho sqrt(|a2 - c2|) v_a, Vtau = rho sqrt(|b2 - c2|) v_b,
//    rho = sqrt(2/5), and the four support points x + Veta cos(th_j) +
//    Vtau sin(th_j), th_j = (2j+1) pi/4 (parcel_ellipsoid.f90:29-34, 216-260)
//  * dB/dt = L B + B L^T on the 5 stored components (:466-507)
#pragma once

#include "common.h"

namespace epic {

// sqrt(2/5), cos/sin((2j+1) pi/4)
#define EPIC_RHO 0.63245553203367586640
#define EPIC_RT2I 0.70710678118654752440

// FP64 reciprocal and reciprocal square root: the hardware estimates
// refined by two Newton steps each (within an ulp or so of the correctly
// rounded values, in a third of the instructions of the IEEE division and
// square root sequences).  x > 0 for the square roots, x != 0 for rcp_nr.
__device__ inline double rcp_nr(double x)
{
    double y = __builtin_amdgcn_rcp(x);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double e = fma(-x, y, 1.0);
        y = fma(y, e, y);
    }
    return y;
}
__device__ inline double rsqrt_nr(double x)
{
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double r = fma(-x * y, y, 1.0);   // 1 - x y^2
        y = fma(0.5 * y, r, y);
    }
    return y;
}
__device__ inline double sqrt_nr(double x) { return x > 0.0 ? x * rsqrt_nr(x) : 0.0; }

__device__ inline bool b33_of(const double B[5], double vol, double &b33)
{
    const double abc = 0.75 * vol * (1.0 / 3.14159265358979323846);
    const double den = B[0] * B[3] - B[1] * B[1];
    if (fabs(den) <= 2.220446049250313e-16) return false;
    b33 = (abc * abc - B[2] * (B[1] * B[4] - B[2] * B[3]) + B[0] * B[4] * B[4] - B[1] * B[2] * B[4]) * rcp_nr(den);
    return true;
}

__device__ inline double dsq3(double a, double b, double c) { return a * a + b * b + c * c; }
__device__ inline double dnorm3(double a, double b, double c) { return sqrt(dsq3(a, b, c)); }

// single-precision literal 1.0e-12 widened, as declared by the reference
#define EPIC_SCH_ATOL ((double)1.0e-12f)

// A: full symmetric matrix a[r][c] (row major); D descending; V[c][i] = i-th
// component of eigenvector c.
__device__ inline void scherzinger(const double A_in[9], double D[3], double V[3][3])
{
    double A[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) A[i] = A_in[i];
    const double off = fabs(A[1]) + fabs(A[2]) + fabs(A[5]);
    if (off < EPIC_SCH_ATOL) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int i = 0; i < 3; ++i) V[c][i] = (c == i) ? 1.0 : 0.0;
        D[0] = A[0];
        D[1] = A[4];
        D[2] = A[8];
    } else {
        // deviatoric part and invariants
        const double tr = A[0] + A[4] + A[8];
        const double t3 = (1.0 / 3.0) * tr;
        A[0] -= t3;
        A[4] -= t3;
        A[8] -= t3;
        // trace(A A) with the upper-triangle product of the reference (dsymm)
        const double aa00 = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
        const double aa11 = A[1] * A[1] + A[4] * A[4] + A[5] * A[5];
        const double aa22 = A[2] * A[2] + A[5] * A[5] + A[8] * A[8];
        const double j2 = 0.5 * (aa00 + aa11 + aa22);
        const double j3 = A[0] * (A[4] * A[8] - A[5] * A[5]) - A[1] * (A[1] * A[8] - A[2] * A[5]) +
                          A[2] * (A[1] * A[5] - A[2] * A[4]);
        // (3/j2)^(3/2) = 3 sqrt(3) j2^(-3/2)
        const double rj = rsqrt_nr(j2);
        double smp = (0.5 * 5.1961524227066318806) * j3 * (rj * rj * rj);
        smp = fmin(fmax(-1.0, smp), 1.0);
        // cos(alpha), alpha = acos(smp)/3, shifted by 2 pi/3 when alpha >
        // pi/6 (scherzinger.f90:175-181): by cos(3a) = 4 cos^3 a - 3 cos a it
        // is the root of 4c^3 - 3c = smp of largest magnitude, with the sign
        // of smp (+sqrt(3)/2 at 0).  That root of 4c^3 - 3c = |smp| lies in
        // [sqrt(3)/2, 1]; a degree-4 least-squares fit of cos(acos(s)/3) on
        // [0, 1] starts within 1e-5 of it, and two Newton steps (error
        // ratio below 1.7 e^2 there) reach it to FP64 rounding; the
        // derivative's reciprocal only needs the hardware estimate (a
        // relative error d in it adds d e to the step's error).
        const double sabs = fabs(smp);
        double cr = 0.8660344784 + sabs * (0.1663758323 + sabs * (-0.04585590254 + sabs * (0.01750308532 + sabs * -0.004064325447)));
#pragma unroll
        for (int it = 0; it < 2; ++it)
            cr = cr - ((4.0 * cr * cr - 3.0) * cr - sabs) * __builtin_amdgcn_rcp(12.0 * cr * cr - 3.0);
        const double cos_alpha = (smp < 0.0) ? -cr : cr;
        // 2 sqrt(j2 / 3) = (2 / sqrt(3)) j2 / sqrt(j2)
        const double eta1 = 1.1547005383792515290 * (j2 * rj) * cos_alpha;

        double r[3][3];   // r[c] = column c of A - eta1 I
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int i = 0; i < 3; ++i) r[c][i] = A[3 * i + c] - (i == c ? eta1 : 0.0);
        // the longest column (first of equals), compared by squared length
        double l0 = dsq3(r[0][0], r[0][1], r[0][2]);
        double l1 = dsq3(r[1][0], r[1][1], r[1][2]);
        double l2 = dsq3(r[2][0], r[2][1], r[2][2]);
        int im = 0;
        double lm = l0;
        if (l1 > lm) { im = 1; lm = l1; }
        if (l2 > lm) { im = 2; lm = l2; }
        double s1[3];
        const double ilm = rsqrt_nr(lm);
#pragma unroll
        for (int i = 0; i < 3; ++i) s1[i] = r[im][i] * ilm;
        if (im == 1) {
#pragma unroll
            for (int i = 0; i < 3; ++i) r[1][i] = r[0][i];
        } else if (im == 2) {
#pragma unroll
            for (int i = 0; i < 3; ++i) r[2][i] = r[0][i];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) r[0][i] = s1[i];
        double t = r[0][0] * r[1][0] + r[0][1] * r[1][1] + r[0][2] * r[1][2];
        double s = r[0][0] * r[2][0] + r[0][1] * r[2][1] + r[0][2] * r[2][2];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            r[1][i] = r[1][i] - t * r[0][i];
            r[2][i] = r[2][i] - s * r[0][i];
        }
        l1 = dsq3(r[1][0], r[1][1], r[1][2]);
        l2 = dsq3(r[2][0], r[2][1], r[2][2]);
        {
            const bool use2 = l2 > l1;
            const double il = rsqrt_nr(use2 ? l2 : l1);
#pragma unroll
            for (int i = 0; i < 3; ++i) r[1][i] = (use2 ? r[2][i] : r[1][i]) * il;
        }
        double a1[3], a2[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            a1[i] = A[3 * i + 0] * r[0][0] + A[3 * i + 1] * r[0][1] + A[3 * i + 2] * r[0][2];
            a2[i] = A[3 * i + 0] * r[1][0] + A[3 * i + 1] * r[1][1] + A[3 * i + 2] * r[1][2];
        }
        const double s1As1 = r[0][0] * a1[0] + r[0][1] * a1[1] + r[0][2] * a1[2];
        const double s2As1 = r[1][0] * a1[0] + r[1][1] * a1[1] + r[1][2] * a1[2];
        const double s2As2 = r[1][0] * a2[0] + r[1][1] * a2[1] + r[1][2] * a2[2];
        t = s1As1 - s2As2;
        s = s1As1 + s2As2;
        const double sg = (t == 0.0) ? 0.0 : copysign(1.0, t);
        const double eta2 = 0.5 * s - 0.5 * sg * sqrt_nr(t * t + 4.0 * s2As1 * s2As1);
        const double eta3 = s - eta2;
        D[0] = eta1 + t3;
        D[1] = eta2 + t3;
        D[2] = eta3 + t3;

        // eigenvectors (scherzinger.f90:60-82): shift by eta2
        A[0] -= eta2;
        A[4] -= eta2;
        A[8] -= eta2;
        V[0][0] = r[0][1] * r[1][2] - r[0][2] * r[1][1];
        V[0][1] = r[0][2] * r[1][0] - r[0][0] * r[1][2];
        V[0][2] = r[0][0] * r[1][1] - r[0][1] * r[1][0];
        double u0[3], u1[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            u0[i] = A[3 * i + 0] * r[0][0] + A[3 * i + 1] * r[0][1] + A[3 * i + 2] * r[0][2];
            u1[i] = A[3 * i + 0] * r[1][0] + A[3 * i + 1] * r[1][1] + A[3 * i + 2] * r[1][2];
        }
        const double m0 = dsq3(u0[0], u0[1], u0[2]);
        const double m1 = dsq3(u1[0], u1[1], u1[2]);
        {
            const bool use1 = m1 > m0;
            const double im_ = rsqrt_nr(use1 ? m1 : m0);
#pragma unroll
            for (int i = 0; i < 3; ++i) V[2][i] = (use1 ? u1[i] : u0[i]) * im_;
        }
        V[1][0] = V[2][1] * V[0][2] - V[2][2] * V[0][1];
        V[1][1] = V[2][2] * V[0][0] - V[2][0] * V[0][2];
        V[1][2] = V[2][0] * V[0][1] - V[2][1] * V[0][0];
        V[2][0] = V[0][1] * V[1][2] - V[0][2] * V[1][1];
        V[2][1] = V[0][2] * V[1][0] - V[0][0] * V[1][2];
        V[2][2] = V[0][0] * V[1][1] - V[0][1] * V[1][0];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double il = rsqrt_nr(dsq3(V[c][0], V[c][1], V[c][2]));
#pragma unroll
            for (int i = 0; i < 3; ++i) V[c][i] *= il;
        }
    }
    // three-swap descending sort (scherzinger.f90:85-110, 270-300)
    auto swp = [&](int a, int b) {
        double t = D[a];
        D[a] = D[b];
        D[b] = t;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double v = V[a][i];
            V[a][i] = V[b][i];
            V[b][i] = v;
        }
    };
    if (D[1] > D[0]) swp(0, 1);
    if (D[2] > D[1]) swp(1, 2);
    if (D[1] > D[0]) swp(0, 1);
}

// half axes of the support-point ellipse; false if B33 is degenerate
__device__ inline bool support_axes(const double B[5], double vol, double veta[3], double vtau[3])
{
    double b33;
    if (!b33_of(B, vol, b33)) return false;
    const double A[9] = {B[0], B[1], B[2], B[1], B[3], B[4], B[2], B[4], b33};
    double D[3], V[3][3];
    scherzinger(A, D, V);
    const double se = sqrt_nr(fabs(D[0] - D[2])) * EPIC_RHO;
    const double st = sqrt_nr(fabs(D[1] - D[2])) * EPIC_RHO;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        veta[i] = se * V[0][i];
        vtau[i] = st * V[1][i];
    }
    return true;
}

// cos and sin of (2j+1) pi/4 for j = 0..3
__device__ inline double support_cos(int j) { return (j == 0 || j == 3) ? EPIC_RT2I : -EPIC_RT2I; }
__device__ inline double support_sin(int j) { return (j < 2) ? EPIC_RT2I : -EPIC_RT2I; }

// dB/dt from the velocity-gradient components S (dudx dudy dudz dvdx dvdy dvdz dwdx dwdy)
__device__ inline bool dBdt(const double B[5], const double S[8], double vol, double out[5])
{
    double b33;
    if (!b33_of(B, vol, b33)) return false;
    const double dwdz = -(S[DUDX] + S[DVDY]);
    out[0] = 2.0 * (S[DUDX] * B[0] + S[DUDY] * B[1] + S[DUDZ] * B[2]);
    out[1] = S[DVDX] * B[0] - dwdz * B[1] + S[DVDZ] * B[2] + S[DUDY] * B[3] + S[DUDZ] * B[4];
    out[2] = S[DWDX] * B[0] + S[DWDY] * B[1] - S[DVDY] * B[2] + S[DUDY] * B[4] + S[DUDZ] * b33;
    out[3] = 2.0 * (S[DVDX] * B[1] + S[DVDY] * B[3] + S[DVDZ] * B[4]);
    out[4] = S[DWDX] * B[1] + S[DVDX] * B[2] + S[DWDY] * B[3] - S[DUDX] * B[4] + S[DVDZ] * b33;
    return true;
}

}  // namespace epic
