/*
 * epic_oracle.c -- FP64 CPU restatement of EPIC 0.14.3's per-timestep
 * parcel-grid core.  TEST INFRASTRUCTURE: the checker for the HIP path and
 * the timed CPU baseline; never part of the product library.
 *
 * Each function cites the reference file:line it restates.  The transforms'
 * defining sums and normalisations follow stafft.f90; their radix internals
 * are not restated (a plain mixed-radix DFT computes the same sums).
 *
 * Parity status of each part (what pins it):
 *   scherzinger_*      golden vectors unit-tests/3d/{B,D,V}.asc (1e-12) and
 *                      the amdflang build of scherzinger.f90 under oracle/_ref
 *   forfft/revfft/dct/dst  round trips, textbook DST-I/DCT-I sums, and the
 *                      amdflang build of stafft.f90 under oracle/_ref
 *   trilinear/par2grid sum rules of test_mpi_trilinear.f90, brute force
 *   grid2par           constant fields exact (test_mpi_grid2par.f90)
 *   vor2vel/vel2vgrad  Taylor-Green closed forms (test_mpi_vor2vel.f90,
 *                      test_mpi_vel2vgrad.f90)
 *   vorticity_tendency Beltrami closed form (test_mpi_vtend.f90)
 *   ls_rk_step         RK4 order of convergence on a closed-form ODE,
 *                      conservation of volume (det B), solenoidality
 *   nearest_cell_sort  brute-force stable sort
 */
#include "epic_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* constants.f90:8-46 */
static const double zero = 0.0, one = 1.0, two = 2.0, three = 3.0, four = 4.0;
#define ORC_PI 3.14159265358979323846264338327950288
static const double f12 = 1.0 / 2.0, f13 = 1.0 / 3.0, f14 = 1.0 / 4.0, f23 = 2.0 / 3.0,
                    f32 = 3.0 / 2.0, f34 = 3.0 / 4.0;

static double now_s(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

static orc_timers g_timers;
void orc_get_timers(orc_timers *t) { *t = g_timers; }
void orc_reset_timers(void) { memset(&g_timers, 0, sizeof g_timers); }

/* ------------------------------------------------------------------------ */
/* domain (parameters.f90:93-165)                                           */
/* ------------------------------------------------------------------------ */
int orc_dom_init(orc_dom *d)
{
    int n3[3] = {d->nx, d->ny, d->nz};
    double vc = 1.0;
    for (int i = 0; i < 3; ++i) {
        d->upper[i] = d->lower[i] + d->extent[i];
        d->dx[i] = d->extent[i] / (double)n3[i];
        d->dxi[i] = one / d->dx[i];
        d->center[i] = f12 * (d->lower[i] + d->upper[i]);
        d->hl[i] = d->extent[i] / two;
        d->hli[i] = one / d->hl[i];
        vc *= d->dx[i];
    }
    d->vcell = vc;
    if (d->vmin <= 0.0) d->vmin = d->vcell / 20.0;
    return ORC_OK;
}

int64_t orc_halo_size(const orc_dom *d)
{
    return (int64_t)(d->nz + 3) * (d->ny + 3) * (d->nx + 3);
}

int64_t orc_spec_size(const orc_dom *d)
{
    return (int64_t)(d->nz + 1) * d->ny * d->nx;
}

/* halo-field index f(iz, iy, ix), iz,iy,ix >= -1 */
#define HIDX(d, iz, iy, ix) \
    ((int64_t)((iz) + 1) + (int64_t)((d)->nz + 3) * ((int64_t)((iy) + 1) + (int64_t)((d)->ny + 3) * ((ix) + 1)))
/* spectral index s(iz, iy, ix), 0-based */
#define SIDX(d, iz, iy, ix) \
    ((int64_t)(iz) + (int64_t)((d)->nz + 1) * ((int64_t)(iy) + (int64_t)(d)->ny * (ix)))

static int nthreads_of(const orc_dom *d)
{
#ifdef _OPENMP
    return d->nthreads > 0 ? d->nthreads : omp_get_max_threads();
#else
    (void)d;
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* linalg.f90:8-71 and scherzinger.f90:31-300                               */
/* ------------------------------------------------------------------------ */
/* matrices as m[r][c] row-major in a flat [9]; all symmetric inputs.       */
#define M(a, r, c) ((a)[3 * (r) + (c)])

static double signum(double v)   /* linalg.f90:8-17 */
{
    double s = copysign(one, v);
    if (v == zero) s = zero;
    return s;
}

static double trace3(const double *a) { return M(a, 0, 0) + M(a, 1, 1) + M(a, 2, 2); }

static double det3(const double *a)   /* linalg.f90:45-53 */
{
    return M(a, 0, 0) * (M(a, 1, 1) * M(a, 2, 2) - M(a, 1, 2) * M(a, 1, 2))
         - M(a, 0, 1) * (M(a, 0, 1) * M(a, 2, 2) - M(a, 0, 2) * M(a, 1, 2))
         + M(a, 0, 2) * (M(a, 0, 1) * M(a, 1, 2) - M(a, 0, 2) * M(a, 1, 1));
}

/* C = A * B for symmetric A, B using only their upper triangles (linalg.f90:58-71) */
static void dsymm3(const double *a, const double *b, double *c)
{
    M(c, 0, 0) = M(a, 0, 0) * M(b, 0, 0) + M(a, 0, 1) * M(b, 0, 1) + M(a, 0, 2) * M(b, 0, 2);
    M(c, 0, 1) = M(a, 0, 0) * M(b, 0, 1) + M(a, 0, 1) * M(b, 1, 1) + M(a, 0, 2) * M(b, 1, 2);
    M(c, 0, 2) = M(a, 0, 0) * M(b, 0, 2) + M(a, 0, 1) * M(b, 1, 2) + M(a, 0, 2) * M(b, 2, 2);
    M(c, 1, 0) = M(a, 0, 1) * M(b, 0, 0) + M(a, 1, 1) * M(b, 0, 1) + M(a, 1, 2) * M(b, 0, 2);
    M(c, 1, 1) = M(a, 0, 1) * M(b, 0, 1) + M(a, 1, 1) * M(b, 1, 1) + M(a, 1, 2) * M(b, 1, 2);
    M(c, 1, 2) = M(a, 0, 1) * M(b, 0, 2) + M(a, 1, 1) * M(b, 1, 2) + M(a, 1, 2) * M(b, 2, 2);
    M(c, 2, 0) = M(a, 0, 2) * M(b, 0, 0) + M(a, 1, 2) * M(b, 0, 1) + M(a, 2, 2) * M(b, 0, 2);
    M(c, 2, 1) = M(a, 0, 2) * M(b, 0, 1) + M(a, 1, 2) * M(b, 1, 1) + M(a, 2, 2) * M(b, 1, 2);
    M(c, 2, 2) = M(a, 0, 2) * M(b, 0, 2) + M(a, 1, 2) * M(b, 1, 2) + M(a, 2, 2) * M(b, 2, 2);
}

static void cross3(const double *a, const double *b, double *c)   /* linalg.f90:22-29 */
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double norm3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

/* the reference declares atol = 1.0e-12 as a default-real literal
 * (scherzinger.f90:16), i.e. the single-precision value widened to double */
static const double SCH_ATOL = (double)1.0e-12f;

/* r columns stored as r[3*c + i] */
static void evaluate_eigenvalues(double *A, double *r, double eta[3], double D[3])
{
    double AA[9];
    double tr = trace3(A);
    double tmp = f13 * tr;
    M(A, 0, 0) -= tmp;
    M(A, 1, 1) -= tmp;
    M(A, 2, 2) -= tmp;

    dsymm3(A, A, AA);
    double j2 = f12 * trace3(AA);
    double j3 = det3(A);

    double smp = f12 * j3 * pow(three / j2, f32);
    smp = fmin(fmax(-one, smp), one);
    double alpha = f13 * acos(smp);
    if (alpha > ORC_PI / 6.0) alpha = alpha + f23 * ORC_PI;

    eta[0] = two * sqrt(f13 * j2) * cos(alpha);

    /* eq 19: r(:,c) = A(:,c) - eta1 e_c */
    for (int c = 0; c < 3; ++c)
        for (int i = 0; i < 3; ++i) r[3 * c + i] = M(A, i, c) - (i == c ? eta[0] : 0.0);

    double rlen[3];
    for (int c = 0; c < 3; ++c) rlen[c] = norm3(r + 3 * c);
    int im = 0;   /* maxloc: first maximum */
    if (rlen[1] > rlen[im]) im = 1;
    if (rlen[2] > rlen[im]) im = 2;
    double s1[3];
    for (int i = 0; i < 3; ++i) s1[i] = r[3 * im + i] / rlen[im];
    if (im == 1) memcpy(r + 3, r, 3 * sizeof(double));
    else if (im == 2) memcpy(r + 6, r, 3 * sizeof(double));
    memcpy(r, s1, 3 * sizeof(double));

    tmp = dot3(r, r + 3);
    smp = dot3(r, r + 6);
    for (int i = 0; i < 3; ++i) {
        r[3 + i] = r[3 + i] - tmp * r[i];
        r[6 + i] = r[6 + i] - smp * r[i];
    }
    for (int c = 0; c < 3; ++c) rlen[c] = norm3(r + 3 * c);
    int i2 = (rlen[2] > rlen[1]) ? 2 : 1;
    for (int i = 0; i < 3; ++i) r[3 + i] = r[3 * i2 + i] / rlen[i2];

    /* eq 22 */
    double a1[3], a2[3];
    for (int i = 0; i < 3; ++i) {
        a1[i] = M(A, i, 0) * r[0] + M(A, i, 1) * r[1] + M(A, i, 2) * r[2];
        a2[i] = M(A, i, 0) * r[3] + M(A, i, 1) * r[4] + M(A, i, 2) * r[5];
    }
    double s1As1 = dot3(r, a1);
    double s2As1 = dot3(r + 3, a1);
    double s1As2 = s2As1;
    double s2As2 = dot3(r + 3, a2);

    tmp = s1As1 - s2As2;
    smp = s1As1 + s2As2;
    eta[1] = f12 * smp - f12 * signum(tmp) * sqrt(tmp * tmp + four * s1As2 * s2As1);
    eta[2] = smp - eta[1];
    for (int i = 0; i < 3; ++i) D[i] = eta[i] + f13 * tr;
}

static void swap_d(double *a, double *b) { double t = *a; *a = *b; *b = t; }
static void swap_col(double *V, int c0, int c1)
{
    for (int i = 0; i < 3; ++i) swap_d(&V[3 * c0 + i], &V[3 * c1 + i]);
}

/* scherzinger.f90:270-300 */
static void sort_descending(double D[3], double *V)
{
    if (D[1] > D[0]) { swap_d(&D[0], &D[1]); swap_col(V, 0, 1); }
    if (D[2] > D[1]) { swap_d(&D[1], &D[2]); swap_col(V, 1, 2); }
    if (D[1] > D[0]) { swap_d(&D[0], &D[1]); swap_col(V, 0, 1); }
}

/* scherzinger.f90:31-112 */
void orc_scherzinger_diagonalise(const double Ain[9], double D[3], double V[9])
{
    double A[9], r[9], eta[3];
    memcpy(A, Ain, sizeof A);
    double tmp = fabs(M(A, 0, 1)) + fabs(M(A, 0, 2)) + fabs(M(A, 1, 2));
    if (tmp < SCH_ATOL) {
        memset(V, 0, 9 * sizeof(double));
        V[0] = V[4] = V[8] = one;
        D[0] = M(A, 0, 0);
        D[1] = M(A, 1, 1);
        D[2] = M(A, 2, 2);
        sort_descending(D, V);
        return;
    }
    evaluate_eigenvalues(A, r, eta, D);
    M(A, 0, 0) -= eta[1];
    M(A, 1, 1) -= eta[1];
    M(A, 2, 2) -= eta[1];

    cross3(r, r + 3, V);   /* V(:,1), eq 21 */

    double u[6];   /* r(:,1:2) = matmul(A, r(:,1:2)), eq 30 */
    for (int c = 0; c < 2; ++c)
        for (int i = 0; i < 3; ++i)
            u[3 * c + i] = M(A, i, 0) * r[3 * c + 0] + M(A, i, 1) * r[3 * c + 1] + M(A, i, 2) * r[3 * c + 2];
    double l0 = norm3(u), l1 = norm3(u + 3);
    int im = (l1 > l0) ? 1 : 0;
    double lim = im ? l1 : l0;
    for (int i = 0; i < 3; ++i) V[6 + i] = u[3 * im + i] / lim;

    cross3(V + 6, V, V + 3);   /* V(:,2) = V(:,3) x V(:,1), eq 31 */
    cross3(V, V + 3, V + 6);   /* V(:,3) = V(:,1) x V(:,2) */
    for (int c = 0; c < 3; ++c) {
        double l = norm3(V + 3 * c);
        for (int i = 0; i < 3; ++i) V[3 * c + i] /= l;
    }
    sort_descending(D, V);
}

/* scherzinger.f90:122-142 */
void orc_scherzinger_eigenvalues(const double Ain[9], double D[3])
{
    double A[9], r[9], eta[3], dummy[9];
    memcpy(A, Ain, sizeof A);
    double tmp = fabs(M(A, 0, 1)) + fabs(M(A, 0, 2)) + fabs(M(A, 1, 2));
    if (tmp < SCH_ATOL) {
        D[0] = M(A, 0, 0);
        D[1] = M(A, 1, 1);
        D[2] = M(A, 2, 2);
    } else {
        evaluate_eigenvalues(A, r, eta, D);
    }
    memset(dummy, 0, sizeof dummy);
    sort_descending(D, dummy);
}

/* ------------------------------------------------------------------------ */
/* parcel_ellipsoid.f90                                                     */
/* ------------------------------------------------------------------------ */
enum { B11 = 0, B12 = 1, B13 = 2, B22 = 3, B23 = 4 };

/* parcel_ellipsoid.f90:383-402 (get_abc :409-415: abc = f34 * volume * fpi) */
int orc_get_B33(const double B[5], double volume, double *B33)
{
    double abc = f34 * volume * (one / ORC_PI);
    double den = B[B11] * B[B22] - B[B12] * B[B12];
    if (fabs(den) <= 2.220446049250313e-16) return ORC_E_B33;   /* epsilon(abc) */
    *B33 = (abc * abc - B[B13] * (B[B12] * B[B23] - B[B13] * B[B22]) + B[B11] * B[B23] * B[B23]
            - B[B12] * B[B13] * B[B23]) / den;
    return ORC_OK;
}

/* support-point geometry, parcel_ellipsoid.f90:29-34 */
static double rho_c(void) { return sqrt(two / 5.0); }

/* parcel_ellipsoid.f90:216-260 (recompute branch) + :359-376 */
int orc_get_points(const double pos[3], const double B[5], double volume,
                   double Veta[3], double Vtau[3], double pts[12])
{
    double B33, A[9], D[3], V[9];
    int rc = orc_get_B33(B, volume, &B33);
    if (rc) return rc;
    M(A, 0, 0) = B[B11]; M(A, 0, 1) = B[B12]; M(A, 0, 2) = B[B13];
    M(A, 1, 0) = B[B12]; M(A, 1, 1) = B[B22]; M(A, 1, 2) = B[B23];
    M(A, 2, 0) = B[B13]; M(A, 2, 1) = B[B23]; M(A, 2, 2) = B33;
    orc_scherzinger_diagonalise(A, D, V);
    double rho = rho_c();
    double se = sqrt(fabs(D[0] - D[2])) * rho, st = sqrt(fabs(D[1] - D[2])) * rho;
    for (int i = 0; i < 3; ++i) {
        Veta[i] = se * V[i];
        Vtau[i] = st * V[3 + i];
    }
    if (pts) {
        const double pi4 = ORC_PI / 4.0;
        for (int j = 0; j < 4; ++j) {
            double th = (double)(2 * j + 1) * pi4;
            double c = cos(th), s = sin(th);
            for (int i = 0; i < 3; ++i) pts[3 * j + i] = pos[i] + Veta[i] * c + Vtau[i] * s;
        }
    }
    return ORC_OK;
}

static void points_from_V(const double pos[3], const double Veta[3], const double Vtau[3], double pts[12])
{
    const double pi4 = ORC_PI / 4.0;
    for (int j = 0; j < 4; ++j) {
        double th = (double)(2 * j + 1) * pi4;
        double c = cos(th), s = sin(th);
        for (int i = 0; i < 3; ++i) pts[3 * j + i] = pos[i] + Veta[i] * c + Vtau[i] * s;
    }
}

/* parcel_ellipsoid.f90:466-507; strain order fields.f90:54-61 */
int orc_get_dBdt(const double B[5], const double S[8], double volume, double dB[5])
{
    enum { DUDX, DUDY, DUDZ, DVDX, DVDY, DVDZ, DWDX, DWDY };
    double B33;
    int rc = orc_get_B33(B, volume, &B33);
    if (rc) return rc;
    double dwdz = -(S[DUDX] + S[DVDY]);
    dB[B11] = two * (S[DUDX] * B[B11] + S[DUDY] * B[B12] + S[DUDZ] * B[B13]);
    dB[B12] = S[DVDX] * B[B11] - dwdz * B[B12] + S[DVDZ] * B[B13] + S[DUDY] * B[B22] + S[DUDZ] * B[B23];
    dB[B13] = S[DWDX] * B[B11] + S[DWDY] * B[B12] - S[DVDY] * B[B13] + S[DUDY] * B[B23] + S[DUDZ] * B33;
    dB[B22] = two * (S[DVDX] * B[B12] + S[DVDY] * B[B22] + S[DVDZ] * B[B23]);
    dB[B23] = S[DWDX] * B[B12] + S[DVDX] * B[B13] + S[DWDY] * B[B22] - S[DUDX] * B[B23] + S[DVDZ] * B33;
    return ORC_OK;
}

/* interpl.f90:17-56; weights w[k*4 + j*2 + i] = ww(k, j, i) */
void orc_trilinear(const orc_dom *d, const double pos[3], int ijk[3], double w[8])
{
    double xyz[3], p[3], pc[3];
    for (int a = 0; a < 3; ++a) {
        xyz[a] = (pos[a] - d->lower[a]) * d->dxi[a];
        ijk[a] = (int)floor(xyz[a]);
        p[a] = xyz[a] - (double)ijk[a];
        pc[a] = one - p[a];
    }
    double w00 = pc[1] * pc[0], w10 = pc[1] * p[0], w01 = p[1] * pc[0], w11 = p[1] * p[0];
    w[0] = pc[2] * w00;   /* (0,0,0) */
    w[1] = pc[2] * w10;   /* (0,0,1) */
    w[2] = pc[2] * w01;   /* (0,1,0) */
    w[3] = pc[2] * w11;   /* (0,1,1) */
    w[4] = p[2] * w00;    /* (1,0,0) */
    w[5] = p[2] * w10;
    w[6] = p[2] * w01;
    w[7] = p[2] * w11;
}

/* fields.f90:165-172 */
void orc_get_index(const orc_dom *d, const double pos[3], int ijk[3])
{
    for (int a = 0; a < 3; ++a) ijk[a] = (int)floor((pos[a] - d->lower[a]) * d->dxi[a]);
}

static int in_halo(const orc_dom *d, const int ijk[3])
{
    /* the 2x2x2 stencil must stay inside -1..n+1 */
    return ijk[0] >= -1 && ijk[0] <= d->nx && ijk[1] >= -1 && ijk[1] <= d->ny && ijk[2] >= -1 && ijk[2] <= d->nz;
}

/* ------------------------------------------------------------------------ */
/* single-rank halo semantics of field_mpi.f90                              */
/* ------------------------------------------------------------------------ */
static int wrapi(int i, int n) { int r = i % n; return r < 0 ? r + n : r; }

/* field_buffer_to_interior (accumulate, field_mpi.f90:965-1017): every x/y
 * halo point is added to its periodic image, over all z */
static void halo_accumulate(const orc_dom *d, double *f)
{
    int nx = d->nx, ny = d->ny, nz = d->nz;
    /* x halos first over interior y, then y halos over interior x, then the
     * corners -- the same grouping as the 8 neighbour buffers */
    for (int ix = -1; ix <= nx + 1; ++ix) {
        for (int iy = -1; iy <= ny + 1; ++iy) {
            int hx = (ix < 0 || ix >= nx), hy = (iy < 0 || iy >= ny);
            if (!hx && !hy) continue;
            int tx = wrapi(ix, nx), ty = wrapi(iy, ny);
            for (int iz = -1; iz <= nz + 1; ++iz) f[HIDX(d, iz, ty, tx)] += f[HIDX(d, iz, iy, ix)];
        }
    }
}

/* field_buffer_to_halo (fill): halo = periodic copy of the interior */
static void halo_fill(const orc_dom *d, double *f)
{
    int nx = d->nx, ny = d->ny, nz = d->nz;
    for (int ix = -1; ix <= nx + 1; ++ix)
        for (int iy = -1; iy <= ny + 1; ++iy) {
            int hx = (ix < 0 || ix >= nx), hy = (iy < 0 || iy >= ny);
            if (!hx && !hy) continue;
            int tx = wrapi(ix, nx), ty = wrapi(iy, ny);
            for (int iz = -1; iz <= nz + 1; ++iz) f[HIDX(d, iz, iy, ix)] = f[HIDX(d, iz, ty, tx)];
        }
}

static void halo_accumulate_int(const orc_dom *d, int *f)
{
    int nx = d->nx, ny = d->ny, nz = d->nz;
    for (int ix = -1; ix <= nx + 1; ++ix)
        for (int iy = -1; iy <= ny + 1; ++iy) {
            int hx = (ix < 0 || ix >= nx), hy = (iy < 0 || iy >= ny);
            if (!hx && !hy) continue;
            int tx = wrapi(ix, nx), ty = wrapi(iy, ny);
            for (int iz = -1; iz <= nz + 1; ++iz) f[HIDX(d, iz, ty, tx)] += f[HIDX(d, iz, iy, ix)];
        }
    for (int ix = -1; ix <= nx + 1; ++ix)
        for (int iy = -1; iy <= ny + 1; ++iy) {
            int hx = (ix < 0 || ix >= nx), hy = (iy < 0 || iy >= ny);
            if (!hx && !hy) continue;
            int tx = wrapi(ix, nx), ty = wrapi(iy, ny);
            for (int iz = -1; iz <= nz + 1; ++iz) f[HIDX(d, iz, iy, ix)] = f[HIDX(d, iz, ty, tx)];
        }
}

/* ------------------------------------------------------------------------ */
/* par2grid (parcel_interpl.f90:146-323)                                    */
/* ------------------------------------------------------------------------ */
static int parcel_points(const orc_dom *d, orc_parcels *p, int64_t n, int l_reuse, double pts[12])
{
    (void)d;
    double *pos = p->position + 3 * n;
    double *Ve = p->Vetas + 3 * n, *Vt = p->Vtaus + 3 * n;
    if (l_reuse == 1) {
        points_from_V(pos, Ve, Vt, pts);
        return ORC_OK;
    }
    return orc_get_points(pos, p->B + 5 * n, p->volume[n], Ve, Vt, pts);
}

int orc_par2grid(const orc_dom *d, orc_parcels *p, orc_fields *f, int l_reuse)
{
    double t0 = now_s();
    const int64_t H = orc_halo_size(d);
    const int nz = d->nz;
    const int nt = nthreads_of(d);
    /* 7 float fields reduced per thread: vol, xi, eta, zeta, tbuoy, dbuoy, hum */
    enum { NF = 7 };
    double *priv = (double *)calloc((size_t)nt * NF * H, sizeof(double));
    int *ipriv = (int *)calloc((size_t)nt * 2 * H, sizeof(int));
    if (!priv || !ipriv) { free(priv); free(ipriv); return ORC_E_ALLOC; }
    int err = ORC_OK;

#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *mine = priv + (int64_t)tid * NF * H;
        int *imine = ipriv + (int64_t)tid * 2 * H;
        int lerr = ORC_OK;
        /* static schedule: contiguous parcel ranges in thread order */
#pragma omp for schedule(static)
        for (int64_t n = 0; n < p->n; ++n) {
            if (lerr) continue;
            double pvol = p->volume[n];
            double z = p->position[3 * n + 2];
            double q_c = p->humidity[n] - d->q0 * exp(d->lambda_c * (d->lower[2] - z));
            q_c = fmax(zero, q_c);
            double btot = p->buoyancy[n] + d->glat * q_c;
            double pts[12];
            int rc = parcel_points(d, p, n, l_reuse, pts);
            if (rc) { lerr = rc; continue; }
            int c[3];
            orc_get_index(d, p->position + 3 * n, c);
            if (c[0] < -1 || c[0] > d->nx + 1 || c[1] < -1 || c[1] > d->ny + 1 || c[2] < -1 || c[2] > nz + 1) {
                lerr = ORC_E_OUT_OF_GRID;
                continue;
            }
            imine[HIDX(d, c[2], c[1], c[0])] += 1;
            if (p->volume[n] <= d->vmin) imine[H + HIDX(d, c[2], c[1], c[0])] += 1;
            for (int q = 0; q < 4; ++q) {
                int ijk[3];
                double w[8];
                orc_trilinear(d, pts + 3 * q, ijk, w);
                if (!in_halo(d, ijk)) { lerr = ORC_E_OUT_OF_GRID; break; }
                for (int kk = 0; kk < 2; ++kk)
                    for (int jj = 0; jj < 2; ++jj)
                        for (int ii = 0; ii < 2; ++ii) {
                            double wt = f14 * pvol * w[kk * 4 + jj * 2 + ii];
                            int64_t h = HIDX(d, ijk[2] + kk, ijk[1] + jj, ijk[0] + ii);
                            mine[1 * H + h] += wt * p->vorticity[3 * n + 0];
                            mine[2 * H + h] += wt * p->vorticity[3 * n + 1];
                            mine[3 * H + h] += wt * p->vorticity[3 * n + 2];
                            mine[5 * H + h] += wt * p->buoyancy[n];
                            mine[6 * H + h] += wt * p->humidity[n];
                            mine[4 * H + h] += wt * btot;
                            mine[0 * H + h] += wt;
                        }
            }
        }
        if (lerr) {
#pragma omp critical
            err = lerr;
        }
    }
    if (err) { free(priv); free(ipriv); return err; }

    /* ordered (deterministic) reduction of the thread-private copies */
    double *outs[NF] = {f->volg, f->vortg, f->vortg + H, f->vortg + 2 * H, f->tbuoyg, f->dbuoyg, f->humg};
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t h = 0; h < H; ++h) {
        for (int k = 0; k < NF; ++k) {
            double s = 0.0;
            for (int t = 0; t < nt; ++t) s += priv[((int64_t)t * NF + k) * H + h];
            outs[k][h] = s;
        }
        int a = 0, b = 0;
        for (int t = 0; t < nt; ++t) {
            a += ipriv[(int64_t)t * 2 * H + h];
            b += ipriv[(int64_t)t * 2 * H + H + h];
        }
        f->nparg[h] = a;
        f->nsparg[h] = b;
    }
    free(priv);
    free(ipriv);

    /* z halo of counts folded into cells 0 and nz-1 (:239-243) */
    for (int ix = -1; ix <= d->nx + 1; ++ix)
        for (int iy = -1; iy <= d->ny + 1; ++iy) {
            f->nparg[HIDX(d, 0, iy, ix)] += f->nparg[HIDX(d, -1, iy, ix)];
            f->nparg[HIDX(d, nz - 1, iy, ix)] += f->nparg[HIDX(d, nz, iy, ix)];
            f->nsparg[HIDX(d, 0, iy, ix)] += f->nsparg[HIDX(d, -1, iy, ix)];
            f->nsparg[HIDX(d, nz - 1, iy, ix)] += f->nsparg[HIDX(d, nz, iy, ix)];
        }
    /* sanity check (:248-250) */
    int64_t total = 0;
    for (int ix = -1; ix <= d->nx + 1; ++ix)
        for (int iy = -1; iy <= d->ny + 1; ++iy)
            for (int iz = 0; iz <= nz - 1; ++iz) total += f->nparg[HIDX(d, iz, iy, ix)];
    if (total != p->n) return ORC_E_PARCEL_COUNT;

    /* par2grid_halo_swap (:326-397), single rank */
    for (int k = 0; k < NF; ++k) { halo_accumulate(d, outs[k]); halo_fill(d, outs[k]); }
    halo_accumulate_int(d, f->nparg);
    halo_accumulate_int(d, f->nsparg);

    /* free-slip folds (:258-287) */
    for (int k = 0; k < NF; ++k) {
        double *g = outs[k];
        for (int ix = -1; ix <= d->nx + 1; ++ix)
            for (int iy = -1; iy <= d->ny + 1; ++iy) {
                g[HIDX(d, 0, iy, ix)] = two * g[HIDX(d, 0, iy, ix)];
                g[HIDX(d, nz, iy, ix)] = two * g[HIDX(d, nz, iy, ix)];
                g[HIDX(d, 1, iy, ix)] = g[HIDX(d, 1, iy, ix)] + g[HIDX(d, -1, iy, ix)];
                g[HIDX(d, nz - 1, iy, ix)] = g[HIDX(d, nz - 1, iy, ix)] + g[HIDX(d, nz + 1, iy, ix)];
            }
    }
    /* normalise, zeta BC, extrapolate (:290-319) */
    for (int ix = -1; ix <= d->nx + 1; ++ix)
        for (int iy = -1; iy <= d->ny + 1; ++iy) {
            for (int iz = 0; iz <= nz; ++iz) {
                int64_t h = HIDX(d, iz, iy, ix);
                double v = f->volg[h];
                for (int c = 0; c < 3; ++c) f->vortg[c * H + h] = f->vortg[c * H + h] / v;
            }
            if (d->zeta_zero[0]) f->vortg[2 * H + HIDX(d, 0, iy, ix)] = zero;
            if (d->zeta_zero[1]) f->vortg[2 * H + HIDX(d, nz, iy, ix)] = zero;
            for (int c = 0; c < 3; ++c) {
                double *g = f->vortg + c * H;
                g[HIDX(d, -1, iy, ix)] = two * g[HIDX(d, 0, iy, ix)] - g[HIDX(d, 1, iy, ix)];
                g[HIDX(d, nz + 1, iy, ix)] = two * g[HIDX(d, nz, iy, ix)] - g[HIDX(d, nz - 1, iy, ix)];
            }
            for (int iz = 0; iz <= nz; ++iz) {
                int64_t h = HIDX(d, iz, iy, ix);
                double v = f->volg[h];
                f->dbuoyg[h] = f->dbuoyg[h] / v;
                f->humg[h] = f->humg[h] / v;
                f->tbuoyg[h] = f->tbuoyg[h] / v;
            }
            f->tbuoyg[HIDX(d, -1, iy, ix)] = two * f->tbuoyg[HIDX(d, 0, iy, ix)] - f->tbuoyg[HIDX(d, 1, iy, ix)];
            f->tbuoyg[HIDX(d, nz + 1, iy, ix)] =
                two * f->tbuoyg[HIDX(d, nz, iy, ix)] - f->tbuoyg[HIDX(d, nz - 1, iy, ix)];
        }
    g_timers.par2grid += now_s() - t0;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* grid2par (parcel_interpl.f90:405-470)                                    */
/* ------------------------------------------------------------------------ */
int orc_grid2par(const orc_dom *d, orc_parcels *p, const orc_fields *f, int add)
{
    double t0 = now_s();
    const int64_t H = orc_halo_size(d);
    int reset = (add != 1);
    int err = ORC_OK;
#pragma omp parallel for schedule(static) num_threads(nthreads_of(d))
    for (int64_t n = 0; n < p->n; ++n) {
        double *dp = p->delta_pos + 3 * n, *dv = p->delta_vor + 3 * n, *st = p->strain + 8 * n;
        if (reset) {
            dp[0] = dp[1] = dp[2] = zero;
            dv[0] = dv[1] = dv[2] = zero;
        }
        for (int l = 0; l < 8; ++l) st[l] = zero;
        double pts[12];
        int rc = parcel_points(d, p, n, -1, pts);
        if (rc) {
#pragma omp critical
            err = rc;
            continue;
        }
        for (int q = 0; q < 4; ++q) {
            int ijk[3];
            double w[8];
            orc_trilinear(d, pts + 3 * q, ijk, w);
            if (!in_halo(d, ijk)) {
#pragma omp critical
                err = ORC_E_OUT_OF_GRID;
                break;
            }
            int64_t hs[8];
            for (int kk = 0; kk < 2; ++kk)
                for (int jj = 0; jj < 2; ++jj)
                    for (int ii = 0; ii < 2; ++ii)
                        hs[kk * 4 + jj * 2 + ii] = HIDX(d, ijk[2] + kk, ijk[1] + jj, ijk[0] + ii);
            /* sum(weights * field(ks:ks+1, js:js+1, is:is+1)) in Fortran
             * array-element order: k fastest, then j, then i */
            static const int ord[8] = {0, 4, 2, 6, 1, 5, 3, 7};
            for (int l = 0; l < 3; ++l) {
                double s = 0.0;
                for (int m = 0; m < 8; ++m) s += w[ord[m]] * f->velog[l * H + hs[ord[m]]];
                dp[l] = dp[l] + f14 * s;
            }
            for (int l = 0; l < 8; ++l) {
                double s = 0.0;
                for (int m = 0; m < 8; ++m) s += w[ord[m]] * f->velgradg[l * H + hs[ord[m]]];
                st[l] = st[l] + f14 * s;
            }
            for (int l = 0; l < 3; ++l) {
                double s = 0.0;
                for (int m = 0; m < 8; ++m) s += w[ord[m]] * f->vtend[l * H + hs[ord[m]]];
                dv[l] = dv[l] + f14 * s;
            }
        }
    }
    g_timers.grid2par += now_s() - t0;
    return err;
}

/* ------------------------------------------------------------------------ */
/* 1D transforms.                                                           */
/* forfft (stafft.f90:196-294): x -> Hermitian form, normalised 1/sqrt(n):  */
/*   h[k]   =  Re X_k / sqrt(n),  k = 0..n/2                                */
/*   h[n-k] =  Im X_k / sqrt(n),  k = 1..(n-1)/2,  X_k = sum x_j e^{-2pi i jk/n} */
/* revfft (:296-405) is its exact inverse.  The radix kernels of the         */
/* reference are not restated: a mixed-radix DFT evaluates the same sums.    */
/* ------------------------------------------------------------------------ */
typedef double complex cplx;

static int factor_ok(int n)
{
    if (n < 2) return 0;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    while (n % 5 == 0) n /= 5;
    return n == 1;
}

static int smallest_factor(int n)
{
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    if (n % 3 == 0) return 3;
    return 5;
}

/* recursive decimation-in-time DFT: out[k] = sum_j in[j*is] W^{jk}, W = e^{-2 pi i/n} */
static void dft_rec(int n, const cplx *in, int is, cplx *out, int N, const cplx *tw)
{
    if (n == 1) { out[0] = in[0]; return; }
    int p = smallest_factor(n), m = n / p;
    for (int r = 0; r < p; ++r) dft_rec(m, in + (int64_t)r * is, is * p, out + r * m, N, tw);
    int step = N / n;
    cplx t[5];
    for (int k = 0; k < m; ++k) {
        for (int r = 0; r < p; ++r) t[r] = out[r * m + k] * tw[((int64_t)r * k * step) % N];
        for (int q = 0; q < p; ++q) {
            cplx s = 0;
            for (int r = 0; r < p; ++r) s += t[r] * tw[((int64_t)r * q * (N / p)) % N];
            out[k + m * q] = s;
        }
    }
}

static cplx *twiddles(int n)
{
    cplx *tw = (cplx *)malloc(sizeof(cplx) * n);
    for (int t = 0; t < n; ++t) {
        double a = 2.0 * ORC_PI * (double)t / (double)n;
        tw[t] = cos(a) - I * sin(a);
    }
    return tw;
}

/* thread-local cache of twiddle tables per length */
#define TW_CACHE 8
static _Thread_local int tw_n[TW_CACHE];
static _Thread_local cplx *tw_t[TW_CACHE];
static const cplx *get_tw(int n)
{
    for (int i = 0; i < TW_CACHE; ++i)
        if (tw_n[i] == n) return tw_t[i];
    for (int i = 0; i < TW_CACHE; ++i)
        if (tw_n[i] == 0) { tw_n[i] = n; tw_t[i] = twiddles(n); return tw_t[i]; }
    free(tw_t[0]);
    tw_n[0] = n;
    tw_t[0] = twiddles(n);
    return tw_t[0];
}

static void dft(int n, const cplx *in, cplx *out)
{
    dft_rec(n, in, 1, out, n, get_tw(n));
}

int orc_forfft(int n, double *x)
{
    if (!factor_ok(n)) return ORC_E_FFT_LEN;
    cplx buf[2 * 1024], *a = buf, *X = buf + n;
    cplx *heap = NULL;
    if (n > 1024) { heap = (cplx *)malloc(sizeof(cplx) * 2 * n); a = heap; X = heap + n; }
    for (int j = 0; j < n; ++j) a[j] = x[j];
    dft(n, a, X);
    double nf = one / sqrt((double)n);
    for (int k = 0; k <= n / 2; ++k) x[k] = creal(X[k]) * nf;
    for (int k = 1; k <= (n - 1) / 2; ++k) x[n - k] = cimag(X[k]) * nf;
    free(heap);
    return ORC_OK;
}

int orc_revfft(int n, double *x)
{
    if (!factor_ok(n)) return ORC_E_FFT_LEN;
    cplx buf[2 * 1024], *X = buf, *a = buf + n;
    cplx *heap = NULL;
    if (n > 1024) { heap = (cplx *)malloc(sizeof(cplx) * 2 * n); X = heap; a = heap + n; }
    double sn = sqrt((double)n);
    X[0] = x[0] * sn;
    for (int k = 1; k <= (n - 1) / 2; ++k) {
        X[k] = (x[k] + I * x[n - k]) * sn;
        X[n - k] = conj(X[k]);
    }
    if (n % 2 == 0) X[n / 2] = x[n / 2] * sn;
    /* inverse DFT via conjugation */
    for (int k = 0; k < n; ++k) X[k] = conj(X[k]);
    dft(n, X, a);
    for (int j = 0; j < n; ++j) x[j] = creal(a[j]) / (double)n;
    free(heap);
    return ORC_OK;
}

/* DCT-I of x[0..n], self-inverse, stafft.f90:410-486 (pre/post-processing
 * of a length-n forfft) */
int orc_dct(int n, double *x)
{
    if (!factor_ok(n)) return ORC_E_FFT_LEN;
    const double rt2 = sqrt(2.0);
    double fpin = ORC_PI / (double)n, rtn = sqrt((double)n);
    double *wk = (double *)malloc(sizeof(double) * n);
    wk[0] = f12 * (x[0] + x[n]);
    for (int j = 1; j <= n - 1; ++j)
        wk[j] = f12 * (x[j] + x[n - j]) - sin((double)j * fpin) * (x[j] - x[n - j]);
    double rowsum = zero;
    rowsum = rowsum + f12 * x[0];
    for (int j = 1; j <= n - 1; ++j) rowsum = rowsum + x[j] * cos((double)j * fpin);
    rowsum = rowsum - f12 * x[n];
    x[n] = rt2 * rowsum / rtn;
    orc_forfft(n, wk);
    x[0] = rt2 * wk[0];
    x[1] = x[n];
    if (n % 2 == 0) {
        int nd2 = n / 2;
        for (int j = 1; j <= nd2 - 1; ++j) {
            x[2 * j] = rt2 * wk[j];
            x[2 * j + 1] = x[2 * j - 1] - rt2 * wk[n - j];
        }
        x[n] = rt2 * wk[nd2];
    } else {
        for (int j = 1; j <= (n - 1) / 2; ++j) {
            x[2 * j] = rt2 * wk[j];
            x[2 * j + 1] = x[2 * j - 1] - rt2 * wk[n - j];
        }
    }
    free(wk);
    return ORC_OK;
}

/* DST-I of x(1..n-1) (x(n) set to zero), self-inverse, stafft.f90:489-555.
 * xs points at x(1): xs[j-1] = x(j). */
int orc_dst(int n, double *xs)
{
    if (!factor_ok(n)) return ORC_E_FFT_LEN;
    const double rt2 = sqrt(2.0);
    double *x = xs - 1;   /* 1-based view */
    double fpin = ORC_PI / (double)n;
    double *wk = (double *)malloc(sizeof(double) * n);
    wk[0] = zero;
    for (int j = 1; j <= n - 1; ++j)
        wk[j] = f12 * (x[j] - x[n - j]) + sin((double)j * fpin) * (x[j] + x[n - j]);
    orc_forfft(n, wk);
    x[1] = wk[0] / rt2;
    if (n % 2 == 0) {
        for (int j = 1; j <= n / 2 - 1; ++j) {
            x[2 * j] = -rt2 * wk[n - j];
            x[2 * j + 1] = rt2 * wk[j] + x[2 * j - 1];
        }
    } else {
        for (int j = 1; j <= (n - 1) / 2 - 1; ++j) {
            x[2 * j] = -rt2 * wk[n - j];
            x[2 * j + 1] = rt2 * wk[j] + x[2 * j - 1];
        }
        x[n - 1] = -rt2 * wk[(n + 1) / 2];
    }
    x[n] = zero;
    free(wk);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* sta3dfft.f90:134-257: 2D real transforms of every z level 0..nz          */
/* ------------------------------------------------------------------------ */
int orc_fftxyp2s(const orc_dom *d, const double *fp, double *fs)
{
    int nx = d->nx, ny = d->ny, nz = d->nz;
    if (!factor_ok(nx) || !factor_ok(ny)) return ORC_E_FFT_LEN;
#pragma omp parallel for schedule(static) num_threads(nthreads_of(d))
    for (int iz = 0; iz <= nz; ++iz) {
        double line[1024];
        /* y transform for each x */
        for (int ix = 0; ix < nx; ++ix) {
            for (int iy = 0; iy < ny; ++iy) line[iy] = fp[HIDX(d, iz, iy, ix)];
            orc_forfft(ny, line);
            for (int ky = 0; ky < ny; ++ky) fs[SIDX(d, iz, ky, ix)] = line[ky];
        }
        /* x transform for each ky */
        for (int ky = 0; ky < ny; ++ky) {
            for (int ix = 0; ix < nx; ++ix) line[ix] = fs[SIDX(d, iz, ky, ix)];
            orc_forfft(nx, line);
            for (int kx = 0; kx < nx; ++kx) fs[SIDX(d, iz, ky, kx)] = line[kx];
        }
    }
    return ORC_OK;
}

/* writes the interior 0..nz, 0..ny-1, 0..nx-1 only; the halo is untouched */
int orc_fftxys2p(const orc_dom *d, const double *fs, double *fp)
{
    int nx = d->nx, ny = d->ny, nz = d->nz;
    if (!factor_ok(nx) || !factor_ok(ny)) return ORC_E_FFT_LEN;
    double *tmp = (double *)malloc(sizeof(double) * orc_spec_size(d));
#pragma omp parallel for schedule(static) num_threads(nthreads_of(d))
    for (int iz = 0; iz <= nz; ++iz) {
        double line[1024];
        for (int ky = 0; ky < ny; ++ky) {
            for (int kx = 0; kx < nx; ++kx) line[kx] = fs[SIDX(d, iz, ky, kx)];
            orc_revfft(nx, line);
            for (int ix = 0; ix < nx; ++ix) tmp[SIDX(d, iz, ky, ix)] = line[ix];
        }
        for (int ix = 0; ix < nx; ++ix) {
            for (int ky = 0; ky < ny; ++ky) line[ky] = tmp[SIDX(d, iz, ky, ix)];
            orc_revfft(ny, line);
            for (int iy = 0; iy < ny; ++iy) fp[HIDX(d, iz, iy, ix)] = line[iy];
        }
    }
    free(tmp);
    return ORC_OK;
}

/* diffx (sta3dfft.f90:301-334) with the single-rank reverse_x of
 * mpi_reverse.f90:332-359, 402-417: gs(kx-1) = fs(nx-kx) */
void orc_diffx(const orc_dom *d, const double *fs, double *ds)
{
    int nx = d->nx, ny = d->ny, nz = d->nz, nwx = nx / 2;
    double sc = ORC_PI / d->extent[0];   /* hrkx(k) = pi k / Lx (sta2dfft.f90:61-64) */
#pragma omp parallel for schedule(static) num_threads(nthreads_of(d))
    for (int kx = 0; kx < nx; ++kx) {
        double f;
        int src = (nx - kx) % nx;
        if (kx == 0) f = 0.0;
        else {
            int dkx = 2 * kx < 2 * (nx - kx) ? 2 * kx : 2 * (nx - kx);
            double si = (kx >= nwx + 1) ? 1.0 : -1.0;
            f = si * (sc * (double)dkx);
        }
        if (nx % 2 == 0 && kx == nwx) f = 0.0;
        for (int ky = 0; ky < ny; ++ky)
            for (int iz = 0; iz <= nz; ++iz) ds[SIDX(d, iz, ky, kx)] = f * fs[SIDX(d, iz, ky, src)];
    }
}

/* diffy (sta3dfft.f90:342-374) */
void orc_diffy(const orc_dom *d, const double *fs, double *ds)
{
    int nx = d->nx, ny = d->ny, nz = d->nz, nwy = ny / 2;
    double sc = ORC_PI / d->extent[1];
#pragma omp parallel for schedule(static) num_threads(nthreads_of(d))
    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky) {
            double f;
            int src = (ny - ky) % ny;
            if (ky == 0) f = 0.0;
            else {
                int dky = 2 * ky < 2 * (ny - ky) ? 2 * ky : 2 * (ny - ky);
                double si = (ky >= nwy + 1) ? 1.0 : -1.0;
                f = si * (sc * (double)dky);
            }
            if (ny % 2 == 0 && ky == nwy) f = 0.0;
            for (int iz = 0; iz <= nz; ++iz) ds[SIDX(d, iz, ky, kx)] = f * fs[SIDX(d, iz, src, kx)];
        }
}

/* ------------------------------------------------------------------------ */
/* inversion tables (inversion_utils.f90:79-275, 305-351)                   */
/* ------------------------------------------------------------------------ */
struct orc_inv {
    int64_t G;
    double *rkx, *rky, *rkz, *rkzi;   /* sta3dfft.f90:88-107 */
    double *k2l2, *k2l2i;             /* (ky, kx): [ky + ny*kx] */
    double *filt, *green;             /* spectral layout */
    double *phim, *phip, *dphim, *dphip, *thetam, *thetap, *dthetam, *dthetap;
    double *gamtop, *gambot;          /* 0..nz */
};

void orc_inv_destroy(orc_inv *v)
{
    if (!v) return;
    free(v->rkx); free(v->rky); free(v->rkz); free(v->rkzi);
    free(v->k2l2); free(v->k2l2i); free(v->filt); free(v->green);
    free(v->phim); free(v->phip); free(v->dphim); free(v->dphip);
    free(v->thetam); free(v->thetap); free(v->dthetam); free(v->dthetap);
    free(v->gamtop); free(v->gambot);
    free(v);
}

int orc_inv_create(const orc_dom *d, orc_inv **out)
{
    int nx = d->nx, ny = d->ny, nz = d->nz;
    if (!factor_ok(nx) || !factor_ok(ny) || !factor_ok(nz)) return ORC_E_FFT_LEN;
    orc_inv *v = (orc_inv *)calloc(1, sizeof *v);
    int64_t G = orc_spec_size(d);
    v->G = G;
    v->rkx = malloc(sizeof(double) * nx);
    v->rky = malloc(sizeof(double) * ny);
    v->rkz = malloc(sizeof(double) * (nz + 1));
    v->rkzi = malloc(sizeof(double) * (nz + 1));
    v->k2l2 = malloc(sizeof(double) * nx * ny);
    v->k2l2i = malloc(sizeof(double) * nx * ny);
    double **tabs[] = {&v->filt, &v->green, &v->phim, &v->phip, &v->dphim, &v->dphip,
                       &v->thetam, &v->thetap, &v->dthetam, &v->dthetap};
    for (unsigned i = 0; i < sizeof tabs / sizeof tabs[0]; ++i) *tabs[i] = calloc(G, sizeof(double));
    v->gamtop = malloc(sizeof(double) * (nz + 1));
    v->gambot = malloc(sizeof(double) * (nz + 1));

    /* wavenumbers: hrkx(k) = pi k / L (sta2dfft.f90:59-64), rkx folded
     * (sta3dfft.f90:88-102), rkz(k) = pi k / Lz (deriv1d.f90:18-23) */
    int nwx = nx / 2, nwy = ny / 2;
    double scx = ORC_PI / d->extent[0], scy = ORC_PI / d->extent[1], scz = ORC_PI / d->extent[2];
    v->rkx[0] = zero;
    for (int k = 1; k <= nwx - 1; ++k) v->rkx[k] = v->rkx[nx - k] = scx * (double)(2 * k);
    v->rkx[nwx] = scx * (double)nx;
    v->rky[0] = zero;
    for (int k = 1; k <= nwy - 1; ++k) v->rky[k] = v->rky[ny - k] = scy * (double)(2 * k);
    v->rky[nwy] = scy * (double)ny;
    v->rkz[0] = zero;
    for (int k = 1; k <= nz; ++k) v->rkz[k] = scz * (double)k;
    for (int k = 1; k <= nz - 1; ++k) v->rkzi[k] = one / v->rkz[k];

    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky) v->k2l2[ky + ny * kx] = v->rkx[kx] * v->rkx[kx] + v->rky[ky] * v->rky[ky];
    v->k2l2[0] = one;
    for (int i = 0; i < nx * ny; ++i) v->k2l2i[i] = one / v->k2l2[i];
    v->k2l2[0] = zero;
    v->k2l2i[0] = zero;

    /* Hou & Li filter (:130-150) */
    double kxmax = 0, kymax = 0, kzmax = 0;
    for (int k = 0; k < nx; ++k) kxmax = fmax(kxmax, v->rkx[k]);
    for (int k = 0; k < ny; ++k) kymax = fmax(kymax, v->rky[k]);
    for (int k = 0; k <= nz; ++k) kzmax = fmax(kzmax, v->rkz[k]);
    double kxmaxi = one / kxmax, kymaxi = one / kymax, kzmaxi = one / kzmax;
    double *skx = malloc(sizeof(double) * nx), *sky = malloc(sizeof(double) * ny), *skz = malloc(sizeof(double) * (nz + 1));
    for (int k = 0; k < nx; ++k) skx[k] = -36.0 * pow(kxmaxi * v->rkx[k], 36);
    for (int k = 0; k < ny; ++k) sky[k] = -36.0 * pow(kymaxi * v->rky[k], 36);
    for (int k = 0; k <= nz; ++k) skz[k] = -36.0 * pow(kzmaxi * v->rkz[k], 36);
    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky) {
            double f0 = exp(skx[kx] + sky[ky]);
            v->filt[SIDX(d, 0, ky, kx)] = f0;
            v->filt[SIDX(d, nz, ky, kx)] = f0;
            for (int kz = 1; kz <= nz - 1; ++kz) v->filt[SIDX(d, kz, ky, kx)] = f0 * exp(skz[kz]);
        }
    for (int kz = 0; kz <= nz; ++kz) v->filt[SIDX(d, kz, 0, 0)] = one;
    free(skx); free(sky); free(skz);

    /* Green function (:169-173) */
    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky) {
            double kl2 = v->k2l2[ky + ny * kx];
            for (int kz = 1; kz <= nz; ++kz) v->green[SIDX(d, kz, ky, kx)] = -one / (kl2 + v->rkz[kz] * v->rkz[kz]);
            v->green[SIDX(d, 0, ky, kx)] = -v->k2l2i[ky + ny * kx];
        }

    double *zm = malloc(sizeof(double) * (nz + 1)), *zp = malloc(sizeof(double) * (nz + 1));
    for (int iz = 0; iz <= nz; ++iz) {
        double z = d->lower[2] + d->dx[2] * (double)iz;
        zm[iz] = d->upper[2] - z;
        zp[iz] = z - d->lower[2];
    }
    /* set_hyperbolic_functions (:305-351), release build (NDEBUG: no clamping) */
    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky) {
            if (kx == 0 && ky == 0) continue;
            double kl = sqrt(v->k2l2[ky + ny * kx]);
            double fac = kl * d->extent[2];
            double ef = exp(-fac);
            double div = one / (one - ef * ef);
            double k2ifac = f12 * v->k2l2i[ky + ny * kx];
            double Q = div * (one + ef * ef);
            double R = div * two * ef;
            for (int iz = 0; iz <= nz; ++iz) {
                double Lm = kl * zm[iz], Lp = kl * zp[iz];
                double ep = exp(-Lp), em = exp(-Lm);
                int64_t s = SIDX(d, iz, ky, kx);
                double pm = div * (ep - ef * em), pp = div * (em - ef * ep);
                double dpm = -kl * div * (ep + ef * em), dpp = kl * div * (em + ef * ep);
                v->phim[s] = pm;
                v->phip[s] = pp;
                v->dphim[s] = dpm;
                v->dphip[s] = dpp;
                v->thetam[s] = k2ifac * (R * Lm * pp - Q * Lp * pm);
                v->thetap[s] = k2ifac * (R * Lp * pm - Q * Lm * pp);
                v->dthetam[s] = -k2ifac * ((Q * Lp - one) * dpm - R * Lm * dpp);
                v->dthetap[s] = -k2ifac * ((Q * Lm - one) * dpp - R * Lp * dpm);
            }
        }
    for (int iz = 0; iz <= nz; ++iz) {
        int64_t s = SIDX(d, iz, 0, 0);
        v->phim[s] = zm[iz] / d->extent[2];
        v->phip[s] = zp[iz] / d->extent[2];
        v->dphim[s] = -one / d->extent[2];
        v->dphip[s] = one / d->extent[2];
        v->thetam[s] = v->thetap[s] = v->dthetam[s] = v->dthetap[s] = zero;
    }
    /* gamtop / gambot (:222-238) */
    for (int iz = 0; iz <= nz; ++iz) {
        double pp = v->phip[SIDX(d, iz, 0, 0)];
        v->gamtop[iz] = f12 * d->extent[2] * (pp * pp - f13);
    }
    for (int iz = 0; iz <= nz; ++iz) v->gambot[iz] = v->gamtop[nz - iz];
    free(zm); free(zp);
    *out = v;
    return ORC_OK;
}

/* field_decompose_semi_spectral (inversion_utils.f90:372-401) */
static void decompose_ss(const orc_dom *d, const orc_inv *v, double *sf)
{
    int nx = d->nx, ny = d->ny, nz = d->nz;
#pragma omp parallel for schedule(static) num_threads(nthreads_of(d))
    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky) {
            double *c = sf + SIDX(d, 0, ky, kx);
            const double *pm = v->phim + SIDX(d, 0, ky, kx), *pp = v->phip + SIDX(d, 0, ky, kx);
            for (int iz = 1; iz <= nz - 1; ++iz) c[iz] = c[iz] - (c[0] * pm[iz] + c[nz] * pp[iz]);
            double top = c[nz];
            orc_dst(nz, c + 1);
            c[nz] = top;
        }
}

/* field_combine_semi_spectral (:426-456) */
static void combine_ss(const orc_dom *d, const orc_inv *v, double *sf)
{
    int nx = d->nx, ny = d->ny, nz = d->nz;
#pragma omp parallel for schedule(static) num_threads(nthreads_of(d))
    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky) {
            double *c = sf + SIDX(d, 0, ky, kx);
            const double *pm = v->phim + SIDX(d, 0, ky, kx), *pp = v->phip + SIDX(d, 0, ky, kx);
            double top = c[nz];
            c[nz] = zero;
            orc_dst(nz, c + 1);
            c[nz] = top;
            for (int iz = 1; iz <= nz - 1; ++iz) c[iz] = c[iz] + c[0] * pm[iz] + c[nz] * pp[iz];
        }
}

/* central_diffz_semi_spectral (:462-482) */
static void central_diffz_ss(const orc_dom *d, const double *fs, double *ds)
{
    int nx = d->nx, ny = d->ny, nz = d->nz;
    double dzi = d->dxi[2], hdzi = f12 * d->dxi[2];
#pragma omp parallel for schedule(static) num_threads(nthreads_of(d))
    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky) {
            const double *f = fs + SIDX(d, 0, ky, kx);
            double *g = ds + SIDX(d, 0, ky, kx);
            g[0] = dzi * (f[1] - f[0]);
            g[nz] = dzi * (f[nz] - f[nz - 1]);
            for (int iz = 1; iz <= nz - 1; ++iz) g[iz] = (f[iz + 1] - f[iz - 1]) * hdzi;
        }
}

/* central_diffz (:488-514) on halo fields */
static void central_diffz(const orc_dom *d, const double *f, double *df)
{
    int nx = d->nx, ny = d->ny, nz = d->nz;
    double dzi = d->dxi[2], hdzi = f12 * d->dxi[2];
#pragma omp parallel for schedule(static) num_threads(nthreads_of(d))
    for (int ix = -1; ix <= nx + 1; ++ix)
        for (int iy = -1; iy <= ny + 1; ++iy) {
            const double *a = f + HIDX(d, -1, iy, ix) + 1;   /* a[iz] = f(iz) */
            double *g = df + HIDX(d, -1, iy, ix) + 1;
            g[0] = dzi * (a[1] - a[0]);
            g[nz] = dzi * (a[nz] - a[nz - 1]);
            for (int iz = 1; iz <= nz - 1; ++iz) g[iz] = (a[iz + 1] - a[iz - 1]) * hdzi;
            g[-1] = two * g[0] - g[1];
            g[nz + 1] = two * g[nz] - g[nz - 1];
        }
}

static void extrap_lin(const orc_dom *d, double *g)   /* g(-1)=2g(0)-g(1), g(nz+1)=2g(nz)-g(nz-1) */
{
    int nz = d->nz;
    for (int ix = -1; ix <= d->nx + 1; ++ix)
        for (int iy = -1; iy <= d->ny + 1; ++iy) {
            g[HIDX(d, -1, iy, ix)] = two * g[HIDX(d, 0, iy, ix)] - g[HIDX(d, 1, iy, ix)];
            g[HIDX(d, nz + 1, iy, ix)] = two * g[HIDX(d, nz, iy, ix)] - g[HIDX(d, nz - 1, iy, ix)];
        }
}

static void extrap_anti(const orc_dom *d, double *g)   /* g(-1)=-g(1), g(nz+1)=-g(nz-1) */
{
    int nz = d->nz;
    for (int ix = -1; ix <= d->nx + 1; ++ix)
        for (int iy = -1; iy <= d->ny + 1; ++iy) {
            g[HIDX(d, -1, iy, ix)] = -g[HIDX(d, 1, iy, ix)];
            g[HIDX(d, nz + 1, iy, ix)] = -g[HIDX(d, nz - 1, iy, ix)];
        }
}

/* vel2vgrad (inversion.f90:259-317); svel = 3 spectral components */
int orc_vel2vgrad(const orc_dom *d, orc_inv *v, const double *svel, orc_fields *f)
{
    (void)v;
    const int64_t G = orc_spec_size(d), H = orc_halo_size(d);
    enum { DUDX, DUDY, DUDZ, DVDX, DVDY, DVDZ, DWDX, DWDY };
    double *ds = malloc(sizeof(double) * G);
    double *vg = f->velgradg;
    orc_diffx(d, svel, ds);          orc_fftxys2p(d, ds, vg + DUDX * H);
    orc_diffy(d, svel, ds);          orc_fftxys2p(d, ds, vg + DUDY * H);
    orc_diffx(d, svel + 2 * G, ds);  orc_fftxys2p(d, ds, vg + DWDX * H);
    extrap_lin(d, vg + DUDX * H);
    extrap_lin(d, vg + DUDY * H);
    extrap_anti(d, vg + DWDX * H);
    orc_diffy(d, svel + G, ds);      orc_fftxys2p(d, ds, vg + DVDY * H);
    orc_diffy(d, svel + 2 * G, ds);  orc_fftxys2p(d, ds, vg + DWDY * H);
    extrap_lin(d, vg + DVDY * H);
    extrap_anti(d, vg + DWDY * H);
    for (int64_t h = 0; h < H; ++h) {
        vg[DUDZ * H + h] = vg[DWDX * H + h] + f->vortg[1 * H + h];
        vg[DVDX * H + h] = vg[DUDY * H + h] + f->vortg[2 * H + h];
        vg[DVDZ * H + h] = vg[DWDY * H + h] - f->vortg[0 * H + h];
    }
    for (int c = 0; c < 8; ++c) halo_fill(d, vg + c * H);
    free(ds);
    return ORC_OK;
}

/* vor2vel (inversion.f90:25-254) */
int orc_vor2vel(const orc_dom *d, orc_inv *v, orc_fields *f)
{
    double t0 = now_s();
    int nx = d->nx, ny = d->ny, nz = d->nz;
    const int64_t G = orc_spec_size(d), H = orc_halo_size(d);
    double *as = malloc(sizeof(double) * G), *bs = malloc(sizeof(double) * G), *cs = malloc(sizeof(double) * G),
           *ds = malloc(sizeof(double) * G), *es = malloc(sizeof(double) * G);
    double *svel = malloc(sizeof(double) * 3 * G), *svor = malloc(sizeof(double) * 3 * G);
    double *ubar = malloc(sizeof(double) * (nz + 1)), *vbar = malloc(sizeof(double) * (nz + 1));
    double *sx = svor, *sy = svor + G, *sz = svor + 2 * G;

    /* decompose and filter (:40-43) */
    for (int c = 0; c < 3; ++c) {
        orc_fftxyp2s(d, f->vortg + c * H, svor + c * G);
        decompose_ss(d, v, svor + c * G);
        for (int64_t s = 0; s < G; ++s) svor[c * G + s] = v->filt[s] * svor[c * G + s];
    }
    /* solenoidality (:50-87) */
    orc_diffx(d, sy, as);
    orc_diffy(d, sx, bs);
    for (int64_t s = 0; s < G; ++s) { ds[s] = as[s] - bs[s]; cs[s] = sz[s]; }
    combine_ss(d, v, cs);
    central_diffz_ss(d, cs, es);
    decompose_ss(d, v, es);
    for (int iz = 0; iz <= nz; ++iz) { ubar[iz] = sx[SIDX(d, iz, 0, 0)]; vbar[iz] = sy[SIDX(d, iz, 0, 0)]; }
    orc_diffx(d, es, sx);
    orc_diffy(d, ds, cs);
    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky)
            for (int iz = 0; iz <= nz; ++iz) {
                int64_t s = SIDX(d, iz, ky, kx);
                sx[s] = v->k2l2i[ky + ny * kx] * (sx[s] + cs[s]);
            }
    orc_diffy(d, es, sy);
    orc_diffx(d, ds, cs);
    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky)
            for (int iz = 0; iz <= nz; ++iz) {
                int64_t s = SIDX(d, iz, ky, kx);
                sy[s] = v->k2l2i[ky + ny * kx] * (sy[s] - cs[s]);
            }
    for (int iz = 0; iz <= nz; ++iz) { sx[SIDX(d, iz, 0, 0)] = ubar[iz]; sy[SIDX(d, iz, 0, 0)] = vbar[iz]; }

    /* combine vorticity in physical space (:91-93) */
    for (int c = 0; c < 3; ++c) {
        memcpy(as, svor + c * G, sizeof(double) * G);
        combine_ss(d, v, as);
        orc_fftxys2p(d, as, f->vortg + c * H);
    }

    /* source for w (:97-115) */
    orc_diffy(d, sx, ds);
    orc_diffx(d, sy, es);
    for (int64_t s = 0; s < G; ++s) ds[s] = ds[s] - es[s];
    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky) {
            int64_t b = SIDX(d, 0, ky, kx);
            for (int iz = 1; iz <= nz - 1; ++iz)
                bs[b + iz] = ds[b] * v->thetam[b + iz] + ds[b + nz] * v->thetap[b + iz];
            for (int iz = 0; iz <= nz; ++iz)
                es[b + iz] = ds[b] * v->dthetam[b + iz] + ds[b + nz] * v->dthetap[b + iz];
            for (int iz = 1; iz <= nz - 1; ++iz) ds[b + iz] = v->green[b + iz] * ds[b + iz];
            as[b] = zero;
            for (int kz = 1; kz <= nz - 1; ++kz) as[b + kz] = v->rkz[kz] * ds[b + kz];
            as[b + nz] = zero;
        }
    /* back to semi-spectral space (:136-143) */
#pragma omp parallel for schedule(static) num_threads(nthreads_of(d))
    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky) {
            int64_t b = SIDX(d, 0, ky, kx);
            orc_dct(nz, as + b);
            orc_dst(nz, ds + b + 1);
        }
    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky) {
            int64_t b = SIDX(d, 0, ky, kx);
            ds[b] = zero;
            for (int iz = 1; iz <= nz - 1; ++iz) ds[b + iz] = ds[b + iz] + bs[b + iz];
            ds[b + nz] = zero;
            for (int iz = 0; iz <= nz; ++iz) es[b + iz] = es[b + iz] + as[b + iz];
        }
    memcpy(cs, sz, sizeof(double) * G);
    combine_ss(d, v, cs);

    /* horizontally-averaged flow (:161-176) */
    ubar[0] = vbar[0] = zero;
    for (int iz = 1; iz <= nz - 1; ++iz) {
        ubar[iz] = -v->rkzi[iz] * sy[SIDX(d, iz, 0, 0)];
        vbar[iz] = v->rkzi[iz] * sx[SIDX(d, iz, 0, 0)];
    }
    ubar[nz] = vbar[nz] = zero;
    orc_dct(nz, ubar);
    orc_dct(nz, vbar);
    for (int iz = 0; iz <= nz; ++iz) {
        ubar[iz] = ubar[iz] + sy[SIDX(d, nz, 0, 0)] * v->gamtop[iz] - sy[SIDX(d, 0, 0, 0)] * v->gambot[iz];
        vbar[iz] = vbar[iz] - sx[SIDX(d, nz, 0, 0)] * v->gamtop[iz] + sx[SIDX(d, 0, 0, 0)] * v->gambot[iz];
    }

    /* u (:180-200) */
    orc_diffx(d, es, as);
    orc_diffy(d, cs, bs);
    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky)
            for (int iz = 0; iz <= nz; ++iz) {
                int64_t s = SIDX(d, iz, ky, kx);
                as[s] = v->k2l2i[ky + ny * kx] * (as[s] + bs[s]);
            }
    for (int iz = 0; iz <= nz; ++iz) as[SIDX(d, iz, 0, 0)] = ubar[iz];
    memcpy(svel, as, sizeof(double) * G);
    orc_fftxys2p(d, as, f->velog);
    /* v (:204-224) */
    orc_diffy(d, es, as);
    orc_diffx(d, cs, bs);
    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky)
            for (int iz = 0; iz <= nz; ++iz) {
                int64_t s = SIDX(d, iz, ky, kx);
                as[s] = v->k2l2i[ky + ny * kx] * (as[s] - bs[s]);
            }
    for (int iz = 0; iz <= nz; ++iz) as[SIDX(d, iz, 0, 0)] = vbar[iz];
    memcpy(svel + G, as, sizeof(double) * G);
    orc_fftxys2p(d, as, f->velog + H);
    /* w (:228-233) */
    memcpy(svel + 2 * G, ds, sizeof(double) * G);
    orc_fftxys2p(d, ds, f->velog + 2 * H);

    orc_vel2vgrad(d, v, svel, f);

    /* z halo of velocity (:241-248) and x/y halo fill (:250) */
    extrap_lin(d, f->velog);
    extrap_lin(d, f->velog + H);
    extrap_anti(d, f->velog + 2 * H);
    for (int c = 0; c < 3; ++c) halo_fill(d, f->velog + c * H);

    free(as); free(bs); free(cs); free(ds); free(es);
    free(svel); free(svor); free(ubar); free(vbar);
    g_timers.vor2vel += now_s() - t0;
    return ORC_OK;
}

/* divergence (inversion.f90:372-404); fx, fy, fzz are halo fields, fx, fy
 * are overwritten */
static void divergence3(const orc_dom *d, double *fx, double *fy, const double *fzz, double *div)
{
    const int64_t G = orc_spec_size(d), H = orc_halo_size(d);
    double *fs = malloc(sizeof(double) * G), *ds = malloc(sizeof(double) * G);
    orc_fftxyp2s(d, fx, fs);
    orc_diffx(d, fs, ds);
    orc_fftxys2p(d, ds, fx);
    extrap_lin(d, fx);
    orc_fftxyp2s(d, fy, fs);
    orc_diffy(d, fs, ds);
    orc_fftxys2p(d, ds, fy);
    extrap_lin(d, fy);
    central_diffz(d, fzz, div);
    for (int64_t h = 0; h < H; ++h) div[h] = fx[h] + fy[h] + div[h];
    free(fs);
    free(ds);
}

/* vorticity_tendency (inversion.f90:321-366) */
int orc_vorticity_tendency(const orc_dom *d, orc_inv *v, orc_fields *f)
{
    (void)v;
    double t0 = now_s();
    const int64_t H = orc_halo_size(d);
    double *fx = malloc(sizeof(double) * H), *fy = malloc(sizeof(double) * H), *fz = malloc(sizeof(double) * H);
    const double *xi = f->vortg, *eta = f->vortg + H, *zeta = f->vortg + 2 * H;
    const double *b = f->tbuoyg;
    for (int c = 0; c < 3; ++c) {
        const double *u = f->velog + c * H;
        for (int64_t h = 0; h < H; ++h) {
            fx[h] = (xi[h] + d->f_cor[0]) * u[h];
            fy[h] = (eta[h] + d->f_cor[1]) * u[h];
            fz[h] = (zeta[h] + d->f_cor[2]) * u[h];
            if (c == 0) fy[h] = fy[h] + b[h];
            if (c == 1) fx[h] = fx[h] - b[h];
        }
        divergence3(d, fx, fy, fz, f->vtend + c * H);
    }
    int nz = d->nz;
    for (int ix = -1; ix <= d->nx + 1; ++ix)
        for (int iy = -1; iy <= d->ny + 1; ++iy) {
            if (d->zeta_zero[0]) f->vtend[2 * H + HIDX(d, 0, iy, ix)] = zero;
            if (d->zeta_zero[1]) f->vtend[2 * H + HIDX(d, nz, iy, ix)] = zero;
        }
    for (int c = 0; c < 3; ++c) { extrap_lin(d, f->vtend + c * H); halo_fill(d, f->vtend + c * H); }
    free(fx); free(fy); free(fz);
    g_timers.vtend += now_s() - t0;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* parcel_bc.f90:17-57                                                      */
/* ------------------------------------------------------------------------ */
void orc_apply_periodic_bc(const orc_dom *d, double pos[3])
{
    /* Fortran int() truncates toward zero */
    pos[0] = pos[0] - d->extent[0] * (double)(int64_t)((pos[0] - d->center[0]) * d->hli[0]);
    pos[1] = pos[1] - d->extent[1] * (double)(int64_t)((pos[1] - d->center[1]) * d->hli[1]);
}

void orc_apply_reflective_bc(const orc_dom *d, double pos[3], double B[5])
{
    if (pos[2] > d->upper[2]) {
        pos[2] = two * d->upper[2] - pos[2];
        B[2] = -B[2];
        B[4] = -B[4];
    } else if (pos[2] < d->lower[2]) {
        pos[2] = two * d->lower[2] - pos[2];
        B[2] = -B[2];
        B[4] = -B[4];
    }
}

/* ------------------------------------------------------------------------ */
/* ls_rk.f90:31-54 coefficients, :88-217 step                               */
/* ------------------------------------------------------------------------ */
static const double cas4[5] = {-567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                               -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0, 0.0};
static const double cbs4[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                               1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                               2277821191437.0 / 14882151754819.0};
static const double cas3[3] = {-5.0 / 9.0, -153.0 / 128.0, 0.0};
static const double cbs3[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

/* One step of y' = lam y (dt = 1, y0 = 1) through the same coefficient
 * tables and the same stage recursion as substep(): stage 1 assigns the
 * tendency, later stages add it to the ca-scaled running delta, every stage
 * pushes y by cb dt delta (ls_rk.f90:151-208).  Returns y after one step:
 * the scheme's stability polynomial R(lam). */
double orc_lsrk_linear(double lam, int rk_order)
{
    const double *ca, *cb;
    int ns;
    if (rk_order == 4) { ca = cas4; cb = cbs4; ns = 5; }
    else if (rk_order == 3) { ca = cas3; cb = cbs3; ns = 3; }
    else return NAN;
    double y = one, delta = zero;
    for (int s = 1; s <= ns; ++s) {
        double rhs = lam * y;
        delta = (s == 1) ? rhs : delta + rhs;
        y = y + cb[s - 1] * one * delta;
        if (s < ns) delta = ca[s - 1] * delta;
    }
    return y;
}

static int substep(const orc_dom *d, orc_inv *v, orc_parcels *p, orc_fields *f, double dt, int step,
                   int nstages, const double *ca_, const double *cb_)
{
    double ca = ca_[step - 1], cb = cb_[step - 1];
    int err = ORC_OK, rc;
    if (step == 1) {
        double t0 = now_s();
#pragma omp parallel for schedule(static) num_threads(nthreads_of(d))
        for (int64_t n = 0; n < p->n; ++n) {
            int r = orc_get_dBdt(p->B + 5 * n, p->strain + 8 * n, p->volume[n], p->delta_b + 5 * n);
            if (r) {
#pragma omp critical
                err = r;
            }
        }
        g_timers.push += now_s() - t0;
    } else {
        if ((rc = orc_vor2vel(d, v, f))) return rc;
        if ((rc = orc_vorticity_tendency(d, v, f))) return rc;
        if ((rc = orc_grid2par(d, p, f, 1))) return rc;
        double t0 = now_s();
#pragma omp parallel for schedule(static) num_threads(nthreads_of(d))
        for (int64_t n = 0; n < p->n; ++n) {
            double dB[5];
            int r = orc_get_dBdt(p->B + 5 * n, p->strain + 8 * n, p->volume[n], dB);
            if (r) {
#pragma omp critical
                err = r;
            }
            for (int l = 0; l < 5; ++l) p->delta_b[5 * n + l] = p->delta_b[5 * n + l] + dB[l];
        }
        g_timers.push += now_s() - t0;
    }
    if (err) return err;
    double t0 = now_s();
    double cbdt = cb * dt;
#pragma omp parallel for schedule(static) num_threads(nthreads_of(d))
    for (int64_t n = 0; n < p->n; ++n) {
        for (int l = 0; l < 3; ++l) {
            p->position[3 * n + l] = p->position[3 * n + l] + cbdt * p->delta_pos[3 * n + l];
            p->vorticity[3 * n + l] = p->vorticity[3 * n + l] + cbdt * p->delta_vor[3 * n + l];
        }
        for (int l = 0; l < 5; ++l) p->B[5 * n + l] = p->B[5 * n + l] + cbdt * p->delta_b[5 * n + l];
    }
    if (step == nstages) {
#pragma omp parallel for schedule(static) num_threads(nthreads_of(d))
        for (int64_t n = 0; n < p->n; ++n) orc_apply_periodic_bc(d, p->position + 3 * n);
        g_timers.push += now_s() - t0;
        return ORC_OK;
    }
#pragma omp parallel for schedule(static) num_threads(nthreads_of(d))
    for (int64_t n = 0; n < p->n; ++n) {
        for (int l = 0; l < 3; ++l) {
            p->delta_pos[3 * n + l] = ca * p->delta_pos[3 * n + l];
            p->delta_vor[3 * n + l] = ca * p->delta_vor[3 * n + l];
        }
        for (int l = 0; l < 5; ++l) p->delta_b[5 * n + l] = ca * p->delta_b[5 * n + l];
        orc_apply_periodic_bc(d, p->position + 3 * n);
    }
    g_timers.push += now_s() - t0;
    return orc_par2grid(d, p, f, -1);
}

/* get_strain (rk_utils.f90:30-66): the symmetrised velocity gradient, dw/dz
 * from incompressibility */
static void get_strain(const double g[8], double S[9])
{
    S[0] = g[0];
    S[1] = f12 * (g[1] + g[3]);
    S[2] = f12 * (g[2] + g[6]);
    S[3] = S[1];
    S[4] = g[4];
    S[5] = f12 * (g[5] + g[7]);
    S[6] = S[2];
    S[7] = S[5];
    S[8] = -(g[0] + g[4]);
}

int orc_get_time_step(const orc_dom *d, const orc_fields *f, double t, const orc_time *tm, double *dt,
                      double *gmax_out, double *bmax_out)
{
    const double eps = 2.220446049250313e-16;   /* epsilon(1d0) */
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    const int64_t H = orc_halo_size(d);
    double gmax = eps, b2max = zero;
#pragma omp parallel for schedule(static) num_threads(nthreads_of(d)) reduction(max : gmax, b2max)
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            for (int iz = 0; iz <= nz; ++iz) {
                double g[8], S[9], D[3];
                for (int c = 0; c < 8; ++c) g[c] = f->velgradg[c * H + HIDX(d, iz, iy, ix)];
                get_strain(g, S);
                orc_scherzinger_eigenvalues(S, D);
                for (int c = 0; c < 3; ++c) gmax = fmax(gmax, fabs(D[c]));
                /* central differences of tbuoyg; x and y periodic (field_halo_fill_scalar) */
                const int xp = (ix + 1) % nx, xm = (ix + nx - 1) % nx;
                const int yp = (iy + 1) % ny, ym = (iy + ny - 1) % ny;
                const double gx = f12 * d->dxi[0] * (f->tbuoyg[HIDX(d, iz, iy, xp)] - f->tbuoyg[HIDX(d, iz, iy, xm)]);
                const double gy = f12 * d->dxi[1] * (f->tbuoyg[HIDX(d, iz, yp, ix)] - f->tbuoyg[HIDX(d, iz, ym, ix)]);
                const double gz = f12 * d->dxi[2] * (f->tbuoyg[HIDX(d, iz + 1, iy, ix)] - f->tbuoyg[HIDX(d, iz - 1, iy, ix)]);
                b2max = fmax(b2max, gx * gx + gy * gy + gz * gz);
            }
    double bmax = fmax(eps, sqrt(sqrt(b2max)));
    if (gmax_out) *gmax_out = gmax;
    if (bmax_out) *bmax_out = bmax;
    double h = tm->fixed ? tm->fixed_dt : fmin(tm->alpha / gmax, tm->alpha / bmax);
    if (tm->precise_stop && (t + h > tm->limit)) h = tm->limit - t;
    *dt = h;
    return (h <= zero) ? ORC_E_DT : ORC_OK;
}

static int rk_step(const orc_dom *d, orc_inv *v, orc_parcels *p, orc_fields *f, double *t, double dt_fixed,
                   const orc_time *tm, double t_initial, int rk_order, double *dt_out)
{
    const double *ca, *cb;
    int ns;
    if (rk_order == 4) { ca = cas4; cb = cbs4; ns = 5; }
    else if (rk_order == 3) { ca = cas3; cb = cbs3; ns = 3; }
    else return ORC_E_RK_ORDER;
    if (!tm && !(dt_fixed > 0.0)) return ORC_E_DT;
    int rc;
    if ((rc = orc_par2grid(d, p, f, (*t > t_initial) ? 1 : 0))) return rc;
    if ((rc = orc_vor2vel(d, v, f))) return rc;
    if ((rc = orc_vorticity_tendency(d, v, f))) return rc;
    double dt = dt_fixed;
    if (tm && (rc = orc_get_time_step(d, f, *t, tm, &dt, NULL, NULL))) return rc;
    if (dt_out) *dt_out = dt;
    if ((rc = orc_grid2par(d, p, f, -1))) return rc;
    for (int s = 1; s <= ns; ++s)
        if ((rc = substep(d, v, p, f, dt, s, ns, ca, cb))) return rc;
    double t0 = now_s();
#pragma omp parallel for schedule(static) num_threads(nthreads_of(d))
    for (int64_t n = 0; n < p->n; ++n) orc_apply_reflective_bc(d, p->position + 3 * n, p->B + 5 * n);
    g_timers.push += now_s() - t0;
    *t = *t + dt;
    return ORC_OK;
}

int orc_ls_rk_step(const orc_dom *d, orc_inv *v, orc_parcels *p, orc_fields *f, double *t, double dt,
                   double t_initial, int rk_order)
{
    return rk_step(d, v, p, f, t, dt, NULL, t_initial, rk_order, NULL);
}

int orc_ls_rk_step_adaptive(const orc_dom *d, orc_inv *v, orc_parcels *p, orc_fields *f, double *t,
                            const orc_time *tm, double t_initial, int rk_order, double *dt)
{
    return rk_step(d, v, p, f, t, 0.0, tm, t_initial, rk_order, dt);
}

/* ------------------------------------------------------------------------ */
/* diagnostics: parcel_diagnostics.f90:51-159, field_diagnostics.f90:40-154  */
/* ------------------------------------------------------------------------ */
void orc_parcel_diagnostics(const orc_dom *d, const orc_parcels *p, const double counters[3], double stats[18])
{
    double ke = zero, ens = zero, lam1 = zero, lam2 = zero, v1 = zero, v2 = zero, rx = zero, ry = zero, rz = zero;
    double nsmall = zero, bmin = zero, bmax = zero;
    const double vmin = d->vmin > 0.0 ? d->vmin : d->vcell / 20.0;
    for (int64_t n = 0; n < p->n; ++n) {
        const double *vel = p->delta_pos + 3 * n, *vor = p->vorticity + 3 * n;
        const double vol = p->volume[n], b = p->buoyancy[n];
        ke += (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]) * vol;
        ens += (vor[0] * vor[0] + vor[1] * vor[1] + vor[2] * vor[2]) * vol;
        double A[9], D[3], b33;
        const double *B = p->B + 5 * n;
        if (orc_get_B33(B, vol, &b33) != ORC_OK) b33 = zero;
        A[0] = B[B11]; A[1] = B[B12]; A[2] = B[B13];
        A[3] = B[B12]; A[4] = B[B22]; A[5] = B[B23];
        A[6] = B[B13]; A[7] = B[B23]; A[8] = b33;
        orc_scherzinger_eigenvalues(A, D);
        const double lam = sqrt(D[0] / D[2]);   /* get_aspect_ratio (parcel_ellipsoid.f90:423-429) */
        lam1 += lam;
        lam2 += lam * lam;
        v1 += vol;
        v2 += vol * vol;
        if (vol <= vmin) nsmall += one;
        bmax = fmax(bmax, b);
        bmin = fmin(bmin, b);
        rx += vol * vor[0] * vor[0];
        ry += vol * vor[1] * vor[1];
        rz += vol * vor[2] * vor[2];
    }
    const double vdomaini = one / (d->extent[0] * d->extent[1] * d->extent[2]);
    const double ntot = (double)p->n, ntoti = one / ntot;
    for (int i = 0; i < 18; ++i) stats[i] = zero;
    stats[1] = f12 * ke * vdomaini;
    stats[2] = nsmall;
    stats[3] = lam1 * ntoti;
    stats[5] = sqrt(fabs(lam2 * ntoti - stats[3] * stats[3]));
    stats[7] = v1;
    stats[8] = sqrt(rx / v1);
    stats[9] = sqrt(ry / v1);
    stats[10] = sqrt(rz / v1);
    stats[11] = ntot;
    stats[12] = f12 * ens * vdomaini;
    stats[13] = counters[0];
    stats[14] = counters[1];
    stats[15] = counters[2];
    stats[4] = v1 * ntoti;
    stats[6] = sqrt(fabs(v2 * ntoti - stats[4] * stats[4]));
    stats[16] = bmin;
    stats[17] = bmax;
}

void orc_field_diagnostics(const orc_dom *d, const orc_fields *f, double stats[11])
{
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    const int64_t H = orc_halo_size(d);
    const double vcell = d->vcell, vcelli = one / vcell;
    double rms = zero, aerr = zero, np = zero, nsp = zero, ke = zero, en = zero;
    double npmax = -1e300, npmin = 1e300, bmax = -1e300, bmin = 1e300;
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            for (int iz = 0; iz <= nz; ++iz) {
                const int64_t o = HIDX(d, iz, iy, ix);
                const double v = f->volg[o];
                rms += (v - vcell) * (v - vcell);
                aerr = fmax(aerr, fabs(v - vcell));
                bmax = fmax(bmax, f->tbuoyg[o]);
                bmin = fmin(bmin, f->tbuoyg[o]);
                if (iz < nz) {
                    np += f->nparg[o];
                    nsp += f->nsparg[o];
                    npmax = fmax(npmax, (double)f->nparg[o]);
                    npmin = fmin(npmin, (double)f->nparg[o]);
                }
                /* half weights for the boundary planes */
                const double w = (iz == 0 || iz == nz) ? f14 : f12;
                double u2 = zero, w2 = zero;
                for (int c = 0; c < 3; ++c) {
                    u2 += f->velog[c * H + o] * f->velog[c * H + o];
                    w2 += f->vortg[c * H + o] * f->vortg[c * H + o];
                }
                ke += w * v * u2;
                en += w * v * w2;
            }
    const double vdomaini = one / (d->extent[0] * d->extent[1] * d->extent[2]);
    const double ngridi = one / ((double)nx * ny * (nz + 1)), ncelli = one / ((double)nx * ny * nz);
    for (int i = 0; i < 11; ++i) stats[i] = zero;
    stats[0] = sqrt(rms * ngridi) * vcelli;
    stats[1] = np * ncelli;
    stats[2] = nsp * ncelli;
    stats[3] = ke * vdomaini;
    stats[5] = en * vdomaini;
    stats[6] = aerr * vcelli;
    stats[7] = npmax;
    stats[8] = bmax;
    stats[9] = npmin;
    stats[10] = bmin;
}

/* ------------------------------------------------------------------------ */
/* cell counting sort (parcel_nearest.f90:196-257, 411-424, 531-541)        */
/* halo box: 1 lower + 2 upper cells in x and y (mpi_layout.f90:118-123)    */
/* ------------------------------------------------------------------------ */
/* halo_ncell = halo_size(1) * halo_size(2) * (halo_size(3) - 1)
 * (mpi_layout.f90:127, 134) = (nx+3)(ny+3)(nz+2) */
int64_t orc_nearest_ncell(const orc_dom *d)
{
    return (int64_t)(d->nx + 3) * (d->ny + 3) * (d->nz + 2);
}

int orc_nearest_cell_sort(const orc_dom *d, int64_t n, const double *position, int *loca, int *nppc, int *kc1,
                          int *node)
{
    int hx = d->nx + 3, hy = d->ny + 3, nz = d->nz;
    int64_t ncell = orc_nearest_ncell(d);
    double hlx = d->lower[0] - d->dx[0], hly = d->lower[1] - d->dx[1];
    for (int64_t c = 0; c < ncell; ++c) nppc[c] = 0;
    for (int64_t i = 0; i < n; ++i) {
        double pos[3] = {position[3 * i], position[3 * i + 1], position[3 * i + 2]};
        /* handle_periodic_edge_parcels (:168-188): single rank, box hlo = -1,
         * hhi = n + 1 */
        int ix = (int)floor((pos[0] - d->lower[0]) * d->dxi[0]);
        int iy = (int)floor((pos[1] - d->lower[1]) * d->dxi[1]);
        if (ix < -1) pos[0] += d->extent[0];
        if (ix > d->nx) pos[0] -= d->extent[0];
        if (iy < -1) pos[1] += d->extent[1];
        if (iy > d->ny) pos[1] -= d->extent[1];
        int jx = (int)(d->dxi[0] * (pos[0] - hlx));
        int jy = (int)(d->dxi[1] * (pos[1] - hly));
        int jz = (int)(d->dxi[2] * (pos[2] - d->lower[2]));
        if (jz > nz - 1) jz = nz - 1;
        int64_t ijk = 1 + jx + (int64_t)hx * jy + (int64_t)hx * hy * jz;
        if (ijk < 1 || ijk > ncell) return ORC_E_OUT_OF_GRID;
        nppc[ijk - 1] += 1;
        loca[i] = (int)ijk;
    }
    kc1[0] = 1;
    for (int64_t c = 1; c < ncell; ++c) kc1[c] = kc1[c - 1] + nppc[c - 1];
    int *kc2 = (int *)malloc(sizeof(int) * ncell);
    for (int64_t c = 0; c < ncell; ++c) kc2[c] = kc1[c] - 1;
    for (int64_t i = 0; i < n; ++i) {
        int64_t c = loca[i] - 1;
        int k = kc2[c] + 1;
        node[k - 1] = (int)i;
        kc2[c] = k;
    }
    free(kc2);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* parcel_split.f90:37-197, single rank in serial order                     */
/* ------------------------------------------------------------------------ */
static int full_matrix(const orc_parcels *p, int64_t n, double A[9])
{
    const double *B = p->B + 5 * n;
    double b33;
    int rc = orc_get_B33(B, p->volume[n], &b33);
    if (rc) return rc;
    A[0] = B[0]; A[1] = B[1]; A[2] = B[2];
    A[3] = B[1]; A[4] = B[3]; A[5] = B[4];
    A[6] = B[2]; A[7] = B[4]; A[8] = b33;
    return ORC_OK;
}

double orc_default_amax(const orc_dom *d)
{
    /* parameters.f90:150: amax = (f34 * fpi) ** f13 * minval(dx) */
    double mdx = d->dx[0];
    if (d->dx[1] < mdx) mdx = d->dx[1];
    if (d->dx[2] < mdx) mdx = d->dx[2];
    return pow(f34 * (one / ORC_PI), f13) * mdx;
}

int orc_parcel_split(const orc_dom *d, orc_parcels *p, int64_t cap, double lambda_max, double amax,
                     int64_t *n_split)
{
    const double dh = f12 * sqrt(three / 5.0);
    if (amax <= zero) amax = orc_default_amax(d);
    const int64_t n0 = p->n;
    /* which parcels split (:56-76): lam = sqrt(D1/D3) (get_aspect_ratio :423-429) */
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n0 > 0 ? n0 : 1));
    int64_t ni = 0;
    for (int64_t n = 0; n < n0; ++n) {
        double A[9], D[3];
        int rc = full_matrix(p, n, A);
        if (rc) {
            free(idx);
            return rc;
        }
        orc_scherzinger_eigenvalues(A, D);
        const double lam = sqrt(D[0] / D[2]);
        if (lam < lambda_max && D[0] < amax * amax) continue;
        idx[ni++] = n;
    }
    if (n0 + ni > cap) {   /* the reference grows its container (:87-96) */
        free(idx);
        return ORC_E_ALLOC;
    }
    for (int64_t i = 0; i < ni; ++i) {
        const int64_t n = idx[i];
        double A[9], D[3], V[9];
        full_matrix(p, n, A);
        orc_scherzinger_diagonalise(A, D, V);
        double *B = p->B + 5 * n;
        const double vol = p->volume[n];
        /* B - 3/4 D1 v1 v1^T (:118-122) */
        B[0] = B[0] - f34 * D[0] * (V[0] * V[0]);
        B[1] = B[1] - f34 * D[0] * V[0] * V[1];
        B[2] = B[2] - f34 * D[0] * V[0] * V[2];
        B[3] = B[3] - f34 * D[0] * (V[1] * V[1]);
        B[4] = B[4] - f34 * D[0] * V[1] * V[2];
        p->volume[n] = f12 * vol;
        const int64_t c = p->n++;
        memcpy(p->B + 5 * c, B, 5 * sizeof(double));
        memcpy(p->vorticity + 3 * c, p->vorticity + 3 * n, 3 * sizeof(double));
        p->volume[c] = p->volume[n];
        p->buoyancy[c] = p->buoyancy[n];
        p->humidity[c] = p->humidity[n];
        for (int k = 0; k < 3; ++k) {
            p->delta_pos[3 * c + k] = zero;
            p->delta_vor[3 * c + k] = zero;
        }
        for (int k = 0; k < 5; ++k) p->delta_b[5 * c + k] = zero;
        /* offsets along the major axis (:147-149) */
        const double sd = sqrt(D[0]);
        for (int k = 0; k < 3; ++k) {
            const double v = V[k] * dh * sd;
            p->position[3 * c + k] = p->position[3 * n + k] - v;
            p->position[3 * n + k] = p->position[3 * n + k] + v;
        }
        orc_apply_reflective_bc(d, p->position + 3 * c, p->B + 5 * c);
        orc_apply_reflective_bc(d, p->position + 3 * n, p->B + 5 * n);
        /* parcel_communicate -> locate_parcels applies the periodic BC to
         * both (parcel_mpi.f90:279-300) */
        orc_apply_periodic_bc(d, p->position + 3 * c);
        orc_apply_periodic_bc(d, p->position + 3 * n);
    }
    free(idx);
    if (n_split) *n_split = ni;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* parcel_merge.f90 + parcel_nearest.f90:375-627, 937-1377 on one rank       */
/*                                                                          */
/* On one rank the eight MPI neighbours are the rank itself: every small    */
/* parcel whose nearest grid index is 0 or n in x or y is sent to itself as */
/* a periodic image (apply_periodic_shift, :564-598), appended after the    */
/* parcels in the receive order of the directions (mpi_tags.f90) and binned */
/* into the halo cells.  Images search like any small parcel; their result  */
/* returns to the source, which keeps it when strictly closer               */
/* (find_closest_parcel_globally, :794-932).                                */
/* ------------------------------------------------------------------------ */
static double fnint(double v) { return (double)lround(v); }   /* Fortran nint: half away from zero */

int orc_parcel_merge(const orc_dom *d, orc_parcels *p, int64_t *n_merge_out)
{
    return orc_parcel_merge_stats(d, p, n_merge_out, NULL);
}

/* parcel_merge with collect_merge_stats (parcel_merge.f90:365-393): stats[0]
 * += the pairs whose target is big (volume > vmin, before the merge),
 * stats[j] += the targets with min(7, pairs into it) = j, j = 1..7
 * (n_big_close and n_way_parcel_mergers) */
int orc_parcel_merge_stats(const orc_dom *d, orc_parcels *p, int64_t *n_merge_out, int64_t *stats)
{
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    const int hx = nx + 3, hy = ny + 3;
    const double vmin = d->vmin > zero ? d->vmin : d->vcell / 20.0;
    const double hlx = d->lower[0] - d->dx[0], hly = d->lower[1] - d->dx[1];
    const int64_t n = p->n;
    if (n_merge_out) *n_merge_out = 0;

    /* handle_periodic_edge_parcels (:172-188) and the local small list */
    int64_t ns_loc = 0;
    for (int64_t i = 0; i < n; ++i) {
        double *x = p->position + 3 * i;
        int g0 = (int)floor((x[0] - d->lower[0]) * d->dxi[0]);
        int g1 = (int)floor((x[1] - d->lower[1]) * d->dxi[1]);
        if (g0 < -1) x[0] += d->extent[0];
        if (g0 > nx) x[0] -= d->extent[0];
        if (g1 < -1) x[1] += d->extent[1];
        if (g1 > ny) x[1] -= d->extent[1];
        if (p->volume[i] < vmin) ++ns_loc;
    }
    if (ns_loc == 0) return ORC_OK;
    int64_t *sml = (int64_t *)malloc(sizeof(int64_t) * ns_loc);
    {
        int64_t j = 0;
        for (int64_t i = 0; i < n; ++i)
            if (p->volume[i] < vmin) sml[j++] = i;
    }
    /* images in receive order n = NORTH..SOUTHEAST; list and shift of each:
     * list (boundary test on the nearest grid index) and shift (dx, dy) */
    static const int lst_x[8] = {-1, -1, 1, 0, 1, 0, 1, 0};     /* ix == nx (1), == 0 (0), any (-1) */
    static const int lst_y[8] = {0, 1, -1, -1, 0, 0, 1, 1};     /* iy == 0 (0), == ny (1), any (-1) */
    static const int sh_x[8] = {0, 0, -1, 1, -1, 1, -1, 1};
    static const int sh_y[8] = {1, -1, 0, 0, 1, 1, -1, -1};
    int64_t cap_img = 8 * ns_loc, nimg = 0;
    double *ipos = (double *)malloc(sizeof(double) * 3 * cap_img);
    int64_t *ipid = (int64_t *)malloc(sizeof(int64_t) * cap_img);
    int64_t *imid = (int64_t *)malloc(sizeof(int64_t) * cap_img);
    int img_dir_start[9];
    for (int dir = 0; dir < 8; ++dir) {
        img_dir_start[dir] = (int)nimg;
        for (int64_t m = 0; m < ns_loc; ++m) {
            const double *x = p->position + 3 * sml[m];
            const int ix = (int)fnint(d->dxi[0] * (x[0] - d->lower[0]));
            const int iy = (int)fnint(d->dxi[1] * (x[1] - d->lower[1]));
            const int in_x = (lst_x[dir] < 0) || (lst_x[dir] == 0 && ix == 0) || (lst_x[dir] == 1 && ix == nx);
            const int in_y = (lst_y[dir] < 0) || (lst_y[dir] == 0 && iy == 0) || (lst_y[dir] == 1 && iy == ny);
            if (!(in_x && in_y)) continue;
            /* apply_periodic_shift: edges tested on the unshifted position */
            const int ex_ = (ix <= 0) || (ix >= nx), ey_ = (iy <= 0) || (iy >= ny);
            double *q = ipos + 3 * nimg;
            q[0] = x[0] + ((sh_x[dir] != 0 && ex_) ? sh_x[dir] * d->extent[0] : zero);
            q[1] = x[1] + ((sh_y[dir] != 0 && ey_) ? sh_y[dir] * d->extent[1] : zero);
            q[2] = x[2];
            ipid[nimg] = sml[m];
            imid[nimg] = m;
            ++nimg;
        }
    }
    img_dir_start[8] = (int)nimg;
    /* bin parcels then images into the halo cells (:196-257) */
    const int64_t ntot = n + nimg;
    const int64_t ncell = orc_nearest_ncell(d);
    int *nppc = (int *)calloc((size_t)ncell, sizeof(int));
    int *kc1 = (int *)malloc(sizeof(int) * ncell), *kc2 = (int *)malloc(sizeof(int) * ncell);
    int64_t *loca = (int64_t *)malloc(sizeof(int64_t) * ntot);
    int64_t *node = (int64_t *)malloc(sizeof(int64_t) * ntot);
    int rc = ORC_OK;
#define PPOS(i) ((i) < n ? p->position + 3 * (i) : ipos + 3 * ((i) - n))
    for (int64_t i = 0; i < ntot && rc == ORC_OK; ++i) {
        double *x = PPOS(i);
        if (i >= n) {   /* handle_periodic_edge_parcels for the images */
            int g0 = (int)floor((x[0] - d->lower[0]) * d->dxi[0]);
            int g1 = (int)floor((x[1] - d->lower[1]) * d->dxi[1]);
            if (g0 < -1) x[0] += d->extent[0];
            if (g0 > nx) x[0] -= d->extent[0];
            if (g1 < -1) x[1] += d->extent[1];
            if (g1 > ny) x[1] -= d->extent[1];
        }
        int jx = (int)(d->dxi[0] * (x[0] - hlx));
        int jy = (int)(d->dxi[1] * (x[1] - hly));
        int jz = (int)(d->dxi[2] * (x[2] - d->lower[2]));
        if (jz > nz - 1) jz = nz - 1;
        int64_t ijk = (int64_t)jx + (int64_t)hx * jy + (int64_t)hx * hy * jz;
        if (ijk < 0 || ijk >= ncell) rc = ORC_E_OUT_OF_GRID;
        else {
            nppc[ijk] += 1;
            loca[i] = ijk;
        }
    }
    kc1[0] = 0;
    for (int64_t c = 1; c < ncell; ++c) kc1[c] = kc1[c - 1] + nppc[c - 1];
    for (int64_t c = 0; c < ncell; ++c) kc2[c] = kc1[c];
    for (int64_t i = 0; i < ntot && rc == ORC_OK; ++i) node[kc2[loca[i]]++] = i;
    /* closest other parcel of every small parcel, local then images (:608-662) */
    const int64_t nsm = ns_loc + nimg;
    int64_t *iclo = (int64_t *)malloc(sizeof(int64_t) * nsm);
    double *dclo = (double *)malloc(sizeof(double) * nsm);
    const double dsq0 = d->extent[0] * d->extent[0] + d->extent[1] * d->extent[1] + d->extent[2] * d->extent[2];
    for (int64_t m = 0; m < nsm && rc == ORC_OK; ++m) {
        const int64_t is = (m < ns_loc) ? sml[m] : n + (m - ns_loc);
        const double *xs = PPOS(is);
        const int ix0 = (int)fnint(d->dxi[0] * (xs[0] - hlx));
        const int iy0 = (int)fnint(d->dxi[1] * (xs[1] - hly));
        const int iz0 = (int)fnint(d->dxi[2] * (xs[2] - d->lower[2]));
        double dmin = dsq0;
        int64_t ic = -1;
        for (int iz = (iz0 - 1 > 0 ? iz0 - 1 : 0); iz <= (iz0 < nz - 1 ? iz0 : nz - 1); ++iz)
            for (int iy = (iy0 - 1 > 0 ? iy0 - 1 : 0); iy <= (iy0 < ny + 1 ? iy0 : ny + 1); ++iy)
                for (int ix = (ix0 - 1 > 0 ? ix0 - 1 : 0); ix <= (ix0 < nx + 1 ? ix0 : nx + 1); ++ix) {
                    const int64_t ijk = (int64_t)ix + (int64_t)hx * iy + (int64_t)hx * hy * iz;
                    for (int k = kc1[ijk]; k < kc2[ijk]; ++k) {
                        const int64_t j = node[k];
                        if (j == is) continue;
                        const double *xj = PPOS(j);
                        const double delz = xj[2] - xs[2];
                        if (delz * delz < dmin) {
                            const double delx = xj[0] - xs[0], dely = xj[1] - xs[1];
                            const double dsq = delx * delx + dely * dely + delz * delz;
                            if (dsq < dmin) {
                                dmin = dsq;
                                ic = j;
                            }
                        }
                    }
                }
        /* update_remote_indices (:746-788) */
        if (is >= n && ic >= n) dmin = HUGE_VAL;
        else if (ic >= n) ic = ipid[ic - n];
        iclo[m] = ic;
        dclo[m] = dmin;
    }
    /* find_closest_parcel_globally: the image results of the directions in
     * the order SOUTH, NORTH, EAST, WEST, SOUTHEAST, SOUTHWEST, NORTHEAST,
     * NORTHWEST (receive tag n' = SEND_NEIGHBOUR_TAG of the image direction) */
    static const int back_order[8] = {1, 0, 3, 2, 7, 6, 5, 4};
    for (int o = 0; o < 8 && rc == ORC_OK; ++o) {
        const int dir = back_order[o];
        for (int k = img_dir_start[dir]; k < img_dir_start[dir + 1]; ++k) {
            const int64_t m = ns_loc + k, ms = imid[k];
            if (dclo[ms] > dclo[m]) {
                dclo[ms] = dclo[m];
                iclo[ms] = iclo[m];
            }
        }
    }
    for (int64_t m = 0; m < ns_loc && rc == ORC_OK; ++m)
        if (iclo[m] < 0) rc = ORC_E_MERGE;   /* 'Merge error: no near enough neighbour found.' */
    /* resolve_tree (:937-1377), one rank */
    unsigned char *leaf = NULL, *avail = NULL, *merged = NULL, *dom_ = NULL, *iso = NULL;
    int64_t nm = 0;
    int64_t *isma = NULL, *ic_m = NULL;
    if (rc == ORC_OK) {
        leaf = (unsigned char *)calloc((size_t)n, 1);
        avail = (unsigned char *)calloc((size_t)n, 1);
        merged = (unsigned char *)calloc((size_t)n, 1);
        dom_ = (unsigned char *)calloc((size_t)ns_loc, 1);
        iso = (unsigned char *)calloc((size_t)ns_loc, 1);
        int cont = 1;
        while (cont) {
            cont = 0;
            for (int64_t m = 0; m < ns_loc; ++m)
                if (!merged[sml[m]]) {
                    leaf[sml[m]] = 1;
                    avail[iclo[m]] = 1;
                }
            for (int64_t m = 0; m < ns_loc; ++m)
                if (!merged[sml[m]]) leaf[iclo[m]] = 0;
            for (int64_t m = 0; m < ns_loc; ++m)
                if (!merged[sml[m]] && !leaf[sml[m]]) avail[iclo[m]] = 0;
            for (int64_t m = 0; m < ns_loc; ++m) {
                const int64_t is = sml[m], ic = iclo[m];
                if (!merged[is] && leaf[is] && avail[ic]) {
                    cont = 1;
                    merged[is] = 1;
                    merged[ic] = 1;
                }
            }
        }
        for (int64_t m = 0; m < ns_loc; ++m)
            if (!merged[sml[m]] && leaf[sml[m]]) avail[iclo[m]] = 1;
        for (int64_t m = 0; m < ns_loc; ++m) {
            const int64_t is = sml[m], ic = iclo[m];
            if (merged[is] && leaf[is]) {
                dom_[m] = 1;
            } else if (!merged[is]) {
                if (leaf[is]) dom_[m] = 1;
                else if (!avail[is]) {
                    if (avail[ic]) dom_[m] = 1;
                    else {
                        iso[m] = 1;
                        avail[is] = 1;   /* the lower rank (here: the only one) makes its parcel available */
                    }
                }
            }
        }
        for (int64_t m = 0; m < ns_loc; ++m)
            if (!dom_[m] && iso[m] && avail[iclo[m]]) dom_[m] = 1;
        isma = (int64_t *)malloc(sizeof(int64_t) * ns_loc);
        ic_m = (int64_t *)malloc(sizeof(int64_t) * ns_loc);
        for (int64_t m = 0; m < ns_loc; ++m)
            if (dom_[m]) {
                isma[nm] = sml[m];
                ic_m[nm] = iclo[m];
                ++nm;
            }
    }
    /* collect_merge_stats (parcel_merge.f90:365-393), before the volumes change */
    if (rc == ORC_OK && nm > 0 && stats) {
        int64_t *loca = (int64_t *)calloc((size_t)n, sizeof(int64_t));
        for (int64_t m = 0; m < nm; ++m) {
            const int64_t ic = ic_m[m];
            stats[0] += (p->volume[ic] > vmin) ? 1 : 0;
            loca[ic] += 1;
        }
        for (int64_t m = 0; m < nm; ++m) {
            const int64_t ic = ic_m[m];
            const int64_t cnt = loca[ic];
            if (cnt > 0) {
                loca[ic] = -1;
                stats[cnt < 7 ? cnt : 7] += 1;
            }
        }
        free(loca);
    }
    /* geometric_merge / do_group_merge (parcel_merge.f90:129-317) */
    if (rc == ORC_OK && nm > 0) {
        int64_t *grp = (int64_t *)malloc(sizeof(int64_t) * n);
        for (int64_t i = 0; i < n; ++i) grp[i] = -1;
        double *vm = (double *)malloc(sizeof(double) * nm), *x0 = (double *)malloc(sizeof(double) * nm);
        double *y0 = (double *)malloc(sizeof(double) * nm), *posm = (double *)malloc(sizeof(double) * 3 * nm);
        double *bm = (double *)malloc(sizeof(double) * nm), *hm = (double *)malloc(sizeof(double) * nm);
        double *wm = (double *)malloc(sizeof(double) * 3 * nm), *Bm = (double *)malloc(sizeof(double) * 6 * nm);
        int64_t ng = 0;
        for (int64_t m = 0; m < nm; ++m) {
            const int64_t ic = ic_m[m], is = isma[m];
            orc_apply_periodic_bc(d, p->position + 3 * is);
            if (grp[ic] < 0) {
                const int64_t l = ng++;
                grp[ic] = l;
                const double v = p->volume[ic];
                vm[l] = v;
                x0[l] = p->position[3 * ic + 0];
                y0[l] = p->position[3 * ic + 1];
                posm[3 * l + 0] = zero;
                posm[3 * l + 1] = zero;
                posm[3 * l + 2] = v * p->position[3 * ic + 2];
                bm[l] = v * p->buoyancy[ic];
                hm[l] = v * p->humidity[ic];
                for (int k = 0; k < 3; ++k) wm[3 * l + k] = v * p->vorticity[3 * ic + k];
                for (int k = 0; k < 6; ++k) Bm[6 * l + k] = zero;
            }
            const int64_t l = grp[ic];
            const double vs = p->volume[is];
            vm[l] = vm[l] + vs;
            double delx = p->position[3 * is + 0] - x0[l];
            delx = delx - d->extent[0] * fnint(delx * (one / d->extent[0]));
            double dely = p->position[3 * is + 1] - y0[l];
            dely = dely - d->extent[1] * fnint(dely * (one / d->extent[1]));
            posm[3 * l + 0] = posm[3 * l + 0] + vs * delx;
            posm[3 * l + 1] = posm[3 * l + 1] + vs * dely;
            posm[3 * l + 2] = posm[3 * l + 2] + vs * p->position[3 * is + 2];
            bm[l] = bm[l] + vs * p->buoyancy[is];
            hm[l] = hm[l] + vs * p->humidity[is];
            for (int k = 0; k < 3; ++k) wm[3 * l + k] = wm[3 * l + k] + vs * p->vorticity[3 * is + k];
        }
        for (int64_t l = 0; l < ng; ++l) {
            const double vi = one / vm[l];
            posm[3 * l + 0] = vi * posm[3 * l + 0];
            posm[3 * l + 1] = vi * posm[3 * l + 1];
            posm[3 * l + 0] = x0[l] + posm[3 * l + 0];
            posm[3 * l + 1] = y0[l] + posm[3 * l + 1];
            posm[3 * l + 2] = vi * posm[3 * l + 2];
            orc_apply_periodic_bc(d, posm + 3 * l);
            bm[l] = vi * bm[l];
            hm[l] = vi * hm[l];
            for (int k = 0; k < 3; ++k) wm[3 * l + k] = vi * wm[3 * l + k];
        }
        for (int64_t i = 0; i < n; ++i) grp[i] = -1;
        int64_t l2 = 0;
        for (int64_t m = 0; m < nm && rc == ORC_OK; ++m) {
            const int64_t ic = ic_m[m], is = isma[m];
            double b33;
            if (grp[ic] < 0) {
                const int64_t l = l2++;
                grp[ic] = l;
                const double vi = one / vm[l];
                rc = orc_get_B33(p->B + 5 * ic, p->volume[ic], &b33);
                if (rc) break;
                double delx = p->position[3 * ic + 0] - posm[3 * l + 0];
                delx = delx - d->extent[0] * fnint(delx * (one / d->extent[0]));
                double dely = p->position[3 * ic + 1] - posm[3 * l + 1];
                dely = dely - d->extent[1] * fnint(dely * (one / d->extent[1]));
                const double delz = p->position[3 * ic + 2] - posm[3 * l + 2];
                const double mu = p->volume[ic] * vi;
                const double *B = p->B + 5 * ic;
                Bm[6 * l + 0] = mu * (5.0 * (delx * delx) + B[0]);
                Bm[6 * l + 1] = mu * (5.0 * delx * dely + B[1]);
                Bm[6 * l + 2] = mu * (5.0 * delx * delz + B[2]);
                Bm[6 * l + 3] = mu * (5.0 * (dely * dely) + B[3]);
                Bm[6 * l + 4] = mu * (5.0 * dely * delz + B[4]);
                Bm[6 * l + 5] = mu * (5.0 * (delz * delz) + b33);
                p->volume[ic] = vm[l];
                for (int k = 0; k < 3; ++k) p->position[3 * ic + k] = posm[3 * l + k];
                p->buoyancy[ic] = bm[l];
                p->humidity[ic] = hm[l];
                for (int k = 0; k < 3; ++k) p->vorticity[3 * ic + k] = wm[3 * l + k];
            }
            const int64_t l = grp[ic];
            const double vi = one / vm[l];
            double delx = p->position[3 * is + 0] - posm[3 * l + 0];
            delx = delx - d->extent[0] * fnint(delx * (one / d->extent[0]));
            double dely = p->position[3 * is + 1] - posm[3 * l + 1];
            dely = dely - d->extent[1] * fnint(dely * (one / d->extent[1]));
            const double delz = p->position[3 * is + 2] - posm[3 * l + 2];
            rc = orc_get_B33(p->B + 5 * is, p->volume[is], &b33);
            if (rc) break;
            const double mu = vi * p->volume[is];
            const double *B = p->B + 5 * is;
            Bm[6 * l + 0] = Bm[6 * l + 0] + mu * (5.0 * (delx * delx) + B[0]);
            Bm[6 * l + 1] = Bm[6 * l + 1] + mu * (5.0 * delx * dely + B[1]);
            Bm[6 * l + 2] = Bm[6 * l + 2] + mu * (5.0 * delx * delz + B[2]);
            Bm[6 * l + 3] = Bm[6 * l + 3] + mu * (5.0 * (dely * dely) + B[3]);
            Bm[6 * l + 4] = Bm[6 * l + 4] + mu * (5.0 * dely * delz + B[4]);
            Bm[6 * l + 5] = Bm[6 * l + 5] + mu * (5.0 * (delz * delz) + b33);
        }
        /* geometric_merge normalisation (:319-352) */
        for (int64_t i = 0; i < n; ++i) grp[i] = -1;
        l2 = 0;
        for (int64_t m = 0; m < nm && rc == ORC_OK; ++m) {
            const int64_t ic = ic_m[m];
            if (grp[ic] >= 0) continue;
            const int64_t l = l2++;
            grp[ic] = l;
            const double *b = Bm + 6 * l;
            const double detB = b[0] * (b[3] * b[5] - b[4] * b[4]) - b[1] * (b[1] * b[5] - b[2] * b[4]) +
                                b[2] * (b[1] * b[4] - b[2] * b[3]);
            const double abc = f34 * vm[l] * (one / ORC_PI);
            const double factor = pow(abc * abc / detB, f13);
            for (int k = 0; k < 5; ++k) p->B[5 * ic + k] = b[k] * factor;
        }
        free(grp); free(vm); free(x0); free(y0); free(posm); free(bm); free(hm); free(wm); free(Bm);
    }
    /* parcels%delete (parcel_container.f90:413-452): holes filled from the end */
    if (rc == ORC_OK && nm > 0) {
        int64_t l = n - 1, k = nm - 1;
        while (k >= 0 && l == isma[k]) {
            --l;
            --k;
        }
        for (int64_t m = 0; m <= k; ++m) {
            const int64_t dst = isma[m], src = l;
            memcpy(p->position + 3 * dst, p->position + 3 * src, 3 * sizeof(double));
            memcpy(p->vorticity + 3 * dst, p->vorticity + 3 * src, 3 * sizeof(double));
            memcpy(p->B + 5 * dst, p->B + 5 * src, 5 * sizeof(double));
            p->volume[dst] = p->volume[src];
            p->buoyancy[dst] = p->buoyancy[src];
            p->humidity[dst] = p->humidity[src];
            memcpy(p->delta_pos + 3 * dst, p->delta_pos + 3 * src, 3 * sizeof(double));
            memcpy(p->delta_vor + 3 * dst, p->delta_vor + 3 * src, 3 * sizeof(double));
            memcpy(p->delta_b + 5 * dst, p->delta_b + 5 * src, 5 * sizeof(double));
            memcpy(p->strain + 8 * dst, p->strain + 8 * src, 8 * sizeof(double));
            memcpy(p->Vetas + 3 * dst, p->Vetas + 3 * src, 3 * sizeof(double));
            memcpy(p->Vtaus + 3 * dst, p->Vtaus + 3 * src, 3 * sizeof(double));
            --l;
            while (k >= 0 && l == isma[k]) {
                --l;
                --k;
            }
        }
        p->n = n - nm;
        /* parcel_communicate -> locate_parcels: periodic BC on every parcel */
        for (int64_t i = 0; i < p->n; ++i) orc_apply_periodic_bc(d, p->position + 3 * i);
    }
#undef PPOS
    if (n_merge_out && rc == ORC_OK) *n_merge_out = nm;
    free(sml); free(ipos); free(ipid); free(imid); free(nppc); free(kc1); free(kc2); free(loca); free(node);
    free(iclo); free(dclo); free(leaf); free(avail); free(merged); free(dom_); free(iso); free(isma); free(ic_m);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* parcel_correction.f90: apply_laplace (:142-181) with diverge             */
/* (inversion.f90:410-455) and lapinv1 (inversion_utils.f90:565-589, tables */
/* :240-271); apply_gradient (:183-259); apply_vortcor (:89-138)            */
/* ------------------------------------------------------------------------ */
/* Lap(phi) = rhs semi-spectrally: 2nd-order differences in z, FP64
 * tridiagonal elimination per horizontal wavenumber; the (0,0) mode is zero */
static void lapinv1(const orc_dom *d, const orc_inv *v, double *fs)
{
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    const double dzisq = d->dxi[2] * d->dxi[2], ap = dzisq;
#pragma omp parallel for schedule(static) num_threads(nthreads_of(d))
    for (int kx = 0; kx < nx; ++kx) {
        double *etdv = (double *)malloc(sizeof(double) * (nz + 1));
        double *htdv = (double *)malloc(sizeof(double) * (nz + 1));
        double *rs = (double *)malloc(sizeof(double) * (nz + 1));
        for (int ky = 0; ky < ny; ++ky) {
            double *f = fs + SIDX(d, 0, ky, kx);
            if (kx == 0 && ky == 0) {
                for (int iz = 0; iz <= nz; ++iz) f[iz] = zero;
                continue;
            }
            const double a0 = -two * dzisq - v->k2l2[ky + ny * kx];
            htdv[0] = one / a0;
            etdv[0] = -two * ap * htdv[0];
            for (int iz = 1; iz <= nz - 1; ++iz) {
                htdv[iz] = one / (a0 + ap * etdv[iz - 1]);
                etdv[iz] = -ap * htdv[iz];
            }
            htdv[nz] = one / (a0 + two * ap * etdv[nz - 1]);
            for (int iz = 0; iz <= nz; ++iz) rs[iz] = f[iz];
            f[0] = rs[0] * htdv[0];
            for (int iz = 1; iz <= nz - 1; ++iz) f[iz] = (rs[iz] - ap * f[iz - 1]) * htdv[iz];
            f[nz] = (rs[nz] - two * ap * f[nz - 1]) * htdv[nz];
            for (int iz = nz - 1; iz >= 0; --iz) f[iz] = etdv[iz] * f[iz + 1] + f[iz];
        }
        free(etdv);
        free(htdv);
        free(rs);
    }
}

/* parcel correction field: grad phi with Lap phi = volg/vcell - 1 */
static int laplace_grad(const orc_dom *d, orc_inv *v, const double *volg, double *grad)
{
    const int64_t H = orc_halo_size(d), G = orc_spec_size(d);
    double *phi = (double *)malloc(sizeof(double) * H);
    double *ds = (double *)malloc(sizeof(double) * G), *ws = (double *)malloc(sizeof(double) * G);
    for (int64_t i = 0; i < H; ++i) phi[i] = volg[i] * (one / d->vcell) - one;
    int rc = orc_fftxyp2s(d, phi, ds);
    if (rc == ORC_OK) {
        for (int iz = 0; iz <= d->nz; ++iz) ds[SIDX(d, iz, 0, 0)] = zero;
        lapinv1(d, v, ds);
        orc_diffx(d, ds, ws);
        orc_fftxys2p(d, ws, grad + 0 * H);
        orc_diffy(d, ds, ws);
        orc_fftxys2p(d, ws, grad + 1 * H);
        central_diffz_ss(d, ds, ws);
        for (int kx = 0; kx < d->nx; ++kx)
            for (int ky = 0; ky < d->ny; ++ky) {
                ws[SIDX(d, 0, ky, kx)] = zero;
                ws[SIDX(d, d->nz, ky, kx)] = zero;
            }
        orc_fftxys2p(d, ws, grad + 2 * H);
        for (int c = 0; c < 3; ++c) {
            halo_fill(d, grad + c * H);
            extrap_lin(d, grad + c * H);   /* grad(-1) = 2 grad(0) - grad(1), likewise at nz+1 */
        }
    }
    free(phi);
    free(ds);
    free(ws);
    return rc;
}

int orc_apply_laplace(const orc_dom *d, orc_inv *v, orc_parcels *p, orc_fields *f, int l_reuse)
{
    int rc = orc_par2grid(d, p, f, l_reuse);   /* volg as vol2grid forms it (parcel_interpl.f90:92-136) */
    if (rc) return rc;
    const int64_t H = orc_halo_size(d);
    double *grad = (double *)calloc((size_t)(3 * H), sizeof(double));
    rc = laplace_grad(d, v, f->volg, grad);
    if (rc == ORC_OK) {
        static const int ord[8] = {0, 4, 2, 6, 1, 5, 3, 7};
        for (int64_t n = 0; n < p->n; ++n) {
            double *x = p->position + 3 * n;
            int ijk[3];
            double w[8];
            orc_trilinear(d, x, ijk, w);
            int64_t hs[8];
            for (int kk = 0; kk < 2; ++kk)
                for (int jj = 0; jj < 2; ++jj)
                    for (int ii = 0; ii < 2; ++ii) hs[kk * 4 + jj * 2 + ii] = HIDX(d, ijk[2] + kk, ijk[1] + jj, ijk[0] + ii);
            double s[3];
            for (int l = 0; l < 3; ++l) {
                s[l] = 0.0;
                for (int m = 0; m < 8; ++m) s[l] += w[ord[m]] * grad[l * H + hs[ord[m]]];
            }
            for (int l = 0; l < 3; ++l) x[l] = x[l] + s[l];
            orc_apply_periodic_bc(d, x);   /* parcel_communicate */
        }
    }
    free(grad);
    return rc;
}

int orc_apply_gradient(const orc_dom *d, orc_parcels *p, orc_fields *f, double prefactor, double max_compression,
                       int l_reuse)
{
    int rc = orc_par2grid(d, p, f, l_reuse);
    if (rc) return rc;
    const double vci = one / d->vcell;
#define PHI(k, j, i) (f->volg[HIDX(d, k, j, i)] * vci - one)
    for (int64_t n = 0; n < p->n; ++n) {
        double *x = p->position + 3 * n;
        int ijk[3];
        double w[8];
        orc_trilinear(d, x, ijk, w);
        const int is = ijk[0], js = ijk[1], ks = ijk[2];
        const double xf = w[1] + w[3] + w[5] + w[7];
        const double yf = w[2] + w[3] + w[6] + w[7];
        const double zf = w[4] + w[5] + w[6] + w[7];
        const double xfc = one - xf, yfc = one - yf, zfc = one - zf;
        double lim_x = d->dx[0] * xf * xfc;
        double xs = -prefactor * lim_x *
                    (zfc * (yfc * (PHI(ks, js, is + 1) - PHI(ks, js, is)) + yf * (PHI(ks, js + 1, is + 1) - PHI(ks, js + 1, is))) +
                     zf * (yfc * (PHI(ks + 1, js, is + 1) - PHI(ks + 1, js, is)) +
                           yf * (PHI(ks + 1, js + 1, is + 1) - PHI(ks + 1, js + 1, is))));
        lim_x = lim_x * max_compression;
        xs = fmax(-lim_x, fmin(xs, lim_x));
        double lim_y = d->dx[1] * yf * yfc;
        double ys = -prefactor * lim_y *
                    (zfc * (xfc * (PHI(ks, js + 1, is) - PHI(ks, js, is)) + xf * (PHI(ks, js + 1, is + 1) - PHI(ks, js, is + 1))) +
                     zf * (xfc * (PHI(ks + 1, js + 1, is) - PHI(ks + 1, js, is)) +
                           xf * (PHI(ks + 1, js + 1, is + 1) - PHI(ks + 1, js, is + 1))));
        lim_y = lim_y * max_compression;
        ys = fmax(-lim_y, fmin(ys, lim_y));
        double lim_z = d->dx[2] * zf * zfc;
        double zs = -prefactor * lim_z *
                    (xfc * (yfc * (PHI(ks + 1, js, is) - PHI(ks, js, is)) + yf * (PHI(ks + 1, js + 1, is) - PHI(ks, js + 1, is))) +
                     xf * (yfc * (PHI(ks + 1, js, is + 1) - PHI(ks, js, is + 1)) +
                           yf * (PHI(ks + 1, js + 1, is + 1) - PHI(ks, js + 1, is + 1))));
        lim_z = lim_z * max_compression;
        zs = fmax(-lim_z, fmin(zs, lim_z));
        x[0] = x[0] + xs;
        x[1] = x[1] + ys;
        x[2] = x[2] + zs;
        orc_apply_periodic_bc(d, x);
    }
#undef PHI
    return ORC_OK;
}

void orc_volume_weighted_vorticity(const orc_parcels *p, double vor_bar[3])
{
    double vs = zero, w[3] = {zero, zero, zero};
    for (int64_t n = 0; n < p->n; ++n) {
        vs = vs + p->volume[n];
        for (int c = 0; c < 3; ++c) w[c] = w[c] + p->vorticity[3 * n + c] * p->volume[n];
    }
    for (int c = 0; c < 3; ++c) vor_bar[c] = w[c] / vs;
}

void orc_apply_vortcor(orc_parcels *p, const double vor_bar[3])
{
    double now[3];
    orc_volume_weighted_vorticity(p, now);
    for (int c = 0; c < 3; ++c) now[c] = now[c] - vor_bar[c];
    for (int64_t n = 0; n < p->n; ++n)
        for (int c = 0; c < 3; ++c) p->vorticity[3 * n + c] = p->vorticity[3 * n + c] - now[c];
}
