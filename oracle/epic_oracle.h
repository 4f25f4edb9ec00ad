/*
 * epic_oracle.h -- FP64 CPU restatement of EPIC's per-timestep parcel-grid
 * core (EPIC 0.14.3).  TEST INFRASTRUCTURE ONLY: this library is the checker
 * for the HIP path (and the timed CPU baseline); it is never linked into,
 * loaded by or called from the product library.
 *
 * Layouts follow the reference's Fortran arrays exactly:
 *   halo field  f(-1:nz+1, -1:ny+1, -1:nx+1[, comp])   z fastest
 *               (src/3d/fields/fields.f90:66-104, single-rank box hlo..hhi;
 *                the x/y halo is 1 lower + 2 upper cells,
 *                src/mpi/mpi_layout.f90:118-123; here one symmetric-size
 *                (n+3) box from -1 to n+1)
 *   spectral    s(0:nz, 0:ny-1, 0:nx-1)                 z fastest
 *   parcels     position(3,n), vorticity(3,n), B(5,n), volume(n), ...
 *               component fastest (src/3d/parcels/parcel_container.f90:29-57)
 *
 * Every function returns 0 on success or a negative ORC_E* code.
 */
#ifndef EPIC_ORACLE_H
#define EPIC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_OK = 0,
    ORC_E_PARCEL_COUNT = -1,   /* par2grid: Wrong total number of parcels (parcel_interpl.f90:248-250) */
    ORC_E_B33 = -2,            /* get_B33: Division by small number (parcel_ellipsoid.f90:391-394) */
    ORC_E_OUT_OF_GRID = -3,    /* a support point lands outside the halo grid */
    ORC_E_RK_ORDER = -4,       /* ls_rk_setup: only 3rd/4th order (ls_rk.f90:76-78) */
    ORC_E_DT = -5,             /* dt <= 0 (rk_utils.f90:179-182) */
    ORC_E_FFT_LEN = -6,        /* length not factorisable by 2,3,5 (stafft.f90:87-96) */
    ORC_E_ALLOC = -7,
    ORC_E_MERGE = -8           /* Merge error: no near enough neighbour found (parcel_nearest.f90:551-556) */
};

typedef struct {
    /* inputs */
    int nx, ny, nz;
    double lower[3], extent[3];
    double glat, lambda_c, q0;     /* physics.f90:170-175 */
    double f_cor[3];
    int zeta_zero[2];              /* l_bndry_zeta_zero (parameters.f90:168-217) */
    double vmin;                   /* <= 0 -> vcell / 20 (options.f90:92, parameters.f90:148) */
    int nthreads;                  /* OpenMP threads (<= 0 -> runtime default) */
    /* derived by orc_dom_init (parameters.f90:93-165) */
    double upper[3], dx[3], dxi[3], vcell, center[3], hl[3], hli[3];
} orc_dom;

int orc_dom_init(orc_dom *d);

/* sizes */
int64_t orc_halo_size(const orc_dom *d);   /* (nz+3)(ny+3)(nx+3) */
int64_t orc_spec_size(const orc_dom *d);   /* (nz+1) ny nx */

/* ---- ellipsoid / eigen (scherzinger.f90, parcel_ellipsoid.f90) ---- */
/* A is a full symmetric 3x3 (row-major here; symmetric so layout-free).
 * D descending, V(:,j) stored as V[3*j + i] (column j contiguous). */
void orc_scherzinger_diagonalise(const double A[9], double D[3], double V[9]);
void orc_scherzinger_eigenvalues(const double A[9], double D[3]);
int orc_get_B33(const double B[5], double volume, double *B33);
int orc_get_points(const double pos[3], const double B[5], double volume,
                   double Veta[3], double Vtau[3], double pts[12]);
int orc_get_dBdt(const double B[5], const double strain[8], double volume, double dB[5]);
void orc_trilinear(const orc_dom *d, const double pos[3], int ijk[3], double w[8]);

/* ---- par2grid / grid2par (parcel_interpl.f90) ---- */
typedef struct {
    int64_t n;
    double *position, *vorticity, *B, *volume, *buoyancy, *humidity;
    double *delta_pos, *delta_vor, *delta_b, *strain;
    double *Vetas, *Vtaus;
} orc_parcels;

typedef struct {
    /* halo layout, H = orc_halo_size */
    double *velog;      /* 3H */
    double *vortg;      /* 3H */
    double *vtend;      /* 3H */
    double *velgradg;   /* 8H */
    double *tbuoyg, *dbuoyg, *humg, *volg;   /* H each */
    int *nparg, *nsparg;                     /* H each */
} orc_fields;

/* l_reuse: -1 absent, 0 false, 1 true (Fortran optional logical) */
int orc_par2grid(const orc_dom *d, orc_parcels *p, orc_fields *f, int l_reuse);
/* add: -1 absent, 0 false, 1 true */
int orc_grid2par(const orc_dom *d, orc_parcels *p, const orc_fields *f, int add);

/* ---- 1D transforms (stafft.f90:196-555), single vector ---- */
int orc_forfft(int n, double *x);
int orc_revfft(int n, double *x);
int orc_dct(int n, double *x);   /* x[0..n] */
int orc_dst(int n, double *x);   /* x[1..n] passed as x pointing at element 1 */

/* ---- 3D spectral wrappers (sta3dfft.f90) ---- */
int orc_fftxyp2s(const orc_dom *d, const double *fp_halo, double *fs);
int orc_fftxys2p(const orc_dom *d, const double *fs, double *fp_halo);
void orc_diffx(const orc_dom *d, const double *fs, double *ds);
void orc_diffy(const orc_dom *d, const double *fs, double *ds);

/* ---- inversion (inversion.f90, inversion_utils.f90) ---- */
typedef struct orc_inv orc_inv;
int orc_inv_create(const orc_dom *d, orc_inv **out);
void orc_inv_destroy(orc_inv *inv);
int orc_vor2vel(const orc_dom *d, orc_inv *inv, orc_fields *f);
int orc_vel2vgrad(const orc_dom *d, orc_inv *inv, const double *svel, orc_fields *f);
int orc_vorticity_tendency(const orc_dom *d, orc_inv *inv, orc_fields *f);

/* ---- stepper (ls_rk.f90) and BCs (parcel_bc.f90) ---- */
void orc_apply_periodic_bc(const orc_dom *d, double pos[3]);
void orc_apply_reflective_bc(const orc_dom *d, double pos[3], double B[5]);
int orc_ls_rk_step(const orc_dom *d, orc_inv *inv, orc_parcels *p, orc_fields *f,
                   double *t, double dt, double t_initial, int rk_order);

/* time options (options.f90:99-107): fixed dt or adaptive with alpha */
typedef struct {
    int fixed;              /* time%l_use_fixed_dt */
    double fixed_dt, alpha, limit;
    int precise_stop;
} orc_time;
/* get_time_step (rk_utils.f90:73-183) on one rank: dt = min(alpha/gmax,
 * alpha/bmax) from the largest strain eigenvalue magnitude and the fourth
 * root of the largest squared buoyancy gradient on the nodes; gmax and bmax
 * (may be NULL) receive the two maxima */
int orc_get_time_step(const orc_dom *d, const orc_fields *f, double t, const orc_time *tm, double *dt,
                      double *gmax, double *bmax);
/* ls_rk_step with the time step of get_time_step (ls_rk.f90:88-136); *dt
 * receives the step taken */
int orc_ls_rk_step_adaptive(const orc_dom *d, orc_inv *inv, orc_parcels *p, orc_fields *f, double *t,
                            const orc_time *tm, double t_initial, int rk_order, double *dt);

/* stability polynomial of the low-storage scheme (test of the coefficients) */
double orc_lsrk_linear(double lam, int rk_order);

/* per-phase timers (seconds, accumulated; epic3d.f90:79-95 names) */
typedef struct { double par2grid, grid2par, vor2vel, vtend, push; } orc_timers;
void orc_get_timers(orc_timers *t);
void orc_reset_timers(void);

/* ---- cell counting sort (parcel_nearest.f90:196-257, 411-424, 531-541) ---- */
/* loca[n] = ijk (1-based halo cell index), node[k] = parcel index (0-based)
 * for k in [kc1[ijk]-1, kc1[ijk]+nppc[ijk]-1), kc1 1-based offsets. */
/* ---- parcel split (parcel_split.f90:37-197) ---- */
double orc_default_amax(const orc_dom *d);   /* (3/(4 pi))^(1/3) min(dx) (parameters.f90:150) */
/* cap: capacity of the arrays of p; amax <= 0 selects the default */
int orc_parcel_split(const orc_dom *d, orc_parcels *p, int64_t cap, double lambda_max, double amax,
                     int64_t *n_split);

/* ---- parcel merge (parcel_merge.f90, parcel_nearest.f90), one rank ---- */
int orc_parcel_merge(const orc_dom *d, orc_parcels *p, int64_t *n_merge);
/* the same, adding collect_merge_stats' counts (parcel_merge.f90:365-393) to
 * stats[8] (may be NULL): [0] n_big_close, [1..7] n_way_parcel_mergers */
int orc_parcel_merge_stats(const orc_dom *d, orc_parcels *p, int64_t *n_merge, int64_t *stats);

/* ---- parcel correction (parcel_correction.f90) ---- */
int orc_apply_laplace(const orc_dom *d, orc_inv *v, orc_parcels *p, orc_fields *f, int l_reuse);
int orc_apply_gradient(const orc_dom *d, orc_parcels *p, orc_fields *f, double prefactor, double max_compression,
                       int l_reuse);
void orc_volume_weighted_vorticity(const orc_parcels *p, double vor_bar[3]);
void orc_apply_vortcor(orc_parcels *p, const double vor_bar[3]);

/* ---- diagnostics (parcel_diagnostics.f90:51-159, field_diagnostics.f90:40-154), one rank,
 * ape_calculation = 'none' (the APE entries stay 0) ---- */
/* stats[0..17] in the order of the reference's IDX_* (1-based there):
 * ape, ke, n_small, avg_lam, avg_vol, std_lam, std_vol, sum_vol, rms_xi,
 * rms_eta, rms_zeta, ntot_par, enstrophy, nsplits, nbig_iclo, nmerges,
 * min_buoy, max_buoy; the three counters are taken from counters[3] */
void orc_parcel_diagnostics(const orc_dom *d, const orc_parcels *p, const double counters[3], double stats[18]);
/* stats[0..10]: rms_v, avg_npar, avg_nspar, keg, apeg, eng, abserr_v,
 * max_npar, max_buoy, min_npar, min_buoy */
void orc_field_diagnostics(const orc_dom *d, const orc_fields *f, double stats[11]);

int64_t orc_nearest_ncell(const orc_dom *d);
int orc_nearest_cell_sort(const orc_dom *d, int64_t n, const double *position,
                          int *loca, int *nppc, int *kc1, int *node);
/* get_index (fields.f90:165-172) */
void orc_get_index(const orc_dom *d, const double pos[3], int ijk[3]);

#ifdef __cplusplus
}
#endif
#endif
