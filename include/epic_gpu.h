/*
 * epic_gpu.h -- C-ABI of the MI355X parcel-grid core (libepic_gpu.so).
 *
 * Drop-in for EPIC's per-timestep core.  The reference's entry points are
 * Fortran module procedures without array arguments that read and write
 * module globals (parcels, field arrays, inversion tables); here one handle
 * owns that state on the GPU.  A Fortran caller binds these functions with
 * ISO_C_BINDING (see epic_amd/fortran/epic_gpu_mod.f90).
 *
 * Conventions
 *  - Return value: 0 on success, a negative EPIC_E_* code otherwise; the
 *    message of the last failure is returned by epic_last_error().  The
 *    reference aborts on these conditions (mpi_utils.f90:15-33); the caller
 *    decides here.
 *  - Host arrays are caller-owned FP64 arrays in the reference's Fortran
 *    layouts, copied on every set/get call; no pointer is kept after return:
 *      parcels  position(3,n), vorticity(3,n), B(5,n), volume(n),
 *               buoyancy(n), humidity(n)      (parcel_container.f90:29-57)
 *      fields   f(-1:nz+1, -1:ny+1, -1:nx+1[, ncomp])   (fields.f90:66-104)
 *  - One host thread per handle.  Work is queued on the handle's HIP stream;
 *    every entry point below returns after its work has completed.
 *  - Device state (parcels, fields, spectra) stays resident between calls.
 */
#ifndef EPIC_GPU_H
#define EPIC_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct epic_member *epic_handle;
typedef struct epic_batch *epic_batch_handle;

/* domain: parameters.f90:93-165 (nx, ny, nz, lower, extent) */
typedef struct {
    int nx, ny, nz;
    double lower[3];
    double extent[3];
} epic_grid_t;

/* physical constants used on the path: physics.f90:170-185 */
typedef struct {
    double glat;        /* g L_v / (c_p theta_0) */
    double lambda_c;    /* 1 / height_c */
    double q0;          /* saturation humidity at ground level */
    double f_cor[3];    /* planetary vorticity */
} epic_phys_t;

/* run options: l_bndry_zeta_zero (parameters.f90:168-217), rk_order
 * (options.f90:32), vmin = vcell / parcel%min_vratio (parameters.f90:148) */
typedef struct {
    int zeta_bndry_zero[2];
    int rk_order;       /* 3 or 4 */
    double vmin;        /* <= 0 selects vcell / 20 */
} epic_flags_t;

enum {
    EPIC_OK = 0,
    EPIC_E_INVALID = -1,       /* bad argument / unsupported grid */
    EPIC_E_HIP = -2,           /* HIP runtime failure */
    EPIC_E_PARCEL_COUNT = -3,  /* "par2grid: Wrong total number of parcels!" (parcel_interpl.f90:248-250) */
    EPIC_E_B33 = -4,           /* "get_B33: Division by small number!" (parcel_ellipsoid.f90:391-394) */
    EPIC_E_OUT_OF_GRID = -5,   /* a support point left the z halo layers, or lies more than 64 cells from its parcel */
    EPIC_E_RK_ORDER = -6,      /* "Only third and fourth order RK supported." (ls_rk.f90:76-78) */
    EPIC_E_DT = -7,            /* time step must be positive (rk_utils.f90:179-182) */
    EPIC_E_NOMEM = -8,
    EPIC_E_CAPACITY = -9,      /* more parcels than max_parcels */
    EPIC_E_MERGE = -10         /* "Merge error: no near enough neighbour found." (parcel_nearest.f90:551-556) */
};

/* field selectors for epic_get_field / epic_set_field (fields.f90:21-51) */
enum {
    EPIC_FIELD_VOLG = 0,      /* 1 comp */
    EPIC_FIELD_VORTG = 1,     /* 3 */
    EPIC_FIELD_TBUOYG = 2,    /* 1 */
    EPIC_FIELD_DBUOYG = 3,    /* 1 */
    EPIC_FIELD_HUMG = 4,      /* 1 */
    EPIC_FIELD_VELOG = 5,     /* 3 */
    EPIC_FIELD_VELGRADG = 6,  /* 8: du/dx du/dy du/dz dv/dx dv/dy dv/dz dw/dx dw/dy */
    EPIC_FIELD_VTEND = 7,     /* 3 */
    EPIC_FIELD_NPARG = 8,     /* integer, 1 */
    EPIC_FIELD_NSPARG = 9     /* integer, 1 */
};

/* Allocates a member: field_alloc (fields.f90:66-104), parcels%allocate
 * (parcel_ellipsoid.f90:84-134), init_inversion (inversion_utils.f90:79-275),
 * ls_rk_setup (ls_rk.f90:64-80).  nx even, nx, ny >= 8; nx/2, ny and nz of
 * the form 2^a 3^b 5^c (the radices of stafft.f90:128-190), nz >= 4;
 * max_parcels <= EPIC_MAX_PARCELS (2^30 - 1).  A transform work-group holds
 * 16 lines of a horizontal FP64 transform, or one vertical column, in LDS:
 * with 160 KB per work-group (gfx950) nx <= 600, ny <= 310 and nz <= 930.
 * Otherwise EPIC_E_INVALID. */
#define EPIC_MAX_PARCELS ((int64_t)0x3fffffff)
int epic_create(const epic_grid_t *grid, const epic_phys_t *phys, const epic_flags_t *flags,
                int64_t max_parcels, epic_handle *out);
/* finalise_inversion + field_dealloc + parcels%deallocate */
int epic_destroy(epic_handle h);
const char *epic_last_error(epic_handle h);

/* parcel state in / out (FP64 host arrays, Fortran layout).  Positions are
 * kept in FP64 on the device, the other attributes in FP32. */
int epic_set_parcels(epic_handle h, int64_t n, const double *position, const double *vorticity, const double *B,
                     const double *volume, const double *buoyancy, const double *humidity);
int epic_get_parcels(epic_handle h, int64_t n, double *position, double *vorticity, double *B, double *volume,
                     double *buoyancy, double *humidity);
/* RK work arrays delta_pos(3,n), delta_vor(3,n), strain(8,n) (any may be NULL) */
int epic_get_parcel_tendencies(epic_handle h, int64_t n, double *delta_pos, double *delta_vor, double *strain);
int epic_set_parcel_tendencies(epic_handle h, int64_t n, const double *delta_pos, const double *delta_vor);

/* gridded fields (FP64 host arrays, Fortran halo layout, components last) */
int epic_get_field(epic_handle h, int which, double *out);
int epic_set_field(epic_handle h, int which, const double *in);
int epic_get_field_int(epic_handle h, int which, int *out);

/* ---- the reference call surface ---- */
/* par2grid(l_reuse) (parcel_interpl.f90:146-323); l_reuse: -1 absent, 0/1 */
int epic_par2grid(epic_handle h, int l_reuse);
/* vor2vel + vel2vgrad (inversion.f90:25-317) */
int epic_vor2vel(epic_handle h);
/* vorticity_tendency (inversion.f90:321-366) */
int epic_vorticity_tendency(epic_handle h);
/* grid2par(add) (parcel_interpl.f90:405-470); add: -1 absent, 0/1 */
int epic_grid2par(epic_handle h, int add);
/* ls_rk_step(t) with a fixed dt (ls_rk.f90:88-136; get_time_step's fixed-dt
 * branch rk_utils.f90:83-88); t is advanced by dt.  t_initial is
 * time%initial: the first par2grid reuses the cached support points when
 * t > t_initial. */
int epic_ls_rk_step(epic_handle h, double *t, double dt, double t_initial);

/* Time options (options.f90:99-107): fixed (time%l_use_fixed_dt) with
 * fixed_dt, else the adaptive step with alpha (default 0.2); with
 * precise_stop the step is cut at limit. */
typedef struct {
    int fixed;
    double fixed_dt;
    double alpha;
    double limit;
    int precise_stop;
} epic_time_t;
/* get_time_step(t) (rk_utils.f90:73-183) from the current velgradg and
 * tbuoyg (after epic_vor2vel and epic_par2grid): dt = min(alpha/gmax,
 * alpha/bmax), gmax the largest |eigenvalue| of the symmetrised velocity
 * gradient over the nodes, bmax the fourth root of the largest squared
 * buoyancy gradient (central differences), both at least epsilon; the fixed
 * branch returns fixed_dt.  dt <= 0 fails with EPIC_E_DT.  gmax and bmax
 * (may be NULL) receive the maxima (adaptive branch). */
int epic_get_time_step(epic_handle h, double t, const epic_time_t *opts, double *dt, double *gmax, double *bmax);
/* ls_rk_step(t) with the time step of get_time_step, taken after the first
 * vorticity_tendency as in ls_rk.f90:88-136; the whole step, time step
 * included, is one captured HIP graph.  *dt (may be NULL) receives the step
 * taken; t is advanced by it. */
int epic_ls_rk_step_adaptive(epic_handle h, double *t, const epic_time_t *opts, double t_initial, double *dt);

/* parcel_split (parcel_split.f90:37-197): every parcel whose aspect ratio
 * sqrt(D1/D3) >= lambda_max or whose largest eigenvalue D1 >= amax^2 splits
 * along its major axis into two parcels of half the volume, with B reduced by
 * 3/4 D1 v1 v1^T, at x -/+ sqrt(3/5)/2 sqrt(D1) v1 (reflective, then periodic
 * BC).  Children are appended after the existing parcels in ascending order
 * of their parent's index.  lambda_max <= 0 selects 4 (options.f90:89),
 * amax <= 0 selects (3/(4 pi))^(1/3) min(dx) (parameters.f90:150).  If the
 * split would exceed max_parcels, fails with EPIC_E_CAPACITY and changes
 * nothing (the reference grows its container instead).  The cached
 * support-point axes are recomputed from the new shapes. */
int epic_parcel_split(epic_handle h, double lambda_max, double amax, int64_t *n_split);

/* parcel_merge (parcel_merge.f90:48-109 with find_nearest,
 * parcel_nearest.f90:375-627, 937-1377) with the semantics of the reference
 * on one MPI rank: every parcel with volume < vmin merges into its nearest
 * neighbour (periodic images of boundary small parcels included) as the
 * tree resolution allows; mergers conserve volume, volume-weighted
 * vorticity, buoyancy and humidity, take the volume-weighted centre and the
 * summed second moments rescaled to det B = (3V/4pi)^2.  Merged parcels are
 * deleted, their slots filled from the end of the container.  The cached
 * support-point axes are recomputed.  *n_merge: number of parcels removed. */
int epic_parcel_merge(epic_handle h, int64_t *n_merge);

/* apply_laplace(l_reuse) (parcel_correction.f90:142-181): vol2grid, then
 * the parcel displacement grad phi with Lap phi = volg/vcell - 1 (diverge,
 * inversion.f90:410-455; horizontal FFTs, FP64 tridiagonal z solve lapinv1,
 * inversion_utils.f90:565-589; the (0,0) mode is zero), extrapolated to the
 * z halo, interpolated trilinearly at each parcel centre and added to its
 * position; periodic BC in x and y.  l_reuse as for epic_par2grid. */
int epic_apply_laplace(epic_handle h, int l_reuse);
/* apply_gradient(prefactor, max_compression, l_reuse)
 * (parcel_correction.f90:183-259): the limited displacement
 * -prefactor lim grad phi from phi = volg/vcell - 1 on the 8 nodes around
 * each parcel centre, clipped to +/- max_compression lim per direction with
 * lim = dx w (1 - w); periodic BC.  EPIC's defaults are prefactor 1.8
 * (gradient_pref) and max_compression 0.5 (options.f90:92-93). */
int epic_apply_gradient(epic_handle h, double prefactor, double max_compression, int l_reuse);
/* init_parcel_correction (parcel_correction.f90:42-83): stores the volume-
 * weighted mean vorticity of the parcels as the target of
 * epic_apply_vortcor; vor_bar (3 doubles, may be NULL) receives it. */
int epic_init_parcel_correction(epic_handle h, double *vor_bar);
/* apply_vortcor (parcel_correction.f90:89-138): shifts every parcel's
 * vorticity so that the volume-weighted mean equals the stored target. */
int epic_apply_vortcor(epic_handle h);

/* calculate_parcel_diagnostics (parcel_diagnostics.f90:51-159) on one rank
 * with ape_calculation = 'none': stats[18] in the reference's IDX order
 * (ape, ke, n_small, avg_lam, avg_vol, std_lam, std_vol, sum_vol, rms_xi,
 * rms_eta, rms_zeta, ntot_par, enstrophy, nsplits, nbig_iclo, nmerges,
 * min_buoy, max_buoy).  The velocity is delta_pos, so call it after
 * epic_grid2par, as ls_rk_step does (ls_rk.f90:108-110).  nsplits and
 * nmerges count epic_parcel_split / epic_parcel_merge since the last reset;
 * nbig_iclo is collect_merge_stats' n_big_close (parcel_merge.f90:365-393),
 * the merged pairs whose target was big; reset_counters != 0 clears the
 * three and the n-way counts (parcel_diagnostics_netcdf.f90:181-184).  Sums
 * are FP64 in a fixed order. */
int epic_parcel_diagnostics(epic_handle h, double *stats, int reset_counters);
/* parcel_merge_stats (parcel_diagnostics.f90:132-133): n_way[j-1] = the
 * merges since the last reset whose target took j small parcels (j = 7:
 * seven or more), collect_merge_stats' n_way_parcel_mergers(7)
 * (parcel_merge.f90:381-391). */
int epic_parcel_merge_stats(epic_handle h, int64_t *n_way);
/* No counterpart in the reference: epic_parcel_diagnostics for callers that
 * cannot allocate device or pinned memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same stats[18] and
 * counters; calls no hipMalloc, hipFree, hipHostMalloc or other allocating
 * routine.  Its reduction partials live in one fixed device array per GPU
 * (1,024 blocks x 12 doubles plus the totals, 98,432 bytes, held until the
 * process ends; a grid-stride loop covers any parcel count up to
 * EPIC_MAX_PARCELS, beyond which it returns EPIC_E_INVALID).  The sums are
 * FP64 in a fixed order, which differs from epic_parcel_diagnostics' above
 * 262,144 parcels (agreement to FP64 rounding).  Calls that overlap on one
 * GPU share that array and race. */
int epic_parcel_diagnostics_noalloc(epic_handle h, double *stats, int reset_counters);
/* calculate_field_diagnostics (field_diagnostics.f90:40-154) on one rank
 * from volg, nparg, nsparg, tbuoyg, velog and vortg (after epic_par2grid and
 * epic_vor2vel): stats[11] = rms_v, avg_npar, avg_nspar, keg, apeg (0),
 * eng, abserr_v, max_npar, max_buoy, min_npar, min_buoy. */
int epic_field_diagnostics(epic_handle h, double *stats);

/* Physical re-sort: the parcel arrays are reordered into the cell order of
 * the current positions, so that the permutation gathers of the step stay
 * coalesced as parcels drift.  Not observable through the API
 * (epic_get_parcels returns the caller's order); the stage tendencies are
 * not kept.  No counterpart in the reference, whose parcels are not
 * sorted.  epic_ls_rk_step and epic_batch_step re-sort a member every
 * `steps` steps (3 by default; 0 never). */
int epic_resort(epic_handle h);
int epic_set_resort_interval(epic_handle h, int steps);

/* The grid, physics and flags the handle was created with. */
int epic_get_config(epic_handle h, epic_grid_t *grid, epic_phys_t *phys, epic_flags_t *flags);

/* ---- NetCDF files (NetCDF classic format; parcel_netcdf.f90, field_netcdf.f90) ----
 * The reference's dimensions, variables, units and long names in the classic
 * binary format (CDF-2, or CDF-5 when a record exceeds 4 GiB), written and
 * read without a netCDF library.  The reference's NetCDF-4 groups
 * ('parameters', 'physical_quantities') become global attributes
 * "<group>.<name>".  NetCDF-4/HDF5 files are not readable here. */
/* write_netcdf_parcels(t) (parcel_netcdf.f90:201-309): one file, one record */
int epic_write_parcels_nc(epic_handle h, const char *path, double t);
/* the same from host arrays (Fortran layouts), without a handle */
int epic_nc_write_parcels(const char *path, const epic_grid_t *grid, const epic_phys_t *phys,
                          const epic_flags_t *flags, int64_t n, const double *position, const double *vorticity,
                          const double *B, const double *volume, const double *buoyancy, const double *humidity,
                          double t);
/* read a parcel file into host arrays of `capacity` parcels; with position
 * NULL only *n, *grid (may be NULL; nx = ny = nz = 0 when the file has no
 * box attributes) and *t (may be NULL) are filled */
int epic_nc_read_parcels(const char *path, int64_t capacity, int64_t *n, epic_grid_t *grid, double *position,
                         double *vorticity, double *B, double *volume, double *buoyancy, double *humidity,
                         double *t);
/* field file from host arrays in the halo layout of epic_get_field
 * (nparg and nsparg as doubles) */
int epic_nc_write_fields(const char *path, const epic_grid_t *grid, const epic_phys_t *phys,
                         const epic_flags_t *flags, double t, const double *velog, const double *vtend,
                         const double *vortg, const double *tbuoyg, const double *dbuoyg, const double *humg,
                         const double *volg, const double *nparg, const double *nsparg);
/* message of the last failure of the epic_nc_* calls (this thread) */
const char *epic_nc_last_error(void);
/* read_netcdf_parcels (parcel_netcdf.f90:311-724) on one rank: the last
 * record; B11..B23, positions and volume required, vorticity, buoyancy and
 * humidity optional (0); a box in the file must match the grid; *t (may
 * be NULL) receives the record's time */
int epic_read_parcels_nc(epic_handle h, const char *path, double *t);
/* write_netcdf_fields(t) (field_netcdf.f90:182-266): the node fields of
 * planes 0..nz with the coordinate axes x, y, z and one time record */
int epic_write_fields_nc(epic_handle h, const char *path, double t);
/* get_file_type, get_time (the last record's t), get_num_steps and
 * get_netcdf_box of a file (netcdf_reader.f90; utils.f90:144-160);
 * file_type (16 chars) is "fields" or "parcels"; any pointer may be NULL */
int epic_nc_file_info(const char *path, char *file_type, double *t, int64_t *n_steps, epic_grid_t *grid);
/* read_netcdf_fields(fname, step) (field_netcdf.f90:306-386) into host
 * arrays in the halo layout of epic_get_field: x/y/z_vorticity, buoyancy
 * (total) and humidity of record `step` (1-based, -1 the last) on planes
 * 0..nz; an absent dataset leaves its field zero */
int epic_nc_read_fields(const char *path, int step, const epic_grid_t *grid, double *vortg, double *tbuoyg,
                        double *humg, double *t);
/* the same into the member's vortg, tbuoyg and humg (a box in the file must
 * match the member's grid) */
int epic_read_fields_nc(epic_handle h, const char *path, int step, double *t);

/* ---- parcel set-up and restart (parcel_init.f90, utils.f90:144-249) ---- */
/* parcel_default (parcel_init.f90:31-100): n_per_cell parcels per cell on
 * the regular lattice of init_regular_positions (:104-152), volume vcell /
 * n_per_cell, spheroids B11 = B22 = (lam abc)^(2/3) with lam the largest cell
 * aspect ratio, split (init_refine, :156-166) while lam >= lambda_max;
 * vorticity, buoyancy and humidity zero.  n_per_cell <= 0 selects 8 and
 * lambda_max <= 0 selects 4 (options.f90:88-89); n_per_cell must be a cube
 * (else EPIC_E_INVALID), the parcels must fit max_parcels (EPIC_E_CAPACITY). */
int epic_parcel_default(epic_handle h, int n_per_cell, double lambda_max);
/* init_parcels_from_grids (parcel_init.f90:174-203): every parcel adds the
 * trilinear interpolation of vortg, tbuoyg and humg at its centre to its
 * vorticity, buoyancy and humidity */
int epic_init_parcels_from_grids(epic_handle h);
/* get_next_write_time (utils.f90:48-54): freq (int(t / freq) + 1) */
double epic_next_write_time(double freq, double t);
/* setup_restart + the restart branch of setup_fields_and_parcels
 * (utils.f90:144-160, 200-216): *t receives the file's last time and
 * file_type (16 chars, may be NULL) its type; a "parcels" file is read into
 * the member, a "fields" file gives parcel_default (8 per cell) with the
 * file's last vorticity, buoyancy and humidity interpolated to the parcels;
 * any other type fails with EPIC_E_INVALID */
int epic_setup_restart(epic_handle h, const char *path, double *t, char *file_type);

/* Cell counting sort of the parcels' current positions with the reference's
 * key ijk = 1 + ix + hx iy + hx hy iz on the halo box
 * (parcel_nearest.f90:196-257, 411-424, 531-541): loca(n) 1-based cell of
 * each parcel, nppc/kc1 (ncell = (nx+3)(ny+3)(nz+2)) counts and 1-based
 * first slots, node(n) parcel indices (0-based) in sorted order. */
int epic_cell_sort(epic_handle h, int *loca, int *nppc, int *kc1, int *node);
int64_t epic_nearest_ncell(epic_handle h);

int64_t epic_num_parcels(epic_handle h);

/* scherzinger_diagonalise(A, D, V) (scherzinger.f90:31-112), as called by
 * diagonalise (parcel_ellipsoid.f90:359-376), on the device in FP64 for n
 * symmetric matrices A(3,3,n): D(3,n) eigenvalues in descending order,
 * V(3,3,n) with V(:,k,i) the unit eigenvector of D(k,i).  This is the solver
 * every par2grid, push, split and merge applies to the parcels' B.  Without
 * a handle; the message of a failure is epic_diagonalise_last_error()
 * (this thread). */
int epic_diagonalise(int64_t n, const double *A, double *D, double *V);
const char *epic_diagonalise_last_error(void);

/* Per-phase GPU timers, the device counterpart of the reference's named
 * wall-clock timers (mpi_timer.f90:53-104, names registered in
 * epic3d.f90:79-95).  With profiling on (on = 1), every phase is bracketed
 * by HIP events on the handle's stream and epic_ls_rk_step launches its
 * kernels eagerly; with on = 2 the step stays one graph replay, captured
 * with a one-thread kernel storing the device wall clock before and after
 * every phase, so the timers measure the replayed step itself (other entry
 * points time as with on = 1).  epic_get_timers returns the accumulated
 * milliseconds and call counts per phase (ms/calls may be NULL) and resets
 * them when reset != 0.  "push" is the fused grid2par + dB/dt + RK push
 * kernel of ls_rk_step; "grid2par" is the stand-alone grid2par. */
enum {
    EPIC_TIMER_PAR2GRID = 0,    /* cell sort + tile scatter + finalise */
    EPIC_TIMER_CELL_SORT = 1,   /* keys, scan, placement, per-cell order (inside par2grid) */
    EPIC_TIMER_VOR2VEL = 2,
    EPIC_TIMER_VTEND = 3,
    EPIC_TIMER_PUSH = 4,
    EPIC_TIMER_GRID2PAR = 5,
    EPIC_TIMER_AXES = 6,        /* support-point half axes (FP64 eigen-solve), once per stage */
    EPIC_TIMER_P2G_TILES = 7,   /* the tile scatter kernel alone (inside par2grid) */
    EPIC_NUM_TIMERS = 8
};
int epic_set_profiling(epic_handle h, int on);
int epic_get_timers(epic_handle h, double *ms, int *calls, int reset);

/* ---- ensembles: batched step, tiling autotuner, admission estimate ---- */
/* Device bytes epic_create allocates for this grid and capacity (the
 * admission figure of a server that keeps members resident in HBM). */
int64_t epic_member_bytes(const epic_grid_t *grid, int64_t max_parcels);
/* Picks the z chunk (cell layers per work-group: 2, 4 or 8) of the
 * par2grid and the fused grid2par/push tile kernels by timing each
 * candidate on this member's parcels; leaves position, vorticity and B unchanged (the
 * RK stage work arrays are overwritten).  Returns the choice. */
int epic_autotune(epic_handle h, int *tz_p2g, int *tz_g2p);
int epic_set_tiling(epic_handle h, int tz_p2g, int tz_g2p);
/* A batch advances many members (uneven grids and parcel counts) through
 * one ls_rk_step each (ls_rk.f90:88-136, per member exactly as
 * epic_ls_rk_step): every kernel of the step is ONE launch over all members
 * and the whole step is one captured HIP graph, replayed per call.  Members
 * stay usable on their own; a member belongs to any number of batches but
 * calls on it must not overlap. */
int epic_batch_create(const epic_handle *members, int count, epic_batch_handle *out);
int epic_batch_destroy(epic_batch_handle b);
const char *epic_batch_last_error(epic_batch_handle b);
/* t[i] += dt[i] for every member whose step succeeded; status[i] (may be
 * NULL) gets each member's code; returns the first failure or EPIC_OK. */
int epic_batch_step(epic_batch_handle b, double *t, const double *dt, const double *t_initial, int *status);
/* epic_batch_step in two halves, so that a server can admit new members
 * while a step runs: epic_batch_launch queues the step of every member on
 * the batch's own stream and returns at once; epic_batch_wait waits for it,
 * then does the rest of epic_batch_step (statuses, t[i] += dt[i], periodic
 * re-sorts).  Until the wait the batch's members must not be used; other
 * members and batches may be, and their work overlaps on the GPU.
 * epic_batch_prepare builds and captures the step graph (done by the first
 * launch otherwise), e.g. while another batch's step is in flight. */
int epic_batch_prepare(epic_batch_handle b);
int epic_batch_launch(epic_batch_handle b, const double *t, const double *dt, const double *t_initial);
int epic_batch_wait(epic_batch_handle b, double *t, int *status);

#ifdef __cplusplus
}
#endif
#endif /* EPIC_GPU_H */
