// Per-launch argument tables and plan builders of the device kernels
// (internal to the library).  Every builder appends ONE grouped launch whose
// work items are the given argument structs (members, components or both).
#pragma once

#include <vector>

#include "common.h"
#include "plan.h"

namespace epic {

// parcels of one item are processed PER_BLOCK at a time by a 256-thread block
constexpr int PER_BLOCK = 1024;

// ---- cell sort (p2g.hip) ----
struct KeysArgs {
    Grid g;
    ParcelsDev p;
    int *keys, *count, *scount, *err;
};
struct ScanArgs {
    const int *in;
    int64_t n;
    int *out;      // n + 1 entries: exclusive prefix and total
    int *bsum;     // scan_scratch_ints(n)
    int *clear;    // n entries set to zero on the way (may be null)
};
struct PlaceArgs {
    int64_t n;
    const int *keys, *cell_start;
    int *cursor, *perm;
    int64_t cap = 0;   // grid size in parcels (0: n)
    const int64_t *n_dev = nullptr;   // when set, the count is read from here instead of n
};
int64_t scan_scratch_ints(int64_t n);
void add_cell_keys(Plan &P, int phase, const std::vector<KeysArgs> &items);
void add_exclusive_scan(Plan &P, int phase, const std::vector<ScanArgs> &items);   // three launches
void add_place(Plan &P, int phase, const std::vector<PlaceArgs> &items);
// order of every cell by caller index (epic_cell_sort only; a direct launch)
void launch_cell_order(int64_t nc, const int *cell_start, const int *perm_tmp, const int *orig, int *perm,
                       hipStream_t s);

// ---- support-point half axes + par2grid (p2g.hip) ----
struct AxesArgs {
    ParcelsDev p;
    float *ax[6];   // Veta (3), Vtau (3)
    int *err;
};
struct P2GLayout {
    int tz, ntx, nty, nchz;
    int64_t staging_elems;   // FP64 values
};
P2GLayout p2g_layout(const Grid &g);
struct P2GArgs {
    Grid g;
    ParcelsDev p;
    const float *axes[6];       // Veta (3), Vtau (3) per parcel, FP32
    const float *axes_alt[6];   // taken instead when *use_alt != 0
    const int *use_alt;         // may be null
    const int *perm;
    const int *cell_start;
    double *staging;            // exact partial node sums (scaled integers)
    int *err;
    const float *tmax;          // P2G_NF + 1 maxima from k_p2g_max (fixed-point scales)
    // support points beyond the neighbour cells of their parcel's cell: their
    // fixed-point terms go to far_acc ([field][plane k+1][j][i], 64-bit) by
    // global atomics; far_flag is set when there are any, far_scale receives
    // the fields' inverse scales for the finalize
    unsigned long long *far_acc;
    int *far_flag;
    double *far_scale;
    int tz, ntx, nty, nchz;
};
// largest |contribution| per field (and |tbuoy - buoy|) over a member's parcels
constexpr int P2G_MAX_BLOCKS = 4096;
struct P2GMaxArgs {
    Grid g;
    ParcelsDev p;
    int nblk;                   // p2g_max_blocks(p.n)
    float *out;                 // P2G_NF + 1, zero on entry; cleared again by the finalize kernel
};
int p2g_max_blocks(int64_t n);
void add_p2g_max(Plan &P, int phase, const std::vector<P2GMaxArgs> &items);
struct FinalizeArgs {
    Grid g;
    const double *staging;
    const int *count, *scount;
    float *fld;     // F_VOLG .. F_HUMG
    int *nparg, *nsparg;
    int *clear_count, *clear_scount;   // histograms cleared for the next push (may be null)
    float *clear_tmax;                 // the tile kernel's maxima, zeroed for the next k_p2g_max (may be null)
    const unsigned long long *far_acc; // the tile kernel's far-point sums (read when *far_flag)
    const int *far_flag;
    const double *far_scale;
    int tz, ntx, nty, nchz;
};
// clears far_acc and then far_flag after a finalize that read them
struct FarClearArgs {
    unsigned long long *far_acc;
    int *far_flag;
    int64_t n;                         // P2G_NF * field_stride
};
void add_axes(Plan &P, int phase, const std::vector<AxesArgs> &items);
void add_p2g(Plan &P, int phase, const std::vector<P2GArgs> &tiles, const std::vector<FinalizeArgs> &fin,
             const std::vector<FarClearArgs> &far);

// ---- grid2par + push (g2p.hip) ----
struct G2PArgs {
    Grid g;
    ParcelsDev p;      // p.veta / p.vtau receive the half axes
    const float *fld;
    int add;
    int *err;
};
struct G2PTileArgs {
    Grid g;
    ParcelsDev p;
    const float *axes[6];
    float *axes_next[6];         // null, or where the half axes of the pushed B go (the next stage's)
    float *tmax_next;            // null, or the next par2grid's maxima (atomicMax; zero on entry)
    const int *perm;
    const int *cell_start;
    const float *fld;
    int stage, nstages, reflect;
    double cb, ca;               // the stage pushes by cb * (*dt)
    const double *dt;            // the step's dt (device)
    int *keys, *count, *scount;
    int *err;
    int tz, ntx, nty, nchz;
};
void add_grid2par(Plan &P, int phase, const std::vector<G2PArgs> &items);
void add_rk_stage_tiles(Plan &P, int phase, const std::vector<G2PTileArgs> &items);

// ---- horizontal FFTs (fft.hip) ----
// fp64: spectra and arithmetic in FP64 (double2 spectra, twiddles tw64);
// every item of one launch has the same precision.  The FP64 transforms are
// the vorticity tendency's (modes 1 forward and 2 inverse) and carry the
// halo planes -1 and nz + 1 at z indices nz + 1 and nz + 2.
struct FftXArgs {
    Grid g;
    FftPlan plan;          // complex length nx/2
    const float2 *tw;      // e^{-2 pi i t / nx}, t < nx
    const double2 *tw64;   // the same in FP64
    int fp64;
    int mode;              // forward: 0 field, 1 tendency flux, 3 volg / vcell - 1;
                           // inverse: 0 store, 2 store + d/dz flux
    int slot;              // field slot read (forward, mode 0) or written (inverse)
    int comp_a, comp_b, sgn;
    const float *fld;      // node fields read
    float *fld_out;        // node fields written (inverse)
    void *out;             // forward: T[j][kx][z] (float2 or double2)
    const void *in;        // inverse: T[j][kx][z]
};
struct FftYArgs {
    Grid g;
    FftPlan plan;          // length ny
    const float2 *tw;      // e^{-2 pi i t / ny}
    const double2 *tw64;
    int fp64;
    int mode;              // 0 forward, 1 inverse, 2 inverse of i kx in + i ky in2
    const void *in, *in2;
    void *out;             // may equal in
};
void add_fft_xf(Plan &P, int phase, const std::vector<FftXArgs> &items);
void add_fft_y(Plan &P, int phase, const std::vector<FftYArgs> &items);
void add_fft_xi(Plan &P, int phase, const std::vector<FftXArgs> &items);

// ---- vertical column kernel + fix-ups (inversion.hip) ----
struct ColArgs {
    Grid g;
    FftPlan zplan;         // length 2 nz
    const double2 *ztw;    // e^{-2 pi i t / (2 nz)}
    const float2 *in[3];
    float2 *out[11];       // xi eta zeta u v w dudx dudy dvdy dwdx dwdy
};
struct FixArgs {
    Grid g;
    float *fld;
};
size_t col_lds_bytes(const Grid &g);
// the largest LDS allocation of the transform work-groups (FP64 x and y
// passes, vertical columns)
size_t transform_lds_bytes(const Grid &g);
void add_vor2vel_columns(Plan &P, int phase, const std::vector<ColArgs> &items);
void add_vor2vel_fixup(Plan &P, int phase, const std::vector<FixArgs> &items);
void add_vtend_fixup(Plan &P, int phase, const std::vector<FixArgs> &items);

// ---- parcel splitting (split.hip) ----
struct SplitArgs {
    Grid g;
    ParcelsDev p;          // p.n: parcel count before the split
    double lambda_max, amax;
    int *flag;             // n entries
    const int *rank;       // exclusive prefix of flag (a child's slot: n + rank)
    const int *crank;      // exclusive prefix of flag in caller-index order (its caller index: n + crank[orig])
    int *err;
};
void add_split_flag(Plan &P, int phase, const std::vector<SplitArgs> &items);
void add_split_apply(Plan &P, int phase, const std::vector<SplitArgs> &items);

// ---- parcel merging (merge.hip); direct launches on ctx.s ----
// device scratch reused across calls: bump allocation from base; a call
// that needs more takes the rest from hipMalloc and records its total in
// want, and the next call grows base to it (freed by the owner)
struct ScratchArena {
    char *base = nullptr;
    size_t cap = 0, want = 0;
};
struct MergeCtx {
    Grid g;
    ParcelsDev p;          // p.n: parcel count before the merge
    double vmin;
    hipStream_t s;
    int64_t ncell;         // (nx+3)(ny+3)(nz+2) halo cells
    int *err;
    float *attr;           // the parcel attribute slab (nattr arrays of cap)
    int64_t cap;
    int nattr;
    ScratchArena *scratch;   // the member's merge scratch
    int64_t *stats;          // host: += collect_merge_stats' counts ([0] n_big_close, [1..7] n-way mergers)
};
// 0: done (*n_merge merged), 1: stopped on a device error bit, -1: HIP failure
int run_merge(const MergeCtx &c, int64_t *n_merge);

// ---- parcel correction (correction.hip); direct launches on s ----
// the FP64 tridiagonal solve of every spectral column and its three gradient
// spectra; fr, fi, et hold (nz + 1) ny (nx/2 + 1) doubles each
void launch_lapinv_columns(const Grid &g, const float2 *in, double *fr, double *fi, double *et, float2 *ox, float2 *oy,
                           float2 *oz, hipStream_t s);
void launch_grad_halo(const Grid &g, float *fld, hipStream_t s);
void launch_laplace_push(const Grid &g, const ParcelsDev &p, const float *fld, int *err, hipStream_t s);
void launch_gradient_push(const Grid &g, const ParcelsDev &p, const float *fld, double prefactor,
                          double max_compression, int *err, hipStream_t s);
int64_t vw_partials(int64_t n);   // doubles of scratch launch_volume_weighted_vorticity needs
// out[0] = sum vol, out[1..3] = sum vol * vorticity (FP64, fixed order)
void launch_volume_weighted_vorticity(const ParcelsDev &p, double *part, double *out, hipStream_t s);
void launch_vortcor(const ParcelsDev &p, const double d[3], hipStream_t s);

// ---- adaptive time step (timestep.hip) ----
struct TsArgs {
    Grid g;
    const float *fld;
    StepTime *tm;
    int *err;
};
// largest strain eigenvalue magnitude and squared buoyancy gradient over the
// nodes, then dt (two launches; nothing happens for a fixed-step item)
void add_time_step(Plan &P, int phase, const std::vector<TsArgs> &items);

// ---- diagnostics (diagnostics.hip); direct launches on s ----
int64_t diag_scratch_doubles(const Grid &g, int64_t cap);
// out[0..11]: sums of vol |u|^2, vol |w|^2, lam, lam^2, vol, vol^2, small,
// vol xi^2, vol eta^2, vol zeta^2; max and min buoyancy
// parcel diagnostics partials (at most max_blocks blocks when positive) and
// totals into out; src (may be null) maps a slot to the slot of its tendencies
void launch_parcel_diagnostics(const Grid &g, const ParcelsDev &p, const int *src, double *scratch, double *out, int *err,
                               hipStream_t s, int max_blocks = 0);
// out[0..9]: sums of (volg - vcell)^2, nparg, nsparg, weighted volg |u|^2,
// volg |w|^2; max |volg - vcell|, nparg, tbuoyg; min nparg, tbuoyg
void launch_field_diagnostics(const Grid &g, const float *fld, const int *nparg, const int *nsparg, double *scratch,
                              double *out, hipStream_t s);

// ---- physical re-sort (resort.hip) ----
// the arrays of a member that outlive a step: positions (8-byte), then the
// 11 attributes, the caller index, the keys and the 6 cached half axes
constexpr int RESORT_W = 3, RESORT_F = 19;
struct ResortArgs {
    const int *perm;
    int64_t n;
    int gather;             // 1: tw/tf[s] = w/f[perm[s]]; 0: w/f[s] = tw/tf[s]
    double *w[RESORT_W];
    float *f[RESORT_F];     // float or int bits
    double *tw[RESORT_W];
    float *tf[RESORT_F];
};
// with_doubles: the RESORT_W double arrays move too; otherwise only the
// RESORT_F float arrays (the stage tendencies: dpos 3, dvor 3, dB 5, strain 8)
void add_resort(Plan &P, int phase, const std::vector<ResortArgs> &items, bool with_doubles = true);
// out[orig[s]] = in[s] for s < n: a per-slot array in caller-index order
struct ToCallerArgs {
    int64_t n;
    const int *orig;
    const int *in;
    int *out;
};
void add_to_caller(Plan &P, int phase, const std::vector<ToCallerArgs> &items);

// ---- utilities (abi.hip) ----
struct FillArgs {
    int *p;
    int64_t n;
    int v;
};
void add_fill_int(Plan &P, int phase, const std::vector<FillArgs> &items);
// moves each item's device error word to out[item] and clears it
struct ErrArgs {
    int *err;
    int *out;
};
void add_gather_errors(Plan &P, const std::vector<ErrArgs> &items);

}  // namespace epic
