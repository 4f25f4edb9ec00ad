// Host-side launchers of the device kernels (internal to the library).
#pragma once

#include "common.h"

namespace epic {

// ---- sort + par2grid (p2g.hip) ----
struct P2GLayout {
    int tz, ntx, nty, nchz;
    int64_t staging_floats;
};
P2GLayout p2g_layout(const Grid &g);
int64_t scan_scratch_ints(int64_t n);
void launch_exclusive_scan(const int *in, int64_t n, int *out, int *bsum, hipStream_t s);
void launch_cell_keys(const Grid &g, const ParcelsDev &p, int *keys, int *count, int *scount, int *err, hipStream_t s);
void launch_place(int64_t n, const int *keys, const int *cell_start, int *cursor, int *perm_tmp, hipStream_t s);
void launch_cell_order(int64_t nc, const int *cell_start, const int *perm_tmp, const int *orig, int *perm,
                       hipStream_t s);
void launch_axes(const Grid &g, const ParcelsDev &p, float *const axes[6], int *err, hipStream_t s);
void launch_p2g(const Grid &g, const ParcelsDev &p, const float *const axes[6], const int *perm,
                const int *cell_start, const int *count, const int *scount, float *staging, float *fld, int *nparg,
                int *nsparg, int *err, hipStream_t s);

// ---- grid2par + push (g2p.hip) ----
void launch_grid2par(const Grid &g, const ParcelsDev &p, const float *fld, int add, int *err, hipStream_t s);
void launch_rk_stage(const Grid &g, const ParcelsDev &p, const float *fld, int stage, int nstages, double cbdt,
                     double ca, int reflect, int store_axes, int *keys, int *count, int *scount, int *err,
                     hipStream_t s);

void launch_rk_stage_tiles(const Grid &g, const ParcelsDev &p, const float *const axes[6], const int *perm,
                           const int *cell_start, const float *fld, int stage, int nstages, double cbdt, double ca,
                           int reflect, int *keys, int *count, int *scount, int *err, hipStream_t s);

// ---- horizontal FFTs (fft.hip) ----
struct FftXArgs {
    Grid g;
    FftPlan plan;          // complex length nx/2
    const float2 *tw;      // e^{-2 pi i t / nx}, t < nx
    int mode;              // forward: 0 field, 1 tendency flux; inverse: 0 store, 2 store + d/dz flux
    int slot;              // field slot read (forward, mode 0) or written (inverse)
    int comp_a, comp_b, sgn;
    const float *fld;      // node fields read
    float *fld_out;        // node fields written (inverse)
    float2 *out;           // forward: T
    const float2 *in;      // inverse: T
};
struct FftYArgs {
    Grid g;
    FftPlan plan;          // length ny
    const float2 *tw;      // e^{-2 pi i t / ny}
    int mode;              // 0 forward, 1 inverse, 2 inverse of i kx in + i ky in2
    const float2 *in, *in2;
    float2 *out;
};
void launch_fft_xf(const FftXArgs &a, hipStream_t s);
void launch_fft_y(const FftYArgs &a, hipStream_t s);
void launch_fft_xi(const FftXArgs &a, hipStream_t s);

// ---- vertical column kernel + fix-ups (inversion.hip) ----
struct ColArgs {
    Grid g;
    FftPlan zplan;         // length 2 nz
    const double2 *ztw;    // e^{-2 pi i t / (2 nz)}
    const float2 *in[3];
    float2 *out[11];       // xi eta zeta u v w dudx dudy dvdy dwdx dwdy
};
size_t col_lds_bytes(const Grid &g);
void launch_vor2vel_columns(const ColArgs &a, hipStream_t s);
void launch_vor2vel_fixup(const Grid &g, float *fld, hipStream_t s);
void launch_vtend_fixup(const Grid &g, float *fld, hipStream_t s);

// ---- utilities (abi.cpp) ----
void launch_fill_int(int *p, int64_t n, int v, hipStream_t s);

}  // namespace epic
