// C-ABI of the parcel-grid core: member state, host <-> device conversion of
// the reference's Fortran layouts, and the per-call orchestration of the
// kernels (the reference's call surface: par2grid, vor2vel,
// vorticity_tendency, grid2par, ls_rk_step).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "epic_gpu.h"
#include "common.h"
#include "fft.h"
#include "launch.h"
#include "plan.h"

using namespace epic;

namespace {

// low-storage RK coefficients (ls_rk.f90:31-54)
const double CAS4[5] = {-567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                        -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0, 0.0};
const double CBS4[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                        1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                        2277821191437.0 / 14882151754819.0};
const double CAS3[3] = {-5.0 / 9.0, -153.0 / 128.0, 0.0};
const double CBS3[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

constexpr int NSPEC = 15;

}  // namespace

struct epic_member {
    Grid g{};
    int rk_order = 4;
    int64_t cap = 0, n = 0;
    hipStream_t s = nullptr;
    // parcels
    double *pos = nullptr;   // 3 cap
    float *attr = nullptr;   // NATTR cap
    int *orig = nullptr;
    float *axbuf = nullptr;            // 12 cap: two sets of half axes
    float *ax_cur[6] = {}, *ax_stale[6] = {};
    ParcelsDev pd{};
    // sort: cnt[0]/scnt[0] hold the histogram of the current positions
    // between calls; inside a step the stages rotate through the three
    int *keys = nullptr, *cnt[3] = {}, *scnt[3] = {};
    int *cell_start = nullptr, *cursor = nullptr, *perm = nullptr, *bsum = nullptr;
    bool keys_valid = false;
    int tz_p2g = 2, tz_g2p = 2;     // z chunk of the tile kernels (autotuned)
    // fields
    float *fld = nullptr;
    int *nparg = nullptr, *nsparg = nullptr;
    double *staging = nullptr;
    // par2grid support points beyond the neighbour cells: 64-bit fixed-point
    // sums [field][plane][j][i], their flag and the fields' inverse scales
    unsigned long long *far_acc = nullptr;
    int *far_flag = nullptr;
    double *far_scale = nullptr;
    // spectra
    float2 *spec[NSPEC] = {};   // NSPEC consecutive spectra of one allocation (spec[0])
    float2 *twx = nullptr, *twy = nullptr;
    double2 *twx64 = nullptr, *twy64 = nullptr;
    double2 *twz = nullptr;
    FftPlan px{}, py{}, pz{};
    // per-phase GPU timers (events on the member stream)
    int profiling = 0;   // 1: eager launches between event pairs; 2: graph replays with wall-clock stamps per phase
    std::vector<hipEvent_t> ev;        // pairs (start, stop) recorded in order
    std::vector<int> ev_phase;
    double timers_ms[EPIC_NUM_TIMERS] = {};
    int timer_calls[EPIC_NUM_TIMERS] = {};
    // errors; flags_d[0] = reuse the cached axes in the first par2grid of a step
    int *err_d = nullptr;
    int *err_h = nullptr;
    int *flags_d = nullptr;
    StepTime *tm_d = nullptr, *tm_h = nullptr;   // the step's time step (device, pinned host copy)
    std::string msg;
    // launch plans of the call surface (built on first use)
    Plan pl_keys, pl_p2g_fresh, pl_p2g_reuse, pl_vor2vel, pl_vtend, pl_g2p[2], pl_step, pl_step_adaptive;
    // the cached half axes (ax_stale) after a change of the parcels that did
    // not recompute them (set_parcels, merge, split): recomputed from the
    // current B before anything reads them (ensure_axes)
    bool axes_dirty = true;
    Plan pl_axes;
    Plan pl_lap[2];   // parcel correction: forward / inverse transforms
    // parcel correction: the target volume-weighted vorticity and the FP64
    // reduction scratch (allocated on first use)
    double vor_bar[3] = {0.0, 0.0, 0.0};
    double *vw_d = nullptr;
    double *vw_h = nullptr;
    // diagnostics: split / merge counts since the last reset, FP64 scratch;
    // collect_merge_stats' counts: [0] n_big_close, [1..7] n-way mergers
    int64_t n_splits = 0, n_merges = 0;
    int64_t merge_stats[8] = {};
    double *diag_d = nullptr, *diag_h = nullptr;
    // physical re-sort into cell order every resort_every steps (0: never)
    int resort_every = 3, steps_since_sort = 0;
    double *sort_tmp = nullptr;
    // the stage tendencies (dpos, dvor, dB, strain) are not moved by a
    // re-sort: while tend_displaced, the tendencies of the parcel in slot s
    // are in slot tend_src[s], until settle_tendencies gathers them
    bool tend_displaced = false;
    int *tend_src = nullptr;
    Plan pl_settle;
    // p2g's fixed-point maxima (P2G_NF + 1)
    float *pmax_out = nullptr;
    Plan pl_resort;
    // the parcel count in device memory (and its pinned host staging word)
    int64_t *n_dev = nullptr, *n_h = nullptr;
    // device scratch of the parcel merge (grown to the largest need seen)
    ScratchArena merge_scratch;
    uint64_t gen = 0;   // bumped whenever plans built over this member go stale
    // the member generation the step plans were built at: a stale step plan
    // is rebuilt, and when the rebuilt launch sequence is the same (a new
    // parcel count, say) only its argument tables are replaced, so the
    // captured graph is kept
    uint64_t step_gen[2] = {0, 0};
    void invalidate_plans()
    {
        ++gen;
        pl_keys.release();
        pl_axes.release();
        pl_p2g_fresh.release();
        pl_p2g_reuse.release();
        pl_vor2vel.release();
        pl_vtend.release();
        pl_g2p[0].release();
        pl_g2p[1].release();
        pl_lap[0].release();
        pl_lap[1].release();
        pl_resort.release();
        pl_settle.release();
    }
    void release_plans()
    {
        invalidate_plans();
        pl_step.release();
        pl_step_adaptive.release();
        pl_axes.release();
    }
    // the parcel count changed (set_parcels, split, merge): the step plans
    // and the keys and axes plans read the count from n_dev and stay valid;
    // the other plans of the single calls hold it in their tables and are
    // rebuilt on their next use
    void counts_changed()
    {
        const uint64_t keep = gen;
        Plan k, a;
        k.swap(pl_keys);
        a.swap(pl_axes);
        invalidate_plans();
        pl_keys.swap(k);
        pl_axes.swap(a);
        gen = keep;
    }
};

namespace {

constexpr int NATTR = 3 + 5 + 3 + 3 + 3 + 5 + 6 + 8;   // vor B vol/buoy/hum dpos dvor dB veta/vtau strain

int fail(epic_member *m, int code, const std::string &what)
{
    if (m) m->msg = what;
    return code;
}

#define HIPOK(m, expr)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return fail(m, EPIC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// zero-filled device array; the fill is queued on the owner's stream (the
// streams are non-blocking, so nothing here waits for other members' work)
template <typename T> hipError_t dalloc(T **p, int64_t count, hipStream_t s)
{
    *p = nullptr;
    if (count <= 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * (size_t)count);
    if (e == hipSuccess) e = hipMemsetAsync(*p, 0, sizeof(T) * (size_t)count, s);
    return e;
}

// fn(lo, hi) over [0, n) on up to 16 host threads (the per-parcel layout
// conversions of set/get; each element is written by one thread)
template <class F> void host_parallel_for(int64_t n, F fn)
{
    const int64_t grain = 1 << 16;
    unsigned hw = std::thread::hardware_concurrency();
    int nt = (int)std::min<int64_t>(std::min<unsigned>(hw ? hw : 1, 16u), (n + grain - 1) / grain);
    if (nt <= 1) {
        fn((int64_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int t = 0; t < nt; ++t) {
        const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
        th.emplace_back([=, &fn]() { fn(lo, hi); });
    }
    for (std::thread &x : th) x.join();
}

// lengths the Stockham stages cover: 2^a 3^b 5^c (stafft.f90:128-190
// factorises 2, 3, 4, 5, 6)
bool factor235(int n)
{
    if (n < 1) return false;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    while (n % 5 == 0) n /= 5;
    return n == 1;
}

// the parcel views of the member's arrays and its count, also into device
// memory (n_dev), which the step plans' kernels read
hipError_t bind_parcels(epic_member *m)
{
    ParcelsDev &p = m->pd;
    const int64_t c = m->cap;
    p.n = m->n;
    p.cap = c;
    p.n_dev = m->n_dev;
    p.x = m->pos;
    p.y = m->pos + c;
    p.z = m->pos + 2 * c;
    float *a = m->attr;
    for (int i = 0; i < 3; ++i) p.vor[i] = a + (0 + i) * c;
    for (int i = 0; i < 5; ++i) p.B[i] = a + (3 + i) * c;
    p.vol = a + 8 * c;
    p.buoy = a + 9 * c;
    p.hum = a + 10 * c;
    for (int i = 0; i < 3; ++i) p.dpos[i] = a + (11 + i) * c;
    for (int i = 0; i < 3; ++i) p.dvor[i] = a + (14 + i) * c;
    for (int i = 0; i < 5; ++i) p.dB[i] = a + (17 + i) * c;
    for (int i = 0; i < 3; ++i) p.veta[i] = a + (22 + i) * c;
    for (int i = 0; i < 3; ++i) p.vtau[i] = a + (25 + i) * c;
    for (int i = 0; i < 8; ++i) p.strain[i] = a + (28 + i) * c;
    p.orig = m->orig;
    // the stale set is the reference's Vetas/Vtaus cache (par2grid l_reuse)
    for (int i = 0; i < 3; ++i) {
        p.veta[i] = m->ax_stale[i];
        p.vtau[i] = m->ax_stale[3 + i];
    }
    *m->n_h = m->n;
    hipError_t e = hipMemcpyAsync(m->n_dev, m->n_h, sizeof(int64_t), hipMemcpyHostToDevice, m->s);
    if (e == hipSuccess) e = hipStreamSynchronize(m->s);
    return e;
}

void collect_timers(epic_member *m);

// maps the device error bits to a status code and message
int error_code(epic_member *m, int e, const char *where)
{
    if (e == 0) return EPIC_OK;
    std::string w(where);
    if (e & ERR_PARCEL_COUNT) return fail(m, EPIC_E_PARCEL_COUNT, w + ": par2grid: Wrong total number of parcels!");
    if (e & ERR_B33) return fail(m, EPIC_E_B33, w + ": get_B33: Division by small number!");
    if (e & ERR_OUT_OF_GRID) return fail(m, EPIC_E_OUT_OF_GRID, w + ": support point outside the node grid");
    if (e & ERR_MERGE) return fail(m, EPIC_E_MERGE, w + ": Merge error: no near enough neighbour found.");
    if (e & ERR_DT) return fail(m, EPIC_E_DT, w + ": Time step <= 0!");
    return fail(m, EPIC_E_INVALID, w + ": internal consistency error");
}

// drain the device error word; maps the first error bit to a code
int check(epic_member *m, const char *where)
{
    HIPOK(m, hipMemcpyAsync(m->err_h, m->err_d, sizeof(int), hipMemcpyDeviceToHost, m->s));
    HIPOK(m, hipStreamSynchronize(m->s));
    HIPOK(m, hipGetLastError());
    const int e = *m->err_h;
    collect_timers(m);
    if (e == 0) return EPIC_OK;
    HIPOK(m, hipMemsetAsync(m->err_d, 0, sizeof(int), m->s));
    HIPOK(m, hipStreamSynchronize(m->s));
    return error_code(m, e, where);
}

__global__ __launch_bounds__(256) void k_fill_int(const FillArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    const FillArgs &a = tab[ib.item];
#pragma unroll
    for (int r = 0; r < PER_BLOCK / 256; ++r) {
        const int64_t i = (int64_t)ib.blk * PER_BLOCK + r * 256 + threadIdx.x;
        if (i < a.n) a.p[i] = a.v;
    }
}

__global__ void k_gather_errors(const ErrArgs *__restrict__ tab, const int2 *__restrict__ bmap)
{
    const ItemBlock ib = item_block(bmap);
    if (threadIdx.x == 0) {
        const ErrArgs &a = tab[ib.item];
        *a.out = *a.err;
        *a.err = 0;
    }
}

// nearest-cell key of parcel_nearest.f90:168-221 (single rank)
__global__ void k_nearest_keys(Grid g, ParcelsDev p, int *keys, int *count, int64_t ncell, int *err)
{
    int e = 0;
    const int hx = g.nx + 3, hy = g.ny + 3;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < p.n; q += (int64_t)gridDim.x * blockDim.x) {
        double x = p.x[q], y = p.y[q];
        const double z = p.z[q];
        const int gx = (int)floor((x - g.lower[0]) * g.dxi[0]);
        const int gy = (int)floor((y - g.lower[1]) * g.dxi[1]);
        if (gx < -1) x = x + g.extent[0];
        if (gx > g.nx) x = x - g.extent[0];
        if (gy < -1) y = y + g.extent[1];
        if (gy > g.ny) y = y - g.extent[1];
        const int ix = (int)(g.dxi[0] * (x - (g.lower[0] - g.dx[0])));
        const int iy = (int)(g.dxi[1] * (y - (g.lower[1] - g.dx[1])));
        const int iz = min((int)(g.dxi[2] * (z - g.lower[2])), g.nz - 1);
        const int64_t ijk = (int64_t)ix + (int64_t)hx * iy + (int64_t)hx * hy * iz;   // 0-based
        if (ijk < 0 || ijk >= ncell) {
            e |= ERR_OUT_OF_GRID;
            keys[q] = -1;
            continue;
        }
        keys[q] = (int)ijk;
        atomicAdd(&count[ijk], 1);
    }
    if (e) atomicOr(err, e);
}

int grid_blocks(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 8192); }

// after a stream synchronisation: fold the recorded event pairs into the
// public timers (a par2grid is its cell sort plus the tile scatter)
static void add_phase_time(epic_member *m, int phase, float ms)
{
            switch (phase) {
            case PH_AXES: m->timers_ms[EPIC_TIMER_AXES] += ms; m->timer_calls[EPIC_TIMER_AXES] += 1; break;
            case PH_SORT:
                m->timers_ms[EPIC_TIMER_CELL_SORT] += ms;
                m->timer_calls[EPIC_TIMER_CELL_SORT] += 1;
                m->timers_ms[EPIC_TIMER_PAR2GRID] += ms;
                break;
            case PH_P2G: m->timers_ms[EPIC_TIMER_PAR2GRID] += ms; m->timer_calls[EPIC_TIMER_PAR2GRID] += 1; break;
            case PH_P2G_TILES:
                m->timers_ms[EPIC_TIMER_P2G_TILES] += ms;
                m->timer_calls[EPIC_TIMER_P2G_TILES] += 1;
                m->timers_ms[EPIC_TIMER_PAR2GRID] += ms;
                break;
            case PH_VOR2VEL: m->timers_ms[EPIC_TIMER_VOR2VEL] += ms; m->timer_calls[EPIC_TIMER_VOR2VEL] += 1; break;
            case PH_VTEND: m->timers_ms[EPIC_TIMER_VTEND] += ms; m->timer_calls[EPIC_TIMER_VTEND] += 1; break;
            case PH_PUSH: m->timers_ms[EPIC_TIMER_PUSH] += ms; m->timer_calls[EPIC_TIMER_PUSH] += 1; break;
            case PH_GRID2PAR: m->timers_ms[EPIC_TIMER_GRID2PAR] += ms; m->timer_calls[EPIC_TIMER_GRID2PAR] += 1; break;
            default: break;
            }
}

void collect_timers(epic_member *m)
{
    for (size_t k = 0; k < m->ev_phase.size(); ++k) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, m->ev[2 * k], m->ev[2 * k + 1]) == hipSuccess) add_phase_time(m, m->ev_phase[k], ms);
        (void)hipEventDestroy(m->ev[2 * k]);
        (void)hipEventDestroy(m->ev[2 * k + 1]);
    }
    m->ev.clear();
    m->ev_phase.clear();
}

// after a synchronised replay of a graph captured with timers
static int collect_graph_timers(epic_member *m, Plan &P)
{
    std::vector<std::pair<int, float>> t;
    HIPOK(m, P.graph_phase_times(t));
    for (const auto &pt : t) add_phase_time(m, pt.first, pt.second);
    return EPIC_OK;
}

// ---------------------------------------------------------------------------
// plan builders over a set of members (one member: the plain call surface;
// many: the batched step).  Every builder appends grouped launches.
// ---------------------------------------------------------------------------
using Members = std::vector<epic_member *>;

P2GLayout layout_tz(const Grid &g, int tz)
{
    P2GLayout L = p2g_layout(g);
    L.tz = tz;
    L.nchz = (g.nz + 2 + tz - 1) / tz;
    L.staging_elems = (int64_t)L.ntx * L.nty * L.nchz * (tz + 3) * P2G_BOX * P2G_BOX * P2G_NF;
    return L;
}

// z chunks the autotuner may choose (par2grid: 64 tz P2G_TPC threads per work-group;
// both stage a node box of tz + 3 planes in LDS); the staging buffer fits
// every par2grid candidate
const int TZ_P2G[3] = {2, 4, 8};
const int TZ_G2P[3] = {2, 4, 8};

int64_t staging_elems_max(const Grid &g)
{
    int64_t mx = 0;
    for (int tz : TZ_P2G) mx = std::max(mx, layout_tz(g, tz).staging_elems);
    return mx;
}

void build_keys(Plan &P, const Members &ms)
{
    std::vector<FillArgs> f;
    std::vector<KeysArgs> k;
    for (epic_member *m : ms) {
        const int64_t nc = ncells(m->g);
        f.push_back(FillArgs{m->cnt[0], nc, 0});
        f.push_back(FillArgs{m->scnt[0], nc, 0});
        k.push_back(KeysArgs{m->g, m->pd, m->keys, m->cnt[0], m->scnt[0], m->err_d});
    }
    add_fill_int(P, PH_SORT, f);
    add_cell_keys(P, PH_SORT, k);
}

void build_axes(Plan &P, const Members &ms, const std::vector<float *const *> &ax)
{
    std::vector<AxesArgs> v;
    for (size_t i = 0; i < ms.size(); ++i) {
        AxesArgs a{};
        a.p = ms[i]->pd;
        for (int c = 0; c < 6; ++c) a.ax[c] = ax[i][c];
        a.err = ms[i]->err_d;
        v.push_back(a);
    }
    add_axes(P, PH_AXES, v);
}

// sort by the histogram rd, scatter with the axes ax (ax_alt when *use_alt),
// finalise; clears the histogram pair clr for the next push (may be null)
struct P2GSel {
    float *const *ax;
    float *const *ax_alt;
    const int *use_alt;
    int rd;
    int clr;                  // -1: none
    bool max_from_push = false;   // the previous push stored the fixed-point maxima: no k_p2g_max
};

void build_p2g(Plan &P, const Members &ms, const std::vector<P2GSel> &sel)
{
    std::vector<ScanArgs> sc;
    std::vector<PlaceArgs> pl;
    std::vector<P2GArgs> ti;
    std::vector<FinalizeArgs> fi;
    std::vector<FarClearArgs> far;
    for (size_t i = 0; i < ms.size(); ++i) {
        epic_member *m = ms[i];
        const P2GSel &q = sel[i];
        const int64_t nc = ncells(m->g);
        sc.push_back(ScanArgs{m->cnt[q.rd], nc, m->cell_start, m->bsum, m->cursor});
        pl.push_back(PlaceArgs{m->n, m->keys, m->cell_start, m->cursor, m->perm, m->cap, m->n_dev});
        const P2GLayout L = layout_tz(m->g, m->tz_p2g);
        P2GArgs a{};
        a.g = m->g;
        a.p = m->pd;
        for (int c = 0; c < 6; ++c) {
            a.axes[c] = q.ax[c];
            a.axes_alt[c] = q.ax_alt ? q.ax_alt[c] : q.ax[c];
        }
        a.use_alt = q.use_alt;
        a.perm = m->perm;
        a.cell_start = m->cell_start;
        a.staging = m->staging;
        a.err = m->err_d;
        a.tmax = m->pmax_out;
        a.far_acc = m->far_acc;
        a.far_flag = m->far_flag;
        a.far_scale = m->far_scale;
        a.tz = L.tz;
        a.ntx = L.ntx;
        a.nty = L.nty;
        a.nchz = L.nchz;
        ti.push_back(a);
        FinalizeArgs f{};
        f.g = m->g;
        f.staging = m->staging;
        f.count = m->cnt[q.rd];
        f.scount = m->scnt[q.rd];
        f.fld = m->fld;
        f.nparg = m->nparg;
        f.nsparg = m->nsparg;
        f.clear_count = q.clr >= 0 ? m->cnt[q.clr] : nullptr;
        f.clear_scount = q.clr >= 0 ? m->scnt[q.clr] : nullptr;
        f.clear_tmax = m->pmax_out;
        f.far_acc = m->far_acc;
        f.far_flag = m->far_flag;
        f.far_scale = m->far_scale;
        far.push_back(FarClearArgs{m->far_acc, m->far_flag, (int64_t)P2G_NF * field_stride(m->g)});
        f.tz = L.tz;
        f.ntx = L.ntx;
        f.nty = L.nty;
        f.nchz = L.nchz;
        fi.push_back(f);
    }
    std::vector<P2GMaxArgs> mx;
    for (size_t i = 0; i < ms.size(); ++i)
        if (!sel[i].max_from_push) mx.push_back(P2GMaxArgs{ms[i]->g, ms[i]->pd, p2g_max_blocks(ms[i]->cap), ms[i]->pmax_out});
    if (!mx.empty()) add_p2g_max(P, PH_P2G, mx);
    add_exclusive_scan(P, PH_SORT, sc);
    add_place(P, PH_SORT, pl);
    add_p2g(P, PH_P2G, ti, fi, far);
}

FftXArgs xargs(epic_member *m)
{
    FftXArgs a{};
    a.g = m->g;
    a.plan = m->px;
    a.tw = m->twx;
    a.tw64 = m->twx64;
    a.fld = m->fld;
    a.fld_out = m->fld;
    return a;
}

FftYArgs yargs(epic_member *m)
{
    FftYArgs a{};
    a.g = m->g;
    a.plan = m->py;
    a.tw = m->twy;
    a.tw64 = m->twy64;
    return a;
}

// vor2vel + vel2vgrad: 3 forward transforms, the column solve, 11 inverse
// transforms, the physical-space completion
void build_vor2vel(Plan &P, const Members &ms)
{
    static const int slots[11] = {F_VORTG + 0, F_VORTG + 1, F_VORTG + 2, F_VELOG + 0, F_VELOG + 1, F_VELOG + 2,
                                  F_VELGRADG + DUDX, F_VELGRADG + DUDY, F_VELGRADG + DVDY, F_VELGRADG + DWDX,
                                  F_VELGRADG + DWDY};
    std::vector<FftXArgs> xf, xi;
    std::vector<FftYArgs> yf, yi;
    std::vector<ColArgs> co;
    std::vector<FixArgs> fx;
    for (epic_member *m : ms) {
        for (int c = 0; c < 3; ++c) {
            FftXArgs x = xargs(m);
            x.mode = 0;
            x.slot = F_VORTG + c;
            x.out = m->spec[c];
            xf.push_back(x);
            FftYArgs y = yargs(m);
            y.mode = 0;
            y.in = m->spec[c];
            y.out = m->spec[c];
            yf.push_back(y);
        }
        ColArgs ca{};
        ca.g = m->g;
        ca.zplan = m->pz;
        ca.ztw = m->twz;
        for (int c = 0; c < 3; ++c) ca.in[c] = m->spec[c];
        for (int o = 0; o < 11; ++o) ca.out[o] = m->spec[3 + o];
        co.push_back(ca);
        for (int o = 0; o < 11; ++o) {
            FftYArgs y = yargs(m);
            y.mode = 1;
            y.in = m->spec[3 + o];
            y.out = m->spec[3 + o];
            yi.push_back(y);
            FftXArgs x = xargs(m);
            x.mode = 0;
            x.slot = slots[o];
            x.in = m->spec[3 + o];
            xi.push_back(x);
        }
        fx.push_back(FixArgs{m->g, m->fld});
    }
    add_fft_xf(P, PH_VOR2VEL, xf);
    add_fft_y(P, PH_VOR2VEL, yf);
    add_vor2vel_columns(P, PH_VOR2VEL, co);
    add_fft_y(P, PH_VOR2VEL, yi);
    add_fft_xi(P, PH_VOR2VEL, xi);
    add_vor2vel_fixup(P, PH_VOR2VEL, fx);
}

// vorticity_tendency: six flux transforms, two combined inverse transforms
// per component (the z derivative is added in the last pass), all in FP64
// from the FP32 node fields: the spectral derivative amplifies white
// rounding noise at the highest wavenumbers, which FP32 products and
// transforms would add.  The FP64 spectrum q takes the float2 spectra 2q and
// 2q + 1 (one allocation).
void build_vtend(Plan &P, const Members &ms)
{
    std::vector<FftXArgs> xf, xi;
    std::vector<FftYArgs> yf, yi;
    std::vector<FixArgs> fx;
    for (epic_member *m : ms) {
        auto sp64 = [&](int q) { return (void *)m->spec[2 * q]; };
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 2; ++a) {
                FftXArgs x = xargs(m);
                x.fp64 = 1;
                x.mode = 1;
                x.comp_a = a;
                x.comp_b = c;
                // f_x of the y tendency carries -b, f_y of the x tendency +b (inversion.f90:327-335)
                x.sgn = (a == 0 && c == 1) ? -1 : ((a == 1 && c == 0) ? 1 : 0);
                x.out = sp64(2 * c + a);
                xf.push_back(x);
                FftYArgs y = yargs(m);
                y.fp64 = 1;
                y.mode = 0;
                y.in = sp64(2 * c + a);
                y.out = sp64(2 * c + a);
                yf.push_back(y);
            }
        for (int c = 0; c < 3; ++c) {
            FftYArgs y = yargs(m);
            y.fp64 = 1;
            y.mode = 2;
            y.in = sp64(2 * c);
            y.in2 = sp64(2 * c + 1);
            y.out = sp64(2 * c);
            yi.push_back(y);
            FftXArgs x = xargs(m);
            x.fp64 = 1;
            x.mode = 2;
            x.comp_b = c;
            x.slot = F_VTEND + c;
            x.in = sp64(2 * c);
            xi.push_back(x);
        }
        fx.push_back(FixArgs{m->g, m->fld});
    }
    add_fft_xf(P, PH_VTEND, xf);
    add_fft_y(P, PH_VTEND, yf);
    add_fft_y(P, PH_VTEND, yi);
    add_fft_xi(P, PH_VTEND, xi);
    add_vtend_fixup(P, PH_VTEND, fx);
}

// one LS-RK step (ls_rk.f90:88-136) of every member.  Histograms: stage s
// sorts by cnt[(s-1)%3] and its push fills cnt[s%3], the last push cnt[0]
// (never the one the stage read, so par2grid can clear it in parallel).
// Half axes: stages 1..ns-1 use ax_cur, the last stage ax_stale, which is
// then the cache the next step's first par2grid takes when *reuse is set.
// Stage 1 computes ax_cur from B; the push of stage s < ns stores the axes of
// the B it writes, which are stage s+1's (ax_cur, or ax_stale for the last).
struct StepItem {
    epic_member *m;
    StepTime *tm;        // device: dt (written by the host, or by the adaptive time step)
    const int *reuse;
    bool adaptive;
};

void build_step(Plan &P, const std::vector<StepItem> &items)
{
    int maxns = 0;
    for (const StepItem &it : items) maxns = std::max(maxns, it.m->rk_order == 4 ? 5 : 3);
    for (int s = 1; s <= maxns; ++s) {
        Members ms;
        std::vector<const StepItem *> act;
        for (const StepItem &it : items)
            if ((it.m->rk_order == 4 ? 5 : 3) >= s) {
                ms.push_back(it.m);
                act.push_back(&it);
            }
        std::vector<float *const *> ax;
        std::vector<P2GSel> sel;
        for (const StepItem *it : act) {
            epic_member *m = it->m;
            const int ns = (m->rk_order == 4) ? 5 : 3;
            const bool last = (s == ns);
            float *const *a = last ? m->ax_stale : m->ax_cur;
            ax.push_back(a);
            const int rd = (s - 1) % 3, wr = last ? 0 : (s % 3);
            sel.push_back(P2GSel{a, m->ax_stale, s == 1 ? it->reuse : nullptr, rd, wr, s > 1});
        }
        if (s == 1) build_axes(P, ms, ax);   // later stages: stored by the previous push
        build_p2g(P, ms, sel);
        build_vor2vel(P, ms);
        build_vtend(P, ms);
        if (s == 1) {   // get_time_step after the first vorticity_tendency (ls_rk.f90:92-101)
            std::vector<TsArgs> ts;
            for (const StepItem *it : act)
                if (it->adaptive) ts.push_back(TsArgs{it->m->g, it->m->fld, it->tm, it->m->err_d});
            if (!ts.empty()) add_time_step(P, PH_OTHER, ts);
        }
        std::vector<G2PTileArgs> gv;
        for (size_t i = 0; i < act.size(); ++i) {
            epic_member *m = act[i]->m;
            const int ns = (m->rk_order == 4) ? 5 : 3;
            const double *ca = (m->rk_order == 4) ? CAS4 : CAS3;
            const double *cb = (m->rk_order == 4) ? CBS4 : CBS3;
            const bool last = (s == ns);
            const int wr = last ? 0 : (s % 3);
            const P2GLayout L = layout_tz(m->g, m->tz_g2p);
            G2PTileArgs a{};
            a.g = m->g;
            a.p = m->pd;
            for (int c = 0; c < 6; ++c) a.axes[c] = ax[i][c];
            if (!last) {
                for (int c = 0; c < 6; ++c) a.axes_next[c] = (s + 1 == ns) ? m->ax_stale[c] : m->ax_cur[c];
                a.tmax_next = m->pmax_out;   // cleared by this stage's finalize
            }
            a.perm = m->perm;
            a.cell_start = m->cell_start;
            a.fld = m->fld;
            a.stage = s;
            a.nstages = ns;
            a.reflect = last ? 1 : 0;
            a.cb = cb[s - 1];
            a.dt = &act[i]->tm->dt;
            a.ca = ca[s - 1];
            a.keys = m->keys;
            a.count = m->cnt[wr];
            a.scount = m->scnt[wr];
            a.err = m->err_d;
            a.tz = L.tz;
            a.ntx = L.ntx;
            a.nty = L.nty;
            a.nchz = L.nchz;
            gv.push_back(a);
        }
        add_rk_stage_tiles(P, PH_PUSH, gv);
    }
}

// parcel correction transforms: phi = volg / vcell - 1 to its spectrum
// (forward), the three gradient spectra back to F_GRAD (inverse)
void build_laplace(epic_member *m)
{
    FftXArgs x = xargs(m);
    x.mode = 3;
    x.out = m->spec[0];
    add_fft_xf(m->pl_lap[0], PH_OTHER, {x});
    FftYArgs y = yargs(m);
    y.mode = 0;
    y.in = m->spec[0];
    y.out = m->spec[0];
    add_fft_y(m->pl_lap[0], PH_OTHER, {y});
    std::vector<FftYArgs> yi;
    std::vector<FftXArgs> xi;
    for (int c = 0; c < 3; ++c) {
        FftYArgs b = yargs(m);
        b.mode = 1;
        b.in = m->spec[1 + c];
        b.out = m->spec[1 + c];
        yi.push_back(b);
        FftXArgs a = xargs(m);
        a.mode = 0;
        a.slot = F_GRAD + c;
        a.in = m->spec[1 + c];
        xi.push_back(a);
    }
    add_fft_y(m->pl_lap[1], PH_OTHER, yi);
    add_fft_xi(m->pl_lap[1], PH_OTHER, xi);
}

// enqueue a member plan on the member stream (event timers when profiling)
int run_plan(epic_member *m, Plan &P)
{
    hipError_t e = m->profiling ? P.run(m->s, &m->ev, &m->ev_phase) : P.run(m->s);
    if (e != hipSuccess) return fail(m, EPIC_E_HIP, std::string("launch: ") + hipGetErrorString(e));
    return EPIC_OK;
}

int ensure_keys(epic_member *m)
{
    if (m->keys_valid) return EPIC_OK;
    if (m->pl_keys.empty()) build_keys(m->pl_keys, {m});
    const int rc = run_plan(m, m->pl_keys);
    if (rc == EPIC_OK) m->keys_valid = true;
    return rc;
}

// the cached half axes from the current B, when stale (enqueued on the
// member stream)
int ensure_axes(epic_member *m)
{
    if (!m->axes_dirty) return EPIC_OK;
    if (m->pl_axes.empty()) build_axes(m->pl_axes, {m}, {m->ax_stale});
    const int rc = run_plan(m, m->pl_axes);
    if (rc == EPIC_OK) m->axes_dirty = false;
    return rc;
}

int field_ncomp(int which)
{
    switch (which) {
    case EPIC_FIELD_VORTG:
    case EPIC_FIELD_VELOG:
    case EPIC_FIELD_VTEND: return 3;
    case EPIC_FIELD_VELGRADG: return 8;
    case EPIC_FIELD_VOLG:
    case EPIC_FIELD_TBUOYG:
    case EPIC_FIELD_DBUOYG:
    case EPIC_FIELD_HUMG:
    case EPIC_FIELD_NPARG:
    case EPIC_FIELD_NSPARG: return 1;
    default: return 0;
    }
}

int field_slot(int which)
{
    switch (which) {
    case EPIC_FIELD_VOLG: return F_VOLG;
    case EPIC_FIELD_VORTG: return F_VORTG;
    case EPIC_FIELD_TBUOYG: return F_TBUOYG;
    case EPIC_FIELD_DBUOYG: return F_DBUOYG;
    case EPIC_FIELD_HUMG: return F_HUMG;
    case EPIC_FIELD_VELOG: return F_VELOG;
    case EPIC_FIELD_VELGRADG: return F_VELGRADG;
    case EPIC_FIELD_VTEND: return F_VTEND;
    default: return -1;
    }
}

template <typename T> int copy_field_out(epic_member *m, const T *dev, int ncomp, double *out)
{
    const Grid &g = m->g;
    const int64_t FS = field_stride(g);
    std::vector<T> h((size_t)(FS * ncomp));
    HIPOK(m, hipMemcpyAsync(h.data(), dev, sizeof(T) * h.size(), hipMemcpyDeviceToHost, m->s));
    HIPOK(m, hipStreamSynchronize(m->s));
    const int64_t H = (int64_t)(g.nz + 3) * (g.ny + 3) * (g.nx + 3);
    for (int c = 0; c < ncomp; ++c)
        for (int ix = -1; ix <= g.nx + 1; ++ix)
            for (int iy = -1; iy <= g.ny + 1; ++iy) {
                const int wx = (ix + g.nx) % g.nx, wy = (iy + g.ny) % g.ny;
                double *o = out + c * H + ((int64_t)(ix + 1) * (g.ny + 3) + (iy + 1)) * (g.nz + 3);
                const T *src = h.data() + c * FS + (int64_t)wy * g.nx + wx;
                for (int iz = -1; iz <= g.nz + 1; ++iz) o[iz + 1] = (double)src[(int64_t)(iz + 1) * g.ny * g.nx];
            }
    return EPIC_OK;
}

}  // namespace

namespace epic {
void add_fill_int(Plan &P, int phase, const std::vector<FillArgs> &items)
{
    std::vector<int> nb;
    for (const FillArgs &a : items) nb.push_back((int)((a.n + PER_BLOCK - 1) / PER_BLOCK));
    P.add((const void *)k_fill_int, 256, 0, phase, items, nb);
}

void add_gather_errors(Plan &P, const std::vector<ErrArgs> &items)
{
    P.add((const void *)k_gather_errors, 64, 0, PH_OTHER, items, std::vector<int>(items.size(), 1));
}
}  // namespace epic

// ===========================================================================
// exported entry points
// ===========================================================================
extern "C" {

static int settle_tendencies(epic_member *m);

int epic_create(const epic_grid_t *grid, const epic_phys_t *phys, const epic_flags_t *flags, int64_t max_parcels,
                epic_handle *out)
{
    if (!grid || !phys || !flags || !out || max_parcels < 0) return EPIC_E_INVALID;
    *out = nullptr;
    if (grid->nx < 8 || grid->ny < 8 || grid->nx % 2 || grid->nz < 4) return EPIC_E_INVALID;
    if (!factor235(grid->nx / 2) || !factor235(grid->ny) || !factor235(2 * grid->nz)) return EPIC_E_INVALID;
    if (grid->nx / 2 > FFT_MAX_N || grid->ny > FFT_MAX_N || 2 * grid->nz > FFT_MAX_N) return EPIC_E_INVALID;
    if (flags->rk_order != 3 && flags->rk_order != 4) return EPIC_E_RK_ORDER;
    for (int i = 0; i < 3; ++i)
        if (!(grid->extent[i] > 0.0)) return EPIC_E_INVALID;
    if (max_parcels > EPIC_MAX_PARCELS) return EPIC_E_INVALID;   // 32-bit byte offsets of the 4-byte arrays

    epic_member *m = new epic_member();
    Grid &g = m->g;
    g.nx = grid->nx;
    g.ny = grid->ny;
    g.nz = grid->nz;
    g.nxh = g.nx / 2 + 1;
    g.nzp = ((g.nz + 3 + ZB - 1) / ZB) * ZB;   // planes 0..nz and two halo planes
    g.zeta_zero[0] = flags->zeta_bndry_zero[0] != 0;
    g.zeta_zero[1] = flags->zeta_bndry_zero[1] != 0;
    const int n3[3] = {g.nx, g.ny, g.nz};
    double vc = 1.0;
    for (int i = 0; i < 3; ++i) {   // parameters.f90:100-162
        g.lower[i] = grid->lower[i];
        g.extent[i] = grid->extent[i];
        g.upper[i] = g.lower[i] + g.extent[i];
        g.dx[i] = g.extent[i] / (double)n3[i];
        g.dxi[i] = 1.0 / g.dx[i];
        g.center[i] = 0.5 * (g.lower[i] + g.upper[i]);
        g.hli[i] = 1.0 / (g.extent[i] / 2.0);
        vc *= g.dx[i];
    }
    g.vcell = vc;
    g.vmin = flags->vmin > 0.0 ? flags->vmin : g.vcell / 20.0;
    g.glat = phys->glat;
    g.lambda_c = phys->lambda_c;
    g.q0 = phys->q0;
    for (int i = 0; i < 3; ++i) g.f_cor[i] = phys->f_cor[i];
    m->rk_order = flags->rk_order;
    m->cap = std::max<int64_t>(max_parcels, 1);
    {   // a transform work-group holds ZB lines (or a column) in LDS
        int dev = 0, lds_max = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) {
            epic_destroy(m);
            return EPIC_E_HIP;
        }
        if (transform_lds_bytes(g) > (size_t)lds_max) {
            epic_destroy(m);
            return EPIC_E_INVALID;
        }
    }

    auto bad = [&](hipError_t e) {
        if (e != hipSuccess) {
            epic_destroy(m);
            return true;
        }
        return false;
    };
    if (bad(hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking))) return EPIC_E_HIP;
    const int64_t c = m->cap, nc = ncells(g);
    if (bad(dalloc(&m->pos, 3 * c, m->s)) || bad(dalloc(&m->attr, (int64_t)NATTR * c, m->s)) || bad(dalloc(&m->orig, c, m->s)) ||
        bad(dalloc(&m->axbuf, 12 * c, m->s)) || bad(dalloc(&m->keys, c, m->s)) || bad(dalloc(&m->perm, c + 1, m->s)) ||
        bad(dalloc(&m->cnt[0], nc, m->s)) || bad(dalloc(&m->scnt[0], nc, m->s)) || bad(dalloc(&m->cnt[1], nc, m->s)) ||
        bad(dalloc(&m->scnt[1], nc, m->s)) || bad(dalloc(&m->cnt[2], nc, m->s)) || bad(dalloc(&m->scnt[2], nc, m->s)) ||
        bad(dalloc(&m->cell_start, nc + 1, m->s)) || bad(dalloc(&m->cursor, nc, m->s)) ||
        bad(dalloc(&m->bsum, scan_scratch_ints(std::max<int64_t>(std::max<int64_t>(nc, c),
                                                                  (int64_t)(g.nx + 3) * (g.ny + 3) * (g.nz + 2))), m->s)))
        return EPIC_E_NOMEM;
    const int64_t FS = field_stride(g);
    if (bad(dalloc(&m->fld, (int64_t)F_COUNT * FS, m->s)) || bad(dalloc(&m->nparg, FS, m->s)) || bad(dalloc(&m->nsparg, FS, m->s)) ||
        bad(dalloc(&m->staging, staging_elems_max(g), m->s)) ||
        bad(dalloc(&m->far_acc, (int64_t)P2G_NF * FS, m->s)) || bad(dalloc(&m->far_flag, 1, m->s)) ||
        bad(dalloc(&m->far_scale, P2G_NF, m->s)))
        return EPIC_E_NOMEM;
    static_assert(NSPEC >= 12, "the six FP64 spectra of vorticity_tendency take twelve float2 spectra");
    if (bad(dalloc(&m->spec[0], NSPEC * spec_size(g), m->s))) return EPIC_E_NOMEM;
    for (int i = 1; i < NSPEC; ++i) m->spec[i] = m->spec[0] + i * spec_size(g);
    if (bad(dalloc(&m->err_d, 4, m->s)) || bad(dalloc(&m->flags_d, 4, m->s)) || bad(dalloc(&m->tm_d, 1, m->s))) return EPIC_E_NOMEM;
    if (bad(dalloc(&m->pmax_out, P2G_NF + 1, m->s))) return EPIC_E_NOMEM;
    if (bad(hipHostMalloc((void **)&m->err_h, sizeof(int) * 4, hipHostMallocDefault)) ||
        bad(hipHostMalloc((void **)&m->tm_h, sizeof(StepTime), hipHostMallocDefault)))
        return EPIC_E_NOMEM;
    *m->tm_h = StepTime{};
    // twiddles (FP64-accurate roots of unity)
    m->px = make_plan(g.nx / 2);
    m->py = make_plan(g.ny);
    m->pz = make_plan(2 * g.nz);
    std::vector<float2> tx(g.nx), ty(g.ny);
    std::vector<double2> tz(2 * g.nz);
    const double tp = 2.0 * 3.14159265358979323846;
    for (int t = 0; t < g.nx; ++t) tx[t] = float2{(float)std::cos(tp * t / g.nx), (float)-std::sin(tp * t / g.nx)};
    for (int t = 0; t < g.ny; ++t) ty[t] = float2{(float)std::cos(tp * t / g.ny), (float)-std::sin(tp * t / g.ny)};
    for (int t = 0; t < 2 * g.nz; ++t) tz[t] = double2{std::cos(tp * t / (2 * g.nz)), -std::sin(tp * t / (2 * g.nz))};
    std::vector<double2> tx64(g.nx), ty64(g.ny);
    for (int t = 0; t < g.nx; ++t) tx64[t] = double2{std::cos(tp * t / g.nx), -std::sin(tp * t / g.nx)};
    for (int t = 0; t < g.ny; ++t) ty64[t] = double2{std::cos(tp * t / g.ny), -std::sin(tp * t / g.ny)};
    if (bad(dalloc(&m->twx, g.nx, m->s)) || bad(dalloc(&m->twy, g.ny, m->s)) || bad(dalloc(&m->twz, 2 * g.nz, m->s)) ||
        bad(dalloc(&m->twx64, g.nx, m->s)) || bad(dalloc(&m->twy64, g.ny, m->s)))
        return EPIC_E_NOMEM;
    if (bad(hipMemcpyAsync(m->twx, tx.data(), sizeof(float2) * tx.size(), hipMemcpyHostToDevice, m->s)) ||
        bad(hipMemcpyAsync(m->twy, ty.data(), sizeof(float2) * ty.size(), hipMemcpyHostToDevice, m->s)) ||
        bad(hipMemcpyAsync(m->twx64, tx64.data(), sizeof(double2) * tx64.size(), hipMemcpyHostToDevice, m->s)) ||
        bad(hipMemcpyAsync(m->twy64, ty64.data(), sizeof(double2) * ty64.size(), hipMemcpyHostToDevice, m->s)) ||
        bad(hipMemcpyAsync(m->twz, tz.data(), sizeof(double2) * tz.size(), hipMemcpyHostToDevice, m->s)) ||
        bad(hipStreamSynchronize(m->s)))
        return EPIC_E_HIP;
    for (int i = 0; i < 6; ++i) {
        m->ax_cur[i] = m->axbuf + i * c;
        m->ax_stale[i] = m->axbuf + (6 + i) * c;
    }
    if (bad(dalloc(&m->n_dev, 1, m->s)) || bad(hipHostMalloc((void **)&m->n_h, sizeof(int64_t), hipHostMallocDefault)))
        return EPIC_E_NOMEM;
    if (bad(bind_parcels(m))) return EPIC_E_HIP;
    *out = m;
    return EPIC_OK;
}

int epic_destroy(epic_handle m)
{
    if (!m) return EPIC_OK;
    if (m->s) (void)hipStreamSynchronize(m->s);
    m->release_plans();   // before the stream goes: released arenas wait for it
    if (m->merge_scratch.base) (void)hipFree(m->merge_scratch.base);
    void *ptrs[] = {m->pos, m->attr, m->orig, m->axbuf, m->keys, m->cnt[0], m->scnt[0], m->cnt[1], m->scnt[1],
                    m->cnt[2], m->scnt[2], m->cell_start, m->cursor, m->perm, m->bsum, m->fld, m->nparg, m->nsparg, m->staging,
                    m->twx, m->twy, m->twz, m->twx64, m->twy64, m->err_d, m->flags_d, m->vw_d, m->tm_d, m->diag_d,
                    m->sort_tmp, m->pmax_out, m->tend_src, m->far_acc, m->far_flag, m->far_scale, m->n_dev};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (m->spec[0]) (void)hipFree(m->spec[0]);
    if (m->err_h) (void)hipHostFree(m->err_h);
    if (m->vw_h) (void)hipHostFree(m->vw_h);
    if (m->tm_h) (void)hipHostFree(m->tm_h);
    if (m->diag_h) (void)hipHostFree(m->diag_h);
    if (m->n_h) (void)hipHostFree(m->n_h);
    if (m->s) (void)hipStreamDestroy(m->s);
    delete m;
    return EPIC_OK;
}

const char *epic_last_error(epic_handle m) { return m ? m->msg.c_str() : "null handle"; }

int64_t epic_num_parcels(epic_handle m) { return m ? m->n : -1; }

int64_t epic_nearest_ncell(epic_handle m) { return m ? (int64_t)(m->g.nx + 3) * (m->g.ny + 3) * (m->g.nz + 2) : -1; }

int epic_set_parcels(epic_handle m, int64_t n, const double *position, const double *vorticity, const double *B,
                     const double *volume, const double *buoyancy, const double *humidity)
{
    if (!m) return EPIC_E_INVALID;
    if (n < 0 || n > m->cap) return fail(m, EPIC_E_CAPACITY, "epic_set_parcels: more parcels than max_parcels");
    if (n > 0 && (!position || !vorticity || !B || !volume || !buoyancy || !humidity))
        return fail(m, EPIC_E_INVALID, "epic_set_parcels: null array");
    const int64_t c = m->cap;
    std::vector<double> hp(3 * (size_t)n);
    std::vector<float> ha((size_t)11 * n);
    std::vector<int> ho(n);
    host_parallel_for(n, [&](int64_t lo, int64_t hi) {
        for (int64_t q = lo; q < hi; ++q) {
            for (int d = 0; d < 3; ++d) hp[d * n + q] = position[3 * q + d];
            for (int d = 0; d < 3; ++d) ha[d * n + q] = (float)vorticity[3 * q + d];
            for (int d = 0; d < 5; ++d) ha[(3 + d) * n + q] = (float)B[5 * q + d];
            ha[8 * n + q] = (float)volume[q];
            ha[9 * n + q] = (float)buoyancy[q];
            ha[10 * n + q] = (float)humidity[q];
            ho[q] = (int)q;
        }
    });
    m->n = n;
    HIPOK(m, bind_parcels(m));
    m->counts_changed();
    for (int d = 0; d < 3 && n; ++d)
        HIPOK(m, hipMemcpyAsync(m->pos + d * c, hp.data() + d * n, sizeof(double) * n, hipMemcpyHostToDevice, m->s));
    for (int a = 0; a < 11 && n; ++a)
        HIPOK(m, hipMemcpyAsync(m->attr + a * c, ha.data() + a * n, sizeof(float) * n, hipMemcpyHostToDevice, m->s));
    if (n) HIPOK(m, hipMemcpyAsync(m->orig, ho.data(), sizeof(int) * n, hipMemcpyHostToDevice, m->s));
    HIPOK(m, hipStreamSynchronize(m->s));
    m->keys_valid = false;
    m->axes_dirty = true;
    m->tend_displaced = false;
    return EPIC_OK;
}

int epic_get_parcels(epic_handle m, int64_t n, double *position, double *vorticity, double *B, double *volume,
                     double *buoyancy, double *humidity)
{
    if (!m) return EPIC_E_INVALID;
    if (n != m->n) return fail(m, EPIC_E_INVALID, "epic_get_parcels: n differs from the stored parcel count");
    const int64_t c = m->cap;
    std::vector<double> hp(3 * (size_t)n);
    std::vector<float> ha((size_t)11 * n);
    std::vector<int> ho(n);
    for (int d = 0; d < 3 && n; ++d)
        HIPOK(m, hipMemcpyAsync(hp.data() + d * n, m->pos + d * c, sizeof(double) * n, hipMemcpyDeviceToHost, m->s));
    for (int a = 0; a < 11 && n; ++a)
        HIPOK(m, hipMemcpyAsync(ha.data() + a * n, m->attr + a * c, sizeof(float) * n, hipMemcpyDeviceToHost, m->s));
    if (n) HIPOK(m, hipMemcpyAsync(ho.data(), m->orig, sizeof(int) * n, hipMemcpyDeviceToHost, m->s));
    HIPOK(m, hipStreamSynchronize(m->s));
    host_parallel_for(n, [&](int64_t lo, int64_t hi) {   // orig is a permutation: one writer per element
        for (int64_t q = lo; q < hi; ++q) {
            const int64_t o = ho[q];
            if (position)
                for (int d = 0; d < 3; ++d) position[3 * o + d] = hp[d * n + q];
            if (vorticity)
                for (int d = 0; d < 3; ++d) vorticity[3 * o + d] = ha[d * n + q];
            if (B)
                for (int d = 0; d < 5; ++d) B[5 * o + d] = ha[(3 + d) * n + q];
            if (volume) volume[o] = ha[8 * n + q];
            if (buoyancy) buoyancy[o] = ha[9 * n + q];
            if (humidity) humidity[o] = ha[10 * n + q];
        }
    });
    return EPIC_OK;
}

int epic_get_parcel_tendencies(epic_handle m, int64_t n, double *delta_pos, double *delta_vor, double *strain)
{
    if (!m) return EPIC_E_INVALID;
    if (n != m->n) return fail(m, EPIC_E_INVALID, "epic_get_parcel_tendencies: n differs from the parcel count");
    {
        const int rc = settle_tendencies(m);
        if (rc) return rc;
    }
    const int64_t c = m->cap;
    std::vector<float> h((size_t)14 * n);
    std::vector<int> ho(n);
    for (int a = 0; a < 6 && n; ++a)
        HIPOK(m, hipMemcpyAsync(h.data() + a * n, m->attr + (11 + a) * c, sizeof(float) * n, hipMemcpyDeviceToHost, m->s));
    for (int a = 0; a < 8 && n; ++a)
        HIPOK(m, hipMemcpyAsync(h.data() + (6 + a) * n, m->attr + (28 + a) * c, sizeof(float) * n,
                                hipMemcpyDeviceToHost, m->s));
    if (n) HIPOK(m, hipMemcpyAsync(ho.data(), m->orig, sizeof(int) * n, hipMemcpyDeviceToHost, m->s));
    HIPOK(m, hipStreamSynchronize(m->s));
    for (int64_t q = 0; q < n; ++q) {
        const int64_t o = ho[q];
        for (int d = 0; d < 3; ++d) {
            if (delta_pos) delta_pos[3 * o + d] = h[d * n + q];
            if (delta_vor) delta_vor[3 * o + d] = h[(3 + d) * n + q];
        }
        if (strain)
            for (int d = 0; d < 8; ++d) strain[8 * o + d] = h[(6 + d) * n + q];
    }
    return EPIC_OK;
}

int epic_set_parcel_tendencies(epic_handle m, int64_t n, const double *delta_pos, const double *delta_vor)
{
    if (!m) return EPIC_E_INVALID;
    if (n != m->n) return fail(m, EPIC_E_INVALID, "epic_set_parcel_tendencies: n differs from the parcel count");
    {   // dB and strain stay with their parcels
        const int rc = settle_tendencies(m);
        if (rc) return rc;
    }
    const int64_t c = m->cap;
    std::vector<int> ho(n);
    if (n) HIPOK(m, hipMemcpyAsync(ho.data(), m->orig, sizeof(int) * n, hipMemcpyDeviceToHost, m->s));
    HIPOK(m, hipStreamSynchronize(m->s));
    std::vector<float> h((size_t)6 * n);
    for (int64_t q = 0; q < n; ++q)
        for (int d = 0; d < 3; ++d) {
            h[d * n + q] = delta_pos ? (float)delta_pos[3 * ho[q] + d] : 0.0f;
            h[(3 + d) * n + q] = delta_vor ? (float)delta_vor[3 * ho[q] + d] : 0.0f;
        }
    for (int a = 0; a < 6 && n; ++a)
        HIPOK(m, hipMemcpyAsync(m->attr + (11 + a) * c, h.data() + a * n, sizeof(float) * n, hipMemcpyHostToDevice, m->s));
    HIPOK(m, hipStreamSynchronize(m->s));
    return EPIC_OK;
}

int epic_get_field(epic_handle m, int which, double *out)
{
    if (!m || !out) return EPIC_E_INVALID;
    const int nc = field_ncomp(which);
    if (nc == 0) return fail(m, EPIC_E_INVALID, "epic_get_field: unknown field");
    if (which == EPIC_FIELD_NPARG || which == EPIC_FIELD_NSPARG)
        return copy_field_out(m, which == EPIC_FIELD_NPARG ? m->nparg : m->nsparg, 1, out);
    return copy_field_out(m, m->fld + field_slot(which) * field_stride(m->g), nc, out);
}

int epic_get_field_int(epic_handle m, int which, int *out)
{
    if (!m || !out) return EPIC_E_INVALID;
    if (which != EPIC_FIELD_NPARG && which != EPIC_FIELD_NSPARG)
        return fail(m, EPIC_E_INVALID, "epic_get_field_int: not an integer field");
    const Grid &g = m->g;
    const int64_t H = (int64_t)(g.nz + 3) * (g.ny + 3) * (g.nx + 3);
    std::vector<double> tmp((size_t)H);
    int rc = copy_field_out(m, which == EPIC_FIELD_NPARG ? m->nparg : m->nsparg, 1, tmp.data());
    if (rc) return rc;
    for (int64_t i = 0; i < H; ++i) out[i] = (int)tmp[i];
    return EPIC_OK;
}

int epic_set_field(epic_handle m, int which, const double *in)
{
    if (!m || !in) return EPIC_E_INVALID;
    const int nc = field_ncomp(which);
    const int slot = field_slot(which);
    if (nc == 0 || slot < 0) return fail(m, EPIC_E_INVALID, "epic_set_field: unknown or read-only field");
    const Grid &g = m->g;
    const int64_t FS = field_stride(g), H = (int64_t)(g.nz + 3) * (g.ny + 3) * (g.nx + 3);
    std::vector<float> h((size_t)(FS * nc));
    for (int c = 0; c < nc; ++c)
        for (int ix = 0; ix < g.nx; ++ix)
            for (int iy = 0; iy < g.ny; ++iy)
                for (int iz = -1; iz <= g.nz + 1; ++iz)
                    h[c * FS + (int64_t)(iz + 1) * g.ny * g.nx + (int64_t)iy * g.nx + ix] =
                        (float)in[c * H + ((int64_t)(ix + 1) * (g.ny + 3) + (iy + 1)) * (g.nz + 3) + (iz + 1)];
    HIPOK(m, hipMemcpyAsync(m->fld + slot * FS, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice, m->s));
    HIPOK(m, hipStreamSynchronize(m->s));
    return EPIC_OK;
}

int epic_par2grid(epic_handle m, int l_reuse)
{
    if (!m) return EPIC_E_INVALID;
    int rc = ensure_keys(m);
    if (rc) return rc;
    // get_points stores the axes it computes (parcel_ellipsoid.f90:216-247);
    // with l_reuse the stored ones are taken
    Plan &P = (l_reuse == 1) ? m->pl_p2g_reuse : m->pl_p2g_fresh;
    if (P.empty()) {
        if (l_reuse != 1) build_axes(P, {m}, {m->ax_stale});
        build_p2g(P, {m}, {P2GSel{m->ax_stale, nullptr, nullptr, 0, -1}});
    }
    if (l_reuse == 1) {
        rc = ensure_axes(m);
        if (rc) return rc;
    }
    rc = run_plan(m, P);
    if (rc) return rc;
    if (l_reuse != 1) m->axes_dirty = false;
    return check(m, "epic_par2grid");
}

int epic_vor2vel(epic_handle m)
{
    if (!m) return EPIC_E_INVALID;
    if (m->pl_vor2vel.empty()) build_vor2vel(m->pl_vor2vel, {m});
    const int rc = run_plan(m, m->pl_vor2vel);
    if (rc) return rc;
    return check(m, "epic_vor2vel");
}

int epic_vorticity_tendency(epic_handle m)
{
    if (!m) return EPIC_E_INVALID;
    if (m->pl_vtend.empty()) build_vtend(m->pl_vtend, {m});
    const int rc = run_plan(m, m->pl_vtend);
    if (rc) return rc;
    return check(m, "epic_vorticity_tendency");
}

int epic_grid2par(epic_handle m, int add)
{
    if (!m) return EPIC_E_INVALID;
    if (add == 1) {   // adds onto the tendencies of the same parcels
        const int rc = settle_tendencies(m);
        if (rc) return rc;
    }
    m->tend_displaced = false;   // rewritten (add) or added onto in place
    Plan &P = m->pl_g2p[add == 1 ? 1 : 0];
    if (P.empty()) {
        G2PArgs a{};
        a.g = m->g;
        a.p = m->pd;
        a.fld = m->fld;
        a.add = (add == 1);
        a.err = m->err_d;
        add_grid2par(P, PH_GRID2PAR, {a});
    }
    const int rc = run_plan(m, P);
    if (rc) return rc;
    return check(m, "epic_grid2par");
}

// the arrays that outlive a step: positions, vorticity, B, volume,
// buoyancy, humidity, caller index, next-stage keys, cached half axes; all
// are gathered into the cell order in sort_tmp by one launch and copied back
// by another (stage deltas and strain are rewritten before they are read
// again)
static_assert(RESORT_W == 3 && RESORT_F == 11 + 2 + 6, "positions; attributes, orig, keys, half axes");
static int64_t resort_tmp_doubles(int64_t cap) { return (RESORT_W + (RESORT_F + 1) / 2) * cap; }

static void build_resort(epic_member *m)
{
    Plan &P = m->pl_resort;
    const int64_t n = m->n, c = m->cap, nc = ncells(m->g);
    add_exclusive_scan(P, PH_SORT, {ScanArgs{m->cnt[0], nc, m->cell_start, m->bsum, m->cursor}});
    add_place(P, PH_SORT, {PlaceArgs{n, m->keys, m->cell_start, m->cursor, m->perm}});
    ResortArgs a{};
    a.perm = m->perm;
    a.n = n;
    for (int d = 0; d < RESORT_W; ++d) {
        a.w[d] = m->pos + d * c;
        a.tw[d] = m->sort_tmp + d * c;
    }
    int k = 0;
    for (int i = 0; i < 11; ++i) a.f[k++] = m->attr + i * c;   // vorticity, B, volume, buoyancy, humidity
    a.f[k++] = (float *)m->orig;
    a.f[k++] = (float *)m->keys;
    for (int i = 0; i < 6; ++i) a.f[k++] = m->ax_stale[i];
    float *tf = (float *)(m->sort_tmp + RESORT_W * c);
    for (int i = 0; i < RESORT_F; ++i) a.tf[i] = tf + i * c;
    a.gather = 1;
    add_resort(P, PH_SORT, {a});
    a.gather = 0;
    add_resort(P, PH_SORT, {a});
}

static int ensure_sort_scratch(epic_member *m)
{
    if (!m->sort_tmp) HIPOK(m, dalloc(&m->sort_tmp, resort_tmp_doubles(m->cap), m->s));
    if (!m->tend_src) HIPOK(m, dalloc(&m->tend_src, m->cap, m->s));
    return EPIC_OK;
}

// the 19 tendency arrays gathered by src (dst[s] = a[src[s]]) through sort_tmp
static void build_tendency_gather(epic_member *m, Plan &P, const int *src)
{
    const int64_t c = m->cap;
    ResortArgs a{};
    a.perm = src;
    a.n = m->n;
    for (int i = 0; i < RESORT_F; ++i) {   // attr slots 11..21 (dpos, dvor, dB) and 28..35 (strain)
        const int slot = i < 11 ? 11 + i : 28 + (i - 11);
        a.f[i] = m->attr + slot * c;
        a.tf[i] = (float *)(m->sort_tmp) + i * c;
    }
    a.gather = 1;
    add_resort(P, PH_SORT, {a}, false);
    a.gather = 0;
    add_resort(P, PH_SORT, {a}, false);
}

// puts every parcel's stage tendencies back into its own slot
static int settle_tendencies(epic_member *m)
{
    if (!m->tend_displaced) return EPIC_OK;
    m->tend_displaced = false;
    if (m->n == 0) return EPIC_OK;
    int rc = ensure_sort_scratch(m);
    if (rc) return rc;
    if (m->pl_settle.empty()) build_tendency_gather(m, m->pl_settle, m->tend_src);
    rc = run_plan(m, m->pl_settle);
    if (rc) return rc;
    return check(m, "settle_tendencies");
}

// after the persistent arrays were gathered by perm (n entries): the
// tendencies stay where they were, so the parcel now in slot s finds them in
// slot perm[s]
static int displace_tendencies(epic_member *m, const int *perm)
{
    if (m->n) HIPOK(m, hipMemcpyAsync(m->tend_src, perm, sizeof(int) * m->n, hipMemcpyDeviceToDevice, m->s));
    m->tend_displaced = true;
    return EPIC_OK;
}

static int resort(epic_member *m)
{
    m->steps_since_sort = 0;
    if (m->n == 0) return EPIC_OK;
    int rc = settle_tendencies(m);   // one displacement at a time
    if (rc) return rc;
    rc = ensure_keys(m);   // keys and the histogram cnt[0] of the current positions
    if (rc) return rc;
    rc = ensure_sort_scratch(m);
    if (rc) return rc;
    if (m->pl_resort.empty()) build_resort(m);
    rc = run_plan(m, m->pl_resort);
    if (rc) return rc;
    rc = displace_tendencies(m, m->perm);
    if (rc) return rc;
    return check(m, "epic_resort");
}


// one RK step of a member: the plan (fixed or adaptive) and its StepTime
// written before the launch; *dt_taken receives the step's dt
static int run_step(epic_member *m, double *t, double t_initial, bool adaptive, double *dt_taken, const char *where)
{
    int rc = ensure_keys(m);
    if (rc == EPIC_OK) rc = ensure_axes(m);
    if (rc) return rc;
    Plan &P = adaptive ? m->pl_step_adaptive : m->pl_step;
    uint64_t &pg = m->step_gen[adaptive ? 1 : 0];
    if (!P.empty() && pg != m->gen) {   // stale: new tables into the same graph, or a new plan
        Plan Q;
        build_step(Q, {StepItem{m, m->tm_d, m->flags_d, adaptive}});
        bool same = false;
        HIPOK(m, P.refresh_from(Q, m->s, &same));
        if (!same) P.swap(Q);   // Q now holds the old plan and releases it
    }
    if (P.empty()) build_step(P, {StepItem{m, m->tm_d, m->flags_d, adaptive}});
    pg = m->gen;
    // the first par2grid of a step reuses the cached axes (ls_rk.f90:93)
    m->err_h[1] = (*t > t_initial) ? 1 : 0;
    m->tm_h->t = *t;
    m->tm_h->adaptive = adaptive ? 1 : 0;
    m->tm_h->gmax_bits = m->tm_h->b2max_bits = 0ull;
    HIPOK(m, hipMemcpyAsync(m->flags_d, m->err_h + 1, sizeof(int), hipMemcpyHostToDevice, m->s));
    HIPOK(m, hipMemcpyAsync(m->tm_d, m->tm_h, sizeof(StepTime), hipMemcpyHostToDevice, m->s));
    if (m->profiling == 1) {
        rc = run_plan(m, P);
        if (rc) return rc;
    } else {
        HIPOK(m, P.run_graph(m->s, m->profiling == 2));
    }
    HIPOK(m, hipMemcpyAsync(m->tm_h, m->tm_d, sizeof(StepTime), hipMemcpyDeviceToHost, m->s));
    m->tend_displaced = false;   // every stage rewrote the tendencies in place
    rc = check(m, where);   // synchronises the stream
    if (rc) return rc;
    if (m->profiling == 2) {
        rc = collect_graph_timers(m, P);
        if (rc) return rc;
    }
    *dt_taken = m->tm_h->dt;
    *t = *t + m->tm_h->dt;
    if (m->resort_every > 0 && ++m->steps_since_sort >= m->resort_every) return resort(m);
    return EPIC_OK;
}

int epic_resort(epic_handle m)
{
    if (!m) return EPIC_E_INVALID;
    return resort(m);
}

int epic_set_resort_interval(epic_handle m, int steps)
{
    if (!m || steps < 0) return EPIC_E_INVALID;
    m->resort_every = steps;
    return EPIC_OK;
}

int epic_ls_rk_step(epic_handle m, double *t, double dt, double t_initial)
{
    if (!m || !t) return EPIC_E_INVALID;
    if (!(dt > 0.0)) return fail(m, EPIC_E_DT, "epic_ls_rk_step: time step must be positive");
    m->tm_h->dt = dt;
    double taken;
    return run_step(m, t, t_initial, false, &taken, "epic_ls_rk_step");
}

int epic_ls_rk_step_adaptive(epic_handle m, double *t, const epic_time_t *opts, double t_initial, double *dt)
{
    if (!m || !t || !opts) return EPIC_E_INVALID;
    if (opts->fixed) {
        if (!(opts->fixed_dt > 0.0)) return fail(m, EPIC_E_DT, "epic_ls_rk_step_adaptive: time step must be positive");
        m->tm_h->dt = opts->fixed_dt;
        double taken;
        const int rc = run_step(m, t, t_initial, false, &taken, "epic_ls_rk_step_adaptive");
        if (rc == EPIC_OK && dt) *dt = taken;
        return rc;
    }
    m->tm_h->alpha = opts->alpha;
    m->tm_h->limit = opts->limit;
    m->tm_h->precise_stop = opts->precise_stop ? 1 : 0;
    double taken = 0.0;
    const int rc = run_step(m, t, t_initial, true, &taken, "epic_ls_rk_step_adaptive");
    if (rc == EPIC_OK && dt) *dt = taken;
    return rc;
}

int epic_get_time_step(epic_handle m, double t, const epic_time_t *opts, double *dt, double *gmax, double *bmax)
{
    if (!m || !opts || !dt) return EPIC_E_INVALID;
    StepTime h{};
    h.t = t;
    h.alpha = opts->alpha;
    h.limit = opts->limit;
    h.precise_stop = opts->precise_stop ? 1 : 0;
    h.adaptive = 1;
    *m->tm_h = h;
    HIPOK(m, hipMemcpyAsync(m->tm_d, m->tm_h, sizeof(StepTime), hipMemcpyHostToDevice, m->s));
    {
        Plan P;
        add_time_step(P, PH_OTHER, {TsArgs{m->g, m->fld, m->tm_d, m->err_d}});
        HIPOK(m, P.run(m->s));
        HIPOK(m, hipMemcpyAsync(m->tm_h, m->tm_d, sizeof(StepTime), hipMemcpyDeviceToHost, m->s));
        const int rc = check(m, "epic_get_time_step");
        if (gmax) *gmax = m->tm_h->gmax;
        if (bmax) *bmax = m->tm_h->bmax;
        if (rc) return rc;
    }
    double h_dt = opts->fixed ? opts->fixed_dt : m->tm_h->dt;
    if (opts->fixed && opts->precise_stop && (t + h_dt > opts->limit)) h_dt = opts->limit - t;
    if (!(h_dt > 0.0)) return fail(m, EPIC_E_DT, "epic_get_time_step: Time step <= 0!");
    *dt = h_dt;
    return EPIC_OK;
}

int epic_set_profiling(epic_handle m, int on)
{
    if (!m || on < 0 || on > 2) return EPIC_E_INVALID;
    m->profiling = on;
    return EPIC_OK;
}

int epic_get_timers(epic_handle m, double *ms, int *calls, int reset)
{
    if (!m) return EPIC_E_INVALID;
    for (int k = 0; k < EPIC_NUM_TIMERS; ++k) {
        if (ms) ms[k] = m->timers_ms[k];
        if (calls) calls[k] = m->timer_calls[k];
        if (reset) {
            m->timers_ms[k] = 0.0;
            m->timer_calls[k] = 0;
        }
    }
    return EPIC_OK;
}

// ---------------------------------------------------------------------------
// batched step, autotuner, memory estimate
// ---------------------------------------------------------------------------
}  // extern "C"

struct epic_batch {
    std::vector<epic_member *> ms;
    hipStream_t s = nullptr;
    int *flags_d = nullptr, *errs_d = nullptr;
    int *flags_h = nullptr, *errs_h = nullptr;
    StepTime *tms_d = nullptr, *tms_h = nullptr;   // one time step per member
    Plan plan;
    std::vector<uint64_t> gens;
    std::string msg;
    // a launched step not yet waited for: each member's dt
    bool in_flight = false;
    std::vector<double> dt_pending;
};

namespace {

int batch_fail(epic_batch *b, int code, const std::string &what)
{
    if (b) b->msg = what;
    return code;
}

#define BHIPOK(b, expr)                                                                         \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return batch_fail(b, EPIC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// device time of one plan replay on the member stream (median of reps)
float time_plan(epic_member *m, Plan &P, int reps)
{
    std::vector<float> t;
    for (int r = 0; r <= reps; ++r) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess) return 1e30f;
        if (hipEventCreate(&b) != hipSuccess) return 1e30f;
        (void)hipEventRecord(a, m->s);
        (void)P.run(m->s);
        (void)hipEventRecord(b, m->s);
        (void)hipEventSynchronize(b);
        float ms = 1e30f;
        (void)hipEventElapsedTime(&ms, a, b);
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        if (r > 0) t.push_back(ms);   // the first replay warms up
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

}  // namespace

extern "C" {

int64_t epic_member_bytes(const epic_grid_t *grid, int64_t max_parcels)
{
    if (!grid || grid->nx < 8 || grid->ny < 8 || grid->nz < 4 || max_parcels < 0) return -1;
    Grid g{};
    g.nx = grid->nx;
    g.ny = grid->ny;
    g.nz = grid->nz;
    g.nxh = g.nx / 2 + 1;
    g.nzp = ((g.nz + 3 + ZB - 1) / ZB) * ZB;   // planes 0..nz and two halo planes
    const int64_t c = std::max<int64_t>(max_parcels, 1), nc = ncells(g), FS = field_stride(g);
    int64_t b = 0;
    b += c * (3 * 8 + (int64_t)NATTR * 4 + 4 + 12 * 4 + 4 + 4) + 4;           // parcels, axes, orig, keys, perm
    b += 4 * (6 * nc + (nc + 1) + nc);                                        // histograms, cell_start, cursor
    b += 4 * scan_scratch_ints(std::max<int64_t>(std::max<int64_t>(nc, c), (int64_t)(g.nx + 3) * (g.ny + 3) * (g.nz + 2)));
    b += 4 * ((int64_t)F_COUNT * FS + 2 * FS) + 8 * staging_elems_max(g);   // fields, counts, staging
    b += 8 * (int64_t)P2G_NF * FS + 4 + 8 * P2G_NF;                          // far-point sums
    b += 20 * c + 4096;                                                       // merge scratch (its first phase)
    b += (int64_t)NSPEC * 8 * spec_size(g);                                  // spectra
    return b;
}

int epic_autotune(epic_handle m, int *tz_p2g, int *tz_g2p)
{
    if (!m) return EPIC_E_INVALID;
    int rc = ensure_keys(m);
    if (rc) return rc;
    {
        Plan P;
        build_axes(P, {m}, {m->ax_cur});
        HIPOK(m, P.run(m->s));
    }
    float best_p = 1e30f, best_g = 1e30f;
    int tp = m->tz_p2g, tg = m->tz_g2p;
    for (int tz : TZ_P2G) {
        if (tz > m->g.nz + 2) continue;
        m->tz_p2g = tz;
        Plan P;
        build_p2g(P, {m}, {P2GSel{m->ax_cur, nullptr, nullptr, 0, -1}});
        const float t = time_plan(m, P, 3);
        if (t < best_p) {
            best_p = t;
            tp = tz;
        }
    }
    m->tz_p2g = tp;
    {   // a p2g with the chosen layout leaves perm / cell_start for the probes below
        Plan P;
        build_p2g(P, {m}, {P2GSel{m->ax_cur, nullptr, nullptr, 0, -1}});
        HIPOK(m, P.run(m->s));
    }
    for (int tz : TZ_G2P) {
        if (tz > m->g.nz + 2) continue;
        // a first stage with cb dt = 0 leaves position, vorticity and B
        // unchanged; it rewrites the stage deltas (reset by every step's first
        // stage) and recounts the current cells into the spare histogram
        const P2GLayout L = layout_tz(m->g, tz);
        G2PTileArgs a{};
        a.g = m->g;
        a.p = m->pd;
        for (int c = 0; c < 6; ++c) a.axes[c] = a.axes_next[c] = m->ax_cur[c];   // rewritten unchanged, as a step's push does
        a.perm = m->perm;
        a.cell_start = m->cell_start;
        a.fld = m->fld;
        a.stage = 1;
        a.nstages = 2;
        a.reflect = 0;
        a.cb = 0.0;
        a.dt = &m->tm_d->dt;
        a.ca = 0.0;
        a.keys = m->keys;
        a.count = m->cnt[1];
        a.scount = m->scnt[1];
        a.err = m->err_d;
        a.tz = L.tz;
        a.ntx = L.ntx;
        a.nty = L.nty;
        a.nchz = L.nchz;
        Plan P;
        add_rk_stage_tiles(P, PH_PUSH, {a});
        const float t = time_plan(m, P, 3);
        if (t < best_g) {
            best_g = t;
            tg = tz;
        }
    }
    m->tz_g2p = tg;
    m->invalidate_plans();
    if (tz_p2g) *tz_p2g = m->tz_p2g;
    if (tz_g2p) *tz_g2p = m->tz_g2p;
    return check(m, "epic_autotune");
}

int epic_set_tiling(epic_handle m, int tz_p2g, int tz_g2p)
{
    if (!m) return EPIC_E_INVALID;
    bool ok_p = false, ok_g = false;
    for (int tz : TZ_P2G) ok_p |= (tz == tz_p2g);
    for (int tz : TZ_G2P) ok_g |= (tz == tz_g2p);
    if (!ok_p || !ok_g)
        return fail(m, EPIC_E_INVALID, "epic_set_tiling: z chunks must be 2, 4 or 8");
    m->tz_p2g = tz_p2g;
    m->tz_g2p = tz_g2p;
    m->invalidate_plans();
    return EPIC_OK;
}

int epic_batch_create(const epic_handle *members, int count, epic_batch_handle *out)
{
    if (!members || count <= 0 || !out) return EPIC_E_INVALID;
    *out = nullptr;
    epic_batch *b = new epic_batch();
    for (int i = 0; i < count; ++i) {
        if (!members[i]) {
            delete b;
            return EPIC_E_INVALID;
        }
        for (int j = 0; j < i; ++j)
            if (members[j] == members[i]) {
                delete b;
                return EPIC_E_INVALID;
            }
        b->ms.push_back(members[i]);
    }
    auto bad = [&](hipError_t e) {
        if (e != hipSuccess) {
            epic_batch_destroy(b);
            return true;
        }
        return false;
    };
    if (bad(hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking))) return EPIC_E_HIP;
    if (bad(dalloc(&b->flags_d, count, b->s)) || bad(dalloc(&b->errs_d, count, b->s)) || bad(dalloc(&b->tms_d, count, b->s)))
        return EPIC_E_NOMEM;
    if (bad(hipHostMalloc((void **)&b->flags_h, sizeof(int) * count, hipHostMallocDefault)) ||
        bad(hipHostMalloc((void **)&b->errs_h, sizeof(int) * count, hipHostMallocDefault)) ||
        bad(hipHostMalloc((void **)&b->tms_h, sizeof(StepTime) * count, hipHostMallocDefault)))
        return EPIC_E_NOMEM;
    *out = b;
    return EPIC_OK;
}

int epic_batch_destroy(epic_batch_handle b)
{
    if (!b) return EPIC_OK;
    if (b->s) (void)hipStreamSynchronize(b->s);
    b->plan.release();
    if (b->flags_d) (void)hipFree(b->flags_d);
    if (b->errs_d) (void)hipFree(b->errs_d);
    if (b->flags_h) (void)hipHostFree(b->flags_h);
    if (b->errs_h) (void)hipHostFree(b->errs_h);
    if (b->tms_d) (void)hipFree(b->tms_d);
    if (b->tms_h) (void)hipHostFree(b->tms_h);
    if (b->s) (void)hipStreamDestroy(b->s);
    delete b;
    return EPIC_OK;
}

const char *epic_batch_last_error(epic_batch_handle b) { return b ? b->msg.c_str() : "null handle"; }

// builds (when membership or a member changed) and captures the batch's
// step graph; keys of fresh members are computed on their own streams
static int batch_prepare(epic_batch *b)
{
    const int nm = (int)b->ms.size();
    for (int i = 0; i < nm; ++i) {   // keys and cached axes of changed members, on their streams
        epic_member *m = b->ms[i];
        if (!m->keys_valid || m->axes_dirty) {
            int rc = ensure_keys(m);
            if (rc == EPIC_OK) rc = ensure_axes(m);
            if (rc) return batch_fail(b, rc, m->msg);
            hipEvent_t ev = nullptr;
            BHIPOK(b, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            hipError_t e = hipEventRecord(ev, m->s);
            if (e == hipSuccess) e = hipStreamWaitEvent(b->s, ev, 0);
            (void)hipEventDestroy(ev);
            BHIPOK(b, e);
        }
    }
    bool same = b->gens.size() == (size_t)nm;
    for (int i = 0; same && i < nm; ++i) same = (b->gens[i] == b->ms[i]->gen);
    if (b->plan.empty() || !same) {
        // a member changed: the step rebuilt; when its launch sequence is the
        // same (new parcel counts, say) the captured graph keeps running on
        // the new argument tables
        Plan Q;
        std::vector<StepItem> items;
        std::vector<ErrArgs> errs;
        for (int i = 0; i < nm; ++i) {
            items.push_back(StepItem{b->ms[i], b->tms_d + i, b->flags_d + i, false});
            errs.push_back(ErrArgs{b->ms[i]->err_d, b->errs_d + i});
        }
        build_step(Q, items);
        add_gather_errors(Q, errs);
        bool kept = false;
        if (!b->plan.empty()) BHIPOK(b, b->plan.refresh_from(Q, b->s, &kept));
        if (!kept) b->plan.swap(Q);
        b->gens.clear();
        for (int i = 0; i < nm; ++i) b->gens.push_back(b->ms[i]->gen);
    }
    BHIPOK(b, b->plan.prepare_graph(b->s));
    return EPIC_OK;
}

int epic_batch_prepare(epic_batch_handle b)
{
    if (!b) return EPIC_E_INVALID;
    if (b->in_flight) return batch_fail(b, EPIC_E_INVALID, "epic_batch_prepare: a step is in flight");
    return batch_prepare(b);
}

int epic_batch_launch(epic_batch_handle b, const double *t, const double *dt, const double *t_initial)
{
    if (!b || !t || !dt || !t_initial) return EPIC_E_INVALID;
    if (b->in_flight) return batch_fail(b, EPIC_E_INVALID, "epic_batch_launch: the previous step was not waited for");
    const int nm = (int)b->ms.size();
    for (int i = 0; i < nm; ++i)
        if (!(dt[i] > 0.0)) return batch_fail(b, EPIC_E_DT, "epic_batch_step: time step must be positive");
    int rc = batch_prepare(b);
    if (rc) return rc;
    // the first par2grid of a step reuses the cached axes (ls_rk.f90:93)
    for (int i = 0; i < nm; ++i) {
        b->flags_h[i] = (t[i] > t_initial[i]) ? 1 : 0;
        b->tms_h[i] = StepTime{};
        b->tms_h[i].dt = dt[i];
        b->tms_h[i].t = t[i];
    }
    BHIPOK(b, hipMemcpyAsync(b->flags_d, b->flags_h, sizeof(int) * nm, hipMemcpyHostToDevice, b->s));
    BHIPOK(b, hipMemcpyAsync(b->tms_d, b->tms_h, sizeof(StepTime) * nm, hipMemcpyHostToDevice, b->s));
    BHIPOK(b, b->plan.run_graph(b->s));
    BHIPOK(b, hipMemcpyAsync(b->errs_h, b->errs_d, sizeof(int) * nm, hipMemcpyDeviceToHost, b->s));
    b->dt_pending.assign(dt, dt + nm);
    b->in_flight = true;
    return EPIC_OK;
}

int epic_batch_wait(epic_batch_handle b, double *t, int *status)
{
    if (!b || !t) return EPIC_E_INVALID;
    if (!b->in_flight) return batch_fail(b, EPIC_E_INVALID, "epic_batch_wait: no step in flight");
    b->in_flight = false;
    BHIPOK(b, hipStreamSynchronize(b->s));
    // The step's launch errors came back from hipGraphLaunch and its
    // execution errors from the synchronisation.  The thread's last-error
    // slot may still hold an error of an unrelated earlier call (a cleanup
    // call whose status is ignored, or one the runtime made while another
    // batch was captured): it is cleared here, not taken as this step's.
    (void)hipGetLastError();
    const int nm = (int)b->ms.size();
    int first = EPIC_OK;
    for (int i = 0; i < nm; ++i) {
        epic_member *m = b->ms[i];
        int rc = error_code(m, b->errs_h[i], "epic_batch_step");
        m->tend_displaced = false;   // every stage rewrote the tendencies in place
        if (rc == EPIC_OK) {
            t[i] += b->dt_pending[i];
            if (m->resort_every > 0 && ++m->steps_since_sort >= m->resort_every) rc = resort(m);
        }
        if (status) status[i] = rc;
        else if (first == EPIC_OK) {
            first = rc;
            b->msg = m->msg;
        }
    }
    return first;
}

int epic_batch_step(epic_batch_handle b, double *t, const double *dt, const double *t_initial, int *status)
{
    const int rc = epic_batch_launch(b, t, dt, t_initial);
    if (rc) return rc;
    return epic_batch_wait(b, t, status);
}

int epic_parcel_split(epic_handle m, double lambda_max, double amax, int64_t *n_split)
{
    if (!m) return EPIC_E_INVALID;
    (void)hipGetLastError();   // only this call's launch errors are read below
    if (n_split) *n_split = 0;
    const Grid &g = m->g;
    if (!(lambda_max > 0.0)) lambda_max = 4.0;   // options.f90:89
    if (!(amax > 0.0))                           // parameters.f90:150
        amax = std::cbrt(0.75 / 3.14159265358979323846) * std::min(g.dx[0], std::min(g.dx[1], g.dx[2]));
    const int64_t n0 = m->n;
    if (n0 == 0) return EPIC_OK;
    {   // the children's tendencies are written in their own slots
        int rc = settle_tendencies(m);
        if (rc == EPIC_OK) rc = ensure_sort_scratch(m);
        if (rc) return rc;
    }
    // flags in keys, child slots (exclusive prefix, total at [n0]) in perm;
    // the children follow their parents' caller index (parcel_split.f90:87-160):
    // the flags moved to caller indices and their exclusive prefix, in sort_tmp
    int *cflag = (int *)m->sort_tmp, *crank = cflag + m->cap + 1;
    SplitArgs a{};
    a.g = g;
    a.p = m->pd;
    a.lambda_max = lambda_max;
    a.amax = amax;
    a.flag = m->keys;
    a.rank = m->perm;
    a.crank = crank;
    a.err = m->err_d;
    {
        Plan P;
        add_split_flag(P, PH_OTHER, {a});
        add_exclusive_scan(P, PH_OTHER, {ScanArgs{m->keys, n0, m->perm, m->bsum, nullptr}});
        add_to_caller(P, PH_OTHER, {ToCallerArgs{n0, m->orig, m->keys, cflag}});
        add_exclusive_scan(P, PH_OTHER, {ScanArgs{cflag, n0, crank, m->bsum, nullptr}});
        HIPOK(m, P.run(m->s));
    }
    m->keys_valid = false;   // the keys array now holds the flags
    int total = 0;
    HIPOK(m, hipMemcpyAsync(&total, m->perm + n0, sizeof(int), hipMemcpyDeviceToHost, m->s));
    int rc = check(m, "epic_parcel_split");
    if (rc) return rc;
    if (n0 + total > m->cap)   // the reference grows its container (parcel_split.f90:87-96)
        return fail(m, EPIC_E_CAPACITY, "epic_parcel_split: more parcels than max_parcels after the split");
    if (total > 0) {
        Plan P;
        add_split_apply(P, PH_OTHER, {a});
        HIPOK(m, P.run(m->s));
        m->n = n0 + total;
        HIPOK(m, bind_parcels(m));
        m->counts_changed();
        // the cached support-point axes of every parcel from its current B
        // (what the next correction's vol2grid stores, parcel_interpl.f90:92-106),
        // computed before they are next read
        m->axes_dirty = true;
        rc = check(m, "epic_parcel_split");
        if (rc) return rc;
    }
    if (n_split) *n_split = total;
    m->n_splits += total;
    return EPIC_OK;
}

int epic_parcel_merge(epic_handle m, int64_t *n_merge)
{
    if (!m) return EPIC_E_INVALID;
    (void)hipGetLastError();   // only this call's launch errors are read below
    if (n_merge) *n_merge = 0;
    {   // the slot compaction moves the tendencies with their parcels
        const int rc = settle_tendencies(m);
        if (rc) return rc;
    }
    MergeCtx c{};
    c.g = m->g;
    c.p = m->pd;
    c.vmin = m->g.vmin;
    c.s = m->s;
    c.ncell = (int64_t)(m->g.nx + 3) * (m->g.ny + 3) * (m->g.nz + 2);
    c.err = m->err_d;
    c.attr = m->attr;
    c.cap = m->cap;
    c.nattr = NATTR;
    c.scratch = &m->merge_scratch;
    c.stats = m->merge_stats;
    int64_t nm = 0;
    const int r = run_merge(c, &nm);
    if (r < 0) return fail(m, EPIC_E_HIP, "epic_parcel_merge: HIP failure");
    int rc = check(m, "epic_parcel_merge");
    if (rc) return rc;
    if (nm > 0) {
        m->n -= nm;
        HIPOK(m, bind_parcels(m));
        m->counts_changed();
        m->keys_valid = false;   // keys, histogram and cached axes of the new
        m->axes_dirty = true;    // parcel set: computed before they are next read
    }
    if (n_merge) *n_merge = nm;
    m->n_merges += nm;
    return EPIC_OK;
}

int epic_parcel_merge_stats(epic_handle m, int64_t *n_way)
{
    if (!m || !n_way) return EPIC_E_INVALID;
    for (int j = 0; j < 7; ++j) n_way[j] = m->merge_stats[1 + j];
    return EPIC_OK;
}

int epic_apply_laplace(epic_handle m, int l_reuse)
{
    if (!m) return EPIC_E_INVALID;
    int rc = epic_par2grid(m, l_reuse);   // vol2grid (parcel_correction.f90:151)
    if (rc) return rc;
    if (m->n == 0) return EPIC_OK;
    const Grid &g = m->g;
    if (m->pl_lap[0].empty()) build_laplace(m);
    rc = run_plan(m, m->pl_lap[0]);
    if (rc) return rc;
    // spectra 4..6 are free between calls: the FP64 tridiagonal scratch
    launch_lapinv_columns(g, m->spec[0], (double *)m->spec[4], (double *)m->spec[5], (double *)m->spec[6], m->spec[1],
                          m->spec[2], m->spec[3], m->s);
    HIPOK(m, hipGetLastError());
    rc = run_plan(m, m->pl_lap[1]);
    if (rc) return rc;
    launch_grad_halo(g, m->fld, m->s);
    launch_laplace_push(g, m->pd, m->fld, m->err_d, m->s);
    HIPOK(m, hipGetLastError());
    m->keys_valid = false;
    return check(m, "epic_apply_laplace");
}

int epic_apply_gradient(epic_handle m, double prefactor, double max_compression, int l_reuse)
{
    if (!m) return EPIC_E_INVALID;
    int rc = epic_par2grid(m, l_reuse);   // vol2grid (parcel_correction.f90:194)
    if (rc) return rc;
    if (m->n == 0) return EPIC_OK;
    launch_gradient_push(m->g, m->pd, m->fld, prefactor, max_compression, m->err_d, m->s);
    HIPOK(m, hipGetLastError());
    m->keys_valid = false;
    return check(m, "epic_apply_gradient");
}

// sum vol and sum vol * vorticity over the parcels, FP64 in a fixed order
static int volume_weighted_sums(epic_member *m, double out[4])
{
    if (!m->vw_d) {
        HIPOK(m, dalloc(&m->vw_d, vw_partials(m->cap) + 4, m->s));
        HIPOK(m, hipHostMalloc((void **)&m->vw_h, sizeof(double) * 4, hipHostMallocDefault));
    }
    double *tot = m->vw_d + vw_partials(m->cap);
    launch_volume_weighted_vorticity(m->pd, m->vw_d, tot, m->s);
    HIPOK(m, hipGetLastError());
    HIPOK(m, hipMemcpyAsync(m->vw_h, tot, sizeof(double) * 4, hipMemcpyDeviceToHost, m->s));
    HIPOK(m, hipStreamSynchronize(m->s));
    for (int c = 0; c < 4; ++c) out[c] = m->vw_h[c];
    return EPIC_OK;
}

int epic_init_parcel_correction(epic_handle m, double *vor_bar)
{
    if (!m) return EPIC_E_INVALID;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    if (m->n > 0) {
        const int rc = volume_weighted_sums(m, s);
        if (rc) return rc;
    }
    for (int c = 0; c < 3; ++c) m->vor_bar[c] = (s[0] > 0.0) ? s[1 + c] / s[0] : 0.0;
    if (vor_bar)
        for (int c = 0; c < 3; ++c) vor_bar[c] = m->vor_bar[c];
    return EPIC_OK;
}

int epic_apply_vortcor(epic_handle m)
{
    if (!m) return EPIC_E_INVALID;
    if (m->n == 0) return EPIC_OK;
    double s[4];
    const int rc = volume_weighted_sums(m, s);
    if (rc) return rc;
    if (!(s[0] > 0.0)) return fail(m, EPIC_E_INVALID, "epic_apply_vortcor: total parcel volume is not positive");
    double d[3];
    for (int c = 0; c < 3; ++c) d[c] = s[1 + c] / s[0] - m->vor_bar[c];
    launch_vortcor(m->pd, d, m->s);
    HIPOK(m, hipGetLastError());
    return check(m, "epic_apply_vortcor");
}

// the reported parcel diagnostics from the 12 reduced totals
// (parcel_diagnostics.f90:51-159)
static void parcel_stats_of(const epic_member *m, const double s[12], double *stats)
{
    const Grid &g = m->g;
    const double vdomaini = 1.0 / (g.extent[0] * g.extent[1] * g.extent[2]);
    const double ntoti = 1.0 / (double)m->n;
    stats[1] = 0.5 * s[0] * vdomaini;
    stats[2] = s[6];
    stats[3] = s[2] * ntoti;
    stats[4] = s[4] * ntoti;
    stats[5] = std::sqrt(std::fabs(s[3] * ntoti - stats[3] * stats[3]));
    stats[6] = std::sqrt(std::fabs(s[5] * ntoti - stats[4] * stats[4]));
    stats[7] = s[4];
    for (int c = 0; c < 3; ++c) stats[8 + c] = std::sqrt(s[7 + c] / s[4]);
    stats[12] = 0.5 * s[1] * vdomaini;
    stats[16] = std::min(0.0, s[11]);   // the extremes start from zero (parcel_diagnostics.f90:60-61)
    stats[17] = std::max(0.0, s[10]);
}

static void parcel_counts_of(epic_member *m, double *stats, int reset_counters)
{
    stats[11] = (double)m->n;
    stats[13] = (double)m->n_splits;
    stats[14] = (double)m->merge_stats[0];   // n_big_close
    stats[15] = (double)m->n_merges;
    if (reset_counters) {   // as after a write (parcel_diagnostics_netcdf.f90:181-184)
        m->n_splits = m->n_merges = 0;
        for (int j = 0; j < 8; ++j) m->merge_stats[j] = 0;
    }
}

// run a diagnostics reduction and copy its nv totals to the host
static int diag_run(epic_member *m, bool parcels, double *out, int nv, const char *where)
{
    const int64_t tot = diag_scratch_doubles(m->g, m->cap);
    if (!m->diag_d) {
        HIPOK(m, dalloc(&m->diag_d, tot, m->s));
        HIPOK(m, hipHostMalloc((void **)&m->diag_h, sizeof(double) * 16, hipHostMallocDefault));
    }
    double *res = m->diag_d + tot - 16;
    if (parcels) launch_parcel_diagnostics(m->g, m->pd, nullptr, m->diag_d, res, m->err_d, m->s);
    else launch_field_diagnostics(m->g, m->fld, m->nparg, m->nsparg, m->diag_d, res, m->s);
    HIPOK(m, hipGetLastError());
    HIPOK(m, hipMemcpyAsync(m->diag_h, res, sizeof(double) * nv, hipMemcpyDeviceToHost, m->s));
    const int rc = check(m, where);   // synchronises
    if (rc) return rc;
    for (int c = 0; c < nv; ++c) out[c] = m->diag_h[c];
    return EPIC_OK;
}

int epic_parcel_diagnostics(epic_handle m, double *stats, int reset_counters)
{
    if (!m || !stats) return EPIC_E_INVALID;
    for (int i = 0; i < 18; ++i) stats[i] = 0.0;
    {   // the velocity is each parcel's delta_pos
        const int rc = settle_tendencies(m);
        if (rc) return rc;
    }
    if (m->n > 0) {
        double s[12];
        const int rc = diag_run(m, true, s, 12, "epic_parcel_diagnostics");
        if (rc) return rc;
        parcel_stats_of(m, s, stats);
    }
    parcel_counts_of(m, stats, reset_counters);
    return EPIC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Parcel diagnostics without allocation: the block partials and totals live
// in one file-scope device array (each GPU holds its own copy until the
// process ends), whose address on the current device is looked up once per
// device with hipGetSymbolAddress.  Every call writes every partial it then
// reads, so nothing needs clearing between calls; calls that overlap on one
// GPU share the array and race.
// ---------------------------------------------------------------------------
namespace {
constexpr int NOALLOC_BLOCKS = 1024;                                  // partial blocks (grid-stride over the parcels)
constexpr int NOALLOC_TOTALS = 16;                                    // the totals after the partials
constexpr int NOALLOC_DOUBLES = NOALLOC_BLOCKS * 12 + NOALLOC_TOTALS;  // 12 partials per block: 98,432 bytes
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex scratch_mu;
double *scratch_dev[NOALLOC_MAX_DEVICES] = {};
}  // namespace

static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

// the device address of scratch on the current device (host code never
// passes the array's name itself: that is a host-side stand-in)
static int scratch_on_device(double **out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) return EPIC_E_HIP;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return EPIC_E_HIP;
        scratch_dev[dev] = (double *)p;
    }
    *out = scratch_dev[dev];
    return EPIC_OK;
}

extern "C" {

int epic_parcel_diagnostics_noalloc(epic_handle m, double *stats, int reset_counters)
{
    if (!m || !stats) return EPIC_E_INVALID;
    if (m->n > EPIC_MAX_PARCELS)
        return fail(m, EPIC_E_INVALID, "epic_parcel_diagnostics_noalloc: more parcels than EPIC_MAX_PARCELS");
    for (int i = 0; i < 18; ++i) stats[i] = 0.0;
    if (m->n > 0) {
        double *sp = nullptr;
        if (scratch_on_device(&sp) != EPIC_OK)
            return fail(m, EPIC_E_HIP, "epic_parcel_diagnostics_noalloc: no device address for the scratch array");
        // tendencies a re-sort left in their slots are read through tend_src
        // (in place of epic_parcel_diagnostics' gather, which allocates)
        const int *src = m->tend_displaced ? m->tend_src : nullptr;
        double *res = sp + NOALLOC_BLOCKS * 12;
        launch_parcel_diagnostics(m->g, m->pd, src, sp, res, m->err_d, m->s, NOALLOC_BLOCKS);
        HIPOK(m, hipGetLastError());
        double s[12];
        HIPOK(m, hipMemcpyAsync(s, res, sizeof(double) * 12, hipMemcpyDeviceToHost, m->s));
        const int rc = check(m, "epic_parcel_diagnostics_noalloc");   // synchronises
        if (rc) return rc;
        parcel_stats_of(m, s, stats);
    }
    parcel_counts_of(m, stats, reset_counters);
    return EPIC_OK;
}

int epic_field_diagnostics(epic_handle m, double *stats)
{
    if (!m || !stats) return EPIC_E_INVALID;
    const Grid &g = m->g;
    double s[10];
    const int rc = diag_run(m, false, s, 10, "epic_field_diagnostics");
    if (rc) return rc;
    const double vdomaini = 1.0 / (g.extent[0] * g.extent[1] * g.extent[2]);
    const double ngridi = 1.0 / ((double)g.nx * g.ny * (g.nz + 1)), ncelli = 1.0 / ((double)g.nx * g.ny * g.nz);
    const double vcelli = 1.0 / g.vcell;
    for (int i = 0; i < 11; ++i) stats[i] = 0.0;
    stats[0] = std::sqrt(s[0] * ngridi) * vcelli;
    stats[1] = s[1] * ncelli;
    stats[2] = s[2] * ncelli;
    stats[3] = s[3] * vdomaini;
    stats[5] = s[4] * vdomaini;
    stats[6] = s[5] * vcelli;
    stats[7] = s[6];
    stats[8] = s[7];
    stats[9] = s[8];
    stats[10] = s[9];
    return EPIC_OK;
}

// ---------------------------------------------------------------------------
// parcel set-up from gridded fields and restart (parcel_init.f90,
// utils.f90:144-249)
// ---------------------------------------------------------------------------
namespace {

// init_parcels_from_grids (parcel_init.f90:174-203): every parcel adds the
// trilinear interpolation (interpl.f90:17-56, FP64 weights, the sum in the
// reference's element order) of vortg, tbuoyg and humg at its centre to
// its vorticity, buoyancy and humidity
__global__ __launch_bounds__(256) void k_init_from_grids(Grid g, ParcelsDev p, const float *__restrict__ fld, int *err)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= p.n) return;
    const double pos[3] = {p.x[q], p.y[q], p.z[q]};
    int ijk[3];
    double fr[3];
    for (int d = 0; d < 3; ++d) {
        const double xyz = (pos[d] - g.lower[d]) * g.dxi[d];
        const double fl = floor(xyz);
        ijk[d] = (int)fl;
        fr[d] = xyz - fl;
    }
    if (ijk[2] < -1 || ijk[2] > g.nz) {
        atomicOr(err, ERR_OUT_OF_GRID);
        return;
    }
    const int64_t FS = field_stride(g), PS = plane_stride(g);
    int64_t off[8];
    double w[8];
    // Fortran element order of weights(0:1, 0:1, 0:1): k fastest, then j, then i
    for (int ii = 0; ii < 2; ++ii)
        for (int jj = 0; jj < 2; ++jj)
            for (int kk = 0; kk < 2; ++kk) {
                const int m = (ii * 2 + jj) * 2 + kk;
                const int i = wrap(ijk[0] + ii, g.nx), j = wrap(ijk[1] + jj, g.ny);
                off[m] = (int64_t)(ijk[2] + kk + 1) * PS + (int64_t)j * g.nx + i;
                w[m] = (ii ? fr[0] : 1.0 - fr[0]) * (jj ? fr[1] : 1.0 - fr[1]) * (kk ? fr[2] : 1.0 - fr[2]);
            }
    auto interp = [&](int slot) {
        double s = 0.0;
        for (int m = 0; m < 8; ++m) s += w[m] * (double)fld[slot * FS + off[m]];
        return s;
    };
    for (int c = 0; c < 3; ++c) p.vor[c][q] = (float)((double)p.vor[c][q] + interp(F_VORTG + c));
    p.buoy[q] = (float)((double)p.buoy[q] + interp(F_TBUOYG));
    p.hum[q] = (float)((double)p.hum[q] + interp(F_HUMG));
}

}  // namespace

int epic_parcel_default(epic_handle m, int n_per_cell, double lambda_max)
{
    if (!m) return EPIC_E_INVALID;
    const Grid &g = m->g;
    if (n_per_cell <= 0) n_per_cell = 8;   // options.f90:88
    if (!(lambda_max > 0.0)) lambda_max = 4.0;   // options.f90:89
    int npd = (int)std::lround(std::cbrt((double)n_per_cell));
    if (npd * npd * npd != n_per_cell)
        return fail(m, EPIC_E_INVALID, "epic_parcel_default: Number of parcels per cell not a cubic.");
    const int64_t ncell = (int64_t)g.nx * g.ny * g.nz, n = n_per_cell * ncell;
    if (n > m->cap) return fail(m, EPIC_E_CAPACITY, "epic_parcel_default: Number of parcels exceeds limit of max_parcels.");
    // init_regular_positions (parcel_init.f90:104-152): cells iz, iy, ix
    // (ix fastest), then k, j, i in the cell (i fastest) at (i - 1/2)/npd
    std::vector<double> pos(3 * (size_t)n), vor(3 * (size_t)n, 0.0), B(5 * (size_t)n, 0.0), vol((size_t)n),
        zero((size_t)n, 0.0);
    const double im = 1.0 / npd;
    // aspect ratio lam = a / c from the cell (parcel_init.f90:60-63)
    const double lam = std::max({g.dx[0] / g.dx[1], g.dx[1] / g.dx[0], g.dx[0] / g.dx[2], g.dx[2] / g.dx[0],
                                 g.dx[1] / g.dx[2], g.dx[2] / g.dx[1]});
    const double v = g.vcell / n_per_cell;
    const double abc = 0.75 * v / 3.14159265358979323846;   // get_abc (parcel_ellipsoid.f90:409-415)
    const double l23 = std::pow(lam * abc, 2.0 / 3.0);
    host_parallel_for(ncell, [&](int64_t lo, int64_t hi) {
        for (int64_t c = lo; c < hi; ++c) {
            const int ix = (int)(c % g.nx), iy = (int)((c / g.nx) % g.ny), iz = (int)(c / ((int64_t)g.nx * g.ny));
            const double corner[3] = {g.lower[0] + ix * g.dx[0], g.lower[1] + iy * g.dx[1], g.lower[2] + iz * g.dx[2]};
            int64_t l = c * n_per_cell;
            for (int k = 1; k <= npd; ++k)
                for (int j = 1; j <= npd; ++j)
                    for (int i = 1; i <= npd; ++i, ++l) {
                        pos[3 * l + 0] = corner[0] + g.dx[0] * (i - 0.5) * im;
                        pos[3 * l + 1] = corner[1] + g.dx[1] * (j - 0.5) * im;
                        pos[3 * l + 2] = corner[2] + g.dx[2] * (k - 0.5) * im;
                        vol[l] = v;
                        B[5 * l + 0] = l23;
                        B[5 * l + 3] = l23;
                    }
        }
    });
    int rc = epic_set_parcels(m, n, pos.data(), vor.data(), B.data(), vol.data(), zero.data(), zero.data());
    if (rc) return rc;
    // init_refine (parcel_init.f90:156-166): split while the aspect ratio
    // reaches lambda_max; every parcel has the same shape
    double lr = lam;
    while (lr >= lambda_max) {
        int64_t ns = 0;
        rc = epic_parcel_split(m, lambda_max, -1.0, &ns);
        if (rc) return rc;
        if (ns == 0) break;
        // the eigenvalues of parcel 1's B (get_eigenvalues): from its matrix
        std::vector<double> all_b(5 * (size_t)m->n), all_p(3 * (size_t)m->n), all_v(3 * (size_t)m->n),
            all_vol((size_t)m->n), all_bu((size_t)m->n), all_h((size_t)m->n);
        rc = epic_get_parcels(m, m->n, all_p.data(), all_v.data(), all_b.data(), all_vol.data(), all_bu.data(),
                              all_h.data());
        if (rc) return rc;
        const double *bb = all_b.data();
        const double a2 = 0.75 * all_vol[0] / 3.14159265358979323846;
        const double den = bb[0] * bb[3] - bb[1] * bb[1];
        const double b33 = (a2 * a2 - bb[2] * (bb[1] * bb[4] - bb[2] * bb[3]) + bb[0] * bb[4] * bb[4] -
                            bb[1] * bb[2] * bb[4]) / den;
        // largest / smallest eigenvalue of the symmetric 3x3 matrix
        double A[9] = {bb[0], bb[1], bb[2], bb[1], bb[3], bb[4], bb[2], bb[4], b33}, D[3], V[9];
        rc = epic_diagonalise(1, A, D, V);
        if (rc) return fail(m, rc, epic_diagonalise_last_error());
        lr = std::sqrt(D[0] / D[2]);
    }
    return EPIC_OK;
}

int epic_init_parcels_from_grids(epic_handle m)
{
    if (!m) return EPIC_E_INVALID;
    if (m->n == 0) return EPIC_OK;
    const int rc = settle_tendencies(m);
    if (rc) return rc;
    hipLaunchKernelGGL(k_init_from_grids, dim3((unsigned)((m->n + 255) / 256)), dim3(256), 0, m->s, m->g, m->pd,
                       (const float *)m->fld, m->err_d);
    HIPOK(m, hipGetLastError());
    return check(m, "epic_init_parcels_from_grids");
}

double epic_next_write_time(double freq, double t)
{
    // get_next_write_time (utils.f90:48-54): freq (int(t / freq) + 1)
    return freq * (double)((long long)(t / freq) + 1);
}

int epic_setup_restart(epic_handle m, const char *path, double *t, char *file_type)
{
    if (!m || !path) return EPIC_E_INVALID;
    char ft[16];
    double tt = 0.0;
    // setup_restart (utils.f90:144-160): the file type and the last time
    int rc = epic_nc_file_info(path, ft, &tt, nullptr, nullptr);
    if (rc) return fail(m, rc, std::string("epic_setup_restart: ") + epic_nc_last_error());
    if (file_type) std::memcpy(file_type, ft, sizeof(ft));
    if (t) *t = tt;
    // setup_fields_and_parcels (utils.f90:200-216)
    if (std::strcmp(ft, "fields") == 0) {
        rc = epic_parcel_default(m, 8, 4.0);
        if (rc) return rc;
        rc = epic_read_fields_nc(m, path, -1, nullptr);
        if (rc) return rc;
        return epic_init_parcels_from_grids(m);
    }
    if (std::strcmp(ft, "parcels") == 0) return epic_read_parcels_nc(m, path, nullptr);
    return fail(m, EPIC_E_INVALID, "epic_setup_restart: Restart file must be of type \"fields\" or \"parcels\".");
}

int epic_get_config(epic_handle m, epic_grid_t *grid, epic_phys_t *phys, epic_flags_t *flags)
{
    if (!m) return EPIC_E_INVALID;
    const Grid &g = m->g;
    if (grid) {
        grid->nx = g.nx;
        grid->ny = g.ny;
        grid->nz = g.nz;
        for (int d = 0; d < 3; ++d) {
            grid->lower[d] = g.lower[d];
            grid->extent[d] = g.extent[d];
        }
    }
    if (phys) {
        phys->glat = g.glat;
        phys->lambda_c = g.lambda_c;
        phys->q0 = g.q0;
        for (int d = 0; d < 3; ++d) phys->f_cor[d] = g.f_cor[d];
    }
    if (flags) {
        flags->zeta_bndry_zero[0] = g.zeta_zero[0];
        flags->zeta_bndry_zero[1] = g.zeta_zero[1];
        flags->rk_order = m->rk_order;
        flags->vmin = g.vmin;
    }
    return EPIC_OK;
}

int epic_cell_sort(epic_handle m, int *loca, int *nppc, int *kc1, int *node)
{
    if (!m) return EPIC_E_INVALID;
    const Grid &g = m->g;
    const int64_t ncell = (int64_t)(g.nx + 3) * (g.ny + 3) * (g.nz + 2), n = m->n;
    int *keys = nullptr, *cnt = nullptr, *start = nullptr, *cur = nullptr, *ptmp = nullptr, *pm = nullptr;
    HIPOK(m, dalloc(&keys, n, m->s));
    HIPOK(m, dalloc(&cnt, ncell, m->s));
    HIPOK(m, dalloc(&start, ncell + 1, m->s));
    HIPOK(m, dalloc(&cur, ncell, m->s));
    HIPOK(m, dalloc(&ptmp, n, m->s));
    HIPOK(m, dalloc(&pm, n, m->s));
    hipLaunchKernelGGL(k_nearest_keys, dim3(grid_blocks(std::max<int64_t>(n, 1))), dim3(256), 0, m->s, g, m->pd, keys,
                       cnt, ncell, m->err_d);
    {
        Plan P;
        add_exclusive_scan(P, PH_OTHER, {ScanArgs{cnt, ncell, start, m->bsum, cur}});
        add_place(P, PH_OTHER, {PlaceArgs{n, keys, start, cur, ptmp}});
        HIPOK(m, P.run(m->s));
        launch_cell_order(ncell, start, ptmp, m->orig, pm, m->s);
        HIPOK(m, hipStreamSynchronize(m->s));
    }
    int rc = check(m, "epic_cell_sort");
    std::vector<int> hk(n), hp(n), ho(n), hs(ncell + 1), hc(ncell);
    if (rc == EPIC_OK) {
        HIPOK(m, hipMemcpyAsync(hk.data(), keys, sizeof(int) * n, hipMemcpyDeviceToHost, m->s));
        HIPOK(m, hipMemcpyAsync(hp.data(), pm, sizeof(int) * n, hipMemcpyDeviceToHost, m->s));
        HIPOK(m, hipMemcpyAsync(ho.data(), m->orig, sizeof(int) * n, hipMemcpyDeviceToHost, m->s));
        HIPOK(m, hipMemcpyAsync(hs.data(), start, sizeof(int) * (ncell + 1), hipMemcpyDeviceToHost, m->s));
        HIPOK(m, hipMemcpyAsync(hc.data(), cnt, sizeof(int) * ncell, hipMemcpyDeviceToHost, m->s));
        HIPOK(m, hipStreamSynchronize(m->s));
        for (int64_t q = 0; q < n; ++q) {
            if (loca) loca[ho[q]] = hk[q] + 1;
            if (node) node[q] = ho[hp[q]];
        }
        for (int64_t c = 0; c < ncell; ++c) {
            if (nppc) nppc[c] = hc[c];
            if (kc1) kc1[c] = hs[c] + 1;
        }
    }
    (void)hipFree(keys);
    (void)hipFree(cnt);
    (void)hipFree(start);
    (void)hipFree(cur);
    (void)hipFree(ptmp);
    (void)hipFree(pm);
    return rc;
}

}  // extern "C"

// failure message on a handle for the host-side file I/O (ncio.cpp)
int epic_internal_fail(epic_handle m, int code, const std::string &what) { return fail(m, code, what); }
