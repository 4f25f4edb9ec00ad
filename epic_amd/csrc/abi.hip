// C-ABI of the parcel-grid core: member state, host <-> device conversion of
// the reference's Fortran layouts, and the per-call orchestration of the
// kernels (the reference's call surface: par2grid, vor2vel,
// vorticity_tendency, grid2par, ls_rk_step).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "epic_gpu.h"
#include "common.h"
#include "fft.h"
#include "launch.h"

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
    // sort
    int *keys = nullptr, *count = nullptr, *scount = nullptr, *count2 = nullptr, *scount2 = nullptr;
    int *cell_start = nullptr, *cursor = nullptr, *perm_tmp = nullptr, *perm = nullptr, *bsum = nullptr;
    bool keys_valid = false;
    // fields
    float *fld = nullptr;
    int *nparg = nullptr, *nsparg = nullptr;
    float *staging = nullptr;
    // spectra
    float2 *spec[NSPEC] = {};
    float2 *T = nullptr;
    float2 *twx = nullptr, *twy = nullptr;
    double2 *twz = nullptr;
    FftPlan px{}, py{}, pz{};
    // per-phase GPU timers (events on the member stream)
    bool profiling = false;
    std::vector<hipEvent_t> ev;        // pairs (start, stop) recorded in order
    std::vector<int> ev_phase;
    double timers_ms[EPIC_NUM_TIMERS] = {};
    int timer_calls[EPIC_NUM_TIMERS] = {};
    // errors
    int *err_d = nullptr;
    int *err_h = nullptr;
    std::string msg;
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

template <typename T> hipError_t dalloc(T **p, int64_t count)
{
    *p = nullptr;
    if (count <= 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * (size_t)count);
    if (e == hipSuccess) e = hipMemset(*p, 0, sizeof(T) * (size_t)count);
    return e;
}

bool factor23(int n)
{
    if (n < 1) return false;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    return n == 1;
}

void bind_parcels(epic_member *m)
{
    ParcelsDev &p = m->pd;
    const int64_t c = m->cap;
    p.n = m->n;
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
}

void collect_timers(epic_member *m);

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
    std::string w(where);
    if (e & ERR_PARCEL_COUNT) return fail(m, EPIC_E_PARCEL_COUNT, w + ": par2grid: Wrong total number of parcels!");
    if (e & ERR_B33) return fail(m, EPIC_E_B33, w + ": get_B33: Division by small number!");
    if (e & ERR_OUT_OF_GRID) return fail(m, EPIC_E_OUT_OF_GRID, w + ": support point outside the node grid");
    return fail(m, EPIC_E_INVALID, w + ": internal consistency error");
}

__global__ void k_fill_int(int *p, int64_t n, int v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
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

// phase timers: record a (start, stop) event pair around a phase
struct PhaseTimer {
    epic_member *m;
    int phase;
    hipEvent_t stop = nullptr;
    PhaseTimer(epic_member *mm, int ph) : m(mm), phase(ph)
    {
        if (!m->profiling) return;
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
        (void)hipEventRecord(a, m->s);
        m->ev.push_back(a);
        m->ev.push_back(b);
        m->ev_phase.push_back(ph);
        stop = b;
    }
    ~PhaseTimer()
    {
        if (stop) (void)hipEventRecord(stop, m->s);
    }
};

// after a stream synchronisation: fold the recorded pairs into the totals
void collect_timers(epic_member *m)
{
    for (size_t k = 0; k < m->ev_phase.size(); ++k) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, m->ev[2 * k], m->ev[2 * k + 1]) == hipSuccess) {
            m->timers_ms[m->ev_phase[k]] += ms;
            m->timer_calls[m->ev_phase[k]] += 1;
        }
        (void)hipEventDestroy(m->ev[2 * k]);
        (void)hipEventDestroy(m->ev[2 * k + 1]);
    }
    m->ev.clear();
    m->ev_phase.clear();
}

// ---------------------------------------------------------------------------
// orchestration
// ---------------------------------------------------------------------------
void ensure_keys(epic_member *m)
{
    if (m->keys_valid) return;
    const int64_t nc = ncells(m->g);
    launch_fill_int(m->count, nc, 0, m->s);
    launch_fill_int(m->scount, nc, 0, m->s);
    launch_cell_keys(m->g, m->pd, m->keys, m->count, m->scount, m->err_d, m->s);
    m->keys_valid = true;
}

// cell sort: histogram -> offsets -> placement (order inside a cell is
// arbitrary here; par2grid orders each cell by caller index in LDS)
void run_sort(epic_member *m)
{
    const int64_t nc = ncells(m->g);
    launch_exclusive_scan(m->count, nc, m->cell_start, m->bsum, m->s);
    launch_fill_int(m->cursor, nc, 0, m->s);
    launch_place(m->n, m->keys, m->cell_start, m->cursor, m->perm, m->s);
}

// par2grid with the given half axes (cur: this stage's, stale: the cached set)
void run_par2grid(epic_member *m, float *const *axes)
{
    PhaseTimer t(m, EPIC_TIMER_PAR2GRID);
    {
        PhaseTimer t2(m, EPIC_TIMER_CELL_SORT);
        ensure_keys(m);
        run_sort(m);
    }
    launch_p2g(m->g, m->pd, axes, m->perm, m->cell_start, m->count, m->scount, m->staging, m->fld, m->nparg,
               m->nsparg, m->err_d, m->s);
}

void run_axes(epic_member *m, float *const *axes)
{
    PhaseTimer t(m, EPIC_TIMER_AXES);
    launch_axes(m->g, m->pd, axes, m->err_d, m->s);
}

FftXArgs xargs(epic_member *m)
{
    FftXArgs a{};
    a.g = m->g;
    a.plan = m->px;
    a.tw = m->twx;
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
    return a;
}

void run_vor2vel(epic_member *m)
{
    PhaseTimer t(m, EPIC_TIMER_VOR2VEL);
    for (int c = 0; c < 3; ++c) {
        FftXArgs x = xargs(m);
        x.mode = 0;
        x.slot = F_VORTG + c;
        x.out = m->T;
        launch_fft_xf(x, m->s);
        FftYArgs y = yargs(m);
        y.mode = 0;
        y.in = m->T;
        y.out = m->spec[c];
        launch_fft_y(y, m->s);
    }
    ColArgs ca{};
    ca.g = m->g;
    ca.zplan = m->pz;
    ca.ztw = m->twz;
    for (int c = 0; c < 3; ++c) ca.in[c] = m->spec[c];
    for (int o = 0; o < 11; ++o) ca.out[o] = m->spec[3 + o];
    launch_vor2vel_columns(ca, m->s);
    static const int slots[11] = {F_VORTG + 0, F_VORTG + 1, F_VORTG + 2, F_VELOG + 0, F_VELOG + 1, F_VELOG + 2,
                                  F_VELGRADG + DUDX, F_VELGRADG + DUDY, F_VELGRADG + DVDY, F_VELGRADG + DWDX,
                                  F_VELGRADG + DWDY};
    for (int o = 0; o < 11; ++o) {
        FftYArgs y = yargs(m);
        y.mode = 1;
        y.in = m->spec[3 + o];
        y.out = m->T;
        launch_fft_y(y, m->s);
        FftXArgs x = xargs(m);
        x.mode = 0;
        x.slot = slots[o];
        x.in = m->T;
        launch_fft_xi(x, m->s);
    }
    launch_vor2vel_fixup(m->g, m->fld, m->s);
}

void run_vtend(epic_member *m)
{
    PhaseTimer t(m, EPIC_TIMER_VTEND);
    for (int c = 0; c < 3; ++c) {
        for (int a = 0; a < 2; ++a) {
            FftXArgs x = xargs(m);
            x.mode = 1;
            x.comp_a = a;
            x.comp_b = c;
            // f_x of the y tendency carries -b, f_y of the x tendency +b (inversion.f90:327-335)
            x.sgn = (a == 0 && c == 1) ? -1 : ((a == 1 && c == 0) ? 1 : 0);
            x.out = m->T;
            launch_fft_xf(x, m->s);
            FftYArgs y = yargs(m);
            y.mode = 0;
            y.in = m->T;
            y.out = m->spec[2 * c + a];
            launch_fft_y(y, m->s);
        }
    }
    for (int c = 0; c < 3; ++c) {
        FftYArgs y = yargs(m);
        y.mode = 2;
        y.in = m->spec[2 * c];
        y.in2 = m->spec[2 * c + 1];
        y.out = m->T;
        launch_fft_y(y, m->s);
        FftXArgs x = xargs(m);
        x.mode = 2;
        x.comp_b = c;
        x.slot = F_VTEND + c;
        x.in = m->T;
        launch_fft_xi(x, m->s);
    }
    launch_vtend_fixup(m->g, m->fld, m->s);
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
void launch_fill_int(int *p, int64_t n, int v, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_int, dim3(grid_blocks(n)), dim3(256), 0, s, p, n, v);
}
}  // namespace epic

// ===========================================================================
// exported entry points
// ===========================================================================
extern "C" {

int epic_create(const epic_grid_t *grid, const epic_phys_t *phys, const epic_flags_t *flags, int64_t max_parcels,
                epic_handle *out)
{
    if (!grid || !phys || !flags || !out || max_parcels < 0) return EPIC_E_INVALID;
    *out = nullptr;
    if (grid->nx < 8 || grid->ny < 8 || grid->nx % 8 || grid->ny % 8 || grid->nz < 4) return EPIC_E_INVALID;
    if (!factor23(grid->nx / 2) || !factor23(grid->ny) || !factor23(2 * grid->nz)) return EPIC_E_INVALID;
    if (flags->rk_order != 3 && flags->rk_order != 4) return EPIC_E_RK_ORDER;
    for (int i = 0; i < 3; ++i)
        if (!(grid->extent[i] > 0.0)) return EPIC_E_INVALID;
    if (max_parcels > (int64_t)0x7fffffff) return EPIC_E_INVALID;

    epic_member *m = new epic_member();
    Grid &g = m->g;
    g.nx = grid->nx;
    g.ny = grid->ny;
    g.nz = grid->nz;
    g.nxh = g.nx / 2 + 1;
    g.nzp = ((g.nz + 1 + ZB - 1) / ZB) * ZB;
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

    auto bad = [&](hipError_t e) {
        if (e != hipSuccess) {
            epic_destroy(m);
            return true;
        }
        return false;
    };
    if (bad(hipStreamCreate(&m->s))) return EPIC_E_HIP;
    const int64_t c = m->cap, nc = ncells(g);
    if (bad(dalloc(&m->pos, 3 * c)) || bad(dalloc(&m->attr, (int64_t)NATTR * c)) || bad(dalloc(&m->orig, c)) ||
        bad(dalloc(&m->axbuf, 12 * c)) ||
        bad(dalloc(&m->keys, c)) || bad(dalloc(&m->perm_tmp, c)) || bad(dalloc(&m->perm, c)) ||
        bad(dalloc(&m->count, nc)) || bad(dalloc(&m->scount, nc)) || bad(dalloc(&m->count2, nc)) ||
        bad(dalloc(&m->scount2, nc)) || bad(dalloc(&m->cell_start, nc + 1)) || bad(dalloc(&m->cursor, nc)) ||
        bad(dalloc(&m->bsum, scan_scratch_ints(std::max<int64_t>(nc, (int64_t)(g.nx + 3) * (g.ny + 3) * (g.nz + 2))))))
        return EPIC_E_NOMEM;
    const int64_t FS = field_stride(g);
    if (bad(dalloc(&m->fld, (int64_t)F_COUNT * FS)) || bad(dalloc(&m->nparg, FS)) || bad(dalloc(&m->nsparg, FS)) ||
        bad(dalloc(&m->staging, p2g_layout(g).staging_floats)))
        return EPIC_E_NOMEM;
    for (int i = 0; i < NSPEC; ++i)
        if (bad(dalloc(&m->spec[i], spec_size(g)))) return EPIC_E_NOMEM;
    if (bad(dalloc(&m->T, spec_size(g))) || bad(dalloc(&m->err_d, 4))) return EPIC_E_NOMEM;
    if (bad(hipHostMalloc((void **)&m->err_h, sizeof(int) * 4, hipHostMallocDefault))) return EPIC_E_NOMEM;
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
    if (bad(dalloc(&m->twx, g.nx)) || bad(dalloc(&m->twy, g.ny)) || bad(dalloc(&m->twz, 2 * g.nz))) return EPIC_E_NOMEM;
    if (bad(hipMemcpy(m->twx, tx.data(), sizeof(float2) * tx.size(), hipMemcpyHostToDevice)) ||
        bad(hipMemcpy(m->twy, ty.data(), sizeof(float2) * ty.size(), hipMemcpyHostToDevice)) ||
        bad(hipMemcpy(m->twz, tz.data(), sizeof(double2) * tz.size(), hipMemcpyHostToDevice)))
        return EPIC_E_HIP;
    for (int i = 0; i < 6; ++i) {
        m->ax_cur[i] = m->axbuf + i * c;
        m->ax_stale[i] = m->axbuf + (6 + i) * c;
    }
    bind_parcels(m);
    *out = m;
    return EPIC_OK;
}

int epic_destroy(epic_handle m)
{
    if (!m) return EPIC_OK;
    if (m->s) (void)hipStreamSynchronize(m->s);
    void *ptrs[] = {m->pos, m->attr, m->orig, m->axbuf, m->keys, m->count, m->scount, m->count2, m->scount2, m->cell_start,
                    m->cursor, m->perm_tmp, m->perm, m->bsum, m->fld, m->nparg, m->nsparg, m->staging, m->T,
                    m->twx, m->twy, m->twz, m->err_d};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (int i = 0; i < NSPEC; ++i)
        if (m->spec[i]) (void)hipFree(m->spec[i]);
    if (m->err_h) (void)hipHostFree(m->err_h);
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
    for (int64_t q = 0; q < n; ++q) {
        for (int d = 0; d < 3; ++d) hp[d * n + q] = position[3 * q + d];
        for (int d = 0; d < 3; ++d) ha[d * n + q] = (float)vorticity[3 * q + d];
        for (int d = 0; d < 5; ++d) ha[(3 + d) * n + q] = (float)B[5 * q + d];
        ha[8 * n + q] = (float)volume[q];
        ha[9 * n + q] = (float)buoyancy[q];
        ha[10 * n + q] = (float)humidity[q];
        ho[q] = (int)q;
    }
    m->n = n;
    bind_parcels(m);
    for (int d = 0; d < 3 && n; ++d)
        HIPOK(m, hipMemcpyAsync(m->pos + d * c, hp.data() + d * n, sizeof(double) * n, hipMemcpyHostToDevice, m->s));
    for (int a = 0; a < 11 && n; ++a)
        HIPOK(m, hipMemcpyAsync(m->attr + a * c, ha.data() + a * n, sizeof(float) * n, hipMemcpyHostToDevice, m->s));
    if (n) HIPOK(m, hipMemcpyAsync(m->orig, ho.data(), sizeof(int) * n, hipMemcpyHostToDevice, m->s));
    HIPOK(m, hipStreamSynchronize(m->s));
    m->keys_valid = false;
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
    for (int64_t q = 0; q < n; ++q) {
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
    return EPIC_OK;
}

int epic_get_parcel_tendencies(epic_handle m, int64_t n, double *delta_pos, double *delta_vor, double *strain)
{
    if (!m) return EPIC_E_INVALID;
    if (n != m->n) return fail(m, EPIC_E_INVALID, "epic_get_parcel_tendencies: n differs from the parcel count");
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
    const int64_t c = m->cap;
    std::vector<int> ho(n);
    if (n) HIPOK(m, hipMemcpy(ho.data(), m->orig, sizeof(int) * n, hipMemcpyDeviceToHost));
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
    // get_points stores the axes it computes (parcel_ellipsoid.f90:216-247)
    if (l_reuse != 1) run_axes(m, m->ax_stale);
    run_par2grid(m, m->ax_stale);
    return check(m, "epic_par2grid");
}

int epic_vor2vel(epic_handle m)
{
    if (!m) return EPIC_E_INVALID;
    run_vor2vel(m);
    return check(m, "epic_vor2vel");
}

int epic_vorticity_tendency(epic_handle m)
{
    if (!m) return EPIC_E_INVALID;
    run_vtend(m);
    return check(m, "epic_vorticity_tendency");
}

int epic_grid2par(epic_handle m, int add)
{
    if (!m) return EPIC_E_INVALID;
    {
        PhaseTimer t(m, EPIC_TIMER_GRID2PAR);
        launch_grid2par(m->g, m->pd, m->fld, add == 1, m->err_d, m->s);
    }
    return check(m, "epic_grid2par");
}

int epic_ls_rk_step(epic_handle m, double *t, double dt, double t_initial)
{
    if (!m || !t) return EPIC_E_INVALID;
    if (!(dt > 0.0)) return fail(m, EPIC_E_DT, "epic_ls_rk_step: time step must be positive");
    const double *ca = (m->rk_order == 4) ? CAS4 : CAS3;
    const double *cb = (m->rk_order == 4) ? CBS4 : CBS3;
    const int ns = (m->rk_order == 4) ? 5 : 3;
    const int64_t nc = ncells(m->g);
    for (int s = 1; s <= ns; ++s) {
        // the first par2grid of a step reuses the cached axes (ls_rk.f90:93);
        // grid2par and later stages use the axes of the current B
        const bool reuse = (s == 1 && *t > t_initial);
        run_axes(m, m->ax_cur);
        run_par2grid(m, reuse ? m->ax_stale : m->ax_cur);
        run_vor2vel(m);
        run_vtend(m);
        launch_fill_int(m->count2, nc, 0, m->s);
        launch_fill_int(m->scount2, nc, 0, m->s);
        {
            PhaseTimer t(m, EPIC_TIMER_PUSH);
            launch_rk_stage_tiles(m->g, m->pd, m->ax_cur, m->perm, m->cell_start, m->fld, s, ns, cb[s - 1] * dt,
                                  ca[s - 1], s == ns ? 1 : 0, m->keys, m->count2, m->scount2, m->err_d, m->s);
        }
        std::swap(m->count, m->count2);
        std::swap(m->scount, m->scount2);
        m->keys_valid = true;
    }
    // the axes of the last stage become the cache of the next step's first
    // par2grid (grid2par stores them, parcel_interpl.f90:442)
    for (int i = 0; i < 6; ++i) std::swap(m->ax_cur[i], m->ax_stale[i]);
    bind_parcels(m);
    const int rc = check(m, "epic_ls_rk_step");
    if (rc == EPIC_OK) *t = *t + dt;
    return rc;
}

int epic_set_profiling(epic_handle m, int on)
{
    if (!m) return EPIC_E_INVALID;
    m->profiling = (on != 0);
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

int epic_cell_sort(epic_handle m, int *loca, int *nppc, int *kc1, int *node)
{
    if (!m) return EPIC_E_INVALID;
    const Grid &g = m->g;
    const int64_t ncell = (int64_t)(g.nx + 3) * (g.ny + 3) * (g.nz + 2), n = m->n;
    int *keys = nullptr, *cnt = nullptr, *start = nullptr, *cur = nullptr, *ptmp = nullptr, *pm = nullptr;
    HIPOK(m, dalloc(&keys, n));
    HIPOK(m, dalloc(&cnt, ncell));
    HIPOK(m, dalloc(&start, ncell + 1));
    HIPOK(m, dalloc(&cur, ncell));
    HIPOK(m, dalloc(&ptmp, n));
    HIPOK(m, dalloc(&pm, n));
    hipLaunchKernelGGL(k_nearest_keys, dim3(grid_blocks(std::max<int64_t>(n, 1))), dim3(256), 0, m->s, g, m->pd, keys,
                       cnt, ncell, m->err_d);
    launch_exclusive_scan(cnt, ncell, start, m->bsum, m->s);
    launch_place(n, keys, start, cur, ptmp, m->s);
    launch_cell_order(ncell, start, ptmp, m->orig, pm, m->s);
    int rc = check(m, "epic_cell_sort");
    std::vector<int> hk(n), hp(n), ho(n), hs(ncell + 1), hc(ncell);
    if (rc == EPIC_OK) {
        HIPOK(m, hipMemcpy(hk.data(), keys, sizeof(int) * n, hipMemcpyDeviceToHost));
        HIPOK(m, hipMemcpy(hp.data(), pm, sizeof(int) * n, hipMemcpyDeviceToHost));
        HIPOK(m, hipMemcpy(ho.data(), m->orig, sizeof(int) * n, hipMemcpyDeviceToHost));
        HIPOK(m, hipMemcpy(hs.data(), start, sizeof(int) * (ncell + 1), hipMemcpyDeviceToHost));
        HIPOK(m, hipMemcpy(hc.data(), cnt, sizeof(int) * ncell, hipMemcpyDeviceToHost));
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
