// Shared definitions of the MI355X parcel-grid core (gfx950, HIP).
//
// Device data layouts (one member = one EPIC domain):
//   parcels  structure of arrays in "physical" order; positions FP64,
//            every other attribute FP32 (see ParcelsDev)
//   fields   FP32 node fields f[k+1][j][i], k in [-1, nz+1], j in [0, ny),
//            i in [0, nx) (x fastest, x/y periodic so no stored x/y halo)
//   spectra  complex FP32 s[ky][kx][z], z in [0, nzp) padded to a multiple
//            of ZB planes, kx in [0, nx/2], ky in [0, ny)
//   cells    parcel-centre cells c = ((k+1) ny + j) nx + i, k in [-1, nz]
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace epic {

constexpr int ZB = 16;           // z planes per FFT work-group
constexpr int P2G_TX = 8;        // p2g tile cells in x
constexpr int P2G_TY = 8;        // p2g tile cells in y
constexpr int P2G_BOX = 11;      // node box per tile side: 1 lower + 8 + 2 upper
constexpr int P2G_NF = 7;        // vol, xi, eta, zeta, tbuoy, dbuoy, hum
constexpr int P2G_TPC = 2;       // par2grid threads per cell (neighbouring lanes)

// error bits reported through the member's device error word
enum : int {
    ERR_PARCEL_COUNT = 1,   // parcel centre outside layers -1..nz
    ERR_B33 = 2,            // |B11 B22 - B12^2| <= eps
    ERR_OUT_OF_GRID = 4,    // support point outside the node grid / neighbour cells
    ERR_CELL_MISMATCH = 8,  // a sorted parcel is not in its sorted cell
    ERR_MERGE = 16,         // a small parcel found no neighbour to merge with
    ERR_DT = 32,            // the adaptive time step is not positive
};

// the time step of one member's RK step, resident in device memory: the
// host writes dt (fixed step) or t, alpha, limit, precise_stop (adaptive)
// before the step; the adaptive time-step kernels fill dt, gmax and bmax
struct StepTime {
    double dt;
    double t, alpha, limit;
    double gmax, bmax;                          // maxima of the last adaptive step
    unsigned long long gmax_bits, b2max_bits;   // running maxima (non-negative doubles as bits), kept at 0
    int precise_stop, adaptive;
};

// field slots inside Member::fld (FP32 node fields)
enum FieldSlot : int {
    F_VOLG = 0,
    F_VORTG = 1,    // 3 components
    F_TBUOYG = 4,
    F_DBUOYG = 5,
    F_HUMG = 6,
    F_VELOG = 7,    // 3
    F_VELGRADG = 10, // 8: dudx dudy dudz dvdx dvdy dvdz dwdx dwdy
    F_VTEND = 18,   // 3
    F_GRAD = 21,    // 3: the parcel-correction displacement field (grad phi)
    F_COUNT = 24
};

enum { DUDX = 0, DUDY, DUDZ, DVDX, DVDY, DVDZ, DWDX, DWDY };

struct Grid {
    int nx, ny, nz;
    int nxh;          // nx/2 + 1
    int nzp;          // nz+3 rounded up to ZB: planes 0..nz and the two halo planes
    int zeta_zero[2];
    double lower[3], extent[3], upper[3], dx[3], dxi[3], center[3], hli[3];
    double vcell, vmin;
    double glat, lambda_c, q0, f_cor[3];
};

// parcel attribute arrays (n entries each)
struct ParcelsDev {
    int64_t n;
    int64_t cap;                // allocated slots: grids over the parcels are sized by it (0: by n), so
                                // a plan stays valid when the count changes
    const int64_t *n_dev;       // the count in device memory (may be null): the kernels of the step
                                // plans read it, so their tables do not hold the count
    double *x, *y, *z;
    float *vor[3];
    float *B[5];
    float *vol, *buoy, *hum;
    float *dpos[3], *dvor[3], *dB[5];
    float *veta[3], *vtau[3];   // support-point half axes cached for l_reuse
    float *strain[8];           // only written by the stand-alone grid2par
    int *orig;                  // caller index of each physical slot
};

// the parcel count a kernel works on: the device copy when the argument
// tables leave it there
__device__ inline int64_t parcel_count(const ParcelsDev &p) { return p.n_dev ? *p.n_dev : p.n; }

// Integer quotients x / d for 0 <= x < 2^16 and 1 <= d <= 2^16 as one high
// product: with m = ceil(2^32 / d) = (2^32 + e) / d, 0 <= e < d, the
// product x m / 2^32 exceeds x / d by x e / (d 2^32) < 2^-16 <= 1/d, which
// never reaches the next integer (x / d is at most q + 1 - 1/d).  d = 1
// (m would need 33 bits) keeps m = 0 and returns x.
struct FastDiv {
    unsigned m = 0;
    int d = 1;
};
inline FastDiv make_fastdiv(int d)
{
    FastDiv f;
    f.d = d;
    f.m = (d > 1) ? (unsigned)((0x100000000ull + (unsigned long long)d - 1) / (unsigned long long)d) : 0u;
    return f;
}
__device__ inline int fdiv(int x, const FastDiv &f) { return f.d == 1 ? x : (int)__umulhi((unsigned)x, f.m); }

// The index arithmetic of the transforms stays below 2^16 (at most ZB lines
// of FFT_MAX_N points).
constexpr int FFT_MAX_N = 2048;
struct FftPlan {
    int n;          // complex length
    int nst;        // number of stages
    int radix[16];
    FastDiv nb[16];   // per stage: n / radix
    FastDiv ns[16];   // per stage: the product of the earlier radices
    int tstep[16];    // per stage: n / (ns radix), the twiddle stride
    FastDiv n1, n2;   // n and 2 n
};

// Global-address-space views of pointers kept in argument tables.  Loads
// through a generic pointer are flat loads, which count in the LDS/scalar
// counter too: every wait for a scalar load of a further table entry (the
// next attribute's pointer) would also drain the loads in flight.  Global
// loads count only in the vector-memory counter, so one parcel's loads all
// stay in flight together.  The host pass only parses device code; there the
// qualifier is dropped so that the runtime's generic-pointer overloads apply.
#ifdef __HIP_DEVICE_COMPILE__
#define EPIC_GLOBAL __attribute__((address_space(1)))
#else
#define EPIC_GLOBAL
#endif
template <class T> __device__ inline EPIC_GLOBAL T *gl(T *p) { return (EPIC_GLOBAL T *)p; }
template <class T> __device__ inline const EPIC_GLOBAL T *gl(const T *p) { return (const EPIC_GLOBAL T *)p; }

// The quantities par2grid accumulates for one parcel, in P2G_NF slots:
// vol/4 and vol/4 times vorticity, total buoyancy (with condensation),
// buoyancy and humidity.  A parcel without humidity has no condensate (the
// saturation humidity is positive), so its exp is skipped.
__device__ inline void p2g_values_of(const Grid &g, float vol, const float vor[3], float buoy, float hum, double z,
                                     float v[P2G_NF])
{
    const double qc = (hum > 0.0f) ? fmax(0.0, (double)hum - g.q0 * exp(g.lambda_c * (g.lower[2] - z))) : 0.0;
    const float v0 = (float)(0.25 * (double)vol);
    v[0] = v0;
    v[1] = v0 * vor[0];
    v[2] = v0 * vor[1];
    v[3] = v0 * vor[2];
    v[4] = v0 * (float)((double)buoy + g.glat * qc);
    v[5] = v0 * buoy;
    v[6] = v0 * hum;
}

// element at a byte offset below 4 GiB from a table pointer: the load or
// store takes the pointer as its scalar base and the offset as a 32-bit
// vector operand, with no 64-bit address arithmetic per access
template <class T> __device__ inline EPIC_GLOBAL T &at(T *p, unsigned off)
{
    return *(EPIC_GLOBAL T *)((EPIC_GLOBAL char *)p + off);
}
template <class T> __device__ inline const EPIC_GLOBAL T &at(const T *p, unsigned off)
{
    return *(const EPIC_GLOBAL T *)((const EPIC_GLOBAL char *)p + off);
}

__host__ __device__ inline int64_t field_stride(const Grid &g) { return (int64_t)(g.nz + 3) * g.ny * g.nx; }
__host__ __device__ inline int64_t plane_stride(const Grid &g) { return (int64_t)g.ny * g.nx; }
__host__ __device__ inline int64_t spec_size(const Grid &g) { return (int64_t)g.ny * g.nxh * g.nzp; }
__host__ __device__ inline int64_t ncells(const Grid &g) { return (int64_t)g.nx * g.ny * (g.nz + 2); }

__device__ inline int wrap(int i, int n)
{
    if ((unsigned)i < (unsigned)n) return i;   // nearly always: skips the integer division
    int r = i % n;
    return r < 0 ? r + n : r;
}

// Fortran int(): truncation toward zero
__device__ inline double ftrunc(double v) { return trunc(v); }

// ---------------------------------------------------------------------------
// wave-aggregated integer tickets: lanes of one wavefront holding the same key
// in consecutive lanes share one atomicAdd on counter[key]; returns the lane's
// ticket (old counter value + its rank within the run).  key < 0: no ticket.
// Must be called by all lanes of the wavefront.
// ---------------------------------------------------------------------------
template <class P> __device__ inline int wave_run_ticket(int key, P counter)
{
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1, 64);
    const bool head = (lane == 0) || (prev != key);
    const unsigned long long heads = __ballot(head);
    // run head: highest head lane <= lane; run end: next head lane
    const unsigned long long below = heads & ((lane == 63) ? ~0ull : ((2ull << lane) - 1));
    const int hl = 63 - __clzll(below);
    const unsigned long long above = heads & ~((2ull << lane) - 1 | (lane == 63 ? ~0ull : 0ull));
    const int next = above ? __ffsll((long long)above) - 1 : 64;
    int base = 0;
    if (head && key >= 0) base = atomicAdd(&counter[key], next - lane);
    base = __shfl(base, hl, 64);
    return base + (lane - hl);
}

// wave-aggregated integer increments: counter[key] += 1 for every lane, one
// atomic per run of equal keys in consecutive lanes, no return value waited
// for.  key < 0: nothing.  Must be called by all lanes of the wavefront.
template <class P> __device__ inline void wave_run_add(int key, P counter)
{
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1, 64);
    const bool head = (lane == 0) || (prev != key);
    const unsigned long long heads = __ballot(head);
    const unsigned long long above = heads & ~((2ull << lane) - 1 | (lane == 63 ? ~0ull : 0ull));
    const int next = above ? __ffsll((long long)above) - 1 : 64;
    if (head && key >= 0) (void)atomicAdd(&counter[key], next - lane);
}

}  // namespace epic
