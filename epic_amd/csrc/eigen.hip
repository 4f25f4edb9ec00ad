// Stand-alone FP64 eigen-decomposition of symmetric 3x3 matrices on the
// device: the same Scherzinger-Dohrmann solver (ellipsoid.h) that every
// par2grid, push, split and merge uses for the parcels' shape matrices,
// exposed over host arrays so that it can be checked on its own.
#include <hip/hip_runtime.h>

#include <string>

#include "epic_gpu.h"
#include "ellipsoid.h"

using namespace epic;

namespace {

// one thread per matrix; A(3,3,n), D(3,n), V(3,3,n) column-major FP64
__global__ __launch_bounds__(256) void k_diagonalise(int64_t n, const double *__restrict__ A, double *__restrict__ D,
                                                     double *__restrict__ V)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a[9], d[3], v[3][3];
    // column-major A(r, c) = A[9 i + r + 3 c]; the solver takes row-major
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) a[3 * r + c] = A[9 * i + r + 3 * c];
    scherzinger(a, d, v);
#pragma unroll
    for (int k = 0; k < 3; ++k) D[3 * i + k] = d[k];
    // V(:, c) = eigenvector c
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) V[9 * i + r + 3 * c] = v[c][r];
}

thread_local std::string g_eig_msg;

int eig_fail(int code, const std::string &m)
{
    g_eig_msg = m;
    return code;
}

}  // namespace

extern "C" int epic_diagonalise(int64_t n, const double *A, double *D, double *V)
{
    if (n < 0 || (n > 0 && (!A || !D || !V))) return eig_fail(EPIC_E_INVALID, "epic_diagonalise: bad arguments");
    if (n == 0) return EPIC_OK;
    if (n > ((int64_t)1 << 26)) return eig_fail(EPIC_E_INVALID, "epic_diagonalise: at most 2^26 matrices per call");
    double *dA = nullptr, *dD = nullptr, *dV = nullptr;
    auto cleanup = [&]() {
        if (dA) (void)hipFree(dA);
        if (dD) (void)hipFree(dD);
        if (dV) (void)hipFree(dV);
    };
    auto hip = [&](hipError_t e, const char *what) {
        if (e == hipSuccess) return false;
        g_eig_msg = std::string("epic_diagonalise: ") + what + ": " + hipGetErrorString(e);
        return true;
    };
    const size_t b9 = sizeof(double) * 9 * (size_t)n, b3 = sizeof(double) * 3 * (size_t)n;
    if (hip(hipMalloc((void **)&dA, b9), "hipMalloc") || hip(hipMalloc((void **)&dD, b3), "hipMalloc") ||
        hip(hipMalloc((void **)&dV, b9), "hipMalloc")) {
        cleanup();
        return EPIC_E_NOMEM;
    }
    int rc = EPIC_OK;
    if (hip(hipMemcpy(dA, A, b9, hipMemcpyHostToDevice), "upload")) rc = EPIC_E_HIP;
    if (rc == EPIC_OK) {
        const unsigned blocks = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(k_diagonalise, dim3(blocks), dim3(256), 0, 0, n, (const double *)dA, dD, dV);
        if (hip(hipGetLastError(), "launch") || hip(hipDeviceSynchronize(), "k_diagonalise")) rc = EPIC_E_HIP;
    }
    if (rc == EPIC_OK && (hip(hipMemcpy(D, dD, b3, hipMemcpyDeviceToHost), "download") ||
                          hip(hipMemcpy(V, dV, b9, hipMemcpyDeviceToHost), "download")))
        rc = EPIC_E_HIP;
    cleanup();
    return rc;
}

extern "C" const char *epic_diagonalise_last_error(void) { return g_eig_msg.c_str(); }
