// Batched per-site SU(3) algebra on the device (the per-site operations of SC/SUNc/SU3/*.cpp and the three
// SU(3) operations the reference declares but never defines).  Used by tests and tools: every kernel of the
// engine inlines the same functions from su3.h.
#include "internal.h"

namespace glasma {
namespace {

// complex matrices interleaved (re, im), column-major i + 3j, 18 doubles per matrix
__device__ __forceinline__ void load_m(const double* p, M3& A) {
    for (int k = 0; k < 9; k++) {
        A.re[k] = p[2 * k];
        A.im[k] = p[2 * k + 1];
    }
}
__device__ __forceinline__ void store_m(double* p, const M3& A) {
    for (int k = 0; k < 9; k++) {
        p[2 * k] = A.re[k];
        p[2 * k + 1] = A.im[k];
    }
}

__global__ void __launch_bounds__(128) k_su3_batch(int op, long n, double c, const double* __restrict__ in, const double* __restrict__ in2,
                            double* __restrict__ out) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    M3 A;
    double v[8], r[8];
    switch (op) {
        case 0:  // MatrixIExp  SC/SUNc/SU3/AlgebraOperations.cpp:48-89
            for (int a = 0; a < 8; a++) v[a] = in[8 * i + a];
            expi(c, v, A);
            store_m(out + 18 * i, A);
            return;
        case 1:  // MatrixILog (declared in the reference, never defined for SU(3))
            load_m(in + 18 * i, A);
            ilog(c, A, r);
            break;
        case 2:  // AdjointMultiplication: components of V (alpha t) V^dag
            load_m(in + 18 * i, A);
            for (int a = 0; a < 8; a++) v[a] = in2[8 * i + a];
            adj_mul(A, v, r);
            break;
        case 3:  // InverseAdjointMultiplication: components of V^dag (alpha t) V
            load_m(in + 18 * i, A);
            for (int a = 0; a < 8; a++) v[a] = in2[8 * i + a];
            inv_adj_mul(A, v, r);
            break;
        default:  // ReTrIGenU  SC/SUNc/SU3/GroupOperations.cpp:93-103
            load_m(in + 18 * i, A);
            retr_igen(A, r);
            break;
    }
    for (int a = 0; a < 8; a++) out[8 * i + a] = r[a];
}

}  // namespace
}  // namespace glasma

extern "C" int glasma_su3_batch(int op, long n, double c, const double* in, const double* in2, double* out) {
    using namespace glasma;
    if (op < 0 || op > 4 || n < 0 || !in || !out || ((op == 2 || op == 3) && !in2)) return GLASMA_ERR_ARG;
    if (op == 1 && c == 0.0) return GLASMA_ERR_ARG;
    if (n == 0) return 0;
    const int win = op == 0 ? 8 : 18, wout = op == 0 ? 18 : 8;
    double *din = nullptr, *din2 = nullptr, *dout = nullptr;
    int rc = 0;
    if (hipMalloc(&din, sizeof(double) * win * n) != hipSuccess || hipMalloc(&dout, sizeof(double) * wout * n) != hipSuccess ||
        ((op == 2 || op == 3) && hipMalloc(&din2, sizeof(double) * 8 * n) != hipSuccess))
        rc = GLASMA_ERR_NOMEM;
    if (rc == 0 && hipMemcpy(din, in, sizeof(double) * win * n, hipMemcpyHostToDevice) != hipSuccess) rc = GLASMA_ERR_HIP;
    if (rc == 0 && din2 && hipMemcpy(din2, in2, sizeof(double) * 8 * n, hipMemcpyHostToDevice) != hipSuccess)
        rc = GLASMA_ERR_HIP;
    if (rc == 0) {
        hipLaunchKernelGGL(k_su3_batch, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, nullptr, op, n, c, din, din2,
                           dout);
        if (hipGetLastError() != hipSuccess) rc = GLASMA_ERR_HIP;
    }
    if (rc == 0 && hipMemcpy(out, dout, sizeof(double) * wout * n, hipMemcpyDeviceToHost) != hipSuccess) rc = GLASMA_ERR_HIP;
    if (din) (void)hipFree(din);
    if (din2) (void)hipFree(din2);
    if (dout) (void)hipFree(dout);
    if (rc) (void)hipGetLastError();  // leave no stale error for the next launch check
    return rc;
}
