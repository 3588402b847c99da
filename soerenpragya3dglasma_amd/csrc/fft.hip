// Complex FP64 FFTs on the device, replacing the FFTW plans of SC/FFT/FFT1D.cpp, FFT2D.cpp and FFT3D.cpp.
//
// Conventions follow the reference's FFTW use: forward = exp(-2 pi i jk/n), backward = exp(+2 pi i jk/n),
// both unnormalised (callers scale by 1/n in momentum space, e.g. SC/INITIALCONDITIONS/SUNCCURRENTS/
// RevSingleNucleus.cpp:150, SC/GAUGETRANSFORMATION/ALGORITHMS/FourierAcceleration.cpp:62).
//
// Design: every transform is a set of independent 1D lines.  One workgroup loads `lpb` whole lines into
// LDS (coalesced: along the line when its elements are contiguous, across lines otherwise), runs a
// self-sorting Stockham transform in LDS (radix 4, 2, 3 kernels, any other prime by a direct DFT), and
// writes the lines back in place.  Multi-dimensional transforms are one pass per axis, so each pass
// reads and writes the array once.  Twiddles come from a table exp(-2 pi i j/n) computed once per length
// in long double and rounded to FP64.
#include <map>
#include <mutex>
#include <vector>
#include <cmath>

#include "internal.h"

namespace glasma {

namespace {

constexpr int kFftLds = 2048;     // complex entries per LDS buffer (two buffers: 64 KiB)
constexpr int kFftThreads = 256;
constexpr int kMaxStages = 24;
constexpr int kMaxGenericRadix = 16;

struct Radices {
    int n, nst;
    int r[kMaxStages];
};

int factorize(int n, Radices& R) {
    R.n = n;
    R.nst = 0;
    int m = n;
    for (int f : {4, 2, 3}) {
        while (m % f == 0) {
            if (R.nst == kMaxStages) return -1;
            R.r[R.nst++] = f;
            m /= f;
        }
    }
    for (int p = 5; m > 1; p += 2) {
        while (m % p == 0) {
            if (p > kMaxGenericRadix || R.nst == kMaxStages) return -1;
            R.r[R.nst++] = p;
            m /= p;
        }
    }
    return 0;
}

typedef double dbl2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ dbl2 cmul(dbl2 a, dbl2 b) { return dbl2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
// twiddle exp(sign * 2 pi i k / n) from the forward table
__device__ __forceinline__ dbl2 twid(const dbl2* tw, int k, int sign) {
    dbl2 w = tw[k];
    return sign < 0 ? w : dbl2{w.x, -w.y};
}

__global__ __launch_bounds__(kFftThreads) void k_fft_lines(double* __restrict__ data, Radices R, long es, long nlines,
                                                            long b1, long s1, long s2, int sign, int lpb, int ld,
                                                            const dbl2* __restrict__ tw) {
    // two LDS buffers of lpb * ld entries each, sized per launch (at most 2 * kFftLds entries), so that short
    // lines do not pay for the longest line's LDS and more workgroups fit on a CU
    extern __shared__ dbl2 lds[];
    dbl2* const buf[2] = {lds, lds + lpb * ld};
    const int n = R.n;
    const long l0 = (long)blockIdx.x * lpb;
    const int nload = lpb * n;
    dbl2* dv = reinterpret_cast<dbl2*>(data);
    // load
    for (int idx = threadIdx.x; idx < nload; idx += blockDim.x) {
        int li, j;
        if (es == 1) { li = idx / n; j = idx - li * n; }
        else { j = idx / lpb; li = idx - j * lpb; }
        long line = l0 + li;
        if (line < nlines) {
            long base = (line % b1) * s1 + (line / b1) * s2;
            buf[0][li * ld + j] = dv[base + (long)j * es];
        }
    }
    __syncthreads();
    int cur = 0, Ns = 1;
    for (int st = 0; st < R.nst; st++) {
        const int r = R.r[st];
        const int nr = n / r;
        const int tstride = n / (Ns * r);
        const dbl2* a = buf[cur];
        dbl2* b = buf[cur ^ 1];
        for (int it = threadIdx.x; it < lpb * nr; it += blockDim.x) {
            const int li = it / nr;
            const int j = it - li * nr;
            const int k = j % Ns;
            const dbl2* al = a + li * ld;
            dbl2* bl = b + li * ld;
            const int od = (j / Ns) * Ns * r + k;
            if (r == 4) {
                dbl2 v0 = al[j], v1 = al[j + nr], v2 = al[j + 2 * nr], v3 = al[j + 3 * nr];
                if (k) {
                    v1 = cmul(v1, twid(tw, k * tstride, sign));
                    v2 = cmul(v2, twid(tw, 2 * k * tstride, sign));
                    v3 = cmul(v3, twid(tw, 3 * k * tstride, sign));
                }
                dbl2 t0 = v0 + v2, t1 = v0 - v2, t2 = v1 + v3, d = v1 - v3;
                // t3 = (sign * i) * d
                dbl2 t3 = sign < 0 ? dbl2{d.y, -d.x} : dbl2{-d.y, d.x};
                bl[od] = t0 + t2;
                bl[od + Ns] = t1 + t3;
                bl[od + 2 * Ns] = t0 - t2;
                bl[od + 3 * Ns] = t1 - t3;
            } else if (r == 2) {
                dbl2 v0 = al[j], v1 = al[j + nr];
                if (k) v1 = cmul(v1, twid(tw, k * tstride, sign));
                bl[od] = v0 + v1;
                bl[od + Ns] = v0 - v1;
            } else if (r == 3) {
                dbl2 v0 = al[j], v1 = al[j + nr], v2 = al[j + 2 * nr];
                if (k) {
                    v1 = cmul(v1, twid(tw, k * tstride, sign));
                    v2 = cmul(v2, twid(tw, 2 * k * tstride, sign));
                }
                const double h = 0.86602540378443864676;  // sqrt(3)/2
                dbl2 t = v1 + v2, d = v1 - v2;
                dbl2 m = v0 - 0.5 * t;
                // (sign * i * h) * d
                dbl2 q = sign < 0 ? dbl2{h * d.y, -h * d.x} : dbl2{-h * d.y, h * d.x};
                bl[od] = v0 + t;
                bl[od + Ns] = m + q;
                bl[od + 2 * Ns] = m - q;
            } else {
                dbl2 v[kMaxGenericRadix];
                for (int q = 0; q < r; q++) {
                    v[q] = al[j + q * nr];
                    if (k && q) v[q] = cmul(v[q], twid(tw, (k * q * tstride) % n, sign));
                }
                for (int p = 0; p < r; p++) {
                    dbl2 s = v[0];
                    for (int q = 1; q < r; q++) s += cmul(v[q], twid(tw, ((p * q) % r) * nr, sign));
                    bl[od + p * Ns] = s;
                }
            }
        }
        __syncthreads();
        cur ^= 1;
        Ns *= r;
    }
    for (int idx = threadIdx.x; idx < nload; idx += blockDim.x) {
        int li, j;
        if (es == 1) { li = idx / n; j = idx - li * n; }
        else { j = idx / lpb; li = idx - j * lpb; }
        long line = l0 + li;
        if (line < nlines) {
            long base = (line % b1) * s1 + (line / b1) * s2;
            dv[base + (long)j * es] = buf[cur][li * ld + j];
        }
    }
}

// forward twiddle tables exp(-2 pi i j / n), j < n, one per (device, n); kept for the process lifetime
std::mutex g_tw_mu;
std::map<std::pair<int, int>, dbl2*> g_tw;

hipError_t twiddles(int n, const dbl2** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto key = std::make_pair(dev, n);
    auto it = g_tw.find(key);
    if (it != g_tw.end()) {
        *out = it->second;
        return hipSuccess;
    }
    std::vector<double> h(2 * (size_t)n);
    const long double tp = 6.283185307179586476925286766559005768L;
    for (int j = 0; j < n; j++) {
        long double ang = tp * (long double)j / (long double)n;
        h[2 * j] = (double)cosl(ang);
        h[2 * j + 1] = (double)(-sinl(ang));
    }
    dbl2* d = nullptr;
    e = hipMalloc(&d, sizeof(dbl2) * (size_t)n);
    if (e != hipSuccess) return e;
    e = hipMemcpy(d, h.data(), sizeof(dbl2) * (size_t)n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    g_tw[key] = d;
    *out = d;
    return hipSuccess;
}

}  // namespace

int fft_lines(hipStream_t st, double* data, int n, long es, long nlines, long b1, long s1, long s2, int sign) {
    if (n <= 1 || nlines <= 0) return 0;
    if (n > kFftLds) return GLASMA_ERR_ARG;
    Radices R;
    if (factorize(n, R) != 0) return GLASMA_ERR_ARG;
    const dbl2* tw = nullptr;
    const hipError_t te = twiddles(n, &tw);
    if (te != hipSuccess) return hip_code(te);
    int lpb = 1;
    while (2 * lpb * (n + 1) <= kFftLds && 2 * lpb <= 64) lpb *= 2;
    const int ld = lpb > 1 ? n + 1 : n;
    const long blocks = (nlines + lpb - 1) / lpb;
    if (blocks > 0x7fffffffL) return GLASMA_ERR_ARG;
    const size_t lds = 2 * sizeof(dbl2) * (size_t)lpb * ld;
    hipLaunchKernelGGL(k_fft_lines, dim3((unsigned)blocks), dim3(kFftThreads), lds, st, data, R, es, nlines, b1, s1,
                       s2, sign, lpb, ld, tw);
    return hipGetLastError() == hipSuccess ? 0 : GLASMA_ERR_HIP;
}

// c[(comp*nzl + z)*P + y*nx + x]
int fft_xy(hipStream_t st, double* c, const Geom& g, int ncomp, int nzl, int sign) {
    const long P = (long)g.nx * g.ny;
    int rc = fft_lines(st, c, g.nx, 1, (long)g.ny * nzl * ncomp, (long)g.ny * nzl * ncomp, g.nx, 0, sign);
    if (rc) return rc;
    return fft_lines(st, c, g.ny, g.nx, (long)g.nx * nzl * ncomp, g.nx, 1, P, sign);
}

int fft_z(hipStream_t st, double* c, const Geom& g, int ncomp, int sign) {
    const long P = (long)g.nx * g.ny;
    return fft_lines(st, c, g.nz, P, P * ncomp, P, 1, P * g.nz, sign);
}

int fft_3d(hipStream_t st, double* c, const Geom& g, int ncomp, int sign) {
    int rc = fft_xy(st, c, g, ncomp, g.nz, sign);
    if (rc) return rc;
    return fft_z(st, c, g, ncomp, sign);
}

}  // namespace glasma

extern "C" int glasma_fft_host(double* data, int nd, const int* dims, long batch, int sign) {
    using namespace glasma;
    if (!data || !dims || nd < 1 || nd > 8 || batch < 1 || (sign != 1 && sign != -1)) return GLASMA_ERR_ARG;
    long total = batch;
    for (int d = 0; d < nd; d++) {
        if (dims[d] < 1) return GLASMA_ERR_ARG;
        total *= dims[d];
    }
    double* dev = nullptr;
    const size_t bytes = sizeof(double) * 2 * (size_t)total;
    const hipError_t ae = hipMalloc(&dev, bytes);
    if (ae != hipSuccess) return hip_code(ae);
    int rc = 0;
    if (hipMemcpy(dev, data, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = GLASMA_ERR_HIP;
    long inner = batch;
    for (int d = 0; d < nd && rc == 0; d++) {
        const long n = dims[d];
        const long outer = total / (inner * n);
        rc = fft_lines(nullptr, dev, (int)n, inner, inner * outer, inner, 1, inner * n, sign);
        inner *= n;
    }
    if (rc == 0 && hipMemcpy(data, dev, bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = GLASMA_ERR_HIP;
    hipFree(dev);
    if (rc) (void)hipGetLastError();  // leave no stale error for the next launch check
    return rc;
}
