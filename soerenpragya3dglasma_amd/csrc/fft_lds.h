// In-LDS FFT core shared by the transform passes (fft.hip) and the kernels that fuse a transform of their own x lines
// (gauge.hip): line factorisation, complex helpers and the self-sorting Stockham stages over lines held in LDS.
#pragma once

#include <hip/hip_runtime.h>

namespace glasma {

constexpr int kMaxStages = 24;
constexpr int kMaxGenericRadix = 16;

struct Radices {
    int n, nst;
    int r[kMaxStages];
};

inline int factorize(int n, Radices& R) {
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
// the same from a two-level table of kTwSplit + ceil(n / kTwSplit) entries: lo[j] = exp(-2 pi i j / n), j < kTwSplit,
// then hi[j] = exp(-2 pi i kTwSplit j / n); exp(-2 pi i k / n) = lo[k % kTwSplit] hi[k / kTwSplit], within two units
// of the last place.  1.5 KiB for a line of 2048, where the full table takes 32 KiB of LDS.
constexpr int kTwSplit = 32;
__device__ __forceinline__ dbl2 twid_split(const dbl2* tw, int k, int sign) {
    dbl2 w = cmul(tw[k & (kTwSplit - 1)], tw[kTwSplit + k / kTwSplit]);
    return sign < 0 ? w : dbl2{w.x, -w.y};
}
template <bool SPLIT>
__device__ __forceinline__ dbl2 twiddle(const dbl2* tw, int k, int sign) {
    return SPLIT ? twid_split(tw, k, sign) : twid(tw, k, sign);
}

// a / d for 0 <= a < 2^20 and 1 <= d < 2^20, with rd = 1/d in FP32: the FP32 product is within 1/8 of a / d,
// so the truncated estimate is off by at most one, and one remainder test corrects it.  The index arithmetic
// below divides by run-time lengths once per element; this replaces the integer division sequence.
__device__ __forceinline__ int qdiv(int a, int d, float rd) {
    int q = (int)((float)a * rd);
    const int r = a - q * d;
    if (r < 0) q--;
    else if (r >= d) q++;
    return q;
}

// the forward twiddle table exp(-2 pi i j / n), j < n, of the current device (fft.hip; kept for the process lifetime)
hipError_t fft_twiddles(int n, const dbl2** out);
// entries of the two-level table of twid_split
__host__ __device__ inline int split_entries(int n) { return kTwSplit + (n + kTwSplit - 1) / kTwSplit; }

// The Stockham stages over the lpb lines in buffer `cur` (line li at lb + cur * bs + li * ld); buffer c starts at
// lb + c * bs.  Returns which buffer holds the transform.  The buffers are addressed as offsets from the LDS base, not through
// a run-time indexed pointer array, so the compiler keeps them in the LDS address space (ds_read / ds_write);
// a generic pointer would give flat instructions, whose waits also wait for every global load in flight.
// transform.  Every stage ends with a barrier.
template <bool SPLIT = false>
__device__ __forceinline__ int stockham(dbl2* const lb, int bs, int cur, const Radices& R, int lpb, int ld,
                                        int sign, const dbl2* tw) {
    const int n = R.n;
    int Ns = 1;
    for (int st = 0; st < R.nst; st++) {
        const int r = R.r[st];
        const int nr = n / r;
        const int tstride = n / (Ns * r);
        const dbl2* a = lb + cur * bs;
        dbl2* b = lb + (cur ^ 1) * bs;
        const float rnr = 1.0f / (float)nr, rns = 1.0f / (float)Ns;
        for (int it = threadIdx.x; it < lpb * nr; it += blockDim.x) {
            const int li = qdiv(it, nr, rnr);
            const int j = it - li * nr;
            const int jq = qdiv(j, Ns, rns);
            const int k = j - jq * Ns;
            const dbl2* al = a + li * ld;
            dbl2* bl = b + li * ld;
            const int od = jq * Ns * r + k;
            if (r == 4) {
                dbl2 v0 = al[j], v1 = al[j + nr], v2 = al[j + 2 * nr], v3 = al[j + 3 * nr];
                if (k) {
                    v1 = cmul(v1, twiddle<SPLIT>(tw, k * tstride, sign));
                    v2 = cmul(v2, twiddle<SPLIT>(tw, 2 * k * tstride, sign));
                    v3 = cmul(v3, twiddle<SPLIT>(tw, 3 * k * tstride, sign));
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
                if (k) v1 = cmul(v1, twiddle<SPLIT>(tw, k * tstride, sign));
                bl[od] = v0 + v1;
                bl[od + Ns] = v0 - v1;
            } else if (r == 3) {
                dbl2 v0 = al[j], v1 = al[j + nr], v2 = al[j + 2 * nr];
                if (k) {
                    v1 = cmul(v1, twiddle<SPLIT>(tw, k * tstride, sign));
                    v2 = cmul(v2, twiddle<SPLIT>(tw, 2 * k * tstride, sign));
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
                // a direct DFT of prime length r <= 13; the inputs are re-read from LDS for each output rather than
                // held in a run-time indexed array (which would live in scratch memory)
                for (int p = 0; p < r; p++) {
                    dbl2 s = al[j];
                    for (int q = 1; q < r; q++) {
                        dbl2 v = al[j + q * nr];
                        if (k) v = cmul(v, twiddle<SPLIT>(tw, (k * q * tstride) % n, sign));
                        s += cmul(v, twiddle<SPLIT>(tw, ((p * q) % r) * nr, sign));
                    }
                    bl[od + p * Ns] = s;
                }
            }
        }
        __syncthreads();
        cur ^= 1;
        Ns *= r;
    }
    return cur;
}

}  // namespace glasma
