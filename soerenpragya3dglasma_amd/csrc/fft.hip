// Complex FP64 FFTs on the device, replacing the FFTW plans of SC/FFT/FFT1D.cpp, FFT2D.cpp and FFT3D.cpp.
//
// Conventions follow the reference's FFTW use: forward = exp(-2 pi i jk/n), backward = exp(+2 pi i jk/n),
// both unnormalised (callers scale by 1/n in momentum space, e.g. SC/INITIALCONDITIONS/SUNCCURRENTS/
// RevSingleNucleus.cpp:150, SC/GAUGETRANSFORMATION/ALGORITHMS/FourierAcceleration.cpp:62).
//
// Design: every transform is a set of independent 1D lines, transformed in groups.  A workgroup loads a group
// into LDS (contiguous lines as one packed run; strided lines across the group, so neighbouring threads read
// neighbouring addresses), runs a self-sorting Stockham transform in LDS (radix 4, 2, 3 kernels, any other
// prime <= 13 by a direct DFT) and writes the lines back in place.  Multi-dimensional transforms are one pass per
// axis, so each pass reads and writes the array once.
// - Strided lines (the y and z passes) go 4 to a group within the device's LDS (160 KiB on gfx950), so each row
//   of a load is 64 contiguous bytes; their groups are dealt to the XCDs so that neighbouring groups run on one
//   XCD and share the rest of each 128-byte line in its L2.
// - Where only one or two workgroups fit per CU, each block walks several groups and prefetches the next one
//   into registers while it transforms the current one.
// - The twiddle table exp(-2 pi i j/n) is computed once per length in long double, rounded to FP64, and copied
//   into LDS with the data where it fits.
// - fft_3d_poisson fuses the forward z transform, the Poisson filter of gauge fixing and the inverse z transform
//   into one pass.
#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>
#include <cmath>

#include "fft_lds.h"
#include "internal.h"

namespace glasma {

namespace {

constexpr int kFftLds = 2048;  // longest line
// complex entries per LDS buffer for contiguous lines (several short lines per workgroup share it; one longer
// line takes what it needs).  Smaller buffers leave room for more workgroups per CU: the x pass of length 256
// at 256^2 x 1024 with 8 components takes 5.3-5.8 ms with 1024 entries (2 lines), 5.8-6.2 ms with 2048 (4) and
// 8.1-8.4 ms with 4096 (8).
#ifndef GLASMA_FFT_CONTIG_ENTRIES
#define GLASMA_FFT_CONTIG_ENTRIES 1024
#endif
constexpr int kContigEntries = GLASMA_FFT_CONTIG_ENTRIES;
constexpr int kFftThreads = 256;
constexpr int kMaxLinesPerBlock = 64;
// strided lines per block at most.  8 would make each row of a load one 128-byte cache line, but fewer leave
// room for more workgroups per CU: at 256^2 x 1024 with 8 components the y pass (length 256) takes 7.6 ms with
// 8 lines (66 KiB of LDS, two workgroups per CU), 5.4 ms with 4 and 5.7 ms with 2; the z pass (length 1024)
// 14.3 ms with 4 and 16.6 ms with 2.
#ifndef GLASMA_FFT_STRIDED_LINES
#define GLASMA_FFT_STRIDED_LINES 4
#endif
constexpr int kStridedLines = GLASMA_FFT_STRIDED_LINES;
// the filtered z pass of fft_3d_poisson: lines per group.  With its two transforms per group the pass gains more
// from a second workgroup per CU than from 64-byte rows: at 256^2 x 1024 (four arrays) 2 lines with twiddles from
// memory (74 KiB of LDS) took 6.9-7.1 ms, 4 lines with the full twiddle table in LDS (155 KiB) 8.2 ms, 4 lines with
// twiddles from memory 9.3 ms.  Its twiddles now come from the two-level table (twid_split) in LDS, which fits
// beside two lines at two workgroups per CU.
#ifndef GLASMA_FFT_FILTER_LINES
#define GLASMA_FFT_FILTER_LINES 2
#endif
constexpr int kFilterLines = GLASMA_FFT_FILTER_LINES;
// the radix-16 filtered z pass (k_zpoisson) where it applies; 0 keeps the generic pass everywhere
#ifndef GLASMA_FFT_ZPOISSON
#define GLASMA_FFT_ZPOISSON 1
#endif
constexpr bool kZPoisson = GLASMA_FFT_ZPOISSON != 0;
static_assert(kStridedLines >= 1 && kStridedLines <= 64 && (kStridedLines & (kStridedLines - 1)) == 0,
              "strided lines per block: a power of two dividing the block size");

// Optional momentum-space filter of a strided z pass (the Poisson step of Fourier-accelerated gauge fixing): the
// lines are transformed forward, element k_z of the line at transverse site (x, y) is multiplied as
// v -> nf * v / p^2 when p^2 = sum_mu (2 - 2 cos(2 pi k_mu / N_mu)) / a_mu^2 > 0 (the zero mode passes
// unscaled), and the lines are transformed back, all in LDS.  The x and y passes of the forward transform run
// before, those of the inverse after (the passes commute).
// Neumann (open z, the links that exist only): the z transform is the DCT-II of the line, computed as a complex FFT of
// length n of the line in even-odd order (Makhoul, IEEE Trans. ASSP 28 (1980) 27-34; element j sits at position
// nperm(j)), with p_z^2 = (2 - 2 cos(pi k_z / n)) / a_z^2 and the inverse a DCT-III: the result is the periodic
// Poisson step of the line doubled by its mirror image, restricted to the line.  The quarter-wave table wq holds
// exp(-i pi k / (2 n)), k < n.
struct LineFilter {
    int on, nx, ny;
    long P;  // transverse sites: the line's site is line % P
    double ax, ay, az, nf;
    int drop_zero;  // the zero mode set to 0 (else passed through unscaled)
    int neumann;
    const dbl2* wq;
};

// position of element j of a line of length n in even-odd order: the even elements ascending, then the odd ones
// descending
__device__ __forceinline__ int nperm(int j, int n) { return (j & 1) ? n - ((j + 1) >> 1) : (j >> 1); }

// One launch transforms `nlines` lines of length n in groups of lpb lines.
// - Packed (es = 1): line l starts at l * n, so a group is one run of lpb * n elements.
// - Strided: line l starts at (l % b1) * s1 + (l / b1) * s2 and steps by es; a group's lines are read across
//   (lpb divides the block size, so thread t always meets line t & (lpb - 1), whose start it finds once per
//   group: the 64-bit divisions stay out of the per-element loops).
// Each block walks several groups and fetches the next group into registers (MAXE elements per thread) while
// it transforms and stores the current one, so a CU keeps loads in flight through the LDS stages.  No block
// waits for another.  TwLds: the twiddle table is copied into LDS once per block.
template <bool TwLds, int MAXE, bool Filt>
__global__ __launch_bounds__(kFftThreads) void k_fft_lines(double* __restrict__ data, Radices R, long es, long nlines,
                                                            long b1, long s1, long s2, int sign, int lpb, int ld,
                                                            long groups, long xcd_per, LineFilter filt,
                                                            const dbl2* __restrict__ twg) {
    // LDS, sized per launch: two buffers of lpb * ld entries, then the twiddles (TwLds: the full table; Filt: the
    // two-level one), then (Filt) the z term of p^2 for each k_z
    extern __shared__ dbl2 lds[];
    const int bs = lpb * ld;  // buffer c starts at lds + c * bs
    dbl2* const twl = lds + 2 * lpb * ld;
    const int n = R.n;
    // the groups of this block: g, g + gstep, ... below gend.  xcd_per > 0: block b runs on XCD b % 8, which owns
    // groups [k * xcd_per, (k + 1) * xcd_per); its blocks walk them side by side, so neighbouring groups (which
    // share cache lines when the lines are strided) are in flight on one XCD together
    long g, gstep, gend;
    if (xcd_per > 0) {
        const long k = blockIdx.x % 8;
        g = k * xcd_per + blockIdx.x / 8;
        gstep = gridDim.x / 8;
        gend = min((k + 1) * xcd_per, groups);
    } else {
        g = blockIdx.x;
        gstep = gridDim.x;
        gend = groups;
    }
    if (g >= gend) return;
    // Filt: twg is the two-level table (split_entries(n) entries), always copied
    const int ntw = Filt ? split_entries(n) : (TwLds ? n : 0);
    for (int k = threadIdx.x; k < ntw; k += blockDim.x) twl[k] = twg[k];
    const dbl2* const tw = (TwLds || Filt) ? twl : twg;
    double* const ztab = reinterpret_cast<double*>(twl + ntw);
    if (Filt)
        for (int k = threadIdx.x; k < n; k += blockDim.x)
            ztab[k] = (2.0 - 2.0 * cos((filt.neumann ? 1.0 : 2.0) * kPi * k / (double)n)) / (filt.az * filt.az);
    // the first entry of buffer 1 takes the writes of elements past the group: nothing reads buffer 1 between
    // those writes and the first stage, which overwrites it, and the previous group's reads of it end at the
    // barrier that closes each iteration
    dbl2* const spare = lds + bs;
    dbl2* const dv = reinterpret_cast<dbl2*>(data);
    const bool packed = es == 1;
    const int nload = lpb * n;
    const int lg_lpb = __builtin_ctz(lpb);
    const int myli = threadIdx.x & (lpb - 1);
    const float rn = 1.0f / (float)n;
    // packed: start of group gg; strided: start of this thread's line in group gg, -1 past the last line
    auto start = [&](long gg) -> long {
        if (packed) return gg * lpb * (long)n;
        const long line = gg * lpb + myli;
        return line < nlines ? (line % b1) * s1 + (line / b1) * s2 : -1;
    };
    // packed: number of elements of group gg that exist (the last group may be short)
    auto valid = [&](long gg) -> long { return packed ? (nlines - gg * lpb) * (long)n : 0; };
    // pre[m] = element threadIdx.x + m * kFftThreads of the group, if bit m of ok is set.  The loads are
    // unconditional (an element that does not exist reads the group's first element, or element 0) and their
    // values are not touched until the LDS writes, so the compiler can count them: those writes wait for these
    // loads only, and the stages and stores run while they are in flight.
    dbl2 pre[MAXE];
    unsigned ok = 0;
    auto fetch = [&](long gg) {
        const long base = start(gg), lim = valid(gg);
        ok = 0;
#pragma unroll
        for (int m = 0; m < MAXE; m++) {
            const int idx = threadIdx.x + m * kFftThreads;
            bool v;
            long at;
            if (packed) {
                v = idx < nload && idx < lim;
                at = base + (v ? idx : 0);
            } else {
                v = idx < nload && base >= 0;
                at = v ? base + (long)(idx >> lg_lpb) * es : 0;
            }
            pre[m] = dv[at];
            ok |= (unsigned)v << m;
        }
    };
    fetch(g);
    while (true) {
        // branch-free, like the loads: an element past the group goes to the spare slot
#pragma unroll
        for (int m = 0; m < MAXE; m++) {
            const int idx = threadIdx.x + m * kFftThreads;
            const int lq = qdiv(idx, n, rn);
            const int li = packed ? lq : myli;
            const int j0 = packed ? idx - lq * n : idx >> lg_lpb;
            const int j = Filt && filt.neumann ? nperm(j0, n) : j0;
            dbl2* const at = idx < nload ? lds + li * ld + j : spare;
            *at = (ok >> m) & 1u ? pre[m] : dbl2{0.0, 0.0};
        }
        __syncthreads();
        const long gn = g + gstep;
        const bool more = gn < gend;
        if (more) fetch(gn);  // in flight during the stages and the stores below
        int cur = stockham<Filt>(lds, bs, 0, R, lpb, ld, sign, tw);
        if (Filt) {  // strided lines only: this thread's line is myli, its positions idx >> lg_lpb
            dbl2* const fb = lds + cur * bs;
            const long t = (g * lpb + myli) % filt.P;
            const int px = (int)(t % filt.nx), py = (int)(t / filt.nx);
            const double pxy = (2.0 - 2.0 * cos(2.0 * kPi * px / (double)filt.nx)) / (filt.ax * filt.ax) +
                               (2.0 - 2.0 * cos(2.0 * kPi * py / (double)filt.ny)) / (filt.ay * filt.ay);
            if (filt.neumann) {
                // V (the FFT of the even-odd line) -> X_k = (w_k V_k + conj(w_k) V_{n-k}) / 2, the DCT-II; Y = h X with
                // h = nf / p^2 (0 at p^2 = 0); Q_k = conj(w_k) (Y_k - i Y_{n-k}), Y_n = 0, whose inverse FFT read
                // in even-odd order is the DCT-III of Y.  One thread takes the pair (k, n - k).
                dbl2* const row = fb + myli * ld;
                for (int k = threadIdx.x >> lg_lpb; 2 * k <= n; k += kFftThreads >> lg_lpb) {
                    const int kr = k == 0 ? 0 : n - k;
                    const dbl2 vk = row[k], vr = row[kr];
                    const dbl2 wk = filt.wq[k], wr = filt.wq[kr];
                    const dbl2 cwk = {wk.x, -wk.y}, cwr = {wr.x, -wr.y};
                    const double pk = pxy + ztab[k], pr = pxy + ztab[kr];
                    const double hk = pk > 0.0 ? filt.nf / pk : 0.0, hr = pr > 0.0 ? filt.nf / pr : 0.0;
                    const dbl2 yk = (0.5 * hk) * (cmul(wk, vk) + cmul(cwk, vr));
                    const dbl2 yr = (0.5 * hr) * (cmul(wr, vr) + cmul(cwr, vk));
                    if (k == 0) {
                        row[0] = yk;
                    } else {
                        row[k] = cmul(cwk, dbl2{yk.x + yr.y, yk.y - yr.x});
                        row[kr] = cmul(cwr, dbl2{yr.x + yk.y, yr.y - yk.x});
                    }
                }
            } else {
#pragma unroll
                for (int m = 0; m < MAXE; m++) {
                    const int idx = threadIdx.x + m * kFftThreads;
                    if (idx < nload) {
                        const int pz = idx >> lg_lpb;
                        const double p2 = pxy + ztab[pz];
                        if (p2 > 0.0) {
                            dbl2 v = fb[myli * ld + pz];
                            v.x = filt.nf * v.x / p2;
                            v.y = filt.nf * v.y / p2;
                            fb[myli * ld + pz] = v;
                        } else if (filt.drop_zero) {
                            fb[myli * ld + pz] = dbl2{0.0, 0.0};
                        }
                    }
                }
            }
            __syncthreads();
            cur = stockham<Filt>(lds, bs, cur, R, lpb, ld, -sign, tw);
        }
        const dbl2* const res = lds + cur * bs;
        const long base = start(g), lim = valid(g);
#pragma unroll
        for (int m = 0; m < MAXE; m++) {
            const int idx = threadIdx.x + m * kFftThreads;
            if (idx < nload) {
                if (packed) {
                    const int li = qdiv(idx, n, rn);
                    if (idx < lim) dv[base + idx] = res[li * ld + idx - li * n];
                } else if (base >= 0) {
                    const int j = idx >> lg_lpb;
                    dv[base + (long)j * es] = res[myli * ld + (Filt && filt.neumann ? nperm(j, n) : j)];
                }
            }
        }
        if (!more) break;
        __syncthreads();  // every read of this group's result precedes the next group's writes to buffer 0
        g = gn;
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

// two-level twiddle tables (twid_split), one per (device, n); kept for the process lifetime
std::map<std::pair<int, int>, dbl2*> g_tws;

hipError_t twiddles_split(int n, const dbl2** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto key = std::make_pair(dev, n);
    auto it = g_tws.find(key);
    if (it != g_tws.end()) {
        *out = it->second;
        return hipSuccess;
    }
    const int ns = split_entries(n);
    std::vector<double> h(2 * (size_t)ns);
    const long double tp = 6.283185307179586476925286766559005768L;
    for (int j = 0; j < ns; j++) {
        const long double ang = tp * (long double)(j < kTwSplit ? j : kTwSplit * (j - kTwSplit)) / (long double)n;
        h[2 * j] = (double)cosl(ang);
        h[2 * j + 1] = (double)(-sinl(ang));
    }
    dbl2* d = nullptr;
    e = hipMalloc(&d, sizeof(dbl2) * (size_t)ns);
    if (e != hipSuccess) return e;
    e = hipMemcpy(d, h.data(), sizeof(dbl2) * (size_t)ns, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return e;
    }
    g_tws[key] = d;
    *out = d;
    return hipSuccess;
}

// quarter-wave tables exp(-i pi j / (2 n)), j < n, of the Neumann filter, one per (device, n); kept for the process
// lifetime
std::map<std::pair<int, int>, dbl2*> g_wq;

hipError_t quarter_wave(int n, const dbl2** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto key = std::make_pair(dev, n);
    auto it = g_wq.find(key);
    if (it != g_wq.end()) {
        *out = it->second;
        return hipSuccess;
    }
    std::vector<double> h(2 * (size_t)n);
    const long double hp = 1.570796326794896619231321691639751442L;  // pi / 2
    for (int j = 0; j < n; j++) {
        long double ang = hp * (long double)j / (long double)n;
        h[2 * j] = (double)cosl(ang);
        h[2 * j + 1] = (double)(-sinl(ang));
    }
    dbl2* d = nullptr;
    e = hipMalloc(&d, sizeof(dbl2) * (size_t)n);
    if (e != hipSuccess) return e;
    e = hipMemcpy(d, h.data(), sizeof(dbl2) * (size_t)n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return e;
    }
    g_wq[key] = d;
    *out = d;
    return hipSuccess;
}

// per device, found once: the dynamic LDS k_fft_lines may use (the device's per-workgroup limit, 160 KiB on
// gfx950, when the runtime grants the kernel that much, else the 64 KiB every CDNA part allows) and the CU count
struct FftDevice {
    size_t lds;
    int cus;
};
std::map<int, FftDevice> g_fft_dev;

template <bool TW, int MAXE>
hipError_t allow_lds(int v) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_lines<TW, MAXE, false>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, v);
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_lines<TW, MAXE, true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, v);
    return e;
}

hipError_t fft_device(FftDevice* out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto it = g_fft_dev.find(dev);
    if (it != g_fft_dev.end()) {
        *out = it->second;
        return hipSuccess;
    }
    int v = 0, cus = 0;
    e = hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    FftDevice d = {65536, cus > 0 ? cus : 1};
    if (v > 65536) {
        if (allow_lds<true, 2>(v) == hipSuccess && allow_lds<true, 4>(v) == hipSuccess &&
            allow_lds<true, 8>(v) == hipSuccess && allow_lds<true, 16>(v) == hipSuccess &&
            allow_lds<false, 2>(v) == hipSuccess && allow_lds<false, 4>(v) == hipSuccess &&
            allow_lds<false, 8>(v) == hipSuccess && allow_lds<false, 16>(v) == hipSuccess)
            d.lds = (size_t)v;
        else
            (void)hipGetLastError();  // not granted: stay at 64 KiB, and leave no stale error
    }
    g_fft_dev[dev] = d;
    *out = d;
    return hipSuccess;
}

// One launch of k_fft_lines<TW, MAXE>.  Where the LDS admits at most GLASMA_FFT_WALK_PER_CU workgroups per CU
// (long strided lines), the grid is what fits on the device at once and each block walks several groups with the
// register prefetch: the z pass of length 1024 at 256^2 x 1024 x 8 colours fell from 14.3 to 11.2 ms.  Elsewhere
// each block takes one group, since many resident workgroups already overlap loads with LDS work, and a walking
// block still waits for its previous group's stores (they are conditional) before it writes the next group to
// LDS: walking the y pass of length 256 (four workgroups per CU) took 6.9 ms against 4.2 ms.  With XCD groups
// the grid is a multiple of 8.
#ifndef GLASMA_FFT_WALK_PER_CU
#define GLASMA_FFT_WALK_PER_CU 2
#endif
template <bool TW, int MAXE, bool F>
hipError_t launch_lines(hipStream_t st, size_t lds, int cus, double* data, const Radices& R, long es, long nlines,
                        long b1, long s1, long s2, int sign, int lpb, int ld, long groups, long xcd_per,
                        const LineFilter& filt, const dbl2* tw) {
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(&k_fft_lines<TW, MAXE, F>), kFftThreads, lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    const long resident = per_cu <= GLASMA_FFT_WALK_PER_CU ? (long)cus * per_cu : groups;
    const long blocks = xcd_per > 0 ? 8 * std::min(xcd_per, (resident + 7) / 8) : std::min(groups, resident);
    if (blocks > 0x7fffffffL) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((k_fft_lines<TW, MAXE, F>), dim3((unsigned)blocks), dim3(kFftThreads), lds, st, data, R, es,
                       nlines, b1, s1, s2, sign, lpb, ld, groups, xcd_per, filt, tw);
    return hipGetLastError();
}

thread_local std::string g_fft_err;

int fft_fail(int code, const std::string& what) {
    g_fft_err = what;
    return code;
}
int fft_fail_hip(hipError_t e, const std::string& what) { return fft_fail(hip_code(e), what + ": " + hipGetErrorString(e)); }

}  // namespace

const char* fft_last_error() { return g_fft_err.c_str(); }

hipError_t fft_twiddles(int n, const dbl2** out) { return twiddles(n, out); }

namespace {

// ---- the filtered z pass for Nz = 256 R3, R3 in {1, 2, 4} -------------------------------------------------------
// One group is 4096 entries: L = 16 / R3 neighbouring z lines (64-byte rows at Nz = 1024), in place in one LDS
// buffer, so that two workgroups fit per CU.  The transforms are Stockham stages of radix 16, 16 and R3 whose
// butterflies run in registers (one radix-16 butterfly per thread), three stages per transform where the generic
// pass takes five radix-4 ones: the first forward stage works on the registers the loads filled, the last inverse
// stage stores from registers, the stages between read all their inputs, wait, and write all their outputs.  The
// plain filter is applied as the last forward stage writes; the Neumann filter (even-odd lines) runs its pair step
// between the transforms.

// DFT of length R in registers, natural order in and out: V_k = sum_m v_m w^(m k), w = exp(sign 2 pi i / R)
__device__ __forceinline__ dbl2 mul_si(dbl2 a, int sign) { return sign < 0 ? dbl2{a.y, -a.x} : dbl2{-a.y, a.x}; }
__device__ __forceinline__ void dft2(dbl2& a, dbl2& b) {
    const dbl2 t = a - b;
    a = a + b;
    b = t;
}
__device__ __forceinline__ void dft4(dbl2& v0, dbl2& v1, dbl2& v2, dbl2& v3, int sign) {
    const dbl2 t0 = v0 + v2, t1 = v0 - v2, t2 = v1 + v3, t3 = mul_si(v1 - v3, sign);
    v0 = t0 + t2;
    v1 = t1 + t3;
    v2 = t0 - t2;
    v3 = t1 - t3;
}
// exp(sign 2 pi i e / 16), e < 16: cos and sin of e pi / 8
__device__ __forceinline__ dbl2 w16(int e, int sign) {
    constexpr double c1 = 0.92387953251128675613, c2 = 0.70710678118654752440, c3 = 0.38268343236508977173;
    const double cs[16] = {1.0, c1, c2, c3, 0.0, -c3, -c2, -c1, -1.0, -c1, -c2, -c3, 0.0, c3, c2, c1};
    const double sn = cs[(e + 12) & 15];  // sin(x) = cos(x - pi / 2)
    return dbl2{cs[e & 15], sign < 0 ? -sn : sn};
}
template <int R>
__device__ __forceinline__ void dft(dbl2* v, int sign);
template <>
__device__ __forceinline__ void dft<1>(dbl2*, int) {}
template <>
__device__ __forceinline__ void dft<2>(dbl2* v, int) { dft2(v[0], v[1]); }
template <>
__device__ __forceinline__ void dft<4>(dbl2* v, int sign) { dft4(v[0], v[1], v[2], v[3], sign); }
// R = 16 as 4 x 4: m = q + 4 m2, k = k1 + 4 k2; inner DFT_4 over m2, twiddle w16^(q k1), outer DFT_4 over q.  The
// outputs stay where the outer transforms leave them: V_k in v[oidx<16>(k)] = v[4 (k % 4) + k / 4].
template <>
__device__ __forceinline__ void dft<16>(dbl2* v, int sign) {
#pragma unroll
    for (int q = 0; q < 4; q++) dft4(v[q], v[q + 4], v[q + 8], v[q + 12], sign);  // a[q][k1] now at v[q + 4 k1]
#pragma unroll
    for (int q = 1; q < 4; q++)
#pragma unroll
        for (int k1 = 1; k1 < 4; k1++) v[q + 4 * k1] = cmul(v[q + 4 * k1], w16(q * k1, sign));
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) dft4(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3], sign);
}
// register of output k after dft<R>
template <int R>
__device__ __forceinline__ constexpr int oidx(int k) {
    return R == 16 ? 4 * (k & 3) + (k >> 2) : k;
}

// v[m] *= w^(m K) for m = 1 .. R - 1, w = exp(sign 2 pi i / n), from the table entries of K, 2K, 4K and 8K (each
// power at most three products from the table)
template <int R>
__device__ __forceinline__ void twiddle_run(dbl2* v, int K, int sign, const dbl2* tw) {
    const dbl2 w1 = twid_split(tw, K, sign);
    if (R == 2) {
        v[1] = cmul(v[1], w1);
        return;
    }
    const dbl2 w2 = twid_split(tw, 2 * K, sign);
    const dbl2 w3 = cmul(w2, w1);
    v[1] = cmul(v[1], w1);
    v[2] = cmul(v[2], w2);
    v[3] = cmul(v[3], w3);
    if (R == 4) return;
    const dbl2 w4 = twid_split(tw, 4 * K, sign);
    v[4] = cmul(v[4], w4);
    v[5] = cmul(v[5], cmul(w4, w1));
    v[6] = cmul(v[6], cmul(w4, w2));
    v[7] = cmul(v[7], cmul(w4, w3));
    if (R == 8) return;
    const dbl2 w8 = twid_split(tw, 8 * K, sign);
    v[8] = cmul(v[8], w8);
    v[9] = cmul(v[9], cmul(w8, w1));
    v[10] = cmul(v[10], cmul(w8, w2));
    v[11] = cmul(v[11], cmul(w8, w3));
    const dbl2 w12 = cmul(w8, w4);
    v[12] = cmul(v[12], w12);
    v[13] = cmul(v[13], cmul(w12, w1));
    v[14] = cmul(v[14], cmul(w12, w2));
    v[15] = cmul(v[15], cmul(w12, w3));
}

// padded LDS position of entry e of a line: one spare entry per 16, so the stride-16 writes of the first stage fall
// on distinct banks
__device__ __forceinline__ int zp_pidx(int e) { return e + (e >> 4); }

// One Stockham stage of radix R over the group in LDS (line li at lds + li * ld): butterfly j < n / R of a line takes
// entries j + m n / R, twiddles them by w^(m k n / (Ns R)), k = j % Ns, transforms them and puts output m at
// (j - k) R + k + m Ns.  Butterflies b = threadIdx.x + 256 i, line b / (n / R).  Reads all, waits, writes all, waits.
// n / R, and Ns unless it is 1, are multiples of 16, so zp_pidx of the m-th entry is that of the first plus m times
// a constant: the offsets fold into the LDS instructions.
template <int R, int n, int Ns>
__device__ __forceinline__ void zp_stage(dbl2* lds, int ld, int sign, const dbl2* tw) {
    constexpr int nr = n / R;
    constexpr int B = 4096 / R / 256;  // butterflies per thread
    static_assert(nr % 16 == 0 && (Ns == 1 || Ns % 16 == 0), "strides of the padded layout");
    constexpr int rs = nr + nr / 16;                  // padded read stride
    constexpr int ws = Ns == 1 ? 1 : Ns + Ns / 16;    // padded write stride (Ns = 1: od is a multiple of R = 16)
    dbl2 v[B][R];
#pragma unroll
    for (int i = 0; i < B; i++) {
        const int b = threadIdx.x + 256 * i, li = b / nr, j = b - li * nr;
        const dbl2* const src = lds + li * ld + zp_pidx(j);
#pragma unroll
        for (int m = 0; m < R; m++) v[i][m] = src[m * rs];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < B; i++) {
        const int b = threadIdx.x + 256 * i, li = b / nr, j = b - li * nr;
        const int k = j % Ns;
        if (Ns > 1 && k) twiddle_run<R>(v[i], k * (n / (Ns * R)), sign, tw);
        dft<R>(v[i], sign);
        dbl2* const dst = lds + li * ld + zp_pidx((j - k) * R + k);
#pragma unroll
        for (int m = 0; m < R; m++) dst[m * ws] = v[i][oidx<R>(m)];
    }
    __syncthreads();
}

// the last stage of a transform (Ns = n / R, so butterfly j leaves its outputs at j + m Ns in natural order):
// butterflies b = threadIdx.x + 256 i with line b % L and j = b / L, the layout of the loads and stores; the
// outputs stay in registers, output m of butterfly i in v[i][oidx<R>(m)]
template <int R, int n, int L>
__device__ __forceinline__ void zp_last(const dbl2* lds, int ld, int sign, const dbl2* tw, dbl2 (*v)[R]) {
    constexpr int nr = n / R;
    constexpr int B = 4096 / R / 256;
    static_assert(nr % 16 == 0, "strides of the padded layout");
    constexpr int rs = nr + nr / 16;
#pragma unroll
    for (int i = 0; i < B; i++) {
        const int b = threadIdx.x + 256 * i, li = b % L, j = b / L;
        const dbl2* const src = lds + li * ld + zp_pidx(j);
#pragma unroll
        for (int m = 0; m < R; m++) v[i][m] = src[m * rs];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < B; i++) {
        const int j = (threadIdx.x + 256 * i) / L;  // k = j
        if (j && R > 1) twiddle_run<R>(v[i], j, sign, tw);
        dft<R>(v[i], sign);
    }
}

// the entry of the natural line held at position e of the even-odd line (the inverse of nperm)
__device__ __forceinline__ int zp_iperm(int e, int n) { return 2 * e < n ? 2 * e : 2 * (n - 1 - e) + 1; }

template <int R3, bool NEU>
__global__ __launch_bounds__(256) void k_zpoisson(dbl2* __restrict__ dv, long P, long groups, long xcd_per,
                                                  LineFilter filt, const dbl2* __restrict__ twg) {
    constexpr int n = 256 * R3;
    constexpr int L = 16 / R3;
    constexpr int nr1 = n / 16;
    constexpr int B3 = 16 / R3;  // last-stage butterflies per thread (radix R3)
    extern __shared__ dbl2 lds[];
    const int ld = n + n / 16 + 4;  // padded line length: lines L apart start 4 banks apart modulo 64
    dbl2* const tw = lds + L * ld;
    double* const ztab = reinterpret_cast<double*>(tw + split_entries(n));
    double* const pxyt = ztab + n;  // the transverse term of p^2 of each line of the group
    long g, gstep, gend;
    if (xcd_per > 0) {
        const long kx = blockIdx.x % 8;
        g = kx * xcd_per + blockIdx.x / 8;
        gstep = gridDim.x / 8;
        gend = min((kx + 1) * xcd_per, groups);
    } else {
        g = blockIdx.x;
        gstep = gridDim.x;
        gend = groups;
    }
    if (g >= gend) return;
    for (int k = threadIdx.x; k < split_entries(n); k += 256) tw[k] = twg[k];
    for (int k = threadIdx.x; k < n; k += 256)
        ztab[k] = (2.0 - 2.0 * cos((NEU ? 1.0 : 2.0) * kPi * k / (double)n)) / (filt.az * filt.az);
    // the register stages: line t % L, butterfly t / L of the radix-16 stage
    const int li0 = threadIdx.x % L, j0 = threadIdx.x / L;
    auto line_base = [&](long gg, int li) -> long {
        const long l = gg * L + li;
        return (l / P) * P * n + l % P;
    };
    auto set_pxy = [&](long gg) {  // before the first barrier of the group; read after it
        if (threadIdx.x < L) {
            const long t = (gg * L + threadIdx.x) % P;
            const int px = (int)(t % filt.nx), py = (int)(t / filt.nx);
            pxyt[threadIdx.x] = (2.0 - 2.0 * cos(2.0 * kPi * px / (double)filt.nx)) / (filt.ax * filt.ax) +
                                (2.0 - 2.0 * cos(2.0 * kPi * py / (double)filt.ny)) / (filt.ay * filt.ay);
        }
    };
    dbl2 pre[16];
    auto fetch = [&](long gg) {
        const long base = line_base(gg, li0);
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const int e = j0 + m * nr1;
            pre[m] = dv[base + (long)(NEU ? zp_iperm(e, n) : e) * P];
        }
    };
    fetch(g);
    __syncthreads();  // the tables
    while (true) {
        // forward stage 1 (radix 16, Ns = 1) on the loaded registers
        set_pxy(g);
        dft<16>(pre, -1);
        {
            dbl2* const dst = lds + li0 * ld + 17 * j0;  // zp_pidx(16 j0 + m) = 17 j0 + m
#pragma unroll
            for (int m = 0; m < 16; m++) dst[m] = pre[oidx<16>(m)];
        }
        const long gn = g + gstep;
        const bool more = gn < gend;
        if (more) fetch(gn);  // in flight during the stages and the stores below
        __syncthreads();
        // forward stage 2 (radix 16, Ns = 16), then the last one (radix R3, Ns = 256) with the plain filter
        if (R3 > 1) {
            zp_stage<16, n, 16>(lds, ld, -1, tw);
            dbl2 v[B3][R3];
            zp_last<R3, n, L>(lds, ld, -1, tw, v);
#pragma unroll
            for (int i = 0; i < B3; i++) {
                const int b = threadIdx.x + 256 * i, li = b % L, j = b / L;
                const double pxy = pxyt[li];
                dbl2* const dst = lds + li * ld + zp_pidx(j);
#pragma unroll
                for (int m = 0; m < R3; m++) {
                    const int e = j + 256 * m;
                    dbl2 w = v[i][oidx<R3>(m)];
                    if (!NEU) {
                        const double p2 = pxy + ztab[e];
                        w = (p2 > 0.0 ? filt.nf / p2 : (filt.drop_zero ? 0.0 : 1.0)) * w;
                    }
                    dst[272 * m] = w;  // zp_pidx(j + 256 m)
                }
            }
        } else {
            dbl2 v[1][16];
            zp_last<16, n, L>(lds, ld, -1, tw, v);
            const int li = threadIdx.x % L, j = threadIdx.x / L;
            const double pxy = pxyt[li];
#pragma unroll
            for (int m = 0; m < 16; m++) {
                const int e = j + 16 * m;
                dbl2 w = v[0][oidx<16>(m)];
                if (!NEU) {
                    const double p2 = pxy + ztab[e];
                    w = (p2 > 0.0 ? filt.nf / p2 : (filt.drop_zero ? 0.0 : 1.0)) * w;
                }
                lds[li * ld + j + 17 * m] = w;  // zp_pidx(j + 16 m), j < 16
            }
        }
        __syncthreads();
        if (NEU) {
            // the pair step of the cosine transform (see the generic pass): X_k = (w_k V_k + conj(w_k) V_{n-k}) / 2,
            // Y = h X, Q_k = conj(w_k) (Y_k - i Y_{n-k})
            constexpr int np = n / 2 + 1;
            for (int it = threadIdx.x; it < L * np; it += 256) {
                const int li = it / np, k = it - li * np;
                const double pxy = pxyt[li];
                dbl2* const row = lds + li * ld;
                const int kr = k == 0 ? 0 : n - k;
                const dbl2 vk = row[zp_pidx(k)], vr = row[zp_pidx(kr)];
                const dbl2 wk = filt.wq[k], wr = filt.wq[kr];
                const dbl2 cwk = {wk.x, -wk.y}, cwr = {wr.x, -wr.y};
                const double pk = pxy + ztab[k], pr = pxy + ztab[kr];
                const double hk = pk > 0.0 ? filt.nf / pk : 0.0, hr = pr > 0.0 ? filt.nf / pr : 0.0;
                const dbl2 yk = (0.5 * hk) * (cmul(wk, vk) + cmul(cwk, vr));
                const dbl2 yr = (0.5 * hr) * (cmul(wr, vr) + cmul(cwr, vk));
                if (k == 0) {
                    row[0] = yk;
                } else {
                    row[zp_pidx(k)] = cmul(cwk, dbl2{yk.x + yr.y, yk.y - yr.x});
                    row[zp_pidx(kr)] = cmul(cwr, dbl2{yr.x + yk.y, yr.y - yk.x});
                }
            }
            __syncthreads();
        }
        // inverse: stage 1 (radix 16, Ns = 1), stage 2 (radix 16, Ns = 16), the last stage from LDS to memory
        zp_stage<16, n, 1>(lds, ld, +1, tw);
        if (R3 > 1) {
            zp_stage<16, n, 16>(lds, ld, +1, tw);
            dbl2 v[B3][R3];
            zp_last<R3, n, L>(lds, ld, +1, tw, v);
#pragma unroll
            for (int i = 0; i < B3; i++) {
                const int b = threadIdx.x + 256 * i, li = b % L, j = b / L;
                const long base = line_base(g, li);
#pragma unroll
                for (int m = 0; m < R3; m++) {
                    const int e = j + 256 * m;
                    dv[base + (long)(NEU ? zp_iperm(e, n) : e) * P] = v[i][oidx<R3>(m)];
                }
            }
        } else {
            dbl2 v[1][16];
            zp_last<16, n, L>(lds, ld, +1, tw, v);
            const int li = threadIdx.x % L, j = threadIdx.x / L;
            const long base = line_base(g, li);
#pragma unroll
            for (int m = 0; m < 16; m++) {
                const int e = j + 16 * m;
                dv[base + (long)(NEU ? zp_iperm(e, n) : e) * P] = v[0][oidx<16>(m)];
            }
        }
        if (!more) break;
        __syncthreads();  // every read of this group's LDS precedes the next group's writes
        g = gn;
    }
}

// ---- strided lines of length n = 256 R3 (R3 in {1, 2, 4}), one transform -----------------------------------------
// The scheme of k_zpoisson without the filter: groups of L = 16 / R3 lines whose starts are neighbours (line l
// starts at (l % b1) s1 + (l / b1) s2, s1 = 1, L dividing b1; rows of 16 L bytes), the first stage on the loaded
// registers, the last one stored from registers.  At 256 x 256 x 1024 the y pass of the four packed arrays
// (length 256, stride 256) is one of these.
template <int R3>
__global__ __launch_bounds__(256) void k_lines16(dbl2* __restrict__ dv, long es, long b1, long s2, long groups,
                                                 long xcd_per, int sign, const dbl2* __restrict__ twg) {
    constexpr int n = 256 * R3;
    constexpr int L = 16 / R3;
    constexpr int nr1 = n / 16;
    constexpr int B3 = 16 / R3;
    extern __shared__ dbl2 lds[];
    const int ld = n + n / 16 + 4;
    dbl2* const tw = lds + L * ld;
    long g, gstep, gend;
    if (xcd_per > 0) {
        const long kx = blockIdx.x % 8;
        g = kx * xcd_per + blockIdx.x / 8;
        gstep = gridDim.x / 8;
        gend = min((kx + 1) * xcd_per, groups);
    } else {
        g = blockIdx.x;
        gstep = gridDim.x;
        gend = groups;
    }
    if (g >= gend) return;
    for (int k = threadIdx.x; k < split_entries(n); k += 256) tw[k] = twg[k];
    const int li0 = threadIdx.x % L, j0 = threadIdx.x / L;
    auto line_start = [&](long gg, int li) -> long {
        const long l = gg * L + li;
        return l % b1 + (l / b1) * s2;
    };
    dbl2 pre[16];
    auto fetch = [&](long gg) {
        const dbl2* const src = dv + line_start(gg, li0) + (long)j0 * es;
#pragma unroll
        for (int m = 0; m < 16; m++) pre[m] = src[(long)m * nr1 * es];
    };
    fetch(g);
    __syncthreads();  // the twiddles
    while (true) {
        dft<16>(pre, sign);
        {
            dbl2* const dst = lds + li0 * ld + 17 * j0;  // zp_pidx(16 j0 + m) = 17 j0 + m
#pragma unroll
            for (int m = 0; m < 16; m++) dst[m] = pre[oidx<16>(m)];
        }
        const long gn = g + gstep;
        const bool more = gn < gend;
        if (more) fetch(gn);
        __syncthreads();
        if (R3 > 1) {
            zp_stage<16, n, 16>(lds, ld, sign, tw);
            dbl2 v[B3][R3];
            zp_last<R3, n, L>(lds, ld, sign, tw, v);
#pragma unroll
            for (int i = 0; i < B3; i++) {
                const int b = threadIdx.x + 256 * i, li = b % L, j = b / L;
                dbl2* const dst = dv + line_start(g, li) + (long)j * es;
#pragma unroll
                for (int m = 0; m < R3; m++) dst[(long)256 * m * es] = v[i][oidx<R3>(m)];
            }
        } else {
            dbl2 v[1][16];
            zp_last<16, n, L>(lds, ld, sign, tw, v);
            const int li = threadIdx.x % L, j = threadIdx.x / L;
            dbl2* const dst = dv + line_start(g, li) + (long)j * es;
#pragma unroll
            for (int m = 0; m < 16; m++) dst[(long)16 * m * es] = v[0][oidx<16>(m)];
        }
        if (!more) break;
        __syncthreads();
        g = gn;
    }
}

// the dynamic LDS limit of the k_zpoisson instantiations raised to the device's, once per device
std::map<int, bool> g_zp_allowed;
template <int R3, bool NEU>
hipError_t zp_allow_one(int v) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_zpoisson<R3, NEU>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, v);
}
template <int R3>
hipError_t l16_allow_one(int v) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_lines16<R3>), hipFuncAttributeMaxDynamicSharedMemorySize,
                               v);
}
hipError_t zp_allow(size_t lds_max) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(g_tw_mu);
    if (g_zp_allowed.count(dev)) return hipSuccess;
    const int v = (int)lds_max;
    for (hipError_t r : {zp_allow_one<1, false>(v), zp_allow_one<1, true>(v), zp_allow_one<2, false>(v),
                         zp_allow_one<2, true>(v), zp_allow_one<4, false>(v), zp_allow_one<4, true>(v),
                         l16_allow_one<1>(v), l16_allow_one<2>(v), l16_allow_one<4>(v)})
        if (r != hipSuccess) return r;
    g_zp_allowed[dev] = true;
    return hipSuccess;
}

template <int R3, bool NEU>
hipError_t launch_zpoisson(hipStream_t st, size_t lds, int cus, double* c, long P, long groups,
                           const LineFilter& f, const dbl2* tw) {
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(&k_zpoisson<R3, NEU>), 256, lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    const long xcd_per = groups >= 64 ? (groups + 7) / 8 : 0;
    const long resident = (long)cus * per_cu;
    const long blocks = xcd_per > 0 ? 8 * std::min(xcd_per, (resident + 7) / 8) : std::min(groups, resident);
    hipLaunchKernelGGL((k_zpoisson<R3, NEU>), dim3((unsigned)blocks), dim3(256), lds, st, reinterpret_cast<dbl2*>(c), P,
                       groups, xcd_per, f, tw);
    return hipGetLastError();
}

// the filtered z pass by k_zpoisson where it applies (Nz in {256, 512, 1024}, P a multiple of 16 / (Nz / 256));
// returns 1 where it does not
int zpoisson(hipStream_t st, double* c, const Geom& g, int ncomp, const LineFilter& f) {
    const int n = g.nz;
    if (n != 256 && n != 512 && n != 1024) return 1;
    const int L = 4096 / n;
    const long P = (long)g.nx * g.ny;
    if (P % L) return 1;
    FftDevice dev;
    const hipError_t de = fft_device(&dev);
    if (de != hipSuccess) return fft_fail_hip(de, "device query");
    const size_t lds = sizeof(dbl2) * ((size_t)L * (n + n / 16 + 4) + split_entries(n)) + sizeof(double) * (n + L);
    if (lds > dev.lds) return 1;
    const hipError_t ae = zp_allow(dev.lds);
    if (ae != hipSuccess) return fft_fail_hip(ae, "k_zpoisson LDS limit");
    const dbl2* tw = nullptr;
    const hipError_t te = twiddles_split(n, &tw);
    if (te != hipSuccess) return fft_fail_hip(te, "twiddle table");
    const long groups = P * ncomp / L;
    hipError_t ke;
    const bool neu = f.neumann != 0;
    if (n == 256) ke = neu ? launch_zpoisson<1, true>(st, lds, dev.cus, c, P, groups, f, tw)
                           : launch_zpoisson<1, false>(st, lds, dev.cus, c, P, groups, f, tw);
    else if (n == 512) ke = neu ? launch_zpoisson<2, true>(st, lds, dev.cus, c, P, groups, f, tw)
                                : launch_zpoisson<2, false>(st, lds, dev.cus, c, P, groups, f, tw);
    else ke = neu ? launch_zpoisson<4, true>(st, lds, dev.cus, c, P, groups, f, tw)
                  : launch_zpoisson<4, false>(st, lds, dev.cus, c, P, groups, f, tw);
    if (ke != hipSuccess)
        return fft_fail_hip(ke, "k_zpoisson launch (length " + std::to_string(n) + ", " + std::to_string(lds) +
                                    " bytes of LDS)");
    return 0;
}

template <int R3>
hipError_t launch_lines16(hipStream_t st, size_t lds, int cus, double* c, long es, long b1, long s2, long groups,
                          int sign, const dbl2* tw) {
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(&k_lines16<R3>), 256, lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    const long xcd_per = groups >= 64 ? (groups + 7) / 8 : 0;
    const long resident = (long)cus * per_cu;
    const long blocks = xcd_per > 0 ? 8 * std::min(xcd_per, (resident + 7) / 8) : std::min(groups, resident);
    hipLaunchKernelGGL((k_lines16<R3>), dim3((unsigned)blocks), dim3(256), lds, st, reinterpret_cast<dbl2*>(c), es, b1,
                       s2, groups, xcd_per, sign, tw);
    return hipGetLastError();
}

// strided lines of length 256, 512 or 1024 by k_lines16 where their starts come in runs of 4096 / n neighbours
// (s1 = 1, b1 a multiple of the run); returns 1 where it does not apply
int lines16(hipStream_t st, double* c, int n, long es, long nlines, long b1, long s1, long s2, int sign) {
    if (!kZPoisson || (n != 256 && n != 512 && n != 1024) || es == 1 || s1 != 1) return 1;
    const int L = 4096 / n;
    if (b1 % L || nlines % L || (b1 < nlines && nlines % b1)) return 1;
    FftDevice dev;
    const hipError_t de = fft_device(&dev);
    if (de != hipSuccess) return fft_fail_hip(de, "device query");
    const size_t lds = sizeof(dbl2) * ((size_t)L * (n + n / 16 + 4) + split_entries(n));
    if (lds > dev.lds) return 1;
    const hipError_t ae = zp_allow(dev.lds);
    if (ae != hipSuccess) return fft_fail_hip(ae, "k_lines16 LDS limit");
    const dbl2* tw = nullptr;
    const hipError_t te = twiddles_split(n, &tw);
    if (te != hipSuccess) return fft_fail_hip(te, "twiddle table");
    const long groups = nlines / L;
    hipError_t ke;
    if (n == 256) ke = launch_lines16<1>(st, lds, dev.cus, c, es, b1, s2, groups, sign, tw);
    else if (n == 512) ke = launch_lines16<2>(st, lds, dev.cus, c, es, b1, s2, groups, sign, tw);
    else ke = launch_lines16<4>(st, lds, dev.cus, c, es, b1, s2, groups, sign, tw);
    if (ke != hipSuccess)
        return fft_fail_hip(ke, "k_lines16 launch (length " + std::to_string(n) + ", " + std::to_string(lds) +
                                    " bytes of LDS)");
    return 0;
}

int fft_lines_filtered(hipStream_t st, double* data, int n, long es, long nlines, long b1, long s1, long s2, int sign,
                       const LineFilter& filt) {
    if (n <= 1 || nlines <= 0) return 0;
    if (filt.on && es == 1) return fft_fail(GLASMA_ERR_ARG, "a filtered pass needs strided lines");
    if (n > kFftLds) return fft_fail(GLASMA_ERR_ARG, "length " + std::to_string(n) + " is above 2048");
    Radices R;
    if (factorize(n, R) != 0)
        return fft_fail(GLASMA_ERR_ARG, "length " + std::to_string(n) + " has a prime factor above 13");
    const dbl2* tw = nullptr;
    const hipError_t te = filt.on ? twiddles_split(n, &tw) : twiddles(n, &tw);
    if (te != hipSuccess) return fft_fail_hip(te, "twiddle table");
    // contiguous lines are transformed as one packed run: line l must start at l * n
    if (es == 1 && !((s1 == n && (b1 >= nlines || s2 == b1 * n)) || (b1 == 1 && s2 == n)))
        return fft_fail(GLASMA_ERR_ARG, "contiguous lines must lie end to end");
    FftDevice dev;
    const hipError_t de = fft_device(&dev);
    if (de != hipSuccess) return fft_fail_hip(de, "device query");
    // lines per group; a group is at most 16 elements per thread (the register prefetch)
    int lpb = 1;
    if (es == 1) {
        while (2 * lpb * (n + 1) <= kContigEntries && 2 * lpb <= kMaxLinesPerBlock) lpb *= 2;
    } else {
        const size_t cap = dev.lds / (2 * sizeof(dbl2));  // entries per buffer
        const int most = filt.on ? kFilterLines : kStridedLines;
        while ((size_t)(2 * lpb) * (n + 1) <= cap && 2 * lpb <= most && 2 * lpb * n <= 16 * kFftThreads)
            lpb *= 2;
    }
    const int ld = lpb > 1 ? n + 1 : n;
    const long groups = (nlines + lpb - 1) / lpb;
    const long xcd_per = es != 1 && groups >= 64 ? (groups + 7) / 8 : 0;
    // two buffers, and with a filter the two-level twiddle table and the table of its z term
    const size_t bytes =
        2 * sizeof(dbl2) * (size_t)lpb * ld + (filt.on ? sizeof(dbl2) * split_entries(n) + sizeof(double) * n : 0);
    if (bytes > dev.lds) return fft_fail(GLASMA_ERR_ARG, "length " + std::to_string(n) + " exceeds the LDS");
    const bool twlds = !filt.on && bytes + sizeof(dbl2) * n <= dev.lds;
    const size_t lds = bytes + (twlds ? sizeof(dbl2) * n : 0);
    const int need = (lpb * n + kFftThreads - 1) / kFftThreads;  // elements per thread
    const int maxe = need <= 2 ? 2 : need <= 4 ? 4 : need <= 8 ? 8 : 16;
    hipError_t ke;
#define GLASMA_FFT_LAUNCH(TW, M, F)                                                                                 \
    ke = launch_lines<TW, M, F>(st, lds, dev.cus, data, R, es, nlines, b1, s1, s2, sign, lpb, ld, groups, xcd_per, \
                                filt, tw)
#define GLASMA_FFT_LAUNCH_M(TW, F)                   \
    switch (maxe) {                                  \
        case 2: GLASMA_FFT_LAUNCH(TW, 2, F); break;  \
        case 4: GLASMA_FFT_LAUNCH(TW, 4, F); break;  \
        case 8: GLASMA_FFT_LAUNCH(TW, 8, F); break;  \
        default: GLASMA_FFT_LAUNCH(TW, 16, F); break; \
    }
    if (twlds) {
        if (filt.on) GLASMA_FFT_LAUNCH_M(true, true) else GLASMA_FFT_LAUNCH_M(true, false)
    } else {
        if (filt.on) GLASMA_FFT_LAUNCH_M(false, true) else GLASMA_FFT_LAUNCH_M(false, false)
    }
#undef GLASMA_FFT_LAUNCH_M
#undef GLASMA_FFT_LAUNCH
    if (ke != hipSuccess)
        return fft_fail_hip(ke, "k_fft_lines launch (length " + std::to_string(n) + ", " + std::to_string(lpb) +
                                    " lines per block, " + std::to_string(lds) + " bytes of LDS)");
    return 0;
}

}  // namespace

int fft_lines(hipStream_t st, double* data, int n, long es, long nlines, long b1, long s1, long s2, int sign) {
    if (nlines > 0) {
        const int rc = lines16(st, data, n, es, nlines, b1, s1, s2, sign);
        if (rc != 1) return rc;
    }
    const LineFilter none = {0, 0, 0, 0, 0.0, 0.0, 0.0, 0.0, 0, 0, nullptr};
    return fft_lines_filtered(st, data, n, es, nlines, b1, s1, s2, sign, none);
}

// c[(comp*nzl + z)*P + y*nx + x]
namespace {
int fft_x(hipStream_t st, double* c, const Geom& g, int ncomp, int nzl, int sign) {
    return fft_lines(st, c, g.nx, 1, (long)g.ny * nzl * ncomp, (long)g.ny * nzl * ncomp, g.nx, 0, sign);
}
int fft_y(hipStream_t st, double* c, const Geom& g, int ncomp, int nzl, int sign) {
    return fft_lines(st, c, g.ny, g.nx, (long)g.nx * nzl * ncomp, g.nx, 1, (long)g.nx * g.ny, sign);
}
}  // namespace

int fft_xy(hipStream_t st, double* c, const Geom& g, int ncomp, int nzl, int sign) {
    const int rc = fft_x(st, c, g, ncomp, nzl, sign);
    return rc ? rc : fft_y(st, c, g, ncomp, nzl, sign);
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

int fft_3d_poisson(hipStream_t st, double* c, const Geom& g, int ncomp, const double a[3], bool drop_zero,
                   bool neumann, bool fwd_x, bool inv_x) {
    if (g.nz < 2) return fft_fail(GLASMA_ERR_ARG, "the fused Poisson filter needs Nz >= 2");
    if (neumann && !drop_zero) return fft_fail(GLASMA_ERR_ARG, "the Neumann filter drops the zero mode");
    const long P = (long)g.nx * g.ny;
    const dbl2* wq = nullptr;
    if (neumann) {
        const hipError_t we = quarter_wave(g.nz, &wq);
        if (we != hipSuccess) return fft_fail_hip(we, "quarter-wave table");
    }
    const LineFilter f = {1, g.nx, g.ny, P, a[0], a[1], a[2], 1.0 / (double)g.S, drop_zero ? 1 : 0,
                          neumann ? 1 : 0, wq};
    int rc = fwd_x ? fft_x(st, c, g, ncomp, g.nz, -1) : 0;
    if (!rc) rc = fft_y(st, c, g, ncomp, g.nz, -1);
    if (!rc && kZPoisson) {
        rc = zpoisson(st, c, g, ncomp, f);
        if (rc == 1) rc = fft_lines_filtered(st, c, g.nz, P, P * ncomp, P, 1, P * g.nz, -1, f);
    } else if (!rc) {
        rc = fft_lines_filtered(st, c, g.nz, P, P * ncomp, P, 1, P * g.nz, -1, f);
    }
    if (!rc) rc = fft_y(st, c, g, ncomp, g.nz, +1);
    if (!rc && inv_x) rc = fft_x(st, c, g, ncomp, g.nz, +1);
    return rc;
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
    (void)hipFree(dev);
    if (rc) (void)hipGetLastError();  // leave no stale error for the next launch check
    return rc;
}
