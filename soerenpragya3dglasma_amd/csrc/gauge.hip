// Coulomb gauge fixing on the device: Fourier-accelerated steepest descent.
//   CoulombGaugeFixing::SetCoulombGauge              SC/GAUGETRANSFORMATION/CoulombGaugeFixing.cpp:32-88
//   FourierAccelerationAlgorithm::UpdateGaugeTransformation
//                                                    SC/GAUGETRANSFORMATION/ALGORITHMS/FourierAcceleration.cpp:15-148
//   GET_LOCAL_GAUGE_VIOLATION                        SC/GAUGETRANSFORMATION/CoulombGaugeDeviation.cpp:23-41
//   GaugeTransformLinks / electric fields            SC/GAUGETRANSFORMATION/GaugeTransformation.cpp:93-164,
//                                                    SC/CALC/GaugeTransformation.cpp:5-17
//
// One iteration: Delta^a(x) = Re tr(i t^a w(x)), w = sum_mu a_mu^-2 (U^G_mu(x) + U^G_mu^dag(x - mu)) (the z
// term only for Nz > 1); Delta -> 3D FFT -> Delta~/(V p^2) for p^2 > 0 (periodic momenta in all three
// directions, the zero mode passed through unscaled) -> inverse FFT -> g <- exp(-i alpha Delta t)^dag g,
// U^G_mu(x) = g(x) U_mu(x) g^dag(x + mu).  Monitors (max / RMS of Delta and of the step, min / mean of the
// functional Re tr w) come from deterministic two-level reductions.
//
// The eight real components travel as four complex arrays, Delta^{2k} + i Delta^{2k+1}: the filter is real and
// even in the momentum, so by linearity the inverse transform returns the two filtered components as its real
// and imaginary parts.  That is half the transforms of one complex array per component (as the reference runs
// them), at a round-off-level difference.
//
// The reference's two other update algorithms (selected there by #define, CoulombGaugeFixing.cpp:4,55-67) share
// w, Delta, the monitors and the link transformation:
//   Jacobi      ALGORITHMS/Jacobi.cpp:13-74      g <- exp(-i alpha Delta t)^dag g at every site, no FFT;
//   Los Alamos  ALGORITHMS/LosAlamos.cpp:13-87   even, then odd checkerboard sites ((x + y + z) % 2):
//               g <- MaxTraceProjection(w)^dag g (seven Cabibbo-Marinari sweeps, SUNc/SU3/GroupOperations.cpp:358-378,
//               SUNc/SU3/util/CabbiboMarinariProjection.cpp), links re-transformed after each half sweep; the
//               monitors cover both halves.  Over-relaxation probability 0 only (the reference's value).
//
// Ghost-layer behaviour of the reference, reproduced: the gauge-fixed links' z ghost layers are never
// written, so U^G(z = -1) is the identity; the gauge transformation's top ghost layer is synchronised
// after each link transformation, so a transformation sees the top layer of the previous iteration.  Here
// that layer is the separate array Gtop (18 planes over one transverse plane).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "fft_lds.h"
#include "internal.h"

namespace glasma {
namespace {

constexpr int kCgBlock = 256;

inline dim3 grid1(long n) { return dim3((unsigned)((n + kCgBlock - 1) / kCgBlock)); }

// Block order of the site kernels that read neighbours (k_cg_dev_fused, k_cg_transform, k_cg_losalamos).  A block
// takes 256 consecutive sites of one z-plane (a unit).  With U = P / 256 units per plane divisible by 8 h, block b
// runs on XCD b % 8 (the hardware deals workgroups to the eight XCDs in turn), which owns units [k U/8, (k+1) U/8)
// of every plane and walks them in chunks of h units, z fastest: a site's y neighbours lie in its own unit or the
// one before or after it on the same XCD, its z neighbours h units back or ahead, so both are still in that XCD's
// L2 (at 256 x 256 x 1024, h = 16 units of links and G are 2.4 MB against 4 MB of L2).  In flat order a unit's y
// neighbours run on other XCDs and its z neighbours a whole plane (37 MB) apart: the deviation kernel then read
// 2.5 times its compulsory bytes.  Otherwise (P not a multiple of 256 sites) blocks take units in flat order.
struct CgOrder {
    long upx;  // units per plane per XCD; 0: flat order
    int h;     // units per chunk
};

CgOrder cg_order(const Geom& g) {
    static const int want = [] {
        const char* e = std::getenv("GLASMA_CG_CHUNK");
        const int v = e ? std::atoi(e) : 16;
        return v >= 1 ? v : 16;
    }();
    if (g.P % kCgBlock) return {0, 0};
    const long U = g.P / kCgBlock;
    int h = want;
    while (h > 1 && U % (8L * h)) h /= 2;
    if (U % (8L * h)) return {0, 0};
    return {U / 8, h};
}

// the first site of this block's unit plus threadIdx.x (callers check s < S for the flat order's last block)
__device__ __forceinline__ long cg_site(const Geom& g, const CgOrder& o) {
    if (o.upx == 0) return (long)blockIdx.x * kCgBlock + threadIdx.x;
    const long b = blockIdx.x;
    const long k = b & 7, i = b >> 3;
    const long per_chunk = (long)g.nz * o.h;
    const long chunk = i / per_chunk, rem = i - chunk * per_chunk;
    const long z = rem / o.h, j = rem - z * o.h;
    return z * g.P + (k * o.upx + chunk * o.h + j) * kCgBlock + threadIdx.x;
}

// U^G_mu(x) = G(x) U_mu(x) G^dag(x + mu); the z neighbour of the top layer is Gtop.  EF (the end of a gauge-fixing
// call): also E^G_mu(x) = G(x) E_mu(x) G^dag(x), with G(x) read once for both.
template <bool EF>
__global__ __launch_bounds__(kCgBlock) void k_cg_transform(const double* __restrict__ U, const double* __restrict__ G,
                                                            const double* __restrict__ Gtop, double* __restrict__ UG,
                                                            const double* __restrict__ E, double* __restrict__ EG,
                                                            Geom g, CgOrder o) {
    const long s = cg_site(g, o);
    if (s >= g.S) return;
    const int z = (int)(s / g.P);
    const long s2 = s - (long)z * g.P;
    const int x = (int)(s2 % g.nx), y = (int)(s2 / g.nx);
    M3 Gx, Um, Gn, T, R;
    ldm<long>(G, kGPlanes, 0, s, Gx);
    for (int mu = 0; mu < 3; mu++) {
        ldm<long>(U, kUPlanes, 18 * mu, s, Um);
        if (mu == 0) ldm<long>(G, kGPlanes, 0, sidx(g, x + 1, y, z), Gn);
        else if (mu == 1) ldm<long>(G, kGPlanes, 0, sidx(g, x, y + 1, z), Gn);
        else if (z + 1 < g.nz) ldm<long>(G, kGPlanes, 0, s + g.P, Gn);
        else ldm<long>(Gtop, kGPlanes, 0, s2, Gn);
        mul_nn(Gx, Um, T);
        mul_nd(T, Gn, R);
        stm<long>(UG, kUPlanes, 18 * mu, s, R);
    }
    if (EF)
        for (int mu = 0; mu < 3; mu++) {
            double e[8], f[8];
            ldv<long>(E, kEPlanes, 8 * mu, s, e);
            adj_mul(Gx, e, f);
            stv<long>(EG, kEPlanes, 8 * mu, s, f);
        }
}

// ghost synchronisation of G's top layer: open -> copy of layer Nz-1, periodic -> copy of layer 0
__global__ void k_cg_sync_top(const double* __restrict__ G, double* __restrict__ Gtop, Geom g) {
    long s2 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s2 >= g.P) return;
    M3 A;
    ldm<long>(G, kGPlanes, 0, s2 + (g.bc ? 0 : (long)(g.nz - 1) * g.P), A);
    stm<long>(Gtop, kGPlanes, 0, s2, A);
}

__global__ void k_cg_identity(double* __restrict__ G, long n) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    M3 A;
    set_unit(A);
    stm<long>(G, kGPlanes, 0, s, A);
}

// per-block partials: [0] sum Delta^2, [1] sum Re tr w, [2] max |Delta^a|, [3] min Re tr w
__device__ void block_partials(double s0, double s1, double m0, double m1, double* out) {
    __shared__ double sh[4][kCgBlock / 64];
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    m0 = wave_max(m0);
    m1 = -wave_max(-m1);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sh[0][w] = s0;
        sh[1][w] = s1;
        sh[2][w] = m0;
        sh[3][w] = m1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0, c = 0, d = sh[3][0];
        for (int k = 0; k < kCgBlock / 64; k++) {
            a += sh[0][k];
            b += sh[1][k];
            c = fmax(c, sh[2][k]);
            d = fmin(d, sh[3][k]);
        }
        out[0] = a;
        out[1] = b;
        out[2] = c;
        out[3] = d;
    }
}

// w(x) = sum_mu c_mu (U^G_mu(x) + U^G_mu(x - mu)^dag), the z term only for Nz > 1; below z = 0 the gauge-fixed
// links' ghost layer is the identity (GET_LOCAL_GAUGE_VIOLATION, CoulombGaugeDeviation.cpp:23-41)
__device__ __forceinline__ void local_w(const double* __restrict__ UG, const Geom& g, long s, int x, int y, int z,
                                        double cx, double cy, double cz, M3& W) {
    M3 A, B;
    const double zf = g.nz > 1 ? 1.0 : 0.0;
    for (int k = 0; k < 9; k++) { W.re[k] = 0.0; W.im[k] = 0.0; }
    for (int mu = 0; mu < 3; mu++) {
        const double c = mu == 0 ? cx : (mu == 1 ? cy : zf * cz);
        ldm<long>(UG, kUPlanes, 18 * mu, s, A);
        if (mu == 0) ldm<long>(UG, kUPlanes, 0, sidx(g, x - 1, y, z), B);
        else if (mu == 1) ldm<long>(UG, kUPlanes, 18, sidx(g, x, y - 1, z), B);
        else if (z > 0) ldm<long>(UG, kUPlanes, 36, s - g.P, B);
        else set_unit(B);
        // w += c (A + B^dag), grouped as the reference adds the three directions
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                W.re[mi(i, j)] += c * (A.re[mi(i, j)] + B.re[mi(j, i)]);
                W.im[mi(i, j)] += c * (A.im[mi(i, j)] - B.im[mi(j, i)]);
            }
    }
}

// The ten linear functions of w that the iteration reads (retr_igen's eight Re tr(i t^a w) and Re tr w) as
// accumulators: q[0..5] = Im(w01 + w10), Re(w01 - w10), Im(w02 + w20), Re(w02 - w20), Im(w12 + w21), Re(w12 - w21);
// q[6..8] = Im w00, Im w11, Im w22; q[9] = Re tr w.  q += c X Y^dag, element by element.
__device__ __forceinline__ void acc_proj_nd(double c, const M3& X, const M3& Y, double* q) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double r = 0.0, m = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                r = fma(X.re[mi(i, k)], Y.re[mi(j, k)], r);
                r = fma(X.im[mi(i, k)], Y.im[mi(j, k)], r);
                m = fma(X.im[mi(i, k)], Y.re[mi(j, k)], m);
                m = fma(-X.re[mi(i, k)], Y.im[mi(j, k)], m);
            }
            if (i == j) {
                q[6 + i] = fma(c, m, q[6 + i]);
                q[9] = fma(c, r, q[9]);
            } else {
                const int p = i + j - 1;  // (0,1) -> 0, (0,2) -> 1, (1,2) -> 2
                q[2 * p] = fma(c, m, q[2 * p]);
                q[2 * p + 1] = fma(i < j ? c : -c, r, q[2 * p + 1]);
            }
        }
}

// local_w with the gauge-fixed links formed on the fly from U and G (the Fourier and Jacobi iterations): the
// iteration then needs no separate link transformation writing U^G and reading it back.  Each term goes straight into
// the accumulators of acc_proj_nd: c (G(x) U_mu(x)) G(x + mu)^dag and c G(x) (G(x - mu) U_mu(x - mu))^dag, so that
// neither the links nor w are held whole and four matrices stay live instead of seven (the deviation kernel had two
// waves per SIMD, now three).  The sums differ from local_w's grouping c (A + B^dag) at round-off.
// CONS: the boundary-consistent update (no counterpart in the reference): only links that exist enter w (open z:
// none from the top layer to its ghost copy and none below layer 0; periodic z: the wrapped links, with the
// current G on both ends instead of the reference's lagging ghost layer and identity below layer 0).
// the direction-mu terms of w at site s = (x, y, z), G(x) in Gx: forward c (G(x) U_mu(x)) G(x + mu)^dag and backward
// c G(x) (G(x - mu) U_mu(x - mu))^dag
template <bool CONS>
__device__ __forceinline__ void w_terms(const double* __restrict__ U, const double* __restrict__ G,
                                        const double* __restrict__ Gtop, const Geom& g, int mu, long s, int x, int y,
                                        int z, double c, const M3& Gx, double* q) {
    M3 Um, Gn, T;
    // U^G_mu(x) = (G(x) U_mu(x)) G^dag(x + mu)
    const bool has_up = !CONS || mu < 2 || (g.nz > 1 && (g.bc || z + 1 < g.nz));
    if (has_up) {
        ldm<long>(U, kUPlanes, 18 * mu, s, Um);
        mul_nn(Gx, Um, T);
        if (mu == 0) ldm<long>(G, kGPlanes, 0, sidx(g, x + 1, y, z), Gn);
        else if (mu == 1) ldm<long>(G, kGPlanes, 0, sidx(g, x, y + 1, z), Gn);
        else if (z + 1 < g.nz) ldm<long>(G, kGPlanes, 0, s + g.P, Gn);
        else if (CONS) ldm<long>(G, kGPlanes, 0, s - (long)z * g.P, Gn);  // periodic: layer 0 now
        else ldm<long>(Gtop, kGPlanes, 0, s - (long)z * g.P, Gn);
        acc_proj_nd(c, T, Gn, q);
    }
    // U^G_mu(x - mu)^dag = G(x) (G(x - mu) U_mu(x - mu))^dag; below z = 0 the identity (CONS: none for open z,
    // the top layer's link for periodic z)
    const bool has_dn = mu < 2 || z > 0 || (CONS && g.bc && g.nz > 1);
    if (has_dn) {
        const long sb = mu == 0 ? sidx(g, x - 1, y, z)
                                : (mu == 1 ? sidx(g, x, y - 1, z) : (z > 0 ? s - g.P : s + (long)(g.nz - 1) * g.P));
        ldm<long>(G, kGPlanes, 0, sb, Gn);
        ldm<long>(U, kUPlanes, 18 * mu, sb, Um);
        mul_nn(Gn, Um, T);
        acc_proj_nd(c, Gx, T, q);
    } else if (!CONS) {
        q[9] += 3.0 * c;
    }
}

template <bool CONS>
__device__ __forceinline__ void local_w_fused(const double* __restrict__ U, const double* __restrict__ G,
                                              const double* __restrict__ Gtop, const Geom& g, int x, int y, int z,
                                              double cx, double cy, double cz, double* q) {
    M3 Gx;
    const double zf = g.nz > 1 ? 1.0 : 0.0;
    const long s = sidx(g, x, y, z);
    ldm<long>(G, kGPlanes, 0, s, Gx);  // G(x), read once: it starts every forward term and ends every backward one
    for (int k = 0; k < 10; k++) q[k] = 0.0;
#pragma unroll 1
    for (int mu = 0; mu < 3; mu++)
        w_terms<CONS>(U, G, Gtop, g, mu, s, x, y, z, mu == 0 ? cx : (mu == 1 ? cy : zf * cz), Gx, q);
}

// Delta^a = Re tr(i t^a w) (retr_igen) from the accumulators of acc_proj_nd
__device__ __forceinline__ void delta_from_proj(const double* q, double* dv) {
    dv[0] = -0.5 * q[0];
    dv[1] = -0.5 * q[1];
    dv[2] = -0.5 * (q[6] - q[7]);
    dv[3] = -0.5 * q[2];
    dv[4] = -0.5 * q[3];
    dv[5] = -0.5 * q[4];
    dv[6] = -0.5 * q[5];
    dv[7] = -0.5 * (q[6] + q[7] - 2.0 * q[8]) * kInvSqrt3;
}

// XF: the forward x transform of the Fourier step fused into the deviation kernel and the inverse one into the update.
// Each block then holds 256 / Nx whole x lines of each of the four packed arrays (Nx divides 256, S a multiple of
// 256), so it transforms them in LDS with the Stockham stages of the transform passes and the arrays make two fewer
// trips through HBM per iteration.
constexpr int kXfEntries = 4 * kCgBlock;  // complex entries of one LDS buffer: the block's lines of the four arrays

// Delta and the functional (CoulombGaugeDeviation.cpp:23-41, FourierAcceleration.cpp:20-52), with the monitors,
// from the ungauged links and G.  XF: the four arrays leave x-transformed (R: the factors of Nx, tw: its twiddles).
template <bool CONS, bool XF>
__global__ __launch_bounds__(kCgBlock, 3) void k_cg_dev_fused(const double* __restrict__ U, const double* __restrict__ G,
                                                            const double* __restrict__ Gtop, double* __restrict__ F,
                                                            double* __restrict__ part, Geom g, double cx, double cy,
                                                            double cz, double fm0, CgOrder o, Radices R,
                                                            const dbl2* __restrict__ tw) {
    const long s = cg_site(g, o);
    __shared__ dbl2 xb[XF ? 2 * kXfEntries : 1];
    double sq = 0.0, fs = 0.0, mx = 0.0, fm = fm0;
    if (s < g.S) {
        const int z = (int)(s / g.P);
        const long s2 = s - (long)z * g.P;
        const int x = (int)(s2 % g.nx), y = (int)(s2 / g.nx);
        double q[10];
        local_w_fused<CONS>(U, G, Gtop, g, x, y, z, cx, cy, cz, q);
        double dv[8];
        delta_from_proj(q, dv);
        for (int a = 0; a < 8; a++) {
            mx = fmax(mx, fabs(dv[a]));
            sq += dv[a] * dv[a];
        }
        // two real components per complex array: Delta^{2k} + i Delta^{2k+1} in array k
        if (XF) {
            for (int k = 0; k < 4; k++) xb[k * kCgBlock + threadIdx.x] = dbl2{dv[2 * k], dv[2 * k + 1]};
        } else {
            for (int k = 0; k < 4; k++) {
                F[2 * ((long)k * g.S + s)] = dv[2 * k];
                F[2 * ((long)k * g.S + s) + 1] = dv[2 * k + 1];
            }
        }
        const double lf = q[9];
        fm = fmin(fm, lf);
        fs = lf;
    }
    if (XF) {  // every block is whole: all threads hold a site
        __syncthreads();
        const int cur = stockham(xb, kXfEntries, 0, R, kXfEntries / R.n, R.n, -1, tw);
        dbl2* const Fv = reinterpret_cast<dbl2*>(F);
        for (int k = 0; k < 4; k++) Fv[(long)k * g.S + s] = xb[cur * kXfEntries + k * kCgBlock + threadIdx.x];
    }
    block_partials(sq, fs, mx, fm, part + 4 * (long)blockIdx.x);
}

// The deviation kernel walking z (P a multiple of 256): block b takes unit u (sites [256 u, 256 u + 256) of a plane)
// in the z chunk [z0, z0 + zc) and walks it plane by plane.  The backward z term at z is the adjoint of the forward z
// term at z - 1, A = U^G_z(z - 1), and the accumulators of A^dag are those of A with the eight anti-Hermitian ones
// negated and Re tr unchanged: the walk carries those ten numbers from one plane to the next instead of reading
// G(z - 1) and U_z(z - 1) and forming the link again, and G(z + 1), read for the forward term, is G(x) of the next
// plane.  That is three matrices read per site less than k_cg_dev_fused, among them the two whose planes lie furthest
// away.  With XCD order (U = P / 256 units a multiple of 8), XCD k (= b % 8) owns units [k U/8, (k+1) U/8) of every
// chunk, so the y neighbour units run on the same XCD.  Two waves per SIMD (at three the kernel spills and was slower
// than k_cg_dev_fused): at 256 x 256 x 1024 the Fourier iteration took 25.2 ms with chunks of 32 planes, 25.6 ms
// with 16, against 27.1 ms with k_cg_dev_fused.
struct CgWalk {
    int zc;   // planes per chunk
    int xcd;  // XCD order
};

template <bool CONS, bool XF>
__global__ __launch_bounds__(kCgBlock, 2) void k_cg_dev_walk(const double* __restrict__ U, const double* __restrict__ G,
                                                           const double* __restrict__ Gtop, double* __restrict__ F,
                                                           double* __restrict__ part, Geom g, double cx, double cy,
                                                           double cz, double fm0, CgWalk o, Radices R,
                                                           const dbl2* __restrict__ tw) {
    __shared__ dbl2 xb[XF ? 2 * kXfEntries : 1];
    const long nu = g.P / kCgBlock;
    long unit, zcid;
    if (o.xcd) {
        const long b = blockIdx.x, upx = nu / 8;
        const long i = b >> 3;
        zcid = i / upx;
        unit = (b & 7) * upx + (i - zcid * upx);
    } else {
        zcid = blockIdx.x / nu;
        unit = blockIdx.x - zcid * nu;
    }
    const int z0 = (int)zcid * o.zc, z1 = min(z0 + o.zc, g.nz);
    const long s2 = unit * kCgBlock + threadIdx.x;
    const int x = (int)(s2 % g.nx), y = (int)(s2 / g.nx);
    const double czf = g.nz > 1 ? cz : 0.0;
    double sq = 0.0, fs = 0.0, mx = 0.0, fm = fm0;
    M3 Gx;
    ldm<long>(G, kGPlanes, 0, s2 + (long)z0 * g.P, Gx);
    double pz[10];  // accumulators of the forward z term of the previous plane
#pragma unroll 1
    for (int z = z0; z < z1; z++) {
        const long s = s2 + (long)z * g.P;
        double q[10];
        for (int k = 0; k < 10; k++) q[k] = 0.0;
        w_terms<CONS>(U, G, Gtop, g, 0, s, x, y, z, cx, Gx, q);
        w_terms<CONS>(U, G, Gtop, g, 1, s, x, y, z, cy, Gx, q);
        // backward z: carried within the chunk (z > z0 >= 0, so the link exists), formed at its first plane
        if (z > z0) {
            for (int k = 0; k < 9; k++) q[k] -= pz[k];
            q[9] += pz[9];
        } else {
            const bool has_dn = z > 0 || (CONS && g.bc && g.nz > 1);
            if (has_dn) {
                M3 Gn, Um, T;
                const long sb = z > 0 ? s - g.P : s + (long)(g.nz - 1) * g.P;
                ldm<long>(G, kGPlanes, 0, sb, Gn);
                ldm<long>(U, kUPlanes, 36, sb, Um);
                mul_nn(Gn, Um, T);
                acc_proj_nd(czf, Gx, T, q);
            } else if (!CONS) {
                q[9] += 3.0 * czf;
            }
        }
        // forward z, last: G(z + 1) becomes G(x) of the next plane
        const bool has_up = !CONS || (g.nz > 1 && (g.bc || z + 1 < g.nz));
        if (has_up) {
            M3 Um, T, Gu;
            ldm<long>(U, kUPlanes, 36, s, Um);
            mul_nn(Gx, Um, T);
            if (z + 1 < g.nz) ldm<long>(G, kGPlanes, 0, s + g.P, Gu);
            else if (CONS) ldm<long>(G, kGPlanes, 0, s2, Gu);  // periodic: layer 0 now
            else ldm<long>(Gtop, kGPlanes, 0, s2, Gu);
            for (int k = 0; k < 10; k++) pz[k] = 0.0;
            acc_proj_nd(czf, T, Gu, pz);
            for (int k = 0; k < 10; k++) q[k] += pz[k];
            Gx = Gu;  // read again from G only where the next plane is another chunk's
        }
        double dv[8];
        delta_from_proj(q, dv);
        for (int a = 0; a < 8; a++) {
            mx = fmax(mx, fabs(dv[a]));
            sq += dv[a] * dv[a];
        }
        const double lf = q[9];
        fm = fmin(fm, lf);
        fs += lf;
        if (XF) {
            for (int k = 0; k < 4; k++) xb[k * kCgBlock + threadIdx.x] = dbl2{dv[2 * k], dv[2 * k + 1]};
            __syncthreads();
            const int cur = stockham(xb, kXfEntries, 0, R, kXfEntries / R.n, R.n, -1, tw);
            dbl2* const Fv = reinterpret_cast<dbl2*>(F);
            for (int k = 0; k < 4; k++) Fv[(long)k * g.S + s] = xb[cur * kXfEntries + k * kCgBlock + threadIdx.x];
            __syncthreads();  // the reads of this plane's result precede the next plane's writes
        } else {
            for (int k = 0; k < 4; k++) {
                F[2 * ((long)k * g.S + s)] = dv[2 * k];
                F[2 * ((long)k * g.S + s) + 1] = dv[2 * k + 1];
            }
        }
    }
    block_partials(sq, fs, mx, fm, part + 4 * (long)blockIdx.x);
}

// the walk where P is a multiple of 256 and chunks of 32 or 16 planes still give 2048 blocks (eight per CU), else
// k_cg_dev_fused (at 64 x 64 x 256, 1024 blocks of 16 planes, the walk took 0.62 ms per iteration against 0.56);
// GLASMA_CG_WALK = planes per chunk overrides, 0 turns the walk off
CgWalk cg_walk(const Geom& g) {
    const char* e = std::getenv("GLASMA_CG_WALK");  // read per call, so that tests can set it
    const int forced = e ? std::atoi(e) : -1;
    if (forced == 0 || g.P % kCgBlock) return {0, 0};
    const long nu = g.P / kCgBlock;
    int zc = forced > 0 ? forced : 0;
    if (!zc)
        for (int c : {32, 16})
            if (nu * ((g.nz + c - 1) / c) >= 2048) {
                zc = c;
                break;
            }
    if (!zc) return {0, 0};
    return {zc, nu % 8 == 0 ? 1 : 0};
}

// rows P and Q of X <- A X, A the identity outside the SU(2) block [[a, b], [c, d]] of rows / columns P, Q
template <int P, int Q>
__device__ __forceinline__ void left_su2(M3& X, double ar, double ai, double br, double bi, double cr, double ci,
                                         double dr, double di) {
    for (int j = 0; j < 3; j++) {
        const double pr = X.re[mi(P, j)], pi = X.im[mi(P, j)], qr = X.re[mi(Q, j)], qi = X.im[mi(Q, j)];
        X.re[mi(P, j)] = (ar * pr - ai * pi) + (br * qr - bi * qi);
        X.im[mi(P, j)] = (ar * pi + ai * pr) + (br * qi + bi * qr);
        X.re[mi(Q, j)] = (cr * pr - ci * pi) + (dr * qr - di * qi);
        X.im[mi(Q, j)] = (cr * pi + ci * pr) + (dr * qi + di * qr);
    }
}

// one Cabibbo-Marinari step in subgroup (P, Q) (CabbiboMarinariProjection.cpp:4-96): the SU(2) element that
// turns VNew's (P, Q) block into a positive multiple of the identity, applied to G and VNew
template <int P, int Q>
__device__ __forceinline__ void cm_step(M3& G, M3& V) {
    const double ar = V.re[mi(P, P)] + V.re[mi(Q, Q)], ai = -V.im[mi(P, P)] + V.im[mi(Q, Q)];  // conj(Vpp) + Vqq
    const double br = V.re[mi(Q, P)] - V.re[mi(P, Q)], bi = -V.im[mi(Q, P)] - V.im[mi(P, Q)];  // conj(Vqp) - Vpq
    const double cr = V.re[mi(P, Q)] - V.re[mi(Q, P)], ci = -V.im[mi(P, Q)] - V.im[mi(Q, P)];  // conj(Vpq) - Vqp
    const double dr = V.re[mi(Q, Q)] + V.re[mi(P, P)], di = -V.im[mi(Q, Q)] + V.im[mi(P, P)];  // conj(Vqq) + Vpp
    const double inv = 1.0 / sqrt((ar * ar + ai * ai) + (br * br + bi * bi));
    left_su2<P, Q>(G, inv * ar, inv * ai, inv * br, inv * bi, inv * cr, inv * ci, inv * dr, inv * di);
    left_su2<P, Q>(V, inv * ar, inv * ai, inv * br, inv * bi, inv * cr, inv * ci, inv * dr, inv * di);
}

// SUNcGroup::Extended::MaxTraceProjection (SUNc/SU3/GroupOperations.cpp:358-378): Vt = G^dag after seven sweeps
__device__ void max_trace_projection(const M3& V, M3& Vt) {
    M3 G, W = V;
    set_unit(G);
#pragma unroll 1
    for (int n = 0; n < 7; n++) {
        cm_step<0, 1>(G, W);
        cm_step<0, 2>(G, W);
        cm_step<1, 2>(G, W);
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            Vt.re[mi(i, j)] = G.re[mi(j, i)];
            Vt.im[mi(i, j)] = -G.im[mi(j, i)];
        }
}

// Los Alamos half sweep over the sites of one checkerboard parity (LosAlamos.cpp:13-67): monitors of Delta and
// of the functional into dpart, of the step |tr(1 - wTilde)| into spart; g <- wTilde^dag g
__global__ __launch_bounds__(kCgBlock) void k_cg_losalamos(const double* __restrict__ UG, double* __restrict__ G,
                                                            double* __restrict__ dpart, double* __restrict__ spart,
                                                            Geom g, double cx, double cy, double cz, double fm0,
                                                            int parity, CgOrder o) {
    const long s = cg_site(g, o);
    double sq = 0.0, fs = 0.0, mx = 0.0, fm = fm0, ssq = 0.0, sm = 0.0;
    if (s < g.S) {
        const int z = (int)(s / g.P);
        const long s2 = s - (long)z * g.P;
        const int x = (int)(s2 % g.nx), y = (int)(s2 / g.nx);
        if (((x + y + z) & 1) == parity) {
            M3 W, Wt, Gx, Gn;
            local_w(UG, g, s, x, y, z, cx, cy, cz, W);
            double dv[8];
            retr_igen(W, dv);
            for (int a = 0; a < 8; a++) {
                mx = fmax(mx, fabs(dv[a]));
                sq += dv[a] * dv[a];
            }
            const double lf = W.re[0] + W.re[4] + W.re[8];
            fm = fmin(fm, lf);
            fs = lf;
            max_trace_projection(W, Wt);
            const double st = hypot(3.0 - (Wt.re[0] + Wt.re[4] + Wt.re[8]), -(Wt.im[0] + Wt.im[4] + Wt.im[8]));
            sm = st;
            ssq = st * st;
            ldm<long>(G, kGPlanes, 0, s, Gx);
            mul_dn(Wt, Gx, Gn);
            stm<long>(G, kGPlanes, 0, s, Gn);
        }
    }
    block_partials(sq, fs, mx, fm, dpart + 4 * (long)blockIdx.x);
    __syncthreads();  // block_partials' shared scratch is reused
    block_partials(ssq, 0.0, sm, 0.0, spart + 4 * (long)blockIdx.x);
}

// Fourier acceleration: F *= 1/(V p^2) for p^2 > 0 (FourierAcceleration.cpp:55-75); the zero mode is passed
// through, or set to 0 with drop_zero (the consistent update)
__global__ void k_cg_scale(double* __restrict__ F, Geom g, double ax, double ay, double az, int drop_zero) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    const int pz = (int)(s / g.P);
    const long s2 = s - (long)pz * g.P;
    const int px = (int)(s2 % g.nx), py = (int)(s2 / g.nx);
    const double p2 = (2.0 - 2.0 * cos(2.0 * kPi * px / (double)g.nx)) / (ax * ax) +
                      (2.0 - 2.0 * cos(2.0 * kPi * py / (double)g.ny)) / (ay * ay) +
                      (2.0 - 2.0 * cos(2.0 * kPi * pz / (double)g.nz)) / (az * az);
    if (!(p2 > 0.0)) {
        if (drop_zero)
            for (int a = 0; a < 4; a++) {
                const long i = (long)a * g.S + s;
                F[2 * i] = 0.0;
                F[2 * i + 1] = 0.0;
            }
        return;
    }
    const double nf = 1.0 / (double)g.S;
    for (int a = 0; a < 4; a++) {  // the four packed arrays
        const long i = (long)a * g.S + s;
        F[2 * i] = nf * F[2 * i] / p2;
        F[2 * i + 1] = nf * F[2 * i + 1] / p2;
    }
}

// g <- exp(-i alpha Delta t)^dag g and the step size |3 - tr exp(...)| (FourierAcceleration.cpp:80-140).  XF: the
// arrays arrive without their inverse x transform, which the block runs on its lines first.
template <bool XF>
__global__ __launch_bounds__(kCgBlock) void k_cg_update(const double* __restrict__ F, double* __restrict__ G,
                                                         double* __restrict__ part, Geom g, double alpha, Radices R,
                                                         const dbl2* __restrict__ tw) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ dbl2 xb[XF ? 2 * kXfEntries : 1];
    double ssq = 0.0, sm = 0.0;
    double ub[8];
    if (XF) {  // every block is whole
        const dbl2* const Fv = reinterpret_cast<const dbl2*>(F);
        for (int k = 0; k < 4; k++) xb[k * kCgBlock + threadIdx.x] = Fv[(long)k * g.S + s];
        __syncthreads();
        const int cur = stockham(xb, kXfEntries, 0, R, kXfEntries / R.n, R.n, +1, tw);
        for (int k = 0; k < 4; k++) {
            const dbl2 v = xb[cur * kXfEntries + k * kCgBlock + threadIdx.x];
            ub[2 * k] = v.x;
            ub[2 * k + 1] = v.y;
        }
    }
    if (s < g.S) {
        if (!XF)
            for (int a = 0; a < 8; a++) ub[a] = F[2 * ((long)(a >> 1) * g.S + s) + (a & 1)];  // as k_cg_dev_fused packs
        M3 Wt, Gx, Gn;
        expi(-alpha, ub, Wt);
        const double tr = 3.0 - (Wt.re[0] + Wt.re[4] + Wt.re[8]);
        const double ti = -(Wt.im[0] + Wt.im[4] + Wt.im[8]);
        const double st = hypot(tr, ti);
        sm = st;
        ssq = st * st;
        ldm<long>(G, kGPlanes, 0, s, Gx);
        mul_dn(Wt, Gx, Gn);
        stm<long>(G, kGPlanes, 0, s, Gn);
    }
    block_partials(ssq, 0.0, sm, 0.0, part + 4 * (long)blockIdx.x);
}

// fixed-order reduction of block partials: block i takes partials [i per, (i + 1) per), strided per-thread sums then the
// block tree, into out + 4 i
__global__ __launch_bounds__(kCgBlock) void k_cg_reduce(const double* __restrict__ part, long nblk, long per,
                                                         double* __restrict__ out, double fm0) {
    double s0 = 0, s1 = 0, m0 = 0, m1 = fm0;
    const long end = min(nblk, (long)(blockIdx.x + 1) * per);
    for (long b = (long)blockIdx.x * per + threadIdx.x; b < end; b += blockDim.x) {
        s0 += part[4 * b];
        s1 += part[4 * b + 1];
        m0 = fmax(m0, part[4 * b + 2]);
        m1 = fmin(m1, part[4 * b + 3]);
    }
    block_partials(s0, s1, m0, m1, out + 4 * (long)blockIdx.x);
}

// the four monitors of nblk block partials into out[0..3]: one block up to 4096 partials, else up to 256 blocks of
// 1024 or more into tmp (4 * 256 doubles) and one block over those (a single block over 262144 partials took 0.25 ms)
constexpr long kReduceTmp = 4 * 256;
void reduce_partials(hipStream_t st, const double* part, long nblk, double* out, double fm0, double* tmp) {
    if (nblk <= 4096) {
        hipLaunchKernelGGL(k_cg_reduce, dim3(1), dim3(kCgBlock), 0, st, part, nblk, nblk, out, fm0);
        return;
    }
    const long nb2 = std::min(256L, (nblk + 1023) / 1024);
    const long per = (nblk + nb2 - 1) / nb2;
    hipLaunchKernelGGL(k_cg_reduce, dim3((unsigned)nb2), dim3(kCgBlock), 0, st, part, nblk, per, tmp, fm0);
    hipLaunchKernelGGL(k_cg_reduce, dim3(1), dim3(kCgBlock), 0, st, tmp, nb2, nb2, out, fm0);
}

// ---- Gauss-law restoration (GAUSSLAWFIXING/GaussRestore.cpp, GaussRestore2D.cpp) -------------------------
// violation G^a(x) = (1/a^3) (sum_mu E_mu(x) - U_mu^dag(x - mu) E_mu(x - mu) U_mu(x - mu) - (JpS + JmS)/sqrt2)
// (COMPUTE_GAUSS_VIOLATION, CALC/GaussViolation.cpp:13-23) into Gv (8 tiled planes), monitors into part
__global__ __launch_bounds__(kCgBlock) void k_gr_violation(const double* __restrict__ U, const double* __restrict__ E,
                                                            const double* __restrict__ J, double* __restrict__ Gv,
                                                            double* __restrict__ part, Geom g, double norm) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double sq = 0.0, mx = 0.0;
    if (s < g.S) {
        const int z = (int)(s / g.P);
        const long s2 = s - (long)z * g.P;
        const int x = (int)(s2 % g.nx), y = (int)(s2 / g.nx);
        const long sm[3] = {sidx(g, x - 1, y, z), sidx(g, x, y - 1, z), sidx(g, x, y, zdn(g, z))};
        double acc[8], e[8], d[8];
        for (int a = 0; a < 8; a++) acc[a] = 0.0;
        for (int mu = 0; mu < 3; mu++) {
            M3 V;
            ldm<long>(U, kUPlanes, 18 * mu, sm[mu], V);
            ldv<long>(E, kEPlanes, 8 * mu, sm[mu], e);
            inv_adj_mul(V, e, d);
            ldv<long>(E, kEPlanes, 8 * mu, s, e);
            for (int a = 0; a < 8; a++) acc[a] += e[a] - d[a];
        }
        double jp[8], jm[8], gv[8];
        ldv<long>(J, kJPlanes, JPS * 8, s, jp);
        ldv<long>(J, kJPlanes, JMS * 8, s, jm);
        for (int a = 0; a < 8; a++) {
            gv[a] = norm * (acc[a] - (jp[a] + jm[a]) / kSqrt2);
            mx = fmax(mx, fabs(gv[a]));
            sq += gv[a] * gv[a];
        }
        stv<long>(Gv, 8, 0, s, gv);
    }
    block_partials(sq, 0.0, mx, 0.0, part + 4 * (long)blockIdx.x);
}

// E_mu(x) -= gamma (G(x) - U_mu(x) G(x + mu) U_mu(x)^dag) for mu < nmu (GaussRestore.cpp:72-115); above the last
// layer G's ghost is a copy of that layer (GaussRestore2D.cpp:65-67)
__global__ __launch_bounds__(kCgBlock) void k_gr_efield(const double* __restrict__ U, double* __restrict__ E,
                                                         const double* __restrict__ Gv, Geom g, double gamma, int nmu) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    const int z = (int)(s / g.P);
    const long s2 = s - (long)z * g.P;
    const int x = (int)(s2 % g.nx), y = (int)(s2 / g.nx);
    const long sp[3] = {sidx(g, x + 1, y, z), sidx(g, x, y + 1, z), z + 1 < g.nz ? s + g.P : s};
    double gx[8], gn[8], up[8], e[8];
    ldv<long>(Gv, 8, 0, s, gx);
    for (int mu = 0; mu < nmu; mu++) {
        M3 V;
        ldm<long>(U, kUPlanes, 18 * mu, s, V);
        ldv<long>(Gv, 8, 0, sp[mu], gn);
        adj_mul(V, gn, up);
        ldv<long>(E, kEPlanes, 8 * mu, s, e);
        for (int a = 0; a < 8; a++) e[a] -= gamma * (gx[a] - up[a]);
        stv<long>(E, kEPlanes, 8 * mu, s, e);
    }
}

// G and its top ghost layer are published in the context together, after both allocations and the
// identity fills were issued, so a refused allocation leaves the context as it was and the call can be
// retried.
int ensure_gauge(glasma_ctx* ctx) {
    if (ctx->G && ctx->Gtop) return 0;
    const long Sp = pad_sites(ctx->g.S), Pp = pad_sites(ctx->g.P);
    double* G = nullptr;
    double* Gtop = nullptr;
    hipError_t e = hipMalloc(&G, (size_t)kGPlanes * Sp * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&Gtop, (size_t)kGPlanes * Pp * sizeof(double));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_cg_identity, grid1(ctx->g.S), dim3(kCgBlock), 0, ctx->stream, G, ctx->g.S);
        hipLaunchKernelGGL(k_cg_identity, grid1(ctx->g.P), dim3(kCgBlock), 0, ctx->stream, Gtop, ctx->g.P);
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        if (G) (void)hipFree(G);
        if (Gtop) (void)hipFree(Gtop);
        return hip_fail(ctx, e, "ensure_gauge");
    }
    ctx->G = G;
    ctx->Gtop = Gtop;
    return 0;
}

int ensure_gauge_fixed(glasma_ctx* ctx) {
    if (!ctx->UG) GL_HIP(ctx, hipMalloc(&ctx->UG, (size_t)kUPlanes * pad_sites(ctx->g.S) * sizeof(double)));
    if (!ctx->EG) GL_HIP(ctx, hipMalloc(&ctx->EG, (size_t)kEPlanes * pad_sites(ctx->g.S) * sizeof(double)));
    return 0;
}

// U^G from U and G (with_e: E^G from E as well), then G's top ghost layer synchronised
int transform_links(glasma_ctx* ctx, bool with_e = false) {
    const Geom& g = ctx->g;
    if (with_e)
        hipLaunchKernelGGL(k_cg_transform<true>, grid1(g.S), dim3(kCgBlock), 0, ctx->stream, ctx->U[ctx->cur], ctx->G,
                           ctx->Gtop, ctx->UG, ctx->E, ctx->EG, g, cg_order(g));
    else
        hipLaunchKernelGGL(k_cg_transform<false>, grid1(g.S), dim3(kCgBlock), 0, ctx->stream, ctx->U[ctx->cur], ctx->G,
                           ctx->Gtop, ctx->UG, nullptr, nullptr, g, cg_order(g));
    hipLaunchKernelGGL(k_cg_sync_top, grid1(g.P), dim3(kCgBlock), 0, ctx->stream, ctx->G, ctx->Gtop, g);
    GL_HIP(ctx, hipGetLastError());
    return 0;
}

}  // namespace
}  // namespace glasma

using namespace glasma;

namespace {

int coulomb_gauge(glasma_ctx* ctx, int algo, double tol, int maxit, double param, double* stats7) {
    if (!ctx) return GLASMA_ERR_ARG;
    if (maxit < 0 || !(tol >= 0.0)) return set_error(ctx, GLASMA_ERR_ARG, "invalid tolerance or iteration cap");
    if (algo < GLASMA_GAUGE_FOURIER || algo > GLASMA_GAUGE_FOURIER_CONSISTENT)
        return set_error(ctx, GLASMA_ERR_ARG, "unknown gauge-fixing algorithm");
    if (algo == GLASMA_GAUGE_LOS_ALAMOS && param != 0.0)
        return set_error(ctx, GLASMA_ERR_ARG,
                         "Los Alamos over-relaxation needs the reference's shared random stream; only probability 0 "
                         "is supported");
    int rc = ensure_gauge(ctx);
    if (!rc) rc = ensure_gauge_fixed(ctx);
    if (rc) return rc;
    const Geom& g = ctx->g;
    const long nblk = (g.S + kCgBlock - 1) / kCgBlock;
    const bool cons = algo == GLASMA_GAUGE_FOURIER_CONSISTENT;
    const CgOrder ord = cg_order(g);
    const CgWalk walk = algo == GLASMA_GAUGE_LOS_ALAMOS ? CgWalk{0, 0} : cg_walk(g);
    const long nwalk = walk.zc ? g.P / kCgBlock * ((g.nz + walk.zc - 1) / walk.zc) : 0;
    // the x transforms fused into the deviation and update kernels (the Fourier steps with Nz >= 2)
    Radices R{};
    const dbl2* tw = nullptr;
    const bool xf = (algo == GLASMA_GAUGE_FOURIER || cons) && g.nz >= 2 && g.S % kCgBlock == 0 &&
                    kCgBlock % g.nx == 0 && factorize(g.nx, R) == 0 && fft_twiddles(g.nx, &tw) == hipSuccess;
    (void)hipGetLastError();
    rc = ensure_red(ctx, 16 * (size_t)nblk + 8 + kReduceTmp);
    if (!rc && algo != GLASMA_GAUGE_LOS_ALAMOS) rc = ensure_fftbuf(ctx, 8 * (size_t)g.S);  // 4 complex per site
    if (rc) return rc;
    const double* a = ctx->p.a;
    // weights (1/a_mu)^2 of w and the functional's starting minimum 6 sum_mu 1/a_mu^2, rounded as the
    // reference writes them
    const double cx = (1.0 / a[0]) * (1.0 / a[0]), cy = (1.0 / a[1]) * (1.0 / a[1]), cz = (1.0 / a[2]) * (1.0 / a[2]);
    const double fm0 = 6.0 * (1.0 / (a[0] * a[0]) + 1.0 / (a[1] * a[1]) + 1.0 / (a[2] * a[2]));
    double* part = ctx->red;                // 2 * nblk blocks of 4 (Los Alamos: both half sweeps)
    double* spart = ctx->red + 8 * nblk;    // Los Alamos step monitors
    double* fin = ctx->red + 16 * nblk;
    double* tmp = fin + 8;
    double* F = ctx->fftbuf;
    // Los Alamos sweeps read U^G; the Fourier and Jacobi iterations form it on the fly (k_cg_dev_fused) and keep
    // the reference's lag of the top ghost layer by synchronising Gtop right after each deviation kernel, where
    // the reference's link transformation at the end of the previous iteration would have
    const bool fused = algo != GLASMA_GAUGE_LOS_ALAMOS;
    if (!fused && (rc = transform_links(ctx))) return rc;
    double gmax = 1.0, gavg = 0.0, smax = 0.0, savg = 0.0, fmin = 0.0, fsum = 0.0;
    int n = 0;
    const double V = (double)g.S;
    while (n < maxit && gmax > tol) {
        double h[8];
        if (algo == GLASMA_GAUGE_LOS_ALAMOS) {
            for (int parity = 0; parity < 2; parity++) {
                hipLaunchKernelGGL(k_cg_losalamos, dim3((unsigned)nblk), dim3(kCgBlock), 0, ctx->stream, ctx->UG, ctx->G,
                                   part + 4 * parity * nblk, spart + 4 * parity * nblk, g, cx, cy, cz, fm0, parity,
                                   ord);
                if ((rc = transform_links(ctx))) return rc;
            }
            reduce_partials(ctx->stream, part, 2 * nblk, fin, fm0, tmp);
            reduce_partials(ctx->stream, spart, 2 * nblk, fin + 4, 0.0, tmp);
        } else {
#define GLASMA_CG_DEV(C, X)                                                                                     \
    if (walk.zc)                                                                                                 \
        hipLaunchKernelGGL((k_cg_dev_walk<C, X>), dim3((unsigned)nwalk), dim3(kCgBlock), 0, ctx->stream,            \
                           ctx->U[ctx->cur], ctx->G, ctx->Gtop, F, part, g, cx, cy, cz, fm0, walk, R, tw);         \
    else                                                                                                         \
        hipLaunchKernelGGL((k_cg_dev_fused<C, X>), dim3((unsigned)nblk), dim3(kCgBlock), 0, ctx->stream,           \
                           ctx->U[ctx->cur], ctx->G, ctx->Gtop, F, part, g, cx, cy, cz, fm0, ord, R, tw)
            if (cons) {
                if (xf) {
                    GLASMA_CG_DEV(true, true);
                } else {
                    GLASMA_CG_DEV(true, false);
                }
            } else {
                if (xf) {
                    GLASMA_CG_DEV(false, true);
                } else {
                    GLASMA_CG_DEV(false, false);
                }
                hipLaunchKernelGGL(k_cg_sync_top, grid1(g.P), dim3(kCgBlock), 0, ctx->stream, ctx->G, ctx->Gtop, g);
            }
#undef GLASMA_CG_DEV
            reduce_partials(ctx->stream, part, walk.zc ? nwalk : nblk, fin, fm0, tmp);
            if (cons) {
                // the Laplacian of the links that exist (open z: the Neumann Laplacian, a cosine transform in z); the
                // zero mode (zero by construction) dropped
                if (g.nz >= 2) {
                    rc = fft_3d_poisson(ctx->stream, F, g, 4, a, true, g.bc == 0, !xf, !xf);
                } else {
                    rc = fft_3d(ctx->stream, F, g, 4, -1);
                    if (!rc) {
                        hipLaunchKernelGGL(k_cg_scale, grid1(g.S), dim3(kCgBlock), 0, ctx->stream, F, g, a[0], a[1],
                                           a[2], 1);
                        rc = fft_3d(ctx->stream, F, g, 4, +1);
                    }
                }
                if (rc) return set_error(ctx, rc, std::string("gauge fixing FFT failed: ") + fft_last_error());
            }
            if (algo == GLASMA_GAUGE_FOURIER) {
                if (g.nz >= 2) {
                    rc = fft_3d_poisson(ctx->stream, F, g, 4, a, false, false, !xf, !xf);
                } else {
                    rc = fft_3d(ctx->stream, F, g, 4, -1);
                    if (!rc) {
                        hipLaunchKernelGGL(k_cg_scale, grid1(g.S), dim3(kCgBlock), 0, ctx->stream, F, g, a[0], a[1],
                                           a[2], 0);
                        rc = fft_3d(ctx->stream, F, g, 4, +1);
                    }
                }
                if (rc) return set_error(ctx, rc, std::string("gauge fixing FFT failed: ") + fft_last_error());
            }
            if (xf)
                hipLaunchKernelGGL(k_cg_update<true>, dim3((unsigned)nblk), dim3(kCgBlock), 0, ctx->stream, F, ctx->G,
                                   part, g, param, R, tw);
            else
                hipLaunchKernelGGL(k_cg_update<false>, dim3((unsigned)nblk), dim3(kCgBlock), 0, ctx->stream, F, ctx->G,
                                   part, g, param, R, tw);
            reduce_partials(ctx->stream, part, nblk, fin + 4, 0.0, tmp);
            GL_HIP(ctx, hipGetLastError());
        }
        GL_HIP(ctx, hipMemcpyAsync(h, fin, 8 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        GL_HIP(ctx, hipStreamSynchronize(ctx->stream));
        gmax = h[2];
        gavg = std::sqrt(h[0] / (V * 8.0));
        fmin = h[3];
        fsum = h[1] / V;
        smax = h[6];
        savg = std::sqrt(h[4] / (V * 8.0));
        n++;
    }
    // the gauge-fixed links, transformed with the final G and (as after the reference's last transformation) the
    // top ghost layer synchronised to it
    if (fused) hipLaunchKernelGGL(k_cg_sync_top, grid1(g.P), dim3(kCgBlock), 0, ctx->stream, ctx->G, ctx->Gtop, g);
    if ((rc = transform_links(ctx, true))) return rc;
    GL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // the iteration cap reached above the tolerance: success as the reference's SetCoulombGauge returns, with a
    // message in glasma_last_error for callers that check (the reference's update cannot converge for Nz > 1)
    if (n >= maxit && gmax > tol && tol > 0.0) {
        char msg[200];
        std::snprintf(msg, sizeof(msg), "gauge fixing stopped at the iteration cap %d with max|Delta| = %.3e above the "
                      "tolerance %.3e", maxit, gmax, tol);
        ctx->err = msg;
    }
    if (stats7) {
        stats7[0] = n;
        stats7[1] = gmax;
        stats7[2] = gavg;
        stats7[3] = smax;
        stats7[4] = savg;
        stats7[5] = fmin;
        stats7[6] = fsum;
    }
    return 0;
}

}  // namespace

extern "C" int glasma_gauss_restore(glasma_ctx* ctx, int transverse_only, double tol, int maxit, double* stats3) {
    if (!ctx) return GLASMA_ERR_ARG;
    if (maxit < 0 || !(tol >= 0.0)) return set_error(ctx, GLASMA_ERR_ARG, "invalid tolerance or iteration cap");
    const Geom& g = ctx->g;
    const long nblk = (g.S + kCgBlock - 1) / kCgBlock;
    int rc = ensure_red(ctx, 4 * (size_t)nblk + 8 + kReduceTmp);
    if (!rc) rc = ensure_fftbuf(ctx, 8 * (size_t)pad_sites(g.S));
    if (rc) return rc;
    const double norm = 1.0 / (ctx->p.a[0] * ctx->p.a[1] * ctx->p.a[2]);
    const int nmu = transverse_only ? 2 : 3;
    double* Gv = ctx->fftbuf;
    double* part = ctx->red;
    double* fin = ctx->red + 4 * nblk;
    const double* U = ctx->U[ctx->cur];
    const double* J = ctx->J[ctx->cur];
    auto deviation = [&]() {
        hipLaunchKernelGGL(k_gr_violation, dim3((unsigned)nblk), dim3(kCgBlock), 0, ctx->stream, U, ctx->E, J, Gv, part,
                           g, norm);
        reduce_partials(ctx->stream, part, nblk, fin, 0.0, fin + 8);
    };
    double h[4];
    auto read = [&]() -> int {
        GL_HIP(ctx, hipMemcpyAsync(h, fin, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        GL_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return 0;
    };
    deviation();
    if ((rc = read())) return rc;
    int n = 0;
    const double gammas[3] = {7.0 / 24.0, 7.0 / 48.0, 7.0 / 72.0};
    while (n < maxit && h[2] > tol) {
        for (int k = 0; k < 3; k++) {
            deviation();
            hipLaunchKernelGGL(k_gr_efield, grid1(g.S), dim3(kCgBlock), 0, ctx->stream, U, ctx->E, Gv, g, gammas[k], nmu);
        }
        GL_HIP(ctx, hipGetLastError());
        if ((rc = read())) return rc;  // the violation of the third step, measured before its update
        n++;
    }
    deviation();
    if ((rc = read())) return rc;
    if (stats3) {
        stats3[0] = n;
        stats3[1] = h[2];
        stats3[2] = std::sqrt(h[0] / (8.0 * (double)g.S));
    }
    return 0;
}

extern "C" int glasma_coulomb_gauge(glasma_ctx* ctx, double tol, int maxit, double alpha, double* stats7) {
    return coulomb_gauge(ctx, GLASMA_GAUGE_FOURIER, tol, maxit, alpha, stats7);
}

extern "C" int glasma_coulomb_gauge_algorithm(glasma_ctx* ctx, int algorithm, double tol, int maxit, double param,
                                              double* stats7) {
    return coulomb_gauge(ctx, algorithm, tol, maxit, param, stats7);
}

extern "C" int glasma_gauge_reset(glasma_ctx* ctx) {
    if (!ctx) return GLASMA_ERR_ARG;
    int rc = ensure_gauge(ctx);
    if (rc) return rc;
    hipLaunchKernelGGL(k_cg_identity, grid1(ctx->g.S), dim3(kCgBlock), 0, ctx->stream, ctx->G, ctx->g.S);
    hipLaunchKernelGGL(k_cg_identity, grid1(ctx->g.P), dim3(kCgBlock), 0, ctx->stream, ctx->Gtop, ctx->g.P);
    GL_HIP(ctx, hipGetLastError());
    GL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

extern "C" int glasma_get_gauge_fixed(glasma_ctx* ctx, double* UG, double* EG, double* G) {
    if (!ctx) return GLASMA_ERR_ARG;
    if ((UG || EG) && !ctx->UG) return set_error(ctx, GLASMA_ERR_STATE, "no gauge-fixed fields yet");
    int rc = 0;
    if (UG && (rc = download_aos(ctx, ctx->UG, UG, kUPlanes))) return rc;
    if (EG && (rc = download_aos(ctx, ctx->EG, EG, kEPlanes))) return rc;
    if (G) {
        if ((rc = ensure_gauge(ctx))) return rc;
        if ((rc = download_aos(ctx, ctx->G, G, kGPlanes))) return rc;
    }
    return 0;
}

extern "C" int glasma_set_gauge_transformation(glasma_ctx* ctx, const double* G) {
    if (!ctx || !G) return GLASMA_ERR_ARG;
    int rc = ensure_gauge(ctx);
    if (rc) return rc;
    if ((rc = upload_aos(ctx, G, ctx->G, kGPlanes))) return rc;
    // the top ghost layer (z = Nz) as given
    const Geom& g = ctx->g;
    std::vector<double> top((size_t)pad_sites(g.P) * kGPlanes);
    const double* src = G + (size_t)(g.nz + 1) * g.P * 18;
    for (long s2 = 0; s2 < g.P; s2++)
        for (int p = 0; p < 18; p++) top[lidx<long>(kGPlanes, p, s2)] = src[18 * s2 + p];
    GL_HIP(ctx, hipMemcpy(ctx->Gtop, top.data(), top.size() * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}
