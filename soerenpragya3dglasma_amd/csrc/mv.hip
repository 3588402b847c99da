// MV-model initial conditions on the device, SU(3):
//   CollidingObjects::SetWilsonLines  SC/COLLIDINGOBJECTS/CollidingObjects.cpp:318-371
//     (SetChargeDensities :34-102, SolveLaplaceEquation :131-176, SetNewWilsonLines :179-226,
//      LongitudinalProfile :30-32)
//   InitialConditions::SingleSheet    SC/INITIALCONDITIONS/SUNCCURRENTS/RevSingleNucleus.cpp:7-210
//     with Dynamics::SUNc::EvolveRhoViaFourier SC/DYNAMICS/SUNcDynamics.cpp:275-350 and
//     CollidingObjects::SetWilsonLinesForFirstStep SC/COLLIDINGOBJECTS/CollidingObjects.cpp:377-411
//   the initial state of Simulation::Run  SC/Simulation.cpp:327-378 (collision) and of the left / right
//     single-nucleus drivers (SC/LeftNucleusSimulation.cpp:343-357, SC/RightNucleusSimulation.cpp:343-357).
//
// The reference walks z sequentially and solves one transverse slice at a time.  Here every transverse
// operation (charge, smoothing, Poisson solve) runs on all z slices at once as batched 2D FFTs; only the
// path-ordered Wilson-line product is sequential in z, one thread per transverse site.
//
// Work arrays (allocated for the call, freed at the end):
//   F     complex, F[(a*Nz + z)*P + y*Nx + x], a = colour 0..7 (the FFT layout)
//   rho   real,    rho[a*S + s]  (s = x + Nx*(y + Ny*z))
//   V     Wilson lines in the tiled matrix layout (18 planes)
#include <cmath>
#include <limits>
#include <vector>

#include "internal.h"
#include "adjoint.h"

namespace glasma {

namespace {

struct Wk {
    double* F = nullptr;
    double* rho = nullptr;
    double* V = nullptr;
    double* xi = nullptr;
    double* prof = nullptr;  // z profiles of a separable charge (mv_sheet): 4 nz doubles
};

__device__ __forceinline__ long fidx(const Geom& g, int a, long s) {
    // complex entry (a, s) of F; s = s2 + P*z so (a*Nz + z)*P + s2 = a*S + s
    return (long)a * g.S + s;
}

__device__ __forceinline__ double psqr2(const Geom& g, int px, int py, double ax, double ay) {
    return (2.0 - 2.0 * cos(2.0 * kPi * px / (double)g.nx)) / (ax * ax) +
           (2.0 - 2.0 * cos(2.0 * kPi * py / (double)g.ny)) / (ay * ay);
}

// SetChargeDensities: rho(z) = mult * sqrt(ax ay) * T(z) * xi with T the unit-area Gaussian in index units
__global__ void k_charge(double* __restrict__ F, const double* __restrict__ xi, Geom g, double mult_norm,
                         int center, double width) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    const int z = (int)(s / g.P);
    const long s2 = s - (long)z * g.P;
    const double u = (z - center) / width;
    const double prof = exp(-0.5 * u * u) / sqrt(2.0 * kPi * width * width);
    const double c = mult_norm * prof;
    for (int a = 0; a < 8; a++) {
        const long i = fidx(g, a, s);
        F[2 * i] = c * xi[a + 8 * s2];
        F[2 * i + 1] = 0.0;
    }
}

// momentum-space factors of one z slice batch.  mode 0: smoothing nf * exp(-p^2/2 Lambda^2) p^2/(p^2+m^2);
// mode 1: Poisson nf / p^2 with the zero mode set to 0.
__global__ void k_transverse_filter(double* __restrict__ F, Geom g, double ax, double ay, double lam, double meff,
                                   int mode) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    const long s2 = s % g.P;
    const int px = (int)(s2 % g.nx), py = (int)(s2 / g.nx);
    const double p2 = psqr2(g, px, py, ax, ay);
    const double nf = 1.0 / ((double)g.nx * g.ny);
    for (int a = 0; a < 8; a++) {
        const long i = fidx(g, a, s);
        double re = F[2 * i], im = F[2 * i + 1];
        if (mode == 0) {
            const double f = exp(-p2 / (2.0 * lam * lam));
            re = nf * re * f * p2 / (p2 + meff * meff);
            im = nf * im * f * p2 / (p2 + meff * meff);
        } else if (px == 0 && py == 0) {
            re = 0.0;
            im = 0.0;
        } else {
            re = nf * re / p2;
            im = nf * im / p2;
        }
        F[2 * i] = re;
        F[2 * i + 1] = im;
    }
}

// rho = Re F (optional), then F = (Re F, 0)
__global__ void k_take_real(double* __restrict__ F, double* __restrict__ rho, Geom g) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    for (int a = 0; a < 8; a++) {
        const long i = fidx(g, a, s);
        if (rho) rho[(long)a * g.S + s] = F[2 * i];
        F[2 * i + 1] = 0.0;
    }
}

// F = (rho, 0)
__global__ void k_load_real(double* __restrict__ F, const double* __restrict__ rho, Geom g) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    for (int a = 0; a < 8; a++) {
        const long i = fidx(g, a, s);
        F[2 * i] = rho[(long)a * g.S + s];
        F[2 * i + 1] = 0.0;
    }
}

// SetNewWilsonLines for all z: one thread per transverse site walks z in the reference's order
// (L: Nz-1 down to 0 with exp(-i A t / sqrt 2); R: 0 up to Nz-1 with exp(+i A t / sqrt 2)),
// W <- exp(+-i A(z) t / sqrt 2) W, V(z) = W.
__global__ void __launch_bounds__(256) k_wilson_scan(const double* __restrict__ F, double* __restrict__ V, Geom g, int right) {
    long s2 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s2 >= g.P) return;
    M3 W, X, T;
    set_unit(W);
    const double c = (right ? 1.0 : -1.0) / kSqrt2;
    for (int k = 0; k < g.nz; k++) {
        const int z = right ? k : g.nz - 1 - k;
        const long s = s2 + g.P * z;
        double col[8];
        for (int a = 0; a < 8; a++) col[a] = F[2 * fidx(g, a, s)];
        expi(c, col, X);
        mul_nn(X, W, T);
        W = T;
        stm<long>(V, kGPlanes, 0, s, W);
    }
}

// transverse links V^dag(x) V(x + mu) for mu = x, y and the unit z link, for z in [zlo, zhi]
__global__ void k_sheet_links(const double* __restrict__ V, double* __restrict__ U, Geom g, int zlo, int zhi) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)(zhi - zlo + 1) * g.P;
    if (i >= n) return;
    const long s = i + (long)zlo * g.P;
    const int z = (int)(s / g.P);
    const long s2 = s - (long)z * g.P;
    const int x = (int)(s2 % g.nx), y = (int)(s2 / g.nx);
    M3 A, B, C;
    ldm<long>(V, kGPlanes, 0, s, A);
    ldm<long>(V, kGPlanes, 0, sidx(g, x + 1, y, z), B);
    mul_dn(A, B, C);
    stm<long>(U, kUPlanes, 0, s, C);
    ldm<long>(V, kGPlanes, 0, sidx(g, x, y + 1, z), B);
    mul_dn(A, B, C);
    stm<long>(U, kUPlanes, 18, s, C);
    set_unit(C);
    stm<long>(U, kUPlanes, 36, s, C);
}

// EvolveRhoViaFourier phase: F(pz) *= nf * exp(i sgn omega(pz) dt), sgn by half-spectrum and direction
__global__ void k_evolve_phase(double* __restrict__ F, Geom g, double az, double dt, int right) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    const int pz = (int)(s / g.P);
    const double om = sqrt((2.0 - 2.0 * cos(2.0 * kPi * (double)pz / (double)g.nz)) / (az * az));
    double sgn;
    if (!right) sgn = (pz < g.nz / 2) ? -1.0 : 1.0;
    else sgn = (pz < g.nz / 2) ? 1.0 : -1.0;
    double sn, cs;
    sincos(sgn * om * dt, &sn, &cs);
    const double nf = 1.0 / g.nz;
    for (int a = 0; a < 8; a++) {
        const long i = fidx(g, a, s);
        const double re = nf * F[2 * i], im = nf * F[2 * i + 1];
        F[2 * i] = cs * re - sn * im;
        F[2 * i + 1] = sn * re + cs * im;
    }
}

// The momentum-space factor of mv_sheet over one transverse round trip: smoothing x Poisson,
// nf exp(-p^2/2 Lambda^2) p^2/(p^2+m^2) / p^2 with the zero mode set to 0, and with `evolve` the z phase of
// k_evolve_phase (1/Nz included).  F holds the transverse (and, with `evolve`, the z) transform of the charge.
__global__ void k_sheet_filter(double* __restrict__ F, Geom g, double ax, double ay, double az, double lam, double meff,
                               double dt, int right, int evolve) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    const int pz = (int)(s / g.P);
    const long s2 = s - (long)pz * g.P;
    const int px = (int)(s2 % g.nx), py = (int)(s2 / g.nx);
    const double p2 = psqr2(g, px, py, ax, ay);
    const double nf = 1.0 / ((double)g.nx * g.ny);
    const double fs = nf * exp(-p2 / (2.0 * lam * lam)) * p2 / (p2 + meff * meff);
    const double f = (px == 0 && py == 0) ? 0.0 : fs / p2;
    double cs = f, sn = 0.0;
    if (evolve) {
        const double om = sqrt((2.0 - 2.0 * cos(2.0 * kPi * (double)pz / (double)g.nz)) / (az * az));
        double sgn;
        if (!right) sgn = (pz < g.nz / 2) ? -1.0 : 1.0;
        else sgn = (pz < g.nz / 2) ? 1.0 : -1.0;
        sincos(sgn * om * dt, &sn, &cs);
        cs *= f / g.nz;
        sn *= f / g.nz;
    }
    for (int a = 0; a < 8; a++) {
        const long i = fidx(g, a, s);
        const double re = F[2 * i], im = F[2 * i + 1];
        F[2 * i] = cs * re - sn * im;
        F[2 * i + 1] = sn * re + cs * im;
    }
}

// one transverse slice of F (complex, component-major: entry (a, s2) at a P + s2) = (xi, 0)
__global__ void k_load_xi(double* __restrict__ F, const double* __restrict__ xi, long P) {
    const long s2 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s2 >= P) return;
    for (int a = 0; a < 8; a++) {
        F[2 * (a * P + s2)] = xi[a + 8 * s2];
        F[2 * (a * P + s2) + 1] = 0.0;
    }
}

// z profiles of the separable MV charge rho(x, y, z) = c(z) xi(x, y): prof[z] = c(z), k_charge's factor; then
// prof[2 nz + 2 pz (+1)] = sum_z c(z) e^{-2 pi i pz z / nz}, and prof[nz + z] = the real part of the inverse transform
// of that times k_evolve_phase's factor (the profile evolved by dt).  Direct sums (nz^2 terms), one thread per output.
__global__ void k_profile_c(double* __restrict__ prof, Geom g, double mult_norm, int center, double width) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= g.nz) return;
    const double u = (z - center) / width;
    prof[z] = mult_norm * (exp(-0.5 * u * u) / sqrt(2.0 * kPi * width * width));
}
__global__ void k_profile_dft(double* __restrict__ prof, Geom g) {
    const int pz = blockIdx.x * blockDim.x + threadIdx.x;
    if (pz >= g.nz) return;
    double re = 0.0, im = 0.0;
    for (int z = 0; z < g.nz; z++) {
        double sn, cs;
        sincos(-2.0 * kPi * (double)(((long)pz * z) % g.nz) / (double)g.nz, &sn, &cs);
        re = fma(prof[z], cs, re);
        im = fma(prof[z], sn, im);
    }
    prof[2 * g.nz + 2 * pz] = re;
    prof[2 * g.nz + 2 * pz + 1] = im;
}
__global__ void k_profile_evolve(double* __restrict__ prof, Geom g, double az, double dt, int right) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= g.nz) return;
    double acc = 0.0;
    for (int pz = 0; pz < g.nz; pz++) {
        const double om = sqrt((2.0 - 2.0 * cos(2.0 * kPi * (double)pz / (double)g.nz)) / (az * az));
        double sgn;
        if (!right) sgn = (pz < g.nz / 2) ? -1.0 : 1.0;
        else sgn = (pz < g.nz / 2) ? 1.0 : -1.0;
        double sp, cp, sn, cs;
        sincos(sgn * om * dt, &sp, &cp);
        sincos(2.0 * kPi * (double)(((long)pz * z) % g.nz) / (double)g.nz, &sn, &cs);
        const double re = prof[2 * g.nz + 2 * pz], im = prof[2 * g.nz + 2 * pz + 1];
        const double er = cp * re - sp * im, ei = sp * re + cp * im;  // evolved coefficient
        acc += er * cs - ei * sn;                                       // real part of it times e^{+i ...}
    }
    prof[g.nz + z] = acc / g.nz;
}

// The path-ordered Wilson lines of k_wilson_scan for a potential separable in z, A(x, y, z) = cz(z) a(x, y), with a
// in the real parts of one transverse slice of F (entry (a, s2) at a P + s2)
__global__ void __launch_bounds__(256) k_wilson_scan_sep(const double* __restrict__ F, const double* __restrict__ cz,
                                                         double* __restrict__ V, Geom g, int right) {
    const long s2 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s2 >= g.P) return;
    double a0[8];
    for (int a = 0; a < 8; a++) a0[a] = F[2 * (a * g.P + s2)];
    M3 W, X, T;
    set_unit(W);
    const double c = (right ? 1.0 : -1.0) / kSqrt2;
    for (int k = 0; k < g.nz; k++) {
        const int z = right ? k : g.nz - 1 - k;
        const double f = cz[z];
        double col[8];
        for (int a = 0; a < 8; a++) col[a] = f * a0[a];
        expi(c, col, X);
        mul_nn(X, W, T);
        W = T;
        stm<long>(V, kGPlanes, 0, s2 + g.P * z, W);
    }
}

// E_perp(t + dt/2) = MatrixILog(coff_mu, U_mu UOld_mu^dag), E_z = 0 for z in [zlo, zhi]  (:66-92)
__global__ void __launch_bounds__(256) k_sheet_efield(const double* __restrict__ U, const double* __restrict__ UO, double* __restrict__ E,
                               Geom g, int zlo, int zhi, double coff0, double coff1) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)(zhi - zlo + 1) * g.P;
    if (i >= n) return;
    const long s = i + (long)zlo * g.P;
    for (int mu = 0; mu < 2; mu++) {
        M3 A, B, C;
        ldm<long>(U, kUPlanes, 18 * mu, s, A);
        ldm<long>(UO, kUPlanes, 18 * mu, s, B);
        mul_nd(A, B, C);
        double e[8];
        ilog(mu == 0 ? coff0 : coff1, C, e);
        stv<long>(E, kEPlanes, 8 * mu, s, e);
    }
    double zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    stv<long>(E, kEPlanes, 16, s, zero);
}

// J_stat = sqrt2 * sum_{i=x,y} (E_i(x) - U_i^dag(x-i) E_i(x-i) U_i(x-i)) inside [zlo, zhi], 0 outside
// (:98-125); also F = (J_stat, 0) for the continuity inversion.
__global__ void __launch_bounds__(256) k_sheet_jstat(const double* __restrict__ U, const double* __restrict__ E, double* __restrict__ J,
                              double* __restrict__ F, Geom g, int zlo, int zhi, int wstat) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    const int z = (int)(s / g.P);
    const long s2 = s - (long)z * g.P;
    const int x = (int)(s2 % g.nx), y = (int)(s2 / g.nx);
    double js[8];
    if (z >= zlo && z <= zhi) {
        const long sx = sidx(g, x - 1, y, z);
        const long sy = sidx(g, x, y - 1, z);
        M3 Ux, Uy;
        double ex[8], ey[8], exm[8], eym[8], xd[8], yd[8];
        ldm<long>(U, kUPlanes, 0, sx, Ux);
        ldm<long>(U, kUPlanes, 18, sy, Uy);
        ldv<long>(E, kEPlanes, 0, s, ex);
        ldv<long>(E, kEPlanes, 8, s, ey);
        ldv<long>(E, kEPlanes, 0, sx, exm);
        ldv<long>(E, kEPlanes, 8, sy, eym);
        inv_adj_mul(Ux, exm, xd);
        inv_adj_mul(Uy, eym, yd);
        for (int a = 0; a < 8; a++) js[a] = kSqrt2 * ((ex[a] - xd[a]) + (ey[a] - yd[a]));
    } else {
        for (int a = 0; a < 8; a++) js[a] = 0.0;
    }
    stv<long>(J, kJPlanes, 8 * wstat, s, js);
    for (int a = 0; a < 8; a++) {
        const long i = fidx(g, a, s);
        F[2 * i] = js[a];
        F[2 * i + 1] = 0.0;
    }
}

// continuity inversion in z momentum space (:131-209):
//   J~(0) = J~stat(0); J~(N/2) = 0; otherwise J~ = J~stat (e^{-+ i omega dt} - 1) / (dt pZD)
// (conj = 1: the complex-conjugate filter, i.e. the transpose of the real-linear map)
__global__ void k_jinv(double* __restrict__ F, Geom g, double az, double dt, int right, int conj) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    const int pz = (int)(s / g.P);
    const int nz = g.nz;
    const double om = sqrt((2.0 - 2.0 * cos(2.0 * kPi * (double)pz / (double)nz)) / (az * az));
    double es, ec;
    sincos(-2.0 * kPi * (double)pz / (double)nz, &es, &ec);
    // pZD = +-(1 - e)/az
    double zr = (1.0 - ec) / az, zi = -es / az;
    double sp, cp, sm, cm;
    sincos(om * dt, &sp, &cp);
    sincos(-om * dt, &sm, &cm);
    double pr, pi_, nr, ni;  // pNpos, pNneg
    if (right) {
        pr = cp - 1.0; pi_ = sp; nr = cm - 1.0; ni = sm;
    } else {
        zr = -zr; zi = -zi;
        pr = cm - 1.0; pi_ = sm; nr = cp - 1.0; ni = sp;
    }
    if (conj) {
        zi = -zi;
        pi_ = -pi_;
        ni = -ni;
    }
    const double dr = dt * zr, di = dt * zi;  // dt * pZD
    const double den = dr * dr + di * di;
    const double nf = 1.0 / nz;
    for (int a = 0; a < 8; a++) {
        const long i = fidx(g, a, s);
        const double jr = nf * F[2 * i], ji = nf * F[2 * i + 1];
        double outr, outi;
        if (pz == 0) {
            outr = jr; outi = ji;
        } else if (pz == nz / 2) {
            outr = 0.0; outi = 0.0;
        } else {
            const double mr = (pz < nz / 2) ? pr : nr, mi = (pz < nz / 2) ? pi_ : ni;
            // (j * m) / d
            const double qr = jr * mr - ji * mi, qi = jr * mi + ji * mr;
            outr = (qr * dr + qi * di) / den;
            outi = (qi * dr - qr * di) / den;
        }
        F[2 * i] = outr;
        F[2 * i + 1] = outi;
    }
}

// J = Re F into planes 8*w of the current-density field
__global__ void k_store_j(const double* __restrict__ F, double* __restrict__ J, Geom g, int w) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    double v[8];
    for (int a = 0; a < 8; a++) v[a] = F[2 * fidx(g, a, s)];
    stv<long>(J, kJPlanes, 8 * w, s, v);
}

// rho (real, a*S + s) -> host AoS with ghosts (a + 8*Index3D)
__global__ void k_rho_to_aos(const double* __restrict__ rho, double* __restrict__ out, Geom g) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 8 * g.S) return;
    const long s = i / 8;
    const int a = (int)(i % 8);
    out[i] = rho[(long)a * g.S + s];
}

inline dim3 grid1(long n) { return dim3((unsigned)((n + 255) / 256)); }

int alloc_work(glasma_ctx* ctx, Wk& w) {
    const long S = ctx->g.S, Sp = pad_sites(S);
    int rc = ensure_fftbuf(ctx, 16 * (size_t)S);
    if (rc) return rc;
    w.F = ctx->fftbuf;
    GL_HIP(ctx, hipMalloc(&w.rho, 8 * (size_t)S * sizeof(double)));
    GL_HIP(ctx, hipMalloc(&w.V, (size_t)kGPlanes * Sp * sizeof(double)));
    GL_HIP(ctx, hipMalloc(&w.xi, 8 * (size_t)ctx->g.P * sizeof(double)));
    GL_HIP(ctx, hipMalloc(&w.prof, 4 * (size_t)ctx->g.nz * sizeof(double)));
    return 0;
}
void free_work(Wk& w) {
    if (w.rho) (void)hipFree(w.rho);
    if (w.V) (void)hipFree(w.V);
    if (w.xi) (void)hipFree(w.xi);
    if (w.prof) (void)hipFree(w.prof);
    w = Wk();
}

#define GL_RC(x)                  \
    do {                          \
        int _rc = (x);            \
        if (_rc) return _rc;      \
    } while (0)
#define GL_FFT(ctx, x)                                                                       \
    do {                                                                                     \
        int _rc = (x);                                                                       \
        if (_rc) return set_error(ctx, _rc, std::string("FFT failed (lengths must be 2^a 3^b p <= 2048, " \
                                                        "prime factors <= 13): ") + fft_last_error()); \
    } while (0)

// Poisson solve of every transverse slice: F (real parts) -> A, imaginary parts of the result dropped
int poisson_all(glasma_ctx* ctx, double* F) {
    const Geom& g = ctx->g;
    GL_FFT(ctx, fft_xy(ctx->stream, F, g, 8, g.nz, -1));
    hipLaunchKernelGGL(k_transverse_filter, grid1(g.S), dim3(256), 0, ctx->stream, F, g, ctx->p.a[0], ctx->p.a[1], 0.0,
                       0.0, 1);
    GL_FFT(ctx, fft_xy(ctx->stream, F, g, 8, g.nz, +1));
    return 0;
}

// the per-z part of CollidingObjects::SetWilsonLines for the charge held in F (real parts): smoothing
// (UV factor exp(-p^2 / 2 Lambda^2), 1 for Lambda = infinity; IR factor p^2 / (p^2 + m^2)), w.rho = the
// smoothed charge, Poisson solve and the path-ordered Wilson lines w.V
int wilson_from_charge(glasma_ctx* ctx, Wk& w, double lambda, double meff, bool right) {
    const Geom& g = ctx->g;
    GL_FFT(ctx, fft_xy(ctx->stream, w.F, g, 8, g.nz, -1));
    hipLaunchKernelGGL(k_transverse_filter, grid1(g.S), dim3(256), 0, ctx->stream, w.F, g, ctx->p.a[0], ctx->p.a[1],
                       lambda, meff, 0);
    GL_FFT(ctx, fft_xy(ctx->stream, w.F, g, 8, g.nz, +1));
    hipLaunchKernelGGL(k_take_real, grid1(g.S), dim3(256), 0, ctx->stream, w.F, w.rho, g);
    GL_RC(poisson_all(ctx, w.F));
    hipLaunchKernelGGL(k_wilson_scan, grid1(g.P), dim3(256), 0, ctx->stream, w.F, w.V, g, right ? 1 : 0);
    GL_HIP(ctx, hipGetLastError());
    return 0;
}

// CollidingObjects::SetWilsonLines for one nucleus: w.rho = smoothed charge, w.V = Wilson lines
int set_wilson_lines(glasma_ctx* ctx, Wk& w, const glasma_mv_params* mv, const double* xi_host, bool right) {
    const Geom& g = ctx->g;
    GL_HIP(ctx, hipMemcpyAsync(w.xi, xi_host, 8 * (size_t)g.P * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    const double norm = std::sqrt(ctx->p.a[0] * ctx->p.a[1]);
    const double mult = right ? mv->nucleus_r : mv->nucleus_l;
    hipLaunchKernelGGL(k_charge, grid1(g.S), dim3(256), 0, ctx->stream, w.F, w.xi, g, mult * norm,
                       right ? mv->center_r : mv->center_l, right ? mv->sigma_r : mv->sigma_l);
    return wilson_from_charge(ctx, w, mv->lambda, mv->m_eff, right);
}

// InitialConditions::SingleSheet for one nucleus over [zlo, zhi]; J-stat into plane group wstat, J into w
int single_sheet(glasma_ctx* ctx, Wk& w, bool right, int zlo, int zhi, int wstat, int wj) {
    const Geom& g = ctx->g;
    const double dt = ctx->p.dt;
    double* U = ctx->U[ctx->cur];
    double* J = ctx->J[ctx->cur];
    const long nsl = (long)(zhi - zlo + 1) * g.P;
    // UOld = links of the sheet at t = 0 (:14-27)
    hipLaunchKernelGGL(k_sheet_links, grid1(nsl), dim3(256), 0, ctx->stream, w.V, ctx->UOld, g, zlo, zhi);
    // rho evolved by dt through the z FFT (:33, SUNcDynamics.cpp:275-350)
    hipLaunchKernelGGL(k_load_real, grid1(g.S), dim3(256), 0, ctx->stream, w.F, w.rho, g);
    GL_FFT(ctx, fft_z(ctx->stream, w.F, g, 8, -1));
    hipLaunchKernelGGL(k_evolve_phase, grid1(g.S), dim3(256), 0, ctx->stream, w.F, g, ctx->p.a[2], dt, right ? 1 : 0);
    GL_FFT(ctx, fft_z(ctx->stream, w.F, g, 8, +1));
    hipLaunchKernelGGL(k_take_real, grid1(g.S), dim3(256), 0, ctx->stream, w.F, (double*)nullptr, g);
    // Wilson lines of the evolved charge (SetWilsonLinesForFirstStep, no smoothing) and U(t + dt) (:40-60)
    GL_RC(poisson_all(ctx, w.F));
    hipLaunchKernelGGL(k_wilson_scan, grid1(g.P), dim3(256), 0, ctx->stream, w.F, w.V, g, right ? 1 : 0);
    hipLaunchKernelGGL(k_sheet_links, grid1(nsl), dim3(256), 0, ctx->stream, w.V, U, g, zlo, zhi);
    // E(t + dt/2) from the logarithm of U UOld^dag (:66-92)
    const double acube = ctx->p.a[0] * ctx->p.a[1] * ctx->p.a[2];
    const double coff0 = -dt * ctx->p.a[0] * ctx->p.a[0] / acube;
    const double coff1 = -dt * ctx->p.a[1] * ctx->p.a[1] / acube;
    hipLaunchKernelGGL(k_sheet_efield, grid1(nsl), dim3(256), 0, ctx->stream, U, ctx->UOld, ctx->E, g, zlo, zhi, coff0,
                       coff1);
    // J-stat from the transverse Gauss law and J from the z continuity inversion (:98-209)
    hipLaunchKernelGGL(k_sheet_jstat, grid1(g.S), dim3(256), 0, ctx->stream, U, ctx->E, J, w.F, g, zlo, zhi, wstat);
    GL_FFT(ctx, fft_z(ctx->stream, w.F, g, 8, -1));
    hipLaunchKernelGGL(k_jinv, grid1(g.S), dim3(256), 0, ctx->stream, w.F, g, ctx->p.a[2], dt, right ? 1 : 0, 0);
    GL_FFT(ctx, fft_z(ctx->stream, w.F, g, 8, +1));
    hipLaunchKernelGGL(k_store_j, grid1(g.S), dim3(256), 0, ctx->stream, w.F, J, g, wj);
    GL_HIP(ctx, hipGetLastError());
    return 0;
}

// CollidingObjects::SetWilsonLines and InitialConditions::SingleSheet for one MV nucleus, using that its charge is
// separable, rho(x, y, z) = c(z) xi(x, y) (SetChargeDensities, k_charge).  The reference smooths the charge (a
// transverse round trip), solves Poisson (another), and for U(t + dt) evolves the smoothed charge along z (a z round
// trip) and solves Poisson again, taking real parts in between.  Those filters are linear and each acts on one set of
// axes, so both potentials factor: A0 = c(z) a(x, y) and A1 = c_ev(z) a(x, y), with a = the smoothing x Poisson of xi
// on one transverse slice (k_sheet_filter) and c_ev = the real part of c evolved by dt (k_profile_*; the filters are
// real and even in p, so the real parts taken in between are exact up to round-off).  The Wilson lines are scanned
// from these factors (k_wilson_scan_sep): no transform of the whole lattice.  m_eff = 0 keeps the separate steps
// (their smoothing factor is 0/0 at p = 0, which the reference's order spreads over the slice).
int mv_sheet(glasma_ctx* ctx, Wk& w, const glasma_mv_params* mv, const double* xi_host, bool right, int zlo, int zhi,
             int wstat, int wj) {
    const Geom& g = ctx->g;
    const double dt = ctx->p.dt;
    const double* a = ctx->p.a;
    GL_HIP(ctx, hipMemcpyAsync(w.xi, xi_host, 8 * (size_t)g.P * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    const double norm = std::sqrt(a[0] * a[1]);
    const double mult = right ? mv->nucleus_r : mv->nucleus_l;
    const int center = right ? mv->center_r : mv->center_l;
    const double sigma = right ? mv->sigma_r : mv->sigma_l;
    const long nsl = (long)(zhi - zlo + 1) * g.P;
    double* U = ctx->U[ctx->cur];
    double* J = ctx->J[ctx->cur];
    // the transverse factor a(x, y) in the first slice of F
    Geom g1 = g;
    g1.nz = 1;
    g1.S = g.P;
    hipLaunchKernelGGL(k_load_xi, grid1(g.P), dim3(256), 0, ctx->stream, w.F, w.xi, g.P);
    GL_FFT(ctx, fft_xy(ctx->stream, w.F, g, 8, 1, -1));
    hipLaunchKernelGGL(k_sheet_filter, grid1(g.P), dim3(256), 0, ctx->stream, w.F, g1, a[0], a[1], a[2], mv->lambda,
                       mv->m_eff, dt, right ? 1 : 0, 0);
    GL_FFT(ctx, fft_xy(ctx->stream, w.F, g, 8, 1, +1));
    // the z profiles c(z) and c_ev(z)
    hipLaunchKernelGGL(k_profile_c, grid1(g.nz), dim3(256), 0, ctx->stream, w.prof, g, mult * norm, center, sigma);
    hipLaunchKernelGGL(k_profile_dft, grid1(g.nz), dim3(256), 0, ctx->stream, w.prof, g);
    hipLaunchKernelGGL(k_profile_evolve, grid1(g.nz), dim3(256), 0, ctx->stream, w.prof, g, a[2], dt, right ? 1 : 0);
    // U(t + dt) from A1 (SetWilsonLinesForFirstStep, :33-60), UOld from A0 (:14-27)
    hipLaunchKernelGGL(k_wilson_scan_sep, grid1(g.P), dim3(256), 0, ctx->stream, w.F, w.prof + g.nz, w.V, g,
                       right ? 1 : 0);
    hipLaunchKernelGGL(k_sheet_links, grid1(nsl), dim3(256), 0, ctx->stream, w.V, U, g, zlo, zhi);
    hipLaunchKernelGGL(k_wilson_scan_sep, grid1(g.P), dim3(256), 0, ctx->stream, w.F, w.prof, w.V, g, right ? 1 : 0);
    hipLaunchKernelGGL(k_sheet_links, grid1(nsl), dim3(256), 0, ctx->stream, w.V, ctx->UOld, g, zlo, zhi);
    // E(t + dt/2), J-stat and J as single_sheet (:66-209)
    const double acube = a[0] * a[1] * a[2];
    const double coff0 = -dt * a[0] * a[0] / acube;
    const double coff1 = -dt * a[1] * a[1] / acube;
    hipLaunchKernelGGL(k_sheet_efield, grid1(nsl), dim3(256), 0, ctx->stream, U, ctx->UOld, ctx->E, g, zlo, zhi, coff0,
                       coff1);
    hipLaunchKernelGGL(k_sheet_jstat, grid1(g.S), dim3(256), 0, ctx->stream, U, ctx->E, J, w.F, g, zlo, zhi, wstat);
    GL_FFT(ctx, fft_z(ctx->stream, w.F, g, 8, -1));
    hipLaunchKernelGGL(k_jinv, grid1(g.S), dim3(256), 0, ctx->stream, w.F, g, a[2], dt, right ? 1 : 0, 0);
    GL_FFT(ctx, fft_z(ctx->stream, w.F, g, 8, +1));
    hipLaunchKernelGGL(k_store_j, grid1(g.S), dim3(256), 0, ctx->stream, w.F, J, g, wj);
    GL_HIP(ctx, hipGetLastError());
    return 0;
}

// ------------------------------------------------------------------------------------------------
// The left-moving sheet from a given charge density, with its intermediates kept, and the reverse-mode
// derivative of that map (cotangents of U, E, J+, J+stat at t = 0 -> cotangent of the charge density).
// Forward: rho_s = smooth(rho); A0 = poisson(rho_s); V0 = scan(A0); UOld = links(V0);
// rho_ev = evolve_L(rho_s); A1 = poisson(rho_ev); V1 = scan(A1); U = links(V1); E = ilog(U UOld^dag);
// J+stat = sqrt2 sum_i (E_i - Ad(U_i^dag(x-i)) E_i(x-i)); J+ = jinv_L(J+stat).
// The smoothing and Poisson filters are real and even in p (self-adjoint maps); the z-evolution and the
// current inversion transpose to their complex-conjugate filters (evolve_L^T = evolve_R).
// ------------------------------------------------------------------------------------------------
// F = (values of tiled planes p0..p0+7 of a field with NP planes, 0)
__global__ void k_load_planes(double* __restrict__ F, const double* __restrict__ X, Geom g, int NP, int p0) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    for (int a = 0; a < 8; a++) {
        const long i = fidx(g, a, s);
        F[2 * i] = X[lidx<long>(NP, p0 + a, s)];
        F[2 * i + 1] = 0.0;
    }
}
// out[a*S + s] = x[a*S + s] + c * Re F
__global__ void k_axpy_real(double* __restrict__ out, const double* __restrict__ x, const double* __restrict__ F,
                            double c, Geom g) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 8 * g.S) return;
    out[i] = (x ? x[i] : 0.0) + c * F[2 * i];
}
__global__ void k_copy_real(double* __restrict__ out, const double* __restrict__ F, Geom g) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 8 * g.S) return;
    out[i] = F[2 * i];
}
__global__ void k_zero(double* __restrict__ p, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// J+stat cotangent jsb (real, a*S + s) -> E_x, E_y and U_x, U_y cotangents (gathered at y from y + i)
__global__ void __launch_bounds__(256) k_ic_jstat_vjp(const double* __restrict__ U, const double* __restrict__ E,
                               const double* __restrict__ jsb, double* __restrict__ Ub, double* __restrict__ Eb,
                               Geom g) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    const int z = (int)(s / g.P);
    const long s2 = s - (long)z * g.P;
    const int x = (int)(s2 % g.nx), y = (int)(s2 / g.nx);
    double own[8];
    for (int a = 0; a < 8; a++) own[a] = jsb[(long)a * g.S + s];
    for (int mu = 0; mu < 2; mu++) {
        const long sn = mu == 0 ? sidx(g, x + 1, y, z) : sidx(g, x, y + 1, z);
        double up[8], e[8], eb[8], t8[8];
        for (int a = 0; a < 8; a++) up[a] = jsb[(long)a * g.S + sn];
        M3 Um, Ubm;
        ldm<long>(U, kUPlanes, 18 * mu, s, Um);
        ldm<long>(Ub, kUPlanes, 18 * mu, s, Ubm);
        ldv<long>(E, kEPlanes, 8 * mu, s, e);
        ldv<long>(Eb, kEPlanes, 8 * mu, s, eb);
        adj_mul(Um, up, t8);
        for (int a = 0; a < 8; a++) eb[a] += kSqrt2 * own[a] - kSqrt2 * t8[a];
        acc_lvr(-kSqrt2, e, Um, up, Ubm);
        stv<long>(Eb, kEPlanes, 8 * mu, s, eb);
        stm<long>(Ub, kUPlanes, 18 * mu, s, Ubm);
    }
}

// 8x8 linear solve by Gaussian elimination with partial pivoting (A is overwritten); returns x in b
__device__ void solve8(double (&A)[8][8], double* b) {
    for (int c = 0; c < 8; c++) {
        int p = c;
        for (int r = c + 1; r < 8; r++)
            if (fabs(A[r][c]) > fabs(A[p][c])) p = r;
        if (p != c) {
            for (int k = 0; k < 8; k++) { const double t = A[c][k]; A[c][k] = A[p][k]; A[p][k] = t; }
            const double t = b[c]; b[c] = b[p]; b[p] = t;
        }
        const double inv = 1.0 / A[c][c];
        for (int r = c + 1; r < 8; r++) {
            const double f = A[r][c] * inv;
            for (int k = c; k < 8; k++) A[r][k] -= f * A[c][k];
            b[r] -= f * b[c];
        }
    }
    for (int c = 7; c >= 0; c--) {
        double v = b[c];
        for (int k = c + 1; k < 8; k++) v -= A[c][k] * b[k];
        b[c] = v / A[c][c];
    }
}

// E_mu = ilog(coff, M), M = U_mu UOld_mu^dag (mu = x, y): with A = i coff E t and the left-trivialised
// derivative G = sum_k (-ad_A)^k/(k+1)!, M^dag dM = G(i coff t dE); in components kappa = -2 Re tr(i t M^dag dM)
// = Ghat coff dE, so kappabar = Ghat^-T Ebar / coff and Mbar = 2 i M H(kappabar).  Ubar += Mbar UOld,
// UOldbar = Mbar^dag U.
__global__ void __launch_bounds__(256) k_ic_ilog_vjp(const double* __restrict__ U, const double* __restrict__ UO,
                              const double* __restrict__ E, const double* __restrict__ Eb, double* __restrict__ Ub,
                              double* __restrict__ UOb, Geom g, double coff0, double coff1) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    for (int mu = 0; mu < 2; mu++) {
        const double coff = mu == 0 ? coff0 : coff1;
        M3 Um, Uo, M;
        ldm<long>(U, kUPlanes, 18 * mu, s, Um);
        ldm<long>(UO, kUPlanes, 18 * mu, s, Uo);
        mul_nd(Um, Uo, M);
        double e[8], eb[8];
        ldv<long>(E, kEPlanes, 8 * mu, s, e);
        ldv<long>(Eb, kEPlanes, 8 * mu, s, eb);
        // row b = column b of Ghat: the coordinates (basis i t^a) of sum_k (-ad_A)^k (i t^b) / (k+1)!, A = i coff e^a t^a,
        // so -ad_A is ad_series' map with c = -coff
        double G[8][8];
        for (int b = 0; b < 8; b++) {
            double unit[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            unit[b] = 1.0;
            ad_series(-coff, e, unit, G[b]);
        }
        double kb[8];
        for (int a = 0; a < 8; a++) kb[a] = eb[a] / coff;
        solve8(G, kb);
        M3 Hk, Mb, T;
        herm_from_alg(1.0, kb, Hk);
        mul_nn(M, Hk, T);
        for (int k = 0; k < 9; k++) { Mb.re[k] = -2.0 * T.im[k]; Mb.im[k] = 2.0 * T.re[k]; }  // 2 i M H
        M3 Ubm, Ta;
        ldm<long>(Ub, kUPlanes, 18 * mu, s, Ubm);
        mul_nn(Mb, Uo, Ta);
        m3_axpy(1.0, Ta, Ubm);
        stm<long>(Ub, kUPlanes, 18 * mu, s, Ubm);
        mul_dn(Mb, Um, Ta);
        stm<long>(UOb, kUPlanes, 18 * mu, s, Ta);
    }
}

// V cotangent from the link cotangents of U_mu = V^dag(x) V(x + mu), mu = x, y (gathered at y)
__global__ void __launch_bounds__(256) k_links_vjp(const double* __restrict__ V, const double* __restrict__ Ub, double* __restrict__ Vb,
                            Geom g) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    const int z = (int)(s / g.P);
    const long s2 = s - (long)z * g.P;
    const int x = (int)(s2 % g.nx), y = (int)(s2 / g.nx);
    M3 acc, Vn, Bm, T;
    m3_zero(acc);
    for (int mu = 0; mu < 2; mu++) {
        const long sp = mu == 0 ? sidx(g, x + 1, y, z) : sidx(g, x, y + 1, z);
        const long sm = mu == 0 ? sidx(g, x - 1, y, z) : sidx(g, x, y - 1, z);
        ldm<long>(V, kGPlanes, 0, sp, Vn);
        ldm<long>(Ub, kUPlanes, 18 * mu, s, Bm);
        mul_nd(Vn, Bm, T);  // V(y+mu) Ubar_mu(y)^dag
        m3_axpy(1.0, T, acc);
        ldm<long>(V, kGPlanes, 0, sm, Vn);
        ldm<long>(Ub, kUPlanes, 18 * mu, sm, Bm);
        mul_nn(Vn, Bm, T);  // V(y-mu) Ubar_mu(y-mu)
        m3_axpy(1.0, T, acc);
    }
    stm<long>(Vb, kGPlanes, 0, s, acc);
}

// Reverse of the left-moving Wilson-line scan V(z) = X(z) V(z+1), X(z) = exp(-i A(z) t / sqrt2), in two kernels.
// The total cotangent of V(z) is its own Vbar(z) plus, through V(z-1) = X(z-1) V(z), X(z-1)^dag times that of V(z-1);
// with X(z') ... X(z-1) = V(z') V(z)^dag (V unitary) this is Wbar(z) = V(z) sum_{z' <= z} V(z')^dag Vbar(z'), a running
// sum along z (k_scan_vjp_sum, one thread per transverse site, Wbar in place of Vbar), after which every layer's
// exponential adjoint is independent (k_scan_vjp_site, one thread per site): Xbar = Wbar(z) V(z+1)^dag.
__global__ void __launch_bounds__(256) k_scan_vjp_sum(const double* __restrict__ V, double* __restrict__ Vb, Geom g) {
    long s2 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s2 >= g.P) return;
    M3 acc, Vz, B, T;
    m3_zero(acc);
    for (int z = 0; z < g.nz; z++) {
        const long s = s2 + g.P * z;
        ldm<long>(V, kGPlanes, 0, s, Vz);
        ldm<long>(Vb, kGPlanes, 0, s, B);
        mul_dn(Vz, B, T);
        m3_axpy(1.0, T, acc);
        mul_nn(Vz, acc, T);
        stm<long>(Vb, kGPlanes, 0, s, T);
    }
}
// Abar (real, a*S + s) from Wbar
__global__ void __launch_bounds__(256) k_scan_vjp_site(const double* __restrict__ A, const double* __restrict__ V,
                                                       const double* __restrict__ Wb, double* __restrict__ Ab, Geom g) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    const double c = -1.0 / kSqrt2;
    const int z = (int)(s / g.P);
    M3 W, Wn, X, Xb;
    ldm<long>(Wb, kGPlanes, 0, s, W);
    if (z + 1 < g.nz) ldm<long>(V, kGPlanes, 0, s + g.P, Wn);
    else set_unit(Wn);
    double a[8], ab[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < 8; q++) a[q] = A[(long)q * g.S + s];
    expi(c, a, X);
    mul_nd(W, Wn, Xb);
    dexp_adjoint(c, a, X, Xb, ab);
    for (int q = 0; q < 8; q++) Ab[(long)q * g.S + s] = ab[q];
}

// ------------------------------------------------------------------------------------------------
// Physical-charge (TMD / GBW) nuclei: SC/TMDNUCLEUS/CorrelationFunction.cpp, GBWSampling.cpp, TMDbased.cpp
// and the initial state of SC/SimulationmitPC.cpp:326-370.
// ------------------------------------------------------------------------------------------------
// F = (noise / sqrt(a^3), 0); noise[a + 8 s] (ChargeFluctuation :29-39)
__global__ void k_tmd_noise(double* __restrict__ F, const double* __restrict__ noise, Geom g, double sqrt_acube) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    for (int a = 0; a < 8; a++) {
        const long i = fidx(g, a, s);
        F[2 * i] = noise[a + 8 * s] / sqrt_acube;
        F[2 * i + 1] = 0.0;
    }
}

// ChargeFluctuation's momentum-space factor (:45-79) on the 3D transform: nf sqrt(GammaK(x, kT^2)) for
// x = |p_z| / (CoMEnergy / sqrt2) <= 1, else 0; |p_z| = sqrt((2 - 2 cos(2 pi n_z / N_z)) / aZInFm^2) / sqrt2;
// GammaK = 4 pi Nc / (Nc^2 - 1) kT^4 / Qs^2 exp(-kT^2 / Qs^2), Qs = Qs0A x^-0.1 (1 - x) / sqrt(ax ay) (:17-22)
__global__ void k_tmd_cf_filter(double* __restrict__ F, Geom g, double ax, double ay, double qs0a, double com_energy,
                                double az_fm) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    const int pz = (int)(s / g.P);
    const long s2 = s - (long)pz * g.P;
    const int px = (int)(s2 % g.nx), py = (int)(s2 / g.nx);
    const double pl = (1 / kSqrt2) * sqrt((2.0 - 2.0 * cos(2.0 * kPi * pz / (double)g.nz)) / (az_fm * az_fm));
    const double kT2 = psqr2(g, px, py, ax, ay);
    const double x = pl / (com_energy / kSqrt2);
    const double nf = 1.0 / ((double)g.nx * g.ny * g.nz);
    double f = 0.0;
    if (x <= 1.0) {
        const double Qs = qs0a * pow(x, -0.1) * (1 - x) / sqrt(ax * ay);
        f = sqrt(4.0 * kPi * 3.0 / (3.0 * 3.0 - 1.0) * kT2 * kT2 / (Qs * Qs) * exp(-kT2 / (Qs * Qs)));
    }
    for (int a = 0; a < 8; a++) {
        const long i = fidx(g, a, s);
        if (x <= 1.0) {
            F[2 * i] = nf * F[2 * i] * f;
            F[2 * i + 1] = nf * F[2 * i + 1] * f;
        } else {
            F[2 * i] = 0.0;
            F[2 * i + 1] = 0.0;
        }
    }
}

struct TmdCharge {
    const double* pos;  // nucleon positions, 3 per nucleon (fm)
    int A;
    double afm[3];
    double gamma, acube, az;
    int zc;
};

// TmdBasedCollision::SetChargeDensities' charge (TMDbased.cpp:27-57): F = (a^3 sqrt(aZInFm T / a_z) Re F, 0), T the
// nucleus profile (GBWSampling.cpp:110-157) at x = (ix - Nx/2) aXInFm, y = (iy - Ny/2) aYInFm, z = (iz - centre) aZInFm
__global__ void k_tmd_charge(double* __restrict__ F, Geom g, TmdCharge C) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    const int z = (int)(s / g.P);
    const long s2 = s - (long)z * g.P;
    const int ix = (int)(s2 % g.nx), iy = (int)(s2 / g.nx);
    const double x = (ix - g.nx / 2) * C.afm[0], y = (iy - g.ny / 2) * C.afm[1], zf = (z - C.zc) * C.afm[2];
    const double BG = 4.0 * 0.197 * 0.197;
    double T = 0.0;
    for (int n = 0; n < C.A; n++) {
        const double dx = x - C.pos[3 * n], dy = y - C.pos[3 * n + 1], dz = C.gamma * zf - C.pos[3 * n + 2];
        const double d2 = dx * dx + dy * dy + dz * dz;
        T += C.gamma * exp(-0.5 * d2 / BG) / sqrt(2.0 * kPi * BG);
    }
    const double c = C.acube * sqrt(C.afm[2] * T / C.az);
    for (int a = 0; a < 8; a++) {
        const long i = fidx(g, a, s);
        F[2 * i] = c * F[2 * i];
        F[2 * i + 1] = 0.0;
    }
}

bool tmd_valid(const glasma_tmd_params* p) {
    return p && p->qs0a > 0.0 && p->m_eff >= 0.0 && p->com_energy > 0.0 && p->a_fm[0] > 0.0 && p->a_fm[1] > 0.0 &&
           p->a_fm[2] > 0.0 && p->atomic_number > 0;
}

// ChargeFluctuation + TmdBasedCollision::SetWilsonLines for one nucleus: w.rho = filtered charge, w.V
int tmd_wilson_lines(glasma_ctx* ctx, Wk& w, const glasma_tmd_params* p, const double* noise_host,
                     const double* pos_host, bool right) {
    const Geom& g = ctx->g;
    const double* a = ctx->p.a;
    // the noise goes through w.rho (8 S doubles), which the filter step overwrites later
    GL_HIP(ctx, hipMemcpyAsync(w.rho, noise_host, 8 * (size_t)g.S * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    double* pos = nullptr;
    GL_HIP(ctx, hipMalloc(&pos, 3 * (size_t)p->atomic_number * sizeof(double)));
    hipError_t e = hipMemcpyAsync(pos, pos_host, 3 * (size_t)p->atomic_number * sizeof(double), hipMemcpyHostToDevice,
                                  ctx->stream);
    int rc = 0;
    if (e != hipSuccess) rc = hip_fail(ctx, e, "tmd nucleon positions");
    if (!rc) {
        hipLaunchKernelGGL(k_tmd_noise, grid1(g.S), dim3(256), 0, ctx->stream, w.F, w.rho, g, std::sqrt(a[0] * a[1] * a[2]));
        rc = fft_3d(ctx->stream, w.F, g, 8, -1);
        if (!rc) {
            hipLaunchKernelGGL(k_tmd_cf_filter, grid1(g.S), dim3(256), 0, ctx->stream, w.F, g, a[0], a[1], p->qs0a,
                               p->com_energy, p->a_fm[2]);
            rc = fft_3d(ctx->stream, w.F, g, 8, +1);
        }
        if (rc)
            rc = set_error(ctx, rc,
                           std::string("FFT failed (lengths must be 2^a 3^b p <= 2048, prime factors <= 13): ") +
                               fft_last_error());
    }
    if (!rc) {
        TmdCharge C;
        C.pos = pos;
        C.A = p->atomic_number;
        for (int m = 0; m < 3; m++) C.afm[m] = p->a_fm[m];
        C.gamma = p->com_energy * 0.197 / 2.0;
        C.acube = a[0] * a[1] * a[2];
        C.az = a[2];
        C.zc = right ? p->center_r : p->center_l;
        hipLaunchKernelGGL(k_tmd_charge, grid1(g.S), dim3(256), 0, ctx->stream, w.F, g, C);
        rc = wilson_from_charge(ctx, w, std::numeric_limits<double>::infinity(), p->m_eff, right);
    }
    (void)hipStreamSynchronize(ctx->stream);  // the positions buffer is freed below
    (void)hipFree(pos);
    return rc;
}

bool mv_valid(const glasma_mv_params* mv) {
    return mv && mv->sigma_l > 0.0 && mv->sigma_r > 0.0 && mv->lambda > 0.0 && mv->m_eff >= 0.0;
}

}  // namespace

// Left-moving sheet over all z from a given charge density rho (device, real, a*S + s; the density
// SetChargeDensities would produce), keeping the intermediates in tape (device arrays of 8*S doubles
// for rho_s, A0, A1 and tiled 18-plane Wilson lines V0, V1).  Resets the fields and the time first.
int ic_left_from_charge(glasma_ctx* ctx, const glasma_mv_params* mv, const double* rho, IcTape& tape) {
    const Geom& g = ctx->g;
    int rc = reset_fields(ctx);
    if (rc) return rc;
    rc = ensure_fftbuf(ctx, 16 * (size_t)g.S);
    if (rc) return rc;
    double* F = ctx->fftbuf;
    hipStream_t st = ctx->stream;
    double* U = ctx->U[ctx->cur];
    double* J = ctx->J[ctx->cur];
    const double dt = ctx->p.dt;
    hipLaunchKernelGGL(k_load_real, grid1(g.S), dim3(256), 0, st, F, rho, g);
    GL_FFT(ctx, fft_xy(st, F, g, 8, g.nz, -1));
    hipLaunchKernelGGL(k_transverse_filter, grid1(g.S), dim3(256), 0, st, F, g, ctx->p.a[0], ctx->p.a[1], mv->lambda,
                       mv->m_eff, 0);
    GL_FFT(ctx, fft_xy(st, F, g, 8, g.nz, +1));
    hipLaunchKernelGGL(k_take_real, grid1(g.S), dim3(256), 0, st, F, tape.rho_s, g);
    GL_RC(poisson_all(ctx, F));
    hipLaunchKernelGGL(k_copy_real, grid1(8 * g.S), dim3(256), 0, st, tape.A0, F, g);
    hipLaunchKernelGGL(k_wilson_scan, grid1(g.P), dim3(256), 0, st, F, tape.V0, g, 0);
    hipLaunchKernelGGL(k_sheet_links, grid1(g.S), dim3(256), 0, st, tape.V0, ctx->UOld, g, 0, g.nz - 1);
    hipLaunchKernelGGL(k_load_real, grid1(g.S), dim3(256), 0, st, F, tape.rho_s, g);
    GL_FFT(ctx, fft_z(st, F, g, 8, -1));
    hipLaunchKernelGGL(k_evolve_phase, grid1(g.S), dim3(256), 0, st, F, g, ctx->p.a[2], dt, 0);
    GL_FFT(ctx, fft_z(st, F, g, 8, +1));
    hipLaunchKernelGGL(k_take_real, grid1(g.S), dim3(256), 0, st, F, (double*)nullptr, g);
    GL_RC(poisson_all(ctx, F));
    hipLaunchKernelGGL(k_copy_real, grid1(8 * g.S), dim3(256), 0, st, tape.A1, F, g);
    hipLaunchKernelGGL(k_wilson_scan, grid1(g.P), dim3(256), 0, st, F, tape.V1, g, 0);
    hipLaunchKernelGGL(k_sheet_links, grid1(g.S), dim3(256), 0, st, tape.V1, U, g, 0, g.nz - 1);
    const double acube = ctx->p.a[0] * ctx->p.a[1] * ctx->p.a[2];
    const double coff0 = -dt * ctx->p.a[0] * ctx->p.a[0] / acube;
    const double coff1 = -dt * ctx->p.a[1] * ctx->p.a[1] / acube;
    hipLaunchKernelGGL(k_sheet_efield, grid1(g.S), dim3(256), 0, st, U, ctx->UOld, ctx->E, g, 0, g.nz - 1, coff0, coff1);
    hipLaunchKernelGGL(k_sheet_jstat, grid1(g.S), dim3(256), 0, st, U, ctx->E, J, F, g, 0, g.nz - 1, JPS);
    GL_FFT(ctx, fft_z(st, F, g, 8, -1));
    hipLaunchKernelGGL(k_jinv, grid1(g.S), dim3(256), 0, st, F, g, ctx->p.a[2], dt, 0, 0);
    GL_FFT(ctx, fft_z(st, F, g, 8, +1));
    hipLaunchKernelGGL(k_store_j, grid1(g.S), dim3(256), 0, st, F, J, g, JP);
    GL_HIP(ctx, hipGetLastError());
    return 0;
}

// Reverse-mode derivative of ic_left_from_charge.  In: cotangents of the sheet's U (Ub, tiled 54 planes;
// modified), E (Eb, tiled 24; modified) and J (Jb, tiled 32: J+ and J+stat used), the fields and tape of
// the same call.  Out: rho_bar (device, real, a*S + s).  work: 54*Sp + 2*18*Sp + 4*8*S doubles.
int ic_left_vjp(glasma_ctx* ctx, const glasma_mv_params* mv, const IcTape& tape, double* Ub, double* Eb,
                const double* Jb, double* rho_bar, double* work) {
    const Geom& g = ctx->g;
    const long Sp = pad_sites(g.S);
    hipStream_t st = ctx->stream;
    int rc = ensure_fftbuf(ctx, 16 * (size_t)g.S);
    if (rc) return rc;
    double* F = ctx->fftbuf;
    double* UOb = work;
    double* V0b = UOb + 54 * Sp;
    double* V1b = V0b + 18 * Sp;
    double* jsb = V1b + 18 * Sp;
    double* A0b = jsb + 8 * g.S;
    double* A1b = A0b + 8 * g.S;
    double* tmp = A1b + 8 * g.S;
    const double* U = ctx->U[ctx->cur];
    const double dt = ctx->p.dt;
    // J+ = jinv(J+stat): J+stat cotangent += jinv^T(J+ cotangent)
    hipLaunchKernelGGL(k_load_planes, grid1(g.S), dim3(256), 0, st, F, Jb, g, kJPlanes, 8 * JP);
    GL_FFT(ctx, fft_z(st, F, g, 8, -1));
    hipLaunchKernelGGL(k_jinv, grid1(g.S), dim3(256), 0, st, F, g, ctx->p.a[2], dt, 0, 1);
    GL_FFT(ctx, fft_z(st, F, g, 8, +1));
    hipLaunchKernelGGL(k_copy_real, grid1(8 * g.S), dim3(256), 0, st, tmp, F, g);
    hipLaunchKernelGGL(k_load_planes, grid1(g.S), dim3(256), 0, st, F, Jb, g, kJPlanes, 8 * JPS);
    hipLaunchKernelGGL(k_axpy_real, grid1(8 * g.S), dim3(256), 0, st, jsb, tmp, F, 1.0, g);
    // J+stat from E and U
    hipLaunchKernelGGL(k_ic_jstat_vjp, grid1(g.S), dim3(256), 0, st, U, ctx->E, jsb, Ub, Eb, g);
    // E = ilog(U UOld^dag)
    const double acube = ctx->p.a[0] * ctx->p.a[1] * ctx->p.a[2];
    const double coff0 = -dt * ctx->p.a[0] * ctx->p.a[0] / acube;
    const double coff1 = -dt * ctx->p.a[1] * ctx->p.a[1] / acube;
    hipLaunchKernelGGL(k_zero, grid1(54 * Sp), dim3(256), 0, st, UOb, 54 * Sp);
    hipLaunchKernelGGL(k_ic_ilog_vjp, grid1(g.S), dim3(256), 0, st, U, ctx->UOld, ctx->E, Eb, Ub, UOb, g, coff0, coff1);
    // links -> Wilson lines -> Poisson solutions
    hipLaunchKernelGGL(k_links_vjp, grid1(g.S), dim3(256), 0, st, tape.V1, Ub, V1b, g);
    hipLaunchKernelGGL(k_links_vjp, grid1(g.S), dim3(256), 0, st, tape.V0, UOb, V0b, g);
    hipLaunchKernelGGL(k_scan_vjp_sum, grid1(g.P), dim3(256), 0, st, tape.V1, V1b, g);
    hipLaunchKernelGGL(k_scan_vjp_sum, grid1(g.P), dim3(256), 0, st, tape.V0, V0b, g);
    hipLaunchKernelGGL(k_scan_vjp_site, grid1(g.S), dim3(256), 0, st, tape.A1, tape.V1, V1b, A1b, g);
    hipLaunchKernelGGL(k_scan_vjp_site, grid1(g.S), dim3(256), 0, st, tape.A0, tape.V0, V0b, A0b, g);
    // rho_ev cotangent = poisson(A1bar); rho_s cotangent = evolve_R(that) + poisson(A0bar)
    hipLaunchKernelGGL(k_load_real, grid1(g.S), dim3(256), 0, st, F, A1b, g);
    GL_RC(poisson_all(ctx, F));
    hipLaunchKernelGGL(k_take_real, grid1(g.S), dim3(256), 0, st, F, (double*)nullptr, g);
    GL_FFT(ctx, fft_z(st, F, g, 8, -1));
    hipLaunchKernelGGL(k_evolve_phase, grid1(g.S), dim3(256), 0, st, F, g, ctx->p.a[2], dt, 1);
    GL_FFT(ctx, fft_z(st, F, g, 8, +1));
    hipLaunchKernelGGL(k_copy_real, grid1(8 * g.S), dim3(256), 0, st, tmp, F, g);
    hipLaunchKernelGGL(k_load_real, grid1(g.S), dim3(256), 0, st, F, A0b, g);
    GL_RC(poisson_all(ctx, F));
    hipLaunchKernelGGL(k_axpy_real, grid1(8 * g.S), dim3(256), 0, st, tmp, tmp, F, 1.0, g);
    // rho cotangent = smooth(rho_s cotangent)
    hipLaunchKernelGGL(k_load_real, grid1(g.S), dim3(256), 0, st, F, tmp, g);
    GL_FFT(ctx, fft_xy(st, F, g, 8, g.nz, -1));
    hipLaunchKernelGGL(k_transverse_filter, grid1(g.S), dim3(256), 0, st, F, g, ctx->p.a[0], ctx->p.a[1], mv->lambda,
                       mv->m_eff, 0);
    GL_FFT(ctx, fft_xy(st, F, g, 8, g.nz, +1));
    hipLaunchKernelGGL(k_copy_real, grid1(8 * g.S), dim3(256), 0, st, rho_bar, F, g);
    GL_HIP(ctx, hipGetLastError());
    return 0;
}

namespace {

}  // namespace

}  // namespace glasma

using namespace glasma;

namespace glasma {
namespace {
// The part of reset_fields a sheet initial state needs.  The sheets write the current links, UOld and E at every
// site (the collision's two sheets cover [0, (Nz-2)/2] and [Nz/2, Nz-1], a single nucleus all of z), and whole planes
// of the current J (k_sheet_jstat, k_store_j); the planes of the nucleus a single-nucleus driver leaves out must be
// zero.  The other buffers of the ping-pong pairs are written by the first step before anything reads them.
int reset_for_sheets(glasma_ctx* ctx, int driver) {
    ctx->cur = 0;
    ctx->t = 0.0;
    ctx->tsteps = 0;
    if (driver != 0)
        GL_HIP(ctx, hipMemsetAsync(ctx->J[0], 0, (size_t)kJPlanes * pad_sites(ctx->g.S) * sizeof(double), ctx->stream));
    return 0;
}
}  // namespace
}  // namespace glasma

extern "C" int glasma_mv_initial_state(glasma_ctx* ctx, int driver, const glasma_mv_params* mv, const double* xi_l,
                                       const double* xi_r) {
    if (!ctx) return GLASMA_ERR_ARG;
    if (!mv_valid(mv)) return set_error(ctx, GLASMA_ERR_ARG, "invalid MV parameters");
    if (driver < 0 || driver > 2) return set_error(ctx, GLASMA_ERR_ARG, "driver must be 0, 1 or 2");
    if ((driver != 2 && !xi_l) || (driver != 1 && !xi_r)) return set_error(ctx, GLASMA_ERR_ARG, "missing charges");
    int rc = reset_for_sheets(ctx, driver);
    if (rc) return rc;
    const int nz = ctx->g.nz;
    Wk w;
    rc = alloc_work(ctx, w);
    if (rc == 0 && driver != 2) {
        // collision: left sheet on [0, (Nz-2)/2] (SC/Simulation.cpp:359); single nucleus: all z
        const int zhi = driver == 0 ? (nz - 2) / 2 : nz - 1;
        if (mv->m_eff > 0.0) {
            rc = mv_sheet(ctx, w, mv, xi_l, false, 0, zhi, JPS, JP);
        } else {
            rc = set_wilson_lines(ctx, w, mv, xi_l, false);
            if (rc == 0) rc = single_sheet(ctx, w, false, 0, zhi, JPS, JP);
        }
    }
    if (rc == 0 && driver != 1) {
        const int zlo = driver == 0 ? nz / 2 : 0;
        if (mv->m_eff > 0.0) {
            rc = mv_sheet(ctx, w, mv, xi_r, true, zlo, nz - 1, JMS, JM);
        } else {
            rc = set_wilson_lines(ctx, w, mv, xi_r, true);
            if (rc == 0) rc = single_sheet(ctx, w, true, zlo, nz - 1, JMS, JM);
        }
    }
    if (rc == 0) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_fail(ctx, e, "mv_initial_state");
    }
    free_work(w);
    return rc;
}

extern "C" int glasma_tmd_initial_state(glasma_ctx* ctx, int driver, const glasma_tmd_params* p, const double* noise_l,
                                        const double* noise_r, const double* nucleons_l, const double* nucleons_r) {
    if (!ctx) return GLASMA_ERR_ARG;
    if (!tmd_valid(p)) return set_error(ctx, GLASMA_ERR_ARG, "invalid physical-charge parameters");
    if (driver < 0 || driver > 2) return set_error(ctx, GLASMA_ERR_ARG, "driver must be 0, 1 or 2");
    if ((driver != 2 && (!noise_l || !nucleons_l)) || (driver != 1 && (!noise_r || !nucleons_r)))
        return set_error(ctx, GLASMA_ERR_ARG, "missing noise or nucleon positions");
    int rc = reset_for_sheets(ctx, driver);
    if (rc) return rc;
    const int nz = ctx->g.nz;
    Wk w;
    rc = alloc_work(ctx, w);
    if (rc == 0 && driver != 2) {
        const int zhi = driver == 0 ? (nz - 2) / 2 : nz - 1;
        rc = tmd_wilson_lines(ctx, w, p, noise_l, nucleons_l, false);
        if (rc == 0) rc = single_sheet(ctx, w, false, 0, zhi, JPS, JP);
    }
    if (rc == 0 && driver != 1) {
        const int zlo = driver == 0 ? nz / 2 : 0;
        rc = tmd_wilson_lines(ctx, w, p, noise_r, nucleons_r, true);
        if (rc == 0) rc = single_sheet(ctx, w, true, zlo, nz - 1, JMS, JM);
    }
    if (rc == 0) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_fail(ctx, e, "tmd_initial_state");
    }
    free_work(w);
    return rc;
}

extern "C" int glasma_set_wilson_lines(glasma_ctx* ctx, char dir, const glasma_mv_params* mv, const double* xi,
                                       double* V_out, double* rho_out) {
    if (!ctx) return GLASMA_ERR_ARG;
    if (!mv_valid(mv) || !xi || (dir != 'L' && dir != 'R'))
        return set_error(ctx, GLASMA_ERR_ARG, "invalid arguments to glasma_set_wilson_lines");
    Wk w;
    int rc = alloc_work(ctx, w);
    if (rc == 0) rc = set_wilson_lines(ctx, w, mv, xi, dir == 'R');
    if (rc == 0 && V_out) rc = download_aos(ctx, w.V, V_out, kGPlanes);
    if (rc == 0 && rho_out) {
        const Geom& g = ctx->g;
        double* tmp = nullptr;
        hipError_t e = hipMalloc(&tmp, 8 * (size_t)g.S * sizeof(double));
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_rho_to_aos, grid1(8 * g.S), dim3(256), 0, ctx->stream, w.rho, tmp, g);
            e = hipMemcpyAsync(rho_out + 8 * g.P, tmp, 8 * (size_t)g.S * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            (void)hipFree(tmp);
        }
        if (e != hipSuccess) rc = hip_fail(ctx, e, "set_wilson_lines rho");
        else host_sync_ghosts(ctx, rho_out, 8);
    }
    free_work(w);
    return rc;
}
