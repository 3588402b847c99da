// Colour charges of the MV model, drawn on the host exactly as the reference draws them:
//   RandomNumberGenerator (SC/MISC/RNG/GSLRandomNumberGenerator.cpp:12-33): gsl_rng_default (mt19937) seeded
//   with the global seed through gsl_rng_set, Gauss() = gsl_ran_gaussian(rng, 1.0);
//   CollidingObjects::SetChargeDensities draw order (SC/COLLIDINGOBJECTS/CollidingObjects.cpp:335-342):
//   for y, for x, for a: xi = QsA * Gauss(); the left nucleus first, then the right one from the same
//   stream (SC/Simulation.cpp:344, :348).
//
// Generator: MT19937 of Matsumoto & Nishimura (ACM TOMACS 8 (1998) 3-30), state initialisation by the
// 2002 init_genrand recursion x_i = 1812433253 (x_{i-1} ^ (x_{i-1} >> 30)) + i; GSL maps seed 0 to 4357.
// Uniform on (0,1): 32-bit output / 2^32, zero rejected (gsl_rng_uniform_pos).
// Normal deviates: Marsaglia's polar method (Knuth, TAOCP vol. 2, 3.4.1, algorithm P) returning one
// deviate per accepted pair, y * sqrt(-2 ln r^2 / r^2) with x, y = 2u - 1 drawn in that order, as
// gsl_ran_gaussian documents.
#include <cmath>
#include <cstdint>

#include "../../include/glasma.h"

// every product and sum rounded separately (no fused multiply-add), as the sequence is specified
#pragma clang fp contract(off)

namespace glasma {
namespace {

class Mt19937 {
public:
    explicit Mt19937(unsigned long seed) {
        uint32_t s = (uint32_t)(seed == 0 ? 4357UL : seed);
        st_[0] = s;
        for (int i = 1; i < kN; i++) st_[i] = 1812433253u * (st_[i - 1] ^ (st_[i - 1] >> 30)) + (uint32_t)i;
        pos_ = kN;
    }
    uint32_t operator()() {
        if (pos_ == kN) regenerate();
        uint32_t y = st_[pos_++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        return y;
    }
    double uniform() { return (double)(*this)() * (1.0 / 4294967296.0); }  // gsl_rng_uniform: [0, 1)
    double uniform_pos() {
        for (;;) {
            const double u = (double)(*this)() * (1.0 / 4294967296.0);
            if (u != 0.0) return u;
        }
    }
    double normal() {
        for (;;) {
            const double x = 2.0 * uniform_pos() - 1.0;
            const double y = 2.0 * uniform_pos() - 1.0;
            const double r2 = x * x + y * y;
            if (r2 <= 1.0 && r2 != 0.0) return y * std::sqrt(-2.0 * std::log(r2) / r2);
        }
    }

private:
    static constexpr int kN = 624, kM = 397;
    uint32_t st_[kN];
    int pos_;
    static uint32_t twist(uint32_t hi, uint32_t lo, uint32_t far) {
        const uint32_t v = (hi & 0x80000000u) | (lo & 0x7fffffffu);
        return far ^ (v >> 1) ^ ((v & 1u) ? 0x9908b0dfu : 0u);
    }
    void regenerate() {
        for (int i = 0; i < kN; i++) st_[i] = twist(st_[i], st_[(i + 1) % kN], st_[(i + kM) % kN]);
        pos_ = 0;
    }
};

}  // namespace
}  // namespace glasma

extern "C" int glasma_draw_charges(unsigned long seed, int nx, int ny, double qs_l, double qs_r, double* xi_l,
                                   double* xi_r) {
    if (nx <= 0 || ny <= 0 || !xi_l || !xi_r) return GLASMA_ERR_ARG;
    glasma::Mt19937 g(seed);
    const long n = 8L * nx * ny;  // y, x, a order = index a + 8*(x + nx*y)
    for (long i = 0; i < n; i++) xi_l[i] = qs_l * g.normal();
    for (long i = 0; i < n; i++) xi_r[i] = qs_r * g.normal();
    return 0;
}

// Woods-Saxon radius and surface thickness (SC/TMDNUCLEUS/NuclearData.cpp:16-87)
static bool woods_saxon(int A, double* R, double* d) {
    switch (A) {
        case 16: *R = 2.608; *d = 0.5130; return true;
        case 54: *R = 5.420; *d = 0.57; return true;
        case 197: *R = 6.38; *d = 0.535; return true;
        case 208: *R = 6.62; *d = 0.546; return true;
        default: return false;
    }
}

// Nucleus::SetNucleonPositions (SC/TMDNUCLEUS/GBWSampling.cpp:21-87): for each nucleon, uniform positions in
// [-4R, 4R)^3 accepted with probability (1 + e^{-R/d}) / (1 + e^{(r - R)/d}) against a fourth uniform; then the
// centre of mass, accumulated as the sum of x_n / A, subtracted.
static void sample_nucleus(glasma::Mt19937& g, int A, double R, double d, double* pos) {
    for (int n = 0; n < A; n++) {
        double* x = pos + 3 * n;
        for (;;) {
            x[0] = 8.0 * R * (g.uniform() - 0.5);
            x[1] = 8.0 * R * (g.uniform() - 0.5);
            x[2] = 8.0 * R * (g.uniform() - 0.5);
            const double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
            const double prob = (1.0 + std::exp(-R / d)) / (1.0 + std::exp((r - R) / d));
            if (g.uniform() < prob) break;
        }
    }
    double bar[3] = {0.0, 0.0, 0.0};
    for (int n = 0; n < A; n++)
        for (int i = 0; i < 3; i++) bar[i] += pos[3 * n + i] / double(A);
    for (int n = 0; n < A; n++)
        for (int i = 0; i < 3; i++) pos[3 * n + i] -= bar[i];
}

extern "C" int glasma_tmd_draw(unsigned long seed, int nx, int ny, int nz, int atomic_number, double* noise_l,
                               double* noise_r, double* nucleons_l, double* nucleons_r) {
    double R, d;
    if (nx <= 0 || ny <= 0 || nz <= 0 || !noise_l || !noise_r || !nucleons_l || !nucleons_r ||
        !woods_saxon(atomic_number, &R, &d))
        return GLASMA_ERR_ARG;
    glasma::Mt19937 g(seed);
    const long n = 8L * nx * ny * nz;
    for (long i = 0; i < n; i++) noise_l[i] = g.normal();
    sample_nucleus(g, atomic_number, R, d, nucleons_l);
    for (long i = 0; i < n; i++) noise_r[i] = g.normal();
    sample_nucleus(g, atomic_number, R, d, nucleons_r);
    return 0;
}
