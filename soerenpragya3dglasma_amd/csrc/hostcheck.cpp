// Host (g++) build of the per-site product code in su3.h / site.h, for CPU tests of the product
// arithmetic against the oracle and scipy.  Not part of libglasma and never a fallback for it: the
// device library is the only execution path of the engine.  Field arrays use the device's tiled layout.
#include <cstring>

#include "adjoint.h"

using namespace glasma;

static void to_m3(const double* in, M3& A) {  // interleaved complex column-major -> M3
    for (int k = 0; k < 9; k++) { A.re[k] = in[2 * k]; A.im[k] = in[2 * k + 1]; }
}
static void from_m3(const M3& A, double* out) {
    for (int k = 0; k < 9; k++) { out[2 * k] = A.re[k]; out[2 * k + 1] = A.im[k]; }
}

extern "C" {

void hc_su3_expi(double c, const double* al, double* X) { M3 m; expi(c, al, m); from_m3(m, X); }
void hc_su3_expi_m1(double c, const double* al, double* X) { M3 m; expi_m1(c, al, m); from_m3(m, X); }
void hc_su3_ilog(double c, const double* U, double* out) { M3 m; to_m3(U, m); ilog(c, m, out); }
void hc_su3_adj_mul(const double* V, const double* al, double* out) { M3 m; to_m3(V, m); adj_mul(m, al, out); }
void hc_su3_inv_adj_mul(const double* V, const double* al, double* out) { M3 m; to_m3(V, m); inv_adj_mul(m, al, out); }
void hc_su3_retrigenu(const double* V, double* out) { M3 m; to_m3(V, m); retr_igen(m, out); }
// z = sum_k ad_A^k(y) / (k+1)! on su(3) coordinates, A = i c e^a t^a (adjoint.h ad_series)
void hc_ad_series(double c, const double* e, const double* y, double* z) {
    double yy[8];
    std::memcpy(yy, y, sizeof yy);
    ad_series(c, e, yy, z);
}

// first site of every block of a step under the tile order (mode, bpb, bw, bh, nbx, tpbp) (site.h tile_site)
void hc_tile_sites(const int* n, int mode, int bpb, int bw, int bh, int nbx, int tpbp, long nblocks, long* out) {
    Geom g;
    g.nx = n[0]; g.ny = n[1]; g.nz = n[2]; g.bc = 0;
    g.P = (long)n[0] * n[1];
    g.S = g.P * n[2];
    TileMap m = {mode, bpb, bw, bh, nbx, tpbp};
    for (long b = 0; b < nblocks; b++) out[b] = tile_site(g, m, b);
}

// One fused leapfrog step over tiled device-layout arrays, sequentially over sites.
void hc_step(const int* n, const double* a, int bc, double dt, const double* U, double* Un, double* E,
             const double* J, double* Jn, double* UOld) {
    StepArgs A;
    A.U = U; A.Un = Un; A.E = E; A.J = J; A.Jn = Jn; A.UOld = UOld;
    A.g.nx = n[0]; A.g.ny = n[1]; A.g.nz = n[2]; A.g.bc = bc;
    A.g.P = (long)n[0] * n[1];
    A.g.S = A.g.P * n[2];
    fill_step_coefficients(A, a, dt);
    for (long s = 0; s < A.g.S; s++) step_site<unsigned long>(A, (unsigned long)s);
}


// one backward step: passes A, B, C over all sites (C accumulates serially)
void hc_step_vjp(const int* n, const double* a, int bc, double dt, const double* U, const double* Uh, const double* E1,
                 const double* J, const double* J1, const double* Ub1, double* Ub, double* Eb, const double* Jb1,
                 double* Jb) {
    StepArgs F;
    fill_step_coefficients(F, a, dt);
    StepAdjArgs A;
    A.U = U; A.Uh = Uh; A.E1 = E1; A.J = J; A.J1 = J1; A.Ub1 = Ub1; A.Ub = Ub; A.Eb = Eb; A.Jb1 = Jb1; A.Jb = Jb;
    A.g.nx = n[0]; A.g.ny = n[1]; A.g.nz = n[2]; A.g.bc = bc;
    A.g.P = (long)n[0] * n[1];
    A.g.S = A.g.P * n[2];
    for (int m = 0; m < 3; m++) { A.gc[m] = F.gc[m]; A.cE[m] = F.cE[m]; A.cU[m] = F.cU[m]; }
    A.dtz = F.dtz;
    for (long s = 0; s < A.g.S; s++) step_vjp_local<unsigned long>(A, (unsigned long)s);
    for (long s = 0; s < A.g.S; s++) step_vjp_down<unsigned long>(A, (unsigned long)s);
    auto add = [&](int mu, unsigned long site, const M3& M) {
        for (int k = 0; k < 9; k++) {
            Ub[lidx<unsigned long>(kUPlanes, 18 * mu + 2 * k, site)] += M.re[k];
            Ub[lidx<unsigned long>(kUPlanes, 18 * mu + 2 * k + 1, site)] += M.im[k];
        }
    };
    for (long s = 0; s < A.g.S; s++) step_vjp_staples<unsigned long>(A, (unsigned long)s, add);
}

static SlabArgs mk_slab(const int* n, const double* a, int bc, const double* U, const double* E, double* Eb, int zc,
                        int half) {
    SlabArgs A;
    A.U = U; A.E = E; A.Eb = Eb;
    A.g.nx = n[0]; A.g.ny = n[1]; A.g.nz = n[2]; A.g.bc = bc;
    A.g.P = (long)n[0] * n[1];
    A.g.S = A.g.P * n[2];
    const double acube = a[0] * a[1] * a[2];
    for (int m = 0; m < 3; m++) A.c[m] = (a[m] / acube) * (a[m] / acube);
    A.zc = zc; A.half = half;
    return A;
}

double hc_slab_energy(const int* n, const double* a, int bc, const double* U, const double* E, int zc, int half) {
    SlabArgs A = mk_slab(n, a, bc, U, E, nullptr, zc, half);
    double v = 0.0;
    for (long s = 0; s < A.g.S; s++) v += slab_energy_site<unsigned long>(A, (unsigned long)s);
    return v;
}

void hc_slab_energy_vjp(const int* n, const double* a, int bc, const double* U, const double* E, int zc, int half,
                        double* Ub, double* Eb) {
    SlabArgs A = mk_slab(n, a, bc, U, E, Eb, zc, half);
    auto add = [&](int mu, unsigned long site, const M3& M) {
        for (int k = 0; k < 9; k++) {
            Ub[lidx<unsigned long>(kUPlanes, 18 * mu + 2 * k, site)] += M.re[k];
            Ub[lidx<unsigned long>(kUPlanes, 18 * mu + 2 * k + 1, site)] += M.im[k];
        }
    };
    for (long s = 0; s < A.g.S; s++) slab_energy_vjp_site<unsigned long>(A, (unsigned long)s, 1.0, add);
}


// one backward step with the gather form of pass C (C1 + C2); Sb: 54-plane scratch
void hc_step_vjp_gather(const int* n, const double* a, int bc, double dt, const double* U, const double* Uh,
                        const double* E1, const double* J, const double* J1, const double* Ub1, double* Ub, double* Eb,
                        const double* Jb1, double* Jb, double* Sb) {
    StepArgs F;
    fill_step_coefficients(F, a, dt);
    StepAdjArgs A;
    A.U = U; A.Uh = Uh; A.E1 = E1; A.J = J; A.J1 = J1; A.Ub1 = Ub1; A.Ub = Ub; A.Eb = Eb; A.Jb1 = Jb1; A.Jb = Jb;
    A.g.nx = n[0]; A.g.ny = n[1]; A.g.nz = n[2]; A.g.bc = bc;
    A.g.P = (long)n[0] * n[1];
    A.g.S = A.g.P * n[2];
    for (int m = 0; m < 3; m++) { A.gc[m] = F.gc[m]; A.cE[m] = F.cE[m]; A.cU[m] = F.cU[m]; }
    A.dtz = F.dtz;
    for (long s = 0; s < A.g.S; s++) step_vjp_local<unsigned long>(A, (unsigned long)s);
    for (long s = 0; s < A.g.S; s++) step_vjp_down<unsigned long>(A, (unsigned long)s);
    for (long s = 0; s < A.g.S; s++) step_vjp_sigbar<unsigned long>(A, (unsigned long)s, Sb);
    for (long s = 0; s < A.g.S; s++) step_vjp_gather<unsigned long>(A, (unsigned long)s);
}

// one backward step in the device's two passes: A with B folded in, then the gather with the own-link term
void hc_step_vjp_fused(const int* n, const double* a, int bc, double dt, const double* U, const double* Uh,
                       const double* E1, const double* J, const double* J1, const double* Ub1, double* Ub, double* Eb,
                       const double* Jb1, double* Jb) {
    StepArgs F;
    fill_step_coefficients(F, a, dt);
    StepAdjArgs A;
    A.U = U; A.Uh = Uh; A.E1 = E1; A.J = J; A.J1 = J1; A.Ub1 = Ub1; A.Ub = Ub; A.Eb = Eb; A.Jb1 = Jb1; A.Jb = Jb;
    A.g.nx = n[0]; A.g.ny = n[1]; A.g.nz = n[2]; A.g.bc = bc;
    A.g.P = (long)n[0] * n[1];
    A.g.S = A.g.P * n[2];
    for (int m = 0; m < 3; m++) { A.gc[m] = F.gc[m]; A.cE[m] = F.cE[m]; A.cU[m] = F.cU[m]; }
    A.dtz = F.dtz;
    for (long s = 0; s < A.g.S; s++) step_vjp_local<unsigned long, true>(A, (unsigned long)s);
    for (long s = 0; s < A.g.S; s++) step_vjp_gather<unsigned long, true>(A, (unsigned long)s);
}

}
