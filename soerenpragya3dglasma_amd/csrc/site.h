// Per-site kernels bodies of the glasma engine, written once for device and host.
//
// The HIP kernels call these functions for their site; the host build (hostcheck.cpp, compiled with
// g++) calls the very same code so that the arithmetic of the product can also be exercised on a CPU.
#pragma once

#include "su3.h"

namespace glasma {

// ------------------------------------------------------------------------------------------------
// Device layout (structure of arrays, no ghost layers).  Site index s = x + Nx*(y + Ny*z), S sites.
//   links    U[(mu*18 + 2k + c) * S + s]    k = i + 3j, c = 0 real / 1 imaginary   (54 planes)
//   E field  E[(mu*8 + a) * S + s]                                                  (24 planes)
//   currents J[(w*8 + a) * S + s]  w = 0 J+, 1 J-, 2 J+stat, 3 J-stat              (32 planes)
// Every plane is contiguous in x, so a 64-lane wavefront reading one plane entry for 64 consecutive
// sites issues one fully coalesced 512-byte access.
// ------------------------------------------------------------------------------------------------
constexpr int kUPlanes = 54, kEPlanes = 24, kJPlanes = 32;
enum { JP = 0, JM = 1, JPS = 2, JMS = 3 };

struct Geom {
    int nx, ny, nz;
    int bc;   // 0 open, 1 periodic in z
    long S;   // nx*ny*nz
    long P;   // nx*ny
};

// z neighbour with the reference's ghost-layer semantics (open: the ghost is a copy of the edge layer;
// periodic: wrap), SC/LATTICE/GaugeLinks.cpp:59-73
GHD int zup(const Geom& g, int z) { return (z + 1 < g.nz) ? z + 1 : (g.bc ? 0 : g.nz - 1); }
GHD int zdn(const Geom& g, int z) { return (z > 0) ? z - 1 : (g.bc ? g.nz - 1 : 0); }
GHD int xwrap(int x, int n) { return x < 0 ? x + n : (x >= n ? x - n : x); }
GHD long sidx(const Geom& g, int x, int y, int z) {
    return (long)xwrap(x, g.nx) + (long)g.nx * ((long)xwrap(y, g.ny) + (long)g.ny * z);
}

GHD void ldm(const double* __restrict__ base, long S, long s, M3& A) {
#pragma unroll
    for (int k = 0; k < 9; k++) {
        A.re[k] = base[(2 * k) * S + s];
        A.im[k] = base[(2 * k + 1) * S + s];
    }
}
GHD void stm(double* __restrict__ base, long S, long s, const M3& A) {
#pragma unroll
    for (int k = 0; k < 9; k++) {
        base[(2 * k) * S + s] = A.re[k];
        base[(2 * k + 1) * S + s] = A.im[k];
    }
}
GHD void ldv(const double* __restrict__ base, long S, long s, double* v) {
#pragma unroll
    for (int a = 0; a < 8; a++) v[a] = base[a * S + s];
}
GHD void stv(double* __restrict__ base, long S, long s, const double* v) {
#pragma unroll
    for (int a = 0; a < 8; a++) base[a * S + s] = v[a];
}

// ------------------------------------------------------------------------------------------------
// The fused leapfrog step of one site: Dynamics::SUNc::Update (SC/DYNAMICS/SUNcDynamics.cpp:353-378).
//
//   E'(x)       = E(x) + gamma c_mu sum_nu c_nu [Re tr(i t P_mu,nu) - Re tr(i t P_mu,-nu)]  (:134-203)
//                  - (dt/a_z) (J+(x) - J-(x)) / sqrt2  on E_z
//   J+-stat'(x) = J+-stat(x) -+ (dt/a_z) [J+-(x) - U_z^dag(x-z) J+-(x-z) U_z(x-z)]          (:205-236)
//   X_mu        = exp(i c_mu E'_mu(x)),  c_mu = -(dt/2) a_mu^2 / a^3                         (:62-106)
//   U_h         = X U,  UOld = U_h                                                             (:108-128)
//   J+-'(x)     = J+-(x) -+ (dt/a_z) [U_h,z(x) J+-stat'(x+z) U_h,z^dag(x) - J+-stat'(x)]       (:238-269)
//   U'          = X U_h
//
// Every stage reads only time-t data in a small neighbourhood, so a site advances itself from
// time-t buffers into time-(t+dt) buffers: J+-stat'(x+z) is recomputed at x from time-t data (it
// needs U_z(x), J(x), J(x+z), J-stat(x+z)), and no grid-wide barrier is needed.  U and J are
// ping-pong buffers; E is updated in place (only site x reads E(x)).  X is computed once and
// applied twice; the reference recomputes it from identical inputs.  It is applied as U + (X - 1) U
// with X - 1 formed directly (expi_m1), which keeps rounding from biasing U away from SU(3).
// ------------------------------------------------------------------------------------------------
struct StepArgs {
    const double* __restrict__ U;  // time t links
    double* __restrict__ Un;       // time t+dt links
    double* __restrict__ E;        // in place
    const double* __restrict__ J;  // time t currents (J+, J-, J+stat, J-stat)
    double* __restrict__ Jn;       // time t+dt currents
    double* __restrict__ UOld;     // written when non-null
    Geom g;
    double gc[3];  // gamma * cE[mu],  gamma = -2 dt a^3  (SUNcDynamics.cpp:144-149)
    double cE[3];  // 1 / a_mu^2
    double cU[3];  // -(dt/2) a_mu^2 / a^3  (SUNcDynamics.cpp:67-69 with dTStep = dt/2)
    double dtz;    // dt / a_z
};

GHD void fill_step_coefficients(StepArgs& A, const double* a, double dt) {
    const double acube = a[0] * a[1] * a[2];
    const double gamma = -2.0 * dt * acube;
    for (int mu = 0; mu < 3; mu++) {
        A.cE[mu] = 1.0 / (a[mu] * a[mu]);
        A.gc[mu] = gamma * A.cE[mu];
        A.cU[mu] = -(0.5 * dt) * (a[mu] * a[mu]) / acube;
    }
    A.dtz = dt / a[2];
}

// r = Re tr(i t^a P) for the four-link products of SC/CALC/Plaquettes.cpp:83-103
// (SC/SUNc/AdvancedOperations.cpp:175-228):
// KIND 0: UUDD(v1,v2,v3,v4) = (v1 v2)(v4 v3)^dag ; 1: UDDU = (v1 v2^dag)(v3^dag v4);
// 2: DUUD = (v1^dag v2)(v3 v4^dag) ; 3: DDUU = (v1^dag v2^dag)(v3 v4)
template <int KIND>
GHD void plaq_mat(const double* __restrict__ U, long S, int m1, long s1, int m2, long s2, int m3, long s3, int m4,
                  long s4, M3& Pq) {
    M3 A, B, L1, L2;
    ldm(U + m1 * 18 * S, S, s1, A);
    ldm(U + m2 * 18 * S, S, s2, B);
    if (KIND == 0) mul_nn(A, B, L1);
    else if (KIND == 1) mul_nd(A, B, L1);
    else if (KIND == 2) mul_dn(A, B, L1);
    else mul_dd(A, B, L1);
    ldm(U + m3 * 18 * S, S, s3, A);
    ldm(U + m4 * 18 * S, S, s4, B);
    if (KIND == 0) {
        mul_nn(B, A, L2);
        mul_nd(L1, L2, Pq);
    } else if (KIND == 1) {
        mul_dn(A, B, L2);
        mul_nn(L1, L2, Pq);
    } else if (KIND == 2) {
        mul_nd(A, B, L2);
        mul_nn(L1, L2, Pq);
    } else {
        mul_nn(A, B, L2);
        mul_nn(L1, L2, Pq);
    }
}
template <int KIND>
GHD void plaq_tr(const double* __restrict__ U, long S, int m1, long s1, int m2, long s2, int m3, long s3, int m4,
                 long s4, double* r) {
    M3 Pq;
    plaq_mat<KIND>(U, S, m1, s1, m2, s2, m3, s3, m4, s4, Pq);
    retr_igen(Pq, r);
}

GHD void step_site(const StepArgs& A, long s) {
    const Geom g = A.g;
    const long S = g.S;
    const int x = (int)(s % g.nx);
    const long r_ = s / g.nx;
    const int y = (int)(r_ % g.ny);
    const int z = (int)(r_ / g.ny);
    const int xp = (x + 1 == g.nx) ? 0 : x + 1, xm = (x == 0) ? g.nx - 1 : x - 1;
    const int yp = (y + 1 == g.ny) ? 0 : y + 1, ym = (y == 0) ? g.ny - 1 : y - 1;
    const int zp = zup(g, z), zm = zdn(g, z);
    const long nx = g.nx, P = g.P;
    const long s0 = s;
    const long sxp = xp + nx * y + P * z, sxm = xm + nx * y + P * z;
    const long syp = x + nx * yp + P * z, sym = x + nx * ym + P * z;
    const long szp = x + nx * y + P * zp, szm = x + nx * y + P * zm;
    const long s_mxpy = xm + nx * yp + P * z, s_xpmy = xp + nx * ym + P * z;
    const long s_ympz = x + nx * ym + P * zp, s_ypmz = x + nx * yp + P * zm;
    const long s_xpmz = xp + nx * y + P * zm, s_mxpz = xm + nx * y + P * zp;
    const double* __restrict__ U = A.U;

    // ---------------- electric field (SC/DYNAMICS/SUNcDynamics.cpp:164-193) ----------------
    double p0[8], p1[8], p2[8];
    {
        double tA[8], tB[8];
        plaq_tr<0>(U, S, 0, s0, 1, sxp, 0, syp, 1, s0, tA);       // Uxy
        plaq_tr<2>(U, S, 1, sym, 0, sym, 1, s_xpmy, 0, s0, tB);   // UxMy
#pragma unroll
        for (int a = 0; a < 8; a++) p0[a] = A.cE[1] * (tA[a] - tB[a]);
        plaq_tr<1>(U, S, 1, s0, 0, s_mxpy, 1, sxm, 0, sxm, tB);   // UMxy
#pragma unroll
        for (int a = 0; a < 8; a++) p1[a] = -A.cE[0] * (tA[a] - tB[a]);
        plaq_tr<0>(U, S, 1, s0, 2, syp, 1, szp, 2, s0, tA);       // Uyz
        plaq_tr<2>(U, S, 2, szm, 1, szm, 2, s_ypmz, 1, s0, tB);   // UyMz
#pragma unroll
        for (int a = 0; a < 8; a++) p1[a] = A.cE[2] * (tA[a] - tB[a]) + p1[a];
        plaq_tr<1>(U, S, 2, s0, 1, s_ympz, 2, sym, 1, sym, tB);   // UMyz
#pragma unroll
        for (int a = 0; a < 8; a++) p2[a] = -A.cE[1] * (tA[a] - tB[a]);
        plaq_tr<0>(U, S, 2, s0, 0, szp, 2, sxp, 0, s0, tA);       // Uzx
        plaq_tr<1>(U, S, 0, s0, 2, s_xpmz, 0, szm, 2, szm, tB);   // UMzx
#pragma unroll
        for (int a = 0; a < 8; a++) p0[a] = p0[a] - A.cE[2] * (tA[a] - tB[a]);
        plaq_tr<2>(U, S, 0, sxm, 2, sxm, 0, s_mxpz, 2, s0, tB);   // UzMx
#pragma unroll
        for (int a = 0; a < 8; a++) p2[a] = A.cE[0] * (tA[a] - tB[a]) + p2[a];
    }
    double Ex[8], Ey[8], Ez[8], jp[8], jm[8];
    ldv(A.E + 0 * 8 * S, S, s0, Ex);
    ldv(A.E + 1 * 8 * S, S, s0, Ey);
    ldv(A.E + 2 * 8 * S, S, s0, Ez);
    ldv(A.J + JP * 8 * S, S, s0, jp);
    ldv(A.J + JM * 8 * S, S, s0, jm);
#pragma unroll
    for (int a = 0; a < 8; a++) {
        Ex[a] += A.gc[0] * p0[a];
        Ey[a] += A.gc[1] * p1[a];
        Ez[a] += A.gc[2] * p2[a];
    }
#pragma unroll
    for (int a = 0; a < 8; a++) Ez[a] += -A.dtz * ((jp[a] - jm[a]) / kSqrt2);
    stv(A.E + 0 * 8 * S, S, s0, Ex);
    stv(A.E + 1 * 8 * S, S, s0, Ey);
    stv(A.E + 2 * 8 * S, S, s0, Ez);

    // ---------------- J-stat advection at x and at x+z (SC/DYNAMICS/SUNcDynamics.cpp:205-236) ----------------
    double jps[8], jms[8];
    {
        M3 Uzm;
        double jpd[8], jmd[8], down[8];
        ldm(U + 2 * 18 * S, S, szm, Uzm);
        ldv(A.J + JP * 8 * S, S, szm, jpd);
        ldv(A.J + JM * 8 * S, S, szm, jmd);
        ldv(A.J + JPS * 8 * S, S, s0, jps);
        ldv(A.J + JMS * 8 * S, S, s0, jms);
        inv_adj_mul(Uzm, jpd, down);
#pragma unroll
        for (int a = 0; a < 8; a++) jps[a] += -A.dtz * (jp[a] - down[a]);
        inv_adj_mul(Uzm, jmd, down);
#pragma unroll
        for (int a = 0; a < 8; a++) jms[a] += A.dtz * (jm[a] - down[a]);
    }
    stv(A.Jn + JPS * 8 * S, S, s0, jps);
    stv(A.Jn + JMS * 8 * S, S, s0, jms);
    double jpsU[8], jmsU[8];  // J+-stat'(x+z)
    if (zp != z) {
        M3 Uz0;
        double jpu[8], jmu[8], down[8];
        ldm(U + 2 * 18 * S, S, s0, Uz0);
        ldv(A.J + JP * 8 * S, S, szp, jpu);
        ldv(A.J + JM * 8 * S, S, szp, jmu);
        ldv(A.J + JPS * 8 * S, S, szp, jpsU);
        ldv(A.J + JMS * 8 * S, S, szp, jmsU);
        inv_adj_mul(Uz0, jp, down);
#pragma unroll
        for (int a = 0; a < 8; a++) jpsU[a] += -A.dtz * (jpu[a] - down[a]);
        inv_adj_mul(Uz0, jm, down);
#pragma unroll
        for (int a = 0; a < 8; a++) jmsU[a] += A.dtz * (jmu[a] - down[a]);
    } else {
#pragma unroll
        for (int a = 0; a < 8; a++) {
            jpsU[a] = jps[a];
            jmsU[a] = jms[a];
        }
    }

    // ---------------- links: two half steps with one exponential (:62-106), UOld (:108-128) ----------------
#pragma unroll
    for (int mu = 0; mu < 3; mu++) {
        M3 D, Uo, Uh;
        expi_m1(A.cU[mu], mu == 0 ? Ex : (mu == 1 ? Ey : Ez), D);  // X - 1
        ldm(U + mu * 18 * S, S, s0, Uo);
        add_mul(D, Uo, Uh);  // X U
        if (A.UOld) stm(A.UOld + mu * 18 * S, S, s0, Uh);
        if (mu == 2) {
            // ---------------- J advection with U_z(t + dt/2) (:238-269) ----------------
            double up[8];
            adj_mul(Uh, jpsU, up);
#pragma unroll
            for (int a = 0; a < 8; a++) jp[a] += -A.dtz * (up[a] - jps[a]);
            adj_mul(Uh, jmsU, up);
#pragma unroll
            for (int a = 0; a < 8; a++) jm[a] += A.dtz * (up[a] - jms[a]);
            stv(A.Jn + JP * 8 * S, S, s0, jp);
            stv(A.Jn + JM * 8 * S, S, s0, jm);
        }
        add_mul(D, Uh, Uo);  // X U_h
        stm(A.Un + mu * 18 * S, S, s0, Uo);
    }
}

}  // namespace glasma
