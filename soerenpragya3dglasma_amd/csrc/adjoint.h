// Reverse-mode derivative (vector-Jacobian product) of the leapfrog step and of the slab energy, per site.
//
// Conventions: a complex matrix M has the cotangent Mbar with dL = <Mbar, dM> = Re tr(Mbar^dag dM); an
// algebra vector v (8 reals) has the cotangent vbar with dL = vbar . dv.  Links are treated as general
// complex matrices (every formula of the step is defined for those), so the products below are exact
// adjoints and finite differences along arbitrary directions check them.
//
// Forward step (site.h): E' = E + gc Re tr(i t U Sigma) - dtz (J+ - J-)/sqrt2 [z],
// J+-stat' = J+-stat -+ dtz (J+- - Ad(U_z^dag(x-z)) J+-(x-z)), X = exp(i cU E' t), U_h = X U,
// J+-' = J+- -+ dtz (Ad(U_h,z) J+-stat'(x+z) - J+-stat'), U' = X U_h.  The adjoint runs in three passes:
//   A (site-local): cotangents of X, U_h, U (direct part), E (through the exponential), J+-stat', J+-;
//   B (gather from x+z): the J+-(x-z), U_z(x-z) terms of the J-stat' update;
//   C (staples): the staple terms of the E update, either scattered into the links of neighbouring sites
//     (step_vjp_staples, with an accumulation functor) or, without atomics, as C1 (step_vjp_sigbar) +
//     C2 (step_vjp_gather).
// The device runs two passes: A with B folded in (step_vjp_local<I, true>), then C2 with C1's own-link term
// (step_vjp_gather<I, true>), which forms each Sigbar it needs from Ebar; the host build also runs the separate passes
// and the scatter form, which the tests compare with it.
// The exponential's adjoint: with A = i cU E' t, dexp_A[H] = e^A sum_k (-ad_A)^k H/(k+1)!, so
// Zbar = sum_k ad_A^k (X^dag Xbar)/(k+1)! and Ebar^a += cU Re tr(i t^a Zbar^dag).
#pragma once

#include "site.h"

namespace glasma {

GHD void m3_zero(M3& A) {
#pragma unroll
    for (int k = 0; k < 9; k++) { A.re[k] = 0.0; A.im[k] = 0.0; }
}
GHD void m3_axpy(double c, const M3& A, M3& C) {
#pragma unroll
    for (int k = 0; k < 9; k++) { C.re[k] = fma(c, A.re[k], C.re[k]); C.im[k] = fma(c, A.im[k], C.im[k]); }
}
// Mbar = -i c H(v)  (cotangent of M from r = c Re tr(i t^a M) with cotangent v)
GHD void m_from_retr_cot(double c, const double* v, M3& Mbar) {
    M3 H;
    herm_from_alg(c, v, H);
#pragma unroll
    for (int k = 0; k < 9; k++) { Mbar.re[k] = H.im[k]; Mbar.im[k] = -H.re[k]; }
}
// r = Re tr(i t^a Z^dag)
GHD void retr_igen_dag(const M3& Z, double* r) {
    M3 D;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) { D.re[mi(i, j)] = Z.re[mi(j, i)]; D.im[mi(i, j)] = -Z.im[mi(j, i)]; }
    retr_igen(D, r);
}
// Vbar += s * 4 L V R  for Hermitian L = H(l), R = H(r): the V-cotangent of adj_mul (L = H(obar), R = H(alpha))
// and of inv_adj_mul (L = H(alpha), R = H(obar)) with output cotangent obar
GHD void acc_lvr(double s, const double* l, const M3& V, const double* r, M3& Vbar) {
    M3 L, R, T, P;
    herm_from_alg(1.0, l, L);
    herm_from_alg(1.0, r, R);
    mul_nn(L, V, T);
    mul_nn(T, R, P);
    m3_axpy(4.0 * s, P, Vbar);
}

// z = sum_k ad_A^k(y) / (k+1)! for A = i c e^a t^a, on coordinates in the basis i t^a of su(3) (y is overwritten).  In
// that basis ad_A acts as y'^r = -c f^{pqr} e^p y^q with the structure constants of su(3), [t^a, t^b] = i f^{abc} t^c:
// each term of the series is 54 multiply-adds on eight numbers instead of two 3 x 3 complex products.
GHD void ad_series(double c, const double* e, double* y, double* z) {
    // the nonzero f^{pqr}, p < q < r (0-based, t^a = lambda^a / 2 in retr_igen's order), and their values
    constexpr int tri[9][3] = {{0, 1, 2}, {0, 3, 6}, {0, 4, 5}, {1, 3, 5}, {1, 4, 6},
                               {2, 3, 4}, {2, 5, 6}, {3, 4, 7}, {5, 6, 7}};
    constexpr double h3 = 0.86602540378443864676;  // sqrt(3) / 2
    constexpr double fv[9] = {1.0, 0.5, -0.5, 0.5, 0.5, 0.5, -0.5, h3, h3};
#pragma unroll
    for (int a = 0; a < 8; a++) z[a] = y[a];
    // w[t] = -c f^{pqr} (e_p, e_q, e_r) of triple t: y'_r += w_p y_q - w_q y_p, y'_p += w_q y_r - w_r y_q,
    // y'_q += w_r y_p - w_p y_r
    double w[9][3];
#pragma unroll
    for (int t = 0; t < 9; t++)
#pragma unroll
        for (int k = 0; k < 3; k++) w[t][k] = -c * fv[t] * e[tri[t][k]];
    double fact = 1.0;
    for (int it = 1; it <= 40; it++) {
        double yn[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < 9; t++) {
            const int p = tri[t][0], q = tri[t][1], r = tri[t][2];
            yn[r] = fma(w[t][0], y[q], fma(-w[t][1], y[p], yn[r]));
            yn[p] = fma(w[t][1], y[r], fma(-w[t][2], y[q], yn[p]));
            yn[q] = fma(w[t][2], y[p], fma(-w[t][0], y[r], yn[q]));
        }
        double mx = 0.0;
#pragma unroll
        for (int a = 0; a < 8; a++) {
            y[a] = yn[a];
            mx = fmax(mx, fabs(yn[a]));
        }
        fact /= (double)(it + 1);
#pragma unroll
        for (int a = 0; a < 8; a++) z[a] = fma(fact, y[a], z[a]);
        if (mx * fact < 1e-19) break;
    }
}

// ebar += cU Re tr(i t^a Zbar^dag), Zbar = sum_k ad_A^k(X^dag Xbar)/(k+1)!, A = i cU e^a t^a.
// Only the traceless anti-Hermitian part of Y = X^dag Xbar reaches ebar (Re tr(i t^a H) = 0 for Hermitian H, and
// the trace commutes with A), and ad_A keeps that part: in the basis i t^b it is y^b = -2 Re tr(i t^b Y).  Then
// Zbar's part is z = ad_series(y) and ebar^a += (cU / 2) z^a.
GHD void dexp_adjoint(double cU, const double* e, const M3& X, const M3& Xbar, double* ebar) {
    double y[8], z[8];
    {
        M3 Y;
        mul_dn(X, Xbar, Y);
        retr_igen(Y, y);
    }
#pragma unroll
    for (int a = 0; a < 8; a++) y[a] *= -2.0;
    ad_series(cU, e, y, z);
#pragma unroll
    for (int a = 0; a < 8; a++) ebar[a] = fma(0.5 * cU, z[a], ebar[a]);
}

// cotangent buffers of one backward step (tiled layouts like the fields)
struct StepAdjArgs {
    const double* __restrict__ U;    // links at time t
    const double* __restrict__ Uh;   // half-step links U_h (the step's UOld)
    const double* __restrict__ E1;   // E at t+dt (= E')
    const double* __restrict__ J;    // currents at t
    const double* __restrict__ J1;   // currents at t+dt
    const double* __restrict__ Ub1;  // cotangent of U(t+dt)
    double* __restrict__ Ub;         // cotangent of U(t): written by pass A, accumulated by B and C
    double* __restrict__ Eb;         // in: cotangent of E(t+dt); out: cotangent of E(t) (in place)
    const double* __restrict__ Jb1;  // cotangent of J(t+dt)
    double* __restrict__ Jb;         // cotangent of J(t)
    Geom g;
    double gc[3], cE[3], cU[3], dtz;
};

// the site whose J+-stat' site s reads as "x+z" (jstat_new_up): above, or s itself on the open top
template <class I>
GHD I up_site(const Geom& g, I s, int z) {
    const int zp = zup(g, z);
    return zp != z ? (I)((long)s + g.P * (long)(zp - z)) : s;
}
// the site whose J+-, U_z site s reads as "x-z" (jstat_new): below, or s itself on the open bottom
template <class I>
GHD I down_site(const Geom& g, I s, int z) {
    const int zm = zdn(g, z);
    return (I)((long)s + g.P * (long)(zm - z));
}

// Cotangents of J+-stat'(s) (layer z): their own J+- terms, and the J advection of the site that reads J+-stat'(s)
// as "x+z" (the site below, and s itself on the open top layer).  jb1p, jb1m: the cotangents of J+-(s) at t+dt.
template <class I>
GHD void jstat_cot(const StepAdjArgs& A, I s, int z, const double* jb1p, const double* jb1m, double* jsbp,
                   double* jsbm) {
    const Geom& g = A.g;
    const double k = A.dtz;
    double t8[8];
    ldv<I>(A.Jb1, kJPlanes, JPS * 8, s, jsbp);
    ldv<I>(A.Jb1, kJPlanes, JMS * 8, s, jsbm);
#pragma unroll
    for (int a = 0; a < 8; a++) {
        jsbp[a] += k * jb1p[a];
        jsbm[a] -= k * jb1m[a];
    }
    const int zm = zdn(g, z);
    if (zm != z) {
        const I sb = (I)((long)s + g.P * (long)(zm - z));
        if (up_site<I>(g, sb, zm) == s) {
            M3 Uhb;
            double jb[8];
            ldm<I>(A.Uh, kUPlanes, 36, sb, Uhb);
            ldv<I>(A.Jb1, kJPlanes, JP * 8, sb, jb);
            inv_adj_mul(Uhb, jb, t8);
            for (int a = 0; a < 8; a++) jsbp[a] -= k * t8[a];
            ldv<I>(A.Jb1, kJPlanes, JM * 8, sb, jb);
            inv_adj_mul(Uhb, jb, t8);
            for (int a = 0; a < 8; a++) jsbm[a] += k * t8[a];
        }
    }
    if (up_site<I>(g, s, z) == s) {
        M3 Uhs;
        ldm<I>(A.Uh, kUPlanes, 36, s, Uhs);
        inv_adj_mul(Uhs, jb1p, t8);
        for (int a = 0; a < 8; a++) jsbp[a] -= k * t8[a];
        inv_adj_mul(Uhs, jb1m, t8);
        for (int a = 0; a < 8; a++) jsbm[a] += k * t8[a];
    }
}

// Pass B's terms at site s, added to the cotangents of J+-(s) and U_z(s): s is read as "x-z" by the site above (its
// J+-stat' cotangents jsb_up) and by itself on the open bottom layer (jsb_own).
GHD void down_terms(const StepAdjArgs& A, const M3& Uz, const double* jp, const double* jm, const double* jsbp,
                    const double* jsbm, double* jbp, double* jbm, M3& Ubz) {
    const double k = A.dtz;
    double t8[8];
    adj_mul(Uz, jsbp, t8);
    for (int a = 0; a < 8; a++) jbp[a] += k * t8[a];
    acc_lvr(k, jp, Uz, jsbp, Ubz);
    adj_mul(Uz, jsbm, t8);
    for (int a = 0; a < 8; a++) jbm[a] -= k * t8[a];
    acc_lvr(-k, jm, Uz, jsbm, Ubz);
}

// Pass A at site s.  FUSED (the device form): pass B's terms at s are added here, with the J+-stat' cotangents of
// the site above formed again from the inputs (jstat_cot).  The sums are those of the separate passes, in the same
// order.
template <class I, bool FUSED = false>
GHD void step_vjp_local(const StepAdjArgs& A, I s) {
    const Geom& g = A.g;
    int x, y, z;
    site_coords<I>(g, s, x, y, z);
    const double k = A.dtz;
    double jbp[8], jbm[8], jsbp[8], jsbm[8];
    ldv<I>(A.Jb1, kJPlanes, JP * 8, s, jbp);
    ldv<I>(A.Jb1, kJPlanes, JM * 8, s, jbm);
    // J+-stat'(s) is read as "x+z" by the site below (and by s itself on the open top layer)
    jstat_cot<I>(A, s, z, jbp, jbm, jsbp, jsbm);
    // U_h,z cotangent from the J advection at s
    M3 Uhbz;
    m3_zero(Uhbz);
    {
        M3 Uhz;
        double su[8];
        ldm<I>(A.Uh, kUPlanes, 36, s, Uhz);
        const I su_s = up_site<I>(g, s, z);
        ldv<I>(A.J1, kJPlanes, JPS * 8, su_s, su);
        acc_lvr(-k, jbp, Uhz, su, Uhbz);
        ldv<I>(A.J1, kJPlanes, JMS * 8, su_s, su);
        acc_lvr(k, jbm, Uhz, su, Uhbz);
    }
    // the three links: U' = X U_h, U_h = X U, X = exp(i cU E' t)
    double ebz[8];
    M3 Ubz;  // FUSED: the cotangent of U_z(s), stored after pass B's terms
    for (int mu = 0; mu < 3; mu++) {
        M3 Ub1, Uh, U0, X, Xb, Uhbar, T;
        double e1[8], eb[8];
        ldm<I>(A.Ub1, kUPlanes, 18 * mu, s, Ub1);
        ldm<I>(A.Uh, kUPlanes, 18 * mu, s, Uh);
        if (!FUSED) ldm<I>(A.U, kUPlanes, 18 * mu, s, U0);
        ldv<I>(A.E1, kEPlanes, 8 * mu, s, e1);
        ldv<I>(A.Eb, kEPlanes, 8 * mu, s, eb);
        expi(A.cU[mu], e1, X);
        // FUSED: U = X^dag U_h (X unitary to round-off) instead of reading the time-t links: this pass is bound by
        // its HBM reads
        if (FUSED) mul_dn(X, Uh, U0);
        mul_nd(Ub1, Uh, Xb);
        mul_dn(X, Ub1, Uhbar);
        if (mu == 2) m3_axpy(1.0, Uhbz, Uhbar);
        mul_nd(Uhbar, U0, T);
        m3_axpy(1.0, T, Xb);
        mul_dn(X, Uhbar, T);
        if (FUSED && mu == 2) Ubz = T;
        else stm<I>(A.Ub, kUPlanes, 18 * mu, s, T);
        dexp_adjoint(A.cU[mu], e1, X, Xb, eb);
        stv<I>(A.Eb, kEPlanes, 8 * mu, s, eb);
        if (mu == 2)
            for (int a = 0; a < 8; a++) ebz[a] = eb[a];
        if (FUSED) sched_fence();
    }
    const double ks = k / kSqrt2;
#pragma unroll
    for (int a = 0; a < 8; a++) {
        jbp[a] += -k * jsbp[a] - ks * ebz[a];
        jbm[a] += k * jsbm[a] + ks * ebz[a];
    }
    if (FUSED) {
        // pass B: s is read as "x-z" by the site above su (have_up) and by itself on the open bottom layer
        // (have_own); the J terms first, then the U_z terms, each in pass B's order
        const int zp = zup(g, z);
        const I su = (I)((long)s + g.P * (long)(zp - z));
        const bool have_up = zp != z && down_site<I>(g, su, zp) == s;
        const bool have_own = down_site<I>(g, s, z) == s;
        double usp[8], usm[8], t8[8];
        if (have_up) {
            double b1p[8], b1m[8];
            ldv<I>(A.Jb1, kJPlanes, JP * 8, su, b1p);
            ldv<I>(A.Jb1, kJPlanes, JM * 8, su, b1m);
            jstat_cot<I>(A, su, zp, b1p, b1m, usp, usm);
        }
        sched_fence();
        M3 Uz;
        ldm<I>(A.U, kUPlanes, 36, s, Uz);
        for (int src = 0; src < 2; src++) {
            if (src == 0 ? !have_up : !have_own) continue;
            const double* qp = src == 0 ? usp : jsbp;
            const double* qm = src == 0 ? usm : jsbm;
            adj_mul(Uz, qp, t8);
            for (int a = 0; a < 8; a++) jbp[a] += k * t8[a];
            adj_mul(Uz, qm, t8);
            for (int a = 0; a < 8; a++) jbm[a] -= k * t8[a];
        }
        stv<I>(A.Jb, kJPlanes, JP * 8, s, jbp);
        stv<I>(A.Jb, kJPlanes, JM * 8, s, jbm);
        sched_fence();
        double jp[8], jm[8];
        ldv<I>(A.J, kJPlanes, JP * 8, s, jp);
        ldv<I>(A.J, kJPlanes, JM * 8, s, jm);
        for (int src = 0; src < 2; src++) {
            if (src == 0 ? !have_up : !have_own) continue;
            acc_lvr(k, jp, Uz, src == 0 ? usp : jsbp, Ubz);
            acc_lvr(-k, jm, Uz, src == 0 ? usm : jsbm, Ubz);
        }
        stm<I>(A.Ub, kUPlanes, 36, s, Ubz);
        stv<I>(A.Jb, kJPlanes, JPS * 8, s, jsbp);
        stv<I>(A.Jb, kJPlanes, JMS * 8, s, jsbm);
        return;
    }
    stv<I>(A.Jb, kJPlanes, JP * 8, s, jbp);
    stv<I>(A.Jb, kJPlanes, JM * 8, s, jbm);
    stv<I>(A.Jb, kJPlanes, JPS * 8, s, jsbp);
    stv<I>(A.Jb, kJPlanes, JMS * 8, s, jsbm);
}

// Pass B at site s: s is read as "x-z" by the site above (and by itself on the open bottom layer).
template <class I>
GHD void step_vjp_down(const StepAdjArgs& A, I s) {
    const Geom& g = A.g;
    int x, y, z;
    site_coords<I>(g, s, x, y, z);
    M3 Uz, Ubz;
    ldm<I>(A.U, kUPlanes, 36, s, Uz);
    ldm<I>(A.Ub, kUPlanes, 36, s, Ubz);
    double jp[8], jm[8], jbp[8], jbm[8];
    ldv<I>(A.J, kJPlanes, JP * 8, s, jp);
    ldv<I>(A.J, kJPlanes, JM * 8, s, jm);
    ldv<I>(A.Jb, kJPlanes, JP * 8, s, jbp);
    ldv<I>(A.Jb, kJPlanes, JM * 8, s, jbm);
    for (int src = 0; src < 2; src++) {
        I sx;
        if (src == 0) {
            const int zp = zup(g, z);
            if (zp == z) continue;
            sx = (I)((long)s + g.P * (long)(zp - z));
            if (down_site<I>(g, sx, zp) != s) continue;
        } else {
            if (down_site<I>(g, s, z) != s) continue;
            sx = s;
        }
        double jsbp[8], jsbm[8];
        ldv<I>(A.Jb, kJPlanes, JPS * 8, sx, jsbp);
        ldv<I>(A.Jb, kJPlanes, JMS * 8, sx, jsbm);
        down_terms(A, Uz, jp, jm, jsbp, jsbm, jbp, jbm, Ubz);
    }
    stv<I>(A.Jb, kJPlanes, JP * 8, s, jbp);
    stv<I>(A.Jb, kJPlanes, JM * 8, s, jbm);
    stm<I>(A.Ub, kUPlanes, 36, s, Ubz);
}

// Pass C at site s: staple terms of the E update, scattered with ADD (atomic on the device).
template <class I, class ADD>
GHD void step_vjp_staples(const StepAdjArgs& A, I s, ADD add) {
    const Geom& g = A.g;
    int x, y, z;
    site_coords<I>(g, s, x, y, z);
    for (int mu = 0; mu < 3; mu++) {
        double eb[8];
        ldv<I>(A.Eb, kEPlanes, 8 * mu, s, eb);
        M3 Mb, Umu, Sig, Sigb, T;
        m_from_retr_cot(A.gc[mu], eb, Mb);
        ldm<I>(A.U, kUPlanes, 18 * mu, s, Umu);
        m3_zero(Sig);
        for (int q = 1; q <= 2; q++) {
            const int nu = (mu + q) % 3;
            M3 P1, P2, Tq;
            ldm<I>(A.U, kUPlanes, 18 * nu, nbr<I>(g, x, y, z, mu, 1, nu, 0), P1);
            ldm<I>(A.U, kUPlanes, 18 * mu, nbr<I>(g, x, y, z, nu, 1, mu, 0), P2);
            mul_nd(P1, P2, Tq);
            ldm<I>(A.U, kUPlanes, 18 * nu, s, P1);
            acc_mul_nd(A.cE[nu], Tq, P1, Sig);
            const I sm = nbr<I>(g, x, y, z, nu, -1, mu, 0);
            ldm<I>(A.U, kUPlanes, 18 * nu, nbr<I>(g, x, y, z, mu, 1, nu, -1), P1);
            ldm<I>(A.U, kUPlanes, 18 * mu, sm, P2);
            mul_dd(P1, P2, Tq);
            ldm<I>(A.U, kUPlanes, 18 * nu, sm, P1);
            acc_mul_nn(A.cE[nu], Tq, P1, Sig);
        }
        mul_nd(Mb, Sig, T);
        add(mu, s, T);
        mul_dn(Umu, Mb, Sigb);
        for (int q = 1; q <= 2; q++) {
            const int nu = (mu + q) % 3;
            M3 Sb, P, Q, R, T1, T2;
            Sb = Sigb;
            for (int kk = 0; kk < 9; kk++) { Sb.re[kk] *= A.cE[nu]; Sb.im[kk] *= A.cE[nu]; }
            // forward staple P Q^dag R^dag: P = U_nu(x+mu), Q = U_mu(x+nu), R = U_nu(x)
            const I sP = nbr<I>(g, x, y, z, mu, 1, nu, 0), sQ = nbr<I>(g, x, y, z, nu, 1, mu, 0);
            ldm<I>(A.U, kUPlanes, 18 * nu, sP, P);
            ldm<I>(A.U, kUPlanes, 18 * mu, sQ, Q);
            ldm<I>(A.U, kUPlanes, 18 * nu, s, R);
            mul_nn(Sb, R, T1);
            mul_nn(T1, Q, T2);
            add(nu, sP, T2);  // Pbar = Sb R Q
            mul_dd(R, Sb, T1);
            mul_nn(T1, P, T2);
            add(mu, sQ, T2);  // Qbar = R^dag Sb^dag P
            mul_dn(Sb, P, T1);
            mul_nd(T1, Q, T2);
            add(nu, s, T2);  // Rbar = Sb^dag P Q^dag
            // backward staple P^dag Q^dag R: P = U_nu(x+mu-nu), Q = U_mu(x-nu), R = U_nu(x-nu)
            const I sP2 = nbr<I>(g, x, y, z, mu, 1, nu, -1), sm = nbr<I>(g, x, y, z, nu, -1, mu, 0);
            ldm<I>(A.U, kUPlanes, 18 * nu, sP2, P);
            ldm<I>(A.U, kUPlanes, 18 * mu, sm, Q);
            ldm<I>(A.U, kUPlanes, 18 * nu, sm, R);
            mul_dn(Q, R, T1);
            mul_nd(T1, Sb, T2);
            add(nu, sP2, T2);  // Pbar = Q^dag R Sb^dag
            mul_nd(R, Sb, T1);
            mul_nd(T1, P, T2);
            add(mu, sm, T2);  // Qbar = R Sb^dag P^dag
            mul_nn(Q, P, T1);
            mul_nn(T1, Sb, T2);
            add(nu, sm, T2);  // Rbar = Q P Sb
        }
    }
}

// ---- pass C without atomics: C1 stores Sigbar_mu(x) = U_mu(x)^dag Mbar_mu(x) and adds the own-link term
// Mbar Sigma^dag at x; C2 gathers, for each link (y, lam), the staple terms of every site x whose E update
// reads it.  The sources of a target are the preimages of the forward neighbour map, so the clamped
// neighbours of open z boundaries contribute exactly as in the scatter form.
// The own-link term of link MU at s, T = Mbar_mu Sigma_mu^dag (Sigma_mu: the staple sum of the forward step).
template <class I, int MU>
GHD void own_link_term(const StepAdjArgs& A, I s, int x, int y, int z, M3& T) {
    const Geom& g = A.g;
    constexpr int mu = MU;
    double eb[8];
    ldv<I>(A.Eb, kEPlanes, 8 * mu, s, eb);
    M3 Mb, Sig;
    m_from_retr_cot(A.gc[mu], eb, Mb);
    m3_zero(Sig);
    for (int q = 1; q <= 2; q++) {
        const int nu = (mu + q) % 3;
        M3 P1, P2, Tq;
        ldm<I>(A.U, kUPlanes, 18 * nu, nbr<I>(g, x, y, z, mu, 1, nu, 0), P1);
        ldm<I>(A.U, kUPlanes, 18 * mu, nbr<I>(g, x, y, z, nu, 1, mu, 0), P2);
        mul_nd(P1, P2, Tq);
        ldm<I>(A.U, kUPlanes, 18 * nu, s, P1);
        acc_mul_nd(A.cE[nu], Tq, P1, Sig);
        const I sm = nbr<I>(g, x, y, z, nu, -1, mu, 0);
        ldm<I>(A.U, kUPlanes, 18 * nu, nbr<I>(g, x, y, z, mu, 1, nu, -1), P1);
        ldm<I>(A.U, kUPlanes, 18 * mu, sm, P2);
        mul_dd(P1, P2, Tq);
        ldm<I>(A.U, kUPlanes, 18 * nu, sm, P1);
        acc_mul_nn(A.cE[nu], Tq, P1, Sig);
    }
    mul_nd(Mb, Sig, T);
}

template <class I, int MU>
GHD void step_vjp_sigbar_link(const StepAdjArgs& A, I s, double* __restrict__ Sb) {
    int x, y, z;
    site_coords<I>(A.g, s, x, y, z);
    constexpr int mu = MU;
    double eb[8];
    ldv<I>(A.Eb, kEPlanes, 8 * mu, s, eb);
    M3 Mb, Umu, T, Ub;
    m_from_retr_cot(A.gc[mu], eb, Mb);
    ldm<I>(A.U, kUPlanes, 18 * mu, s, Umu);
    mul_dn(Umu, Mb, T);
    stm<I>(Sb, kUPlanes, 18 * mu, s, T);
    own_link_term<I, MU>(A, s, x, y, z, T);
    ldm<I>(A.Ub, kUPlanes, 18 * mu, s, Ub);
    m3_axpy(1.0, T, Ub);
    stm<I>(A.Ub, kUPlanes, 18 * mu, s, Ub);
}

template <class I>
GHD void step_vjp_sigbar(const StepAdjArgs& A, I s, double* __restrict__ Sb) {
    step_vjp_sigbar_link<I, 0>(A, s, Sb);
    step_vjp_sigbar_link<I, 1>(A, s, Sb);
    step_vjp_sigbar_link<I, 2>(A, s, Sb);
}

// source z layers of a target layer zy under a z shift dz of the forward neighbour map (up to two with
// open boundaries: the regular one and the clamped layer itself)
GHD int z_sources(const Geom& g, int zy, int dz, int* zs) {
    if (dz == 0) {
        zs[0] = zy;
        return 1;
    }
    int n = 0;
    const int c0 = dz > 0 ? (zy > 0 ? zy - 1 : g.nz - 1) : (zy + 1 < g.nz ? zy + 1 : 0);
    const int cand[2] = {c0, zy};
    for (int k = 0; k < 2; k++) {
        const int c = cand[k];
        const int t = dz > 0 ? zup(g, c) : zdn(g, c);
        if (t == zy && (n == 0 || zs[0] != c)) zs[n++] = c;
    }
    return n;
}

// Mbar_mu(s) = -i gc_mu H(Ebar_mu(s)), the cotangent of U_mu Sigma_mu; Sigbar_mu = U_mu^dag Mbar_mu
template <class I>
GHD void mbar_at(const StepAdjArgs& A, int mu, I s, M3& Mb) {
    double eb[8];
    ldv<I>(A.Eb, kEPlanes, 8 * mu, s, eb);
    m_from_retr_cot(A.gc[mu], eb, Mb);
}
template <class I>
GHD void sigbar_at(const StepAdjArgs& A, int mu, I s, M3& Sbar) {
    M3 U, Mb;
    ldm<I>(A.U, kUPlanes, 18 * mu, s, U);
    mbar_at<I>(A, mu, s, Mb);
    mul_dn(U, Mb, Sbar);
}
// C += c W^dag
GHD void m3_axpy_dag(double c, const M3& W, M3& C) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            C.re[mi(i, j)] = fma(c, W.re[mi(j, i)], C.re[mi(i, j)]);
            C.im[mi(i, j)] = fma(-c, W.im[mi(j, i)], C.im[mi(i, j)]);
        }
}

// The gather with the own-link term on a layer where every staple term has its source at the plain neighbour position
// (periodic z, or 1 <= z <= nz - 2 with open z; z_sources gives exactly one source there).  For each transverse
// direction d the six links a = U_lam(x+d), b = U_lam(x-d), c = U_d(x+lam), e = U_d(x-d+lam), f = U_d(x-d),
// h = U_d(x) are all that the six staple terms of d and d's part of the staple sum Sigma_lam(x) read, and each
// Sigbar_mu(x') = U_mu(x')^dag Mbar_mu(x') a term reads has its link among them, so the terms need only Mbar, formed from
// Ebar (8 numbers) at the source site:
//   E_d staples (coefficient cE[lam]): Sbar_d(x)^dag a c^dag = Mbar_d(x)^dag W, (h a) Sbar_d(x+lam) = W Mbar_d(x+lam),
//     Sbar_d(x-d) b e = f^dag Mbar_d(x-d) V, (f^dag b) Sbar_d(x-d+lam)^dag = (f^dag b) Mbar_d(x-d+lam)^dag e;
//   E_lam staples (cE[d]): f^dag Sbar_lam(x-d)^dag e = f^dag Mbar_lam(x-d)^dag V, h Sbar_lam(x+d)^dag c^dag =
//     h Mbar_lam(x+d)^dag a c^dag;
//   Sigma_lam += cE[d] (c (h a)^dag + e^dag (f^dag b)^dag) = cE[d] (W^dag + e^dag (f^dag b)^dag);
// with W = h a c^dag and V = b e, and the own-link term Mbar_lam(x) Sigma_lam^dag added last.  The two terms at x-d
// share one product: Mbar is linear in Ebar, so cE[lam] Mbar_d + cE[d] Mbar_lam^dag = Mbar of the combined vector
// (Mbar^dag = -Mbar).  Each link reads 12 links and 13 Ebar (against 25 links and 24 Sigbar in the plain form).
template <class I, int LAM>
GHD void gather_link_interior(const StepAdjArgs& A, I s, int x, int y, int z) {
    const Geom& g = A.g;
    constexpr int lam = LAM;
    const double cl = A.cE[lam];
    M3 acc, Sig;
    ldm<I>(A.Ub, kUPlanes, 18 * lam, s, acc);
    m3_zero(Sig);
    // not unrolled: unrolled, the compiler interleaves the two directions and runs out of registers
    for (int k = 1; k <= 2; k++) {
        const int d = (lam + k) % 3;
        const double cd = A.cE[d];
        const I s_pd = nbr<I>(g, x, y, z, d, 1, lam, 0), s_md = nbr<I>(g, x, y, z, d, -1, lam, 0);
        const I s_pl = nbr<I>(g, x, y, z, lam, 1, d, 0), s_mdpl = nbr<I>(g, x, y, z, lam, 1, d, -1);
        {
            M3 a, h, c, W, T1, Mb;
            ldm<I>(A.U, kUPlanes, 18 * lam, s_pd, a);
            ldm<I>(A.U, kUPlanes, 18 * d, s, h);
            ldm<I>(A.U, kUPlanes, 18 * d, s_pl, c);
            mul_nn(h, a, T1);
            mul_nd(T1, c, W);
            m3_axpy_dag(cd, W, Sig);
            mbar_at<I>(A, d, s, Mb);
            mul_dn(Mb, W, T1);
            m3_axpy(cl, T1, acc);
            mbar_at<I>(A, d, s_pl, Mb);
            acc_mul_nn(cl, W, Mb, acc);
            mul_nd(a, c, T1);
            mbar_at<I>(A, lam, s_pd, Mb);
            mul_dn(Mb, T1, W);
            acc_mul_nn(cd, h, W, acc);
        }
        sched_fence();
        {
            M3 b, f, e, V, T1, Mb;
            ldm<I>(A.U, kUPlanes, 18 * lam, s_md, b);
            ldm<I>(A.U, kUPlanes, 18 * d, s_md, f);
            ldm<I>(A.U, kUPlanes, 18 * d, s_mdpl, e);
            mul_nn(b, e, V);
            {
                double ed[8], el[8], w[8];
                ldv<I>(A.Eb, kEPlanes, 8 * d, s_md, ed);
                ldv<I>(A.Eb, kEPlanes, 8 * lam, s_md, el);
                for (int q = 0; q < 8; q++) w[q] = cl * A.gc[d] * ed[q] - cd * A.gc[lam] * el[q];
                m_from_retr_cot(1.0, w, Mb);
            }
            mul_nn(Mb, V, T1);
            mul_dn(f, T1, V);
            m3_axpy(1.0, V, acc);
            mul_dn(f, b, T1);
            mul_dd(e, T1, V);
            m3_axpy(cd, V, Sig);
            mbar_at<I>(A, d, s_mdpl, Mb);
            mul_nd(T1, Mb, V);
            acc_mul_nn(cl, V, e, acc);
        }
        sched_fence();
    }
    M3 Mb, T;
    mbar_at<I>(A, lam, s, Mb);
    mul_nd(Mb, Sig, T);
    m3_axpy(1.0, T, acc);
    stm<I>(A.Ub, kUPlanes, 18 * lam, s, acc);
}

// the layers where gather_link_interior applies
GHD bool gather_interior(const Geom& g, int z) { return g.bc || (z >= 1 && z + 2 <= g.nz); }

// The gather for any layer, each staple term from its sources (z_sources).  OWN (the device form, after the fused pass
// A): the own-link term of C1 (own_link_term) is added first, as C1 would have added it to Ub.
template <class I, int LAM, bool OWN>
GHD void gather_link_general(const StepAdjArgs& A, I s, int x, int y, int z) {
    const Geom& g = A.g;
    constexpr int lam = LAM;
    {
        M3 acc, Sbar, P, Q, R, T1;
        ldm<I>(A.Ub, kUPlanes, 18 * lam, s, acc);
        if (OWN) {
            own_link_term<I, LAM>(A, s, x, y, z, T1);
            m3_axpy(1.0, T1, acc);
        }
        for (int k = 1; k <= 2; k++) {
            const int mu = (lam + k) % 3;  // staples of E_mu in which U_lam is a nu = lam link
            const double c = A.cE[lam];
            // forward R: x = y; Rbar = Sb^dag P Q^dag, P = U_lam(x+mu), Q = U_mu(x+lam)
            sigbar_at<I>(A, mu, s, Sbar);
            ldm<I>(A.U, kUPlanes, 18 * lam, nbr<I>(g, x, y, z, mu, 1, lam, 0), P);
            ldm<I>(A.U, kUPlanes, 18 * mu, nbr<I>(g, x, y, z, lam, 1, mu, 0), Q);
            mul_dn(Sbar, P, T1);
            acc_mul_nd(c, T1, Q, acc);
            int zs[2];
            // forward P: y = x + mu; Pbar = Sb R Q, R = U_lam(x), Q = U_mu(x+lam)
            {
                const int nzs = z_sources(g, z, mu == 2 ? 1 : 0, zs);
                const int xs = mu == 0 ? x - 1 : x, ys = mu == 1 ? y - 1 : y;
                for (int iz = 0; iz < nzs; iz++) {
                    const int xx = xwrap(xs, g.nx), yy = xwrap(ys, g.ny), zz = zs[iz];
                    const I sx = (I)sidx(g, xx, yy, zz);
                    sigbar_at<I>(A, mu, sx, Sbar);
                    ldm<I>(A.U, kUPlanes, 18 * lam, sx, R);
                    ldm<I>(A.U, kUPlanes, 18 * mu, nbr<I>(g, xx, yy, zz, lam, 1, mu, 0), Q);
                    mul_nn(Sbar, R, T1);
                    acc_mul_nn(c, T1, Q, acc);
                }
            }
            // backward P: y = x + mu - lam; Pbar = Q^dag R Sb^dag, Q = U_mu(x-lam), R = U_lam(x-lam)
            {
                const int dz = (mu == 2 ? 1 : 0) + (lam == 2 ? -1 : 0);
                const int nzs = z_sources(g, z, dz, zs);
                const int xs = x - (mu == 0 ? 1 : 0) + (lam == 0 ? 1 : 0), ys = y - (mu == 1 ? 1 : 0) + (lam == 1 ? 1 : 0);
                for (int iz = 0; iz < nzs; iz++) {
                    const int xx = xwrap(xs, g.nx), yy = xwrap(ys, g.ny), zz = zs[iz];
                    const I sx = (I)sidx(g, xx, yy, zz);
                    const I sm = nbr<I>(g, xx, yy, zz, lam, -1, mu, 0);
                    sigbar_at<I>(A, mu, sx, Sbar);
                    ldm<I>(A.U, kUPlanes, 18 * mu, sm, Q);
                    ldm<I>(A.U, kUPlanes, 18 * lam, sm, R);
                    mul_dn(Q, R, T1);
                    acc_mul_nd(c, T1, Sbar, acc);
                }
            }
            // backward R: y = x - lam; Rbar = Q P Sb, Q = U_mu(x-lam), P = U_lam(x+mu-lam)
            {
                const int nzs = z_sources(g, z, lam == 2 ? -1 : 0, zs);
                const int xs = lam == 0 ? x + 1 : x, ys = lam == 1 ? y + 1 : y;
                for (int iz = 0; iz < nzs; iz++) {
                    const int xx = xwrap(xs, g.nx), yy = xwrap(ys, g.ny), zz = zs[iz];
                    const I sx = (I)sidx(g, xx, yy, zz);
                    sigbar_at<I>(A, mu, sx, Sbar);
                    ldm<I>(A.U, kUPlanes, 18 * mu, nbr<I>(g, xx, yy, zz, lam, -1, mu, 0), Q);
                    ldm<I>(A.U, kUPlanes, 18 * lam, nbr<I>(g, xx, yy, zz, mu, 1, lam, -1), P);
                    mul_nn(Q, P, T1);
                    acc_mul_nn(c, T1, Sbar, acc);
                }
            }
        }
        for (int k = 1; k <= 2; k++) {
            const int nu = (lam + k) % 3;  // staples of E_lam in which U_lam(x +- nu) is the Q link
            const double c = A.cE[nu];
            int zs[2];
            // forward Q: y = x + nu; Qbar = R^dag Sb^dag P, R = U_nu(x), P = U_nu(x+lam)
            {
                const int nzs = z_sources(g, z, nu == 2 ? 1 : 0, zs);
                const int xs = nu == 0 ? x - 1 : x, ys = nu == 1 ? y - 1 : y;
                for (int iz = 0; iz < nzs; iz++) {
                    const int xx = xwrap(xs, g.nx), yy = xwrap(ys, g.ny), zz = zs[iz];
                    const I sx = (I)sidx(g, xx, yy, zz);
                    sigbar_at<I>(A, lam, sx, Sbar);
                    ldm<I>(A.U, kUPlanes, 18 * nu, sx, R);
                    ldm<I>(A.U, kUPlanes, 18 * nu, nbr<I>(g, xx, yy, zz, lam, 1, nu, 0), P);
                    mul_dd(R, Sbar, T1);
                    acc_mul_nn(c, T1, P, acc);
                }
            }
            // backward Q: y = x - nu; Qbar = R Sb^dag P^dag, R = U_nu(x-nu), P = U_nu(x+lam-nu)
            {
                const int nzs = z_sources(g, z, nu == 2 ? -1 : 0, zs);
                const int xs = nu == 0 ? x + 1 : x, ys = nu == 1 ? y + 1 : y;
                for (int iz = 0; iz < nzs; iz++) {
                    const int xx = xwrap(xs, g.nx), yy = xwrap(ys, g.ny), zz = zs[iz];
                    const I sx = (I)sidx(g, xx, yy, zz);
                    sigbar_at<I>(A, lam, sx, Sbar);
                    ldm<I>(A.U, kUPlanes, 18 * nu, nbr<I>(g, xx, yy, zz, nu, -1, lam, 0), R);
                    ldm<I>(A.U, kUPlanes, 18 * nu, nbr<I>(g, xx, yy, zz, lam, 1, nu, -1), P);
                    mul_nd(R, Sbar, T1);
                    acc_mul_nd(c, T1, P, acc);
                }
            }
        }
        stm<I>(A.Ub, kUPlanes, 18 * lam, s, acc);
    }
}

// OWN: the device form (interior layers by gather_link_interior, the others by gather_link_general with the own-link
// term); else the C2 pass of the separate form
template <class I, int LAM, bool OWN = false>
GHD void step_vjp_gather_link(const StepAdjArgs& A, I s) {
    int x, y, z;
    site_coords<I>(A.g, s, x, y, z);
    if (OWN && gather_interior(A.g, z)) gather_link_interior<I, LAM>(A, s, x, y, z);
    else gather_link_general<I, LAM, OWN>(A, s, x, y, z);
}

template <class I, bool OWN = false>
GHD void step_vjp_gather(const StepAdjArgs& A, I s) {
    step_vjp_gather_link<I, 0, OWN>(A, s);
    step_vjp_gather_link<I, 1, OWN>(A, s);
    step_vjp_gather_link<I, 2, OWN>(A, s);
}

// Slab energy 1/2 sum_{|z - zc| < half} sum_mu c_mu (E_mu^2 + 4 Re tr(1 - P_mu)), c_mu = (a_mu/a^3)^2,
// P_0 = P_yz, P_1 = P_zx, P_2 = P_xy rooted at x (the local form of Observables::Bulk's T00,
// SC/OBSERVABLES/BulkObservables.cpp:22-97): value at site s and, with ADD, its cotangents
// (Ebar written at s, links scattered).
struct SlabArgs {
    const double* __restrict__ U;
    const double* __restrict__ E;
    double* __restrict__ Eb;
    Geom g;
    double c[3];
    int zc, half;
};
GHD bool in_slab(const SlabArgs& A, int z) { return (z - A.zc) < A.half && (A.zc - z) < A.half; }

template <class I>
GHD double slab_energy_site(const SlabArgs& A, I s) {
    const Geom& g = A.g;
    int x, y, z;
    site_coords<I>(g, s, x, y, z);
    if (!in_slab(A, z)) return 0.0;
    double v = 0.0;
    for (int mu = 0; mu < 3; mu++) {
        double e[8];
        ldv<I>(A.E, kEPlanes, 8 * mu, s, e);
        double e2 = 0.0;
        for (int a = 0; a < 8; a++) e2 = fma(e[a], e[a], e2);
        const int j = (mu + 1) % 3, kk = (mu + 2) % 3;
        M3 Pq;
        plaq_mat<I>(A.U, 0, j, s, kk, nbr<I>(g, x, y, z, j, 1, kk, 0), j, nbr<I>(g, x, y, z, kk, 1, j, 0), kk, s, Pq);
        v += 0.5 * A.c[mu] * (e2 + 4.0 * retr_one_minus(Pq));
    }
    return v;
}

template <class I, class ADD>
GHD void slab_energy_vjp_site(const SlabArgs& A, I s, double scale, ADD add) {
    const Geom& g = A.g;
    int x, y, z;
    site_coords<I>(g, s, x, y, z);
    const bool in = in_slab(A, z);
    for (int mu = 0; mu < 3; mu++) {
        double e[8];
        ldv<I>(A.E, kEPlanes, 8 * mu, s, e);
        for (int a = 0; a < 8; a++) e[a] = in ? scale * A.c[mu] * e[a] : 0.0;
        stv<I>(A.Eb, kEPlanes, 8 * mu, s, e);
        if (!in) continue;
        // P = a b c^dag d^dag: a = U_j(x), b = U_k(x+j), c = U_j(x+k), d = U_k(x); Pbar = -2 c_mu scale 1
        const int j = (mu + 1) % 3, kk = (mu + 2) % 3;
        const I sb = nbr<I>(g, x, y, z, j, 1, kk, 0), sc = nbr<I>(g, x, y, z, kk, 1, j, 0);
        M3 Ma, Mb, Mc, Md, T1, T2, T3;
        ldm<I>(A.U, kUPlanes, 18 * j, s, Ma);
        ldm<I>(A.U, kUPlanes, 18 * kk, sb, Mb);
        ldm<I>(A.U, kUPlanes, 18 * j, sc, Mc);
        ldm<I>(A.U, kUPlanes, 18 * kk, s, Md);
        const double w = -2.0 * A.c[mu] * scale;
        // abar = w d c b^dag
        mul_nn(Md, Mc, T1);
        mul_nd(T1, Mb, T2);
        for (int q = 0; q < 9; q++) { T2.re[q] *= w; T2.im[q] *= w; }
        add(j, s, T2);
        // bbar = w a^dag d c
        mul_dn(Ma, Md, T1);
        mul_nn(T1, Mc, T2);
        for (int q = 0; q < 9; q++) { T2.re[q] *= w; T2.im[q] *= w; }
        add(kk, sb, T2);
        // cbar = w d^dag a b
        mul_nn(Ma, Mb, T3);
        mul_dn(Md, T3, T2);
        for (int q = 0; q < 9; q++) { T2.re[q] *= w; T2.im[q] *= w; }
        add(j, sc, T2);
        // dbar = w a b c^dag
        mul_nd(T3, Mc, T2);
        for (int q = 0; q < 9; q++) { T2.re[q] *= w; T2.im[q] *= w; }
        add(kk, s, T2);
    }
}

}  // namespace glasma
