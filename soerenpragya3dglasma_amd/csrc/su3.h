// Per-site SU(3) algebra for the MI355X glasma engine (FP64), usable on device and host.
//
// Conventions are the reference's: matrices are 3x3 complex, entry (i,j) at k = i + 3j
// (SC/SUNc/SU3/GroupOperations.cpp:13-15); algebra vectors alpha[8] in Gell-Mann order with
// generators t^a = lambda^a / 2 (SC/SUNc/SU3/AlgebraOperations.cpp:6-45).
// Real and imaginary parts are kept in separate arrays so that a matrix maps onto 18 scalar
// registers and onto 18 structure-of-arrays planes in HBM.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define GHD __host__ __device__ __forceinline__
#else
#define GHD inline
#endif

namespace glasma {

constexpr double kSqrt2 = 1.41421356237309504880168872420969808;
constexpr double kSqrt3 = 1.73205080756887729352744634150587237;
constexpr double kInvSqrt3 = 0.577350269189625764509148780501957456;
constexpr double kPi = 3.14159265358979323846264338327950288;

struct M3 {
    double re[9], im[9];
};

GHD constexpr int mi(int i, int j) { return i + 3 * j; }

GHD void set_unit(M3& A) {
#pragma unroll
    for (int k = 0; k < 9; k++) { A.re[k] = 0.0; A.im[k] = 0.0; }
    A.re[0] = A.re[4] = A.re[8] = 1.0;
}

// C = A * B
GHD void mul_nn(const M3& A, const M3& B, M3& C) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double r = 0.0, m = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                r = fma(A.re[mi(i, k)], B.re[mi(k, j)], r);
                r = fma(-A.im[mi(i, k)], B.im[mi(k, j)], r);
                m = fma(A.re[mi(i, k)], B.im[mi(k, j)], m);
                m = fma(A.im[mi(i, k)], B.re[mi(k, j)], m);
            }
            C.re[mi(i, j)] = r;
            C.im[mi(i, j)] = m;
        }
}
// C = A * B^dagger
GHD void mul_nd(const M3& A, const M3& B, M3& C) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double r = 0.0, m = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                // A(i,k) * conj(B(j,k))
                r = fma(A.re[mi(i, k)], B.re[mi(j, k)], r);
                r = fma(A.im[mi(i, k)], B.im[mi(j, k)], r);
                m = fma(A.im[mi(i, k)], B.re[mi(j, k)], m);
                m = fma(-A.re[mi(i, k)], B.im[mi(j, k)], m);
            }
            C.re[mi(i, j)] = r;
            C.im[mi(i, j)] = m;
        }
}
// C = A^dagger * B
GHD void mul_dn(const M3& A, const M3& B, M3& C) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double r = 0.0, m = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                // conj(A(k,i)) * B(k,j)
                r = fma(A.re[mi(k, i)], B.re[mi(k, j)], r);
                r = fma(A.im[mi(k, i)], B.im[mi(k, j)], r);
                m = fma(A.re[mi(k, i)], B.im[mi(k, j)], m);
                m = fma(-A.im[mi(k, i)], B.re[mi(k, j)], m);
            }
            C.re[mi(i, j)] = r;
            C.im[mi(i, j)] = m;
        }
}
// C = A^dagger * B^dagger = (B A)^dagger
GHD void mul_dd(const M3& A, const M3& B, M3& C) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double r = 0.0, m = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                // conj(A(k,i)) * conj(B(j,k))
                r = fma(A.re[mi(k, i)], B.re[mi(j, k)], r);
                r = fma(-A.im[mi(k, i)], B.im[mi(j, k)], r);
                m = fma(-A.re[mi(k, i)], B.im[mi(j, k)], m);
                m = fma(-A.im[mi(k, i)], B.re[mi(j, k)], m);
            }
            C.re[mi(i, j)] = r;
            C.im[mi(i, j)] = m;
        }
}

// Re tr(i t^a V)  (SC/SUNc/SU3/GroupOperations.cpp:93-103)
GHD void retr_igen(const M3& V, double* r) {
    r[0] = -0.5 * (V.im[mi(0, 1)] + V.im[mi(1, 0)]);
    r[1] = -0.5 * (V.re[mi(0, 1)] - V.re[mi(1, 0)]);
    r[2] = -0.5 * (V.im[mi(0, 0)] - V.im[mi(1, 1)]);
    r[3] = -0.5 * (V.im[mi(0, 2)] + V.im[mi(2, 0)]);
    r[4] = -0.5 * (V.re[mi(0, 2)] - V.re[mi(2, 0)]);
    r[5] = -0.5 * (V.im[mi(1, 2)] + V.im[mi(2, 1)]);
    r[6] = -0.5 * (V.re[mi(1, 2)] - V.re[mi(2, 1)]);
    r[7] = -0.5 * (V.im[mi(0, 0)] + V.im[mi(1, 1)] - 2.0 * V.im[mi(2, 2)]) * kInvSqrt3;
}

// Re tr(1 - V)
GHD double retr_one_minus(const M3& V) { return 3.0 - V.re[0] - V.re[4] - V.re[8]; }

// ||V V^dag - 1||_F
GHD double unitarity_norm(const M3& V) {
    M3 C;
    mul_nd(V, V, C);
    double n = 0.0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        double d = C.re[k] - ((k % 4 == 0) ? 1.0 : 0.0);
        n = fma(d, d, n);
        n = fma(C.im[k], C.im[k], n);
    }
    return sqrt(n);
}

// Hermitian matrix c * alpha^a t^a
GHD void herm_from_alg(double c, const double* al, M3& Q) {
    const double h = 0.5 * c;
    Q.re[mi(0, 0)] = h * (al[2] + al[7] * kInvSqrt3);
    Q.re[mi(1, 1)] = -h * (al[2] - al[7] * kInvSqrt3);
    Q.re[mi(2, 2)] = -c * al[7] * kInvSqrt3;
    Q.im[mi(0, 0)] = Q.im[mi(1, 1)] = Q.im[mi(2, 2)] = 0.0;
    Q.re[mi(1, 0)] = h * al[0]; Q.im[mi(1, 0)] = h * al[1];
    Q.re[mi(2, 0)] = h * al[3]; Q.im[mi(2, 0)] = h * al[4];
    Q.re[mi(2, 1)] = h * al[5]; Q.im[mi(2, 1)] = h * al[6];
    Q.re[mi(0, 1)] = Q.re[mi(1, 0)]; Q.im[mi(0, 1)] = -Q.im[mi(1, 0)];
    Q.re[mi(0, 2)] = Q.re[mi(2, 0)]; Q.im[mi(0, 2)] = -Q.im[mi(2, 0)];
    Q.re[mi(1, 2)] = Q.re[mi(2, 1)]; Q.im[mi(1, 2)] = -Q.im[mi(2, 1)];
}

// alpha^a = 2 tr(t^a H) of a Hermitian H, from its diagonal and lower triangle
GHD void alg_from_herm(const M3& H, double* al) {
    al[0] = 2.0 * H.re[mi(1, 0)];
    al[1] = 2.0 * H.im[mi(1, 0)];
    al[2] = H.re[mi(0, 0)] - H.re[mi(1, 1)];
    al[3] = 2.0 * H.re[mi(2, 0)];
    al[4] = 2.0 * H.im[mi(2, 0)];
    al[5] = 2.0 * H.re[mi(2, 1)];
    al[6] = 2.0 * H.im[mi(2, 1)];
    al[7] = (H.re[mi(0, 0)] + H.re[mi(1, 1)] - 2.0 * H.re[mi(2, 2)]) * kInvSqrt3;
}

// D = exp(i c alpha^a t^a) - 1.
//
// Q = c alpha^a t^a is Hermitian and traceless, so Q^3 = c1 Q + c0 with c1 = tr(Q^2)/2, c0 = det Q
// (Cayley-Hamilton), and every power is Q^k = a_k + b_k Q + c_k Q^2 with
// (a, b, c)_{k+1} = (c_k c0, a_k + c_k c1, b_k).  For small Q (c1 <= 1/16, i.e. eigenvalues below ~0.3)
// D = sum_{k>=1} (i^k / k!) Q^k is summed through these real coefficients (14 terms reach 1e-19 relative),
// so D carries an error relative to |D| rather than to 1: a link update U + D U then adds no
// systematic drift of U away from SU(3) (a bias of ~1e-16 per exponential accumulates visibly in
// plaquette sums over many steps).  Larger Q use the closed form of C. Morningstar and M. Peardon,
// Phys. Rev. D 69, 054501 (2004), eqs. (23)-(34): exp(iQ) = f0 + f1 Q + f2 Q^2, c0 >= 0 taken by
// f_j(-c0) = (-1)^j conj(f_j(c0)).
GHD void expi_m1(double c, const double* al, M3& D) {
    M3 Q;
    herm_from_alg(c, al, Q);
    M3 Q2;  // Hermitian: diagonal and lower triangle, upper by conjugation
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j <= i; j++) {
            double r = 0.0, m = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                r = fma(Q.re[mi(i, k)], Q.re[mi(k, j)], r);
                r = fma(-Q.im[mi(i, k)], Q.im[mi(k, j)], r);
                m = fma(Q.re[mi(i, k)], Q.im[mi(k, j)], m);
                m = fma(Q.im[mi(i, k)], Q.re[mi(k, j)], m);
            }
            Q2.re[mi(i, j)] = r;
            Q2.im[mi(i, j)] = (i == j) ? 0.0 : m;
            Q2.re[mi(j, i)] = r;
            Q2.im[mi(j, i)] = (i == j) ? 0.0 : -m;
        }
    const double c1 = 0.5 * (Q2.re[0] + Q2.re[4] + Q2.re[8]);
    if (!(c1 > 0.0)) {
#pragma unroll
        for (int k = 0; k < 9; k++) { D.re[k] = 0.0; D.im[k] = 0.0; }
        return;
    }
    // det Q for Hermitian Q: q00 q11 q22 + 2 Re(Q01 Q12 Q20) - q00|Q12|^2 - q11|Q20|^2 - q22|Q01|^2
    const double q00 = Q.re[0], q11 = Q.re[4], q22 = Q.re[8];
    const double a01r = Q.re[mi(0, 1)], a01i = Q.im[mi(0, 1)];
    const double a12r = Q.re[mi(1, 2)], a12i = Q.im[mi(1, 2)];
    const double a20r = Q.re[mi(2, 0)], a20i = Q.im[mi(2, 0)];
    const double t_r = a01r * a12r - a01i * a12i;
    const double t_i = a01r * a12i + a01i * a12r;
    const double re3 = t_r * a20r - t_i * a20i;
    double c0 = q00 * q11 * q22 + 2.0 * re3 - q00 * (a12r * a12r + a12i * a12i) - q11 * (a20r * a20r + a20i * a20i) -
                q22 * (a01r * a01r + a01i * a01i);
    double f0r, f0i, f1r, f1i, f2r, f2i;  // D = f0 + f1 Q + f2 Q^2
    if (c1 <= 0.0625) {
        // k = 1: i Q ; k = 2: -Q^2/2
        f0r = 0.0; f0i = 0.0;
        f1r = 0.0; f1i = 1.0;
        f2r = -0.5; f2i = 0.0;
        double ak = 0.0, bk = 0.0, ck = 1.0;  // Q^2
        double fact = 0.5;                    // 1/2!
#pragma unroll
        for (int k = 3; k <= 16; k++) {
            const double an = ck * c0, bn = fma(ck, c1, ak), cn = bk;
            ak = an; bk = bn; ck = cn;
            fact *= 1.0 / (double)k;
            const double ta = fact * ak, tb = fact * bk, tc = fact * ck;
            switch (k & 3) {  // i^k
                case 0: f0r += ta; f1r += tb; f2r += tc; break;
                case 1: f0i += ta; f1i += tb; f2i += tc; break;
                case 2: f0r -= ta; f1r -= tb; f2r -= tc; break;
                default: f0i -= ta; f1i -= tb; f2i -= tc; break;
            }
        }
    } else {
        const bool neg = c0 < 0.0;
        if (neg) c0 = -c0;
        const double s13 = sqrt(c1 * (1.0 / 3.0));
        const double c0max = 2.0 * s13 * s13 * s13;
        double ratio = (c0max > 0.0) ? c0 / c0max : 1.0;
        ratio = ratio > 1.0 ? 1.0 : ratio;
        const double theta = acos(ratio);
        double st, ct;
        sincos(theta * (1.0 / 3.0), &st, &ct);
        const double u = s13 * ct;
        const double w = sqrt(c1) * st;
        const double u2 = u * u, w2 = w * w;
        double sw, cw;
        sincos(w, &sw, &cw);
        double xi0;
        if (fabs(w) < 0.05) xi0 = 1.0 - w2 * (1.0 / 6.0) * (1.0 - w2 * (1.0 / 20.0) * (1.0 - w2 * (1.0 / 42.0)));
        else xi0 = sw / w;
        double su, cu;
        sincos(u, &su, &cu);
        const double emr = cu, emi = -su;                        // e^{-iu}
        const double e2r = cu * cu - su * su, e2i = 2.0 * su * cu;  // e^{2iu}
        const double b0r = 8.0 * u2 * cw, b0i = 2.0 * u * (3.0 * u2 + w2) * xi0;
        const double h0r = (u2 - w2) * e2r + (emr * b0r - emi * b0i);
        const double h0i = (u2 - w2) * e2i + (emr * b0i + emi * b0r);
        const double b1r = 2.0 * u * cw, b1i = -(3.0 * u2 - w2) * xi0;
        const double h1r = 2.0 * u * e2r - (emr * b1r - emi * b1i);
        const double h1i = 2.0 * u * e2i - (emr * b1i + emi * b1r);
        const double b2r = cw, b2i = 3.0 * u * xi0;
        const double h2r = e2r - (emr * b2r - emi * b2i);
        const double h2i = e2i - (emr * b2i + emi * b2r);
        const double den = 1.0 / (9.0 * u2 - w2);
        f0r = h0r * den; f0i = h0i * den;
        f1r = h1r * den; f1i = h1i * den;
        f2r = h2r * den; f2i = h2i * den;
        if (neg) { f0i = -f0i; f1r = -f1r; f2i = -f2i; }
        f0r -= 1.0;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {
        double xr = f1r * Q.re[k] - f1i * Q.im[k] + f2r * Q2.re[k] - f2i * Q2.im[k];
        double xi = f1r * Q.im[k] + f1i * Q.re[k] + f2r * Q2.im[k] + f2i * Q2.re[k];
        if (k % 4 == 0) { xr += f0r; xi += f0i; }
        D.re[k] = xr;
        D.im[k] = xi;
    }
}

// X = exp(i c alpha^a t^a)
GHD void expi(double c, const double* al, M3& X) {
    expi_m1(c, al, X);
    X.re[0] += 1.0;
    X.re[4] += 1.0;
    X.re[8] += 1.0;
}

// C = B + D B  (the product D B is formed first, then added once)
GHD void add_mul(const M3& D, const M3& B, M3& C) {
    M3 T;
    mul_nn(D, B, T);
#pragma unroll
    for (int k = 0; k < 9; k++) {
        C.re[k] = B.re[k] + T.re[k];
        C.im[k] = B.im[k] + T.im[k];
    }
}

// Sig += c * T C^dag
GHD void acc_mul_nd(double c, const M3& T, const M3& C, M3& Sig) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double r = 0.0, m = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                r = fma(T.re[mi(i, k)], C.re[mi(j, k)], r);
                r = fma(T.im[mi(i, k)], C.im[mi(j, k)], r);
                m = fma(T.im[mi(i, k)], C.re[mi(j, k)], m);
                m = fma(-T.re[mi(i, k)], C.im[mi(j, k)], m);
            }
            Sig.re[mi(i, j)] = fma(c, r, Sig.re[mi(i, j)]);
            Sig.im[mi(i, j)] = fma(c, m, Sig.im[mi(i, j)]);
        }
}
// Sig += c * T C
GHD void acc_mul_nn(double c, const M3& T, const M3& C, M3& Sig) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double r = 0.0, m = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                r = fma(T.re[mi(i, k)], C.re[mi(k, j)], r);
                r = fma(-T.im[mi(i, k)], C.im[mi(k, j)], r);
                m = fma(T.re[mi(i, k)], C.im[mi(k, j)], m);
                m = fma(T.im[mi(i, k)], C.re[mi(k, j)], m);
            }
            Sig.re[mi(i, j)] = fma(c, r, Sig.re[mi(i, j)]);
            Sig.im[mi(i, j)] = fma(c, m, Sig.im[mi(i, j)]);
        }
}
// r = Re tr(i t^a A B): only the entries that the projection reads
GHD void retr_igen_prod(const M3& A, const M3& B, double* r) {
    M3 P;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double re = 0.0, im = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                re = fma(A.re[mi(i, k)], B.re[mi(k, j)], re);
                re = fma(-A.im[mi(i, k)], B.im[mi(k, j)], re);
                im = fma(A.re[mi(i, k)], B.im[mi(k, j)], im);
                im = fma(A.im[mi(i, k)], B.re[mi(k, j)], im);
            }
            P.re[mi(i, j)] = re;
            P.im[mi(i, j)] = im;
        }
    retr_igen(P, r);
}

// components of V (alpha^b t^b) V^dagger  (meaning of SC/SUNc/SU2/AlgebraOperations.cpp:110-116)
GHD void adj_mul(const M3& V, const double* al, double* out) {
    M3 A, T, H;
    herm_from_alg(1.0, al, A);
    mul_nn(V, A, T);
    // H = T V^dag : diagonal (real) and lower triangle only
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j <= i; j++) {
            double r = 0.0, m = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                r = fma(T.re[mi(i, k)], V.re[mi(j, k)], r);
                r = fma(T.im[mi(i, k)], V.im[mi(j, k)], r);
                m = fma(T.im[mi(i, k)], V.re[mi(j, k)], m);
                m = fma(-T.re[mi(i, k)], V.im[mi(j, k)], m);
            }
            H.re[mi(i, j)] = r;
            H.im[mi(i, j)] = m;
        }
    alg_from_herm(H, out);
}

// components of V^dagger (alpha^b t^b) V  (meaning of SC/SUNc/SU2/AlgebraOperations.cpp:119-124)
GHD void inv_adj_mul(const M3& V, const double* al, double* out) {
    M3 A, T, H;
    herm_from_alg(1.0, al, A);
    mul_nn(A, V, T);
    // H = V^dag T
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j <= i; j++) {
            double r = 0.0, m = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                r = fma(V.re[mi(k, i)], T.re[mi(k, j)], r);
                r = fma(V.im[mi(k, i)], T.im[mi(k, j)], r);
                m = fma(V.re[mi(k, i)], T.im[mi(k, j)], m);
                m = fma(-V.im[mi(k, i)], T.re[mi(k, j)], m);
            }
            H.re[mi(i, j)] = r;
            H.im[mi(i, j)] = m;
        }
    alg_from_herm(H, out);
}

// ---------------------------------------------------------------------------------------------
// Hermitian 3x3 eigen-decomposition by cyclic complex Jacobi rotations (used by the logarithm,
// which is needed once per site in the initial conditions, never in the time step).
// On exit A = W diag(ev) W^dagger for the input A.
// ---------------------------------------------------------------------------------------------
GHD void jacobi_rotate(M3& A, M3& W, int p, int q) {
    const double zr = A.re[mi(p, q)], zi = A.im[mi(p, q)];
    const double r = sqrt(zr * zr + zi * zi);
    if (r == 0.0) return;
    const double phr = zr / r, phi = zi / r;
    const double th = (A.re[mi(q, q)] - A.re[mi(p, p)]) / (2.0 * r);
    const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
    const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
    // R = 1 except R_pp = R_qq = c, R_pq = s e^{i phi}, R_qp = -s e^{-i phi};  A <- R^dag A R, W <- W R
    double cr[3], ci[3], dr[3], di[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {  // columns p,q of A R
        const double apr = A.re[mi(i, p)], api = A.im[mi(i, p)];
        const double aqr = A.re[mi(i, q)], aqi = A.im[mi(i, q)];
        // (AR)_ip = c A_ip - s e^{-i phi} A_iq
        cr[i] = c * apr - s * (phr * aqr + phi * aqi);
        ci[i] = c * api - s * (phr * aqi - phi * aqr);
        // (AR)_iq = s e^{i phi} A_ip + c A_iq
        dr[i] = s * (phr * apr - phi * api) + c * aqr;
        di[i] = s * (phr * api + phi * apr) + c * aqi;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        A.re[mi(i, p)] = cr[i]; A.im[mi(i, p)] = ci[i];
        A.re[mi(i, q)] = dr[i]; A.im[mi(i, q)] = di[i];
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {  // rows p,q of R^dag (A R)
        const double apr = A.re[mi(p, j)], api = A.im[mi(p, j)];
        const double aqr = A.re[mi(q, j)], aqi = A.im[mi(q, j)];
        // row p: c A_pj - s e^{i phi} A_qj ; row q: s e^{-i phi} A_pj + c A_qj
        cr[j] = c * apr - s * (phr * aqr - phi * aqi);
        ci[j] = c * api - s * (phr * aqi + phi * aqr);
        dr[j] = s * (phr * apr + phi * api) + c * aqr;
        di[j] = s * (phr * api - phi * apr) + c * aqi;
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        A.re[mi(p, j)] = cr[j]; A.im[mi(p, j)] = ci[j];
        A.re[mi(q, j)] = dr[j]; A.im[mi(q, j)] = di[j];
    }
    A.re[mi(p, q)] = A.im[mi(p, q)] = A.re[mi(q, p)] = A.im[mi(q, p)] = 0.0;
    A.im[mi(p, p)] = A.im[mi(q, q)] = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {  // W <- W R
        const double wpr = W.re[mi(i, p)], wpi = W.im[mi(i, p)];
        const double wqr = W.re[mi(i, q)], wqi = W.im[mi(i, q)];
        W.re[mi(i, p)] = c * wpr - s * (phr * wqr + phi * wqi);
        W.im[mi(i, p)] = c * wpi - s * (phr * wqi - phi * wqr);
        W.re[mi(i, q)] = s * (phr * wpr - phi * wpi) + c * wqr;
        W.im[mi(i, q)] = s * (phr * wpi + phi * wpr) + c * wqi;
    }
}

GHD void jacobi_herm(M3& A, M3& W) {
    set_unit(W);
    for (int sweep = 0; sweep < 40; sweep++) {
        double off = 0.0, dia = 0.0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const double n2 = A.re[k] * A.re[k] + A.im[k] * A.im[k];
            if (k % 4 == 0) dia += n2; else off += n2;
        }
        if (off == 0.0 || off <= 1e-34 * dia) break;
        jacobi_rotate(A, W, 0, 1);
        jacobi_rotate(A, W, 0, 2);
        jacobi_rotate(A, W, 1, 2);
    }
}

// complex determinant and inverse (adjugate / det) of a 3x3 matrix
GHD void det3(const M3& A, double& dr, double& di) {
    dr = 0.0;
    di = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        // A_0j * (A_1j1 A_2j2 - A_1j2 A_2j1)
        const double mr = A.re[mi(1, j1)] * A.re[mi(2, j2)] - A.im[mi(1, j1)] * A.im[mi(2, j2)] -
                          (A.re[mi(1, j2)] * A.re[mi(2, j1)] - A.im[mi(1, j2)] * A.im[mi(2, j1)]);
        const double mim = A.re[mi(1, j1)] * A.im[mi(2, j2)] + A.im[mi(1, j1)] * A.re[mi(2, j2)] -
                           (A.re[mi(1, j2)] * A.im[mi(2, j1)] + A.im[mi(1, j2)] * A.re[mi(2, j1)]);
        dr += A.re[mi(0, j)] * mr - A.im[mi(0, j)] * mim;
        di += A.re[mi(0, j)] * mim + A.im[mi(0, j)] * mr;
    }
}
GHD void inv3(const M3& A, M3& B) {
    double dr, di;
    det3(A, dr, di);
    const double n = dr * dr + di * di;
    const double ir = dr / n, ii = -di / n;  // 1/det
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            // B_ij = cofactor(j, i) / det
            const int r0 = (j + 1) % 3, r1 = (j + 2) % 3, c0 = (i + 1) % 3, c1 = (i + 2) % 3;
            const double cr = A.re[mi(r0, c0)] * A.re[mi(r1, c1)] - A.im[mi(r0, c0)] * A.im[mi(r1, c1)] -
                              (A.re[mi(r0, c1)] * A.re[mi(r1, c0)] - A.im[mi(r0, c1)] * A.im[mi(r1, c0)]);
            const double ci = A.re[mi(r0, c0)] * A.im[mi(r1, c1)] + A.im[mi(r0, c0)] * A.re[mi(r1, c1)] -
                              (A.re[mi(r0, c1)] * A.im[mi(r1, c0)] + A.im[mi(r0, c1)] * A.re[mi(r1, c0)]);
            B.re[mi(i, j)] = cr * ir - ci * ii;
            B.im[mi(i, j)] = cr * ii + ci * ir;
        }
}

// Principal traceless logarithm: returns A^a with exp(i C A^a t^a) = U (the meaning of the SU(2)
// MatrixILog, SC/SUNc/SU2/AlgebraOperations.cpp:92-107).  Through the Cayley transform: Y = i^k U with
// k in {0,1,2,3} maximising |det(1 + Y)| (no eigenvalue of Y near -1; one of the four always keeps
// |det| >= 0.44), C = i (1 - Y)(1 + Y)^-1 is Hermitian with eigenvalues tan(psi_j / 2), monotonic in
// Y's eigenphases psi_j, so distinct phases stay distinct in C and its Jacobi eigenvectors are accurate
// even near the identity.  Phases psi_j - k pi/2 wrapped to (-pi, pi]; when they sum to +-2 pi the
// largest (smallest) is moved by 2 pi.
GHD void ilog(double C, const M3& U, double* out) {
    int kb = 0;
    double db = -1.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        M3 A;
#pragma unroll
        for (int e = 0; e < 9; e++) {
            // i^k U
            const double r = U.re[e], m = U.im[e];
            A.re[e] = k == 0 ? r : (k == 1 ? -m : (k == 2 ? -r : m));
            A.im[e] = k == 0 ? m : (k == 1 ? r : (k == 2 ? -m : -r));
        }
        A.re[0] += 1.0;
        A.re[4] += 1.0;
        A.re[8] += 1.0;
        double dr, di;
        det3(A, dr, di);
        const double d = dr * dr + di * di;
        if (d > db) { db = d; kb = k; }
    }
    M3 Y, Ap, Am, Ai, Cm, T;
#pragma unroll
    for (int e = 0; e < 9; e++) {
        const double r = U.re[e], m = U.im[e];
        Y.re[e] = kb == 0 ? r : (kb == 1 ? -m : (kb == 2 ? -r : m));
        Y.im[e] = kb == 0 ? m : (kb == 1 ? r : (kb == 2 ? -m : -r));
        Ap.re[e] = Y.re[e];
        Ap.im[e] = Y.im[e];
        Am.re[e] = -Y.re[e];
        Am.im[e] = -Y.im[e];
    }
#pragma unroll
    for (int d = 0; d < 9; d += 4) {
        Ap.re[d] += 1.0;
        Am.re[d] += 1.0;
    }
    inv3(Ap, Ai);
    M3 P;
    mul_nn(Am, Ai, P);
    // Cm = i P, made exactly Hermitian
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            // (i P)_ij = -Im P_ij + i Re P_ij ; Hermitian part (X + X^dag)/2
            const double xr = -P.im[mi(i, j)], xi = P.re[mi(i, j)];
            const double yr = -P.im[mi(j, i)], yi = P.re[mi(j, i)];
            Cm.re[mi(i, j)] = 0.5 * (xr + yr);
            Cm.im[mi(i, j)] = 0.5 * (xi - yi);
        }
    jacobi_herm(Cm, T);
    // phases of the diagonal of T^dag Y T
    M3 YT, D;
    mul_nn(Y, T, YT);
    mul_dn(T, YT, D);
    double ph[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        double p = atan2(D.im[mi(j, j)], D.re[mi(j, j)]) - kb * (0.5 * kPi);
        if (p > kPi) p -= 2.0 * kPi;
        if (p <= -kPi) p += 2.0 * kPi;
        ph[j] = p;
    }
    const double s = ph[0] + ph[1] + ph[2];
    if (s > kPi) {
        int m = (ph[0] >= ph[1] && ph[0] >= ph[2]) ? 0 : (ph[1] >= ph[2] ? 1 : 2);
        ph[m] -= 2.0 * kPi;
    } else if (s < -kPi) {
        int m = (ph[0] <= ph[1] && ph[0] <= ph[2]) ? 0 : (ph[1] <= ph[2] ? 1 : 2);
        ph[m] += 2.0 * kPi;
    }
    // H = T diag(ph) T^dag (diagonal and lower triangle)
    M3 H;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k <= i; k++) {
            double r = 0.0, m = 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                // T_ij ph_j conj(T_kj)
                const double pr = T.re[mi(i, j)] * T.re[mi(k, j)] + T.im[mi(i, j)] * T.im[mi(k, j)];
                const double pi_ = T.im[mi(i, j)] * T.re[mi(k, j)] - T.re[mi(i, j)] * T.im[mi(k, j)];
                r = fma(ph[j], pr, r);
                m = fma(ph[j], pi_, m);
            }
            H.re[mi(i, k)] = r;
            H.im[mi(i, k)] = m;
        }
    alg_from_herm(H, out);
    const double ic = 1.0 / C;
#pragma unroll
    for (int a = 0; a < 8; a++) out[a] *= ic;
}

}  // namespace glasma
