// Device field layout and the per-site bodies of the glasma engine, written once for device and host.
//
// The HIP kernels call these functions for their site; the host build (hostcheck.cpp, compiled with
// g++) runs the very same code so that the product's arithmetic can also be exercised on a CPU.
#pragma once

#include "su3.h"

namespace glasma {

// ------------------------------------------------------------------------------------------------
// Device layout: tiles of 64 consecutive sites (s = x + Nx*(y + Ny*z)).  A field with NP scalar
// planes per site (NP even) groups the planes in pairs (2q, 2q+1); inside a tile, pair q of the 64 sites
// is stored contiguously with the two members of a pair adjacent:
//        plane p of site s at  ((s / 64) * NP/2 + p/2) * 128 + (s % 64) * 2 + p % 2.
// A lane then loads a pair (the real and imaginary part of a matrix entry, or two colour components)
// with ONE 16-byte access, a wavefront issues one aligned 1 KiB access per pair, and all planes of a
// tile are contiguous (a site's 3 links: 27.6 KB per tile), so a wavefront touches a few pages.  The
// number of vector-memory instructions, which bounds this kernel, is half that of one plane per
// access.  The site count is padded to a multiple of 64.
//   links    NP = 54: plane mu*18 + 2k + c, k = i + 3j (column-major), c = 0 real / 1 imaginary
//   E field  NP = 24: plane mu*8 + a
//   currents NP = 32: plane w*8 + a, w = 0 J+, 1 J-, 2 J+stat, 3 J-stat
//   one matrix per site (Wilson lines, gauge transformations): NP = 18
// ------------------------------------------------------------------------------------------------
constexpr int kTile = 64;
constexpr int kUPlanes = 54, kEPlanes = 24, kJPlanes = 32, kGPlanes = 18;
enum { JP = 0, JM = 1, JPS = 2, JMS = 3 };

GHD long pad_sites(long S) { return (S + kTile - 1) / kTile * kTile; }
template <class I>
GHD long lidx(int NP, int p, I s) {
    return ((long)(s / (I)kTile) * (NP >> 1) + (p >> 1)) * (2 * kTile) + 2 * (long)(s % (I)kTile) + (p & 1);
}

#if defined(__clang__)
typedef double d2_t __attribute__((ext_vector_type(2)));
#else
struct d2_t {
    double x, y;
};
#endif

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

// Address of (even) plane p0 of site s.  On the device the address is passed through an empty asm:
// each access then starts from an opaque base, so identical loads in different stages of a kernel are
// not merged (merging keeps whole matrices live across stages and exhausts the register file); the
// pair offsets k*1024 bytes are folded into the instructions' immediate offsets.
template <class I>
GHD const double* site_ptr(const double* base, int NP, int p0, I s) {
    const double* q = base + lidx<I>(NP, p0, s);
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(q));
#endif
    return q;
}
template <class I>
GHD double* site_ptr(double* base, int NP, int p0, I s) {
    double* q = base + lidx<I>(NP, p0, s);
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(q));
#endif
    return q;
}
// keeps the compiler from moving instructions across this point (device only): bounds how much of the next
// stage's loads it hoists, and so the live registers
GHD void sched_fence() {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_sched_barrier(0);
#endif
}
// streamed outputs are written non-temporally so that they do not evict neighbour data from L2 / MALL
GHD void st_nt(double* p, double v0, double v1) {
#if defined(__HIP_DEVICE_COMPILE__) && !defined(GLASMA_NO_NT_STORES)
    d2_t v;
    v.x = v0;
    v.y = v1;
    __builtin_nontemporal_store(v, (__attribute__((address_space(1))) d2_t*)p);
#else
    p[0] = v0;
    p[1] = v1;
#endif
}
GHD d2_t ld2(const double* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const __attribute__((address_space(1))) d2_t*)p;
#else
    d2_t v;
    v.x = p[0];
    v.y = p[1];
    return v;
#endif
}

// non-temporal load, for data read once per step (E): keeps L2 for the neighbour links, which are
// re-read by the tiles around a site (82 against 86 GB of reads per step at 256 x 256 x 1024)
GHD d2_t ld2_nt(const double* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_nontemporal_load((const __attribute__((address_space(1))) d2_t*)p);
#else
    return ld2(p);
#endif
}
template <class I>
GHD void ldv_nt(const double* __restrict__ base, int NP, int p0, I s, double* v) {
    const double* q = site_ptr<I>(base, NP, p0, s);
#pragma unroll
    for (int a = 0; a < 8; a += 2) {
        const d2_t w = ld2_nt(q + (a >> 1) * 2 * kTile);
        v[a] = w.x;
        v[a + 1] = w.y;
    }
}
template <class I>
GHD void ldm(const double* __restrict__ base, int NP, int p0, I s, M3& A) {
    const double* q = site_ptr<I>(base, NP, p0, s);
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const d2_t v = ld2(q + k * 2 * kTile);
        A.re[k] = v.x;
        A.im[k] = v.y;
    }
}
template <class I>
GHD void stm(double* __restrict__ base, int NP, int p0, I s, const M3& A) {
    double* q = site_ptr<I>(base, NP, p0, s);
#pragma unroll
    for (int k = 0; k < 9; k++) st_nt(q + k * 2 * kTile, A.re[k], A.im[k]);
}
template <class I>
GHD void ldv(const double* __restrict__ base, int NP, int p0, I s, double* v) {
    const double* q = site_ptr<I>(base, NP, p0, s);
#pragma unroll
    for (int a = 0; a < 8; a += 2) {
        const d2_t w = ld2(q + (a >> 1) * 2 * kTile);
        v[a] = w.x;
        v[a + 1] = w.y;
    }
}
template <class I>
GHD void stv(double* __restrict__ base, int NP, int p0, I s, const double* v) {
    double* q = site_ptr<I>(base, NP, p0, s);
#pragma unroll
    for (int a = 0; a < 8; a += 2) st_nt(q + (a >> 1) * 2 * kTile, v[a], v[a + 1]);
}

template <class I>
GHD void site_coords(const Geom& g, I s, int& x, int& y, int& z) {
    x = (int)(s % (I)g.nx);
    const I r_ = s / (I)g.nx;
    y = (int)(r_ % (I)g.ny);
    z = (int)(r_ / (I)g.ny);
}

// site index of (x,y,z) moved by s1 along d1 and s2 along d2 (0 = no move); x, y periodic, z with the
// reference's ghost semantics
template <class I>
GHD I nbr(const Geom& g, int x, int y, int z, int d1, int s1, int d2, int s2) {
    int xx = x + (d1 == 0 ? s1 : 0) + (d2 == 0 ? s2 : 0);
    int yy = y + (d1 == 1 ? s1 : 0) + (d2 == 1 ? s2 : 0);
    const int dz = (d1 == 2 ? s1 : 0) + (d2 == 2 ? s2 : 0);
    xx = xwrap(xx, g.nx);
    yy = xwrap(yy, g.ny);
    const int zz = dz > 0 ? zup(g, z) : (dz < 0 ? zdn(g, z) : z);
    return (I)xx + (I)g.nx * ((I)yy + (I)g.ny * (I)zz);
}

// four-link products of SC/CALC/Plaquettes.cpp:83-103 (SC/SUNc/AdvancedOperations.cpp:175-228):
// KIND 0: UUDD(v1,v2,v3,v4) = (v1 v2)(v4 v3)^dag ; 1: UDDU = (v1 v2^dag)(v3^dag v4);
// 2: DUUD = (v1^dag v2)(v3 v4^dag) ; 3: DDUU = (v1^dag v2^dag)(v3 v4).  U has 54 planes, link mu at
// plane 18*mu.
template <class I>
GHD void plaq_mat(const double* __restrict__ U, int kind, int m1, I s1, int m2, I s2, int m3, I s3, int m4, I s4,
                  M3& Pq) {
    M3 A, B, L1, L2;
    ldm<I>(U, kUPlanes, 18 * m1, s1, A);
    ldm<I>(U, kUPlanes, 18 * m2, s2, B);
    if (kind == 0) mul_nn(A, B, L1);
    else if (kind == 1) mul_nd(A, B, L1);
    else if (kind == 2) mul_dn(A, B, L1);
    else mul_dd(A, B, L1);
    ldm<I>(U, kUPlanes, 18 * m3, s3, A);
    ldm<I>(U, kUPlanes, 18 * m4, s4, B);
    if (kind == 0) {
        mul_nn(B, A, L2);
        mul_nd(L1, L2, Pq);
    } else if (kind == 1) {
        mul_dn(A, B, L2);
        mul_nn(L1, L2, Pq);
    } else if (kind == 2) {
        mul_nd(A, B, L2);
        mul_nn(L1, L2, Pq);
    } else {
        mul_nn(A, B, L2);
        mul_nn(L1, L2, Pq);
    }
}

// ------------------------------------------------------------------------------------------------
// The leapfrog step: Dynamics::SUNc::Update (SC/DYNAMICS/SUNcDynamics.cpp:353-378) fused per site.
//
//   E'_mu(x)    = E_mu(x) + gamma c_mu Re tr(i t^a U_mu(x) Sigma_mu(x))                       (:134-203)
//                 (- (dt/a_z)(J+(x) - J-(x))/sqrt2 on E_z), with the staple sum
//   Sigma_mu    = sum_{nu != mu} c_nu [ U_nu(x+mu) U_mu^dag(x+nu) U_nu^dag(x)
//                                      + U_nu^dag(x+mu-nu) U_mu^dag(x-nu) U_nu(x-nu) ]
//   J+-stat'(x) = J+-stat(x) -+ (dt/a_z) [J+-(x) - U_z^dag(x-z) J+-(x-z) U_z(x-z)]            (:205-236)
//   X_mu        = exp(i c_mu E'_mu(x)),  c_mu = -(dt/2) a_mu^2 / a^3                           (:62-106)
//   U_h         = X U,  UOld = U_h                                                               (:108-128)
//   J+-'(x)     = J+-(x) -+ (dt/a_z) [U_h,z(x) J+-stat'(x+z) U_h,z^dag(x) - J+-stat'(x)]         (:238-269)
//   U'          = X U_h
//
// The staple form is the reference's nine plaquettes regrouped: since Re tr(i t^a X^dag) =
// -Re tr(i t^a X), the six plaquettes of :181-183 that contain U_mu(x) combine into U_mu(x) Sigma_mu
// (the same 27 matrix products per site).  Every stage reads only time-t data in a small
// neighbourhood, so a site advances itself from time-t buffers into time-(t+dt) buffers (U and J are
// ping-pong buffers; E is updated in place, only site x reads E(x)); J+-stat'(x+z) is recomputed at x
// from time-t data, so no grid-wide barrier is needed.  X is computed once and applied twice (the
// reference recomputes it from identical inputs), as U + (X - 1) U with X - 1 formed directly, which
// keeps rounding from biasing U away from SU(3).
// ------------------------------------------------------------------------------------------------
// Order in which the blocks of a step visit the tiles.  Workgroups are dealt to the 8 XCDs round-robin
// (block b runs on XCD b % 8), each XCD with its own L2.
//   mode 0: linear, tile b.
//   mode 1: XCD k owns the k-th band of P/8 sites of every z plane and walks the planes in order.
//   mode 2: column bands of bw tiles (64 x-sites each) by bh rows; XCD k takes bands k, k+8, ... and
//           walks each band through all z planes before the next one, so the few planes of one band that
//           are in flight at once share their z neighbours in L2.
//   mode 3: z slabs: XCD k takes the k-th eighth of the tiles in linear order (bpb tiles: whole planes and their
//           y neighbours on one XCD, for planes whose tile count is not a multiple of 8).
constexpr int kXcds = 8;

// whether the step can take its row operands from an LDS copy of the tile: nx a multiple of 64 (a tile is part
// of one row), a divisor of 64 (whole rows), or a tile spans at most two row segments (every tile start x0 =
// 64 t mod nx, a multiple of gcd(64, nx), has x0 + 64 <= 2 nx, i.e. nx >= 64 - gcd(64, nx); nx = 48 and 96
// among others) with no partial last tile
GHD int gcd64(int n) {
    int g = kTile;
    while (n % g) g >>= 1;
    return g;
}
// 0: the plain form; 1: the row-staged form with whole rows or one row segment per tile; 2: two segments
GHD int row_form(const Geom& g) {
    if (g.nx % kTile == 0 || kTile % g.nx == 0) return 1;
    return g.nx >= kTile - gcd64(g.nx) && g.S % kTile == 0 ? 2 : 0;
}
GHD bool row_tiles(const Geom& g) { return row_form(g) != 0; }
struct TileMap {
    int mode, bpb;
    int bw, bh, nbx, tpbp;
};

// first site of block b
GHD long tile_site(const Geom& g, const TileMap& m, long b) {
    if (m.mode == 0) return b * kTile;
    const long k = b % kXcds, i = b / kXcds;
    if (m.mode == 3) return (k * m.bpb + i) * kTile;
    if (m.mode == 1) {
        const long z = i / m.bpb, j = i % m.bpb;
        return z * g.P + k * (g.P / kXcds) + j * kTile;
    }
    const long per_band = (long)g.nz * m.tpbp;
    const long mb = i / per_band, r = i - mb * per_band;
    const long z = r / m.tpbp, t = r - z * m.tpbp;
    const long band = mb * kXcds + k;
    const long bxo = (band % m.nbx) * m.bw, byo = (band / m.nbx) * m.bh;
    const long tx = t % m.bw, ty = t / m.bw;
    return z * g.P + (byo + ty) * g.nx + (bxo + tx) * kTile;
}

struct StepArgs {
    const double* __restrict__ U;  // time t links
    double* __restrict__ Un;       // time t+dt links
    double* __restrict__ E;        // time t electric fields (updated in place when En is null)
    double* __restrict__ En = nullptr;  // time t+dt electric fields, when not in place
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

// J+-stat' of site xs from time-t data, xd = xs - z (SC/DYNAMICS/SUNcDynamics.cpp:205-236)
template <class I>
GHD void jstat_core(const StepArgs& A, I xs, I xd, const M3& Ud, double* jps, double* jms) {
    double jv[8], jdn[8], down[8];
    ldv<I>(A.J, kJPlanes, JPS * 8, xs, jps);
    ldv<I>(A.J, kJPlanes, JP * 8, xd, jdn);
    inv_adj_mul(Ud, jdn, down);
    ldv<I>(A.J, kJPlanes, JP * 8, xs, jv);
#pragma unroll
    for (int a = 0; a < 8; a++) jps[a] += -A.dtz * (jv[a] - down[a]);
    ldv<I>(A.J, kJPlanes, JMS * 8, xs, jms);
    ldv<I>(A.J, kJPlanes, JM * 8, xd, jdn);
    inv_adj_mul(Ud, jdn, down);
    ldv<I>(A.J, kJPlanes, JM * 8, xs, jv);
#pragma unroll
    for (int a = 0; a < 8; a++) jms[a] += A.dtz * (jv[a] - down[a]);
}
template <class I>
GHD void jstat_new(const StepArgs& A, I xs, I xd, double* jps, double* jms) {
    M3 Ud;
    ldm<I>(A.U, kUPlanes, 36, xd, Ud);
    jstat_core<I>(A, xs, xd, Ud, jps, jms);
}

// jstat_new at xs = x that also returns the J+-(x) it read, for the J advection of the same site
template <class I>
GHD void jstat_new_keep(const StepArgs& A, I xs, I xd, double* jps, double* jms, double* jp, double* jm) {
    M3 Ud;
    double jdn[8], down[8];
    ldm<I>(A.U, kUPlanes, 36, xd, Ud);
    ldv<I>(A.J, kJPlanes, JPS * 8, xs, jps);
    ldv<I>(A.J, kJPlanes, JP * 8, xd, jdn);
    inv_adj_mul(Ud, jdn, down);
    ldv<I>(A.J, kJPlanes, JP * 8, xs, jp);
#pragma unroll
    for (int a = 0; a < 8; a++) jps[a] += -A.dtz * (jp[a] - down[a]);
    ldv<I>(A.J, kJPlanes, JMS * 8, xs, jms);
    ldv<I>(A.J, kJPlanes, JM * 8, xd, jdn);
    inv_adj_mul(Ud, jdn, down);
    ldv<I>(A.J, kJPlanes, JM * 8, xs, jm);
#pragma unroll
    for (int a = 0; a < 8; a++) jms[a] += A.dtz * (jm[a] - down[a]);
}

// J+-stat' at the site above (x+z), or at x itself on the open top layer (the reference's ghost copy)
template <class I>
GHD void jstat_new_up(const StepArgs& A, I s, int z, double* jpsU, double* jmsU) {
    const int zp = zup(A.g, z);
    if (zp != z) {
        jstat_new<I>(A, (I)((long)s + A.g.P * (long)(zp - z)), s, jpsU, jmsU);
    } else {
        const int zm = zdn(A.g, z);
        jstat_new<I>(A, s, (I)((long)s + A.g.P * (long)(zm - z)), jpsU, jmsU);
    }
}

// The rest of link_update once the staple sum Sig is formed: E'_MU, the exponential, U_h (and UOld),
// U'_MU, with Uo = U_MU(x) and, for the z link, J+-(x) given.
template <class I, int MU>
GHD void link_finish_j(const StepArgs& A, I s, M3& Uo, const M3& Sig, M3& Uh, const double* jp, const double* jm) {
    double Em[8];
    {
        double F[8];
        retr_igen_prod(Uo, Sig, F);
        ldv_nt<I>(A.E, kEPlanes, 8 * MU, s, Em);
#pragma unroll
        for (int a = 0; a < 8; a++) Em[a] += A.gc[MU] * F[a];
    }
    if (MU == 2) {
#pragma unroll
        for (int a = 0; a < 8; a++) Em[a] += -A.dtz * ((jp[a] - jm[a]) / kSqrt2);
    }
    stv<I>(A.En ? A.En : A.E, kEPlanes, 8 * MU, s, Em);
    M3 D;
    expi_m1(A.cU[MU], Em, D);
    add_mul(D, Uo, Uh);
    if (A.UOld) stm<I>(A.UOld, kUPlanes, 18 * MU, s, Uh);
    add_mul(D, Uh, Uo);
    stm<I>(A.Un, kUPlanes, 18 * MU, s, Uo);
}
// the same with J+-(x) (the E source of the z link) read here
template <class I, int MU>
GHD void link_finish(const StepArgs& A, I s, M3& Uo, const M3& Sig, M3& Uh) {
    double jp[8] = {}, jm[8] = {};
    if (MU == 2) {
        ldv<I>(A.J, kJPlanes, JP * 8, s, jp);
        ldv<I>(A.J, kJPlanes, JM * 8, s, jm);
    }
    link_finish_j<I, MU>(A, s, Uo, Sig, Uh, jp, jm);
}

// Link direction MU of site s: E'_MU, the exponential, U_h (and UOld), U'_MU.  Returns U_h in Uh.
template <class I, int MU>
GHD void link_update(const StepArgs& A, I s, int x, int y, int z, M3& Uh) {
    const Geom& g = A.g;
    M3 Sig;
#pragma unroll
    for (int k = 0; k < 9; k++) { Sig.re[k] = 0.0; Sig.im[k] = 0.0; }
#pragma unroll
    for (int q = 1; q <= 2; q++) {
        const int nu = (MU + q) % 3;
        const double cn = A.cE[nu];
        M3 P1, P2, T;
        // forward staple U_nu(x+mu) U_mu^dag(x+nu) U_nu^dag(x)
        ldm<I>(A.U, kUPlanes, 18 * nu, nbr<I>(g, x, y, z, MU, 1, nu, 0), P1);
        ldm<I>(A.U, kUPlanes, 18 * MU, nbr<I>(g, x, y, z, nu, 1, MU, 0), P2);
        mul_nd(P1, P2, T);
        ldm<I>(A.U, kUPlanes, 18 * nu, s, P1);
        acc_mul_nd(cn, T, P1, Sig);
        // backward staple U_nu^dag(x+mu-nu) U_mu^dag(x-nu) U_nu(x-nu)
        const I sm = nbr<I>(g, x, y, z, nu, -1, MU, 0);
        ldm<I>(A.U, kUPlanes, 18 * nu, nbr<I>(g, x, y, z, MU, 1, nu, -1), P1);
        ldm<I>(A.U, kUPlanes, 18 * MU, sm, P2);
        mul_dd(P1, P2, T);
        ldm<I>(A.U, kUPlanes, 18 * nu, sm, P1);
        acc_mul_nn(cn, T, P1, Sig);
    }
    M3 Uo;
    ldm<I>(A.U, kUPlanes, 18 * MU, s, Uo);
    link_finish<I, MU>(A, s, Uo, Sig, Uh);
}


// J+-'(x) = J+-(x) -+ (dt/a_z) [U_h,z(x) J+-stat'(x+z) U_h,z^dag(x) - J+-stat'(x)]  (:238-269)
// the same with J+-(x) already at hand (jp, jm are overwritten)
template <class I>
GHD void j_advect_with(const StepArgs& A, I s, const M3& Uhz, const double* jps, const double* jms,
                       const double* jpsU, const double* jmsU, double* jp, double* jm) {
    double up[8];
    adj_mul(Uhz, jpsU, up);
#pragma unroll
    for (int a = 0; a < 8; a++) jp[a] += -A.dtz * (up[a] - jps[a]);
    stv<I>(A.Jn, kJPlanes, JP * 8, s, jp);
    adj_mul(Uhz, jmsU, up);
#pragma unroll
    for (int a = 0; a < 8; a++) jm[a] += A.dtz * (up[a] - jms[a]);
    stv<I>(A.Jn, kJPlanes, JM * 8, s, jm);
}

template <class I>
GHD void j_advect(const StepArgs& A, I s, const M3& Uhz, const double* jps, const double* jms, const double* jpsU,
                  const double* jmsU) {
    double jp[8], jm[8];
    ldv<I>(A.J, kJPlanes, JP * 8, s, jp);
    ldv<I>(A.J, kJPlanes, JM * 8, s, jm);
    j_advect_with<I>(A, s, Uhz, jps, jms, jpsU, jmsU, jp, jm);
}

// Whole step of one site, in the order the direction-split kernel performs it (host equivalent).
template <class I>
GHD void step_site(const StepArgs& A, I s) {
    int x, y, z;
    site_coords<I>(A.g, s, x, y, z);
    double jps[8], jms[8], jpsU[8], jmsU[8];
    const int zm = zdn(A.g, z);
    jstat_new<I>(A, s, (I)((long)s + A.g.P * (long)(zm - z)), jps, jms);
    stv<I>(A.Jn, kJPlanes, JPS * 8, s, jps);
    stv<I>(A.Jn, kJPlanes, JMS * 8, s, jms);
    jstat_new_up<I>(A, s, z, jpsU, jmsU);
    M3 Uh;
    link_update<I, 0>(A, s, x, y, z, Uh);
    link_update<I, 1>(A, s, x, y, z, Uh);
    link_update<I, 2>(A, s, x, y, z, Uh);
    j_advect<I>(A, s, Uh, jps, jms, jpsU, jmsU);
}

}  // namespace glasma
