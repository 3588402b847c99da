// libglasma core: context lifetime, field transfer between the reference AoS layout and the device
// SoA layout, the time loop entry points and the observable reductions.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "internal.h"

namespace glasma {

int launch_step(glasma_ctx* ctx, bool write_old);

int set_error(glasma_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}
int hip_code(hipError_t e) {
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? GLASMA_ERR_NOMEM : GLASMA_ERR_HIP;
}
int hip_fail(glasma_ctx* ctx, hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return set_error(ctx, hip_code(e), m);
}
int ensure_red(glasma_ctx* ctx, size_t doubles) {
    if (ctx->red_cap >= doubles) return 0;
    if (ctx->red) (void)hipFree(ctx->red);
    ctx->red = nullptr;
    ctx->red_cap = 0;
    GL_HIP(ctx, hipMalloc(&ctx->red, doubles * sizeof(double)));
    ctx->red_cap = doubles;
    return 0;
}
int ensure_fftbuf(glasma_ctx* ctx, size_t doubles) {
    if (ctx->fftbuf_cap >= doubles) return 0;
    if (ctx->fftbuf) (void)hipFree(ctx->fftbuf);
    ctx->fftbuf = nullptr;
    ctx->fftbuf_cap = 0;
    GL_HIP(ctx, hipMalloc(&ctx->fftbuf, doubles * sizeof(double)));
    ctx->fftbuf_cap = doubles;
    return 0;
}

// ------------------------------------------------------------------------------------------------
// AoS <-> SoA transposition (record of R doubles per site)
// ------------------------------------------------------------------------------------------------
// record of R doubles per site (AoS) <-> planes p0 .. p0+R-1 of a tiled field with NP planes
__global__ void k_aos_to_soa(const double* __restrict__ aos, double* __restrict__ soa, long S, int R, int NP, int p0) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S * R) return;
    long s = i / R;
    int j = (int)(i % R);
    soa[lidx<long>(NP, p0 + j, s)] = aos[i];
}
__global__ void k_soa_to_aos(const double* __restrict__ soa, double* __restrict__ aos, long S, int R, int NP, int p0) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S * R) return;
    long s = i / R;
    int j = (int)(i % R);
    aos[i] = soa[lidx<long>(NP, p0 + j, s)];
}

int upload_aos(glasma_ctx* ctx, const double* host_with_ghosts, double* dev_soa, int R, int NP, int p0) {
    if (NP < 0) NP = R;
    const long S = ctx->g.S, P = ctx->g.P;
    double* tmp = nullptr;
    GL_HIP(ctx, hipMalloc(&tmp, (size_t)S * R * sizeof(double)));
    hipError_t e = hipMemcpyAsync(tmp, host_with_ghosts + (size_t)P * R, (size_t)S * R * sizeof(double),
                                  hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        long n = S * R;
        hipLaunchKernelGGL(k_aos_to_soa, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, tmp, dev_soa, S, R,
                           NP, p0);
        e = hipStreamSynchronize(ctx->stream);
    }
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(ctx, e, "upload_aos");
    return 0;
}

// fill the two ghost layers of a host AoS array as SynchronizeGhostCells does
void host_sync_ghosts(const glasma_ctx* ctx, double* h, int R) {
    const long P = ctx->g.P * R;
    const int nz = ctx->g.nz;
    if (ctx->p.bc_z == 0) {
        std::memcpy(h, h + P, P * sizeof(double));
        std::memcpy(h + (long)(nz + 1) * P, h + (long)nz * P, P * sizeof(double));
    } else {
        std::memcpy(h, h + (long)nz * P, P * sizeof(double));
        std::memcpy(h + (long)(nz + 1) * P, h + P, P * sizeof(double));
    }
}

int download_aos(glasma_ctx* ctx, const double* dev_soa, double* host_with_ghosts, int R, int NP, int p0) {
    if (NP < 0) NP = R;
    const long S = ctx->g.S, P = ctx->g.P;
    double* tmp = nullptr;
    GL_HIP(ctx, hipMalloc(&tmp, (size_t)S * R * sizeof(double)));
    long n = S * R;
    hipLaunchKernelGGL(k_soa_to_aos, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, dev_soa, tmp, S, R, NP,
                       p0);
    hipError_t e = hipMemcpyAsync(host_with_ghosts + (size_t)P * R, tmp, (size_t)S * R * sizeof(double),
                                  hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(ctx, e, "download_aos");
    host_sync_ghosts(ctx, host_with_ghosts, R);
    return 0;
}

static int upload_j(glasma_ctx* ctx, const double* h, double* J, int w) { return upload_aos(ctx, h, J, 8, kJPlanes, 8 * w); }
static int download_j(glasma_ctx* ctx, const double* J, double* h, int w) {
    return download_aos(ctx, J, h, 8, kJPlanes, 8 * w);
}

// identity links / zero fields
// identity matrices in all nmat matrix slots of a field with 18*nmat planes
__global__ void k_set_identity(double* __restrict__ U, long S, int nmat) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    for (int m = 0; m < nmat; m++)
        for (int e = 0; e < 18; e++)
            U[lidx<long>(18 * nmat, 18 * m + e, s)] = (e == 0 || e == 8 || e == 16) ? 1.0 : 0.0;
}

// counter-based uniform in (0,1): splitmix64 finaliser of (seed, index)
__device__ __forceinline__ double hash_uniform(unsigned long long seed, unsigned long long i) {
    unsigned long long z = seed * 0x9E3779B97F4A7C15ull + i * 0xD1B54A32D192ED03ull + 0x632BE59BD9B4E019ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
// uniform in [-1, 1) per component: links exp(i s_U alpha t) with alpha_a uniform, E and J uniform
__global__ void __launch_bounds__(256) k_fill_random(double* __restrict__ U, double* __restrict__ E, double* __restrict__ J, long S,
                              unsigned long long seed, double su, double se, double sj) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    unsigned long long base = (unsigned long long)s * 64ull;
    for (int mu = 0; mu < 3; mu++) {
        double al[8];
        for (int a = 0; a < 8; a++) al[a] = 2.0 * hash_uniform(seed, base + mu * 8 + a) - 1.0;
        M3 X;
        expi(su, al, X);
        stm<long>(U, kUPlanes, 18 * mu, s, X);
        for (int a = 0; a < 8; a++)
            E[lidx<long>(kEPlanes, mu * 8 + a, s)] = se * (2.0 * hash_uniform(seed ^ 0x5555ull, base + mu * 8 + a) - 1.0);
    }
    for (int w = 0; w < 4; w++)
        for (int a = 0; a < 8; a++)
            J[lidx<long>(kJPlanes, w * 8 + a, s)] = sj * (2.0 * hash_uniform(seed ^ 0xAAAAull, base + w * 8 + a) - 1.0);
}

int reset_fields(glasma_ctx* ctx) {
    const long S = ctx->g.S;
    dim3 grid((unsigned)((S + 255) / 256));
    hipLaunchKernelGGL(k_set_identity, grid, dim3(256), 0, ctx->stream, ctx->U[0], S, 3);
    hipLaunchKernelGGL(k_set_identity, grid, dim3(256), 0, ctx->stream, ctx->U[1], S, 3);
    hipLaunchKernelGGL(k_set_identity, grid, dim3(256), 0, ctx->stream, ctx->UOld, S, 3);
    const long Sp = pad_sites(S);
    GL_HIP(ctx, hipMemsetAsync(ctx->E, 0, (size_t)kEPlanes * Sp * sizeof(double), ctx->stream));
    GL_HIP(ctx, hipMemsetAsync(ctx->J[0], 0, (size_t)kJPlanes * Sp * sizeof(double), ctx->stream));
    GL_HIP(ctx, hipMemsetAsync(ctx->J[1], 0, (size_t)kJPlanes * Sp * sizeof(double), ctx->stream));
    GL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->cur = 0;
    ctx->t = 0.0;
    ctx->tsteps = 0;
    return 0;
}

// ------------------------------------------------------------------------------------------------
// observables
// ------------------------------------------------------------------------------------------------
// One block of 256 threads per (chunk of a z plane); partial sums of K quantities per block are
// written to red[(z*bpp + b)*K + k] and summed afterwards in a fixed order: results do not depend on
// scheduling and are reproducible bit for bit.
template <int K>
__device__ __forceinline__ void block_store(double (&v)[K], double* out) {
    __shared__ double sh[4][K];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) sh[wid][k] = v[k];
    __syncthreads();
    if (threadIdx.x < K) {
        const int k = threadIdx.x;
        out[k] = ((sh[0][k] + sh[1][k]) + sh[2][k]) + sh[3][k];
    }
}
template <int K>
__device__ __forceinline__ void block_store_max(double (&v)[K], double* out) {
    __shared__ double shm[4][K];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = wave_max(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) shm[wid][k] = v[k];
    __syncthreads();
    if (threadIdx.x < K) {
        const int k = threadIdx.x;
        out[k] = fmax(fmax(shm[0][k], shm[1][k]), fmax(shm[2][k], shm[3][k]));
    }
}

struct ObsArgs {
    const double* __restrict__ U;
    const double* __restrict__ UOld;
    const double* __restrict__ E;
    const double* __restrict__ J;
    double* __restrict__ red;
    Geom g;
    int bpp;  // blocks per plane
    double a[3];
};

__device__ __forceinline__ bool site_of(const ObsArgs& A, int& x, int& y, int& z, long& s) {
    z = blockIdx.y;
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= A.g.P) return false;
    x = (int)(i % A.g.nx);
    y = (int)(i / A.g.nx);
    s = i + A.g.P * z;
    return true;
}

__device__ __forceinline__ void plaq_full(const double* __restrict__ U, long S, int kind, int m1, long s1, int m2,
                                          long s2, int m3, long s3, int m4, long s4, M3& Pq) {
    (void)S;
    plaq_mat<long>(U, kind, m1, s1, m2, s2, m3, s3, m4, s4, Pq);
}
__device__ __forceinline__ double Ev(const double* __restrict__ E, int mu, int a, long s) {
    return E[lidx<long>(kEPlanes, mu * 8 + a, s)];
}
__device__ __forceinline__ double Jv(const double* __restrict__ J, int w, int a, long s) {
    return J[lidx<long>(kJPlanes, w * 8 + a, s)];
}

// Observables::Bulk (SC/OBSERVABLES/BulkObservables.cpp:65-169): K = 9 (E^2[3], B^2[3], E.J[3])
__global__ void __launch_bounds__(256) k_bulk(ObsArgs A) {
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int x, y, z;
    long s;
    if (site_of(A, x, y, z, s)) {
        const Geom& g = A.g;
        const long S = g.S;
        const long sxp = sidx(g, x + 1, y, z), syp = sidx(g, x, y + 1, z), szp = sidx(g, x, y, zup(g, z));
        M3 Pq;
        plaq_full(A.U, S, 0, 0, s, 1, sxp, 0, syp, 1, s, Pq);
        v[5] = 4.0 * retr_one_minus(Pq);  // B2 from Uxy
        plaq_full(A.U, S, 0, 1, s, 2, syp, 1, szp, 2, s, Pq);
        v[3] = 4.0 * retr_one_minus(Pq);  // B0 from Uyz
        plaq_full(A.U, S, 0, 2, s, 0, szp, 2, sxp, 0, s, Pq);
        v[4] = 4.0 * retr_one_minus(Pq);  // B1 from Uzx
        for (int a = 0; a < 8; a++) {
            const double e0 = Ev(A.E, 0, a, s), e1 = Ev(A.E, 1, a, s), e2 = Ev(A.E, 2, a, s);
            v[0] += e0 * e0;
            v[1] += e1 * e1;
            v[2] += e2 * e2;
            v[8] += e2 * (Jv(A.J, JP, a, s) - Jv(A.J, JM, a, s)) / kSqrt2;
        }
    }
    block_store<9>(v, A.red + ((long)blockIdx.y * A.bpp + blockIdx.x) * 9);
}

// COMPUTE_GAUSS_VIOLATION (SC/CALC/GaussViolation.cpp:13-23): K = 2 sums (sum G^2) + 1 max (|G^a|)
__global__ void __launch_bounds__(256) k_gauss(ObsArgs A) {
    double v[1] = {0.0};
    double m[1] = {0.0};
    int x, y, z;
    long s;
    if (site_of(A, x, y, z, s)) {
        const Geom& g = A.g;
        const long sxm = sidx(g, x - 1, y, z), sym = sidx(g, x, y - 1, z), szm = sidx(g, x, y, zdn(g, z));
        const double norm = 1.0 / (A.a[0] * A.a[1] * A.a[2]);
        M3 V;
        double e[8], xd[8], yd[8], zd[8];
        ldm<long>(A.U, kUPlanes, 0, sxm, V);
        ldv<long>(A.E, kEPlanes, 0, sxm, e);
        inv_adj_mul(V, e, xd);
        ldm<long>(A.U, kUPlanes, 18, sym, V);
        ldv<long>(A.E, kEPlanes, 8, sym, e);
        inv_adj_mul(V, e, yd);
        ldm<long>(A.U, kUPlanes, 36, szm, V);
        ldv<long>(A.E, kEPlanes, 16, szm, e);
        inv_adj_mul(V, e, zd);
        for (int a = 0; a < 8; a++) {
            const double G = norm * ((Ev(A.E, 0, a, s) - xd[a]) + (Ev(A.E, 1, a, s) - yd[a]) + (Ev(A.E, 2, a, s) - zd[a]) -
                                     (Jv(A.J, JPS, a, s) + Jv(A.J, JMS, a, s)) / kSqrt2);
            v[0] += G * G;
            m[0] = fmax(m[0], fabs(G));
        }
    }
    double* out = A.red + ((long)blockIdx.y * A.bpp + blockIdx.x) * 2;
    block_store<1>(v, out);
    block_store_max<1>(m, out + 1);
}

// Unitarity (SC/OBSERVABLES/UnitarityViolation.cpp:13-51): sum of squares and max
__global__ void __launch_bounds__(256) k_unitarity(ObsArgs A) {
    double v[1] = {0.0}, m[1] = {0.0};
    int x, y, z;
    long s;
    if (site_of(A, x, y, z, s)) {
        for (int mu = 0; mu < 3; mu++) {
            M3 V;
            ldm<long>(A.U, kUPlanes, 18 * mu, s, V);
            const double n = unitarity_norm(V);
            v[0] += n * n;
            m[0] = fmax(m[0], n);
        }
    }
    double* out = A.red + ((long)blockIdx.y * A.bpp + blockIdx.x) * 2;
    block_store<1>(v, out);
    block_store_max<1>(m, out + 1);
}

// EnergyMomentumTensor::CreateMap (SC/OBSERVABLES/EnergyMomentumTensor.cpp:92-170) with
// COMPUTE_AVG_FIELD_STRENGTH (SC/CALC/AvgFieldStrength.cpp:68-145): clover B on UOld, E averaged over
// x and x - mu (colour vectors parallel transported with the current U).  K = 10.
__global__ void __launch_bounds__(256) k_emt(ObsArgs A) {
    double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    int x, y, z;
    long s;
    if (site_of(A, x, y, z, s)) {
        const Geom& g = A.g;
        const long S = g.S;
        const int zp = zup(g, z), zm = zdn(g, z);
        const long s0 = s;
        const long sxp = sidx(g, x + 1, y, z), sxm = sidx(g, x - 1, y, z), syp = sidx(g, x, y + 1, z),
                   sym = sidx(g, x, y - 1, z), szp = sidx(g, x, y, zp), szm = sidx(g, x, y, zm);
        const long s_mxpy = sidx(g, x - 1, y + 1, z), s_xpmy = sidx(g, x + 1, y - 1, z);
        const long s_ympz = sidx(g, x, y - 1, zp), s_ypmz = sidx(g, x, y + 1, zm);
        const long s_xpmz = sidx(g, x + 1, y, zm), s_mxpz = sidx(g, x - 1, y, zp);
        const long s_mxmy = sidx(g, x - 1, y - 1, z), s_mymz = sidx(g, x, y - 1, zm), s_mxmz = sidx(g, x - 1, y, zm);
        const double* W = A.UOld;
        M3 Pq, Axy, Ayz, Azx;
        // clover averages, SC/SUNc/AdvancedOperations.cpp:35-40 (0.25 * (P1 + P2 + P3 + P4))
        auto acc4 = [&](M3& Av, const M3& p1, const M3& p2, const M3& p3, const M3& p4) {
            for (int k = 0; k < 9; k++) {
                Av.re[k] = 0.25 * (p1.re[k] + p2.re[k] + p3.re[k] + p4.re[k]);
                Av.im[k] = 0.25 * (p1.im[k] + p2.im[k] + p3.im[k] + p4.im[k]);
            }
        };
        {
            M3 a1, a2, a3, a4;
            plaq_full(W, S, 0, 0, s0, 1, sxp, 0, syp, 1, s0, a1);        // Uxy
            plaq_full(W, S, 1, 1, s0, 0, s_mxpy, 1, sxm, 0, sxm, a2);    // UMxy
            plaq_full(W, S, 2, 1, sym, 0, sym, 1, s_xpmy, 0, s0, a3);    // UxMy
            plaq_full(W, S, 3, 0, sxm, 1, s_mxmy, 0, s_mxmy, 1, sym, a4);  // UMxMy
            acc4(Axy, a1, a2, a3, a4);
            plaq_full(W, S, 0, 1, s0, 2, syp, 1, szp, 2, s0, a1);        // Uyz
            plaq_full(W, S, 1, 2, s0, 1, s_ympz, 2, sym, 1, sym, a2);    // UMyz
            plaq_full(W, S, 2, 2, szm, 1, szm, 2, s_ypmz, 1, s0, a3);    // UyMz
            plaq_full(W, S, 3, 1, sym, 2, s_mymz, 1, s_mymz, 2, szm, a4);  // UMyMz
            acc4(Ayz, a1, a2, a3, a4);
            plaq_full(W, S, 0, 2, s0, 0, szp, 2, sxp, 0, s0, a1);        // Uzx
            plaq_full(W, S, 1, 0, s0, 2, s_xpmz, 0, szm, 2, szm, a2);    // UMzx
            plaq_full(W, S, 2, 0, sxm, 2, sxm, 0, s_mxpz, 2, s0, a3);    // UzMx
            plaq_full(W, S, 3, 2, szm, 0, s_mxmz, 2, s_mxmz, 0, sxm, a4);  // UMzMx
            acc4(Azx, a1, a2, a3, a4);
        }
        (void)Pq;
        double B0[8], B1[8], B2[8];
        retr_igen(Ayz, B0);
        retr_igen(Azx, B1);
        retr_igen(Axy, B2);
        for (int a = 0; a < 8; a++) { B0[a] *= 2.0; B1[a] *= 2.0; B2[a] *= 2.0; }
        const double B0S = 4.0 * retr_one_minus(Ayz), B1S = 4.0 * retr_one_minus(Azx), B2S = 4.0 * retr_one_minus(Axy);
        double E0[8], E1[8], E2[8], ed[8], e[8];
        double E0S = 0, E1S = 0, E2S = 0;
        M3 V;
        const long sm[3] = {sxm, sym, szm};
        for (int mu = 0; mu < 3; mu++) {
            double* Em = mu == 0 ? E0 : (mu == 1 ? E1 : E2);
            ldm<long>(A.U, kUPlanes, 18 * mu, sm[mu], V);
            ldv<long>(A.E, kEPlanes, 8 * mu, sm[mu], e);
            inv_adj_mul(V, e, ed);
            double es = 0.0;
            for (int a = 0; a < 8; a++) {
                const double eh = Ev(A.E, mu, a, s0);
                Em[a] = 0.5 * (eh + ed[a]);
            }
            for (int a = 0; a < 8; a++) {
                const double eh = Ev(A.E, mu, a, s0);
                es += 0.5 * (eh * eh + e[a] * e[a]);
            }
            if (mu == 0) E0S = es; else if (mu == 1) E1S = es; else E2S = es;
        }
        double q0 = A.a[0] / (A.a[0] * A.a[1] * A.a[2]), q1 = A.a[1] / (A.a[0] * A.a[1] * A.a[2]),
               q2 = A.a[2] / (A.a[0] * A.a[1] * A.a[2]);
        const double inv = 1.0 / (double)g.P;
        auto dot = [](const double* p, const double* r) {
            double acc = 0.0;
            for (int a = 0; a < 8; a++) acc += p[a] * r[a];
            return acc;
        };
        const double Q0 = q0 * q0, Q1 = q1 * q1, Q2 = q2 * q2;
        v[0] = (0.5 * (Q0 * E0S + Q1 * E1S + Q2 * E2S + Q0 * B0S + Q1 * B1S + Q2 * B2S)) * inv;
        v[1] = (0.5 * ((Q1 * E1S + Q2 * E2S + Q1 * B1S + Q2 * B2S) - (Q0 * E0S + Q0 * B0S))) * inv;
        v[2] = (0.5 * ((Q0 * E0S + Q2 * E2S + Q0 * B0S + Q2 * B2S) - (Q1 * E1S + Q1 * B1S))) * inv;
        v[3] = (0.5 * ((Q0 * E0S + Q1 * E1S + Q0 * B0S + Q1 * B1S) - (Q2 * E2S + Q2 * B2S))) * inv;
        v[4] = (q1 * q2 * dot(E1, B2) - q2 * q1 * dot(E2, B1)) * inv;
        v[5] = (q2 * q0 * dot(E2, B0) - q0 * q2 * dot(E0, B2)) * inv;
        v[6] = (q0 * q1 * dot(E0, B1) - q1 * q0 * dot(E1, B0)) * inv;
        v[7] = (q0 * q1 * dot(E0, E1) + q0 * q1 * dot(B0, B1)) * inv;
        v[8] = (q1 * q2 * dot(E1, E2) + q1 * q2 * dot(B1, B2)) * inv;
        v[9] = (q2 * q0 * dot(E2, E0) + q2 * q0 * dot(B2, B0)) * inv;
    }
    block_store<10>(v, A.red + ((long)blockIdx.y * A.bpp + blockIdx.x) * 10);
}

// IO::SaveTransverselyIntegratedCurrents (SC/IO/SaveConfiguration.cpp:160-204): per layer
// sum_{x,y,a} of Jm^2, JmStat^2, Jp^2, JpStat^2.  K = 4.
__global__ void __launch_bounds__(256) k_currents(ObsArgs A) {
    double v[4] = {0, 0, 0, 0};
    int x, y, z;
    long s;
    if (site_of(A, x, y, z, s)) {
        for (int a = 0; a < 8; a++) {
            const double jm = Jv(A.J, JM, a, s), jms = Jv(A.J, JMS, a, s), jp = Jv(A.J, JP, a, s),
                         jps = Jv(A.J, JPS, a, s);
            v[0] += jm * jm;
            v[1] += jms * jms;
            v[2] += jp * jp;
            v[3] += jps * jps;
        }
    }
    block_store<4>(v, A.red + ((long)blockIdx.y * A.bpp + blockIdx.x) * 4);
}

// EnergyMomentumTensor::CreateMapBfCollision (SC/OBSERVABLES/EnergyMomentumTensor.cpp:10-77): for each
// transverse site the energy density summed over z, split at z <= (Nz - 2) / 2 (left) and above (right),
// from the elementary plaquettes of the current links and the local E.  One thread per (x, y) and chunk of
// kMapChunk layers; part[(chunk * P + i) * 2 + side].
constexpr int kMapChunk = 32;

__global__ void __launch_bounds__(256) k_tmap(ObsArgs A, double* __restrict__ part) {
    const Geom& g = A.g;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= g.P) return;
    const int x = (int)(i % g.nx), y = (int)(i / g.nx);
    const int z0 = blockIdx.y * kMapChunk, z1 = min(z0 + kMapChunk, g.nz);
    const int zsplit = (g.nz - 2) / 2;  // (zHigh - 1) / 2 with zHigh = Nz - 1
    const double ac = A.a[0] * A.a[1] * A.a[2];
    double c[3];
    for (int m = 0; m < 3; m++) c[m] = (A.a[m] / ac) * (A.a[m] / ac);  // cB = cE = (a_mu aScale^2 / aCube)^2
    const double inv_nz = 1.0 / (double)g.nz;
    double left = 0.0, right = 0.0;
    for (int z = z0; z < z1; z++) {
        const long s = i + g.P * z;
        const long sxp = sidx(g, x + 1, y, z), syp = sidx(g, x, y + 1, z), szp = sidx(g, x, y, zup(g, z));
        M3 Pq;
        plaq_full(A.U, g.S, 0, 0, s, 1, sxp, 0, syp, 1, s, Pq);
        const double B2 = 4.0 * retr_one_minus(Pq);
        plaq_full(A.U, g.S, 0, 1, s, 2, syp, 1, szp, 2, s, Pq);
        const double B0 = 4.0 * retr_one_minus(Pq);
        plaq_full(A.U, g.S, 0, 2, s, 0, szp, 2, sxp, 0, s, Pq);
        const double B1 = 4.0 * retr_one_minus(Pq);
        double E0 = 0.0, E1 = 0.0, E2 = 0.0;
        for (int a = 0; a < 8; a++) {
            const double e0 = Ev(A.E, 0, a, s), e1 = Ev(A.E, 1, a, s), e2 = Ev(A.E, 2, a, s);
            E0 += e0 * e0;
            E1 += e1 * e1;
            E2 += e2 * e2;
        }
        const double t = 0.5 * (c[0] * E0 + c[1] * E1 + c[2] * E2 + c[0] * B0 + c[1] * B1 + c[2] * B2) * inv_nz;
        if (z <= zsplit) left += t; else right += t;
    }
    double* o = part + ((long)blockIdx.y * g.P + i) * 2;
    o[0] = left;
    o[1] = right;
}

// fixed-order sum of the chunks: out[i * 2 + side] = sum_c part[(c * P + i) * 2 + side]
__global__ void k_tmap_finalize(const double* __restrict__ part, double* __restrict__ out, long P, int nchunk) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * P) return;
    double acc = 0.0;
    for (int c = 0; c < nchunk; c++) acc += part[(long)c * 2 * P + i];
    out[i] = acc;
}

// per-z fixed-order finalisation: out[z*K + k] = sum_b red[(z*bpp + b)*K + k] (or max for the
// entries flagged in maxmask)
__global__ void k_finalize(const double* __restrict__ red, double* __restrict__ out, int nz, int bpp, int K,
                           unsigned maxmask) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nz * K) return;
    int z = i / K, k = i % K;
    const bool mx = (maxmask >> k) & 1u;
    double acc = 0.0;
    for (int b = 0; b < bpp; b++) {
        double v = red[((long)z * bpp + b) * K + k];
        acc = mx ? fmax(acc, v) : acc + v;
    }
    out[i] = acc;
}

// ------------------------------------------------------------------------------------------------
// bitwise state checksum: sum over the physical sites' words of mix(bits ^ index-dependent key), wrapping
// 64-bit addition.  Integer addition is associative and commutative, so the result does not depend on the
// order of the per-block atomics; a change of any bit of any word changes the sum unless two 64-bit mixes
// collide.  The word index is the (plane, site) position, so permutations are detected too.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {  // SplitMix64 finaliser (Steele et al.)
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) k_checksum(const double* __restrict__ f, int NP, long S,
                                                  unsigned long long* __restrict__ out) {
    const long n = (long)NP * pad_sites(S);
    unsigned long long acc = 0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        // inverse of lidx: tile, plane pair, site in tile, member of the pair
        const long tile = i / ((long)NP * kTile);
        const int r = (int)(i - tile * NP * kTile);
        const long s = tile * kTile + ((r & 127) >> 1);
        if (s >= S) continue;
        const int p = 2 * (r >> 7) + (r & 1);
        const unsigned long long key = (unsigned long long)(s * NP + p) * 0x9e3779b97f4a7c15ull;
        acc += mix64((unsigned long long)__double_as_longlong(f[i]) ^ key);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    __shared__ unsigned long long sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, sh[0] + sh[1] + sh[2] + sh[3]);
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// ------------------------------------------------------------------------------------------------
// Fixed device scratch of glasma_bulk_noalloc: the block partials, the per-layer sums and the 9 totals.
// Every device the library runs on holds its own copy of the array, at its own address, until the process
// ends.  Host code never passes the array's name to a kernel (on the host it only stands for the symbol):
// it uses the device address that hipGetSymbolAddress returns, looked up once per device with that device
// current.
// ------------------------------------------------------------------------------------------------
constexpr long kScratchDoubles = 1L << 20;  // 8 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static double* g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_ptr(glasma_ctx* ctx, double** out) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return set_error(ctx, GLASMA_ERR_ARG, "glasma_bulk_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[ctx->device]) {
        int cur = 0;
        GL_HIP(ctx, hipGetDevice(&cur));
        if (cur != ctx->device) GL_HIP(ctx, hipSetDevice(ctx->device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != ctx->device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipGetSymbolAddress(scratch)");
        g_scratch[ctx->device] = static_cast<double*>(p);
    }
    *out = g_scratch[ctx->device];
    return 0;
}

// out[k] = the per-layer sums fin[z*K + k] added in increasing z, the order of glasma_bulk's host loop
__global__ void k_sum_layers(const double* __restrict__ fin, double* __restrict__ out, int nz, int K) {
    const int k = threadIdx.x;
    if (k >= K) return;
    double acc = 0.0;
    for (int z = 0; z < nz; z++) acc += fin[(long)z * K + k];
    out[k] = acc;
}

static int run_obs(glasma_ctx* ctx, void (*kern)(ObsArgs), int K, unsigned maxmask, std::vector<double>& per_z,
                   bool use_old) {
    const Geom& g = ctx->g;
    const int bpp = (int)((g.P + 255) / 256);
    int rc = ensure_red(ctx, (size_t)g.nz * bpp * K + (size_t)g.nz * K);
    if (rc) return rc;
    ObsArgs A;
    A.U = ctx->U[ctx->cur];
    A.UOld = use_old ? ctx->UOld : nullptr;
    A.E = ctx->E;
    A.J = ctx->J[ctx->cur];
    A.red = ctx->red;
    A.g = g;
    A.bpp = bpp;
    for (int m = 0; m < 3; m++) A.a[m] = ctx->p.a[m];
    hipLaunchKernelGGL(kern, dim3(bpp, g.nz), dim3(256), 0, ctx->stream, A);
    double* fin = ctx->red + (size_t)g.nz * bpp * K;
    const int n = g.nz * K;
    hipLaunchKernelGGL(k_finalize, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, ctx->red, fin, g.nz, bpp, K, maxmask);
    per_z.resize(n);
    GL_HIP(ctx, hipMemcpyAsync(per_z.data(), fin, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    GL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

}  // namespace glasma

using namespace glasma;

// ==================================================================================================
// C ABI
// ==================================================================================================
extern "C" {

const char* glasma_version(void) { return "glasma-mi355x 0.1 (gfx950, FP64)"; }

// the message of the last failed glasma_create on this thread (there is no context to hold it)
static thread_local std::string g_create_err;

static int create_fail(int code, const std::string& msg) {
    g_create_err = msg;
    return code;
}

const char* glasma_last_error(const glasma_ctx* ctx) {
    if (!ctx) return g_create_err.c_str();
    return ctx->err.c_str();
}

int glasma_create(const glasma_params* p, glasma_ctx** out) {
    if (!p || !out) return create_fail(GLASMA_ERR_ARG, "glasma_create: null argument");
    *out = nullptr;
    if (p->nc != 3) return create_fail(GLASMA_ERR_ARG, "glasma_create: only SU(3) (nc = 3) is supported");
    for (int m = 0; m < 3; m++)
        if (p->n[m] <= 0 || !(p->a[m] > 0.0))
            return create_fail(GLASMA_ERR_ARG, "glasma_create: lattice sizes and spacings must be positive");
    if (p->bc_z != 0 && p->bc_z != 1) return create_fail(GLASMA_ERR_ARG, "glasma_create: bc_z must be 0 or 1");
    glasma_ctx* ctx = new glasma_ctx();
    ctx->p = *p;
    ctx->g.nx = p->n[0];
    ctx->g.ny = p->n[1];
    ctx->g.nz = p->n[2];
    ctx->g.bc = p->bc_z;
    ctx->g.P = (long)p->n[0] * p->n[1];
    ctx->g.S = ctx->g.P * p->n[2];
    const long S = pad_sites(ctx->g.S);
    hipError_t e = hipGetDevice(&ctx->device);
    if (e == hipSuccess) e = hipStreamCreate(&ctx->stream);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
    if (e == hipSuccess) e = hipMalloc(&ctx->U[0], (size_t)kUPlanes * S * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&ctx->U[1], (size_t)kUPlanes * S * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&ctx->UOld, (size_t)kUPlanes * S * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&ctx->E, (size_t)kEPlanes * S * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&ctx->J[0], (size_t)kJPlanes * S * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&ctx->J[1], (size_t)kJPlanes * S * sizeof(double));
    if (e != hipSuccess) {
        const int code = hip_code(e);
        glasma_destroy(ctx);
        return create_fail(code, std::string("glasma_create: ") + hipGetErrorString(e));
    }
    int rc = reset_fields(ctx);
    if (rc) {
        const std::string msg = std::string("glasma_create: ") + ctx->err;
        glasma_destroy(ctx);
        return create_fail(rc, msg);
    }
    g_create_err.clear();
    *out = ctx;
    return 0;
}

int glasma_destroy(glasma_ctx* ctx) {
    if (!ctx) return 0;
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    if (ctx->stream2) hipStreamSynchronize(ctx->stream2);
    double* bufs[] = {ctx->U[0], ctx->U[1], ctx->UOld, ctx->E, ctx->J[0], ctx->J[1], ctx->red, ctx->G, ctx->Gtop,
                      ctx->UG, ctx->EG, ctx->fftbuf};
    for (double* b : bufs)
        if (b) (void)hipFree(b);
    for (auto& w : ctx->wpool)
        if (w.second) (void)hipFree(w.second);
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    for (hipEvent_t ev : ctx->seg_ev)
        if (ev) hipEventDestroy(ev);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    if (ctx->stream2) hipStreamDestroy(ctx->stream2);
    delete ctx;
    return 0;
}

int glasma_memory_bytes(const glasma_ctx* ctx, size_t* bytes) {
    if (!ctx || !bytes) return GLASMA_ERR_ARG;
    const size_t S = ctx->g.S;
    size_t b = (3 * kUPlanes + kEPlanes + 2 * kJPlanes) * S * sizeof(double);
    b += ctx->red_cap * sizeof(double) + ctx->fftbuf_cap * sizeof(double);
    if (ctx->G) b += (18 * S + 18 * ctx->g.P) * sizeof(double);
    if (ctx->UG) b += kUPlanes * S * sizeof(double);
    if (ctx->EG) b += kEPlanes * S * sizeof(double);
    for (const auto& w : ctx->wpool) b += w.first * sizeof(double);
    *bytes = b;
    return 0;
}

int glasma_set_fields(glasma_ctx* ctx, const double* U, const double* E, const double* Jp, const double* Jm,
                      const double* JpStat, const double* JmStat) {
    if (!ctx) return GLASMA_ERR_ARG;
    int rc;
    if (U && (rc = upload_aos(ctx, U, ctx->U[ctx->cur], kUPlanes))) return rc;
    if (E && (rc = upload_aos(ctx, E, ctx->E, kEPlanes))) return rc;
    const double* Js[4] = {Jp, Jm, JpStat, JmStat};
    for (int w = 0; w < 4; w++)
        if (Js[w] && (rc = upload_j(ctx, Js[w], ctx->J[ctx->cur], w))) return rc;
    return 0;
}

int glasma_get_fields(glasma_ctx* ctx, double* U, double* E, double* Jp, double* Jm, double* JpStat, double* JmStat) {
    if (!ctx) return GLASMA_ERR_ARG;
    int rc;
    if (U && (rc = download_aos(ctx, ctx->U[ctx->cur], U, kUPlanes))) return rc;
    if (E && (rc = download_aos(ctx, ctx->E, E, kEPlanes))) return rc;
    double* Js[4] = {Jp, Jm, JpStat, JmStat};
    for (int w = 0; w < 4; w++)
        if (Js[w] && (rc = download_j(ctx, ctx->J[ctx->cur], Js[w], w))) return rc;
    return 0;
}

int glasma_set_uold(glasma_ctx* ctx, const double* UOld) {
    if (!ctx || !UOld) return GLASMA_ERR_ARG;
    return upload_aos(ctx, UOld, ctx->UOld, kUPlanes);
}
int glasma_get_uold(glasma_ctx* ctx, double* UOld) {
    if (!ctx || !UOld) return GLASMA_ERR_ARG;
    return download_aos(ctx, ctx->UOld, UOld, kUPlanes);
}

int glasma_get_time(const glasma_ctx* ctx, double* t, long* tsteps) {
    if (!ctx) return GLASMA_ERR_ARG;
    if (t) *t = ctx->t;
    if (tsteps) *tsteps = ctx->tsteps;
    return 0;
}
int glasma_set_time(glasma_ctx* ctx, double t, long tsteps) {
    if (!ctx) return GLASMA_ERR_ARG;
    ctx->t = t;
    ctx->tsteps = tsteps;
    return 0;
}

int glasma_update_async(glasma_ctx* ctx, int nsteps) {
    if (!ctx || nsteps < 0) return GLASMA_ERR_ARG;
    if (nsteps == 0) return 0;
    GL_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (int i = 0; i < nsteps; i++) {
        launch_step(ctx, i == nsteps - 1);
        ctx->tsteps++;
        ctx->t += ctx->p.dt;
    }
    GL_HIP(ctx, hipGetLastError());
    GL_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    ctx->last_step_ms = -(double)nsteps;  // resolved by glasma_last_step_kernel_ms
    return 0;
}

int glasma_synchronize(glasma_ctx* ctx) {
    if (!ctx) return GLASMA_ERR_ARG;
    GL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int glasma_update(glasma_ctx* ctx, int nsteps) {
    int rc = glasma_update_async(ctx, nsteps);
    if (rc) return rc;
    return glasma_synchronize(ctx);
}

int glasma_last_step_kernel_ms(const glasma_ctx* ctx, double* ms) {
    if (!ctx || !ms) return GLASMA_ERR_ARG;
    glasma_ctx* c = const_cast<glasma_ctx*>(ctx);
    if (c->last_step_ms < 0) {
        const double n = -c->last_step_ms;
        float el = 0.f;
        GL_HIP(c, hipEventSynchronize(c->ev1));
        GL_HIP(c, hipEventElapsedTime(&el, c->ev0, c->ev1));
        c->last_step_ms = (double)el / n;
    }
    *ms = c->last_step_ms;
    return 0;
}

int glasma_fill_random(glasma_ctx* ctx, unsigned long seed, double link_scale, double e_scale, double j_scale) {
    if (!ctx) return GLASMA_ERR_ARG;
    const long S = ctx->g.S;
    hipLaunchKernelGGL(k_fill_random, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, ctx->stream, ctx->U[ctx->cur],
                       ctx->E, ctx->J[ctx->cur], S, (unsigned long long)seed, link_scale, e_scale, j_scale);
    GL_HIP(ctx, hipGetLastError());
    GL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int glasma_bulk(glasma_ctx* ctx, double* out9) {
    if (!ctx || !out9) return GLASMA_ERR_ARG;
    std::vector<double> pz;
    int rc = run_obs(ctx, k_bulk, 9, 0u, pz, false);
    if (rc) return rc;
    for (int k = 0; k < 9; k++) out9[k] = 0.0;
    for (int z = 0; z < ctx->g.nz; z++)
        for (int k = 0; k < 9; k++) out9[k] += pz[(size_t)z * 9 + k];
    for (int k = 0; k < 9; k++)
        if (!std::isfinite(out9[k]))
            return set_error(ctx, GLASMA_ERR_NUMERIC, "non-finite fields: the evolution diverged or a step's current "
                                                      "hand-over timed out (sums returned as computed)");
    return 0;
}

int glasma_bulk_noalloc(glasma_ctx* ctx, double* out9) {
    if (!ctx || !out9) return GLASMA_ERR_ARG;
    const Geom& g = ctx->g;
    const int bpp = (int)((g.P + 255) / 256);
    const long need = (long)g.nz * bpp * 9 + (long)g.nz * 9 + 9;
    if (need > kScratchDoubles) {
        char msg[200];
        std::snprintf(msg, sizeof(msg), "glasma_bulk_noalloc: this lattice needs %ld doubles of scratch, more than the "
                      "fixed %ld (use glasma_bulk)", need, kScratchDoubles);
        return set_error(ctx, GLASMA_ERR_ARG, msg);
    }
    double* ptr = nullptr;
    int rc = scratch_ptr(ctx, &ptr);
    if (rc) return rc;
    // the kernels write every entry they later read (all blocks store their partials), so nothing is cleared
    ObsArgs A;
    A.U = ctx->U[ctx->cur];
    A.UOld = nullptr;
    A.E = ctx->E;
    A.J = ctx->J[ctx->cur];
    A.red = ptr;
    A.g = g;
    A.bpp = bpp;
    for (int m = 0; m < 3; m++) A.a[m] = ctx->p.a[m];
    double* fin = ptr + (long)g.nz * bpp * 9;
    double* tot = fin + (long)g.nz * 9;
    hipLaunchKernelGGL(k_bulk, dim3(bpp, g.nz), dim3(256), 0, ctx->stream, A);
    hipLaunchKernelGGL(k_finalize, dim3((g.nz * 9 + 255) / 256), dim3(256), 0, ctx->stream, ptr, fin, g.nz, bpp, 9, 0u);
    hipLaunchKernelGGL(k_sum_layers, dim3(1), dim3(64), 0, ctx->stream, fin, tot, g.nz, 9);
    GL_HIP(ctx, hipGetLastError());
    GL_HIP(ctx, hipMemcpyAsync(out9, tot, 9 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    GL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < 9; k++)
        if (!std::isfinite(out9[k]))
            return set_error(ctx, GLASMA_ERR_NUMERIC, "non-finite fields: the evolution diverged or a step's current "
                                                      "hand-over timed out (sums returned as computed)");
    return 0;
}

int glasma_gauss(glasma_ctx* ctx, double* per_z, double* gmax, double* grms) {
    if (!ctx) return GLASMA_ERR_ARG;
    std::vector<double> pz;
    int rc = run_obs(ctx, k_gauss, 2, 2u, pz, false);
    if (rc) return rc;
    double sq = 0.0, mx = 0.0;
    for (int z = 0; z < ctx->g.nz; z++) {
        if (per_z) per_z[z] = pz[2 * z] / (double)ctx->g.P;
        sq += pz[2 * z];
        mx = fmax(mx, pz[2 * z + 1]);
    }
    if (gmax) *gmax = mx;
    if (grms) *grms = sqrt(sq / (8.0 * (double)ctx->g.S));
    return 0;
}

int glasma_unitarity(glasma_ctx* ctx, double* umax, double* urms) {
    if (!ctx) return GLASMA_ERR_ARG;
    std::vector<double> pz;
    int rc = run_obs(ctx, k_unitarity, 2, 2u, pz, false);
    if (rc) return rc;
    double sq = 0.0, mx = 0.0;
    for (int z = 0; z < ctx->g.nz; z++) {
        sq += pz[2 * z];
        mx = fmax(mx, pz[2 * z + 1]);
    }
    if (umax) *umax = mx;
    if (urms) *urms = sqrt(sq / (3.0 * (double)ctx->g.S));
    return 0;
}

int glasma_checksum(glasma_ctx* ctx, uint64_t* out4) {
    if (!ctx || !out4) return GLASMA_ERR_ARG;
    int rc = ensure_red(ctx, 4);
    if (rc) return rc;
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(ctx->red);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ctx->stream, acc, 4);
    const double* fields[4] = {ctx->U[ctx->cur], ctx->UOld, ctx->E, ctx->J[ctx->cur]};
    const int np[4] = {kUPlanes, kUPlanes, kEPlanes, kJPlanes};
    for (int k = 0; k < 4; k++) {
        const long n = (long)np[k] * pad_sites(ctx->g.S);
        const unsigned nb = (unsigned)std::min<long>(4096, (n + 255) / 256);
        hipLaunchKernelGGL(k_checksum, dim3(nb), dim3(256), 0, ctx->stream, fields[k], np[k], ctx->g.S, acc + k);
    }
    GL_HIP(ctx, hipGetLastError());
    unsigned long long h[4];
    GL_HIP(ctx, hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    GL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < 4; k++) out4[k] = h[k];
    return 0;
}

int glasma_emt(glasma_ctx* ctx, double* out) {
    if (!ctx || !out) return GLASMA_ERR_ARG;
    std::vector<double> pz;
    int rc = run_obs(ctx, k_emt, 10, 0u, pz, true);
    if (rc) return rc;
    std::memcpy(out, pz.data(), pz.size() * sizeof(double));
    return 0;
}

int glasma_field_profile(glasma_ctx* ctx, double* out) {
    if (!ctx || !out) return GLASMA_ERR_ARG;
    std::vector<double> pz;
    int rc = run_obs(ctx, k_bulk, 9, 0u, pz, false);
    if (rc) return rc;
    const double inv = 1.0 / (double)ctx->g.P;
    for (int z = 0; z < ctx->g.nz; z++) {
        const double* r = &pz[(size_t)z * 9];  // E0^2 E1^2 E2^2 B0 B1 B2 (sums over the layer)
        out[4 * z + 0] = 0.5 * (r[3] * inv + r[4] * inv);
        out[4 * z + 1] = 0.5 * (r[5] * inv);
        out[4 * z + 2] = 0.5 * (r[0] * inv + r[1] * inv);
        out[4 * z + 3] = 0.5 * (r[2] * inv);
    }
    return 0;
}

int glasma_current_profile(glasma_ctx* ctx, double* out) {
    if (!ctx || !out) return GLASMA_ERR_ARG;
    std::vector<double> pz;
    int rc = run_obs(ctx, k_currents, 4, 0u, pz, false);
    if (rc) return rc;
    const int nz = ctx->g.nz;
    const double inv = 1.0 / (double)ctx->g.P;
    // rows z = -1 .. nz; the ghost layers hold what SynchronizeGhostCells puts there
    for (int r = 0; r < nz + 2; r++) {
        const int z = r - 1;
        const int src = z < 0 ? (ctx->g.bc ? nz - 1 : 0) : (z >= nz ? (ctx->g.bc ? 0 : nz - 1) : z);
        for (int k = 0; k < 4; k++) out[4 * r + k] = 0.5 * pz[(size_t)src * 4 + k] * inv;
    }
    return 0;
}

int glasma_transverse_energy_map(glasma_ctx* ctx, double* out) {
    if (!ctx || !out) return GLASMA_ERR_ARG;
    const Geom& g = ctx->g;
    const int bpp = (int)((g.P + 255) / 256);
    const int nchunk = (g.nz + kMapChunk - 1) / kMapChunk;
    int rc = ensure_red(ctx, (size_t)(nchunk + 1) * 2 * g.P);
    if (rc) return rc;
    ObsArgs A;
    A.U = ctx->U[ctx->cur];
    A.UOld = nullptr;
    A.E = ctx->E;
    A.J = ctx->J[ctx->cur];
    A.red = nullptr;
    A.g = g;
    A.bpp = bpp;
    for (int m = 0; m < 3; m++) A.a[m] = ctx->p.a[m];
    double* fin = ctx->red + (size_t)nchunk * 2 * g.P;
    hipLaunchKernelGGL(k_tmap, dim3(bpp, nchunk), dim3(256), 0, ctx->stream, A, ctx->red);
    hipLaunchKernelGGL(k_tmap_finalize, dim3((unsigned)((2 * g.P + 255) / 256)), dim3(256), 0, ctx->stream, ctx->red,
                       fin, (long)g.P, nchunk);
    GL_HIP(ctx, hipMemcpyAsync(out, fin, 2 * g.P * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    GL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

}  // extern "C"
