// The fused FP64 leapfrog kernel: one launch per Dynamics::SUNc::Update (SC/DYNAMICS/SUNcDynamics.cpp:353-378).
//
// A block of 192 threads advances one tile of 64 sites; wavefront w (0, 1, 2) owns link direction w
// of every site (link_update in site.h).  The x wave also forms J+-stat'(x) and the y wave
// J+-stat'(x+z); they hand them to the z wave through LDS, which applies the J advection with its
// half-step link after one barrier.  Splitting a site over three wavefronts keeps each wavefront's
// live state to a few matrices, so two of them fit on a SIMD without spilling.
//
// HBM traffic per site-update: read U (432 B) + E (192 B) + J (256 B), write the same: 1760 B, the
// compulsory minimum.  The neighbour link matrices are re-read through L2 / MALL, so the order in
// which the 8 XCDs (each with its own 4 MB L2) visit the lattice matters: workgroups are dealt to the
// XCDs round-robin, so with the XCD-aware mapping XCD k owns the k-th band of rows of every z plane
// and walks the planes in order; rows y-1, y, y+1 are then processed by the same XCD close in time
// (L2 hits) and a plane's band is re-read by the next plane shortly afterwards (MALL hits).
//
// GLASMA_STEP_WAVES (2 or 3: register budget of 256 or 168 per lane) and GLASMA_STEP_XCD (0/1)
// select variants for measurement sweeps; the defaults are the fastest measured on MI355X.
#include <cstdlib>

#include "internal.h"

namespace glasma {

constexpr int kXcds = 8;

// first site of this block: with bpb > 0, XCD-aware band mapping with bpb tiles per band per plane
__device__ __forceinline__ long tile_first_site(const Geom& g, int bpb) {
    const long b = blockIdx.x;
    if (bpb <= 0) return b * kTile;
    const long k = b % kXcds, i = b / kXcds;
    const long z = i / bpb, j = i % bpb;
    return z * g.P + k * (g.P / kXcds) + j * kTile;
}

template <class I, int W>
__global__ void __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(W, W))) k_step(StepArgs A, int bpb) {
    __shared__ double sh[4][8][kTile];  // J+stat'(x), J-stat'(x), J+stat'(x+z), J-stat'(x+z)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long sl = tile_first_site(A.g, bpb) + lane;
    const bool valid = sl < A.g.S;
    const I s = (I)(valid ? sl : 0);
    int x, y, z;
    site_coords<I>(A.g, s, x, y, z);
    M3 Uh;
    if (wave == 0) {
        if (valid) {
            link_update<I, 0>(A, s, x, y, z, Uh);
            double jps[8], jms[8];
            const int zm = zdn(A.g, z);
            jstat_new<I>(A, s, (I)(sl + A.g.P * (long)(zm - z)), jps, jms);
            stv<I>(A.Jn, kJPlanes, JPS * 8, s, jps);
            stv<I>(A.Jn, kJPlanes, JMS * 8, s, jms);
            for (int a = 0; a < 8; a++) {
                sh[0][a][lane] = jps[a];
                sh[1][a][lane] = jms[a];
            }
        }
    } else if (wave == 1) {
        if (valid) {
            link_update<I, 1>(A, s, x, y, z, Uh);
            double ju[8], jv[8];
            jstat_new_up<I>(A, s, z, ju, jv);
            for (int a = 0; a < 8; a++) {
                sh[2][a][lane] = ju[a];
                sh[3][a][lane] = jv[a];
            }
        }
    } else {
        if (valid) link_update<I, 2>(A, s, x, y, z, Uh);
    }
    __syncthreads();
    if (wave == 2 && valid) {
        double jps[8], jms[8], jpsU[8], jmsU[8];
        for (int a = 0; a < 8; a++) {
            jps[a] = sh[0][a][lane];
            jms[a] = sh[1][a][lane];
            jpsU[a] = sh[2][a][lane];
            jmsU[a] = sh[3][a][lane];
        }
        j_advect<I>(A, s, Uh, jps, jms, jpsU, jmsU);
    }
}

// Four-wavefront variant: waves 0-2 own the link directions, wave 3 forms J+-stat'(x) and J+-stat'(x+z)
// and, after the barrier, applies the J advection with U_h,z handed over through LDS.  Two such blocks
// fill the 8 wavefront slots of a CU at two waves per SIMD.
template <class I, int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W))) k_step4(StepArgs A, int bpb) {
    __shared__ double shU[18][kTile];  // U_h,z
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long sl = tile_first_site(A.g, bpb) + lane;
    const bool valid = sl < A.g.S;
    const I s = (I)(valid ? sl : 0);
    int x, y, z;
    site_coords<I>(A.g, s, x, y, z);
    double jps[8], jms[8], jpsU[8], jmsU[8];
    M3 Uh;
    if (wave == 0) {
        if (valid) link_update<I, 0>(A, s, x, y, z, Uh);
    } else if (wave == 1) {
        if (valid) link_update<I, 1>(A, s, x, y, z, Uh);
    } else if (wave == 2) {
        if (valid) {
            link_update<I, 2>(A, s, x, y, z, Uh);
            for (int k = 0; k < 9; k++) {
                shU[2 * k][lane] = Uh.re[k];
                shU[2 * k + 1][lane] = Uh.im[k];
            }
        }
    } else {
        if (valid) {
            const int zm = zdn(A.g, z);
            jstat_new<I>(A, s, (I)(sl + A.g.P * (long)(zm - z)), jps, jms);
            stv<I>(A.Jn, kJPlanes, JPS * 8, s, jps);
            stv<I>(A.Jn, kJPlanes, JMS * 8, s, jms);
            jstat_new_up<I>(A, s, z, jpsU, jmsU);
        }
    }
    __syncthreads();
    if (wave == 3 && valid) {
        for (int k = 0; k < 9; k++) {
            Uh.re[k] = shU[2 * k][lane];
            Uh.im[k] = shU[2 * k + 1][lane];
        }
        j_advect<I>(A, s, Uh, jps, jms, jpsU, jmsU);
    }
}

static int env_int(const char* name, int dflt) {
    const char* v = std::getenv(name);
    return v ? std::atoi(v) : dflt;
}

template <class I>
static void launch(int waves, int four, int bpb, const StepArgs& A, hipStream_t st) {
    dim3 grid((unsigned)(pad_sites(A.g.S) / kTile));
    if (four) {
        hipLaunchKernelGGL((k_step4<I, 2>), grid, dim3(256), 0, st, A, bpb);
        return;
    }
    if (waves == 3) hipLaunchKernelGGL((k_step<I, 3>), grid, dim3(192), 0, st, A, bpb);
    else hipLaunchKernelGGL((k_step<I, 2>), grid, dim3(192), 0, st, A, bpb);
}

int launch_step(glasma_ctx* ctx, bool write_old) {
    static const int waves = env_int("GLASMA_STEP_WAVES", 2);
    static const int xcd = env_int("GLASMA_STEP_XCD", 1);
    static const int four = env_int("GLASMA_STEP_FOUR", 1);
    StepArgs A;
    A.U = ctx->U[ctx->cur];
    A.Un = ctx->U[ctx->cur ^ 1];
    A.E = ctx->E;
    A.J = ctx->J[ctx->cur];
    A.Jn = ctx->J[ctx->cur ^ 1];
    A.UOld = write_old ? ctx->UOld : nullptr;
    A.g = ctx->g;
    fill_step_coefficients(A, ctx->p.a, ctx->p.dt);
    // the band mapping needs whole tiles per band
    const int bpb = (xcd && ctx->g.P % (kXcds * kTile) == 0) ? (int)(ctx->g.P / (kXcds * kTile)) : 0;
    if (A.g.S < (1L << 32)) launch<unsigned int>(waves, four, bpb, A, ctx->stream);
    else launch<unsigned long>(waves, four, bpb, A, ctx->stream);
    ctx->cur ^= 1;
    return 0;
}

}  // namespace glasma
