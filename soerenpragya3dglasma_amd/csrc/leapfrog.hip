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
// GLASMA_STEP_WAVES (2 or 3: register budget of 256 or 168 per lane), GLASMA_STEP_FOUR and
// GLASMA_STEP_MAP / _BW / _BH select variants for measurement sweeps; the defaults are the fastest
// measured on MI355X.
#include <cstdlib>

#include "internal.h"

namespace glasma {


template <class I, int W>
__global__ void __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(W, W))) k_step(StepArgs A, TileMap tm) {
    __shared__ double sh[4][8][kTile];  // J+stat'(x), J-stat'(x), J+stat'(x+z), J-stat'(x+z)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long sl = tile_site(A.g, tm, blockIdx.x) + lane;
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
// the body of the four-wavefront step for the tile starting at site `first`
template <class I>
__device__ __forceinline__ void step_tile4(const StepArgs& A, long first) {
    __shared__ double shU[18][kTile];  // U_h,z
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long sl = first + lane;
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

// Four-wavefront variant: waves 0-2 own the link directions, wave 3 forms J+-stat'(x) and J+-stat'(x+z)
// and, after the barrier, applies the J advection with U_h,z handed over through LDS.  Two such blocks
// fill the 8 wavefront slots of a CU at two waves per SIMD.
template <class I, int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W))) k_step4(StepArgs A, TileMap tm) {
    step_tile4<I>(A, tile_site(A.g, tm, blockIdx.x));
}

// Ensemble step: one launch advances every event of a batch.  Block b works on the tile starting at site
// block_site[b] of event e = block_event[b] (each event's blocks in the order tile_site gives it alone);
// the event's arguments come from a device table (uniform per block, so they are read with scalar
// loads).  The per-site arithmetic is that of k_step4.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_step_batch(const StepArgs* __restrict__ table, const int* __restrict__ block_event,
             const unsigned* __restrict__ block_site) {
    const int e = block_event[blockIdx.x];
    const StepArgs A = table[e];
    step_tile4<unsigned int>(A, (long)block_site[blockIdx.x]);
}

int launch_step_batch(hipStream_t st, const StepArgs* table, const int* block_event, const unsigned* block_site,
                      long nblocks) {
    if (nblocks <= 0 || nblocks > 0x7fffffffL) return GLASMA_ERR_ARG;
    hipLaunchKernelGGL(k_step_batch, dim3((unsigned)nblocks), dim3(256), 0, st, table, block_event, block_site);
    return hipGetLastError() == hipSuccess ? 0 : GLASMA_ERR_HIP;
}

// the arguments of one step of a context (current buffers, optional UOld)
StepArgs step_args(const glasma_ctx* ctx, int cur, bool write_old) {
    StepArgs A;
    A.U = ctx->U[cur];
    A.Un = ctx->U[cur ^ 1];
    A.E = ctx->E;
    A.J = ctx->J[cur];
    A.Jn = ctx->J[cur ^ 1];
    A.UOld = write_old ? ctx->UOld : nullptr;
    A.g = ctx->g;
    fill_step_coefficients(A, ctx->p.a, ctx->p.dt);
    return A;
}

static int env_int(const char* name, int dflt) {
    const char* v = std::getenv(name);
    return v ? std::atoi(v) : dflt;
}

template <class I>
static void launch(int waves, int four, const TileMap& tm, const StepArgs& A, hipStream_t st) {
    dim3 grid((unsigned)(pad_sites(A.g.S) / kTile));
    if (four) {
        if (waves == 3) hipLaunchKernelGGL((k_step4<I, 3>), grid, dim3(256), 0, st, A, tm);
        else if (waves == 4) hipLaunchKernelGGL((k_step4<I, 4>), grid, dim3(256), 0, st, A, tm);
        else hipLaunchKernelGGL((k_step4<I, 2>), grid, dim3(256), 0, st, A, tm);
        return;
    }
    if (waves == 3) hipLaunchKernelGGL((k_step<I, 3>), grid, dim3(192), 0, st, A, tm);
    else hipLaunchKernelGGL((k_step<I, 2>), grid, dim3(192), 0, st, A, tm);
}

// tile order for a geometry: the requested mode when the lattice fits it, else the next simpler one
TileMap tile_map(const Geom& g, int mode, int bw, int bh) {
    TileMap m = {0, 0, 0, 0, 0, 0};
    if (mode == 2 && g.nx % kTile == 0 && bw > 0 && bh > 0) {
        const int tpr = g.nx / kTile;
        if (tpr % bw == 0 && g.ny % bh == 0) {
            const long bands = (long)(tpr / bw) * (g.ny / bh);
            if (bands % kXcds == 0) {
                m.mode = 2;
                m.bw = bw;
                m.bh = bh;
                m.nbx = tpr / bw;
                m.tpbp = bw * bh;
                return m;
            }
        }
    }
    if (mode >= 1 && g.P % (kXcds * kTile) == 0) {
        m.mode = 1;
        m.bpb = (int)(g.P / (kXcds * kTile));
    }
    return m;
}

TileMap default_tile_map(const Geom& g) {
    // GLASMA_STEP_MAP = mode (0, 1, 2) and, for mode 2, the band width in tiles and height in rows
    // (mode 2 with 4 x 4 bands measured fastest at 256 x 256 x 1024: 33.9 ms against 36.6 ms for mode 1)
    static const int map_mode = env_int("GLASMA_STEP_MAP", 2);
    static const int map_bw = env_int("GLASMA_STEP_BW", 4);
    static const int map_bh = env_int("GLASMA_STEP_BH", 4);
    return tile_map(g, map_mode, map_bw, map_bh);
}

int launch_step(glasma_ctx* ctx, bool write_old) {
    static const int waves = env_int("GLASMA_STEP_WAVES", 2);
    static const int four = env_int("GLASMA_STEP_FOUR", 1);
    const StepArgs A = step_args(ctx, ctx->cur, write_old);
    const TileMap tm = default_tile_map(ctx->g);
    if (A.g.S < (1L << 32)) launch<unsigned int>(waves, four, tm, A, ctx->stream);
    else launch<unsigned long>(waves, four, tm, A, ctx->stream);
    ctx->cur ^= 1;
    return 0;
}

}  // namespace glasma
