// Internal definitions shared by the HIP translation units of libglasma.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/glasma.h"
#include "site.h"
#include "su3.h"

namespace glasma {

// ------------------------------------------------------------------------------------------------
// deterministic block reductions (256 threads = 4 wavefronts of 64)
// ------------------------------------------------------------------------------------------------
constexpr int kRedBlock = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

}  // namespace glasma

// ------------------------------------------------------------------------------------------------
// the context
// ------------------------------------------------------------------------------------------------
struct glasma_ctx {
    glasma_params p;
    glasma::Geom g;
    int device = 0;
    hipStream_t stream = nullptr;
    double* U[2] = {nullptr, nullptr};
    double* J[2] = {nullptr, nullptr};
    double* E = nullptr;
    double* UOld = nullptr;
    int cur = 0;  // which of U[.] / J[.] holds the current state
    double t = 0.0;
    long tsteps = 0;
    // reduction scratch
    double* red = nullptr;
    size_t red_cap = 0;  // doubles
    // timing
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_step_ms = 0.0;
    // gauge fixing state (allocated on first use)
    double* G = nullptr;       // gauge transformation, 18 planes
    double* Gtop = nullptr;    // the top ghost layer of G as the reference sees it (18 planes of P)
    double* UG = nullptr;      // gauge-fixed links
    double* EG = nullptr;      // gauge-fixed electric field
    double* fftbuf = nullptr;  // complex work array (2 doubles per entry)
    size_t fftbuf_cap = 0;     // doubles
    // block -> tile order of the step, when glasma_tune_step has chosen one (else the default order)
    glasma::TileMap tm{};
    bool tm_tuned = false;
    // work arrays of the slab-energy and gradient calls (doubles, pointer), handed out in request order and kept for
    // the next call, so that repeated calls of one size allocate nothing; freed by glasma_destroy
    std::vector<std::pair<size_t, double*>> wpool;
    // second stream and events of the gradient call (created on first use, default priority): a checkpoint segment's
    // recompute there overlaps the previous segment's backward sweep
    hipStream_t stream2 = nullptr;
    hipEvent_t seg_ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    std::string err;
};

namespace glasma {

int set_error(glasma_ctx* ctx, int code, const std::string& msg);
int hip_fail(glasma_ctx* ctx, hipError_t e, const char* what);
// Status code for a failed HIP call.  It also clears the thread's last-error state: HIP keeps the error of a failed
// API call (a hipMalloc past the device's memory, say) until hipGetLastError reads it, and the launch checks of the
// next entry point would otherwise report that stale error as their own.
int hip_code(hipError_t e);
int ensure_red(glasma_ctx* ctx, size_t doubles);
int ensure_fftbuf(glasma_ctx* ctx, size_t doubles);
int reset_fields(glasma_ctx* ctx);
// host array in the reference AoS layout with z ghosts (R doubles per site) <-> planes p0..p0+R-1 of a
// tiled device field with NP planes (NP < 0: NP = R)
int upload_aos(glasma_ctx* ctx, const double* host_with_ghosts, double* dev_soa, int R, int NP = -1, int p0 = 0);
int download_aos(glasma_ctx* ctx, const double* dev_soa, double* host_with_ghosts, int R, int NP = -1, int p0 = 0);
void host_sync_ghosts(const glasma_ctx* ctx, double* h, int R);
__global__ void k_set_identity(double* __restrict__ U, long S, int nmat);

// FFT over complex arrays in device memory: lines of length n with element stride `es`; line l
// starts at (l % b1) * s1 + (l / b1) * s2 (units of complex entries).  sign -1 forward, +1 backward.
// leapfrog launches (leapfrog.hip)
int launch_step(glasma_ctx* ctx, bool write_old);
// explicit buffers; ctx->cur is left alone; on stream st (the context's stream when null)
int launch_step_args(glasma_ctx* ctx, const StepArgs& A, hipStream_t st = nullptr);
StepArgs step_args(const glasma_ctx* ctx, int cur, bool write_old);
// ensemble step over nrow row-form blocks (whole rows or one segment), nseg two-segment row-form blocks and
// nplain plain-form blocks, in that order
int launch_step_batch(hipStream_t st, const StepArgs* table, const int* block_event, const unsigned* block_site,
                      long nrow, long nseg, long nplain);
TileMap default_tile_map(const Geom& g);
TileMap tile_map(const Geom& g, int mode, int bw, int bh);  // the requested order, or the next simpler one
TileMap ctx_tile_map(const glasma_ctx* ctx);                // the context's tuned order, or the default
int launch_step_map(glasma_ctx* ctx, const StepArgs& A, const TileMap& tm, hipStream_t st = nullptr);

// the left-moving MV sheet from a given charge density, and its reverse-mode derivative (mv.hip)
struct IcTape {
    double* rho_s;  // 8*S
    double* A0;     // 8*S
    double* A1;     // 8*S
    double* V0;     // 18 planes, tiled
    double* V1;     // 18 planes, tiled
};
int ic_left_from_charge(glasma_ctx* ctx, const glasma_mv_params* mv, const double* rho, IcTape& tape);
int ic_left_vjp(glasma_ctx* ctx, const glasma_mv_params* mv, const IcTape& tape, double* Ub, double* Eb,
                const double* Jb, double* rho_bar, double* work);

int fft_lines(hipStream_t st, double* data, int n, long es, long nlines, long b1, long s1, long s2, int sign);
// why the calling thread's last failed FFT call failed
const char* fft_last_error();
// complex array c[(comp*nz + z)*P + y*nx + x] of `ncomp` components: 1D along z, 2D in (x,y), 3D.
int fft_z(hipStream_t st, double* c, const Geom& g, int ncomp, int sign);
int fft_xy(hipStream_t st, double* c, const Geom& g, int ncomp, int nzl, int sign);
int fft_3d(hipStream_t st, double* c, const Geom& g, int ncomp, int sign);
// forward 3D transform, c~ -> c~ / (V p^2) for p^2 = sum_mu (2 - 2 cos(2 pi k_mu / N_mu)) / a_mu^2 > 0 (the zero
// mode unscaled, or set to 0 with drop_zero), inverse 3D transform; the z transforms and the filter run in one pass
// (Nz >= 2).  neumann (with drop_zero): in z the cosine transform, p_z^2 = (2 - 2 cos(pi k_z / N_z)) / a_z^2, which
// is the periodic step of the lattice doubled by its mirror image in z, restricted to the lattice.  fwd_x / inv_x
// false: the caller runs the forward / inverse x transform itself (fused into the kernels either side).
int fft_3d_poisson(hipStream_t st, double* c, const Geom& g, int ncomp, const double a[3], bool drop_zero = false,
                   bool neumann = false, bool fwd_x = true, bool inv_x = true);

}  // namespace glasma

#define GL_HIP(ctx, expr)                                         \
    do {                                                          \
        hipError_t _e = (expr);                                   \
        if (_e != hipSuccess) return glasma::hip_fail(ctx, _e, #expr); \
    } while (0)
