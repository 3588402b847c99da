// The fused FP64 leapfrog kernel: one launch per Dynamics::SUNc::Update (SC/DYNAMICS/SUNcDynamics.cpp:353-378).
// The per-site arithmetic is step_site() in site.h.
//
// HBM traffic per site-update: read U (432 B) + E (192 B) + J (256 B), write the same: 1760 B, the
// compulsory minimum.  The 24 neighbour link matrices (13 sites) are re-read through L2 / MALL.
#include "internal.h"

namespace glasma {

__global__ void __launch_bounds__(256) k_step(StepArgs A) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= A.g.S) return;
    step_site(A, s);
}

int launch_step(glasma_ctx* ctx, bool write_old) {
    StepArgs A;
    A.U = ctx->U[ctx->cur];
    A.Un = ctx->U[ctx->cur ^ 1];
    A.E = ctx->E;
    A.J = ctx->J[ctx->cur];
    A.Jn = ctx->J[ctx->cur ^ 1];
    A.UOld = write_old ? ctx->UOld : nullptr;
    A.g = ctx->g;
    fill_step_coefficients(A, ctx->p.a, ctx->p.dt);
    const long S = ctx->g.S;
    hipLaunchKernelGGL(k_step, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, ctx->stream, A);
    ctx->cur ^= 1;
    return 0;
}

}  // namespace glasma
