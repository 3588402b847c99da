// Reverse-mode derivatives on the device (no counterpart in the reference, which has no gradient code):
// the leapfrog step (two passes per step, no atomics: A with B folded in, then the staple gather), the slab
// energy, and the whole map charge density -> left-moving MV sheet -> nsteps leapfrog steps -> slab energy, with
// the forward run checkpointed every ckpt steps and each segment but the last recomputed once during the backward
// sweep.
#include <cstdlib>
#include <cstring>
#include <vector>

#include "internal.h"
#include "adjoint.h"

namespace glasma {
namespace {

constexpr int kAdjBlock = 256;
inline dim3 agrid(long n) { return dim3((unsigned)((n + kAdjBlock - 1) / kAdjBlock)); }

// Block order of the backward passes.  nd directions per site block (1, or 3 for the staple passes, which give each
// thread one link direction); block b runs on XCD b % 8.  With the site blocks a multiple of 8, XCD k takes site
// blocks [k nsb/8, (k+1) nsb/8) and, for each, its nd directions in turn: a site's y and z neighbours (a row and a
// plane away) are then read by the same XCD at about the same time, and the three directions of a pass share them in
// its L2.  In dispatch order (all sites of direction 0, then 1, then 2) the staple gather read 7 times its
// compulsory bytes from beyond L2.  Otherwise the directions still alternate, without the XCD ranges.
__device__ __forceinline__ void adj_block(long nsb, int nd, long& sb, int& dir) {
    const long b = blockIdx.x;
    if (nsb % 8 == 0) {
        const long i = b >> 3;
        dir = (int)(i % nd);
        sb = (b & 7) * (nsb >> 3) + i / nd;
    } else {
        dir = (int)(b % nd);
        sb = b / nd;
    }
}
inline dim3 adj_grid(long n, int nd) { return dim3((unsigned)(nd * ((n + kAdjBlock - 1) / kAdjBlock))); }

struct AtomicAddU {
    double* Ub;
    __device__ void operator()(int mu, unsigned s, const M3& M) const {
        double* q = Ub + lidx<unsigned>(kUPlanes, 18 * mu, s);
        for (int k = 0; k < 9; k++) {
            unsafeAtomicAdd(q + k * 2 * kTile, M.re[k]);
            unsafeAtomicAdd(q + k * 2 * kTile + 1, M.im[k]);
        }
    }
};

// pass A with pass B's terms and the Sigbar store folded in (step_vjp_local<I, true>)
__global__ void __launch_bounds__(kAdjBlock) k_vjp_local(StepAdjArgs A, long nsb) {
    long sb;
    int dir;
    adj_block(nsb, 1, sb, dir);
    const long s = sb * kAdjBlock + threadIdx.x;
    if (s < A.g.S) step_vjp_local<unsigned, true>(A, (unsigned)s);
}
// The staple pass gives each thread one link direction (a third of the work per thread; the three directions run
// in one launch): the own-link term and the twelve staple terms of the neighbours' E updates.
// Measured at 48 x 48 x 128: 283 us at one wave per SIMD; the interior layers at two waves (252 registers) in one
// launch and the two open-z boundary layers in a second took 274 + 33 us.
__global__ void __launch_bounds__(kAdjBlock) k_vjp_gather(StepAdjArgs A, long nsb) {
    long sb;
    int dir;
    adj_block(nsb, 3, sb, dir);
    const long s = sb * kAdjBlock + threadIdx.x;
    if (s >= A.g.S) return;
    switch (dir) {
        case 0: step_vjp_gather_link<unsigned, 0, true>(A, (unsigned)s); break;
        case 1: step_vjp_gather_link<unsigned, 1, true>(A, (unsigned)s); break;
        default: step_vjp_gather_link<unsigned, 2, true>(A, (unsigned)s); break;
    }
}

// slab energy: per-block partial sums, then one block sums them in a fixed order
__global__ void __launch_bounds__(kAdjBlock) k_slab_energy(SlabArgs A, double* part) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double v = s < A.g.S ? slab_energy_site<unsigned>(A, (unsigned)s) : 0.0;
    __shared__ double sh[kAdjBlock / 64];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kAdjBlock / 64; w++) t += sh[w];
        part[blockIdx.x] = t;
    }
}
__global__ void __launch_bounds__(kAdjBlock) k_sum_parts(const double* part, long n, double* out) {
    double v = 0.0;
    for (long i = threadIdx.x; i < n; i += blockDim.x) v += part[i];
    __shared__ double sh[kAdjBlock / 64];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kAdjBlock / 64; w++) t += sh[w];
        *out = t;
    }
}
__global__ void __launch_bounds__(kAdjBlock) k_slab_vjp(SlabArgs A, double* Ub) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < A.g.S) slab_energy_vjp_site<unsigned>(A, (unsigned)s, 1.0, AtomicAddU{Ub});
}
__global__ void k_fill0(double* p, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

SlabArgs slab_args(const glasma_ctx* ctx, int zc, int half) {
    SlabArgs A;
    A.U = ctx->U[ctx->cur];
    A.E = ctx->E;
    A.Eb = nullptr;
    A.g = ctx->g;
    const double* a = ctx->p.a;
    const double acube = a[0] * a[1] * a[2];
    for (int m = 0; m < 3; m++) A.c[m] = (a[m] / acube) * (a[m] / acube);
    A.zc = zc;
    A.half = half;
    return A;
}

// slab energy of the fields U, E (the context's current fields when null)
int slab_energy(glasma_ctx* ctx, int zc, int half, double* out, const double* U = nullptr, const double* E = nullptr) {
    const long nb = (ctx->g.S + kAdjBlock - 1) / kAdjBlock;
    int rc = ensure_red(ctx, (size_t)nb + 1);
    if (rc) return rc;
    SlabArgs A = slab_args(ctx, zc, half);
    if (U) A.U = U;
    if (E) A.E = E;
    hipLaunchKernelGGL(k_slab_energy, dim3((unsigned)nb), dim3(kAdjBlock), 0, ctx->stream, A, ctx->red);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kAdjBlock), 0, ctx->stream, ctx->red, nb, ctx->red + nb);
    GL_HIP(ctx, hipMemcpyAsync(out, ctx->red + nb, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    GL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

StepAdjArgs adj_args(const glasma_ctx* ctx) {
    StepArgs F = step_args(ctx, ctx->cur, false);
    StepAdjArgs A;
    std::memset(&A, 0, sizeof(A));
    A.g = ctx->g;
    for (int m = 0; m < 3; m++) {
        A.gc[m] = F.gc[m];
        A.cE[m] = F.cE[m];
        A.cU[m] = F.cU[m];
    }
    A.dtz = F.dtz;
    return A;
}

// one backward step in two launches (A with B folded in, then C2 with C1's own-link term)
int run_vjp(glasma_ctx* ctx, const StepAdjArgs& A) {
    const long S = ctx->g.S;
    const long nsb = (S + kAdjBlock - 1) / kAdjBlock;
    hipLaunchKernelGGL(k_vjp_local, adj_grid(S, 1), dim3(kAdjBlock), 0, ctx->stream, A, nsb);
    hipLaunchKernelGGL(k_vjp_gather, adj_grid(S, 3), dim3(kAdjBlock), 0, ctx->stream, A, nsb);
    GL_HIP(ctx, hipGetLastError());
    return 0;
}

// device buffers of one field state (tiled U, E, J) and helpers to move the context's state in and out
struct State {
    double *U, *E, *J;
};
int save_state(glasma_ctx* ctx, const State& s) {
    const size_t Sp = pad_sites(ctx->g.S);
    GL_HIP(ctx, hipMemcpyAsync(s.U, ctx->U[ctx->cur], kUPlanes * Sp * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    GL_HIP(ctx, hipMemcpyAsync(s.E, ctx->E, kEPlanes * Sp * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    GL_HIP(ctx, hipMemcpyAsync(s.J, ctx->J[ctx->cur], kJPlanes * Sp * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return 0;
}
int load_state(glasma_ctx* ctx, const State& s) {
    const size_t Sp = pad_sites(ctx->g.S);
    GL_HIP(ctx, hipMemcpyAsync(ctx->U[ctx->cur], s.U, kUPlanes * Sp * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    GL_HIP(ctx, hipMemcpyAsync(ctx->E, s.E, kEPlanes * Sp * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    GL_HIP(ctx, hipMemcpyAsync(ctx->J[ctx->cur], s.J, kJPlanes * Sp * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return 0;
}

// steps of one checkpoint segment from seg[0], each state into the next slot (U, E and J out of place) and U_h
// into Uh[i], on stream st (the context's stream when null)
int run_segment(glasma_ctx* ctx, const std::vector<State>& seg, const std::vector<double*>& Uh, int len,
                hipStream_t st = nullptr) {
    for (int i = 0; i < len; i++) {
        StepArgs F = step_args(ctx, ctx->cur, true);
        F.U = seg[i].U;
        F.Un = seg[i + 1].U;
        F.E = seg[i].E;
        F.En = seg[i + 1].E;
        F.J = seg[i].J;
        F.Jn = seg[i + 1].J;
        F.UOld = Uh[i];
        const int rc = launch_step_args(ctx, F, st);
        if (rc) return rc;
    }
    return 0;
}

// the gradient's second stream and its events (default priority, kept in the context)
int ensure_seg_stream(glasma_ctx* ctx) {
    if (!ctx->stream2) GL_HIP(ctx, hipStreamCreate(&ctx->stream2));
    for (hipEvent_t& ev : ctx->seg_ev)
        if (!ev) GL_HIP(ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    return 0;
}

// device work memory of one call from the context's pool: the i-th request of a call takes the pool's i-th array when
// it is large enough, else that array is replaced.  A call of the same sizes as the previous one allocates nothing
// (each gradient call at 48 x 48 x 128 otherwise made about a hundred allocations of 4.6 GB in all, 45 ms).
struct DevBuf {
    size_t next = 0;
    double* get(glasma_ctx* ctx, size_t doubles, int* rc) {
        auto& pool = ctx->wpool;
        // a gradient call that returned early on an error can leave a recompute queued on the second stream, still
        // writing into the pool's arrays: wait for it before this call hands them out again
        if (next == 0 && ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
        if (next < pool.size() && pool[next].first >= doubles) return pool[next++].second;
        double* q = nullptr;
        const hipError_t e = hipMalloc(&q, doubles * sizeof(double));
        if (e != hipSuccess) {
            *rc = set_error(ctx, hip_code(e), "device allocation failed in the adjoint");
            return nullptr;
        }
        if (next < pool.size()) {
            (void)hipStreamSynchronize(ctx->stream);  // the old array may still be read by queued work
            if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
            (void)hipFree(pool[next].second);
            pool[next] = {doubles, q};
        } else {
            pool.push_back({doubles, q});
        }
        return pool[next++].second;
    }
};

// host charge density (a + 8*Index3D with ghosts) <-> device (a*S + s)
void rho_to_dev_layout(const glasma_ctx* ctx, const double* h, std::vector<double>& out) {
    const long S = ctx->g.S, P = ctx->g.P;
    out.resize(8 * (size_t)S);
    for (long s = 0; s < S; s++)
        for (int a = 0; a < 8; a++) out[(size_t)a * S + s] = h[8 * (s + P) + a];
}
void rho_from_dev_layout(const glasma_ctx* ctx, const std::vector<double>& in, double* h) {
    const long S = ctx->g.S, P = ctx->g.P;
    std::memset(h, 0, sizeof(double) * 8 * (size_t)(S + 2 * P));
    for (long s = 0; s < S; s++)
        for (int a = 0; a < 8; a++) h[8 * (s + P) + a] = in[(size_t)a * S + s];
}

bool mv_ok(const glasma_mv_params* mv) { return mv && mv->lambda > 0.0 && mv->m_eff >= 0.0; }

}  // namespace
}  // namespace glasma

using namespace glasma;

extern "C" {

int glasma_slab_energy(glasma_ctx* ctx, int zc, int half, double* energy) {
    if (!ctx || !energy || half < 0) return GLASMA_ERR_ARG;
    return slab_energy(ctx, zc, half, energy);
}

int glasma_step_vjp(glasma_ctx* ctx, double* Ub, double* Eb, double* Jpb, double* Jmb, double* JpSb, double* JmSb) {
    if (!ctx || !Ub || !Eb || !Jpb || !Jmb || !JpSb || !JmSb) return GLASMA_ERR_ARG;
    const size_t Sp = pad_sites(ctx->g.S);
    int rc = 0;
    DevBuf b;
    double* s0U = b.get(ctx, kUPlanes * Sp, &rc);
    double* s0E = rc ? nullptr : b.get(ctx, kEPlanes * Sp, &rc);
    double* s0J = rc ? nullptr : b.get(ctx, kJPlanes * Sp, &rc);
    double* cU1 = rc ? nullptr : b.get(ctx, kUPlanes * Sp, &rc);
    double* cU0 = rc ? nullptr : b.get(ctx, kUPlanes * Sp, &rc);
    double* cE = rc ? nullptr : b.get(ctx, kEPlanes * Sp, &rc);
    double* cJ1 = rc ? nullptr : b.get(ctx, kJPlanes * Sp, &rc);
    double* cJ0 = rc ? nullptr : b.get(ctx, kJPlanes * Sp, &rc);
    if (rc) return rc;
    GL_HIP(ctx, hipMemsetAsync(cU1, 0, kUPlanes * Sp * sizeof(double), ctx->stream));
    GL_HIP(ctx, hipMemsetAsync(cE, 0, kEPlanes * Sp * sizeof(double), ctx->stream));
    GL_HIP(ctx, hipMemsetAsync(cJ1, 0, kJPlanes * Sp * sizeof(double), ctx->stream));
    if ((rc = upload_aos(ctx, Ub, cU1, kUPlanes))) return rc;
    if ((rc = upload_aos(ctx, Eb, cE, kEPlanes))) return rc;
    const double* jin[4] = {Jpb, Jmb, JpSb, JmSb};
    for (int w = 0; w < 4; w++)
        if ((rc = upload_aos(ctx, jin[w], cJ1, 8, kJPlanes, 8 * w))) return rc;
    const State s0 = {s0U, s0E, s0J};
    if ((rc = save_state(ctx, s0))) return rc;
    const double t = ctx->t;
    const long ts = ctx->tsteps;
    launch_step(ctx, true);
    StepAdjArgs A = adj_args(ctx);
    A.U = s0U;
    A.Uh = ctx->UOld;
    A.E1 = ctx->E;
    A.J = s0J;
    A.J1 = ctx->J[ctx->cur];
    A.Ub1 = cU1;
    A.Ub = cU0;
    A.Eb = cE;
    A.Jb1 = cJ1;
    A.Jb = cJ0;
    if ((rc = run_vjp(ctx, A))) return rc;
    // back to the state at t
    ctx->cur ^= 1;
    if ((rc = load_state(ctx, s0))) return rc;
    ctx->t = t;
    ctx->tsteps = ts;
    GL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if ((rc = download_aos(ctx, cU0, Ub, kUPlanes))) return rc;
    if ((rc = download_aos(ctx, cE, Eb, kEPlanes))) return rc;
    double* jout[4] = {Jpb, Jmb, JpSb, JmSb};
    for (int w = 0; w < 4; w++)
        if ((rc = download_aos(ctx, cJ0, jout[w], 8, kJPlanes, 8 * w))) return rc;
    return 0;
}

int glasma_slab_energy_from_charge(glasma_ctx* ctx, const glasma_mv_params* mv, const double* rho, int nsteps, int zc,
                                   int half, double* energy) {
    if (!ctx || !rho || !energy || nsteps < 0 || half < 0 || !mv_ok(mv)) return GLASMA_ERR_ARG;
    const long S = ctx->g.S, Sp = pad_sites(S);
    int rc = 0;
    DevBuf b;
    IcTape tape;
    double* rho_d = b.get(ctx, 8 * S, &rc);
    tape.rho_s = rc ? nullptr : b.get(ctx, 8 * S, &rc);
    tape.A0 = rc ? nullptr : b.get(ctx, 8 * S, &rc);
    tape.A1 = rc ? nullptr : b.get(ctx, 8 * S, &rc);
    tape.V0 = rc ? nullptr : b.get(ctx, 18 * Sp, &rc);
    tape.V1 = rc ? nullptr : b.get(ctx, 18 * Sp, &rc);
    if (rc) return rc;
    std::vector<double> h;
    rho_to_dev_layout(ctx, rho, h);
    GL_HIP(ctx, hipMemcpy(rho_d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    if ((rc = ic_left_from_charge(ctx, mv, rho_d, tape))) return rc;
    for (int n = 0; n < nsteps; n++) {
        launch_step(ctx, n == nsteps - 1);
        ctx->tsteps++;
        ctx->t += ctx->p.dt;
    }
    GL_HIP(ctx, hipGetLastError());
    return slab_energy(ctx, zc, half, energy);
}

int glasma_slab_energy_gradient(glasma_ctx* ctx, const glasma_mv_params* mv, const double* rho, int nsteps, int ckpt,
                                int zc, int half, double* energy, double* grad_rho) {
    if (!ctx || !rho || !energy || !grad_rho || nsteps < 0 || ckpt <= 0 || half < 0 || !mv_ok(mv)) return GLASMA_ERR_ARG;
    const long S = ctx->g.S, Sp = pad_sites(S);
    const int nseg = (nsteps + ckpt - 1) / ckpt;
    const int seglen = nsteps < ckpt ? nsteps : ckpt;
    int rc = 0;
    DevBuf b;
    IcTape tape;
    double* rho_d = b.get(ctx, 8 * S, &rc);
    tape.rho_s = rc ? nullptr : b.get(ctx, 8 * S, &rc);
    tape.A0 = rc ? nullptr : b.get(ctx, 8 * S, &rc);
    tape.A1 = rc ? nullptr : b.get(ctx, 8 * S, &rc);
    tape.V0 = rc ? nullptr : b.get(ctx, 18 * Sp, &rc);
    tape.V1 = rc ? nullptr : b.get(ctx, 18 * Sp, &rc);
    // checkpoints at the starts of all segments but the last, whose states the forward run leaves in the segment
    // slots; the slots in two sets (segment sg uses set sg % 2), so that one segment's recompute, on the second
    // stream, overlaps the backward sweep of the segment after it, when the second set takes at most a quarter of the
    // device memory (else one set, each segment recomputed just before its sweep)
    int nset = 1;
    if (nseg > 1) {
        size_t freeb = 0, totalb = 0;
        const size_t set_bytes = (size_t)seglen * (2 * kUPlanes + kEPlanes + kJPlanes) * Sp * sizeof(double);
        if (hipMemGetInfo(&freeb, &totalb) == hipSuccess && set_bytes <= totalb / 4) nset = 2;
        const char* env = std::getenv("GLASMA_GRAD_SETS");  // 1: one set always (read on every call)
        if (env && std::atoi(env) == 1) nset = 1;
    }
    // ck[0] is the initial state in the context's fields, which no step writes; tmp: the forward run's other states
    std::vector<State> ck(nseg > 1 ? nseg - 1 : 0), seg[2], tmp(nseg > 1 ? 2 : 0);
    std::vector<double*> Uh[2];
    for (size_t k = 1; k < ck.size(); k++) {
        ck[k].U = rc ? nullptr : b.get(ctx, kUPlanes * Sp, &rc);
        ck[k].E = rc ? nullptr : b.get(ctx, kEPlanes * Sp, &rc);
        ck[k].J = rc ? nullptr : b.get(ctx, kJPlanes * Sp, &rc);
    }
    for (auto& s : tmp) {
        s.U = rc ? nullptr : b.get(ctx, kUPlanes * Sp, &rc);
        s.E = rc ? nullptr : b.get(ctx, kEPlanes * Sp, &rc);
        s.J = rc ? nullptr : b.get(ctx, kJPlanes * Sp, &rc);
    }
    for (int q = 0; q < nset; q++) {
        seg[q].resize(seglen + 1);
        for (size_t i = 1; i < seg[q].size(); i++) {  // seg[q][0] is the segment's first state itself
            seg[q][i].U = rc ? nullptr : b.get(ctx, kUPlanes * Sp, &rc);
            seg[q][i].E = rc ? nullptr : b.get(ctx, kEPlanes * Sp, &rc);
            seg[q][i].J = rc ? nullptr : b.get(ctx, kJPlanes * Sp, &rc);
        }
        Uh[q].resize(seglen > 0 ? seglen : 0);
        for (auto& p : Uh[q]) p = rc ? nullptr : b.get(ctx, kUPlanes * Sp, &rc);
    }
    double* cU[2];
    double* cJ[2];
    cU[0] = rc ? nullptr : b.get(ctx, kUPlanes * Sp, &rc);
    cU[1] = rc ? nullptr : b.get(ctx, kUPlanes * Sp, &rc);
    cJ[0] = rc ? nullptr : b.get(ctx, kJPlanes * Sp, &rc);
    cJ[1] = rc ? nullptr : b.get(ctx, kJPlanes * Sp, &rc);
    double* cE = rc ? nullptr : b.get(ctx, kEPlanes * Sp, &rc);
    double* rho_bar = rc ? nullptr : b.get(ctx, 8 * S, &rc);
    double* work = rc ? nullptr : b.get(ctx, 54 * Sp + 36 * Sp + 32 * S, &rc);
    if (rc) return rc;
    if (nset > 1 && (rc = ensure_seg_stream(ctx))) return rc;
    std::vector<double> h;
    rho_to_dev_layout(ctx, rho, h);
    GL_HIP(ctx, hipMemcpy(rho_d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    // forward, every step out of place: the segments before the last from the initial state in the context's fields
    // (left as they are), each checkpoint written by the step that reaches it and the other states alternating between
    // the two tmp states; then the last segment into the segment slots, so that its backward sweep recomputes nothing
    if ((rc = ic_left_from_charge(ctx, mv, rho_d, tape))) return rc;
    const int nlast = nseg > 0 ? (nseg - 1) * ckpt : 0;
    State fin = {ctx->U[ctx->cur], ctx->E, ctx->J[ctx->cur]};
    if (nseg > 1) ck[0] = fin;
    for (int n = 0; n < nlast; n++) {
        const State nx = ((n + 1) % ckpt == 0 && n + 1 < nlast) ? ck[(n + 1) / ckpt] : tmp[n & 1];
        StepArgs F = step_args(ctx, ctx->cur, false);
        F.U = fin.U;
        F.Un = nx.U;
        F.E = fin.E;
        F.En = nx.E;
        F.J = fin.J;
        F.Jn = nx.J;
        if ((rc = launch_step_args(ctx, F))) return rc;
        fin = nx;
    }
    if (nseg > 0) {
        const int q = (nseg - 1) % nset;
        seg[q][0] = fin;
        if ((rc = run_segment(ctx, seg[q], Uh[q], nsteps - nlast))) return rc;
        fin = seg[q][nsteps - nlast];
    }
    GL_HIP(ctx, hipGetLastError());
    if ((rc = slab_energy(ctx, zc, half, energy, fin.U, fin.E))) return rc;
    // cotangent of the final state
    hipLaunchKernelGGL(k_fill0, agrid(kUPlanes * Sp), dim3(kAdjBlock), 0, ctx->stream, cU[0], kUPlanes * Sp);
    hipLaunchKernelGGL(k_fill0, agrid(kJPlanes * Sp), dim3(kAdjBlock), 0, ctx->stream, cJ[0], kJPlanes * Sp);
    {
        SlabArgs A = slab_args(ctx, zc, half);
        A.U = fin.U;
        A.E = fin.E;
        A.Eb = cE;
        hipLaunchKernelGGL(k_slab_vjp, agrid(S), dim3(kAdjBlock), 0, ctx->stream, A, cU[0]);
    }
    // events: seg_ev[q] the recompute into set q done (second stream), seg_ev[2 + q] the sweep over set q done (the
    // context's stream), seg_ev[4] the forward run done (the checkpoints written)
    hipEvent_t* ev = ctx->seg_ev;
    if (nset > 1) {
        GL_HIP(ctx, hipEventRecord(ev[4], ctx->stream));
        GL_HIP(ctx, hipStreamWaitEvent(ctx->stream2, ev[4], 0));
    }
    int cb = 0;
    // backward sweep, one segment at a time; with two sets the recompute of segment sg - 1 is queued on the second
    // stream first
    for (int sg = nseg - 1; sg >= 0; sg--) {
        const int n0 = sg * ckpt, n1 = (n0 + ckpt < nsteps) ? n0 + ckpt : nsteps;
        const int q = sg % nset;
        if (nset == 1 && sg < nseg - 1) {
            seg[0][0] = ck[sg];
            if ((rc = run_segment(ctx, seg[0], Uh[0], ckpt))) return rc;
        }
        if (nset > 1 && sg >= 1) {
            const int q1 = (sg - 1) & 1;  // the set the sweep of segment sg + 1 read
            if (sg + 1 <= nseg - 1) GL_HIP(ctx, hipStreamWaitEvent(ctx->stream2, ev[2 + q1], 0));
            seg[q1][0] = ck[sg - 1];
            if ((rc = run_segment(ctx, seg[q1], Uh[q1], ckpt, ctx->stream2))) return rc;
            GL_HIP(ctx, hipEventRecord(ev[q1], ctx->stream2));
        }
        if (nset > 1 && sg < nseg - 1) GL_HIP(ctx, hipStreamWaitEvent(ctx->stream, ev[q], 0));
        for (int n = n1 - 1; n >= n0; n--) {
            const int i = n - n0;
            StepAdjArgs A = adj_args(ctx);
            A.U = seg[q][i].U;
            A.Uh = Uh[q][i];
            A.E1 = seg[q][i + 1].E;
            A.J = seg[q][i].J;
            A.J1 = seg[q][i + 1].J;
            A.Ub1 = cU[cb];
            A.Ub = cU[cb ^ 1];
            A.Eb = cE;
            A.Jb1 = cJ[cb];
            A.Jb = cJ[cb ^ 1];
            if ((rc = run_vjp(ctx, A))) return rc;
            cb ^= 1;
        }
        if (nset > 1 && sg >= 2) GL_HIP(ctx, hipEventRecord(ev[2 + q], ctx->stream));  // set q: next for sg - 2
    }
    // the context's fields and UOld still hold the initial state (the linearisation point of the initial-condition
    // map): no step above wrote them
    if ((rc = ic_left_vjp(ctx, mv, tape, cU[cb], cE, cJ[cb], rho_bar, work))) return rc;
    GL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    GL_HIP(ctx, hipMemcpy(h.data(), rho_bar, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    rho_from_dev_layout(ctx, h, grad_rho);
    return 0;
}

}  // extern "C"
