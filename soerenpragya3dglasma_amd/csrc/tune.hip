// Tile-order autotuner of the leapfrog step (no counterpart in the reference, which has no GPU kernels).
//
// The order in which the step's blocks visit the lattice decides how much of the neighbour-link traffic
// the 8 XCD-local L2 caches absorb, and the best order depends on the lattice.  glasma_tune_step times
// the step kernel on the context's current fields under every order that fits the lattice -- XCD
// column bands of several shapes walked in z, one band of every z plane per XCD, linear -- and keeps the
// fastest for the context's later steps.  The timed launches write to scratch buffers (E is updated in
// place, so it is copied first), so the fields and the clock are unchanged; the order does not change
// any result, only the time.
#include <vector>

#include "internal.h"

using namespace glasma;

namespace {

struct Cand {
    int mode, bw, bh;
};

// the orders worth timing for this lattice, each once
std::vector<TileMap> candidates(const Geom& g) {
    static const Cand want[] = {{2, 4, 4}, {2, 2, 8}, {2, 1, 16}, {2, 1, 8}, {2, 4, 8}, {2, 2, 4}, {2, 4, 2},
                                {2, 8, 4}, {2, 8, 2}, {2, 2, 16}, {1, 0, 0}, {0, 0, 0}};
    std::vector<TileMap> out;
    for (const Cand& c : want) {
        const TileMap m = tile_map(g, c.mode, c.bw, c.bh);
        if (m.mode != c.mode) continue;  // the lattice does not fit this order
        bool dup = false;
        for (const TileMap& o : out) dup = dup || (o.mode == m.mode && o.bw == m.bw && o.bh == m.bh);
        if (!dup) out.push_back(m);
    }
    return out;
}

int code_of(const TileMap& m) { return m.mode * 10000 + m.bw * 100 + m.bh; }

}  // namespace

extern "C" {

int glasma_tune_step(glasma_ctx* ctx, int reps, int* ncand, int* codes, double* ms, int* best) {
    if (!ctx || reps <= 0) return GLASMA_ERR_ARG;
    const std::vector<TileMap> cand = candidates(ctx->g);
    if (cand.size() > (size_t)GLASMA_TUNE_MAX) return set_error(ctx, GLASMA_ERR_ARG, "too many tile orders");
    const size_t Sp = pad_sites(ctx->g.S);
    double *sU = nullptr, *sE = nullptr, *sJ = nullptr;
    hipError_t e = hipMalloc(&sU, (size_t)kUPlanes * Sp * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&sE, (size_t)kEPlanes * Sp * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&sJ, (size_t)kJPlanes * Sp * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpyAsync(sE, ctx->E, (size_t)kEPlanes * Sp * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream);
    std::vector<double> t(cand.size(), 0.0);
    StepArgs A = step_args(ctx, ctx->cur, false);
    A.Un = sU;
    A.E = sE;
    A.Jn = sJ;
    A.UOld = nullptr;
    for (size_t i = 0; i < cand.size() && e == hipSuccess; i++) {
        launch_step_map(ctx, A, cand[i]);  // warm-up
        e = hipEventRecord(ctx->ev0, ctx->stream);
        for (int r = 0; r < reps && e == hipSuccess; r++) launch_step_map(ctx, A, cand[i]);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
        float f = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&f, ctx->ev0, ctx->ev1);
        t[i] = f / reps;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(sU);
    hipFree(sE);
    hipFree(sJ);
    if (e != hipSuccess) {
        const int code = hip_code(e);
        return set_error(ctx, code, std::string("glasma_tune_step: ") + hipGetErrorString(e));
    }
    size_t bi = 0;
    for (size_t i = 1; i < cand.size(); i++)
        if (t[i] < t[bi]) bi = i;
    ctx->tm = cand[bi];
    ctx->tm_tuned = true;
    if (ncand) *ncand = (int)cand.size();
    for (size_t i = 0; i < cand.size(); i++) {
        if (codes) codes[i] = code_of(cand[i]);
        if (ms) ms[i] = t[i];
    }
    if (best) *best = (int)bi;
    return 0;
}

int glasma_tile_order(const glasma_ctx* ctx, int* code, int* tuned) {
    if (!ctx || !code) return GLASMA_ERR_ARG;
    *code = code_of(ctx_tile_map(ctx));
    if (tuned) *tuned = ctx->tm_tuned ? 1 : 0;
    return 0;
}

}  // extern "C"
