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
                                {2, 8, 4}, {2, 8, 2}, {2, 2, 16}, {1, 0, 0}, {3, 0, 0}, {0, 0, 0}};
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

// The step's compulsory traffic alone: one 64-site tile per block of 256 threads, each block reading the tile's
// U (27 plane pairs), E (12) and currents (16) once with non-temporal loads, all issued before the first store,
// and writing them once to the destination arrays with non-temporal stores (880 + 880 bytes per site, the same
// tiled layout and the same grid as the step).
typedef double cp2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ cp2 cp_ld(const cp2* p) {
    return __builtin_nontemporal_load((const __attribute__((address_space(1))) cp2*)p);
}
__device__ __forceinline__ void cp_st(cp2* p, cp2 v) {
    __builtin_nontemporal_store(v, (__attribute__((address_space(1))) cp2*)p);
}
constexpr int kUPairs = kUPlanes / 2 * kTile, kEPairs = kEPlanes / 2 * kTile, kJPairs = kJPlanes / 2 * kTile;
__global__ void __launch_bounds__(256) k_tile_copy(const cp2* __restrict__ U, const cp2* __restrict__ E,
                                                   const cp2* __restrict__ J, cp2* __restrict__ Uo,
                                                   cp2* __restrict__ Eo, cp2* __restrict__ Jo) {
    static_assert(kUPairs == 6 * 256 + 192 && kEPairs == 3 * 256 && kJPairs == 4 * 256, "tile sizes");
    const long t = blockIdx.x;
    const int i = threadIdx.x;
    const cp2* u = U + t * kUPairs;
    const cp2* e = E + t * kEPairs;
    const cp2* j = J + t * kJPairs;
    cp2 r[14];
#pragma unroll
    for (int k = 0; k < 6; k++) r[k] = cp_ld(u + k * 256 + i);
    r[6] = i < 192 ? cp_ld(u + 1536 + i) : cp2{0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 3; k++) r[7 + k] = cp_ld(e + k * 256 + i);
#pragma unroll
    for (int k = 0; k < 4; k++) r[10 + k] = cp_ld(j + k * 256 + i);
    cp2* uo = Uo + t * kUPairs;
    cp2* eo = Eo + t * kEPairs;
    cp2* jo = Jo + t * kJPairs;
#pragma unroll
    for (int k = 0; k < 6; k++) cp_st(uo + k * 256 + i, r[k]);
    if (i < 192) cp_st(uo + 1536 + i, r[6]);
#pragma unroll
    for (int k = 0; k < 3; k++) cp_st(eo + k * 256 + i, r[7 + k]);
#pragma unroll
    for (int k = 0; k < 4; k++) cp_st(jo + k * 256 + i, r[10 + k]);
}

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
    (void)hipFree(sU);
    (void)hipFree(sE);
    (void)hipFree(sJ);
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

int glasma_copy_ceiling(glasma_ctx* ctx, int reps, double* ms) {
    if (!ctx || reps <= 0 || !ms) return GLASMA_ERR_ARG;
    const size_t Sp = pad_sites(ctx->g.S);
    const long tiles = (long)(Sp / kTile);
    if (tiles > 0x7fffffffL) return set_error(ctx, GLASMA_ERR_ARG, "glasma_copy_ceiling: too many tiles");
    double *sU = nullptr, *sE = nullptr, *sJ = nullptr;
    hipError_t e = hipMalloc(&sU, (size_t)kUPlanes * Sp * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&sE, (size_t)kEPlanes * Sp * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&sJ, (size_t)kJPlanes * Sp * sizeof(double));
    const cp2* U = reinterpret_cast<const cp2*>(ctx->U[ctx->cur]);
    const cp2* E = reinterpret_cast<const cp2*>(ctx->E);
    const cp2* J = reinterpret_cast<const cp2*>(ctx->J[ctx->cur]);
    float f = 0.f;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_tile_copy, dim3((unsigned)tiles), dim3(256), 0, ctx->stream, U, E, J,
                           reinterpret_cast<cp2*>(sU), reinterpret_cast<cp2*>(sE), reinterpret_cast<cp2*>(sJ));
        e = hipEventRecord(ctx->ev0, ctx->stream);
    }
    for (int r = 0; r < reps && e == hipSuccess; r++)
        hipLaunchKernelGGL(k_tile_copy, dim3((unsigned)tiles), dim3(256), 0, ctx->stream, U, E, J,
                           reinterpret_cast<cp2*>(sU), reinterpret_cast<cp2*>(sE), reinterpret_cast<cp2*>(sJ));
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    if (e == hipSuccess) e = hipEventElapsedTime(&f, ctx->ev0, ctx->ev1);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(sU);
    (void)hipFree(sE);
    (void)hipFree(sJ);
    if (e != hipSuccess) return set_error(ctx, hip_code(e), std::string("glasma_copy_ceiling: ") + hipGetErrorString(e));
    *ms = f / reps;
    return 0;
}

int glasma_tile_order(const glasma_ctx* ctx, int* code, int* tuned) {
    if (!ctx || !code) return GLASMA_ERR_ARG;
    *code = code_of(ctx_tile_map(ctx));
    if (tuned) *tuned = ctx->tm_tuned ? 1 : 0;
    return 0;
}

}  // extern "C"
