// Ensembles of independent events advanced together (no counterpart in the reference, which runs one
// event per process): every event is an ordinary context -- fields, initial conditions and observables
// go through the single-event entry points -- and glasma_batch_update advances all of them with one
// kernel launch per step and step form over the combined (event, tile) work list.  The nsteps launches are captured
// once into a HIP graph per step count and replayed, so a call costs one graph launch.
//
// The per-event step arguments live in small device tables rebuilt from each context's current buffer
// parity at every call; the graph reads them at run time, so a graph stays valid when an event was
// advanced on its own in between.
#include <chrono>
#include <map>
#include <vector>

#include "internal.h"

struct glasma_batch {
    std::vector<glasma_ctx*> ev;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    glasma::StepArgs* table = nullptr;  // [phase 0/1][write_old 0/1][event]
    int* block_event = nullptr;
    unsigned* block_site = nullptr;
    long nblocks = 0;
    std::vector<long> t0;  // first block of each event
    long nrow = 0;  // the first nrow blocks: events of row_form 1 (row-staged, whole rows or one segment)
    long nseg = 0;  // the next nseg blocks: row_form 2 (row-staged, two row segments); the rest the plain form
    std::map<int, hipGraphExec_t> graphs;  // by step count
    double last_ms = 0.0;
    std::string err;
};

namespace glasma {
namespace {

int bfail(glasma_batch* b, int code, const std::string& m) {
    b->err = m;
    return code;
}

#define GB_HIP(b, expr)                                                                          \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess) return bfail(b, GLASMA_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

int upload_tables(glasma_batch* b) {
    const size_t n = b->ev.size();
    std::vector<StepArgs> h(4 * n);
    for (int ph = 0; ph < 2; ph++)
        for (int old = 0; old < 2; old++)
            for (size_t e = 0; e < n; e++) h[(2 * ph + old) * n + e] = step_args(b->ev[e], b->ev[e]->cur ^ ph, old != 0);
    GB_HIP(b, hipMemcpyAsync(b->table, h.data(), h.size() * sizeof(StepArgs), hipMemcpyHostToDevice, b->stream));
    GB_HIP(b, hipStreamSynchronize(b->stream));
    return 0;
}

int graph_for(glasma_batch* b, int nsteps, hipGraphExec_t* out) {
    auto it = b->graphs.find(nsteps);
    if (it != b->graphs.end()) {
        *out = it->second;
        return 0;
    }
    const size_t n = b->ev.size();
    GB_HIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int i = 0; i < nsteps && rc == 0; i++) {
        const int ph = i & 1, old = (i == nsteps - 1) ? 1 : 0;
        rc = launch_step_batch(b->stream, b->table + (2 * ph + old) * n, b->block_event, b->block_site, b->nrow,
                               b->nseg, b->nblocks - b->nrow - b->nseg);
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return bfail(b, rc, "kernel launch failed during graph capture");
    }
    if (e != hipSuccess) return bfail(b, GLASMA_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return bfail(b, GLASMA_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    b->graphs[nsteps] = x;
    *out = x;
    return 0;
}

// the (event, tile) work list: each event's blocks in the tile order it takes alone (its context's order)
hipError_t upload_work_list(glasma_batch* b) {
    std::vector<int> be(b->nblocks);
    std::vector<unsigned> bs(b->nblocks);
    for (size_t e = 0; e < b->ev.size(); e++) {
        const Geom& g = b->ev[e]->g;
        const TileMap tm = ctx_tile_map(b->ev[e]);
        const long ntile = pad_sites(g.S) / kTile;
        for (long k = 0; k < ntile; k++) {
            be[b->t0[e] + k] = (int)e;
            bs[b->t0[e] + k] = (unsigned)tile_site(g, tm, k);
        }
    }
    hipError_t err = hipMemcpy(b->block_event, be.data(), be.size() * sizeof(int), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(b->block_site, bs.data(), bs.size() * sizeof(unsigned), hipMemcpyHostToDevice);
    return err;
}

void advance_clocks(glasma_batch* b, int nsteps) {
    for (glasma_ctx* c : b->ev) {
        c->cur ^= (nsteps & 1);
        c->tsteps += nsteps;
        for (int i = 0; i < nsteps; i++) c->t += c->p.dt;
    }
}

}  // namespace
}  // namespace glasma

using namespace glasma;

extern "C" {

int glasma_batch_destroy(glasma_batch* b) {
    if (!b) return 0;
    if (b->stream) hipStreamSynchronize(b->stream);
    for (auto& kv : b->graphs) hipGraphExecDestroy(kv.second);
    for (glasma_ctx* c : b->ev) glasma_destroy(c);
    if (b->table) (void)hipFree(b->table);
    if (b->block_event) (void)hipFree(b->block_event);
    if (b->block_site) (void)hipFree(b->block_site);
    if (b->e0) hipEventDestroy(b->e0);
    if (b->e1) hipEventDestroy(b->e1);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

int glasma_batch_create(const glasma_params* p, int nevents, glasma_batch** out) {
    if (!p || !out || nevents <= 0) return GLASMA_ERR_ARG;
    *out = nullptr;
    for (int e = 0; e < nevents; e++)
        if ((long)p[e].n[0] * p[e].n[1] * p[e].n[2] >= (1L << 32)) return GLASMA_ERR_ARG;
    glasma_batch* b = new glasma_batch();
    for (int e = 0; e < nevents; e++) {
        glasma_ctx* c = nullptr;
        int rc = glasma_create(&p[e], &c);
        if (rc) {
            glasma_batch_destroy(b);
            return rc;
        }
        b->ev.push_back(c);
    }
    // events of row_form 1 first, then row_form 2 (both the row-staged form), then the plain form
    std::vector<long> t0(nevents);
    long nb = 0, nrow = 0, nseg = 0;
    const int order[3] = {1, 2, 0};
    for (int pass = 0; pass < 3; pass++) {
        for (int e = 0; e < nevents; e++) {
            if (row_form(b->ev[e]->g) != order[pass]) continue;
            t0[e] = nb;
            nb += pad_sites(b->ev[e]->g.S) / kTile;
        }
        if (pass == 0) nrow = nb;
        if (pass == 1) nseg = nb - nrow;
    }
    b->t0 = t0;
    b->nblocks = nb;
    b->nrow = nrow;
    b->nseg = nseg;
    hipError_t err = hipStreamCreate(&b->stream);
    if (err == hipSuccess) err = hipEventCreate(&b->e0);
    if (err == hipSuccess) err = hipEventCreate(&b->e1);
    if (err == hipSuccess) err = hipMalloc(&b->table, 4 * (size_t)nevents * sizeof(StepArgs));
    if (err == hipSuccess) err = hipMalloc(&b->block_event, (size_t)nb * sizeof(int));
    if (err == hipSuccess) err = hipMalloc(&b->block_site, (size_t)nb * sizeof(unsigned));
    if (err == hipSuccess) err = upload_work_list(b);
    if (err != hipSuccess) {
        const int code = hip_code(err);
        glasma_batch_destroy(b);
        return code;
    }
    *out = b;
    return 0;
}

int glasma_batch_size(const glasma_batch* b, int* nevents) {
    if (!b || !nevents) return GLASMA_ERR_ARG;
    *nevents = (int)b->ev.size();
    return 0;
}

int glasma_batch_context(glasma_batch* b, int e, glasma_ctx** ctx) {
    if (!b || !ctx || e < 0 || e >= (int)b->ev.size()) return GLASMA_ERR_ARG;
    *ctx = b->ev[e];
    return 0;
}

const char* glasma_batch_last_error(const glasma_batch* b) { return b ? b->err.c_str() : ""; }

int glasma_batch_update(glasma_batch* b, int nsteps) {
    if (!b || nsteps < 0) return GLASMA_ERR_ARG;
    if (nsteps == 0) return 0;
    for (glasma_ctx* c : b->ev) GB_HIP(b, hipStreamSynchronize(c->stream));
    int rc = upload_tables(b);
    if (rc) return rc;
    hipGraphExec_t x = nullptr;
    if ((rc = graph_for(b, nsteps, &x))) return rc;
    GB_HIP(b, hipEventRecord(b->e0, b->stream));
    GB_HIP(b, hipGraphLaunch(x, b->stream));
    GB_HIP(b, hipEventRecord(b->e1, b->stream));
    GB_HIP(b, hipStreamSynchronize(b->stream));
    float ms = 0.f;
    GB_HIP(b, hipEventElapsedTime(&ms, b->e0, b->e1));
    b->last_ms = ms;
    advance_clocks(b, nsteps);
    return 0;
}

int glasma_batch_update_serial(glasma_batch* b, int nsteps) {
    if (!b || nsteps < 0) return GLASMA_ERR_ARG;
    const auto t0 = std::chrono::steady_clock::now();
    for (glasma_ctx* c : b->ev) {
        int rc = glasma_update(c, nsteps);
        if (rc) return bfail(b, rc, std::string("event update failed: ") + glasma_last_error(c));
    }
    b->last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

int glasma_batch_tune(glasma_batch* b, int reps) {
    if (!b || reps <= 0) return GLASMA_ERR_ARG;
    // one tuning per distinct lattice; events of the same lattice share its order
    for (size_t e = 0; e < b->ev.size(); e++) {
        glasma_ctx* c = b->ev[e];
        bool done = false;
        for (size_t f = 0; f < e && !done; f++) {
            const Geom& g = b->ev[f]->g;
            if (g.nx == c->g.nx && g.ny == c->g.ny && g.nz == c->g.nz) {
                c->tm = b->ev[f]->tm;
                c->tm_tuned = b->ev[f]->tm_tuned;
                done = true;
            }
        }
        if (done) continue;
        const int rc = glasma_tune_step(c, reps, nullptr, nullptr, nullptr, nullptr);
        if (rc) return bfail(b, rc, std::string("tuning event failed: ") + glasma_last_error(c));
    }
    GB_HIP(b, hipStreamSynchronize(b->stream));
    GB_HIP(b, upload_work_list(b));
    return 0;
}

int glasma_batch_last_ms(const glasma_batch* b, double* ms) {
    if (!b || !ms) return GLASMA_ERR_ARG;
    *ms = b->last_ms;
    return 0;
}

}  // extern "C"
