// ab_graph.hip -- several contexts' steps captured into one HIP graph (no
// counterpart in the reference, whose time loop launches ~20 kernels per
// step from the host with three host synchronisations).
//
// The capture forks from a private stream into every context's stream and
// joins back: each instance is an independent branch of the graph, so a
// replay runs the instances concurrently and a small instance no longer
// pays a host launch per kernel. The captured step is exactly the eager
// step (same kernels, arguments and order), so a replay gives bit-identical
// results; what the capture cannot hold is refused up front: a step with
// host synchronisation (stats, profiling), training-row recording, or a
// table pipeline whose coefficients would be rebuilt for a new dt.
#include "ab_internal.hpp"

#include <new>
#include <string>
#include <vector>

struct ab_graph
{
    std::vector<ab_ctx *> ctxs;
    std::vector<unsigned long long> gens;   // each context's generation at capture
    std::vector<unsigned long long> tgens;  // and its 2-D table arrays' generation
    double dt = 0.0;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> events;         // the capture's fork and join
    std::vector<hipEvent_t> order;          // a launch's ordering against the contexts' streams
};

namespace ab {
int graph_step(ab_ctx *c, double dt);   // ab_api.hip: the eager step, no host work
int graph_prepare(ab_ctx *c, double dt);
int graph_coefficients(ab_ctx *c, double dt);
}  // namespace ab

namespace {

void graph_free(ab_graph *g)
{
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    for (auto e : g->events) if (e) (void)hipEventDestroy(e);
    for (auto e : g->order) if (e) (void)hipEventDestroy(e);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

}  // namespace

extern "C" {

int ab_graph_create(ab_ctx *const *ctxs, int n, double dt, int nsteps, ab_graph **out)
{
    if (!ctxs || n <= 0 || nsteps <= 0 || !out || !(dt > 0.0)) return AB_ERR_ARG;
    *out = nullptr;
    for (int k = 0; k < n; ++k)
    {
        if (!ctxs[k]) return AB_ERR_ARG;
        for (int q = 0; q < k; ++q) if (ctxs[q] == ctxs[k]) return AB_ERR_ARG;
        int rc = ab::graph_prepare(ctxs[k], dt);
        if (rc) return rc;
    }
    ab_graph *g = new (std::nothrow) ab_graph;
    if (!g) return AB_ERR_NOMEM;
    g->ctxs.assign(ctxs, ctxs + n);
    for (int k = 0; k < n; ++k)
    {
        g->gens.push_back(ctxs[k]->gen);
        g->tgens.push_back(ctxs[k]->tgen);
    }
    g->dt = dt;
    g->events.assign(n + 1, nullptr);
    g->order.assign(n + 1, nullptr);
    bool ok = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) == hipSuccess;
    for (auto &e : g->events)
        ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    for (auto &e : g->order)
        ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    // everything enqueued so far completes before the capture starts
    for (int k = 0; k < n && ok; ++k) ok = hipStreamSynchronize(ctxs[k]->stream) == hipSuccess;
    if (!ok) { graph_free(g); return AB_ERR_HIP; }

    int rc = AB_OK;
    ok = hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal) == hipSuccess &&
         hipEventRecord(g->events[n], g->stream) == hipSuccess;
    for (int k = 0; k < n && ok; ++k)
    {
        ab_ctx *c = ctxs[k];
        ok = hipStreamWaitEvent(c->stream, g->events[n], 0) == hipSuccess;
        for (int s = 0; s < nsteps && ok && !rc; ++s) rc = ab::graph_step(c, dt);
        ok = ok && !rc && hipEventRecord(g->events[k], c->stream) == hipSuccess &&
             hipStreamWaitEvent(g->stream, g->events[k], 0) == hipSuccess;
    }
    hipGraph_t graph = nullptr;
    const bool ended = hipStreamEndCapture(g->stream, &graph) == hipSuccess;
    g->graph = graph;
    if (!ok || rc || !ended || !graph)
    {
        graph_free(g);
        return rc ? rc : AB_ERR_HIP;
    }
    if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
    {
        graph_free(g);
        return AB_ERR_HIP;
    }
    *out = g;
    return AB_OK;
}

int ab_graph_launch(ab_graph *g)
{
    if (!g || !g->exec) return AB_ERR_ARG;
    // the graph holds the device addresses and sizes of the clouds it was
    // captured on: refuse a replay once any of them was re-uploaded or
    // re-laid out
    for (size_t k = 0; k < g->ctxs.size(); ++k)
    {
        if (g->ctxs[k]->gen != g->gens[k])
        {
            g->ctxs[k]->err = "ab_graph_launch: the cloud changed after the capture (upload or "
                              "cloud management); create the graph again";
            return AB_ERR_STATE;
        }
        if (g->ctxs[k]->tgen != g->tgens[k])
        {
            g->ctxs[k]->err = "ab_graph_launch: the 2-D table transport was set up again after the "
                              "capture (neighbour rebuild, wall state or run length); create the "
                              "graph again";
            return AB_ERR_STATE;
        }
    }
    // the captured steps apply the table coefficients without building them:
    // build them for the graph's dt where an eager step at another dt or a
    // wall-state change left them stale (on the context's stream, so before
    // the replay)
    for (size_t k = 0; k < g->ctxs.size(); ++k)
    {
        const int rc = ab::graph_coefficients(g->ctxs[k], g->dt);
        if (rc) return rc;
    }
    // the replay runs after the work already enqueued on every context's
    // stream, and what is enqueued on them afterwards runs after the replay,
    // as if the steps had been enqueued eagerly
    const int n = (int)g->ctxs.size();
    for (int k = 0; k < n; ++k)
        if (hipEventRecord(g->order[k], g->ctxs[k]->stream) != hipSuccess ||
            hipStreamWaitEvent(g->stream, g->order[k], 0) != hipSuccess)
            return AB_ERR_HIP;
    if (hipGraphLaunch(g->exec, g->stream) != hipSuccess ||
        hipEventRecord(g->order[n], g->stream) != hipSuccess)
        return AB_ERR_HIP;
    for (int k = 0; k < n; ++k)
        if (hipStreamWaitEvent(g->ctxs[k]->stream, g->order[n], 0) != hipSuccess) return AB_ERR_HIP;
    return AB_OK;
}

int ab_graph_synchronize(ab_graph *g)
{
    if (!g) return AB_ERR_ARG;
    return hipStreamSynchronize(g->stream) == hipSuccess ? AB_OK : AB_ERR_HIP;
}

int ab_graph_destroy(ab_graph *g)
{
    if (!g) return AB_ERR_ARG;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    graph_free(g);
    return AB_OK;
}

}  // extern "C"
