// Batches of novel-view solves (include/lfrv_batch.h).
#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "../../include/lfrv_batch.h"
#include "lfrv_internal.hpp"

namespace lfrv {

// done flag of every solve of the batch (the poll after each chunk)
__global__ void k_batch_done(const CgBatchItem *__restrict__ items, int n, int *__restrict__ done)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) done[i] = items[i].st->done;
}

}  // namespace lfrv

using namespace lfrv;

struct lfrv_batch {
    lfrv_batch_params p{};
    std::string err;
    struct Entry {
        lfrv_ctx *ctx;
        int kind;
        double tol;
        int maxit, steps;
    };
    std::vector<Entry> entries;
    hipStream_t stream = nullptr;
    // device copies for the CG part
    CgBatchItem *d_items = nullptr;
    int2 *d_work = nullptr;
    int *d_done = nullptr, *h_done = nullptr;
    size_t cap_items = 0, cap_work = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_blocks = -1, graph_items = -1;
    std::vector<hipEvent_t> events;
    // the IBR_color steps of all batchable solves: one launch per step and footprint
    hipStream_t pd_stream = nullptr;
    SrBatchItem *d_sr_items = nullptr;
    int2 *d_sr_work = nullptr;
    float *d_theta = nullptr;
    size_t cap_sr_items = 0, cap_sr_work = 0, cap_theta = 0;
};

namespace {

lfrv_status bfail(lfrv_batch *b, lfrv_status st, const std::string &m)
{
    b->err = m;
    return st;
}

#define B_HIP(b, call)                                                                                   \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess) return bfail(b, LFRV_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define B_CHECK(b, ctx, call)                                                              \
    do {                                                                                   \
        lfrv_status s_ = (call);                                                           \
        if (s_ != LFRV_OK) return bfail(b, s_, std::string(#call) + ": " + lfrv_last_error(ctx)); \
    } while (0)

bool has_depth(lfrv_ctx *ctx)
{
    bool d = ctx->have_target_cam;
    for (int v = 0; v < ctx->V; v++) d = d && ctx->have_depth[v];
    return d;
}

hipEvent_t event_of(lfrv_batch *b, size_t i)
{
    while (b->events.size() <= i) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
        b->events.push_back(e);
    }
    return b->events[i];
}

}  // namespace

extern "C" lfrv_status lfrv_batch_create(const lfrv_batch_params *p, lfrv_batch **out)
{
    if (!out) return LFRV_E_ARG;
    *out = nullptr;
    lfrv_batch *b = new lfrv_batch;
    if (p) b->p = *p;
    if (b->p.chunk <= 0) b->p.chunk = 8;
    if (b->p.chunk % 2) b->p.chunk++;  // even: a replay starts every solve on the parity it left
    if (hipStreamCreate(&b->stream) != hipSuccess) {
        delete b;
        return LFRV_E_HIP;
    }
    if (hipStreamCreate(&b->pd_stream) != hipSuccess) {
        (void)hipStreamDestroy(b->stream);
        delete b;
        return LFRV_E_HIP;
    }
    *out = b;
    return LFRV_OK;
}

extern "C" void lfrv_batch_destroy(lfrv_batch *b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->pd_stream) (void)hipStreamSynchronize(b->pd_stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (void *q : {(void *)b->d_items, (void *)b->d_work, (void *)b->d_done, (void *)b->d_sr_items,
                    (void *)b->d_sr_work, (void *)b->d_theta})
        if (q) (void)hipFree(q);
    if (b->h_done) (void)hipHostFree(b->h_done);
    for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->pd_stream) (void)hipStreamDestroy(b->pd_stream);
    delete b;
}

extern "C" const char *lfrv_batch_last_error(const lfrv_batch *b) { return b ? b->err.c_str() : "null batch"; }

extern "C" lfrv_status lfrv_batch_add(lfrv_batch *b, lfrv_ctx *ctx, int kind, double tol, int maxit, int steps)
{
    if (!b || !ctx) return LFRV_E_ARG;
    if (kind == LFRV_BATCH_CG) {
        if (ctx->p.mode != LFRV_MODE_GRADIENT_IBR || ctx->dsf != 1 || !(tol > 0.0) || maxit < 0)
            return bfail(b, LFRV_E_ARG, "a CG entry takes a dsf 1 IBR_gradient context, tol > 0, maxit >= 0");
    } else if (kind == LFRV_BATCH_PD) {
        if (ctx->p.mode != LFRV_MODE_COLOR_SR || steps < 0)
            return bfail(b, LFRV_E_ARG, "a PD entry takes an IBR_color context and steps >= 0");
    } else {
        return bfail(b, LFRV_E_ARG, "unknown entry kind");
    }
    b->entries.push_back({ctx, kind, tol, maxit, steps});
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_batch_run(lfrv_batch *b, lfrv_batch_result *res)
{
    if (!b) return LFRV_E_ARG;
    const auto t0 = std::chrono::steady_clock::now();
    auto ms_now = [&]() { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); };
    const size_t ne = b->entries.size();
    std::vector<lfrv_batch_result> out(ne);
    std::vector<int> cg_index;  // entry of each CG item
    std::vector<CgBatchItem> items;
    std::vector<int> sr_index;  // entry of each batched IBR_color item
    std::vector<SrBatchItem> sr_items;
    std::vector<float> theta_host;
    std::vector<size_t> sorted_theta;
    // uploaded with asynchronous copies: they live until the batch has run
    std::vector<SrBatchItem> sr_sorted;
    std::vector<int2> sr_work;
    // 1. every solve's set-up on its own stream, in two rounds over the solves:
    // the hole filling of an IBR_gradient solve waits on the host for its
    // visibility pass's hole count, so every solve's warps, splat structure,
    // visibility and weights are queued first, and that wait then finds them
    // running (or done) on the other streams
    for (size_t i = 0; i < ne; i++) {
        auto &e = b->entries[i];
        lfrv_ctx *ctx = e.ctx;
        out[i] = lfrv_batch_result{e.kind, LFRV_OK, 0, 0.0, 0.0};
        B_CHECK(b, ctx, lfrv_clear_solution(ctx));
        if (has_depth(ctx)) B_CHECK(b, ctx, lfrv_warps_from_depth(ctx));
        B_CHECK(b, ctx, lfrv_prepare(ctx));
        if (e.kind == LFRV_BATCH_CG) {
            B_CHECK(b, ctx, lfrv_setup_visibility(ctx));
            B_CHECK(b, ctx, lfrv_compute_weights(ctx));
        } else {
            B_CHECK(b, ctx, lfrv_compute_weights(ctx));
            B_CHECK(b, ctx, lfrv_setup_visibility(ctx));
        }
    }
    // the rest of the set-up; the PD solves' steps queued behind it
    for (size_t i = 0; i < ne; i++) {
        auto &e = b->entries[i];
        lfrv_ctx *ctx = e.ctx;
        B_CHECK(b, ctx, lfrv_compute_initial_image(ctx, nullptr));
        if (e.kind == LFRV_BATCH_CG) {
            B_CHECK(b, ctx, lfrv_init_u_gradient(ctx));
            CgBatchItem it{};
            B_CHECK(b, ctx, cg_batch_begin(ctx, e.tol, e.maxit, &it));
            items.push_back(it);
            cg_index.push_back((int)i);
        } else {
            B_CHECK(b, ctx, lfrv_sr_init_regularizer_weight(ctx));
            if (sr_batchable(ctx) && e.steps > 0) {
                SrBatchItem it{};
                theta_host.resize(theta_host.size() + e.steps);
                B_CHECK(b, ctx, sr_batch_begin(ctx, e.steps, &it, theta_host.data() + theta_host.size() - e.steps));
                sr_items.push_back(it);
                sr_index.push_back((int)i);
            } else {
                B_CHECK(b, ctx, lfrv_sr_iterate(ctx, e.steps));
            }
            out[i].iterations = e.steps;
        }
        hipEvent_t ev = event_of(b, i);
        if (!ev) return bfail(b, LFRV_E_HIP, "event creation failed");
        B_HIP(b, hipEventRecord(ev, ctx->stream));
    }
    // the batched IBR_color steps: after every such solve's set-up, on their
    // own stream beside the PCG iterations; step s of all solves of one
    // footprint in one launch.  Each solve's stream then waits for them, and
    // its event marks their end.
    if (!sr_items.empty()) {
        const int ns = (int)sr_items.size();
        for (int k = 0; k < ns; k++) B_HIP(b, hipStreamWaitEvent(b->pd_stream, b->events[sr_index[k]], 0));
        // items grouped by footprint, each group's work list (item, local block)
        std::vector<SrBatchItem> &items_sorted = sr_sorted;
        std::vector<int2> &work = sr_work;
        struct Group { int ks; bool grad; int block0, nblocks; };
        std::vector<Group> groups;
        int max_steps = 0;
        size_t toff = 0;
        std::vector<size_t> theta_off(ns);
        for (int k = 0; k < ns; k++) {
            theta_off[k] = toff;
            toff += (size_t)sr_items[k].steps;
        }
        for (int gk = 0; gk < 6; gk++) {
            const int ks = 3 + gk / 2;
            const bool grad = (gk & 1) != 0;
            const int b0 = (int)work.size();
            for (int k = 0; k < ns; k++) {
                if (b->entries[sr_index[k]].ctx->ks != ks || (sr_items[k].da.beta > 0.0f) != grad) continue;
                const int idx = (int)items_sorted.size();
                items_sorted.push_back(sr_items[k]);
                items_sorted.back().theta = nullptr;  // set below, once the device array exists
                items_sorted.back().NA = sr_items[k].NA;
                for (int j = 0; j < sr_items[k].ntiles; j++) work.push_back(make_int2(idx, j));
                max_steps = std::max(max_steps, sr_items[k].steps);
                sorted_theta.push_back(theta_off[k]);
            }
            if ((int)work.size() > b0) groups.push_back({ks, grad, b0, (int)work.size() - b0});
        }
        const size_t nit = items_sorted.size();
        if (nit > b->cap_sr_items) {
            if (b->d_sr_items) B_HIP(b, hipFree(b->d_sr_items));
            B_HIP(b, hipMalloc((void **)&b->d_sr_items, nit * sizeof(SrBatchItem)));
            b->cap_sr_items = nit;
        }
        if (work.size() > b->cap_sr_work) {
            if (b->d_sr_work) B_HIP(b, hipFree(b->d_sr_work));
            B_HIP(b, hipMalloc((void **)&b->d_sr_work, work.size() * sizeof(int2)));
            b->cap_sr_work = work.size();
        }
        if (toff > b->cap_theta) {
            if (b->d_theta) B_HIP(b, hipFree(b->d_theta));
            B_HIP(b, hipMalloc((void **)&b->d_theta, toff * sizeof(float)));
            b->cap_theta = toff;
        }
        for (size_t k = 0; k < nit; k++) items_sorted[k].theta = b->d_theta + sorted_theta[k];
        B_HIP(b, hipMemcpyAsync(b->d_theta, theta_host.data(), toff * sizeof(float), hipMemcpyHostToDevice,
                                b->pd_stream));
        B_HIP(b, hipMemcpyAsync(b->d_sr_items, items_sorted.data(), nit * sizeof(SrBatchItem), hipMemcpyHostToDevice,
                                b->pd_stream));
        B_HIP(b, hipMemcpyAsync(b->d_sr_work, work.data(), work.size() * sizeof(int2), hipMemcpyHostToDevice,
                                b->pd_stream));
        for (int st = 0; st < max_steps; st++)
            for (const Group &gr : groups)
                B_HIP(b, launch_sr_batch_step(gr.ks, gr.grad, b->d_sr_items, b->d_sr_work + gr.block0, gr.nblocks,
                                              st, b->pd_stream));
        for (int k = 0; k < ns; k++) {
            lfrv_ctx *ctx = b->entries[sr_index[k]].ctx;
            B_HIP(b, hipEventRecord(b->events[sr_index[k]], b->pd_stream));
            B_HIP(b, hipStreamWaitEvent(ctx->stream, b->events[sr_index[k]], 0));
            sr_batch_end(ctx, sr_items[k].steps);
        }
    }
    const int ni = (int)items.size();
    std::vector<char> seen(ne, 0);
    if (ni > 0) {
        // 2. the batch stream waits for the CG solves' set-up
        for (int k = 0; k < ni; k++) B_HIP(b, hipStreamWaitEvent(b->stream, b->events[cg_index[k]], 0));
        std::vector<int2> work;
        for (int k = 0; k < ni; k++)
            for (int t = 0; t < items[k].a[0].ntiles; t++) work.push_back(make_int2(k, t));
        const int nblocks = (int)work.size();
        if ((size_t)ni > b->cap_items) {
            if (b->d_items) B_HIP(b, hipFree(b->d_items));
            if (b->d_done) B_HIP(b, hipFree(b->d_done));
            if (b->h_done) B_HIP(b, hipHostFree(b->h_done));
            B_HIP(b, hipMalloc((void **)&b->d_items, ni * sizeof(CgBatchItem)));
            B_HIP(b, hipMalloc((void **)&b->d_done, ni * sizeof(int)));
            B_HIP(b, hipHostMalloc((void **)&b->h_done, ni * sizeof(int), hipHostMallocDefault));
            b->cap_items = ni;
            b->graph_blocks = -1;
        }
        if ((size_t)nblocks > b->cap_work) {
            if (b->d_work) B_HIP(b, hipFree(b->d_work));
            B_HIP(b, hipMalloc((void **)&b->d_work, nblocks * sizeof(int2)));
            b->cap_work = nblocks;
            b->graph_blocks = -1;
        }
        B_HIP(b, hipMemcpyAsync(b->d_items, items.data(), ni * sizeof(CgBatchItem), hipMemcpyHostToDevice, b->stream));
        B_HIP(b, hipMemcpyAsync(b->d_work, work.data(), nblocks * sizeof(int2), hipMemcpyHostToDevice, b->stream));
        // 3. the chunk of iterations, captured once per work-list shape
        const int chunk = b->p.chunk;
        auto launch_chunk = [&]() -> hipError_t {
            for (int k = 0; k < chunk; k++) {
                hipError_t e = launch_cg_batch_iteration(b->d_items, b->d_work, nblocks, k & 1, b->stream);
                if (e == hipSuccess) e = launch_cg_batch_step(b->d_items, ni, b->stream);
                if (e != hipSuccess) return e;
            }
            return hipSuccess;
        };
        if (b->p.use_graph && (b->graph_blocks != nblocks || b->graph_items != ni || !b->exec)) {
            if (b->exec) B_HIP(b, hipGraphExecDestroy(b->exec));
            if (b->graph) B_HIP(b, hipGraphDestroy(b->graph));
            b->exec = nullptr;
            b->graph = nullptr;
            B_HIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            const hipError_t ce = launch_chunk();
            hipGraph_t g = nullptr;
            const hipError_t ee = hipStreamEndCapture(b->stream, &g);
            if (ce != hipSuccess || ee != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return bfail(b, LFRV_E_HIP, "graph capture of the batched iterations failed");
            }
            b->graph = g;
            B_HIP(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->graph_blocks = nblocks;
            b->graph_items = ni;
        }
        int maxit = 0;
        for (int k = 0; k < ni; k++) maxit = std::max(maxit, b->entries[cg_index[k]].maxit);
        int remaining = ni, ran = 0;
        // 4. replay until every solve is done; each finished solve's end pass runs at once
        while (remaining > 0) {
            if (b->p.use_graph) B_HIP(b, hipGraphLaunch(b->exec, b->stream));
            else B_HIP(b, launch_chunk());
            ran += chunk;
            k_batch_done<<<(ni + 255) / 256, 256, 0, b->stream>>>(b->d_items, ni, b->d_done);
            B_HIP(b, hipGetLastError());
            B_HIP(b, hipMemcpyAsync(b->h_done, b->d_done, ni * sizeof(int), hipMemcpyDeviceToHost, b->stream));
            B_HIP(b, hipStreamSynchronize(b->stream));
            for (int k = 0; k < ni; k++) {
                const int i = cg_index[k];
                if (seen[i] || !(b->h_done[k] || ran >= maxit + chunk)) continue;
                lfrv_ctx *ctx = b->entries[i].ctx;
                lfrv_cg_stats st{};
                const lfrv_status s = cg_batch_end(ctx, &st);
                out[i].status = s;
                out[i].iterations = std::max(st.iterations[0], std::max(st.iterations[1], st.iterations[2]));
                out[i].true_rel_residual =
                    std::max(st.true_rel_residual[0], std::max(st.true_rel_residual[1], st.true_rel_residual[2]));
                out[i].done_ms = ms_now();
                seen[i] = 1;
                remaining--;
            }
            for (size_t i = 0; i < ne; i++)
                if (!seen[i] && b->entries[i].kind == LFRV_BATCH_PD && hipEventQuery(b->events[i]) == hipSuccess) {
                    out[i].done_ms = ms_now();
                    seen[i] = 1;
                }
        }
    }
    // 5. the PD solves still running
    for (size_t i = 0; i < ne; i++) {
        if (seen[i]) continue;
        B_HIP(b, hipEventSynchronize(b->events[i]));
        out[i].done_ms = ms_now();
        seen[i] = 1;
    }
    if (res)
        for (size_t i = 0; i < ne; i++) res[i] = out[i];
    lfrv_status worst = LFRV_OK;
    for (size_t i = 0; i < ne; i++)
        if (out[i].status != LFRV_OK) worst = (lfrv_status)out[i].status;
    if (worst != LFRV_OK) return bfail(b, worst, "a solve of the batch did not finish cleanly (see its result)");
    return LFRV_OK;
}
