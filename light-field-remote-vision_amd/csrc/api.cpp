// C-ABI of the lfrv library: context lifecycle, inputs, and the solver steps.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <tuple>

#include "tile_common.hpp"

namespace lfrv {

lfrv_status fail(lfrv_ctx *ctx, lfrv_status st, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return st;
}

lfrv_status hip_fail(lfrv_ctx *ctx, hipError_t e, const char *what)
{
    if (ctx) ctx->err = std::string(what) + ": " + hipGetErrorString(e);
    return LFRV_E_HIP;
}

// ---- event timing ------------------------------------------------------------

static hipEvent_t get_event(lfrv_ctx *ctx)
{
    if (!ctx->ev_pool.empty()) {
        hipEvent_t e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

TimedRegion::TimedRegion(lfrv_ctx *c, int kk) : ctx(c), k(kk)
{
    if (!ctx->timing) return;
    // event records are queue packets the stream waits on: sampled like the
    // kernel events, one region in timing_stride of each class
    if (ctx->timing_seq[k]++ % ctx->timing_stride != 0) return;
    a = get_event(ctx);
    if (a) (void)hipEventRecord(a, ctx->stream);
}

TimedRegion::~TimedRegion()
{
    if (!ctx->timing || !a) return;
    b = get_event(ctx);
    if (!b) return;
    (void)hipEventRecord(b, ctx->stream);
    ctx->ev_pending.emplace_back(k, a, b);
    if (ctx->ev_pending.size() > 4096) flush_timing(ctx);
}

KernelEvents kernel_events(lfrv_ctx *ctx, int k)
{
    KernelEvents ev;
    if (!ctx->timing) return ev;
    // a timestamped dispatch waits for its completion signal (several us on a
    // short kernel), so only one launch in timing_stride of each class carries one
    if (ctx->timing_seq[k]++ % ctx->timing_stride != 0) return ev;
    ev.a = get_event(ctx);
    ev.b = ev.a ? get_event(ctx) : nullptr;
    if (!ev.b) ev.a = nullptr;
    return ev;
}

void record_kernel_events(lfrv_ctx *ctx, int k, const KernelEvents &ev)
{
    if (!ev.a || !ev.b) return;
    ctx->ev_pending.emplace_back(k, ev.a, ev.b);
    if (ctx->ev_pending.size() > 4096) flush_timing(ctx);
}

}  // namespace lfrv

namespace lfrv {

void flush_timing(lfrv_ctx *ctx)
{
    if (ctx->ev_pending.empty()) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &t : ctx->ev_pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, std::get<1>(t), std::get<2>(t)) == hipSuccess) {
            ctx->tslot[std::get<0>(t)].ms += ms;
            ctx->tslot[std::get<0>(t)].launches += 1;
        }
        ctx->ev_pool.push_back(std::get<1>(t));
        ctx->ev_pool.push_back(std::get<2>(t));
    }
    ctx->ev_pending.clear();
}

}  // namespace lfrv

using namespace lfrv;

// ---- lifecycle ----------------------------------------------------------------

extern "C" void lfrv_default_params(lfrv_params *p, int mode, int nviews, int w, int h, int dsf)
{
    std::memset(p, 0, sizeof(*p));
    p->mode = mode;
    p->nviews = nviews;
    p->w = w;
    p->h = h;
    p->dsf = dsf;
    // Config_data::Config_data (config.cpp:11-64)
    p->sigma_sensor = 1.0f / 255.0f;
    p->ugrad_threshold = 1.0f;
    p->dw_type = 0;
    p->gw_type = 0;
    p->dt_alpha = 1.0f;
    p->dt_beta = 0.0f;
    p->gradient_step = 0.1f;
    p->lambda = 0.0f;
    p->lambda_max_factor = 0.0f;
    p->inner_iterations = 10;  // coco_vtv_alloc (vtv.cu:37)
    p->occlusion_tol_px = 10.0f;
    p->sigma_z = 10.0f;
    p->backproject_source_cam = 0;
    p->tile_w = 0;
    p->tile_h = 0;
    p->pull_lds_kb = 0;
}

template <class T>
static hipError_t dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e == hipSuccess) e = hipMemset(*p, 0, count * sizeof(T));
    return e;
}

extern "C" lfrv_status lfrv_create(const lfrv_params *p, lfrv_ctx **out)
{
    if (!p || !out) return LFRV_E_ARG;
    *out = nullptr;
    if (p->nviews < 1 || p->w < 3 || p->h < 3 || p->dsf < 1 || p->dsf > 4) return LFRV_E_ARG;
    if (p->mode != LFRV_MODE_GRADIENT_IBR && p->mode != LFRV_MODE_COLOR_SR) return LFRV_E_ARG;
    if (p->dw_type != 0) return LFRV_E_ARG;  // the reference asserts on 1 and 2 (gradientIBR.cu:636-641)
    if (p->mode == LFRV_MODE_COLOR_SR && !(p->lambda > 0.0f)) return LFRV_E_ARG;
    lfrv_ctx *ctx = new lfrv_ctx();
    ctx->p = *p;
    ctx->V = p->nviews;
    ctx->w = p->w;
    ctx->h = p->h;
    ctx->dsf = p->dsf;
    ctx->ks = p->dsf + 1;
    ctx->W = p->w * p->dsf;
    ctx->H = p->h * p->dsf;
    ctx->n = (size_t)ctx->w * ctx->h;
    ctx->N = (size_t)ctx->W * ctx->H;
    ctx->tile_w = p->tile_w > 0 ? p->tile_w : 32;
    ctx->tile_h = p->tile_h > 0 ? p->tile_h : 8;
    if (p->pull_lds_kb < 0 || p->pull_lds_kb > 160) {
        delete ctx;
        return LFRV_E_ARG;
    }
    ctx->pull_budget = (p->pull_lds_kb > 0 ? p->pull_lds_kb : 48) * 1024;
    if (ctx->tile_w < ctx->ks || ctx->tile_h < ctx->ks || ctx->tile_w * ctx->tile_h > kPullMaxThreads ||
        (ctx->tile_w + ctx->ks - 1) * (ctx->tile_h + ctx->ks - 1) > kPullMaxThreads ||
        ctx->tile_h + ctx->ks - 1 > 64) {  // the window rows are scanned by one wave
        delete ctx;
        return LFRV_E_ARG;
    }
    {
        const int cells = (ctx->tile_w + ctx->ks - 1) * (ctx->tile_h + ctx->ks - 1);
        const int need = std::max(cells, ctx->tile_w * ctx->tile_h);
        ctx->pull_bs = std::min(kPullMaxThreads, (need + 63) / 64 * 64);
    }
    ctx->ntx = (ctx->W + ctx->tile_w - 1) / ctx->tile_w;
    ctx->nty = (ctx->H + ctx->tile_h - 1) / ctx->tile_h;
    ctx->ntiles = ctx->ntx * ctx->nty;
    ctx->P_views.assign((size_t)ctx->V * 12, 0.0);
    ctx->have_depth.assign(ctx->V, 0);
    const size_t V = ctx->V, n = ctx->n, N = ctx->N;
    hipError_t e = hipSuccess;
    auto A = [&](auto **ptr, size_t cnt) {
        if (e == hipSuccess) e = dalloc(ptr, cnt);
    };
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->side_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->side_join, hipEventDisableTiming);
    A(&ctx->views, 3 * V * n);
    A(&ctx->tau_x, V * n);
    A(&ctx->tau_y, V * n);
    A(&ctx->dpart_x, V * n);
    A(&ctx->dpart_y, V * n);
    A(&ctx->omega, V * n);
    A(&ctx->depth, V * n);
    A(&ctx->view_lap, 3 * V * n);
    A(&ctx->view_grad, 2 * V * n);
    ctx->ncells = (ctx->W + ctx->ks - 1) * (ctx->H + ctx->ks - 1);
    A(&ctx->cell_count, (size_t)ctx->ncells + 1);
    A(&ctx->cell_start, (size_t)ctx->ncells + 1);
    A(&ctx->cell_cursor, (size_t)ctx->ncells);
    A(&ctx->scan_tmp, (size_t)ctx->ncells / 2048 + 16);
    ctx->e_cap = V * n;
    A(&ctx->e_src, ctx->e_cap);
    A(&ctx->e_tx, ctx->e_cap);
    A(&ctx->e_ty, ctx->e_cap);
    A(&ctx->e_w, ctx->e_cap);
    if (p->mode == LFRV_MODE_COLOR_SR && ctx->ks >= 3) {
        A(&ctx->e_pay, 6 * ctx->e_cap);
        A(&ctx->sr_ops, kSrOpsPlanes * (size_t)(ctx->W + ctx->ks - 1) * (ctx->H + ctx->ks - 1));
    }
    for (float **b : {&ctx->U, &ctx->Uq, &ctx->G, &ctx->PX, &ctx->PY, &ctx->U2, &ctx->P2X, &ctx->P2Y, &ctx->tmp3N,
                      &ctx->cg_r, &ctx->cg_d, &ctx->cg_q, &ctx->cg_b, &ctx->cache_num})
        A(b, 3 * N);
    for (float **b : {&ctx->reg_g, &ctx->norm_mask, &ctx->coverage, &ctx->u_grad_x, &ctx->u_grad_y,
                      &ctx->target_mask, &ctx->visibility, &ctx->target_depth, &ctx->tmpN, &ctx->cg_minv,
                      &ctx->cache_sq})
        A(b, N);
    A(&ctx->hole_list, N);
    if (ctx->ks == 2) {
        A(&ctx->gram, 5 * N);
        A(&ctx->cg_extra, 9 * N);
    } else if (p->mode == LFRV_MODE_GRADIENT_IBR) {
        // the box operator's coefficients, and the second r, s, w buffers and u
        A(&ctx->box_ops, 5 * (size_t)(ctx->W + ctx->ks - 1) * (ctx->H + ctx->ks - 1));
        A(&ctx->cg_extra, 12 * N);
    }
    ctx->pp_cap = 3 * (N + N / 2) + 3 * N + N / 4 + 64;
    A(&ctx->pp_scratch, ctx->pp_cap);
    A(&ctx->cg_state, 1);
    A(&ctx->sr_loop, 1);
    A(&ctx->counters, 32);
    A(&ctx->dcams, V);
    A(&ctx->scal, 64);
    const size_t cam_doubles = (V * sizeof(WarpCam) + 7) / 8;
    const size_t st2_tiles = (size_t)((ctx->W + 63) / 64) * ((ctx->H + 3) / 4);  // stencil operator tiles
    ctx->partials_cap = std::max<size_t>({(size_t)ctx->ntiles * 9, st2_tiles * 12, 2048 * 9, cam_doubles});
    A(&ctx->partials, ctx->partials_cap);
    if (e == hipSuccess) e = hipHostMalloc((void **)&ctx->cg_state_host, sizeof(CgState), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&ctx->scal_host, 64 * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&ctx->hole_count_host, sizeof(unsigned int), hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->hole_ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipHostMalloc((void **)&ctx->sr_loop_host, 3 * sizeof(SrLoop), hipHostMallocDefault);
    for (hipEvent_t &ev : ctx->sr_ev)
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        lfrv_destroy(ctx);
        return LFRV_E_HIP;
    }
    *out = ctx;
    return LFRV_OK;
}

extern "C" void lfrv_destroy(lfrv_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    flush_timing(ctx);
    for (void *p : std::initializer_list<void *>{
             ctx->views, ctx->tau_x, ctx->tau_y, ctx->dpart_x, ctx->dpart_y, ctx->omega, ctx->depth, ctx->view_lap,
             ctx->view_grad, ctx->cell_count, ctx->cell_start, ctx->cell_cursor, ctx->scan_tmp, ctx->e_src, ctx->e_tx, ctx->e_ty,
             ctx->e_w, ctx->e_pay, ctx->sr_ops, ctx->U, ctx->Uq, ctx->G, ctx->PX, ctx->PY, ctx->U2, ctx->P2X, ctx->P2Y, ctx->tmp3N, ctx->cg_r,
             ctx->cg_d, ctx->cg_q, ctx->cg_b, ctx->reg_g, ctx->norm_mask, ctx->coverage, ctx->u_grad_x,
             ctx->u_grad_y, ctx->target_mask, ctx->visibility, ctx->target_depth, ctx->tmpN, ctx->cg_minv,
             ctx->hole_list, ctx->pp_scratch, ctx->cg_state, ctx->counters, ctx->dcams, ctx->scal, ctx->partials,
             ctx->cache_num, ctx->cache_sq, ctx->gram, ctx->cg_extra, ctx->sr_loop, ctx->sr_prev, ctx->depth_grad,
             ctx->tau_weight, ctx->gl_big, ctx->box_ops})
        if (p) (void)hipFree(p);
    if (ctx->cg_state_host) (void)hipHostFree(ctx->cg_state_host);
    if (ctx->scal_host) (void)hipHostFree(ctx->scal_host);
    if (ctx->hole_count_host) (void)hipHostFree(ctx->hole_count_host);
    if (ctx->hole_ev) (void)hipEventDestroy(ctx->hole_ev);
    if (ctx->sr_loop_host) (void)hipHostFree(ctx->sr_loop_host);
    for (hipEvent_t ev : ctx->sr_ev)
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : ctx->ev_pool) (void)hipEventDestroy(ev);
    if (ctx->side_fork) (void)hipEventDestroy(ctx->side_fork);
    if (ctx->side_join) (void)hipEventDestroy(ctx->side_join);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" const char *lfrv_last_error(const lfrv_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" lfrv_status lfrv_sync(lfrv_ctx *ctx)
{
    if (!ctx) return LFRV_E_ARG;
    LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    flush_timing(ctx);
    return LFRV_OK;
}

// ---- inputs -------------------------------------------------------------------

__global__ static void k_zero(size_t n, float *__restrict__ p)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0.0f;
}

static lfrv_status copy_in(lfrv_ctx *ctx, float *dst, const float *src, size_t count)
{
    if (!src) return fail(ctx, LFRV_E_ARG, "null input pointer");
    LFRV_HIP(ctx, hipMemcpyAsync(dst, src, count * sizeof(float), hipMemcpyDefault, ctx->stream));
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_set_view(lfrv_ctx *ctx, int i, const float *rgb)
{
    if (!ctx || i < 0 || i >= ctx->V) return ctx ? fail(ctx, LFRV_E_ARG, "view index out of range") : LFRV_E_ARG;
    lfrv_status st = copy_in(ctx, ctx->views + (size_t)i * 3 * ctx->n, rgb, 3 * ctx->n);
    if (st != LFRV_OK) return st;
    ctx->prepared = false;  // view stencils are rebuilt by lfrv_prepare
    ctx->epoch++;
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_set_tau(lfrv_ctx *ctx, int i, const float *tau_x, const float *tau_y)
{
    if (!ctx || i < 0 || i >= ctx->V) return ctx ? fail(ctx, LFRV_E_ARG, "view index out of range") : LFRV_E_ARG;
    lfrv_status st = copy_in(ctx, ctx->tau_x + (size_t)i * ctx->n, tau_x, ctx->n);
    if (st == LFRV_OK) st = copy_in(ctx, ctx->tau_y + (size_t)i * ctx->n, tau_y, ctx->n);
    ctx->prepared = false;
    ctx->counts_valid = false;
    ctx->epoch++;
    return st;
}

__global__ static void k_mul(size_t n, const float *__restrict__ a, float *__restrict__ b)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) b[k] *= a[k];
}

extern "C" lfrv_status lfrv_set_partial_tau(lfrv_ctx *ctx, int i, const float *sigma_z, const float *dtdz_x,
                                            const float *dtdz_y)
{
    if (!ctx || i < 0 || i >= ctx->V) return ctx ? fail(ctx, LFRV_E_ARG, "view index out of range") : LFRV_E_ARG;
    const size_t n = ctx->n;
    float *dx = ctx->dpart_x + (size_t)i * n, *dy = ctx->dpart_y + (size_t)i * n;
    lfrv_status st = materialize_dpart(ctx);  // the other views keep theirs
    if (st == LFRV_OK) st = copy_in(ctx, ctx->tmpN, sigma_z, n);
    if (st == LFRV_OK) st = copy_in(ctx, dx, dtdz_x, n);
    if (st == LFRV_OK) st = copy_in(ctx, dy, dtdz_y, n);
    if (st != LFRV_OK) return st;
    // read_partial_tau multiplies both derivatives by sigma_z (gradientIBR.cu:709-718)
    k_mul<<<blocks_for(n), kBlock, 0, ctx->stream>>>(n, ctx->tmpN, dx);
    k_mul<<<blocks_for(n), kBlock, 0, ctx->stream>>>(n, ctx->tmpN, dy);
    LFRV_LAUNCH(ctx, "partial tau");
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_set_depth(lfrv_ctx *ctx, int i, const float *depth, const double *P)
{
    if (!ctx || i < 0 || i >= ctx->V || !P) return ctx ? fail(ctx, LFRV_E_ARG, "bad lfrv_set_depth arguments") : LFRV_E_ARG;
    // dpart still owed for the current warps is formed from the depth they came from
    lfrv_status st = materialize_dpart(ctx);
    if (st == LFRV_OK) st = copy_in(ctx, ctx->depth + (size_t)i * ctx->n, depth, ctx->n);
    if (st != LFRV_OK) return st;
    std::memcpy(ctx->P_views.data() + 12 * i, P, 12 * sizeof(double));
    ctx->have_depth[i] = 1;
    ctx->cams_valid = false;
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_set_depth_gradient(lfrv_ctx *ctx, int i, const float *dzdx, const float *dzdy)
{
    if (!ctx || i < 0 || i >= ctx->V) return ctx ? fail(ctx, LFRV_E_ARG, "bad lfrv_set_depth_gradient arguments") : LFRV_E_ARG;
    if ((dzdx == nullptr) != (dzdy == nullptr)) return fail(ctx, LFRV_E_ARG, "lfrv_set_depth_gradient: both planes or neither");
    const size_t Vn = (size_t)ctx->V * ctx->n;
    if (!ctx->depth_grad) {
        if (!dzdx) return LFRV_OK;  // no gradients anywhere: the kernels take them as zero
        LFRV_HIP(ctx, hipMalloc((void **)&ctx->depth_grad, 2 * Vn * sizeof(float)));
        k_zero<<<blocks_for(2 * Vn), kBlock, 0, ctx->stream>>>(2 * Vn, ctx->depth_grad);
        LFRV_LAUNCH(ctx, "clear depth gradients");
    }
    float *gx = ctx->depth_grad + (size_t)i * ctx->n, *gy = gx + Vn;
    if (!dzdx) {
        k_zero<<<blocks_for(ctx->n), kBlock, 0, ctx->stream>>>(ctx->n, gx);
        k_zero<<<blocks_for(ctx->n), kBlock, 0, ctx->stream>>>(ctx->n, gy);
        LFRV_LAUNCH(ctx, "clear depth gradients");
        return LFRV_OK;
    }
    lfrv_status st = copy_in(ctx, gx, dzdx, ctx->n);
    if (st == LFRV_OK) st = copy_in(ctx, gy, dzdy, ctx->n);
    return st;
}

extern "C" lfrv_status lfrv_set_target_camera(lfrv_ctx *ctx, const double *P)
{
    if (!ctx || !P) return LFRV_E_ARG;
    std::memcpy(ctx->P_target, P, 12 * sizeof(double));
    ctx->have_target_cam = true;
    ctx->cams_valid = false;
    return LFRV_OK;
}

// ---- steps --------------------------------------------------------------------

extern "C" lfrv_status lfrv_warps_from_depth(lfrv_ctx *ctx)
{
    if (!ctx) return LFRV_E_ARG;
    return run_warps_from_depth(ctx);
}

extern "C" lfrv_status lfrv_warps_from_target_depth(lfrv_ctx *ctx, int skip_view)
{
    if (!ctx) return LFRV_E_ARG;
    if (skip_view < -1 || skip_view >= ctx->V) return fail(ctx, LFRV_E_ARG, "lfrv_warps_from_target_depth: bad view");
    return run_warps_from_target_depth(ctx, skip_view);
}

extern "C" lfrv_status lfrv_prepare(lfrv_ctx *ctx)
{
    if (!ctx) return LFRV_E_ARG;
    return run_prepare(ctx);
}

extern "C" lfrv_status lfrv_setup_visibility(lfrv_ctx *ctx)
{
    if (!ctx) return LFRV_E_ARG;
    lfrv_status st = run_visibility(ctx);
    if (st == LFRV_OK) ctx->vis_ready = true;
    return st;
}

extern "C" lfrv_status lfrv_compute_weights(lfrv_ctx *ctx)
{
    if (!ctx) return LFRV_E_ARG;
    return run_compute_weights(ctx);
}

extern "C" lfrv_status lfrv_compute_initial_image(lfrv_ctx *ctx, const float *u0)
{
    if (!ctx) return LFRV_E_ARG;
    ctx->u_zero = false;
    if (u0) return copy_in(ctx, ctx->U, u0, 3 * ctx->N);
    if (!ctx->vis_ready) return fail(ctx, LFRV_E_STATE, "lfrv_compute_initial_image before lfrv_setup_visibility");
    lfrv_status st = run_initial_image(ctx);
    if (st != LFRV_OK) return st;
    if (ctx->p.mode == LFRV_MODE_GRADIENT_IBR) {
        int ran = 0;
        st = run_hole_filling(ctx, 500, 1e-3, &ran);  // gradientIBR.cu:102, :10
        ctx->hole_iters = ran;
        return st;
    }
    return run_push_pull(ctx, 10);  // optimization.cu:574
}

extern "C" lfrv_status lfrv_get_hole_stats(lfrv_ctx *ctx, int *holes, int *sweeps)
{
    if (!ctx) return LFRV_E_ARG;
    if (holes) *holes = ctx->n_holes;
    if (sweeps) *sweeps = ctx->hole_iters;
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_get_solution(lfrv_ctx *ctx, float *rgb)
{
    if (!ctx || !rgb) return LFRV_E_ARG;
    LFRV_HIP(ctx, hipMemcpyAsync(rgb, ctx->U, 3 * ctx->N * sizeof(float), hipMemcpyDefault, ctx->stream));
    LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_set_solution(lfrv_ctx *ctx, const float *rgb)
{
    if (!ctx) return LFRV_E_ARG;
    ctx->u_zero = false;
    return copy_in(ctx, ctx->U, rgb, 3 * ctx->N);
}

extern "C" lfrv_status lfrv_clear_solution(lfrv_ctx *ctx)
{
    if (!ctx) return LFRV_E_ARG;
    k_zero<<<blocks_for(3 * ctx->N), kBlock, 0, ctx->stream>>>(3 * ctx->N, ctx->U);
    if (ctx->p.mode == LFRV_MODE_COLOR_SR) {
        // a new view starts from the state of a fresh workspace: coco_vtv_alloc
        // zeroes Uq, G, the duals and the normalisation buffer (vtv.cu:60-89),
        // coco_vtv_initialize sets alpha = 1 (vtv.cu:241-263)
        for (float *b : {ctx->Uq, ctx->G, ctx->PX, ctx->PY})
            k_zero<<<blocks_for(3 * ctx->N), kBlock, 0, ctx->stream>>>(3 * ctx->N, b);
        k_zero<<<blocks_for(ctx->N), kBlock, 0, ctx->stream>>>(ctx->N, ctx->norm_mask);
        ctx->fista_alpha = 1.0;
        ctx->has_reg_g = false;
    }
    LFRV_LAUNCH(ctx, "clear solution");
    ctx->u_zero = true;
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_init_u_gradient(lfrv_ctx *ctx)
{
    if (!ctx) return LFRV_E_ARG;
    return run_u_gradient(ctx);
}

extern "C" lfrv_status lfrv_cg_solve(lfrv_ctx *ctx, double tol, int maxit, lfrv_cg_stats *st)
{
    if (!ctx || !(tol > 0.0) || maxit < 0) return ctx ? fail(ctx, LFRV_E_ARG, "bad CG arguments") : LFRV_E_ARG;
    ctx->u_zero = false;
    return run_cg(ctx, ctx->p.dt_alpha, 0.0, ctx->p.dt_beta, true, true, false, tol, maxit, st);
}

extern "C" lfrv_status lfrv_ibr_gradient(lfrv_ctx *ctx, double tol, int maxit, lfrv_cg_stats *st)
{
    if (!ctx) return LFRV_E_ARG;
    if (ctx->p.mode != LFRV_MODE_GRADIENT_IBR) return fail(ctx, LFRV_E_STATE, "lfrv_ibr_gradient on an IBR_color context");
    if (!(tol > 0.0) || maxit < 0) return fail(ctx, LFRV_E_ARG, "bad CG arguments");
    bool depth = ctx->have_target_cam;
    for (int v = 0; v < ctx->V; v++) depth = depth && ctx->have_depth[v];
    // init_data's U = 0 only enters through the weights, which take the
    // zero-gradient form for it; compute_initial_image then writes all of U
    ctx->u_zero = true;
    lfrv_status s = LFRV_OK;
    if (depth) s = run_warps_from_depth(ctx);  // else the tau / dpart already set
    if (s == LFRV_OK) s = run_prepare(ctx);
    if (s == LFRV_OK) s = lfrv_setup_visibility(ctx);
    if (s == LFRV_OK) s = run_compute_weights(ctx);
    if (s == LFRV_OK) s = lfrv_compute_initial_image(ctx, nullptr);
    if (s == LFRV_OK) s = run_u_gradient(ctx);
    if (s == LFRV_OK) s = lfrv_cg_solve(ctx, tol, maxit, st);
    return s;
}

extern "C" lfrv_status lfrv_poisson_cg(lfrv_ctx *ctx, double lambda, double tol, int maxit, int use_precond,
                                       lfrv_cg_stats *st)
{
    if (!ctx || !(tol > 0.0) || maxit < 0 || lambda < 0.0) return ctx ? fail(ctx, LFRV_E_ARG, "bad CG arguments") : LFRV_E_ARG;
    ctx->u_zero = false;
    return run_cg(ctx, 0.0, lambda, 1.0, false, use_precond != 0, true, tol, maxit, st);
}

extern "C" lfrv_status lfrv_gibr_energy(lfrv_ctx *ctx, double *energy)
{
    if (!ctx || !energy) return LFRV_E_ARG;
    return run_gibr_energy(ctx, energy);
}

extern "C" lfrv_status lfrv_sr_init_regularizer_weight(lfrv_ctx *ctx)
{
    if (!ctx) return LFRV_E_ARG;
    return run_regularizer_weight(ctx);
}

extern "C" lfrv_status lfrv_sr_iteration_fista(lfrv_ctx *ctx)
{
    if (!ctx) return LFRV_E_ARG;
    if (ctx->p.mode != LFRV_MODE_COLOR_SR) return fail(ctx, LFRV_E_STATE, "FISTA step needs LFRV_MODE_COLOR_SR");
    ctx->u_zero = false;
    return run_fista(ctx);
}

extern "C" lfrv_status lfrv_sr_iterate(lfrv_ctx *ctx, int steps)
{
    if (!ctx) return LFRV_E_ARG;
    if (ctx->p.mode != LFRV_MODE_COLOR_SR) return fail(ctx, LFRV_E_STATE, "FISTA steps need LFRV_MODE_COLOR_SR");
    if (steps < 0) return fail(ctx, LFRV_E_ARG, "negative step count");
    ctx->u_zero = false;
    for (int k = 0; k < steps; k++) {
        const lfrv_status st = run_fista(ctx, k == steps - 1);
        if (st != LFRV_OK) return st;
    }
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_sr_primal_energy(lfrv_ctx *ctx, double *energy, double *parts)
{
    if (!ctx || !energy) return LFRV_E_ARG;
    if (!ctx->prepared) return fail(ctx, LFRV_E_STATE, "energy before lfrv_prepare");
    return run_sr_energy(ctx, energy, parts);
}

extern "C" lfrv_status lfrv_sr_count_moving(lfrv_ctx *ctx, const float *prev, float max_diff, int64_t *moving)
{
    if (!ctx || !prev || !moving) return LFRV_E_ARG;
    lfrv_status st = copy_in(ctx, ctx->tmp3N, prev, 3 * ctx->N);
    if (st != LFRV_OK) return st;
    return run_count_moving(ctx, ctx->tmp3N, max_diff, moving);
}

extern "C" lfrv_status lfrv_sr_solve(lfrv_ctx *ctx, int niter, int nmeta, lfrv_sr_stats *st)
{
    if (!ctx) return LFRV_E_ARG;
    if (ctx->p.mode != LFRV_MODE_COLOR_SR) return fail(ctx, LFRV_E_STATE, "lfrv_sr_solve needs LFRV_MODE_COLOR_SR");
    if (niter < 0 || nmeta < 0) return fail(ctx, LFRV_E_ARG, "negative iteration counts");
    if (!ctx->prepared) return fail(ctx, LFRV_E_STATE, "lfrv_sr_solve before lfrv_prepare");
    if (!ctx->vis_ready) return fail(ctx, LFRV_E_STATE, "lfrv_sr_solve before lfrv_setup_visibility");
    ctx->u_zero = false;
    return run_sr_solve(ctx, niter, nmeta, st);
}

// ---- inspection -----------------------------------------------------------------

static float *buffer_ptr(lfrv_ctx *ctx, int which, size_t *count)
{
    const size_t Vn = (size_t)ctx->V * ctx->n, N = ctx->N;
    switch (which) {
    case LFRV_BUF_TAU_X: *count = Vn; return ctx->tau_x;
    case LFRV_BUF_TAU_Y: *count = Vn; return ctx->tau_y;
    case LFRV_BUF_DPART_X: *count = Vn; return ctx->dpart_x;
    case LFRV_BUF_DPART_Y: *count = Vn; return ctx->dpart_y;
    case LFRV_BUF_OMEGA: *count = Vn; return ctx->omega;
    case LFRV_BUF_U: *count = 3 * N; return ctx->U;
    case LFRV_BUF_UQ: *count = 3 * N; return ctx->Uq;
    case LFRV_BUF_G: *count = 3 * N; return ctx->G;
    case LFRV_BUF_PX: *count = 3 * N; return ctx->PX;
    case LFRV_BUF_PY: *count = 3 * N; return ctx->PY;
    case LFRV_BUF_REG_G: *count = N; return ctx->reg_g;
    case LFRV_BUF_NORM_MASK: *count = N; return ctx->norm_mask;
    case LFRV_BUF_COVERAGE: *count = N; return ctx->coverage;
    case LFRV_BUF_U_GRAD_X: *count = N; return ctx->u_grad_x;
    case LFRV_BUF_U_GRAD_Y: *count = N; return ctx->u_grad_y;
    case LFRV_BUF_TARGET_MASK: *count = N; return ctx->target_mask;
    case LFRV_BUF_VISIBILITY: *count = N; return ctx->visibility;
    case LFRV_BUF_TARGET_DEPTH: *count = N; return ctx->target_depth;
    case LFRV_BUF_CG_B: *count = 3 * N; return ctx->cg_b;
    case LFRV_BUF_CG_MINV: *count = N; return ctx->cg_minv;
    case LFRV_BUF_TAU_WEIGHT: *count = ctx->tau_weight ? Vn : 0; return ctx->tau_weight;
    default: *count = 0; return nullptr;
    }
}

extern "C" lfrv_status lfrv_get_buffer(lfrv_ctx *ctx, int which, float *out, size_t *count)
{
    if (!ctx || !count) return LFRV_E_ARG;
    if (which == LFRV_BUF_DPART_X || which == LFRV_BUF_DPART_Y) {
        const lfrv_status ds = materialize_dpart(ctx);
        if (ds != LFRV_OK) return ds;
    }
    float *p = buffer_ptr(ctx, which, count);
    if (!p) return fail(ctx, LFRV_E_ARG, "unknown buffer");
    if (!out) return LFRV_OK;
    LFRV_HIP(ctx, hipMemcpyAsync(out, p, *count * sizeof(float), hipMemcpyDefault, ctx->stream));
    LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_set_buffer(lfrv_ctx *ctx, int which, const float *in)
{
    if (!ctx) return LFRV_E_ARG;
    if (which == LFRV_BUF_DPART_X || which == LFRV_BUF_DPART_Y) {
        const lfrv_status ds = materialize_dpart(ctx);  // the other plane stays valid
        if (ds != LFRV_OK) return ds;
    }
    size_t count = 0;
    float *p = buffer_ptr(ctx, which, &count);
    if (!p) return fail(ctx, LFRV_E_ARG, "unknown buffer");
    if (which == LFRV_BUF_U) ctx->u_zero = false;
    lfrv_status st = copy_in(ctx, p, in, count);
    if (st != LFRV_OK) return st;
    if (which == LFRV_BUF_TAU_X || which == LFRV_BUF_TAU_Y) {
        ctx->prepared = false;
        ctx->counts_valid = false;
    }
    if (which == LFRV_BUF_OMEGA) ctx->entry_weights_valid = false;
    ctx->epoch++;  // any overwritten input invalidates the set-up sums
    if (which == LFRV_BUF_REG_G) ctx->has_reg_g = true;
    if (which == LFRV_BUF_VISIBILITY) {
        ctx->vis_ready = true;
        ctx->holes_listed = false;  // the hole list follows the visibility
    }
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_set_timing(lfrv_ctx *ctx, int enable)
{
    if (!ctx) return LFRV_E_ARG;
    if (enable < 0) return fail(ctx, LFRV_E_ARG, "lfrv_set_timing: negative stride");
    flush_timing(ctx);
    ctx->timing = enable != 0;
    ctx->timing_stride = enable > 0 ? enable : 1;
    for (auto &s : ctx->tslot) s = TimingSlot();
    for (int64_t &q : ctx->timing_seq) q = 0;
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_get_timing(lfrv_ctx *ctx, int kclass, double *total_ms, int64_t *launches)
{
    if (!ctx || kclass < 0 || kclass >= 6) return LFRV_E_ARG;
    flush_timing(ctx);
    if (total_ms) *total_ms = ctx->tslot[kclass].ms;
    if (launches) *launches = ctx->tslot[kclass].launches;
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_get_bin_stats(lfrv_ctx *ctx, int64_t *entries, int *tiles)
{
    if (!ctx) return LFRV_E_ARG;
    if (entries) {
        int total = 0;
        if (ctx->prepared) {
            LFRV_HIP(ctx, hipMemcpyAsync(&total, ctx->cell_start + ctx->ncells, sizeof(int), hipMemcpyDeviceToHost,
                                         ctx->stream));
            LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
        }
        *entries = total;
    }
    if (tiles) *tiles = ctx->ntiles;
    return LFRV_OK;
}

// ---- stateless operators -----------------------------------------------------------

extern "C" lfrv_status lfrv_op_gather(int W, int H, int w, int h, int dsf, const float *tau_x, const float *tau_y,
                                      const float *u, int C, float *out, void *stream)
{
    if (W < 1 || H < 1 || w < 1 || h < 1 || dsf < 1 || dsf > 4 || C < 1 || C > 4 || !tau_x || !tau_y || !u || !out)
        return LFRV_E_ARG;
    return launch_gather(W, H, w, h, dsf, tau_x, tau_y, u, C, out, (hipStream_t)stream) == hipSuccess ? LFRV_OK
                                                                                                     : LFRV_E_HIP;
}

extern "C" lfrv_status lfrv_op_splat_atomic(int W, int H, int w, int h, int dsf, const float *tau_x,
                                            const float *tau_y, const float *in, int C, const float *wgt, float *out,
                                            float *wsum, void *stream)
{
    if (W < 1 || H < 1 || w < 1 || h < 1 || dsf < 1 || dsf > 4 || C < 0 || C > 4 || !tau_x || !tau_y) return LFRV_E_ARG;
    if (out && (!in || C < 1)) return LFRV_E_ARG;
    return launch_splat_atomic(W, H, w, h, dsf, tau_x, tau_y, in, C, wgt, out, wsum, (hipStream_t)stream) == hipSuccess
               ? LFRV_OK
               : LFRV_E_HIP;
}

extern "C" lfrv_status lfrv_op_sobel(int W, int H, const float *rgb, float *gx, float *gy, const uint8_t *vis,
                                     void *stream)
{
    if (W < 1 || H < 1 || !rgb || !gx || !gy) return LFRV_E_ARG;
    const size_t N = (size_t)W * H;
    return launch_sobel(W, H, rgb, rgb + N, rgb + 2 * N, gx, gy, nullptr, vis, (hipStream_t)stream) == hipSuccess
               ? LFRV_OK
               : LFRV_E_HIP;
}

extern "C" lfrv_status lfrv_op_radial_to_ortho(int W, int H, const double *K, float *z, float *dzdx, float *dzdy,
                                               void *stream)
{
    if (W < 1 || H < 1 || !K || !z) return LFRV_E_ARG;
    return launch_radial_to_ortho(W, H, K, z, dzdx, dzdy, (hipStream_t)stream) == hipSuccess ? LFRV_OK : LFRV_E_ARG;
}

// ---- device memory helpers -------------------------------------------------------------

extern "C" lfrv_status lfrv_dev_alloc(size_t bytes, void **ptr)
{
    if (!ptr) return LFRV_E_ARG;
    *ptr = nullptr;
    if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) return LFRV_E_HIP;
    return hipMemset(*ptr, 0, bytes ? bytes : 1) == hipSuccess ? LFRV_OK : LFRV_E_HIP;
}

extern "C" lfrv_status lfrv_dev_free(void *ptr) { return hipFree(ptr) == hipSuccess ? LFRV_OK : LFRV_E_HIP; }

extern "C" lfrv_status lfrv_memcpy(void *dst, const void *src, size_t bytes)
{
    if ((!dst || !src) && bytes) return LFRV_E_ARG;
    return hipMemcpy(dst, src, bytes, hipMemcpyDefault) == hipSuccess ? LFRV_OK : LFRV_E_HIP;
}

// Plain streaming copy: 16-byte loads and stores, grid-stride.
__global__ static void __launch_bounds__(256) k_copy16(size_t n16, const uint4 *__restrict__ src, uint4 *__restrict__ dst)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" lfrv_status lfrv_dev_copy_kernel(void *dst, const void *src, size_t bytes, void *stream)
{
    if (!dst || !src || bytes % 16 != 0 || ((uintptr_t)dst | (uintptr_t)src) % 16 != 0) return LFRV_E_ARG;
    const size_t n16 = bytes / 16;
    if (n16 == 0) return LFRV_OK;
    const size_t want = (n16 + 255) / 256;
    const unsigned int nb = (unsigned int)(want < 8192 ? want : 8192);
    k_copy16<<<nb, 256, 0, (hipStream_t)stream>>>(n16, (const uint4 *)src, (uint4 *)dst);
    return hipGetLastError() == hipSuccess ? LFRV_OK : LFRV_E_HIP;
}

extern "C" lfrv_status lfrv_device_sync(void) { return hipDeviceSynchronize() == hipSuccess ? LFRV_OK : LFRV_E_HIP; }

// ---- quality metric ------------------------------------------------------------------

extern "C" double lfrv_psnr(int W, int H, const float *a, const float *b)
{
    // gsl_image_psnr (gsl_image.cpp:991-1025): mean of the per-channel MSE over
    // x, y in [2, W-2) x [2, H-2), vmax = 1.
    if (W <= 4 || H <= 4 || !a || !b) return 0.0;
    const size_t N = (size_t)W * H;
    double mse_t = 0.0;
    for (int c = 0; c < 3; c++) {
        double mse = 0.0;
        for (int y = 2; y < H - 2; y++)
            for (int x = 2; x < W - 2; x++) {
                const double d = (double)a[c * N + (size_t)y * W + x] - (double)b[c * N + (size_t)y * W + x];
                mse += d * d;
            }
        mse_t += mse / ((double)(W - 4) * (H - 4));
    }
    mse_t /= 3.0;
    if (mse_t == 0.0) return DBL_MAX;  // coco::psnr (gsl_image.cpp:1009-1015)
    return 10.0 * std::log(1.0 / mse_t) / std::log(10.0);
}
