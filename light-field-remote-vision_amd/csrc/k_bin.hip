// Splat structure: source pixels binned by target tile, and the tile kernels
// that accumulate splats in LDS (no global atomics, one store per target pixel).
//
// The reference makes its scatter race-free by colouring source pixels on the
// host into segments whose targets are >= 1 + ks/2 pixels apart and launching
// one kernel per segment (gradientIBR.cu:383-521, optimization.cu:882-1025).
// Here every cell-valid source pixel of every view is appended to the list of
// each target tile its footprint touches (at most 2 x 2 tiles); a workgroup
// owns one tile, walks its list, and accumulates the footprint taps that fall
// inside the tile with LDS float atomics.
#include "tile_common.hpp"

namespace lfrv {

namespace {

template <int KS>
__device__ __forceinline__ void tile_range(float t0, float t1, const TileGeom &g, int &cx0,
                                           int &cx1, int &cy0, int &cy1)
{
    constexpr float hd = 0.5f * (float)(KS - 1);
    const int x0 = (int)floorf(t0 - 0.5f - hd + 0.5f);
    const int y0 = (int)floorf(t1 - 0.5f - hd + 0.5f);
    const int xa = max(x0, 0), xb = min(x0 + KS - 1, g.W - 1);
    const int ya = max(y0, 0), yb = min(y0 + KS - 1, g.H - 1);
    cx0 = xa / g.tw;
    cx1 = xb / g.tw;
    cy0 = ya / g.th;
    cy1 = yb / g.th;
}

template <int KS>
__global__ void k_bin_count(int V, size_t n, TileGeom g, const float *__restrict__ tx,
                            const float *__restrict__ ty, int *__restrict__ count)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)V * n) return;
    const float t0 = tx[k], t1 = ty[k];
    if (!cell_valid(t0, t1, g.W, g.H)) return;
    int cx0, cx1, cy0, cy1;
    tile_range<KS>(t0, t1, g, cx0, cx1, cy0, cy1);
    for (int cy = cy0; cy <= cy1; cy++)
        for (int cx = cx0; cx <= cx1; cx++) atomicAdd(count + cy * g.ntx + cx, 1);
}

template <int KS>
__global__ void k_bin_fill(int V, size_t n, TileGeom g, const float *__restrict__ tx,
                           const float *__restrict__ ty, int *__restrict__ cursor,
                           int *__restrict__ e_src, float *__restrict__ e_tx, float *__restrict__ e_ty)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)V * n) return;
    const float t0 = tx[k], t1 = ty[k];
    if (!cell_valid(t0, t1, g.W, g.H)) return;
    int cx0, cx1, cy0, cy1;
    tile_range<KS>(t0, t1, g, cx0, cx1, cy0, cy1);
    for (int cy = cy0; cy <= cy1; cy++)
        for (int cx = cx0; cx <= cx1; cx++) {
            const int pos = atomicAdd(cursor + cy * g.ntx + cx, 1);
            e_src[pos] = (int)k;
            e_tx[pos] = t0;
            e_ty[pos] = t1;
        }
}

// Exclusive scan of the tile counts in one workgroup of 1024 threads.
__global__ void k_scan_tiles(int ntiles, const int *__restrict__ count, int *__restrict__ start,
                             int *__restrict__ cursor)
{
    __shared__ int part[1024];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int per = (ntiles + nt - 1) / nt;
    const int a = tid * per, b = min(a + per, ntiles);
    int s = 0;
    for (int i = a; i < b; i++) s += count[i];
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < nt; off <<= 1) {
        const int v = (tid >= off) ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = part[tid] - s;
    for (int i = a; i < b; i++) {
        start[i] = run;
        cursor[i] = run;
        run += count[i];
    }
    if (tid == nt - 1) start[ntiles] = part[nt - 1];
}

__global__ void k_refresh_weights(int64_t ne, const int *__restrict__ e_src,
                                  const float *__restrict__ omega, float *__restrict__ e_w)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < ne) e_w[e] = omega[e_src[e]];
}

// ---- generic tile splat ---------------------------------------------------
// For every entry of the tile, vals(e, src, out[NC]) produces NC values that
// are splatted with the footprint weights (times `scale`) into NC LDS planes;
// the epilogue then sees the accumulated planes per pixel.

template <int KS, int NC, class Vals, class Epi>
__device__ __forceinline__ void tile_splat(const TileGeom &g, const TileEntries &te, Vals vals,
                                           Epi epi, float *acc)
{
    const int t = xcd_tile(blockIdx.x, g.ntiles);
    const int tx0 = (t % g.ntx) * g.tw, ty0 = (t / g.ntx) * g.th;
    const int tp = g.tw * g.th;
    for (int i = threadIdx.x; i < NC * tp; i += blockDim.x) acc[i] = 0.0f;
    __syncthreads();
    const int e0 = te.active ? te.start[t] : 0, e1 = te.active ? te.start[t + 1] : 0;
    for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const float t0 = te.tx[e], t1 = te.ty[e];
        int x0, y0;
        float wx[KS], wy[KS];
        if (!footprint<KS>(t0, t1, g.W, g.H, x0, y0, wx, wy)) continue;
        float v[NC];
        vals(e, v);
#pragma unroll
        for (int i = 0; i < KS; i++) {
            const int ly = y0 + i - ty0;
            if (ly < 0 || ly >= g.th) continue;
#pragma unroll
            for (int j = 0; j < KS; j++) {
                const int lx = x0 + j - tx0;
                if (lx < 0 || lx >= g.tw) continue;
                const float a = wy[i] * wx[j];
                if (a == 0.0f) continue;
#pragma unroll
                for (int c = 0; c < NC; c++) atomicAdd(acc + c * tp + ly * g.tw + lx, a * v[c]);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < tp; i += blockDim.x) {
        const int lx = i % g.tw, ly = i / g.tw;
        const int x = tx0 + lx, y = ty0 + ly;
        if (x >= g.W || y >= g.H) continue;
        float a[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) a[c] = acc[c * tp + i];
        epi((size_t)y * g.W + x, x, y, a);
    }
}

// coverage = sum_i B_i^T 1
template <int KS>
__global__ void k_tile_coverage(TileGeom g, TileEntries te, float *__restrict__ coverage)
{
    extern __shared__ float acc[];
    tile_splat<KS, 1>(
        g, te, [&](int, float *v) { v[0] = 1.0f; },
        [&](size_t o, int, int, const float *a) { coverage[o] = a[0]; }, acc);
}

// Weighted initial image (gradientIBR.cu:764-811, optimization.cu:1303-1349):
// U = sum B^T (omega v) / sum B^T omega where the weight is non-zero, else 0;
// norm_mask = sum B^T omega.
template <int KS>
__global__ void k_tile_init_image(TileGeom g, TileEntries te, size_t n,
                                  const float *__restrict__ views, float *__restrict__ U,
                                  float *__restrict__ norm_mask)
{
    extern __shared__ float acc[];
    const size_t N = (size_t)g.W * g.H;
    tile_splat<KS, 4>(
        g, te,
        [&](int e, float *v) {
            const int k = te.src[e];
            const size_t vi = (size_t)k / n, m = (size_t)k - vi * n;
            const float om = te.w[e];
            const float *vp = views + vi * 3 * n + m;
            v[0] = om * vp[0];
            v[1] = om * vp[n];
            v[2] = om * vp[2 * n];
            v[3] = om;
        },
        [&](size_t o, int, int, const float *a) {
            const float wsum = a[3];
            for (int c = 0; c < 3; c++) U[c * N + o] = (wsum != 0.0f) ? a[c] / wsum : a[c];
            norm_mask[o] = wsum;
        },
        acc);
}

// Target gradient (init_u_gradient, gradientIBR.cu:858-913): splat (gx, gy,
// omega, 1) with the footprint weights, normalise by the last plane.
template <int KS>
__global__ void k_tile_u_gradient(TileGeom g, TileEntries te, size_t n, const float *__restrict__ vgrad,
                                  float sigma2, float *__restrict__ ugx, float *__restrict__ ugy,
                                  float *__restrict__ tmask, float *__restrict__ norm_mask)
{
    extern __shared__ float acc[];
    tile_splat<KS, 4>(
        g, te,
        [&](int e, float *v) {
            const int k = te.src[e];
            const size_t vi = (size_t)k / n, m = (size_t)k - vi * n;
            v[0] = vgrad[vi * 2 * n + m];
            v[1] = vgrad[vi * 2 * n + n + m];
            v[2] = te.w[e];
            v[3] = 1.0f;
        },
        [&](size_t o, int, int, const float *a) {
            const float nm = a[3];
            float gx = a[0], gy = a[1], tm = a[2];
            if (nm != 0.0f) {
                gx /= nm;
                gy /= nm;
                tm /= nm;
            }
            ugx[o] = gx;
            ugy[o] = gy;
            tmask[o] = tm * sigma2;
            norm_mask[o] = nm;
        },
        acc);
}

// Regularizer weight, splat part (optimization.cu:2264-2324 with
// vtv_sr_kernels.cu:1172-1194): the reference accumulates sum B^T 1 on top of
// whatever its normalisation buffer held, so norm_mask += coverage here.
template <int KS>
__global__ void k_tile_reg_weight(TileGeom g, TileEntries te, const float *__restrict__ coverage,
                                  float sigma2, float lmax, float lmin, float nviews,
                                  float *__restrict__ norm_mask, float *__restrict__ graw)
{
    extern __shared__ float acc[];
    tile_splat<KS, 1>(
        g, te, [&](int e, float *v) { v[0] = te.w[e]; },
        [&](size_t o, int, int, const float *a) {
            const float nm = norm_mask[o] + coverage[o];
            norm_mask[o] = nm;
            float m = a[0];
            if (nm != 0.0f) m /= nm;
            m *= sigma2;
            graw[o] = (m == 0.0f) ? lmax : fmaxf(lmin, lmax - m / nviews);
        },
        acc);
}

// CG right-hand side: b = alpha sum B^T (omega sigma^2 v) + c0 u0 + beta D^T g.
template <int KS>
__global__ void k_tile_rhs(TileGeom g, TileEntries te, size_t n, const float *__restrict__ views,
                           float wscale, float c0, const float *__restrict__ u0, float beta,
                           const float *__restrict__ gx, const float *__restrict__ gy,
                           float *__restrict__ b)
{
    extern __shared__ float acc[];
    const size_t N = (size_t)g.W * g.H;
    tile_splat<KS, 3>(
        g, te,
        [&](int e, float *v) {
            const int k = te.src[e];
            const size_t vi = (size_t)k / n, m = (size_t)k - vi * n;
            const float om = te.w[e] * wscale;
            const float *vp = views + vi * 3 * n + m;
            v[0] = om * vp[0];
            v[1] = om * vp[n];
            v[2] = om * vp[2 * n];
        },
        [&](size_t o, int x, int y, const float *a) {
            float dtg = 0.0f;
            if (beta != 0.0f) {
                if (x > 0) dtg += gx[o - 1];
                if (x < g.W - 1) dtg -= gx[o];
                if (y > 0) dtg += gy[o - g.W];
                if (y < g.H - 1) dtg -= gy[o];
                dtg *= beta;
            }
            for (int c = 0; c < 3; c++) b[c * N + o] = a[c] + (c0 != 0.0f ? c0 * u0[c * N + o] : 0.0f) + dtg;
        },
        acc);
}

// Jacobi preconditioner 1 / diag(A), diag = alpha sum omega sigma^2 B^2 + c0 + beta deg.
template <int KS>
__global__ void k_tile_diag(TileGeom g, TileEntries te, float wscale, float c0, float beta,
                            float *__restrict__ minv)
{
    extern __shared__ float acc[];
    const int t = xcd_tile(blockIdx.x, g.ntiles);
    const int tx0 = (t % g.ntx) * g.tw, ty0 = (t / g.ntx) * g.th;
    const int tp = g.tw * g.th;
    for (int i = threadIdx.x; i < tp; i += blockDim.x) acc[i] = 0.0f;
    __syncthreads();
    if (wscale != 0.0f && te.active)
        for (int e = te.start[t] + threadIdx.x; e < te.start[t + 1]; e += blockDim.x) {
            int x0, y0;
            float wx[KS], wy[KS];
            if (!footprint<KS>(te.tx[e], te.ty[e], g.W, g.H, x0, y0, wx, wy)) continue;
            const float om = te.w[e] * wscale;
            for (int i = 0; i < KS; i++) {
                const int ly = y0 + i - ty0;
                if (ly < 0 || ly >= g.th) continue;
                for (int j = 0; j < KS; j++) {
                    const int lx = x0 + j - tx0;
                    if (lx < 0 || lx >= g.tw) continue;
                    const float a = wy[i] * wx[j];
                    if (a != 0.0f) atomicAdd(acc + ly * g.tw + lx, om * a * a);
                }
            }
        }
    __syncthreads();
    for (int i = threadIdx.x; i < tp; i += blockDim.x) {
        const int x = tx0 + i % g.tw, y = ty0 + i / g.tw;
        if (x >= g.W || y >= g.H) continue;
        const int deg = (g.W > 1 ? ((x == 0 || x == g.W - 1) ? 1 : 2) : 0) +
                        (g.H > 1 ? ((y == 0 || y == g.H - 1) ? 1 : 2) : 0);
        const float d = acc[i] + c0 + beta * (float)deg;
        minv[(size_t)y * g.W + x] = d > 0.0f ? 1.0f / d : 1.0f;
    }
}

__global__ void k_fill(float *__restrict__ p, size_t N, float v)
{
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o < N) p[o] = v;
}

// Visibility from the coverage: GRADIENT_IBR marks any footprint hit
// (coverage > 0, setup_visibility_mask gradientIBR_kernels.cuh:4-54 sets the
// flag for dx*dy != 0 and the normalised weights are positive exactly there);
// COLOR_SR requires coverage >= 1 (vtv_sr_kernels_unstructured.cu:76-93).
__global__ void k_visibility(size_t N, const float *__restrict__ coverage, int color, float *__restrict__ vis)
{
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= N) return;
    const float c = coverage[o];
    vis[o] = color ? (c < 1.0f ? 0.0f : 1.0f) : (c > 0.0f ? 1.0f : 0.0f);
}

__global__ void k_add_to(size_t N, const float *__restrict__ src, float *__restrict__ dst)
{
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o < N) dst[o] += src[o];
}

}  // namespace

size_t tile_lds_bytes(const lfrv_ctx *ctx, int planes) { return sizeof(float) * (size_t)planes * ctx->tile_w * ctx->tile_h; }

TileGeom tile_geom(const lfrv_ctx *ctx)
{
    TileGeom g;
    g.W = ctx->W;
    g.H = ctx->H;
    g.tw = ctx->tile_w;
    g.th = ctx->tile_h;
    g.ntx = ctx->ntx;
    g.ntiles = ctx->ntiles;
    return g;
}

TileEntries tile_entries(const lfrv_ctx *ctx)
{
    TileEntries t;
    t.start = ctx->tile_start;
    t.src = ctx->e_src;
    t.tx = ctx->e_tx;
    t.ty = ctx->e_ty;
    t.w = ctx->e_w;
    t.active = 1;
    return t;
}

#define LFRV_KS_DISPATCH(ks, MACRO) \
    switch (ks) {                   \
    case 2: MACRO(2); break;        \
    case 3: MACRO(3); break;        \
    case 4: MACRO(4); break;        \
    case 5: MACRO(5); break;        \
    default: return fail(ctx, LFRV_E_ARG, "dsf out of range"); \
    }

lfrv_status run_view_stencils(lfrv_ctx *ctx);  // k_warp.hip

lfrv_status run_prepare(lfrv_ctx *ctx)
{
    const TileGeom g = tile_geom(ctx);
    const size_t tot = (size_t)ctx->V * ctx->n;
    TimedRegion tr(ctx, 5);
    LFRV_HIP(ctx, hipMemsetAsync(ctx->tile_count, 0, sizeof(int) * ctx->ntiles, ctx->stream));
#define BIN_COUNT(K) k_bin_count<K><<<blocks_for(tot), kBlock, 0, ctx->stream>>>(ctx->V, ctx->n, g, ctx->tau_x, ctx->tau_y, ctx->tile_count)
    LFRV_KS_DISPATCH(ctx->ks, BIN_COUNT);
#undef BIN_COUNT
    k_scan_tiles<<<1, 1024, 0, ctx->stream>>>(ctx->ntiles, ctx->tile_count, ctx->tile_start, ctx->tile_cursor);
    int total = 0;
    LFRV_HIP(ctx, hipMemcpyAsync(&total, ctx->tile_start + ctx->ntiles, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if ((size_t)total > ctx->e_cap) return fail(ctx, LFRV_E_ARG, "splat structure overflow");
    ctx->e_total = total;
#define BIN_FILL(K) k_bin_fill<K><<<blocks_for(tot), kBlock, 0, ctx->stream>>>(ctx->V, ctx->n, g, ctx->tau_x, ctx->tau_y, ctx->tile_cursor, ctx->e_src, ctx->e_tx, ctx->e_ty)
    LFRV_KS_DISPATCH(ctx->ks, BIN_FILL);
#undef BIN_FILL
    LFRV_LAUNCH(ctx, "binning");
    const TileEntries te = tile_entries(ctx);
    const size_t lds = tile_lds_bytes(ctx, 1);
#define COV(K) k_tile_coverage<K><<<ctx->ntiles, kBlock, lds, ctx->stream>>>(g, te, ctx->coverage)
    LFRV_KS_DISPATCH(ctx->ks, COV);
#undef COV
    LFRV_LAUNCH(ctx, "coverage");
    lfrv_status st = run_view_stencils(ctx);
    if (st != LFRV_OK) return st;
    ctx->prepared = true;
    ctx->entry_weights_valid = false;
    return LFRV_OK;
}

lfrv_status run_refresh_entry_weights(lfrv_ctx *ctx, float)
{
    if (ctx->entry_weights_valid) return LFRV_OK;
    if (ctx->e_total > 0)
        k_refresh_weights<<<blocks_for((size_t)ctx->e_total), kBlock, 0, ctx->stream>>>(ctx->e_total, ctx->e_src,
                                                                                       ctx->omega, ctx->e_w);
    LFRV_LAUNCH(ctx, "refresh weights");
    ctx->entry_weights_valid = true;
    return LFRV_OK;
}

lfrv_status run_visibility(lfrv_ctx *ctx)
{
    if (!ctx->prepared) return fail(ctx, LFRV_E_STATE, "lfrv_setup_visibility before lfrv_prepare");
    const int color = ctx->p.mode == LFRV_MODE_COLOR_SR;
    k_visibility<<<blocks_for(ctx->N), kBlock, 0, ctx->stream>>>(ctx->N, ctx->coverage, color, ctx->visibility);
    if (color)  // the reference splats into its normalisation buffer without clearing it (optimization.cu:108-117)
        k_add_to<<<blocks_for(ctx->N), kBlock, 0, ctx->stream>>>(ctx->N, ctx->coverage, ctx->norm_mask);
    LFRV_LAUNCH(ctx, "visibility");
    return LFRV_OK;
}

lfrv_status run_initial_image(lfrv_ctx *ctx)
{
    if (!ctx->prepared) return fail(ctx, LFRV_E_STATE, "initial image before lfrv_prepare");
    lfrv_status st = run_refresh_entry_weights(ctx, 1.0f);
    if (st != LFRV_OK) return st;
    const TileGeom g = tile_geom(ctx);
    const TileEntries te = tile_entries(ctx);
    const size_t lds = tile_lds_bytes(ctx, 4);
    TimedRegion tr(ctx, 5);
#define INIT(K) k_tile_init_image<K><<<ctx->ntiles, kBlock, lds, ctx->stream>>>(g, te, ctx->n, ctx->views, ctx->U, ctx->norm_mask)
    LFRV_KS_DISPATCH(ctx->ks, INIT);
#undef INIT
    LFRV_LAUNCH(ctx, "initial image");
    return LFRV_OK;
}

lfrv_status run_u_gradient(lfrv_ctx *ctx)
{
    if (!ctx->prepared) return fail(ctx, LFRV_E_STATE, "lfrv_init_u_gradient before lfrv_prepare");
    lfrv_status st = run_refresh_entry_weights(ctx, 1.0f);
    if (st != LFRV_OK) return st;
    const TileGeom g = tile_geom(ctx);
    const TileEntries te = tile_entries(ctx);
    const size_t lds = tile_lds_bytes(ctx, 4);
    const float s2 = ctx->p.sigma_sensor * ctx->p.sigma_sensor;
    TimedRegion tr(ctx, 5);
#define UG(K) k_tile_u_gradient<K><<<ctx->ntiles, kBlock, lds, ctx->stream>>>(g, te, ctx->n, ctx->view_grad, s2, ctx->u_grad_x, ctx->u_grad_y, ctx->target_mask, ctx->norm_mask)
    LFRV_KS_DISPATCH(ctx->ks, UG);
#undef UG
    LFRV_LAUNCH(ctx, "u gradient");
    return LFRV_OK;
}

lfrv_status run_gauss_smooth(lfrv_ctx *ctx, const float *in, float *out, float *tmp);  // k_stencil.hip

lfrv_status run_regularizer_weight(lfrv_ctx *ctx)
{
    if (!ctx->prepared) return fail(ctx, LFRV_E_STATE, "regularizer weight before lfrv_prepare");
    lfrv_status st = run_refresh_entry_weights(ctx, 1.0f);
    if (st != LFRV_OK) return st;
    const TileGeom g = tile_geom(ctx);
    const TileEntries te = tile_entries(ctx);
    const size_t lds = tile_lds_bytes(ctx, 1);
    const float s2 = ctx->p.sigma_sensor * ctx->p.sigma_sensor;
    const float lmax = ctx->p.lambda_max_factor * ctx->p.lambda, lmin = ctx->p.lambda;
    float *graw = ctx->tmpN;
#define RW(K) k_tile_reg_weight<K><<<ctx->ntiles, kBlock, lds, ctx->stream>>>(g, te, ctx->coverage, s2, lmax, lmin, (float)ctx->V, ctx->norm_mask, graw)
    LFRV_KS_DISPATCH(ctx->ks, RW);
#undef RW
    LFRV_LAUNCH(ctx, "regularizer weight");
    st = run_gauss_smooth(ctx, graw, ctx->reg_g, ctx->tmp3N);
    if (st != LFRV_OK) return st;
    ctx->has_reg_g = true;
    return LFRV_OK;
}

lfrv_status run_rhs_and_diag(lfrv_ctx *ctx, double alpha, double c0, double beta, bool use_views,
                             bool use_precond, const float *u0)
{
    const TileGeom g = tile_geom(ctx);
    const TileEntries te = tile_entries(ctx);
    const float s2 = ctx->p.sigma_sensor * ctx->p.sigma_sensor;
    const float wscale = use_views ? (float)(alpha * s2) : 0.0f;
    TimedRegion tr(ctx, 5);
    if (use_views) {
        const size_t lds = tile_lds_bytes(ctx, 3);
#define RHS(K) k_tile_rhs<K><<<ctx->ntiles, kBlock, lds, ctx->stream>>>(g, te, ctx->n, ctx->views, wscale, (float)c0, u0, (float)beta, ctx->u_grad_x, ctx->u_grad_y, ctx->cg_b)
        LFRV_KS_DISPATCH(ctx->ks, RHS);
#undef RHS
    } else {
        TileEntries none = te;
        none.active = 0;
        const size_t lds = tile_lds_bytes(ctx, 3);
#define RHS0(K) k_tile_rhs<K><<<ctx->ntiles, kBlock, lds, ctx->stream>>>(g, none, ctx->n, ctx->views, 0.0f, (float)c0, u0, (float)beta, ctx->u_grad_x, ctx->u_grad_y, ctx->cg_b)
        LFRV_KS_DISPATCH(ctx->ks, RHS0);
#undef RHS0
    }
    if (use_precond) {
        const size_t lds = tile_lds_bytes(ctx, 1);
#define DG(K) k_tile_diag<K><<<ctx->ntiles, kBlock, lds, ctx->stream>>>(g, te, wscale, (float)c0, (float)beta, ctx->cg_minv)
        LFRV_KS_DISPATCH(ctx->ks, DG);
#undef DG
    } else {
        k_fill<<<blocks_for(ctx->N), kBlock, 0, ctx->stream>>>(ctx->cg_minv, ctx->N, 1.0f);
    }
    LFRV_LAUNCH(ctx, "rhs / diag");
    return LFRV_OK;
}

}  // namespace lfrv
