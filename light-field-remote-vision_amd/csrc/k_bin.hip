// Splat structure (entries sorted by footprint cell) and the pull kernels.
//
// The reference makes its scatter race-free by colouring source pixels on the
// host into segments whose targets are >= 1 + ks/2 pixels apart and launching
// one kernel per segment (gradientIBR.cu:383-521, optimization.cu:882-1025).
// Here the source pixels are counting-sorted on the GPU by the target cell of
// their footprint's top-left tap and every splat becomes a pull (see
// tile_common.hpp): no atomics in the splats, deterministic sums.
#include "tile_common.hpp"

namespace lfrv {

namespace {

template <int KS>
__device__ __forceinline__ int cell_of(float t0, float t1, int Wc)
{
    constexpr float hd = 0.5f * (float)(KS - 1);
    const int x0 = (int)floorf(t0 - 0.5f - hd + 0.5f);
    const int y0 = (int)floorf(t1 - 0.5f - hd + 0.5f);
    return (y0 + KS - 1) * Wc + (x0 + KS - 1);
}

template <int KS>
__global__ void k_cell_count(size_t tot, int W, int H, int Wc, const float *__restrict__ tx,
                             const float *__restrict__ ty, int *__restrict__ count)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= tot) return;
    const float t0 = tx[k], t1 = ty[k];
    if (!cell_valid(t0, t1, W, H)) return;
    atomicAdd(count + cell_of<KS>(t0, t1, Wc), 1);
}

template <int KS>
__global__ void k_cell_fill(size_t tot, int W, int H, int Wc, const float *__restrict__ tx,
                            const float *__restrict__ ty, const int *__restrict__ start,
                            int *__restrict__ cursor, int *__restrict__ e_src)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= tot) return;
    const float t0 = tx[k], t1 = ty[k];
    if (!cell_valid(t0, t1, W, H)) return;
    const int c = cell_of<KS>(t0, t1, Wc);
    const int pos = start[c] + atomicAdd(cursor + c, 1);
    e_src[pos] = (int)k;
}

// Order every cell's entries by source index (insertion sort; cells hold a
// handful of entries) and gather the per-entry tau.
__global__ void k_cell_sort(int ncells, const int *__restrict__ start, int *__restrict__ e_src,
                            const float *__restrict__ tx, const float *__restrict__ ty,
                            float *__restrict__ e_tx, float *__restrict__ e_ty)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const int a = start[c], b = start[c + 1];
    for (int i = a + 1; i < b; i++) {
        const int key = e_src[i];
        int j = i - 1;
        while (j >= a && e_src[j] > key) {
            e_src[j + 1] = e_src[j];
            j--;
        }
        e_src[j + 1] = key;
    }
    for (int i = a; i < b; i++) {
        const int k = e_src[i];
        e_tx[i] = tx[k];
        e_ty[i] = ty[k];
    }
}

// Device-wide exclusive scan of `in` (n ints) into `out` (n + 1 ints):
// per-block scan with block sums, a single-block scan of the sums, offsets.
constexpr int kScanPer = 8;  // elements per thread

__global__ void k_scan_blocks(int n, const int *__restrict__ in, int *__restrict__ out, int *__restrict__ bsum)
{
    __shared__ int tmp[256];
    const int base = blockIdx.x * 256 * kScanPer + threadIdx.x * kScanPer;
    int v[kScanPer], s = 0;
    for (int i = 0; i < kScanPer; i++) {
        v[i] = (base + i < n) ? in[base + i] : 0;
        s += v[i];
    }
    tmp[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int t = (threadIdx.x >= off) ? tmp[threadIdx.x - off] : 0;
        __syncthreads();
        tmp[threadIdx.x] += t;
        __syncthreads();
    }
    int run = tmp[threadIdx.x] - s;
    for (int i = 0; i < kScanPer; i++) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (threadIdx.x == 255) bsum[blockIdx.x] = tmp[255];
}

__global__ void k_scan_sums(int nb, int *__restrict__ bsum, int *__restrict__ total_out)
{
    __shared__ int tmp[1024];
    const int per = (nb + 1023) / 1024;
    const int b0 = threadIdx.x * per, b1 = min(b0 + per, nb);
    int s = 0;
    for (int i = b0; i < b1; i++) s += bsum[i];
    tmp[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int t = (threadIdx.x >= off) ? tmp[threadIdx.x - off] : 0;
        __syncthreads();
        tmp[threadIdx.x] += t;
        __syncthreads();
    }
    int run = tmp[threadIdx.x] - s;
    for (int i = b0; i < b1; i++) {
        const int c = bsum[i];
        bsum[i] = run;
        run += c;
    }
    if (threadIdx.x == 1023) *total_out = tmp[1023];
}

__global__ void k_scan_add(int n, int *__restrict__ out, const int *__restrict__ bsum)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += bsum[i / (256 * kScanPer)];
}

__global__ void k_refresh_weights(int64_t ne, const int *__restrict__ e_src,
                                  const float *__restrict__ omega, float *__restrict__ e_w)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < ne) e_w[e] = omega[e_src[e]];
}

// ---- pull kernels -----------------------------------------------------------

struct PullCfg {
    int cap;  // LDS slots
};

#define PULL_PROLOGUE(NV)                     \
    extern __shared__ int smem[];             \
    const int t = xcd_tile(blockIdx.x, g.ntiles);

// coverage = sum_i B_i^T 1
template <int KS>
__global__ void __launch_bounds__(1024) k_pull_coverage(TileGeom g, TileEntries te, PullCfg pc, float *__restrict__ coverage)
{
    PULL_PROLOGUE(1)
    any_pull<KS, 1, false>(
        g, te, t, smem, pc.cap,
        [&](int, int, int, const float(&)[KS], const float(&)[KS], float(&v)[1]) { v[0] = 1.0f; },
        [&](size_t o, int, int, const float(&a)[1]) { coverage[o] = a[0]; });
}

// Weighted initial image (gradientIBR.cu:764-811, optimization.cu:1303-1349):
// U = sum B^T (omega v) / sum B^T omega where the weight is non-zero, else 0;
// norm_mask = sum B^T omega.
template <int KS>
__global__ void __launch_bounds__(1024) k_pull_init_image(TileGeom g, TileEntries te, PullCfg pc, size_t n,
                                                         const float *__restrict__ views, float *__restrict__ U,
                                                         float *__restrict__ norm_mask)
{
    PULL_PROLOGUE(4)
    const size_t N = (size_t)g.W * g.H;
    any_pull<KS, 4, false>(
        g, te, t, smem, pc.cap,
        [&](int e, int, int, const float(&)[KS], const float(&)[KS], float(&v)[4]) {
            const int k = te.src[e];
            const size_t vi = (size_t)k / n, m = (size_t)k - vi * n;
            const float om = te.w[e];
            const float *vp = views + vi * 3 * n + m;
            v[0] = om * vp[0];
            v[1] = om * vp[n];
            v[2] = om * vp[2 * n];
            v[3] = om;
        },
        [&](size_t o, int, int, const float(&a)[4]) {
            const float wsum = a[3];
            for (int c = 0; c < 3; c++) U[c * N + o] = (wsum != 0.0f) ? a[c] / wsum : a[c];
            norm_mask[o] = wsum;
        });
}

// Target gradient (init_u_gradient, gradientIBR.cu:858-913): pull (gx, gy,
// omega, 1) with the footprint weights, normalise by the last plane.
template <int KS>
__global__ void __launch_bounds__(1024) k_pull_u_gradient(TileGeom g, TileEntries te, PullCfg pc, size_t n,
                                                         const float *__restrict__ vgrad, float sigma2,
                                                         float *__restrict__ ugx, float *__restrict__ ugy,
                                                         float *__restrict__ tmask, float *__restrict__ norm_mask)
{
    PULL_PROLOGUE(4)
    any_pull<KS, 4, false>(
        g, te, t, smem, pc.cap,
        [&](int e, int, int, const float(&)[KS], const float(&)[KS], float(&v)[4]) {
            const int k = te.src[e];
            const size_t vi = (size_t)k / n, m = (size_t)k - vi * n;
            v[0] = vgrad[vi * 2 * n + m];
            v[1] = vgrad[vi * 2 * n + n + m];
            v[2] = te.w[e];
            v[3] = 1.0f;
        },
        [&](size_t o, int, int, const float(&a)[4]) {
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
        });
}

// Regularizer weight, splat part (optimization.cu:2264-2324 with
// vtv_sr_kernels.cu:1172-1194): the reference accumulates sum B^T 1 on top of
// whatever its normalisation buffer held, so norm_mask += coverage here.
template <int KS>
__global__ void __launch_bounds__(1024) k_pull_reg_weight(TileGeom g, TileEntries te, PullCfg pc,
                                                         const float *__restrict__ coverage, float sigma2,
                                                         float lmax, float lmin, float nviews,
                                                         float *__restrict__ norm_mask, float *__restrict__ graw)
{
    PULL_PROLOGUE(1)
    any_pull<KS, 1, false>(
        g, te, t, smem, pc.cap,
        [&](int e, int, int, const float(&)[KS], const float(&)[KS], float(&v)[1]) { v[0] = te.w[e]; },
        [&](size_t o, int, int, const float(&a)[1]) {
            const float nm = norm_mask[o] + coverage[o];
            norm_mask[o] = nm;
            float m = a[0];
            if (nm != 0.0f) m /= nm;
            m *= sigma2;
            graw[o] = (m == 0.0f) ? lmax : fmaxf(lmin, lmax - m / nviews);
        });
}

// CG right-hand side: b = alpha sum B^T (omega sigma^2 v) + c0 u0 + beta D^T g.
template <int KS>
__global__ void __launch_bounds__(1024) k_pull_rhs(TileGeom g, TileEntries te, PullCfg pc, size_t n,
                                                  const float *__restrict__ views, float wscale, float c0,
                                                  const float *__restrict__ u0, float beta,
                                                  const float *__restrict__ gx, const float *__restrict__ gy,
                                                  float *__restrict__ b)
{
    PULL_PROLOGUE(3)
    const size_t N = (size_t)g.W * g.H;
    any_pull<KS, 3, false>(
        g, te, t, smem, pc.cap,
        [&](int e, int, int, const float(&)[KS], const float(&)[KS], float(&v)[3]) {
            const int k = te.src[e];
            const size_t vi = (size_t)k / n, m = (size_t)k - vi * n;
            const float om = te.w[e] * wscale;
            const float *vp = views + vi * 3 * n + m;
            v[0] = om * vp[0];
            v[1] = om * vp[n];
            v[2] = om * vp[2 * n];
        },
        [&](size_t o, int x, int y, const float(&a)[3]) {
            float dtg = 0.0f;
            if (beta != 0.0f) {
                if (x > 0) dtg += gx[o - 1];
                if (x < g.W - 1) dtg -= gx[o];
                if (y > 0) dtg += gy[o - g.W];
                if (y < g.H - 1) dtg -= gy[o];
                dtg *= beta;
            }
            for (int c = 0; c < 3; c++) b[c * N + o] = a[c] + (c0 != 0.0f ? c0 * u0[c * N + o] : 0.0f) + dtg;
        });
}

// Jacobi preconditioner 1 / diag(A), diag = alpha sum omega sigma^2 B^2 + c0 + beta deg.
template <int KS>
__global__ void __launch_bounds__(1024) k_pull_diag(TileGeom g, TileEntries te, PullCfg pc, float wscale, float c0,
                                                   float beta, float *__restrict__ minv)
{
    PULL_PROLOGUE(1)
    any_pull<KS, 1, true>(
        g, te, t, smem, pc.cap,
        [&](int e, int, int, const float(&)[KS], const float(&)[KS], float(&v)[1]) { v[0] = te.w[e] * wscale; },
        [&](size_t o, int x, int y, const float(&a)[1]) {
            const int deg = (g.W > 1 ? ((x == 0 || x == g.W - 1) ? 1 : 2) : 0) +
                            (g.H > 1 ? ((y == 0 || y == g.H - 1) ? 1 : 2) : 0);
            const float d = a[0] + c0 + beta * (float)deg;
            minv[o] = d > 0.0f ? 1.0f / d : 1.0f;
        });
}

#undef PULL_PROLOGUE

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

TileGeom tile_geom(const lfrv_ctx *ctx)
{
    TileGeom g;
    g.W = ctx->W;
    g.H = ctx->H;
    g.tw = ctx->tile_w;
    g.th = ctx->tile_h;
    g.ntx = ctx->ntx;
    g.ntiles = ctx->ntiles;
    g.Wc = ctx->W + ctx->ks - 1;
    g.Hc = ctx->H + ctx->ks - 1;
    return g;
}

TileEntries tile_entries(const lfrv_ctx *ctx)
{
    TileEntries t;
    t.cell_start = ctx->cell_start;
    t.src = ctx->e_src;
    t.tx = ctx->e_tx;
    t.ty = ctx->e_ty;
    t.w = ctx->e_w;
    t.active = 1;
    return t;
}

// Largest slot count whose LDS fits `budget` bytes besides `extra` bytes.
int pull_cap(const lfrv_ctx *ctx, int nv, size_t extra, size_t budget)
{
    const size_t meta = sizeof(int) * (size_t)pull_meta_ints(ctx->tile_w, ctx->tile_h, ctx->ks, ctx->pull_bs);
    const long avail = (long)budget - (long)meta - (long)extra;
    const long cap = avail / (long)(sizeof(float) * (nv + 2));
    return (int)std::max(64L, std::min(cap, 16384L));
}

size_t pull_lds(const lfrv_ctx *ctx, int nv, int cap)
{
    if (ctx->ks <= 3) {
        const size_t cells = (size_t)(ctx->tile_w + ctx->ks - 1) * (ctx->tile_h + ctx->ks - 1);
        return sizeof(float) * (size_t)ctx->ks * ctx->ks * nv * cells;
    }
    return pull_lds_bytes(ctx->tile_w, ctx->tile_h, ctx->ks, ctx->pull_bs, nv, cap);
}

#define LFRV_KS_DISPATCH(ks, MACRO) \
    switch (ks) {                   \
    case 2: MACRO(2); break;        \
    case 3: MACRO(3); break;        \
    case 4: MACRO(4); break;        \
    case 5: MACRO(5); break;        \
    default: return fail(ctx, LFRV_E_ARG, "dsf out of range"); \
    }

static constexpr size_t kPullBudget = 64 * 1024;

lfrv_status run_prepare(lfrv_ctx *ctx)
{
    const TileGeom g = tile_geom(ctx);
    const size_t tot = (size_t)ctx->V * ctx->n;
    const int nc = ctx->ncells;
    TimedRegion tr(ctx, 5);
    LFRV_HIP(ctx, hipMemsetAsync(ctx->cell_count, 0, sizeof(int) * (nc + 1), ctx->stream));
    LFRV_HIP(ctx, hipMemsetAsync(ctx->cell_cursor, 0, sizeof(int) * nc, ctx->stream));
#define CC(K) k_cell_count<K><<<blocks_for(tot), kBlock, 0, ctx->stream>>>(tot, ctx->W, ctx->H, g.Wc, ctx->tau_x, ctx->tau_y, ctx->cell_count)
    LFRV_KS_DISPATCH(ctx->ks, CC);
#undef CC
    const int nb = (nc + 256 * kScanPer - 1) / (256 * kScanPer);
    k_scan_blocks<<<nb, 256, 0, ctx->stream>>>(nc, ctx->cell_count, ctx->cell_start, ctx->scan_tmp);
    k_scan_sums<<<1, 1024, 0, ctx->stream>>>(nb, ctx->scan_tmp, ctx->cell_start + nc);
    k_scan_add<<<blocks_for(nc), kBlock, 0, ctx->stream>>>(nc, ctx->cell_start, ctx->scan_tmp);
#define CF(K) k_cell_fill<K><<<blocks_for(tot), kBlock, 0, ctx->stream>>>(tot, ctx->W, ctx->H, g.Wc, ctx->tau_x, ctx->tau_y, ctx->cell_start, ctx->cell_cursor, ctx->e_src)
    LFRV_KS_DISPATCH(ctx->ks, CF);
#undef CF
    k_cell_sort<<<blocks_for(nc), kBlock, 0, ctx->stream>>>(nc, ctx->cell_start, ctx->e_src, ctx->tau_x, ctx->tau_y,
                                                           ctx->e_tx, ctx->e_ty);
    LFRV_LAUNCH(ctx, "splat structure");
    int total = 0;
    LFRV_HIP(ctx, hipMemcpyAsync(&total, ctx->cell_start + nc, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if ((size_t)total > ctx->e_cap) return fail(ctx, LFRV_E_ARG, "splat structure overflow");
    ctx->e_total = total;
    const TileEntries te = tile_entries(ctx);
    PullCfg pc{pull_cap(ctx, 1, 0, kPullBudget)};
#define COV(K) k_pull_coverage<K><<<ctx->ntiles, ctx->pull_bs, pull_lds(ctx, 1, pc.cap), ctx->stream>>>(g, te, pc, ctx->coverage)
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
    PullCfg pc{pull_cap(ctx, 4, 0, kPullBudget)};
    TimedRegion tr(ctx, 5);
#define INIT(K) k_pull_init_image<K><<<ctx->ntiles, ctx->pull_bs, pull_lds(ctx, 4, pc.cap), ctx->stream>>>(g, te, pc, ctx->n, ctx->views, ctx->U, ctx->norm_mask)
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
    PullCfg pc{pull_cap(ctx, 4, 0, kPullBudget)};
    const float s2 = ctx->p.sigma_sensor * ctx->p.sigma_sensor;
    TimedRegion tr(ctx, 5);
#define UG(K) k_pull_u_gradient<K><<<ctx->ntiles, ctx->pull_bs, pull_lds(ctx, 4, pc.cap), ctx->stream>>>(g, te, pc, ctx->n, ctx->view_grad, s2, ctx->u_grad_x, ctx->u_grad_y, ctx->target_mask, ctx->norm_mask)
    LFRV_KS_DISPATCH(ctx->ks, UG);
#undef UG
    LFRV_LAUNCH(ctx, "u gradient");
    return LFRV_OK;
}

lfrv_status run_regularizer_weight(lfrv_ctx *ctx)
{
    if (!ctx->prepared) return fail(ctx, LFRV_E_STATE, "regularizer weight before lfrv_prepare");
    lfrv_status st = run_refresh_entry_weights(ctx, 1.0f);
    if (st != LFRV_OK) return st;
    const TileGeom g = tile_geom(ctx);
    const TileEntries te = tile_entries(ctx);
    PullCfg pc{pull_cap(ctx, 1, 0, kPullBudget)};
    const float s2 = ctx->p.sigma_sensor * ctx->p.sigma_sensor;
    const float lmax = ctx->p.lambda_max_factor * ctx->p.lambda, lmin = ctx->p.lambda;
    float *graw = ctx->tmpN;
#define RW(K) k_pull_reg_weight<K><<<ctx->ntiles, ctx->pull_bs, pull_lds(ctx, 1, pc.cap), ctx->stream>>>(g, te, pc, ctx->coverage, s2, lmax, lmin, (float)ctx->V, ctx->norm_mask, graw)
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
    TileEntries te = tile_entries(ctx);
    te.active = use_views ? 1 : 0;
    const float s2 = ctx->p.sigma_sensor * ctx->p.sigma_sensor;
    const float wscale = use_views ? (float)(alpha * s2) : 0.0f;
    TimedRegion tr(ctx, 5);
    {
        PullCfg pc{pull_cap(ctx, 3, 0, kPullBudget)};
#define RHS(K) k_pull_rhs<K><<<ctx->ntiles, ctx->pull_bs, pull_lds(ctx, 3, pc.cap), ctx->stream>>>(g, te, pc, ctx->n, ctx->views, wscale, (float)c0, u0, (float)beta, ctx->u_grad_x, ctx->u_grad_y, ctx->cg_b)
        LFRV_KS_DISPATCH(ctx->ks, RHS);
#undef RHS
    }
    if (use_precond) {
        PullCfg pc{pull_cap(ctx, 1, 0, kPullBudget)};
#define DG(K) k_pull_diag<K><<<ctx->ntiles, ctx->pull_bs, pull_lds(ctx, 1, pc.cap), ctx->stream>>>(g, te, pc, wscale, (float)c0, (float)beta, ctx->cg_minv)
        LFRV_KS_DISPATCH(ctx->ks, DG);
#undef DG
    } else {
        k_fill<<<blocks_for(ctx->N), kBlock, 0, ctx->stream>>>(ctx->cg_minv, ctx->N, 1.0f);
    }
    LFRV_LAUNCH(ctx, "rhs / diag");
    return LFRV_OK;
}

}  // namespace lfrv
