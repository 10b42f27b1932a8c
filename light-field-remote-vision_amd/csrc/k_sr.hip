// IBR_color (TV super-resolution with warped data terms): fused data-term
// gradient, the outer proximal step, the primal energy.
#include "tile_common.hpp"

#include <algorithm>
#include <cmath>

namespace lfrv {

lfrv_status run_pd_steps(lfrv_ctx *ctx, int iters, const float *F);  // k_stencil.hip

namespace {

struct DataTermArgs {
    float alpha, beta, sigma2, grad_step, dsf2, prox_factor;  // prox_factor = -1 / (lambda L)
    int gw_type;
};

// Data-term gradient and the clamped gradient step for one target tile
// (vtv_sr_ista_step_unstructured, optimization.cu:1775-1812):
//   G_int  = sum_i B_i^T (omega (S_i U - v_i)) / cov * sigma^2           (:1835-1966)
//   G_grad = sum_i B_i^T ((dsf^2 S_i lap U - lap v_i) [* omega]) / cov [* sigma^2] * (-step)  (:1969-2116)
//   G = clamp(alpha G_int + beta G_grad, -1, 1);  F = clamp(U + prox_factor G, 0, 1)
// U (tile + halo KS) and lap U (tile + halo KS-1) are staged in LDS; every
// binned entry gathers both, forms the residuals and splats them back.
// Writes F (into G), Uq = U and norm_mask = cov (the reference's buffer ends
// holding the coverage).
template <int KS>
__global__ void __launch_bounds__(kPullMaxThreads) k_sr_dataterm(TileGeom g, TileEntries te, int cap, size_t n, DataTermArgs da,
                                                     const float *__restrict__ U, const float *__restrict__ views,
                                                     const float *__restrict__ vlap, const float *__restrict__ coverage,
                                                     float *__restrict__ Gout, float *__restrict__ Uq,
                                                     float *__restrict__ norm_mask)
{
    constexpr int HAL = KS - 1;
    extern __shared__ int smem[];
    const int t = xcd_tile(blockIdx.x, g.ntiles);
    const int tx0 = (t % g.ntx) * g.tw, ty0 = (t / g.ntx) * g.th;
    const int uw = g.tw + 2 * HAL + 2, uh = g.th + 2 * HAL + 2, up = uw * uh;  // U: halo HAL+1
    const int lw = g.tw + 2 * HAL, lh = g.th + 2 * HAL, lpp = lw * lh;         // lap U: halo HAL
    float *ul = reinterpret_cast<float *>(smem);  // 3 * up
    float *ll = ul + 3 * up;                      // 3 * lpp
    int *plds = reinterpret_cast<int *>(ll + 3 * lpp);
    const size_t N = (size_t)g.W * g.H;
    for (int i = threadIdx.x; i < up; i += blockDim.x) {
        const int x = tx0 - HAL - 1 + i % uw, y = ty0 - HAL - 1 + i / uw;
        const bool in = x >= 0 && x < g.W && y >= 0 && y < g.H;
        const size_t o = in ? (size_t)y * g.W + x : 0;
        for (int c = 0; c < 3; c++) ul[c * up + i] = in ? U[c * N + o] : 0.0f;
    }
    __syncthreads();
    const bool need_lap = da.beta > 0.0f;
    if (need_lap) {
        // cuda_laplacian_device (vtv_sr_kernels_unstructured.cu:720-761), no hole gating
        for (int i = threadIdx.x; i < lpp; i += blockDim.x) {
            const int lx = i % lw, ly = i / lw;
            const int x = tx0 - HAL + lx, y = ty0 - HAL + ly;
            const int ui = (ly + 1) * uw + lx + 1;
            for (int c = 0; c < 3; c++) {
                float v = 0.0f;
                if (x >= 0 && x < g.W && y >= 0 && y < g.H) {
                    const float *u = ul + c * up;
                    float s = 0.0f, wsum = 0.0f;
                    if (x > 0) { s += u[ui - 1]; wsum += 1.0f; }
                    if (y > 0) { s += u[ui - uw]; wsum += 1.0f; }
                    if (x < g.W - 1) { s += u[ui + 1]; wsum += 1.0f; }
                    if (y < g.H - 1) { s += u[ui + uw]; wsum += 1.0f; }
                    v = s / wsum - u[ui];
                }
                ll[c * lpp + i] = v;
            }
        }
    }
    // (the pull synchronises before it reads the staged tiles)
    any_pull<KS, 6, 0>(
        g, te, t, plds, cap,
        [&](int e, int x0, int y0, const float(&wx)[KS], const float(&wy)[KS], float om, float(&r)[6]) {
            const int k = te.src[e];
            const unsigned vi = (unsigned)k / (unsigned)n, m = (unsigned)k - vi * (unsigned)n;
            const int bx = x0 - tx0 + HAL, by = y0 - ty0 + HAL;
            float su[3] = {0, 0, 0}, sl[3] = {0, 0, 0};
#pragma unroll
            for (int i = 0; i < KS; i++)
#pragma unroll
                for (int j = 0; j < KS; j++) {
                    const float a = wy[i] * wx[j];
                    const int uidx = (by + i + 1) * uw + bx + j + 1;
                    const int lidx = (by + i) * lw + bx + j;
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        su[c] += a * ul[c * up + uidx];
                        if (need_lap) sl[c] += a * ll[c * lpp + lidx];
                    }
                }
            const float *vp = views + vi * 3 * n + m;
            const float *lp = vlap + vi * 3 * n + m;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                r[c] = (da.alpha > 0.0f) ? (su[c] - vp[c * n]) * om : 0.0f;
                float rg = 0.0f;
                if (need_lap) {
                    rg = sl[c] * da.dsf2 - lp[c * n];
                    if (da.gw_type != 0) rg *= om;
                }
                r[3 + c] = rg;
            }
        },
        [&](size_t o, int x, int y, const float(&acc)[6]) {
            const float cov = coverage[o];
            const int ui = (y - ty0 + HAL + 1) * uw + (x - tx0) + HAL + 1;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                float G = 0.0f;
                if (da.alpha > 0.0f) {
                    float gi = acc[c];
                    if (cov != 0.0f) gi /= cov;
                    gi *= da.sigma2;
                    G += da.alpha * gi;
                }
                if (need_lap) {
                    float gg = acc[3 + c];
                    if (cov != 0.0f) gg /= cov;
                    if (da.gw_type != 0) gg *= da.sigma2;
                    gg *= -da.grad_step;
                    G += da.beta * gg;
                }
                G = fminf(1.0f, fmaxf(-1.0f, G));
                const float uv = ul[c * up + ui];
                G *= da.prox_factor;
                G += uv;
                G = fminf(1.0f, fmaxf(0.0f, G));
                Gout[c * N + o] = G;
                Uq[c * N + o] = uv;
            }
            if (da.alpha > 0.0f || need_lap) norm_mask[o] = cov;
        });
}

// The same data term for dsf >= 2 in box form (see box_pull): the gathers
// S_i U and S_i lap U of an entry are 2 x 2 gathers, at its anchors, of the
// dsf x dsf box sums of U and lap U (zero outside the image), and the splat
// of its residuals is a box pull.  U (the box region of the tile's anchors
// plus a 1-pixel ring) and lap U are staged in the pull's region first; the
// box sums stay in LDS beside it.
template <int KS>
__global__ void __launch_bounds__(kPullMaxThreads)
    k_sr_dataterm_box(TileGeom g, TileEntries te, int cap, size_t n, DataTermArgs da, const float *__restrict__ U,
                      const float *__restrict__ views, const float *__restrict__ vlap,
                      const float *__restrict__ coverage, float *__restrict__ Gout, float *__restrict__ Uq,
                      float *__restrict__ norm_mask)
{
    constexpr int D = KS - 1;
    extern __shared__ int smem[];
    const int t = xcd_tile(blockIdx.x, g.ntiles);
    const int tx0 = (t % g.ntx) * g.tw, ty0 = (t / g.ntx) * g.th;
    const int ww = g.tw + KS - 1, wh = g.th + KS - 1;
    const int aw = ww + 1, ah = wh + 1, ap = aw * ah;  // anchors (ac, ar): box start (tx0 + ac - D, ty0 + ar - D)
    const int uw = g.tw + 2 * D + 2, uh = g.th + 2 * D + 2, up = uw * uh;  // U from (tx0 - D - 1, ty0 - D - 1)
    const int lw = g.tw + 2 * D, lh = g.th + 2 * D, lpp = lw * lh;        // lap U from (tx0 - D, ty0 - D)
    float *bu = reinterpret_cast<float *>(smem);  // 3 * ap box sums of U
    float *bl = bu + 3 * ap;                      // 3 * ap box sums of lap U
    int *pull = reinterpret_cast<int *>(bl + 3 * ap);
    float *us = reinterpret_cast<float *>(pull + 2 * wh + 2);  // staging, inside the pull's region
    float *ls = us + 3 * up;
    const size_t N = (size_t)g.W * g.H;
    const bool need_lap = da.beta > 0.0f;
    for (int i = threadIdx.x; i < up; i += blockDim.x) {
        const int x = tx0 - D - 1 + i % uw, y = ty0 - D - 1 + i / uw;
        const bool in = x >= 0 && x < g.W && y >= 0 && y < g.H;
        const size_t o = in ? (size_t)y * g.W + x : 0;
#pragma unroll
        for (int c = 0; c < 3; c++) us[c * up + i] = in ? U[c * N + o] : 0.0f;
    }
    __syncthreads();
    if (need_lap) {
        // cuda_laplacian_device (vtv_sr_kernels_unstructured.cu:720-761), no hole gating
        for (int i = threadIdx.x; i < lpp; i += blockDim.x) {
            const int lx = i % lw, ly = i / lw;
            const int x = tx0 - D + lx, y = ty0 - D + ly;
            const int ui = (ly + 1) * uw + lx + 1;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                float v = 0.0f;
                if (x >= 0 && x < g.W && y >= 0 && y < g.H) {
                    const float *u = us + c * up;
                    float sm = 0.0f, wsum = 0.0f;
                    if (x > 0) { sm += u[ui - 1]; wsum += 1.0f; }
                    if (y > 0) { sm += u[ui - uw]; wsum += 1.0f; }
                    if (x < g.W - 1) { sm += u[ui + 1]; wsum += 1.0f; }
                    if (y < g.H - 1) { sm += u[ui + uw]; wsum += 1.0f; }
                    v = sm / wsum - u[ui];
                }
                ls[c * lpp + i] = v;
            }
        }
        __syncthreads();
    }
    // box sums at the anchors: anchor (ac, ar) covers U-staging pixels
    // (ac + 1 .. ac + D, ar + 1 .. ar + D) and lap-staging (ac .. ac + D - 1, ...)
    for (int a = threadIdx.x; a < ap; a += blockDim.x) {
        const int ar = a / aw, ac = a - ar * aw;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float su = 0.0f, sl = 0.0f;
#pragma unroll
            for (int l = 0; l < D; l++)
#pragma unroll
                for (int k = 0; k < D; k++) {
                    su += us[c * up + (ar + l + 1) * uw + ac + k + 1];
                    if (need_lap) sl += ls[c * lpp + (ar + l) * lw + ac + k];
                }
            bu[c * ap + a] = su;
            bl[c * ap + a] = sl;
        }
    }
    // (the pull's first barrier orders these sums before the walk and the
    // staging before its own use of the region)
    box_pull<KS, 6, 6>(
        g, te, t, pull, cap,
        [&](int, int x0, int y0, const float(&ax)[2], const float(&ay)[2], float om, const float(&pay)[6],
            float(&r)[6]) {
            const int a0 = (y0 - ty0 + D) * aw + (x0 - tx0 + D);
            float su[3] = {0, 0, 0}, sl[3] = {0, 0, 0};
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    const float w = ay[i] * ax[j];
                    const int a = a0 + i * aw + j;
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        su[c] += w * bu[c * ap + a];
                        if (need_lap) sl[c] += w * bl[c * ap + a];
                    }
                }
#pragma unroll
            for (int c = 0; c < 3; c++) {
                r[c] = (da.alpha > 0.0f) ? (su[c] - pay[c]) * om : 0.0f;
                float rg = 0.0f;
                if (need_lap) {
                    rg = sl[c] * da.dsf2 - pay[3 + c];
                    if (da.gw_type != 0) rg *= om;
                }
                r[3 + c] = rg;
            }
        },
        [&](size_t o, int, int, const float(&acc)[6]) {
            const float cov = coverage[o];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                float G = 0.0f;
                if (da.alpha > 0.0f) {
                    float gi = acc[c];
                    if (cov != 0.0f) gi /= cov;
                    gi *= da.sigma2;
                    G += da.alpha * gi;
                }
                if (need_lap) {
                    float gg = acc[3 + c];
                    if (cov != 0.0f) gg /= cov;
                    if (da.gw_type != 0) gg *= da.sigma2;
                    gg *= -da.grad_step;
                    G += da.beta * gg;
                }
                G = fminf(1.0f, fmaxf(-1.0f, G));
                const float uv = U[c * N + o];
                G *= da.prox_factor;
                G += uv;
                G = fminf(1.0f, fmaxf(0.0f, G));
                Gout[c * N + o] = G;
                Uq[c * N + o] = uv;
            }
            if (da.alpha > 0.0f || need_lap) norm_mask[o] = cov;
        });
}

// theta over-relaxation (vtv_rof.cu:28-45): uq <- u + theta (uq - u); u <- uq_old
__global__ void k_overrelax(size_t n3, float theta, float *__restrict__ uq, float *__restrict__ u)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n3) return;
    const float uv = u[k], uqv = uq[k];
    uq[k] = uv + theta * (uqv - uv);
    u[k] = uqv;
}

// E_TV over pixels whose 3x3 neighbourhood is visible, and the ungated Sobel
// of Y(U) for the gradient term (optimization.cu:1505-1534, 1622-1625).
__global__ void __launch_bounds__(256) k_sr_energy_tv(int W, int H, const float *__restrict__ U,
                                                      const float *__restrict__ vis, float *__restrict__ ugx,
                                                      float *__restrict__ ugy, double *partials)
{
    __shared__ double red[16];
    const size_t N = (size_t)W * H;
    double v[1] = {0.0};
    for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < N; o += (size_t)gridDim.x * blockDim.x) {
        const int ox = (int)(o % W), oy = (int)(o / W);
        float GX = 0.0f, GY = 0.0f;
        bool ok = false;
        if (0 < ox && ox < W - 1 && 0 < oy && oy < H - 1) {
            const float *r = U, *g = U + N, *b = U + 2 * N;
            auto Y = [&](long d) { return 0.2126f * r[o + d] + 0.7152f * g[o + d] + 0.0722f * b[o + d]; };
            const float Ypp = Y(-W - 1), Ypc = Y(-W), Ypn = Y(-W + 1);
            const float Ycp = Y(-1), Ycn = Y(1);
            const float Ynp = Y(W - 1), Ync = Y(W), Ynn = Y(W + 1);
            GX = (Ypn + 2.0f * Ycn + Ynn - Ypp - 2.0f * Ycp - Ynp) * 0.125f;
            GY = (Ynp + 2.0f * Ync + Ynn - Ypp - 2.0f * Ypc - Ypn) * 0.125f;
            ok = true;
            for (int dy = -1; dy <= 1 && ok; dy++)
                for (int dx = -1; dx <= 1; dx++)
                    if (vis[o + (long)dy * W + dx] == 0.0f) ok = false;
        }
        ugx[o] = GX;
        ugy[o] = GY;
        if (ok) v[0] += fabs((double)GX) + fabs((double)GY);
    }
    block_partials<1>(v, partials, red);
}

// E_int and E_grad over every (view, pixel) with gather validity
// (optimization.cu:1549-1606, 1611-1704).
template <int KS>
__global__ void __launch_bounds__(256) k_sr_energy_views(int V, int W, int H, int w, int h,
                                                         const float *__restrict__ tx, const float *__restrict__ ty,
                                                         const float *__restrict__ omega, const float *__restrict__ views,
                                                         const float *__restrict__ vgrad, const float *__restrict__ U,
                                                         const float *__restrict__ ugx, const float *__restrict__ ugy,
                                                         float sigma2, int gw_type, double *partials)
{
    __shared__ double red[2 * 16];
    const size_t n = (size_t)w * h, N = (size_t)W * H;
    double v[2] = {0.0, 0.0};
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)V * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t vi = k / n, m = k - vi * n;
        const float t0 = tx[k], t1 = ty[k];
        float s[5] = {0, 0, 0, 0, 0};
        int x0, y0;
        float wx[KS], wy[KS];
        if (gather_valid(t0, t1, W, H) && footprint<KS>(t0, t1, W, H, x0, y0, wx, wy)) {
            for (int i = 0; i < KS; i++) {
                const int y = y0 + i;
                if (y < 0 || y > H - 1) continue;
                for (int j = 0; j < KS; j++) {
                    const int x = x0 + j;
                    if (x < 0 || x > W - 1) continue;
                    const float a = wy[i] * wx[j];
                    const size_t o = (size_t)y * W + x;
                    s[0] += a * U[o];
                    s[1] += a * U[N + o];
                    s[2] += a * U[2 * N + o];
                    s[3] += a * ugx[o];
                    s[4] += a * ugy[o];
                }
            }
        }
        const float om = omega[k];
        for (int c = 0; c < 3; c++) {
            const float r = s[c] - views[vi * 3 * n + c * n + m];
            v[0] += (double)(r * r) * (double)om * (double)sigma2;
        }
        float a = s[3] - vgrad[vi * 2 * n + m], b = s[4] - vgrad[vi * 2 * n + n + m];
        a *= a;
        b *= b;
        if (gw_type != 0) {
            a = a * om * sigma2;
            b = b * om * sigma2;
        }
        v[1] += (double)a + (double)b;
    }
    block_partials<2>(v, partials, red);
}

__global__ void __launch_bounds__(256) k_count_moving(size_t N, const float *__restrict__ U,
                                                      const float *__restrict__ prev, float thr,
                                                      double *partials)
{
    __shared__ double red[16];
    double v[1] = {0.0};
    for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < N; o += (size_t)gridDim.x * blockDim.x) {
        // the difference in double, as sr_compare_images takes it of its double planes
        double mx = 0.0;
        for (int c = 0; c < 3; c++) mx = fmax(mx, fabs((double)U[c * N + o] - (double)prev[c * N + o]));
        if (!(mx < (double)thr)) v[0] += 1.0;
    }
    block_partials<1>(v, partials, red);
}

unsigned int rblocks(size_t N)
{
    const size_t b = (N + kBlock - 1) / kBlock;
    return (unsigned int)(b < 2048 ? (b == 0 ? 1 : b) : 2048);
}

}  // namespace

lfrv_status run_fista(lfrv_ctx *ctx)
{
    if (!ctx->prepared) return fail(ctx, LFRV_E_STATE, "FISTA step before lfrv_prepare");
    lfrv_status st = run_refresh_entry_weights(ctx, 1.0f);
    if (st != LFRV_OK) return st;
    const TileGeom g = tile_geom(ctx);
    const TileEntries te = tile_entries(ctx);
    DataTermArgs da;
    da.alpha = ctx->p.dt_alpha;
    da.beta = ctx->p.dt_beta;
    da.sigma2 = ctx->p.sigma_sensor * ctx->p.sigma_sensor;
    da.grad_step = ctx->p.gradient_step;
    da.dsf2 = (float)(ctx->dsf * ctx->dsf);
    const float L = (float)(1.0 / (double)ctx->p.lambda);  // optimization.cu:1741
    da.prox_factor = (float)(-1.0 / ((double)ctx->p.lambda * (double)L));
    da.gw_type = ctx->p.gw_type;
    const int HAL = ctx->ks - 1;
    if (ctx->ks == 2) {
        const size_t up = (size_t)(ctx->tile_w + 2 * HAL + 2) * (ctx->tile_h + 2 * HAL + 2);
        const size_t lpp = (size_t)(ctx->tile_w + 2 * HAL) * (ctx->tile_h + 2 * HAL);
        const size_t tiles = sizeof(float) * (3 * up + 3 * lpp);
        const int cap = pull_cap(ctx, 6, tiles, ctx->pull_budget, PullKind::Cell);
        const size_t lds = tiles + pull_lds(ctx, 6, cap, PullKind::Cell);
        TimedRegion tr(ctx, 3);
        k_sr_dataterm<2><<<ctx->ntiles, ctx->pull_bs, lds, ctx->stream>>>(g, te, cap, ctx->n, da, ctx->U, ctx->views,
                                                                          ctx->view_lap, ctx->coverage, ctx->G, ctx->Uq,
                                                                          ctx->norm_mask);
        LFRV_LAUNCH(ctx, "data term");
    } else {
        // box form: box sums of U and lap U at the anchors beside the pull,
        // whose region first holds the staged U and lap U
        const int D = HAL, ww = ctx->tile_w + ctx->ks - 1, wh = ctx->tile_h + ctx->ks - 1;
        const size_t ap = (size_t)(ww + 1) * (wh + 1);
        const size_t sums = sizeof(float) * 6 * ap;
        const size_t up = (size_t)(ctx->tile_w + 2 * D + 2) * (ctx->tile_h + 2 * D + 2);
        const size_t lpp = (size_t)(ctx->tile_w + 2 * D) * (ctx->tile_h + 2 * D);
        const size_t stage = sizeof(int) * (2 * wh + 2) + sizeof(float) * (3 * up + 3 * lpp);
        // entries staged with their 6 payload floats; up to 64 KiB of LDS
        const size_t budget = std::max<size_t>(ctx->pull_budget, 65536 - 512);
        const int cap = pull_cap(ctx, 6, sums, budget, PullKind::Box, 6);
        const size_t lds = sums + std::max(stage, pull_lds(ctx, 6, cap, PullKind::Box, 6));
        TimedRegion tr(ctx, 3);
#define DT(K) k_sr_dataterm_box<K><<<ctx->ntiles, ctx->pull_bs, lds, ctx->stream>>>(g, te, cap, ctx->n, da, ctx->U, ctx->views, ctx->view_lap, ctx->coverage, ctx->G, ctx->Uq, ctx->norm_mask)
        switch (ctx->ks) {
        case 3: DT(3); break;
        case 4: DT(4); break;
        case 5: DT(5); break;
        default: return fail(ctx, LFRV_E_ARG, "dsf out of range");
        }
#undef DT
        LFRV_LAUNCH(ctx, "data term");
    }
    st = run_pd_steps(ctx, ctx->p.inner_iterations, ctx->G);
    if (st != LFRV_OK) return st;
    // FISTA weights (optimization.cu:1744-1746)
    const double an = 0.5 * (1.0 + std::sqrt(1.0 + 4.0 * std::pow(ctx->fista_alpha, 2.0)));
    const float theta = (float)((ctx->fista_alpha - 1.0) / an);
    k_overrelax<<<blocks_for(3 * ctx->N), kBlock, 0, ctx->stream>>>(3 * ctx->N, theta, ctx->Uq, ctx->U);
    LFRV_LAUNCH(ctx, "overrelaxation");
    ctx->fista_alpha = an;
    return LFRV_OK;
}

lfrv_status run_sr_energy(lfrv_ctx *ctx, double *e, double *parts)
{
    const size_t N = ctx->N;
    float *ugx = ctx->tmp3N, *ugy = ctx->tmp3N + N;
    k_sr_energy_tv<<<rblocks(N), 256, 0, ctx->stream>>>(ctx->W, ctx->H, ctx->U, ctx->visibility, ugx, ugy,
                                                        ctx->partials);
    k_reduce_out<1><<<1, kFinalizeThreads, 0, ctx->stream>>>(ctx->partials, (int)rblocks(N), ctx->scal + 8);
    const size_t tot = (size_t)ctx->V * ctx->n;
    const float s2 = ctx->p.sigma_sensor * ctx->p.sigma_sensor;
#define EV(K) k_sr_energy_views<K><<<rblocks(tot), 256, 0, ctx->stream>>>(ctx->V, ctx->W, ctx->H, ctx->w, ctx->h, ctx->tau_x, ctx->tau_y, ctx->omega, ctx->views, ctx->view_grad, ctx->U, ugx, ugy, s2, ctx->p.gw_type, ctx->partials)
    switch (ctx->ks) {
    case 2: EV(2); break;
    case 3: EV(3); break;
    case 4: EV(4); break;
    case 5: EV(5); break;
    default: return fail(ctx, LFRV_E_ARG, "dsf out of range");
    }
#undef EV
    k_reduce_out<2><<<1, kFinalizeThreads, 0, ctx->stream>>>(ctx->partials, (int)rblocks(tot), ctx->scal + 9);
    LFRV_LAUNCH(ctx, "sr energy");
    LFRV_HIP(ctx, hipMemcpyAsync(ctx->scal_host + 8, ctx->scal + 8, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double etv = ctx->scal_host[8], eint = ctx->scal_host[9], egrad = ctx->scal_host[10];
    if (parts) {
        parts[0] = etv;
        parts[1] = eint;
        parts[2] = egrad;
    }
    *e = (double)ctx->p.lambda * etv + (double)ctx->p.dt_alpha * eint + (double)ctx->p.dt_beta * egrad;
    return LFRV_OK;
}

lfrv_status run_count_moving(lfrv_ctx *ctx, const float *prev_dev, float max_diff, int64_t *moving)
{
    k_count_moving<<<rblocks(ctx->N), 256, 0, ctx->stream>>>(ctx->N, ctx->U, prev_dev, max_diff, ctx->partials);
    k_reduce_out<1><<<1, kFinalizeThreads, 0, ctx->stream>>>(ctx->partials, (int)rblocks(ctx->N), ctx->scal + 12);
    LFRV_LAUNCH(ctx, "count moving");
    LFRV_HIP(ctx, hipMemcpyAsync(ctx->scal_host + 12, ctx->scal + 12, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *moving = (int64_t)std::llround(ctx->scal_host[12]);
    return LFRV_OK;
}

}  // namespace lfrv
