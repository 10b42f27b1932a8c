// Stencil kernels: hole filling (Poisson diffusion, push-pull), the Gaussian
// smoothing of the TV weight, and the fused primal-dual ROF step.
#include "tile_common.hpp"

#include <hip/hip_ext.h>

#include <cmath>

namespace lfrv {

namespace {

// ---- Poisson diffusion in the holes (gradientIBR.cu:87-152) ---------------

__global__ void k_hole_list(size_t N, const float *__restrict__ vis, int *__restrict__ list,
                            unsigned int *__restrict__ count)
{
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= N || vis[o] != 0.0f) return;
    const unsigned int k = atomicAdd(count, 1u);
    list[k] = (int)o;
}

struct HoleState {  // read by the host after each chunk of sweeps
    double e;
    int iters;
    int done;
};

// The diffusion's bookkeeping between sweeps: eh[j] holds the energy of sweep
// j (eh[0] = 0, the energy the first sweep is compared with); *stopped the
// number of sweeps run once the stopping rule has fired.
struct HoleLoop {
    int maxit;
    double eps;
    double *eh;      // maxit + 1
    int *stopped;
    int G;           // blocks of a sweep (its partials per sweep)
};

__global__ void k_hole_init(HoleLoop hl)
{
    if (threadIdx.x == 0) {
        hl.eh[0] = 0.0;
        *hl.stopped = 0;
    }
}

// The stopping rule of gradientIBR.cu:139-143 after sweep j (>= 1), from that
// sweep's block partials: |E_{j-1} - E_j| < eps, or j >= maxit.  Every block
// evaluates it the same way (a deterministic reduction of G partials).
__device__ __forceinline__ bool hole_rule(const HoleLoop &hl, int j, const double *pj, double *red, double &e)
{
    double v[1];
    reduce_partials<1>(pj, hl.G, v, red);
    __shared__ double eb;
    if (threadIdx.x == 0) eb = v[0];  // the sum is valid in thread 0
    __syncthreads();
    e = eb;
    return fabs(hl.eh[j - 1] - e) < hl.eps || j >= hl.maxit;
}

// *hl.stopped read once per block (every thread then takes the same branch)
__device__ __forceinline__ int hole_stopped(const HoleLoop &hl)
{
    __shared__ int sb;
    if (threadIdx.x == 0) sb = *hl.stopped;
    __syncthreads();
    return sb;
}

// Sweep k (1-based) of the Poisson diffusion: first the stopping rule of sweep
// k - 1 (so the rule needs no kernel of its own), then one Jacobi sweep
// u <- mean of the 4-neighbours (the reference's laplacian then add_to) on the
// hole pixels, all three channels, with the energy sum(lap * u_new) of the
// last channel as this sweep's block partials.  pin / pout: the partials of
// sweeps k - 1 and k.
__global__ void __launch_bounds__(256) k_hole_sweep(int W, int H, int nholes, const int *__restrict__ list,
                                                    const float *__restrict__ uin, float *__restrict__ uout,
                                                    HoleLoop hl, int k, const double *pin, double *pout)
{
    __shared__ double red[16];
    if (hole_stopped(hl)) return;
    if (k >= 2) {
        double e;
        const bool done = hole_rule(hl, k - 1, pin, red, e);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            hl.eh[k - 1] = e;
            if (done) *hl.stopped = k - 1;
        }
        if (done) return;
    }
    const size_t N = (size_t)W * H;
    double v[1] = {0.0};
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nholes; k += gridDim.x * blockDim.x) {
        const int o = list[k];
        const int x = o % W, y = o / W;
        for (int c = 0; c < 3; c++) {
            const float *u = uin + c * N;
            float s = 0.0f, wsum = 0.0f;
            if (x > 0) { s += u[o - 1]; wsum += 1.0f; }
            if (y > 0) { s += u[o - W]; wsum += 1.0f; }
            if (x < W - 1) { s += u[o + 1]; wsum += 1.0f; }
            if (y < H - 1) { s += u[o + W]; wsum += 1.0f; }
            const float lap = s / wsum - u[o];
            const float un = u[o] + lap;
            uout[c * N + o] = un;
            if (c == 2) v[0] += (double)lap * (double)un;
        }
    }
    block_partials<1>(v, pout, red);
}

// The state after the last sweep `last` of a chunk, for the host: the sweeps
// run and whether the rule has fired (one block).
__global__ void __launch_bounds__(kFinalizeThreads) k_hole_state(HoleLoop hl, int last, const double *plast,
                                                                 HoleState *st)
{
    __shared__ double red[16];
    int iters = hole_stopped(hl), done = iters != 0;
    double e = 0.0;
    if (!done) {
        done = hole_rule(hl, last, plast, red, e);
        iters = last;
    }
    if (threadIdx.x == 0) {
        st->iters = iters;
        st->done = done;
        st->e = done && *hl.stopped ? hl.eh[iters] : e;
    }
}

// ---- separable Gaussian, clamp to edge (cuda_convolutions.cu:110-237) -------

struct Taps {
    float k[15];
    int r;
};

__global__ void k_conv_rows(int W, int H, Taps t, const float *__restrict__ in, float *__restrict__ out)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    float s = 0.0f;
    for (int k = -t.r; k <= t.r; k++) {
        const int xx = min(max(x + k, 0), W - 1);
        s += in[(size_t)y * W + xx] * t.k[t.r - k];
    }
    out[(size_t)y * W + x] = s;
}

__global__ void k_conv_cols(int W, int H, Taps t, const float *__restrict__ in, float *__restrict__ out)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    float s = 0.0f;
    for (int k = -t.r; k <= t.r; k++) {
        const int yy = min(max(y + k, 0), H - 1);
        s += in[(size_t)yy * W + x] * t.k[t.r - k];
    }
    out[(size_t)y * W + x] = s;
}

// ---- push-pull (optimization.cu:210-651), three channels per launch -------
// Images of level s live at img + c*lvl_stride + off[s]; visibility at vis + off[s].
// The centre tap of expand reads row/column i/2 of the coarser level; for an
// odd-sized level that index runs one past the coarser image in the
// reference (optimization.cu:455, 492) and is clamped here.

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__global__ void k_pp_reduce_rows(int Ws, int Hs, int hs, const float *__restrict__ in, float *__restrict__ tmp,
                                 size_t in_cstride, size_t tmp_cstride)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y, c = blockIdx.z;
    if (j >= Ws || i >= hs) return;
    const float a = 0.4f;
    const float k[5] = {0.25f - 0.5f * a, 0.25f, a, 0.25f, 0.25f - 0.5f * a};
    const int rows[5] = {max(2 * i - 2, 0), max(2 * i - 1, 0), 2 * i, min(2 * i + 1, Hs - 1), min(2 * i + 2, Hs - 1)};
    const float *p = in + c * in_cstride;
    float s = 0.0f, wt = 0.0f;
    for (int q = 0; q < 5; q++) {
        const float v = p[(size_t)rows[q] * Ws + j];
        if (v != 0.0f) { s += k[q] * v; wt += k[q]; }
    }
    tmp[c * tmp_cstride + (size_t)i * Ws + j] = (wt == 0.0f) ? 0.0f : s / wt;
}

__global__ void k_pp_reduce_cols(int Ws, int hs, int ws, const float *__restrict__ tmp, float *__restrict__ out,
                                 size_t tmp_cstride, size_t out_cstride)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y, c = blockIdx.z;
    if (j >= ws || i >= hs) return;
    const float a = 0.4f;
    const float k[5] = {0.25f - 0.5f * a, 0.25f, a, 0.25f, 0.25f - 0.5f * a};
    const int cols[5] = {max(2 * j - 2, 0), max(2 * j - 1, 0), 2 * j, min(2 * j + 1, Ws - 1), min(2 * j + 2, Ws - 1)};
    const float *p = tmp + c * tmp_cstride + (size_t)i * Ws;
    float s = 0.0f, wt = 0.0f;
    for (int q = 0; q < 5; q++) {
        const float v = p[cols[q]];
        if (v != 0.0f) { s += k[q] * v; wt += k[q]; }
    }
    out[c * out_cstride + (size_t)i * ws + j] = (wt == 0.0f) ? 0.0f : s / wt;
}

// visibility pyramid: OR over the same 5 x 5 support
__global__ void k_pp_reduce_vis(int Ws, int Hs, int ws, int hs, const uint8_t *__restrict__ in, uint8_t *__restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= ws || i >= hs) return;
    uint8_t v = 0;
    const int rows[5] = {max(2 * i - 2, 0), max(2 * i - 1, 0), 2 * i, min(2 * i + 1, Hs - 1), min(2 * i + 2, Hs - 1)};
    const int cols[5] = {max(2 * j - 2, 0), max(2 * j - 1, 0), 2 * j, min(2 * j + 1, Ws - 1), min(2 * j + 2, Ws - 1)};
    for (int a = 0; a < 5; a++)
        for (int b = 0; b < 5; b++) v |= in[(size_t)rows[a] * Ws + cols[b]] ? 1 : 0;
    out[(size_t)i * ws + j] = v;
}

__global__ void k_pp_expand_rows(int ws, int hs, int Hs, const float *__restrict__ in, float *__restrict__ tmp,
                                 size_t in_cstride, size_t tmp_cstride)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y, c = blockIdx.z;
    if (j >= ws || i >= Hs) return;
    const float a = 0.4f;
    const float k[5] = {0.25f - 0.5f * a, 0.25f, a, 0.25f, 0.25f - 0.5f * a};
    const float *p = in + c * in_cstride;
    float s = 0.0f;
    if (i % 2 == 0) {
        const int r0 = max((i - 2) / 2, 0), r2 = min((i + 2) / 2, hs - 1);
        s += 2.0f * k[0] * p[(size_t)r0 * ws + j];
        s += 2.0f * k[2] * p[(size_t)min(i / 2, hs - 1) * ws + j];
        s += 2.0f * k[4] * p[(size_t)r2 * ws + j];
    } else {
        const int r1 = max((i - 1) / 2, 0), r3 = min((i + 1) / 2, hs - 1);
        s += 2.0f * k[1] * p[(size_t)r1 * ws + j];
        s += 2.0f * k[3] * p[(size_t)r3 * ws + j];
    }
    tmp[c * tmp_cstride + (size_t)i * ws + j] = s;
}

__global__ void k_pp_expand_cols(int Ws, int Hs, int ws, const float *__restrict__ tmp, float *__restrict__ out,
                                 const uint8_t *__restrict__ vis, size_t tmp_cstride, size_t out_cstride)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y, c = blockIdx.z;
    if (j >= Ws || i >= Hs) return;
    if (vis[(size_t)i * Ws + j]) return;
    const float a = 0.4f;
    const float k[5] = {0.25f - 0.5f * a, 0.25f, a, 0.25f, 0.25f - 0.5f * a};
    const float *p = tmp + c * tmp_cstride + (size_t)i * ws;
    float s = 0.0f;
    if (j % 2 == 0) {
        const int c0 = max((j - 2) / 2, 0), c2 = min((j + 2) / 2, ws - 1);
        s += 2.0f * k[0] * p[c0];
        s += 2.0f * k[2] * p[min(j / 2, ws - 1)];
        s += 2.0f * k[4] * p[c2];
    } else {
        const int c1 = max((j - 1) / 2, 0), c3 = min((j + 1) / 2, ws - 1);
        s += 2.0f * k[1] * p[c1];
        s += 2.0f * k[3] * p[c3];
    }
    out[c * out_cstride + (size_t)i * Ws + j] = s;
}

__global__ void k_vis_to_u8(size_t N, const float *__restrict__ vis, uint8_t *__restrict__ out)
{
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o < N) out[o] = vis[o] != 0.0f ? 1 : 0;
}

// U <- filled - U (optimization.cu:629-646)
__global__ void k_pp_finish(size_t N, const float *__restrict__ filled, float *__restrict__ U)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < 3 * N) U[k] = filled[k] - U[k];
}

// ---- fused primal-dual ROF step --------------------------------------------
// One inner iteration of vtv_sr_ista_step_unstructured (optimization.cu:1817-1828):
//   dual:   p_c <- p_c + sigma grad+ Uq_c           (vtv_rof.cu:118-147)
//   proj:   joint TV_F projection onto |p| <= g      (vtv_rof.cu:280-321; radius 1 without g, :237-275)
//   primal: Uq_c <- (tau F_c + lam (Uq_c + tau div- p_c)) / (tau + lam)   (cuda_kernels.cu:504-532)
// computed per tile: p_new on the tile plus its left/top halo, then Uq.
// Inputs and outputs are separate buffers (neighbouring tiles read halos).
constexpr int PD_TW = 64, PD_TH = 4;

// With u_final (the last inner step of an outer step) the over-relaxation of
// the outer step is fused in (vtv_rof.cu:28-45): u_final <- the new primal,
// uout <- u_final_old + theta (new primal - u_final_old).
// A non-null `halt` that holds nonzero makes the launch a no-op (a step
// enqueued ahead of a stopping rule that has fired).
__global__ void __launch_bounds__(256) k_pd_step(int W, int H, float tau, float lam, float sigma,
                                                 const float *__restrict__ uin, float *__restrict__ uout,
                                                 const float *__restrict__ pxin, const float *__restrict__ pyin,
                                                 float *__restrict__ pxout, float *__restrict__ pyout,
                                                 const float *__restrict__ F, const float *__restrict__ g,
                                                 float *__restrict__ u_final, float theta, const int *halt)
{
    if (halt && *halt) return;
    constexpr int UW = PD_TW + 2, UH = PD_TH + 2;  // Uq region [x0-1, x0+TW] x [y0-1, y0+TH]
    constexpr int PW = PD_TW + 1, PH = PD_TH + 1;  // p region  [x0-1, x0+TW-1] x [y0-1, y0+TH-1]
    __shared__ float us[3][UH][UW];
    __shared__ float ps[6][PH][PW];
    const size_t N = (size_t)W * H;
    const int x0 = blockIdx.x * PD_TW, y0 = blockIdx.y * PD_TH;
    for (int i = threadIdx.x; i < UW * UH; i += blockDim.x) {
        const int lx = i % UW, ly = i / UW;
        const int x = x0 - 1 + lx, y = y0 - 1 + ly;
        const bool in = x >= 0 && x < W && y >= 0 && y < H;
        const size_t o = in ? (size_t)y * W + x : 0;
        for (int c = 0; c < 3; c++) us[c][ly][lx] = in ? uin[c * N + o] : 0.0f;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < PW * PH; i += blockDim.x) {
        const int lx = i % PW, ly = i / PW;
        const int x = x0 - 1 + lx, y = y0 - 1 + ly;
        float a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
        if (x >= 0 && x < W && y >= 0 && y < H) {
            const size_t o = (size_t)y * W + x;
            float nn = 0.0f;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                // the p region and the Uq region share the origin (x0-1, y0-1)
                const float uc = us[c][ly][lx];
                const float gx = (x < W - 1) ? us[c][ly][lx + 1] - uc : 0.0f;
                const float gy = (y < H - 1) ? us[c][ly + 1][lx] - uc : 0.0f;
                a[c] = pxin[c * N + o] + sigma * gx;
                b[c] = pyin[c * N + o] + sigma * gy;
                nn += a[c] * a[c] + b[c] * b[c];
            }
            nn = sqrtf(nn);
            const float rad = g ? g[o] : 1.0f;
            if (nn > rad) {
                const float s = rad / nn;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    a[c] *= s;
                    b[c] *= s;
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            ps[c][ly][lx] = a[c];
            ps[3 + c][ly][lx] = b[c];
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < PD_TW * PD_TH; i += blockDim.x) {
        const int lx = i % PD_TW, ly = i / PD_TW;
        const int x = x0 + lx, y = y0 + ly;
        if (x >= W || y >= H) continue;
        const size_t o = (size_t)y * W + x;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float pxv = ps[c][ly + 1][lx + 1], pyv = ps[3 + c][ly + 1][lx + 1];
            pxout[c * N + o] = pxv;
            pyout[c * N + o] = pyv;
            float step = pxv + pyv;
            if (x > 0) step -= ps[c][ly + 1][lx];
            if (y > 0) step -= ps[3 + c][ly][lx + 1];
            const float unew = us[c][ly + 1][lx + 1] + tau * step;
            const float v = (tau * F[c * N + o] + lam * unew) / (tau + lam);
            if (u_final) {
                const float uold = u_final[c * N + o];
                u_final[c * N + o] = v;
                uout[c * N + o] = uold + theta * (v - uold);
            } else {
                uout[c * N + o] = v;
            }
        }
    }
}

}  // namespace

lfrv_status run_hole_filling(lfrv_ctx *ctx, int iters, double eps, int *ran)
{
    const size_t N = ctx->N;
    unsigned int nh = 0;
    if (ctx->holes_listed) {
        // listed by the visibility pass; its length arrived behind hole_ev
        LFRV_HIP(ctx, hipEventSynchronize(ctx->hole_ev));
        nh = *ctx->hole_count_host;
    } else {
        unsigned int *cnt = ctx->counters + 10;
        LFRV_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(unsigned int), ctx->stream));
        k_hole_list<<<blocks_for(N), kBlock, 0, ctx->stream>>>(N, ctx->visibility, ctx->hole_list, cnt);
        LFRV_HIP(ctx, hipMemcpyAsync(&nh, cnt, sizeof nh, hipMemcpyDeviceToHost, ctx->stream));
        LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    ctx->n_holes = (int)nh;
    if (ran) *ran = 0;
    if (nh == 0 || iters <= 0) return LFRV_OK;
    HoleState *dst = reinterpret_cast<HoleState *>(ctx->scal + 16);
    float *bufs[2] = {ctx->U, ctx->tmp3N};
    LFRV_HIP(ctx, hipMemcpyAsync(ctx->tmp3N, ctx->U, 3 * N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    unsigned int nb = blocks_for(nh);
    if (nb > 1024) nb = 1024;
    // partials of odd / even sweeps, the energy history and the stop flag
    HoleLoop hl;
    hl.maxit = iters;
    hl.eps = eps;
    hl.G = (int)nb;
    double *psw[2] = {ctx->partials, ctx->partials + 1024};
    hl.eh = ctx->partials + 2048;  // iters + 1 <= partials_cap - 2048, checked below
    hl.stopped = reinterpret_cast<int *>(ctx->counters + 12);
    if ((size_t)iters + 1 > ctx->partials_cap - 2048) return fail(ctx, LFRV_E_ARG, "hole filling: too many sweeps");
    k_hole_init<<<1, 64, 0, ctx->stream>>>(hl);
    int launched = 0;
    HoleState host{};
    while (launched < iters) {
        const int chunk = std::min(16, iters - launched);
        for (int j = 0; j < chunk; j++) {
            const int k = launched + j + 1;  // sweep k reads bufs[(k - 1) & 1]
            k_hole_sweep<<<nb, 256, 0, ctx->stream>>>(ctx->W, ctx->H, (int)nh, ctx->hole_list, bufs[(k - 1) & 1],
                                                      bufs[k & 1], hl, k, psw[(k - 1) & 1], psw[k & 1]);
        }
        launched += chunk;
        k_hole_state<<<1, kFinalizeThreads, 0, ctx->stream>>>(hl, launched, psw[launched & 1], dst);
        LFRV_LAUNCH(ctx, "hole filling");
        LFRV_HIP(ctx, hipMemcpyAsync(&host, dst, sizeof host, hipMemcpyDeviceToHost, ctx->stream));
        LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (host.done) break;
    }
    // the result of sweep k (1-based) is in bufs[k & 1]
    if (host.iters & 1)
        LFRV_HIP(ctx, hipMemcpyAsync(ctx->U, ctx->tmp3N, 3 * N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    if (ran) *ran = host.iters;
    return LFRV_OK;
}

lfrv_status run_gauss_smooth(lfrv_ctx *ctx, const float *in, float *out, float *tmp)
{
    // gsl_kernel_gauss_1xn(11, 2.0) (gsl_matrix_convolutions.cpp:42-63)
    Taps t{};
    t.r = 5;
    double s = 0.0, k[11];
    for (int x = -5; x <= 5; x++) {
        k[x + 5] = std::exp(-0.5 * x * x / 4.0);
        s += k[x + 5];
    }
    for (int i = 0; i < 11; i++) t.k[i] = (float)(k[i] / s);
    dim3 bs(256), gs((ctx->W + 255) / 256, ctx->H);
    k_conv_rows<<<gs, bs, 0, ctx->stream>>>(ctx->W, ctx->H, t, in, tmp);
    k_conv_cols<<<gs, bs, 0, ctx->stream>>>(ctx->W, ctx->H, t, tmp, out);
    LFRV_LAUNCH(ctx, "gaussian");
    return LFRV_OK;
}

lfrv_status run_push_pull(lfrv_ctx *ctx, int levels)
{
    const int W = ctx->W, H = ctx->H;
    const size_t N = ctx->N;
    // level offsets inside the pyramid scratch: level 0 is a copy of U
    std::vector<size_t> off(levels + 1, 0);
    std::vector<int> ws(levels + 1, 0), hs(levels + 1, 0);
    size_t tot = 0;
    int top = 0;
    ws[0] = W;
    hs[0] = H;
    off[0] = 0;
    tot = N;
    for (int s = 1; s <= levels; s++) {
        const int wsn = W >> s, hsn = H >> s;
        if (wsn <= 0 || hsn <= 0) break;
        ws[s] = wsn;
        hs[s] = hsn;
        off[s] = tot;
        tot += (size_t)wsn * hsn;
        top = s;
    }
    // scratch: pyramid images (3 * tot floats) + row temp (3 * N floats) + vis pyramid (tot bytes)
    const size_t need = 3 * tot + 3 * N + (tot + 3) / 4 + 1;
    if (need > ctx->pp_cap) return fail(ctx, LFRV_E_ARG, "push-pull scratch too small");
    float *img = ctx->pp_scratch, *tmp = img + 3 * tot;
    uint8_t *vis = reinterpret_cast<uint8_t *>(tmp + 3 * N);
    for (int c = 0; c < 3; c++)
        LFRV_HIP(ctx, hipMemcpyAsync(img + c * tot, ctx->U + c * N, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    k_vis_to_u8<<<blocks_for(N), kBlock, 0, ctx->stream>>>(N, ctx->visibility, vis);
    for (int s = 1; s <= top; s++) {
        const int Ws = ws[s - 1], Hs = hs[s - 1], wsn = ws[s], hsn = hs[s];
        dim3 g1((Ws + 255) / 256, hsn, 3), g2((wsn + 255) / 256, hsn, 3), g3((wsn + 255) / 256, hsn, 1);
        k_pp_reduce_rows<<<g1, 256, 0, ctx->stream>>>(Ws, Hs, hsn, img + off[s - 1], tmp, tot, N);
        k_pp_reduce_cols<<<g2, 256, 0, ctx->stream>>>(Ws, hsn, wsn, tmp, img + off[s], N, tot);
        k_pp_reduce_vis<<<g3, 256, 0, ctx->stream>>>(Ws, Hs, wsn, hsn, vis + off[s - 1], vis + off[s]);
    }
    for (int s = top; s > 0; s--) {
        const int Ws = ws[s - 1], Hs = hs[s - 1], wsn = ws[s], hsn = hs[s];
        dim3 g1((wsn + 255) / 256, Hs, 3), g2((Ws + 255) / 256, Hs, 3);
        k_pp_expand_rows<<<g1, 256, 0, ctx->stream>>>(wsn, hsn, Hs, img + off[s], tmp, tot, N);
        k_pp_expand_cols<<<g2, 256, 0, ctx->stream>>>(Ws, Hs, wsn, tmp, img + off[s - 1], vis + off[s - 1], N, tot);
    }
    // gather the three level-0 planes contiguously, then U <- filled - U
    for (int c = 0; c < 3; c++)
        LFRV_HIP(ctx, hipMemcpyAsync(tmp + c * N, img + c * tot, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    k_pp_finish<<<blocks_for(3 * N), kBlock, 0, ctx->stream>>>(N, tmp, ctx->U);
    LFRV_LAUNCH(ctx, "push-pull");
    return LFRV_OK;
}

lfrv_status run_pd_steps(lfrv_ctx *ctx, int iters, const float *F, bool overrelax, float theta, const int *halt)
{
    const int W = ctx->W, H = ctx->H;
    const float tau = (float)(0.3 / std::sqrt(8.0)), sigma = (float)(0.3 / std::sqrt(8.0));
    const float L = (float)(1.0 / (double)ctx->p.lambda);
    const float lam = (float)(1.0 / (double)L);
    dim3 gs((W + PD_TW - 1) / PD_TW, (H + PD_TH - 1) / PD_TH);
    const float *g = ctx->has_reg_g ? ctx->reg_g : nullptr;
    for (int k = 0; k < iters; k++) {
        float *uin = ctx->Uq, *uout = ctx->U2;
        float *pxin = ctx->PX, *pyin = ctx->PY, *pxout = ctx->P2X, *pyout = ctx->P2Y;
        float *ufin = (overrelax && k == iters - 1) ? ctx->U : nullptr;
        const KernelEvents ev = kernel_events(ctx, 2);  // stamped by the dispatch itself
        hipExtLaunchKernelGGL(k_pd_step, gs, dim3(256), 0, ctx->stream, ev.a, ev.b, 0, W, H, tau, lam, sigma,
                              (const float *)uin, uout, (const float *)pxin, (const float *)pyin, pxout, pyout, F, g,
                              ufin, theta, halt);
        record_kernel_events(ctx, 2, ev);
        std::swap(ctx->Uq, ctx->U2);
        std::swap(ctx->PX, ctx->P2X);
        std::swap(ctx->PY, ctx->P2Y);
    }
    LFRV_LAUNCH(ctx, "pd step");
    return LFRV_OK;
}

}  // namespace lfrv
