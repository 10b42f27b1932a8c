// Jacobi-preconditioned CG of the gradient-constraint normal equations
//   A u = (alpha sum_i S_i^T W_i S_i + c0 I + beta D^T D) u = b
// per RGB channel, FP32 vectors, FP64 reductions, scalars on the device.
//
// One iteration is two streaming launches and two one-block launches:
//   k_cg_op     d = z + beta d_prev (z = Minv r) over the tile and its halo,
//               q = A d with the binned gather/splat in LDS, per-block d.q;
//               a one-block kernel sums them and computes alpha.
//   k_cg_update x += alpha d, r -= alpha q, per-block r.r and r.z; a one-block
//               kernel sums them, tests convergence and computes beta.
// d is double-buffered because neighbouring tiles read it in their halos.
#include "tile_common.hpp"

#include <hip/hip_ext.h>

#include <chrono>

namespace lfrv {

namespace {

// q = A * din (mode 0) or, with the fused direction update (mode 1),
// d = Minv r + beta d_prev, q = A d, written to dout / q, d.q reduced.
// LDS: d tile with halo ks-1 (3 planes), then the pull's meta and slots.
template <int KS>
__global__ void __launch_bounds__(kPullMaxThreads) k_cg_op(TileGeom g, TileEntries te, int cap, float wscale, float c0,
                                               float beta_reg, int mode, const float *__restrict__ din,
                                               const float *__restrict__ r, const float *__restrict__ minv,
                                               float *__restrict__ dout, float *__restrict__ q,
                                               const CgState *st, double *partials)
{
    constexpr int HAL = KS - 1;
    extern __shared__ int smem[];
    __shared__ double red[3 * 16];
    if (mode == 1 && st->done) return;
    const int t = xcd_tile(blockIdx.x, g.ntiles);
    const int tx0 = (t % g.ntx) * g.tw, ty0 = (t / g.ntx) * g.th;
    const int lw = g.tw + 2 * HAL, lh = g.th + 2 * HAL, lp = lw * lh;
    float *dl = reinterpret_cast<float *>(smem);  // 3 planes of lw x lh
    int *plds = smem + 3 * lp;
    const size_t N = (size_t)g.W * g.H;

    float bc[3] = {0.0f, 0.0f, 0.0f};
    int first = 1;
    if (mode == 1) {
        first = st->first;
        for (int c = 0; c < 3; c++) bc[c] = (float)st->beta[c];
    }
    // stage d (tile + halo); outside the image d = 0
    for (int i = threadIdx.x; i < lp; i += blockDim.x) {
        const int lx = i % lw, ly = i / lw;
        const int x = tx0 - HAL + lx, y = ty0 - HAL + ly;
        const bool in = (x >= 0 && x < g.W && y >= 0 && y < g.H);
        const size_t o = in ? (size_t)y * g.W + x : 0;
        float v[3] = {0.0f, 0.0f, 0.0f};
        if (in) {
            if (mode == 0) {
                for (int c = 0; c < 3; c++) v[c] = din[c * N + o];
            } else {
                const float mi = minv[o];
                for (int c = 0; c < 3; c++) {
                    const float z = mi * r[c * N + o];
                    v[c] = first ? z : z + bc[c] * din[c * N + o];
                }
                const bool interior = lx >= HAL && lx < HAL + g.tw && ly >= HAL && ly < HAL + g.th;
                if (interior)
                    for (int c = 0; c < 3; c++) dout[c * N + o] = v[c];
            }
        }
        for (int c = 0; c < 3; c++) dl[c * lp + i] = v[c];
    }
    // (the pull synchronises before it reads the staged tile)
    TileEntries tea = te;
    tea.active = te.active && wscale != 0.0f;
    double dq[3] = {0.0, 0.0, 0.0};
    any_pull<KS, 3, 0>(
        g, tea, t, plds, cap,
        [&](int, int x0, int y0, const float(&wx)[KS], const float(&wy)[KS], float om_e, float(&v)[3]) {
            const int bx = x0 - tx0 + HAL, by = y0 - ty0 + HAL;
            float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
#pragma unroll
            for (int i = 0; i < KS; i++)
#pragma unroll
                for (int j = 0; j < KS; j++) {
                    const float a = wy[i] * wx[j];
                    const int li = (by + i) * lw + bx + j;
                    s0 += a * dl[li];
                    s1 += a * dl[lp + li];
                    s2 += a * dl[2 * lp + li];
                }
            const float om = om_e * wscale;
            v[0] = om * s0;
            v[1] = om * s1;
            v[2] = om * s2;
        },
        [&](size_t o, int x, int y, const float(&a)[3]) {
            const int li = (y - ty0 + HAL) * lw + (x - tx0) + HAL;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float *dp = dl + c * lp;
                const float dv = dp[li];
                float lap = 0.0f;
                if (g.W > 1) {
                    if (x == 0) lap += dv - dp[li + 1];
                    else if (x < g.W - 1) lap += 2.0f * dv - dp[li - 1] - dp[li + 1];
                    else lap += dv - dp[li - 1];
                }
                if (g.H > 1) {
                    if (y == 0) lap += dv - dp[li + lw];
                    else if (y < g.H - 1) lap += 2.0f * dv - dp[li - lw] - dp[li + lw];
                    else lap += dv - dp[li - lw];
                }
                const float qv = a[c] + c0 * dv + beta_reg * lap;
                q[c * N + o] = qv;
                dq[c] += (double)dv * (double)qv;
            }
        });
    if (mode == 0) return;
    block_partials<3>(dq, partials, red);
}

// alpha = r.z / d.q from the operator's partial sums (one block).
__global__ void __launch_bounds__(kFinalizeThreads) k_cg_alpha(CgState *st, const double *__restrict__ partials, int G)
{
    __shared__ double red[3 * 16];
    if (st->done) return;
    double dq[3];
    reduce_partials<3>(partials, G, dq, red);
    if (threadIdx.x == 0) {
        for (int c = 0; c < 3; c++) {
            st->dq[c] = dq[c];
            if (st->active[c]) {
                if (!(dq[c] > 0.0)) {
                    st->breakdown[c] = 1;
                    st->active[c] = 0;
                    st->alpha[c] = 0.0;
                } else {
                    st->alpha[c] = st->rz[c] / dq[c];
                }
            } else {
                st->alpha[c] = 0.0;
            }
        }
    }
}

// ---- dsf = 1: the data operator assembled into a 9-point stencil -----------
//
// With 2 x 2 footprints, sum_i S_i^T W_i S_i couples a pixel with its eight
// neighbours only.  Its raw coefficients (without alpha sigma^2) are the Gram
// sums sum_e omega_e a_e(p) a_e(p') over the entries whose footprint holds
// both p and p'; five planes per pixel by symmetry: D (p, p), E (p, p+(1,0)),
// S (p, p+(0,1)), SE (p, p+(1,1)), SW (p, p+(-1,1)).  They are fixed during a
// solve, so the CG operator becomes a dense stencil pass.

// Assembly: each window cell sums its 4 squared and 6 cross tap products,
// each pixel then adds the cells that hold it (cell (x0, y0) has taps
// TL = (x0, y0), TR, BL, BR).
__global__ void __launch_bounds__(kPullMaxThreads) k_gram2(TileGeom g, TileEntries te, int cap, float *__restrict__ gD,
                                               float *__restrict__ gE, float *__restrict__ gS,
                                               float *__restrict__ gSE, float *__restrict__ gSW)
{
    extern __shared__ int smem[];
    const int t = xcd_tile(blockIdx.x, g.ntiles);
    const int tx0 = (t % g.ntx) * g.tw, ty0 = (t / g.ntx) * g.th;
    const int ww = g.tw + 1, wh = g.th + 1, ncw = ww * wh;
    float acc[10];
    cell_accumulate<2, 10>(g, te, t, smem, cap, acc,
                           [&](int, int, int, const float(&wx)[2], const float(&wy)[2], float om, float(&a)[10]) {
                               const float tl = wy[0] * wx[0], tr = wy[0] * wx[1];
                               const float bl = wy[1] * wx[0], br = wy[1] * wx[1];
                               a[0] += om * tl * tl;
                               a[1] += om * tr * tr;
                               a[2] += om * bl * bl;
                               a[3] += om * br * br;
                               a[4] += om * tl * tr;  // E of TL
                               a[5] += om * bl * br;  // E of BL
                               a[6] += om * tl * bl;  // S of TL
                               a[7] += om * tr * br;  // S of TR
                               a[8] += om * tl * br;  // SE of TL
                               a[9] += om * tr * bl;  // SW of TR
                           });
    float *part = reinterpret_cast<float *>(smem + 2 * wh + 2);
    const int c = threadIdx.x;
    if (c < ncw) {
#pragma unroll
        for (int k = 0; k < 10; k++) part[k * ncw + c] = acc[k];
    }
    __syncthreads();
    for (int p = threadIdx.x; p < g.tw * g.th; p += blockDim.x) {
        const int lx = p % g.tw, ly = p / g.tw;
        const int x = tx0 + lx, y = ty0 + ly;
        if (x >= g.W || y >= g.H) continue;
        // the cell whose top-left tap is (x - dx, y - dy) sits at local (lx + 1 - dx, ly + 1 - dy)
        const int c00 = (ly + 1) * ww + lx + 1, c10 = c00 - 1, c01 = c00 - ww, c11 = c00 - ww - 1;
        const size_t o = (size_t)y * g.W + x;
        gD[o] = part[0 * ncw + c00] + part[1 * ncw + c10] + part[2 * ncw + c01] + part[3 * ncw + c11];
        gE[o] = part[4 * ncw + c00] + part[5 * ncw + c01];
        gS[o] = part[6 * ncw + c00] + part[7 * ncw + c10];
        gSE[o] = part[8 * ncw + c00];
        gSW[o] = part[9 * ncw + c10];
    }
}

constexpr int kSt2W = 64, kSt2H = 8;  // tile of the stencil kernels, one thread per pixel (512 threads)

// ---- dsf = 1: single-reduction PCG on the assembled stencil -----------------
//
// The Chronopoulos-Gear form of preconditioned CG needs one global reduction
// per iteration instead of two: with u = Minv r, w = A u, gamma = (r, u),
// delta = (w, u),
//   beta_i  = gamma_i / gamma_{i-1},  alpha_i = gamma_i / (delta_i - beta_i gamma_i / alpha_{i-1}),
//   p_i = u_i + beta_i p_{i-1},  s_i = w_i + beta_i s_{i-1},
//   x_{i+1} = x_i + alpha_i p_i,  r_{i+1} = r_i - alpha_i s_i.
// (Same iterates as standard PCG in exact arithmetic.)  One kernel does a
// whole iteration: it rebuilds r_{i+1} and u_{i+1} on its tile and a 1-pixel
// halo (s_i and r_{i+1} of the halo are recomputed from w_i, s_{i-1}, r_i),
// applies the stencil to u_{i+1}, writes p_i, x_{i+1}, s_i, r_{i+1}, w_{i+1}
// and reduces gamma, delta and r.r of the new residual.  r, s and w are read
// across tiles, so they are double-buffered.
//
// The solve starts with one pass that forms r_0, u_0, w_0 = A u_0 and their
// sums (k_cgcg_begin_st2).  (A variant that lets every block of the next
// launch finish the reduction itself measured slower: the redundant per-block
// sums cost more than the one-block launch they replace.)
//
// The five coefficient planes of the assembled stencil: centre and the E, S,
// SE, SW couplings of each pixel (the W, N, NW, NE ones are its neighbours').
// w = A u at pixel (x, y) from the LDS tile `up` (row stride LW, centre li):
// the assembled 9-point stencil cf = {D, E, S, SE, SW, E(x-1), S(y-1),
// SE(x-1,y-1), SW(x+1,y-1)} (zero where the neighbour is outside), the
// diagonal shift c0 and the Neumann Laplacian D^T D.
__device__ __forceinline__ float st2_apply(const float *up, int li, int LW, const float (&cf)[9], int x, int y, int W,
                                           int H, float wscale, float c0, float beta_reg)
{
    const float uv = up[li];
    const float sten = cf[0] * uv + cf[1] * up[li + 1] + cf[5] * up[li - 1] + cf[2] * up[li + LW] +
                       cf[6] * up[li - LW] + cf[3] * up[li + LW + 1] + cf[7] * up[li - LW - 1] +
                       cf[4] * up[li + LW - 1] + cf[8] * up[li - LW + 1];
    float lap = 0.0f;
    if (W > 1) {
        if (x == 0) lap += uv - up[li + 1];
        else if (x < W - 1) lap += 2.0f * uv - up[li - 1] - up[li + 1];
        else lap += uv - up[li - 1];
    }
    if (H > 1) {
        if (y == 0) lap += uv - up[li + LW];
        else if (y < H - 1) lap += 2.0f * uv - up[li - LW] - up[li + LW];
        else lap += uv - up[li - LW];
    }
    return wscale * sten + c0 * uv + beta_reg * lap;
}

// Coefficients of the stencil at pixel o = (xc, yc) (clamped into the image);
// every load is issued unconditionally at an in-bounds address and masked
// afterwards, so they all go out before the caller's barrier.
__device__ __forceinline__ void st2_coeffs(const St2Coeffs &g, int W, size_t o, int xc, int yc, float (&cf)[9])
{
    const bool xm = xc > 0, ym = yc > 0, xp = xc < W - 1;
    cf[0] = g.gD[o];
    cf[1] = g.gE[o];
    cf[2] = g.gS[o];
    cf[3] = g.gSE[o];
    cf[4] = g.gSW[o];
    const float e = g.gE[xm ? o - 1 : o];
    const float sm = g.gS[ym ? o - W : o];
    const float se = g.gSE[(xm && ym) ? o - W - 1 : o];
    const float sw = g.gSW[(xp && ym) ? o - W + 1 : o];
    cf[5] = xm ? e : 0.0f;
    cf[6] = ym ? sm : 0.0f;
    cf[7] = (xm && ym) ? se : 0.0f;
    cf[8] = (xp && ym) ? sw : 0.0f;
}

// One thread per tile pixel; the first 2 (LW + TH) threads also own one cell
// of the halo ring, whose r_{i+1} and u_{i+1} are recomputed from r_i, w_i,
// s_{i-1} (never written).  All global loads of the pass (the pixel's
// vectors, its halo cell's, the stencil coefficients) are issued before the
// one barrier.

// The scalar state is moved through LDS by one-block kernels: its words are
// loaded by as many threads at once, together with the partial sums, and
// written back the same way (thread 0 working on the global copy would pay
// the memory latency once per field it reads after a write).
constexpr int kCgStateWords = (int)(sizeof(CgState) / 4);
static_assert(sizeof(CgState) % 4 == 0 && kCgStateWords <= kFinalizeThreads, "CgState moved by words");

// Channel c of the scalar step between two passes (thread c of the finalize;
// the caller sets the done flag from every channel's active flag).  After the
// start pass: alpha_0 = gamma_0 / delta_0.  After an iteration: convergence of
// the new residual (rho), then beta and alpha of the next iteration.
// `pv` = gamma[3], delta[3], rho[3].
__device__ __forceinline__ void cgcg_scalars_channel(CgState &s, const double (&pv)[9], bool after_start, int c)
{
    const double gam = pv[c], del = pv[3 + c];
    if (after_start) {
        s.rz[c] = gam;
        s.dq[c] = del;
        s.beta[c] = 0.0;
        s.alpha[c] = 0.0;
        if (s.active[c]) {
            if (!(del > 0.0)) {
                s.breakdown[c] = 1;
                s.active[c] = 0;
            } else {
                s.alpha[c] = gam / del;
            }
        }
    } else if (s.active[c]) {
        s.iter[c] += 1;
        s.rr[c] = pv[6 + c];
        if (pv[6 + c] <= s.tol2 * s.bb[c] || s.iter[c] >= s.maxit) {
            s.active[c] = 0;
            s.alpha[c] = 0.0;
        } else {
            const double bt = gam / s.rz[c];
            const double den = del - bt * gam / s.alpha[c];
            if (!(den > 0.0)) {
                s.breakdown[c] = 1;
                s.active[c] = 0;
                s.alpha[c] = 0.0;
            } else {
                s.beta[c] = bt;
                s.alpha[c] = gam / den;
                s.rz[c] = gam;
                s.dq[c] = del;
            }
        }
    }
}


// Scalars after iteration i: convergence on r_{i+1}.r_{i+1}, then beta and
// alpha of the next iteration (a solve already done is left as it is).
__device__ __forceinline__ void cgcg_step_body(CgState *st, const double *partials, int G)
{
    __shared__ double red[9 * 16];
    __shared__ CgState sst;
    const int t = threadIdx.x;
    const uint32_t word = t < kCgStateWords ? reinterpret_cast<const uint32_t *>(st)[t] : 0u;
    double v[9];
    reduce_partials<9>(partials, G, v, red);
    if (t < kCgStateWords) reinterpret_cast<uint32_t *>(&sst)[t] = word;
    __syncthreads();
    if (sst.done) return;
    if (t < 3) cgcg_scalars_channel(sst, v, false, t);
    __syncthreads();
    if (t == 0) {
        sst.done = !(sst.active[0] | sst.active[1] | sst.active[2]);
        sst.first = 0;
    }
    __syncthreads();
    if (t < kCgStateWords) reinterpret_cast<uint32_t *>(st)[t] = reinterpret_cast<const uint32_t *>(&sst)[t];
}

// The pass on tile t; its partial sums go to slot `slot` of G per quantity.
__device__ __forceinline__ void cgcg_st2_tile(const CgcgArgs &a, const CgState *st, double *partials, int t, int slot,
                                              int G)
{
    constexpr int LW = kSt2W + 2, LH = kSt2H + 2, LP = LW * LH, NH = 2 * LW + 2 * kSt2H;
    static_assert(NH <= kSt2W * kSt2H, "one thread per halo cell");
    __shared__ float ul[3 * LP];
    __shared__ double red[9 * 16];
    const int tx0 = (t % a.ntx) * kSt2W, ty0 = (t / a.ntx) * kSt2H;
    const int W = a.W, H = a.H;
    const size_t N = (size_t)W * H;
    // every scalar in one batch of loads, then the done test
    float al[3], be[3];
    const int done = st->done;
    const int first = st->first;
    for (int c = 0; c < 3; c++) {
        al[c] = (float)st->alpha[c];
        be[c] = (float)st->beta[c];
    }
    if (done) return;
    // this thread's pixel
    const int lx = threadIdx.x % kSt2W, ly = threadIdx.x / kSt2W;
    const int x = tx0 + lx, y = ty0 + ly;
    const bool pin = x < W && y < H;
    const int xc = min(x, W - 1), yc = min(y, H - 1);
    const size_t o = (size_t)yc * W + xc;
    float cf[9];
    st2_coeffs(a.g, W, o, xc, yc, cf);
    const float mi = a.minv[o];
    float rv[3], wv[3], sv[3], pv[3], xv[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const size_t k = c * N + o;
        rv[c] = a.r_in[k];
        wv[c] = a.w_in[k];
        xv[c] = a.x[k];
        sv[c] = pv[c] = 0.0f;
        if (!first) {
            sv[c] = a.s_in[k];
            pv[c] = a.p[k];
        }
    }
    // this thread's halo cell
    const int h = threadIdx.x;
    const bool hact = h < NH;
    int hlx, hly;
    if (h < LW) {
        hlx = h;
        hly = 0;
    } else if (h < 2 * LW) {
        hlx = h - LW;
        hly = LH - 1;
    } else if (h < 2 * LW + kSt2H) {
        hlx = 0;
        hly = 1 + (h - 2 * LW);
    } else {
        hlx = LW - 1;
        hly = 1 + (h - 2 * LW - kSt2H);
    }
    const int hx = tx0 - 1 + hlx, hy = ty0 - 1 + hly;
    const bool hin = hact && hx >= 0 && hx < W && hy >= 0 && hy < H;
    float hm = 0.0f, hr[3] = {0.0f, 0.0f, 0.0f}, hw[3] = {0.0f, 0.0f, 0.0f}, hs[3] = {0.0f, 0.0f, 0.0f};
    if (hact) {
        const size_t ho = (size_t)min(max(hy, 0), H - 1) * W + min(max(hx, 0), W - 1);
        hm = a.minv[ho];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            hr[c] = a.r_in[c * N + ho];
            hw[c] = a.w_in[c * N + ho];
            if (!first) hs[c] = a.s_in[c * N + ho];
        }
    }
    // vector updates: s_i, p_i, x_{i+1}, r_{i+1} on the pixel; u_{i+1} to LDS
    float rn[3], un[3];
    const int li = (ly + 1) * LW + lx + 1;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        {
            const float s = first ? wv[c] : wv[c] + be[c] * sv[c];
            rn[c] = rv[c] - al[c] * s;
            const float uo = mi * rv[c];
            const float pn = first ? uo : uo + be[c] * pv[c];
            if (pin) {
                const size_t k = c * N + o;
                a.p[k] = pn;
                a.x[k] = xv[c] + al[c] * pn;
                a.s_out[k] = s;
                a.r_out[k] = rn[c];
            }
        }
        un[c] = pin ? mi * rn[c] : 0.0f;
        ul[c * LP + li] = un[c];
    }
    if (hact) {
        const int hl = hly * LW + hlx;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float r1 = hr[c] - al[c] * (first ? hw[c] : hw[c] + be[c] * hs[c]);
            ul[c * LP + hl] = hin ? hm * r1 : 0.0f;
        }
    }
    __syncthreads();
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (pin) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float wn = st2_apply(ul + c * LP, li, LW, cf, x, y, W, H, a.wscale, a.c0, a.beta_reg);
            a.w_out[c * N + o] = wn;
            v[c] += (double)rn[c] * (double)un[c];
            v[3 + c] += (double)wn * (double)un[c];
            v[6 + c] += (double)rn[c] * (double)rn[c];
        }
    }
    block_partials_at<9>(v, partials, red, slot, G);
}

__global__ void __launch_bounds__(kSt2W * kSt2H) k_cgcg_st2(CgcgArgs a, const CgState *st, double *partials)
{
    cgcg_st2_tile(a, st, partials, xcd_tile(blockIdx.x, a.ntiles), blockIdx.x, gridDim.x);
}

// Many solves in one launch: block b is block work[b].y of solve work[b].x
// in that solve's own launch (the same tile and partial-sum slot), with its
// arguments of this iteration's parity and its scalar state.
__global__ void __launch_bounds__(kSt2W * kSt2H) k_cgcg_st2_batch(const CgBatchItem *__restrict__ items,
                                                                   const int2 *__restrict__ work, int parity)
{
    const int2 wk = work[blockIdx.x];
    const CgBatchItem &it = items[wk.x];
    const int G = it.a[parity].ntiles;
    cgcg_st2_tile(it.a[parity], it.st, it.partials, xcd_tile(wk.y, G), wk.y, G);
}

// The scalar step of every solve of the batch: block q sums solve q's tile
// partials in tile order and updates its state (k_cgcg_step per solve).
__global__ void __launch_bounds__(kFinalizeThreads) k_cgcg_step_batch(const CgBatchItem *__restrict__ items)
{
    const CgBatchItem &it = items[blockIdx.x];
    cgcg_step_body(it.st, it.partials, it.a[0].ntiles);
}

// The true residual of a stencil-path solve: partials ||b - A x||^2 per
// channel in FP64 (same halo-ring layout as k_cgcg_st2; nothing written).
__global__ void __launch_bounds__(kSt2W * kSt2H) k_st2_residual(int W, int H, int ntx, int ntiles, float wscale,
                                                                float c0, float beta_reg, St2Coeffs g,
                                                                const float *__restrict__ x,
                                                                const float *__restrict__ b, double *partials)
{
    constexpr int LW = kSt2W + 2, LH = kSt2H + 2, LP = LW * LH, NH = 2 * LW + 2 * kSt2H;
    constexpr int NQ = 3;
    __shared__ float xl[3 * LP];
    __shared__ double red[NQ * 16];
    const int t = xcd_tile(blockIdx.x, ntiles);
    const int tx0 = (t % ntx) * kSt2W, ty0 = (t / ntx) * kSt2H;
    const size_t N = (size_t)W * H;
    const int lx = threadIdx.x % kSt2W, ly = threadIdx.x / kSt2W;
    const int px = tx0 + lx, py = ty0 + ly;
    const bool pin = px < W && py < H;
    const int xc = min(px, W - 1), yc = min(py, H - 1);
    const size_t o = (size_t)yc * W + xc;
    float cf[9];
    st2_coeffs(g, W, o, xc, yc, cf);
    float xv[3], bv[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        xv[c] = x[c * N + o];
        bv[c] = b[c * N + o];
    }
    const int h = threadIdx.x;
    const bool hact = h < NH;
    int hlx, hly;
    if (h < LW) {
        hlx = h;
        hly = 0;
    } else if (h < 2 * LW) {
        hlx = h - LW;
        hly = LH - 1;
    } else if (h < 2 * LW + kSt2H) {
        hlx = 0;
        hly = 1 + (h - 2 * LW);
    } else {
        hlx = LW - 1;
        hly = 1 + (h - 2 * LW - kSt2H);
    }
    const int hx = tx0 - 1 + hlx, hy = ty0 - 1 + hly;
    const bool hin = hact && hx >= 0 && hx < W && hy >= 0 && hy < H;
    float hxv[3] = {0.0f, 0.0f, 0.0f};
    if (hact) {
        const size_t ho = (size_t)min(max(hy, 0), H - 1) * W + min(max(hx, 0), W - 1);
#pragma unroll
        for (int c = 0; c < 3; c++) hxv[c] = x[c * N + ho];
    }
    const int li = (ly + 1) * LW + lx + 1;
#pragma unroll
    for (int c = 0; c < 3; c++) xl[c * LP + li] = pin ? xv[c] : 0.0f;
    if (hact) {
        const int hl = hly * LW + hlx;
#pragma unroll
        for (int c = 0; c < 3; c++) xl[c * LP + hl] = hin ? hxv[c] : 0.0f;
    }
    __syncthreads();
    double v[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) v[q] = 0.0;
    if (pin) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float q = st2_apply(xl + c * LP, li, LW, cf, px, py, W, H, wscale, c0, beta_reg);
            const double d = (double)bv[c] - (double)q;
            v[c] += d * d;
        }
    }
    block_partials<NQ>(v, partials, red);
}


// The start of a stencil-path solve in one pass: r_0 = b - A x_0 and
// u_0 = Minv r_0 on the tile and a 1-pixel ring (x_0 staged with a 2-pixel
// ring), w_0 = A u_0 on the tile; writes r_0 and w_0 and the partial sums
// r.r, r.(Minv r), b.b and w.u per channel.  Every global load is issued
// before the first barrier.
__global__ void __launch_bounds__(kSt2W * kSt2H) k_cgcg_begin_st2(int W, int H, int ntx, int ntiles, float wscale,
                                                                  float c0, float beta_reg, St2Coeffs g,
                                                                  const float *__restrict__ x,
                                                                  const float *__restrict__ b,
                                                                  const float *__restrict__ minv,
                                                                  float *__restrict__ r_out,
                                                                  float *__restrict__ w_out, double *partials)
{
    constexpr int NT = kSt2W * kSt2H;
    constexpr int XW = kSt2W + 4, XH = kSt2H + 4, XP = XW * XH;  // x from (tx0 - 2, ty0 - 2)
    constexpr int UW = kSt2W + 2, UH = kSt2H + 2, UP = UW * UH;  // r, u from (tx0 - 1, ty0 - 1)
    constexpr int NXT = (XP + NT - 1) / NT, NUT = (UP + NT - 1) / NT;
    __shared__ float xl[3 * XP];
    __shared__ float ul[3 * UP];
    __shared__ double red[12 * 16];
    const int t = xcd_tile(blockIdx.x, ntiles);
    const int tx0 = (t % ntx) * kSt2W, ty0 = (t / ntx) * kSt2H;
    const size_t N = (size_t)W * H;
    const int tid = threadIdx.x;
    // ---- global loads ----
    float xv[NXT][3];
#pragma unroll
    for (int k = 0; k < NXT; k++) {
        const int i = tid + k * NT;
        const int px = tx0 - 2 + i % XW, py = ty0 - 2 + i / XW;
        const bool in = i < XP && px >= 0 && px < W && py >= 0 && py < H;
        const size_t o = in ? (size_t)py * W + px : 0;
#pragma unroll
        for (int c = 0; c < 3; c++) xv[k][c] = in ? x[c * N + o] : 0.0f;
    }
    float rcf[NUT][9], rb[NUT][3], rmi[NUT];
    bool rin[NUT];
#pragma unroll
    for (int k = 0; k < NUT; k++) {
        const int i = tid + k * NT;
        const int px = tx0 - 1 + i % UW, py = ty0 - 1 + i / UW;
        rin[k] = i < UP && px >= 0 && px < W && py >= 0 && py < H;
        const int xc = min(max(px, 0), W - 1), yc = min(max(py, 0), H - 1);
        const size_t o = (size_t)yc * W + xc;
        st2_coeffs(g, W, o, xc, yc, rcf[k]);
        rmi[k] = minv[o];
#pragma unroll
        for (int c = 0; c < 3; c++) rb[k][c] = b[c * N + o];
    }
    const int lx = tid % kSt2W, ly = tid / kSt2W;
    const int px = tx0 + lx, py = ty0 + ly;
    const bool pin = px < W && py < H;
    const int pxc = min(px, W - 1), pyc = min(py, H - 1);
    float cf[9];
    st2_coeffs(g, W, (size_t)pyc * W + pxc, pxc, pyc, cf);
    // ---- x, then r and u on the ring ----
#pragma unroll
    for (int k = 0; k < NXT; k++) {
        const int i = tid + k * NT;
        if (i < XP)
#pragma unroll
            for (int c = 0; c < 3; c++) xl[c * XP + i] = xv[k][c];
    }
    __syncthreads();
    double v[12];
#pragma unroll
    for (int q = 0; q < 12; q++) v[q] = 0.0;
#pragma unroll
    for (int k = 0; k < NUT; k++) {
        const int i = tid + k * NT;
        if (i >= UP) continue;
        const int ux = i % UW, uy = i / UW;
        const int gx = tx0 - 1 + ux, gy = ty0 - 1 + uy;
        const int xi = (uy + 1) * XW + ux + 1;  // the point in the x tile
        const bool interior = ux >= 1 && ux <= kSt2W && uy >= 1 && uy <= kSt2H;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float u = 0.0f;
            if (rin[k]) {
                const float q = st2_apply(xl + c * XP, xi, XW, rcf[k], gx, gy, W, H, wscale, c0, beta_reg);
                const float rv = rb[k][c] - q;
                u = rmi[k] * rv;
                if (interior) {
                    r_out[c * N + (size_t)gy * W + gx] = rv;
                    v[c] += (double)rv * rv;
                    v[3 + c] += (double)rv * (double)u;
                    v[6 + c] += (double)rb[k][c] * rb[k][c];
                }
            }
            ul[c * UP + i] = u;
        }
    }
    __syncthreads();
    // ---- w = A u on the tile ----
    if (pin) {
        const int li = (ly + 1) * UW + lx + 1;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float wv = st2_apply(ul + c * UP, li, UW, cf, px, py, W, H, wscale, c0, beta_reg);
            w_out[c * N + (size_t)py * W + px] = wv;
            v[9 + c] += (double)wv * (double)ul[c * UP + li];
        }
    }
    block_partials<12>(v, partials, red);
}

// Scalars of a new solve from the begin pass: the state of k_cg_init_state,
// then alpha_0 = gamma_0 / delta_0 as after the start pass.
__global__ void __launch_bounds__(kFinalizeThreads) k_cgcg_begin(CgState *st, const double *__restrict__ partials,
                                                                 int G, double tol2, int maxit)
{
    __shared__ double red[12 * 16];
    __shared__ CgState sst;
    const int t = threadIdx.x;
    // the fields this kernel does not set (resid) keep their values
    const uint32_t word = t < kCgStateWords ? reinterpret_cast<const uint32_t *>(st)[t] : 0u;
    double v[12];
    reduce_partials<12>(partials, G, v, red);
    if (t < kCgStateWords) reinterpret_cast<uint32_t *>(&sst)[t] = word;
    __syncthreads();
    // thread c sets channel c (every thread holds the sums)
    bool any = false;
#pragma unroll
    for (int c = 0; c < 3; c++) any |= (v[c] > tol2 * v[6 + c]) && maxit > 0;
    if (t < 3) {
        CgState &s = sst;
        const int c = t;
        s.rr[c] = v[c];
        s.rz[c] = v[3 + c];
        s.bb[c] = v[6 + c];
        s.alpha[c] = 0.0;
        s.beta[c] = 0.0;
        s.iter[c] = 0;
        s.breakdown[c] = 0;
        s.active[c] = (v[c] > tol2 * v[6 + c]) && maxit > 0;
        if (any) {
            const double pv[9] = {v[3], v[4], v[5], v[9], v[10], v[11], v[0], v[1], v[2]};
            cgcg_scalars_channel(s, pv, true, c);
        }
    }
    if (t == 0) {
        sst.tol2 = tol2;
        sst.maxit = maxit;
        sst.first = 1;
    }
    __syncthreads();
    if (t == 0) sst.done = !(sst.active[0] | sst.active[1] | sst.active[2]);
    __syncthreads();
    if (t < kCgStateWords) reinterpret_cast<uint32_t *>(st)[t] = reinterpret_cast<const uint32_t *>(&sst)[t];
}

// The scalar step of one solve between two passes.
__global__ void __launch_bounds__(kFinalizeThreads) k_cgcg_step(CgState *st, const double *__restrict__ partials,
                                                                int G)
{
    cgcg_step_body(st, partials, G);
}

// r = b - q (q = A x0), r.r, r.z, b.b; initialises the scalar state.
__global__ void __launch_bounds__(256) k_cg_init(size_t N, const float *__restrict__ b,
                                                 const float *__restrict__ q, float *__restrict__ r,
                                                 const float *__restrict__ minv, double *partials)
{
    __shared__ double red[9 * 16];
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < N; o += (size_t)gridDim.x * blockDim.x) {
        const float mi = minv[o];
        for (int c = 0; c < 3; c++) {
            const float bv = b[c * N + o];
            const float rv = bv - q[c * N + o];
            r[c * N + o] = rv;
            v[c] += (double)rv * rv;
            v[3 + c] += (double)rv * (double)(mi * rv);
            v[6 + c] += (double)bv * bv;
        }
    }
    block_partials<9>(v, partials, red);
}

// Scalar state of a new solve from the partial sums of k_cg_init (one block).
__global__ void __launch_bounds__(kFinalizeThreads) k_cg_init_state(CgState *st, const double *__restrict__ partials, int G,
                                                       double tol2, int maxit)
{
    __shared__ double red[9 * 16];
    double v[9];
    reduce_partials<9>(partials, G, v, red);
    if (threadIdx.x == 0) {
        int any = 0;
        for (int c = 0; c < 3; c++) {
            st->rr[c] = v[c];
            st->rz[c] = v[3 + c];
            st->bb[c] = v[6 + c];
            st->alpha[c] = 0.0;
            st->beta[c] = 0.0;
            st->iter[c] = 0;
            st->breakdown[c] = 0;
            st->active[c] = (v[c] > tol2 * v[6 + c]) && maxit > 0;
            any |= st->active[c];
        }
        st->tol2 = tol2;
        st->maxit = maxit;
        st->first = 1;
        st->done = !any;
    }
}

// x += alpha d; r -= alpha q; reduce r.r and r.(Minv r); convergence + beta.
__global__ void __launch_bounds__(256) k_cg_update(size_t N, float *__restrict__ x,
                                                   const float *__restrict__ d, const float *__restrict__ q,
                                                   float *__restrict__ r, const float *__restrict__ minv,
                                                   const CgState *st, double *partials)
{
    __shared__ double red[6 * 16];
    if (st->done) return;
    const float a[3] = {(float)st->alpha[0], (float)st->alpha[1], (float)st->alpha[2]};
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < N; o += (size_t)gridDim.x * blockDim.x) {
        const float mi = minv[o];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const size_t k = c * N + o;
            x[k] += a[c] * d[k];
            const float rv = r[k] - a[c] * q[k];
            r[k] = rv;
            v[c] += (double)rv * rv;
            v[3 + c] += (double)rv * (double)(mi * rv);
        }
    }
    block_partials<6>(v, partials, red);
}

// Convergence test and beta = r.z_new / r.z_old (one block).
__global__ void __launch_bounds__(kFinalizeThreads) k_cg_beta(CgState *st, const double *__restrict__ partials, int G)
{
    __shared__ double red[6 * 16];
    if (st->done) return;
    double v[6];
    reduce_partials<6>(partials, G, v, red);
    if (threadIdx.x == 0) {
        int any = 0;
        for (int c = 0; c < 3; c++) {
            if (st->active[c]) {
                st->iter[c] += 1;
                st->rr[c] = v[c];
                if (v[c] <= st->tol2 * st->bb[c]) {
                    st->active[c] = 0;
                    st->beta[c] = 0.0;
                } else {
                    st->beta[c] = v[3 + c] / st->rz[c];
                    st->rz[c] = v[3 + c];
                    if (st->iter[c] >= st->maxit) st->active[c] = 0;
                }
            }
            any |= st->active[c];
        }
        st->first = 0;
        st->done = !any;
    }
}

// ||b - q||^2 per channel (true residual check).
__global__ void __launch_bounds__(256) k_resid(size_t N, const float *__restrict__ b,
                                               const float *__restrict__ q, double *partials)
{
    __shared__ double red[3 * 16];
    double v[3] = {0, 0, 0};
    for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < N; o += (size_t)gridDim.x * blockDim.x)
        for (int c = 0; c < 3; c++) {
            const double t = (double)b[c * N + o] - (double)q[c * N + o];
            v[c] += t * t;
        }
    block_partials<3>(v, partials, red);
}

// Energy of the normal equations, data part: sum over binned entries would
// count duplicates, so this walks the source pixels directly.
template <int KS>
__global__ void __launch_bounds__(256) k_gibr_energy_data(int V, int W, int H, int w, int h,
                                                          const float *__restrict__ tx, const float *__restrict__ ty,
                                                          const float *__restrict__ omega, float wscale,
                                                          const float *__restrict__ views,
                                                          const float *__restrict__ U, double *partials)
{
    __shared__ double red[16];
    const size_t n = (size_t)w * h, N = (size_t)W * H;
    double v[1] = {0.0};
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)V * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const float om = omega[k] * wscale;
        const float t0 = tx[k], t1 = ty[k];
        if (om == 0.0f || !cell_valid(t0, t1, W, H)) continue;
        int x0, y0;
        float wx[KS], wy[KS];
        if (!footprint<KS>(t0, t1, W, H, x0, y0, wx, wy)) continue;
        const size_t vi = k / n, m = k - vi * n;
        for (int c = 0; c < 3; c++) {
            double s = 0.0;
            for (int i = 0; i < KS; i++) {
                const int y = y0 + i;
                if (y < 0 || y > H - 1) continue;
                for (int j = 0; j < KS; j++) {
                    const int x = x0 + j;
                    if (x < 0 || x > W - 1) continue;
                    s += (double)(wy[i] * wx[j]) * U[c * N + (size_t)y * W + x];
                }
            }
            const double rr = s - views[vi * 3 * n + c * n + m];
            v[0] += (double)om * rr * rr;
        }
    }
    block_partials<1>(v, partials, red);
}

__global__ void __launch_bounds__(256) k_gibr_energy_grad(int W, int H, const float *__restrict__ U,
                                                          const float *__restrict__ gx, const float *__restrict__ gy,
                                                          double *partials)
{
    __shared__ double red[16];
    const size_t N = (size_t)W * H;
    double v[1] = {0.0};
    for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < N; o += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(o % W), y = (int)(o / W);
        for (int c = 0; c < 3; c++) {
            const float *u = U + c * N;
            if (x < W - 1) { const double t = ((double)u[o + 1] - u[o]) - gx[o]; v[0] += t * t; }
            if (y < H - 1) { const double t = ((double)u[o + W] - u[o]) - gy[o]; v[0] += t * t; }
        }
    }
    block_partials<1>(v, partials, red);
}

unsigned int reduce_blocks(size_t N)
{
    const size_t b = (N + kBlock - 1) / kBlock;
    return (unsigned int)(b < 2048 ? (b == 0 ? 1 : b) : 2048);
}


// ---------------------------------------------------------------------------
// dsf >= 2: the data operator sum_e omega_e S_e^T S_e in assembled form.
//
// With D = dsf, an entry's footprint weights are the 2 x 2 anchor taps t_e of
// its box start followed by a D x D box (box_weights).  So
//   S^T diag(omega) S u = Box^T G Box u,   G = sum_e omega_e t_e t_e^T,
// Box u at an anchor the sum of u over its D x D box, and G couples each
// anchor with its 8 neighbours (5 planes by symmetry: centre, E, S, SE, SW).
// G is summed once per set of weights (k_gram_box, a pull over the entries);
// a PCG operator application is then a tiled stencil (box_apply_tile): the
// box sums of u in LDS (rows, then columns), G at the anchors, the box sums
// back onto the pixels, plus c0 u and the Neumann Laplacian of u.  Anchor
// plane index a holds the cell a + 1 of the splat structure, whose box covers
// pixels a - D + 1 .. a.
// ---------------------------------------------------------------------------
template <int KS>
__global__ void __launch_bounds__(kPullMaxThreads) k_gram_box(TileGeom g, TileEntries te, int cap, size_t NA,
                                                              float *__restrict__ ops)
{
    extern __shared__ int smem[];
    constexpr int NC = 10;
    const int t = xcd_tile(blockIdx.x, g.ntiles);
    const int tx0 = (t % g.ntx) * g.tw, ty0 = (t / g.ntx) * g.th;
    const int ww = g.tw + 1, wh = g.th + 1, ncw = ww * wh;
    float acc[NC];
    cell_accumulate<KS, NC, false, true, 0, 1>(
        g, te, t, smem, cap, acc,
        [&](int, int, int, const float(&ax)[2], const float(&ay)[2], float om, float(&a)[NC]) {
            const float w[4] = {ay[0] * ax[0], ay[0] * ax[1], ay[1] * ax[0], ay[1] * ax[1]};  // tl tr bl br
            a[0] += om * w[0] * w[0];
            a[1] += om * w[1] * w[1];
            a[2] += om * w[2] * w[2];
            a[3] += om * w[3] * w[3];
            a[4] += om * w[0] * w[1];  // tl - tr (E of tl)
            a[5] += om * w[2] * w[3];  // bl - br (E of bl)
            a[6] += om * w[0] * w[2];  // tl - bl (S of tl)
            a[7] += om * w[1] * w[3];  // tr - br (S of tr)
            a[8] += om * w[0] * w[3];  // tl - br (SE of tl)
            a[9] += om * w[1] * w[2];  // tr - bl (SW of tr)
        });
    float *part = reinterpret_cast<float *>(smem + 2 * wh + 2);
    const int c = threadIdx.x;
    if (c < ncw) {
#pragma unroll
        for (int k = 0; k < NC; k++) part[k * ncw + c] = acc[k];
    }
    __syncthreads();
    for (int p = threadIdx.x; p < g.tw * g.th; p += blockDim.x) {
        const int lx = p % g.tw, ly = p / g.tw;
        const int x = tx0 + lx, y = ty0 + ly;
        if (x >= g.Wc || y >= g.Hc) continue;
        // the anchor is the tl tap of its own cell, tr of the left cell, bl of
        // the upper cell and br of the upper-left cell
        const int c00 = (ly + 1) * ww + lx + 1, c10 = c00 - 1, c01 = c00 - ww, c11 = c00 - ww - 1;
        const size_t o = (size_t)y * g.Wc + x;
        ops[o] = part[0 * ncw + c00] + part[1 * ncw + c10] + part[2 * ncw + c01] + part[3 * ncw + c11];
        ops[NA + o] = part[4 * ncw + c00] + part[5 * ncw + c01];
        ops[2 * NA + o] = part[6 * ncw + c00] + part[7 * ncw + c10];
        ops[3 * NA + o] = part[8 * ncw + c00];
        ops[4 * NA + o] = part[9 * ncw + c10];
    }
}

constexpr int kBoxTW = 32, kBoxTH = 16;

template <int KS>
struct BoxDims {
    static constexpr int D = KS - 1, TW = kBoxTW, TH = kBoxTH, NT = TW * TH;
    static constexpr int LW = TW + 2 * D, LH = TH + 2 * D;  // u from (tx0 - D, ty0 - D)
    static constexpr int QW = TW + D - 1, QH = TH + D - 1;  // anchors tx0 .. tx0 + QW - 1
    static constexpr int BW = QW + 2, BH = QH + 2;          // box sums at anchors tx0 - 1 + bx
    static constexpr int NU = (LW * LH + NT - 1) / NT, NQ = (QW * QH + NT - 1) / NT;
};

// Coefficients of anchor (xq, yq): own centre, E, S, SE, SW and the
// neighbours' couplings to it (zero beyond the anchor grid).
__device__ __forceinline__ void box_coeffs(const float *__restrict__ ops, size_t NA, int Wc, int Hc, int xq, int yq,
                                           float (&f)[9])
{
    const bool in = xq >= 0 && yq >= 0 && xq < Wc && yq < Hc;
    const int xc = min(max(xq, 0), Wc - 1), yc = min(max(yq, 0), Hc - 1);
    const size_t o = (size_t)yc * Wc + xc;
    const bool xm = xc > 0, ym = yc > 0, xp = xc < Wc - 1;
    const size_t oE = xm ? o - 1 : o, oS = ym ? o - Wc : o, oSE = (xm && ym) ? o - Wc - 1 : o,
                 oSW = (xp && ym) ? o - Wc + 1 : o;
    f[0] = ops[o];
    f[1] = ops[NA + o];
    f[2] = ops[2 * NA + o];
    f[3] = ops[3 * NA + o];
    f[4] = ops[4 * NA + o];
    const float e = ops[NA + oE], sn = ops[2 * NA + oS], se = ops[3 * NA + oSE], sw = ops[4 * NA + oSW];
    f[5] = xm ? e : 0.0f;
    f[6] = ym ? sn : 0.0f;
    f[7] = (xm && ym) ? se : 0.0f;
    f[8] = (xp && ym) ? sw : 0.0f;
    if (!in)
#pragma unroll
        for (int q = 0; q < 9; q++) f[q] = 0.0f;
}

// A u at this thread's pixel of the tile (tx0, ty0): wscale Box^T G Box u +
// c0 u + beta_reg (Neumann Laplacian of u).  All threads of the block call it.
// uc receives u at the pixel.  The caller provides the LDS.
template <int KS>
__device__ __forceinline__ void box_apply_tile(int W, int H, int Wc, int Hc, int tx0, int ty0, float wscale, float c0,
                                               float beta_reg, const float *__restrict__ ops, size_t NA,
                                               const float *__restrict__ u, float *su, float *sh, float *sb,
                                               float *sq, float (&out)[3], float (&uc)[3])
{
    using S = BoxDims<KS>;
    constexpr int D = S::D, TW = S::TW, NT = S::NT, LW = S::LW, LH = S::LH, QW = S::QW, QH = S::QH, BW = S::BW,
                  BH = S::BH;
    const size_t N = (size_t)W * H;
    const int tid = threadIdx.x;
    // global loads first: u on the tile and its D-ring, the anchors' coefficients
    float uv[S::NU][3];
#pragma unroll
    for (int k = 0; k < S::NU; k++) {
        const int i = tid + k * NT;
        const int x = tx0 - D + i % LW, y = ty0 - D + i / LW;
        const bool in = i < LW * LH && x >= 0 && x < W && y >= 0 && y < H;
        const size_t o = in ? (size_t)y * W + x : 0;
#pragma unroll
        for (int c = 0; c < 3; c++) uv[k][c] = in ? u[c * N + o] : 0.0f;
    }
    float cf[S::NQ][9];
#pragma unroll
    for (int k = 0; k < S::NQ; k++) {
        const int i = tid + k * NT;
        box_coeffs(ops, NA, Wc, Hc, i < QW * QH ? tx0 + i % QW : -1, ty0 + i / QW, cf[k]);
    }
#pragma unroll
    for (int k = 0; k < S::NU; k++) {
        const int i = tid + k * NT;
        if (i < LW * LH)
#pragma unroll
            for (int c = 0; c < 3; c++) su[c * LW * LH + i] = uv[k][c];
    }
    __syncthreads();
    // box sums of u: rows (anchor column bx covers tile columns bx .. bx + D - 1), then columns
    for (int i = tid; i < LH * BW; i += NT) {
        const int ly = i / BW, bx = i - ly * BW;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float sm = 0.0f;
#pragma unroll
            for (int l = 0; l < D; l++) sm += su[c * LW * LH + ly * LW + bx + l];
            sh[c * LH * BW + i] = sm;
        }
    }
    __syncthreads();
    for (int i = tid; i < BH * BW; i += NT) {
        const int by = i / BW, bx = i - by * BW;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float sm = 0.0f;
#pragma unroll
            for (int l = 0; l < D; l++) sm += sh[c * LH * BW + (by + l) * BW + bx];
            sb[c * BH * BW + i] = sm;
        }
    }
    __syncthreads();
    // G at the anchors
#pragma unroll
    for (int k = 0; k < S::NQ; k++) {
        const int i = tid + k * NT;
        if (i >= QW * QH) continue;
        const int qy = i / QW, qx = i - qy * QW;
        const int b = (qy + 1) * BW + qx + 1;
        const float(&f)[9] = cf[k];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float *B = sb + c * BH * BW;
            sq[c * QW * QH + i] = f[0] * B[b] + f[1] * B[b + 1] + f[5] * B[b - 1] + f[2] * B[b + BW] +
                                  f[6] * B[b - BW] + f[3] * B[b + BW + 1] + f[7] * B[b - BW - 1] +
                                  f[4] * B[b + BW - 1] + f[8] * B[b - BW + 1];
        }
    }
    __syncthreads();
    // box sums back: pixel lx takes anchors lx .. lx + D - 1; rows into sh, then columns
    for (int i = tid; i < QH * TW; i += NT) {
        const int qy = i / TW, lx = i - qy * TW;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float sm = 0.0f;
#pragma unroll
            for (int l = 0; l < D; l++) sm += sq[c * QW * QH + qy * QW + lx + l];
            sh[c * QH * TW + i] = sm;
        }
    }
    __syncthreads();
    const int lx = tid % TW, ly = tid / TW;
    const int x = tx0 + lx, y = ty0 + ly;
    const int li = (ly + D) * LW + lx + D;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float acc = 0.0f;
#pragma unroll
        for (int l = 0; l < D; l++) acc += sh[c * QH * TW + (ly + l) * TW + lx];
        const float *dp = su + c * LW * LH;
        const float dv = dp[li];
        float lap = 0.0f;
        if (W > 1) {
            if (x == 0) lap += dv - dp[li + 1];
            else if (x < W - 1) lap += 2.0f * dv - dp[li - 1] - dp[li + 1];
            else lap += dv - dp[li - 1];
        }
        if (H > 1) {
            if (y == 0) lap += dv - dp[li + LW];
            else if (y < H - 1) lap += 2.0f * dv - dp[li - LW] - dp[li + LW];
            else lap += dv - dp[li - LW];
        }
        out[c] = wscale * acc + c0 * dv + beta_reg * lap;
        uc[c] = dv;
    }
    __syncthreads();  // the LDS is free again
}

// LDS floats of box_apply_tile: u, the row / column box sums, the anchor values
template <int KS>
constexpr int box_lds_floats()
{
    using S = BoxDims<KS>;
    return 3 * (S::LW * S::LH + S::LH * S::BW + S::BH * S::BW + S::QW * S::QH);
}

enum : int { kBoxResInit = 0, kBoxBegin = 1, kBoxOp = 2, kBoxResid = 3 };

// One box-operator pass over the tiles.
//   kBoxResInit: r = b - A x, u = Minv r (no sums)
//   kBoxBegin:   w = A u; 12 sums r.r, r.u, b.b, w.u (the begin pass)
//   kBoxOp:      w = A u; 9 sums r.u, w.u, r.r (a PCG iteration)
//   kBoxResid:   3 sums |b - A x|^2 (the true residual)
// `in` is x (ResInit, Resid) or u (Begin, Op).  A solve already done returns.
template <int KS, int MODE>
__global__ void __launch_bounds__(kBoxTW * kBoxTH) k_box_pass(int W, int H, int Wc, int Hc, int ntx, int ntiles,
                                                                float wscale, float c0, float beta_reg,
                                                                const float *__restrict__ ops, size_t NA,
                                                                const float *__restrict__ in,
                                                                const float *__restrict__ b,
                                                                const float *__restrict__ r,
                                                                const float *__restrict__ minv,
                                                                float *__restrict__ out, float *__restrict__ out2,
                                                                const CgState *st, double *partials)
{
    constexpr int NL = box_lds_floats<KS>();
    __shared__ float lds[NL];
    __shared__ double red[12 * 16];
    using S = BoxDims<KS>;
    if (MODE == kBoxOp && st->done) return;
    float *su = lds, *sh = su + 3 * S::LW * S::LH, *sb = sh + 3 * S::LH * S::BW, *sq = sb + 3 * S::BH * S::BW;
    const int t = xcd_tile(blockIdx.x, ntiles);
    const int tx0 = (t % ntx) * S::TW, ty0 = (t / ntx) * S::TH;
    const int x = tx0 + threadIdx.x % S::TW, y = ty0 + threadIdx.x / S::TW;
    const bool pin = x < W && y < H;
    const size_t N = (size_t)W * H, o = pin ? (size_t)y * W + x : 0;
    float av[3], uv[3];
    box_apply_tile<KS>(W, H, Wc, Hc, tx0, ty0, wscale, c0, beta_reg, ops, NA, in, su, sh, sb, sq, av, uv);
    constexpr int NQS = MODE == kBoxBegin ? 12 : MODE == kBoxOp ? 9 : 3;
    double v[NQS];
#pragma unroll
    for (int q = 0; q < NQS; q++) v[q] = 0.0;
    if (pin) {
        const float mi = (MODE == kBoxResInit) ? minv[o] : 0.0f;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const size_t k = c * N + o;
            if (MODE == kBoxResInit) {
                const float rv = b[k] - av[c];
                out[k] = rv;
                out2[k] = mi * rv;
            } else if (MODE == kBoxResid) {
                const double rv = (double)b[k] - (double)av[c];
                v[c] += rv * rv;
            } else {
                out[k] = av[c];
                const double rr = r[k], uu = uv[c], ww = av[c];
                if (MODE == kBoxBegin) {
                    const double bb = b[k];
                    v[c] += rr * rr;
                    v[3 + c] += rr * uu;
                    v[6 + c] += bb * bb;
                    v[9 + c] += ww * uu;
                } else {
                    v[c] += rr * uu;
                    v[3 + c] += ww * uu;
                    v[6 + c] += rr * rr;
                }
            }
        }
    }
    if (MODE != kBoxResInit) block_partials<NQS>(v, partials, red);
}

// The vector updates of a PCG iteration from the scalars of the state:
// s_i = w_i + beta s_{i-1}, p_i = Minv r_i + beta p_{i-1}, x += alpha p_i,
// r_{i+1} = r_i - alpha s_i, u_{i+1} = Minv r_{i+1} (the operand of the next
// box pass), per pixel and channel (the same updates as k_cgcg_st2's).
__global__ void __launch_bounds__(256) k_box_update(size_t N, const CgState *st, const float *__restrict__ minv,
                                                    const float *__restrict__ r_in, const float *__restrict__ w_in,
                                                    const float *__restrict__ s_in, float *__restrict__ r_out,
                                                    float *__restrict__ s_out, float *__restrict__ p,
                                                    float *__restrict__ x, float *__restrict__ u)
{
    const int done = st->done, first = st->first;
    float al[3], be[3];
    for (int c = 0; c < 3; c++) {
        al[c] = (float)st->alpha[c];
        be[c] = (float)st->beta[c];
    }
    if (done) return;
    for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < N; o += (size_t)gridDim.x * blockDim.x) {
        const float mi = minv[o];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const size_t k = c * N + o;
            const float rv = r_in[k];
            const float s = first ? w_in[k] : w_in[k] + be[c] * s_in[k];
            const float uo = mi * rv;
            const float pn = first ? uo : uo + be[c] * p[k];
            const float rn = rv - al[c] * s;
            p[k] = pn;
            x[k] += al[c] * pn;
            s_out[k] = s;
            r_out[k] = rn;
            u[k] = mi * rn;
        }
    }
}

}  // namespace

static lfrv_status cg_finish_stats(lfrv_ctx *ctx, lfrv_cg_stats *st, std::chrono::steady_clock::time_point t_start);

lfrv_status run_cg(lfrv_ctx *ctx, double alpha, double c0, double beta, bool use_views, bool use_precond,
                   bool poisson_rhs, double tol, int maxit, lfrv_cg_stats *st)
{
    auto t_start = std::chrono::steady_clock::now();
    if (!ctx->prepared) return fail(ctx, LFRV_E_STATE, "CG before lfrv_prepare");
    // the sorted entry weights are needed unless the stencil is already assembled
    if (use_views && !(ctx->ks == 2 && ctx->gram && ctx->gram_epoch == ctx->epoch)) {
        lfrv_status s = run_refresh_entry_weights(ctx, 1.0f);
        if (s != LFRV_OK) return s;
    }
    const size_t N = ctx->N;
    const float s2 = ctx->p.sigma_sensor * ctx->p.sigma_sensor;
    const float wscale = use_views ? (float)(alpha * s2) : 0.0f;
    // right-hand side and preconditioner; the Poisson form uses u0 = U
    float *u0 = ctx->tmp3N;
    if (poisson_rhs) LFRV_HIP(ctx, hipMemcpyAsync(u0, ctx->U, 3 * N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    lfrv_status s = run_rhs_and_diag(ctx, alpha, c0, beta, use_views, use_precond, u0);
    if (s != LFRV_OK) return s;

    const TileGeom g = tile_geom(ctx);
    TileEntries te = tile_entries(ctx);
    te.active = use_views ? 1 : 0;
    const int HAL = ctx->ks - 1;
    const size_t dtile = sizeof(float) * 3 * (size_t)(ctx->tile_w + 2 * HAL) * (ctx->tile_h + 2 * HAL);
    const int cap = pull_cap(ctx, 3, dtile, ctx->pull_budget, walk_kind(ctx));
    const size_t lds = dtile + pull_lds(ctx, 3, cap, walk_kind(ctx));
    const unsigned int rb = reduce_blocks(N);
    float *x = ctx->U, *r = ctx->cg_r, *q = ctx->cg_q, *minv = ctx->cg_minv;
    float *dbuf[2] = {ctx->cg_d, ctx->tmp3N};
    CgState *stt = ctx->cg_state;
    double *part = ctx->partials;

    // dsf = 1 with views: assemble the 9-point stencil once per entries / weights
    const bool st2 = ctx->ks == 2 && use_views && wscale != 0.0f && ctx->gram != nullptr;
    const int st2_ntx = (ctx->W + kSt2W - 1) / kSt2W, st2_tiles = st2_ntx * ((ctx->H + kSt2H - 1) / kSt2H);
    float *gD = ctx->gram, *gE = gD + N, *gS = gE + N, *gSE = gS + N, *gSW = gSE + N;
    if (st2 && ctx->gram_epoch != ctx->epoch) {
        TimedRegion tr(ctx, 5);
        const int gcap = pull_cap(ctx, 3, 0, ctx->pull_budget, PullKind::Cell);
        const size_t glds = pull_lds(ctx, 3, gcap, PullKind::Cell);  // 12 floats per window cell >= the 10 Gram sums
        k_gram2<<<ctx->ntiles, ctx->pull_bs, glds, ctx->stream>>>(g, te, gcap, gD, gE, gS, gSE, gSW);
        LFRV_LAUNCH(ctx, "stencil assembly");
        ctx->gram_epoch = ctx->epoch;
    }
    const int op_blocks = st2 ? st2_tiles : ctx->ntiles;
    // dsf >= 2 with views: the assembled box operator
    const bool boxp = !st2 && ctx->ks >= 3 && use_views && wscale != 0.0f && ctx->box_ops != nullptr;
    const int Wc = ctx->W + ctx->ks - 1, Hc = ctx->H + ctx->ks - 1;
    const size_t NA = (size_t)Wc * Hc;
    const int box_ntx = (ctx->W + kBoxTW - 1) / kBoxTW, box_tiles = box_ntx * ((ctx->H + kBoxTH - 1) / kBoxTH);
    if (boxp && ctx->box_epoch != ctx->epoch) {
        TimedRegion tr(ctx, 5);
        TileGeom ga = g;
        ga.ntx = (Wc + g.tw - 1) / g.tw;
        ga.ntiles = ga.ntx * ((Hc + g.th - 1) / g.th);
        const int wh = g.th + 1, ncw = (g.tw + 1) * wh;
        const int gcap = std::min(1024, kStagePer * ctx->pull_bs);
        const size_t glds = sizeof(int) * (2 * wh + 2) + sizeof(float) * std::max<size_t>(10 * (size_t)ncw, 3 * (size_t)gcap);
#define GB(K) k_gram_box<K><<<ga.ntiles, ctx->pull_bs, glds, ctx->stream>>>(ga, te, gcap, NA, ctx->box_ops)
        switch (ctx->ks) {
        case 3: GB(3); break;
        case 4: GB(4); break;
        case 5: GB(5); break;
        default: return fail(ctx, LFRV_E_ARG, "dsf out of range");
        }
#undef GB
        LFRV_LAUNCH(ctx, "box operator assembly");
        ctx->box_epoch = ctx->epoch;
    }

#define OPK(K, MODE, DIN, DOUT)                                                                          \
    k_cg_op<K><<<ctx->ntiles, ctx->pull_bs, lds, ctx->stream>>>(g, te, cap, wscale, (float)c0, (float)beta, \
                                                                MODE, DIN, r, minv, DOUT, q, stt, part)
#define LAUNCH_OP(MODE, DIN, DOUT)                     \
    switch (ctx->ks) {                                 \
    case 2: OPK(2, MODE, DIN, DOUT); break;            \
    case 3: OPK(3, MODE, DIN, DOUT); break;            \
    case 4: OPK(4, MODE, DIN, DOUT); break;            \
    case 5: OPK(5, MODE, DIN, DOUT); break;            \
    default: return fail(ctx, LFRV_E_ARG, "dsf out of range"); \
    }

    const St2Coeffs gco{gD, gE, gS, gSE, gSW};
    if (st2) {
        // r_0 = b - A x_0, u_0 = Minv r_0, w_0 = A u_0 and the start sums in one pass
        {
            TimedRegion tr(ctx, 5);
            k_cgcg_begin_st2<<<st2_tiles, kSt2W * kSt2H, 0, ctx->stream>>>(
                ctx->W, ctx->H, st2_ntx, st2_tiles, wscale, (float)c0, (float)beta, gco, x, ctx->cg_b, minv, r,
                ctx->cg_q, part);
        }
        k_cgcg_begin<<<1, kFinalizeThreads, 0, ctx->stream>>>(stt, part, st2_tiles, tol * tol, maxit);
    } else if (boxp) {
        // r_0 = b - A x_0 and u_0 = Minv r_0, then w_0 = A u_0 with the start sums
        float *u = ctx->cg_extra + 9 * N;
        TimedRegion tr(ctx, 5);
#define BP(K, M, IN, R, OUT, OUT2) k_box_pass<K, M><<<box_tiles, kBoxTW * kBoxTH, 0, ctx->stream>>>(ctx->W, ctx->H, Wc, Hc, box_ntx, box_tiles, wscale, (float)c0, (float)beta, ctx->box_ops, NA, IN, ctx->cg_b, R, minv, OUT, OUT2, stt, part)
        switch (ctx->ks) {
        case 3: BP(3, kBoxResInit, x, nullptr, r, u); BP(3, kBoxBegin, u, r, ctx->cg_q, nullptr); break;
        case 4: BP(4, kBoxResInit, x, nullptr, r, u); BP(4, kBoxBegin, u, r, ctx->cg_q, nullptr); break;
        case 5: BP(5, kBoxResInit, x, nullptr, r, u); BP(5, kBoxBegin, u, r, ctx->cg_q, nullptr); break;
        }
        k_cgcg_begin<<<1, kFinalizeThreads, 0, ctx->stream>>>(stt, part, box_tiles, tol * tol, maxit);
    } else {
        {
            TimedRegion tr(ctx, 5);
            LAUNCH_OP(0, x, nullptr);
        }
        k_cg_init<<<rb, 256, 0, ctx->stream>>>(N, ctx->cg_b, q, r, minv, part);
        k_cg_init_state<<<1, kFinalizeThreads, 0, ctx->stream>>>(stt, part, (int)rb, tol * tol, maxit);
    }
    LFRV_LAUNCH(ctx, "cg init");
    int it = 0, cur = 0;
    const int chunk = 8;
    bool done = false, final_ready = false;
    if (st2) {
        // single-reduction iterations on the stencil; r, s, w double-buffered
        float *rbuf[2] = {ctx->cg_r, ctx->cg_extra};
        float *sbuf[2] = {ctx->cg_extra + 3 * N, ctx->tmp3N};
        float *wbuf[2] = {ctx->cg_q, ctx->cg_extra + 6 * N};
        CgcgArgs a{};
        a.W = ctx->W;
        a.H = ctx->H;
        a.ntx = st2_ntx;
        a.ntiles = st2_tiles;
        a.wscale = wscale;
        a.c0 = (float)c0;
        a.beta_reg = (float)beta;
        a.minv = minv;
        a.p = ctx->cg_d;
        a.x = x;
        a.g = gco;
        // true residual ||b - A x||^2 of the current x, then the state to the host
        auto st2_poll = [&]() -> lfrv_status {
            {
                TimedRegion tr(ctx, 5);
                k_st2_residual<<<st2_tiles, kSt2W * kSt2H, 0, ctx->stream>>>(
                    ctx->W, ctx->H, st2_ntx, st2_tiles, wscale, (float)c0, (float)beta, gco, x, ctx->cg_b, part);
            }
            k_reduce_out<3><<<1, kFinalizeThreads, 0, ctx->stream>>>(part, st2_tiles, stt->resid);
            LFRV_LAUNCH(ctx, "cg residual");
            LFRV_HIP(ctx, hipMemcpyAsync(ctx->cg_state_host, stt, sizeof(CgState), hipMemcpyDeviceToHost, ctx->stream));
            LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
            final_ready = true;
            return LFRV_OK;
        };
        LFRV_LAUNCH(ctx, "cg start");
        // launch as many iterations as the last solve needed before the first
        // look at the convergence flag, then small chunks
        int next_chunk = std::max(1, ctx->cg_iter_hint);
        while (it < maxit && !done) {
            const int nk = std::min(next_chunk, maxit - it);
            next_chunk = 4;
            for (int k = 0; k < nk; k++) {
                a.r_in = rbuf[cur];
                a.w_in = wbuf[cur];
                a.s_in = sbuf[cur];
                a.r_out = rbuf[cur ^ 1];
                a.w_out = wbuf[cur ^ 1];
                a.s_out = sbuf[cur ^ 1];
                {
                    const KernelEvents ev = kernel_events(ctx, 0);
                    hipExtLaunchKernelGGL(k_cgcg_st2, dim3(st2_tiles), dim3(kSt2W * kSt2H), 0, ctx->stream,
                                          ev.a, ev.b, 0, a, (const CgState *)stt, part);
                    record_kernel_events(ctx, 0, ev);
                }
                k_cgcg_step<<<1, kFinalizeThreads, 0, ctx->stream>>>(stt, part, st2_tiles);
                cur ^= 1;
            }
            LFRV_LAUNCH(ctx, "cg iteration");
            it += nk;
            // the true residual of the current x goes out with the poll: if the
            // chunk converged it is the answer, and the host waits once
            lfrv_status ps = st2_poll();
            if (ps != LFRV_OK) return ps;
            done = ctx->cg_state_host->done != 0;
        }
        if (final_ready) {
            const CgState &hs = *ctx->cg_state_host;
            ctx->cg_iter_hint = std::max(hs.iter[0], std::max(hs.iter[1], hs.iter[2]));
        } else {
            // maxit == 0: no iteration ran; the state and true residual of x_0
            // come from the same stencil residual pass
            lfrv_status ps = st2_poll();
            if (ps != LFRV_OK) return ps;
        }
        done = true;
    }
    if (boxp) {
        // single-reduction iterations on the box operator: the vector updates,
        // then w = A u with the iteration's sums, then the scalar step; r, s, w
        // double-buffered as in the dsf = 1 stencil path
        float *rbuf[2] = {ctx->cg_r, ctx->cg_extra};
        float *sbuf[2] = {ctx->cg_extra + 3 * N, ctx->tmp3N};
        float *wbuf[2] = {ctx->cg_q, ctx->cg_extra + 6 * N};
        float *u = ctx->cg_extra + 9 * N;
        const unsigned int ub = (unsigned int)std::min<size_t>((N + 255) / 256, 8192);
        auto box_poll = [&]() -> lfrv_status {
            {
                TimedRegion tr(ctx, 5);
                switch (ctx->ks) {
                case 3: BP(3, kBoxResid, x, nullptr, nullptr, nullptr); break;
                case 4: BP(4, kBoxResid, x, nullptr, nullptr, nullptr); break;
                case 5: BP(5, kBoxResid, x, nullptr, nullptr, nullptr); break;
                }
            }
            k_reduce_out<3><<<1, kFinalizeThreads, 0, ctx->stream>>>(part, box_tiles, stt->resid);
            LFRV_LAUNCH(ctx, "cg residual");
            LFRV_HIP(ctx, hipMemcpyAsync(ctx->cg_state_host, stt, sizeof(CgState), hipMemcpyDeviceToHost, ctx->stream));
            LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
            final_ready = true;
            return LFRV_OK;
        };
        LFRV_LAUNCH(ctx, "cg start");
        int next_chunk = std::max(1, ctx->cg_iter_hint);
        while (it < maxit && !done) {
            const int nk = std::min(next_chunk, maxit - it);
            next_chunk = 4;
            for (int k = 0; k < nk; k++) {
                {
                    TimedRegion tr(ctx, 1);
                    k_box_update<<<ub, 256, 0, ctx->stream>>>(N, stt, minv, rbuf[cur], wbuf[cur], sbuf[cur],
                                                              rbuf[cur ^ 1], sbuf[cur ^ 1], ctx->cg_d, x, u);
                }
                {
                    TimedRegion tr(ctx, 0);
                    switch (ctx->ks) {
                    case 3: BP(3, kBoxOp, u, rbuf[cur ^ 1], wbuf[cur ^ 1], nullptr); break;
                    case 4: BP(4, kBoxOp, u, rbuf[cur ^ 1], wbuf[cur ^ 1], nullptr); break;
                    case 5: BP(5, kBoxOp, u, rbuf[cur ^ 1], wbuf[cur ^ 1], nullptr); break;
                    }
                }
                k_cgcg_step<<<1, kFinalizeThreads, 0, ctx->stream>>>(stt, part, box_tiles);
                cur ^= 1;
            }
            LFRV_LAUNCH(ctx, "cg iteration");
            it += nk;
            lfrv_status ps = box_poll();
            if (ps != LFRV_OK) return ps;
            done = ctx->cg_state_host->done != 0;
        }
        if (final_ready) {
            const CgState &hs = *ctx->cg_state_host;
            ctx->cg_iter_hint = std::max(hs.iter[0], std::max(hs.iter[1], hs.iter[2]));
        } else {
            lfrv_status ps = box_poll();  // maxit == 0
            if (ps != LFRV_OK) return ps;
        }
        done = true;
    }
#undef BP
    while (it < maxit && !done) {
        const int nk = std::min(chunk, maxit - it);
        for (int k = 0; k < nk; k++) {
            float *dprev = dbuf[cur], *dnew = dbuf[cur ^ 1];
            {
                TimedRegion tr(ctx, 0);
                LAUNCH_OP(1, dprev, dnew);
            }
            k_cg_alpha<<<1, kFinalizeThreads, 0, ctx->stream>>>(stt, part, op_blocks);
            {
                TimedRegion tr(ctx, 1);
                k_cg_update<<<rb, 256, 0, ctx->stream>>>(N, x, dnew, q, r, minv, stt, part);
            }
            k_cg_beta<<<1, kFinalizeThreads, 0, ctx->stream>>>(stt, part, (int)rb);
            cur ^= 1;
        }
        LFRV_LAUNCH(ctx, "cg iteration");
        it += nk;
        LFRV_HIP(ctx, hipMemcpyAsync(ctx->cg_state_host, stt, sizeof(CgState), hipMemcpyDeviceToHost, ctx->stream));
        LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
        done = ctx->cg_state_host->done != 0;
    }
#undef LAUNCH_OP
    // final state and true residual (the stencil path has it with its last poll)
    if (!final_ready) {
        {
            TimedRegion tr(ctx, 5);
            switch (ctx->ks) {
            case 2: OPK(2, 0, x, nullptr); break;
            case 3: OPK(3, 0, x, nullptr); break;
            case 4: OPK(4, 0, x, nullptr); break;
            case 5: OPK(5, 0, x, nullptr); break;
            }
        }
        k_resid<<<rb, 256, 0, ctx->stream>>>(N, ctx->cg_b, q, part);
        k_reduce_out<3><<<1, kFinalizeThreads, 0, ctx->stream>>>(part, (int)rb, stt->resid);
        LFRV_LAUNCH(ctx, "cg residual");
        LFRV_HIP(ctx, hipMemcpyAsync(ctx->cg_state_host, stt, sizeof(CgState), hipMemcpyDeviceToHost, ctx->stream));
        LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
#undef OPK
    return cg_finish_stats(ctx, st, t_start);
}

// Statistics of a finished solve from the state copied to the host.
static lfrv_status cg_finish_stats(lfrv_ctx *ctx, lfrv_cg_stats *st,
                                   std::chrono::steady_clock::time_point t_start)
{
    flush_timing(ctx);
    const CgState &hs = *ctx->cg_state_host;
    lfrv_status status = LFRV_OK;
    for (int c = 0; c < 3; c++) {
        const bool conv = hs.rr[c] <= hs.tol2 * hs.bb[c];
        if (st) {
            st->iterations[c] = hs.iter[c];
            st->converged[c] = conv ? 1 : 0;
            st->rel_residual[c] = hs.bb[c] > 0 ? std::sqrt(hs.rr[c] / hs.bb[c]) : std::sqrt(hs.rr[c]);
            st->true_rel_residual[c] = hs.bb[c] > 0 ? std::sqrt(hs.resid[c] / hs.bb[c]) : std::sqrt(hs.resid[c]);
        }
        if (hs.breakdown[c]) status = LFRV_E_BREAKDOWN;
        else if (!conv && status == LFRV_OK) status = LFRV_E_NONCONVERGED;
        if (!std::isfinite(hs.rr[c])) status = LFRV_E_NAN;
    }
    if (st) {
        st->status = status;
        st->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    }
    if (status != LFRV_OK) {
        const char *msg = status == LFRV_E_BREAKDOWN ? "CG breakdown (d.q <= 0)"
                          : status == LFRV_E_NAN     ? "CG produced a non-finite residual"
                                                     : "CG did not reach the tolerance";
        return fail(ctx, status, msg);
    }
    return LFRV_OK;
}

// ---- batches of dsf = 1 solves ------------------------------------------------

lfrv_status cg_batch_begin(lfrv_ctx *ctx, double tol, int maxit, CgBatchItem *item)
{
    if (!ctx->prepared) return fail(ctx, LFRV_E_STATE, "batched CG before lfrv_prepare");
    const double alpha = ctx->p.dt_alpha, beta = ctx->p.dt_beta;
    const float s2 = ctx->p.sigma_sensor * ctx->p.sigma_sensor;
    const float wscale = (float)(alpha * s2);
    if (ctx->ks != 2 || !ctx->gram || wscale == 0.0f)
        return fail(ctx, LFRV_E_ARG, "the batched CG takes dsf 1 solves with the data term");
    ctx->u_zero = false;
    if (!(ctx->gram_epoch == ctx->epoch)) {
        lfrv_status s = run_refresh_entry_weights(ctx, 1.0f);
        if (s != LFRV_OK) return s;
    }
    lfrv_status s = run_rhs_and_diag(ctx, alpha, 0.0, beta, true, true, ctx->tmp3N);
    if (s != LFRV_OK) return s;
    const size_t N = ctx->N;
    const int ntx = (ctx->W + kSt2W - 1) / kSt2W, ntiles = ntx * ((ctx->H + kSt2H - 1) / kSt2H);
    float *gD = ctx->gram, *gE = gD + N, *gS = gE + N, *gSE = gS + N, *gSW = gSE + N;
    if (ctx->gram_epoch != ctx->epoch) {
        const TileGeom g = tile_geom(ctx);
        TileEntries te = tile_entries(ctx);
        te.active = 1;
        const int gcap = pull_cap(ctx, 3, 0, ctx->pull_budget, PullKind::Cell);
        const size_t glds = pull_lds(ctx, 3, gcap, PullKind::Cell);
        k_gram2<<<ctx->ntiles, ctx->pull_bs, glds, ctx->stream>>>(g, te, gcap, gD, gE, gS, gSE, gSW);
        LFRV_LAUNCH(ctx, "stencil assembly");
        ctx->gram_epoch = ctx->epoch;
    }
    const St2Coeffs gco{gD, gE, gS, gSE, gSW};
    float *x = ctx->U, *minv = ctx->cg_minv;
    k_cgcg_begin_st2<<<ntiles, kSt2W * kSt2H, 0, ctx->stream>>>(ctx->W, ctx->H, ntx, ntiles, wscale, 0.0f,
                                                                 (float)beta, gco, x, ctx->cg_b, minv, ctx->cg_r,
                                                                 ctx->cg_q, ctx->partials);
    k_cgcg_begin<<<1, kFinalizeThreads, 0, ctx->stream>>>(ctx->cg_state, ctx->partials, ntiles, tol * tol, maxit);
    LFRV_LAUNCH(ctx, "batched cg start");
    float *rbuf[2] = {ctx->cg_r, ctx->cg_extra};
    float *sbuf[2] = {ctx->cg_extra + 3 * N, ctx->tmp3N};
    float *wbuf[2] = {ctx->cg_q, ctx->cg_extra + 6 * N};
    for (int par = 0; par < 2; par++) {
        CgcgArgs &a = item->a[par];
        a = CgcgArgs{};
        a.W = ctx->W;
        a.H = ctx->H;
        a.ntx = ntx;
        a.ntiles = ntiles;
        a.wscale = wscale;
        a.c0 = 0.0f;
        a.beta_reg = (float)beta;
        a.minv = minv;
        a.p = ctx->cg_d;
        a.x = x;
        a.g = gco;
        a.r_in = rbuf[par];
        a.w_in = wbuf[par];
        a.s_in = sbuf[par];
        a.r_out = rbuf[par ^ 1];
        a.w_out = wbuf[par ^ 1];
        a.s_out = sbuf[par ^ 1];
    }
    item->st = ctx->cg_state;
    item->partials = ctx->partials;
    return LFRV_OK;
}

hipError_t launch_cg_batch_iteration(const CgBatchItem *items, const int2 *work, int nblocks, int parity, hipStream_t s)
{
    k_cgcg_st2_batch<<<nblocks, kSt2W * kSt2H, 0, s>>>(items, work, parity);
    return hipGetLastError();
}

hipError_t launch_cg_batch_step(const CgBatchItem *items, int nitems, hipStream_t s)
{
    k_cgcg_step_batch<<<nitems, kFinalizeThreads, 0, s>>>(items);
    return hipGetLastError();
}

lfrv_status cg_batch_end(lfrv_ctx *ctx, lfrv_cg_stats *st)
{
    const auto t0 = std::chrono::steady_clock::now();
    const size_t N = ctx->N;
    const float s2 = ctx->p.sigma_sensor * ctx->p.sigma_sensor;
    const float wscale = (float)(ctx->p.dt_alpha * s2);
    const int ntx = (ctx->W + kSt2W - 1) / kSt2W, ntiles = ntx * ((ctx->H + kSt2H - 1) / kSt2H);
    float *gD = ctx->gram, *gE = gD + N, *gS = gE + N, *gSE = gS + N, *gSW = gSE + N;
    const St2Coeffs gco{gD, gE, gS, gSE, gSW};
    k_st2_residual<<<ntiles, kSt2W * kSt2H, 0, ctx->stream>>>(ctx->W, ctx->H, ntx, ntiles, wscale, 0.0f,
                                                               (float)ctx->p.dt_beta, gco, ctx->U, ctx->cg_b,
                                                               ctx->partials);
    k_reduce_out<3><<<1, kFinalizeThreads, 0, ctx->stream>>>(ctx->partials, ntiles, ctx->cg_state->resid);
    LFRV_LAUNCH(ctx, "batched cg residual");
    LFRV_HIP(ctx, hipMemcpyAsync(ctx->cg_state_host, ctx->cg_state, sizeof(CgState), hipMemcpyDeviceToHost, ctx->stream));
    LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const CgState &hs = *ctx->cg_state_host;
    ctx->cg_iter_hint = std::max(hs.iter[0], std::max(hs.iter[1], hs.iter[2]));
    return cg_finish_stats(ctx, st, t0);
}

lfrv_status run_gibr_energy(lfrv_ctx *ctx, double *e)
{
    const float s2 = ctx->p.sigma_sensor * ctx->p.sigma_sensor;
    const size_t tot = (size_t)ctx->V * ctx->n;
    const unsigned int rb1 = reduce_blocks(tot), rb2 = reduce_blocks(ctx->N);
#define EN(K) k_gibr_energy_data<K><<<rb1, 256, 0, ctx->stream>>>(ctx->V, ctx->W, ctx->H, ctx->w, ctx->h, ctx->tau_x, ctx->tau_y, ctx->omega, s2, ctx->views, ctx->U, ctx->partials)
    switch (ctx->ks) {
    case 2: EN(2); break;
    case 3: EN(3); break;
    case 4: EN(4); break;
    case 5: EN(5); break;
    default: return fail(ctx, LFRV_E_ARG, "dsf out of range");
    }
#undef EN
    k_reduce_out<1><<<1, kFinalizeThreads, 0, ctx->stream>>>(ctx->partials, (int)rb1, ctx->scal + 4);
    k_gibr_energy_grad<<<rb2, 256, 0, ctx->stream>>>(ctx->W, ctx->H, ctx->U, ctx->u_grad_x, ctx->u_grad_y,
                                                     ctx->partials);
    k_reduce_out<1><<<1, kFinalizeThreads, 0, ctx->stream>>>(ctx->partials, (int)rb2, ctx->scal + 5);
    LFRV_LAUNCH(ctx, "gibr energy");
    LFRV_HIP(ctx, hipMemcpyAsync(ctx->scal_host + 4, ctx->scal + 4, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *e = 0.5 * ((double)ctx->p.dt_alpha * ctx->scal_host[4] + (double)ctx->p.dt_beta * ctx->scal_host[5]);
    return LFRV_OK;
}

}  // namespace lfrv
