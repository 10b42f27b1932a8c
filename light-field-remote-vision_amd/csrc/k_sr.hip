// IBR_color (TV super-resolution with warped data terms): the data-term
// gradient and the outer proximal step (dsf = 1: an entry pull; dsf >= 2: the
// assembled anchor-grid operator as a stencil, fused with the first
// primal-dual iteration), the primal energy, and the re-weighting loop with
// its stopping rules on the device.
#include "tile_common.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

namespace lfrv {

// k_stencil.hip; with `overrelax` the last step also does the over-relaxation
lfrv_status run_pd_steps(lfrv_ctx *ctx, int iters, const float *F, bool overrelax = false, float theta = 0.0f,
                         const int *halt = nullptr);

namespace {


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
                                                     float *__restrict__ norm_mask, const int *halt)
{
    if (halt && *halt) return;  // a step enqueued behind a stopping rule that fired
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

// ---- dsf >= 2: the data term as an assembled operator -------------------------
//
// With t_e the 2 x 2 anchor taps of entry e (box_weights) and BU, BL the box
// sums of U and lap U at the anchors, the data-term sums of one outer step are
// linear in U while omega is fixed:
//   sum_e omega t_e (t_e^T BU - v_e)                 = G9 BU - bI
//   sum_e t_e (dsf^2 t_e^T BL - lap v_e) [* omega]     = dsf^2 G9' BL - bG
// G9 = sum omega t t^T couples each anchor with its 8 neighbours (centre, E,
// S, SE, SW planes; the others are the neighbours' couplings), G9' is the same
// without omega (gw_type 0), bI = sum omega t v, bG = sum t lap v [* omega].
// They are assembled once per set of weights (k_sr_assemble); every outer
// step is then a tiled stencil (k_sr_dataterm_st): box sums, the 9-point
// couplings, and the box sums back onto the pixels.  The step weighs the two
// terms' back-projections by constants of the parameters before adding them
// per pixel (cI = alpha sigma^2, cG = -beta gradient_step [sigma^2]), and the
// back-projection is linear, so with the gradient term the step back-projects
// one residual, cI (G9 BU - bI) + cG (dsf^2 G9' BL - bG), and the assembly
// stores the constant part as one plane set b = cI bI + cG bG.
//
// Anchor grid: point x' in [0, W + dsf) has box start x' + 1 - dsf; it is the
// top-left tap of the cell x' + 1 (cells hold box start cx - dsf).  Pixel x
// lies in the boxes of the points x .. x + dsf - 1.
// (kSrOpsPlanes planes: G9 (5), bI (3) or b (3) with the gradient term, G9' (5))

template <int KS>
__global__ void __launch_bounds__(kPullMaxThreads) k_sr_assemble(TileGeom g, TileEntries te, int cap, int grad,
                                                                 int gw, float ci, float cg, size_t NA,
                                                                 float *__restrict__ ops)
{
    extern __shared__ int smem[];
    constexpr int NC = 44;  // 10 weighted couplings, 12 bI taps, 12 bG taps, 10 unweighted couplings
    const int t = xcd_tile(blockIdx.x, g.ntiles);
    const int tx0 = (t % g.ntx) * g.tw, ty0 = (t / g.ntx) * g.th;
    const int ww = g.tw + 1, wh = g.th + 1, ncw = ww * wh;
    float acc[NC];
    cell_accumulate<KS, NC, false, true, 6, 1>(
        g, te, t, smem, cap, acc,
        [&](int, int, int, const float(&ax)[2], const float(&ay)[2], float om, const float(&pay)[6],
            float(&a)[NC]) {
            const float w[4] = {ay[0] * ax[0], ay[0] * ax[1], ay[1] * ax[0], ay[1] * ax[1]};  // tl tr bl br
            const float pr[10] = {w[0] * w[0], w[1] * w[1], w[2] * w[2], w[3] * w[3], w[0] * w[1],
                                  w[2] * w[3], w[0] * w[2], w[1] * w[3], w[0] * w[3], w[1] * w[2]};
#pragma unroll
            for (int k = 0; k < 10; k++) a[k] += om * pr[k];
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int c = 0; c < 3; c++) a[10 + k * 3 + c] += om * w[k] * pay[c];
            if (grad) {
                const float og = gw ? om : 1.0f;
#pragma unroll
                for (int k = 0; k < 4; k++)
#pragma unroll
                    for (int c = 0; c < 3; c++) a[22 + k * 3 + c] += og * w[k] * pay[3 + c];
#pragma unroll
                for (int k = 0; k < 10; k++) a[34 + k] += pr[k];
            }
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
        // own cell (its tl tap), left cell (tr), upper cell (bl), upper-left cell (br)
        const int c00 = (ly + 1) * ww + lx + 1, c10 = c00 - 1, c01 = c00 - ww, c11 = c00 - ww - 1;
        const size_t o = (size_t)y * g.Wc + x;
        auto couple = [&](int base, int plane) {
            ops[(size_t)(plane + 0) * NA + o] = part[(base + 0) * ncw + c00] + part[(base + 1) * ncw + c10] +
                                                part[(base + 2) * ncw + c01] + part[(base + 3) * ncw + c11];
            ops[(size_t)(plane + 1) * NA + o] = part[(base + 4) * ncw + c00] + part[(base + 5) * ncw + c01];
            ops[(size_t)(plane + 2) * NA + o] = part[(base + 6) * ncw + c00] + part[(base + 7) * ncw + c10];
            ops[(size_t)(plane + 3) * NA + o] = part[(base + 8) * ncw + c00];
            ops[(size_t)(plane + 4) * NA + o] = part[(base + 9) * ncw + c10];
        };
        auto tap = [&](int base, int ch) {
            return part[(base + 0 * 3 + ch) * ncw + c00] + part[(base + 1 * 3 + ch) * ncw + c10] +
                   part[(base + 2 * 3 + ch) * ncw + c01] + part[(base + 3 * 3 + ch) * ncw + c11];
        };
        couple(0, 0);
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            const float bI = tap(10, ch);
            ops[(size_t)(5 + ch) * NA + o] = grad ? ci * bI + cg * tap(22, ch) : bI;
        }
        if (grad) couple(34, 8);
    }
}

// One outer step's data term on a 32 x 16 pixel tile from the assembled
// operator.  Every global load of the block (U with its halo, the couplings
// and right-hand sides of the tile's anchors, coverage) is issued first; the
// rest runs in LDS: lap U, the box sums BU / BL at the anchors (+1 ring), the
// anchor residuals QI = G9 BU - bI (with the gradient term the weighted sum
// cI G9 BU + cG dsf^2 G9' BL - b), then the box sums back onto the pixels,
// split into rows and columns.
constexpr int kStTW = 32, kStTH = 16;

// Every LDS plane of the tile has the row pitch P of the widest one (U with
// its ring) and every pass maps consecutive threads to consecutive points of
// that pitch (idle where x is past the pass's width), so a wave's accesses are
// contiguous and free of bank conflicts; a pitch per plane made them jump at
// each row wrap.
// GRAD: the gradient-constraint term is on (dt_beta > 0): lap U, its box
// sums and residuals are formed too (NC = 6 planes instead of 3).
template <int KS, bool GRAD = true>
struct StDims {
    static constexpr int D = KS - 1, TW = kStTW, TH = kStTH, NT = TW * TH;
    static constexpr int NC = GRAD ? 6 : 3;
    static constexpr int QW = TW + D - 1, QH = TH + D - 1;  // anchor points x' = tx0 + qx
    static constexpr int BW = QW + 2, BH = QH + 2;          // box sums at x' = tx0 - 1 + bx
    static constexpr int LW = BW + D - 1, LH = BH + D - 1;  // pixels tx0 - D + lx (box sums' support)
    static constexpr int SW = LW + 2, SH = LH + 2;          // U from tx0 - D - 1 (the Laplacian's ring)
    static constexpr int P = SW;                            // the common row pitch
    // The two LDS regions take turns: sa holds U and lap U, sb their row sums
    // (D > 2); then the box sums BU, BL go to the region nobody reads any more
    // (sa for D > 2, sb for D <= 2), the anchor residuals to the other one, and
    // (D > 2) the residuals' row sums to sa again; sb ends with the duals of
    // the primal-dual step (6 planes of (TH + 1) x (TW + 2)).
    static constexpr int A1 = 3 * SH * P + (GRAD ? 3 * LH * P : 0), A2 = 3 * QH * P;  // (3 residual planes)
    static constexpr int AREA = A1 > A2 ? A1 : A2;
    static constexpr int HR = BH + D - 1;  // rows of the horizontal box sums
    // D > 2: the row sums keep only the BW columns the column sums read (pitch
    // BW: a wave's active lanes then write and read consecutive words);
    // D <= 2: the box sums themselves
    static constexpr int B1 = D > 2 ? NC * HR * BW : NC * BH * P, B2 = 6 * (TH + 1) * (TW + 2);
    static constexpr int BAREA = B1 > B2 ? B1 : B2;  // sb
    static constexpr int NQP = QH * P;               // anchor slots (qx < QW used)
    static constexpr int NQT = (NQP + NT - 1) / NT, NST = (SH * P + NT - 1) / NT;
};

// The data term of one 32 x 16 tile.  Every thread of the block calls it; on
// return each thread holds F (the clamped gradient step) and U at its pixel,
// and the pixel's coverage; sb is free for the caller at once, sa only after
// a barrier.  With ur, the U values of
// [tx0-1, tx0+TW] x [ty0-1, ty0+TH] (3 planes, pitch TW+2) are also left there.
template <int KS, bool GRAD>
__device__ __forceinline__ void sr_dataterm_tile(int W, int H, int Wc, int Hc, int tx0, int ty0, const DataTermArgs &da,
                                                 const float *__restrict__ U, const float *__restrict__ ops, size_t NA,
                                                 const float *__restrict__ coverage, float *sa, float *sb,
                                                 float (&F)[3], float (&up)[3], float &cov, float *ur = nullptr)
{
    using S = StDims<KS, GRAD>;
    constexpr int D = S::D, TW = S::TW, NT = S::NT, QW = S::QW, QH = S::QH, BW = S::BW, BH = S::BH, LW = S::LW,
                  LH = S::LH, SH = S::SH, P = S::P, NQP = S::NQP, NQT = S::NQT, NST = S::NST;
    constexpr bool SEP = D > 2;  // separable box sums
    static_assert(S::NC * BH * P <= (SEP ? S::AREA : S::BAREA) && 3 * NQP <= (SEP ? S::BAREA : S::AREA) &&
                      3 * QH * TW <= S::AREA,
                  "the box sums, residuals and row sums fit their LDS regions");
    float *us = sa, *ls = sa + 3 * SH * P;
    float *bus = SEP ? sa : sb, *bls = bus + 3 * BH * P;
    float *qs = SEP ? sb : sa;
    float *rs = sa;  // D > 2: the residuals' row sums, once BU and BL are read
    const size_t N = (size_t)W * H;
    constexpr bool grad = GRAD;  // da.beta > 0
    const float *g9 = da.gw_type ? ops : ops + 8 * NA;  // couplings of the gradient term
    const float ci = dt_intensity_weight(da), cgd = dt_gradient_weight(da) * da.dsf2;
    const int tid = threadIdx.x;

    // ---- all global loads ----
    float uv[NST][3];
#pragma unroll
    for (int k = 0; k < NST; k++) {
        const int i = tid + k * NT;
        const int x = tx0 - D - 1 + i % P, y = ty0 - D - 1 + i / P;
        const bool in = i < SH * P && x >= 0 && x < W && y >= 0 && y < H;
        const size_t o = in ? (size_t)y * W + x : 0;
#pragma unroll
        for (int c = 0; c < 3; c++) uv[k][c] = in ? U[c * N + o] : 0.0f;
    }
    float cf[NQT][9], cg[NQT][9], bi[NQT][3];  // bi: bI, or b with the gradient term
#pragma unroll
    for (int k = 0; k < NQT; k++) {
        const int i = tid + k * NT;
        const int xq = tx0 + i % P, yq = ty0 + i / P;
        const bool in = i < NQP && i % P < QW && xq < Wc && yq < Hc;
        const int xc = min(xq, Wc - 1), yc = min(yq, Hc - 1);
        const size_t o = (size_t)yc * Wc + xc;
        const bool xm = xc > 0, ym = yc > 0, xp = xc < Wc - 1;
        const size_t oE = xm ? o - 1 : o, oS = ym ? o - Wc : o, oSE = (xm && ym) ? o - Wc - 1 : o,
                     oSW = (xp && ym) ? o - Wc + 1 : o;
        auto load9 = [&](const float *Q, float(&f)[9]) {
            f[0] = Q[o];
            f[1] = Q[NA + o];
            f[2] = Q[2 * NA + o];
            f[3] = Q[3 * NA + o];
            f[4] = Q[4 * NA + o];
            const float e = Q[NA + oE], sn = Q[2 * NA + oS], se = Q[3 * NA + oSE], sw = Q[4 * NA + oSW];
            f[5] = xm ? e : 0.0f;
            f[6] = ym ? sn : 0.0f;
            f[7] = (xm && ym) ? se : 0.0f;
            f[8] = (xp && ym) ? sw : 0.0f;
            if (!in)
#pragma unroll
                for (int q = 0; q < 9; q++) f[q] = 0.0f;
        };
        load9(ops, cf[k]);
#pragma unroll
        for (int c = 0; c < 3; c++) bi[k][c] = in ? ops[(size_t)(5 + c) * NA + o] : 0.0f;
        if (grad) load9(g9, cg[k]);
    }
    const int px = tx0 + tid % TW, py = ty0 + tid / TW;
    const bool pin = px < W && py < H;
    const size_t po = pin ? (size_t)py * W + px : 0;
    cov = pin ? coverage[po] : 0.0f;
#pragma unroll
    for (int c = 0; c < 3; c++) up[c] = pin ? U[c * N + po] : 0.0f;

    // ---- U, lap U, box sums ----
#pragma unroll
    for (int k = 0; k < NST; k++) {
        const int i = tid + k * NT;
        if (i < SH * P)
#pragma unroll
            for (int c = 0; c < 3; c++) us[c * SH * P + i] = uv[k][c];
        if (ur) {
            // the ring of the primal-dual step: U tile coordinates (D, D) on
            constexpr int RW = TW + 2, RH = S::TH + 2;
            const int rx = i % P - D, ry = i / P - D;
            if (i < SH * P && rx >= 0 && rx < RW && ry >= 0 && ry < RH)
#pragma unroll
                for (int c = 0; c < 3; c++) ur[(c * RH + ry) * RW + rx] = uv[k][c];
        }
    }
    __syncthreads();
    if (grad) {
        // cuda_laplacian_device (vtv_sr_kernels_unstructured.cu:720-761), no hole gating
        for (int i = tid; i < LH * P; i += NT) {
            const int lx = i % P, ly = i / P;
            if (lx >= LW) continue;
            const int x = tx0 - D + lx, y = ty0 - D + ly;
            const int ui = i + P + 1;  // the pixel in the U tile
#pragma unroll
            for (int c = 0; c < 3; c++) {
                float v = 0.0f;
                if (x >= 0 && x < W && y >= 0 && y < H) {
                    const float *u = us + c * SH * P;
                    float sm = 0.0f, wsum = 0.0f;
                    if (x > 0) { sm += u[ui - 1]; wsum += 1.0f; }
                    if (y > 0) { sm += u[ui - P]; wsum += 1.0f; }
                    if (x < W - 1) { sm += u[ui + 1]; wsum += 1.0f; }
                    if (y < H - 1) { sm += u[ui + P]; wsum += 1.0f; }
                    v = sm / wsum - u[ui];
                }
                ls[c * LH * P + i] = v;
            }
        }
        __syncthreads();
    }
    // D x D box sums.  D <= 2: each box point reads its D x D values directly
    // (as many reads as the separable form, one barrier instead of three; the
    // sums are added in the separable form's order).  Larger D: separably,
    // the row sums of U (rows 1.. of the U tile) and lap U (rows 0..) into sb,
    // then their column sums into sa, which nothing reads any more.
    if constexpr (D <= 2) {
        for (int i = tid; i < BH * P; i += NT) {
            if (i % P >= BW) continue;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                float su = 0.0f, sl = 0.0f;
#pragma unroll
                for (int l = 0; l < D; l++) {
                    float ru = 0.0f, rl = 0.0f;
#pragma unroll
                    for (int k = 0; k < D; k++) {
                        ru += us[c * SH * P + i + (l + 1) * P + k + 1];
                        if (grad) rl += ls[c * LH * P + i + l * P + k];
                    }
                    su += ru;
                    sl += rl;
                }
                bus[c * BH * P + i] = su;
                if (grad) bls[c * BH * P + i] = sl;
            }
        }
        __syncthreads();
    } else {
    constexpr int HR = S::HR, NBT = (BH * P + NT - 1) / NT;
    float *hsu = sb, *hsl = sb + 3 * HR * BW;  // pitch BW
    for (int i = tid; i < HR * P; i += NT) {
        if (i % P >= BW) continue;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float su = 0.0f, sl = 0.0f;
#pragma unroll
            for (int k = 0; k < D; k++) {
                su += us[c * SH * P + i + P + k + 1];
                if (grad) sl += ls[c * LH * P + i + k];
            }
            const int h = i / P * BW + i % P;
            hsu[c * HR * BW + h] = su;
            if (grad) hsl[c * HR * BW + h] = sl;
        }
    }
    __syncthreads();
    float bsum[NBT][6];
#pragma unroll
    for (int k = 0; k < NBT; k++) {
        const int i = tid + k * NT;
#pragma unroll
        for (int c = 0; c < 6; c++) bsum[k][c] = 0.0f;
        if (i >= BH * P || i % P >= BW) continue;
        const int h = i / P * BW + i % P;
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int l = 0; l < D; l++) {
                bsum[k][c] += hsu[c * HR * BW + h + l * BW];
                if (grad) bsum[k][3 + c] += hsl[c * HR * BW + h + l * BW];
            }
    }
    // BU and BL go to sa, which the row sums no longer read: no barrier here
#pragma unroll
    for (int k = 0; k < NBT; k++) {
        const int i = tid + k * NT;
        if (i >= BH * P || i % P >= BW) continue;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            bus[c * BH * P + i] = bsum[k][c];
            if (grad) bls[c * BH * P + i] = bsum[k][3 + c];  // (no room for BL without the gradient term)
        }
    }
    __syncthreads();
    }

    // ---- anchor residuals ----
#pragma unroll
    for (int k = 0; k < NQT; k++) {
        const int i = tid + k * NT;
        if (i >= NQP || i % P >= QW) continue;
        const int b = i + P + 1;  // the point in the box-sum tile
        auto st9 = [&](const float *B, const float(&f)[9]) {
            return f[0] * B[b] + f[1] * B[b + 1] + f[5] * B[b - 1] + f[2] * B[b + P] + f[6] * B[b - P] +
                   f[3] * B[b + P + 1] + f[7] * B[b - P - 1] + f[4] * B[b + P - 1] + f[8] * B[b - P + 1];
        };
#pragma unroll
        for (int c = 0; c < 3; c++) {
            if (grad)
                qs[c * NQP + i] = ci * st9(bus + c * BH * P, cf[k]) + cgd * st9(bls + c * BH * P, cg[k]) - bi[k][c];
            else
                qs[c * NQP + i] = st9(bus + c * BH * P, cf[k]) - bi[k][c];
        }
    }
    __syncthreads();
    // rows: r(x, qy) = sum_k q(x + k, qy), stored with pitch TW for the pixel
    // pass (D <= 2: the pixel pass reads its D x D residuals directly)
    if constexpr (D > 2) {
    for (int i = tid; i < QH * P; i += NT) {
        const int lx = i % P, qy = i / P;
        if (lx >= TW) continue;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float sm = 0.0f;
#pragma unroll
            for (int k = 0; k < D; k++) sm += qs[c * NQP + i + k];
            rs[(c * QH + qy) * TW + lx] = sm;
        }
    }
    __syncthreads();
    }

    // ---- pixel: columns of the row sums, then the step of vtv_sr_ista_step_unstructured ----
    const int lx = tid % TW, ly = tid / TW;
    float acc[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float sm = 0.0f;
        if constexpr (D <= 2) {
#pragma unroll
            for (int l = 0; l < D; l++) {
                float row = 0.0f;
#pragma unroll
                for (int k = 0; k < D; k++) row += qs[c * NQP + (ly + l) * P + lx + k];
                sm += row;
            }
        } else {
#pragma unroll
            for (int l = 0; l < D; l++) sm += rs[(c * QH + ly + l) * TW + lx];
        }
        acc[c] = sm;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float G = 0.0f;
        if (grad) {  // both terms, weighted in the residual
            G = acc[c];
            if (cov != 0.0f) G /= cov;
        } else if (da.alpha > 0.0f) {
            float gi = acc[c];
            if (cov != 0.0f) gi /= cov;
            gi *= da.sigma2;
            G += da.alpha * gi;
        }
        G = fminf(1.0f, fmaxf(-1.0f, G));
        G *= da.prox_factor;
        G += up[c];
        F[c] = fminf(1.0f, fmaxf(0.0f, G));
    }
    // no barrier here: sb is free since the residual phase's barrier, while
    // other threads may still read sa
}

template <int KS, bool GRAD>
__global__ void __launch_bounds__(kStTW * kStTH) k_sr_dataterm_st(int W, int H, int Wc, int Hc, int ntx, int ntiles,
                                                                   DataTermArgs da, const float *__restrict__ U,
                                                                   const float *__restrict__ ops, size_t NA,
                                                                   const float *__restrict__ coverage,
                                                                   float *__restrict__ Gout, float *__restrict__ Uq,
                                                                   float *__restrict__ norm_mask, const int *halt)
{
    if (halt && *halt) return;  // a step enqueued behind a stopping rule that fired
    __shared__ float sa[StDims<KS, GRAD>::AREA];
    __shared__ float sb[StDims<KS, GRAD>::BAREA];
    const int t = xcd_tile(blockIdx.x, ntiles);
    const int tx0 = (t % ntx) * kStTW, ty0 = (t / ntx) * kStTH;
    float F[3], up[3], cov;
    sr_dataterm_tile<KS, GRAD>(W, H, Wc, Hc, tx0, ty0, da, U, ops, NA, coverage, sa, sb, F, up, cov);
    const int px = tx0 + threadIdx.x % kStTW, py = ty0 + threadIdx.x / kStTW;
    if (px >= W || py >= H) return;
    const size_t N = (size_t)W * H, po = (size_t)py * W + px;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        Gout[c * N + po] = F[c];
        Uq[c * N + po] = up[c];
    }
    if (da.alpha > 0.0f || da.beta > 0.0f) norm_mask[po] = cov;
}


// One outer step for dsf >= 2 in one pass: the data term of the tile, then
// the first primal-dual iteration on it (k_pd_step's update, with Uq = U as the
// step sets it, optimization.cu:1765-1767).  U is only read; the new primal
// goes to unew.  With uq_relax (the step's only inner iteration) the
// over-relaxation of vtv_rof.cu:28-45 is fused too: uq_relax <- U + theta (new -
// U).  Gout (nullable) receives F for further inner iterations.  uq_relax and
// norm_mask are nullable: a run of fixed steps reads neither between its steps
// (the next step starts from the new primal, and the mask is the coverage
// again), so only its last step writes them.
// The step on tile t (of the W x H target's ntx-wide tile grid); every
// thread of the block calls it.
template <int KS, bool GRAD>
__device__ __forceinline__ void sr_step_tile(int t, int W, int H, int Wc, int Hc, int ntx, const DataTermArgs &da,
                                             const PdArgs &pa, const float *__restrict__ U,
                                             const float *__restrict__ ops, size_t NA,
                                             const float *__restrict__ coverage, const float *__restrict__ pxin,
                                             const float *__restrict__ pyin, const float *__restrict__ g,
                                             float *__restrict__ pxout, float *__restrict__ pyout,
                                             float *__restrict__ unew, float *__restrict__ uq_relax,
                                             float *__restrict__ Gout, float *__restrict__ norm_mask)
{
    constexpr int TW = kStTW, TH = kStTH, NT = TW * TH;
    constexpr int UW = TW + 2, UH = TH + 2;  // U region [x0-1, x0+TW] x [y0-1, y0+TH]
    constexpr int PW = TW + 1, PH = TH + 1;  // p region [x0-1, x0+TW-1] x [y0-1, y0+TH-1]
    constexpr int NPP = PH * UW;             // p slots with the U region's pitch (x < PW used)
    constexpr int NPT = (NPP + NT - 1) / NT;
    static_assert(6 * NPP <= StDims<KS, GRAD>::BAREA, "primal-dual tiles exceed the data term's second LDS region");
    static_assert(UW == kStTW + 2 && UH == kStTH + 2, "the ring the data term leaves");
    __shared__ float sa[StDims<KS, GRAD>::AREA];
    __shared__ float sb[StDims<KS, GRAD>::BAREA];
    __shared__ float us[3 * UH * UW];  // U on the primal-dual step's ring, left by the data term's staging
    const int tx0 = (t % ntx) * TW, ty0 = (t / ntx) * TH;
    const size_t N = (size_t)W * H;
    const int tid = threadIdx.x;
    // the dual inputs of this thread's p points and the TV weight there, loaded
    // first so that they arrive while the data term runs
    float pa_in[NPT][6], gr[NPT];
#pragma unroll
    for (int k = 0; k < NPT; k++) {
        const int i = tid + k * NT;
        const int x = tx0 - 1 + i % UW, y = ty0 - 1 + i / UW;
        const bool in = i < NPP && i % UW < PW && x >= 0 && x < W && y >= 0 && y < H;
        const size_t o = in ? (size_t)y * W + x : 0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            pa_in[k][c] = in ? pxin[c * N + o] : 0.0f;
            pa_in[k][3 + c] = in ? pyin[c * N + o] : 0.0f;
        }
        gr[k] = (in && g) ? g[o] : 1.0f;
    }
    float F[3], up[3], cov;
    sr_dataterm_tile<KS, GRAD>(W, H, Wc, Hc, tx0, ty0, da, U, ops, NA, coverage, sa, sb, F, up, cov, us);

    // ---- the primal-dual iteration (vtv_rof.cu:118-147, 280-321; cuda_kernels.cu:504-532) ----
    float *ps = sb;  // free once the data term returns (sa may still be read)
#pragma unroll
    for (int k = 0; k < NPT; k++) {
        const int i = tid + k * NT;
        if (i >= NPP || i % UW >= PW) continue;
        const int lx = i % UW, ly = i / UW;
        const int x = tx0 - 1 + lx, y = ty0 - 1 + ly;
        float a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
        if (x >= 0 && x < W && y >= 0 && y < H) {
            float nn = 0.0f;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                // the p region and the U region share the origin (x0-1, y0-1)
                const float *u = us + c * UH * UW;
                const float uc = u[i];
                const float gx = (x < W - 1) ? u[i + 1] - uc : 0.0f;
                const float gy = (y < H - 1) ? u[i + UW] - uc : 0.0f;
                a[c] = pa_in[k][c] + pa.sigma * gx;
                b[c] = pa_in[k][3 + c] + pa.sigma * gy;
                nn += a[c] * a[c] + b[c] * b[c];
            }
            nn = sqrtf(nn);
            if (nn > gr[k]) {
                const float sc = gr[k] / nn;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    a[c] *= sc;
                    b[c] *= sc;
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            ps[c * NPP + i] = a[c];
            ps[(3 + c) * NPP + i] = b[c];
        }
    }
    __syncthreads();
    const int lx = tid % TW, ly = tid / TW;
    const int x = tx0 + lx, y = ty0 + ly;
    if (x >= W || y >= H) return;
    const size_t o = (size_t)y * W + x;
    const int pi = (ly + 1) * UW + lx + 1;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float *pxs = ps + c * NPP, *pys = ps + (3 + c) * NPP;
        const float pxv = pxs[pi], pyv = pys[pi];
        pxout[c * N + o] = pxv;
        pyout[c * N + o] = pyv;
        float step = pxv + pyv;
        if (x > 0) step -= pxs[pi - 1];
        if (y > 0) step -= pys[pi - UW];
        const float un = up[c] + pa.tau * step;
        const float v = (pa.tau * F[c] + pa.lam * un) / (pa.tau + pa.lam);
        unew[c * N + o] = v;
        if (uq_relax) uq_relax[c * N + o] = up[c] + pa.theta * (v - up[c]);
        if (Gout) Gout[c * N + o] = F[c];
    }
    if (norm_mask && (da.alpha > 0.0f || da.beta > 0.0f)) norm_mask[o] = cov;
}

template <int KS, bool GRAD>
__global__ void __launch_bounds__(kStTW * kStTH) k_sr_step_st(int W, int H, int Wc, int Hc, int ntx, int ntiles,
                                                               DataTermArgs da, PdArgs pa, const float *__restrict__ U,
                                                               const float *__restrict__ ops, size_t NA,
                                                               const float *__restrict__ coverage,
                                                               const float *__restrict__ pxin,
                                                               const float *__restrict__ pyin,
                                                               const float *__restrict__ g, float *__restrict__ pxout,
                                                               float *__restrict__ pyout, float *__restrict__ unew,
                                                               float *__restrict__ uq_relax, float *__restrict__ Gout,
                                                               float *__restrict__ norm_mask, const int *halt)
{
    if (halt && *halt) return;  // a step enqueued behind a stopping rule that fired
    sr_step_tile<KS, GRAD>(xcd_tile(blockIdx.x, ntiles), W, H, Wc, Hc, ntx, da, pa, U, ops, NA, coverage, pxin, pyin, g,
                     pxout, pyout, unew, uq_relax, Gout, norm_mask);
}

// Step s of every IBR_color solve of a batch in one launch: block b runs tile
// xcd_tile(j, ntiles) of solve k, (k, j) = work[b], exactly as block j of
// that solve's own k_sr_step_st launch (the same tile, the same arithmetic),
// on the buffers of step parity s & 1; a solve with fewer steps returns.
// Only a solve's last step writes Uq and the norm mask, as in its own run.
template <int KS, bool GRAD>
__global__ void __launch_bounds__(kStTW * kStTH) k_sr_step_st_batch(const SrBatchItem *__restrict__ items,
                                                                     const int2 *__restrict__ work, int s)
{
    const int2 wk = work[blockIdx.x];
    const SrBatchItem &it = items[wk.x];
    if (s >= it.steps) return;
    const int p = s & 1;
    PdArgs pa = it.pa;
    pa.theta = it.theta[s];
    const bool last = s == it.steps - 1;
    sr_step_tile<KS, GRAD>(xcd_tile(wk.y, it.ntiles), it.W, it.H, it.Wc, it.Hc, it.ntx, it.da, pa, it.U[p], it.ops, it.NA,
                     it.coverage, it.PX[p], it.PY[p], it.g, it.PX[p ^ 1], it.PY[p ^ 1], it.U[p ^ 1],
                     last ? it.uq : nullptr, nullptr, last ? it.norm_mask : nullptr);
}

// theta over-relaxation (vtv_rof.cu:28-45): uq <- u + theta (uq - u); u <- uq_old
__global__ void k_overrelax(size_t n3, float theta, float *__restrict__ uq, float *__restrict__ u, const int *halt)
{
    if (halt && *halt) return;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n3) return;
    const float uv = u[k], uqv = uq[k];
    uq[k] = uv + theta * (uqv - uv);
    u[k] = uqv;
}

// E_TV over pixels whose 3x3 neighbourhood is visible, and the ungated Sobel
// of Y(U) for the gradient term (optimization.cu:1505-1534, 1622-1625).
// With prev2 (two 3N buffers, the loop state L saying which one holds the
// previous solution), the same pass also counts the pixels of U that moved
// against it (the count of k_count_moving, block partials at pmove) and
// writes U into the other buffer: if the step's rules keep it, that buffer
// becomes the previous solution (k_sr_loop_step flips L->cur).
__global__ void __launch_bounds__(256) k_sr_energy_tv(int W, int H, const float *__restrict__ U,
                                                      const float *__restrict__ vis, float *__restrict__ ugx,
                                                      float *__restrict__ ugy, double *partials, const int *halt,
                                                      float *__restrict__ prev2, const SrLoop *L, float thr,
                                                      double *pmove)
{
    if (halt && *halt) return;
    __shared__ double red[16], red_m[16];
    const size_t N = (size_t)W * H;
    const int cur = prev2 ? L->cur : 0;
    const float *prev = prev2 ? prev2 + (size_t)cur * 3 * N : nullptr;
    float *next = prev2 ? prev2 + (size_t)(cur ^ 1) * 3 * N : nullptr;
    double v[1] = {0.0}, vm[1] = {0.0};
    for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < N; o += (size_t)gridDim.x * blockDim.x) {
        if (prev2) {
            double mx = 0.0;
            for (int c = 0; c < 3; c++) {
                const float u = U[c * N + o];
                mx = fmax(mx, fabs((double)u - (double)prev[c * N + o]));
                next[c * N + o] = u;
            }
            if (!(mx < (double)thr)) vm[0] += 1.0;
        }
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
    if (prev2) block_partials<1>(vm, pmove, red_m);
}

// E_int and E_grad over every (view, pixel) with gather validity
// (optimization.cu:1549-1606, 1611-1704).
template <int KS>
__global__ void __launch_bounds__(256) k_sr_energy_views(int V, int W, int H, int w, int h,
                                                         const float *__restrict__ tx, const float *__restrict__ ty,
                                                         const float *__restrict__ omega, const float *__restrict__ views,
                                                         const float *__restrict__ vgrad, const float *__restrict__ U,
                                                         const float *__restrict__ ugx, const float *__restrict__ ugy,
                                                         float sigma2, int gw_type, double *partials,
                                                         const int *halt)
{
    if (halt && *halt) return;
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
                                                      double *partials, const int *halt)
{
    if (halt && *halt) return;
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

// The energy's block partials: E_TV at partials[0, 2048), E_int and E_grad
// at partials + 2048 (2 x up to 2048 blocks).
constexpr size_t kViewPartials = 2048;

struct EnergyCoef {
    double lambda, alpha, beta;
};

// The energy of the driver, E = lambda E_TV + dt_alpha E_int + dt_beta E_grad
// (optimization.cu:1710), from the block partials; the driver keeps it in
// float (super-resolution.cpp:181, 230).  The sum in the host's order,
// without contraction into FMAs.
__device__ __forceinline__ float sr_loop_energy(const double (&tv)[1], const double (&dv)[2], const EnergyCoef &c)
{
    return (float)__dadd_rn(__dadd_rn(__dmul_rn(c.lambda, tv[0]), __dmul_rn(c.alpha, dv[0])),
                            __dmul_rn(c.beta, dv[1]));
}

// The pass's starting energy (super-resolution.cpp:207-209), which also opens
// the pass.
__global__ void __launch_bounds__(kFinalizeThreads) k_sr_loop_begin(SrLoop *L, const double *__restrict__ ptv, int Gtv,
                                                                  const double *__restrict__ pviews, int Gv,
                                                                  EnergyCoef c)
{
    __shared__ double red[2 * 16];
    double tv[1], dv[2];
    reduce_partials<1>(ptv, Gtv, tv, red);
    reduce_partials<2>(pviews, Gv, dv, red);
    if (threadIdx.x != 0) return;
    const float E = sr_loop_energy(tv, dv, c);
    L->parts[0] = tv[0];
    L->parts[1] = dv[0];
    L->parts[2] = dv[1];
    L->energy = E;
    L->halt = 0;
    L->steps = 0;
    L->stop = LFRV_SR_STOP_MAXIT;
    L->prev_energy = E;
}

// The rules after a FISTA step, in the driver's order: E < 1e-4 stops;
// |(E - E_prev) / E| < 1e-4 stops; else E_prev = E (super-resolution.cpp:
// 230-244), and fewer than pct of the pixels moved against the previous
// step's solution stops (:246-257), on the count the step's energy pass took
// (k_sr_energy_tv with prev2; its block partials at pmove); a step the rules
// keep becomes the previous solution.
__global__ void __launch_bounds__(kFinalizeThreads) k_sr_loop_step(SrLoop *L, const double *__restrict__ ptv, int Gtv,
                                                                 const double *__restrict__ pviews, int Gv,
                                                                 EnergyCoef c, float eps,
                                                                 const double *__restrict__ pmove, int Gm,
                                                                 float npix, float pct)
{
    if (L->halt) return;
    __shared__ double red[2 * 16];
    double tv[1], dv[2], m[1];
    reduce_partials<1>(ptv, Gtv, tv, red);
    reduce_partials<2>(pviews, Gv, dv, red);
    reduce_partials<1>(pmove, Gm, m, red);
    if (threadIdx.x != 0) return;
    const float E = sr_loop_energy(tv, dv, c);
    L->parts[0] = tv[0];
    L->parts[1] = dv[0];
    L->parts[2] = dv[1];
    L->energy = E;
    L->steps += 1;
    if (E < eps) {
        L->halt = 1;
        L->stop = LFRV_SR_STOP_ENERGY;
    } else if (fabsf((E - L->prev_energy) / E) < eps) {
        L->halt = 1;
        L->stop = LFRV_SR_STOP_RELATIVE;
    } else {
        L->prev_energy = E;
        if ((float)m[0] / npix < pct) {
            L->halt = 1;
            L->stop = LFRV_SR_STOP_MOVING;
        } else {
            L->cur ^= 1;  // prev <- U (super-resolution.cpp:255): the energy pass wrote it
        }
    }
}

}  // namespace

namespace {

DataTermArgs data_term_args(const lfrv_ctx *ctx)
{
    DataTermArgs da;
    da.alpha = ctx->p.dt_alpha;
    da.beta = ctx->p.dt_beta;
    da.sigma2 = ctx->p.sigma_sensor * ctx->p.sigma_sensor;
    da.grad_step = ctx->p.gradient_step;
    da.dsf2 = (float)(ctx->dsf * ctx->dsf);
    const float L = (float)(1.0 / (double)ctx->p.lambda);  // optimization.cu:1741
    da.prox_factor = (float)(-1.0 / ((double)ctx->p.lambda * (double)L));
    da.gw_type = ctx->p.gw_type;
    return da;
}

// FISTA weights (optimization.cu:1744-1746): alpha_new is a cuflt (float)
// there, theta the float of the double quotient.  Returns theta of the step
// that starts at alpha; `an` receives alpha_new.
float fista_theta(double alpha, float &an)
{
    an = (float)(0.5 * (1.0 + std::sqrt(1.0 + 4.0 * std::pow(alpha, 2.0))));
    return (float)((alpha - 1.0) / (double)an);
}

// The primal-dual constants of the step fused with the data term
// (vtv_rof.cu:118-147), theta left for the caller.
PdArgs pd_args(const lfrv_ctx *ctx)
{
    PdArgs pa;
    pa.tau = pa.sigma = (float)(0.3 / std::sqrt(8.0));
    const float Lp = (float)(1.0 / (double)ctx->p.lambda);
    pa.lam = (float)(1.0 / (double)Lp);
    pa.theta = 0.0f;
    return pa;
}

// The assembled anchor-grid operator of the data term (dsf >= 2), rebuilt
// when the weights or the entries changed.
lfrv_status ensure_sr_ops(lfrv_ctx *ctx)
{
    if (ctx->sr_ops_epoch == ctx->epoch) return LFRV_OK;
    const TileGeom g = tile_geom(ctx);
    const TileEntries te = tile_entries(ctx);
    const int Wc = ctx->W + ctx->ks - 1, Hc = ctx->H + ctx->ks - 1;
    const size_t NA = (size_t)Wc * Hc;
    const int grad = ctx->p.dt_beta > 0.0f ? 1 : 0;
    const DataTermArgs da = data_term_args(ctx);
    TileGeom ga = g;
    ga.ntx = (Wc + g.tw - 1) / g.tw;
    ga.ntiles = ga.ntx * ((Hc + g.th - 1) / g.th);
    const int wh = g.th + 1, ncw = (g.tw + 1) * wh;
    const int cap = std::min(1024, kStagePer * ctx->pull_bs);
    const size_t lds = sizeof(int) * (2 * wh + 2) + sizeof(float) * std::max<size_t>(44 * (size_t)ncw, 9 * (size_t)cap);
    TimedRegion tr(ctx, 5);
#define AS(K) k_sr_assemble<K><<<ga.ntiles, ctx->pull_bs, lds, ctx->stream>>>(ga, te, cap, grad, ctx->p.gw_type != 0, dt_intensity_weight(da), dt_gradient_weight(da), NA, ctx->sr_ops)
    switch (ctx->ks) {
    case 3: AS(3); break;
    case 4: AS(4); break;
    case 5: AS(5); break;
    default: return fail(ctx, LFRV_E_ARG, "dsf out of range");
    }
#undef AS
    LFRV_LAUNCH(ctx, "data-term assembly");
    ctx->sr_ops_epoch = ctx->epoch;
    return LFRV_OK;
}

// One FISTA step on the stream; with `halt` every launch of the step is a
// no-op once a stopping rule has fired on the device.  A step that is not the
// last of a fixed run (last = false) leaves Uq and the norm mask to the run's
// last step where one pass does the whole step (dsf >= 2, one inner iteration).
lfrv_status enqueue_fista(lfrv_ctx *ctx, const int *halt, bool last = true)
{
    if (!ctx->prepared) return fail(ctx, LFRV_E_STATE, "FISTA step before lfrv_prepare");
    lfrv_status st = run_refresh_entry_weights(ctx, 1.0f);
    if (st != LFRV_OK) return st;
    const TileGeom g = tile_geom(ctx);
    const TileEntries te = tile_entries(ctx);
    const DataTermArgs da = data_term_args(ctx);
    // the over-relaxation is fused into the last primal-dual step
    float an;
    const float theta = fista_theta(ctx->fista_alpha, an);
    const int inner = ctx->p.inner_iterations;
    bool pd_done = false;  // the first primal-dual iteration ran with the data term
    const int HAL = ctx->ks - 1;
    if (ctx->ks == 2) {
        const size_t up = (size_t)(ctx->tile_w + 2 * HAL + 2) * (ctx->tile_h + 2 * HAL + 2);
        const size_t lpp = (size_t)(ctx->tile_w + 2 * HAL) * (ctx->tile_h + 2 * HAL);
        const size_t tiles = sizeof(float) * (3 * up + 3 * lpp);
        const int cap = pull_cap(ctx, 6, tiles, ctx->pull_budget, PullKind::Cell);
        const size_t lds = tiles + pull_lds(ctx, 6, cap, PullKind::Cell);
        const KernelEvents ev = kernel_events(ctx, 3);
        hipExtLaunchKernelGGL(k_sr_dataterm<2>, dim3(ctx->ntiles), dim3(ctx->pull_bs), lds, ctx->stream, ev.a, ev.b, 0,
                              g, te, cap, ctx->n, da, (const float *)ctx->U, (const float *)ctx->views,
                              (const float *)ctx->view_lap, (const float *)ctx->coverage, ctx->G, ctx->Uq,
                              ctx->norm_mask, halt);
        record_kernel_events(ctx, 3, ev);
        LFRV_LAUNCH(ctx, "data term");
    } else {
        // assembled operator on the anchor grid, rebuilt when the weights or
        // the entries changed; the step itself is a stencil
        const int Wc = ctx->W + ctx->ks - 1, Hc = ctx->H + ctx->ks - 1;
        const size_t NA = (size_t)Wc * Hc;
        st = ensure_sr_ops(ctx);
        if (st != LFRV_OK) return st;
        const int ntx = (ctx->W + kStTW - 1) / kStTW, ntiles = ntx * ((ctx->H + kStTH - 1) / kStTH);
        const KernelEvents ev = kernel_events(ctx, 3);  // stamped by the dispatch itself
        if (inner > 0) {
            // the data term and the first primal-dual iteration in one pass: U is
            // read, the new primal goes to U2 (with one inner iteration the
            // over-relaxed one to Uq), the duals to P2X / P2Y
            PdArgs pa = pd_args(ctx);
            pa.theta = theta;
            const float *gw = ctx->has_reg_g ? ctx->reg_g : nullptr;
            float *uq_relax = (inner == 1 && last) ? ctx->Uq : nullptr;
            float *gout = inner > 1 ? ctx->G : nullptr;
            float *nmask = (inner > 1 || last) ? ctx->norm_mask : nullptr;
#define ST1(K, GR) hipExtLaunchKernelGGL(k_sr_step_st<K, GR>, dim3(ntiles), dim3(kStTW * kStTH), 0, ctx->stream, ev.a, ev.b, 0, ctx->W, ctx->H, Wc, Hc, ntx, ntiles, da, pa, (const float *)ctx->U, (const float *)ctx->sr_ops, NA, (const float *)ctx->coverage, (const float *)ctx->PX, (const float *)ctx->PY, gw, ctx->P2X, ctx->P2Y, ctx->U2, uq_relax, gout, nmask, halt)
#define ST(K) do { if (da.beta > 0.0f) ST1(K, true); else ST1(K, false); } while (0)
            switch (ctx->ks) {
            case 3: ST(3); break;
            case 4: ST(4); break;
            case 5: ST(5); break;
            default: return fail(ctx, LFRV_E_ARG, "dsf out of range");
            }
#undef ST
#undef ST1
            record_kernel_events(ctx, 3, ev);
            LFRV_LAUNCH(ctx, "data term + primal-dual step");
            std::swap(ctx->PX, ctx->P2X);
            std::swap(ctx->PY, ctx->P2Y);
            if (inner == 1)
                std::swap(ctx->U, ctx->U2);  // U the new primal, Uq the over-relaxed one
            else
                std::swap(ctx->Uq, ctx->U2);  // the remaining iterations start from Uq
            pd_done = true;
        } else {
#define DT1(K, GR) hipExtLaunchKernelGGL(k_sr_dataterm_st<K, GR>, dim3(ntiles), dim3(kStTW * kStTH), 0, ctx->stream, ev.a, ev.b, 0, ctx->W, ctx->H, Wc, Hc, ntx, ntiles, da, (const float *)ctx->U, (const float *)ctx->sr_ops, NA, (const float *)ctx->coverage, ctx->G, ctx->Uq, ctx->norm_mask, halt)
#define DT(K) do { if (da.beta > 0.0f) DT1(K, true); else DT1(K, false); } while (0)
            switch (ctx->ks) {
            case 3: DT(3); break;
            case 4: DT(4); break;
            case 5: DT(5); break;
            default: return fail(ctx, LFRV_E_ARG, "dsf out of range");
            }
#undef DT
#undef DT1
            record_kernel_events(ctx, 3, ev);
            LFRV_LAUNCH(ctx, "data term");
        }
    }
    if (pd_done) {
        if (inner > 1) {
            st = run_pd_steps(ctx, inner - 1, ctx->G, true, theta, halt);
            if (st != LFRV_OK) return st;
        }
    } else if (inner > 0) {
        st = run_pd_steps(ctx, inner, ctx->G, true, theta, halt);
        if (st != LFRV_OK) return st;
    } else {
        k_overrelax<<<blocks_for(3 * ctx->N), kBlock, 0, ctx->stream>>>(3 * ctx->N, theta, ctx->Uq, ctx->U, halt);
        LFRV_LAUNCH(ctx, "overrelaxation");
    }
    ctx->fista_alpha = an;
    return LFRV_OK;
}

// The energy's block partials (tv at partials[0], views at
// partials + kViewPartials); ugx / ugy in tmp3N.
lfrv_status enqueue_sr_energy(lfrv_ctx *ctx, const int *halt, float *prev2 = nullptr, float thr = 0.0f,
                              double *pmove = nullptr)
{
    const size_t N = ctx->N;
    float *ugx = ctx->tmp3N, *ugy = ctx->tmp3N + N;
    k_sr_energy_tv<<<rblocks(N), 256, 0, ctx->stream>>>(ctx->W, ctx->H, ctx->U, ctx->visibility, ugx, ugy,
                                                        ctx->partials, halt, prev2, ctx->sr_loop, thr, pmove);
    const size_t tot = (size_t)ctx->V * ctx->n;
    const float s2 = ctx->p.sigma_sensor * ctx->p.sigma_sensor;
    double *pv = ctx->partials + kViewPartials;
#define EV(K) k_sr_energy_views<K><<<rblocks(tot), 256, 0, ctx->stream>>>(ctx->V, ctx->W, ctx->H, ctx->w, ctx->h, ctx->tau_x, ctx->tau_y, ctx->omega, ctx->views, ctx->view_grad, ctx->U, ugx, ugy, s2, ctx->p.gw_type, pv, halt)
    switch (ctx->ks) {
    case 2: EV(2); break;
    case 3: EV(3); break;
    case 4: EV(4); break;
    case 5: EV(5); break;
    default: return fail(ctx, LFRV_E_ARG, "dsf out of range");
    }
#undef EV
    LFRV_LAUNCH(ctx, "sr energy");
    return LFRV_OK;
}

EnergyCoef energy_coef(const lfrv_ctx *ctx)
{
    return EnergyCoef{(double)ctx->p.lambda, (double)ctx->p.dt_alpha, (double)ctx->p.dt_beta};
}

}  // namespace

lfrv_status run_fista(lfrv_ctx *ctx, bool last)
{
    return enqueue_fista(ctx, nullptr, last);
}

bool sr_batchable(const lfrv_ctx *ctx)
{
    return ctx->p.mode == LFRV_MODE_COLOR_SR && ctx->ks >= 3 && ctx->ks <= 5 && ctx->p.inner_iterations == 1;
}

lfrv_status sr_batch_begin(lfrv_ctx *ctx, int steps, SrBatchItem *it, float *theta_host)
{
    if (!ctx->prepared) return fail(ctx, LFRV_E_STATE, "FISTA step before lfrv_prepare");
    if (!sr_batchable(ctx)) return fail(ctx, LFRV_E_ARG, "batched FISTA steps need dsf >= 2 and one inner iteration");
    lfrv_status st = run_refresh_entry_weights(ctx, 1.0f);
    if (st == LFRV_OK) st = ensure_sr_ops(ctx);
    if (st != LFRV_OK) return st;
    const int Wc = ctx->W + ctx->ks - 1, Hc = ctx->H + ctx->ks - 1;
    it->W = ctx->W;
    it->H = ctx->H;
    it->Wc = Wc;
    it->Hc = Hc;
    it->ntx = (ctx->W + kStTW - 1) / kStTW;
    it->ntiles = it->ntx * ((ctx->H + kStTH - 1) / kStTH);
    it->steps = steps;
    it->da = data_term_args(ctx);
    it->pa = pd_args(ctx);
    it->theta = nullptr;
    it->ops = ctx->sr_ops;
    it->NA = (size_t)Wc * Hc;
    it->coverage = ctx->coverage;
    it->g = ctx->has_reg_g ? ctx->reg_g : nullptr;
    it->U[0] = ctx->U;
    it->U[1] = ctx->U2;
    it->PX[0] = ctx->PX;
    it->PX[1] = ctx->P2X;
    it->PY[0] = ctx->PY;
    it->PY[1] = ctx->P2Y;
    it->uq = ctx->Uq;
    it->norm_mask = ctx->norm_mask;
    double a = ctx->fista_alpha;
    for (int k = 0; k < steps; k++) {
        float an;
        theta_host[k] = fista_theta(a, an);
        a = an;
    }
    ctx->u_zero = false;
    return LFRV_OK;
}

void sr_batch_end(lfrv_ctx *ctx, int steps)
{
    double a = ctx->fista_alpha;
    for (int k = 0; k < steps; k++) {
        float an;
        (void)fista_theta(a, an);
        a = an;
    }
    ctx->fista_alpha = a;
    if (steps & 1) {  // the solve's own steps swap after every step
        std::swap(ctx->U, ctx->U2);
        std::swap(ctx->PX, ctx->P2X);
        std::swap(ctx->PY, ctx->P2Y);
    }
}

hipError_t launch_sr_batch_step(int ks, bool grad, const SrBatchItem *items, const int2 *work, int nblocks, int s,
                                hipStream_t stream)
{
#define SB(K, G) k_sr_step_st_batch<K, G><<<nblocks, kStTW * kStTH, 0, stream>>>(items, work, s)
    switch (ks * 2 + (grad ? 1 : 0)) {
    case 6: SB(3, false); break;
    case 7: SB(3, true); break;
    case 8: SB(4, false); break;
    case 9: SB(4, true); break;
    case 10: SB(5, false); break;
    case 11: SB(5, true); break;
    default: return hipErrorInvalidValue;
    }
#undef SB
    return hipGetLastError();
}

lfrv_status run_sr_energy(lfrv_ctx *ctx, double *e, double *parts)
{
    lfrv_status st = enqueue_sr_energy(ctx, nullptr);
    if (st != LFRV_OK) return st;
    const size_t tot = (size_t)ctx->V * ctx->n;
    k_reduce_out<1><<<1, kFinalizeThreads, 0, ctx->stream>>>(ctx->partials, (int)rblocks(ctx->N), ctx->scal + 8);
    k_reduce_out<2><<<1, kFinalizeThreads, 0, ctx->stream>>>(ctx->partials + kViewPartials, (int)rblocks(tot),
                                                             ctx->scal + 9);
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
    k_count_moving<<<rblocks(ctx->N), 256, 0, ctx->stream>>>(ctx->N, ctx->U, prev_dev, max_diff, ctx->partials,
                                                             nullptr);
    k_reduce_out<1><<<1, kFinalizeThreads, 0, ctx->stream>>>(ctx->partials, (int)rblocks(ctx->N), ctx->scal + 12);
    LFRV_LAUNCH(ctx, "count moving");
    LFRV_HIP(ctx, hipMemcpyAsync(ctx->scal_host + 12, ctx->scal + 12, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *moving = (int64_t)std::llround(ctx->scal_host[12]);
    return LFRV_OK;
}

// The re-weighted least-squares loop of super-resolution.cpp:177-300.  Each
// FISTA step is followed on the stream by its energy pass (which also counts
// the moving pixels against the previous step's solution and writes U into
// the spare one of the two previous-step buffers) and one rule kernel (the
// energy rules, then the moving rule; when the step is kept it flips
// SrLoop::cur, so the spare buffer becomes the previous solution without a
// copy); once a rule fires, the launches enqueued after it return at once.  The host enqueues kSrPollSteps steps at
// a time and reads the loop state of the chunk before the one it has just
// enqueued, so the device runs ahead of the poll.
lfrv_status run_sr_solve(lfrv_ctx *ctx, int niter, int nmeta, lfrv_sr_stats *out)
{
    constexpr int kSrPollSteps = 8;
    const float EPSILON_TOL = 1e-4f, MAX_DIFF_THRESHOLD = 1e-3f, PERCENT_MOVING = 0.01f;
    const auto t0 = std::chrono::steady_clock::now();
    const size_t N = ctx->N, n3 = 3 * N;
    // the previous step's solution (two buffers, SrLoop::cur picks) and the previous pass's
    if (!ctx->sr_prev) LFRV_HIP(ctx, hipMalloc((void **)&ctx->sr_prev, 3 * n3 * sizeof(float)));
    float *prev2 = ctx->sr_prev, *prev_meta = ctx->sr_prev + 2 * n3;
    LFRV_HIP(ctx, hipMemcpyAsync(prev2, ctx->U, n3 * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    LFRV_HIP(ctx, hipMemsetAsync(&ctx->sr_loop->cur, 0, sizeof(int), ctx->stream));
    LFRV_HIP(ctx, hipMemcpyAsync(prev_meta, ctx->U, n3 * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    const int *halt = &ctx->sr_loop->halt;
    const EnergyCoef ec = energy_coef(ctx);
    const int Gtv = (int)rblocks(N), Gv = (int)rblocks((size_t)ctx->V * ctx->n), Gm = (int)rblocks(N);
    const float npix = (float)N;
    // the moving count's block partials, after the energy's (E_TV: 2048, views: 2 x 2048)
    double *pmove = ctx->partials + 3 * kViewPartials;
    if (3 * kViewPartials + (size_t)Gm > ctx->partials_cap) return fail(ctx, LFRV_E_STATE, "partials buffer too small");
    lfrv_sr_stats s{};
    float energy = 0.0f, previous_energy_reweight = 0.0f;
    int r;
    for (r = 0; r < nmeta; r++) {
        lfrv_status st = run_compute_weights(ctx);
        if (st == LFRV_OK) st = run_regularizer_weight(ctx);
        if (st == LFRV_OK) st = enqueue_sr_energy(ctx, nullptr);
        if (st != LFRV_OK) return st;
        k_sr_loop_begin<<<1, kFinalizeThreads, 0, ctx->stream>>>(ctx->sr_loop, ctx->partials, Gtv,
                                                                 ctx->partials + kViewPartials, Gv, ec);
        LFRV_LAUNCH(ctx, "sr loop begin");
        // alpha and the ping-pong buffers after each enqueued step, to undo the
        // host-side effects of the steps the device skipped
        struct StepState {
            double alpha;
            float *U, *Uq, *U2, *PX, *PY, *P2X, *P2Y;
        };
        auto snap = [ctx]() {
            return StepState{ctx->fista_alpha, ctx->U, ctx->Uq, ctx->U2, ctx->PX, ctx->PY, ctx->P2X, ctx->P2Y};
        };
        std::vector<StepState> states{snap()};
        int enq = 0, polls = 0;
        bool halted = false;
        while (enq < niter && !halted) {
            const int chunk = std::min(kSrPollSteps, niter - enq);
            for (int j = 0; j < chunk; j++) {
                st = enqueue_fista(ctx, halt);
                // the energy pass also takes the moving count (rblocks(N) == Gm block partials)
                if (st == LFRV_OK) st = enqueue_sr_energy(ctx, halt, prev2, MAX_DIFF_THRESHOLD, pmove);
                if (st != LFRV_OK) return st;
                states.push_back(snap());
                k_sr_loop_step<<<1, kFinalizeThreads, 0, ctx->stream>>>(
                    ctx->sr_loop, ctx->partials, Gtv, ctx->partials + kViewPartials, Gv, ec, EPSILON_TOL, pmove, Gm,
                    npix, PERCENT_MOVING);
                LFRV_LAUNCH(ctx, "sr loop step");
            }
            enq += chunk;
            const int slot = polls & 1;
            LFRV_HIP(ctx, hipMemcpyAsync(ctx->sr_loop_host + slot, ctx->sr_loop, sizeof(SrLoop), hipMemcpyDeviceToHost,
                                         ctx->stream));
            LFRV_HIP(ctx, hipEventRecord(ctx->sr_ev[slot], ctx->stream));
            if (++polls >= 2) {
                LFRV_HIP(ctx, hipEventSynchronize(ctx->sr_ev[slot ^ 1]));
                halted = ctx->sr_loop_host[slot ^ 1].halt != 0;
            }
        }
        LFRV_HIP(ctx, hipMemcpyAsync(ctx->sr_loop_host + 2, ctx->sr_loop, sizeof(SrLoop), hipMemcpyDeviceToHost,
                                     ctx->stream));
        LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const SrLoop L = ctx->sr_loop_host[2];
        if (L.steps < 0 || L.steps > enq) return fail(ctx, LFRV_E_STATE, "FISTA loop state out of range");
        const StepState &ss = states[L.steps];
        ctx->fista_alpha = ss.alpha;
        ctx->U = ss.U;
        ctx->Uq = ss.Uq;
        ctx->U2 = ss.U2;
        ctx->PX = ss.PX;
        ctx->PY = ss.PY;
        ctx->P2X = ss.P2X;
        ctx->P2Y = ss.P2Y;
        energy = L.energy;
        s.fista_steps += L.steps;
        s.last_stop = L.stop;
        // the re-weighting rules (super-resolution.cpp:272-297)
        if (energy < EPSILON_TOL) {
            if (previous_energy_reweight < EPSILON_TOL) {
                s.converged = 1;
                break;
            }
            previous_energy_reweight = energy;
            continue;
        }
        int64_t moving = 0;
        st = run_count_moving(ctx, prev_meta, MAX_DIFF_THRESHOLD, &moving);
        if (st != LFRV_OK) return st;
        if ((float)moving / npix < PERCENT_MOVING) {
            s.converged = 1;
            break;
        }
        LFRV_HIP(ctx, hipMemcpyAsync(prev_meta, ctx->U, n3 * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    }
    LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s.meta_iterations = r < nmeta ? r + 1 : nmeta;
    s.energy = energy;
    s.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (out) *out = s;
    return LFRV_OK;
}

}  // namespace lfrv
