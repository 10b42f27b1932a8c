// Shared pieces of the tile (binned splat) kernels.
//
// Splat structure.  Every cell-valid source pixel (view v, pixel m) is an
// entry; its footprint covers the ks x ks target pixels x0..x0+ks-1,
// y0..y0+ks-1 (set_B_bilinear, gradientIBR_kernels.cuh:447-507).  Entries are
// sorted by their top-left tap (x0, y0) into a CSR over an extended cell grid
// of (W+ks-1) x (H+ks-1) cells, cell (cx, cy) holding x0 = cx-(ks-1),
// y0 = cy-(ks-1); inside a cell entries are ordered by source index, so every
// sum below has a fixed order.
//
// A splat S^T v is then a pull: target pixel x collects the entries of the
// ks x ks cells cx in [x, x+ks-1] with tap index j = ks-1-(cx-x).  A
// workgroup owns a tile of target pixels; phase 1 computes each entry's
// values once (into LDS), phase 2 lets each thread pull its pixel's cells.
// No atomics, one store per target pixel.
#pragma once

#include "lfrv_internal.hpp"

#include <algorithm>

namespace lfrv {

// Target tiling and the extended cell grid.
struct TileGeom {
    int W, H, tw, th, ntx, ntiles;
    int Wc, Hc;  // cell grid: W + ks - 1, H + ks - 1
};

// Entry arrays sorted by cell: entries cell_start[c] .. cell_start[c+1]-1 of
// cell c; src = view * n + m; (tx, ty) = tau; w = omega of the source pixel.
struct TileEntries {
    const int *cell_start;
    const int *src;
    const float *tx, *ty, *w;
    const float *omega;  // per-source weights (for pulls that read omega[src] instead of w)
    const float *pay;    // per-entry payload planes (pay[q * pay_stride + e]), entry order
    size_t pay_stride;
    int active;          // 0: the tile kernels see no entries
};

TileGeom tile_geom(const lfrv_ctx *ctx);
TileEntries tile_entries(const lfrv_ctx *ctx);

// The three pulls (device side below): the cell-owner pull, the box pull
// (ks >= 3, plain weights) and the slot pull.  Kernels that only splat use
// splat_pull (cell-owner for ks = 2, box for larger footprints; when a value
// sums squared weights, cell-owner up to ks = 3 and slot beyond); kernels
// whose values gather through the
// footprint use any_pull (cell-owner for ks <= 3, slot beyond).
enum class PullKind { Cell, Box, Slot };
inline PullKind splat_kind(const lfrv_ctx *ctx, bool sq)
{
    return ctx->ks <= 2 ? PullKind::Cell : (!sq ? PullKind::Box : (ctx->ks <= 3 ? PullKind::Cell : PullKind::Slot));
}
inline PullKind walk_kind(const lfrv_ctx *ctx) { return ctx->ks <= 3 ? PullKind::Cell : PullKind::Slot; }
// Staged-entry (Cell, Box) or slot (Slot) capacity within `budget` bytes of
// LDS, `extra` of them taken by the kernel's own tiles; the pull's LDS bytes.
// `npay`: payload floats staged per entry besides tau and omega (box pull).
int pull_cap(const lfrv_ctx *ctx, int nv, size_t extra, size_t budget, PullKind kind, int npay = 0);
size_t pull_lds(const lfrv_ctx *ctx, int nv, int cap, PullKind kind, int npay = 0);

lfrv_status run_rhs_and_diag(lfrv_ctx *ctx, double alpha, double c0, double beta, bool use_views,
                             bool use_precond, const float *u0);
lfrv_status run_gauss_smooth(lfrv_ctx *ctx, const float *in, float *out, float *tmp);
lfrv_status run_view_stencils(lfrv_ctx *ctx, hipStream_t s);

// ---------------------------------------------------------------------------
// Device side of the pull
// ---------------------------------------------------------------------------

// In-place exclusive scan of a[0..n) in LDS by the whole block; a[n] = total.
__device__ __forceinline__ void block_exclusive_scan(int *a, int n, int *tmp /* blockDim ints */)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    const int per = (n + nt - 1) / nt;
    const int b0 = tid * per, b1 = min(b0 + per, n);
    int s = 0;
    for (int i = b0; i < b1; i++) s += a[i];
    tmp[tid] = s;
    __syncthreads();
    for (int off = 1; off < nt; off <<= 1) {
        const int v = (tid >= off) ? tmp[tid - off] : 0;
        __syncthreads();
        tmp[tid] += v;
        __syncthreads();
    }
    int run = tmp[tid] - s;
    for (int i = b0; i < b1; i++) {
        const int c = a[i];
        a[i] = run;
        run += c;
    }
    if (tid == nt - 1) a[n] = tmp[nt - 1];
    __syncthreads();
}

// LDS words the slot pull needs besides the slots: lstart (ncw + 1), row
// bases (wh), scan scratch (blockDim).
__host__ __device__ __forceinline__ int pull_window_w(int tw, int ks) { return tw + ks - 1; }
__host__ __device__ __forceinline__ int pull_meta_ints(int tw, int th, int ks, int bs)
{
    return (tw + ks - 1) * (th + ks - 1) + 1 + (th + ks - 1) + bs;
}

// acc += a v, with a squared for the values whose bit is set in SQ (the
// Jacobi diagonal sum_e a_e^2 w_e).
template <int NV, int SQ>
__device__ __forceinline__ void tap_accumulate(float (&acc)[NV], float a, const float (&v)[NV])
{
#pragma unroll
    for (int q = 0; q < NV; q++) acc[q] += (((SQ >> q) & 1) ? a * a : a) * v[q];
}

// The two anchor weights per axis: ax = ((1-f)/sx, f/sx), as d_0 / sx and
// d_dsf / sx of footprint<KS> (same arithmetic).  False when every in-image
// tap has zero weight.
template <int KS>
__device__ __forceinline__ bool box_weights(float tx, float ty, int W, int H, int &x0, int &y0, float (&ax)[2],
                                            float (&ay)[2])
{
    constexpr int D = KS - 1;
    constexpr float hd = 0.5f * (float)D;
    const float px = tx - 0.5f, py = ty - 0.5f;
    x0 = (int)floorf(px - hd + 0.5f);
    y0 = (int)floorf(py - hd + 0.5f);
    // end taps as footprint<KS>; the D - 1 interior taps weigh 1 each
    const float fx0 = ((float)x0 + 0.5f) - (px - hd), fx1 = (px + hd) - ((float)(x0 + D) - 0.5f);
    const float fy0 = ((float)y0 + 0.5f) - (py - hd), fy1 = (py + hd) - ((float)(y0 + D) - 0.5f);
    const int nix = max(0, min(x0 + D - 1, W - 1) - max(x0 + 1, 0) + 1);  // interior taps inside the image
    const int niy = max(0, min(y0 + D - 1, H - 1) - max(y0 + 1, 0) + 1);
    const float sx = (x0 >= 0 && x0 <= W - 1 ? fx0 : 0.0f) + (float)nix + (x0 + D >= 0 && x0 + D <= W - 1 ? fx1 : 0.0f);
    const float sy = (y0 >= 0 && y0 <= H - 1 ? fy0 : 0.0f) + (float)niy + (y0 + D >= 0 && y0 + D <= H - 1 ? fy1 : 0.0f);
    if (sx * sy == 0.0f) return false;
    const float isx = 1.0f / sx, isy = 1.0f / sy;
    ax[0] = fx0 * isx;
    ax[1] = fx1 * isx;
    ay[0] = fy0 * isy;
    ay[1] = fy1 * isy;
    return true;
}

// Slot pull (ks >= 4: few entries per cell).  `vals(e, x0, y0, wx, wy, om, v)`
// produces the NV values of entry e (om = its omega); `epi(o, x, y, acc)`
// receives the pulled sums of pixel (x, y).  Phase 1 stages the tile's entries
// and their values in LDS slots (coalesced over the window rows); phase 2 lets
// each thread pull its pixel's cells.  Slots beyond `cap` are not cached:
// phase 2 recomputes their values.
template <int KS, int NV, int SQ, class Vals, class Epi>
__device__ __forceinline__ void tile_pull(const TileGeom &g, const TileEntries &te, int t, int *meta,
                                          float *slots, int cap, Vals vals, Epi epi)
{
    const int tx0 = (t % g.ntx) * g.tw, ty0 = (t / g.ntx) * g.th;
    const int ww = g.tw + KS - 1, wh = g.th + KS - 1, ncw = ww * wh;
    int *lstart = meta;              // ncw + 1
    int *grow = meta + ncw + 1;      // wh: global entry index of each window row's first cell
    int *scan_tmp = grow + wh;       // blockDim
    constexpr int SW = NV + 2;       // slot: tx, ty, values

    __syncthreads();  // the caller's staged tiles are complete
    // phase 0: window cell counts -> local starts
    for (int c = threadIdx.x; c < ncw; c += blockDim.x) {
        const int r = c / ww, col = c - r * ww;
        const int cx = tx0 + col, cy = ty0 + r;
        int cnt = 0;
        if (te.active && cx < g.Wc && cy < g.Hc) {
            const int gc = cy * g.Wc + cx;
            cnt = te.cell_start[gc + 1] - te.cell_start[gc];
        }
        lstart[c] = cnt;
    }
    for (int r = threadIdx.x; r < wh; r += blockDim.x) {
        const int cy = ty0 + r;
        grow[r] = (te.active && cy < g.Hc) ? te.cell_start[cy * g.Wc + min(tx0, g.Wc - 1)] : 0;
    }
    __syncthreads();
    block_exclusive_scan(lstart, ncw, scan_tmp);
    const int total = lstart[ncw];

    // phase 1: entry values into LDS slots
    const int ncache = min(total, cap);
    for (int s = threadIdx.x; s < ncache; s += blockDim.x) {
        int r = 0;
        while (r + 1 < wh && lstart[(r + 1) * ww] <= s) r++;
        const int e = grow[r] + (s - lstart[r * ww]);
        const float t0 = te.tx[e], t1 = te.ty[e];
        int x0, y0;
        float wx[KS], wy[KS];
        float v[NV];
        if (footprint<KS>(t0, t1, g.W, g.H, x0, y0, wx, wy)) {
            vals(e, x0, y0, wx, wy, te.w[e], v);
        } else {
#pragma unroll
            for (int q = 0; q < NV; q++) v[q] = 0.0f;
        }
        float *sl = slots + (size_t)s * SW;
        sl[0] = t0;
        sl[1] = t1;
#pragma unroll
        for (int q = 0; q < NV; q++) sl[2 + q] = v[q];
    }
    __syncthreads();

    // phase 2: every thread pulls its pixels
    for (int p = threadIdx.x; p < g.tw * g.th; p += blockDim.x) {
        const int lx = p % g.tw, ly = p / g.tw;
        const int x = tx0 + lx, y = ty0 + ly;
        if (x >= g.W || y >= g.H) continue;
        float acc[NV];
#pragma unroll
        for (int q = 0; q < NV; q++) acc[q] = 0.0f;
#pragma unroll
        for (int b = 0; b < KS; b++) {
            const int row = ly + b;
            const int rbase = lstart[row * ww];
            const int i = KS - 1 - b;
#pragma unroll
            for (int a = 0; a < KS; a++) {
                const int cl = row * ww + lx + a;
                const int s0 = lstart[cl], s1 = lstart[cl + 1];
                const int j = KS - 1 - a;
                for (int s = s0; s < s1; s++) {
                    float t0, t1, v[NV];
                    if (s < cap) {
                        const float *sl = slots + (size_t)s * SW;
                        t0 = sl[0];
                        t1 = sl[1];
#pragma unroll
                        for (int q = 0; q < NV; q++) v[q] = sl[2 + q];
                    } else {
                        const int e = grow[row] + (s - rbase);
                        t0 = te.tx[e];
                        t1 = te.ty[e];
                        int xx0, yy0;
                        float wx2[KS], wy2[KS];
                        if (footprint<KS>(t0, t1, g.W, g.H, xx0, yy0, wx2, wy2)) {
                            vals(e, xx0, yy0, wx2, wy2, te.w[e], v);
                        } else {
#pragma unroll
                            for (int q = 0; q < NV; q++) v[q] = 0.0f;
                        }
                    }
                    int x0, y0;
                    float wx[KS], wy[KS];
                    if (!footprint<KS>(t0, t1, g.W, g.H, x0, y0, wx, wy)) continue;
                    tap_accumulate<NV, SQ>(acc, wx[j] * wy[i], v);
                }
            }
        }
        epi((size_t)y * g.W + x, x, y, acc);
    }
}

// Cell-owner pull for small footprints (ks <= 3).
//
// Stage: the window's entries (tau_x, tau_y, omega) are copied into LDS with
// coalesced loads -- the cells of one window row are contiguous in the CSR,
// so the window is wh contiguous runs.  Walk: thread c owns window cell c,
// walks the cell's staged entries, forms their values and accumulates the
// per-tap sums sum_e a_e(i,j) v_e in registers.  Publish: after a barrier
// the same LDS holds part[(i*KS+j)*NV + q][cell]; each pixel finally adds the
// KS*KS tap sums of the cells that cover it.  A window holding more than
// `ecap` entries walks them from global memory instead (same order, same
// sums).  Requires blockDim.x >= ncw.
constexpr int kStagePer = 16;  // staged entries per thread at most

template <int KS, int NV>
__host__ __device__ __forceinline__ size_t cell_pull_region_floats(int tw, int th, int ecap)
{
    const size_t part = (size_t)KS * KS * NV * (tw + KS - 1) * (th + KS - 1);
    const size_t stage = 3 * (size_t)ecap;
    return part > stage ? part : stage;
}

// Stage + walk of the cell pull: returns, in acc, the sums that `entry_fn(e,
// x0, y0, wx, wy, om, acc)` accumulates over the entries of the window cell
// owned by this thread (threadIdx.x < ncw; other threads get zeros).  On
// return every thread has finished reading the staged entries, so the region
// may be overwritten.  Requires blockDim.x >= ncw and the caller's LDS tiles
// complete (the first barrier below orders them).
// WEXT: the window is (tw + WEXT) x (th + WEXT) cells from the tile origin
// (KS - 1 for target-pixel tiles; 1 for tiles of box anchors, where the cell
// of box start b holds the anchor taps of b and b + 1).
template <int KS, int NC, bool W_FROM_SRC = false, bool ANCHOR = false, int NPAY = 0, int WEXT = KS - 1,
          class EntryFn>
__device__ __forceinline__ void cell_accumulate(const TileGeom &g, const TileEntries &te, int t, int *meta, int ecap,
                                                float (&acc)[NC], EntryFn entry_fn)
{
    const int tx0 = (t % g.ntx) * g.tw, ty0 = (t / g.ntx) * g.th;
    const int ww = g.tw + WEXT, wh = g.th + WEXT, ncw = ww * wh;
    int *rbeg = meta;       // wh: CSR index of the row's first entry
    int *roff = meta + wh;  // wh + 1: staged offset of the row
    float *region = reinterpret_cast<float *>(meta + 2 * wh + 2);
    float *stx = region, *sty = region + ecap, *sw = region + 2 * ecap;
    float *spay = region + 3 * ecap;  // NPAY planes of ecap

    // global loads first (they do not touch LDS): the window rows' CSR ranges
    // and the owned cell's range
    const int c = threadIdx.x;
    int rb = 0, rcnt = 0, e0 = 0, e1 = 0;
    if ((int)threadIdx.x < wh) {
        const int cy = ty0 + (int)threadIdx.x;
        if (te.active && cy < g.Hc) {
            rb = te.cell_start[cy * g.Wc + tx0];
            rcnt = te.cell_start[cy * g.Wc + min(tx0 + ww, g.Wc)] - rb;
        }
    }
    const int crow = c / ww, ccol = c - crow * ww;
    if (c < ncw && te.active && tx0 + ccol < g.Wc && ty0 + crow < g.Hc) {
        const int gc = (ty0 + crow) * g.Wc + tx0 + ccol;
        e0 = te.cell_start[gc];
        e1 = te.cell_start[gc + 1];
    }
    __syncthreads();  // the caller's LDS tiles are complete
    if (threadIdx.x < 64) {
        // staged offsets of the rows: exclusive scan of the counts by wave 0
        int v = (int)threadIdx.x < wh ? rcnt : 0;
        int incl = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int u = __shfl_up(incl, off, 64);
            if ((int)threadIdx.x >= off) incl += u;
        }
        if ((int)threadIdx.x < wh) {
            rbeg[threadIdx.x] = rb;
            roff[threadIdx.x] = incl - v;
        }
        if ((int)threadIdx.x == wh - 1) roff[wh] = incl;
    }
    __syncthreads();
    const int total = roff[wh];
    // the first ecap entries of the window are staged, the rest read in place
    const int nst = min(total, ecap);
    if (NPAY == 0) {
        float a[kStagePer], b[kStagePer], w[kStagePer];
        int r = 0;
#pragma unroll
        for (int k = 0; k < kStagePer; k++) {
            const int i = threadIdx.x + k * blockDim.x;
            if (i < nst) {
                while (roff[r + 1] <= i) r++;
                const int e = rbeg[r] + (i - roff[r]);
                a[k] = te.tx[e];
                b[k] = te.ty[e];
                // W_FROM_SRC: the third staged word is the source index
                w[k] = W_FROM_SRC ? __int_as_float(te.src[e]) : te.w[e];
            }
        }
#pragma unroll
        for (int k = 0; k < kStagePer; k++) {
            const int i = threadIdx.x + k * blockDim.x;
            if (i < nst) {
                stx[i] = a[k];
                sty[i] = b[k];
                sw[i] = w[k];
            }
        }
    } else {
        // with a payload: fewer entries per round, every load of a round in flight
        constexpr int SP = kStagePer / 4;
        int r = 0;
        for (int k0 = 0; k0 < kStagePer; k0 += SP) {
            float a[SP], b[SP], w[SP], pv[SP][NPAY > 0 ? NPAY : 1];
#pragma unroll
            for (int k = 0; k < SP; k++) {
                const int i = threadIdx.x + (k0 + k) * blockDim.x;
                if (i < nst) {
                    while (roff[r + 1] <= i) r++;
                    const int e = rbeg[r] + (i - roff[r]);
                    a[k] = te.tx[e];
                    b[k] = te.ty[e];
                    w[k] = te.w[e];
#pragma unroll
                    for (int q = 0; q < NPAY; q++) pv[k][q] = te.pay[q * te.pay_stride + e];
                }
            }
#pragma unroll
            for (int k = 0; k < SP; k++) {
                const int i = threadIdx.x + (k0 + k) * blockDim.x;
                if (i < nst) {
                    stx[i] = a[k];
                    sty[i] = b[k];
                    sw[i] = w[k];
#pragma unroll
                    for (int q = 0; q < NPAY; q++) spay[q * ecap + i] = pv[k][q];
                }
            }
        }
    }
    __syncthreads();

#pragma unroll
    for (int k = 0; k < NC; k++) acc[k] = 0.0f;
    if (c < ncw) {
        {
            const int lbase = roff[crow] - rbeg[crow];
            for (int e = e0; e < e1; e++) {
                float t0, t1, om;
                float pv[NPAY > 0 ? NPAY : 1];
                int id = e;  // entry index, or the source index with W_FROM_SRC
                const int li = lbase + e;
                if (li < nst) {
                    t0 = stx[li];
                    t1 = sty[li];
                    if (W_FROM_SRC) id = __float_as_int(sw[li]);
                    else om = sw[li];
#pragma unroll
                    for (int q = 0; q < NPAY; q++) pv[q] = spay[q * ecap + li];
                } else {
                    t0 = te.tx[e];
                    t1 = te.ty[e];
                    if (W_FROM_SRC) id = te.src[e];
                    else om = te.w[e];
#pragma unroll
                    for (int q = 0; q < NPAY; q++) pv[q] = te.pay[q * te.pay_stride + e];
                }
                if (W_FROM_SRC) om = te.omega[id];
                int x0, y0;
                if constexpr (ANCHOR) {
                    float ax[2], ay[2];
                    if (!box_weights<KS>(t0, t1, g.W, g.H, x0, y0, ax, ay)) continue;
                    if constexpr (NPAY > 0) entry_fn(id, x0, y0, ax, ay, om, pv, acc);
                    else entry_fn(id, x0, y0, ax, ay, om, acc);
                } else {
                    float wx[KS], wy[KS];
                    if (!footprint<KS>(t0, t1, g.W, g.H, x0, y0, wx, wy)) continue;
                    entry_fn(id, x0, y0, wx, wy, om, acc);
                }
            }
        }
    }
    __syncthreads();  // every walk is done with the staged entries
}

template <int KS, int NV, int SQ, class Vals, class Epi>
__device__ __forceinline__ void cell_pull(const TileGeom &g, const TileEntries &te, int t, int *meta,
                                          int ecap, Vals vals, Epi epi)
{
    const int tx0 = (t % g.ntx) * g.tw, ty0 = (t / g.ntx) * g.th;
    const int ww = g.tw + KS - 1, wh = g.th + KS - 1, ncw = ww * wh;
    float acc[KS * KS * NV];
    cell_accumulate<KS, KS * KS * NV>(
        g, te, t, meta, ecap, acc,
        [&](int e, int x0, int y0, const float(&wx)[KS], const float(&wy)[KS], float om,
            float(&a)[KS * KS * NV]) {
            float v[NV];
            vals(e, x0, y0, wx, wy, om, v);
#pragma unroll
            for (int i = 0; i < KS; i++)
#pragma unroll
                for (int j = 0; j < KS; j++) {
                    const float w = wy[i] * wx[j];
#pragma unroll
                    for (int q = 0; q < NV; q++) a[(i * KS + j) * NV + q] += (((SQ >> q) & 1) ? w * w : w) * v[q];
                }
        });
    float *part = reinterpret_cast<float *>(meta + 2 * wh + 2);
    const int c = threadIdx.x;
    if (c < ncw) {
#pragma unroll
        for (int k = 0; k < KS * KS * NV; k++) part[k * ncw + c] = acc[k];
    }
    __syncthreads();
    for (int p = threadIdx.x; p < g.tw * g.th; p += blockDim.x) {
        const int lx = p % g.tw, ly = p / g.tw;
        const int x = tx0 + lx, y = ty0 + ly;
        if (x >= g.W || y >= g.H) continue;
        float sum[NV];
#pragma unroll
        for (int q = 0; q < NV; q++) sum[q] = 0.0f;
#pragma unroll
        for (int i = 0; i < KS; i++)
#pragma unroll
            for (int j = 0; j < KS; j++) {
                // pixel = x0 + j with x0 = cx - (KS-1)  =>  local cell column lx - j + KS - 1
                const int cc = (ly - i + KS - 1) * ww + (lx - j + KS - 1);
#pragma unroll
                for (int q = 0; q < NV; q++) sum[q] += part[((i * KS + j) * NV + q) * ncw + cc];
            }
        epi((size_t)y * g.W + x, x, y, sum);
    }
}

// ---- box decomposition of large footprints ---------------------------------
//
// Along x the footprint weights of an entry are d = (1-f, 1, ..., 1, f) on the
// ks = dsf + 1 taps x0 .. x0+dsf (f = the fraction of the box start), divided
// by their in-image sum sx; out-of-image taps are 0.  That is
//   (1-f)/sx * box(x0) + f/sx * box(x0+1),  box(b) = 1 on b .. b+dsf-1,
// restricted to the image, and the 2-D weights are the product of the x and y
// forms.  So a splat is a 2 x 2 (bilinear-like) splat of the entry values
// onto box starts (anchors) followed by a dsf x dsf box sum, and a gather is a
// 2 x 2 gather of the box-summed image (zero outside) at the anchors.  Per
// entry 4 taps instead of ks^2.  The anchor of cell (cx, cy) is the entry's
// box start x0 = cx - dsf; tap (i, j) goes to anchor (x0 + j, y0 + i).

// LDS floats of the box pull's region (beside its 2 wh + 2 meta ints): the
// staged entries (tau, omega and NPAY payload floats each), then the taps
// shared between neighbouring cells (2 NV per cell) and the anchor sums (NV).
template <int KS, int NV, int NPAY = 0>
__host__ __device__ __forceinline__ size_t box_pull_region_floats(int tw, int th, int ecap)
{
    const size_t cells = (size_t)(tw + KS - 1) * (th + KS - 1);
    const size_t part = 3 * (size_t)NV * cells;
    const size_t stage = (3 + NPAY) * (size_t)ecap;
    return part > stage ? part : stage;
}

// Box pull (ks >= 3, plain weights).  Walk: thread c owns window cell c and
// accumulates, over its entries, the 2 x 2 anchor taps of vals(e, x0, y0, ax,
// ay, om, [pay,] v).  Cell (col, row) puts tap (i, j) on anchor (col + j,
// row + i): each anchor adds its own cell's (0, 0) tap, the left cell's
// (0, 1), the upper cell's (1, 0) and the upper-left cell's (1, 1) -- first
// along x (the (., 1) taps through LDS), then along y.  Each pixel finally
// adds the dsf x dsf anchors whose box covers it.  Requires blockDim.x >= ncw.
template <int KS, int NV, int NPAY = 0, class Vals, class Epi>
__device__ __forceinline__ void box_pull(const TileGeom &g, const TileEntries &te, int t, int *meta, int ecap,
                                         Vals vals, Epi epi)
{
    constexpr int D = KS - 1;
    const int tx0 = (t % g.ntx) * g.tw, ty0 = (t / g.ntx) * g.th;
    const int ww = g.tw + KS - 1, wh = g.th + KS - 1, ncw = ww * wh;
    float acc[4 * NV];
    auto add_taps = [&](const float(&ax)[2], const float(&ay)[2], const float(&v)[NV], float(&a)[4 * NV]) {
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const float w = ay[i] * ax[j];
#pragma unroll
                for (int q = 0; q < NV; q++) a[(i * 2 + j) * NV + q] += w * v[q];
            }
    };
    if constexpr (NPAY > 0) {
        cell_accumulate<KS, 4 * NV, false, true, NPAY>(
            g, te, t, meta, ecap, acc,
            [&](int e, int x0, int y0, const float(&ax)[2], const float(&ay)[2], float om,
                const float(&pay)[NPAY], float(&a)[4 * NV]) {
                float v[NV];
                vals(e, x0, y0, ax, ay, om, pay, v);
                add_taps(ax, ay, v, a);
            });
    } else {
        cell_accumulate<KS, 4 * NV, false, true>(
            g, te, t, meta, ecap, acc,
            [&](int e, int x0, int y0, const float(&ax)[2], const float(&ay)[2], float om, float(&a)[4 * NV]) {
                float v[NV];
                vals(e, x0, y0, ax, ay, om, v);
                add_taps(ax, ay, v, a);
            });
    }
    float *sh = reinterpret_cast<float *>(meta + 2 * wh + 2);  // 2 NV * ncw
    float *anc = sh + 2 * NV * ncw;                             // NV * ncw
    const int c = threadIdx.x;
    const int row = c / ww, col = c - row * ww;
    const bool own = c < ncw;
    // along x: h_i = own (i, 0) tap + the left cell's (i, 1) tap
    if (own) {
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int q = 0; q < NV; q++) sh[(i * NV + q) * ncw + c] = acc[(i * 2 + 1) * NV + q];
    }
    __syncthreads();
    float h[2 * NV];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int q = 0; q < NV; q++)
            h[i * NV + q] = acc[(i * 2) * NV + q] + ((own && col > 0) ? sh[(i * NV + q) * ncw + c - 1] : 0.0f);
    __syncthreads();
    // along y: anchor = own h_0 + the upper cell's h_1
    if (own) {
#pragma unroll
        for (int q = 0; q < NV; q++) sh[q * ncw + c] = h[NV + q];
    }
    __syncthreads();
    if (own) {
#pragma unroll
        for (int q = 0; q < NV; q++) anc[q * ncw + c] = h[q] + (row > 0 ? sh[q * ncw + c - ww] : 0.0f);
    }
    __syncthreads();
    // pixel (x, y) = the anchors ac in [lx + 1, lx + D], ar in [ly + 1, ly + D]
    for (int p = threadIdx.x; p < g.tw * g.th; p += blockDim.x) {
        const int lx = p % g.tw, ly = p / g.tw;
        const int x = tx0 + lx, y = ty0 + ly;
        if (x >= g.W || y >= g.H) continue;
        float sum[NV];
#pragma unroll
        for (int q = 0; q < NV; q++) sum[q] = 0.0f;
#pragma unroll 1
        for (int i = 1; i <= D; i++)
#pragma unroll
            for (int j = 1; j <= D; j++) {
                const int a = (ly + i) * ww + lx + j;
#pragma unroll
                for (int q = 0; q < NV; q++) sum[q] += anc[q * ncw + a];
            }
        epi((size_t)y * g.W + x, x, y, sum);
    }
}

// The pull of the splat-only kernels: `vals` must not use the weights it is
// given (their number depends on the pull); the choice is splat_kind's.
template <int KS, int NV, int SQ, class Vals, class Epi>
__device__ __forceinline__ void splat_pull(const TileGeom &g, const TileEntries &te, int t, int *lds, int cap,
                                           Vals vals, Epi epi)
{
    if constexpr (KS <= 2 || (SQ != 0 && KS <= 3)) {
        cell_pull<KS, NV, SQ>(g, te, t, lds, cap, vals, epi);
    } else if constexpr (SQ == 0) {
        box_pull<KS, NV>(g, te, t, lds, cap, vals, epi);
    } else {
        float *slots = reinterpret_cast<float *>(lds + pull_meta_ints(g.tw, g.th, KS, blockDim.x));
        tile_pull<KS, NV, SQ>(g, te, t, lds, slots, cap, vals, epi);
    }
}

// The pull used by every splat kernel: cell-owner with LDS staging for
// ks <= 3, LDS slots for larger footprints (whose cells hold few entries).
// `lds` is the kernel's pull region; `cap` is the staged-entry capacity
// (ks <= 3) or the slot capacity (ks >= 4).
template <int KS, int NV, int SQ, class Vals, class Epi>
__device__ __forceinline__ void any_pull(const TileGeom &g, const TileEntries &te, int t, int *lds, int cap,
                                         Vals vals, Epi epi)
{
    if constexpr (KS <= 3) {
        cell_pull<KS, NV, SQ>(g, te, t, lds, cap, vals, epi);
    } else {
        float *slots = reinterpret_cast<float *>(lds + pull_meta_ints(g.tw, g.th, KS, blockDim.x));
        tile_pull<KS, NV, SQ>(g, te, t, lds, slots, cap, vals, epi);
    }
}

// LDS bytes of a slot pull with NV values per slot and `cap` slots.
inline size_t pull_lds_bytes(int tw, int th, int ks, int bs, int nv, int cap)
{
    return sizeof(int) * (size_t)pull_meta_ints(tw, th, ks, bs) + sizeof(float) * (size_t)(nv + 2) * cap;
}

}  // namespace lfrv
