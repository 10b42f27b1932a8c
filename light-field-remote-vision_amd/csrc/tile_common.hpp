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
    int active;  // 0: the tile kernels see no entries
};

TileGeom tile_geom(const lfrv_ctx *ctx);
TileEntries tile_entries(const lfrv_ctx *ctx);

int pull_cap(const lfrv_ctx *ctx, int nv, size_t extra, size_t budget);
size_t pull_lds(const lfrv_ctx *ctx, int nv, int cap);

lfrv_status run_rhs_and_diag(lfrv_ctx *ctx, double alpha, double c0, double beta, bool use_views,
                             bool use_precond, const float *u0);
lfrv_status run_gauss_smooth(lfrv_ctx *ctx, const float *in, float *out, float *tmp);
lfrv_status run_view_stencils(lfrv_ctx *ctx);

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

// LDS words the pull needs besides the slots: lstart (ncw + 1), row bases
// (wh), scan scratch (blockDim).
__host__ __device__ __forceinline__ int pull_window_w(int tw, int ks) { return tw + ks - 1; }
__host__ __device__ __forceinline__ int pull_meta_ints(int tw, int th, int ks, int bs)
{
    return (tw + ks - 1) * (th + ks - 1) + 1 + (th + ks - 1) + bs;
}

// The pull.  `vals(e, x0, y0, wx, wy, v)` produces the NV values of entry e;
// `epi(o, x, y, acc)` receives the pulled sums of pixel (x, y).  With
// SQUARED the footprint weight enters squared (a^2 v, the Jacobi diagonal).
// Slots beyond `cap` are not cached: phase 2 recomputes their values.
template <int KS, int NV, bool SQUARED, class Vals, class Epi>
__device__ __forceinline__ void tile_pull(const TileGeom &g, const TileEntries &te, int t, int *meta,
                                          float *slots, int cap, Vals vals, Epi epi)
{
    const int tx0 = (t % g.ntx) * g.tw, ty0 = (t / g.ntx) * g.th;
    const int ww = g.tw + KS - 1, wh = g.th + KS - 1, ncw = ww * wh;
    int *lstart = meta;              // ncw + 1
    int *grow = meta + ncw + 1;      // wh: global entry index of each window row's first cell
    int *scan_tmp = grow + wh;       // blockDim
    constexpr int SW = NV + 2;       // slot: tx, ty, values

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
            vals(e, x0, y0, wx, wy, v);
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
                            vals(e, xx0, yy0, wx2, wy2, v);
                        } else {
#pragma unroll
                            for (int q = 0; q < NV; q++) v[q] = 0.0f;
                        }
                    }
                    int x0, y0;
                    float wx[KS], wy[KS];
                    if (!footprint<KS>(t0, t1, g.W, g.H, x0, y0, wx, wy)) continue;
                    float a2 = wx[j] * wy[i];
                    if (SQUARED) a2 *= a2;
#pragma unroll
                    for (int q = 0; q < NV; q++) acc[q] += a2 * v[q];
                }
            }
        }
        epi((size_t)y * g.W + x, x, y, acc);
    }
}

// Cell-owner pull for small footprints (ks <= 3).  Thread c owns window cell
// c: it walks the cell's entries, forms their values and accumulates the
// per-tap sums sum_e a_e(i,j) v_e in registers, then publishes them to LDS
// part[(i*KS+j)*NV + q][cell]; each pixel finally adds the KS*KS tap sums of
// the cells that cover it.  LDS: KS*KS*NV*ncw floats + ncw+1 ints.
template <int KS>
__host__ __device__ __forceinline__ int cell_pull_cells(int tw, int th)
{
    return (tw + KS - 1) * (th + KS - 1);
}

template <int KS, int NV, bool SQUARED, class Vals, class Epi>
__device__ __forceinline__ void cell_pull(const TileGeom &g, const TileEntries &te, int t, float *part, Vals vals,
                                          Epi epi)
{
    const int tx0 = (t % g.ntx) * g.tw, ty0 = (t / g.ntx) * g.th;
    const int ww = g.tw + KS - 1, wh = g.th + KS - 1, ncw = ww * wh;
    __syncthreads();  // the caller's staged tiles are complete
    for (int c = threadIdx.x; c < ncw; c += blockDim.x) {
        const int r = c / ww, col = c - r * ww;
        const int cx = tx0 + col, cy = ty0 + r;
        float acc[KS * KS][NV];
#pragma unroll
        for (int k = 0; k < KS * KS; k++)
#pragma unroll
            for (int q = 0; q < NV; q++) acc[k][q] = 0.0f;
        if (te.active && cx < g.Wc && cy < g.Hc) {
            const int gc = cy * g.Wc + cx;
            const int e0 = te.cell_start[gc], e1 = te.cell_start[gc + 1];
            for (int e = e0; e < e1; e++) {
                int x0, y0;
                float wx[KS], wy[KS];
                if (!footprint<KS>(te.tx[e], te.ty[e], g.W, g.H, x0, y0, wx, wy)) continue;
                float v[NV];
                vals(e, x0, y0, wx, wy, v);
#pragma unroll
                for (int i = 0; i < KS; i++)
#pragma unroll
                    for (int j = 0; j < KS; j++) {
                        float a = wy[i] * wx[j];
                        if (SQUARED) a *= a;
#pragma unroll
                        for (int q = 0; q < NV; q++) acc[i * KS + j][q] += a * v[q];
                    }
            }
        }
#pragma unroll
        for (int k = 0; k < KS * KS; k++)
#pragma unroll
            for (int q = 0; q < NV; q++) part[(k * NV + q) * ncw + c] = acc[k][q];
    }
    __syncthreads();
    for (int p = threadIdx.x; p < g.tw * g.th; p += blockDim.x) {
        const int lx = p % g.tw, ly = p / g.tw;
        const int x = tx0 + lx, y = ty0 + ly;
        if (x >= g.W || y >= g.H) continue;
        float acc[NV];
#pragma unroll
        for (int q = 0; q < NV; q++) acc[q] = 0.0f;
#pragma unroll
        for (int i = 0; i < KS; i++)
#pragma unroll
            for (int j = 0; j < KS; j++) {
                // pixel = x0 + j with x0 = cx - (KS-1)  =>  local cell column lx - j + KS - 1
                const int c = (ly - i + KS - 1) * ww + (lx - j + KS - 1);
#pragma unroll
                for (int q = 0; q < NV; q++) acc[q] += part[((i * KS + j) * NV + q) * ncw + c];
            }
        epi((size_t)y * g.W + x, x, y, acc);
    }
}

template <int KS, int NV>
__host__ __device__ __forceinline__ size_t cell_pull_lds(int tw, int th)
{
    return sizeof(float) * (size_t)KS * KS * NV * cell_pull_cells<KS>(tw, th);
}

// The pull used by every splat kernel: cell-owner for ks <= 3, LDS slots for
// larger footprints (whose cells hold few entries).  `lds` is the kernel's
// pull region (cell_pull: part planes; tile_pull: meta ints then slots).
template <int KS, int NV, bool SQUARED, class Vals, class Epi>
__device__ __forceinline__ void any_pull(const TileGeom &g, const TileEntries &te, int t, int *lds, int cap,
                                         Vals vals, Epi epi)
{
    if constexpr (KS <= 3) {
        cell_pull<KS, NV, SQUARED>(g, te, t, reinterpret_cast<float *>(lds), vals, epi);
    } else {
        float *slots = reinterpret_cast<float *>(lds + pull_meta_ints(g.tw, g.th, KS, blockDim.x));
        tile_pull<KS, NV, SQUARED>(g, te, t, lds, slots, cap, vals, epi);
    }
}

// LDS bytes of a pull with NV values per slot and `cap` slots.
inline size_t pull_lds_bytes(int tw, int th, int ks, int bs, int nv, int cap)
{
    return sizeof(int) * (size_t)pull_meta_ints(tw, th, ks, bs) + sizeof(float) * (size_t)(nv + 2) * cap;
}

}  // namespace lfrv
