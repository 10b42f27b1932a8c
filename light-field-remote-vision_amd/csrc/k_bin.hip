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

// Binning: one thread per source pixel position, looping over the views, so
// the loads are coalesced and the V entries that land in the same target
// neighbourhood are counted / stored by the same wave close in time.
constexpr int kViewBatch = 4;  // views whose tau is loaded together by the binning kernels

template <int KS>
__global__ void k_cell_count(int V, int w, int h, int W, int H, int Wc, const float *__restrict__ tx,
                             const float *__restrict__ ty, int *__restrict__ count)
{
    const int mx = blockIdx.x * blockDim.x + threadIdx.x, my = blockIdx.y * blockDim.y + threadIdx.y;
    if (mx >= w || my >= h) return;
    const size_t n = (size_t)w * h, m = (size_t)my * w + mx;
    for (int v0 = 0; v0 < V; v0 += kViewBatch) {
        float t0[kViewBatch], t1[kViewBatch];
#pragma unroll
        for (int b = 0; b < kViewBatch; b++) {
            const size_t k = (size_t)min(v0 + b, V - 1) * n + m;
            t0[b] = tx[k];
            t1[b] = ty[k];
        }
#pragma unroll
        for (int b = 0; b < kViewBatch; b++)
            if (v0 + b < V && cell_valid(t0[b], t1[b], W, H)) atomicAdd(count + cell_of<KS>(t0[b], t1[b], Wc), 1);
    }
}

// The fill's blocks are numbered so that each XCD (blocks are dealt to the 8
// XCDs round-robin) takes a contiguous band of source rows: the cells its
// entries land in are then mostly its own, and the partial cache lines of
// the scattered entry writes are completed in one L2 instead of several.
template <int KS>
__global__ void k_cell_fill(int V, int w, int h, int W, int H, int Wc, const float *__restrict__ tx,
                            const float *__restrict__ ty, int *__restrict__ cursor, int *__restrict__ e_src)
{
    const int nbx = (w + blockDim.x - 1) / blockDim.x, nb = nbx * ((h + blockDim.y - 1) / blockDim.y);
    const int lb = xcd_tile(blockIdx.x, nb), bx = lb % nbx, by = lb / nbx;
    const int mx = bx * blockDim.x + threadIdx.x, my = by * blockDim.y + threadIdx.y;
    if (mx >= w || my >= h) return;
    const size_t n = (size_t)w * h, m = (size_t)my * w + mx;
    for (int v0 = 0; v0 < V; v0 += kViewBatch) {
        float t0[kViewBatch], t1[kViewBatch];
#pragma unroll
        for (int b = 0; b < kViewBatch; b++) {
            const size_t k = (size_t)min(v0 + b, V - 1) * n + m;
            t0[b] = tx[k];
            t1[b] = ty[k];
        }
        int c[kViewBatch], pos[kViewBatch];
#pragma unroll
        for (int b = 0; b < kViewBatch; b++) {
            const bool ok = v0 + b < V && cell_valid(t0[b], t1[b], W, H);
            c[b] = ok ? cell_of<KS>(t0[b], t1[b], Wc) : -1;
            pos[b] = ok ? atomicAdd(cursor + c[b], 1) : 0;  // the cursors start at the cell starts
        }
#pragma unroll
        for (int b = 0; b < kViewBatch; b++)
            if (c[b] >= 0) e_src[pos[b]] = (int)((size_t)(v0 + b) * n + m);
    }
}

// Order every cell's entries by source index and gather their tau.  A block
// of kSortCells threads owns `cpb` (<= kSortCells) consecutive cells, whose
// entries are one contiguous run: the run is staged in LDS with coalesced
// loads, each owner thread insertion-sorts its cell's segment there (cells
// hold a handful of entries), and the sorted run is written back coalesced.
// The host picks cpb so that the expected run (cpb times the mean entries per
// cell, with a margin) fits kSortCap: with many views per cell (dsf 1, 25
// views: 25 entries per cell) a block of 256 cells would overflow the stage.
// Runs longer than kSortCap are sorted in place in global memory.
constexpr int kSortCells = 256;
constexpr int kSortCap = 4096;

// Cells per sort block for a mean of `mean` entries per cell.
inline int sort_cells_per_block(double mean)
{
    const double fit = (double)kSortCap / (1.5 * std::max(mean, 1.0));
    const int c = (int)(fit / 32.0) * 32;
    return std::max(32, std::min(kSortCells, c));
}

// Per-cell footprint tap sums sum_e a_e(i, j) of the cell's sorted entries
// (the coverage pull's per-cell phase), written as tap planes [KS*KS][ncells].
template <int KS>
__device__ __forceinline__ void cell_tap_sums(int c, int a, int b, const int *keys, const float *__restrict__ tx,
                                              const float *__restrict__ ty, int W, int H, int ncells,
                                              float *__restrict__ cell_tap)
{
    float acc[KS * KS];
#pragma unroll
    for (int q = 0; q < KS * KS; q++) acc[q] = 0.0f;
    for (int i = a; i < b; i++) {
        const int k = keys[i];
        int x0, y0;
        float wx[KS], wy[KS];
        if (!footprint<KS>(tx[k], ty[k], W, H, x0, y0, wx, wy)) continue;
#pragma unroll
        for (int ii = 0; ii < KS; ii++)
#pragma unroll
            for (int jj = 0; jj < KS; jj++) acc[ii * KS + jj] += wy[ii] * wx[jj];
    }
#pragma unroll
    for (int q = 0; q < KS * KS; q++) cell_tap[(size_t)q * ncells + c] = acc[q];
}

template <int KS>
__global__ void __launch_bounds__(kSortCells) k_cell_sort(int ncells, int cpb, int W, int H, const int *__restrict__ start,
                                                         int *__restrict__ e_src, const float *__restrict__ tx,
                                                         const float *__restrict__ ty, float *__restrict__ e_tx,
                                                         float *__restrict__ e_ty, float *__restrict__ cell_tap)
{
    __shared__ int keys[kSortCap];
    const int c0 = blockIdx.x * cpb;
    const int c1 = min(c0 + cpb, ncells);
    const int base = start[c0], total = start[c1] - base;
    const int c = c0 + (int)threadIdx.x;
    if (total <= kSortCap) {
        for (int i = threadIdx.x; i < total; i += blockDim.x) keys[i] = e_src[base + i];
        __syncthreads();
        if (c < c1) {
            const int a = start[c] - base, b = start[c + 1] - base;
            for (int i = a + 1; i < b; i++) {
                const int key = keys[i];
                int j = i - 1;
                while (j >= a && keys[j] > key) {
                    keys[j + 1] = keys[j];
                    j--;
                }
                keys[j + 1] = key;
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < total; i += blockDim.x) {
            const int k = keys[i];
            e_src[base + i] = k;
            e_tx[base + i] = tx[k];
            e_ty[base + i] = ty[k];
        }
        if (cell_tap && c < c1) cell_tap_sums<KS>(c, start[c] - base, start[c + 1] - base, keys, tx, ty, W, H, ncells,
                                                  cell_tap);
        return;
    }
    if (c >= c1) return;
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
    if (cell_tap) cell_tap_sums<KS>(c, a, b, e_src, tx, ty, W, H, ncells, cell_tap);
}

// coverage(p) = sum of the tap sums of the cells whose footprint holds p, in
// the order of the pull's pixel phase (taps (i, j) row-major, cell
// (x + KS-1-j, y + KS-1-i)).
template <int KS>
__global__ void k_coverage_from_taps(int W, int H, int Wc, int ncells, const float *__restrict__ cell_tap,
                                     float *__restrict__ coverage)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    float v[KS * KS];
#pragma unroll
    for (int i = 0; i < KS; i++)
#pragma unroll
        for (int j = 0; j < KS; j++)
            v[i * KS + j] = cell_tap[(size_t)(i * KS + j) * ncells + (size_t)(y + KS - 1 - i) * Wc + (x + KS - 1 - j)];
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < KS * KS; q++) s += v[q];
    coverage[(size_t)y * W + x] = s;
}

// Device-wide exclusive scan of the cell counts `in` (n ints) into the cell
// starts `out` (n + 1 ints) in two launches: the sums of the blocks of
// kScanBlock counts, then per block the sum of the blocks before it (at most
// a few hundred ints, read by every block) and the scan of its own counts.
// The fill cursors start at the cell starts (the fill's atomicAdd then
// returns the entry's slot itself).
constexpr int kScanPer = 8;                  // counts per thread
constexpr int kScanBlock = 256 * kScanPer;  // counts per block

__device__ __forceinline__ int wave_incl_scan(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    return v;
}

__global__ void __launch_bounds__(256) k_scan_reduce(int n, const int *__restrict__ in, int *__restrict__ bsum)
{
    __shared__ int ws[4];
    const int base = blockIdx.x * kScanBlock + threadIdx.x * kScanPer;
    int s = 0;
#pragma unroll
    for (int i = 0; i < kScanPer; i++) s += (base + i < n) ? in[base + i] : 0;
    s = wave_incl_scan(s);
    if ((threadIdx.x & 63) == 63) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) bsum[blockIdx.x] = (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

__global__ void __launch_bounds__(256) k_scan_apply(int n, const int *__restrict__ in, const int *__restrict__ bsum,
                                                    int *__restrict__ out, int *__restrict__ cursor)
{
    __shared__ int ws[4], wo[4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    // this block's offset: the sums of the blocks before it
    int o = 0;
    for (int i = threadIdx.x; i < (int)blockIdx.x; i += 256) o += bsum[i];
    const int base = blockIdx.x * kScanBlock + threadIdx.x * kScanPer;
    int v[kScanPer], s = 0;
#pragma unroll
    for (int i = 0; i < kScanPer; i++) {
        v[i] = (base + i < n) ? in[base + i] : 0;
        s += v[i];
    }
    o = wave_incl_scan(o);
    const int incl = wave_incl_scan(s);
    if (lane == 63) {
        wo[wid] = o;
        ws[wid] = incl;
    }
    __syncthreads();
    int run = (wo[0] + wo[1]) + (wo[2] + wo[3]) + incl - s;
    for (int w = 0; w < wid; w++) run += ws[w];
#pragma unroll
    for (int i = 0; i < kScanPer; i++) {
        if (base + i < n) {
            out[base + i] = run;
            cursor[base + i] = run;
        }
        run += v[i];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 255) out[n] = run;
}

// The payload of the IBR_color data term for the first *total entries: the
// entry's source pixel in its view (3 planes) and in its view's Laplacian.
__global__ void k_entry_payload(int64_t cap, const int *__restrict__ total, const int *__restrict__ e_src, size_t n,
                                const float *__restrict__ views, const float *__restrict__ vlap,
                                float *__restrict__ pay)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= cap || e >= *total) return;
    const int k = e_src[e];
    const unsigned vi = (unsigned)k / (unsigned)n, m = (unsigned)k - vi * (unsigned)n;
    const float *vp = views + (size_t)vi * 3 * n + m, *lp = vlap + (size_t)vi * 3 * n + m;
    float v[6];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        v[c] = vp[c * n];
        v[3 + c] = lp[c * n];
    }
#pragma unroll
    for (int q = 0; q < 6; q++) pay[q * cap + e] = v[q];
}

// e_w = omega[e_src] for the first *total entries (total read on the device).
__global__ void k_refresh_weights(int64_t cap, const int *__restrict__ total, const int *__restrict__ e_src,
                                  const float *__restrict__ omega, float *__restrict__ e_w)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < cap && e < *total) e_w[e] = omega[e_src[e]];
}

// ---- pull kernels -----------------------------------------------------------

struct PullCfg {
    int cap;  // staged entries (ks <= 3) or LDS slots (ks >= 4)
};

#define PULL_PROLOGUE()               \
    extern __shared__ int smem[];     \
    const int t = xcd_tile(blockIdx.x, g.ntiles);

// (the footprint weights are not used by a splat-only pull's values)
#define PULL_VALS(NV) [&](int e, int, int, const auto &, const auto &, float om, float(&v)[NV])

// coverage = sum_i B_i^T 1
template <int KS>
__global__ void __launch_bounds__(kPullMaxThreads) k_pull_coverage(TileGeom g, TileEntries te, PullCfg pc, float *__restrict__ coverage)
{
    PULL_PROLOGUE()
    splat_pull<KS, 1, 0>(
        g, te, t, smem, pc.cap, PULL_VALS(1) { v[0] = 1.0f; },
        [&](size_t o, int, int, const float(&a)[1]) { coverage[o] = a[0]; });
}

// Weighted initial image (gradientIBR.cu:764-811, optimization.cu:1303-1349):
// U = sum B^T (omega v) / sum B^T omega where the weight is non-zero, else 0;
// norm_mask = sum B^T omega.
template <int KS>
__global__ void __launch_bounds__(kPullMaxThreads) k_pull_init_image(TileGeom g, TileEntries te, PullCfg pc, size_t n,
                                                         const float *__restrict__ views, float *__restrict__ U,
                                                         float *__restrict__ norm_mask)
{
    PULL_PROLOGUE()
    const size_t N = (size_t)g.W * g.H;
    splat_pull<KS, 4, 0>(
        g, te, t, smem, pc.cap,
        PULL_VALS(4) {
            const int k = te.src[e];
            const unsigned vi = (unsigned)k / (unsigned)n, m = (unsigned)k - vi * (unsigned)n;
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

// The gradient-IBR set-up in one pull: every sum the set-up needs from the
// same entries, read once.
//   initial image   U = sum B^T(omega v) / sum B^T omega, norm_mask = sum B^T omega
//                   (gradientIBR.cu:764-811)
//   target gradient g = sum B^T Sobel(Y v) / sum B^T 1, target mask
//                   sigma^2 sum B^T omega / sum B^T 1 (gradientIBR.cu:858-913)
//   and, kept for the CG right-hand side and preconditioner, the raw sums
//   num = sum B^T(omega v) (3 planes) and sq = sum (B^2)^T omega.
// sum B^T 1 is the coverage plane of lfrv_prepare (the same sums in the same
// order as the u-gradient pull's own normaliser).
template <int KS>
__global__ void __launch_bounds__(kPullMaxThreads) k_pull_gibr_setup(TileGeom g, TileEntries te, PullCfg pc, size_t n,
                                                         const float *__restrict__ views,
                                                         const float *__restrict__ vgrad,
                                                         const float *__restrict__ coverage, float sigma2,
                                                         float *__restrict__ U, float *__restrict__ norm_mask,
                                                         float *__restrict__ ugx, float *__restrict__ ugy,
                                                         float *__restrict__ tmask, float *__restrict__ num,
                                                         float *__restrict__ sq)
{
    PULL_PROLOGUE()
    const size_t N = (size_t)g.W * g.H;
    splat_pull<KS, 7, (1 << 6)>(
        g, te, t, smem, pc.cap,
        PULL_VALS(7) {
            const int k = te.src[e];
            const unsigned vi = (unsigned)k / (unsigned)n, m = (unsigned)k - vi * (unsigned)n;
            const float *vp = views + vi * 3 * n + m;
            v[0] = om * vp[0];
            v[1] = om * vp[n];
            v[2] = om * vp[2 * n];
            v[3] = om;
            v[4] = vgrad[vi * 2 * n + m];
            v[5] = vgrad[vi * 2 * n + n + m];
            v[6] = om;
        },
        [&](size_t o, int, int, const float(&a)[7]) {
            const float wsum = a[3];
            for (int c = 0; c < 3; c++) {
                U[c * N + o] = (wsum != 0.0f) ? a[c] / wsum : a[c];
                num[c * N + o] = a[c];
            }
            norm_mask[o] = wsum;
            const float nm = coverage[o];
            float gx = a[4], gy = a[5], tm = a[3];
            if (nm != 0.0f) {
                gx /= nm;
                gy /= nm;
                tm /= nm;
            }
            ugx[o] = gx;
            ugy[o] = gy;
            tmask[o] = tm * sigma2;
            sq[o] = a[6];
        });
}

// dsf = 1: the set-up pull above fused with the assembly of the CG stencil
// (the six cross Gram sums of each 2 x 2 footprint; the diagonal is the sq
// sum).  omega is read through the staged source index, so the sorted entry
// weights need no refresh.  gram = 5 planes D, E, S, SE, SW; D is sq.
__global__ void __launch_bounds__(kPullMaxThreads) k_gibr_setup_st2(TileGeom g, TileEntries te, PullCfg pc, size_t n,
                                                        const float *__restrict__ views,
                                                        const float *__restrict__ vgrad,
                                                        const float *__restrict__ coverage, float sigma2,
                                                        float *__restrict__ U, float *__restrict__ norm_mask,
                                                        float *__restrict__ ugx, float *__restrict__ ugy,
                                                        float *__restrict__ tmask, float *__restrict__ num,
                                                        float *__restrict__ gram)
{
    constexpr int NV = 7, NT = 4 * NV, NC = NT + 6;
    extern __shared__ int smem[];
    const int t = xcd_tile(blockIdx.x, g.ntiles);
    const int tx0 = (t % g.ntx) * g.tw, ty0 = (t / g.ntx) * g.th;
    const int ww = g.tw + 1, wh = g.th + 1, ncw = ww * wh;
    const unsigned nu = (unsigned)n;
    float acc[NC];
    cell_accumulate<2, NC, true>(
        g, te, t, smem, pc.cap, acc,
        [&](int k, int, int, const float(&wx)[2], const float(&wy)[2], float om, float(&a)[NC]) {
            const unsigned vi = (unsigned)k / nu, m = (unsigned)k - vi * nu;
            const float *vp = views + (size_t)vi * 3 * n + m;
            const float v[NV] = {om * vp[0], om * vp[n], om * vp[2 * n], om, vgrad[(size_t)vi * 2 * n + m],
                                 vgrad[(size_t)vi * 2 * n + n + m], om};
            const float tl = wy[0] * wx[0], tr = wy[0] * wx[1], bl = wy[1] * wx[0], br = wy[1] * wx[1];
            const float tap[4] = {tl, tr, bl, br};
#pragma unroll
            for (int k2 = 0; k2 < 4; k2++)
#pragma unroll
                for (int q = 0; q < NV; q++) a[k2 * NV + q] += (q == 6 ? tap[k2] * tap[k2] : tap[k2]) * v[q];
            a[NT + 0] += om * tl * tr;  // E of TL
            a[NT + 1] += om * bl * br;  // E of BL
            a[NT + 2] += om * tl * bl;  // S of TL
            a[NT + 3] += om * tr * br;  // S of TR
            a[NT + 4] += om * tl * br;  // SE of TL
            a[NT + 5] += om * tr * bl;  // SW of TR
        });
    float *part = reinterpret_cast<float *>(smem + 2 * wh + 2);
    const int c = threadIdx.x;
    if (c < ncw) {
#pragma unroll
        for (int k = 0; k < NC; k++) part[k * ncw + c] = acc[k];
    }
    __syncthreads();
    const size_t N = (size_t)g.W * g.H;
    float *gD = gram, *gE = gram + N, *gS = gE + N, *gSE = gS + N, *gSW = gSE + N;
    for (int p = threadIdx.x; p < g.tw * g.th; p += blockDim.x) {
        const int lx = p % g.tw, ly = p / g.tw;
        const int x = tx0 + lx, y = ty0 + ly;
        if (x >= g.W || y >= g.H) continue;
        // the cell whose top-left tap is (x - dx, y - dy) sits at local (lx + 1 - dx, ly + 1 - dy);
        // there p is tap (i = dy, j = dx)
        const int c00 = (ly + 1) * ww + lx + 1, c10 = c00 - 1, c01 = c00 - ww, c11 = c00 - ww - 1;
        float sum[NV];
#pragma unroll
        for (int q = 0; q < NV; q++)
            sum[q] = part[(0 * NV + q) * ncw + c00] + part[(1 * NV + q) * ncw + c10] +
                     part[(2 * NV + q) * ncw + c01] + part[(3 * NV + q) * ncw + c11];
        const size_t o = (size_t)y * g.W + x;
        const float wsum = sum[3];
        for (int ch = 0; ch < 3; ch++) {
            U[ch * N + o] = (wsum != 0.0f) ? sum[ch] / wsum : sum[ch];
            num[ch * N + o] = sum[ch];
        }
        norm_mask[o] = wsum;
        const float nm = coverage[o];
        float gx = sum[4], gy = sum[5], tm = sum[3];
        if (nm != 0.0f) {
            gx /= nm;
            gy /= nm;
            tm /= nm;
        }
        ugx[o] = gx;
        ugy[o] = gy;
        tmask[o] = tm * sigma2;
        gD[o] = sum[6];
        gE[o] = part[(NT + 0) * ncw + c00] + part[(NT + 1) * ncw + c01];
        gS[o] = part[(NT + 2) * ncw + c00] + part[(NT + 3) * ncw + c10];
        gSE[o] = part[(NT + 4) * ncw + c00];
        gSW[o] = part[(NT + 5) * ncw + c10];
    }
}

// Target gradient (init_u_gradient, gradientIBR.cu:858-913): pull (gx, gy,
// omega, 1) with the footprint weights, normalise by the last plane.
template <int KS>
__global__ void __launch_bounds__(kPullMaxThreads) k_pull_u_gradient(TileGeom g, TileEntries te, PullCfg pc, size_t n,
                                                         const float *__restrict__ vgrad, float sigma2,
                                                         float *__restrict__ ugx, float *__restrict__ ugy,
                                                         float *__restrict__ tmask, float *__restrict__ norm_mask)
{
    PULL_PROLOGUE()
    splat_pull<KS, 4, 0>(
        g, te, t, smem, pc.cap,
        PULL_VALS(4) {
            const int k = te.src[e];
            const unsigned vi = (unsigned)k / (unsigned)n, m = (unsigned)k - vi * (unsigned)n;
            v[0] = vgrad[vi * 2 * n + m];
            v[1] = vgrad[vi * 2 * n + n + m];
            v[2] = om;
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
__global__ void __launch_bounds__(kPullMaxThreads) k_pull_reg_weight(TileGeom g, TileEntries te, PullCfg pc,
                                                         const float *__restrict__ coverage, float sigma2,
                                                         float lmax, float lmin, float nviews,
                                                         float *__restrict__ norm_mask, float *__restrict__ graw)
{
    PULL_PROLOGUE()
    splat_pull<KS, 1, 0>(
        g, te, t, smem, pc.cap, PULL_VALS(1) { v[0] = om; },
        [&](size_t o, int, int, const float(&a)[1]) {
            const float nm = norm_mask[o] + coverage[o];
            norm_mask[o] = nm;
            float m = a[0];
            if (nm != 0.0f) m /= nm;
            m *= sigma2;
            graw[o] = (m == 0.0f) ? lmax : fmaxf(lmin, lmax - m / nviews);
        });
}

// D^T g at pixel o (initConjGrad, solver.cu:384-467).
__device__ __forceinline__ float dtg_at(const float *__restrict__ gx, const float *__restrict__ gy, size_t o, int x,
                                        int y, int W, int H)
{
    float v = 0.0f;
    if (x > 0) v += gx[o - 1];
    if (x < W - 1) v -= gx[o];
    if (y > 0) v += gy[o - W];
    if (y < H - 1) v -= gy[o];
    return v;
}

__device__ __forceinline__ int dtd_degree(int x, int y, int W, int H)
{
    return (W > 1 ? ((x == 0 || x == W - 1) ? 1 : 2) : 0) + (H > 1 ? ((y == 0 || y == H - 1) ? 1 : 2) : 0);
}

// CG right-hand side: b = alpha sum B^T (omega sigma^2 v) + c0 u0 + beta D^T g.
template <int KS>
__global__ void __launch_bounds__(kPullMaxThreads) k_pull_rhs(TileGeom g, TileEntries te, PullCfg pc, size_t n,
                                                  const float *__restrict__ views, float wscale, float c0,
                                                  const float *__restrict__ u0, float beta,
                                                  const float *__restrict__ gx, const float *__restrict__ gy,
                                                  float *__restrict__ b)
{
    PULL_PROLOGUE()
    const size_t N = (size_t)g.W * g.H;
    splat_pull<KS, 3, 0>(
        g, te, t, smem, pc.cap,
        PULL_VALS(3) {
            const int k = te.src[e];
            const unsigned vi = (unsigned)k / (unsigned)n, m = (unsigned)k - vi * (unsigned)n;
            const float ow = om * wscale;
            const float *vp = views + vi * 3 * n + m;
            v[0] = ow * vp[0];
            v[1] = ow * vp[n];
            v[2] = ow * vp[2 * n];
        },
        [&](size_t o, int x, int y, const float(&a)[3]) {
            const float dtg = beta != 0.0f ? beta * dtg_at(gx, gy, o, x, y, g.W, g.H) : 0.0f;
            for (int c = 0; c < 3; c++) b[c * N + o] = a[c] + (c0 != 0.0f ? c0 * u0[c * N + o] : 0.0f) + dtg;
        });
}

// Jacobi preconditioner 1 / diag(A), diag = alpha sum omega sigma^2 B^2 + c0 + beta deg.
template <int KS>
__global__ void __launch_bounds__(kPullMaxThreads) k_pull_diag(TileGeom g, TileEntries te, PullCfg pc, float wscale, float c0,
                                                   float beta, float *__restrict__ minv)
{
    PULL_PROLOGUE()
    splat_pull<KS, 1, 1>(
        g, te, t, smem, pc.cap, PULL_VALS(1) { v[0] = om * wscale; },
        [&](size_t o, int x, int y, const float(&a)[1]) {
            const float d = a[0] + c0 + beta * (float)dtd_degree(x, y, g.W, g.H);
            minv[o] = d > 0.0f ? 1.0f / d : 1.0f;
        });
}

// Right-hand side and preconditioner from the sums kept by the set-up pull:
// b = wscale num + c0 u0 + beta D^T g, Minv = 1 / (wscale sq + c0 + beta deg).
__global__ void k_rhs_diag_cached(int W, int H, float wscale, float c0, const float *__restrict__ u0, float beta,
                                  const float *__restrict__ gx, const float *__restrict__ gy,
                                  const float *__restrict__ num, const float *__restrict__ sq, int precond,
                                  float *__restrict__ b, float *__restrict__ minv)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    const size_t N = (size_t)W * H, o = (size_t)y * W + x;
    // neighbours loaded unconditionally (clamped), masked afterwards
    const float gxl = gx[x > 0 ? o - 1 : o], gxc = gx[o], gyu = gy[y > 0 ? o - W : o], gyc = gy[o];
    const float sqv = sq[o];
    float nm[3], uv[3] = {0.0f, 0.0f, 0.0f};
    for (int c = 0; c < 3; c++) {
        nm[c] = num[c * N + o];
        if (c0 != 0.0f) uv[c] = u0[c * N + o];
    }
    const float dtg = (x > 0 ? gxl : 0.0f) - (x < W - 1 ? gxc : 0.0f) + (y > 0 ? gyu : 0.0f) - (y < H - 1 ? gyc : 0.0f);
    const float bdt = beta != 0.0f ? beta * dtg : 0.0f;
    for (int c = 0; c < 3; c++) b[c * N + o] = wscale * nm[c] + (c0 != 0.0f ? c0 * uv[c] : 0.0f) + bdt;
    if (precond) {
        const float d = sqv * wscale + c0 + beta * (float)dtd_degree(x, y, W, H);
        minv[o] = d > 0.0f ? 1.0f / d : 1.0f;
    } else {
        minv[o] = 1.0f;
    }
}

#undef PULL_PROLOGUE
#undef PULL_VALS

__global__ void k_fill(float *__restrict__ p, size_t N, float v)
{
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o < N) p[o] = v;
}

// Visibility from the coverage: GRADIENT_IBR marks any footprint hit
// (coverage > 0, setup_visibility_mask gradientIBR_kernels.cuh:4-54 sets the
// flag for dx*dy != 0 and the normalised weights are positive exactly there);
// COLOR_SR requires coverage >= 1 (vtv_sr_kernels_unstructured.cu:76-93).
// Visibility from the coverage (IBR_color: coverage >= 1; IBR_gradient: any
// splat).  IBR_gradient also lists the invisible pixels (the holes) into
// `list` / `count`, and clears `count_next` for the following pass.
__global__ void k_visibility(size_t N, const float *__restrict__ coverage, int color, float *__restrict__ vis,
                             int *__restrict__ list, unsigned int *__restrict__ count,
                             unsigned int *__restrict__ count_next)
{
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (list && o == 0) *count_next = 0u;
    if (o >= N) return;
    const float c = coverage[o];
    const float v = color ? (c < 1.0f ? 0.0f : 1.0f) : (c > 0.0f ? 1.0f : 0.0f);
    vis[o] = v;
    if (list && v == 0.0f) list[atomicAdd(count, 1u)] = (int)o;
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
    t.omega = ctx->omega;
    t.pay = ctx->e_pay;
    t.pay_stride = ctx->e_cap;
    t.active = 1;
    return t;
}

// Capacity of a pull's LDS cache within `budget` bytes besides `extra` bytes:
// staged entries for the cell pull (ks <= 3), slots for the slot pull.
int pull_cap(const lfrv_ctx *ctx, int nv, size_t extra, size_t budget, PullKind kind, int npay)
{
    if (kind != PullKind::Slot) {
        const int wh = ctx->tile_h + ctx->ks - 1;
        const long avail = (long)budget - (long)extra - (long)sizeof(int) * (2 * wh + 2);
        const long ecap = avail / (long)((3 + npay) * sizeof(float));
        // the box pull reads what does not fit in place, so it stages at most
        // a few times a window's typical entry count (more blocks per CU)
        const long lim = kind == PullKind::Box ? std::min(2048L, (long)kStagePer * ctx->pull_bs)
                                               : (long)kStagePer * ctx->pull_bs;
        return (int)std::max(0L, std::min(ecap, lim));
    }
    const size_t meta = sizeof(int) * (size_t)pull_meta_ints(ctx->tile_w, ctx->tile_h, ctx->ks, ctx->pull_bs);
    const long avail = (long)budget - (long)meta - (long)extra;
    const long cap = avail / (long)(sizeof(float) * (nv + 2));
    return (int)std::max(64L, std::min(cap, 16384L));
}

size_t pull_lds(const lfrv_ctx *ctx, int nv, int cap, PullKind kind, int npay)
{
    if (kind != PullKind::Slot) {
        const int wh = ctx->tile_h + ctx->ks - 1;
        const size_t cells = (size_t)(ctx->tile_w + ctx->ks - 1) * wh;
        const size_t part = (kind == PullKind::Box ? 3 : (size_t)ctx->ks * ctx->ks) * nv * cells;
        return sizeof(int) * (size_t)(2 * wh + 2) + sizeof(float) * std::max(part, (size_t)(3 + npay) * cap);
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


lfrv_status run_prepare(lfrv_ctx *ctx)
{
    const TileGeom g = tile_geom(ctx);
    const int nc = ctx->ncells;
    TimedRegion tr(ctx, 5);
    // the views' stencils depend on the views only: on the side stream,
    // concurrent with the splat structure, joined before the entry payload
    LFRV_HIP(ctx, hipEventRecord(ctx->side_fork, ctx->stream));
    LFRV_HIP(ctx, hipStreamWaitEvent(ctx->side, ctx->side_fork, 0));
    lfrv_status vs = run_view_stencils(ctx, ctx->side);
    if (vs != LFRV_OK) return vs;
    LFRV_HIP(ctx, hipEventRecord(ctx->side_join, ctx->side));
    const dim3 bgrid((ctx->w + 63) / 64, (ctx->h + 3) / 4), bblock(64, 4);
    if (!ctx->counts_valid) {
        // (lfrv_warps_from_depth counts the cells of the warps it produces)
        LFRV_HIP(ctx, hipMemsetAsync(ctx->cell_count, 0, sizeof(int) * (nc + 1), ctx->stream));
#define CC(K) k_cell_count<K><<<bgrid, bblock, 0, ctx->stream>>>(ctx->V, ctx->w, ctx->h, ctx->W, ctx->H, g.Wc, ctx->tau_x, ctx->tau_y, ctx->cell_count)
        LFRV_KS_DISPATCH(ctx->ks, CC);
#undef CC
        ctx->counts_valid = true;
    }
    const int nb = (nc + kScanBlock - 1) / kScanBlock;
    k_scan_reduce<<<nb, 256, 0, ctx->stream>>>(nc, ctx->cell_count, ctx->scan_tmp);
    k_scan_apply<<<nb, 256, 0, ctx->stream>>>(nc, ctx->cell_count, ctx->scan_tmp, ctx->cell_start, ctx->cell_cursor);
#define CF(K) k_cell_fill<K><<<bgrid.x * bgrid.y, bblock, 0, ctx->stream>>>(ctx->V, ctx->w, ctx->h, ctx->W, ctx->H, g.Wc, ctx->tau_x, ctx->tau_y, ctx->cell_cursor, ctx->e_src)
    LFRV_KS_DISPATCH(ctx->ks, CF);
#undef CF
    // dsf = 1: the sort also forms the per-cell tap sums of the coverage
    // (scratch: the single-reduction CG buffers, idle here)
    float *cell_tap = (ctx->ks == 2 && ctx->cg_extra) ? ctx->cg_extra : nullptr;
    // the mean entries per cell, from the source pixels (invalid ones drop out)
    const int cpb = sort_cells_per_block((double)ctx->V * (double)ctx->n / (double)nc);
    const int sort_blocks = (nc + cpb - 1) / cpb;
#define CS(K) k_cell_sort<K><<<sort_blocks, kSortCells, 0, ctx->stream>>>(nc, cpb, ctx->W, ctx->H, ctx->cell_start, ctx->e_src, ctx->tau_x, ctx->tau_y, ctx->e_tx, ctx->e_ty, cell_tap)
    LFRV_KS_DISPATCH(ctx->ks, CS);
#undef CS
    LFRV_LAUNCH(ctx, "splat structure");
    if (cell_tap) {
        k_coverage_from_taps<2><<<dim3((ctx->W + 63) / 64, (ctx->H + 3) / 4), dim3(64, 4), 0, ctx->stream>>>(
            ctx->W, ctx->H, g.Wc, nc, cell_tap, ctx->coverage);
    } else {
        const TileEntries te = tile_entries(ctx);
        const PullKind pk = splat_kind(ctx, false);
        PullCfg pc{pull_cap(ctx, 1, 0, ctx->pull_budget, pk)};
#define COV(K) k_pull_coverage<K><<<ctx->ntiles, ctx->pull_bs, pull_lds(ctx, 1, pc.cap, pk), ctx->stream>>>(g, te, pc, ctx->coverage)
        LFRV_KS_DISPATCH(ctx->ks, COV);
#undef COV
    }
    LFRV_LAUNCH(ctx, "coverage");
    LFRV_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->side_join, 0));
    if (ctx->e_pay) {
        k_entry_payload<<<blocks_for(ctx->e_cap), kBlock, 0, ctx->stream>>>(
            (int64_t)ctx->e_cap, ctx->cell_start + ctx->ncells, ctx->e_src, ctx->n, ctx->views, ctx->view_lap,
            ctx->e_pay);
        LFRV_LAUNCH(ctx, "entry payload");
    }
    ctx->prepared = true;
    ctx->entry_weights_valid = false;
    ctx->epoch++;
    return LFRV_OK;
}

lfrv_status run_refresh_entry_weights(lfrv_ctx *ctx, float)
{
    if (ctx->entry_weights_valid) return LFRV_OK;
    k_refresh_weights<<<blocks_for(ctx->e_cap), kBlock, 0, ctx->stream>>>(
        (int64_t)ctx->e_cap, ctx->cell_start + ctx->ncells, ctx->e_src, ctx->omega, ctx->e_w);
    LFRV_LAUNCH(ctx, "refresh weights");
    ctx->entry_weights_valid = true;
    return LFRV_OK;
}

lfrv_status run_visibility(lfrv_ctx *ctx)
{
    if (!ctx->prepared) return fail(ctx, LFRV_E_STATE, "lfrv_setup_visibility before lfrv_prepare");
    const int color = ctx->p.mode == LFRV_MODE_COLOR_SR;
    unsigned int *cnt = ctx->counters + 8 + ctx->hole_parity, *cnt_next = ctx->counters + 8 + (ctx->hole_parity ^ 1);
    k_visibility<<<blocks_for(ctx->N), kBlock, 0, ctx->stream>>>(ctx->N, ctx->coverage, color, ctx->visibility,
                                                                 color ? nullptr : ctx->hole_list, cnt, cnt_next);
    if (color)  // the reference splats into its normalisation buffer without clearing it (optimization.cu:108-117)
        k_add_to<<<blocks_for(ctx->N), kBlock, 0, ctx->stream>>>(ctx->N, ctx->coverage, ctx->norm_mask);
    LFRV_LAUNCH(ctx, "visibility");
    ctx->holes_listed = false;
    if (!color) {
        LFRV_HIP(ctx, hipMemcpyAsync(ctx->hole_count_host, cnt, sizeof(unsigned int), hipMemcpyDeviceToHost,
                                     ctx->stream));
        LFRV_HIP(ctx, hipEventRecord(ctx->hole_ev, ctx->stream));
        ctx->hole_parity ^= 1;
        ctx->holes_listed = true;
    }
    return LFRV_OK;
}

lfrv_status run_initial_image(lfrv_ctx *ctx)
{
    if (!ctx->prepared) return fail(ctx, LFRV_E_STATE, "initial image before lfrv_prepare");
    const TileGeom g = tile_geom(ctx);
    const TileEntries te = tile_entries(ctx);
    TimedRegion tr(ctx, 5);
    if (ctx->p.mode == LFRV_MODE_GRADIENT_IBR && ctx->ks == 2 && ctx->gram) {
        // one pass over the entries: initial image, target gradient, CG sums
        // and the CG stencil (omega read through the source index)
        TileEntries tes = te;
        PullCfg pc{pull_cap(ctx, 9, 0, ctx->pull_budget, PullKind::Cell)};  // 4 * 9 = 36 >= 34 sums per window cell
        const float s2 = ctx->p.sigma_sensor * ctx->p.sigma_sensor;
        k_gibr_setup_st2<<<ctx->ntiles, ctx->pull_bs, pull_lds(ctx, 9, pc.cap, PullKind::Cell), ctx->stream>>>(
            g, tes, pc, ctx->n, ctx->views, ctx->view_grad, ctx->coverage, s2, ctx->U, ctx->norm_mask,
            ctx->u_grad_x, ctx->u_grad_y, ctx->target_mask, ctx->cache_num, ctx->gram);
        LFRV_LAUNCH(ctx, "set-up pull");
        ctx->setup_epoch = ctx->epoch;
        ctx->gram_epoch = ctx->epoch;
        return LFRV_OK;
    }
    lfrv_status st = run_refresh_entry_weights(ctx, 1.0f);
    if (st != LFRV_OK) return st;
    if (ctx->p.mode == LFRV_MODE_GRADIENT_IBR) {
        // one pull for the initial image, the target gradient and the CG sums
        const PullKind pk = splat_kind(ctx, true);
        PullCfg pc{pull_cap(ctx, 7, 0, ctx->pull_budget, pk)};
        const float s2 = ctx->p.sigma_sensor * ctx->p.sigma_sensor;
#define SETUP(K) k_pull_gibr_setup<K><<<ctx->ntiles, ctx->pull_bs, pull_lds(ctx, 7, pc.cap, pk), ctx->stream>>>(g, te, pc, ctx->n, ctx->views, ctx->view_grad, ctx->coverage, s2, ctx->U, ctx->norm_mask, ctx->u_grad_x, ctx->u_grad_y, ctx->target_mask, ctx->cache_num, ctx->cache_sq)
        LFRV_KS_DISPATCH(ctx->ks, SETUP);
#undef SETUP
        LFRV_LAUNCH(ctx, "set-up pull");
        ctx->setup_epoch = ctx->epoch;
        return LFRV_OK;
    }
    const PullKind pk = splat_kind(ctx, false);
    PullCfg pc{pull_cap(ctx, 4, 0, ctx->pull_budget, pk)};
#define INIT(K) k_pull_init_image<K><<<ctx->ntiles, ctx->pull_bs, pull_lds(ctx, 4, pc.cap, pk), ctx->stream>>>(g, te, pc, ctx->n, ctx->views, ctx->U, ctx->norm_mask)
    LFRV_KS_DISPATCH(ctx->ks, INIT);
#undef INIT
    LFRV_LAUNCH(ctx, "initial image");
    return LFRV_OK;
}

// True when the set-up pull ran on the current entries, weights and views.
static bool setup_cache_valid(const lfrv_ctx *ctx)
{
    return ctx->setup_epoch == ctx->epoch && ctx->prepared;
}

lfrv_status run_u_gradient(lfrv_ctx *ctx)
{
    if (!ctx->prepared) return fail(ctx, LFRV_E_STATE, "lfrv_init_u_gradient before lfrv_prepare");
    if (setup_cache_valid(ctx)) {
        // g and the target mask came out of the set-up pull; the reference
        // leaves sum B^T 1 in its normalisation buffer here
        LFRV_HIP(ctx, hipMemcpyAsync(ctx->norm_mask, ctx->coverage, ctx->N * sizeof(float), hipMemcpyDeviceToDevice,
                                     ctx->stream));
        return LFRV_OK;
    }
    lfrv_status st = run_refresh_entry_weights(ctx, 1.0f);
    if (st != LFRV_OK) return st;
    const TileGeom g = tile_geom(ctx);
    const TileEntries te = tile_entries(ctx);
    const PullKind pk = splat_kind(ctx, false);
    PullCfg pc{pull_cap(ctx, 4, 0, ctx->pull_budget, pk)};
    const float s2 = ctx->p.sigma_sensor * ctx->p.sigma_sensor;
    TimedRegion tr(ctx, 5);
#define UG(K) k_pull_u_gradient<K><<<ctx->ntiles, ctx->pull_bs, pull_lds(ctx, 4, pc.cap, pk), ctx->stream>>>(g, te, pc, ctx->n, ctx->view_grad, s2, ctx->u_grad_x, ctx->u_grad_y, ctx->target_mask, ctx->norm_mask)
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
    const PullKind pk = splat_kind(ctx, false);
    PullCfg pc{pull_cap(ctx, 1, 0, ctx->pull_budget, pk)};
    const float s2 = ctx->p.sigma_sensor * ctx->p.sigma_sensor;
    const float lmax = ctx->p.lambda_max_factor * ctx->p.lambda, lmin = ctx->p.lambda;
    float *graw = ctx->tmpN;
#define RW(K) k_pull_reg_weight<K><<<ctx->ntiles, ctx->pull_bs, pull_lds(ctx, 1, pc.cap, pk), ctx->stream>>>(g, te, pc, ctx->coverage, s2, lmax, lmin, (float)ctx->V, ctx->norm_mask, graw)
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
    if (use_views && setup_cache_valid(ctx)) {
        k_rhs_diag_cached<<<dim3((ctx->W + 63) / 64, (ctx->H + 3) / 4), dim3(64, 4), 0, ctx->stream>>>(
            ctx->W, ctx->H, wscale, (float)c0, u0, (float)beta, ctx->u_grad_x, ctx->u_grad_y, ctx->cache_num,
            ctx->ks == 2 && ctx->gram ? ctx->gram : ctx->cache_sq,  // dsf = 1: sq is the stencil diagonal
            use_precond ? 1 : 0, ctx->cg_b, ctx->cg_minv);
        LFRV_LAUNCH(ctx, "rhs / diag");
        return LFRV_OK;
    }
    if (use_views) {
        lfrv_status st = run_refresh_entry_weights(ctx, 1.0f);
        if (st != LFRV_OK) return st;
    }
    {
        const PullKind pk = splat_kind(ctx, false);
        PullCfg pc{pull_cap(ctx, 3, 0, ctx->pull_budget, pk)};
#define RHS(K) k_pull_rhs<K><<<ctx->ntiles, ctx->pull_bs, pull_lds(ctx, 3, pc.cap, pk), ctx->stream>>>(g, te, pc, ctx->n, ctx->views, wscale, (float)c0, u0, (float)beta, ctx->u_grad_x, ctx->u_grad_y, ctx->cg_b)
        LFRV_KS_DISPATCH(ctx->ks, RHS);
#undef RHS
    }
    if (use_precond) {
        const PullKind pk = splat_kind(ctx, true);
        PullCfg pc{pull_cap(ctx, 1, 0, ctx->pull_budget, pk)};
#define DG(K) k_pull_diag<K><<<ctx->ntiles, ctx->pull_bs, pull_lds(ctx, 1, pc.cap, pk), ctx->stream>>>(g, te, pc, wscale, (float)c0, (float)beta, ctx->cg_minv)
        LFRV_KS_DISPATCH(ctx->ks, DG);
#undef DG
    } else {
        k_fill<<<blocks_for(ctx->N), kBlock, 0, ctx->stream>>>(ctx->cg_minv, ctx->N, 1.0f);
    }
    LFRV_LAUNCH(ctx, "rhs / diag");
    return LFRV_OK;
}

}  // namespace lfrv
