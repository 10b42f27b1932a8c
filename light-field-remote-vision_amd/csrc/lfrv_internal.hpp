// Internal declarations of the lfrv HIP library (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/lfrv.h"

namespace lfrv {

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

constexpr int kMaxKs = 5;       // dsf <= 4
constexpr int kBlock = 256;     // threads per block of the elementwise kernels
constexpr int kNumXcd = 8;      // MI355X: 8 XCDs, blocks are dealt round-robin
constexpr int kPullMaxThreads = 512;  // threads of the pull kernels at most (window cells, tile pixels)

// Map a linear block id to a tile id so that each XCD receives a contiguous
// run of tiles (neighbouring tiles then share that XCD's L2 for their halos).
__device__ __forceinline__ int xcd_tile(int bid, int ntiles)
{
    const int q = ntiles / kNumXcd, r = ntiles % kNumXcd;
    const int x = bid % kNumXcd, k = bid / kNumXcd;
    return (x < r) ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// Gather validity (gradientIBR_kernels.cuh:216-219).
__host__ __device__ __forceinline__ bool gather_valid(float tx, float ty, int W, int H)
{
    return !(tx < 0.0f || ty < 0.0f || tx > (float)W || ty > (float)H);
}

// Splat validity = membership of the reference's cell list
// (gradientIBR.cu:404-426): gather-valid and floor(tau - .5) inside the target.
__host__ __device__ __forceinline__ bool cell_valid(float tx, float ty, int W, int H)
{
    if (!gather_valid(tx, ty, W, H)) return false;
    const int px = (int)floorf(tx - 0.5f), py = (int)floorf(ty - 0.5f);
    return !(px < 0 || py < 0 || px > W - 1 || py > H - 1);
}

// Cell of a cell-valid warped pixel in the extended grid of (W+KS-1) x
// (H+KS-1) cells: the cell of footprint origin (x0, y0) is (x0+KS-1, y0+KS-1).
template <int KS>
__device__ __forceinline__ int cell_of(float t0, float t1, int Wc)
{
    constexpr float hd = 0.5f * (float)(KS - 1);
    const int x0 = (int)floorf(t0 - 0.5f - hd + 0.5f);
    const int y0 = (int)floorf(t1 - 0.5f - hd + 0.5f);
    return (y0 + KS - 1) * Wc + (x0 + KS - 1);
}

// Footprint weights of one warped pixel (set_B_bilinear,
// gradientIBR_kernels.cuh:447-507).  The 2-D weights are wy[i]*wx[j]; both
// factors come out normalised over the in-image taps (the 2-D normaliser of
// the reference factorises).  Out-of-image taps get weight 0.  Returns false
// when every in-image tap has zero weight.
template <int KS>
__device__ __forceinline__ bool footprint(float tx, float ty, int W, int H, int &x0, int &y0,
                                          float (&wx)[KS], float (&wy)[KS])
{
    constexpr float hd = 0.5f * (float)(KS - 1);
    const float px = tx - 0.5f, py = ty - 0.5f;
    x0 = (int)floorf(px - hd + 0.5f);
    y0 = (int)floorf(py - hd + 0.5f);
    float sx = 0.0f, sy = 0.0f;
#pragma unroll
    for (int j = 0; j < KS; j++) {
        const int x = x0 + j;
        float d = 1.0f;
        if (j == 0) d = ((float)x + 0.5f) - (px - hd);
        else if (j == KS - 1) d = (px + hd) - ((float)x - 0.5f);
        d = (x >= 0 && x <= W - 1) ? d : 0.0f;
        wx[j] = d;
        sx += d;
    }
#pragma unroll
    for (int i = 0; i < KS; i++) {
        const int y = y0 + i;
        float d = 1.0f;
        if (i == 0) d = ((float)y + 0.5f) - (py - hd);
        else if (i == KS - 1) d = (py + hd) - ((float)y - 0.5f);
        d = (y >= 0 && y <= H - 1) ? d : 0.0f;
        wy[i] = d;
        sy += d;
    }
    const float nrm = sx * sy;
    if (nrm == 0.0f) return false;
    const float isx = 1.0f / sx, isy = 1.0f / sy;
#pragma unroll
    for (int j = 0; j < KS; j++) wx[j] *= isx;
#pragma unroll
    for (int i = 0; i < KS; i++) wy[i] *= isy;
    return true;
}

// v of the lane `SHIFT` below in the same row of 16 lanes, 0 for the first
// SHIFT lanes of a row (DPP row_shr on both halves of the double).
template <int SHIFT>
__device__ __forceinline__ double dpp_row_shr(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x110 + SHIFT, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x110 + SHIFT, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Wave-level FP64 sum (wave64), fixed order.  With every lane active the rows
// of 16 lanes are summed by DPP shifts (lane 15 of a row ends with the row's
// sum) and the four row sums added in every lane; a partial wave (a block
// size that is not a multiple of 64) falls back to shuffles.  The sum is valid
// in lane 0 at least.
__device__ __forceinline__ double wave_sum(double v)
{
    if (__builtin_amdgcn_read_exec() == ~0ull) {
        v += dpp_row_shr<1>(v);
        v += dpp_row_shr<2>(v);
        v += dpp_row_shr<4>(v);
        v += dpp_row_shr<8>(v);
        return (__shfl(v, 15, 64) + __shfl(v, 31, 64)) + (__shfl(v, 47, 64) + __shfl(v, 63, 64));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Block-level sum through LDS: every thread must call it; on return every
// thread holds the sums (per-wave sums added by thread q in wave order).
template <int NQ>
__device__ __forceinline__ void block_sum(double (&v)[NQ], double *lds /* >= NQ * 16 */)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const double s = wave_sum(v[q]);
        if (lane == 0) lds[q * 16 + wid] = s;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        const int q = threadIdx.x;
        double s = 0.0;
        for (int k = 0; k < nw; k++) s += lds[q * 16 + k];
        lds[q * 16] = s;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; q++) v[q] = lds[q * 16];
    __syncthreads();
}

// Grid-wide reductions without device-scope fences.  The L2 caches of the
// XCDs are not coherent with each other, so a device-scope release after
// every block (the classic last-block pattern) writes back an L2 per block:
// measured on the 1024 x 1024 PCG iteration (2048 blocks), that pattern made
// the pass 80 us long, against 32 us plus an 11 us finalize launch.
// Instead each block stores its NQ partial sums (layout [q][block], plain
// stores) and a one-block finalize kernel launched after it on the same
// stream (the kernel boundary orders the stores) reduces them in block order.
// The per-wave sums of value q are added by thread q (in wave order) and
// stored by it, so no thread walks all NQ columns.
template <int NQ>
__device__ __forceinline__ void block_partials(double (&v)[NQ], double *partials, double *lds /* >= NQ * 16 */)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const double s = wave_sum(v[q]);
        if (lane == 0) lds[q * 16 + wid] = s;
    }
    __syncthreads();
    const int q = threadIdx.x;
    if (q < NQ) {
        double s = 0.0;
        for (int k = 0; k < nw; k++) s += lds[q * 16 + k];
        partials[(size_t)q * gridDim.x + blockIdx.x] = s;
    }
}

// block_partials with an explicit slot and stride: value q of this block goes
// to partials[q * G + slot].
template <int NQ>
__device__ __forceinline__ void block_partials_at(double (&v)[NQ], double *partials, double *lds, int slot, int G)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const double s = wave_sum(v[q]);
        if (lane == 0) lds[q * 16 + wid] = s;
    }
    __syncthreads();
    const int q = threadIdx.x;
    if (q < NQ) {
        double s = 0.0;
        for (int k = 0; k < nw; k++) s += lds[q * 16 + k];
        partials[(size_t)q * G + slot] = s;
    }
}

// Deterministic sum of G block partials (layout [q][block]) by one block;
// every thread holds the result.  Up to 2 x blocks-size partials, all
// loads of a thread are independent and in flight at once (the partials come
// from other XCDs' L2s, so a dependent load chain would pay the full memory
// latency per element); beyond that, four blocks' worth of every value per
// round of loads.
template <int NQ>
__device__ __forceinline__ void reduce_partials(const double *partials, int G, double (&out)[NQ], double *lds)
{
    const int bs = blockDim.x, t = threadIdx.x;
    if (G <= bs) {
#pragma unroll
        for (int q = 0; q < NQ; q++) out[q] = t < G ? partials[(size_t)q * G + t] : 0.0;
    } else if (G <= 2 * bs) {
        double hi[NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            out[q] = partials[(size_t)q * G + t];
            hi[q] = t + bs < G ? partials[(size_t)q * G + t + bs] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < NQ; q++) out[q] += hi[q];
    } else {
        // four blocks' worth of every quantity in flight at once
#pragma unroll
        for (int q = 0; q < NQ; q++) out[q] = 0.0;
        int b = t;
        for (; b + 3 * bs < G; b += 4 * bs) {
            double x[4][NQ];
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int q = 0; q < NQ; q++) x[k][q] = partials[(size_t)q * G + b + k * bs];
#pragma unroll
            for (int q = 0; q < NQ; q++) out[q] += (x[0][q] + x[1][q]) + (x[2][q] + x[3][q]);
        }
        for (; b < G; b += bs)
#pragma unroll
            for (int q = 0; q < NQ; q++) out[q] += partials[(size_t)q * G + b];
    }
    block_sum<NQ>(out, lds);
}

// Threads of the one-block finalize kernels.  A finalize after a full-GPU pass
// is latency-bound (an empty one-block launch there takes 4.5 us on MI355X):
// 256 threads with four strided loads each beat 1024 with one or two.
constexpr int kFinalizeThreads = 256;
constexpr int kSrOpsPlanes = 13;        // planes of the assembled IBR_color data term

// Finalize kernel of the plain sums.
template <int NQ>
__global__ void __launch_bounds__(kFinalizeThreads) k_reduce_out(const double *__restrict__ partials, int G, double *out)
{
    __shared__ double red[NQ * 16];
    double v[NQ];
    reduce_partials<NQ>(partials, G, v, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; q++) out[q] = v[q];
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

// Per-view camera data for the depth warp, in FP32 (built in FP64 on the host).
struct WarpCam {
    float M[9];   // K_u R_u R_v^T K_v^-1 (import.cpp:301-320 composed)
    float c[3];   // P_u (C_v, 1)
    float Nb[9];  // P_v3 R_b^T K_u^-1 (back-projection of a target pixel, import.cpp:412-430)
    float e[3];   // P_v (C_b, 1)
};

// CG scalars, device resident (one problem, three RGB channels).
struct CgState {
    double rz[3], dq[3], rr[3], bb[3];
    double alpha[3], beta[3];
    double tol2;
    int active[3];
    int iter[3];
    int first;    // the next operator launch builds d = z (no previous direction)
    int done;
    int maxit;
    int breakdown[3];
    double resid[3];  // true residual sums of the final iterate (filled with each poll)
};

// Device state of the IBR_color FISTA loop (the stopping rules of
// super-resolution.cpp:215-252 evaluated on the device, one step at a time).
struct SrLoop {
    int halt;    // a rule fired: the steps enqueued after it return at once
    int steps;   // FISTA steps run in this pass
    int stop;    // LFRV_SR_STOP_* of the pass
    int cur;     // which of the two previous-solution buffers holds the last step the rules kept
    float energy, prev_energy;  // float, as the driver keeps them
    double parts[3];            // E_TV, E_int, E_grad of the last energy
};

// The five coefficient planes of the assembled dsf = 1 stencil and the
// arguments of one single-reduction PCG pass over it (k_cg.hip).
struct St2Coeffs {
    const float *gD, *gE, *gS, *gSE, *gSW;
};

struct CgcgArgs {
    int W, H, ntx, ntiles;
    float wscale, c0, beta_reg;
    const float *r_in, *w_in, *s_in;
    float *r_out, *w_out, *s_out, *p, *x;
    const float *minv;
    St2Coeffs g;
};


// One solve of a batch of PCG solves (lfrv_batch): its iteration arguments for even and odd iterations
// (r, w, s swap buffers), its device scalars and its partial sums.
struct CgBatchItem {
    CgcgArgs a[2];
    CgState *st;
    double *partials;
};

struct TimingSlot {
    double ms = 0.0;
    int64_t launches = 0;
};

}  // namespace lfrv

struct lfrv_ctx {
    lfrv_params p{};
    int V = 0, w = 0, h = 0, dsf = 1, ks = 2, W = 0, H = 0;
    size_t n = 0, N = 0;
    hipStream_t stream = nullptr;
    // second stream for work of lfrv_prepare that does not depend on the warps
    // (the views' stencils), forked from and joined to `stream` by two events
    hipStream_t side = nullptr;
    hipEvent_t side_fork = nullptr, side_join = nullptr;
    std::string err;

    // views (contiguous over views)
    float *views = nullptr;    // V*3n
    float *tau_x = nullptr, *tau_y = nullptr;      // V*n
    float *dpart_x = nullptr, *dpart_y = nullptr;  // V*n
    float *omega = nullptr;    // V*n
    float *depth = nullptr;    // V*n
    float *depth_grad = nullptr;  // 2*V*n dz/dx then dz/dy (lfrv_set_depth_gradient; null: zero)
    float *tau_weight = nullptr;  // V*n deformation weights of the OpenGL warp producer (allocated on first use)
    unsigned int *gl_big = nullptr;  // 1 + V*n: the OpenGL producer's count and list of big quads (first use)
    float *view_lap = nullptr; // V*3n Laplacian of the views (IBR_color gradient term)
    float *view_grad = nullptr;// V*2n Sobel of the views' luminance
    std::vector<double> P_views;  // V*12
    double P_target[12] = {0};
    bool have_target_cam = false;
    lfrv::WarpCam *dcams = nullptr;  // V per-view camera records on the device
    bool cams_valid = false;      // dcams match P_views / P_target
    std::vector<int> have_depth;

    // splat structure: entries sorted by footprint cell, target tiling of the pulls
    int tile_w = 32, tile_h = 8, ntx = 0, nty = 0, ntiles = 0;
    int ncells = 0;                       // (W + ks - 1) * (H + ks - 1)
    int pull_bs = 256;                    // threads of the pull kernels: >= window cells and tile pixels
    int *cell_count = nullptr, *cell_start = nullptr, *cell_cursor = nullptr;  // ncells(+1)
    int *scan_tmp = nullptr;              // block sums of the cell scan
    int *e_src = nullptr;       // view * n + m
    float *e_tx = nullptr, *e_ty = nullptr, *e_w = nullptr;
    float *e_pay = nullptr;  // IBR_color, dsf >= 2: per entry its view's RGB and Laplacian (6 planes)
    float *sr_ops = nullptr;   // IBR_color, dsf >= 2: the assembled data-term operator (kSrOpsPlanes anchor planes)
    uint64_t sr_ops_epoch = 0;
    size_t e_cap = 0;
    bool prepared = false;
    bool entry_weights_valid = false;
    bool counts_valid = false;  // cell_count holds the cell counts of the current warps
    // dpart_x / dpart_y not yet written for the current import warps: IBR_gradient's
    // first weights do not read them, so they are formed from the depth when a
    // reader asks (materialize_dpart)
    bool dpart_pending = false;
    uint64_t epoch = 1;        // bumped whenever entries, weights or views change
    uint64_t setup_epoch = 0;  // epoch of the last gradient-IBR set-up pull
    float *cache_num = nullptr, *cache_sq = nullptr;  // its raw sums: 3N and N
    float *gram = nullptr;     // dsf = 1: 5N raw stencil coefficients of sum_i S_i^T omega_i S_i
    float *cg_extra = nullptr; // dsf = 1: 9N, the second r, s and w buffers of the single-reduction PCG
    uint64_t gram_epoch = 0;
    float *box_ops = nullptr;  // IBR_gradient, dsf >= 2: 5 anchor planes of sum_e omega_e t_e t_e^T
    uint64_t box_epoch = 0;
    int cg_iter_hint = 8;      // iterations of the last single-reduction solve (first look at the flag)
    int pull_budget = 48 * 1024;  // LDS bytes of a pull's entry cache

    // target side (3 planes each unless noted)
    float *U = nullptr, *Uq = nullptr, *G = nullptr;
    float *PX = nullptr, *PY = nullptr;      // dual variables
    float *U2 = nullptr, *P2X = nullptr, *P2Y = nullptr;  // ping-pong partners of the PD step
    float *reg_g = nullptr;        // N (valid once has_reg_g)
    bool has_reg_g = false;
    float *norm_mask = nullptr;    // N
    float *coverage = nullptr;     // N
    float *u_grad_x = nullptr, *u_grad_y = nullptr, *target_mask = nullptr;  // N
    float *visibility = nullptr;   // N (0/1)
    float *target_depth = nullptr; // N
    float *tmpN = nullptr;         // N scratch
    float *tmp3N = nullptr;        // 3N scratch
    double fista_alpha = 1.0;
    lfrv::SrLoop *sr_loop = nullptr;       // device loop state of lfrv_sr_solve
    lfrv::SrLoop *sr_loop_host = nullptr;  // pinned, 3 slots: two alternating polls, the final read
    hipEvent_t sr_ev[2] = {nullptr, nullptr};
    // 9N (allocated on first use): two 3N buffers of the previous step's solution, double-buffered
    // (SrLoop::cur names the one the rules kept; the energy pass writes U into the other), then the
    // previous pass's 3N solution
    float *sr_prev = nullptr;

    // CG
    float *cg_r = nullptr, *cg_d = nullptr, *cg_q = nullptr, *cg_b = nullptr;  // 3N
    float *cg_minv = nullptr;      // N
    lfrv::CgState *cg_state = nullptr;
    lfrv::CgState *cg_state_host = nullptr;   // pinned
    double *partials = nullptr;    // per-block partial sums
    size_t partials_cap = 0;       // doubles
    unsigned int *counters = nullptr;  // small device counters (hole-list length)
    double *scal = nullptr;        // device scalars (reduction outputs)
    double *scal_host = nullptr;   // pinned mirror
    int *hole_list = nullptr;      // N
    // hole list built by the visibility pass (IBR_gradient): its length is
    // copied to pinned memory behind an event, so the hole filling reads it
    // without draining the stream; the two device counters alternate, each
    // pass clearing the one the next pass uses
    unsigned int *hole_count_host = nullptr;  // pinned
    hipEvent_t hole_ev = nullptr;
    int hole_parity = 0;
    bool holes_listed = false;
    int n_holes = 0;
    int hole_iters = 0;            // sweeps run by the last hole filling
    float *pp_scratch = nullptr;   // push-pull pyramid scratch
    size_t pp_cap = 0;             // floats
    bool vis_ready = false;

    // U holds the cleared (all-zero) image: set by lfrv_clear_solution, reset by
    // every entry point that writes U
    bool u_zero = false;

    // timing
    bool timing = false;
    int timing_stride = 1;          // kernel-event classes: one launch in timing_stride is timed
    int64_t timing_seq[6] = {};     // launches seen per class while timing
    lfrv::TimingSlot tslot[6];
    std::vector<hipEvent_t> ev_pool;
    // recorded (class, start, stop) event pairs not yet read back; owned by the
    // context, so contexts driven from different host threads share nothing
    std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> ev_pending;
};

namespace lfrv {

// error helpers
lfrv_status fail(lfrv_ctx *ctx, lfrv_status st, const std::string &msg);
lfrv_status hip_fail(lfrv_ctx *ctx, hipError_t e, const char *what);

#define LFRV_HIP(ctx, call)                                       \
    do {                                                          \
        hipError_t e_ = (call);                                   \
        if (e_ != hipSuccess) return lfrv::hip_fail(ctx, e_, #call); \
    } while (0)

#define LFRV_LAUNCH(ctx, what)                                    \
    do {                                                          \
        hipError_t e_ = hipGetLastError();                        \
        if (e_ != hipSuccess) return lfrv::hip_fail(ctx, e_, what); \
    } while (0)

// Timed launch bracket (class k); no-ops unless timing is on.
struct TimedRegion {
    lfrv_ctx *ctx;
    int k;
    hipEvent_t a = nullptr, b = nullptr;
    TimedRegion(lfrv_ctx *c, int kk);
    ~TimedRegion();
};
void flush_timing(lfrv_ctx *ctx);

// Kernel-exact timing: an event pair handed to hipExtLaunchKernelGGL, which
// stamps the kernel's own start and stop (no dispatch overhead in between).
struct KernelEvents {
    hipEvent_t a = nullptr, b = nullptr;
};
KernelEvents kernel_events(lfrv_ctx *ctx, int k);   // empty pair unless timing is on and launch k's turn
void record_kernel_events(lfrv_ctx *ctx, int k, const KernelEvents &ev);

inline unsigned int blocks_for(size_t n, int bs = kBlock) { return (unsigned int)((n + bs - 1) / bs); }

// kernels-side entry points (defined in the .hip files)
lfrv_status run_warps_from_depth(lfrv_ctx *ctx);
lfrv_status run_warps_from_target_depth(lfrv_ctx *ctx, int skip);
lfrv_status run_prepare(lfrv_ctx *ctx);
lfrv_status run_refresh_entry_weights(lfrv_ctx *ctx, float scale);
lfrv_status run_visibility(lfrv_ctx *ctx);
// s: the stream to launch on (nullptr: the context's own)
lfrv_status run_compute_weights(lfrv_ctx *ctx, hipStream_t s = nullptr);
lfrv_status materialize_dpart(lfrv_ctx *ctx);

// ---- IBR_color steps of a batch of solves (k_sr.hip) ----
struct DataTermArgs {
    float alpha, beta, sigma2, grad_step, dsf2, prox_factor;  // prox_factor = -1 / (lambda L)
    int gw_type;
};
// The weights of the two data terms' back-projections in the gradient step
// (optimization.cu:1835-2116): alpha sigma^2 for the intensity
// term (0 when alpha <= 0), -beta gradient_step [sigma^2 with gw_type] for the
// gradient term.
__host__ __device__ __forceinline__ float dt_intensity_weight(const DataTermArgs &da)
{
    return da.alpha > 0.0f ? da.alpha * da.sigma2 : 0.0f;
}
__host__ __device__ __forceinline__ float dt_gradient_weight(const DataTermArgs &da)
{
    return da.beta * -da.grad_step * (da.gw_type != 0 ? da.sigma2 : 1.0f);
}
struct PdArgs {
    float tau, lam, sigma, theta;
};
// One solve's FISTA steps inside a batched launch: step s reads U[s & 1],
// PX[s & 1], PY[s & 1] and writes the other parity (the pointer swaps of
// the solve's own steps), with theta[s] of its FISTA sequence.
struct SrBatchItem {
    int W, H, Wc, Hc, ntx, ntiles, steps;
    DataTermArgs da;
    PdArgs pa;
    const float *theta;  // device, `steps` entries
    const float *ops;
    size_t NA;
    const float *coverage, *g;
    float *U[2], *PX[2], *PY[2];
    float *uq, *norm_mask;
};
// Whether a context's FISTA steps can run in a batched launch (dsf >= 2, one
// primal-dual iteration per step).
bool sr_batchable(const lfrv_ctx *ctx);
// Readies `steps` steps of ctx for a batched launch on ctx's stream (the
// assembled operator, the entry weights) and fills `it` (all but `theta`)
// and theta_host[0 .. steps).
lfrv_status sr_batch_begin(lfrv_ctx *ctx, int steps, SrBatchItem *it, float *theta_host);
// After the batched steps: the context's buffers and FISTA state as after
// `steps` steps of its own.
void sr_batch_end(lfrv_ctx *ctx, int steps);
// Step s of the batch (items of footprint ks, with the gradient term or
// without, work list of (item, local block)).
hipError_t launch_sr_batch_step(int ks, bool grad, const SrBatchItem *items, const int2 *work, int nblocks, int s,
                                hipStream_t stream);
lfrv_status run_initial_image(lfrv_ctx *ctx);
lfrv_status run_hole_filling(lfrv_ctx *ctx, int iters, double eps, int *ran);
lfrv_status run_push_pull(lfrv_ctx *ctx, int levels);
lfrv_status run_u_gradient(lfrv_ctx *ctx);
lfrv_status run_cg(lfrv_ctx *ctx, double alpha, double c0, double beta, bool use_views,
                   bool use_precond, bool poisson_rhs, double tol, int maxit, lfrv_cg_stats *st);
lfrv_status run_gibr_energy(lfrv_ctx *ctx, double *e);
lfrv_status run_regularizer_weight(lfrv_ctx *ctx);
// last = false: a step followed by another of the same run (Uq and the norm
// mask are then left to the last one)
lfrv_status run_fista(lfrv_ctx *ctx, bool last = true);
lfrv_status run_sr_energy(lfrv_ctx *ctx, double *e, double *parts);
lfrv_status run_count_moving(lfrv_ctx *ctx, const float *prev_dev, float max_diff, int64_t *moving);
lfrv_status run_sr_solve(lfrv_ctx *ctx, int niter, int nmeta, lfrv_sr_stats *st);

// batched stencil PCG (k_cg.hip): the part of lfrv_cg_solve before the
// iterations for one dsf = 1 solve (its CgBatchItem filled, work queued on the
// context's stream), the iterations of many solves (block b on tile
// work[b].y of solve work[b].x), the scalar step of every solve, and the
// part after the iterations (true residual, statistics; waits for the stream)
lfrv_status cg_batch_begin(lfrv_ctx *ctx, double tol, int maxit, CgBatchItem *item);
hipError_t launch_cg_batch_iteration(const CgBatchItem *items, const int2 *work, int nblocks, int parity, hipStream_t s);
hipError_t launch_cg_batch_step(const CgBatchItem *items, int nitems, hipStream_t s);
lfrv_status cg_batch_end(lfrv_ctx *ctx, lfrv_cg_stats *st);

// device PSNR (k_quality.hip): interior pixels per partial-sum block, the
// number of blocks for a W x H image, and the two launches over `partials`
constexpr int kPsnrChunk = 4096;
int psnr_blocks(int W, int H);
lfrv_status psnr_launch(int W, int H, const float *a, const float *b, double *out, double *partials, int G,
                        hipStream_t s);

// stateless kernels
hipError_t launch_gather(int W, int H, int w, int h, int dsf, const float *tx, const float *ty,
                         const float *u, int C, float *out, hipStream_t s);
hipError_t launch_splat_atomic(int W, int H, int w, int h, int dsf, const float *tx,
                               const float *ty, const float *in, int C, const float *wgt,
                               float *out, float *wsum, hipStream_t s);
hipError_t launch_radial_to_ortho(int W, int H, const double *K, float *z, float *gx, float *gy, hipStream_t s);
hipError_t launch_sobel(int W, int H, const float *r, const float *g, const float *b, float *gx,
                        float *gy, const float *visf, const uint8_t *visb, hipStream_t s);

}  // namespace lfrv
