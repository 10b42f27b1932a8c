// Kernels of the warm-start CNN: 3x3 convolutions on the bf16 MFMA units of
// CDNA4 with FP32 accumulation, their backward passes, and the elementwise
// pieces of a U-Net (2x2 average pooling, nearest 2x upsampling, ReLU, the
// squared-error loss, Adam).  No counterpart in the reference (its CNN
// training, sources/CNN/, is a separate TensorFlow tool outside this path).
//
// Layouts.  Activations are NHWC bf16: [B][H][W][C], channels contiguous.
// Weights are [Cout][9][Cin] bf16 (tap t = 3 (dy + 1) + (dx + 1)), so one
// output channel's weights are contiguous along K = 9 Cin.  Gradients of
// activations and weights are FP32.  Input channel counts are multiples of
// 32 (the first layer's are zero-padded), so a K chunk of 32 never straddles
// a tap.
//
// The convolution is an implicit GEMM: C[pixel][cout] = sum_k A[pixel][k]
// W[cout][k] with A[p][(t, ci)] = X[p + offset(t)][ci] (0 outside the
// image).  k_conv3x3 walks K by (tap, 32-channel chunk) over 128 pixels in
// row order; k_conv3x3_halo (64 input channels and more) by 32-channel chunk
// over a 2-D pixel tile whose halo serves all 9 taps.  Both run
// v_mfma_f32_16x16x32_bf16 on 16-pixel x 16-channel tiles.  The backward
// pass with respect to the input is the
// same kernel on the output gradient with the weights flipped in space and
// transposed in channels; the weight gradient is a second MFMA kernel whose
// K runs over pixels.
#include "lfrv_internal.hpp"
#include "tile_common.hpp"

namespace lfrv {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);                      // round to nearest even
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr int kCvBM = 128, kCvBK = 32, kCvPitch = 40;  // LDS rows of 40 bf16 (80 B, 16-byte aligned)

typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

enum : int { kEpiRelu = 1, kEpiSkip = 2, kEpiF32 = 4, kEpiAccum = 8, kEpiBias = 16 };

// Y[p][co] = act(sum_{t,ci} W[co][t][ci] X[p + off(t)][ci] + bias[co] + skip[p][co]).
// EPI: kEpiRelu, kEpiSkip (bf16 skip), kEpiBias, and the output kind: bf16,
// or FP32 (kEpiF32), optionally added to what Y holds (kEpiAccum).
// A block owns 128 pixels x 32 WN output channels; its 4 waves are arranged
// (4 / WN) x WN, each on (128 WN / 4) pixels x 32 channels, i.e. 2 WN x 2
// tiles of v_mfma_f32_16x16x32_bf16 (gfx950).  Per K chunk of 32 (one tap,
// 32 input channels) the next chunk's tiles are loaded into registers while
// the current one is multiplied out of LDS, then stored into the other LDS
// buffer: one barrier per chunk.  An MFMA operand's 8 K values per lane are
// 8 consecutive input channels in both the pixel and the weight tile, so
// any K order the instruction uses inside pairs them alike.
template <int EPI, int WN>
__global__ void __launch_bounds__(256)
k_conv3x3(const uint16_t *__restrict__ X, const uint16_t *__restrict__ Wt, const float *__restrict__ bias,
          const uint16_t *__restrict__ skip, void *__restrict__ Y, int B, int H, int W, int Cin, int Cout)
{
    constexpr int BN = 32 * WN, WM = 4 / WN, FM = kCvBM / WM / 16;
    __shared__ __align__(16) uint16_t As[2][kCvBM * kCvPitch];
    __shared__ __align__(16) uint16_t Bs[2][BN * kCvPitch];
    const int HW = H * W, P = B * HW;
    const int p0 = blockIdx.x * kCvBM, co0 = blockIdx.y * BN;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, wm = wv / WN, wn = wv % WN;
    // staging: A rows sr and sr + 64 (8 channels each), B row sr (threads < 4 BN)
    const int sr = tid >> 2, sc = (tid & 3) * 8;
    int sb[2], sy[2], sx[2];
    bool prow[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int sp = p0 + sr + 64 * h;
        prow[h] = sp < P;
        sb[h] = prow[h] ? sp / HW : 0;
        const int rem = prow[h] ? sp - sb[h] * HW : 0;
        sy[h] = rem / W;
        sx[h] = rem - sy[h] * W;
    }
    const int sco = co0 + sr;
    const bool bload = tid < 4 * BN && sco < Cout;
    const int ncc = Cin / kCvBK, nk = 9 * ncc;
    // loads always read a valid address (pixel 0 / row 0 when the tap falls
    // outside) and zero the value afterwards: a select between the load and a
    // zero constant would put the constant in scratch memory and make every
    // load a flat load through it
    u32x4 a0, a1, b0;
    bool m0, m1;
#define CV_LOAD(kc)                                                                                         \
    do {                                                                                                    \
        const int t_ = (kc) / ncc, ci0_ = ((kc) - t_ * ncc) * kCvBK;                                       \
        const int dy_ = t_ / 3 - 1, dx_ = t_ % 3 - 1;                                                       \
        const int y0_ = sy[0] + dy_, x0_ = sx[0] + dx_, y1_ = sy[1] + dy_, x1_ = sx[1] + dx_;               \
        const bool v0_ = prow[0] && y0_ >= 0 && y0_ < H && x0_ >= 0 && x0_ < W;                             \
        const bool v1_ = prow[1] && y1_ >= 0 && y1_ < H && x1_ >= 0 && x1_ < W;                             \
        const size_t o0_ = v0_ ? ((size_t)(sb[0] * H + y0_) * W + x0_) * Cin : 0;                          \
        const size_t o1_ = v1_ ? ((size_t)(sb[1] * H + y1_) * W + x1_) * Cin : 0;                          \
        a0 = *(const u32x4 *)(X + o0_ + ci0_ + sc);                                                         \
        a1 = *(const u32x4 *)(X + o1_ + ci0_ + sc);                                                         \
        b0 = *(const u32x4 *)(Wt + (bload ? ((size_t)sco * 9 + t_) * Cin : 0) + ci0_ + sc);                 \
        m0 = v0_;                                                                                           \
        m1 = v1_;                                                                                           \
    } while (0)
#define CV_STORE(buf)                                                                                       \
    do {                                                                                                    \
        const u32x4 z_ = {0u, 0u, 0u, 0u};                                                                  \
        *(u32x4 *)(&As[buf][sr * kCvPitch + sc]) = m0 ? a0 : z_;                                            \
        *(u32x4 *)(&As[buf][(sr + 64) * kCvPitch + sc]) = m1 ? a1 : z_;                                     \
        if (tid < 4 * BN) *(u32x4 *)(&Bs[buf][sr * kCvPitch + sc]) = bload ? b0 : z_;                       \
    } while (0)
    f4v acc[FM][2];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};
    CV_LOAD(0);
    CV_STORE(0);
    __syncthreads();
    for (int kc = 0; kc < nk; kc++) {
        const int buf = kc & 1;
        if (kc + 1 < nk) CV_LOAD(kc + 1);
        bf8v a[FM], b[2];
#pragma unroll
        for (int i = 0; i < FM; i++)
            a[i] = *(const bf8v *)(&As[buf][(wm * (kCvBM / WM) + i * 16 + (lane & 15)) * kCvPitch + 8 * (lane >> 4)]);
#pragma unroll
        for (int j = 0; j < 2; j++)
            b[j] = *(const bf8v *)(&Bs[buf][(wn * 32 + j * 16 + (lane & 15)) * kCvPitch + 8 * (lane >> 4)]);
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        if (kc + 1 < nk) CV_STORE(buf ^ 1);
        __syncthreads();
    }
#undef CV_LOAD
#undef CV_STORE
    // lane holds rows 4 (lane / 16) + r, column lane % 16 of each 16 x 16 tile
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int co = co0 + wn * 32 + j * 16 + (lane & 15);
        if (co >= Cout) continue;
        const float bco = (EPI & kEpiBias) ? bias[co] : 0.0f;
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int p = p0 + wm * (kCvBM / WM) + i * 16 + 4 * (lane >> 4) + r;
                if (p >= P) continue;
                const size_t o = (size_t)p * Cout + co;
                float v = acc[i][j][r] + bco;
                if (EPI & kEpiSkip) v += bf2f(skip[o]);
                if (EPI & kEpiRelu) v = fmaxf(v, 0.0f);
                if (EPI & kEpiF32) {
                    float *yf = (float *)Y;
                    yf[o] = (EPI & kEpiAccum) ? yf[o] + v : v;
                } else {
                    ((uint16_t *)Y)[o] = f2bf(v);
                }
            }
    }
}

// The same convolution for layers with 64 or more input channels, staged by
// chunk of 32 input channels instead of by (tap, chunk): a block owns a
// TH x 16 pixel tile of one image and 32 WN output channels.  Per chunk the
// (TH + 2) x 18 halo of the tile (32 channels per pixel) and the weights of
// all 9 taps go to LDS once; the 9 taps then read the halo at their offsets,
// so a pixel's input is fetched (TH + 2) 18 / (16 TH) times per chunk instead
// of 9, and a wave runs 9 FM 2 MFMAs between barriers instead of FM 2.  The
// next chunk's halo and weights are loaded into registers before the 9 taps
// are multiplied out and stored after them (two barriers: the weights have
// one LDS buffer, the halo two).  With a single chunk (32 input channels)
// nothing is left to overlap, and k_conv3x3 is the faster form.
template <int EPI, int WN, int TH>
__global__ void __launch_bounds__(256)
k_conv3x3_halo(const uint16_t *__restrict__ X, const uint16_t *__restrict__ Wt, const float *__restrict__ bias,
             const uint16_t *__restrict__ skip, void *__restrict__ Y, int B, int H, int W, int Cin, int Cout)
{
    constexpr int TW = 16, HX = TW + 2, HP = HX * (TH + 2);
    constexpr int BN = 32 * WN, WM = 4 / WN, FM = TH / WM;
    static_assert(TH % WM == 0, "tile rows split evenly over the waves");
    constexpr int NA = (HP * 4 + 255) / 256, NB = (36 * BN + 255) / 256;
    __shared__ __align__(16) uint16_t As[2][HP * kCvPitch];
    __shared__ __align__(16) uint16_t Bs[9 * BN * kCvPitch];
    const int ntx = (W + TW - 1) / TW, nty = (H + TH - 1) / TH;
    const int img = blockIdx.x / (ntx * nty), trem = blockIdx.x - img * ntx * nty;
    const int ty = trem / ntx, tx = trem - ty * ntx;
    const int y0 = ty * TH, x0 = tx * TW, co0 = blockIdx.y * BN;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, wm = wv / WN, wn = wv % WN;
    const int ncc = Cin / kCvBK;
    size_t aoff[NA];
    int alds[NA];
    bool am[NA];
#pragma unroll
    for (int j = 0; j < NA; j++) {
        const int idx = tid + 256 * j, hp = idx >> 2, q = idx & 3;
        const bool slot = idx < HP * 4;
        const int hy = hp / HX, hx = hp - hy * HX, y = y0 - 1 + hy, x = x0 - 1 + hx;
        am[j] = slot && y >= 0 && y < H && x >= 0 && x < W;
        aoff[j] = am[j] ? ((size_t)(img * H + y) * W + x) * Cin + 8 * q : 0;
        alds[j] = slot ? hp * kCvPitch + 8 * q : -1;
    }
    // weight slots: idx -> tap idx / (4 BN), channel row (idx / 4) % BN, 8-channel group idx % 4
    size_t boff[NB];
    int blds[NB];
    bool bm[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) {
        const int idx = tid + 256 * j, t = idx / (4 * BN), r = (idx >> 2) - t * BN, q = idx & 3;
        const bool slot = idx < 36 * BN;
        bm[j] = slot && co0 + r < Cout;
        boff[j] = bm[j] ? ((size_t)(co0 + r) * 9 + t) * Cin + 8 * q : 0;
        blds[j] = slot ? (t * BN + r) * kCvPitch + 8 * q : -1;
    }
    const u32x4 z = {0u, 0u, 0u, 0u};
    u32x4 ar[NA], br[NB];
    auto load = [&](int cc) {
#pragma unroll
        for (int j = 0; j < NA; j++) ar[j] = *(const u32x4 *)(X + aoff[j] + cc * kCvBK);
#pragma unroll
        for (int j = 0; j < NB; j++) br[j] = *(const u32x4 *)(Wt + boff[j] + cc * kCvBK);
    };
    auto store = [&](int abuf) {
#pragma unroll
        for (int j = 0; j < NA; j++)
            if (alds[j] >= 0) *(u32x4 *)(&As[abuf][alds[j]]) = am[j] ? ar[j] : z;
#pragma unroll
        for (int j = 0; j < NB; j++)
            if (blds[j] >= 0) *(u32x4 *)(&Bs[blds[j]]) = bm[j] ? br[j] : z;
    };
    f4v acc[FM][2];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};
    load(0);
    store(0);
    __syncthreads();
    for (int cc = 0; cc < ncc; cc++) {
        const int abuf = cc & 1;
        if (cc + 1 < ncc) load(cc + 1);
#pragma unroll 3
        for (int t = 0; t < 9; t++) {
            const int dy = t / 3, dx = t - 3 * dy;
            bf8v a[FM], b[2];
#pragma unroll
            for (int i = 0; i < FM; i++) {
                const int hp = (wm * FM + i + dy) * HX + (lane & 15) + dx;
                a[i] = *(const bf8v *)(&As[abuf][hp * kCvPitch + 8 * (lane >> 4)]);
            }
#pragma unroll
            for (int j = 0; j < 2; j++)
                b[j] = *(const bf8v *)(&Bs[(t * BN + wn * 32 + j * 16 + (lane & 15)) * kCvPitch + 8 * (lane >> 4)]);
#pragma unroll
            for (int i = 0; i < FM; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (cc + 1 < ncc) {
            __syncthreads();  // every wave is done with the weights of this chunk
            store(abuf ^ 1);
            __syncthreads();
        }
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int co = co0 + wn * 32 + j * 16 + (lane & 15);
        if (co >= Cout) continue;
        const float bco = (EPI & kEpiBias) ? bias[co] : 0.0f;
#pragma unroll
        for (int i = 0; i < FM; i++) {
            const int y = y0 + wm * FM + i;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int x = x0 + 4 * (lane >> 4) + r;
                if (y >= H || x >= W) continue;
                const size_t o = ((size_t)(img * H + y) * W + x) * Cout + co;
                float v = acc[i][j][r] + bco;
                if (EPI & kEpiSkip) v += bf2f(skip[o]);
                if (EPI & kEpiRelu) v = fmaxf(v, 0.0f);
                if (EPI & kEpiF32) {
                    float *yf = (float *)Y;
                    yf[o] = (EPI & kEpiAccum) ? yf[o] + v : v;
                } else {
                    ((uint16_t *)Y)[o] = f2bf(v);
                }
            }
        }
    }
}

// Weight gradient, split over pixel chunks:
//   dW[co][t][ci] += sum_{p in chunk} dU[p][co] X[p + off(t)][ci].
// GEMM with M = co (64 per block), N = ci of one tap (32 per block), K =
// pixels (32 per LDS stage, kWgChunk per block).  Both tiles are stored
// transposed in LDS ([co][p], [ci][p]) so an MFMA fragment's 4 consecutive K
// values are contiguous.  Each wave owns 16 rows of co and both 16-column
// tiles of ci; the block adds its 64 x 32 partial to dW with FP32 atomics.
constexpr int kWgChunk = 512;
constexpr int kWgPitch = 36;  // bf16 per LDS row of 32 pixels (+4: rows spread over the banks)

__global__ void __launch_bounds__(256)
k_conv3x3_wgrad(const uint16_t *__restrict__ dU, const uint16_t *__restrict__ X, float *__restrict__ dW, int B,
                int H, int W, int Cin, int Cout)
{
    __shared__ __align__(16) uint16_t Ds[64 * kWgPitch];
    __shared__ __align__(16) uint16_t Xs[32 * kWgPitch];
    const int HW = H * W, P = B * HW;
    const int pc0 = blockIdx.x * kWgChunk;
    const int co0 = blockIdx.y * 64;
    const int t = blockIdx.z / (Cin / 32), ci0 = (blockIdx.z % (Cin / 32)) * 32;
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    f4v acc[2] = {f4v{0.f, 0.f, 0.f, 0.f}, f4v{0.f, 0.f, 0.f, 0.f}};
    // staging: dU tile 32 pixels x 64 co = 256 x 8 bf16 (one 16-byte load per
    // thread); X tile 32 pixels x 32 ci = 128 x 8 bf16 (threads 0..127)
    const int dp = tid >> 3, dc = (tid & 7) * 8;
    const int xp = (tid & 127) >> 2, xc = (tid & 3) * 8;
    const uint4 zero = make_uint4(0, 0, 0, 0);
    const int pend = min(pc0 + kWgChunk, P);
    for (int pb = pc0; pb < pend; pb += 32) {
        uint4 dv = zero, xv = zero;
        {
            const int p = pb + dp;
            if (p < pend && co0 + dc < Cout) dv = *(const uint4 *)(dU + (size_t)p * Cout + co0 + dc);
        }
        if (tid < 128) {
            const int p = pb + xp;
            if (p < pend) {
                const int b = p / HW, r = p - b * HW, y = r / W + dy, x = r % W + dx;
                if (y >= 0 && y < H && x >= 0 && x < W)
                    xv = *(const uint4 *)(X + ((size_t)(b * H + y) * W + x) * Cin + ci0 + xc);
            }
        }
        __syncthreads();
        {
            const uint16_t *d = (const uint16_t *)&dv;
#pragma unroll
            for (int k = 0; k < 8; k++) Ds[(dc + k) * kWgPitch + dp] = d[k];
        }
        if (tid < 128) {
            const uint16_t *d = (const uint16_t *)&xv;
#pragma unroll
            for (int k = 0; k < 8; k++) Xs[(xc + k) * kWgPitch + xp] = d[k];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 32; kk += 16) {
            const s4v a = *(const s4v *)(Ds + (wv * 16 + (lane & 15)) * kWgPitch + kk + 4 * (lane >> 4));
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const s4v b = *(const s4v *)(Xs + (j * 16 + (lane & 15)) * kWgPitch + kk + 4 * (lane >> 4));
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int ci = ci0 + j * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int co = co0 + wv * 16 + 4 * (lane >> 4) + r;
            if (co < Cout) atomicAdd(dW + ((size_t)co * 9 + t) * Cin + ci, acc[j][r]);
        }
    }
}

// db[c] += sum_p dU[p][c] over a chunk of pixels per block (C <= 256).
__global__ void __launch_bounds__(256)
k_bias_grad(const uint16_t *__restrict__ dU, float *__restrict__ db, int P, int C)
{
    const int c = threadIdx.x;
    if (c >= C) return;
    const int p0 = blockIdx.x * 256, p1 = min(p0 + 256, P);
    float s = 0.0f;
    for (int p = p0; p < p1; p++) s += bf2f(dU[(size_t)p * C + c]);
    atomicAdd(db + c, s);
}

// dU = bf16(dY * [y > 0]) (y: the layer's ReLU output; relu = 0: no mask);
// with dskip: dskip += the masked gradient (the skip input's share).
__global__ void k_relu_bwd(const float *__restrict__ dY, const uint16_t *__restrict__ y, uint16_t *__restrict__ dU,
                           float *__restrict__ dskip, size_t n, int relu)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float g = dY[i];
        if (relu && !(bf2f(y[i]) > 0.0f)) g = 0.0f;
        dU[i] = f2bf(g);
        if (dskip) dskip[i] += g;
    }
}

// 2x2 average pooling of NHWC bf16 (H, W even): out [B][H/2][W/2][C].
__global__ void k_pool2(const uint16_t *__restrict__ x, uint16_t *__restrict__ y, int B, int H, int W, int C)
{
    const int Ho = H / 2, Wo = W / 2;
    const size_t n = (size_t)B * Ho * Wo * C;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        size_t q = i / C;
        const int xo = (int)(q % Wo);
        q /= Wo;
        const int yo = (int)(q % Ho);
        const int b = (int)(q / Ho);
        const size_t r0 = ((size_t)(b * H + 2 * yo) * W + 2 * xo) * C + c, r1 = r0 + (size_t)W * C;
        y[i] = f2bf(0.25f * ((bf2f(x[r0]) + bf2f(x[r0 + C])) + (bf2f(x[r1]) + bf2f(x[r1 + C]))));
    }
}

// Gradient of the pooling: dx[4 children] += dy / 4 (FP32).
__global__ void k_pool2_bwd(const float *__restrict__ dy, float *__restrict__ dx, int B, int H, int W, int C)
{
    const int Ho = H / 2, Wo = W / 2;
    const size_t n = (size_t)B * Ho * Wo * C;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        size_t q = i / C;
        const int xo = (int)(q % Wo);
        q /= Wo;
        const int yo = (int)(q % Ho);
        const int b = (int)(q / Ho);
        const float g = 0.25f * dy[i];
        const size_t r0 = ((size_t)(b * H + 2 * yo) * W + 2 * xo) * C + c, r1 = r0 + (size_t)W * C;
        dx[r0] += g;
        dx[r0 + C] += g;
        dx[r1] += g;
        dx[r1 + C] += g;
    }
}

// Nearest 2x upsampling of NHWC bf16: out [B][2H][2W][C].
__global__ void k_up2(const uint16_t *__restrict__ x, uint16_t *__restrict__ y, int B, int H, int W, int C)
{
    const int Ho = 2 * H, Wo = 2 * W;
    const size_t n = (size_t)B * Ho * Wo * C;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        size_t q = i / C;
        const int xo = (int)(q % Wo);
        q /= Wo;
        const int yo = (int)(q % Ho);
        const int b = (int)(q / Ho);
        y[i] = x[((size_t)(b * H + yo / 2) * W + xo / 2) * C + c];
    }
}

// Gradient of the upsampling: dx += the sum of the 4 children of dy (FP32).
__global__ void k_up2_bwd(const float *__restrict__ dy, float *__restrict__ dx, int B, int H, int W, int C)
{
    const size_t n = (size_t)B * H * W * C;
    const int Wo = 2 * W;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        size_t q = i / C;
        const int x = (int)(q % W);
        q /= W;
        const int y = (int)(q % H);
        const int b = (int)(q / H);
        const size_t r0 = ((size_t)(b * 2 * H + 2 * y) * Wo + 2 * x) * C + c, r1 = r0 + (size_t)Wo * C;
        dx[i] += (dy[r0] + dy[r0 + C]) + (dy[r1] + dy[r1 + C]);
    }
}

__global__ void k_fill_f32(float *__restrict__ a, size_t n, float v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}

__global__ void k_f32_to_bf16(const float *__restrict__ a, uint16_t *__restrict__ b, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = f2bf(a[i]);
}

// Wf[ci][t][co] = W[co][8 - t][ci]: the weights of the input-gradient convolution.
__global__ void k_flip_weights(const uint16_t *__restrict__ Wt, uint16_t *__restrict__ Wf, int Cin, int Cout)
{
    const size_t n = (size_t)Cout * 9 * Cin;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int ci = (int)(i % Cin);
        const size_t q = i / Cin;
        const int t = (int)(q % 9);
        const int co = (int)(q / 9);
        Wf[((size_t)ci * 9 + (8 - t)) * Cout + co] = Wt[i];
    }
}

// Loss of the prediction base + out[., 0..2] against the target (planar FP32
// [B][3][H][W] for base and target; out NHWC FP32 with `co` channels):
// per-block partial sums of the squared error (FP64) and the gradient
// d loss / d out = 2 (pred - target) / count (0 on the padding channels).
__global__ void __launch_bounds__(256)
k_sq_loss(const float *__restrict__ out, const float *__restrict__ base, const float *__restrict__ target,
          float *__restrict__ dout, double *__restrict__ partials, int B, int HW, int co, float scale)
{
    double s = 0.0;
    const size_t n = (size_t)B * HW;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
        const size_t b = p / HW, o = p - b * HW;
        for (int c = 0; c < co; c++) {
            float g = 0.0f;
            if (c < 3) {
                const size_t q = (b * 3 + c) * HW + o;
                const float d = base[q] + out[p * co + c] - target[q];
                s += (double)d * d;
                g = scale * d;
            }
            if (dout) dout[p * co + c] = g;
        }
    }
    __shared__ double red[16];
    double v[1] = {s};
    block_partials<1>(v, partials, red);
}

// pred (planar FP32 [B][3][H][W]) = base + out[., 0..2].
__global__ void k_add_prediction(const float *__restrict__ out, const float *__restrict__ base, float *__restrict__ pred,
                                 int B, int HW, int co)
{
    const size_t n = (size_t)B * 3 * HW;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t o = i % HW, bc = i / HW, b = bc / 3, c = bc % 3;
        pred[i] = base[i] + out[(b * HW + o) * co + c];
    }
}

// Adam on FP32 master weights (Kingma & Ba, bias-corrected), then the bf16
// copy the convolutions read.  mask (nullable) zeroes the update of padding
// weights.
__global__ void k_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v, const float *__restrict__ g,
                       uint16_t *__restrict__ wb, size_t n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (wb) wb[i] = f2bf(wi);
    }
}

// The warped stack (the network's input): per view, the view's colours
// splatted into the target with the solver's footprint weights and
// normalised by the splatted weight, and min(1, weight); 0 where no footprint
// reached the pixel; channels 4 v .. 4 v + 3 of a bf16 NHWC [N][cpad] tensor,
// the channels beyond 4 V zero.
//
// A footprint splat is a 2 x 2 splat of the entry onto box starts (anchors)
// followed by a D x D box sum (D = dsf, box_weights): per source pixel 4
// anchor taps instead of (D + 1)^2, which is what the atomics cost.  Views go
// in groups of kStackGroup: their anchors (r, g, b, weight sums on a
// (W + D + 1) x (H + D + 1) grid, anchor x0 at column x0 + D) by float
// atomics, then one pass per target tile that box-sums them from LDS and
// writes the group's 4 kStackGroup channels of each pixel in two 16-byte
// stores.
constexpr int kStackGroup = 4;
constexpr int kStackTW = 32, kStackTH = 8;

template <int KS>
__global__ void k_stack_anchor(int w, int h, int W, int H, int WA, int HA, const float *__restrict__ tx,
                               const float *__restrict__ ty, const float *__restrict__ views, int v0, int nv,
                               float *__restrict__ anchors)
{
    constexpr int D = KS - 1;
    const size_t n = (size_t)w * h, NA = (size_t)WA * HA;
    const size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int g = blockIdx.y;
    if (m >= n || g >= nv) return;
    const size_t k = (size_t)(v0 + g) * n + m;
    const float t0 = tx[k], t1 = ty[k];
    if (!cell_valid(t0, t1, W, H)) return;
    int x0, y0;
    float ax[2], ay[2];
    if (!box_weights<KS>(t0, t1, W, H, x0, y0, ax, ay)) return;
    const float *vp = views + (size_t)(v0 + g) * 3 * n + m;
    const float val[4] = {vp[0], vp[n], vp[2 * n], 1.0f};
    float *A = anchors + (size_t)g * 4 * NA;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const size_t o = (size_t)(y0 + i + D) * WA + (x0 + j + D);
            const float a = ay[i] * ax[j];
#pragma unroll
            for (int q = 0; q < 4; q++) atomicAdd(A + q * NA + o, a * val[q]);
        }
}

template <int D>
__global__ void __launch_bounds__(kStackTW * kStackTH)
k_stack_box_pack(int W, int H, int WA, int HA, const float *__restrict__ anchors, int nv, uint16_t *__restrict__ out,
                 int cpad, int chan0, int zero_from)
{
    constexpr int LW = kStackTW + D - 1, LH = kStackTH + D - 1, LP = LW * LH, NPL = 4 * kStackGroup;
    __shared__ float win[NPL * LP];
    const size_t NA = (size_t)WA * HA;
    const int x0 = blockIdx.x * kStackTW, y0 = blockIdx.y * kStackTH;
    // pixel (x, y) sums the anchors of box starts x - D + 1 .. x, i.e. anchor
    // columns x + 1 .. x + D; the window starts at column x0 + 1, row y0 + 1
    for (int i = threadIdx.x; i < NPL * LP; i += blockDim.x) {
        const int pl = i / LP, r = i - pl * LP, ly = r / LW, lx = r - ly * LW;
        const int ax = x0 + 1 + lx, ay = y0 + 1 + ly;
        win[i] = (pl < 4 * nv && ax < WA && ay < HA) ? anchors[(size_t)pl * NA + (size_t)ay * WA + ax] : 0.0f;
    }
    __syncthreads();
    const int lx = threadIdx.x % kStackTW, ly = threadIdx.x / kStackTW;
    const int x = x0 + lx, y = y0 + ly;
    if (x >= W || y >= H) return;
    uint16_t ch[NPL];
#pragma unroll
    for (int g = 0; g < kStackGroup; g++) {
        float sum[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float *pw = win + (4 * g + q) * LP;
            float acc = 0.0f;
#pragma unroll
            for (int i = 0; i < D; i++)
#pragma unroll
                for (int j = 0; j < D; j++) acc += pw[(ly + i) * LW + lx + j];
            sum[q] = acc;
        }
        const float iw = sum[3] > 0.0f ? 1.0f / sum[3] : 0.0f;
        const bool live = g < nv;
        ch[4 * g + 0] = live ? f2bf(sum[0] * iw) : 0;
        ch[4 * g + 1] = live ? f2bf(sum[1] * iw) : 0;
        ch[4 * g + 2] = live ? f2bf(sum[2] * iw) : 0;
        ch[4 * g + 3] = live ? f2bf(fminf(sum[3], 1.0f)) : 0;
    }
    uint16_t *q = out + ((size_t)y * W + x) * cpad + chan0;
    u32x4 lo, hi;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        lo[k] = (uint32_t)ch[2 * k] | ((uint32_t)ch[2 * k + 1] << 16);
        hi[k] = (uint32_t)ch[8 + 2 * k] | ((uint32_t)ch[8 + 2 * k + 1] << 16);
    }
    *(u32x4 *)q = lo;
    *(u32x4 *)(q + 8) = hi;
    // after the last group: the channels beyond the views, zero
    if (zero_from >= 0)
        for (int c = zero_from; c < cpad; c += 8) *(u32x4 *)(out + ((size_t)y * W + x) * cpad + c) = u32x4{0, 0, 0, 0};
}

// ---- launchers --------------------------------------------------------------

inline unsigned int ew_blocks(size_t n) { return (unsigned int)std::min<size_t>((n + 255) / 256, 16384); }

hipError_t cnn_conv_fwd(const uint16_t *X, const uint16_t *Wt, const float *bias, const uint16_t *skip, void *Y,
                        int B, int H, int W, int Cin, int Cout, int epi, hipStream_t s)
{
    if (Cin % kCvBK != 0) return hipErrorInvalidValue;
    const unsigned int gm = (unsigned int)(((size_t)B * H * W + kCvBM - 1) / kCvBM);
    // 64-channel blocks when there are more than 32 output channels, else 32
    const bool wide = Cout > 32;
    const unsigned int gn = (unsigned int)((Cout + (wide ? 63 : 31)) / (wide ? 64 : 32));
    const dim3 grid(gm, gn);
    // halo tiles of 8 x 16 pixels from 64 input channels up
    const bool halo = Cin >= 2 * kCvBK;
    const dim3 hgrid((unsigned int)((size_t)B * ((W + 15) / 16) * ((H + 7) / 8)), gn);
#define CV(E)                                                                                                 \
    do {                                                                                                      \
        if (halo && wide) k_conv3x3_halo<E, 2, 8><<<hgrid, 256, 0, s>>>(X, Wt, bias, skip, Y, B, H, W, Cin, Cout); \
        else if (halo) k_conv3x3_halo<E, 1, 8><<<hgrid, 256, 0, s>>>(X, Wt, bias, skip, Y, B, H, W, Cin, Cout); \
        else if (wide) k_conv3x3<E, 2><<<grid, 256, 0, s>>>(X, Wt, bias, skip, Y, B, H, W, Cin, Cout);        \
        else k_conv3x3<E, 1><<<grid, 256, 0, s>>>(X, Wt, bias, skip, Y, B, H, W, Cin, Cout);                  \
    } while (0)
    switch (epi) {
    case kEpiBias | kEpiRelu: CV(kEpiBias | kEpiRelu); break;
    case kEpiBias | kEpiRelu | kEpiSkip: CV(kEpiBias | kEpiRelu | kEpiSkip); break;
    case kEpiBias: CV(kEpiBias); break;
    case kEpiBias | kEpiF32: CV(kEpiBias | kEpiF32); break;
    case kEpiF32: CV(kEpiF32); break;
    case kEpiF32 | kEpiAccum: CV(kEpiF32 | kEpiAccum); break;
    case 0: CV(0); break;
    default: return hipErrorInvalidValue;
    }
#undef CV
    return hipGetLastError();
}

hipError_t cnn_conv_wgrad(const uint16_t *dU, const uint16_t *X, float *dW, int B, int H, int W, int Cin, int Cout,
                          hipStream_t s)
{
    if (Cin % 32 != 0 || Cout % 8 != 0) return hipErrorInvalidValue;
    const size_t P = (size_t)B * H * W;
    const dim3 grid((unsigned int)((P + kWgChunk - 1) / kWgChunk), (unsigned int)((Cout + 63) / 64),
                    (unsigned int)(9 * (Cin / 32)));
    k_conv3x3_wgrad<<<grid, 256, 0, s>>>(dU, X, dW, B, H, W, Cin, Cout);
    return hipGetLastError();
}

hipError_t cnn_bias_grad(const uint16_t *dU, float *db, int P, int C, hipStream_t s)
{
    if (C > 256) return hipErrorInvalidValue;
    k_bias_grad<<<(P + 255) / 256, 256, 0, s>>>(dU, db, P, C);
    return hipGetLastError();
}

hipError_t cnn_relu_bwd(const float *dY, const uint16_t *y, uint16_t *dU, float *dskip, size_t n, int relu,
                        hipStream_t s)
{
    k_relu_bwd<<<ew_blocks(n), 256, 0, s>>>(dY, y, dU, dskip, n, relu);
    return hipGetLastError();
}

hipError_t cnn_pool2(const uint16_t *x, uint16_t *y, int B, int H, int W, int C, hipStream_t s)
{
    k_pool2<<<ew_blocks((size_t)B * H * W * C / 4), 256, 0, s>>>(x, y, B, H, W, C);
    return hipGetLastError();
}

hipError_t cnn_pool2_bwd(const float *dy, float *dx, int B, int H, int W, int C, hipStream_t s)
{
    k_pool2_bwd<<<ew_blocks((size_t)B * H * W * C / 4), 256, 0, s>>>(dy, dx, B, H, W, C);
    return hipGetLastError();
}

hipError_t cnn_up2(const uint16_t *x, uint16_t *y, int B, int H, int W, int C, hipStream_t s)
{
    k_up2<<<ew_blocks((size_t)B * H * W * C * 4), 256, 0, s>>>(x, y, B, H, W, C);
    return hipGetLastError();
}

hipError_t cnn_up2_bwd(const float *dy, float *dx, int B, int H, int W, int C, hipStream_t s)
{
    k_up2_bwd<<<ew_blocks((size_t)B * H * W * C), 256, 0, s>>>(dy, dx, B, H, W, C);
    return hipGetLastError();
}

hipError_t cnn_fill(float *a, size_t n, float v, hipStream_t s)
{
    k_fill_f32<<<ew_blocks(n), 256, 0, s>>>(a, n, v);
    return hipGetLastError();
}

hipError_t cnn_to_bf16(const float *a, uint16_t *b, size_t n, hipStream_t s)
{
    k_f32_to_bf16<<<ew_blocks(n), 256, 0, s>>>(a, b, n);
    return hipGetLastError();
}

hipError_t cnn_flip_weights(const uint16_t *Wt, uint16_t *Wf, int Cin, int Cout, hipStream_t s)
{
    k_flip_weights<<<ew_blocks((size_t)Cout * 9 * Cin), 256, 0, s>>>(Wt, Wf, Cin, Cout);
    return hipGetLastError();
}

hipError_t cnn_sq_loss(const float *out, const float *base, const float *target, float *dout, double *partials,
                       int nblocks, int B, int HW, int co, float scale, hipStream_t s)
{
    k_sq_loss<<<nblocks, 256, 0, s>>>(out, base, target, dout, partials, B, HW, co, scale);
    return hipGetLastError();
}

hipError_t cnn_add_prediction(const float *out, const float *base, float *pred, int B, int HW, int co, hipStream_t s)
{
    k_add_prediction<<<ew_blocks((size_t)B * 3 * HW), 256, 0, s>>>(out, base, pred, B, HW, co);
    return hipGetLastError();
}

int cnn_stack_group() { return kStackGroup; }

size_t cnn_stack_anchor_floats(int W, int H, int dsf) { return (size_t)kStackGroup * 4 * (W + dsf + 1) * (H + dsf + 1); }

// One group of views of the warped stack: anchors (cleared here) then the box
// pass writing channels 4 v0 .. 4 v0 + 4 kStackGroup - 1 (zero for the views
// beyond nv); zero_from >= 0 also clears channels zero_from .. cpad - 1.
hipError_t cnn_stack_group_pass(int dsf, int w, int h, int W, int H, const float *tx, const float *ty,
                                const float *views, int v0, int nv, float *anchors, uint16_t *out, int cpad,
                                int zero_from, hipStream_t s)
{
    const int WA = W + dsf + 1, HA = H + dsf + 1;
    if (nv < 1 || nv > kStackGroup || (4 * v0) % 16 || cpad % 8) return hipErrorInvalidValue;
    hipError_t e = hipMemsetAsync(anchors, 0, cnn_stack_anchor_floats(W, H, dsf) * sizeof(float), s);
    if (e != hipSuccess) return e;
    const size_t n = (size_t)w * h;
    const dim3 ga((unsigned int)((n + 255) / 256), (unsigned int)nv);
    const dim3 gb((W + kStackTW - 1) / kStackTW, (H + kStackTH - 1) / kStackTH);
    switch (dsf) {
#define ST(D)                                                                                                   \
    case D:                                                                                                     \
        k_stack_anchor<D + 1><<<ga, 256, 0, s>>>(w, h, W, H, WA, HA, tx, ty, views, v0, nv, anchors);          \
        k_stack_box_pack<D><<<gb, kStackTW * kStackTH, 0, s>>>(W, H, WA, HA, anchors, nv, out, cpad, 4 * v0,    \
                                                                zero_from);                                     \
        break;
        ST(1) ST(2) ST(3) ST(4)
#undef ST
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t cnn_adam(float *w, float *m, float *v, const float *g, uint16_t *wb, size_t n, float lr, float b1,
                    float b2, float eps, int step, hipStream_t s)
{
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    k_adam<<<ew_blocks(n), 256, 0, s>>>(w, m, v, g, wb, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

}  // namespace lfrv
