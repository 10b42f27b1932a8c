// The warm-start U-Net (include/lfrv_cnn.h): buffers, forward and backward
// schedules, Adam, and the warm start on a solver context.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lfrv_cnn.h"
#include "lfrv_internal.hpp"

namespace lfrv {

// launchers (k_cnn.hip)
hipError_t cnn_conv_fwd(const uint16_t *X, const uint16_t *Wt, const float *bias, const uint16_t *skip, void *Y,
                        int B, int H, int W, int Cin, int Cout, int epi, hipStream_t s);
hipError_t cnn_conv_wgrad(const uint16_t *dU, const uint16_t *X, float *dW, int B, int H, int W, int Cin, int Cout,
                          hipStream_t s);
hipError_t cnn_bias_grad(const uint16_t *dU, float *db, int P, int C, hipStream_t s);
hipError_t cnn_relu_bwd(const float *dY, const uint16_t *y, uint16_t *dU, float *dskip, size_t n, int relu,
                        hipStream_t s);
hipError_t cnn_pool2(const uint16_t *x, uint16_t *y, int B, int H, int W, int C, hipStream_t s);
hipError_t cnn_pool2_bwd(const float *dy, float *dx, int B, int H, int W, int C, hipStream_t s);
hipError_t cnn_up2(const uint16_t *x, uint16_t *y, int B, int H, int W, int C, hipStream_t s);
hipError_t cnn_up2_bwd(const float *dy, float *dx, int B, int H, int W, int C, hipStream_t s);
hipError_t cnn_fill(float *a, size_t n, float v, hipStream_t s);
hipError_t cnn_to_bf16(const float *a, uint16_t *b, size_t n, hipStream_t s);
hipError_t cnn_flip_weights(const uint16_t *Wt, uint16_t *Wf, int Cin, int Cout, hipStream_t s);
hipError_t cnn_sq_loss(const float *out, const float *base, const float *target, float *dout, double *partials,
                       int nblocks, int B, int HW, int co, float scale, hipStream_t s);
hipError_t cnn_add_prediction(const float *out, const float *base, float *pred, int B, int HW, int co, hipStream_t s);
hipError_t cnn_adam(float *w, float *m, float *v, const float *g, uint16_t *wb, size_t n, float lr, float b1,
                    float b2, float eps, int step, hipStream_t s);
int cnn_stack_group();
size_t cnn_stack_anchor_floats(int W, int H, int dsf);
hipError_t cnn_stack_group_pass(int dsf, int w, int h, int W, int H, const float *tx, const float *ty,
                                const float *views, int v0, int nv, float *anchors, uint16_t *out, int cpad,
                                int zero_from, hipStream_t s);

constexpr int kEpiRelu = 1, kEpiSkip = 2, kEpiF32 = 4, kEpiAccum = 8, kEpiBias = 16;
constexpr int kHeadCout = 32;     // the head's 3 channels, computed as 32
constexpr int kLossBlocks = 256;  // partial sums of the loss

// One tensor of the network: NHWC, [B][h][w][c] at its level.
struct Tensor {
    int level = 0, c = 0;
    bool f32 = false;       // the head's output
    void *data = nullptr;   // bf16 (uint16_t) or float
    float *grad = nullptr;  // FP32 gradient (nullptr for the input)
};

enum OpKind { OP_CONV, OP_POOL, OP_UP };

struct Op {
    OpKind kind;
    int in, out, skip = -1;  // tensor indices
    int cin = 0, cout = 0;
    bool relu = false;
    size_t woff = 0, boff = 0;  // offsets of W and b in the parameter vector
};

}  // namespace lfrv

using namespace lfrv;

struct lfrv_cnn {
    lfrv_cnn_params p{};
    int cpad = 0;  // input channels padded to 32
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<Tensor> T;
    std::vector<Op> ops;
    size_t nparams = 0;
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // FP32 master, Adam moments, gradient
    uint16_t *wb = nullptr, *wf = nullptr;  // bf16 weights and the flipped copies of the input-gradient convs
    uint16_t *dU = nullptr;                 // bf16 masked output gradient of the conv being differentiated
    float *base = nullptr, *target = nullptr, *pred = nullptr;
    double *partials = nullptr, *partials_host = nullptr;
    int B = 0, H = 0, W = 0;
    int step = 0;
    bool have_target = false;
    // timing
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    std::vector<double> ev_flops;
    double conv_ms = 0.0, conv_flops = 0.0;
    int64_t conv_launches = 0;
};

namespace {

lfrv_status cfail(lfrv_cnn *c, lfrv_status st, const std::string &msg)
{
    if (c) c->err = msg;
    return st;
}

#define CNN_HIP(c, call)                                                                      \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) return cfail(c, LFRV_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

size_t level_pixels(const lfrv_cnn *c, int level, int B, int H, int W)
{
    return (size_t)B * (size_t)(H >> level) * (size_t)(W >> level);
}

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x51ull) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() { return std::sqrt(-2.0 * std::log(uni())) * std::cos(6.283185307179586 * uni()); }
};

// The layer list of the U-Net described in lfrv_cnn.h.
void build_graph(lfrv_cnn *c)
{
    const int L = c->p.levels;
    auto ch = [&](int l) { return std::min(c->p.base_channels << l, 256); };
    auto tensor = [&](int level, int chn, bool f32 = false) {
        Tensor t;
        t.level = level;
        t.c = chn;
        t.f32 = f32;
        c->T.push_back(t);
        return (int)c->T.size() - 1;
    };
    size_t off = 0;
    auto conv = [&](int in, int cout, bool relu, int skip, bool f32 = false) {
        Op o;
        o.kind = OP_CONV;
        o.in = in;
        o.cin = c->T[in].c;
        o.cout = cout;
        o.relu = relu;
        o.skip = skip;
        o.out = tensor(c->T[in].level, cout, f32);
        o.woff = off;
        off += (size_t)cout * 9 * o.cin;
        o.boff = off;
        off += (size_t)cout;
        c->ops.push_back(o);
        return o.out;
    };
    auto simple = [&](OpKind k, int in) {
        Op o;
        o.kind = k;
        o.in = in;
        o.out = tensor(c->T[in].level + (k == OP_POOL ? 1 : -1), c->T[in].c);
        c->ops.push_back(o);
        return o.out;
    };
    int x = tensor(0, c->cpad);
    std::vector<int> enc(L);
    for (int l = 0; l < L; l++) {
        if (l > 0) x = simple(OP_POOL, x);
        x = conv(x, ch(l), true, -1);
        x = conv(x, ch(l), true, -1);
        enc[l] = x;
    }
    for (int l = L - 2; l >= 0; l--) {
        x = simple(OP_UP, x);
        x = conv(x, ch(l), true, enc[l]);
        x = conv(x, ch(l), true, -1);
    }
    conv(x, kHeadCout, false, -1, true);
    c->nparams = off;
}

lfrv_status alloc_buffers(lfrv_cnn *c)
{
    const int B = c->p.max_batch, H = c->p.max_h, W = c->p.max_w;
    for (size_t i = 0; i < c->T.size(); i++) {
        Tensor &t = c->T[i];
        const size_t n = level_pixels(c, t.level, B, H, W) * t.c;
        CNN_HIP(c, hipMalloc(&t.data, n * (t.f32 ? 4 : 2)));
        if (i > 0) CNN_HIP(c, hipMalloc((void **)&t.grad, n * 4));
    }
    size_t max_out = 0, max_w = 0;
    for (const Op &o : c->ops)
        if (o.kind == OP_CONV) {
            max_out = std::max(max_out, level_pixels(c, c->T[o.out].level, B, H, W) * o.cout);
            max_w = std::max(max_w, (size_t)o.cout * 9 * o.cin);
        }
    const size_t np = c->nparams;
    CNN_HIP(c, hipMalloc((void **)&c->w, np * 4));
    CNN_HIP(c, hipMalloc((void **)&c->m, np * 4));
    CNN_HIP(c, hipMalloc((void **)&c->v, np * 4));
    CNN_HIP(c, hipMalloc((void **)&c->g, np * 4));
    CNN_HIP(c, hipMalloc((void **)&c->wb, np * 2));
    CNN_HIP(c, hipMalloc((void **)&c->wf, np * 2));
    CNN_HIP(c, hipMalloc((void **)&c->dU, max_out * 2));
    const size_t nimg = (size_t)B * 3 * H * W;
    CNN_HIP(c, hipMalloc((void **)&c->base, nimg * 4));
    CNN_HIP(c, hipMalloc((void **)&c->target, nimg * 4));
    CNN_HIP(c, hipMalloc((void **)&c->pred, nimg * 4));
    CNN_HIP(c, hipMalloc((void **)&c->partials, kLossBlocks * sizeof(double)));
    CNN_HIP(c, hipHostMalloc((void **)&c->partials_host, kLossBlocks * sizeof(double), hipHostMallocDefault));
    (void)max_w;
    return LFRV_OK;
}

lfrv_status init_params(lfrv_cnn *c)
{
    std::vector<float> w(c->nparams, 0.0f);
    Rng rng(c->p.seed);
    const Op *head = nullptr;
    for (const Op &o : c->ops)
        if (o.kind == OP_CONV) head = &o;
    for (const Op &o : c->ops) {
        if (o.kind != OP_CONV) continue;
        const int real_cin = (o.in == 0) ? c->p.in_channels : o.cin;
        const int real_cout = (&o == head) ? 3 : o.cout;
        double sd = std::sqrt(2.0 / (9.0 * real_cin));
        if (&o == head) sd *= 1e-2;
        for (int co = 0; co < real_cout; co++)
            for (int t = 0; t < 9; t++)
                for (int ci = 0; ci < real_cin; ci++)
                    w[o.woff + ((size_t)co * 9 + t) * o.cin + ci] = (float)(sd * rng.normal());
    }
    return lfrv_cnn_set_params(c, w.data());
}

void ev_begin(lfrv_cnn *c, hipEvent_t *a)
{
    *a = nullptr;
    if (!c->timing) return;
    if (hipEventCreate(a) == hipSuccess) (void)hipEventRecord(*a, c->stream);
}

void ev_end(lfrv_cnn *c, hipEvent_t a, double flops)
{
    if (!a) return;
    hipEvent_t b = nullptr;
    if (hipEventCreate(&b) != hipSuccess) return;
    (void)hipEventRecord(b, c->stream);
    c->ev.emplace_back(a, b);
    c->ev_flops.push_back(flops);
}

void ev_flush(lfrv_cnn *c)
{
    if (c->ev.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    for (size_t i = 0; i < c->ev.size(); i++) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, c->ev[i].first, c->ev[i].second) == hipSuccess) {
            c->conv_ms += ms;
            c->conv_flops += c->ev_flops[i];
            c->conv_launches++;
        }
        (void)hipEventDestroy(c->ev[i].first);
        (void)hipEventDestroy(c->ev[i].second);
    }
    c->ev.clear();
    c->ev_flops.clear();
}

double conv_flops(const lfrv_cnn *c, const Op &o)
{
    return 2.0 * 9.0 * o.cin * o.cout * (double)level_pixels(c, c->T[o.out].level, c->B, c->H, c->W);
}

lfrv_status forward(lfrv_cnn *c)
{
    const hipStream_t s = c->stream;
    const int B = c->B, H = c->H, W = c->W;
    for (const Op &o : c->ops) {
        const Tensor &ti = c->T[o.in], &to = c->T[o.out];
        const int h = H >> ti.level, w = W >> ti.level;
        if (o.kind == OP_POOL) {
            CNN_HIP(c, cnn_pool2((const uint16_t *)ti.data, (uint16_t *)to.data, B, h, w, ti.c, s));
        } else if (o.kind == OP_UP) {
            CNN_HIP(c, cnn_up2((const uint16_t *)ti.data, (uint16_t *)to.data, B, h, w, ti.c, s));
        } else {
            int epi = kEpiBias;
            if (o.relu) epi |= kEpiRelu;
            if (o.skip >= 0) epi |= kEpiSkip;
            if (to.f32) epi |= kEpiF32;
            hipEvent_t a;
            ev_begin(c, &a);
            CNN_HIP(c, cnn_conv_fwd((const uint16_t *)ti.data, c->wb + o.woff, c->w + o.boff,
                                    o.skip >= 0 ? (const uint16_t *)c->T[o.skip].data : nullptr, to.data, B, h, w,
                                    o.cin, o.cout, epi, s));
            ev_end(c, a, conv_flops(c, o));
        }
    }
    return LFRV_OK;
}

lfrv_status backward(lfrv_cnn *c)
{
    const hipStream_t s = c->stream;
    const int B = c->B, H = c->H, W = c->W;
    for (size_t i = 1; i + 1 < c->T.size(); i++) {
        const Tensor &t = c->T[i];
        CNN_HIP(c, cnn_fill(t.grad, level_pixels(c, t.level, B, H, W) * t.c, 0.0f, s));
    }
    CNN_HIP(c, cnn_fill(c->g, c->nparams, 0.0f, s));
    for (const Op &o : c->ops)
        if (o.kind == OP_CONV && o.in != 0) CNN_HIP(c, cnn_flip_weights(c->wb + o.woff, c->wf + o.woff, o.cin, o.cout, s));
    for (auto it = c->ops.rbegin(); it != c->ops.rend(); ++it) {
        const Op &o = *it;
        const Tensor &ti = c->T[o.in], &to = c->T[o.out];
        const int h = H >> ti.level, w = W >> ti.level;
        if (o.kind == OP_POOL) {
            CNN_HIP(c, cnn_pool2_bwd(to.grad, ti.grad, B, h, w, ti.c, s));
            continue;
        }
        if (o.kind == OP_UP) {
            CNN_HIP(c, cnn_up2_bwd(to.grad, ti.grad, B, h, w, ti.c, s));
            continue;
        }
        const size_t n = (size_t)B * h * w * o.cout;
        CNN_HIP(c, cnn_relu_bwd(to.grad, to.f32 ? nullptr : (const uint16_t *)to.data, c->dU,
                                o.skip >= 0 ? c->T[o.skip].grad : nullptr, n, o.relu ? 1 : 0, s));
        hipEvent_t a;
        ev_begin(c, &a);
        CNN_HIP(c, cnn_conv_wgrad(c->dU, (const uint16_t *)ti.data, c->g + o.woff, B, h, w, o.cin, o.cout, s));
        ev_end(c, a, conv_flops(c, o));
        CNN_HIP(c, cnn_bias_grad(c->dU, c->g + o.boff, B * h * w, o.cout, s));
        if (o.in != 0) {
            ev_begin(c, &a);
            CNN_HIP(c, cnn_conv_fwd(c->dU, c->wf + o.woff, nullptr, nullptr, ti.grad, B, h, w, o.cout, o.cin,
                                    kEpiF32 | kEpiAccum, s));
            ev_end(c, a, conv_flops(c, o));
        }
    }
    return LFRV_OK;
}

lfrv_status loss(lfrv_cnn *c, bool with_grad, double *out)
{
    const int HW = c->H * c->W;
    const Tensor &head = c->T.back();
    const float scale = 2.0f / (float)((double)c->B * 3 * HW);
    CNN_HIP(c, cnn_sq_loss((const float *)head.data, c->base, c->target, with_grad ? head.grad : nullptr, c->partials,
                           kLossBlocks, c->B, HW, kHeadCout, scale, c->stream));
    CNN_HIP(c, hipMemcpyAsync(c->partials_host, c->partials, kLossBlocks * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    CNN_HIP(c, hipStreamSynchronize(c->stream));
    double sum = 0.0;
    for (int i = 0; i < kLossBlocks; i++) sum += c->partials_host[i];
    if (out) *out = sum / ((double)c->B * 3 * HW);
    return std::isfinite(sum) ? LFRV_OK : cfail(c, LFRV_E_NAN, "non-finite loss");
}

lfrv_status check_batch(lfrv_cnn *c, int B, int H, int W)
{
    const int q = 1 << (c->p.levels - 1);
    if (B < 1 || B > c->p.max_batch || H < q || W < q || H > c->p.max_h || W > c->p.max_w || H % q || W % q)
        return cfail(c, LFRV_E_ARG, "batch size or image size out of range (H, W multiples of 2^(levels-1))");
    return LFRV_OK;
}

}  // namespace

// stack packing kernel (k_cnn_stack in this file's companion launcher)
namespace lfrv {
lfrv_status run_gibr_energy(lfrv_ctx *ctx, double *e);
}

extern "C" lfrv_status lfrv_cnn_create(const lfrv_cnn_params *p, lfrv_cnn **out)
{
    if (!p || !out) return LFRV_E_ARG;
    *out = nullptr;
    if (p->in_channels < 1 || p->levels < 1 || p->levels > 6 || p->base_channels < 32 || p->base_channels % 32 ||
        p->max_batch < 1 || p->max_h < 1 || p->max_w < 1)
        return LFRV_E_ARG;
    lfrv_cnn *c = new lfrv_cnn;
    c->p = *p;
    c->cpad = (p->in_channels + 31) / 32 * 32;
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return LFRV_E_HIP;
    }
    build_graph(c);
    lfrv_status st = alloc_buffers(c);
    if (st == LFRV_OK) st = init_params(c);
    if (st == LFRV_OK && hipStreamSynchronize(c->stream) != hipSuccess) st = LFRV_E_HIP;
    if (st != LFRV_OK) {
        lfrv_cnn_destroy(c);
        return st;
    }
    *out = c;
    return LFRV_OK;
}

extern "C" void lfrv_cnn_destroy(lfrv_cnn *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    ev_flush(c);
    for (Tensor &t : c->T) {
        if (t.data) (void)hipFree(t.data);
        if (t.grad) (void)hipFree(t.grad);
    }
    for (void *q : {(void *)c->w, (void *)c->m, (void *)c->v, (void *)c->g, (void *)c->wb, (void *)c->wf,
                    (void *)c->dU, (void *)c->base, (void *)c->target, (void *)c->pred, (void *)c->partials})
        if (q) (void)hipFree(q);
    if (c->partials_host) (void)hipHostFree(c->partials_host);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char *lfrv_cnn_last_error(const lfrv_cnn *c) { return c ? c->err.c_str() : "null network"; }

extern "C" lfrv_status lfrv_cnn_num_params(lfrv_cnn *c, size_t *n)
{
    if (!c || !n) return LFRV_E_ARG;
    *n = c->nparams;
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_cnn_get_params(lfrv_cnn *c, float *out)
{
    if (!c || !out) return LFRV_E_ARG;
    CNN_HIP(c, hipMemcpyAsync(out, c->w, c->nparams * 4, hipMemcpyDefault, c->stream));
    CNN_HIP(c, hipStreamSynchronize(c->stream));
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_cnn_set_params(lfrv_cnn *c, const float *in)
{
    if (!c || !in) return LFRV_E_ARG;
    CNN_HIP(c, hipMemcpyAsync(c->w, in, c->nparams * 4, hipMemcpyDefault, c->stream));
    CNN_HIP(c, cnn_to_bf16(c->w, c->wb, c->nparams, c->stream));
    CNN_HIP(c, cnn_fill(c->m, c->nparams, 0.0f, c->stream));
    CNN_HIP(c, cnn_fill(c->v, c->nparams, 0.0f, c->stream));
    CNN_HIP(c, hipStreamSynchronize(c->stream));
    c->step = 0;
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_cnn_get_grads(lfrv_cnn *c, float *out)
{
    if (!c || !out) return LFRV_E_ARG;
    CNN_HIP(c, hipMemcpyAsync(out, c->g, c->nparams * 4, hipMemcpyDefault, c->stream));
    CNN_HIP(c, hipStreamSynchronize(c->stream));
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_cnn_set_batch(lfrv_cnn *c, int B, int H, int W, const void *const *x,
                                          const float *const *base, const float *const *target)
{
    if (!c || !x || !base) return LFRV_E_ARG;
    lfrv_status st = check_batch(c, B, H, W);
    if (st != LFRV_OK) return st;
    const size_t HW = (size_t)H * W, xs = HW * c->cpad * 2, is = 3 * HW * 4;
    for (int i = 0; i < B; i++) {
        if (!x[i] || !base[i]) return cfail(c, LFRV_E_ARG, "null input of the batch");
        CNN_HIP(c, hipMemcpyAsync((char *)c->T[0].data + i * xs, x[i], xs, hipMemcpyDefault, c->stream));
        CNN_HIP(c, hipMemcpyAsync((char *)c->base + i * is, base[i], is, hipMemcpyDefault, c->stream));
        if (target) {
            if (!target[i]) return cfail(c, LFRV_E_ARG, "null target of the batch");
            CNN_HIP(c, hipMemcpyAsync((char *)c->target + i * is, target[i], is, hipMemcpyDefault, c->stream));
        }
    }
    c->B = B;
    c->H = H;
    c->W = W;
    c->have_target = target != nullptr;
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_cnn_set_batch_host(lfrv_cnn *c, int B, int H, int W, const float *x, const float *base,
                                               const float *target)
{
    if (!c || !x || !base) return LFRV_E_ARG;
    lfrv_status st = check_batch(c, B, H, W);
    if (st != LFRV_OK) return st;
    const size_t HW = (size_t)H * W, P = (size_t)B * HW;
    std::vector<float> xp(P * c->cpad, 0.0f);
    for (size_t q = 0; q < P; q++)
        std::memcpy(&xp[q * c->cpad], x + q * c->p.in_channels, c->p.in_channels * sizeof(float));
    float *tmp = nullptr;
    CNN_HIP(c, hipMalloc((void **)&tmp, xp.size() * 4));
    hipError_t e = hipMemcpy(tmp, xp.data(), xp.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = cnn_to_bf16(tmp, (uint16_t *)c->T[0].data, xp.size(), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return cfail(c, LFRV_E_HIP, std::string("input upload: ") + hipGetErrorString(e));
    CNN_HIP(c, hipMemcpy(c->base, base, 3 * P * 4, hipMemcpyHostToDevice));
    if (target) CNN_HIP(c, hipMemcpy(c->target, target, 3 * P * 4, hipMemcpyHostToDevice));
    c->B = B;
    c->H = H;
    c->W = W;
    c->have_target = target != nullptr;
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_cnn_forward(lfrv_cnn *c, float *pred)
{
    if (!c) return LFRV_E_ARG;
    if (c->B == 0) return cfail(c, LFRV_E_STATE, "no batch set");
    lfrv_status st = forward(c);
    if (st != LFRV_OK) return st;
    const int HW = c->H * c->W;
    CNN_HIP(c, cnn_add_prediction((const float *)c->T.back().data, c->base, c->pred, c->B, HW, kHeadCout, c->stream));
    if (pred) {
        CNN_HIP(c, hipMemcpyAsync(pred, c->pred, (size_t)c->B * 3 * HW * 4, hipMemcpyDefault, c->stream));
        CNN_HIP(c, hipStreamSynchronize(c->stream));
    }
    ev_flush(c);
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_cnn_loss_grad(lfrv_cnn *c, double *lossv)
{
    if (!c) return LFRV_E_ARG;
    if (c->B == 0 || !c->have_target) return cfail(c, LFRV_E_STATE, "no batch with targets set");
    lfrv_status st = forward(c);
    if (st == LFRV_OK) st = loss(c, true, lossv);
    if (st == LFRV_OK) st = backward(c);
    ev_flush(c);
    return st;
}

extern "C" lfrv_status lfrv_cnn_train_step(lfrv_cnn *c, float lr, double *lossv)
{
    lfrv_status st = lfrv_cnn_loss_grad(c, lossv);
    if (st != LFRV_OK) return st;
    c->step++;
    CNN_HIP(c, cnn_adam(c->w, c->m, c->v, c->g, c->wb, c->nparams, lr, 0.9f, 0.999f, 1e-8f, c->step, c->stream));
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_cnn_forward_flops(lfrv_cnn *c, double *flops)
{
    if (!c || !flops) return LFRV_E_ARG;
    double f = 0.0;
    for (const Op &o : c->ops)
        if (o.kind == OP_CONV) f += conv_flops(c, o);
    *flops = f;
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_cnn_set_timing(lfrv_cnn *c, int enable)
{
    if (!c) return LFRV_E_ARG;
    ev_flush(c);
    c->timing = enable != 0;
    c->conv_ms = 0.0;
    c->conv_flops = 0.0;
    c->conv_launches = 0;
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_cnn_get_timing(lfrv_cnn *c, double *ms, int64_t *launches, double *flops)
{
    if (!c) return LFRV_E_ARG;
    ev_flush(c);
    if (ms) *ms = c->conv_ms;
    if (launches) *launches = c->conv_launches;
    if (flops) *flops = c->conv_flops;
    return LFRV_OK;
}

// ---- stateless operators ---------------------------------------------------------------

extern "C" lfrv_status lfrv_op_conv3x3(int B, int H, int W, int Cin, int Cout, const void *X, const void *Wt,
                                       const float *bias, float *Y, void *stream)
{
    if (B < 1 || H < 1 || W < 1 || Cin < 32 || Cin % 32 || Cout < 1 || !X || !Wt || !Y) return LFRV_E_ARG;
    const hipError_t e = cnn_conv_fwd((const uint16_t *)X, (const uint16_t *)Wt, bias, nullptr, Y, B, H, W, Cin, Cout,
                                      bias ? (kEpiF32 | kEpiBias) : kEpiF32, (hipStream_t)stream);
    return e == hipSuccess ? LFRV_OK : LFRV_E_HIP;
}

extern "C" lfrv_status lfrv_op_conv3x3_wgrad(int B, int H, int W, int Cin, int Cout, const void *dU, const void *X,
                                             float *dW, void *stream)
{
    if (B < 1 || H < 1 || W < 1 || Cin < 32 || Cin % 32 || Cout < 8 || Cout % 8 || !dU || !X || !dW) return LFRV_E_ARG;
    const hipError_t e = cnn_conv_wgrad((const uint16_t *)dU, (const uint16_t *)X, dW, B, H, W, Cin, Cout,
                                        (hipStream_t)stream);
    return e == hipSuccess ? LFRV_OK : LFRV_E_HIP;
}

// ---- the warm start on a solver context -----------------------------------------------

extern "C" lfrv_status lfrv_warped_stack(lfrv_ctx *ctx, void *out, int cpad)
{
    if (!ctx || !out) return LFRV_E_ARG;
    if (cpad < 4 * ctx->V || cpad % 32) return fail(ctx, LFRV_E_ARG, "cpad must hold 4 channels per view and be a multiple of 32");
    if (ctx->dsf > 4) return fail(ctx, LFRV_E_ARG, "the warped stack is built for dsf <= 4");
    float *anchors = nullptr;
    LFRV_HIP(ctx, hipMallocAsync(reinterpret_cast<void **>(&anchors),
                                 cnn_stack_anchor_floats(ctx->W, ctx->H, ctx->dsf) * sizeof(float), ctx->stream));
    const int G = cnn_stack_group();
    hipError_t e = hipSuccess;
    for (int v0 = 0; v0 < ctx->V && e == hipSuccess; v0 += G) {
        const int nv = std::min(G, ctx->V - v0);
        const bool last = v0 + G >= ctx->V;
        // the last group's pass also clears the channels beyond its own 4 G
        e = cnn_stack_group_pass(ctx->dsf, ctx->w, ctx->h, ctx->W, ctx->H, ctx->tau_x, ctx->tau_y, ctx->views, v0, nv,
                                 anchors, (uint16_t *)out, cpad, last ? 4 * (v0 + G) : -1, ctx->stream);
    }
    const hipError_t ef = hipFreeAsync(anchors, ctx->stream);
    LFRV_HIP(ctx, e);
    LFRV_HIP(ctx, ef);
    return LFRV_OK;
}

extern "C" lfrv_status lfrv_warm_start(lfrv_ctx *ctx, lfrv_cnn *c, lfrv_warm_stats *st)
{
    if (!ctx || !c) return ctx ? fail(ctx, LFRV_E_ARG, "null network") : LFRV_E_ARG;
    const auto t0 = std::chrono::steady_clock::now();
    if (ctx->p.mode != LFRV_MODE_GRADIENT_IBR) return fail(ctx, LFRV_E_STATE, "lfrv_warm_start on an IBR_color context");
    if (c->cpad < 4 * ctx->V || c->p.in_channels != 4 * ctx->V)
        return fail(ctx, LFRV_E_ARG, "the network's input channels are not 4 per view of this context");
    if (check_batch(c, 1, ctx->H, ctx->W) != LFRV_OK) return fail(ctx, LFRV_E_ARG, c->err);
    double e_plain = 0.0, e_net = 0.0;
    lfrv_status s = run_gibr_energy(ctx, &e_plain);
    if (s != LFRV_OK) return s;
    // the stack into the network's input, the current solution as its base
    s = lfrv_warped_stack(ctx, c->T[0].data, c->cpad);
    if (s != LFRV_OK) return s;
    const size_t img = 3 * ctx->N * 4;
    LFRV_HIP(ctx, hipMemcpyAsync(c->base, ctx->U, img, hipMemcpyDeviceToDevice, ctx->stream));
    LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    c->B = 1;
    c->H = ctx->H;
    c->W = ctx->W;
    c->have_target = false;
    if (lfrv_cnn_forward(c, nullptr) != LFRV_OK) return fail(ctx, LFRV_E_HIP, "network: " + c->err);
    LFRV_HIP(ctx, hipStreamSynchronize(c->stream));
    // the network's start; the splatted average stays in c->base
    LFRV_HIP(ctx, hipMemcpyAsync(ctx->U, c->pred, img, hipMemcpyDeviceToDevice, ctx->stream));
    s = run_gibr_energy(ctx, &e_net);
    if (s != LFRV_OK) return s;
    const bool use = std::isfinite(e_net) && e_net < e_plain;
    if (!use) LFRV_HIP(ctx, hipMemcpyAsync(ctx->U, c->base, img, hipMemcpyDeviceToDevice, ctx->stream));
    ctx->u_zero = false;
    LFRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (st) {
        st->energy_plain = e_plain;
        st->energy_net = e_net;
        st->used_net = use ? 1 : 0;
        st->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return LFRV_OK;
}
