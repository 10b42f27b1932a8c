/* FP64 CPU restatement of the warm-start U-Net (test infrastructure only:
 * the tests compare the GPU network's outputs and gradients with it; nothing
 * in the library calls it).
 *
 * The network of include/lfrv_cnn.h: 3x3 convolutions with bias and zero
 * padding, NHWC layout, weights [Cout][9][Cin] with tap t = 3 (dy+1) + (dx+1);
 * encoder conv-ReLU-conv-ReLU per level with 2x2 average pooling between
 * levels; decoder nearest 2x upsampling, conv + the encoder output of the
 * level, ReLU, conv-ReLU; head conv to 32 channels (the first 3 used), no
 * activation; prediction = base + head; loss = mean over the batch, the 3
 * channels and the pixels of (prediction - target)^2.  Parameters in forward
 * order, W then b for each convolution.
 *
 * With round_bf16 != 0 the values the GPU stores as bf16 are rounded the
 * same way (round to nearest even): the weights, every activation after its
 * layer (not the head), the pooled values, and the masked output gradient of
 * each convolution before it enters the weight and input gradients.  With
 * round_bf16 = 0 everything is exact FP64 (the form checked against finite
 * differences).  The reference has no counterpart (its CNN, sources/CNN/, is
 * a separate TensorFlow tool). */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static double bf16_round(double x)
{
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return x;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

static void maybe_round(double *a, size_t n, int on)
{
    if (!on) return;
    for (size_t i = 0; i < n; i++) a[i] = bf16_round(a[i]);
}

/* Y[p][co] = sum_{t,ci} W[co][t][ci] X[p + off(t)][ci] (+ b[co]) */
void orc_conv3x3(int B, int H, int W, int Cin, int Cout, const double *X, const double *Wt, const double *b,
                 double *Y)
{
    for (int n = 0; n < B; n++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                const size_t p = ((size_t)n * H + y) * W + x;
                for (int co = 0; co < Cout; co++) {
                    double s = b ? b[co] : 0.0;
                    for (int t = 0; t < 9; t++) {
                        const int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
                        if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
                        const double *xr = X + (((size_t)n * H + yy) * W + xx) * Cin;
                        const double *wr = Wt + ((size_t)co * 9 + t) * Cin;
                        for (int ci = 0; ci < Cin; ci++) s += wr[ci] * xr[ci];
                    }
                    Y[p * Cout + co] = s;
                }
            }
}

/* dX[q][ci] += sum_{t,co} W[co][t][ci] dY[q - off(t)][co];
 * dW[co][t][ci] += sum_p dY[p][co] X[p + off(t)][ci]; db[co] += sum_p dY[p][co] */
static void conv_bwd(int B, int H, int W, int Cin, int Cout, const double *X, const double *Wt, const double *dY,
                     double *dX, double *dW, double *db)
{
    for (int n = 0; n < B; n++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                const size_t p = ((size_t)n * H + y) * W + x;
                for (int co = 0; co < Cout; co++) {
                    const double g = dY[p * Cout + co];
                    if (g == 0.0) continue;
                    if (db) db[co] += g;
                    for (int t = 0; t < 9; t++) {
                        const int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
                        if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
                        const size_t q = ((size_t)n * H + yy) * W + xx;
                        const double *wr = Wt + ((size_t)co * 9 + t) * Cin;
                        double *dwr = dW ? dW + ((size_t)co * 9 + t) * Cin : NULL;
                        for (int ci = 0; ci < Cin; ci++) {
                            if (dwr) dwr[ci] += g * X[q * Cin + ci];
                            if (dX) dX[q * Cin + ci] += g * wr[ci];
                        }
                    }
                }
            }
}

typedef struct {
    int kind; /* 0 conv, 1 pool, 2 up */
    int in, out, skip, cin, cout, relu;
    size_t woff, boff;
} OrcOp;

typedef struct {
    int level, c, f32;
} OrcTensor;

#define ORC_MAX_T 64
#define ORC_MAX_O 64

typedef struct {
    OrcTensor T[ORC_MAX_T];
    OrcOp O[ORC_MAX_O];
    int nt, no;
    size_t np;
} OrcNet;

static int orc_tensor(OrcNet *g, int level, int c, int f32)
{
    g->T[g->nt].level = level;
    g->T[g->nt].c = c;
    g->T[g->nt].f32 = f32;
    return g->nt++;
}

static int orc_conv(OrcNet *g, int in, int cout, int relu, int skip, int f32)
{
    OrcOp *o = &g->O[g->no++];
    o->kind = 0;
    o->in = in;
    o->cin = g->T[in].c;
    o->cout = cout;
    o->relu = relu;
    o->skip = skip;
    o->out = orc_tensor(g, g->T[in].level, cout, f32);
    o->woff = g->np;
    g->np += (size_t)cout * 9 * o->cin;
    o->boff = g->np;
    g->np += (size_t)cout;
    return o->out;
}

static int orc_simple(OrcNet *g, int kind, int in)
{
    OrcOp *o = &g->O[g->no++];
    o->kind = kind;
    o->in = in;
    o->skip = -1;
    o->out = orc_tensor(g, g->T[in].level + (kind == 1 ? 1 : -1), g->T[in].c, 0);
    return o->out;
}

static void orc_build(OrcNet *g, int levels, int base, int cpad)
{
    memset(g, 0, sizeof(*g));
    int skip_t[8];
    int x = orc_tensor(g, 0, cpad, 0);
    for (int l = 0; l < levels; l++) {
        const int c = (base << l) < 256 ? (base << l) : 256;
        if (l > 0) x = orc_simple(g, 1, x);
        x = orc_conv(g, x, c, 1, -1, 0);
        x = orc_conv(g, x, c, 1, -1, 0);
        skip_t[l] = x;
    }
    for (int l = levels - 2; l >= 0; l--) {
        const int c = (base << l) < 256 ? (base << l) : 256;
        x = orc_simple(g, 2, x);
        x = orc_conv(g, x, c, 1, skip_t[l], 0);
        x = orc_conv(g, x, c, 1, -1, 0);
    }
    orc_conv(g, x, 32, 0, -1, 1);
}

size_t orc_cnn_num_params(int levels, int base, int cpad)
{
    OrcNet g;
    orc_build(&g, levels, base, cpad);
    return g.np;
}

/* Forward (and with grads != NULL the backward) of a batch: x [B][H][W][cpad],
 * base and target [B][3][H][W], params in the network's order.  Returns the
 * loss; pred (nullable) receives base + head [B][3][H][W]. */
double orc_cnn_forward_backward(int levels, int base_ch, int cpad, int B, int H, int W, const double *x,
                                const double *params, const double *base, const double *target, int round_bf16,
                                double *pred, double *grads)
{
    OrcNet g;
    orc_build(&g, levels, base_ch, cpad);
    double *w = (double *)malloc(g.np * sizeof(double));
    memcpy(w, params, g.np * sizeof(double));
    maybe_round(w, g.np, round_bf16);
    double *A[ORC_MAX_T] = {0}, *dA[ORC_MAX_T] = {0};
    size_t nel[ORC_MAX_T] = {0};
    for (int i = 0; i < g.nt; i++) {
        nel[i] = (size_t)B * (H >> g.T[i].level) * (W >> g.T[i].level) * g.T[i].c;
        A[i] = (double *)calloc(nel[i], sizeof(double));
        dA[i] = (double *)calloc(nel[i], sizeof(double));
    }
    memcpy(A[0], x, nel[0] * sizeof(double));
    maybe_round(A[0], nel[0], round_bf16);
    for (int k = 0; k < g.no; k++) {
        const OrcOp *o = &g.O[k];
        const int h = H >> g.T[o->in].level, wd = W >> g.T[o->in].level;
        const int C = g.T[o->in].c;
        if (o->kind == 0) {
            orc_conv3x3(B, h, wd, o->cin, o->cout, A[o->in], w + o->woff, w + o->boff, A[o->out]);
            for (size_t i = 0; i < nel[o->out]; i++) {
                double v = A[o->out][i];
                if (o->skip >= 0) v += A[o->skip][i];
                if (o->relu && v < 0.0) v = 0.0;
                A[o->out][i] = v;
            }
            if (!g.T[o->out].f32) maybe_round(A[o->out], nel[o->out], round_bf16);
        } else if (o->kind == 1) {
            for (int n = 0; n < B; n++)
                for (int y = 0; y < h / 2; y++)
                    for (int xx = 0; xx < wd / 2; xx++)
                        for (int c = 0; c < C; c++) {
                            const double *r0 = A[o->in] + (((size_t)n * h + 2 * y) * wd + 2 * xx) * C + c;
                            const double *r1 = r0 + (size_t)wd * C;
                            A[o->out][(((size_t)n * (h / 2) + y) * (wd / 2) + xx) * C + c] =
                                0.25 * ((r0[0] + r0[C]) + (r1[0] + r1[C]));
                        }
            maybe_round(A[o->out], nel[o->out], round_bf16);
        } else {
            for (int n = 0; n < B; n++)
                for (int y = 0; y < 2 * h; y++)
                    for (int xx = 0; xx < 2 * wd; xx++)
                        for (int c = 0; c < C; c++)
                            A[o->out][(((size_t)n * 2 * h + y) * 2 * wd + xx) * C + c] =
                                A[o->in][(((size_t)n * h + y / 2) * wd + xx / 2) * C + c];
        }
    }
    /* loss */
    const int head = g.O[g.no - 1].out;
    const size_t HW = (size_t)H * W;
    double loss = 0.0;
    for (int n = 0; n < B; n++)
        for (int c = 0; c < 3; c++)
            for (size_t q = 0; q < HW; q++) {
                const size_t i = ((size_t)n * 3 + c) * HW + q;
                const double pv = base[i] + A[head][((size_t)n * HW + q) * 32 + c];
                if (pred) pred[i] = pv;
                const double d = pv - target[i];
                loss += d * d;
                dA[head][((size_t)n * HW + q) * 32 + c] = 2.0 * d / ((double)B * 3 * HW);
            }
    loss /= (double)B * 3 * HW;
    if (grads) {
        memset(grads, 0, g.np * sizeof(double));
        for (int k = g.no - 1; k >= 0; k--) {
            const OrcOp *o = &g.O[k];
            const int h = H >> g.T[o->in].level, wd = W >> g.T[o->in].level;
            const int C = g.T[o->in].c;
            if (o->kind == 1) {
                for (int n = 0; n < B; n++)
                    for (int y = 0; y < h / 2; y++)
                        for (int xx = 0; xx < wd / 2; xx++)
                            for (int c = 0; c < C; c++) {
                                const double gq = 0.25 * dA[o->out][(((size_t)n * (h / 2) + y) * (wd / 2) + xx) * C + c];
                                double *r0 = dA[o->in] + (((size_t)n * h + 2 * y) * wd + 2 * xx) * C + c;
                                double *r1 = r0 + (size_t)wd * C;
                                r0[0] += gq;
                                r0[C] += gq;
                                r1[0] += gq;
                                r1[C] += gq;
                            }
            } else if (o->kind == 2) {
                for (int n = 0; n < B; n++)
                    for (int y = 0; y < 2 * h; y++)
                        for (int xx = 0; xx < 2 * wd; xx++)
                            for (int c = 0; c < C; c++)
                                dA[o->in][(((size_t)n * h + y / 2) * wd + xx / 2) * C + c] +=
                                    dA[o->out][(((size_t)n * 2 * h + y) * 2 * wd + xx) * C + c];
            } else {
                double *dU = (double *)malloc(nel[o->out] * sizeof(double));
                for (size_t i = 0; i < nel[o->out]; i++) {
                    double gq = dA[o->out][i];
                    if (o->relu && !(A[o->out][i] > 0.0)) gq = 0.0;
                    if (o->skip >= 0) dA[o->skip][i] += gq;
                    dU[i] = gq;
                }
                maybe_round(dU, nel[o->out], round_bf16);
                conv_bwd(B, h, wd, o->cin, o->cout, A[o->in], w + o->woff, dU, o->in == 0 ? NULL : dA[o->in],
                         grads + o->woff, grads + o->boff);
                free(dU);
            }
        }
    }
    for (int i = 0; i < g.nt; i++) {
        free(A[i]);
        free(dA[i]);
    }
    free(w);
    return loss;
}
