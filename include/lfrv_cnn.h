/*
 * lfrv_cnn.h -- the learned warm start: a small U-Net on the bf16 MFMA
 * units, trained on the GPU on solutions the solver itself produces, that
 * predicts the starting image of a novel-view solve from the stack of
 * warped views.
 *
 * The reference's counterpart of the hook is init_name: a starting image
 * read from a PFM in place of the splatted average (gradientIBR.cu:737-762,
 * optimization.cu:1277-1299).  Here the start comes from the network, and
 * the solver still runs to its own tolerance from it; the network's start is
 * kept only when it lowers the energy of the normal equations, else the
 * splatted average (the reference's start) is used.
 *
 * Network (all 3x3 convolutions with bias, "same" zero padding):
 *   level l = 0 .. levels-1 has c_l = min(base << l, 256) channels and
 *   resolution H / 2^l.  Encoder: conv-ReLU-conv-ReLU per level, 2x2 average
 *   pooling between levels.  Decoder: nearest 2x upsampling, conv, + the
 *   encoder output of that level, ReLU, conv-ReLU.  Head: conv to 3 channels
 *   (computed as 32, the rest zero), no activation; the prediction is
 *   base + head, with base the solver's own starting image.
 * Input: in_channels planes per pixel (25 views x (R, G, B, validity) = 100
 * for 5 x 5 views), zero-padded to a multiple of 32.  Activations are NHWC
 * bf16, weights bf16 [Cout][9][Cin] with FP32 master copies, FP32
 * accumulation.  H and W must be multiples of 2^(levels-1).
 *
 * Parameter order (lfrv_cnn_get_params / set_params / get_grads), for each
 * convolution in forward order: W [Cout][9][Cin_padded] then b [Cout], with
 * the head's Cout = 32 (channels 3.. are zero and stay zero).
 */
#ifndef LFRV_CNN_H
#define LFRV_CNN_H

#include <stddef.h>
#include <stdint.h>

#include "lfrv.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lfrv_cnn lfrv_cnn;

typedef struct {
    int in_channels;   /* input planes per pixel (100 for 25 views) */
    int levels;        /* resolution levels (4) */
    int base_channels; /* channels of level 0 (32; a multiple of 32) */
    int max_batch;     /* images per call at most */
    int max_h, max_w;  /* largest input size */
    uint64_t seed;     /* weight initialisation (He normal; the head 1e-2 of it) */
} lfrv_cnn_params;

lfrv_status lfrv_cnn_create(const lfrv_cnn_params *p, lfrv_cnn **out);
void lfrv_cnn_destroy(lfrv_cnn *cnn);
const char *lfrv_cnn_last_error(const lfrv_cnn *cnn);

/* Number of FP32 parameters, and their copies (host or device pointers). */
lfrv_status lfrv_cnn_num_params(lfrv_cnn *cnn, size_t *n);
lfrv_status lfrv_cnn_get_params(lfrv_cnn *cnn, float *out);
lfrv_status lfrv_cnn_set_params(lfrv_cnn *cnn, const float *in);
lfrv_status lfrv_cnn_get_grads(lfrv_cnn *cnn, float *out);

/* The batch of the next forward / training call: B images of H x W, image i
 * from x[i] (bf16 NHWC [H][W][Cpad] on the device, Cpad = in_channels
 * rounded up to 32, as lfrv_warped_stack writes it), base[i] and target[i]
 * (planar FP32 [3][H][W] on the device; target may be NULL for inference). */
lfrv_status lfrv_cnn_set_batch(lfrv_cnn *cnn, int B, int H, int W, const void *const *x, const float *const *base,
                               const float *const *target);

/* The same from host FP32 arrays: x [B][H][W][in_channels] (converted to
 * padded bf16), base and target [B][3][H][W] (target nullable). */
lfrv_status lfrv_cnn_set_batch_host(lfrv_cnn *cnn, int B, int H, int W, const float *x, const float *base,
                                    const float *target);

/* Forward pass of the batch; pred (nullable, host or device) receives
 * base + head as planar FP32 [B][3][H][W]. */
lfrv_status lfrv_cnn_forward(lfrv_cnn *cnn, float *pred);

/* Forward, squared-error loss mean((pred - target)^2) and backward: the
 * gradients of every parameter (lfrv_cnn_get_grads).  loss nullable. */
lfrv_status lfrv_cnn_loss_grad(lfrv_cnn *cnn, double *loss);

/* lfrv_cnn_loss_grad then one Adam step (beta1 0.9, beta2 0.999, eps 1e-8)
 * with learning rate lr. */
lfrv_status lfrv_cnn_train_step(lfrv_cnn *cnn, float lr, double *loss);

/* Floating-point operations of one forward pass of the current batch (the
 * convolutions' multiply-adds x 2, padded channels included). */
lfrv_status lfrv_cnn_forward_flops(lfrv_cnn *cnn, double *flops);

/* Device time of the convolution kernels of the calls since the last reset
 * (events around each convolution launch when enable != 0). */
lfrv_status lfrv_cnn_set_timing(lfrv_cnn *cnn, int enable);
lfrv_status lfrv_cnn_get_timing(lfrv_cnn *cnn, double *conv_ms, int64_t *conv_launches, double *conv_flops);

/* ---- the convolutions as stateless operators (device pointers) ------------ */

/* Y (FP32 NHWC [B][H][W][Cout]) = conv3x3(X) + bias: X bf16 NHWC [B][H][W][Cin],
 * Wt bf16 [Cout][9][Cin], bias FP32 (nullable).  Cin a multiple of 32. */
lfrv_status lfrv_op_conv3x3(int B, int H, int W, int Cin, int Cout, const void *X, const void *Wt, const float *bias,
                            float *Y, void *stream);

/* dW (FP32 [Cout][9][Cin]) += sum_p dU[p][co] X[p + off(t)][ci]: dU bf16 NHWC
 * [B][H][W][Cout], X bf16 NHWC [B][H][W][Cin].  Cin a multiple of 32, Cout of 8. */
lfrv_status lfrv_op_conv3x3_wgrad(int B, int H, int W, int Cin, int Cout, const void *dU, const void *X, float *dW,
                                  void *stream);

/* ---- the warm start on a solver context ------------------------------------ */

/* The stack of warped views of a context: for each view i, its colours
 * splatted into the target (sum_m B_i[m,o] v_i(m) / sum_m B_i[m,o], the
 * footprint weights of the current warps) and its validity
 * min(1, sum_m B_i[m,o]), as bf16 NHWC [H][W][cpad] with channel 4i + c
 * (c = R, G, B, validity) and zero padding channels (cpad >= 4 nviews, a
 * multiple of 32).  out is device memory.  Needs lfrv_prepare (or set tau). */
lfrv_status lfrv_warped_stack(lfrv_ctx *ctx, void *out, int cpad);

typedef struct {
    double energy_plain; /* normal-equation energy of the splatted-average start */
    double energy_net;   /* ... of the network's start */
    int used_net;        /* 1: the network's start was kept (it lowered the energy) */
    double ms;           /* wall time of the call (host clock) */
} lfrv_warm_stats;

/* Replace the context's current solution (the start lfrv_compute_initial_image
 * made, with lfrv_init_u_gradient done) by the network's prediction when that
 * lowers the energy of the normal equations (lfrv_gibr_energy); otherwise keep
 * it.  The network's input is lfrv_warped_stack of the context, its base the
 * current solution; the context's target size must be one the network takes.
 * Afterwards lfrv_cg_solve runs from whichever start was kept. */
lfrv_status lfrv_warm_start(lfrv_ctx *ctx, lfrv_cnn *cnn, lfrv_warm_stats *st);

#ifdef __cplusplus
}
#endif
#endif /* LFRV_CNN_H */
