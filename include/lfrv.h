/*
 * lfrv.h -- C-ABI of the MI355X-native variational light-field solver.
 *
 * One context (lfrv_ctx) holds one novel-view problem on one GPU: the input
 * views, their warps tau into the target, the per-view weights omega and the
 * target-side state of the two reference solvers:
 *
 *   LFRV_MODE_GRADIENT_IBR : gradient-domain IBR ("IBR_gradient",
 *                            sources/gradientIBR.cu, sources/solver.cu)
 *   LFRV_MODE_COLOR_SR     : TV super-resolution with warped data terms
 *                            ("IBR_color", sources/optimization.cu,
 *                            cocolib/vtv/vtv_rof.cu)
 *
 * Conventions (identical to the reference):
 *   - images are planar FP32, RRR..GGG..BBB.., rows contiguous, top-left
 *     origin (gradientIBR.cuh:16-18);
 *   - views are w x h, the target is W x H = (w*dsf) x (h*dsf);
 *   - tau is given per view pixel in target pixel units with a corner origin
 *     (target pixel x spans [x, x+1)); tau < 0 or tau > W/H marks an invalid
 *     pixel (gradientIBR_kernels.cuh:216-225);
 *   - sigma_sensor is already divided by 255 (config.cpp:118-119).
 *
 * Pointers passed in may be host or device memory (hipMemcpyDefault); the
 * library copies them and keeps nothing of the caller's after return.  All
 * device memory is owned by the context.  A context is used from one host
 * thread at a time.  Calls are asynchronous on the context's stream unless
 * they return data to the host (get_* / stats) or say otherwise.
 *
 * Errors are returned as lfrv_status codes; lfrv_last_error() gives text.
 * Nothing in the library aborts the process (the reference asserted,
 * cuda_helper.h:38-44).
 */
#ifndef LFRV_H
#define LFRV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    LFRV_OK = 0,
    LFRV_E_ARG = 1,          /* invalid argument or size */
    LFRV_E_HIP = 2,          /* HIP runtime error */
    LFRV_E_NONCONVERGED = 3, /* iterative solve hit its iteration limit */
    LFRV_E_NAN = 4,          /* non-finite value detected */
    LFRV_E_STATE = 5,        /* call out of order (e.g. solve before prepare) */
    LFRV_E_BREAKDOWN = 6     /* CG breakdown: d.q <= 0 */
} lfrv_status;

typedef enum { LFRV_MODE_GRADIENT_IBR = 0, LFRV_MODE_COLOR_SR = 1 } lfrv_mode;

/* Solver parameters; field names follow Config_data (config.h:11-122). */
typedef struct {
    int mode;                /* lfrv_mode */
    int nviews;              /* _nviews */
    int w, h;                /* _w, _h: input view size */
    int dsf;                 /* _dsf: target = (w*dsf) x (h*dsf) */
    float sigma_sensor;      /* _sigma_sensor (already /255) */
    float ugrad_threshold;   /* _ugrad_threshold */
    int dw_type;             /* _dw_type (only 0 is implemented, as in the reference) */
    int gw_type;             /* _gw_type */
    float dt_alpha;          /* _dt_alpha: intensity data term weight */
    float dt_beta;           /* _dt_beta: gradient term weight */
    float gradient_step;     /* _gradient_step */
    float lambda;            /* _lambda: TV weight (IBR_color) */
    float lambda_max_factor; /* _lambda_max_factor */
    int inner_iterations;    /* ROF inner iterations per outer step (coco_vtv_data::_inner_iterations) */
    float occlusion_tol_px;  /* round-trip tolerance of the depth visibility test (import.cpp:436-441: 10) */
    float sigma_z;           /* depth std-dev for partial tau (tauPartial.frag:25: 10) */
    int backproject_source_cam; /* 1: import.cpp:423-424 literal (view R,C); 0: target camera */
    int tile_w, tile_h;      /* target tile of the splat (pull) kernels; 0 = default 32 x 8 */
    int pull_lds_kb;         /* LDS budget of a pull's entry cache in KiB; 0 = default 48 */
} lfrv_params;

/* Fill *p with the reference defaults (config.cpp:11-64) for a problem of the
 * given mode and size. */
void lfrv_default_params(lfrv_params *p, int mode, int nviews, int w, int h, int dsf);

typedef struct lfrv_ctx lfrv_ctx;

/* ---- lifecycle ------------------------------------------------------------ */

/* Allocate every device buffer of the problem.  Replaces init_data
 * (gradientIBR.cu:155-299) and coco_vtv_alloc + coco_vtv_sr_init_unstructured
 * (vtv.cu:14-136, optimization.cu:654-788). */
lfrv_status lfrv_create(const lfrv_params *p, lfrv_ctx **out);

/* Release everything.  Replaces free_data (gradientIBR.cu:302-348),
 * coco_vtv_sr_free_unstructured + coco_vtv_free. */
void lfrv_destroy(lfrv_ctx *ctx);

/* Text of the last error on this context (never NULL). */
const char *lfrv_last_error(const lfrv_ctx *ctx);

/* Wait for all work queued on the context's stream. */
lfrv_status lfrv_sync(lfrv_ctx *ctx);

/* ---- inputs --------------------------------------------------------------- */

/* View i, planar RGB w*h*3.  Replaces create_view (gradientIBR.cu:524-555),
 * coco_vtv_sr_create_view_unstructured (optimization.cu:1028-1063). */
lfrv_status lfrv_set_view(lfrv_ctx *ctx, int i, const float *rgb);

/* Warp of view i: tau_x, tau_y (w*h each).  Replaces read_tau
 * (gradientIBR.cu:651-679), coco_vtv_sr_read_tau (optimization.cu:1170-1205). */
lfrv_status lfrv_set_tau(lfrv_ctx *ctx, int i, const float *tau_x, const float *tau_y);

/* Partial tau of view i: the three PFM channels (sigma_z, dtau_x/dz,
 * dtau_y/dz); stored as sigma_z * dtau/dz.  Replaces read_partial_tau
 * (gradientIBR.cu:682-721), coco_vtv_sr_read_partial_tau (optimization.cu:1208-1256). */
lfrv_status lfrv_set_partial_tau(lfrv_ctx *ctx, int i, const float *sigma_z, const float *dtdz_x,
                                 const float *dtdz_y);

/* Depth map of view i (w*h, z-depth) and its 3x4 row-major projection
 * matrix P (PMVS CONTOUR layout, import.cpp:219-240). */
lfrv_status lfrv_set_depth(lfrv_ctx *ctx, int i, const float *depth, const double *P);

/* 3x4 row-major projection matrix of the target camera (import.cpp:255-260). */
lfrv_status lfrv_set_target_camera(lfrv_ctx *ctx, const double *P);

/* Depth gradients of view i (dz/dx, dz/dy per pixel, w*h each; both NULL: zero),
 * the two normal-map channels the OpenGL producer reads with each depth map
 * (InputView::setDepthAndNormals, view.cpp:608-636; zero when no normals,
 * initDepthAndNormals :639-661).  Only lfrv_warps_from_target_depth uses them. */
lfrv_status lfrv_set_depth_gradient(lfrv_ctx *ctx, int i, const float *dzdx, const float *dzdy);

/* ---- warp and splat structure -------------------------------------------- */

/* tau and sigma_z*dtau/dz of every view from its depth map and the cameras,
 * with target z-buffer visibility.  Replaces import_warps_from_depth
 * (import.cpp:243-450) plus the partial-tau producer (import.cpp:658-680,
 * tauPartial.frag:23-47); runs on the GPU.  In LFRV_MODE_GRADIENT_IBR the
 * sigma_z*dtau/dz planes are formed from the same depth and cameras when first
 * read (lfrv_get_buffer, weights of a non-zero solution) or before an input they
 * depend on changes (lfrv_set_depth, lfrv_set_partial_tau, lfrv_set_buffer):
 * the values are the same, the first weights of a solve do not read them. */
lfrv_status lfrv_warps_from_depth(lfrv_ctx *ctx);

/* The OpenGL warp producer (the reference's fromMVE2coco path, dsf = 1 there),
 * on the GPU: Scene::computeTargetDepthMap (scene.cpp:1506-1544) splats every
 * view but skip_view (-1: none) as one quad per pixel of valid depth (z < 500),
 * corners at the pixel centre +- 0.5 at depths z +- 0.5 dz/dx +- 0.5 dz/dy
 * (depthSplatting.geom:21-61), into the target z-buffer; each covered target
 * pixel centre keeps the target depth of the nearest quad's centre
 * (depthSplatting.frag), 0 elsewhere (LFRV_BUF_TARGET_DEPTH).  Then
 * InputView::exportWarp (view.cpp:663-721) for every view: tau = x_u, or -2
 * outside (0, W) x (0, H) or at invalid depth, or -1 where the target depth map
 * (bilinear, clamped) is nearer by more than 0.1 (tauWarp.frag:25-74); the
 * deformation weight (LFRV_BUF_TAU_WEIGHT) and sigma_z dtau/dz
 * (tauPartial.frag:23-47).  Rasterisation: pixel centres inside either
 * triangle of the strip, edges owned by one triangle, depth test on 1/z
 * interpolated in the image; quads spanning more than 64 target pixels or
 * crossing the near plane z = 0.01 are skipped (GL would clip them). */
lfrv_status lfrv_warps_from_target_depth(lfrv_ctx *ctx, int skip_view);

/* Build the splat structure from the current warps: bins every cell-valid
 * source pixel of every view by target tile, computes the target coverage
 * sum_i B_i^T 1 and the per-view Laplacian / Sobel planes of the views.
 * Replaces compute_sparse_matrix (gradientIBR.cu:351-378, optimization.cu:838-877;
 * the footprint weights are recomputed from tau in every kernel instead of
 * being stored) and init_forward_warp_structure (gradientIBR.cu:383-521,
 * optimization.cu:882-1025; the host segment colouring is replaced by
 * tile binning + LDS accumulation). */
lfrv_status lfrv_prepare(lfrv_ctx *ctx);

/* ---- shared steps ---------------------------------------------------------- */

/* Backward visibility mask.  GRADIENT_IBR: any footprint hit
 * (setup_target_visibility, gradientIBR.cu:40-83).  COLOR_SR: coverage >= 1
 * (coco_vtv_setup_visibility_mask, optimization.cu:33-135). */
lfrv_status lfrv_setup_visibility(lfrv_ctx *ctx);

/* Weights omega_i = angular x deformation from the current solution.
 * Replaces compute_weights (gradientIBR.cu:558-648) and
 * coco_vtv_sr_compute_weights_unstructured (optimization.cu:1066-1167). */
lfrv_status lfrv_compute_weights(lfrv_ctx *ctx);

/* Starting image.  u0 != NULL loads it (planar RGB W*H; the reference's
 * init_name PFM).  Otherwise the omega-weighted splat average of the views,
 * then hole filling: Poisson diffusion (GRADIENT_IBR, gradientIBR.cu:87-152)
 * or push-pull (COLOR_SR, optimization.cu:544-651).  Replaces
 * compute_initial_image (gradientIBR.cu:724-819) and
 * coco_vtv_sr_compute_initial_image (optimization.cu:1259-1360). */
lfrv_status lfrv_compute_initial_image(lfrv_ctx *ctx, const float *u0);

/* The last Poisson hole filling of lfrv_compute_initial_image: the number of
 * hole pixels (not visible, setup_target_visibility) and of diffusion sweeps
 * run before its stopping rule (gradientIBR.cu:87-152: 500 at most, relative
 * change < 1e-3). */
lfrv_status lfrv_get_hole_stats(lfrv_ctx *ctx, int *holes, int *sweeps);

/* Current solution, planar RGB W*H.  Replaces get_solution
 * (gradientIBR.cu:20-37) and coco_vtv_get_solution (vtv.cu:268-287). */
lfrv_status lfrv_get_solution(lfrv_ctx *ctx, float *rgb);

/* Overwrite the current solution (planar RGB W*H). */
lfrv_status lfrv_set_solution(lfrv_ctx *ctx, const float *rgb);

/* Zero the current solution on the device (the cleared U of init_data,
 * gradientIBR.cu:281-287, and of compute_initial_image, :732-735).  For
 * IBR_color it also starts a new view from the state of a fresh workspace,
 * as coco_vtv_alloc and coco_vtv_initialize(ZERO) leave it (vtv.cu:60-89,
 * 241-263): Uq, G, the dual variables and the normalisation buffer zeroed,
 * the FISTA weight alpha = 1, no TV weight. */
lfrv_status lfrv_clear_solution(lfrv_ctx *ctx);

/* ---- gradient-domain IBR --------------------------------------------------- */

/* Target gradient g and target mask by splatting the views' Sobel gradients.
 * Replaces init_u_gradient (gradientIBR.cu:833-920). */
lfrv_status lfrv_init_u_gradient(lfrv_ctx *ctx);

typedef struct {
    int iterations[3];          /* CG iterations run per RGB channel */
    int converged[3];
    double rel_residual[3];     /* recursive ||r|| / ||b|| at exit (FP64) */
    double true_rel_residual[3];/* ||b - A x|| / ||b|| recomputed at exit */
    int status;                 /* lfrv_status of the solve */
    double ms;                  /* wall time of the solve (host clock) */
} lfrv_cg_stats;

/* Solve the gradient-constraint normal equations per RGB channel,
 *   (dt_alpha sum_i S_i^T W_i S_i + dt_beta D^T D) u = dt_alpha sum_i S_i^T W_i v_i + dt_beta D^T g,
 * W_i = omega_i sigma^2, D forward differences with Neumann boundary, by
 * Jacobi-preconditioned CG from the current solution to ||r|| <= tol ||b||.
 * FP32 vectors, FP64 reductions, device-side scalars.  The CG of
 * possion_conj_grad (solver.cu:526-699) with its data term replaced by the
 * warped intensity term of vtv_sr_dataterm_intensities (optimization.cu:1835-1966). */
lfrv_status lfrv_cg_solve(lfrv_ctx *ctx, double tol, int maxit, lfrv_cg_stats *st);

/* The device-resident part of the IBR_gradient driver (gradientIBR.cpp:12-150
 * without its file I/O), in one call: U = 0 (init_data), warps from depth when
 * every view's depth and the target camera are set (else the tau / dpart set
 * with lfrv_set_tau / lfrv_set_partial_tau), lfrv_prepare,
 * lfrv_setup_visibility, lfrv_compute_weights, lfrv_compute_initial_image
 * (NULL), lfrv_init_u_gradient and lfrv_cg_solve (tol, maxit, st).  Same
 * results as those calls made one by one. */
lfrv_status lfrv_ibr_gradient(lfrv_ctx *ctx, double tol, int maxit, lfrv_cg_stats *st);

/* The screened Poisson system possion_conj_grad sets up (solver.cu:361-505):
 * (lambda I + D^T D) u = lambda u0 + D^T g, u0 = the current solution,
 * plain CG (use_precond = 0) or Jacobi PCG, to ||r|| <= tol ||b||. */
lfrv_status lfrv_poisson_cg(lfrv_ctx *ctx, double lambda, double tol, int maxit, int use_precond,
                            lfrv_cg_stats *st);

/* Energy of the current solution for the normal equations of lfrv_cg_solve:
 * E = dt_alpha/2 sum ||W^1/2 (S u - v)||^2 + dt_beta/2 ||D u - g||^2 (FP64). */
lfrv_status lfrv_gibr_energy(lfrv_ctx *ctx, double *energy);

/* ---- IBR_color: TV primal-dual -------------------------------------------- */

/* TV weight g from the splatted omega.  Replaces
 * coco_vtv_sr_init_regularizer_weight_unstructured (optimization.cu:2251-2338). */
lfrv_status lfrv_sr_init_regularizer_weight(lfrv_ctx *ctx);

/* One outer step: data-term gradient, F = clamp(U - G), inner_iterations
 * primal-dual ROF steps, over-relaxation.  Replaces
 * coco_vtv_sr_iteration_fista_unstructured (optimization.cu:1736-1748) and
 * vtv_sr_ista_step_unstructured (:1751-1832). */
lfrv_status lfrv_sr_iteration_fista(lfrv_ctx *ctx);

/* `steps` outer steps in one call (queued back to back on the stream, no host
 * wait): the loop body of super-resolution.cpp:218-227 without its energy and
 * moving-pixel checks.  Same results as `steps` lfrv_sr_iteration_fista calls. */
lfrv_status lfrv_sr_iterate(lfrv_ctx *ctx, int steps);

/* Primal energy in FP64; parts (nullable) receives E_TV, E_int, E_grad.
 * Replaces coco_vtv_sr_primal_energy_unstructured (optimization.cu:1477-1717). */
lfrv_status lfrv_sr_primal_energy(lfrv_ctx *ctx, double *energy, double *parts);

/* Pixels whose max-channel change against `prev` (planar RGB W*H, host or
 * device) is >= max_diff, i.e. the "moving" count of sr_compare_images
 * (super-resolution.cpp:16-51). */
lfrv_status lfrv_sr_count_moving(lfrv_ctx *ctx, const float *prev, float max_diff, int64_t *moving);

typedef struct lfrv_sr_stats {
    int meta_iterations;  /* re-weighting passes started */
    int fista_steps;      /* FISTA steps run, all passes */
    int last_stop;        /* how the last pass's FISTA loop ended: LFRV_SR_STOP_* */
    int converged;        /* 1 when a re-weighting rule ended the solve before nmeta */
    float energy;         /* energy after the last step, in float as the driver keeps it */
    double ms;            /* wall time (host clock) */
} lfrv_sr_stats;

enum {
    LFRV_SR_STOP_MAXIT = 0,     /* niter steps ran */
    LFRV_SR_STOP_ENERGY = 1,    /* energy < 1e-4 */
    LFRV_SR_STOP_RELATIVE = 2,  /* |E - E_prev| / E < 1e-4 */
    LFRV_SR_STOP_MOVING = 3     /* < 1% of the pixels changed by >= 1e-3 */
};

/* The re-weighted least-squares loop of the IBR_color driver
 * (super-resolution.cpp:177-300) from the current solution: up to nmeta
 * passes of lfrv_compute_weights, lfrv_sr_init_regularizer_weight and up to
 * niter FISTA steps, with the driver's stopping rules in float (EPSILON_TOL
 * 1e-4, MAX_DIFF_THRESHOLD 1e-3, PERCENT_MOVING 0.01) and its previous
 * solutions kept on the device.  Each step's energy and moving-pixel test
 * run on the device, which skips the steps enqueued after a rule fired; the
 * host reads the loop state every few steps.  Same results as the loop made
 * of lfrv_sr_iteration_fista, lfrv_sr_primal_energy and lfrv_sr_count_moving
 * calls. */
lfrv_status lfrv_sr_solve(lfrv_ctx *ctx, int niter, int nmeta, lfrv_sr_stats *st);

/* ---- inspection ------------------------------------------------------------ */

typedef enum {
    LFRV_BUF_TAU_X = 0,      /* V*w*h */
    LFRV_BUF_TAU_Y = 1,      /* V*w*h */
    LFRV_BUF_DPART_X = 2,    /* V*w*h */
    LFRV_BUF_DPART_Y = 3,    /* V*w*h */
    LFRV_BUF_OMEGA = 4,      /* V*w*h */
    LFRV_BUF_U = 5,          /* 3*W*H */
    LFRV_BUF_UQ = 6,         /* 3*W*H */
    LFRV_BUF_G = 7,          /* 3*W*H: data-term gradient / F of the last step */
    LFRV_BUF_PX = 8,         /* 3*W*H dual x */
    LFRV_BUF_PY = 9,         /* 3*W*H dual y */
    LFRV_BUF_REG_G = 10,     /* W*H TV weight */
    LFRV_BUF_NORM_MASK = 11, /* W*H */
    LFRV_BUF_COVERAGE = 12,  /* W*H */
    LFRV_BUF_U_GRAD_X = 13,  /* W*H */
    LFRV_BUF_U_GRAD_Y = 14,  /* W*H */
    LFRV_BUF_TARGET_MASK = 15, /* W*H */
    LFRV_BUF_VISIBILITY = 16,  /* W*H, 0/1 as float */
    LFRV_BUF_TARGET_DEPTH = 17,/* W*H z-buffer of lfrv_warps_from_depth */
    LFRV_BUF_CG_B = 18,      /* 3*W*H right-hand side of the last CG solve */
    LFRV_BUF_CG_MINV = 19,   /* W*H Jacobi preconditioner of the last CG solve */
    LFRV_BUF_TAU_WEIGHT = 20 /* V*w*h deformation weights of lfrv_warps_from_target_depth (the B channel of
                                the tau PFMs InputView::exportWarp writes) */
} lfrv_buffer;

/* Copy an internal buffer to `out` (host or device, float); *count receives
 * the number of floats.  out == NULL only queries the size. */
lfrv_status lfrv_get_buffer(lfrv_ctx *ctx, int which, float *out, size_t *count);

/* Overwrite an internal buffer (same layouts as lfrv_get_buffer). */
lfrv_status lfrv_set_buffer(lfrv_ctx *ctx, int which, const float *in);

/* Per-kernel-class GPU time recorded with HIP events on the context stream
 * while timing is enabled: class 0 = CG operator, 1 = CG vector update,
 * 2 = PD step, 3 = data term, 4 = warp from depth, 5 = splat/setup.  enable = 0
 * turns it off; enable = n >= 1 times one launch (classes 0-3, with events the
 * dispatch itself stamps) or one call (classes 4-5, events recorded around it)
 * in n of each class: a timed dispatch or an event record waits for its
 * completion signal, several microseconds each.  lfrv_get_timing's launch
 * count is the number timed. */
lfrv_status lfrv_set_timing(lfrv_ctx *ctx, int enable);
lfrv_status lfrv_get_timing(lfrv_ctx *ctx, int kclass, double *total_ms, int64_t *launches);

/* Number of binned (view, pixel, tile) entries of the splat structure. */
lfrv_status lfrv_get_bin_stats(lfrv_ctx *ctx, int64_t *entries, int *tiles);

/* ---- stateless operators (device pointers, caller's stream) ---------------- */

/* v = S u for one view: gather C planes of the W x H image u at the footprint
 * of tau (gradientIBR_kernels.cuh:196-290).  All pointers are device memory. */
lfrv_status lfrv_op_gather(int W, int H, int w, int h, int dsf, const float *tau_x,
                           const float *tau_y, const float *u, int C, float *out, void *stream);

/* out += S^T (wgt * in) for one view with float atomics (the reference's
 * deconvolution_nonsep_param, gradientIBR_kernels.cuh:293-388, in atomic
 * form); wsum += S^T wgt.  wgt, out, wsum may be NULL.  Device pointers. */
lfrv_status lfrv_op_splat_atomic(int W, int H, int w, int h, int dsf, const float *tau_x,
                                 const float *tau_y, const float *in, int C, const float *wgt,
                                 float *out, float *wsum, void *stream);

/* Sobel/8 of the luminance of a planar RGB image (gradientIBR_kernels.cuh:57-99);
 * vis (uint8, nullable) enables the 3x3 gate of vtv_sr_kernels_unstructured.cu:120-126
 * (gated pixels are left unwritten).  Device pointers. */
lfrv_status lfrv_op_sobel(int W, int H, const float *rgb, float *gx, float *gy,
                          const uint8_t *vis, void *stream);

/* The OpenGL path's depth-map filter: radial depth (distance along the
 * pixel's ray, as MVE stores it) to orthographic depth, with its x / y
 * derivatives, in place (fromRadial2Ortho.frag, applied by
 * Scene::filterDepthMaps, scene.cpp:1605-1624, view.cpp:723-733).  K is the
 * view's row-major 3x3 intrinsic matrix; dzdx, dzdy may be NULL.  Device
 * pointers, W*H floats each.  LFRV_E_ARG for a singular K. */
lfrv_status lfrv_op_radial_to_ortho(int W, int H, const double *K, float *z, float *dzdx, float *dzdy,
                                    void *stream);

/* ---- device memory helpers ------------------------------------------------- */
/* Plumbing for callers that hold no device allocator of their own (no
 * counterpart in the reference, whose buffers are private to Data /
 * coco_vtv_workspace): allocate, free and copy device memory on the
 * library's HIP runtime, so the stateless operators above can be fed without
 * a second runtime in the process. */
lfrv_status lfrv_dev_alloc(size_t bytes, void **ptr);
lfrv_status lfrv_dev_free(void *ptr);
/* Synchronous copy between any host / device pointers. */
lfrv_status lfrv_memcpy(void *dst, const void *src, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (16-byte vector loads and
 * stores; bytes and both pointers 16-byte aligned).  Its memory-counter
 * readings calibrate the HBM-traffic figures of the other kernels. */
lfrv_status lfrv_dev_copy_kernel(void *dst, const void *src, size_t bytes, void *stream);
/* Wait for the default stream (the stateless operators' NULL stream). */
lfrv_status lfrv_device_sync(void);

/* ---- quality metrics ------------------------------------------------------- */

/* PSNR, vmax = 1, 2-px border excluded, RGB-averaged MSE; DBL_MAX for equal
 * images (gsl_image_psnr, gsl_image.cpp:991-1025, psnr_main.cpp).  Host
 * pointers, planar RGB. */
double lfrv_psnr(int W, int H, const float *a, const float *b);

/* SSIM as the reference computes it (gsl_image_ssim, gsl_image.cpp:1027-1037,
 * gsl_matrix_ssim, gsl_matrix_helper.cpp:799-914): the mean over R, G, B of
 * the mean over all pixels of ((2 mu1 mu2 + C1) + (2 s12 + C2)) /
 * ((mu1^2 + mu2^2 + C1) + (s1^2 + s2^2 + C2)), local statistics by an 11-tap
 * Gaussian of sigma 1.5 renormalised at the borders, C1 = (0.01 range)^2,
 * C2 = (0.03 range)^2; the factors are added, not multiplied, as in the
 * reference.  Host pointers, planar RGB, FP64. */
double lfrv_ssim(int W, int H, const float *a, const float *b, double dynamic_range);

/* PSNR of gsl_image_psnr (as lfrv_psnr) of two device-resident planar RGB
 * images, written as a double to the device pointer `out` on `stream`
 * (FP64 sums, one partial per 4096 interior pixels, summed in a fixed
 * order).  Its scratch is stream-ordered memory it allocates and frees on
 * that stream.  W, H > 4; W*H <= 2^31.  No counterpart kernel in the
 * reference (its tool reads two image files, psnr_main.cpp:12-20). */
lfrv_status lfrv_op_psnr(int W, int H, const float *a, const float *b, double *out, void *stream);

/* lfrv_op_psnr with its partial sums in a fixed device array (128 KiB per
 * GPU, held until the process ends) instead of memory it allocates: no
 * counterpart in the reference; for callers that cannot allocate (code under
 * HIP graph capture, serving loops that forbid hipMalloc).  The same kernels
 * and result as lfrv_op_psnr.  Images whose interior (W-4)(H-4) exceeds
 * 16384 * 4096 pixels are refused (LFRV_E_ARG).  Calls that overlap on one
 * GPU (two streams or two host threads) share the array and race. */
lfrv_status lfrv_op_psnr_noalloc(int W, int H, const float *a, const float *b, double *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* LFRV_H */
