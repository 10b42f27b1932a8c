/*
 * lfrv_synth.h -- seeded synthetic light fields for tests and benchmarks.
 *
 * These entry points have no counterpart in the reference (which reads PNG /
 * PFM light fields produced by Blender or its OpenGL renderer); they produce
 * the synthetic inputs the benchmark configurations are defined on:
 *
 *   scene kind 1: one fronto-parallel textured plane (light-field
 *                 super-resolution: 3x3 views, box-downsampled, noisy);
 *   scene kind 2: 8 random tilted textured planes at depth [zmin, zmax]
 *                 (plane 0 unbounded background, planes 1..7 bounded
 *                 rectangles), views z-buffered analytically.
 *
 * Textures are 32 random sinusoids plus 16 random rectangles per plane in
 * plane coordinates.  Cameras are pinhole, R = I, centres on a regular grid
 * (baseline per step); an optional principal-point shift per grid step sets
 * a zero-disparity plane.  All output is host memory, planar RGB.
 */
#ifndef LFRV_SYNTH_H
#define LFRV_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int kind;             /* 1 fronto-parallel plane, 2 tilted planes */
    uint64_t seed;
    int grid;             /* views per side; V = grid*grid */
    int w, h;             /* input view size */
    int dsf;              /* target = (w*dsf) x (h*dsf) */
    int render_scale;     /* views rendered at (w,h)*render_scale then box-averaged */
    double focal;         /* focal length in input-view pixels */
    double baseline;      /* camera spacing per grid step */
    double pp_shift;      /* principal-point shift per grid step, input pixels */
    double plane_depth;   /* kind 1: plane depth */
    int nplanes;          /* kind 2: number of planes */
    double zmin, zmax;    /* kind 2: depth range */
    int random_target;    /* 1: target centre uniform in the grid's interior; 0: grid centre */
    double noise_sigma;   /* Gaussian noise on the views, then clamped to [0,1] */
    double tex_min_period, tex_max_period; /* sinusoid periods, target pixels at plane depth */
} lfrv_synth_desc;

/* Presets: config 1 (3x3, 128^2 -> 256^2, one plane, disparity 0.7 px/step
 * in the 256^2 frame, noise 0.01), 2 (3x3, 1024^2, 8 planes at depth 2-10,
 * f = 1000, baseline 0.05, random interior target), 3 (5x5, 512^2 -> 2048^2,
 * 8 planes, disparity +-2 px/step in the 512^2 frame).  Any other value of
 * `config` gives config 2's scene at w = h = 64 (a small CPU-checkable case). */
void lfrv_synth_preset(lfrv_synth_desc *d, int config, uint64_t seed);

/* Number of views of a description (grid*grid). */
int lfrv_synth_nviews(const lfrv_synth_desc *d);

/* Render: views (V*3*w*h), depths (V*w*h, z-depth at input-pixel centres),
 * P_views (V*12, row-major 3x4 for the w x h images), P_target (12, for the
 * W x H target), target_rgb (3*W*H ground truth rendered at the target
 * camera; nullable), target_depth (W*H; nullable).  Returns 0 on success. */
int lfrv_synth_generate(const lfrv_synth_desc *d, float *views, float *depths, double *P_views,
                        double *P_target, float *target_rgb, float *target_depth);

/* The same scene rendered on the GPU: views, depths, target_rgb and
 * target_depth are device pointers (each nullable), P_views / P_target host
 * pointers as above; work is queued on `stream` and finished when the call
 * returns.  The ray tracer is the host one in FP64, so the images agree with
 * lfrv_synth_generate's to float rounding.  Noise is not supported (returns 1
 * when noise_sigma > 0).  Returns 0, 1 for a bad description, 2 for a HIP
 * error. */
int lfrv_synth_generate_device(const lfrv_synth_desc *d, float *views, float *depths, double *P_views,
                               double *P_target, float *target_rgb, float *target_depth, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* LFRV_SYNTH_H */
