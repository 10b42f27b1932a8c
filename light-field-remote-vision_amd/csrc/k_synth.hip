// The seeded synthetic light fields of include/lfrv_synth.h rendered on the
// GPU: the scene is drawn on the host exactly as lfrv_synth_generate draws it
// (build_scene), then every view pixel (with its render_scale^2 supersamples)
// and every target pixel is one thread running the same FP64 ray tracer.
// For the training sets of the warm start and the batched workloads, whose
// thousands of scenes would take hours on the host.
#include <vector>

#include "../../include/lfrv_synth.h"
#include "lfrv_internal.hpp"
#include "synth_scene.hpp"

namespace lfrv {

using lfrv_synth::Scene;

__global__ void __launch_bounds__(256)
k_synth_views(const Scene *__restrict__ sc, int V, int w, int h, int rs, float *__restrict__ views,
              float *__restrict__ depths)
{
    const size_t n = (size_t)w * h;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)V * n) return;
    const int v = (int)(i / n);
    const size_t m = i - (size_t)v * n;
    const int y = (int)(m / w), x = (int)(m % w);
    const lfrv_synth::Cam &c = sc->cams[v];
    double dd;
    if (depths) {
        lfrv_synth::trace(*sc, c, x + 0.5, y + 0.5, &dd, nullptr);
        depths[i] = (float)dd;
    }
    if (views) {
        double acc[3] = {0, 0, 0};
        for (int sy = 0; sy < rs; sy++)
            for (int sx = 0; sx < rs; sx++) {
                float rgb[3];
                lfrv_synth::trace(*sc, c, x + (sx + 0.5) / rs, y + (sy + 0.5) / rs, &dd, rgb);
                acc[0] += rgb[0];
                acc[1] += rgb[1];
                acc[2] += rgb[2];
            }
        float *img = views + (size_t)v * 3 * n;
        for (int q = 0; q < 3; q++) img[q * n + m] = (float)(acc[q] / (rs * rs));
    }
}

__global__ void __launch_bounds__(256)
k_synth_target(const Scene *__restrict__ sc, int W, int H, float *__restrict__ rgb_out, float *__restrict__ depth_out)
{
    const size_t N = (size_t)W * H;
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= N) return;
    const int y = (int)(o / W), x = (int)(o % W);
    double dd;
    float rgb[3];
    lfrv_synth::trace(*sc, sc->target, x + 0.5, y + 0.5, &dd, rgb);
    if (rgb_out)
        for (int q = 0; q < 3; q++) rgb_out[q * N + o] = rgb[q];
    if (depth_out) depth_out[o] = (float)dd;
}

}  // namespace lfrv

using namespace lfrv;

extern "C" int lfrv_synth_generate_device(const lfrv_synth_desc *d, float *views, float *depths, double *P_views,
                                          double *P_target, float *target_rgb, float *target_depth, void *stream)
{
    if (!d || d->noise_sigma > 0.0) return 1;
    std::vector<Scene> hold(1);
    if (lfrv_synth::build_scene(d, &hold[0], P_views, P_target) != 0) return 1;
    const hipStream_t s = (hipStream_t)stream;
    Scene *dsc = nullptr;
    if (hipMallocAsync((void **)&dsc, sizeof(Scene), s) != hipSuccess) return 2;
    int rc = 0;
    if (hipMemcpyAsync(dsc, &hold[0], sizeof(Scene), hipMemcpyHostToDevice, s) != hipSuccess) rc = 2;
    const int V = hold[0].V;
    const size_t n = (size_t)d->w * d->h, N = n * d->dsf * d->dsf;
    if (!rc && (views || depths)) {
        k_synth_views<<<(unsigned int)((V * n + 255) / 256), 256, 0, s>>>(dsc, V, d->w, d->h, d->render_scale, views,
                                                                           depths);
        if (hipGetLastError() != hipSuccess) rc = 2;
    }
    if (!rc && (target_rgb || target_depth)) {
        k_synth_target<<<(unsigned int)((N + 255) / 256), 256, 0, s>>>(dsc, d->w * d->dsf, d->h * d->dsf, target_rgb,
                                                                        target_depth);
        if (hipGetLastError() != hipSuccess) rc = 2;
    }
    // the host copy of the scene must outlive the upload
    if (hipStreamSynchronize(s) != hipSuccess) rc = 2;
    if (hipFreeAsync(dsc, s) != hipSuccess) rc = 2;
    return rc;
}
