// The seeded synthetic scene of include/lfrv_synth.h as plain data, and the
// ray tracer that renders it, shared by the host renderer (synth.cpp) and the
// device renderer (k_synth.hip): textured planes seen by pinhole cameras with
// R = I.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/lfrv_synth.h"

namespace lfrv_synth {

constexpr int kSin = 32;        // sinusoids per texture
constexpr int kRect = 16;       // rectangles per texture
constexpr int kMaxPlanes = 16;
constexpr int kMaxViews = 64;   // grid <= 8

struct Texture {
    double base[3];
    double amp[kSin][3], kx[kSin], ky[kSin], ph[kSin];
    double rc[kRect][2], rh[kRect][2], col[kRect][3];

    __host__ __device__ void eval(double u, double v, float *rgb) const
    {
        double c[3] = {base[0], base[1], base[2]};
        for (int k = 0; k < kSin; k++) {
            const double s = sin(kx[k] * u + ky[k] * v + ph[k]);
            c[0] += amp[k][0] * s;
            c[1] += amp[k][1] * s;
            c[2] += amp[k][2] * s;
        }
        for (int k = 0; k < kRect; k++) {
            if (fabs(u - rc[k][0]) <= rh[k][0] && fabs(v - rc[k][1]) <= rh[k][1]) {
                for (int q = 0; q < 3; q++) c[q] = 0.5 * c[q] + 0.5 * col[k][q];
            }
        }
        for (int q = 0; q < 3; q++) rgb[q] = (float)(c[q] < 0.0 ? 0.0 : (c[q] > 1.0 ? 1.0 : c[q]));
    }
};

struct Plane {
    double n[3], d;  // n . X = d
    double o[3], e1[3], e2[3];
    double h1, h2;
    int bounded;
    Texture tex;
};

struct Cam {
    double f, cx, cy;  // K = [[f,0,cx],[0,f,cy],[0,0,1]], R = I
    double C[3];
};

struct Scene {
    int nplanes;
    Plane planes[kMaxPlanes];
    int V;
    Cam cams[kMaxViews];
    Cam target;
};

// Nearest surface along the ray through image point (x, y) (corner-origin
// pixel coordinates).  Returns false when nothing is hit (depth 0, black).
__host__ __device__ inline bool trace(const Scene &sc, const Cam &cam, double x, double y, double *depth, float *rgb)
{
    const double r[3] = {(x - cam.cx) / cam.f, (y - cam.cy) / cam.f, 1.0};
    double best = 1e30;
    int bi = -1;
    double bu = 0, bv = 0;
    for (int i = 0; i < sc.nplanes; i++) {
        const Plane &p = sc.planes[i];
        const double nr = p.n[0] * r[0] + p.n[1] * r[1] + p.n[2] * r[2];
        if (fabs(nr) < 1e-12) continue;
        const double t = (p.d - (p.n[0] * cam.C[0] + p.n[1] * cam.C[1] + p.n[2] * cam.C[2])) / nr;
        if (!(t > 1e-6) || t >= best) continue;
        const double X[3] = {cam.C[0] + t * r[0] - p.o[0], cam.C[1] + t * r[1] - p.o[1],
                             cam.C[2] + t * r[2] - p.o[2]};
        const double u = X[0] * p.e1[0] + X[1] * p.e1[1] + X[2] * p.e1[2];
        const double v = X[0] * p.e2[0] + X[1] * p.e2[1] + X[2] * p.e2[2];
        if (p.bounded && (fabs(u) > p.h1 || fabs(v) > p.h2)) continue;
        best = t;  // r_z = 1 and R = I, so t is the z-depth
        bi = i;
        bu = u;
        bv = v;
    }
    if (bi < 0) {
        *depth = 0.0;
        if (rgb) rgb[0] = rgb[1] = rgb[2] = 0.0f;
        return false;
    }
    *depth = best;
    if (rgb) sc.planes[bi].tex.eval(bu, bv, rgb);
    return true;
}

// The planes, textures and cameras of a description: the seeded draws in a
// fixed order (synth.cpp).  P_views (V*12) and P_target (12) are nullable.
// Returns 0, or 1 for a description out of range.
int build_scene(const lfrv_synth_desc *d, Scene *sc, double *P_views, double *P_target);

}  // namespace lfrv_synth
