// Seeded synthetic light fields (host code).  See include/lfrv_synth.h.
#include "../../include/lfrv_synth.h"
#include "synth_scene.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

using namespace lfrv_synth;

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x1234567ull) {}
    uint64_t next()
    {  // splitmix64
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uni(double a, double b) { return a + (b - a) * uni(); }
    double normal()
    {
        double u1 = uni(), u2 = uni();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

void make_texture(Texture &t, Rng &rng, double world_per_px, double pmin, double pmax, double extent)
{
    for (int q = 0; q < 3; q++) t.base[q] = rng.uni(0.35, 0.65);
    for (int k = 0; k < kSin; k++) {
        const double period_px = pmin * std::exp(rng.uni() * std::log(pmax / pmin));
        const double kk = 2.0 * M_PI / (period_px * world_per_px);
        const double th = rng.uni(0.0, M_PI);
        t.kx[k] = kk * std::cos(th);
        t.ky[k] = kk * std::sin(th);
        t.ph[k] = rng.uni(0.0, 2.0 * M_PI);
        for (int q = 0; q < 3; q++) t.amp[k][q] = rng.uni(0.01, 0.05);
    }
    for (int k = 0; k < kRect; k++) {
        t.rc[k][0] = rng.uni(-extent, extent);
        t.rc[k][1] = rng.uni(-extent, extent);
        t.rh[k][0] = rng.uni(6.0, 60.0) * world_per_px;
        t.rh[k][1] = rng.uni(6.0, 60.0) * world_per_px;
        for (int q = 0; q < 3; q++) t.col[k][q] = rng.uni(0.0, 1.0);
    }
}

void normalize3(double *v)
{
    const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= l;
    v[1] /= l;
    v[2] /= l;
}

void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

template <class F>
void parallel_rows(int rows, F fn)
{
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    if ((int)nt > rows) nt = (unsigned)std::max(rows, 1);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t]() {
            for (int r = (int)t; r < rows; r += (int)nt) fn(r);
        });
    for (auto &x : th) x.join();
}

void cam_P(const Cam &c, double *P)
{
    // P = K [I | -C]
    const double K[9] = {c.f, 0, c.cx, 0, c.f, c.cy, 0, 0, 1};
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) P[4 * i + j] = K[3 * i + j];
        P[4 * i + 3] = -(K[3 * i + 0] * c.C[0] + K[3 * i + 1] * c.C[1] + K[3 * i + 2] * c.C[2]);
    }
}

}  // namespace

extern "C" void lfrv_synth_preset(lfrv_synth_desc *d, int config, uint64_t seed)
{
    std::memset(d, 0, sizeof(*d));
    d->seed = seed;
    d->tex_min_period = 4.0;
    d->tex_max_period = 64.0;
    d->nplanes = 8;
    d->zmin = 2.0;
    d->zmax = 10.0;
    switch (config) {
    case 1:  // 3x3 views 128^2 from 256^2 renders, one plane at 0.7 px/step in the 256^2 frame
        d->kind = 1;
        d->grid = 3;
        d->w = d->h = 128;
        d->dsf = 2;
        d->render_scale = 2;
        d->focal = 128.0;
        d->plane_depth = 4.0;
        d->baseline = 0.7 * 4.0 / 256.0;
        d->noise_sigma = 0.01;
        d->random_target = 0;
        break;
    case 3:  // 5x5 views 512^2 from 2048^2 renders, disparity +-2 px/step in the 512^2 frame
        d->kind = 2;
        d->grid = 5;
        d->w = d->h = 512;
        d->dsf = 4;
        d->render_scale = 4;
        d->focal = 512.0;
        d->baseline = 10.0 / 512.0;  // f*b = 10: disparity f*b/Z - 3 spans [-2, 2] on Z in [2, 10]
        d->pp_shift = 3.0;
        d->random_target = 0;
        break;
    case 2:  // 3x3 views 1024^2, f = 1000, baseline 0.05, random interior target
    default:
        d->kind = 2;
        d->grid = 3;
        d->w = d->h = (config == 2) ? 1024 : 64;
        d->dsf = 1;
        d->render_scale = 1;
        d->focal = 1000.0 * d->w / 1024.0;
        d->baseline = 0.05;
        d->random_target = 1;
        break;
    }
}

extern "C" int lfrv_synth_nviews(const lfrv_synth_desc *d) { return d->grid * d->grid; }

namespace lfrv_synth {

int build_scene(const lfrv_synth_desc *d, Scene *sc, double *P_views, double *P_target)
{
    if (!d || d->grid < 1 || d->grid * d->grid > kMaxViews || d->w < 1 || d->h < 1 || d->dsf < 1 ||
        d->render_scale < 1 || d->focal <= 0 || d->nplanes > kMaxPlanes)
        return 1;
    Rng rng(d->seed);
    const int g = d->grid, V = g * g, w = d->w, h = d->h;
    const int W = w * d->dsf, H = h * d->dsf;
    const double ft = d->focal * d->dsf;  // target focal, target pixels
    const double half = 0.5 * (g - 1);

    // target position
    double ta = 0.0, tb = 0.0;
    if (d->random_target && g > 1) {
        ta = rng.uni(-half, half);
        tb = rng.uni(-half, half);
    }

    // scene
    sc->nplanes = 0;
    if (d->kind == 1) {
        Plane &p = sc->planes[sc->nplanes++];
        p = Plane{};
        p.n[0] = 0; p.n[1] = 0; p.n[2] = 1;
        p.d = d->plane_depth;
        p.o[0] = 0; p.o[1] = 0; p.o[2] = d->plane_depth;
        p.e1[0] = 1; p.e1[1] = 0; p.e1[2] = 0;
        p.e2[0] = 0; p.e2[1] = 1; p.e2[2] = 0;
        p.bounded = 0;
        const double wpp = d->plane_depth / ft;
        make_texture(p.tex, rng, wpp, d->tex_min_period, d->tex_max_period, 0.5 * W * wpp);
    } else {
        const int np = std::max(1, d->nplanes);
        for (int k = 0; k < np; k++) {
            Plane &p = sc->planes[sc->nplanes++];
            p = Plane{};
            double z0, tilt, az;
            double ix, iy;
            if (k == 0) {
                z0 = d->zmax * 0.97;
                tilt = rng.uni(0.0, 10.0) * M_PI / 180.0;
                ix = 0.5;
                iy = 0.5;
            } else {
                z0 = rng.uni(d->zmin + 0.25 * (d->zmax - d->zmin) * 0.2, d->zmin + 0.8 * (d->zmax - d->zmin));
                tilt = rng.uni(0.0, 35.0) * M_PI / 180.0;
                ix = rng.uni(0.15, 0.85);
                iy = rng.uni(0.15, 0.85);
            }
            az = rng.uni(0.0, 2.0 * M_PI);
            p.n[0] = std::sin(tilt) * std::cos(az);
            p.n[1] = std::sin(tilt) * std::sin(az);
            p.n[2] = std::cos(tilt);
            normalize3(p.n);
            // centre point: on the ray of the target-grid-centre camera through (ix, iy)
            p.o[0] = (ix * W - 0.5 * W) * z0 / ft;
            p.o[1] = (iy * H - 0.5 * H) * z0 / ft;
            p.o[2] = z0;
            p.d = p.n[0] * p.o[0] + p.n[1] * p.o[1] + p.n[2] * p.o[2];
            const double up[3] = {0.0, 1.0, 0.0};
            cross3(up, p.n, p.e1);
            normalize3(p.e1);
            cross3(p.n, p.e1, p.e2);
            normalize3(p.e2);
            const double wpp = z0 / ft;
            if (k == 0) {
                p.bounded = 0;
                p.h1 = p.h2 = 0.0;
            } else {
                p.bounded = 1;
                p.h1 = rng.uni(0.08, 0.25) * W * wpp;
                p.h2 = rng.uni(0.08, 0.25) * H * wpp;
            }
            make_texture(p.tex, rng, wpp, d->tex_min_period, d->tex_max_period,
                         p.bounded ? std::max(p.h1, p.h2) : 0.6 * W * wpp);
        }
    }

    // cameras (low-res views, target at W x H)
    sc->V = V;
    for (int t = 0; t < g; t++)
        for (int s = 0; s < g; s++) {
            const double ss = s - half, tt = t - half;
            Cam &c = sc->cams[t * g + s];
            c.f = d->focal;
            c.cx = 0.5 * w + ss * d->pp_shift;
            c.cy = 0.5 * h + tt * d->pp_shift;
            c.C[0] = d->baseline * ss;
            c.C[1] = d->baseline * tt;
            c.C[2] = 0.0;
            if (P_views) cam_P(c, P_views + 12 * (t * g + s));
        }
    Cam &tc = sc->target;
    tc.f = ft;
    tc.cx = (0.5 * w + ta * d->pp_shift) * d->dsf;
    tc.cy = (0.5 * h + tb * d->pp_shift) * d->dsf;
    tc.C[0] = d->baseline * ta;
    tc.C[1] = d->baseline * tb;
    tc.C[2] = 0.0;
    if (P_target) cam_P(tc, P_target);
    return 0;
}

}  // namespace lfrv_synth

using namespace lfrv_synth;

extern "C" int lfrv_synth_generate(const lfrv_synth_desc *d, float *views, float *depths,
                                   double *P_views, double *P_target, float *target_rgb,
                                   float *target_depth)
{
    std::vector<Scene> hold(1);
    Scene &sc = hold[0];
    if (lfrv_synth::build_scene(d, &sc, P_views, P_target) != 0) return 1;
    const int V = sc.V, w = d->w, h = d->h, rs = d->render_scale;
    const int W = w * d->dsf, H = h * d->dsf;
    const size_t n = (size_t)w * h;
    for (int v = 0; v < V; v++) {
        const Cam &c = sc.cams[v];
        float *img = views ? views + (size_t)v * 3 * n : nullptr;
        float *dep = depths ? depths + (size_t)v * n : nullptr;
        parallel_rows(h, [&](int y) {
            for (int x = 0; x < w; x++) {
                const size_t m = (size_t)y * w + x;
                double dd;
                if (dep) {
                    trace(sc, c, x + 0.5, y + 0.5, &dd, nullptr);
                    dep[m] = (float)dd;
                }
                if (img) {
                    double acc[3] = {0, 0, 0};
                    for (int sy = 0; sy < rs; sy++)
                        for (int sx = 0; sx < rs; sx++) {
                            float rgb[3];
                            trace(sc, c, x + (sx + 0.5) / rs, y + (sy + 0.5) / rs, &dd, rgb);
                            acc[0] += rgb[0];
                            acc[1] += rgb[1];
                            acc[2] += rgb[2];
                        }
                    for (int q = 0; q < 3; q++) img[q * n + m] = (float)(acc[q] / (rs * rs));
                }
            }
        });
        if (img && d->noise_sigma > 0.0) {
            Rng nr(d->seed * 1315423911ull + 977ull * (uint64_t)(v + 1));
            for (size_t k = 0; k < 3 * n; k++) {
                const double val = img[k] + d->noise_sigma * nr.normal();
                img[k] = (float)std::min(1.0, std::max(0.0, val));
            }
        }
    }
    if (target_rgb || target_depth) {
        const size_t N = (size_t)W * H;
        parallel_rows(H, [&](int y) {
            for (int x = 0; x < W; x++) {
                const size_t o = (size_t)y * W + x;
                double dd;
                float rgb[3];
                trace(sc, sc.target, x + 0.5, y + 0.5, &dd, rgb);
                if (target_rgb)
                    for (int q = 0; q < 3; q++) target_rgb[q * N + o] = rgb[q];
                if (target_depth) target_depth[o] = (float)dd;
            }
        });
    }
    return 0;
}
