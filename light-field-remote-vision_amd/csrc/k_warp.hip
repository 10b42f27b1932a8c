// Warps from depth (the import path and the OpenGL path), gather (warp) and
// per-view weight kernels.
#include "lfrv_internal.hpp"

#include <cstring>

namespace lfrv {

namespace {

// Per-source-pixel kernels run on grid (x tiles, y tiles, views) with 2-D
// blocks: pixel (mx, my) of view v, m = my * w + mx, k = v * n + m.
#define SOURCE_INDEX()                                                      \
    const int mx = blockIdx.x * blockDim.x + threadIdx.x;                   \
    const int my = blockIdx.y * blockDim.y + threadIdx.y;                   \
    const int v = blockIdx.z;                                               \
    if (mx >= w || my >= h) return;                                         \
    const size_t n = (size_t)w * h;                                         \
    const int m = my * w + mx;                                              \
    const size_t k = (size_t)v * n + (size_t)m;                             \
    (void)k;

constexpr float kInvalidDepth = 1000.0f;  // import.cpp:23

// Order-preserving map float -> int for atomicMin on signed floats.
__device__ __forceinline__ int f2key(float f)
{
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float key2f(int k) { return __int_as_float(k >= 0 ? k : k ^ 0x7FFFFFFF); }

// Pass 1 of import_warps_from_depth (import.cpp:288-339) for every view:
// tau = pi(d M x + c), partial tau (tauPartial.frag:38-47, import.cpp:658-680)
// and the target z-buffer over a dsf x dsf footprint.  The block is 64 x 4:
// one wave per row of 64 source pixels.  For dsf >= 2 the wave writes its
// footprints row by row, dsf consecutive lanes on one source pixel's dsf
// consecutive target pixels (64 / dsf source pixels per instruction), so an
// atomic instruction touches 2-3 cache lines instead of one per source
// pixel: device-scope atomics are paid per cache line.  The set of
// (pixel, key) minima is the same in either order.
__global__ void k_tau_from_depth(int V, int w, int h, int dsf, const float *__restrict__ depth,
                                 const WarpCam *__restrict__ cams, float sigma_z,
                                 float *__restrict__ tx, float *__restrict__ ty,
                                 float *__restrict__ dpx, float *__restrict__ dpy,
                                 int *__restrict__ zkey)
{
    const int mx = blockIdx.x * blockDim.x + threadIdx.x, my = blockIdx.y * blockDim.y + threadIdx.y;
    const int v = blockIdx.z;
    const int W = w * dsf, H = h * dsf;
    float tux = 0.0f, tuy = 0.0f;
    int key = 0, live = 0;
    if (mx < w && my < h) {
        const size_t k = (size_t)v * w * h + (size_t)my * w + mx;
        const WarpCam &cam = cams[v];
        const float d = depth[k];
        if (d < 1.0f / kInvalidDepth || kInvalidDepth < d) {
            tx[k] = -(float)w;
            ty[k] = -(float)h;
            if (dpx) {
                dpx[k] = 0.0f;
                dpy[k] = 0.0f;
            }
        } else {
            const float xx = (float)mx + 0.5f, yy = (float)my + 0.5f;
            const float Mx0 = cam.M[0] * xx + cam.M[1] * yy + cam.M[2];
            const float Mx1 = cam.M[3] * xx + cam.M[4] * yy + cam.M[5];
            const float Mx2 = cam.M[6] * xx + cam.M[7] * yy + cam.M[8];
            const float X0 = d * Mx0 + cam.c[0], X1 = d * Mx1 + cam.c[1], X2 = d * Mx2 + cam.c[2];
            tux = X0 / X2;
            tuy = X1 / X2;
            tx[k] = tux;
            ty[k] = tuy;
            if (dpx) {
                dpx[k] = sigma_z * (Mx0 - tux * Mx2) / X2;
                dpy[k] = sigma_z * (Mx1 - tuy * Mx2) / X2;
            }
            key = f2key(X2);
            live = 1;
        }
    }
    if (dsf == 1) {
        const int px = (int)tux, py = (int)tuy;
        if (live && 0 <= px && px <= W - 1 && 0 <= py && py <= H - 1) atomicMin(zkey + (size_t)py * W + px, key);
        return;
    }
    // every lane of the wave takes part in the shuffles below
    const int lane = threadIdx.x;  // blockDim.x == 64
    const int per = 64 / dsf, s_off = lane / dsf, j = lane - s_off * dsf;
    for (int g0 = 0; g0 < 64; g0 += per) {
        const int src = g0 + s_off;
        const bool act = s_off < per && src < 64;
        const int sl = act ? src : lane;
        const float gx = __shfl(tux, sl, 64), gy = __shfl(tuy, sl, 64);
        const int gk = __shfl(key, sl, 64), gl = __shfl(live, sl, 64);
        if (!act || !gl) continue;
        const int px = (int)gx - dsf / 2 + j;
        if (px < 0 || px > W - 1) continue;
        for (int i = 0; i < dsf; i++) {
            const int py = (int)gy - dsf / 2 + i;
            if (0 <= py && py <= H - 1) atomicMin(zkey + (size_t)py * W + px, gk);
        }
    }
}

// sigma_z dtau/dz of the import warps (the dpart k_tau_from_depth leaves out
// when no reader needs them yet): the same arithmetic on the same depth and
// cameras, so the values are those it would have written.
__global__ void k_dpart_from_depth(int V, int w, int h, const float *__restrict__ depth,
                                   const WarpCam *__restrict__ cams, float sigma_z, float *__restrict__ dpx,
                                   float *__restrict__ dpy)
{
    const int mx = blockIdx.x * blockDim.x + threadIdx.x, my = blockIdx.y * blockDim.y + threadIdx.y;
    const int v = blockIdx.z;
    if (mx >= w || my >= h) return;
    const size_t k = (size_t)v * w * h + (size_t)my * w + mx;
    const WarpCam &cam = cams[v];
    const float d = depth[k];
    if (d < 1.0f / kInvalidDepth || kInvalidDepth < d) {
        dpx[k] = 0.0f;
        dpy[k] = 0.0f;
        return;
    }
    const float xx = (float)mx + 0.5f, yy = (float)my + 0.5f;
    const float Mx0 = cam.M[0] * xx + cam.M[1] * yy + cam.M[2];
    const float Mx1 = cam.M[3] * xx + cam.M[4] * yy + cam.M[5];
    const float Mx2 = cam.M[6] * xx + cam.M[7] * yy + cam.M[8];
    const float X0 = d * Mx0 + cam.c[0], X1 = d * Mx1 + cam.c[1], X2 = d * Mx2 + cam.c[2];
    const float tux = X0 / X2, tuy = X1 / X2;
    dpx[k] = sigma_z * (Mx0 - tux * Mx2) / X2;
    dpy[k] = sigma_z * (Mx1 - tuy * Mx2) / X2;
}

// Fills the z-buffer keys; also clears the ncnt cell counters.
__global__ void k_fill_key(int *__restrict__ zkey, size_t N, int key, int *__restrict__ cnt, size_t ncnt)
{
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o < N) zkey[o] = key;
    if (o < ncnt) cnt[o] = 0;
}

__global__ void k_key_to_depth(const int *__restrict__ zkey, float *__restrict__ depth, size_t N)
{
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o < N) depth[o] = key2f(zkey[o]);
}

// Pass 2 (import.cpp:351-442): footprint-averaged target depth at tau, round
// trip of the target pixel int(tau)+.5 into view v; a miss by more than tol
// pixels marks the source pixel occluded.
template <int KS>
__global__ void k_tau_visibility(int V, int w, int h, const float *__restrict__ ud,
                                 const WarpCam *__restrict__ cams, float tol,
                                 float *__restrict__ tx, float *__restrict__ ty, int *__restrict__ cell_count)
{
    SOURCE_INDEX()
    constexpr int dsf = KS - 1;
    const int W = w * dsf, H = h * dsf;
    const WarpCam &cam = cams[v];
    const float tux = tx[k], tuy = ty[k];
    float u_d = 0.0f, u_norm = 0.0f;
    const float hd = 0.5f * (float)dsf;
    const int bx = (int)floorf(tux - hd + 0.5f), by = (int)floorf(tuy - hd + 0.5f);
    // the footprint's depths are loaded unconditionally (clamped addresses,
    // masked below) so that they are in flight together
    float zt[KS * KS];
#pragma unroll
    for (int i = 0; i < KS; i++)
#pragma unroll
        for (int j = 0; j < KS; j++) {
            const int px = min(max(bx + j, 0), W - 1), py = min(max(by + i, 0), H - 1);
            zt[i * KS + j] = ud[(size_t)py * W + px];
        }
#pragma unroll
    for (int i = 0; i < KS; i++)
#pragma unroll
        for (int j = 0; j < KS; j++) {
            const int px = bx + j, py = by + i;
            if (0 <= px && px <= W - 1 && 0 <= py && py <= H - 1) {
                const float z = zt[i * KS + j];
                if (z == kInvalidDepth) continue;
                float dx = 1.0f, dy = 1.0f;
                if (j == 0) dx = ((float)px + 0.5f) - (tux - hd);
                else if (j == dsf) dx = (tux + hd) - ((float)px - 0.5f);
                if (i == 0) dy = ((float)py + 0.5f) - (tuy - hd);
                else if (i == dsf) dy = (tuy + hd) - ((float)py - 0.5f);
                u_d += dx * dy * z;
                u_norm += dx * dy;
            }
        }
    if (u_norm == 0.0f) {
        tx[k] = -(float)w;
        ty[k] = -(float)h;
        return;
    }
    u_d /= u_norm;
    const float xx = (float)(int)tux + 0.5f, yy = (float)(int)tuy + 0.5f;
    const float X0 = u_d * (cam.Nb[0] * xx + cam.Nb[1] * yy + cam.Nb[2]) + cam.e[0];
    const float X1 = u_d * (cam.Nb[3] * xx + cam.Nb[4] * yy + cam.Nb[5]) + cam.e[1];
    const float X2 = u_d * (cam.Nb[6] * xx + cam.Nb[7] * yy + cam.Nb[8]) + cam.e[2];
    const int ox = (int)(X0 / X2), oy = (int)(X1 / X2);
    const float ddx = (float)(ox - mx), ddy = (float)(oy - my);
    if (ddx < -tol || tol < ddx || ddy < -tol || tol < ddy) {
        tx[k] = -(float)w * 2.0f;
        ty[k] = -(float)h * 2.0f;
        return;
    }
    // the final warp is known here: count it into the splat structure's cells
    if (cell_valid(tux, tuy, W, H)) atomicAdd(cell_count + cell_of<KS>(tux, tuy, W + KS - 1), 1);
}

// ---- the OpenGL warp producer (Scene::computeTargetDepthMap + exportWarps) ----

constexpr float kGlZNear = 1e-2f;  // near plane of the render camera, pinholeCamera.cpp:133
constexpr int kGlMaxSpan = 64;     // quads wider or taller than this go to the overflow list (k_gl_splat_big)

__device__ __forceinline__ float gl_edge(const float *a, const float *b, float px, float py)
{
    return (b[0] - a[0]) * (py - a[1]) - (b[1] - a[1]) * (px - a[0]);
}

// A pixel centre on an edge belongs to the triangle owning the edge; of the two
// directions a shared edge is walked in, exactly one owns it.
__device__ __forceinline__ bool gl_owns(const float *a, const float *b)
{
    const float dx = b[0] - a[0], dy = b[1] - a[1];
    return dy > 0.0f || (dy == 0.0f && dx < 0.0f);
}

// Triangle (a, b, c) = (x, y, depth) at pixel centre p: inside, and the depth
// as the depth test interpolates it (1/z linear in the image).
__device__ __forceinline__ bool gl_tri(const float *a, const float *b, const float *c, float px, float py, float &zi)
{
    float area = gl_edge(a, b, c[0], c[1]);
    if (area == 0.0f) return false;
    if (area < 0.0f) {
        const float *t = b;
        b = c;
        c = t;
        area = -area;
    }
    const float ea = gl_edge(b, c, px, py), eb = gl_edge(c, a, px, py), ec = gl_edge(a, b, px, py);
    if (ea < 0.0f || (ea == 0.0f && !gl_owns(b, c))) return false;
    if (eb < 0.0f || (eb == 0.0f && !gl_owns(c, a))) return false;
    if (ec < 0.0f || (ec == 0.0f && !gl_owns(a, b))) return false;
    zi = 1.0f / ((ea / area) / a[2] + (eb / area) / b[2] + (ec / area) / c[2]);
    return true;
}

__global__ void k_zero_u32(unsigned int *__restrict__ p) { *p = 0u; }

__global__ void k_fill_u64(unsigned long long *__restrict__ p, size_t n, unsigned long long v)
{
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o < n) p[o] = v;
}

// depthSplatting.vert / .geom / .frag (scene.cpp:1506-1544): every pixel of
// valid depth of every view but `skip` is a quad, corners at the pixel centre
// +- 0.5 with depths z +- 0.5 dz/dx +- 0.5 dz/dy, drawn as the strip BL, TL,
// BR, TR into the target.  The depth test keeps the nearest quad at each
// covered pixel centre: a 64-bit atomicMin of (interpolated depth, target
// depth of the quad's centre).
//
// The quad of source pixel k = (v, mx, my): its corners (x, y, h2) in the
// target, the low word of its depth-test key and the box of target pixels it
// may cover, [px0, px1] x [py0, py1] clipped to the target (empty when
// px0 > px1 or py0 > py1); false when the quad is not drawn (a corner at or
// behind the near plane, a non-finite corner, or a box outside the target).
// `big`: the quad spans more than kGlMaxSpan target pixels in x or y; such
// quads are rasterised by a block each (k_gl_splat_big), not by one lane.
__device__ __forceinline__ bool gl_quad(int V, int W, int H, size_t n, size_t k, int mx, int my,
                                        const float *__restrict__ depth, const float *__restrict__ grad,
                                        const WarpCam &cam, float (&q)[4][3], unsigned int &lo, int &px0,
                                        int &px1, int &py0, int &py1, bool &big)
{
    const float z = depth[k];
    if (z >= 0.5f * kInvalidDepth) return false;
    const float gx = grad ? grad[k] : 0.0f, gy = grad ? grad[(size_t)V * n + k] : 0.0f;
    const float x = (float)mx + 0.5f, y = (float)my + 0.5f;
    bool ok = true;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const float sx = c < 2 ? -0.5f : 0.5f, sy = (c & 1) ? 0.5f : -0.5f;  // BL, TL, BR, TR
        const float zc = z + sx * gx + sy * gy, xc = x + sx, yc = y + sy;
        const float h0 = zc * (cam.M[0] * xc + cam.M[1] * yc + cam.M[2]) + cam.c[0];
        const float h1 = zc * (cam.M[3] * xc + cam.M[4] * yc + cam.M[5]) + cam.c[1];
        const float h2 = zc * (cam.M[6] * xc + cam.M[7] * yc + cam.M[8]) + cam.c[2];
        ok = ok && h2 > kGlZNear;
        q[c][0] = h0 / h2;
        q[c][1] = h1 / h2;
        q[c][2] = h2;
    }
    if (!ok) return false;
    const float zval = z * (cam.M[6] * x + cam.M[7] * y + cam.M[8]) + cam.c[2];
    float u0 = q[0][0], u1 = q[0][0], v0 = q[0][1], v1 = q[0][1];
#pragma unroll
    for (int c = 1; c < 4; c++) {
        u0 = fminf(u0, q[c][0]);
        u1 = fmaxf(u1, q[c][0]);
        v0 = fminf(v0, q[c][1]);
        v1 = fmaxf(v1, q[c][1]);
    }
    if (!(isfinite(u0) && isfinite(u1) && isfinite(v0) && isfinite(v1))) return false;
    if (u1 < 0.0f || v1 < 0.0f || u0 > (float)W || v0 > (float)H) return false;
    big = !(u1 - u0 <= (float)kGlMaxSpan && v1 - v0 <= (float)kGlMaxSpan);
    // clipped in float: a quad near the camera plane can project far beyond the int range
    px0 = (int)fmaxf(ceilf(u0 - 0.5f), 0.0f);
    px1 = (int)fminf(floorf(u1 - 0.5f), (float)(W - 1));
    py0 = (int)fmaxf(ceilf(v0 - 0.5f), 0.0f);
    py1 = (int)fminf(floorf(v1 - 0.5f), (float)(H - 1));
    lo = __float_as_uint(zval);
    return true;
}

// A big quad waits in the overflow list for k_gl_splat_big.
__device__ __forceinline__ void gl_defer(size_t k, unsigned int *__restrict__ big_count, unsigned int *__restrict__ big_list,
                                         unsigned int cap)
{
    const unsigned int i = atomicAdd(big_count, 1u);
    if (i < cap) big_list[i] = (unsigned int)k;
}

// Target pixel (px, py) against the quad's two triangles; the depth test.
__device__ __forceinline__ void gl_cover(const float (&q)[4][3], unsigned int lo, int px, int py, int W,
                                         unsigned long long *__restrict__ key)
{
    float zi;
    const float cx = (float)px + 0.5f, cy = (float)py + 0.5f;
    if (!gl_tri(q[0], q[1], q[2], cx, cy, zi) && !gl_tri(q[2], q[1], q[3], cx, cy, zi)) return;
    atomicMin(key + (size_t)py * W + px, ((unsigned long long)__float_as_uint(zi) << 32) | lo);
}

// One quad per thread (dsf 1: a quad covers one or two target pixels, and
// the lanes of a wave already write neighbouring pixels).
__global__ void k_gl_splat(int V, int w, int h, int W, int H, int skip, const float *__restrict__ depth,
                           const float *__restrict__ grad, const WarpCam *__restrict__ cams,
                           unsigned long long *__restrict__ key, unsigned int *__restrict__ big_count,
                           unsigned int *__restrict__ big_list, unsigned int big_cap)
{
    SOURCE_INDEX()
    if (v == skip) return;
    float q[4][3];
    unsigned int lo;
    int px0, px1, py0, py1;
    bool big = false;
    if (!gl_quad(V, W, H, n, k, mx, my, depth, grad, cams[v], q, lo, px0, px1, py0, py1, big)) return;
    if (big) {
        gl_defer(k, big_count, big_list, big_cap);
        return;
    }
    for (int py = py0; py <= py1; py++)
        for (int px = px0; px <= px1; px++) gl_cover(q, lo, px, py, W, key);
}

// dsf >= 2: a quad's box is about dsf + 1 pixels wide.  The columns of the 64
// boxes of a wave (one row of source pixels, blockDim.x == 64) are dealt to
// the lanes in order, 64 at a time, and each lane walks its column's rows, so
// an atomic instruction writes consecutive target pixels of a row: a few
// cache lines instead of one per quad (device-scope atomics are paid per
// line).  Every (quad, pixel) pair of k_gl_splat is tested once, with the
// same arithmetic: the same minima.
__global__ void k_gl_splat_wave(int V, int w, int h, int W, int H, int skip, const float *__restrict__ depth,
                                const float *__restrict__ grad, const WarpCam *__restrict__ cams,
                                unsigned long long *__restrict__ key, unsigned int *__restrict__ big_count,
                                unsigned int *__restrict__ big_list, unsigned int big_cap)
{
    const int mx = blockIdx.x * blockDim.x + threadIdx.x, my = blockIdx.y * blockDim.y + threadIdx.y;
    const int v = blockIdx.z, lane = threadIdx.x;
    const size_t n = (size_t)w * h;
    float q[4][3] = {};
    unsigned int lo = 0;
    int px0 = 0, px1 = -1, py0 = 0, py1 = -1;
    bool ok = mx < w && my < h && v != skip, big = false;
    if (ok)
        ok = gl_quad(V, W, H, n, (size_t)v * n + (size_t)my * w + mx, mx, my, depth, grad, cams[v], q, lo, px0,
                     px1, py0, py1, big);
    if (ok && big) {
        gl_defer((size_t)v * n + (size_t)my * w + mx, big_count, big_list, big_cap);
        ok = false;
    }
    const int bw = ok && py0 <= py1 ? max(px1 - px0 + 1, 0) : 0;
    // exclusive offsets of the boxes' columns in lane order (every lane takes
    // part in the shuffles below)
    int inc = bw;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(inc, off, 64);
        if (lane >= off) inc += u;
    }
    const int exc = inc - bw, total = __shfl(inc, 63, 64);
    for (int base = 0; base < total; base += 64) {
        const int f = base + lane;
        // the owner: the last lane whose offset is <= f (it has columns, and f is one of them)
        int o = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1)
            if (__shfl(exc, o + step, 64) <= f) o += step;
        float oq[4][3];
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int d = 0; d < 3; d++) oq[c][d] = __shfl(q[c][d], o, 64);
        const unsigned int olo = (unsigned int)__shfl((int)lo, o, 64);
        const int px = __shfl(px0, o, 64) + (f - __shfl(exc, o, 64));
        const int oy0 = __shfl(py0, o, 64), oy1 = __shfl(py1, o, 64);
        if (f >= total) continue;
        for (int py = oy0; py <= oy1; py++) gl_cover(oq, olo, px, py, W, key);
    }
}

// The quads of the overflow list, a block each: the threads walk the quad's
// box (clipped to the target) in row-major steps, so an atomic instruction
// writes consecutive target pixels.  Same arithmetic as k_gl_splat.
__global__ void __launch_bounds__(256)
k_gl_splat_big(int V, int w, int h, int W, int H, const float *__restrict__ depth, const float *__restrict__ grad,
               const WarpCam *__restrict__ cams, unsigned long long *__restrict__ key,
               const unsigned int *__restrict__ big_count, const unsigned int *__restrict__ big_list,
               unsigned int big_cap)
{
    const size_t n = (size_t)w * h;
    const unsigned int cnt = min(*big_count, big_cap);
    for (unsigned int e = blockIdx.x; e < cnt; e += gridDim.x) {
        const size_t k = big_list[e];
        const int v = (int)(k / n);
        const size_t m = k - (size_t)v * n;
        const int mx = (int)(m % w), my = (int)(m / w);
        float q[4][3];
        unsigned int lo;
        int px0, px1, py0, py1;
        bool big = false;
        if (!gl_quad(V, W, H, n, k, mx, my, depth, grad, cams[v], q, lo, px0, px1, py0, py1, big)) continue;
        if (px0 > px1 || py0 > py1) continue;
        const int bw = px1 - px0 + 1;
        const long long area = (long long)bw * (py1 - py0 + 1);
        for (long long i = threadIdx.x; i < area; i += blockDim.x)
            gl_cover(q, lo, px0 + (int)(i % bw), py0 + (int)(i / bw), W, key);
    }
}

// The target depth map: the winning quad's centre depth, 0 where no quad
// landed (the colour clear of scene.cpp:1515-1517).
__global__ void k_gl_depth(const unsigned long long *__restrict__ key, float *__restrict__ depth, size_t N)
{
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= N) return;
    const unsigned long long kv = key[o];
    depth[o] = kv == ~0ull ? 0.0f : __uint_as_float((unsigned int)(kv & 0xffffffffull));
}

// tauWarp.frag:25-74 and tauPartial.frag:23-47 for every pixel of every view:
// tau = x_u, or (-2, -2) outside (0, W) x (0, H) or at invalid depth, or
// (-1, -1) when the target depth map (bilinear, clamped edges, texel centres
// at +0.5) is nearer than the point by more than 0.1; wdef the deformation
// weight 1 / |det dtau/dx| capped at 30 (0 when invalid); dpart = sigma_z dtau/dz.
__global__ void k_gl_tau(int V, int w, int h, int W, int H, const float *__restrict__ depth,
                         const float *__restrict__ grad, const WarpCam *__restrict__ cams,
                         const float *__restrict__ ud, float sigma_z, float *__restrict__ tx,
                         float *__restrict__ ty, float *__restrict__ wdef, float *__restrict__ dpx,
                         float *__restrict__ dpy)
{
    SOURCE_INDEX()
    const WarpCam &cam = cams[v];
    const float z = depth[k];
    const float gx = grad ? grad[k] : 0.0f, gy = grad ? grad[(size_t)V * n + k] : 0.0f;
    const float x = (float)mx + 0.5f, y = (float)my + 0.5f;
    float Mx[3], hh[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        Mx[i] = cam.M[3 * i] * x + cam.M[3 * i + 1] * y + cam.M[3 * i + 2];
        hh[i] = z * Mx[i] + cam.c[i];
    }
    float u = 0.0f, vv = 0.0f;
    if (hh[2] != 0.0f) {
        u = hh[0] / hh[2];
        vv = hh[1] / hh[2];
    }
    dpx[k] = sigma_z * (Mx[0] - u * Mx[2]) / hh[2];
    dpy[k] = sigma_z * (Mx[1] - vv * Mx[2]) / hh[2];
    const float t0[3] = {x * gx + z, y * gx, gx}, t1[3] = {x * gy, y * gy + z, gy};
    float a0[3], a1[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        a0[i] = cam.M[3 * i] * t0[0] + cam.M[3 * i + 1] * t0[1] + cam.M[3 * i + 2] * t0[2];
        a1[i] = cam.M[3 * i] * t1[0] + cam.M[3 * i + 1] * t1[1] + cam.M[3 * i + 2] * t1[2];
    }
    const float j00 = (a0[0] - u * a0[2]) / hh[2], j10 = (a0[1] - vv * a0[2]) / hh[2];
    const float j01 = (a1[0] - u * a1[2]) / hh[2], j11 = (a1[1] - vv * a1[2]) / hh[2];
    const float wd = fminf(1.0f / fabsf(j00 * j11 - j10 * j01), 30.0f);
    if (u <= 0.0f || (float)W <= u || vv <= 0.0f || (float)H <= vv || z >= 0.5f * kInvalidDepth) {
        tx[k] = ty[k] = -2.0f;
        wdef[k] = 0.0f;
        return;
    }
    const float s = u - 0.5f, t = vv - 0.5f;
    const int i0 = (int)floorf(s), j0 = (int)floorf(t);
    const float fx = s - (float)i0, fy = t - (float)j0;
    const int xa = min(max(i0, 0), W - 1), xb = min(max(i0 + 1, 0), W - 1);
    const int ya = min(max(j0, 0), H - 1), yb = min(max(j0 + 1, 0), H - 1);
    const float du = (1.0f - fy) * ((1.0f - fx) * ud[(size_t)ya * W + xa] + fx * ud[(size_t)ya * W + xb]) +
                     fy * ((1.0f - fx) * ud[(size_t)yb * W + xa] + fx * ud[(size_t)yb * W + xb]);
    if (hh[2] > du + 0.1f) {
        tx[k] = ty[k] = -1.0f;
        wdef[k] = 0.0f;
        return;
    }
    tx[k] = u;
    ty[k] = vv;
    wdef[k] = wd;
}

// v = S u (gradientIBR_kernels.cuh:196-290), C <= 4 planes.
template <int KS>
__global__ void k_gather(int W, int H, int w, int h, const float *__restrict__ tx,
                         const float *__restrict__ ty, const float *__restrict__ u, int C,
                         float *__restrict__ out)
{
    const size_t n = (size_t)w * h, N = (size_t)W * H;
    const size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    const float t0 = tx[m], t1 = ty[m];
    int x0, y0;
    float wx[KS], wy[KS];
    if (gather_valid(t0, t1, W, H) && footprint<KS>(t0, t1, W, H, x0, y0, wx, wy)) {
#pragma unroll
        for (int i = 0; i < KS; i++) {
            const int y = y0 + i;
            if (y < 0 || y > H - 1) continue;
#pragma unroll
            for (int j = 0; j < KS; j++) {
                const int x = x0 + j;
                if (x < 0 || x > W - 1) continue;
                const float a = wy[i] * wx[j];
                for (int c = 0; c < C; c++) acc[c] += a * u[c * N + (size_t)y * W + x];
            }
        }
    }
    for (int c = 0; c < C; c++) out[c * n + m] = acc[c];
}

// out += S^T (wgt * in), wsum += S^T wgt with global float atomics; the
// reference's deconvolution_nonsep_param (gradientIBR_kernels.cuh:293-444) in
// atomic form.  Only cell-valid source pixels contribute.
template <int KS>
__global__ void k_splat_atomic(int W, int H, int w, int h, const float *__restrict__ tx,
                               const float *__restrict__ ty, const float *__restrict__ in, int C,
                               const float *__restrict__ wgt, float *out, float *wsum)
{
    const size_t n = (size_t)w * h, N = (size_t)W * H;
    const size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const float t0 = tx[m], t1 = ty[m];
    if (!cell_valid(t0, t1, W, H)) return;
    int x0, y0;
    float wx[KS], wy[KS];
    if (!footprint<KS>(t0, t1, W, H, x0, y0, wx, wy)) return;
    const float g = wgt ? wgt[m] : 1.0f;
    float val[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (out)
        for (int c = 0; c < C; c++) val[c] = in[c * n + m];
#pragma unroll
    for (int i = 0; i < KS; i++) {
        const int y = y0 + i;
        if (y < 0 || y > H - 1) continue;
#pragma unroll
        for (int j = 0; j < KS; j++) {
            const int x = x0 + j;
            if (x < 0 || x > W - 1) continue;
            const float a = wy[i] * wx[j] * g;
            const size_t o = (size_t)y * W + x;
            if (wsum) atomicAdd(wsum + o, a);
            if (out)
                for (int c = 0; c < C; c++) atomicAdd(out + c * N + o, a * val[c]);
        }
    }
}

// Sobel/8 of the Rec.709 luminance (gradientIBR_kernels.cuh:57-99); with a
// visibility plane, pixels with an invisible 3x3 neighbour are left unwritten
// (vtv_sr_kernels_unstructured.cu:120-126).
__global__ void k_sobel(int W, int H, const float *__restrict__ r, const float *__restrict__ g,
                        const float *__restrict__ b, float *gx, float *gy,
                        const float *__restrict__ visf, const uint8_t *__restrict__ visb)
{
    const int ox = blockIdx.x * blockDim.x + threadIdx.x;
    const int oy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ox >= W || oy >= H) return;
    const size_t o = (size_t)oy * W + ox;
    float GX = 0.0f, GY = 0.0f;
    if (0 < ox && ox < W - 1 && 0 < oy && oy < H - 1) {
        if (visf || visb) {
            bool ok = true;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    const size_t q = o + (long)dy * W + dx;
                    const bool vv = visf ? (visf[q] != 0.0f) : (visb[q] != 0);
                    ok = ok && vv;
                }
            if (!ok) return;
        }
        auto Y = [&](long d) { return 0.2126f * r[o + d] + 0.7152f * g[o + d] + 0.0722f * b[o + d]; };
        const float Ypp = Y(-W - 1), Ypc = Y(-W), Ypn = Y(-W + 1);
        const float Ycp = Y(-1), Ycn = Y(1);
        const float Ynp = Y(W - 1), Ync = Y(W), Ynn = Y(W + 1);
        GX = (Ypn + 2.0f * Ycn + Ynn - Ypp - 2.0f * Ycp - Ynp) * 0.125f;
        GY = (Ynp + 2.0f * Ync + Ynn - Ypp - 2.0f * Ypc - Ypn) * 0.125f;
    }
    gx[o] = GX;
    gy[o] = GY;
}

// Per-view Sobel of the luminance (V images of w x h) and per-view Laplacian
// of each channel (cuda_laplacian_device, vtv_sr_kernels_unstructured.cu:720-761).
__global__ void k_view_stencils(int V, int w, int h, const float *__restrict__ views,
                                float *__restrict__ vgrad, float *__restrict__ vlap)
{
    SOURCE_INDEX()
    const int ox = mx, oy = my;
    const float *r = views + (size_t)v * 3 * n, *g = r + n, *b = g + n;
    float GX = 0.0f, GY = 0.0f;
    if (0 < ox && ox < w - 1 && 0 < oy && oy < h - 1) {
        auto Y = [&](long d) { return 0.2126f * r[m + d] + 0.7152f * g[m + d] + 0.0722f * b[m + d]; };
        const float Ypp = Y(-w - 1), Ypc = Y(-w), Ypn = Y(-w + 1);
        const float Ycp = Y(-1), Ycn = Y(1);
        const float Ynp = Y(w - 1), Ync = Y(w), Ynn = Y(w + 1);
        GX = (Ypn + 2.0f * Ycn + Ynn - Ypp - 2.0f * Ycp - Ynp) * 0.125f;
        GY = (Ynp + 2.0f * Ync + Ynn - Ypp - 2.0f * Ypc - Ypn) * 0.125f;
    }
    vgrad[(size_t)v * 2 * n + m] = GX;
    vgrad[(size_t)v * 2 * n + n + m] = GY;
    for (int c = 0; c < 3; c++) {
        const float *u = views + (size_t)v * 3 * n + c * n;
        float s = 0.0f, wsum = 0.0f;
        if (ox > 0) { s += u[m - 1]; wsum += 1.0f; }
        if (oy > 0) { s += u[m - w]; wsum += 1.0f; }
        if (ox < w - 1) { s += u[m + 1]; wsum += 1.0f; }
        if (oy < h - 1) { s += u[m + w]; wsum += 1.0f; }
        vlap[(size_t)v * 3 * n + c * n + m] = s / wsum - u[m];
    }
}

// Sobel/8 of the luminance of every view (gradientIBR_kernels.cuh:57-99),
// LDS-tiled: a 64 x 16 tile computes Y once per pixel of the tile and its
// 1-pixel halo, then the 3 x 3 stencil reads LDS.  grid = (tiles x, tiles y, V).
constexpr int kVsW = 64, kVsH = 16;

__global__ void __launch_bounds__(kVsW * kVsH) k_view_sobel(int w, int h, const float *__restrict__ views,
                                                           float *__restrict__ vgrad)
{
    constexpr int LW = kVsW + 2, LH = kVsH + 2;
    __shared__ float Yl[LH * LW];
    const size_t n = (size_t)w * h;
    const int v = blockIdx.z;
    const int x0 = blockIdx.x * kVsW, y0 = blockIdx.y * kVsH;
    const float *r = views + (size_t)v * 3 * n, *g = r + n, *b = g + n;
    for (int i = threadIdx.x; i < LH * LW; i += blockDim.x) {
        const int x = x0 - 1 + i % LW, y = y0 - 1 + i / LW;
        float Y = 0.0f;
        if (x >= 0 && x < w && y >= 0 && y < h) {
            const size_t m = (size_t)y * w + x;
            Y = 0.2126f * r[m] + 0.7152f * g[m] + 0.0722f * b[m];
        }
        Yl[i] = Y;
    }
    __syncthreads();
    const int lx = threadIdx.x % kVsW, ly = threadIdx.x / kVsW;
    const int ox = x0 + lx, oy = y0 + ly;
    if (ox >= w || oy >= h) return;
    float GX = 0.0f, GY = 0.0f;
    if (0 < ox && ox < w - 1 && 0 < oy && oy < h - 1) {
        const int c = (ly + 1) * LW + lx + 1;
        const float Ypp = Yl[c - LW - 1], Ypc = Yl[c - LW], Ypn = Yl[c - LW + 1];
        const float Ycp = Yl[c - 1], Ycn = Yl[c + 1];
        const float Ynp = Yl[c + LW - 1], Ync = Yl[c + LW], Ynn = Yl[c + LW + 1];
        GX = (Ypn + 2.0f * Ycn + Ynn - Ypp - 2.0f * Ycp - Ynp) * 0.125f;
        GY = (Ynp + 2.0f * Ync + Ynn - Ypp - 2.0f * Ypc - Ypn) * 0.125f;
    }
    const size_t m = (size_t)oy * w + ox;
    vgrad[(size_t)v * 2 * n + m] = GX;
    vgrad[(size_t)v * 2 * n + n + m] = GY;
}

// omega = angular * deformation weight for every view pixel
// (compute_weights, gradientIBR.cu:558-648; angular_weights and
// gold_deform_weights, gradientIBR_kernels.cuh:102-194).  The warped target
// gradient is gathered from (gxs, gys) at the footprint of tau.
template <int KS, bool ZERO_GRAD>
__global__ void k_weights(int V, int W, int H, int w, int h, const float *__restrict__ tx,
                          const float *__restrict__ ty, const float *__restrict__ gxs,
                          const float *__restrict__ gys, const float *__restrict__ dpx,
                          const float *__restrict__ dpy, float sigma, float thr,
                          float *__restrict__ omega)
{
    SOURCE_INDEX()
    const float *vtx = tx + (size_t)v * n, *vty = ty + (size_t)v * n;
    // warped u gradient (convolution_nonsep_param); identically zero, and not
    // gathered, when the solution is the cleared image
    float gxv = 0.0f, gyv = 0.0f;
    const float t0 = vtx[m], t1 = vty[m];
    int x0, y0;
    float wx[KS], wy[KS];
    if (!ZERO_GRAD && gather_valid(t0, t1, W, H) && footprint<KS>(t0, t1, W, H, x0, y0, wx, wy)) {
#pragma unroll
        for (int i = 0; i < KS; i++) {
            const int y = y0 + i;
            if (y < 0 || y > H - 1) continue;
#pragma unroll
            for (int j = 0; j < KS; j++) {
                const int x = x0 + j;
                if (x < 0 || x > W - 1) continue;
                const float a = wy[i] * wx[j];
                gxv += a * gxs[(size_t)y * W + x];
                gyv += a * gys[(size_t)y * W + x];
            }
        }
    }
    // angular weight
    const float nrm = sqrtf(gxv * gxv + gyv * gyv);
    if (nrm > thr) {
        gxv = gxv / nrm * thr;
        gyv = gyv / nrm * thr;
    }
    float sg = ZERO_GRAD ? 0.0f : gxv * dpx[k] + gyv * dpy[k];
    if (sg > 0.5f) sg = 0.5f;
    const float ang = 1.0f / (sg * sg + sigma * sigma);
    // deformation weight; the 4-neighbourhood is loaded up front (clamped at
    // the border, where it is not used) so the loads are not serialised
    // behind the validity branches
    const int ml = my * w + (mx > 0 ? mx - 1 : mx), mr = my * w + (mx < w - 1 ? mx + 1 : mx);
    const int mu = (my > 0 ? my - 1 : my) * w + mx, md = (my < h - 1 ? my + 1 : my) * w + mx;
    const float xc = vtx[m], yc = vty[m];
    const float xl = vtx[ml], yl = vty[ml], xr = vtx[mr], yr = vty[mr];
    const float xu = vtx[mu], yu = vty[mu], xd = vtx[md], yd = vty[md];
    // non-short-circuit tests: every load above is consumed unconditionally,
    // which keeps the compiler from sinking them into dependent branches
    const float Wf = (float)W, Hf = (float)H;
    auto inval = [&](float a, float b) { return (a < 0.0f) | (a > Wf) | (b < 0.0f) | (b > Hf); };
    const bool bad_c = inval(xc, yc);
    const bool border = (mx == 0) | (my == 0) | (mx == w - 1) | (my == h - 1);
    const bool bad_n = inval(xl, yl) | inval(xr, yr) | inval(xu, yu) | inval(xd, yd);
    const float Dxx = 0.5f * (xr - xl);
    const float Dyx = 0.5f * (yr - yl);
    const float Dxy = 0.5f * (xd - xu);
    const float Dyy = 0.5f * (yd - yu);
    const float Nd = fabsf(Dxx * Dyy - Dxy * Dyx);
    const float dwi = (Nd > 1.0f) ? 1.0f / Nd : 1.0f;
    const float dw = bad_c ? 0.0f : ((border | bad_n) ? 1.0f : dwi);
    omega[k] = ang * dw;
}

const dim3 kSourceBlock(64, 4);

inline dim3 source_grid(const lfrv_ctx *ctx)
{
    return dim3((ctx->w + 63) / 64, (ctx->h + 3) / 4, ctx->V);
}

}  // namespace

// ---------------------------------------------------------------------------

static void mat3_mul(const double *A, const double *B, double *C)
{
    double T[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += A[3 * i + k] * B[3 * k + j];
            T[3 * i + j] = s;
        }
    for (int i = 0; i < 9; i++) C[i] = T[i];
}

void decompose_P(const double *P, double *K, double *R, double *t, double *C);  // camera.cpp

static void kinv_ref(const double *K, double *Ki)
{
    for (int i = 0; i < 9; i++) Ki[i] = 0.0;
    Ki[0] = 1.0 / K[0];
    Ki[4] = 1.0 / K[4];
    Ki[2] = -K[2] / K[0];
    Ki[5] = -K[5] / K[4];
    Ki[8] = 1.0;
}

// Camera algebra of the warps in FP64 on the host (import.cpp:301-320,
// 412-430), uploaded to ctx->dcams; redone only when a camera changed.
static lfrv_status upload_warp_cams(lfrv_ctx *ctx)
{
    if (!ctx->cams_valid) {
        std::vector<WarpCam> hc(ctx->V);
        double Ku[9], Ru[9], tu[3], Cu[3], Kui[9];
        decompose_P(ctx->P_target, Ku, Ru, tu, Cu);
        kinv_ref(Ku, Kui);
        const double *Pu = ctx->P_target;
        for (int v = 0; v < ctx->V; v++) {
            const double *Pv = ctx->P_views.data() + 12 * v;
            double Kv[9], Rv[9], tv[3], Cv[3], Kvi[9], RvT[9], Pu3[9], M[9], Pv3[9], RbT[9], Nb[9];
            decompose_P(Pv, Kv, Rv, tv, Cv);
            kinv_ref(Kv, Kvi);
            const double *Rb = ctx->p.backproject_source_cam ? Rv : Ru;
            const double *Cb = ctx->p.backproject_source_cam ? Cv : Cu;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    RvT[3 * i + j] = Rv[3 * j + i];
                    RbT[3 * i + j] = Rb[3 * j + i];
                    Pu3[3 * i + j] = Pu[4 * i + j];
                    Pv3[3 * i + j] = Pv[4 * i + j];
                }
            mat3_mul(RvT, Kvi, M);
            mat3_mul(Pu3, M, M);
            mat3_mul(RbT, Kui, Nb);
            mat3_mul(Pv3, Nb, Nb);
            WarpCam &c = hc[v];
            for (int i = 0; i < 9; i++) {
                c.M[i] = (float)M[i];
                c.Nb[i] = (float)Nb[i];
            }
            for (int i = 0; i < 3; i++) {
                c.c[i] = (float)(Pu[4 * i] * Cv[0] + Pu[4 * i + 1] * Cv[1] + Pu[4 * i + 2] * Cv[2] + Pu[4 * i + 3]);
                c.e[i] = (float)(Pv[4 * i] * Cb[0] + Pv[4 * i + 1] * Cb[1] + Pv[4 * i + 2] * Cb[2] + Pv[4 * i + 3]);
            }
        }
        LFRV_HIP(ctx, hipMemcpyAsync(ctx->dcams, hc.data(), hc.size() * sizeof(WarpCam), hipMemcpyHostToDevice,
                                     ctx->stream));
        ctx->cams_valid = true;
    }
    return LFRV_OK;
}

static lfrv_status check_depth_inputs(lfrv_ctx *ctx, const char *who)
{
    if (!ctx->have_target_cam) return fail(ctx, LFRV_E_STATE, std::string(who) + ": target camera not set");
    for (int v = 0; v < ctx->V; v++)
        if (!ctx->have_depth[v]) return fail(ctx, LFRV_E_STATE, std::string(who) + ": depth of a view not set");
    return LFRV_OK;
}

lfrv_status run_warps_from_depth(lfrv_ctx *ctx)
{
    lfrv_status st = check_depth_inputs(ctx, "lfrv_warps_from_depth");
    if (st == LFRV_OK) st = upload_warp_cams(ctx);
    if (st != LFRV_OK) return st;
    const WarpCam *dcams = ctx->dcams;
    int *zkey = reinterpret_cast<int *>(ctx->tmpN);
    TimedRegion tr(ctx, 4);
    int inv_key;
    {
        const float f = kInvalidDepth;
        std::memcpy(&inv_key, &f, sizeof inv_key);  // positive float: key == bit pattern
    }
    const size_t ncnt = (size_t)ctx->ncells + 1;
    k_fill_key<<<blocks_for(std::max(ctx->N, ncnt)), kBlock, 0, ctx->stream>>>(zkey, ctx->N, inv_key, ctx->cell_count,
                                                                              ncnt);
    // IBR_gradient's first weights take the zero-gradient form, which does not
    // read dpart: it is formed later, when something reads it
    const bool lazy = ctx->p.mode == LFRV_MODE_GRADIENT_IBR;
    k_tau_from_depth<<<source_grid(ctx), kSourceBlock, 0, ctx->stream>>>(
        ctx->V, ctx->w, ctx->h, ctx->dsf, ctx->depth, dcams, ctx->p.sigma_z, ctx->tau_x, ctx->tau_y,
        lazy ? nullptr : ctx->dpart_x, lazy ? nullptr : ctx->dpart_y, zkey);
    ctx->dpart_pending = lazy;
    k_key_to_depth<<<blocks_for(ctx->N), kBlock, 0, ctx->stream>>>(zkey, ctx->target_depth, ctx->N);
#define TV(K) k_tau_visibility<K><<<source_grid(ctx), kSourceBlock, 0, ctx->stream>>>(ctx->V, ctx->w, ctx->h, ctx->target_depth, dcams, ctx->p.occlusion_tol_px, ctx->tau_x, ctx->tau_y, ctx->cell_count)
    switch (ctx->ks) {
    case 2: TV(2); break;
    case 3: TV(3); break;
    case 4: TV(4); break;
    case 5: TV(5); break;
    default: return fail(ctx, LFRV_E_ARG, "dsf out of range");
    }
#undef TV
    LFRV_LAUNCH(ctx, "warps_from_depth kernels");
    ctx->prepared = false;
    ctx->epoch++;
    ctx->counts_valid = true;
    return LFRV_OK;
}

lfrv_status run_warps_from_target_depth(lfrv_ctx *ctx, int skip)
{
    lfrv_status st = check_depth_inputs(ctx, "lfrv_warps_from_target_depth");
    if (st == LFRV_OK) st = upload_warp_cams(ctx);
    if (st != LFRV_OK) return st;
    if (!ctx->tau_weight) LFRV_HIP(ctx, hipMalloc((void **)&ctx->tau_weight, (size_t)ctx->V * ctx->n * sizeof(float)));
    // the depth-test keys in the 3N scratch (8 bytes per target pixel)
    unsigned long long *key = reinterpret_cast<unsigned long long *>(ctx->tmp3N);
    TimedRegion tr(ctx, 4);
    // the overflow list of big quads: one slot per source pixel, and its count
    const size_t vn = (size_t)ctx->V * ctx->n;
    if (!ctx->gl_big) LFRV_HIP(ctx, hipMalloc((void **)&ctx->gl_big, (vn + 1) * sizeof(unsigned int)));
    unsigned int *big_count = ctx->gl_big, *big_list = ctx->gl_big + 1;
    const unsigned int big_cap = (unsigned int)std::min<size_t>(vn, 0xffffffffu);
    k_fill_u64<<<blocks_for(ctx->N), kBlock, 0, ctx->stream>>>(key, ctx->N, ~0ull);
    k_zero_u32<<<1, 1, 0, ctx->stream>>>(big_count);
    if (ctx->dsf == 1)
        k_gl_splat<<<source_grid(ctx), kSourceBlock, 0, ctx->stream>>>(ctx->V, ctx->w, ctx->h, ctx->W, ctx->H, skip,
                                                                       ctx->depth, ctx->depth_grad, ctx->dcams, key,
                                                                       big_count, big_list, big_cap);
    else
        k_gl_splat_wave<<<source_grid(ctx), kSourceBlock, 0, ctx->stream>>>(ctx->V, ctx->w, ctx->h, ctx->W, ctx->H, skip,
                                                                       ctx->depth, ctx->depth_grad, ctx->dcams, key,
                                                                       big_count, big_list, big_cap);
    k_gl_splat_big<<<1024, 256, 0, ctx->stream>>>(ctx->V, ctx->w, ctx->h, ctx->W, ctx->H, ctx->depth, ctx->depth_grad,
                                                  ctx->dcams, key, big_count, big_list, big_cap);
    k_gl_depth<<<blocks_for(ctx->N), kBlock, 0, ctx->stream>>>(key, ctx->target_depth, ctx->N);
    k_gl_tau<<<source_grid(ctx), kSourceBlock, 0, ctx->stream>>>(
        ctx->V, ctx->w, ctx->h, ctx->W, ctx->H, ctx->depth, ctx->depth_grad, ctx->dcams, ctx->target_depth,
        ctx->p.sigma_z, ctx->tau_x, ctx->tau_y, ctx->tau_weight, ctx->dpart_x, ctx->dpart_y);
    LFRV_LAUNCH(ctx, "warps_from_target_depth kernels");
    ctx->dpart_pending = false;
    ctx->prepared = false;
    ctx->epoch++;
    ctx->counts_valid = false;  // the splat structure counts its cells itself
    return LFRV_OK;
}

lfrv_status materialize_dpart(lfrv_ctx *ctx)
{
    if (!ctx->dpart_pending) return LFRV_OK;
    k_dpart_from_depth<<<source_grid(ctx), kSourceBlock, 0, ctx->stream>>>(
        ctx->V, ctx->w, ctx->h, ctx->depth, ctx->dcams, ctx->p.sigma_z, ctx->dpart_x, ctx->dpart_y);
    LFRV_LAUNCH(ctx, "dpart from depth");
    ctx->dpart_pending = false;
    return LFRV_OK;
}

template <int KS>
static void launch_weights_ks(lfrv_ctx *ctx, const float *gxs, const float *gys, bool zero_grad, hipStream_t st)
{
    if (zero_grad)
        k_weights<KS, true><<<source_grid(ctx), kSourceBlock, 0, st>>>(
            ctx->V, ctx->W, ctx->H, ctx->w, ctx->h, ctx->tau_x, ctx->tau_y, gxs, gys, ctx->dpart_x,
            ctx->dpart_y, ctx->p.sigma_sensor, ctx->p.ugrad_threshold, ctx->omega);
    else
        k_weights<KS, false><<<source_grid(ctx), kSourceBlock, 0, st>>>(
            ctx->V, ctx->W, ctx->H, ctx->w, ctx->h, ctx->tau_x, ctx->tau_y, gxs, gys, ctx->dpart_x,
            ctx->dpart_y, ctx->p.sigma_sensor, ctx->p.ugrad_threshold, ctx->omega);
}

lfrv_status run_compute_weights(lfrv_ctx *ctx, hipStream_t st)
{
    if (!st) st = ctx->stream;
    const float *gxs, *gys;
    dim3 bs(32, 8), gs((ctx->W + 31) / 32, (ctx->H + 7) / 8);
    const size_t N = ctx->N;
    if (ctx->p.mode == LFRV_MODE_COLOR_SR) {
        // vtv_sr_compute_gradient_device writes the gated Sobel of U into the
        // dual variable of channel 0 (optimization.cu:1090-1091); kept as is.
        float *px0 = ctx->PX, *py0 = ctx->PY;  // the current dual buffers
        k_sobel<<<gs, bs, 0, st>>>(ctx->W, ctx->H, ctx->U, ctx->U + N, ctx->U + 2 * N, px0, py0,
                                            ctx->visibility, nullptr);
        gxs = px0;
        gys = py0;
    } else {
        // compute_gradient over the whole target (gradientIBR.cu:578-580 launched
        // it on the low-res grid only; the build covers W x H).  The ungated
        // Sobel of the cleared image is zero everywhere, so is its warp: the
        // weights then take the zero-gradient form and need no Sobel.
        if (!ctx->u_zero)
            k_sobel<<<gs, bs, 0, st>>>(ctx->W, ctx->H, ctx->U, ctx->U + N, ctx->U + 2 * N,
                                                ctx->u_grad_x, ctx->u_grad_y, nullptr, nullptr);
        gxs = ctx->u_grad_x;
        gys = ctx->u_grad_y;
    }
    const bool zero_grad = ctx->p.mode == LFRV_MODE_GRADIENT_IBR && ctx->u_zero;
    if (!zero_grad) {
        const lfrv_status ds = materialize_dpart(ctx);
        if (ds != LFRV_OK) return ds;
    }
    switch (ctx->ks) {
    case 2: launch_weights_ks<2>(ctx, gxs, gys, zero_grad, st); break;
    case 3: launch_weights_ks<3>(ctx, gxs, gys, zero_grad, st); break;
    case 4: launch_weights_ks<4>(ctx, gxs, gys, zero_grad, st); break;
    case 5: launch_weights_ks<5>(ctx, gxs, gys, zero_grad, st); break;
    default: return fail(ctx, LFRV_E_ARG, "dsf out of range");
    }
    LFRV_LAUNCH(ctx, "compute_weights kernels");
    ctx->entry_weights_valid = false;
    ctx->epoch++;
    return LFRV_OK;
}

lfrv_status run_view_stencils(lfrv_ctx *ctx, hipStream_t s)
{
    if (ctx->p.mode == LFRV_MODE_GRADIENT_IBR) {
        // IBR_gradient needs the views' Sobel only (init_u_gradient)
        dim3 gs((ctx->w + kVsW - 1) / kVsW, (ctx->h + kVsH - 1) / kVsH, ctx->V);
        k_view_sobel<<<gs, kVsW * kVsH, 0, s>>>(ctx->w, ctx->h, ctx->views, ctx->view_grad);
        LFRV_LAUNCH(ctx, "view stencils");
        return LFRV_OK;
    }
    k_view_stencils<<<source_grid(ctx), kSourceBlock, 0, s>>>(ctx->V, ctx->w, ctx->h, ctx->views, ctx->view_grad,
                                                              ctx->view_lap);
    LFRV_LAUNCH(ctx, "view stencils");
    return LFRV_OK;
}

hipError_t launch_gather(int W, int H, int w, int h, int dsf, const float *tx, const float *ty,
                         const float *u, int C, float *out, hipStream_t s)
{
    const size_t n = (size_t)w * h;
    switch (dsf + 1) {
    case 2: k_gather<2><<<blocks_for(n), kBlock, 0, s>>>(W, H, w, h, tx, ty, u, C, out); break;
    case 3: k_gather<3><<<blocks_for(n), kBlock, 0, s>>>(W, H, w, h, tx, ty, u, C, out); break;
    case 4: k_gather<4><<<blocks_for(n), kBlock, 0, s>>>(W, H, w, h, tx, ty, u, C, out); break;
    case 5: k_gather<5><<<blocks_for(n), kBlock, 0, s>>>(W, H, w, h, tx, ty, u, C, out); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_splat_atomic(int W, int H, int w, int h, int dsf, const float *tx,
                               const float *ty, const float *in, int C, const float *wgt,
                               float *out, float *wsum, hipStream_t s)
{
    const size_t n = (size_t)w * h;
    switch (dsf + 1) {
    case 2: k_splat_atomic<2><<<blocks_for(n), kBlock, 0, s>>>(W, H, w, h, tx, ty, in, C, wgt, out, wsum); break;
    case 3: k_splat_atomic<3><<<blocks_for(n), kBlock, 0, s>>>(W, H, w, h, tx, ty, in, C, wgt, out, wsum); break;
    case 4: k_splat_atomic<4><<<blocks_for(n), kBlock, 0, s>>>(W, H, w, h, tx, ty, in, C, wgt, out, wsum); break;
    case 5: k_splat_atomic<5><<<blocks_for(n), kBlock, 0, s>>>(W, H, w, h, tx, ty, in, C, wgt, out, wsum); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_sobel(int W, int H, const float *r, const float *g, const float *b, float *gx,
                        float *gy, const float *visf, const uint8_t *visb, hipStream_t s)
{
    dim3 bs(32, 8), gs((W + 31) / 32, (H + 7) / 8);
    k_sobel<<<gs, bs, 0, s>>>(W, H, r, g, b, gx, gy, visf, visb);
    return hipGetLastError();
}


// Radial to orthographic depth, the depth-map filter of the OpenGL path
// (fromRadial2Ortho.frag, run per view by Scene::filterDepthMaps,
// scene.cpp:1605-1624, through InputView::filterDepthmap, view.cpp:723-733,
// and FrameBuffer::fromRadial2Ortho, frameBuffer.cpp:263-315).  The fragment
// of pixel (x, y) sits at its centre: u = K^-1 (x + 0.5, y + 0.5, 1), d = |u|,
// and the depth and its derivatives become z / d and
// dz / d - z (u . K^-1 e_x or e_y) / d^3, in float as the shader computes.
// Every pixel is converted, the invalid ones (z = 1000) included.
struct Mat3f {
    float m[9];  // row-major
};

__global__ void k_radial_to_ortho(int W, int H, Mat3f ki, float *__restrict__ z, float *__restrict__ gx,
                                  float *__restrict__ gy)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    const size_t o = (size_t)y * W + x;
    const float px = (float)x + 0.5f, py = (float)y + 0.5f;
    const float u0 = ki.m[0] * px + ki.m[1] * py + ki.m[2];
    const float u1 = ki.m[3] * px + ki.m[4] * py + ki.m[5];
    const float u2 = ki.m[6] * px + ki.m[7] * py + ki.m[8];
    const float d = sqrtf(u0 * u0 + u1 * u1 + u2 * u2);
    const float d3 = d * d * d;
    // u . (column 0 of K^-1) and u . (column 1 of K^-1)
    const float c0 = u0 * ki.m[0] + u1 * ki.m[3] + u2 * ki.m[6];
    const float c1 = u0 * ki.m[1] + u1 * ki.m[4] + u2 * ki.m[7];
    const float zr = z[o];
    z[o] = zr / d;
    if (gx) gx[o] = gx[o] / d - zr * c0 / d3;
    if (gy) gy[o] = gy[o] / d - zr * c1 / d3;
}

hipError_t launch_radial_to_ortho(int W, int H, const double *K, float *z, float *gx, float *gy, hipStream_t s)
{
    // K^-1 in double, handed to the kernel in float (the shader's mat3 uniform)
    const double a = K[0], b = K[1], c = K[2], d = K[3], e = K[4], f = K[5], g = K[6], h = K[7], i = K[8];
    const double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
    if (!(det != 0.0) || !std::isfinite(det)) return hipErrorInvalidValue;
    const double inv[9] = {(e * i - f * h) / det, (c * h - b * i) / det, (b * f - c * e) / det,
                           (f * g - d * i) / det, (a * i - c * g) / det, (c * d - a * f) / det,
                           (d * h - e * g) / det, (b * g - a * h) / det, (a * e - b * d) / det};
    Mat3f ki;
    for (int k = 0; k < 9; k++) ki.m[k] = (float)inv[k];
    const dim3 bs(32, 8), gs((W + 31) / 32, (H + 7) / 8);
    k_radial_to_ortho<<<gs, bs, 0, s>>>(W, H, ki, z, gx, gy);
    return hipGetLastError();
}

}  // namespace lfrv
