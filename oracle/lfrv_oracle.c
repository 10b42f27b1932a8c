/*
 * lfrv_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * FP64 CPU restatement of the reference's variational light-field hot path
 * (Nighteye/light-field-remote-vision).  It is the checker for the HIP path
 * and the CPU baseline timed by bench.py.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library.
 *
 * Every function names the reference file:line it follows.  Arithmetic is
 * double throughout; inputs that the reference holds in float (tau, images)
 * are taken as float and widened, so tap selection (floor of tau) is made on
 * the same float values the GPU sees.
 *
 * Parity pins (checks that do not retype these formulas) live in
 * tests/test_oracle_*.py: closed forms (identity/integer-shift warps, ramps,
 * constants, fronto-parallel disparity), adjoint identities, dense brute-force
 * solves, idempotence of projections, hand-computed PSNR.
 *
 * Layout: images are planar, channel c of an image of N pixels starts at c*N;
 * rows are contiguous, pixel (x,y) is at y*W + x, top-left origin.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_INVALID_DEPTH 1000.0 /* sources/import.cpp:23 */
#define ORC_MAXKS 9

/* ------------------------------------------------------------------------- */
/* Footprint matrix B                                                        */
/* ------------------------------------------------------------------------- */

/* Weights of the dsf-wide box footprint of one warped pixel.
 * gradientIBR_kernels.cuh:447-507 (set_B_bilinear),
 * vtv_sr_kernels_unstructured.cu:658-718 (cuda_set_A_bilinear).
 * Taps are x0+j, y0+i for i,j < ks = dsf+1; A[ks*i+j] = wy[i]*wx[j].
 * The reference normalises the 2-D products by their sum over in-image taps;
 * that sum factorises into (sum wx)(sum wy), so the 1-D factors are returned
 * normalised.  Returns ks, or 0 when every in-image tap has zero weight. */
int orc_footprint(float tx, float ty, int W, int H, int dsf, int *x0, int *y0,
                  double *wx, double *wy)
{
    const int ks = dsf + 1;
    const double px = (double)tx - 0.5, py = (double)ty - 0.5;
    const double hd = 0.5 * (double)dsf;
    const int bx = (int)floor(px - hd + 0.5);
    const int by = (int)floor(py - hd + 0.5);
    double sx = 0.0, sy = 0.0;
    for (int j = 0; j < ks; j++) {
        const int x = bx + j;
        double d = 0.0;
        if (x >= 0 && x <= W - 1) {
            d = 1.0;
            if (j == 0) d = ((double)x + 0.5) - (px - hd);
            else if (j == ks - 1) d = (px + hd) - ((double)x - 0.5);
        }
        wx[j] = d;
        sx += d;
    }
    for (int i = 0; i < ks; i++) {
        const int y = by + i;
        double d = 0.0;
        if (y >= 0 && y <= H - 1) {
            d = 1.0;
            if (i == 0) d = ((double)y + 0.5) - (py - hd);
            else if (i == ks - 1) d = (py + hd) - ((double)y - 0.5);
        }
        wy[i] = d;
        sy += d;
    }
    *x0 = bx;
    *y0 = by;
    if (sx * sy == 0.0) {
        for (int j = 0; j < ks; j++) wx[j] = 0.0;
        for (int i = 0; i < ks; i++) wy[i] = 0.0;
        return 0;
    }
    for (int j = 0; j < ks; j++) wx[j] /= sx;
    for (int i = 0; i < ks; i++) wy[i] /= sy;
    return ks;
}

/* Gather validity: gradientIBR_kernels.cuh:216-219 (tau < 0 or > W/H is invalid). */
static int gather_valid(float tx, float ty, int W, int H)
{
    return !(tx < 0.0f || ty < 0.0f || tx > (float)W || ty > (float)H);
}

/* Splat validity: a source pixel is splatted only if it is in the cell list,
 * gradientIBR.cu:404-426 and optimization.cu:907-929: gather-valid and
 * floor(tau - 0.5) inside the target. */
int orc_cell_valid(float tx, float ty, int W, int H)
{
    if (!gather_valid(tx, ty, W, H)) return 0;
    const float ux = (float)((double)tx - 0.5), uy = (float)((double)ty - 0.5);
    const int px = (int)floorf(ux), py = (int)floorf(uy);
    if (px < 0 || py < 0 || px > W - 1 || py > H - 1) return 0;
    return 1;
}

/* Warp / gather  v(m) = sum_k A[m,k] u(x_k).
 * gradientIBR_kernels.cuh:196-290, vtv_sr_kernels_unstructured.cu:268-362.
 * u: C planes of W*H; out: C planes of w*h. */
void orc_gather(int W, int H, int w, int h, int dsf, const float *tx, const float *ty,
                const double *u, int C, double *out)
{
    const size_t N = (size_t)W * H, n = (size_t)w * h;
#pragma omp parallel for schedule(static)
    for (long m = 0; m < (long)n; m++) {
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (gather_valid(tx[m], ty[m], W, H)) {
            int x0, y0;
            double wx[ORC_MAXKS], wy[ORC_MAXKS];
            const int ks = orc_footprint(tx[m], ty[m], W, H, dsf, &x0, &y0, wx, wy);
            for (int i = 0; i < ks; i++) {
                const int y = y0 + i;
                if (y < 0 || y > H - 1) continue;
                for (int j = 0; j < ks; j++) {
                    const int x = x0 + j;
                    if (x < 0 || x > W - 1) continue;
                    const double a = wy[i] * wx[j];
                    for (int c = 0; c < C; c++) acc[c] += a * u[c * N + (size_t)y * W + x];
                }
            }
        }
        for (int c = 0; c < C; c++) out[c * n + m] = acc[c];
    }
}

/* Splat / adjoint  O[x_k] += A[m,k] * wgt(m) * I(m),  wsum[x_k] += A[m,k] * wgt(m).
 * gradientIBR_kernels.cuh:293-388 (deconvolution_nonsep_param),
 * :391-444 (weighted_deconvolution_nonsep_param),
 * vtv_sr_kernels_unstructured.cu:365-516.  Only cell-valid source pixels are
 * splatted (the reference iterates its cell list).  wgt == NULL means 1.
 * out / wsum may be NULL; both are accumulated into (not cleared). */
void orc_splat(int W, int H, int w, int h, int dsf, const float *tx, const float *ty,
               const double *in, int C, const double *wgt, double *out, double *wsum)
{
    const size_t N = (size_t)W * H, n = (size_t)w * h;
    for (size_t m = 0; m < n; m++) {
        if (!orc_cell_valid(tx[m], ty[m], W, H)) continue;
        int x0, y0;
        double wx[ORC_MAXKS], wy[ORC_MAXKS];
        const int ks = orc_footprint(tx[m], ty[m], W, H, dsf, &x0, &y0, wx, wy);
        const double g = wgt ? wgt[m] : 1.0;
        for (int i = 0; i < ks; i++) {
            const int y = y0 + i;
            if (y < 0 || y > H - 1) continue;
            for (int j = 0; j < ks; j++) {
                const int x = x0 + j;
                if (x < 0 || x > W - 1) continue;
                const double a = wy[i] * wx[j] * g;
                const size_t o = (size_t)y * W + x;
                if (wsum) wsum[o] += a;
                if (out)
                    for (int c = 0; c < C; c++) out[c * N + o] += a * in[c * n + m];
            }
        }
    }
}

/* Raw (unnormalised) footprint hit test used by the target visibility mask:
 * gradientIBR_kernels.cuh:4-54 marks x_k visible when dx*dy != 0. */
void orc_target_visibility(int W, int H, int w, int h, int dsf, const float *tx,
                           const float *ty, unsigned char *vis)
{
    const size_t n = (size_t)w * h;
    const int ks = dsf + 1;
    for (size_t m = 0; m < n; m++) {
        if (!orc_cell_valid(tx[m], ty[m], W, H)) continue;
        const double px = (double)tx[m] - 0.5, py = (double)ty[m] - 0.5, hd = 0.5 * dsf;
        const int bx = (int)floor(px - hd + 0.5), by = (int)floor(py - hd + 0.5);
        for (int j = 0; j < ks; j++)
            for (int i = 0; i < ks; i++) {
                const int x = bx + j, y = by + i;
                if (x < 0 || x > W - 1 || y < 0 || y > H - 1) continue;
                double dx = 1.0, dy = 1.0;
                if (j == 0) dx = (x + 0.5) - (px - hd);
                else if (j == ks - 1) dx = (px + hd) - (x - 0.5);
                if (i == 0) dy = (y + 0.5) - (py - hd);
                else if (i == ks - 1) dy = (py + hd) - (y - 0.5);
                if (dx * dy != 0.0) vis[(size_t)y * W + x] = 1;
            }
    }
}

/* ------------------------------------------------------------------------- */
/* Stencils                                                                  */
/* ------------------------------------------------------------------------- */

/* Sobel/8 of the Rec.709 luminance, zero on the 1-px border.
 * gradientIBR_kernels.cuh:57-99 (compute_gradient) and, with the 3x3
 * visibility gate, vtv_sr_kernels_unstructured.cu:96-147: an interior pixel
 * whose 3x3 neighbourhood holds an invisible pixel is left unwritten. */
void orc_sobel(int W, int H, const double *r, const double *g, const double *b,
               double *gx, double *gy, const unsigned char *vis)
{
#pragma omp parallel for schedule(static)
    for (long oy = 0; oy < H; oy++)
        for (long ox = 0; ox < W; ox++) {
            const size_t o = (size_t)oy * W + ox;
            double GX = 0.0, GY = 0.0;
            if (0 < ox && ox < W - 1 && 0 < oy && oy < H - 1) {
                if (vis) {
                    int ok = 1;
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++)
                            if (!vis[o + dy * W + dx]) ok = 0;
                    if (!ok) continue;
                }
#define Y_(d) (0.2126 * r[o + (d)] + 0.7152 * g[o + (d)] + 0.0722 * b[o + (d)])
                const double Ypp = Y_(-W - 1), Ypc = Y_(-W), Ypn = Y_(-W + 1);
                const double Ycp = Y_(-1), Ycn = Y_(1);
                const double Ynp = Y_(W - 1), Ync = Y_(W), Ynn = Y_(W + 1);
#undef Y_
                GX = (Ypn + 2 * Ycn + Ynn - Ypp - 2 * Ycp - Ynp) / 8.0;
                GY = (Ynp + 2 * Ync + Ynn - Ypp - 2 * Ypc - Ypn) / 8.0;
            }
            gx[o] = GX;
            gy[o] = GY;
        }
}

/* Mean of the 4-neighbours minus the centre.
 * mode 0: cuda_laplacian_device (vtv_sr_kernels_unstructured.cu:720-761), vis NULL.
 * mode 1: compute_laplacian (gradientIBR_kernels.cuh:509-550): output 0 where vis[o]. */
void orc_laplacian(int W, int H, const double *in, double *out, const unsigned char *vis)
{
#pragma omp parallel for schedule(static)
    for (long oy = 0; oy < H; oy++)
        for (long ox = 0; ox < W; ox++) {
            const size_t o = (size_t)oy * W + ox;
            if (vis && vis[o]) {
                out[o] = 0.0;
                continue;
            }
            double s = 0.0, wsum = 0.0;
            if (ox > 0) { s += in[o - 1]; wsum += 1.0; }
            if (oy > 0) { s += in[o - W]; wsum += 1.0; }
            if (ox < W - 1) { s += in[o + 1]; wsum += 1.0; }
            if (oy < H - 1) { s += in[o + W]; wsum += 1.0; }
            out[o] = s / wsum - in[o];
        }
}

/* Forward-difference Neumann D^T D u (solver.cu:229-266) */
static double dtd_at(const double *u, int W, int H, int ox, int oy)
{
    const size_t o = (size_t)oy * W + ox;
    double lx = 0.0, ly = 0.0;
    if (W > 1) {
        if (ox == 0) lx = u[o] - u[o + 1];
        else if (ox < W - 1) lx = 2 * u[o] - u[o - 1] - u[o + 1];
        else lx = u[o] - u[o - 1];
    }
    if (H > 1) {
        if (oy == 0) ly = u[o] - u[o + W];
        else if (oy < H - 1) ly = 2 * u[o] - u[o - W] - u[o + W];
        else ly = u[o] - u[o - W];
    }
    return lx + ly;
}

/* D^T g for forward differences: (initConjGrad, solver.cu:384-467) */
static double dtg_at(const double *gxv, const double *gyv, int W, int H, int ox, int oy)
{
    const size_t o = (size_t)oy * W + ox;
    double v = 0.0;
    if (ox > 0) v += gxv[o - 1];
    if (ox < W - 1) v -= gxv[o];
    if (oy > 0) v += gyv[o - W];
    if (oy < H - 1) v -= gyv[o];
    return v;
}

void orc_dtd(int W, int H, const double *u, double *out)
{
#pragma omp parallel for schedule(static)
    for (long oy = 0; oy < H; oy++)
        for (long ox = 0; ox < W; ox++) out[(size_t)oy * W + ox] = dtd_at(u, W, H, (int)ox, (int)oy);
}

void orc_dtg(int W, int H, const double *gxv, const double *gyv, double *out)
{
#pragma omp parallel for schedule(static)
    for (long oy = 0; oy < H; oy++)
        for (long ox = 0; ox < W; ox++)
            out[(size_t)oy * W + ox] = dtg_at(gxv, gyv, W, H, (int)ox, (int)oy);
}

/* ------------------------------------------------------------------------- */
/* Weights                                                                   */
/* ------------------------------------------------------------------------- */

/* omega = 1 / (min(sigma_geom, .5)^2 + sigma_s^2); gradientIBR_kernels.cuh:102-143 */
void orc_angular_weights(int w, int h, const double *gx, const double *gy, double sigma,
                         const float *dpx, const float *dpy, double thr, double *out)
{
    const size_t n = (size_t)w * h;
    for (size_t m = 0; m < n; m++) {
        double ax = gx[m], ay = gy[m];
        const double nrm = sqrt(ax * ax + ay * ay);
        if (nrm > thr) {
            ax = ax / nrm * thr;
            ay = ay / nrm * thr;
        }
        double sg = ax * dpx[m] + ay * dpy[m];
        if (sg > 0.5) sg = 0.5;
        out[m] = 1.0 / (sg * sg + sigma * sigma);
    }
}

/* |det D tau|^-1 capped at 1; gradientIBR_kernels.cuh:146-194 */
void orc_deform_weights(int W, int H, int w, int h, const float *tx, const float *ty, double *out)
{
    for (long my = 0; my < h; my++)
        for (long mx = 0; mx < w; mx++) {
            const size_t m = (size_t)my * w + mx;
#define INV_(k) (tx[k] < 0 || tx[k] > W || ty[k] < 0 || ty[k] > H)
            if (INV_(m)) { out[m] = 0.0; continue; }
            if (mx == 0 || my == 0 || mx == w - 1 || my == h - 1) { out[m] = 1.0; continue; }
            if (INV_(m - 1) || INV_(m + 1) || INV_(m - w) || INV_(m + w)) { out[m] = 1.0; continue; }
#undef INV_
            const double Dxx = 0.5 * ((double)tx[m + 1] - tx[m - 1]);
            const double Dyx = 0.5 * ((double)ty[m + 1] - ty[m - 1]);
            const double Dxy = 0.5 * ((double)tx[m + w] - tx[m - w]);
            const double Dyy = 0.5 * ((double)ty[m + w] - ty[m - w]);
            const double Nd = fabs(Dxx * Dyy - Dxy * Dyx);
            out[m] = (Nd > 1.0) ? 1.0 / Nd : 1.0;
        }
}

/* ------------------------------------------------------------------------- */
/* Cameras and warps from depth                                              */
/* ------------------------------------------------------------------------- */

static void mat3_mul(const double *A, const double *B, double *C) /* row-major 3x3 */
{
    double T[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += A[3 * i + k] * B[3 * k + j];
            T[3 * i + j] = s;
        }
    memcpy(C, T, sizeof T);
}

/* RQ decomposition of P = K [R | t] by three Givens rotations, sign fixes and
 * K(2,2) = 1 scaling: sources/pinholeCamera.cpp:17-94.  P is row-major 3x4.
 * Outputs row-major K, R; t; C = -R^T t. */
void orc_decompose_P(const double *P, double *K, double *R, double *t, double *Cc)
{
    double k[9], Q[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) k[3 * i + j] = P[4 * i + j];
    /* zero K(2,1) */
    if (k[7] != 0) {
        double c = -k[8], s = k[7], l = sqrt(c * c + s * s);
        c /= l; s /= l;
        const double Qx[9] = {1, 0, 0, 0, c, -s, 0, s, c};
        double QxT[9] = {1, 0, 0, 0, c, s, 0, -s, c};
        mat3_mul(k, Qx, k);
        mat3_mul(QxT, Q, Q);
    }
    /* zero K(2,0) */
    if (k[6] != 0) {
        double c = k[8], s = k[6], l = sqrt(c * c + s * s);
        c /= l; s /= l;
        const double Qy[9] = {c, 0, s, 0, 1, 0, -s, 0, c};
        double QyT[9] = {c, 0, -s, 0, 1, 0, s, 0, c};
        mat3_mul(k, Qy, k);
        mat3_mul(QyT, Q, Q);
    }
    /* zero K(1,0) */
    if (k[3] != 0) {
        double c = -k[4], s = k[3], l = sqrt(c * c + s * s);
        c /= l; s /= l;
        const double Qz[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
        double QzT[9] = {c, s, 0, -s, c, 0, 0, 0, 1};
        mat3_mul(k, Qz, k);
        mat3_mul(QzT, Q, Q);
    }
    double r[9];
    memcpy(r, Q, sizeof r);
    if (k[8] < 0) {
        for (int i = 0; i < 9; i++) { k[i] = -k[i]; r[i] = -r[i]; }
    }
    if (k[4] < 0) { /* K = K*S, R = S*R with S = diag(1,-1,1) */
        for (int i = 0; i < 3; i++) k[3 * i + 1] = -k[3 * i + 1];
        for (int j = 0; j < 3; j++) r[3 + j] = -r[3 + j];
    }
    if (k[0] < 0) {
        for (int i = 0; i < 3; i++) k[3 * i + 0] = -k[3 * i + 0];
        for (int j = 0; j < 3; j++) r[j] = -r[j];
    }
    /* t = K^-1 P[:,3] */
    {
        const double a = k[0], b = k[1], c = k[2], d = k[4], e = k[5], f = k[8];
        const double p0 = P[3], p1 = P[7], p2 = P[11];
        const double z = p2 / f, y = (p1 - e * z) / d, x = (p0 - b * y - c * z) / a;
        t[0] = x; t[1] = y; t[2] = z;
    }
    const double det = r[0] * (r[4] * r[8] - r[5] * r[7]) - r[1] * (r[3] * r[8] - r[5] * r[6]) +
                       r[2] * (r[3] * r[7] - r[4] * r[6]);
    if (det < 0) {
        for (int i = 0; i < 9; i++) r[i] = -r[i];
        for (int i = 0; i < 3; i++) t[i] = -t[i];
    }
    const double s22 = k[8];
    for (int i = 0; i < 9; i++) k[i] /= s22;
    memcpy(K, k, sizeof k);
    memcpy(R, r, sizeof r);
    for (int i = 0; i < 3; i++) Cc[i] = -(r[0 * 3 + i] * t[0] + r[1 * 3 + i] * t[1] + r[2 * 3 + i] * t[2]);
}

/* Inverse intrinsics exactly as import.cpp:302-308 builds it (skew ignored). */
static void kinv_ref(const double *K, double *Ki)
{
    memset(Ki, 0, 9 * sizeof(double));
    Ki[0] = 1.0 / K[0];
    Ki[4] = 1.0 / K[4];
    Ki[2] = -K[2] / K[0];
    Ki[5] = -K[5] / K[4];
    Ki[8] = 1.0;
}

/* Warps from depth: sources/import.cpp:243-450.
 * Pass 1: tau = pi(P_u X), X = R_v^T (d K_v^-1 [x+.5, y+.5, 1]) + C_v; invalid
 * depth -> (-w,-h); z-buffer the target depth over a dsf x dsf footprint.
 * Pass 2: footprint-averaged target depth at tau; back-project the target
 * pixel int(tau)+.5 at that depth and reproject into view i; a round trip
 * farther than occl_tol px marks tau = (-2w,-2h).
 * backproject_source_cam = 1 reproduces import.cpp:423-424 literally (the
 * view's R and C are used for the target pixel); 0 uses the target camera.
 * Partial tau (tauPartial.frag:23-47, import.cpp:658-680):
 * dpart = sigma_z * Je * K_u R_u R_v^T K_v^-1 x, 0 where depth is invalid.
 * depth: V planes of w*h; P: V*12 row-major; Pu: 12. */
void orc_warps_from_depth(int V, int w, int h, int dsf, const float *depth, const double *P,
                          const double *Pu, double occl_tol, int backproject_source_cam,
                          double sigma_z, float *tx, float *ty, float *dpx, float *dpy,
                          float *u_depth_out)
{
    const int W = w * dsf, H = h * dsf;
    const size_t n = (size_t)w * h, N = (size_t)W * H;
    double Ku[9], Ru[9], tu[3], Cu[3], Kui[9];
    orc_decompose_P(Pu, Ku, Ru, tu, Cu);
    kinv_ref(Ku, Kui);
    float *ud = (float *)malloc(N * sizeof(float));
    for (size_t p = 0; p < N; p++) ud[p] = (float)ORC_INVALID_DEPTH;

    for (int v = 0; v < V; v++) {
        const double *Pv = P + 12 * v;
        double Kv[9], Rv[9], tv[3], Cv[3], Kvi[9], RvT[9], Pu3[9], M[9], c[3];
        orc_decompose_P(Pv, Kv, Rv, tv, Cv);
        kinv_ref(Kv, Kvi);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                RvT[3 * i + j] = Rv[3 * j + i];
                Pu3[3 * i + j] = Pu[4 * i + j];
            }
        mat3_mul(RvT, Kvi, M);
        mat3_mul(Pu3, M, M); /* M = P_u3 R_v^T K_v^-1 = K_u R_u R_v^T K_v^-1 */
        for (int i = 0; i < 3; i++) c[i] = Pu[4 * i] * Cv[0] + Pu[4 * i + 1] * Cv[1] + Pu[4 * i + 2] * Cv[2] + Pu[4 * i + 3];
        for (size_t m = 0; m < n; m++) {
            const size_t o = (size_t)v * n + m;
            const float d = depth[o];
            const double xx = (double)(m % w) + 0.5, yy = (double)(m / w) + 0.5;
            const double Mx0 = M[0] * xx + M[1] * yy + M[2];
            const double Mx1 = M[3] * xx + M[4] * yy + M[5];
            const double Mx2 = M[6] * xx + M[7] * yy + M[8];
            if (d < 1.0 / ORC_INVALID_DEPTH || ORC_INVALID_DEPTH < d) {
                tx[o] = (float)(-(double)w);
                ty[o] = (float)(-(double)h);
                dpx[o] = 0.0f;
                dpy[o] = 0.0f;
                continue;
            }
            const double X0 = d * Mx0 + c[0], X1 = d * Mx1 + c[1], X2 = d * Mx2 + c[2];
            const double tux = X0 / X2, tuy = X1 / X2;
            tx[o] = (float)tux;
            ty[o] = (float)tuy;
            /* partial tau */
            dpx[o] = (float)(sigma_z * (Mx0 - tux * Mx2) / X2);
            dpy[o] = (float)(sigma_z * (Mx1 - tuy * Mx2) / X2);
            for (int i = 0; i < dsf; i++)
                for (int j = 0; j < dsf; j++) {
                    const int px = (int)tux - dsf / 2 + j;
                    const int py = (int)tuy - dsf / 2 + i;
                    if (0 <= px && px <= W - 1 && 0 <= py && py <= H - 1) {
                        const size_t p = (size_t)py * W + px;
                        if (X2 < ud[p]) ud[p] = (float)X2;
                    }
                }
        }
    }

    for (int v = 0; v < V; v++) {
        const double *Pv = P + 12 * v;
        double Kv[9], Rv[9], tv[3], Cv[3], N3[9], RT[9], Pv3[9], e[3];
        orc_decompose_P(Pv, Kv, Rv, tv, Cv);
        const double *Rb = backproject_source_cam ? Rv : Ru;
        const double *Cb = backproject_source_cam ? Cv : Cu;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                RT[3 * i + j] = Rb[3 * j + i];
                Pv3[3 * i + j] = Pv[4 * i + j];
            }
        mat3_mul(RT, Kui, N3);
        mat3_mul(Pv3, N3, N3);
        for (int i = 0; i < 3; i++) e[i] = Pv[4 * i] * Cb[0] + Pv[4 * i + 1] * Cb[1] + Pv[4 * i + 2] * Cb[2] + Pv[4 * i + 3];
        for (size_t m = 0; m < n; m++) {
            const size_t o = (size_t)v * n + m;
            const double tux = tx[o], tuy = ty[o];
            double u_d = 0.0, u_norm = 0.0;
            for (int i = 0; i < dsf + 1; i++)
                for (int j = 0; j < dsf + 1; j++) {
                    const int px = (int)floor(tux - (double)dsf * 0.5 + 0.5) + j;
                    const int py = (int)floor(tuy - (double)dsf * 0.5 + 0.5) + i;
                    if (0 <= px && px <= W - 1 && 0 <= py && py <= H - 1) {
                        const size_t p = (size_t)py * W + px;
                        if (ud[p] == (float)ORC_INVALID_DEPTH) continue;
                        double dx = 1.0, dy = 1.0;
                        if (j == 0) dx = ((double)px + 0.5) - (tux - 0.5 * dsf);
                        else if (j == dsf) dx = (tux + 0.5 * dsf) - ((double)px - 0.5);
                        if (i == 0) dy = ((double)py + 0.5) - (tuy - 0.5 * dsf);
                        else if (i == dsf) dy = (tuy + 0.5 * dsf) - ((double)py - 0.5);
                        u_d += dx * dy * ud[p];
                        u_norm += dx * dy;
                    }
                }
            if (u_norm == 0.0) {
                tx[o] = (float)(-(double)w);
                ty[o] = (float)(-(double)h);
                continue;
            }
            u_d /= u_norm;
            const double xx = (double)(int)tux + 0.5, yy = (double)(int)tuy + 0.5;
            const double X0 = u_d * (N3[0] * xx + N3[1] * yy + N3[2]) + e[0];
            const double X1 = u_d * (N3[3] * xx + N3[4] * yy + N3[5]) + e[1];
            const double X2 = u_d * (N3[6] * xx + N3[7] * yy + N3[8]) + e[2];
            const int ox = (int)(X0 / X2), oy = (int)(X1 / X2);
            const int mx = (int)(m % w), my = (int)(m / w);
            if (ox - mx < -occl_tol || occl_tol < ox - mx || oy - my < -occl_tol || occl_tol < oy - my) {
                tx[o] = (float)(-(double)w * 2.0);
                ty[o] = (float)(-(double)h * 2.0);
            }
        }
    }
    if (u_depth_out) memcpy(u_depth_out, ud, N * sizeof(float));
    free(ud);
}

/* ------------------------------------------------------------------------- */
/* The OpenGL warp producer: target depth by depth splatting, tau from it    */
/* ------------------------------------------------------------------------- */

#define ORC_GL_ZNEAR 1e-2  /* near plane of the render camera, pinholeCamera.cpp:133 */

/* M = P_u3 R_v^T K_v^-1 and c = P_u (C_v, 1): x_u ~ z M (x, y, 1) + c for a
 * source pixel (x, y) at depth z (tauWarp.frag:30-32, depthSplatting.geom:34-57
 * with depthSplatting.frag:14). */
static void gl_cam(const double *Pv, const double *Pu, double *M, double *c)
{
    double Kv[9], Rv[9], tv[3], Cv[3], Kvi[9], RvT[9], Pu3[9];
    orc_decompose_P(Pv, Kv, Rv, tv, Cv);
    kinv_ref(Kv, Kvi);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            RvT[3 * i + j] = Rv[3 * j + i];
            Pu3[3 * i + j] = Pu[4 * i + j];
        }
    mat3_mul(RvT, Kvi, M);
    mat3_mul(Pu3, M, M);
    for (int i = 0; i < 3; i++) c[i] = Pu[4 * i] * Cv[0] + Pu[4 * i + 1] * Cv[1] + Pu[4 * i + 2] * Cv[2] + Pu[4 * i + 3];
}

/* Edge function of a -> b at p; positive inside a positively oriented triangle. */
static double gl_edge(const double *a, const double *b, double px, double py)
{
    return (b[0] - a[0]) * (py - a[1]) - (b[1] - a[1]) * (px - a[0]);
}

/* A pixel centre on an edge belongs to the triangle that owns the edge: of the
 * two directions a shared edge is walked in, exactly one satisfies this. */
static int gl_owns(const double *a, const double *b)
{
    const double dx = b[0] - a[0], dy = b[1] - a[1];
    return dy > 0.0 || (dy == 0.0 && dx < 0.0);
}

/* Rasterises triangle (a, b, c) (x, y, depth) at pixel centre p: 1 and the
 * depth interpolated as the depth test sees it (1/z linear in the image) when
 * p is inside. */
static int gl_tri(const double *a, const double *b, const double *c, double px, double py, double *zi)
{
    double area = gl_edge(a, b, c[0], c[1]);
    if (area == 0.0) return 0;
    if (area < 0.0) {
        const double *t = b;
        b = c;
        c = t;
        area = -area;
    }
    const double ea = gl_edge(b, c, px, py), eb = gl_edge(c, a, px, py), ec = gl_edge(a, b, px, py);
    if (ea < 0.0 || (ea == 0.0 && !gl_owns(b, c))) return 0;
    if (eb < 0.0 || (eb == 0.0 && !gl_owns(c, a))) return 0;
    if (ec < 0.0 || (ec == 0.0 && !gl_owns(a, b))) return 0;
    const double inv = (ea / area) / a[2] + (eb / area) / b[2] + (ec / area) / c[2];
    *zi = 1.0 / inv;
    return 1;
}


/* Radial to orthographic depth of one view, in place: fromRadial2Ortho.frag
 * (run per view by Scene::filterDepthMaps, scene.cpp:1605-1624, through
 * InputView::filterDepthmap, view.cpp:723-733).  Kinv is the row-major
 * inverse of the view's K; the fragment of pixel (x, y) is its centre, with
 * u = Kinv (x + 0.5, y + 0.5, 1) and d = |u|:
 *   z <- z / d,  dz/dx <- dz/dx / d - z (u . Kinv e_x) / d^3  (e_y for dz/dy).
 * dzdx / dzdy may be NULL. */
void orc_radial_to_ortho(int W, int H, const double *Kinv, float *z, float *dzdx, float *dzdy)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const double px = x + 0.5, py = y + 0.5;
            double u[3];
            for (int r = 0; r < 3; r++) u[r] = Kinv[3 * r] * px + Kinv[3 * r + 1] * py + Kinv[3 * r + 2];
            const double d = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            const double cx = u[0] * Kinv[0] + u[1] * Kinv[3] + u[2] * Kinv[6];
            const double cy = u[0] * Kinv[1] + u[1] * Kinv[4] + u[2] * Kinv[7];
            const size_t o = (size_t)y * W + x;
            const double zr = z[o];
            z[o] = (float)(zr / d);
            if (dzdx) dzdx[o] = (float)(dzdx[o] / d - zr * cx / (d * d * d));
            if (dzdy) dzdy[o] = (float)(dzdy[o] / d - zr * cy / (d * d * d));
        }
}

/* Scene::computeTargetDepthMap (scene.cpp:1506-1544; depthSplatting.vert,
 * .geom, .frag): every view but `skip` is drawn as one quad per pixel of valid
 * depth (z < INVALID_DEPTH / 2), corners at the pixel centre +- 0.5 with depths
 * z +- 0.5 dz/dx +- 0.5 dz/dy, as the triangle strip BL, TL, BR, TR.  At each
 * covered target pixel centre the nearest quad wins the depth test and writes
 * the target depth of its centre; uncovered pixels keep the cleared 0
 * (glClearColor 0, scene.cpp:1515-1517).  dzdx / dzdy may be NULL (zero). */
void orc_gl_target_depth(int V, int w, int h, int W, int H, const float *depth, const float *dzdx,
                         const float *dzdy, const double *P, const double *Pu, int skip, float *out)
{
    const size_t n = (size_t)w * h, N = (size_t)W * H;
    double *zb = (double *)malloc(N * sizeof(double)), *val = (double *)malloc(N * sizeof(double));
    for (size_t p = 0; p < N; p++) {
        zb[p] = INFINITY;
        val[p] = 0.0;
    }
    for (int v = 0; v < V; v++) {
        if (v == skip) continue;
        double M[9], c[3];
        gl_cam(P + 12 * v, Pu, M, c);
        for (size_t m = 0; m < n; m++) {
            const size_t o = (size_t)v * n + m;
            const double z = depth[o];
            if (z >= 0.5 * ORC_INVALID_DEPTH) continue;
            const double gx = dzdx ? dzdx[o] : 0.0, gy = dzdy ? dzdy[o] : 0.0;
            const double x = (double)(m % w) + 0.5, y = (double)(m / w) + 0.5;
            /* BL, TL, BR, TR */
            const double sx[4] = {-0.5, -0.5, 0.5, 0.5}, sy[4] = {-0.5, 0.5, -0.5, 0.5};
            double q[4][3];
            int ok = 1;
            for (int k = 0; k < 4; k++) {
                const double zc = z + sx[k] * gx + sy[k] * gy, xc = x + sx[k], yc = y + sy[k];
                double hh[3];
                for (int i = 0; i < 3; i++) hh[i] = zc * (M[3 * i] * xc + M[3 * i + 1] * yc + M[3 * i + 2]) + c[i];
                if (!(hh[2] > ORC_GL_ZNEAR)) ok = 0;
                q[k][0] = hh[0] / hh[2];
                q[k][1] = hh[1] / hh[2];
                q[k][2] = hh[2];
            }
            if (!ok) continue;
            const double zval = z * (M[6] * x + M[7] * y + M[8]) + c[2];
            double u0 = q[0][0], u1 = q[0][0], v0 = q[0][1], v1 = q[0][1];
            for (int k = 1; k < 4; k++) {
                u0 = fmin(u0, q[k][0]);
                u1 = fmax(u1, q[k][0]);
                v0 = fmin(v0, q[k][1]);
                v1 = fmax(v1, q[k][1]);
            }
            if (!(isfinite(u0) && isfinite(u1) && isfinite(v0) && isfinite(v1))) continue;
            /* every quad is rasterised, over its box clipped to the target (GL
             * clips primitives to the viewport, not by size) */
            if (u1 < 0.0 || v1 < 0.0 || u0 > W || v0 > H) continue;
            const int px0 = (int)fmax(ceil(u0 - 0.5), 0.0), px1 = (int)fmin(floor(u1 - 0.5), W - 1.0);
            const int py0 = (int)fmax(ceil(v0 - 0.5), 0.0), py1 = (int)fmin(floor(v1 - 0.5), H - 1.0);
            for (int py = py0; py <= py1; py++)
                for (int px = px0; px <= px1; px++) {
                    double zi;
                    const double cx = px + 0.5, cy = py + 0.5;
                    if (!gl_tri(q[0], q[1], q[2], cx, cy, &zi) && !gl_tri(q[2], q[1], q[3], cx, cy, &zi)) continue;
                    const size_t p = (size_t)py * W + px;
                    if (zi < zb[p] || (zi == zb[p] && zval < val[p])) {
                        zb[p] = zi;
                        val[p] = zval;
                    }
                }
        }
    }
    for (size_t p = 0; p < N; p++) out[p] = zb[p] < INFINITY ? (float)val[p] : 0.0f;
    free(zb);
    free(val);
}

/* InputView::exportWarp's two passes (view.cpp:663-721): tauWarp.frag:25-74
 * and tauPartial.frag:23-47 for every pixel of every view against the target
 * depth map ud.  tau = x_u, or (-2, -2) outside (0, W) x (0, H) or at invalid
 * depth, or (-1, -1) when the target depth map, sampled bilinearly with
 * clamped edges at x_u, is nearer than the point by more than 0.1; wdef the
 * deformation weight 1 / |det d tau / d x| capped at 30 (0 when invalid);
 * dpart = sigma_z d tau / dz. */
void orc_gl_warps(int V, int w, int h, int W, int H, const float *depth, const float *dzdx, const float *dzdy,
                  const double *P, const double *Pu, const float *ud, double sigma_z, float *tx, float *ty,
                  float *wdef, float *dpx, float *dpy)
{
    const size_t n = (size_t)w * h;
    for (int v = 0; v < V; v++) {
        double M[9], c[3];
        gl_cam(P + 12 * v, Pu, M, c);
        for (size_t m = 0; m < n; m++) {
            const size_t o = (size_t)v * n + m;
            const double z = depth[o];
            const double gx = dzdx ? dzdx[o] : 0.0, gy = dzdy ? dzdy[o] : 0.0;
            const double x = (double)(m % w) + 0.5, y = (double)(m / w) + 0.5;
            double Mx[3], hh[3];
            for (int i = 0; i < 3; i++) {
                Mx[i] = M[3 * i] * x + M[3 * i + 1] * y + M[3 * i + 2];
                hh[i] = z * Mx[i] + c[i];
            }
            double u = 0.0, vv = 0.0;
            if (hh[2] != 0.0) {
                u = hh[0] / hh[2];
                vv = hh[1] / hh[2];
            }
            dpx[o] = (float)(sigma_z * (Mx[0] - u * Mx[2]) / hh[2]);
            dpy[o] = (float)(sigma_z * (Mx[1] - vv * Mx[2]) / hh[2]);
            /* d tau / d x = Je M temp, temp's columns d(z x_k) / dx and / dy */
            const double t0[3] = {x * gx + z, y * gx, gx}, t1[3] = {x * gy, y * gy + z, gy};
            double a0[3], a1[3];
            for (int i = 0; i < 3; i++) {
                a0[i] = M[3 * i] * t0[0] + M[3 * i + 1] * t0[1] + M[3 * i + 2] * t0[2];
                a1[i] = M[3 * i] * t1[0] + M[3 * i + 1] * t1[1] + M[3 * i + 2] * t1[2];
            }
            const double j00 = (a0[0] - u * a0[2]) / hh[2], j10 = (a0[1] - vv * a0[2]) / hh[2];
            const double j01 = (a1[0] - u * a1[2]) / hh[2], j11 = (a1[1] - vv * a1[2]) / hh[2];
            double wd = 1.0 / fabs(j00 * j11 - j10 * j01);
            if (wd >= 30.0) wd = 30.0;
            if (u <= 0.0 || W <= u || vv <= 0.0 || H <= vv || z >= 0.5 * ORC_INVALID_DEPTH) {
                tx[o] = ty[o] = -2.0f;
                wdef[o] = 0.0f;
                continue;
            }
            /* GL_LINEAR with GL_CLAMP_TO_EDGE (texture.cpp:182-185): texel centres at +0.5 */
            const double s = u - 0.5, t = vv - 0.5;
            const int i0 = (int)floor(s), j0 = (int)floor(t);
            const double fx = s - i0, fy = t - j0;
            const int xa = i0 < 0 ? 0 : (i0 > W - 1 ? W - 1 : i0), xb = i0 + 1 < 0 ? 0 : (i0 + 1 > W - 1 ? W - 1 : i0 + 1);
            const int ya = j0 < 0 ? 0 : (j0 > H - 1 ? H - 1 : j0), yb = j0 + 1 < 0 ? 0 : (j0 + 1 > H - 1 ? H - 1 : j0 + 1);
            const double du = (1.0 - fy) * ((1.0 - fx) * ud[(size_t)ya * W + xa] + fx * ud[(size_t)ya * W + xb]) +
                              fy * ((1.0 - fx) * ud[(size_t)yb * W + xa] + fx * ud[(size_t)yb * W + xb]);
            if (hh[2] > du + 0.1) {
                tx[o] = ty[o] = -1.0f;
                wdef[o] = 0.0f;
                continue;
            }
            tx[o] = (float)u;
            ty[o] = (float)vv;
            wdef[o] = (float)wd;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Gradient-domain IBR (IBR_gradient)                                        */
/* ------------------------------------------------------------------------- */

/* Weighted initial image sum_i A^T (w_i v_i) / sum_i A^T w_i, zero-weight
 * pixels left at 0: gradientIBR.cu:764-811 with cuda_kernels.cu:322-337. */
void orc_initial_image(int V, int W, int H, int w, int h, int dsf, const float *tx, const float *ty,
                       const float *views, const double *omega, double *U, double *norm)
{
    const size_t N = (size_t)W * H, n = (size_t)w * h;
    memset(U, 0, 3 * N * sizeof(double));
    memset(norm, 0, N * sizeof(double));
    double *vin = (double *)malloc(3 * n * sizeof(double));
    for (int v = 0; v < V; v++) {
        for (size_t k = 0; k < 3 * n; k++) vin[k] = views[(size_t)v * 3 * n + k];
        orc_splat(W, H, w, h, dsf, tx + v * n, ty + v * n, vin, 3, omega + v * n, U, norm);
    }
    for (size_t o = 0; o < N; o++)
        if (norm[o] != 0.0)
            for (int c = 0; c < 3; c++) U[c * N + o] /= norm[o];
    free(vin);
}

/* Poisson diffusion in the holes: gradientIBR.cu:87-152.  Up to `iters`
 * Jacobi sweeps; after each sweep the energy sum(lap * u) of the last channel
 * is compared with the previous sweep's and the loop stops when the change is
 * below eps.  Returns the number of sweeps run. */
int orc_hole_filling(int W, int H, double *U, const unsigned char *vis, int iters, double eps)
{
    const size_t N = (size_t)W * H;
    double *lap = (double *)malloc(N * sizeof(double));
    double energy = 0.0, prev = 0.0;
    int k;
    for (k = 0; k < iters; k++) {
        prev = energy;
        for (int c = 0; c < 3; c++) {
            double *u = U + c * N;
            orc_laplacian(W, H, u, lap, vis);
            double e = 0.0;
            for (size_t o = 0; o < N; o++) {
                u[o] += lap[o];
                e += lap[o] * u[o];
            }
            energy = e;
        }
        if (fabs(prev - energy) < eps) {
            k++;
            break;
        }
    }
    free(lap);
    return k;
}

/* Target gradient g = sum_i A^T Sobel(Y(v_i)) / sum_i A^T 1 and target mask
 * sigma^2 * sum_i A^T omega_i / sum_i A^T 1: gradientIBR.cu:833-920.  The
 * splat is weighted by A only (deconvolution_nonsep_param with the omega
 * plane as third channel). */
void orc_u_gradient(int V, int W, int H, int w, int h, int dsf, const float *tx, const float *ty,
                    const float *views, const double *omega, double sigma, double *gxo, double *gyo,
                    double *mask, double *norm)
{
    const size_t N = (size_t)W * H, n = (size_t)w * h;
    double *vv = (double *)malloc(3 * n * sizeof(double));
    double *in3 = (double *)malloc(3 * n * sizeof(double));
    double *out3 = (double *)calloc(3 * N, sizeof(double));
    memset(norm, 0, N * sizeof(double));
    for (int v = 0; v < V; v++) {
        for (size_t k = 0; k < 3 * n; k++) vv[k] = views[(size_t)v * 3 * n + k];
        orc_sobel(w, h, vv, vv + n, vv + 2 * n, in3, in3 + n, NULL);
        memcpy(in3 + 2 * n, omega + v * n, n * sizeof(double));
        orc_splat(W, H, w, h, dsf, tx + v * n, ty + v * n, in3, 3, NULL, out3, norm);
    }
    for (size_t o = 0; o < N; o++) {
        double a = out3[o], b = out3[N + o], c = out3[2 * N + o];
        if (norm[o] != 0.0) { a /= norm[o]; b /= norm[o]; c /= norm[o]; }
        gxo[o] = a;
        gyo[o] = b;
        mask[o] = c * sigma * sigma;
    }
    free(vv);
    free(in3);
    free(out3);
}

/* Normal-equation operator of the gradient-constraint solve,
 *   q = alpha * sum_i S_i^T W_i S_i d + c0 * d + beta * D^T D d,
 * with S_i the footprint warp restricted to cell-valid pixels and
 * W_i = omega_i * sigma^2 (optimization.cu:1835-1966 data term, solver.cu:201-310
 * gradient term).  d, q: one channel of N.  Per-thread accumulators make the
 * scatter parallel; their reduction order is fixed. */
void orc_apply_A(int V, int W, int H, int w, int h, int dsf, const float *tx, const float *ty,
                 const double *wdata, double alpha, double c0, double beta, const double *d, double *q)
{
    const size_t N = (size_t)W * H, n = (size_t)w * h;
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    double *acc = (double *)calloc((size_t)nt * N, sizeof(double));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *a = acc + (size_t)tid * N;
#pragma omp for schedule(static)
        for (long k = 0; k < (long)((size_t)V * n); k++) {
            const size_t m = (size_t)k;
            if (alpha == 0.0 || wdata[m] == 0.0) continue;
            if (!orc_cell_valid(tx[m], ty[m], W, H)) continue;
            int x0, y0;
            double wx[ORC_MAXKS], wy[ORC_MAXKS];
            const int ks = orc_footprint(tx[m], ty[m], W, H, dsf, &x0, &y0, wx, wy);
            double s = 0.0;
            for (int i = 0; i < ks; i++) {
                const int y = y0 + i;
                if (y < 0 || y > H - 1) continue;
                for (int j = 0; j < ks; j++) {
                    const int x = x0 + j;
                    if (x < 0 || x > W - 1) continue;
                    s += wy[i] * wx[j] * d[(size_t)y * W + x];
                }
            }
            s *= wdata[m] * alpha;
            for (int i = 0; i < ks; i++) {
                const int y = y0 + i;
                if (y < 0 || y > H - 1) continue;
                for (int j = 0; j < ks; j++) {
                    const int x = x0 + j;
                    if (x < 0 || x > W - 1) continue;
                    a[(size_t)y * W + x] += wy[i] * wx[j] * s;
                }
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (long oy = 0; oy < H; oy++)
        for (long ox = 0; ox < W; ox++) {
            const size_t o = (size_t)oy * W + ox;
            double s = 0.0;
            for (int t = 0; t < nt; t++) s += acc[(size_t)t * N + o];
            s += c0 * d[o];
            if (beta != 0.0) s += beta * dtd_at(d, W, H, (int)ox, (int)oy);
            q[o] = s;
        }
    free(acc);
}

/* Jacobi preconditioner diag(A) = alpha sum_i sum_k W_i A_k^2 + c0 + beta deg(x). */
void orc_diag_A(int V, int W, int H, int w, int h, int dsf, const float *tx, const float *ty,
                const double *wdata, double alpha, double c0, double beta, double *diag)
{
    const size_t N = (size_t)W * H, n = (size_t)w * h;
    memset(diag, 0, N * sizeof(double));
    for (size_t m = 0; m < (size_t)V * n; m++) {
        if (alpha == 0.0 || wdata[m] == 0.0) continue;
        if (!orc_cell_valid(tx[m], ty[m], W, H)) continue;
        int x0, y0;
        double wx[ORC_MAXKS], wy[ORC_MAXKS];
        const int ks = orc_footprint(tx[m], ty[m], W, H, dsf, &x0, &y0, wx, wy);
        for (int i = 0; i < ks; i++) {
            const int y = y0 + i;
            if (y < 0 || y > H - 1) continue;
            for (int j = 0; j < ks; j++) {
                const int x = x0 + j;
                if (x < 0 || x > W - 1) continue;
                const double a = wy[i] * wx[j];
                diag[(size_t)y * W + x] += alpha * wdata[m] * a * a;
            }
        }
    }
    for (long oy = 0; oy < H; oy++)
        for (long ox = 0; ox < W; ox++) {
            const int deg = (W > 1 ? (ox == 0 || ox == W - 1 ? 1 : 2) : 0) +
                            (H > 1 ? (oy == 0 || oy == H - 1 ? 1 : 2) : 0);
            diag[(size_t)oy * W + ox] += c0 + beta * deg;
        }
}

/* Right-hand side b = alpha sum_i S_i^T W_i v_i + c0 u0 + beta D^T g (all three channels,
 * the same luminance g for every channel as in solver.cu:387-425). */
void orc_rhs(int V, int W, int H, int w, int h, int dsf, const float *tx, const float *ty,
             const float *views, const double *wdata, double alpha, double c0, const double *u0,
             double beta, const double *gxv, const double *gyv, double *b)
{
    const size_t N = (size_t)W * H, n = (size_t)w * h;
    memset(b, 0, 3 * N * sizeof(double));
    double *vin = (double *)malloc(3 * n * sizeof(double));
    if (alpha != 0.0)
        for (int v = 0; v < V; v++) {
            for (size_t k = 0; k < 3 * n; k++) vin[k] = views[(size_t)v * 3 * n + k];
            double *wv = (double *)malloc(n * sizeof(double));
            for (size_t m = 0; m < n; m++) wv[m] = alpha * wdata[(size_t)v * n + m];
            orc_splat(W, H, w, h, dsf, tx + v * n, ty + v * n, vin, 3, wv, b, NULL);
            free(wv);
        }
    for (int c = 0; c < 3; c++)
        for (long oy = 0; oy < H; oy++)
            for (long ox = 0; ox < W; ox++) {
                const size_t o = (size_t)oy * W + ox;
                double s = 0.0;
                if (c0 != 0.0 && u0) s += c0 * u0[c * N + o];
                if (beta != 0.0) s += beta * dtg_at(gxv, gyv, W, H, (int)ox, (int)oy);
                b[c * N + o] += s;
            }
    free(vin);
}

typedef struct {
    int iterations[3];
    int converged[3];
    double rel_residual[3];   /* recursive residual ||r|| / ||b|| at exit */
    double true_rel_residual[3]; /* ||b - A x|| / ||b|| recomputed at exit */
    int status;               /* 0 ok, 1 not converged, 2 breakdown */
} orc_cg_stats;

/* Jacobi-preconditioned CG per RGB channel in FP64 (the intended solver.cu:526-699,
 * with its reduction and scalar bugs corrected).  Stops a channel when
 * ||r|| <= tol * ||b||.  X holds x0 on entry and the solution on exit.
 * use_precond == 0 runs plain CG. */
int orc_pcg(int V, int W, int H, int w, int h, int dsf, const float *tx, const float *ty,
            const double *wdata, double alpha, double c0, double beta, const double *B, double *X,
            double tol, int maxit, int use_precond, orc_cg_stats *st)
{
    const size_t N = (size_t)W * H;
    double *r = (double *)malloc(N * sizeof(double)), *z = (double *)malloc(N * sizeof(double));
    double *d = (double *)malloc(N * sizeof(double)), *q = (double *)malloc(N * sizeof(double));
    double *minv = (double *)malloc(N * sizeof(double));
    orc_diag_A(V, W, H, w, h, dsf, tx, ty, wdata, alpha, c0, beta, minv);
    for (size_t o = 0; o < N; o++) minv[o] = (use_precond && minv[o] > 0.0) ? 1.0 / minv[o] : 1.0;
    int status = 0;
    for (int c = 0; c < 3; c++) {
        double *x = X + c * N;
        const double *bb = B + c * N;
        orc_apply_A(V, W, H, w, h, dsf, tx, ty, wdata, alpha, c0, beta, x, q);
        double bnorm2 = 0.0, rr = 0.0, rz = 0.0;
        for (size_t o = 0; o < N; o++) {
            r[o] = bb[o] - q[o];
            z[o] = minv[o] * r[o];
            d[o] = z[o];
            bnorm2 += bb[o] * bb[o];
            rr += r[o] * r[o];
            rz += r[o] * z[o];
        }
        const double thr2 = tol * tol * bnorm2;
        int it = 0, conv = (rr <= thr2);
        while (!conv && it < maxit) {
            orc_apply_A(V, W, H, w, h, dsf, tx, ty, wdata, alpha, c0, beta, d, q);
            double dq = 0.0;
            for (size_t o = 0; o < N; o++) dq += d[o] * q[o];
            if (!(dq > 0.0)) { status = 2; break; }
            const double a = rz / dq;
            rr = 0.0;
            for (size_t o = 0; o < N; o++) {
                x[o] += a * d[o];
                r[o] -= a * q[o];
                rr += r[o] * r[o];
            }
            it++;
            if (rr <= thr2) { conv = 1; break; }
            double rzn = 0.0;
            for (size_t o = 0; o < N; o++) {
                z[o] = minv[o] * r[o];
                rzn += r[o] * z[o];
            }
            const double bt = rzn / rz;
            rz = rzn;
            for (size_t o = 0; o < N; o++) d[o] = z[o] + bt * d[o];
        }
        if (!conv && status == 0) status = 1;
        if (st) {
            st->iterations[c] = it;
            st->converged[c] = conv;
            st->rel_residual[c] = bnorm2 > 0 ? sqrt(rr / bnorm2) : sqrt(rr);
            orc_apply_A(V, W, H, w, h, dsf, tx, ty, wdata, alpha, c0, beta, x, q);
            double tr = 0.0;
            for (size_t o = 0; o < N; o++) tr += (bb[o] - q[o]) * (bb[o] - q[o]);
            st->true_rel_residual[c] = bnorm2 > 0 ? sqrt(tr / bnorm2) : sqrt(tr);
        }
    }
    if (st) st->status = status;
    free(r); free(z); free(d); free(q); free(minv);
    return status;
}

/* Quadratic objective whose normal equations orc_pcg solves, per channel summed:
 * E = alpha/2 sum_i ||W_i^1/2 (S_i u - v_i)||^2 + c0/2 ||u - u0||^2 + beta/2 ||D u - g||^2.
 * S_i uses the same cell-valid restriction as orc_apply_A. */
double orc_gibr_energy(int V, int W, int H, int w, int h, int dsf, const float *tx, const float *ty,
                       const float *views, const double *wdata, double alpha, double c0,
                       const double *u0, double beta, const double *gxv, const double *gyv,
                       const double *U)
{
    const size_t N = (size_t)W * H, n = (size_t)w * h;
    double e_data = 0.0, e_grad = 0.0, e_prox = 0.0;
    for (int v = 0; v < V; v++)
        for (size_t m = 0; m < n; m++) {
            const size_t k = (size_t)v * n + m;
            if (wdata[k] == 0.0 || !orc_cell_valid(tx[k], ty[k], W, H)) continue;
            int x0, y0;
            double wx[ORC_MAXKS], wy[ORC_MAXKS];
            const int ks = orc_footprint(tx[k], ty[k], W, H, dsf, &x0, &y0, wx, wy);
            for (int c = 0; c < 3; c++) {
                double s = 0.0;
                for (int i = 0; i < ks; i++) {
                    const int y = y0 + i;
                    if (y < 0 || y > H - 1) continue;
                    for (int j = 0; j < ks; j++) {
                        const int x = x0 + j;
                        if (x < 0 || x > W - 1) continue;
                        s += wy[i] * wx[j] * U[c * N + (size_t)y * W + x];
                    }
                }
                const double rr = s - views[(size_t)v * 3 * n + c * n + m];
                e_data += wdata[k] * rr * rr;
            }
        }
    for (int c = 0; c < 3; c++) {
        const double *u = U + c * N;
        for (long oy = 0; oy < H; oy++)
            for (long ox = 0; ox < W; ox++) {
                const size_t o = (size_t)oy * W + ox;
                if (ox < W - 1) { const double t = (u[o + 1] - u[o]) - gxv[o]; e_grad += t * t; }
                if (oy < H - 1) { const double t = (u[o + W] - u[o]) - gyv[o]; e_grad += t * t; }
                if (c0 != 0.0 && u0) { const double t = u[o] - u0[c * N + o]; e_prox += t * t; }
            }
    }
    return 0.5 * (alpha * e_data + c0 * e_prox + beta * e_grad);
}

/* ------------------------------------------------------------------------- */
/* Colour IBR with gradient constraints (IBR_color): TV primal-dual          */
/* ------------------------------------------------------------------------- */

/* p <- p + sigma grad+ u (forward differences, 0 on the last column/row):
 * vtv_rof.cu:118-147; then joint RGB Frobenius projection onto the ball of
 * radius g (vtv_rof.cu:280-321) or radius 1 when g is NULL (vtv_rof.cu:237-275).
 * U: 3 planes; PX, PY: 3 planes each. */
void orc_pd_dual(int W, int H, double sigma, const double *U, double *PX, double *PY, const double *g)
{
    const size_t N = (size_t)W * H;
    for (long oy = 0; oy < H; oy++)
        for (long ox = 0; ox < W; ox++) {
            const size_t o = (size_t)oy * W + ox;
            double a[3], b[3], nn = 0.0;
            for (int c = 0; c < 3; c++) {
                const double *u = U + c * N;
                const double gx = (ox < W - 1) ? u[o + 1] - u[o] : 0.0;
                const double gy = (oy < H - 1) ? u[o + W] - u[o] : 0.0;
                a[c] = PX[c * N + o] + sigma * gx;
                b[c] = PY[c * N + o] + sigma * gy;
                nn += a[c] * a[c] + b[c] * b[c];
            }
            nn = sqrt(nn);
            const double rad = g ? g[o] : 1.0;
            const double s = (nn > rad) ? rad / nn : 1.0;
            for (int c = 0; c < 3; c++) {
                PX[c * N + o] = (nn > rad) ? a[c] * s : a[c];
                PY[c * N + o] = (nn > rad) ? b[c] * s : b[c];
            }
        }
}

/* u <- (tau f + lambda (u + tau div- p)) / (tau + lambda), backward differences
 * with the x=0 / y=0 terms dropped: cuda_kernels.cu:504-532. One channel. */
void orc_pd_primal(int W, int H, double tau, double lambda, double *u, const double *f,
                   const double *px, const double *py)
{
    const size_t N = (size_t)W * H;
    double *un = (double *)malloc(N * sizeof(double));
    for (long oy = 0; oy < H; oy++)
        for (long ox = 0; ox < W; ox++) {
            const size_t o = (size_t)oy * W + ox;
            double step = px[o] + py[o];
            if (ox > 0) step -= px[o - 1];
            if (oy > 0) step -= py[o - W];
            const double unew = u[o] + tau * step;
            un[o] = (tau * f[o] + lambda * unew) / (tau + lambda);
        }
    memcpy(u, un, N * sizeof(double));
    free(un);
}

/* Separable 11-tap Gaussian, sigma 2, clamp-to-edge, rows then columns:
 * gsl_matrix_convolutions.cpp:42-63 (taps) and cuda_convolutions.cu:110-237,310-359. */
void orc_gauss_smooth(int W, int H, int taps, double sg, const double *in, double *out)
{
    const int o2 = taps / 2;
    double k[64], s = 0.0;
    for (int x = -o2; x <= o2; x++) { k[x + o2] = exp(-0.5 * x * x / (sg * sg)); s += k[x + o2]; }
    for (int i = 0; i < taps; i++) k[i] /= s;
    const size_t N = (size_t)W * H;
    double *tmp = (double *)malloc(N * sizeof(double));
    for (long y = 0; y < H; y++)
        for (long x = 0; x < W; x++) {
            double a = 0.0;
            for (int t = -o2; t <= o2; t++) {
                long xx = x + t;
                if (xx < 0) xx = 0;
                if (xx > W - 1) xx = W - 1;
                a += in[(size_t)y * W + xx] * k[o2 - t];
            }
            tmp[(size_t)y * W + x] = a;
        }
    for (long y = 0; y < H; y++)
        for (long x = 0; x < W; x++) {
            double a = 0.0;
            for (int t = -o2; t <= o2; t++) {
                long yy = y + t;
                if (yy < 0) yy = 0;
                if (yy > H - 1) yy = H - 1;
                a += tmp[(size_t)yy * W + x] * k[o2 - t];
            }
            out[(size_t)y * W + x] = a;
        }
    free(tmp);
}

/* TV weight g(x): optimization.cu:2251-2338 with vtv_sr_kernels.cu:1172-1194.
 * mask = sigma^2 * (sum_i A^T omega_i) / norm_mask where norm_mask is the
 * caller's accumulator *plus* sum_i A^T 1 (the reference does not clear it,
 * optimization.cu:2264-2305); then g = lmax where mask == 0, else
 * max(lmin, lmax - mask/V); then the 11-tap sigma-2 Gaussian.
 * norm_mask is updated in place like the reference's buffer. */
void orc_regularizer_weight(int V, int W, int H, int w, int h, int dsf, const float *tx,
                            const float *ty, const double *omega, double sigma, double lmax,
                            double lmin, double *norm_mask, double *g)
{
    const size_t N = (size_t)W * H, n = (size_t)w * h;
    double *acc = (double *)calloc(N, sizeof(double));
    for (int v = 0; v < V; v++)
        orc_splat(W, H, w, h, dsf, tx + v * n, ty + v * n, omega + v * n, 1, NULL, acc, norm_mask);
    double *gw = (double *)malloc(N * sizeof(double));
    for (size_t o = 0; o < N; o++) {
        double m = acc[o];
        if (norm_mask[o] != 0.0) m /= norm_mask[o];
        m *= sigma * sigma;
        const float mf = (float)m;
        if (mf == 0.0f) gw[o] = lmax;
        else gw[o] = fmax(lmin, lmax - m / (double)V);
    }
    orc_gauss_smooth(W, H, 11, 2.0, gw, g);
    free(acc);
    free(gw);
}

/* One proximal-gradient step of IBR_color: optimization.cu:1736-1832
 * (vtv_sr_ista_step_unstructured) with the data terms of :1835-1966 and
 * :1969-2116, followed by the over-relaxation of vtv_rof.cu:28-57.
 *   G = alpha*sigma^2*G_int/cov + beta*(-step)*(sigma^2)*G_grad/cov, clamped to [-1,1]
 *   F = clamp(U - G/(lambda L), 0, 1), L = 1/lambda
 *   Uq <- U; niter x { dual step + TV_F projection; primal prox }
 *   theta = (a_k - 1)/a_{k+1}; Uq <- U + theta (Uq - U); U <- Uq_old
 * The reference's FISTA extrapolation is overwritten by Uq <- U at the next
 * step (the build keeps that).  vis gates nothing here.  norm_mask is the
 * reference's shared accumulator; on exit it holds the coverage sum_i A^T 1.
 * Uq, PX, PY (3 planes each) and *fista_alpha are state. */
void orc_fista_step(int V, int W, int H, int w, int h, int dsf, const float *tx, const float *ty,
                    const float *views, const double *omega, double sigma, double dt_alpha,
                    double dt_beta, double gradient_step, int gw_type, double lambda,
                    int inner_iters, const double *g, double *U, double *Uq, double *PX, double *PY,
                    double *norm_mask, double *fista_alpha, double *Gout)
{
    const size_t N = (size_t)W * H, n = (size_t)w * h;
    double *G = (double *)calloc(3 * N, sizeof(double));
    double *tmp = (double *)malloc(3 * n * sizeof(double));
    double *acc = (double *)calloc(3 * N, sizeof(double));
    double *lapU = (double *)malloc(3 * N * sizeof(double));
    double *lv = (double *)malloc(n * sizeof(double));
    double *vv = (double *)malloc(n * sizeof(double));
    memcpy(Uq, U, 3 * N * sizeof(double));
    if (dt_alpha > 0.0) {
        memset(norm_mask, 0, N * sizeof(double));
        memset(acc, 0, 3 * N * sizeof(double));
        for (int v = 0; v < V; v++) {
            orc_gather(W, H, w, h, dsf, tx + v * n, ty + v * n, U, 3, tmp);
            for (int c = 0; c < 3; c++)
                for (size_t m = 0; m < n; m++)
                    tmp[c * n + m] = (tmp[c * n + m] - views[(size_t)v * 3 * n + c * n + m]) * omega[v * n + m];
            orc_splat(W, H, w, h, dsf, tx + v * n, ty + v * n, tmp, 3, NULL, acc, norm_mask);
        }
        for (int c = 0; c < 3; c++)
            for (size_t o = 0; o < N; o++) {
                double a = acc[c * N + o];
                if (norm_mask[o] != 0.0) a /= norm_mask[o];
                G[c * N + o] += dt_alpha * (a * sigma * sigma);
            }
    }
    if (dt_beta > 0.0) {
        memset(norm_mask, 0, N * sizeof(double));
        memset(acc, 0, 3 * N * sizeof(double));
        for (int c = 0; c < 3; c++) orc_laplacian(W, H, U + c * N, lapU + c * N, NULL);
        for (int v = 0; v < V; v++) {
            orc_gather(W, H, w, h, dsf, tx + v * n, ty + v * n, lapU, 3, tmp);
            for (int c = 0; c < 3; c++) {
                for (size_t m = 0; m < n; m++) vv[m] = views[(size_t)v * 3 * n + c * n + m];
                orc_laplacian(w, h, vv, lv, NULL);
                for (size_t m = 0; m < n; m++) {
                    double r = tmp[c * n + m] * (double)(dsf * dsf) - lv[m];
                    if (gw_type != 0) r *= omega[v * n + m];
                    tmp[c * n + m] = r;
                }
            }
            orc_splat(W, H, w, h, dsf, tx + v * n, ty + v * n, tmp, 3, NULL, acc, norm_mask);
        }
        for (int c = 0; c < 3; c++)
            for (size_t o = 0; o < N; o++) {
                double a = acc[c * N + o];
                if (norm_mask[o] != 0.0) a /= norm_mask[o];
                if (gw_type != 0) a *= sigma * sigma;
                a *= -gradient_step;
                G[c * N + o] += dt_beta * a;
            }
    }
    const double L = 1.0 / lambda;
    for (size_t k = 0; k < 3 * N; k++) {
        double gg = G[k];
        if (gg < -1.0) gg = -1.0;
        if (gg > 1.0) gg = 1.0;
        gg *= -1.0 / (lambda * L);
        gg += Uq[k];
        if (gg < 0.0) gg = 0.0;
        if (gg > 1.0) gg = 1.0;
        G[k] = gg;
    }
    const double tau = 0.3 / sqrt(8.0), sg = 0.3 / sqrt(8.0);
    for (int k = 0; k < inner_iters; k++) {
        orc_pd_dual(W, H, sg, Uq, PX, PY, g);
        for (int c = 0; c < 3; c++) orc_pd_primal(W, H, tau, 1.0 / L, Uq + c * N, G + c * N, PX + c * N, PY + c * N);
    }
    const double an = 0.5 * (1.0 + sqrt(1.0 + 4.0 * (*fista_alpha) * (*fista_alpha)));
    const double theta = (*fista_alpha - 1.0) / an;
    for (size_t k = 0; k < 3 * N; k++) {
        const double uv = U[k], uqv = Uq[k];
        Uq[k] = uv + theta * (uqv - uv);
        U[k] = uqv;
    }
    *fista_alpha = an;
    if (Gout) memcpy(Gout, G, 3 * N * sizeof(double));
    free(G); free(tmp); free(acc); free(lapU); free(lv); free(vv);
}

/* Primal energy of IBR_color: optimization.cu:1477-1717 in FP64.
 *   E_TV  = sum |Sobel_x Y(U)| + |Sobel_y Y(U)| over pixels whose 3x3 neighbourhood is visible
 *   E_int = sum_i sum_c sum_m omega (S_i U - v_i)^2 sigma^2   (gather-valid, S U = 0 elsewhere)
 *   E_grad= sum_i sum_m (S_i Sobel(Y U) - Sobel(Y v_i))^2 (* omega sigma^2 if gw_type)
 * parts[0..2] = E_TV, E_int, E_grad.  Returns lambda E_TV + alpha E_int + beta E_grad. */
double orc_sr_energy(int V, int W, int H, int w, int h, int dsf, const float *tx, const float *ty,
                     const float *views, const double *omega, double sigma, double lambda,
                     double dt_alpha, double dt_beta, int gw_type, const unsigned char *vis,
                     const double *U, double *parts)
{
    const size_t N = (size_t)W * H, n = (size_t)w * h;
    double *gx = (double *)calloc(N, sizeof(double)), *gy = (double *)calloc(N, sizeof(double));
    orc_sobel(W, H, U, U + N, U + 2 * N, gx, gy, vis);
    double e_tv = 0.0;
    for (size_t o = 0; o < N; o++) e_tv += fabs(gx[o]) + fabs(gy[o]);
    double *tmp = (double *)malloc(3 * n * sizeof(double));
    double e_int = 0.0;
    for (int v = 0; v < V; v++) {
        orc_gather(W, H, w, h, dsf, tx + v * n, ty + v * n, U, 3, tmp);
        for (int c = 0; c < 3; c++)
            for (size_t m = 0; m < n; m++) {
                const double r = tmp[c * n + m] - views[(size_t)v * 3 * n + c * n + m];
                e_int += r * r * omega[v * n + m] * sigma * sigma;
            }
    }
    double e_grad = 0.0;
    double *ugx = (double *)malloc(N * sizeof(double)), *ugy = (double *)malloc(N * sizeof(double));
    orc_sobel(W, H, U, U + N, U + 2 * N, ugx, ugy, NULL);
    double *vv = (double *)malloc(3 * n * sizeof(double));
    double *vgx = (double *)malloc(n * sizeof(double)), *vgy = (double *)malloc(n * sizeof(double));
    double *wg = (double *)malloc(2 * n * sizeof(double));
    double *ug2 = (double *)malloc(2 * N * sizeof(double));
    memcpy(ug2, ugx, N * sizeof(double));
    memcpy(ug2 + N, ugy, N * sizeof(double));
    for (int v = 0; v < V; v++) {
        for (size_t k = 0; k < 3 * n; k++) vv[k] = views[(size_t)v * 3 * n + k];
        orc_sobel(w, h, vv, vv + n, vv + 2 * n, vgx, vgy, NULL);
        orc_gather(W, H, w, h, dsf, tx + v * n, ty + v * n, ug2, 2, wg);
        for (size_t m = 0; m < n; m++) {
            double a = wg[m] - vgx[m], b = wg[n + m] - vgy[m];
            a *= a;
            b *= b;
            if (gw_type != 0) {
                a *= omega[v * n + m] * sigma * sigma;
                b *= omega[v * n + m] * sigma * sigma;
            }
            e_grad += a + b;
        }
    }
    if (parts) { parts[0] = e_tv; parts[1] = e_int; parts[2] = e_grad; }
    free(gx); free(gy); free(tmp); free(ugx); free(ugy); free(vv); free(vgx); free(vgy); free(wg); free(ug2);
    return lambda * e_tv + dt_alpha * e_int + dt_beta * e_grad;
}

/* Push-pull hole filling, one channel: optimization.cu:210-517 (reduce / expand
 * with the 5-tap a = 0.4 kernel, zero-valued samples skipped on the way down,
 * visibility OR-reduced) and :544-651 (10 levels; the channel becomes
 * filled - original, :629-646).  Levels of zero size are skipped, and the
 * centre tap of expand is clamped to the coarser level (the reference reads
 * past it for odd sizes, :455 and :492). */
static void pp_reduce_f(int W, int H, int w, int h, const double *in, double *out)
{
    const double a = 0.4;
    const double k[5] = {0.25 - 0.5 * a, 0.25, a, 0.25, 0.25 - 0.5 * a};
    double *t = (double *)malloc((size_t)W * (h > 0 ? h : 1) * sizeof(double));
    for (int i = 0; i < h; i++)
        for (int j = 0; j < W; j++) {
            double s = 0.0, wt = 0.0;
            const int rows[5] = {2 * i - 2 < 0 ? 0 : 2 * i - 2, 2 * i - 1 < 0 ? 0 : 2 * i - 1, 2 * i,
                                 H - 1 < 2 * i + 1 ? H - 1 : 2 * i + 1, H - 1 < 2 * i + 2 ? H - 1 : 2 * i + 2};
            for (int q = 0; q < 5; q++) {
                const double v = in[(size_t)rows[q] * W + j];
                if (v != 0.0) { s += k[q] * v; wt += k[q]; }
            }
            t[(size_t)i * W + j] = (wt == 0.0) ? 0.0 : s / wt;
        }
    for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++) {
            double s = 0.0, wt = 0.0;
            const int cols[5] = {2 * j - 2 < 0 ? 0 : 2 * j - 2, 2 * j - 1 < 0 ? 0 : 2 * j - 1, 2 * j,
                                 W - 1 < 2 * j + 1 ? W - 1 : 2 * j + 1, W - 1 < 2 * j + 2 ? W - 1 : 2 * j + 2};
            for (int q = 0; q < 5; q++) {
                const double v = t[(size_t)i * W + cols[q]];
                if (v != 0.0) { s += k[q] * v; wt += k[q]; }
            }
            out[(size_t)i * w + j] = (wt == 0.0) ? 0.0 : s / wt;
        }
    free(t);
}

static void pp_reduce_b(int W, int H, int w, int h, const unsigned char *in, unsigned char *out)
{
    unsigned char *t = (unsigned char *)malloc((size_t)W * (h > 0 ? h : 1));
    for (int i = 0; i < h; i++)
        for (int j = 0; j < W; j++) {
            const int rows[5] = {2 * i - 2 < 0 ? 0 : 2 * i - 2, 2 * i - 1 < 0 ? 0 : 2 * i - 1, 2 * i,
                                 H - 1 < 2 * i + 1 ? H - 1 : 2 * i + 1, H - 1 < 2 * i + 2 ? H - 1 : 2 * i + 2};
            unsigned char v = 0;
            for (int q = 0; q < 5; q++) v |= in[(size_t)rows[q] * W + j] ? 1 : 0;
            t[(size_t)i * W + j] = v;
        }
    for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++) {
            const int cols[5] = {2 * j - 2 < 0 ? 0 : 2 * j - 2, 2 * j - 1 < 0 ? 0 : 2 * j - 1, 2 * j,
                                 W - 1 < 2 * j + 1 ? W - 1 : 2 * j + 1, W - 1 < 2 * j + 2 ? W - 1 : 2 * j + 2};
            unsigned char v = 0;
            for (int q = 0; q < 5; q++) v |= t[(size_t)i * W + cols[q]] ? 1 : 0;
            out[(size_t)i * w + j] = v;
        }
    free(t);
}

static void pp_expand(int W, int H, int w, int h, const double *in, double *out, const unsigned char *vis)
{
    const double a = 0.4;
    const double k[5] = {0.25 - 0.5 * a, 0.25, a, 0.25, 0.25 - 0.5 * a};
    double *t = (double *)malloc((size_t)w * H * sizeof(double));
    for (int i = 0; i < H; i++)
        for (int j = 0; j < w; j++) {
            double s = 0.0;
            if (i % 2 == 0) {
                const int r0 = (i - 2) / 2 < 0 ? 0 : (i - 2) / 2;
                const int r2 = h - 1 < (i + 2) / 2 ? h - 1 : (i + 2) / 2;
                s += 2 * k[0] * in[(size_t)r0 * w + j];
                s += 2 * k[2] * in[(size_t)(i / 2 < h - 1 ? i / 2 : h - 1) * w + j];
                s += 2 * k[4] * in[(size_t)r2 * w + j];
            } else {
                const int r1 = (i - 1) / 2 < 0 ? 0 : (i - 1) / 2;
                const int r3 = h - 1 < (i + 1) / 2 ? h - 1 : (i + 1) / 2;
                s += 2 * k[1] * in[(size_t)r1 * w + j];
                s += 2 * k[3] * in[(size_t)r3 * w + j];
            }
            t[(size_t)i * w + j] = s;
        }
    for (int i = 0; i < H; i++)
        for (int j = 0; j < W; j++) {
            if (vis[(size_t)i * W + j]) continue;
            double s = 0.0;
            if (j % 2 == 0) {
                const int c0 = (j - 2) / 2 < 0 ? 0 : (j - 2) / 2;
                const int c2 = w - 1 < (j + 2) / 2 ? w - 1 : (j + 2) / 2;
                s += 2 * k[0] * t[(size_t)i * w + c0];
                s += 2 * k[2] * t[(size_t)i * w + (j / 2 < w - 1 ? j / 2 : w - 1)];
                s += 2 * k[4] * t[(size_t)i * w + c2];
            } else {
                const int c1 = (j - 1) / 2 < 0 ? 0 : (j - 1) / 2;
                const int c3 = w - 1 < (j + 1) / 2 ? w - 1 : (j + 1) / 2;
                s += 2 * k[1] * t[(size_t)i * w + c1];
                s += 2 * k[3] * t[(size_t)i * w + c3];
            }
            out[(size_t)i * W + j] = s;
        }
    free(t);
}

void orc_push_pull(int W, int H, int levels, double *U, const unsigned char *vis)
{
    const size_t N = (size_t)W * H;
    for (int c = 0; c < 3; c++) {
        double **img = (double **)calloc(levels + 1, sizeof(double *));
        unsigned char **vv = (unsigned char **)calloc(levels + 1, sizeof(unsigned char *));
        int *ws = (int *)calloc(levels + 1, sizeof(int)), *hs = (int *)calloc(levels + 1, sizeof(int));
        img[0] = (double *)malloc(N * sizeof(double));
        memcpy(img[0], U + c * N, N * sizeof(double));
        vv[0] = (unsigned char *)malloc(N);
        memcpy(vv[0], vis, N);
        ws[0] = W; hs[0] = H;
        int top = 0;
        for (int s = 1; s <= levels; s++) {
            const int Ws = W >> (s - 1), Hs = H >> (s - 1), wsn = W >> s, hsn = H >> s;
            if (wsn <= 0 || hsn <= 0) break;
            img[s] = (double *)calloc((size_t)wsn * hsn, sizeof(double));
            vv[s] = (unsigned char *)calloc((size_t)wsn * hsn, 1);
            ws[s] = wsn; hs[s] = hsn;
            pp_reduce_f(Ws, Hs, wsn, hsn, img[s - 1], img[s]);
            pp_reduce_b(Ws, Hs, wsn, hsn, vv[s - 1], vv[s]);
            top = s;
        }
        for (int s = top; s > 0; s--) pp_expand(ws[s - 1], hs[s - 1], ws[s], hs[s], img[s], img[s - 1], vv[s - 1]);
        for (size_t o = 0; o < N; o++) U[c * N + o] = img[0][o] - U[c * N + o];
        for (int s = 0; s <= levels; s++) { free(img[s]); free(vv[s]); }
        free(img); free(vv); free(ws); free(hs);
    }
}

/* ------------------------------------------------------------------------- */
/* Quality metric                                                            */
/* ------------------------------------------------------------------------- */

/* PSNR with vmax = 1, excluding a 2-px border, MSE averaged over RGB:
 * cocolib/common/gsl_image.cpp:991-1025, psnr/psnr_main.cpp:17-20. */
double orc_psnr(int W, int H, const double *A, const double *B)
{
    const int border = 2;
    const size_t N = (size_t)W * H;
    double mse = 0.0;
    long cnt = 0;
    for (int c = 0; c < 3; c++)
        for (int y = border; y < H - border; y++)
            for (int x = border; x < W - border; x++) {
                const double d = A[c * N + (size_t)y * W + x] - B[c * N + (size_t)y * W + x];
                mse += d * d;
                cnt++;
            }
    mse /= (double)cnt;
    if (mse == 0.0) return DBL_MAX; /* coco::psnr (gsl_image.cpp:1009-1015) */
    return 10.0 * log10(1.0 / mse);
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int t)
{
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}
