// The MVE front of the reference's OpenGL warp path (algorithm fromMVE2coco,
// main.cpp:313-346): Scene::importViews (importScene.cpp:105-230), the
// depth-map filter Scene::filterDepthMaps (scene.cpp:1605-1624), then the
// target depth map and the warps as export_warps computes them.
#include "../../include/lfrv_ref.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>

namespace lfrv_ref {

namespace {

constexpr float kInvalidDepth = 1000.0f;  // frameBuffer.h:13

void check(lfrv_ctx *ctx, lfrv_status st, const char *what)
{
    if (st == LFRV_OK) return;
    throw Error(std::string(what) + ": " + (ctx ? lfrv_last_error(ctx) : "invalid argument"));
}

std::string view_file(const std::string &mve, int s, const char *leaf)
{
    char dir[64];
    std::snprintf(dir, sizeof dir, "/views/view_%04i.mve/", s);
    return mve + dir + leaf;
}

}  // namespace

bool import_mve_camera(const std::string &path, int W, int H, MveCamera &cam)
{
    std::ifstream in(path);
    if (!in) return false;
    // the words up to "[camera]"; a "[view]" first, or 100 words without
    // either, means structure from motion gave this view no camera
    std::string tok;
    int words = 0;
    while (tok != "[camera]" && tok != "[view]" && words < 100) {
        if (!(in >> tok)) return false;
        words++;
    }
    if (tok != "[camera]") return false;
    std::string key, eq;
    double focal = 0.0, aspect = 0.0, pp[2] = {0.0, 0.0}, R[9], t[3];
    // "name = value..." in this order (view.cpp:244-248)
    in >> key >> eq >> focal >> key >> eq >> aspect >> key >> eq >> pp[0] >> pp[1] >> key >> eq;
    for (double &r : R) in >> r;
    in >> key >> eq >> t[0] >> t[1] >> t[2];
    if (!in || aspect == 0.0) return false;
    // focal length in units of the larger side, principal point in units of the sides
    const double fx = (W >= H ? W : H) * focal;
    const double K[9] = {fx, 0.0, W * pp[0], 0.0, fx / aspect, H * pp[1], 0.0, 0.0, 1.0};
    std::memcpy(cam.K, K, sizeof K);
    std::memcpy(cam.R, R, sizeof R);
    std::memcpy(cam.t, t, sizeof t);
    // P = K [R | t]  (pinholeCamera.cpp:5-15)
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 4; c++) {
            double v = 0.0;
            for (int k = 0; k < 3; k++) v += K[3 * r + k] * (c < 3 ? R[3 * k + c] : t[k]);
            cam.P[4 * r + c] = v;
        }
    return true;
}

bool load_mvei(const std::string &path, int &w, int &h, std::vector<std::vector<float>> &planes)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    std::string signature;
    std::getline(in, signature);  // "\x89MVE_IMAGE"
    uint32_t hdr[4];              // width, height, channels, type
    if (!in.read(reinterpret_cast<char *>(hdr), sizeof hdr)) return false;
    if (hdr[0] == 0 || hdr[1] == 0 || hdr[2] == 0 || hdr[0] > 65536 || hdr[1] > 65536 || hdr[2] > 16) return false;
    w = (int)hdr[0];
    h = (int)hdr[1];
    const size_t n = (size_t)w * h, C = hdr[2];
    std::vector<float> raw(n * C);  // interleaved, rows top to bottom
    if (!in.read(reinterpret_cast<char *>(raw.data()), (std::streamsize)(raw.size() * sizeof(float)))) return false;
    planes.assign(C, std::vector<float>(n));
    for (size_t m = 0; m < n; m++)
        for (size_t c = 0; c < C; c++) planes[c][m] = raw[m * C + c];
    return true;
}

bool import_mve_depth(const std::string &mve, int s, int scale_min, int scale_max, int W, int H, Image &zdz)
{
    zdz = Image((size_t)W, (size_t)H);
    bool any = false;
    for (int scale = scale_min; scale <= scale_max; scale++) {
        char leaf[32];
        std::vector<std::vector<float>> d, g;
        int dw = 0, dh = 0, gw = 0, gh = 0;
        std::snprintf(leaf, sizeof leaf, "depth-L%01i.mvei", scale);
        const bool okd = load_mvei(view_file(mve, s, leaf), dw, dh, d);
        std::snprintf(leaf, sizeof leaf, "dz-L%01i.mvei", scale);
        const bool okg = load_mvei(view_file(mve, s, leaf), gw, gh, g);
        if (!okd || !okg || d.empty() || g.size() < 2 || dw != gw || dh != gh) return false;
        // nearest upsampling by W / w (texture.cpp:218-278): a pixel whose
        // three values are all 0 becomes (1000, 0, 0)
        const int f = W / dw;
        if (f < 1 || dw * f != W || dh * f != H) return false;
        // InputView::addDepthScale clears the accumulated map before every
        // scale (view.cpp:735-741), so the last scale read is the one kept,
        // its depth derivatives divided by 2^scale (addDepthScale.frag)
        const float sc = std::ldexp(1.0f, scale);
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                const size_t q = (size_t)(y / f) * dw + x / f, o = (size_t)y * W + x;
                const float z = d[0][q], gx = g[0][q], gy = g[1][q];
                const bool zero = z == 0.0f && gx == 0.0f && gy == 0.0f;
                zdz._r[o] = zero ? kInvalidDepth : z;
                zdz._g[o] = zero ? 0.0f : gx / sc;
                zdz._b[o] = zero ? 0.0f : gy / sc;
            }
        any = true;
    }
    return any;
}

void from_mve_to_coco(ConfigData *cd)
{
    if (cd->_dsf != 1) throw Error("fromMVE2coco: the OpenGL path renders the target at the views' size (dsf 1)");
    if (cd->_mve_name.empty()) throw Error("fromMVE2coco needs mve_name");
    const int W = (int)cd->_w, H = (int)cd->_h;
    if (W < 1 || H < 1) throw Error("fromMVE2coco needs the view size (width, height)");
    if (cd->_scale_min < 0 || cd->_scale_max < cd->_scale_min) throw Error("fromMVE2coco needs scale_min <= scale_max");
    // importViews: views with an image, a camera and every depth scale, in
    // order; the target is the view s_rmv, kept among the inputs
    std::vector<Image> views, zdz;
    std::vector<MveCamera> cams;
    int target = -1;
    for (int s = cd->_s_min; s <= cd->_s_max; s++) {
        Image img, d;
        MveCamera cam;
        const bool ok = load_image(view_file(cd->_mve_name, s, "undistorted.png"), img) &&
                        import_mve_camera(view_file(cd->_mve_name, s, "meta.ini"), W, H, cam) &&
                        import_mve_depth(cd->_mve_name, s, cd->_scale_min, cd->_scale_max, W, H, d);
        if (!ok) {
            if (s == cd->_s_rmv) throw Error("fromMVE2coco: the target view " + std::to_string(s) + " is incomplete");
            continue;
        }
        if ((int)img._w != W || (int)img._h != H) throw Error("fromMVE2coco: view " + std::to_string(s) + " is not w x h");
        if (s == cd->_s_rmv) target = (int)views.size();
        views.push_back(std::move(img));
        zdz.push_back(std::move(d));
        cams.push_back(cam);
    }
    if (target < 0) throw Error("fromMVE2coco: s_rmv names no view in [s_min, s_max]");
    const int nv = (int)views.size();
    const size_t n = (size_t)W * H;
    lfrv_params p;
    lfrv_default_params(&p, LFRV_MODE_GRADIENT_IBR, nv, W, H, 1);
    lfrv_ctx *raw = nullptr;
    if (lfrv_create(&p, &raw) != LFRV_OK) throw Error("lfrv_create failed");
    std::unique_ptr<lfrv_ctx, void (*)(lfrv_ctx *)> ctx(raw, lfrv_destroy);
    // filterDepthMaps on the GPU, then the views' depth and its derivatives
    float *dev = nullptr;
    if (lfrv_dev_alloc(3 * n * sizeof(float), reinterpret_cast<void **>(&dev)) != LFRV_OK) throw Error("device memory");
    std::unique_ptr<float, void (*)(float *)> hold(dev, [](float *q) { lfrv_dev_free(q); });
    std::vector<float> zf(n), gxf(n), gyf(n);
    for (int i = 0; i < nv; i++) {
        check(nullptr, lfrv_memcpy(dev, zdz[i]._r.data(), n * sizeof(float)), "copy");
        check(nullptr, lfrv_memcpy(dev + n, zdz[i]._g.data(), n * sizeof(float)), "copy");
        check(nullptr, lfrv_memcpy(dev + 2 * n, zdz[i]._b.data(), n * sizeof(float)), "copy");
        check(nullptr, lfrv_op_radial_to_ortho(W, H, cams[i].K, dev, dev + n, dev + 2 * n, nullptr), "filterDepthMaps");
        check(nullptr, lfrv_memcpy(zf.data(), dev, n * sizeof(float)), "copy");
        check(nullptr, lfrv_memcpy(gxf.data(), dev + n, n * sizeof(float)), "copy");
        check(nullptr, lfrv_memcpy(gyf.data(), dev + 2 * n, n * sizeof(float)), "copy");
        check(ctx.get(), lfrv_set_depth(ctx.get(), i, zf.data(), cams[i].P), "depth");
        check(ctx.get(), lfrv_set_depth_gradient(ctx.get(), i, gxf.data(), gyf.data()), "depth gradient");
    }
    check(ctx.get(), lfrv_set_target_camera(ctx.get(), cams[target].P), "target camera");
    check(ctx.get(), lfrv_warps_from_target_depth(ctx.get(), target), "lfrv_warps_from_target_depth");
    // exportWarps: file index = position among the imported views (scene.cpp:1546-1585)
    write_gl_warps(ctx.get(), cd, nv, W, H, p.sigma_z, [](int i) { return i; });
    for (int i = 0; i < nv; i++)
        if (!cd->_lf_name.empty() && !save_image(pattern_name(cd->_lf_name, i), views[i]))
            throw Error("cannot write " + pattern_name(cd->_lf_name, i));
    std::printf("fromMVE2coco: %d views %dx%d, target view %d (index %d)\n", nv, W, H, cd->_s_rmv, target);
}

}  // namespace lfrv_ref
