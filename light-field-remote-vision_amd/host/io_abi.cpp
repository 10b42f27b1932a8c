// C-ABI of the front end (include/lfrv_io.h).
#include <algorithm>
#include "../../include/lfrv_io.h"

#include <cstdio>
#include <cstring>
#include <exception>
#include <sstream>
#include <vector>

#include "../../include/lfrv_ref_config.hpp"
#include "../../include/lfrv_ref.hpp"
#include "../../include/lfrv_ref_image.hpp"

using namespace lfrv_ref;

extern "C" int lfrv_io_load_image(const char *path, size_t *w, size_t *h, float *rgb)
{
    if (!path || !w || !h) return -1;
    Image img;
    if (!load_image(path, img)) return -1;
    *w = img._w;
    *h = img._h;
    if (rgb) {
        const std::vector<float> v = img.planar();
        std::memcpy(rgb, v.data(), v.size() * sizeof(float));
    }
    return 0;
}

extern "C" int lfrv_io_save_image(const char *path, size_t w, size_t h, const float *rgb)
{
    if (!path || !rgb || w == 0 || h == 0) return -1;
    Image img(w, h);
    img.set_planar(rgb);
    return save_image(path, img) ? 0 : -1;
}

extern "C" int lfrv_io_config_data(int argc, const char *const *argv, char *out, size_t cap)
{
    if (argc < 1 || !argv) return -1;
    std::vector<std::string> args(argv, argv + argc);
    std::vector<char *> av;
    for (auto &a : args) av.push_back(&a[0]);
    Config cfg;
    cfg.parse_command_line(argc, av.data());
    ConfigData cd;
    cd.parse_config(cfg);
    std::ostringstream o;
    o.precision(9);
    o << "s_min " << cd._s_min << "\ns_max " << cd._s_max << "\ns_rmv " << cd._s_rmv << "\nt_min " << cd._t_min
      << "\nt_max " << cd._t_max << "\nt_rmv " << cd._t_rmv << "\nlf_name " << cd._lf_name << "\ninit_name "
      << cd._init_name << "\ntau_name " << cd._tau_name << "\ndpart_name " << cd._dpart_name << "\ndepth_name "
      << cd._depth_name << "\ncam_name " << cd._cam_name << "\noutdir " << cd._outdir << "\noutfile " << cd._outfile
      << "\ndsf " << cd._dsf << "\nwidth " << cd._w << "\nheight " << cd._h << "\nugrad_threshold "
      << cd._ugrad_threshold << "\nsigma_sensor " << cd._sigma_sensor << "\nlambda " << cd._lambda
      << "\nlambda_max_factor " << cd._lambda_max_factor << "\ndw_type " << cd._dw_type << "\ngw_type " << cd._gw_type
      << "\ndt_alpha " << cd._dt_alpha << "\ndt_beta " << cd._dt_beta << "\nniter " << cd._niter << "\nnmeta_iter "
      << cd._nmeta_iter << "\ngradient_step " << cd._gradient_step << "\nibr_system " << cd._ibr_system << "\n";
    std::string algo;
    if (cfg.get_switch("algorithm", algo)) o << "algorithm " << algo << "\n";
    const std::string s = o.str();
    if (out && cap > 0) {
        const size_t n = std::min(cap - 1, s.size());
        std::memcpy(out, s.data(), n);
        out[n] = 0;
    }
    return (int)s.size() + 1;
}

extern "C" int lfrv_io_export_warps(int argc, const char *const *argv)
{
    if (argc < 1 || !argv) return -1;
    std::vector<std::string> args(argv, argv + argc);
    std::vector<char *> av;
    for (auto &a : args) av.push_back(&a[0]);
    try {
        Config cfg;
        cfg.parse_command_line(argc, av.data());
        ConfigData cd;
        cd.parse_config(cfg);
        export_warps(&cd);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "lfrv_io_export_warps: %s\n", e.what());
        return -1;
    }
    return 0;
}

extern "C" int lfrv_io_from_mve(int argc, const char *const *argv)
{
    if (argc < 1 || !argv) return -1;
    std::vector<std::string> args(argv, argv + argc);
    std::vector<char *> av;
    for (auto &a : args) av.push_back(&a[0]);
    try {
        Config cfg;
        cfg.parse_command_line(argc, av.data());
        ConfigData cd;
        cd.parse_config(cfg);
        from_mve_to_coco(&cd);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "lfrv_io_from_mve: %s\n", e.what());
        return -1;
    }
    return 0;
}

extern "C" int lfrv_io_mve_camera(const char *meta_ini, int w, int h, double *K, double *R, double *t, double *P)
{
    if (!meta_ini) return -1;
    MveCamera cam;
    if (!import_mve_camera(meta_ini, w, h, cam)) return -1;
    if (K) std::copy(cam.K, cam.K + 9, K);
    if (R) std::copy(cam.R, cam.R + 9, R);
    if (t) std::copy(cam.t, cam.t + 3, t);
    if (P) std::copy(cam.P, cam.P + 12, P);
    return 0;
}

extern "C" int lfrv_io_mve_depth(const char *mve_name, int s, int scale_min, int scale_max, int w, int h, float *z,
                                 float *dzdx, float *dzdy)
{
    if (!mve_name || w < 1 || h < 1) return -1;
    Image d;
    if (!import_mve_depth(mve_name, s, scale_min, scale_max, w, h, d)) return -1;
    const size_t n = (size_t)w * h;
    if (z) std::copy(d._r.begin(), d._r.begin() + n, z);
    if (dzdx) std::copy(d._g.begin(), d._g.begin() + n, dzdx);
    if (dzdy) std::copy(d._b.begin(), d._b.begin() + n, dzdy);
    return 0;
}
