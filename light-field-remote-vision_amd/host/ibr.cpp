// The reference's host interface over the lfrv C-ABI; see include/lfrv_ref.hpp.
#include "../../include/lfrv_ref.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <functional>
#include <memory>

namespace lfrv_ref {

static void check(lfrv_ctx *ctx, lfrv_status st, const char *what)
{
    if (st == LFRV_OK) return;
    throw Error(std::string(what) + ": " + (ctx ? lfrv_last_error(ctx) : "invalid argument"));
}

// ---- front end ---------------------------------------------------------------------

std::vector<Image> import_lightfield(ConfigData *cd)
{
    std::vector<Image> lf;
    if (cd->_w * cd->_h == 0) {  // the size of the image at index 0 (import.cpp:131-136)
        Image first;
        if (load_image(pattern_name(cd->_lf_name, 0), first)) {
            cd->_w = (unsigned)first._w;
            cd->_h = (unsigned)first._h;
        }
    }
    for (int s = cd->_s_min; s <= cd->_s_max; s++) {
        if (s == cd->_s_rmv) continue;  // the view to synthesise
        const std::string name = pattern_name(cd->_lf_name, s);
        Image img;
        if (!load_image(name, img)) throw Error("cannot read view " + name);
        if (img._w != cd->_w || img._h != cd->_h) throw Error("view " + name + " has the wrong size");
        lf.push_back(std::move(img));
    }
    if (cd->_dsf <= 0) throw Error("dsf must be positive");
    cd->_nviews = lf.size();
    cd->_W = cd->_w * cd->_dsf;
    cd->_H = cd->_h * cd->_dsf;
    return lf;
}

std::vector<Image> import_warps(ConfigData *cd, const std::string &pattern)
{
    std::vector<Image> out;
    for (int s = cd->_s_min; s <= cd->_s_max; s++) {
        if (s == cd->_s_rmv) continue;
        const std::string name = pattern_name(pattern, s);
        Image img;
        if (!load_pfm(name, img)) throw Error("cannot read " + name);
        if (img._w != cd->_w || img._h != cd->_h) throw Error(name + " has the wrong size");
        out.push_back(std::move(img));
    }
    return out;
}

std::vector<double> import_cam(const std::string &pattern, int index)
{
    const std::string name = pattern_name(pattern, index);
    std::ifstream in(name);
    std::string header;
    std::vector<double> P(12);
    if (!(in >> header) || header != "CONTOUR") throw Error("camera file " + name + " lacks the CONTOUR header");
    for (int k = 0; k < 12; k++)
        if (!(in >> P[k])) throw Error("camera file " + name + " is truncated");
    return P;
}

static std::vector<Image *> ptrs(std::vector<Image> &v)
{
    std::vector<Image *> p;
    for (auto &i : v) p.push_back(&i);
    return p;
}

// Warps of every view from depth maps + cameras on the GPU (import.cpp:243-450
// with the partial tau of tauPartial.frag): depth_name / cam_name patterns at
// the view index s, the target camera at index nviews (+1 when a view is
// removed), as import.cpp:254-258 reads it.  When tau_name is set the warps
// are also written there (R = tau_x, G = B = tau_y), as import.cpp:446 does.
static void warps_from_depth(lfrv_ctx *ctx, ConfigData *cd)
{
    int i = 0;
    for (int s = cd->_s_min; s <= cd->_s_max; s++) {
        if (s == cd->_s_rmv) continue;
        Image depth;
        const std::string dn = pattern_name(cd->_depth_name, s);
        if (!load_pfm(dn, depth)) throw Error("cannot read depth map " + dn);
        if (depth._w != cd->_w || depth._h != cd->_h) throw Error(dn + " has the wrong size");
        const std::vector<double> P = import_cam(cd->_cam_name, s);
        check(ctx, lfrv_set_depth(ctx, i, depth._r.data(), P.data()), "lfrv_set_depth");
        i++;
    }
    const int u_index = (int)cd->_nviews + (cd->_s_rmv != -1 ? 1 : 0);
    const std::vector<double> Pu = import_cam(cd->_cam_name, u_index);
    check(ctx, lfrv_set_target_camera(ctx, Pu.data()), "lfrv_set_target_camera");
    check(ctx, lfrv_warps_from_depth(ctx), "lfrv_warps_from_depth");
    if (!cd->_tau_name.empty()) {
        const size_t n = (size_t)cd->_w * cd->_h;
        std::vector<float> tx(cd->_nviews * n), ty(cd->_nviews * n);
        size_t cnt = 0;
        check(ctx, lfrv_get_buffer(ctx, LFRV_BUF_TAU_X, tx.data(), &cnt), "tau");
        check(ctx, lfrv_get_buffer(ctx, LFRV_BUF_TAU_Y, ty.data(), &cnt), "tau");
        int v = 0;
        for (int s = cd->_s_min; s <= cd->_s_max; s++) {
            if (s == cd->_s_rmv) continue;
            Image t(cd->_w, cd->_h);
            std::copy(tx.begin() + v * n, tx.begin() + (v + 1) * n, t._r.begin());
            std::copy(ty.begin() + v * n, ty.begin() + (v + 1) * n, t._g.begin());
            t._b = t._g;
            save_pfm(pattern_name(cd->_tau_name, s), t);
            v++;
        }
    }
}

// Config_data asserts three channels (config.cpp:115-116); checked here.
static void check_channels(const ConfigData *cd)
{
    if (cd->_nchannels != 3) throw Error("nchannels must be 3 (grayscale is not handled)");
}

static void load_warps(lfrv_ctx *ctx, ConfigData *cd, std::vector<Image> &tau, std::vector<Image> &dpart)
{
    if (!cd->_depth_name.empty() && !cd->_ply_name.empty() && !cd->_cam_name.empty())
        throw Error("depth maps from a point cloud (ply_name) are not provided; give depth_name + cam_name");
    if (!cd->_depth_name.empty() && !cd->_cam_name.empty()) {
        warps_from_depth(ctx, cd);
        return;
    }
    tau = import_warps(cd, cd->_tau_name);
    dpart = import_warps(cd, cd->_dpart_name);
}

static std::vector<std::vector<float>> split(const std::vector<float> &planar, size_t N)
{
    std::vector<std::vector<float>> U(3);
    for (int c = 0; c < 3; c++) U[c].assign(planar.begin() + c * N, planar.begin() + (c + 1) * N);
    return U;
}

static Image solution_image(lfrv_ctx *ctx, size_t W, size_t H)
{
    Image R(W, H);
    std::vector<float> u(3 * W * H);
    check(ctx, lfrv_get_solution(ctx, u.data()), "lfrv_get_solution");
    R.set_planar(u.data());
    return R;
}

// ---- IBR_gradient ---------------------------------------------------------------------

Data *init_data(ConfigData *cd)
{
    lfrv_params p;
    lfrv_default_params(&p, LFRV_MODE_GRADIENT_IBR, (int)cd->_nviews, (int)cd->_w, (int)cd->_h, cd->_dsf);
    p.sigma_sensor = cd->_sigma_sensor;
    p.ugrad_threshold = cd->_ugrad_threshold;
    p.dw_type = cd->_dw_type;
    p.gw_type = cd->_gw_type;
    p.dt_alpha = cd->_dt_alpha;
    p.dt_beta = cd->_dt_beta;
    p.gradient_step = cd->_gradient_step;
    std::unique_ptr<Data> d(new Data);
    d->_nviews = cd->_nviews;
    d->_dsf = (size_t)cd->_dsf;
    d->_w = cd->_w;
    d->_h = cd->_h;
    d->_W = d->_w * d->_dsf;
    d->_H = d->_h * d->_dsf;
    d->_ks = d->_dsf + 1;
    d->_sigma_sensor = cd->_sigma_sensor;
    d->_ugrad_threshold = cd->_ugrad_threshold;
    d->_dw_type = cd->_dw_type;
    d->_gw_type = cd->_gw_type;
    d->_dt_alpha = cd->_dt_alpha;
    d->_dt_beta = cd->_dt_beta;
    d->_gradient_step = cd->_gradient_step;
    d->_outdir = cd->_outdir;
    d->_init_name = cd->_init_name;
    if (cd->_ibr_system == "normal_equations") d->_normal_equations = 1;
    else if (cd->_ibr_system != "poisson")
        throw Error("ibr_system must be poisson or normal_equations, not '" + cd->_ibr_system + "'");
    check(nullptr, lfrv_create(&p, &d->ctx), "lfrv_create");
    check(d->ctx, lfrv_clear_solution(d->ctx), "lfrv_clear_solution");  // U = 0 (gradientIBR.cu:281-287)
    return d.release();
}

void free_data(Data *data)
{
    if (!data) return;
    lfrv_destroy(data->ctx);
    delete data;
}

void create_view(Data *d, size_t nview, const Image *I)
{
    const std::vector<float> v = I->planar();
    check(d->ctx, lfrv_set_view(d->ctx, (int)nview, v.data()), "create_view");
}

void read_tau(Data *d, Image **tau)
{
    for (size_t i = 0; i < d->_nviews; i++)
        check(d->ctx, lfrv_set_tau(d->ctx, (int)i, tau[i]->_r.data(), tau[i]->_g.data()), "read_tau");
}

void read_partial_tau(Data *d, Image **dp)
{
    for (size_t i = 0; i < d->_nviews; i++)
        check(d->ctx, lfrv_set_partial_tau(d->ctx, (int)i, dp[i]->_r.data(), dp[i]->_g.data(), dp[i]->_b.data()),
              "read_partial_tau");
}

void compute_sparse_matrix(Data *d) { check(d->ctx, lfrv_prepare(d->ctx), "compute_sparse_matrix"); }
void init_forward_warp_structure(Data *, size_t) {}  // built by compute_sparse_matrix for every view
void setup_target_visibility(Data *d) { check(d->ctx, lfrv_setup_visibility(d->ctx), "setup_target_visibility"); }
void compute_weights(Data *d) { check(d->ctx, lfrv_compute_weights(d->ctx), "compute_weights"); }

void compute_initial_image(Data *d)
{
    // a readable init_name PFM replaces the average (gradientIBR.cu:737-762)
    Image init;
    if (!d->_init_name.empty() && load_pfm(d->_init_name, init) && init._w == d->_W && init._h == d->_H) {
        const std::vector<float> v = init.planar();
        check(d->ctx, lfrv_compute_initial_image(d->ctx, v.data()), "compute_initial_image");
        return;
    }
    check(d->ctx, lfrv_compute_initial_image(d->ctx, nullptr), "compute_initial_image");
}

void hole_filling(Data *) {}  // run by compute_initial_image
void init_u_gradient(Data *d) { check(d->ctx, lfrv_init_u_gradient(d->ctx), "init_u_gradient"); }

void possion_conj_grad(Data *d)
{
    // The reference's system: lambda = 1 (solver.cu:541), u0 the current
    // (hole-filled) solution, the luminance gradient target for every channel,
    // plain CG as in solver.cu:526-699 (its bugs fixed), run to a tolerance
    // instead of its 5 iterations.  dt_alpha / dt_beta play no part in it.
    const lfrv_status st = d->_normal_equations ? lfrv_cg_solve(d->ctx, 1e-5, 5000, &d->cg)
                                                : lfrv_poisson_cg(d->ctx, 1.0, 1e-5, 5000, 0, &d->cg);
    if (st == LFRV_E_NONCONVERGED) {
        // the iterate is kept (the caller may still write it); the driver reports it
        std::fprintf(stderr, "possion_conj_grad: CG did not reach 1e-5 in 5000 iterations (true rel. residual "
                             "%.3g %.3g %.3g)\n", d->cg.true_rel_residual[0], d->cg.true_rel_residual[1],
                     d->cg.true_rel_residual[2]);
        return;
    }
    check(d->ctx, st, "possion_conj_grad");
}

void get_solution(Data *d, std::vector<std::vector<float>> &U)
{
    std::vector<float> u(3 * d->_W * d->_H);
    check(d->ctx, lfrv_get_solution(d->ctx, u.data()), "get_solution");
    U = split(u, d->_W * d->_H);
}

void write_pfm_solution(Data *d)
{
    if (!save_pfm(d->_outdir + "/output.pfm", solution_image(d->ctx, d->_W, d->_H)))
        throw Error("cannot write " + d->_outdir + "/output.pfm");
}

void IBR_gradient(ConfigData *cd)
{
    check_channels(cd);
    std::vector<Image> lf = import_lightfield(cd);
    Data *data = init_data(cd);
    try {
        for (size_t v = 0; v < lf.size(); v++) create_view(data, v, &lf[v]);
        std::vector<Image> tau, dpart;
        load_warps(data->ctx, cd, tau, dpart);
        if (!tau.empty()) {
            read_tau(data, ptrs(tau).data());
            read_partial_tau(data, ptrs(dpart).data());
        }
        compute_sparse_matrix(data);
        for (size_t v = 0; v < data->_nviews; v++) init_forward_warp_structure(data, v);
        setup_target_visibility(data);
        compute_weights(data);
        compute_initial_image(data);
        save_png(data->_outdir + "/u_init_filled.png", solution_image(data->ctx, data->_W, data->_H));
        init_u_gradient(data);
        possion_conj_grad(data);
        save_image(data->_outdir + "/" + cd->_outfile, solution_image(data->ctx, data->_W, data->_H));
        write_pfm_solution(data);
        std::printf("IBR_gradient (%s): %zu views %ux%u -> %zux%zu, CG iterations %d %d %d, rel. residual %.3g %.3g "
                    "%.3g\n", data->_normal_equations ? "normal equations" : "screened Poisson", data->_nviews,
                    cd->_w, cd->_h, data->_W, data->_H, data->cg.iterations[0], data->cg.iterations[1],
                    data->cg.iterations[2], data->cg.true_rel_residual[0], data->cg.true_rel_residual[1],
                    data->cg.true_rel_residual[2]);
        if (data->cg.status == LFRV_E_NONCONVERGED)
            throw Error("possion_conj_grad did not converge (outputs written from the last iterate)");
    } catch (...) {
        free_data(data);
        throw;
    }
    free_data(data);
}

// ---- IBR_color ------------------------------------------------------------------------

coco_vtv_data *coco_vtv_alloc(size_t W, size_t H)
{
    coco_vtv_data *d = new coco_vtv_data;
    d->_W = W;
    d->_H = H;
    return d;
}

void coco_vtv_free(coco_vtv_data *d) { delete d; }

bool coco_vtv_sr_init_unstructured(coco_vtv_data *d, ConfigData *cd)
{
    lfrv_params p;
    lfrv_default_params(&p, LFRV_MODE_COLOR_SR, (int)cd->_nviews, (int)cd->_w, (int)cd->_h, cd->_dsf);
    p.sigma_sensor = cd->_sigma_sensor;
    p.ugrad_threshold = cd->_ugrad_threshold;
    p.dw_type = cd->_dw_type;
    p.gw_type = cd->_gw_type;
    p.dt_alpha = cd->_dt_alpha;
    p.dt_beta = cd->_dt_beta;
    p.gradient_step = cd->_gradient_step;
    p.lambda = d->_lambda;
    p.lambda_max_factor = d->_lambda_max_factor;
    p.inner_iterations = d->_inner_iterations;
    d->_nviews = cd->_nviews;
    d->_dsf = (size_t)cd->_dsf;
    check(nullptr, lfrv_create(&p, &d->ctx), "coco_vtv_sr_init_unstructured");
    check(d->ctx, lfrv_clear_solution(d->ctx), "coco_vtv_initialize");
    return true;
}

bool coco_vtv_sr_free_unstructured(coco_vtv_data *d)
{
    lfrv_destroy(d->ctx);
    d->ctx = nullptr;
    return true;
}

bool coco_vtv_sr_create_view_unstructured(coco_vtv_data *d, size_t nview, const Image *I)
{
    const std::vector<float> v = I->planar();
    check(d->ctx, lfrv_set_view(d->ctx, (int)nview, v.data()), "coco_vtv_sr_create_view_unstructured");
    return true;
}

bool coco_vtv_sr_read_tau(coco_vtv_data *d, Image **tau)
{
    for (size_t i = 0; i < d->_nviews; i++)
        check(d->ctx, lfrv_set_tau(d->ctx, (int)i, tau[i]->_r.data(), tau[i]->_g.data()), "coco_vtv_sr_read_tau");
    return true;
}

bool coco_vtv_sr_read_partial_tau(coco_vtv_data *d, Image **dp)
{
    for (size_t i = 0; i < d->_nviews; i++)
        check(d->ctx, lfrv_set_partial_tau(d->ctx, (int)i, dp[i]->_r.data(), dp[i]->_g.data(), dp[i]->_b.data()),
              "coco_vtv_sr_read_partial_tau");
    return true;
}

bool coco_vtv_sr_compute_sparse_matrix(coco_vtv_data *d)
{
    check(d->ctx, lfrv_prepare(d->ctx), "coco_vtv_sr_compute_sparse_matrix");
    return true;
}

bool vtv_sr_init_forward_warp_structure_unstructured(coco_vtv_data *, size_t) { return true; }

bool coco_vtv_sr_compute_weights_unstructured(coco_vtv_data *d)
{
    check(d->ctx, lfrv_compute_weights(d->ctx), "coco_vtv_sr_compute_weights_unstructured");
    return true;
}

void coco_vtv_setup_visibility_mask(coco_vtv_data *d)
{
    check(d->ctx, lfrv_setup_visibility(d->ctx), "coco_vtv_setup_visibility_mask");
}

void coco_vtv_sr_compute_initial_image(coco_vtv_data *d, ConfigData *cd)
{
    Image init;
    if (!cd->_init_name.empty() && load_pfm(cd->_init_name, init) && init._w == d->_W && init._h == d->_H) {
        const std::vector<float> v = init.planar();
        check(d->ctx, lfrv_compute_initial_image(d->ctx, v.data()), "coco_vtv_sr_compute_initial_image");
        return;
    }
    check(d->ctx, lfrv_compute_initial_image(d->ctx, nullptr), "coco_vtv_sr_compute_initial_image");
}

double coco_vtv_sr_primal_energy_unstructured(coco_vtv_data *d, float)
{
    double e = 0.0;
    check(d->ctx, lfrv_sr_primal_energy(d->ctx, &e, nullptr), "coco_vtv_sr_primal_energy_unstructured");
    return e;
}

bool coco_vtv_sr_iteration_fista_unstructured(coco_vtv_data *d, bool)
{
    check(d->ctx, lfrv_sr_iteration_fista(d->ctx), "coco_vtv_sr_iteration_fista_unstructured");
    return true;
}

bool coco_vtv_sr_init_regularizer_weight_unstructured(coco_vtv_data *d)
{
    check(d->ctx, lfrv_sr_init_regularizer_weight(d->ctx), "coco_vtv_sr_init_regularizer_weight_unstructured");
    return true;
}

void coco_vtv_push_pull(coco_vtv_data *) {}  // run by coco_vtv_sr_compute_initial_image

void coco_vtv_get_solution(coco_vtv_data *d, std::vector<std::vector<float>> &U)
{
    std::vector<float> u(3 * d->_W * d->_H);
    check(d->ctx, lfrv_get_solution(d->ctx, u.data()), "coco_vtv_get_solution");
    U = split(u, d->_W * d->_H);
}

bool coco_vtv_sr_write_pfm_solution(coco_vtv_data *d)
{
    return save_pfm(d->_basedir + "/output.pfm", solution_image(d->ctx, d->_W, d->_H));
}

void sr_synthesize_view(ConfigData *cd)
{
    check_channels(cd);
    std::vector<Image> lf = import_lightfield(cd);
    const size_t W = cd->_W, H = cd->_H;
    std::unique_ptr<coco_vtv_data, void (*)(coco_vtv_data *)> mtv(coco_vtv_alloc(W, H), [](coco_vtv_data *d) {
        if (d->ctx) coco_vtv_sr_free_unstructured(d);
        coco_vtv_free(d);
    });
    mtv->_basedir = cd->_outdir;
    mtv->_lambda = cd->_lambda;
    mtv->_lambda_max_factor = cd->_lambda_max_factor;
    if (cd->_niter <= 0) throw Error("niter must be positive");
    mtv->_inner_iterations = cd->_niter;  // niter also counts the outer steps (reproduced)
    coco_vtv_sr_init_unstructured(mtv.get(), cd);
    for (size_t v = 0; v < lf.size(); v++) coco_vtv_sr_create_view_unstructured(mtv.get(), v, &lf[v]);
    std::vector<Image> tau, dpart;
    load_warps(mtv->ctx, cd, tau, dpart);
    if (!tau.empty()) {
        coco_vtv_sr_read_tau(mtv.get(), ptrs(tau).data());
        coco_vtv_sr_read_partial_tau(mtv.get(), ptrs(dpart).data());
    }
    coco_vtv_sr_compute_sparse_matrix(mtv.get());
    for (size_t v = 0; v < mtv->_nviews; v++) vtv_sr_init_forward_warp_structure_unstructured(mtv.get(), v);
    coco_vtv_sr_compute_weights_unstructured(mtv.get());
    coco_vtv_setup_visibility_mask(mtv.get());
    coco_vtv_sr_compute_initial_image(mtv.get(), cd);

    // the re-weighted FISTA loop of super-resolution.cpp:177-300, its stopping
    // rules (in the reference's float arithmetic) evaluated on the device
    lfrv_sr_stats st{};
    check(mtv->ctx, lfrv_sr_solve(mtv->ctx, cd->_niter, cd->_nmeta_iter, &st), "IBR_color solve");
    save_image(cd->_outdir + "/" + cd->_outfile, solution_image(mtv->ctx, W, H));
    coco_vtv_sr_write_pfm_solution(mtv.get());
    std::printf("IBR_color: %zu views %ux%u -> %zux%zu, %d meta-iterations, %d FISTA steps, energy %.6g\n",
                mtv->_nviews, cd->_w, cd->_h, W, H, st.meta_iterations, st.fista_steps, (double)st.energy);
}

void write_gl_warps(lfrv_ctx *ctx, ConfigData *cd, int nv, int w, int h, float sigma_z,
                    const std::function<int(int)> &file_index)
{
    const size_t n = (size_t)w * h;
    auto get = [&](int which, size_t count) {
        std::vector<float> v(count);
        size_t cnt = 0;
        check(ctx, lfrv_get_buffer(ctx, which, v.data(), &cnt), "warp buffers");
        return v;
    };
    const std::vector<float> tx = get(LFRV_BUF_TAU_X, nv * n), ty = get(LFRV_BUF_TAU_Y, nv * n),
                             wd = get(LFRV_BUF_TAU_WEIGHT, nv * n), dx = get(LFRV_BUF_DPART_X, nv * n),
                             dy = get(LFRV_BUF_DPART_Y, nv * n), ud = get(LFRV_BUF_TARGET_DEPTH, n);
    for (int i = 0; i < nv; i++) {
        const int s = file_index(i);
        if (!cd->_tau_name.empty()) {
            Image t(w, h);
            std::copy(tx.begin() + i * n, tx.begin() + (i + 1) * n, t._r.begin());
            std::copy(ty.begin() + i * n, ty.begin() + (i + 1) * n, t._g.begin());
            std::copy(wd.begin() + i * n, wd.begin() + (i + 1) * n, t._b.begin());
            if (!save_pfm(pattern_name(cd->_tau_name, s), t)) throw Error("cannot write " + pattern_name(cd->_tau_name, s));
        }
        if (!cd->_dpart_name.empty()) {
            // tauPartial.frag:46 writes (sigma, dtau/dz); the buffers hold sigma dtau/dz
            Image t(w, h);
            for (size_t m = 0; m < n; m++) {
                t._r[m] = sigma_z;
                t._g[m] = dx[i * n + m] / sigma_z;
                t._b[m] = dy[i * n + m] / sigma_z;
            }
            if (!save_pfm(pattern_name(cd->_dpart_name, s), t)) throw Error("cannot write " + pattern_name(cd->_dpart_name, s));
        }
    }
    if (!cd->_uDepth_name.empty()) {
        Image t(w, h);
        t._r = ud;
        std::fill(t._g.begin(), t._g.end(), 0.0f);
        std::fill(t._b.begin(), t._b.end(), 0.0f);
        if (!save_pfm(cd->_uDepth_name, t)) throw Error("cannot write " + cd->_uDepth_name);
    }
}

void export_warps(ConfigData *cd)
{
    if (cd->_dsf != 1) throw Error("export_warps: the OpenGL path renders the target at the views' size (dsf 1)");
    if (cd->_depth_name.empty() || cd->_cam_name.empty()) throw Error("export_warps needs depth_name and cam_name");
    if (cd->_s_rmv < cd->_s_min || cd->_s_rmv > cd->_s_max)
        throw Error("export_warps: s_rmv names the target view and must lie in [s_min, s_max]");
    const int nv = cd->_s_max - cd->_s_min + 1;
    std::vector<Image> depth(nv);
    std::vector<std::vector<double>> cams(nv);
    for (int i = 0; i < nv; i++) {
        const std::string dn = pattern_name(cd->_depth_name, cd->_s_min + i);
        if (!load_pfm(dn, depth[i])) throw Error("cannot read depth map " + dn);
        if (depth[i]._w != depth[0]._w || depth[i]._h != depth[0]._h) throw Error(dn + " has the wrong size");
        cams[i] = import_cam(cd->_cam_name, cd->_s_min + i);
    }
    const size_t w = depth[0]._w, h = depth[0]._h;
    lfrv_params p;
    lfrv_default_params(&p, LFRV_MODE_GRADIENT_IBR, nv, (int)w, (int)h, 1);
    lfrv_ctx *raw = nullptr;
    if (lfrv_create(&p, &raw) != LFRV_OK) throw Error("lfrv_create failed");
    std::unique_ptr<lfrv_ctx, void (*)(lfrv_ctx *)> ctx(raw, lfrv_destroy);
    for (int i = 0; i < nv; i++) check(ctx.get(), lfrv_set_depth(ctx.get(), i, depth[i]._r.data(), cams[i].data()), "depth");
    const int target = cd->_s_rmv - cd->_s_min;
    check(ctx.get(), lfrv_set_target_camera(ctx.get(), cams[target].data()), "target camera");
    check(ctx.get(), lfrv_warps_from_target_depth(ctx.get(), target), "lfrv_warps_from_target_depth");
    write_gl_warps(ctx.get(), cd, nv, (int)w, (int)h, p.sigma_z, [cd](int i) { return cd->_s_min + i; });
    std::printf("export_warps: %d views %zux%zu, target view %d\n", nv, w, h, cd->_s_rmv);
}

}  // namespace lfrv_ref
