/*
 * lfrv_ref.hpp -- the reference's host interface for the novel-view solve,
 * re-exported over the lfrv C-ABI (C++).
 *
 * Names, argument meaning and call order follow the reference so that
 * gradientIBR.cpp- and super-resolution.cpp-shaped drivers port line by line:
 *   IBR_gradient: gradientIBR.cuh:128-173 and solver.cuh:7
 *   IBR_color:    optimization.cuh:111-175 and vtv.h:85-103
 * The reference's coco::gsl_image / gsl_matrix / Config_data arguments are
 * replaced by lfrv_ref::Image (planar float RGB) and lfrv_ref::ConfigData
 * (Config_data's fields).  Where the reference asserts, these functions
 * throw lfrv_ref::Error carrying lfrv_last_error(); nothing aborts.
 * Built into liblfrv_host.so with the config / image front end and the
 * lfrv_cli driver (light-field-remote-vision_amd/host/).
 */
#ifndef LFRV_REF_HPP
#define LFRV_REF_HPP

#include <stdexcept>
#include <functional>
#include <string>
#include <vector>

#include "lfrv.h"
#include "lfrv_ref_config.hpp"
#include "lfrv_ref_image.hpp"

namespace lfrv_ref {

struct Error : std::runtime_error {
    explicit Error(const std::string &m) : std::runtime_error(m) {}
};

/* ---- IBR_gradient (gradientIBR.cuh:14-126) ------------------------------- */

/* The reference's Data: sizes and parameters as there; the device buffers are
 * owned by the lfrv context. */
struct Data {
    size_t _nviews = 0, _dsf = 1, _w = 0, _h = 0, _W = 0, _H = 0, _ks = 2;
    float _sigma_sensor = 1.0f / 255.0f, _ugrad_threshold = 1.0f;
    int _dw_type = 0, _gw_type = 0;
    float _dt_alpha = 1.0f, _dt_beta = 0.0f, _gradient_step = 0.1f;
    std::string _outdir, _init_name;
    int _normal_equations = 0;  /* ConfigData::_ibr_system == "normal_equations" (an extension) */
    lfrv_ctx *ctx = nullptr;
    lfrv_cg_stats cg{};  /* statistics of the last possion_conj_grad */
};

Data *init_data(ConfigData *config_data);                       /* gradientIBR.cu:155-299 */
void free_data(Data *data);                                     /* :302-348 */
void create_view(Data *data, size_t nview, const Image *I);     /* :524-555 */
void read_tau(Data *data, Image **tau_warps);                   /* :651-679 (R = tau_x, G = tau_y) */
void read_partial_tau(Data *data, Image **partial_tau);         /* :682-721 (R = sigma_z, G/B = dtau/dz) */
void compute_sparse_matrix(Data *data);                         /* :351-378 (+ the cell structure) */
void init_forward_warp_structure(Data *data, size_t nview);     /* :383-521 (folded into the above) */
void setup_target_visibility(Data *data);                       /* :40-83 */
void compute_weights(Data *data);                               /* :558-648 */
void compute_initial_image(Data *data);                         /* :724-819, hole filling included */
void hole_filling(Data *data);                                  /* :87-152 (run by compute_initial_image) */
void init_u_gradient(Data *data);                               /* :833-920 */
void possion_conj_grad(Data *data);                             /* solver.cu:526-699: (I + D^T D) u = u + D^T g by a
                                                                   corrected CG to 1e-5; the normal equations
                                                                   when _normal_equations */
void get_solution(Data *data, std::vector<std::vector<float>> &U); /* :20-37, U[c] of W*H */
void write_pfm_solution(Data *data);                            /* :822-830, <outdir>/output.pfm */

/* The driver of gradientIBR.cpp:12-150: light field, warps (tau / dpart PFMs,
 * or depth maps + cameras), the solve and the output files. */
void IBR_gradient(ConfigData *config_data);

/* ---- IBR_color (optimization.cuh:17-175, vtv.h) --------------------------- */

struct coco_vtv_data {
    size_t _W = 0, _H = 0, _nviews = 0, _dsf = 1;
    float _lambda = 0.0f, _lambda_max_factor = 0.0f;
    int _inner_iterations = 0;
    std::string _basedir;
    lfrv_ctx *ctx = nullptr;
};

coco_vtv_data *coco_vtv_alloc(size_t W, size_t H);                                /* vtv.cu:14-136 */
void coco_vtv_free(coco_vtv_data *data);
bool coco_vtv_sr_init_unstructured(coco_vtv_data *data, ConfigData *config_data);  /* optimization.cu:654-788 */
bool coco_vtv_sr_free_unstructured(coco_vtv_data *data);
bool coco_vtv_sr_create_view_unstructured(coco_vtv_data *data, size_t nview, const Image *I);
bool coco_vtv_sr_read_tau(coco_vtv_data *data, Image **tau_warps);
bool coco_vtv_sr_read_partial_tau(coco_vtv_data *data, Image **partial_tau);
bool coco_vtv_sr_compute_sparse_matrix(coco_vtv_data *data);
bool vtv_sr_init_forward_warp_structure_unstructured(coco_vtv_data *data, size_t nview);
bool coco_vtv_sr_compute_weights_unstructured(coco_vtv_data *data);
void coco_vtv_setup_visibility_mask(coco_vtv_data *data);
void coco_vtv_sr_compute_initial_image(coco_vtv_data *data, ConfigData *config_data);
double coco_vtv_sr_primal_energy_unstructured(coco_vtv_data *data, float previous_energy = 0.0f);
bool coco_vtv_sr_iteration_fista_unstructured(coco_vtv_data *data, bool data_term = true);
bool coco_vtv_sr_init_regularizer_weight_unstructured(coco_vtv_data *data);
void coco_vtv_push_pull(coco_vtv_data *data);
void coco_vtv_get_solution(coco_vtv_data *data, std::vector<std::vector<float>> &U);
bool coco_vtv_sr_write_pfm_solution(coco_vtv_data *data);

/* The driver of super-resolution.cpp:54-338 (meta-iterations of weights and
 * TV weight, FISTA steps with the energy / moving-pixel stopping rules). */
void sr_synthesize_view(ConfigData *config_data);

/* The warp export of the reference's OpenGL path, Scene::computeTargetDepthMap
 * + Scene::exportWarps (scene.cpp:1506-1585, view.cpp:663-721), on the GPU
 * (lfrv_warps_from_target_depth) from depth maps (depth_name) and cameras
 * (cam_name) at every index s_min..s_max, the target being view s_rmv (its own
 * depth map not splatted), dsf 1: writes tau_name (R tau_x, G tau_y, B the
 * deformation weight; -1 occluded, -2 outside), dpart_name (R sigma_z, G / B
 * dtau / dz) and uDepth_name (R the target depth map) where those are set. */
void export_warps(ConfigData *config_data);

/* The fromMVE2coco algorithm of the reference's command line (main.cpp:313-346,
 * the OpenGL path): the views of an MVE scene (mve_name/views/view_%04i.mve/:
 * undistorted.png, meta.ini, depth-L<scale>.mvei and dz-L<scale>.mvei for
 * scale_min..scale_max) for s_min..s_max as Scene::importViews reads them
 * (importScene.cpp:105-230; views lacking an image, a camera or a depth scale
 * are left out), their depth maps filtered from radial to orthographic depth on
 * the GPU (Scene::filterDepthMaps), then the target depth map and the warps of
 * export_warps with the target camera of view s_rmv.  Files are numbered by
 * the position among the imported views (scene.cpp:1546-1585): lf_name (the
 * view images), tau_name, dpart_name, and uDepth_name.  The reference also
 * asserts a mask file at a fixed path of its authors' machine
 * (importScene.cpp:117, 138), used only for a printed statistic; it is not read. */
void from_mve_to_coco(ConfigData *config_data);

/* An MVE camera (view.cpp:214-270): K (row-major; focal length and principal
 * point scaled by the view size w x h), R (row-major, as meta.ini lists it), t,
 * and P = K [R | t]. */
struct MveCamera {
    double K[9], R[9], t[3], P[12];
};
/* The [camera] section of an MVE meta.ini; false when the file is missing or
 * holds no camera (a [view] section first). */
bool import_mve_camera(const std::string &path, int w, int h, MveCamera &cam);
/* An MVE .mvei image of float channels (view.cpp:279-326): planes[c][y*w + x]. */
bool load_mvei(const std::string &path, int &w, int &h, std::vector<std::vector<float>> &planes);
/* The depth map of view s as importViews leaves it (view.cpp:607-634,
 * texture.cpp:218-278, addDepthScale.frag): the last scale of
 * scale_min..scale_max upsampled by nearest neighbour to w x h, its derivatives
 * divided by 2^scale, pixels whose three values are 0 set to (1000, 0, 0).
 * zdz: R depth, G dz/dx, B dz/dy.  False when a file is missing. */
bool import_mve_depth(const std::string &mve_name, int s, int scale_min, int scale_max, int w, int h, Image &zdz);
/* The file output of export_warps / from_mve_to_coco for a context whose warps
 * lfrv_warps_from_target_depth computed: view i's files take number file_index(i). */
void write_gl_warps(lfrv_ctx *ctx, ConfigData *config_data, int nviews, int w, int h, float sigma_z,
                    const std::function<int(int)> &file_index);

/* ---- front end helpers ------------------------------------------------------ */

/* import_lightfield (import.cpp:116-170): views s_min..s_max except s_rmv of
 * the lf_name pattern; fills _w, _h (if 0), _nviews, _W, _H. */
std::vector<Image> import_lightfield(ConfigData *config_data);
/* import_warps (import.cpp:173-206): the PFMs of a pattern for the same views. */
std::vector<Image> import_warps(ConfigData *config_data, const std::string &pattern);
/* import_cam (import.cpp:219-240): "CONTOUR" + 12 numbers, row-major 3x4. */
std::vector<double> import_cam(const std::string &pattern, int index);

}  // namespace lfrv_ref

#endif /* LFRV_REF_HPP */
