/*
 * lfrv_io.h -- C-ABI of the reference-format front end (liblfrv_host.so).
 *
 * The file formats and configuration semantics the reference's drivers read
 * and write, for callers that drive the solver through lfrv.h directly and
 * for tests.  No GPU work happens here except in lfrv_io_export_warps and
 * lfrv_io_from_mve.
 *
 *   images: PFM (gsl_image.cpp:416-502, import.cpp:695-775) and 8-bit PNG
 *           (gsl_image.cpp:539-637), planar float RGB, rows top to bottom
 *   config: coco::config (cocolib/common/parse_config.cpp:43-196) and
 *           Config_data::parse_config (sources/config.cpp:67-128)
 */
#ifndef LFRV_IO_H
#define LFRV_IO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Load a .pfm or .png file into planar RGB (rgb holds 3*w*h floats).  With
 * rgb == NULL only *w and *h are set.  Returns 0 on success, -1 on failure. */
int lfrv_io_load_image(const char *path, size_t *w, size_t *h, float *rgb);

/* Save planar RGB of w*h as .pfm (little-endian, bottom-up rows) or .png
 * (clamp(int(v*255), 0, 255)), chosen by the file extension.  0 / -1. */
int lfrv_io_save_image(const char *path, size_t w, size_t h, const float *rgb);

/* Parse a reference command line (argv[0] is the program name) into
 * Config_data and print its fields as "name value" lines into out (cap
 * bytes, NUL-terminated).  Returns the length needed, or -1. */
int lfrv_io_config_data(int argc, const char *const *argv, char *out, size_t cap);

/* lfrv_ref::export_warps for a reference command line (argv[0] the program
 * name; -config file and switches as lfrv_cli takes them): the OpenGL path's
 * target depth map and warps computed on the GPU and written as PFMs.  The one
 * entry point here that uses the GPU.  Returns 0, or -1 with the reason on
 * stderr. */
int lfrv_io_export_warps(int argc, const char *const *argv);

/* lfrv_ref::from_mve_to_coco for a reference command line (the fromMVE2coco
 * algorithm, main.cpp:313-346): an MVE scene's views, cameras and depth maps,
 * the depth filter and the warps on the GPU, written in the coco formats.
 * Returns 0, or -1 with the reason on stderr. */
int lfrv_io_from_mve(int argc, const char *const *argv);

/* The [camera] of an MVE meta.ini for views of w x h (view.cpp:214-270):
 * K, R row-major 3x3, t, P = K [R | t] row-major 3x4 (each nullable).  0 / -1. */
int lfrv_io_mve_camera(const char *meta_ini, int w, int h, double *K, double *R, double *t, double *P);

/* View s's depth map of an MVE scene as Scene::importViews leaves it before the
 * depth filter (view.cpp:607-634, texture.cpp:218-278, addDepthScale.frag):
 * w*h floats each (nullable).  0 / -1. */
int lfrv_io_mve_depth(const char *mve_name, int s, int scale_min, int scale_max, int w, int h, float *z,
                      float *dzdx, float *dzdy);

#ifdef __cplusplus
}
#endif

#endif /* LFRV_IO_H */
