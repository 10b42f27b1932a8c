/*
 * lfrv_io.h -- C-ABI of the reference-format front end (liblfrv_host.so).
 *
 * The file formats and configuration semantics the reference's drivers read
 * and write, for callers that drive the solver through lfrv.h directly and
 * for tests.  No GPU work happens here except in lfrv_io_export_warps.
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

#ifdef __cplusplus
}
#endif

#endif /* LFRV_IO_H */
