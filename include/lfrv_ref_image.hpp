// Image files of the reference's front end.
//
// PFM: "PF" (3 channels) / "Pf" (1 channel) header, little-endian when the
// scale is negative, rows stored bottom-up; loaded top-down as the reference
// does (gsl_image.cpp:416-502, import.cpp:695-775).  A 1-channel PFM loads
// into all three planes.
// PNG: 8-bit grey / grey+alpha / RGB / RGBA, non-interlaced; values map to
// v / 255 on load and to clamp(int(v * 255), 0, 255) on save, as the
// reference's QImage conversions (gsl_image.cpp:539-637).
#pragma once

#include <string>
#include <vector>

namespace lfrv_ref {

// Planar RGB image (the role of coco::gsl_image): r, g, b planes of w*h,
// rows top to bottom.
struct Image {
    size_t _w = 0, _h = 0;
    std::vector<float> _r, _g, _b;
    Image() = default;
    Image(size_t w, size_t h) : _w(w), _h(h), _r(w * h), _g(w * h), _b(w * h) {}
    std::vector<float> planar() const;        // RRR..GGG..BBB..
    void set_planar(const float *rgb);
};

bool load_pfm(const std::string &path, Image &img, int *channels = nullptr);
bool save_pfm(const std::string &path, const Image &img);
bool save_pfm_gray(const std::string &path, size_t w, size_t h, const float *v);
bool load_png(const std::string &path, Image &img);
bool save_png(const std::string &path, const Image &img);
// by extension: .pfm or .png
bool load_image(const std::string &path, Image &img);
bool save_image(const std::string &path, const Image &img);

// printf-style file pattern with one integer (v_%02i.png)
std::string pattern_name(const std::string &pattern, int index);

}  // namespace lfrv_ref
