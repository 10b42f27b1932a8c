// The reference's quality tool (psnr/psnr_main.cpp:12-20):
//   lfrv_psnr_tool reference.png candidate.png
// prints "& PSNR & 10000 (1 - SSIM) / 2" as the reference's LaTeX-table
// harness reads it (TS.sh:223-268): PSNR of gsl_image_psnr (vmax 1, 2-px
// border excluded, RGB-averaged MSE) and SSIM of gsl_image_ssim (range 1).
// Images are PNG or PFM, read as the reference's front end reads them.
#include <cstdio>

#include "../../include/lfrv.h"
#include "../../include/lfrv_ref_image.hpp"

int main(int argc, char **argv)
{
    using namespace lfrv_ref;
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s reference candidate\n", argv[0]);
        return 2;
    }
    Image ref, cand;
    if (!load_image(argv[1], ref) || !load_image(argv[2], cand)) {
        std::fprintf(stderr, "lfrv_psnr_tool: cannot read %s or %s\n", argv[1], argv[2]);
        return 1;
    }
    if (ref._w != cand._w || ref._h != cand._h) {
        std::fprintf(stderr, "lfrv_psnr_tool: sizes differ\n");
        return 1;
    }
    const std::vector<float> a = ref.planar(), b = cand.planar();
    const int W = (int)ref._w, H = (int)ref._h;
    const double psnr = lfrv_psnr(W, H, a.data(), b.data());
    const double ssim = lfrv_ssim(W, H, a.data(), b.data(), 1.0);
    std::printf("& %.4g & %d\n", psnr, (int)(10000 * (1 - ssim) / 2));
    return 0;
}
