// Structural similarity of the reference's quality harness (psnr tool,
// psnr/psnr_main.cpp:12-20): gsl_image_ssim (cocolib/common/gsl_image.cpp:
// 1027-1037) = the mean over R, G, B of gsl_matrix_ssim
// (cocolib/common/gsl_matrix_helper.cpp:799-914), in FP64 on the host.
//
// Per channel: mu = G * A, G * B and sigma from G * A^2, G * B^2, G * AB,
// with G the separable 11-tap Gaussian of sigma 1.5 (gsl_kernel_gauss_1xn,
// gsl_matrix_convolutions.cpp:42-63) applied along x then y, each output
// normalised by the taps that fall inside the image
// (gsl_matrix_convolution_1xn / _nx1, :396-472).  The map is averaged over
// every pixel.  As the reference computes it, the two factors of the SSIM
// numerator and of its denominator are ADDED (gsl_matrix_add_scaled with
// weights 1, 1; the comments there say ".*"):
//     ((2 mu1 mu2 + C1) + (2 s12 + C2)) / ((mu1^2 + mu2^2 + C1) + (s1^2 + s2^2 + C2)),
// C1 = (0.01 range)^2, C2 = (0.03 range)^2.  That sum is reproduced, so the
// harness prints the reference's numbers.
#include <cmath>
#include <vector>

#include "../../include/lfrv.h"

namespace {

std::vector<double> gauss_taps(int n, double sigma)
{
    std::vector<double> k(n);
    const int o = n / 2;
    double s = 0.0;
    for (int x = -o; x <= o; x++) s += (k[x + o] = std::exp(-0.5 * x * x / (sigma * sigma)));
    for (double &v : k) v /= s;
    return k;
}

// Separable filter along x then y, renormalised by the in-image taps.
void gauss_filter(int W, int H, const std::vector<double> &k, const double *in, double *out)
{
    const int o = (int)k.size() / 2;
    std::vector<double> t((size_t)W * H);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            double s = 0.0, wsum = 0.0;
            for (int i = -o; i <= o; i++) {
                const int xx = x + i;
                if (xx < 0 || xx >= W) continue;
                s += k[i + o] * in[(size_t)y * W + xx];
                wsum += k[i + o];
            }
            t[(size_t)y * W + x] = s / wsum;
        }
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            double s = 0.0, wsum = 0.0;
            for (int i = -o; i <= o; i++) {
                const int yy = y + i;
                if (yy < 0 || yy >= H) continue;
                s += k[i + o] * t[(size_t)yy * W + x];
                wsum += k[i + o];
            }
            out[(size_t)y * W + x] = s / wsum;
        }
}

double ssim_channel(int W, int H, const float *a, const float *b, double range)
{
    const double C1 = std::pow(0.01 * range, 2.0), C2 = std::pow(0.03 * range, 2.0);
    const size_t N = (size_t)W * H;
    const std::vector<double> k = gauss_taps(11, 1.5);
    std::vector<double> A(N), B(N), AA(N), BB(N), AB(N);
    for (size_t i = 0; i < N; i++) {
        A[i] = a[i];
        B[i] = b[i];
        AA[i] = A[i] * A[i];
        BB[i] = B[i] * B[i];
        AB[i] = A[i] * B[i];
    }
    std::vector<double> mu1(N), mu2(N), s11(N), s22(N), s12(N);
    gauss_filter(W, H, k, A.data(), mu1.data());
    gauss_filter(W, H, k, B.data(), mu2.data());
    gauss_filter(W, H, k, AA.data(), s11.data());
    gauss_filter(W, H, k, BB.data(), s22.data());
    gauss_filter(W, H, k, AB.data(), s12.data());
    double sum = 0.0;
    for (size_t i = 0; i < N; i++) {
        const double m11 = mu1[i] * mu1[i], m22 = mu2[i] * mu2[i], m12 = mu1[i] * mu2[i];
        const double v1 = s11[i] - m11, v2 = s22[i] - m22, v12 = s12[i] - m12;
        const double num = (2.0 * m12 + C1) + (2.0 * v12 + C2);
        const double den = (m11 + m22 + C1) + (v1 + v2 + C2);
        sum += num / den;
    }
    return sum / (double)N;
}

}  // namespace

extern "C" double lfrv_ssim(int W, int H, const float *a, const float *b, double dynamic_range)
{
    if (W <= 0 || H <= 0 || !a || !b) return 0.0;
    const size_t N = (size_t)W * H;
    double s = 0.0;
    for (int c = 0; c < 3; c++) s += ssim_channel(W, H, a + c * N, b + c * N, dynamic_range);
    return s / 3.0;
}
