// Quality metric on the device: the PSNR of gsl_image_psnr
// (cocolib/common/gsl_image.cpp:991-1025) over two device-resident planar
// RGB images, for callers that keep their solutions on the GPU.
//
// One pass: block b sums (a - b)^2 over its chunk of the interior pixels
// (x, y in [2, W-2) x [2, H-2), all three channels) in FP64 and stores one
// partial; a one-block finalize sums the partials in block order and writes
// 10 log10(1 / mse) (DBL_MAX when mse = 0) to the caller's device double.
// The partials live in stream-ordered memory the call allocates and frees
// on the caller's stream.
#include <cfloat>

#include "lfrv_internal.hpp"

namespace lfrv {

constexpr int kPsnrThreads = 256;
constexpr int kPsnrPerThread = 16;
static_assert(kPsnrThreads * kPsnrPerThread == kPsnrChunk, "interior pixels per block");

__global__ void __launch_bounds__(kPsnrThreads)
k_psnr_partials(int W, int H, const float *__restrict__ a, const float *__restrict__ b, double *partials)
{
    const int wi = W - 4;
    const int64_t nint = (int64_t)wi * (H - 4), N = (int64_t)W * H;
    double s = 0.0;
    const int64_t base = (int64_t)blockIdx.x * kPsnrChunk;
#pragma unroll 4
    for (int k = 0; k < kPsnrPerThread; k++) {
        const int64_t i = base + (int64_t)k * kPsnrThreads + threadIdx.x;  // consecutive lanes, consecutive pixels
        if (i < nint) {
            const int y = (int)(i / wi) + 2, x = (int)(i % wi) + 2;
            const int64_t o = (int64_t)y * W + x;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double d = (double)a[c * N + o] - (double)b[c * N + o];
                s += d * d;
            }
        }
    }
    __shared__ double red[16];
    double v[1] = {s};
    block_partials<1>(v, partials, red);
}

__global__ void __launch_bounds__(kFinalizeThreads)
k_psnr_finalize(const double *__restrict__ partials, int G, double nint3, double *out)
{
    __shared__ double red[16];
    double v[1];
    reduce_partials<1>(partials, G, v, red);
    if (threadIdx.x == 0) {
        const double mse = v[0] / nint3;
        out[0] = mse == 0.0 ? DBL_MAX : 10.0 * log(1.0 / mse) / log(10.0);
    }
}

int psnr_blocks(int W, int H)
{
    const int64_t nint = (int64_t)(W - 4) * (H - 4);
    return (int)((nint + kPsnrChunk - 1) / kPsnrChunk);
}

lfrv_status psnr_launch(int W, int H, const float *a, const float *b, double *out, double *partials,
                               int G, hipStream_t s)
{
    k_psnr_partials<<<G, kPsnrThreads, 0, s>>>(W, H, a, b, partials);
    k_psnr_finalize<<<1, kFinalizeThreads, 0, s>>>(partials, G, 3.0 * (double)(W - 4) * (double)(H - 4), out);
    return hipGetLastError() == hipSuccess ? LFRV_OK : LFRV_E_HIP;
}

}  // namespace lfrv

using namespace lfrv;

extern "C" lfrv_status lfrv_op_psnr(int W, int H, const float *a, const float *b, double *out, void *stream)
{
    if (W <= 4 || H <= 4 || !a || !b || !out || (int64_t)W * H > (int64_t)1 << 31) return LFRV_E_ARG;
    const hipStream_t s = (hipStream_t)stream;
    const int G = psnr_blocks(W, H);
    double *partials = nullptr;
    if (hipMallocAsync((void **)&partials, (size_t)G * sizeof(double), s) != hipSuccess) return LFRV_E_HIP;
    lfrv_status st = psnr_launch(W, H, a, b, out, partials, G, s);
    if (hipFreeAsync(partials, s) != hipSuccess && st == LFRV_OK) st = LFRV_E_HIP;
    return st;
}
