// lfrv_op_psnr_noalloc: the device PSNR of lfrv_op_psnr (k_quality.hip) with
// its per-block partial sums in a fixed device array instead of memory it
// allocates, for callers that cannot allocate (code under HIP graph capture,
// serving loops that forbid hipMalloc).
//
// The array `scratch` holds kPsnrScratchBlocks doubles (128 KiB); every GPU
// the library runs on keeps its own copy until the process ends.  Its device
// address is taken once per device with hipGetSymbolAddress, with that
// device current, and only that address reaches the kernels (in host code
// the name is a stand-in, not a device pointer).  Calls on one GPU share
// the array: calls that overlap on one GPU (two streams, or two host
// threads) race on it.  Every partial the finalize reads is written by the
// same call's first kernel, so nothing needs clearing between calls.
#include <mutex>

#include "lfrv_internal.hpp"

namespace lfrv {

constexpr int kPsnrScratchBlocks = 16384;  // partials: images up to 16384 * 4096 interior pixels

static __device__ __align__(256) double scratch[kPsnrScratchBlocks];

constexpr int kMaxDevices = 64;
static double *g_scratch_ptr[kMaxDevices];
static hipError_t g_scratch_err[kMaxDevices];
static std::once_flag g_scratch_once[kMaxDevices];

// The device address of this device's `scratch` (nullptr and *err set when
// hipGetSymbolAddress fails); taken on the first call on each device.
static double *scratch_for(int dev, hipError_t *err)
{
    std::call_once(g_scratch_once[dev], [dev]() {
        void *p = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_ptr[dev] = g_scratch_err[dev] == hipSuccess ? static_cast<double *>(p) : nullptr;
    });
    *err = g_scratch_err[dev];
    return g_scratch_ptr[dev];
}

}  // namespace lfrv

using namespace lfrv;

extern "C" lfrv_status lfrv_op_psnr_noalloc(int W, int H, const float *a, const float *b, double *out,
                                            void *stream)
{
    if (W <= 4 || H <= 4 || !a || !b || !out) return LFRV_E_ARG;
    if ((int64_t)(W - 4) * (H - 4) > (int64_t)kPsnrScratchBlocks * kPsnrChunk) return LFRV_E_ARG;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return LFRV_E_HIP;
    hipError_t e = hipSuccess;
    double *partials = scratch_for(dev, &e);
    if (!partials) return LFRV_E_HIP;
    return psnr_launch(W, H, a, b, out, partials, psnr_blocks(W, H), (hipStream_t)stream);
}
