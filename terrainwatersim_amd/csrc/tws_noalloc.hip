// tws_noalloc.hip -- entry points that allocate no device memory, for callers
// that cannot call hipMalloc (code under HIP graph capture, tight serving
// loops, harnesses that forbid it).
//
// tws_volume_noalloc computes what tws_volume computes for a handle's grid,
// with the same kernel (k_volume) and the same fixed-order reduction, but
// keeps the kernel's partial sums in `scratch`, a file-scope device array of
// the library, instead of a buffer of the handle.  Each device holds its own
// copy of the array (at its own address) until the process ends.  Host code
// never passes the array's name to a kernel or copy: it resolves the
// device's address once with hipGetSymbolAddress and uses that pointer.

#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "tws.h"
#include "tws_internal.hpp"
#include "tws_runtime.hpp"

namespace tws {

// The partial sums of one volume reduction: every call writes all
// kVolumeParts of them before reading any, so nothing needs clearing.
static __device__ __align__(256) double scratch[kVolumeParts];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, resolved on the first call on
// that device (with that device current) and kept.
static int scratch_on_current_device(double** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "noalloc: hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices)
        return fail(TWS_ERR_DEVICE, "noalloc: device %d beyond the %d supported", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "noalloc: hipGetSymbolAddress(scratch)");
        if (!p) return fail(TWS_ERR_DEVICE, "noalloc: scratch has no device address");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return TWS_OK;
}

}  // namespace tws

using namespace tws;

extern "C" {

int tws_volume_noalloc(const float* depth, int32_t width, int32_t height, int64_t pitch,
                       void* stream, double* out) {
    if (!depth || !out) return fail(TWS_ERR_INVALID, "tws_volume_noalloc: NULL argument");
    if (width <= 0 || height <= 0 || width > (1 << 20) || height > (1 << 20) || pitch < width ||
        pitch > (int64_t)1 << 30)
        return fail(TWS_ERR_INVALID,
                    "tws_volume_noalloc: invalid plane %d x %d, pitch %lld (width and height "
                    "1..2^20, width <= pitch <= 2^30)", width, height, (long long)pitch);
    double* partials = nullptr;
    int rc = scratch_on_current_device(&partials);
    if (rc != TWS_OK) return rc;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    // k_volume reads d[parity], W, H and pitch of its descriptor and nothing else
    GridDesc g{};
    g.d[0] = const_cast<float*>(depth);
    g.W = width;
    g.H = height;
    g.pitch = pitch;
    hipError_t e = launch_volume_partials(g, 0, partials, kVolumeParts, s);
    if (e != hipSuccess) return hip_fail(e, "tws_volume_noalloc: kernel");
    double host[kVolumeParts];
    e = hipMemcpyAsync(host, partials, sizeof(host), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "tws_volume_noalloc: copy");
    double acc = 0.0;
    for (int i = 0; i < kVolumeParts; ++i) acc += host[i];
    *out = acc;
    return TWS_OK;
}

}  // extern "C"
