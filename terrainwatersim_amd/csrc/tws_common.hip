// tws_common.hip -- error reporting, parameter derivation and diagnostics of
// the C ABI (tws.h).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>

#include "tws.h"
#include "tws_internal.hpp"
#include "tws_runtime.hpp"

namespace tws {

// A fixed buffer: recording an error allocates nothing and cannot throw.
static thread_local char g_last_error[512];

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(e == hipErrorOutOfMemory ? TWS_ERR_NOMEM : TWS_ERR_DEVICE, "%s: %s", what,
                hipGetErrorString(e));
}

bool valid_params(const tws_params* p) {
    if (!p) return false;
    if (!std::isfinite(p->flow_friction) || !std::isfinite(p->water_acceleration) ||
        !std::isfinite(p->cell_area_inv_timescaled) || !std::isfinite(p->ghost_alpha))
        return false;
    // the reference derives all three from non-negative quantities
    if (p->flow_friction < 0.0f || p->water_acceleration < 0.0f ||
        p->cell_area_inv_timescaled < 0.0f)
        return false;
    // Closed walls: a wall ghost of height 1e30 emits max(0, (1e30 - h) * K),
    // scaled to 0 by its zero depth; that flux must stay finite (Inf * 0 is
    // NaN), so K is bounded there.
    if (p->boundary == TWS_BOUNDARY_CLOSED) return p->water_acceleration <= kMaxClosedK;
    return p->boundary == TWS_BOUNDARY_ZERO || p->boundary == TWS_BOUNDARY_REFERENCE_DISPATCH;
}

void fill_ghosts(GridDesc& g, const tws_params& p) {
    g.boundary = p.boundary;
    g.F = p.flow_friction;
    g.K = p.water_acceleration;
    g.C = p.cell_area_inv_timescaled;
    if (p.boundary == TWS_BOUNDARY_CLOSED) {
        g.ghost_b = kWallHeight;
        g.ghost_d = 0.0f;
        g.ghost_fw = 0.0f;
    } else {
        g.ghost_b = 0.0f;
        g.ghost_d = p.ghost_alpha;
        g.ghost_fw = p.ghost_alpha;
    }
}

int64_t pitch_for(int width) { return ((int64_t)width + 63) / 64 * 64; }

int device_cus() {
    static std::mutex mu;
    static int cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    std::lock_guard<std::mutex> lk(mu);
    if (!cached[dev]) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            return 0;
        cached[dev] = cus;
    }
    return cached[dev];
}

}  // namespace tws

using namespace tws;

#define TWS_HIP(call, what)                              \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

extern "C" {

int tws_abi_version(void) { return TWS_ABI_VERSION; }

const char* tws_last_error(void) { return tws::g_last_error; }

int tws_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int tws_params_physical(float dt, float gravity, float cell_size, float pipe_area, float damping,
                        int32_t boundary, tws_params* out) {
    if (!out) return fail(TWS_ERR_INVALID, "tws_params_physical: out is NULL");
    if (!(dt > 0.0f) || !(cell_size > 0.0f) || !std::isfinite(gravity) ||
        !std::isfinite(pipe_area) || !(damping >= 0.0f))
        return fail(TWS_ERR_INVALID, "tws_params_physical: invalid physical parameters");
    if (boundary != TWS_BOUNDARY_CLOSED && boundary != TWS_BOUNDARY_ZERO &&
        boundary != TWS_BOUNDARY_REFERENCE_DISPATCH)
        return fail(TWS_ERR_INVALID, "tws_params_physical: unknown boundary %d", boundary);
    tws_params p{};
    p.flow_friction = (float)std::pow((double)damping, (double)dt);
    p.water_acceleration =
        (float)((double)dt * (double)pipe_area * (double)gravity / (double)cell_size);
    p.cell_area_inv_timescaled = (float)((double)dt / ((double)cell_size * (double)cell_size));
    p.boundary = boundary;
    p.ghost_alpha = 0.0f;
    *out = p;
    return TWS_OK;
}

int tws_params_reference(float steps_per_second, float damping, float acceleration,
                         float grid_world_size, uint32_t grid_resolution, int32_t boundary,
                         tws_params* out) {
    if (!out) return fail(TWS_ERR_INVALID, "tws_params_reference: out is NULL");
    if (!(steps_per_second > 0.0f) || grid_resolution == 0 || !(grid_world_size > 0.0f))
        return fail(TWS_ERR_INVALID, "tws_params_reference: invalid settings");
    if (boundary != TWS_BOUNDARY_CLOSED && boundary != TWS_BOUNDARY_ZERO &&
        boundary != TWS_BOUNDARY_REFERENCE_DISPATCH)
        return fail(TWS_ERR_INVALID, "tws_params_reference: unknown boundary %d", boundary);
    // Terrain.cpp:177 -- the step length is the float 1/steps, held as double.
    const double dt = (double)(1.0f / steps_per_second);
    const float cell = grid_world_size / (float)grid_resolution;  // :183, :196
    tws_params p{};
    p.flow_friction = powf(damping, (float)dt);                                 // :190
    p.water_acceleration = (float)(dt * (double)acceleration * (double)cell);   // :197
    p.cell_area_inv_timescaled = (float)(dt / (double)(cell * cell));           // :184
    p.boundary = boundary;
    p.ghost_alpha = 0.0f;
    *out = p;
    return TWS_OK;
}

int tws_float_to_half_rtz(const float* in, uint16_t* out, int64_t n) {
    if ((!in || !out) && n > 0) return fail(TWS_ERR_INVALID, "tws_float_to_half_rtz: NULL");
    for (int64_t i = 0; i < n; ++i) out[i] = half_rtz(in[i]);
    return TWS_OK;
}

int tws_stream_copy_bench(int64_t bytes, int32_t iterations, int32_t shape, double* gbps) {
    if (!gbps || bytes < (1 << 20) || iterations <= 0 || shape < -1 ||
        shape >= stream_copy_shapes())
        return fail(TWS_ERR_INVALID, "tws_stream_copy_bench: invalid argument");
    bytes &= ~(int64_t)255;
    void *src = nullptr, *dst = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&src, (size_t)bytes);
    if (e == hipSuccess) e = hipMalloc(&dst, (size_t)bytes);
    if (e == hipSuccess) e = hipMemset(src, 0, (size_t)bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const int first = shape < 0 ? 0 : shape, last = shape < 0 ? stream_copy_shapes() - 1 : shape;
    double best = 0.0;
    for (int sh = first; e == hipSuccess && sh <= last; ++sh) {
        e = launch_stream_copy(src, dst, bytes, sh, nullptr);  // warm
        if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
        for (int i = 0; e == hipSuccess && i < iterations; ++i)
            e = launch_stream_copy(src, dst, bytes, sh, nullptr);
        if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && ms > 0.0f)
            best = std::max(best, 2.0 * (double)bytes * iterations / (ms * 1e-3) / 1e9);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (e != hipSuccess) return hip_fail(e, "tws_stream_copy_bench");
    *gbps = best;
    return TWS_OK;
}

int tws_stream_copy_shapes(void) { return stream_copy_shapes(); }

int tws_debug_lane_shift_probe(int32_t* out128) {
    if (!out128) return fail(TWS_ERR_INVALID, "tws_debug_lane_shift_probe: NULL");
    int* dev = nullptr;
    TWS_HIP(hipMalloc(&dev, 128 * sizeof(int)), "probe: hipMalloc");
    hipError_t e = launch_lane_shift_probe(dev, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out128, dev, 128 * sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(dev);
    if (e != hipSuccess) return hip_fail(e, "probe");
    return TWS_OK;
}

}  // extern "C"
