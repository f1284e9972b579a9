// tws_api.hip -- C ABI of one simulation (tws.h): device state, uploads and
// downloads, parameter derivation and the step planner.
//
// Device layout of one simulation: one allocation of 11 planes, each
// height x pitch floats with pitch = width rounded up to 64 (256 B rows):
//   b | d0 | d1 | f0[+x,-x,+y,-y] | f1[+x,-x,+y,-y]
// d and f are ping-pong sets; `parity` names the set holding the current
// state.  The idle set doubles as staging for interleaved host layouts, so no
// call after tws_create allocates device memory (the velocity download
// included).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "tws.h"
#include "tws_internal.hpp"
#include "tws_runtime.hpp"

using namespace tws;

namespace tws {

static thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
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
    if (p->boundary != TWS_BOUNDARY_CLOSED && p->boundary != TWS_BOUNDARY_ZERO) return false;
    return true;
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

}  // namespace tws

struct tws_sim {
    int W = 0, H = 0;
    int64_t pitch = 0;
    tws_params params{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;
    float* base = nullptr;      // 11 planes
    float* b = nullptr;
    float* d[2] = {nullptr, nullptr};
    float* f[2][4] = {};
    unsigned* status = nullptr;  // device flag word
    double* partials = nullptr;  // device, kVolumeParts
    int parity = 0;
    int rain_on = 0;
    float rain[4] = {0, 0, 0, 0};
    int kernel = TWS_KERNEL_TILE_128x128;
    int k = 4;
    bool auto_launch = true;
    double time_acc = 0.0;  // PerformSimulationStep accumulator
};

static GridDesc desc_of(const tws_sim* s) {
    GridDesc g{};
    g.b = s->b;
    for (int p = 0; p < 2; ++p) {
        g.d[p] = s->d[p];
        for (int c = 0; c < 4; ++c) g.f[p][c] = s->f[p][c];
    }
    g.pitch = s->pitch;
    g.W = s->W;
    g.H = s->H;
    fill_ghosts(g, s->params);
    g.rain_on = s->rain_on;
    g.rain_cx = s->rain[0];
    g.rain_cy = s->rain[1];
    g.rain_I = s->rain[2];
    g.rain_R2 = s->rain[3];
    g.status = s->status;
    g.parity0 = s->parity;
    return g;
}

static float* staging_of(tws_sim* s) { return s->f[s->parity ^ 1][0]; }

#define TWS_HIP(call, what)                                  \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return hip_fail(e_, what);     \
    } while (0)

extern "C" {

int tws_abi_version(void) { return TWS_ABI_VERSION; }

const char* tws_last_error(void) { return tws::g_last_error.c_str(); }

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
    if (boundary != TWS_BOUNDARY_CLOSED && boundary != TWS_BOUNDARY_ZERO)
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
    if (boundary != TWS_BOUNDARY_CLOSED && boundary != TWS_BOUNDARY_ZERO)
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

int tws_create(int32_t width, int32_t height, const tws_params* params, void* stream,
               tws_sim** out) {
    if (!out) return fail(TWS_ERR_INVALID, "tws_create: out is NULL");
    *out = nullptr;
    if (width <= 0 || height <= 0 || width > (1 << 20) || height > (1 << 20))
        return fail(TWS_ERR_INVALID, "tws_create: invalid size %d x %d", width, height);
    if (!valid_params(params)) return fail(TWS_ERR_INVALID, "tws_create: invalid params");
    tws_sim* s = new (std::nothrow) tws_sim();
    if (!s) return fail(TWS_ERR_NOMEM, "tws_create: host allocation failed");
    s->W = width;
    s->H = height;
    s->pitch = pitch_for(width);
    s->params = *params;
    hipError_t e = hipGetDevice(&s->device);
    if (e != hipSuccess) { delete s; return hip_fail(e, "tws_create: hipGetDevice"); }
    if (stream) {
        s->stream = (hipStream_t)stream;
    } else {
        e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
        if (e != hipSuccess) { delete s; return hip_fail(e, "tws_create: hipStreamCreate"); }
        s->own_stream = true;
    }
    const int64_t plane = s->pitch * (int64_t)height;
    e = hipMalloc(&s->base, sizeof(float) * (size_t)(plane * 11));
    if (e != hipSuccess) { tws_destroy(s); return hip_fail(e, "tws_create: hipMalloc state"); }
    e = hipMalloc(&s->partials, sizeof(double) * kVolumeParts + 256);
    if (e != hipSuccess) { tws_destroy(s); return hip_fail(e, "tws_create: hipMalloc aux"); }
    s->status = reinterpret_cast<unsigned*>(s->partials + kVolumeParts);
    s->b = s->base;
    s->d[0] = s->base + plane;
    s->d[1] = s->base + 2 * plane;
    for (int p = 0; p < 2; ++p)
        for (int c = 0; c < 4; ++c) s->f[p][c] = s->base + (3 + 4 * p + c) * plane;
    e = launch_zero_fill(s->base, plane * 11, s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(s->status, 0, 256, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) { tws_destroy(s); return hip_fail(e, "tws_create: clear state"); }
    choose_launch(s->W, s->H, &s->kernel, &s->k);
    *out = s;
    return TWS_OK;
}

int tws_destroy(tws_sim* s) {
    if (!s) return TWS_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->base) (void)hipFree(s->base);
    if (s->partials) (void)hipFree(s->partials);
    if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return TWS_OK;
}

int tws_set_params(tws_sim* s, const tws_params* p) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_set_params: NULL handle");
    if (!valid_params(p)) return fail(TWS_ERR_INVALID, "tws_set_params: invalid params");
    s->params = *p;
    return TWS_OK;
}

int tws_get_params(const tws_sim* s, tws_params* out) {
    if (!s || !out) return fail(TWS_ERR_INVALID, "tws_get_params: NULL argument");
    *out = s->params;
    return TWS_OK;
}

int tws_upload(tws_sim* s, const float* terrain, const float* depth, const float* flux4) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_upload: NULL handle");
    const size_t row = sizeof(float) * (size_t)s->W, dpitch = sizeof(float) * (size_t)s->pitch;
    if (terrain)
        TWS_HIP(hipMemcpy2DAsync(s->b, dpitch, terrain, row, row, s->H, hipMemcpyHostToDevice,
                                 s->stream), "tws_upload: terrain");
    if (depth)
        TWS_HIP(hipMemcpy2DAsync(s->d[s->parity], dpitch, depth, row, row, s->H,
                                 hipMemcpyHostToDevice, s->stream), "tws_upload: depth");
    const GridDesc g = desc_of(s);
    if (flux4) {
        float* st = staging_of(s);
        TWS_HIP(hipMemcpyAsync(st, flux4, sizeof(float) * 4 * (size_t)s->W * s->H,
                               hipMemcpyHostToDevice, s->stream), "tws_upload: flux");
        TWS_HIP(launch_deinterleave(g, s->parity, st, s->stream), "tws_upload: deinterleave");
    } else {
        TWS_HIP(launch_zero_fill(s->f[s->parity][0], 4 * s->pitch * (int64_t)s->H, s->stream),
                "tws_upload: clear flux");
    }
    TWS_HIP(launch_check_finite(g, s->parity, s->stream), "tws_upload: finite check");
    TWS_HIP(hipStreamSynchronize(s->stream), "tws_upload: sync");
    return TWS_OK;
}

int tws_download(tws_sim* s, float* depth, float* flux4, float* velocity2) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_download: NULL handle");
    const size_t row = sizeof(float) * (size_t)s->W, spitch = sizeof(float) * (size_t)s->pitch;
    const GridDesc g = desc_of(s);
    float* st = staging_of(s);
    if (depth) {
        TWS_HIP(hipMemcpy2DAsync(depth, row, s->d[s->parity], spitch, row, s->H,
                                 hipMemcpyDeviceToHost, s->stream), "tws_download: depth");
    }
    if (flux4) {
        TWS_HIP(launch_interleave(g, s->parity, st, s->stream), "tws_download: interleave");
        TWS_HIP(hipMemcpyAsync(flux4, st, sizeof(float) * 4 * (size_t)s->W * s->H,
                               hipMemcpyDeviceToHost, s->stream), "tws_download: flux");
        TWS_HIP(hipStreamSynchronize(s->stream), "tws_download: sync");
    }
    if (velocity2) {
        TWS_HIP(launch_velocity(g, s->parity, st, s->stream), "tws_download: velocity");
        TWS_HIP(hipMemcpyAsync(velocity2, st, sizeof(float) * 2 * (size_t)s->W * s->H,
                               hipMemcpyDeviceToHost, s->stream), "tws_download: velocity copy");
    }
    TWS_HIP(hipStreamSynchronize(s->stream), "tws_download: sync");
    return TWS_OK;
}

int tws_upload_reference_layout(tws_sim* s, const float* terrain_rgba32f,
                                const float* flow_rgba32f) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_upload_reference_layout: NULL handle");
    const GridDesc g = desc_of(s);
    float* st = staging_of(s);
    const size_t bytes = sizeof(float) * 4 * (size_t)s->W * s->H;
    if (terrain_rgba32f) {
        TWS_HIP(hipMemcpyAsync(st, terrain_rgba32f, bytes, hipMemcpyHostToDevice, s->stream),
                "tws_upload_reference_layout: terrain");
        TWS_HIP(launch_ref_unpack(g, s->parity, st, nullptr, s->b, s->stream),
                "tws_upload_reference_layout: unpack terrain");
    }
    if (flow_rgba32f) {
        TWS_HIP(hipMemcpyAsync(st, flow_rgba32f, bytes, hipMemcpyHostToDevice, s->stream),
                "tws_upload_reference_layout: flow");
        TWS_HIP(launch_ref_unpack(g, s->parity, nullptr, st, s->b, s->stream),
                "tws_upload_reference_layout: unpack flow");
    }
    TWS_HIP(launch_check_finite(g, s->parity, s->stream), "tws_upload_reference_layout: check");
    TWS_HIP(hipStreamSynchronize(s->stream), "tws_upload_reference_layout: sync");
    return TWS_OK;
}

int tws_download_reference_layout(tws_sim* s, float* terrain_rgba32f, float gb_fill,
                                  float* flow_rgba32f, uint16_t* flowmap_rg16f) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_download_reference_layout: NULL handle");
    const GridDesc g = desc_of(s);
    float* st = staging_of(s);
    const size_t cells = (size_t)s->W * s->H;
    if (terrain_rgba32f) {
        TWS_HIP(launch_ref_pack(g, s->parity, gb_fill, st, nullptr, nullptr, s->stream),
                "tws_download_reference_layout: pack terrain");
        TWS_HIP(hipMemcpyAsync(terrain_rgba32f, st, sizeof(float) * 4 * cells,
                               hipMemcpyDeviceToHost, s->stream), "copy terrain");
        TWS_HIP(hipStreamSynchronize(s->stream), "tws_download_reference_layout: sync");
    }
    if (flow_rgba32f) {
        TWS_HIP(launch_ref_pack(g, s->parity, gb_fill, nullptr, st, nullptr, s->stream),
                "tws_download_reference_layout: pack flow");
        TWS_HIP(hipMemcpyAsync(flow_rgba32f, st, sizeof(float) * 4 * cells, hipMemcpyDeviceToHost,
                               s->stream), "copy flow");
        TWS_HIP(hipStreamSynchronize(s->stream), "tws_download_reference_layout: sync");
    }
    if (flowmap_rg16f) {
        uint16_t* st16 = reinterpret_cast<uint16_t*>(st);
        TWS_HIP(launch_ref_pack(g, s->parity, gb_fill, nullptr, nullptr, st16, s->stream),
                "tws_download_reference_layout: pack flowmap");
        TWS_HIP(hipMemcpyAsync(flowmap_rg16f, st16, sizeof(uint16_t) * 2 * cells,
                               hipMemcpyDeviceToHost, s->stream), "copy flowmap");
    }
    TWS_HIP(hipStreamSynchronize(s->stream), "tws_download_reference_layout: sync");
    return TWS_OK;
}

int tws_apply_brush(tws_sim* s, float cx, float cy, float intensity, float radius_sq) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_apply_brush: NULL handle");
    if (!(radius_sq > 0.0f) || !std::isfinite(cx) || !std::isfinite(cy) ||
        !std::isfinite(intensity))
        return fail(TWS_ERR_INVALID, "tws_apply_brush: invalid brush");
    TWS_HIP(launch_brush(desc_of(s), s->parity, cx, cy, intensity, radius_sq, s->stream),
            "tws_apply_brush");
    return TWS_OK;
}

int tws_set_rain(tws_sim* s, float cx, float cy, float intensity, float radius_sq) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_set_rain: NULL handle");
    if (intensity == 0.0f) {
        s->rain_on = 0;
        return TWS_OK;
    }
    if (!(radius_sq > 0.0f) || !std::isfinite(cx) || !std::isfinite(cy) ||
        !std::isfinite(intensity))
        return fail(TWS_ERR_INVALID, "tws_set_rain: invalid brush");
    s->rain_on = 1;
    s->rain[0] = cx;
    s->rain[1] = cy;
    s->rain[2] = intensity;
    s->rain[3] = radius_sq;
    return TWS_OK;
}

// Runs n steps as a sequence of launches.  With `ev` (2 events per launch),
// each launch is bracketed by an event pair on the handle's stream.
static int step_impl(tws_sim* s, int32_t n, hipEvent_t* ev, int ev_pairs, int* launches) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_step: NULL handle");
    if (n < 0) return fail(TWS_ERR_INVALID, "tws_step: negative step count %d", n);
    int nl = 0;
    auto rec = [&](int which) -> hipError_t {
        if (!ev || nl >= ev_pairs) return hipSuccess;
        return hipEventRecord(ev[2 * nl + which], s->stream);
    };
    if (s->kernel == TWS_KERNEL_TWOPASS) {
        for (int i = 0; i < n; ++i) {
            GridDesc g = desc_of(s);
            TWS_HIP(rec(0), "tws_step: event");
            TWS_HIP(launch_twopass(nullptr, g, s->parity, s->stream), "tws_step: twopass");
            TWS_HIP(rec(1), "tws_step: event");
            s->parity ^= 1;
            ++nl;
        }
        if (launches) *launches = nl;
        return TWS_OK;
    }
    int left = n;
    while (left > 0) {
        // full launches at the configured k; the tail uses the smallest
        // template that holds it, so its halo is no wider than needed
        int kk = s->k;
        if (left < kk) {
            static const int ladder[] = {1, 2, 3, 4, 6, 8};
            for (int c : ladder)
                if (c >= left) { kk = c; break; }
        }
        const int levels = left < kk ? left : kk;
        TileShape sh{};
        if (!tile_shape(s->kernel, kk, &sh))
            return fail(TWS_ERR_INVALID, "tws_step: bad launch config %d/%d", s->kernel, kk);
        GridDesc g = desc_of(s);
        g.tiles_x = (s->W + sh.ox - 1) / sh.ox;
        g.tiles_y = (s->H + sh.oy - 1) / sh.oy;
        g.tile_begin = 0;
        g.steps = levels;
        g.parity0 = s->parity;
        TWS_HIP(rec(0), "tws_step: event");
        TWS_HIP(launch_tile_single(s->kernel, kk, g, 0, s->stream), "tws_step: tile kernel");
        TWS_HIP(rec(1), "tws_step: event");
        s->parity ^= 1;
        left -= levels;
        ++nl;
    }
    if (launches) *launches = nl;
    return TWS_OK;
}

int tws_step(tws_sim* s, int32_t n) { return step_impl(s, n, nullptr, 0, nullptr); }

int tws_step_profiled(tws_sim* s, int32_t n, float* launch_ms, int32_t capacity,
                      int32_t* launches) {
    if (!s || !launch_ms || capacity <= 0 || !launches)
        return fail(TWS_ERR_INVALID, "tws_step_profiled: invalid argument");
    std::vector<hipEvent_t> ev(2 * (size_t)capacity, nullptr);
    for (auto& e : ev) {
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) {
            for (auto& x : ev) if (x) (void)hipEventDestroy(x);
            return hip_fail(r, "tws_step_profiled: hipEventCreate");
        }
    }
    int nl = 0;
    int rc = step_impl(s, n, ev.data(), capacity, &nl);
    hipError_t e = hipStreamSynchronize(s->stream);
    const int m = nl < capacity ? nl : capacity;
    for (int i = 0; rc == TWS_OK && e == hipSuccess && i < m; ++i)
        e = hipEventElapsedTime(&launch_ms[i], ev[2 * i], ev[2 * i + 1]);
    for (auto& x : ev) (void)hipEventDestroy(x);
    if (rc != TWS_OK) return rc;
    if (e != hipSuccess) return hip_fail(e, "tws_step_profiled: timing");
    *launches = nl;
    return TWS_OK;
}

int tws_perform_simulation_step(tws_sim* s, double frame_seconds, double step_seconds,
                                int32_t* steps_done) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_perform_simulation_step: NULL handle");
    if (!(step_seconds > 0.0) || !(frame_seconds >= 0.0))
        return fail(TWS_ERR_INVALID, "tws_perform_simulation_step: invalid times");
    // Terrain.cpp:242-247: accumulate, take whole steps, then clamp to 10
    // (time beyond the clamp is dropped, as in the reference).
    s->time_acc += frame_seconds;
    uint32_t n = (uint32_t)(s->time_acc / step_seconds);
    s->time_acc -= step_seconds * n;
    if (n > 10) n = 10;
    if (steps_done) *steps_done = (int32_t)n;
    return tws_step(s, (int32_t)n);
}

int tws_volume(tws_sim* s, double* out) {
    if (!s || !out) return fail(TWS_ERR_INVALID, "tws_volume: NULL argument");
    TWS_HIP(launch_volume_partials(desc_of(s), s->parity, s->partials, kVolumeParts, s->stream),
            "tws_volume: kernel");
    double host[kVolumeParts];
    TWS_HIP(hipMemcpyAsync(host, s->partials, sizeof(host), hipMemcpyDeviceToHost, s->stream),
            "tws_volume: copy");
    TWS_HIP(hipStreamSynchronize(s->stream), "tws_volume: sync");
    double acc = 0.0;
    for (int i = 0; i < kVolumeParts; ++i) acc += host[i];
    *out = acc;
    return TWS_OK;
}

int tws_check_finite(tws_sim* s) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_check_finite: NULL handle");
    unsigned flag = 0;
    TWS_HIP(hipMemcpyAsync(&flag, s->status, sizeof(flag), hipMemcpyDeviceToHost, s->stream),
            "tws_check_finite: copy");
    TWS_HIP(hipMemsetAsync(s->status, 0, sizeof(unsigned), s->stream), "tws_check_finite: clear");
    TWS_HIP(hipStreamSynchronize(s->stream), "tws_check_finite: sync");
    if (flag & 1u) return fail(TWS_ERR_NONFINITE, "tws_check_finite: non-finite state seen");
    return TWS_OK;
}

int tws_sync(tws_sim* s) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_sync: NULL handle");
    TWS_HIP(hipStreamSynchronize(s->stream), "tws_sync");
    return TWS_OK;
}

int tws_get_device_planes(const tws_sim* s, tws_device_planes* out) {
    if (!s || !out) return fail(TWS_ERR_INVALID, "tws_get_device_planes: NULL argument");
    out->terrain = s->b;
    out->depth = s->d[s->parity];
    for (int c = 0; c < 4; ++c) out->flux[c] = s->f[s->parity][c];
    out->pitch = s->pitch;
    out->width = s->W;
    out->height = s->H;
    return TWS_OK;
}

int tws_set_launch(tws_sim* s, int32_t kernel, int32_t k) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_set_launch: NULL handle");
    if (kernel == TWS_KERNEL_AUTO) {
        s->auto_launch = true;
        choose_launch(s->W, s->H, &s->kernel, &s->k);
        return TWS_OK;
    }
    if (kernel == TWS_KERNEL_TWOPASS) {
        if (k != 1) return fail(TWS_ERR_INVALID, "tws_set_launch: TWOPASS runs 1 step per launch");
    } else {
        TileShape sh{};
        if (!tile_shape(kernel, k, &sh))
            return fail(TWS_ERR_INVALID, "tws_set_launch: unknown kernel %d or k %d", kernel, k);
    }
    s->auto_launch = false;
    s->kernel = kernel;
    s->k = k;
    return TWS_OK;
}

int tws_get_launch(const tws_sim* s, int32_t* kernel, int32_t* k) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_get_launch: NULL handle");
    if (kernel) *kernel = s->kernel;
    if (k) *k = s->k;
    return TWS_OK;
}

int tws_float_to_half_rtz(const float* in, uint16_t* out, int64_t n) {
    if ((!in || !out) && n > 0) return fail(TWS_ERR_INVALID, "tws_float_to_half_rtz: NULL");
    for (int64_t i = 0; i < n; ++i) out[i] = half_rtz(in[i]);
    return TWS_OK;
}

int tws_stream_copy_bench(int64_t bytes, int32_t iterations, double* gbps) {
    if (!gbps || bytes < (1 << 20) || iterations <= 0)
        return fail(TWS_ERR_INVALID, "tws_stream_copy_bench: invalid argument");
    bytes &= ~(int64_t)255;
    void *src = nullptr, *dst = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&src, (size_t)bytes);
    if (e == hipSuccess) e = hipMalloc(&dst, (size_t)bytes);
    if (e == hipSuccess) e = hipMemset(src, 0, (size_t)bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = launch_stream_copy(src, dst, bytes, nullptr);  // warm
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    for (int i = 0; e == hipSuccess && i < iterations; ++i)
        e = launch_stream_copy(src, dst, bytes, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (e != hipSuccess) return hip_fail(e, "tws_stream_copy_bench");
    *gbps = 2.0 * (double)bytes * iterations / (ms * 1e-3) / 1e9;
    return TWS_OK;
}

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
