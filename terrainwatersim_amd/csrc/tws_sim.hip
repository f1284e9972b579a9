// tws_sim.hip -- the single-grid handle of the C ABI (tws.h): one grid's
// device state (tws_grid.hpp), its stream and the step planner.

#include <hip/hip_runtime.h>

#include <cmath>
#include <new>
#include <vector>

#include "tws.h"
#include "tws_grid.hpp"
#include "tws_runtime.hpp"

using namespace tws;

struct tws_sim {
    GridState g;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    float* base = nullptr;       // the grid's planes
    double* aux = nullptr;       // volume partials + status word
    int kernel = TWS_KERNEL_TILE_128x64;
    int k = 4;
    double time_acc = 0.0;       // PerformSimulationStep accumulator
    float* frame_mips = nullptr; // mip chain kept current per frame (caller's buffer)
    size_t frame_mips_cap = 0;
    float frame_mips_gb = 0.3f;
};

#define TWS_HIP(call, what)                              \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

namespace tws {

bool stream_usable(const GridState& gs) {
    return gs.params.boundary == TWS_BOUNDARY_CLOSED && gs.pitch * gs.H < ((int64_t)1 << 29);
}

// Streaming launches (tws_stream.hip): full launches of k steps, then the
// remainder as launches of the largest built step counts that fit it (each
// launch runs exactly its template's number of levels).
static int step_stream(GridState& gs, int k, int32_t n, hipStream_t stream, hipEvent_t* ev,
                       int ev_pairs, int* launches) {
    int nl = 0;
    const int cus = device_cus();
    if (cus <= 0) return fail(TWS_ERR_DEVICE, "step: cannot query the device");
    int left = n;
    while (left > 0) {
        int kk = k;
        if (left < kk) {
            static const int ladder[] = {8, 6, 5, 4, 3, 2, 1};
            for (int c : ladder)
                if (c <= left) { kk = c; break; }
        }
        const StreamPlan plan = stream_plan(gs.W, gs.H, kk, cus);
        if (plan.blocks <= 0) return fail(TWS_ERR_INVALID, "step: bad stream launch (k %d)", kk);
        GridDesc g = gs.desc();
        g.steps = kk;
        if (ev && nl < ev_pairs) TWS_HIP(hipEventRecord(ev[2 * nl], stream), "step: event");
        TWS_HIP(launch_stream_single(kk, g, plan, gs.parity, stream), "step: stream kernel");
        if (ev && nl < ev_pairs) TWS_HIP(hipEventRecord(ev[2 * nl + 1], stream), "step: event");
        gs.parity ^= 1;
        left -= kk;
        ++nl;
    }
    if (launches) *launches = nl;
    return TWS_OK;
}

// Runs n steps of one grid as a sequence of launches: full launches of k
// steps, then the tail on the smallest template that holds it.  With `ev`
// (2 events per launch) each launch is bracketed by an event pair.
int step_grid(GridState& gs, int kernel, int k, int32_t n, hipStream_t stream, hipEvent_t* ev,
              int ev_pairs, int* launches) {
    if (n < 0) return fail(TWS_ERR_INVALID, "step: negative step count %d", n);
    int nl = 0;
    auto rec = [&](int which) -> hipError_t {
        if (!ev || nl >= ev_pairs) return hipSuccess;
        return hipEventRecord(ev[2 * nl + which], stream);
    };
    if (kernel == TWS_KERNEL_STREAM_128) {
        // closed walls and planes under 2 GiB (its stores address a plane
        // with 31-bit offsets); other grids take the tile kernel's launch
        if (stream_usable(gs))
            return step_stream(gs, k, n, stream, ev, ev_pairs, launches);
        tile_launch(gs.W, gs.H, &kernel, &k);
    }
    if (kernel == TWS_KERNEL_TWOPASS || needs_twopass(gs.desc())) {
        for (int i = 0; i < n; ++i) {
            const GridDesc g = gs.desc();
            TWS_HIP(rec(0), "step: event");
            TWS_HIP(launch_twopass(nullptr, g, gs.parity, stream), "step: two-pass kernels");
            TWS_HIP(rec(1), "step: event");
            gs.parity ^= 1;
            ++nl;
        }
        if (launches) *launches = nl;
        return TWS_OK;
    }
    int left = n;
    while (left > 0) {
        int kk = k;
        if (left < kk) {
            static const int ladder[] = {1, 2, 3, 4, 5, 6, 8};
            for (int c : ladder)
                if (c >= left) { kk = c; break; }
        }
        const int levels = left < kk ? left : kk;
        GridDesc g = gs.desc();
        if (!grid_tiles(gs.W, gs.H, kernel, kk, &g.tiles_x, &g.tiles_y))
            return fail(TWS_ERR_INVALID, "step: bad launch config %d/%d", kernel, kk);
        g.tile_begin = 0;
        g.steps = levels;
        TWS_HIP(rec(0), "step: event");
        TWS_HIP(launch_tile_single(kernel, kk, g, 0, stream), "step: tile kernel");
        TWS_HIP(rec(1), "step: event");
        gs.parity ^= 1;
        left -= levels;
        ++nl;
    }
    if (launches) *launches = nl;
    return TWS_OK;
}

}  // namespace tws

extern "C" {

int tws_create(int32_t width, int32_t height, const tws_params* params, void* stream,
               tws_sim** out) {
    if (!out) return fail(TWS_ERR_INVALID, "tws_create: out is NULL");
    *out = nullptr;
    if (width <= 0 || height <= 0 || width > (1 << 20) || height > (1 << 20))
        return fail(TWS_ERR_INVALID, "tws_create: invalid size %d x %d", width, height);
    if (!valid_params(params)) return fail(TWS_ERR_INVALID, "tws_create: invalid params");
    tws_sim* s = new (std::nothrow) tws_sim();
    if (!s) return fail(TWS_ERR_NOMEM, "tws_create: host allocation failed");
    hipError_t e = hipSuccess;
    if (stream) {
        s->stream = (hipStream_t)stream;
    } else {
        e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
        if (e != hipSuccess) { delete s; return hip_fail(e, "tws_create: hipStreamCreate"); }
        s->own_stream = true;
    }
    e = hipMalloc(&s->base, GridState::bytes_for(width, height));
    if (e != hipSuccess) { tws_destroy(s); return hip_fail(e, "tws_create: hipMalloc state"); }
    e = hipMalloc(&s->aux, sizeof(double) * kVolumeParts + 256);
    if (e != hipSuccess) { tws_destroy(s); return hip_fail(e, "tws_create: hipMalloc aux"); }
    s->g.bind(s->base, width, height);
    s->g.params = *params;
    s->g.status = reinterpret_cast<unsigned*>(s->aux + kVolumeParts);
    int rc = grid_clear(s->g, s->stream);
    if (rc == TWS_OK) {
        e = hipMemsetAsync(s->g.status, 0, 256, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) rc = hip_fail(e, "tws_create: clear state");
    }
    if (rc != TWS_OK) { tws_destroy(s); return rc; }
    choose_launch(width, height, &s->kernel, &s->k);
    s->g.tile_group = default_tile_group(width, height, s->kernel, s->k);
    *out = s;
    return TWS_OK;
}

int tws_destroy(tws_sim* s) {
    if (!s) return TWS_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->base) (void)hipFree(s->base);
    if (s->aux) (void)hipFree(s->aux);
    if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return TWS_OK;
}

int tws_set_params(tws_sim* s, const tws_params* p) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_set_params: NULL handle");
    if (!valid_params(p)) return fail(TWS_ERR_INVALID, "tws_set_params: invalid params");
    s->g.params = *p;
    return TWS_OK;
}

int tws_get_params(const tws_sim* s, tws_params* out) {
    if (!s || !out) return fail(TWS_ERR_INVALID, "tws_get_params: NULL argument");
    *out = s->g.params;
    return TWS_OK;
}

int tws_upload(tws_sim* s, const float* terrain, const float* depth, const float* flux4) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_upload: NULL handle");
    return grid_upload(s->g, terrain, depth, flux4, s->stream);
}

int tws_download(tws_sim* s, float* depth, float* flux4, float* velocity2) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_download: NULL handle");
    return grid_download(s->g, depth, flux4, velocity2, s->stream);
}

int tws_upload_reference_layout(tws_sim* s, const float* terrain_rgba32f,
                                const float* flow_rgba32f) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_upload_reference_layout: NULL handle");
    return grid_upload_reference(s->g, terrain_rgba32f, flow_rgba32f, s->stream);
}

int tws_download_reference_layout(tws_sim* s, float* terrain_rgba32f, float gb_fill,
                                  float* flow_rgba32f, uint16_t* flowmap_rg16f) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_download_reference_layout: NULL handle");
    return grid_download_reference(s->g, terrain_rgba32f, gb_fill, flow_rgba32f, flowmap_rg16f,
                                   s->stream);
}

int tws_apply_brush(tws_sim* s, float cx, float cy, float intensity, float radius_sq) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_apply_brush: NULL handle");
    return grid_brush(s->g, cx, cy, intensity, radius_sq, s->stream);
}

int tws_set_rain(tws_sim* s, float cx, float cy, float intensity, float radius_sq) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_set_rain: NULL handle");
    return grid_set_rain(s->g, cx, cy, intensity, radius_sq);
}

int tws_step(tws_sim* s, int32_t n) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_step: NULL handle");
    return step_grid(s->g, s->kernel, s->k, n, s->stream, nullptr, 0, nullptr);
}

int tws_step_profiled(tws_sim* s, int32_t n, float* launch_ms, int32_t capacity,
                      int32_t* launches) {
    if (!s || !launch_ms || capacity <= 0 || !launches)
        return fail(TWS_ERR_INVALID, "tws_step_profiled: invalid argument");
    std::vector<hipEvent_t> ev;
    try {
        ev.assign(2 * (size_t)capacity, nullptr);
    } catch (...) {
        return fail(TWS_ERR_NOMEM, "tws_step_profiled: host allocation failed");
    }
    for (auto& e : ev) {
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) {
            for (auto& x : ev) if (x) (void)hipEventDestroy(x);
            return hip_fail(r, "tws_step_profiled: hipEventCreate");
        }
    }
    int nl = 0;
    int rc = step_grid(s->g, s->kernel, s->k, n, s->stream, ev.data(), capacity, &nl);
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
    if (!(step_seconds > 0.0) || !(frame_seconds >= 0.0) || !std::isfinite(step_seconds) ||
        !std::isfinite(frame_seconds))
        return fail(TWS_ERR_INVALID, "tws_perform_simulation_step: invalid times");
    // Terrain.cpp:242-247: accumulate, take whole steps, then clamp to 10
    // (time beyond the clamp is dropped, as in the reference).  The whole
    // step count is taken in double, so a long frame cannot overflow the
    // integer conversion; rounding can leave the accumulator an ulp below
    // zero, which counts as no step.
    s->time_acc += frame_seconds;
    const double whole = std::fmax(0.0, std::floor(s->time_acc / step_seconds));
    s->time_acc -= step_seconds * whole;
    const uint32_t n = whole > 10.0 ? 10u : (uint32_t)whole;
    if (steps_done) *steps_done = (int32_t)n;
    const int rc = tws_step(s, (int32_t)n);
    if (rc != TWS_OK || n == 0 || !s->frame_mips) return rc;
    // Terrain.cpp:270-276: the mip chain follows every frame with steps
    return grid_mips(s->g, s->frame_mips_gb, s->frame_mips, s->frame_mips_cap, s->stream);
}

int tws_mip_chain_size(int32_t width, int32_t height, int32_t* levels, size_t* floats) {
    if (width <= 0 || height <= 0)
        return fail(TWS_ERR_INVALID, "tws_mip_chain_size: invalid size %dx%d", width, height);
    int n = 0;
    mip_chain(width, height, &n, floats);
    if (levels) *levels = n;
    return TWS_OK;
}

int tws_generate_mips(tws_sim* s, float gb_fill, float* device_rgba, size_t capacity_floats) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_generate_mips: NULL handle");
    return grid_mips(s->g, gb_fill, device_rgba, capacity_floats, s->stream);
}

int tws_download_mips(tws_sim* s, float gb_fill, float* host_rgba, size_t capacity_floats) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_download_mips: NULL handle");
    size_t need = 0;
    mip_chain(s->g.W, s->g.H, nullptr, &need);
    if (need == 0) return TWS_OK;
    if (!host_rgba || capacity_floats < need)
        return fail(TWS_ERR_INVALID, "tws_download_mips: buffer holds %zu floats, the chain needs %zu",
                    host_rgba ? capacity_floats : (size_t)0, need);
    // the idle flux set: 4 * pitch * H floats, more than any chain of the grid
    float* st = s->g.staging();
    const int rc = grid_mips(s->g, gb_fill, st, 4 * (size_t)s->g.pitch * s->g.H, s->stream);
    if (rc != TWS_OK) return rc;
    TWS_HIP(hipMemcpyAsync(host_rgba, st, sizeof(float) * need, hipMemcpyDeviceToHost, s->stream),
            "tws_download_mips: copy");
    TWS_HIP(hipStreamSynchronize(s->stream), "tws_download_mips: sync");
    return TWS_OK;
}

int tws_set_frame_mips(tws_sim* s, float* device_rgba, size_t capacity_floats, float gb_fill) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_set_frame_mips: NULL handle");
    if (device_rgba) {
        size_t need = 0;
        mip_chain(s->g.W, s->g.H, nullptr, &need);
        if (capacity_floats < need)
            return fail(TWS_ERR_INVALID, "tws_set_frame_mips: buffer holds %zu floats, the chain needs %zu",
                        capacity_floats, need);
        if (reinterpret_cast<uintptr_t>(device_rgba) % 16 != 0 || !std::isfinite(gb_fill))
            return fail(TWS_ERR_INVALID, "tws_set_frame_mips: unaligned buffer or non-finite fill");
    }
    s->frame_mips = device_rgba;
    s->frame_mips_cap = device_rgba ? capacity_floats : 0;
    s->frame_mips_gb = gb_fill;
    return TWS_OK;
}

int tws_volume(tws_sim* s, double* out) {
    if (!s || !out) return fail(TWS_ERR_INVALID, "tws_volume: NULL argument");
    return grid_volume(s->g, s->aux, out, s->stream);
}

int tws_check_finite(tws_sim* s) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_check_finite: NULL handle");
    return grid_check_finite(s->g, s->stream);
}

int tws_sync(tws_sim* s) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_sync: NULL handle");
    TWS_HIP(hipStreamSynchronize(s->stream), "tws_sync");
    return TWS_OK;
}

int tws_get_device_planes(const tws_sim* s, tws_device_planes* out) {
    if (!s || !out) return fail(TWS_ERR_INVALID, "tws_get_device_planes: NULL argument");
    out->terrain = s->g.b;
    out->depth = s->g.d[s->g.parity];
    for (int c = 0; c < 4; ++c) out->flux[c] = s->g.f[s->g.parity][c];
    out->pitch = s->g.pitch;
    out->width = s->g.W;
    out->height = s->g.H;
    return TWS_OK;
}

int tws_set_launch(tws_sim* s, int32_t kernel, int32_t k) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_set_launch: NULL handle");
    if (kernel == TWS_KERNEL_AUTO) {
        choose_launch(s->g.W, s->g.H, &s->kernel, &s->k);
        return TWS_OK;
    }
    if (kernel == TWS_KERNEL_TWOPASS) {
        if (k != 1) return fail(TWS_ERR_INVALID, "tws_set_launch: TWOPASS runs 1 step per launch");
    } else if (kernel == TWS_KERNEL_STREAM_128) {
        int ox = 0;
        if (k < 2 || !stream_shape(k, &ox))
            return fail(TWS_ERR_INVALID, "tws_set_launch: STREAM_128 runs k = 2..6 or 8, not %d", k);
    } else {
        TileShape sh{};
        if (!tile_shape(kernel, k, &sh))
            return fail(TWS_ERR_INVALID, "tws_set_launch: unknown kernel %d or k %d", kernel, k);
    }
    s->kernel = kernel;
    s->k = k;
    return TWS_OK;
}

int tws_init_terrain_fbm(tws_sim* s, uint64_t seed, int32_t octaves, int32_t base_frequency,
                         float persistence, float height_scale) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_init_terrain_fbm: NULL handle");
    return grid_init_terrain_fbm(s->g, seed, octaves, base_frequency, persistence, height_scale,
                                 s->stream);
}

int tws_init_water(tws_sim* s, float depth, int32_t x0, int32_t y0, int32_t x1, int32_t y1,
                   float rect_depth) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_init_water: NULL handle");
    return grid_init_water(s->g, depth, x0, y0, x1, y1, rect_depth, s->stream);
}

int tws_init_reference_scene(tws_sim* s, float height_scale, uint32_t seed, int32_t scene_index) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_init_reference_scene: NULL handle");
    return grid_init_reference_scene(s->g, height_scale, seed, scene_index, s->stream);
}

int tws_set_tile_order(tws_sim* s, int32_t group_columns) {
    if (!s || group_columns < 0) return fail(TWS_ERR_INVALID, "tws_set_tile_order: bad argument");
    s->g.tile_group = group_columns;
    return TWS_OK;
}

int tws_get_launch(const tws_sim* s, int32_t* kernel, int32_t* k) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_get_launch: NULL handle");
    if (kernel) *kernel = s->kernel;
    if (k) *k = s->k;
    return TWS_OK;
}

}  // extern "C"
