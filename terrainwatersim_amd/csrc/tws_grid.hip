// tws_grid.hip -- operations on the device state of one grid (tws_grid.hpp).

#include <hip/hip_runtime.h>

#include <cmath>

#include "tws_fbm.hpp"
#include "tws_grid.hpp"
#include "tws_runtime.hpp"

namespace tws {

#define TWS_HIP(call, what)                              \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

void GridState::bind(float* base, int w, int h) {
    W = w;
    H = h;
    pitch = pitch_of(w);
    const int64_t plane = pitch * (int64_t)h;
    b = base;
    d[0] = base + plane;
    d[1] = base + 2 * plane;
    for (int p = 0; p < 2; ++p)
        for (int c = 0; c < 4; ++c) f[p][c] = base + (3 + 4 * p + c) * plane;
    parity = 0;
}

GridDesc GridState::desc() const {
    GridDesc g{};
    g.b = b;
    for (int p = 0; p < 2; ++p) {
        g.d[p] = d[p];
        for (int c = 0; c < 4; ++c) g.f[p][c] = f[p][c];
    }
    g.pitch = pitch;
    g.W = W;
    g.H = H;
    fill_ghosts(g, params);
    g.rain_on = rain_on;
    g.rain_cx = rain[0];
    g.rain_cy = rain[1];
    g.rain_I = rain[2];
    g.rain_R2 = rain[3];
    g.status = status;
    g.parity0 = parity;
    g.tile_group = tile_group;
    const bool dispatch = params.boundary == TWS_BOUNDARY_REFERENCE_DISPATCH;
    g.upd_w = dispatch ? W / 16 * 16 : W;
    g.upd_h = dispatch ? H / 16 * 16 : H;
    g.brush_w = dispatch ? W / 32 * 32 : W;
    g.brush_h = dispatch ? H / 32 * 32 : H;
    return g;
}

bool grid_tiles(int W, int H, int kernel, int k, int* tiles_x, int* tiles_y) {
    TileShape sh{};
    if (!tile_shape(kernel, k, &sh)) return false;
    *tiles_x = (W + sh.ox - 1) / sh.ox;
    *tiles_y = (H + sh.oy - 1) / sh.oy;
    return true;
}

// One plane between a packed host array and its pitched device rows.  The
// copy across PCIe is always one contiguous transfer: pageable 2D copies run
// at ~2 GB/s against ~17 GB/s for 1D ones.  A plane whose rows are padded
// goes through `st` (>= W*H floats of device scratch) and a 2D copy on the
// device.
static hipError_t plane_h2d(float* plane, int64_t pitch, const float* host, int W, int H,
                            float* st, hipStream_t s) {
    const size_t row = sizeof(float) * (size_t)W;
    if (pitch == W) return hipMemcpyAsync(plane, host, row * H, hipMemcpyHostToDevice, s);
    hipError_t e = hipMemcpyAsync(st, host, row * H, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(plane, sizeof(float) * (size_t)pitch, st, row, row, H,
                             hipMemcpyDeviceToDevice, s);
    return e;
}

static hipError_t plane_d2h(float* host, const float* plane, int64_t pitch, int W, int H,
                            float* st, hipStream_t s) {
    const size_t row = sizeof(float) * (size_t)W;
    if (pitch == W) return hipMemcpyAsync(host, plane, row * H, hipMemcpyDeviceToHost, s);
    hipError_t e = hipMemcpy2DAsync(st, row, plane, sizeof(float) * (size_t)pitch, row, H,
                                    hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(host, st, row * H, hipMemcpyDeviceToHost, s);
    return e;
}

int grid_clear(GridState& g, hipStream_t s) {
    TWS_HIP(launch_zero_fill(g.b, (int64_t)kPlanes * g.pitch * g.H, s), "clear state");
    g.parity = 0;
    return TWS_OK;
}

int grid_upload(GridState& g, const float* terrain, const float* depth, const float* flux4,
                hipStream_t s) {
    float* st = g.staging();
    if (terrain)
        TWS_HIP(plane_h2d(g.b, g.pitch, terrain, g.W, g.H, st, s), "upload: terrain");
    if (depth)
        TWS_HIP(plane_h2d(g.d[g.parity], g.pitch, depth, g.W, g.H, st, s), "upload: depth");
    const GridDesc gd = g.desc();
    if (flux4) {
        TWS_HIP(hipMemcpyAsync(st, flux4, sizeof(float) * 4 * (size_t)g.W * g.H,
                               hipMemcpyHostToDevice, s), "upload: flux");
        TWS_HIP(launch_deinterleave(gd, g.parity, st, s), "upload: deinterleave");
    } else {
        TWS_HIP(launch_zero_fill(g.f[g.parity][0], 4 * g.pitch * (int64_t)g.H, s),
                "upload: clear flux");
    }
    TWS_HIP(launch_check_finite(gd, g.parity, s), "upload: finite check");
    TWS_HIP(hipStreamSynchronize(s), "upload: sync");
    return TWS_OK;
}

int grid_upload_async(GridState& g, const float* terrain, const float* depth, hipStream_t s) {
    // a grid being admitted is not stepping: its idle flux set is free scratch
    float* st = g.staging();
    if (terrain)
        TWS_HIP(plane_h2d(g.b, g.pitch, terrain, g.W, g.H, st, s), "upload: terrain");
    if (depth)
        TWS_HIP(plane_h2d(g.d[g.parity], g.pitch, depth, g.W, g.H, st, s), "upload: depth");
    TWS_HIP(launch_zero_fill(g.f[g.parity][0], 4 * g.pitch * (int64_t)g.H, s), "upload: flux");
    TWS_HIP(launch_check_finite(g.desc(), g.parity, s), "upload: finite check");
    return TWS_OK;
}

int grid_download_depth_async(GridState& g, float* depth, hipStream_t s) {
    TWS_HIP(plane_d2h(depth, g.d[g.parity], g.pitch, g.W, g.H, g.staging(), s), "download: depth");
    return TWS_OK;
}

int grid_download(GridState& g, float* depth, float* flux4, float* velocity2, hipStream_t s) {
    const GridDesc gd = g.desc();
    float* st = g.staging();
    if (depth) {
        TWS_HIP(plane_d2h(depth, g.d[g.parity], g.pitch, g.W, g.H, st, s), "download: depth");
        TWS_HIP(hipStreamSynchronize(s), "download: sync");  // st is reused below
    }
    if (flux4) {
        TWS_HIP(launch_interleave(gd, g.parity, st, s), "download: interleave");
        TWS_HIP(hipMemcpyAsync(flux4, st, sizeof(float) * 4 * (size_t)g.W * g.H,
                               hipMemcpyDeviceToHost, s), "download: flux");
        TWS_HIP(hipStreamSynchronize(s), "download: sync");
    }
    if (velocity2) {
        TWS_HIP(launch_velocity(gd, g.parity, st, s), "download: velocity");
        TWS_HIP(hipMemcpyAsync(velocity2, st, sizeof(float) * 2 * (size_t)g.W * g.H,
                               hipMemcpyDeviceToHost, s), "download: velocity copy");
    }
    TWS_HIP(hipStreamSynchronize(s), "download: sync");
    return TWS_OK;
}

int grid_upload_reference(GridState& g, const float* terrain_rgba, const float* flow_rgba,
                          hipStream_t s) {
    const GridDesc gd = g.desc();
    float* st = g.staging();
    const size_t bytes = sizeof(float) * 4 * (size_t)g.W * g.H;
    if (terrain_rgba) {
        TWS_HIP(hipMemcpyAsync(st, terrain_rgba, bytes, hipMemcpyHostToDevice, s),
                "upload_reference: terrain");
        TWS_HIP(launch_ref_unpack(gd, g.parity, st, nullptr, g.b, s),
                "upload_reference: unpack terrain");
    }
    if (flow_rgba) {
        TWS_HIP(hipMemcpyAsync(st, flow_rgba, bytes, hipMemcpyHostToDevice, s),
                "upload_reference: flow");
        TWS_HIP(launch_ref_unpack(gd, g.parity, nullptr, st, g.b, s),
                "upload_reference: unpack flow");
    }
    TWS_HIP(launch_check_finite(gd, g.parity, s), "upload_reference: finite check");
    TWS_HIP(hipStreamSynchronize(s), "upload_reference: sync");
    return TWS_OK;
}

int grid_download_reference(GridState& g, float* terrain_rgba, float gb_fill, float* flow_rgba,
                            uint16_t* flowmap, hipStream_t s) {
    const GridDesc gd = g.desc();
    float* st = g.staging();
    const size_t cells = (size_t)g.W * g.H;
    if (terrain_rgba) {
        TWS_HIP(launch_ref_pack(gd, g.parity, gb_fill, st, nullptr, nullptr, s),
                "download_reference: pack terrain");
        TWS_HIP(hipMemcpyAsync(terrain_rgba, st, sizeof(float) * 4 * cells,
                               hipMemcpyDeviceToHost, s), "download_reference: terrain");
        TWS_HIP(hipStreamSynchronize(s), "download_reference: sync");
    }
    if (flow_rgba) {
        TWS_HIP(launch_ref_pack(gd, g.parity, gb_fill, nullptr, st, nullptr, s),
                "download_reference: pack flow");
        TWS_HIP(hipMemcpyAsync(flow_rgba, st, sizeof(float) * 4 * cells, hipMemcpyDeviceToHost,
                               s), "download_reference: flow");
        TWS_HIP(hipStreamSynchronize(s), "download_reference: sync");
    }
    if (flowmap) {
        uint16_t* st16 = reinterpret_cast<uint16_t*>(st);
        TWS_HIP(launch_ref_pack(gd, g.parity, gb_fill, nullptr, nullptr, st16, s),
                "download_reference: pack flowmap");
        TWS_HIP(hipMemcpyAsync(flowmap, st16, sizeof(uint16_t) * 2 * cells,
                               hipMemcpyDeviceToHost, s), "download_reference: flowmap");
    }
    TWS_HIP(hipStreamSynchronize(s), "download_reference: sync");
    return TWS_OK;
}

int grid_brush(GridState& g, float cx, float cy, float I, float R2, hipStream_t s) {
    if (!(R2 > 0.0f) || !std::isfinite(cx) || !std::isfinite(cy) || !std::isfinite(I))
        return fail(TWS_ERR_INVALID, "brush: invalid centre, intensity or radius");
    TWS_HIP(launch_brush(g.desc(), g.parity, cx, cy, I, R2, s), "brush");
    return TWS_OK;
}

int grid_set_rain(GridState& g, float cx, float cy, float I, float R2) {
    if (I == 0.0f) {
        g.rain_on = 0;
        return TWS_OK;
    }
    if (!(R2 > 0.0f) || !std::isfinite(cx) || !std::isfinite(cy) || !std::isfinite(I))
        return fail(TWS_ERR_INVALID, "rain: invalid centre, intensity or radius");
    g.rain_on = 1;
    g.rain[0] = cx;
    g.rain[1] = cy;
    g.rain[2] = I;
    g.rain[3] = R2;
    return TWS_OK;
}

int grid_volume(GridState& g, double* partials_dev, double* out, hipStream_t s) {
    TWS_HIP(launch_volume_partials(g.desc(), g.parity, partials_dev, kVolumeParts, s),
            "volume: kernel");
    double host[kVolumeParts];
    TWS_HIP(hipMemcpyAsync(host, partials_dev, sizeof(host), hipMemcpyDeviceToHost, s),
            "volume: copy");
    TWS_HIP(hipStreamSynchronize(s), "volume: sync");
    double acc = 0.0;
    for (int i = 0; i < kVolumeParts; ++i) acc += host[i];
    *out = acc;
    return TWS_OK;
}

int grid_init_terrain_fbm(GridState& g, uint64_t seed, int octaves, int base, float persistence,
                          float scale, hipStream_t s) {
    if (!fbm_args_ok(g.W, g.H, octaves, base))
        return fail(TWS_ERR_INVALID, "init_terrain_fbm: invalid octaves/base frequency");
    TWS_HIP(launch_fbm(g.b, g.desc(), seed, octaves, base, (double)persistence, (double)scale, s),
            "init_terrain_fbm");
    return TWS_OK;
}

int grid_init_water(GridState& g, float depth, int x0, int y0, int x1, int y1, float rect_depth,
                    hipStream_t s) {
    if (!std::isfinite(depth) || !std::isfinite(rect_depth))
        return fail(TWS_ERR_INVALID, "init_water: non-finite depth");
    TWS_HIP(launch_init_water(g.desc(), g.parity, depth, x0, y0, x1, y1, rect_depth, s),
            "init_water");
    return TWS_OK;
}

int grid_check_finite(GridState& g, hipStream_t s) {
    unsigned flag = 0;
    TWS_HIP(hipMemcpyAsync(&flag, g.status, sizeof(flag), hipMemcpyDeviceToHost, s),
            "check_finite: copy");
    TWS_HIP(hipMemsetAsync(g.status, 0, sizeof(unsigned), s), "check_finite: clear");
    TWS_HIP(hipStreamSynchronize(s), "check_finite: sync");
    if (flag & 1u) return fail(TWS_ERR_NONFINITE, "non-finite depth or flux seen");
    return TWS_OK;
}

}  // namespace tws
