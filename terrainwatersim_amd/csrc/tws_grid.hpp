// tws_grid.hpp -- device state of one grid and the operations on it, shared
// by the single-grid handle (tws_sim) and the slots of a batch (tws_batch).
//
// Layout: 11 planes of height x pitch floats (pitch = width rounded up to 64,
// i.e. 256-byte rows) in one block of device memory:
//   b | d0 | d1 | f0[+x,-x,+y,-y] | f1[+x,-x,+y,-y]
// d and f are ping-pong sets; `parity` names the set holding the current
// state.  The idle set doubles as staging for the interleaved host layouts,
// so no operation here allocates device memory.
#pragma once

#include <cstddef>
#include <cstdint>

#include "tws.h"
#include "tws_internal.hpp"

namespace tws {

constexpr int kPlanes = 11;

struct GridState {
    int W = 0, H = 0;
    int64_t pitch = 0;
    tws_params params{};
    float* b = nullptr;
    float* d[2] = {nullptr, nullptr};
    float* f[2][4] = {};
    unsigned* status = nullptr;  // device flag word (bit 0: non-finite seen)
    int parity = 0;
    int rain_on = 0;
    float rain[4] = {0, 0, 0, 0};
    // tile order of the step kernel: groups of 16 tile columns (0 = row-major);
    // 16 measured best at 4096^2 (+1.6% at k = 4, +2.8% at k = 6)
    int tile_group = 16;

    static size_t bytes_for(int W, int H) {
        return sizeof(float) * (size_t)kPlanes * (size_t)pitch_of(W) * (size_t)H;
    }
    static int64_t pitch_of(int W) { return ((int64_t)W + 63) / 64 * 64; }
    void bind(float* base, int w, int h);   // plane pointers inside `base`
    GridDesc desc() const;                  // kernel view, parity0 = parity
    float* staging() const { return f[parity ^ 1][0]; }  // >= 4*W*H floats
};

int grid_clear(GridState& g, hipStream_t s);
// n steps of one grid on a kernel variant (tws_sim.hip): the streaming kernel
// where it applies (else the tile table's launch), tile launches of k steps,
// or the two-pass kernels; optional events around each launch.
int step_grid(GridState& gs, int kernel, int k, int32_t n, hipStream_t stream, hipEvent_t* ev,
              int ev_pairs, int* launches);
// Whether the streaming kernel can step this grid: closed walls, and each
// plane smaller than 2 GiB (its stores address a plane with 31-bit offsets
// and use offset 2^31 for lanes that must not write).
bool stream_usable(const GridState& gs);
int grid_upload(GridState& g, const float* terrain, const float* depth, const float* flux4,
                hipStream_t s);
int grid_download(GridState& g, float* depth, float* flux4, float* velocity2, hipStream_t s);
int grid_upload_reference(GridState& g, const float* terrain_rgba, const float* flow_rgba,
                          hipStream_t s);
int grid_download_reference(GridState& g, float* terrain_rgba, float gb_fill, float* flow_rgba,
                            uint16_t* flowmap, hipStream_t s);
int grid_brush(GridState& g, float cx, float cy, float I, float R2, hipStream_t s);
int grid_set_rain(GridState& g, float cx, float cy, float I, float R2);
int grid_volume(GridState& g, double* partials_dev, double* out, hipStream_t s);
int grid_check_finite(GridState& g, hipStream_t s);  // reads and clears the flag
int grid_init_terrain_fbm(GridState& g, uint64_t seed, int octaves, int base, float persistence,
                          float scale, hipStream_t s);
int grid_init_water(GridState& g, float depth, int x0, int y0, int x1, int y1, float rect_depth,
                    hipStream_t s);
// The reference's default scene (tws_refscene.cpp): b, the lake depth and
// zero flux of a square grid; synchronises the stream.
int grid_init_reference_scene(GridState& g, float height_scale, uint32_t seed,
                              int32_t scene_index, hipStream_t s);

// Terrain mip chain (tws_mips.hip): level count and floats of levels 1..,
// and the chain written into device memory `dst` on stream s.
int mip_chain(int W, int H, int* levels, size_t* floats);
int grid_mips(GridState& g, float gb_fill, float* dst, size_t capacity_floats, hipStream_t s);

// Tiles of a grid for a (kernel variant, steps per launch) pair.
bool grid_tiles(int W, int H, int kernel, int k, int* tiles_x, int* tiles_y);

// Asynchronous transfers for callers that synchronise themselves: the host
// buffers must be pinned (hipHostMalloc) for the copies to overlap anything.
// Upload: terrain and depth (either may be NULL), flux reset to zero.
int grid_upload_async(GridState& g, const float* terrain, const float* depth, hipStream_t s);
// Download of the depth plane into a dense host buffer.
int grid_download_depth_async(GridState& g, float* depth, hipStream_t s);

// Device state of grid `id` of a batch (null if unknown), and the batch's
// stream (where its launch sequences and grid clears run).
GridState* batch_grid(tws_batch* b, int32_t id);
hipStream_t batch_stream(tws_batch* b);

}  // namespace tws
