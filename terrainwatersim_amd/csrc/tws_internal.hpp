// tws_internal.hpp -- device-side grid descriptor and kernel launchers shared
// by the C-ABI front ends (single simulation, batch, server).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace tws {

// Terrain height given to ghost cells in closed-wall mode.  A cell next to
// such a ghost computes a flux toward it of max(0, f*F + (h - kWallHeight)*K)
// = +0 for every K >= 0 once its own stored wall-ward f is 0 (the step
// kernels zero it when they load an edge cell), and the ghost itself emits
// (BIG * K) * (0 / S) = 0: closed walls, exactly.
constexpr float kWallHeight = 1.0e30f;
// Largest K accepted with closed walls: kWallHeight * K stays finite.
constexpr float kMaxClosedK = 1.0e8f;

constexpr int kWave = 64;

// float -> IEEE half, rounding toward zero: the conversion the reference's
// rg16f FlowMap store performs when the shaders run on Mesa
// (flowApply.comp:52).  Finite values beyond the half range give +-65504.
__host__ __device__ inline uint16_t half_rtz(float x) {
    const uint32_t u = __builtin_bit_cast(uint32_t, x);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const uint32_t a = u & 0x7fffffffu;
    if (a >= 0x7f800000u)  // Inf, NaN (quiet, top payload bits kept)
        return (uint16_t)(a == 0x7f800000u ? sign | 0x7c00u : sign | 0x7e00u | ((a >> 13) & 0x3ffu));
    if (a >= 0x47800000u) return (uint16_t)(sign | 0x7bffu);  // |x| >= 65536
    const int E = (int)(a >> 23);
    if (E >= 113) return (uint16_t)(sign | ((uint32_t)(E - 112) << 10) | ((a >> 13) & 0x3ffu));
    if (E == 0) return (uint16_t)sign;  // float denormal: below the half range
    const int shift = 126 - E;          // >= 14
    if (shift >= 32) return (uint16_t)sign;
    return (uint16_t)(sign | (((a & 0x7fffffu) | 0x800000u) >> shift));
}

// One grid as the kernels see it.  All grids of a launch live in one device
// array; blocks find theirs by tile_begin.
struct GridDesc {
    const float* b;          // terrain plane
    float* d[2];             // depth planes, ping-pong
    float* f[2][4];          // flux planes (+x,-x,+y,-y), ping-pong
    int64_t pitch;           // row pitch in floats (multiple of 64)
    int32_t W, H;
    int32_t tiles_x, tiles_y;
    int32_t tile_begin;      // first flattened tile index of this grid
    int32_t steps;           // steps this launch sequence advances the grid
    int32_t parity0;         // buffer set holding the state before the sequence
    int32_t boundary;        // TWS_BOUNDARY_*
    float F, K, C;
    float ghost_b, ghost_d, ghost_fw;  // ghost cell: terrain, depth, -y flux
    float rain_cx, rain_cy, rain_I, rain_R2;
    int32_t rain_on;
    int32_t tile_group;      // tile order: columns per group (0 = row-major)
    // Cells updated by a step ([0, upd_w) x [0, upd_h)) and reached by the
    // rain and the brush ([0, brush_w) x [0, brush_h)): the whole grid but in
    // TWS_BOUNDARY_REFERENCE_DISPATCH mode, where the reference's truncating
    // dispatch shapes leave the tail untouched.
    int32_t upd_w, upd_h, brush_w, brush_h;
    // Streaming kernel in a batch: output rows per segment (tiles_x strips by
    // tiles_y segments, tasks numbered from tile_begin).
    int32_t seg_rows;
    unsigned int* status;    // bit 0: non-finite value seen
};

// Whether a grid must step on the two-pass kernels: the tile kernels update
// every cell and rain on every cell, so they serve a reference-dispatch grid
// only when its extents cover the whole grid.
inline bool needs_twopass(const GridDesc& g) {
    return g.upd_w < g.W || g.upd_h < g.H ||
           (g.rain_on != 0 && (g.brush_w < g.W || g.brush_h < g.H));
}

// Tile geometry of the fused kernel.  CPL columns per lane, RPW rows per
// wave, NW waves, at most KH steps per launch.
template <int CPL, int RPW, int NW, int KH>
struct TileGeom {
    static constexpr int EXW = CPL * kWave;                       // tile width incl. halo
    static constexpr int EXH = RPW * NW;                          // tile height incl. halo
    static constexpr int HALOX = ((2 * KH + CPL - 1) / CPL) * CPL;  // keeps vector alignment
    static constexpr int HALOY = 2 * KH;
    static constexpr int OX = EXW - 2 * HALOX;                    // output columns per tile
    static constexpr int OY = EXH - 2 * HALOY;                    // output rows per tile
    static_assert(OX > 0 && OY > 0, "halo larger than tile");
};

// Output tile size of a (variant, k) pair; used to lay out tile lists.
struct TileShape {
    int ox, oy, threads;
};
bool tile_shape(int kernel, int k, TileShape* out);

// Step-kernel modes: what a launch compiles in.  Plain: closed walls without
// rain.  Rain: closed walls, rain on some grids.  General: any edge mode.
// Each mode runs every grid a lower mode can.
constexpr int kModePlain = 0, kModeRain = 1, kModeGeneral = 2;

// Launchers (stream-ordered, return hipError_t).
// Batch launch: hdr_dev = {number of grids, number of tiles, offset of the
// first grid in grids_dev}, the grids ordered by tile_begin; `blocks` is an
// upper bound of the tile count (extra blocks exit).  `mode` must be at least
// kernel_mode(g) of every grid of the table.
int kernel_mode(const GridDesc& g);
hipError_t launch_tile_batch(int kernel, int k, int mode, const GridDesc* grids_dev,
                             const int* hdr_dev, int blocks, int launch_idx, hipStream_t s);
hipError_t launch_twopass(const GridDesc* g_dev, const GridDesc& g_host, int parity,
                          hipStream_t s);
hipError_t launch_brush(const GridDesc& g, int parity, float cx, float cy, float I, float R2,
                        hipStream_t s);
hipError_t launch_velocity(const GridDesc& g, int parity, float* v2, hipStream_t s);
hipError_t launch_tile_single(int kernel, int k, const GridDesc& g, int launch_idx,
                              hipStream_t s);
hipError_t launch_interleave(const GridDesc& g, int parity, float* staging4, hipStream_t s);
hipError_t launch_deinterleave(const GridDesc& g, int parity, const float* staging4,
                               hipStream_t s);
hipError_t launch_ref_unpack(const GridDesc& g, int parity, const float* terrain_rgba,
                             const float* flow_rgba, float* b_out, hipStream_t s);
hipError_t launch_ref_pack(const GridDesc& g, int parity, float gb_fill, float* terrain_rgba,
                           float* flow_rgba, uint16_t* flowmap, hipStream_t s);
hipError_t launch_volume_partials(const GridDesc& g, int parity, double* partials, int nparts,
                                  hipStream_t s);
hipError_t launch_zero_fill(float* p, int64_t n, hipStream_t s);
hipError_t launch_lane_shift_probe(int* out128, hipStream_t s);
hipError_t launch_check_finite(const GridDesc& g, int parity, hipStream_t s);
hipError_t launch_fbm(float* b, const GridDesc& g, uint64_t seed, int octaves, int base,
                      double persistence, double scale, hipStream_t s);
hipError_t launch_init_water(const GridDesc& g, int parity, float depth, int x0, int y0, int x1,
                             int y1, float rect_depth, hipStream_t s);
// Copy of `bytes` (a multiple of 16) in launch shape `shape`, 0 <= shape <
// stream_copy_shapes() (loads in flight per lane, non-temporal hints, work
// groups per CU; shape 0 is the one-load-per-lane grid-stride loop).
int stream_copy_shapes();
hipError_t launch_stream_copy(const void* src, void* dst, int64_t bytes, int shape,
                              hipStream_t s);

// The streaming step kernel (tws_stream.hip): one wavefront per strip
// segment of `seg_rows` output rows and stream_shape() output columns.
struct StreamPlan {
    int strips = 0, segs = 0, seg_rows = 0, blocks = 0;
};
// Output columns per strip at k steps per launch; false if k is not built.
bool stream_shape(int k, int* ox);
// A grid's tasks in a batched launch of the streaming kernel: strips of
// stream_shape() output columns by segments of seg_rows rows, seg_rows the
// largest even split of H into segments of at most `cap` rows.
bool stream_batch_plan(int W, int H, int k, int cap, int* strips, int* segs, int* seg_rows);
// Segment cap (rows) for a batched launch of the streaming kernel over the
// grids shapes[2i] x shapes[2i+1], i < n, in task order, on a device of `cus`
// compute units: the cap whose tasks, dealt to the XCDs and run on their
// resident work-group slots in order, finish soonest.  May throw bad_alloc.
int stream_batch_cap(const int* shapes, int n, int k, int cus);
// Batched launch launch_idx of the streaming kernel (closed walls; `mode`
// kModePlain or kModeRain, the highest of the table's grids): hdr_dev =
// {grids, tasks, offset}; `blocks` is at least the work groups the tasks need.
hipError_t launch_stream_batch(int k, int mode, const GridDesc* grids_dev, const int* hdr_dev,
                               int blocks, int launch_idx, hipStream_t s);
// Segments of a W x H grid on a device of `cus` compute units: as many as
// fill its wavefront slots once.
StreamPlan stream_plan(int W, int H, int k, int cus);
// One launch of k steps (g.steps of them) from set `parity`; closed walls only.
hipError_t launch_stream_single(int k, const GridDesc& g, const StreamPlan& p, int parity,
                                hipStream_t s);

constexpr int kVolumeParts = 512;

}  // namespace tws
