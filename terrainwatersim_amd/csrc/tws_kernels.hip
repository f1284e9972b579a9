// tws_kernels.hip -- gfx950 kernels of the virtual-pipe water step.
//
// Hot path: k_tile, the fused and temporally blocked step.  One work group
// owns an EXW x EXH tile of cells (the output tile plus a halo of 2 cells per
// step on each side) and advances it up to KH steps without touching HBM:
//   - each lane keeps CPL adjacent columns x RPW rows of depth d and flux f in
//     VGPRs for the whole launch; terrain b sits in LDS;
//   - x-neighbours come from the adjacent lane through DPP wave shifts,
//     y-neighbours from the lane's own registers, and only the first and last
//     row of each wave go through LDS;
//   - per step, two work-group barriers separate the flux phase
//     (flowUpdate.comp:29-62) from the depth phase (flowApply.comp:32-41).
// The halo region goes stale by 2 cells per step (flux needs h at radius 1,
// depth needs flux at radius 1), so after KH steps exactly the interior
// OX x OY cells are valid and stored.  HBM traffic per launch is one read of
// b, d, f and one write of d, f per output cell (plus the halo re-read),
// shared by KH steps.
//
// Every expression keeps the shaders' operation order and the library is
// compiled with -ffp-contract=off, so each cell's result is bit-identical to
// the FP32 restatement regardless of tile shape or steps per launch.

#include <algorithm>

#include "tws_internal.hpp"
#include "tws.h"
#include "tws_fbm.hpp"
#include "tws_kernel_util.hpp"

namespace tws {

// ---------------------------------------------------------------------------
// Tile (tx, ty) of a grid's tile list position lt: groups of tile_group tile
// columns, row by row within a group, so the tiles in flight at once form a
// compact block whose shared halos meet in L2 (tile_group 0: row-major).
__device__ __forceinline__ void tile_coords(const GridDesc& g, int lt, int& tx, int& ty) {
    if (g.tile_group > 0) {
        const int per_group = g.tile_group * g.tiles_y;
        const int grp = lt / per_group, in_grp = lt - grp * per_group;
        const int first = grp * g.tile_group;
        const int gw = min(g.tiles_x - first, g.tile_group);
        ty = in_grp / gw;
        tx = first + (in_grp - ty * gw);
    } else {
        ty = lt / g.tiles_x;
        tx = lt - ty * g.tiles_x;
    }
}

// One tile: load its EXW x EXH block of state set `parity`, advance it
// `levels` steps, store the valid interior into the other set.  Returns
// whether this lane stored a non-finite value.  MODE (kModePlain,
// kModeRain, kModeGeneral) compiles in only what the launch's grids use:
// the plain kernel has neither the rain source nor the reference-edge
// masking (no per-cell domain masks, so no scalar spills and 111 instead of
// 128 VGPRs at 128x64: +2.6% at k = 4, +7% at k = 6); the rain kernel has
// the rain but not the edge masking (6 instead of 23 scalar spills: +0.6% to
// +1.8% on closed-wall grids with rain).
template <int CPL, int RPW, int NW, int KH, int MODE>
__device__ __forceinline__ bool tile_step(const GridDesc& g, int tx, int ty, int levels,
                                          int parity) {
    constexpr bool GEN = MODE != kModePlain;   // rain source compiled in
    constexpr bool EDGE = MODE == kModeGeneral;  // reference-edge masking compiled in
    using G = TileGeom<CPL, RPW, NW, KH>;
    constexpr int EXW = G::EXW, EXH = G::EXH;
    __shared__ float s_b[EXH * EXW];
    __shared__ float s_xh[2][NW][CPL][kWave];  // [0]: first row of a wave, [1]: last row
    __shared__ float s_xf[2][NW][CPL][kWave];

    const int ox = tx * G::OX - G::HALOX;
    const int oy = ty * G::OY - G::HALOY;
    const int W = g.W, H = g.H;
    const int64_t pitch = g.pitch;
    const float F = g.F, K = g.K, C = g.C;

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lx = lane * CPL;   // tile-local column of this lane's first cell
    const int ly = wave * RPW;   // tile-local row of this wave's first row
    const int x0 = ox + lx;
    // a wave whose rows lie wholly outside the rain disc skips the brush
    // (wave-uniform, decided before the state occupies the registers)
    const bool rain =
        GEN && g.rain_on != 0 &&
        __builtin_amdgcn_readfirstlane(brush_reaches(g.rain_cx, g.rain_cy, g.rain_R2, ox,
                                                     ox + EXW - 1, oy + ly, oy + ly + RPW - 1)
                                           ? 1 : 0) != 0;

    float d[RPW][CPL], f0[RPW][CPL], f1[RPW][CPL], f2[RPW][CPL], f3[RPW][CPL];
    unsigned inside = 0;  // bit r*CPL+c: cell inside the domain
    bool bad = false;

    // ---- load -------------------------------------------------------------
    {
        const gfloat* __restrict__ pb = gptr(g.b);
        const gfloat* __restrict__ pd = gptr(parity ? g.d[1] : g.d[0]);
        const gfloat* __restrict__ p0 = gptr(parity ? g.f[1][0] : g.f[0][0]);
        const gfloat* __restrict__ p1 = gptr(parity ? g.f[1][1] : g.f[0][1]);
        const gfloat* __restrict__ p2 = gptr(parity ? g.f[1][2] : g.f[0][2]);
        const gfloat* __restrict__ p3 = gptr(parity ? g.f[1][3] : g.f[0][3]);
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int y = oy + ly + r;
            const bool vin = (unsigned)y < (unsigned)H && x0 >= 0 && x0 < W;
            float vb[CPL], vd[CPL], v0[CPL], v1[CPL], v2[CPL], v3[CPL];
#pragma unroll
            for (int c = 0; c < CPL; ++c) vb[c] = vd[c] = v0[c] = v1[c] = v2[c] = v3[c] = 0.0f;
            if (vin) {
                const int64_t off = (int64_t)y * pitch + x0;
                vload<CPL>(pb + off, vb);
                vload<CPL>(pd + off, vd);
                vload<CPL>(p0 + off, v0);
                vload<CPL>(p1 + off, v1);
                vload<CPL>(p2 + off, v2);
                vload<CPL>(p3 + off, v3);
            }
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const bool in = vin && (x0 + c < W);
                inside |= (in ? 1u : 0u) << (r * CPL + c);
                s_b[(ly + r) * EXW + lx + c] = in ? vb[c] : g.ghost_b;
                d[r][c] = in ? vd[c] : g.ghost_d;
                f0[r][c] = in ? v0[c] : 0.0f;
                f1[r][c] = in ? v1[c] : 0.0f;
                f2[r][c] = in ? v2[c] : 0.0f;
                f3[r][c] = in ? v3[c] : g.ghost_fw;
            }
        }
    }
    const bool edge_tile = ox < 0 || oy < 0 || ox + EXW > W || oy + EXH > H;
    const bool need_mask = EDGE && g.boundary != TWS_BOUNDARY_CLOSED && edge_tile;
    // Closed walls carry no flux either way.  The wall ghost makes every
    // wall-ward flux the kernel computes max(0, f*F + (h - 1e30)*K) = +0 as
    // long as the cell's own stored wall-ward f is 0; an uploaded state may
    // hold another value there (and with K = 0 it would pass the wall), so
    // edge cells drop it as they load.
    if ((!EDGE || g.boundary == TWS_BOUNDARY_CLOSED) && edge_tile) {
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int y = oy + ly + r;
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const int x = x0 + c;
                if (x == W - 1) f0[r][c] = 0.0f;
                if (x == 0) f1[r][c] = 0.0f;
                if (y == H - 1) f2[r][c] = 0.0f;
                if (y == 0) f3[r][c] = 0.0f;
            }
        }
    }
    // The brush amounts do not change from level to level: computed once, in
    // registers across the loop (recomputing them per level measured 13%
    // slower at 256^2, where waves under the disc are on the critical path),
    // and only by waves under the disc.  Left to itself the compiler hoisted
    // the divisions out of the loop into every wave.
    float wet[RPW][CPL];
    if (rain) {
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int c = 0; c < CPL; ++c)
                wet[r][c] = brush_amount(g.rain_cx, g.rain_cy, g.rain_I, g.rain_R2, x0 + c,
                                         oy + ly + r);
    } else {
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int c = 0; c < CPL; ++c) wet[r][c] = 0.0f;
    }
    __syncthreads();  // s_b complete

    for (int lev = 0; lev < levels; ++lev) {
        if (rain) {
#pragma unroll
            for (int r = 0; r < RPW; ++r)
#pragma unroll
                for (int c = 0; c < CPL; ++c)
                    if (inside & (1u << (r * CPL + c))) d[r][c] = d[r][c] + wet[r][c];
        }

        // ---- flux phase: flowUpdate.comp:34-62 ------------------------------
        float h[RPW][CPL];
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int c = 0; c < CPL; ++c) h[r][c] = d[r][c] + s_b[(ly + r) * EXW + lx + c];
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            s_xh[0][wave][c][lane] = h[0][c];
            s_xh[1][wave][c][lane] = h[RPW - 1][c];
        }
        __syncthreads();
        float h_up[CPL], h_dn[CPL];  // rows y-1 of the first row and y+1 of the last row
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            h_up[c] = wave > 0 ? s_xh[1][wave - 1][c][lane] : h[0][c];
            h_dn[c] = wave < NW - 1 ? s_xh[0][wave + 1][c][lane] : h[RPW - 1][c];
        }
        // Rows that can still reach the output: the flux of level lev is
        // needed on rows [2 lev + 1, EXH - 2 lev - 1), the depth on
        // [2 lev + 2, EXH - 2 lev - 2).  Rows outside only feed rows outside,
        // so a wave may skip them (wave-uniform branches).  Measured: +12-14%
        // for 64x64 tiles at 256^2-512^2, +2-5% for the general kernel at
        // 2048^2-4096^2, but -2% for the plain kernels whose lanes hold 16
        // cells (the branches cost more than the rows they save there).
        constexpr bool kSkipRows = GEN || CPL * RPW < 16;
        const int flo = 2 * lev + 1 - ly, fhi = EXH - 2 * lev - 1 - ly;
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            if (kSkipRows && (r < flo || r >= fhi)) continue;
            const float h_left = from_left_lane_or(h[r][CPL - 1], h[r][0]);
            const float h_right = from_right_lane_or(h[r][0], h[r][CPL - 1]);
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const float own = h[r][c];
                const float hx1 = c < CPL - 1 ? h[r][c + 1] : h_right;
                const float hx0 = c > 0 ? h[r][c - 1] : h_left;
                const float hy1 = r < RPW - 1 ? h[r + 1][c] : h_dn[c];
                const float hy0 = r > 0 ? h[r - 1][c] : h_up[c];
                flux_update(F, K, C, d[r][c], own - hx1, own - hx0, own - hy1, own - hy0,
                            f0[r][c], f1[r][c], f2[r][c], f3[r][c]);
            }
        }
        if (need_mask) {
#pragma unroll
            for (int r = 0; r < RPW; ++r)
#pragma unroll
                for (int c = 0; c < CPL; ++c)
                    if (!(inside & (1u << (r * CPL + c)))) {
                        f0[r][c] = 0.0f; f1[r][c] = 0.0f; f2[r][c] = 0.0f;
                        f3[r][c] = g.ghost_fw;
                    }
        }

        // ---- depth phase: flowApply.comp:32-41 -------------------------------
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            s_xf[0][wave][c][lane] = f3[0][c];        // -y flux of the first row
            s_xf[1][wave][c][lane] = f2[RPW - 1][c];  // +y flux of the last row
        }
        __syncthreads();
        float f_up[CPL], f_dn[CPL];  // +y flux of row y-1, -y flux of row y+1
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            f_up[c] = wave > 0 ? s_xf[1][wave - 1][c][lane] : 0.0f;
            f_dn[c] = wave < NW - 1 ? s_xf[0][wave + 1][c][lane] : 0.0f;
        }
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            if (kSkipRows && (r < flo + 1 || r >= fhi - 1)) continue;
            const float in_left = from_left_lane(f0[r][CPL - 1]);   // X0.x
            const float in_right = from_right_lane(f1[r][0]);      // X1.y
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const float X1y = c < CPL - 1 ? f1[r][c + 1] : in_right;
                const float X0x = c > 0 ? f0[r][c - 1] : in_left;
                const float Y1w = r < RPW - 1 ? f3[r + 1][c] : f_dn[c];
                const float Y0z = r > 0 ? f2[r - 1][c] : f_up[c];
                const float in = ((X1y + X0x) + Y1w) + Y0z;
                const float out = ((f0[r][c] + f1[r][c]) + f2[r][c]) + f3[r][c];
                d[r][c] = fmaxf(0.0f, d[r][c] + (in - out) * C);
            }
        }
        if (need_mask) {
#pragma unroll
            for (int r = 0; r < RPW; ++r)
#pragma unroll
                for (int c = 0; c < CPL; ++c)
                    if (!(inside & (1u << (r * CPL + c)))) d[r][c] = g.ghost_d;
        }
    }

    // ---- store the valid interior ------------------------------------------
    {
        gfloat* __restrict__ pd = gptr(parity ? g.d[0] : g.d[1]);
        gfloat* __restrict__ p0 = gptr(parity ? g.f[0][0] : g.f[1][0]);
        gfloat* __restrict__ p1 = gptr(parity ? g.f[0][1] : g.f[1][1]);
        gfloat* __restrict__ p2 = gptr(parity ? g.f[0][2] : g.f[1][2]);
        gfloat* __restrict__ p3 = gptr(parity ? g.f[0][3] : g.f[1][3]);
        const bool col_ok = lx >= G::HALOX && lx < EXW - G::HALOX && x0 < W;
        const bool full = x0 + CPL <= W;
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int yl = ly + r;
            const int y = oy + yl;
            if (!(col_ok && yl >= G::HALOY && yl < EXH - G::HALOY && y < H)) continue;
            const int64_t off = (int64_t)y * pitch + x0;
#pragma unroll
            for (int c = 0; c < CPL; ++c)
                bad |= (full || x0 + c < W) &
                       nonfinite5(d[r][c], f0[r][c], f1[r][c], f2[r][c], f3[r][c]);
            if (full) {
                vstore<CPL>(pd + off, d[r]);
                vstore<CPL>(p0 + off, f0[r]);
                vstore<CPL>(p1 + off, f1[r]);
                vstore<CPL>(p2 + off, f2[r]);
                vstore<CPL>(p3 + off, f3[r]);
            } else {
#pragma unroll
                for (int c = 0; c < CPL; ++c)
                    if (x0 + c < W) {
                        pd[off + c] = d[r][c];
                        p0[off + c] = f0[r][c];
                        p1[off + c] = f1[r][c];
                        p2[off + c] = f2[r][c];
                        p3[off + c] = f3[r][c];
                    }
            }
        }
    }
    return bad;
}

// ---------------------------------------------------------------------------
// Fused, temporally blocked step.  Grid: one block per tile of the flattened
// tile list.  Single grid: `grids` is null, g1 (by value) is the grid and
// the tile count is gridDim.x.  Batch: hdr = {number of grids, number of
// tiles, offset} and grids[offset..] (ordered by tile_begin) live in device
// memory, so a captured launch stays valid while grids come and go; blocks past the tile
// count exit at once.  MODE is the highest kernel_mode of the launch's grids
// (chosen on the host from the descriptors).
// At least 4 waves per SIMD (at most 128 VGPRs): 2 work groups of 512 per CU.
template <int CPL, int RPW, int NW, int KH, int MODE>
__global__ void __launch_bounds__(NW * kWave, 4)
k_tile(GridDesc g1, const GridDesc* __restrict__ grids, const int* __restrict__ hdr,
       int launch_idx) {
    // XCD-aware order: the hardware deals blocks to the 8 XCDs round robin;
    // give each XCD a contiguous run of tiles so neighbouring tiles (whose
    // halos overlap) share one L2.
    const int ngrids = grids ? hdr[0] : 0;
    const int total_tiles = grids ? hdr[1] : (int)gridDim.x;
    if (grids) grids += hdr[2];  // this launch's segment of the batch table
    const int bid = blockIdx.x;
    if (bid >= total_tiles) return;
    const int xcd = bid & 7, idx = bid >> 3;
    const int q = total_tiles >> 3, rmd = total_tiles & 7;
    const int tile = xcd * q + (xcd < rmd ? xcd : rmd) + idx;

    GridDesc g;
    if (!grids) {
        g = g1;
    } else {
        int lo = 0, hi = ngrids - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (grids[mid].tile_begin <= tile) lo = mid; else hi = mid - 1;
        }
        g = grids[lo];
    }
    const int levels = min(KH, g.steps - launch_idx * KH);
    if (levels <= 0) return;
    const int parity = (g.parity0 + launch_idx) & 1;
    // Every other launch walks the grid's tiles backwards, so it starts on the
    // tiles the launch before wrote last, whose state may still sit in the
    // memory-side cache.  Tiles of one launch are independent: the order does
    // not change results.
    int lt = tile - g.tile_begin;
    if (parity) lt = g.tiles_x * g.tiles_y - 1 - lt;
    int tx, ty;
    tile_coords(g, lt, tx, ty);
    const bool bad = tile_step<CPL, RPW, NW, KH, MODE>(g, tx, ty, levels, parity);
    if (__ballot(bad) != 0 && (threadIdx.x & (kWave - 1)) == 0) atomicOr(g.status, 1u);
}

// ---------------------------------------------------------------------------
// Two-pass step with the reference's structure: a flux kernel and an apply
// kernel per step, neighbours read from memory.  Reads set `parity`, writes
// flux to set parity^1 (flux pass) and depth to set parity^1 (apply pass).
__global__ void __launch_bounds__(256) k_flux_twopass(GridDesc g, int parity) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= g.W || y >= g.H) return;
    const int64_t p = g.pitch, i = (int64_t)y * p + x;
    if (x >= g.upd_w || y >= g.upd_h) {  // no reference work group covers it
        for (int c = 0; c < 4; ++c) g.f[parity ^ 1][c][i] = g.f[parity][c][i];
        return;
    }
    const float* b = g.b;
    const float* d = g.d[parity];
    const bool in_x1 = x + 1 < g.W, in_x0 = x > 0, in_y1 = y + 1 < g.H, in_y0 = y > 0;
    const float gh = g.ghost_d + g.ghost_b;  // never used in closed mode
    const float own = d[i] + b[i];
    const float hx1 = in_x1 ? d[i + 1] + b[i + 1] : gh;
    const float hx0 = in_x0 ? d[i - 1] + b[i - 1] : gh;
    const float hy1 = in_y1 ? d[i + p] + b[i + p] : gh;
    const float hy0 = in_y0 ? d[i - p] + b[i - p] : gh;
    float o0 = g.f[parity][0][i], o1 = g.f[parity][1][i], o2 = g.f[parity][2][i],
          o3 = g.f[parity][3][i];
    if (g.boundary == TWS_BOUNDARY_CLOSED) {
        // the flux toward a wall is clamped to zero before the scaling sum
        float t0 = fmaxf(0.0f, o0 * g.F + (own - hx1) * g.K);
        float t1 = fmaxf(0.0f, o1 * g.F + (own - hx0) * g.K);
        float t2 = fmaxf(0.0f, o2 * g.F + (own - hy1) * g.K);
        float t3 = fmaxf(0.0f, o3 * g.F + (own - hy0) * g.K);
        if (!in_x1) t0 = 0.0f;
        if (!in_x0) t1 = 0.0f;
        if (!in_y1) t2 = 0.0f;
        if (!in_y0) t3 = 0.0f;
        const float total = ((t0 + t1) + (t2 + t3)) * g.C;
        if (total > d[i]) {
            const float s = d[i] / total;
            t0 = t0 * s; t1 = t1 * s; t2 = t2 * s; t3 = t3 * s;
        }
        o0 = t0; o1 = t1; o2 = t2; o3 = t3;
    } else {
        flux_update(g.F, g.K, g.C, d[i], own - hx1, own - hx0, own - hy1, own - hy0, o0, o1, o2,
                    o3);
    }
    g.f[parity ^ 1][0][i] = o0;
    g.f[parity ^ 1][1][i] = o1;
    g.f[parity ^ 1][2][i] = o2;
    g.f[parity ^ 1][3][i] = o3;
}

__global__ void __launch_bounds__(256) k_apply_twopass(GridDesc g, int parity) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= g.W || y >= g.H) return;
    const int64_t p = g.pitch, i = (int64_t)y * p + x;
    if (x >= g.upd_w || y >= g.upd_h) {  // no reference work group covers it
        g.d[parity ^ 1][i] = g.d[parity][i];
        return;
    }
    float* const* fo = g.f[parity ^ 1];
    const float gz = g.boundary == TWS_BOUNDARY_CLOSED ? 0.0f : g.ghost_fw;
    const float X1y = x + 1 < g.W ? fo[1][i + 1] : 0.0f;
    const float X0x = x > 0 ? fo[0][i - 1] : 0.0f;
    const float Y1w = y + 1 < g.H ? fo[3][i + p] : gz;
    const float Y0z = y > 0 ? fo[2][i - p] : 0.0f;
    const float in = ((X1y + X0x) + Y1w) + Y0z;
    const float out = ((fo[0][i] + fo[1][i]) + fo[2][i]) + fo[3][i];
    g.d[parity ^ 1][i] = fmaxf(0.0f, g.d[parity][i] + (in - out) * g.C);
}

// One brush application to the depth plane of set `parity`, in place.
__global__ void __launch_bounds__(256) k_brush(GridDesc g, int parity, float cx, float cy,
                                                float I, float R2) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= g.brush_w || y >= g.brush_h) return;
    const int64_t i = (int64_t)y * g.pitch + x;
    float* d = g.d[parity];
    d[i] = d[i] + brush_amount(cx, cy, I, R2, x, y);
}

// Directed flow from the current flux, flowApply.comp:45-46; dense output
// v2[(x + y*W)*2 + {0,1}].
__global__ void __launch_bounds__(256) k_velocity(GridDesc g, int parity, float* v2) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= g.W || y >= g.H) return;
    const int64_t p = g.pitch, i = (int64_t)y * p + x;
    float* const* f = g.f[parity];
    const float gz = g.boundary == TWS_BOUNDARY_CLOSED ? 0.0f : g.ghost_fw;
    const float X1y = x + 1 < g.W ? f[1][i + 1] : 0.0f;
    const float X0x = x > 0 ? f[0][i - 1] : 0.0f;
    const float Y1w = y + 1 < g.H ? f[3][i + p] : gz;
    const float Y0z = y > 0 ? f[2][i - p] : 0.0f;
    const int64_t o = (int64_t)y * g.W + x;
    // a cell no reference work group covers never receives a FlowMap value
    const bool upd = x < g.upd_w && y < g.upd_h;
    v2[o * 2 + 0] = upd ? (X1y - f[0][i]) - (X0x - f[1][i]) : 0.0f;
    v2[o * 2 + 1] = upd ? (Y1w - f[2][i]) - (Y0z - f[3][i]) : 0.0f;
}

// Planar flux of set `parity` <-> dense interleaved staging[(x + y*W)*4 + c].
__global__ void __launch_bounds__(256) k_interleave(GridDesc g, int parity, float* st) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= g.W || y >= g.H) return;
    const int64_t i = (int64_t)y * g.pitch + x, o = (int64_t)y * g.W + x;
    *reinterpret_cast<float4*>(st + o * 4) =
        make_float4(g.f[parity][0][i], g.f[parity][1][i], g.f[parity][2][i], g.f[parity][3][i]);
}
__global__ void __launch_bounds__(256) k_deinterleave(GridDesc g, int parity, const float* st) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= g.W || y >= g.H) return;
    const int64_t i = (int64_t)y * g.pitch + x, o = (int64_t)y * g.W + x;
    const float4 v = *reinterpret_cast<const float4*>(st + o * 4);
    g.f[parity][0][i] = v.x; g.f[parity][1][i] = v.y;
    g.f[parity][2][i] = v.z; g.f[parity][3][i] = v.w;
}

// Reference interleaved layout (Terrain.cpp:205-237) -> planes.
__global__ void __launch_bounds__(256) k_ref_unpack(GridDesc g, int parity, const float* terrain,
                                                     const float* flow, float* b_out) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= g.W || y >= g.H) return;
    const int64_t i = (int64_t)y * g.pitch + x, o = (int64_t)y * g.W + x;
    if (terrain) {
        const float4 t = *reinterpret_cast<const float4*>(terrain + o * 4);
        b_out[i] = t.x;
        g.d[parity][i] = t.w;
    }
    if (flow) {
        const float4 v = *reinterpret_cast<const float4*>(flow + o * 4);
        g.f[parity][0][i] = v.x; g.f[parity][1][i] = v.y;
        g.f[parity][2][i] = v.z; g.f[parity][3][i] = v.w;
    }
}

// Planes -> reference layout; FlowMap as rg16f bits (flowApply.comp:52).
__global__ void __launch_bounds__(256) k_ref_pack(GridDesc g, int parity, float gb_fill,
                                                   float* terrain, float* flow, uint16_t* flowmap) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= g.W || y >= g.H) return;
    const int64_t p = g.pitch, i = (int64_t)y * p + x, o = (int64_t)y * g.W + x;
    float* const* f = g.f[parity];
    if (terrain)
        *reinterpret_cast<float4*>(terrain + o * 4) = make_float4(g.b[i], gb_fill, gb_fill, g.d[parity][i]);
    if (flow)
        *reinterpret_cast<float4*>(flow + o * 4) = make_float4(f[0][i], f[1][i], f[2][i], f[3][i]);
    if (flowmap) {
        const float gz = g.boundary == TWS_BOUNDARY_CLOSED ? 0.0f : g.ghost_fw;
        const float X1y = x + 1 < g.W ? f[1][i + 1] : 0.0f;
        const float X0x = x > 0 ? f[0][i - 1] : 0.0f;
        const float Y1w = y + 1 < g.H ? f[3][i + p] : gz;
        const float Y0z = y > 0 ? f[2][i - p] : 0.0f;
        const bool upd = x < g.upd_w && y < g.upd_h;  // see k_velocity
        flowmap[o * 2 + 0] = upd ? half_rtz((X1y - f[0][i]) - (X0x - f[1][i])) : (uint16_t)0;
        flowmap[o * 2 + 1] = upd ? half_rtz((Y1w - f[2][i]) - (Y0z - f[3][i])) : (uint16_t)0;
    }
}

// FP64 partial sums of depth: block p takes rows p, p+nparts, ...; threads
// stride the columns; a fixed tree combines them.  Deterministic.
__global__ void __launch_bounds__(256) k_volume(GridDesc g, int parity, double* partials,
                                                 int nparts) {
    __shared__ double s[256];
    double acc = 0.0;
    const float* d = g.d[parity];
    for (int y = blockIdx.x; y < g.H; y += nparts)
        for (int x = threadIdx.x; x < g.W; x += 256) acc += (double)d[(int64_t)y * g.pitch + x];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(256) k_zero_fill(float* p, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0.0f;
}

// Finiteness of the state of set `parity` (b, d, f) after an upload: sets
// bit 0 of *status if any value inside the domain is NaN or Inf.  The step
// kernel itself only checks what it stores (its max(0, .) clamps would turn
// a NaN input into 0).
__global__ void __launch_bounds__(256) k_check_finite(GridDesc g, int parity) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    bool bad = false;
    if (x < g.W && y < g.H) {
        const int64_t i = (int64_t)y * g.pitch + x;
        bad = nonfinite5(g.b[i], g.d[parity][i], g.f[parity][0][i], g.f[parity][1][i],
                         g.f[parity][2][i]) |
              !__builtin_isfinite(g.f[parity][3][i]);
    }
    if (__ballot(bad) != 0 && (threadIdx.x & 63) == 0) atomicOr(g.status, 1u);
}

// Seeded fBm terrain straight into the b plane (same bits as the host
// generator: tws_fbm.hpp).
__global__ void __launch_bounds__(256) k_fbm(float* b, int64_t pitch, int W, int H,
                                              uint64_t seed, int octaves, int base,
                                              double persistence, double scale) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    b[(int64_t)y * pitch + x] = fbm_cell(x, y, W, H, seed, octaves, base, persistence, scale);
}

// Water of set `parity`: rect_depth inside [x0, x1) x [y0, y1), `depth`
// elsewhere; flux zero.
__global__ void __launch_bounds__(256) k_init_water(GridDesc g, int parity, float depth, int x0,
                                                     int y0, int x1, int y1, float rect_depth) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= g.W || y >= g.H) return;
    const int64_t i = (int64_t)y * g.pitch + x;
    const bool in = x >= x0 && x < x1 && y >= y0 && y < y1;
    g.d[parity][i] = in ? rect_depth : depth;
    for (int c = 0; c < 4; ++c) g.f[parity][c][i] = 0.0f;
}

// Device copy of n float4 (STREAM copy): the measured HBM ceiling, and the
// known-size reference that scales the FETCH_SIZE / WRITE_SIZE counters.
// Each lane keeps U loads in flight (issued back to back, then U stores);
// NT marks them non-temporal.  Chunks of U * blockDim float4 are dealt out
// grid-stride; the tail chunk is bounds-checked per element.
template <int U, bool NT>
__global__ void __launch_bounds__(256) k_stream_copy(const vf4* __restrict__ src,
                                                      vf4* __restrict__ dst, int64_t n) {
    const int64_t chunk = (int64_t)U * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * chunk; base < n; base += (int64_t)gridDim.x * chunk) {
        vf4 v[U];
        if (base + chunk <= n) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t i = base + (int64_t)u * blockDim.x + threadIdx.x;
                v[u] = NT ? __builtin_nontemporal_load(src + i) : src[i];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t i = base + (int64_t)u * blockDim.x + threadIdx.x;
                if (NT) __builtin_nontemporal_store(v[u], dst + i); else dst[i] = v[u];
            }
        } else {
            for (int u = 0; u < U; ++u) {
                const int64_t i = base + (int64_t)u * blockDim.x + threadIdx.x;
                if (i < n) dst[i] = src[i];
            }
        }
    }
}

// Lane-layout check of the DPP wave shifts used by k_tile.
__global__ void __launch_bounds__(64) k_lane_shift_probe(int* out) {
    const int lane = threadIdx.x;
    // lane i holds i + 1, so a lane without a source (which reads 0) shows as 0
    out[lane] = __float_as_int(from_left_lane(__int_as_float(lane + 1)));
    out[64 + lane] = __float_as_int(from_right_lane(__int_as_float(lane + 1)));
}

// ---------------------------------------------------------------------------
// Host-side launch helpers.

static dim3 cell_grid(const GridDesc& g) { return dim3((g.W + 15) / 16, (g.H + 15) / 16); }

template <int CPL, int RPW, int NW, int MODE>
static hipError_t launch_tile_kg(int k, const GridDesc* grids_dev, const int* hdr,
                                 const GridDesc& g1, int blocks, int launch_idx, hipStream_t s) {
    const dim3 grid(blocks), block(NW * kWave);
    switch (k) {
        case 1: hipLaunchKernelGGL((k_tile<CPL, RPW, NW, 1, MODE>), grid, block, 0, s, g1, grids_dev, hdr, launch_idx); break;
        case 2: hipLaunchKernelGGL((k_tile<CPL, RPW, NW, 2, MODE>), grid, block, 0, s, g1, grids_dev, hdr, launch_idx); break;
        case 3: hipLaunchKernelGGL((k_tile<CPL, RPW, NW, 3, MODE>), grid, block, 0, s, g1, grids_dev, hdr, launch_idx); break;
        case 4: hipLaunchKernelGGL((k_tile<CPL, RPW, NW, 4, MODE>), grid, block, 0, s, g1, grids_dev, hdr, launch_idx); break;
        case 5: hipLaunchKernelGGL((k_tile<CPL, RPW, NW, 5, MODE>), grid, block, 0, s, g1, grids_dev, hdr, launch_idx); break;
        case 6: hipLaunchKernelGGL((k_tile<CPL, RPW, NW, 6, MODE>), grid, block, 0, s, g1, grids_dev, hdr, launch_idx); break;
        case 8: hipLaunchKernelGGL((k_tile<CPL, RPW, NW, 8, MODE>), grid, block, 0, s, g1, grids_dev, hdr, launch_idx); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int CPL, int RPW, int NW>
static hipError_t launch_tile_k(int k, int mode, const GridDesc* grids_dev, const int* hdr,
                                const GridDesc* single, int blocks, int launch_idx,
                                hipStream_t s) {
    GridDesc g1{};
    if (single) g1 = *single;
    switch (mode) {
        case kModePlain: return launch_tile_kg<CPL, RPW, NW, kModePlain>(k, grids_dev, hdr, g1, blocks, launch_idx, s);
        case kModeRain: return launch_tile_kg<CPL, RPW, NW, kModeRain>(k, grids_dev, hdr, g1, blocks, launch_idx, s);
        default: return launch_tile_kg<CPL, RPW, NW, kModeGeneral>(k, grids_dev, hdr, g1, blocks, launch_idx, s);
    }
}

template <int CPL, int RPW, int NW>
static void shape_of(int k, TileShape* out) {
    int ox = 0, oy = 0;
    switch (k) {
        case 1: ox = TileGeom<CPL, RPW, NW, 1>::OX; oy = TileGeom<CPL, RPW, NW, 1>::OY; break;
        case 2: ox = TileGeom<CPL, RPW, NW, 2>::OX; oy = TileGeom<CPL, RPW, NW, 2>::OY; break;
        case 4: ox = TileGeom<CPL, RPW, NW, 4>::OX; oy = TileGeom<CPL, RPW, NW, 4>::OY; break;
        case 3: ox = TileGeom<CPL, RPW, NW, 3>::OX; oy = TileGeom<CPL, RPW, NW, 3>::OY; break;
        case 5: ox = TileGeom<CPL, RPW, NW, 5>::OX; oy = TileGeom<CPL, RPW, NW, 5>::OY; break;
        case 6: ox = TileGeom<CPL, RPW, NW, 6>::OX; oy = TileGeom<CPL, RPW, NW, 6>::OY; break;
        case 8: ox = TileGeom<CPL, RPW, NW, 8>::OX; oy = TileGeom<CPL, RPW, NW, 8>::OY; break;
    }
    out->ox = ox; out->oy = oy; out->threads = NW * kWave;
}

bool tile_shape(int kernel, int k, TileShape* out) {
    if (k < 1 || k > 8 || k == 7) return false;
    switch (kernel) {
        case TWS_KERNEL_TILE_128x128: shape_of<2, 8, 16>(k, out); return true;
        case TWS_KERNEL_TILE_128x64: shape_of<2, 8, 8>(k, out); return true;
        case TWS_KERNEL_TILE_256x64: shape_of<4, 4, 16>(k, out); return true;
        case TWS_KERNEL_TILE_64x64: shape_of<1, 8, 8>(k, out); return true;
        case TWS_KERNEL_TILE_64x64_16W: shape_of<1, 4, 16>(k, out); return true;
        case TWS_KERNEL_TILE_64x48_16W: shape_of<1, 3, 16>(k, out); return true;
        default: return false;
    }
}

static hipError_t launch_tile_any(int kernel, int k, int mode, const GridDesc* grids_dev,
                                  const int* hdr, const GridDesc* single, int blocks,
                                  int launch_idx, hipStream_t s) {
    if (blocks <= 0) return hipSuccess;
    switch (kernel) {
        case TWS_KERNEL_TILE_128x128: return launch_tile_k<2, 8, 16>(k, mode, grids_dev, hdr, single, blocks, launch_idx, s);
        case TWS_KERNEL_TILE_128x64: return launch_tile_k<2, 8, 8>(k, mode, grids_dev, hdr, single, blocks, launch_idx, s);
        case TWS_KERNEL_TILE_256x64: return launch_tile_k<4, 4, 16>(k, mode, grids_dev, hdr, single, blocks, launch_idx, s);
        case TWS_KERNEL_TILE_64x64: return launch_tile_k<1, 8, 8>(k, mode, grids_dev, hdr, single, blocks, launch_idx, s);
        case TWS_KERNEL_TILE_64x64_16W: return launch_tile_k<1, 4, 16>(k, mode, grids_dev, hdr, single, blocks, launch_idx, s);
        case TWS_KERNEL_TILE_64x48_16W: return launch_tile_k<1, 3, 16>(k, mode, grids_dev, hdr, single, blocks, launch_idx, s);
        default: return hipErrorInvalidValue;
    }
}

int kernel_mode(const GridDesc& g) {
    if (g.boundary != TWS_BOUNDARY_CLOSED) return kModeGeneral;
    return g.rain_on != 0 ? kModeRain : kModePlain;
}

hipError_t launch_tile_batch(int kernel, int k, int mode, const GridDesc* grids_dev,
                             const int* hdr_dev, int blocks, int launch_idx, hipStream_t s) {
    return launch_tile_any(kernel, k, mode, grids_dev, hdr_dev, nullptr, blocks, launch_idx, s);
}

hipError_t launch_tile_single(int kernel, int k, const GridDesc& g, int launch_idx,
                              hipStream_t s) {
    return launch_tile_any(kernel, k, kernel_mode(g), nullptr, nullptr, &g,
                           g.tiles_x * g.tiles_y, launch_idx, s);
}

hipError_t launch_twopass(const GridDesc* /*g_dev*/, const GridDesc& g, int parity,
                          hipStream_t s) {
    if (g.rain_on && g.brush_w > 0 && g.brush_h > 0)
        hipLaunchKernelGGL(k_brush, cell_grid(g), dim3(256), 0, s, g, parity, g.rain_cx, g.rain_cy,
                           g.rain_I, g.rain_R2);
    hipLaunchKernelGGL(k_flux_twopass, cell_grid(g), dim3(256), 0, s, g, parity);
    hipLaunchKernelGGL(k_apply_twopass, cell_grid(g), dim3(256), 0, s, g, parity);
    return hipGetLastError();
}

hipError_t launch_brush(const GridDesc& g, int parity, float cx, float cy, float I, float R2,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_brush, cell_grid(g), dim3(256), 0, s, g, parity, cx, cy, I, R2);
    return hipGetLastError();
}

hipError_t launch_velocity(const GridDesc& g, int parity, float* v2, hipStream_t s) {
    hipLaunchKernelGGL(k_velocity, cell_grid(g), dim3(256), 0, s, g, parity, v2);
    return hipGetLastError();
}

hipError_t launch_interleave(const GridDesc& g, int parity, float* staging4, hipStream_t s) {
    hipLaunchKernelGGL(k_interleave, cell_grid(g), dim3(256), 0, s, g, parity, staging4);
    return hipGetLastError();
}

hipError_t launch_deinterleave(const GridDesc& g, int parity, const float* staging4,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_deinterleave, cell_grid(g), dim3(256), 0, s, g, parity, staging4);
    return hipGetLastError();
}

hipError_t launch_ref_unpack(const GridDesc& g, int parity, const float* terrain_rgba,
                             const float* flow_rgba, float* b_out, hipStream_t s) {
    hipLaunchKernelGGL(k_ref_unpack, cell_grid(g), dim3(256), 0, s, g, parity, terrain_rgba,
                       flow_rgba, b_out);
    return hipGetLastError();
}

hipError_t launch_ref_pack(const GridDesc& g, int parity, float gb_fill, float* terrain_rgba,
                           float* flow_rgba, uint16_t* flowmap, hipStream_t s) {
    hipLaunchKernelGGL(k_ref_pack, cell_grid(g), dim3(256), 0, s, g, parity, gb_fill, terrain_rgba,
                       flow_rgba, flowmap);
    return hipGetLastError();
}

hipError_t launch_volume_partials(const GridDesc& g, int parity, double* partials, int nparts,
                                  hipStream_t s) {
    hipLaunchKernelGGL(k_volume, dim3(nparts), dim3(256), 0, s, g, parity, partials, nparts);
    return hipGetLastError();
}

hipError_t launch_zero_fill(float* p, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_zero_fill, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_fbm(float* b, const GridDesc& g, uint64_t seed, int octaves, int base,
                      double persistence, double scale, hipStream_t s) {
    hipLaunchKernelGGL(k_fbm, cell_grid(g), dim3(256), 0, s, b, g.pitch, g.W, g.H, seed, octaves,
                       base, persistence, scale);
    return hipGetLastError();
}

hipError_t launch_init_water(const GridDesc& g, int parity, float depth, int x0, int y0, int x1,
                             int y1, float rect_depth, hipStream_t s) {
    hipLaunchKernelGGL(k_init_water, cell_grid(g), dim3(256), 0, s, g, parity, depth, x0, y0, x1,
                       y1, rect_depth);
    return hipGetLastError();
}

hipError_t launch_check_finite(const GridDesc& g, int parity, hipStream_t s) {
    hipLaunchKernelGGL(k_check_finite, cell_grid(g), dim3(256), 0, s, g, parity);
    return hipGetLastError();
}

// Launch shapes of the copy: (loads in flight per lane, non-temporal,
// work groups of 256 per CU).
struct CopyShape {
    int unroll;
    bool nt;
    int groups_per_cu;
};
static const CopyShape kCopyShapes[] = {
    {1, false, 64},  // one load per lane and iteration
    {4, false, 8}, {4, true, 8}, {8, true, 4}, {8, true, 8}, {4, true, 16}, {8, false, 8},
};
int stream_copy_shapes() { return (int)(sizeof(kCopyShapes) / sizeof(kCopyShapes[0])); }

hipError_t launch_stream_copy(const void* src, void* dst, int64_t bytes, int shape,
                              hipStream_t s) {
    if (shape < 0 || shape >= stream_copy_shapes()) return hipErrorInvalidValue;
    const CopyShape& c = kCopyShapes[shape];
    const int64_t n = bytes / 16;
    if (n < 1) return hipSuccess;
    int64_t blocks = (n + 256LL * c.unroll - 1) / (256LL * c.unroll);
    blocks = std::min<int64_t>(blocks, 256LL * c.groups_per_cu);
    const vf4* sv = static_cast<const vf4*>(src);
    vf4* dv = static_cast<vf4*>(dst);
    const dim3 grid((unsigned)blocks), block(256);
    if (c.unroll == 1) hipLaunchKernelGGL((k_stream_copy<1, false>), grid, block, 0, s, sv, dv, n);
    else if (c.unroll == 4 && !c.nt) hipLaunchKernelGGL((k_stream_copy<4, false>), grid, block, 0, s, sv, dv, n);
    else if (c.unroll == 4) hipLaunchKernelGGL((k_stream_copy<4, true>), grid, block, 0, s, sv, dv, n);
    else if (!c.nt) hipLaunchKernelGGL((k_stream_copy<8, false>), grid, block, 0, s, sv, dv, n);
    else hipLaunchKernelGGL((k_stream_copy<8, true>), grid, block, 0, s, sv, dv, n);
    return hipGetLastError();
}

hipError_t launch_lane_shift_probe(int* out128, hipStream_t s) {
    hipLaunchKernelGGL(k_lane_shift_probe, dim3(1), dim3(64), 0, s, out128);
    return hipGetLastError();
}

}  // namespace tws
