// tws_mips.hip -- the terrain mip chain the reference's renderer samples.
//
// After every frame with at least one step, Terrain::PerformSimulationStep
// regenerates the full mip chain of TerrainData (Terrain.cpp:270-276 ->
// Texture2D::GenMipMaps, Texture2D.cpp:64-68, glGenerateMipmap).  The chain
// has floor(log2(max(W,H))) + 1 levels (Texture.cpp:28-44); level l is
// max(1, W>>l) x max(1, H>>l) RGBA32F texels (r = terrain, g = b = the
// constant fill, a = water depth).
//
// Reduction: the bilinear blit Mesa's llvmpipe uses for glGenerateMipmap.
// Each texel of level l+1 samples level l at its own centre,
// u = (i + 0.5) * ws / wd - 0.5, with lerp(a, b, t) = a + t * (b - a), along
// x first, then y, clamped to the edge.  Halving (ws = 2 wd) is t = 0.5
// between texels 2i and 2i+1 and matches llvmpipe bit for bit; other ratios
// use the exact sample position (double, t rounded once to float).
//
// Level 1 reads the b and d planes of the current state (8 B per level-0
// cell) and writes 16 B per level-1 texel; level l > 1 reads level l-1.  A
// wave covers 64 consecutive texels of one row, so the source rows are read
// as contiguous 512-byte runs.

#include <hip/hip_runtime.h>

#include <cmath>

#include "tws.h"
#include "tws_grid.hpp"
#include "tws_runtime.hpp"

namespace tws {

namespace {

__host__ __device__ inline void mip_axis(int i, int ws, int wd, int& i0, int& i1, float& t) {
    if (ws == 2 * wd) {
        i0 = 2 * i;
        i1 = 2 * i + 1;
        t = 0.5f;
        return;
    }
    const double u = ((double)i + 0.5) * (double)ws / (double)wd - 0.5;
    double fl = floor(u);
    if (fl < 0.0) fl = 0.0;
    int a = (int)fl;
    if (a > ws - 1) a = ws - 1;
    i0 = a;
    i1 = a + 1 < ws ? a + 1 : ws - 1;
    t = (float)(u - (double)a);
    if (t < 0.0f) t = 0.0f;
}

__device__ __forceinline__ float mip_lerp(float a, float b, float t) { return a + t * (b - a); }

__device__ __forceinline__ float bilerp(float s00, float s01, float s10, float s11, float tx,
                                        float ty) {
    return mip_lerp(mip_lerp(s00, s01, tx), mip_lerp(s10, s11, tx), ty);
}

constexpr int kMipBX = 64, kMipBY = 4;

// Level 1 from the state planes b (-> r) and d of set `parity` (-> a).
__global__ void __launch_bounds__(kMipBX * kMipBY)
k_mip_first(GridDesc g, int parity, float gb, float4* __restrict__ dst, int wd, int hd) {
    const int i = blockIdx.x * kMipBX + (threadIdx.x & (kMipBX - 1));
    const int j = blockIdx.y * kMipBY + (threadIdx.x / kMipBX);
    if (i >= wd || j >= hd) return;
    int x0, x1, y0, y1;
    float tx, ty;
    mip_axis(i, g.W, wd, x0, x1, tx);
    mip_axis(j, g.H, hd, y0, y1, ty);
    const int64_t r0 = (int64_t)y0 * g.pitch, r1 = (int64_t)y1 * g.pitch;
    const float* b = g.b;
    const float* d = g.d[parity];
    const float r = bilerp(b[r0 + x0], b[r0 + x1], b[r1 + x0], b[r1 + x1], tx, ty);
    const float a = bilerp(d[r0 + x0], d[r0 + x1], d[r1 + x0], d[r1 + x1], tx, ty);
    const float c = bilerp(gb, gb, gb, gb, tx, ty);
    dst[(int64_t)j * wd + i] = make_float4(r, c, c, a);
}

// Level l+1 from level l (both dense RGBA).
__global__ void __launch_bounds__(kMipBX * kMipBY)
k_mip_next(const float4* __restrict__ src, int ws, int hs, float4* __restrict__ dst, int wd,
           int hd) {
    const int i = blockIdx.x * kMipBX + (threadIdx.x & (kMipBX - 1));
    const int j = blockIdx.y * kMipBY + (threadIdx.x / kMipBX);
    if (i >= wd || j >= hd) return;
    int x0, x1, y0, y1;
    float tx, ty;
    mip_axis(i, ws, wd, x0, x1, tx);
    mip_axis(j, hs, hd, y0, y1, ty);
    const float4 s00 = src[(int64_t)y0 * ws + x0], s01 = src[(int64_t)y0 * ws + x1];
    const float4 s10 = src[(int64_t)y1 * ws + x0], s11 = src[(int64_t)y1 * ws + x1];
    dst[(int64_t)j * wd + i] = make_float4(bilerp(s00.x, s01.x, s10.x, s11.x, tx, ty),
                                           bilerp(s00.y, s01.y, s10.y, s11.y, tx, ty),
                                           bilerp(s00.z, s01.z, s10.z, s11.z, tx, ty),
                                           bilerp(s00.w, s01.w, s10.w, s11.w, tx, ty));
}

// The small end of the chain in one work group: every level from one of at
// most kTailTexels texels down to 1x1, level after level, each read back
// from the chain buffer that this work group wrote before the barrier.  The
// per-texel arithmetic is k_mip_next's; the levels cost one launch instead of
// one each (they are bound by launch latency, not by bytes).
constexpr int kTailThreads = 1024;
constexpr int kTailTexels = 64 * 64;

__global__ void __launch_bounds__(kTailThreads)
k_mip_tail(const float4* src, int ws, int hs, float4* dst) {
    while (ws > 1 || hs > 1) {
        const int wd = ws / 2 > 0 ? ws / 2 : 1, hd = hs / 2 > 0 ? hs / 2 : 1;
        for (int k = threadIdx.x; k < wd * hd; k += kTailThreads) {
            const int j = k / wd, i = k - j * wd;
            int x0, x1, y0, y1;
            float tx, ty;
            mip_axis(i, ws, wd, x0, x1, tx);
            mip_axis(j, hs, hd, y0, y1, ty);
            const float4 s00 = src[y0 * ws + x0], s01 = src[y0 * ws + x1];
            const float4 s10 = src[y1 * ws + x0], s11 = src[y1 * ws + x1];
            dst[k] = make_float4(bilerp(s00.x, s01.x, s10.x, s11.x, tx, ty),
                                 bilerp(s00.y, s01.y, s10.y, s11.y, tx, ty),
                                 bilerp(s00.z, s01.z, s10.z, s11.z, tx, ty),
                                 bilerp(s00.w, s01.w, s10.w, s11.w, tx, ty));
        }
        __syncthreads();  // this level is complete and visible to the whole group
        src = dst;
        dst += wd * hd;
        ws = wd;
        hs = hd;
    }
}

dim3 mip_grid(int wd, int hd) { return dim3((wd + kMipBX - 1) / kMipBX, (hd + kMipBY - 1) / kMipBY); }

}  // namespace

int mip_chain(int W, int H, int* levels, size_t* floats) {
    int n = 0;
    for (int w = W, h = H; w > 0 || h > 0; w /= 2, h /= 2) ++n;
    size_t total = 0;
    for (int l = 1; l < n; ++l) {
        const size_t w = (size_t)(W >> l > 0 ? W >> l : 1), h = (size_t)(H >> l > 0 ? H >> l : 1);
        total += 4 * w * h;
    }
    if (levels) *levels = n;
    if (floats) *floats = total;
    return TWS_OK;
}

int grid_mips(GridState& g, float gb_fill, float* dst, size_t capacity_floats, hipStream_t s) {
    int levels = 0;
    size_t need = 0;
    mip_chain(g.W, g.H, &levels, &need);
    if (capacity_floats < need)
        return fail(TWS_ERR_INVALID, "mips: buffer holds %zu floats, the chain needs %zu",
                    capacity_floats, need);
    if (!std::isfinite(gb_fill)) return fail(TWS_ERR_INVALID, "mips: non-finite fill value");
    if (need > 0 && (!dst || reinterpret_cast<uintptr_t>(dst) % 16 != 0))
        return fail(TWS_ERR_INVALID, "mips: output must be a 16-byte aligned device buffer");
    const GridDesc gd = g.desc();
    float4* out = reinterpret_cast<float4*>(dst);
    const float4* prev = nullptr;
    int ws = g.W, hs = g.H;
    for (int l = 1; l < levels; ++l) {
        const int wd = g.W >> l > 0 ? g.W >> l : 1, hd = g.H >> l > 0 ? g.H >> l : 1;
        if (l > 1 && (int64_t)wd * hd <= kTailTexels) {
            // levels l .. levels-1 in one launch (their sizes follow the same
            // halving rule, so the tail writes them where the loop would)
            hipLaunchKernelGGL(k_mip_tail, dim3(1), dim3(kTailThreads), 0, s, prev, ws, hs, out);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return hip_fail(e, "mips: launch");
            return TWS_OK;
        }
        if (l == 1)
            hipLaunchKernelGGL(k_mip_first, mip_grid(wd, hd), dim3(kMipBX * kMipBY), 0, s, gd,
                               g.parity, gb_fill, out, wd, hd);
        else
            hipLaunchKernelGGL(k_mip_next, mip_grid(wd, hd), dim3(kMipBX * kMipBY), 0, s, prev,
                               ws, hs, out, wd, hd);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "mips: launch");
        prev = out;
        out += (size_t)wd * hd;
        ws = wd;
        hs = hd;
    }
    return TWS_OK;
}

}  // namespace tws
