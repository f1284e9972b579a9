// tws_kernel_util.hpp -- device helpers shared by the step kernels
// (tws_kernels.hip, tws_stream.hip): cross-lane moves, vector loads and
// stores of the state planes, the brush and the per-cell flux update.
#pragma once

#include <hip/hip_runtime.h>

#include "tws_internal.hpp"

namespace tws {

// ---------------------------------------------------------------------------
// Cross-lane moves.  DPP wave_shr:1 (lane i reads lane i-1) and wave_shl:1
// (lane i reads lane i+1).  The lane without a source (0 resp. 63, which sit
// in the tile halo) reads 0, or `fill` in the *_or variants: heads are filled
// with the lane's own head so the halo edge sees no gradient and does not
// send every wave down the flux-scaling division.
__device__ __forceinline__ float from_left_lane(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float from_right_lane(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xF, 0xF, true));
}
__device__ __forceinline__ float from_left_lane_or(float v, float fill) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float from_right_lane_or(float v, float fill) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v), 0x130, 0xF, 0xF, false));
}

// Turns at the SIMD.  Wavefronts of different work groups that share a SIMD
// are issued oldest first whenever several are ready, so of two groups with
// the same work started together the younger falls behind and runs alone,
// latencies exposed, after the older one ends.  take_turn() alternates the
// issue priority between them in epochs of 2^SHIFT ticks of the 100 MHz
// real-time counter: `parity` tells the groups of a SIMD apart (turn_parity:
// the wave slot, in units of the group's wavefronts per SIMD).  It pays in
// the streaming kernel, whose wavefronts never wait for each other; in the
// tile kernels, whose levels end in work-group barriers, it measured 1-7%
// slower.
template <int WAVES_PER_SIMD>
__device__ __forceinline__ unsigned turn_parity() {
    constexpr int kHwRegHwId = 4, kWaveIdBits = 4;  // HW_ID bits [3:0]: wave slot in the SIMD
    const unsigned slot =
        __builtin_amdgcn_s_getreg(kHwRegHwId | (0 << 6) | ((kWaveIdBits - 1) << 11));
    return (slot / WAVES_PER_SIMD) & 1u;
}
template <int SHIFT>
__device__ __forceinline__ void take_turn(unsigned parity) {
    const unsigned epoch = (unsigned)(__builtin_amdgcn_s_memrealtime() >> SHIFT);
    if ((epoch ^ parity) & 1u) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}

// Global-address-space views of the state planes: loads and stores through
// them are global_* instructions rather than flat_* ones.
typedef __attribute__((address_space(1))) float gfloat;
__device__ __forceinline__ const gfloat* gptr(const float* p) { return (const gfloat*)p; }
__device__ __forceinline__ gfloat* gptr(float* p) { return (gfloat*)p; }

typedef float vf2 __attribute__((ext_vector_type(2)));
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) vf2 gfloat2;
typedef __attribute__((address_space(1))) vf4 gfloat4;

template <int N>
__device__ __forceinline__ void vload(const gfloat* p, float (&o)[N]);
template <>
__device__ __forceinline__ void vload<1>(const gfloat* p, float (&o)[1]) { o[0] = p[0]; }
template <>
__device__ __forceinline__ void vload<2>(const gfloat* p, float (&o)[2]) {
    const vf2 v = *reinterpret_cast<const gfloat2*>(p);
    o[0] = v.x; o[1] = v.y;
}
template <>
__device__ __forceinline__ void vload<4>(const gfloat* p, float (&o)[4]) {
    const vf4 v = *reinterpret_cast<const gfloat4*>(p);
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}
// The tile kernels' stores of the new state are marked non-temporal: no tile
// of the same launch reads them back, so they need not displace the halos
// other tiles are about to re-read from L2.  MI355X, against default-policy
// stores: a batch of 32 x 512^2 grids 2.80 vs 2.93 ms per 100 steps; 128x64
// tiles on one 4096^2 grid (a zero-edge grid; closed ones take the streaming
// kernel) lose 2.5%.
template <int N>
__device__ __forceinline__ void vstore(gfloat* p, const float (&o)[N]);
template <>
__device__ __forceinline__ void vstore<1>(gfloat* p, const float (&o)[1]) { p[0] = o[0]; }
template <>
__device__ __forceinline__ void vstore<2>(gfloat* p, const float (&o)[2]) {
    __builtin_nontemporal_store(vf2{o[0], o[1]}, reinterpret_cast<gfloat2*>(p));
}
template <>
__device__ __forceinline__ void vstore<4>(gfloat* p, const float (&o)[4]) {
    __builtin_nontemporal_store(vf4{o[0], o[1], o[2], o[3]}, reinterpret_cast<gfloat4*>(p));
}

// Non-short-circuit finiteness test (no branches).
__device__ __forceinline__ bool nonfinite5(float a, float b, float c, float d, float e) {
    return !__builtin_isfinite(a) | !__builtin_isfinite(b) | !__builtin_isfinite(c) |
           !__builtin_isfinite(d) | !__builtin_isfinite(e);
}

// waterBrush.comp:26-28, one texel: saturate(1 - dot(c-p, c-p)/R^2) * I.
__device__ __forceinline__ float brush_amount(float cx, float cy, float I, float R2, int x, int y) {
    const float tx = cx - (float)x;
    const float ty = cy - (float)y;
    const float dist = tx * tx + ty * ty;
    const float q = dist / R2;
    float w = 1.0f - q;
    w = w < 0.0f ? 0.0f : (w > 1.0f ? 1.0f : w);
    return w * I;
}

// Whether the brush can reach any cell of columns [x_lo, x_hi] x rows
// [y_lo, y_hi].  brush_amount is +0 * I, which leaves every depth unchanged
// but for the sign of a zero, wherever the FP32 squared distance is at least
// R^2.  That distance is within 4 roundings (a relative 2.4e-7) of the exact
// one, so a block whose exact nearest squared distance is at least
// R^2 * (1 + 1e-6) gets +0 * I at every cell.
__device__ __forceinline__ bool brush_reaches(float cx, float cy, float R2, int x_lo, int x_hi,
                                              int y_lo, int y_hi) {
    const double dx = cx < (float)x_lo ? (double)x_lo - cx : (cx > (float)x_hi ? (double)cx - x_hi : 0.0);
    const double dy = cy < (float)y_lo ? (double)y_lo - cy : (cy > (float)y_hi ? (double)cy - y_hi : 0.0);
    return dx * dx + dy * dy < (double)R2 * (1.0 + 1e-6);
}

// flowUpdate.comp:53-59 for one cell: new outgoing flux from the products
// p_i = g_i * K of the four head differences g_i (own head minus the
// neighbour's, +x, -x, +y, -y), clamped at zero and scaled so the cell cannot
// drain below 0.
__device__ __forceinline__ void flux_update_p(float F, float C, float d, float p0, float p1,
                                              float p2, float p3, float& f0, float& f1,
                                              float& f2, float& f3) {
    float o0 = fmaxf(0.0f, f0 * F + p0);
    float o1 = fmaxf(0.0f, f1 * F + p1);
    float o2 = fmaxf(0.0f, f2 * F + p2);
    float o3 = fmaxf(0.0f, f3 * F + p3);
    // flowUpdate.comp:57 summed pairwise, the order Mesa executes the reference in
    const float total = ((o0 + o1) + (o2 + o3)) * C;
    if (total > d) {
        // flowUpdate.comp:58-59.  A dry cell on a slope (d = +0) gets s = +0;
        // an uploaded d = -0 gets s = -0 and signs the scaled fluxes, as the
        // division in the shader does.
        const float s = d / total;
        o0 = o0 * s; o1 = o1 * s; o2 = o2 * s; o3 = o3 * s;
    }
    f0 = o0; f1 = o1; f2 = o2; f3 = o3;
}

// The same from the head differences g_i themselves.
__device__ __forceinline__ void flux_update(float F, float K, float C, float d, float g0,
                                            float g1, float g2, float g3, float& f0, float& f1,
                                            float& f2, float& f3) {
    flux_update_p(F, C, d, g0 * K, g1 * K, g2 * K, g3 * K, f0, f1, f2, f3);
}

}  // namespace tws
