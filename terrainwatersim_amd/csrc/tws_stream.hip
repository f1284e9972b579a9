// tws_stream.hip -- the streaming, temporally blocked step kernel (k_stream).
//
// Each wavefront owns one strip segment of a grid: EXW = 128 columns (2 per
// lane) by a run of output rows, and sweeps it top to bottom once per
// launch, advancing every row KH steps on the way.  Level j (the state after
// j steps) trails level j-1 by two rows: at row-step t, level j computes the
// flux of row a = t - 2(j-1) (flowUpdate.comp:34-62) and the depth of row
// a-1 (flowApply.comp:32-41), which needs level j-1 only at rows a-1..a+1,
// all of them finished by then.  So each level holds a window of three rows
// of depth/head and two of flux in VGPRs, and
//   - nothing is recomputed in y except the 2 KH rows a segment's cone
//     reaches into its neighbours (the tile kernels recompute 2 KH rows at
//     both ends of every 64-row tile);
//   - no work-group barrier: the wavefronts of a work group are independent,
//     x-neighbours come from DPP lane shifts, y-neighbours from the lane's
//     own window;
//   - terrain b, read by each level at its own row, sits in a wave-private
//     ring of 16 rows in LDS (and the rain amounts beside it);
//   - the next input row is loaded one row-step ahead.
// The x-halo is the tile kernels': HALOX = 2 KH columns on each side go stale
// and are recomputed by the neighbouring strip.  Out-of-domain cells are the
// closed-wall ghosts of the tile kernels (b = 1e30, d = f = 0), so the
// arithmetic per cell is theirs, in the same order: results are bit-identical
// to the FP32 restatement for any segment length and steps per launch.

#include <algorithm>

#include "tws.h"
#include "tws_internal.hpp"
#include "tws_kernel_util.hpp"

namespace tws {

namespace {

constexpr int kSCPL = 2;                 // columns per lane
constexpr int kSEXW = kSCPL * kWave;     // strip width incl. halo
constexpr int kRing = 16;                // rows of b (and rain) kept per wave
constexpr int kStreamWaves = 4;          // independent wavefronts per work group

template <int KH>
struct StreamGeom {
    // Wavefronts per SIMD: a level's window takes about 36 VGPRs, so up to 5
    // levels fit 256 VGPRs (2 per SIMD); 6 and 8 levels take 1 per SIMD
    // with up to 512 VGPRs instead of spilling.
    static constexpr int WAVES_PER_SIMD = KH <= 5 ? 2 : 1;
    static constexpr int HALOX = ((2 * KH + kSCPL - 1) / kSCPL) * kSCPL;
    static constexpr int OX = kSEXW - 2 * HALOX;
    static_assert(2 * KH - 1 <= kRing, "ring too short for KH");
    static_assert(OX > 0, "halo wider than the strip");
};

// One level's window: flux of rows a-1 (F1) and a-2 (F2), depth and head of
// rows a-2 (D1, H1) and a-3 (D2, H2), a being the level's flux row at the
// current row-step.  Depth and head already include the next level's rain.
struct Level {
    float F1[4][kSCPL], F2[4][kSCPL];
    float D1[kSCPL], D2[kSCPL], H1[kSCPL], H2[kSCPL];
};

// The six planes of one row for this lane's columns.
struct Row {
    float b[kSCPL], d[kSCPL], f[4][kSCPL];
};

}  // namespace

// Loads row y of set `parity` for columns x0..x0+kSCPL-1 (ghost values
// outside the domain).  Closed walls: an edge cell's flux toward the wall is
// dropped (the walls carry nothing, whatever was uploaded).
__device__ __forceinline__ void stream_load(const GridDesc& g, int parity, int y, int x0, Row& r) {
    const bool vin = (unsigned)y < (unsigned)g.H && x0 >= 0 && x0 < g.W;
    float vb[kSCPL], vd[kSCPL], v0[kSCPL], v1[kSCPL], v2[kSCPL], v3[kSCPL];
#pragma unroll
    for (int c = 0; c < kSCPL; ++c) vb[c] = vd[c] = v0[c] = v1[c] = v2[c] = v3[c] = 0.0f;
    if (vin) {
        const int64_t off = (int64_t)y * g.pitch + x0;
        vload<kSCPL>(gptr(g.b) + off, vb);
        vload<kSCPL>(gptr(g.d[parity]) + off, vd);
        vload<kSCPL>(gptr(g.f[parity][0]) + off, v0);
        vload<kSCPL>(gptr(g.f[parity][1]) + off, v1);
        vload<kSCPL>(gptr(g.f[parity][2]) + off, v2);
        vload<kSCPL>(gptr(g.f[parity][3]) + off, v3);
    }
#pragma unroll
    for (int c = 0; c < kSCPL; ++c) {
        const bool in = vin && (x0 + c < g.W);
        r.b[c] = in ? vb[c] : g.ghost_b;
        r.d[c] = in ? vd[c] : g.ghost_d;
        r.f[0][c] = in && x0 + c != g.W - 1 ? v0[c] : 0.0f;
        r.f[1][c] = in && x0 + c != 0 ? v1[c] : 0.0f;
        r.f[2][c] = in && y != g.H - 1 ? v2[c] : 0.0f;
        r.f[3][c] = in && y != 0 ? v3[c] : 0.0f;
    }
}

// k_stream: grid of ceil(tasks / kStreamWaves) work groups; task = segment *
// strips + strip, segments of seg_rows output rows.  Closed walls only
// (MODE kModePlain or kModeRain).
template <int KH, int MODE>
__global__ void __launch_bounds__(kStreamWaves * kWave, StreamGeom<KH>::WAVES_PER_SIMD)
k_stream(GridDesc g, int strips, int segs, int seg_rows, int parity) {
    constexpr bool RAIN = MODE == kModeRain;
    using G = StreamGeom<KH>;
    __shared__ float s_b[kStreamWaves][kRing][kSEXW];
    __shared__ float s_wet[RAIN ? kStreamWaves : 1][RAIN ? kRing : 1][kSEXW];

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // XCD-aware order: work groups are dealt to the 8 XCDs round robin; each
    // XCD gets a contiguous run of tasks, so neighbouring strips (which
    // re-read each other's halo columns) meet in one L2.
    const int nb = (int)gridDim.x, bid = blockIdx.x;
    const int xcd = bid & 7, idx = bid >> 3;
    const int q = nb >> 3, rmd = nb & 7;
    const int blk = xcd * q + (xcd < rmd ? xcd : rmd) + idx;
    const int task = blk * kStreamWaves + wave;
    if (task >= strips * segs) return;  // whole wavefront: no barrier follows
    const int seg = task / strips, strip = task - seg * strips;
    const int Y0 = seg * seg_rows;
    const int Y1 = min(g.H, Y0 + seg_rows);
    if (Y0 >= Y1) return;
    const int ox = strip * G::OX - G::HALOX;
    const int x0 = ox + lane * kSCPL;
    const int lx = lane * kSCPL;
    const float F = g.F, K = g.K, C = g.C;
    const int wr = RAIN ? wave : 0;

    // stored columns of this lane
    const bool col_ok = lx >= G::HALOX && lx < G::HALOX + G::OX && x0 < g.W;
    const bool full = x0 + kSCPL <= g.W;
    gfloat* const pd = gptr(g.d[parity ^ 1]);
    gfloat* const p0 = gptr(g.f[parity ^ 1][0]);
    gfloat* const p1 = gptr(g.f[parity ^ 1][1]);
    gfloat* const p2 = gptr(g.f[parity ^ 1][2]);
    gfloat* const p3 = gptr(g.f[parity ^ 1][3]);
    bool bad = false;

    // rows of the ring that the rain reaches (bit r & 15), wave-uniform
    unsigned rain_rows = 0;
    const bool rain_on = RAIN && g.rain_on != 0;

    // Accepts input row y: writes b (and the rain amount) into the ring,
    // returns d with the first level's rain added and its head.
    auto accept = [&](int y, const Row& r, float (&dn)[kSCPL], float (&hn)[kSCPL]) {
        const int slot = y & (kRing - 1);
#pragma unroll
        for (int c = 0; c < kSCPL; ++c) s_b[wave][slot][lx + c] = r.b[c];
        bool wet_row = false;
        if (rain_on) {
            wet_row = __builtin_amdgcn_readfirstlane(
                          brush_reaches(g.rain_cx, g.rain_cy, g.rain_R2, ox, ox + kSEXW - 1, y, y)
                              ? 1 : 0) != 0 &&
                      (unsigned)y < (unsigned)g.H;
        }
        rain_rows = (rain_rows & ~(1u << slot)) | ((wet_row ? 1u : 0u) << slot);
#pragma unroll
        for (int c = 0; c < kSCPL; ++c) dn[c] = r.d[c];
        if (wet_row) {
#pragma unroll
            for (int c = 0; c < kSCPL; ++c) {
                const bool in = x0 + c >= 0 && x0 + c < g.W;
                const float w = in ? brush_amount(g.rain_cx, g.rain_cy, g.rain_I, g.rain_R2,
                                                  x0 + c, y)
                                   : 0.0f;
                s_wet[wr][slot][lx + c] = w;
                if (in) dn[c] = dn[c] + w;
            }
        }
#pragma unroll
        for (int c = 0; c < kSCPL; ++c) hn[c] = dn[c] + r.b[c];
    };

    // ---- state at the first row-step -------------------------------------
    const int t_begin = Y0 - 2 * KH + 1;
    const int t_end = Y1 + 2 * KH - 2;
#pragma unroll
    for (int s = 0; s < kRing; ++s)
#pragma unroll
        for (int c = 0; c < kSCPL; ++c) s_b[wave][s][lx + c] = 0.0f;
    Level L[KH + 1];
#pragma unroll
    for (int j = 0; j <= KH; ++j)
#pragma unroll
        for (int c = 0; c < kSCPL; ++c) {
#pragma unroll
            for (int ch = 0; ch < 4; ++ch) L[j].F1[ch][c] = L[j].F2[ch][c] = 0.0f;
            L[j].D1[c] = L[j].D2[c] = L[j].H1[c] = L[j].H2[c] = 0.0f;
        }
    float Fown0[4][kSCPL];  // input flux of row t (the first level's own flux)
    {
        Row r;
        stream_load(g, parity, t_begin - 1, x0, r);
        accept(t_begin - 1, r, L[0].D2, L[0].H2);
        stream_load(g, parity, t_begin, x0, r);
        accept(t_begin, r, L[0].D1, L[0].H1);
#pragma unroll
        for (int ch = 0; ch < 4; ++ch)
#pragma unroll
            for (int c = 0; c < kSCPL; ++c) Fown0[ch][c] = r.f[ch][c];
    }
    Row next;
    stream_load(g, parity, t_begin + 1, x0, next);

    // ---- one row-step ----------------------------------------------------
    auto step = [&](int t) {
        // level 0: input row t+1, loaded one row-step ago; load row t+2
        float dN0[kSCPL], hN0[kSCPL], fN0[4][kSCPL];
        accept(t + 1, next, dN0, hN0);
#pragma unroll
        for (int ch = 0; ch < 4; ++ch)
#pragma unroll
            for (int c = 0; c < kSCPL; ++c) fN0[ch][c] = next.f[ch][c];
        stream_load(g, parity, t + 2, x0, next);

        // inputs of level 1: level 0 at rows t-1, t, t+1
        float hm1[kSCPL], h0[kSCPL], hp1[kSCPL], dOwn[kSCPL], dBase[kSCPL], fOwn[4][kSCPL];
#pragma unroll
        for (int c = 0; c < kSCPL; ++c) {
            hm1[c] = L[0].H2[c];
            h0[c] = L[0].H1[c];
            hp1[c] = hN0[c];
            dOwn[c] = L[0].D1[c];
            dBase[c] = L[0].D2[c];
#pragma unroll
            for (int ch = 0; ch < 4; ++ch) fOwn[ch][c] = Fown0[ch][c];
        }
#pragma unroll
        for (int c = 0; c < kSCPL; ++c) {
            L[0].D2[c] = L[0].D1[c];
            L[0].D1[c] = dN0[c];
            L[0].H2[c] = L[0].H1[c];
            L[0].H1[c] = hN0[c];
#pragma unroll
            for (int ch = 0; ch < 4; ++ch) Fown0[ch][c] = fN0[ch][c];
        }

#pragma unroll
        for (int j = 1; j <= KH; ++j) {
            const int a = t - 2 * (j - 1);  // flux row of level j
            Level& Lj = L[j];
            // flux of row a: flowUpdate.comp:34-62
            float FN[4][kSCPL];
            {
                const float h_left = from_left_lane_or(h0[kSCPL - 1], h0[0]);
                const float h_right = from_right_lane_or(h0[0], h0[kSCPL - 1]);
#pragma unroll
                for (int c = 0; c < kSCPL; ++c) {
                    const float own = h0[c];
                    const float hx1 = c < kSCPL - 1 ? h0[c + 1] : h_right;
                    const float hx0 = c > 0 ? h0[c - 1] : h_left;
#pragma unroll
                    for (int ch = 0; ch < 4; ++ch) FN[ch][c] = fOwn[ch][c];
                    flux_update(F, K, C, dOwn[c], own - hx1, own - hx0, own - hp1[c],
                                own - hm1[c], FN[0][c], FN[1][c], FN[2][c], FN[3][c]);
                }
            }
            // depth of row a-1: flowApply.comp:32-41
            float dN[kSCPL], hN[kSCPL];
            {
                const float in_left = from_left_lane(Lj.F1[0][kSCPL - 1]);   // X0.x
                const float in_right = from_right_lane(Lj.F1[1][0]);        // X1.y
#pragma unroll
                for (int c = 0; c < kSCPL; ++c) {
                    const float X1y = c < kSCPL - 1 ? Lj.F1[1][c + 1] : in_right;
                    const float X0x = c > 0 ? Lj.F1[0][c - 1] : in_left;
                    const float Y1w = FN[3][c];
                    const float Y0z = Lj.F2[2][c];
                    const float in = ((X1y + X0x) + Y1w) + Y0z;
                    const float out = ((Lj.F1[0][c] + Lj.F1[1][c]) + Lj.F1[2][c]) + Lj.F1[3][c];
                    dN[c] = fmaxf(0.0f, dBase[c] + (in - out) * C);
                }
            }
            const int yd = a - 1;  // row of the new depth
            if (j < KH) {
                // the next level's rain, then its head
                const int slot = yd & (kRing - 1);
                if (RAIN && ((rain_rows >> slot) & 1u)) {
#pragma unroll
                    for (int c = 0; c < kSCPL; ++c) {
                        const bool in = x0 + c >= 0 && x0 + c < g.W;
                        if (in) dN[c] = dN[c] + s_wet[wr][slot][lx + c];
                    }
                }
#pragma unroll
                for (int c = 0; c < kSCPL; ++c) hN[c] = dN[c] + s_b[wave][slot][lx + c];
                // inputs of level j+1 (level j at rows a-3, a-2, a-1), then
                // the window moves down one row
#pragma unroll
                for (int c = 0; c < kSCPL; ++c) {
                    hm1[c] = Lj.H2[c];
                    h0[c] = Lj.H1[c];
                    hp1[c] = hN[c];
                    dOwn[c] = Lj.D1[c];
                    dBase[c] = Lj.D2[c];
#pragma unroll
                    for (int ch = 0; ch < 4; ++ch) fOwn[ch][c] = Lj.F2[ch][c];
                }
#pragma unroll
                for (int c = 0; c < kSCPL; ++c) {
                    Lj.D2[c] = Lj.D1[c];
                    Lj.D1[c] = dN[c];
                    Lj.H2[c] = Lj.H1[c];
                    Lj.H1[c] = hN[c];
                }
            } else if (yd >= Y0 && yd < Y1 && col_ok) {
                // last level: store row a-1 (its flux is F1, final since the
                // previous row-step)
                const int64_t off = (int64_t)yd * g.pitch + x0;
#pragma unroll
                for (int c = 0; c < kSCPL; ++c)
                    bad |= (full || x0 + c < g.W) &
                           nonfinite5(dN[c], Lj.F1[0][c], Lj.F1[1][c], Lj.F1[2][c], Lj.F1[3][c]);
                if (full) {
                    vstore<kSCPL>(pd + off, dN);
                    vstore<kSCPL>(p0 + off, Lj.F1[0]);
                    vstore<kSCPL>(p1 + off, Lj.F1[1]);
                    vstore<kSCPL>(p2 + off, Lj.F1[2]);
                    vstore<kSCPL>(p3 + off, Lj.F1[3]);
                } else {
#pragma unroll
                    for (int c = 0; c < kSCPL; ++c)
                        if (x0 + c < g.W) {
                            pd[off + c] = dN[c];
                            p0[off + c] = Lj.F1[0][c];
                            p1[off + c] = Lj.F1[1][c];
                            p2[off + c] = Lj.F1[2][c];
                            p3[off + c] = Lj.F1[3][c];
                        }
                }
            }
#pragma unroll
            for (int ch = 0; ch < 4; ++ch)
#pragma unroll
                for (int c = 0; c < kSCPL; ++c) {
                    Lj.F2[ch][c] = Lj.F1[ch][c];
                    Lj.F1[ch][c] = FN[ch][c];
                }
        }
    };

    // Six row-steps per iteration: every window rotates with period 2 or 3,
    // so the unrolled body renames registers instead of moving them.
    for (int t = t_begin; t <= t_end; t += 6) {
        step(t);
        if (t + 1 > t_end) break;
        step(t + 1);
        if (t + 2 > t_end) break;
        step(t + 2);
        if (t + 3 > t_end) break;
        step(t + 3);
        if (t + 4 > t_end) break;
        step(t + 4);
        if (t + 5 > t_end) break;
        step(t + 5);
    }
    if (__ballot(bad) != 0 && lane == 0) atomicOr(g.status, 1u);
}

// ---------------------------------------------------------------------------

template <int KH>
static void geom_of(int* ox, int* waves_per_simd) {
    *ox = StreamGeom<KH>::OX;
    *waves_per_simd = StreamGeom<KH>::WAVES_PER_SIMD;
}

static bool stream_geom(int k, int* ox, int* waves_per_simd) {
    switch (k) {
        case 1: geom_of<1>(ox, waves_per_simd); return true;
        case 2: geom_of<2>(ox, waves_per_simd); return true;
        case 3: geom_of<3>(ox, waves_per_simd); return true;
        case 4: geom_of<4>(ox, waves_per_simd); return true;
        case 5: geom_of<5>(ox, waves_per_simd); return true;
        case 6: geom_of<6>(ox, waves_per_simd); return true;
        case 8: geom_of<8>(ox, waves_per_simd); return true;
        default: return false;
    }
}

bool stream_shape(int k, int* ox) {
    int w = 0;
    return stream_geom(k, ox, &w);
}

StreamPlan stream_plan(int W, int H, int k, int cus) {
    StreamPlan p{};
    int ox = 0, per_simd = 0;
    if (!stream_geom(k, &ox, &per_simd) || W <= 0 || H <= 0 || cus <= 0) return p;
    const int resident_waves = cus * 4 * per_simd;
    p.strips = (W + ox - 1) / ox;
    // one wavefront per segment, as many segments as fill the resident
    // wavefront slots once, but no segment shorter than 4 k rows (its cone
    // reaches 2 k rows into each neighbour)
    int segs = std::max(1, resident_waves / p.strips);
    segs = std::min(segs, std::max(1, H / (4 * k)));
    p.seg_rows = (H + segs - 1) / segs;
    p.segs = (H + p.seg_rows - 1) / p.seg_rows;
    p.blocks = (p.strips * p.segs + kStreamWaves - 1) / kStreamWaves;
    return p;
}

template <int MODE>
static hipError_t launch_stream_m(int k, const GridDesc& g, const StreamPlan& p, int parity,
                                  hipStream_t s) {
    const dim3 grid(p.blocks), block(kStreamWaves * kWave);
    switch (k) {
        case 1: hipLaunchKernelGGL((k_stream<1, MODE>), grid, block, 0, s, g, p.strips, p.segs, p.seg_rows, parity); break;
        case 2: hipLaunchKernelGGL((k_stream<2, MODE>), grid, block, 0, s, g, p.strips, p.segs, p.seg_rows, parity); break;
        case 3: hipLaunchKernelGGL((k_stream<3, MODE>), grid, block, 0, s, g, p.strips, p.segs, p.seg_rows, parity); break;
        case 4: hipLaunchKernelGGL((k_stream<4, MODE>), grid, block, 0, s, g, p.strips, p.segs, p.seg_rows, parity); break;
        case 5: hipLaunchKernelGGL((k_stream<5, MODE>), grid, block, 0, s, g, p.strips, p.segs, p.seg_rows, parity); break;
        case 6: hipLaunchKernelGGL((k_stream<6, MODE>), grid, block, 0, s, g, p.strips, p.segs, p.seg_rows, parity); break;
        case 8: hipLaunchKernelGGL((k_stream<8, MODE>), grid, block, 0, s, g, p.strips, p.segs, p.seg_rows, parity); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_stream_single(int k, const GridDesc& g, const StreamPlan& p, int parity,
                                hipStream_t s) {
    if (p.blocks <= 0) return hipErrorInvalidValue;
    if (g.boundary != TWS_BOUNDARY_CLOSED) return hipErrorInvalidValue;
    return g.rain_on ? launch_stream_m<kModeRain>(k, g, p, parity, s)
                     : launch_stream_m<kModePlain>(k, g, p, parity, s);
}

}  // namespace tws
