// tws_stream.hip -- the streaming, temporally blocked step kernel (k_stream).
//
// Each wavefront owns one strip segment of a grid: EXW = 128 columns (2 per
// lane) by a run of output rows, and sweeps it top to bottom once per
// launch, advancing every row KH steps on the way.  Level j (the state after
// j steps) trails level j-1 by two rows: at row-step t, level j computes the
// flux of row a = t - 2(j-1) (flowUpdate.comp:34-62) and the depth of row
// a-1 (flowApply.comp:32-41), which needs level j-1 only at rows a-1..a+1,
// all of them finished by then.  So each level holds a window of three rows
// of depth and three of flux in VGPRs (heads are recomputed from the terrain
// as they are read), and
//   - nothing is recomputed in y except the 2 KH rows a segment's cone
//     reaches into its neighbours (the tile kernels recompute 2 KH rows at
//     both ends of every 64-row tile);
//   - no work-group barrier: the wavefronts of a work group are independent,
//     x-neighbours come from DPP lane shifts, y-neighbours from the lane's
//     own window;
//   - terrain b, read by each level at its own rows, sits in a wave-private
//     ring of 16 (32 for 8 levels) rows in LDS, the rain amounts beside it;
//     a level's rows are read while the level before it computes;
//   - each head difference serves the two cells it separates (negated for
//     the second: fl(a - b) = -fl(b - a));
//   - the input rows are loaded two row-steps ahead, and the new state is
//     stored with the default cache policy (stores share the memory counter
//     that the row-steps' waits on those loads read);
//   - the two wavefronts sharing a SIMD take turns at the higher issue
//     priority, so neither runs alone at the end of the launch.
// The x-halo is the tile kernels': HALOX = 2 KH columns on each side go stale
// and are recomputed by the neighbouring strip.  Out-of-domain cells are the
// closed-wall ghosts of the tile kernels (b = 1e30, d = f = 0), so the
// arithmetic per cell is theirs, in the same order: results are bit-identical
// to the FP32 restatement for any segment length and steps per launch.

#include <algorithm>
#include <cstdint>
#include <functional>
#include <vector>
#include <type_traits>

#include "tws.h"
#include "tws_internal.hpp"
#include "tws_kernel_util.hpp"

namespace tws {

namespace {

constexpr int kSCPL = 2;                 // columns per lane
constexpr int kSEXW = kSCPL * kWave;     // strip width incl. halo
constexpr int kStreamWaves = 4;          // independent wavefronts per work group

template <int KH>
struct StreamGeom {
    // Wavefronts per SIMD: a level's window takes about 34 VGPRs, so up to 3
    // levels fit 168 VGPRs (3 per SIMD), 6 levels 256 (2 per SIMD); 8 levels
    // take 1 per SIMD, and past the 256 architectural VGPRs of a wavefront
    // the compiler keeps values in accumulation registers.
    static constexpr int WAVES_PER_SIMD = KH <= 3 ? 3 : (KH <= 6 ? 2 : 1);
    static constexpr int HALOX = ((2 * KH + kSCPL - 1) / kSCPL) * kSCPL;
    static constexpr int OX = kSEXW - 2 * HALOX;
    // rows of b (and rain) kept per wave: the levels read rows t-2KH+1 ..
    // t+1 at row-step t
    static constexpr int RING = 2 * KH + 1 <= 16 ? 16 : 32;
    static_assert(OX > 0, "halo wider than the strip");
};

// Index of a ring slot: a value of row r lives in slot (r mod n) of its ring.
// Rows are counted by their phase relative to the first row-step, so that
// the six row-steps of the unrolled loop body use compile-time slots.
__device__ constexpr int slot_of(int phase, int n) { return ((phase % n) + n) % n; }

// Row loads are unconditional, from an address clamped into the grid's
// rows (x0 is even and at most pitch - 2, so both columns lie in the row's
// pitch), so no branch surrounds them and the loads of later rows stay in
// flight while a row-step computes.  Values of cells outside the domain are
// replaced where they are consumed (bd_fix, f_fix).
// The row part of an address is wave-uniform (scalar registers) and the
// column part a per-lane constant `xc` (x0 clamped), so a load needs no
// per-lane address arithmetic.
__device__ __forceinline__ int64_t row_base(const GridDesc& g, int y) {
    return (int64_t)min(max(y, 0), g.H - 1) * g.pitch;
}

// A lane's element of a row: the row's address, made a scalar (it is the
// same in every lane), plus the lane's byte offset zero-extended from 32
// bits, so the load takes its base from scalar registers and needs no
// per-lane 64-bit address.
__device__ __forceinline__ const gfloat* at(const float* plane, int64_t row, uint32_t xbytes) {
    const uint64_t a = reinterpret_cast<uint64_t>(plane + row);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    const char __attribute__((address_space(1)))* p =
        (const char __attribute__((address_space(1)))*)(((uint64_t)hi << 32) | lo);
    return (const gfloat*)(p + xbytes);
}

__device__ __forceinline__ void load_bd(const GridDesc& g, int parity, int y, uint32_t xb,
                                        float (&b)[kSCPL], float (&d)[kSCPL]) {
    const int64_t r = row_base(g, y);
    vload<kSCPL>(at(g.b, r, xb), b);
    vload<kSCPL>(at(g.d[parity], r, xb), d);
}

__device__ __forceinline__ void load_f(const GridDesc& g, int parity, int y, uint32_t xb,
                                       float (&f)[4][kSCPL]) {
    const int64_t r = row_base(g, y);
#pragma unroll
    for (int ch = 0; ch < 4; ++ch) vload<kSCPL>(at(g.f[parity][ch], r, xb), f[ch]);
}

// Stores of the new state are buffer stores that always issue: a lane whose
// cell is not this wavefront's to write gives an offset past the plane's end
// and the hardware drops the write.  With no branch around them, the waits
// on the loads in flight (which share the memory counter with the stores)
// can count the stores, instead of assuming they were skipped and waiting
// for loads issued one row-step ago.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr uint32_t kNoStore = 0x80000000u;  // past every plane the kernel is given

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(float* plane, uint32_t bytes) {
    // raw buffer (stride 0); the last descriptor word as for all gfx9 raw buffers
    return __builtin_amdgcn_make_buffer_rsrc(plane, (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ void store2(__amdgpu_buffer_rsrc_t r, uint32_t off, float a, float b) {
    // Default cache policy.  Stores share the memory counter with the loads
    // that the row-steps wait on, so a store's acknowledgement delays those
    // waits: MI355X, ms per launch of 6 steps at 4096^2 against non-temporal
    // stores (policy 2), 0.211 vs 0.225; cell-updates/s at 8192^2 5.30e11
    // vs 5.06e11, at 16384^2 5.60e11 vs 5.40e11; the batched mix the same.
    __builtin_amdgcn_raw_buffer_store_b64(u32x2{__float_as_uint(a), __float_as_uint(b)}, r,
                                          (int)off, 0, 0);
}

// Ghost terrain and depth for cells of row y outside the domain (the
// closed-wall ghost b = 1e30, d = 0).  `inner` (wave-uniform): every column
// of the strip lies inside the domain.
__device__ __forceinline__ void bd_fix(const GridDesc& g, int y, int x0, bool inner,
                                       float (&b)[kSCPL], float (&d)[kSCPL]) {
    if (inner && (unsigned)y < (unsigned)g.H) return;
#pragma unroll
    for (int c = 0; c < kSCPL; ++c) {
        const bool in = (unsigned)y < (unsigned)g.H && (unsigned)(x0 + c) < (unsigned)g.W;
        b[c] = in ? b[c] : g.ghost_b;
        d[c] = in ? d[c] : g.ghost_d;
    }
}

// Zero flux outside the domain, and no flux toward a closed wall from the
// cells along it (the walls carry nothing, whatever was uploaded).  `inner`
// (wave-uniform): no column of the strip lies outside or on the first or
// last column of the domain.
__device__ __forceinline__ void f_fix(const GridDesc& g, int y, int x0, bool inner,
                                      float (&f)[4][kSCPL]) {
    if (inner && y >= 1 && y < g.H - 1) return;
#pragma unroll
    for (int c = 0; c < kSCPL; ++c) {
        const int x = x0 + c;
        const bool in = (unsigned)y < (unsigned)g.H && (unsigned)x < (unsigned)g.W;
        f[0][c] = in && x != g.W - 1 ? f[0][c] : 0.0f;
        f[1][c] = in && x != 0 ? f[1][c] : 0.0f;
        f[2][c] = in && y != g.H - 1 ? f[2][c] : 0.0f;
        f[3][c] = in && y != 0 ? f[3][c] : 0.0f;
    }
}

}  // namespace

// k_stream: grid of ceil(tasks / kStreamWaves) work groups; task = segment *
// strips + strip, segments of seg_rows output rows.  Closed walls only
// (MODE kModePlain or kModeRain).  Single grid (BATCH false): g1 (by value)
// is the grid.  Batch: hdr = {grids, tasks, offset} and grids[offset..]
// (ordered by tile_begin, the grid's first task; tiles_x strips by tiles_y
// segments of seg_rows rows) live in device memory; launch launch_idx
// advances the grids with at least (launch_idx + 1) KH steps by KH steps.
//
// Registers, per lane and column: level j >= 1 keeps the flux of its rows
// a, a-1, a-2 in a 3-slot ring FS[j] (the new row a overwrites a-3) and the
// +y product of its last flux row in GY[j], levels 0..KH-1 the depth of
// their last three output rows in DS[j] (heads are recomputed from the b
// ring as the next level reads them: 6 VGPRs per level fewer); the input
// rows come through PB (b and d, two rows ahead), BR (the terrain of the last
// two input rows, which level 1 reads) and PF (flux, the row level 1 reads
// plus one ahead).
template <int KH, int MODE, bool BATCH>
__global__ void __launch_bounds__(kStreamWaves * kWave, StreamGeom<KH>::WAVES_PER_SIMD)
k_stream(GridDesc g1, const GridDesc* __restrict__ grids, const int* __restrict__ hdr,
         int launch_idx, int strips1, int segs1, int seg_rows1, int parity1) {
    constexpr bool RAIN = MODE == kModeRain;
    using G = StreamGeom<KH>;
    constexpr int kRing = G::RING;
    __shared__ float s_b[kStreamWaves][kRing][kSEXW];
    __shared__ float s_wet[RAIN ? kStreamWaves : 1][RAIN ? kRing : 1][kSEXW];

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // XCD-aware order: work groups are dealt to the 8 XCDs round robin; each
    // XCD gets a contiguous run of tasks, so neighbouring strips (which
    // re-read each other's halo columns) meet in one L2.  The runs cover the
    // work groups that hold tasks, so the spare ones of a batch launch (its
    // block count is rounded up) fall evenly on the XCDs.
    const int ntasks = BATCH ? hdr[1] : strips1 * segs1;
    const int nb = min((int)gridDim.x, (ntasks + kStreamWaves - 1) / kStreamWaves);
    const int bid = blockIdx.x;
    const int xcd = bid & 7, idx = bid >> 3;
    const int q = nb >> 3, rmd = nb & 7;
    if (idx >= q + (xcd < rmd ? 1 : 0)) return;  // whole work group: no barrier follows
    const int blk = xcd * q + (xcd < rmd ? xcd : rmd) + idx;
    const int task = blk * kStreamWaves + wave;
    if (task >= ntasks) return;  // whole wavefront: no barrier follows
    GridDesc g;
    int strips, seg_rows, parity, local;
    if constexpr (!BATCH) {
        g = g1;
        strips = strips1;
        seg_rows = seg_rows1;
        parity = parity1;
        local = task;
    } else {
        const GridDesc* table = grids + hdr[2];
        int lo = 0, hi = hdr[0] - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (table[mid].tile_begin <= task) lo = mid; else hi = mid - 1;
        }
        g = table[__builtin_amdgcn_readfirstlane(lo)];  // wave-uniform: scalar loads
        if (g.steps < (launch_idx + 1) * KH) return;  // its remaining steps run elsewhere
        strips = g.tiles_x;
        seg_rows = g.seg_rows;
        local = task - g.tile_begin;
        // put this launch's input set at index 0, so that every index into
        // the copy's ping-pong arrays is a constant and the copy can live in
        // registers (a runtime index would keep it in private memory)
        const bool odd = ((g.parity0 + launch_idx) & 1) != 0;
        float* const d_in = odd ? g.d[1] : g.d[0];
        float* const d_out = odd ? g.d[0] : g.d[1];
        g.d[0] = d_in;
        g.d[1] = d_out;
#pragma unroll
        for (int ch = 0; ch < 4; ++ch) {
            float* const f_in = odd ? g.f[1][ch] : g.f[0][ch];
            float* const f_out = odd ? g.f[0][ch] : g.f[1][ch];
            g.f[0][ch] = f_in;
            g.f[1][ch] = f_out;
        }
        parity = 0;
    }
    const int seg = local / strips, strip = local - seg * strips;
    const int Y0 = seg * seg_rows;
    const int Y1 = min(g.H, Y0 + seg_rows);
    if (Y0 >= Y1) return;
    const int ox = strip * G::OX - G::HALOX;
    const int x0 = ox + lane * kSCPL;
    const int lx = lane * kSCPL;
    // x0 is even and at most pitch - 2: both columns lie in the row's pitch
    const uint32_t xc = (uint32_t)min(max(x0, 0), (int)g.pitch - kSCPL) * 4u;  // bytes
    const float F = g.F, K = g.K, C = g.C;
    const int wr = RAIN ? wave : 0;
    // every column of the strip inside the domain (b, d), and none on its
    // first or last column either (f)
    const bool inner_bd = ox >= 0 && ox + kSEXW <= g.W;
    const bool inner_f = ox >= 1 && ox + kSEXW <= g.W - 1;

    // stored columns of this lane (both, or only the first one when the grid
    // ends between them: the second store then lands in the row padding,
    // which nothing reads)
    const bool col_ok = lx >= G::HALOX && lx < G::HALOX + G::OX && x0 < g.W;
    const bool second_in = x0 + 1 < g.W;
    const uint32_t plane_bytes = (uint32_t)(g.pitch * g.H * 4);  // < 2^31 (stream_usable)
    const __amdgpu_buffer_rsrc_t rd = plane_rsrc(g.d[parity ^ 1], plane_bytes);
    const __amdgpu_buffer_rsrc_t r0 = plane_rsrc(g.f[parity ^ 1][0], plane_bytes);
    const __amdgpu_buffer_rsrc_t r1 = plane_rsrc(g.f[parity ^ 1][1], plane_bytes);
    const __amdgpu_buffer_rsrc_t r2 = plane_rsrc(g.f[parity ^ 1][2], plane_bytes);
    const __amdgpu_buffer_rsrc_t r3 = plane_rsrc(g.f[parity ^ 1][3], plane_bytes);
    bool bad = false;

    // rows of the b ring that the rain reaches (bit r & (kRing - 1)), wave-uniform
    unsigned rain_rows = 0;
    const bool rain_on = RAIN && g.rain_on != 0;

    // Accepts input row y: writes b (and the rain amount) into the ring and
    // returns d with the first level's rain added.
    auto accept = [&](int y, float (&b)[kSCPL], float (&d)[kSCPL], float (&dn)[kSCPL]) {
        bd_fix(g, y, x0, inner_bd, b, d);
        const int slot = y & (kRing - 1);
#pragma unroll
        for (int c = 0; c < kSCPL; ++c) s_b[wave][slot][lx + c] = b[c];
        bool wet_row = false;
        if (rain_on) {
            wet_row = __builtin_amdgcn_readfirstlane(
                          brush_reaches(g.rain_cx, g.rain_cy, g.rain_R2, ox, ox + kSEXW - 1, y, y)
                              ? 1 : 0) != 0 &&
                      (unsigned)y < (unsigned)g.H;
        }
        rain_rows = (rain_rows & ~(1u << slot)) | ((wet_row ? 1u : 0u) << slot);
#pragma unroll
        for (int c = 0; c < kSCPL; ++c) dn[c] = d[c];
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
    };

    const int t_begin = Y0 - 2 * KH + 1;
    const int t_end = Y1 + 2 * KH - 2;
#pragma unroll
    for (int s = 0; s < kRing; ++s)
#pragma unroll
        for (int c = 0; c < kSCPL; ++c) s_b[wave][s][lx + c] = 0.0f;
    float FS[KH + 1][3][4][kSCPL];
    float DS[KH][3][kSCPL];
    // (h(a) - h(a+1)) * K of each level's last flux row a: the next row's -y
    // product is its negation (fl(x - y) = -fl(y - x) exactly)
    float GY[KH + 1][kSCPL];
    float PB[2][2][kSCPL];   // [slot][b, d][column]
    float BR[2][kSCPL];      // terrain of the last two input rows (level 1 reads them)
    float PF[2][4][kSCPL];   // [slot][channel][column]
#pragma unroll
    for (int j = 0; j <= KH; ++j)
#pragma unroll
        for (int c = 0; c < kSCPL; ++c) {
            GY[j][c] = 0.0f;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
#pragma unroll
                for (int ch = 0; ch < 4; ++ch) FS[j][r][ch][c] = 0.0f;
                if (j < KH) DS[j][r][c] = 0.0f;
            }
        }
    // Rows relative to row-step t = t_begin (phase 0): level 0 holds rows
    // t-1 (slot 2) and t (slot 0); PF rows t, t+1, t+2 (slots 0, 1, 2); PB rows
    // t+1 (slot 1) and t+2 (slot 0).
    {
        float b[kSCPL], d[kSCPL];
        load_bd(g, parity, t_begin - 1, xc, b, d);
        accept(t_begin - 1, b, d, DS[0][2]);
        load_bd(g, parity, t_begin, xc, b, d);
        accept(t_begin, b, d, DS[0][0]);
#pragma unroll
        for (int c = 0; c < kSCPL; ++c) BR[0][c] = b[c];
    }
    // level 1's product for row t_begin - 1, which its first row reads
    {
        const int rm = (t_begin - 1) & (kRing - 1), ro = t_begin & (kRing - 1);
#pragma unroll
        for (int c = 0; c < kSCPL; ++c)
            GY[1][c] = ((DS[0][2][c] + s_b[wave][rm][lx + c]) - (DS[0][0][c] + s_b[wave][ro][lx + c])) * K;
    }
    load_f(g, parity, t_begin, xc, PF[0]);
    load_f(g, parity, t_begin + 1, xc, PF[1]);
    load_bd(g, parity, t_begin + 1, xc, PB[1][0], PB[1][1]);
    load_bd(g, parity, t_begin + 2, xc, PB[0][0], PB[0][1]);

    // ---- one row-step at phase P (t = t_begin + P mod 6) ------------------
    auto step = [&](auto phase, int t) {
        constexpr int P = decltype(phase)::value;
        // level 0: input row t+1, then the loads of row t+3
        accept(t + 1, PB[slot_of(P + 1, 2)][0], PB[slot_of(P + 1, 2)][1],
               DS[0][slot_of(P + 1, 3)]);
#pragma unroll
        for (int c = 0; c < kSCPL; ++c) BR[slot_of(P + 1, 2)][c] = PB[slot_of(P + 1, 2)][0][c];
        load_bd(g, parity, t + 3, xc, PB[slot_of(P + 1, 2)][0], PB[slot_of(P + 1, 2)][1]);
        // input flux of row t, for level 1
        f_fix(g, t, x0, inner_f, PF[slot_of(P, 2)]);

        float dLast[kSCPL] = {0.0f, 0.0f};  // the last level's new depth row
        // terrain of the rows a, a+1 the next level reads: level 1's rows t,
        // t+1 from registers (the last two input rows), each later level's
        // read from the ring while the level before it computes
        float bn[2][kSCPL];
#pragma unroll
        for (int c = 0; c < kSCPL; ++c) {
            bn[0][c] = BR[slot_of(P, 2)][c];
            bn[1][c] = BR[slot_of(P + 1, 2)][c];
        }
#pragma unroll
        for (int j = 1; j <= KH; ++j) {
            const int xj = P - 2 * (j - 1);          // phase of the level's flux row a
            const int a = t - 2 * (j - 1);
            // level j-1 at rows a-1 (m), a (o), a+1 (p); its own flux of row a
            const int m = slot_of(xj + 2 - 3, 3), o = slot_of(xj + 2 - 2, 3), pp = slot_of(xj + 2 - 1, 3);
            // heads of level j-1 at rows a, a+1: its depth plus b (the rain
            // kernel, short of registers, reads b from the ring here)
            if (RAIN && j > 1) {
                const int ro = a & (kRing - 1), rp = (a + 1) & (kRing - 1);
#pragma unroll
                for (int c = 0; c < kSCPL; ++c) {
                    bn[0][c] = s_b[wave][ro][lx + c];
                    bn[1][c] = s_b[wave][rp][lx + c];
                }
            }
            float h0[kSCPL], hp1[kSCPL];
#pragma unroll
            for (int c = 0; c < kSCPL; ++c) {
                h0[c] = DS[j - 1][o][c] + bn[0][c];
                hp1[c] = DS[j - 1][pp][c] + bn[1][c];
            }
            if (!RAIN && j < KH) {
                const int ro = (a - 2) & (kRing - 1), rp = (a - 1) & (kRing - 1);
#pragma unroll
                for (int c = 0; c < kSCPL; ++c) {
                    bn[0][c] = s_b[wave][ro][lx + c];
                    bn[1][c] = s_b[wave][rp][lx + c];
                }
            }
            // y products: +y of row a, and -y as the negated +y of row a-1
            // (kept from the previous row-step, also while the level ramps up)
            float py[kSCPL], pym[kSCPL];
#pragma unroll
            for (int c = 0; c < kSCPL; ++c) {
                pym[c] = -GY[j][c];
                py[c] = (h0[c] - hp1[c]) * K;
                GY[j][c] = py[c];
            }
            const float (&dOwn)[kSCPL] = DS[j - 1][o];
            const float (&dBase)[kSCPL] = DS[j - 1][m];
            const float (&fOwn)[4][kSCPL] = j == 1 ? PF[slot_of(P, 2)] : FS[j - 1][slot_of(xj + 2 - 2, 3)];
            float (&FN)[4][kSCPL] = FS[j][slot_of(xj, 3)];
            const float (&F1)[4][kSCPL] = FS[j][slot_of(xj - 1, 3)];
            const float (&F2)[4][kSCPL] = FS[j][slot_of(xj - 2, 3)];
            // levels start one after the other: level j's first row-step
            // that can reach the output is t_begin + 4 (j - 1)
            if (j > 1 && t < t_begin + 4 * (j - 1)) continue;
            // flux of row a: flowUpdate.comp:34-62.  x products: +x of column
            // c, whose negation is -x of column c+1 (the last column's from
            // the right lane's first head; the first column's -x from the
            // left lane's last +x product).  Lanes at the strip's edges read
            // 0 for the missing neighbour: their values are wrong whatever
            // the fill, and reach 2 columns further per level, i.e. the
            // 2 KH halo columns and no output column.  (A zero fill lets the
            // lane shift fold into the subtraction.)
            {
                const float h_right = from_right_lane(h0[0]);
                float px[kSCPL];
#pragma unroll
                for (int c = 0; c < kSCPL; ++c)
                    px[c] = (h0[c] - (c < kSCPL - 1 ? h0[c + 1] : h_right)) * K;
                const float pl = from_left_lane_or(px[kSCPL - 1], 0.0f);
#pragma unroll
                for (int c = 0; c < kSCPL; ++c) {
                    float o0 = fOwn[0][c], o1 = fOwn[1][c], o2 = fOwn[2][c], o3 = fOwn[3][c];
                    flux_update_p(F, C, dOwn[c], px[c], -(c > 0 ? px[c - 1] : pl), py[c],
                                  pym[c], o0, o1, o2, o3);
                    FN[0][c] = o0; FN[1][c] = o1; FN[2][c] = o2; FN[3][c] = o3;
                }
            }
            if (j == 1) load_f(g, parity, t + 2, xc, PF[slot_of(P, 2)]);
            // depth of row a-1: flowApply.comp:32-41
            float dN[kSCPL];
            {
                const float in_left = from_left_lane(F1[0][kSCPL - 1]);   // X0.x
                const float in_right = from_right_lane(F1[1][0]);        // X1.y
#pragma unroll
                for (int c = 0; c < kSCPL; ++c) {
                    const float X1y = c < kSCPL - 1 ? F1[1][c + 1] : in_right;
                    const float X0x = c > 0 ? F1[0][c - 1] : in_left;
                    const float Y1w = FN[3][c];
                    const float Y0z = F2[2][c];
                    const float in = ((X1y + X0x) + Y1w) + Y0z;
                    const float out = ((F1[0][c] + F1[1][c]) + F1[2][c]) + F1[3][c];
                    dN[c] = fmaxf(0.0f, dBase[c] + (in - out) * C);
                }
            }
            const int yd = a - 1;  // row of the new depth
            if (j < KH) {
                // the next level's rain, then its head
                const int ring = yd & (kRing - 1);
                if (RAIN && ((rain_rows >> ring) & 1u)) {
#pragma unroll
                    for (int c = 0; c < kSCPL; ++c) {
                        const bool in = x0 + c >= 0 && x0 + c < g.W;
                        if (in) dN[c] = dN[c] + s_wet[wr][ring][lx + c];
                    }
                }
                const int ds = slot_of(xj - 1, 3);
#pragma unroll
                for (int c = 0; c < kSCPL; ++c) DS[j][ds][c] = dN[c];
            } else {
#pragma unroll
                for (int c = 0; c < kSCPL; ++c) dLast[c] = dN[c];
            }
        }
        // Last level: store its depth row a-1 = t - 2 KH + 1 and the flux
        // of that row (F1, final since the previous row-step).  Outside the
        // ramp-up predicate above, so the store always issues; lanes outside
        // the segment (every lane during the ramp) or the strip's output
        // columns store past the plane, which the hardware drops.
        {
            const int yd = t - 2 * KH + 1;
            const float (&F1)[4][kSCPL] = FS[KH][slot_of(P - 2 * (KH - 1) - 1, 3)];
            const bool mine = col_ok && yd >= Y0 && yd < Y1;
            const uint32_t off = mine ? (uint32_t)((int64_t)yd * g.pitch + x0) * 4u : kNoStore;
            bad |= mine && (nonfinite5(dLast[0], F1[0][0], F1[1][0], F1[2][0], F1[3][0]) ||
                            (second_in &&
                             nonfinite5(dLast[1], F1[0][1], F1[1][1], F1[2][1], F1[3][1])));
            store2(rd, off, dLast[0], dLast[1]);
            store2(r0, off, F1[0][0], F1[0][1]);
            store2(r1, off, F1[1][0], F1[1][1]);
            store2(r2, off, F1[2][0], F1[2][1]);
            store2(r3, off, F1[3][0], F1[3][1]);
        }
    };

    // Wait once for the loads above: the loop's waits are then sized by the
    // loads the loop itself keeps in flight (two and three rows ahead), not
    // by the order in which the preloads happened to issue.  vmcnt(0), other
    // counters left at their maxima (gfx9 encoding: expcnt 7, lgkmcnt 15).
    __builtin_amdgcn_s_waitcnt(0x0F70);
    // Even progress of the wavefronts that share a SIMD (one from each of
    // two work groups with the same work): measured at 4096^2, the younger
    // one finished 19% later without turns.  Epochs of 2^11 ticks (20.5 us),
    // re-evaluated every three row-steps.
    const unsigned turn = turn_parity<1>();
    auto share_simd = [&]() { take_turn<11>(turn); };
    // Six row-steps per iteration: every ring has 2 or 3 slots, so each of
    // the six phases addresses its slots at compile time.  The loop body has
    // no exit (the remaining 0..5 row-steps follow it), so the waits on the
    // loads in flight see one path around the loop.
    int t = t_begin;
    for (; t + 5 <= t_end; t += 6) {
        share_simd();
        step(std::integral_constant<int, 0>{}, t);
        step(std::integral_constant<int, 1>{}, t + 1);
        step(std::integral_constant<int, 2>{}, t + 2);
        share_simd();
        step(std::integral_constant<int, 3>{}, t + 3);
        step(std::integral_constant<int, 4>{}, t + 4);
        step(std::integral_constant<int, 5>{}, t + 5);
    }
    if (t <= t_end) step(std::integral_constant<int, 0>{}, t);
    if (t + 1 <= t_end) step(std::integral_constant<int, 1>{}, t + 1);
    if (t + 2 <= t_end) step(std::integral_constant<int, 2>{}, t + 2);
    if (t + 3 <= t_end) step(std::integral_constant<int, 3>{}, t + 3);
    if (t + 4 <= t_end) step(std::integral_constant<int, 4>{}, t + 4);
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

// Longest segment worth a single round of wavefronts (rows).
static constexpr int kStreamSegRows = 320;

StreamPlan stream_plan(int W, int H, int k, int cus) {
    StreamPlan p{};
    int ox = 0, per_simd = 0;
    if (!stream_geom(k, &ox, &per_simd) || W <= 0 || H <= 0 || cus <= 0) return p;
    const int resident_waves = cus * 4 * per_simd;
    p.strips = (W + ox - 1) / ox;
    // one wavefront per segment, as many segments as fill the resident
    // wavefront slots once -- or a whole number of times, so that no segment
    // runs past kStreamSegRows rows -- but no segment shorter than 4 k rows
    // (its cone reaches 2 k rows into each neighbour).  MI355X, closed sheet,
    // k = 6, cell-updates/s by segment rows: 4096^2 81 rows (one round)
    // 3.98e11, ahead of 41-162 rows; 8192^2 161 rows 4.61e11, 328 rows 4.55e11, 107
    // rows 4.49e11; 16384^2 322 rows 5.24e11, 432 rows 5.17e11, 1366 rows
    // 4.68e11; 256 rows in 4.94 rounds 5.24e11 but 274 rows in 4.63 rounds
    // 4.98e11.
    const int segs1 = std::max(1, resident_waves / p.strips);
    const int rows1 = (H + segs1 - 1) / segs1;
    const int rounds = std::max(1, (rows1 + kStreamSegRows - 1) / kStreamSegRows);
    // whole rounds: the last round of a fractional count idles its slots
    int segs = std::max(1, resident_waves * rounds / p.strips);
    segs = std::min(segs, std::max(1, H / (4 * k)));
    p.seg_rows = (H + segs - 1) / segs;
    p.segs = (H + p.seg_rows - 1) / p.seg_rows;
    p.blocks = (p.strips * p.segs + kStreamWaves - 1) / kStreamWaves;
    return p;
}

// One launch of k levels: a single grid (grids null; g1, the plan and the
// parity by value) or a batch (the table, launch launch_idx).
template <int MODE, bool BATCH>
static hipError_t launch_stream_m(int k, const GridDesc& g1, const GridDesc* grids,
                                  const int* hdr, int launch_idx, const StreamPlan& p,
                                  int parity, hipStream_t s) {
    const dim3 grid(p.blocks), block(kStreamWaves * kWave);
#define TWS_STREAM_CASE(KK)                                                                   \
    case KK:                                                                                  \
        hipLaunchKernelGGL((k_stream<KK, MODE, BATCH>), grid, block, 0, s, g1, grids, hdr,    \
                           launch_idx, p.strips, p.segs, p.seg_rows, parity);                 \
        break;
    switch (k) {
        TWS_STREAM_CASE(1)
        TWS_STREAM_CASE(2)
        TWS_STREAM_CASE(3)
        TWS_STREAM_CASE(4)
        TWS_STREAM_CASE(5)
        TWS_STREAM_CASE(6)
        TWS_STREAM_CASE(8)
        default: return hipErrorInvalidValue;
    }
#undef TWS_STREAM_CASE
    return hipGetLastError();
}

hipError_t launch_stream_single(int k, const GridDesc& g, const StreamPlan& p, int parity,
                                hipStream_t s) {
    if (p.blocks <= 0) return hipErrorInvalidValue;
    if (g.boundary != TWS_BOUNDARY_CLOSED) return hipErrorInvalidValue;
    return g.rain_on ? launch_stream_m<kModeRain, false>(k, g, nullptr, nullptr, 0, p, parity, s)
                     : launch_stream_m<kModePlain, false>(k, g, nullptr, nullptr, 0, p, parity, s);
}

// Rows per segment of an H-row grid in a batch: the largest even split into
// segments of at most `cap` rows, but none shorter than 4 k rows when the
// grid has more than one (as stream_plan keeps them).
static int batch_seg_rows(int H, int k, int cap) {
    const int n = std::min((H + cap - 1) / cap, std::max(1, H / (4 * k)));
    return (H + n - 1) / n;
}

bool stream_batch_plan(int W, int H, int k, int cap, int* strips, int* segs, int* seg_rows) {
    int ox = 0;
    if (!stream_shape(k, &ox) || W <= 0 || H <= 0 || cap <= 0) return false;
    *seg_rows = batch_seg_rows(H, k, cap);
    *segs = (H + *seg_rows - 1) / *seg_rows;
    *strips = (W + ox - 1) / ox;
    return true;
}

// Caps tried for a batch, in rows: kStreamSegRows down to 4 k.
static constexpr int kBatchCapStep = 8;

int stream_batch_cap(const int* shapes, int n, int k, int cus) {
    int ox = 0, per_simd = 0;
    if (!stream_geom(k, &ox, &per_simd) || !shapes || n <= 0 || cus < 8) return kStreamSegRows;
    // A batch launch is a few rounds of tasks of unequal length (a grid
    // shorter than the cap is one segment of its own height), and a last
    // round that is nearly empty costs nearly a whole round: MI355X, 256
    // grids of 128^2..2047^2 with rain, k = 6, cell-updates/s by cap: 96 rows
    // 3.97e11, 128 rows 4.01e11, 192 rows 4.22e11, 320 rows 3.72e11 (3.15
    // rounds).  So each cap is priced by replaying the dispatch: the work
    // groups (4 consecutive tasks; one lasts as long as its longest task plus
    // a ramp of 4 k rows) are dealt to the 8 XCDs in contiguous runs, as the
    // kernel maps them, and each XCD starts its run in order on its resident
    // slots.  The cap whose slowest XCD ends first wins; ties keep the
    // longer segments.  Shorter caps only add tasks, so the search stops
    // once the tasks fill 8 rounds, past which the last round's idle slots
    // cost less than the ramps that more segments add.  A candidate's replay
    // stops as soon as some slot ends no earlier than the best cap's slowest
    // XCD: it cannot win any more.
    const int slots = std::max(1, cus / 8 * per_simd);  // work groups resident per XCD
    const int64_t max_tasks = (int64_t)8 * 8 * slots * kStreamWaves;
    struct Run {
        int64_t count;
        int rows;
    };
    std::vector<Run> runs;  // the tasks in order: runs of equal segment rows
    std::vector<int64_t> heap;
    int best = kStreamSegRows;
    int64_t best_span = INT64_MAX;
    for (int cap = kStreamSegRows; cap >= 4 * k; cap -= kBatchCapStep) {
        runs.clear();
        int64_t nt = 0;
        for (int i = 0; i < n; ++i) {
            const int W = shapes[2 * i], H = shapes[2 * i + 1];
            if (W <= 0 || H <= 0) continue;
            const int sr = batch_seg_rows(H, k, cap), segs = (H + sr - 1) / sr;
            const int strips = (W + ox - 1) / ox;
            if (segs > 1) runs.push_back(Run{(int64_t)strips * (segs - 1), sr});
            runs.push_back(Run{strips, H - (segs - 1) * sr});
            nt += (int64_t)strips * segs;
        }
        if (nt == 0) break;
        if (nt > max_tasks && cap < kStreamSegRows) break;
        const int64_t nb = (nt + kStreamWaves - 1) / kStreamWaves;
        const int64_t q = nb / 8, rmd = nb % 8;
        size_t ri = 0;
        int64_t left = runs[0].count;
        int64_t span = 0;
        for (int x = 0; x < 8 && span < best_span; ++x) {
            const int64_t cnt = q + (x < rmd ? 1 : 0);
            heap.assign((size_t)std::min<int64_t>(slots, std::max<int64_t>(cnt, 1)), 0);
            for (int64_t j = 0; j < cnt && span < best_span; ++j) {
                // the work group's tasks: the next kStreamWaves of the runs
                int d = 0;
                for (int take = kStreamWaves; take > 0 && ri < runs.size();) {
                    const int64_t t = std::min<int64_t>(take, left);
                    d = std::max(d, runs[ri].rows);
                    take -= (int)t;
                    left -= t;
                    if (left == 0 && ++ri < runs.size()) left = runs[ri].count;
                }
                std::pop_heap(heap.begin(), heap.end(), std::greater<int64_t>());
                heap.back() += d + 4 * k;
                span = std::max(span, heap.back());
                std::push_heap(heap.begin(), heap.end(), std::greater<int64_t>());
            }
        }
        if (span < best_span) {
            best_span = span;
            best = cap;
        }
    }
    return best;
}

hipError_t launch_stream_batch(int k, int mode, const GridDesc* grids_dev, const int* hdr_dev,
                               int blocks, int launch_idx, hipStream_t s) {
    if (blocks <= 0) return hipErrorInvalidValue;
    StreamPlan p{};
    p.blocks = blocks;
    const GridDesc g1{};
    switch (mode) {
        case kModePlain:
            return launch_stream_m<kModePlain, true>(k, g1, grids_dev, hdr_dev, launch_idx, p, 0, s);
        case kModeRain:
            return launch_stream_m<kModeRain, true>(k, g1, grids_dev, hdr_dev, launch_idx, p, 0, s);
        default:
            return hipErrorInvalidValue;
    }
}

}  // namespace tws
