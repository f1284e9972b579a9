// tws_tune.cpp -- launch configuration of the fused step kernel: the default
// table and the measuring autotuner.
//
// choose_launch() maps a grid size to (kernel variant, steps per launch).
// The built-in table comes from an MI355X sweep over sizes 128^2..16384^2:
// 64 x 64 tiles (16 waves up to 512^2, 8 waves up to 2816^2: small grids
// are bound by the latency of one tile's levels, so more and smaller tiles
// finish sooner), 128 x 64 tiles beyond (two work groups per CU so one loads
// while the other computes), 5 steps per launch up to 4096^2 and 4 above;
// grids above 2176^2 in closed walls take the streaming kernel at k = 6.
// tws_autotune() times every candidate (the tile variants and the streaming
// kernel, each at k = 2..6 and 8) on scratch grids of the requested shapes
// advanced as one batch and returns the winner; only tws_tune_record() enters a launch into
// the per-size-class table (log2 of the cell count) that choose_launch()
// consults first, so tuning a batch never changes what a later plain grid
// of the process runs.

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "tws.h"
#include "tws_runtime.hpp"

namespace tws {

static std::mutex g_tune_mu;
static std::map<int, std::pair<int, int>> g_tuned;  // size class -> (kernel, k)

static int size_class(int64_t cells) {
    int c = 0;
    while (cells > 1) { cells >>= 1; ++c; }
    return c;
}

// 512^2 < cells <= 832^2: a CU holding two work groups of a 16-wave tile
// runs both at half speed, so the best launch follows the number of rounds
// of tiles on 256 CUs.  Each band below keeps one round count for its launch
// over the whole band (square grids).  Same-run sweep, 1e9 cell-updates/s with a rain disc
// (without), against 64x64 at k = 4:
//   576^2  64x48_16W k=4: 108 (109) vs 81 (83)      1 round
//   640^2  64x48_16W k=6: 102 (126) vs 95 (98)      2 rounds
//   704^2  64x64_16W k=8: 119 (142) vs 111 (113)    2 rounds
//   768^2  64x48_16W k=5: 133 (148) vs 126 (128)    2 rounds
//   832^2  64x48_16W k=4: 142 (149) vs 120 (123)    2 rounds
// Above 832^2 64x64 at k = 4 leads again (896^2: 142 vs at most 138 with rain).
static void band_launch(int64_t cells, int* kernel, int* k) {
    if (cells <= 576 * 576) {
        *kernel = TWS_KERNEL_TILE_64x48_16W;
        *k = 4;
    } else if (cells <= 640 * 640) {
        *kernel = TWS_KERNEL_TILE_64x48_16W;
        *k = 6;
    } else if (cells <= 704 * 704) {
        *kernel = TWS_KERNEL_TILE_64x64_16W;
        *k = 8;
    } else if (cells <= 768 * 768) {
        *kernel = TWS_KERNEL_TILE_64x48_16W;
        *k = 5;
    } else {
        *kernel = TWS_KERNEL_TILE_64x48_16W;
        *k = 4;
    }
}

void choose_launch(int width, int height, int* kernel, int* k) {
    const int64_t cells = (int64_t)width * height;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(size_class(cells));
        if (it != g_tuned.end()) {
            *kernel = it->second.first;
            *k = it->second.second;
            return;
        }
    }
    // Above 2176^2 the streaming kernel at k = 6 leads (closed walls; other
    // boundaries step on tile_launch's choice).  MI355X, 0.5 m sheet on fBm
    // terrain in closed walls, cell-updates/s stream (k = 6) vs the tile
    // table: 2048^2 2.71e11 (k = 5: 2.89e11) vs 3.02e11, 2304^2 3.19e11
    // (k = 5: 3.22e11) vs 2.95e11, 2560^2 3.50e11 vs 3.15e11, 2816^2 3.70e11
    // vs 3.29e11, 3072^2 3.96e11 vs 3.40e11; 4096^2 4.42e11 vs 3.67e11, 8192^2
    // 4.61e11 vs 3.98e11 and 16384^2 5.3e11 vs 4.26e11 in other runs.
    if (cells > 2176 * 2176) {
        *kernel = TWS_KERNEL_STREAM_128;
        *k = 6;
        return;
    }
    tile_launch(width, height, kernel, k);
}

void tile_launch(int width, int height, int* kernel, int* k) {
    const int64_t cells = (int64_t)width * height;
    // MI355X sweep (cell-updates/s; 0.5 m sheet on fBm terrain, with and without
    // a rain disc; stale halo rows skipped in the 64-wide tiles):
    //   <= 384^2   latency-bound, few tiles: 64x48 tiles over 16 waves of 3
    //              rows put the grid on more CUs; ms per 1000 steps with rain:
    //              256^2 1.70 (64x48, k = 8; 128 tiles) vs 2.10 (64x64, k = 8);
    //              128^2 1.65 vs 2.07; 384^2 2.12 (64x48, k = 6; 160 tiles) vs
    //              2.23 (64x64, k = 8) vs 3.11 (64x48, k = 8: 288 tiles, more
    //              than one per CU)
    //   <= 512^2   latency-bound: 64x64 tiles spread the grid over more CUs, and
    //              16 waves of 4 rows (instead of 8 of 8) halve each wave's
    //              serial work per level: 512^2 1.16e11 (64x64 16 waves, k = 8)
    //              vs 8.7e10 (8 waves, k = 6) vs 5.1e10 (128x64, k = 5);
    //              64x48 tiles: 2.93 ms per 1000 steps at best vs 2.38
    //   <= 832^2   bands by tile rounds (band_launch)
    //   <= 2304^2  64x64 tiles, 8 waves, k = 4: 2048^2 2.98e11 with rain
    //              (3.31e11 without) vs 2.67e11 (2.91e11) for 128x64, k = 5;
    //              1024^2 1.88e11 with rain vs 1.72e11; 2304^2 zero edge, ms
    //              per 100 steps without (with) rain: 1.61 (1.62) vs 1.91
    //              (1.81) for 128x64, k = 5
    //   <= 2816^2  64x64, k = 5 (zero edge, ms per 100 steps without (with)
    //              rain): 2560^2 2.09 (2.10) vs 2.31 (2.15) for 128x64, k = 5
    //              and 2.14 (2.16) at k = 4; 2816^2 2.50 (2.53) vs 2.57 (2.53)
    //   <= 4096^2  128x64, k = 5: 3072^2 2.93 (2.90) ms vs 2.98 (3.01) for 64x64,
    //              k = 5; 4096^2 3.36e11 with rain (3.59e11 without) vs
    //              3.08e11 (3.42e11) for 64x64, k = 4
    //   larger     128x64, k = 4: the 16384^2 dam-break 4.30e11 vs 4.10e11 at
    //              k = 5 (on an all-wet 16384^2 sheet k = 5 leads by 4%).
    if (cells <= 256 * 256) {
        *kernel = TWS_KERNEL_TILE_64x48_16W;
        *k = 8;
    } else if (cells <= 384 * 384) {
        *kernel = TWS_KERNEL_TILE_64x48_16W;
        *k = 6;
    } else if (cells <= 512 * 512) {
        *kernel = TWS_KERNEL_TILE_64x64_16W;
        *k = 8;
    } else if (cells <= 832 * 832) {
        band_launch(cells, kernel, k);
    } else if (cells <= 2304 * 2304) {
        *kernel = TWS_KERNEL_TILE_64x64;
        *k = 4;
    } else if (cells <= 2816 * 2816) {
        *kernel = TWS_KERNEL_TILE_64x64;
        *k = 5;
    } else if (cells <= 4096 * 4096) {
        *kernel = TWS_KERNEL_TILE_128x64;
        *k = 5;
    } else {
        *kernel = TWS_KERNEL_TILE_128x64;
        *k = 4;
    }
}

// Tile order of the step kernel for a grid's default launch: groups of
// tile columns walked row by row (tws_set_tile_order).  MI355X, 4096^2 plain
// sheet at 128x64 / k = 5 (38 tile columns): 3.81e11 cell-updates/s with
// groups of 20 against 3.74e11 with 16, 3.78e11 with 24, 3.73e11 row-major;
// other sizes keep 16 (at 16384^2 and k = 4, widths of 21-37 and 16 came within
// 2% of each other and 25 lost 3%).
int default_tile_group(int width, int height, int kernel, int k) {
    const int64_t cells = (int64_t)width * height;
    if (kernel == TWS_KERNEL_STREAM_128) tile_launch(width, height, &kernel, &k);
    if (kernel == TWS_KERNEL_TILE_128x64 && k == 5 && cells > 2048 * 2048 && cells <= 4096 * 4096)
        return 20;
    return 16;
}

}  // namespace tws

using namespace tws;

extern "C" {

int tws_autotune(const int32_t* widths, const int32_t* heights, int32_t n, int32_t steps,
                 tws_tune_entry* best, tws_tune_entry* results, int32_t capacity,
                 int32_t* count) {
    if (!widths || !heights || n <= 0 || steps <= 0 || !best)
        return fail(TWS_ERR_INVALID, "tws_autotune: invalid argument");
    int64_t bytes = 0, cells = 0;
    for (int i = 0; i < n; ++i) {
        if (widths[i] <= 0 || heights[i] <= 0)
            return fail(TWS_ERR_INVALID, "tws_autotune: invalid shape %d", i);
        bytes += tws_grid_footprint(widths[i], heights[i]);
        cells += (int64_t)widths[i] * heights[i];
    }
    tws_params p{};
    int rc = tws_params_physical(0.02f, 9.81f, 1.0f, 1.0f, 0.995f, TWS_BOUNDARY_CLOSED, &p);
    if (rc != TWS_OK) return rc;
    tws_batch* b = nullptr;
    rc = tws_batch_create(n, bytes, nullptr, &b);
    if (rc != TWS_OK) return rc;
    // synthetic wet state: fBm terrain 0..100 m under a 0.5 m sheet, and a
    // rain disc (r = 16) at the centre as in the batched and served
    // workloads, so the kernel timed is the one with the rain source
    std::vector<float> terrain, depth;
    for (int i = 0; i < n && rc == TWS_OK; ++i) {
        int32_t id = -1;
        rc = tws_batch_add(b, widths[i], heights[i], &p, &id);
        if (rc != TWS_OK) break;
        const size_t c = (size_t)widths[i] * heights[i];
        try {
            terrain.resize(c);
            depth.assign(c, 0.5f);
        } catch (...) {
            rc = fail(TWS_ERR_NOMEM, "tws_autotune: host allocation failed");
            break;
        }
        rc = tws_generate_fbm_terrain(widths[i], heights[i], 1000u + (uint64_t)i, 6, 4, 0.43f,
                                      100.0f, terrain.data());
        if (rc == TWS_OK) rc = tws_batch_upload(b, id, terrain.data(), depth.data(), nullptr);
        if (rc == TWS_OK)
            rc = tws_batch_set_rain(b, id, 0.5f * widths[i], 0.5f * heights[i], 1e-3f, 256.0f);
    }
    static const int kernels[] = {TWS_KERNEL_TILE_128x64,    TWS_KERNEL_TILE_128x128,
                                  TWS_KERNEL_TILE_256x64,    TWS_KERNEL_TILE_64x64,
                                  TWS_KERNEL_TILE_64x64_16W, TWS_KERNEL_TILE_64x48_16W,
                                  TWS_KERNEL_STREAM_128};
    static const int ks[] = {2, 3, 4, 5, 6, 8};
    tws_tune_entry top{0, 0, 0.0};
    int m = 0;
    for (int kern : kernels) {
        for (int kk : ks) {
            if (rc != TWS_OK) break;
            rc = tws_batch_set_launch(b, kern, kk);
            // warm with the same step count: the timed run then replays the
            // captured launch sequence instead of capturing one
            if (rc == TWS_OK) rc = tws_batch_step(b, steps);
            if (rc == TWS_OK) rc = tws_batch_sync(b);
            const auto t0 = std::chrono::steady_clock::now();
            if (rc == TWS_OK) rc = tws_batch_step(b, steps);
            if (rc == TWS_OK) rc = tws_batch_sync(b);
            const double sec =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (rc != TWS_OK) break;
            const tws_tune_entry e{kern, kk, (double)cells * steps / sec};
            if (results && m < capacity) results[m] = e;
            ++m;
            if (e.cell_updates_per_s > top.cell_updates_per_s) top = e;
        }
    }
    tws_batch_destroy(b);
    if (rc != TWS_OK) return rc;
    if (count) *count = m;
    *best = top;
    return TWS_OK;
}

int tws_tune_record(int32_t width, int32_t height, int32_t kernel, int32_t k) {
    if (width <= 0 || height <= 0)
        return fail(TWS_ERR_INVALID, "tws_tune_record: invalid size %d x %d", width, height);
    TileShape sh{};
    int ox = 0;
    if (!(kernel == TWS_KERNEL_STREAM_128 ? k >= 2 && stream_shape(k, &ox) : tile_shape(kernel, k, &sh)))
        return fail(TWS_ERR_INVALID, "tws_tune_record: unknown kernel %d or k %d", kernel, k);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    try {
        g_tuned[size_class((int64_t)width * height)] = {kernel, k};
    } catch (...) {
        return fail(TWS_ERR_NOMEM, "tws_tune_record: host allocation failed");
    }
    return TWS_OK;
}

int tws_tune_forget(void) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned.clear();
    return TWS_OK;
}

}  // extern "C"
