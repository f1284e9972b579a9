// tws_runtime.hpp -- host runtime pieces shared by the C-ABI front ends.
#pragma once

#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

#include "tws.h"
#include "tws_internal.hpp"

namespace tws {

int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
int64_t pitch_for(int width);
void fill_ghosts(GridDesc& g, const tws_params& p);
bool valid_params(const tws_params* p);
// Whether a batch can step a W x H grid with these parameters: a
// reference-dispatch grid whose dispatch shapes leave a tail (W or H not a
// multiple of 32) steps on the single-grid path only.
inline bool batchable(int W, int H, const tws_params& p) {
    return p.boundary != TWS_BOUNDARY_REFERENCE_DISPATCH || (W % 32 == 0 && H % 32 == 0);
}

// Compute units of the current device, or 0 if it cannot be queried.
int device_cus();

// Launch configuration (kernel variant, steps per launch) for a grid size,
// from the tuning table in tws_tune.cpp.
void choose_launch(int width, int height, int* kernel, int* k);
// The best tile-kernel launch for a grid size: choose_launch's choice for
// grids it does not give the streaming kernel, and the fallback for grids
// the streaming kernel cannot step.
void tile_launch(int width, int height, int* kernel, int* k);
// Tile order (columns per group) for a grid's default launch.
int default_tile_group(int width, int height, int kernel, int k);

// Runs body(y0, y1) over the rows [0, n) on up to max_threads host threads.
// Rows whose thread cannot be started run on the calling thread, so no
// exception leaves it (the C ABI must not throw).
template <class Body>
void parallel_rows(int n, unsigned max_threads, Body&& body) noexcept {
    unsigned nt = std::max(1u, std::min(max_threads, std::thread::hardware_concurrency()));
    nt = std::min<unsigned>(nt, (unsigned)std::max(n, 1));
    std::vector<std::thread> pool;
    try {
        pool.reserve(nt);
    } catch (...) {
        nt = 1;
    }
    for (unsigned t = 0; t < nt; ++t) {
        const int y0 = (int)((int64_t)n * t / nt), y1 = (int)((int64_t)n * (t + 1) / nt);
        if (y0 >= y1) continue;
        bool started = false;
        if (t + 1 < nt) {
            try {
                pool.emplace_back(body, y0, y1);
                started = true;
            } catch (...) {
            }
        }
        if (!started) body(y0, y1);
    }
    for (auto& th : pool) th.join();
}

}  // namespace tws
