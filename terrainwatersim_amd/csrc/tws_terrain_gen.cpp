// tws_terrain_gen.cpp -- host side of the seeded fBm heightfield (the cell
// function is shared with the device initialiser: tws_fbm.hpp).

#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

#include "tws_fbm.hpp"
#include "tws_runtime.hpp"

namespace {

void fbm_rows(int W, int H, uint64_t seed, int octaves, int base, double persistence,
              double scale, float* out, int y0, int y1) {
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < W; ++x)
            out[(int64_t)y * W + x] = tws::fbm_cell(x, y, W, H, seed, octaves, base, persistence,
                                                    scale);
}

}  // namespace

extern "C" int tws_generate_fbm_terrain(int32_t width, int32_t height, uint64_t seed,
                                        int32_t octaves, int32_t base_frequency,
                                        float persistence, float height_scale, float* out) {
    if (!out || !tws::fbm_args_ok(width, height, octaves, base_frequency))
        return tws::fail(TWS_ERR_INVALID, "tws_generate_fbm_terrain: invalid arguments");
    const int64_t cells = (int64_t)width * height;
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    nt = std::min<unsigned>(nt, 16);
    if (cells < (1 << 16)) nt = 1;
    nt = std::min<unsigned>(nt, (unsigned)height);
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t) {
        const int y0 = (int)((int64_t)height * t / nt), y1 = (int)((int64_t)height * (t + 1) / nt);
        pool.emplace_back(fbm_rows, width, height, seed, octaves, base_frequency,
                          (double)persistence, (double)height_scale, out, y0, y1);
    }
    for (auto& th : pool) th.join();
    return TWS_OK;
}
