// tws_terrain_gen.cpp -- host side of the seeded fBm heightfield (the cell
// function is shared with the device initialiser: tws_fbm.hpp).

#include <cstdint>

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
    tws::parallel_rows(height, cells < (1 << 16) ? 1u : 16u, [&](int y0, int y1) {
        fbm_rows(width, height, seed, octaves, base_frequency, (double)persistence,
                 (double)height_scale, out, y0, y1);
    });
    return TWS_OK;
}
