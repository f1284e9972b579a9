// tws_terrain_gen.cpp -- seeded value-noise fBm heightfield for initial states.
//
// The reference builds its terrain from a 3-D value-noise fBm with a quintic
// fade over a period-wrapped lattice (NoiseGenerator.cpp:12-97, fade at
// NoiseGenerator.h:22; persistence 0.43 at Terrain.cpp:216).  This generator
// keeps those choices -- quintic fade, lattice wrapped to the domain,
// persistence per octave -- in 2-D, with lattice values from splitmix64 of
// (seed, octave, i, j) so that any size is reproducible from the seed alone.
// It does not reproduce the reference's PRNG (Random.cpp); both sides of every
// comparison receive the same arrays.
//
// Octave o has base_frequency * 2^o lattice cells across the domain and
// amplitude persistence^o; the sum is divided by the total amplitude (range
// [-1, 1]), mapped to [0, 1] and scaled by height_scale.  All arithmetic is
// double (+, -, *, floor only) and the result is rounded once to float, so the
// output is identical on every host.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "tws_runtime.hpp"

namespace {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Lattice value in [-1, 1).
inline double lattice(uint64_t seed, int octave, int64_t i, int64_t j) {
    const uint64_t key = ((uint64_t)octave << 56) ^ ((uint64_t)i << 28) ^ (uint64_t)j;
    const uint64_t h = splitmix64(seed ^ splitmix64(key));
    return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

inline double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }

void fbm_rows(int W, int H, uint64_t seed, int octaves, int base, double persistence,
              double scale, float* out, int y0, int y1) {
    double amp_total = 0.0, amp = 1.0;
    for (int o = 0; o < octaves; ++o) { amp_total += amp; amp *= persistence; }
    for (int y = y0; y < y1; ++y) {
        for (int x = 0; x < W; ++x) {
            double sum = 0.0;
            amp = 1.0;
            for (int o = 0; o < octaves; ++o) {
                const int64_t freq = (int64_t)base << o;
                const double u = (double)x * (double)freq / (double)W;
                const double v = (double)y * (double)freq / (double)H;
                const double fu = std::floor(u), fv = std::floor(v);
                const int64_t i0 = (int64_t)fu % freq, j0 = (int64_t)fv % freq;
                const int64_t i1 = (i0 + 1) % freq, j1 = (j0 + 1) % freq;
                const double su = fade(u - fu), sv = fade(v - fv);
                const double a = lattice(seed, o, i0, j0), b = lattice(seed, o, i1, j0);
                const double c = lattice(seed, o, i0, j1), d = lattice(seed, o, i1, j1);
                const double ab = a + (b - a) * su;
                const double cd = c + (d - c) * su;
                sum += amp * (ab + (cd - ab) * sv);
                amp *= persistence;
            }
            const double n = sum / amp_total;
            out[(int64_t)y * W + x] = (float)((n * 0.5 + 0.5) * scale);
        }
    }
}

}  // namespace

extern "C" int tws_generate_fbm_terrain(int32_t width, int32_t height, uint64_t seed,
                                        int32_t octaves, int32_t base_frequency,
                                        float persistence, float height_scale, float* out) {
    if (!out || width <= 0 || height <= 0 || octaves <= 0 || octaves > 24 ||
        base_frequency <= 0 || ((int64_t)base_frequency << (octaves - 1)) > (1ll << 26))
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
