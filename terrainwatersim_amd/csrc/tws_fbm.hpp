// tws_fbm.hpp -- the seeded value-noise fBm of one cell, shared by the host
// generator (tws_terrain_gen.cpp) and the device initialiser (k_fbm), so
// both produce the same bits.
//
// The reference builds terrain from a 3-D value-noise fBm with a quintic fade
// over a period-wrapped lattice (NoiseGenerator.cpp:12-97, fade at
// NoiseGenerator.h:22; persistence 0.43 at Terrain.cpp:216).  This keeps those
// choices in 2-D, with lattice values from splitmix64 of (seed, octave, i, j)
// so any size is reproducible from the seed alone; it does not reproduce the
// reference's PRNG (Random.cpp).  Octave o has base*2^o lattice cells across
// the domain and amplitude persistence^o; the sum over the total amplitude
// ([-1, 1]) is mapped to [0, 1] and scaled.  Arithmetic is double (+, -, *,
// /, floor, all correctly rounded, built without contraction) rounded once to
// float, so host and device agree bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tws {

__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Lattice value in [-1, 1).
__host__ __device__ inline double fbm_lattice(uint64_t seed, int octave, int64_t i, int64_t j) {
    const uint64_t key = ((uint64_t)octave << 56) ^ ((uint64_t)i << 28) ^ (uint64_t)j;
    const uint64_t h = splitmix64(seed ^ splitmix64(key));
    return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

__host__ __device__ inline double fbm_fade(double t) {
    return t * t * t * (t * (t * 6.0 - 15.0) + 10.0);
}

__host__ __device__ inline float fbm_cell(int x, int y, int W, int H, uint64_t seed, int octaves,
                                          int base, double persistence, double scale) {
    double amp_total = 0.0, amp = 1.0;
    for (int o = 0; o < octaves; ++o) {
        amp_total += amp;
        amp *= persistence;
    }
    double sum = 0.0;
    amp = 1.0;
    for (int o = 0; o < octaves; ++o) {
        const int64_t freq = (int64_t)base << o;
        const double u = (double)x * (double)freq / (double)W;
        const double v = (double)y * (double)freq / (double)H;
        const double fu = floor(u), fv = floor(v);
        const int64_t i0 = (int64_t)fu % freq, j0 = (int64_t)fv % freq;
        const int64_t i1 = (i0 + 1) % freq, j1 = (j0 + 1) % freq;
        const double su = fbm_fade(u - fu), sv = fbm_fade(v - fv);
        const double a = fbm_lattice(seed, o, i0, j0), b = fbm_lattice(seed, o, i1, j0);
        const double c = fbm_lattice(seed, o, i0, j1), d = fbm_lattice(seed, o, i1, j1);
        const double ab = a + (b - a) * su;
        const double cd = c + (d - c) * su;
        sum += amp * (ab + (cd - ab) * sv);
        amp *= persistence;
    }
    const double n = sum / amp_total;
    return (float)((n * 0.5 + 0.5) * scale);
}

inline bool fbm_args_ok(int W, int H, int octaves, int base) {
    return W > 0 && H > 0 && octaves > 0 && octaves <= 24 && base > 0 &&
           ((int64_t)base << (octaves - 1)) <= (1ll << 26);
}

}  // namespace tws
