// tws_refscene.cpp -- the reference's own default scene, bit for bit: its
// Mersenne-Twister variant (Random.cpp:22-56), its 3-D value-noise fBm
// (NoiseGenerator.cpp:6-97) and the terrain + lake of
// Terrain::CreateHeightmapFromNoiseAndResetSim (Terrain.cpp:200-222).
//
// Every float expression keeps the reference's operation order and rounds
// each operation to float (the library is built without multiply-add
// contraction), as the reference's x64 SSE build evaluates it.

#include <algorithm>
#include <cstdint>

#include "tws.h"
#include "tws_runtime.hpp"

namespace {

// Random.cpp:4-18: the generator's constants.
constexpr int kN = 624, kM = 397;
constexpr uint32_t kMatrixA[2] = {0u, 0x9908B0DFu};
constexpr uint32_t kUpper = 0x80000000u, kLower = 0x7FFFFFFFu;

// The generator of Random.cpp, as an object instead of process-wide state.
class RefRandom {
public:
    // Random.cpp:22-39: a seeded fill, then one pass of the twist.
    explicit RefRandom(uint32_t seed) {
        for (int i = 0; i < kN; ++i)
            mt_[i] = (i % 2) ? seed + (uint32_t)(i * 527) : (2135u + seed * 74111u) * (uint32_t)i;
        for (int i = 0; i < kN; ++i) {
            const uint32_t y = (mt_[i] & kUpper) | (mt_[i + 1 == kN ? 0 : i + 1] & kLower);
            mt_[i] = mt_[(i + kM) % kN] ^ (y >> 1) ^ kMatrixA[y & 1];
        }
    }
    // Random.cpp:42-56: twist one word, temper it, map to [-1, 1].
    float next_float() {
        const int next = idx_ + 1 == kN ? 0 : idx_ + 1;
        uint32_t y = (mt_[idx_] & kUpper) | (mt_[next] & kLower);
        mt_[idx_] = mt_[(idx_ + kM) % kN] ^ (y >> 1) ^ kMatrixA[y & 1];
        y = mt_[idx_];
        idx_ = next;
        y ^= y >> 11;
        y ^= (y << 7) & 0x9D2C5680u;
        y ^= (y << 15) & 0xEFC60000u;
        y ^= y >> 18;
        return (float)((double)y * 4.656612874e-10 - 1.0);
    }

private:
    uint32_t mt_[kN];
    int idx_ = 0;
};

// NoiseGenerator.h:22-24.
inline float smooth(float f) { return f * f * f * (f * (f * 6.0f - 15.0f) + 10.0f); }
inline int floor_int(float a) {
    const int r = (int)a;
    return r - (int)((a < 0) && (a - (float)r != 0.0f));
}

class RefNoise {
public:
    // NoiseGenerator.cpp:6-10: 4096 draws of the generator.
    explicit RefNoise(RefRandom& rnd) {
        for (float& w : white_) w = rnd.next_float();
    }

    // NoiseGenerator.cpp:12-38 (without the gradient, which the scene does
    // not request).
    float value_noise(float cx, float cy, float cz, int low, int high, float persistence,
                      bool periodic) const {
        float res = 0.0f;
        float amplitude = 1.0f;
        float frequency = (float)(1 << low);
        for (int i = low; i <= high; ++i) {
            const float n = noise3d(cx * frequency, cy * frequency, cz * frequency,
                                    periodic ? (int)frequency : 16);
            res += amplitude * (n * 0.5f + 0.5f);
            amplitude *= persistence;
            frequency *= 2.0f;
        }
        return res * 2.0f * (1.0f - persistence) / (1.0f - amplitude) - 1.0f;
    }

private:
    // NoiseGenerator.cpp:40-97: trilinear value noise with the quintic fade,
    // written as the polynomial k0..k7 of the reference.
    float noise3d(float x, float y, float z, int period) const {
        period = (int)std::min<uint32_t>((uint32_t)period, 16u);
        const int mod = period - 1;
        int x0 = floor_int(x), y0 = floor_int(y), z0 = floor_int(z);
        const float fx = x - (float)x0, fy = y - (float)y0, fz = z - (float)z0;
        x0 = (x0 % period + period) & mod;
        y0 = (y0 % period + period) & mod;
        z0 = (z0 % period + period) & mod;
        const int x1 = (x0 + 1) & mod, y1 = (y0 + 1) & mod, z1 = (z0 + 1) & mod;
        const float s000 = white_[x0 + 16 * (y0 + 16 * z0)];
        const float s100 = white_[x1 + 16 * (y0 + 16 * z0)];
        const float s010 = white_[x0 + 16 * (y1 + 16 * z0)];
        const float s110 = white_[x1 + 16 * (y1 + 16 * z0)];
        const float s001 = white_[x0 + 16 * (y0 + 16 * z1)];
        const float s101 = white_[x1 + 16 * (y0 + 16 * z1)];
        const float s011 = white_[x0 + 16 * (y1 + 16 * z1)];
        const float s111 = white_[x1 + 16 * (y1 + 16 * z1)];
        const float u = smooth(fx), v = smooth(fy), w = smooth(fz);
        const float uv = u * v, uw = u * w, vw = v * w;
        const float k0 = s000;
        const float k1 = s100 - s000;
        const float k2 = s010 - s000;
        const float k3 = s001 - s000;
        const float k4 = s110 - s010 - k1;
        const float k5 = s000 - s010 - s001 + s011;
        const float k6 = -k1 - s001 + s101;
        const float k7 = -k4 + s001 - s101 - s011 + s111;
        return k0 + k1 * u + k2 * v + k3 * w + k4 * uv + k5 * vw + k6 * uw + k7 * uv * w;
    }

    float white_[4096];
};

}  // namespace

extern "C" int tws_generate_reference_scene(int32_t resolution, float height_scale, uint32_t seed,
                                            int32_t scene_index, float* terrain, float* depth) {
    if (resolution < 2 || resolution > (1 << 16) || scene_index < 0 || !terrain ||
        !(height_scale == height_scale))
        return tws::fail(TWS_ERR_INVALID, "tws_generate_reference_scene: invalid argument");
    // Application.cpp:57 seeds the generator once; every scene builds a
    // NoiseGenerator, which draws the next 4096 values (NoiseGenerator.cpp:6-10).
    RefRandom rnd(seed);
    for (int64_t i = 0; i < (int64_t)scene_index * 4096; ++i) (void)rnd.next_float();
    const RefNoise noise(rnd);
    const int N = resolution;
    const float m = 1.0f / (float)(N - 1);
    // Terrain.cpp:211-222, rows in parallel as the reference's OpenMP loop
    auto rows = [&](int y0, int y1) {
        for (int y = y0; y < y1; ++y)
            for (int x = 0; x < N; ++x) {
                const size_t i = (size_t)x + (size_t)y * (size_t)N;
                const float b =
                    (noise.value_noise(m * (float)x, m * (float)y, 0.0f, 2, 10, 0.43f, true) * 0.5f +
                     0.5f) * height_scale;
                terrain[i] = b;
                if (depth) {
                    const float dx = (float)x * m - 0.5f, dy = (float)y * m - 0.5f;
                    const float r2 = dx * dx + dy * dy;
                    // pow(r2, 2.0f) of the reference, taken as the correctly rounded square
                    depth[i] = std::max(0.0f, (0.45f - r2 * r2 * 800.0f) * height_scale - b);
                }
            }
    };
    tws::parallel_rows(N, 16u, rows);
    return TWS_OK;
}
