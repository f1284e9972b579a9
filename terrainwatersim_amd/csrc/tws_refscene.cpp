// tws_refscene.cpp -- the reference's own default scene, bit for bit: its
// Mersenne-Twister variant (Random.cpp:22-56), its 3-D value-noise fBm
// (NoiseGenerator.cpp:6-97) and the terrain + lake of
// Terrain::CreateHeightmapFromNoiseAndResetSim (Terrain.cpp:200-222).
//
// Every float expression keeps the reference's operation order and rounds
// each operation to float (the library is built without multiply-add
// contraction), as the reference's x64 SSE build evaluates it.  The noise and
// the per-cell scene are __host__ __device__, so the host generator and the
// device initialiser run the same source and give the same bits; only the
// generator's 4096-value noise table is drawn on the host.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "tws.h"
#include "tws_grid.hpp"
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

constexpr int kWhite = 4096;  // NoiseGenerator.h: 16^3 lattice values

// NoiseGenerator.cpp:6-10: the table of the (scene_index+1)-th noise
// generator after Random::Init(seed); each generator draws the next 4096.
void white_table(uint32_t seed, int scene_index, float* white) {
    RefRandom rnd(seed);
    for (int64_t i = 0; i < (int64_t)scene_index * kWhite; ++i) (void)rnd.next_float();
    for (int i = 0; i < kWhite; ++i) white[i] = rnd.next_float();
}

// NoiseGenerator.h:22-24.
__host__ __device__ inline float smooth(float f) {
    return f * f * f * (f * (f * 6.0f - 15.0f) + 10.0f);
}
__host__ __device__ inline int floor_int(float a) {
    const int r = (int)a;
    return r - (int)((a < 0) && (a - (float)r != 0.0f));
}

// NoiseGenerator.cpp:40-97: trilinear value noise with the quintic fade,
// written as the polynomial k0..k7 of the reference.
__host__ __device__ inline float noise3d(const float* white, float x, float y, float z,
                                         int period) {
    period = (uint32_t)period < 16u ? period : 16;
    const int mod = period - 1;
    int x0 = floor_int(x), y0 = floor_int(y), z0 = floor_int(z);
    const float fx = x - (float)x0, fy = y - (float)y0, fz = z - (float)z0;
    x0 = (x0 % period + period) & mod;
    y0 = (y0 % period + period) & mod;
    z0 = (z0 % period + period) & mod;
    const int x1 = (x0 + 1) & mod, y1 = (y0 + 1) & mod, z1 = (z0 + 1) & mod;
    const float s000 = white[x0 + 16 * (y0 + 16 * z0)];
    const float s100 = white[x1 + 16 * (y0 + 16 * z0)];
    const float s010 = white[x0 + 16 * (y1 + 16 * z0)];
    const float s110 = white[x1 + 16 * (y1 + 16 * z0)];
    const float s001 = white[x0 + 16 * (y0 + 16 * z1)];
    const float s101 = white[x1 + 16 * (y0 + 16 * z1)];
    const float s011 = white[x0 + 16 * (y1 + 16 * z1)];
    const float s111 = white[x1 + 16 * (y1 + 16 * z1)];
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

// NoiseGenerator.cpp:12-38 (without the gradient, which the scene does not
// request).
__host__ __device__ inline float value_noise(const float* white, float cx, float cy, float cz,
                                             int low, int high, float persistence,
                                             bool periodic) {
    float res = 0.0f;
    float amplitude = 1.0f;
    float frequency = (float)(1 << low);
    for (int i = low; i <= high; ++i) {
        const float n = noise3d(white, cx * frequency, cy * frequency, cz * frequency,
                                periodic ? (int)frequency : 16);
        res += amplitude * (n * 0.5f + 0.5f);
        amplitude *= persistence;
        frequency *= 2.0f;
    }
    return res * 2.0f * (1.0f - persistence) / (1.0f - amplitude) - 1.0f;
}

// Terrain.cpp:216-220 for cell (x, y), m = 1/(N-1): the terrain height and
// the lake depth.
__host__ __device__ inline void scene_cell(const float* white, int x, int y, float m,
                                           float height_scale, float& b, float& d) {
    b = (value_noise(white, m * (float)x, m * (float)y, 0.0f, 2, 10, 0.43f, true) * 0.5f + 0.5f) *
        height_scale;
    const float dx = (float)x * m - 0.5f, dy = (float)y * m - 0.5f;
    const float r2 = dx * dx + dy * dy;
    // pow(r2, 2.0f) of the reference, taken as the correctly rounded square
    const float lake = (0.45f - r2 * r2 * 800.0f) * height_scale - b;
    d = lake > 0.0f ? lake : 0.0f;
}

// The scene straight into a grid's planes: b, the depth of set `parity`
// and zero flux.  The noise table is staged in LDS for the gathers.
__global__ void __launch_bounds__(256) k_reference_scene(const float* __restrict__ white_dev,
                                                         int N, float m, float height_scale,
                                                         int64_t pitch, float* b, float* d,
                                                         float* f0, float* f1, float* f2,
                                                         float* f3) {
    __shared__ float white[kWhite];
    for (int i = threadIdx.x; i < kWhite; i += 256) white[i] = white_dev[i];
    __syncthreads();
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= N || y >= N) return;
    float bb, dd;
    scene_cell(white, x, y, m, height_scale, bb, dd);
    const int64_t i = (int64_t)y * pitch + x;
    b[i] = bb;
    d[i] = dd;
    f0[i] = 0.0f;
    f1[i] = 0.0f;
    f2[i] = 0.0f;
    f3[i] = 0.0f;
}

bool scene_args_ok(int32_t resolution, float height_scale, int32_t scene_index) {
    return resolution >= 2 && resolution <= (1 << 16) && scene_index >= 0 &&
           height_scale == height_scale;
}

}  // namespace

extern "C" int tws_generate_reference_scene(int32_t resolution, float height_scale, uint32_t seed,
                                            int32_t scene_index, float* terrain, float* depth) {
    if (!scene_args_ok(resolution, height_scale, scene_index) || !terrain)
        return tws::fail(TWS_ERR_INVALID, "tws_generate_reference_scene: invalid argument");
    // Application.cpp:57 seeds the generator once; every scene builds a
    // NoiseGenerator, which draws the next 4096 values (NoiseGenerator.cpp:6-10).
    std::vector<float> white;
    try {
        white.resize(kWhite);
    } catch (...) {
        return tws::fail(TWS_ERR_NOMEM, "tws_generate_reference_scene: out of host memory");
    }
    white_table(seed, scene_index, white.data());
    const int N = resolution;
    const float m = 1.0f / (float)(N - 1);
    // Terrain.cpp:211-222, rows in parallel as the reference's OpenMP loop
    auto rows = [&](int y0, int y1) {
        for (int y = y0; y < y1; ++y)
            for (int x = 0; x < N; ++x) {
                const size_t i = (size_t)x + (size_t)y * (size_t)N;
                float b, d;
                scene_cell(white.data(), x, y, m, height_scale, b, d);
                terrain[i] = b;
                if (depth) depth[i] = d;
            }
    };
    tws::parallel_rows(N, 16u, rows);
    return TWS_OK;
}

namespace tws {

#define TWS_HIP(call, what)                              \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

int grid_init_reference_scene(GridState& g, float height_scale, uint32_t seed,
                              int32_t scene_index, hipStream_t s) {
    if (g.W != g.H || !scene_args_ok(g.W, height_scale, scene_index))
        return fail(TWS_ERR_INVALID,
                    "init_reference_scene: needs a square grid, a finite height scale and a "
                    "scene index >= 0");
    // The table goes through the idle flux set (4 * pitch * H floats).
    if ((int64_t)4 * g.pitch * g.H < kWhite)
        return fail(TWS_ERR_INVALID,
                    "init_reference_scene: grids below 16 x 16 have no room to stage the noise "
                    "table; use tws_generate_reference_scene and tws_upload");
    std::vector<float> white;
    try {
        white.resize(kWhite);
    } catch (...) {
        return fail(TWS_ERR_NOMEM, "init_reference_scene: out of host memory");
    }
    white_table(seed, scene_index, white.data());
    float* table = g.staging();
    TWS_HIP(hipMemcpyAsync(table, white.data(), sizeof(float) * kWhite, hipMemcpyHostToDevice, s),
            "init_reference_scene: table");
    const int N = g.W, p = g.parity;
    const dim3 grid((N + 15) / 16, (N + 15) / 16);
    hipLaunchKernelGGL(k_reference_scene, grid, dim3(256), 0, s, table, N, 1.0f / (float)(N - 1),
                       height_scale, g.pitch, g.b, g.d[p], g.f[p][0], g.f[p][1], g.f[p][2],
                       g.f[p][3]);
    TWS_HIP(hipGetLastError(), "init_reference_scene: launch");
    // the host table must outlive the copy
    TWS_HIP(hipStreamSynchronize(s), "init_reference_scene: sync");
    return TWS_OK;
}

}  // namespace tws
