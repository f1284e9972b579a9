// Drives the C++ mirror of the reference's Terrain class (include/tws_terrain.hpp)
// and checks it against the same steps issued through the C ABI directly.
// Built by __graft_entry__.build(); run by tests/test_terrain_cpp.py.
//   test_terrain [N [out_prefix]]
// N: grid resolution (default 256).  With out_prefix, the initial terrain and
// depth and the final depth and flux are written as raw float32 files
// <out_prefix>{b0,d0,d,f}.f32 for a comparison with the CPU restatement.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tws_terrain.hpp"

static int failures = 0;
#define EXPECT(c)                                                    \
    do {                                                             \
        if (!(c)) { std::printf("FAILED: %s (line %d)\n", #c, __LINE__); ++failures; } \
    } while (0)

static bool write_f32(const std::string& path, const std::vector<float>& v) {
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) return false;
    const bool ok = std::fwrite(v.data(), sizeof(float), v.size(), fp) == v.size();
    return std::fclose(fp) == 0 && ok;
}

int main(int argc, char** argv) {
    const int N = argc > 1 ? std::atoi(argv[1]) : 256;
    const char* out_prefix = argc > 2 ? argv[2] : nullptr;
    if (N < 16) { std::printf("N must be at least 16\n"); return 2; }
    tws::Terrain terrain(N, 1024.0f);
    EXPECT(terrain.Ok());
    // Terrain.h:76 returns (float)(1/step + 0.5) with step = (double)(1.0f/60):
    // 60.499996 at 60 steps/s, as in the reference
    EXPECT(terrain.GetSimulationStepsPerSecond() == (float)(1.0 / (double)(1.0f / 60.0f) + 0.5));
    EXPECT(terrain.GetFlowDamping() == 0.98f && terrain.GetFlowAcceleration() == 10.0f);

    // initial state, captured through the reference layout
    std::vector<float> t0, f0;
    terrain.ReadReferenceTextures(&t0, &f0, nullptr);
    int flux_nonzero = 0;  // the scene starts at rest (Terrain.cpp:231-233)
    for (float v : f0) flux_nonzero += v != 0.0f;
    EXPECT(flux_nonzero == 0);
    std::vector<float> b(N * N), d(N * N);
    for (int i = 0; i < N * N; ++i) { b[i] = t0[i * 4 + 0]; d[i] = t0[i * 4 + 3]; }
    // default lake: water only inside 800 r^4 < 0.45 (r from the grid centre)
    double lake = 0;
    int outside_wet = 0;
    for (int y = 0; y < N; ++y)
        for (int x = 0; x < N; ++x) {
            const double dx = x / double(N - 1) - 0.5, dy = y / double(N - 1) - 0.5;
            const double r2 = dx * dx + dy * dy;
            lake += d[y * N + x];
            if (800.0 * r2 * r2 > 0.4501 && d[y * N + x] > 0.0f) ++outside_wet;
        }
    EXPECT(lake > 0.0);
    EXPECT(outside_wet == 0);
    // ... and it is the reference's own scene (seed of Application.cpp:57)
    std::vector<float> rb(N * N), rd(N * N);
    EXPECT(tws_generate_reference_scene(N, 300.0f, 231656522u, 0, rb.data(), rd.data()) == TWS_OK);
    EXPECT(std::memcmp(rb.data(), b.data(), sizeof(float) * N * N) == 0);
    EXPECT(std::memcmp(rd.data(), d.data(), sizeof(float) * N * N) == 0);

    // 51 ms at 60 steps/s: 3 steps (the step length is the float 1/60 held as
    // double, 0.01666666754 s, so exactly 50 ms would give only 2); then a
    // brush; then a 1 s frame, clamped to 10 steps
    terrain.PerformSimulationStep(0.051);
    EXPECT(terrain.LastStepCount() == 3);
    terrain.ApplyRadialWaterBrush(1024.0f + 100.0f, -20.0f, 2.0f);  // wraps into the grid
    terrain.PerformSimulationStep(1.0);
    EXPECT(terrain.LastStepCount() == 10);
    terrain.SetFlowDamping(0.9f);
    terrain.PerformSimulationStep(1.0 / 60.0 + 1e-9);
    EXPECT(terrain.LastStepCount() == 1);

    // the same through the C ABI
    tws_params p{};
    // the mirror's default edge: the shaders' own, with their dispatch shapes
    EXPECT(terrain.Params().boundary == TWS_BOUNDARY_REFERENCE_DISPATCH);
    EXPECT(tws_params_reference(60.0f, 0.98f, 10.0f, 1024.0f, N, TWS_BOUNDARY_REFERENCE_DISPATCH,
                                &p) == TWS_OK);
    tws_sim* s = nullptr;
    EXPECT(tws_create(N, N, &p, nullptr, &s) == TWS_OK);
    EXPECT(tws_upload(s, b.data(), d.data(), nullptr) == TWS_OK);
    EXPECT(tws_step(s, 3) == TWS_OK);
    const float u = 100.0f / 1024.0f, v = 1.0f - 20.0f / 1024.0f;
    EXPECT(tws_apply_brush(s, u * N, (v - 0.0f) * N, 2.0f, 32.0f) == TWS_OK);
    EXPECT(tws_step(s, 10) == TWS_OK);
    tws_params p2{};
    EXPECT(tws_params_reference(60.0f, 0.9f, 10.0f, 1024.0f, N, TWS_BOUNDARY_REFERENCE_DISPATCH,
                                &p2) == TWS_OK);
    EXPECT(tws_set_params(s, &p2) == TWS_OK);
    EXPECT(tws_step(s, 1) == TWS_OK);
    std::vector<float> dd(N * N), ff(N * N * 4);
    EXPECT(tws_download(s, dd.data(), ff.data(), nullptr) == TWS_OK);
    tws_destroy(s);

    std::vector<float> t1, f1;
    terrain.ReadReferenceTextures(&t1, &f1, nullptr);
    int diff = 0;
    for (int i = 0; i < N * N; ++i) {
        if (t1[i * 4 + 3] != dd[i]) ++diff;
        for (int c = 0; c < 4; ++c)
            if (f1[i * 4 + c] != ff[i * 4 + c]) ++diff;
    }
    EXPECT(diff == 0);
    if (out_prefix) {
        std::vector<float> f_planar(ff);
        const std::string pre(out_prefix);
        EXPECT(write_f32(pre + "b0.f32", b) && write_f32(pre + "d0.f32", d) &&
               write_f32(pre + "d.f32", dd) && write_f32(pre + "f.f32", f_planar));
    }

    // mip chain of the current TerrainData: floor(log2 N) + 1 levels (Texture.cpp:28-44),
    // level 1 = 128^2 texels of (r, 0.3, 0.3, a), each the half-weight bilinear
    // blend of its 2x2 block of level 0
    std::vector<float> mips;
    terrain.ReadMipChain(&mips);
    size_t want = 0;
    int levels = 1;
    for (int m = N; m > 1; m >>= 1) ++levels;
    for (int l = 1; l < levels; ++l)
        want += 4u * (size_t)std::max(1, N >> l) * (size_t)std::max(1, N >> l);
    EXPECT(terrain.MipChainFloats() == want && mips.size() == want);
    int mip_bad = 0;
    auto lerp = [](float a, float b2) { return a + 0.5f * (b2 - a); };
    for (int j = 0; j < N / 2; j += 17)
        for (int i = 0; i < N / 2; i += 13)
            for (int c = 0; c < 4; ++c) {
                const float* q = t1.data();
                const float s00 = q[((2 * j) * N + 2 * i) * 4 + c], s01 = q[((2 * j) * N + 2 * i + 1) * 4 + c];
                const float s10 = q[((2 * j + 1) * N + 2 * i) * 4 + c];
                const float s11 = q[((2 * j + 1) * N + 2 * i + 1) * 4 + c];
                if (mips[((size_t)j * (N / 2) + i) * 4 + c] != lerp(lerp(s00, s01), lerp(s10, s11)))
                    ++mip_bad;
            }
    EXPECT(mip_bad == 0);
    EXPECT(terrain.Ok());
    std::printf(failures ? "FAILED %d\n" : "OK\n", failures);
    return failures ? 1 : 0;
}
