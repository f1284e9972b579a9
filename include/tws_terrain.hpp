// tws_terrain.hpp -- C++ host mirror of the simulation half of terrainwatersim's
// Terrain class (terrainwatersim/source/scene/Terrain.h:15-153), over the C
// ABI in tws.h.  Same method names, argument meanings and defaults; the GPU
// work runs through libtws.so instead of GL compute dispatches.
//
// Behaviour kept from the reference:
//   - constructor defaults: 1024 x 1024 grid over 1024 world units, 60 steps
//     per second, damping 0.98, acceleration 10 (Terrain.cpp:21-30, 95);
//   - SetSimulationStepsPerSecond / SetFlowDamping / SetFlowAcceleration
//     re-derive the per-step constants (Terrain.cpp:175-198);
//   - PerformSimulationStep accumulates frame time, runs whole steps, clamps
//     to 10 per call and drops the excess time (Terrain.cpp:240-249);
//   - ApplyRadialWaterBrush wraps the world position into the grid and uses
//     radius^2 = 32 texels (Terrain.cpp:150-168);
//   - CreateHeightmapFromNoiseAndResetSim builds the reference's own scene,
//     bit for bit: terrain (0..300 m) from its NoiseGenerator and PRNG
//     (seed 231656522 as Application.cpp:57, the n-th call drawing the n-th
//     noise table) and the default lake depth max(0, (0.45 - 800 r^4) * 300 - b)
//     (Terrain.cpp:200-238), through tws_generate_reference_scene.
//   - after a frame with steps the TerrainData mip chain is regenerated into
//     the target set by SetMipChainTarget (Terrain.cpp:270-276).
//   - the edge: by default TWS_BOUNDARY_REFERENCE_DISPATCH, i.e. the shaders'
//     own out-of-image reads (water drains off the edges) with their
//     dispatch shapes, so a resolution that is not a multiple of 32 leaves the
//     same tail untouched as Terrain.cpp:167,258,264 do.  At multiples of 32
//     that is TWS_BOUNDARY_ZERO and runs the fused kernels; TWS_BOUNDARY_CLOSED
//     (walls) can be asked for.
// Error behaviour: the reference's methods return void and report failures
// through its log; here they return void as well, print the library's
// message to stderr and keep it for Ok()/LastError().
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "tws.h"

namespace tws {

class Terrain {
public:
    explicit Terrain(uint32_t grid_resolution = 1024, float grid_world_size = 1024.0f,
                     void* hip_stream = nullptr, uint64_t seed = 231656522u,
                     int32_t boundary = TWS_BOUNDARY_REFERENCE_DISPATCH)
        : resolution_(grid_resolution), world_size_(grid_world_size), seed_(seed),
          boundary_(boundary) {
        DeriveParams();
        Check(tws_create((int32_t)resolution_, (int32_t)resolution_, &params_, hip_stream, &sim_));
        if (sim_) CreateHeightmapFromNoiseAndResetSim();
    }
    ~Terrain() { tws_destroy(sim_); }
    Terrain(const Terrain&) = delete;
    Terrain& operator=(const Terrain&) = delete;

    // Updates the simulation by the frame's duration (seconds).
    void PerformSimulationStep(double last_frame_duration_seconds) {
        int32_t n = 0;
        Check(tws_perform_simulation_step(sim_, last_frame_duration_seconds, step_length_, &n));
        last_steps_ = n;
    }

    // Terrain from noise, lake in the middle, zero flux: the reference's scene,
    // computed on the device (below 16 x 16 on the host, then uploaded).
    void CreateHeightmapFromNoiseAndResetSim() {
        const int N = (int)resolution_;
        if (N >= 16) {
            Check(tws_init_reference_scene(sim_, height_scale_, (uint32_t)seed_, scenes_++));
            return;
        }
        std::vector<float> b((size_t)N * N), d((size_t)N * N);
        if (!Check(tws_generate_reference_scene(N, height_scale_, (uint32_t)seed_, scenes_++,
                                                b.data(), d.data())))
            return;
        Check(tws_upload(sim_, b.data(), d.data(), nullptr));
    }

    // Adds water with a radial falloff at a world position (x, z plane).
    void ApplyRadialWaterBrush(float world_x, float world_z, float strength) {
        float u = world_x / world_size_, v = world_z / world_size_;
        u -= std::floor(u);  // ezMath::Fraction
        v -= std::floor(v);
        Check(tws_apply_brush(sim_, u * (float)resolution_, v * (float)resolution_, strength, 32.0f));
    }

    float GetTerrainWorldSize() const { return world_size_; }

    float GetSimulationStepsPerSecond() const { return (float)(1.0 / step_length_ + 0.5); }
    void SetSimulationStepsPerSecond(float steps_per_second) {
        steps_per_second_ = steps_per_second;
        DeriveParams();
        if (sim_) Check(tws_set_params(sim_, &params_));
    }
    float GetFlowDamping() const { return damping_; }
    void SetFlowDamping(float flow_damping) {
        damping_ = flow_damping;
        DeriveParams();
        if (sim_) Check(tws_set_params(sim_, &params_));
    }
    float GetFlowAcceleration() const { return acceleration_; }
    void SetFlowAcceleration(float flow_acceleration) {
        acceleration_ = flow_acceleration;
        DeriveParams();
        if (sim_) Check(tws_set_params(sim_, &params_));
    }

    // Renderer-facing state: TerrainData rgba32f, Flow rgba32f and FlowMap
    // rg16f in the reference's texture layout (Terrain.cpp:205-237).
    void ReadReferenceTextures(std::vector<float>* terrain_rgba, std::vector<float>* flow_rgba,
                               std::vector<uint16_t>* flowmap_rg16) {
        const size_t cells = (size_t)resolution_ * resolution_;
        if (terrain_rgba) terrain_rgba->resize(cells * 4);
        if (flow_rgba) flow_rgba->resize(cells * 4);
        if (flowmap_rg16) flowmap_rg16->resize(cells * 2);
        Check(tws_download_reference_layout(sim_, terrain_rgba ? terrain_rgba->data() : nullptr,
                                            0.3f, flow_rgba ? flow_rgba->data() : nullptr,
                                            flowmap_rg16 ? flowmap_rg16->data() : nullptr));
    }
    // TerrainData mip chain (levels 1.., Terrain.cpp:270-276).  With a device
    // target set, PerformSimulationStep regenerates it after every frame with
    // steps, as the reference does with glGenerateMipmap; the buffer (at
    // least MipChainFloats() floats, 16-byte aligned) stays the caller's.
    size_t MipChainFloats() const {
        size_t f = 0;
        tws_mip_chain_size((int32_t)resolution_, (int32_t)resolution_, nullptr, &f);
        return f;
    }
    void SetMipChainTarget(float* device_rgba, size_t capacity_floats) {
        Check(tws_set_frame_mips(sim_, device_rgba, capacity_floats, 0.3f));
    }
    void ReadMipChain(std::vector<float>* rgba) {
        rgba->resize(MipChainFloats());
        Check(tws_download_mips(sim_, 0.3f, rgba->data(), rgba->size()));
    }
    // Device planes for interop with device-side consumers.
    tws_device_planes DevicePlanes() const {
        tws_device_planes p{};
        tws_get_device_planes(sim_, &p);
        return p;
    }

    tws_sim* Handle() const { return sim_; }
    const tws_params& Params() const { return params_; }
    int32_t LastStepCount() const { return last_steps_; }
    bool Ok() const { return error_.empty(); }
    const std::string& LastError() const { return error_; }

private:
    void DeriveParams() {
        const float ss = steps_per_second_;
        step_length_ = (double)(1.0f / ss);  // Terrain.cpp:177
        Check(tws_params_reference(ss, damping_, acceleration_, world_size_, resolution_,
                                   boundary_, &params_));
    }
    bool Check(int rc) {
        if (rc == TWS_OK) return true;
        error_ = tws_last_error();
        std::fprintf(stderr, "tws::Terrain: %s\n", error_.c_str());
        return false;
    }

    uint32_t resolution_;
    float world_size_;
    uint64_t seed_;
    int32_t boundary_;
    float height_scale_ = 300.0f;
    int32_t scenes_ = 0;  // noise tables drawn so far (one per scene)
    float steps_per_second_ = 60.0f;
    float damping_ = 0.98f;
    float acceleration_ = 10.0f;
    double step_length_ = 1.0 / 60.0;
    tws_params params_{};
    tws_sim* sim_ = nullptr;
    int32_t last_steps_ = 0;
    std::string error_;
};

}  // namespace tws
