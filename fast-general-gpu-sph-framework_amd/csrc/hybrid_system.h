// hybrid_system.h - drop-in host class of the hybrid SPH step.
//
// Mirrors sph::HybridSystem (reference src/simulation/sph_hybrid_system.h:46-137):
// the same constructor arguments, tick(), runBenchmark(frames), setPause(),
// isRunning(), getNumParticles() and getPosition(idx), the same stdout report.
// It runs the step through the C ABI (include/gsph.h); the scene comes from the
// reference scene JSON (default "assets/scene_default.json") or one of the
// synthetic workloads.
#pragma once

#include <string>
#include <vector>

#include "gsph.h"

namespace sph {

struct float3h {
    float x, y, z;
};

struct SceneSource {
    std::string json_path = "assets/scene_default.json";
    int config = -1;  // >= 0: synthetic workload instead of the JSON file
};

class HybridSystem {
public:
    HybridSystem(const float3h &real_world_side, const float3h &sim_origin,
                 bool headless = false, const SceneSource &src = SceneSource());
    ~HybridSystem();
    HybridSystem(const HybridSystem &) = delete;
    HybridSystem &operator=(const HybridSystem &) = delete;

    void tick();
    void runBenchmark(int frames);
    void setPause();
    bool isRunning();
    unsigned getNumParticles();
    float3h getPosition(unsigned idx);
    // type 1 toggles the particle emitter: while it is on, every tick first
    // appends one emitter block (action1) through gsph_add_particles
    void insertParticles(unsigned type);

    bool ok() const { return ctx_ != nullptr; }
    gsph_ctx *context() { return ctx_; }

private:
    void action1();

    gsph_ctx *ctx_ = nullptr;
    bool action1_ = false;
    gsph_params params_{};
    unsigned nump_ = 0;
    bool is_running_ = false;
    bool headless_ = true;
    float3h sim_ratio_{1, 1, 1}, sim_origin_{0, 0, 0};
    gsph_stage_ms last_{};
    std::vector<float> host_pos_;
};

}  // namespace sph
