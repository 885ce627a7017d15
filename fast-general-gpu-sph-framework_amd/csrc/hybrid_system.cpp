// hybrid_system.cpp - sph::HybridSystem over the C ABI.
//
// runBenchmark follows the reference (sph_hybrid_system.cpp:341-430): up to
// 5 warm-up frames with per-stage events, the remaining frames back to back
// under a wall clock with one synchronisation at the end, then the FP64 state
// checksum, the hybrid split, the judgeTask pair statistics and the per-stage
// averages, printed with the reference's labels.
#include "hybrid_system.h"

#include <chrono>
#include <cstdio>
#include <iomanip>
#include <iostream>

namespace sph {

HybridSystem::HybridSystem(const float3h &real_world_side, const float3h &sim_origin,
                           bool headless, const SceneSource &src)
    : headless_(headless) {
    int n = 0;
    int rc;
    if (src.config >= 0)
        rc = gsph_scene_config(src.config, &params_, &n, nullptr, 0);
    else
        rc = gsph_scene_json(src.json_path.c_str(), &params_, &n, nullptr, 0);
    if (rc) {
        std::cerr << gsph_last_error(nullptr) << std::endl;
        return;
    }
    std::vector<float> pos(4 * (size_t)n);
    if (src.config >= 0)
        rc = gsph_scene_config(src.config, &params_, &n, pos.data(), n);
    else
        rc = gsph_scene_json(src.json_path.c_str(), &params_, &n, pos.data(), n);
    if (rc) {
        std::cerr << gsph_last_error(nullptr) << std::endl;
        return;
    }
    for (int a = 0; a < 3; ++a) {
        const float w = params_.world_size[a];
        (&sim_ratio_.x)[a] = w > 0 ? (&real_world_side.x)[a] / w : 1.f;
    }
    sim_origin_ = sim_origin;
    gsph_options opt;
    gsph_options_default(&opt);
    rc = gsph_create(&params_, &opt, n, pos.data(), nullptr, &ctx_);
    if (rc) {
        std::cerr << gsph_last_error(nullptr) << std::endl;
        ctx_ = nullptr;
        return;
    }
    nump_ = (unsigned)n;
    std::cout << "Number of particles: " << nump_ << std::endl;
}

HybridSystem::~HybridSystem() { gsph_destroy(ctx_); }

// HybridSystem::insertParticles (sph_hybrid_system.cpp:576-582).
void HybridSystem::insertParticles(unsigned type) {
    if (type == 1) action1_ = !action1_;
}

// The emitter of sph_hybrid_system.cpp:653-676: a block x, z in [0.48, 0.52),
// y in [0.6, 0.7) moving at (0, -6, 0), with ev and acc of every particle
// reset as its BuffInit over all particles does.  The reference steps the
// lattice by spacing_fluid, which it never sets (line 196 is commented out);
// here the step is h/2, the lattice spacing of the scene files.
void HybridSystem::action1() {
    const float lo[3] = {0.48f, 0.6f, 0.48f}, hi[3] = {0.52f, 0.7f, 0.52f};
    const float step = 0.5f * params_.kernel_h;
    std::vector<float> pos, vel;
    for (float x = lo[0]; x < hi[0]; x += step)
        for (float y = lo[1]; y < hi[1]; y += step)
            for (float z = lo[2]; z < hi[2]; z += step) {
                pos.insert(pos.end(), {x, y, z, 0.f});
                vel.insert(vel.end(), {0.f, -6.f, 0.f});
            }
    const int count = (int)(pos.size() / 4);
    if (gsph_add_particles(ctx_, count, pos.data(), vel.data(), 1) != GSPH_OK) {
        std::cerr << "insert: " << gsph_last_error(ctx_) << std::endl;
        action1_ = false;
        return;
    }
    nump_ += (unsigned)count;
}

void HybridSystem::tick() {
    if (!is_running_ || !ctx_) return;
    if (action1_) action1();
    if (gsph_step(ctx_, 1, &last_) != GSPH_OK)
        std::cerr << "tick: " << gsph_last_error(ctx_) << std::endl;
}

void HybridSystem::runBenchmark(int frames) {
    if (frames <= 0 || !ctx_) return;
    if (!is_running_) setPause();
    const int warmup = frames < 5 ? frames : 5;
    double t_arr = 0, t_den = 0, t_for = 0, t_tot = 0;
    for (int i = 0; i < warmup; ++i) {
        tick();
        t_arr += last_.arrange;
        t_den += last_.density;
        t_for += last_.force;
        t_tot += last_.total;
    }
    const int measured = frames - warmup;
    const auto t0 = std::chrono::steady_clock::now();
    if (measured > 0) {
        if (gsph_step(ctx_, measured, nullptr) != GSPH_OK ||
            gsph_sync(ctx_) != GSPH_OK)
            std::cerr << "benchmark: " << gsph_last_error(ctx_) << std::endl;
    }
    const auto t1 = std::chrono::steady_clock::now();
    const double elapsed_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    const double avg_frame = measured > 0 ? elapsed_ms / measured : t_tot / warmup;
    const double fps = 1000.0 / avg_frame;
    double cs[5] = {0, 0, 0, 0, 0};
    gsph_checksum(ctx_, cs);
    int cells = 0, tasks = 0, mid = 0, tra = 0;
    gsph_grid_info(ctx_, &cells, &tasks, &mid, &tra);
    gsph_pair_stats ps;
    gsph_get_pair_stats(ctx_, &ps);
    std::cout << "\n========== Headless benchmark (" << measured << " measured + " << warmup
              << " warm-up frames) ==========\n";
    std::cout << "Total wall time: " << elapsed_ms << " ms\n";
    std::cout << "Average frame time: " << avg_frame << " ms\n";
    std::cout << "FPS: " << fps << "\n";
    std::cout << std::setprecision(17);
    std::cout << "State checksum: " << cs[0] << ", " << cs[1] << ", " << cs[2] << ", " << cs[3]
              << ", " << cs[4] << "\n";
    std::cout << std::setprecision(6);
    std::cout << "Hybrid split: " << tra << " TRA particles, " << tasks << " SMS tasks\n";
    std::cout << "judgeTask isSame SMS pairs: " << ps.is_same_pairs << " / " << ps.total_pairs
              << " (same-cell pairs: " << ps.same_cell_pairs << ", full: " << ps.full_is_same_pairs
              << ", partial: " << ps.partial_is_same_pairs
              << ", compact: " << ps.compact_is_same_pairs
              << ", widened: " << ps.widened_is_same_pairs << ")\n";
    std::cout << "Per-stage averages (HIP events, warm-up frames):\n";
    std::cout << "  arrange/grid : " << (t_arr / warmup) << " ms\n";
    std::cout << "  density      : " << (t_den / warmup) << " ms\n";
    std::cout << "  force        : " << (t_for / warmup) << " ms\n";
    std::cout << "  total kernel : " << (t_tot / warmup) << " ms\n";
    std::cout << "Particle-steps/s: " << (double)nump_ * 1000.0 / avg_frame << "\n";
    std::cout << "==================================================\n" << std::endl;
}

void HybridSystem::setPause() { is_running_ = !is_running_; }

bool HybridSystem::isRunning() { return is_running_; }

unsigned HybridSystem::getNumParticles() { return nump_; }

// The reference returns the render-space position (final_position =
// position * sim_ratio + sim_origin, sph_hybrid_system.cpp:624).
float3h HybridSystem::getPosition(unsigned idx) {
    if (!ctx_ || idx >= nump_) return {0, 0, 0};
    if (host_pos_.size() != 4 * (size_t)nump_) host_pos_.resize(4 * (size_t)nump_);
    gsph_read_state(ctx_, host_pos_.data(), nullptr, nullptr, nullptr);
    const float *p = &host_pos_[4 * (size_t)idx];
    return {p[0] * sim_ratio_.x + sim_origin_.x, p[1] * sim_ratio_.y + sim_origin_.y,
            p[2] * sim_ratio_.z + sim_origin_.z};
}

}  // namespace sph
