// main.cpp - the gsph CLI (reference src/main.cpp:42-85).
//   gsph --benchmark N | --headless N   run N frames and print the report
//   --scene PATH                        reference scene JSON (default assets/scene_default.json)
//   --config K                          synthetic workload K (0, 1, 2) instead of a JSON scene
//   --autotune                          pick launch options and the SMS/TRA split by timing
//                                       the scene first (no counterpart in the reference)
// The reference's SMS build options (CMakeLists.txt:59-100) as run-time flags:
//   --sms-legacy                        GSPH_USE_REGISTER_SMS=OFF: legacy shared staging
//   --no-sms-local-merge                GSPH_ENABLE_SMS_LOCAL_MERGE=OFF
//   --same-cell-pair-force              GSPH_ENABLE_SAME_CELL_PAIR_FORCE=ON
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>

#include "gsph.h"
#include "hybrid_system.h"

namespace {

// The reference reads assets/ relative to the working directory.  When that
// file is absent (and no --scene was given), the copy shipped beside the
// executable's package (<exe dir>/../assets/...) is used instead.
std::string default_scene(const std::string &path) {
    if (std::ifstream(path).good()) return path;
    char buf[4096];
    const ssize_t k = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
    if (k <= 0) return path;
    buf[k] = '\0';
    const std::string exe(buf);
    const size_t slash = exe.rfind('/');
    if (slash == std::string::npos) return path;
    const std::string alt = exe.substr(0, slash) + "/../" + path;
    return std::ifstream(alt).good() ? alt : path;
}

}  // namespace

int main(int argc, char **argv) {
    int frames = 200;  // headless default (main.cpp:44)
    sph::SceneSource src;
    bool scene_given = false, autotune = false;
    int sms_legacy = 0, local_merge = 1, same_cell = 0, lean = 1;
    for (int i = 1; i < argc; ++i) {
        const bool has_value = i + 1 < argc;
        if (!std::strcmp(argv[i], "--autotune")) {
            autotune = true;
        } else if (!std::strcmp(argv[i], "--sms-legacy")) {
            sms_legacy = 1;
        } else if (!std::strcmp(argv[i], "--no-sms-local-merge")) {
            local_merge = 0;
        } else if (!std::strcmp(argv[i], "--same-cell-pair-force")) {
            same_cell = 1;
        } else if (!std::strcmp(argv[i], "--no-lean-passes")) {
            lean = 0;
        } else if (has_value &&
                   (!std::strcmp(argv[i], "--headless") || !std::strcmp(argv[i], "--benchmark"))) {
            frames = std::atoi(argv[++i]);
        } else if (has_value && !std::strcmp(argv[i], "--scene")) {
            src.json_path = argv[++i];
            scene_given = true;
        } else if (has_value && !std::strcmp(argv[i], "--config")) {
            src.config = std::atoi(argv[++i]);
        }
    }
    if (!scene_given) src.json_path = default_scene(src.json_path);
    if (frames <= 0) {
        std::fprintf(stderr, "Benchmark frame count must be positive.\n");
        return EXIT_FAILURE;
    }
    int ndev = 0;
    if (gsph_device_count(&ndev) != GSPH_OK || ndev == 0) {
        std::fprintf(stderr, "There is no HIP device.\n");
        return EXIT_FAILURE;
    }
    std::printf("HIP initialized.\n");
    const sph::float3h origin{-40.f, -40.f, -40.f};
    const sph::float3h side{80.f, 80.f, 80.f};
    sph::HybridSystem system(side, origin, true, src);
    if (!system.ok()) return EXIT_FAILURE;
    {
        gsph_options o;
        if (gsph_get_options(system.context(), &o) != GSPH_OK) return EXIT_FAILURE;
        o.sms_legacy_staging = sms_legacy;
        o.sms_local_merge = local_merge;
        o.same_cell_pair_force = same_cell;
        o.lean_passes = lean;
        if (gsph_set_options(system.context(), &o) != GSPH_OK) {
            std::fprintf(stderr, "%s\n", gsph_last_error(system.context()));
            return EXIT_FAILURE;
        }
    }
    if (autotune) {
        gsph_options chosen;
        float ms = 0.f;
        if (gsph_autotune(system.context(), 2, 1, &chosen, &ms) != GSPH_OK) {
            std::fprintf(stderr, "%s\n", gsph_last_error(system.context()));
            return EXIT_FAILURE;
        }
        std::printf("Autotuned: %d blocks per CU, staging depth %d, pair masks %d, split %s, "
                    "lean passes %d (%.3f ms per trial step)\n",
                    chosen.blocks_per_cu, chosen.stage_depth, chosen.pair_masks,
                    chosen.split_mode == 0 ? "reference" : "tuned", chosen.lean_passes, ms);
    }
    system.runBenchmark(frames);
    return EXIT_SUCCESS;
}
