// main.cpp - the gsph CLI (reference src/main.cpp:42-85).
//   gsph --benchmark N | --headless N   run N frames and print the report
//   --scene PATH                        reference scene JSON (default assets/scene_default.json)
//   --config K                          synthetic workload K (0, 1, 2) instead of a JSON scene
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "gsph.h"
#include "hybrid_system.h"

int main(int argc, char **argv) {
    int frames = 200;  // headless default (main.cpp:44)
    sph::SceneSource src;
    for (int i = 1; i + 1 < argc; ++i) {
        if (!std::strcmp(argv[i], "--headless") || !std::strcmp(argv[i], "--benchmark"))
            frames = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "--scene"))
            src.json_path = argv[++i];
        else if (!std::strcmp(argv[i], "--config"))
            src.config = std::atoi(argv[++i]);
    }
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
    system.runBenchmark(frames);
    return EXIT_SUCCESS;
}
