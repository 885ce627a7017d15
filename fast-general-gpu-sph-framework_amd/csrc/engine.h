// engine.h - internal hooks that let the serving loop re-compose a batch
// context (admit / evict scenes) without rebuilding it.
#pragma once

#include <vector>

#include "common.h"
#include "gsph.h"

namespace gsph {

struct EngineScene {
    float world[3];
    int count;
    int old_begin;          // first state slot of the scene in the current layout, or -1 (new)
    const float *pos4;      // new scenes: host state
    const float *vel3;      // may be null (zero)
};

size_t engine_bytes(long long n, long long cells);
int engine_create(const gsph_params *p, const gsph_options *opt, int cap_n, int cap_c,
                  int cap_s, gsph_ctx **out);
void engine_capacity(const gsph_ctx *ctx, int *cap_n, int *cap_c, int *cap_s);
SceneGeom engine_geom(const gsph_ctx *ctx, const float world[3], int count);
int engine_recompose(gsph_ctx *src, gsph_ctx *dst, const std::vector<EngineScene> &scenes);
int engine_read_scene(gsph_ctx *ctx, int s, float *pos4, float *vel3, float *ev4);
// First state slot of scene s in the current layout.
int engine_scene_begin(const gsph_ctx *ctx, int s);
int engine_take_errors(gsph_ctx *ctx, int *flags);

}  // namespace gsph
