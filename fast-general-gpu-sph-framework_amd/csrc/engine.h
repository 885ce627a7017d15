// engine.h - internal hooks that let the serving loop re-lay out a batch
// context (admit / retire scenes) and run it without host synchronisation.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "common.h"
#include "gsph.h"

namespace gsph {

struct EngineScene {
    float world[3];
    int count;
    int old_begin;          // first state slot in the current layout, or -1 (new)
    int old_cell_base;      // first cell id in the current layout (continuing scenes)
    const float *pos4;      // new scenes: host state (pinned for an asynchronous upload)
    const float *vel3;      // may be null (zero)
};

// Device bytes of the buffers of a context with these capacities (exactly
// what alloc_buffers allocates).
size_t buffer_bytes(const gsph_options &o, size_t N, size_t C, size_t S);
size_t engine_bytes(const gsph_options &o, long long cap_n, long long cap_c, long long cap_s);
int engine_create(const gsph_params *p, const gsph_options *opt, int cap_n, int cap_c,
                  int cap_s, gsph_ctx **out);
void engine_capacity(const gsph_ctx *ctx, int *cap_n, int *cap_c, int *cap_s);
size_t engine_allocated(const gsph_ctx *ctx);
long long engine_cells(const gsph_ctx *ctx, const float world[3]);
hipStream_t engine_stream(const gsph_ctx *ctx);
int engine_relayout(gsph_ctx *src, gsph_ctx *dst, const std::vector<EngineScene> &scenes);
int engine_scene_layout(const gsph_ctx *ctx, int s, int *first, int *cell_base);
int engine_retire_async(gsph_ctx *ctx, int s, float *pos4, float *vel3, float *ev4,
                        hipEvent_t done);
int engine_flags_async(gsph_ctx *ctx, int *flags_pinned);
int engine_step(gsph_ctx *ctx, bool graph);
int engine_read_scene(gsph_ctx *ctx, int s, float *pos4, float *vel3, float *ev4);

}  // namespace gsph
