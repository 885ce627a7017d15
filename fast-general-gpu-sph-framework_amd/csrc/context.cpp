// context.cpp - the C ABI: contexts, device buffers, the step pipeline,
// timing, graphs and read-back.
//
// One context holds one or more scenes (the plain path is a context with one
// scene; gsph_batch_create builds a context whose scenes share every launch).
// Device memory per particle: two copies of the state (A resident, B sorted),
// acceleration, two radix key/value buffers, compact index; per cell: dense
// begin/count plus sparse-written x-layer starts, req and TRA/SMS bases.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "gsph.h"
#include "host_util.h"
#include "engine.h"
#include "kernels.h"

using namespace gsph;

namespace {
thread_local std::string g_thread_error;
}

struct KernelTimer {
    std::vector<std::string> names;
    std::vector<double> ms;
    std::vector<int> launches;
    struct Pending {
        int id;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
};

struct gsph_ctx {
    gsph_params params;
    gsph_options opt;
    gsph_constants c;
    bool batch = false;
    std::vector<SceneGeom> scenes;
    int n = 0, numc = 0, passes = 0;
    int sort_tiles = 0, cell_tiles = 0, nchunks = 0, split_tiles = 0, task_cap = 0;
    int cap_n = 0, cap_c = 0, cap_s = 0;  // allocated capacity
    int grid_blocks = 0;
    int cus = 256;            // compute units of the device
    int device = 0;
    hipStream_t stream = nullptr;
    // device buffers
    float4 *pos[2] = {nullptr, nullptr};
    f3 *vel[2] = {nullptr, nullptr};
    float4 *ev[2] = {nullptr, nullptr};
    f3 *acc = nullptr;
    uint32_t *hash = nullptr, *skey = nullptr;   // sorted keys; stayer keys
    int *perm = nullptr, *sidx = nullptr;        // sorted -> slot; stayer slots
    uint32_t *mkey[2] = {nullptr, nullptr};      // mover list (radix ping-pong)
    int *mval[2] = {nullptr, nullptr};
    unsigned long long *mflag = nullptr, *status = nullptr;
    uint32_t *mpre = nullptr, *tile_m = nullptr, *tile_occ = nullptr, *dhist = nullptr;
    uint32_t *bcnt = nullptr, *split_sum = nullptr;
    int *cells = nullptr, *cstart = nullptr, *chunk_k0 = nullptr, *fill_val = nullptr;
    uint8_t *tile_defer = nullptr;
    int *scene_tab = nullptr;        // scene of every 64th sorted particle (batched lookups)
    int *cell_scene_tab = nullptr;   // and of every 64th cell
    int *begin = nullptr, *req = nullptr, *tra_base = nullptr, *sms_base = nullptr;
    uint2 *xcum = nullptr;
    int *compact = nullptr;
    uint32_t *nbmask = nullptr;   // [19][cap_n] accepted-pair masks and row plans
    uint32_t *sms_pool = nullptr; // SMS accepted-pair masks, rows of 64 words
    int sms_pool_rows = 0;
    int *sms_off = nullptr;       // first pool row of each SMS item
    float *scf = nullptr;         // [10][cap_n] same-cell pair sums (option same_cell_pair_force)
    gsph_task *tasks = nullptr;
    int *dc = nullptr;
    SceneGeom *scenes_d = nullptr;
    double *sum_partial = nullptr, *sum_out = nullptr;
    std::vector<void *> allocs;
    size_t alloc_bytes = 0;       // device bytes of the buffers above
    RelayoutSeg *segs = nullptr;  // [cap_s] scene moves of a serving re-layout
    // pinned host staging of asynchronous layout uploads (serving): a ring of
    // slots, each reused only once the event of its last upload has completed
    struct Staging {
        void *host = nullptr;
        size_t bytes = 0;
        hipEvent_t done = nullptr;
    };
    Staging staging[4];
    int staging_next = 0;
    // state tracking
    int cur = 0;              // buffer holding the current state (0 = A, 1 = B)
    bool grid_valid = false;  // arrange outputs match the state in B
    bool masks_valid = false; // nbmask holds the pair masks of this arrangement
    long long steps = 0;      // completed steps (integrations)
    std::vector<std::array<float, 3>> worlds;  // box of each scene (plain and batch contexts)
    // graph replay of one full step
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // timing
    bool ktiming = false;
    KernelTimer kt;
    hipEvent_t stage_ev[5] = {};
    int inject_after = -1;
    std::string err;
};

namespace gsph {
void set_error(gsph_ctx *ctx, const std::string &msg) {
    if (ctx)
        ctx->err = msg;
    else
        g_thread_error = msg;
}
}  // namespace gsph

namespace {

#define HIP_TRY(ctx, expr)                                                        \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess)                                                     \
            return fail(ctx, GSPH_ERR_HIP,                                        \
                        std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

template <class T>
int dalloc(gsph_ctx *ctx, T **p, size_t count) {
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess)
        return fail(ctx, GSPH_ERR_HIP, "hipMalloc(" + std::to_string(bytes) +
                                           " B): " + hipGetErrorString(e));
    ctx->allocs.push_back(q);
    ctx->alloc_bytes += bytes;
    *p = static_cast<T *>(q);
    return GSPH_OK;
}

int check_launch(gsph_ctx *ctx, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(ctx, GSPH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return GSPH_OK;
}

int consume_injection(gsph_ctx *ctx) {
    if (ctx->inject_after < 0) return GSPH_OK;
    if (ctx->inject_after == 0) {
        ctx->inject_after = -1;
        return fail(ctx, GSPH_ERR_INJECTED, "injected fault");
    }
    --ctx->inject_after;
    return GSPH_OK;
}

// Kernel timing: brackets a launch with events when enabled.
struct Timed {
    gsph_ctx *ctx;
    int id = -1;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(gsph_ctx *c, const char *name) : ctx(c) {
        if (!ctx->ktiming) return;
        KernelTimer &kt = ctx->kt;
        for (size_t i = 0; i < kt.names.size(); ++i)
            if (kt.names[i] == name) id = (int)i;
        if (id < 0) {
            id = (int)kt.names.size();
            kt.names.push_back(name);
            kt.ms.push_back(0.0);
            kt.launches.push_back(0);
        }
        for (hipEvent_t *e : {&a, &b}) {
            if (!kt.pool.empty()) {
                *e = kt.pool.back();
                kt.pool.pop_back();
            } else {
                (void)hipEventCreate(e);
            }
        }
        (void)hipEventRecord(a, ctx->stream);
    }
    ~Timed() {
        if (id < 0) return;
        (void)hipEventRecord(b, ctx->stream);
        ctx->kt.pending.push_back({id, a, b});
    }
};

void drain_timer(gsph_ctx *ctx) {
    KernelTimer &kt = ctx->kt;
    if (kt.pending.empty()) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &p : kt.pending) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, p.a, p.b);
        kt.ms[p.id] += ms;
        kt.launches[p.id] += 1;
        kt.pool.push_back(p.a);
        kt.pool.push_back(p.b);
    }
    kt.pending.clear();
}

ArrangeArgs arrange_args(gsph_ctx *ctx) {
    ArrangeArgs a;
    std::memset(&a, 0, sizeof(a));
    a.n = ctx->n;
    a.numc = ctx->numc;
    a.nscenes = (int)ctx->scenes.size();
    a.scenes = ctx->scenes_d;
    a.scene_tab = ctx->scene_tab;
    a.cell_scene_tab = ctx->cell_scene_tab;
    a.inv_cell_size = ctx->c.inv_cell_size;
    a.split_mode = ctx->opt.split_mode;
    a.sms_min_self = ctx->opt.sms_min_self;
    a.sms_min_total = ctx->opt.sms_min_total;
    a.pos_a = ctx->pos[0];
    a.vel_a = ctx->vel[0];
    a.ev_a = ctx->ev[0];
    a.pos_b = ctx->pos[1];
    a.vel_b = ctx->vel[1];
    a.ev_b = ctx->ev[1];
    a.hash = ctx->hash;
    a.perm = ctx->perm;
    a.skey = ctx->skey;
    a.sidx = ctx->sidx;
    for (int k = 0; k < 2; ++k) {
        a.mkey[k] = ctx->mkey[k];
        a.mval[k] = ctx->mval[k];
    }
    a.mflag = ctx->mflag;
    a.mpre = ctx->mpre;
    a.tile_m = ctx->tile_m;
    a.tile_occ = ctx->tile_occ;
    a.dhist = ctx->dhist;
    a.status = ctx->status;
    a.sample_stride = sample_stride(ctx->n);
    a.msort_tiles = msort_tiles(ctx->n);
    a.passes = ctx->passes;
    a.sort_tiles = ctx->sort_tiles;
    a.cell_tiles = ctx->cell_tiles;
    a.bcnt = ctx->bcnt;
    a.cells = ctx->cells;
    a.cstart = ctx->cstart;
    a.nchunks = ctx->nchunks;
    a.chunk_k0 = ctx->chunk_k0;
    a.fill_val = ctx->fill_val;
    a.begin = ctx->begin;
    a.xcum = ctx->xcum;
    a.req = ctx->req;
    a.tra_base = ctx->tra_base;
    a.sms_base = ctx->sms_base;
    a.split_tiles = ctx->split_tiles;
    a.split_sum = ctx->split_sum;
    a.compact = ctx->compact;
    a.tasks = ctx->tasks;
    a.task_cap = ctx->task_cap;
    a.dc = ctx->dc;
    return a;
}

PhysicsArgs physics_args(gsph_ctx *ctx) {
    PhysicsArgs p;
    std::memset(&p, 0, sizeof(p));
    p.n = ctx->n;
    p.nscenes = (int)ctx->scenes.size();
    p.scenes = ctx->scenes_d;
    p.c = ctx->c;
    p.pos = ctx->pos[1];
    p.vel = ctx->vel[1];
    p.ev = ctx->ev[1];
    p.acc = ctx->acc;
    p.pos_out = ctx->pos[0];
    p.vel_out = ctx->vel[0];
    p.ev_out = ctx->ev[0];
    p.hash = ctx->hash;
    p.begin = ctx->begin;
    p.xcum = ctx->xcum;
    p.compact = ctx->compact;
    p.tasks = ctx->tasks;
    p.dc = ctx->dc;
    p.grid_blocks = ctx->grid_blocks;
    p.row_prune = ctx->opt.row_prune;
    p.sms_local_merge = ctx->opt.sms_local_merge;
    p.sms_legacy = ctx->opt.sms_legacy_staging;
    p.lean = ctx->opt.lean_passes;
    p.scf = ctx->opt.same_cell_pair_force ? ctx->scf : nullptr;
    p.scf_stride = ctx->cap_n;
    p.nbmask = ctx->nbmask;
    p.mask_stride = ctx->cap_n;
    p.use_masks = 0;
    p.stage_depth = ctx->opt.stage_depth;
    p.sms_pool = ctx->sms_pool;
    p.sms_pool_rows = ctx->sms_pool_rows;
    p.sms_off = ctx->sms_off;
    p.tile_defer = ctx->tile_defer;
    p.scene_tab = ctx->scene_tab;
    p.dcw = ctx->dc;
    p.wall_margin = (float)((double)ctx->c.kernel_h + 0.01);
    p.fuse_integrate = 0;
    p.next_keys = ctx->mkey[1];
    return p;
}

int do_arrange(gsph_ctx *ctx) {
    const ArrangeArgs a = arrange_args(ctx);
    hipStream_t s = ctx->stream;
    if (ctx->n == 0) {
        { Timed t(ctx, "empty_grid"); launch_empty_grid(a, s); }
        { Timed t(ctx, "fill"); launch_fill(a, s); }
        { Timed t(ctx, "compact"); launch_compact(a, s); }
    } else {
        { Timed t(ctx, "hash"); launch_hash(a, ctx->batch, s); }
        { Timed t(ctx, "movers"); launch_movers(a, s); }
        for (int p = 0; p < ctx->passes; ++p) {
            Timed t(ctx, "msort");
            launch_msort(a, p, s);
        }
        { Timed t(ctx, "mplace"); launch_mplace(a, s); }
        { Timed t(ctx, "merge"); launch_merge(a, s); }
        { Timed t(ctx, "cells"); launch_cells(a, s); }
        { Timed t(ctx, "fill"); launch_fill(a, s); }
        { Timed t(ctx, "split_count"); launch_split_count(a, ctx->batch, s); }
        { Timed t(ctx, "split_apply"); launch_split_apply(a, ctx->batch, s); }
        { Timed t(ctx, "compact"); launch_compact(a, s); }
    }
    int rc = check_launch(ctx, "arrange");
    if (rc) return rc;
    ctx->cur = 1;
    ctx->grid_valid = true;
    ctx->masks_valid = false;
    return GSPH_OK;
}

int do_density(gsph_ctx *ctx) {
    if (ctx->n == 0) return GSPH_OK;
    PhysicsArgs p = physics_args(ctx);
    p.use_masks = ctx->opt.pair_masks != 0;
    { Timed t(ctx, "density"); launch_density(p, ctx->batch, ctx->stream); }
    const int rc = check_launch(ctx, "density");
    ctx->masks_valid = !rc && p.use_masks;
    return rc;
}

// With fuse the force kernel also integrates each particle (the full step);
// do_integrate(ctx, true) then only does the bookkeeping.
int do_force(gsph_ctx *ctx, bool fuse = false) {
    if (ctx->n == 0) return GSPH_OK;
    PhysicsArgs p = physics_args(ctx);
    p.use_masks = ctx->masks_valid && ctx->opt.pair_masks != 0;
    p.fuse_integrate = fuse ? 1 : 0;
    if (p.scf) { Timed t(ctx, "same_cell"); launch_same_cell(p, ctx->stream); }
    { Timed t(ctx, "force"); launch_force(p, ctx->batch, ctx->stream); }
    return check_launch(ctx, "force");
}

int do_integrate(gsph_ctx *ctx, bool fused = false) {
    if (ctx->n > 0 && !fused) {
        const PhysicsArgs p = physics_args(ctx);
        { Timed t(ctx, "integrate"); launch_integrate(p, ctx->batch, ctx->stream); }
        int rc = check_launch(ctx, "integrate");
        if (rc) return rc;
    }
    ctx->cur = 0;
    ctx->grid_valid = false;
    ctx->masks_valid = false;
    ++ctx->steps;
    return GSPH_OK;
}

int do_step(gsph_ctx *ctx) {
    int rc = do_arrange(ctx);
    if (!rc) rc = do_density(ctx);
    if (!rc) rc = do_force(ctx, true);
    if (!rc) rc = do_integrate(ctx, true);
    return rc;
}

// The slot order no longer follows hash[] (state or layout set from the
// host): the next arrange treats every particle as moved.  Host API paths
// only, never inside a captured step.
int invalidate_order(gsph_ctx *ctx) {
    HIP_TRY(ctx, hipMemsetAsync(ctx->dc + kCntOrder, 0, sizeof(int), ctx->stream));
    HIP_TRY(ctx, hipMemsetAsync(ctx->dc + kCntKeysReady, 0, sizeof(int), ctx->stream));
    return GSPH_OK;
}

// Reports one raised device error flag and clears only that flag, so that a
// second condition raised in the same step (say a capacity overflow next to a
// particle that left its box) is reported by the next call instead of lost.
int collect_errors(gsph_ctx *ctx) {
    int flags = 0;
    HIP_TRY(ctx, hipMemcpyAsync(&flags, ctx->dc + kCntErr, sizeof(int),
                                hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (!flags) return GSPH_OK;
    const int bit = (flags & kErrOutOfDomain) ? kErrOutOfDomain
                    : (flags & kErrCapacity)  ? kErrCapacity
                    : (flags & kErrSortStall) ? kErrSortStall
                                              : flags;
    const int rest = flags & ~bit;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->dc + kCntErr, &rest, sizeof(int),
                                hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (bit == kErrOutOfDomain)
        return fail(ctx, GSPH_ERR_OUT_OF_DOMAIN, "a particle left the grid (or is not finite)");
    if (bit == kErrCapacity)
        return fail(ctx, GSPH_ERR_CAPACITY,
                    "task capacity exceeded or a cell holds more than 65535 particles");
    if (bit == kErrSortStall)
        return fail(ctx, GSPH_ERR_SPIN_TIMEOUT, "a sort tile waited too long for its predecessor");
    return fail(ctx, GSPH_ERR_HIP, "unknown device error flag " + std::to_string(bit));
}

int bits_for(long long maxval) {
    int b = 1;
    while (b < 40 && (1LL << b) <= maxval) ++b;
    return b;
}

// Assigns cell bases to the scenes and derives every launch size from the
// active particle / cell counts.  Fails if they exceed the allocated capacity.
int configure_layout(gsph_ctx *ctx, std::vector<SceneGeom> scenes, int n) {
    long long numc = 0, pb = 0;
    for (auto &g : scenes) {
        if (g.gx <= 0 || g.gy <= 0 || g.gz <= 0 || g.gx > 16383 || g.gy > 16383 || g.gz > 16383)
            return fail(ctx, GSPH_ERR_INVALID_ARG, "grid dimensions must be in [1, 16383]");
        g.cell_base = (int)numc;
        g.p_begin = (int)pb;
        numc += (long long)g.gx * g.gy * g.gz;
        pb += g.p_count;
        if (numc > (1LL << 26))
            return fail(ctx, GSPH_ERR_CAPACITY,
                        "more than 2^26 grid cells (micro ids must fit 32 bits)");
    }
    if (pb != n) return fail(ctx, GSPH_ERR_INVALID_ARG, "scene counts do not add up");
    if (n > ctx->cap_n || numc > ctx->cap_c || (int)scenes.size() > ctx->cap_s)
        return fail(ctx, GSPH_ERR_CAPACITY, "layout exceeds the context capacity");
    ctx->scenes = scenes;
    ctx->n = n;
    ctx->numc = (int)numc;
    ctx->passes = (bits_for(std::max(numc * 64 - 1, 1LL)) + 7) / 8;
    ctx->sort_tiles = std::max(1, (n + kSortTile - 1) / kSortTile);
    ctx->cell_tiles = std::max(1, (n + kCellTile - 1) / kCellTile);
    ctx->nchunks = std::max(1, (ctx->numc + kFillChunk - 1) / kFillChunk);
    ctx->split_tiles = std::max(1, (n + kSplitTile - 1) / kSplitTile);
    ctx->task_cap = (n + 31) / 32 + std::min(ctx->numc, n) + 2;
    // a new layout: the fill writes every chunk again (a chunk's extent and
    // the meaning of its cells may have changed)
    HIP_TRY(ctx, hipMemsetAsync(ctx->fill_val, 0xff, sizeof(int) * (ctx->nchunks + 1), ctx->stream));
    if (!scenes.empty()) {
        HIP_TRY(ctx, hipMemcpyAsync(ctx->scenes_d, ctx->scenes.data(),
                                    sizeof(SceneGeom) * ctx->scenes.size(),
                                    hipMemcpyHostToDevice, ctx->stream));
        launch_scene_tabs(ctx->scenes_d, (int)scenes.size(), ctx->scene_tab, n, ctx->cell_scene_tab,
                          ctx->numc, ctx->stream);
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    }
    // launch sizes changed: a captured step graph is stale
    if (ctx->graph_exec) {
        (void)hipGraphExecDestroy(ctx->graph_exec);
        ctx->graph_exec = nullptr;
    }
    if (ctx->graph) {
        (void)hipGraphDestroy(ctx->graph);
        ctx->graph = nullptr;
    }
    ctx->grid_valid = false;
    ctx->masks_valid = false;
    return GSPH_OK;
}

// SMS / long-row mask pool rows for a particle capacity: about one
// 32-candidate sub-batch row (64 lanes x 4 B) per 256 bytes of
// sms_mask_bytes per particle.
size_t pool_rows_for(const gsph_options &o, size_t N) {
    const size_t rows = o.sms_mask_bytes > 0 ? N * (size_t)o.sms_mask_bytes / 256 : 0;
    return std::min<size_t>(rows, (size_t)INT32_MAX / 2);
}

}  // namespace

namespace gsph {
// Device bytes that alloc_buffers allocates for capacities N particles, C
// cells and S scenes under options o (every buffer, as sized there; each
// allocation at least one element).  alloc_buffers checks its own total
// against this, so the serving loop's HBM accounting is exact.
size_t buffer_bytes(const gsph_options &o, size_t N, size_t C, size_t S) {
    S = std::max<size_t>(S, 1);
    const size_t sort_tiles = std::max<size_t>(1, (N + kSortTile - 1) / kSortTile);
    const size_t cell_tiles = std::max<size_t>(1, (N + kCellTile - 1) / kCellTile);
    const size_t split_tiles = std::max<size_t>(1, (N + kSplitTile - 1) / kSplitTile);
    const size_t nchunks = std::max<size_t>(1, (C + kFillChunk - 1) / kFillChunk);
    const size_t task_cap = (N + 31) / 32 + std::min(C, N) + 2;
    const size_t rows = pool_rows_for(o, N);
    auto b = [](size_t count, size_t elem) { return std::max<size_t>(count, 1) * elem; };
    size_t t = 0;
    t += 2 * (b(N, 16) + b(N, 12) + b(N, 16) + b(N, 4) + b(N, 4));   // pos vel ev mkey mval
    t += 4 * b(N, 4);                                                 // hash perm skey sidx
    t += b((N + 63) / 64, 8) + b((N + 63) / 64, 4);                   // mflag mpre
    t += 2 * b(sort_tiles, 4) + b(4 * 256, 4);                        // tile_m tile_occ dhist
    t += b(4 * 256 * (size_t)msort_tiles((int)N), 8);                 // status
    t += b(N, 12);                                                    // acc
    t += b(cell_tiles, 4) + b(3 * split_tiles, 4);                    // bcnt split_sum
    t += 2 * b(N + 1, 4) + 2 * b(nchunks + 1, 4);                     // cells cstart chunk_k0 fill_val
    t += b(C + 1, 4) + 3 * b(C, 4) + b(C, 8);                         // begin req tra sms xcum
    t += b((C + 63) / 64, 4);                                         // cell_scene_tab
    t += b(N, 4) + b((size_t)kMaskWords * N, 4);                      // compact nbmask
    t += b((N + 63) / 64, 1) + b((N + 63) / 64, 4);                   // tile_defer scene_tab
    t += b(rows * 64, 4) + b(task_cap / 2 + 2, 4);                    // sms_pool sms_off
    t += b(task_cap, sizeof(gsph_task)) + b(kCntSlots, 4);            // tasks dc
    t += b(S, sizeof(SceneGeom)) + b(S, sizeof(RelayoutSeg));         // scenes segs
    if (o.same_cell_pair_force) t += b(10 * N, 4);                    // scf
    t += b(5 * 256, 8) + b(5, 8);                                     // checksum
    return t;
}
}  // namespace gsph

namespace {

// All device buffers, sized by capacity (particles, cells, scenes).
int alloc_buffers(gsph_ctx *ctx, int cap_n, int cap_c, int cap_s) {
    ctx->cap_n = cap_n;
    ctx->cap_c = cap_c;
    ctx->cap_s = std::max(cap_s, 1);
    const size_t N = (size_t)cap_n, C = (size_t)cap_c;
    const size_t sort_tiles = std::max<size_t>(1, (N + kSortTile - 1) / kSortTile);
    const size_t cell_tiles = std::max<size_t>(1, (N + kCellTile - 1) / kCellTile);
    const size_t split_tiles = std::max<size_t>(1, (N + kSplitTile - 1) / kSplitTile);
    const size_t nchunks = std::max<size_t>(1, (C + kFillChunk - 1) / kFillChunk);
    const size_t task_cap = (N + 31) / 32 + std::min(C, N) + 2;
    int rc;
#define ALLOC(ptr, cnt)                          \
    do {                                         \
        rc = dalloc(ctx, &(ptr), (size_t)(cnt)); \
        if (rc) return rc;                       \
    } while (0)
    for (int k = 0; k < 2; ++k) {
        ALLOC(ctx->pos[k], N);
        ALLOC(ctx->vel[k], N);
        ALLOC(ctx->ev[k], N);
        ALLOC(ctx->mkey[k], N);
        ALLOC(ctx->mval[k], N);
    }
    ALLOC(ctx->hash, N);
    ALLOC(ctx->perm, N);
    ALLOC(ctx->skey, N);
    ALLOC(ctx->sidx, N);
    ALLOC(ctx->mflag, (N + 63) / 64);
    ALLOC(ctx->mpre, (N + 63) / 64);
    ALLOC(ctx->tile_m, sort_tiles);
    ALLOC(ctx->tile_occ, sort_tiles);
    ALLOC(ctx->dhist, 4 * 256);
    ALLOC(ctx->status, 4 * 256 * (size_t)msort_tiles((int)N));
    ALLOC(ctx->acc, N);
    ALLOC(ctx->bcnt, cell_tiles);
    ALLOC(ctx->split_sum, 3 * split_tiles);
    ALLOC(ctx->cells, N + 1);
    ALLOC(ctx->cstart, N + 1);
    ALLOC(ctx->chunk_k0, nchunks + 1);
    ALLOC(ctx->fill_val, nchunks + 1);
    ALLOC(ctx->begin, C + 1);
    ALLOC(ctx->req, C);
    ALLOC(ctx->tra_base, C);
    ALLOC(ctx->sms_base, C);
    ALLOC(ctx->xcum, C);
    ALLOC(ctx->compact, N);
    ALLOC(ctx->tile_defer, (N + 63) / 64);
    ALLOC(ctx->scene_tab, (N + 63) / 64);
    ALLOC(ctx->cell_scene_tab, (C + 63) / 64);
    ALLOC(ctx->nbmask, (size_t)kMaskWords * N);
    // SMS masks: about one 32-candidate sub-batch row per 8 bytes of budget
    // per particle (rows hold 64 lanes x 4 B)
    ctx->sms_pool_rows = (int)pool_rows_for(ctx->opt, N);
    ALLOC(ctx->sms_pool, (size_t)ctx->sms_pool_rows * 64);
    ALLOC(ctx->sms_off, task_cap / 2 + 2);
    ALLOC(ctx->tasks, task_cap);
    ALLOC(ctx->dc, kCntSlots);
    ALLOC(ctx->scenes_d, ctx->cap_s);
    ALLOC(ctx->segs, ctx->cap_s);
    ctx->scf = nullptr;
    if (ctx->opt.same_cell_pair_force) ALLOC(ctx->scf, 10 * N);
    ALLOC(ctx->sum_partial, 5 * 256);
    ALLOC(ctx->sum_out, 5);
#undef ALLOC
    if (ctx->alloc_bytes != buffer_bytes(ctx->opt, N, C, (size_t)ctx->cap_s))
        return fail(ctx, GSPH_ERR_HIP, "internal: buffer accounting differs from the allocation");
    HIP_TRY(ctx, hipMemsetAsync(ctx->dc, 0, sizeof(int) * kCntSlots, ctx->stream));
    HIP_TRY(ctx, hipMemsetAsync(ctx->dhist, 0, sizeof(uint32_t) * 4 * 256, ctx->stream));
    HIP_TRY(ctx, hipMemsetAsync(ctx->status, 0,
                                sizeof(unsigned long long) * 4 * 256 * (size_t)msort_tiles((int)N),
                                ctx->stream));
    HIP_TRY(ctx, hipMemsetAsync(ctx->tasks, 0, sizeof(gsph_task) * task_cap, ctx->stream));
    // no chunk of begin[] is known uniform yet: the first fill writes every chunk
    HIP_TRY(ctx, hipMemsetAsync(ctx->fill_val, 0xff, sizeof(int) * (nchunks + 1), ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return GSPH_OK;
}

// An empty context: constants, stream, events; no buffers yet.
// Persistent density/force grid from blocks_per_cu: a multiple of 8 blocks
// so that every XCD gets the same share.
void set_grid(gsph_ctx *ctx) {
    const int bpc = ctx->opt.blocks_per_cu > 0 ? ctx->opt.blocks_per_cu : 32;
    ctx->grid_blocks = std::max(8, (ctx->cus * bpc + 7) / 8 * 8);
}

// Validates o; 0 selects the default of blocks_per_cu and stage_depth.
int check_options(gsph_ctx *ctx, gsph_options &o) {
    if (o.blocks_per_cu == 0) o.blocks_per_cu = 32;
    if (o.stage_depth == 0) o.stage_depth = 128;
    if (o.stage_depth != 64 && o.stage_depth != 128 && o.stage_depth != 256)
        return fail(ctx, GSPH_ERR_INVALID_ARG, "stage_depth must be 64, 128 or 256");
    if (o.blocks_per_cu < 1 || o.blocks_per_cu > 32)
        return fail(ctx, GSPH_ERR_INVALID_ARG, "blocks_per_cu must be in [1, 32]");
    if (o.split_mode != 0 && o.split_mode != 1)
        return fail(ctx, GSPH_ERR_INVALID_ARG, "split_mode must be 0 or 1");
    if (o.sms_mask_bytes < 0 || o.sms_mask_bytes > 65536)
        return fail(ctx, GSPH_ERR_INVALID_ARG, "sms_mask_bytes must be in [0, 65536]");
    if (o.sms_local_merge != 0 && o.sms_local_merge != 1)
        return fail(ctx, GSPH_ERR_INVALID_ARG, "sms_local_merge must be 0 or 1");
    if (o.sms_legacy_staging != 0 && o.sms_legacy_staging != 1)
        return fail(ctx, GSPH_ERR_INVALID_ARG, "sms_legacy_staging must be 0 or 1");
    if (o.same_cell_pair_force != 0 && o.same_cell_pair_force != 1)
        return fail(ctx, GSPH_ERR_INVALID_ARG, "same_cell_pair_force must be 0 or 1");
    if (o.lean_passes != 0 && o.lean_passes != 1)
        return fail(ctx, GSPH_ERR_INVALID_ARG, "lean_passes must be 0 or 1");
    return GSPH_OK;
}

int new_context(const gsph_params *p, const gsph_options *opt, bool batch, gsph_ctx **out) {
    if (!p || !out) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(nullptr, GSPH_ERR_NO_DEVICE, "There is no HIP device.");
    gsph_ctx *ctx = new gsph_ctx();
    ctx->params = *p;
    if (opt)
        ctx->opt = *opt;
    else
        gsph_options_default(&ctx->opt);
    if (check_options(nullptr, ctx->opt)) {
        delete ctx;
        return GSPH_ERR_INVALID_ARG;
    }
    int rc = gsph_derive_constants(p, &ctx->c);
    if (rc) {
        delete ctx;
        return rc;
    }
    ctx->batch = batch;
    hipDeviceProp_t prop;
    (void)hipGetDevice(&ctx->device);
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess)
        ctx->cus = prop.multiProcessorCount;
    set_grid(ctx);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return fail(nullptr, GSPH_ERR_HIP, "hipStreamCreate failed");
    }
    for (auto &e : ctx->stage_ev) (void)hipEventCreate(&e);
    *out = ctx;
    return GSPH_OK;
}

int build_context(const gsph_params *p, const gsph_options *opt,
                  const std::vector<SceneGeom> &scenes_in, int n, const float *pos4,
                  const float *vel3, bool batch, gsph_ctx **out) {
    if (!p || !out) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null argument");
    if (n < 0) return fail(nullptr, GSPH_ERR_INVALID_ARG, "negative particle count");
    if (n > 0 && !pos4) return fail(nullptr, GSPH_ERR_INVALID_ARG, "pos4 is null");
    gsph_ctx *ctx = nullptr;
    int rc = new_context(p, opt, batch, &ctx);
    if (rc) return rc;
    long long numc = 0;
    for (auto &g : scenes_in) numc += (long long)std::max(g.gx, 0) * std::max(g.gy, 0) * std::max(g.gz, 0);
    if (numc > (1LL << 26)) {
        gsph_destroy(ctx);
        return fail(nullptr, GSPH_ERR_CAPACITY, "more than 2^26 grid cells (micro ids must fit 32 bits)");
    }
    rc = alloc_buffers(ctx, n, (int)numc, (int)scenes_in.size());
    if (!rc) rc = configure_layout(ctx, scenes_in, n);
    if (!rc && n > 0) {
        hipError_t e = hipMemcpyAsync(ctx->pos[0], pos4, sizeof(float4) * (size_t)n,
                                      hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) {
            if (vel3)
                e = hipMemcpyAsync(ctx->vel[0], vel3, sizeof(f3) * (size_t)n,
                                   hipMemcpyHostToDevice, ctx->stream);
            else
                e = hipMemsetAsync(ctx->vel[0], 0, sizeof(f3) * (size_t)n, ctx->stream);
        }
        if (e == hipSuccess) {
            launch_init_state(ctx->ev[0], ctx->acc, n, ctx->stream);
            e = hipStreamSynchronize(ctx->stream);
            if (e == hipSuccess) e = hipGetLastError();
        }
        if (e != hipSuccess) rc = fail(ctx, GSPH_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
    }
    if (rc) {
        set_error(nullptr, ctx->err);
        gsph_destroy(ctx);
        return rc;
    }
    ctx->cur = 0;
    *out = ctx;
    return GSPH_OK;
}

SceneGeom geom_for(const gsph_constants &c, const float world[3], int p_begin, int p_count) {
    SceneGeom g;
    std::memset(&g, 0, sizeof(g));
    for (int a = 0; a < 3; ++a) {
        const float cells = world[a] / c.cell_size;
        const int gs = (int)std::ceil(cells);
        (&g.gx)[a] = gs;
        g.bound_min[a] = c.kernel_h;
        g.bound_max[a] = world[a] - c.kernel_h;
    }
    g.p_begin = p_begin;
    g.p_count = p_count;
    return g;
}

int capture_graph(gsph_ctx *ctx) {
    if (ctx->graph_exec) {
        (void)hipGraphExecDestroy(ctx->graph_exec);
        ctx->graph_exec = nullptr;
    }
    if (ctx->graph) {
        (void)hipGraphDestroy(ctx->graph);
        ctx->graph = nullptr;
    }
    HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const bool kt = ctx->ktiming;
    ctx->ktiming = false;
    int rc = do_step(ctx);
    ctx->ktiming = kt;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(ctx, GSPH_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
    ctx->graph = g;
    HIP_TRY(ctx, hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0));
    return GSPH_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Engine hooks of the serving loop (engine.h).
namespace {

// A pinned staging slot with at least `bytes` bytes whose previous upload has
// completed.
void *staging_slot(gsph_ctx *ctx, size_t bytes, hipEvent_t *done) {
    gsph_ctx::Staging &st = ctx->staging[ctx->staging_next];
    ctx->staging_next = (ctx->staging_next + 1) % 4;
    if (st.done) {
        if (hipEventSynchronize(st.done) != hipSuccess) return nullptr;
    } else if (hipEventCreateWithFlags(&st.done, hipEventDisableTiming) != hipSuccess) {
        st.done = nullptr;
        return nullptr;
    }
    if (st.bytes < bytes) {
        if (st.host) (void)hipHostFree(st.host);
        st.host = nullptr;
        st.bytes = 0;
        const size_t want = std::max<size_t>(bytes, 4096);
        if (hipHostMalloc(&st.host, want, hipHostMallocDefault) != hipSuccess) {
            st.host = nullptr;
            return nullptr;
        }
        st.bytes = want;
    }
    *done = st.done;
    return st.host;
}

// configure_layout without a host synchronisation: the scene table (and the
// re-layout segments) go up from a pinned staging slot.
int configure_layout_async(gsph_ctx *ctx, std::vector<SceneGeom> scenes, int n,
                           const std::vector<RelayoutSeg> &segs) {
    long long numc = 0, pb = 0;
    for (auto &g : scenes) {
        if (g.gx <= 0 || g.gy <= 0 || g.gz <= 0 || g.gx > 16383 || g.gy > 16383 || g.gz > 16383)
            return fail(ctx, GSPH_ERR_INVALID_ARG, "grid dimensions must be in [1, 16383]");
        g.cell_base = (int)numc;
        g.p_begin = (int)pb;
        numc += (long long)g.gx * g.gy * g.gz;
        pb += g.p_count;
    }
    if (numc > (1LL << 26))
        return fail(ctx, GSPH_ERR_CAPACITY, "more than 2^26 grid cells (micro ids must fit 32 bits)");
    if (pb != n) return fail(ctx, GSPH_ERR_INVALID_ARG, "scene counts do not add up");
    if (n > ctx->cap_n || numc > ctx->cap_c || (int)scenes.size() > ctx->cap_s ||
        (int)segs.size() > ctx->cap_s)
        return fail(ctx, GSPH_ERR_CAPACITY, "layout exceeds the context capacity");
    const size_t gb = sizeof(SceneGeom) * scenes.size(), sb = sizeof(RelayoutSeg) * segs.size();
    if (gb + sb > 0) {
        hipEvent_t done = nullptr;
        char *host = static_cast<char *>(staging_slot(ctx, gb + sb, &done));
        if (!host) return fail(ctx, GSPH_ERR_HIP, "pinned staging allocation failed");
        if (gb) std::memcpy(host, scenes.data(), gb);
        if (sb) std::memcpy(host + gb, segs.data(), sb);
        if (gb) HIP_TRY(ctx, hipMemcpyAsync(ctx->scenes_d, host, gb, hipMemcpyHostToDevice, ctx->stream));
        if (sb) HIP_TRY(ctx, hipMemcpyAsync(ctx->segs, host + gb, sb, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(ctx, hipEventRecord(done, ctx->stream));
    }
    if (gb)
        launch_scene_tabs(ctx->scenes_d, (int)scenes.size(), ctx->scene_tab, n, ctx->cell_scene_tab,
                          (int)numc, ctx->stream);
    ctx->scenes = scenes;
    ctx->n = n;
    ctx->numc = (int)numc;
    ctx->passes = (bits_for(std::max(numc * 64 - 1, 1LL)) + 7) / 8;
    ctx->sort_tiles = std::max(1, (n + kSortTile - 1) / kSortTile);
    ctx->cell_tiles = std::max(1, (n + kCellTile - 1) / kCellTile);
    ctx->nchunks = std::max(1, (ctx->numc + kFillChunk - 1) / kFillChunk);
    ctx->split_tiles = std::max(1, (n + kSplitTile - 1) / kSplitTile);
    ctx->task_cap = (n + 31) / 32 + std::min(ctx->numc, n) + 2;
    // a new layout: the fill writes every chunk again (a chunk's extent and
    // the meaning of its cells may have changed)
    HIP_TRY(ctx, hipMemsetAsync(ctx->fill_val, 0xff, sizeof(int) * (ctx->nchunks + 1), ctx->stream));
    if (ctx->graph_exec) {
        (void)hipGraphExecDestroy(ctx->graph_exec);
        ctx->graph_exec = nullptr;
    }
    if (ctx->graph) {
        (void)hipGraphDestroy(ctx->graph);
        ctx->graph = nullptr;
    }
    ctx->grid_valid = false;
    ctx->masks_valid = false;
    return GSPH_OK;
}

}  // namespace

namespace gsph {

size_t engine_bytes(const gsph_options &o, long long cap_n, long long cap_c, long long cap_s) {
    return buffer_bytes(o, (size_t)std::max(cap_n, 0LL), (size_t)std::max(cap_c, 0LL),
                        (size_t)std::max(cap_s, 1LL));
}

int engine_create(const gsph_params *p, const gsph_options *opt, int cap_n, int cap_c,
                  int cap_s, gsph_ctx **out) {
    gsph_ctx *ctx = nullptr;
    int rc = new_context(p, opt, true, &ctx);
    if (rc) return rc;
    rc = alloc_buffers(ctx, cap_n, cap_c, cap_s);
    if (!rc) rc = configure_layout(ctx, {}, 0);
    if (rc) {
        set_error(nullptr, ctx->err);
        gsph_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return GSPH_OK;
}

void engine_capacity(const gsph_ctx *ctx, int *cap_n, int *cap_c, int *cap_s) {
    *cap_n = ctx->cap_n;
    *cap_c = ctx->cap_c;
    *cap_s = ctx->cap_s;
}

size_t engine_allocated(const gsph_ctx *ctx) { return ctx->alloc_bytes; }

long long engine_cells(const gsph_ctx *ctx, const float world[3]) {
    const SceneGeom g = geom_for(ctx->c, world, 0, 0);
    return (long long)g.gx * g.gy * g.gz;
}

hipStream_t engine_stream(const gsph_ctx *ctx) { return ctx->stream; }

// Order-preserving re-layout.  Continuing scenes (old_begin >= 0) come first
// and in their current relative order: their slots are moved (one kernel)
// with their sorted keys shifted by the change of their first cell id, so the
// moved keys stay sorted and the next arrange sorts only the particles that
// change micro cell, as after any step.  New scenes follow; their state goes
// up from pinned host memory with keys that match no micro id (all movers).
// Nothing waits on the host unless dst != src (a grown engine), where the
// copy finishes before returning so that src can be destroyed.
int engine_relayout(gsph_ctx *src, gsph_ctx *dst, const std::vector<EngineScene> &scenes) {
    std::vector<SceneGeom> geo;
    std::vector<RelayoutSeg> segs;
    long long n = 0, cells = 0;
    int last_old = -1;
    bool new_seen = false;
    for (const auto &e : scenes) {
        SceneGeom g = geom_for(dst->c, e.world, (int)n, e.count);
        const long long gc = (long long)g.gx * g.gy * g.gz;
        if (e.old_begin >= 0) {
            if (new_seen || e.old_begin < last_old || !src)
                return fail(dst, GSPH_ERR_INVALID_ARG, "continuing scenes must come first, in order");
            last_old = e.old_begin;
            if (e.count > 0) {
                RelayoutSeg sg;
                sg.src = e.old_begin;
                sg.dst = (int)n;
                sg.count = e.count;
                sg.key_delta = (int)((cells - (long long)e.old_cell_base) * 64);
                segs.push_back(sg);
            }
        } else {
            new_seen = true;
        }
        geo.push_back(g);
        n += e.count;
        cells += gc;
    }
    if (n > dst->cap_n || cells > dst->cap_c || (long long)scenes.size() > dst->cap_s)
        return fail(dst, GSPH_ERR_CAPACITY, "engine capacity");
    if (src && src->cur != 0) return fail(dst, GSPH_ERR_INVALID_ARG, "re-layout needs integrated state");
    hipStream_t st = dst->stream;
    if (src && src != dst) {
        hipEvent_t e = nullptr;
        HIP_TRY(dst, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        hipError_t er = hipEventRecord(e, src->stream);
        if (er == hipSuccess) er = hipStreamWaitEvent(st, e, 0);
        (void)hipEventDestroy(e);
        if (er != hipSuccess) return fail(dst, GSPH_ERR_HIP, std::string("re-layout: ") + hipGetErrorString(er));
    }
    // the scene table and segments first (the kernel reads the segments)
    int rc = configure_layout_async(dst, geo, (int)n, segs);
    if (rc) return rc;
    long long cont = 0;
    for (const auto &sg : segs) cont = std::max<long long>(cont, (long long)sg.dst + sg.count);
    if (src)
        launch_relayout(src->pos[0], src->vel[0], src->ev[0], src->hash, dst->pos[1], dst->vel[1],
                        dst->ev[1], dst->skey, dst->segs, (int)segs.size(), (int)cont, st);
    rc = check_launch(dst, "relayout");
    if (rc) return rc;
    long long off = 0;
    for (const auto &e : scenes) {
        const size_t c = (size_t)e.count;
        if (c && e.old_begin < 0) {
            HIP_TRY(dst, hipMemcpyAsync(dst->pos[1] + off, e.pos4, 16 * c, hipMemcpyHostToDevice, st));
            if (e.vel3)
                HIP_TRY(dst, hipMemcpyAsync(dst->vel[1] + off, e.vel3, 12 * c, hipMemcpyHostToDevice, st));
            else
                HIP_TRY(dst, hipMemsetAsync(dst->vel[1] + off, 0, 12 * c, st));
            HIP_TRY(dst, hipMemsetAsync(dst->ev[1] + off, 0, 16 * c, st));
            HIP_TRY(dst, hipMemsetAsync(dst->skey + off, 0xff, 4 * c, st));
        }
        off += (long long)c;
    }
    std::swap(dst->pos[0], dst->pos[1]);
    std::swap(dst->vel[0], dst->vel[1]);
    std::swap(dst->ev[0], dst->ev[1]);
    std::swap(dst->hash, dst->skey);
    // the slot order is the moved keys' sorted order; the fused integration's
    // keys are for the old slots
    launch_set_order(dst->dc, 1, 0, st);
    rc = check_launch(dst, "relayout order");
    if (rc) return rc;
    dst->cur = 0;
    if (src && src != dst) HIP_TRY(dst, hipStreamSynchronize(st));
    return GSPH_OK;
}

int engine_scene_layout(const gsph_ctx *ctx, int s, int *first, int *cell_base) {
    if (s < 0 || s >= (int)ctx->scenes.size()) return GSPH_ERR_INVALID_ARG;
    *first = ctx->scenes[s].p_begin;
    *cell_base = ctx->scenes[s].cell_base;
    return GSPH_OK;
}

// Enqueues the copy of scene s's state (current slot order) into host memory
// (pinned for an asynchronous copy) and records `done` after it.
int engine_retire_async(gsph_ctx *ctx, int s, float *pos4, float *vel3, float *ev4,
                        hipEvent_t done) {
    if (s < 0 || s >= (int)ctx->scenes.size()) return fail(ctx, GSPH_ERR_INVALID_ARG, "scene index");
    if (ctx->cur != 0) return fail(ctx, GSPH_ERR_INVALID_ARG, "retire needs integrated state");
    const SceneGeom &g = ctx->scenes[s];
    const size_t c = (size_t)g.p_count;
    if (c) {
        HIP_TRY(ctx, hipMemcpyAsync(pos4, ctx->pos[0] + g.p_begin, 16 * c, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(vel3, ctx->vel[0] + g.p_begin, 12 * c, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(ev4, ctx->ev[0] + g.p_begin, 16 * c, hipMemcpyDeviceToHost, ctx->stream));
    }
    HIP_TRY(ctx, hipEventRecord(done, ctx->stream));
    return GSPH_OK;
}

// Enqueues the copy of the device error flags into pinned host memory, then
// clears them on the device (nothing else runs on the stream in between).
int engine_flags_async(gsph_ctx *ctx, int *flags_pinned) {
    launch_take_flags(ctx->dc, flags_pinned, ctx->stream);
    return check_launch(ctx, "flags");
}

// One step of every resident scene; with `graph` the step is replayed from
// a hipGraph captured on first use (dropped by the next re-layout).
int engine_step(gsph_ctx *ctx, bool graph) {
    if (!graph) return do_step(ctx);
    if (!ctx->graph_exec) {
        const int rc = capture_graph(ctx);
        if (rc) return rc;
    }
    HIP_TRY(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
    ++ctx->steps;
    ctx->cur = 0;
    ctx->grid_valid = false;
    ctx->masks_valid = false;
    return GSPH_OK;
}

int engine_read_scene(gsph_ctx *ctx, int s, float *pos4, float *vel3, float *ev4) {
    if (s < 0 || s >= (int)ctx->scenes.size()) return fail(ctx, GSPH_ERR_INVALID_ARG, "scene index");
    const SceneGeom &g = ctx->scenes[s];
    const size_t c = (size_t)g.p_count;
    const int b = ctx->cur;
    if (c) {
        if (pos4) HIP_TRY(ctx, hipMemcpyAsync(pos4, ctx->pos[b] + g.p_begin, 16 * c, hipMemcpyDeviceToHost, ctx->stream));
        if (vel3) HIP_TRY(ctx, hipMemcpyAsync(vel3, ctx->vel[b] + g.p_begin, 12 * c, hipMemcpyDeviceToHost, ctx->stream));
        if (ev4) HIP_TRY(ctx, hipMemcpyAsync(ev4, ctx->ev[b] + g.p_begin, 16 * c, hipMemcpyDeviceToHost, ctx->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return GSPH_OK;
}

}  // namespace gsph

// ---------------------------------------------------------------------------
extern "C" {

const char *gsph_last_error(const gsph_ctx *ctx) {
    return ctx ? ctx->err.c_str() : g_thread_error.c_str();
}

int gsph_device_count(int *count) {
    if (!count) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null argument");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return GSPH_OK;
}

int gsph_create(const gsph_params *p, const gsph_options *opt, int n, const float *pos4,
                const float *vel3, gsph_ctx **out) {
    if (!p) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null params");
    gsph_constants c;
    int rc = gsph_derive_constants(p, &c);
    if (rc) return rc;
    std::vector<SceneGeom> sc(1, geom_for(c, p->world_size, 0, n));
    rc = build_context(p, opt, sc, n, pos4, vel3, false, out);
    if (!rc) (*out)->worlds.push_back({p->world_size[0], p->world_size[1], p->world_size[2]});
    return rc;
}

int gsph_batch_create(const gsph_params *p, const gsph_options *opt, int nscenes,
                      const float *world3, const int *counts, const float *pos4,
                      const float *vel3, gsph_ctx **out) {
    if (!p || nscenes <= 0 || !world3 || !counts)
        return fail(nullptr, GSPH_ERR_INVALID_ARG, "bad batch arguments");
    gsph_constants c;
    int rc = gsph_derive_constants(p, &c);
    if (rc) return rc;
    std::vector<SceneGeom> sc;
    long long total = 0;
    for (int s = 0; s < nscenes; ++s) {
        if (counts[s] < 0) return fail(nullptr, GSPH_ERR_INVALID_ARG, "negative scene count");
        sc.push_back(geom_for(c, world3 + 3 * s, (int)total, counts[s]));
        total += counts[s];
        if (total > 0x7fffffffLL) return fail(nullptr, GSPH_ERR_CAPACITY, "too many particles");
    }
    rc = build_context(p, opt, sc, (int)total, pos4, vel3, true, out);
    if (!rc)
        for (int s = 0; s < nscenes; ++s)
            (*out)->worlds.push_back({world3[3 * s], world3[3 * s + 1], world3[3 * s + 2]});
    return rc;
}

int gsph_get_params(const gsph_ctx *ctx, gsph_params *out) {
    if (!ctx || !out) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null argument");
    *out = ctx->params;
    return GSPH_OK;
}

int gsph_scene_world(const gsph_ctx *ctx, int scene, float world[3]) {
    if (!ctx || !world) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null argument");
    if (ctx->worlds.size() != ctx->scenes.size())
        return fail(nullptr, GSPH_ERR_INVALID_ARG, "scene boxes are not kept by this context");
    if (scene < 0 || scene >= (int)ctx->worlds.size())
        return fail(nullptr, GSPH_ERR_INVALID_ARG, "bad scene index");
    for (int a = 0; a < 3; ++a) world[a] = ctx->worlds[scene][a];
    return GSPH_OK;
}

int gsph_is_batched(const gsph_ctx *ctx, int *batched) {
    if (!ctx || !batched) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null argument");
    *batched = ctx->batch ? 1 : 0;
    return GSPH_OK;
}

int gsph_step_count(const gsph_ctx *ctx, long long *steps) {
    if (!ctx || !steps) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null argument");
    *steps = ctx->steps;
    return GSPH_OK;
}

int gsph_set_step_count(gsph_ctx *ctx, long long steps) {
    if (!ctx || steps < 0) return fail(ctx, GSPH_ERR_INVALID_ARG, "bad arguments");
    ctx->steps = steps;
    return GSPH_OK;
}

int gsph_scene_count(const gsph_ctx *ctx, int *nscenes) {
    if (!ctx || !nscenes) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null argument");
    *nscenes = (int)ctx->scenes.size();
    return GSPH_OK;
}

int gsph_scene_range(const gsph_ctx *ctx, int s, int *first, int *count) {
    if (!ctx || !first || !count || s < 0 || s >= (int)ctx->scenes.size())
        return fail(nullptr, GSPH_ERR_INVALID_ARG, "bad scene index");
    *first = ctx->scenes[s].p_begin;
    *count = ctx->scenes[s].p_count;
    return GSPH_OK;
}

int gsph_destroy(gsph_ctx *ctx) {
    if (!ctx) return GSPH_OK;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    for (void *p : ctx->allocs) (void)hipFree(p);
    for (auto &st : ctx->staging) {
        if (st.done) (void)hipEventDestroy(st.done);
        if (st.host) (void)hipHostFree(st.host);
    }
    for (auto &e : ctx->stage_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &p : ctx->kt.pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : ctx->kt.pool) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return GSPH_OK;
}

int gsph_num_particles(const gsph_ctx *ctx, int *n) {
    if (!ctx || !n) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null argument");
    *n = ctx->n;
    return GSPH_OK;
}

int gsph_get_constants(const gsph_ctx *ctx, gsph_constants *out) {
    if (!ctx || !out) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null argument");
    *out = ctx->c;
    return GSPH_OK;
}

int gsph_set_state(gsph_ctx *ctx, const float *pos4, const float *vel3, const float *ev4) {
    if (!ctx || (ctx->n > 0 && !pos4)) return fail(ctx, GSPH_ERR_INVALID_ARG, "null argument");
    const size_t n = (size_t)ctx->n;
    if (n) {
        HIP_TRY(ctx, hipMemcpyAsync(ctx->pos[0], pos4, sizeof(float4) * n,
                                    hipMemcpyHostToDevice, ctx->stream));
        if (vel3)
            HIP_TRY(ctx, hipMemcpyAsync(ctx->vel[0], vel3, sizeof(f3) * n,
                                        hipMemcpyHostToDevice, ctx->stream));
        else
            HIP_TRY(ctx, hipMemsetAsync(ctx->vel[0], 0, sizeof(f3) * n, ctx->stream));
        if (ev4)
            HIP_TRY(ctx, hipMemcpyAsync(ctx->ev[0], ev4, sizeof(float4) * n,
                                        hipMemcpyHostToDevice, ctx->stream));
        else
            HIP_TRY(ctx, hipMemsetAsync(ctx->ev[0], 0, sizeof(float4) * n, ctx->stream));
    }
    int rc = invalidate_order(ctx);
    if (rc) return rc;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->cur = 0;
    ctx->grid_valid = false;
    ctx->masks_valid = false;
    return GSPH_OK;
}

namespace {

// Reallocates every buffer for cap particles and copies the resident state
// (pos, vel, ev, acc of slots [0, n)).  On failure the context keeps its old
// buffers untouched.
int grow_particles(gsph_ctx *ctx, int cap) {
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    const gsph_ctx saved = *ctx;   // host copy: pointers, sizes, bookkeeping
    ctx->allocs.clear();
    ctx->alloc_bytes = 0;
    int rc = alloc_buffers(ctx, cap, saved.cap_c, saved.cap_s);
    const size_t n = (size_t)saved.n;
    hipStream_t s = ctx->stream;
    if (!rc && n) {
        hipError_t e = hipMemcpyAsync(ctx->pos[0], saved.pos[0], 16 * n, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ctx->vel[0], saved.vel[0], 12 * n, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ctx->ev[0], saved.ev[0], 16 * n, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ctx->acc, saved.acc, 12 * n, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = fail(ctx, GSPH_ERR_HIP, std::string("grow: ") + hipGetErrorString(e));
    }
    if (rc) {
        (void)hipStreamSynchronize(s);
        for (void *p : ctx->allocs) (void)hipFree(p);
        const std::string err = ctx->err;
        *ctx = saved;
        ctx->err = err;
        return rc;
    }
    for (void *p : saved.allocs) (void)hipFree(p);
    // a captured step refers to the old buffers: never replay it
    if (ctx->graph_exec) {
        (void)hipGraphExecDestroy(ctx->graph_exec);
        ctx->graph_exec = nullptr;
    }
    if (ctx->graph) {
        (void)hipGraphDestroy(ctx->graph);
        ctx->graph = nullptr;
    }
    ctx->grid_valid = false;
    ctx->masks_valid = false;
    return GSPH_OK;
}

}  // namespace

int gsph_add_particles(gsph_ctx *ctx, int count, const float *pos4, const float *vel3,
                       int reset_all) {
    if (!ctx) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null ctx");
    if (ctx->batch)
        return fail(ctx, GSPH_ERR_INVALID_ARG, "particles are added to plain contexts only");
    if (count < 0 || (count > 0 && !pos4))
        return fail(ctx, GSPH_ERR_INVALID_ARG, "bad particle count or null pos4");
    if (ctx->cur != 0)
        return fail(ctx, GSPH_ERR_INVALID_ARG,
                    "particles are added between steps (call integrate first)");
    constexpr long long kMaxParticles = 0x7fffff00LL;
    const long long n_new = (long long)ctx->n + count;
    if (n_new > kMaxParticles) return fail(ctx, GSPH_ERR_CAPACITY, "too many particles");
    if (ctx->worlds.empty() || ctx->scenes.empty())
        return fail(ctx, GSPH_ERR_INVALID_ARG, "context has no scene box");
    const auto &w = ctx->worlds[0];
    std::vector<float> p4((size_t)count * 4);
    for (int k = 0; k < count; ++k) {
        for (int a = 0; a < 3; ++a) {
            const float x = pos4[4 * (size_t)k + a];
            if (!std::isfinite(x) || x < 0.f || x >= w[a])
                return fail(ctx, GSPH_ERR_OUT_OF_DOMAIN,
                            "particle " + std::to_string(k) + " is outside the box");
            p4[4 * (size_t)k + a] = x;
        }
        p4[4 * (size_t)k + 3] = 0.f;
    }
    // errors raised by earlier steps are reported before the counters are
    // replaced
    int rc = collect_errors(ctx);
    if (rc) return rc;
    if (n_new > ctx->cap_n) {
        const long long cap = std::min(kMaxParticles, std::max(n_new, 2LL * ctx->cap_n));
        rc = grow_particles(ctx, (int)cap);
        if (rc) return rc;
    }
    const size_t n = (size_t)ctx->n, c = (size_t)count;
    hipStream_t s = ctx->stream;
    if (c) {
        HIP_TRY(ctx, hipMemcpyAsync(ctx->pos[0] + n, p4.data(), 16 * c, hipMemcpyHostToDevice, s));
        if (vel3)
            HIP_TRY(ctx, hipMemcpyAsync(ctx->vel[0] + n, vel3, 12 * c, hipMemcpyHostToDevice, s));
        else
            HIP_TRY(ctx, hipMemsetAsync(ctx->vel[0] + n, 0, 12 * c, s));
        launch_init_state(ctx->ev[0] + n, ctx->acc + n, count, s);
    }
    if (reset_all && n_new > 0) launch_init_state(ctx->ev[0], ctx->acc, (int)n_new, s);
    rc = check_launch(ctx, "init_state");
    if (rc) return rc;
    SceneGeom g = ctx->scenes[0];
    g.p_count = (int)n_new;
    rc = configure_layout(ctx, {g}, (int)n_new);
    if (!rc) rc = invalidate_order(ctx);
    if (rc) return rc;
    HIP_TRY(ctx, hipStreamSynchronize(s));
    return GSPH_OK;
}

int gsph_arrange(gsph_ctx *ctx) {
    if (!ctx) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null ctx");
    int rc = consume_injection(ctx);
    if (rc) return rc;
    if (ctx->cur != 0)
        return fail(ctx, GSPH_ERR_INVALID_ARG,
                    "arrange expects the integrated state (call integrate first)");
    return do_arrange(ctx);
}

int gsph_density(gsph_ctx *ctx) {
    if (!ctx) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null ctx");
    int rc = consume_injection(ctx);
    if (rc) return rc;
    if (!ctx->grid_valid) return fail(ctx, GSPH_ERR_INVALID_ARG, "density needs arrange first");
    return do_density(ctx);
}

int gsph_force(gsph_ctx *ctx) {
    if (!ctx) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null ctx");
    int rc = consume_injection(ctx);
    if (rc) return rc;
    if (!ctx->grid_valid) return fail(ctx, GSPH_ERR_INVALID_ARG, "force needs arrange first");
    return do_force(ctx);
}

int gsph_integrate(gsph_ctx *ctx) {
    if (!ctx) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null ctx");
    int rc = consume_injection(ctx);
    if (rc) return rc;
    if (!ctx->grid_valid) return fail(ctx, GSPH_ERR_INVALID_ARG, "integrate needs arrange first");
    return do_integrate(ctx);
}

int gsph_step(gsph_ctx *ctx, int nsteps, gsph_stage_ms *per_step_ms) {
    if (!ctx || nsteps < 0) return fail(ctx, GSPH_ERR_INVALID_ARG, "bad arguments");
    int rc = consume_injection(ctx);
    if (rc) return rc;
    if (ctx->cur != 0)
        return fail(ctx, GSPH_ERR_INVALID_ARG, "step expects the integrated state");
    for (int s = 0; s < nsteps; ++s) {
        if (per_step_ms) {
            hipEvent_t *e = ctx->stage_ev;
            (void)hipEventRecord(e[0], ctx->stream);
            rc = do_arrange(ctx);
            (void)hipEventRecord(e[1], ctx->stream);
            if (!rc) rc = do_density(ctx);
            (void)hipEventRecord(e[2], ctx->stream);
            if (!rc) rc = do_force(ctx, true);
            (void)hipEventRecord(e[3], ctx->stream);
            if (!rc) rc = do_integrate(ctx, true);
            (void)hipEventRecord(e[4], ctx->stream);
            if (rc) return rc;
            HIP_TRY(ctx, hipEventSynchronize(e[4]));
            gsph_stage_ms &o = per_step_ms[s];
            (void)hipEventElapsedTime(&o.arrange, e[0], e[1]);
            (void)hipEventElapsedTime(&o.density, e[1], e[2]);
            (void)hipEventElapsedTime(&o.force, e[2], e[3]);
            (void)hipEventElapsedTime(&o.integrate, e[3], e[4]);
            (void)hipEventElapsedTime(&o.total, e[0], e[4]);
        } else if (ctx->use_graph && !ctx->ktiming) {
            if (!ctx->graph_exec) {
                rc = capture_graph(ctx);
                if (rc) return rc;
            }
            HIP_TRY(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
            ++ctx->steps;
            ctx->cur = 0;
            ctx->grid_valid = false;
            ctx->masks_valid = false;
        } else {
            rc = do_step(ctx);
            if (rc) return rc;
        }
    }
    return GSPH_OK;
}

namespace {

// Times `steps` steps of the context's scene from the saved state under
// options o (one untimed step first); restores the saved state afterwards.
int time_trial(gsph_ctx *ctx, const gsph_options &o, const float4 *pos0, const f3 *vel0,
               const float4 *ev0, int steps, float *ms) {
    ctx->opt = o;
    set_grid(ctx);
    const size_t n = (size_t)ctx->n;
    hipStream_t s = ctx->stream;
    auto restore = [&]() -> int {
        HIP_TRY(ctx, hipMemcpyAsync(ctx->pos[0], pos0, n * sizeof(float4), hipMemcpyDeviceToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->vel[0], vel0, n * sizeof(f3), hipMemcpyDeviceToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->ev[0], ev0, n * sizeof(float4), hipMemcpyDeviceToDevice, s));
        const int rc0 = invalidate_order(ctx);
        if (rc0) return rc0;
        ctx->cur = 0;
        ctx->grid_valid = false;
        ctx->masks_valid = false;
        return GSPH_OK;
    };
    int rc = restore();
    if (!rc) rc = do_step(ctx);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(ctx, GSPH_ERR_HIP, "hipEventCreate failed");
    if (!rc) {
        (void)hipEventRecord(e0, s);
        for (int k = 0; k < steps && !rc; ++k) rc = do_step(ctx);
        (void)hipEventRecord(e1, s);
        if (!rc && hipEventSynchronize(e1) != hipSuccess)
            rc = fail(ctx, GSPH_ERR_HIP, "trial synchronisation failed");
        if (!rc) (void)hipEventElapsedTime(ms, e0, e1);
        *ms /= (float)std::max(steps, 1);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (!rc) rc = collect_errors(ctx);
    const int rc2 = restore();
    return rc ? rc : rc2;
}

}  // namespace

int gsph_autotune(gsph_ctx *ctx, int trial_steps, int tune_split, gsph_options *chosen,
                  float *best_ms) {
    if (!ctx) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null ctx");
    if (ctx->cur != 0) return fail(ctx, GSPH_ERR_INVALID_ARG, "autotune expects the integrated state");
    const gsph_options start = ctx->opt;
    const long long steps0 = ctx->steps;
    if (ctx->n == 0) {
        if (chosen) *chosen = start;
        if (best_ms) *best_ms = 0.f;
        return GSPH_OK;
    }
    const int steps = std::min(std::max(trial_steps, 1), 100);
    const size_t n = (size_t)ctx->n;
    float4 *pos0 = nullptr, *ev0 = nullptr;
    f3 *vel0 = nullptr;
    HIP_TRY(ctx, hipMalloc(&pos0, n * sizeof(float4)));
    if (hipMalloc(&vel0, n * sizeof(f3)) != hipSuccess || hipMalloc(&ev0, n * sizeof(float4)) != hipSuccess) {
        (void)hipFree(pos0);
        if (vel0) (void)hipFree(vel0);
        return fail(ctx, GSPH_ERR_HIP, "hipMalloc of the autotune snapshot failed");
    }
    hipStream_t s = ctx->stream;
    int rc = GSPH_OK;
    if (hipMemcpyAsync(pos0, ctx->pos[0], n * sizeof(float4), hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipMemcpyAsync(vel0, ctx->vel[0], n * sizeof(f3), hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipMemcpyAsync(ev0, ctx->ev[0], n * sizeof(float4), hipMemcpyDeviceToDevice, s) != hipSuccess)
        rc = fail(ctx, GSPH_ERR_HIP, "autotune snapshot copy failed");
    // coordinate search: launch shape (grid x staging depth x pair masks)
    // under the current split, then the split rule under the best shape, then
    // the shape again if the split changed
    gsph_options best = start;
    float best_t = 1e30f;
    auto trial = [&](const gsph_options &o) {
        if (rc) return;
        float t = 0.f;
        rc = time_trial(ctx, o, pos0, vel0, ev0, steps, &t);
        if (!rc && t < best_t) {
            best_t = t;
            best = o;
        }
    };
    auto shape_search = [&](const gsph_options split) {   // a copy: trial() updates best
        for (int bpc : {4, 8, 16, 32})
            for (int depth : {64, 128, 256})
                for (int masks : {1, 0}) {
                    gsph_options o = split;
                    o.blocks_per_cu = bpc;
                    o.stage_depth = depth;
                    o.pair_masks = masks;
                    trial(o);
                }
    };
    shape_search(start);
    if (tune_split) {
        const gsph_options shape = best;
        gsph_options o = shape;
        o.split_mode = 0;
        trial(o);
        // tuned (n, 7-cell total) thresholds, the last one "never SMS" (every
        // cell on the TRA path, with pool masks for long rows)
        const int rules[][2] = {{15, 96}, {24, 128}, {32, 192}, {48, 256}, {64, 384},
                                {1 << 30, 1 << 30}};
        for (const auto &r : rules) {
            o = shape;
            o.split_mode = 1;
            o.sms_min_self = r[0];
            o.sms_min_total = r[1];
            trial(o);
        }
        // a different split moves work between the paths: tune the shape again
        if (best.split_mode != shape.split_mode || best.sms_min_self != shape.sms_min_self ||
            best.sms_min_total != shape.sms_min_total)
            shape_search(best);
    }
    {   // shared staging of isSame pairs on / off (no effect without SMS tasks)
        gsph_options o = best;
        o.sms_local_merge ^= 1;
        trial(o);
    }
    {   // the legacy SMS staging on / off (same results; no effect without SMS tasks)
        gsph_options o = best;
        o.sms_legacy_staging ^= 1;
        trial(o);
    }
    {   // the lean density/force passes on / off (same results)
        gsph_options o = best;
        o.lean_passes ^= 1;
        trial(o);
    }
    (void)hipFree(pos0);
    (void)hipFree(vel0);
    (void)hipFree(ev0);
    ctx->opt = rc ? start : best;
    ctx->steps = steps0;   // trials do not count as steps of the simulation
    set_grid(ctx);
    if (ctx->graph_exec) {   // captured with the old launch shape
        (void)hipGraphExecDestroy(ctx->graph_exec);
        ctx->graph_exec = nullptr;
    }
    if (rc) return rc;
    if (chosen) *chosen = best;
    if (best_ms) *best_ms = best_t;
    return GSPH_OK;
}

int gsph_set_options(gsph_ctx *ctx, const gsph_options *opt) {
    if (!ctx || !opt) return fail(ctx, GSPH_ERR_INVALID_ARG, "null argument");
    gsph_options o = *opt;
    int rc = check_options(ctx, o);
    if (rc) return rc;
    if (o.same_cell_pair_force && !ctx->scf && ctx->cap_n > 0) {
        // the same-cell sums buffer, allocated the first time the option is on
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        rc = dalloc(ctx, &ctx->scf, (size_t)10 * ctx->cap_n);
        if (rc) return rc;
    }
    ctx->opt = o;
    set_grid(ctx);
    // density's SMS masks depend on the staging variant: force tests its
    // candidates itself until the next density pass
    ctx->masks_valid = false;
    if (ctx->graph_exec) {
        (void)hipGraphExecDestroy(ctx->graph_exec);
        ctx->graph_exec = nullptr;
    }
    return GSPH_OK;
}

int gsph_get_options(gsph_ctx *ctx, gsph_options *out) {
    if (!ctx || !out) return fail(ctx, GSPH_ERR_INVALID_ARG, "null argument");
    *out = ctx->opt;
    return GSPH_OK;
}

int gsph_enable_graph(gsph_ctx *ctx, int enable) {
    if (!ctx) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null ctx");
    ctx->use_graph = enable != 0;
    return GSPH_OK;
}

int gsph_sync(gsph_ctx *ctx) {
    if (!ctx) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null ctx");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return collect_errors(ctx);
}

int gsph_read_state(gsph_ctx *ctx, float *pos4, float *vel3, float *ev4, float *acc3) {
    if (!ctx) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null ctx");
    const size_t n = (size_t)ctx->n;
    const int b = ctx->cur;
    if (n) {
        if (pos4)
            HIP_TRY(ctx, hipMemcpyAsync(pos4, ctx->pos[b], sizeof(float4) * n,
                                        hipMemcpyDeviceToHost, ctx->stream));
        if (vel3)
            HIP_TRY(ctx, hipMemcpyAsync(vel3, ctx->vel[b], sizeof(f3) * n,
                                        hipMemcpyDeviceToHost, ctx->stream));
        if (ev4)
            HIP_TRY(ctx, hipMemcpyAsync(ev4, ctx->ev[b], sizeof(float4) * n,
                                        hipMemcpyDeviceToHost, ctx->stream));
        if (acc3)
            HIP_TRY(ctx, hipMemcpyAsync(acc3, ctx->acc, sizeof(f3) * n,
                                        hipMemcpyDeviceToHost, ctx->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return collect_errors(ctx);
}

int gsph_grid_info(gsph_ctx *ctx, int *cells, int *tasks, int *middle_raw, int *tra) {
    if (!ctx) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null ctx");
    int dc[kCntSlots];
    HIP_TRY(ctx, hipMemcpyAsync(dc, ctx->dc, sizeof(dc), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (cells) *cells = ctx->numc;
    if (tasks) *tasks = dc[kCntT];
    if (middle_raw) *middle_raw = dc[kCntMiddle];
    if (tra) *tra = dc[kCntTra];
    return collect_errors(ctx);
}

int gsph_read_grid(gsph_ctx *ctx, int *hash, int *perm, int *begin, int *count, int *req,
                   int *compact, gsph_task *tasks) {
    if (!ctx) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null ctx");
    if (!ctx->grid_valid) return fail(ctx, GSPH_ERR_INVALID_ARG, "no arranged grid to read");
    const size_t n = (size_t)ctx->n, C = (size_t)ctx->numc;
    hipStream_t s = ctx->stream;
    int dc[kCntSlots];
    HIP_TRY(ctx, hipMemcpyAsync(dc, ctx->dc, sizeof(dc), hipMemcpyDeviceToHost, s));
    if (n && hash) HIP_TRY(ctx, hipMemcpyAsync(hash, ctx->hash, 4 * n, hipMemcpyDeviceToHost, s));
    if (n && perm) HIP_TRY(ctx, hipMemcpyAsync(perm, ctx->perm, 4 * n, hipMemcpyDeviceToHost, s));
    // counts are begin differences (no dense count array on the device)
    std::vector<int> bg_tmp;
    int *bg = nullptr;
    if (begin || count || req) {
        bg_tmp.resize(C + 1);
        bg = bg_tmp.data();
        HIP_TRY(ctx, hipMemcpyAsync(bg, ctx->begin, 4 * (C + 1), hipMemcpyDeviceToHost, s));
    }
    if (req) HIP_TRY(ctx, hipMemcpyAsync(req, ctx->req, 4 * C, hipMemcpyDeviceToHost, s));
    if (n && compact)
        HIP_TRY(ctx, hipMemcpyAsync(compact, ctx->compact, 4 * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    if (begin) std::copy(bg, bg + C, begin);
    if (count)
        for (size_t c = 0; c < C; ++c) count[c] = bg[c + 1] - bg[c];
    if (req)
        for (size_t c = 0; c < C; ++c)
            if (bg[c + 1] == bg[c]) req[c] = 0;  // the sparse table is only valid in occupied cells
    if (tasks) {
        const int T = dc[kCntT];
        const int nd = T + (T & 1);
        if (nd > 0)
            HIP_TRY(ctx, hipMemcpy(tasks, ctx->tasks, sizeof(gsph_task) * nd, hipMemcpyDeviceToHost));
    }
    return collect_errors(ctx);
}

int gsph_checksum_range(gsph_ctx *ctx, int first, int count, double out[5]) {
    if (!ctx || !out || first < 0 || count < 0 || first + count > ctx->n)
        return fail(ctx, GSPH_ERR_INVALID_ARG, "bad checksum range");
    launch_checksum(ctx->pos[ctx->cur] + first, count, ctx->sum_partial, 256, ctx->sum_out,
                    ctx->stream);
    int rc = check_launch(ctx, "checksum");
    if (rc) return rc;
    HIP_TRY(ctx, hipMemcpyAsync(out, ctx->sum_out, 5 * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return collect_errors(ctx);
}

int gsph_checksum_ranges_noalloc(gsph_ctx *ctx, int nranges, const int *first, const int *count,
                                 double *out5_device) {
    if (!ctx) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null ctx");
    if (nranges < 0 || (nranges > 0 && (!first || !count || !out5_device)))
        return fail(ctx, GSPH_ERR_INVALID_ARG, "bad ranges or output");
    if (nranges > kNoAllocRanges)
        return fail(ctx, GSPH_ERR_CAPACITY,
                    "more than " + std::to_string(kNoAllocRanges) +
                        " ranges per call (the size of the static scratch)");
    for (int k = 0; k < nranges; ++k)
        if (first[k] < 0 || count[k] < 0 || (long long)first[k] + count[k] > ctx->n)
            return fail(ctx, GSPH_ERR_INVALID_ARG, "bad checksum range " + std::to_string(k));
    int dev = -1;
    HIP_TRY(ctx, hipGetDevice(&dev));
    if (dev != ctx->device)
        return fail(ctx, GSPH_ERR_INVALID_ARG, "the context's device is not the current device");
    double *scr = nullptr;
    HIP_TRY(ctx, checksum_scratch(&scr));
    for (int k = 0; k < nranges; ++k)
        launch_checksum(ctx->pos[ctx->cur] + first[k], count[k],
                        scr + (size_t)k * kChecksumBlocks * 5, kChecksumBlocks, out5_device + 5 * k,
                        ctx->stream);
    return check_launch(ctx, "checksum");
}

int gsph_checksum(gsph_ctx *ctx, double out[5]) {
    if (!ctx) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null ctx");
    return gsph_checksum_range(ctx, 0, ctx->n, out);
}

int gsph_mask_pool_info(gsph_ctx *ctx, int *used_rows, int *capacity_rows) {
    if (!ctx) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null ctx");
    int dc[kCntSlots];
    HIP_TRY(ctx, hipMemcpyAsync(dc, ctx->dc, sizeof(dc), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (used_rows) *used_rows = std::min(dc[kCntSmsPool], ctx->sms_pool_rows);
    if (capacity_rows) *capacity_rows = ctx->sms_pool_rows;
    return GSPH_OK;
}

int gsph_get_pair_stats(gsph_ctx *ctx, gsph_pair_stats *out) {
    if (!ctx || !out) return fail(ctx, GSPH_ERR_INVALID_ARG, "null argument");
    std::memset(out, 0, sizeof(*out));
    int dc[kCntSlots];
    HIP_TRY(ctx, hipMemcpyAsync(dc, ctx->dc, sizeof(dc), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    const int T = dc[kCntT];
    if (T <= 0) return GSPH_OK;
    std::vector<gsph_task> tk(T + (T & 1));
    HIP_TRY(ctx, hipMemcpy(tk.data(), ctx->tasks, sizeof(gsph_task) * tk.size(),
                           hipMemcpyDeviceToHost));
    out->total_pairs = (T + 1) / 2;
    for (int t = 0; t < T; t += 2) {
        const gsph_task &f = tk[t], &s = tk[t + 1];
        if (f.is_same) {
            ++out->is_same_pairs;
            if (f.cell_particle_count - s.p_offset >= kSmsTaskParticles)
                ++out->full_is_same_pairs;
            else
                ++out->partial_is_same_pairs;
            if (f.xxi == s.xxx)
                ++out->compact_is_same_pairs;
            else
                ++out->widened_is_same_pairs;
        }
        if (f.cellid == s.cellid) ++out->same_cell_pairs;
    }
    return GSPH_OK;
}

int gsph_kernel_timing(gsph_ctx *ctx, int enable) {
    if (!ctx) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null ctx");
    drain_timer(ctx);
    ctx->ktiming = enable != 0;
    if (enable) {
        std::fill(ctx->kt.ms.begin(), ctx->kt.ms.end(), 0.0);
        std::fill(ctx->kt.launches.begin(), ctx->kt.launches.end(), 0);
    }
    return GSPH_OK;
}

int gsph_kernel_times(gsph_ctx *ctx, char *names, int names_cap, float *ms, int *launches,
                      int cap, int *count) {
    if (!ctx || !count) return fail(ctx, GSPH_ERR_INVALID_ARG, "null argument");
    drain_timer(ctx);
    const KernelTimer &kt = ctx->kt;
    *count = (int)kt.names.size();
    std::string joined;
    for (size_t i = 0; i < kt.names.size(); ++i) {
        if (i) joined += ';';
        joined += kt.names[i];
        if ((int)i < cap) {
            if (ms) ms[i] = (float)kt.ms[i];
            if (launches) launches[i] = kt.launches[i];
        }
    }
    if (names && names_cap > 0) {
        std::strncpy(names, joined.c_str(), (size_t)names_cap - 1);
        names[names_cap - 1] = 0;
    }
    return GSPH_OK;
}

int gsph_inject_fault(gsph_ctx *ctx, int after_calls) {
    if (!ctx) return fail(nullptr, GSPH_ERR_INVALID_ARG, "null ctx");
    ctx->inject_after = after_calls;
    return GSPH_OK;
}

}  // extern "C"
