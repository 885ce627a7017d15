// kernels.h - argument blocks and host launch wrappers of the HIP kernels.
// The context layer (context.cpp) fills these structs; arrange.hip and
// physics.hip define the kernels and the launchers declared here.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"
#include "gsph.h"

namespace gsph {

struct f3 {
    float x, y, z;
};

constexpr int kSortTile = 4096;     // keys per radix tile (256 threads x 16)
constexpr int kCellTile = 4096;     // sorted particles per cell-table tile
constexpr int kFillChunk = 2048;    // cells per fill block
constexpr int kSplitTile = 1024;    // occupied cells per split tile
constexpr int kMaskWords = 19;       // accepted-pair mask + row plan words per TRA particle

// One scene's move in a serving re-layout: slots [src, src + count) of the
// resident state go to [dst, dst + count), and its sorted keys shift by
// key_delta (the change of the scene's first cell id, times 64).
struct RelayoutSeg {
    int src, dst, count;
    int key_delta;
};

// Everything the arrange kernels touch.  Device pointers.
struct ArrangeArgs {
    int n;                 // particles
    int numc;              // cells of the (virtual) grid
    int nscenes;
    const SceneGeom *scenes;
    const int *scene_tab;       // [ceil(n / 64)] scene of particle 64 t (batched contexts)
    const int *cell_scene_tab;  // [ceil(numc / 64)] scene of cell 64 t
    float inv_cell_size;
    int split_mode;        // 0 reference heuristic, 1 tuned thresholds
    int sms_min_self, sms_min_total;
    // state: A = resident (input of the sort), B = sorted working copy
    const float4 *pos_a;
    const f3 *vel_a;
    const float4 *ev_a;
    float4 *pos_b;
    f3 *vel_b;
    float4 *ev_b;
    // sort by merging (k_hash .. k_merge)
    uint32_t *hash;        // [n] sorted keys; kept between steps (previous order)
    int *perm;             // [n] sorted position -> slot
    uint32_t *skey;        // [n] stayer list: key, by rank
    int *sidx;             //     slot, by rank
    uint32_t *mkey[2];     // [n] mover list keys (radix ping-pong; [1] also
                           //     holds new keys by slot, later the positions)
    int *mval[2];          // [n] mover list slots
    unsigned long long *mflag;   // [ceil(n/64)] mover flags, one bit per slot
    uint32_t *mpre;        // [ceil(n/64)] movers before each flag word
    uint32_t *tile_m;      // [sort_tiles] movers per slot tile
    uint32_t *tile_occ;    // [sort_tiles] movers per sorted tile
    uint32_t *dhist;       // [4][256] movers' digit counts, zero between steps
    unsigned long long *status;  // [4][msort_tiles][256] look-back words
    int msort_tiles;       // tiles of the mover radix passes (1024 movers each)
    int sample_stride;     // hash[] sample spacing of k_mplace
    int passes;            // 8-bit digits of the keys (at most 4)
    int sort_tiles;
    int cell_tiles;
    uint32_t *bcnt;        // [cell_tiles] boundary counts
    int *cells;            // [n+1] occupied cell ids
    int *cstart;           // [n+1] first sorted particle of each occupied cell
    int nchunks;
    int *chunk_k0;         // [nchunks] first occupied cell index of each chunk
    int *fill_val;         // [nchunks] the value every begin[] entry of a chunk holds
                           // when the last fill wrote it uniform (no occupied cell), else -1
    int *begin;            // [numc+1]; count of cell c = begin[c+1] - begin[c]
    uint2 *xcum;           // [numc] x-layer starts (valid where count > 0)
    int *req;              // [numc] valid where count > 0
    int *tra_base;         // [numc] valid where count > 0
    int *sms_base;         // [numc] valid where count > 0
    int split_tiles;
    uint32_t *split_sum;   // [3][split_tiles]
    int *compact;          // [n]
    gsph_task *tasks;      // [task_cap]
    int task_cap;
    int *dc;               // counters (CounterSlot)
};

// Density / force / integrate.
struct PhysicsArgs {
    int n;
    int nscenes;
    const SceneGeom *scenes;
    const int *scene_tab;  // [ceil(n / 64)] scene of particle 64 t (batched contexts)
    gsph_constants c;
    float4 *pos;           // sorted state (B)
    f3 *vel;
    float4 *ev;
    f3 *acc;
    float4 *pos_out;       // integrate destination (A)
    f3 *vel_out;
    float4 *ev_out;
    const uint32_t *hash;
    const int *begin;
    const uint2 *xcum;
    const int *compact;
    const gsph_task *tasks;
    const int *dc;
    int grid_blocks;       // persistent grid of density / force
    int row_prune;
    int sms_local_merge;   // isSame pairs of one cell stage their rows once
    int sms_legacy;        // SMS items stage z-pruned layer segments (legacy staging)
    // same-cell pair force: per-particle sums of the pairs inside its cell,
    // 10 floats (pressure xyz, viscosity xyz, gradient xyz, laplacian) with
    // stride scf_stride, written by k_same_cell for cells of SMS tasks with at
    // most kSameCellCap particles; null when the option is off
    float *scf;
    int scf_stride;
    // accepted-pair masks and row plans of TRA particles: kMaskWords words per
    // entry of compact, word-major with stride mask_stride; written by
    // density, read by force
    uint32_t *nbmask;
    int mask_stride;
    int use_masks;
    int lean;              // 1: the lean density/force passes run first (option lean_passes)
    int stage_depth;       // LDS staging depth of TRA rows: 64, 128 or 256
    // accepted-pair masks of SMS items: a pool of rows of 64 words (one per
    // lane); density takes W rows per item (W = 32-candidate sub-batches of
    // its longest half) with an atomic on dc[kCntSmsPool] and records the
    // first row in sms_off[item] (-1: pool full, force tests that item)
    uint32_t *sms_pool;
    int sms_pool_rows;
    int *sms_off;          // [task_cap / 2 + 1]
    uint8_t *tile_defer;   // [ceil(n / 64)] 1 for the TRA tiles the lean density pass
                           // leaves to the general one (written every density pass)
    int *dcw;              // writable alias of dc
    float wall_margin;     // the integrator's wall margin h + 0.01: the reference adds
                           // the double literal to the float h and stores the sum as a
                           // float (integration_kernels.cu:256)
    int fuse_integrate;    // force also integrates (the full step); else k_integrate
    uint32_t *next_keys;   // [n] with fuse_integrate: micro id of each integrated particle
};

// arrange.hip
void launch_hash(const ArrangeArgs &a, bool batch, hipStream_t s);
void launch_movers(const ArrangeArgs &a, hipStream_t s);
void launch_msort(const ArrangeArgs &a, int pass, hipStream_t s);
void launch_mplace(const ArrangeArgs &a, hipStream_t s);
void launch_merge(const ArrangeArgs &a, hipStream_t s);
int sample_stride(int n);
int msort_tiles(int n);
void launch_cells(const ArrangeArgs &a, hipStream_t s);
void launch_fill(const ArrangeArgs &a, hipStream_t s);
void launch_split_count(const ArrangeArgs &a, bool batch, hipStream_t s);
void launch_split_apply(const ArrangeArgs &a, bool batch, hipStream_t s);
void launch_compact(const ArrangeArgs &a, hipStream_t s);
void launch_empty_grid(const ArrangeArgs &a, hipStream_t s);
// Serving re-layout: the first `total` slots of the destination state are the
// continuing scenes' moved slots (segments sorted by dst, nseg of them),
// keys shifted per segment; order flags set for the next arrange (the slot
// order is the sorted order of the moved keys; new scenes carry keys that
// never match, so they are sorted as movers).
void launch_relayout(const float4 *pos_s, const f3 *vel_s, const float4 *ev_s,
                     const uint32_t *hash_s, float4 *pos_d, f3 *vel_d, float4 *ev_d,
                     uint32_t *hash_d, const RelayoutSeg *segs, int nseg, int total,
                     hipStream_t s);
void launch_set_order(int *dc, int order, int keys_ready, hipStream_t s);
void launch_take_flags(int *dc, int *out_host, hipStream_t s);

// physics.hip
void launch_density(const PhysicsArgs &p, bool batch, hipStream_t s);
void launch_force(const PhysicsArgs &p, bool batch, hipStream_t s);
void launch_integrate(const PhysicsArgs &p, bool batch, hipStream_t s);
void launch_scene_tabs(const SceneGeom *scenes, int nscenes, int *tab, int n, int *ctab,
                       int numc, hipStream_t s);
constexpr int kSameCellCap = 512;      // particles of a cell the same-cell pass takes
void launch_same_cell(const PhysicsArgs &p, hipStream_t s);
constexpr int kChecksumBlocks = 256;   // blocks of the FP64 checksum's first pass
constexpr int kNoAllocRanges = 64;     // ranges per call of the no-allocation checksum
constexpr int kMaxDevices = 64;
// Device address (current device) of the no-allocation checksum's static
// scratch: kNoAllocRanges * kChecksumBlocks * 5 doubles.
hipError_t checksum_scratch(double **ptr);
void launch_checksum(const float4 *pos, int n, double *partial, int nblocks,
                     double *out5, hipStream_t s);
void launch_init_state(float4 *ev, f3 *acc, int n, hipStream_t s);

}  // namespace gsph
