// physics.hip - density, force and integration kernels of one SPH step on
// gfx950, plus the FP64 state checksum.
//
// Work items of the density and force kernels (one 64-lane wave each):
//   TRA tile  - 64 consecutive entries of compact[0, tra): one particle per
//               lane, each lane walks its own 9 neighbour rows (the reference's
//               TRA path, density_kernels.cu:323-367, force_kernels.cu:749-818);
//   SMS pair  - tasks 2q and 2q+1: each half-wave owns one 32-particle task and
//               streams its rows through LDS 32 candidates at a time, every lane
//               testing all staged candidates against its particle (the
//               reference's register SMS path, density_kernels.cu:370-545).
// A row (dy, dz) is the sorted range from the left cell's micro-x layer L to
// the right cell's layer R (findTraNeighborParticleRange,
// kernel_common.cuh:69-112); cells x-1, x, x+1 of a row are contiguous in
// sorted order.  With row_prune on, a row is skipped when its (y, z) slab lies
// farther than h from the particle, and L/R shrink to the x reach left by the
// slab distance: a superset of the interacting set, smaller than the
// reference's.
//
// Grid: a persistent grid; item i goes to XCD (i * 8 / items) so that the
// waves sharing one L2 work on one contiguous eighth of space, round-robin
// among that XCD's waves.
#include "kernels.h"

#include <algorithm>
#include <mutex>

// No implicit fusion of a * b + c anywhere in this file: the pair bodies are
// inlined into several paths (TRA rows, SMS batches, pool rounds, fused and
// separate integration), and each copy must round exactly like the others
// whatever the compiler does with the code around it.  The fused
// multiply-adds the kernels want are written as fmaf.
#pragma clang fp contract(off)

namespace gsph {
namespace {

constexpr int kNT = 256;

__device__ __forceinline__ float powf7(float b) {
    return b * b * b * b * b * b * b;
}

template <bool kBatch>
__device__ __forceinline__ int scene_of_particle(const SceneGeom *sc, int ns,
                                                 int i) {
    if (!kBatch) return 0;
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
        const int m = (lo + hi + 1) >> 1;
        if (sc[m].p_begin <= i)
            lo = m;
        else
            hi = m - 1;
    }
    return lo;
}

// The scene of sorted particle i in a batched context: the table (built by
// launch_scene_tabs with each layout) gives the scene of particle
// 64 (i / 64), and the scenes that begin between there and i (none, but where
// a scene starts inside those 64) are stepped over.  The same answer as
// scene_of_particle's search, in two loads instead of one per halving of the
// scene count.
template <bool kBatch>
__device__ __forceinline__ int scene_of(const PhysicsArgs &P, int i) {
    if (!kBatch) return 0;
    int s = P.scene_tab[i >> 6];
    while (s + 1 < P.nscenes && P.scenes[s + 1].p_begin <= i) ++s;
    return s;
}

// #particles of a cell (begin b0, next cell's begin b1, x-layer starts x)
// whose micro x layer is < L (L in 0..4).  x is only meaningful for an
// occupied cell.
__device__ __forceinline__ int layer_count(int b0, int b1, uint2 x, int L) {
    const int cnt = b1 - b0;
    const int v = L == 1 ? (int)(x.x & 0xffffu) : L == 2 ? (int)(x.x >> 16) : (int)x.y;
    return L <= 0 ? 0 : L >= 4 ? cnt : cnt == 0 ? 0 : v;
}

// Sorted ranges [lo, hi) of nine rows: plan(r, ny, nz, gl, gr) gives row r's
// cell row (ny, nz), its micro-x layers [gl, gr] (in range) and whether it is
// live; rows that are not live, or all rows when !on, get lo = hi = 0.  A row
// is the left cell's layer gl to the right cell's layer gr
// (findTraNeighborParticleRange, kernel_common.cuh:69-112).  The loads of a
// batch of rows are all issued before any is used (no branch between them):
// begin of the left and right cells with the next begin, and the one x-layer
// start each side needs; rows that are not live read a clamped in-range cell.
// Three batches of three rows bound the registers the loads hold, and each
// batch's bounds are formed before the next batch's loads issue, so that its
// loaded words are dead by then.
template <typename Plan>
__device__ __forceinline__ void row_ranges(const SceneGeom &g, const int *begin, const uint2 *xcum,
                                           bool on, Plan plan, int lo[9], int hi[9]) {
    const uint32_t *xw = reinterpret_cast<const uint32_t *>(xcum);
    auto batch = [&](int r0, int r1) __attribute__((always_inline)) {
        int bl0[3], bl1[3], br0[3], br1[3], sel[3];
        uint32_t xl[3], xr[3];
#pragma unroll
        for (int r = r0; r < r1; ++r) {
            const int k = r - r0;
            int ny, nz, gl, gr;
            const bool live = plan(r, ny, nz, gl, gr);
            ny = min(max(ny, 0), g.gy - 1);
            nz = min(max(nz, 0), g.gz - 1);
            const int rowc = g.cell_base + g.gx * (ny + g.gy * nz);
            const int lc = rowc + (gl >> 2), rc = rowc + (gr >> 2);
            const int Ll = gl & 3, Lr = (gr & 3) + 1;
            bl0[k] = begin[lc];
            bl1[k] = begin[lc + 1];
            br0[k] = begin[rc];
            br1[k] = begin[rc + 1];
            xl[k] = xw[2 * lc + (Ll == 3 ? 1 : 0)];
            xr[k] = xw[2 * rc + (Lr == 3 ? 1 : 0)];
            // bit-field of the layer start each side reads (offset, width):
            // layer L = 1: bits 0-15 of word 0, 2: bits 16-31, 3: word 1 (its
            // low 31 bits: v_bfe takes the width mod 32); L = 0 (left) reads
            // nothing, L = 4 (right) takes the next begin
            const int offl = Ll == 2 ? 16 : 0, wl = Ll == 0 ? 0 : Ll == 3 ? 31 : 16;
            const int offr = Lr == 2 ? 16 : 0, wr = Lr == 3 ? 31 : 16;
            sel[k] = offl | (wl << 5) | (offr << 11) | (wr << 16) | (Lr == 4 ? 1 << 22 : 0) |
                     ((on && live) ? 1 << 23 : 0);
        }
#pragma unroll
        for (int r = r0; r < r1; ++r) {
            const int k = r - r0;
            const uint32_t f = (uint32_t)sel[k];
            const int vl = (int)__builtin_amdgcn_ubfe(xl[k], f & 31u, (f >> 5) & 63u);
            const int vr = (int)__builtin_amdgcn_ubfe(xr[k], (f >> 11) & 31u, (f >> 16) & 63u);
            const int l = bl0[k] + (bl1[k] == bl0[k] ? 0 : vl);
            const int h = (f >> 22) & 1u ? br1[k] : br0[k] + (br1[k] == br0[k] ? 0 : vr);
            const bool live = (f >> 23) & 1u;
            lo[r] = live ? l : 0;
            hi[r] = live ? h : 0;
        }
    };
    batch(0, 3);
    asm volatile("" ::"v"(lo[0]), "v"(hi[0]), "v"(lo[1]), "v"(hi[1]), "v"(lo[2]), "v"(hi[2])
                 : "memory");
    batch(3, 6);
    asm volatile("" ::"v"(lo[3]), "v"(hi[3]), "v"(lo[4]), "v"(hi[4]), "v"(lo[5]), "v"(hi[5])
                 : "memory");
    batch(6, 9);
}

// Row r (dy = r % 3 - 1, dz = r / 3 - 1) of a particle at (x, y, z) in
// micro cell (mx, my, mz): cell row (ny, nz) and micro-x layers [gl, gr];
// returns whether the row is live.  Reference rows: layers [mx - 4, mx + 4].
// Pruned rows: the slab distance d of the row in (y, z) gives the x reach
// sqrt(h^2 - d^2); layers outside [x - reach, x + reach] are dropped (with a
// small relative margin, so that a pair the FP32 test accepts is never
// pruned).
struct RowSlack {
    float sx, sy, sz;
};

// Rounding slack of the slab distances and the reach.  The hash puts a cell
// boundary where fl(y * rat) crosses a multiple of 4, which differs from
// fl(cy * h) by a few ulp of y; fl(x -+ reach) rounds by half an ulp of x.
// 2^-20 of the coordinate (8 ulp) plus the fixed margins of plan_row (1e-3
// relative on h^2, 1e-3 on the reach, 1e-4 h) keep the pruned set a superset
// of the pairs the FP32 test accepts for every grid up to 16383 cells per
// axis.
__device__ __forceinline__ RowSlack row_slack(float x, float y, float z, float h) {
    return {(fabsf(x) + h) * 0x1p-20f, (fabsf(y) + h) * 0x1p-20f, (fabsf(z) + h) * 0x1p-20f};
}

__device__ __forceinline__ bool plan_row(const SceneGeom &g, const gsph_constants &c, float x,
                                         float y, float z, int mx, int my, int mz, bool prune,
                                         RowSlack sl, int r, int &ny, int &nz, int &gl, int &gr) {
    const int cy = my >> 2, cz = mz >> 2;
    const float h = c.cell_size;
    const float rat = 4.f * c.inv_cell_size;
    const int dy = r % 3 - 1, dz = r / 3 - 1;
    ny = cy + dy;
    nz = cz + dz;
    bool live = ny >= 0 && ny < g.gy && nz >= 0 && nz < g.gz;
    gl = mx - 4;
    gr = mx + 4;
    if (prune) {   // uniform; selects below, so the row loads can follow at once
        const float ey = dy < 0 ? y - (float)cy * h - sl.sy : dy > 0 ? (float)(cy + 1) * h - y - sl.sy : 0.f;
        const float ez = dz < 0 ? z - (float)cz * h - sl.sz : dz > 0 ? (float)(cz + 1) * h - z - sl.sz : 0.f;
        const float d2 = fmaxf(ey, 0.f) * fmaxf(ey, 0.f) + fmaxf(ez, 0.f) * fmaxf(ez, 0.f);
        const float lim = c.kernel_2 * 1.001f;
        // the hardware square root (about 1 ulp) is well inside the margins
        const float reach = __builtin_amdgcn_sqrtf(fmaxf(lim - d2, 0.f)) * 1.001f + 1e-4f * h + sl.sx;
        live = live && !(d2 > lim);
        gl = max(gl, (int)floorf((x - reach) * rat));
        gr = min(gr, (int)floorf((x + reach) * rat));
    }
    gl = max(gl, 0);
    gr = min(gr, g.gx * 4 - 1);
    return live && gl <= gr;
}

struct ItemMap {
    int first, stride, end;
};

// XCD-aware static item schedule of the persistent grid.
__device__ __forceinline__ ItemMap item_map(int items) {
    const int nb = gridDim.x;
    const int xcd = blockIdx.x & 7;
    const int per_xcd_blocks = (nb + 7 - xcd) >> 3;      // blocks on this XCD
    const int local = blockIdx.x >> 3;
    const int waves_here = per_xcd_blocks * (kNT / 64);
    const int lo = (int)(((long long)items * xcd) >> 3);
    const int hi = (int)(((long long)items * (xcd + 1)) >> 3);
    ItemMap m;
    m.first = lo + local * (kNT / 64) + (int)(threadIdx.x >> 6);
    m.stride = waves_here;
    m.end = hi;
    return m;
}

// Wave-wide minimum with DPP lane moves (quad permutes, row mirrors, row
// broadcasts 15/31), read back from lane 63: no LDS round trips.
__device__ __forceinline__ int wave_min(int v) {
    constexpr int kMax = 0x7fffffff;
    v = min(v, __builtin_amdgcn_update_dpp(kMax, v, 0xb1, 0xf, 0xf, false));  // quad [1,0,3,2]
    v = min(v, __builtin_amdgcn_update_dpp(kMax, v, 0x4e, 0xf, 0xf, false));  // quad [2,3,0,1]
    v = min(v, __builtin_amdgcn_update_dpp(kMax, v, 0x141, 0xf, 0xf, false)); // row half mirror
    v = min(v, __builtin_amdgcn_update_dpp(kMax, v, 0x140, 0xf, 0xf, false)); // row mirror
    v = min(v, __builtin_amdgcn_update_dpp(kMax, v, 0x142, 0xa, 0xf, false)); // row bcast 15
    v = min(v, __builtin_amdgcn_update_dpp(kMax, v, 0x143, 0xc, 0xf, false)); // row bcast 31
    return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ int wave_max_neg(int v) {   // max of v >= -2^31 + 1, kNone-free
    return -wave_min(-v);
}

// Rows of the mask pool for one wave (lane 0 takes them with one atomic;
// -1 when the pool cannot hold them).  No increment once the pool is full, so
// the counter cannot run past it by more than one row set per wave.
__device__ __forceinline__ int pool_take(const PhysicsArgs &P, int rows, int cap, int lane) {
    int b = -1;
    if (lane == 0) {
        const int limit = cap - rows;
        if (rows > 0 && limit >= 0 &&
            __hip_atomic_load(&P.dcw[kCntSmsPool], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <=
                limit) {
            b = atomicAdd(&P.dcw[kCntSmsPool], rows);
            if (b < 0 || b > limit) b = -1;
        }
    }
    return __builtin_amdgcn_readfirstlane(b);
}

// candidates staged per wave and round: a launch option (64, 128 or 256)
constexpr int kStageDefault = 128;
constexpr int kNone = 0x7fffffff;

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The pair test, evaluated on the density kernel's own variable
// t = h^2 - r^2: one FMA chain from h^2, t = fl(fl(fl(h^2 - dx^2) - dy^2) -
// dz^2), and the pair is accepted when 0 <= t < h^2, one unsigned compare of
// bit patterns (non-negative floats order like their bits; a negative t, -0
// and NaN carry the sign bit and fall outside).  This is the reference's
// 1e-12 <= r^2 <= h^2 (density_kernels.cu:166-171) up to FP32 rounding:
//  - a pair within about an ulp of h^2 of the kernel edge may fall on the
//    other side; its weight (h^2 - r^2)^3 is below 1e-20 of the largest;
//  - t == h^2 exactly (the particle itself, coincident particles, and pairs
//    whose squared components each lie below half an ulp of h^2) is rejected
//    where the reference rejects r^2 < 1e-12.
// Density sums t^3 straight from the test's t.  Every kernel evaluates the
// test with these roundings (force consumes density's accepted pairs).
struct PairTest {
    float h2;
    uint32_t bits;   // bit pattern of h2
};

__device__ __forceinline__ PairTest pair_test(float k2) { return {k2, __float_as_uint(k2)}; }

__device__ __forceinline__ float kernel_t(const float4 &a, const float4 &b, float h2) {
    const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return fmaf(-dz, dz, fmaf(-dy, dy, fmaf(-dx, dx, h2)));
}

__device__ __forceinline__ bool accept_t(float t, PairTest pt) {
    return __float_as_uint(t) < pt.bits;
}

__device__ __forceinline__ bool accept(const float4 &a, const float4 &b, PairTest pt) {
    return accept_t(kernel_t(a, b, pt.h2), pt);
}

// Accepted-pair masks and row plan of TRA particles (entry idx of compact,
// word-major with stride mask_stride).  Word r < 9 holds bit k for candidate
// lo_r + k of row r (rows of at most 32 candidates); word 9 has bit 0 set
// when some row is longer (force then tests that particle itself) and bit 1
// when words 10..18 hold the row plan, so that force does not plan the rows
// again: word 10 + r = (lo_r - i) as 24-bit two's complement | length << 24,
// i the particle's sorted index (stored only when every row fits; an empty
// row is stored as 0).  kMaskWords words per entry in all.
constexpr int kMaskRowMax = 32;
// bit 2 (kFlagPool): the tile has a row longer than one word and keeps its
// masks in the mask pool instead, one word per 32 positions of each staging
// round (rounds in walk order; rows of 64 words, one per lane); the pool row
// of the tile's first word is the flag word >> 3
constexpr uint32_t kFlagLongRow = 1u, kFlagPlan = 2u, kFlagPool = 4u;
constexpr int kPoolBaseMax = 1 << 29;

__device__ __forceinline__ uint32_t finish_mask(uint32_t shifted, int cnt) {
    // bits entered at the bottom in candidate order; reverse into bit k
    return cnt > 0 ? __builtin_bitreverse32(shifted) >> (32 - cnt) : 0u;
}

// The accept mask of cnt candidates from density_pair's rejection bits.
__device__ __forceinline__ uint32_t finish_rejected(uint32_t rej, int cnt) {
    return finish_mask(~rej, cnt);
}

// One SMS work item: tasks t0 = 2q and t0 + 1, one per half-wave.  When both
// are tasks of one cell (a judgeTask isSame pair) their rows are the same
// cell rows, and the wave stages the union of their x-layer ranges once, 64
// candidates per batch.  The layers only one task needs lie more than 4
// layers (h) beyond every particle of the other task, so each lane still
// accepts exactly its own pairs, in the same index order.
struct SmsItem {
    bool has_task, active, shared;
    int self, si, gidx, cx, cy, cz, gl, gr;
    int cell_begin, cell_count;   // the task's cell
    int zmin, zmax;               // micro-z bounds of the staged particles (legacy staging)
};

template <bool kBatch>
__device__ __forceinline__ SmsItem sms_item(const PhysicsArgs &P, int t0, int T, int lane) {
    SmsItem m;
    const int task = t0 + (lane >> 5);
    m.has_task = task < T;
    gsph_task tk;
    if (m.has_task) tk = P.tasks[task];
    const int cell_end = m.has_task ? tk.cell_begin + tk.cell_particle_count : 0;
    m.self = m.has_task ? tk.cell_begin + tk.p_offset + (lane & 31) : 0;
    m.active = m.has_task && m.self < cell_end;
    m.si = m.has_task ? min(m.self, cell_end - 1) : 0;
    m.gidx = m.has_task ? scene_of<kBatch>(P, tk.cell_begin) : 0;
    m.cell_begin = m.has_task ? tk.cell_begin : 0;
    m.cell_count = m.has_task ? tk.cell_particle_count : 0;
    m.cx = m.has_task ? tk.cell_pos[0] : 0;
    m.cy = m.has_task ? tk.cell_pos[1] : 0;
    m.cz = m.has_task ? tk.cell_pos[2] : 0;
    int xi = m.has_task ? tk.xxi : 0, xa = m.has_task ? tk.xxx : 0;
    // the legacy staging always shares an isSame pair's rows (its iterator
    // runs over the pair's paired_bounds, kernel_common.cuh:289, 992-1001)
    m.shared = (P.sms_local_merge || P.sms_legacy) && t0 + 1 < T &&
               P.tasks[t0].cellid == P.tasks[t0 + 1].cellid;
    int pb = m.has_task ? tk.paired_bounds : 0;
    if (m.shared) {
        xi = min((int)P.tasks[t0].xxi, (int)P.tasks[t0 + 1].xxi);
        xa = max((int)P.tasks[t0].xxx, (int)P.tasks[t0 + 1].xxx);
        pb = P.tasks[t0].paired_bounds;   // judgeTask's z bounds of the pair
    }
    m.zmin = (pb >> 8) & 3;
    m.zmax = (pb >> 10) & 3;
    m.gl = 4 * m.cx - 4 + xi;
    m.gr = 4 * m.cx + 4 + xa;
    return m;
}

// Legacy SMS staging (the reference's GSPH_USE_REGISTER_SMS=0 path: the
// shared-memory iterators SimDenSharedData128 / SimForSharedData128,
// kernel_common.cuh:282-519 and 886-1151).  A row (dy, dz) below the task's
// cell (dz = -1) whose x-pruned range holds more than 22 candidates per
// micro-x layer is read layer by layer, each layer from micro z = zmin on;
// one above it (dz = +1), each layer up to micro z = zmax.  The slices left
// out lie five or more micro layers (more than h) away in z from every
// particle the item stages.  Within a layer the keys ascend in micro z, so
// a binary search over the sorted keys gives the cut (no dense micro table).
// Each segment is staged in batches of its own, as the reference's read32Data
// reads within one segment per batch.  Writes the segments of row r to tab
// ([lo, hi) pairs, one per layer, by lanes sl < layers) and returns their
// number; an unpruned row is one segment.  Uniform over the lanes that share
// the item's staging (the half-wave, or the wave when shared).
constexpr int kLegacyPerLayer = 22;

// The nine rows of an SMS task: the task's cell row +-1 in y and z, micro-x
// layers [gl, gr] of the task's (or the merged pair's) particles.
__device__ __forceinline__ void sms_rows(const SceneGeom &g, const PhysicsArgs &P, const SmsItem &sm,
                                         int lo[9], int hi[9]) {
    row_ranges(g, P.begin, P.xcum, sm.has_task,
               [&](int r, int &ny, int &nz, int &gl, int &gr) __attribute__((always_inline)) {
                   ny = sm.cy + r % 3 - 1;
                   nz = sm.cz + r / 3 - 1;
                   gl = max(sm.gl, 0);
                   gr = min(sm.gr, g.gx * 4 - 1);
                   return ny >= 0 && ny < g.gy && nz >= 0 && nz < g.gz;
               },
               lo, hi);
}

// Layers of row r when the legacy staging cuts it, else 0.
__device__ __forceinline__ int legacy_cut_layers(const SceneGeom &g, const SmsItem &sm, int r,
                                                 int lo, int hi) {
    const int dz = r / 3 - 1;
    const int layers = min(sm.gr, g.gx * 4 - 1) - max(sm.gl, 0) + 1;
    const bool zcut = (dz < 0 && sm.zmin > 0) || (dz > 0 && sm.zmax < 3);
    return zcut && hi - lo > kLegacyPerLayer * layers ? layers : 0;
}

__device__ __forceinline__ int legacy_segments(const PhysicsArgs &P, const SceneGeom &g,
                                               const SmsItem &sm, int r, int lo, int hi, int sl,
                                               int *tab) {
    const int dz = r / 3 - 1;
    const int gl = max(sm.gl, 0);
    const int layers = legacy_cut_layers(g, sm, r, lo, hi);
    wave_sync_lds();   // the previous row's reads of tab precede these writes
    if (layers == 0) {
        if (sl == 0) {
            tab[0] = lo;
            tab[1] = hi;
        }
        wave_sync_lds();
        return 1;
    }
    if (sl < layers) {
        const int ny = sm.cy + r % 3 - 1, nz = sm.cz + dz;   // a live row (hi > lo)
        const int gx = gl + sl;
        const int c = g.cell_base + (gx >> 2) + g.gx * (ny + g.gy * nz);
        const int b0 = P.begin[c], b1 = P.begin[c + 1];
        const uint2 xc = P.xcum[c];
        int a = b0 + layer_count(b0, b1, xc, gx & 3);
        int e = b0 + layer_count(b0, b1, xc, (gx & 3) + 1);
        // first key of the layer with micro z >= zt
        const int zt = dz < 0 ? sm.zmin : sm.zmax + 1;
        int u = a, v = e;
        while (u < v) {
            const int mid = (u + v) >> 1;
            if ((int)((P.hash[mid] >> 2) & 3u) < zt)
                u = mid + 1;
            else
                v = mid;
        }
        if (dz < 0)
            a = u;
        else
            e = u;
        tab[2 * sl] = a;
        tab[2 * sl + 1] = e;
    }
    wave_sync_lds();
    return layers;
}

// ---------------------------------------------------------------------------
// Density.  rho = m*poly6 * sum (h^2 - r^2)^3 + self_density over
// 1e-12 <= r^2 <= h^2; w = 1/rho into pos.w, p = ((rho/rho0)^7 - 1)*gas into
// ev.w (density_kernels.cu:353-358, 536-544).
// One candidate of the density sum: rs += (h^2 - r^2)^3 if the pair is
// accepted (a rejected one adds an exact zero), and its acceptance bit is
// shifted into mk.
// One density pair: rs += t^3 for an accepted pair; rej collects one bit per
// candidate (bits enter at the bottom), 1 for a rejected pair.  A rejected
// pair's t is -0, which adds nothing (the sums start at +0 and never become
// -0) and whose sign bit is the rejection bit, shifted in with one
// v_alignbit; an accepted t is non-negative.
__device__ __forceinline__ void density_pair(const float4 &a, const float4 &b, PairTest pt,
                                             float &rs, uint32_t &rej) {
    const float t0 = kernel_t(a, b, pt.h2);
    const float t = accept_t(t0, pt) ? t0 : -0.f;
    rs = fmaf(t * t, t, rs);
    rej = __builtin_amdgcn_alignbit(rej, __float_as_uint(t), 31);
}

__device__ __forceinline__ void density_store(const PhysicsArgs &P, int i, float s) {
    const gsph_constants &c = P.c;
    const float rho = s * c.density_scale + c.self_density;
    const float stored = rho < kFloatSmall ? c.rest_density : rho;
    reinterpret_cast<float *>(&P.pos[i])[3] = 1.0f / stored;
    reinterpret_cast<float *>(&P.ev[i])[3] =
        (powf7(rho * c.inv_rest_density) - 1.f) * c.gas_constant;
}

// The rows of one TRA lane, a particle whose scene g and micro coordinates
// are known: the sorted range per row (empty when the row is not live).
__device__ __forceinline__ void tra_rows_at(const PhysicsArgs &P, const float4 &pi, bool active,
                                            const SceneGeom &g, int mx, int my, int mz,
                                            int lo[9], int hi[9]) {
    const RowSlack sl = row_slack(pi.x, pi.y, pi.z, P.c.cell_size);
    const bool prune = P.row_prune != 0;
    row_ranges(g, P.begin, P.xcum, active,
               [&](int r, int &ny, int &nz, int &gl, int &gr) __attribute__((always_inline)) {
                   return plan_row(g, P.c, pi.x, pi.y, pi.z, mx, my, mz, prune, sl, r, ny, nz, gl, gr);
               },
               lo, hi);
}

// The same from the particle's index i (scene lookup and micro coordinates
// here), optionally with its scene's first particle and its cell's count.
template <bool kBatch>
__device__ __forceinline__ void tra_rows(const PhysicsArgs &P, const float4 &pi, bool active,
                                         int i, int lo[9], int hi[9], int *p_begin = nullptr,
                                         int *cell_count = nullptr) {
    const SceneGeom g = P.scenes[scene_of<kBatch>(P, i)];
    if (p_begin) *p_begin = g.p_begin;
    int mx, my, mz;
    micro_coords(pi.x, pi.y, pi.z, P.c.inv_cell_size, g, mx, my, mz);
    if (cell_count) {
        const int c = g.cell_base + (mx >> 2) + g.gx * ((my >> 2) + g.gy * (mz >> 2));
        *cell_count = P.begin[c + 1] - P.begin[c];
    }
    tra_rows_at(P, pi, active, g, mx, my, mz, lo, hi);
}

// The order in which every path visits a particle's 9 rows (r = 3 (dz + 1)
// + dy + 1): mirrored pairs (0, 8), (1, 7), (2, 6), (3, 5), then the centre
// row 4.  A particle near one face of its cell has many pairs in the rows on
// that side and few in the mirrored ones, so a pair's total varies much less
// from lane to lane than one row's count: the staged force walk takes a pair of
// rows as one per-lane loop.  Density sums each row apart and adds the row
// sums in this order; force accumulates its pairs in this order; the TRA and
// SMS paths both follow it, so a particle gets the same bits on either path.
__device__ __forceinline__ constexpr int row_of(int s) {
    return (s & 1) ? 8 - (s >> 1) : (s >> 1);
}

// Union span of each row over the tile's lanes: [L[r], L[r] + S[r]) in the
// sorted order (S = 0 when no lane has a candidate in the row).  Wave-uniform.
__device__ __forceinline__ void row_spans(const int lo[9], const int hi[9], int L[9], int S[9]) {
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const bool live = lo[r] < hi[r];
        const int lmin = wave_min(live ? lo[r] : kNone);
        const int hmax = lmin == kNone ? 0 : wave_max_neg(live ? hi[r] : lmin);
        L[r] = lmin == kNone ? 0 : lmin;
        S[r] = lmin == kNone ? 0 : hmax - lmin;
    }
}

// Copies src[L, L + S) into LDS at dst with LDS-direct loads
// (global_load_lds_dwordx4: lane l of a wave writes dst + 16 l; no registers
// are involved).  dst is wave-uniform; completion is tracked by vmcnt, so the
// caller waits with s_waitcnt before reading dst.
__device__ __forceinline__ void dma_span(const float4 *src, int L, int S, float4 *dst, int lane) {
    for (int c = 0; c < S; c += 64) {
        if (c + lane < S)
            __builtin_amdgcn_global_load_lds((const void *)(src + L + c + lane),
                                             (__attribute__((address_space(3))) void *)(dst + c),
                                             16, 0, 0);
    }
}

__device__ __forceinline__ void wait_dma() {
    __builtin_amdgcn_s_waitcnt(0);   // vmcnt, expcnt, lgkmcnt all zero
    wave_sync_lds();
}

// Upper bound of a TRA tile's staging rounds (next_round below): each live
// row's rounds start at the wave's lowest lo and advance by at least kStage
// while some lane's range reaches past the round, so a row takes at most
// ceil((max hi - min lo) / kStage).  Wave-uniform; all lanes must call it.
template <int kStage>
__device__ __forceinline__ int tra_round_bound(const int lo[9], const int hi[9]) {
    int rounds = 0;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const bool live = lo[r] < hi[r];
        const int lmin = wave_min(live ? lo[r] : kNone);
        if (lmin != kNone) {
            const int hmax = wave_max_neg(live ? hi[r] : lmin);
            rounds += (hmax - lmin + kStage - 1) / kStage;
        }
    }
    return rounds;
}

// All 9 rows of a TRA tile walked as one sequence of staging rounds (rows in
// the canonical order row_of(0..8); within a row, the lanes' ranges [lo, hi)
// are covered by rounds of kStage consecutive particles loaded into LDS with
// coalesced loads, gaps between lanes' ranges skipped; each lane reads its
// own candidates from LDS in index order), software-pipelined by the callers:
// while round t is visited, the loads of round t + 1 (possibly the next
// row's first) are in flight.  Rows empty for every lane are skipped.  The
// row bounds come from LDS (rlo/rhi: row r of lane l at r * 64 + l), so the
// wave-uniform row index never indexes a register array.  This walk serves
// the tiles the staged-row path (stage_rows below) cannot take: rows longer
// than its buffers, and tiles whose masks live in the pool.
//
// Advances (s, cur) to the next round (s: position in the row order);
// cur = kNone when all rows are done.
template <int kStage>
__device__ __forceinline__ void next_round(int &s, int &cur, int &lo_r, int &hi_r,
                                           const int *rlo, const int *rhi, int lane) {
    int nxt = kNone;
    // only lanes with candidates in the row steer the walk, so the rounds
    // depend on the live ranges alone (density and force must walk the same
    // rounds: the pool masks are indexed by round).  The next round of the
    // row starts right after this one when some lane has candidates there,
    // else at the lowest candidate left (past a gap between the lanes'
    // ranges); two ballots settle the common cases (the row is done, or it
    // goes on at once) without a wave-wide minimum.
    if (s >= 0) {
        const int nb = cur + kStage;
        const bool more = lo_r < hi_r && hi_r > nb;
        if (__ballot(more) == 0)
            nxt = kNone;
        else if (__ballot(more && lo_r < nb + kStage) != 0)
            nxt = nb;
        else
            nxt = wave_min(more ? max(lo_r, nb) : kNone);
    }
    while (nxt == kNone && s < 8) {
        ++s;
        const int r = row_of(s);
        lo_r = rlo[r * 64 + lane];
        hi_r = rhi[r * 64 + lane];
        nxt = wave_min(lo_r < hi_r ? lo_r : kNone);
    }
    cur = nxt;
}

// One staging round in registers: entries lane + 64 u, u < kStage / 64.
template <int kStage>
__device__ __forceinline__ void load_round(const float4 *src, int base, int n, int lane,
                                           float4 &v0, float4 &v1, float4 &v2, float4 &v3) {
    const int k = base + lane;
    if (k < n) v0 = src[k];
    if (kStage > 64 && k + 64 < n) v1 = src[k + 64];
    if (kStage > 128 && k + 128 < n) v2 = src[k + 128];
    if (kStage > 128 && k + 192 < n) v3 = src[k + 192];
}

template <int kStage>
__device__ __forceinline__ void store_round(float4 *dst, int lane, const float4 &v0,
                                            const float4 &v1, const float4 &v2,
                                            const float4 &v3) {
    dst[lane] = v0;
    if (kStage > 64) dst[lane + 64] = v1;
    if (kStage > 128) {
        dst[lane + 128] = v2;
        dst[lane + 192] = v3;
    }
}

// LDS of the density kernel, per wave: the round walk's row bounds and
// staging round, or an SMS item's batch (64 entries).
template <int kStage>
struct DensLds {
    static constexpr int kWords = 2 * 9 * 64 / 4 + kStage;    // float4s: rows + round
};

// Per-row density sums.  A particle in a cell of at most kFourChainCell
// particles sums each row in one sequential chain (the reference's order).
// In denser cells a row holds ~100+ candidates (~130 at the reference
// scene's density), and one chain lets the FP32 rounding error grow with that
// length, which the 7th power of the equation of state amplifies 7x in the
// pressure; there the row is summed in four interleaved chains: the candidate
// with place j in its scene (sorted index minus the scene's first) goes to
// chain j & 3, and the row's sum is (a0 + a1) + (a2 + a3).  Four chains a
// quarter as long keep density and pressure within 1e-5 of FP64 at the
// reference scene's density.  Both the choice (the particle's own cell
// count) and the chain of a candidate depend on the particle and the
// candidate alone, and rejected candidates add exact zeros, so every path
// (TRA windows of any depth, pruned or not, SMS batches, pool sub-windows) and
// every place of the scene in a batch gets the same bits.
constexpr int kFourChainCell = 48;

struct RowSum {
    float s1;                  // one chain
    float a0, a1, a2, a3;      // four chains
};

__device__ __forceinline__ RowSum row_sum_zero() { return {0.f, 0.f, 0.f, 0.f, 0.f}; }

__device__ __forceinline__ float row_sum_value(const RowSum &r, bool four) {
    return four ? (r.a0 + r.a1) + (r.a2 + r.a3) : r.s1;
}

// (b0..b3) = (a_q, a_{q+1}, a_{q+2}, a_{q+3}) indices mod 4 (q = 0..3, per
// lane): two conditional swaps
__device__ __forceinline__ void chains_rotate(float &a0, float &a1, float &a2, float &a3, int q) {
    if (q & 2) {
        const float t0 = a0, t1 = a1;
        a0 = a2;
        a1 = a3;
        a2 = t0;
        a3 = t1;
    }
    if (q & 1) {
        const float t = a0;
        a0 = a1;
        a1 = a2;
        a2 = a3;
        a3 = t;
    }
}

// Candidates q[0, cnt) of a row, the first at place off of its scene;
// rejection bits mk (in candidate order, entering at the bottom; see
// density_pair).
__device__ __forceinline__ void density_row(const float4 &pi, const float4 *q, int cnt, int off,
                                            bool four, PairTest pt, RowSum &rs,
                                            uint32_t &mk) {
    if (!four) {
        int k = 0;
        for (; k + 4 <= cnt; k += 4) {
#pragma unroll
            for (int u = 0; u < 4; ++u) density_pair(pi, q[k + u], pt, rs.s1, mk);
        }
        for (; k < cnt; ++k) density_pair(pi, q[k], pt, rs.s1, mk);
    } else {
        const int r = off & 3;
        float a0 = rs.a0, a1 = rs.a1, a2 = rs.a2, a3 = rs.a3;
        chains_rotate(a0, a1, a2, a3, r);
        int k = 0;
        for (; k + 4 <= cnt; k += 4) {
            density_pair(pi, q[k], pt, a0, mk);
            density_pair(pi, q[k + 1], pt, a1, mk);
            density_pair(pi, q[k + 2], pt, a2, mk);
            density_pair(pi, q[k + 3], pt, a3, mk);
        }
        if (k < cnt) density_pair(pi, q[k], pt, a0, mk);
        if (k + 1 < cnt) density_pair(pi, q[k + 1], pt, a1, mk);
        if (k + 2 < cnt) density_pair(pi, q[k + 2], pt, a2, mk);
        chains_rotate(a0, a1, a2, a3, (4 - r) & 3);
        rs.a0 = a0;
        rs.a1 = a1;
        rs.a2 = a2;
        rs.a3 = a3;
    }
}

// The particle of TRA entry e: compact[e], which is e itself when no cell
// has SMS tasks (ident: every particle on the TRA path, in sorted order;
// k_compact then writes the identity), so that the load is skipped there.
// Kernels read the condition once (tra_ident) and pass it to their tiles.
__device__ __forceinline__ bool tra_ident(const PhysicsArgs &P) { return P.dc[kCntSms] == 0; }

__device__ __forceinline__ int tra_particle(const PhysicsArgs &P, int e, bool ident) {
    return ident ? e : P.compact[e];
}

// A tile the lean pass leaves: flagged for the general pass, which takes
// TRA tiles at all only in an arrange epoch where some tile was left (the
// epoch is stored once, not by every such tile).
__device__ __forceinline__ void defer_tile(const PhysicsArgs &P, int it, int lane) {
    if (lane == 0) {
        P.tile_defer[it] = 1;
        const int ep = P.dc[kCntEpoch];
        if (*reinterpret_cast<const volatile int *>(&P.dc[kCntAnyDefer]) != ep) P.dcw[kCntAnyDefer] = ep;
    }
}

// One TRA tile of the density pass: 64 consecutive entries of compact[0,
// tra), one particle per lane.  kLean (the lean pass) takes a tile only when
// no lane sums in four chains and, with pair masks on, no row exceeds one
// mask word (no pool rounds); it marks every tile it leaves in tile_defer for
// the general pass.  Without those paths the lean pass fits 80
// registers (six waves per SIMD instead of five); both give the same bits.
template <bool kBatch, int kStage, bool kLean>
__device__ __forceinline__ void density_tile(const PhysicsArgs &P, int it, int tra, int lane,
                                             float4 *wl, PairTest pt, bool ident) {
    // round walk: row bounds (ints) then the staging round
    int *rlo = reinterpret_cast<int *>(wl), *rhi = rlo + 9 * 64;
    float4 *stg = wl + 2 * 9 * 64 / 4;
        // all lanes stay in the wave (the row walk is cooperative);
        // lanes past the TRA range carry empty rows
        const int idx = it * 64 + lane;
        const bool active = idx < tra;
        const int i = tra_particle(P, active ? idx : tra - 1, ident);
        const float4 pi = P.pos[i];
        int lo[9], hi[9], pb = 0, own = 0;
        if (kLean) {
            // a lane in a dense cell (four chains) leaves the tile to the
            // general pass; its cell's count is loaded ahead of the row
            // lookups and tested after them, so the rows wait for no extra
            // round trip
            int mx, my, mz;
            const SceneGeom g = P.scenes[scene_of<kBatch>(P, i)];
            micro_coords(pi.x, pi.y, pi.z, P.c.inv_cell_size, g, mx, my, mz);
            const int c = g.cell_base + (mx >> 2) + g.gx * ((my >> 2) + g.gy * (mz >> 2));
            const int own0 = P.begin[c], own1 = P.begin[c + 1];
            tra_rows_at(P, pi, active, g, mx, my, mz, lo, hi);
            if (__ballot(active && own1 - own0 > kFourChainCell) != 0) {
                defer_tile(P, it, lane);
                return;
            }
        } else {
            tra_rows<kBatch>(P, pi, active, i, lo, hi, &pb, &own);
        }
        bool lng = false;
#pragma unroll
        for (int r = 0; r < 9; ++r) lng |= active && hi[r] - lo[r] > kMaskRowMax;
        const bool any_long = __ballot(lng) != 0;
        if (kLean) {
            if (P.use_masks && any_long) {   // (four-chain lanes left above)
                defer_tile(P, it, lane);
                return;
            }
            if (lane == 0) P.tile_defer[it] = 0;
        }
        const bool four = !kLean && own > kFourChainCell;
        float s = 0.f;
        bool long_row = false;
        int pbase = -1;
        bool pool = false;
        {
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                rlo[r * 64 + lane] = lo[r];
                rhi[r * 64 + lane] = hi[r];
            }
            // a tile with a row longer than one mask word keeps its masks in
            // the pool: rows for an upper bound of its staging rounds (each
            // row's rounds advance by at least kStage over the wave's span)
            int pool_rounds = 0;
            if (!kLean && P.use_masks && any_long) {
                pool_rounds = tra_round_bound<kStage>(lo, hi);
                pbase = pool_take(P, pool_rounds * (kStage / 32), min(P.sms_pool_rows, kPoolBaseMax),
                                  lane);
            }
            pool = !kLean && pbase >= 0;
            int round = 0;
            // per-row partial sums (rejected candidates add exact zeros, so
            // pruning cannot change the result) and per-row rejection bits
            RowSum rs = row_sum_zero();
            uint32_t mk = 0;
            int done = 0;   // rows at order positions [0, done) are finished
            auto finish_rows = [&](int upto) __attribute__((always_inline)) {
                for (; done < upto; ++done) {
                    s += row_sum_value(rs, four);
                    rs = row_sum_zero();
                    if (P.use_masks && !pool) {
                        const int r = row_of(done);
                        const int cnt = rhi[r * 64 + lane] - rlo[r * 64 + lane];
                        long_row |= cnt > kMaskRowMax;
                        if (active)
                            P.nbmask[(size_t)r * P.mask_stride + idx] =
                                cnt > kMaskRowMax ? 0u : finish_rejected(mk, cnt);
                    }
                    mk = 0;
                }
            };
            float4 v0 = pi, v1 = pi, v2 = pi, v3 = pi;
            int q = -1, cur = kNone, lo_r = 0, hi_r = 0;
            next_round<kStage>(q, cur, lo_r, hi_r, rlo, rhi, lane);
            if (cur != kNone) load_round<kStage>(P.pos, cur, P.n, lane, v0, v1, v2, v3);
            while (cur != kNone) {
                const int qq = q, base = cur, lo_c = lo_r, hi_c = hi_r;
                store_round<kStage>(stg, lane, v0, v1, v2, v3);
                wave_sync_lds();
                next_round<kStage>(q, cur, lo_r, hi_r, rlo, rhi, lane);
                if (cur != kNone) load_round<kStage>(P.pos, cur, P.n, lane, v0, v1, v2, v3);
                finish_rows(qq);
                const int jb = max(lo_c, base), je = min(hi_c, base + kStage);
                const int cnt = je - jb;
                if (pool) {
                    // the same candidates in the same order, 32 window
                    // positions per mask word
#pragma unroll
                    for (int c = 0; c < kStage / 32; ++c) {
                        const int cb = max(jb, base + 32 * c), ce = min(je, base + 32 * c + 32);
                        const int nc = ce - cb;
                        uint32_t w = 0;
                        if (nc > 0) {
                            density_row(pi, stg + (cb - base), nc, cb - pb, four, pt, rs, w);
                            w = finish_rejected(w, nc) << (cb - base - 32 * c);
                        }
                        if (round < pool_rounds)   // always, by the bound
                            P.sms_pool[(size_t)(pbase + round * (kStage / 32) + c) * 64 + lane] = w;
                    }
                    ++round;
                } else if (cnt > 0) {
                    density_row(pi, stg + (jb - base), cnt, jb - pb, four, pt, rs, mk);
                }
                wave_sync_lds();
            }
            finish_rows(9);
        }
        if (active) {
            density_store(P, i, s);
            if (P.use_masks) {
                bool fits_plan = true;
#pragma unroll
                for (int r = 0; r < 9; ++r) {
                    const int len = hi[r] - lo[r], off = len > 0 ? lo[r] - i : 0;
                    fits_plan &= len <= 255 && off >= -(1 << 23) && off < (1 << 23);
                    P.nbmask[(size_t)(10 + r) * P.mask_stride + idx] =
                        ((uint32_t)off & 0xffffffu) | ((uint32_t)len << 24);
                }
                P.nbmask[(size_t)9 * P.mask_stride + idx] =
                    pool ? (kFlagPool | ((uint32_t)pbase << 3) | (fits_plan ? kFlagPlan : 0u))
                         : ((long_row ? kFlagLongRow : 0u) | (fits_plan ? kFlagPlan : 0u));
            }
        }
}

// The lean density pass: TRA tiles without four-chain lanes or pool rounds.
template <bool kBatch, int kStage>
__global__ __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(kStage > 128 ? 4 : 6))) void k_density_lean(PhysicsArgs P) {
    __shared__ float4 s_lds[kNT / 64][DensLds<kStage>::kWords];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float4 *wl = s_lds[wave];
    const int tra = P.dc[kCntTra];
    const ItemMap m = item_map((tra + 63) >> 6);
    const PairTest pt = pair_test(P.c.kernel_2);
    const bool ident = tra_ident(P);
    for (int it = m.first; it < m.end; it += m.stride) {
        wave_sync_lds();   // the previous item's LDS reads precede this item's writes
        density_tile<kBatch, kStage, true>(P, it, tra, lane, wl, pt, ident);
    }
}

// kLegacy: SMS items use the legacy staging (P.sms_legacy), a variant of its
// own so that the default kernels carry none of its registers
template <bool kBatch, int kStage, bool kLegacy>
__global__ __launch_bounds__(kNT) void k_density(PhysicsArgs P) {
    __shared__ float4 s_lds[kNT / 64][DensLds<kStage>::kWords];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float4 *wl = s_lds[wave];
    const int tra = P.dc[kCntTra];
    const int T = P.dc[kCntT];
    // without the lean pass every TRA tile; with it, only when it left some
    // in this arrange epoch
    const int tra_tiles = !P.lean || P.dc[kCntAnyDefer] == P.dc[kCntEpoch] ? (tra + 63) >> 6 : 0;
    const int items = tra_tiles + ((T + 1) >> 1);
    const ItemMap m = item_map(items);
    const PairTest pt = pair_test(P.c.kernel_2);
    const bool ident = tra_ident(P);
    for (int it = m.first; it < m.end; it += m.stride) {
        wave_sync_lds();   // the previous item's LDS reads precede this item's writes
        if (it < tra_tiles) {
            // the tiles the lean pass left (tile_defer, written by it)
            if (!P.lean || P.tile_defer[it]) density_tile<kBatch, kStage, false>(P, it, tra, lane, wl, pt, ident);
        } else {
            const int item = it - tra_tiles;
            const SmsItem sm = sms_item<kBatch>(P, item << 1, T, lane);
            const float4 pi = P.pos[sm.si];
            const SceneGeom g = P.scenes[sm.gidx];
            const bool four = sm.cell_count > kFourChainCell;
            const int width = sm.shared ? 64 : 32;
            const int sl = sm.shared ? lane : (lane & 31);
            float4 *lds = sm.shared ? wl : wl + (lane & 32);
            int lo[9], hi[9], subs = 0;
            sms_rows(g, P, sm, lo, hi);
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                subs += (hi[r] - lo[r] + 31) >> 5;
                // a cut row's segments each start a sub-batch: at most one more
                // per layer (an upper bound; force walks the same sub-batches)
                if (kLegacy) subs += legacy_cut_layers(g, sm, r, lo[r], hi[r]);
            }
            // pool rows of this item: the longer half's 32-candidate sub-batches
            const int rows = max(__builtin_amdgcn_readlane(subs, 0),
                                 __builtin_amdgcn_readlane(subs, 32));
            int base = -1;
            if (P.use_masks) {
                int b = -1;
                if (lane == 0) {
                    // no increment once the pool is full, so the counter
                    // cannot run past it by more than one row set per wave
                    const int limit = P.sms_pool_rows - rows;
                    if (rows > 0 && limit >= 0 &&
                        __hip_atomic_load(&P.dcw[kCntSmsPool], __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT) <= limit) {
                        b = atomicAdd(&P.dcw[kCntSmsPool], rows);
                        if (b < 0 || b > limit) b = -1;
                    }
                    P.sms_off[item] = b;
                }
                base = __builtin_amdgcn_readfirstlane(b);
            }
            float s = 0.f;
            int w = 0;
            int *tab = reinterpret_cast<int *>(wl + 64) + (sm.shared ? 0 : (lane & 32));
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                const int r = row_of(q);
                RowSum rs = row_sum_zero();
                const int nseg = kLegacy ? legacy_segments(P, g, sm, r, lo[r], hi[r], sl, tab) : 1;
                for (int sg = 0; sg < nseg; ++sg) {
                    const int slo = kLegacy ? tab[2 * sg] : lo[r];
                    const int shi = kLegacy ? tab[2 * sg + 1] : hi[r];
                    // the next batch's candidate is loaded while this one computes
                    float4 nx = pi;
                    if (sl < min(width, shi - slo)) nx = P.pos[slo + sl];
                    for (int b0 = slo; b0 < shi; b0 += width) {
                        const int cnt = min(width, shi - b0);
                        if (sl < cnt) lds[sl] = nx;
                        wave_sync_lds();
                        if (sl < min(width, shi - b0 - width)) nx = P.pos[b0 + width + sl];
                        for (int k0 = 0; k0 < cnt; k0 += 32, ++w) {
                            const int n = min(32, cnt - k0);
                            uint32_t mk = 0;
                            density_row(pi, lds + k0, n, b0 + k0 - g.p_begin, four, pt, rs, mk);
                            if (base >= 0) P.sms_pool[(size_t)(base + w) * 64 + lane] = finish_rejected(mk, n);
                        }
                        wave_sync_lds();
                    }
                }
                s += row_sum_value(rs, four);
            }
            if (sm.active) density_store(P, sm.self, s);
        }
    }
}

// ---------------------------------------------------------------------------
// Force (force_kernels.cu:525-561 summand, 818-852 / 1019-1051 epilogue).
struct ForceAcc {
    float px, py, pz, vx, vy, vz, gx, gy, gz, l;
};

// h^2 - r^2 in FP32, correct to about one rounding of the result.  Near
// r = h the plain difference h^2 - fl(r^2) cancels: its error is ~1 ulp of h^2,
// i.e. relative h^2/(h^2 - r^2), which makes the kr^2 and (h^2 - r^2)^2 terms of
// a pair near the kernel edge inaccurate.  Here every rounding is recovered:
// the coordinate differences of neighbours are exact (Sterbenz); each square
// is split exactly into a + ea by one FMA; the two sums of the large parts are
// error-free transformations (Fast2Sum after ordering by max/min, all terms
// non-negative); h^2 is the exact pair hh.hi + hh.lo; hh.hi - S is exact by
// Sterbenz whenever S lies in [h^2/2, 2h^2] (and where it does not, t is large
// and its rounding harmless); the small parts are added last.  S, the rounded
// r^2, is returned for the reciprocal square root.
struct HH {
    float hi, lo;   // h^2 = hi + lo exactly
};

__device__ __forceinline__ HH exact_h2(float h) {
    const float hi = h * h;
    return {hi, fmaf(h, h, -hi)};
}

__device__ __forceinline__ float h2_minus_r2(float rx, float ry, float rz, HH hh, float &S) {
    const float a = rx * rx, ea = fmaf(rx, rx, -a);
    const float b = ry * ry, eb = fmaf(ry, ry, -b);
    const float c = rz * rz, ec = fmaf(rz, rz, -c);
    const float p = fmaxf(a, b), q = fminf(a, b);
    const float s = p + q, es = q - (s - p);
    const float p2 = fmaxf(s, c), q2 = fminf(s, c);
    S = p2 + q2;
    const float eS = q2 - (S - p2);
    const float u = hh.hi - S;
    return u + ((((hh.lo - ea) - eb) - ec) - (es + eS));
}

// Pressure, viscosity and colour-field terms of one accepted pair.
__device__ __forceinline__ void force_pair(const float4 &pi, const float4 &ei,
                                           const float4 &pj, const float4 &ej,
                                           const gsph_constants &c, HH hh,
                                           ForceAcc &A) {
    const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
    float d2;
    const float t = h2_minus_r2(rx, ry, rz, hh, d2);
    const float inv = __frsqrt_rn(d2);
    const float V = pj.w;
    // h - r without cancellation: one Newton step on kr (h + r) = t from
    // h - r, with 1/(2h) for 1/(h + r); it leaves dr kr / h of the error dr of
    // r, about 2 ulp of kr over the whole range (r = d2 / sqrt(d2))
    const float kr0 = fmaf(-d2, inv, c.kernel_h);
    const float kr = fmaf(fmaf(-kr0, fmaf(d2, inv, c.kernel_h), t), 0.5f / c.kernel_h, kr0);
    const float Vkr = V * kr;
    const float pw = Vkr * (ei.w + ej.w) * kr * inv;
    A.px = fmaf(-rx, pw, A.px);
    A.py = fmaf(-ry, pw, A.py);
    A.pz = fmaf(-rz, pw, A.pz);
    A.vx = fmaf(ej.x - ei.x, Vkr, A.vx);
    A.vy = fmaf(ej.y - ei.y, Vkr, A.vy);
    A.vz = fmaf(ej.z - ei.z, Vkr, A.vz);
    const float Vt = V * t;
    const float gw = Vt * t;
    A.gx = fmaf(rx, gw, A.gx);
    A.gy = fmaf(ry, gw, A.gy);
    A.gz = fmaf(rz, gw, A.gz);
    A.l = fmaf(Vt, fmaf(-0.75f, t, d2), A.l);
}

// One particle's integration (knIntegrateVelocityE, integration_kernels.cu:
// 235-336): penalty walls, 1/rho and gravity, acceleration and speed clamps,
// Euler step, wall clamp, leapfrog evaluated velocity.  Reads the sorted
// state (p, e, vel vv, acceleration aa) and writes slot i of the output
// state.  Shared by k_integrate and the force kernel's fused store, so both
// give the same bits.
__device__ __forceinline__ float3 integrate_one(const PhysicsArgs &P, int i, const SceneGeom &g,
                                                const float4 &p, const float4 &e, const f3 &vv,
                                                const f3 &aa) {
    const gsph_constants &c = P.c;
    float v[3] = {vv.x, vv.y, vv.z};
    float a[3] = {aa.x, aa.y, aa.z};
    float x[3] = {p.x, p.y, p.z};
    const float alpha = 20000000.f, beta = 200000.f;
    const float kern = P.wall_margin;
    const int order[3] = {1, 0, 2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int ax = order[k];
        float diff = kern - (x[ax] - g.bound_min[ax]);
        if (diff > 0.f) a[ax] += alpha * diff - beta * v[ax];
        diff = kern - (g.bound_max[ax] - x[ax]);
        if (diff > 0.f) a[ax] -= alpha * diff + beta * v[ax];
    }
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) a[ax] = a[ax] * p.w + c.gravity[ax];
    float sp = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    if (sp > 3000000.f) {
        const float f = sqrtf(3000000.f / sp);
        a[0] *= f;
        a[1] *= f;
        a[2] *= f;
    }
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) v[ax] += a[ax] * c.time_step;
    sp = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (sp > 36.f) {
        const float f = sqrtf(36.f / sp);
        v[0] *= f;
        v[1] *= f;
        v[2] *= f;
    }
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) x[ax] += v[ax] * c.time_step;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        if (x[ax] >= g.bound_max[ax]) {
            v[ax] *= c.wall_damping;
            x[ax] = g.bound_max[ax] - kern;
        }
        if (x[ax] <= g.bound_min[ax]) {
            v[ax] *= c.wall_damping;
            x[ax] = g.bound_min[ax] + kern;
        }
    }
    P.pos_out[i] = make_float4(x[0], x[1], x[2], p.w);
    f3 vo;
    vo.x = v[0];
    vo.y = v[1];
    vo.z = v[2];
    P.vel_out[i] = vo;
    P.ev_out[i] = make_float4((v[0] + e.x) * 0.5f, (v[1] + e.y) * 0.5f,
                              (v[2] + e.z) * 0.5f, e.w);
    return make_float3(x[0], x[1], x[2]);
}

// Acceleration of particle i from its sums; with P.fuse_integrate the
// particle is integrated here too (the full step), else k_integrate does it.
template <bool kBatch>
__device__ __forceinline__ void force_store(const PhysicsArgs &P, int i, const float4 &pi,
                                            const float4 &ei, ForceAcc A) {
    const gsph_constants &c = P.c;
    const float wi = pi.w;
    const float fx = A.px * c.half_spiky_value + A.vx * c.viscosity_visco_value;
    const float fy = A.py * c.half_spiky_value + A.vy * c.viscosity_visco_value;
    const float fz = A.pz * c.half_spiky_value + A.vz * c.viscosity_visco_value;
    const float gx = A.gx * c.grad_color_scale;
    const float gy = A.gy * c.grad_color_scale;
    const float gz = A.gz * c.grad_color_scale;
    const float l = A.l * c.lplc_color_scale * wi;
    const float gn = sqrtf(gx * gx + gy * gy + gz * gz);
    float sx = 0.f, sy = 0.f, sz = 0.f;
    if (gn > c.surface_normal) {
        const float k = c.surface_coe * l / gn;
        sx = gx * k;
        sy = gy * k;
        sz = gz * k;
    }
    f3 a;
    a.x = fx * c.mass + sx;
    a.y = fy * c.mass + sy;
    a.z = fz * c.mass + sz;
    P.acc[i] = a;
    if (P.fuse_integrate) {
        const SceneGeom g = P.scenes[scene_of<kBatch>(P, i)];
        const float3 x = integrate_one(P, i, g, pi, ei, P.vel[i], a);
        // the next arrange's micro id of the new position, as k_hash computes it
        int mx, my, mz;
        uint32_t key;
        if (micro_coords(x.x, x.y, x.z, P.c.inv_cell_size, g, mx, my, mz)) {
            key = micro_key(mx, my, mz, g);
        } else {
            atomicOr(&P.dcw[kCntErr], kErrOutOfDomain);
            key = (uint32_t)g.cell_base << 6;
        }
        P.next_keys[i] = key;
    }
}

// LDS of the force kernel, per wave: the staged-row path's two group buffers
// (pos then ev, kForceCap entries each), or the round walk's row bounds and
// staging round (pos, ev), or an SMS item's batch.
constexpr int kForceCap = 160;   // entries of one row pair (group) per buffer
template <int kStage>
struct ForceLds {
    static constexpr int kWalk = 2 * 9 * 64 / 4 + 2 * kStage;
    static constexpr int kWords = (4 * kForceCap > kWalk ? 4 * kForceCap : kWalk);
};

// Accept mask of cnt <= 32 staged candidates q[0, cnt): bit k for candidate k.
__device__ __forceinline__ uint32_t test_row(const float4 &pi, const float4 *q, int cnt,
                                             PairTest pt) {
    uint32_t mask = 0;   // bits enter at the bottom (reversed order, undone below)
    int k = 0;
    for (; k + 4 <= cnt; k += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) mask += mask + (accept(pi, q[k + u], pt) ? 1u : 0u);
    }
    for (; k < cnt; ++k) mask += mask + (accept(pi, q[k], pt) ? 1u : 0u);
    return finish_mask(mask, cnt);
}

// The lean force passes take the TRA tiles where every lane has the density
// pass's row plan and accepted-pair masks (no row longer than one mask word,
// no pool rounds): the staged-row path alone, in one group buffer of
// kForceCap entries (5 KB of LDS).  The first (kWin = false) takes the tiles
// whose mirrored row pairs each fit the buffer, in 72 registers, seven waves
// per SIMD; the second (kWin = true) the planned tiles the first left (tiles
// across cell rows of small scenes), each pair streamed through the buffer in
// windows.  Every other tile is flagged in tile_defer for the general pass
// (k_force); all give the same bits.  slot: the counter that records, once
// per arrange epoch, that a pass left a tile (the windowed pass, always the
// last lean one, uses the one the general pass reads).
__device__ __forceinline__ void defer_force_tile(const PhysicsArgs &P, int it, int lane, int slot) {
    if (lane == 0) {
        P.tile_defer[it] = 1;
        const int ep = P.dc[kCntEpoch];
        if (*reinterpret_cast<const volatile int *>(&P.dc[slot]) != ep) P.dcw[slot] = ep;
    }
}

// Bits k of mask m (candidate lo + k, at position o + k of a row pair's
// staged span) whose position lies in the window [ws, we).
__device__ __forceinline__ uint32_t window_bits(uint32_t m, int o, int ws, int we) {
    const int k0 = max(ws - o, 0), k1 = min(we - o, 32);
    if (k1 <= k0) return 0u;
    const uint32_t upto = k1 >= 32 ? 0xffffffffu : (1u << k1) - 1u;
    return m & upto & ~((1u << k0) - 1u);
}

template <bool kBatch, bool kWin>
__device__ __forceinline__ void force_tile_lean(const PhysicsArgs &P, int it, int tra, int lane,
                                                float4 *wl, HH hh, bool ident) {
    const int slot = kWin ? kCntForceAnyDefer2 : kCntForceAnyDefer;
    const int idx = it * 64 + lane;
    const bool active = idx < tra;
    const int i = tra_particle(P, active ? idx : tra - 1, ident);
    const uint32_t flag = P.use_masks && active ? P.nbmask[(size_t)9 * P.mask_stride + idx] : 0u;
    const bool planned = (flag & kFlagPlan) && !(flag & (kFlagLongRow | kFlagPool));
    if (!P.use_masks || __ballot(active && !planned) != 0) {
        defer_force_tile(P, it, lane, slot);
        return;
    }
    const float4 pi = P.pos[i];
    const float4 ei = P.ev[i];
    int lo[9], hi[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const uint32_t w = active ? P.nbmask[(size_t)(10 + r) * P.mask_stride + idx] : 0u;
        lo[r] = active ? i + ((int)(w << 8) >> 8) : 0;
        hi[r] = active ? lo[r] + (int)(w >> 24) : 0;
    }
    int L[9], S[9];
    row_spans(lo, hi, L, S);
    if (!kWin) {
        bool fits = true;
#pragma unroll
        for (int g = 0; g < 5; ++g)
            fits = fits && S[row_of(2 * g)] + (g < 4 ? S[row_of(2 * g + 1)] : 0) <= kForceCap;
        if (!fits) {
            defer_force_tile(P, it, lane, slot);
            return;
        }
    }
    if (lane == 0) P.tile_defer[it] = 0;
    ForceAcc A = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    float4 *bp = wl, *be = wl + kForceCap;
#pragma unroll
    for (int g = 0; g < 5; ++g) {
        const int ra = row_of(2 * g), rb = g < 4 ? row_of(2 * g + 1) : 0;
        const int sa = S[ra], sb = g < 4 ? S[rb] : 0;
        // the pair's span: row a at [0, sa), row b at [sa, sa + sb)
        const int oa = lo[ra] - L[ra];
        const int ob = sa + lo[rb] - L[rb];
        const uint32_t fa = active ? P.nbmask[(size_t)ra * P.mask_stride + idx] : 0u;
        const uint32_t fb = active && g < 4 ? P.nbmask[(size_t)rb * P.mask_stride + idx] : 0u;
        // one window when the pair fits (always, in the first pass)
        const int span = kWin ? sa + sb : 1;
        for (int ws = 0; ws < span; ws += kForceCap) {   // wave-uniform
            const int we = kWin ? min(ws + kForceCap, sa + sb) : sa + sb;
            if (kWin || g > 0) wave_sync_lds();   // earlier reads of the buffer precede these writes
            const int a1 = min(we, sa), b0 = max(ws, sa);
            if (a1 > ws) {
                dma_span(P.pos, L[ra] + ws, a1 - ws, bp, lane);
                dma_span(P.ev, L[ra] + ws, a1 - ws, be, lane);
            }
            if (we > b0) {
                dma_span(P.pos, L[rb] + b0 - sa, we - b0, bp + (b0 - ws), lane);
                dma_span(P.ev, L[rb] + b0 - sa, we - b0, be + (b0 - ws), lane);
            }
            wait_dma();
            // row a's pairs, then row b's (all of row a's positions come first)
            uint32_t ma = kWin ? window_bits(fa, oa, ws, we) : fa;
            uint32_t mb = kWin ? window_bits(fb, ob, ws, we) : fb;
            while (ma | mb) {
                const bool ua = ma != 0;
                const int j = (ua ? oa : ob) - ws + __builtin_ctz(ua ? ma : mb);
                if (ua)
                    ma &= ma - 1u;
                else
                    mb &= mb - 1u;
                force_pair(pi, ei, bp[j], be[j], P.c, hh, A);
            }
        }
    }
    if (active) force_store<kBatch>(P, i, pi, ei, A);
}

// kAll: the pass takes every TRA tile (the first pass); else only the tiles
// the first pass left.  Plain contexts run the fit-only pass and then the
// windowed one; batched contexts, whose small scenes put many tiles across
// cell rows, run the windowed pass over every tile.
template <bool kBatch, bool kWin, bool kAll>
__global__ __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(kWin ? 6 : 7)))
void k_force_lean(PhysicsArgs P) {
    __shared__ float4 s_lds[kNT / 64][2 * kForceCap];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float4 *wl = s_lds[wave];
    const int tra = P.dc[kCntTra];
    // integrated here: the state goes back in sorted order (see k_integrate)
    if (P.fuse_integrate && blockIdx.x == 0 && threadIdx.x == 0) {
        P.dcw[kCntOrder] = 1;
        P.dcw[kCntKeysReady] = 1;
    }
    // a second pass only in an epoch where the first left a tile
    const int tiles = kAll || P.dc[kCntForceAnyDefer] == P.dc[kCntEpoch] ? (tra + 63) >> 6 : 0;
    const ItemMap m = item_map(tiles);
    const HH hh = exact_h2(P.c.kernel_h);
    const bool ident = tra_ident(P);
    for (int it = m.first; it < m.end; it += m.stride) {
        if (!kAll && !P.tile_defer[it]) continue;   // the first pass took it
        wave_sync_lds();   // the previous item's LDS reads precede this item's writes
        force_tile_lean<kBatch, kWin>(P, it, tra, lane, wl, hh, ident);
    }
}

template <bool kBatch, int kStage, bool kLegacy>
__global__ __launch_bounds__(kNT) void k_force(PhysicsArgs P) {
    __shared__ float4 s_lds[kNT / 64][ForceLds<kStage>::kWords];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float4 *wl = s_lds[wave];
    int *rlo = reinterpret_cast<int *>(wl), *rhi = rlo + 9 * 64;
    float4 *sp = wl + 2 * 9 * 64 / 4;   // round walk: pos, then ev
    float4 *se = sp + kStage;
    const int tra = P.dc[kCntTra];
    const int T = P.dc[kCntT];
    // integrated here: the state goes back in sorted order (see k_integrate)
    // and next_keys holds its micro ids
    if (P.fuse_integrate && blockIdx.x == 0 && threadIdx.x == 0) {
        P.dcw[kCntOrder] = 1;
        P.dcw[kCntKeysReady] = 1;
    }
    // without the lean pass every TRA tile; with it, only when it left some
    // in this arrange epoch
    const int tra_tiles = !P.lean || P.dc[kCntForceAnyDefer2] == P.dc[kCntEpoch] ? (tra + 63) >> 6 : 0;
    const int items = tra_tiles + ((T + 1) >> 1);
    const ItemMap m = item_map(items);
    const HH hh = exact_h2(P.c.kernel_h);
    const bool ident = tra_ident(P);
    const PairTest pt = pair_test(P.c.kernel_2);
    for (int it = m.first; it < m.end; it += m.stride) {
        wave_sync_lds();   // the previous item's LDS reads precede this item's writes
        if (it < tra_tiles && P.lean && !P.tile_defer[it]) continue;   // the lean pass took it
        if (it < tra_tiles) {
            const int idx = it * 64 + lane;
            const bool active = idx < tra;
            const int i = tra_particle(P, active ? idx : tra - 1, ident);
            const float4 pi = P.pos[i];
            const float4 ei = P.ev[i];
            // the density pass's row plan when it stored one, else plan here
            const uint32_t flag = P.use_masks && active ? P.nbmask[(size_t)9 * P.mask_stride + idx] : 0u;
            int lo[9], hi[9];
            if (flag & kFlagPlan) {
#pragma unroll
                for (int r = 0; r < 9; ++r) {
                    const uint32_t w = P.nbmask[(size_t)(10 + r) * P.mask_stride + idx];
                    lo[r] = i + ((int)(w << 8) >> 8);
                    hi[r] = lo[r] + (int)(w >> 24);
                }
            } else {
                tra_rows<kBatch>(P, pi, active, i, lo, hi);
            }
            ForceAcc A = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            // Lanes whose density pass recorded accepted-pair masks run the
            // pair body over them; the others (masks off, or a row longer
            // than one mask word) test the staged candidates first.  Same
            // pairs, same order.  Tiles with long rows read their masks per
            // staging round from the pool (lane 0 is active in every tile: its
            // flag is the tile's).
            const uint32_t flag0 = __builtin_amdgcn_readfirstlane(flag);
            const bool pool = P.use_masks && (flag0 & kFlagPool) != 0;
            const bool self_test = !P.use_masks || (active && (flag & kFlagLongRow) != 0);
            const bool use_mask = P.use_masks && active && !self_test && !pool;
            int L[9], S[9];
            row_spans(lo, hi, L, S);
            bool lng = false;
#pragma unroll
            for (int r = 0; r < 9; ++r) lng |= self_test && hi[r] - lo[r] > kMaskRowMax;
            bool fits = !pool && __ballot(lng) == 0;
#pragma unroll
            for (int g = 0; g < 5; ++g)
                fits = fits && S[row_of(2 * g)] + (g < 4 ? S[row_of(2 * g + 1)] : 0) <= kForceCap;
            if (fits) {
                // Staged-row path: the rows in mirrored pairs (groups), both
                // rows of a group in LDS at once (pos and ev by LDS-direct
                // loads, the next group's while this one computes), and each
                // lane walks its accepted pairs of the two rows as one loop:
                // row a in index order, then row b.
                // a group's staged rows and its two mask words are loaded
                // while the previous group computes
                uint32_t nma = 0u, nmb = 0u;
                auto issue = [&](int g) __attribute__((always_inline)) {
                    float4 *bp = wl + (g & 1) * 2 * kForceCap, *be = bp + kForceCap;
                    const int ra = row_of(2 * g), rb = g < 4 ? row_of(2 * g + 1) : 0;
                    const int sb = g < 4 ? S[rb] : 0;
                    dma_span(P.pos, L[ra], S[ra], bp, lane);
                    dma_span(P.ev, L[ra], S[ra], be, lane);
                    dma_span(P.pos, L[rb], sb, bp + S[ra], lane);
                    dma_span(P.ev, L[rb], sb, be + S[ra], lane);
                    nma = use_mask ? P.nbmask[(size_t)ra * P.mask_stride + idx] : 0u;
                    nmb = use_mask && g < 4 ? P.nbmask[(size_t)rb * P.mask_stride + idx] : 0u;
                };
                issue(0);
#pragma unroll
                for (int g = 0; g < 5; ++g) {
                    wait_dma();
                    uint32_t ma = nma, mb = nmb;
                    if (g + 1 < 5) issue(g + 1);
                    const float4 *bp = wl + (g & 1) * 2 * kForceCap, *be = bp + kForceCap;
                    const int ra = row_of(2 * g), rb = g < 4 ? row_of(2 * g + 1) : 0;
                    const int oa = lo[ra] - L[ra];
                    const int ob = g < 4 ? S[ra] + lo[rb] - L[rb] : 0;
                    if (self_test) {
                        ma = test_row(pi, bp + oa, hi[ra] - lo[ra], pt);
                        if (g < 4) mb = test_row(pi, bp + ob, hi[rb] - lo[rb], pt);
                    }
                    while (ma | mb) {
                        const bool ua = ma != 0;
                        const int j = (ua ? oa : ob) + __builtin_ctz(ua ? ma : mb);
                        if (ua)
                            ma &= ma - 1u;
                        else
                            mb &= mb - 1u;
                        force_pair(pi, ei, bp[j], be[j], P.c, hh, A);
                    }
                }
            } else {
                const int pbase = (int)(flag0 >> 3);
                const int pool_rounds = pool ? tra_round_bound<kStage>(lo, hi) : 0;   // density's bound
#pragma unroll
                for (int r = 0; r < 9; ++r) {
                    rlo[r * 64 + lane] = lo[r];
                    rhi[r * 64 + lane] = hi[r];
                }
                float4 p0 = pi, p1 = pi, p2 = pi, p3 = pi, e0 = ei, e1 = ei, e2 = ei, e3 = ei;
                uint32_t m_ld = 0, m_st = 0;   // row mask of the loaded / the staged round
                int q_ld = -1;
                int q = -1, cur = kNone, lo_r = 0, hi_r = 0;
                uint32_t pw_ld[kStage / 32], pw_st[kStage / 32];   // pool words of a round
                int round_ld = 0;
                auto load = [&]() __attribute__((always_inline)) {
                    load_round<kStage>(P.pos, cur, P.n, lane, p0, p1, p2, p3);
                    load_round<kStage>(P.ev, cur, P.n, lane, e0, e1, e2, e3);
                    if (q != q_ld) {
                        q_ld = q;
                        m_ld = use_mask ? P.nbmask[(size_t)row_of(q) * P.mask_stride + idx] : 0u;
                    }
                    if (pool) {
#pragma unroll
                        for (int c = 0; c < kStage / 32; ++c)
                            pw_ld[c] = round_ld < pool_rounds
                                           ? P.sms_pool[(size_t)(pbase + round_ld * (kStage / 32) + c) * 64 + lane]
                                           : 0u;
                        ++round_ld;
                    }
                };
                next_round<kStage>(q, cur, lo_r, hi_r, rlo, rhi, lane);
                if (cur != kNone) load();
                while (cur != kNone) {
                    const int base = cur, lo_c = lo_r, hi_c = hi_r;
                    store_round<kStage>(sp, lane, p0, p1, p2, p3);
                    store_round<kStage>(se, lane, e0, e1, e2, e3);
                    m_st = m_ld;
#pragma unroll
                    for (int c = 0; c < kStage / 32; ++c) pw_st[c] = pool ? pw_ld[c] : 0u;
                    wave_sync_lds();
                    next_round<kStage>(q, cur, lo_r, hi_r, rlo, rhi, lane);
                    if (cur != kNone) load();
                    const int jb = max(lo_c, base), je = min(hi_c, base + kStage);
                    const float4 *qp = sp + (jb - base);
                    const float4 *qe = se + (jb - base);
                    if (pool) {
                        // bit k of word c: window position 32 c + k (zero outside
                        // this lane's range)
#pragma unroll
                        for (int c = 0; c < kStage / 32; ++c) {
                            uint32_t mm = pw_st[c];
                            while (mm) {
                                const int kk = 32 * c + __builtin_ctz(mm);
                                mm &= mm - 1u;
                                force_pair(pi, ei, sp[kk], se[kk], P.c, hh, A);
                            }
                        }
                    } else if (!self_test) {
                        const int cnt = je - jb;   // <= 32, and jb - lo < 32
                        uint32_t mm = 0;
                        if (cnt > 0) {
                            mm = m_st >> (jb - lo_c);
                            if (cnt < 32) mm &= (1u << cnt) - 1u;
                        }
                        while (mm) {
                            const int kk = __builtin_ctz(mm);
                            mm &= mm - 1u;
                            force_pair(pi, ei, qp[kk], qe[kk], P.c, hh, A);
                        }
                    } else {
                        // 32 candidates at a time: accept mask, then the pair
                        // body over the accepted ones (in order)
                        for (int b0 = 0; b0 < je - jb; b0 += 32) {
                            uint32_t mask = test_row(pi, qp + b0, min(32, je - jb - b0), pt);
                            while (mask) {
                                const int kk = b0 + __builtin_ctz(mask);
                                mask &= mask - 1u;
                                force_pair(pi, ei, qp[kk], qe[kk], P.c, hh, A);
                            }
                        }
                    }
                    wave_sync_lds();
                }
            }
            if (active) force_store<kBatch>(P, i, pi, ei, A);
        } else {
            const int item = it - tra_tiles;
            const SmsItem sm = sms_item<kBatch>(P, item << 1, T, lane);
            const float4 pi = P.pos[sm.si];
            const float4 ei = P.ev[sm.si];
            // first pool row of the density pass's masks (-1: test here)
            const int base = P.use_masks ? P.sms_off[item] : -1;
            int w = 0;
            const SceneGeom g = P.scenes[sm.gidx];
            const int width = sm.shared ? 64 : 32;
            const int sl = sm.shared ? lane : (lane & 31);
            float4 *lp = sm.shared ? wl : wl + (lane & 32);
            float4 *le = lp + 64;
            int *tab = reinterpret_cast<int *>(wl + 128) + (sm.shared ? 0 : (lane & 32));
            ForceAcc A = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            // same-cell pair force: the pairs inside the particle's own cell
            // come summed from k_same_cell, and its own cell's candidates
            // (one sorted range inside the centre row) are skipped here
            // (force_kernels.cu:858-866, 392-398)
            const bool same_cell = P.scf && sm.has_task && sm.cell_count <= kSameCellCap;
            if (same_cell && sm.active) {
                const float *f = P.scf + sm.self;
                const size_t st = (size_t)P.scf_stride;
                A = {f[0], f[st], f[2 * st], f[3 * st], f[4 * st], f[5 * st], f[6 * st], f[7 * st],
                     f[8 * st], f[9 * st]};
            }
            const int own_lo = same_cell ? sm.cell_begin : 0;
            const int own_hi = same_cell ? sm.cell_begin + sm.cell_count : 0;
            // the nine rows' bounds, parked in this lane's LDS slots (after
            // the batch and the legacy segment table) for the row loop
            static_assert(ForceLds<kStage>::kWords >= 160 + 2 * 9 * 64 / 4, "SMS row slots fit");
            int *srow = reinterpret_cast<int *>(wl + 160);
            {
                int lo9[9], hi9[9];
                sms_rows(g, P, sm, lo9, hi9);
#pragma unroll
                for (int r = 0; r < 9; ++r) {
                    srow[r * 64 + lane] = lo9[r];
                    srow[(9 + r) * 64 + lane] = hi9[r];
                }
            }
            for (int q = 0; q < 9; ++q) {
                const int r = row_of(q);
                const int lo = srow[r * 64 + lane], hi = srow[(9 + r) * 64 + lane];
                const int nseg = kLegacy ? legacy_segments(P, g, sm, r, lo, hi, sl, tab) : 1;
                for (int sg = 0; sg < nseg; ++sg) {
                    const int slo = kLegacy ? tab[2 * sg] : lo;
                    const int shi = kLegacy ? tab[2 * sg + 1] : hi;
                    // the next batch's candidate (and its masks) are loaded while
                    // this batch computes
                    float4 np = pi, ne = ei;
                    uint32_t nm0 = 0, nm1 = 0;
                    auto fetch = [&](int b, int wb) __attribute__((always_inline)) {
                        const int c = min(width, shi - b);
                        if (sl < c) {
                            np = P.pos[b + sl];
                            ne = P.ev[b + sl];
                        }
                        if (base >= 0 && c > 0) {
                            const size_t row0 = (size_t)(base + wb) * 64 + lane;
                            nm0 = P.sms_pool[row0];
                            if (c > 32) nm1 = P.sms_pool[row0 + 64];
                        }
                    };
                    fetch(slo, w);
                    for (int b0 = slo; b0 < shi; b0 += width) {
                        const int cnt = min(width, shi - b0);
                        if (sl < cnt) {
                            lp[sl] = np;
                            le[sl] = ne;
                        }
                        const uint32_t cm0 = nm0, cm1 = nm1;
                        wave_sync_lds();
                        fetch(b0 + width, w + ((cnt + 31) >> 5));
                        // 32 candidates at a time: accept mask, then the pair body
                        for (int k0 = 0; k0 < cnt; k0 += 32, ++w) {
                            const int n = min(32, cnt - k0);
                            uint32_t mask = base >= 0 ? (k0 == 0 ? cm0 : cm1) : test_row(pi, lp + k0, n, pt);
                            if (same_cell && r == 4) {
                                // candidates b0 + k0 + [ol, oh) are the own cell's
                                const int ol = min(max(own_lo - (b0 + k0), 0), 32);
                                const int oh = min(max(own_hi - (b0 + k0), 0), 32);
                                const uint32_t below = ol >= 32 ? ~0u : (1u << ol) - 1u;
                                const uint32_t upto = oh >= 32 ? ~0u : (1u << oh) - 1u;
                                mask &= ~(upto & ~below);
                            }
                            while (mask) {
                                const int kk = k0 + __builtin_ctz(mask);
                                mask &= mask - 1u;
                                force_pair(pi, ei, lp[kk], le[kk], P.c, hh, A);
                            }
                        }
                        wave_sync_lds();
                    }
                }
            }
            if (sm.active) force_store<kBatch>(P, sm.self, pi, ei, A);
        }
    }
}

// ---------------------------------------------------------------------------
// Same-cell pair force (the reference's GSPH_ENABLE_SAME_CELL_PAIR_FORCE
// option: computeSameCellPairForceKernel, force_kernels.cu:9-131).  Every
// pair of particles inside one SMS cell of at most kSameCellCap particles is
// computed once and both sides are accumulated; the SMS force pass starts
// each particle of such a cell from these sums and skips its own cell's
// candidates.  One workgroup of kSameCellNT threads per cell (the cell of
// each task with p_offset 0).  Instead of atomics, the pairs are visited in
// rounds d = 1 .. n/2: in round d particle i pairs with (i + d) mod n (for
// even n, in round n/2 only i < n/2), which covers each pair once and, within
// a round, sends the partner's share of every pair to a different particle.
// A thread keeps the sums of its own particles (i = tid + kSameCellNT u) in
// registers and adds the partner shares into LDS, the rounds separated by
// barriers, so the result does not depend on timing.  The terms of each side
// are those of force_pair.
constexpr int kSameCellNT = 256;
constexpr int kSameCellOwn = kSameCellCap / kSameCellNT;

struct SameCellLds {
    float4 pos[kSameCellCap];
    float4 ev[kSameCellCap];
    float part[10][kSameCellCap];   // partner shares
};

template <int kOwn>
__device__ __forceinline__ void same_cell_block(const PhysicsArgs &P, SameCellLds &L, int cb, int n,
                                                HH hh, PairTest pt) {
    const int tid = threadIdx.x;
    for (int k = tid; k < n; k += kSameCellNT) {
        L.pos[k] = P.pos[cb + k];
        L.ev[k] = P.ev[cb + k];
#pragma unroll
        for (int f = 0; f < 10; ++f) L.part[f][k] = 0.f;
    }
    __syncthreads();
    ForceAcc own[kOwn];
#pragma unroll
    for (int u = 0; u < kOwn; ++u) own[u] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const gsph_constants &c = P.c;
    const int half = n >> 1;
    for (int d = 1; d <= half; ++d) {
        const int lim = 2 * d == n ? half : n;   // particles that start a pair this round
#pragma unroll
        for (int u = 0; u < kOwn; ++u) {
            const int i = tid + kSameCellNT * u;
            if (i < lim) {
                int j = i + d;
                if (j >= n) j -= n;
                const float4 pi = L.pos[i], ei = L.ev[i], pj = L.pos[j], ej = L.ev[j];
                if (accept(pi, pj, pt)) {
                    const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
                    float d2;
                    const float t = h2_minus_r2(rx, ry, rz, hh, d2);
                    const float inv = __frsqrt_rn(d2);
                    const float kr0 = fmaf(-d2, inv, c.kernel_h);
                    const float kr =
                        fmaf(fmaf(-kr0, fmaf(d2, inv, c.kernel_h), t), 0.5f / c.kernel_h, kr0);
                    const float psum = ei.w + ej.w;
                    const float lap = fmaf(-0.75f, t, d2);
                    // side i (V = w_j)
                    ForceAcc &A = own[u];
                    const float Vkr = pj.w * kr;
                    const float pw = Vkr * psum * kr * inv;
                    A.px = fmaf(-rx, pw, A.px);
                    A.py = fmaf(-ry, pw, A.py);
                    A.pz = fmaf(-rz, pw, A.pz);
                    A.vx = fmaf(ej.x - ei.x, Vkr, A.vx);
                    A.vy = fmaf(ej.y - ei.y, Vkr, A.vy);
                    A.vz = fmaf(ej.z - ei.z, Vkr, A.vz);
                    const float Vt = pj.w * t;
                    const float gw = Vt * t;
                    A.gx = fmaf(rx, gw, A.gx);
                    A.gy = fmaf(ry, gw, A.gy);
                    A.gz = fmaf(rz, gw, A.gz);
                    A.l = fmaf(Vt, lap, A.l);
                    // side j (V = w_i, the relative vectors negated)
                    const float Wkr = pi.w * kr;
                    const float qw = Wkr * psum * kr * inv;
                    const float Wt = pi.w * t;
                    const float hw = Wt * t;
                    L.part[0][j] = fmaf(rx, qw, L.part[0][j]);
                    L.part[1][j] = fmaf(ry, qw, L.part[1][j]);
                    L.part[2][j] = fmaf(rz, qw, L.part[2][j]);
                    L.part[3][j] = fmaf(ei.x - ej.x, Wkr, L.part[3][j]);
                    L.part[4][j] = fmaf(ei.y - ej.y, Wkr, L.part[4][j]);
                    L.part[5][j] = fmaf(ei.z - ej.z, Wkr, L.part[5][j]);
                    L.part[6][j] = fmaf(-rx, hw, L.part[6][j]);
                    L.part[7][j] = fmaf(-ry, hw, L.part[7][j]);
                    L.part[8][j] = fmaf(-rz, hw, L.part[8][j]);
                    L.part[9][j] = fmaf(Wt, lap, L.part[9][j]);
                }
            }
        }
        __syncthreads();   // this round's partner shares precede the next round's
    }
#pragma unroll
    for (int u = 0; u < kOwn; ++u) {
        const int i = tid + kSameCellNT * u;
        if (i < n) {
            const ForceAcc &A = own[u];
            const float v[10] = {A.px, A.py, A.pz, A.vx, A.vy, A.vz, A.gx, A.gy, A.gz, A.l};
#pragma unroll
            for (int f = 0; f < 10; ++f)
                P.scf[(size_t)f * P.scf_stride + cb + i] = v[f] + L.part[f][i];
        }
    }
    __syncthreads();   // the reads above precede the next cell's staging
}

__global__ __launch_bounds__(kSameCellNT) void k_same_cell(PhysicsArgs P) {
    __shared__ SameCellLds L;
    const int T = P.dc[kCntT];
    const HH hh = exact_h2(P.c.kernel_h);
    const PairTest pt = pair_test(P.c.kernel_2);
    for (int t = blockIdx.x; t < T; t += gridDim.x) {
        const gsph_task tk = P.tasks[t];
        const int n = tk.cell_particle_count;
        if (tk.p_offset != 0 || n > kSameCellCap) continue;   // block-uniform
        if (n <= kSameCellNT)
            same_cell_block<1>(P, L, tk.cell_begin, n, hh, pt);
        else
            same_cell_block<kSameCellOwn>(P, L, tk.cell_begin, n, hh, pt);
    }
}

// ---------------------------------------------------------------------------
// knIntegrateVelocityE (src/solver/integration_kernels.cu:235-336): penalty
// walls, a*w + g, |a| and |v| clamps, Euler step, wall clamps, leapfrog ev.
// Reads the sorted buffer, writes the resident one.
template <bool kBatch>
__global__ __launch_bounds__(kNT) void k_integrate(PhysicsArgs P) {
    const int i = blockIdx.x * kNT + threadIdx.x;
    // the state goes back in sorted order: the slot order is hash[]'s again
    // (the next keys are not computed here)
    if (i == 0) {
        P.dcw[kCntOrder] = 1;
        P.dcw[kCntKeysReady] = 0;
    }
    if (i >= P.n) return;
    const SceneGeom g = P.scenes[scene_of<kBatch>(P, i)];
    (void)integrate_one(P, i, g, P.pos[i], P.ev[i], P.vel[i], P.acc[i]);
}

// knInit (integration_kernels.cu:6-17): ev = 0, acc = 0.
__global__ void k_init_state(float4 *ev, f3 *acc, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ev[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    f3 z;
    z.x = z.y = z.z = 0.f;
    acc[i] = z;
}

// FP64 checksum: block b sums slots [b*chunk, (b+1)*chunk) in a fixed tree
// order; the final block adds the partials in block order.
constexpr int kSumBlocks = kChecksumBlocks;

__global__ __launch_bounds__(kNT) void k_checksum_partial(const float4 *pos, int n,
                                                          double *partial) {
    __shared__ double s[5][kNT];
    const int chunk = (n + gridDim.x - 1) / gridDim.x;
    const int b0 = blockIdx.x * chunk;
    const int b1 = min(n, b0 + chunk);
    double acc[5] = {0, 0, 0, 0, 0};
    for (int i = b0 + threadIdx.x; i < b1; i += kNT) {
        const float4 q = pos[i];
        acc[0] += q.x;
        acc[1] += q.y;
        acc[2] += q.z;
        acc[3] += q.w;
        acc[4] += (double)((i & 1023) + 1) * ((double)q.x + 3.0 * q.y + 7.0 * q.z);
    }
    for (int k = 0; k < 5; ++k) s[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = kNT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 5; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 5) partial[blockIdx.x * 5 + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void k_checksum_final(const double *partial, int nblocks, double *out) {
    if (threadIdx.x >= 5) return;
    double t = 0;
    for (int b = 0; b < nblocks; ++b) t += partial[b * 5 + threadIdx.x];
    out[threadIdx.x] = t;
}

}  // namespace

template <bool B, int S>
void launch_density_t(const PhysicsArgs &p, hipStream_t s) {
    if (p.lean) hipLaunchKernelGGL((k_density_lean<B, S>), dim3(p.grid_blocks), dim3(kNT), 0, s, p);
    if (p.sms_legacy)
        hipLaunchKernelGGL((k_density<B, S, true>), dim3(p.grid_blocks), dim3(kNT), 0, s, p);
    else
        hipLaunchKernelGGL((k_density<B, S, false>), dim3(p.grid_blocks), dim3(kNT), 0, s, p);
}

template <bool B, int S>
void launch_force_t(const PhysicsArgs &p, hipStream_t s) {
    if (p.lean && B) {
        hipLaunchKernelGGL((k_force_lean<B, true, true>), dim3(p.grid_blocks), dim3(kNT), 0, s, p);
    } else if (p.lean) {
        hipLaunchKernelGGL((k_force_lean<B, false, true>), dim3(p.grid_blocks), dim3(kNT), 0, s, p);
        // the windowed pass takes only the planned tiles the first left (few
        // or none, never SMS items): a quarter of the grid, still more blocks
        // than every CU holds at once, costs a quarter of the block dispatch
        // (the item schedule serves any multiple of 8 blocks)
        const int gb = std::max(8, (p.grid_blocks / 4) & ~7);
        hipLaunchKernelGGL((k_force_lean<B, true, false>), dim3(gb), dim3(kNT), 0, s, p);
    }
    if (p.sms_legacy)
        hipLaunchKernelGGL((k_force<B, S, true>), dim3(p.grid_blocks), dim3(kNT), 0, s, p);
    else
        hipLaunchKernelGGL((k_force<B, S, false>), dim3(p.grid_blocks), dim3(kNT), 0, s, p);
}

template <bool B>
void launch_density_b(const PhysicsArgs &p, hipStream_t s) {
    if (p.stage_depth == 64)
        launch_density_t<B, 64>(p, s);
    else if (p.stage_depth == 256)
        launch_density_t<B, 256>(p, s);
    else
        launch_density_t<B, kStageDefault>(p, s);
}

template <bool B>
void launch_force_b(const PhysicsArgs &p, hipStream_t s) {
    if (p.stage_depth == 64)
        launch_force_t<B, 64>(p, s);
    else if (p.stage_depth == 256)
        launch_force_t<B, 256>(p, s);
    else
        launch_force_t<B, kStageDefault>(p, s);
}

void launch_density(const PhysicsArgs &p, bool batch, hipStream_t s) {
    if (batch)
        launch_density_b<true>(p, s);
    else
        launch_density_b<false>(p, s);
}

void launch_force(const PhysicsArgs &p, bool batch, hipStream_t s) {
    if (batch)
        launch_force_b<true>(p, s);
    else
        launch_force_b<false>(p, s);
}

void launch_same_cell(const PhysicsArgs &p, hipStream_t s) {
    // 36 KB of LDS per block, four blocks per CU; a persistent grid over the tasks
    hipLaunchKernelGGL(k_same_cell, dim3(1024), dim3(kSameCellNT), 0, s, p);
}

void launch_integrate(const PhysicsArgs &p, bool batch, hipStream_t s) {
    const int blocks = (p.n + kNT - 1) / kNT;
    if (blocks == 0) return;
    if (batch)
        hipLaunchKernelGGL(k_integrate<true>, dim3(blocks), dim3(kNT), 0, s, p);
    else
        hipLaunchKernelGGL(k_integrate<false>, dim3(blocks), dim3(kNT), 0, s, p);
}

// ---------------------------------------------------------------------------
// Scratch of the no-allocation checksum entry point: the per-block partial
// sums of up to kNoAllocRanges ranges, in a file-scope device array (each GPU
// holds its own copy for the life of the process).  Host code never passes
// the array's name to a kernel: it passes the address hipGetSymbolAddress
// returns for the current device, looked up once per device.
namespace {
__device__ __align__(256) double scratch[kNoAllocRanges * kChecksumBlocks * 5];

std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t checksum_scratch(double **ptr) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *ptr = g_scratch_ptr[dev];
    return hipSuccess;
}

void launch_checksum(const float4 *pos, int n, double *partial, int nblocks,
                     double *out5, hipStream_t s) {
    (void)nblocks;
    hipLaunchKernelGGL(k_checksum_partial, dim3(kSumBlocks), dim3(kNT), 0, s, pos, n, partial);
    hipLaunchKernelGGL(k_checksum_final, dim3(1), dim3(64), 0, s, partial, kSumBlocks, out5);
}

void launch_init_state(float4 *ev, f3 *acc, int n, hipStream_t s) {
    const int blocks = (n + kNT - 1) / kNT;
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_init_state, dim3(blocks), dim3(kNT), 0, s, ev, acc, n);
}

}  // namespace gsph
