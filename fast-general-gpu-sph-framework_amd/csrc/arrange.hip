// arrange.hip - the grid stage of one SPH step on gfx950.
//
// Replaces Arrangement::arrangeHybridFrame (reference
// src/grid/sph_arrangement.cu:1493-1506) with a stable, deterministic design:
//
//   k_hash        micro hash of every particle + radix-pass-0 tile histograms
//   k_upsweep     tile histograms of later radix passes
//   k_rowscan     per digit, exclusive scan of the tile histograms over tiles
//   k_downsweep   stable scatter (wave64 ballot ranking, tiles in index order)
//   k_gather      permute pos/vel/ev into the sorted buffer; cell boundaries
//   k_cells       occupied-cell list (id, first particle), chunk cursors
//   k_fill        dense cell begin/count (the reference's d_cell_offset_ and
//                 d_cell_nump_) and per-cell x-layer starts
//   k_split_*     split heuristic, TRA/SMS counts, task generation + pairing
//   k_compact     compact index (TRA first), odd-tail pad, middle value
//
// The LSD radix sort on (micro id, previous slot) makes the permutation the
// stable one: ties keep the pre-sort order, which is one of the outcomes the
// reference's atomicAdd ranks (sph_arrangement.cu:1339) can produce.  Every
// device-wide scan is reduce-then-scan, so no block waits on another.
#include "kernels.h"

namespace gsph {
namespace {

constexpr int kNT = 256;          // threads of the tile kernels (4 waves)
constexpr int kItems = 16;        // items per thread: 4096-key tiles
constexpr int kCompactItems = 4;  // particles per thread of k_compact

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

template <bool kBatch>
__device__ __forceinline__ int scene_of_cell(const SceneGeom *sc, int ns,
                                             int c) {
    if (!kBatch) return 0;
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
        const int m = (lo + hi + 1) >> 1;
        if (sc[m].cell_base <= c)
            lo = m;
        else
            hi = m - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// Hash + pass-0 histogram.  Tile t covers keys [4096 t, 4096 t + 4096); wave w
// of the block owns 1024 consecutive keys, slot s / lane l is key
// 4096 t + 1024 w + 64 s + l, so (wave, slot, lane) order is index order.
template <bool kBatch>
__global__ __launch_bounds__(kNT) void k_hash(ArrangeArgs a) {
    __shared__ uint32_t s_cnt[4][256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int k = tid; k < 4 * 256; k += kNT) (&s_cnt[0][0])[k] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortTile + wave * 1024;
    for (int s = 0; s < kItems; ++s) {
        const int i = base + s * 64 + lane;
        const bool valid = i < a.n;
        uint32_t key = 0;
        if (valid) {
            const SceneGeom g = a.scenes[scene_of_particle<kBatch>(a.scenes, a.nscenes, i)];
            const float4 p = a.pos_a[i];
            int mx, my, mz;
            if (micro_coords(p.x, p.y, p.z, a.inv_cell_size, g, mx, my, mz)) {
                key = micro_key(mx, my, mz, g);
            } else {
                // Out-of-domain particle: reported to the host, parked in the
                // scene's first cell so that every later index stays in range.
                atomicOr(&a.dc[kCntErr], kErrOutOfDomain);
                key = (uint32_t)g.cell_base << 6;
            }
            a.keys[0][i] = key;
        }
        if (valid) atomicAdd(&s_cnt[wave][key & 255u], 1u);
    }
    __syncthreads();
    const uint32_t tot = s_cnt[0][tid] + s_cnt[1][tid] + s_cnt[2][tid] + s_cnt[3][tid];
    a.tilecnt[(size_t)tid * a.sort_tiles + blockIdx.x] = tot;
}

__global__ __launch_bounds__(kNT) void k_upsweep(ArrangeArgs a, int pass) {
    __shared__ uint32_t s_cnt[4][256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t *keys = a.keys[pass & 1];
    const int shift = 8 * pass;
    for (int k = tid; k < 4 * 256; k += kNT) (&s_cnt[0][0])[k] = 0;
    const int base = blockIdx.x * kSortTile + wave * 1024;
    uint32_t kv[kItems];
#pragma unroll
    for (int s = 0; s < kItems; ++s) {
        const int i = base + s * 64 + lane;
        kv[s] = i < a.n ? keys[i] : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kItems; ++s) {
        const int i = base + s * 64 + lane;
        const bool valid = i < a.n;
        if (valid) atomicAdd(&s_cnt[wave][(kv[s] >> shift) & 255u], 1u);
    }
    __syncthreads();
    const uint32_t tot = s_cnt[0][tid] + s_cnt[1][tid] + s_cnt[2][tid] + s_cnt[3][tid];
    a.tilecnt[(size_t)tid * a.sort_tiles + blockIdx.x] = tot;
}

// One block per digit: exclusive scan of the digit's row over tiles, in place;
// the row total goes to digit_total.
__global__ __launch_bounds__(kNT) void k_rowscan(ArrangeArgs a) {
    __shared__ uint32_t s_tmp[8];
    uint32_t *row = a.tilecnt + (size_t)blockIdx.x * a.sort_tiles;
    uint32_t carry = 0;
    for (int c0 = 0; c0 < a.sort_tiles; c0 += kNT) {
        const int t = c0 + threadIdx.x;
        const uint32_t v = t < a.sort_tiles ? row[t] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan<kNT>(v, s_tmp, tot);
        if (t < a.sort_tiles) row[t] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) a.digit_total[blockIdx.x] = carry;
}

// Stable scatter of one radix pass.  dest = (keys with a smaller digit) +
// (same digit in earlier tiles) + (same digit earlier in this tile).  The tile
// is first reordered by digit in LDS, then written out run by run so that
// consecutive threads store to consecutive addresses.
template <bool kFirst>
__global__ __launch_bounds__(kNT) void k_downsweep(ArrangeArgs a, int pass) {
    __shared__ uint32_t s_w[4][256];
    __shared__ uint32_t s_base[256];      // global destination of the tile's digit run
    __shared__ uint32_t s_local[256];     // start of the digit run inside the tile
    __shared__ uint32_t s_tmp[8];
    __shared__ uint32_t s_key[kSortTile];
    __shared__ __align__(8) int s_val[kSortTile];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tile = blockIdx.x;
    // lane masks of the digits while the tile is ranked (in s_val's space,
    // which is free until the reorder): 256 words per wave
    unsigned long long *s_peer = reinterpret_cast<unsigned long long *>(s_val);
    const int shift = 8 * pass;
    const uint32_t *kin = a.keys[pass & 1];
    const int *vin = a.vals[pass & 1];
    uint32_t *kout = a.keys[(pass + 1) & 1];
    int *vout = a.vals[(pass + 1) & 1];
    for (int k = tid; k < 4 * 256; k += kNT) {
        (&s_w[0][0])[k] = 0;
        s_peer[k] = 0ull;
    }
    {
        uint32_t tot;
        const uint32_t ex = block_excl_scan<kNT>(a.digit_total[tid], s_tmp, tot);
        s_base[tid] = ex + a.tilecnt[(size_t)tid * a.sort_tiles + tile];
    }
    __syncthreads();
    const int base = tile * kSortTile + wave * 1024;
    const int tile_n = min(kSortTile, a.n - tile * kSortTile);
    uint32_t key[kItems];
    int val[kItems];
    uint32_t rank[kItems];
#pragma unroll
    for (int s = 0; s < kItems; ++s) {
        const int i = base + s * 64 + lane;
        if (i < a.n) {
            key[s] = kin[i];
            val[s] = kFirst ? i : vin[i];
        } else {
            key[s] = 0xffffffffu;
            val[s] = 0;
        }
    }
#pragma unroll
    for (int s = 0; s < kItems; ++s) {
        const int i = base + s * 64 + lane;
        const bool valid = i < a.n;
        const uint32_t d = (key[s] >> shift) & 255u;
        // lanes of the wave with the same digit: each sets its bit in the
        // digit's word, then reads the word back (one wave's LDS operations
        // complete in order, so the read sees every lane's bit); the first
        // of them clears it again for the next slot
        unsigned long long *pm = s_peer + wave * 256 + d;
        if (valid) atomicOr(pm, 1ull << lane);
        const unsigned long long peers =
            valid ? __hip_atomic_load(pm, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) : 0ull;
        const uint32_t lt = (uint32_t)__popcll(peers & lanemask_lt());
        uint32_t old = 0;
        if (valid) old = s_w[wave][d];
        rank[s] = old + lt;
        if (valid && lt == 0) {
            s_w[wave][d] = old + (uint32_t)__popcll(peers);
            __hip_atomic_store(pm, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        }
    }
    __syncthreads();
    uint32_t cnt_d;
    {
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const uint32_t c = s_w[w][tid];
            s_w[w][tid] = run;
            run += c;
        }
        cnt_d = run;
    }
    {
        uint32_t tot;
        s_local[tid] = block_excl_scan<kNT>(cnt_d, s_tmp, tot);
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kItems; ++s) {
        const int i = base + s * 64 + lane;
        if (i < a.n) {
            const uint32_t d = (key[s] >> shift) & 255u;
            const uint32_t lp = s_local[d] + s_w[wave][d] + rank[s];
            s_key[lp] = key[s];
            s_val[lp] = val[s];
        }
    }
    __syncthreads();
    for (int k = tid; k < tile_n; k += kNT) {
        const uint32_t kk = s_key[k];
        const uint32_t d = (kk >> shift) & 255u;
        const uint32_t dest = s_base[d] + ((uint32_t)k - s_local[d]);
        kout[dest] = kk;
        vout[dest] = s_val[k];
    }
}

// Permute the state into the sorted buffer and count coarse-cell boundaries
// (first particle of each occupied cell) per tile.
__global__ __launch_bounds__(kNT) void k_gather(ArrangeArgs a) {
    __shared__ uint32_t s_tmp[8];
    const int tid = threadIdx.x;
    constexpr int kSlots = kCellTile / kNT;
    const int t0 = blockIdx.x * kCellTile + tid;
    int src[kSlots];
    uint32_t nb = 0;
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
        const int i = t0 + s * kNT;
        src[s] = i < a.n ? a.perm[i] : -1;
        const uint32_t h = i < a.n ? a.hash[i] >> 6 : 0u;
        const uint32_t hp = i < a.n && i > 0 ? a.hash[i - 1] >> 6 : ~0u;
        nb += (i < a.n && (i == 0 || hp != h)) ? 1u : 0u;
    }
    // four slots' loads in flight together, then their stores
    auto move4 = [&](int s0) __attribute__((always_inline)) {
        const int j0 = max(src[s0], 0), j1 = max(src[s0 + 1], 0);
        const int j2 = max(src[s0 + 2], 0), j3 = max(src[s0 + 3], 0);
        const float4 p0 = a.pos_a[j0], p1 = a.pos_a[j1], p2 = a.pos_a[j2], p3 = a.pos_a[j3];
        const float *va = &a.vel_a[0].x;
        float *vb = &a.vel_b[0].x;
        const float v0x = va[3 * j0], v0y = va[3 * j0 + 1], v0z = va[3 * j0 + 2];
        const float v1x = va[3 * j1], v1y = va[3 * j1 + 1], v1z = va[3 * j1 + 2];
        const float v2x = va[3 * j2], v2y = va[3 * j2 + 1], v2z = va[3 * j2 + 2];
        const float v3x = va[3 * j3], v3y = va[3 * j3 + 1], v3z = va[3 * j3 + 2];
        const float4 e0 = a.ev_a[j0], e1 = a.ev_a[j1], e2 = a.ev_a[j2], e3 = a.ev_a[j3];
        auto put = [&](int i, const float4 &p, float vx, float vy, float vz, const float4 &e)
            __attribute__((always_inline)) {
            a.pos_b[i] = p;
            vb[3 * i] = vx;
            vb[3 * i + 1] = vy;
            vb[3 * i + 2] = vz;
            a.ev_b[i] = e;
        };
        const int i = t0 + s0 * kNT;
        if (src[s0] >= 0) put(i, p0, v0x, v0y, v0z, e0);
        if (src[s0 + 1] >= 0) put(i + kNT, p1, v1x, v1y, v1z, e1);
        if (src[s0 + 2] >= 0) put(i + 2 * kNT, p2, v2x, v2y, v2z, e2);
        if (src[s0 + 3] >= 0) put(i + 3 * kNT, p3, v3x, v3y, v3z, e3);
    };
    static_assert(kSlots % 4 == 0, "slots come in fours");
#pragma unroll
    for (int s0 = 0; s0 < kSlots; s0 += 4) move4(s0);
    uint32_t tot;
    block_excl_scan<kNT>(nb, s_tmp, tot);
    if (tid == 0) a.bcnt[blockIdx.x] = tot;
}

// Occupied-cell list.  The tile prefix is re-derived by every block from the
// per-tile boundary counts (a few thousand at most).  Items are wave-striped
// (slot s, thread t -> particle 4096 tile + 256 s + t, coalesced) and ranked
// by ballots: the (slot, wave, lane) order is the particle order.
__global__ __launch_bounds__(kNT) void k_cells(ArrangeArgs a) {
    __shared__ uint32_t s_tmp[8];
    __shared__ uint32_t s_base[kCellTile / kNT][kNT / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int kSlots = kCellTile / kNT;
    uint32_t pre = 0;
    for (int t = tid; t < (int)blockIdx.x; t += kNT) pre += a.bcnt[t];
    {
        uint32_t tot;
        block_excl_scan<kNT>(pre, s_tmp, tot);
        pre = tot;
    }
    const int t0 = blockIdx.x * kCellTile;
    uint32_t bits = 0;
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
        const int i = t0 + s * kNT + tid;
        bool b = false;
        if (i < a.n) b = i == 0 || (a.hash[i - 1] >> 6) != (a.hash[i] >> 6);
        bits |= (b ? 1u : 0u) << s;
        const uint64_t m = __ballot(b);
        if (lane == 0) s_base[s][wave] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    static_assert(kSlots * (kNT / 64) == 64, "one wave scans the slot counts");
    if (wave == 0) {   // (slot, wave) order is particle order: one wave scan
        const uint32_t v = (&s_base[0][0])[lane];
        uint32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        (&s_base[0][0])[lane] = pre + x - v;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
        const int i = t0 + s * kNT + tid;
        const bool b = (bits >> s) & 1u;
        const uint64_t m = __ballot(b);
        if (i >= a.n) continue;
        const int cell = (int)(a.hash[i] >> 6);
        if (b) {
            const int k = (int)(s_base[s][wave] + (uint32_t)__popcll(m & lanemask_lt()));
            const int prev = i == 0 ? -1 : (int)(a.hash[i - 1] >> 6);
            a.cells[k] = cell;
            a.cstart[k] = i;
            // interior chunks whose first cell lies in (prev, cell] start at k;
            // chunks before the first / after the last occupied cell are
            // resolved by k_fill itself
            if (prev >= 0)
                for (int q = prev / kFillChunk + 1; q <= cell / kFillChunk; ++q) a.chunk_k0[q] = k;
        }
        if (i == a.n - 1) {
            // the last particle closes the list (it sees every boundary before it)
            const int K = (int)(s_base[s][wave] + (uint32_t)__popcll(m & lanemask_lt())) + (b ? 1 : 0);
            a.cstart[K] = a.n;
            a.dc[kCntK] = K;
        }
    }
}

// First index j in [lo, hi) whose micro x layer is >= L (keys ascend there).
__device__ __forceinline__ int layer_bound(const uint32_t *hash, int lo, int hi,
                                           int L) {
    const uint32_t th = (uint32_t)L << 4;
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if ((hash[m] & 63u) < th)
            lo = m + 1;
        else
            hi = m;
    }
    return lo;
}

// Dense begin[numc+1] (a cell's count is begin[c+1] - begin[c]; no dense
// count array is written).
// Block q fills cells [4096 q, 4096 q + 4096) in 4 rounds of 1024; each
// thread owns 4 consecutive cells (16-byte stores).  The occupied cells of a
// round are marked in a 1024-bit LDS map; a thread's cursor into the
// occupied-cell list is the map's population below its cells.
__global__ __launch_bounds__(kNT) void k_fill(ArrangeArgs a) {
    constexpr int kRound = 4 * kNT;
    __shared__ uint32_t s_bits[kRound / 32];
    __shared__ uint32_t s_pre[kRound / 32 + 1];
    const int tid = threadIdx.x;
    const int q = blockIdx.x;
    const int K = a.dc[kCntK];
    int kcur;
    {
        const int qb = q * kFillChunk;
        if (K == 0 || qb <= a.cells[0])
            kcur = 0;                      // at or before the first occupied cell
        else if (qb > a.cells[K - 1])
            kcur = K;                      // after the last one
        else
            kcur = a.chunk_k0[q];          // written by k_cells
    }
    // empty chunk (most of a sparsely filled box): begin is the next
    // occupied cell's start - no map, no barriers
    const int chunk_end = min(q * kFillChunk + kFillChunk, a.numc);
    if (kcur >= K || a.cells[kcur] >= chunk_end) {
        const int bg = kcur < K ? a.cstart[kcur] : a.n;
        const int4 b4 = make_int4(bg, bg, bg, bg);
        for (int c0 = q * kFillChunk + 4 * tid; c0 < chunk_end; c0 += 4 * kNT) {
            if (c0 + 3 < a.numc) {
                *reinterpret_cast<int4 *>(a.begin + c0) = b4;
            } else {
                for (int u = 0; u < 4 && c0 + u < a.numc; ++u) a.begin[c0 + u] = bg;
            }
        }
        if (q == a.nchunks - 1 && tid == 0) a.begin[a.numc] = a.n;
        return;
    }
    for (int it = 0; it < kFillChunk / kRound; ++it) {
        const int base = q * kFillChunk + it * kRound;
        if (base >= a.numc) break;
        if (tid < kRound / 32) s_bits[tid] = 0;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int kk = kcur + tid + kNT * j;
            if (kk < K) {
                const int cid = a.cells[kk];
                if (cid < base + kRound) {
                    const int off = cid - base;
                    atomicOr(&s_bits[off >> 5], 1u << (off & 31));
                }
            }
        }
        __syncthreads();
        if (tid < 64) {   // prefix of the 32 word populations (first wave)
            const uint32_t pc = tid < kRound / 32 ? (uint32_t)__popc(s_bits[tid]) : 0u;
            uint32_t x = pc;
#pragma unroll
            for (int o = 1; o < 32; o <<= 1) {
                const uint32_t y = __shfl_up(x, o, 64);
                if (tid >= o) x += y;
            }
            if (tid < kRound / 32) s_pre[tid] = x - pc;
            if (tid == kRound / 32 - 1) s_pre[kRound / 32] = x;
        }
        __syncthreads();
        const int w = tid >> 3, sh = (tid & 7) * 4;
        const uint32_t word = s_bits[w];
        const uint32_t nib = (word >> sh) & 15u;
        int k = kcur + (int)s_pre[w] + __popc(word & ((1u << sh) - 1u));
        const int c0 = base + 4 * tid;
        int bg[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const bool occ = (nib >> u) & 1u;
            bg[u] = k < K ? a.cstart[k] : a.n;
            if (occ && c0 + u < a.numc) {
                const int cnt = a.cstart[k + 1] - bg[u];
                if (cnt > 65535) atomicOr(&a.dc[kCntErr], kErrCapacity);
            }
            k += occ ? 1 : 0;
        }
        if (c0 + 3 < a.numc) {
            *reinterpret_cast<int4 *>(a.begin + c0) = make_int4(bg[0], bg[1], bg[2], bg[3]);
        } else {
            for (int u = 0; u < 4; ++u)
                if (c0 + u < a.numc) a.begin[c0 + u] = bg[u];
        }
        kcur += (int)s_pre[kRound / 32];
        __syncthreads();
    }
    if (q == a.nchunks - 1 && tid == 0) a.begin[a.numc] = a.n;
}

// knCalculateBlockRequirementHybridMode (src/io/gpu_model.cu:137-181): SMS
// tasks of a cell from its count and the counts of its 6 face neighbours.
// split_mode 1 replaces the (15, 96) thresholds by tuned ones.
template <bool kBatch>
__device__ __forceinline__ int cell_req(const ArrangeArgs &a, int c, int n) {
    if (n == 0) return 0;
    const SceneGeom g = a.scenes[scene_of_cell<kBatch>(a.scenes, a.nscenes, c)];
    const int local = c - g.cell_base;
    const int slice = g.gx * g.gy;
    const int nc = slice * g.gz;
    const int x = local % g.gx;
    const int y = (local / g.gx) % g.gy;
    auto count = [&](int d) { return a.begin[d + 1] - a.begin[d]; };
    int tot = n;
    if (x > 0) tot += count(c - 1);
    if (x + 1 < g.gx) tot += count(c + 1);
    if (y > 0) tot += count(c - g.gx);
    if (y + 1 < g.gy) tot += count(c + g.gx);
    if (local >= slice) tot += count(c - slice);
    if (local + slice < nc) tot += count(c + slice);
    const int min_self = a.split_mode ? a.sms_min_self : 15;
    const int min_tot = a.split_mode ? a.sms_min_total : 96;
    if (tot < min_tot && n < min_self) return 0;
    if (tot < 60) return (n + 27) >> 5;
    return (n + 31) >> 5;
}

// x-layer starts of an occupied cell: the number of its particles (sorted
// keys hash[pb, pb + n)) whose micro x layer is < 1, < 2, < 3, packed
// 16 + 16 | 16 bits.  Small cells are counted over their keys, 8 loads in
// flight at a time; large ones by binary search.
__device__ __forceinline__ uint2 x_layer_starts(const uint32_t *hash, int pb, int n) {
    int l1 = 0, l2 = 0, l3 = 0;
    if (n <= 32) {
        for (int j0 = 0; j0 < n; j0 += 8) {
            uint32_t h[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) h[u] = j0 + u < n ? hash[pb + j0 + u] : 0xffffffffu;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t L = (h[u] >> 4) & 3u;
                const bool in = j0 + u < n;
                l1 += in && L < 1u;
                l2 += in && L < 2u;
                l3 += in && L < 3u;
            }
        }
    } else {
        l1 = layer_bound(hash, pb, pb + n, 1) - pb;
        l2 = layer_bound(hash, pb + l1, pb + n, 2) - pb;
        l3 = layer_bound(hash, pb + l2, pb + n, 3) - pb;
    }
    return make_uint2((uint32_t)l1 | ((uint32_t)l2 << 16), (uint32_t)l3);
}

template <bool kBatch>
__global__ __launch_bounds__(kNT) void k_split_count(ArrangeArgs a) {
    __shared__ uint32_t s_tmp[8];
    const int K = a.dc[kCntK];
    if (blockIdx.x * kSplitTile >= K) return;
    uint32_t st = 0, ss = 0, sr = 0;
    for (int s = 0; s < kSplitTile / kNT; ++s) {
        const int k = blockIdx.x * kSplitTile + s * kNT + threadIdx.x;
        if (k >= K) break;
        const int c = a.cells[k];
        const int pb = a.cstart[k];
        const int n = a.cstart[k + 1] - pb;
        a.xcum[c] = x_layer_starts(a.hash, pb, n);
        const int r = cell_req<kBatch>(a, c, n);
        a.req[c] = r;
        const int nsms = min(n, r * kSmsTaskParticles);
        st += (uint32_t)(n - nsms);
        ss += (uint32_t)nsms;
        sr += (uint32_t)r;
    }
    uint32_t tot;
    block_excl_scan<kNT>(st, s_tmp, tot);
    if (threadIdx.x == 0) a.split_sum[blockIdx.x] = tot;
    block_excl_scan<kNT>(ss, s_tmp, tot);
    if (threadIdx.x == 0) a.split_sum[a.split_tiles + blockIdx.x] = tot;
    block_excl_scan<kNT>(sr, s_tmp, tot);
    if (threadIdx.x == 0) a.split_sum[2 * a.split_tiles + blockIdx.x] = tot;
}

// Task bounds of the 32 particles starting at offset po of a cell
// (knArrangeSmsTasks, sph_arrangement.cu:910-956).
__device__ __forceinline__ void task_bounds(const uint32_t *hash, int pb, int pc,
                                            int po, int8_t b[6]) {
    const uint32_t fh = hash[pb + po];
    const uint32_t lh = hash[pb + min(po + kSmsTaskParticles - 1, pc - 1)];
    const int min_x = (fh >> 4) & 3, max_x = (lh >> 4) & 3;
    const int min_z = (fh >> 2) & 3, max_z = (lh >> 2) & 3;
    const int min_y = fh & 3, max_y = lh & 3;
    b[0] = (int8_t)min_x;
    b[1] = (int8_t)max_x;
    if (min_x == max_x) {
        b[4] = (int8_t)min_z;
        b[5] = (int8_t)max_z;
        if (min_z == max_z) {
            b[2] = (int8_t)min_y;
            b[3] = (int8_t)max_y;
        } else {
            b[2] = 0;
            b[3] = 3;
        }
    } else {
        b[2] = 0;
        b[3] = 3;
        b[4] = 0;
        b[5] = 3;
    }
}

__device__ __forceinline__ uint16_t pack_bounds(int x0, int x1, int y0, int y1,
                                                int z0, int z1) {
    return (uint16_t)(x0 | (x1 << 2) | (y0 << 4) | (y1 << 6) | (z0 << 8) |
                      (z1 << 10));
}

// Per occupied cell: req, TRA/SMS bases, and the cell's BlockTasks.  Pairs
// (2k, 2k+1) inside one cell get the judgeTask treatment here
// (sph_arrangement.cu:966-1022); pairs across cells stay isSame = 0 and the
// odd tail is fixed by k_compact once the task total is known.
template <bool kBatch>
__global__ __launch_bounds__(kNT) void k_split_apply(ArrangeArgs a) {
    __shared__ uint32_t s_tmp[8];
    __shared__ uint32_t s_pre[3];
    const int K = a.dc[kCntK];
    if (blockIdx.x * kSplitTile >= K) return;
    const int tid = threadIdx.x;
    for (int ch = 0; ch < 3; ++ch) {
        uint32_t v = 0;
        for (int t = tid; t < (int)blockIdx.x; t += kNT)
            v += a.split_sum[ch * a.split_tiles + t];
        uint32_t tot;
        block_excl_scan<kNT>(v, s_tmp, tot);
        if (tid == 0) s_pre[ch] = tot;
    }
    __syncthreads();
    // each thread owns 4 consecutive occupied cells
    const int k0 = blockIdx.x * kSplitTile + tid * (kSplitTile / kNT);
    int nn[4], rr[4], cc[4], pb4[5];
    uint32_t st = 0, ss = 0, sr = 0;
#pragma unroll
    for (int s = 0; s < 5; ++s) pb4[s] = k0 + s <= K ? a.cstart[k0 + s] : 0;
#pragma unroll
    for (int s = 0; s < 4; ++s) cc[s] = k0 + s < K ? a.cells[k0 + s] : 0;
#pragma unroll
    for (int s = 0; s < 4; ++s) rr[s] = k0 + s < K ? a.req[cc[s]] : 0;   // written by k_split_count
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int k = k0 + s;
        nn[s] = 0;
        if (k < K) {
            nn[s] = pb4[s + 1] - pb4[s];
            const int nsms = min(nn[s], rr[s] * kSmsTaskParticles);
            st += (uint32_t)(nn[s] - nsms);
            ss += (uint32_t)nsms;
            sr += (uint32_t)rr[s];
        }
    }
    uint32_t tot;
    uint32_t tra = s_pre[0] + block_excl_scan<kNT>(st, s_tmp, tot);
    uint32_t sms = s_pre[1] + block_excl_scan<kNT>(ss, s_tmp, tot);
    uint32_t tb = s_pre[2] + block_excl_scan<kNT>(sr, s_tmp, tot);
    for (int s = 0; s < 4; ++s) {
        const int k = k0 + s;
        if (k >= K) break;
        const int c = cc[s];
        const int n = nn[s], r = rr[s];
        const int nsms = min(n, r * kSmsTaskParticles);
        a.tra_base[c] = (int)tra;
        a.sms_base[c] = (int)sms;
        if (r > 0) {
            const SceneGeom g = a.scenes[scene_of_cell<kBatch>(a.scenes, a.nscenes, c)];
            const int local = c - g.cell_base;
            const uint16_t px = (uint16_t)(local % g.gx);
            const uint16_t py = (uint16_t)(local / g.gx % g.gy);
            const uint16_t pz = (uint16_t)(local / g.gx / g.gy);
            const int pb = pb4[s];
            for (int lt = 0; lt < r; ++lt) {
                const int gi = (int)tb + lt;
                if (gi >= a.task_cap) {
                    atomicOr(&a.dc[kCntErr], kErrCapacity);
                    break;
                }
                int8_t b[6];
                task_bounds(a.hash, pb, n, lt * kSmsTaskParticles, b);
                gsph_task t;
                t.is_same = 0;
                t.pad0 = 0;
                t.cellid = c;
                t.cell_pos[0] = px;
                t.cell_pos[1] = py;
                t.cell_pos[2] = pz;
                t.p_offset = (uint16_t)(lt * kSmsTaskParticles);
                t.xxi = b[0];
                t.xxx = b[1];
                t.yyi = b[2];
                t.yyy = b[3];
                t.zzi = b[4];
                t.zzz = b[5];
                t.pad1[0] = 0;
                t.pad1[1] = 0;
                t.cell_begin = pb;
                t.cell_particle_count = n;
                t.paired_bounds = pack_bounds(b[0], b[1], b[2], b[3], b[4], b[5]);
                const int partner = gi ^ 1;
                const bool same = partner >= (int)tb && partner < (int)tb + r;
                if (same) {
                    t.is_same = 1;
                    if ((gi & 1) == 0) {
                        int8_t b2[6];
                        task_bounds(a.hash, pb, n, (lt + 1) * kSmsTaskParticles, b2);
                        int zlo = b[4], zhi = b[5];
                        if (b[0] == b2[1]) {
                            zhi = b2[5];
                        } else {
                            zlo = 0;
                            zhi = 3;
                        }
                        t.paired_bounds = pack_bounds(b[0], b2[1], b[2], b[3], zlo, zhi);
                    }
                }
                a.tasks[gi] = t;
            }
        }
        tra += (uint32_t)(n - nsms);
        sms += (uint32_t)nsms;
        tb += (uint32_t)r;
        if (k == K - 1) {
            a.dc[kCntT] = (int)tb;
            a.dc[kCntTra] = (int)tra;
            a.dc[kCntSms] = (int)sms;
        }
    }
}

// compact[] (CountingSort_Result_two9, sph_arrangement.cu:193-206): TRA
// particles first in cell order, then SMS particles in cell order; inside a
// bucket the sorted order.  Block 0 also applies judgeTask's odd-tail rule
// and writes the reference's d_middle_value_.
__global__ __launch_bounds__(kNT) void k_compact(ArrangeArgs a) {
    const int i0 = blockIdx.x * (kNT * kCompactItems) + threadIdx.x;
    const int total_tra = a.dc[kCntTra];
    // kCompactItems particles per thread, each step's loads issued together
    int c[kCompactItems], bg[kCompactItems], rq[kCompactItems], sb[kCompactItems],
        tb[kCompactItems];
#pragma unroll
    for (int u = 0; u < kCompactItems; ++u) {
        const int i = i0 + u * kNT;
        c[u] = i < a.n ? (int)(a.hash[i] >> 6) : 0;
    }
#pragma unroll
    for (int u = 0; u < kCompactItems; ++u) {
        const bool in = i0 + u * kNT < a.n;
        bg[u] = in ? a.begin[c[u]] : 0;
        rq[u] = in ? a.req[c[u]] : 0;
        sb[u] = in ? a.sms_base[c[u]] : 0;
        tb[u] = in ? a.tra_base[c[u]] : 0;
    }
#pragma unroll
    for (int u = 0; u < kCompactItems; ++u) {
        const int i = i0 + u * kNT;
        if (i < a.n) {
            const int rank = i - bg[u];
            const int r32 = rq[u] * kSmsTaskParticles;
            if (rank < r32)
                a.compact[total_tra + sb[u] + rank] = i;
            else
                a.compact[tb[u] + rank - r32] = i;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const int T = a.dc[kCntT];
        if ((T & 1) && T < a.task_cap) {
            a.tasks[T - 1].is_same = 1;
            gsph_task pad = a.tasks[T - 1];
            pad.p_offset = (uint16_t)(pad.p_offset + kSmsTaskParticles);
            a.tasks[T] = pad;
        }
        a.dc[kCntMiddle] = a.dc[kCntSms] > 0 ? total_tra : -1;
        a.dc[kCntSmsPool] = 0;   // density allocates this step's SMS masks afresh
    }
}

__global__ void k_empty_grid(ArrangeArgs a) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < a.nchunks) a.chunk_k0[q] = 0;
    if (q == 0) {
        a.cstart[0] = 0;
        a.dc[kCntK] = 0;
        a.dc[kCntT] = 0;
        a.dc[kCntTra] = 0;
        a.dc[kCntSms] = 0;
        a.dc[kCntMiddle] = -1;
        a.dc[kCntSmsPool] = 0;
    }
}

}  // namespace

void launch_hash(const ArrangeArgs &a, bool batch, hipStream_t s) {
    if (batch)
        hipLaunchKernelGGL(k_hash<true>, dim3(a.sort_tiles), dim3(kNT), 0, s, a);
    else
        hipLaunchKernelGGL(k_hash<false>, dim3(a.sort_tiles), dim3(kNT), 0, s, a);
}

void launch_upsweep(const ArrangeArgs &a, int pass, hipStream_t s) {
    hipLaunchKernelGGL(k_upsweep, dim3(a.sort_tiles), dim3(kNT), 0, s, a, pass);
}

void launch_rowscan(const ArrangeArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_rowscan, dim3(256), dim3(kNT), 0, s, a);
}

void launch_downsweep(const ArrangeArgs &a, int pass, hipStream_t s) {
    if (pass == 0)
        hipLaunchKernelGGL(k_downsweep<true>, dim3(a.sort_tiles), dim3(kNT), 0, s, a, pass);
    else
        hipLaunchKernelGGL(k_downsweep<false>, dim3(a.sort_tiles), dim3(kNT), 0, s, a, pass);
}

void launch_gather(const ArrangeArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_gather, dim3(a.cell_tiles), dim3(kNT), 0, s, a);
}

void launch_cells(const ArrangeArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_cells, dim3(a.cell_tiles), dim3(kNT), 0, s, a);
}

void launch_fill(const ArrangeArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_fill, dim3(a.nchunks), dim3(kNT), 0, s, a);
}

void launch_split_count(const ArrangeArgs &a, bool batch, hipStream_t s) {
    if (batch)
        hipLaunchKernelGGL(k_split_count<true>, dim3(a.split_tiles), dim3(kNT), 0, s, a);
    else
        hipLaunchKernelGGL(k_split_count<false>, dim3(a.split_tiles), dim3(kNT), 0, s, a);
}

void launch_split_apply(const ArrangeArgs &a, bool batch, hipStream_t s) {
    if (batch)
        hipLaunchKernelGGL(k_split_apply<true>, dim3(a.split_tiles), dim3(kNT), 0, s, a);
    else
        hipLaunchKernelGGL(k_split_apply<false>, dim3(a.split_tiles), dim3(kNT), 0, s, a);
}

void launch_compact(const ArrangeArgs &a, hipStream_t s) {
    const int per_block = kNT * kCompactItems;
    const int blocks = a.n > 0 ? (a.n + per_block - 1) / per_block : 1;
    hipLaunchKernelGGL(k_compact, dim3(blocks), dim3(kNT), 0, s, a);
}

void launch_empty_grid(const ArrangeArgs &a, hipStream_t s) {
    const int blocks = (a.nchunks + kNT - 1) / kNT;
    hipLaunchKernelGGL(k_empty_grid, dim3(blocks > 0 ? blocks : 1), dim3(kNT), 0, s, a);
}

}  // namespace gsph
