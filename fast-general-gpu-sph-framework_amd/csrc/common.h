// common.h - device-side building blocks shared by the arrange and physics
// kernels: scene geometry, the bit-exact micro hash, wave64 ballot helpers and
// block-level scans.  Device-wide scans are reduce-then-scan (per-tile totals
// written by one kernel, prefixes formed by the next), except the radix
// passes over the moved particles, whose tiles look back at their
// predecessors' published counts with a bounded wait (k_msort).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gsph.h"

namespace gsph {

constexpr int kWave = 64;
constexpr int kSmsTaskParticles = 32;     // reference: src/core/sph_utils.cuh:35
constexpr float kFloatSmall = 1e-12f;     // reference: src/core/cuda_math.cuh:14

// Error bits collected in DevCounters::err and reported by the host.
constexpr int kErrOutOfDomain = 1;
constexpr int kErrCapacity = 4;
constexpr int kErrSortStall = 8;     // a look-back wait reached its bound

// Slots of the small device counter block each context keeps.
enum CounterSlot {
    kCntK = 0,          // occupied cells
    kCntT = 1,          // SMS tasks
    kCntTra = 2,        // TRA particles
    kCntSms = 3,        // SMS particles
    kCntMiddle = 4,     // reference d_middle_value_ (-1 without SMS)
    kCntErr = 5,
    kCntSmsPool = 6,    // rows of the SMS mask pool taken this step
    kCntAnyDefer = 7,   // the arrange epoch in which the lean density pass last left a tile
    kCntTicket = 8,     // 8..11: tile tickets of the mover radix passes 0..3
    kCntMovers = 12,    // particles whose key changed this step
    kCntOrder = 13,     // 1: the slot order is the order of hash[] (set by integrate)
    kCntEpoch = 14,     // arranges so far (tags the sort's look-back words)
    kCntKeysReady = 15, // 1: next_keys holds the micro ids of the slot order's state
    kCntForceAnyDefer = 16,  // the same for the first lean force pass
    kCntForceAnyDefer2 = 17, // and for the second (windowed) one
    kCntAnySms = 18,    // the arrange epoch in which some occupied cell last had SMS tasks
    kCntSlots = 19,
};

// Geometry of one scene inside the (possibly batched) virtual grid.  In the
// plain single-scene path there is exactly one, with cell_base = 0.
struct SceneGeom {
    int gx, gy, gz;
    int cell_base;       // first global cell id of the scene
    int p_begin;         // first sorted particle of the scene
    int p_count;
    float bound_min[3];
    float bound_max[3];
};

// ParticlePos2CellPosM / CellPos2CellIdxM (reference src/core/sph_utils.cuh:53-105):
// micro coordinate floor(p * (4 * (1/h))) per axis; coarse = micro >> 2;
// id = coarse << 6 | (mx&3) << 4 | (mz&3) << 2 | (my&3).  Returns false when
// the particle lies outside the scene grid (or is not finite).
__device__ __forceinline__ bool micro_coords(float x, float y, float z,
                                             float inv_cell_size,
                                             const SceneGeom &g, int &mx,
                                             int &my, int &mz) {
    const float rat = 4.f * inv_cell_size;
    const float fx = floorf(__fmul_rn(x, rat));
    const float fy = floorf(__fmul_rn(y, rat));
    const float fz = floorf(__fmul_rn(z, rat));
    // bitwise ands: one mask, no branches
    const bool ok = (fx >= 0.f) & (fy >= 0.f) & (fz >= 0.f) & (fx < (float)(g.gx * 4)) &
                    (fy < (float)(g.gy * 4)) & (fz < (float)(g.gz * 4));
    mx = ok ? (int)fx : 0;
    my = ok ? (int)fy : 0;
    mz = ok ? (int)fz : 0;
    return ok;
}

__device__ __forceinline__ uint32_t micro_key(int mx, int my, int mz,
                                              const SceneGeom &g) {
    const int cell = g.cell_base + (mx >> 2) + g.gx * ((my >> 2) + g.gy * (mz >> 2));
    const int idi = (my & 3) + (((mz & 3) + ((mx & 3) << 2)) << 2);
    return ((uint32_t)cell << 6) | (uint32_t)idi;
}

// ---- wave64 helpers --------------------------------------------------------
__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Inclusive scan over the 64 lanes of a wave (all lanes active) with DPP, in
// registers: Hillis-Steele over row_shr 1, 2, 4, 8 inside each row of 16
// lanes (lanes whose source falls below the row read 0), then lane 15 of
// rows 0 and 2 broadcast into rows 1 and 3 (row_bcast:15) and lane 31 into
// rows 2 and 3 (row_bcast:31); rows outside the row mask add 0.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true);   // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true);   // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true);   // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true);   // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}

// Block-wide exclusive scan of one uint32 per thread (blockDim.x = NT, a
// multiple of 64).  Returns the exclusive prefix, writes the block total.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *s_tmp,
                                                    uint32_t &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t x = wave_incl_scan(v);
    if (lane == 63) s_tmp[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int w = 0; w < NT / 64; ++w) {
            const uint32_t t = s_tmp[w];
            s_tmp[w] = run;
            run += t;
        }
        s_tmp[NT / 64] = run;
    }
    __syncthreads();
    total = s_tmp[NT / 64];
    const uint32_t r = s_tmp[wave] + x - v;
    __syncthreads();
    return r;
}

}  // namespace gsph
