// oracle.cpp - CPU restatement of the reference hybrid SPH step.
//
// TEST INFRASTRUCTURE ONLY.  Loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker (or the timed CPU baseline);
// the product library never links or calls it.
//
// Three parts:
//   1. Integer arrange stage, restated literally from the reference so that it
//      is bit-exact by construction: micro hash, micro counting sort with ranks
//      taken in index order (the one stable outcome of the reference's
//      atomics), cell info, split heuristic, TRA/SMS bucket sort, middle value,
//      task generation and judgeTask pairing.
//   2. FP64 density / force / integration over the exact neighbour set.
//   3. An FP32 OpenMP step that follows the reference's TRA traversal; it is
//      the CPU baseline that bench.py times.
// Each function cites the reference file:line it follows.
//
// Build: see oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off).

#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr float kFloatSmall = 1e-12f;  // src/core/cuda_math.cuh:14
constexpr int kSmsTaskParticles = 32;   // src/core/sph_utils.cuh:35

struct Grid {
    int gx, gy, gz;
    long long numc;
};

Grid grid_of(const orc_constants &c) {
    Grid g{c.grid_size[0], c.grid_size[1], c.grid_size[2], 0};
    g.numc = (long long)g.gx * g.gy * g.gz;
    return g;
}

// ParticlePos2CellPosM + CellPos2CellIdxM (src/core/sph_utils.cuh:53-105).
// Returns -1 (kInvalidCellIdx) outside the grid or for a non-finite /
// negative coordinate (the reference's float->ushort conversion is undefined
// there; the product reports such particles as out of domain).
inline int micro_hash(float x, float y, float z, float inv_cell_size,
                      const Grid &g) {
    const float rat = 4.f * inv_cell_size;
    const float fx = std::floor(x * rat);
    const float fy = std::floor(y * rat);
    const float fz = std::floor(z * rat);
    if (!(fx >= 0.f && fy >= 0.f && fz >= 0.f)) return -1;
    if (!(fx < (float)(g.gx * 4) && fy < (float)(g.gy * 4) &&
          fz < (float)(g.gz * 4)))
        return -1;
    const int mx = (int)fx, my = (int)fy, mz = (int)fz;
    const int idc = (mx >> 2) + g.gx * ((my >> 2) + g.gy * (mz >> 2));
    const int idi = (my & 3) + (((mz & 3) + ((mx & 3) << 2)) << 2);
    return (idc << 6) | idi;
}

// packSmsTaskBounds (src/particle/particle_buffer.h:89-95).
inline uint16_t pack_bounds(int min_x, int max_x, int min_y, int max_y,
                            int min_z, int max_z) {
    return (uint16_t)(min_x | (max_x << 2) | (min_y << 4) | (max_y << 6) |
                      (min_z << 8) | (max_z << 10));
}

}  // namespace

// Definitions take C linkage from the declarations in oracle.h.

// defaultInitializeSPHSysPara (src/simulation/sph_hybrid_system.cpp:143-193):
// derived coefficients evaluated in double with M_PI/pow, stored as float.
int orc_derive(const orc_params *p, orc_constants *c) {
    if (!p || !c) return -1;
    std::memset(c, 0, sizeof(*c));
    const float h = p->kernel_h;
    c->kernel_h = h;
    c->kernel_2 = h * h;
    c->cell_size = h;
    c->inv_cell_size = 1.0f / c->cell_size;
    for (int a = 0; a < 3; ++a) {
        c->world_size[a] = p->world_size[a];
        c->grid_size[a] = (int)std::ceil(p->world_size[a] / c->cell_size);
        c->gravity[a] = p->gravity[a];
    }
    c->mass = p->mass;
    c->rest_density = p->rest_density;
    c->gas_constant = p->gas_constant;
    c->viscosity = p->viscosity;
    c->time_step = p->time_step;
    c->wall_damping = p->wall_damping;
    c->surface_normal = p->surface_normal;
    c->surface_coe = p->surface_coe;
    const double hd = (double)h;
    c->poly6_value = (float)(315.0 / (64.0 * M_PI * std::pow(hd, 9)));
    c->spiky_value = (float)(-45.0 / (M_PI * std::pow(hd, 6)));
    c->visco_value = (float)(45.0 / (M_PI * std::pow(hd, 6)));
    c->grad_poly6 = (float)(-945 / (32 * M_PI * std::pow(hd, 9)));
    c->lplc_poly6 = (float)(945 / (8 * M_PI * std::pow(hd, 9)));
    c->self_density =
        (float)((double)(c->mass * c->poly6_value) * std::pow(hd, 6));
    c->inv_rest_density = 1.0f / c->rest_density;
    c->density_scale = c->mass * c->poly6_value;
    c->half_spiky_value = c->spiky_value * 0.5f;
    c->viscosity_visco_value = c->viscosity * c->visco_value;
    c->grad_color_scale = c->grad_poly6 * c->mass;
    c->lplc_color_scale = c->lplc_poly6 * c->mass;
    for (int a = 0; a < 3; ++a) {
        c->bound_min[a] = h;
        c->bound_max[a] = p->world_size[a] - h;
    }
    return 0;
}

// SplitMix64 (Steele, Lea, Flood 2014; the constants of Vigna's public
// reference implementation).
uint64_t orc_splitmix64(uint64_t *state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Synthetic jittered lattice block: corner + (i + 1/2) * spacing in double,
// jitter (2u - 1) * jitter_frac * spacing per axis with u a 24-bit uniform,
// axis order x, y, z; particle order x outer, z inner (the order of
// initializeScene's loops, sph_hybrid_system.cpp:454-462).
int orc_scene_block(const float corner[3], int nx, int ny, int nz,
                    float spacing, float jitter_frac, uint64_t seed,
                    float *pos4) {
    uint64_t st = seed;
    const double s = spacing;
    long long k = 0;
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            for (int iz = 0; iz < nz; ++iz) {
                const int idx[3] = {ix, iy, iz};
                for (int a = 0; a < 3; ++a) {
                    const double u =
                        (double)(orc_splitmix64(&st) >> 40) * (1.0 / 16777216.0);
                    const double base = (double)corner[a] + (idx[a] + 0.5) * s;
                    const double jit = (2.0 * u - 1.0) * (double)jitter_frac * s;
                    pos4[4 * k + a] = (float)(base + jit);
                }
                pos4[4 * k + 3] = 0.f;
                ++k;
            }
    return (int)k;
}

// initializeScene lattice loops (sph_hybrid_system.cpp:441-464): FP32
// accumulation x += kernel * interval, x outer, z inner, w = 0.
int orc_scene_lattice(const float begin[3], const float end[3], float kernel_h,
                      float interval, float *pos4, int capacity) {
    const float step = kernel_h * interval;
    int k = 0;
    for (float x = begin[0]; x < end[0]; x += step)
        for (float y = begin[1]; y < end[1]; y += step)
            for (float z = begin[2]; z < end[2]; z += step) {
                if (pos4 && k < capacity) {
                    pos4[4 * k + 0] = x;
                    pos4[4 * k + 1] = y;
                    pos4[4 * k + 2] = z;
                    pos4[4 * k + 3] = 0.f;
                }
                ++k;
            }
    return k;
}

int orc_hash(const orc_params *p, int n, const float *pos4, int *key) {
    orc_constants c;
    orc_derive(p, &c);
    const Grid g = grid_of(c);
    int bad = 0;
    for (int i = 0; i < n; ++i) {
        key[i] = micro_hash(pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2],
                            c.inv_cell_size, g);
        if (key[i] < 0) ++bad;
    }
    return bad;
}

// knCalculateBlockRequirementHybridMode (src/io/gpu_model.cu:137-181).
int orc_split_req(const int *count, const int grid[3], int *req) {
    const long long gx = grid[0], gy = grid[1], gz = grid[2];
    const long long slice = gx * gy, numc = slice * gz;
    for (long long idx = 0; idx < numc; ++idx) {
        const int self = count[idx];
        if (self == 0) {
            req[idx] = 0;
            continue;
        }
        int tot = self;
        const long long x = idx % gx, y = (idx / gx) % gy;
        if (x > 0) tot += count[idx - 1];
        if (x + 1 < gx) tot += count[idx + 1];
        if (y > 0) tot += count[idx - gx];
        if (y + 1 < gy) tot += count[idx + gx];
        if (idx >= slice) tot += count[idx - slice];
        if (idx + slice < numc) tot += count[idx + slice];
        if (tot < 96 && self < 15)
            req[idx] = 0;
        else if (tot < 60)
            req[idx] = (self + 27) >> 5;
        else
            req[idx] = (self + 31) >> 5;
    }
    return 0;
}

// Arrangement::arrangeHybridFrame (src/grid/sph_arrangement.cu:1493-1506):
//   CountingSort_O_M           :1352-1395 (+ kernels :1309-1351)
//   split heuristic            src/io/gpu_model.cu:137-181
//   CountingSortCUDA_Two9_M    :1459-1482 (+ kernels :87-107, :193-206, :358-384)
//   task offsets ExclusiveSum  :1499
//   knArrangeSmsTasks          :885-965
//   judgeTask                  :966-1022
// The reference ranks particles inside a micro cell / bucket with atomicAdd;
// here the ranks follow the index order, the one stable outcome.
int orc_arrange(const orc_params *p, int n, const float *pos4, int *hash,
                int *perm, int *begin, int *count, int *req, int *compact,
                int *middle_raw, int *num_tasks, orc_task *tasks,
                int task_cap) {
    orc_constants c;
    orc_derive(p, &c);
    const Grid g = grid_of(c);
    const long long numc = g.numc;
    const long long M = numc * 64;
    std::vector<int> key(n), rank(n);
    std::vector<int> cntM(M + 1, 0), offM(M + 1, 0);
    for (int i = 0; i < n; ++i) {
        key[i] = micro_hash(pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2],
                            c.inv_cell_size, g);
        if (key[i] < 0) return -3;
        rank[i] = cntM[key[i]]++;
    }
    long long run = 0;
    for (long long m = 0; m <= M; ++m) {
        offM[m] = (int)run;
        run += cntM[m];
    }
    std::vector<int> shash(n), sperm(n);
    for (int i = 0; i < n; ++i) {
        const int dest = offM[key[i]] + rank[i];
        shash[dest] = key[i];
        sperm[dest] = i;
    }
    std::vector<int> cbeg(numc), ccnt(numc), creq(numc);
    for (long long cc = 0; cc < numc; ++cc) {
        ccnt[cc] = offM[(cc << 6) + 64] - offM[cc << 6];
        cbeg[cc] = offM[cc << 6];
    }
    orc_split_req(ccnt.data(), c.grid_size, creq.data());
    // bucket counting sort: bucket = coarse + numc when SMS, coarse when TRA
    std::vector<int> cnt2(2 * numc, 0), bucket(n), rank2(n);
    for (int q = 0; q < n; ++q) {
        const int coarse = shash[q] >> 6;
        const int off = q - cbeg[coarse];
        const long long b =
            creq[coarse] * kSmsTaskParticles > off ? coarse + numc : coarse;
        bucket[q] = (int)b;
        rank2[q] = cnt2[b]++;
    }
    std::vector<int> incl(2 * numc);
    run = 0;
    for (long long b = 0; b < 2 * numc; ++b) {
        run += cnt2[b];
        incl[b] = (int)run;
    }
    std::vector<int> ccompact(n), cbucket(n);
    for (int q = 0; q < n; ++q) {
        const int b = bucket[q];
        const int start = b == 0 ? 0 : incl[b - 1];
        ccompact[start + rank2[q]] = q;
        cbucket[start + rank2[q]] = b;
    }
    int mid = -1;
    for (int q = 0; q < n; ++q) {
        const int prior = q == 0 ? -1 : cbucket[q - 1];
        if (cbucket[q] != prior && cbucket[q] / numc > 0 &&
            (q == 0 || prior / numc == 0))
            mid = q;
    }
    // task offsets + knArrangeSmsTasks
    std::vector<int> toff(numc);
    run = 0;
    for (long long cc = 0; cc < numc; ++cc) {
        toff[cc] = (int)run;
        run += creq[cc];
    }
    const int T = (int)run;
    std::vector<orc_task> tk(T + 1);
    std::memset(tk.data(), 0, sizeof(orc_task) * (T + 1));
    for (long long cc = 0; cc < numc; ++cc) {
        const int tc = creq[cc];
        if (tc == 0) continue;
        const int pb = cbeg[cc], pc = ccnt[cc];
        const uint16_t px = (uint16_t)(cc % g.gx);
        const uint16_t py = (uint16_t)(cc / g.gx % g.gy);
        const uint16_t pz = (uint16_t)(cc / g.gx / g.gy);
        for (int lt = 0; lt < tc; ++lt) {
            const int po = lt * kSmsTaskParticles;
            const int fh = shash[pb + po];
            const int lo = std::min(po + kSmsTaskParticles - 1, pc - 1);
            const int lh = shash[pb + lo];
            const int min_x = (fh >> 4) & 3, max_x = (lh >> 4) & 3;
            const int min_z = (fh >> 2) & 3, max_z = (lh >> 2) & 3;
            const int min_y = fh & 3, max_y = lh & 3;
            orc_task t;
            std::memset(&t, 0, sizeof(t));
            t.cellid = (int)cc;
            t.cell_pos[0] = px;
            t.cell_pos[1] = py;
            t.cell_pos[2] = pz;
            t.cell_begin = pb;
            t.cell_particle_count = pc;
            t.p_offset = (uint16_t)po;
            t.is_same = 0;
            t.xxi = (int8_t)min_x;
            t.xxx = (int8_t)max_x;
            if (min_x == max_x) {
                t.zzi = (int8_t)min_z;
                t.zzz = (int8_t)max_z;
                if (min_z == max_z) {
                    t.yyi = (int8_t)min_y;
                    t.yyy = (int8_t)max_y;
                } else {
                    t.yyi = 0;
                    t.yyy = 3;
                }
            } else {
                t.zzi = 0;
                t.zzz = 3;
                t.yyi = 0;
                t.yyy = 3;
            }
            t.paired_bounds = pack_bounds(t.xxi, t.xxx, t.yyi, t.yyy, t.zzi, t.zzz);
            tk[toff[cc] + lt] = t;
        }
    }
    // judgeTask
    for (int f = 0; f < T; f += 2) {
        const int s = f + 1;
        if (s >= T) {
            // judgeTask's odd tail (:976-991) sets is_same, cellid,
            // cell_pos, cell_begin, the count and p_offset + 32 of the pad and
            // leaves its other bytes as they were in device memory (stale).
            // Those bytes are defined here as padIndependentSmsTasks
            // (:1024-1032, the build without the historical judgeTask)
            // defines them: a copy of the first task.  The pair statistics
            // read the pad's xxx (compact vs widened).
            tk[f].is_same = 1;
            tk[s] = tk[f];
            tk[s].p_offset = (uint16_t)(tk[f].p_offset + kSmsTaskParticles);
            continue;
        }
        if (tk[f].cellid != tk[s].cellid) {
            tk[f].is_same = 0;
            tk[s].is_same = 0;
            continue;
        }
        tk[f].is_same = 1;
        tk[s].is_same = 1;
        int zlo = tk[f].zzi, zhi = tk[f].zzz;
        if (tk[f].xxi == tk[s].xxx) {
            zhi = tk[s].zzz;
        } else {
            zlo = 0;
            zhi = 3;
        }
        tk[f].paired_bounds =
            pack_bounds(tk[f].xxi, tk[s].xxx, tk[f].yyi, tk[f].yyy, zlo, zhi);
    }
    if (hash) std::memcpy(hash, shash.data(), sizeof(int) * n);
    if (perm) std::memcpy(perm, sperm.data(), sizeof(int) * n);
    if (begin) std::memcpy(begin, cbeg.data(), sizeof(int) * numc);
    if (count) std::memcpy(count, ccnt.data(), sizeof(int) * numc);
    if (req) std::memcpy(req, creq.data(), sizeof(int) * numc);
    if (compact) std::memcpy(compact, ccompact.data(), sizeof(int) * n);
    if (middle_raw) *middle_raw = mid;
    if (num_tasks) *num_tasks = T;
    if (tasks) {
        const int ndesc = T + (T & 1);
        if (ndesc > task_cap) return -4;
        std::memcpy(tasks, tk.data(), sizeof(orc_task) * ndesc);
    }
    return 0;
}

// Arrangement::getSmsTaskPairStats (src/grid/sph_arrangement.cu:1595-1636).
int orc_pair_stats_of(const orc_task *tasks, int T, const int *count,
                      orc_pair_stats *out) {
    std::memset(out, 0, sizeof(*out));
    if (T <= 0) return 0;
    out->total_pairs = (T + 1) / 2;
    for (int t = 0; t < T; t += 2) {
        const orc_task &f = tasks[t];
        const orc_task &s = tasks[t + 1];
        if (f.is_same != 0) {
            ++out->is_same_pairs;
            const int second = count[f.cellid] - s.p_offset;
            if (second >= kSmsTaskParticles)
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
    return 0;
}

// ---------------------------------------------------------------------------
// FP64 physics.  Neighbour set: every j != i with 1e-12 <= r^2 <= h^2
// (the acceptance test of density_kernels.cu:153,172 and force_kernels.cu:
// 397-401,538), found through a cell-size-h grid (27 cells).
namespace {

struct NbrGrid {
    double h;
    int gx, gy, gz;
    std::vector<int> start, items;
};

void build_nbr_grid(const orc_constants &c, int n, const float *pos4,
                    NbrGrid &ng) {
    ng.h = c.kernel_h;
    ng.gx = c.grid_size[0];
    ng.gy = c.grid_size[1];
    ng.gz = c.grid_size[2];
    const long long numc = (long long)ng.gx * ng.gy * ng.gz;
    std::vector<int> cell(n);
    ng.start.assign(numc + 1, 0);
    for (int i = 0; i < n; ++i) {
        int ix = (int)std::floor((double)pos4[4 * i] / ng.h);
        int iy = (int)std::floor((double)pos4[4 * i + 1] / ng.h);
        int iz = (int)std::floor((double)pos4[4 * i + 2] / ng.h);
        ix = std::min(std::max(ix, 0), ng.gx - 1);
        iy = std::min(std::max(iy, 0), ng.gy - 1);
        iz = std::min(std::max(iz, 0), ng.gz - 1);
        cell[i] = ix + ng.gx * (iy + ng.gy * iz);
        ng.start[cell[i] + 1]++;
    }
    for (long long k = 0; k < numc; ++k) ng.start[k + 1] += ng.start[k];
    ng.items.resize(n);
    std::vector<int> fill(ng.start.begin(), ng.start.end() - 1);
    for (int i = 0; i < n; ++i) ng.items[fill[cell[i]]++] = i;
}

template <class F>
void for_neighbours(const NbrGrid &ng, const float *pos4, int i, F &&f) {
    const double xi = pos4[4 * i], yi = pos4[4 * i + 1], zi = pos4[4 * i + 2];
    const double h2 = ng.h * ng.h;
    const int ix = std::min(std::max((int)std::floor(xi / ng.h), 0), ng.gx - 1);
    const int iy = std::min(std::max((int)std::floor(yi / ng.h), 0), ng.gy - 1);
    const int iz = std::min(std::max((int)std::floor(zi / ng.h), 0), ng.gz - 1);
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int x = ix + dx, y = iy + dy, z = iz + dz;
                if (x < 0 || y < 0 || z < 0 || x >= ng.gx || y >= ng.gy ||
                    z >= ng.gz)
                    continue;
                const long long cc = x + (long long)ng.gx * (y + (long long)ng.gy * z);
                for (int k = ng.start[cc]; k < ng.start[cc + 1]; ++k) {
                    const int j = ng.items[k];
                    const double rx = xi - pos4[4 * j];
                    const double ry = yi - pos4[4 * j + 1];
                    const double rz = zi - pos4[4 * j + 2];
                    const double r2 = rx * rx + ry * ry + rz * rz;
                    if (r2 < 1e-12 || r2 > h2) continue;
                    f(j, rx, ry, rz, r2);
                }
            }
}

struct DConst {
    double h, h2, mass, rho0, gas, poly6, self_density, half_spiky, visc_visco,
        grad_scale, lplc_scale, surface_normal, surface_coe, dt;
    double gravity[3], bound_min[3], bound_max[3], wall_damping;
};

// The same derivation as orc_derive but kept in double from the FP32 inputs.
DConst dconst(const orc_params *p) {
    DConst d;
    d.h = p->kernel_h;
    d.h2 = d.h * d.h;
    d.mass = p->mass;
    d.rho0 = p->rest_density;
    d.gas = p->gas_constant;
    d.poly6 = 315.0 / (64.0 * M_PI * std::pow(d.h, 9));
    d.self_density = d.mass * d.poly6 * std::pow(d.h, 6);
    d.half_spiky = -45.0 / (M_PI * std::pow(d.h, 6)) * 0.5;
    d.visc_visco = (double)p->viscosity * 45.0 / (M_PI * std::pow(d.h, 6));
    d.grad_scale = -945.0 / (32 * M_PI * std::pow(d.h, 9)) * d.mass;
    d.lplc_scale = 945.0 / (8 * M_PI * std::pow(d.h, 9)) * d.mass;
    d.surface_normal = p->surface_normal;
    d.surface_coe = p->surface_coe;
    d.dt = p->time_step;
    for (int a = 0; a < 3; ++a) {
        d.gravity[a] = p->gravity[a];
        d.bound_min[a] = (double)p->kernel_h;
        d.bound_max[a] = (double)(p->world_size[a] - p->kernel_h);
    }
    d.wall_damping = p->wall_damping;
    return d;
}

}  // namespace

// Density (density_kernels.cu:160-179 summand, :353-358 / :536-544 epilogue):
// rho = m*poly6*sum (h^2-r^2)^3 + self_density, p = ((rho/rho0)^7 - 1)*gas.
int orc_density_f64(const orc_params *p, int n, const float *pos4, double *rho,
                    double *pressure) {
    orc_constants c;
    orc_derive(p, &c);
    const DConst d = dconst(p);
    NbrGrid ng;
    build_nbr_grid(c, n, pos4, ng);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
        double s = 0;
        for_neighbours(ng, pos4, i,
                       [&](int, double, double, double, double r2) {
                           const double t = d.h2 - r2;
                           s += t * t * t;
                       });
        const double r = s * d.mass * d.poly6 + d.self_density;
        rho[i] = r;
        if (pressure) pressure[i] = (std::pow(r / d.rho0, 7) - 1.0) * d.gas;
    }
    return 0;
}

// O(N^2) density, pins the grid search of orc_density_f64 on small inputs.
int orc_density_brute_f64(const orc_params *p, int n, const float *pos4,
                          double *rho) {
    const DConst d = dconst(p);
    for (int i = 0; i < n; ++i) {
        double s = 0;
        for (int j = 0; j < n; ++j) {
            const double rx = (double)pos4[4 * i] - pos4[4 * j];
            const double ry = (double)pos4[4 * i + 1] - pos4[4 * j + 1];
            const double rz = (double)pos4[4 * i + 2] - pos4[4 * j + 2];
            const double r2 = rx * rx + ry * ry + rz * rz;
            if (r2 < 1e-12 || r2 > d.h2) continue;
            const double t = d.h2 - r2;
            s += t * t * t;
        }
        rho[i] = s * d.mass * d.poly6 + d.self_density;
    }
    return 0;
}

// Force (force_kernels.cu:525-561 summand, :818-852 epilogue).  Inputs:
// positions and evaluated velocities (xyz) of the step, FP64 rho and p.
// acc_scale[i] is the magnitude the FP32 result's rounding error is relative
// to: the same sums taken over |term|, with the pressure of each particle
// replaced by its uncancelled part gas*(rho/rho0)^7 (the "- 1" of the state
// equation cancels, so an FP32 p carries an absolute error relative to that
// part), plus the surface force and its per-neighbour contributions.
int orc_force_f64(const orc_params *p, int n, const float *pos4,
                  const float *ev4, const double *rho, const double *pressure,
                  double *acc, double *acc_scale, double *acc_alt) {
    // the relative bar of the physics parity: an FP32 grad c (FP32 1/rho_j,
    // a few roundings per term and the sum) is within it of the FP64 one
    constexpr double kBranchTol = 1e-5;
    orc_constants c;
    orc_derive(p, &c);
    const DConst d = dconst(p);
    NbrGrid ng;
    build_nbr_grid(c, n, pos4, ng);
    std::vector<double> pbar(n);
    for (int i = 0; i < n; ++i)
        pbar[i] = std::fabs(pressure[i]) + d.gas * std::pow(rho[i] / d.rho0, 7);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
        double P[3] = {0, 0, 0}, V[3] = {0, 0, 0}, G[3] = {0, 0, 0}, L = 0;
        double Ps = 0, Vs = 0, Gs = 0, Ls = 0;
        const double evi[3] = {ev4[4 * i], ev4[4 * i + 1], ev4[4 * i + 2]};
        for_neighbours(ng, pos4, i, [&](int j, double rx, double ry, double rz,
                                        double r2) {
            const double r = std::sqrt(r2);
            const double w = 1.0 / rho[j];
            const double kr = d.h - r;
            const double pw = w * (pressure[i] + pressure[j]) * kr * kr / r;
            P[0] -= rx * pw;
            P[1] -= ry * pw;
            P[2] -= rz * pw;
            Ps += w * (pbar[i] + pbar[j]) * kr * kr;
            const double dv[3] = {ev4[4 * j] - evi[0], ev4[4 * j + 1] - evi[1],
                                  ev4[4 * j + 2] - evi[2]};
            for (int a = 0; a < 3; ++a) V[a] += dv[a] * w * kr;
            Vs += std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]) * w * kr;
            const double t = d.h2 - r2;
            G[0] += rx * w * t * t;
            G[1] += ry * w * t * t;
            G[2] += rz * w * t * t;
            Gs += r * w * t * t;
            L += w * t * (r2 - 0.75 * t);
            Ls += std::fabs(w * t * (r2 - 0.75 * t));
        });
        double F[3];
        for (int a = 0; a < 3; ++a) F[a] = P[a] * d.half_spiky + V[a] * d.visc_visco;
        for (int a = 0; a < 3; ++a) G[a] *= d.grad_scale;
        L *= d.lplc_scale / rho[i];
        const double gn = std::sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]);
        double S[3] = {0, 0, 0};
        if (gn > d.surface_normal)
            for (int a = 0; a < 3; ++a) S[a] = G[a] * d.surface_coe * L / gn;
        for (int a = 0; a < 3; ++a) acc[3 * i + a] = F[a] * d.mass + S[a];
        if (acc_alt) {
            const double margin = kBranchTol * std::fabs(d.grad_scale) * Gs;
            const bool near = std::fabs(gn - d.surface_normal) <= margin && gn > 0;
            for (int a = 0; a < 3; ++a) {
                // the other branch: surface force off if it was on, and on if off
                const double s_alt =
                    gn > d.surface_normal ? 0.0 : G[a] * d.surface_coe * L / gn;
                acc_alt[3 * i + a] = near ? F[a] * d.mass + s_alt : acc[3 * i + a];
            }
        }
        if (acc_scale) {
            // surface force G coe L / |G|: its per-neighbour contributions are
            // coe |L_j| (the Laplacian sum cancels between r < 0.655 h and
            // r > 0.655 h) and coe |L| |G_j| / |G| (its direction)
            double smag = std::sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
            if (gn > d.surface_normal)
                smag += d.surface_coe * (Ls * d.lplc_scale / rho[i] +
                                         std::fabs(L) * Gs * std::fabs(d.grad_scale) / gn);
            acc_scale[i] = d.mass * (std::fabs(d.half_spiky) * Ps +
                                     d.visc_visco * Vs) +
                           smag;
        }
    }
    return 0;
}

// knIntegrateVelocityE (src/solver/integration_kernels.cu:235-336) in FP64.
int orc_integrate_f64(const orc_params *p, int n, const float *pos4,
                      const float *vel3, const float *ev4, const double *acc,
                      const double *rho, double *pos_out, double *vel_out,
                      double *ev_out) {
    const DConst d = dconst(p);
    const double alpha = 20000000, beta = 200000;
    // knIntegrateVelocityE: float kernel = kDevSysPara.kernel + 0.01 (double sum)
    const double kern = (double)(float)((double)p->kernel_h + 0.01);
    const double acc_limit = 3000000, vel_limit = 36;
    for (int i = 0; i < n; ++i) {
        double x[3] = {pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2]};
        double v[3] = {vel3[3 * i], vel3[3 * i + 1], vel3[3 * i + 2]};
        double a[3] = {acc[3 * i], acc[3 * i + 1], acc[3 * i + 2]};
        const int order[3] = {1, 0, 2};  // the reference checks y, then x, then z
        for (int k = 0; k < 3; ++k) {
            const int ax = order[k];
            double diff = kern - (x[ax] - d.bound_min[ax]);
            if (diff > 0.0) a[ax] += alpha * diff - beta * v[ax];
            diff = kern - (d.bound_max[ax] - x[ax]);
            if (diff > 0.0) a[ax] -= alpha * diff + beta * v[ax];
        }
        const double w = 1.0 / rho[i];
        for (int ax = 0; ax < 3; ++ax) a[ax] = a[ax] * w + d.gravity[ax];
        double sp = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
        if (sp > acc_limit) {
            const double f = std::sqrt(acc_limit / sp);
            for (int ax = 0; ax < 3; ++ax) a[ax] *= f;
        }
        for (int ax = 0; ax < 3; ++ax) v[ax] += a[ax] * d.dt;
        sp = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        if (sp > vel_limit) {
            const double f = std::sqrt(vel_limit / sp);
            for (int ax = 0; ax < 3; ++ax) v[ax] *= f;
        }
        for (int ax = 0; ax < 3; ++ax) x[ax] += v[ax] * d.dt;
        for (int ax = 0; ax < 3; ++ax) {
            if (x[ax] >= d.bound_max[ax]) {
                v[ax] *= d.wall_damping;
                x[ax] = d.bound_max[ax] - kern;
            }
            if (x[ax] <= d.bound_min[ax]) {
                v[ax] *= d.wall_damping;
                x[ax] = d.bound_min[ax] + kern;
            }
        }
        for (int ax = 0; ax < 3; ++ax) {
            pos_out[4 * i + ax] = x[ax];
            vel_out[3 * i + ax] = v[ax];
            ev_out[4 * i + ax] = (v[ax] + ev4[4 * i + ax]) * 0.5;
        }
        pos_out[4 * i + 3] = w;
        ev_out[4 * i + 3] = ev4[4 * i + 3];
    }
    return 0;
}

// ---------------------------------------------------------------------------
// FP32 reference step, OpenMP.  Arrange: micro hash and a stable sort by micro
// id (coarse counting sort, then micro order inside each cell; the same
// permutation as the reference's micro counting sort without its dense
// 64*C table).  Density and force: the reference's TRA traversal for every
// particle (9 rows, x pruned to micro layers; findTraNeighborParticleRange,
// kernel_common.cuh:69-112; knComputeCellDensityTRA9 density_kernels.cu:
// 160-179; knComputeCellForceTRA9 force_kernels.cu:525-561).  Integrate:
// knIntegrateVelocityE.
namespace {

inline float powf7(float b) { return b * b * b * b * b * b * b; }

struct F32Step {
    orc_constants c;
    Grid g;
    int n;
    std::vector<int> key, tmp_key, order, cnt, beg;   // per cell: beg[C+1]
    std::vector<uint16_t> xl;                          // per particle-cell: layer starts
    std::vector<float> pos2, vel2, ev2;
};

void f32_arrange(F32Step &S, float *pos4, float *vel3, float *ev4) {
    const int n = S.n;
    const long long C = S.g.numc;
    S.key.resize(n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        S.key[i] = micro_hash(pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2],
                              S.c.inv_cell_size, S.g);
    S.beg.assign(C + 1, 0);
    for (int i = 0; i < n; ++i) S.beg[(S.key[i] >> 6) + 1]++;
    for (long long k = 0; k < C; ++k) S.beg[k + 1] += S.beg[k];
    S.order.resize(n);
    {
        std::vector<int> fill(S.beg.begin(), S.beg.end() - 1);
        for (int i = 0; i < n; ++i) S.order[fill[S.key[i] >> 6]++] = i;
    }
    // stable order by micro sub-id inside each occupied cell
#pragma omp parallel for schedule(dynamic, 1024)
    for (long long k = 0; k < C; ++k) {
        const int b = S.beg[k], e = S.beg[k + 1];
        if (e - b < 2) continue;
        std::stable_sort(S.order.begin() + b, S.order.begin() + e,
                         [&](int a, int bb) { return S.key[a] < S.key[bb]; });
    }
    S.pos2.resize(4 * (size_t)n);
    S.vel2.resize(3 * (size_t)n);
    S.ev2.resize(4 * (size_t)n);
    S.tmp_key.resize(n);
#pragma omp parallel for schedule(static)
    for (int q = 0; q < n; ++q) {
        const int i = S.order[q];
        std::memcpy(&S.pos2[4 * (size_t)q], pos4 + 4 * (size_t)i, 16);
        std::memcpy(&S.vel2[3 * (size_t)q], vel3 + 3 * (size_t)i, 12);
        std::memcpy(&S.ev2[4 * (size_t)q], ev4 + 4 * (size_t)i, 16);
        S.tmp_key[q] = S.key[i];
    }
    std::memcpy(pos4, S.pos2.data(), 16 * (size_t)n);
    std::memcpy(vel3, S.vel2.data(), 12 * (size_t)n);
    std::memcpy(ev4, S.ev2.data(), 16 * (size_t)n);
    S.key.swap(S.tmp_key);
}

// #particles of cell c with micro x < L, from the sorted keys (L in 0..4).
inline int layer_start(const F32Step &S, long long c, int L) {
    const int b = S.beg[c], e = S.beg[c + 1];
    if (L <= 0) return b;
    if (L >= 4) return e;
    int lo = b, hi = e;
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (((S.key[m] >> 4) & 3) < L)
            lo = m + 1;
        else
            hi = m;
    }
    return lo;
}

// findTraNeighborParticleRange (kernel_common.cuh:69-112).
inline void tra_row(const F32Step &S, int cx, int cy, int cz, int lmx, int dy,
                    int dz, int &b, int &e) {
    const int ny = cy + dy, nz = cz + dz;
    b = e = 0;
    if (ny < 0 || ny >= S.g.gy || nz < 0 || nz >= S.g.gz) return;
    const long long mid = cx + (long long)S.g.gx * (ny + (long long)S.g.gy * nz);
    b = cx > 0 ? layer_start(S, mid - 1, lmx) : S.beg[mid];
    e = cx + 1 < S.g.gx ? layer_start(S, mid + 1, lmx + 1) : S.beg[mid + 1];
}

void f32_density(F32Step &S, float *pos4, float *ev4) {
    const orc_constants &c = S.c;
#pragma omp parallel for schedule(dynamic, 512)
    for (int i = 0; i < S.n; ++i) {
        const float x = pos4[4 * i], y = pos4[4 * i + 1], z = pos4[4 * i + 2];
        const int k = S.key[i];
        const int cc = k >> 6;
        const int cx = cc % S.g.gx, cy = cc / S.g.gx % S.g.gy,
                  cz = cc / S.g.gx / S.g.gy;
        const int lmx = (k >> 4) & 3;
        float s = 0.f;
        for (int r = 0; r < 9; ++r) {
            int b, e;
            tra_row(S, cx, cy, cz, lmx, r % 3 - 1, r / 3 - 1, b, e);
            for (int j = b; j < e; ++j) {
                const float rx = x - pos4[4 * j], ry = y - pos4[4 * j + 1],
                            rz = z - pos4[4 * j + 2];
                const float d2 = rx * rx + ry * ry + rz * rz;
                if (d2 < kFloatSmall || d2 > c.kernel_2) continue;
                const float t = c.kernel_2 - d2;
                s += t * t * t;
            }
        }
        float rho = s * c.density_scale + c.self_density;
        ev4[4 * i + 3] = (powf7(rho * c.inv_rest_density) - 1) * c.gas_constant;
        S.pos2[4 * (size_t)i + 3] = 1.0f / rho;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < S.n; ++i) pos4[4 * i + 3] = S.pos2[4 * (size_t)i + 3];
}

void f32_force(F32Step &S, const float *pos4, const float *ev4, float *acc3) {
    const orc_constants &c = S.c;
#pragma omp parallel for schedule(dynamic, 512)
    for (int i = 0; i < S.n; ++i) {
        const float x = pos4[4 * i], y = pos4[4 * i + 1], z = pos4[4 * i + 2];
        const float wi = pos4[4 * i + 3];
        const float evx = ev4[4 * i], evy = ev4[4 * i + 1], evz = ev4[4 * i + 2],
                    pi = ev4[4 * i + 3];
        const int k = S.key[i];
        const int cc = k >> 6;
        const int cx = cc % S.g.gx, cy = cc / S.g.gx % S.g.gy,
                  cz = cc / S.g.gx / S.g.gy;
        const int lmx = (k >> 4) & 3;
        float P[3] = {0, 0, 0}, V[3] = {0, 0, 0}, G[3] = {0, 0, 0}, L = 0;
        for (int r = 0; r < 9; ++r) {
            int b, e;
            tra_row(S, cx, cy, cz, lmx, r % 3 - 1, r / 3 - 1, b, e);
            for (int j = b; j < e; ++j) {
                const float rx = x - pos4[4 * j], ry = y - pos4[4 * j + 1],
                            rz = z - pos4[4 * j + 2];
                const float d2 = rx * rx + ry * ry + rz * rz;
                if (d2 < kFloatSmall || d2 > c.kernel_2) continue;
                const float inv = 1.0f / std::sqrt(d2);
                const float dist = d2 * inv;
                const float Vj = pos4[4 * j + 3];
                const float kr = c.kernel_h - dist;
                const float tp = Vj * (pi + ev4[4 * j + 3]) * kr * kr;
                P[0] -= rx * (tp * inv);
                P[1] -= ry * (tp * inv);
                P[2] -= rz * (tp * inv);
                const float tv = Vj * kr;
                V[0] += (ev4[4 * j] - evx) * tv;
                V[1] += (ev4[4 * j + 1] - evy) * tv;
                V[2] += (ev4[4 * j + 2] - evz) * tv;
                const float t = c.kernel_2 - d2;
                const float tg = Vj * t * t;
                G[0] += rx * tg;
                G[1] += ry * tg;
                G[2] += rz * tg;
                L += Vj * t * (d2 - 0.75f * t);
            }
        }
        float F[3];
        for (int a = 0; a < 3; ++a)
            F[a] = P[a] * c.half_spiky_value + V[a] * c.viscosity_visco_value;
        for (int a = 0; a < 3; ++a) G[a] *= c.grad_color_scale;
        L *= c.lplc_color_scale;
        L *= wi;
        const float gn = std::sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]);
        for (int a = 0; a < 3; ++a) {
            float sf = 0.f;
            if (gn > c.surface_normal) sf = G[a] * c.surface_coe * L / gn;
            acc3[3 * i + a] = F[a] * c.mass + sf;
        }
    }
}

void f32_integrate(const F32Step &S, float *pos4, float *vel3, float *ev4,
                   const float *acc3) {
    const orc_constants &c = S.c;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < S.n; ++i) {
        float x[3] = {pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2]};
        const float w = pos4[4 * i + 3];
        float v[3] = {vel3[3 * i], vel3[3 * i + 1], vel3[3 * i + 2]};
        float a[3] = {acc3[3 * i], acc3[3 * i + 1], acc3[3 * i + 2]};
        const float alpha = 20000000, beta = 200000;
        const float kern = (float)((double)c.kernel_h + 0.01);
        const int order[3] = {1, 0, 2};
        for (int k = 0; k < 3; ++k) {
            const int ax = order[k];
            float diff = kern - (x[ax] - c.bound_min[ax]);
            if (diff > 0.0f) a[ax] += alpha * diff - beta * v[ax];
            diff = kern - (c.bound_max[ax] - x[ax]);
            if (diff > 0.0f) a[ax] -= alpha * diff + beta * v[ax];
        }
        for (int ax = 0; ax < 3; ++ax) a[ax] = a[ax] * w + c.gravity[ax];
        float sp = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
        if (sp > 3000000.f) {
            const float f = std::sqrt(3000000.f / sp);
            for (int ax = 0; ax < 3; ++ax) a[ax] *= f;
        }
        for (int ax = 0; ax < 3; ++ax) v[ax] += a[ax] * c.time_step;
        sp = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        if (sp > 36.f) {
            const float f = std::sqrt(36.f / sp);
            for (int ax = 0; ax < 3; ++ax) v[ax] *= f;
        }
        for (int ax = 0; ax < 3; ++ax) x[ax] += v[ax] * c.time_step;
        for (int ax = 0; ax < 3; ++ax) {
            if (x[ax] >= c.bound_max[ax]) {
                v[ax] *= c.wall_damping;
                x[ax] = c.bound_max[ax] - kern;
            }
            if (x[ax] <= c.bound_min[ax]) {
                v[ax] *= c.wall_damping;
                x[ax] = c.bound_min[ax] + kern;
            }
        }
        for (int ax = 0; ax < 3; ++ax) {
            pos4[4 * i + ax] = x[ax];
            vel3[3 * i + ax] = v[ax];
            ev4[4 * i + ax] = (v[ax] + ev4[4 * i + ax]) * 0.5f;
        }
    }
}

}  // namespace

int orc_step_f32(const orc_params *p, int n, float *pos4, float *vel3,
                 float *ev4, float *acc3, int nsteps, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    F32Step S;
    orc_derive(p, &S.c);
    S.g = grid_of(S.c);
    S.n = n;
    for (int s = 0; s < nsteps; ++s) {
        f32_arrange(S, pos4, vel3, ev4);
        for (int i = 0; i < n; ++i)
            if (S.key[i] < 0) return -3;
        f32_density(S, pos4, ev4);
        f32_force(S, pos4, ev4, acc3);
        f32_integrate(S, pos4, vel3, ev4, acc3);
    }
    return 0;
}

// computeBenchmarkStateChecksum (sph_hybrid_system.cpp:43-62), sequential.
int orc_checksum(int n, const float *pos4, double out[5]) {
    double s[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
        const float *q = pos4 + 4 * (size_t)i;
        s[0] += q[0];
        s[1] += q[1];
        s[2] += q[2];
        s[3] += q[3];
        s[4] += (double)((i & 1023u) + 1u) * (q[0] + 3.0 * q[1] + 7.0 * q[2]);
    }
    for (int k = 0; k < 5; ++k) out[k] = s[k];
    return 0;
}

