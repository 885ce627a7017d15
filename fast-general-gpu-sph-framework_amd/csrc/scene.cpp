// scene.cpp - host-side scene input: default parameters, derived constants,
// the reference scene JSON format and the synthetic workloads.
//
// Compiled with -ffp-contract=off: the lattice generator reproduces the
// reference's FP32 accumulation (x += kernel * interval) bit for bit.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "gsph.h"
#include "host_util.h"

namespace {

// ---- a small JSON reader (objects, arrays, numbers, strings, literals) -----
struct JVal {
    enum Kind { Null, Num, Str, Bool, Arr, Obj } kind = Null;
    double num = 0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal *get(const char *k) const {
        for (auto &kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct JParser {
    const char *p, *e;
    bool ok = true;
    void ws() {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool lit(const char *s) {
        const size_t n = std::strlen(s);
        if ((size_t)(e - p) >= n && std::strncmp(p, s, n) == 0) {
            p += n;
            return true;
        }
        return false;
    }
    std::string string() {
        std::string out;
        ++p;  // opening quote
        while (p < e && *p != '"') {
            if (*p == '\\' && p + 1 < e) {
                ++p;
                const char c = *p;
                out += c == 'n' ? '\n' : c == 't' ? '\t' : c;
            } else {
                out += *p;
            }
            ++p;
        }
        if (p >= e) {
            ok = false;
            return out;
        }
        ++p;
        return out;
    }
    JVal value(int depth) {
        JVal v;
        ws();
        if (p >= e || depth > 64) {
            ok = false;
            return v;
        }
        if (*p == '{') {
            v.kind = JVal::Obj;
            ++p;
            ws();
            if (p < e && *p == '}') {
                ++p;
                return v;
            }
            while (ok) {
                ws();
                if (p >= e || *p != '"') {
                    ok = false;
                    break;
                }
                std::string k = string();
                ws();
                if (p >= e || *p != ':') {
                    ok = false;
                    break;
                }
                ++p;
                v.obj.emplace_back(k, value(depth + 1));
                ws();
                if (p < e && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < e && *p == '}') {
                    ++p;
                    break;
                }
                ok = false;
            }
        } else if (*p == '[') {
            v.kind = JVal::Arr;
            ++p;
            ws();
            if (p < e && *p == ']') {
                ++p;
                return v;
            }
            while (ok) {
                v.arr.push_back(value(depth + 1));
                ws();
                if (p < e && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < e && *p == ']') {
                    ++p;
                    break;
                }
                ok = false;
            }
        } else if (*p == '"') {
            v.kind = JVal::Str;
            v.str = string();
        } else if (lit("true")) {
            v.kind = JVal::Bool;
            v.num = 1;
        } else if (lit("false")) {
            v.kind = JVal::Bool;
        } else if (lit("null")) {
            v.kind = JVal::Null;
        } else {
            char *end = nullptr;
            v.num = std::strtod(p, &end);
            if (end == p) {
                ok = false;
                return v;
            }
            v.kind = JVal::Num;
            p = end;
        }
        return v;
    }
};

// jsoncpp's asFloat is a double -> float cast of the parsed number.
float as_float(const JVal *v, float dflt) {
    return (v && (v->kind == JVal::Num || v->kind == JVal::Bool)) ? (float)v->num : dflt;
}

uint64_t splitmix64(uint64_t &state) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// One jittered lattice point per axis: (corner + (i + 1/2) s) + (2u - 1) f s,
// evaluated in double and rounded once to float.
inline float lattice_coord(uint64_t &st, float corner, int i, float s, float f) {
    const double u = (double)(splitmix64(st) >> 40) * (1.0 / 16777216.0);
    const double base = (double)corner + (i + 0.5) * (double)s;
    const double jit = (2.0 * u - 1.0) * (double)f * (double)s;
    return (float)(base + jit);
}

// Synthetic physics shared by the configs: h = 2 spacing, rho0 = 1000,
// g = -9.81, dt = 1e-4, m = rho0 s^3, the reference's other constants.
void synthetic_params(gsph_params *p, double spacing, float wx, float wy, float wz) {
    gsph_params_default(p);
    p->kernel_h = 2.0f * (float)spacing;
    p->rest_density = 1000.0f;
    p->mass = (float)(1000.0 * spacing * spacing * spacing);
    p->gravity[0] = 0.0f;
    p->gravity[1] = -9.81f;
    p->gravity[2] = 0.0f;
    p->time_step = 1e-4f;
    p->world_size[0] = wx;
    p->world_size[1] = wy;
    p->world_size[2] = wz;
}

}  // namespace

extern "C" {

const char *gsph_version(void) { return "gsph-mi355x 0.1"; }

// Reference defaults: defaultInitializeSPHSysPara (sph_hybrid_system.cpp:
// 136-193), Scene defaults (sph_hybrid_system.h:37-44) and the world of
// assets/scene_default.json.
void gsph_params_default(gsph_params *out) {
    if (!out) return;
    std::memset(out, 0, sizeof(*out));
    out->kernel_h = 0.03f;
    out->mass = 0.02f;
    out->rest_density = 1000.0f;
    out->gas_constant = 1.0f;
    out->viscosity = 6.5f;
    out->time_step = 0.003f;
    out->gravity[1] = -9.8f;
    out->wall_damping = -0.5f;
    out->surface_normal = 0.1f;
    out->surface_coe = 0.2f;
    out->world_size[0] = out->world_size[1] = out->world_size[2] = 1.6f;
}

void gsph_options_default(gsph_options *out) {
    if (!out) return;
    out->split_mode = 0;
    out->sms_min_self = 15;
    out->sms_min_total = 96;
    out->blocks_per_cu = 32;
    out->pair_masks = 1;
    out->row_prune = 1;
    out->stage_depth = 128;
    out->sms_mask_bytes = 1024;
    out->sms_local_merge = 1;
    out->sms_legacy_staging = 0;
    out->same_cell_pair_force = 0;
    out->lean_passes = 1;
}

int gsph_derive_constants(const gsph_params *p, gsph_constants *c) {
    if (!p || !c) return gsph::fail(nullptr, GSPH_ERR_INVALID_ARG, "null argument");
    // h^2 >= 1e-12: below it the reference's pair test accepts nothing
    if (!(p->kernel_h >= 1e-6f) || !(p->kernel_h * p->kernel_h >= 1e-12f))
        return gsph::fail(nullptr, GSPH_ERR_INVALID_ARG, "kernel_h must be >= 1e-6");
    std::memset(c, 0, sizeof(*c));
    const float h = p->kernel_h;
    c->kernel_h = h;
    c->kernel_2 = h * h;
    c->cell_size = h;
    c->inv_cell_size = 1.0f / h;
    for (int a = 0; a < 3; ++a) {
        c->world_size[a] = p->world_size[a];
        const float cells = p->world_size[a] / c->cell_size;
        c->grid_size[a] = (int)std::ceil(cells);
        c->gravity[a] = p->gravity[a];
        c->bound_min[a] = h;
        c->bound_max[a] = p->world_size[a] - h;
    }
    c->mass = p->mass;
    c->rest_density = p->rest_density;
    c->gas_constant = p->gas_constant;
    c->viscosity = p->viscosity;
    c->time_step = p->time_step;
    c->wall_damping = p->wall_damping;
    c->surface_normal = p->surface_normal;
    c->surface_coe = p->surface_coe;
    // kernel coefficients: double arithmetic (M_PI, pow), stored as float
    const double hd = h;
    const double h6 = std::pow(hd, 6), h9 = std::pow(hd, 9);
    c->poly6_value = (float)(315.0 / (64.0 * M_PI * h9));
    c->spiky_value = (float)(-45.0 / (M_PI * h6));
    c->visco_value = (float)(45.0 / (M_PI * h6));
    c->grad_poly6 = (float)(-945.0 / (32.0 * M_PI * h9));
    c->lplc_poly6 = (float)(945.0 / (8.0 * M_PI * h9));
    c->self_density = (float)((double)(c->mass * c->poly6_value) * h6);
    c->inv_rest_density = 1.0f / c->rest_density;
    c->density_scale = c->mass * c->poly6_value;
    c->half_spiky_value = c->spiky_value * 0.5f;
    c->viscosity_visco_value = c->viscosity * c->visco_value;
    c->grad_color_scale = c->grad_poly6 * c->mass;
    c->lplc_color_scale = c->lplc_poly6 * c->mass;
    return GSPH_OK;
}

int gsph_scene_block(const float corner[3], int nx, int ny, int nz, float spacing,
                     float jitter_frac, uint64_t seed, float *pos4) {
    if (!corner || nx < 0 || ny < 0 || nz < 0 || !pos4)
        return gsph::fail(nullptr, GSPH_ERR_INVALID_ARG, "bad block arguments");
    uint64_t st = seed;
    size_t k = 0;
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            for (int iz = 0; iz < nz; ++iz) {
                pos4[4 * k + 0] = lattice_coord(st, corner[0], ix, spacing, jitter_frac);
                pos4[4 * k + 1] = lattice_coord(st, corner[1], iy, spacing, jitter_frac);
                pos4[4 * k + 2] = lattice_coord(st, corner[2], iz, spacing, jitter_frac);
                pos4[4 * k + 3] = 0.f;
                ++k;
            }
    return GSPH_OK;
}

// readSceneFromJsonFile + initializeScene (sph_hybrid_system.cpp:69-126,
// 432-473): lattice loops x += kernel * interval over each fluid block.
int gsph_scene_json(const char *path, gsph_params *params, int *count, float *pos4,
                    int capacity) {
    if (!path || !params || !count)
        return gsph::fail(nullptr, GSPH_ERR_INVALID_ARG, "null argument");
    std::ifstream in(path, std::ios::binary);
    if (!in) return gsph::fail(nullptr, GSPH_ERR_IO, std::string("can not open ") + path);
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string text = ss.str();
    JParser jp{text.data(), text.data() + text.size()};
    const JVal root = jp.value(0);
    if (!jp.ok || root.kind != JVal::Obj)
        return gsph::fail(nullptr, GSPH_ERR_IO, std::string("can not parse scene file ") + path);
    gsph_params p;
    gsph_params_default(&p);
    p.world_size[0] = as_float(root.get("xx"), 0.f);
    p.world_size[1] = as_float(root.get("yy"), 0.f);
    p.world_size[2] = as_float(root.get("zz"), 0.f);
    p.mass = as_float(root.get("mass"), 0.02f);
    const float interval = as_float(root.get("interval"), 0.5f);
    p.kernel_h = as_float(root.get("kernel_h"), p.kernel_h);
    p.gravity[1] = as_float(root.get("gravity_y"), p.gravity[1]);
    p.time_step = as_float(root.get("time_step"), p.time_step);
    p.rest_density = as_float(root.get("rest_density"), p.rest_density);
    const float step = p.kernel_h * interval;
    if (!(step > 0.f)) return gsph::fail(nullptr, GSPH_ERR_INVALID_ARG, "non-positive lattice step");
    long long k = 0;
    const JVal *blocks = root.get("fluid_block");
    if (blocks && blocks->kind == JVal::Arr) {
        for (const JVal &b : blocks->arr) {
            const float bx = as_float(b.get("begin_x"), 0.f), by = as_float(b.get("begin_y"), 0.f),
                        bz = as_float(b.get("begin_z"), 0.f);
            const float ex = as_float(b.get("end_x"), 0.f), ey = as_float(b.get("end_y"), 0.f),
                        ez = as_float(b.get("end_z"), 0.f);
            for (float x = bx; x < ex; x += step)
                for (float y = by; y < ey; y += step)
                    for (float z = bz; z < ez; z += step) {
                        if (pos4 && k < capacity) {
                            pos4[4 * k + 0] = x;
                            pos4[4 * k + 1] = y;
                            pos4[4 * k + 2] = z;
                            pos4[4 * k + 3] = 0.f;
                        }
                        ++k;
                        if (k > 0x7fffffffLL)
                            return gsph::fail(nullptr, GSPH_ERR_CAPACITY, "scene too large");
                    }
        }
    }
    *params = p;
    *count = (int)k;
    if (pos4 && k > capacity)
        return gsph::fail(nullptr, GSPH_ERR_CAPACITY, "pos4 buffer too small");
    return GSPH_OK;
}

// Synthetic workloads (seed 20261203 + config):
//  0: 32^3 block, spacing 0.01, corner 0.06, box 0.64^3 (grid 32^3).
//  1: 158^3 column, spacing 0.01, box 6.32^3 (4x the column width, grid
//     316^3), column at x = y = 0.06 (resting on the floor, -x side),
//     centred in z; it collapses into the empty +x half.
//  2: pool 256x64x256 at spacing 0.01 (corner 0.06) plus a sphere of radius
//     0.49 m at spacing 0.005 whose bottom is 1 m above the pool surface;
//     box 2.68 x 2.80 x 2.68.  One global mass (rho0 * 0.01^3).
int gsph_scene_config(int config, gsph_params *params, int *count, float *pos4,
                      int capacity) {
    if (!params || !count) return gsph::fail(nullptr, GSPH_ERR_INVALID_ARG, "null argument");
    const uint64_t seed = 20261203ull + (uint64_t)config;
    const float jit = 0.01f;
    if (config == 0 || config == 1) {
        const int m = config == 0 ? 32 : 158;
        const float world = config == 0 ? 0.64f : 6.32f;
        synthetic_params(params, 0.01, world, world, world);
        const float zc = config == 0 ? 0.06f : (float)((6.32 - 1.58) / 2.0);
        const float corner[3] = {0.06f, 0.06f, zc};
        *count = m * m * m;
        if (!pos4) return GSPH_OK;
        if (capacity < *count) return gsph::fail(nullptr, GSPH_ERR_CAPACITY, "pos4 buffer too small");
        return gsph_scene_block(corner, m, m, m, 0.01f, jit, seed, pos4);
    }
    if (config == 2) {
        synthetic_params(params, 0.01, 2.68f, 2.80f, 2.68f);
        const int px = 256, py = 64, pz = 256;
        const long long npool = (long long)px * py * pz;
        const double r = 0.49, s = 0.005;
        const double cx = 1.34, cy = 0.70 + 1.0 + r, cz = 1.34;
        const int m = (int)std::ceil(2 * r / s);
        const double c0x = cx - r, c0y = cy - r, c0z = cz - r;
        long long nsph = 0;
        for (int ix = 0; ix < m; ++ix)
            for (int iy = 0; iy < m; ++iy)
                for (int iz = 0; iz < m; ++iz) {
                    const double dx = c0x + (ix + 0.5) * s - cx, dy = c0y + (iy + 0.5) * s - cy,
                                 dz = c0z + (iz + 0.5) * s - cz;
                    if (dx * dx + dy * dy + dz * dz < r * r) ++nsph;
                }
        *count = (int)(npool + nsph);
        if (!pos4) return GSPH_OK;
        if (capacity < *count) return gsph::fail(nullptr, GSPH_ERR_CAPACITY, "pos4 buffer too small");
        const float corner[3] = {0.06f, 0.06f, 0.06f};
        int rc = gsph_scene_block(corner, px, py, pz, 0.01f, jit, seed, pos4);
        if (rc) return rc;
        uint64_t st = seed + 1000003ull;
        long long k = npool;
        for (int ix = 0; ix < m; ++ix)
            for (int iy = 0; iy < m; ++iy)
                for (int iz = 0; iz < m; ++iz) {
                    const double dx = c0x + (ix + 0.5) * s - cx, dy = c0y + (iy + 0.5) * s - cy,
                                 dz = c0z + (iz + 0.5) * s - cz;
                    if (!(dx * dx + dy * dy + dz * dz < r * r)) continue;
                    pos4[4 * k + 0] = lattice_coord(st, (float)c0x, ix, (float)s, jit);
                    pos4[4 * k + 1] = lattice_coord(st, (float)c0y, iy, (float)s, jit);
                    pos4[4 * k + 2] = lattice_coord(st, (float)c0z, iz, (float)s, jit);
                    pos4[4 * k + 3] = 0.f;
                    ++k;
                }
        return GSPH_OK;
    }
    return gsph::fail(nullptr, GSPH_ERR_INVALID_ARG, "unknown config (0, 1 or 2)");
}

}  // extern "C"
