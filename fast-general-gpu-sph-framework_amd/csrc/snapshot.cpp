// snapshot.cpp - checkpoint / resume of a context's state to a file.
//
// The reference has no snapshot format; its state moves only between device
// and host buffers (ParticleBufferObject::transfer, particle_buffer.cpp:
// 369-381).  A snapshot here is what restarting a simulation needs: the
// physics parameters, the launch options, the scene layout (box and
// particle count of every scene), the completed step count and the state
// (pos4 = xyz + 1/rho, vel3, ev4 = evaluated velocity + pressure) in the
// context's current slot order.  A context restored from it continues
// bit-exactly: the next arrange sorts the same keys from the same slot order.
//
// File layout (little-endian), all sizes in bytes:
//   0   char[8]  "GSPHSNAP"
//   8   u32      format version (1)
//   12  u32      header size (64)
//   16  u32      sizeof(gsph_params)     20  u32  sizeof(gsph_options)
//   24  u32      flags (bit 0: batched context)
//   28  u32      scene count S
//   32  u64      particle count n        40  u64  completed steps
//   48  u64      FNV-1a 64 of every byte after the header
//   56  u64      0
//   64  gsph_params, gsph_options, S x {f32 world[3], i32 count},
//       pos4 n x 16, vel3 n x 12, ev4 n x 16
// Host-only code: it uses the public ABI and plain file I/O.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gsph.h"
#include "host_util.h"

namespace {

constexpr char kMagic[8] = {'G', 'S', 'P', 'H', 'S', 'N', 'A', 'P'};
constexpr uint32_t kVersion = 1;
constexpr uint32_t kHeader = 64;

struct Header {
    char magic[8];
    uint32_t version, header_bytes, params_bytes, options_bytes, flags, nscenes;
    uint64_t n, steps, hash, zero;
};
static_assert(sizeof(Header) == kHeader, "snapshot header is 64 bytes");

struct SceneRec {
    float world[3];
    int32_t count;
};

uint64_t fnv1a(uint64_t h, const void *p, size_t len) {
    const auto *b = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < len; ++i) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

constexpr uint64_t kFnvBasis = 14695981039346656037ull;

struct File {
    FILE *f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

int io_fail(const std::string &msg) { return gsph::fail(nullptr, GSPH_ERR_IO, msg); }

// Reads and validates a snapshot; fills the layout and, when state is
// non-null, the payload.
int read_snapshot(const char *path, Header &h, gsph_params *params, gsph_options *opt,
                  std::vector<SceneRec> *scenes, std::vector<float> *pos4,
                  std::vector<float> *vel3, std::vector<float> *ev4) {
    if (!path) return gsph::fail(nullptr, GSPH_ERR_INVALID_ARG, "null path");
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) return io_fail(std::string("cannot open ") + path);
    if (std::fread(&h, sizeof(h), 1, fh.f) != 1) return io_fail("truncated snapshot header");
    if (std::memcmp(h.magic, kMagic, 8) != 0) return io_fail("not a gsph snapshot (bad magic)");
    if (h.version != kVersion || h.header_bytes != kHeader)
        return io_fail("unsupported snapshot version");
    if (h.params_bytes != sizeof(gsph_params) || h.options_bytes != sizeof(gsph_options))
        return io_fail("snapshot written by an incompatible library (struct sizes differ)");
    if (h.nscenes == 0 || h.nscenes > (1u << 24) || h.n > 0x7fffffffull)
        return io_fail("snapshot layout out of range");
    const uint64_t body = sizeof(gsph_params) + sizeof(gsph_options) +
                          (uint64_t)h.nscenes * sizeof(SceneRec) + h.n * (16 + 12 + 16);
    if (std::fseek(fh.f, 0, SEEK_END) != 0) return io_fail("cannot seek snapshot");
    const long len = std::ftell(fh.f);
    if (len < 0 || (uint64_t)len != kHeader + body)
        return io_fail("snapshot size does not match its header");
    if (std::fseek(fh.f, kHeader, SEEK_SET) != 0) return io_fail("cannot seek snapshot");
    gsph_params p;
    gsph_options o;
    std::vector<SceneRec> sc(h.nscenes);
    std::vector<float> a(4 * h.n), b(3 * h.n), c(4 * h.n);
    bool ok = std::fread(&p, sizeof(p), 1, fh.f) == 1 && std::fread(&o, sizeof(o), 1, fh.f) == 1 &&
              std::fread(sc.data(), sizeof(SceneRec), sc.size(), fh.f) == sc.size();
    if (ok && h.n) {
        ok = std::fread(a.data(), 16, h.n, fh.f) == h.n && std::fread(b.data(), 12, h.n, fh.f) == h.n &&
             std::fread(c.data(), 16, h.n, fh.f) == h.n;
    }
    if (!ok) return io_fail("truncated snapshot body");
    uint64_t hash = kFnvBasis;
    hash = fnv1a(hash, &p, sizeof(p));
    hash = fnv1a(hash, &o, sizeof(o));
    hash = fnv1a(hash, sc.data(), sc.size() * sizeof(SceneRec));
    hash = fnv1a(hash, a.data(), 16 * h.n);
    hash = fnv1a(hash, b.data(), 12 * h.n);
    hash = fnv1a(hash, c.data(), 16 * h.n);
    if (hash != h.hash) return io_fail("snapshot checksum mismatch (corrupted file)");
    uint64_t total = 0;
    for (const auto &s : sc) {
        if (s.count < 0) return io_fail("negative scene count in snapshot");
        total += (uint64_t)s.count;
    }
    if (total != h.n) return io_fail("scene counts do not add up to the particle count");
    if (params) *params = p;
    if (opt) *opt = o;
    if (scenes) *scenes = std::move(sc);
    if (pos4) *pos4 = std::move(a);
    if (vel3) *vel3 = std::move(b);
    if (ev4) *ev4 = std::move(c);
    return GSPH_OK;
}

}  // namespace

extern "C" {

int gsph_snapshot_save(gsph_ctx *ctx, const char *path) {
    if (!ctx || !path) return gsph::fail(ctx, GSPH_ERR_INVALID_ARG, "null argument");
    int n = 0, ns = 0;
    long long steps = 0;
    gsph_params p;
    gsph_options o;
    int rc = gsph_num_particles(ctx, &n);
    if (!rc) rc = gsph_scene_count(ctx, &ns);
    if (!rc) rc = gsph_step_count(ctx, &steps);
    if (!rc) rc = gsph_get_params(ctx, &p);
    if (!rc) rc = gsph_get_options(ctx, &o);
    if (rc) return rc;
    std::vector<SceneRec> sc(ns);
    for (int s = 0; s < ns; ++s) {
        int first = 0, count = 0;
        rc = gsph_scene_world(ctx, s, sc[s].world);
        if (!rc) rc = gsph_scene_range(ctx, s, &first, &count);
        if (rc) return rc;
        sc[s].count = count;
    }
    std::vector<float> a(4 * (size_t)n), b(3 * (size_t)n), c(4 * (size_t)n);
    rc = gsph_read_state(ctx, a.data(), b.data(), c.data(), nullptr);
    if (rc) return rc;
    Header h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.header_bytes = kHeader;
    h.params_bytes = sizeof(gsph_params);
    h.options_bytes = sizeof(gsph_options);
    int batched = 0;
    rc = gsph_is_batched(ctx, &batched);
    if (rc) return rc;
    h.flags = batched ? 1u : 0u;
    h.nscenes = (uint32_t)ns;
    h.n = (uint64_t)n;
    h.steps = (uint64_t)steps;
    uint64_t hash = kFnvBasis;
    hash = fnv1a(hash, &p, sizeof(p));
    hash = fnv1a(hash, &o, sizeof(o));
    hash = fnv1a(hash, sc.data(), sc.size() * sizeof(SceneRec));
    hash = fnv1a(hash, a.data(), 16 * (size_t)n);
    hash = fnv1a(hash, b.data(), 12 * (size_t)n);
    hash = fnv1a(hash, c.data(), 16 * (size_t)n);
    h.hash = hash;
    // write to a temporary name and rename, so a crash never leaves a
    // half-written snapshot under the final name
    const std::string tmp = std::string(path) + ".tmp";
    {
        File fh;
        fh.f = std::fopen(tmp.c_str(), "wb");
        if (!fh.f) return gsph::fail(ctx, GSPH_ERR_IO, "cannot create " + tmp);
        bool ok = std::fwrite(&h, sizeof(h), 1, fh.f) == 1 && std::fwrite(&p, sizeof(p), 1, fh.f) == 1 &&
                  std::fwrite(&o, sizeof(o), 1, fh.f) == 1 &&
                  std::fwrite(sc.data(), sizeof(SceneRec), sc.size(), fh.f) == sc.size();
        if (ok && n) {
            ok = std::fwrite(a.data(), 16, n, fh.f) == (size_t)n &&
                 std::fwrite(b.data(), 12, n, fh.f) == (size_t)n &&
                 std::fwrite(c.data(), 16, n, fh.f) == (size_t)n;
        }
        ok = ok && std::fflush(fh.f) == 0;
        if (!ok) {
            std::remove(tmp.c_str());
            return gsph::fail(ctx, GSPH_ERR_IO, "short write to " + tmp);
        }
    }
    if (std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return gsph::fail(ctx, GSPH_ERR_IO, std::string("cannot rename snapshot to ") + path);
    }
    return GSPH_OK;
}

int gsph_snapshot_check(const char *path, long long *n, int *nscenes, long long *steps) {
    Header h;
    const int rc = read_snapshot(path, h, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
    if (n) *n = (long long)h.n;
    if (nscenes) *nscenes = (int)h.nscenes;
    if (steps) *steps = (long long)h.steps;
    return GSPH_OK;
}

int gsph_snapshot_load(const char *path, const gsph_options *opt, gsph_ctx **out) {
    if (!out) return gsph::fail(nullptr, GSPH_ERR_INVALID_ARG, "null out");
    Header h;
    gsph_params p;
    gsph_options o;
    std::vector<SceneRec> sc;
    std::vector<float> a, b, c;
    int rc = read_snapshot(path, h, &p, &o, &sc, &a, &b, &c);
    if (rc) return rc;
    if (opt) o = *opt;
    gsph_ctx *ctx = nullptr;
    if (h.flags & 1u) {
        std::vector<float> worlds(3 * sc.size());
        std::vector<int> counts(sc.size());
        for (size_t s = 0; s < sc.size(); ++s) {
            for (int k = 0; k < 3; ++k) worlds[3 * s + k] = sc[s].world[k];
            counts[s] = sc[s].count;
        }
        rc = gsph_batch_create(&p, &o, (int)sc.size(), worlds.data(), counts.data(), a.data(),
                               b.data(), &ctx);
    } else {
        for (int k = 0; k < 3; ++k) p.world_size[k] = sc[0].world[k];
        rc = gsph_create(&p, &o, (int)h.n, a.data(), b.data(), &ctx);
    }
    if (rc) return rc;
    rc = gsph_set_state(ctx, a.data(), b.data(), c.data());
    if (!rc) rc = gsph_set_step_count(ctx, (long long)h.steps);
    if (rc) {
        gsph::set_error(nullptr, gsph_last_error(ctx));
        gsph_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return GSPH_OK;
}

}  // extern "C"
