// serve.cpp - continuous-batching request loop over one batch context.
//
// No counterpart in the reference (it runs one scene per process).  Requests
// (a scene plus a step count) arrive over time.  Each engine iteration:
//   1. admits arrived requests in arrival order while the HBM budget and the
//      scene slots allow it (a request larger than the whole budget is
//      rejected at submission);
//   2. if the set of resident scenes changed, re-composes the batch context:
//      continuing scenes keep their device state (device-to-device moves),
//      new scenes are uploaded;
//   3. advances every resident scene by one step with one launch per kernel;
//   4. retires scenes whose step count is reached (state copied out) and
//      fails scenes whose state left their box or stopped being finite,
//      without disturbing the others.
// Clock: realtime = 1 compares arrival times (s) with the wall clock;
// realtime = 0 counts arrival in engine iterations (deterministic tests).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "engine.h"
#include "gsph.h"
#include "host_util.h"

using namespace gsph;

namespace {

enum ReqState { kQueued = 0, kActive = 1, kDone = 2, kFailed = 3, kRejected = 4 };

struct Request {
    int id;
    float world[3];
    int n;
    std::vector<float> pos, vel;
    int steps, left;
    double arrival, admitted = -1, finished = -1;
    int state = kQueued;
    int slot = -1;         // scene index in the current layout
    long long cells = 0;
    size_t bytes = 0;
    std::vector<float> out_pos, out_vel, out_ev;
    std::string error;
};

}  // namespace

struct gsph_server {
    gsph_params params;
    gsph_options opt;
    double budget;
    int max_scenes;
    gsph_ctx *ctx = nullptr;
    std::vector<Request> reqs;
    std::deque<int> queue;        // queued request ids in arrival order
    std::vector<int> active;      // resident request ids in layout order
    size_t used = 0, peak = 0;
    long long engine_steps = 0, recompositions = 0, particle_steps = 0;
    int max_active = 0;
    double wall = 0;
    std::string err;
};

namespace {

int sfail(gsph_server *s, int code, const std::string &m) {
    if (s) s->err = m;
    set_error(nullptr, m);
    return code;
}

long long cells_of(const gsph_constants &c, const float w[3]) {
    long long t = 1;
    for (int a = 0; a < 3; ++a) t *= (long long)std::ceil(w[a] / c.cell_size);
    return t;
}

// Rebuilds the device layout for the current active set; grows the engine
// when its capacity is exceeded.
int recompose(gsph_server *s, const std::vector<int> &old_slots) {
    std::vector<EngineScene> es;
    long long n = 0, cells = 0;
    for (size_t k = 0; k < s->active.size(); ++k) {
        Request &r = s->reqs[s->active[k]];
        EngineScene e;
        std::memcpy(e.world, r.world, sizeof(e.world));
        e.count = r.n;
        e.old_begin = old_slots[k];
        e.pos4 = r.pos.data();
        e.vel3 = r.vel.empty() ? nullptr : r.vel.data();
        es.push_back(e);
        n += r.n;
        cells += r.cells;
    }
    int cn, cc, cs;
    engine_capacity(s->ctx, &cn, &cc, &cs);
    gsph_ctx *dst = s->ctx;
    if (n > cn || cells > cc || (long long)es.size() > cs) {
        const int nn = (int)std::max<long long>(n, (long long)cn * 3 / 2);
        const int nc = (int)std::min<long long>(std::max<long long>(cells, (long long)cc * 3 / 2), 1LL << 26);
        const int ns = std::max<int>((int)es.size(), cs * 3 / 2);
        int rc = engine_create(&s->params, &s->opt, nn, nc, ns, &dst);
        if (rc) return sfail(s, rc, std::string("engine grow: ") + gsph_last_error(nullptr));
    }
    int rc = engine_recompose(s->ctx, dst, es);
    if (rc) {
        const std::string m = gsph_last_error(dst);
        if (dst != s->ctx) gsph_destroy(dst);
        return sfail(s, rc, "recompose: " + m);
    }
    if (dst != s->ctx) {
        gsph_destroy(s->ctx);
        s->ctx = dst;
    }
    for (size_t k = 0; k < s->active.size(); ++k) {
        Request &r = s->reqs[s->active[k]];
        r.slot = (int)k;
        r.pos.clear();
        r.pos.shrink_to_fit();
        r.vel.clear();
        r.vel.shrink_to_fit();
    }
    ++s->recompositions;
    return GSPH_OK;
}

bool scene_ok(const float *pos, int n, const gsph_constants &c, const float w[3]) {
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            const float x = pos[4 * i + a];
            if (!(x >= 0.f && x < w[a]) || !std::isfinite(pos[4 * i + 3])) return false;
        }
    (void)c;
    return true;
}

}  // namespace

extern "C" {

int gsph_server_create(const gsph_params *p, const gsph_options *opt, double hbm_budget_bytes,
                       int max_scenes, gsph_server **out) {
    if (!p || !out || hbm_budget_bytes <= 0 || max_scenes <= 0)
        return sfail(nullptr, GSPH_ERR_INVALID_ARG, "bad server arguments");
    gsph_server *s = new gsph_server();
    s->params = *p;
    if (opt)
        s->opt = *opt;
    else
        gsph_options_default(&s->opt);
    s->budget = hbm_budget_bytes;
    s->max_scenes = max_scenes;
    int rc = engine_create(p, &s->opt, 1 << 16, 1 << 16, 16, &s->ctx);
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return GSPH_OK;
}

int gsph_server_destroy(gsph_server *s) {
    if (!s) return GSPH_OK;
    gsph_destroy(s->ctx);
    delete s;
    return GSPH_OK;
}

int gsph_server_submit(gsph_server *s, const float world3[3], int n, const float *pos4,
                       const float *vel3, int steps, double arrival, int *id) {
    if (!s || !world3 || n < 0 || (n > 0 && !pos4) || steps < 0 || !id)
        return sfail(s, GSPH_ERR_INVALID_ARG, "bad request");
    Request r;
    r.id = (int)s->reqs.size();
    std::memcpy(r.world, world3, sizeof(r.world));
    r.n = n;
    r.pos.assign(pos4, pos4 + 4 * (size_t)n);
    if (vel3) r.vel.assign(vel3, vel3 + 3 * (size_t)n);
    r.steps = r.left = steps;
    r.arrival = arrival;
    gsph_constants c;
    gsph_params p = s->params;
    std::memcpy(p.world_size, world3, sizeof(p.world_size));
    int rc = gsph_derive_constants(&p, &c);
    if (rc) return rc;
    r.cells = cells_of(c, world3);
    r.bytes = engine_bytes(n, r.cells);
    if ((double)r.bytes > s->budget) {
        r.state = kRejected;
        r.error = "request exceeds the HBM budget";
    } else if (!scene_ok(pos4, n, c, world3)) {
        r.state = kRejected;
        r.error = "particles outside the box or not finite";
    }
    if (r.state == kQueued && steps == 0) {   // nothing to simulate: the input is the result
        r.state = kDone;
        r.admitted = r.finished = arrival;
        r.out_pos = r.pos;
        r.out_vel = r.vel.empty() ? std::vector<float>(3 * (size_t)n, 0.f) : r.vel;
        r.out_ev.assign(4 * (size_t)n, 0.f);
    }
    *id = r.id;
    s->reqs.push_back(std::move(r));
    if (s->reqs.back().state == kQueued) s->queue.push_back(*id);
    return GSPH_OK;
}

int gsph_server_run(gsph_server *s, int realtime) {
    if (!s) return sfail(nullptr, GSPH_ERR_INVALID_ARG, "null server");
    std::stable_sort(s->queue.begin(), s->queue.end(), [&](int a, int b) {
        return s->reqs[a].arrival < s->reqs[b].arrival;
    });
    const auto t0 = std::chrono::steady_clock::now();
    auto clock = [&]() -> double {
        if (realtime)
            return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return (double)s->engine_steps;
    };
    bool changed = false;
    std::vector<int> old_slots;
    while (!s->queue.empty() || !s->active.empty()) {
        const double now = clock();
        // 1. admission, arrival order, budget and slots permitting
        while (!s->queue.empty()) {
            Request &r = s->reqs[s->queue.front()];
            if (r.arrival > now) break;
            if ((int)s->active.size() >= s->max_scenes || (double)(s->used + r.bytes) > s->budget)
                break;
            s->queue.pop_front();
            r.state = kActive;
            r.admitted = now;
            r.slot = -1;
            s->used += r.bytes;
            s->peak = std::max(s->peak, s->used);
            s->active.push_back(r.id);
            changed = true;
        }
        if (s->active.empty()) {
            if (s->queue.empty()) break;
            const double next = s->reqs[s->queue.front()].arrival;
            if (realtime) {
                const double wait = next - clock();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
            } else {
                s->engine_steps = (long long)std::ceil(next);
            }
            continue;
        }
        // 2. re-compose
        if (changed) {
            old_slots.clear();
            for (int id : s->active) {
                const int slot = s->reqs[id].slot;
                old_slots.push_back(slot < 0 ? -1 : engine_scene_begin(s->ctx, slot));
            }
            int rc = recompose(s, old_slots);
            if (rc) return rc;
            changed = false;
        }
        s->max_active = std::max(s->max_active, (int)s->active.size());
        // 3. one step for every resident scene
        int rc = gsph_step(s->ctx, 1, nullptr);
        if (rc == GSPH_OK) {
            rc = gsph_sync(s->ctx);
            if (rc == GSPH_ERR_OUT_OF_DOMAIN) rc = 1;   // find the scene(s) below
        }
        if (rc < 0) return sfail(s, rc, std::string("engine step: ") + gsph_last_error(s->ctx));
        ++s->engine_steps;
        // 4. retire finished scenes, fail broken ones
        std::vector<int> keep;
        for (int id : s->active) {
            Request &r = s->reqs[id];
            --r.left;
            s->particle_steps += r.n;
            bool broken = false;
            if (rc == 1 || r.left == 0) {
                std::vector<float> pos(4 * (size_t)r.n), vel(3 * (size_t)r.n), ev(4 * (size_t)r.n);
                int e = engine_read_scene(s->ctx, r.slot, pos.data(), vel.data(), ev.data());
                if (e) return sfail(s, e, gsph_last_error(s->ctx));
                if (rc == 1) broken = !scene_ok(pos.data(), r.n, gsph_constants(), r.world);
                if (!broken && r.left == 0) {
                    r.out_pos.swap(pos);
                    r.out_vel.swap(vel);
                    r.out_ev.swap(ev);
                }
            }
            if (broken) {
                r.state = kFailed;
                r.error = "state left the box or is not finite";
            } else if (r.left == 0) {
                r.state = kDone;
            }
            if (r.state == kDone || r.state == kFailed) {
                r.finished = clock();
                s->used -= r.bytes;
                changed = true;
            } else {
                keep.push_back(id);
            }
        }
        if (rc == 1) {
            int flags = 0;
            engine_take_errors(s->ctx, &flags);
        }
        s->active.swap(keep);
    }
    s->wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return GSPH_OK;
}

int gsph_server_result(gsph_server *s, int id, float *pos4, float *vel3, float *ev4,
                       int *state, double *latency, double *queued) {
    if (!s || id < 0 || id >= (int)s->reqs.size())
        return sfail(s, GSPH_ERR_INVALID_ARG, "bad request id");
    const Request &r = s->reqs[id];
    if (state) *state = r.state;
    if (latency) *latency = r.finished >= 0 ? r.finished - r.arrival : -1;
    if (queued) *queued = r.admitted >= 0 ? r.admitted - r.arrival : -1;
    if (r.state == kDone) {
        const size_t n = (size_t)r.n;
        if (pos4) std::memcpy(pos4, r.out_pos.data(), 16 * n);
        if (vel3) std::memcpy(vel3, r.out_vel.data(), 12 * n);
        if (ev4) std::memcpy(ev4, r.out_ev.data(), 16 * n);
    }
    return GSPH_OK;
}

int gsph_server_get_stats(gsph_server *s, gsph_server_stats *out) {
    if (!s || !out) return sfail(s, GSPH_ERR_INVALID_ARG, "null argument");
    std::memset(out, 0, sizeof(*out));
    std::vector<double> lat;
    for (const auto &r : s->reqs) {
        if (r.state == kDone) {
            ++out->completed;
            lat.push_back(r.finished - r.arrival);
        } else if (r.state == kFailed) {
            ++out->failed;
        } else if (r.state == kRejected) {
            ++out->rejected;
        }
    }
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double q) {
        if (lat.empty()) return 0.0;
        const double pos = q * (lat.size() - 1);
        const size_t i = (size_t)pos;
        const double f = pos - i;
        return i + 1 < lat.size() ? lat[i] * (1 - f) + lat[i + 1] * f : lat[i];
    };
    out->p50 = pct(0.5);
    out->p99 = pct(0.99);
    out->wall = s->wall;
    out->particle_steps = (double)s->particle_steps;
    out->engine_steps = s->engine_steps;
    out->recompositions = s->recompositions;
    out->peak_bytes = (double)s->peak;
    out->max_active = s->max_active;
    return GSPH_OK;
}

const char *gsph_server_last_error(const gsph_server *s) {
    return s ? s->err.c_str() : gsph_last_error(nullptr);
}

}  // extern "C"
