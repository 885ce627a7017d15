// serve.cpp - continuous-batching request loop over one batch context.
//
// No counterpart in the reference (it runs one scene per process).  Requests
// (a scene plus a step count) arrive over time.  Each engine iteration:
//   1. harvests finished requests whose state has reached host memory;
//   2. admits arrived requests in arrival order while the HBM budget, the
//      scene slots and the 2^26-cell grid limit allow it;
//   3. if the set of resident scenes changed, re-lays out the batch context
//      without a host synchronisation: continuing scenes keep their device
//      state and their sorted order (one move kernel, keys shifted), new
//      scenes are uploaded from pinned memory and appended;
//   4. enqueues one step for every resident scene (one launch per kernel);
//   5. scenes whose step count is reached are copied out asynchronously
//      (pinned memory, an event per request) and leave the layout.
// At most kInFlight steps are enqueued ahead of the GPU.  Device error flags
// are copied to pinned memory after every step and read once that step has
// completed: a scene whose state left its box or stopped being finite fails
// alone (found by reading the resident states); a capacity or sort error ends
// the run.  A finished scene's output is checked on the host (finite, inside
// its box) by a worker thread, so that the check of a large scene does not
// hold back the next steps; a scene that fails it is reported failed when
// the run returns.
// HBM accounting is the exact byte count of the engine's buffers at its
// allocated capacity (engine_bytes), including both engines while a grown one
// replaces the old.
// Clock: realtime = 1 compares arrival times (s) with the wall clock;
// realtime = 0 counts arrival in engine iterations (deterministic tests).
#include <algorithm>
#include <climits>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine.h"
#include "gsph.h"
#include "host_util.h"

using namespace gsph;

namespace {

enum ReqState { kQueued = 0, kActive = 1, kDone = 2, kFailed = 3, kRejected = 4 };
constexpr int kInFlight = 3;                 // engine steps enqueued ahead of the GPU
constexpr int kGraphAfter = 4;               // steps on one layout before it is captured
constexpr int kGraphMinRun = 16;             // ... and steps it is known to have left
constexpr long long kMaxCells = 1LL << 26;   // micro ids must fit 32 bits
constexpr long long kMaxParticles = 0x7fffff00LL;

struct Request {
    int id = -1;
    float world[3] = {0, 0, 0};
    int n = 0;
    int steps = 0, left = 0;
    double arrival = 0, admitted = -1, finished = -1;
    int state = kQueued;
    long long cells = 0;
    bool retiring = false;   // output copy in flight
    float *in_pos = nullptr, *in_vel = nullptr;                        // pinned inputs
    float *out_pos = nullptr, *out_vel = nullptr, *out_ev = nullptr;   // pinned outputs
    hipEvent_t out_done = nullptr;
    int valid = -1;          // host check of the output: -1 pending, 0 failed, 1 passed
    std::string error;
};

// Checks finished outputs on a thread of its own: harvest queues request ids,
// finish() waits for the queue to drain and joins.  The worker writes only
// Request::valid, which the loop reads after finish().
class OutputChecker {
public:
    template <class F>
    explicit OutputChecker(F check) : th_([this, check] {
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_.wait(lk, [this] { return stop_ || !q_.empty(); });
            if (q_.empty()) return;
            const int id = q_.front();
            q_.pop_front();
            lk.unlock();
            check(id);
            lk.lock();
        }
    }) {}
    ~OutputChecker() { finish(); }
    void push(int id) {
        {
            std::lock_guard<std::mutex> g(m_);
            q_.push_back(id);
        }
        cv_.notify_one();
    }
    void finish() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_one();
        if (th_.joinable()) th_.join();
    }

private:
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<int> q_;
    bool stop_ = false;
    std::thread th_;
};

struct StepRec {
    hipEvent_t done = nullptr;
    int *flags = nullptr;   // pinned
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
    std::vector<int> layout;      // request ids of the engine's scenes, in layout order
    std::vector<int> active;      // request ids that step: layout order, then newly admitted
    std::vector<int> retiring;    // output copies in flight
    long long act_n = 0, act_cells = 0;
    size_t peak = 0;
    long long engine_steps = 0, recompositions = 0, particle_steps = 0, grows = 0;
    int max_active = 0;
    double wall = 0;
    std::vector<StepRec> ring;    // kInFlight + 1 step records
    std::deque<int> pending;      // ring slots of enqueued, unchecked steps
    int ring_next = 0;
    long long since_layout = 0;   // steps enqueued on the current layout
    long long layout_run = 0;     // steps until the first resident scene finishes
    bool graphed = false;         // the current layout's step graph exists
    std::string err;
};

namespace {

int sfail(gsph_server *s, int code, const std::string &m) {
    if (s) s->err = m;
    set_error(nullptr, m);
    return code;
}

bool scene_ok(const float *pos, int n, const float w[3]) {
    for (int i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            const float x = pos[4 * i + a];
            if (!(x >= 0.f && x < w[a])) return false;
        }
        if (!std::isfinite(pos[4 * i + 3])) return false;
    }
    return true;
}

void free_request(Request &r) {
    for (float **p : {&r.in_pos, &r.in_vel, &r.out_pos, &r.out_vel, &r.out_ev})
        if (*p) {
            (void)hipHostFree(*p);
            *p = nullptr;
        }
    if (r.out_done) {
        (void)hipEventDestroy(r.out_done);
        r.out_done = nullptr;
    }
}

size_t engine_now(const gsph_server *s) { return s->ctx ? engine_allocated(s->ctx) : 0; }

void note_peak(gsph_server *s, size_t bytes) { s->peak = std::max(s->peak, bytes); }

// Capacity the engine needs for n particles, c cells, k scenes: the current
// one on every axis that suffices, else 1.5x growth (at least what is asked).
void grown_caps(const gsph_server *s, long long n, long long c, long long k, long long caps[3]) {
    int cn, cc, cs;
    engine_capacity(s->ctx, &cn, &cc, &cs);
    const long long cur[3] = {cn, cc, cs}, want[3] = {n, c, k};
    for (int a = 0; a < 3; ++a)
        caps[a] = want[a] <= cur[a] ? cur[a] : std::max(want[a], cur[a] * 3 / 2);
    caps[0] = std::min(caps[0], kMaxParticles);
    caps[1] = std::min(caps[1], kMaxCells);
}

bool fits_engine(const gsph_server *s, long long n, long long c, long long k) {
    int cn, cc, cs;
    engine_capacity(s->ctx, &cn, &cc, &cs);
    return n <= cn && c <= cc && k <= cs;
}

// Device bytes held while an engine of capacity caps replaces the current
// one: both are allocated during the move, unless no scene continues (then
// the old engine is freed first).
double growth_bytes(const gsph_server *s, const long long caps[3], bool continuing) {
    return (double)engine_bytes(s->opt, caps[0], caps[1], caps[2]) +
           (continuing ? (double)engine_now(s) : 0.0);
}

// Whether request r can join the active set now: grid limit, scene slots and
// the HBM budget (with the transient double allocation of a growth).
bool admissible(const gsph_server *s, const Request &r) {
    const long long n = s->act_n + r.n, c = s->act_cells + r.cells;
    const long long k = (long long)s->active.size() + 1;
    if (c > kMaxCells || n > kMaxParticles || k > s->max_scenes) return false;
    if (fits_engine(s, n, c, k)) return true;
    const bool cont = !s->active.empty();
    long long caps[3];
    grown_caps(s, n, c, k, caps);
    if (growth_bytes(s, caps, cont) <= s->budget) return true;
    const long long exact[3] = {n, c, k};
    return growth_bytes(s, exact, cont) <= s->budget;
}

// Re-lays out the engine for the active list: the continuing scenes in their
// layout order, then the newly admitted ones.  Grows the engine (1.5x, or
// exactly what is needed when the budget allows no more) when its capacity
// is exceeded.
int relayout(gsph_server *s) {
    std::vector<EngineScene> es;
    std::vector<int> next;
    std::vector<char> is_active(s->reqs.size(), 0);
    for (int id : s->active) is_active[id] = 1;
    std::vector<char> in_layout(s->reqs.size(), 0);
    for (size_t k = 0; k < s->layout.size(); ++k) {
        const int id = s->layout[k];
        in_layout[id] = 1;
        if (!is_active[id]) continue;
        const Request &r = s->reqs[id];
        EngineScene e;
        std::memcpy(e.world, r.world, sizeof(e.world));
        e.count = r.n;
        int first = 0, cb = 0;
        engine_scene_layout(s->ctx, (int)k, &first, &cb);
        e.old_begin = first;
        e.old_cell_base = cb;
        e.pos4 = nullptr;
        e.vel3 = nullptr;
        es.push_back(e);
        next.push_back(id);
    }
    const bool continuing = !es.empty();
    for (int id : s->active) {
        if (in_layout[id]) continue;
        const Request &r = s->reqs[id];
        EngineScene e;
        std::memcpy(e.world, r.world, sizeof(e.world));
        e.count = r.n;
        e.old_begin = -1;
        e.old_cell_base = 0;
        e.pos4 = r.in_pos;
        e.vel3 = r.in_vel;
        es.push_back(e);
        next.push_back(id);
    }
    const long long k = (long long)es.size();
    gsph_ctx *dst = s->ctx;
    if (!fits_engine(s, s->act_n, s->act_cells, k)) {
        long long caps[3];
        grown_caps(s, s->act_n, s->act_cells, k, caps);
        if (growth_bytes(s, caps, continuing) > s->budget) {
            caps[0] = s->act_n;
            caps[1] = s->act_cells;
            caps[2] = k;
        }
        note_peak(s, (size_t)growth_bytes(s, caps, continuing));
        if (!continuing) {   // nothing to keep: free the old engine first
            gsph_destroy(s->ctx);
            s->ctx = nullptr;
        }
        int rc = engine_create(&s->params, &s->opt, (int)caps[0], (int)caps[1], (int)caps[2], &dst);
        if (rc) return sfail(s, rc, std::string("engine grow: ") + gsph_last_error(nullptr));
        ++s->grows;
    }
    int rc = engine_relayout(continuing ? s->ctx : nullptr, dst, es);
    if (rc) {
        const std::string m = gsph_last_error(dst);
        if (dst != s->ctx) gsph_destroy(dst);
        return sfail(s, rc, "re-layout: " + m);
    }
    if (dst != s->ctx) {
        if (s->ctx) gsph_destroy(s->ctx);
        s->ctx = dst;
    }
    s->layout = next;
    s->active = next;
    s->since_layout = 0;
    s->graphed = false;
    s->layout_run = LLONG_MAX;
    for (int id : next) s->layout_run = std::min<long long>(s->layout_run, s->reqs[id].left);
    note_peak(s, engine_now(s));
    ++s->recompositions;
    return GSPH_OK;
}

// Reads every resident scene's state and fails those that left their box or
// are not finite (after a device out-of-domain flag).  Synchronous; rare.
int fail_broken_scenes(gsph_server *s, double now, bool &changed) {
    std::vector<int> keep;
    for (size_t k = 0; k < s->layout.size(); ++k) {
        const int id = s->layout[k];
        Request &r = s->reqs[id];
        if (r.state != kActive || r.retiring) continue;   // a retiring output is checked on harvest
        std::vector<float> pos(4 * (size_t)r.n);
        int e = engine_read_scene(s->ctx, (int)k, pos.data(), nullptr, nullptr);
        if (e) return sfail(s, e, gsph_last_error(s->ctx));
        if (!scene_ok(pos.data(), r.n, r.world)) {
            r.state = kFailed;
            r.error = "state left the box or is not finite";
            r.finished = now;
            s->act_n -= r.n;
            s->act_cells -= r.cells;
            changed = true;
        }
    }
    for (int id : s->active)
        if (s->reqs[id].state == kActive) keep.push_back(id);
    s->active.swap(keep);
    return GSPH_OK;
}

// Checks the error flags of the oldest enqueued step once it has completed
// (blocking on it when `wait`).  Returns < 0 on a fatal device error.
int check_oldest_step(gsph_server *s, bool wait, double now, bool &changed) {
    if (s->pending.empty()) return GSPH_OK;
    StepRec &sr = s->ring[s->pending.front()];
    if (wait) {
        if (hipEventSynchronize(sr.done) != hipSuccess)
            return sfail(s, GSPH_ERR_HIP, "engine step failed");
    } else if (hipEventQuery(sr.done) != hipSuccess) {
        return GSPH_OK;
    }
    s->pending.pop_front();
    const int flags = *sr.flags;
    if (flags & (kErrCapacity | kErrSortStall))   // the whole batch is unusable
        return sfail(s, (flags & kErrCapacity) ? GSPH_ERR_CAPACITY : GSPH_ERR_SPIN_TIMEOUT,
                     "engine step: device error flags " + std::to_string(flags));
    if (flags & kErrOutOfDomain) {   // find the scene(s)
        const int rc = fail_broken_scenes(s, now, changed);
        if (rc) return rc;
    }
    return GSPH_OK;
}

// Finishes requests whose output copy has completed; their host check runs
// on the checker thread.
void harvest(gsph_server *s, bool wait_one, double now, OutputChecker &checker) {
    std::vector<int> keep;
    bool waited = false;
    for (int id : s->retiring) {
        Request &r = s->reqs[id];
        bool done = hipEventQuery(r.out_done) == hipSuccess;
        if (!done && wait_one && !waited) {
            done = hipEventSynchronize(r.out_done) == hipSuccess;
            waited = true;
        }
        if (!done) {
            keep.push_back(id);
            continue;
        }
        r.retiring = false;
        r.finished = now;
        r.state = kDone;   // until the checker reports otherwise
        checker.push(id);
        // the request's pinned input staging stays until the server is
        // destroyed: freeing pinned memory waits for the whole device, which
        // would drain the steps in flight
    }
    s->retiring.swap(keep);
}

}  // namespace

extern "C" {

int gsph_engine_bytes(const gsph_options *opt, long long particles, long long cells, int scenes,
                      double *bytes) {
    if (!bytes || particles < 0 || cells < 0 || scenes < 0)
        return fail(nullptr, GSPH_ERR_INVALID_ARG, "bad capacity");
    gsph_options o;
    if (opt)
        o = *opt;
    else
        gsph_options_default(&o);
    if (o.sms_mask_bytes < 0 || o.sms_mask_bytes > 65536)
        return fail(nullptr, GSPH_ERR_INVALID_ARG, "sms_mask_bytes must be in [0, 65536]");
    *bytes = (double)engine_bytes(o, particles, cells, scenes);
    return GSPH_OK;
}

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
    const int cap = 1 << 14;
    if ((double)engine_bytes(s->opt, cap, cap, 4) > hbm_budget_bytes) {
        delete s;
        return sfail(nullptr, GSPH_ERR_BUDGET, "the budget does not hold the smallest engine");
    }
    int rc = engine_create(p, &s->opt, cap, cap, 4, &s->ctx);
    if (rc) {
        delete s;
        return rc;
    }
    s->ring.resize(kInFlight + 1);
    for (auto &sr : s->ring) {
        if (hipEventCreateWithFlags(&sr.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void **)&sr.flags, sizeof(int), hipHostMallocDefault) != hipSuccess) {
            gsph_server_destroy(s);
            return sfail(nullptr, GSPH_ERR_HIP, "server event / pinned allocation failed");
        }
        *sr.flags = 0;
    }
    note_peak(s, engine_now(s));
    *out = s;
    return GSPH_OK;
}

int gsph_server_destroy(gsph_server *s) {
    if (!s) return GSPH_OK;
    if (s->ctx) (void)hipStreamSynchronize(engine_stream(s->ctx));
    for (auto &r : s->reqs) free_request(r);
    for (auto &sr : s->ring) {
        if (sr.done) (void)hipEventDestroy(sr.done);
        if (sr.flags) (void)hipHostFree(sr.flags);
    }
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
    r.steps = r.left = steps;
    r.arrival = arrival;
    gsph_constants c;
    gsph_params p = s->params;
    std::memcpy(p.world_size, world3, sizeof(p.world_size));
    int rc = gsph_derive_constants(&p, &c);
    if (rc) return rc;
    r.cells = engine_cells(s->ctx, world3);
    if (r.cells > kMaxCells) {
        r.state = kRejected;
        r.error = "more than 2^26 grid cells";
    } else if ((double)engine_bytes(s->opt, n, r.cells, 1) > s->budget) {
        r.state = kRejected;
        r.error = "request exceeds the HBM budget";
    } else if (!scene_ok(pos4, n, world3)) {
        r.state = kRejected;
        r.error = "particles outside the box or not finite";
    }
    if (r.state == kQueued) {
        const size_t nn = (size_t)std::max(n, 1);
        if (hipHostMalloc((void **)&r.in_pos, 16 * nn, hipHostMallocDefault) != hipSuccess ||
            (vel3 && hipHostMalloc((void **)&r.in_vel, 12 * nn, hipHostMallocDefault) != hipSuccess) ||
            hipHostMalloc((void **)&r.out_pos, 16 * nn, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&r.out_vel, 12 * nn, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&r.out_ev, 16 * nn, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&r.out_done, hipEventDisableTiming) != hipSuccess) {
            free_request(r);
            return sfail(s, GSPH_ERR_HIP, "pinned host allocation for a request failed");
        }
        std::memcpy(r.in_pos, pos4, 16 * (size_t)n);
        if (vel3) std::memcpy(r.in_vel, vel3, 12 * (size_t)n);
        if (steps == 0) {   // nothing to simulate: the input is the result
            r.state = kDone;
            r.admitted = r.finished = arrival;
            std::memcpy(r.out_pos, pos4, 16 * (size_t)n);
            if (vel3)
                std::memcpy(r.out_vel, vel3, 12 * (size_t)n);
            else
                std::memset(r.out_vel, 0, 12 * (size_t)n);
            std::memset(r.out_ev, 0, 16 * (size_t)n);
        }
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
    OutputChecker checker([s](int id) {
        Request &r = s->reqs[id];
        r.valid = scene_ok(r.out_pos, r.n, r.world) ? 1 : 0;
    });
    // applies the checker's verdicts (after it has joined), also on an early return
    struct Verdicts {
        gsph_server *s;
        OutputChecker &c;
        ~Verdicts() {
            c.finish();
            for (auto &r : s->reqs)
                if (r.state == kDone && r.valid == 0) {
                    r.state = kFailed;
                    r.error = "state left the box or is not finite";
                }
        }
    } verdicts{s, checker};
    auto clock = [&]() -> double {
        if (realtime)
            return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return (double)s->engine_steps;
    };
    bool changed = false;
    while (!s->queue.empty() || !s->active.empty() || !s->retiring.empty() ||
           !s->pending.empty()) {
        double now = clock();
        // 1. completed steps' error flags, finished outputs
        int rc = check_oldest_step(s, false, now, changed);
        if (rc) return rc;
        harvest(s, false, now, checker);
        // 2. admission, arrival order, budget / slots / grid permitting
        while (!s->queue.empty()) {
            Request &r = s->reqs[s->queue.front()];
            if (r.arrival > now || !admissible(s, r)) break;
            s->queue.pop_front();
            r.state = kActive;
            r.admitted = now;
            s->active.push_back(r.id);
            s->act_n += r.n;
            s->act_cells += r.cells;
            changed = true;
        }
        if (s->active.empty()) {
            if (!s->pending.empty()) {
                rc = check_oldest_step(s, true, clock(), changed);
                if (rc) return rc;
                continue;
            }
            if (!s->retiring.empty()) {
                harvest(s, true, clock(), checker);
                continue;
            }
            if (s->queue.empty()) break;
            const double next = s->reqs[s->queue.front()].arrival;
            if (next <= now) {
                // alone, the head request always fits: its engine was checked
                // against the budget at submission, and an engine with no
                // continuing scene is freed before its replacement is made
                Request &r = s->reqs[s->queue.front()];
                s->queue.pop_front();
                r.state = kActive;
                r.admitted = now;
                s->active.push_back(r.id);
                s->act_n += r.n;
                s->act_cells += r.cells;
                changed = true;
            } else if (realtime) {
                const double wait = next - clock();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
                continue;
            } else {
                s->engine_steps = (long long)std::ceil(next);
                continue;
            }
        }
        // 3. re-layout
        if (changed) {
            rc = relayout(s);
            if (rc) return rc;
            changed = false;
        }
        s->max_active = std::max(s->max_active, (int)s->active.size());
        // 4. one step for every resident scene, at most kInFlight ahead
        while ((int)s->pending.size() >= kInFlight) {
            rc = check_oldest_step(s, true, clock(), changed);
            if (rc) return rc;
        }
        if (changed) continue;   // a scene failed: lay out again before stepping
        // a layout that has lasted a few steps, and lasts at least kGraphMinRun
        // more unless an arrival changes it, is replayed from a graph (one
        // launch instead of one per kernel); a re-layout drops it.  Capturing
        // and instantiating a graph costs about as much as a few eager steps,
        // so a layout about to change (a resident scene finishing) stays eager.
        s->graphed = s->graphed || (s->since_layout >= kGraphAfter &&
                                    s->layout_run - s->since_layout >= kGraphMinRun);
        rc = engine_step(s->ctx, s->graphed);
        ++s->since_layout;
        if (rc) return sfail(s, rc, std::string("engine step: ") + gsph_last_error(s->ctx));
        StepRec &sr = s->ring[s->ring_next];
        rc = engine_flags_async(s->ctx, sr.flags);
        if (!rc && hipEventRecord(sr.done, engine_stream(s->ctx)) != hipSuccess) rc = GSPH_ERR_HIP;
        if (rc) return sfail(s, rc, "engine step bookkeeping failed");
        s->pending.push_back(s->ring_next);
        s->ring_next = (s->ring_next + 1) % (int)s->ring.size();
        ++s->engine_steps;
        // 5. retire finished scenes
        std::vector<int> keep;
        for (size_t k = 0; k < s->layout.size(); ++k) {
            const int id = s->layout[k];
            Request &r = s->reqs[id];
            if (r.state != kActive) continue;
            --r.left;
            s->particle_steps += r.n;
            if (r.left > 0) {
                keep.push_back(id);
                continue;
            }
            rc = engine_retire_async(s->ctx, (int)k, r.out_pos, r.out_vel, r.out_ev, r.out_done);
            if (rc) return sfail(s, rc, gsph_last_error(s->ctx));
            r.retiring = true;
            s->retiring.push_back(id);
            s->act_n -= r.n;
            s->act_cells -= r.cells;
            changed = true;
        }
        s->active.swap(keep);
    }
    if (s->ctx) (void)hipStreamSynchronize(engine_stream(s->ctx));
    checker.finish();   // the last checks belong to the run
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
        if (pos4) std::memcpy(pos4, r.out_pos, 16 * n);
        if (vel3) std::memcpy(vel3, r.out_vel, 12 * n);
        if (ev4) std::memcpy(ev4, r.out_ev, 16 * n);
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
