// rb_server.hip -- continuous-batching request server (host code).
//
// One scheduler thread per server owns every HIP call of the serving path:
// it admits queued requests while a slot (HIP stream + event + pinned
// staging) is free and the request's estimated device memory fits the
// budget, launches the request's passes on the slot's stream, polls the
// slot events for completion, copies results out, and records latency
// (submit -> result ready).  Admission order: earliest virtual deadline
// first (deadline = submit time + 4 x the request's estimated GPU time, so
// short requests overtake long ones but nothing waits forever), or plain
// FIFO; requests that do not fit the budget wait without blocking the ones
// behind them (bounded bypass); a request larger than the whole budget is
// rejected at submit.  Results do not depend on what else runs: every
// random number is keyed by (request seed, pixel, sample, bounce).
//
// Two ways to run the admitted requests:
//  * streams (default): each request's whole pass is enqueued on its slot's
//    stream at admission, and the streams' kernels overlap on the GPU;
//  * lockstep: every request is a render job (make_render_job) and all of
//    them advance together on one stream: per round, each job enqueues its
//    integrator kernels up to its next trace, then one batched launch
//    (launch_trace_sets) traces the ray sets of every in-flight request, so
//    a round pays one persistent launch's ramp-up and tail instead of one
//    per request.  The scheduler keeps at most kRoundsAhead rounds queued on
//    the device, so a request admitted meanwhile joins within a round or
//    two.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "rb_device.h"

namespace rb {
namespace {

using Clock = std::chrono::steady_clock;

struct Req {
    uint64_t id = 0;
    rb_request r{};
    rb_bvh_t bvh = nullptr;
    uint64_t bytes = 0;
    double est_rays = 0;
    Clock::time_point deadline;
    int state = RB_REQ_QUEUED;
    rb_status result = RB_OK;
    bool deferred = false;
    Clock::time_point t_submit, t_admit, t_done;
    float* d_out = nullptr;
    size_t out_bytes = 0;
    uint64_t rays = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    float* pinned = nullptr;
    size_t pinned_cap = 0;
    std::vector<float*> retired;  // outgrown staging, freed with the server (freeing pinned memory syncs the device)
    unsigned long long* d_counter = nullptr;
    unsigned long long* h_counter = nullptr;  // pinned
    char* arena = nullptr;  // device memory for the output and scratch of this slot's requests
    size_t arena_cap = 0;
    Req* req = nullptr;
    bool tail = false;                // the request's result copies and event are enqueued
    std::unique_ptr<RenderJob> job;   // lockstep: the request's remaining steps
};

// Lockstep rounds queued on the device at most.
constexpr uint32_t kRoundsAhead = 2;

// Arena headroom over a request's estimate: 256-byte alignment of each
// scratch allocation.
constexpr size_t kArenaSlack = (size_t)64 << 10;
constexpr size_t kArenaGrain = (size_t)2 << 20;

struct SceneEntry {
    rb_bvh_t bvh;
    uint32_t refs;
};

}  // namespace

uint64_t request_bytes(const rb_request& r) {
    if (r.kind == RB_REQ_TRACE) return 64;
    const uint64_t out = (uint64_t)r.w * r.h * 12;
    return out + render_scratch_bytes(r.kind, r.w, r.h, &r.params) + 4096;
}

// Rays a request is expected to trace (scheduling estimate only), from the
// per-sample ray counts of the passes at their default depths: AO 1 + n_ao,
// path tracing ~7.5 at depth 8, BDPT ~(2 + connections) per camera vertex
// plus the light pass.
double estimate_rays(const rb_request& r) {
    const double n = (double)r.w * r.h * (r.params.spp ? r.params.spp : 1);
    const double depth = (double)r.params.max_depth + 1.0;
    switch (r.kind) {
        case RB_REQ_AO: return n * (1.0 + r.params.ao_samples);
        case RB_REQ_RAYCAST: return n;
        case RB_REQ_PT: return n * 7.5 * depth / 9.0;
        case RB_REQ_BDPT: {
            const double D = r.params.light_max_depth ? r.params.light_max_depth : 1;
            const double nl = r.params.light_paths ? r.params.light_paths : r.params.lvc_capacity / D;
            return n * (2.0 + r.params.connections) * depth / 3.0 + nl * D;
        }
        default: return (double)r.n;
    }
}

}  // namespace rb

struct rb_server_s {
    rb_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::multimap<rb::Clock::time_point, rb::Req*> queue;  // by admission key (deadline or submit time)
    std::unordered_map<uint64_t, std::unique_ptr<rb::Req>> reqs;
    std::map<uint32_t, rb::SceneEntry> scenes;
    uint32_t next_scene = 0;
    uint64_t next_id = 1;
    std::vector<rb::Slot> slots;
    uint64_t used = 0, peak = 0;
    uint32_t inject_failures = 0;
    std::atomic<uint64_t> reserved{0};          // arena bytes held (changed by the scheduler thread only)
    uint64_t peak_reserved = 0, arena_misses = 0;  // under mu
    bool stop = false;
    std::thread th;
    // stats
    uint64_t submitted = 0, completed = 0, failed = 0, rejected = 0, deferrals = 0, rays = 0;
    std::vector<double> lat_ms;
    bool have_first = false;
    rb::Clock::time_point first_submit, last_done;
    // lockstep mode
    hipStream_t lstream = nullptr;
    hipEvent_t round_ev[rb::kRoundsAhead] = {};
    uint32_t round_head = 0, rounds_out = 0;
    uint64_t rounds = 0;
    std::vector<rb::BatchSet> round_sets;

    void loop();
    rb_status prepare(rb::Slot& sl, rb::Req* q, rb::Arena* arena);
    void launch(rb::Slot& sl, rb::Req* q);
    void launch_lockstep(rb::Slot& sl, rb::Req* q);
    rb_status tail(rb::Slot& sl);
    bool round();
    void finish(rb::Slot& sl, rb_status result);
    rb_status ensure_arena(rb::Slot& sl, size_t need);
    void drop_arena(rb::Slot& sl);
};

using namespace rb;

void rb_server_s::finish(Slot& sl, rb_status result) {
    Req* q = sl.req;
    if (result == RB_OK) {
        if (q->r.out_host && q->d_out) std::memcpy(q->r.out_host, sl.pinned, q->out_bytes);
        q->rays = *sl.h_counter;
    }
    q->d_out = nullptr;  // lives in the slot's arena
    std::lock_guard<std::mutex> g(mu);
    q->t_done = Clock::now();
    q->result = result;
    q->state = result == RB_OK ? RB_REQ_DONE : RB_REQ_FAILED;
    used -= q->bytes;
    auto it = scenes.find(q->r.scene);
    if (it != scenes.end()) it->second.refs--;
    if (result == RB_OK) {
        completed++;
        rays += q->rays;
        lat_ms.push_back(std::chrono::duration<double, std::milli>(q->t_done - q->t_submit).count());
        last_done = q->t_done;
    } else {
        failed++;
    }
    sl.req = nullptr;
    sl.tail = false;
    sl.job.reset();
    cv_done.notify_all();
}

void rb_server_s::drop_arena(Slot& sl) {
    if (!sl.arena) return;
    (void)hipFreeAsync(sl.arena, sl.stream);
    reserved -= sl.arena_cap;
    sl.arena = nullptr;
    sl.arena_cap = 0;
}

// Each slot keeps one device arena for its requests' output and scratch.
// Only the slot's stream uses it, in order, so a request reuses the memory of
// the previous one on that slot without an allocator call and without
// waiting for other streams (a shared stream-ordered pool may order a stream
// behind the one that freed the block it hands out).  The arena grows to the
// largest request the slot has run; when the arenas together would pass the
// budget, idle slots give theirs back first.
rb_status rb_server_s::ensure_arena(Slot& sl, size_t need) {
    if (sl.arena_cap >= need) return RB_OK;
    const size_t cap = (need + kArenaGrain - 1) / kArenaGrain * kArenaGrain;
    if (reserved - sl.arena_cap + cap > cfg.hbm_budget)
        for (Slot& o : slots)
            if (&o != &sl && !o.req) drop_arena(o);
    drop_arena(sl);
    void* p = nullptr;
    hipError_t e = hipMallocAsync(&p, cap, sl.stream);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(RB_ERR_OOM, "server arena (%zu bytes): %s", cap, hipGetErrorString(e));
    }
    sl.arena = (char*)p;
    sl.arena_cap = cap;
    reserved += cap;
    std::lock_guard<std::mutex> g(mu);
    peak_reserved = std::max<uint64_t>(peak_reserved, reserved.load());
    return RB_OK;
}

// Counter reset, output and scratch arena, pinned staging of a request
// admitted to slot sl (arena: where its scratch comes from).
rb_status rb_server_s::prepare(Slot& sl, Req* q, Arena* arena) {
    sl.req = q;
    sl.tail = false;
    rb_status st = RB_OK;
    hipError_t e = hipMemsetAsync(sl.d_counter, 0, sizeof(unsigned long long), sl.stream);
    if (e != hipSuccess) st = hip_fail(e, "server counter reset");
    if (st == RB_OK && q->r.kind != RB_REQ_TRACE) {
        q->out_bytes = (size_t)q->r.w * q->r.h * 12;
        bool inject = false;
        {
            std::lock_guard<std::mutex> g(mu);
            if (inject_failures) {
                inject_failures--;
                inject = true;
            }
        }
        if (inject) {
            st = fail(RB_ERR_OOM, "request %llu: device allocation failed (injected)", (unsigned long long)q->id);
        } else if ((st = ensure_arena(sl, q->bytes + kArenaSlack)) == RB_OK) {
            *arena = Arena{sl.arena, sl.arena_cap, 0};
            q->d_out = (float*)arena->take(q->out_bytes);  // first in the arena: always fits
        }
        if (st == RB_OK && q->r.out_host && sl.pinned_cap < q->out_bytes) {
            // grow geometrically from 8 MB; the old buffer is kept until the
            // server goes, since freeing pinned memory synchronises the device
            size_t cap = sl.pinned_cap ? 2 * sl.pinned_cap : ((size_t)8 << 20);
            while (cap < q->out_bytes) cap *= 2;
            float* np_ = nullptr;
            if (hipHostMalloc((void**)&np_, cap, hipHostMallocDefault) != hipSuccess) {
                st = fail(RB_ERR_OOM, "pinned staging allocation failed");
            } else {
                if (sl.pinned) sl.retired.push_back(sl.pinned);
                sl.pinned = np_;
                sl.pinned_cap = cap;
            }
        }
    }
    return st;
}

// The request's result and ray count to the host, then its event.
rb_status rb_server_s::tail(Slot& sl) {
    Req* q = sl.req;
    hipError_t e;
    if (q->r.out_host && q->d_out &&
        (e = hipMemcpyAsync(sl.pinned, q->d_out, q->out_bytes, hipMemcpyDeviceToHost, sl.stream)) != hipSuccess)
        return hip_fail(e, "server result copy");
    if ((e = hipMemcpyAsync(sl.h_counter, sl.d_counter, sizeof(unsigned long long), hipMemcpyDeviceToHost, sl.stream)) !=
        hipSuccess)
        return hip_fail(e, "server ray count copy");
    if ((e = hipEventRecord(sl.ev, sl.stream)) != hipSuccess) return hip_fail(e, "server event");
    sl.tail = true;
    return RB_OK;
}

void rb_server_s::launch(Slot& sl, Req* q) {
    Arena arena;
    rb_status st = prepare(sl, q, &arena);
    hipError_t e;
    if (st == RB_OK) {
        rb_render_params p = q->r.params;
        p.ray_counter = sl.d_counter;
        p.path_counter = p.light_path_counter = nullptr;
        t_arena = q->r.kind == RB_REQ_TRACE ? nullptr : &arena;
        switch (q->r.kind) {
            case RB_REQ_AO: st = render_ao(q->bvh, &q->r.camera, q->r.x0, q->r.y0, q->r.w, q->r.h, &p, q->d_out, sl.stream); break;
            case RB_REQ_PT: st = render_pt(q->bvh, &q->r.camera, q->r.x0, q->r.y0, q->r.w, q->r.h, &p, q->d_out, sl.stream); break;
            case RB_REQ_BDPT: st = render_bdpt(q->bvh, &q->r.camera, q->r.x0, q->r.y0, q->r.w, q->r.h, &p, q->d_out, sl.stream); break;
            case RB_REQ_RAYCAST: st = render_raycast(q->bvh, &q->r.camera, q->r.x0, q->r.y0, q->r.w, q->r.h, &p, q->d_out, sl.stream); break;
            case RB_REQ_TRACE:
                st = launch_trace(q->bvh, q->r.rays, q->r.n, q->r.flags & RB_ANY, nullptr, q->r.hits, nullptr, sl.stream,
                                  dev().cfg);
                if (st == RB_OK && (e = hipMemcpyAsync(sl.d_counter, &q->r.n, 4, hipMemcpyHostToDevice, sl.stream)) != hipSuccess)
                    st = hip_fail(e, "server trace count");
                break;
            default: st = fail(RB_ERR_INVALID, "unknown request kind %u", q->r.kind);
        }
        t_arena = nullptr;
        if (arena.misses) {
            std::lock_guard<std::mutex> g(mu);
            arena_misses += arena.misses;
        }
    }
    if (st == RB_OK) st = tail(sl);
    if (st != RB_OK) {
        (void)hipStreamSynchronize(sl.stream);
        finish(sl, st);
    }
}

// Lockstep admission: the request becomes a render job whose steps the
// rounds run (a trace request is traced in the next round).
void rb_server_s::launch_lockstep(Slot& sl, Req* q) {
    Arena arena;
    rb_status st = prepare(sl, q, &arena);
    if (st == RB_OK && q->r.kind != RB_REQ_TRACE) {
        rb_render_params p = q->r.params;
        p.ray_counter = sl.d_counter;
        p.path_counter = p.light_path_counter = nullptr;
        t_arena = &arena;
        st = make_render_job(q->r.kind, q->bvh, &q->r.camera, q->r.x0, q->r.y0, q->r.w, q->r.h, &p, q->d_out,
                             sl.stream, &sl.job);
        t_arena = nullptr;
        if (arena.misses) {
            std::lock_guard<std::mutex> g(mu);
            arena_misses += arena.misses;
        }
    }
    if (st != RB_OK) {
        sl.job.reset();
        (void)hipStreamSynchronize(sl.stream);
        finish(sl, st);
    }
}

// One lockstep round: every running job's next step, then one batched trace
// of all their ray sets (and of the trace requests).  Returns whether any
// work was enqueued.
bool rb_server_s::round() {
    while (rounds_out && hipEventQuery(round_ev[(round_head + kRoundsAhead - rounds_out) % kRoundsAhead]) == hipSuccess)
        rounds_out--;
    if (rounds_out >= kRoundsAhead) return false;
    round_sets.clear();
    std::vector<Slot*> traced;
    bool work = false;
    for (Slot& sl : slots) {
        if (!sl.req || sl.tail) continue;
        Req* q = sl.req;
        work = true;
        if (q->r.kind == RB_REQ_TRACE) {
            if (q->r.n)
                round_sets.push_back(BatchSet{q->bvh, reinterpret_cast<const float4*>(q->r.rays),
                                              reinterpret_cast<int2*>(q->r.hits), q->r.n, nullptr,
                                              (q->r.flags & RB_ANY) != 0, nullptr});
            traced.push_back(&sl);
            continue;
        }
        TraceSet ts[2];
        uint32_t ns = 0;
        bool done = false;
        rb_status st = sl.job->step(lstream, ts, &ns, &done);
        if (st == RB_OK && done) {
            sl.job.reset();  // its scratch lives in the slot's arena
            st = tail(sl);
        } else if (st == RB_OK) {
            for (uint32_t k = 0; k < ns; k++)
                round_sets.push_back(BatchSet{q->bvh, ts[k].rays, ts[k].hits, ts[k].n, ts[k].n_dev,
                                              (ts[k].flags & RB_ANY) != 0,
                                              (ts[k].flags & RB_MASKED) ? ts[k].mask : nullptr,
                                              (ts[k].flags & kOcclusion) != 0});
        }
        if (st != RB_OK) {
            sl.job.reset();
            (void)hipStreamSynchronize(lstream);
            finish(sl, st);
        }
    }
    if (!round_sets.empty()) {
        const rb_status st = launch_trace_sets(round_sets.data(), (uint32_t)round_sets.size(), lstream, dev().cfg);
        if (st != RB_OK) {  // every job of the round stops with the error
            (void)hipStreamSynchronize(lstream);
            for (Slot& sl : slots)
                if (sl.req && !sl.tail) finish(sl, st);
            return true;
        }
    }
    for (Slot* sl : traced) {
        hipError_t e = hipMemcpyAsync(sl->d_counter, &sl->req->r.n, 4, hipMemcpyHostToDevice, lstream);
        rb_status st = e == hipSuccess ? tail(*sl) : hip_fail(e, "server trace count");
        if (st != RB_OK) {
            (void)hipStreamSynchronize(lstream);
            finish(*sl, st);
        }
    }
    if (work && hipEventRecord(round_ev[round_head], lstream) == hipSuccess) {
        round_head = (round_head + 1) % kRoundsAhead;
        rounds_out++;
        rounds++;
    }
    return work;
}

void rb_server_s::loop() {
    (void)hipSetDevice(dev().device);  // the device rb_server_create validated
    std::vector<std::pair<Slot*, Req*>> to_launch;
    while (true) {
        // completions
        bool busy = false;
        for (Slot& sl : slots) {
            if (!sl.req) continue;
            if (!sl.tail) {  // lockstep: steps to go
                busy = true;
                continue;
            }
            hipError_t e = hipEventQuery(sl.ev);
            if (e == hipSuccess) finish(sl, RB_OK);
            else if (e != hipErrorNotReady) {
                (void)hipGetLastError();
                finish(sl, hip_fail(e, "request execution"));
            } else busy = true;
        }
        // admission
        to_launch.clear();
        {
            std::unique_lock<std::mutex> lk(mu);
            if (stop && queue.empty() && !busy) break;
            size_t looked = 0;
            for (auto it = queue.begin(); it != queue.end() && looked < 16;) {  // earliest key first
                Slot* free_slot = nullptr;
                for (Slot& sl : slots)
                    if (!sl.req) {
                        bool taken = false;
                        for (auto& p : to_launch) taken |= p.first == &sl;
                        if (!taken) {
                            free_slot = &sl;
                            break;
                        }
                    }
                if (!free_slot) break;
                Req* q = it->second;
                if (used + q->bytes <= cfg.hbm_budget) {
                    used += q->bytes;
                    peak = std::max(peak, used);
                    q->state = RB_REQ_RUNNING;
                    q->t_admit = Clock::now();
                    to_launch.emplace_back(free_slot, q);
                    it = queue.erase(it);
                } else {
                    if (!q->deferred) {
                        q->deferred = true;
                        deferrals++;
                    }
                    ++it;
                    ++looked;
                }
            }
            if (to_launch.empty() && !busy) {
                if (stop && queue.empty()) break;
                cv_work.wait_for(lk, std::chrono::milliseconds(queue.empty() ? 50 : 1));
                continue;
            }
        }
        const bool lockstep = cfg.mode == RB_SERVE_LOCKSTEP;
        for (auto& p : to_launch) {
            if (lockstep) launch_lockstep(*p.first, p.second);
            else launch(*p.first, p.second);
        }
        const bool stepped = lockstep && round();
        if (to_launch.empty() && !stepped) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

extern "C" {

rb_status rb_server_request_bytes(const rb_request* req, uint64_t* bytes) {
    if (!req || !bytes) return fail(RB_ERR_INVALID, "NULL argument");
    *bytes = request_bytes(*req);
    return RB_OK;
}

rb_status rb_server_create(const rb_server_config* cfg, rb_server_t* out) {
    if (!out) return fail(RB_ERR_INVALID, "NULL out");
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    std::unique_ptr<rb_server_s> S(new rb_server_s());
    if (cfg) S->cfg = *cfg;
    if (S->cfg.max_inflight == 0) S->cfg.max_inflight = 8;
    if (S->cfg.max_inflight > 32) return fail(RB_ERR_INVALID, "max_inflight must be <= 32");
    if (S->cfg.max_queue == 0) S->cfg.max_queue = 65536;
    if (S->cfg.policy > RB_SCHED_FIFO) return fail(RB_ERR_INVALID, "unknown scheduling policy %u", S->cfg.policy);
    if (S->cfg.mode > RB_SERVE_LOCKSTEP) return fail(RB_ERR_INVALID, "unknown serving mode %u", S->cfg.mode);
    if (S->cfg.hbm_budget == 0) {
        size_t f = 0, t = 0;
        if (hipMemGetInfo(&f, &t) != hipSuccess) return fail(RB_ERR_DEVICE, "hipMemGetInfo failed");
        S->cfg.hbm_budget = (uint64_t)(f * 0.6);
    }
    S->slots.resize(S->cfg.max_inflight);
    if (S->cfg.mode == RB_SERVE_LOCKSTEP) {  // one stream for every request
        if (hipStreamCreateWithFlags(&S->lstream, hipStreamNonBlocking) != hipSuccess)
            return fail(RB_ERR_DEVICE, "server stream creation failed");
        for (hipEvent_t& ev : S->round_ev)
            if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
                return fail(RB_ERR_DEVICE, "server event creation failed");
    }
    for (Slot& sl : S->slots) {
        if (S->lstream) sl.stream = S->lstream;
        if ((!S->lstream && hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess) ||
            hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess ||
            hipMalloc((void**)&sl.d_counter, sizeof(unsigned long long)) != hipSuccess ||
            hipHostMalloc((void**)&sl.h_counter, sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
            return fail(RB_ERR_DEVICE, "server slot creation failed");
        *sl.h_counter = 0;
    }
    rb_server_s* raw = S.get();
    S->th = std::thread([raw] { raw->loop(); });
    *out = S.release();
    return RB_OK;
}

rb_status rb_server_destroy(rb_server_t s) {
    if (!s) return RB_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        for (auto& kv : s->queue) {
            Req* q = kv.second;
            q->state = RB_REQ_FAILED;
            q->result = RB_ERR_BUSY;
            auto it = s->scenes.find(q->r.scene);
            if (it != s->scenes.end()) it->second.refs--;
        }
        s->queue.clear();
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    if (s->th.joinable()) s->th.join();
    for (Slot& sl : s->slots) {
        sl.job.reset();
        if (sl.arena) (void)hipFreeAsync(sl.arena, sl.stream);
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.stream && sl.stream != s->lstream) (void)hipStreamDestroy(sl.stream);
        if (sl.d_counter) (void)hipFree(sl.d_counter);
        if (sl.h_counter) (void)hipHostFree(sl.h_counter);
        if (sl.pinned) (void)hipHostFree(sl.pinned);
        for (float* r : sl.retired) (void)hipHostFree(r);
    }
    for (hipEvent_t ev : s->round_ev)
        if (ev) (void)hipEventDestroy(ev);
    if (s->lstream) (void)hipStreamDestroy(s->lstream);
    delete s;
    return RB_OK;
}

rb_status rb_server_add_scene(rb_server_t s, rb_bvh_t bvh, uint32_t* id) {
    if (!s || !bvh || !id) return fail(RB_ERR_INVALID, "NULL argument");
    std::lock_guard<std::mutex> g(s->mu);
    *id = s->next_scene++;
    s->scenes[*id] = SceneEntry{bvh, 0};
    return RB_OK;
}

rb_status rb_server_remove_scene(rb_server_t s, uint32_t id) {
    if (!s) return fail(RB_ERR_INVALID, "NULL server");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->scenes.find(id);
    if (it == s->scenes.end()) return fail(RB_ERR_INVALID, "unknown scene %u", id);
    if (it->second.refs) return fail(RB_ERR_BUSY, "scene %u has %u queued or running requests", id, it->second.refs);
    s->scenes.erase(it);
    return RB_OK;
}

rb_status rb_server_submit(rb_server_t s, const rb_request* req, uint64_t* id) {
    if (!s || !req) return fail(RB_ERR_INVALID, "NULL argument");
    if (req->kind > RB_REQ_RAYCAST) return fail(RB_ERR_INVALID, "unknown request kind %u", req->kind);
    std::unique_ptr<Req> q(new Req());
    q->r = *req;
    q->bytes = request_bytes(*req);
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->scenes.find(req->scene);
    if (it == s->scenes.end()) return fail(RB_ERR_INVALID, "unknown scene %u", req->scene);
    if (req->kind != RB_REQ_TRACE) {
        if (req->w == 0 || req->h == 0 || (uint64_t)req->x0 + req->w > req->camera.width ||
            (uint64_t)req->y0 + req->h > req->camera.height)
            return fail(RB_ERR_INVALID, "bad tile");
        if (req->params.spp == 0) return fail(RB_ERR_INVALID, "spp must be >= 1");
        if (!it->second.bvh->tri_v) return fail(RB_ERR_INVALID, "scene %u has no shading data", req->scene);
    } else if (req->n && (!req->rays || !req->hits)) {
        return fail(RB_ERR_INVALID, "trace request without rays/hits");
    }
    if (q->bytes > s->cfg.hbm_budget) {
        s->rejected++;
        return fail(RB_ERR_OOM, "request needs %llu bytes, the server budget is %llu",
                    (unsigned long long)q->bytes, (unsigned long long)s->cfg.hbm_budget);
    }
    if (s->queue.size() >= s->cfg.max_queue) {
        s->rejected++;
        return fail(RB_ERR_BUSY, "server queue full (%u)", s->cfg.max_queue);
    }
    q->bvh = it->second.bvh;
    it->second.refs++;
    q->id = s->next_id++;
    q->t_submit = Clock::now();
    q->est_rays = estimate_rays(*req);
    // virtual deadline: 4x the estimated time at 1 Grays/s, plus 1 ms
    q->deadline = q->t_submit + std::chrono::microseconds((int64_t)(1000.0 + 4.0 * q->est_rays / 1.0e3));
    if (!s->have_first) {
        s->have_first = true;
        s->first_submit = q->t_submit;
    }
    s->submitted++;
    if (id) *id = q->id;
    s->queue.emplace(s->cfg.policy == RB_SCHED_FIFO ? q->t_submit : q->deadline, q.get());
    s->reqs[q->id] = std::move(q);
    s->cv_work.notify_one();
    return RB_OK;
}

static void fill_status(const Req& q, rb_request_status* st) {
    st->state = q.state;
    st->result = q.result;
    st->rays = q.rays;
    st->queue_ms = q.state >= RB_REQ_RUNNING && q.t_admit.time_since_epoch().count()
                       ? std::chrono::duration<double, std::milli>(q.t_admit - q.t_submit).count()
                       : 0.0;
    st->latency_ms = q.state >= RB_REQ_DONE ? std::chrono::duration<double, std::milli>(q.t_done - q.t_submit).count()
                                            : 0.0;
}

rb_status rb_server_poll(rb_server_t s, uint64_t id, rb_request_status* st) {
    if (!s || !st) return fail(RB_ERR_INVALID, "NULL argument");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(RB_ERR_INVALID, "unknown request %llu", (unsigned long long)id);
    fill_status(*it->second, st);
    return st->state >= RB_REQ_DONE ? RB_OK : RB_ERR_NOT_READY;
}

rb_status rb_server_wait(rb_server_t s, uint64_t id, rb_request_status* st) {
    if (!s) return fail(RB_ERR_INVALID, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(RB_ERR_INVALID, "unknown request %llu", (unsigned long long)id);
    Req* q = it->second.get();
    s->cv_done.wait(lk, [&] { return q->state >= RB_REQ_DONE; });
    if (st) fill_status(*q, st);
    return RB_OK;
}

rb_status rb_server_drain(rb_server_t s) {
    if (!s) return fail(RB_ERR_INVALID, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] {
        if (!s->queue.empty()) return false;
        for (auto& kv : s->reqs)
            if (kv.second->state < RB_REQ_DONE) return false;
        return true;
    });
    return RB_OK;
}

rb_status rb_server_get_stats(rb_server_t s, rb_server_stats* out) {
    if (!s || !out) return fail(RB_ERR_INVALID, "NULL argument");
    std::lock_guard<std::mutex> g(s->mu);
    std::memset(out, 0, sizeof(*out));
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->failed = s->failed;
    out->rejected = s->rejected;
    out->deferrals = s->deferrals;
    out->rays = s->rays;
    out->peak_budget_bytes = s->peak;
    out->peak_arena_bytes = s->peak_reserved;
    out->arena_misses = s->arena_misses;
    if (!s->lat_ms.empty()) {
        std::vector<double> v = s->lat_ms;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            double pos = p * (v.size() - 1);
            size_t i = (size_t)pos;
            double f = pos - i;
            return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
        };
        out->p50_ms = pct(0.5);
        out->p99_ms = pct(0.99);
        double sum = 0;
        for (double x : v) sum += x;
        out->mean_ms = sum / v.size();
        out->max_ms = v.back();
        out->first_submit_to_last_done_s = std::chrono::duration<double>(s->last_done - s->first_submit).count();
    }
    return RB_OK;
}

rb_status rb_server_reset_stats(rb_server_t s) {
    if (!s) return fail(RB_ERR_INVALID, "NULL server");
    std::lock_guard<std::mutex> g(s->mu);
    s->submitted = s->completed = s->failed = s->rejected = s->deferrals = s->rays = 0;
    s->lat_ms.clear();
    s->have_first = false;
    s->peak = s->used;
    s->peak_reserved = s->reserved.load();
    s->arena_misses = 0;
    for (auto it = s->reqs.begin(); it != s->reqs.end();)
        it = it->second->state >= RB_REQ_DONE ? s->reqs.erase(it) : std::next(it);
    return RB_OK;
}

rb_status rb_server_inject_alloc_failures(rb_server_t s, uint32_t k) {
    if (!s) return fail(RB_ERR_INVALID, "NULL server");
    std::lock_guard<std::mutex> g(s->mu);
    s->inject_failures = k;
    return RB_OK;
}

}  // extern "C"
