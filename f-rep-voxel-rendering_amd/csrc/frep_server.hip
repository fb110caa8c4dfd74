// frep_server.hip -- request server: voxelize + render requests with
// continuous batching on one GPU.
//
// The reference has no serving layer: its application voxelizes and renders
// one scene per frame (vulkan/src/application.cpp:60-93).  This server keeps
// that path's semantics per request -- a request's tree is the tree
// frep_voxelize builds and its image is what frep_render returns -- and
// changes only how requests share the device:
//
//   * Scenes (tape + grid) are registered once; their trees stay resident in
//     device memory and later requests only render them.
//   * A dedicated host thread drains the request queue into batches.  The
//     trees a batch needs are built by the batched level kernels (one
//     persistent launch per level for all of them), then every request of
//     the batch is rendered, all on one of several streams ("slots").
//   * While a batch is on the device the thread keeps admitting: arrivals
//     form the next batch on a free slot, so batches overlap on the GPU.
//   * Admission is by device memory: a scene is built only if the resident
//     trees plus those being built plus its estimate fit the budget; the
//     least recently used idle trees are evicted to make room (and rebuilt
//     if requested again).  First capacities come from the largest tree
//     seen so far for the same grid, so most builds need no regrowth.
//   * Latency (submit -> outputs written) of every completed request feeds
//     p50 / p99 statistics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/frep.h"
#include "frep_host.hpp"
#include "frep_internal.hpp"

using namespace frep;

namespace {

using Clock = std::chrono::steady_clock;

constexpr uint32_t kMaxSlots = 4;

struct Scene {
    frep_tape_t tape = 0;
    frep_grid grid{};
    frep_tree_t tree = 0;          // created on first build
    bool resident = false;         // tree built and its buffers held
    bool building = false;         // a batch in flight is building it
    uint32_t pins = 0;             // in-flight requests reading the tree
    uint64_t bytes = 0;            // device bytes charged to the budget
    uint64_t last_use = 0;
    float built_time = 0.0f;       // tape time the resident tree was built at
};

inline bool same_time(float a, float b) { return std::memcmp(&a, &b, 4) == 0; }

// Output layout of one request: rgba, hit ids, normals at 256-byte aligned
// offsets of one block (the same layout in the slot's device buffer, at
// `dbase`, and in the request's pinned block).
struct OutLayout {
    size_t o_rgba = 0, o_hit = 0, o_nrm = 0, total = 0;
    size_t b_rgba = 0, b_hit = 0, b_nrm = 0;
};

struct Ticket {
    frep_request req{};
    Clock::time_point submitted;
    bool done = false;
    frep_result result{};
    OutLayout out;
    void* h_block = nullptr;             // pinned outputs, copied out by the waiter
};

// A request as its batch carries it (copied from the ticket and the scene
// at admission, so the launch needs no lock on the server's tables).
struct SlotReq {
    uint64_t ticket;
    frep_scene_t scene;
    frep_tree_t tree;
    frep_camera cam;
    size_t dbase;                        // offset in the slot's device buffer
    OutLayout out;
    void* h_block;                       // pinned outputs of the request
};

// A tree the batch builds.
struct SlotJob {
    frep_scene_t scene;
    frep_tape_t tape;
    frep_tree_t tree;
    frep_grid grid;
    bool has_hint;
    CapHint hint;
    uint64_t est;                        // budget charged while building
};

// Renders of a slot run on side streams beside each other (one request per
// stream in turn), forked from and joined back into the slot's stream.
constexpr uint32_t kRenderLanes = 4;

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    hipStream_t rstream[kRenderLanes] = {};
    hipEvent_t rfork = nullptr;
    hipEvent_t rjoin[kRenderLanes] = {};
    bool busy = false;
    BatchStage stage;
    std::vector<SlotJob> build;          // trees this batch builds
    std::vector<Job> jobs;               // their voxelization jobs (same order)
    std::vector<SlotReq> reqs;
    size_t out_need = 0;                 // device output bytes of the batch
    void* d_out = nullptr;
    size_t out_bytes = 0;
    uint64_t seq = 0;                    // launch order
};

std::string dims_key(const frep_grid& g) {
    std::string k;
    for (uint32_t L = 0; L < g.n_levels; ++L) k += std::to_string(g.dims[L]) + ",";
    return k;
}

}  // namespace

struct frep_server {
    frep_ctx* ctx = nullptr;
    frep_server_config cfg{};
    std::string err;

    std::mutex mu;
    std::condition_variable cv;          // queue / stop changes
    std::condition_variable done_cv;     // a ticket completed
    bool stop = false;
    std::thread th;

    std::vector<Scene> scenes;
    std::deque<uint64_t> queue;          // tickets not yet launched
    std::unordered_map<uint64_t, Ticket> tickets;
    uint64_t next_ticket = 1;
    Slot slots[kMaxSlots];
    uint32_t n_slots = 2;
    uint64_t launch_seq = 0;
    uint64_t use_clock = 0;
    std::map<std::string, CapHint> hints;

    // statistics
    uint64_t resident_bytes = 0;         // built trees + estimates of trees being built
    uint64_t submitted = 0, completed = 0, failed = 0, batches = 0, voxelizations = 0,
             evictions = 0, batched_reqs = 0;
    std::vector<float> latencies;
    bool have_first = false;
    Clock::time_point first_submit, last_done;
};

namespace {

frep_status srv_fail(frep_server* srv, frep_status s, const std::string& msg) {
    srv->err = msg;
    return s;
}

// Evicts the least recently used idle resident scene; false if none.
// Caller holds srv->mu.
bool evict_one(frep_server* srv) {
    Scene* victim = nullptr;
    for (Scene& sc : srv->scenes)
        if (sc.resident && !sc.building && sc.pins == 0 && (!victim || sc.last_use < victim->last_use))
            victim = &sc;
    if (!victim) return false;
    frep_ctx* ctx = srv->ctx;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        TreeObj* t = get_tree(ctx, victim->tree);
        if (t) {
            free_tree_buffers(ctx, t);
            t->built = false;
        }
    }
    victim->resident = false;
    srv->resident_bytes -= std::min(srv->resident_bytes, victim->bytes);
    victim->bytes = 0;
    ++srv->evictions;
    return true;
}

void complete(frep_server* srv, uint64_t ticket, frep_status st, uint32_t voxelized,
              uint32_t batch_size) {
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end()) return;
    Ticket& tk = it->second;
    const Clock::time_point now = Clock::now();
    tk.done = true;
    tk.result.status = st;
    tk.result.voxelized = voxelized;
    tk.result.batch_size = batch_size;
    tk.result.latency_ms =
        (float)std::chrono::duration<double, std::milli>(now - tk.submitted).count();
    if (st == FREP_OK) {
        ++srv->completed;
        srv->latencies.push_back(tk.result.latency_ms);
    } else {
        ++srv->failed;
    }
    srv->last_done = now;
}

OutLayout layout_outputs(const frep_request& r) {
    const frep_camera& c = r.camera;
    const uint32_t spp = c.spp ? c.spp : 1;
    const size_t px = (size_t)c.width * c.height;
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    OutLayout o;
    o.b_rgba = r.rgba ? px * 4 : 0;
    o.b_hit = r.hit_ids ? px * spp * 8 : 0;
    o.b_nrm = r.normals ? px * spp * 12 : 0;
    o.o_rgba = 0;
    o.o_hit = al(o.b_rgba);
    o.o_nrm = al(o.o_hit + o.b_hit);
    o.total = al(o.o_nrm + o.b_nrm);
    return o;
}

// Enqueues the renders of the slot's requests (only those of scene `only`
// when it is not 0) and the copies of their outputs to the pinned blocks.
frep_status enqueue_renders(frep_server* srv, Slot& S, frep_scene_t only) {
    frep_ctx* ctx = srv->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    // fork: every render lane starts after the slot stream's work so far
    // (the batch that builds the trees)
    HIP_TRY(hipEventRecord(S.rfork, S.stream));
    bool used[kRenderLanes] = {};
    uint32_t next = 0;
    for (const SlotReq& r : S.reqs) {
        if (only && r.scene != only) continue;
        const uint32_t ln = next++ % kRenderLanes;
        const hipStream_t rs = S.rstream[ln];
        if (!used[ln]) HIP_TRY(hipStreamWaitEvent(rs, S.rfork, 0));
        used[ln] = true;
        const TreeObj* t = get_tree(ctx, r.tree);
        char* d = static_cast<char*>(S.d_out) + r.dbase;
        char* h = static_cast<char*>(r.h_block);
        const OutLayout& o = r.out;
        const RenderArgs a = render_args(
            t, &r.cam, o.b_rgba ? (uint8_t*)(d + o.o_rgba) : nullptr,
            o.b_hit ? (unsigned long long*)(d + o.o_hit) : nullptr,
            o.b_nrm ? (float*)(d + o.o_nrm) : nullptr);
        HIP_TRY(launch_render(rs, a, t->maxs, t->slot_count));
        if (o.b_rgba)
            HIP_TRY(hipMemcpyAsync(h + o.o_rgba, d + o.o_rgba, o.b_rgba, hipMemcpyDeviceToHost, rs));
        if (o.b_hit)
            HIP_TRY(hipMemcpyAsync(h + o.o_hit, d + o.o_hit, o.b_hit, hipMemcpyDeviceToHost, rs));
        if (o.b_nrm)
            HIP_TRY(hipMemcpyAsync(h + o.o_nrm, d + o.o_nrm, o.b_nrm, hipMemcpyDeviceToHost, rs));
    }
    // join: the slot stream (and its completion event) waits for every lane
    for (uint32_t k = 0; k < kRenderLanes; ++k)
        if (used[k]) {
            HIP_TRY(hipEventRecord(S.rjoin[k], S.rstream[k]));
            HIP_TRY(hipStreamWaitEvent(S.stream, S.rjoin[k], 0));
        }
    return FREP_OK;
}

// Admission: moves requests from the queue into slot S in FIFO order.  A
// request renders its scene's tree at its tape time (its own with
// FREP_REQUEST_AT_TIME, else the scene's): it joins a build of that time in
// this batch, or renders the resident tree if it was built at that time, or
// asks for a build -- which must wait while a batch in flight builds or
// reads the tree, or while the tree does not fit the budget.  A request left
// waiting holds back the later requests of its scene, so each scene's
// requests are admitted in submission order.  Caller holds srv->mu.
// Returns the number admitted.
size_t admit(frep_server* srv, Slot& S) {
    frep_ctx* ctx = srv->ctx;
    S.build.clear();
    S.jobs.clear();
    S.reqs.clear();
    S.out_need = 0;
    auto job_of = [&](frep_scene_t sid) -> const SlotJob* {
        for (const SlotJob& j : S.build)
            if (j.scene == sid) return &j;
        return nullptr;
    };
    std::vector<frep_scene_t> held;      // scenes with an earlier request left waiting
    auto is_held = [&](frep_scene_t sid) {
        return std::find(held.begin(), held.end(), sid) != held.end();
    };
    for (auto it = srv->queue.begin(); it != srv->queue.end();) {
        if (S.reqs.size() >= srv->cfg.max_batch) break;
        const uint64_t tid = *it;
        Ticket& tk = srv->tickets[tid];
        const frep_scene_t sid = tk.req.scene;
        Scene& sc = srv->scenes[sid - 1];
        if (is_held(sid)) {
            ++it;
            continue;
        }
        const float te = (tk.req.flags & FREP_REQUEST_AT_TIME) ? tk.req.tape_time
                                                               : sc.grid.tape_time;
        const SlotJob* mj = job_of(sid);
        bool build = false;
        if (mj) {
            if (!same_time(mj->grid.tape_time, te)) {   // another time: a later batch
                held.push_back(sid);
                ++it;
                continue;
            }
        } else if (sc.building) {                        // being built by a batch in flight
            held.push_back(sid);
            ++it;
            continue;
        } else if (!(sc.resident && same_time(sc.built_time, te))) {
            if (sc.resident && sc.pins > 0) {            // requests in flight read the tree
                held.push_back(sid);
                ++it;
                continue;
            }
            build = true;
        }
        if (build) {
            SlotJob j{};
            j.scene = sid;
            j.tape = sc.tape;
            j.grid = sc.grid;
            j.grid.tape_time = te;
            auto h = srv->hints.find(dims_key(sc.grid));
            j.has_hint = h != srv->hints.end();
            if (j.has_hint) j.hint = h->second;
            {
                std::lock_guard<std::mutex> lk(ctx->mu);
                j.est = estimate_job_bytes(get_tape(ctx, sc.tape), &j.grid,
                                           j.has_hint ? &j.hint : nullptr);
            }
            // a resident tree rebuilt at another time gives back its bytes;
            // marked as building so that eviction leaves it alone
            const uint64_t credit = sc.resident ? std::min(sc.bytes, srv->resident_bytes) : 0;
            sc.building = true;
            while (srv->resident_bytes - credit + j.est > srv->cfg.device_budget_bytes &&
                   evict_one(srv)) {
            }
            if (srv->resident_bytes - credit + j.est > srv->cfg.device_budget_bytes) {
                sc.building = false;
                // does not fit now; if nothing is in flight it never will
                bool any_busy = false;
                for (uint32_t k = 0; k < srv->n_slots; ++k) any_busy = any_busy || srv->slots[k].busy;
                if (!any_busy && S.reqs.empty()) {
                    complete(srv, tid, FREP_OUT_OF_MEMORY, 0, 0);
                    srv->err = "scene does not fit the device budget";
                    it = srv->queue.erase(it);
                    srv->done_cv.notify_all();
                    continue;
                }
                held.push_back(sid);
                ++it;
                continue;
            }
            if (sc.resident) {
                srv->resident_bytes -= credit;
                sc.bytes = 0;
                sc.resident = false;
            }
            if (!sc.tree) {
                std::lock_guard<std::mutex> lk(ctx->mu);
                if (new_tree(ctx, &sc.tree) != FREP_OK) {
                    sc.building = false;
                    complete(srv, tid, FREP_OUT_OF_MEMORY, 0, 0);
                    it = srv->queue.erase(it);
                    srv->done_cv.notify_all();
                    continue;
                }
            }
            j.tree = sc.tree;
            srv->resident_bytes += j.est;
            sc.building = true;
            S.build.push_back(j);
        }
        sc.pins += 1;
        sc.last_use = ++srv->use_clock;
        SlotReq r{};
        r.ticket = tid;
        r.scene = sid;
        r.tree = sc.tree;
        r.cam = tk.req.camera;
        r.out = layout_outputs(tk.req);
        r.dbase = S.out_need;
        r.h_block = nullptr;
        if (r.out.total && ctx->hpool.alloc(r.out.total, &r.h_block) != hipSuccess) {
            // cannot take outputs: fail this request
            sc.pins -= 1;
            complete(srv, tid, FREP_OUT_OF_MEMORY, 0, 0);
            it = srv->queue.erase(it);
            srv->done_cv.notify_all();
            continue;
        }
        tk.out = r.out;
        tk.h_block = r.h_block;
        S.out_need += r.out.total;
        S.reqs.push_back(r);
        it = srv->queue.erase(it);
    }
    if (S.reqs.empty()) {
        // scenes reserved for builds with no request left cannot happen: every
        // build is pushed together with its first request
        return 0;
    }
    S.busy = true;          // reserved; the work is enqueued without srv->mu
    S.seq = ++srv->launch_seq;
    ++srv->batches;
    srv->batched_reqs += S.reqs.size();
    srv->voxelizations += S.build.size();
    return S.reqs.size();
}

// Enqueues the admitted batch of slot S: one batched voxelization of the
// trees to build, the renders, and the completion event.  Runs without
// srv->mu (it touches only the slot and ctx, under ctx->mu).
frep_status enqueue_slot(frep_server* srv, Slot& S) {
    frep_ctx* ctx = srv->ctx;
    if (S.out_bytes < S.out_need) {
        if (S.d_out) HIP_TRY(hipFree(S.d_out));
        S.d_out = nullptr;
        S.out_bytes = 0;
        const size_t cap = S.out_need + S.out_need / 2 + 4096;
        HIP_TRY(hipMalloc(&S.d_out, cap));
        S.out_bytes = cap;
    }
    if (!S.build.empty()) {
        std::lock_guard<std::mutex> lk(ctx->mu);
        S.jobs.assign(S.build.size(), Job{});
        std::vector<uint32_t> idx(S.jobs.size());
        for (size_t k = 0; k < S.jobs.size(); ++k) {
            const SlotJob& j = S.build[k];
            const frep_status s =
                prepare_job(ctx, get_tape(ctx, j.tape), get_tree(ctx, j.tree), &j.grid, &S.jobs[k],
                            S.stream, j.has_hint ? &j.hint : nullptr);
            if (s != FREP_OK) return s;
            const frep_status e = ensure_job(ctx, &S.jobs[k]);
            if (e != FREP_OK) return e;
            idx[k] = (uint32_t)k;
        }
        const BatchPlan P = plan_batch(S.jobs, idx, 1, ctx->root_ops, ctx->cu_count);
        const frep_status es = ensure_stage(ctx, &S.stage, P);
        if (es != FREP_OK) return es;
        fill_batch(P, S.jobs, idx, static_cast<uint8_t*>(S.stage.h_stage));
        const std::vector<uint32_t> grid(P.launches.size(), ctx->level_grid);
        const frep_status q = enqueue_batch(ctx, &S.stage, S.stream, P, grid, nullptr);
        if (q != FREP_OK) return q;
    }
    {
        const frep_status s = enqueue_renders(srv, S, 0);
        if (s != FREP_OK) return s;
    }
    HIP_TRY(hipEventRecord(S.done, S.stream));
    return FREP_OK;
}

// Rebuilds job k of slot S alone until its tree fits, then re-renders the
// requests of its scene (synchronously on the slot's stream).
frep_status rebuild(frep_server* srv, Slot& S, size_t k) {
    frep_ctx* ctx = srv->ctx;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        std::vector<Job> one(1, S.jobs[k]);
        const std::vector<uint32_t> idx(1, 0u);
        BatchStage tmp;
        frep_status st = FREP_OK;
        for (uint32_t attempt = 0;; ++attempt) {
            st = regrow_job(ctx, &one[0], attempt);
            if (st != FREP_OK) break;
            st = ensure_job(ctx, &one[0]);
            if (st != FREP_OK) break;
            const BatchPlan P = plan_batch(one, idx, 1, ctx->root_ops, ctx->cu_count);
            st = ensure_stage(ctx, &tmp, P);
            if (st != FREP_OK) break;
            fill_batch(P, one, idx, static_cast<uint8_t*>(tmp.h_stage));
            const std::vector<uint32_t> grid(P.launches.size(), ctx->level_grid);
            st = enqueue_batch(ctx, &tmp, S.stream, P, grid, nullptr);
            if (st != FREP_OK) break;
            const hipError_t e = hipStreamSynchronize(S.stream);
            if (e != hipSuccess) {
                st = fail(ctx, FREP_DEVICE, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
                break;
            }
            std::memcpy(one[0].t->h_counters, tmp.h_cnt, CNT_WORDS * 4);
            if (one[0].t->h_counters[CNT_OVERFLOW] == 0) break;
        }
        free_stage(&tmp);
        S.jobs[k] = one[0];
        if (st != FREP_OK) return st;
    }
    const frep_status s = enqueue_renders(srv, S, S.build[k].scene);
    if (s != FREP_OK) return s;
    HIP_TRY(hipStreamSynchronize(S.stream));
    return FREP_OK;
}

bool builds(const Slot& S, frep_scene_t sid) {
    for (const SlotJob& j : S.build)
        if (j.scene == sid) return true;
    return false;
}

// Completes the batch on slot S (its event has fired).  Caller holds srv->mu.
void finalize(frep_server* srv, Slot& S) {
    frep_ctx* ctx = srv->ctx;
    std::vector<frep_scene_t> failed_scenes;
    for (size_t k = 0; k < S.jobs.size(); ++k) {
        const SlotJob& j = S.build[k];
        Scene& sc = srv->scenes[j.scene - 1];
        TreeObj* t = S.jobs[k].t;
        std::memcpy(t->h_counters, S.stage.h_cnt + k * CNT_WORDS, CNT_WORDS * 4);
        frep_status st = FREP_OK;
        if (t->h_counters[CNT_OVERFLOW] != 0) st = rebuild(srv, S, k);
        srv->resident_bytes -= std::min(srv->resident_bytes, j.est);
        sc.building = false;
        if (st != FREP_OK) {
            failed_scenes.push_back(j.scene);
            std::lock_guard<std::mutex> lk(ctx->mu);
            free_tree_buffers(ctx, t);
            t->built = false;
            continue;
        }
        t->built = true;
        sc.resident = true;
        sc.built_time = j.grid.tape_time;
        {
            std::lock_guard<std::mutex> lk(ctx->mu);
            sc.bytes = tree_device_bytes(t);
        }
        srv->resident_bytes += sc.bytes;
        // capacities for the next tree of this grid
        CapHint& h = srv->hints[dims_key(sc.grid)];
        for (uint32_t L = 0; L < sc.grid.n_levels; ++L)
            h.nodes[L] = std::max(h.nodes[L], host_nodes(t, L));
        h.pool_words = std::max<uint64_t>(h.pool_words, host_pool_top(t));
    }
    const uint32_t bs = (uint32_t)S.reqs.size();
    for (const SlotReq& r : S.reqs) {
        Scene& sc = srv->scenes[r.scene - 1];
        if (sc.pins) --sc.pins;
        const bool bad = std::find(failed_scenes.begin(), failed_scenes.end(), r.scene) !=
                         failed_scenes.end();
        // outputs stay in the request's pinned block until its waiter copies them
        complete(srv, r.ticket, bad ? FREP_OUT_OF_MEMORY : FREP_OK, builds(S, r.scene) ? 1u : 0u,
                 bs);
    }
    S.busy = false;
    S.reqs.clear();
    S.build.clear();
    S.jobs.clear();
    srv->done_cv.notify_all();
}

// An error while enqueueing fails the slot's requests.  Caller holds srv->mu.
void fail_slot(frep_server* srv, Slot& S, frep_status st) {
    for (const SlotJob& j : S.build) {
        Scene& sc = srv->scenes[j.scene - 1];
        sc.building = false;
        srv->resident_bytes -= std::min(srv->resident_bytes, j.est);
    }
    for (const SlotReq& r : S.reqs) {
        Scene& sc = srv->scenes[r.scene - 1];
        if (sc.pins) --sc.pins;
        complete(srv, r.ticket, st, 0, (uint32_t)S.reqs.size());
    }
    S.reqs.clear();
    S.jobs.clear();
    S.build.clear();
    S.busy = false;
    srv->done_cv.notify_all();
}

void server_loop(frep_server* srv) {
    (void)hipSetDevice(srv->ctx->device);
    std::unique_lock<std::mutex> lk(srv->mu);
    for (;;) {
        // complete every batch whose work has finished
        for (uint32_t k = 0; k < srv->n_slots; ++k) {
            Slot& S = srv->slots[k];
            if (S.busy && hipEventQuery(S.done) == hipSuccess) finalize(srv, S);
        }
        bool any_busy = false;
        Slot* free_slot = nullptr;
        for (uint32_t k = 0; k < srv->n_slots; ++k) {
            any_busy = any_busy || srv->slots[k].busy;
            if (!srv->slots[k].busy && !free_slot) free_slot = &srv->slots[k];
        }
        if (srv->stop && srv->queue.empty() && !any_busy) break;
        if (free_slot && !srv->queue.empty()) {
            const size_t before = srv->queue.size();
            if (admit(srv, *free_slot)) {
                // enqueue without the server lock, so submitters never wait on
                // HIP calls; only this thread touches the slots
                lk.unlock();
                const frep_status s = enqueue_slot(srv, *free_slot);
                lk.lock();
                if (s != FREP_OK) fail_slot(srv, *free_slot, s);
                continue;
            }
            if (srv->queue.size() != before) continue;
        }
        if (!free_slot && !srv->queue.empty()) {
            // every slot busy: wait for the oldest batch
            Slot* oldest = nullptr;
            for (uint32_t k = 0; k < srv->n_slots; ++k)
                if (srv->slots[k].busy && (!oldest || srv->slots[k].seq < oldest->seq))
                    oldest = &srv->slots[k];
            const hipEvent_t e = oldest->done;
            lk.unlock();
            (void)hipEventSynchronize(e);
            lk.lock();
            continue;
        }
        if (any_busy) {
            // poll the batches in flight; a new request wakes the loop
            srv->cv.wait_for(lk, std::chrono::microseconds(100));
        } else {
            srv->cv.wait(lk, [&] { return srv->stop || !srv->queue.empty(); });
        }
    }
}

}  // namespace

extern "C" {

frep_status frep_server_create(frep_ctx* ctx, const frep_server_config* cfg, frep_server** out) {
    if (!ctx || !out) return FREP_INVALID_ARG;
    *out = nullptr;
    std::unique_ptr<frep_server> srv(new frep_server());
    srv->ctx = ctx;
    if (cfg) srv->cfg = *cfg;
    if (srv->cfg.device_budget_bytes == 0) srv->cfg.device_budget_bytes = 64ull << 30;
    if (srv->cfg.max_batch == 0) srv->cfg.max_batch = kBatchMax;
    if (srv->cfg.max_batch > kBatchMax) return fail(ctx, FREP_INVALID_ARG, "max_batch above 64");
    if (srv->cfg.max_inflight == 0) srv->cfg.max_inflight = 2;
    if (srv->cfg.max_inflight > kMaxSlots)
        return fail(ctx, FREP_INVALID_ARG, "max_inflight above 4");
    srv->n_slots = srv->cfg.max_inflight;
    HIP_TRY(hipSetDevice(ctx->device));
    for (uint32_t k = 0; k < srv->n_slots; ++k) {
        HIP_TRY(hipStreamCreateWithFlags(&srv->slots[k].stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&srv->slots[k].done, hipEventDisableTiming));
        for (auto& st : srv->slots[k].rstream)
            HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&srv->slots[k].rfork, hipEventDisableTiming));
        for (auto& e : srv->slots[k].rjoin) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    frep_server* p = srv.get();
    srv->th = std::thread(server_loop, p);
    *out = srv.release();
    return FREP_OK;
}

frep_status frep_server_destroy(frep_server* srv) {
    if (!srv) return FREP_INVALID_ARG;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv.notify_all();
    if (srv->th.joinable()) srv->th.join();
    frep_ctx* ctx = srv->ctx;
    (void)hipSetDevice(ctx->device);
    for (uint32_t k = 0; k < srv->n_slots; ++k) {
        Slot& S = srv->slots[k];
        if (S.stream) (void)hipStreamSynchronize(S.stream);
        free_stage(&S.stage);
        if (S.d_out) (void)hipFree(S.d_out);
        if (S.done) (void)hipEventDestroy(S.done);
        for (auto& st : S.rstream)
            if (st) {
                (void)hipStreamSynchronize(st);
                (void)hipStreamDestroy(st);
            }
        if (S.rfork) (void)hipEventDestroy(S.rfork);
        for (auto& e : S.rjoin)
            if (e) (void)hipEventDestroy(e);
        if (S.stream) (void)hipStreamDestroy(S.stream);
    }
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        for (Scene& sc : srv->scenes) {
            TreeObj* t = get_tree(ctx, sc.tree);
            if (!t) continue;
            free_tree_buffers(ctx, t);
            ctx->dpool.free(t->d_counters);
            ctx->hpool.free(t->h_counters);
            ctx->trees[sc.tree - 1].reset();
        }
    }
    delete srv;
    return FREP_OK;
}

const char* frep_server_last_error(const frep_server* srv) {
    return srv ? srv->err.c_str() : "null server";
}

frep_status frep_server_scene(frep_server* srv, frep_tape_t tape, const frep_grid* grid,
                              frep_scene_t* out) {
    if (!srv || !grid || !out) return FREP_INVALID_ARG;
    frep_ctx* ctx = srv->ctx;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        const TapeObj* tp = get_tape(ctx, tape);
        if (!tp) return srv_fail(srv, FREP_INVALID_ARG, "unknown tape handle");
        const frep_status gs = check_grid(ctx, grid);
        if (gs != FREP_OK) return srv_fail(srv, gs, ctx->err);
        if (grid->n_levels > 1 && std::max<uint32_t>(tp->n_mm, 1) > max_minmax(ctx->lds_per_block))
            return srv_fail(srv, FREP_UNSUPPORTED, "tape has too many MIN/MAX instructions");
    }
    std::lock_guard<std::mutex> lk(srv->mu);
    Scene sc;
    sc.tape = tape;
    sc.grid = *grid;
    srv->scenes.push_back(sc);
    *out = (frep_scene_t)srv->scenes.size();
    return FREP_OK;
}

frep_status frep_server_submit(frep_server* srv, const frep_request* req, uint64_t* ticket) {
    if (!srv || !req || !ticket) return FREP_INVALID_ARG;
    const frep_camera& cam = req->camera;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (cam.width == 0 || cam.height == 0 || cam.width > 16384 || cam.height > 16384)
            return srv_fail(srv, FREP_INVALID_ARG, "image size out of range");
        if (cam.light_count > 4) return srv_fail(srv, FREP_INVALID_ARG, "at most 4 lights");
        if (req->scene == 0 || req->scene > srv->scenes.size())
            return srv_fail(srv, FREP_INVALID_ARG, "unknown scene");
        if (srv->stop) return srv_fail(srv, FREP_INVALID_ARG, "server is stopping");
        const uint64_t id = srv->next_ticket++;
        Ticket& tk = srv->tickets[id];
        tk.req = *req;
        tk.submitted = Clock::now();
        if (!srv->have_first) {
            srv->have_first = true;
            srv->first_submit = tk.submitted;
        }
        ++srv->submitted;
        srv->queue.push_back(id);
        *ticket = id;
    }
    srv->cv.notify_all();
    return FREP_OK;
}

// Copies a completed request's outputs from its pinned block to the caller's
// buffers (on the caller's thread) and releases the block.
static void deliver(frep_server* srv, const Ticket& tk) {
    const char* h = static_cast<const char*>(tk.h_block);
    if (h && tk.result.status == FREP_OK) {
        const OutLayout& o = tk.out;
        if (o.b_rgba) std::memcpy(tk.req.rgba, h + o.o_rgba, o.b_rgba);
        if (o.b_hit) std::memcpy(tk.req.hit_ids, h + o.o_hit, o.b_hit);
        if (o.b_nrm) std::memcpy(tk.req.normals, h + o.o_nrm, o.b_nrm);
    }
    srv->ctx->hpool.free(tk.h_block);
}

frep_status frep_server_poll(frep_server* srv, uint64_t ticket, frep_result* out, int* done) {
    if (!srv || !done) return FREP_INVALID_ARG;
    Ticket tk;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        auto it = srv->tickets.find(ticket);
        if (it == srv->tickets.end()) return srv_fail(srv, FREP_INVALID_ARG, "unknown ticket");
        *done = it->second.done ? 1 : 0;
        if (!it->second.done) return FREP_OK;
        tk = it->second;
        srv->tickets.erase(it);
    }
    deliver(srv, tk);
    if (out) *out = tk.result;
    return FREP_OK;
}

frep_status frep_server_wait(frep_server* srv, uint64_t ticket, frep_result* out) {
    if (!srv) return FREP_INVALID_ARG;
    Ticket tk;
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        auto it = srv->tickets.find(ticket);
        if (it == srv->tickets.end()) return srv_fail(srv, FREP_INVALID_ARG, "unknown ticket");
        srv->done_cv.wait(lk, [&] { return srv->tickets[ticket].done; });
        it = srv->tickets.find(ticket);
        tk = it->second;
        srv->tickets.erase(it);
    }
    deliver(srv, tk);
    if (out) *out = tk.result;
    return FREP_OK;
}

frep_status frep_server_stats_get(frep_server* srv, frep_server_stats* out) {
    if (!srv || !out) return FREP_INVALID_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    *out = frep_server_stats{};
    out->submitted = srv->submitted;
    out->completed = srv->completed;
    out->failed = srv->failed;
    out->batches = srv->batches;
    out->voxelizations = srv->voxelizations;
    out->evictions = srv->evictions;
    out->resident_bytes = srv->resident_bytes;
    out->budget_bytes = srv->cfg.device_budget_bytes;
    for (const Scene& sc : srv->scenes) out->resident_scenes += sc.resident ? 1u : 0u;
    out->queued = (uint32_t)srv->queue.size();
    if (!srv->latencies.empty()) {
        std::vector<float> v = srv->latencies;
        std::sort(v.begin(), v.end());
        auto pct = [&](double q) {
            const size_t i = (size_t)std::min<double>(v.size() - 1, std::ceil(q * v.size()) - 1);
            return v[i];
        };
        out->p50_ms = pct(0.50);
        out->p99_ms = pct(0.99);
        double sum = 0;
        for (float x : v) sum += x;
        out->mean_ms = (float)(sum / v.size());
        out->max_ms = v.back();
    }
    out->mean_batch = srv->batches ? (float)((double)srv->batched_reqs / srv->batches) : 0.0f;
    if (srv->have_first && srv->completed)
        out->elapsed_s =
            (float)std::chrono::duration<double>(srv->last_done - srv->first_submit).count();
    return FREP_OK;
}

frep_status frep_server_reset_stats(frep_server* srv) {
    if (!srv) return FREP_INVALID_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->submitted = srv->completed = srv->failed = srv->batches = srv->voxelizations = 0;
    srv->evictions = srv->batched_reqs = 0;
    srv->latencies.clear();
    srv->have_first = false;
    return FREP_OK;
}

}  // extern "C"
