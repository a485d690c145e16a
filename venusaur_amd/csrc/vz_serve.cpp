// vz_serve.cpp -- the serving form and the autotuner.
//
// Serving: a host thread owns a wavefront engine and a non-blocking stream.
// Requests are queued FIFO by vzServerSubmit and admitted into the engine's
// request table whenever a slot is free and the request's buffers fit the
// HBM budget; admitted requests join the running wavefront at the next
// iteration, so new work overlaps with requests already on the GPU
// (continuous batching).  BVHs stay resident (vzScene objects owned by the
// caller) and request buffers are allocated stream-ordered, so admission and
// release never synchronise the device.  The nearest reference analogue is
// the interactive loop that keeps the GAS and accumulator resident across
// frames (Core/Core.cpp:358-430, Core/Renderer.h:37-54).
//
// Autotuner: times the caller's batch over candidate queue capacities and
// camera-ray tile shapes and keeps the fastest (segments per millisecond),
// cached per context.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>

#include "vz_engine.h"
#include "vz_internal.h"

using namespace vz;

namespace {

struct Pending {
    uint64_t id;
    const vzScene_st* scene;
    vzLaunchParams params;
    size_t bytes;
    double submit_ms;
};

struct Active {
    uint64_t id;
    const vzScene_st* scene;
    uint32_t slot;
    size_t bytes;
    uint32_t first_step;  // first engine_step whose status includes the slot
    double submit_ms, admit_ms;
};

}  // namespace

struct vzServer_st {
    vzContext ctx = nullptr;
    vzServeConfig cfg{};
    Engine* engine = nullptr;
    hipStream_t stream = nullptr;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Pending> pending;
    std::unordered_map<uint64_t, vzRequestResult> results;
    std::vector<Active> active;
    uint64_t next_id = 1;
    size_t used_bytes = 0;
    bool stop = false;
    vzStatus fatal = VZ_SUCCESS;
    std::string fatal_msg;
    std::chrono::steady_clock::time_point t0;

    double now_ms() const
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    void loop();
    // A request no longer needs its scene (finished, rejected or dropped).
    static void release_scene(const vzScene_st* sc)
    {
        const_cast<vzScene_st*>(sc)->server_refs.fetch_sub(1);
    }
};

void vzServer_st::loop()
{
    (void)hipSetDevice(ctx->device);
    uint32_t k = 0;          // steps enqueued
    uint32_t harvested = 0;  // steps whose status has been read
    hipEvent_t ev[2];
    (void)hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
    (void)hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
    auto fail = [&](vzStatus s) {
        std::lock_guard<std::mutex> lk(mu);
        fatal = s;
        fatal_msg = vzGetLastError();
        stop = true;
        cv_done.notify_all();
    };
    auto reject = [&](const Pending& p, uint32_t status) {
        release_scene(p.scene);
        vzRequestResult r{};
        r.request_id = p.id;
        r.submit_ms = p.submit_ms;
        r.admit_ms = r.finish_ms = now_ms();
        r.status = status;
        results[p.id] = r;
    };
    for (;;) {
        // 1. admit queued requests while a slot is free and the HBM budget allows
        {
            std::unique_lock<std::mutex> lk(mu);
            if (k == harvested && active.empty() && pending.empty()) {
                cv_done.notify_all();
                cv_work.wait(lk, [&] { return stop || !pending.empty(); });
                if (stop && pending.empty()) break;
            }
            while (!pending.empty() && active.size() < cfg.max_active) {
                Pending& p = pending.front();
                if (p.bytes > cfg.hbm_budget_bytes) {  // can never fit
                    reject(p, VZ_ERROR_CAPACITY);
                    pending.pop_front();
                    continue;
                }
                if (used_bytes + p.bytes > cfg.hbm_budget_bytes) break;  // wait for memory
                RunItem it;
                it.scene = p.scene;
                it.params = p.params;
                uint32_t slot = 0;
                vzStatus s = engine_admit(engine, it, stream, &slot);
                if (s == VZ_ERROR_CAPACITY || s == VZ_ERROR_OUT_OF_MEMORY) {
                    if (active.empty() && k == harvested) {  // nothing will free up
                        reject(p, s);
                        pending.pop_front();
                        continue;
                    }
                    break;
                }
                if (s != VZ_SUCCESS) {
                    fatal = s;
                    fatal_msg = vzGetLastError();
                    stop = true;
                    cv_done.notify_all();
                    break;
                }
                Active a;
                a.id = p.id;
                a.scene = p.scene;
                a.slot = slot;
                a.bytes = p.bytes;
                a.first_step = k;
                a.submit_ms = p.submit_ms;
                a.admit_ms = now_ms();
                used_bytes += p.bytes;
                active.push_back(a);
                pending.pop_front();
            }
            if (fatal != VZ_SUCCESS) break;
        }
        // 2. keep up to two blocks of iterations in flight
        if (!active.empty() && k - harvested < 2u) {
            if (engine_step(engine, stream, k) != VZ_SUCCESS || hipEventRecord(ev[k & 1u], stream) != hipSuccess) {
                fail(VZ_ERROR_HIP);
                break;
            }
            k++;
        }
        // 3. harvest the oldest block once two are queued (or nothing is left to queue)
        if (k > harvested && (k - harvested >= 2u || active.empty())) {
            uint32_t h = harvested;
            if (hipEventSynchronize(ev[h & 1u]) != hipSuccess) {
                fail(VZ_ERROR_HIP);
                break;
            }
            harvested++;
            double t = now_ms();
            std::vector<Active> still;
            for (const Active& a : active) {
                bool act = true;
                unsigned long long seg = 0;
                if (h >= a.first_step) engine_slot_status(engine, h, a.slot, &act, &seg);
                if (act) {
                    still.push_back(a);
                    continue;
                }
                if (engine_release_slot(engine, a.slot, stream) != VZ_SUCCESS) {
                    fail(VZ_ERROR_HIP);
                    break;
                }
                release_scene(a.scene);
                std::lock_guard<std::mutex> lk(mu);
                vzRequestResult r{};
                r.request_id = a.id;
                r.submit_ms = a.submit_ms;
                r.admit_ms = a.admit_ms;
                r.finish_ms = t;
                r.segments = seg;
                r.status = VZ_SUCCESS;
                results[a.id] = r;
                used_bytes -= a.bytes;
            }
            std::lock_guard<std::mutex> lk(mu);
            active.swap(still);
            if (fatal != VZ_SUCCESS) break;
            cv_done.notify_all();
        }
    }
    (void)hipStreamSynchronize(stream);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
}

extern "C" {

vzStatus vzServerCreate(vzContext ctx, const vzServeConfig* cfg, vzServer* out)
{
    if (!ctx || !cfg || !out) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    *out = nullptr;
    VZ_HIP_TRY(hipSetDevice(ctx->device));
    vzServer s = new vzServer_st();
    s->ctx = ctx;
    s->cfg = *cfg;
    if (s->cfg.max_active == 0 || s->cfg.max_active > kMaxRequests) s->cfg.max_active = kMaxRequests;
    if (s->cfg.hbm_budget_bytes == 0) s->cfg.hbm_budget_bytes = ctx->hbm_bytes / 2;
    EngineConfig ec;
    // Serving defaults: a 48M-ray queue (4.8 GB for its two parities), 8x8
    // camera tiles and the queue split 4 ways per resident step workgroup
    // (kServeSplit).  On the first 600 requests of the 2000-request stream,
    // all in flight at once: 5648 Mray-segments/s against 5576 for 32M, 5303
    // for 24M and 5690 for 64M, all at split 4 (16x4 tiles at 24M: 5020).
    ec.queue_capacity = 12u << 22;
    ec.tile_w = 8;
    ec.tile_h = 8;
    if (s->cfg.tuning.queue_capacity) ec.queue_capacity = s->cfg.tuning.queue_capacity;
    if (s->cfg.tuning.tile_w) ec.tile_w = s->cfg.tuning.tile_w;
    if (s->cfg.tuning.tile_h) ec.tile_h = s->cfg.tuning.tile_h;
    if (s->cfg.tuning.iters_per_graph >= 2) ec.iters_per_graph = s->cfg.tuning.iters_per_graph;
    ec.refill_lanes = std::min<uint32_t>(s->cfg.tuning.refill_lanes, 65u);  // 0: the kernel variant's default
    vzStatus rs = engine_create(ctx, ec, &s->engine);
    if (rs == VZ_SUCCESS && hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        rs = set_error(VZ_ERROR_HIP, "hipStreamCreateWithFlags failed");
    const uint32_t split = s->cfg.tuning.slice_split ? s->cfg.tuning.slice_split : kServeSplit;
    if (rs == VZ_SUCCESS) rs = engine_serve_init(s->engine, s->stream, split);
    if (rs != VZ_SUCCESS) {
        if (s->engine) engine_destroy(s->engine);
        if (s->stream) (void)hipStreamDestroy(s->stream);
        delete s;
        return rs;
    }
    s->t0 = std::chrono::steady_clock::now();
    s->worker = std::thread([s] { s->loop(); });
    ctx->live_servers++;
    *out = s;
    return VZ_SUCCESS;
}

vzStatus vzServerSubmit(vzServer srv, vzScene scene, const vzLaunchParams* params, uint64_t* request_id)
{
    if (!srv || !scene || !params || !request_id) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    if (scene->ctx != srv->ctx) return set_error(VZ_ERROR_INVALID_VALUE, "scene belongs to another context");
    VZ_TRY(check_launch_params(params));
    if (params->flags & VZ_MEGAKERNEL) return set_error(VZ_ERROR_INVALID_VALUE, "VZ_MEGAKERNEL is not a serving mode");
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->fatal != VZ_SUCCESS) return set_error(srv->fatal, "server stopped: " + srv->fatal_msg);
    Pending p;
    p.id = srv->next_id++;
    p.scene = scene;
    scene->server_refs.fetch_add(1);  // vzSceneDestroy refuses until the request is done
    p.params = *params;
    p.bytes = request_bytes(*params);
    p.submit_ms = srv->now_ms();
    srv->pending.push_back(p);
    *request_id = p.id;
    srv->cv_work.notify_one();
    return VZ_SUCCESS;
}

vzStatus vzServerPoll(vzServer srv, uint64_t request_id, vzRequestResult* out)
{
    if (!srv || !out) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->results.find(request_id);
    if (it == srv->results.end()) {
        if (request_id == 0 || request_id >= srv->next_id) return set_error(VZ_ERROR_INVALID_VALUE, "unknown request id");
        if (srv->fatal != VZ_SUCCESS) return set_error(srv->fatal, "server stopped: " + srv->fatal_msg);
        return VZ_ERROR_NOT_READY;
    }
    *out = it->second;
    return VZ_SUCCESS;
}

vzStatus vzServerDrain(vzServer srv)
{
    if (!srv) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    std::unique_lock<std::mutex> lk(srv->mu);
    srv->cv_done.wait(lk, [&] {
        return srv->fatal != VZ_SUCCESS || (srv->pending.empty() && srv->active.empty());
    });
    if (srv->fatal != VZ_SUCCESS) return set_error(srv->fatal, "server stopped: " + srv->fatal_msg);
    return VZ_SUCCESS;
}

vzStatus vzServerDestroy(vzServer srv)
{
    if (!srv) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
        srv->cv_work.notify_all();
    }
    if (srv->worker.joinable()) srv->worker.join();
    (void)hipStreamSynchronize(srv->stream);
    // after a fatal error requests may remain; they will never run
    for (const Pending& p : srv->pending) vzServer_st::release_scene(p.scene);
    for (const Active& a : srv->active) vzServer_st::release_scene(a.scene);
    srv->pending.clear();
    srv->active.clear();
    engine_destroy(srv->engine);
    (void)hipStreamDestroy(srv->stream);
    srv->ctx->live_servers--;
    delete srv;
    return VZ_SUCCESS;
}

vzStatus vzAutotune(vzContext ctx, const vzBatchItem* items, uint32_t n, vzStream stream, vzTuning* out)
{
    if (!ctx || !items || n == 0 || !out) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument or empty batch");
    // Times the batch itself, twice per candidate, and keeps each
    // candidate's faster run: the candidates are often within a few percent,
    // and a single run let the noise pick.  The engine's render time starts
    // after the queue allocation and graph capture, so no warm-up run is
    // needed.  Reduced-sample probes shrank each request's pass ring and
    // favoured small queues that lost on the real batch.
    std::vector<vzBatchItem> probe(items, items + n);
    for (auto& it : probe) it.params.flags &= ~(uint32_t)(VZ_COUNT_VISITS | VZ_TIME_STAGES);
    vzTuning best{};
    double best_rate = -1.0;
    auto time = [&](const vzTuning& cand) -> vzStatus {
        double rate = 0.0;
        for (int rep = 0; rep < 2; rep++) {
            vzStats st{};
            VZ_TRY(vzBatchLaunch(ctx, probe.data(), n, &cand, stream, &st));
            if (st.render_ms > 0.0) rate = std::max(rate, (double)st.segments / st.render_ms);
        }
        if (rate > best_rate) {
            best_rate = rate;
            best = cand;
        }
        return VZ_SUCCESS;
    };
    // stage 1: queue capacity x camera tile at the default refill and split
    const uint32_t caps[] = {1u << 22, 1u << 23, 3u << 22, 6u << 22, 8u << 22, 12u << 22};
    const uint32_t tiles[][2] = {{32, 2}, {16, 4}, {8, 8}};
    for (uint32_t cap : caps)
        for (auto& t : tiles) {
            vzTuning cand{};
            cand.queue_capacity = cap;
            cand.tile_w = t[0];
            cand.tile_h = t[1];
            cand.iters_per_graph = 16;
            cand.refill_lanes = 0;
            cand.slice_split = kBatchSplit;
            VZ_TRY(time(cand));
        }
    // stage 2: at the best queue and tile, the refill threshold, then the split
    const vzTuning base = best;
    for (uint32_t refill : {25u, 33u, 49u, 57u}) {
        vzTuning cand = base;
        cand.refill_lanes = refill;
        VZ_TRY(time(cand));
    }
    const vzTuning base2 = best;
    for (uint32_t split : {1u, 2u, 8u}) {
        vzTuning cand = base2;
        cand.slice_split = split;
        VZ_TRY(time(cand));
    }
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->tuning = best;
        ctx->tuned = true;
    }
    tuning_cache_store(ctx, best);
    *out = best;
    return VZ_SUCCESS;
}

}  // extern "C"
