// vz_api.cpp -- the C ABI (include/vz_api.h).  Never throws; every failure
// returns a vzStatus and leaves a message for vzGetLastError.  The C++ layer
// in venusaur_amd/core turns these into Exception throws, matching the
// reference's CHECK-macro convention (Core/Exception.h:16-27, 90-154).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <string>

#include "vz_engine.h"
#include "vz_internal.h"

namespace vz {

static thread_local std::string g_last_error;
static std::atomic<uint64_t> g_next_scene_uid{1};

vzStatus set_error(vzStatus s, const std::string& msg)
{
    g_last_error = msg;
    return s;
}

static hipStream_t as_stream(vzStream s) { return reinterpret_cast<hipStream_t>(s); }

static bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

vzStatus check_spheres(const vzSphere* s, uint32_t n)
{
    for (uint32_t i = 0; i < n; i++) {
        if (!finite3(s[i].center) || !std::isfinite(s[i].radius) || s[i].radius == 0.0f)
            return set_error(VZ_ERROR_INVALID_VALUE,
                             "sphere " + std::to_string(i) + " has a non-finite centre or a zero/non-finite radius");
    }
    return VZ_SUCCESS;
}

vzStatus check_launch_params(const vzLaunchParams* p)
{
    if (!p) return set_error(VZ_ERROR_INVALID_VALUE, "params is NULL");
    if (p->width < 2 || p->height < 2)
        return set_error(VZ_ERROR_INVALID_VALUE, "width and height must be >= 2 (the camera divides by w-1, h-1)");
    if ((uint64_t)p->width * p->height > (1ull << 31))
        return set_error(VZ_ERROR_INVALID_VALUE, "image too large");
    if (p->samples_per_pixel < 1) return set_error(VZ_ERROR_INVALID_VALUE, "samples_per_pixel must be >= 1");
    if (p->max_depth < 1 || p->max_depth > 255) return set_error(VZ_ERROR_INVALID_VALUE, "max_depth must be in [1, 255]");
    if (!p->accum) return set_error(VZ_ERROR_INVALID_VALUE, "accum buffer is NULL");
    if (p->samples_per_pixel > 1024 && !(p->flags & VZ_ACCUM_FP64))
        return set_error(VZ_ERROR_INVALID_VALUE, "more than 1024 samples per pixel need VZ_ACCUM_FP64");
    if (!finite3(p->origin) || !finite3(p->u) || !finite3(p->v) || !finite3(p->w) || !std::isfinite(p->lens_radius))
        return set_error(VZ_ERROR_INVALID_VALUE, "camera frame is not finite");
    return VZ_SUCCESS;
}

// Autotuned configurations per kind of device (name and CU count), for the
// life of the process: a context that has not tuned itself starts from the
// choice another context made on the same kind of device.
static std::mutex g_tune_mu;
static std::map<std::string, vzTuning> g_tune_cache;

static std::string device_key(vzContext ctx) { return ctx->name + "/" + std::to_string(ctx->num_cus); }

void tuning_cache_store(vzContext ctx, const vzTuning& t)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[device_key(ctx)] = t;
}

// Adopts the device's cached tuning if this context has none (ctx->mu held).
static void adopt_cached_tuning(vzContext ctx)
{
    if (ctx->tuned) return;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(device_key(ctx));
    if (it != g_tune_cache.end()) {
        ctx->tuning = it->second;
        ctx->tuned = true;
    }
}

static vzStatus get_engine(vzContext ctx, Engine** out)
{
    if (!ctx->engine) {
        EngineConfig cfg;
        adopt_cached_tuning(ctx);
        if (ctx->tuned) {
            cfg.queue_capacity = ctx->tuning.queue_capacity;
            cfg.tile_w = ctx->tuning.tile_w;
            cfg.tile_h = ctx->tuning.tile_h;
            cfg.iters_per_graph = ctx->tuning.iters_per_graph;
            cfg.refill_lanes = ctx->tuning.refill_lanes;
        }
        VZ_TRY(engine_create(ctx, cfg, &ctx->engine));
    }
    *out = ctx->engine;
    return VZ_SUCCESS;
}

struct DevCounters {   // per-call counters for the non-engine kernels
    unsigned long long* d = nullptr;
    ~DevCounters()
    {
        if (d) (void)hipFree(d);
    }
};

}  // namespace vz

using namespace vz;

extern "C" {

const char* vzStatusString(vzStatus s)
{
    switch (s) {
    case VZ_SUCCESS: return "VZ_SUCCESS";
    case VZ_ERROR_INVALID_VALUE: return "VZ_ERROR_INVALID_VALUE";
    case VZ_ERROR_OUT_OF_MEMORY: return "VZ_ERROR_OUT_OF_MEMORY";
    case VZ_ERROR_HIP: return "VZ_ERROR_HIP";
    case VZ_ERROR_NOT_READY: return "VZ_ERROR_NOT_READY";
    case VZ_ERROR_CAPACITY: return "VZ_ERROR_CAPACITY";
    case VZ_ERROR_UNSUPPORTED: return "VZ_ERROR_UNSUPPORTED";
    case VZ_ERROR_INTERNAL: return "VZ_ERROR_INTERNAL";
    case VZ_ERROR_NO_DEVICE: return "VZ_ERROR_NO_DEVICE";
    }
    return "VZ_ERROR_UNKNOWN";
}

const char* vzGetLastError(void) { return g_last_error.c_str(); }

vzStatus vzCreate(int device, vzContext* out)
{
    if (!out) return set_error(VZ_ERROR_INVALID_VALUE, "out is NULL");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return set_error(VZ_ERROR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= count) return set_error(VZ_ERROR_INVALID_VALUE, "device index out of range");
    VZ_HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    VZ_HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_error(VZ_ERROR_NO_DEVICE, std::string("libvz is built for gfx950, device is ") + prop.gcnArchName);
    unsigned long long* scratch = nullptr;
    VZ_TRY(trace_scratch_address(&scratch));  // this device's copy, looked up once with the device current
    vzContext c = new vzContext_st();
    c->device = device;
    c->trace_scratch = scratch;
    c->num_cus = (uint32_t)prop.multiProcessorCount;
    c->hbm_bytes = (uint64_t)prop.totalGlobalMem;
    // the marketing name can be empty (no entry in the driver's device-id
    // table); the architecture name is always set
    c->name = prop.name[0] ? std::string(prop.name) : std::string("AMD ") + prop.gcnArchName;
    *out = c;
    return VZ_SUCCESS;
}

vzStatus vzDestroy(vzContext ctx)
{
    if (!ctx) return set_error(VZ_ERROR_INVALID_VALUE, "ctx is NULL");
    const uint32_t ns = ctx->live_scenes.load(), nsrv = ctx->live_servers.load();
    if (ns || nsrv)
        return set_error(VZ_ERROR_INVALID_VALUE, std::to_string(ns) + " scene(s) and " + std::to_string(nsrv) +
                                                     " server(s) of this context are still alive");
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    if (ctx->engine) engine_destroy(ctx->engine);
    delete ctx;
    return VZ_SUCCESS;
}

vzStatus vzDeviceInfo(vzContext ctx, uint32_t* num_cus, uint64_t* hbm_bytes, char* name, size_t name_len)
{
    if (!ctx) return set_error(VZ_ERROR_INVALID_VALUE, "ctx is NULL");
    if (num_cus) *num_cus = ctx->num_cus;
    if (hbm_bytes) *hbm_bytes = ctx->hbm_bytes;
    if (name && name_len) {
        std::strncpy(name, ctx->name.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    return VZ_SUCCESS;
}

vzStatus vzSceneCreate(vzContext ctx, const vzSphere* spheres, const vzMaterial* materials, uint32_t n,
                       vzStream stream, vzScene* out)
{
    if (!ctx || !spheres || !materials || !out) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    *out = nullptr;
    if (n == 0 || n >= (1u << 30)) return set_error(VZ_ERROR_INVALID_VALUE, "sphere count must be in [1, 2^30)");
    VZ_TRY(check_spheres(spheres, n));
    for (uint32_t i = 0; i < n; i++)
        if (materials[i].type > VZ_DIELECTRIC)
            return set_error(VZ_ERROR_INVALID_VALUE, "material " + std::to_string(i) + " has an unknown type");
    VZ_HIP_TRY(hipSetDevice(ctx->device));
    vzScene s = new vzScene_st();
    s->ctx = ctx;
    s->n = n;
    s->uid = g_next_scene_uid.fetch_add(1);
    vzStatus rs = lbvh_alloc(s);
    hipStream_t st = as_stream(stream);
    if (rs == VZ_SUCCESS && (hipMemcpyAsync(s->d_sph_in, spheres, sizeof(vzSphere) * n, hipMemcpyHostToDevice, st) != hipSuccess ||
                             hipMemcpyAsync(s->d_mat_in, materials, sizeof(vzMaterial) * n, hipMemcpyHostToDevice, st) != hipSuccess))
        rs = set_error(VZ_ERROR_HIP, "scene upload failed");
    if (rs == VZ_SUCCESS) rs = lbvh_build(s, st);
    if (rs != VZ_SUCCESS) {
        (void)hipStreamSynchronize(st);
        lbvh_free(s);
        delete s;
        return rs;
    }
    ctx->live_scenes++;
    *out = s;
    return VZ_SUCCESS;
}

vzStatus vzSceneCreateDevice(vzContext ctx, const vzSphere* spheres, const vzMaterial* materials, uint32_t n,
                             vzStream stream, vzScene* out)
{
    if (!ctx || !spheres || !materials || !out) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    *out = nullptr;
    if (n == 0 || n >= (1u << 30)) return set_error(VZ_ERROR_INVALID_VALUE, "sphere count must be in [1, 2^30)");
    VZ_HIP_TRY(hipSetDevice(ctx->device));
    vzScene s = new vzScene_st();
    s->ctx = ctx;
    s->n = n;
    s->uid = g_next_scene_uid.fetch_add(1);
    vzStatus rs = lbvh_alloc(s);
    hipStream_t st = as_stream(stream);
    if (rs == VZ_SUCCESS &&
        (hipMemcpyAsync(s->d_sph_in, spheres, sizeof(vzSphere) * n, hipMemcpyDeviceToDevice, st) != hipSuccess ||
         hipMemcpyAsync(s->d_mat_in, materials, sizeof(vzMaterial) * n, hipMemcpyDeviceToDevice, st) != hipSuccess))
        rs = set_error(VZ_ERROR_HIP, "scene copy failed");
    uint32_t bad_s = 0xFFFFFFFFu, bad_m = 0xFFFFFFFFu;
    if (rs == VZ_SUCCESS) rs = lbvh_check_input(s, s->d_mat_in, st, &bad_s, &bad_m);
    if (rs == VZ_SUCCESS && bad_s != 0xFFFFFFFFu)
        rs = set_error(VZ_ERROR_INVALID_VALUE,
                       "sphere " + std::to_string(bad_s) + " has a non-finite centre or a zero/non-finite radius");
    if (rs == VZ_SUCCESS && bad_m != 0xFFFFFFFFu)
        rs = set_error(VZ_ERROR_INVALID_VALUE, "material " + std::to_string(bad_m) + " has an unknown type");
    if (rs == VZ_SUCCESS) rs = lbvh_build(s, st);
    if (rs != VZ_SUCCESS) {
        (void)hipStreamSynchronize(st);
        lbvh_free(s);
        delete s;
        return rs;
    }
    ctx->live_scenes++;
    *out = s;
    return VZ_SUCCESS;
}

vzStatus vzSceneDestroy(vzScene scene)
{
    if (!scene) return set_error(VZ_ERROR_INVALID_VALUE, "scene is NULL");
    const uint32_t refs = scene->server_refs.load();
    if (refs)
        return set_error(VZ_ERROR_INVALID_VALUE, "scene is named by " + std::to_string(refs) +
                                                     " queued or running server request(s)");
    (void)hipSetDevice(scene->ctx->device);
    (void)hipDeviceSynchronize();
    scene->ctx->live_scenes--;
    lbvh_free(scene);
    delete scene;
    return VZ_SUCCESS;
}

vzStatus vzAccelBuild(vzScene scene, vzStream stream)
{
    if (!scene) return set_error(VZ_ERROR_INVALID_VALUE, "scene is NULL");
    return lbvh_build(scene, as_stream(stream));
}

vzStatus vzSceneUpdateSpheres(vzScene scene, const vzSphere* spheres, vzStream stream)
{
    if (!scene || !spheres) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    VZ_TRY(check_spheres(spheres, scene->n));
    VZ_HIP_TRY(hipMemcpyAsync(scene->d_sph_in, spheres, sizeof(vzSphere) * scene->n, hipMemcpyHostToDevice,
                              as_stream(stream)));
    return lbvh_refit(scene, as_stream(stream));
}

vzStatus vzSceneUpdateSpheresDevice(vzScene scene, const vzSphere* spheres, vzStream stream)
{
    if (!scene || !spheres) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    hipStream_t st = as_stream(stream);
    // check a copy first, so a rejected update leaves the scene unchanged
    vzSphere* tmp = nullptr;
    VZ_HIP_TRY(hipMallocAsync((void**)&tmp, sizeof(vzSphere) * scene->n, st));
    vzStatus rs = VZ_SUCCESS;
    if (hipMemcpyAsync(tmp, spheres, sizeof(vzSphere) * scene->n, hipMemcpyDeviceToDevice, st) != hipSuccess)
        rs = set_error(VZ_ERROR_HIP, "sphere copy failed");
    float4* keep = scene->d_sph_in;
    uint32_t bad_s = 0xFFFFFFFFu, bad_m = 0xFFFFFFFFu;
    if (rs == VZ_SUCCESS) {
        scene->d_sph_in = reinterpret_cast<float4*>(tmp);
        rs = lbvh_check_input(scene, nullptr, st, &bad_s, &bad_m);
        scene->d_sph_in = keep;
    }
    if (rs == VZ_SUCCESS && bad_s != 0xFFFFFFFFu)
        rs = set_error(VZ_ERROR_INVALID_VALUE,
                       "sphere " + std::to_string(bad_s) + " has a non-finite centre or a zero/non-finite radius");
    if (rs == VZ_SUCCESS &&
        hipMemcpyAsync(scene->d_sph_in, tmp, sizeof(vzSphere) * scene->n, hipMemcpyDeviceToDevice, st) != hipSuccess)
        rs = set_error(VZ_ERROR_HIP, "sphere copy failed");
    (void)hipFreeAsync(tmp, st);
    VZ_TRY(rs);
    return lbvh_refit(scene, st);
}

vzStatus vzAccelRefit(vzScene scene, vzStream stream)
{
    if (!scene) return set_error(VZ_ERROR_INVALID_VALUE, "scene is NULL");
    return lbvh_refit(scene, as_stream(stream));
}

vzStatus vzAccelInfoGet(vzScene scene, vzAccelInfo* out)
{
    if (!scene || !out) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    out->num_prims = scene->n;
    out->num_nodes = scene->n > 1 ? scene->n - 1 : 0;
    out->depth = scene->depth;
    out->node_bytes = 32;
    out->device_bytes = scene->device_bytes;
    out->build_ms = scene->build_ms;
    out->num_wide_nodes = scene->nw;
    out->wide_depth = scene->wdepth;
    out->builder = scene->builder;
    return VZ_SUCCESS;
}

vzStatus vzAccelDownloadWide(vzScene scene, void* wide_nodes, vzStream stream)
{
    if (!scene || !wide_nodes) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    hipStream_t st = as_stream(stream);
    if (scene->nw)
        VZ_HIP_TRY(hipMemcpyAsync(wide_nodes, scene->d_wnodes, (size_t)scene->nw * 64, hipMemcpyDeviceToHost, st));
    VZ_HIP_TRY(hipStreamSynchronize(st));
    return VZ_SUCCESS;
}

vzStatus vzAccelDownload(vzScene scene, void* nodes, uint32_t* leaf_prims, uint32_t* morton, vzStream stream)
{
    if (!scene) return set_error(VZ_ERROR_INVALID_VALUE, "scene is NULL");
    hipStream_t st = as_stream(stream);
    uint32_t n = scene->n;
    if (nodes && n > 1) VZ_HIP_TRY(hipMemcpyAsync(nodes, scene->d_nodes, (size_t)(n - 1) * 32, hipMemcpyDeviceToHost, st));
    if (leaf_prims) VZ_HIP_TRY(hipMemcpyAsync(leaf_prims, scene->d_leaf_prim, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    if (morton) VZ_HIP_TRY(hipMemcpyAsync(morton, scene->d_morton, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    VZ_HIP_TRY(hipStreamSynchronize(st));
    return VZ_SUCCESS;
}

vzStatus vzTraceClosest(vzScene scene, const vzRay* rays, uint32_t n, vzHit* hits, vzStream stream, vzStats* stats)
{
    if (!scene || (n && (!rays || !hits))) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    hipStream_t st = as_stream(stream);
    // per-workgroup partial counters, then the 4 summed counters
    const size_t nparts = (size_t)trace_blocks(n) * kTraceParts;
    DevCounters c;
    VZ_HIP_TRY(hipMalloc(&c.d, (nparts + 8) * sizeof(unsigned long long)));
    Event e0, e1;
    VZ_EVENT_TRY(e0);
    VZ_EVENT_TRY(e1);
    VZ_HIP_TRY(hipEventRecord(e0.e, st));
    vzStatus rs = trace_rays(scene->view(), rays, n, hits, st, stats != nullptr, c.d, c.d + nparts);
    unsigned long long h[8] = {};
    if (rs == VZ_SUCCESS && hipEventRecord(e1.e, st) != hipSuccess)
        rs = set_error(VZ_ERROR_HIP, "hipEventRecord failed");
    if (rs == VZ_SUCCESS && stats &&
        hipMemcpyAsync(h, c.d + nparts, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st) != hipSuccess)
        rs = set_error(VZ_ERROR_HIP, "hipMemcpyAsync of the trace counters failed");
    // c.d is freed on return: the kernel must be done with it, on every path
    if (hipStreamSynchronize(st) != hipSuccess && rs == VZ_SUCCESS)
        rs = set_error(VZ_ERROR_HIP, "hipStreamSynchronize failed");
    VZ_TRY(rs);
    if (stats) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0.e, e1.e);
        std::memset(stats, 0, sizeof(*stats));
        stats->segments = h[0];
        stats->node_visits = h[2];
        stats->sphere_tests = h[3];
        stats->render_ms = ms;
        stats->trace_ms = ms;
        stats->trace_launches = 1;
    }
    return VZ_SUCCESS;
}

vzStatus vzTraceClosestNoAlloc(vzScene scene, const vzRay* rays, uint32_t n, vzHit* hits, vzStream stream,
                               vzStats* stats)
{
    if (!scene || (n && (!rays || !hits))) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    if (n > kTraceScratchBlocks * kTraceBlockSize)
        return set_error(VZ_ERROR_CAPACITY, "vzTraceClosestNoAlloc traces at most " +
                                                std::to_string(kTraceScratchBlocks * kTraceBlockSize) +
                                                " rays per call (its fixed scratch); split the call or use "
                                                "vzTraceClosest");
    vzContext ctx = scene->ctx;
    int cur = -1;
    VZ_HIP_TRY(hipGetDevice(&cur));
    if (cur != ctx->device)
        return set_error(VZ_ERROR_INVALID_VALUE, "the current device is not the scene's device");
    if (!ctx->trace_scratch) return set_error(VZ_ERROR_INTERNAL, "trace scratch address not set");
    hipStream_t st = as_stream(stream);
    unsigned long long* parts = ctx->trace_scratch;
    unsigned long long* counters = parts + (size_t)kTraceScratchBlocks * kTraceParts;
    VZ_TRY(trace_rays(scene->view(), rays, n, hits, st, stats != nullptr, parts, counters));
    if (stats) {
        unsigned long long h[4] = {};
        VZ_HIP_TRY(hipMemcpyAsync(h, counters, sizeof(h), hipMemcpyDeviceToHost, st));
        VZ_HIP_TRY(hipStreamSynchronize(st));
        std::memset(stats, 0, sizeof(*stats));
        stats->segments = h[0];
        stats->node_visits = h[2];
        stats->sphere_tests = h[3];
        stats->trace_launches = 1;
    }
    return VZ_SUCCESS;
}

vzStatus vzLaunch(vzScene scene, const vzLaunchParams* params, vzStream stream, vzStats* stats)
{
    if (!scene) return set_error(VZ_ERROR_INVALID_VALUE, "scene is NULL");
    VZ_TRY(check_launch_params(params));
    hipStream_t st = as_stream(stream);
    vzContext ctx = scene->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (params->flags & VZ_MEGAKERNEL) {
        DevCounters c;
        VZ_HIP_TRY(hipMalloc(&c.d, 8 * sizeof(unsigned long long)));
        VZ_HIP_TRY(hipMemsetAsync(c.d, 0, 8 * sizeof(unsigned long long), st));
        Event e0, e1;
        VZ_EVENT_TRY(e0);
        VZ_EVENT_TRY(e1);
        VZ_HIP_TRY(hipEventRecord(e0.e, st));
        vzStatus rs = megakernel_launch(scene->view(), *params, st, c.d, (params->flags & VZ_COUNT_VISITS) != 0);
        unsigned long long h[8] = {};
        if (rs == VZ_SUCCESS && hipEventRecord(e1.e, st) != hipSuccess)
            rs = set_error(VZ_ERROR_HIP, "hipEventRecord failed");
        if (rs == VZ_SUCCESS && hipMemcpyAsync(h, c.d, sizeof(h), hipMemcpyDeviceToHost, st) != hipSuccess)
            rs = set_error(VZ_ERROR_HIP, "hipMemcpyAsync of the counters failed");
        // c.d is freed on return: the kernel must be done with it, on every path
        if (hipStreamSynchronize(st) != hipSuccess && rs == VZ_SUCCESS)
            rs = set_error(VZ_ERROR_HIP, "hipStreamSynchronize failed");
        VZ_TRY(rs);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0.e, e1.e);
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->segments = h[0];
            stats->paths = h[1];
            stats->node_visits = h[2];
            stats->sphere_tests = h[3];
            stats->render_ms = ms;
            stats->trace_ms = ms;
            stats->trace_launches = 1;
            stats->iterations = 1;
        }
        return VZ_SUCCESS;
    }
    Engine* e = nullptr;
    VZ_TRY(get_engine(ctx, &e));
    std::vector<RunItem> items(1);
    items[0].scene = scene;
    items[0].params = *params;
    RunOptions opt;
    opt.use_graph = false;
    opt.count_visits = (params->flags & VZ_COUNT_VISITS) != 0;
    opt.time_trace = (params->flags & VZ_TIME_STAGES) != 0;
    // Rays in flight for one request: four image passes, at least 1M, at
    // most the queue.  Measured on MI355X (same box, plain launches one after
    // another): a 120-request mix of the serving distribution 2345 against
    // 1929 Mray-segments/s with the whole 8M queue in use, the 64-scene batch
    // mix 2106 against 1998; 1080p frames fill the 8M queue either way (book
    // scene 11.2 Gray-segments/s).  Six passes, or two and three (which also
    // cut 1080p below 8M), were slower; floors from 256k to 1M were equal.
    constexpr uint64_t kPassesInFlight = 4, kMinRaysInFlight = 1u << 20;
    opt.rays_in_flight =
        std::max<uint64_t>(kMinRaysInFlight, kPassesInFlight * (uint64_t)params->width * params->height);
    RunResult rr;
    VZ_TRY(engine_run(e, items, st, opt, &rr));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->segments = rr.segments;
        stats->paths = rr.paths;
        stats->node_visits = rr.node_visits;
        stats->sphere_tests = rr.sphere_tests;
        stats->hits = rr.hits;
        stats->render_ms = rr.render_ms;
        stats->trace_ms = rr.trace_ms;
        stats->trace_launches = rr.trace_launches;
        stats->iterations = rr.iterations;
    }
    return VZ_SUCCESS;
}

vzStatus vzDefaultTuning(vzContext ctx, vzTuning* out)
{
    if (!ctx || !out) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    adopt_cached_tuning(ctx);
    if (ctx->tuned) {
        *out = ctx->tuning;
        return VZ_SUCCESS;
    }
    EngineConfig cfg;
    out->queue_capacity = cfg.queue_capacity;
    out->tile_w = cfg.tile_w;
    out->tile_h = cfg.tile_h;
    out->iters_per_graph = cfg.iters_per_graph;
    out->refill_lanes = cfg.refill_lanes;
    out->slice_split = kBatchSplit;
    return VZ_SUCCESS;
}

static bool same_tuning(const EngineConfig& c, const vzTuning& t)
{
    return c.queue_capacity == t.queue_capacity && c.tile_w == t.tile_w && c.tile_h == t.tile_h &&
           c.iters_per_graph == t.iters_per_graph && c.refill_lanes == t.refill_lanes;
}

vzStatus vzBatchLaunch(vzContext ctx, const vzBatchItem* items, uint32_t n, const vzTuning* tuning, vzStream stream,
                       vzStats* stats)
{
    if (!ctx || (n && !items)) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    if (n > kMaxRequests) return set_error(VZ_ERROR_CAPACITY, "at most 1024 items per batch");
    std::vector<RunItem> run(n);
    for (uint32_t i = 0; i < n; i++) {
        if (!items[i].scene || items[i].scene->ctx != ctx)
            return set_error(VZ_ERROR_INVALID_VALUE, "batch item " + std::to_string(i) + " has no scene of this context");
        VZ_TRY(check_launch_params(&items[i].params));
        if (items[i].params.flags & VZ_MEGAKERNEL)
            return set_error(VZ_ERROR_INVALID_VALUE, "VZ_MEGAKERNEL is not a batched mode");
        run[i].scene = items[i].scene;
        run[i].params = items[i].params;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    vzTuning t;
    if (tuning) t = *tuning; else VZ_TRY(vzDefaultTuning(ctx, &t));
    if (t.queue_capacity == 0) t.queue_capacity = EngineConfig().queue_capacity;
    if (t.tile_w == 0) t.tile_w = EngineConfig().tile_w;
    if (t.tile_h == 0) t.tile_h = EngineConfig().tile_h;
    if (t.iters_per_graph < 2) t.iters_per_graph = EngineConfig().iters_per_graph;
    if (t.refill_lanes > 65) t.refill_lanes = 65;
    if (ctx->engine && !same_tuning(ctx->engine->cfg, t)) {
        (void)hipStreamSynchronize(as_stream(stream));
        engine_destroy(ctx->engine);
        ctx->engine = nullptr;
    }
    if (!ctx->engine) {
        EngineConfig cfg;
        cfg.queue_capacity = t.queue_capacity;
        cfg.tile_w = t.tile_w;
        cfg.tile_h = t.tile_h;
        cfg.iters_per_graph = t.iters_per_graph;
        cfg.refill_lanes = t.refill_lanes;
        VZ_TRY(engine_create(ctx, cfg, &ctx->engine));
    }
    RunOptions opt;
    opt.use_graph = true;
    bool count = false, timing = false;
    for (uint32_t i = 0; i < n; i++) {
        count |= (items[i].params.flags & VZ_COUNT_VISITS) != 0;
        timing |= (items[i].params.flags & VZ_TIME_STAGES) != 0;
    }
    opt.count_visits = count;
    opt.time_trace = timing;
    opt.slice_split = t.slice_split ? t.slice_split : kBatchSplit;
    RunResult rr;
    VZ_TRY(engine_run(ctx->engine, run, as_stream(stream), opt, &rr));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->segments = rr.segments;
        stats->paths = rr.paths;
        stats->node_visits = rr.node_visits;
        stats->sphere_tests = rr.sphere_tests;
        stats->hits = rr.hits;
        stats->render_ms = rr.render_ms;
        stats->trace_ms = rr.trace_ms;
        stats->trace_launches = rr.trace_launches;
        stats->iterations = rr.iterations;
    }
    return VZ_SUCCESS;
}

uint64_t vzRequestFootprint(const vzLaunchParams* params)
{
    if (!params) return 0;
    return (uint64_t)request_bytes(*params);
}

// ---- host helpers -------------------------------------------------------------
// The uniform float the scene generators use: the top 24 bits of one
// std::mt19937 output, exact in [0, 1).
static float uniform01(std::mt19937& g) { return (float)(g() >> 8) * (1.0f / 16777216.0f); }

static void material_draw(std::mt19937& g, float choose, vzMaterial* m)
{
    std::memset(m, 0, sizeof(*m));
    if ((double)choose < 0.8) {
        float a[3], b[3];
        for (int k = 0; k < 3; k++) a[k] = uniform01(g);
        for (int k = 0; k < 3; k++) b[k] = uniform01(g);
        m->type = VZ_LAMBERTIAN;
        for (int k = 0; k < 3; k++) m->albedo[k] = a[k] * b[k];
    } else if ((double)choose < 0.95) {
        m->type = VZ_METAL;
        for (int k = 0; k < 3; k++) m->albedo[k] = 0.5f + (1.0f - 0.5f) * uniform01(g);
        m->fuzz = 0.0f + (0.5f - 0.0f) * uniform01(g);
    } else {
        m->type = VZ_DIELECTRIC;
        m->ir = 1.5f;
    }
}

vzStatus vzGenerateBookScene(uint32_t seed, vzSphere* spheres, vzMaterial* materials, uint32_t cap, uint32_t* n_out)
{
    if (!n_out) return set_error(VZ_ERROR_INVALID_VALUE, "n_out is NULL");
    std::mt19937 g(seed);
    uint32_t n = 0;
    auto push = [&](float x, float y, float z, float r, const vzMaterial& m) {
        if (n < cap && spheres && materials) {
            spheres[n].center[0] = x;
            spheres[n].center[1] = y;
            spheres[n].center[2] = z;
            spheres[n].radius = r;
            materials[n] = m;
        }
        n++;
    };
    vzMaterial m;
    std::memset(&m, 0, sizeof(m));
    m.type = VZ_LAMBERTIAN;
    m.albedo[0] = m.albedo[1] = m.albedo[2] = 0.5f;
    push(0.0f, -1000.0f, 0.0f, 1000.0f, m);
    for (int a = -11; a < 11; a++) {
        for (int b = -11; b < 11; b++) {
            float choose = uniform01(g);
            float ux = uniform01(g);
            float uz = uniform01(g);
            float cx = (float)((double)a + 0.9 * (double)ux);
            float cy = 0.2f;
            float cz = (float)((double)b + 0.9 * (double)uz);
            float dx = cx - 4.0f, dy = cy - 0.2f, dz = cz - 0.0f;
            float len = std::sqrt(dx * dx + dy * dy + dz * dz);
            if ((double)len > 0.9) {
                material_draw(g, choose, &m);
                push(cx, cy, cz, 0.2f, m);
            }
        }
    }
    std::memset(&m, 0, sizeof(m));
    m.type = VZ_DIELECTRIC;
    m.ir = 1.5f;
    push(0.0f, 1.0f, 0.0f, 1.0f, m);
    std::memset(&m, 0, sizeof(m));
    m.type = VZ_LAMBERTIAN;
    m.albedo[0] = 0.4f; m.albedo[1] = 0.2f; m.albedo[2] = 0.1f;
    push(-4.0f, 1.0f, 0.0f, 1.0f, m);
    std::memset(&m, 0, sizeof(m));
    m.type = VZ_METAL;
    m.albedo[0] = 0.7f; m.albedo[1] = 0.6f; m.albedo[2] = 0.5f;
    push(4.0f, 1.0f, 0.0f, 1.0f, m);
    *n_out = n;
    if (n > cap) return set_error(VZ_ERROR_CAPACITY, "book scene needs " + std::to_string(n) + " spheres");
    return VZ_SUCCESS;
}

vzStatus vzGenerateBoxScene(uint32_t n_small, uint32_t seed, vzSphere* spheres, vzMaterial* materials)
{
    if (!spheres || !materials) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    const float ln_lo = -0x1.26bb1cp+2f;  // logf(0.01f)
    const float ln_hi = -0x1.62e430p-1f;  // logf(0.5f)
    std::mt19937 g(seed);
    std::memset(&materials[0], 0, sizeof(vzMaterial));
    spheres[0].center[0] = 0.0f;
    spheres[0].center[1] = -1000.0f;
    spheres[0].center[2] = 0.0f;
    spheres[0].radius = 1000.0f;
    materials[0].type = VZ_LAMBERTIAN;
    materials[0].albedo[0] = materials[0].albedo[1] = materials[0].albedo[2] = 0.5f;
    for (uint32_t i = 1; i <= n_small; i++) {
        float x = -100.0f + 200.0f * uniform01(g);
        float y = 0.0f + 20.0f * uniform01(g);
        float z = -100.0f + 200.0f * uniform01(g);
        float r = std::exp(ln_lo + (ln_hi - ln_lo) * uniform01(g));
        float choose = uniform01(g);
        spheres[i].center[0] = x;
        spheres[i].center[1] = y;
        spheres[i].center[2] = z;
        spheres[i].radius = r;
        material_draw(g, choose, &materials[i]);
    }
    return VZ_SUCCESS;
}

vzStatus vzCameraLookAt(const float lookfrom[3], const float lookat[3], float vfov_deg, float aspect, float aperture,
                        float focus_dist, vzLaunchParams* params)
{
    if (!lookfrom || !lookat || !params) return set_error(VZ_ERROR_INVALID_VALUE, "NULL argument");
    auto nrm = [](float v[3]) {
        float inv = 1.0f / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        v[0] *= inv;
        v[1] *= inv;
        v[2] *= inv;
    };
    auto cross = [](const float a[3], const float b[3], float o[3]) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    float f[3] = {lookat[0] - lookfrom[0], lookat[1] - lookfrom[1], lookat[2] - lookfrom[2]};
    nrm(f);
    float W[3] = {f[0] * focus_dist, f[1] * focus_dist, f[2] * focus_dist};
    const float up[3] = {0.0f, 1.0f, 0.0f};
    float U[3], V[3];
    cross(W, up, U);
    nrm(U);
    cross(U, W, V);
    nrm(V);
    float theta = vfov_deg * 0.01745329251994329577f;
    float h = std::tan(theta * 0.5f);
    float vh = 2.0f * h;
    float vw = aspect * vh;
    float su = focus_dist * vw, sv = focus_dist * vh;
    for (int k = 0; k < 3; k++) {
        params->origin[k] = lookfrom[k];
        params->u[k] = U[k] * su;
        params->v[k] = V[k] * sv;
        params->w[k] = W[k];
    }
    params->lens_radius = aperture * 0.5f;
    return VZ_SUCCESS;
}

}  // extern "C"
