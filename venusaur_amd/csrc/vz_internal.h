// vz_internal.h -- host-side internals of libvz shared across translation
// units: context, scene, error reporting, kernel-launch entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/vz_api.h"

namespace vz {

// Largest BVH depth the traversal stack supports (levels from the root to
// the deepest leaf).  Karras trees over 30-bit codes with index tie-breaks
// stay well below this for any scene that fits in memory.
constexpr uint32_t kMaxDepth = 64;
// Traversal stack of the 4-wide tree: up to 3 entries per wide level above
// the node being visited, kWideStackS of them per lane in LDS and the oldest
// kWideOverflow in private memory.  Scenes whose wide depth needs more are
// refused at build time.
constexpr uint32_t kWideStackS = 8;
constexpr uint32_t kWideOverflow = 96;
constexpr uint32_t kMaxWideDepth = (kWideOverflow + kWideStackS) / 3 + 1;
constexpr uint32_t kMaxRequests = 1024;  // device request-table size
// vzTraceClosest: workgroup size, counters per workgroup (segments, node
// visits, sphere tests), and the workgroups the no-allocation entry point's
// fixed partials array holds (rays per call <= 65536 x 256).
constexpr uint32_t kTraceBlockSize = 256;
constexpr int kTraceParts = 3;
constexpr uint32_t kTraceScratchBlocks = 65536;

// Records a message for vzGetLastError and returns the status.
vzStatus set_error(vzStatus s, const std::string& msg);
// Validates a launch's parameters (sizes, samples, depth, buffers, a finite
// camera frame); shared by vzLaunch, vzBatchLaunch and vzServerSubmit.
vzStatus check_launch_params(const vzLaunchParams* p);
// Rejects spheres with a non-finite centre or a zero / non-finite radius.
vzStatus check_spheres(const vzSphere* s, uint32_t n);

#define VZ_HIP_TRY(expr)                                                              \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return ::vz::set_error(VZ_ERROR_HIP, std::string(#expr) + ": " +          \
                                                     hipGetErrorString(e_));          \
    } while (0)

#define VZ_TRY(expr)                                  \
    do {                                              \
        vzStatus s_ = (expr);                         \
        if (s_ != VZ_SUCCESS) return s_;              \
    } while (0)

struct SceneView {           // what kernels need to traverse one scene
    const uint4* nodes;      // 4-wide nodes, 4 uint4 (64 B) each, root = 0
    const float4* spheres;   // leaf order {c, r}
    const uint32_t* leaf_info;  // leaf -> prim id | material type << 30
    const float4* mats;      // leaf order {albedo rgb, fuzz | ior}
    uint32_t n;              // primitives (leaves)
    uint32_t nw;             // 4-wide nodes (0 for a single sphere)
};

struct Engine;

// A HIP event owned by the enclosing scope, so early error returns do not
// leak it.  ok() is false when creation failed.
struct Event {
    hipEvent_t e = nullptr;
    explicit Event(unsigned flags = hipEventDefault)
    {
        if (hipEventCreateWithFlags(&e, flags) != hipSuccess) e = nullptr;
    }
    ~Event()
    {
        if (e) (void)hipEventDestroy(e);
    }
    Event(const Event&) = delete;
    Event& operator=(const Event&) = delete;
    Event(Event&& o) noexcept : e(o.e) { o.e = nullptr; }
    Event& operator=(Event&& o) noexcept
    {
        std::swap(e, o.e);
        return *this;
    }
    bool ok() const { return e != nullptr; }
};

#define VZ_EVENT_TRY(ev)                                                                  \
    do {                                                                                  \
        if (!(ev).ok()) return ::vz::set_error(VZ_ERROR_HIP, "hipEventCreate failed");    \
    } while (0)

}  // namespace vz

struct vzContext_st {
    int device = 0;
    uint32_t num_cus = 0;
    uint64_t hbm_bytes = 0;
    std::string name;
    bool tuned = false;
    vzTuning tuning{};
    std::mutex mu;
    vz::Engine* engine = nullptr;  // lazily created wavefront engine
    std::atomic<uint32_t> live_scenes{0}, live_servers{0};  // vzDestroy refuses while any remain
    unsigned long long* trace_scratch = nullptr;  // this device's fixed trace scratch (vzTraceClosestNoAlloc)
};

struct vzScene_st {
    vzContext ctx = nullptr;
    uint32_t n = 0;
    // Unique over the process lifetime: tables that cache a scene's device
    // view key it by this id, never by the object's address (a destroyed
    // scene's address can be reused by the next one created).
    uint64_t uid = 0;
    // Server requests (queued or running) that name this scene; while any
    // remain, vzSceneDestroy refuses.
    std::atomic<uint32_t> server_refs{0};
    // inputs (original order)
    float4* d_sph_in = nullptr;
    vzMaterial* d_mat_in = nullptr;
    // build scratch
    uint32_t* d_keys[2] = {nullptr, nullptr};
    uint32_t* d_vals[2] = {nullptr, nullptr};
    uint32_t* d_hist = nullptr;
    float* d_bounds = nullptr;        // partial min/max per block + final 9 floats
    uint32_t hist_blocks = 0;
    // result (leaf order)
    float4* d_sph = nullptr;
    uint32_t* d_leaf_info = nullptr;
    float4* d_mat = nullptr;
    uint32_t* d_morton = nullptr;     // sorted keys (points into d_keys[0])
    uint32_t* d_leaf_prim = nullptr;  // sorted vals (points into d_vals[0])
    uint4* d_nodes = nullptr;
    float* d_boxes = nullptr;
    uint32_t* d_child = nullptr;      // g | flags << 30
    uint32_t* d_parent_int = nullptr;
    uint32_t* d_parent_leaf = nullptr;
    uint32_t* d_visit = nullptr;
    uint32_t* d_height = nullptr;
    uint32_t* d_frontier = nullptr;   // refit: finished children of workgroup-crossing nodes
    uint32_t* d_fcount = nullptr;
    // the 4-wide tree collapsed from the binary one (traversed by every kernel)
    uint4* d_wnodes = nullptr;        // 4 uint4 per wide node, at most n-1 of them
    uint32_t* d_wkids = nullptr;      // 4 binary child references per wide root
    uint32_t* d_wflag = nullptr;      // per binary node: wide-root bit and child count
    uint32_t* d_widx = nullptr;       // per binary node: its wide node index (roots only)
    uint32_t* d_wfront[2] = {nullptr, nullptr};  // level frontiers
    uint32_t* d_wcnt = nullptr;       // per level frontier counts, then nw and the wide depth
    uint32_t nw = 0;                  // wide nodes
    uint32_t wdepth = 0;              // wide levels from the root to the deepest leaf
    uint32_t builder = 0;             // VZ_BUILD_LBVH or VZ_BUILD_SAH
    uint32_t depth = 0;
    uint64_t device_bytes = 0;
    double build_ms = 0.0;
    vz::SceneView view() const
    {
        vz::SceneView v;
        v.nodes = d_wnodes;
        v.spheres = d_sph;
        v.leaf_info = d_leaf_info;
        v.mats = d_mat;
        v.n = n;
        v.nw = nw;
        return v;
    }
};

namespace vz {
// vz_lbvh.hip
vzStatus lbvh_alloc(vzScene s);
vzStatus lbvh_build(vzScene s, hipStream_t st);   // full build from d_sph_in
vzStatus lbvh_refit(vzScene s, hipStream_t st);   // same topology, new d_sph_in
// Checks d_sph_in (and `mat`, device memory, when not NULL) on the GPU;
// returns the first bad sphere / material index, 0xFFFFFFFF for none.
// Synchronises the stream.
vzStatus lbvh_check_input(vzScene s, const vzMaterial* mat, hipStream_t st, uint32_t* bad_sphere,
                          uint32_t* bad_material);
void lbvh_free(vzScene s);

// vz_trace.hip
// Workgroups of a trace of n rays (each writes kTraceParts partial counters).
uint32_t trace_blocks(uint32_t n);
// Closest hits of n rays; parts holds trace_blocks(n) * kTraceParts words,
// counters 4 words (segments, 0, node visits, sphere tests), both device
// memory written (not accumulated) by the call.
vzStatus trace_rays(const SceneView& sc, const vzRay* rays, uint32_t n, vzHit* hits,
                    hipStream_t st, bool count, unsigned long long* parts, unsigned long long* counters);
// Device address of the current device's fixed trace scratch:
// kTraceScratchBlocks * kTraceParts partial words, then 8 counter words.
vzStatus trace_scratch_address(unsigned long long** out);

// vz_api.cpp: the process-wide autotune cache, keyed by device name and CU count
void tuning_cache_store(vzContext ctx, const vzTuning& t);

// vz_megakernel.hip
vzStatus megakernel_launch(const SceneView& sc, const vzLaunchParams& p, hipStream_t st,
                           unsigned long long* d_counters, bool count);
}  // namespace vz
