// vz_engine.h -- the wavefront integrator shared by the plain launch
// (vzLaunch), the batched form (vzBatchLaunch) and the serving loop.
//
// Layout.  The ray queue (two parities) is split into NB block-private
// slices of R rays, one per step-kernel workgroup: as many as fit on the
// device at once (plain launches), or 4x that many smaller ones, which the
// dispatcher hands out as workgroups finish (batches and servers, whose
// slices differ in cost).  A workgroup only ever touches its own slice, so
// claiming work and compacting survivors use LDS atomics and wave ballots,
// never a device-wide atomic (on MI355X a device-scope atomic to one address
// serialises across the 8 XCDs).  Each wavefront iteration runs three launches
// on one stream (the last with two stages on disjoint workgroups):
//   step    one slice per workgroup: waves pull rays of their slice, trace them
//           (LBVH traversal), run the miss stage (throughput * sky) or the
//           shade stage (material scatter) in registers, and compact the
//           surviving paths into the slice of the next parity
//   control one workgroup: finds passes whose paths all ended, plans their
//           accumulation, grants free queue space to requests and splits
//           the granted camera rays over the slices' free room
//   finish  apply: adds finished passes into per-pixel sums in sample order
//           and writes accum/image when a request's last pass lands;
//           regen: writes the granted camera rays into each slice's free room
// A "pass" is sample s of every pixel of one request; its radiances land in
// one of K ring slots, so per-pixel sums are formed in sample order whatever
// order the paths finish in.  Iterations carry no host round trip; the host
// polls a status copy between blocks of iterations (or hipGraphs of them).
#pragma once
#include <vector>

#include "vz_device.h"
#include "vz_internal.h"

namespace vz {

constexpr uint32_t kMaxBlocks = 8192;   // step-kernel workgroups (queue slices)
constexpr uint32_t kMaxSplit = 8;       // queue slices per resident step workgroup, at most
// Batches and servers split the queue 4 ways per resident step workgroup: on
// the 64-scene batch (12M queue, 8x8 tiles) 3833 Mray-segments/s against 3422
// for 2 and 2712 for 1; plain launches of one scene keep 1 (1M box scene 2434
// against 2259 at 4).
constexpr uint32_t kBatchSplit = 4;
// Servers split 4 ways: on the 2000-request stream's mix (5% 512x512 2048-spp
// FP64 requests among 256-1024 px ones), all 600 first requests in flight at
// once, with small requests visiting their tiles strided, a 48M-ray queue
// served 5648 Mray-segments/s at split 4 against 5518 at split 8; 32M: 5576
// at 4, 5424 at 8, 5367 at 2; 24M: 5303 at 4, 5087 at 8.
constexpr uint32_t kServeSplit = 4;
constexpr uint32_t kDeadRay = 0xFFFFFFFFu;  // meta of a padding item: skipped
// Scene table entries, allocated once when the engine is created, so that the
// table's address never changes: captured hipGraphs bake it into their step
// launches.  A batch names at most kMaxRequests distinct scenes; a server
// holds at most max_active referenced ones and caches the rest until their
// entries are needed.
constexpr uint32_t kMaxScenes = 4096;

struct ReqDesc {             // device table entry, written by the host
    Cam cam;
    uint32_t scene;          // index into the scene table
    uint32_t spp, subframe, max_depth, flags;
    uint32_t npix, tw, th, tiles_x, npad;
    uint32_t K;              // pass ring slots
    double rd_tsz, rd_tw, rd_tx, rd_K;  // 1 / (tw * th), 1 / tw, 1 / tiles_x, 1 / K (udiv_r)
    uint32_t ntiles, tstride;  // tiles per pass; the stride of the tile visiting order
    double rd_ntiles;
    uint32_t seq;            // arrival number: free queue room goes to the lowest first
    unsigned long long total_items;  // spp * npad
    float* passbuf;          // K * 3 * npix, planar per slot
    float* sum32;            // 3 * npix
    double* sum64;           // 3 * npix (FP64 accumulation)
    uint32_t* camseed;       // npix
    uint32_t* done;          // K pass counters (paths ended)
    unsigned long long* segsum;  // K pass segment sums
    void* accum;
    uint32_t* image;
};

struct ReqState {            // device, advanced by the control kernel
    unsigned long long next_item;
    uint32_t acc_next;
    uint32_t active;
    uint32_t finish_iter;
    uint32_t pad;
    unsigned long long segments;  // segments of the passes accumulated so far
};

struct Counters {            // written by the control kernel, except fill (host, per run)
    uint32_t iteration;
    uint32_t active;         // requests still running
    uint32_t n_slots;        // request slots in use (table prefix)
    uint32_t total_regen;
    uint32_t fill;           // rays a slice is filled to (0: its capacity R)
};

struct BlockStats {          // one per queue slice: its step workgroup writes hits..segments, the regen workgroups add paths atomically
    unsigned long long hits, node_visits, sphere_tests, paths, segments;
};

struct Plan {
    uint32_t n_acc, total_acc_pix, n_regen, total_regen;
    uint32_t acc_req[kMaxRequests];
    uint32_t acc_first[kMaxRequests];
    uint32_t acc_cnt[kMaxRequests];
    uint32_t acc_prefix[kMaxRequests];
    uint32_t regen_req[kMaxRequests];
    uint32_t regen_prefix[kMaxRequests];   // first granted index of each regen entry
    uint32_t regen_cnt[kMaxRequests];      // items granted to it (may span several passes)
    uint32_t run_prefix[kMaxRequests];     // first pixel run of it: min(cnt, npad) runs each
    unsigned long long regen_base[kMaxRequests];  // its first item (pass * npad + padded pixel)
    uint32_t regen_s0[kMaxRequests];       // that item's pass
    uint32_t regen_i0[kMaxRequests];       // and padded pixel
    uint32_t total_runs;
    uint32_t blk_first[kMaxBlocks];   // first granted index of each slice
    uint32_t blk_count[kMaxBlocks];   // granted items for each slice
    uint32_t blk_pos0[kMaxBlocks];    // the slice's ray count before regen
};

struct EngineConfig {
    uint32_t queue_capacity = 1u << 23;  // rays in flight (8M)
    uint32_t tile_w = 32, tile_h = 2;   // one wave of camera rays = 32x2 pixels
    uint32_t iters_per_graph = 16;
    uint32_t refill_lanes = 0;          // 0: per-variant default (kernel side)
};

struct DeviceBlock {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct Engine {
    vzContext ctx = nullptr;
    EngineConfig cfg;
    uint32_t Q = 0;                        // rays in flight (NB * R)
    uint32_t NB = 0, R = 0;                // slices and slice capacity of the current partition
    uint32_t NB_res = 0, R_base = 0;       // one slice per resident step workgroup (split 1)
    uint32_t split = 1;                    // current partition: NB = NB_res * split
    float4* q_o[2] = {nullptr, nullptr};   // origin xyz, pixel
    float4* q_d[2] = {nullptr, nullptr};   // dir xyz, seed
    float4* q_t[2] = {nullptr, nullptr};   // throughput xyz, req<<16 | slot<<8 | depth
    uint32_t* d_bcnt = nullptr;            // 2 * NB slice counts
    BlockStats* d_bstats = nullptr;        // NB partial stats
    ReqDesc* d_req = nullptr;
    ReqState* d_state = nullptr;
    SceneView* d_scenes = nullptr;         // kMaxScenes entries, never reallocated
    Counters* d_cnt = nullptr;
    Plan* d_plan = nullptr;
    Counters* h_cnt = nullptr;             // pinned copies for polling (2)
    ReqState* h_state = nullptr;           // pinned copy of the request states
    ReqDesc* h_req = nullptr;              // pinned staging for the table
    SceneView* h_scenes = nullptr;
    BlockStats* h_bstats = nullptr;
    uint32_t grid_small = 0;
    hipGraph_t graph_tmpl = nullptr;       // the captured graph, kept as long as its instance
    hipGraphExec_t graph = nullptr;
    hipStream_t graph_stream = nullptr;
    int graph_key = -1;
    // the control kernel runs 256 threads instead of 1024: at most 256
    // request slots and 1024 queue slices (plain launches, small batches)
    bool ctl_small = false;
    const SceneView* graph_scenes = nullptr;  // scene table the captured launches read
    // per-slot buffers, reused when large enough
    std::vector<DeviceBlock> slot_mem;
    uint64_t slot_bytes_total = 0;
    // serving: stream-ordered per-request buffers, slot occupancy, scene table
    bool stream_alloc = false;
    std::vector<uint8_t> slot_used;
    uint32_t n_slots = 0;
    uint32_t next_seq = 0;                 // arrival numbers of served requests
    ReqState* h_state_stage = nullptr;     // pinned staging for state uploads
    ReqState* h_status = nullptr;          // pinned, 2 x kMaxRequests status copies
    uint32_t* h_nslots = nullptr;          // pinned staging for the slot count
    // serving scene table: entry i caches the device view of scene uid[i];
    // refs counts the admitted requests that use it.  Entries with no refs
    // stay cached for later requests of the same scene and are reused,
    // least recently used first, when a new scene needs one.
    std::vector<uint64_t> scene_uid;       // 0: never used
    std::vector<uint32_t> scene_refs;
    std::vector<uint64_t> scene_last_use;
    std::vector<uint32_t> slot_scene;      // request slot -> scene entry
    uint64_t scene_clock = 0;

    ~Engine();
};

// Creates an engine for ctx with the given configuration.
vzStatus engine_create(vzContext ctx, const EngineConfig& cfg, Engine** out);
void engine_destroy(Engine* e);

struct RunItem {
    const vzScene_st* scene;
    vzLaunchParams params;
};

struct RunOptions {
    bool use_graph = false;
    bool count_visits = false;
    bool time_trace = false;
    uint64_t rays_in_flight = 0;  // cap on queued rays for this run (0: the whole queue)
    uint32_t slice_split = 1;     // queue slices per resident step workgroup (1, 2 or 4)
};

struct RunResult {
    unsigned long long segments = 0, paths = 0, node_visits = 0, sphere_tests = 0, hits = 0;
    double render_ms = 0.0, trace_ms = 0.0;
    uint32_t trace_launches = 0, iterations = 0;
    std::vector<unsigned long long> req_segments;
    std::vector<uint32_t> req_finish_iter;
};

// Renders every item to completion (synchronises the stream).
vzStatus engine_run(Engine* e, const std::vector<RunItem>& items, hipStream_t st,
                    const RunOptions& opt, RunResult* res);

// ---- incremental interface for the serving loop -------------------------
// Prepares the engine for incremental use (empty tables, stream-ordered
// request buffers).
vzStatus engine_serve_init(Engine* e, hipStream_t st, uint32_t slice_split);
// Installs one request into a free slot (written to the device tables in
// stream order, so it joins the wavefront at the next iteration).
vzStatus engine_admit(Engine* e, const RunItem& item, hipStream_t st, uint32_t* slot);
// Enqueues one block of wavefront iterations (the captured hipGraph) and a
// copy of the request states into pinned memory (buffer `k & 1`).
vzStatus engine_step(Engine* e, hipStream_t st, uint32_t k);
// Reads slot state from status buffer `k & 1` (valid once the stream passed
// the engine_step that wrote it).
void engine_slot_status(Engine* e, uint32_t k, uint32_t slot, bool* active,
                        unsigned long long* segments);
// Frees a finished slot's buffers (stream-ordered), drops its reference on
// its scene entry and marks the slot reusable.
vzStatus engine_release_slot(Engine* e, uint32_t slot, hipStream_t st);
// Scene entries currently referenced by admitted requests (for tests).
uint32_t engine_scenes_referenced(const Engine* e);
size_t request_bytes(const vzLaunchParams& p);

Cam make_cam(const vzLaunchParams& p);

}  // namespace vz
