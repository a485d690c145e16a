/*
 * vz_api.h -- C ABI of the MI355X-native sphere path tracer (libvz.so).
 *
 * This is the drop-in boundary for the OptiX launch of TienYoung/Venusaur.
 * Plain C: POD structs, plain pointers and sizes, status codes.  Streams are
 * passed as `vzStream` (a hipStream_t cast to void*; NULL = the null stream).
 * Every call that enqueues GPU work is stream-ordered; calls that return data
 * to the host synchronise the stream they were given.
 *
 * Reference interfaces replaced (file:line in the reference tree):
 *   Renderer::Init / CreateContext ........ Core/Renderer.h:25-33, 144-158
 *   Renderer::BuildAccelerationStructures . Core/Renderer.h:160-255 (optixAccelBuild)
 *   Renderer::CreateSBT ................... Core/Renderer.h:452-520 (per-sphere records)
 *   Renderer::Draw ........................ Core/Renderer.h:35-78   (optixLaunch, line 75)
 *   Renderer::Cleanup ..................... Core/Renderer.h:80-97
 *   Params / SphereHitGroupData / MaterialData  Core/RayTracer.h:3-45
 *   optixTrace (closest hit over the GAS) . Core/RayTracer.cu:190-202
 *   Scene::Scene (book scene) ............. Core/Scene.h:13-112
 *   Camera::UVWFrame ...................... Core/Camera.cpp:16-37
 */
#ifndef VZ_API_H
#define VZ_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VZ_API_VERSION 1

typedef enum vzStatus {
    VZ_SUCCESS = 0,
    VZ_ERROR_INVALID_VALUE = 1,  /* bad argument (null pointer, size, range) */
    VZ_ERROR_OUT_OF_MEMORY = 2,  /* device or host allocation failed */
    VZ_ERROR_HIP = 3,            /* a HIP runtime call failed */
    VZ_ERROR_NOT_READY = 4,      /* serving: result not finished yet */
    VZ_ERROR_CAPACITY = 5,       /* request does not fit (HBM budget, queue, table) */
    VZ_ERROR_UNSUPPORTED = 6,    /* feature not available (e.g. GL interop) */
    VZ_ERROR_INTERNAL = 7,       /* internal invariant violated */
    VZ_ERROR_NO_DEVICE = 8       /* no usable gfx950 device */
} vzStatus;

typedef void* vzStream;                    /* hipStream_t */
typedef struct vzContext_st* vzContext;    /* device context (Renderer State, Renderer.h:111-132) */
typedef struct vzScene_st* vzScene;        /* resident scene + BVH (the GAS + SBT) */
typedef struct vzServer_st* vzServer;      /* serving loop */

enum { VZ_LAMBERTIAN = 0, VZ_METAL = 1, VZ_DIELECTRIC = 2 }; /* material.h:9-14 */

/* SphereHitGroupData {center, radius} (RayTracer.h:40-45) */
typedef struct vzSphere { float center[3]; float radius; } vzSphere;
/* Material (material.h:16-32) / MaterialData (RayTracer.h:27-38), unpacked */
typedef struct vzMaterial { uint32_t type; float albedo[3]; float fuzz; float ir; } vzMaterial;
/* One closest-hit query, the arguments of optixTrace (RayTracer.cu:190-202) */
typedef struct vzRay { float origin[3]; float tmin; float dir[3]; float tmax; } vzRay;
/* Closest hit: t of the hit (tmax on a miss) and the primitive id
 * (index into the sphere array given at scene creation), VZ_MISS on a miss. */
typedef struct vzHit { float t; uint32_t prim; } vzHit;
#define VZ_MISS 0xFFFFFFFFu

/* Launch flags */
#define VZ_ACCUM_FP64 0x1u     /* accum is double4 (required above 1024 spp) */
#define VZ_MEGAKERNEL 0x2u     /* one-thread-per-pixel megakernel instead of the wavefront */
#define VZ_COUNT_VISITS 0x4u   /* count BVH node fetches and sphere tests into vzStats */
#define VZ_TIME_STAGES 0x8u    /* time every trace-kernel launch with HIP events */

/* Params (RayTracer.h:3-17).  image/accum are device pointers.  The
 * traversable handle is replaced by the vzScene argument of vzLaunch. */
typedef struct vzLaunchParams {
    void* image;                 /* uchar4[width*height] or NULL */
    void* accum;                 /* float4[width*height] (double4 with VZ_ACCUM_FP64), required */
    uint32_t width, height;      /* >= 2 each (RayTracer.cu:173-174 divides by w-1, h-1) */
    uint32_t samples_per_pixel;  /* >= 1 */
    uint32_t subframe_index;     /* frame number since reset, >= 1; seeds tea<4>(pixel, frame) */
    float origin[3], u[3], v[3], w[3];
    float lens_radius;
    uint32_t max_depth;          /* segments per path, 1..255 (the reference hard-codes 4) */
    uint32_t flags;              /* VZ_* flags above */
} vzLaunchParams;

typedef struct vzStats {
    uint64_t segments;       /* closest-hit queries traced (the metric's unit) */
    uint64_t paths;          /* camera paths started */
    uint64_t node_visits;    /* BVH node fetches (VZ_COUNT_VISITS only) */
    uint64_t sphere_tests;   /* leaf sphere tests (VZ_COUNT_VISITS only) */
    double render_ms;        /* wall time of the launch (HIP events) */
    double trace_ms;         /* sum of trace-kernel times (VZ_TIME_STAGES only) */
    uint32_t trace_launches; /* trace-kernel launches (timed with VZ_TIME_STAGES) */
    uint32_t iterations;     /* wavefront iterations */
    uint64_t hits;           /* segments that hit a sphere (the rest missed) */
} vzStats;

typedef struct vzAccelInfo {
    uint32_t num_prims;
    uint32_t num_nodes;       /* binary internal nodes, num_prims - 1 */
    uint32_t depth;           /* binary levels from the root to the deepest leaf */
    uint32_t node_bytes;      /* bytes per binary node (32) */
    uint64_t device_bytes;    /* resident scene + BVH footprint */
    double build_ms;          /* last build/refit time (HIP events) */
    uint32_t num_wide_nodes;  /* nodes of the 4-wide tree traversal uses (64 bytes each) */
    uint32_t wide_depth;      /* its levels from the root to the deepest leaf */
    uint32_t builder;         /* VZ_BUILD_LBVH, or VZ_BUILD_SAH for 2 to 1024 spheres */
} vzAccelInfo;

/* Which build made the binary tree.  Scenes of 2 to 1024 spheres get a
 * binned-SAH tree built top down by one workgroup (leaves in split order,
 * the morton array then holds each leaf's key unsorted); larger ones an
 * LBVH (Morton-sorted leaves, Karras linking).  Both use one node layout and
 * give the same closest hits. */
enum { VZ_BUILD_LBVH = 0, VZ_BUILD_SAH = 1 };

/* ---- context -------------------------------------------------------- */
vzStatus vzCreate(int device, vzContext* out);
/* Fails with VZ_ERROR_INVALID_VALUE while scenes or servers of the context
 * are alive: destroy them first. */
vzStatus vzDestroy(vzContext ctx);
/* Device properties used by the autotuner/admission: CU count, HBM bytes. */
vzStatus vzDeviceInfo(vzContext ctx, uint32_t* num_cus, uint64_t* hbm_bytes, char* name,
                      size_t name_len);

/* ---- scene and acceleration structure ------------------------------ */
/* Copies n spheres/materials from host memory, uploads them and builds the
 * BVH on the GPU (replaces Renderer.h:160-255 and 452-520). */
vzStatus vzSceneCreate(vzContext ctx, const vzSphere* spheres, const vzMaterial* materials,
                       uint32_t n, vzStream stream, vzScene* out);
/* As vzSceneCreate, with spheres and materials already in device memory
 * (copied device to device and validated on the GPU; synchronises the
 * stream).  For callers that generate or keep scenes in HBM. */
vzStatus vzSceneCreateDevice(vzContext ctx, const vzSphere* spheres, const vzMaterial* materials,
                             uint32_t n, vzStream stream, vzScene* out);
/* Fails with VZ_ERROR_INVALID_VALUE while requests that name the scene are
 * queued or running on a server (vzServerSubmit); the scene is unchanged. */
vzStatus vzSceneDestroy(vzScene scene);
/* Rebuild the BVH from the current spheres. */
vzStatus vzAccelBuild(vzScene scene, vzStream stream);
/* Replace sphere geometry (same count, same materials) and refit the
 * existing topology: boxes and node quantisation are recomputed bottom-up. */
vzStatus vzSceneUpdateSpheres(vzScene scene, const vzSphere* spheres, vzStream stream);
/* As vzSceneUpdateSpheres, with the spheres in device memory (validated on
 * the GPU; a rejected update leaves the scene unchanged). */
vzStatus vzSceneUpdateSpheresDevice(vzScene scene, const vzSphere* spheres, vzStream stream);
vzStatus vzAccelRefit(vzScene scene, vzStream stream);
vzStatus vzAccelInfoGet(vzScene scene, vzAccelInfo* out);
/* Copies the BVH to host memory for inspection (synchronises the stream).
 * nodes: num_nodes * 32 bytes; leaf_prims, morton: num_prims words each.
 * Any pointer may be NULL. */
vzStatus vzAccelDownload(vzScene scene, void* nodes, uint32_t* leaf_prims, uint32_t* morton,
                         vzStream stream);
/* Copies the 4-wide tree that every traversal walks (num_wide_nodes * 64
 * bytes; synchronises the stream).  It is collapsed from the binary tree on
 * every build and refit.  Node layout, 16 words: the node box's lo corner
 * (3 floats); meta = per-axis scale exponents (bytes 0-2), the slots' leaf
 * flags (bits 24-27) and valid flags (bits 28-31); the quantised lo planes
 * of the 4 slots for x, y, z (byte c = slot c), then the hi planes; the 4
 * children (leaf index or wide node index); 2 zero words. */
vzStatus vzAccelDownloadWide(vzScene scene, void* wide_nodes, vzStream stream);

/* ---- tracing and rendering ----------------------------------------- */
/* Closest hit for n rays (device arrays).  Equivalent of one optixTrace per
 * ray with the given [tmin, tmax]; ties resolve to the smaller t, then the
 * smaller primitive id. */
vzStatus vzTraceClosest(vzScene scene, const vzRay* rays, uint32_t n, vzHit* hits,
                        vzStream stream, vzStats* stats);
/* As vzTraceClosest, allocating nothing: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under hipGraph
 * capture, serving loops and harnesses that forbid hipMalloc).  Its
 * per-workgroup counters live in a fixed 1.5 MiB array in the library's
 * device memory (one per GPU, held until the process ends), so it traces at
 * most 16,777,216 rays per call (VZ_ERROR_CAPACITY above that).  Calls that
 * overlap on one GPU share that array and race: order them on one stream.
 * With stats == NULL it only enqueues (capturable); with stats it copies the
 * counters back and synchronises the stream, and leaves the times 0.  The
 * current device must be the scene's. */
vzStatus vzTraceClosestNoAlloc(vzScene scene, const vzRay* rays, uint32_t n, vzHit* hits,
                               vzStream stream, vzStats* stats);
/* One frame: the OptiX launch of Renderer::Draw (Renderer.h:75) plus the
 * accumulation of __raygen__rg (RayTracer.cu:163-217).  Synchronises the
 * stream when stats != NULL, otherwise returns after enqueueing. */
vzStatus vzLaunch(vzScene scene, const vzLaunchParams* params, vzStream stream, vzStats* stats);

/* ---- batched form --------------------------------------------------- */
typedef struct vzBatchItem {
    vzScene scene;
    vzLaunchParams params;
} vzBatchItem;

typedef struct vzTuning {
    uint32_t queue_capacity;  /* rays in flight (0 = default) */
    uint32_t tile_w, tile_h;  /* camera-ray tile: tile_w*tile_h pixels per block of items */
    uint32_t iters_per_graph; /* wavefront iterations captured per hipGraph (even) */
    uint32_t refill_lanes;    /* a wave fetches new rays once fewer than this many of its lanes
                                 still trace: 1 = only when all are done, 65 = after every
                                 node step, 0 = the default for the kernel variant: 9 for
                                 one small scene with its nodes in LDS, 41 for one scene with
                                 its nodes in memory, 33 for several scenes in one run */
    uint32_t slice_split;     /* queue slices per resident step workgroup: 1, 2, 4 or 8 (0 = 4 for
                                 batches and servers).  Slices of uneven cost (many scenes
                                 in one run) balance better when split; vzLaunch always uses 1 */
} vzTuning;

/* Renders every item in one wavefront run (its iterations replayed as
 * hipGraphs); the items share the ray queue.  Synchronises the stream. */
vzStatus vzBatchLaunch(vzContext ctx, const vzBatchItem* items, uint32_t n,
                       const vzTuning* tuning, vzStream stream, vzStats* stats);
/* Renders the batch twice per candidate tuning, timing the wavefront only,
 * and returns the candidate with the fastest run.  Stage 1: queue capacity
 * 4M/8M/12M/24M/32M/48M rays x camera tiles 32x2/16x4/8x8 (default refill, queue
 * split 4); stage 2, at the best of those: refill_lanes 25/33/49/57, then
 * slice_split 1/2/8.  The choice becomes this context's default and is cached
 * for the process per kind of device (name and CU count): contexts on the
 * same kind of device that have not tuned themselves start from it.  The
 * items' accumulators are written (as by vzBatchLaunch). */
vzStatus vzAutotune(vzContext ctx, const vzBatchItem* items, uint32_t n, vzStream stream,
                    vzTuning* out);
vzStatus vzDefaultTuning(vzContext ctx, vzTuning* out);

/* ---- serving form --------------------------------------------------- */
typedef struct vzServeConfig {
    uint64_t hbm_budget_bytes;   /* admission budget: the per-request device buffers the server
                                    allocates (vzRequestFootprint: pass ring, per-pixel sums,
                                    camera seeds).  Scenes are the caller's and already resident;
                                    they are not counted (vzAccelInfo.device_bytes gives theirs) */
    uint32_t max_active;         /* max requests in the device table (<= 1024) */
    vzTuning tuning;             /* 0 fields take the serving defaults: 48M-ray queue, 8x8 tiles,
                                    queue split 4 */
} vzServeConfig;

typedef struct vzRequestResult {
    uint64_t request_id;
    double submit_ms, admit_ms, finish_ms;  /* host clock, ms since server start */
    uint64_t segments;
    uint32_t status;                        /* VZ_SUCCESS or an error */
} vzRequestResult;

vzStatus vzServerCreate(vzContext ctx, const vzServeConfig* cfg, vzServer* out);
/* Enqueue a request; the accum/image buffers in params must stay valid until
 * the request completes.  Returns immediately with its id.  The scene is held
 * until the request has finished: vzSceneDestroy refuses it meanwhile. */
vzStatus vzServerSubmit(vzServer srv, vzScene scene, const vzLaunchParams* params,
                        uint64_t* request_id);
/* Result of a finished request (VZ_ERROR_NOT_READY if still in flight).
 * Results stay available, to any number of polls, until vzServerDestroy
 * (about 48 bytes per request). */
vzStatus vzServerPoll(vzServer srv, uint64_t request_id, vzRequestResult* out);
/* Block until every submitted request finished. */
vzStatus vzServerDrain(vzServer srv);
/* Finishes every submitted request, then stops the loop and frees the
 * server's engine and stream. */
vzStatus vzServerDestroy(vzServer srv);
/* Bytes of device memory a request of these params needs (admission). */
uint64_t vzRequestFootprint(const vzLaunchParams* params);

/* ---- wire format (input of the serving form) ------------------------
 * The reference has no scene file: its scene is code (Core/Scene.h:13-80)
 * and its request is the Params a Draw fills in (Core/RayTracer.h:3-17,
 * Core/Renderer.h:37-75).  These records carry the same content between
 * processes.  Little-endian, 4-byte aligned:
 *
 *   scene record (24 + 40 n bytes)
 *     0  "VZSC"   4  u32 version (1)   8  u32 n   12  u32 0
 *     16 u64 FNV-1a 64 of the payload
 *     24 n x vzSphere (16 B), then n x vzMaterial (24 B)
 *   request record (96 bytes, then scene_bytes of an inline scene record)
 *     0  "VZRQ"   4  u32 version (1)   8  u64 tag   16  u64 scene_key
 *     24 u32 width, height, samples_per_pixel, max_depth
 *     40 u32 flags (VZ_ACCUM_FP64 or 0)   44 u32 subframe_index
 *     48 f32 lookfrom[3]   60 f32 lookat[3]
 *     72 f32 vfov_deg, aperture, focus_dist, arrival_ms
 *     88 u32 scene_bytes (0: the key names a scene already resident)   92 u32 0
 *
 * Decoders check magic, version, sizes, the checksum and every value a
 * launch would reject, and never read past `size`. */
#define VZ_WIRE_SCENE_HEADER 24u
#define VZ_WIRE_REQUEST_HEADER 96u

typedef struct vzRequestDesc {
    uint64_t tag;                 /* caller's id, echoed back */
    uint64_t scene_key;           /* names the scene; requests may share one */
    uint32_t width, height, samples_per_pixel, max_depth;
    uint32_t flags;               /* 0 or VZ_ACCUM_FP64 */
    uint32_t subframe_index;      /* >= 1 */
    float lookfrom[3], lookat[3]; /* camera (Camera.cpp:24-37), up = +y */
    float vfov_deg, aperture, focus_dist;
    float arrival_ms;             /* offset from the start of the stream (replay) */
} vzRequestDesc;

/* Bytes of a scene record of n spheres. */
size_t vzWireSceneSize(uint32_t n);
/* Writes a scene record into buf (cap bytes); *written = its size. */
vzStatus vzWireEncodeScene(const vzSphere* spheres, const vzMaterial* materials, uint32_t n, void* buf,
                           size_t cap, size_t* written);
/* Reads the scene record at the start of buf.  *n = its sphere count; with
 * spheres and materials non-NULL (room for cap each) also copies them out.
 * *consumed = the record's size. */
vzStatus vzWireDecodeScene(const void* buf, size_t size, vzSphere* spheres, vzMaterial* materials,
                           uint32_t cap, uint32_t* n, size_t* consumed);
/* Writes a request record, followed by a scene record of n spheres when
 * n > 0 (n = 0: the request names a resident scene by its key). */
vzStatus vzWireEncodeRequest(const vzRequestDesc* req, const vzSphere* spheres, const vzMaterial* materials,
                             uint32_t n, void* buf, size_t cap, size_t* written);
/* Reads the request record at the start of buf.  *scene_bytes = the size of
 * its inline scene record (0 if none), which starts at buf + 96; *consumed =
 * 96 + *scene_bytes. */
vzStatus vzWireDecodeRequest(const void* buf, size_t size, vzRequestDesc* out, size_t* scene_bytes,
                             size_t* consumed);
/* Launch params of a request (camera by vzCameraLookAt, aspect = width /
 * height in float) for the given device buffers. */
vzStatus vzRequestLaunchParams(const vzRequestDesc* req, void* image, void* accum, vzLaunchParams* out);

/* ---- host helpers (scene generators and camera, Scene.h / Camera.cpp) -- */
/* Book final scene (Scene.h:13-80) from std::mt19937(seed), draws in
 * x -> y -> z order.  Writes at most cap spheres; *n_out = count. */
vzStatus vzGenerateBookScene(uint32_t seed, vzSphere* spheres, vzMaterial* materials,
                             uint32_t cap, uint32_t* n_out);
/* Ground sphere plus n_small random spheres in [-100,100]x[0,20]x[-100,100],
 * radii log-uniform in [0.01, 0.5].  Writes n_small + 1 spheres. */
vzStatus vzGenerateBoxScene(uint32_t n_small, uint32_t seed, vzSphere* spheres,
                            vzMaterial* materials);
/* Camera frame (Camera.cpp:24-37) into params->origin/u/v/w/lens_radius. */
vzStatus vzCameraLookAt(const float lookfrom[3], const float lookat[3], float vfov_deg,
                        float aspect, float aperture, float focus_dist, vzLaunchParams* params);

/* ---- errors ----------------------------------------------------------- */
const char* vzStatusString(vzStatus s);
/* Message of the last failing call on this thread. */
const char* vzGetLastError(void);

#ifdef __cplusplus
}
#endif
#endif
