/*
 * vz_oracle.h -- CPU oracle for the sphere path tracer (TEST INFRASTRUCTURE).
 *
 * This library is the checker, never the product: only tests/, the smoke()
 * entry point and bench.py's cpu_baseline leg may load it.  It restates, in
 * plain C with -ffp-contract=off, the algorithm of the reference's OptiX
 * launch:
 *
 *   RNG ........... Core/random.cuh:31-67 (tea<4>, lcg, rnd)
 *   camera ray .... Core/RayTracer.cu:141-177, Core/Camera.cpp:24-37
 *   sphere test ... Core/RayTracer.cu:229-270 (with the FMA-compensated
 *                   discriminant and stable root form the build mandates)
 *   materials ..... Core/RayTracer.cu:272-440, Core/vec_math.h:545-570
 *   miss .......... Core/RayTracer.cu:442-450
 *   accumulate .... Core/RayTracer.cu:203-216 and make_color 16-47
 *   scene ......... Core/Scene.h:13-112 (explicit x->y->z draw order)
 *
 * and the build's own BVH (for more than 1024 spheres an LBVH: Morton codes,
 * stable sort, Karras 2012 linking; for 2 to 1024 a binned-SAH tree in the
 * same node layout; 32-byte quantised AABB-pair nodes, then the 4-wide tree
 * collapsed from them) plus its short-stack traversal, which replace
 * OptiX's closed acceleration structure.  The GPU library must match
 * this code bit for bit (BVH arrays, hit records, segment counts) and within
 * RMSE 1e-3 on images.  An FP64 variant of the integrator is the accuracy
 * oracle.
 *
 * Parity pinning (what checks each function, see tests/test_oracle.py):
 *   tea4/lcg/rnd ........ KATs compiled from the reference header itself
 *                         (oracle/_ref) and the committed KAT table
 *                         (tests/golden/rng_kat.json)
 *   mt19937 ............. C++ standard check value (10000th output)
 *   book scene .......... committed sphere-list fixture + structural checks
 *   camera .............. closed forms (centre ray = W, corners W+-U/2+-V/2)
 *   sphere test ......... closed forms and the FP64 test
 *   BVH ................. invariants + brute-force closest hit equality
 *   materials ........... invariants (Snell, |reflect|, Schlick r0, TIR)
 *   integrator .......... known answers (empty scene = sky, depth 1) and the
 *                         FP64 integrator
 *   box scene (1M) ...... parity unpinned against the reference (the
 *                         reference has no such generator); structural checks
 */
#ifndef VZ_ORACLE_H
#define VZ_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float cx, cy, cz, r; } vzo_sphere;                     /* 16 B */
typedef struct { uint32_t type; float albedo[3]; float fuzz; float ir; } vzo_material; /* 24 B */
typedef struct { float ox, oy, oz, tmin, dx, dy, dz, tmax; } vzo_ray;    /* 32 B */
typedef struct { float t; uint32_t prim; } vzo_hit;                      /* 8 B, prim=~0 on miss */
typedef struct { float o[3]; uint32_t meta; uint32_t q[3]; uint32_t child; } vzo_node; /* 32 B */
/* 4-wide node (64 B): the node box's lo corner, per-axis scale exponents
 * (meta bytes 0-2), slot leaf flags (meta bits 24-27) and valid flags (bits
 * 28-31); per axis the four slots' quantised lo planes (byte c = slot c) and
 * hi planes; per slot the child: a leaf index (Morton order) or a wide node
 * index. */
typedef struct { float o[3]; uint32_t meta; uint32_t qlo[3]; uint32_t qhi[3]; uint32_t child[4]; uint32_t pad[2]; } vzo_wnode;
typedef struct { float origin[3], u[3], v[3], w[3]; float lens_radius; } vzo_camera;

enum { VZO_LAMBERTIAN = 0, VZO_METAL = 1, VZO_DIELECTRIC = 2 };
#define VZO_MISS 0xFFFFFFFFu

/* RNG (random.cuh:31-67) */
uint32_t vzo_tea4(uint32_t v0, uint32_t v1);
uint32_t vzo_lcg(uint32_t* state);
float vzo_rnd(uint32_t* state);

/* std::mt19937 restated; uniform float u = (x >> 8) * 2^-24 */
uint32_t vzo_mt19937_nth(uint32_t seed, uint32_t n);  /* n-th output, 1-based */

/* Scene generators.  book: returns the sphere count (<= cap).  box: writes
 * n_small + 1 spheres (ground first). */
uint32_t vzo_book_scene(uint32_t seed, vzo_sphere* s, vzo_material* m, uint32_t cap);
void vzo_box_scene(uint32_t n_small, uint32_t seed, vzo_sphere* s, vzo_material* m);

/* Camera frame (Camera.cpp:24-37 with forward = normalize(at - from)). */
void vzo_camera_lookat(const float from[3], const float at[3], float vfov_deg, float aspect,
                       float aperture, float focus, vzo_camera* out);

/* Single sphere tests.  Return 1 and write t on a hit in [tmin, tmax]. */
int vzo_hit_sphere_f32(const float o[3], const float d[3], float tmin, float tmax,
                       const vzo_sphere* s, float* t);
int vzo_hit_sphere_f64(const double o[3], const double d[3], double tmin, double tmax,
                       const vzo_sphere* s, double* t);

/* LBVH */
typedef struct vzo_bvh vzo_bvh;
vzo_bvh* vzo_bvh_build(const vzo_sphere* s, uint32_t n);
void vzo_bvh_free(vzo_bvh* b);
uint32_t vzo_bvh_num_prims(const vzo_bvh* b);
uint32_t vzo_bvh_num_nodes(const vzo_bvh* b);
const vzo_node* vzo_bvh_nodes(const vzo_bvh* b);
const uint32_t* vzo_bvh_leaf_prims(const vzo_bvh* b);    /* leaf -> primitive id */
const uint32_t* vzo_bvh_morton(const vzo_bvh* b);        /* Morton key of each leaf's sphere (sorted for an LBVH) */
const float* vzo_bvh_node_boxes(const vzo_bvh* b);       /* internal boxes, 6 floats each */
uint32_t vzo_bvh_depth(const vzo_bvh* b);
/* The 4-wide tree collapsed from the binary one (root = wide node 0). */
uint32_t vzo_bvh_num_wide_nodes(const vzo_bvh* b);
const vzo_wnode* vzo_bvh_wide_nodes(const vzo_bvh* b);
uint32_t vzo_bvh_wide_depth(const vzo_bvh* b);
/* Which build made the binary tree: an LBVH, or (2 to 1024 spheres) the
 * binned-SAH build. */
enum { VZO_BUILD_LBVH = 0, VZO_BUILD_SAH = 1 };
uint32_t vzo_bvh_builder(const vzo_bvh* b);
/* One wide node's slot boxes against a ray (float): masks[0]/tn_fast by the
 * traversal's one-fma-per-plane test (planes moved out by a rounding bound),
 * masks[1]/tn_exact by the exact-form robust slab test on the dequantised
 * boxes.  The first must contain the second. */
void vzo_wide_box_check(const vzo_wnode* w, const float o[3], const float d[3], float tmin, float tmax,
                        uint32_t masks[2], float tn_fast[4], float tn_exact[4]);
/* Returns 0 when every structural invariant holds (binary and wide trees),
 * else a nonzero code. */
int vzo_bvh_validate(const vzo_bvh* b, const vzo_sphere* s);

/* Closest hit over a ray array (tmin/tmax from each ray), traversing the
 * 4-wide tree; node_visits counts wide nodes.  vzo_trace_binary traverses
 * the binary tree instead (same hits, its own counts). */
void vzo_trace(const vzo_bvh* b, const vzo_sphere* s, const vzo_ray* rays, uint32_t n,
               vzo_hit* hits, uint64_t* node_visits, uint64_t* sphere_tests);
void vzo_trace_binary(const vzo_bvh* b, const vzo_sphere* s, const vzo_ray* rays, uint32_t n,
                      vzo_hit* hits, uint64_t* node_visits, uint64_t* sphere_tests);
void vzo_trace_brute(const vzo_sphere* s, uint32_t nsph, const vzo_ray* rays, uint32_t n,
                     vzo_hit* hits);
void vzo_trace_brute_f64(const vzo_sphere* s, uint32_t nsph, const vzo_ray* rays, uint32_t n,
                         double* t_out, uint32_t* prim_out);

/* Scatter kernels, exposed for invariant tests.  Return 1 if the path
 * continues, 0 if absorbed.  seed is advanced exactly as in the integrator. */
int vzo_scatter_f32(const vzo_material* m, const float d[3], const float n_out[3], int front,
                    uint32_t* seed, float new_dir[3]);
void vzo_sky_f32(const float d[3], float rgb[3]);

/* Integrator */
typedef struct {
    uint32_t width, height, spp, subframe, max_depth;
    vzo_camera cam;
    uint32_t row_start, row_step;   /* render rows y = row_start + k*row_step */
    uint32_t threads;               /* 0 = the CPUs this process may run on */
    uint32_t accum_fp64;            /* accum buffer is double4 instead of float4 */
    uint32_t binary_tree;           /* traverse the binary tree instead of the 4-wide one (same
                                       hits; the visit counts of a small scene rendered alone on
                                       the GPU, whose binary tree is staged in LDS) */
} vzo_render_params;

typedef struct {
    uint64_t segments, paths, hits, node_visits, sphere_tests;
    double seconds;
    uint32_t threads;
} vzo_stats;

/* FP32 integrator (the parity target).  accum: float4 or double4 per pixel
 * (read when subframe > 1, written always); image: uchar4 per pixel; either
 * may be NULL.  Uses the BVH when b != NULL, else brute force. */
void vzo_render(const vzo_bvh* b, const vzo_sphere* s, const vzo_material* m, uint32_t nsph,
                const vzo_render_params* p, void* accum, uint8_t* image, vzo_stats* st);
/* FP64 integrator (accuracy oracle): every path computed in double, accum is
 * double4. */
void vzo_render_f64(const vzo_bvh* b, const vzo_sphere* s, const vzo_material* m, uint32_t nsph,
                    const vzo_render_params* p, double* accum, vzo_stats* st);

/* Fixed ray set: the first n_int segments the FP32 integrator traces (pixel
 * order, sample order, bounce order) followed by n_iso isotropic rays from
 * points in the AABB of spheres 1..nsph-1, from mt19937(iso_seed). */
void vzo_make_rayset(const vzo_bvh* b, const vzo_sphere* s, const vzo_material* m, uint32_t nsph,
                     const vzo_render_params* p, uint32_t n_int, uint32_t n_iso, uint32_t iso_seed,
                     vzo_ray* out);

/* sRGB encode of one linear value (RayTracer.cu:16-36). */
uint8_t vzo_encode8(float c);

#ifdef __cplusplus
}
#endif
#endif
