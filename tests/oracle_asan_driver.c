/* Exercises every entry point of the CPU oracle (oracle/vz_oracle.c) on
 * small inputs, for a host build under AddressSanitizer and UBSan: an
 * out-of-bounds access or undefined behaviour aborts the run.  Exact-size
 * heap buffers throughout, so a one-past-the-end write is caught. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vz_oracle.h"

#define CHECK(c)                                                           \
    do {                                                                   \
        if (!(c)) {                                                        \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            return 1;                                                      \
        }                                                                  \
    } while (0)

static void* xalloc(size_t n) { return calloc(n ? n : 1, 1); }

static int run_scene(vzo_sphere* s, vzo_material* m, uint32_t n, int fp64_accum)
{
    vzo_bvh* b = vzo_bvh_build(s, n);
    CHECK(b != NULL);
    CHECK(vzo_bvh_validate(b, s) == 0);
    CHECK(vzo_bvh_num_prims(b) == n);
    (void)vzo_bvh_depth(b);
    vzo_camera cam;
    const float from[3] = {0.0f, 15.0f, -120.0f}, at[3] = {0.0f, 0.0f, 0.0f};
    vzo_camera_lookat(from, at, 40.0f, 2.0f, 0.1f, 10.0f, &cam);
    vzo_render_params p;
    memset(&p, 0, sizeof(p));
    p.width = 7;
    p.height = 5;
    p.spp = 3;
    p.subframe = 2;
    p.max_depth = 6;
    p.cam = cam;
    p.row_step = 1;
    p.threads = 2;
    p.accum_fp64 = (uint32_t)fp64_accum;
    p.binary_tree = (uint32_t)fp64_accum;
    const size_t npx = (size_t)p.width * p.height;
    void* acc = xalloc(npx * 4 * (fp64_accum ? sizeof(double) : sizeof(float)));
    uint8_t* img = (uint8_t*)xalloc(npx * 4);
    vzo_stats st;
    vzo_render(b, s, m, n, &p, acc, img, &st);
    CHECK(st.paths == npx * p.spp);
    vzo_render(NULL, s, m, n, &p, acc, NULL, &st);
    double* acc64 = (double*)xalloc(npx * 4 * sizeof(double));
    vzo_render_f64(b, s, m, n, &p, acc64, &st);
    p.row_start = 1;
    p.row_step = 2;
    vzo_render(b, s, m, n, &p, acc, img, &st);
    const uint32_t n_int = 300, n_iso = 200;
    vzo_ray* rays = (vzo_ray*)xalloc(sizeof(vzo_ray) * (n_int + n_iso));
    vzo_make_rayset(b, s, m, n, &p, n_int, n_iso, 9, rays);
    vzo_hit* h1 = (vzo_hit*)xalloc(sizeof(vzo_hit) * (n_int + n_iso));
    vzo_hit* h2 = (vzo_hit*)xalloc(sizeof(vzo_hit) * (n_int + n_iso));
    uint64_t nv = 0, ns = 0;
    vzo_trace(b, s, rays, n_int + n_iso, h1, &nv, &ns);
    vzo_trace_brute(s, n, rays, n_int + n_iso, h2);
    CHECK(memcmp(h1, h2, sizeof(vzo_hit) * (n_int + n_iso)) == 0);
    double* t64 = (double*)xalloc(sizeof(double) * (n_int + n_iso));
    uint32_t* p64 = (uint32_t*)xalloc(sizeof(uint32_t) * (n_int + n_iso));
    vzo_trace_brute_f64(s, n, rays, n_int + n_iso, t64, p64);
    free(acc);
    free(img);
    free(acc64);
    free(rays);
    free(h1);
    free(h2);
    free(t64);
    free(p64);
    vzo_bvh_free(b);
    return 0;
}

int main(void)
{
    /* RNG */
    uint32_t st = vzo_tea4(12345u, 7u);
    for (int i = 0; i < 100; i++) (void)vzo_rnd(&st);
    (void)vzo_lcg(&st);
    CHECK(vzo_mt19937_nth(5489u, 10000u) == 4123659995u);
    /* book scene: count first, then an exact-size buffer */
    uint32_t nb = vzo_book_scene(5489u, NULL, NULL, 0);
    vzo_sphere* bs = (vzo_sphere*)xalloc(sizeof(vzo_sphere) * nb);
    vzo_material* bm = (vzo_material*)xalloc(sizeof(vzo_material) * nb);
    CHECK(vzo_book_scene(5489u, bs, bm, nb) == nb);
    if (run_scene(bs, bm, nb, 0)) return 1;
    free(bs);
    free(bm);
    /* box scenes of several sizes, including 1 and 2 spheres, duplicates,
     * the largest SAH-built scene (1024) and an LBVH with SAH treelets */
    const uint32_t sizes[] = {0, 1, 2, 3, 64, 2049, 5000, 1023, 131071};
    for (unsigned k = 0; k < sizeof(sizes) / sizeof(sizes[0]); k++) {
        uint32_t n = sizes[k] + 1;
        vzo_sphere* s = (vzo_sphere*)xalloc(sizeof(vzo_sphere) * n);
        vzo_material* m = (vzo_material*)xalloc(sizeof(vzo_material) * n);
        vzo_box_scene(sizes[k], 100u + k, s, m);
        if (k == 6)
            for (uint32_t i = 10; i < 500; i++) s[i] = s[9];
        if (run_scene(s, m, n, (int)(k & 1u))) return 1;
        free(s);
        free(m);
    }
    /* single-function entry points */
    const float o[3] = {0.0f, 0.5f, -5.0f}, d[3] = {0.01f, 0.0f, 1.0f};
    vzo_sphere one = {0.0f, 0.0f, 0.0f, 1.0f};
    float t = 0.0f;
    CHECK(vzo_hit_sphere_f32(o, d, 0.001f, 1e16f, &one, &t) == 1);
    const double od[3] = {0.0, 0.5, -5.0}, dd[3] = {0.01, 0.0, 1.0};
    double td = 0.0;
    CHECK(vzo_hit_sphere_f64(od, dd, 0.001, 1e16, &one, &td) == 1);
    for (uint32_t type = 0; type < 3; type++) {
        vzo_material mt = {type, {0.5f, 0.5f, 0.5f}, 0.3f, 1.5f};
        const float n_out[3] = {0.0f, 1.0f, 0.0f}, din[3] = {0.3f, -1.0f, 0.1f};
        float nd[3];
        for (uint32_t seed = 0; seed < 200; seed++) {
            uint32_t sd = seed;
            (void)vzo_scatter_f32(&mt, din, n_out, (int)(seed & 1u), &sd, nd);
        }
    }
    float rgb[3];
    vzo_sky_f32(d, rgb);
    for (int i = -10; i < 300; i++) (void)vzo_encode8((float)i / 256.0f);
    printf("oracle sanitizer run: ok\n");
    return 0;
}
