/*
 * vz_oracle.c -- CPU oracle for the sphere path tracer.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never a
 * fallback for it.  See vz_oracle.h for what each function restates and what
 * pins it.  Build: oracle/Makefile (gcc -O3 -ffp-contract=off).
 */
#define _GNU_SOURCE
#include "vz_oracle.h"

#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#define VZO_MAX_STACK 128

struct vzo_bvh {
    uint32_t n;           /* primitives */
    vzo_node* nodes;      /* n-1 internal nodes, root = 0 */
    uint32_t* leaf_prim;  /* n */
    uint32_t* morton;     /* n, sorted */
    float* boxes;         /* (n-1)*6 internal-node boxes lo xyz, hi xyz */
    uint32_t depth;
    /* the 4-wide tree (see collapse_wide) */
    uint32_t nw;          /* wide nodes */
    vzo_wnode* wnodes;
    uint32_t wdepth;      /* wide levels from the root to the deepest leaf */
    uint32_t builder;     /* VZO_BUILD_LBVH or VZO_BUILD_SAH */
};

/* ---------------------------------------------------------------- RNG --- */
/* random.cuh:31-46: TEA with 4 rounds; returns v0 */
uint32_t vzo_tea4(uint32_t v0, uint32_t v1)
{
    uint32_t s0 = 0;
    for (int n = 0; n < 4; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}

/* random.cuh:48-55: Numerical Recipes LCG, low 24 bits returned */
uint32_t vzo_lcg(uint32_t* state)
{
    *state = 1664525u * *state + 1013904223u;
    return *state & 0x00FFFFFFu;
}

/* random.cuh:63-67: 24-bit integer / 2^24, exact in float */
float vzo_rnd(uint32_t* state) { return (float)vzo_lcg(state) / (float)0x01000000; }

/* ------------------------------------------------------------ mt19937 --- */
/* Matsumoto & Nishimura 1998, the parameters std::mt19937 fixes. */
typedef struct {
    uint32_t mt[624];
    int idx;
} vzo_mt;

static void mt_seed(vzo_mt* g, uint32_t seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < 624; i++) g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
}

static uint32_t mt_next(vzo_mt* g)
{
    if (g->idx >= 624) {
        for (int i = 0; i < 624; i++) {
            uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            g->mt[i] = g->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        g->idx = 0;
    }
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* The build's uniform float: the top 24 bits of one output, exact in [0,1). */
static float mt_uniform(vzo_mt* g) { return (float)(mt_next(g) >> 8) * (1.0f / 16777216.0f); }

uint32_t vzo_mt19937_nth(uint32_t seed, uint32_t n)
{
    vzo_mt g;
    mt_seed(&g, seed);
    uint32_t v = 0;
    for (uint32_t i = 0; i < n; i++) v = mt_next(&g);
    return v;
}

/* ------------------------------------------------------------- scenes --- */
static void set_mat(vzo_material* m, uint32_t type, float r, float g, float b, float fuzz, float ir)
{
    m->type = type;
    m->albedo[0] = r;
    m->albedo[1] = g;
    m->albedo[2] = b;
    m->fuzz = fuzz;
    m->ir = ir;
}

static void set_sph(vzo_sphere* s, float x, float y, float z, float r)
{
    s->cx = x;
    s->cy = y;
    s->cz = z;
    s->r = r;
}

/* Draw the material of one random sphere (Scene.h:31-57), x->y->z order. */
static void draw_material(vzo_mt* g, float choose, vzo_material* m)
{
    if ((double)choose < 0.8) {
        float a0 = mt_uniform(g), a1 = mt_uniform(g), a2 = mt_uniform(g);
        float b0 = mt_uniform(g), b1 = mt_uniform(g), b2 = mt_uniform(g);
        set_mat(m, VZO_LAMBERTIAN, a0 * b0, a1 * b1, a2 * b2, 0.0f, 0.0f);
    } else if ((double)choose < 0.95) {
        float r = 0.5f + (1.0f - 0.5f) * mt_uniform(g);
        float gg = 0.5f + (1.0f - 0.5f) * mt_uniform(g);
        float b = 0.5f + (1.0f - 0.5f) * mt_uniform(g);
        float fuzz = 0.0f + (0.5f - 0.0f) * mt_uniform(g);
        set_mat(m, VZO_METAL, r, gg, b, fuzz, 0.0f);
    } else {
        set_mat(m, VZO_DIELECTRIC, 0.0f, 0.0f, 0.0f, 0.0f, 1.5f);
    }
}

/* Scene.h:13-80.  choose, then centre x, then centre z are drawn for every
 * grid cell; material draws only for kept spheres. */
uint32_t vzo_book_scene(uint32_t seed, vzo_sphere* s, vzo_material* m, uint32_t cap)
{
    vzo_mt g;
    mt_seed(&g, seed);
    uint32_t n = 0;
#define PUSH(X, Y, Z, R)                         \
    do {                                         \
        if (n < cap) set_sph(&s[n], X, Y, Z, R); \
    } while (0)
    PUSH(0.0f, -1000.0f, 0.0f, 1000.0f);
    if (n < cap) set_mat(&m[n], VZO_LAMBERTIAN, 0.5f, 0.5f, 0.5f, 0.0f, 0.0f);
    n++;
    for (int a = -11; a < 11; a++) {
        for (int b = -11; b < 11; b++) {
            float choose = mt_uniform(&g);
            float ux = mt_uniform(&g);
            float uz = mt_uniform(&g);
            float cx = (float)((double)a + 0.9 * (double)ux);
            float cy = 0.2f;
            float cz = (float)((double)b + 0.9 * (double)uz);
            float dx = cx - 4.0f, dy = cy - 0.2f, dz = cz - 0.0f;
            float len = sqrtf(dx * dx + dy * dy + dz * dz);
            if ((double)len > 0.9) {
                vzo_material mm;
                draw_material(&g, choose, &mm);
                PUSH(cx, cy, cz, 0.2f);
                if (n < cap) m[n] = mm;
                n++;
            }
        }
    }
    PUSH(0.0f, 1.0f, 0.0f, 1.0f);
    if (n < cap) set_mat(&m[n], VZO_DIELECTRIC, 0.0f, 0.0f, 0.0f, 0.0f, 1.5f);
    n++;
    PUSH(-4.0f, 1.0f, 0.0f, 1.0f);
    if (n < cap) set_mat(&m[n], VZO_LAMBERTIAN, 0.4f, 0.2f, 0.1f, 0.0f, 0.0f);
    n++;
    PUSH(4.0f, 1.0f, 0.0f, 1.0f);
    if (n < cap) set_mat(&m[n], VZO_METAL, 0.7f, 0.6f, 0.5f, 0.0f, 0.0f);
    n++;
#undef PUSH
    return n;
}

/* Random-box scene: ground sphere, then n_small spheres with centres uniform
 * in [-100,100]x[0,20]x[-100,100] and radii exp(U[ln 0.01, ln 0.5]).  Per
 * sphere the draws are x, y, z, radius, choose, then the material draws. */
void vzo_box_scene(uint32_t n_small, uint32_t seed, vzo_sphere* s, vzo_material* m)
{
    const float ln_lo = -0x1.26bb1cp+2f; /* logf(0.01f) */
    const float ln_hi = -0x1.62e430p-1f; /* logf(0.5f)  */
    vzo_mt g;
    mt_seed(&g, seed);
    set_sph(&s[0], 0.0f, -1000.0f, 0.0f, 1000.0f);
    set_mat(&m[0], VZO_LAMBERTIAN, 0.5f, 0.5f, 0.5f, 0.0f, 0.0f);
    for (uint32_t i = 1; i <= n_small; i++) {
        float x = -100.0f + 200.0f * mt_uniform(&g);
        float y = 0.0f + 20.0f * mt_uniform(&g);
        float z = -100.0f + 200.0f * mt_uniform(&g);
        float r = expf(ln_lo + (ln_hi - ln_lo) * mt_uniform(&g));
        float choose = mt_uniform(&g);
        set_sph(&s[i], x, y, z, r);
        draw_material(&g, choose, &m[i]);
    }
}

/* ------------------------------------------------------------- camera --- */
static void n3(float v[3])
{
    float inv = 1.0f / sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] *= inv;
    v[1] *= inv;
    v[2] *= inv;
}

static void cross3(const float a[3], const float b[3], float out[3])
{
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

/* Camera.cpp:24-37, camera.h:20-24: W = forward*focus, U = norm(W x up),
 * V = norm(U x W), then U *= focus*aspect*2h, V *= focus*2h. */
void vzo_camera_lookat(const float from[3], const float at[3], float vfov_deg, float aspect,
                       float aperture, float focus, vzo_camera* out)
{
    float f[3] = {at[0] - from[0], at[1] - from[1], at[2] - from[2]};
    n3(f);
    float W[3] = {f[0] * focus, f[1] * focus, f[2] * focus};
    const float up[3] = {0.0f, 1.0f, 0.0f};
    float U[3], V[3];
    cross3(W, up, U);
    n3(U);
    cross3(U, W, V);
    n3(V);
    float theta = vfov_deg * 0.01745329251994329577f;
    float h = tanf(theta * 0.5f);
    float vh = 2.0f * h;
    float vw = aspect * vh;
    float su = focus * vw, sv = focus * vh;
    for (int k = 0; k < 3; k++) {
        out->origin[k] = from[k];
        out->u[k] = U[k] * su;
        out->v[k] = V[k] * sv;
        out->w[k] = W[k];
    }
    out->lens_radius = aperture * 0.5f;
}

/* ------------------------------------------------------ node decoding --- */
/* Per-axis scale 2^(e-127), e = biased exponent byte k of meta. */
static inline float vzo_scale_of(uint32_t meta, int k)
{
    uint32_t e = (meta >> (8 * k)) & 0xFFu;
    union {
        uint32_t u;
        float f;
    } cv;
    cv.u = e << 23;
    return cv.f;
}

/* -------------------------------------------- real-generic instances --- */
#define REAL float
#define SUF _f32
#define SQRT sqrtf
#define FMA fmaf
#define FMIN fminf
#define FMAX fmaxf
#define FABS fabsf
#define COPYSIGN copysignf
#define ROBUST (1.0f + 0x1p-20f)
#include "vz_oracle_real.inc"
#undef REAL
#undef SUF
#undef SQRT
#undef FMA
#undef FMIN
#undef FMAX
#undef FABS
#undef COPYSIGN
#undef ROBUST

#define REAL double
#define SUF _f64
#define SQRT sqrt
#define FMA fma
#define FMIN fmin
#define FMAX fmax
#define FABS fabs
#define COPYSIGN copysign
#define ROBUST (1.0 + 0x1p-48)
#include "vz_oracle_real.inc"
#undef REAL
#undef SUF
#undef SQRT
#undef FMA
#undef FMIN
#undef FMAX
#undef FABS
#undef COPYSIGN
#undef ROBUST

/* Both slot-box tests of one wide node for one ray (float): masks[0] and
 * tn_fast from the traversal's test, masks[1] and tn_exact from the
 * exact-form robust test it must contain. */
void vzo_wide_box_check(const vzo_wnode* w, const float o[3], const float d[3], float tmin, float tmax,
                        uint32_t masks[2], float tn_fast[4], float tn_exact[4])
{
    float inv[3];
    const float org[3] = {o[0], o[1], o[2]};
    inv_dir_f32(mk_f32(d[0], d[1], d[2]), inv);
    masks[0] = wide_box_f32(w, org, inv, tmin, tmax * (1.0f + 0x1p-20f), tn_fast);
    masks[1] = wide_box_exact_f32(w, org, inv, tmin, tmax, tn_exact);
}

int vzo_hit_sphere_f32(const float o[3], const float d[3], float tmin, float tmax,
                       const vzo_sphere* s, float* t)
{
    return hit_sphere_f32(mk_f32(o[0], o[1], o[2]), mk_f32(d[0], d[1], d[2]), tmin, tmax, s->cx,
                          s->cy, s->cz, s->r, t);
}

int vzo_hit_sphere_f64(const double o[3], const double d[3], double tmin, double tmax,
                       const vzo_sphere* s, double* t)
{
    return hit_sphere_f64(mk_f64(o[0], o[1], o[2]), mk_f64(d[0], d[1], d[2]), tmin, tmax, s->cx,
                          s->cy, s->cz, s->r, t);
}

int vzo_scatter_f32(const vzo_material* m, const float d[3], const float n_out[3], int front,
                    uint32_t* seed, float new_dir[3])
{
    v3_f32 nd = mk_f32(0.0f, 0.0f, 0.0f), att = nd;
    int ok = scatter_f32(m, mk_f32(d[0], d[1], d[2]), mk_f32(n_out[0], n_out[1], n_out[2]), front,
                         seed, &nd, &att);
    new_dir[0] = nd.x;
    new_dir[1] = nd.y;
    new_dir[2] = nd.z;
    return ok;
}

void vzo_sky_f32(const float d[3], float rgb[3])
{
    v3_f32 c = sky_f32(mk_f32(d[0], d[1], d[2]));
    rgb[0] = c.x;
    rgb[1] = c.y;
    rgb[2] = c.z;
}

/* --------------------------------------------------------------- LBVH --- */
static uint32_t expand10(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

/* 10 bits per axis of the centre relative to the centroid bounds:
 * q = (int)((c - lo) * (1024 / ext)) clamped to [0, 1023]; x in the high
 * bit of each triple. */
static uint32_t morton_of(const vzo_sphere* s, const float lo[3], const float sc[3])
{
    float c[3] = {s->cx, s->cy, s->cz};
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
        float f = (c[k] - lo[k]) * sc[k];
        int qi = (int)f;
        if (qi < 0) qi = 0;
        if (qi > 1023) qi = 1023;
        q[k] = (uint32_t)qi;
    }
    return (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
}

#define VZO_LARGE_KEY 0x40000000u

static inline int is_large(const vzo_sphere* s, float thr) { return 2.0f * fabsf(s->r) > thr; }

/* Centroid bounds lo/hi and Morton scale sc over the spheres with
 * 2|r| <= thr (all of them when thr < 0).  An axis with no sphere counted
 * gets lo = hi = 0.  iso = 0: sc = ext > 0 ? 1024/ext : 0 per axis; iso = 1:
 * one scale 1024/max ext for every axis (cubic cells, so a thin slab of
 * spheres is not split along its thin axis as often as along the others). */
static void centroid_bounds(const vzo_sphere* s, uint32_t n, float thr, int iso, float lo[3], float hi[3],
                            float sc[3])
{
    for (int k = 0; k < 3; k++) {
        lo[k] = 3.4e38f;
        hi[k] = -3.4e38f;
    }
    for (uint32_t i = 0; i < n; i++) {
        if (thr >= 0.0f && is_large(&s[i], thr)) continue;
        const float c[3] = {s[i].cx, s[i].cy, s[i].cz};
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], c[k]);
            hi[k] = fmaxf(hi[k], c[k]);
        }
    }
    float extmax = 0.0f;
    for (int k = 0; k < 3; k++) {
        if (lo[k] > hi[k]) lo[k] = hi[k] = 0.0f;
        const float ext = hi[k] - lo[k];
        sc[k] = ext > 0.0f ? 1024.0f / ext : 0.0f;
        extmax = fmaxf(extmax, ext);
    }
    if (iso)
        for (int k = 0; k < 3; k++) sc[k] = extmax > 0.0f ? 1024.0f / extmax : 0.0f;
}

static inline int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }

/* Karras 2012 delta: common prefix length of keys i and j, with the index
 * appended to break ties; -1 outside [0, n). */
static inline int kdelta(const uint32_t* k, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz32(a ^ b);
}

/* Stable LSD radix sort of (key, value), 8-bit digits. */
static void radix_sort(uint32_t* key, uint32_t* val, uint32_t n)
{
    uint32_t* k2 = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* v2 = (uint32_t*)malloc(sizeof(uint32_t) * n);
    for (int pass = 0; pass < 4; pass++) {
        uint32_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        int sh = pass * 8;
        for (uint32_t i = 0; i < n; i++) cnt[((key[i] >> sh) & 0xFF) + 1]++;
        for (int d = 0; d < 256; d++) cnt[d + 1] += cnt[d];
        for (uint32_t i = 0; i < n; i++) {
            uint32_t d = (key[i] >> sh) & 0xFF;
            uint32_t pos = cnt[d]++;
            k2[pos] = key[i];
            v2[pos] = val[i];
        }
        memcpy(key, k2, sizeof(uint32_t) * n);
        memcpy(val, v2, sizeof(uint32_t) * n);
    }
    free(k2);
    free(v2);
}

static float f_as(uint32_t u)
{
    union {
        uint32_t u;
        float f;
    } cv;
    cv.u = u;
    return cv.f;
}
static uint32_t u_as(float f)
{
    union {
        uint32_t u;
        float f;
    } cv;
    cv.f = f;
    return cv.u;
}

/* Biased exponent e (>= 1) of the smallest power of two s with
 * fl(lo + s*255) >= hi. */
static uint32_t pick_exp(float lo, float hi)
{
    float ext = hi - lo;
    uint32_t e = 1;
    if (ext > 0.0f) {
        float m = ext / 255.0f;
        uint32_t b = u_as(m);
        uint32_t ex = b >> 23, mant = b & 0x7FFFFFu;
        e = (ex == 0) ? 1u : ex + (mant != 0 ? 1u : 0u);
        if (e < 1) e = 1;
    }
    while (e < 254 && fmaf(f_as(e << 23), 255.0f, lo) < hi) e++;
    return e;
}

static uint32_t q_lo(float v, float lo, float s)
{
    float qf = floorf((v - lo) / s);
    int q = (int)qf;
    if (q < 0) q = 0;
    if (q > 255) q = 255;
    while (q > 0 && fmaf(s, (float)q, lo) > v) q--;
    return (uint32_t)q;
}

static uint32_t q_hi(float v, float lo, float s)
{
    float qf = ceilf((v - lo) / s);
    int q = (int)qf;
    if (q < 0) q = 0;
    if (q > 255) q = 255;
    while (q < 255 && fmaf(s, (float)q, lo) < v) q++;
    return (uint32_t)q;
}

/* Pack internal node i's two children (boxes cb[0], cb[1]) relative to its
 * own box nb into the 32-byte node. */
static void pack_node(vzo_node* N, const float nb[6], const float cb0[6], const float cb1[6],
                      uint32_t child, uint32_t flags)
{
    uint32_t meta = 0;
    for (int k = 0; k < 3; k++) {
        float lo = nb[k], hi = nb[3 + k];
        uint32_t e = pick_exp(lo, hi);
        float s = f_as(e << 23);
        meta |= e << (8 * k);
        uint32_t l0 = q_lo(cb0[k], lo, s), h0 = q_hi(cb0[3 + k], lo, s);
        uint32_t l1 = q_lo(cb1[k], lo, s), h1 = q_hi(cb1[3 + k], lo, s);
        N->o[k] = lo;
        N->q[k] = l0 | (h0 << 8) | (l1 << 16) | (h1 << 24);
    }
    N->meta = meta | (flags << 24);
    N->child = child;
}

static void leaf_box(const vzo_sphere* s, float b[6])
{
    float r = fabsf(s->r);
    b[0] = s->cx - r;
    b[1] = s->cy - r;
    b[2] = s->cz - r;
    b[3] = s->cx + r;
    b[4] = s->cy + r;
    b[5] = s->cz + r;
}

/* Surface-area measure of a box: dx*dy + dy*dz + dz*dx, in that order. */
static float box_area(const float b[6])
{
    float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return dx * dy + dy * dz + dz * dx;
}

#define VZO_WIDE 4
#define VZO_LEAF_REF 0x80000000u

/* Children of the wide node rooted at binary internal node r: start with r's
 * two children; while fewer than 4, open the internal child of largest box
 * area (the first such slot on a tie) into its two children, in place, so
 * the slots keep left-to-right leaf order.  Children are binary references:
 * a leaf index | VZO_LEAF_REF, or an internal node index.  Returns the count. */
static int wide_children(const uint32_t* gam, const uint32_t* flg, const float* boxes, uint32_t r,
                         uint32_t out[VZO_WIDE])
{
    int cnt = 2;
    for (int c = 0; c < 2; c++)
        out[c] = ((flg[r] >> c) & 1u) ? ((gam[r] + (uint32_t)c) | VZO_LEAF_REF) : gam[r] + (uint32_t)c;
    while (cnt < VZO_WIDE) {
        int best = -1;
        float best_a = 0.0f;
        for (int k = 0; k < cnt; k++) {
            if (out[k] & VZO_LEAF_REF) continue;
            float a = box_area(&boxes[out[k] * 6]);
            if (best < 0 || a > best_a) {
                best = k;
                best_a = a;
            }
        }
        if (best < 0) break;
        uint32_t b = out[best];
        for (int k = cnt - 1; k > best; k--) out[k + 1] = out[k];
        out[best] = ((flg[b] & 1u) ? (gam[b] | VZO_LEAF_REF) : gam[b]);
        out[best + 1] = ((flg[b] & 2u) ? ((gam[b] + 1u) | VZO_LEAF_REF) : gam[b] + 1u);
        cnt++;
    }
    return cnt;
}

/* The 4-wide tree: wide roots are binary node 0 and every internal child of a
 * wide root; wide node indices number the roots in binary index order.  Each
 * wide node is quantised like a binary node: relative to its own box (the
 * binary box of its root) with a power-of-two scale per axis, planes rounded
 * outward. */
static void collapse_wide(vzo_bvh* b, const vzo_sphere* s, const uint32_t* gam, const uint32_t* flg)
{
    uint32_t ni = b->n - 1;
    uint32_t* kids = (uint32_t*)calloc((size_t)ni * VZO_WIDE, sizeof(uint32_t));
    uint8_t* cnt = (uint8_t*)calloc(ni, 1);
    uint8_t* root = (uint8_t*)calloc(ni, 1);
    uint32_t* level = (uint32_t*)calloc(ni, sizeof(uint32_t));
    uint32_t* widx = (uint32_t*)calloc(ni, sizeof(uint32_t));
    uint32_t* queue = (uint32_t*)malloc(sizeof(uint32_t) * ni);
    uint32_t qh = 0, qt = 0, maxlev = 0;
    root[0] = 1;
    level[0] = 1;
    queue[qt++] = 0;
    while (qh < qt) {
        uint32_t r = queue[qh++];
        cnt[r] = (uint8_t)wide_children(gam, flg, b->boxes, r, &kids[(size_t)r * VZO_WIDE]);
        if (level[r] > maxlev) maxlev = level[r];
        for (int c = 0; c < cnt[r]; c++) {
            uint32_t ch = kids[(size_t)r * VZO_WIDE + c];
            if (ch & VZO_LEAF_REF) continue;
            root[ch] = 1;
            level[ch] = level[r] + 1;
            queue[qt++] = ch;
        }
    }
    uint32_t nw = 0;
    for (uint32_t i = 0; i < ni; i++)
        if (root[i]) widx[i] = nw++;
    b->nw = nw;
    b->wdepth = maxlev + 1;  /* wide node levels plus the leaf level */
    b->wnodes = (vzo_wnode*)calloc(nw, sizeof(vzo_wnode));
    for (uint32_t r = 0; r < ni; r++) {
        if (!root[r]) continue;
        vzo_wnode* W = &b->wnodes[widx[r]];
        const float* nb = &b->boxes[r * 6];
        uint32_t meta = 0;
        float sc[3];
        for (int k = 0; k < 3; k++) {
            uint32_t e = pick_exp(nb[k], nb[3 + k]);
            sc[k] = f_as(e << 23);
            meta |= e << (8 * k);
            W->o[k] = nb[k];
        }
        for (int c = 0; c < cnt[r]; c++) {
            uint32_t ch = kids[(size_t)r * VZO_WIDE + c];
            float cb[6];
            if (ch & VZO_LEAF_REF) {
                uint32_t leaf = ch & ~VZO_LEAF_REF;
                leaf_box(&s[b->leaf_prim[leaf]], cb);
                W->child[c] = leaf;
                meta |= 1u << (24 + c);
            } else {
                memcpy(cb, &b->boxes[ch * 6], sizeof(cb));
                W->child[c] = widx[ch];
            }
            meta |= 1u << (28 + c);
            for (int k = 0; k < 3; k++) {
                W->qlo[k] |= q_lo(cb[k], nb[k], sc[k]) << (8 * c);
                W->qhi[k] |= q_hi(cb[3 + k], nb[k], sc[k]) << (8 * c);
            }
        }
        W->meta = meta;
    }
    free(kids);
    free(cnt);
    free(root);
    free(level);
    free(widx);
    free(queue);
}

/* ---- binned-SAH build of a small scene (n <= VZO_SAH_MAX_PRIMS) -------------
 * Top down over ranges of a leaf permutation, root range [0, n-1].  A range
 * [a, b] at depth d < VZO_SAH_DEPTH_CAP is split by the binned surface-area
 * heuristic: per axis k whose centroid extent ext = hi - lo over the range is
 * positive, each sphere falls in bin min(NB - 1, (int)(((c_k - lo) / ext) *
 * NB)) (NB = VZO_SAH_BINS); with L and R the unions of the leaf boxes of bins
 * <= j and > j and nL, nR their sphere counts, the split after bin j costs
 * area(L) * nL + area(R) * nR (area = dx*dy + dy*dz + dz*dx), over splits with
 * both sides non-empty.  The cheapest wins, the lower axis and then the lower
 * bin on a tie; the range is partitioned stably (each side keeps its order).
 * With no axis of positive extent, or at depth >= VZO_SAH_DEPTH_CAP, the
 * range splits after its first floor(m/2) spheres as they stand.  The depth
 * cap bounds the tree: at most VZO_SAH_DEPTH_CAP + 10 internal levels for
 * 1024 spheres, so at most 35 wide levels, what the GPU traversal's stack
 * holds.
 * Nodes use the Karras layout, so every later stage (boxes, quantisation,
 * refit, collapse) is shared with the LBVH: the split after leaf position g
 * makes the children g (covering [a, g]) and g + 1 (covering [g + 1, b]),
 * each a leaf when its range has one sphere, and the root is node 0. */
#define VZO_SAH_MAX_PRIMS 1024u
#define VZO_SAH_BINS 32
#define VZO_SAH_DEPTH_CAP 24u
/* An LBVH of at least VZO_TREELET_MIN_PRIMS spheres has its treelets rebuilt
 * the same way: every maximal subtree of at most VZO_SAH_MAX_PRIMS leaves
 * whose root lies above VZO_SAH_DEPTH_CAP is replaced by the binned-SAH tree
 * over the same leaf range (same root index, same index set in the Karras
 * layout, so the nodes above it do not change), its splits counting depth
 * from the whole tree's root.  Smaller LBVHs keep Morton order all the way
 * down: on box scenes of a few thousand spheres the SAH treelets visited
 * more nodes, from about 2e5 spheres up they visit 4-6% fewer. */
#define VZO_TREELET_MIN_PRIMS 131072u

typedef struct {
    const vzo_sphere* s;
    uint32_t* perm;  /* leaf position -> primitive */
    uint32_t* tmp;
    uint32_t *gam, *flg, *parent_int, *parent_leaf;
} sah_ctx;

static inline float sah_cen(const vzo_sphere* s, int k) { return k == 0 ? s->cx : (k == 1 ? s->cy : s->cz); }

static inline int sah_bin(float c, float lo, float ext)
{
    int j = (int)(((c - lo) / ext) * (float)VZO_SAH_BINS);
    return j > VZO_SAH_BINS - 1 ? VZO_SAH_BINS - 1 : j;
}

static void sah_grow(float B[6], const float b[6])
{
    for (int k = 0; k < 3; k++) {
        B[k] = fminf(B[k], b[k]);
        B[3 + k] = fmaxf(B[3 + k], b[3 + k]);
    }
}

static void sah_range(sah_ctx* c, uint32_t a, uint32_t b, uint32_t self, uint32_t depth)
{
    const uint32_t m = b - a + 1;
    float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (uint32_t i = a; i <= b; i++)
        for (int k = 0; k < 3; k++) {
            const float v = sah_cen(&c->s[c->perm[i]], k);
            lo[k] = fminf(lo[k], v);
            hi[k] = fmaxf(hi[k], v);
        }
    int bk = -1, bj = -1;
    float bcost = 0.0f;
    if (depth < VZO_SAH_DEPTH_CAP) {
        for (int k = 0; k < 3; k++) {
            const float ext = hi[k] - lo[k];
            if (!(ext > 0.0f)) continue;
            float box[VZO_SAH_BINS][6];
            uint32_t cnt[VZO_SAH_BINS];
            for (int j = 0; j < VZO_SAH_BINS; j++) {
                for (int q = 0; q < 3; q++) {
                    box[j][q] = 3.4e38f;
                    box[j][3 + q] = -3.4e38f;
                }
                cnt[j] = 0;
            }
            for (uint32_t i = a; i <= b; i++) {
                const vzo_sphere* sp = &c->s[c->perm[i]];
                const int j = sah_bin(sah_cen(sp, k), lo[k], ext);
                float lb[6];
                leaf_box(sp, lb);
                sah_grow(box[j], lb);
                cnt[j]++;
            }
            float ra[VZO_SAH_BINS], R[6] = {3.4e38f, 3.4e38f, 3.4e38f, -3.4e38f, -3.4e38f, -3.4e38f};
            uint32_t rn[VZO_SAH_BINS], nr = 0;
            for (int j = VZO_SAH_BINS - 1; j >= 1; j--) {
                sah_grow(R, box[j]);
                nr += cnt[j];
                ra[j] = box_area(R);
                rn[j] = nr;
            }
            float L[6] = {3.4e38f, 3.4e38f, 3.4e38f, -3.4e38f, -3.4e38f, -3.4e38f};
            uint32_t nl = 0;
            for (int j = 0; j < VZO_SAH_BINS - 1; j++) {
                sah_grow(L, box[j]);
                nl += cnt[j];
                if (nl == 0 || rn[j + 1] == 0) continue;
                const float cost = box_area(L) * (float)nl + ra[j + 1] * (float)rn[j + 1];
                if (bk < 0 || cost < bcost) {
                    bk = k;
                    bj = j;
                    bcost = cost;
                }
            }
        }
    }
    uint32_t g;
    if (bk < 0) {
        g = a + m / 2 - 1;
    } else {
        const float ext = hi[bk] - lo[bk];
        uint32_t nl = 0, w = a;
        for (uint32_t i = a; i <= b; i++)
            if (sah_bin(sah_cen(&c->s[c->perm[i]], bk), lo[bk], ext) <= bj) c->tmp[w++] = c->perm[i];
        nl = w - a;
        for (uint32_t i = a; i <= b; i++)
            if (sah_bin(sah_cen(&c->s[c->perm[i]], bk), lo[bk], ext) > bj) c->tmp[w++] = c->perm[i];
        memcpy(&c->perm[a], &c->tmp[a], sizeof(uint32_t) * m);
        g = a + nl - 1;
    }
    uint32_t f = 0;
    if (g == a) {
        f |= 1u;
        c->parent_leaf[g] = self;
    } else {
        c->parent_int[g] = self;
    }
    if (g + 1 == b) {
        f |= 2u;
        c->parent_leaf[g + 1] = self;
    } else {
        c->parent_int[g + 1] = self;
    }
    c->gam[self] = g;
    c->flg[self] = f;
    if (g > a) sah_range(c, a, g, g, depth + 1);
    if (g + 1 < b) sah_range(c, g + 1, b, g + 1, depth + 1);
}

vzo_bvh* vzo_bvh_build(const vzo_sphere* s, uint32_t n)
{
    if (n == 0) return NULL;
    vzo_bvh* b = (vzo_bvh*)calloc(1, sizeof(vzo_bvh));
    b->n = n;
    b->leaf_prim = (uint32_t*)malloc(sizeof(uint32_t) * n);
    b->morton = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t ni = n > 1 ? n - 1 : 0;
    b->nodes = (vzo_node*)calloc(ni ? ni : 1, sizeof(vzo_node));
    b->boxes = (float*)calloc((ni ? ni : 1) * 6, sizeof(float));
    /* Centroid bounds of all spheres (grid A), the size threshold, and the
     * centroid bounds of the spheres at or below it (grid B).  A sphere
     * whose diameter exceeds a quarter of the largest centroid extent is
     * "large": its key is its Morton code on grid A with bit 30 set, so the
     * first split of the tree separates it from the rest, and its huge box
     * (the r = 1000 ground) does not inflate a whole root-to-leaf path.  The
     * other spheres get their Morton code on grid B: their own bounds, cubic
     * cells. */
    float loA[3], hiA[3], scA[3], loB[3], hiB[3], scB[3];
    centroid_bounds(s, n, -1.0f, 0, loA, hiA, scA);
    float extmax = 0.0f;
    for (int k = 0; k < 3; k++) extmax = fmaxf(extmax, hiA[k] - loA[k]);
    const float thr = 0.25f * extmax;
    centroid_bounds(s, n, thr, 1, loB, hiB, scB);
    for (uint32_t i = 0; i < n; i++) {
        b->morton[i] = is_large(&s[i], thr) ? (morton_of(&s[i], loA, scA) | VZO_LARGE_KEY)
                                            : morton_of(&s[i], loB, scB);
        b->leaf_prim[i] = i;
    }
    b->depth = 1;
    b->builder = (n >= 2 && n <= VZO_SAH_MAX_PRIMS) ? VZO_BUILD_SAH : VZO_BUILD_LBVH;
    uint32_t* key_of_prim = (uint32_t*)malloc(sizeof(uint32_t) * n);
    memcpy(key_of_prim, b->morton, sizeof(uint32_t) * n);
    if (b->builder == VZO_BUILD_LBVH) radix_sort(b->morton, b->leaf_prim, n);
    if (n == 1) {
        free(key_of_prim);
        return b;
    }

    uint32_t* parent_int = (uint32_t*)malloc(sizeof(uint32_t) * ni);
    uint32_t* parent_leaf = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* gam = (uint32_t*)malloc(sizeof(uint32_t) * ni);
    uint32_t* flg = (uint32_t*)malloc(sizeof(uint32_t) * ni);
    const uint32_t* K = b->morton;
    int nn = (int)n;
    parent_int[0] = 0xFFFFFFFFu;
    if (b->builder == VZO_BUILD_SAH) {
        /* leaves in split order; morton[i] keeps the key of leaf i's sphere */
        uint32_t* tmp = (uint32_t*)malloc(sizeof(uint32_t) * n);
        sah_ctx c = {s, b->leaf_prim, tmp, gam, flg, parent_int, parent_leaf};
        sah_range(&c, 0, n - 1, 0, 0);
        free(tmp);
        nn = 0;  /* no Karras linking */
    }
    /* Karras 2012, algorithm 4 */
    for (int i = 0; i < nn - 1; i++) {
        int d = (kdelta(K, nn, i, i + 1) - kdelta(K, nn, i, i - 1)) > 0 ? 1 : -1;
        int dmin = kdelta(K, nn, i, i - d);
        int lmax = 2;
        while (kdelta(K, nn, i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (kdelta(K, nn, i, i + (l + t) * d) > dmin) l += t;
        int j = i + l * d;
        int dnode = kdelta(K, nn, i, j);
        int sx = 0;
        int div = 2;
        int t = (l + div - 1) / div;
        for (;;) {
            if (kdelta(K, nn, i, i + (sx + t) * d) > dnode) sx += t;
            if (t == 1) break;
            div *= 2;
            t = (l + div - 1) / div;
        }
        int g = i + sx * d + (d < 0 ? d : 0);
        int mn = i < j ? i : j, mx = i < j ? j : i;
        uint32_t f = 0;
        if (mn == g) {
            f |= 1u;
            parent_leaf[g] = (uint32_t)i;
        } else {
            parent_int[g] = (uint32_t)i;
        }
        if (mx == g + 1) {
            f |= 2u;
            parent_leaf[g + 1] = (uint32_t)i;
        } else {
            parent_int[g + 1] = (uint32_t)i;
        }
        gam[i] = (uint32_t)g;
        flg[i] = f;
    }
    if (b->builder == VZO_BUILD_LBVH && n >= VZO_TREELET_MIN_PRIMS) {
        /* maximal subtrees of <= VZO_SAH_MAX_PRIMS leaves above the depth cap,
         * found top down; each is rebuilt over its own leaf range */
        uint32_t* tmp = (uint32_t*)malloc(sizeof(uint32_t) * n);
        sah_ctx c = {s, b->leaf_prim, tmp, gam, flg, parent_int, parent_leaf};
        uint32_t stk[VZO_MAX_STACK * 2], sf[VZO_MAX_STACK * 2], sl[VZO_MAX_STACK * 2], sd[VZO_MAX_STACK * 2];
        int sp = 1;
        stk[0] = 0;
        sf[0] = 0;
        sl[0] = n - 1;
        sd[0] = 0;
        while (sp > 0) {
            sp--;
            const uint32_t nd = stk[sp], f = sf[sp], l = sl[sp], d = sd[sp];
            if (d >= VZO_SAH_DEPTH_CAP) continue;
            if (l - f + 1 <= VZO_SAH_MAX_PRIMS) {
                sah_range(&c, f, l, nd, d);
                continue;
            }
            const uint32_t g = gam[nd];
            if (!(flg[nd] & 2u) && sp < VZO_MAX_STACK * 2) {
                stk[sp] = g + 1; sf[sp] = g + 1; sl[sp] = l; sd[sp] = d + 1; sp++;
            }
            if (!(flg[nd] & 1u) && sp < VZO_MAX_STACK * 2) {
                stk[sp] = g; sf[sp] = f; sl[sp] = g; sd[sp] = d + 1; sp++;
            }
        }
        free(tmp);
    }
    /* morton[i]: the key of leaf i's sphere (sorted unless SAH reordered) */
    for (uint32_t i = 0; i < n; i++) b->morton[i] = key_of_prim[b->leaf_prim[i]];
    free(key_of_prim);
    /* bottom-up boxes: a node is finished when both children are */
    uint8_t* seen = (uint8_t*)calloc(ni, 1);
    for (uint32_t leaf = 0; leaf < n; leaf++) {
        uint32_t p = parent_leaf[leaf];
        for (;;) {
            if (++seen[p] < 2) break;
            float cb[2][6];
            for (int c = 0; c < 2; c++) {
                uint32_t ch = gam[p] + (uint32_t)c;
                if ((flg[p] >> c) & 1u)
                    leaf_box(&s[b->leaf_prim[ch]], cb[c]);
                else
                    memcpy(cb[c], &b->boxes[ch * 6], sizeof(float) * 6);
            }
            float* nb = &b->boxes[p * 6];
            for (int k = 0; k < 3; k++) {
                nb[k] = fminf(cb[0][k], cb[1][k]);
                nb[3 + k] = fmaxf(cb[0][3 + k], cb[1][3 + k]);
            }
            pack_node(&b->nodes[p], nb, cb[0], cb[1], gam[p], flg[p]);
            if (p == 0) break;
            p = parent_int[p];
        }
    }
    /* depth */
    {
        uint32_t stk[VZO_MAX_STACK * 2], dep[VZO_MAX_STACK * 2];
        int sp = 0;
        stk[sp] = 0;
        dep[sp] = 1;
        sp++;
        uint32_t maxd = 1;
        while (sp > 0) {
            sp--;
            uint32_t nd = stk[sp], dd = dep[sp];
            if (dd + 1 > maxd) maxd = dd + 1;
            for (int c = 0; c < 2; c++)
                if (!((flg[nd] >> c) & 1u) && sp < VZO_MAX_STACK * 2) {
                    stk[sp] = gam[nd] + (uint32_t)c;
                    dep[sp] = dd + 1;
                    sp++;
                }
        }
        b->depth = maxd;
    }
    collapse_wide(b, s, gam, flg);
    free(seen);
    free(parent_int);
    free(parent_leaf);
    free(gam);
    free(flg);
    return b;
}

void vzo_bvh_free(vzo_bvh* b)
{
    if (!b) return;
    free(b->nodes);
    free(b->leaf_prim);
    free(b->morton);
    free(b->boxes);
    free(b->wnodes);
    free(b);
}

uint32_t vzo_bvh_num_prims(const vzo_bvh* b) { return b->n; }
uint32_t vzo_bvh_num_nodes(const vzo_bvh* b) { return b->n > 1 ? b->n - 1 : 0; }
const vzo_node* vzo_bvh_nodes(const vzo_bvh* b) { return b->nodes; }
const uint32_t* vzo_bvh_leaf_prims(const vzo_bvh* b) { return b->leaf_prim; }
const uint32_t* vzo_bvh_morton(const vzo_bvh* b) { return b->morton; }
const float* vzo_bvh_node_boxes(const vzo_bvh* b) { return b->boxes; }
uint32_t vzo_bvh_depth(const vzo_bvh* b) { return b->depth; }
uint32_t vzo_bvh_num_wide_nodes(const vzo_bvh* b) { return b->nw; }
const vzo_wnode* vzo_bvh_wide_nodes(const vzo_bvh* b) { return b->wnodes; }
uint32_t vzo_bvh_wide_depth(const vzo_bvh* b) { return b->n > 1 ? b->wdepth : 1; }
uint32_t vzo_bvh_builder(const vzo_bvh* b) { return b->builder; }

/* Wide-tree invariants: node 0 is the root; every leaf and every other wide
 * node is one slot's child exactly once; slots are filled from 0 (2 to 4 of
 * them), empty slots are zero; each slot's dequantised box contains the
 * exact box of its child (a leaf's sphere box, or the union of the child
 * node's leaves); children cover consecutive leaf ranges in slot order.
 * Returns 0 or a code from 20 up. */
static int validate_wide(const vzo_bvh* b, const vzo_sphere* s, uint32_t w, uint32_t first, uint32_t* last,
                         float box[6], uint8_t* ref_leaf, uint8_t* ref_w, int depth)
{
    if (depth > 256) return 20;
    const vzo_wnode* W = &b->wnodes[w];
    uint32_t valid = W->meta >> 28, leafm = (W->meta >> 24) & 0xFu;
    if (valid != 0x3u && valid != 0x7u && valid != 0xFu) return 21;
    if (leafm & ~valid) return 22;
    uint32_t pos = first;
    for (int k = 0; k < 3; k++) {
        box[k] = 3.4e38f;
        box[3 + k] = -3.4e38f;
    }
    for (int c = 0; c < VZO_WIDE; c++) {
        if (!((valid >> c) & 1u)) {
            if (W->child[c] != 0) return 23;
            for (int k = 0; k < 3; k++)
                if (((W->qlo[k] >> (8 * c)) & 0xFFu) || ((W->qhi[k] >> (8 * c)) & 0xFFu)) return 23;
            continue;
        }
        float cb[6];
        uint32_t ch = W->child[c];
        if ((leafm >> c) & 1u) {
            if (ch >= b->n || ch != pos) return 24;
            ref_leaf[ch]++;
            leaf_box(&s[b->leaf_prim[ch]], cb);
            pos++;
        } else {
            if (ch == 0 || ch >= b->nw) return 25;
            ref_w[ch]++;
            uint32_t l;
            int rc = validate_wide(b, s, ch, pos, &l, cb, ref_leaf, ref_w, depth + 1);
            if (rc) return rc;
            pos = l + 1;
        }
        for (int k = 0; k < 3; k++) {
            float sc = vzo_scale_of(W->meta, k);
            float dl = fmaf(sc, (float)((W->qlo[k] >> (8 * c)) & 0xFFu), W->o[k]);
            float dh = fmaf(sc, (float)((W->qhi[k] >> (8 * c)) & 0xFFu), W->o[k]);
            if (!(dl <= cb[k] && dh >= cb[3 + k])) return 26;
            box[k] = fminf(box[k], cb[k]);
            box[3 + k] = fmaxf(box[3 + k], cb[3 + k]);
        }
    }
    for (int k = 0; k < 3; k++)
        if (W->o[k] != box[k]) return 27;  /* the quantisation origin is the node box's lo corner */
    *last = pos - 1;
    return 0;
}

/* Largest and smallest sphere centre per axis over leaves [f, l]. */
static void leaf_centre_bounds(const vzo_bvh* b, const vzo_sphere* s, uint32_t f, uint32_t l, float lo[3],
                               float hi[3])
{
    for (int k = 0; k < 3; k++) {
        lo[k] = 3.4e38f;
        hi[k] = -3.4e38f;
    }
    for (uint32_t i = f; i <= l; i++)
        for (int k = 0; k < 3; k++) {
            const float v = sah_cen(&s[b->leaf_prim[i]], k);
            lo[k] = fminf(lo[k], v);
            hi[k] = fmaxf(hi[k], v);
        }
}

/* Structural invariants: each leaf and each non-root internal node is
 * referenced exactly once; parent boxes contain child boxes; dequantised
 * child boxes contain the exact child boxes; the leaf order is a
 * permutation.  LBVH nodes: the Morton order (treelets sorted back) is
 * sorted, and each split is where the node's common prefix ends.  Nodes of
 * an SAH tree or treelet: each node's index is an end of its leaf range (the
 * Karras layout); above the depth cap, every split separates the two sides'
 * sphere centres strictly along some axis, or the range's centres all
 * coincide. */
static int key_prim_cmp(const void* x, const void* y)
{
    const uint64_t a = *(const uint64_t*)x, c = *(const uint64_t*)y;
    return a < c ? -1 : (a > c ? 1 : 0);
}

/* Is internal node nd (leaves [f, l], depth d) where a binned-SAH tree
 * starts: the root of an SAH build, or the root of an LBVH treelet. */
static int sah_root(const vzo_bvh* b, uint32_t nd, uint32_t f, uint32_t l, uint32_t d)
{
    if (b->builder == VZO_BUILD_SAH) return nd == 0;
    return b->n >= VZO_TREELET_MIN_PRIMS && d < VZO_SAH_DEPTH_CAP && l - f + 1 <= VZO_SAH_MAX_PRIMS;
}

int vzo_bvh_validate(const vzo_bvh* b, const vzo_sphere* s)
{
    uint32_t n = b->n;
    {
        uint8_t* seen = (uint8_t*)calloc(n, 1);
        int bad = 0;
        for (uint32_t i = 0; i < n && !bad; i++) {
            if (b->leaf_prim[i] >= n || seen[b->leaf_prim[i]]) bad = 1;
            else seen[b->leaf_prim[i]] = 1;
        }
        free(seen);
        if (bad) return 14;
    }
    /* the Morton order an LBVH was linked in: treelets keep their leaves'
     * (key, prim) pairs but not their order, so each treelet range is sorted
     * back first */
    uint32_t* rk = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* rp = (uint32_t*)malloc(sizeof(uint32_t) * n);
    memcpy(rk, b->morton, sizeof(uint32_t) * n);
    memcpy(rp, b->leaf_prim, sizeof(uint32_t) * n);
    if (b->builder == VZO_BUILD_LBVH && n >= VZO_TREELET_MIN_PRIMS) {
        uint64_t* kp = (uint64_t*)malloc(sizeof(uint64_t) * VZO_SAH_MAX_PRIMS);
        uint32_t stk[256], fst[256], lst[256], dep[256];
        int sp = 1;
        stk[0] = 0; fst[0] = 0; lst[0] = n - 1; dep[0] = 0;
        while (sp > 0) {
            sp--;
            const uint32_t nd = stk[sp], f = fst[sp], l = lst[sp], d = dep[sp];
            if (sah_root(b, nd, f, l, d)) {
                for (uint32_t i = f; i <= l; i++) kp[i - f] = ((uint64_t)b->morton[i] << 32) | b->leaf_prim[i];
                qsort(kp, l - f + 1, sizeof(uint64_t), key_prim_cmp);
                for (uint32_t i = f; i <= l; i++) {
                    rk[i] = (uint32_t)(kp[i - f] >> 32);
                    rp[i] = (uint32_t)kp[i - f];
                }
                continue;
            }
            const uint32_t g = b->nodes[nd].child, flags = b->nodes[nd].meta >> 24;
            if (!(g >= f && g < l)) break;  /* reported by the walk below */
            if (!(flags & 1u) && sp < 254) { stk[sp] = g; fst[sp] = f; lst[sp] = g; dep[sp] = d + 1; sp++; }
            if (!(flags & 2u) && sp < 254) { stk[sp] = g + 1; fst[sp] = g + 1; lst[sp] = l; dep[sp] = d + 1; sp++; }
        }
        free(kp);
    }
    int order = 0;
    if (b->builder == VZO_BUILD_LBVH)
        for (uint32_t i = 1; i < n && !order; i++) {
            if (rk[i - 1] > rk[i]) order = 1;
            else if (rk[i - 1] == rk[i] && rp[i - 1] > rp[i]) order = 2;
        }
    free(rp);
    if (order || n == 1) {
        free(rk);
        return order;
    }
    uint32_t ni = n - 1;
    uint8_t* ref_leaf = (uint8_t*)calloc(n, 1);
    uint8_t* ref_int = (uint8_t*)calloc(ni, 1);
    int rc = 0;
    for (uint32_t i = 0; i < ni && !rc; i++) {
        const vzo_node* N = &b->nodes[i];
        uint32_t flags = N->meta >> 24;
        for (int c = 0; c < 2; c++) {
            uint32_t ch = N->child + (uint32_t)c;
            float cb[6];
            if ((flags >> c) & 1u) {
                if (ch >= n) { rc = 3; break; }
                ref_leaf[ch]++;
                leaf_box(&s[b->leaf_prim[ch]], cb);
            } else {
                if (ch >= ni || ch == 0) { rc = 4; break; }
                ref_int[ch]++;
                memcpy(cb, &b->boxes[ch * 6], sizeof(cb));
            }
            for (int k = 0; k < 3; k++) {
                float sc = vzo_scale_of(N->meta, k);
                float dl = fmaf(sc, (float)((N->q[k] >> (16 * c)) & 0xFF), N->o[k]);
                float dh = fmaf(sc, (float)((N->q[k] >> (16 * c + 8)) & 0xFF), N->o[k]);
                if (!(dl <= cb[k] && dh >= cb[3 + k])) rc = 5;
                if (!(b->boxes[i * 6 + k] <= cb[k] && b->boxes[i * 6 + 3 + k] >= cb[3 + k])) rc = 6;
            }
        }
    }
    for (uint32_t i = 0; i < n && !rc; i++)
        if (ref_leaf[i] != 1) rc = 7;
    for (uint32_t i = 1; i < ni && !rc; i++)
        if (ref_int[i] != 1) rc = 8;
    /* split property via leaf ranges */
    if (!rc) {
        uint32_t stk[256], fst[256], lst[256], dep[256];
        uint8_t insah[256];
        int sp = 0;
        stk[0] = 0;
        fst[0] = 0;
        lst[0] = n - 1;
        dep[0] = 0;
        insah[0] = 0;
        sp = 1;
        int nn = (int)n;
        while (sp > 0 && !rc) {
            sp--;
            uint32_t nd = stk[sp], f = fst[sp], l = lst[sp], d = dep[sp];
            const int in = insah[sp] || sah_root(b, nd, f, l, d);
            const vzo_node* N = &b->nodes[nd];
            uint32_t g = N->child;
            if (!(g >= f && g < l)) { rc = 9; break; }
            if (!in) {
                int dn = kdelta(rk, nn, (int)f, (int)l);
                if (kdelta(rk, nn, (int)g, (int)g + 1) != dn) { rc = 10; break; }
            } else {
                if (nd != f && nd != l) { rc = 15; break; }
                if (d < VZO_SAH_DEPTH_CAP) {
                    float llo[3], lhi[3], rlo[3], rhi[3];
                    leaf_centre_bounds(b, s, f, g, llo, lhi);
                    leaf_centre_bounds(b, s, g + 1, l, rlo, rhi);
                    int sep = 0, same = 1;
                    for (int k = 0; k < 3; k++) {
                        if (lhi[k] < rlo[k]) sep = 1;
                        if (!(llo[k] == lhi[k] && rlo[k] == rhi[k] && lhi[k] == rlo[k])) same = 0;
                    }
                    if (!sep && !same) { rc = 16; break; }
                }
            }
            uint32_t flags = N->meta >> 24;
            if (((flags & 1u) != 0) != (g == f)) { rc = 11; break; }
            if (((flags & 2u) != 0) != (g + 1 == l)) { rc = 12; break; }
            if (!(flags & 1u)) {
                if (sp >= 255) { rc = 13; break; }
                stk[sp] = g; fst[sp] = f; lst[sp] = g; dep[sp] = d + 1; insah[sp] = (uint8_t)in; sp++;
            }
            if (!(flags & 2u)) {
                if (sp >= 255) { rc = 13; break; }
                stk[sp] = g + 1; fst[sp] = g + 1; lst[sp] = l; dep[sp] = d + 1; insah[sp] = (uint8_t)in; sp++;
            }
        }
    }
    free(rk);
    free(ref_leaf);
    free(ref_int);
    if (rc) return rc;
    {
        uint8_t* rl = (uint8_t*)calloc(n, 1);
        uint8_t* rw = (uint8_t*)calloc(b->nw, 1);
        uint32_t last = 0;
        float box[6];
        rc = b->nw == 0 ? 28 : validate_wide(b, s, 0, 0, &last, box, rl, rw, 0);
        if (!rc && last != n - 1) rc = 29;
        for (uint32_t i = 0; i < n && !rc; i++)
            if (rl[i] != 1) rc = 30;
        for (uint32_t i = 1; i < b->nw && !rc; i++)
            if (rw[i] != 1) rc = 31;
        free(rl);
        free(rw);
    }
    return rc;
}

void vzo_trace_binary(const vzo_bvh* b, const vzo_sphere* s, const vzo_ray* rays, uint32_t n,
                      vzo_hit* hits, uint64_t* node_visits, uint64_t* sphere_tests)
{
    uint64_t nv = 0, ns = 0;
    for (uint32_t i = 0; i < n; i++) {
        const vzo_ray* r = &rays[i];
        float t;
        uint32_t p;
        closest_bin_f32(b, s, mk_f32(r->ox, r->oy, r->oz), mk_f32(r->dx, r->dy, r->dz), r->tmin, r->tmax,
                        &t, &p, &nv, &ns);
        hits[i].t = (p == VZO_MISS) ? r->tmax : t;
        hits[i].prim = p;
    }
    if (node_visits) *node_visits = nv;
    if (sphere_tests) *sphere_tests = ns;
}

void vzo_trace(const vzo_bvh* b, const vzo_sphere* s, const vzo_ray* rays, uint32_t n,
               vzo_hit* hits, uint64_t* node_visits, uint64_t* sphere_tests)
{
    uint64_t nv = 0, ns = 0;
    for (uint32_t i = 0; i < n; i++) {
        const vzo_ray* r = &rays[i];
        float t;
        uint32_t p;
        closest_f32(b, s, mk_f32(r->ox, r->oy, r->oz), mk_f32(r->dx, r->dy, r->dz), r->tmin, r->tmax,
                    &t, &p, &nv, &ns);
        hits[i].t = (p == VZO_MISS) ? r->tmax : t;
        hits[i].prim = p;
    }
    if (node_visits) *node_visits = nv;
    if (sphere_tests) *sphere_tests = ns;
}

void vzo_trace_brute(const vzo_sphere* s, uint32_t nsph, const vzo_ray* rays, uint32_t n,
                     vzo_hit* hits)
{
    for (uint32_t i = 0; i < n; i++) {
        const vzo_ray* r = &rays[i];
        float best = r->tmax;
        uint32_t bp = VZO_MISS;
        v3_f32 o = mk_f32(r->ox, r->oy, r->oz), d = mk_f32(r->dx, r->dy, r->dz);
        float inv[3];
        inv_dir_f32(d, inv);
        for (uint32_t j = 0; j < nsph; j++) {
            float t;
            if (hit_gated_f32(o, d, inv, r->tmin, best, &s[j], &t)) {
                if (t < best || j < bp) {
                    best = t;
                    bp = j;
                }
            }
        }
        hits[i].t = best;
        hits[i].prim = bp;
    }
}

void vzo_trace_brute_f64(const vzo_sphere* s, uint32_t nsph, const vzo_ray* rays, uint32_t n,
                         double* t_out, uint32_t* prim_out)
{
    for (uint32_t i = 0; i < n; i++) {
        const vzo_ray* r = &rays[i];
        double best = r->tmax;
        uint32_t bp = VZO_MISS;
        v3_f64 o = mk_f64(r->ox, r->oy, r->oz), d = mk_f64(r->dx, r->dy, r->dz);
        double inv[3];
        inv_dir_f64(d, inv);
        for (uint32_t j = 0; j < nsph; j++) {
            double t;
            if (hit_gated_f64(o, d, inv, r->tmin, best, &s[j], &t)) {
                if (t < best || j < bp) {
                    best = t;
                    bp = j;
                }
            }
        }
        t_out[i] = best;
        prim_out[i] = bp;
    }
}

/* ------------------------------------------------------ sRGB encoding --- */
/* RayTracer.cu:16-43: clamp, sRGB with exponent 1/2.4, x*256 truncated,
 * clamped to 255. */
uint8_t vzo_encode8(float c)
{
    c = fmaxf(0.0f, fminf(c, 1.0f));
    float s = c < 0.0031308f ? 12.92f * c : 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
    s = fmaxf(0.0f, fminf(s, 1.0f));
    unsigned int q = (unsigned int)(s * 256.0f);
    return (uint8_t)(q < 255u ? q : 255u);
}

/* ---------------------------------------------------------- rendering --- */
typedef struct {
    const vzo_bvh* b;
    const vzo_sphere* s;
    const vzo_material* m;
    uint32_t nsph;
    const vzo_render_params* p;
    void* accum;
    uint8_t* image;
    int f64;
    uint32_t ntiles_x, ntiles_y, nrows;
    volatile uint32_t next;
    vzo_stats st[256];
} job_t;

typedef struct {
    job_t* j;
    int id;
} worker_t;

/* Progressive blend (RayTracer.cu:206-216): avg = sum * (1/spp); for
 * subframe > 1 the result is prev + (1/subframe)*(avg - prev), else avg. */
static void finish_pixel(job_t* j, uint32_t idx, const float sum32[3], const double sum64[3],
                         const double sumf64[3])
{
    const vzo_render_params* p = j->p;
    float outc[3];
    if (j->f64) {
        double* acc = (double*)j->accum + 4 * (size_t)idx;
        double avg[3];
        for (int k = 0; k < 3; k++) avg[k] = sumf64[k] / (double)p->spp;
        if (p->subframe > 1) {
            double a = 1.0 / (double)p->subframe;
            for (int k = 0; k < 3; k++) avg[k] = acc[k] + a * (avg[k] - acc[k]);
        }
        for (int k = 0; k < 3; k++) acc[k] = avg[k];
        acc[3] = 1.0;
        for (int k = 0; k < 3; k++) outc[k] = (float)avg[k];
    } else if (p->accum_fp64) {
        double avg[3];
        for (int k = 0; k < 3; k++) avg[k] = sum64[k] / (double)p->spp;
        double* acc = (double*)j->accum + 4 * (size_t)idx;
        if (p->subframe > 1 && acc) {
            double a = 1.0 / (double)p->subframe;
            for (int k = 0; k < 3; k++) avg[k] = acc[k] + a * (avg[k] - acc[k]);
        }
        if (j->accum) {
            for (int k = 0; k < 3; k++) acc[k] = avg[k];
            acc[3] = 1.0;
        }
        for (int k = 0; k < 3; k++) outc[k] = (float)avg[k];
    } else {
        float inv = 1.0f / (float)p->spp;
        float avg[3] = {sum32[0] * inv, sum32[1] * inv, sum32[2] * inv};
        float* acc = j->accum ? (float*)j->accum + 4 * (size_t)idx : NULL;
        if (p->subframe > 1 && acc) {
            float a = 1.0f / (float)p->subframe;
            for (int k = 0; k < 3; k++) avg[k] = acc[k] + a * (avg[k] - acc[k]);
        }
        if (acc) {
            for (int k = 0; k < 3; k++) acc[k] = avg[k];
            acc[3] = 1.0f;
        }
        for (int k = 0; k < 3; k++) outc[k] = avg[k];
    }
    if (j->image) {
        uint8_t* px = j->image + 4 * (size_t)idx;
        px[0] = vzo_encode8(outc[0]);
        px[1] = vzo_encode8(outc[1]);
        px[2] = vzo_encode8(outc[2]);
        px[3] = 255;
    }
}

static void* worker(void* arg)
{
    worker_t* w = (worker_t*)arg;
    job_t* j = w->j;
    const vzo_render_params* p = j->p;
    vzo_stats* st = &j->st[w->id];
    pctx_f32 c32 = {j->b, j->s, j->m, j->nsph, p->max_depth, p->binary_tree, NULL, 0, 0};
    pctx_f64 c64 = {j->b, j->s, j->m, j->nsph, p->max_depth, p->binary_tree, NULL, 0, 0};
    uint32_t ntiles = j->ntiles_x * j->ntiles_y;
    for (;;) {
        uint32_t t = __atomic_fetch_add(&j->next, 1u, __ATOMIC_RELAXED);
        if (t >= ntiles) break;
        uint32_t tx = t % j->ntiles_x, ty = t / j->ntiles_x;
        for (uint32_t ry = ty * 16; ry < ty * 16 + 16 && ry < j->nrows; ry++) {
            uint32_t y = p->row_start + ry * p->row_step;
            for (uint32_t x = tx * 16; x < tx * 16 + 16 && x < p->width; x++) {
                uint32_t idx = y * p->width + x;
                float s32[3] = {0, 0, 0};
                double s64[3] = {0, 0, 0}, sf64[3] = {0, 0, 0};
                if (j->f64) {
                    double d[3];
                    pixel_f64(&c64, p, x, y, sf64, d, st);
                } else {
                    pixel_f32(&c32, p, x, y, s32, s64, st);
                }
                finish_pixel(j, idx, s32, s64, sf64);
            }
        }
    }
    return NULL;
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void run_job(job_t* j, vzo_stats* out)
{
    const vzo_render_params* p = j->p;
    uint32_t step = p->row_step ? p->row_step : 1;
    j->nrows = p->row_start < p->height ? (p->height - p->row_start + step - 1) / step : 0;
    j->ntiles_x = (p->width + 15) / 16;
    j->ntiles_y = (j->nrows + 15) / 16;
    j->next = 0;
    int nt = (int)p->threads;
    if (nt <= 0) { /* the CPUs this process may run on, not every CPU of the host */
        cpu_set_t cs;
        nt = sched_getaffinity(0, sizeof(cs), &cs) == 0 ? CPU_COUNT(&cs) : (int)sysconf(_SC_NPROCESSORS_ONLN);
    }
    if (nt < 1) nt = 1;
    if (nt > 256) nt = 256;
    memset(j->st, 0, sizeof(j->st));
    pthread_t th[256];
    worker_t w[256];
    double t0 = now_s();
    for (int i = 0; i < nt; i++) {
        w[i].j = j;
        w[i].id = i;
        if (i > 0) pthread_create(&th[i], NULL, worker, &w[i]);
    }
    worker(&w[0]);
    for (int i = 1; i < nt; i++) pthread_join(th[i], NULL);
    double t1 = now_s();
    if (out) {
        memset(out, 0, sizeof(*out));
        for (int i = 0; i < nt; i++) {
            out->segments += j->st[i].segments;
            out->paths += j->st[i].paths;
            out->hits += j->st[i].hits;
            out->node_visits += j->st[i].node_visits;
            out->sphere_tests += j->st[i].sphere_tests;
        }
        out->seconds = t1 - t0;
        out->threads = (uint32_t)nt;
    }
}

static int params_ok(const vzo_render_params* p)
{
    return p->width >= 2 && p->height >= 2 && p->spp >= 1 && p->max_depth >= 1;
}

void vzo_render(const vzo_bvh* b, const vzo_sphere* s, const vzo_material* m, uint32_t nsph,
                const vzo_render_params* p, void* accum, uint8_t* image, vzo_stats* st)
{
    if (!params_ok(p)) return;
    job_t* j = (job_t*)calloc(1, sizeof(job_t));
    j->b = b;
    j->s = s;
    j->m = m;
    j->nsph = nsph;
    j->p = p;
    j->accum = accum;
    j->image = image;
    j->f64 = 0;
    run_job(j, st);
    free(j);
}

void vzo_render_f64(const vzo_bvh* b, const vzo_sphere* s, const vzo_material* m, uint32_t nsph,
                    const vzo_render_params* p, double* accum, vzo_stats* st)
{
    if (!params_ok(p) || !accum) return;
    job_t* j = (job_t*)calloc(1, sizeof(job_t));
    j->b = b;
    j->s = s;
    j->m = m;
    j->nsph = nsph;
    j->p = p;
    j->accum = accum;
    j->image = NULL;
    j->f64 = 1;
    run_job(j, st);
    free(j);
}

void vzo_make_rayset(const vzo_bvh* b, const vzo_sphere* s, const vzo_material* m, uint32_t nsph,
                     const vzo_render_params* p, uint32_t n_int, uint32_t n_iso, uint32_t iso_seed,
                     vzo_ray* out)
{
    pctx_f32 c = {b, s, m, nsph, p->max_depth, p->binary_tree, out, 0, n_int};
    vzo_stats st;
    memset(&st, 0, sizeof(st));
    const uint32_t step = p->row_step ? p->row_step : 1u;
    for (uint32_t y = p->row_start; y < p->height && c.rec_n < n_int; y += step)
        for (uint32_t x = 0; x < p->width && c.rec_n < n_int; x++) {
            float s32[3];
            double s64[3];
            pixel_f32(&c, p, x, y, s32, s64, &st);
        }
    /* isotropic rays from points in the AABB of spheres 1..nsph-1 */
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (uint32_t i = (nsph > 1 ? 1 : 0); i < nsph; i++) {
        float bx[6];
        leaf_box(&s[i], bx);
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], bx[k]);
            hi[k] = fmaxf(hi[k], bx[3 + k]);
        }
    }
    vzo_mt g;
    mt_seed(&g, iso_seed);
    for (uint32_t i = 0; i < n_iso; i++) {
        vzo_ray* r = &out[n_int + i];
        r->ox = lo[0] + (hi[0] - lo[0]) * mt_uniform(&g);
        r->oy = lo[1] + (hi[1] - lo[1]) * mt_uniform(&g);
        r->oz = lo[2] + (hi[2] - lo[2]) * mt_uniform(&g);
        float d[3];
        for (;;) {
            d[0] = -1.0f + 2.0f * mt_uniform(&g);
            d[1] = -1.0f + 2.0f * mt_uniform(&g);
            d[2] = -1.0f + 2.0f * mt_uniform(&g);
            float l2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (l2 < 1.0f && l2 > 1e-6f) break;
        }
        n3(d);
        r->dx = d[0];
        r->dy = d[1];
        r->dz = d[2];
        r->tmin = 0.001f;
        r->tmax = 1e16f;
    }
}
