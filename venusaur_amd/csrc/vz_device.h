// vz_device.h -- device-side math of the path tracer (gfx950).
//
// Every expression is evaluated in the order written (the library is built
// with -ffp-contract=off and IEEE-rounded divide/sqrt), so the GPU computes
// the same float results as the CPU integrator that checks it.  Reference
// semantics followed:
//   RNG ............ Core/random.cuh:31-67
//   vec maths ...... Core/vec_math.h:412-570
//   samplers ....... Core/RayTracer.cu:87-149
//   camera ray ..... Core/RayTracer.cu:151-177
//   sphere test .... Core/RayTracer.cu:229-270 (compensated discriminant,
//                    stable roots)
//   materials ...... Core/RayTracer.cu:272-440
//   miss ........... Core/RayTracer.cu:442-450
//   make_color ..... Core/RayTracer.cu:16-47
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define VZ_DEV __device__ __forceinline__

namespace vz {

struct v3 {
    float x, y, z;
};

// Loads through a global-address-space pointer.  Pointers read out of a
// struct in memory are generic, so the compiler would emit a flat load, whose
// wait also covers LDS traffic; every buffer this library hands its kernels is
// in device (global) memory.
typedef float raw_float4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const raw_float4 global_float4;
VZ_DEV float4 ldg(const float4* p)
{
    const raw_float4 v = *__builtin_bit_cast(global_float4*, p);  // a C-style cast would be dropped
    return make_float4(v.x, v.y, v.z, v.w);
}
typedef unsigned int raw_uint4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const raw_uint4 global_uint4;
VZ_DEV uint4 ldg(const uint4* p)
{
    const raw_uint4 v = *__builtin_bit_cast(global_uint4*, p);
    return make_uint4(v.x, v.y, v.z, v.w);
}
// Two consecutive float4 in one load (16-byte aligned), so both halves are in
// flight before the first wait.
typedef float raw_float8 __attribute__((ext_vector_type(8), aligned(16)));
typedef __attribute__((address_space(1))) const raw_float8 global_float8;
VZ_DEV void ldg2(const float4* p, float4& a, float4& b)
{
    const raw_float8 v = *__builtin_bit_cast(global_float8*, p);
    a = make_float4(v.s0, v.s1, v.s2, v.s3);
    b = make_float4(v.s4, v.s5, v.s6, v.s7);
}
VZ_DEV uint32_t ldg(const uint32_t* p)
{
    return *__builtin_bit_cast(__attribute__((address_space(1))) const uint32_t*, p);
}

VZ_DEV v3 mk(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
VZ_DEV v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
VZ_DEV v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
VZ_DEV v3 mul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
VZ_DEV v3 scl(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
VZ_DEV v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
VZ_DEV float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// 1/x rounded to nearest even.  For |x| in [2^-100, 2^100) the hardware
// reciprocal (v_rcp_f32, within 1 ulp) refined by one Newton-Raphson step in
// fma (Markstein's method: r + (1 - x r) r) equals the IEEE quotient
// 1.0f / x: checked on gfx950 for every float in that range (3,355,443,200
// values, both signs; the unrefined v_rcp_f32 differs in 359,070,000 of
// them).  Inputs, the result and the step's terms are normal there, so the
// denormal mode plays no part.  Elsewhere (zero, denormals, huge values,
// inf, nan) the IEEE division runs.  Four VALU instructions instead of the
// division's ten.
VZ_DEV float recip_rn(float x)
{
    const float ax = fabsf(x);
    if (ax >= 0x1p-100f && ax < 0x1p100f) {
        const float r = __builtin_amdgcn_rcpf(x);
        return __builtin_fmaf(__builtin_fmaf(-x, r, 1.0f), r, r);
    }
    return 1.0f / x;
}
VZ_DEV v3 normalize(v3 v) { float inv = recip_rn(sqrtf(dot(v, v))); return scl(v, inv); }
VZ_DEV v3 reflect(v3 i, v3 n) { return sub(i, scl(scl(n, 2.0f), dot(n, i))); }
VZ_DEV v3 refract(v3 uv, v3 n, float eta)
{
    float ct = fminf(dot(neg(uv), n), 1.0f);
    v3 perp = scl(add(uv, scl(n, ct)), eta);
    float k = -sqrtf(fabsf(1.0f - dot(perp, perp)));
    return add(perp, scl(n, k));
}

// ---- RNG (random.cuh:31-67) -------------------------------------------
VZ_DEV uint32_t tea4(uint32_t v0, uint32_t v1)
{
    uint32_t s0 = 0;
#pragma unroll
    for (int n = 0; n < 4; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}
// The LCG s <- 1664525 s + 1013904223 (mod 2^32), tracked mod 2^24: a draw
// reads only the state's low 24 bits, and those depend only on the low 24
// bits of the previous state (1664525 < 2^24), so the full-rate 24-bit
// multiply-add v_mad_u32_u24 gives every draw of the 32-bit recurrence bit
// for bit, where a 32-bit multiply (v_mul_lo_u32) issues at a quarter of the
// rate.  The high 8 bits of s are not the recurrence's and nothing reads
// them.  Written as one VALU instruction because the compiler, seeing that
// only the low 24 bits are used, turns a 24-bit multiply back into a 32-bit
// one; VOP3 takes no literal here, so the multiplier comes from an SGPR.
VZ_DEV uint32_t lcg_next(uint32_t s)
{
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(s), "s"(1664525u), "v"(1013904223u));
    return r;
}
VZ_DEV float rnd(uint32_t& s)
{
    s = lcg_next(s);
    return (float)(s & 0x00FFFFFFu) * (1.0f / 16777216.0f);  // exact: 24-bit int * 2^-24
}
// -1 + 2 * rnd(s): q * 2^-24 and its double are exact, so the one rounding
// of the sum is the fma's
VZ_DEV float rnd_m11(uint32_t& s)
{
    s = lcg_next(s);
    return __builtin_fmaf((float)(s & 0x00FFFFFFu), 0x1p-23f, -1.0f);
}

VZ_DEV v3 in_unit_sphere(uint32_t& s)
{
    for (;;) {
        float x = rnd_m11(s);
        float y = rnd_m11(s);
        float z = rnd_m11(s);
        v3 p = mk(x, y, z);
        if (dot(p, p) >= 1.0f) continue;
        return p;
    }
}

// |x| < 1e-8 in double  <=>  |x| <= 1e-8f in float (1e-8f rounds below 1e-8)
VZ_DEV bool near_zero(v3 e)
{
    return (fabsf(e.x) <= 1e-8f) && (fabsf(e.y) <= 1e-8f) && (fabsf(e.z) <= 1e-8f);
}

// ---- sphere test ---------------------------------------------------------
// Kahan's FMA difference of products for hb*hb - a*cc, stable root pair
// q/a and cc/q, nearer root in [tmin, tmax] else the farther one.
VZ_DEV bool hit_sphere(v3 o, v3 d, float tmin, float tmax, float4 sph, float& tout)
{
    v3 oc = sub(o, mk(sph.x, sph.y, sph.z));
    float a = dot(d, d);
    float hb = dot(oc, d);
    float cc = dot(oc, oc) - sph.w * sph.w;
    float w = a * cc;
    float e = __builtin_fmaf(-a, cc, w);
    float f = __builtin_fmaf(hb, hb, -w);
    float disc = f + e;
    if (disc < 0.0f) return false;
    float sq = sqrtf(disc);
    float q = -(hb + copysignf(sq, hb));
    if (q == 0.0f) return false;
    // The roots are q/a and cc/q.  With hb < 0 (sign bit set) q > 0 and cc/q
    // is the nearer; otherwise q < 0 and the nearer, q/a, is <= 0, which a
    // positive tmin rejects without dividing.  The farther root is divided
    // out only when the nearer one is rejected.
    float t;
    if (__builtin_signbit(hb)) {
        t = cc / q;
        if (t < tmin || tmax < t) {
            t = q / a;
            if (t < tmin || tmax < t) return false;
        }
    } else {
        bool near_ok = false;
        if (!(tmin > 0.0f)) {
            t = q / a;
            near_ok = !(t < tmin || tmax < t);
        }
        if (!near_ok) {
            t = cc / q;
            if (t < tmin || tmax < t) return false;
        }
    }
    tout = t;
    return true;
}

// ---- miss (RayTracer.cu:442-450) -----------------------------------------
VZ_DEV v3 sky(v3 d)
{
    v3 u = normalize(d);
    float t = 0.5f * (u.y + 1.0f);
    return mk(1.0f + t * (0.5f - 1.0f), 1.0f + t * (0.7f - 1.0f), 1.0f + t * (1.0f - 1.0f));
}

// ---- materials (RayTracer.cu:272-440) --------------------------------------
// mat = {albedo rgb, fuzz (metal) or ior (dielectric)}.  Returns false when
// the path is absorbed.  att is the throughput multiplier.
VZ_DEV bool scatter(uint32_t type, float4 mat, v3 d, v3 n, bool front, uint32_t& seed, v3& nd,
                    v3& att)
{
    // Lambertian and metal both draw a point in the unit sphere (metal's
    // reflection draws nothing) and metal and dielectric both normalise d:
    // each is done once for the lanes that need it, so a wave holding several
    // materials runs one rejection loop and one normalisation instead of two
    v3 p = mk(0.0f, 0.0f, 0.0f);
    if (type != 2u) p = in_unit_sphere(seed);
    if (type == 0u) {
        v3 sd = add(n, normalize(p));
        if (near_zero(sd)) sd = n;
        nd = sd;
        att = mk(mat.x, mat.y, mat.z);
        return true;
    }
    const v3 u = normalize(d);
    if (type == 1u) {
        v3 sd = add(reflect(u, n), scl(p, mat.w));
        if (dot(sd, n) > 0.0f) {
            nd = sd;
            att = mk(mat.x, mat.y, mat.z);
            return true;
        }
        return false;
    } else {
        float ir = mat.w;
        float ratio = front ? recip_rn(ir) : ir;
        float ct = fminf(dot(neg(u), n), 1.0f);
        float st = sqrtf(1.0f - ct * ct);
        bool cannot = ratio * st > 1.0f;
        bool refl = cannot;
        if (!cannot) {
            float r0 = (1.0f - ratio) / (1.0f + ratio);
            r0 = r0 * r0;
            float x = 1.0f - ct;
            float x2 = x * x;
            float x5 = (x2 * x2) * x;
            float sch = r0 + (1.0f - r0) * x5;
            refl = sch > rnd(seed);
        }
        nd = refl ? reflect(u, n) : refract(u, n, ratio);
        att = mk(1.0f, 1.0f, 1.0f);
        return true;
    }
}

// ---- camera (RayTracer.cu:141-177) ---------------------------------------
struct Cam {
    v3 O, U, V, W, un, vn;
    float lr;
    uint32_t width, height;
};

// Draws jitter and the lens sample from seed (advancing it) and returns the
// camera ray of pixel (x, y).
VZ_DEV void camera_ray(const Cam& c, uint32_t x, uint32_t y, uint32_t& seed, v3& org, v3& dir)
{
    float jx = rnd(seed);
    float jy = rnd(seed);
    float u = 2.0f * ((float)x + jx) / (float)(c.width - 1) - 1.0f;
    float v = 2.0f * ((float)y + jy) / (float)(c.height - 1) - 1.0f;
    float dx, dy;
    for (;;) {
        dx = rnd_m11(seed);
        dy = rnd_m11(seed);
        if (dx * dx + dy * dy + 0.0f * 0.0f >= 1.0f) continue;
        break;
    }
    v3 rd = mk(dx * c.lr, dy * c.lr, 0.0f * c.lr);
    v3 off = add(scl(c.un, rd.x), scl(c.vn, rd.y));
    org = add(c.O, off);
    dir = sub(add(add(c.W, scl(scl(c.U, u), 0.5f)), scl(scl(c.V, v), 0.5f)), off);
}

// ---- make_color (RayTracer.cu:16-47) -------------------------------------
VZ_DEV uint32_t encode8(float c)
{
    c = fmaxf(0.0f, fminf(c, 1.0f));
    float s = c < 0.0031308f ? 12.92f * c : 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
    s = fmaxf(0.0f, fminf(s, 1.0f));
    uint32_t q = (uint32_t)(s * 256.0f);
    return q < 255u ? q : 255u;
}
VZ_DEV uint32_t make_color(float r, float g, float b)
{
    return encode8(r) | (encode8(g) << 8) | (encode8(b) << 16) | (255u << 24);
}

// ---- LBVH node (32 bytes) --------------------------------------------------
// word0..2: float origin of the node's quantisation grid (its own box min)
// word3   : biased exponents ex | ey<<8 | ez<<16 | leaf flags<<24
// word4..6: per axis, bytes lo0 | hi0<<8 | lo1<<16 | hi1<<24 (child boxes
//           on the grid origin + 2^(e-127) * q)
// word7   : first child index g; children are g and g+1, each a leaf index
//           (flag set) or an internal node index
VZ_DEV float scale_of(uint32_t meta, int k)
{
    return __uint_as_float(((meta >> (8 * k)) & 0xFFu) << 23);
}

}  // namespace vz
