// vz_traverse.h -- closest-hit traversal of the BVH on gfx950, replacing
// the traversal inside optixTrace (Core/RayTracer.cu:190-202) and the
// __intersection__hit_sphere program it calls (Core/RayTracer.cu:229-270).
//
// The tree walked is the 4-wide one collapsed from the binary LBVH: 64-byte
// nodes holding four quantised child boxes, so a visit is one node fetch
// (four 16-byte loads that lanes on the same node share) and four box tests
// that every lane runs together.  Per lane: the ray in registers, a short
// stack of (node, entry distance) in LDS laid out [slot][lane] so a wave's
// pushes and pops hit distinct banks, and a private overflow array that holds
// the oldest entries when the LDS ring is full.  Per visit: the four slot
// boxes against [tmin, best]; the leaf slots hit, tested in slot order; then
// the internal slots hit whose entry distance is still within the best hit,
// nearest first (lower slot on a tie): the nearest is visited next and the
// others pushed farthest first.  Popped entries beyond the best hit are
// dropped.  Ties between hits resolve to the smaller t, then the smaller
// primitive id, so the result is the exact lexicographic minimum over all
// spheres whatever the visit order (the CPU oracle checks this against brute
// force and follows the same visit order, so node-visit and sphere-test
// counts match it too).
#pragma once
#include "vz_device.h"
#include "vz_internal.h"

namespace vz {

constexpr float kRobust = 1.0f + 0x1p-20f;  // >= 1 + 2*gamma_3 (Ize 2013)

struct TraceResult {
    float t;
    uint32_t leaf;  // leaf (Morton-order) index, 0xFFFFFFFF on a miss
};

// Where the 64-byte nodes come from: the scene's array in HBM, or a copy a
// workgroup staged in LDS (a small scene rendered alone).
struct LdsNodes {
    const uint4* p;  // points into __shared__ memory
    VZ_DEV void fetch(uint32_t i, uint4& a, uint4& b, uint4& c, uint4& d) const
    {
        a = p[4 * i];
        b = p[4 * i + 1];
        c = p[4 * i + 2];
        d = p[4 * i + 3];
    }
};
struct GlobalNodes {
    const uint4* p;
    VZ_DEV void fetch(uint32_t i, uint4& a, uint4& b, uint4& c, uint4& d) const
    {
        const uint4* q = p + 4 * (size_t)i;
        a = ldg(q);
        b = ldg(q + 1);
        c = ldg(q + 2);
        d = ldg(q + 3);
    }
};

// The oldest stack entries of a lane once its LDS ring is full.  Kept apart
// from Traversal so the traversal's scalars stay in registers: the compiler
// only splits an aggregate into registers when nothing in it is indexed
// dynamically.
struct Overflow {
    uint32_t node[kWideOverflow];
    float t[kWideOverflow];
};

// Resumable traversal of one ray: init() sets it up, every step() visits
// one wide node.  A wave can therefore refill a lane with a new ray as soon
// as its traversal ends instead of waiting for the slowest lane.
// S: LDS slots per lane (power of two); lds_node/lds_t point at this lane's
// column, consecutive slots `stride` words apart.
// DEFER: leaf slots that a visit hits are queued (the lane's column of
// lds_pend, at most kPend entries) instead of tested at once, and flush()
// tests them; the caller flushes when many lanes of the wave hold some, so a
// wave's sphere tests run with more lanes at a time.  Culling meanwhile uses
// the best hit found so far, which is no smaller than it would otherwise be:
// no node that holds the closest hit is culled, and the closest hit (the
// least (t, prim) over the spheres tested) does not depend on the order of
// the tests.  Node-visit and sphere-test counts do depend on when the wave
// flushes, so the counting kernels test at once.
constexpr uint32_t kPend = 8;
template <int S, bool COUNT, bool DEFER = false>
struct Traversal {
    v3 o, d;
    float inv[3];
    float tmin, best;
    uint32_t negm;  // bit k: the ray runs towards -k, so a slot's near plane on axis k is its hi plane
    uint32_t bleaf, bprim, node, sp, bottom;
    uint32_t npend = 0;  // DEFER: queued leaf tests (0 also for a lane that never gets a ray)

    // Returns true when the ray is already finished (single-sphere scene).
    VZ_DEV bool init(const SceneView& sc, v3 o_, v3 d_, float tmin_, float tmax_, uint32_t& ns)
    {
        o = o_;
        d = d_;
        tmin = tmin_;
        best = tmax_;
        bleaf = 0xFFFFFFFFu;
        bprim = 0xFFFFFFFFu;
        npend = 0;
        const float dd[3] = {d.x, d.y, d.z};
        negm = 0u;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float dk = dd[k];
            if (!(fabsf(dk) > 1e-20f)) dk = copysignf(1e-20f, dk);
            inv[k] = recip_rn(dk);
            // lo <= hi and rounding is monotone, so (lo - o) * inv <= (hi - o) * inv
            // exactly when inv > 0: the near plane is lo for a positive
            // direction, hi for a negative one
            negm |= (inv[k] < 0.0f ? 1u : 0u) << k;
        }
        if (sc.n == 1u) {
            float t;
            if (COUNT) ns++;
            if (hit_gated(ldg(sc.spheres), t)) {
                best = t;
                bleaf = 0;
            }
            return true;
        }
        sp = 0;
        bottom = 0;
        node = 0;
        // the build checks every child index and that each node has one
        // parent, so a walk from the root ends: no guards per visit
        return false;
    }

    // The sphere test (root in [tmin, best]) gated by the sphere's own box:
    // the root counts only if the ray enters [c - |r|, c + |r|] by t, by the
    // same robust slab test as the nodes.  OptiX runs an intersection program
    // only for primitives whose box the ray enters; the gate also drops the
    // spurious grazing roots that float rounding of |oc|^2 - r^2 gives tiny
    // distant spheres, and it makes the closest hit independent of the tree:
    // every node box contains its leaves' boxes, so it passes whenever the
    // gate does, and any BVH reaches every gated root.
    VZ_DEV bool hit_gated(float4 s, float& t) const
    {
        if (!hit_sphere(o, d, tmin, best, s, t)) return false;
        const float r = fabsf(s.w);
        const float c[3] = {s.x, s.y, s.z}, org[3] = {o.x, o.y, o.z};
        float lo_t = tmin, hi_t = t;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            // the near plane is c - r for a positive inv, c + r for a negative
            // one: with sr = copysign(r, inv), c - sr and c + sr are exactly
            // those planes (x - (-y) rounds as x + y), and rounding is
            // monotone, so these are min(t0, t1) and max(t0, t1) of the two
            // plane distances bit for bit
            const float sr = copysignf(r, inv[k]);
            lo_t = fmaxf(lo_t, ((c[k] - sr) - org[k]) * inv[k]);
            hi_t = fminf(hi_t, ((c[k] + sr) - org[k]) * inv[k]);
        }
        return lo_t <= hi_t * kRobust;
    }

    // Sphere test of leaf `leaf` (sphere s); keeps the lexicographic minimum
    // of (t, prim).
    VZ_DEV void test_leaf(const SceneView& sc, uint32_t leaf, float4 s, uint32_t& ns)
    {
        float t;
        if (COUNT) ns++;
        if (hit_gated(s, t)) {
            if (t < best) {
                best = t;
                bleaf = leaf;
                bprim = 0xFFFFFFFFu;  // looked up lazily on a tie
            } else {
                const uint32_t pid = ldg(sc.leaf_info + leaf) & 0x3FFFFFFFu;
                if (bprim == 0xFFFFFFFFu && bleaf != 0xFFFFFFFFu) bprim = ldg(sc.leaf_info + bleaf) & 0x3FFFFFFFu;
                if (pid < bprim) {
                    best = t;
                    bleaf = leaf;
                    bprim = pid;
                }
            }
        }
    }

    // Pops the nearest stacked node that can still hold a closer hit.
    // Returns true when none is left (the traversal has finished).
    VZ_DEV bool pop(uint32_t* lds_node, float* lds_t, uint32_t stride, const Overflow& ov)
    {
        while (sp > 0u) {
            sp--;
            uint32_t nd;
            float t;
            if (sp < bottom) {
                nd = ov.node[sp];
                t = ov.t[sp];
                bottom = sp;
            } else {
                nd = lds_node[(sp & (S - 1)) * stride];
                t = lds_t[(sp & (S - 1)) * stride];
            }
            if (t <= best * kRobust) {
                node = nd;
                return false;
            }
        }
        return true;
    }

    // DEFER: tests the queued leaves (the lane's column of lds_pend).
    VZ_DEV void flush(const SceneView& sc, const uint32_t* lds_pend, uint32_t stride, uint32_t& ns)
    {
        while (npend) {
            npend--;
            const uint32_t leaf = lds_pend[npend * stride];
            test_leaf(sc, leaf, ldg(sc.spheres + leaf), ns);
        }
    }

    // One wide node visit; returns true when the traversal has finished.
    template <class NodeSrc>
    VZ_DEV bool step(const SceneView& sc, const NodeSrc& nodes, uint32_t* lds_node, float* lds_t,
                     uint32_t stride, Overflow& ov, uint32_t& nv, uint32_t& ns, uint32_t* lds_pend = nullptr)
    {
        uint4 A, B, C, D;
        nodes.fetch(node, A, B, C, D);
        return visit(sc, A, B, C, D, lds_node, lds_t, stride, ov, nv, ns, lds_pend);
    }

    // The visit of the current node, whose 64 bytes are A..D.
    VZ_DEV bool visit(const SceneView& sc, const uint4& A, const uint4& B, const uint4& C, const uint4& D,
                      uint32_t* lds_node, float* lds_t, uint32_t stride, Overflow& ov, uint32_t& nv, uint32_t& ns,
                      uint32_t* lds_pend)
    {
        if (COUNT) nv++;
        const float no[3] = {__uint_as_float(A.x), __uint_as_float(A.y), __uint_as_float(A.z)};
        const float org[3] = {o.x, o.y, o.z};
        const uint32_t meta = A.w;
        const uint32_t lo_w[3] = {B.x, B.y, B.z}, hi_w[3] = {B.w, C.x, C.y};
        const uint32_t ch[4] = {C.z, C.w, D.x, D.y};
        // Slab distances of the slots' planes s*q + no (q a plane byte) as
        // one fma per plane, fma(q, s*inv, (no - o)*inv), moved outwards by
        // e = (255*s + |no| + |no - o|) * |inv| * 2^-19 (near planes) and 2e
        // (far planes).  The exact-form distance fl(fl(fl(s*q + no) - o) * inv)
        // differs from the unmoved fma value by less than 10 u (255*s + |no| +
        // |no - o|) |inv| (u = 2^-24), so every near distance here is <= the
        // exact form's and every far distance >= the exact form's times
        // kRobust: the test passes whenever the exact-form robust test does,
        // which passes whenever a contained sphere's gate does (monotone
        // rounding), so hits stay independent of the tree.
        float a[3], bn[3], bf[3];
        uint32_t qn[3], qf[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float s = scale_of(meta, k);
            const float dd = no[k] - org[k];
            const float b = dd * inv[k];
            const float w = (__builtin_fmaf(255.0f, s, fabsf(no[k])) + fabsf(dd)) * fabsf(inv[k]);
            const float e = fmaxf(w * 0x1p-19f, 0x1p-126f);
            a[k] = s * inv[k];
            bn[k] = b - e;
            bf[k] = __builtin_fmaf(2.0f, e, b);
            // the near plane is the lo plane for a positive direction: a
            // whole-word select per axis instead of a min and max per slot
            const bool ng = ((negm >> k) & 1u) != 0u;
            qn[k] = ng ? hi_w[k] : lo_w[k];
            qf[k] = ng ? lo_w[k] : hi_w[k];
        }
        const float lim_t = best * kRobust;
        float tn[4];
        uint32_t hitm = 0;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            float lo_t = tmin, hi_t = lim_t;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                lo_t = fmaxf(lo_t, __builtin_fmaf((float)((qn[k] >> (8 * c)) & 0xFFu), a[k], bn[k]));
                hi_t = fminf(hi_t, __builtin_fmaf((float)((qf[k] >> (8 * c)) & 0xFFu), a[k], bf[k]));
            }
            tn[c] = lo_t;
            hitm |= (lo_t <= hi_t ? 1u : 0u) << c;
        }
        hitm &= meta >> 28;
        const uint32_t leafm = (meta >> 24) & 0xFu;
        // leaf slots that were hit, in slot order
        uint32_t lm = hitm & leafm;
        while (lm) {
            const uint32_t c = (uint32_t)__builtin_ctz(lm);
            const uint32_t lo = (c & 1u) ? ch[1] : ch[0], hi = (c & 1u) ? ch[3] : ch[2];
            const uint32_t leaf = (c & 2u) ? hi : lo;  // selects, not branches
            if (DEFER) {
                lds_pend[npend * stride] = leaf;
                npend++;
            } else {
                test_leaf(sc, leaf, ldg(sc.spheres + leaf), ns);
            }
            lm &= lm - 1u;
        }
        // internal slots that were hit and can still hold a closer hit
        const float lim = best * kRobust;
        const uint32_t im = hitm & ~leafm;
        bool cand[4];
        uint32_t m = 0;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            cand[c] = ((im >> c) & 1u) != 0u && tn[c] <= lim;
            m += cand[c] ? 1u : 0u;
        }
        if (m == 0u) return pop(lds_node, lds_t, stride, ov);
        // rank of each candidate in (entry distance, slot) order; the other
        // slots get an infinite key, which only ever raises their own ranks
        // (a candidate's key is at most best * kRobust, finite)
        float key[4];
#pragma unroll
        for (int c = 0; c < 4; c++) key[c] = cand[c] ? tn[c] : __builtin_huge_valf();
        // per pair a < b one comparison: key[a] <= key[b] ranks b after a,
        // otherwise a after b
        const uint32_t l01 = key[0] <= key[1] ? 1u : 0u, l02 = key[0] <= key[2] ? 1u : 0u;
        const uint32_t l03 = key[0] <= key[3] ? 1u : 0u, l12 = key[1] <= key[2] ? 1u : 0u;
        const uint32_t l13 = key[1] <= key[3] ? 1u : 0u, l23 = key[2] <= key[3] ? 1u : 0u;
        const uint32_t rk[4] = {3u - (l01 + l02 + l03), 2u + l01 - (l12 + l13), 1u + l02 + l12 - l23,
                                l03 + l13 + l23};
        uint32_t next = 0u;
#pragma unroll
        for (int c = 0; c < 4; c++)
            if (cand[c] && rk[c] == 0u) next = ch[c];
        if (m > 1u) {
            // room in the LDS ring for m - 1 entries: the oldest go to the overflow
            while (sp + (m - 1u) - bottom > (uint32_t)S) {
                if (bottom >= kWideOverflow) return true;
                ov.node[bottom] = lds_node[(bottom & (S - 1)) * stride];
                ov.t[bottom] = lds_t[(bottom & (S - 1)) * stride];
                bottom++;
            }
            // the farthest lands deepest: rank r at sp + (m - 1 - r)
#pragma unroll
            for (int c = 0; c < 4; c++) {
                if (cand[c] && rk[c] != 0u) {
                    const uint32_t pos = sp + (m - 1u) - rk[c];
                    lds_node[(pos & (S - 1)) * stride] = ch[c];
                    lds_t[(pos & (S - 1)) * stride] = tn[c];
                }
            }
            sp += m - 1u;
        }
        node = next;
        return false;
    }
};

// Whole-ray traversal: init, then step until done.
template <int S, bool COUNT, class NodeSrc>
VZ_DEV TraceResult closest_hit(const SceneView& sc, const NodeSrc& nodes, v3 o, v3 d, float tmin,
                               float tmax, uint32_t* lds_node, float* lds_t, uint32_t stride,
                               uint32_t& nv, uint32_t& ns)
{
    Traversal<S, COUNT> tr;
    Overflow ov;
    if (!tr.init(sc, o, d, tmin, tmax, ns))
        while (!tr.step(sc, nodes, lds_node, lds_t, stride, ov, nv, ns)) {
        }
    return TraceResult{tr.best, tr.bleaf};
}

}  // namespace vz
