// vz_traverse_bin.h -- closest-hit traversal of the binary LBVH, for a
// scene rendered alone whose binary tree (at most 512 internal nodes, 16 KB)
// a workgroup stages in LDS.  There, with node fetches as cheap as LDS
// reads, a binary visit's two box tests beat the 4-wide tree's four: on the
// 1080p book scene the 4-wide tree cut node visits per segment from 12.0 to
// 7.3 but ran 4% slower (10844 against 11313 Mray-segments/s), while on
// scenes whose nodes come from memory it is faster (vz_traverse.h).  Same
// semantics and tie rules as the 4-wide traversal: the closest hit is the
// exact lexicographic minimum of (t, prim), whatever the tree.
//
// Per lane: the ray in registers, a short stack of (node, tnear) in LDS laid
// out [slot][lane], and a private overflow array for the oldest entries.
// Near-first order, leaf children tested as soon as their box is hit, popped
// entries culled against the current best hit.
#pragma once
#include "vz_traverse.h"

namespace vz {

// The 32-byte binary nodes, staged in LDS by the workgroup.  kLeafLoop: a
// visit's two leaf children are tested by one copy of the sphere test run
// once per pending leaf (see BinTraversal::leaves), the faster form with LDS
// nodes on the 1080p book scene.
struct LdsBinNodes {
    static constexpr bool kLeafLoop = true;
    const uint4* p;  // points into __shared__ memory
    VZ_DEV void fetch(uint32_t i, uint4& a, uint4& b) const
    {
        a = p[2 * i];
        b = p[2 * i + 1];
    }
};

// The oldest stack entries of a lane once its LDS ring is full.  Kept apart
// from Traversal so the traversal's scalars stay in registers: the compiler
// only splits an aggregate into registers when nothing in it is indexed
// dynamically.
struct BinOverflow {
    uint32_t node[kMaxDepth];
    float t[kMaxDepth];
};

// Resumable traversal of one ray: init() sets it up, every step() visits
// one internal node.  A wave can therefore refill a lane with a new ray as
// soon as its traversal ends instead of waiting for the slowest lane.
// S: LDS slots per lane (power of two); lds_node/lds_t point at this lane's
// column, consecutive slots `stride` words apart.
template <int S, bool COUNT>
struct BinTraversal {
    v3 o, d;
    float inv[3];
    uint32_t sel[3];  // per axis: byte selector that puts the near plane first
    float tmin, best;
    uint32_t bleaf, bprim, node, sp, bottom, budget;
    uint32_t pend, pg;  // the visit waiting for leaves() / decide() (see visit())
    float ptn0, ptn1;

    // Returns true when the ray is already finished (single-sphere scene).
    VZ_DEV bool init(const SceneView& sc, v3 o_, v3 d_, float tmin_, float tmax_, uint32_t& ns)
    {
        o = o_;
        d = d_;
        tmin = tmin_;
        best = tmax_;
        bleaf = 0xFFFFFFFFu;
        bprim = 0xFFFFFFFFu;
        const float dd[3] = {d.x, d.y, d.z};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float dk = dd[k];
            if (!(fabsf(dk) > 1e-20f)) dk = copysignf(1e-20f, dk);
            inv[k] = 1.0f / dk;
            // lo <= hi and rounding is monotone, so (lo - o) * inv <= (hi - o) * inv
            // exactly when inv > 0: the near plane of each child is its lo
            // byte for a positive direction, its hi byte for a negative one
            sel[k] = inv[k] < 0.0f ? 0x02030001u : 0x03020100u;
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
        // a well-formed tree visits each of its n-1 internal nodes at most
        // once; this and the index guards below only matter for a corrupted
        // node array
        budget = sc.n;
        pend = 0u;
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
            const float t0 = ((c[k] - r) - org[k]) * inv[k];
            const float t1 = ((c[k] + r) - org[k]) * inv[k];
            lo_t = fmaxf(lo_t, fminf(t0, t1));
            hi_t = fminf(hi_t, fmaxf(t0, t1));
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

    // A node visit in three parts, so that a wave can run the sphere tests of
    // many lanes together (Aila & Laine 2009, "while-while"): visit() does the
    // box tests and records what the node leads to; leaves() tests its leaf
    // children; decide() picks the next node with the best hit as it stands
    // after those tests.  Each lane runs the same sequence as step(), so the
    // result and every count are the same whenever leaves() runs.
    //
    // visit(): returns true when the traversal has finished (a guard fired).
    // Otherwise pend holds the leaf children to test (bits 0-1) and the
    // internal children to follow (bits 2-3); pg is the node's first child
    // and ptn the two children's entry distances.
    template <class NodeSrc>
    VZ_DEV bool visit(const SceneView& sc, const NodeSrc& nodes, uint32_t& nv)
    {
        uint4 A, B;
        nodes.fetch(node, A, B);
        if (COUNT) nv++;
        const float no[3] = {__uint_as_float(A.x), __uint_as_float(A.y), __uint_as_float(A.z)};
        const float org[3] = {o.x, o.y, o.z};
        const uint32_t meta = A.w;
        const uint32_t qw[3] = {B.x, B.y, B.z};
        const uint32_t g = B.w;
        float tn[2];
        bool hit[2];
        // per axis, the two children's (lo, hi) bytes reordered to (near, far):
        // one byte permute replaces a min and a max per child and axis, with
        // the same result bit for bit
        uint32_t qs[3];
#pragma unroll
        for (int k = 0; k < 3; k++) qs[k] = __builtin_amdgcn_perm(qw[k], qw[k], sel[k]);
#pragma unroll
        for (int c = 0; c < 2; c++) {
            float lo_t = tmin, hi_t = best;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float s = scale_of(meta, k);
                const float pn = __builtin_fmaf(s, (float)((qs[k] >> (16 * c)) & 0xFFu), no[k]);
                const float pf = __builtin_fmaf(s, (float)((qs[k] >> (16 * c + 8)) & 0xFFu), no[k]);
                lo_t = fmaxf(lo_t, (pn - org[k]) * inv[k]);
                hi_t = fminf(hi_t, (pf - org[k]) * inv[k]);
            }
            tn[c] = lo_t;
            hit[c] = lo_t <= hi_t * kRobust;
        }
        // guards: children must be valid indices, and a well-formed tree
        // visits each of its n-1 internal nodes at most once.  Folded in as a
        // select rather than an early return, so that the compiler cannot
        // move the box arithmetic, and with it the first node load, behind a
        // wait for the second.
        const bool bad = g + 1u >= sc.n || budget == 0u;
        budget--;
        const uint32_t leaf = (meta >> 24) & 3u;
        const uint32_t h = (hit[0] ? 1u : 0u) | (hit[1] ? 2u : 0u);
        pend = bad ? 0u : (h & leaf) | ((h & ~leaf & 3u) << 2);
        pg = g;
        ptn0 = tn[0];
        ptn1 = tn[1];
        return bad;
    }
    VZ_DEV bool has_leaves() const { return (pend & 3u) != 0u; }

    // Sphere tests of the pending visit's leaf children, in child order.  Both
    // children's spheres come in one load: pg + 1 < n holds after a visit that
    // did not end the traversal.  LOOP: one copy of the test run once per
    // pending leaf, so a wave whose lanes hold leaves at different child
    // positions runs it once, not twice (faster with LDS nodes); otherwise a
    // copy per child (faster with HBM nodes, whose waits it overlaps).
    template <bool LOOP>
    VZ_DEV void leaves(const SceneView& sc, uint32_t& ns)
    {
        if ((pend & 3u) == 0u) return;
        float4 s0, s1;
        ldg2(sc.spheres + pg, s0, s1);
        if (LOOP) {
            uint32_t p = pend & 3u;
#pragma clang loop unroll(disable)
            do {
                const bool first = (p & 1u) != 0u;
                test_leaf(sc, first ? pg : pg + 1u, first ? s0 : s1, ns);
                p &= p - 1u;
            } while (p != 0u);
        } else {
            if (pend & 1u) test_leaf(sc, pg, s0, ns);
            if (pend & 2u) test_leaf(sc, pg + 1u, s1, ns);
        }
    }

    // Next node after the pending visit: the near internal child (the far one
    // pushed), else the only one, else the nearest stacked node that can
    // still hold a closer hit.  Returns true when the traversal has finished.
    VZ_DEV bool decide(const SceneView& sc, uint32_t* lds_node, float* lds_t, uint32_t stride, BinOverflow& ov)
    {
        const bool in0 = (pend & 4u) != 0u, in1 = (pend & 8u) != 0u;
        const uint32_t g = pg;
        pend = 0u;
        if (in1 && g + 2u >= sc.n) return true;  // internal child g+1 must be <= n-2
        if (in0 && in1) {
            const int nr = (ptn1 < ptn0) ? 1 : 0;
            const uint32_t far_node = g + (uint32_t)(1 - nr);
            const float far_t = nr ? ptn0 : ptn1;
            if (sp - bottom == (uint32_t)S) {  // ring full: oldest entry to the overflow
                if (bottom >= kMaxDepth) return true;
                ov.node[bottom] = lds_node[(bottom & (S - 1)) * stride];
                ov.t[bottom] = lds_t[(bottom & (S - 1)) * stride];
                bottom++;
            }
            lds_node[(sp & (S - 1)) * stride] = far_node;
            lds_t[(sp & (S - 1)) * stride] = far_t;
            sp++;
            node = g + (uint32_t)nr;
            return false;
        } else if (in0) {
            node = g;
            return false;
        } else if (in1) {
            node = g + 1u;
            return false;
        }
        while (sp > 0u) {  // pop, culling entries beyond the best hit
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

    // One whole node visit; returns true when the traversal has finished.
    template <class NodeSrc>
    VZ_DEV bool step(const SceneView& sc, const NodeSrc& nodes, uint32_t* lds_node, float* lds_t,
                     uint32_t stride, BinOverflow& ov, uint32_t& nv, uint32_t& ns)
    {
        if (visit(sc, nodes, nv)) return true;
        leaves<NodeSrc::kLeafLoop>(sc, ns);
        return decide(sc, lds_node, lds_t, stride, ov);
    }
};

}  // namespace vz
