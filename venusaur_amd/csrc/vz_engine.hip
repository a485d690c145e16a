// vz_engine.hip -- wavefront integrator (see vz_engine.h for the layout and
// the stage list).  Replaces the recursive optixTrace chain of the
// reference's OptiX programs (Core/RayTracer.cu:163-450) with queue-driven
// persistent kernels, after Laine, Karras and Aila, "Megakernels Considered
// Harmful" (HPG 2013).
#include <algorithm>
#include <cstring>
#include <string>

#include <type_traits>

#include "vz_engine.h"
#include "vz_traverse.h"

namespace vz {
namespace {

constexpr int kStepBlk = 512;   // step-kernel workgroup (8 waves)
constexpr int kStackS = kWideStackS;  // LDS stack slots per lane
constexpr int kTbl = 128;       // LDS (request, pass) tally table entries
constexpr uint32_t kLdsNodes = 256;  // a scene of at most this many 4-wide nodes is staged in LDS (16 KB)
// A wave goes back for new rays once fewer than this many of its lanes still
// trace (1 = only when all are done, 65 = after every node step).  Defaults per
// kernel variant, chosen by sweeping: one small scene with its nodes in LDS
// (1080p book scene: 9), one scene with its nodes in memory (1M-sphere box
// scene: 41), several scenes in one run (64-scene batch at 48M / 8x8 / split
// 4: 5448 Mray-segments/s at 25, 5440 at 33, 5383 at 17, 5356 at 41; the
// 2000-request stream's first 600 requests: 5756 at 33, 5744 and 5702 at 25,
// 5669 and 5616 at 41).
constexpr uint32_t kRefillLds = 9;
constexpr uint32_t kRefillGlobal = 41;
constexpr uint32_t kRefillMulti = 33;
constexpr uint32_t kStrideMaxTiles = 24576;  // requests with fewer tiles visit them strided
constexpr uint32_t kFlushDen = 2;  // queued sphere tests run once 1/kFlushDen of the tracing lanes hold some
constexpr int kRegenBlk = 256;
constexpr uint32_t kRegenOver = 8;  // regen workgroups per queue slice: ~one pixel run per thread at 1080p
constexpr int kBlk = 256;
constexpr uint32_t kAccBatch = 8;  // finished passes accumulated together when nothing waits on them
constexpr uint32_t kEmpty = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t mbcnt(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Adds (count, segments) for `key` to the workgroup's LDS tally.  Returns
// false when the table is full (the caller then goes to global memory).
__device__ __forceinline__ bool tally_add(uint32_t* keys, uint32_t* cnt, unsigned long long* seg,
                                          uint32_t key, uint32_t c, unsigned long long s)
{
    uint32_t h = (key * 2654435761u) >> 25;  // 7 bits: kTbl == 128
    for (int probe = 0; probe < kTbl; probe++) {
        uint32_t prev = atomicCAS(&keys[h], kEmpty, key);
        if (prev == kEmpty || prev == key) {
            atomicAdd(&cnt[h], c);
            atomicAdd(&seg[h], s);
            return true;
        }
        h = (h + 1u) & (kTbl - 1);
    }
    return false;
}

__device__ __forceinline__ void terminate(const ReqDesc& D, uint32_t slot, uint32_t pixel, v3 L)
{
    float* pb = D.passbuf + (size_t)slot * 3u * D.npix;
    pb[pixel] = L.x;
    pb[D.npix + pixel] = L.y;
    pb[2u * D.npix + pixel] = L.z;
}

// One node visit: the 4-wide tree staged in LDS (LDSN) or read from memory.
template <bool LDSN, int S, bool COUNT, bool DEFER>
__device__ __forceinline__ bool step_node(Traversal<S, COUNT, DEFER>& tr, const SceneView& sc, const uint4* s_bvh,
                                          uint32_t* lds_node, float* lds_t, Overflow& ov, uint32_t& nv,
                                          uint32_t& ns, uint32_t* lds_pend)
{
    if (LDSN) return tr.step(sc, LdsNodes{s_bvh}, lds_node, lds_t, kStepBlk, ov, nv, ns, lds_pend);
    return tr.step(sc, GlobalNodes{sc.nodes}, lds_node, lds_t, kStepBlk, ov, nv, ns, lds_pend);
}

// ---- step: trace + miss/shade stages + compaction, one slice per workgroup ----
// Persistent traversal with per-lane refill (Aila & Laine, "Understanding the
// Efficiency of Ray Traversal on GPUs", HPG 2009): a lane whose ray has
// finished runs the miss or shade stage for it and then pulls the next ray of
// the slice, so a wave is not held by its slowest ray.  SINGLE: every ray of
// the launch belongs to scene 0 (the scene descriptor lives in scalars);
// LDSN: that scene's node array is staged in LDS.
template <bool COUNT, bool LDSN, bool SINGLE>
__global__ void __launch_bounds__(kStepBlk) __attribute__((amdgpu_waves_per_eu(6))) k_wf_step(const SceneView* __restrict__ scenes,
                                                      const ReqDesc* __restrict__ req,
                                                      const float4* __restrict__ q_o,
                                                      const float4* __restrict__ q_d,
                                                      const float4* __restrict__ q_t,
                                                      float4* __restrict__ n_o, float4* __restrict__ n_d,
                                                      float4* __restrict__ n_t,
                                                      const uint32_t* __restrict__ cnt_cur,
                                                      uint32_t* __restrict__ cnt_next,
                                                      BlockStats* __restrict__ bstats, uint32_t R,
                                                      uint32_t refill)
{
    __shared__ uint32_t s_node[kStackS * kStepBlk];
    __shared__ float s_t[kStackS * kStepBlk];
    __shared__ uint32_t t_key[kTbl], t_cnt[kTbl];
    __shared__ unsigned long long t_seg[kTbl];
    __shared__ uint32_t s_head, s_out;
    __shared__ unsigned long long s_red[4][kStepBlk / 64];
    __shared__ uint4 s_bvh[LDSN ? 4 * kLdsNodes : 1];
    // nodes from memory, timed runs: sphere tests queued per lane and run by
    // the whole wave together (counting runs test at once, so their counts
    // follow each ray alone)
    constexpr bool DEFER = !COUNT && !LDSN;
    __shared__ uint32_t s_pend[DEFER ? kPend * kStepBlk : 1];
    const uint32_t tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
    if (LDSN) {  // one small scene: stage its node array in LDS
        const uint32_t nn = 4u * scenes[0].nw;
        const uint4* src = scenes[0].nodes;
        for (uint32_t i = tid; i < nn; i += kStepBlk) s_bvh[i] = src[i];
    }
    for (uint32_t i = tid; i < (uint32_t)kTbl; i += kStepBlk) {
        t_key[i] = kEmpty;
        t_cnt[i] = 0u;
        t_seg[i] = 0ull;
    }
    if (tid == 0) {
        s_head = 0u;
        s_out = 0u;
    }
    __syncthreads();
    const uint32_t b = blockIdx.x;
    const size_t base_off = (size_t)b * R;
    const uint32_t count = cnt_cur[b];
    uint32_t nv = 0, ns = 0, nh = 0, seg = 0;
    SceneView sc;
    if (SINGLE) sc = scenes[0];
    Traversal<kStackS, COUNT, DEFER> tr;
    Overflow ov;
    bool active = false, tdone = false;
    size_t qi = 0;
    uint32_t meta = 0;
    for (;;) {
        // refill: idle lanes claim the next rays of the slice, one LDS atomic per wave
        const bool need = !active;
        const uint64_t m = __ballot(need);
        bool exhausted = false;
        if (m) {
            const int leader = __ffsll((unsigned long long)m) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(&s_head, (uint32_t)__popcll(m));
            base = __shfl(base, leader);
            exhausted = base + (uint32_t)__popcll(m) >= count;
            const uint32_t local = base + mbcnt(m);
            if (need && local < count) {
                qi = base_off + local;
                meta = __float_as_uint(q_t[qi].w);
                if (meta != kDeadRay) {
                    const float4 o4 = q_o[qi], d4 = q_d[qi];
                    if (!SINGLE) sc = scenes[req[meta >> 16].scene];
                    tdone = tr.init(sc, mk(o4.x, o4.y, o4.z), mk(d4.x, d4.y, d4.z), 0.001f, 1e16f, ns);
                    active = true;
                }
            }
        }
        if (__ballot(active) == 0ull) {
            if (exhausted) break;
            continue;  // claimed only padding items
        }
        // trace stage: advance every busy lane until enough of them finish.
        // (Parking lanes whose node has leaf children, to run the sphere tests
        // of many lanes together, was measured slower at every threshold.)
        uint64_t busy = __ballot(active && !tdone);
        while (busy) {
            if (active && !tdone)
                tdone = step_node<LDSN>(tr, sc, s_bvh, s_node + tid, s_t + tid, ov, nv, ns, s_pend + tid);
            busy = __ballot(active && !tdone);
            if (DEFER) {
                // flush once enough lanes hold queued tests, relative to the
                // lanes still tracing, or before a queue could overflow (a
                // visit queues at most 4)
                const uint64_t pm = __ballot(tr.npend != 0u);
                if (pm && (__ballot(tr.npend + 4u > kPend) != 0ull ||
                           kFlushDen * (uint32_t)__popcll(pm) >= (uint32_t)__popcll(busy)))
                    tr.flush(sc, s_pend + tid, kStepBlk, ns);
            }
            // refill <= 1: the whole wave finishes before going back for rays
            if (!exhausted && (uint32_t)__popcll(busy) < refill) break;
        }
        // every queued test has run before a lane's hit is shaded
        if (DEFER) tr.flush(sc, s_pend + tid, kStepBlk, ns);
        // miss and shade stages for the lanes whose ray finished
        const bool fin = active && tdone;
        if (__ballot(fin) == 0ull) continue;
        bool ended = false, survive = false;
        uint32_t key = 0, pathseg = 0, seed = 0;
        v3 pos = mk(0.0f, 0.0f, 0.0f), nd = pos, thr = pos;
        float4 o4 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (fin) {
            o4 = q_o[qi];
            const float4 d4 = q_d[qi], t4 = q_t[qi];
            const uint32_t rq = meta >> 16, slot = (meta >> 8) & 0xFFu, depth = meta & 0xFFu;
            const ReqDesc* D = &req[rq];
            const v3 o = mk(o4.x, o4.y, o4.z), d = mk(d4.x, d4.y, d4.z);
            seg++;
            seed = __float_as_uint(d4.w);
            v3 L = mk(0.0f, 0.0f, 0.0f);
            ended = true;
            if (tr.bleaf == 0xFFFFFFFFu) {
                // miss stage (RayTracer.cu:442-450)
                L = mul(mk(t4.x, t4.y, t4.z), sky(d));
            } else {
                nh++;
                if (depth > 0u) {
                    // shade stage (RayTracer.cu:272-440)
                    const float4 s = ldg(sc.spheres + tr.bleaf);
                    const uint32_t type = ldg(sc.leaf_info + tr.bleaf) >> 30;
                    const float4 mat = ldg(sc.mats + tr.bleaf);
                    pos = add(o, scl(d, tr.best));
                    v3 n = scl(sub(pos, mk(s.x, s.y, s.z)), recip_rn(s.w));
                    const bool front = dot(d, n) < 0.0f;
                    if (!front) n = neg(n);
                    v3 att;
                    if (scatter(type, mat, d, n, front, seed, nd, att)) {
                        thr = mul(mk(t4.x, t4.y, t4.z), att);
                        meta = meta - 1u;
                        survive = true;
                        ended = false;
                    }
                }
            }
            if (ended) {
                terminate(*D, slot, __float_as_uint(o4.w), L);
                key = (rq << 8) | slot;
                pathseg = D->max_depth - depth;
            }
            active = false;
        }
        // survivors: one LDS atomic per wave, slots in lane order
        const uint64_t sm = __ballot(survive);
        if (sm) {
            uint32_t p0 = 0;
            if (lane == (uint32_t)(__ffsll((unsigned long long)sm) - 1)) p0 = atomicAdd(&s_out, (uint32_t)__popcll(sm));
            p0 = __shfl(p0, __ffsll((unsigned long long)sm) - 1);
            if (survive) {
                const size_t k = base_off + p0 + mbcnt(sm);
                n_o[k] = make_float4(pos.x, pos.y, pos.z, o4.w);
                n_d[k] = make_float4(nd.x, nd.y, nd.z, __uint_as_float(seed));
                n_t[k] = make_float4(thr.x, thr.y, thr.z, __uint_as_float(meta));
            }
        }
        // ended paths: per distinct (request, pass) one tally update per wave
        uint64_t todo = __ballot(ended);
        while (todo) {
            const int leader = __ffsll((unsigned long long)todo) - 1;
            const uint32_t lk = __shfl(key, leader);
            const bool mine = ended && key == lk;
            const uint64_t same = __ballot(mine) & todo;
            const uint32_t sg = wave_sum<uint32_t>(mine ? pathseg : 0u);
            if ((int)lane == leader) {
                const uint32_t c = (uint32_t)__popcll(same);
                if (!tally_add(t_key, t_cnt, t_seg, lk, c, sg)) {
                    atomicAdd(&req[lk >> 8].done[lk & 0xFFu], c);
                    atomicAdd(&req[lk >> 8].segsum[lk & 0xFFu], (unsigned long long)sg);
                }
            }
            todo &= ~same;
        }
    }
    // workgroup epilogue: flush the tally, publish the slice count and stats
    const uint32_t w_seg = wave_sum(seg), w_nh = wave_sum(nh);
    const uint32_t w_nv = COUNT ? wave_sum(nv) : 0u, w_ns = COUNT ? wave_sum(ns) : 0u;
    if (lane == 0) {
        s_red[0][wave] = w_seg;
        s_red[1][wave] = w_nh;
        s_red[2][wave] = w_nv;
        s_red[3][wave] = w_ns;
    }
    __syncthreads();
    for (uint32_t i = tid; i < (uint32_t)kTbl; i += kStepBlk) {
        const uint32_t k = t_key[i];
        if (k != kEmpty) {
            atomicAdd(&req[k >> 8].done[k & 0xFFu], t_cnt[i]);
            atomicAdd(&req[k >> 8].segsum[k & 0xFFu], t_seg[i]);
        }
    }
    if (tid == 0) {
        cnt_next[b] = s_out;
        unsigned long long a[4] = {0, 0, 0, 0};
        for (int w = 0; w < kStepBlk / 64; w++)
            for (int q = 0; q < 4; q++) a[q] += s_red[q][w];
        BlockStats& B = bstats[b];
        B.segments += a[0];
        B.hits += a[1];
        B.node_visits += a[2];
        B.sphere_tests += a[3];
    }
}

// ---- control (one workgroup of 1024 threads) -----------------------------------
// Exclusive scan over the control workgroup of NW waves (16 or 4): an
// inclusive scan inside each wave by shuffles, the NW wave totals scanned by
// the first wave, two barriers in all.
// N independent scans share one set of barriers and shuffle steps: out[k]
// is the exclusive prefix of v[k], tot[k] the workgroup total.  s needs
// 33 * N words.
template <int N, uint32_t NW = 16>
__device__ void block_excl_scan_n(const uint32_t (&v)[N], uint32_t (&out)[N], uint32_t (&tot)[N], uint32_t* s)
{
    static_assert(NW >= 1 && NW <= 16, "the wave totals fit one 16-entry row");
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    uint32_t incl[N];
#pragma unroll
    for (int k = 0; k < N; k++) incl[k] = v[k];
#pragma unroll
    for (uint32_t off = 1; off < 64u; off <<= 1) {
#pragma unroll
        for (int k = 0; k < N; k++) {
            const uint32_t x = __shfl_up(incl[k], off);
            if (lane >= off) incl[k] += x;
        }
    }
    if (lane == 63u)
        for (int k = 0; k < N; k++) s[k * 33 + wave] = incl[k];
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int k = 0; k < N; k++) {
            const uint32_t w = lane < NW ? s[k * 33 + lane] : 0u;
            uint32_t wi = w;
#pragma unroll
            for (uint32_t off = 1; off < NW; off <<= 1) {
                const uint32_t x = __shfl_up(wi, off);
                if (lane >= off) wi += x;
            }
            if (lane < NW) s[k * 33 + 16 + lane] = wi - w;  // exclusive prefix of the wave totals
            if (lane == NW - 1u) s[k * 33 + 32] = wi;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; k++) {
        out[k] = s[k * 33 + 16 + wave] + incl[k] - v[k];
        tot[k] = s[k * 33 + 32];
    }
    __syncthreads();  // s is reused by the next scan
}

template <uint32_t NW = 16>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s, uint32_t* total)
{
    const uint32_t in[1] = {v};
    uint32_t out[1], tot[1];
    block_excl_scan_n<1, NW>(in, out, tot, s);
    *total = tot[0];
    return out[0];
}



// NT threads, one per request slot (n_slots <= NT); SPT: queue slices per
// thread (NB <= NT * SPT).  1024 threads with the smallest SPT that covers
// the run's slices, or 256 threads with SPT 4 for runs of at most 256
// requests and 1024 slices, whose scans and barriers span 4 waves instead
// of 16.
template <uint32_t SPT, uint32_t NT = 1024>
__global__ void __launch_bounds__(NT) k_wf_control(const ReqDesc* __restrict__ req,
                                                     ReqState* __restrict__ st, Counters* cnt,
                                                     Plan* __restrict__ plan,
                                                     uint32_t* __restrict__ cnt_next, uint32_t NB,
                                                     uint32_t R)
{
    __shared__ uint32_t s_scan[1024];
    __shared__ uint32_t s_quota[1024], s_seq[1024];
    const uint32_t r = threadIdx.x;
    const uint32_t nslots = cnt->n_slots;
    const uint32_t iter = cnt->iteration;
    // 1. free room of the slices this thread owns (SPT * r + q)
    // A run may cap the rays in flight below the queue's capacity: every
    // slice is then filled to Rf <= R (slice counts never exceed Rf, since
    // grants stop there and a step only removes rays).
    const uint32_t Rf = (cnt->fill != 0u && cnt->fill < R) ? cnt->fill : R;
    uint32_t fr[SPT];
    uint32_t my_free = 0;
    for (uint32_t q = 0; q < SPT; q++) {
        const uint32_t sb = r * SPT + q;
        fr[q] = (sb < NB && cnt_next[sb] < Rf) ? Rf - cnt_next[sb] : 0u;
        my_free += fr[q];
    }
    uint32_t freeq;
    (void)block_excl_scan<NT / 64>(my_free, s_scan, &freeq);
    // a request may start as many passes per iteration as its pass ring
    // allows; grants go to requests in slot (arrival) order, so the oldest
    // requests fill the queue first.  Fewer scenes in flight keep their BVHs
    // in cache (64 box scenes: 1.73 Gray-segments/s against 1.56 with the free
    // room shared evenly) and the oldest requests finish first (served p50
    // 1.6 s against 5.0 s on the 120-request stream).
    // 2. per request: passes whose paths all ended, and the camera-ray quota
    uint32_t acc_first = 0, acc_cnt = 0, quota = 0, npix = 0, npad = 0;
    bool running = false, still = false;
    unsigned long long next_item = 0;
    if (r < nslots && st[r].active) {
        const ReqDesc& D = req[r];
        ReqState S = st[r];
        running = true;
        npix = D.npix;
        npad = D.npad;
        uint32_t a = S.acc_next, c = 0;
        while (a + c < D.spp && (unsigned long long)(a + c + 1) * D.npad <= S.next_item &&
               D.done[(a + c) % D.K] == D.npix)
            c++;
        // Fewer than kAccBatch finished passes wait for more, so that one
        // apply reads several pass buffers and touches the sums once, unless
        // this is the request's last pass or the ring room without them
        // could bind a grant (it no longer covers the queue's free room).
        // The sums are formed in sample order either way.
        if (c != 0u && c < kAccBatch && a + c < D.spp &&
            (unsigned long long)(a + D.K) * D.npad - S.next_item >= freeq)
            c = 0u;
        for (uint32_t q = 0; q < c; q++) {
            const uint32_t slot = (a + q) % D.K;
            S.segments += D.segsum[slot];
            D.done[slot] = 0u;
            D.segsum[slot] = 0ull;
        }
        acc_first = a;
        acc_cnt = c;
        S.acc_next = a + c;
        if (S.acc_next == D.spp) {
            S.active = 0u;
            S.finish_iter = iter;
        } else {
            still = true;
            unsigned long long rem = D.total_items - S.next_item;
            unsigned long long ring = (unsigned long long)(S.acc_next + D.K) * D.npad - S.next_item;
            const unsigned long long q = rem < ring ? rem : ring;
            quota = (uint32_t)(q < 0xFFFFFFFFull ? q : 0xFFFFFFFFull);
        }
        next_item = S.next_item;
        st[r].acc_next = S.acc_next;
        st[r].active = S.active;
        st[r].finish_iter = S.finish_iter;
        st[r].segments = S.segments;
    }
    // free room goes to the requests in arrival order: a request gets what
    // is left after every earlier (lower seq) request's quota.  Slots are
    // reused by the serving loop, so slot order is not arrival order.
    s_quota[r] = quota;
    s_seq[r] = running ? req[r].seq : 0u;
    __syncthreads();
    unsigned long long before = 0;
    if (quota) {
        const uint32_t my_seq = s_seq[r];
        for (uint32_t j = 0; j < nslots; j++)
            if ((int32_t)(s_seq[j] - my_seq) < 0) before += s_quota[j];  // wrap-safe order
    }
    __syncthreads();
    const uint32_t grant = before >= freeq ? 0u : (uint32_t)min((unsigned long long)quota, freeq - before);
    if (running && grant) st[r].next_item = next_item + grant;
    // positions of the apply and regen entries, their pixel and ray
    // prefixes, the pixels the grants cover, the requests still running
    const uint32_t sv[6] = {acc_cnt ? 1u : 0u, acc_cnt ? npix : 0u, grant ? 1u : 0u, grant, min(grant, npad),
                            still ? 1u : 0u};
    uint32_t sp[6], stot[6];
    block_excl_scan_n<6, NT / 64>(sv, sp, stot, s_scan);
    const uint32_t posA = sp[0], pixA = sp[1], posR = sp[2], preR = sp[3], preRun = sp[4];
    const uint32_t nA = stot[0], totA = stot[1], nR = stot[2], totR = stot[3], totRuns = stot[4], nact = stot[5];
    if (acc_cnt) {
        plan->acc_req[posA] = r;
        plan->acc_first[posA] = acc_first;
        plan->acc_cnt[posA] = acc_cnt;
        plan->acc_prefix[posA] = pixA;
    }
    if (grant) {
        plan->regen_req[posR] = r;
        plan->regen_prefix[posR] = preR;
        plan->regen_cnt[posR] = grant;
        plan->run_prefix[posR] = preRun;
        plan->regen_base[posR] = next_item;
        const unsigned long long s0 = next_item / npad;
        plan->regen_s0[posR] = (uint32_t)s0;
        plan->regen_i0[posR] = (uint32_t)(next_item - s0 * npad);
    }
    // 3. split the granted camera rays over the slices so that every slice
    // ends up near the same ray count: the next step launch lasts as long as
    // its fullest slice.  Level L = ceil(rays after regen / NB) <= Rf; a slice
    // below L gets a share of the new rays proportional to its deficit.
    const unsigned long long held = (unsigned long long)NB * Rf - freeq;
    const uint32_t level = (uint32_t)((held + totR + NB - 1u) / NB);
    uint32_t def[SPT], my_def = 0;
    for (uint32_t q = 0; q < SPT; q++) {
        const uint32_t c = Rf - fr[q];
        def[q] = (r * SPT + q < NB && level > c) ? level - c : 0u;  // <= fr[q]: level <= Rf
        my_def += def[q];
    }
    uint32_t tdef;
    (void)block_excl_scan<NT / 64>(my_def, s_scan, &tdef);  // tdef >= totR
    uint32_t g[SPT], my_g = 0, my_short = 0;
    for (uint32_t q = 0; q < SPT; q++) {
        g[q] = tdef ? (uint32_t)((unsigned long long)totR * def[q] / tdef) : 0u;
        my_g += g[q];
        my_short += g[q] < def[q] ? 1u : 0u;
    }
    const uint32_t gv[2] = {my_g, my_short};
    uint32_t gp[2], gtot[2];
    block_excl_scan_n<2, NT / 64>(gv, gp, gtot, s_scan);
    const uint32_t tg = gtot[0], short_pre = gp[1];
    // the rounding leftover (< number of short slices): one more ray each to
    // the first short slices.  The slices before this thread's got
    // min(short_pre, left) of them, so its first slice starts there after
    // the proportional shares: no further scan.
    const uint32_t left = totR - tg;
    uint32_t k = short_pre;
    for (uint32_t q = 0; q < SPT; q++) {
        if (g[q] < def[q]) {
            if (k < left) g[q]++;
            k++;
        }
    }
    uint32_t run = gp[0] + min(short_pre, left);
    for (uint32_t q = 0; q < SPT; q++) {
        const uint32_t sb = r * SPT + q;
        if (sb < NB) {
            const uint32_t c = Rf - fr[q];
            plan->blk_first[sb] = run;
            plan->blk_count[sb] = g[q];
            plan->blk_pos0[sb] = c;
            cnt_next[sb] = c + g[q];
        }
        run += g[q];
    }
    if (r == 0) {
        plan->n_acc = nA;
        plan->total_acc_pix = totA;
        plan->n_regen = nR;
        plan->total_regen = totR;
        plan->total_runs = totRuns;
        cnt->active = nact;
        cnt->total_regen = totR;
        cnt->iteration = iter + 1u;
    }
}

// floor(n / d) for d >= 1 from rd = 1 / d rounded to double: n * rd is
// within n / d * 2^-51 < 1 of n / d, so its truncation is off by at most one
// either way, and one remainder test corrects it.  Exact for every 32-bit n;
// a few instructions where an integer division takes dozens.
__device__ __forceinline__ uint32_t udiv_r(uint32_t n, uint32_t d, double rd)
{
    uint32_t q = (uint32_t)((double)n * rd);
    const long long r = (long long)n - (long long)q * d;
    if (r < 0) q--;
    else if (r >= (long long)d) q++;
    return q;
}

// n mod d for n < 2^53, d >= 1, rd = 1 / d rounded: n is exact in double and
// n * rd is within n / d * 2^-52 < 1 of n / d (n / d < 2^52), so one
// correction either way gives the exact remainder.
__device__ __forceinline__ uint32_t umod_r(unsigned long long n, uint32_t d, double rd)
{
    const long long q = (long long)((double)n * rd);
    long long r = (long long)n - q * (long long)d;
    if (r < 0) r += d;
    else if (r >= (long long)d) r -= d;
    return (uint32_t)r;
}

// Largest e with prefix[e] <= j (prefix is ascending, prefix[0] == 0).
__device__ __forceinline__ uint32_t find_entry(const uint32_t* prefix, uint32_t n, uint32_t j)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= j) lo = mid; else hi = mid;
    }
    return lo;
}

// ---- apply: ordered per-pixel accumulation and the final blend/encode --------
// Pixel px of apply entry e: adds the entry's finished passes in sample order
// to the running sums, and on the request's last pass blends the average
// into the accumulator and encodes the image.
__device__ __forceinline__ void apply_pixel(const ReqDesc& D, const Plan* __restrict__ plan, uint32_t e, uint32_t px)
{
    const uint32_t first = plan->acc_first[e], c = plan->acc_cnt[e];
    const bool fin = first + c == D.spp;
    float oc[3];
    if (D.flags & VZ_ACCUM_FP64) {
        double s[3] = {0.0, 0.0, 0.0};
        if (first) for (int k = 0; k < 3; k++) s[k] = D.sum64[(size_t)k * D.npix + px];
        for (uint32_t q = 0, slot = first - udiv_r(first, D.K, D.rd_K) * D.K; q < c; q++) {
            const float* pb = D.passbuf + (size_t)slot * 3u * D.npix;
            for (int k = 0; k < 3; k++) s[k] += (double)pb[(size_t)k * D.npix + px];
            slot = slot + 1u == D.K ? 0u : slot + 1u;
        }
        if (!fin) {
            for (int k = 0; k < 3; k++) D.sum64[(size_t)k * D.npix + px] = s[k];
            return;
        }
        double* acc = (double*)D.accum + 4 * (size_t)px;
        double avg[3];
        for (int k = 0; k < 3; k++) avg[k] = s[k] / (double)D.spp;
        if (D.subframe > 1u) {
            double a = 1.0 / (double)D.subframe;
            for (int k = 0; k < 3; k++) avg[k] = acc[k] + a * (avg[k] - acc[k]);
        }
        reinterpret_cast<double2*>(acc)[0] = make_double2(avg[0], avg[1]);
        reinterpret_cast<double2*>(acc)[1] = make_double2(avg[2], 1.0);
        for (int k = 0; k < 3; k++) oc[k] = (float)avg[k];
    } else {
        float s[3] = {0.0f, 0.0f, 0.0f};
        if (first) for (int k = 0; k < 3; k++) s[k] = D.sum32[(size_t)k * D.npix + px];
        for (uint32_t q = 0, slot = first - udiv_r(first, D.K, D.rd_K) * D.K; q < c; q++) {
            const float* pb = D.passbuf + (size_t)slot * 3u * D.npix;
            for (int k = 0; k < 3; k++) s[k] = s[k] + pb[(size_t)k * D.npix + px];
            slot = slot + 1u == D.K ? 0u : slot + 1u;
        }
        if (!fin) {
            for (int k = 0; k < 3; k++) D.sum32[(size_t)k * D.npix + px] = s[k];
            return;
        }
        float4* acc = (float4*)D.accum + px;
        float inv = 1.0f / (float)D.spp;
        float avg[3] = {s[0] * inv, s[1] * inv, s[2] * inv};
        if (D.subframe > 1u) {
            float4 prev = *acc;
            float a = 1.0f / (float)D.subframe;
            avg[0] = prev.x + a * (avg[0] - prev.x);
            avg[1] = prev.y + a * (avg[1] - prev.y);
            avg[2] = prev.z + a * (avg[2] - prev.z);
        }
        *acc = make_float4(avg[0], avg[1], avg[2], 1.0f);
        for (int k = 0; k < 3; k++) oc[k] = avg[k];
    }
    if (D.image) D.image[px] = make_color(oc[0], oc[1], oc[2]);
}

// The first lane's value of v among the lanes where v == that value, as a
// wave-uniform value: the caller runs a waterfall loop in which every active
// lane has v == readfirstlane(v), and the compiler, knowing that equality,
// would otherwise use the per-lane v; the empty asm hides the copy it reads.
__device__ __forceinline__ uint32_t uniform_copy(uint32_t v)
{
    asm volatile("" : "+v"(v));
    return (uint32_t)__builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ void apply_passes(const ReqDesc* __restrict__ req, const Plan* __restrict__ plan,
                                             uint32_t bid, uint32_t nblk)
{
    const uint32_t total = plan->total_acc_pix, n = plan->n_acc;
    for (uint32_t j = bid * kBlk + threadIdx.x; j < total; j += nblk * kBlk) {
        const uint32_t e_lane = find_entry(plan->acc_prefix, n, j);
        // one pass per distinct entry among the wave's lanes, with the entry
        // and its request descriptor in SGPRs (waterfall bound: see regen)
        bool todo = true;
        for (uint32_t pass = 0; pass < 64u && todo; pass++) {
            if (e_lane != (uint32_t)__builtin_amdgcn_readfirstlane(e_lane)) continue;
            todo = false;
            const uint32_t e = uniform_copy(e_lane);
            apply_pixel(req[plan->acc_req[e]], plan, e, j - plan->acc_prefix[e]);
        }
    }
}

// ---- regen: camera rays for the items granted to each slice ------------------------
// Camera rays for the granted items.  One thread per pixel run: a request's
// grant [base, base + cnt) in pass-major item order covers each padded pixel
// for one or more consecutive passes, and the camera seed of pass s+1 follows
// from the camera draws of pass s (RayTracer.cu:163-204), so one thread walks
// a pixel's passes in order from a single seed read.  Item j of the grant
// goes to granted index prefix + j, which the control kernel mapped to a
// slice and a position in it.
__device__ __forceinline__ void regen_rays(const ReqDesc* __restrict__ req, const Plan* __restrict__ plan,
                                           uint32_t NB, float4* __restrict__ n_o, float4* __restrict__ n_d,
                                           float4* __restrict__ n_t, BlockStats* __restrict__ bstats, uint32_t R,
                                           uint32_t bid, uint32_t nblk, uint32_t* s_first, uint32_t* s_paths)
{
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < NB; i += kRegenBlk) s_first[i] = plan->blk_first[i];
    __syncthreads();
    const uint32_t nreg = plan->n_regen, runs = plan->total_runs;
    uint32_t paths = 0;
    for (uint32_t r = bid * kRegenBlk + tid; r < runs; r += nblk * kRegenBlk) {
        const uint32_t e_lane = find_entry(plan->run_prefix, nreg, r);
        // One pass per distinct entry among the wave's lanes (one, unless the
        // wave straddles two requests' grants): the entry index is then
        // wave-uniform, so the plan entry and the request descriptor are read
        // by scalar loads into SGPRs instead of into every lane's VGPRs.
        // Each pass retires at least the first remaining lane, so 64 passes
        // cover a wave.  The bound also matters to the compiler: seen one
        // thread at a time, an unbounded loop whose lane never matches would
        // spin without side effects, which it may assume does not happen,
        // and it then drops the comparison (it did so here, and every lane
        // took the first lane's entry).
        bool todo = true;
        for (uint32_t pass = 0; pass < 64u && todo; pass++) {
            if (e_lane != (uint32_t)__builtin_amdgcn_readfirstlane(e_lane)) continue;
            todo = false;
            const uint32_t e = uniform_copy(e_lane);
            const uint32_t rq = plan->regen_req[e];
            const ReqDesc& D = req[rq];
            const unsigned long long base = plan->regen_base[e];
            const unsigned long long end = base + plan->regen_cnt[e];
            const uint32_t jprefix = plan->regen_prefix[e];
            // item base + k = pass s, padded pixel i (k < npad: at most one wrap)
            const uint32_t k = r - plan->run_prefix[e];
            unsigned long long item = base + k;
            uint32_t s = plan->regen_s0[e], i = plan->regen_i0[e] + k;
            if (i >= D.npad) {
                i -= D.npad;
                s++;
            }
            const uint32_t tsz = D.tw * D.th;
            const uint32_t t0 = udiv_r(i, tsz, D.rd_tsz), w = i - t0 * tsz;
            // the t0-th tile visited
            const uint32_t tile = umod_r((unsigned long long)t0 * D.tstride, D.ntiles, D.rd_ntiles);
            const uint32_t ty = udiv_r(tile, D.tiles_x, D.rd_tx), wy = udiv_r(w, D.tw, D.rd_tw);
            const uint32_t x = (tile - ty * D.tiles_x) * D.tw + (w - wy * D.tw);
            const uint32_t y = ty * D.th + wy;
            const bool valid = x < D.cam.width && y < D.cam.height;
            const uint32_t pixel = valid ? y * D.cam.width + x : 0u;
            uint32_t seed = 0;
            if (valid) seed = s == 0u ? tea4(pixel, D.subframe) : D.camseed[pixel];
            uint32_t slot = s - udiv_r(s, D.K, D.rd_K) * D.K;
            for (; item < end; item += D.npad, s++, slot = slot + 1u == D.K ? 0u : slot + 1u) {
                // granted index J lies in the last slice whose first index <= J
                const uint32_t J = jprefix + (uint32_t)(item - base);
                uint32_t lo = 0, hi = NB;
                while (hi - lo > 1u) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_first[mid] <= J) lo = mid; else hi = mid;
                }
                const size_t at = (size_t)lo * R + plan->blk_pos0[lo] + (J - s_first[lo]);
                if (valid) {
                    v3 o, d;
                    camera_ray(D.cam, x, y, seed, o, d);
                    const uint32_t meta = (rq << 16) | (slot << 8) | (D.max_depth - 1u);
                    n_o[at] = make_float4(o.x, o.y, o.z, __uint_as_float(pixel));
                    n_d[at] = make_float4(d.x, d.y, d.z, __uint_as_float(seed));
                    n_t[at] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(meta));
                    paths++;
                } else {
                    n_t[at] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(kDeadRay));
                }
            }
            if (valid) D.camseed[pixel] = seed;
        }
    }
    paths = wave_sum(paths);
    if (__lane_id() == 0) s_paths[tid >> 6] = paths;
    __syncthreads();
    if (tid == 0) {
        unsigned long long p = 0;
        for (int w = 0; w < kRegenBlk / 64; w++) p += s_paths[w];
        // kRegenOver regen workgroups share each slice's stats entry
        if (p) atomicAdd(&bstats[bid % NB].paths, p);
    }
}

// ---- finish: apply and regen in one launch --------------------------------------
// Both stages read the control kernel's plan and touch disjoint memory (apply:
// finished passes' buffers, sums, accum and image; regen: the next queue
// parity, camera seeds and path counts), so one launch runs both: the first
// n_apply workgroups apply, the rest write camera rays, and the two overlap
// instead of running back to back.  The kernel is bound by the latency of
// its dependent loads, so it runs 8 waves per SIMD (52 VGPRs; a few SGPRs
// spill to VGPR lanes): 1080p book scene +2.5% over 6 waves with the request
// descriptor in VGPRs.
static_assert(kBlk == kRegenBlk, "apply and regen share the workgroup size");
__global__ void __launch_bounds__(kBlk) __attribute__((amdgpu_waves_per_eu(8))) k_wf_finish(const ReqDesc* __restrict__ req,
                                                    const Plan* __restrict__ plan, uint32_t NB,
                                                    float4* __restrict__ n_o, float4* __restrict__ n_d,
                                                    float4* __restrict__ n_t,
                                                    BlockStats* __restrict__ bstats, uint32_t R, uint32_t n_apply)
{
    extern __shared__ uint32_t s_first[];  // NB words (dynamic LDS, sized by the launch)
    __shared__ uint32_t s_paths[kRegenBlk / 64];
    if (blockIdx.x < n_apply)
        apply_passes(req, plan, blockIdx.x, n_apply);
    else
        regen_rays(req, plan, NB, n_o, n_d, n_t, bstats, R, blockIdx.x - n_apply, gridDim.x - n_apply, s_first,
                   s_paths);
}

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace

Cam make_cam(const vzLaunchParams& p)
{
    auto nrm = [](v3 v) {
        float inv = 1.0f / sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
        v3 r;
        r.x = v.x * inv;
        r.y = v.y * inv;
        r.z = v.z * inv;
        return r;
    };
    Cam c;
    c.O.x = p.origin[0]; c.O.y = p.origin[1]; c.O.z = p.origin[2];
    c.U.x = p.u[0]; c.U.y = p.u[1]; c.U.z = p.u[2];
    c.V.x = p.v[0]; c.V.y = p.v[1]; c.V.z = p.v[2];
    c.W.x = p.w[0]; c.W.y = p.w[1]; c.W.z = p.w[2];
    c.un = nrm(c.U);
    c.vn = nrm(c.V);
    c.lr = p.lens_radius;
    c.width = p.width;
    c.height = p.height;
    return c;
}

static uint32_t pass_slots(const vzLaunchParams& p)
{
    uint32_t k = p.max_depth + 2u;
    return std::min<uint32_t>(std::min<uint32_t>(k, p.samples_per_pixel), 255u);
}

size_t request_bytes(const vzLaunchParams& p)
{
    size_t npix = (size_t)p.width * p.height;
    size_t K = pass_slots(p);
    bool fp64 = (p.flags & VZ_ACCUM_FP64) != 0;
    return align_up(K * 3 * npix * 4, 256) + align_up(3 * npix * (fp64 ? 8 : 4), 256) +
           align_up(npix * 4, 256) + align_up(K * 4, 256) + align_up(K * 8, 256) + 1024;
}

Engine::~Engine()
{
    if (graph) (void)hipGraphExecDestroy(graph);
    if (graph_tmpl) (void)hipGraphDestroy(graph_tmpl);
    if (graph_stream) (void)hipStreamDestroy(graph_stream);
    for (int i = 0; i < 2; i++) {
        if (q_o[i]) (void)hipFree(q_o[i]);
        if (q_d[i]) (void)hipFree(q_d[i]);
        if (q_t[i]) (void)hipFree(q_t[i]);
    }
    void* dev[] = {d_bcnt, d_bstats, d_req, d_state, d_scenes, d_cnt, d_plan};
    for (void* ptr : dev)
        if (ptr) (void)hipFree(ptr);
    void* host[] = {h_cnt, h_state, h_req, h_scenes, h_bstats, h_state_stage, h_status, h_nslots};
    for (void* ptr : host)
        if (ptr) (void)hipHostFree(ptr);
    for (auto& b : slot_mem)
        if (b.ptr) (void)hipFree(b.ptr);
}

vzStatus engine_create(vzContext ctx, const EngineConfig& cfg, Engine** out)
{
    Engine* e = new Engine();
    e->ctx = ctx;
    e->cfg = cfg;
    if (e->cfg.tile_w == 0) e->cfg.tile_w = EngineConfig().tile_w;
    if (e->cfg.tile_h == 0) e->cfg.tile_h = EngineConfig().tile_h;
    if (e->cfg.iters_per_graph < 2) e->cfg.iters_per_graph = 2;
    e->cfg.iters_per_graph &= ~1u;
    auto fail = [&](vzStatus s, const char* what) {
        delete e;
        return set_error(s, what);
    };
    int occ = 0;
    int occ_lds = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(k_wf_step<false, false, false>),
                                                     kStepBlk, 0) != hipSuccess || occ < 1)
        occ = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_lds, reinterpret_cast<const void*>(k_wf_step<false, true, true>),
                                                     kStepBlk, 0) == hipSuccess && occ_lds >= 1)
        occ = std::min(occ, occ_lds);
    uint32_t cus = ctx->num_cus ? ctx->num_cus : 256;
    // one slice per resident step workgroup; R_base is a multiple of
    // 64 * kMaxSplit so that every split keeps 64-ray slices
    e->NB_res = std::min<uint32_t>(kMaxBlocks, cus * (uint32_t)occ);
    uint32_t want = std::max<uint32_t>(cfg.queue_capacity, e->NB_res * 64u * kMaxSplit);
    e->R_base = (uint32_t)align_up((want + e->NB_res - 1) / e->NB_res, 64 * kMaxSplit);
    e->NB = e->NB_res;
    e->R = e->R_base;
    e->split = 1;
    e->Q = e->NB * e->R;
    const size_t nb_max = (size_t)e->NB_res * kMaxSplit;
    e->cfg.queue_capacity = e->Q;
    size_t Q = e->Q;
    for (int i = 0; i < 2; i++) {
        if (hipMalloc(&e->q_o[i], Q * sizeof(float4)) != hipSuccess ||
            hipMalloc(&e->q_d[i], Q * sizeof(float4)) != hipSuccess ||
            hipMalloc(&e->q_t[i], Q * sizeof(float4)) != hipSuccess)
            return fail(VZ_ERROR_OUT_OF_MEMORY, "engine: ray queue allocation failed");
    }
    if (hipMalloc(&e->d_bcnt, 2 * nb_max * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&e->d_bstats, nb_max * sizeof(BlockStats)) != hipSuccess ||
        hipMalloc(&e->d_req, kMaxRequests * sizeof(ReqDesc)) != hipSuccess ||
        hipMalloc(&e->d_state, kMaxRequests * sizeof(ReqState)) != hipSuccess ||
        hipMalloc(&e->d_cnt, sizeof(Counters)) != hipSuccess ||
        hipMalloc(&e->d_plan, sizeof(Plan)) != hipSuccess ||
        hipMalloc(&e->d_scenes, kMaxScenes * sizeof(SceneView)) != hipSuccess)
        return fail(VZ_ERROR_OUT_OF_MEMORY, "engine: table allocation failed");
    if (hipHostMalloc(&e->h_cnt, 2 * sizeof(Counters)) != hipSuccess ||
        hipHostMalloc(&e->h_state, kMaxRequests * sizeof(ReqState)) != hipSuccess ||
        hipHostMalloc(&e->h_req, kMaxRequests * sizeof(ReqDesc)) != hipSuccess ||
        hipHostMalloc(&e->h_bstats, nb_max * sizeof(BlockStats)) != hipSuccess ||
        hipHostMalloc(&e->h_scenes, kMaxScenes * sizeof(SceneView)) != hipSuccess)
        return fail(VZ_ERROR_OUT_OF_MEMORY, "engine: pinned allocation failed");
    std::memset(e->h_cnt, 0, 2 * sizeof(Counters));
    std::memset(e->h_state, 0, kMaxRequests * sizeof(ReqState));
    std::memset(e->h_req, 0, kMaxRequests * sizeof(ReqDesc));
    std::memset(e->h_scenes, 0, kMaxScenes * sizeof(SceneView));
    if (hipMemset(e->d_state, 0, kMaxRequests * sizeof(ReqState)) != hipSuccess ||
        hipMemset(e->d_scenes, 0, kMaxScenes * sizeof(SceneView)) != hipSuccess ||
        hipMemset(e->d_cnt, 0, sizeof(Counters)) != hipSuccess ||
        hipMemset(e->d_plan, 0, sizeof(Plan)) != hipSuccess ||
        hipMemset(e->d_bcnt, 0, 2 * nb_max * sizeof(uint32_t)) != hipSuccess ||
        hipMemset(e->d_bstats, 0, nb_max * sizeof(BlockStats)) != hipSuccess)
        return fail(VZ_ERROR_HIP, "engine: initial memset failed");
    e->grid_small = cus * 4;
    e->slot_mem.resize(kMaxRequests);
    *out = e;
    return VZ_SUCCESS;
}

void engine_destroy(Engine* e) { delete e; }

// Enqueues one wavefront iteration of parity p (reads slices p, writes 1-p).
static void enqueue_iteration(Engine* e, hipStream_t st, int p, bool count, bool lds, bool single, hipEvent_t ev0,
                              hipEvent_t ev1)
{
    const int n = 1 - p;
    uint32_t* cnt_cur = e->d_bcnt + (size_t)p * e->NB;
    uint32_t* cnt_nxt = e->d_bcnt + (size_t)n * e->NB;
    if (ev0) (void)hipEventRecord(ev0, st);
#define VZ_STEP(C, L, S1)                                                                                  \
    hipLaunchKernelGGL((k_wf_step<C, L, S1>), dim3(e->NB), dim3(kStepBlk), 0, st, e->d_scenes, e->d_req,        \
                       e->q_o[p], e->q_d[p], e->q_t[p], e->q_o[n], e->q_d[n], e->q_t[n], cnt_cur, cnt_nxt,      \
                       e->d_bstats, e->R, refill)
    const uint32_t refill = e->cfg.refill_lanes ? std::min<uint32_t>(e->cfg.refill_lanes, 65u)
                                                : (lds ? kRefillLds : (single ? kRefillGlobal : kRefillMulti));
    if (count) {
        if (lds) VZ_STEP(true, true, true); else if (single) VZ_STEP(true, false, true); else VZ_STEP(true, false, false);
    } else {
        if (lds) VZ_STEP(false, true, true); else if (single) VZ_STEP(false, false, true); else VZ_STEP(false, false, false);
    }
#undef VZ_STEP
    if (ev1) (void)hipEventRecord(ev1, st);
    if (e->ctl_small)
        hipLaunchKernelGGL((k_wf_control<4, 256>), dim3(1), dim3(256), 0, st, e->d_req, e->d_state, e->d_cnt,
                           e->d_plan, cnt_nxt, e->NB, e->R);
    else if (e->NB <= 4096u)
        hipLaunchKernelGGL(k_wf_control<4>, dim3(1), dim3(1024), 0, st, e->d_req, e->d_state, e->d_cnt, e->d_plan,
                           cnt_nxt, e->NB, e->R);
    else
        hipLaunchKernelGGL(k_wf_control<kMaxBlocks / 1024>, dim3(1), dim3(1024), 0, st, e->d_req, e->d_state,
                           e->d_cnt, e->d_plan, cnt_nxt, e->NB, e->R);
    hipLaunchKernelGGL(k_wf_finish, dim3(e->grid_small + kRegenOver * e->NB_res), dim3(kBlk),
                       e->NB * sizeof(uint32_t), st, e->d_req, e->d_plan, e->NB, e->q_o[n], e->q_d[n], e->q_t[n],
                       e->d_bstats, e->R, e->grid_small);
}

static vzStatus ensure_graph(Engine* e, bool count, bool lds, bool single)
{
    // Every buffer the captured launches read is allocated once in
    // engine_create; what varies between runs (the queue partition and the
    // kernel variant) is in the key.
    const int key = (count ? 1 : 0) | (lds ? 2 : 0) | (single ? 4 : 0) | (int)(e->split << 3) |
                    (e->ctl_small ? 1 << 7 : 0);
    if (e->graph && e->graph_key == key) return VZ_SUCCESS;
    if (e->graph) {
        (void)hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
    }
    if (e->graph_tmpl) {
        (void)hipGraphDestroy(e->graph_tmpl);
        e->graph_tmpl = nullptr;
    }
    e->graph_key = -1;
    if (!e->graph_stream) VZ_HIP_TRY(hipStreamCreateWithFlags(&e->graph_stream, hipStreamNonBlocking));
    VZ_HIP_TRY(hipStreamBeginCapture(e->graph_stream, hipStreamCaptureModeThreadLocal));
    for (uint32_t it = 0; it < e->cfg.iters_per_graph; it++)
        enqueue_iteration(e, e->graph_stream, (int)(it & 1u), count, lds, single, nullptr, nullptr);
    VZ_HIP_TRY(hipStreamEndCapture(e->graph_stream, &e->graph_tmpl));
    hipError_t err = hipGraphInstantiate(&e->graph, e->graph_tmpl, nullptr, nullptr, 0);
    if (err != hipSuccess) {
        e->graph = nullptr;
        return set_error(VZ_ERROR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(err));
    }
    e->graph_key = key;
    e->graph_scenes = e->d_scenes;
    return VZ_SUCCESS;
}

// Launches the captured iterations.  The launches read the scene table at the
// address it had at capture, so check that it is still the engine's table.
static vzStatus launch_graph(Engine* e, hipStream_t st)
{
    if (!e->graph || e->graph_scenes != e->d_scenes)
        return set_error(VZ_ERROR_INTERNAL, "captured graph does not match the engine's scene table");
    VZ_HIP_TRY(hipGraphLaunch(e->graph, st));
    return VZ_SUCCESS;
}

// Fills slot `slot` of the host table for one request; allocates its buffers.
static vzStatus setup_slot(Engine* e, uint32_t slot, uint32_t scene_idx, const vzLaunchParams& p, uint32_t seq,
                           hipStream_t st)
{
    ReqDesc& D = e->h_req[slot];
    std::memset(&D, 0, sizeof(D));
    D.cam = make_cam(p);
    D.scene = scene_idx;
    D.spp = p.samples_per_pixel;
    D.subframe = p.subframe_index;
    D.max_depth = p.max_depth;
    D.flags = p.flags;
    D.npix = p.width * p.height;
    D.tw = e->cfg.tile_w;
    D.th = e->cfg.tile_h;
    D.tiles_x = (p.width + D.tw - 1) / D.tw;
    uint32_t tiles_y = (p.height + D.th - 1) / D.th;
    D.npad = D.tiles_x * D.tw * tiles_y * D.th;
    // Small requests visit their tiles in the order t * tstride mod ntiles,
    // a stride near ntiles / golden ratio and coprime to ntiles: the run of
    // consecutive items one slice receives then samples the whole image
    // instead of one band of it, so slices of sky and slices of ground cost
    // alike (64-scene batch of 256-1024 px requests, at most 16384 tiles:
    // plain launches +14%, batched +7%).  Large requests keep tile order:
    // their slices hold many bands each, and neighbouring tiles in one slice
    // trace more coherently (1080p, 32400 tiles: 1-2.5% slower strided).
    D.ntiles = D.tiles_x * tiles_y;
    D.tstride = 1;
    if (D.ntiles > 2u && D.ntiles < kStrideMaxTiles) {
        uint32_t st0 = (uint32_t)((double)D.ntiles * 0.6180339887498949) | 1u;
        auto gcd = [](uint32_t a, uint32_t b) { while (b) { uint32_t t = a % b; a = b; b = t; } return a; };
        while (gcd(st0, D.ntiles) != 1u) st0++;
        D.tstride = st0 % D.ntiles;
    }
    D.rd_ntiles = 1.0 / (double)D.ntiles;
    D.K = pass_slots(p);
    D.rd_tsz = 1.0 / ((double)D.tw * D.th);
    D.rd_tw = 1.0 / (double)D.tw;
    D.rd_tx = 1.0 / (double)D.tiles_x;
    D.rd_K = 1.0 / (double)D.K;
    D.seq = seq;
    D.total_items = (unsigned long long)D.spp * D.npad;
    D.accum = p.accum;
    D.image = (uint32_t*)p.image;
    size_t need = request_bytes(p);
    DeviceBlock& b = e->slot_mem[slot];
    if (e->stream_alloc) {
        // serving: exact-size buffer from the stream-ordered pool, no device sync
        if (b.ptr) return set_error(VZ_ERROR_INTERNAL, "slot buffer still held");
        if (hipMallocAsync(&b.ptr, need, st) != hipSuccess) {
            b.ptr = nullptr;
            return set_error(VZ_ERROR_OUT_OF_MEMORY, "engine: request buffers (" + std::to_string(need) +
                                                         " bytes) do not fit");
        }
        b.bytes = need;
        e->slot_bytes_total += need;
    } else if (b.bytes < need) {
        if (b.ptr) {
            VZ_HIP_TRY(hipStreamSynchronize(st));
            (void)hipFree(b.ptr);
            e->slot_bytes_total -= b.bytes;
            b.ptr = nullptr;
            b.bytes = 0;
        }
        if (hipMalloc(&b.ptr, need) != hipSuccess)
            return set_error(VZ_ERROR_OUT_OF_MEMORY, "engine: request buffers (" + std::to_string(need) +
                                                         " bytes) do not fit");
        b.bytes = need;
        e->slot_bytes_total += need;
    }
    char* base = (char*)b.ptr;
    size_t npix = D.npix;
    bool fp64 = (p.flags & VZ_ACCUM_FP64) != 0;
    D.passbuf = (float*)base;
    base += align_up((size_t)D.K * 3 * npix * 4, 256);
    if (fp64) D.sum64 = (double*)base; else D.sum32 = (float*)base;
    base += align_up(3 * npix * (fp64 ? 8 : 4), 256);
    D.camseed = (uint32_t*)base;
    base += align_up(npix * 4, 256);
    D.done = (uint32_t*)base;
    base += align_up((size_t)D.K * 4, 256);
    D.segsum = (unsigned long long*)base;
    VZ_HIP_TRY(hipMemsetAsync(D.done, 0, (size_t)D.K * 4, st));
    VZ_HIP_TRY(hipMemsetAsync(D.segsum, 0, (size_t)D.K * 8, st));
    ReqState& S = e->h_state[slot];
    std::memset(&S, 0, sizeof(S));
    S.active = 1u;
    return VZ_SUCCESS;
}

static vzStatus upload_scenes(Engine* e, const std::vector<const vzScene_st*>& scenes, hipStream_t st)
{
    uint32_t ns = (uint32_t)scenes.size();
    if (ns > kMaxScenes) return set_error(VZ_ERROR_CAPACITY, "too many distinct scenes in one batch");
    for (uint32_t i = 0; i < ns; i++) e->h_scenes[i] = scenes[i]->view();
    VZ_HIP_TRY(hipMemcpyAsync(e->d_scenes, e->h_scenes, ns * sizeof(SceneView), hipMemcpyHostToDevice, st));
    return VZ_SUCCESS;
}

// Partitions the queue into NB_res * f slices of R_base / f rays (f = 1, 2 or
// 4, halved while NB would exceed kMaxBlocks).  More slices than resident
// step workgroups let the dispatcher balance slices of uneven cost (many
// scenes in one run); one per workgroup costs least when every ray costs
// about the same.
static void set_split(Engine* e, uint32_t f)
{
    f = f >= 8u ? 8u : (f >= 4u ? 4u : (f >= 2u ? 2u : 1u));
    while (f > 1u && (size_t)e->NB_res * f > kMaxBlocks) f >>= 1;
    e->split = f;
    e->NB = e->NB_res * f;
    e->R = e->R_base / f;
}

vzStatus engine_run(Engine* e, const std::vector<RunItem>& items, hipStream_t st, const RunOptions& opt,
                    RunResult* res)
{
    uint32_t n = (uint32_t)items.size();
    if (n == 0) return VZ_SUCCESS;
    if (n > kMaxRequests) return set_error(VZ_ERROR_CAPACITY, "too many requests in one batch");
    if (e->stream_alloc) return set_error(VZ_ERROR_INTERNAL, "engine is in serving mode");
    // the whole table is reset: wait for anything still using it
    VZ_HIP_TRY(hipStreamSynchronize(st));
    std::vector<const vzScene_st*> scenes;
    std::vector<uint32_t> scene_idx(n);
    for (uint32_t i = 0; i < n; i++) {
        auto it = std::find(scenes.begin(), scenes.end(), items[i].scene);
        if (it == scenes.end()) {
            scene_idx[i] = (uint32_t)scenes.size();
            scenes.push_back(items[i].scene);
        } else {
            scene_idx[i] = (uint32_t)(it - scenes.begin());
        }
    }
    set_split(e, opt.slice_split);
    e->ctl_small = n <= 256u && e->NB <= 1024u;
    VZ_TRY(upload_scenes(e, scenes, st));
    for (uint32_t i = 0; i < n; i++) VZ_TRY(setup_slot(e, i, scene_idx[i], items[i].params, i, st));
    VZ_HIP_TRY(hipMemcpyAsync(e->d_req, e->h_req, n * sizeof(ReqDesc), hipMemcpyHostToDevice, st));
    VZ_HIP_TRY(hipMemcpyAsync(e->d_state, e->h_state, n * sizeof(ReqState), hipMemcpyHostToDevice, st));
    Counters c0;
    std::memset(&c0, 0, sizeof(c0));
    c0.n_slots = n;
    c0.active = n;
    if (opt.rays_in_flight)
        c0.fill = (uint32_t)std::min<uint64_t>(e->R, std::max<uint64_t>(64, (opt.rays_in_flight + e->NB - 1) / e->NB));
    e->h_cnt[0] = c0;
    VZ_HIP_TRY(hipMemcpyAsync(e->d_cnt, &e->h_cnt[0], sizeof(Counters), hipMemcpyHostToDevice, st));
    VZ_HIP_TRY(hipMemsetAsync(e->d_bcnt, 0, 2 * e->NB * sizeof(uint32_t), st));
    VZ_HIP_TRY(hipMemsetAsync(e->d_bstats, 0, e->NB * sizeof(BlockStats), st));
    VZ_HIP_TRY(hipStreamSynchronize(st));  // staging buffers are reused below

    const bool use_graph = opt.use_graph && !opt.time_trace;
    // a single small scene is staged in LDS by every step workgroup
    const bool single = scenes.size() == 1;
    const bool lds = single && scenes[0]->n >= 2 && scenes[0]->nw <= kLdsNodes;
    if (use_graph) VZ_TRY(ensure_graph(e, opt.count_visits, lds, single));
    const uint32_t B = e->cfg.iters_per_graph;
    Event t0, t1, blk0(hipEventDisableTiming), blk1(hipEventDisableTiming);
    VZ_EVENT_TRY(t0);
    VZ_EVENT_TRY(t1);
    VZ_EVENT_TRY(blk0);
    VZ_EVENT_TRY(blk1);
    const hipEvent_t blk_ev[2] = {blk0.e, blk1.e};
    std::vector<Event> tev[2];  // per block parity: a pair of events around each step launch
    if (opt.time_trace) {
        for (int k = 0; k < 2; k++) {
            tev[k].reserve(2 * B);
            for (uint32_t i = 0; i < 2 * B; i++) {
                tev[k].emplace_back();
                VZ_EVENT_TRY(tev[k].back());
            }
        }
    }
    double trace_ms = 0.0;
    uint32_t trace_launches = 0;
    auto launch_block = [&](uint32_t k) -> vzStatus {
        if (use_graph) {
            VZ_TRY(launch_graph(e, st));
        } else {
            for (uint32_t it = 0; it < B; it++) {
                hipEvent_t a = opt.time_trace ? tev[k & 1][2 * it].e : nullptr;
                hipEvent_t b = opt.time_trace ? tev[k & 1][2 * it + 1].e : nullptr;
                enqueue_iteration(e, st, (int)(it & 1u), opt.count_visits, lds, single, a, b);
            }
            VZ_HIP_TRY(hipGetLastError());
        }
        VZ_HIP_TRY(hipMemcpyAsync(&e->h_cnt[k & 1], e->d_cnt, sizeof(Counters), hipMemcpyDeviceToHost, st));
        VZ_HIP_TRY(hipEventRecord(blk_ev[k & 1], st));
        return VZ_SUCCESS;
    };
    auto harvest = [&](uint32_t k) -> vzStatus {
        VZ_HIP_TRY(hipEventSynchronize(blk_ev[k & 1]));
        if (opt.time_trace) {
            for (uint32_t it = 0; it < B; it++) {
                float ms = 0.0f;
                VZ_HIP_TRY(hipEventElapsedTime(&ms, tev[k & 1][2 * it].e, tev[k & 1][2 * it + 1].e));
                trace_ms += ms;
                trace_launches++;
            }
        }
        return VZ_SUCCESS;
    };
    VZ_HIP_TRY(hipEventRecord(t0.e, st));
    vzStatus rs = launch_block(0);
    uint32_t k = 0;
    const uint32_t max_blocks = 1u << 20;
    while (rs == VZ_SUCCESS) {
        rs = launch_block(k + 1);
        if (rs != VZ_SUCCESS) break;
        rs = harvest(k);
        if (rs != VZ_SUCCESS) break;
        bool done = e->h_cnt[k & 1].active == 0 && e->h_cnt[k & 1].iteration > 0;
        k++;
        if (done) {
            rs = harvest(k);
            break;
        }
        if (k >= max_blocks) {
            rs = set_error(VZ_ERROR_INTERNAL, "wavefront did not converge");
            break;
        }
    }
    if (rs == VZ_SUCCESS && (hipEventRecord(t1.e, st) != hipSuccess || hipEventSynchronize(t1.e) != hipSuccess))
        rs = set_error(VZ_ERROR_HIP, "hipEventRecord/Synchronize failed");
    if (rs != VZ_SUCCESS) (void)hipStreamSynchronize(st);  // nothing in flight uses the events below
    VZ_TRY(rs);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, t0.e, t1.e);
    Counters fin;
    VZ_HIP_TRY(hipMemcpy(&fin, e->d_cnt, sizeof(Counters), hipMemcpyDeviceToHost));
    VZ_HIP_TRY(hipMemcpy(e->h_state, e->d_state, n * sizeof(ReqState), hipMemcpyDeviceToHost));
    VZ_HIP_TRY(hipMemcpy(e->h_bstats, e->d_bstats, e->NB * sizeof(BlockStats), hipMemcpyDeviceToHost));
    if (res) {
        *res = RunResult();
        for (uint32_t b = 0; b < e->NB; b++) {
            res->segments += e->h_bstats[b].segments;
            res->hits += e->h_bstats[b].hits;
            res->node_visits += e->h_bstats[b].node_visits;
            res->sphere_tests += e->h_bstats[b].sphere_tests;
            res->paths += e->h_bstats[b].paths;
        }
        res->render_ms = ms;
        res->trace_ms = trace_ms;
        res->trace_launches = trace_launches;
        res->iterations = fin.iteration;
        res->req_segments.resize(n);
        res->req_finish_iter.resize(n);
        for (uint32_t i = 0; i < n; i++) {
            res->req_segments[i] = e->h_state[i].segments;
            res->req_finish_iter[i] = e->h_state[i].finish_iter;
        }
    }
    return VZ_SUCCESS;
}

// ---- serving -----------------------------------------------------------------------
vzStatus engine_serve_init(Engine* e, hipStream_t st, uint32_t slice_split)
{
    set_split(e, slice_split);
    e->ctl_small = false;  // up to kMaxRequests request slots
    e->stream_alloc = true;
    e->slot_used.assign(kMaxRequests, 0);
    e->slot_scene.assign(kMaxRequests, 0);
    e->n_slots = 0;
    e->next_seq = 0;
    e->scene_uid.assign(kMaxScenes, 0);
    e->scene_refs.assign(kMaxScenes, 0);
    e->scene_last_use.assign(kMaxScenes, 0);
    e->scene_clock = 0;
    if (!e->h_state_stage) VZ_HIP_TRY(hipHostMalloc(&e->h_state_stage, kMaxRequests * sizeof(ReqState)));
    if (!e->h_status) VZ_HIP_TRY(hipHostMalloc(&e->h_status, 2 * kMaxRequests * sizeof(ReqState)));
    if (!e->h_nslots) VZ_HIP_TRY(hipHostMalloc(&e->h_nslots, 64));
    std::memset(e->h_status, 0, 2 * kMaxRequests * sizeof(ReqState));
    VZ_HIP_TRY(hipMemsetAsync(e->d_state, 0, kMaxRequests * sizeof(ReqState), st));
    VZ_HIP_TRY(hipMemsetAsync(e->d_cnt, 0, sizeof(Counters), st));
    VZ_HIP_TRY(hipMemsetAsync(e->d_bcnt, 0, 2 * e->NB * sizeof(uint32_t), st));
    VZ_HIP_TRY(hipMemsetAsync(e->d_bstats, 0, e->NB * sizeof(BlockStats), st));
    VZ_TRY(ensure_graph(e, false, false, false));
    VZ_HIP_TRY(hipStreamSynchronize(st));
    return VZ_SUCCESS;
}

// Scene entry for a request of `scene`: the cached one if the scene has one,
// else an unused entry, else the least recently used entry that no admitted
// request references.  An entry without references belongs to no ray in the
// queue (a request is released only once all its paths have ended), so
// rewriting its view in stream order is safe.
static vzStatus scene_entry(Engine* e, const vzScene_st* scene, hipStream_t st, uint32_t* out)
{
    uint32_t pick = kMaxScenes;
    uint64_t oldest = ~0ull;
    for (uint32_t i = 0; i < kMaxScenes; i++) {
        if (e->scene_uid[i] == scene->uid) {
            *out = i;
            return VZ_SUCCESS;
        }
        if (e->scene_refs[i] == 0u && e->scene_last_use[i] < oldest) {  // unused entries have last_use 0
            oldest = e->scene_last_use[i];
            pick = i;
        }
    }
    if (pick == kMaxScenes) return set_error(VZ_ERROR_CAPACITY, "scene table full");
    e->scene_uid[pick] = scene->uid;
    e->h_scenes[pick] = scene->view();
    VZ_HIP_TRY(hipMemcpyAsync(e->d_scenes + pick, e->h_scenes + pick, sizeof(SceneView), hipMemcpyHostToDevice, st));
    *out = pick;
    return VZ_SUCCESS;
}

vzStatus engine_admit(Engine* e, const RunItem& item, hipStream_t st, uint32_t* slot_out)
{
    uint32_t slot = kMaxRequests;
    for (uint32_t i = 0; i < kMaxRequests; i++)
        if (!e->slot_used[i]) {
            slot = i;
            break;
        }
    if (slot == kMaxRequests) return set_error(VZ_ERROR_CAPACITY, "request table full");
    uint32_t sidx = 0;
    VZ_TRY(scene_entry(e, item.scene, st, &sidx));
    VZ_TRY(setup_slot(e, slot, sidx, item.params, e->next_seq++, st));
    e->h_state_stage[slot] = e->h_state[slot];
    VZ_HIP_TRY(hipMemcpyAsync(e->d_req + slot, e->h_req + slot, sizeof(ReqDesc), hipMemcpyHostToDevice, st));
    VZ_HIP_TRY(hipMemcpyAsync(e->d_state + slot, e->h_state_stage + slot, sizeof(ReqState), hipMemcpyHostToDevice,
                              st));
    e->slot_used[slot] = 1;
    e->slot_scene[slot] = sidx;
    e->scene_refs[sidx]++;
    e->scene_last_use[sidx] = ++e->scene_clock;
    if (slot + 1 > e->n_slots) {
        e->n_slots = slot + 1;
        // the slot count only grows while serving, so a later value read by
        // an earlier copy is still correct
        *e->h_nslots = e->n_slots;
        VZ_HIP_TRY(hipMemcpyAsync(&e->d_cnt->n_slots, e->h_nslots, sizeof(uint32_t), hipMemcpyHostToDevice, st));
    }
    *slot_out = slot;
    return VZ_SUCCESS;
}

vzStatus engine_step(Engine* e, hipStream_t st, uint32_t k)
{
    VZ_TRY(launch_graph(e, st));
    if (e->n_slots)
        VZ_HIP_TRY(hipMemcpyAsync(e->h_status + (k & 1u) * kMaxRequests, e->d_state, e->n_slots * sizeof(ReqState),
                                  hipMemcpyDeviceToHost, st));
    return VZ_SUCCESS;
}

void engine_slot_status(Engine* e, uint32_t k, uint32_t slot, bool* active, unsigned long long* segments)
{
    const ReqState& s = e->h_status[(k & 1u) * kMaxRequests + slot];
    *active = s.active != 0;
    *segments = s.segments;
}

vzStatus engine_release_slot(Engine* e, uint32_t slot, hipStream_t st)
{
    DeviceBlock& b = e->slot_mem[slot];
    if (b.ptr) {
        VZ_HIP_TRY(hipFreeAsync(b.ptr, st));
        e->slot_bytes_total -= b.bytes;
        b.ptr = nullptr;
        b.bytes = 0;
    }
    if (e->slot_used[slot] && !e->slot_scene.empty()) {
        const uint32_t sidx = e->slot_scene[slot];
        if (e->scene_refs[sidx]) e->scene_refs[sidx]--;
    }
    e->slot_used[slot] = 0;
    return VZ_SUCCESS;
}

uint32_t engine_scenes_referenced(const Engine* e)
{
    uint32_t n = 0;
    for (uint32_t r : e->scene_refs) n += r ? 1u : 0u;
    return n;
}

}  // namespace vz
