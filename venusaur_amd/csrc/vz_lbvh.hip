// vz_lbvh.hip -- LBVH build on the GPU, replacing optixAccelBuild over
// custom-primitive AABBs (Core/Renderer.h:160-255).
//
// Published methods: Lauterbach et al. 2009 (LBVH from Morton codes),
// Karras 2012 "Maximizing Parallelism in the Construction of BVHs, Octrees,
// and k-d Trees" (parent linking, atomic-flag bottom-up boxes).  The output
// is bit-identical to the CPU build in oracle/: same Morton quantisation,
// stable sort, Karras topology with index tie-breaks, min/max boxes and node
// quantisation.
//
// Stages (all on one stream):
//   1. centre bounds       two-pass min/max reduction
//   2. Morton codes        10 bits per axis of the centre in the bounds
//   3. radix sort          4 passes of 8-bit digits; per-tile histograms,
//                          one scan, stable scatter ranked with wave ballots
//                          (match-any over the digit bits + mbcnt) and staged
//                          through LDS so global writes are runs per digit
//   4. leaf reorder        spheres/materials into Morton (leaf) order
//   5. Karras linking      one thread per internal node
//   6. refit + pack        one thread per leaf walks up; the second arrival at
//                          a node computes its box and writes the 32-byte node
//   7. 4-wide collapse     level by level, then a structural check
// A scene of 2 to kSahMaxPrims spheres replaces stages 3 and 5 by a
// binned-SAH build in one workgroup (k_sah_build) that writes the same node
// layout, so stages 4, 6 and 7 are shared.
#include "vz_device.h"
#include "vz_internal.h"

namespace vz {
namespace {

constexpr int kBlk = 256;
constexpr int kItems = 8;
constexpr int kTile = kBlk * kItems;  // keys per radix tile
constexpr uint32_t kLocalBit = 0x80000000u;  // parent link flag (see k_karras)
constexpr uint32_t kLargeKey = 0x40000000u;  // sort-key bit of a large sphere (see k_morton)
constexpr uint32_t kRefitLeaves = 2048;       // leaves per k_refit_local workgroup

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint32_t mbcnt(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// ---- 1. bounds ---------------------------------------------------------------
// Centroid bounds, per workgroup.  With `grid_a` set only the spheres at or
// below its size threshold (2|r| <= grid_a[9]) count.
__global__ void __launch_bounds__(kBlk) k_bounds_partial(const float4* __restrict__ sph, uint32_t n,
                                                         const float* __restrict__ grid_a,
                                                         float* __restrict__ partial)
{
    __shared__ float s[6][kBlk];
    float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    const float thr = grid_a ? grid_a[9] : 0.0f;
    for (uint32_t i = blockIdx.x * kBlk + threadIdx.x; i < n; i += gridDim.x * kBlk) {
        float4 c = sph[i];
        if (grid_a && 2.0f * fabsf(c.w) > thr) continue;
        lo[0] = fminf(lo[0], c.x); lo[1] = fminf(lo[1], c.y); lo[2] = fminf(lo[2], c.z);
        hi[0] = fmaxf(hi[0], c.x); hi[1] = fmaxf(hi[1], c.y); hi[2] = fmaxf(hi[2], c.z);
    }
    for (int k = 0; k < 3; k++) { s[k][threadIdx.x] = lo[k]; s[3 + k][threadIdx.x] = hi[k]; }
    __syncthreads();
    for (int w = kBlk / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int k = 0; k < 3; k++) {
                s[k][threadIdx.x] = fminf(s[k][threadIdx.x], s[k][threadIdx.x + w]);
                s[3 + k][threadIdx.x] = fmaxf(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// out: lo[3], hi[3], scale[3] with scale = ext > 0 ? 1024/ext : 0 (iso: one
// scale 1024/max ext for every axis), and the large-sphere threshold
// out[9] = max extent / 4; an axis that no sphere reached gets lo = hi = 0.  min and max are exact, so the reduction order
// does not change the result.
__global__ void __launch_bounds__(kBlk) k_bounds_final(const float* __restrict__ partial, uint32_t nb,
                                                       uint32_t iso, float* __restrict__ out)
{
    __shared__ float s[6][kBlk];
    float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (uint32_t b = threadIdx.x; b < nb; b += kBlk)
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], partial[b * 6 + k]);
            hi[k] = fmaxf(hi[k], partial[b * 6 + 3 + k]);
        }
    for (int k = 0; k < 3; k++) { s[k][threadIdx.x] = lo[k]; s[3 + k][threadIdx.x] = hi[k]; }
    __syncthreads();
    for (int w = kBlk / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; k++) {
                s[k][threadIdx.x] = fminf(s[k][threadIdx.x], s[k][threadIdx.x + w]);
                s[3 + k][threadIdx.x] = fmaxf(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float extmax = 0.0f;
        for (int k = 0; k < 3; k++) {
            float l = s[k][0], h = s[3 + k][0];
            if (l > h) l = h = 0.0f;
            const float ext = h - l;
            out[k] = l;
            out[3 + k] = h;
            out[6 + k] = ext > 0.0f ? 1024.0f / ext : 0.0f;
            extmax = fmaxf(extmax, ext);
        }
        if (iso)  // one scale for every axis: cubic Morton cells
            for (int k = 0; k < 3; k++) out[6 + k] = extmax > 0.0f ? 1024.0f / extmax : 0.0f;
        out[9] = 0.25f * extmax;
    }
}

// ---- 2. Morton ---------------------------------------------------------------
__device__ __forceinline__ uint32_t expand10(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// Sort key of sphere i.  A sphere whose diameter exceeds grid_a's threshold
// (a quarter of the largest centroid extent, e.g. the r = 1000 ground) is
// "large": its key is its Morton code on grid A (all centroids) with bit 30
// set, so the first split of the tree separates it from the rest and its box
// does not inflate a whole root-to-leaf path.  Every other sphere gets its
// Morton code on grid B: the centroids of the small spheres only, with cubic
// cells, so a thin slab of spheres is not split along its thin axis as often
// as along the others.
__global__ void __launch_bounds__(kBlk) k_morton(const float4* __restrict__ sph, uint32_t n,
                                                 const float* __restrict__ grid_a,
                                                 const float* __restrict__ grid_b,
                                                 uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    uint32_t i = blockIdx.x * kBlk + threadIdx.x;
    if (i >= n) return;
    float4 c = sph[i];
    const bool large = 2.0f * fabsf(c.w) > grid_a[9];
    const float* bnd = large ? grid_a : grid_b;
    float cc[3] = {c.x, c.y, c.z};
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
        float f = (cc[k] - bnd[k]) * bnd[6 + k];
        int qi = (int)f;
        qi = qi < 0 ? 0 : (qi > 1023 ? 1023 : qi);
        q[k] = (uint32_t)qi;
    }
    keys[i] = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]) | (large ? kLargeKey : 0u);
    vals[i] = i;
}

// ---- 3. radix sort -------------------------------------------------------------
// Exclusive scan of one value per thread over a kBlk-thread workgroup.
__device__ __forceinline__ uint32_t block_scan_excl(uint32_t v, uint32_t* s)
{
    s[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t off = 1; off < (uint32_t)kBlk; off <<= 1) {
        const uint32_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    const uint32_t r = s[threadIdx.x] - v;
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kBlk) k_radix_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                     int shift, uint32_t nblocks,
                                                     uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kTile;
    for (int r = 0; r < kItems; r++) {
        uint32_t i = base + r * kBlk + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Per digit (one workgroup each): exclusive scan in place of the digit's
// per-tile counts, row total to dtot[digit].  The scatter adds the digit's
// base, the exclusive scan of dtot.
__global__ void __launch_bounds__(kBlk) k_scan_rows(uint32_t* __restrict__ hist, uint32_t nblocks,
                                                    uint32_t* __restrict__ dtot)
{
    __shared__ uint32_t part[kBlk];
    uint32_t* row = hist + (size_t)blockIdx.x * nblocks;
    const uint32_t chunk = (nblocks + kBlk - 1) / kBlk;
    const uint32_t b = threadIdx.x * chunk, e = min(b + chunk, nblocks);
    uint32_t sum = 0;
    for (uint32_t i = b; i < e; i++) sum += row[i];
    const uint32_t excl = block_scan_excl(sum, part);
    uint32_t run = excl;
    for (uint32_t i = b; i < e; i++) {
        const uint32_t v = row[i];
        row[i] = run;
        run += v;
    }
    if (threadIdx.x == kBlk - 1) dtot[blockIdx.x] = excl + sum;
}

__global__ void __launch_bounds__(kBlk) k_radix_scatter(const uint32_t* __restrict__ kin,
                                                        const uint32_t* __restrict__ vin, uint32_t n,
                                                        int shift, uint32_t nblocks,
                                                        const uint32_t* __restrict__ hist,
                                                        const uint32_t* __restrict__ dtot,
                                                        uint32_t* __restrict__ kout,
                                                        uint32_t* __restrict__ vout)
{
    __shared__ uint32_t s_dbase[256];   // first output slot of each digit
    __shared__ uint32_t s_count[256];   // running per-digit count in this tile
    __shared__ uint32_t s_start[256];   // tile-local start of each digit
    __shared__ uint32_t s_wcnt[4][256];
    __shared__ uint32_t s_wpre[4][256];
    __shared__ uint32_t s_key[kTile];
    __shared__ uint32_t s_val[kTile];
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = lane_id();
    s_dbase[tid] = block_scan_excl(dtot[tid], s_count);  // s_count as scratch
    s_count[tid] = 0;
    for (int w = 0; w < 4; w++) s_wcnt[w][tid] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kTile;
    uint32_t key[kItems], val[kItems], rank[kItems];
    for (int r = 0; r < kItems; r++) {
        uint32_t i = base + r * kBlk + tid;
        bool valid = i < n;
        key[r] = valid ? kin[i] : 0u;
        val[r] = valid ? vin[i] : 0u;
        uint32_t d = (key[r] >> shift) & 0xFFu;
        // lanes of this wave holding the same digit (match-any over 8 bits)
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            uint64_t m = __ballot((d >> bit) & 1u);
            peers &= ((d >> bit) & 1u) ? m : ~m;
        }
        uint32_t rk = mbcnt(peers);
        if (valid && (uint32_t)(__ffsll((unsigned long long)peers) - 1) == lane)
            s_wcnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {   // thread t owns digit t: prefix over the 4 waves, in wave order
            uint32_t run = s_count[tid];
            for (int w = 0; w < 4; w++) {
                uint32_t c = s_wcnt[w][tid];
                s_wpre[w][tid] = run;
                run += c;
                s_wcnt[w][tid] = 0;
            }
            s_count[tid] = run;
        }
        __syncthreads();
        rank[r] = valid ? s_wpre[wave][d] + rk : 0xFFFFFFFFu;
        __syncthreads();
    }
    // exclusive scan of the tile's digit counts (256 entries, one per thread)
    {
        uint32_t v = s_count[tid];
        s_start[tid] = v;
        __syncthreads();
        for (uint32_t off = 1; off < 256; off <<= 1) {
            uint32_t t = tid >= off ? s_start[tid - off] : 0u;
            __syncthreads();
            s_start[tid] += t;
            __syncthreads();
        }
        s_start[tid] -= v;
        __syncthreads();
    }
    for (int r = 0; r < kItems; r++) {
        if (rank[r] != 0xFFFFFFFFu) {
            uint32_t d = (key[r] >> shift) & 0xFFu;
            uint32_t pos = s_start[d] + rank[r];
            s_key[pos] = key[r];
            s_val[pos] = val[r];
        }
    }
    __syncthreads();
    uint32_t cnt = n - base < (uint32_t)kTile ? n - base : (uint32_t)kTile;
    for (uint32_t i = tid; i < cnt; i += kBlk) {
        uint32_t k = s_key[i];
        uint32_t d = (k >> shift) & 0xFFu;
        uint32_t g = s_dbase[d] + hist[d * nblocks + blockIdx.x] + (i - s_start[d]);
        kout[g] = k;
        vout[g] = s_val[i];
    }
}

// ---- 4. leaf reorder -----------------------------------------------------------
__global__ void __launch_bounds__(kBlk) k_reorder(const float4* __restrict__ sph_in,
                                                  const vzMaterial* __restrict__ mat_in,
                                                  const uint32_t* __restrict__ perm, uint32_t n,
                                                  float4* __restrict__ sph, uint32_t* __restrict__ info,
                                                  float4* __restrict__ mat)
{
    uint32_t i = blockIdx.x * kBlk + threadIdx.x;
    if (i >= n) return;
    uint32_t p = perm[i];
    sph[i] = sph_in[p];
    vzMaterial m = mat_in[p];
    uint32_t type = m.type > 2u ? 0u : m.type;
    info[i] = p | (type << 30);
    float param = type == 1u ? m.fuzz : (type == 2u ? m.ir : 0.0f);
    mat[i] = make_float4(m.albedo[0], m.albedo[1], m.albedo[2], param);
}

// refit reads spheres in leaf order but geometry updates arrive in prim order
__global__ void __launch_bounds__(kBlk) k_reorder_sph(const float4* __restrict__ sph_in,
                                                      const uint32_t* __restrict__ perm, uint32_t n,
                                                      float4* __restrict__ sph)
{
    uint32_t i = blockIdx.x * kBlk + threadIdx.x;
    if (i < n) sph[i] = sph_in[perm[i]];
}

// ---- 5. Karras linking -----------------------------------------------------------
__device__ __forceinline__ int kdelta(const uint32_t* __restrict__ k, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
    return __clz((int)(a ^ b));
}

__global__ void __launch_bounds__(kBlk) k_karras(const uint32_t* __restrict__ K, uint32_t n_,
                                                 uint32_t* __restrict__ child,
                                                 uint32_t* __restrict__ parent_int,
                                                 uint32_t* __restrict__ parent_leaf,
                                                 uint32_t* __restrict__ visit, uint2* __restrict__ range)
{
    int n = (int)n_;
    int i = (int)(blockIdx.x * kBlk + threadIdx.x);
    if (i >= n - 1) return;
    visit[i] = 0u;
    if (i == 0) parent_int[0] = 0xFFFFFFFFu;
    int d = (kdelta(K, n, i, i + 1) - kdelta(K, n, i, i - 1)) > 0 ? 1 : -1;
    int dmin = kdelta(K, n, i, i - d);
    int lmax = 2;
    while (kdelta(K, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (kdelta(K, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = kdelta(K, n, i, j);
    int s = 0, div = 2;
    int t = (l + div - 1) / div;
    for (;;) {
        if (kdelta(K, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
        div *= 2;
        t = (l + div - 1) / div;
    }
    int g = i + s * d + (d < 0 ? d : 0);
    int mn = min(i, j), mx = max(i, j);
    uint32_t f = 0;
    // bit 31 of a parent link: every leaf of the parent's range lies in one
    // refit workgroup, so the refit synchronises on it at workgroup scope
    const uint32_t up = (uint32_t)i | ((mn / (int)kRefitLeaves == mx / (int)kRefitLeaves) ? kLocalBit : 0u);
    if (mn == g) { f |= 1u; parent_leaf[g] = up; } else { parent_int[g] = up; }
    if (mx == g + 1) { f |= 2u; parent_leaf[g + 1] = up; } else { parent_int[g + 1] = up; }
    child[i] = (uint32_t)g | (f << 30);
    range[i] = make_uint2((uint32_t)mn, (uint32_t)mx);
}

// ---- 5'. binned-SAH build of a small scene -------------------------------------------
// For 2 .. kSahMaxPrims spheres the Morton order serves large spheres badly
// (the book scene's ground and its three big spheres), so the binary tree is
// built top down by the surface-area heuristic instead.  A range [a, b] of
// the leaf permutation at depth < kSahDepthCap: per axis whose centre extent
// ext is positive, sphere centre c falls in bin min(31, (int)(((c - lo) /
// ext) * 32)); the split after bin j costs area(L) * nL + area(R) * nR over
// the leaf-box unions and counts of bins <= j and > j (both non-empty); the
// cheapest wins, lower axis then lower bin on a tie, and the range is
// partitioned stably.  With no positive extent, or below the depth cap, the
// range splits after its first floor(m / 2) spheres.  The cap bounds the
// tree at kSahDepthCap + 10 internal levels for 1024 spheres, so the 4-wide
// tree stays within the traversal's kMaxWideDepth.
// Nodes use the Karras layout (children g and g + 1 for a split after leaf
// position g, the root is node 0), so refit, collapse and traversal are the
// LBVH's.  One workgroup, level by level: each wave takes one range of the
// level per round.  Bin boxes are kept as order-preserving integer keys and
// merged with LDS integer min/max, counts with adds, so the bins do not
// depend on the order lanes ran in; the sweep and the partition are
// sequential in bin and leaf order.  The result equals the oracle's
// sah_range bit for bit.
constexpr uint32_t kSahMaxPrims = 1024;
constexpr int kSahBins = 32;
constexpr uint32_t kSahDepthCap = 24;
// An LBVH of at least kTreeletMinPrims spheres has its treelets rebuilt by
// the same SAH build: every maximal subtree of at most kSahMaxPrims leaves
// whose root lies above kSahDepthCap, one workgroup each, splits counting
// depth from the whole tree's root.  On box scenes of a few thousand
// spheres SAH treelets visited more nodes than Morton order; from about 2e5
// spheres up they visit 4-6% fewer.
constexpr uint32_t kTreeletMinPrims = 131072;
constexpr uint32_t kTreeletBlocks = 2048;  // grid of k_sah_treelets (grid-stride over the treelets)
constexpr int kSahWaves = 8;
constexpr int kSahThreads = 64 * kSahWaves;
constexpr int kSahBinWords = 3 * kSahBins * 7;  // per wave: [axis][bin] lo xyz, hi xyz keys, count
static_assert(kSahMaxPrims <= kRefitLeaves, "every SAH range lies in one refit workgroup");
static_assert(kSahWaves * kSahBinWords >= (int)kSahMaxPrims, "the bins hold a treelet's keys at the end");
static_assert(kSahDepthCap + 10 + 1 <= kMaxWideDepth, "SAH tree depth within the traversal stack");

__device__ __forceinline__ uint32_t ord_key(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord_val(uint32_t k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}
__device__ __forceinline__ float sah_cen(const float4 s, int k) { return k == 0 ? s.x : (k == 1 ? s.y : s.z); }
__device__ __forceinline__ int sah_bin(float c, float lo, float ext)
{
    const int j = (int)(((c - lo) / ext) * (float)kSahBins);
    return j > kSahBins - 1 ? kSahBins - 1 : j;
}
__device__ __forceinline__ float box_area6(const float* b)
{
    const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return dx * dy + dy * dz + dz * dx;
}
__device__ __forceinline__ float wave_min(float v)
{
    for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v)
{
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// LDS of one treelet build: the range's spheres by local index (gathered
// once: every phase of every round reads them), the local permutation, two
// level lists of ranges, and per wave the bins.
struct SahShared {
    float4 sph[kSahMaxPrims];
    uint32_t perm[kSahMaxPrims], tmp[kSahMaxPrims];
    uint2 task[2][kSahMaxPrims / 2];  // (a | b << 10 local, node index)
    uint32_t ntask[2];
    uint32_t bin[kSahWaves][kSahBinWords];
};

// The SAH split of a range of two spheres [x, x + 1] (node `self`: the
// range's right end for a left child, its left end for a right child) at
// depth d, in closed form: the two leaf boxes are the two sides whatever the axis,
// so every axis of positive centre extent costs area(A) + area(B) alike and
// the first one wins; the sphere with the smaller centre on it goes left.
// With no such axis, or from the depth cap on, the order stays.  This is
// what the binned sweep gives for two spheres (they fall in bins 0 and 31).
__device__ __forceinline__ void sah_pair(uint32_t x, uint32_t f, uint32_t self, uint32_t d,
                                         uint32_t* __restrict__ child, uint32_t* __restrict__ parent_leaf,
                                         SahShared& sh)
{
    const uint32_t X = f + x;  // split after leaf X: both children leaves
    const float4 s0 = sh.sph[sh.perm[x]], s1 = sh.sph[sh.perm[x + 1u]];
    if (d < kSahDepthCap) {
        for (int k = 0; k < 3; k++) {
            const float c0 = sah_cen(s0, k), c1 = sah_cen(s1, k);
            if (c0 != c1) {
                if (c0 > c1) {
                    const uint32_t t = sh.perm[x];
                    sh.perm[x] = sh.perm[x + 1u];
                    sh.perm[x + 1u] = t;
                }
                break;
            }
        }
    }
    const uint32_t up = self | (((X / kRefitLeaves) == ((X + 1u) / kRefitLeaves)) ? kLocalBit : 0u);
    parent_leaf[X] = up;
    parent_leaf[X + 1u] = up;
    child[self] = X | (3u << 30);
}

// Rebuilds the subtree over leaves [f, f + m) (2 <= m <= kSahMaxPrims), root
// node r at depth d0, in place: leaf_prim and keys are permuted within the
// range, and the child words and parent links of the subtree's nodes are
// rewritten.  In the Karras layout a subtree's node indices are fixed by its
// leaf range and root index, so the nodes above it are untouched.  Called
// by a whole workgroup.
__device__ __forceinline__ void sah_treelet(const float4* __restrict__ sph, uint32_t f, uint32_t m, uint32_t r,
                                            uint32_t d0, uint32_t* __restrict__ leaf_prim,
                                            uint32_t* __restrict__ keys, uint32_t* __restrict__ child,
                                            uint32_t* __restrict__ parent_int, uint32_t* __restrict__ parent_leaf,
                                            SahShared& sh)
{
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    for (uint32_t i = tid; i < m; i += kSahThreads) {
        sh.sph[i] = sph[leaf_prim[f + i]];
        sh.perm[i] = i;
    }
    if (tid == 0) {
        sh.task[0][0] = make_uint2((m - 1u) << 10, r);
        sh.ntask[0] = 1u;
        sh.ntask[1] = 0u;
    }
    uint32_t cur = 0;
    for (uint32_t level = 0; level < 64u; level++) {
        __syncthreads();
        const uint32_t nt = sh.ntask[cur];
        if (nt == 0u) break;
        for (uint32_t round = 0; round < nt; round += kSahWaves) {
            const uint32_t t = round + w;
            const bool has = t < nt;
            const uint2 tw = has ? sh.task[cur][t] : make_uint2(0u, 0u);
            const uint32_t a = tw.x & 0x3FFu, b = (tw.x >> 10) & 0x3FFu, self = tw.y;
            // centre bounds of the range
            float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
            if (has)
                for (uint32_t i = a + lane; i <= b; i += 64u) {
                    const float4 sp = sh.sph[sh.perm[i]];
                    for (int k = 0; k < 3; k++) {
                        lo[k] = fminf(lo[k], sah_cen(sp, k));
                        hi[k] = fmaxf(hi[k], sah_cen(sp, k));
                    }
                }
            float ext[3];
            bool any = false;
            for (int k = 0; k < 3; k++) {
                lo[k] = wave_min(lo[k]);
                hi[k] = wave_max(hi[k]);
                ext[k] = hi[k] - lo[k];
                any = any || ext[k] > 0.0f;
            }
            const bool sah = has && d0 + level < kSahDepthCap && any;
            if (sah)
                for (int x = (int)lane; x < kSahBinWords; x += 64) sh.bin[w][x] = x % 7 < 3 ? 0xFFFFFFFFu : 0u;
            __syncthreads();
            if (sah)
                for (uint32_t i = a + lane; i <= b; i += 64u) {
                    const float4 sp = sh.sph[sh.perm[i]];
                    const float rr = fabsf(sp.w);
                    const float lb[6] = {sp.x - rr, sp.y - rr, sp.z - rr, sp.x + rr, sp.y + rr, sp.z + rr};
                    for (int k = 0; k < 3; k++) {
                        if (!(ext[k] > 0.0f)) continue;
                        uint32_t* bn = &sh.bin[w][(k * kSahBins + sah_bin(sah_cen(sp, k), lo[k], ext[k])) * 7];
                        for (int q = 0; q < 3; q++) {
                            atomicMin(&bn[q], ord_key(lb[q]));
                            atomicMax(&bn[3 + q], ord_key(lb[3 + q]));
                        }
                        atomicAdd(&bn[6], 1u);
                    }
                }
            __syncthreads();
            // the cheapest split, per axis with one lane per bin: inclusive
            // prefix and suffix unions (min/max, exact in any order) and
            // counts over the 32 bins, the cost of the split after each bin,
            // then the first minimum.  The unions start from the box
            // (3.4e38, -3.4e38) as the sequential sweep of the oracle does.
            int bk = -1, bj = -1;
            if (sah) {
                float bc = 0.0f;
                const int j = (int)(lane & 31u);
                for (int k = 0; k < 3; k++) {
                    if (!(ext[k] > 0.0f)) continue;
                    const uint32_t* bn = &sh.bin[w][(k * kSahBins + j) * 7];
                    const uint32_t c = bn[6];
                    float P[6], S[6];
                    for (int q = 0; q < 3; q++) {
                        P[q] = S[q] = c ? ord_val(bn[q]) : 3.4e38f;
                        P[3 + q] = S[3 + q] = c ? ord_val(bn[3 + q]) : -3.4e38f;
                    }
                    uint32_t pc = c, sc = c;
                    for (int d = 1; d < kSahBins; d <<= 1) {
                        for (int q = 0; q < 6; q++) {
                            const float up = __shfl_up(P[q], d, kSahBins), dn = __shfl_down(S[q], d, kSahBins);
                            if (j >= d) P[q] = q < 3 ? fminf(P[q], up) : fmaxf(P[q], up);
                            if (j + d < kSahBins) S[q] = q < 3 ? fminf(S[q], dn) : fmaxf(S[q], dn);
                        }
                        const uint32_t upc = __shfl_up(pc, d, kSahBins), dnc = __shfl_down(sc, d, kSahBins);
                        if (j >= d) pc += upc;
                        if (j + d < kSahBins) sc += dnc;
                    }
                    float L[6], R[6];
                    for (int q = 0; q < 6; q++) {
                        const float r1 = __shfl_down(S[q], 1, kSahBins);  // bins j+1 .. 31
                        L[q] = q < 3 ? fminf(3.4e38f, P[q]) : fmaxf(-3.4e38f, P[q]);
                        R[q] = q < 3 ? fminf(3.4e38f, r1) : fmaxf(-3.4e38f, r1);
                    }
                    const uint32_t rn = __shfl_down(sc, 1, kSahBins);
                    bool ok = j < kSahBins - 1 && pc > 0u && rn > 0u;
                    float cost = ok ? box_area6(L) * (float)pc + box_area6(R) * (float)rn : 0.0f;
                    int cj = j;
                    for (int o = kSahBins / 2; o >= 1; o >>= 1) {
                        const bool ook = __shfl_xor(ok ? 1 : 0, o, kSahBins) != 0;
                        const float oc = __shfl_xor(cost, o, kSahBins);
                        const int oj = __shfl_xor(cj, o, kSahBins);
                        if ((ook && !ok) || (ook == ok && (oc < cost || (oc == cost && oj < cj)))) {
                            ok = ook;
                            cost = oc;
                            cj = oj;
                        }
                    }
                    if (ok && (bk < 0 || cost < bc)) {
                        bk = k;
                        bj = cj;
                        bc = cost;
                    }
                }
            }
            uint32_t g = a + (b - a + 1u) / 2u - 1u;
            if (bk >= 0) {
                uint32_t cl = 0;
                for (uint32_t i = a + lane; i <= b; i += 64u)
                    cl += sah_bin(sah_cen(sh.sph[sh.perm[i]], bk), lo[bk], ext[bk]) <= bj ? 1u : 0u;
                const uint32_t nl = wave_sum(cl);
                uint32_t offl = a, offr = a + nl;
                for (uint32_t i0 = a; i0 <= b; i0 += 64u) {
                    const uint32_t i = i0 + lane;
                    const bool valid = i <= b;
                    const uint32_t p = valid ? sh.perm[i] : 0u;
                    const bool left = valid && sah_bin(sah_cen(sh.sph[p], bk), lo[bk], ext[bk]) <= bj;
                    const uint64_t ml = __ballot(left), mr = __ballot(valid && !left);
                    if (left) sh.tmp[offl + mbcnt(ml)] = p;
                    else if (valid) sh.tmp[offr + mbcnt(mr)] = p;
                    offl += (uint32_t)__popcll(ml);
                    offr += (uint32_t)__popcll(mr);
                }
                g = a + nl - 1u;
            }
            __syncthreads();
            if (bk >= 0)
                for (uint32_t i = a + lane; i <= b; i += 64u) sh.perm[i] = sh.tmp[i];
            if (has && lane == 0u) {
                const uint32_t G = f + g;  // split after leaf G: children G and G + 1
                const uint32_t up = self | (((f + a) / kRefitLeaves == (f + b) / kRefitLeaves) ? kLocalBit : 0u);
                uint32_t fl = 0;
                if (g == a) { fl |= 1u; parent_leaf[G] = up; } else { parent_int[G] = up; }
                if (g + 1u == b) { fl |= 2u; parent_leaf[G + 1u] = up; } else { parent_int[G + 1u] = up; }
                child[self] = G | (fl << 30);
                // children of two spheres are finished here; larger ones are
                // the next level's ranges
                const uint32_t ca[2] = {a, g + 1u}, cb[2] = {g, b};
                for (int c = 0; c < 2; c++) {
                    if (cb[c] == ca[c]) continue;
                    const uint32_t node = f + (c == 0 ? g : g + 1u);  // Karras layout index
                    if (cb[c] == ca[c] + 1u)
                        sah_pair(ca[c], f, node, d0 + level + 1u, child, parent_leaf, sh);
                    else
                        sh.task[cur ^ 1u][atomicAdd(&sh.ntask[cur ^ 1u], 1u)] =
                            make_uint2(ca[c] | (cb[c] << 10), node);
                }
            }
        }
        __syncthreads();
        if (tid == 0) sh.ntask[cur] = 0u;
        cur ^= 1u;
    }
    __syncthreads();
    // permute the range's primitive ids and keys in place, staged through
    // LDS (the bins are free now) so every global read has landed before the
    // barrier that precedes the writes
    uint32_t* kbuf = &sh.bin[0][0];
    for (uint32_t i = tid; i < m; i += kSahThreads) {
        const uint32_t p = sh.perm[i];
        sh.tmp[i] = leaf_prim[f + p];
        kbuf[i] = keys[f + p];
    }
    __syncthreads();
    for (uint32_t i = tid; i < m; i += kSahThreads) {
        leaf_prim[f + i] = sh.tmp[i];
        keys[f + i] = kbuf[i];
    }
    __syncthreads();  // the LDS is reused by the next treelet
}

// A small scene: the whole tree is one SAH build (keys and identity
// permutation in sphere order in, leaf order out).
__global__ void __launch_bounds__(kSahThreads) k_sah_build(const float4* __restrict__ sph, uint32_t n,
                                                           uint32_t* __restrict__ leaf_prim,
                                                           uint32_t* __restrict__ keys,
                                                           uint32_t* __restrict__ child,
                                                           uint32_t* __restrict__ parent_int,
                                                           uint32_t* __restrict__ parent_leaf)
{
    __shared__ SahShared sh;
    if (threadIdx.x == 0) parent_int[0] = 0xFFFFFFFFu;
    sah_treelet(sph, 0u, n, 0u, 0u, leaf_prim, keys, child, parent_int, parent_leaf, sh);
}

// A large LBVH: the roots of its treelets, the maximal subtrees of at most
// kSahMaxPrims leaves whose root lies above kSahDepthCap.  Internal node i
// (leaves range[i]) is one when it is small enough and its parent is not;
// its depth is the length of its parent chain.  Appended as (first leaf,
// leaves, node, depth) in any order: the treelets are disjoint.
__global__ void __launch_bounds__(kBlk) k_treelet_roots(const uint2* __restrict__ range,
                                                        const uint32_t* __restrict__ parent_int, uint32_t ni,
                                                        uint4* __restrict__ list, uint32_t* __restrict__ count)
{
    const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
    if (i == 0u || i >= ni) return;
    const uint2 ri = range[i];
    const uint32_t m = ri.y - ri.x + 1u;
    const uint32_t p = parent_int[i] & ~kLocalBit;
    const uint2 rp = range[p];
    if (m > kSahMaxPrims || rp.y - rp.x + 1u <= kSahMaxPrims) return;
    uint32_t d = 1, q = p;
    while (q != 0u && d <= kMaxDepth) {
        q = parent_int[q] & ~kLocalBit;
        d++;
    }
    if (d >= kSahDepthCap) return;
    list[atomicAdd(count, 1u)] = make_uint4(ri.x, m, i, d);
}

__global__ void __launch_bounds__(kSahThreads) k_sah_treelets(const float4* __restrict__ sph,
                                                              const uint4* __restrict__ list,
                                                              const uint32_t* __restrict__ count,
                                                              uint32_t* __restrict__ leaf_prim,
                                                              uint32_t* __restrict__ keys,
                                                              uint32_t* __restrict__ child,
                                                              uint32_t* __restrict__ parent_int,
                                                              uint32_t* __restrict__ parent_leaf)
{
    __shared__ SahShared sh;
    const uint32_t cnt = *count;
    for (uint32_t t = blockIdx.x; t < cnt; t += gridDim.x) {
        const uint4 e = list[t];
        sah_treelet(sph, e.x, e.y, e.z, e.w, leaf_prim, keys, child, parent_int, parent_leaf, sh);
    }
}

// ---- 6. refit + pack ---------------------------------------------------------------
__device__ __forceinline__ uint32_t pick_exp(float lo, float hi)
{
    float ext = hi - lo;
    uint32_t e = 1;
    if (ext > 0.0f) {
        float m = ext / 255.0f;
        uint32_t b = __float_as_uint(m);
        uint32_t ex = b >> 23, mant = b & 0x7FFFFFu;
        e = (ex == 0u) ? 1u : ex + (mant != 0u ? 1u : 0u);
        if (e < 1u) e = 1u;
    }
    while (e < 254u && __builtin_fmaf(__uint_as_float(e << 23), 255.0f, lo) < hi) e++;
    return e;
}

__device__ __forceinline__ uint32_t quant_lo(float v, float lo, float s)
{
    int q = (int)floorf((v - lo) / s);
    q = q < 0 ? 0 : (q > 255 ? 255 : q);
    while (q > 0 && __builtin_fmaf(s, (float)q, lo) > v) q--;
    return (uint32_t)q;
}

__device__ __forceinline__ uint32_t quant_hi(float v, float lo, float s)
{
    int q = (int)ceilf((v - lo) / s);
    q = q < 0 ? 0 : (q > 255 ? 255 : q);
    while (q < 255 && __builtin_fmaf(s, (float)q, lo) < v) q++;
    return (uint32_t)q;
}


__device__ __forceinline__ void leaf_box(const float4 sp, float* b)
{
    const float r = fabsf(sp.w);
    b[0] = sp.x - r; b[1] = sp.y - r; b[2] = sp.z - r;
    b[3] = sp.x + r; b[4] = sp.y + r; b[5] = sp.z + r;
}

// Union of the two child boxes into nb, and the 32-byte packed node p.
__device__ __forceinline__ void pack_node(uint32_t p, const float (&cb)[2][6], uint32_t g, uint32_t flags,
                                          float (&nb)[6], uint4* __restrict__ nodes)
{
    for (int k = 0; k < 3; k++) {
        nb[k] = fminf(cb[0][k], cb[1][k]);
        nb[3 + k] = fmaxf(cb[0][3 + k], cb[1][3 + k]);
    }
    uint32_t meta = 0, q[3];
    float o[3];
    for (int k = 0; k < 3; k++) {
        const float lo = nb[k], hi = nb[3 + k];
        const uint32_t e = pick_exp(lo, hi);
        const float sc = __uint_as_float(e << 23);
        meta |= e << (8 * k);
        const uint32_t l0 = quant_lo(cb[0][k], lo, sc), h0 = quant_hi(cb[0][3 + k], lo, sc);
        const uint32_t l1 = quant_lo(cb[1][k], lo, sc), h1 = quant_hi(cb[1][3 + k], lo, sc);
        o[k] = lo;
        q[k] = l0 | (h0 << 8) | (l1 << 16) | (h1 << 24);
    }
    meta |= flags << 24;
    nodes[2 * p] = make_uint4(__float_as_uint(o[0]), __float_as_uint(o[1]), __float_as_uint(o[2]), meta);
    nodes[2 * p + 1] = make_uint4(q[0], q[1], q[2], g);
}

// Bottom-up refit (Karras 2012, section 4): one thread per leaf climbs, and
// the second thread to reach a node refits it.  Split in two kernels so that
// the bulk of the nodes needs no cross-XCD coherence:
//   k_refit_local  one workgroup per kRefitLeaves leaves.  Every node whose
//                  leaf range lies inside the workgroup's range also has its
//                  index there, so its child word, parent link, arrival
//                  counter, box and height all live in LDS.  A climb stops
//                  below the first node whose range crosses workgroups and
//                  publishes the finished child (a leaf or a local node) on
//                  the frontier list.
//   k_refit_top    climbs from every frontier entry through the crossing
//                  nodes (about n / kRefitLeaves * log2 kRefitLeaves of them)
//                  in global memory at agent scope.  Starting from a clean L2
//                  (the kernel boundary wrote the local results back), each
//                  release only writes back the few lines this kernel
//                  dirtied: 0.8 ms for the whole 1M-sphere build against
//                  1.2 ms with the top done by one workgroup at workgroup
//                  scope, and 10 ms with every node at agent scope in one
//                  kernel.
constexpr uint32_t kLeafItem = 0x80000000u;  // frontier entry is a leaf index
constexpr uint32_t kRefitThreads = 1024;
constexpr uint32_t kLeavesPerThread = kRefitLeaves / kRefitThreads;

__global__ void __launch_bounds__(kRefitThreads) k_refit_local(const float4* __restrict__ sph, uint32_t n,
                                                               const uint32_t* __restrict__ child,
                                                               const uint32_t* __restrict__ parent_int,
                                                               const uint32_t* __restrict__ parent_leaf,
                                                               float* __restrict__ boxes,
                                                               uint32_t* __restrict__ height,
                                                               uint4* __restrict__ nodes,
                                                               uint32_t* __restrict__ frontier,
                                                               uint32_t* __restrict__ fcount)
{
    __shared__ uint32_t s_visit[kRefitLeaves];
    __shared__ uint32_t s_child[kRefitLeaves];
    __shared__ uint32_t s_up[kRefitLeaves];
    __shared__ uint32_t s_h[kRefitLeaves];
    __shared__ float s_box[6][kRefitLeaves];
    const uint32_t tid = threadIdx.x, base = blockIdx.x * kRefitLeaves;
    for (uint32_t i = tid; i < kRefitLeaves; i += kRefitThreads) {
        s_visit[i] = 0u;
        const uint32_t g = base + i;
        if (g + 1u < n) {  // internal nodes are 0 .. n-2
            s_child[i] = child[g];
            s_up[i] = parent_int[g];
        }
    }
    __syncthreads();
    for (uint32_t r = 0; r < kLeavesPerThread; r++) {
        const uint32_t leaf = base + r * kRefitThreads + tid;
        if (leaf >= n) break;
        uint32_t up = parent_leaf[leaf];
        uint32_t done = leaf | kLeafItem;  // the highest node this climb has finished
        float nb[6];
        uint32_t nh = 0;
        while (up & kLocalBit) {
            const uint32_t p = up & ~kLocalBit, slot = p - base;
            if (__hip_atomic_fetch_add(&s_visit[slot], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u)
                break;
            const uint32_t cw = s_child[slot];
            const uint32_t g = cw & 0x3FFFFFFFu, flags = cw >> 30;
            float cb[2][6];
            uint32_t h = 0;
            for (int c = 0; c < 2; c++) {
                const uint32_t ch = g + (uint32_t)c;
                if ((flags >> c) & 1u) {
                    leaf_box(sph[ch], cb[c]);
                } else {
                    const uint32_t cs = ch - base;
                    for (int k = 0; k < 6; k++)
                        cb[c][k] = __hip_atomic_load(&s_box[k][cs], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    h = max(h, __hip_atomic_load(&s_h[cs], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
                }
            }
            pack_node(p, cb, g, flags, nb, nodes);
            nh = h + 1u;
            for (int k = 0; k < 6; k++) {
                __hip_atomic_store(&s_box[k][slot], nb[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                boxes[(size_t)p * 6 + k] = nb[k];  // every node's exact box: the 4-wide collapse reads them
            }
            __hip_atomic_store(&s_h[slot], nh, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            done = p;
            if (p == 0u) break;  // the whole tree was local
            up = s_up[slot];
        }
        if (up & kLocalBit) {
            // stopped inside the workgroup: first arrival, or the local root
            if (done != 0u) continue;
        }
        if (!(done & kLeafItem)) {  // a finished local node: publish its box and height
            for (int k = 0; k < 6; k++) boxes[done * 6 + k] = nb[k];
            height[done] = nh;
        }
        if (done != 0u) frontier[atomicAdd(fcount, 1u)] = done;
    }
}

constexpr uint32_t kTopBlocks = 64;

__global__ void __launch_bounds__(kRefitThreads) k_refit_top(const float4* __restrict__ sph,
                                                    const uint32_t* __restrict__ child,
                                                    const uint32_t* __restrict__ parent_int,
                                                    const uint32_t* __restrict__ parent_leaf,
                                                    const uint32_t* __restrict__ frontier,
                                                    const uint32_t* __restrict__ fcount, uint32_t* visit,
                                                    float* boxes, uint32_t* height, uint4* __restrict__ nodes)
{
    const uint32_t count = *fcount;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t item = frontier[i];
        uint32_t p = ((item & kLeafItem) ? parent_leaf[item & ~kLeafItem] : parent_int[item]) & ~kLocalBit;
        for (;;) {
            if (__hip_atomic_fetch_add(&visit[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == 0u) break;
            const uint32_t cw = child[p];
            const uint32_t g = cw & 0x3FFFFFFFu, flags = cw >> 30;
            float cb[2][6];
            uint32_t h = 0;
            for (int c = 0; c < 2; c++) {
                const uint32_t ch = g + (uint32_t)c;
                if ((flags >> c) & 1u) {
                    leaf_box(sph[ch], cb[c]);
                } else {
                    for (int k = 0; k < 6; k++)
                        cb[c][k] = __hip_atomic_load(&boxes[ch * 6 + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    h = max(h, __hip_atomic_load(&height[ch], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                }
            }
            float nb[6];
            pack_node(p, cb, g, flags, nb, nodes);
            for (int k = 0; k < 6; k++) __hip_atomic_store(&boxes[p * 6 + k], nb[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&height[p], h + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (p == 0u) break;
            p = parent_int[p] & ~kLocalBit;
        }
    }
}

// ---- 7. the 4-wide tree ---------------------------------------------------------
// Collapsed top down from the binary tree, level by level: a wide root's
// children start as its binary node's two children; while there are fewer
// than 4, the internal child of largest box area (the first on a tie) is
// opened into its two children in place, so slots keep left-to-right leaf
// order.  Every internal child becomes a wide root of the next level.  Wide
// indices number the roots in binary index order (an exclusive scan of the
// root flags), so the tree does not depend on the order threads ran in.
// Each wide node (64 bytes) is quantised relative to its own box like the
// binary nodes: lo corner, per-axis power-of-two scale, planes rounded
// outward; meta holds the scale exponents, the slots' leaf flags (bits
// 24-27) and valid flags (28-31).
constexpr uint32_t kWideRoot = 0x80000000u;  // wide-root flag of a binary node
constexpr uint32_t kLeafRef = 0x80000000u;   // a wide child that is a leaf
constexpr uint32_t kScanItems = 4;            // root flags per thread in the index scan
constexpr uint32_t kScanTile = kBlk * kScanItems;

__device__ __forceinline__ float box_area(const float* b)
{
    const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return dx * dy + dy * dz + dz * dx;
}

__device__ __forceinline__ uint32_t child_ref(uint32_t cw, int c)
{
    const uint32_t g = (cw & 0x3FFFFFFFu) + (uint32_t)c;
    return ((cw >> (30 + c)) & 1u) ? (g | kLeafRef) : g;
}

__global__ void __launch_bounds__(kBlk) k_wide_init(uint32_t* __restrict__ wflag, uint32_t* __restrict__ front,
                                                    uint32_t* __restrict__ cnt, uint32_t ncnt)
{
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < ncnt; i += kBlk) cnt[i] = i == 0u ? 1u : 0u;
    if (t == 0) {
        wflag[0] = kWideRoot;
        front[0] = 0u;
    }
}

// One level: every wide root of the frontier picks its children; internal
// children are the next level's frontier.  cnt[level + 1] starts at zero.
__global__ void __launch_bounds__(kBlk) k_wide_level(const uint32_t* __restrict__ child,
                                                     const float* __restrict__ boxes,
                                                     const uint32_t* __restrict__ front_in,
                                                     uint32_t* __restrict__ front_out, uint32_t* __restrict__ cnt,
                                                     uint32_t level, uint32_t* __restrict__ wkids,
                                                     uint32_t* __restrict__ wflag, uint32_t* __restrict__ wdepth)
{
    const uint32_t count = cnt[level];
    for (uint32_t i = blockIdx.x * kBlk + threadIdx.x; i < count; i += gridDim.x * kBlk) {
        const uint32_t r = front_in[i];
        const uint32_t cw = child[r];
        uint32_t ks[4] = {child_ref(cw, 0), child_ref(cw, 1), 0u, 0u};
        uint32_t n_k = 2;
        while (n_k < 4u) {
            int best = -1;
            float ba = 0.0f;
            for (uint32_t k = 0; k < n_k; k++) {
                if (ks[k] & kLeafRef) continue;
                const float a = box_area(boxes + (size_t)ks[k] * 6);
                if (best < 0 || a > ba) {
                    best = (int)k;
                    ba = a;
                }
            }
            if (best < 0) break;
            const uint32_t ow = child[ks[best]];
            for (int k = (int)n_k - 1; k > best; k--) ks[k + 1] = ks[k];
            ks[best] = child_ref(ow, 0);
            ks[best + 1] = child_ref(ow, 1);
            n_k++;
        }
        reinterpret_cast<uint4*>(wkids)[r] = make_uint4(ks[0], ks[1], ks[2], ks[3]);
        wflag[r] = kWideRoot | n_k;
        for (uint32_t k = 0; k < n_k; k++) {
            const uint32_t c = ks[k];
            if (c & kLeafRef) continue;
            wflag[c] = kWideRoot;
            front_out[atomicAdd(&cnt[level + 1], 1u)] = c;
        }
        atomicMax(wdepth, level + 2u);  // wide levels 1..level+1, plus the leaf level
    }
}

// Wide indices: an exclusive scan of the root flags over the binary nodes,
// in tiles of kScanTile (tile totals, their scan, then each tile's own).
__global__ void __launch_bounds__(kBlk) k_wscan_tiles(const uint32_t* __restrict__ wflag, uint32_t ni,
                                                      uint32_t* __restrict__ tsum)
{
    __shared__ uint32_t s[kBlk];
    uint32_t c = 0;
    const uint32_t base = blockIdx.x * kScanTile;
    for (uint32_t r = 0; r < kScanItems; r++) {
        const uint32_t i = base + r * kBlk + threadIdx.x;
        if (i < ni && (wflag[i] & kWideRoot)) c++;
    }
    const uint32_t e = block_scan_excl(c, s);
    if (threadIdx.x == kBlk - 1) tsum[blockIdx.x] = e + c;
}

__global__ void __launch_bounds__(kBlk) k_wscan_top(uint32_t* __restrict__ tsum, uint32_t nt,
                                                    uint32_t* __restrict__ total)
{
    __shared__ uint32_t part[kBlk];
    const uint32_t chunk = (nt + kBlk - 1) / kBlk;
    const uint32_t b = threadIdx.x * chunk, e = min(b + chunk, nt);
    uint32_t sum = 0;
    for (uint32_t i = b; i < e; i++) sum += tsum[i];
    const uint32_t excl = block_scan_excl(sum, part);
    uint32_t run = excl;
    for (uint32_t i = b; i < e; i++) {
        const uint32_t v = tsum[i];
        tsum[i] = run;
        run += v;
    }
    if (threadIdx.x == kBlk - 1) *total = excl + sum;
}

__global__ void __launch_bounds__(kBlk) k_wscan_apply(const uint32_t* __restrict__ wflag, uint32_t ni,
                                                      const uint32_t* __restrict__ tsum,
                                                      uint32_t* __restrict__ widx)
{
    __shared__ uint32_t s[kBlk];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;  // items contiguous per thread
    uint32_t f[kScanItems], c = 0;
    for (uint32_t r = 0; r < kScanItems; r++) {
        const uint32_t i = base + r;
        f[r] = (i < ni && (wflag[i] & kWideRoot)) ? 1u : 0u;
        c += f[r];
    }
    uint32_t run = tsum[blockIdx.x] + block_scan_excl(c, s);
    for (uint32_t r = 0; r < kScanItems; r++) {
        const uint32_t i = base + r;
        if (f[r]) widx[i] = run;
        run += f[r];
    }
}

__global__ void __launch_bounds__(kBlk) k_wide_pack(const float4* __restrict__ sph, const uint32_t* __restrict__ wkids,
                                                    const uint32_t* __restrict__ wflag,
                                                    const uint32_t* __restrict__ widx,
                                                    const float* __restrict__ boxes, uint32_t ni,
                                                    uint4* __restrict__ wnodes)
{
    const uint32_t r = blockIdx.x * kBlk + threadIdx.x;
    if (r >= ni) return;
    const uint32_t fl = wflag[r];
    if (!(fl & kWideRoot)) return;
    const uint32_t cnt = fl & 7u;
    const float* nb = boxes + (size_t)r * 6;
    uint32_t meta = 0, qlo[3] = {0u, 0u, 0u}, qhi[3] = {0u, 0u, 0u}, ch[4] = {0u, 0u, 0u, 0u};
    float sc[3];
    for (int k = 0; k < 3; k++) {
        const uint32_t e = pick_exp(nb[k], nb[3 + k]);
        sc[k] = __uint_as_float(e << 23);
        meta |= e << (8 * k);
    }
    for (uint32_t c = 0; c < cnt; c++) {
        const uint32_t ref = wkids[(size_t)r * 4 + c];
        float cb[6];
        if (ref & kLeafRef) {
            const uint32_t leaf = ref & ~kLeafRef;
            leaf_box(sph[leaf], cb);
            ch[c] = leaf;
            meta |= 1u << (24 + c);
        } else {
            for (int k = 0; k < 6; k++) cb[k] = boxes[(size_t)ref * 6 + k];
            ch[c] = widx[ref];
        }
        meta |= 1u << (28 + c);
        for (int k = 0; k < 3; k++) {
            qlo[k] |= quant_lo(cb[k], nb[k], sc[k]) << (8 * c);
            qhi[k] |= quant_hi(cb[3 + k], nb[k], sc[k]) << (8 * c);
        }
    }
    uint4* w = wnodes + (size_t)widx[r] * 4;
    w[0] = make_uint4(__float_as_uint(nb[0]), __float_as_uint(nb[1]), __float_as_uint(nb[2]), meta);
    w[1] = make_uint4(qlo[0], qlo[1], qlo[2], qhi[0]);
    w[2] = make_uint4(qhi[1], qhi[2], ch[0], ch[1]);
    w[3] = make_uint4(ch[2], ch[3], 0u, 0u);
}

__global__ void k_zero(uint32_t* a, uint32_t n)
{
    uint32_t i = blockIdx.x * kBlk + threadIdx.x;
    if (i < n) a[i] = 0u;
}

__global__ void k_fill(uint32_t* a, uint32_t n, uint32_t v)
{
    uint32_t i = blockIdx.x * kBlk + threadIdx.x;
    if (i < n) a[i] = v;
}

// Validation of device-resident input: first[0] = lowest index of a sphere
// with a non-finite centre or a zero/non-finite radius, first[1] = lowest
// index of a material with an unknown type (0xFFFFFFFF: none).
__global__ void __launch_bounds__(kBlk) k_check_input(const float4* __restrict__ sph,
                                                      const vzMaterial* __restrict__ mat, uint32_t n,
                                                      uint32_t* __restrict__ first)
{
    for (uint32_t i = blockIdx.x * kBlk + threadIdx.x; i < n; i += gridDim.x * kBlk) {
        const float4 c = sph[i];
        if (!(isfinite(c.x) && isfinite(c.y) && isfinite(c.z) && isfinite(c.w) && c.w != 0.0f))
            atomicMin(&first[0], i);
        if (mat && mat[i].type > VZ_DIELECTRIC) atomicMin(&first[1], i);
    }
}

inline uint32_t blocks_for(uint32_t n, uint32_t per) { return (n + per - 1) / per; }

}  // namespace

vzStatus lbvh_alloc(vzScene s)
{
    uint32_t n = s->n, ni = n > 1 ? n - 1 : 1;
    s->hist_blocks = blocks_for(n, kTile);
    uint64_t bytes = 0;
    auto al = [&](void** p, size_t b) -> vzStatus {
        hipError_t e = hipMalloc(p, b);
        if (e != hipSuccess) return set_error(VZ_ERROR_OUT_OF_MEMORY, "hipMalloc of scene buffers failed");
        bytes += b;
        return VZ_SUCCESS;
    };
    VZ_TRY(al((void**)&s->d_sph_in, sizeof(float4) * n));
    VZ_TRY(al((void**)&s->d_mat_in, sizeof(vzMaterial) * n));
    for (int i = 0; i < 2; i++) {
        VZ_TRY(al((void**)&s->d_keys[i], sizeof(uint32_t) * n));
        VZ_TRY(al((void**)&s->d_vals[i], sizeof(uint32_t) * n));
    }
    VZ_TRY(al((void**)&s->d_hist, sizeof(uint32_t) * 256 * (s->hist_blocks + 1)));
    VZ_TRY(al((void**)&s->d_bounds, sizeof(float) * (6 * 1024 + 32)));  // partials, grid A, grid B
    VZ_TRY(al((void**)&s->d_sph, sizeof(float4) * n));
    VZ_TRY(al((void**)&s->d_leaf_info, sizeof(uint32_t) * n));
    VZ_TRY(al((void**)&s->d_mat, sizeof(float4) * n));
    VZ_TRY(al((void**)&s->d_nodes, sizeof(uint4) * 2 * ni));
    VZ_TRY(al((void**)&s->d_boxes, sizeof(float) * 6 * ni));
    VZ_TRY(al((void**)&s->d_child, sizeof(uint32_t) * ni));
    VZ_TRY(al((void**)&s->d_parent_int, sizeof(uint32_t) * ni));
    VZ_TRY(al((void**)&s->d_parent_leaf, sizeof(uint32_t) * n));
    VZ_TRY(al((void**)&s->d_visit, sizeof(uint32_t) * ni));
    VZ_TRY(al((void**)&s->d_height, sizeof(uint32_t) * ni));
    VZ_TRY(al((void**)&s->d_frontier, sizeof(uint32_t) * n));
    VZ_TRY(al((void**)&s->d_fcount, 256));
    VZ_TRY(al((void**)&s->d_wnodes, sizeof(uint4) * 4 * ni));
    VZ_TRY(al((void**)&s->d_wkids, sizeof(uint32_t) * 4 * ni));
    VZ_TRY(al((void**)&s->d_wflag, sizeof(uint32_t) * ni));
    VZ_TRY(al((void**)&s->d_widx, sizeof(uint32_t) * (ni + blocks_for(ni, kScanTile) + 1)));  // + tile sums
    VZ_TRY(al((void**)&s->d_wfront[0], sizeof(uint32_t) * ni));
    VZ_TRY(al((void**)&s->d_wfront[1], sizeof(uint32_t) * ni));
    VZ_TRY(al((void**)&s->d_wcnt, sizeof(uint32_t) * (kMaxDepth + 8)));
    s->device_bytes = bytes;
    return VZ_SUCCESS;
}

void lbvh_free(vzScene s)
{
    void* ptrs[] = {s->d_sph_in, s->d_mat_in, s->d_keys[0], s->d_keys[1], s->d_vals[0], s->d_vals[1],
                    s->d_hist, s->d_bounds, s->d_sph, s->d_leaf_info, s->d_mat, s->d_nodes,
                    s->d_boxes, s->d_child, s->d_parent_int, s->d_parent_leaf, s->d_visit,
                    s->d_height, s->d_frontier, s->d_fcount, s->d_wnodes, s->d_wkids, s->d_wflag,
                    s->d_widx, s->d_wfront[0], s->d_wfront[1], s->d_wcnt};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
}

// Structural check of the packed 4-wide tree, so that traversal needs no
// per-visit guards: every child index in range, every leaf referenced by
// exactly one slot, every wide node but the root by exactly one slot and the
// root by none.  Then a walk from the root reaches each node at most once
// and ends.  Any violation sets *err.
__global__ void __launch_bounds__(kBlk) k_wide_refs(const uint4* __restrict__ wnodes, const uint32_t* __restrict__ nwp,
                                                    uint32_t n, uint32_t* __restrict__ wref,
                                                    uint32_t* __restrict__ lref, uint32_t* __restrict__ err)
{
    const uint32_t nw = *nwp;
    for (uint32_t i = blockIdx.x * kBlk + threadIdx.x; i < nw; i += gridDim.x * kBlk) {
        const uint4 a = wnodes[4 * (size_t)i], c = wnodes[4 * (size_t)i + 2], d = wnodes[4 * (size_t)i + 3];
        const uint32_t valid = a.w >> 28, leafm = (a.w >> 24) & 0xFu;
        const uint32_t ch[4] = {c.z, c.w, d.x, d.y};
        for (int k = 0; k < 4; k++) {
            if (!((valid >> k) & 1u)) continue;
            if ((leafm >> k) & 1u) {
                if (ch[k] < n) atomicAdd(&lref[ch[k]], 1u);
                else *err = 1u;
            } else {
                if (ch[k] != 0u && ch[k] < nw) atomicAdd(&wref[ch[k]], 1u);
                else *err = 1u;
            }
        }
    }
}

__global__ void __launch_bounds__(kBlk) k_wide_verify(const uint32_t* __restrict__ nwp, uint32_t n,
                                                      const uint32_t* __restrict__ wref,
                                                      const uint32_t* __restrict__ lref, uint32_t* __restrict__ err)
{
    const uint32_t nw = *nwp;
    for (uint32_t i = blockIdx.x * kBlk + threadIdx.x; i < n; i += gridDim.x * kBlk) {
        if (lref[i] != 1u) *err = 1u;
        if (i < nw && wref[i] != (i == 0u ? 0u : 1u)) *err = 1u;
    }
}

// The 4-wide tree of the current binary tree (after every build and refit).
// One launch per wide level: the binary depth bounds the wide depth, and
// levels past the deepest find an empty frontier.
static vzStatus collapse_wide(vzScene s, hipStream_t st)
{
    const uint32_t n = s->n, ni = n - 1;
    uint32_t* cnt = s->d_wcnt;  // [0, kMaxDepth]: per level; then nw, the wide depth
    uint32_t* tsum = s->d_widx + ni;
    const uint32_t ntile = blocks_for(ni, kScanTile);
    hipLaunchKernelGGL(k_zero, dim3(blocks_for(ni, kBlk)), dim3(kBlk), 0, st, s->d_wflag, ni);
    hipLaunchKernelGGL(k_wide_init, dim3(1), dim3(kBlk), 0, st, s->d_wflag, s->d_wfront[0], cnt, kMaxDepth + 8);
    const uint32_t grid = std::min<uint32_t>(1024u, blocks_for(ni, kBlk));
    for (uint32_t L = 0; L + 1 < s->depth && L < kMaxDepth; L++)
        hipLaunchKernelGGL(k_wide_level, dim3(grid), dim3(kBlk), 0, st, s->d_child, s->d_boxes,
                           s->d_wfront[L & 1u], s->d_wfront[(L + 1) & 1u], cnt, L, s->d_wkids, s->d_wflag,
                           cnt + kMaxDepth + 2);
    hipLaunchKernelGGL(k_wscan_tiles, dim3(ntile), dim3(kBlk), 0, st, s->d_wflag, ni, tsum);
    hipLaunchKernelGGL(k_wscan_top, dim3(1), dim3(kBlk), 0, st, tsum, ntile, cnt + kMaxDepth + 1);
    hipLaunchKernelGGL(k_wscan_apply, dim3(ntile), dim3(kBlk), 0, st, s->d_wflag, ni, tsum, s->d_widx);
    hipLaunchKernelGGL(k_wide_pack, dim3(blocks_for(ni, kBlk)), dim3(kBlk), 0, st, s->d_sph, s->d_wkids,
                       s->d_wflag, s->d_widx, s->d_boxes, ni, s->d_wnodes);
    // reference counts in buffers free at this point: the level fronts and
    // the refit frontier (both rebuilt by every refit)
    uint32_t* wref = s->d_wfront[0];
    uint32_t* lref = s->d_frontier;
    hipLaunchKernelGGL(k_zero, dim3(blocks_for(ni, kBlk)), dim3(kBlk), 0, st, wref, ni);
    hipLaunchKernelGGL(k_zero, dim3(blocks_for(n, kBlk)), dim3(kBlk), 0, st, lref, n);
    hipLaunchKernelGGL(k_wide_refs, dim3(grid), dim3(kBlk), 0, st, s->d_wnodes, cnt + kMaxDepth + 1, n, wref, lref,
                       cnt + kMaxDepth + 3);
    hipLaunchKernelGGL(k_wide_verify, dim3(grid), dim3(kBlk), 0, st, cnt + kMaxDepth + 1, n, wref, lref,
                       cnt + kMaxDepth + 3);
    VZ_HIP_TRY(hipGetLastError());
    uint32_t h[3] = {0u, 0u, 0u};
    VZ_HIP_TRY(hipMemcpyAsync(h, cnt + kMaxDepth + 1, sizeof(h), hipMemcpyDeviceToHost, st));
    VZ_HIP_TRY(hipStreamSynchronize(st));
    s->nw = h[0];
    s->wdepth = h[1];
    if (h[2] != 0u) return set_error(VZ_ERROR_INTERNAL, "4-wide BVH failed its structural check");
    if (s->wdepth > kMaxWideDepth)
        return set_error(VZ_ERROR_CAPACITY, "4-wide BVH depth " + std::to_string(s->wdepth) +
                                                " exceeds the traversal stack (" + std::to_string(kMaxWideDepth) + ")");
    return VZ_SUCCESS;
}

static vzStatus refit_core(vzScene s, hipStream_t st)
{
    uint32_t n = s->n;
    if (n < 2) return VZ_SUCCESS;
    hipLaunchKernelGGL(k_zero, dim3(blocks_for(n - 1, kBlk)), dim3(kBlk), 0, st, s->d_visit, n - 1);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(kBlk), 0, st, s->d_fcount, 1u);
    hipLaunchKernelGGL(k_refit_local, dim3(blocks_for(n, kRefitLeaves)), dim3(kRefitThreads), 0, st, s->d_sph,
                       n, s->d_child, s->d_parent_int, s->d_parent_leaf, s->d_boxes, s->d_height, s->d_nodes,
                       s->d_frontier, s->d_fcount);
    hipLaunchKernelGGL(k_refit_top, dim3(kTopBlocks), dim3(kRefitThreads), 0, st, s->d_sph, s->d_child,
                       s->d_parent_int, s->d_parent_leaf, s->d_frontier, s->d_fcount, s->d_visit, s->d_boxes,
                       s->d_height, s->d_nodes);
    VZ_HIP_TRY(hipGetLastError());
    uint32_t h = 0;
    VZ_HIP_TRY(hipMemcpyAsync(&h, s->d_height, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    VZ_HIP_TRY(hipStreamSynchronize(st));
    s->depth = h + 1;
    if (s->depth > kMaxDepth)
        return set_error(VZ_ERROR_CAPACITY, "BVH depth " + std::to_string(s->depth) +
                                                " exceeds the traversal stack (" +
                                                std::to_string(kMaxDepth) + ")");
    return collapse_wide(s, st);
}

vzStatus lbvh_build(vzScene s, hipStream_t st)
{
    uint32_t n = s->n;
    uint32_t nb = std::min<uint32_t>(1024, blocks_for(n, kBlk));
    Event e0, e1;
    VZ_EVENT_TRY(e0);
    VZ_EVENT_TRY(e1);
    VZ_HIP_TRY(hipEventRecord(e0.e, st));
    float* partial = s->d_bounds;
    float* fin = s->d_bounds + 6 * 1024;
    float* grid_b = fin + 16;
    hipLaunchKernelGGL(k_bounds_partial, dim3(nb), dim3(kBlk), 0, st, s->d_sph_in, n, nullptr, partial);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(kBlk), 0, st, partial, nb, 0u, fin);
    hipLaunchKernelGGL(k_bounds_partial, dim3(nb), dim3(kBlk), 0, st, s->d_sph_in, n, fin, partial);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(kBlk), 0, st, partial, nb, 1u, grid_b);
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(n, kBlk)), dim3(kBlk), 0, st, s->d_sph_in, n, fin, grid_b,
                       s->d_keys[0], s->d_vals[0]);
    s->builder = (n >= 2 && n <= kSahMaxPrims) ? VZ_BUILD_SAH : VZ_BUILD_LBVH;
    // keys and leaf order end in buffers 0 either way
    s->d_morton = s->d_keys[0];
    s->d_leaf_prim = s->d_vals[0];
    if (s->builder == VZ_BUILD_SAH) {
        hipLaunchKernelGGL(k_sah_build, dim3(1), dim3(kSahThreads), 0, st, s->d_sph_in, n, s->d_leaf_prim,
                           s->d_morton, s->d_child, s->d_parent_int, s->d_parent_leaf);
    } else {
        uint32_t tiles = s->hist_blocks;
        uint32_t* dtot = s->d_hist + 256 * tiles;
        for (int pass = 0; pass < 4; pass++) {  // four passes: the result is back in buffers 0
            int src = pass & 1, dst = src ^ 1;
            hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kBlk), 0, st, s->d_keys[src], n, pass * 8,
                               tiles, s->d_hist);
            hipLaunchKernelGGL(k_scan_rows, dim3(256), dim3(kBlk), 0, st, s->d_hist, tiles, dtot);
            hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kBlk), 0, st, s->d_keys[src],
                               s->d_vals[src], n, pass * 8, tiles, s->d_hist, dtot, s->d_keys[dst], s->d_vals[dst]);
        }
        if (n >= 2) {
            // node leaf ranges go to the wide-children buffer, which the collapse rewrites later
            uint2* range = reinterpret_cast<uint2*>(s->d_wkids);
            hipLaunchKernelGGL(k_karras, dim3(blocks_for(n - 1, kBlk)), dim3(kBlk), 0, st, s->d_morton, n,
                               s->d_child, s->d_parent_int, s->d_parent_leaf, s->d_visit, range);
            if (n >= kTreeletMinPrims) {
                // the treelet list uses the node-box buffer, which the refit fills afterwards
                uint4* list = reinterpret_cast<uint4*>(s->d_boxes);
                uint32_t* count = s->d_fcount + 8;
                hipLaunchKernelGGL(k_zero, dim3(1), dim3(kBlk), 0, st, count, 1u);
                hipLaunchKernelGGL(k_treelet_roots, dim3(blocks_for(n - 1, kBlk)), dim3(kBlk), 0, st, range,
                                   s->d_parent_int, n - 1, list, count);
                hipLaunchKernelGGL(k_sah_treelets, dim3(kTreeletBlocks), dim3(kSahThreads), 0, st, s->d_sph_in,
                                   list, count, s->d_leaf_prim, s->d_morton, s->d_child, s->d_parent_int,
                                   s->d_parent_leaf);
            }
        }
    }
    hipLaunchKernelGGL(k_reorder, dim3(blocks_for(n, kBlk)), dim3(kBlk), 0, st, s->d_sph_in,
                       s->d_mat_in, s->d_leaf_prim, n, s->d_sph, s->d_leaf_info, s->d_mat);
    VZ_HIP_TRY(hipGetLastError());
    vzStatus rs = refit_core(s, st);
    VZ_HIP_TRY(hipEventRecord(e1.e, st));
    VZ_HIP_TRY(hipEventSynchronize(e1.e));
    float ms = 0.0f;
    VZ_HIP_TRY(hipEventElapsedTime(&ms, e0.e, e1.e));
    s->build_ms = ms;
    if (n < 2) {
        s->depth = 1;
        s->nw = 0;
        s->wdepth = 1;
    }
    return rs;
}

vzStatus lbvh_check_input(vzScene s, const vzMaterial* mat, hipStream_t st, uint32_t* bad_sphere,
                          uint32_t* bad_material)
{
    uint32_t* first = s->d_fcount + 2;  // words 2-3 of the refit counter block
    uint32_t h[2] = {0u, 0u};
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(kBlk), 0, st, first, 2u, 0xFFFFFFFFu);
    const uint32_t nb = std::min<uint32_t>(1024u, blocks_for(s->n, kBlk));
    hipLaunchKernelGGL(k_check_input, dim3(nb), dim3(kBlk), 0, st, s->d_sph_in, mat, s->n, first);
    VZ_HIP_TRY(hipGetLastError());
    VZ_HIP_TRY(hipMemcpyAsync(h, first, sizeof(h), hipMemcpyDeviceToHost, st));
    VZ_HIP_TRY(hipStreamSynchronize(st));
    *bad_sphere = h[0];
    *bad_material = h[1];
    return VZ_SUCCESS;
}

vzStatus lbvh_refit(vzScene s, hipStream_t st)
{
    uint32_t n = s->n;
    Event e0, e1;
    VZ_EVENT_TRY(e0);
    VZ_EVENT_TRY(e1);
    VZ_HIP_TRY(hipEventRecord(e0.e, st));
    hipLaunchKernelGGL(k_reorder_sph, dim3(blocks_for(n, kBlk)), dim3(kBlk), 0, st, s->d_sph_in,
                       s->d_leaf_prim, n, s->d_sph);
    VZ_HIP_TRY(hipGetLastError());
    vzStatus rs = refit_core(s, st);
    VZ_HIP_TRY(hipEventRecord(e1.e, st));
    VZ_HIP_TRY(hipEventSynchronize(e1.e));
    float ms = 0.0f;
    VZ_HIP_TRY(hipEventElapsedTime(&ms, e0.e, e1.e));
    s->build_ms = ms;
    return rs;
}

}  // namespace vz
