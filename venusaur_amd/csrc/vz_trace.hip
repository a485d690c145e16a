// vz_trace.hip -- closest-hit queries over a ray array (vzTraceClosest): the
// equivalent of one optixTrace per ray (Core/RayTracer.cu:190-202) without
// shading.  Used for the bit-exact ray-set check and as a traversal
// microbenchmark.
//
// Counters.  Each workgroup writes its totals (segments, node visits, sphere
// tests) to its own slot of a partials array; a one-workgroup kernel then
// sums the slots in slot order.  No two workgroups touch one address, so
// there is no device-scope atomic (on MI355X those serialise across the 8
// XCDs) and nothing needs clearing before a call: every slot the reduction
// reads is written by the same call.  vzTraceClosest allocates the array per
// call; vzTraceClosestNoAlloc uses `scratch` below, a fixed array in the
// library's device memory.
#include "vz_traverse.h"

namespace vz {

// Partials of up to kTraceScratchBlocks workgroups (the no-allocation entry
// point's limit on rays per call: 65536 x 256 = 16,777,216) plus the summed
// counters.  Each GPU the library runs on holds its own copy until the
// process ends; calls on one GPU share it.
static __device__ __align__(256) unsigned long long scratch[kTraceScratchBlocks * kTraceParts + 8];

namespace {

constexpr int kTraceBlk = kTraceBlockSize;

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

template <int S, bool COUNT>
__global__ void __launch_bounds__(kTraceBlk) k_trace(SceneView sc, const vzRay* __restrict__ rays,
                                                     uint32_t n, vzHit* __restrict__ hits,
                                                     unsigned long long* __restrict__ parts)
{
    __shared__ uint32_t s_node[S * kTraceBlk];
    __shared__ float s_t[S * kTraceBlk];
    __shared__ unsigned long long s_red[kTraceParts][kTraceBlk / 64];
    uint32_t i = blockIdx.x * kTraceBlk + threadIdx.x;
    uint32_t nv = 0, ns = 0, seg = 0;
    if (i < n) {
        const float4* rp = reinterpret_cast<const float4*>(rays + i);
        float4 a = rp[0], b = rp[1];
        TraceResult r = closest_hit<S, COUNT>(sc, GlobalNodes{sc.nodes}, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w, b.w,
                                              s_node + threadIdx.x, s_t + threadIdx.x, kTraceBlk, nv, ns);
        uint32_t prim = r.leaf == 0xFFFFFFFFu ? VZ_MISS : (sc.leaf_info[r.leaf] & 0x3FFFFFFFu);
        reinterpret_cast<uint2*>(hits)[i] = make_uint2(__float_as_uint(r.t), prim);
        seg = 1;
    }
    seg = wave_sum(seg);
    nv = COUNT ? wave_sum(nv) : 0u;
    ns = COUNT ? wave_sum(ns) : 0u;
    const uint32_t wave = threadIdx.x >> 6;
    if (__lane_id() == 0) {
        s_red[0][wave] = seg;
        s_red[1][wave] = nv;
        s_red[2][wave] = ns;
    }
    __syncthreads();
    if (threadIdx.x < kTraceParts) {
        unsigned long long t = 0;
        for (int w = 0; w < kTraceBlk / 64; w++) t += s_red[threadIdx.x][w];
        parts[(size_t)blockIdx.x * kTraceParts + threadIdx.x] = t;
    }
}

// Sums the nblk partial slots into out[0] (segments), out[2] (node visits)
// and out[3] (sphere tests), the layout of vzTraceClosest's counters.
__global__ void __launch_bounds__(kTraceBlk) k_trace_reduce(const unsigned long long* __restrict__ parts,
                                                            uint32_t nblk, unsigned long long* __restrict__ out)
{
    __shared__ unsigned long long s[kTraceParts][kTraceBlk];
    unsigned long long acc[kTraceParts] = {0, 0, 0};
    for (uint32_t b = threadIdx.x; b < nblk; b += kTraceBlk)
        for (int k = 0; k < kTraceParts; k++) acc[k] += parts[(size_t)b * kTraceParts + k];
    for (int k = 0; k < kTraceParts; k++) s[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (uint32_t half = kTraceBlk / 2; half > 0; half >>= 1) {
        if (threadIdx.x < half)
            for (int k = 0; k < kTraceParts; k++) s[k][threadIdx.x] += s[k][threadIdx.x + half];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s[0][0];
        out[1] = 0ull;
        out[2] = s[1][0];
        out[3] = s[2][0];
    }
}

}  // namespace

uint32_t trace_blocks(uint32_t n) { return (n + kTraceBlk - 1) / kTraceBlk; }

vzStatus trace_rays(const SceneView& sc, const vzRay* rays, uint32_t n, vzHit* hits, hipStream_t st,
                    bool count, unsigned long long* parts, unsigned long long* counters)
{
    const uint32_t nblk = trace_blocks(n);
    if (nblk) {
        if (count)
            hipLaunchKernelGGL((k_trace<8, true>), dim3(nblk), dim3(kTraceBlk), 0, st, sc, rays, n, hits, parts);
        else
            hipLaunchKernelGGL((k_trace<8, false>), dim3(nblk), dim3(kTraceBlk), 0, st, sc, rays, n, hits, parts);
    }
    hipLaunchKernelGGL(k_trace_reduce, dim3(1), dim3(kTraceBlk), 0, st, parts, nblk, counters);
    VZ_HIP_TRY(hipGetLastError());
    return VZ_SUCCESS;
}

// The device address of this GPU's `scratch` (the current device's copy).
// The array's name in host code is only a stand-in for the symbol; kernels
// must be given this address instead.
vzStatus trace_scratch_address(unsigned long long** out)
{
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(VZ_ERROR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    *out = static_cast<unsigned long long*>(p);
    return VZ_SUCCESS;
}

}  // namespace vz
