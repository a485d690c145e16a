// vz_trace.hip -- closest-hit queries over a ray array (vzTraceClosest): the
// equivalent of one optixTrace per ray (Core/RayTracer.cu:190-202) without
// shading.  Used for the bit-exact ray-set check and as a traversal
// microbenchmark.
#include "vz_traverse.h"

namespace vz {
namespace {

constexpr int kTraceBlk = 256;

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

template <int S, bool COUNT>
__global__ void __launch_bounds__(kTraceBlk) k_trace(SceneView sc, const vzRay* __restrict__ rays,
                                                     uint32_t n, vzHit* __restrict__ hits,
                                                     unsigned long long* __restrict__ counters)
{
    __shared__ uint32_t s_node[S * kTraceBlk];
    __shared__ float s_t[S * kTraceBlk];
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
    if (COUNT) {
        nv = wave_sum(nv);
        ns = wave_sum(ns);
    }
    if (__lane_id() == 0 && seg) {
        atomicAdd(&counters[0], (unsigned long long)seg);
        if (COUNT) {
            atomicAdd(&counters[2], (unsigned long long)nv);
            atomicAdd(&counters[3], (unsigned long long)ns);
        }
    }
}

}  // namespace

vzStatus trace_rays(const SceneView& sc, const vzRay* rays, uint32_t n, vzHit* hits, hipStream_t st,
                    bool count, unsigned long long* d_counters)
{
    if (n == 0) return VZ_SUCCESS;
    dim3 grid((n + kTraceBlk - 1) / kTraceBlk);
    if (count)
        hipLaunchKernelGGL((k_trace<8, true>), grid, dim3(kTraceBlk), 0, st, sc, rays, n, hits, d_counters);
    else
        hipLaunchKernelGGL((k_trace<8, false>), grid, dim3(kTraceBlk), 0, st, sc, rays, n, hits, d_counters);
    VZ_HIP_TRY(hipGetLastError());
    return VZ_SUCCESS;
}

}  // namespace vz
