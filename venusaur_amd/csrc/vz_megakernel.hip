// vz_megakernel.hip -- one thread per pixel runs every sample's whole path:
// __raygen__rg (Core/RayTracer.cu:163-217) with the recursive optixTrace /
// closest-hit chain (RayTracer.cu:190, 299, 347, 422) unrolled into a loop.
// Kept beside the wavefront integrator as a second, structurally different
// GPU path: both must agree with the CPU integrator bit for bit.
#include "vz_traverse.h"

namespace vz {
namespace {

constexpr int kMegaBlk = 256;  // 16x16 pixel tile

struct MegaParams {
    Cam cam;
    uint32_t spp, subframe, max_depth, flags;
    void* accum;
    uint32_t* image;
};

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// One path from the camera ray (o, d) with its own seed; returns radiance.
template <int S, bool COUNT>
__device__ __forceinline__ v3 trace_path(const SceneView& sc, v3 o, v3 d, uint32_t seed,
                                         uint32_t max_depth, uint32_t* ln, float* lt, uint32_t& seg,
                                         uint32_t& nv, uint32_t& ns)
{
    v3 thr = mk(1.0f, 1.0f, 1.0f);
    int depth = (int)max_depth - 1;
    for (;;) {
        seg++;
        TraceResult r = closest_hit<S, COUNT>(sc, GlobalNodes{sc.nodes}, o, d, 0.001f, 1e16f, ln, lt, kMegaBlk, nv, ns);
        if (r.leaf == 0xFFFFFFFFu) return mul(thr, sky(d));
        if (depth <= 0) return mk(0.0f, 0.0f, 0.0f);
        float4 s = sc.spheres[r.leaf];
        uint32_t type = sc.leaf_info[r.leaf] >> 30;
        float4 mat = sc.mats[r.leaf];
        v3 p = add(o, scl(d, r.t));
        v3 n = scl(sub(p, mk(s.x, s.y, s.z)), recip_rn(s.w));
        bool front = dot(d, n) < 0.0f;
        if (!front) n = neg(n);
        v3 nd, att;
        if (!scatter(type, mat, d, n, front, seed, nd, att)) return mk(0.0f, 0.0f, 0.0f);
        thr = mul(thr, att);
        depth -= 1;
        o = p;
        d = nd;
    }
}

template <int S, bool COUNT, bool FP64>
__global__ void __launch_bounds__(kMegaBlk) k_mega(SceneView sc, MegaParams P,
                                                   unsigned long long* __restrict__ counters)
{
    __shared__ uint32_t s_node[S * kMegaBlk];
    __shared__ float s_t[S * kMegaBlk];
    const uint32_t W = P.cam.width, H = P.cam.height;
    const uint32_t tiles_x = (W + 15) / 16;
    const uint32_t x = (blockIdx.x % tiles_x) * 16 + (threadIdx.x & 15);
    const uint32_t y = (blockIdx.x / tiles_x) * 16 + (threadIdx.x >> 4);
    uint32_t seg = 0, nv = 0, ns = 0, paths = 0;
    if (x < W && y < H) {
        const uint32_t idx = y * W + x;
        uint32_t seed = tea4(idx, P.subframe);
        v3 sum = mk(0.0f, 0.0f, 0.0f);
        double s64[3] = {0.0, 0.0, 0.0};
        for (uint32_t s = 0; s < P.spp; s++) {
            v3 o, d;
            camera_ray(P.cam, x, y, seed, o, d);
            v3 L = trace_path<S, COUNT>(sc, o, d, seed, P.max_depth, s_node + threadIdx.x,
                                        s_t + threadIdx.x, seg, nv, ns);
            if (FP64) {
                s64[0] += (double)L.x;
                s64[1] += (double)L.y;
                s64[2] += (double)L.z;
            } else {
                sum = add(sum, L);
            }
            paths++;
        }
        float oc[3];
        if (FP64) {
            double* acc = (double*)P.accum + 4 * (size_t)idx;
            double avg[3];
            for (int k = 0; k < 3; k++) avg[k] = s64[k] / (double)P.spp;
            if (P.subframe > 1u) {
                double a = 1.0 / (double)P.subframe;
                for (int k = 0; k < 3; k++) avg[k] = acc[k] + a * (avg[k] - acc[k]);
            }
            reinterpret_cast<double2*>(acc)[0] = make_double2(avg[0], avg[1]);
            reinterpret_cast<double2*>(acc)[1] = make_double2(avg[2], 1.0);
            for (int k = 0; k < 3; k++) oc[k] = (float)avg[k];
        } else {
            float4* acc = (float4*)P.accum + idx;
            float inv = 1.0f / (float)P.spp;
            float avg[3] = {sum.x * inv, sum.y * inv, sum.z * inv};
            if (P.subframe > 1u) {
                float4 prev = *acc;
                float a = 1.0f / (float)P.subframe;
                avg[0] = prev.x + a * (avg[0] - prev.x);
                avg[1] = prev.y + a * (avg[1] - prev.y);
                avg[2] = prev.z + a * (avg[2] - prev.z);
            }
            *acc = make_float4(avg[0], avg[1], avg[2], 1.0f);
            for (int k = 0; k < 3; k++) oc[k] = avg[k];
        }
        if (P.image) P.image[idx] = make_color(oc[0], oc[1], oc[2]);
    }
    seg = wave_sum(seg);
    paths = wave_sum(paths);
    if (COUNT) {
        nv = wave_sum(nv);
        ns = wave_sum(ns);
    }
    if (__lane_id() == 0 && seg) {
        atomicAdd(&counters[0], (unsigned long long)seg);
        atomicAdd(&counters[1], (unsigned long long)paths);
        if (COUNT) {
            atomicAdd(&counters[2], (unsigned long long)nv);
            atomicAdd(&counters[3], (unsigned long long)ns);
        }
    }
}

}  // namespace

Cam make_cam(const vzLaunchParams& p);

vzStatus megakernel_launch(const SceneView& sc, const vzLaunchParams& p, hipStream_t st,
                           unsigned long long* d_counters, bool count)
{
    MegaParams P;
    P.cam = make_cam(p);
    P.spp = p.samples_per_pixel;
    P.subframe = p.subframe_index;
    P.max_depth = p.max_depth;
    P.flags = p.flags;
    P.accum = p.accum;
    P.image = (uint32_t*)p.image;
    dim3 grid(((p.width + 15) / 16) * ((p.height + 15) / 16));
    bool fp64 = (p.flags & VZ_ACCUM_FP64) != 0;
#define VZ_MEGA(C, F) hipLaunchKernelGGL((k_mega<8, C, F>), grid, dim3(kMegaBlk), 0, st, sc, P, d_counters)
    if (count) {
        if (fp64) VZ_MEGA(true, true); else VZ_MEGA(true, false);
    } else {
        if (fp64) VZ_MEGA(false, true); else VZ_MEGA(false, false);
    }
#undef VZ_MEGA
    VZ_HIP_TRY(hipGetLastError());
    return VZ_SUCCESS;
}

}  // namespace vz
