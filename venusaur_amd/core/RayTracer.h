// RayTracer.h -- launch parameters and per-sphere records with the
// reference's field names (Core/RayTracer.h:3-45).  The OptiX traversable
// handle has no equivalent: the scene handle is an argument of vzLaunch.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/vz_api.h"

struct Params {
    uchar4* image = nullptr;
    float4* accum = nullptr;  // double4* when accum_fp64
    unsigned int width = 0;
    unsigned int height = 0;
    unsigned int samples_per_pixel = 16;  // Renderer.h:53
    unsigned int subframe_index = 0;
    float3 origin{0.0f, 0.0f, 0.0f};
    float3 u{0.0f, 0.0f, 0.0f}, v{0.0f, 0.0f, 0.0f}, w{0.0f, 0.0f, 0.0f};
    float lens_radius = 0.0f;
    unsigned int max_depth = 4;  // RayTracer.cu:172
    bool accum_fp64 = false;

    vzLaunchParams ToLaunchParams() const
    {
        vzLaunchParams p{};
        p.image = image;
        p.accum = accum;
        p.width = width;
        p.height = height;
        p.samples_per_pixel = samples_per_pixel;
        p.subframe_index = subframe_index;
        p.origin[0] = origin.x; p.origin[1] = origin.y; p.origin[2] = origin.z;
        p.u[0] = u.x; p.u[1] = u.y; p.u[2] = u.z;
        p.v[0] = v.x; p.v[1] = v.y; p.v[2] = v.z;
        p.w[0] = w.x; p.w[1] = w.y; p.w[2] = w.z;
        p.lens_radius = lens_radius;
        p.max_depth = max_depth;
        p.flags = accum_fp64 ? VZ_ACCUM_FP64 : 0u;
        return p;
    }
};

struct MaterialData {  // RayTracer.h:27-38, unpacked
    float3 albedo;
    float fuzz;
    float ir;
};

struct SphereHitGroupData {  // RayTracer.h:40-45
    float3 center;
    float radius;
    MaterialData mat;
};
