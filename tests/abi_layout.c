/* Prints the size and field offsets of the C-ABI structs of include/vz_api.h
 * as JSON, so tests/test_abi.py can check the ctypes mirror against what a C
 * compiler lays out. */
#include <stddef.h>
#include <stdio.h>

#include "vz_api.h"

#define F(T, f) printf("\"%s.%s\": %zu, ", #T, #f, offsetof(T, f))
#define S(T) printf("\"%s\": %zu, ", #T, sizeof(T))

int main(void)
{
    printf("{");
    S(vzSphere); S(vzMaterial); S(vzRay); S(vzHit); S(vzLaunchParams); S(vzStats);
    S(vzAccelInfo); S(vzBatchItem); S(vzTuning); S(vzServeConfig); S(vzRequestResult); S(vzRequestDesc);
    F(vzLaunchParams, image); F(vzLaunchParams, accum); F(vzLaunchParams, width);
    F(vzLaunchParams, samples_per_pixel); F(vzLaunchParams, subframe_index); F(vzLaunchParams, origin);
    F(vzLaunchParams, u); F(vzLaunchParams, w); F(vzLaunchParams, lens_radius);
    F(vzLaunchParams, max_depth); F(vzLaunchParams, flags);
    F(vzStats, segments); F(vzStats, render_ms); F(vzStats, trace_launches); F(vzStats, hits);
    F(vzMaterial, albedo); F(vzMaterial, fuzz); F(vzMaterial, ir);
    F(vzRay, tmin); F(vzRay, dir); F(vzRay, tmax);
    F(vzBatchItem, params); F(vzServeConfig, tuning); F(vzRequestResult, segments);
    F(vzRequestResult, status);
    F(vzRequestDesc, scene_key); F(vzRequestDesc, width); F(vzRequestDesc, flags);
    F(vzRequestDesc, subframe_index); F(vzRequestDesc, lookfrom); F(vzRequestDesc, lookat);
    F(vzRequestDesc, vfov_deg); F(vzRequestDesc, focus_dist); F(vzRequestDesc, arrival_ms);
    F(vzAccelInfo, build_ms); F(vzAccelInfo, num_wide_nodes); F(vzAccelInfo, wide_depth);
    printf("\"end\": 0}\n");
    return 0;
}
