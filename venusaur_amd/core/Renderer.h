// Renderer.h -- the reference's Renderer (Core/Renderer.h:22-97) over libvz.
//
//   Init(scene, ptxSource)  context, scene upload and GPU LBVH build
//                           (Renderer.h:25-33).  ptxSource is accepted for
//                           source compatibility and ignored: the HIP code
//                           objects are linked into libvz.so.
//   Draw(camera, buffer)    one progressive frame (Renderer.h:35-78)
//   Cleanup()               releases device state (Renderer.h:80-97)
//
// Differences that fix reference defects (documented, deliberate):
//   * the accumulator is allocated once per size and the first frame after a
//     reset overwrites it instead of blending with uninitialised memory
//     (Renderer.h:37-45 with RayTracer.cu:208-213);
//   * one stream for the renderer's life, no per-frame leaks
//     (Renderer.h:65-73, 87);
//   * samples per pixel and max depth are settable (Renderer.h:53,
//     RayTracer.cu:172 hard-code 16 and 4, which stay the defaults).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "Exception.h"
#include "HIPOutputBuffer.h"
#include "RayTracer.h"
#include "Scene.h"
#include "camera.h"

class Renderer {
public:
    Renderer() = default;
    ~Renderer()
    {
        try {
            Cleanup();
        } catch (std::exception& e) {
            std::cerr << "Renderer destructor caught exception: " << e.what() << std::endl;
        }
    }
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;

    void Init(const Scene& scene, const std::string ptxSource = std::string(), int device = 0)
    {
        (void)ptxSource;
        Cleanup();
        VZ_CHECK(vzCreate(device, &m_ctx));
        HIP_CHECK(hipStreamCreateWithFlags(&m_stream, hipStreamNonBlocking));
        std::vector<vzSphere> s;
        std::vector<vzMaterial> m;
        scene.ToArrays(s, m);
        VZ_CHECK(vzSceneCreate(m_ctx, s.data(), m.data(), (uint32_t)s.size(), m_stream, &m_scene));
    }

    void Draw(Camera& camera, HIPOutputBuffer<uchar4>& outputBuffer)
    {
        if (!m_scene) throw Exception(VZ_ERROR_INVALID_VALUE, "Renderer::Draw before Init");
        const uint32_t w = (uint32_t)outputBuffer.width(), h = (uint32_t)outputBuffer.height();
        const size_t px_bytes = m_params.accum_fp64 ? 32 : 16;
        if (w != m_params.width || h != m_params.height || px_bytes != m_accum_px_bytes) {
            if (m_accum) HIP_CHECK(hipFree(m_accum));
            HIP_CHECK(hipMalloc(&m_accum, (size_t)w * h * px_bytes));
            m_accum_px_bytes = px_bytes;
            m_params.width = w;
            m_params.height = h;
            m_params.subframe_index = 0;
        }
        if (camera.Changed()) m_params.subframe_index = 0;
        outputBuffer.setStream(m_stream);
        m_params.image = outputBuffer.map();
        m_params.accum = reinterpret_cast<float4*>(m_accum);
        m_params.subframe_index++;
        const vec3 o = camera.GetPosition();
        vec3 u, v, wv;
        camera.UVWFrame(u, v, wv);
        m_params.origin = make_float3(o.x, o.y, o.z);
        m_params.u = make_float3(u.x, u.y, u.z);
        m_params.v = make_float3(v.x, v.y, v.z);
        m_params.w = make_float3(wv.x, wv.y, wv.z);
        m_params.lens_radius = camera.GetLensRadius();
        const vzLaunchParams lp = m_params.ToLaunchParams();
        VZ_CHECK(vzLaunch(m_scene, &lp, m_stream, &m_stats));
        outputBuffer.unmap();
        HIP_SYNC_CHECK();
    }

    void Cleanup()
    {
        if (m_accum) HIP_CHECK_NOTHROW(hipFree(m_accum));
        m_accum = nullptr;
        m_accum_px_bytes = 0;
        m_params.width = m_params.height = 0;
        if (m_scene) VZ_CHECK_NOTHROW(vzSceneDestroy(m_scene));
        m_scene = nullptr;
        if (m_stream) HIP_CHECK_NOTHROW(hipStreamDestroy(m_stream));
        m_stream = nullptr;
        if (m_ctx) VZ_CHECK_NOTHROW(vzDestroy(m_ctx));
        m_ctx = nullptr;
    }

    void SetSamplesPerPixel(unsigned int spp) { m_params.samples_per_pixel = spp; }
    void SetMaxDepth(unsigned int d) { m_params.max_depth = d; }
    void SetAccumFP64(bool on) { m_params.accum_fp64 = on; }
    unsigned int SubframeIndex() const { return m_params.subframe_index; }
    const vzStats& LastStats() const { return m_stats; }
    // Linear radiance of the progressive average (float4, or double4 with FP64).
    const void* AccumDevicePointer() const { return m_accum; }
    vzScene SceneHandle() const { return m_scene; }

private:
    vzContext m_ctx = nullptr;
    vzScene m_scene = nullptr;
    hipStream_t m_stream = nullptr;
    void* m_accum = nullptr;
    size_t m_accum_px_bytes = 0;
    Params m_params;
    vzStats m_stats{};
};
