// HIPOutputBuffer.h -- the pixel buffer the renderer draws into, with the
// surface of the reference's CUDAOutputBuffer (Core/CUDAOutputBuffer.h:56-105):
// type, setDevice, setStream, resize, map/unmap, width/height,
// getHostPointer.  HIP_DEVICE is hipMalloc'd memory; ZERO_COPY is mapped
// pinned host memory the GPU writes directly; HIP_P2P behaves as HIP_DEVICE.
// GL_INTEROP and the PBO calls need a display, which a headless MI355X does
// not have: they throw.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "Exception.h"

enum class HIPOutputBufferType { HIP_DEVICE = 0, GL_INTEROP = 1, ZERO_COPY = 2, HIP_P2P = 3 };

template <typename PIXEL_FORMAT>
class HIPOutputBuffer {
public:
    HIPOutputBuffer(HIPOutputBufferType type, int32_t width, int32_t height) : m_type(type)
    {
        if (type == HIPOutputBufferType::GL_INTEROP)
            throw Exception(VZ_ERROR_UNSUPPORTED, "GL interop output buffers need a display (headless GPU)");
        resize(width, height);
    }
    ~HIPOutputBuffer()
    {
        try {
            release();
        } catch (std::exception& e) {
            std::cerr << "HIPOutputBuffer destructor caught exception: " << e.what() << std::endl;
        }
    }
    HIPOutputBuffer(const HIPOutputBuffer&) = delete;
    HIPOutputBuffer& operator=(const HIPOutputBuffer&) = delete;

    void setDevice(int32_t device_idx) { m_device_idx = device_idx; }
    void setStream(hipStream_t stream) { m_stream = stream; }

    // Sizes below 1 are clamped to 1 (CUDAOutputBuffer.h:40-54).
    void resize(int32_t width, int32_t height)
    {
        if (width < 1) width = 1;
        if (height < 1) height = 1;
        if (width == m_width && height == m_height) return;
        makeCurrent();
        release();
        m_width = width;
        m_height = height;
        size_t bytes = (size_t)width * height * sizeof(PIXEL_FORMAT);
        if (m_type == HIPOutputBufferType::ZERO_COPY) {
            HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&m_host_zcopy_pixels), bytes, hipHostMallocMapped));
            HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&m_device_pixels), m_host_zcopy_pixels, 0));
        } else {
            HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&m_device_pixels), bytes));
        }
        if (!m_host_pixels.empty()) m_host_pixels.resize((size_t)m_width * m_height);
    }

    PIXEL_FORMAT* map() { return m_device_pixels; }
    void unmap()
    {
        makeCurrent();
        HIP_CHECK(hipStreamSynchronize(m_stream));
    }

    int32_t width() const { return m_width; }
    int32_t height() const { return m_height; }

    PIXEL_FORMAT* getHostPointer()
    {
        if (m_type == HIPOutputBufferType::ZERO_COPY) {
            HIP_CHECK(hipStreamSynchronize(m_stream));
            return m_host_zcopy_pixels;
        }
        m_host_pixels.resize((size_t)m_width * m_height);
        makeCurrent();
        HIP_CHECK(hipMemcpy(m_host_pixels.data(), map(), (size_t)m_width * m_height * sizeof(PIXEL_FORMAT),
                            hipMemcpyDeviceToHost));
        unmap();
        return m_host_pixels.data();
    }

    unsigned int getPBO() { throw Exception(VZ_ERROR_UNSUPPORTED, "getPBO needs OpenGL (headless GPU)"); }
    void deletePBO() { throw Exception(VZ_ERROR_UNSUPPORTED, "deletePBO needs OpenGL (headless GPU)"); }

private:
    void makeCurrent() { HIP_CHECK(hipSetDevice(m_device_idx)); }
    void release()
    {
        if (m_type == HIPOutputBufferType::ZERO_COPY) {
            if (m_host_zcopy_pixels) HIP_CHECK(hipHostFree(m_host_zcopy_pixels));
        } else if (m_device_pixels) {
            HIP_CHECK(hipFree(m_device_pixels));
        }
        m_host_zcopy_pixels = nullptr;
        m_device_pixels = nullptr;
    }

    HIPOutputBufferType m_type;
    int32_t m_width = 0, m_height = 0;
    PIXEL_FORMAT* m_device_pixels = nullptr;
    PIXEL_FORMAT* m_host_zcopy_pixels = nullptr;
    std::vector<PIXEL_FORMAT> m_host_pixels;
    hipStream_t m_stream = nullptr;
    int32_t m_device_idx = 0;
};
