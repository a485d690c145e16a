// Exception.h -- error convention of the C++ API: every failure throws
// Exception (a std::runtime_error carrying file:line and the failing call),
// as the reference's OPTIX_CHECK / CUDA_CHECK / CUDA_SYNC_CHECK do
// (Core/Exception.h:16-27, 90-118, 136-154).  Destructors use the NOTHROW
// variants, which log instead (Core/Exception.h:72-82, 123-134).
#pragma once

#include <hip/hip_runtime.h>

#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "../../include/vz_api.h"

class Exception : public std::runtime_error {
public:
    explicit Exception(const char* msg) : std::runtime_error(msg) {}
    explicit Exception(const std::string& msg) : std::runtime_error(msg) {}
    Exception(vzStatus status, const std::string& msg) : std::runtime_error(msg), m_status(status) {}
    vzStatus status() const { return m_status; }

private:
    vzStatus m_status = VZ_ERROR_INTERNAL;
};

#define VZ_CHECK(call)                                                                                   \
    do {                                                                                                 \
        vzStatus vz_s_ = (call);                                                                         \
        if (vz_s_ != VZ_SUCCESS) {                                                                       \
            std::stringstream ss_;                                                                       \
            ss_ << "libvz call (" << #call << ") failed with " << vzStatusString(vz_s_) << ": "           \
                << vzGetLastError() << " (" << __FILE__ << ":" << __LINE__ << ")";                       \
            throw Exception(vz_s_, ss_.str());                                                           \
        }                                                                                                \
    } while (0)

#define HIP_CHECK(call)                                                                                  \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess) {                                                                          \
            std::stringstream ss_;                                                                       \
            ss_ << "HIP call (" << #call << ") failed with error: '" << hipGetErrorString(e_) << "' ("   \
                << __FILE__ << ":" << __LINE__ << ")";                                                   \
            throw Exception(VZ_ERROR_HIP, ss_.str());                                                    \
        }                                                                                                \
    } while (0)

#define HIP_SYNC_CHECK()                                                                                 \
    do {                                                                                                 \
        hipError_t e_ = hipDeviceSynchronize();                                                          \
        if (e_ == hipSuccess) e_ = hipGetLastError();                                                    \
        if (e_ != hipSuccess) {                                                                          \
            std::stringstream ss_;                                                                       \
            ss_ << "HIP error on synchronize with error '" << hipGetErrorString(e_) << "' (" << __FILE__ \
                << ":" << __LINE__ << ")";                                                               \
            throw Exception(VZ_ERROR_HIP, ss_.str());                                                    \
        }                                                                                                \
    } while (0)

#define HIP_CHECK_NOTHROW(call)                                                                          \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            std::cerr << "HIP call (" << #call << ") failed with error: '" << hipGetErrorString(e_)       \
                      << "' (" << __FILE__ << ":" << __LINE__ << ")\n";                                  \
    } while (0)

#define VZ_CHECK_NOTHROW(call)                                                                           \
    do {                                                                                                 \
        vzStatus vz_s_ = (call);                                                                         \
        if (vz_s_ != VZ_SUCCESS)                                                                         \
            std::cerr << "libvz call (" << #call << ") failed with " << vzStatusString(vz_s_) << ": "     \
                      << vzGetLastError() << " (" << __FILE__ << ":" << __LINE__ << ")\n";               \
    } while (0)
