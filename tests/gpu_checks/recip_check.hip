// Exhaustive check of vz::recip_rn (venusaur_amd/csrc/vz_device.h), the
// hardware reciprocal refined by one fma Newton-Raphson step, against the
// IEEE quotient 1.0f / x for every float with |x| in [2^-100, 2^100): both
// signs, all 200 exponents, all 2^23 significands.  The quotient is also
// compared with (float)(1.0 / (double)x), which rounds once correctly
// (53 >= 2 * 24 + 2), and the unrefined v_rcp_f32 is counted as a control
// that the comparison can fail.  Test infrastructure: loaded only by
// tests/test_gpu_recip.py.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../venusaur_amd/csrc/vz_device.h"

namespace {

constexpr unsigned long long kExps = 200ull;  // biased exponents 27 .. 226: 2^-100 .. 2^99
constexpr unsigned long long kValues = 2ull * kExps << 23;

// out: [0] quotient != double reference, [1] recip_rn != quotient, [2] raw
// v_rcp_f32 != quotient, [3] values visited, [4] slots of bad taken; bad:
// the first mismatches of [1]
__global__ void k_recip(unsigned long long* out, uint32_t* bad)
{
    unsigned long long d_ref = 0, d_rn = 0, d_raw = 0, seen = 0;
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < kValues;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        const uint32_t m = (uint32_t)(i & 0x7FFFFFull);
        const uint32_t ex = (uint32_t)((i >> 23) % kExps) + 27u;
        const uint32_t sg = (uint32_t)(i / (kExps << 23)) << 31;
        const uint32_t bits = sg | (ex << 23) | m;
        const float x = __uint_as_float(bits);
        const float q = 1.0f / x;
        d_ref += __float_as_uint(q) != __float_as_uint((float)(1.0 / (double)x)) ? 1ull : 0ull;
        if (__float_as_uint(vz::recip_rn(x)) != __float_as_uint(q)) {
            const unsigned long long k = atomicAdd(&out[4], 1ull);
            if (k < 4ull) bad[k] = bits;
            d_rn++;
        }
        d_raw += __float_as_uint(__builtin_amdgcn_rcpf(x)) != __float_as_uint(q) ? 1ull : 0ull;
        seen++;
    }
    atomicAdd(&out[0], d_ref);
    atomicAdd(&out[1], d_rn);
    atomicAdd(&out[2], d_raw);
    atomicAdd(&out[3], seen);
}

}  // namespace

// Runs the check; out receives {quotient != double, recip_rn != quotient,
// raw rcp != quotient, values visited, values expected} and bad up to four
// mismatching inputs.  Returns 0, or a HIP error code.
extern "C" int vz_check_recip_rn(unsigned long long* out, uint32_t* bad)
{
    unsigned long long* d_out = nullptr;
    uint32_t* d_bad = nullptr;
    hipError_t err = hipMalloc(&d_out, 5 * sizeof(unsigned long long));
    if (err == hipSuccess) err = hipMalloc(&d_bad, 4 * sizeof(uint32_t));
    if (err == hipSuccess) err = hipMemset(d_out, 0, 5 * sizeof(unsigned long long));
    if (err == hipSuccess) err = hipMemset(d_bad, 0, 4 * sizeof(uint32_t));
    if (err == hipSuccess) {
        hipLaunchKernelGGL(k_recip, dim3(8192), dim3(256), 0, 0, d_out, d_bad);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipDeviceSynchronize();
    if (err == hipSuccess) err = hipMemcpy(out, d_out, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(bad, d_bad, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost);
    out[4] = kValues;
    if (d_out) (void)hipFree(d_out);
    if (d_bad) (void)hipFree(d_bad);
    return (int)err;
}
