// dpe_rcp_check.hip -- exhaustive check of the tap projection's reciprocal:
// recip_refined(w) (one Newton step from v_rcp_f32) against the correctly
// rounded 1.0f / w for every float w with |w| in [2^-40, 2^40), both signs
// (80 exponents x 2^23 mantissas x 2 signs).  Prints the count of mismatches and
// exits 0 only when there are none; the GPU tests run it on the box's GPU.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "dpe_device.h"

namespace {

__global__ void k_rcp_check(int e0, unsigned long long *bad) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;  // mantissa
    if (m >= (1u << 23)) return;
    const int e = e0 + (int)blockIdx.y;                         // unbiased exponent
    const uint32_t sign = blockIdx.z ? 0x80000000u : 0u;
    const float w = __uint_as_float(sign | ((uint32_t)(e + 127) << 23) | m);
    const bool mismatch = __float_as_uint(dpe::recip_refined(w)) != __float_as_uint(1.0f / w);
    const unsigned long long b = __ballot(mismatch);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(bad, (unsigned long long)__popcll(b));
}

}  // namespace

int main() {
    unsigned long long *bad = nullptr, h = 0;
    if (hipMalloc(&bad, sizeof(h)) != hipSuccess || hipMemset(bad, 0, sizeof(h)) != hipSuccess) {
        std::fprintf(stderr, "dpe_rcp_check: no HIP device memory\n");
        return 2;
    }
    hipLaunchKernelGGL(k_rcp_check, dim3((1u << 23) / 256, 80, 2), dim3(256), 0, 0, -40, bad);
    if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(&h, bad, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) {
        std::fprintf(stderr, "dpe_rcp_check: kernel failed\n");
        return 2;
    }
    (void)hipFree(bad);
    std::printf("recip_refined vs 1.0f / w over |w| in [2^-40, 2^40), both signs: %llu values, %llu mismatches\n",
                80ull * 2ull * (1ull << 23), h);
    return h == 0 ? 0 : 1;
}
