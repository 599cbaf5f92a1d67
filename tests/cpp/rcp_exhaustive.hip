// rcp_exhaustive.hip -- checks pmkd::rcp_rn (pm_device.h), the reciprocal
// the Moller-Trumbore tests use for 1/det, against the IEEE quotient
// 1.0f / x for every one of the 2^32 float bit patterns on the current GPU.
// Prints "checked N mismatches M" (M counts results whose bits differ) and
// exits nonzero when M > 0.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../../extensions-photonmapping_amd/csrc/pm_device.h"

__global__ void check_kernel(unsigned long long base, unsigned long long count, unsigned long long *bad,
                             unsigned *example) {
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += stride) {
        const unsigned bits = (unsigned)(base + i);
        const float x = __uint_as_float(bits);
        if (__float_as_uint(pmkd::rcp_rn(x)) != __float_as_uint(1.0f / x)) {
            atomicAdd(bad, 1ull);
            atomicExch(example, bits);
        }
    }
}

#define CHECK(call)                                                              \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) {                                                  \
            std::fprintf(stderr, "%s: %s\n", #call, hipGetErrorString(e_));      \
            return 2;                                                            \
        }                                                                        \
    } while (0)

int main() {
    unsigned long long *bad = nullptr;
    unsigned *example = nullptr;
    CHECK(hipMalloc(&bad, sizeof(unsigned long long)));
    CHECK(hipMalloc(&example, sizeof(unsigned)));
    CHECK(hipMemset(bad, 0, sizeof(unsigned long long)));
    CHECK(hipMemset(example, 0, sizeof(unsigned)));
    const unsigned long long total = 1ull << 32, step = 1ull << 30;
    for (unsigned long long b = 0; b < total; b += step) {
        check_kernel<<<4096, 256>>>(b, step, bad, example);
        CHECK(hipGetLastError());
        CHECK(hipDeviceSynchronize());
    }
    unsigned long long h = 0;
    unsigned ex = 0;
    CHECK(hipMemcpy(&h, bad, sizeof(h), hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(&ex, example, sizeof(ex), hipMemcpyDeviceToHost));
    std::printf("checked %llu mismatches %llu example 0x%08x\n", total, h, ex);
    return h ? 1 : 0;
}
