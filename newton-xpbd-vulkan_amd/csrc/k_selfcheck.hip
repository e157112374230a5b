// k_selfcheck.hip -- a self-check of the gfx950 arithmetic the contact
// projection relies on: rcp_cr and sqrt_cr (xpbd_device.h) against the
// compiler's correctly rounded 1.0f / d and sqrtf(a) on every float of the
// ranges where the solver uses them.  The kernel only reads its loop index
// and adds to three counters.
#include "devbuf.h"
#include "xpbd_kernels.h"

namespace xpbd {

// mode 0: rcp_cr against 1.0f / d on bit patterns [lo, lo + count);
// mode 1: sqrt_cr against sqrtf
__global__ __launch_bounds__(256) void k_check_arith(uint32_t lo, uint64_t count, int mode,
                                                     unsigned long long* bad) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    unsigned long long miss = 0;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
        const float x = __uint_as_float(lo + uint32_t(i));
        float ref, got;
        if (mode == 0) {
            float xx = x;
            asm volatile("" : "+v"(xx));   // the division itself, not a folded form
            ref = 1.0f / xx;
            got = rcp_cr(x);
        } else {
            float xx = x;
            asm volatile("" : "+v"(xx));
            ref = sqrtf(xx);
            got = sqrt_cr(x);
        }
        miss += __float_as_uint(ref) != __float_as_uint(got) ? 1ull : 0ull;
    }
    if (miss) atomicAdd(bad, miss);
}

// out[0]: reciprocal mismatches for |d| in [2^-126, 2^126) (both signs);
// out[1]: square-root mismatches for a in [2^-96, +inf); out[2]: for a = +0
void check_arith(uint64_t out[3], hipStream_t s) {
    unsigned long long* bad = nullptr;
    HIPCHECK(hipMalloc(reinterpret_cast<void**>(&bad), 3 * sizeof(unsigned long long)));
    HIPCHECK(hipMemsetAsync(bad, 0, 3 * sizeof(unsigned long long), s));
    const dim3 g(8192), b(256);
    hipLaunchKernelGGL(k_check_arith, g, b, 0, s, 0x00800000u, uint64_t(0x7e800000u - 0x00800000u), 0, bad);
    hipLaunchKernelGGL(k_check_arith, g, b, 0, s, 0x80800000u, uint64_t(0xfe800000u - 0x80800000u), 0, bad);
    hipLaunchKernelGGL(k_check_arith, g, b, 0, s, 0x0f800000u, uint64_t(0x7f800000u - 0x0f800000u), 1, bad + 1);
    hipLaunchKernelGGL(k_check_arith, dim3(1), dim3(64), 0, s, 0u, uint64_t(1), 1, bad + 2);
    HIPCHECK(hipGetLastError());
    unsigned long long h[3];
    HIPCHECK(hipMemcpyAsync(h, bad, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    HIPCHECK(hipFree(bad));
    for (int k = 0; k < 3; k++) out[k] = h[k];
}

}  // namespace xpbd
