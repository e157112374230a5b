// devbuf.h -- device allocation helpers shared by the solver's host files.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "solver.h"

#define HIPCHECK(x)                                                                                          \
    do {                                                                                                     \
        hipError_t e_ = (x);                                                                                 \
        if (e_ != hipSuccess)                                                                                \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + __FILE__ + \
                                     ":" + std::to_string(__LINE__));                                        \
    } while (0)

namespace xpbd {

// Fills every byte of b's allocation and waits for the fill: hipMemset runs on
// the null stream, which the solvers' non-blocking streams do not wait for,
// so work queued on them next could otherwise run before the fill lands (or
// be overwritten by it).
template <class T> inline void dfill(DevBuf<T>& b, int byte) {
    HIPCHECK(hipMemset(b.p, byte, (b.n ? b.n : 1) * sizeof(T)));
    HIPCHECK(hipDeviceSynchronize());
}
// (Re)allocates b to n elements, zero filled (one element for n = 0).
template <class T> inline void dalloc(DevBuf<T>& b, size_t n) {
    if (b.p) { (void)hipFree(b.p); b.p = nullptr; b.n = 0; }
    size_t m = n ? n : 1;
    HIPCHECK(hipMalloc(reinterpret_cast<void**>(&b.p), m * sizeof(T)));
    b.n = n;
    dfill(b, 0);
}
template <class T> inline void dfree(DevBuf<T>& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
}
template <class T> inline void upload(DevBuf<T>& b, const T* src, size_t n) {
    dalloc(b, n);
    if (n) HIPCHECK(hipMemcpy(b.p, src, n * sizeof(T), hipMemcpyHostToDevice));
}

struct DeviceGuard {
    explicit DeviceGuard(int dev) { (void)hipSetDevice(dev); }
};

}  // namespace xpbd
