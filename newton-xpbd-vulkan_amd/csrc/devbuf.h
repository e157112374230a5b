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

// (Re)allocates b to n elements, zero filled (one element for n = 0).
template <class T> inline void dalloc(DevBuf<T>& b, size_t n) {
    if (b.p) { (void)hipFree(b.p); b.p = nullptr; b.n = 0; }
    size_t m = n ? n : 1;
    HIPCHECK(hipMalloc(reinterpret_cast<void**>(&b.p), m * sizeof(T)));
    HIPCHECK(hipMemset(b.p, 0, m * sizeof(T)));
    b.n = n;
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
