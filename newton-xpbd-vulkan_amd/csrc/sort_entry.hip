// sort_entry.hip -- the stable radix sort of (key, value) pairs as a
// stand-alone entry point on caller-owned device buffers: the counterpart of
// the reference's reusable GpuSort (src/gpu_sort.h:18-129, shaders
// sort_histogram / sort_scan_a/b/c / sort_scatter), which sorts (cell, id)
// pairs into its own ping-pong buffers with its own histogram scratch.  The
// passes are the solver's (k_sort.hip): per pass an upsweep of tile
// histograms, a scan of each bin's row over the tiles and a stable scatter.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <string>

#include "sort_entry.h"
#include "xpbd_kernels.h"

namespace xpbd {

namespace {

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Scratch of the allocation-free variant: the temporaries of one sort of up
// to kNoAllocMaxPairs pairs (sort_scratch_bytes is at most 8.5 MiB there for
// every key width, digits being at most 8 bits wide).  Every device the
// library runs on holds its own copy until the process ends.  In host code
// `scratch` names a host-side stand-in: its device address comes from
// hipGetSymbolAddress, per device.
constexpr size_t kNoAllocScratchBytes = size_t(9) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call there (the one set-up step; it allocates nothing).
void* scratch_on_current_device() {
    int dev = -1;
    check(hipGetDevice(&dev), "sort_pairs_noalloc: current device");
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("sort_pairs_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "sort_pairs_noalloc: scratch address");
        if (!p) throw std::runtime_error("sort_pairs_noalloc: scratch address is null");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace

SortShape sort_shape(size_t n, int key_bits) {
    SortShape s;
    s.passes = (key_bits + kSortUseBits - 1) / kSortUseBits;
    s.bits = std::max(7, (key_bits + s.passes - 1) / s.passes);
    s.items = sort_items_for((long long)n);
    const size_t tile = size_t(256) * size_t(s.items);
    s.tiles = (n + tile - 1) / tile;
    return s;
}

size_t sort_scratch_bytes(size_t n, int key_bits) {
    const SortShape s = sort_shape(n, key_bits);
    const size_t bins = size_t(1) << s.bits;
    // temporary keys and values, tile histograms, bin totals (each 256-byte aligned)
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    return 2 * up(n * sizeof(uint32_t)) + up(bins * std::max<size_t>(s.tiles, 1) * sizeof(uint32_t)) +
           up(bins * sizeof(uint32_t));
}

void sort_validate(const void* keys_in, const void* vals_in, const void* keys_out, const void* vals_out, size_t n,
                   int key_bits) {
    if (key_bits < 1 || key_bits > 32) throw std::runtime_error("sort_pairs: key_bits must be in [1, 32]");
    if (n > size_t(0x7FFFFFFF)) throw std::runtime_error("sort_pairs: at most 2^31 - 1 pairs");
    if (n && (!keys_in || !vals_in || !keys_out || !vals_out)) throw std::runtime_error("sort_pairs: null buffer");
}

// The passes over a scratch block laid out as sort_scratch_bytes describes.
// The last pass lands in (keys_out, vals_out); the input is never written.
void sort_run(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out, size_t n,
              int key_bits, void* scratch, hipStream_t stream) {
    if (n == 0) return;
    const SortShape s = sort_shape(n, key_bits);
    const size_t bins = size_t(1) << s.bits;
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    char* p = static_cast<char*>(scratch);
    uint32_t* tk = reinterpret_cast<uint32_t*>(p);
    p += up(n * sizeof(uint32_t));
    uint32_t* tv = reinterpret_cast<uint32_t*>(p);
    p += up(n * sizeof(uint32_t));
    uint32_t* hist = reinterpret_cast<uint32_t*>(p);
    p += up(bins * std::max<size_t>(s.tiles, 1) * sizeof(uint32_t));
    uint32_t* bintot = reinterpret_cast<uint32_t*>(p);
    const int tiles = int(s.tiles), nn = int(n);
    const uint32_t* kin = keys_in;
    const uint32_t* vin = vals_in;
    for (int pass = 0; pass < s.passes; pass++) {
        // passes alternate between the output and the temporaries so that
        // the last one writes the output
        const bool to_out = ((s.passes - 1 - pass) & 1) == 0;
        uint32_t* ko = to_out ? keys_out : tk;
        uint32_t* vo = to_out ? vals_out : tv;
        const int shift = pass * s.bits;
        launch_sort_upsweep(s.bits, s.items, nn, kin, nullptr, shift, hist, tiles, SortPass{}, SortReset{}, stream);
        launch_sort_binscan(s.bits, hist, bintot, tiles, stream);
        launch_sort_scatter(s.bits, s.items, nn, kin, vin, nullptr, ko, vo, shift, hist, bintot, tiles, nullptr, SortPass{},
                            stream);
        kin = ko;
        vin = vo;
    }
    check(hipGetLastError(), "sort_pairs launch");
}

void sort_pairs_alloc(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                      size_t n, int key_bits, hipStream_t stream) {
    sort_validate(keys_in, vals_in, keys_out, vals_out, n, key_bits);
    if (n == 0) return;
    void* scratch = nullptr;
    check(hipMalloc(&scratch, sort_scratch_bytes(n, key_bits)), "sort_pairs: scratch allocation");
    try {
        sort_run(keys_in, vals_in, keys_out, vals_out, n, key_bits, scratch, stream);
        check(hipStreamSynchronize(stream), "sort_pairs");
    } catch (...) {
        (void)hipStreamSynchronize(stream);
        (void)hipFree(scratch);
        throw;
    }
    check(hipFree(scratch), "sort_pairs: scratch release");
}

void sort_pairs_noalloc(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                        size_t n, int key_bits, hipStream_t stream) {
    sort_validate(keys_in, vals_in, keys_out, vals_out, n, key_bits);
    if (n > kNoAllocMaxPairs)
        throw std::runtime_error("sort_pairs_noalloc: n = " + std::to_string(n) + " exceeds the maximum of " +
                                 std::to_string(kNoAllocMaxPairs) + " pairs of its static scratch");
    if (sort_scratch_bytes(n, key_bits) > kNoAllocScratchBytes)
        throw std::runtime_error("sort_pairs_noalloc: scratch too small for this sort");
    // the set-up runs on every call (cheap after the first on a device), so a
    // call with n = 0 prepares a device before, say, a graph capture
    void* s = scratch_on_current_device();
    // nothing in the scratch has to start at zero: every pass writes its
    // histograms, bin totals and temporaries in full before reading them
    sort_run(keys_in, vals_in, keys_out, vals_out, n, key_bits, s, stream);
}

}  // namespace xpbd
