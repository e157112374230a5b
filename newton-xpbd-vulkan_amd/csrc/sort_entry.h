// sort_entry.h -- stand-alone stable radix sort of (key, value) pairs on
// caller-owned device buffers (see sort_entry.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace xpbd {

struct SortShape {
    int passes = 0, bits = 0, items = 0;
    size_t tiles = 0;
};

// Digit passes for keys below 2^key_bits: the fewest of at most 8 bits, the
// width spread evenly over them (never below 7 bits); tiles of 256 x items keys.
SortShape sort_shape(size_t n, int key_bits);
// Scratch bytes one sort of n pairs uses (temporary keys and values, tile
// histograms, bin totals).
size_t sort_scratch_bytes(size_t n, int key_bits);
void sort_validate(const void* keys_in, const void* vals_in, const void* keys_out, const void* vals_out, size_t n,
                   int key_bits);
// Launches the passes on stream with the given scratch block.
void sort_run(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out, size_t n,
              int key_bits, void* scratch, hipStream_t stream);
// Allocates the scratch, sorts, waits for the stream and frees the scratch.
// Exhaustive check of rcp_cr / sqrt_cr against the correctly rounded
// operations (k_selfcheck.hip): mismatch counts per range.
void check_arith(uint64_t out[3], hipStream_t s);
void sort_pairs_alloc(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                      size_t n, int key_bits, hipStream_t stream);

// As sort_pairs_alloc, with the library's static device scratch instead of an
// allocation (n <= kNoAllocMaxPairs); launches on stream and returns without
// waiting.  Calls that overlap on one device share the scratch and race.
constexpr size_t kNoAllocMaxPairs = size_t(1) << 20;
void sort_pairs_noalloc(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                        size_t n, int key_bits, hipStream_t stream);

}  // namespace xpbd
