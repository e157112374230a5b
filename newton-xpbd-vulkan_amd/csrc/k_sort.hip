// k_sort.hip -- stable LSD radix sort of (cell key, particle id) pairs for
// gfx950: the replacement of sort_histogram / sort_scan_a/b/c / sort_scatter
// (shaders/sort_*.comp, driven from src/gpu_sort.h:110-129 and
// src/scene.cpp:885-914).  Same contract: a stable sort by key, so equal cells
// keep increasing particle ids; the result lands in the "A" pair.
//
// Per 7-bit pass (three passes cover the 21-bit cell id):
//   upsweep  -- per-tile 128-bin histogram in LDS, written bin-major
//               hist[bin][tile] (no global atomics);
//   binscan  -- one workgroup per bin scans its row over tiles in place and
//               writes the bin total;
//   scatter  -- stable in-tile ranks from wave64 ballot matching (7 ballots
//               per key give each lane its same-digit peers), per-wave bin
//               counters in LDS, an LDS-staged reorder of the tile, then
//               bin-run-contiguous global stores.
// The first pass reads (key, slot) pairs in particle-id order (k_integrate
// writes them so), which makes equal keys come out in increasing particle id.
#include "xpbd_kernels.h"

namespace xpbd {

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    return x;
}

// keys, or the x of (key, value) pairs when pairs is non-null (the first pass)
__global__ __launch_bounds__(256) void k_sort_upsweep(int n, const uint32_t* __restrict__ keys,
                                                      const uint2* __restrict__ pairs, int shift,
                                                      uint32_t* __restrict__ hist, int tiles) {
    __shared__ uint32_t h[kSortBins];
    const int tile = blockIdx.x, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (threadIdx.x < kSortBins) h[threadIdx.x] = 0;
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kSortItems; it++) {
        int idx = tile * kSortTile + w * (kSortItems * kWave) + it * kWave + lane;
        if (idx < n) atomicAdd(&h[((pairs ? pairs[idx].x : keys[idx]) >> shift) & (kSortBins - 1)], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kSortBins) hist[threadIdx.x * tiles + tile] = h[threadIdx.x];
}

// One workgroup per bin: exclusive scan of hist[bin][0..tiles) in place.
__global__ __launch_bounds__(256) void k_sort_binscan(uint32_t* __restrict__ hist, uint32_t* __restrict__ bintot,
                                                      int tiles) {
    __shared__ uint32_t wsum[4];
    const int b = blockIdx.x, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t* row = hist + size_t(b) * tiles;
    uint32_t carry = 0;
    for (int base = 0; base < tiles; base += 256) {
        int t = base + threadIdx.x;
        uint32_t x = t < tiles ? row[t] : 0u;
        uint32_t inc = wave_incl_scan(x);
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        uint32_t wpre = 0, tot = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            uint32_t s = wsum[k];
            if (k < w) wpre += s;
            tot += s;
        }
        if (t < tiles) row[t] = carry + wpre + inc - x;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) bintot[b] = carry;
}

// GATHER (last pass only): the values are slots of the previous order; the
// particle's predicted position, velocity, flags and id move to the sorted
// slot (the new order of the state: the particle-particle solve reads
// neighbours as contiguous runs), and rank records previous slot -> new slot.
template <bool PAIRS, bool GATHER>
__global__ __launch_bounds__(256) void k_sort_scatter(int n, const uint32_t* __restrict__ keys_in,
                                                      const uint32_t* __restrict__ vals_in,
                                                      const uint2* __restrict__ pairs_in,
                                                      uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
                                                      int shift, const uint32_t* __restrict__ hist,
                                                      const uint32_t* __restrict__ bintot, int tiles, SortGather g) {
    __shared__ uint32_t s_cnt[4][kSortBins];    // per-wave bin counters, then per-wave bin offsets
    __shared__ uint32_t s_excl[kSortBins];      // tile-local exclusive bin starts (staging order)
    __shared__ uint32_t s_gbase[kSortBins];     // global destination of the tile's first key per bin
    __shared__ uint32_t s_keys[kSortTile];
    __shared__ uint32_t s_vals[kSortTile];
    __shared__ uint32_t s_part[2];
    const int tile = blockIdx.x, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int k = threadIdx.x; k < 4 * kSortBins; k += 256) (&s_cnt[0][0])[k] = 0;
    __syncthreads();

    uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
#pragma unroll
    for (int it = 0; it < kSortItems; it++) {
        const int idx = tile * kSortTile + w * (kSortItems * kWave) + it * kWave + lane;
        const bool valid = idx < n;
        if (PAIRS) {
            const uint2 kv = valid ? pairs_in[idx] : make_uint2(0xFFFFFFFFu, 0u);
            key[it] = kv.x;
            val[it] = kv.y;
        } else {
            key[it] = valid ? keys_in[idx] : 0xFFFFFFFFu;
            val[it] = valid ? vals_in[idx] : 0u;
        }
        const uint32_t d = (key[it] >> shift) & (kSortBins - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kSortBits; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t before = 0;
        if (valid) before = s_cnt[w][d];
        rank[it] = before + uint32_t(__popcll(peers & lt));
        // the lowest peer bumps the counter; every lane of the wave read it above
        if (valid && (peers & lt) == 0) s_cnt[w][d] = before + uint32_t(__popcll(peers));
    }
    __syncthreads();
    // per-bin offsets of each wave inside the tile, tile totals, staging starts
    uint32_t total = 0;
    if (threadIdx.x < kSortBins) {
        const int b = threadIdx.x;
        const uint32_t c0 = s_cnt[0][b], c1 = s_cnt[1][b], c2 = s_cnt[2][b], c3 = s_cnt[3][b];
        s_cnt[0][b] = 0; s_cnt[1][b] = c0; s_cnt[2][b] = c0 + c1; s_cnt[3][b] = c0 + c1 + c2;
        total = c0 + c1 + c2 + c3;
    }
    // exclusive scan of 128 tile totals (threads 0..127 = waves 0,1) and of bin totals
    uint32_t bt = threadIdx.x < kSortBins ? bintot[threadIdx.x] : 0u;
    uint32_t inc_t = wave_incl_scan(total);
    uint32_t inc_b = wave_incl_scan(bt);
    if (threadIdx.x == 63) { s_part[0] = inc_t; s_part[1] = inc_b; }
    __syncthreads();
    if (threadIdx.x < kSortBins) {
        const int b = threadIdx.x;
        const uint32_t add_t = (b >= 64) ? s_part[0] : 0u, add_b = (b >= 64) ? s_part[1] : 0u;
        s_excl[b] = add_t + inc_t - total;
        s_gbase[b] = (add_b + inc_b - bt) + hist[size_t(b) * tiles + tile];
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kSortItems; it++) {
        const int idx = tile * kSortTile + w * (kSortItems * kWave) + it * kWave + lane;
        if (idx < n) {
            const uint32_t d = (key[it] >> shift) & (kSortBins - 1);
            const uint32_t pos = s_excl[d] + s_cnt[w][d] + rank[it];
            if (pos < uint32_t(kSortTile)) {
                s_keys[pos] = key[it];
                s_vals[pos] = val[it];
            }
        }
    }
    __syncthreads();
    const int cnt = min(kSortTile, n - tile * kSortTile);
    for (int e = threadIdx.x; e < cnt; e += 256) {
        const uint32_t k = s_keys[e];
        const uint32_t d = (k >> shift) & (kSortBins - 1);
        const uint32_t dst = s_gbase[d] + (uint32_t(e) - s_excl[d]);
        if (dst < uint32_t(n)) {
            const uint32_t v = s_vals[e];
            keys_out[dst] = k;
            vals_out[dst] = v;
            if (GATHER) {
                g.Xs[dst] = g.P[v];
                g.Vs[dst] = g.V[v];
                g.Fs[dst] = g.flags[v];
                g.ids_out[dst] = g.ids[v];
                g.rank[v] = dst;
            }
        }
    }
}

void launch_sort_upsweep(int n, const uint32_t* keys, const uint2* pairs, int shift, uint32_t* hist, int tiles,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_sort_upsweep, dim3(tiles), dim3(256), 0, s, n, keys, pairs, shift, hist, tiles);
}
void launch_sort_binscan(uint32_t* hist, uint32_t* bintot, int tiles, hipStream_t s) {
    hipLaunchKernelGGL(k_sort_binscan, dim3(kSortBins), dim3(256), 0, s, hist, bintot, tiles);
}
void launch_sort_scatter(int n, const uint32_t* keys_in, const uint32_t* vals_in, const uint2* pairs_in,
                         uint32_t* keys_out, uint32_t* vals_out, int shift, const uint32_t* hist, const uint32_t* bintot,
                         int tiles, const SortGather* gather, hipStream_t s) {
    const SortGather g = gather ? *gather : SortGather{};
    if (pairs_in) {
        if (gather)
            hipLaunchKernelGGL((k_sort_scatter<true, true>), dim3(tiles), dim3(256), 0, s, n, keys_in, vals_in, pairs_in,
                               keys_out, vals_out, shift, hist, bintot, tiles, g);
        else
            hipLaunchKernelGGL((k_sort_scatter<true, false>), dim3(tiles), dim3(256), 0, s, n, keys_in, vals_in, pairs_in,
                               keys_out, vals_out, shift, hist, bintot, tiles, g);
    } else {
        if (gather)
            hipLaunchKernelGGL((k_sort_scatter<false, true>), dim3(tiles), dim3(256), 0, s, n, keys_in, vals_in, pairs_in,
                               keys_out, vals_out, shift, hist, bintot, tiles, g);
        else
            hipLaunchKernelGGL((k_sort_scatter<false, false>), dim3(tiles), dim3(256), 0, s, n, keys_in, vals_in, pairs_in,
                               keys_out, vals_out, shift, hist, bintot, tiles, g);
    }
}

}  // namespace xpbd
