// k_sort.hip -- stable LSD radix sort of (cell key, slot) pairs for gfx950:
// the replacement of sort_histogram / sort_scan_a/b/c / sort_scatter
// (shaders/sort_*.comp, driven from src/gpu_sort.h:110-129 and
// src/scene.cpp:885-914).  Same contract: a stable sort by key, so equal
// cells keep increasing particle ids; the result lands in the "A" pair.
//
// Per pass of BITS-wide digits (7 to 10 bits, 2^BITS bins; the host picks
// the fewest passes of at most 8 bits that cover the key, e.g. 3 x 7 for the
// 21-bit cell id, 3 x 8 with one wrap-plane bit per axis):
//   upsweep  -- per-tile histogram in LDS, written bin-major hist[bin][tile];
//   binscan  -- one workgroup per bin scans its row over tiles in place and
//               writes the bin total;
//   scatter  -- stable in-tile ranks from wave64 ballot matching (BITS
//               ballots per key give each lane its same-digit peers), per-wave
//               bin counters in LDS, an LDS-staged reorder of the tile, then
//               bin-run-contiguous global stores.
// Chunked passes (SortPass, at most 8-bit digits, two passes or more) have
// no binscan: the upsweep writes hist[bin][tile] (rows padded to whole
// chunks) and adds its histogram into its chunk of kSortChunkTiles tiles
// (chunk[c][bin]: one device atomic per bin, a workgroup's atomics on
// consecutive words); each scatter thread then sums, for its bin, the chunk
// totals (the bin's total, and its count before the tile's chunk) and the
// tile counts before it inside its chunk: eight 16-byte loads of its row
// segment plus one load per chunk, issued behind the tile's own keys so that
// the latencies overlap, instead of a launch.  Pass p's
// scatter zeroes the chunk totals of pass (p + 1) mod P, so every chunked
// sort starts and ends with pass 0's totals at zero (zero from allocation on).
// The first pass reads (key, slot) pairs in particle-id order (k_integrate
// writes them so), which makes equal keys come out in increasing particle id.
#include "xpbd_kernels.h"

namespace xpbd {

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) { return uint32_t(wave_incl_scan_i(int(x))); }

// keys, or the x of (key, value) pairs when pairs is non-null (the first pass)
template <int BITS, int ITEMS>
__global__ __launch_bounds__(256) void k_sort_upsweep(int n, const uint32_t* __restrict__ keys,
                                                      const uint2* __restrict__ pairs, int shift,
                                                      uint32_t* __restrict__ hist, int tiles, SortPass ch,
                                                      SortReset rs) {
    constexpr int B = 1 << BITS;
    __shared__ uint32_t h[B];
    const int tile = xcd_tile(blockIdx.x, gridDim.x), w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int b = threadIdx.x; b < B; b += 256) h[b] = 0;
    if (rs.cc_next != nullptr && blockIdx.x == 0 && threadIdx.x == 0) *rs.cc_next = 0;
    int first = tile * (256 * ITEMS), lim = n;   // the tile's keys: [first, lim)
    if (ch.seg != nullptr) {
        const int4 sg = ch.seg[tile];
        first = sg.x;
        lim = sg.x + sg.y;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        int idx = first + w * (ITEMS * kWave) + it * kWave + lane;
        if (idx < lim) atomicAdd(&h[((pairs ? pairs[idx].x : keys[idx]) >> shift) & (B - 1)], 1u);
        if (rs.rkeys != nullptr && idx < lim) {   // the last build's rows (k_integrate's reset)
            const uint32_t c = rs.rkeys[idx];
            if (idx == 0 || rs.rkeys[idx - 1] != c) *reinterpret_cast<int2*>(rs.range + size_t(2) * c) = make_int2(0, 0);
        }
        if (rs.rslot != nullptr && idx < lim) {   // a batched solver's row table, as k_integrate empties it
            const int r = rs.rslot[idx];
            if (r >= 0) rs.rows[r] = make_uint2(0u, 0u);
        }
    }
    __syncthreads();
    if (ch.chunk != nullptr) {   // padded rows, and the tile's share of its chunk's totals
        for (int b = threadIdx.x; b < B; b += 256) {
            const uint32_t c = h[b];
            hist[size_t(b) * ch.tstride + tile] = c;
            if (c != 0u) atomicAdd(ch.chunk + size_t(tile / kSortChunkTiles) * B + b, c);
        }
        return;
    }
    for (int b = threadIdx.x; b < B; b += 256) hist[size_t(b) * tiles + tile] = h[b];
}

// One 512-thread workgroup per bin: exclusive scan of hist[bin][0..tiles) in
// place (one step for up to 512 tiles, the 1,048,576-key sort).
__global__ __launch_bounds__(512) void k_sort_binscan(uint32_t* __restrict__ hist, uint32_t* __restrict__ bintot,
                                                      int tiles) {
    __shared__ uint32_t wsum[8];
    const int b = blockIdx.x, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t* row = hist + size_t(b) * tiles;
    uint32_t carry = 0;
    for (int base = 0; base < tiles; base += 512) {
        int t = base + threadIdx.x;
        uint32_t x = t < tiles ? row[t] : 0u;
        uint32_t inc = wave_incl_scan(x);
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        uint32_t wpre = 0, tot = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
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
// particle's substep-start state is integrated (integrate_particle, the same
// code k_integrate keyed it with) and stored with its flags and id at the
// sorted slot (the new order of the state: the particle-particle solve reads
// neighbours as contiguous runs), and rank records previous slot -> new slot.
template <int BITS, bool PAIRS, bool GATHER, int ITEMS>
__global__ __launch_bounds__(256) void k_sort_scatter(int n, const uint32_t* __restrict__ keys_in,
                                                      const uint32_t* __restrict__ vals_in,
                                                      const uint2* __restrict__ pairs_in,
                                                      uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
                                                      int shift, const uint32_t* __restrict__ hist,
                                                      const uint32_t* __restrict__ bintot, int tiles, SortGather g,
                                                      SortPass ch) {
    constexpr int B = 1 << BITS;
    constexpr int TILE = 256 * ITEMS;
    constexpr int BPT = B > 256 ? B / 256 : 1;   // bins per thread in the tile-level scans
    __shared__ uint32_t s_cnt[4][B];    // per-wave bin counters, then per-wave bin offsets
    __shared__ uint32_t s_excl[B];      // tile-local exclusive bin starts (staging order)
    __shared__ uint32_t s_gbase[B];     // global destination of the tile's first key per bin
    __shared__ uint32_t s_keys[TILE];
    __shared__ uint32_t s_vals[TILE];
    __shared__ uint32_t s_wsum[2][4];
    const int tile = xcd_tile(blockIdx.x, gridDim.x), w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int k = threadIdx.x; k < 4 * B; k += 256) (&s_cnt[0][0])[k] = 0;
    if (ch.chunk != nullptr)
        for (int e = blockIdx.x * 256 + threadIdx.x; e < ch.zero_count; e += gridDim.x * 256) ch.zero_next[e] = 0u;
    int first = tile * TILE, lim = n;   // the tile's keys: [first, lim)
    int4 sg = make_int4(0, 0, 0, 0);
    if (ch.seg != nullptr) {
        sg = ch.seg[tile];
        first = sg.x;
        lim = sg.x + sg.y;
    }
    __syncthreads();

    uint32_t key[ITEMS], val[ITEMS], rank[ITEMS];
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        const int idx = first + w * (ITEMS * kWave) + it * kWave + lane;
        const bool valid = idx < lim;
        if (PAIRS) {
            const uint2 kv = valid ? pairs_in[idx] : make_uint2(0xFFFFFFFFu, 0u);
            key[it] = kv.x;
            val[it] = kv.y;
        } else {
            key[it] = valid ? keys_in[idx] : 0xFFFFFFFFu;
            val[it] = valid ? vals_in[idx] : 0u;
        }
    }
    // chunked pass: each thread's bins' totals and counts before this tile
    // (row segments and chunk totals in 16-byte loads, requested behind the
    // keys, consumed after the ranking)
    uint32_t gt[BPT], pre[BPT];
    if (ch.chunk != nullptr) {
        const int ct = tile / kSortChunkTiles, m = tile - ct * kSortChunkTiles;
#pragma unroll
        for (int j = 0; j < BPT; j++) {
            const int b = min(threadIdx.x * BPT + j, B - 1);
            const uint4* hq = reinterpret_cast<const uint4*>(hist + size_t(b) * ch.tstride + size_t(ct) * kSortChunkTiles);
            uint4 hv[kSortChunkTiles / 4];
#pragma unroll
            for (int q = 0; q < kSortChunkTiles / 4; q++) hv[q] = hq[q];
            const uint32_t* cb = ch.chunk + b;
            uint32_t t = 0, p = 0;
            for (int c = 0; c < ch.chunks; c += 4) {
                uint32_t v[4];
#pragma unroll
                for (int k = 0; k < 4; k++) v[k] = cb[size_t(min(c + k, ch.chunks - 1)) * B];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t x = c + k < ch.chunks ? v[k] : 0u;
                    t += x;
                    p += c + k < ct ? x : 0u;
                }
            }
#pragma unroll
            for (int q = 0; q < kSortChunkTiles / 4; q++) {
                const int k = 4 * q;
                p += (k < m ? hv[q].x : 0u) + (k + 1 < m ? hv[q].y : 0u) + (k + 2 < m ? hv[q].z : 0u) +
                     (k + 3 < m ? hv[q].w : 0u);
            }
            gt[j] = t;
            pre[j] = p;
        }
    }
    // segmented pass: each thread's bins' global prefixes at this tile, at
    // the segment's first tile and at the next segment's first tile (the bin
    // total past the last), requested behind the keys
    uint32_t spt[BPT], sp0[BPT], sp1[BPT];
    uint32_t segfirst = 0;
    if (ch.seg != nullptr) {
        segfirst = uint32_t(ch.seg[sg.z].x);
#pragma unroll
        for (int j = 0; j < BPT; j++) {
            const int b = min(threadIdx.x * BPT + j, B - 1);
            const uint32_t* row = hist + size_t(b) * tiles;
            spt[j] = row[tile];
            sp0[j] = row[sg.z];
            sp1[j] = sg.w < tiles ? row[sg.w] : bintot[b];
        }
    }
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        const int idx = first + w * (ITEMS * kWave) + it * kWave + lane;
        const bool valid = idx < lim;
        const uint32_t d = (key[it] >> shift) & (B - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < BITS; b++) {
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
    // per-bin offsets of each wave inside the tile; each thread then owns
    // BPT consecutive bins for the block-wide scans of the tile totals (the
    // staging starts) and of the bin totals (the global starts)
    uint32_t tot[BPT], bt[BPT];
    uint32_t tsum = 0, bsum = 0;
#pragma unroll
    for (int j = 0; j < BPT; j++) {
        const int b = threadIdx.x * BPT + j;
        tot[j] = 0;
        bt[j] = 0;
        if (b < B) {
            const uint32_t c0 = s_cnt[0][b], c1 = s_cnt[1][b], c2 = s_cnt[2][b], c3 = s_cnt[3][b];
            s_cnt[0][b] = 0; s_cnt[1][b] = c0; s_cnt[2][b] = c0 + c1; s_cnt[3][b] = c0 + c1 + c2;
            tot[j] = c0 + c1 + c2 + c3;
            bt[j] = ch.seg != nullptr ? sp1[j] - sp0[j] : (ch.chunk != nullptr ? gt[j] : bintot[b]);
        }
        tsum += tot[j];
        bsum += bt[j];
    }
    const uint32_t inc_t = wave_incl_scan(tsum), inc_b = wave_incl_scan(bsum);
    if (lane == 63) { s_wsum[0][w] = inc_t; s_wsum[1][w] = inc_b; }
    __syncthreads();
    uint32_t et = inc_t - tsum, eb = inc_b - bsum;
    for (int k = 0; k < w; k++) { et += s_wsum[0][k]; eb += s_wsum[1][k]; }
#pragma unroll
    for (int j = 0; j < BPT; j++) {
        const int b = threadIdx.x * BPT + j;
        if (b < B) {
            s_excl[b] = et;
            if (ch.seg != nullptr) s_gbase[b] = segfirst + eb + (spt[j] - sp0[j]);
            else s_gbase[b] = eb + (ch.chunk != nullptr ? pre[j] : hist[size_t(b) * tiles + tile]);
        }
        et += tot[j];
        eb += bt[j];
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        const int idx = first + w * (ITEMS * kWave) + it * kWave + lane;
        if (idx < lim) {
            const uint32_t d = (key[it] >> shift) & (B - 1);
            const uint32_t pos = s_excl[d] + s_cnt[w][d] + rank[it];
            if (pos < uint32_t(TILE)) {
                s_keys[pos] = key[it];
                s_vals[pos] = val[it];
            }
        }
    }
    __syncthreads();
    const int cnt = min(TILE, lim - first);
    for (int e = threadIdx.x; e < cnt; e += 256) {
        const uint32_t k = s_keys[e];
        const uint32_t d = (k >> shift) & (B - 1);
        const uint32_t dst = s_gbase[d] + (uint32_t(e) - s_excl[d]);
        if (dst < uint32_t(n)) {
            const uint32_t v = s_vals[e];
            keys_out[dst] = k;
            vals_out[dst] = v;
            if (GATHER) {
                float4 P = g.P[v], V = g.V[v];
                const int f = g.flags[v];
                integrate_particle(P, V, f, g.dt, g.vmax, g.gx, g.gy, g.gz);
                g.Xs[dst] = P;
                g.Vs[dst] = V;
                g.Fs[dst] = f;
                g.ids_out[dst] = g.ids[v];
                g.rank[v] = dst;
            }
        }
    }
}

template <int BITS, bool PAIRS, bool GATHER, int ITEMS>
static void scatter(int n, const uint32_t* keys_in, const uint32_t* vals_in, const uint2* pairs_in, uint32_t* keys_out,
                    uint32_t* vals_out, int shift, const uint32_t* hist, const uint32_t* bintot, int tiles,
                    const SortGather& g, const SortPass& ch, hipStream_t s) {
    hipLaunchKernelGGL((k_sort_scatter<BITS, PAIRS, GATHER, ITEMS>), dim3(tiles), dim3(256), 0, s, n, keys_in, vals_in,
                       pairs_in, keys_out, vals_out, shift, hist, bintot, tiles, g, ch);
}
template <int BITS, int ITEMS>
static void scatter_bits(int n, const uint32_t* keys_in, const uint32_t* vals_in, const uint2* pairs_in,
                         uint32_t* keys_out, uint32_t* vals_out, int shift, const uint32_t* hist, const uint32_t* bintot,
                         int tiles, const SortGather* gather, const SortPass& ch, hipStream_t s) {
    const SortGather g = gather ? *gather : SortGather{};
    if (pairs_in) {
        if (gather) scatter<BITS, true, true, ITEMS>(n, keys_in, vals_in, pairs_in, keys_out, vals_out, shift, hist, bintot, tiles, g, ch, s);
        else scatter<BITS, true, false, ITEMS>(n, keys_in, vals_in, pairs_in, keys_out, vals_out, shift, hist, bintot, tiles, g, ch, s);
    } else {
        if (gather) scatter<BITS, false, true, ITEMS>(n, keys_in, vals_in, pairs_in, keys_out, vals_out, shift, hist, bintot, tiles, g, ch, s);
        else scatter<BITS, false, false, ITEMS>(n, keys_in, vals_in, pairs_in, keys_out, vals_out, shift, hist, bintot, tiles, g, ch, s);
    }
}
template <int ITEMS>
static void scatter_items(int bits, int n, const uint32_t* keys_in, const uint32_t* vals_in, const uint2* pairs_in,
                          uint32_t* keys_out, uint32_t* vals_out, int shift, const uint32_t* hist, const uint32_t* bintot,
                          int tiles, const SortGather* gather, const SortPass& ch, hipStream_t s) {
    switch (bits) {
        case 7: scatter_bits<7, ITEMS>(n, keys_in, vals_in, pairs_in, keys_out, vals_out, shift, hist, bintot, tiles, gather, ch, s); break;
        case 8: scatter_bits<8, ITEMS>(n, keys_in, vals_in, pairs_in, keys_out, vals_out, shift, hist, bintot, tiles, gather, ch, s); break;
        case 9: scatter_bits<9, ITEMS>(n, keys_in, vals_in, pairs_in, keys_out, vals_out, shift, hist, bintot, tiles, gather, ch, s); break;
        default: scatter_bits<10, ITEMS>(n, keys_in, vals_in, pairs_in, keys_out, vals_out, shift, hist, bintot, tiles, gather, ch, s);
    }
}
template <int ITEMS>
static void upsweep_items(int bits, int n, const uint32_t* keys, const uint2* pairs, int shift, uint32_t* hist, int tiles,
                          const SortPass& ch, const SortReset& rs, hipStream_t s) {
    switch (bits) {
        case 7: hipLaunchKernelGGL((k_sort_upsweep<7, ITEMS>), dim3(tiles), dim3(256), 0, s, n, keys, pairs, shift, hist, tiles, ch, rs); break;
        case 8: hipLaunchKernelGGL((k_sort_upsweep<8, ITEMS>), dim3(tiles), dim3(256), 0, s, n, keys, pairs, shift, hist, tiles, ch, rs); break;
        case 9: hipLaunchKernelGGL((k_sort_upsweep<9, ITEMS>), dim3(tiles), dim3(256), 0, s, n, keys, pairs, shift, hist, tiles, ch, rs); break;
        default: hipLaunchKernelGGL((k_sort_upsweep<10, ITEMS>), dim3(tiles), dim3(256), 0, s, n, keys, pairs, shift, hist, tiles, ch, rs);
    }
}

void launch_sort_upsweep(int bits, int items, int n, const uint32_t* keys, const uint2* pairs, int shift, uint32_t* hist,
                         int tiles, const SortPass& ch, const SortReset& rs, hipStream_t s) {
    if (items == 8) upsweep_items<8>(bits, n, keys, pairs, shift, hist, tiles, ch, rs, s);
    else upsweep_items<4>(bits, n, keys, pairs, shift, hist, tiles, ch, rs, s);
}
void launch_sort_binscan(int bits, uint32_t* hist, uint32_t* bintot, int tiles, hipStream_t s) {
    hipLaunchKernelGGL(k_sort_binscan, dim3(1 << bits), dim3(512), 0, s, hist, bintot, tiles);
}
void launch_sort_scatter(int bits, int items, int n, const uint32_t* keys_in, const uint32_t* vals_in, const uint2* pairs_in,
                         uint32_t* keys_out, uint32_t* vals_out, int shift, const uint32_t* hist, const uint32_t* bintot,
                         int tiles, const SortGather* gather, const SortPass& ch, hipStream_t s) {
    if (items == 8) scatter_items<8>(bits, n, keys_in, vals_in, pairs_in, keys_out, vals_out, shift, hist, bintot, tiles, gather, ch, s);
    else scatter_items<4>(bits, n, keys_in, vals_in, pairs_in, keys_out, vals_out, shift, hist, bintot, tiles, gather, ch, s);
}

}  // namespace xpbd
