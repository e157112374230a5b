// k_particles.hip -- per-particle and per-contact kernels of the XPBD substep.
//
// Layout in HBM (one solver): positions as float4 (x, y, z, radius) and
// velocities as float4 (vx, vy, vz, inverse mass), three rotating buffers each
// (state and two iteration buffers), so a neighbour read is one 16-byte load and no
// copy-back is ever needed.  Flags are a separate int32 array.
//
// Kernel map (reference shader -> here):
//   generate_plane_contacts + integrate_particles + grid_cell_indices +
//   sort_histogram (pass 0)                 -> k_integrate
//   generate_plane_contacts (collide_only)  -> k_plane_contacts
//   grid_offsets                            -> k_cell_ranges (no table clear)
//   solve_particle_shape_contacts           -> k_shape_solve
//   solve_particle_particle_contacts +
//   apply_particle_deltas                   -> k_pp_apply (+ k_shape_apply for
//                                              particles with shape deltas)
//   apply_particle_shape_restitution        -> k_restitution
#include "xpbd_kernels.h"

#include <algorithm>

namespace xpbd {

// k_integrate's blocking: kIntItems slots per thread, wave-striped (wave w of
// a 256-thread block owns slots [w*64*kIntItems, (w+1)*64*kIntItems) of the
// block's range, 64 per step).
constexpr int kIntItems = 4;
constexpr int kIntTile = 256 * kIntItems;
__device__ __forceinline__ int tile_item(int tile, int it) {
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    return tile * kIntTile + w * (kIntItems * kWave) + it * kWave + lane;
}

// ---------------------------------------------------------------------------
// integrate_particles.comp:26-60 fused with generate_plane_contacts.comp:27-54
// (evaluated on the pre-integration position, as the reference orders them:
// scene.cpp:853-871) and grid_cell_indices.comp:17-32 on the predicted
// position, which only decides the key (the last radix pass integrates the
// state again where it stores it in cell order).  The state is stored in slot order (the previous substep's sorted
// order); the radix input is written in particle-id order, (key, slot) at
// index id, so the stable sort orders equal cells by particle id exactly as
// the reference's does.  Plane contacts name the slot.
// The force buffer and world index of the reference have no setter and are
// always zero (scene.cpp:82-85), so v1 = v0 + (0*w + g*[w>0])*dt.
__global__ __launch_bounds__(256) void k_integrate(IntegrateArgs a) {
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (a.cc_next != nullptr) *a.cc_next = 0;
        if (a.rcc_next != nullptr) *a.rcc_next = 0;
        if (a.mesh_next != nullptr) { a.mesh_next[0] = 0; a.mesh_next[1] = 0; }
    }
    float4 P[kIntItems], V[kIntItems];
    int f[kIntItems];
#pragma unroll
    for (int it = 0; it < kIntItems; it++) {   // issue every load of the tile first
        const int idx = tile_item(tile, it);
        P[it] = make_float4(0.f, 0.f, 0.f, 0.f);
        V[it] = P[it];
        f[it] = 0;
        if (idx < a.n) { P[it] = a.Ps[idx]; V[it] = a.Vs[idx]; f[it] = a.flags[idx]; }
    }
    // plane contacts of this thread's items (on the pre-integration position),
    // slots reserved once per block
    bool want[kIntItems];
    int mine = 0;
#pragma unroll
    for (int it = 0; it < kIntItems; it++) {
        const int idx = tile_item(tile, it);
        want[it] = false;
        if (a.has_plane && idx < a.n && (f[it] & 1)) {
            const float dist = P[it].z - a.altitude;
            want[it] = dist < a.margin + P[it].w;
        }
        mine += want[it] ? 1 : 0;
    }
    if (a.cc != nullptr) {
        __shared__ int s_tmp[256 / 64 + 1];
        int slot = block_reserve(a.cc, mine, s_tmp);
#pragma unroll
        for (int it = 0; it < kIntItems; it++) {
            if (!want[it]) continue;
            if (slot < a.cmax) {
                a.cpart[slot] = tile_item(tile, it);
                a.cshape[slot] = 0;
                a.cbpos[slot] = make_float4(P[it].x, P[it].y, a.altitude, 0.f);
                a.cbvel[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
                a.cnorm[slot] = make_float4(0.f, 0.f, 1.f, 0.f);
            }
            slot++;
        }
    }
#pragma unroll
    for (int it = 0; it < kIntItems; it++) {
        const int idx = tile_item(tile, it);
        const bool valid = idx < a.n;
        if (!valid) continue;
        // the predicted state only decides the key here: the last radix
        // pass integrates again where it stores the state in cell order
        float4 Pn = P[it], Vn = V[it];
        integrate_particle(Pn, Vn, f[it], a.dt, a.vmax, a.gx, a.gy, a.gz);
        if (a.rkeys != nullptr) {   // k_clear_ranges, folded in
            const uint32_t c = a.rkeys[idx];
            if (idx == 0 || a.rkeys[idx - 1] != c) *reinterpret_cast<int2*>(a.range + size_t(2) * c) = make_int2(0, 0);
        }
        if (a.rslot != nullptr) {   // k_clear_rows, folded in
            const int r = a.rslot[idx];
            if (r >= 0) a.rows[r] = make_uint2(0u, 0u);
        }
        uint32_t key = cell_of(Pn.x, Pn.y, Pn.z, a.inv_cell);
        if (a.scene != nullptr) key |= uint32_t(a.scene[idx]) << kCellBits;   // scenes keep their slot ranges
        else if (a.planes) key |= plane_of(Pn.x, Pn.y, Pn.z, a.inv_cell, a.planes) << kCellBits;
        a.kv[a.ids[idx]] = make_uint2(key, uint32_t(idx));
    }
}

// generate_plane_contacts.comp:27-54 on the current positions (collide_only).
__global__ __launch_bounds__(256) void k_plane_contacts(PlaneArgs a) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    bool want = false;
    float4 P = make_float4(0.f, 0.f, 0.f, 0.f);
    if (idx < a.n && (a.flags[idx] & 1)) {
        P = a.P[idx];
        want = (P.z - a.altitude) < a.margin + P.w;
    }
    int slot = wave_reserve(a.cc, want);
    if (want && slot < a.cmax) {
        a.cpart[slot] = idx;
        a.cshape[slot] = 0;
        a.cbpos[slot] = make_float4(P.x, P.y, a.altitude, 0.f);
        a.cbvel[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
        a.cnorm[slot] = make_float4(0.f, 0.f, 1.f, 0.f);
    }
}

// grid_offsets.comp:16-35 into one interleaved (start, end) table.  Instead
// of the reference's two 8 MiB table fills before every build
// (scene.cpp:918-923), rkeys records the keys whose rows this build writes,
// and exactly those rows are reset before the next build (by k_integrate at
// a frame's first substep, by the previous substep's k_shape_apply or the
// first radix upsweep at the others, by k_clear_ranges outside a substep),
// so every row not written by the current build reads (0, 0).
__global__ __launch_bounds__(256) void k_cell_ranges(int n, const uint32_t* __restrict__ keys, int* __restrict__ range,
                                                     uint32_t* __restrict__ rkeys) {
    int t = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t c = keys[t];
    const size_t row = size_t(2) * c;   // batched keys reach 2^32 - 1
    if (t == 0 || keys[t - 1] != c) range[row] = t;
    if (t == n - 1 || keys[t + 1] != c) range[row + 1] = t + 1;
    if (rkeys) rkeys[t] = c;
}

// k_cell_ranges plus the move of the state into the sorted order, one sorted
// slot t per thread: the particle at previous slot vals[t] is integrated
// (integrate_particle, the code k_integrate keyed it with) and stored at t
// with its flags and id, and rank records previous slot -> t.  The last radix
// pass then only moves keys and slots.
__global__ __launch_bounds__(256) void k_cell_ranges_gather(int n, const uint32_t* __restrict__ keys,
                                                            const uint32_t* __restrict__ vals, int* __restrict__ range,
                                                            uint32_t* __restrict__ rkeys, SortGather g) {
    int t = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t v = vals[t];
    float4 P = g.P[v], V = g.V[v];
    const int f = g.flags[v];
    const uint32_t id = g.ids[v];
    const uint32_t c = keys[t];
    const size_t row = size_t(2) * c;
    if (t == 0 || keys[t - 1] != c) range[row] = t;
    if (t == n - 1 || keys[t + 1] != c) range[row + 1] = t + 1;
    rkeys[t] = c;
    integrate_particle(P, V, f, g.dt, g.vmax, g.gx, g.gy, g.gz);
    g.Xs[t] = P;
    g.Vs[t] = V;
    g.Fs[t] = f;
    g.ids_out[t] = id;
    g.rank[v] = uint32_t(t);
}

// Zeroes the rows k_cell_ranges wrote from the sorted keys rkeys: with wrap
// planes the table has 2^27 rows, too many to clear whole every substep.
__global__ __launch_bounds__(256) void k_clear_ranges(int n, const uint32_t* __restrict__ rkeys, int* __restrict__ range) {
    int t = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t c = rkeys[t];
    if (t == 0 || rkeys[t - 1] != c) *reinterpret_cast<int2*>(range + size_t(2) * c) = make_int2(0, 0);
}

// grid_offsets.comp:16-35 for a batched solver, as the row table: the
// occupied groups of 8 cells are numbered in sorted order (a scan of the
// groups' first positions over tiles of kGroupTile positions), then each
// position writes what it decides.
constexpr int kGroupTile = 1024;   // 256 threads x 4 positions

__device__ __forceinline__ bool group_first(const uint32_t* __restrict__ keys, int t) {
    return t == 0 || (keys[t - 1] >> 3) != (keys[t] >> 3);
}

// Group starts per tile.
__global__ __launch_bounds__(256) void k_group_counts(int n, const uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ tilecnt) {
    __shared__ uint32_t s_w[4];
    const int tile = blockIdx.x, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int t = tile * kGroupTile + i * 256 + threadIdx.x;
        c += uint32_t(__popcll(__ballot(t < n && group_first(keys, t))));
    }
    if (lane == 0) s_w[w] = c;
    __syncthreads();
    if (threadIdx.x == 0) tilecnt[tile] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// Exclusive scan of the tile counts (one workgroup).
__global__ __launch_bounds__(1024) void k_group_scan(int tiles, const uint32_t* __restrict__ tilecnt,
                                                     uint32_t* __restrict__ tileoff) {
    __shared__ uint32_t s_w[16];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t carry = 0;
    for (int b = 0; b < tiles; b += 1024) {
        const int t = b + threadIdx.x;
        const uint32_t x = t < tiles ? tilecnt[t] : 0u;
        const uint32_t inc = uint32_t(wave_incl_scan_i(int(x)));
        if (lane == 63) s_w[w] = inc;
        __syncthreads();
        uint32_t pre = 0, tot = 0;
        for (int k = 0; k < 16; k++) {
            const uint32_t v = s_w[k];
            if (k < w) pre += v;
            tot += v;
        }
        if (t < tiles) tileoff[t] = carry + pre + inc - x;
        carry += tot;
        __syncthreads();
    }
}

// Every position: its group's number gi (the group starts at or before it,
// minus one); a cell starting at t is the first with key >= group * 8 + j for
// each j after the group's previous cell, up to its own; the group's last
// position ends every j after its cell; the last position writes the
// sentinel.  A row's mask gathers its groups' bits in LDS at the row's first
// position, which then stores the row's entry (first group, mask) and
// records the row so the next build empties it.  Only rows that cross a tile
// boundary meet in global memory, with atomics.
// With g.Xs set, each position t also moves its particle into the sorted
// order, as k_cell_ranges_gather does for a plain solver: the particle at
// previous slot vals[t] is integrated and stored at t with its flags and id,
// and rank records previous slot -> t (the last radix pass then moves only
// keys and slots).
__global__ __launch_bounds__(256) void k_group_build(int n, const uint32_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ tileoff, uint2* __restrict__ rows,
                                                     uint32_t* __restrict__ glb, int* __restrict__ rslot,
                                                     const uint32_t* __restrict__ vals, SortGather g) {
    __shared__ uint32_t s_c[16];       // group starts per chunk (chunk = 4 i + w, in position order)
    __shared__ int s_rlast[16];        // per chunk: local index of its last row start, or -1
    __shared__ uint32_t s_mask[kGroupTile];
    const int tile = blockIdx.x, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int t0 = tile * kGroupTile;
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    // the state's move: every load requested first, so that it is in flight
    // during the row table's work
    float4 gP[4], gV[4];
    int gf[4];
    uint32_t gv[4], gid[4];
    if (g.Xs != nullptr) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int t = t0 + i * 256 + threadIdx.x;
            gv[i] = t < n ? vals[t] : 0u;
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int t = t0 + i * 256 + threadIdx.x;
            gP[i] = gV[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            gf[i] = 0;
            gid[i] = 0u;
            if (t < n) {
                gP[i] = g.P[gv[i]];
                gV[i] = g.V[gv[i]];
                gf[i] = g.flags[gv[i]];
                gid[i] = g.ids[gv[i]];
            }
        }
    }
    uint32_t k[4], kp[4], kn[4], gi[4];
    uint64_t fm[4], rm[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int t = t0 + i * 256 + threadIdx.x;
        k[i] = kp[i] = kn[i] = 0u;
        if (t < n) {
            k[i] = keys[t];
            if (t > 0) kp[i] = keys[t - 1];
            if (t + 1 < n) kn[i] = keys[t + 1];
        }
        fm[i] = __ballot(t < n && (t == 0 || (kp[i] >> 3) != (k[i] >> 3)));
        rm[i] = __ballot(t < n && (t == 0 || (kp[i] >> 7) != (k[i] >> 7)));
        if (lane == 0) {
            s_c[4 * i + w] = uint32_t(__popcll(fm[i]));
            s_rlast[4 * i + w] = rm[i] ? i * 256 + w * 64 + 63 - int(__clzll((long long)rm[i])) : -1;
        }
        s_mask[i * 256 + threadIdx.x] = 0u;
    }
    __syncthreads();
    const uint32_t pre = tileoff[tile];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int c = 4 * i + w;
        uint32_t before = pre;   // group starts in the tile's earlier chunks
        for (int q = 0; q < c; q++) before += s_c[q];
        gi[i] = before + uint32_t(__popcll(fm[i] & upto)) - 1u;
        const int t = t0 + i * 256 + threadIdx.x;
        if (t >= n) continue;
        const uint32_t key = k[i], grp = key >> 3, row = key >> 7;
        const bool first = (fm[i] >> lane) & 1ull;
        const bool starts = first || kp[i] != key;
        const bool last = t + 1 == n || (kn[i] >> 3) != grp;
        uint32_t* lb = glb + 8 * size_t(gi[i]);
        const int j = int(key & 7u);
        if (starts)
            for (int q = first ? 0 : int(kp[i] & 7u) + 1; q <= j; q++) lb[q] = uint32_t(t);
        if (last)
            for (int q = j + 1; q < 8; q++) lb[q] = uint32_t(t + 1);
        if (first) {
            // the row's first position in this tile: in this chunk, else the
            // last of an earlier chunk; none when the row began in an earlier tile
            const uint64_t m = rm[i] & upto;
            int lead = m ? i * 256 + w * 64 + 63 - int(__clzll((long long)m)) : -1;
            for (int q = c - 1; lead < 0 && q >= 0; q--) lead = s_rlast[q];
            if (lead >= 0) atomicOr(&s_mask[lead], 1u << (grp & 15u));
            else atomicOr(&rows[row].y, 1u << (grp & 15u));
        }
        rslot[t] = ((rm[i] >> lane) & 1ull) ? int(row) : -1;
        if (t + 1 == n) glb[8 * size_t(gi[i] + 1)] = uint32_t(n);   // the sentinel group
    }
    __syncthreads();
    const bool spill_next = t0 + kGroupTile < n;   // the next tile may continue this tile's last row
    const uint32_t next_row = spill_next ? keys[t0 + kGroupTile] >> 7 : 0xFFFFFFFFu;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (!((rm[i] >> lane) & 1ull)) continue;
        const uint32_t row = k[i] >> 7, mask = s_mask[i * 256 + threadIdx.x];
        if (spill_next && row == next_row) {   // later tiles add their groups' bits
            rows[row].x = gi[i];
            atomicOr(&rows[row].y, mask);
        } else {
            rows[row] = make_uint2(gi[i], mask);
        }
    }
    if (g.Xs != nullptr) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int t = t0 + i * 256 + threadIdx.x;
            if (t >= n) continue;
            integrate_particle(gP[i], gV[i], gf[i], g.dt, g.vmax, g.gx, g.gy, g.gz);
            g.Xs[t] = gP[i];
            g.Vs[t] = gV[i];
            g.Fs[t] = gf[i];
            g.ids_out[t] = gid[i];
            g.rank[gv[i]] = uint32_t(t);
        }
    }
}

// Empties the rows the last build filled (outside a substep).
__global__ __launch_bounds__(256) void k_clear_rows(int n, const int* __restrict__ rslot, uint2* __restrict__ rows) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int r = rslot[t];
    if (r >= 0) rows[r] = make_uint2(0u, 0u);
}

// Dense tables exactly as the reference leaves them after grid_offsets
// (get_cell_starts / get_cell_ends); tables are zeroed by the caller.
__global__ __launch_bounds__(256) void k_cell_tables(int n, const uint32_t* __restrict__ keys, int* __restrict__ starts,
                                                     int* __restrict__ ends) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int c = int(keys[t]);
    if (t == 0) starts[c] = 0;
    else {
        int p = int(keys[t - 1]);
        if (c != p) { starts[c] = t; ends[p] = t; }
    }
    if (t == n - 1) ends[c] = t + 1;
}

// The ground plane's contact of a particle: generate_plane_contacts.comp:
// 27-54 emits it from the substep-start position x0 (active particles within
// margin + radius of the plane; a proxy's plane contact adds nothing,
// solve_particle_shape_contacts.comp:115-116),
// and shape_contact solves it on the iteration's state (P, Vv).
__device__ __forceinline__ bool plane_contact(float4 x0, float4 P, float4 Vv, int pf, float altitude, float margin,
                                              float mu, float ka, float dt, float relax, f3& fd) {
    if ((pf & 3) != 1 || !(x0.z - altitude < margin + x0.w)) return false;
    const BodyArgs none{};
    f3 dtot, r;
    return shape_contact(P, Vv, -1, none, mu, f3{x0.x, x0.y, altitude}, f3{0.f, 0.f, 0.f}, f3{0.f, 0.f, 1.f}, ka, dt,
                         relax, fd, dtot, r);
}

// solve_particle_shape_contacts.comp:104-200 over the contact buffer
// (shape_solve_one per contact; deterministic mode stores each contact's
// deltas instead).  skip_plane: the ground plane's contacts are taken by the
// particle-pair solve and the shape-delta apply instead.
__global__ __launch_bounds__(256) void k_shape_solve(ShapeArgs a) {
    const int count = min(a.cmax, *a.cc);
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x) {
        const int pi = a.cpart[t];
        int s = a.cshape[t];
        if (a.skip_plane && s == 0) continue;
        if (a.cfd == nullptr) {
            shape_solve_one(a, t, pi, s, xyz(a.cbpos[t]), xyz(a.cbvel[t]), xyz(a.cnorm[t]));
            continue;
        }
        int body = a.body.shape_body[s];
        int pf = a.flags[pi];
        // deterministic mode: no contribution unless stored below
        a.cfd[t] = make_float4(0.f, 0.f, 0.f, 0.f);
        a.cbd[2 * size_t(t)] = make_float4(0.f, 0.f, 0.f, 0.f);
        if ((pf & 1) == 0) continue;
        if (pf & 2) {
            if (body < 0) continue;
            if (a.body.flags[body] & 4) continue;
            if (a.body.invm[body] == 0.0f) continue;
        }
        float4 P = a.P[pi], Vv = a.V[pi];
        if (a.from_old) integrate_particle(P, Vv, pf, a.g_dt, a.g_vmax, a.gx, a.gy, a.gz);
        const float mu = 0.5f * (a.pmu + a.body.mat_mu[s]);
        f3 fd, dtot, r;
        if (!shape_contact(P, Vv, body, a.body, mu, xyz(a.cbpos[t]), xyz(a.cbvel[t]), xyz(a.cnorm[t]), a.ka, a.dt,
                           a.relax, fd, dtot, r))
            continue;
        a.cfd[t] = make_float4(fd.x, fd.y, fd.z, 1.f);
        if (body >= 0 && a.body.invm[body] > 0.0f) {
            f3 dv = {dtot.x / a.dt, dtot.y / a.dt, dtot.z / a.dt};
            f3 dw = cross3(r, dv);
            a.cbd[2 * size_t(t)] = make_float4(-dv.x, -dv.y, -dv.z, 1.f);
            a.cbd[2 * size_t(t) + 1] = make_float4(-dw.x, -dw.y, -dw.z, 0.f);
        }
    }
}

// One neighbour candidate of solve_particle_particle_contacts.comp:150-210.
// The candidate at sorted slot k has passed the cheap distance reject; the
// rest of the reference's tests and the projection follow in its order
// (self, inactive, proxy, box, refined distance).  Every test is a pure
// reject, so testing the distance first changes nothing.
template <bool DIAG>
__device__ __forceinline__ void pp_pair(const PPArgs& a, int t, int k, float4 Pj, float4 Vj, int nf, int i, int proxy, f3 x,
                                        f3 v, float rad, float w1, float R, f3& acc) {
    // the first four tests in one branch, so the flags and both state loads
    // are in flight together
    const f3 rd = x - xyz(Pj);
    const int skip = int(k == t) | int((nf & 1) == 0) | (int(proxy != 0) & (int((nf & 2) != 0) | int(Vj.w <= 0.0f))) |
                     int(fabsf(rd.x) > R) | int(fabsf(rd.y) > R) | int(fabsf(rd.z) > R);
    if (skip) return;
    const RootInv dl = exact_length_inv_fast(rd);
    const float d = dl.d;
    const float err = d - rad - Pj.w;
    if (err > a.coh || d <= 0.0f) return;
    const float w2 = Vj.w;
    const float denom = w1 + w2;
    if (denom <= 0.0f) return;
    const f3 nn = exact_div_by(rd, d, dl.inv);
    const f3 vrel = v - xyz(Vj);
    const float lambda_n = err;
    const float vn = exact_dot(nn, vrel);
    const f3 vt = {fma_(vn, -nn.x, vrel.x), fma_(vn, -nn.y, vrel.y), fma_(vn, -nn.z, vrel.z)};
    const RootInv vl = exact_length_inv_fast(vt);
    const float vt_len = vl.d;
    const float lambda_f = gmax(a.mu * lambda_n, -vt_len * a.dt);
    const f3 vt_dir = vt_len > 0.0f ? exact_div_by(vt, vt_len, vl.inv) : f3{0.f, 0.f, 0.f};
    const f3 p2 = nn * err;
    const f3 df = {fma_(vt_dir.x, lambda_f, -p2.x), fma_(vt_dir.y, lambda_f, -p2.y), fma_(vt_dir.z, lambda_f, -p2.z)};
    const f3 pair_d = exact_div_by(df, denom, recip_fast(denom));
    acc = acc + pair_d;
    if (DIAG) {
        const uint32_t slot = atomicAdd(a.diag_header, 1u);
        if (slot < a.max_diag_pairs) {
            const f3 dw = {fma_(pair_d.x, w1, 0.f), fma_(pair_d.y, w1, 0.f), fma_(pair_d.z, w1, 0.f)};
            const f3 di = {fma_(dw.x, a.relax, 0.f), fma_(dw.y, a.relax, 0.f), fma_(dw.z, a.relax, 0.f)};
            float* rec = a.diag_pairs + 20 * size_t(slot);
            reinterpret_cast<uint32_t*>(rec)[0] = uint32_t(i);
            reinterpret_cast<uint32_t*>(rec)[1] = a.ids[k];
            rec[2] = d; rec[3] = err;
            rec[4] = nn.x; rec[5] = nn.y; rec[6] = nn.z;
            rec[7] = vrel.x; rec[8] = vrel.y; rec[9] = vrel.z;
            rec[10] = vn; rec[11] = vt_len; rec[12] = denom; rec[13] = lambda_n; rec[14] = lambda_f;
            rec[15] = di.x; rec[16] = di.y; rec[17] = di.z;
            reinterpret_cast<uint32_t*>(rec)[18] = a.iter_index;
            reinterpret_cast<uint32_t*>(rec)[19] = 0u;
        } else {
            atomicAdd(a.diag_header + 1, 1u);
        }
    }
}

// The next substep's radix input for the particle finished at sorted slot t
// with state (P, V) and flags f: what k_integrate computes at the start of
// the next substep (integrate_particle, the cell key, the wrap plane or a
// batched solver's scene above it), stored as (key, t) at the particle's id.
__device__ __forceinline__ void write_next_key(const NextKey& nk, int t, float4 P, float4 V, int f, float dt,
                                               float vmax) {
    integrate_particle(P, V, f, dt, vmax, nk.gx, nk.gy, nk.gz);
    uint32_t key = cell_of(P.x, P.y, P.z, nk.inv_cell);
    if (nk.keys != nullptr) key |= nk.keys[t] & ~((1u << kCellBits) - 1u);   // the particle's scene
    else if (nk.planes) key |= plane_of(P.x, P.y, P.z, nk.inv_cell, nk.planes) << kCellBits;
    nk.kv[nk.ids[t]] = make_uint2(key, uint32_t(t));
}

// apply_particle_deltas.comp:44-57 for sorted slot t: position x (radius
// rad, inverse mass w1) moved by its accumulated delta dl from the
// substep-start position x0; with nk.kv set, the next substep's key of the
// result too (flags f).
__device__ __forceinline__ void apply_delta(f3 x0, int t, f3 x, float rad, float w1, f3 dl, float dt, float vmax,
                                            float4* Po, float4* Vo, const NextKey& nk, int f) {
    f3 xn = x + dl;
    f3 vnew = markstein_div(xn - x0, dt);
    const float mag = len3(vnew);
    if (mag > vmax) {
        vnew = vnew * (vmax / mag);
        xn = x0 + vnew * dt;
    }
    const float4 P = make_float4(xn.x, xn.y, xn.z, rad), V = make_float4(vnew.x, vnew.y, vnew.z, w1);
    Po[t] = P;
    Vo[t] = V;
    if (nk.kv != nullptr) write_next_key(nk, t, P, V, f, dt, vmax);
}

// solve_particle_particle_contacts.comp:107-218 fused with
// apply_particle_deltas.comp:38-58.  One thread per sorted slot t (the
// reference's thread order), particle i = ids[t]; the output state and deltas
// live in sorted order, the substep-start positions at slot vals[t] of the
// previous order.  Particle i's delta is (shape contributions) + (its own
// particle-particle sum): the reference accumulates exactly these with float
// atomics into a zeroed buffer, the shape dispatch first, and with at most two
// shape contacts per particle their sum is order independent.  This kernel
// finishes every particle as if its shape sum were +0 (the sum never rounds
// to -0, so 0 + pp is what the reference adds), independently of the shape
// solve; k_shape_apply then redoes the few particles whose sum is not.
//
// Neighbour state is read from the sorted copies: the query box visits cells z
// -> y -> x -> sorted slot, and the x-adjacent cells of one row are consecutive
// cell ids, hence one contiguous run of sorted slots in the same order.  A row
// is split where the x index wraps at 128; a virtual x below the clamp (the
// reference maps every such cell to x = 0) is visited cell by cell.
constexpr int kSurv = 16;   // recorded survivors per lane before a drain

// Measured on the 1,048,576-particle pile (rocprofv3 SQ counters): the
// vector ALUs are busy about two thirds of the kernel's time and the waves
// wait on their chain of dependent loads (own state, cell rows, candidate
// batches, survivors) most of the rest.  Candidates are read four at a time,
// the last batch of a run masked, so a run of up to four costs one round
// trip.  Six waves per SIMD (80 VGPRs) measured faster than five with wider
// batches or paired survivor drains.
constexpr int kScanBatch = 4;

template <bool DIAG, bool BATCHED, int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_pp_apply(PPArgs a) {
    constexpr int SB = kScanBatch;
    __shared__ uint32_t s_surv[kSurv][BLOCK];
    // Every lane stays to the end (the final drain shares the wave's
    // survivors over all of its lanes): a lane past the particle count
    // shadows the last slot, and it and an inactive particle's lane scan an
    // empty box and write nothing more.
    const int tb = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x;
    const bool in_range = tb + int(threadIdx.x) < a.n;
    const int t = in_range ? tb + int(threadIdx.x) : a.n - 1;
    const int pf = a.Fs[t];
    const float4 Pi = a.Xs[t], Vi = a.Vs[t];
    const bool live = in_range && (pf & 1) != 0;
    if (in_range && !live) {   // inactive: passthrough (the reference leaves x_out unwritten)
        a.delta[3 * t] = 0.0f; a.delta[3 * t + 1] = 0.0f; a.delta[3 * t + 2] = 0.0f;
        a.Po[t] = Pi;
        a.Vo[t] = Vi;
        if (a.nk.kv != nullptr) write_next_key(a.nk, t, Pi, Vi, pf, a.dt, a.vmax);
    }
    const int i = DIAG ? int(a.ids[t]) : 0;
    const int proxy = pf & 2;
    const f3 x = xyz(Pi), v = xyz(Vi);
    const float rad = Pi.w, w1 = Vi.w;
    const float R = rad + a.max_r + a.coh;
    const float inv = a.inv_cell;
    const int xs = (int)floorf((x.x - R) * inv), ys = (int)floorf((x.y - R) * inv), zs = (int)floorf((x.z - R) * inv);
    const int xe = min((int)floorf((x.x + R) * inv), xs + kGridDim - 1);
    const int ye = min((int)floorf((x.y + R) * inv), ys + kGridDim - 1);
    const int ze = live ? min((int)floorf((x.z + R) * inv), zs + kGridDim - 1) : zs - 1;
    const int2* range = reinterpret_cast<const int2*>(a.range);
    // batched mode: the particle's scene bits complete the key of each
    // visited cell; with wrap planes each visited virtual cell selects its
    // plane's table.  Every table holds exactly the current substep's cells:
    // the previous ones are reset before each build.
    const uint32_t sbase = BATCHED ? (a.keys[t] & ~((1u << kCellBits) - 1u)) : 0u;
    auto table = [&](int cx, int cy, int cz) -> uint32_t {
        return BATCHED ? sbase : (a.planes ? wrap_plane(cx, cy, cz, a.planes) << kCellBits : 0u);
    };
    auto raw = [&](uint32_t base, int cell) -> int2 { return range[size_t(base) + size_t(cell)]; };
    // cheap reject before the refined length: a pair the exact test accepts
    // has |rd| <= (r_i + r_j + coh)(1 + 5 ulp) <= (r_i + r_max + |coh|)(1 + 5 ulp)
    // (r_max bounds every radius), and this squared length (the refined
    // length's own fma sum) is within 2 ulp of |rd|^2, far inside the 1e-4
    // relative slack
    const float lim = rad + a.rmax + fabsf(a.coh);
    const float lim2 = lim * lim * 1.0001f + 1e-30f;
    auto near = [&](f3 Pj) -> bool {
        const float dx = x.x - Pj.x, dy = x.y - Pj.y, dz = x.z - Pj.z;
        return !(fma_(dz, dz, fma_(dx, dx, dy * dy)) > lim2);
    };
    // the scan needs positions only (12-byte loads; the radius enters
    // through r_max), addressed by a 32-bit byte offset from the array base
    // (16 bytes per slot, at most 2^28 slots)
    auto pos = [&](int k) -> f3 {
        const float3 p = *reinterpret_cast<const float3*>(reinterpret_cast<const char*>(a.Xs) + (uint32_t(k) << 4));
        return f3{p.x, p.y, p.z};
    };
    f3 acc = {0.f, 0.f, 0.f};
    // Two phases keep the lanes of a wave together: the scan applies the
    // cheap reject to every candidate and records the survivors' slots (in
    // visiting order) in this lane's LDS column; the projection then runs over
    // the recorded slots, where each lane's j-th survivor is processed at the
    // same step, its three loads issued together.  A full column is drained
    // before the scan continues.
    int cnt = 0;
    auto drain = [&]() {
        for (int m = 0; m < cnt; m++) {
            const int k = int(s_surv[m][threadIdx.x]);
            const float4 Pj = a.Xs[k], Vj = a.Vs[k];
            const int nf = a.Fs[k];
            pp_pair<DIAG>(a, t, k, Pj, Vj, nf, i, proxy, x, v, rad, w1, R, acc);
        }
        cnt = 0;
    };
    // a lane about to overflow drains together with every lane of the wave
    // at the same loop point, so the projections stay in lockstep
    auto room = [&](int need) {
        if (__any(cnt > kSurv - need)) drain();
    };
    auto scan = [&](int s0, int e0) {
        // SB loads issue before the tests; the last batch of a run reads its
        // last slot in place of the slots past the run and tests only those
        // inside it (one round trip instead of one per remaining candidate).
        // Each candidate is written to the next column slot unconditionally
        // and kept (cnt advanced) if it is inside the run, not the particle
        // itself and near: no branch per candidate.  room() leaves SB free
        // slots, so the write stays inside the column.
        for (int k = s0; k < e0; k += SB) {
            f3 P[SB];
#pragma unroll
            for (int j = 0; j < SB; j++) P[j] = pos(min(k + j, e0 - 1));
            room(SB);
#pragma unroll
            for (int j = 0; j < SB; j++) {
                const int kk = k + j;
                const bool keep = (kk < e0) & (kk != t) & near(P[j]);
                s_surv[cnt][threadIdx.x] = uint32_t(kk);
                cnt += keep ? 1 : 0;
            }
        }
    };
    // One x-run of a row: cells [cx, cx + len) inside one wrap, hence
    // consecutive cell ids and one contiguous run of sorted slots.  The first
    // (up to) three cells' table rows may come in already loaded.
    auto run = [&](int cx, int cy, int cz, int len, int2 p0, int2 p1, int2 p2, bool pre) {
        const int c0 = cell_index(cx, cy, cz);
        const uint32_t base = table(cx, cy, cz);
        if constexpr (BATCHED) {
            const int2 sp = row_span(a.rowtab, base + uint32_t(c0), len);
            scan(sp.x, sp.y);
            return;
        }
        int s0 = 0, e0 = 0;
        bool any = false;
        for (int m = 0; m < len; m++) {
            int2 rg;
            if (pre && m < 3) rg = m == 0 ? p0 : (m == 1 ? p1 : p2);
            else rg = raw(base, c0 + m);
            if (rg.y > rg.x) {
                if (!any) s0 = rg.x;
                e0 = rg.y;
                any = true;
            }
        }
        scan(s0, e0);
    };
    const int vx0 = xs + kGridOrigin;
    const bool clamp0 = vx0 < 0;   // clamped virtual cells: one physical cell per run
    const int len0 = clamp0 ? 1 : min(xe - xs + 1, kGridDim - vx0 % kGridDim);
    // the rest of a row after its first x-run: past an x wrap, or clamped cells
    auto rest = [&](int cy, int cz) {
        int cx = xs + len0;
        while (cx <= xe) {
            const int vx = cx + kGridOrigin;
            const int len = vx < 0 ? 1 : min(xe - cx + 1, kGridDim - vx % kGridDim);
            run(cx, cy, cz, len, make_int2(0, 0), make_int2(0, 0), make_int2(0, 0), false);
            cx += len;
        }
    };
    if ((xe - xs <= 2) & (ye - ys <= 2) & (ze - zs <= 2) & !clamp0 & (ys + kGridOrigin >= 0) &
        (zs + kGridOrigin >= 0)) {
        // A query box of at most 3 x 3 x 3 cells (every query when R is at
        // most one cell edge).  A row's cells [xs, xe] are one run of sorted
        // slots, or two when the row crosses an x wrap (piece a: the first
        // len0 cells, piece b: the rest, in the next wrap's table); all rows
        // split at the same cell.  A piece's run is [last end - total length,
        // last end) of its cells (unoccupied cells read (0, 0)).  Each z slab
        // is scanned slab by slab as two streams of at most three runs --
        // (row 0, row 1, row 2) and nothing, or (a0, b0, a1) and (b1, a2, b2)
        // when the rows split -- in batches of SB: flat index f maps to its
        // run's slot by two compares against the runs' cumulative lengths.
        // The visiting order, hence every sum, is the general path's.  No
        // virtual coordinate of the box is clamped, so a cell's key is its
        // wrap plane's bits over the low seven bits of each coordinate: a
        // slab's z part, a row's y part and a piece's x part, combined.
        const int nxc = xe - xs + 1;
        const bool split = len0 < nxc;
        auto stream = [&](int2 u0, int2 u1, int2 u2) {   // runs as (first slot, length)
            const int c1 = u0.y, c2 = c1 + u1.y, c3 = c2 + u2.y;
            const int b0 = u0.x, b1 = u1.x - c1, b2 = u2.x - c2;
            for (int f = 0; f < c3; f += SB) {
                int kk[SB];
                f3 P[SB];
#pragma unroll
                for (int j = 0; j < SB; j++) {
                    const int fj = f + j;
                    const int k = (fj < c1 ? b0 : (fj < c2 ? b1 : b2)) + fj;
                    kk[j] = fj < c3 ? k : t;
                    P[j] = pos(kk[j]);
                }
                room(SB);
#pragma unroll
                for (int j = 0; j < SB; j++) {
                    const bool keep = (kk[j] != t) & near(P[j]);
                    s_surv[cnt][threadIdx.x] = uint32_t(kk[j]);
                    cnt += keep ? 1 : 0;
                }
            }
        };
        const int pbits = a.planes;
        const uint32_t pm = pbits ? (1u << pbits) - 1u : 0u;
        auto part = [&](uint32_t v, int shift, int pshift) -> uint32_t {   // one coordinate's key bits
            return (((v >> 7) & pm) << (kCellBits + pshift)) | ((v & 127u) << shift);
        };
        const uint32_t xka = part(uint32_t(vx0), 0, 0), xkb = part(uint32_t(vx0 + len0), 0, 0);
        const uint32_t vy0 = uint32_t(ys + kGridOrigin), vz0 = uint32_t(zs + kGridOrigin);
        for (int dz = 0; dz < 3; dz++) {
            if (zs + dz > ze) break;
            int2 pa[3], pb[3];   // per row: pieces a and b as (first slot, length)
            if constexpr (BATCHED) {
                // a batched solver's row (scene, z, y) is one row-table entry;
                // piece a is its cells [xa, xa + len0), piece b (past the x
                // wrap) its cells [0, nxc - len0); a row outside the box
                // reads the empty entry
                const RowTable& T = a.rowtab;
                const int xa = vx0 & (kGridDim - 1);
                uint2 e[3];
#pragma unroll
                for (int dy = 0; dy < 3; dy++)
                    e[dy] = ys + dy <= ye ? T.rows[(sbase + uint32_t(cell_index(xs, ys + dy, zs + dz))) >> 7]
                                          : make_uint2(0u, 0u);
#pragma unroll
                for (int dy = 0; dy < 3; dy++) {
                    const int a0 = int(row_lb(T, e[dy], xa)), a1 = int(row_lb(T, e[dy], xa + len0));
                    pa[dy] = make_int2(a0, a1 - a0);
                    pb[dy] = make_int2(0, 0);
                    if (split) {
                        const int b0 = int(row_lb(T, e[dy], 0)), b1 = int(row_lb(T, e[dy], nxc - len0));
                        pb[dy] = make_int2(b0, b1 - b0);
                    }
                }
            } else {
            const uint32_t zk = part(vz0 + uint32_t(dz), 14, 2 * pbits);
#pragma unroll
            for (int dy = 0; dy < 3; dy++) {
                const bool in = ys + dy <= ye;
                const uint32_t yk = zk | part(vy0 + uint32_t(in ? dy : 0), 7, pbits);
                const uint32_t ka = yk | xka, kb = yk | xkb;
                int la = 0, ea = 0, lb = 0, eb = 0;
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    const bool inb = m >= len0;
                    const int2 rg = m < nxc ? range[size_t(inb ? kb + uint32_t(m - len0) : ka + uint32_t(m))] : make_int2(0, 0);
                    const int l = rg.y - rg.x;
                    if (inb) { lb += l; eb = max(eb, rg.y); } else { la += l; ea = max(ea, rg.y); }
                }
                pa[dy] = in ? make_int2(ea - la, la) : make_int2(0, 0);
                pb[dy] = in ? make_int2(eb - lb, lb) : make_int2(0, 0);
            }
            }
            const int2 z0 = make_int2(0, 0);
            stream(pa[0], split ? pb[0] : pa[1], split ? pa[1] : pa[2]);
            stream(split ? pb[1] : z0, split ? pa[2] : z0, split ? pb[2] : z0);
        }
    } else if constexpr (BATCHED) {
        // Rows (cz, cy) in the reference's order.  Every row's first x-run
        // starts at xs, at the same offset xa in every row; its slots are two
        // bounds behind the row's table entry.  Entries are requested two
        // rows ahead and bounds one row ahead, so neither wait falls on a
        // scan.
        const RowTable& T = a.rowtab;
        auto rowkey = [&](int cy, int cz) -> uint32_t { return sbase + uint32_t(cell_index(xs, cy, cz)); };
        auto advance = [&](int& cy, int& cz) {
            if (++cy > ye) { cy = ys; cz++; }
        };
        int cy = ys, cz = zs;
        int cy1 = cy, cz1 = cz;
        advance(cy1, cz1);
        int cy2 = cy1, cz2 = cz1;
        advance(cy2, cz2);
        const uint32_t k0 = rowkey(cy, cz);
        const int xa = int(k0 & 127u);
        const uint2 e0 = T.rows[k0 >> 7];
        int2 sp = make_int2(int(row_lb(T, e0, xa)), int(row_lb(T, e0, xa + len0)));
        uint2 e1 = cz1 <= ze ? T.rows[rowkey(cy1, cz1) >> 7] : make_uint2(0u, 0u);
        while (cz <= ze) {
            int2 sp1 = make_int2(0, 0);
            if (cz1 <= ze) sp1 = make_int2(int(row_lb(T, e1, xa)), int(row_lb(T, e1, xa + len0)));
            uint2 e2 = make_uint2(0u, 0u);
            if (cz2 <= ze) e2 = T.rows[rowkey(cy2, cz2) >> 7];
            scan(sp.x, sp.y);
            rest(cy, cz);
            sp = sp1;
            e1 = e2;
            cy = cy1; cz = cz1;
            cy1 = cy2; cz1 = cz2;
            advance(cy2, cz2);
        }
    } else {
        // Each lane walks its x-runs -- rows (cz, cy) in the reference's
        // order, each row's runs left to right -- as one stream of
        // candidates, in batches of SB that may take their last slots from
        // the next run, so a lane does not wait for the wave's longest run at
        // every row (about a third fewer batches on the dense pile).  The
        // candidate order, hence every sum, is unchanged.  The next run's
        // (first three) table rows are requested as soon as a lane enters a
        // run, so they arrive while it scans.
        int gx = xs, gy = ys, gz = zs;   // the next run to request
        int rc0 = 0, rlen = 0;           // the requested run: first cell, length (0: none left) ...
        uint32_t rbase = 0;              // ... table, and first three rows
        int2 q0 = make_int2(0, 0), q1 = q0, q2 = q0;
        auto request = [&]() {
            rlen = 0;
            if (gz > ze) return;
            const int vx = gx + kGridOrigin;
            rlen = gx == xs ? len0 : (vx < 0 ? 1 : min(xe - gx + 1, kGridDim - vx % kGridDim));
            rc0 = cell_index(gx, gy, gz);
            rbase = table(gx, gy, gz);
            q0 = raw(rbase, rc0);
            q1 = rlen > 1 ? raw(rbase, rc0 + 1) : make_int2(0, 0);
            q2 = rlen > 2 ? raw(rbase, rc0 + 2) : make_int2(0, 0);
            gx += rlen;
            if (gx > xe) {
                gx = xs;
                if (++gy > ye) { gy = ys; gz++; }
            }
        };
        // the requested run's slots: its first occupied cell's start to its
        // last occupied cell's end
        auto resolve = [&]() -> int2 {
            int s0 = 0, e0 = 0;
            bool any = false;
            for (int m = 0; m < rlen; m++) {
                const int2 rg = m == 0 ? q0 : (m == 1 ? q1 : (m == 2 ? q2 : raw(rbase, rc0 + m)));
                if (rg.y > rg.x) {
                    if (!any) s0 = rg.x;
                    e0 = rg.y;
                    any = true;
                }
            }
            return make_int2(s0, e0);
        };
        request();
        const int2 first = resolve();
        int cur = first.x, end = first.y;   // the run being scanned
        request();
        while (cur < end || rlen > 0) {
            const bool cross = cur + SB > end && rlen > 0;
            int2 nx = make_int2(0, 0);
            if (cross) nx = resolve();
            int kk[SB];
            bool in[SB];
            f3 P[SB];
#pragma unroll
            for (int j = 0; j < SB; j++) {
                const int idx = cur + j, k2 = nx.x + (idx - end);
                const bool in1 = idx < end, in2 = !in1 && cross && k2 < nx.y;
                kk[j] = in1 ? idx : (in2 ? k2 : t);
                in[j] = in1 || in2;
                P[j] = pos(kk[j]);
            }
            room(SB);
#pragma unroll
            for (int j = 0; j < SB; j++) {
                const bool keep = in[j] & (kk[j] != t) & near(P[j]);
                s_surv[cnt][threadIdx.x] = uint32_t(kk[j]);
                cnt += keep ? 1 : 0;
            }
            if (cur + SB <= end) {
                cur += SB;
            } else if (cross) {   // continue in the next run and request the one after
                cur = min(nx.x + (cur + SB - end), nx.y);
                end = nx.y;
                request();
            } else {
                cur = end;
            }
        }
    }
    // (slots and lanes share one 32-bit word: up to 2^26 particles)
    if (BLOCK == 64 && a.n <= (1 << 26)) {
        // Final drain over the whole wave.  The lanes' survivors, in the flat
        // order (m-th survivor of every lane, lanes ascending, m ascending),
        // are dealt to the 64 lanes a round at a time; each projection runs
        // with its owner's state (from LDS) and its 0 + pair delta returns to
        // the owner, which adds its own in visiting order (the owner's state
        // comes over by lane shuffles).  acc is never -0
        // (it starts at +0 and round-to-nearest sums are -0 only for two -0
        // terms), so adding +0 for a rejected candidate or 0 + d in place of
        // d leaves every sum as the lane's own loop would.
        __shared__ uint4 s_lvl[kSurv];   // per level m: ballot of cnt > m, first flat position
        const int lane = int(threadIdx.x);
        const uint64_t below = (1ull << lane) - 1ull;
        uint32_t* flat = &s_surv[0][0];
        // compaction in place: position(lane, m) <= 64 m + lane, the entry's
        // own index, and level m is read before it is written
        uint32_t total = 0;
        for (int m = 0;; m++) {
            const bool has = cnt > m;
            const uint64_t mk = __ballot(has);
            if (mk == 0) break;
            const uint32_t e = has ? s_surv[m][lane] : 0u;
            if (lane == 0) s_lvl[m] = make_uint4(uint32_t(mk), uint32_t(mk >> 32), total, 0u);
            if (has) flat[total + uint32_t(__popcll(mk & below))] = e | (uint32_t(lane) << 26);
            total += uint32_t(__popcll(mk));
        }
        __syncthreads();
        int mn = 0;   // this lane's next survivor to collect
        for (uint32_t base = 0; base < total; base += 64) {
            f3 pd = {0.f, 0.f, 0.f};
            const uint32_t q = base + uint32_t(lane);
            const uint32_t e = q < total ? flat[q] : 0u;
            const int ol = int(e >> 26), k = int(e & 0x3FFFFFFu);
            const f3 ox = {__shfl(x.x, ol), __shfl(x.y, ol), __shfl(x.z, ol)};
            const f3 ov = {__shfl(v.x, ol), __shfl(v.y, ol), __shfl(v.z, ol)};
            const float orad = __shfl(rad, ol), ow1 = __shfl(w1, ol);
            const int oprox = __shfl(proxy, ol);
            if (q < total) {
                const int ot = tb + ol;
                const float4 Pj = a.Xs[k], Vj = a.Vs[k];
                const int nf = a.Fs[k];
                const int oi = DIAG ? int(a.ids[ot]) : 0;
                pp_pair<DIAG>(a, ot, k, Pj, Vj, nf, oi, oprox, ox, ov, orad, ow1, orad + a.max_r + a.coh, pd);
            }
            while (true) {
                bool take = false;
                int src = lane;
                if (mn < cnt) {
                    const uint4 L = s_lvl[mn];
                    const uint64_t mk = uint64_t(L.x) | (uint64_t(L.y) << 32);
                    const uint32_t pos = L.z + uint32_t(__popcll(mk & below));
                    if (pos < base + 64u) { take = true; src = int(pos - base); }
                }
                if (!__any(take)) break;
                const float rx = __shfl(pd.x, src), ry = __shfl(pd.y, src), rz = __shfl(pd.z, src);
                if (take) {
                    acc = acc + f3{rx, ry, rz};
                    mn++;
                }
            }
        }
        cnt = 0;
    } else {
        drain();
    }
    if (!live) return;
    const f3 dw = {fma_(acc.x, w1, 0.f), fma_(acc.y, w1, 0.f), fma_(acc.z, w1, 0.f)};
    const f3 tot = {fma_(dw.x, a.relax, 0.f), fma_(dw.y, a.relax, 0.f), fma_(dw.z, a.relax, 0.f)};
    const float4 x0 = a.Ps[a.vals[t]];
    f3 dl = f3{0.f, 0.f, 0.f} + tot;
    if (a.plane_fold) {
        // the reference's delta buffer: 0 + the plane's delta (the shape
        // dispatch), then + this particle's pair sum
        f3 fd;
        if (plane_contact(x0, Pi, Vi, pf, a.altitude, a.margin, 0.5f * (a.pmu + a.mat_mu0), a.ka, a.dt, a.relax, fd)) {
            a.ppd[3 * t] = dl.x; a.ppd[3 * t + 1] = dl.y; a.ppd[3 * t + 2] = dl.z;
            dl = (f3{0.f, 0.f, 0.f} + fd) + tot;
        }
    }
    a.delta[3 * t] = dl.x; a.delta[3 * t + 1] = dl.y; a.delta[3 * t + 2] = dl.z;
    apply_delta(xyz(x0), t, x, rad, w1, dl, a.dt, a.vmax, a.Po, a.Vo, a.nk, pf);
}

// The shape half of the delta apply: runs after both k_shape_solve and
// k_pp_apply.  Each particle that received a shape delta is redone by one
// thread, its owner's (the largest contact index that added to it): delta =
// (shape sum) + (the particle-particle delta k_pp_apply stored, 0 + pp, which
// adds to any shape sum exactly as pp does, the sum never being -0), applied
// as k_pp_apply applies it.  The owner clears the particle's shape sum and
// mark for the next solve.  After a solve in the previous slot order (remap
// non-null) every contact is mapped to its particle's sorted slot here.
__global__ __launch_bounds__(256) void k_shape_apply(ShapeApplyArgs a) {
    if (a.rk != nullptr)   // the particle-pair solve, the last reader of these rows, has finished
        for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < a.n; t += gridDim.x * blockDim.x) {
            const uint32_t c = a.rk[t];
            if (t == 0 || a.rk[t - 1] != c) *reinterpret_cast<int2*>(a.range + size_t(2) * c) = make_int2(0, 0);
        }
    const int count = min(a.cmax, *a.cc);
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x) {
        const int pi = a.cpart[t];   // the slot dshape and owner use
        int ps = pi;                 // the particle's slot in the sorted order
        if (a.remap) {
            ps = int(a.remap[pi]);
            a.cpart[t] = ps;
        }
        if (a.owner[pi] != t) continue;
        a.owner[pi] = -1;
        float* ds = a.dshape + 3 * pi;
        float* dp = a.delta + 3 * ps;
        const float4 Pi = a.Xs[ps], Vi = a.Vs[ps];
        const float4 x0 = a.Ps[a.vals[ps]];
        f3 sh = f3{ds[0], ds[1], ds[2]};
        f3 pp = f3{dp[0], dp[1], dp[2]};
        if (a.plane_fold) {   // its plane contact was added by the particle-pair solve
            f3 fd;
            if (plane_contact(x0, Pi, Vi, a.flags[ps], a.altitude, a.margin, 0.5f * (a.pmu + a.mat_mu0), a.ka, a.dt,
                              a.relax, fd)) {
                sh = sh + fd;
                pp = f3{a.ppd[3 * ps], a.ppd[3 * ps + 1], a.ppd[3 * ps + 2]};
            }
        }
        const f3 dl = sh + pp;
        ds[0] = 0.0f; ds[1] = 0.0f; ds[2] = 0.0f;
        dp[0] = dl.x; dp[1] = dl.y; dp[2] = dl.z;
        apply_delta(xyz(x0), ps, xyz(Pi), Pi.w, Vi.w, dl, a.dt, a.vmax, a.Po, a.Vo, a.nk, a.flags[ps]);
    }
}

// apply_particle_shape_restitution.comp:48-116.  Body previous pose/velocity
// are the current ones (scene.cpp:570-574 binds the same buffers twice).
__global__ __launch_bounds__(256) void k_restitution(RestArgs a) {
    const int count = min(a.cmax, *a.cc);
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x) {
        int s = a.cshape[t];
        int body = a.body.shape_body[s];
        int pi = a.cpart[t];
        if ((a.flags[pi] & 1) == 0) continue;
        float4 Vn = a.V[pi];
        f3 vnew = xyz(Vn);
        const uint32_t po = a.old[pi];
        float4 X0 = a.P0[po];
        f3 px = xyz(X0);
        f3 vold = xyz(a.V0[po]);
        BodyFrame F = load_body(a.body, body);
        f3 cb = xyz(a.cbpos[t]);
        f3 bx = F.pos + quat_rotate_sc(F.rot, cb);
        f3 n = xyz(a.cnorm[t]);
        f3 d = px - bx;
        float c = fma_(n.z, d.z, fma_(n.x, d.x, n.y * d.y)) - X0.w;
        if (c > a.ka) continue;
        f3 bxp = F.pos + quat_rotate_sc(F.rot, cb);
        f3 r = bxp - (F.pos + quat_rotate_sc(F.rot, F.com));
        f3 bvc = quat_rotate_sc(F.rot, xyz(a.cbvel[t]));
        f3 bvo = bvc, bvn = bvc;
        if (body >= 0) {
            bvo = (F.v + cross3(F.w, r)) + bvc;
            bvn = (F.v + cross3(F.w, r)) + bvc;
        }
        f3 dvo = vold - bvo;
        float rvo = fma_(n.z, dvo.z, fma_(n.y, dvo.y, n.x * dvo.x));
        f3 dvn = vnew - bvn;
        float rvn = fma_(n.z, dvn.z, fma_(n.x, dvn.x, n.y * dvn.y));
        if (rvo < 0.0f) {
            float rt = gmax((-a.restitution) * rvo, 0.0f);
            float sc = (-rvn) + rt;
            float* vp = reinterpret_cast<float*>(a.V + pi);
            atomicAdd(vp + 0, n.x * sc);
            atomicAdd(vp + 1, n.y * sc);
            atomicAdd(vp + 2, n.z * sc);
        }
    }
}

__global__ __launch_bounds__(256) void k_fill_zero(int4* __restrict__ p, size_t n4) {
    for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n4; t += size_t(gridDim.x) * blockDim.x)
        p[t] = make_int4(0, 0, 0, 0);
}

__global__ void k_zero_int(int* p, int count) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) p[t] = 0;
}

// Zeroes the counters a frame's first substep uses: particle contacts cc[0],
// rigid contacts cc[4], mesh narrowphase queues cc[6], cc[7].
__global__ void k_frame_begin(int* cc, uint32_t* mcnt, int nmcnt) {
    if (threadIdx.x == 0) { cc[0] = 0; cc[4] = 0; cc[6] = 0; cc[7] = 0; }
    for (int k = threadIdx.x; k < nmcnt; k += blockDim.x) mcnt[k] = 0u;
}

// ------------------------------------------------------------- launchers --
static inline unsigned blocks_for(long long n, int bs) { return unsigned((n + bs - 1) / bs); }

void launch_integrate(const IntegrateArgs& a, hipStream_t s) {
    if (a.n > 0) hipLaunchKernelGGL(k_integrate, dim3(blocks_for(a.n, kIntTile)), dim3(256), 0, s, a);
}
void launch_plane_contacts(const PlaneArgs& a, hipStream_t s) {
    if (a.n > 0) hipLaunchKernelGGL(k_plane_contacts, dim3(blocks_for(a.n, 256)), dim3(256), 0, s, a);
}
void launch_cell_ranges(int n, const uint32_t* keys, int* range, uint32_t* rkeys, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_cell_ranges, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, keys, range, rkeys);
}
void launch_cell_ranges_gather(int n, const uint32_t* keys, const uint32_t* vals, int* range, uint32_t* rkeys,
                               const SortGather& g, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_cell_ranges_gather, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, keys, vals, range, rkeys, g);
}
// Contact particle indices through a slot map (previous slot -> new slot
// after a sort, or slot -> particle id when the layout is reset).
__global__ __launch_bounds__(256) void k_remap_contacts(int* __restrict__ cpart, const int* __restrict__ cc, int cmax,
                                                        const uint32_t* __restrict__ map) {
    const int count = min(cmax, *cc);
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x)
        cpart[t] = int(map[cpart[t]]);
}
// dst[ids[t]] = src[t].xyz: slot order -> particle-order float3 rows.
__global__ __launch_bounds__(256) void k_scatter_f3(float* __restrict__ dst, const float4* __restrict__ src,
                                                    const uint32_t* __restrict__ ids, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = ids[t];
    const float4 v = src[t];
    dst[3 * i] = v.x; dst[3 * i + 1] = v.y; dst[3 * i + 2] = v.z;
}
// dst[3 ids[t] + c] = src[3 t + c]: slot-order float triples -> particle order.
__global__ __launch_bounds__(256) void k_scatter3(float* __restrict__ dst, const float* __restrict__ src,
                                                  const uint32_t* __restrict__ ids, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = ids[t];
    dst[3 * size_t(i)] = src[3 * size_t(t)];
    dst[3 * size_t(i) + 1] = src[3 * size_t(t) + 1];
    dst[3 * size_t(i) + 2] = src[3 * size_t(t) + 2];
}
// dst[k] = ids[src[k]] for the contact list's slots (slots outside [0, n) kept).
__global__ __launch_bounds__(256) void k_slots_to_ids(int* __restrict__ dst, const int* __restrict__ src,
                                                      const uint32_t* __restrict__ ids, int count, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int s = src[k];
    dst[k] = (s >= 0 && s < n) ? int(ids[s]) : s;
}
// dst[3 k + c] = src[k].xyz: float4 rows -> packed float triples.
__global__ __launch_bounds__(256) void k_pack3(float* __restrict__ dst, const float4* __restrict__ src, int count) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const float4 v = src[k];
    dst[3 * size_t(k)] = v.x;
    dst[3 * size_t(k) + 1] = v.y;
    dst[3 * size_t(k) + 2] = v.z;
}
__global__ __launch_bounds__(256) void k_iota(uint32_t* __restrict__ p, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = uint32_t(t);
}
// dst[t].xyz = src[ids[t]] (w kept): particle-order float3 rows into slot order.
__global__ __launch_bounds__(256) void k_gather_f3(float4* __restrict__ dst, const float* __restrict__ src,
                                                   const uint32_t* __restrict__ ids, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = ids[t];
    const float4 o = dst[t];
    dst[t] = make_float4(src[3 * i], src[3 * i + 1], src[3 * i + 2], o.w);
}
void launch_remap_contacts(int* cpart, const int* cc, int cmax, const uint32_t* map, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_remap_contacts, dim3(grid), dim3(256), 0, s, cpart, cc, cmax, map);
}
void launch_scatter_f3(float* dst, const float4* src, const uint32_t* ids, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_scatter_f3, dim3(blocks_for(n, 256)), dim3(256), 0, s, dst, src, ids, n);
}
void launch_scatter3(float* dst, const float* src, const uint32_t* ids, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_scatter3, dim3(blocks_for(n, 256)), dim3(256), 0, s, dst, src, ids, n);
}
void launch_slots_to_ids(int* dst, const int* src, const uint32_t* ids, int count, int n, hipStream_t s) {
    if (count > 0) hipLaunchKernelGGL(k_slots_to_ids, dim3(blocks_for(count, 256)), dim3(256), 0, s, dst, src, ids, count, n);
}
void launch_pack3(float* dst, const float4* src, int count, hipStream_t s) {
    if (count > 0) hipLaunchKernelGGL(k_pack3, dim3(blocks_for(count, 256)), dim3(256), 0, s, dst, src, count);
}
void launch_iota(uint32_t* p, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_iota, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n);
}
void launch_gather_f3(float4* dst, const float* src, const uint32_t* ids, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_gather_f3, dim3(blocks_for(n, 256)), dim3(256), 0, s, dst, src, ids, n);
}
int group_tiles(int n) { return (n + kGroupTile - 1) / kGroupTile; }
void launch_row_table(int n, const uint32_t* keys, uint32_t* tilecnt, uint32_t* tileoff, uint2* rows, uint32_t* glb,
                      int* rslot, const uint32_t* vals, const SortGather* gather, hipStream_t s) {
    if (n <= 0) return;
    const int tiles = group_tiles(n);
    const SortGather g = gather ? *gather : SortGather{};
    hipLaunchKernelGGL(k_group_counts, dim3(tiles), dim3(256), 0, s, n, keys, tilecnt);
    hipLaunchKernelGGL(k_group_scan, dim3(1), dim3(1024), 0, s, tiles, tilecnt, tileoff);
    hipLaunchKernelGGL(k_group_build, dim3(tiles), dim3(256), 0, s, n, keys, tileoff, rows, glb, rslot, vals, g);
}
void launch_clear_rows(int n, const int* rslot, uint2* rows, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_clear_rows, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, rslot, rows);
}
void launch_clear_ranges(int n, const uint32_t* rkeys, int* range, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_clear_ranges, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, rkeys, range);
}
void launch_cell_tables(int n, const uint32_t* keys, int* starts, int* ends, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_cell_tables, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, keys, starts, ends);
}
void launch_shape_solve(const ShapeArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_shape_solve, dim3(grid), dim3(256), 0, s, a);
}
void launch_shape_apply(const ShapeApplyArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_shape_apply, dim3(grid), dim3(256), 0, s, a);
}
void launch_pp_apply(const PPArgs& a, bool diag, bool batched, int block, hipStream_t s) {
    if (a.n <= 0) return;
    const int bs = (block == 64 || block == 128) ? block : 256;
    const dim3 g(blocks_for(a.n, bs)), b(bs);
    auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, g, b, 0, s, a); };
    if (bs == 64) {
        if (diag) { if (batched) go(k_pp_apply<true, true, 64>); else go(k_pp_apply<true, false, 64>); }
        else { if (batched) go(k_pp_apply<false, true, 64>); else go(k_pp_apply<false, false, 64>); }
    } else if (bs == 128) {
        if (diag) { if (batched) go(k_pp_apply<true, true, 128>); else go(k_pp_apply<true, false, 128>); }
        else { if (batched) go(k_pp_apply<false, true, 128>); else go(k_pp_apply<false, false, 128>); }
    } else {
        if (diag) { if (batched) go(k_pp_apply<true, true, 256>); else go(k_pp_apply<true, false, 256>); }
        else { if (batched) go(k_pp_apply<false, true, 256>); else go(k_pp_apply<false, false, 256>); }
    }
}
void launch_restitution(const RestArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_restitution, dim3(grid), dim3(256), 0, s, a);
}
// count must be a multiple of 4 ints (16-byte stores)
void launch_fill_zero(int* p, size_t count, hipStream_t s) {
    const size_t n4 = count / 4;
    if (n4 == 0) return;
    const unsigned grid = unsigned(std::min<size_t>((n4 + 255) / 256, 4096));
    hipLaunchKernelGGL(k_fill_zero, dim3(grid), dim3(256), 0, s, reinterpret_cast<int4*>(p), n4);
}
void launch_frame_begin(int* cc, uint32_t* mcnt, int nmcnt, hipStream_t s) {
    hipLaunchKernelGGL(k_frame_begin, dim3(1), dim3(64), 0, s, cc, mcnt, nmcnt);
}
void launch_zero_int(int* p, int count, hipStream_t s) {
    if (count > 0) hipLaunchKernelGGL(k_zero_int, dim3(blocks_for(count, 256)), dim3(256), 0, s, p, count);
}

}  // namespace xpbd
