// xpbd_device.h -- device-side arithmetic shared by the gfx950 kernels.
//
// All kernels are compiled with -ffp-contract=off and correctly rounded fp32
// division/sqrt, so a plain a*b+c is two roundings and only the explicit
// __fmaf_rn calls fuse.  This reproduces the reference shaders' evaluation:
// GLSL fma()/`precise` sites become __fmaf_rn, everything else is evaluated
// left to right (solve_particle_particle_contacts.comp:81-105,
// generate_mesh_contacts.comp:29-75, apply_particle_deltas.comp:30-36).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xpbd {

// Block b of a grid of T blocks -> the tile it processes, chosen so that each
// XCD (workgroups are dealt to the eight XCDs round-robin) works through one
// contiguous range of tiles: writes that land in the same cache line (the
// sort's bin runs, the cell table rows, the id-ordered radix input) then meet
// in one L2 instead of leaving it as partial lines from several.  A bijection
// of [0, T) for every T.
__device__ __forceinline__ int xcd_tile(int b, int T) {
    const int q = T >> 3, r = T & 7, x = b & 7, j = b >> 3;
    return x * q + min(x, r) + j;
}


constexpr int kGridDim = 128;
constexpr int kNumCells = kGridDim * kGridDim * kGridDim;
constexpr int kGridOrigin = 1 << 20;
constexpr int kWave = 64;

struct f3 { float x, y, z; };
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 xyz(float4 a) { return f3{a.x, a.y, a.z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ float fma_(float a, float b, float c) { return __fmaf_rn(a, b, c); }
__device__ __forceinline__ float gmax(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float gmin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float len3(f3 a) { return sqrtf(dot3(a, a)); }

// integrate_particles.comp:44-57 for one particle with flags f: x0, v0 ->
// x1, v1 in place (w = inverse mass in V.w; the path's forces are zero, and
// f * w keeps the reference's arithmetic).  k_integrate derives the cell key
// from it and the last radix pass stores its result in cell order, so both
// evaluate this same code.
__device__ __forceinline__ void integrate_particle(float4& P, float4& V, int f, float dt, float vmax, float gx, float gy,
                                                   float gz) {
    if ((f & 1) == 0) return;
    const float w = V.w;
    const float gs = w > 0.0f ? 1.0f : 0.0f;
    const float fx = 0.0f * w, fy = 0.0f * w, fz = 0.0f * w;
    f3 v1 = {V.x + (fx + gx * gs) * dt, V.y + (fy + gy * gs) * dt, V.z + (fz + gz * gs) * dt};
    const float mag = len3(v1);
    if (mag > vmax) v1 = v1 * (vmax / mag);
    P = make_float4(P.x + v1.x * dt, P.y + v1.y * dt, P.z + v1.z * dt, P.w);
    V = make_float4(v1.x, v1.y, v1.z, V.w);
}
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ f3 normalize3(f3 a) {
    float inv = 1.0f / sqrtf(dot3(a, a));
    return a * inv;
}
__device__ __forceinline__ float comp3(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// Newton-refined sqrt / length / dot / division used by the contact shaders.
__device__ __forceinline__ float exact_sqrt(float a) {
    float d = sqrtf(a);
    float r = fma_(-d, d, a);
    return fma_(r, 0.5f / d, d);
}
__device__ __forceinline__ float exact_length(f3 a) {
    float yy = a.y * a.y;
    float xx_yy = fma_(a.x, a.x, yy);
    float sq = fma_(a.z, a.z, xx_yy);
    return exact_sqrt(sq);
}
__device__ __forceinline__ float exact_dot(f3 a, f3 b) {
    float y = a.y * b.y;
    float xy = fma_(a.x, b.x, y);
    return fma_(a.z, b.z, xy);
}
__device__ __forceinline__ f3 exact_div_by(f3 v, float d, float inv) {   // inv = 1.0f / d
    f3 q = v * inv;
    f3 rem = {fma_(-d, q.x, v.x), fma_(-d, q.y, v.y), fma_(-d, q.z, v.z)};
    return {fma_(rem.x, inv, q.x), fma_(rem.y, inv, q.y), fma_(rem.z, inv, q.z)};
}
__device__ __forceinline__ f3 exact_div(f3 v, float d) { return exact_div_by(v, d, 1.0f / d); }
// exact_sqrt together with 1.0f / result, for an exact_div by the root.  The
// correction's 0.5f / d0 equals 0.5f * (1.0f / d0) exactly: d0, the root of a
// float, lies in [2^-75, 2^64], so neither quotient is subnormal or overflows
// and halving is exact (0, inf and NaN give the same values too).  One
// correctly rounded division then serves the correction and, whenever the
// corrected root is d0 itself, the reciprocal of the root.
struct RootInv {
    float d, inv;
};
__device__ __forceinline__ RootInv exact_sqrt_inv(float a) {
    const float d0 = sqrtf(a);
    const float i0 = 1.0f / d0;
    const float r = fma_(-d0, d0, a);
    const float d = fma_(r, 0.5f * i0, d0);
    float inv = i0;
    if (d != d0) {   // a near-tie moved the root (about 2^-23 of roots): divide again
        float dd = d;
        asm volatile("" : "+v"(dd));   // keeps the division inside the rarely taken branch
        inv = 1.0f / dd;
    }
    return {d, inv};
}
__device__ __forceinline__ RootInv exact_length_inv(f3 a) {
    float yy = a.y * a.y;
    float xx_yy = fma_(a.x, a.x, yy);
    float sq = fma_(a.z, a.z, xx_yy);
    return exact_sqrt_inv(sq);
}
// Shorter sequences for the correctly rounded reciprocal and square root on
// gfx950, bit-identical to the compiler's 1.0f / d and sqrtf(a) (built with
// -fhip-fp32-correctly-rounded-divide-sqrt) inside the ranges named below:
// both were compared with them on every float of those ranges on an MI355X.
// The compiler's sequences add operand scaling for subnormal and extreme
// exponents and special-value fix-ups that are identities inside them.
//  * rcp_cr: v_rcp_f32 and one fma Newton step, for |d| in [2^-126, 2^126).
//  * sqrt_cr: v_sqrt_f32, then the neighbour whose fma residual brackets a
//    (the compiler's own correction), for a = +0 and a in [2^-96, +inf).
__device__ __forceinline__ float rcp_cr(float d) {
    const float r = __builtin_amdgcn_rcpf(d);
    return fma_(fma_(-d, r, 1.0f), r, r);
}
__device__ __forceinline__ bool rcp_cr_ok(float d) {
    const float m = fabsf(d);
    return m >= 0x1p-126f && m < 0x1p126f;
}
__device__ __forceinline__ float sqrt_cr(float a) {
    const float s = __builtin_amdgcn_sqrtf(a);
    const float sdn = __uint_as_float(__float_as_uint(s) - 1u);
    const float sup = __uint_as_float(__float_as_uint(s) + 1u);
    const float t = fma_(-sdn, s, a) <= 0.0f ? sdn : s;
    return fma_(-sup, s, a) > 0.0f ? sup : t;
}
__device__ __forceinline__ bool sqrt_cr_ok(float a) { return a == 0.0f || (a >= 0x1p-96f && a < __builtin_inff()); }
// exact_sqrt_inv with the sequences above where they apply (every a the
// contact solve meets but subnormal, infinite or NaN squared lengths); the
// same values bit for bit: for a = 0 the reciprocal of the root is +inf, as
// the division gives, and the correction then runs the same operations.
__device__ __forceinline__ RootInv exact_sqrt_inv_fast(float a) {
    if (__builtin_expect(!sqrt_cr_ok(a), 0)) return exact_sqrt_inv(a);
    const float d0 = sqrt_cr(a);
    const float i0 = d0 == 0.0f ? __builtin_inff() : rcp_cr(d0);
    const float r = fma_(-d0, d0, a);
    const float d = fma_(r, 0.5f * i0, d0);
    float inv = i0;
    if (d != d0) {
        float dd = d;
        asm volatile("" : "+v"(dd));
        inv = 1.0f / dd;
    }
    return {d, inv};
}
__device__ __forceinline__ RootInv exact_length_inv_fast(f3 a) {
    float yy = a.y * a.y;
    float xx_yy = fma_(a.x, a.x, yy);
    float sq = fma_(a.z, a.z, xx_yy);
    return exact_sqrt_inv_fast(sq);
}
// 1.0f / d, correctly rounded
__device__ __forceinline__ float recip_fast(float d) {
    if (__builtin_expect(!rcp_cr_ok(d), 0)) {
        float dd = d;
        asm volatile("" : "+v"(dd));
        return 1.0f / dd;
    }
    return rcp_cr(d);
}
__device__ __forceinline__ float exact_div1(float v, float d) {
    float inv = 1.0f / d;
    float q = v * inv;
    float rem = fma_(-d, q, v);
    return fma_(rem, inv, q);
}
// Markstein-refined reciprocal division of apply_particle_deltas.comp:30-36.
__device__ __forceinline__ f3 markstein_div(f3 a, float b) {
    float y = 1.0f / b;
    float t = fma_(-b, y, 1.0f);
    y = t * y + y;
    f3 q = a * y;
    f3 r = {fma_(-q.x, b, a.x), fma_(-q.y, b, a.y), fma_(-q.z, b, a.z)};
    return {fma_(r.x, y, q.x), fma_(r.y, y, q.y), fma_(r.z, y, q.z)};
}

// Quaternion (x,y,z,w) rotation as written in solve_particle_shape_contacts.comp:50-68.
__device__ __forceinline__ f3 quat_rotate_sc(float4 q, f3 x) {
    float c = 2.0f * q.w * q.w - 1.0f;
    float d = 2.0f * (q.x * x.x + q.y * x.y + q.z * x.z);
    return {x.x * c + q.x * d + (q.y * x.z - q.z * x.y) * q.w * 2.0f,
            x.y * c + q.y * d + (q.z * x.x - q.x * x.z) * q.w * 2.0f,
            x.z * c + q.z * d + (q.x * x.y - q.y * x.x) * q.w * 2.0f};
}
__device__ __forceinline__ f3 quat_rotate_inv_sc(float4 q, f3 x) {
    float c = 2.0f * q.w * q.w - 1.0f;
    float d = 2.0f * (q.x * x.x + q.y * x.y + q.z * x.z);
    return {x.x * c + q.x * d - (q.y * x.z - q.z * x.y) * q.w * 2.0f,
            x.y * c + q.y * d - (q.z * x.x - q.x * x.z) * q.w * 2.0f,
            x.z * c + q.z * d - (q.x * x.y - q.y * x.x) * q.w * 2.0f};
}

// Hash-grid cell of a position (grid_cell_indices.comp:17-25): floor first,
// then the 2^20 origin offset, clamp at 0, wrap to 128 per axis, z-major.
__device__ __forceinline__ uint32_t cell_of(float px, float py, float pz, float inv) {
    int gx = (int)floorf(px * inv) + kGridOrigin;
    int gy = (int)floorf(py * inv) + kGridOrigin;
    int gz = (int)floorf(pz * inv) + kGridOrigin;
    gx = gx < 0 ? 0 : gx;
    gy = gy < 0 ? 0 : gy;
    gz = gz < 0 ? 0 : gz;
    return (uint32_t)((gz % kGridDim) * (kGridDim * kGridDim) + (gy % kGridDim) * kGridDim + (gx % kGridDim));
}
// Virtual cell coordinates -> physical cell (solve_particle_particle_contacts.comp:66-79).
__device__ __forceinline__ int cell_index(int x, int y, int z) {
    x += kGridOrigin; y += kGridOrigin; z += kGridOrigin;
    x = x < 0 ? 0 : x; y = y < 0 ? 0 : y; z = z < 0 ? 0 : z;
    return (z % kGridDim) * (kGridDim * kGridDim) + (y % kGridDim) * kGridDim + (x % kGridDim);
}

// Wrap plane of a virtual cell: which 128-cell wrap of the hash grid it lies
// in, pb bits per axis (1 or 2), after the same origin offset and clamp as
// cell_index.  Particles that share a physical cell but lie in different
// wraps are at least 128 - 3 cells apart; when no accepted pair can be that
// long, such particles never interact, and a sort key of plane << 21 | cell
// keeps each wrap's particles in a run of their own (in particle order, as
// the plain key does).
__device__ __forceinline__ uint32_t wrap_plane(int x, int y, int z, int pb) {
    x += kGridOrigin; y += kGridOrigin; z += kGridOrigin;
    x = x < 0 ? 0 : x; y = y < 0 ? 0 : y; z = z < 0 ? 0 : z;
    const int m = (1 << pb) - 1;
    return (uint32_t((z >> 7) & m) << (2 * pb)) | (uint32_t((y >> 7) & m) << pb) | uint32_t((x >> 7) & m);
}
static_assert(kGridDim == 128, "wrap_plane shifts by log2(kGridDim)");
// Wrap plane of the cell holding a position (wrap_plane of its virtual cell).
__device__ __forceinline__ uint32_t plane_of(float px, float py, float pz, float inv, int pb) {
    return wrap_plane((int)floorf(px * inv), (int)floorf(py * inv), (int)floorf(pz * inv), pb);
}

// Cell ranges of a batched solver, without a table over every cell.  Keys
// are scene << 21 | cell, cell = (z * 128 + y) * 128 + x, so a row (scene, z,
// y) is key >> 7 and the sorted slots of any x-run of cells in one row are
// contiguous.  A group is 8 consecutive cells of a row (key >> 3; 16 per
// row).  The occupied groups are numbered in sorted order; glb holds 8 lower
// bounds per group, glb[8g + j] = the first sorted position whose key is at
// least group * 8 + j, plus one sentinel group whose first word is the
// particle count.  A dense row table (8 bytes per row of every scene) gives
// each row's first occupied group and a 16-bit mask of its occupied groups.
// The first position at or past cell x of a row (x = 0..128) is then the
// group's bound when x's group is occupied, else the first bound of the
// row's next occupied group -- which, groups being numbered in sorted order,
// is also right past the row's last group (the next row's first group, or
// the sentinel).  No probing, no branches.
struct RowTable {
    const uint2* rows;   // per row: (first group, occupied-group mask); (0, 0) when empty
    const uint32_t* glb;
};
__device__ __forceinline__ uint32_t row_lb(const RowTable& T, uint2 e, int x) {
    const int gx = x >> 3;
    const uint32_t below = e.y & ((1u << gx) - 1u);
    const uint32_t g = e.x + uint32_t(__popc(below));
    const int j = ((e.y >> gx) & 1u) ? (x & 7) : 0;
    return T.glb[8 * size_t(g) + size_t(j)];
}
// The slots of the cells [key0, key0 + len) of one row.
__device__ __forceinline__ int2 row_span(const RowTable& T, uint32_t key0, int len) {
    const uint2 e = T.rows[key0 >> 7];
    const int x = int(key0 & 127u);
    return make_int2(int(row_lb(T, e, x)), int(row_lb(T, e, x + len)));
}
// The slots of one cell, (s, s) when it is empty.
__device__ __forceinline__ int2 row_cell(const RowTable& T, uint32_t key) { return row_span(T, key, 1); }

// Wave-aggregated slot reservation on a global counter: one atomic per wave
// for all active lanes with want == true.  Returns the lane's slot (or -1).
__device__ __forceinline__ int wave_reserve(int* counter, bool want) {
    uint64_t mask = __ballot(want);
    if (!want) return -1;
    int lane = __lane_id();
    int leader = __ffsll((unsigned long long)mask) - 1;
    int total = __popcll(mask);
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, total);
    base = __shfl(base, leader);
    uint64_t below = mask & ((1ull << lane) - 1ull);
    return base + __popcll(below);
}

// Inclusive prefix sum over the 64 lanes of a wave (all lanes active), on
// DPP lane moves: row_shr 1, 2, 4 and 8 scan each row of 16 lanes (a lane
// whose source lies before its row start takes the 0 passed as old), then
// row_bcast:15 adds row 0's total to row 1 and row 2's to row 3, and
// row_bcast:31 adds the total of rows 0-1 to rows 2 and 3.
__device__ __forceinline__ int wave_incl_scan_i(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, false);   // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, false);   // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false);   // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, false);   // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false);   // row_bcast:15 into rows 1, 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false);   // row_bcast:31 into rows 2, 3
    return x;
}

// Block-wide slot reservation: every thread of the block calls this once with
// the number of slots it needs; it returns the thread's first slot.  One
// global atomic per block (none if the block needs nothing): same-address
// atomics serialise across the chip, so aggregating them is what keeps
// contact emission from throttling the kernels that emit.
// s_tmp must hold blockDim.x / 64 + 1 ints.
__device__ __forceinline__ int block_reserve(int* counter, int count, int* s_tmp) {
    const int lane = __lane_id(), w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int incl = wave_incl_scan_i(count);
    if (lane == 63) s_tmp[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int k = 0; k < nw; k++) { const int c = s_tmp[k]; s_tmp[k] = tot; tot += c; }
        s_tmp[nw] = tot > 0 ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    const int base = s_tmp[nw] + s_tmp[w] + (incl - count);
    __syncthreads();
    return base;
}

// Block-wide exclusive prefix sum of one int per thread, in thread order.
// Returns the thread's offset; *total receives the block sum.  s_tmp must
// hold blockDim.x / 64 + 1 ints.
__device__ __forceinline__ int block_exclusive_scan(int count, int* s_tmp, int* total) {
    const int lane = __lane_id(), w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int incl = wave_incl_scan_i(count);
    if (lane == 63) s_tmp[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int k = 0; k < nw; k++) { const int c = s_tmp[k]; s_tmp[k] = tot; tot += c; }
        s_tmp[nw] = tot;
    }
    __syncthreads();
    const int base = s_tmp[w] + (incl - count);
    *total = s_tmp[nw];
    __syncthreads();
    return base;
}

}  // namespace xpbd
