// k_det.hip -- the deterministic mode's segmented reductions.
//
// The reference accumulates shape deltas into particles and bodies, and
// restitution into velocities, with float atomics
// (solve_particle_shape_contacts.comp:189-199, apply_body_deltas.comp:85-86,
// apply_particle_shape_restitution.comp:112-114), so wherever several
// contributions meet the result depends on the order the GPU happens to run
// them in.  In deterministic mode each contribution is written to its own
// slot instead, the contacts are ordered by a canonical key with the radix
// sort, and each sum is taken in that order:
//   * a particle's shape deltas (and restitution updates) in ascending shape
//     index: the plane, the meshes, then the boxes, the order in which the
//     reference's contact generators emit one particle's contacts;
//   * a box body's deltas over its particle contacts in ascending particle
//     id, then over its corner contacts with the plane in corner order.
// That is the order of a sequential pass over the contacts as the generators
// emit them, so results no longer depend on scheduling, and they equal the
// oracle's sequential sums bit for bit.
#include "xpbd_kernels.h"

namespace xpbd {

// Canonical keys of the contact list (slots of the current sorted order):
// key1 = slot << sbits | shape (a particle's contacts, by shape) and, for
// contacts with a box, key2 = box << idbits | particle id (a box's contacts,
// by particle id).  Entries past the count carry the padding key, which
// sorts last.  On the first iteration of a substep the contacts still name
// slots of the previous order and are mapped through remap here.
__global__ __launch_bounds__(256) void k_det_keys(DetKeyArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.cmax) return;
    const int count = min(a.cmax, *a.cc);
    uint32_t k1 = a.pad1, k2 = a.pad2;
    if (t < count) {
        int slot = a.cpart[t];
        if (a.remap) {
            slot = int(a.remap[slot]);
            a.cpart[t] = slot;
        }
        const int s = a.cshape[t];
        k1 = (uint32_t(slot) << a.sbits) | uint32_t(s);
        if (s >= a.box_shape0 && s < a.box_shape0 + a.nbox)
            k2 = (uint32_t(s - a.box_shape0) << a.idbits) | a.ids[slot];
    }
    a.key1[t] = k1;
    a.val1[t] = uint32_t(t);
    if (a.key2) {
        a.key2[t] = k2;
        a.val2[t] = uint32_t(t);
    }
}

// Per box: the range [start, end) of its contacts in the key2 order.
__global__ __launch_bounds__(256) void k_det_box_ranges(int cmax, const uint32_t* __restrict__ key2s, uint32_t pad,
                                                        int idbits, int* __restrict__ seg) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= cmax) return;
    const uint32_t k = key2s[u];
    if (k == pad) return;
    const uint32_t b = k >> idbits;
    if (u == 0 || (key2s[u - 1] >> idbits) != b) seg[2 * b] = u;
    if (u == cmax - 1 || key2s[u + 1] == pad || (key2s[u + 1] >> idbits) != b) seg[2 * b + 1] = u + 1;
}

// k_shape_apply of deterministic mode: the first contact of each particle in
// key1 order sums the particle's shape deltas in shape order, adds the
// particle-particle delta k_pp_apply stored (0 + pp, so the sum is the
// reference's), and applies the particle's delta again.
__global__ __launch_bounds__(256) void k_det_shape_apply(DetApplyArgs a) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= a.cmax) return;
    const uint32_t k = a.key1s[u];
    if (k == a.pad1) return;
    const uint32_t slot = k >> a.sbits;
    if (u > 0 && (a.key1s[u - 1] >> a.sbits) == slot) return;
    f3 sum = {0.f, 0.f, 0.f};
    for (int w = u; w < a.cmax; w++) {
        const uint32_t kw = a.key1s[w];
        if (kw == a.pad1 || (kw >> a.sbits) != slot) break;
        const float4 c = a.cfd[a.val1s[w]];
        if (c.w != 0.0f) sum = sum + xyz(c);
    }
    float* dp = a.delta + 3 * size_t(slot);
    const f3 dl = sum + f3{dp[0], dp[1], dp[2]};
    dp[0] = dl.x; dp[1] = dl.y; dp[2] = dl.z;
    const float4 Pi = a.Xs[slot], Vi = a.Vs[slot];
    const f3 x0 = xyz(a.Ps[a.vals[slot]]);
    f3 xn = xyz(Pi) + dl;
    f3 vnew = markstein_div(xn - x0, a.dt);
    const float mag = len3(vnew);
    if (mag > a.vmax) {
        vnew = vnew * (a.vmax / mag);
        xn = x0 + vnew * a.dt;
    }
    a.Po[slot] = make_float4(xn.x, xn.y, xn.z, Pi.w);
    a.Vo[slot] = make_float4(vnew.x, vnew.y, vnew.z, Vi.w);
}

// A box body's delta: its particle contacts' contributions in particle-id
// order, then its corner contacts' in corner order, summed from zero (the
// reference zeroes the body deltas before each iteration's solves).
__global__ __launch_bounds__(64) void k_det_body_reduce(DetBodyArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nbox) return;
    f3 lin = {0.f, 0.f, 0.f}, ang = {0.f, 0.f, 0.f};
    if (a.seg) {
        const int s0 = a.seg[2 * k], e0 = a.seg[2 * k + 1];
        for (int u = s0; u < e0; u++) {
            const uint32_t t = a.val2s[u];
            const float4 l = a.cbd[2 * size_t(t)], g = a.cbd[2 * size_t(t) + 1];
            if (l.w != 0.0f) {
                lin = lin + xyz(l);
                ang = ang + xyz(g);
            }
        }
    }
    if (a.rcd) {
        for (int c = 0; c < 8; c++) {
            const float4 l = a.rcd[2 * (8 * k + c)], g = a.rcd[2 * (8 * k + c) + 1];
            if (l.w != 0.0f) {
                lin = lin + xyz(l);
                ang = ang + xyz(g);
            }
        }
    }
    float* d8 = a.delta + 8 * size_t(a.b0 + k);
    d8[0] = lin.x; d8[1] = lin.y; d8[2] = lin.z;
    d8[4] = ang.x; d8[5] = ang.y; d8[6] = ang.z;
}

// apply_particle_shape_restitution.comp:48-116 in deterministic mode: the
// first contact of each particle in key1 order runs the particle's contacts
// in shape order, each reading the velocity the previous one left (the
// reference's atomics on the velocity it also reads, run in sequence).
__global__ __launch_bounds__(256) void k_det_restitution(RestArgs r, DetApplyArgs a) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= a.cmax) return;
    const uint32_t k = a.key1s[u];
    if (k == a.pad1) return;
    const uint32_t slot = k >> a.sbits;
    if (u > 0 && (a.key1s[u - 1] >> a.sbits) == slot) return;
    const int pi = int(slot);
    if ((r.flags[pi] & 1) == 0) return;
    f3 vcur = xyz(r.V[pi]);
    const uint32_t po = r.old[pi];
    const float4 X0 = r.P0[po];
    const f3 px = xyz(X0);
    const f3 vold = xyz(r.V0[po]);
    bool changed = false;
    for (int w = u; w < a.cmax; w++) {
        const uint32_t kw = a.key1s[w];
        if (kw == a.pad1 || (kw >> a.sbits) != slot) break;
        const int t = int(a.val1s[w]);
        const int s = r.cshape[t];
        const int body = r.body.shape_body[s];
        BodyFrame F = load_body(r.body, body);
        const f3 cb = xyz(r.cbpos[t]);
        const f3 bx = F.pos + quat_rotate_sc(F.rot, cb);
        const f3 n = xyz(r.cnorm[t]);
        const f3 d = px - bx;
        const float c = fma_(n.z, d.z, fma_(n.x, d.x, n.y * d.y)) - X0.w;
        if (c > r.ka) continue;
        const f3 bxp = F.pos + quat_rotate_sc(F.rot, cb);
        const f3 rr = bxp - (F.pos + quat_rotate_sc(F.rot, F.com));
        const f3 bvc = quat_rotate_sc(F.rot, xyz(r.cbvel[t]));
        f3 bvo = bvc, bvn = bvc;
        if (body >= 0) {
            bvo = (F.v + cross3(F.w, rr)) + bvc;
            bvn = (F.v + cross3(F.w, rr)) + bvc;
        }
        const f3 dvo = vold - bvo;
        const float rvo = fma_(n.z, dvo.z, fma_(n.y, dvo.y, n.x * dvo.x));
        const f3 dvn = vcur - bvn;
        const float rvn = fma_(n.z, dvn.z, fma_(n.x, dvn.x, n.y * dvn.y));
        if (rvo < 0.0f) {
            const float rt = gmax((-r.restitution) * rvo, 0.0f);
            const float sc = (-rvn) + rt;
            vcur = {vcur.x + n.x * sc, vcur.y + n.y * sc, vcur.z + n.z * sc};
            changed = true;
        }
    }
    if (changed) {
        const float w = r.V[pi].w;
        r.V[pi] = make_float4(vcur.x, vcur.y, vcur.z, w);
    }
}

static inline unsigned det_blocks(long long n, int bs) { return unsigned((n + bs - 1) / bs); }

void launch_det_keys(const DetKeyArgs& a, hipStream_t s) {
    if (a.cmax > 0) hipLaunchKernelGGL(k_det_keys, dim3(det_blocks(a.cmax, 256)), dim3(256), 0, s, a);
}
void launch_det_box_ranges(int cmax, const uint32_t* key2s, uint32_t pad, int idbits, int* seg, hipStream_t s) {
    if (cmax > 0)
        hipLaunchKernelGGL(k_det_box_ranges, dim3(det_blocks(cmax, 256)), dim3(256), 0, s, cmax, key2s, pad, idbits, seg);
}
void launch_det_shape_apply(const DetApplyArgs& a, hipStream_t s) {
    if (a.cmax > 0) hipLaunchKernelGGL(k_det_shape_apply, dim3(det_blocks(a.cmax, 256)), dim3(256), 0, s, a);
}
void launch_det_body_reduce(const DetBodyArgs& a, hipStream_t s) {
    if (a.nbox > 0) hipLaunchKernelGGL(k_det_body_reduce, dim3(det_blocks(a.nbox, 64)), dim3(64), 0, s, a);
}
void launch_det_restitution(const RestArgs& r, const DetApplyArgs& a, hipStream_t s) {
    if (a.cmax > 0) hipLaunchKernelGGL(k_det_restitution, dim3(det_blocks(a.cmax, 256)), dim3(256), 0, s, r, a);
}

}  // namespace xpbd
