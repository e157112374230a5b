// k_bodies.hip -- rigid-body stages of the substep: integrate_bodies,
// solve_body_contact_positions and apply_body_deltas (shipped unwired by the
// reference: shaders/integrate_bodies.comp, solve_body_contact_positions.comp,
// apply_body_deltas.comp), plus the two contact generators the reference
// lacks: box corners against the ground plane (rigid contacts) and particles
// against oriented boxes (particle-shape contacts).
//
// Layout decisions where the reference's shaders disagree with each other:
//   * body deltas are 8 floats per body (linear at +0..2, angular at +4..6), the
//     layout the wired particle-shape solve writes (solve_particle_shape_contacts
//     .comp:77-81); the body kernels read the same layout;
//   * inertia tensors are symmetric, so row- and column-major reads agree;
//   * body com is 3 floats per body.
// Bodies integrate and apply in place: each body's state is read and written by
// exactly one thread.  apply re-zeroes its body's delta and contact weight for
// the next iteration.
#include "xpbd_kernels.h"

namespace xpbd {

// integrate_bodies.comp:62-71 (cross-product quaternion rotation)
__device__ __forceinline__ f3 qrot(float4 q, f3 v) {
    const f3 u = {q.x, q.y, q.z};
    return v + 2.0f * cross3(u, cross3(u, v) + v * q.w);
}
__device__ __forceinline__ f3 qrot_inv(float4 q, f3 v) { return qrot(make_float4(-q.x, -q.y, -q.z, q.w), v); }
__device__ __forceinline__ float4 qmul(float4 p, float4 q) {   // integrate_bodies.comp:55-60
    const f3 pu = {p.x, p.y, p.z}, qu = {q.x, q.y, q.z};
    const float w = p.w * q.w - dot3(pu, qu);
    const f3 xyz = p.w * qu + q.w * pu + cross3(pu, qu);
    return make_float4(xyz.x, xyz.y, xyz.z, w);
}
__device__ __forceinline__ float4 qnormalize(float4 q) {
    const float d = ((q.x * q.x + q.y * q.y) + q.z * q.z) + q.w * q.w;
    const float inv = 1.0f / sqrtf(d);
    return make_float4(q.x * inv, q.y * inv, q.z * inv, q.w * inv);
}
__device__ __forceinline__ f3 mcol(const float* m, f3 v) {   // column-major 3x3 * v
    return f3{m[0], m[1], m[2]} * v.x + f3{m[3], m[4], m[5]} * v.y + f3{m[6], m[7], m[8]} * v.z;
}
// solve_body_contact_positions.comp:58-67
__device__ __forceinline__ f3 tf_point(f3 pos, float4 q, f3 p) {
    const f3 u = {q.x, q.y, q.z};
    return p + 2.0f * cross3(u, cross3(u, p) + p * q.w) + pos;
}
__device__ __forceinline__ f3 qrot_inv_b(float4 q, f3 v) {
    const f3 nq = {-q.x, -q.y, -q.z};
    return v + 2.0f * cross3(nq, cross3(nq, v) + v * q.w);
}

// integrate_bodies.comp:73-130 over the dynamic bodies [b0, b1).  Force,
// torque and world index have no setter and are zero.
__global__ __launch_bounds__(64) void k_integrate_bodies(BodyStepArgs a) {
    const int b = a.b0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.b1) return;
    if (a.flags[b] & 2) return;   // KINEMATIC: pass through
    const f3 x0 = xyz(a.q[2 * b]), v0 = xyz(a.qd[2 * b]), w0 = xyz(a.qd[2 * b + 1]);
    const float4 r0 = a.q[2 * b + 1];
    const f3 com = {a.com[3 * b], a.com[3 * b + 1], a.com[3 * b + 2]};
    const float im = a.invm[b];
    const f3 f0 = {0.f, 0.f, 0.f}, t0 = {0.f, 0.f, 0.f};
    const f3 g = {a.gx, a.gy, a.gz};
    const f3 xc = x0 + qrot(r0, com);
    const float ms = im != 0.0f ? 1.0f : 0.0f;
    const f3 v1 = v0 + (f0 * im + g * ms) * a.dt;
    const f3 x1 = xc + v1 * a.dt;
    const f3 wb = qrot_inv(r0, w0);
    const f3 Iwb = mcol(a.I + 9 * b, wb);
    const f3 tb = qrot_inv(r0, t0) - cross3(wb, Iwb);
    const f3 dwb = mcol(a.invI + 9 * b, tb) * a.dt;
    f3 w1 = qrot(r0, wb + dwb);
    const float4 wq = qmul(make_float4(w1.x, w1.y, w1.z, 0.0f), r0);
    const float hdt = 0.5f * a.dt;
    const float4 r1 = qnormalize(make_float4(r0.x + wq.x * hdt, r0.y + wq.y * hdt, r0.z + wq.z * hdt, r0.w + wq.w * hdt));
    w1 = w1 * (1.0f - a.angular_damping * a.dt);
    const f3 x = x1 - qrot(r1, com);
    a.q[2 * b] = make_float4(x.x, x.y, x.z, 0.0f);
    a.q[2 * b + 1] = r1;
    a.qd[2 * b] = make_float4(v1.x, v1.y, v1.z, 0.0f);
    a.qd[2 * b + 1] = make_float4(w1.x, w1.y, w1.z, 0.0f);
}

// Rigid contacts: the 8 corners of each box vs the ground plane.  shape0 =
// box, shape1 = the plane (shape 0, body -1), normal from shape0 to shape1.
__global__ __launch_bounds__(64) void k_body_plane_contacts(BodyContactGenArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = t >> 3, c = t & 7;
    bool want = false;
    f3 lc = {0.f, 0.f, 0.f}, wc = {0.f, 0.f, 0.f};
    if (k < a.nbox) {
        const int b = a.b0 + k;
        const float4 h = a.half[b];
        lc = {(c & 1) ? h.x : -h.x, (c & 2) ? h.y : -h.y, (c & 4) ? h.z : -h.z};
        wc = tf_point(xyz(a.q[2 * b]), a.q[2 * b + 1], lc);
        want = wc.z - a.altitude < a.margin;
    }
    if (a.fixed) {   // deterministic mode: slot 8k + c, in the order the corners are enumerated
        if (t == 0) *a.rcc = a.rcmax;
        if (k < a.nbox && t < a.rcmax) {
            a.s0[t] = want ? a.shape0 + k : 0;
            a.s1[t] = 0;
            a.p0[t] = make_float4(lc.x, lc.y, lc.z, 0.f);
            a.p1[t] = make_float4(wc.x, wc.y, a.altitude, 0.f);
            a.n[t] = make_float4(0.f, 0.f, -1.f, 0.f);
        }
        return;
    }
    const int slot = wave_reserve(a.rcc, want);
    if (want && slot < a.rcmax) {
        a.s0[slot] = a.shape0 + k;
        a.s1[slot] = 0;
        a.p0[slot] = make_float4(lc.x, lc.y, lc.z, 0.f);
        a.p1[slot] = make_float4(wc.x, wc.y, a.altitude, 0.f);
        a.n[slot] = make_float4(0.f, 0.f, -1.f, 0.f);
    }
}

// Particle vs oriented box: closest point of the box to the particle centre.
__global__ __launch_bounds__(256) void k_particle_box_contacts(ParticleBoxArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    if ((a.flags[i] & 1) == 0) return;
    const float4 P = a.P[i];
    const f3 p = xyz(P);
    const float rad = P.w;
    for (int k = 0; k < a.nbox; k++) {
        const int b = a.b0 + k;
        const float4 h4 = a.half[b];
        const f3 h = {h4.x, h4.y, h4.z};
        const f3 pos = xyz(a.q[2 * b]);
        const float4 rot = a.q[2 * b + 1];
        const f3 rel = p - pos;
        const float reach = ((h.x + h.y) + h.z) + a.margin + rad;
        bool want = false;
        f3 cp = {0.f, 0.f, 0.f}, nw = {0.f, 0.f, 0.f};
        if (!(dot3(rel, rel) > reach * reach)) {
            const f3 l = qrot_inv(rot, rel);
            cp = {gmin(gmax(l.x, -h.x), h.x), gmin(gmax(l.y, -h.y), h.y), gmin(gmax(l.z, -h.z), h.z)};
            f3 nl = {0.f, 0.f, 0.f};
            float sd = 0.0f;
            bool ok = true;
            const bool inside = fabsf(l.x) < h.x && fabsf(l.y) < h.y && fabsf(l.z) < h.z;
            if (!inside) {
                const f3 d = l - cp;
                sd = len3(d);
                if (!(sd > 0.0f)) ok = false;
                else nl = d * (1.0f / sd);
            } else {
                const float gx = h.x - fabsf(l.x), gy = h.y - fabsf(l.y), gz = h.z - fabsf(l.z);
                const int ax = (gx <= gy && gx <= gz) ? 0 : (gy <= gz ? 1 : 2);
                const float sg = comp3(l, ax) < 0.0f ? -1.0f : 1.0f;
                nl = {ax == 0 ? sg : 0.0f, ax == 1 ? sg : 0.0f, ax == 2 ? sg : 0.0f};
                if (ax == 0) { cp.x = sg * h.x; sd = -gx; }
                else if (ax == 1) { cp.y = sg * h.y; sd = -gy; }
                else { cp.z = sg * h.z; sd = -gz; }
            }
            if (ok && sd < a.margin + rad) {
                want = true;
                nw = qrot(rot, nl);
            }
        }
        const int slot = wave_reserve(a.cc, want);
        if (want && slot < a.cmax) {
            a.cpart[slot] = i;
            a.cshape[slot] = a.shape0 + k;
            a.cbpos[slot] = make_float4(cp.x, cp.y, cp.z, 0.f);
            a.cbvel[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
            a.cnorm[slot] = make_float4(nw.x, nw.y, nw.z, 0.f);
        }
    }
}

struct MassProps { float invm; const float* invI; };
__device__ __forceinline__ float ccd(float err, float4 qa, float4 qb, float mia, float mib, const float* Ia,
                                     const float* Ib, f3 la, f3 lb, f3 aa, f3 ab, float relax, float dt) {
    float den = 0.0f;
    den += dot3(la, la) * mia;
    den += dot3(lb, lb) * mib;
    const f3 ra = qrot_inv_b(qa, aa), rb = qrot_inv_b(qb, ab);
    den += dot3(ra, mcol(Ia, ra));
    den += dot3(rb, mcol(Ib, rb));
    float dl = -err;
    if (den > 0.0f) dl /= dt * den;
    return dl * relax;
}

// solve_body_contact_positions.comp:97-279 (offsets and thicknesses zero).
__global__ __launch_bounds__(64) void k_body_contact_solve(BodySolveArgs a) {
    __shared__ float zero9[9];
    if (threadIdx.x < 9) zero9[threadIdx.x] = 0.0f;
    __syncthreads();
    const int count = min(a.rcmax, *a.rcc);
    for (int ci = blockIdx.x * blockDim.x + threadIdx.x; ci < count; ci += gridDim.x * blockDim.x) {
        const int sa = a.s0[ci], sb = a.s1[ci];
        if (a.rcd) a.rcd[2 * size_t(ci)] = make_float4(0.f, 0.f, 0.f, 0.f);   // no contribution unless stored
        if (sa == sb) continue;
        const int ba = sa >= 0 ? a.shape_body[sa] : -1, bb = sb >= 0 ? a.shape_body[sb] : -1;
        if (ba == bb) continue;
        f3 pa = {0.f, 0.f, 0.f}, pb = {0.f, 0.f, 0.f};
        float4 qa = make_float4(0.f, 0.f, 0.f, 1.f), qb = qa;
        if (ba >= 0) { pa = xyz(a.q[2 * ba]); qa = a.q[2 * ba + 1]; }
        if (bb >= 0) { pb = xyz(a.q[2 * bb]); qb = a.q[2 * bb + 1]; }
        const f3 p0 = xyz(a.p0[ci]), p1 = xyz(a.p1[ci]);
        const f3 bxa = tf_point(pa, qa, p0), bxb = tf_point(pb, qb, p1);
        const f3 nn = xyz(a.n[ci]);
        const float d = dot3(nn, bxb - bxa) - (0.0f + 0.0f);
        if (d >= 0.0f) continue;
        float mia = 0.f, mib = 0.f;
        const float* Ia = zero9;
        const float* Ib = zero9;
        f3 ca = {0.f, 0.f, 0.f}, cb = ca, wa = ca, wb = ca;
        const f3 oa = {0.f, 0.f, 0.f}, ob = oa;
        if (ba >= 0) { mia = a.invm[ba]; Ia = a.invI + 9 * ba; ca = {a.com[3 * ba], a.com[3 * ba + 1], a.com[3 * ba + 2]}; wa = xyz(a.qd[2 * ba + 1]); }
        if (bb >= 0) { mib = a.invm[bb]; Ib = a.invI + 9 * bb; cb = {a.com[3 * bb], a.com[3 * bb + 1], a.com[3 * bb + 2]}; wb = xyz(a.qd[2 * bb + 1]); }
        int nz = 0;
        float mu = 0.f, mut = 0.f, mur = 0.f;
        if (sa >= 0) { nz += 1; mu += a.mu[sa]; mut += a.mu_t[sa]; mur += a.mu_r[sa]; }
        if (sb >= 0) { nz += 1; mu += a.mu[sb]; mut += a.mu_t[sb]; mur += a.mu_r[sb]; }
        if (nz > 0) { mu /= float(nz); mut /= float(nz); mur /= float(nz); }
        const f3 ra = bxa - tf_point(pa, qa, ca), rb = bxb - tf_point(pb, qb, cb);
        const f3 anga = -cross3(ra, nn), angb = cross3(rb, nn);
        if (ba >= 0) atomicAdd(a.invw + ba, 1.0f);
        if (bb >= 0) atomicAdd(a.invw + bb, 1.0f);
        const float ln = ccd(d, qa, qb, mia, mib, Ia, Ib, -nn, nn, anga, angb, a.relax, a.dt);
        f3 lda = -nn * ln, ldb = nn * ln, ada = anga * ln, adb = angb * ln;
        if (mu > 0.0f) {
            const f3 sxa = tf_point(pa, qa, p0 + oa), sxb = tf_point(pb, qb, p1 + ob);
            const f3 dl = sxb - sxa;
            f3 fd = dl - dot3(nn, dl) * nn;
            const f3 fra = sxa - tf_point(pa, qa, ca), frb = sxb - tf_point(pb, qb, cb);
            f3 rv = {0.f, 0.f, 0.f};
            if (ba >= 0 && (a.flags[ba] & 2)) {
                const f3 va = xyz(a.qd[2 * ba]) + cross3(wa, fra);
                rv = rv - (va - dot3(nn, va) * nn);
            }
            if (bb >= 0 && (a.flags[bb] & 2)) {
                const f3 vb = xyz(a.qd[2 * bb]) + cross3(wb, frb);
                rv = rv + (vb - dot3(nn, vb) * nn);
            }
            fd = fd + rv * a.dt;
            const f3 perp = normalize3(fd);
            const f3 afa = -cross3(fra, perp), afb = cross3(frb, perp);
            const float err = len3(fd);
            if (err > 0.0f) {
                float lf = ccd(err, qa, qb, mia, mib, Ia, Ib, -perp, perp, afa, afb, a.relax, a.dt);
                lf = gmax(lf, -ln * mu);
                lda = lda - perp * lf; ldb = ldb + perp * lf;
                ada = ada + afa * lf; adb = adb + afb * lf;
            }
        }
        f3 dom = wb - wa;
        if (mut > 0.0f) {
            const float err = dot3(dom, nn) * a.dt;
            if (fabsf(err) > 0.0f) {
                float lt = ccd(err, qa, qb, mia, mib, Ia, Ib, f3{0.f, 0.f, 0.f}, f3{0.f, 0.f, 0.f}, -nn, nn, a.relax, a.dt);
                lt = gmin(gmax(lt, -ln * mut), ln * mut);
                ada = ada - nn * lt; adb = adb + nn * lt;
            }
        }
        if (mur > 0.0f) {
            dom = dom - dot3(nn, dom) * nn;
            const float err = len3(dom) * a.dt;
            if (err > 0.0f) {
                const f3 rn = normalize3(dom);
                float lr = ccd(err, qa, qb, mia, mib, Ia, Ib, f3{0.f, 0.f, 0.f}, f3{0.f, 0.f, 0.f}, -rn, rn, a.relax, a.dt);
                lr = gmax(lr, -ln * mur);
                ada = ada - rn * lr; adb = adb + rn * lr;
            }
        }
        if (a.rcd) {   // deterministic mode: the corner contacts only move their box (body b is the plane)
            if (ba >= 0) {
                a.rcd[2 * size_t(ci)] = make_float4(lda.x, lda.y, lda.z, 1.f);
                a.rcd[2 * size_t(ci) + 1] = make_float4(ada.x, ada.y, ada.z, 0.f);
            }
        } else if (ba >= 0) {
            float* d8 = a.delta + 8 * ba;
            atomicAdd(d8 + 0, lda.x); atomicAdd(d8 + 1, lda.y); atomicAdd(d8 + 2, lda.z);
            atomicAdd(d8 + 4, ada.x); atomicAdd(d8 + 5, ada.y); atomicAdd(d8 + 6, ada.z);
        }
        if (bb >= 0) {
            float* d8 = a.delta + 8 * bb;
            atomicAdd(d8 + 0, ldb.x); atomicAdd(d8 + 1, ldb.y); atomicAdd(d8 + 2, ldb.z);
            atomicAdd(d8 + 4, adb.x); atomicAdd(d8 + 5, adb.y); atomicAdd(d8 + 6, adb.z);
        }
        float* imp = a.impulse + 6 * ci;
        atomicAdd(imp + 0, lda.x); atomicAdd(imp + 1, lda.y); atomicAdd(imp + 2, lda.z);
        atomicAdd(imp + 3, ada.x); atomicAdd(imp + 4, ada.y); atomicAdd(imp + 5, ada.z);
    }
}

// apply_body_deltas.comp:58-119 in place; consumes and re-zeroes the body's
// delta and contact weight.
__global__ __launch_bounds__(64) void k_apply_body_deltas(BodyStepArgs a) {
    const int b = a.b0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.b1) return;
    float* d8 = a.delta + 8 * b;
    const f3 dp0 = {d8[0], d8[1], d8[2]}, dq = {d8[4], d8[5], d8[6]};
    const float iw = a.invw[b];
    d8[0] = 0.f; d8[1] = 0.f; d8[2] = 0.f; d8[4] = 0.f; d8[5] = 0.f; d8[6] = 0.f;
    a.invw[b] = 0.f;
    const float im = a.invm[b];
    if (im == 0.0f) return;
    const f3 p0 = xyz(a.q[2 * b]), v0 = xyz(a.qd[2 * b]), w0 = xyz(a.qd[2 * b + 1]);
    const float4 q0 = a.q[2 * b + 1];
    const f3 com = {a.com[3 * b], a.com[3 * b + 1], a.com[3 * b + 2]};
    float wgt = 1.0f;
    if (a.use_weights && iw > 0.0f) wgt = 1.0f / iw;
    const f3 dp = dp0 * (im * wgt);
    const f3 dqw = dq * wgt;
    const f3 wb = qrot_inv(q0, w0);
    const f3 dwb = mcol(a.invI + 9 * b, qrot_inv(q0, dqw));
    const f3 tb = cross3(dwb, mcol(a.I + 9 * b, wb + dwb)) + cross3(wb, mcol(a.I + 9 * b, dwb));
    const f3 dw1 = qrot(q0, dwb - a.dt * mcol(a.invI + 9 * b, tb));
    const f3 dwdt = dw1 * a.dt;
    const float4 wq = qmul(make_float4(dwdt.x, dwdt.y, dwdt.z, 0.0f), q0);
    const float4 q1 = qnormalize(make_float4(q0.x + wq.x * 0.5f, q0.y + wq.y * 0.5f, q0.z + wq.z * 0.5f, q0.w + wq.w * 0.5f));
    const f3 xc = p0 + qrot(q0, com);
    f3 p1 = xc + dp * a.dt;
    p1 = p1 - qrot(q1, com);
    f3 v1 = v0 + dp, w1 = w0 + dw1;
    if (len3(v1) < 1e-4f) v1 = {0.f, 0.f, 0.f};
    if (len3(w1) < 1e-4f) w1 = {0.f, 0.f, 0.f};
    a.q[2 * b] = make_float4(p1.x, p1.y, p1.z, 0.f);
    a.q[2 * b + 1] = q1;
    a.qd[2 * b] = make_float4(v1.x, v1.y, v1.z, 0.f);
    a.qd[2 * b + 1] = make_float4(w1.x, w1.y, w1.z, 0.f);
}

void launch_integrate_bodies(const BodyStepArgs& a, hipStream_t s) {
    const int nb = a.b1 - a.b0;
    if (nb > 0) hipLaunchKernelGGL(k_integrate_bodies, dim3((nb + 63) / 64), dim3(64), 0, s, a);
}
void launch_apply_body_deltas(const BodyStepArgs& a, hipStream_t s) {
    const int nb = a.b1 - a.b0;
    if (nb > 0) hipLaunchKernelGGL(k_apply_body_deltas, dim3((nb + 63) / 64), dim3(64), 0, s, a);
}
void launch_body_plane_contacts(const BodyContactGenArgs& a, hipStream_t s) {
    if (a.nbox > 0) hipLaunchKernelGGL(k_body_plane_contacts, dim3((8 * a.nbox + 63) / 64), dim3(64), 0, s, a);
}
void launch_particle_box_contacts(const ParticleBoxArgs& a, hipStream_t s) {
    if (a.n > 0 && a.nbox > 0)
        hipLaunchKernelGGL(k_particle_box_contacts, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
}
void launch_body_contact_solve(const BodySolveArgs& a, hipStream_t s) {
    const int grid = a.rcmax > 0 ? (a.rcmax + 63) / 64 : 1;
    hipLaunchKernelGGL(k_body_contact_solve, dim3(grid < 1024 ? grid : 1024), dim3(64), 0, s, a);
}

}  // namespace xpbd
