// k_mesh.hip -- particle-vs-triangle-mesh narrowphase over the packed LBVH,
// the replacement of shaders/generate_mesh_contacts.comp:181-414.
//
// Per particle and mesh: depth-first closest-triangle search with the DFS
// stack in LDS (one 32-deep column per thread, bank = thread), identical
// pruning, push order and tie-breaking to the reference so the chosen face is
// the same.  For watertight meshes the inside/outside sign is the parity of
// ray/triangle hits along the reference's fixed ray; the reference tests every
// triangle of the mesh (:372-388), here the same watertight ray test runs only
// on triangles whose LBVH boxes the ray can reach.  Box tests are padded far
// beyond the test's rounding error, so no triangle the test would report is
// culled and the hit count, hence the sign, is unchanged.
#include "xpbd_kernels.h"

namespace xpbd {

constexpr int kMeshBlock = 128;
constexpr int kStack = 32;

__device__ __forceinline__ f3 ldv(const float* v, uint32_t k) { return {v[3 * k], v[3 * k + 1], v[3 * k + 2]}; }

__device__ __forceinline__ float dist_sq_aabb(f3 p, f3 lo, f3 hi) {
    f3 d = {gmax(gmax(lo.x - p.x, 0.0f), p.x - hi.x), gmax(gmax(lo.y - p.y, 0.0f), p.y - hi.y),
            gmax(gmax(lo.z - p.z, 0.0f), p.z - hi.z)};
    return dot3(d, d);
}
__device__ __forceinline__ float exact_dot3(f3 u, f3 v) {
    float py = u.y * v.y;
    float pxy = fma_(u.x, v.x, py);
    return fma_(u.z, v.z, pxy);
}
__device__ __forceinline__ float diff_product(float a, float b, float c, float d) {
    float cd = c * d;
    float diff = fma_(a, b, -cd);
    float err = fma_(-c, d, cd);
    return diff + err;
}

// Watertight ray/triangle test of generate_mesh_contacts.comp:77-113 with the
// per-ray constants (kx, ky, kz, Sx, Sy, Sz) precomputed.
struct Ray { f3 p; int kx, ky, kz; float Sx, Sy, Sz; };
__device__ __forceinline__ bool ray_tri(const Ray& r, f3 a, f3 b, f3 c) {
    f3 A = a - r.p, B = b - r.p, C = c - r.p;
    float Akz = comp3(A, r.kz), Bkz = comp3(B, r.kz), Ckz = comp3(C, r.kz);
    float Ax = comp3(A, r.kx) - r.Sx * Akz, Ay = comp3(A, r.ky) - r.Sy * Akz;
    float Bx = comp3(B, r.kx) - r.Sx * Bkz, By = comp3(B, r.ky) - r.Sy * Bkz;
    float Cx = comp3(C, r.kx) - r.Sx * Ckz, Cy = comp3(C, r.ky) - r.Sy * Ckz;
    float U = diff_product(Cx, By, Cy, Bx);
    float V = diff_product(Ax, Cy, Ay, Cx);
    float W = diff_product(Bx, Ay, By, Ax);
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    if ((T < 0.0f && det > 0.0f) || (T > 0.0f && det < 0.0f)) return false;
    return true;
}

// Conservative ray/box reachability: the box is grown by a pad of 3e-5 times
// the coordinate scale (the watertight test's rounding error is a few ulps of
// that scale), and the ray is accepted if it meets the grown box at t >= 0 or
// starts inside it.
__device__ __forceinline__ bool ray_box(f3 p, f3 dir, f3 lo, f3 hi) {
    float sc = fmaxf(fmaxf(fmaxf(fabsf(p.x), fabsf(p.y)), fabsf(p.z)),
                     fmaxf(fmaxf(fmaxf(fabsf(lo.x), fabsf(lo.y)), fabsf(lo.z)),
                           fmaxf(fmaxf(fabsf(hi.x), fabsf(hi.y)), fabsf(hi.z))));
    float pad = 3e-5f * (1.0f + sc);
    lo = {lo.x - pad, lo.y - pad, lo.z - pad};
    hi = {hi.x + pad, hi.y + pad, hi.z + pad};
    float tmin = 0.0f, tmax = 3.0e38f;
    float pc[3] = {p.x, p.y, p.z}, dc[3] = {dir.x, dir.y, dir.z};
    float lc[3] = {lo.x, lo.y, lo.z}, hc[3] = {hi.x, hi.y, hi.z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (dc[k] == 0.0f) {
            if (pc[k] < lc[k] || pc[k] > hc[k]) return false;
        } else {
            float t0 = (lc[k] - pc[k]) / dc[k], t1 = (hc[k] - pc[k]) / dc[k];
            if (t0 > t1) { float tt = t0; t0 = t1; t1 = tt; }
            tmin = fmaxf(tmin, t0);
            tmax = fminf(tmax, t1);
        }
    }
    return tmin <= tmax;
}

// closest_point_triangle of generate_mesh_contacts.comp:115-179.
__device__ void closest_point_triangle(f3 p, f3 a, f3 b, f3 c, f3& cp, f3& tn, f3& bary) {
    f3 ab = b - a, ac = c - a, ap = p - a;
    tn = normalize3(cross3(ab, ac));
    float u, v;
    float d1 = exact_dot3(ab, ap), d2 = exact_dot3(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) { v = 0.0f; u = 1.0f; }
    else {
        f3 bp = p - b;
        float d3 = exact_dot3(ab, bp), d4 = exact_dot3(ac, bp);
        if (d3 >= 0.0f && d4 <= d3) { v = 1.0f; u = 0.0f; }
        else {
            float vc = fma_(d1, d4, -(d2 * d3));
            if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) { v = exact_div1(d1, d1 - d3); u = 1.0f - v; }
            else {
                f3 cpp = p - c;
                float d5 = exact_dot3(ab, cpp), d6 = exact_dot3(ac, cpp);
                if (d6 >= 0.0f && d5 <= d6) { v = 0.0f; u = 0.0f; }
                else {
                    float vb = fma_(d5, d2, -(d1 * d6));
                    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
                        float wv = exact_div1(d2, d2 - d6);
                        v = 0.0f; u = 1.0f - wv;
                    } else {
                        float va = fma_(d3, d6, -(d5 * d4));
                        if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
                            float wv = exact_div1(d4 - d3, (d4 - d3) + (d5 - d6));
                            v = 1.0f - wv;
                            u = (1.0f - v) - wv;
                        } else {
                            float s_ab = va + vb;
                            float s_all = vc + s_ab;
                            float r0 = 1.0f / s_all;
                            float rem = fma_(-s_all, r0, 1.0f);
                            float den = fma_(rem, r0, r0);
                            v = vb * den;
                            u = fma_(-vc, den, 1.0f - v);
                        }
                    }
                }
            }
        }
    }
    float w = (1.0f - u) - v;
    bary = {u, v, w};
    cp = {(u * a.x + v * b.x) + w * c.x, (u * a.y + v * b.y) + w * c.y, (u * a.z + v * b.z) + w * c.z};
}

__global__ __launch_bounds__(kMeshBlock) void k_mesh_contacts(MeshArgs a) {
    __shared__ uint32_t stk[kStack][kMeshBlock];
    const int tid = blockIdx.x * kMeshBlock + threadIdx.x;
    const int lt = threadIdx.x;
    if (tid >= a.n) return;
    if ((a.flags[tid] & 1) == 0) return;
    const float4 P = a.P[tid];
    const f3 p = xyz(P);
    const float rad = P.w;
    const float eps = a.eps, eps2 = eps * eps;
    const float PI = 3.14159265359f;
    for (int ci = 0; ci < a.nmesh; ci++) {
        const MeshInfo mi = a.info[ci];
        const float max_dist = (a.margin + mi.margin) + rad;
        const bool wt = mi.watertight != 0u;
        int sp = 0;
        stk[sp++][lt] = mi.root;
        float min_dist = max_dist;
        f3 acc_n = {0.f, 0.f, 0.f};
        uint32_t min_face = 0;
        float sv = 0.f, sw = 0.f;
        bool found = false;
        float min_dist_sq = max_dist * max_dist;
        while (sp > 0) {
            const uint32_t ni = stk[--sp][lt];
            const float4 L = a.nl[ni], U = a.nu[ni];
            const float nds = dist_sq_aabb(p, xyz(L), xyz(U));
            if (wt) { if (nds > min_dist_sq) continue; }
            else if (nds > (min_dist + eps) * (min_dist + eps)) continue;
            const uint32_t lp = __float_as_uint(L.w);
            const uint32_t li = lp & 0x3FFFFFFFu, leaf = (lp >> 30) & 3u, ri = __float_as_uint(U.w) & 0x3FFFFFFFu;
            if (leaf) {
                for (uint32_t pc = li; pc < ri; pc++) {
                    const uint32_t t = a.prim[pc];
                    const f3 va = ldv(a.verts, a.tris[3 * t]), vb = ldv(a.verts, a.tris[3 * t + 1]),
                             vc = ldv(a.verts, a.tris[3 * t + 2]);
                    const f3 e0 = vb - va, e1 = vc - va, e2 = vc - vb;
                    const f3 nrm = cross3(e0, e1);
                    const float e0n = dot3(e0, e0), e1n = dot3(e1, e1), e2n = dot3(e2, e2);
                    if (len3(nrm) / ((e0n + e1n) + e2n) < 1.0e-6f) continue;
                    f3 cp, tn, bc;
                    closest_point_triangle(p, va, vb, vc, cp, tn, bc);
                    if (wt) {
                        const f3 df = cp - p;
                        const float dsq = dot3(df, df);
                        if (dsq < min_dist_sq) { min_dist_sq = dsq; min_face = t; sv = bc.y; sw = bc.z; found = true; }
                    } else {
                        const float dist = exact_length(cp - p);
                        if (dist < min_dist + eps) {
                            float weight = 0.0f;
                            const f3 cpa = cp - va, cqb = cp - vb, crc = cp - vc;
                            const float lp2 = dot3(cpa, cpa), lq2 = dot3(cqb, cqb), lr2 = dot3(crc, crc);
                            if (lp2 < eps2) weight = acosf(gmin(gmax(dot3(normalize3(e0), normalize3(e1)), -1.0f), 1.0f));
                            else if (lq2 < eps2) weight = acosf(gmin(gmax(dot3(normalize3(e2), normalize3(-e0)), -1.0f), 1.0f));
                            else if (lr2 < eps2) weight = acosf(gmin(gmax(dot3(normalize3(-e1), normalize3(-e2)), -1.0f), 1.0f));
                            else {
                                const float e0cp = dot3(e0, cpa), e2cq = dot3(e2, cqb), e1cp = dot3(e1, cpa);
                                if ((lp2 * e0n - e0cp * e0cp < eps2 * e0n) || (lq2 * e2n - e2cq * e2cq < eps2 * e2n) ||
                                    (lp2 * e1n - e1cp * e1cp < eps2 * e1n))
                                    weight = PI;
                                else weight = 2.0f * PI;
                            }
                            if (dist > min_dist - eps) {
                                acc_n = acc_n + weight * tn;
                                if (dist < min_dist) { min_dist = dist; min_face = t; sv = bc.y; sw = bc.z; }
                            } else {
                                min_dist = dist; min_face = t; sv = bc.y; sw = bc.z;
                                acc_n = weight * tn;
                            }
                            found = true;
                        }
                    }
                }
            } else {
                const float4 LL = a.nl[li], LU = a.nu[li], RL = a.nl[ri], RU = a.nu[ri];
                const float lds = dist_sq_aabb(p, xyz(LL), xyz(LU));
                const float rds = dist_sq_aabb(p, xyz(RL), xyz(RU));
                const float th = wt ? min_dist_sq : (min_dist + eps) * (min_dist + eps);
                if (sp + 2 > kStack) { atomicOr(a.err, 1u); continue; }
                if (lds < rds) {
                    if (rds < th) stk[sp++][lt] = ri;
                    if (lds < th) stk[sp++][lt] = li;
                } else {
                    if (lds < th) stk[sp++][lt] = li;
                    if (rds < th) stk[sp++][lt] = ri;
                }
            }
        }
        if (wt) min_dist = exact_sqrt(min_dist_sq);
        const bool emit = found && min_dist < max_dist;
        f3 bp = {0.f, 0.f, 0.f}, svel = {0.f, 0.f, 0.f}, nn = {0.f, 0.f, 1.f};
        if (emit) {
            const float u_m = (1.0f - sv) - sw, v_m = sv;
            const uint32_t i0 = a.tris[3 * min_face], i1 = a.tris[3 * min_face + 1], i2 = a.tris[3 * min_face + 2];
            const f3 v0 = ldv(a.verts, i0), v1 = ldv(a.verts, i1), v2 = ldv(a.verts, i2);
            const float w_e = (1.0f - u_m) - v_m;
            const f3 tb = v1 * v_m;
            const f3 tab = {fma_(v0.x, u_m, tb.x), fma_(v0.y, u_m, tb.y), fma_(v0.z, u_m, tb.z)};
            const f3 shp = {fma_(v2.x, w_e, tab.x), fma_(v2.y, w_e, tab.y), fma_(v2.z, w_e, tab.z)};
            const f3 w0 = ldv(a.vvel, i0), w1 = ldv(a.vvel, i1), w2 = ldv(a.vvel, i2);
            const f3 vb_ = w1 * v_m;
            const f3 vab = {fma_(w0.x, u_m, vb_.x), fma_(w0.y, u_m, vb_.y), fma_(w0.z, u_m, vb_.z)};
            svel = {fma_(w2.x, w_e, vab.x), fma_(w2.y, w_e, vab.y), fma_(w2.z, w_e, vab.z)};
            float sign;
            if (wt) {
                const f3 dir = {1.029700994f, 1.086472273f, 1.034694910f};
                Ray ray;
                ray.p = p;
                const f3 ad = {fabsf(dir.x), fabsf(dir.y), fabsf(dir.z)};
                int kz = 0;
                if (ad.y > ad.x) kz = 1;
                if (ad.z > (kz == 1 ? ad.y : ad.x)) kz = 2;
                int kx = kz + 1; if (kx == 3) kx = 0;
                int ky = kx + 1; if (ky == 3) ky = 0;
                if (comp3(dir, kz) < 0.0f) { int tmp = kx; kx = ky; ky = tmp; }
                ray.kx = kx; ray.ky = ky; ray.kz = kz;
                ray.Sx = comp3(dir, kx) / comp3(dir, kz);
                ray.Sy = comp3(dir, ky) / comp3(dir, kz);
                ray.Sz = 1.0f / comp3(dir, kz);
                uint32_t hits = 0;
                sp = 0;
                stk[sp++][lt] = mi.root;
                while (sp > 0) {
                    const uint32_t ni = stk[--sp][lt];
                    const float4 L = a.nl[ni], U = a.nu[ni];
                    if (!ray_box(p, dir, xyz(L), xyz(U))) continue;
                    const uint32_t lp = __float_as_uint(L.w);
                    const uint32_t li = lp & 0x3FFFFFFFu, ri = __float_as_uint(U.w) & 0x3FFFFFFFu;
                    if ((lp >> 30) & 3u) {
                        for (uint32_t pc = li; pc < ri; pc++) {
                            const uint32_t t = a.prim[pc];
                            if (ray_tri(ray, ldv(a.verts, a.tris[3 * t]), ldv(a.verts, a.tris[3 * t + 1]),
                                        ldv(a.verts, a.tris[3 * t + 2])))
                                hits++;
                        }
                    } else {
                        if (sp + 2 > kStack) { atomicOr(a.err, 2u); continue; }
                        stk[sp++][lt] = li;
                        stk[sp++][lt] = ri;
                    }
                }
                sign = (hits & 1u) ? -1.0f : 1.0f;
            } else {
                sign = dot3(acc_n, p - shp) > 0.0f ? 1.0f : -1.0f;
            }
            const f3 dl = p - shp;
            const float dlen = exact_length(dl);
            nn = dlen > 0.0f ? exact_div(dl, dlen) * sign : f3{0.f, 0.f, 1.f};
            const float d = dlen * sign;
            bp = {fma_(-nn.x, d, p.x), fma_(-nn.y, d, p.y), fma_(-nn.z, d, p.z)};
        }
        const int slot = wave_reserve(a.cc, emit);
        if (emit && slot < a.cmax) {
            a.cpart[slot] = tid;
            a.cshape[slot] = ci + 1;
            a.cbpos[slot] = make_float4(bp.x, bp.y, bp.z, 0.f);
            a.cbvel[slot] = make_float4(svel.x, svel.y, svel.z, 0.f);
            a.cnorm[slot] = make_float4(nn.x, nn.y, nn.z, 0.f);
        }
    }
}

void launch_mesh_contacts(const MeshArgs& a, hipStream_t s) {
    if (a.n <= 0 || a.nmesh <= 0) return;
    hipLaunchKernelGGL(k_mesh_contacts, dim3((a.n + kMeshBlock - 1) / kMeshBlock), dim3(kMeshBlock), 0, s, a);
}

}  // namespace xpbd
