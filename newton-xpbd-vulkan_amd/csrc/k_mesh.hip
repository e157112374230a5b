// k_mesh.hip -- particle-vs-triangle-mesh narrowphase over the packed LBVH,
// the replacement of shaders/generate_mesh_contacts.comp:181-414.
//
// Two launches per mesh, one block-level slot reservation per block for the
// contacts emitted (same-address global atomics serialise, so per-wave
// reservations throttle the kernels):
//   * early outs: the root box and a grid of distance lower bounds; the
//     particles left are queued for the tree walk;
//   * closest triangle: depth-first search with the DFS stack in LDS (one
//     32-deep column per thread, bank = thread), the reference's pruning, push
//     order and strict-less tie rule, so the chosen face is the same; then,
//     in the same wave, the watertight sign: parity of ray/triangle hits
//     along the reference's fixed ray.  The reference tests every triangle of
//     the mesh (:372-388); here the same watertight ray test runs only on the
//     triangles registered in the origin's cell of a grid of triangle
//     projections along that ray direction.  Every ray has the reference's
//     direction, so a triangle can only be hit by a ray whose origin projects
//     into the triangle's projection.  Registration is padded far beyond the
//     test's rounding error, so no triangle the test would report is skipped
//     and the hit count, hence the sign, is unchanged;
//   * non-watertight sign: the angle-weighted pseudo-normal (:266-313, :390).
// Node boxes are interleaved (lower, upper) float4 pairs: one 32-byte read per
// node.
#include "xpbd_kernels.h"

#include <algorithm>

namespace xpbd {

constexpr int kMeshBlock = 64;   // tree-walk workgroup: one wave
constexpr uint32_t kXcds = 8;    // MI355X: 8 XCDs, each with its own L2
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

// closest_point_triangle of generate_mesh_contacts.comp:115-179 (NORMAL:
// also the unit face normal, which only the non-watertight sign uses).
template <bool NORMAL>
__device__ void closest_point_triangle(f3 p, f3 a, f3 b, f3 c, f3& cp, f3& tn, f3& bary) {
    f3 ab = b - a, ac = c - a, ap = p - a;
    if (NORMAL) tn = normalize3(cross3(ab, ac));
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


// The reference's fixed ray direction (generate_mesh_contacts.comp:375) and
// the watertight test's per-ray constants.
__device__ __forceinline__ Ray make_ray(f3 p) {
    const f3 dir = {1.029700994f, 1.086472273f, 1.034694910f};
    Ray ray;
    ray.p = p;
    const f3 ad = {fabsf(dir.x), fabsf(dir.y), fabsf(dir.z)};
    int kz = 0;
    if (ad.y > ad.x) kz = 1;
    if (ad.z > (kz == 1 ? ad.y : ad.x)) kz = 2;
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (comp3(dir, kz) < 0.0f) { const int tmp = kx; kx = ky; ky = tmp; }
    ray.kx = kx; ray.ky = ky; ray.kz = kz;
    ray.Sx = comp3(dir, kx) / comp3(dir, kz);
    ray.Sy = comp3(dir, ky) / comp3(dir, kz);
    ray.Sz = 1.0f / comp3(dir, kz);
    return ray;
}

__device__ __forceinline__ float4 node_lo(const float4* nodes, uint32_t i) { return nodes[2 * i]; }
__device__ __forceinline__ float4 node_hi(const float4* nodes, uint32_t i) { return nodes[2 * i + 1]; }

// Closest triangle of generate_mesh_contacts.comp:201-331 for one particle and
// mesh: LDS stack column lt, the reference's pruning, push order and
// strict-less tie rule.  Returns whether a contact is emitted, with the face,
// barycentrics (v, w) and, for non-watertight meshes, the pseudo-normal sign.
template <bool STATS>
__device__ bool closest_face(const MeshArgs& a, const MeshInfo& mi, uint32_t start, f3 p, float rad,
                             uint32_t (*stk)[kMeshBlock], int lt, uint32_t& face, float& sv_out, float& sw_out,
                             float& sign, float& dsq_out, uint32_t& pops, uint32_t& tests) {
    const float eps = a.eps, eps2 = eps * eps;
    const float PI = 3.14159265359f;
    const float max_dist = (a.margin + mi.margin) + rad;
    const bool wt = mi.watertight != 0u;
    int sp = 0;
    stk[sp++][lt] = start;
    float min_dist = max_dist;
    f3 acc_n = {0.f, 0.f, 0.f};
    uint32_t min_face = 0;
    float sv = 0.f, sw = 0.f;
    bool found = false;
    float min_dist_sq = max_dist * max_dist;
    // Two nested loops so the lanes of a wave test triangles together: each
    // lane pops and expands internal nodes until it pops a leaf (or its stack
    // empties), then every lane holding a leaf tests it.  Each lane's own
    // sequence of pops, tests and threshold updates is the single loop's, so
    // the chosen face is too; lanes no longer run the triangle tests of one
    // iteration beside the node expansion of another.
    while (true) {
        uint32_t word = 0;
        bool leaf = false;
        while (sp > 0) {
            word = stk[--sp][lt];
            if (STATS) pops++;
            if (word >> 31) { leaf = true; break; }
            // internal node: one read of a record holding both children's
            // boxes and, for internal children, their children's boxes.  The
            // node's own pop-time box test is not repeated: a node it would
            // prune has children at least as far (their boxes lie inside its
            // box), so neither is pushed and nothing changes.  The child the
            // reference pops next (the last one pushed) is, when internal,
            // expanded right away from the same record: no triangle is tested
            // between its push and its pop, so its pop-time test passes and
            // the threshold is unchanged.
            const float4* rec = a.cnodes + 12 * size_t(word);
            float4 R[12];
#pragma unroll
            for (int q = 0; q < 12; q++) R[q] = rec[q];
            const float th = wt ? min_dist_sq : (min_dist + eps) * (min_dist + eps);
            const float lds = dist_sq_aabb(p, xyz(R[0]), xyz(R[1]));
            const float rds = dist_sq_aabb(p, xyz(R[2]), xyz(R[3]));
            const bool pl = lds < th, pr = rds < th;
            if (!pl && !pr) continue;
            if (sp + 2 > kStack) { atomicOr(a.err, 1u); continue; }
            const uint32_t lw = __float_as_uint(R[0].w), rw = __float_as_uint(R[2].w);
            // push order of the reference: (far, near) with near = the smaller
            // distance, left on ties going first (so right is popped first)
            bool near_left;
            if (lds < rds) {
                if (pl && pr) stk[sp++][lt] = rw;
                near_left = pl;
            } else {
                if (pl && pr) stk[sp++][lt] = lw;
                near_left = !pr;
            }
            const uint32_t nw = near_left ? lw : rw;
            if (nw >> 31) { stk[sp++][lt] = nw; continue; }
            const int g = near_left ? 4 : 8;
            const float cl = dist_sq_aabb(p, xyz(R[g]), xyz(R[g + 1]));
            const float cr = dist_sq_aabb(p, xyz(R[g + 2]), xyz(R[g + 3]));
            const bool ql = cl < th, qr = cr < th;
            if (STATS) pops++;
            if (!ql && !qr) continue;
            if (sp + 2 > kStack) { atomicOr(a.err, 1u); continue; }
            const uint32_t clw = __float_as_uint(R[g].w), crw = __float_as_uint(R[g + 2].w);
            if (cl < cr) {
                if (qr) stk[sp++][lt] = crw;
                if (ql) stk[sp++][lt] = clw;
            } else {
                if (ql) stk[sp++][lt] = clw;
                if (qr) stk[sp++][lt] = crw;
            }
        }
        if (!leaf) break;
        // leaf: the reference's pop-time box test, then its triangles
        // short leaves carry their primitive range in the word, so the
        // box and the first triangle load together; others go through
        // their node
        const uint32_t cnt = (word >> 27) & 15u;
        uint32_t li, ri;
        float4 L, U;
        float4 nA = make_float4(0.f, 0.f, 0.f, 0.f), nB = nA, nC = nA;
        if (cnt) {
            li = word & 0x07FFFFFFu;
            ri = li + cnt;
            nA = a.tsoup[3 * li]; nB = a.tsoup[3 * li + 1]; nC = a.tsoup[3 * li + 2];
            L = a.lbox[2 * size_t(li)]; U = a.lbox[2 * size_t(li) + 1];
        } else {
            const uint32_t ni = word & 0x07FFFFFFu;
            L = node_lo(a.nodes, ni); U = node_hi(a.nodes, ni);
            li = __float_as_uint(L.w) & 0x3FFFFFFFu; ri = __float_as_uint(U.w) & 0x3FFFFFFFu;
            if (li < ri) { nA = a.tsoup[3 * li]; nB = a.tsoup[3 * li + 1]; nC = a.tsoup[3 * li + 2]; }
        }
        const float nds = dist_sq_aabb(p, xyz(L), xyz(U));
        if (wt) { if (nds > min_dist_sq) continue; }
        else if (nds > (min_dist + eps) * (min_dist + eps)) continue;
        // the next triangle's corners load while this one is tested
        for (uint32_t pc = li; pc < ri; pc++) {
            const float4 A = nA, B = nB, C = nC;
            if (pc + 1 < ri) { nA = a.tsoup[3 * pc + 3]; nB = a.tsoup[3 * pc + 4]; nC = a.tsoup[3 * pc + 5]; }
            const uint32_t t = __float_as_uint(A.w);
            if (STATS) tests++;
            const f3 va = xyz(A), vb = xyz(B), vc = xyz(C);
            const f3 e0 = vb - va, e1 = vc - va, e2 = vc - vb;
            const float e0n = dot3(e0, e0), e1n = dot3(e1, e1), e2n = dot3(e2, e2);
            f3 cp, tn, bc;
            if (wt) {
                // the degenerate-triangle skip (:250-251) only matters for a
                // triangle that would become the closest: test it there
                closest_point_triangle<false>(p, va, vb, vc, cp, tn, bc);
                const f3 df = cp - p;
                const float dsq = dot3(df, df);
                if (dsq < min_dist_sq) {
                    if (len3(cross3(e0, e1)) / ((e0n + e1n) + e2n) < 1.0e-6f) continue;
                    min_dist_sq = dsq; min_face = t; sv = bc.y; sw = bc.z; found = true;
                }
            } else {
                if (len3(cross3(e0, e1)) / ((e0n + e1n) + e2n) < 1.0e-6f) continue;
                closest_point_triangle<true>(p, va, vb, vc, cp, tn, bc);
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
    }
    dsq_out = min_dist_sq;
    if (wt) min_dist = exact_sqrt(min_dist_sq);
    if (!(found && min_dist < max_dist)) return false;
    face = min_face;
    sv_out = sv;
    sw_out = sw;
    sign = 1.0f;
    if (!wt) {
        const float u_m = (1.0f - sv) - sw, v_m = sv;
        const f3 v0 = ldv(a.verts, a.tris[3 * min_face]), v1 = ldv(a.verts, a.tris[3 * min_face + 1]),
                 v2 = ldv(a.verts, a.tris[3 * min_face + 2]);
        const float w_e = (1.0f - u_m) - v_m;
        const f3 tb = v1 * v_m;
        const f3 tab = {fma_(v0.x, u_m, tb.x), fma_(v0.y, u_m, tb.y), fma_(v0.z, u_m, tb.z)};
        const f3 shp = {fma_(v2.x, w_e, tab.x), fma_(v2.y, w_e, tab.y), fma_(v2.z, w_e, tab.z)};
        sign = dot3(acc_n, p - shp) > 0.0f ? 1.0f : -1.0f;
    }
    return true;
}

// Conservative early out from the mesh's distance lower-bound grid: returns
// the node the closest-face search starts at, or kNoNode if no triangle lies
// within thr of p.  The bounds are valid below the grid's cap; beyond the
// grid every triangle is farther than the cap.  The pad (2e-5 at unit
// coordinates) is far above the rounding of the device's distance and cell
// computations.  The start node is the root unless the grid applies.
constexpr uint32_t kNoNode = 0xFFFFFFFFu;   // (node words carry a leaf flag in bit 31)
__device__ __forceinline__ uint32_t search_start(const MeshArgs& a, const MeshInfo& mi, f3 p, float thr) {
    const float sc = fmaxf(fmaxf(fabsf(p.x), fabsf(p.y)), fabsf(p.z));
    const float lim = thr + 1e-5f * (1.0f + sc);
    if (!(lim < mi.lb_cap)) return mi.root_word;
    const float fx = floorf((p.x - mi.lb_ox) * mi.lb_inv), fy = floorf((p.y - mi.lb_oy) * mi.lb_inv),
                fz = floorf((p.z - mi.lb_oz) * mi.lb_inv);
    if (!(fx >= 0.0f && fy >= 0.0f && fz >= 0.0f && fx < float(mi.lb_nx) && fy < float(mi.lb_ny) && fz < float(mi.lb_nz)))
        return kNoNode;
    const uint32_t c = mi.lb_off + (uint32_t(fz) * mi.lb_ny + uint32_t(fy)) * mi.lb_nx + uint32_t(fx);
    return a.lb[c] > lim ? kNoNode : a.lb_node[c];
}

// STATS: query counters for tuning (active particle x mesh pairs, particles
// that need the tree walk, node pops, closest-point tests, parity triangle
// tests, contacts), summed per wave into a.stats.
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Candidate test of a particle (position and radius P, flags f) against mesh
// mi: the reference's first pop (the root box against the search distance)
// and the distance-bound early out.  Returns the search start node word or
// kNoNode.
__device__ __forceinline__ uint32_t mesh_candidate(const MeshArgs& a, const MeshInfo& mi, float4 P, int f) {
    if (mi.prim_cnt == 0u || (f & 1) == 0) return kNoNode;   // an empty mesh has no nodes
    const f3 p = xyz(P);
    const float max_dist = (a.margin + mi.margin) + P.w;
    const float thr = mi.watertight ? max_dist : max_dist + a.eps;
    const float nds = dist_sq_aabb(p, xyz(node_lo(a.nodes, mi.root)), xyz(node_hi(a.nodes, mi.root)));
    if (nds > thr * thr) return kNoNode;
    return search_start(a, mi, p, thr);
}

// Queue counters of mesh ci for this substep: (queue length, unused).  Substeps alternate between two pairs; each substep's enqueue
// resets the other pair for the next one (the frame start resets both).
__device__ __forceinline__ uint32_t* mesh_counters(const MeshArgs& a, int ci, int slot) {
    return a.mcnt + 2 * (size_t(slot) * size_t(a.nmesh) + size_t(ci));
}

// Pass 1: candidates into the queue, one slot reservation per block (queue
// order is arbitrary; every queued particle's result is its own).  Each
// thread tests kEnqItems particles of its block's range (wave-striped), their
// loads issued together.  One per thread (1,024 blocks at 1,048,576
// particles) finishes the pass sooner beside the radix passes than two or
// four, which the walk behind it gains in the dense frames.
constexpr int kEnqItems = 1;
constexpr int kEnqBlock = 1024;   // fewer blocks, fewer same-address queue reservations
template <bool STATS>
__global__ __launch_bounds__(kEnqBlock) void k_mesh_enqueue(MeshArgs a, int ci) {
    __shared__ int s_tmp[kEnqBlock / 64 + 1];
    const MeshInfo mi = a.info[ci];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        uint32_t* o = mesh_counters(a, ci, a.slot ^ 1);
        o[0] = 0u;
        o[1] = 0u;
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i0 = blockIdx.x * (kEnqBlock * kEnqItems) + w * (64 * kEnqItems) + lane;
    float4 P[kEnqItems];
    int f[kEnqItems];
#pragma unroll
    for (int it = 0; it < kEnqItems; it++) {
        const int i = i0 + 64 * it;
        P[it] = make_float4(0.f, 0.f, 0.f, 0.f);
        f[it] = 0;
        if (i < a.n) { P[it] = a.P[i]; f[it] = a.flags[i]; }
    }
    uint32_t start[kEnqItems];
    int mine = 0;
#pragma unroll
    for (int it = 0; it < kEnqItems; it++) {
        start[it] = mesh_candidate(a, mi, P[it], f[it]);
        mine += start[it] != kNoNode ? 1 : 0;
    }
    int pos = block_reserve(reinterpret_cast<int*>(mesh_counters(a, ci, a.slot)), mine, s_tmp);
#pragma unroll
    for (int it = 0; it < kEnqItems; it++)
        if (start[it] != kNoNode) a.mq[pos++] = make_int2(i0 + 64 * it, int(start[it]));
    if (a.plane && ci == 0) {
        // generate_plane_contacts.comp:27-54: active particles within margin
        // + radius of the plane, one contact slot reservation per block
        bool want[kEnqItems];
        int np = 0;
#pragma unroll
        for (int it = 0; it < kEnqItems; it++) {
            want[it] = i0 + 64 * it < a.n && (f[it] & 1) && P[it].z - a.altitude < a.margin + P[it].w;
            np += want[it] ? 1 : 0;
        }
        int slot = block_reserve(a.cc, np, s_tmp);
#pragma unroll
        for (int it = 0; it < kEnqItems; it++) {
            if (!want[it]) continue;
            if (slot < a.cmax) {
                a.cpart[slot] = i0 + 64 * it;
                a.cshape[slot] = 0;
                a.cbpos[slot] = make_float4(P[it].x, P[it].y, a.altitude, 0.f);
                a.cbvel[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
                a.cnorm[slot] = make_float4(0.f, 0.f, 1.f, 0.f);
            }
            slot++;
        }
    }
    if (STATS) {
        uint32_t act = 0;
#pragma unroll
        for (int it = 0; it < kEnqItems; it++) act += (f[it] & 1) ? 1u : 0u;
        const uint64_t tot = wave_sum_u64(act);
        if (__lane_id() == 0 && tot) atomicAdd(reinterpret_cast<unsigned long long*>(a.stats), (unsigned long long)tot);
    }
}

// Up to four subtree roots below `start`, in the order the depth-first
// search of closest_face reaches them with threshold th (the search's
// initial one): the near child's children (expanded from this record, near
// first), then the far child's children (popped later and expanded by the
// same rule as any popped node).  A leaf child is a root itself.
__device__ int split_roots(const MeshArgs& a, uint32_t start, f3 p, float th, uint32_t roots[4]) {
    if (start >> 31) { roots[0] = start; return 1; }
    const float4* rec = a.cnodes + 12 * size_t(start);
    float4 R[12];
#pragma unroll
    for (int q = 0; q < 12; q++) R[q] = rec[q];
    const float lds = dist_sq_aabb(p, xyz(R[0]), xyz(R[1]));
    const float rds = dist_sq_aabb(p, xyz(R[2]), xyz(R[3]));
    const bool pl = lds < th, pr = rds < th;
    if (!pl && !pr) return 0;
    const bool near_left = lds < rds ? pl : !pr;
    int n = 0;
    // the near child, expanded at once from this record
    {
        const uint32_t w = __float_as_uint(near_left ? R[0].w : R[2].w);
        if (w >> 31) {
            roots[n++] = w;
        } else {
            const int g = near_left ? 4 : 8;
            const float cl = dist_sq_aabb(p, xyz(R[g]), xyz(R[g + 1]));
            const float cr = dist_sq_aabb(p, xyz(R[g + 2]), xyz(R[g + 3]));
            const uint32_t clw = __float_as_uint(R[g].w), crw = __float_as_uint(R[g + 2].w);
            if (cl < cr) {
                if (cl < th) roots[n++] = clw;
                if (cr < th) roots[n++] = crw;
            } else {
                if (cr < th) roots[n++] = crw;
                if (cl < th) roots[n++] = clw;
            }
        }
    }
    // the far child, when both were pushed: a popped node's rule on its children
    if (pl && pr) {
        const uint32_t w = __float_as_uint(near_left ? R[2].w : R[0].w);
        if (w >> 31) {
            roots[n++] = w;
        } else {
            const int g = near_left ? 8 : 4;
            const float l2 = dist_sq_aabb(p, xyz(R[g]), xyz(R[g + 1]));
            const float r2 = dist_sq_aabb(p, xyz(R[g + 2]), xyz(R[g + 3]));
            const bool ql = l2 < th, qr = r2 < th;
            const uint32_t lw2 = __float_as_uint(R[g].w), rw2 = __float_as_uint(R[g + 2].w);
            if (ql || qr) {
                const bool nl = l2 < r2 ? ql : !qr;
                roots[n++] = nl ? lw2 : rw2;
                if (ql && qr) roots[n++] = nl ? rw2 : lw2;
            }
        }
    }
    return n;
}

// Hits of the reference's fixed ray from p (generate_mesh_contacts.comp:
// 372-388) on the triangles registered in p's cell of mesh mi's projected
// parity grid, counted by the G consecutive lanes of a group (lane j of the
// group takes every G-th triangle; G is 1, 2 or 4) and summed over the group.
// The watertight test against every triangle of the mesh hits exactly these:
// a triangle the ray can hit has its padded projection registered in the
// cell of the ray's origin.  Every lane of the wave calls this; lanes with
// need == false test nothing and add 0.
template <int G>
__device__ __forceinline__ uint32_t ray_hits(const MeshArgs& a, const MeshInfo& mi, f3 p, bool need, int j,
                                             uint32_t& tested) {
    uint32_t b = 0, e = 0;
    const Ray ray = make_ray(p);
    if (need) {
        const float u = comp3(p, ray.kx) - ray.Sx * comp3(p, ray.kz);
        const float w = comp3(p, ray.ky) - ray.Sy * comp3(p, ray.kz);
        const float fu = floorf((u - mi.pg_u0) * mi.pg_inv), fw = floorf((w - mi.pg_w0) * mi.pg_inv);
        if (fu >= 0.0f && fw >= 0.0f && fu < float(mi.pg_dim) && fw < float(mi.pg_dim)) {
            const uint32_t cell = mi.pg_cell_off + uint32_t(fw) * mi.pg_dim + uint32_t(fu);
            b = a.pg_start[cell];
            e = a.pg_start[cell + 1];
        }
    }
    uint32_t hits = 0;
    uint32_t t = b + uint32_t(j);
    for (; t + G < e; t += 2 * G) {   // two triangles' loads in flight
        const float4 A0 = a.pg_soup[3 * t], B0 = a.pg_soup[3 * t + 1], C0 = a.pg_soup[3 * t + 2];
        const float4 A1 = a.pg_soup[3 * (t + G)], B1 = a.pg_soup[3 * (t + G) + 1], C1 = a.pg_soup[3 * (t + G) + 2];
        hits += ray_tri(ray, xyz(A0), xyz(B0), xyz(C0)) ? 1u : 0u;
        hits += ray_tri(ray, xyz(A1), xyz(B1), xyz(C1)) ? 1u : 0u;
    }
    if (t < e) {
        const float4 A = a.pg_soup[3 * t], B = a.pg_soup[3 * t + 1], C = a.pg_soup[3 * t + 2];
        hits += ray_tri(ray, xyz(A), xyz(B), xyz(C)) ? 1u : 0u;
    }
    if (j == 0) tested += e - b;
#pragma unroll
    for (int o = 1; o < G; o <<= 1) hits += __shfl_xor(hits, o);
    return hits;
}

// Pass 2: the tree walk, one wave per workgroup so the walks spread over
// every SIMD; contact slots reserved once per block, then the payloads
// (generate_mesh_contacts.comp:333-411).  Grid-stride over the queue with a
// block-uniform trip count.  G = 1: one queued particle per lane.  G = 4
// (watertight meshes): four lanes per particle, each searching one of the
// subtree roots of split_roots with its own threshold; the group keeps the
// smallest squared distance, ties to the earliest root.  A lane prunes only
// what the single search would prune (its threshold is never below the
// single search's at the same node), and the single search keeps the first
// face in its visiting order among those at the minimum, which is the
// earliest root's: the face is the same.  A watertight mesh's sign is then
// the parity of ray hits (ray_hits, the group's lanes sharing the
// triangles), decided before the contact is written.
template <bool STATS, int G>
__global__ __launch_bounds__(kMeshBlock) void k_mesh_walk(MeshArgs a, int ci) {
    __shared__ uint32_t stk[kStack][kMeshBlock];
    __shared__ int s_tmp[kMeshBlock / 64 + 1];
    constexpr uint32_t PER = kMeshBlock / G;   // queue entries per block step
    uint32_t st[6] = {0, 0, 0, 0, 0, 0};
    const MeshInfo mi = a.info[ci];
    const int lt = threadIdx.x, r = lt % G;
    uint32_t* const cnt = mesh_counters(a, ci, a.slot);
    const uint32_t total = cnt[0];
    // Workgroups are dealt to the eight XCDs round-robin; give each XCD one
    // contiguous eighth of the queue (spatially coherent particles), so the
    // tree nodes its walks share stay in that XCD's L2.
    const uint32_t xcd = blockIdx.x % kXcds, nbx = gridDim.x / kXcds;
    const uint32_t chunk = (total + kXcds - 1) / kXcds;
    const uint32_t lo = min(total, xcd * chunk), hi = min(total, lo + chunk);
    for (uint32_t base = lo + (blockIdx.x / kXcds) * PER; base < hi; base += nbx * PER) {
        const uint32_t q = base + uint32_t(lt / G);
        bool hit = false;
        uint32_t face = 0;
        float sv = 0.f, sw = 0.f, sg = 1.0f, dsq = 0.f;
        int pi = 0;
        float4 P = make_float4(0.f, 0.f, 0.f, 0.f);
        if (q < hi) {
            const int2 e = a.mq[q];
            pi = e.x;
            P = a.P[pi];
            const f3 p = xyz(P);
            if (STATS && r == 0) st[1]++;
            if (G == 1) {
                hit = closest_face<STATS>(a, mi, uint32_t(e.y), p, P.w, stk, lt, face, sv, sw, sg, dsq, st[2], st[3]);
            } else {
                const float max_dist = (a.margin + mi.margin) + P.w;
                uint32_t roots[4];
                const int nr = split_roots(a, uint32_t(e.y), p, max_dist * max_dist, roots);
                if (r < nr)
                    hit = closest_face<STATS>(a, mi, roots[r], p, P.w, stk, lt, face, sv, sw, sg, dsq, st[2], st[3]);
            }
        }
        if (G > 1) {   // butterfly over the group: (hit, dsq, rank) order, payload carried along
            int rank = r;
#pragma unroll
            for (int o = 1; o < G; o <<= 1) {
                const bool h2 = __shfl_xor(int(hit), o) != 0;
                const float d2 = __shfl_xor(dsq, o);
                const int r2 = __shfl_xor(rank, o);
                const uint32_t f2 = uint32_t(__shfl_xor(int(face), o));
                const float sv2 = __shfl_xor(sv, o), sw2 = __shfl_xor(sw, o);
                if (h2 && (!hit || d2 < dsq || (d2 == dsq && r2 < rank))) {
                    hit = true; dsq = d2; rank = r2; face = f2; sv = sv2; sw = sw2;
                }
            }
        }
        // the contact's surface point and offset, in every lane of a hit group
        const f3 p = xyz(P);
        f3 shp = {0.f, 0.f, 0.f}, dl = {0.f, 0.f, 0.f};
        float dlen = 0.0f;
        uint32_t i0 = 0, i1 = 0, i2 = 0;
        const float u_m = (1.0f - sv) - sw, v_m = sv;
        const float w_e = (1.0f - u_m) - v_m;
        if (hit) {
            i0 = a.tris[3 * face]; i1 = a.tris[3 * face + 1]; i2 = a.tris[3 * face + 2];
            const f3 v0 = ldv(a.verts, i0), v1 = ldv(a.verts, i1), v2 = ldv(a.verts, i2);
            const f3 tb = v1 * v_m;
            const f3 tab = {fma_(v0.x, u_m, tb.x), fma_(v0.y, u_m, tb.y), fma_(v0.z, u_m, tb.z)};
            shp = {fma_(v2.x, w_e, tab.x), fma_(v2.y, w_e, tab.y), fma_(v2.z, w_e, tab.z)};
            dl = p - shp;
            dlen = exact_length(dl);
        }
        // watertight: the outside sign, flipped when the ray's hit count is
        // odd (the reference tests the sign only for a nonzero offset)
        if (mi.watertight) {
            const bool need = hit && dlen > 0.0f;
            const uint32_t h = ray_hits<G>(a, mi, p, need, r, st[4]);
            if (need && (h & 1u)) sg = -sg;
        }
        if (G > 1) hit = hit && r == 0;   // the group's first lane writes the contact
        if (STATS) st[5] += hit ? 1u : 0u;
        const int slot = block_reserve(a.cc, hit ? 1 : 0, s_tmp);
        if (hit && slot < a.cmax) {
            const f3 w0 = ldv(a.vvel, i0), w1 = ldv(a.vvel, i1), w2 = ldv(a.vvel, i2);
            const f3 vb_ = w1 * v_m;
            const f3 vab = {fma_(w0.x, u_m, vb_.x), fma_(w0.y, u_m, vb_.y), fma_(w0.z, u_m, vb_.z)};
            const f3 svel = {fma_(w2.x, w_e, vab.x), fma_(w2.y, w_e, vab.y), fma_(w2.z, w_e, vab.z)};
            const f3 nn = dlen > 0.0f ? exact_div(dl, dlen) * sg : f3{0.f, 0.f, 1.f};
            const float d = dlen * sg;
            const f3 bp = {fma_(-nn.x, d, p.x), fma_(-nn.y, d, p.y), fma_(-nn.z, d, p.z)};
            a.cpart[slot] = pi;
            a.cshape[slot] = ci + 1;
            a.cbpos[slot] = make_float4(bp.x, bp.y, bp.z, 0.f);
            a.cbvel[slot] = make_float4(svel.x, svel.y, svel.z, 0.f);
            a.cnorm[slot] = make_float4(nn.x, nn.y, nn.z, 0.f);
            if (a.solve_inline) shape_solve_one(a.solve, slot, pi, ci + 1, bp, svel, nn);
        }
    }
    if (STATS) {
#pragma unroll
        for (int k = 1; k < 6; k++) {
            const uint64_t v = wave_sum_u64(st[k]);
            if (__lane_id() == 0 && v) atomicAdd(reinterpret_cast<unsigned long long*>(a.stats + k), (unsigned long long)v);
        }
    }
}

void launch_mesh_contacts(const MeshArgs& a, const uint8_t* watertight, bool beside, hipStream_t s) {
    if (a.n <= 0 || a.nmesh <= 0) return;
    const int nb = (a.n + kEnqBlock * kEnqItems - 1) / (kEnqBlock * kEnqItems);
    // walk grids are multiples of the XCD count, so every XCD gets the same
    // number of workgroups.  The four-lane walk holds about 90 VGPRs per lane
    // for its whole run; beside the radix passes (beside = true) 2,048
    // one-wave workgroups (two per SIMD) leave them registers and still finish
    // within their time (4,096 made the frame 2% slower, 1,024 9%); alone it
    // takes 4,096.
    const int cap4 = beside ? 2048 : 4096;
    const unsigned walk_grid1 = unsigned(std::max(1, std::min(4096, (a.n + kMeshBlock - 1) / kMeshBlock) / 8)) * 8u;
    const unsigned walk_grid4 = unsigned(std::max(1, std::min(cap4, (a.n + kMeshBlock / 4 - 1) / (kMeshBlock / 4)) / 8)) * 8u;
    for (int ci = 0; ci < a.nmesh; ci++) {
        const bool split = watertight[ci] != 0;
        if (a.stats) {
            hipLaunchKernelGGL(k_mesh_enqueue<true>, dim3(nb), dim3(kEnqBlock), 0, s, a, ci);
            if (split) hipLaunchKernelGGL((k_mesh_walk<true, 4>), dim3(walk_grid4), dim3(kMeshBlock), 0, s, a, ci);
            else hipLaunchKernelGGL((k_mesh_walk<true, 1>), dim3(walk_grid1), dim3(kMeshBlock), 0, s, a, ci);
        } else {
            hipLaunchKernelGGL(k_mesh_enqueue<false>, dim3(nb), dim3(kEnqBlock), 0, s, a, ci);
            if (split) hipLaunchKernelGGL((k_mesh_walk<false, 4>), dim3(walk_grid4), dim3(kMeshBlock), 0, s, a, ci);
            else hipLaunchKernelGGL((k_mesh_walk<false, 1>), dim3(walk_grid1), dim3(kMeshBlock), 0, s, a, ci);
        }
    }
}

}  // namespace xpbd
