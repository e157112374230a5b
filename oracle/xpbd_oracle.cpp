// xpbd_oracle.cpp -- CPU restatement of the reference XPBD GPU step.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product links, imports or calls
// this file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load it (through ctypes), and there only as the checker.
//
// What it restates (all citations relative to the reference repository root):
//   * XpbdSolver::step ordering, push-constant values, ping-pong and odd-K
//     copy-back                             src/scene.cpp:698-992
//   * collide_only / step_solve_only / set_contact_data
//                                           src/scene.cpp:1094-1245,1304-1315
//   * add_mesh watertight detection          src/scene.cpp:154-202
//   * finalize_meshes packing (shape/body tables, node re-basing, meta,
//     params, avg-edge from mesh 0)          src/scene.cpp:204-390
//   * update_mesh CPU refit + re-pack        src/scene.cpp:617-696, src/mesh_bvh.h:142-181
//   * LBVH build                             src/mesh_lbvh.h:49-223
//   * shaders: grid_cell_indices.comp:17-32, grid_offsets.comp:16-35,
//     generate_plane_contacts.comp:27-54, generate_mesh_contacts.comp:29-414,
//     integrate_particles.comp:26-60, solve_particle_shape_contacts.comp:50-200,
//     solve_particle_particle_contacts.comp:66-218,
//     apply_particle_deltas.comp:30-58, apply_particle_shape_restitution.comp:33-116.
//   * the radix sort (sort_*.comp) only by its semantics: a stable sort of
//     (cell, id) pairs by cell.
//
// Arithmetic conventions (the reference leaves these to the Vulkan driver;
// they are the restatement's definition, "parity unpinned" against lavapipe):
//   * compiled with -ffp-contract=off: a non-`precise` mul+add is two
//     roundings; every GLSL fma() is std::fma.
//   * dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z ; length(a) = sqrt(dot(a,a));
//     normalize(a) = a * (1/sqrt(dot(a,a))); division and sqrt correctly
//     rounded; acos = std::acos.
//   * Template parameter R = float gives the fp32 restatement, R = double the
//     FP64 host step used to measure absolute error.
//
// Documented deviations from the reference (each also applied by the product):
//   * apply_particle_deltas writes nothing for inactive particles
//     (apply_particle_deltas.comp:41), so after the odd-K copy-back they read
//     uninitialised q_out.  Here inactive particles pass through unchanged.
//   * set_contact_data updates only the host count (scene.cpp:1304-1315) while
//     the solve reads the device counter; here the injected count is the one
//     the next solve uses.
//   * contact lists are emitted in (particle, shape) order; the reference's
//     order is whatever its global atomic counter produced.
#include <algorithm>
#include <array>
#include <atomic>
#include <cstdlib>
#include <thread>
#include <type_traits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace orc {

static constexpr int GRID_DIM = 128;
static constexpr int NUM_CELLS = GRID_DIM * GRID_DIM * GRID_DIM;
static constexpr int GRID_ORIGIN = 1 << 20;

// Worker threads of the restatement's per-particle loops: ORC_THREADS, else
// the host's hardware threads (at most 16).  Every parallel loop below writes
// only its own particle's results (or per-chunk lists joined in chunk order),
// so the output is bit-identical to the sequential loop for any thread count.
static int orc_threads() {
    const char* e = std::getenv("ORC_THREADS");
    int t = e ? std::atoi(e) : int(std::thread::hardware_concurrency());
    return std::max(1, std::min(t, 16));
}

// f(begin, end, chunk) over [0, n) in chunks of `chunk` items, chunks handed
// out dynamically to orc_threads() threads.
template <class F> static void par_chunks(int64_t n, int64_t chunk, F&& f) {
    const int64_t nchunks = (n + chunk - 1) / chunk;
    const int nt = int(std::min<int64_t>(orc_threads(), nchunks));
    if (nt <= 1) {
        for (int64_t c = 0; c < nchunks; c++) f(c * chunk, std::min(n, (c + 1) * chunk), c);
        return;
    }
    std::atomic<int64_t> next{0};
    auto work = [&] {
        for (int64_t c; (c = next.fetch_add(1)) < nchunks;) f(c * chunk, std::min(n, (c + 1) * chunk), c);
    };
    std::vector<std::thread> th;
    for (int k = 1; k < nt; k++) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

template <class R> struct V3 { R x, y, z; };
template <class R> static inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> static inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> static inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> static inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> static inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> static inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> static inline R fmaR(R a, R b, R c) { return std::fma(a, b, c); }
template <class R> static inline R gmax(R a, R b) { return a < b ? b : a; }   // GLSL max
template <class R> static inline R gmin(R a, R b) { return b < a ? b : a; }   // GLSL min
template <class R> static inline R dot(V3<R> a, V3<R> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class R> static inline R length(V3<R> a) { return std::sqrt(dot(a, a)); }
template <class R> static inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> static inline V3<R> normalize(V3<R> a) {
    R inv = R(1) / std::sqrt(dot(a, a));
    return a * inv;
}
template <class R> static inline R comp(V3<R> a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// solve_particle_particle_contacts.comp:81-105 (also generate_mesh_contacts.comp:36-55)
template <class R> static inline R exact_sqrt(R a) {
    R d = std::sqrt(a);
    R r = fmaR(-d, d, a);
    return fmaR(r, R(0.5) / d, d);
}
template <class R> static inline R exact_length(V3<R> a) {
    R yy = a.y * a.y;
    R xx_yy = fmaR(a.x, a.x, yy);
    R sq = fmaR(a.z, a.z, xx_yy);
    return exact_sqrt(sq);
}
template <class R> static inline R exact_dot(V3<R> a, V3<R> b) {
    R y = a.y * b.y;
    R x_y = fmaR(a.x, b.x, y);
    return fmaR(a.z, b.z, x_y);
}
template <class R> static inline V3<R> exact_div(V3<R> v, R d) {
    R inv_d = R(1) / d;
    V3<R> q = v * inv_d;
    V3<R> rem = {fmaR(-d, q.x, v.x), fmaR(-d, q.y, v.y), fmaR(-d, q.z, v.z)};
    return {fmaR(rem.x, inv_d, q.x), fmaR(rem.y, inv_d, q.y), fmaR(rem.z, inv_d, q.z)};
}
template <class R> static inline R exact_div1(R v, R d) {   // generate_mesh_contacts.comp:57-62
    R inv_d = R(1) / d;
    R q = v * inv_d;
    R rem = fmaR(-d, q, v);
    return fmaR(rem, inv_d, q);
}
// apply_particle_deltas.comp:30-36 (Markstein refinement of the reciprocal)
template <class R> static inline V3<R> apply_exact_div(V3<R> a, R b) {
    R y = R(1) / b;
    R t = fmaR(-b, y, R(1));
    y = t * y + y;
    V3<R> q = a * y;
    V3<R> r = {fmaR(-q.x, b, a.x), fmaR(-q.y, b, a.y), fmaR(-q.z, b, a.z)};
    return {fmaR(r.x, y, q.x), fmaR(r.y, y, q.y), fmaR(r.z, y, q.z)};
}

// ---------------------------------------------------------------- config --
struct OrcConfig {
    float particle_v_max, soft_contact_mu, shape_material_mu, soft_contact_relaxation;
    float particle_adhesion, particle_mu, particle_cohesion, particle_max_radius;
    float search_radius, soft_contact_margin, soft_contact_restitution, ground_plane_altitude;
    float gravity[3];
    int32_t enable_restitution, has_ground_plane;
};

// --------------------------------------------------------------- LBVH ----
// src/mesh_bvh.h:14-20 packed node half: xyz + packed u32.
struct NodeHalf { float x, y, z; uint32_t packed; };

struct Lbvh {
    std::vector<NodeHalf> lowers, uppers;
    std::vector<uint32_t> prim;
    uint32_t root = 0;
};

static inline uint32_t spread_bits10(uint32_t n) {   // mesh_lbvh.h:12-19
    n &= 0x000003ffu;
    n = (n ^ (n << 16)) & 0xff0000ffu;
    n = (n ^ (n << 8)) & 0x0300f00fu;
    n = (n ^ (n << 4)) & 0x030c30c3u;
    n = (n ^ (n << 2)) & 0x09249249u;
    return n;
}
static inline uint32_t morton30(float x, float y, float z) {   // mesh_lbvh.h:21-29
    int ix = int(x * 1024.0f), iy = int(y * 1024.0f), iz = int(z * 1024.0f);
    auto cl = [](int v) { return uint32_t(std::max(0, std::min(1023, v))); };
    return (spread_bits10(cl(iz)) << 2) | (spread_bits10(cl(iy)) << 1) | spread_bits10(cl(ix));
}

// mesh_lbvh.h:49-223 restated.
static void lbvh_build(const float* pts, const uint32_t* tris, uint32_t nt, Lbvh& out) {
    struct B { float lo[3], hi[3]; };
    std::vector<B> tb(nt);
    float tlo[3] = {1e30f, 1e30f, 1e30f}, thi[3] = {-1e30f, -1e30f, -1e30f};
    for (uint32_t t = 0; t < nt; t++) {
        for (int a = 0; a < 3; a++) tb[t].lo[a] = tb[t].hi[a] = pts[tris[t * 3] * 3 + a];
        for (int v = 1; v < 3; v++)
            for (int a = 0; a < 3; a++) {
                float c = pts[tris[t * 3 + v] * 3 + a];
                tb[t].lo[a] = std::min(tb[t].lo[a], c);
                tb[t].hi[a] = std::max(tb[t].hi[a], c);
            }
        for (int a = 0; a < 3; a++) {
            tlo[a] = std::min(tlo[a], tb[t].lo[a]);
            thi[a] = std::max(thi[a], tb[t].hi[a]);
        }
    }
    float inv_e[3];
    for (int a = 0; a < 3; a++) inv_e[a] = 1.0f / ((thi[a] - tlo[a]) + 0.0001f);
    std::vector<std::pair<uint64_t, uint32_t>> items(nt);
    for (uint32_t t = 0; t < nt; t++) {
        float loc[3];
        for (int a = 0; a < 3; a++) {
            float c = 0.5f * (tb[t].lo[a] + tb[t].hi[a]);
            loc[a] = (c - tlo[a]) * inv_e[a];
        }
        items[t] = {uint64_t(morton30(loc[0], loc[1], loc[2])), t};
    }
    std::stable_sort(items.begin(), items.end(),
                     [](const std::pair<uint64_t, uint32_t>& a, const std::pair<uint64_t, uint32_t>& b) { return a.first < b.first; });
    out.prim.resize(nt);
    std::vector<uint64_t> keys(nt);
    for (uint32_t i = 0; i < nt; i++) { keys[i] = items[i].first; out.prim[i] = items[i].second; }
    std::vector<int> deltas(nt > 1 ? nt - 1 : 1);
    for (uint32_t i = 0; i + 1 < nt; i++) {
        uint64_t x = keys[i] ^ keys[i + 1];
        deltas[i] = x == 0 ? 64 : __builtin_clzll(x);
    }
    const uint32_t maxn = 2 * nt;
    out.lowers.assign(maxn, NodeHalf{0, 0, 0, 0});
    out.uppers.assign(maxn, NodeHalf{0, 0, 0, 0});
    std::vector<int> parent(maxn, -1), nchild(maxn, 0), rl(maxn, 0), rr(maxn, 0);
    for (uint32_t i = 0; i < nt; i++) {
        const B& b = tb[out.prim[i]];
        out.lowers[i] = NodeHalf{b.lo[0], b.lo[1], b.lo[2], i | (1u << 30)};
        out.uppers[i] = NodeHalf{b.hi[0], b.hi[1], b.hi[2], i + 1};
        rl[i] = rr[i] = int(i);
    }
    const int off = int(nt);
    for (uint32_t i = 0; i < nt; i++) {
        int cur = int(i);
        for (;;) {
            int l = rl[cur], r = rr[cur];
            if (l == 0 && r == int(nt - 1)) { out.root = uint32_t(cur); parent[cur] = -1; break; }
            bool pr;
            if (l == 0) pr = true;
            else if (r != int(nt - 1) && deltas[r] >= deltas[l - 1]) {
                if (deltas[r] == deltas[l - 1]) pr = bool((out.prim[l - 1] % 2) ^ (out.prim[r] % 2));
                else pr = true;
            } else pr = false;
            int p;
            if (pr) { p = r + off; parent[cur] = p; out.lowers[p].packed = uint32_t(cur); rl[p] = l; }
            else { p = l + off - 1; parent[cur] = p; out.uppers[p].packed = uint32_t(cur); rr[p] = r; }
            nchild[p]++;
            if (nchild[p] == 1) break;
            int lc = int(out.lowers[p].packed & 0x3FFFFFFFu), rc = int(out.uppers[p].packed & 0x3FFFFFFFu);
            out.lowers[p].x = std::min(out.lowers[lc].x, out.lowers[rc].x);
            out.lowers[p].y = std::min(out.lowers[lc].y, out.lowers[rc].y);
            out.lowers[p].z = std::min(out.lowers[lc].z, out.lowers[rc].z);
            out.uppers[p].x = std::max(out.uppers[lc].x, out.uppers[rc].x);
            out.uppers[p].y = std::max(out.uppers[lc].y, out.uppers[rc].y);
            out.uppers[p].z = std::max(out.uppers[lc].z, out.uppers[rc].z);
            cur = p;
        }
    }
    for (uint32_t ni = 0; ni < maxn; ni++) {   // mark_packed_leaf_nodes, leaf size 4
        if (nchild[ni] > 0 || ni < nt) {
            int depth = 1;
            for (int p = parent[ni]; p != -1; p = parent[p]) depth++;
            int l = rl[ni], r = rr[ni] + 1;
            if (r - l <= 4 || depth >= 32) {
                out.lowers[ni].packed = uint32_t(l) | (1u << 30);
                out.uppers[ni].packed = uint32_t(r);
            }
        }
    }
}

// src/mesh_bvh.h:145-181 (recursive refit from the root over the current points)
static void lbvh_refit(Lbvh& b, const float* pts, const uint32_t* tris, uint32_t ni) {
    uint32_t lp = b.lowers[ni].packed;
    if ((lp >> 30) & 3u) {
        uint32_t lo_i = lp & 0x3FFFFFFFu, hi_i = b.uppers[ni].packed & 0x3FFFFFFFu;
        float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
        for (uint32_t pi = lo_i; pi < hi_i; pi++) {
            uint32_t t = b.prim[pi];
            for (int v = 0; v < 3; v++)
                for (int a = 0; a < 3; a++) {
                    float c = pts[tris[t * 3 + v] * 3 + a];
                    lo[a] = std::min(lo[a], c);
                    hi[a] = std::max(hi[a], c);
                }
        }
        b.lowers[ni].x = lo[0]; b.lowers[ni].y = lo[1]; b.lowers[ni].z = lo[2];
        b.uppers[ni].x = hi[0]; b.uppers[ni].y = hi[1]; b.uppers[ni].z = hi[2];
    } else {
        uint32_t l = lp & 0x3FFFFFFFu, r = b.uppers[ni].packed;
        lbvh_refit(b, pts, tris, l);
        lbvh_refit(b, pts, tris, r);
        b.lowers[ni].x = std::min(b.lowers[l].x, b.lowers[r].x);
        b.lowers[ni].y = std::min(b.lowers[l].y, b.lowers[r].y);
        b.lowers[ni].z = std::min(b.lowers[l].z, b.lowers[r].z);
        b.uppers[ni].x = std::max(b.uppers[l].x, b.uppers[r].x);
        b.uppers[ni].y = std::max(b.uppers[l].y, b.uppers[r].y);
        b.uppers[ni].z = std::max(b.uppers[l].z, b.uppers[r].z);
    }
}

// ----------------------------------------------------------- mesh query ---
template <class R> static inline R dist_sq_aabb(V3<R> p, V3<R> lo, V3<R> hi) {   // generate_mesh_contacts.comp:31-34
    V3<R> d = {gmax(gmax(lo.x - p.x, R(0)), p.x - hi.x), gmax(gmax(lo.y - p.y, R(0)), p.y - hi.y),
               gmax(gmax(lo.z - p.z, R(0)), p.z - hi.z)};
    return dot(d, d);
}
template <class R> static inline R exact_dot3(V3<R> u, V3<R> v) {   // :64-68
    R py = u.y * v.y;
    R pxy = fmaR(u.x, v.x, py);
    return fmaR(u.z, v.z, pxy);
}
template <class R> static inline R diff_product(R a, R b, R c, R d) {   // :70-75
    R cd = c * d;
    R diff = fmaR(a, b, -cd);
    R err = fmaR(-c, d, cd);
    return diff + err;
}
// :77-113 watertight ray/triangle test (Woop et al.)
template <class R> static bool ray_tri_woop(V3<R> p, V3<R> dir, V3<R> a, V3<R> b, V3<R> c) {
    V3<R> ad = {std::fabs(dir.x), std::fabs(dir.y), std::fabs(dir.z)};
    int kz = 0;
    if (ad.y > ad.x) kz = 1;
    if (ad.z > (kz == 1 ? ad.y : ad.x)) kz = 2;
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (comp(dir, kz) < R(0)) std::swap(kx, ky);
    R Sx = comp(dir, kx) / comp(dir, kz), Sy = comp(dir, ky) / comp(dir, kz), Sz = R(1) / comp(dir, kz);
    V3<R> A = a - p, B = b - p, C = c - p;
    R Ax = comp(A, kx) - Sx * comp(A, kz), Ay = comp(A, ky) - Sy * comp(A, kz);
    R Bx = comp(B, kx) - Sx * comp(B, kz), By = comp(B, ky) - Sy * comp(B, kz);
    R Cx = comp(C, kx) - Sx * comp(C, kz), Cy = comp(C, ky) - Sy * comp(C, kz);
    R U = diff_product(Cx, By, Cy, Bx);
    R V = diff_product(Ax, Cy, Ay, Cx);
    R W = diff_product(Bx, Ay, By, Ax);
    if ((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0)) return false;
    R det = U + V + W;
    if (det == 0) return false;
    R Az = Sz * comp(A, kz), Bz = Sz * comp(B, kz), Cz = Sz * comp(C, kz);
    R T = U * Az + V * Bz + W * Cz;
    if ((T < 0 && det > 0) || (T > 0 && det < 0)) return false;
    return true;
}
// :115-179 closest point on triangle with barycentrics (u,v,w) for (a,b,c).
template <class R>
static void closest_point_triangle(V3<R> p, V3<R> a, V3<R> b, V3<R> c, V3<R>& cp, V3<R>& tn, V3<R>& bary) {
    V3<R> ab = b - a, ac = c - a, ap = p - a;
    tn = normalize(cross(ab, ac));
    R u, v;
    R d1 = exact_dot3(ab, ap), d2 = exact_dot3(ac, ap);
    if (d1 <= 0 && d2 <= 0) { v = 0; u = 1; }
    else {
        V3<R> bp = p - b;
        R d3 = exact_dot3(ab, bp), d4 = exact_dot3(ac, bp);
        if (d3 >= 0 && d4 <= d3) { v = 1; u = 0; }
        else {
            R vc = fmaR(d1, d4, -(d2 * d3));
            if (vc <= 0 && d1 >= 0 && d3 <= 0) { v = exact_div1(d1, d1 - d3); u = R(1) - v; }
            else {
                V3<R> cpp = p - c;
                R d5 = exact_dot3(ab, cpp), d6 = exact_dot3(ac, cpp);
                if (d6 >= 0 && d5 <= d6) { v = 0; u = 0; }
                else {
                    R vb = fmaR(d5, d2, -(d1 * d6));
                    if (vb <= 0 && d2 >= 0 && d6 <= 0) { R wv = exact_div1(d2, d2 - d6); v = 0; u = R(1) - wv; }
                    else {
                        R va = fmaR(d3, d6, -(d5 * d4));
                        if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
                            R wv = exact_div1(d4 - d3, (d4 - d3) + (d5 - d6));
                            v = R(1) - wv;
                            u = (R(1) - v) - wv;
                        } else {
                            R s_ab = va + vb;
                            R s_all = vc + s_ab;
                            R r0 = R(1) / s_all;
                            R rem = fmaR(-s_all, r0, R(1));
                            R den = fmaR(rem, r0, r0);
                            v = vb * den;
                            u = fmaR(-vc, den, R(1) - v);
                        }
                    }
                }
            }
        }
    }
    R w = (R(1) - u) - v;
    bary = {u, v, w};
    cp = {(u * a.x + v * b.x) + w * c.x, (u * a.y + v * b.y) + w * c.y, (u * a.z + v * b.z) + w * c.z};
}

// ---------------------------------------------------------------- solver --
template <class R> struct Solver {
    OrcConfig cfg{};
    int64_t n = 0, cmax = 0;
    bool initialized = false, finalized = false;
    // particle state (the reference's vec4 w component is always 0 and omitted)
    std::vector<V3<R>> q, qd, q_init, qd_init, q_out, qd_out;
    std::vector<R> invm, radius;
    std::vector<int32_t> flags;
    std::vector<R> delta;   // 3 per particle
    // grid
    std::vector<uint32_t> cells, ids, skeys, svals;
    std::vector<int32_t> cstart, cend;
    // contacts (SoA, as the reference)
    int32_t cc = 0;          // device counter
    int32_t last_count = 0;  // host-visible count
    std::vector<int32_t> cpart, cshape;
    std::vector<V3<R>> cbpos, cbvel, cnorm;
    // meshes
    struct Mesh { std::vector<float> verts; std::vector<uint32_t> tris; float margin; bool watertight; };
    std::vector<Mesh> meshes;
    std::vector<Lbvh> bvh;
    std::vector<float> all_verts, all_vels;
    std::vector<uint32_t> all_tris;
    std::vector<uint32_t> vert_off, tri_off;
    std::vector<NodeHalf> nl, nu;   // packed, re-based
    std::vector<uint32_t> gprim, roots, prim_off, prim_cnt;
    float avg_edge = 0;
    // bodies: one static body per mesh with identity pose (scene.cpp:220-238),
    // then the dynamic boxes added with add_box_body (builder extension).
    // Shapes: 0 = ground plane (body -1), 1..M = meshes, M+1.. = boxes.
    std::vector<int32_t> shape_body;
    std::vector<R> mat_mu, mat_mu_t, mat_mu_r;
    struct Box { float pos[3], rot[4], half[3], density; int32_t flags; };
    std::vector<Box> boxes;
    std::vector<V3<R>> b_pos, b_v, b_w, b_com, b_half;
    std::vector<std::array<R, 4>> b_rot;
    std::vector<std::array<R, 9>> b_I, b_invI;
    std::vector<R> b_invm;
    std::vector<int32_t> b_flags;
    std::vector<V3<R>> b_pos0, b_v0, b_w0;   // state at substep start
    std::vector<std::array<R, 4>> b_rot0;
    std::vector<R> b_delta;                  // 8 per body: lin at +0..2, ang at +4..6
    std::vector<R> b_invw;                   // rigid contacts touching each body
    // rigid contacts (box corners vs ground plane)
    int32_t rcc = 0;
    std::vector<int32_t> rc_s0, rc_s1;
    std::vector<V3<R>> rc_p0, rc_p1, rc_n;
    std::vector<R> rc_impulse;
    // body parameters (builder choices, settable)
    R angular_damping = R(0.05f), rigid_relax = R(0.8f), mu_torsional = 0, mu_rolling = 0;
    int32_t use_weights = 1;

    // contact relaxation as used by the solves (the config's value unless a
    // differentiable rollout overrides it in the solver's precision)
    R relax_r = R(0);
    std::vector<std::pair<int32_t, int32_t>>* pair_log = nullptr;   // contributing (i, j) pairs

    void init(const OrcConfig& c, int64_t np, int64_t mc) {
        cfg = c; n = np; cmax = mc ? mc : np * 2;
        relax_r = R(cfg.soft_contact_relaxation);
        q.assign(n, {0, 0, 0}); qd = q_init = qd_init = q_out = qd_out = q;
        invm.assign(n, 0); radius.assign(n, 0); flags.assign(n, 0);
        delta.assign(3 * n, 0);
        cells.assign(n, 0); ids.assign(n, 0); skeys.assign(n, 0); svals.assign(n, 0);
        cstart.assign(NUM_CELLS, 0); cend.assign(NUM_CELLS, 0);
        cpart.assign(cmax, 0); cshape.assign(cmax, 0);
        cbpos.assign(cmax, {0, 0, 0}); cbvel = cnorm = cbpos;
        initialized = true; finalized = false;
    }

    void add_mesh(const float* v, int64_t nv, const uint32_t* t, int64_t ntri, float margin, int wt) {
        Mesh m;
        m.verts.assign(v, v + nv * 3);
        m.tris.assign(t, t + ntri * 3);
        m.margin = margin;
        if (wt >= 0) m.watertight = wt != 0;
        else if (ntri == 0 || nv == 0) m.watertight = false;
        else {   // scene.cpp:170-198: every undirected edge shared by exactly 2 triangles
            std::vector<std::pair<uint32_t, uint32_t>> e;
            for (int64_t k = 0; k < ntri; k++)
                for (int s = 0; s < 3; s++) {
                    uint32_t a = t[k * 3 + s], b = t[k * 3 + (s + 1) % 3];
                    e.push_back({std::min(a, b), std::max(a, b)});
                }
            std::sort(e.begin(), e.end());
            bool ok = true;
            for (size_t i = 0; i < e.size();) {
                size_t j = i + 1;
                while (j < e.size() && e[j] == e[i]) j++;
                if (j - i != 2) { ok = false; break; }
                i = j;
            }
            m.watertight = ok;
        }
        meshes.push_back(std::move(m));
        finalized = false;
    }

    void pack_nodes() {
        size_t total = 0;
        for (auto& b : bvh) total += b.lowers.size();
        nl.assign(total, NodeHalf{0, 0, 0, 0}); nu = nl;
        size_t nb = 0, pb = 0;
        for (size_t c = 0; c < bvh.size(); c++) {
            for (size_t i = 0; i < bvh[c].lowers.size(); i++) {
                NodeHalf L = bvh[c].lowers[i], U = bvh[c].uppers[i];
                if ((L.packed >> 30) & 3u) {
                    L.packed = ((L.packed & 0x3FFFFFFFu) + uint32_t(pb)) | (1u << 30);
                    U.packed = (U.packed & 0x3FFFFFFFu) + uint32_t(pb);
                } else {
                    L.packed = (L.packed & 0x3FFFFFFFu) + uint32_t(nb);
                    U.packed = (U.packed & 0x3FFFFFFFu) + uint32_t(nb);
                }
                nl[nb + i] = L; nu[nb + i] = U;
            }
            nb += bvh[c].lowers.size();
            pb += bvh[c].prim.size();
        }
    }

    void finalize_meshes() {   // scene.cpp:204-390
        if (!initialized) throw std::runtime_error("XpbdSolver::init must be called before finalize_meshes");
        size_t M = meshes.size(), NB = boxes.size();
        shape_body.assign(M + 1 + NB, -1);
        for (size_t c = 0; c < M + NB; c++) shape_body[c + 1] = int32_t(c);
        mat_mu.assign(M + 1 + NB, R(cfg.shape_material_mu));
        mat_mu_t.assign(M + 1 + NB, mu_torsional);
        mat_mu_r.assign(M + 1 + NB, mu_rolling);
        const size_t B = M + NB;
        b_pos.assign(B, {0, 0, 0}); b_v = b_w = b_com = b_half = b_pos;
        b_rot.assign(B, {R(0), R(0), R(0), R(1)});
        b_I.assign(B, std::array<R, 9>{}); b_invI = b_I;
        b_invm.assign(B, R(0)); b_flags.assign(B, 0);
        for (size_t k = 0; k < NB; k++) {   // solid box: m = rho*8abc, I = m/3*(b^2+c^2, ...)
            const Box& bx = boxes[k];
            size_t b = M + k;
            b_pos[b] = {R(bx.pos[0]), R(bx.pos[1]), R(bx.pos[2])};
            b_rot[b] = {R(bx.rot[0]), R(bx.rot[1]), R(bx.rot[2]), R(bx.rot[3])};
            b_half[b] = {R(bx.half[0]), R(bx.half[1]), R(bx.half[2])};
            b_flags[b] = bx.flags;
            float hx = bx.half[0], hy = bx.half[1], hz = bx.half[2];
            float m = bx.density * 8.0f * hx * hy * hz;
            float ix = m / 3.0f * (hy * hy + hz * hz), iy = m / 3.0f * (hx * hx + hz * hz), iz = m / 3.0f * (hx * hx + hy * hy);
            b_invm[b] = m > 0.0f ? R(1.0f / m) : R(0);
            b_I[b] = {R(ix), R(0), R(0), R(0), R(iy), R(0), R(0), R(0), R(iz)};
            b_invI[b] = {ix > 0 ? R(1.0f / ix) : R(0), R(0), R(0), R(0), iy > 0 ? R(1.0f / iy) : R(0), R(0), R(0), R(0),
                         iz > 0 ? R(1.0f / iz) : R(0)};
        }
        b_delta.assign(8 * B, R(0));
        b_invw.assign(B, R(0));
        rc_s0.assign(8 * NB, 0); rc_s1 = rc_s0;
        rc_p0.assign(8 * NB, {0, 0, 0}); rc_p1 = rc_n = rc_p0;
        rc_impulse.assign(6 * 8 * NB, R(0));
        all_verts.clear(); all_tris.clear(); vert_off.clear(); tri_off.clear(); bvh.assign(M, Lbvh{});
        gprim.clear(); roots.clear(); prim_off.clear(); prim_cnt.clear();
        size_t nb = 0;
        for (size_t c = 0; c < M; c++) {
            uint32_t vo = uint32_t(all_verts.size() / 3);
            vert_off.push_back(vo);
            uint32_t to = uint32_t(all_tris.size() / 3);
            tri_off.push_back(to);
            all_verts.insert(all_verts.end(), meshes[c].verts.begin(), meshes[c].verts.end());
            for (uint32_t k : meshes[c].tris) all_tris.push_back(k + vo);
            uint32_t ntri = uint32_t(meshes[c].tris.size() / 3);
            lbvh_build(all_verts.data(), all_tris.data() + size_t(to) * 3, ntri, bvh[c]);
            if (c == 0) {   // scene.cpp:274-281,325: average edge length of mesh 0 only
                double s = 0.0;
                const uint32_t* tt = all_tris.data() + size_t(to) * 3;
                for (uint32_t k = 0; k < ntri; k++)
                    for (int e = 0; e < 3; e++) {
                        uint32_t i0 = tt[k * 3 + e], i1 = tt[k * 3 + (e + 1) % 3];
                        float dx = all_verts[i1 * 3] - all_verts[i0 * 3], dy = all_verts[i1 * 3 + 1] - all_verts[i0 * 3 + 1],
                              dz = all_verts[i1 * 3 + 2] - all_verts[i0 * 3 + 2];
                        s += std::sqrt(dx * dx + dy * dy + dz * dz);
                    }
                avg_edge = float(s / double(size_t(ntri) * 3));
            }
            roots.push_back(uint32_t(nb + bvh[c].root));
            prim_off.push_back(uint32_t(gprim.size()));
            prim_cnt.push_back(uint32_t(bvh[c].prim.size()));
            for (uint32_t p : bvh[c].prim) gprim.push_back(p + to);
            nb += bvh[c].lowers.size();
        }
        all_vels.assign(all_verts.size(), 0.0f);
        pack_nodes();
        finalized = true;
    }

    void update_mesh(size_t mi, const float* v, int64_t nv, const float* vel) {   // scene.cpp:617-696
        if (!finalized) throw std::runtime_error("update_mesh: call finalize_meshes() first");
        if (mi >= meshes.size()) throw std::runtime_error("update_mesh: mesh_index out of range");
        if (size_t(nv) != meshes[mi].verts.size() / 3) throw std::runtime_error("update_mesh: vertex count mismatch");
        size_t vs = vert_off[mi];
        for (int64_t i = 0; i < nv * 3; i++) all_verts[vs * 3 + i] = v[i];
        for (int64_t i = 0; i < nv * 3; i++) all_vels[vs * 3 + i] = vel ? vel[i] : 0.0f;
        std::copy(v, v + nv * 3, meshes[mi].verts.begin());
        // the per-mesh LBVH holds mesh-local triangle ids over global vertex ids
        if (!bvh[mi].lowers.empty())
            lbvh_refit(bvh[mi], all_verts.data(), all_tris.data() + size_t(tri_off[mi]) * 3, bvh[mi].root);
        pack_nodes();
    }

    void set_particles(const float* qi, const float* qdi, const float* w, const float* r, const int32_t* f, int64_t np) {
        if (!initialized) throw std::runtime_error("XpbdSolver not initialized");
        if (np != n) throw std::runtime_error("Particle count mismatch");
        for (int64_t i = 0; i < n; i++) {
            q[i] = {R(qi[i * 3]), R(qi[i * 3 + 1]), R(qi[i * 3 + 2])};
            qd[i] = qdi ? V3<R>{R(qdi[i * 3]), R(qdi[i * 3 + 1]), R(qdi[i * 3 + 2])} : V3<R>{0, 0, 0};
        }
        if (w) for (int64_t i = 0; i < n; i++) invm[i] = R(w[i]);
        if (r) for (int64_t i = 0; i < n; i++) radius[i] = R(r[i]);
        if (f) for (int64_t i = 0; i < n; i++) flags[i] = f[i];
    }

    // ---------------- kernels --------------------------------------------
    void emit(int32_t p, int32_t s, V3<R> bp, V3<R> bv, V3<R> nn) {
        int32_t slot = cc++;
        if (slot >= cmax) return;
        cpart[slot] = p; cshape[slot] = s; cbpos[slot] = bp; cbvel[slot] = bv; cnorm[slot] = nn;
    }

    void plane_contacts() {   // generate_plane_contacts.comp:27-54
        R margin = R(cfg.soft_contact_margin), alt = R(cfg.ground_plane_altitude);
        for (int64_t i = 0; i < n; i++) {
            if ((flags[i] & 1) == 0) continue;
            V3<R> p = q[i];
            R dist = p.z - alt;
            if (!(dist < margin + radius[i])) continue;
            emit(int32_t(i), 0, {p.x, p.y, alt}, {0, 0, 0}, {0, 0, 1});
        }
    }

    V3<R> vert(uint32_t k) const { return {R(all_verts[k * 3]), R(all_verts[k * 3 + 1]), R(all_verts[k * 3 + 2])}; }
    V3<R> vvel(uint32_t k) const { return {R(all_vels[k * 3]), R(all_vels[k * 3 + 1]), R(all_vels[k * 3 + 2])}; }
    V3<R> nlo(uint32_t i) const { return {R(nl[i].x), R(nl[i].y), R(nl[i].z)}; }
    V3<R> nhi(uint32_t i) const { return {R(nu[i].x), R(nu[i].y), R(nu[i].z)}; }

    struct MeshHit {
        int32_t p, s;
        V3<R> bp, bv, nn;
    };
    // generate_mesh_contacts.comp:181-414 for particle i; appends its contacts
    // (one per mesh at most) to out.
    void mesh_contacts_particle(int64_t i, std::vector<MeshHit>& out) const {
        const R margin = R(cfg.soft_contact_margin);
        const R eps = R(avg_edge) * R(1e-3f);
        const R eps2 = eps * eps;
        const R PI = R(3.14159265359f);
        V3<R> p = q[i];
        R rad = radius[i];
        for (size_t ci = 0; ci < meshes.size(); ci++) {
            R max_dist = (margin + R(meshes[ci].margin)) + rad;
            uint32_t stack[32];
            int sp = 0;
            stack[sp++] = roots[ci];
            R min_dist = max_dist;
            V3<R> acc_n = {0, 0, 0};
            uint32_t min_face = 0;
            R sv = 0, sw = 0;
            bool found = false;
            bool wt = meshes[ci].watertight;
            R min_dist_sq = max_dist * max_dist;
            while (sp > 0) {
                uint32_t ni = stack[--sp];
                R nds = dist_sq_aabb(p, nlo(ni), nhi(ni));
                if (wt) { if (nds > min_dist_sq) continue; }
                else if (nds > (min_dist + eps) * (min_dist + eps)) continue;
                uint32_t lp = nl[ni].packed;
                uint32_t li = lp & 0x3FFFFFFFu, leaf = (lp >> 30) & 3u, ri = nu[ni].packed & 0x3FFFFFFFu;
                if (leaf) {
                    for (uint32_t pc = li; pc < ri; pc++) {
                        uint32_t t = gprim[pc];
                        V3<R> a = vert(all_tris[t * 3]), b = vert(all_tris[t * 3 + 1]), c = vert(all_tris[t * 3 + 2]);
                        V3<R> e0 = b - a, e1 = c - a, e2 = c - b;
                        V3<R> nrm = cross(e0, e1);
                        R e0n = dot(e0, e0), e1n = dot(e1, e1), e2n = dot(e2, e2);
                        if (length(nrm) / ((e0n + e1n) + e2n) < R(1.0e-6f)) continue;
                        V3<R> cp, tn, bc;
                        closest_point_triangle(p, a, b, c, cp, tn, bc);
                        if (wt) {
                            V3<R> df = cp - p;
                            R dsq = dot(df, df);
                            if (dsq < min_dist_sq) { min_dist_sq = dsq; min_face = t; sv = bc.y; sw = bc.z; found = true; }
                        } else {
                            R dist = exact_length(cp - p);
                            if (dist < min_dist + eps) {
                                R weight = 0;
                                V3<R> cpa = cp - a, cqb = cp - b, crc = cp - c;
                                R lp2 = dot(cpa, cpa), lq2 = dot(cqb, cqb), lr2 = dot(crc, crc);
                                auto cl = [](R x) { return gmin(gmax(x, R(-1)), R(1)); };
                                if (lp2 < eps2) weight = std::acos(cl(dot(normalize(e0), normalize(e1))));
                                else if (lq2 < eps2) weight = std::acos(cl(dot(normalize(e2), normalize(-e0))));
                                else if (lr2 < eps2) weight = std::acos(cl(dot(normalize(-e1), normalize(-e2))));
                                else {
                                    R e0cp = dot(e0, cpa), e2cq = dot(e2, cqb), e1cp = dot(e1, cpa);
                                    if ((lp2 * e0n - e0cp * e0cp < eps2 * e0n) || (lq2 * e2n - e2cq * e2cq < eps2 * e2n) ||
                                        (lp2 * e1n - e1cp * e1cp < eps2 * e1n))
                                        weight = PI;
                                    else weight = R(2) * PI;
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
                    R lds = dist_sq_aabb(p, nlo(li), nhi(li));
                    R rds = dist_sq_aabb(p, nlo(ri), nhi(ri));
                    R th = wt ? min_dist_sq : (min_dist + eps) * (min_dist + eps);
                    if (lds < rds) {
                        if (rds < th) stack[sp++] = ri;
                        if (lds < th) stack[sp++] = li;
                    } else {
                        if (lds < th) stack[sp++] = li;
                        if (rds < th) stack[sp++] = ri;
                    }
                }
            }
            if (wt) min_dist = exact_sqrt(min_dist_sq);
            if (!(found && min_dist < max_dist)) continue;
            R u_m = (R(1) - sv) - sw;
            R v_m = sv;
            uint32_t i0 = all_tris[min_face * 3], i1 = all_tris[min_face * 3 + 1], i2 = all_tris[min_face * 3 + 2];
            V3<R> v0 = vert(i0), v1 = vert(i1), v2 = vert(i2);
            R w_e = (R(1) - u_m) - v_m;
            V3<R> tb = v1 * v_m;
            V3<R> tab = {fmaR(v0.x, u_m, tb.x), fmaR(v0.y, u_m, tb.y), fmaR(v0.z, u_m, tb.z)};
            V3<R> sp_ = {fmaR(v2.x, w_e, tab.x), fmaR(v2.y, w_e, tab.y), fmaR(v2.z, w_e, tab.z)};
            V3<R> w0 = vvel(i0), w1 = vvel(i1), w2 = vvel(i2);
            V3<R> vb = w1 * v_m;
            V3<R> vab = {fmaR(w0.x, u_m, vb.x), fmaR(w0.y, u_m, vb.y), fmaR(w0.z, u_m, vb.z)};
            V3<R> sv_ = {fmaR(w2.x, w_e, vab.x), fmaR(w2.y, w_e, vab.y), fmaR(w2.z, w_e, vab.z)};
            R sign;
            if (wt) {   // brute-force ray parity over every triangle of this mesh (:372-388)
                const V3<R> dir = {R(1.029700994f), R(1.086472273f), R(1.034694910f)};
                uint32_t hits = 0;
                for (uint32_t k = 0; k < prim_cnt[ci]; k++) {
                    uint32_t t = gprim[prim_off[ci] + k];
                    if (ray_tri_woop(p, dir, vert(all_tris[t * 3]), vert(all_tris[t * 3 + 1]), vert(all_tris[t * 3 + 2]))) hits++;
                }
                sign = (hits & 1u) ? R(-1) : R(1);
            } else {
                sign = dot(acc_n, p - sp_) > 0 ? R(1) : R(-1);
            }
            V3<R> dl = p - sp_;
            R dlen = exact_length(dl);
            V3<R> nn = dlen > 0 ? exact_div(dl, dlen) * sign : V3<R>{0, 0, 1};
            R d = dlen * sign;
            V3<R> bp = {fmaR(-nn.x, d, p.x), fmaR(-nn.y, d, p.y), fmaR(-nn.z, d, p.z)};
            out.push_back(MeshHit{int32_t(i), int32_t(ci) + 1, bp, sv_, nn});
        }
    }

    void mesh_contacts() {
        if (meshes.empty()) return;
        constexpr int64_t kChunk = 512;
        std::vector<std::vector<MeshHit>> hits(size_t((n + kChunk - 1) / kChunk));
        par_chunks(n, kChunk, [&](int64_t a, int64_t b, int64_t c) {
            for (int64_t i = a; i < b; i++)
                if (flags[i] & 1) mesh_contacts_particle(i, hits[size_t(c)]);
        });
        for (const auto& hv : hits)   // particle order, as the sequential loop emits them
            for (const MeshHit& h : hv) emit(h.p, h.s, h.bp, h.bv, h.nn);
    }

    void integrate(R dt) {   // integrate_particles.comp:26-60 (force and world index are always 0)
        V3<R> g = {R(cfg.gravity[0]), R(cfg.gravity[1]), R(cfg.gravity[2])};
        R vmax = R(cfg.particle_v_max);
        par_chunks(n, 16384, [&](int64_t i0, int64_t i1, int64_t) {
        for (int64_t i = i0; i < i1; i++) {
            V3<R> x0 = q[i], v0 = qd[i];
            q_init[i] = x0; qd_init[i] = v0;
            if ((flags[i] & 1) == 0) continue;
            R w = invm[i];
            R gs = w > 0 ? R(1) : R(0);
            V3<R> f0 = {0, 0, 0};
            V3<R> v1 = v0 + (f0 * w + g * gs) * dt;
            R mag = length(v1);
            if (mag > vmax) v1 = v1 * (vmax / mag);
            q[i] = x0 + v1 * dt;
            qd[i] = v1;
        }
        });
    }

    void grid_and_sort() {
        R inv = R(1.0f / cfg.search_radius);   // scene.cpp:702 (fp32 reciprocal)
        par_chunks(n, 16384, [&](int64_t i0, int64_t i1, int64_t) {
        for (int64_t i = i0; i < i1; i++) {    // grid_cell_indices.comp:17-32
            V3<R> p = q[i];
            int gx = int(std::floor(p.x * inv)) + GRID_ORIGIN;
            int gy = int(std::floor(p.y * inv)) + GRID_ORIGIN;
            int gz = int(std::floor(p.z * inv)) + GRID_ORIGIN;
            gx = std::max(gx, 0); gy = std::max(gy, 0); gz = std::max(gz, 0);
            cells[i] = uint32_t((gz % GRID_DIM) * GRID_DIM * GRID_DIM + (gy % GRID_DIM) * GRID_DIM + (gx % GRID_DIM));
            ids[i] = uint32_t(i);
        }
        });
        std::vector<uint32_t> order(n);
        for (int64_t i = 0; i < n; i++) order[i] = uint32_t(i);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cells[a] < cells[b]; });
        for (int64_t k = 0; k < n; k++) { skeys[k] = cells[order[k]]; svals[k] = ids[order[k]]; }
        std::fill(cstart.begin(), cstart.end(), 0);   // scene.cpp:918-923
        std::fill(cend.begin(), cend.end(), 0);
        for (int64_t t = 0; t < n; t++) {             // grid_offsets.comp:16-35
            int c = int(skeys[t]);
            if (t == 0) cstart[c] = 0;
            else {
                int pc = int(skeys[t - 1]);
                if (c != pc) { cstart[c] = int(t); cend[pc] = int(t); }
            }
            if (t == n - 1) cend[c] = int(t) + 1;
        }
    }

    static V3<R> quat_rotate(const R qr[4], V3<R> x) {   // solve_particle_shape_contacts.comp:50-58
        R c = R(2) * qr[3] * qr[3] - R(1);
        R d = R(2) * (qr[0] * x.x + qr[1] * x.y + qr[2] * x.z);
        return {x.x * c + qr[0] * d + (qr[1] * x.z - qr[2] * x.y) * qr[3] * R(2),
                x.y * c + qr[1] * d + (qr[2] * x.x - qr[0] * x.z) * qr[3] * R(2),
                x.z * c + qr[2] * d + (qr[0] * x.y - qr[1] * x.x) * qr[3] * R(2)};
    }
    static V3<R> quat_rotate_inv(const R qr[4], V3<R> x) {   // :60-68
        R c = R(2) * qr[3] * qr[3] - R(1);
        R d = R(2) * (qr[0] * x.x + qr[1] * x.y + qr[2] * x.z);
        return {x.x * c + qr[0] * d - (qr[1] * x.z - qr[2] * x.y) * qr[3] * R(2),
                x.y * c + qr[1] * d - (qr[2] * x.x - qr[0] * x.z) * qr[3] * R(2),
                x.z * c + qr[2] * d - (qr[0] * x.y - qr[1] * x.x) * qr[3] * R(2)};
    }

    // body frame of a shape's body; identity / zeros for body < 0 (the plane)
    // (solve_particle_shape_contacts.comp:123-131, :137-141)
    struct Frame { V3<R> pos; R rot[4]; V3<R> com, v, w; };
    Frame frame(int body) const {
        Frame F{{0, 0, 0}, {0, 0, 0, 1}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        if (body >= 0) {
            F.pos = b_pos[body];
            for (int k = 0; k < 4; k++) F.rot[k] = b_rot[body][k];
            F.com = b_com[body]; F.v = b_v[body]; F.w = b_w[body];
        }
        return F;
    }

    // solve_particle_shape_contacts.comp:104-200
    void shape_solve(const std::vector<V3<R>>& X, const std::vector<V3<R>>& Vv, R dt) {
        int32_t count = std::min(int32_t(cmax), cc);
        R relax = relax_r, ka = R(cfg.particle_adhesion), pmu = R(cfg.soft_contact_mu);
        for (int32_t t = 0; t < count; t++) {
            int s = cshape[t];
            int body = shape_body[s];
            int pi = cpart[t];
            int pf = flags[pi];
            if ((pf & 1) == 0) continue;
            if (pf & 2) {
                if (body < 0) continue;
                if (b_flags[body] & 4) continue;
                if (b_invm[body] == 0) continue;
            }
            V3<R> px = X[pi], pv = Vv[pi];
            Frame F = frame(body);
            V3<R> bx = F.pos + quat_rotate(F.rot, cbpos[t]);
            V3<R> r = bx - (F.pos + quat_rotate(F.rot, F.com));
            V3<R> nn = cnorm[t];
            V3<R> dp = px - bx;
            R c = fmaR(nn.z, dp.z, fmaR(nn.x, dp.x, nn.y * dp.y)) - radius[pi];
            if (c > ka) continue;
            R mu = R(0.5) * (pmu + mat_mu[s]);
            V3<R> bv = F.v + cross(F.w, r) + quat_rotate(F.rot, cbvel[t]);
            V3<R> v = pv - bv;
            R lambda_n = c;
            V3<R> delta_n = nn * lambda_n;
            R vn = nn.y * v.y;
            vn = fmaR(nn.x, v.x, vn);
            vn = fmaR(nn.z, v.z, vn);
            V3<R> vt = {fmaR(-nn.x, vn, v.x), fmaR(-nn.y, vn, v.y), fmaR(-nn.z, vn, v.z)};
            R w1 = invm[pi];
            R w2 = 0;
            if (body >= 0) {
                V3<R> ang = cross(r, nn);
                V3<R> ra = quat_rotate_inv(F.rot, ang);
                const std::array<R, 9>& I = b_invI[body];   // row-major read, as this shader does
                V3<R> Ira = {I[0] * ra.x + I[1] * ra.y + I[2] * ra.z, I[3] * ra.x + I[4] * ra.y + I[5] * ra.z,
                             I[6] * ra.x + I[7] * ra.y + I[8] * ra.z};
                w2 = b_invm[body] + dot(ra, Ira);
            }
            R denom = w1 + w2;
            if (denom == 0) continue;
            R l2 = vt.y * vt.y;
            l2 = fmaR(vt.x, vt.x, l2);
            l2 = fmaR(vt.z, vt.z, l2);
            R vt_len = exact_sqrt(l2);
            R lambda_f = gmax(mu * lambda_n, -vt_len * dt);
            V3<R> vt_dir = vt_len > 0 ? exact_div(vt, vt_len) : V3<R>{0, 0, 0};
            V3<R> diff = {fmaR(vt_dir.x, lambda_f, -delta_n.x), fmaR(vt_dir.y, lambda_f, -delta_n.y),
                          fmaR(vt_dir.z, lambda_f, -delta_n.z)};
            V3<R> dtot = exact_div(diff, denom) * relax;
            V3<R> fd = w1 * dtot;
            delta[pi * 3 + 0] += fd.x;
            delta[pi * 3 + 1] += fd.y;
            delta[pi * 3 + 2] += fd.z;
            if (body >= 0 && b_invm[body] > 0) {
                V3<R> dv = {dtot.x / dt, dtot.y / dt, dtot.z / dt};
                V3<R> dw = cross(r, dv);
                R* bd = &b_delta[8 * body];
                bd[0] += -dv.x; bd[1] += -dv.y; bd[2] += -dv.z;
                bd[4] += -dw.x; bd[5] += -dw.y; bd[6] += -dw.z;
            }
        }
    }

    // ---- rigid bodies (integrate_bodies.comp, solve_body_contact_positions.comp,
    //      apply_body_deltas.comp); the reference ships these unwired ----
    static V3<R> qrot(const std::array<R, 4>& q, V3<R> v) {   // integrate_bodies.comp:62-65
        V3<R> u = {q[0], q[1], q[2]};
        return v + R(2) * cross(u, cross(u, v) + v * q[3]);
    }
    static V3<R> qrot_inv(const std::array<R, 4>& q, V3<R> v) { return qrot({-q[0], -q[1], -q[2], q[3]}, v); }
    static std::array<R, 4> qmul(const std::array<R, 4>& p, const std::array<R, 4>& q) {   // :55-60
        V3<R> pu = {p[0], p[1], p[2]}, qu = {q[0], q[1], q[2]};
        R w = p[3] * q[3] - dot(pu, qu);
        V3<R> xyz = p[3] * qu + q[3] * pu + cross(pu, qu);
        return {xyz.x, xyz.y, xyz.z, w};
    }
    static std::array<R, 4> qnormalize(const std::array<R, 4>& q) {
        R d = ((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3];
        R inv = R(1) / std::sqrt(d);
        return {q[0] * inv, q[1] * inv, q[2] * inv, q[3] * inv};
    }
    static V3<R> mcol(const std::array<R, 9>& m, V3<R> v) {   // column-major 3x3 times v
        return V3<R>{m[0], m[1], m[2]} * v.x + V3<R>{m[3], m[4], m[5]} * v.y + V3<R>{m[6], m[7], m[8]} * v.z;
    }
    static V3<R> tf_point(V3<R> pos, const std::array<R, 4>& q, V3<R> p) {   // solve_body_contact_positions.comp:58-62
        V3<R> u = {q[0], q[1], q[2]};
        return p + R(2) * cross(u, cross(u, p) + p * q[3]) + pos;
    }
    static V3<R> qrot_inv_b(const std::array<R, 4>& q, V3<R> v) {   // :64-67
        V3<R> nq = {-q[0], -q[1], -q[2]};
        return v + R(2) * cross(nq, cross(nq, v) + v * q[3]);
    }

    void integrate_bodies(R dt) {   // integrate_bodies.comp:73-130 (force, torque, world index zero)
        V3<R> g = {R(cfg.gravity[0]), R(cfg.gravity[1]), R(cfg.gravity[2])};
        for (size_t b = meshes.size(); b < b_pos.size(); b++) {
            if (b_flags[b] & 2) continue;   // KINEMATIC: pass through
            V3<R> x0 = b_pos[b], v0 = b_v[b], w0 = b_w[b], com = b_com[b];
            std::array<R, 4> r0 = b_rot[b];
            R im = b_invm[b];
            V3<R> f0 = {0, 0, 0}, t0 = {0, 0, 0};
            V3<R> xc = x0 + qrot(r0, com);
            R ms = im != 0 ? R(1) : R(0);
            V3<R> v1 = v0 + (f0 * im + g * ms) * dt;
            V3<R> x1 = xc + v1 * dt;
            V3<R> wb = qrot_inv(r0, w0);
            V3<R> Iwb = mcol(b_I[b], wb);
            V3<R> tb = qrot_inv(r0, t0) - cross(wb, Iwb);
            V3<R> dwb = mcol(b_invI[b], tb) * dt;
            V3<R> w1 = qrot(r0, wb + dwb);
            std::array<R, 4> wq = qmul({w1.x, w1.y, w1.z, R(0)}, r0);
            R hdt = R(0.5) * dt;
            std::array<R, 4> r1 = qnormalize({r0[0] + wq[0] * hdt, r0[1] + wq[1] * hdt, r0[2] + wq[2] * hdt, r0[3] + wq[3] * hdt});
            w1 = w1 * (R(1) - angular_damping * dt);
            b_pos[b] = x1 - qrot(r1, com);
            b_rot[b] = r1;
            b_v[b] = v1;
            b_w[b] = w1;
        }
    }

    // Builder-written rigid contact generator: the 8 corners of every box
    // against the ground plane; shape0 = box, shape1 = plane (shape 0, body
    // -1), normal from shape0 to shape1 = (0,0,-1), point0 in the box frame,
    // point1 in world.
    void body_plane_contacts() {
        rcc = 0;
        if (!cfg.has_ground_plane) return;
        R alt = R(cfg.ground_plane_altitude), margin = R(cfg.soft_contact_margin);
        size_t M = meshes.size();
        for (size_t k = 0; k < boxes.size(); k++) {
            size_t b = M + k;
            for (int c = 0; c < 8; c++) {
                V3<R> h = b_half[b];
                V3<R> lc = {(c & 1) ? h.x : -h.x, (c & 2) ? h.y : -h.y, (c & 4) ? h.z : -h.z};
                V3<R> wc = tf_point(b_pos[b], b_rot[b], lc);
                if (!(wc.z - alt < margin)) continue;
                int32_t slot = rcc++;
                rc_s0[slot] = int32_t(M + 1 + k);
                rc_s1[slot] = 0;
                rc_p0[slot] = lc;
                rc_p1[slot] = {wc.x, wc.y, alt};
                rc_n[slot] = {0, 0, -1};
            }
        }
    }

    // Builder-written particle-box contact generator: closest point of the
    // oriented box to the particle centre; contact when the signed distance is
    // below soft_contact_margin + radius.  body_pos in the box frame, normal in
    // world from the box surface towards the particle, body_vel 0.
    void particle_box_contacts() {
        if (boxes.empty()) return;
        R margin = R(cfg.soft_contact_margin);
        size_t M = meshes.size();
        for (int64_t i = 0; i < n; i++) {
            if ((flags[i] & 1) == 0) continue;
            V3<R> p = q[i];
            for (size_t k = 0; k < boxes.size(); k++) {
                size_t b = M + k;
                V3<R> h = b_half[b];
                V3<R> rel = p - b_pos[b];
                R reach = (h.x + h.y + h.z) + margin + radius[i];
                if (dot(rel, rel) > reach * reach) continue;
                V3<R> l = qrot_inv(b_rot[b], rel);
                V3<R> cp = {gmin(gmax(l.x, -h.x), h.x), gmin(gmax(l.y, -h.y), h.y), gmin(gmax(l.z, -h.z), h.z)};
                V3<R> nl;
                R sd;
                bool inside = std::fabs(l.x) < h.x && std::fabs(l.y) < h.y && std::fabs(l.z) < h.z;
                if (!inside) {
                    V3<R> d = l - cp;
                    sd = length(d);
                    if (!(sd > 0)) continue;
                    nl = d * (R(1) / sd);
                } else {
                    R gx = h.x - std::fabs(l.x), gy = h.y - std::fabs(l.y), gz = h.z - std::fabs(l.z);
                    int ax = (gx <= gy && gx <= gz) ? 0 : (gy <= gz ? 1 : 2);
                    R sg = comp(l, ax) < 0 ? R(-1) : R(1);
                    nl = {ax == 0 ? sg : R(0), ax == 1 ? sg : R(0), ax == 2 ? sg : R(0)};
                    if (ax == 0) { cp.x = sg * h.x; sd = -gx; }
                    else if (ax == 1) { cp.y = sg * h.y; sd = -gy; }
                    else { cp.z = sg * h.z; sd = -gz; }
                }
                if (!(sd < margin + radius[i])) continue;
                emit(int32_t(i), int32_t(M + 1 + k), cp, {0, 0, 0}, qrot(b_rot[b], nl));
            }
        }
    }

    // solve_body_contact_positions.comp:70-85
    R ccd(R err, const std::array<R, 4>& qa, const std::array<R, 4>& qb, R mia, R mib, const std::array<R, 9>& Ia,
          const std::array<R, 9>& Ib, V3<R> la, V3<R> lb, V3<R> aa, V3<R> ab, R relax, R dt) const {
        R den = 0;
        den += dot(la, la) * mia;
        den += dot(lb, lb) * mib;
        V3<R> ra = qrot_inv_b(qa, aa), rb = qrot_inv_b(qb, ab);
        den += dot(ra, mcol(Ia, ra));
        den += dot(rb, mcol(Ib, rb));
        R dl = -err;
        if (den > 0) dl /= dt * den;
        return dl * relax;
    }

    // solve_body_contact_positions.comp:97-279 (offsets and thicknesses zero)
    void body_contact_solve(R dt) {
        const std::array<R, 4> ident = {0, 0, 0, 1};
        const std::array<R, 9> zero9{};
        for (int32_t ci = 0; ci < rcc; ci++) {
            int sa = rc_s0[ci], sb = rc_s1[ci];
            if (sa == sb) continue;
            int ba = sa >= 0 ? shape_body[sa] : -1, bb = sb >= 0 ? shape_body[sb] : -1;
            if (ba == bb) continue;
            V3<R> pa = {0, 0, 0}, pb = {0, 0, 0};
            std::array<R, 4> qa = ident, qb = ident;
            if (ba >= 0) { pa = b_pos[ba]; qa = b_rot[ba]; }
            if (bb >= 0) { pb = b_pos[bb]; qb = b_rot[bb]; }
            V3<R> p0 = rc_p0[ci], p1 = rc_p1[ci];
            V3<R> bxa = tf_point(pa, qa, p0), bxb = tf_point(pb, qb, p1);
            V3<R> nn = rc_n[ci];
            R d = dot(nn, bxb - bxa) - (R(0) + R(0));
            if (d >= 0) continue;
            R mia = 0, mib = 0;
            std::array<R, 9> Ia = zero9, Ib = zero9;
            V3<R> ca = {0, 0, 0}, cb = {0, 0, 0}, wa = {0, 0, 0}, wb = {0, 0, 0};
            V3<R> oa = {0, 0, 0}, ob = {0, 0, 0};
            if (ba >= 0) { mia = b_invm[ba]; Ia = b_invI[ba]; ca = b_com[ba]; wa = b_w[ba]; }
            if (bb >= 0) { mib = b_invm[bb]; Ib = b_invI[bb]; cb = b_com[bb]; wb = b_w[bb]; }
            int nz = 0;
            R mu = 0, mut = 0, mur = 0;
            if (sa >= 0) { nz += 1; mu += mat_mu[sa]; mut += mat_mu_t[sa]; mur += mat_mu_r[sa]; }
            if (sb >= 0) { nz += 1; mu += mat_mu[sb]; mut += mat_mu_t[sb]; mur += mat_mu_r[sb]; }
            if (nz > 0) { mu /= R(nz); mut /= R(nz); mur /= R(nz); }
            V3<R> ra = bxa - tf_point(pa, qa, ca), rb = bxb - tf_point(pb, qb, cb);
            V3<R> anga = -cross(ra, nn), angb = cross(rb, nn);
            if (ba >= 0) b_invw[ba] += R(1);
            if (bb >= 0) b_invw[bb] += R(1);
            R ln = ccd(d, qa, qb, mia, mib, Ia, Ib, -nn, nn, anga, angb, rigid_relax, dt);
            V3<R> lda = -nn * ln, ldb = nn * ln, ada = anga * ln, adb = angb * ln;
            if (mu > 0) {
                V3<R> sxa = tf_point(pa, qa, p0 + oa), sxb = tf_point(pb, qb, p1 + ob);
                V3<R> dl = sxb - sxa;
                V3<R> fd = dl - dot(nn, dl) * nn;
                V3<R> fra = sxa - tf_point(pa, qa, ca), frb = sxb - tf_point(pb, qb, cb);
                V3<R> rv = {0, 0, 0};
                if (ba >= 0 && (b_flags[ba] & 2)) {
                    V3<R> va = b_v[ba] + cross(wa, fra);
                    rv = rv - (va - dot(nn, va) * nn);
                }
                if (bb >= 0 && (b_flags[bb] & 2)) {
                    V3<R> vb = b_v[bb] + cross(wb, frb);
                    rv = rv + (vb - dot(nn, vb) * nn);
                }
                fd = fd + rv * dt;
                V3<R> perp = normalize(fd);
                V3<R> afa = -cross(fra, perp), afb = cross(frb, perp);
                R err = length(fd);
                if (err > 0) {
                    R lf = ccd(err, qa, qb, mia, mib, Ia, Ib, -perp, perp, afa, afb, rigid_relax, dt);
                    lf = gmax(lf, -ln * mu);
                    lda = lda - perp * lf; ldb = ldb + perp * lf;
                    ada = ada + afa * lf; adb = adb + afb * lf;
                }
            }
            V3<R> dom = wb - wa;
            if (mut > 0) {
                R err = dot(dom, nn) * dt;
                if (std::fabs(err) > 0) {
                    R lt = ccd(err, qa, qb, mia, mib, Ia, Ib, {0, 0, 0}, {0, 0, 0}, -nn, nn, rigid_relax, dt);
                    lt = gmin(gmax(lt, -ln * mut), ln * mut);
                    ada = ada - nn * lt; adb = adb + nn * lt;
                }
            }
            if (mur > 0) {
                dom = dom - dot(nn, dom) * nn;
                R err = length(dom) * dt;
                if (err > 0) {
                    V3<R> rn = normalize(dom);
                    R lr = ccd(err, qa, qb, mia, mib, Ia, Ib, {0, 0, 0}, {0, 0, 0}, -rn, rn, rigid_relax, dt);
                    lr = gmax(lr, -ln * mur);
                    ada = ada - rn * lr; adb = adb + rn * lr;
                }
            }
            if (ba >= 0) {
                R* d8 = &b_delta[8 * ba];
                d8[0] += lda.x; d8[1] += lda.y; d8[2] += lda.z; d8[4] += ada.x; d8[5] += ada.y; d8[6] += ada.z;
            }
            if (bb >= 0) {
                R* d8 = &b_delta[8 * bb];
                d8[0] += ldb.x; d8[1] += ldb.y; d8[2] += ldb.z; d8[4] += adb.x; d8[5] += adb.y; d8[6] += adb.z;
            }
            R* imp = &rc_impulse[6 * ci];
            imp[0] += lda.x; imp[1] += lda.y; imp[2] += lda.z; imp[3] += ada.x; imp[4] += ada.y; imp[5] += ada.z;
        }
    }

    void apply_body_deltas(R dt) {   // apply_body_deltas.comp:58-119
        for (size_t b = meshes.size(); b < b_pos.size(); b++) {
            R im = b_invm[b];
            if (im == 0) continue;
            V3<R> p0 = b_pos[b], v0 = b_v[b], w0 = b_w[b], com = b_com[b];
            std::array<R, 4> q0 = b_rot[b];
            V3<R> dp = {b_delta[8 * b], b_delta[8 * b + 1], b_delta[8 * b + 2]};
            V3<R> dq = {b_delta[8 * b + 4], b_delta[8 * b + 5], b_delta[8 * b + 6]};
            R wgt = 1;
            if (use_weights && b_invw[b] > 0) wgt = R(1) / b_invw[b];
            dp = dp * (im * wgt);
            V3<R> dqw = dq * wgt;
            V3<R> wb = qrot_inv(q0, w0);
            V3<R> dwb = mcol(b_invI[b], qrot_inv(q0, dqw));
            V3<R> tb = cross(dwb, mcol(b_I[b], wb + dwb)) + cross(wb, mcol(b_I[b], dwb));
            V3<R> dw1 = qrot(q0, dwb - dt * mcol(b_invI[b], tb));
            V3<R> dwdt = dw1 * dt;
            std::array<R, 4> wq = qmul({dwdt.x, dwdt.y, dwdt.z, R(0)}, q0);
            std::array<R, 4> q1 = qnormalize({q0[0] + wq[0] * R(0.5), q0[1] + wq[1] * R(0.5), q0[2] + wq[2] * R(0.5),
                                              q0[3] + wq[3] * R(0.5)});
            V3<R> xc = p0 + qrot(q0, com);
            V3<R> p1 = xc + dp * dt;
            p1 = p1 - qrot(q1, com);
            V3<R> v1 = v0 + dp, w1 = w0 + dw1;
            if (length(v1) < R(1e-4f)) v1 = {0, 0, 0};
            if (length(w1) < R(1e-4f)) w1 = {0, 0, 0};
            b_pos[b] = p1; b_rot[b] = q1; b_v[b] = v1; b_w[b] = w1;
        }
    }

    static int cell_index(int x, int y, int z) {   // solve_particle_particle_contacts.comp:66-79
        x += GRID_ORIGIN; y += GRID_ORIGIN; z += GRID_ORIGIN;
        x = std::max(0, x); y = std::max(0, y); z = std::max(0, z);
        return (z % GRID_DIM) * GRID_DIM * GRID_DIM + (y % GRID_DIM) * GRID_DIM + (x % GRID_DIM);
    }

    // solve_particle_particle_contacts.comp:107-218
    void pp_solve(const std::vector<V3<R>>& X, const std::vector<V3<R>>& Vv, R dt) {
        R kmu = R(cfg.particle_mu), coh = R(cfg.particle_cohesion), maxr = R(cfg.particle_max_radius);
        R relax = relax_r;
        R inv = R(1.0f / cfg.search_radius);
        // each sorted slot adds only to its own particle's delta; the taped
        // pair list is gathered per chunk and joined in slot order
        constexpr int64_t kChunk = 2048;
        std::vector<std::vector<std::pair<int32_t, int32_t>>> chunk_pairs(pair_log ? size_t((n + kChunk - 1) / kChunk) : 0);
        par_chunks(n, kChunk, [&](int64_t t0, int64_t t1, int64_t ch) {
        std::vector<std::pair<int32_t, int32_t>>* plog = pair_log ? &chunk_pairs[size_t(ch)] : nullptr;
        for (int64_t tid = t0; tid < t1; tid++) {
            int i = int(svals[tid]);
            int pf = flags[i];
            if ((pf & 1) == 0) continue;
            int proxy = pf & 2;
            V3<R> x = X[i], v = Vv[i];
            R rad = radius[i], w1 = invm[i];
            R Rq = rad + maxr + coh;
            int xs = int(std::floor((x.x - Rq) * inv)), ys = int(std::floor((x.y - Rq) * inv)), zs = int(std::floor((x.z - Rq) * inv));
            int xe = std::min(int(std::floor((x.x + Rq) * inv)), xs + GRID_DIM - 1);
            int ye = std::min(int(std::floor((x.y + Rq) * inv)), ys + GRID_DIM - 1);
            int ze = std::min(int(std::floor((x.z + Rq) * inv)), zs + GRID_DIM - 1);
            V3<R> acc = {0, 0, 0};
            for (int cz = zs; cz <= ze; cz++)
                for (int cy = ys; cy <= ye; cy++)
                    for (int cx = xs; cx <= xe; cx++) {
                        int cell = cell_index(cx, cy, cz);
                        for (int k = cstart[cell]; k < cend[cell]; k++) {
                            int j = int(svals[k]);
                            if (j == i) continue;
                            int nf = flags[j];
                            if ((nf & 1) == 0) continue;
                            if (proxy != 0 && (((nf & 2) != 0) || (invm[j] <= 0))) continue;
                            V3<R> rd = x - X[j];
                            if (std::fabs(rd.x) > Rq || std::fabs(rd.y) > Rq || std::fabs(rd.z) > Rq) continue;
                            R d = exact_length(rd);
                            R err = d - rad - radius[j];
                            if (err > coh || d <= 0) continue;
                            R w2 = invm[j];
                            R denom = w1 + w2;
                            if (denom <= 0) continue;
                            V3<R> nn = exact_div(rd, d);
                            V3<R> vrel = v - Vv[j];
                            R lambda_n = err;
                            R vn = exact_dot(nn, vrel);
                            V3<R> vt = {fmaR(vn, -nn.x, vrel.x), fmaR(vn, -nn.y, vrel.y), fmaR(vn, -nn.z, vrel.z)};
                            R vt_len = exact_length(vt);
                            R lambda_f = gmax(kmu * lambda_n, -vt_len * dt);
                            V3<R> vt_dir = vt_len > 0 ? exact_div(vt, vt_len) : V3<R>{0, 0, 0};
                            V3<R> p2 = nn * err;
                            V3<R> df = {fmaR(vt_dir.x, lambda_f, -p2.x), fmaR(vt_dir.y, lambda_f, -p2.y),
                                        fmaR(vt_dir.z, lambda_f, -p2.z)};
                            acc = acc + exact_div(df, denom);
                            if (plog) plog->push_back({int32_t(i), int32_t(j)});
                        }
                    }
            V3<R> dw = {fmaR(acc.x, w1, R(0)), fmaR(acc.y, w1, R(0)), fmaR(acc.z, w1, R(0))};
            V3<R> tot = {fmaR(dw.x, relax, R(0)), fmaR(dw.y, relax, R(0)), fmaR(dw.z, relax, R(0))};
            delta[i * 3 + 0] += tot.x;
            delta[i * 3 + 1] += tot.y;
            delta[i * 3 + 2] += tot.z;
        }
        });
        for (const auto& v : chunk_pairs) pair_log->insert(pair_log->end(), v.begin(), v.end());
    }

    // apply_particle_deltas.comp:38-58 (inactive: passthrough of x_pred / v_pred)
    void apply(const std::vector<V3<R>>& Xp, const std::vector<V3<R>>& Vp, std::vector<V3<R>>& Xo, std::vector<V3<R>>& Vo, R dt) {
        R vmax = R(cfg.particle_v_max);
        par_chunks(n, 16384, [&](int64_t i0, int64_t i1, int64_t) {
        for (int64_t i = i0; i < i1; i++) {
            if ((flags[i] & 1) == 0) { Xo[i] = Xp[i]; Vo[i] = Vp[i]; continue; }
            V3<R> x0 = q_init[i], xp = Xp[i];
            V3<R> d = {delta[i * 3], delta[i * 3 + 1], delta[i * 3 + 2]};
            V3<R> xn = xp + d;
            V3<R> vn = apply_exact_div(xn - x0, dt);
            R mag = length(vn);
            if (mag > vmax) {
                vn = vn * (vmax / mag);
                xn = x0 + vn * dt;
            }
            Xo[i] = xn; Vo[i] = vn;
        }
        });
    }

    // apply_particle_shape_restitution.comp:48-116.  The reference reads and
    // atomically updates the same qd buffer; here sequential in contact order.
    // Previous body pose / velocity are the current ones (scene.cpp:570-574
    // binds the same buffers twice).
    void restitution() {
        int32_t count = std::min(int32_t(cmax), cc);
        R ka = R(cfg.particle_adhesion), rest = R(cfg.soft_contact_restitution);
        for (int32_t t = 0; t < count; t++) {
            int s = cshape[t], body = shape_body[s], pi = cpart[t];
            if ((flags[pi] & 1) == 0) continue;
            V3<R> vnew = qd[pi], px = q_init[pi], vold = qd_init[pi];
            Frame F = frame(body);
            V3<R> bx = F.pos + quat_rotate(F.rot, cbpos[t]);
            V3<R> nn = cnorm[t];
            V3<R> dp = px - bx;
            R c = fmaR(nn.z, dp.z, fmaR(nn.x, dp.x, nn.y * dp.y)) - radius[pi];
            if (c > ka) continue;
            V3<R> bxp = F.pos + quat_rotate(F.rot, cbpos[t]);
            V3<R> r = bxp - (F.pos + quat_rotate(F.rot, F.com));
            V3<R> bvc = quat_rotate(F.rot, cbvel[t]);
            V3<R> bvo = bvc, bvn = bvc;
            if (body >= 0) {
                bvo = (F.v + cross(F.w, r)) + bvc;
                bvn = (F.v + cross(F.w, r)) + bvc;
            }
            V3<R> dvo = vold - bvo;
            R rvo = fmaR(nn.z, dvo.z, fmaR(nn.y, dvo.y, nn.x * dvo.x));
            V3<R> dvn = vnew - bvn;
            R rvn = fmaR(nn.z, dvn.z, fmaR(nn.x, dvn.x, nn.y * dvn.y));
            if (rvo < 0) {
                R rt = gmax((-rest) * rvo, R(0));
                R sc = (-rvn) + rt;
                qd[pi] = qd[pi] + V3<R>{nn.x * sc, nn.y * sc, nn.z * sc};
            }
        }
    }

    bool has_dynamic_bodies() const { return !boxes.empty(); }

    void iterations_and_finish(R sdt, int iters, bool pingpong) {
        for (int it = 0; it < iters; it++) {
            bool flip = pingpong && (it % 2 != 0);
            std::vector<V3<R>>& X = flip ? q_out : q;
            std::vector<V3<R>>& Vv = flip ? qd_out : qd;
            std::vector<V3<R>>& Xo = flip ? q : q_out;
            std::vector<V3<R>>& Vo = flip ? qd : qd_out;
            std::fill(delta.begin(), delta.end(), R(0));
            std::fill(b_delta.begin(), b_delta.end(), R(0));
            std::fill(b_invw.begin(), b_invw.end(), R(0));
            shape_solve(X, Vv, sdt);
            pp_solve(X, Vv, sdt);
            apply(X, Vv, Xo, Vo, sdt);
            if (!pingpong) { q = q_out; qd = qd_out; }
            if (has_dynamic_bodies()) {
                body_contact_solve(sdt);
                apply_body_deltas(sdt);
            }
        }
        if (pingpong && iters % 2 != 0) { q = q_out; qd = qd_out; }
        if (cfg.enable_restitution) restitution();
    }

    void generate_contacts(bool rigid) {
        cc = 0;
        if (cfg.has_ground_plane) plane_contacts();
        mesh_contacts();
        particle_box_contacts();
        if (rigid) {
            body_plane_contacts();
            std::fill(rc_impulse.begin(), rc_impulse.end(), R(0));
        }
    }

    // scene.cpp:698-1092, extended with the rigid-body stages (builder order:
    // rigid contacts are generated with the particle contacts; bodies are
    // integrated after the particles; each iteration solves rigid contacts and
    // applies body deltas after the particle apply).
    void step(float dt, int substeps, int iters) {
        if (!finalized) finalize_meshes();
        R sdt = R(dt / float(substeps));
        for (int sub = 0; sub < substeps; sub++) {
            generate_contacts(true);
            integrate(sdt);
            integrate_bodies(sdt);
            grid_and_sort();
            iterations_and_finish(sdt, iters, true);
        }
        last_count = std::min<int32_t>(cc, int32_t(cmax));
    }

    void step_solve_only(float dt, int substeps, int iters) {   // scene.cpp:1094-1227
        if (!finalized) finalize_meshes();
        R sdt = R(dt / float(substeps));
        for (int sub = 0; sub < substeps; sub++) {
            integrate(sdt);
            integrate_bodies(sdt);
            grid_and_sort();
            iterations_and_finish(sdt, iters, false);
        }
    }


    // ---- differentiable rollout (builder extension; no reference counterpart) ----
    // A rollout is `substeps` substeps of length sdt with one iteration each,
    // recorded on a tape: per substep the state at its start (x0, v0), the
    // predicted state (x1, v1), the contacts generated on x0, the contributing
    // particle pairs and the final deltas.  The adjoint treats the contact
    // set and each contact's body point/normal/velocity as constants (the
    // contact set is frozen from the forward pass) and differentiates every
    // other operation of integrate, the shape and particle-particle
    // projections and the delta apply, in double precision.
    struct SubTape {
        std::vector<V3<R>> x0, v0, x1, v1;
        std::vector<R> delta;
        std::vector<int32_t> cpart, cshape;
        std::vector<V3<R>> cbpos, cbvel, cnorm;
        std::vector<std::pair<int32_t, int32_t>> pairs;
    };
    std::vector<SubTape> tape;
    R tape_dt = R(0);

    void rollout(const double* v0, double relax, R sdt, int substeps, bool keep_tape) {
        if (!finalized) finalize_meshes();
        if (!boxes.empty()) throw std::runtime_error("rollout: dynamic bodies are not differentiated");
        if (cfg.enable_restitution) throw std::runtime_error("rollout: restitution is not differentiated");
        relax_r = R(relax);
        if (v0) for (int64_t i = 0; i < n; i++) qd[i] = {R(v0[3 * i]), R(v0[3 * i + 1]), R(v0[3 * i + 2])};
        tape.clear();
        tape_dt = sdt;
        for (int sub = 0; sub < substeps; sub++) {
            generate_contacts(false);
            SubTape T;
            if (keep_tape) {
                T.x0 = q; T.v0 = qd;
                const int32_t c = std::min<int32_t>(cc, int32_t(cmax));
                T.cpart.assign(cpart.begin(), cpart.begin() + c);
                T.cshape.assign(cshape.begin(), cshape.begin() + c);
                T.cbpos.assign(cbpos.begin(), cbpos.begin() + c);
                T.cbvel.assign(cbvel.begin(), cbvel.begin() + c);
                T.cnorm.assign(cnorm.begin(), cnorm.begin() + c);
            }
            integrate(sdt);
            if (keep_tape) { T.x1 = q; T.v1 = qd; }
            grid_and_sort();
            pair_log = keep_tape ? &T.pairs : nullptr;
            iterations_and_finish(sdt, 1, true);
            pair_log = nullptr;
            if (keep_tape) { T.delta = delta; tape.push_back(std::move(T)); }
        }
        last_count = std::min<int32_t>(cc, int32_t(cmax));
    }

    using D3 = V3<double>;
    static D3 dv(V3<R> a) { return {double(a.x), double(a.y), double(a.z)}; }
    static D3 perp(D3 u, D3 nn) { return u - nn * dot(nn, u); }   // (I - n n^T) u

    // Friction-and-normal term g(rd, vrel) of a particle pair
    // (solve_particle_particle_contacts.comp:170-205) and its vector-Jacobian
    // product with cotangent C.  Returns g.
    static D3 pair_vjp(D3 rd, D3 vrel, double ri, double rj, double kmu, double h, D3 C, D3& g_rd, D3& g_vrel) {
        const double d = length(rd);
        const D3 nn = rd * (1.0 / d);
        const double err = d - ri - rj;
        const double vn = dot(nn, vrel);
        const D3 vt = vrel - nn * vn;
        const double L = length(vt);
        const double la = kmu * err, lb = -L * h;
        const D3 pc = perp(C, nn);
        const double nc = dot(nn, C);
        if (!(la < lb)) {   // friction bounded by the cone: lambda_f = kmu * err
            const D3 t = L > 0 ? vt * (1.0 / L) : D3{0, 0, 0};
            g_rd = nn * (kmu * dot(t, C) - nc) - pc * (err / d);
            g_vrel = {0, 0, 0};
            if (L > 0) {
                const D3 u = (C - t * dot(t, C)) * (kmu * err / L);
                const D3 pu = perp(u, nn);
                g_vrel = pu;
                g_rd = g_rd - (vt * dot(nn, u) + pu * vn) * (1.0 / d);
            }
            return t * la - nn * err;
        }
        g_vrel = pc * (-h);
        g_rd = (vt * nc + pc * vn) * (h / d) - nn * nc - pc * (err / d);
        return vt * (-h) - nn * err;
    }

    // Shape-contact term (solve_particle_shape_contacts.comp:143-180) with a
    // constant normal: c = n.(x1 - bx) - r, v = v1 - bv.  VJP with cotangent C.
    static D3 shape_vjp(double c, D3 v, D3 nn, double mu, double h, D3 C, D3& g_x, D3& g_v) {
        const double vn = dot(nn, v);
        const D3 vt = v - nn * vn;
        const double L = length(vt);
        const double la = mu * c, lb = -L * h;
        if (!(la < lb)) {
            const D3 t = L > 0 ? vt * (1.0 / L) : D3{0, 0, 0};
            g_x = nn * (mu * dot(t, C) - dot(nn, C));
            g_v = L > 0 ? perp((C - t * dot(t, C)) * (mu * c / L), nn) : D3{0, 0, 0};
            return t * la - nn * c;
        }
        g_x = nn * (-dot(nn, C));
        g_v = perp(C, nn) * (-h);
        return vt * (-h) - nn * c;
    }

    // Reverse pass over the tape: cotangents of the final positions gx and
    // velocities gv (3N doubles each, gv may be null) -> cotangents of the
    // initial positions and velocities and of the contact relaxation.
    void rollout_backward(const double* gx, const double* gv, double* gx0, double* gv0, double* grelax) {
        if (tape.empty()) throw std::runtime_error("rollout_backward: no taped rollout");
        const double h = double(tape_dt), vmax = double(cfg.particle_v_max);
        const double relax = double(relax_r), kmu = double(cfg.particle_mu), ka = double(cfg.particle_adhesion);
        const double pmu = double(cfg.soft_contact_mu);
        const D3 grav = {double(cfg.gravity[0]), double(cfg.gravity[1]), double(cfg.gravity[2])};
        std::vector<D3> lx(n), lv(n), a(n), lx1(n), lv1(n);
        for (int64_t i = 0; i < n; i++) {
            lx[i] = {gx[3 * i], gx[3 * i + 1], gx[3 * i + 2]};
            lv[i] = gv ? D3{gv[3 * i], gv[3 * i + 1], gv[3 * i + 2]} : D3{0, 0, 0};
        }
        double gr = 0.0;
        const R vmax_r = R(cfg.particle_v_max);
        for (size_t s = tape.size(); s-- > 0;) {
            const SubTape& T = tape[s];
            std::vector<D3> lx0(n);
            // apply (apply_particle_deltas.comp:44-57): x2 = xn or clamped
            for (int64_t i = 0; i < n; i++) {
                if ((flags[i] & 1) == 0) { a[i] = {0, 0, 0}; continue; }
                const V3<R> dl = {T.delta[3 * i], T.delta[3 * i + 1], T.delta[3 * i + 2]};
                const V3<R> xn = T.x1[i] + dl;
                const V3<R> vr = apply_exact_div(xn - T.x0[i], R(tape_dt));
                if (length(vr) > vmax_r) {
                    const D3 u = (dv(xn) - dv(T.x0[i])) * (1.0 / h);
                    const double m = length(u);
                    const D3 uh = u * (1.0 / m);
                    const D3 lt = lv[i] + lx[i] * h;
                    const D3 lu = (lt - uh * dot(uh, lt)) * (vmax / m);
                    a[i] = lu * (1.0 / h);
                    lx0[i] = lx[i] - lu * (1.0 / h);
                } else {
                    a[i] = lx[i] + lv[i] * (1.0 / h);
                    lx0[i] = lv[i] * (-1.0 / h);
                }
            }
            for (int64_t i = 0; i < n; i++) { lx1[i] = a[i]; lv1[i] = {0, 0, 0}; }
            // shape contacts
            for (size_t t = 0; t < T.cpart.size(); t++) {
                const int si = T.cshape[t], body = shape_body[si], pi = T.cpart[t];
                const int pf = flags[pi];
                if ((pf & 1) == 0) continue;
                if (pf & 2) {
                    if (body < 0) continue;
                    if (b_flags[body] & 4) continue;
                    if (b_invm[body] == 0) continue;
                }
                Frame F = frame(body);
                const V3<R> bx = F.pos + quat_rotate(F.rot, T.cbpos[t]);
                const V3<R> rr = bx - (F.pos + quat_rotate(F.rot, F.com));
                const V3<R> nnr = T.cnorm[t];
                const V3<R> dp = T.x1[pi] - bx;
                const R cr = fmaR(nnr.z, dp.z, fmaR(nnr.x, dp.x, nnr.y * dp.y)) - radius[pi];
                if (cr > R(ka)) continue;
                double w2 = 0;
                if (body >= 0) {
                    const V3<R> ang = cross(rr, nnr);
                    const V3<R> ra = quat_rotate_inv(F.rot, ang);
                    const std::array<R, 9>& I = b_invI[body];
                    const V3<R> Ira = {I[0] * ra.x + I[1] * ra.y + I[2] * ra.z, I[3] * ra.x + I[4] * ra.y + I[5] * ra.z,
                                       I[6] * ra.x + I[7] * ra.y + I[8] * ra.z};
                    w2 = double(b_invm[body] + dot(ra, Ira));
                }
                const double w1 = double(invm[pi]);
                if (w1 + w2 == 0) continue;
                const V3<R> bvr = F.v + cross(F.w, rr) + quat_rotate(F.rot, T.cbvel[t]);
                const D3 nn = dv(nnr);
                const double c = dot(nn, dv(T.x1[pi]) - dv(bx)) - double(radius[pi]);
                const D3 v = dv(T.v1[pi]) - dv(bvr);
                const double mu = 0.5 * (pmu + double(mat_mu[si]));
                const double sc = w1 / (w1 + w2);
                D3 g_x, g_v;
                const D3 g = shape_vjp(c, v, nn, mu, h, a[pi] * (sc * relax), g_x, g_v);
                lx1[pi] = lx1[pi] + g_x;
                lv1[pi] = lv1[pi] + g_v;
                gr += dot(a[pi], g) * sc;
            }
            // particle pairs (scatter form of the gather in the forward)
            for (const auto& pr : T.pairs) {
                const int i = pr.first, j = pr.second;
                const double wi = double(invm[i]), wj = double(invm[j]);
                const double sc = wi / (wi + wj);
                D3 g_rd, g_vr;
                const D3 g = pair_vjp(dv(T.x1[i]) - dv(T.x1[j]), dv(T.v1[i]) - dv(T.v1[j]), double(radius[i]),
                                      double(radius[j]), kmu, h, a[i] * (sc * relax), g_rd, g_vr);
                lx1[i] = lx1[i] + g_rd; lx1[j] = lx1[j] - g_rd;
                lv1[i] = lv1[i] + g_vr; lv1[j] = lv1[j] - g_vr;
                gr += dot(a[i], g) * sc;
            }
            // integrate (integrate_particles.comp:44-57)
            for (int64_t i = 0; i < n; i++) {
                if ((flags[i] & 1) == 0) continue;   // passes through unchanged
                lx0[i] = lx0[i] + lx1[i];
                const D3 lvv = lv1[i] + lx1[i] * h;
                const R gs = invm[i] > 0 ? R(1) : R(0);
                const V3<R> gr_ = {R(cfg.gravity[0]), R(cfg.gravity[1]), R(cfg.gravity[2])};
                const V3<R> w = T.v0[i] + (V3<R>{0, 0, 0} * invm[i] + gr_ * gs) * R(tape_dt);
                if (length(w) > vmax_r) {
                    const D3 wd = dv(T.v0[i]) + grav * (double(gs) * h);
                    const double m = length(wd);
                    const D3 wh = wd * (1.0 / m);
                    lv[i] = (lvv - wh * dot(wh, lvv)) * (vmax / m);
                } else {
                    lv[i] = lvv;
                }
                lx[i] = lx0[i];
            }
        }
        for (int64_t i = 0; i < n; i++) {
            if (gx0) { gx0[3 * i] = lx[i].x; gx0[3 * i + 1] = lx[i].y; gx0[3 * i + 2] = lx[i].z; }
            if (gv0) { gv0[3 * i] = lv[i].x; gv0[3 * i + 1] = lv[i].y; gv0[3 * i + 2] = lv[i].z; }
        }
        if (grelax) *grelax = gr;
    }

    void collide_only() {   // scene.cpp:1229-1245
        if (!finalized) finalize_meshes();
        generate_contacts(false);
        last_count = std::min<int32_t>(cc, int32_t(cmax));
    }
};

struct Handle {
    int fp64;
    Solver<float> f;
    Solver<double> d;
    std::string err;
};

}  // namespace orc

using orc::Handle;
using orc::OrcConfig;

#define DISPATCH(h, call) ((h)->fp64 ? (h)->d.call : (h)->f.call)

template <class F> static int guarded(Handle* h, F&& fn) {
    try { fn(); return 0; } catch (const std::exception& e) { h->err = e.what(); return 1; }
}

template <class R> static void get_v3(const std::vector<orc::V3<R>>& v, double* out, int64_t n) {
    for (int64_t i = 0; i < n; i++) { out[i * 3] = double(v[i].x); out[i * 3 + 1] = double(v[i].y); out[i * 3 + 2] = double(v[i].z); }
}

extern "C" {

void* orc_create(int fp64) { Handle* h = new Handle(); h->fp64 = fp64; return h; }
void orc_destroy(void* p) { delete static_cast<Handle*>(p); }
const char* orc_last_error(void* p) { return static_cast<Handle*>(p)->err.c_str(); }

int orc_init(void* p, const OrcConfig* c, int64_t n, int64_t mc) {
    Handle* h = static_cast<Handle*>(p);
    return guarded(h, [&] { DISPATCH(h, init(*c, n, mc)); });
}
int orc_add_mesh(void* p, const float* v, int64_t nv, const uint32_t* t, int64_t nt, float margin, int wt) {
    Handle* h = static_cast<Handle*>(p);
    return guarded(h, [&] { DISPATCH(h, add_mesh(v, nv, t, nt, margin, wt)); });
}
int orc_finalize_meshes(void* p) {
    Handle* h = static_cast<Handle*>(p);
    return guarded(h, [&] { DISPATCH(h, finalize_meshes()); });
}
int orc_update_mesh(void* p, int64_t mi, const float* v, int64_t nv, const float* vel) {
    Handle* h = static_cast<Handle*>(p);
    return guarded(h, [&] { DISPATCH(h, update_mesh(size_t(mi), v, nv, vel)); });
}
int orc_set_particles(void* p, const float* q, const float* qd, const float* w, const float* r, const int32_t* f, int64_t n) {
    Handle* h = static_cast<Handle*>(p);
    return guarded(h, [&] { DISPATCH(h, set_particles(q, qd, w, r, f, n)); });
}
int orc_step(void* p, float dt, int s, int it) {
    Handle* h = static_cast<Handle*>(p);
    return guarded(h, [&] { DISPATCH(h, step(dt, s, it)); });
}
int orc_step_solve_only(void* p, float dt, int s, int it) {
    Handle* h = static_cast<Handle*>(p);
    return guarded(h, [&] { DISPATCH(h, step_solve_only(dt, s, it)); });
}
int orc_collide_only(void* p) {
    Handle* h = static_cast<Handle*>(p);
    return guarded(h, [&] { DISPATCH(h, collide_only()); });
}


// Differentiable rollout (builder extension): v0 3N doubles (NULL keeps the
// current velocities), relax the contact relaxation, `substeps` substeps of
// sdt with one iteration each; keep_tape records what rollout_backward needs.
int orc_rollout(void* p, const double* v0, double relax, double sdt, int substeps, int keep_tape) {
    Handle* h = static_cast<Handle*>(p);
    return guarded(h, [&] {
        if (h->fp64) h->d.rollout(v0, relax, double(float(sdt)), substeps, keep_tape != 0);
        else h->f.rollout(v0, relax, float(sdt), substeps, keep_tape != 0);
    });
}
int orc_rollout_backward(void* p, const double* gx, const double* gv, double* gx0, double* gv0, double* grelax) {
    Handle* h = static_cast<Handle*>(p);
    return guarded(h, [&] { DISPATCH(h, rollout_backward(gx, gv, gx0, gv0, grelax)); });
}

// Getters return doubles (exact for the fp32 state) so one ctypes signature
// serves both precisions.
void orc_get_positions(void* p, double* out) { Handle* h = static_cast<Handle*>(p); if (h->fp64) get_v3(h->d.q, out, h->d.n); else get_v3(h->f.q, out, h->f.n); }
void orc_get_velocities(void* p, double* out) { Handle* h = static_cast<Handle*>(p); if (h->fp64) get_v3(h->d.qd, out, h->d.n); else get_v3(h->f.qd, out, h->f.n); }
void orc_get_deltas(void* p, double* out) {
    Handle* h = static_cast<Handle*>(p);
    if (h->fp64) for (int64_t i = 0; i < 3 * h->d.n; i++) out[i] = h->d.delta[i];
    else for (int64_t i = 0; i < 3 * h->f.n; i++) out[i] = h->f.delta[i];
}
int32_t orc_contact_count(void* p) { Handle* h = static_cast<Handle*>(p); return h->fp64 ? h->d.last_count : h->f.last_count; }
void orc_get_contacts(void* p, int32_t* part, int32_t* shape, double* bpos, double* bvel, double* nrm) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& s) {
        for (int32_t k = 0; k < s.last_count; k++) {
            part[k] = s.cpart[k]; shape[k] = s.cshape[k];
            bpos[k * 3] = s.cbpos[k].x; bpos[k * 3 + 1] = s.cbpos[k].y; bpos[k * 3 + 2] = s.cbpos[k].z;
            bvel[k * 3] = s.cbvel[k].x; bvel[k * 3 + 1] = s.cbvel[k].y; bvel[k * 3 + 2] = s.cbvel[k].z;
            nrm[k * 3] = s.cnorm[k].x; nrm[k * 3 + 1] = s.cnorm[k].y; nrm[k * 3 + 2] = s.cnorm[k].z;
        }
    };
    if (h->fp64) go(h->d); else go(h->f);
}
int orc_set_contact_data(void* p, const int32_t* part, const int32_t* shape, const float* bpos, const float* nrm, int32_t count) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& s) {
        using RR = decltype(s.cbpos[0].x);
        s.last_count = count;
        s.cc = count;
        for (int32_t k = 0; k < count && k < s.cmax; k++) {
            s.cpart[k] = part[k]; s.cshape[k] = shape[k];
            s.cbpos[k] = {RR(bpos[k * 3]), RR(bpos[k * 3 + 1]), RR(bpos[k * 3 + 2])};
            s.cnorm[k] = {RR(nrm[k * 3]), RR(nrm[k * 3 + 1]), RR(nrm[k * 3 + 2])};
        }
    };
    if (h->fp64) go(h->d); else go(h->f);
    return 0;
}
void orc_get_sorted(void* p, uint32_t* keys, uint32_t* vals) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& s) { for (int64_t i = 0; i < s.n; i++) { keys[i] = s.skeys[i]; vals[i] = s.svals[i]; } };
    if (h->fp64) go(h->d); else go(h->f);
}
void orc_get_cell_tables(void* p, int32_t* starts, int32_t* ends) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& s) { std::memcpy(starts, s.cstart.data(), 4 * orc::NUM_CELLS); std::memcpy(ends, s.cend.data(), 4 * orc::NUM_CELLS); };
    if (h->fp64) go(h->d); else go(h->f);
}
int64_t orc_mesh_node_count(void* p, int64_t mi) { Handle* h = static_cast<Handle*>(p); auto& s = h->f; return mi < int64_t(s.bvh.size()) ? int64_t(s.bvh[mi].lowers.size()) : 0; }
void orc_get_mesh_bvh(void* p, int64_t mi, float* lowers, float* uppers, uint32_t* prim, uint32_t* root) {
    Handle* h = static_cast<Handle*>(p);
    auto& s = h->f;
    const orc::Lbvh& b = s.bvh[mi];
    std::memcpy(lowers, b.lowers.data(), 16 * b.lowers.size());
    std::memcpy(uppers, b.uppers.data(), 16 * b.uppers.size());
    std::memcpy(prim, b.prim.data(), 4 * b.prim.size());
    *root = b.root;
}
int orc_add_box_body(void* p, const float* pos, const float* rot, const float* half, float density, int32_t flags) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& s) {
        typename std::remove_reference<decltype(s)>::type::Box b;
        for (int k = 0; k < 3; k++) { b.pos[k] = pos[k]; b.half[k] = half[k]; }
        for (int k = 0; k < 4; k++) b.rot[k] = rot[k];
        b.density = density; b.flags = flags;
        s.boxes.push_back(b);
        s.finalized = false;
    };
    if (h->fp64) go(h->d); else go(h->f);
    return 0;
}
void orc_set_body_params(void* p, float angular_damping, float rigid_relax, float mu_t, float mu_r, int use_weights) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& s) {
        using RR = decltype(s.rigid_relax);
        s.angular_damping = RR(angular_damping); s.rigid_relax = RR(rigid_relax);
        s.mu_torsional = RR(mu_t); s.mu_rolling = RR(mu_r); s.use_weights = use_weights;
        for (auto& x : s.mat_mu_t) x = RR(mu_t);
        for (auto& x : s.mat_mu_r) x = RR(mu_r);
    };
    if (h->fp64) go(h->d); else go(h->f);
}
// dynamic bodies only (boxes), in add order: q = 7 (pos, quat xyzw), qd = 6 (v, w)
int64_t orc_get_bodies(void* p, double* q7, double* qd6) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& s) -> int64_t {
        size_t M = s.meshes.size();
        int64_t nb = int64_t(s.boxes.size());
        if (s.b_pos.size() < M + s.boxes.size()) return 0;
        for (int64_t k = 0; k < nb; k++) {
            size_t b = M + size_t(k);
            if (q7) {
                q7[7 * k] = s.b_pos[b].x; q7[7 * k + 1] = s.b_pos[b].y; q7[7 * k + 2] = s.b_pos[b].z;
                for (int c = 0; c < 4; c++) q7[7 * k + 3 + c] = s.b_rot[b][c];
            }
            if (qd6) {
                qd6[6 * k] = s.b_v[b].x; qd6[6 * k + 1] = s.b_v[b].y; qd6[6 * k + 2] = s.b_v[b].z;
                qd6[6 * k + 3] = s.b_w[b].x; qd6[6 * k + 4] = s.b_w[b].y; qd6[6 * k + 5] = s.b_w[b].z;
            }
        }
        return nb;
    };
    return h->fp64 ? go(h->d) : go(h->f);
}
void orc_set_body_velocities(void* p, const float* qd6) {
    Handle* h = static_cast<Handle*>(p);
    auto go = [&](auto& s) {
        using RR = decltype(s.rigid_relax);
        size_t M = s.meshes.size();
        for (size_t k = 0; k < s.boxes.size(); k++) {
            s.b_v[M + k] = {RR(qd6[6 * k]), RR(qd6[6 * k + 1]), RR(qd6[6 * k + 2])};
            s.b_w[M + k] = {RR(qd6[6 * k + 3]), RR(qd6[6 * k + 4]), RR(qd6[6 * k + 5])};
        }
    };
    if (h->fp64) go(h->d); else go(h->f);
}
int orc_mesh_watertight(void* p, int64_t mi) { Handle* h = static_cast<Handle*>(p); return h->f.meshes[mi].watertight ? 1 : 0; }

// ---- stage-level entry points (stateless), used by the pinning tests ----
void orc_lbvh_build(const float* pts, const uint32_t* tris, uint32_t nt, float* lowers, float* uppers, uint32_t* prim, uint32_t* root) {
    orc::Lbvh b;
    orc::lbvh_build(pts, tris, nt, b);
    std::memcpy(lowers, b.lowers.data(), 16 * b.lowers.size());
    std::memcpy(uppers, b.uppers.data(), 16 * b.uppers.size());
    std::memcpy(prim, b.prim.data(), 4 * b.prim.size());
    *root = b.root;
}
int orc_ray_tri_woop_f(const float* p, const float* dir, const float* a, const float* b, const float* c) {
    using V = orc::V3<float>;
    return orc::ray_tri_woop<float>(V{p[0], p[1], p[2]}, V{dir[0], dir[1], dir[2]}, V{a[0], a[1], a[2]}, V{b[0], b[1], b[2]},
                                    V{c[0], c[1], c[2]}) ? 1 : 0;
}
// Hits of the reference's fixed-direction parity ray (generate_mesh_contacts.comp:372-388)
// from each point over all triangles: out_hits[k] = number of hits; hit ids
// (point-major) written to ids up to ids_cap.  Returns the total hit count.
int64_t orc_ray_hit_lists(const float* verts, const uint32_t* tris, int64_t nt, const float* pts, int64_t np,
                          int32_t* out_hits, uint32_t* ids, int64_t ids_cap) {
    using V = orc::V3<float>;
    const V dir{1.029700994f, 1.086472273f, 1.034694910f};
    int64_t w = 0;
    for (int64_t k = 0; k < np; k++) {
        const V p{pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]};
        int32_t h = 0;
        for (int64_t t = 0; t < nt; t++) {
            const float* a = verts + 3 * size_t(tris[3 * t]);
            const float* b = verts + 3 * size_t(tris[3 * t + 1]);
            const float* c = verts + 3 * size_t(tris[3 * t + 2]);
            if (orc::ray_tri_woop<float>(p, dir, V{a[0], a[1], a[2]}, V{b[0], b[1], b[2]}, V{c[0], c[1], c[2]})) {
                if (w < ids_cap) ids[w] = uint32_t(t);
                w++;
                h++;
            }
        }
        out_hits[k] = h;
    }
    return w;
}
void orc_closest_point_triangle_f(const float* p, const float* a, const float* b, const float* c, float* cp, float* bary) {
    using V = orc::V3<float>;
    V o, tn, bc;
    orc::closest_point_triangle<float>(V{p[0], p[1], p[2]}, V{a[0], a[1], a[2]}, V{b[0], b[1], b[2]}, V{c[0], c[1], c[2]}, o, tn, bc);
    cp[0] = o.x; cp[1] = o.y; cp[2] = o.z; bary[0] = bc.x; bary[1] = bc.y; bary[2] = bc.z;
}

}  // extern "C"
