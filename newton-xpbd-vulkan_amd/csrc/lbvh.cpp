// lbvh.cpp -- host LBVH construction and refit for the mesh narrowphase.
//
// Construction is the agglomerative bottom-up LBVH (Apetrei 2014) that
// src/mesh_lbvh.h:49-223 uses; the topology it yields depends only on the
// sorted Morton codes, so it is computed here in three passes: (1) codes +
// stable LSD radix sort, (2) parent links by climbing from every leaf with a
// per-node arrival count, (3) boxes and leaf collapsing top-down.
#include "lbvh.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace xpbd {

namespace {

inline uint32_t interleave3(uint32_t v) {   // 10 bits -> every third bit
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0xFF0000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

inline uint32_t quantize1024(float t) {
    int q = int(t * 1024.0f);
    return uint32_t(q < 0 ? 0 : (q > 1023 ? 1023 : q));
}

inline int prefix_len(uint64_t a, uint64_t b) {
    uint64_t x = a ^ b;
    return x ? __builtin_clzll(x) : 64;
}

}  // namespace

void lbvh_build(const float* pts, const uint32_t* tris, uint32_t nt, MeshBvhHost& out) {
    out.lo.assign(size_t(2) * nt, NodeHalf{0.f, 0.f, 0.f, 0u});
    out.hi.assign(size_t(2) * nt, NodeHalf{0.f, 0.f, 0.f, 0u});
    out.prim.assign(nt, 0u);
    out.root = 0;
    if (nt == 0) return;

    // triangle boxes and scene box
    std::vector<float> bmin(size_t(3) * nt), bmax(size_t(3) * nt);
    float smin[3] = {1e30f, 1e30f, 1e30f}, smax[3] = {-1e30f, -1e30f, -1e30f};
    for (uint32_t t = 0; t < nt; t++) {
        for (int ax = 0; ax < 3; ax++) {
            float a = pts[size_t(tris[3 * t]) * 3 + ax], b = pts[size_t(tris[3 * t + 1]) * 3 + ax],
                  c = pts[size_t(tris[3 * t + 2]) * 3 + ax];
            float lo = std::min(std::min(a, b), c), hi = std::max(std::max(a, b), c);
            bmin[3 * t + ax] = lo;
            bmax[3 * t + ax] = hi;
            smin[ax] = std::min(smin[ax], lo);
            smax[ax] = std::max(smax[ax], hi);
        }
    }
    float inv[3];
    for (int ax = 0; ax < 3; ax++) inv[ax] = 1.0f / ((smax[ax] - smin[ax]) + 0.0001f);

    // Morton codes, then a stable 3 x 10-bit LSD counting sort of (code, id)
    std::vector<uint32_t> code(nt), order(nt), tmp(nt);
    for (uint32_t t = 0; t < nt; t++) {
        uint32_t q[3];
        for (int ax = 0; ax < 3; ax++) {
            float c = 0.5f * (bmin[3 * t + ax] + bmax[3 * t + ax]);
            q[ax] = quantize1024((c - smin[ax]) * inv[ax]);
        }
        code[t] = (interleave3(q[2]) << 2) | (interleave3(q[1]) << 1) | interleave3(q[0]);
        order[t] = t;
    }
    std::vector<uint32_t> cnt(1024);
    for (int pass = 0; pass < 3; pass++) {
        const int sh = 10 * pass;
        std::fill(cnt.begin(), cnt.end(), 0u);
        for (uint32_t k = 0; k < nt; k++) cnt[(code[order[k]] >> sh) & 1023u]++;
        uint32_t run = 0;
        for (auto& c : cnt) { uint32_t v = c; c = run; run += v; }
        for (uint32_t k = 0; k < nt; k++) tmp[cnt[(code[order[k]] >> sh) & 1023u]++] = order[k];
        order.swap(tmp);
    }
    std::vector<uint64_t> key(nt);
    for (uint32_t k = 0; k < nt; k++) { key[k] = code[order[k]]; out.prim[k] = order[k]; }

    // gap g separates sorted leaves g and g+1; its internal node is n + g
    const int n = int(nt);
    std::vector<int> gap(n > 1 ? n - 1 : 1, 0);
    for (int g = 0; g + 1 < n; g++) gap[g] = prefix_len(key[g], key[g + 1]);

    const int total = 2 * n;
    std::vector<int> parent(total, -1), left(total, -1), right(total, -1), lo_rng(total, 0), hi_rng(total, 0),
        arrivals(total, 0);
    for (int i = 0; i < n; i++) { lo_rng[i] = hi_rng[i] = i; }
    for (int i = 0; i < n; i++) {
        int node = i;
        while (true) {
            const int l = lo_rng[node], r = hi_rng[node];
            if (l == 0 && r == n - 1) { out.root = uint32_t(node); break; }
            // merge towards the neighbouring gap with the longer common prefix;
            // equal prefixes go by the parity of the two bordering primitive ids
            bool merge_right;
            if (l == 0) merge_right = true;
            else if (r == n - 1) merge_right = false;
            else if (gap[r] > gap[l - 1]) merge_right = true;
            else if (gap[r] < gap[l - 1]) merge_right = false;
            else merge_right = ((out.prim[l - 1] & 1u) != (out.prim[r] & 1u));
            int p;
            if (merge_right) { p = n + r; left[p] = node; lo_rng[p] = l; }
            else { p = n + l - 1; right[p] = node; hi_rng[p] = r; }
            parent[node] = p;
            if (++arrivals[p] == 1) break;   // the sibling is still climbing
            node = p;
        }
    }

    // depths from the root, then boxes bottom-up in reverse BFS order
    std::vector<int> depth(total, 0), bfs;
    bfs.reserve(total);
    bfs.push_back(int(out.root));
    depth[out.root] = 1;
    for (size_t k = 0; k < bfs.size(); k++) {
        int v = bfs[k];
        if (v < n) continue;
        for (int c : {left[v], right[v]}) { depth[c] = depth[v] + 1; bfs.push_back(c); }
    }
    for (int i = 0; i < n; i++) {
        const uint32_t t = out.prim[i];
        out.lo[i] = NodeHalf{bmin[3 * t], bmin[3 * t + 1], bmin[3 * t + 2], uint32_t(i) | (1u << 30)};
        out.hi[i] = NodeHalf{bmax[3 * t], bmax[3 * t + 1], bmax[3 * t + 2], uint32_t(i) + 1u};
    }
    for (size_t k = bfs.size(); k-- > 0;) {
        const int v = bfs[k];
        if (v < n) continue;
        const NodeHalf &a = out.lo[left[v]], &b = out.lo[right[v]], &c = out.hi[left[v]], &d = out.hi[right[v]];
        out.lo[v] = NodeHalf{std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z), uint32_t(left[v])};
        out.hi[v] = NodeHalf{std::max(c.x, d.x), std::max(c.y, d.y), std::max(c.z, d.z), uint32_t(right[v])};
    }
    // collapse small or deep subtrees into packed leaves over [lo, hi)
    for (int v : bfs) {
        const int span = hi_rng[v] + 1 - lo_rng[v];
        if (span <= 4 || depth[v] >= 32) {
            out.lo[v].packed = uint32_t(lo_rng[v]) | (1u << 30);
            out.hi[v].packed = uint32_t(hi_rng[v] + 1);
        }
    }
}

namespace {
void refit_node(MeshBvhHost& b, const float* pts, const uint32_t* tris, uint32_t v) {
    const uint32_t lp = b.lo[v].packed;
    if ((lp >> 30) & 3u) {
        float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
        const uint32_t e = b.hi[v].packed & 0x3FFFFFFFu;
        for (uint32_t k = lp & 0x3FFFFFFFu; k < e; k++) {
            const uint32_t t = b.prim[k];
            for (int c = 0; c < 3; c++)
                for (int ax = 0; ax < 3; ax++) {
                    const float x = pts[size_t(tris[3 * t + c]) * 3 + ax];
                    mn[ax] = std::min(mn[ax], x);
                    mx[ax] = std::max(mx[ax], x);
                }
        }
        b.lo[v].x = mn[0]; b.lo[v].y = mn[1]; b.lo[v].z = mn[2];
        b.hi[v].x = mx[0]; b.hi[v].y = mx[1]; b.hi[v].z = mx[2];
        return;
    }
    const uint32_t l = lp & 0x3FFFFFFFu, r = b.hi[v].packed;
    refit_node(b, pts, tris, l);
    refit_node(b, pts, tris, r);
    b.lo[v].x = std::min(b.lo[l].x, b.lo[r].x);
    b.lo[v].y = std::min(b.lo[l].y, b.lo[r].y);
    b.lo[v].z = std::min(b.lo[l].z, b.lo[r].z);
    b.hi[v].x = std::max(b.hi[l].x, b.hi[r].x);
    b.hi[v].y = std::max(b.hi[l].y, b.hi[r].y);
    b.hi[v].z = std::max(b.hi[l].z, b.hi[r].z);
}
}  // namespace

void lbvh_refit(MeshBvhHost& b, const float* pts, const uint32_t* tris) {
    if (!b.lo.empty()) refit_node(b, pts, tris, b.root);
}

RayAxes parity_axes() {
    const float dir[3] = {1.029700994f, 1.086472273f, 1.034694910f};
    const float ad[3] = {std::fabs(dir[0]), std::fabs(dir[1]), std::fabs(dir[2])};
    int kz = 0;
    if (ad[1] > ad[0]) kz = 1;
    if (ad[2] > (kz == 1 ? ad[1] : ad[0])) kz = 2;
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (dir[kz] < 0.0f) std::swap(kx, ky);
    return RayAxes{kx, ky, kz, dir[kx] / dir[kz], dir[ky] / dir[kz]};
}

void dist_grid_build(const float* pts, const uint32_t* tris, uint32_t nt, float cap, DistGrid& g, const MeshBvhHost* bvh) {
    g = DistGrid{};
    if (nt == 0 || !(cap > 0.0f)) return;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    std::vector<double> tb(size_t(6) * nt);
    for (uint32_t t = 0; t < nt; t++) {
        double a[3] = {1e300, 1e300, 1e300}, b[3] = {-1e300, -1e300, -1e300};
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < 3; k++) {
                const double x = pts[size_t(tris[3 * t + c]) * 3 + k];
                a[k] = std::min(a[k], x);
                b[k] = std::max(b[k], x);
            }
        for (int k = 0; k < 3; k++) {
            tb[6 * t + k] = a[k]; tb[6 * t + 3 + k] = b[k];
            lo[k] = std::min(lo[k], a[k]); hi[k] = std::max(hi[k], b[k]);
        }
    }
    // cell edge: half the cap, coarsened so no axis exceeds 256 cells
    double ext = 0.0;
    for (int k = 0; k < 3; k++) { lo[k] -= cap; hi[k] += cap; ext = std::max(ext, hi[k] - lo[k]); }
    const double h = std::max(double(cap) * 0.5, ext / 256.0);
    const uint32_t n[3] = {uint32_t(std::ceil((hi[0] - lo[0]) / h)) + 1, uint32_t(std::ceil((hi[1] - lo[1]) / h)) + 1,
                           uint32_t(std::ceil((hi[2] - lo[2]) / h)) + 1};
    g.ox = float(lo[0]); g.oy = float(lo[1]); g.oz = float(lo[2]);
    g.inv = float(1.0 / h);
    g.h = h;
    g.cap = cap;
    g.nx = n[0]; g.ny = n[1]; g.nz = n[2];
    // cell (i, j, k) covers [o + c*h', o + (c+1)*h'] with h' = 1/inv as the
    // device computes cells; bounds use that box grown by a relative 1e-6
    // so float rounding of the device's cell index cannot leave the box
    const double o[3] = {double(g.ox), double(g.oy), double(g.oz)};
    const double hc = 1.0 / double(g.inv);
    std::vector<double> best(size_t(n[0]) * n[1] * n[2], double(cap));
    // leaf-order span of the triangles within cap of each cell
    std::vector<uint32_t> pos, span_lo, span_hi;
    if (bvh) {
        pos.assign(nt, 0);
        for (uint32_t k = 0; k < nt; k++) pos[bvh->prim[k]] = k;
        span_lo.assign(best.size(), 0xFFFFFFFFu);
        span_hi.assign(best.size(), 0u);
    }
    for (uint32_t t = 0; t < nt; t++) {
        long c0[3], c1[3];
        for (int k = 0; k < 3; k++) {
            c0[k] = std::max<long>(0, long(std::floor((tb[6 * t + k] - cap - o[k]) / hc)) - 1);
            c1[k] = std::min<long>(long(n[k]) - 1, long(std::floor((tb[6 * t + 3 + k] + cap - o[k]) / hc)) + 1);
        }
        for (long z = c0[2]; z <= c1[2]; z++)
            for (long y = c0[1]; y <= c1[1]; y++)
                for (long x = c0[0]; x <= c1[0]; x++) {
                    const long cc[3] = {x, y, z};
                    double d2 = 0.0;
                    for (int k = 0; k < 3; k++) {
                        const double slack = 1e-6 * (std::fabs(o[k]) + double(n[k]) * hc) + 1e-9;
                        const double blo = o[k] + double(cc[k]) * hc - slack, bhi = o[k] + double(cc[k] + 1) * hc + slack;
                        const double gap = std::max(0.0, std::max(tb[6 * t + k] - bhi, blo - tb[6 * t + 3 + k]));
                        d2 += gap * gap;
                    }
                    const size_t ci = (size_t(z) * n[1] + size_t(y)) * n[0] + size_t(x);
                    const double d = std::sqrt(d2);
                    if (d < best[ci]) best[ci] = d;
                    if (bvh && d <= double(cap)) {
                        span_lo[ci] = std::min(span_lo[ci], pos[t]);
                        span_hi[ci] = std::max(span_hi[ci], pos[t]);
                    }
                }
    }
    if (bvh) {
        // leaf-order range [first, last] covered by every node reachable from the root
        const size_t nn = bvh->lo.size();
        std::vector<uint32_t> first(nn, 0), last(nn, 0);
        std::vector<uint32_t> order;
        order.reserve(nn);
        order.push_back(bvh->root);
        for (size_t k = 0; k < order.size(); k++) {
            const uint32_t v = order[k];
            const uint32_t lp = bvh->lo[v].packed;
            if (!((lp >> 30) & 3u)) { order.push_back(lp & 0x3FFFFFFFu); order.push_back(bvh->hi[v].packed & 0x3FFFFFFFu); }
        }
        for (size_t k = order.size(); k-- > 0;) {
            const uint32_t v = order[k];
            const uint32_t lp = bvh->lo[v].packed;
            if ((lp >> 30) & 3u) { first[v] = lp & 0x3FFFFFFFu; last[v] = (bvh->hi[v].packed & 0x3FFFFFFFu) - 1u; }
            else {
                const uint32_t l = lp & 0x3FFFFFFFu, r = bvh->hi[v].packed & 0x3FFFFFFFu;
                first[v] = std::min(first[l], first[r]);
                last[v] = std::max(last[l], last[r]);
            }
        }
        g.start.assign(best.size(), bvh->root);
        for (size_t ci = 0; ci < best.size(); ci++) {
            if (span_lo[ci] > span_hi[ci]) continue;
            uint32_t v = bvh->root;
            for (;;) {
                const uint32_t lp = bvh->lo[v].packed;
                if ((lp >> 30) & 3u) break;
                const uint32_t l = lp & 0x3FFFFFFFu, r = bvh->hi[v].packed & 0x3FFFFFFFu;
                if (first[l] <= span_lo[ci] && span_hi[ci] <= last[l]) v = l;
                else if (first[r] <= span_lo[ci] && span_hi[ci] <= last[r]) v = r;
                else break;
            }
            g.start[ci] = v;
        }
    }
    g.lb.resize(best.size());
    for (size_t k = 0; k < best.size(); k++) {
        float f = float(best[k]);
        if (double(f) > best[k]) f = std::nextafter(f, 0.0f);   // round down
        g.lb[k] = f;
    }
}

float dist_grid_bound(const DistGrid& g, const float p[3]) {
    const float fx = std::floor((p[0] - g.ox) * g.inv), fy = std::floor((p[1] - g.oy) * g.inv),
                fz = std::floor((p[2] - g.oz) * g.inv);
    if (!(fx >= 0.0f && fy >= 0.0f && fz >= 0.0f && fx < float(g.nx) && fy < float(g.ny) && fz < float(g.nz))) return -1.0f;
    return g.lb[(size_t(fz) * g.ny + size_t(fy)) * g.nx + size_t(fx)];
}

long parity_grid_cell(const ParityGrid& g, const float p[3]) {
    const RayAxes r = parity_axes();
    const float u = p[r.kx] - r.Sx * p[r.kz];
    const float w = p[r.ky] - r.Sy * p[r.kz];
    const float fu = std::floor((u - g.u0) * g.inv), fw = std::floor((w - g.w0) * g.inv);
    if (!(fu >= 0.0f && fw >= 0.0f && fu < float(g.dim) && fw < float(g.dim))) return -1;
    return long(fw) * long(g.dim) + long(fu);
}

void parity_grid_build(const float* pts, const uint32_t* tris, uint32_t nt, uint32_t tri_base, ParityGrid& g) {
    g = ParityGrid{};
    if (nt == 0) return;
    const RayAxes r = parity_axes();
    const int kx = r.kx, ky = r.ky, kz = r.kz;
    const double Sx = double(r.Sx), Sy = double(r.Sy);
    double scale = 0.0;
    std::vector<double> bu0(nt), bu1(nt), bw0(nt), bw1(nt);
    double umin = 1e300, umax = -1e300, wmin = 1e300, wmax = -1e300;
    for (uint32_t t = 0; t < nt; t++) {
        double lo_u = 1e300, hi_u = -1e300, lo_w = 1e300, hi_w = -1e300;
        for (int c = 0; c < 3; c++) {
            const float* v = pts + size_t(tris[3 * t + c]) * 3;
            const double u = double(v[kx]) - Sx * double(v[kz]), w = double(v[ky]) - Sy * double(v[kz]);
            lo_u = std::min(lo_u, u); hi_u = std::max(hi_u, u);
            lo_w = std::min(lo_w, w); hi_w = std::max(hi_w, w);
            for (int k = 0; k < 3; k++) scale = std::max(scale, double(std::fabs(v[k])));
        }
        bu0[t] = lo_u; bu1[t] = hi_u; bw0[t] = lo_w; bw1[t] = hi_w;
    }
    // rounding of the device's projection and of the ray test is a few ulps of
    // the coordinate scale; the pad is orders of magnitude larger
    const double pad = 1e-4 * (1.0 + scale);
    for (uint32_t t = 0; t < nt; t++) {
        bu0[t] -= pad; bu1[t] += pad; bw0[t] -= pad; bw1[t] += pad;
        umin = std::min(umin, bu0[t]); umax = std::max(umax, bu1[t]);
        wmin = std::min(wmin, bw0[t]); wmax = std::max(wmax, bw1[t]);
    }
    uint32_t dim = uint32_t(std::ceil(std::sqrt(double(nt)) / 2.0));
    dim = std::max<uint32_t>(1, std::min<uint32_t>(dim, 2048));
    const double ext = std::max(umax - umin, wmax - wmin);
    const float cell = float(ext / double(dim)) * 1.0001f;
    g.u0 = float(umin);
    g.w0 = float(wmin);
    g.inv = 1.0f / cell;
    g.dim = dim;
    auto cell_of = [&](double x, float o) {
        const long c = long(std::floor((x - double(o)) * double(g.inv)));
        return uint32_t(std::max<long>(0, std::min<long>(long(dim) - 1, c)));
    };
    std::vector<uint32_t> cnt(size_t(dim) * dim + 1, 0);
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) {
            uint32_t run = 0;
            for (auto& c : cnt) { const uint32_t v = c; c = run; run += v; }
            g.start = cnt;
            g.tri.assign(run, 0);
        }
        for (uint32_t t = 0; t < nt; t++) {
            const uint32_t a0 = cell_of(bu0[t], g.u0), a1 = cell_of(bu1[t], g.u0);
            const uint32_t b0 = cell_of(bw0[t], g.w0), b1 = cell_of(bw1[t], g.w0);
            for (uint32_t b = b0; b <= b1; b++)
                for (uint32_t a = a0; a <= a1; a++) {
                    const size_t c = size_t(b) * dim + a;
                    if (pass == 0) cnt[c]++;
                    else g.tri[cnt[c]++] = tri_base + t;
                }
        }
    }
}

}  // namespace xpbd
