// k_meshupd.hip -- animated colliders on the device: after update_mesh
// uploads a mesh's new vertices, these kernels refit its BVH, refresh every
// copy of its node boxes and triangle corners the narrowphase reads, and
// rebuild its distance-bound and parity grids in place, so no buffer is
// reallocated and captured frames stay valid.
//
// The reference refits on the CPU and re-uploads the whole node arrays
// (MeshBvh::refit, src/mesh_bvh.h:142-181, called from scene.cpp:617-696).
#include "xpbd_kernels.h"

namespace xpbd {

__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }
// std::min / std::max as the host refit evaluates them: min(a, b) = b < a ? b
// : a, max(a, b) = a < b ? b : a (the same bits for signed zeros).
__device__ __forceinline__ float hmin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float hmax(float a, float b) { return a < b ? b : a; }

// MeshBvh::refit (mesh_bvh.h:142-181) for the nodes of one height: a leaf's
// box over its triangles' corners, an internal node's box over its
// children's.  Internal nodes also rewrite their walk record (the boxes of
// their children and grandchildren, all final at this height; the stack words
// in the lower w are kept), leaves their box by first primitive position.
__global__ __launch_bounds__(256) void k_refit_level(MeshRefitArgs a, const uint32_t* __restrict__ list, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t g = list[i];
    const float4 L = a.nodes[2 * size_t(g)], U = a.nodes[2 * size_t(g) + 1];
    const uint32_t lw = f2u(L.w), uw = f2u(U.w);
    float mn[3], mx[3];
    if ((lw >> 30) & 3u) {
        mn[0] = mn[1] = mn[2] = 1e30f;
        mx[0] = mx[1] = mx[2] = -1e30f;
        const uint32_t b = lw & 0x3FFFFFFFu, e = uw & 0x3FFFFFFFu;
        for (uint32_t k = b; k < e; k++) {
            const uint32_t t = a.prim[k];
            for (int c = 0; c < 3; c++) {
                const float* p = a.verts + 3 * size_t(a.tris[3 * size_t(t) + c]);
                for (int ax = 0; ax < 3; ax++) {
                    mn[ax] = hmin(mn[ax], p[ax]);
                    mx[ax] = hmax(mx[ax], p[ax]);
                }
            }
        }
        if (b < a.nprim) {
            a.lbox[2 * size_t(b)] = make_float4(mn[0], mn[1], mn[2], 0.f);
            a.lbox[2 * size_t(b) + 1] = make_float4(mx[0], mx[1], mx[2], 0.f);
        }
    } else {
        const uint32_t l = lw & 0x3FFFFFFFu, r = uw & 0x3FFFFFFFu;
        const float4 Ll = a.nodes[2 * size_t(l)], Ul = a.nodes[2 * size_t(l) + 1];
        const float4 Lr = a.nodes[2 * size_t(r)], Ur = a.nodes[2 * size_t(r) + 1];
        mn[0] = hmin(Ll.x, Lr.x); mn[1] = hmin(Ll.y, Lr.y); mn[2] = hmin(Ll.z, Lr.z);
        mx[0] = hmax(Ul.x, Ur.x); mx[1] = hmax(Ul.y, Ur.y); mx[2] = hmax(Ul.z, Ur.z);
        float4* rec = a.rec + 12 * size_t(g);
        const uint32_t ch[2] = {l, r};
        for (int k = 0; k < 2; k++) {
            const float4 CL = a.nodes[2 * size_t(ch[k])], CU = a.nodes[2 * size_t(ch[k]) + 1];
            rec[2 * k] = make_float4(CL.x, CL.y, CL.z, rec[2 * k].w);
            rec[2 * k + 1] = make_float4(CU.x, CU.y, CU.z, 0.f);
            const uint32_t cw = f2u(CL.w);
            if ((cw >> 30) & 3u) continue;
            const uint32_t gc[2] = {cw & 0x3FFFFFFFu, f2u(CU.w) & 0x3FFFFFFFu};
            for (int m = 0; m < 2; m++) {
                const float4 GL = a.nodes[2 * size_t(gc[m])], GU = a.nodes[2 * size_t(gc[m]) + 1];
                float4* d = rec + 4 + 4 * k + 2 * m;
                d[0] = make_float4(GL.x, GL.y, GL.z, d[0].w);
                d[1] = make_float4(GU.x, GU.y, GU.z, 0.f);
            }
        }
    }
    a.nodes[2 * size_t(g)] = make_float4(mn[0], mn[1], mn[2], L.w);
    a.nodes[2 * size_t(g) + 1] = make_float4(mx[0], mx[1], mx[2], U.w);
}

// Triangle corners in leaf order for positions [p0, p0 + count); the first
// corner's w carries the global triangle id.
__global__ __launch_bounds__(256) void k_tri_soup(MeshRefitArgs a, uint32_t p0, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t k = p0 + uint32_t(i), t = a.prim[k];
    for (int q = 0; q < 3; q++) {
        const float* p = a.verts + 3 * size_t(a.tris[3 * size_t(t) + q]);
        a.tsoup[3 * size_t(k) + q] = make_float4(p[0], p[1], p[2], q == 0 ? __uint_as_float(t) : 0.f);
    }
}

// Distance-bound grid (dist_grid_build, in the same double arithmetic): every
// cell starts at the cap with an empty leaf-order span ...
__global__ __launch_bounds__(256) void k_dist_fill(DistGridArgs a) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(a.nx) * a.ny * a.nz) return;
    a.lb[i] = a.cap;
    a.slo[i] = 0xFFFFFFFFu;
    a.shi[i] = 0u;
}

// ... each triangle (one wave each) lowers the bound of every cell near its
// box to the box-to-box distance, rounded down, and widens the leaf-order
// span of the cells within the cap ...
__global__ __launch_bounds__(256) void k_dist_scatter(DistGridArgs a) {
    const int t = int(blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64);
    if (t >= a.ntri) return;
    const int lane = threadIdx.x & 63;
    const uint32_t gt = uint32_t(a.tri0 + t);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int c = 0; c < 3; c++) {
        const float* p = a.verts + 3 * size_t(a.tris[3 * size_t(gt) + c]);
        for (int k = 0; k < 3; k++) {
            lo[k] = fmin(lo[k], double(p[k]));
            hi[k] = fmax(hi[k], double(p[k]));
        }
    }
    const double o[3] = {a.ox, a.oy, a.oz}, cap = double(a.cap);
    const uint32_t n[3] = {a.nx, a.ny, a.nz};
    long c0[3], c1[3];
    double slack[3];
    for (int k = 0; k < 3; k++) {
        c0[k] = max(0L, long(floor((lo[k] - cap - o[k]) / a.hc)) - 1);
        c1[k] = min(long(n[k]) - 1, long(floor((hi[k] + cap - o[k]) / a.hc)) + 1);
        slack[k] = 1e-6 * (fabs(o[k]) + double(n[k]) * a.hc) + 1e-9;
    }
    if (c1[0] < c0[0] || c1[1] < c0[1] || c1[2] < c0[2]) return;
    const long ex = c1[0] - c0[0] + 1, ey = c1[1] - c0[1] + 1, ez = c1[2] - c0[2] + 1;
    const long total = ex * ey * ez;
    const uint32_t pos = a.tpos[gt];
    for (long q = lane; q < total; q += 64) {
        const long cc[3] = {c0[0] + q % ex, c0[1] + (q / ex) % ey, c0[2] + q / (ex * ey)};
        double d2 = 0.0;
        for (int k = 0; k < 3; k++) {
            const double blo = o[k] + double(cc[k]) * a.hc - slack[k], bhi = o[k] + double(cc[k] + 1) * a.hc + slack[k];
            const double gap = fmax(0.0, fmax(lo[k] - bhi, blo - hi[k]));
            d2 += gap * gap;
        }
        const double d = sqrt(d2);
        const size_t ci = (size_t(cc[2]) * n[1] + size_t(cc[1])) * n[0] + size_t(cc[0]);
        if (d < cap) atomicMin(reinterpret_cast<unsigned int*>(a.lb) + ci, f2u(__double2float_rd(d)));
        if (d <= cap) {
            atomicMin(a.slo + ci, pos);
            atomicMax(a.shi + ci, pos);
        }
    }
}

// ... then each cell's search start: the deepest node whose leaf range holds
// its span (the root when the span is empty), as its stack word.
__global__ __launch_bounds__(256) void k_dist_start(DistGridArgs a) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(a.nx) * a.ny * a.nz) return;
    const uint32_t lo = a.slo[i], hi = a.shi[i];
    uint32_t v = a.root;
    if (lo <= hi) {
        for (;;) {
            const uint32_t lw = f2u(a.nodes[2 * size_t(v)].w);
            if ((lw >> 30) & 3u) break;
            const uint32_t l = lw & 0x3FFFFFFFu, r = f2u(a.nodes[2 * size_t(v) + 1].w) & 0x3FFFFFFFu;
            if (a.nfirst[l] <= lo && hi <= a.nlast[l]) v = l;
            else if (a.nfirst[r] <= lo && hi <= a.nlast[r]) v = r;
            else break;
        }
    }
    a.lb_node[i] = a.nword[v];
}

// Parity grid (parity_grid_build, in the same double arithmetic): the cells a
// triangle's padded projected box touches.
struct ParCells { uint32_t a0, a1, b0, b1; };
__device__ __forceinline__ ParCells par_cells(const ParGridArgs& a, uint32_t gt) {
    double lu = 1e300, hu = -1e300, lw = 1e300, hw = -1e300;
    for (int c = 0; c < 3; c++) {
        const float* p = a.verts + 3 * size_t(a.tris[3 * size_t(gt) + c]);
        const double u = double(p[a.kx]) - a.Sx * double(p[a.kz]), w = double(p[a.ky]) - a.Sy * double(p[a.kz]);
        lu = fmin(lu, u); hu = fmax(hu, u);
        lw = fmin(lw, w); hw = fmax(hw, w);
    }
    lu -= a.pad; hu += a.pad; lw -= a.pad; hw += a.pad;
    auto cell = [&](double x, float o) -> uint32_t {
        const long c = long(floor((x - double(o)) * double(a.inv)));
        return uint32_t(max(0L, min(long(a.dim) - 1, c)));
    };
    return ParCells{cell(lu, a.u0), cell(hu, a.u0), cell(lw, a.w0), cell(hw, a.w0)};
}

__global__ __launch_bounds__(256) void k_par_count(ParGridArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ntri) return;
    const ParCells c = par_cells(a, uint32_t(a.tri0 + t));
    for (uint32_t b = c.b0; b <= c.b1; b++)
        for (uint32_t x = c.a0; x <= c.a1; x++) atomicAdd(a.cnt + size_t(b) * a.dim + x, 1u);
}

// One workgroup: the cells' exclusive prefix (local, into cnt, as the fill's
// cursors) and the mesh's cell starts (global entry positions, dim^2 + 1 of
// them); total receives the entry count.
__global__ __launch_bounds__(1024) void k_par_scan(ParGridArgs a) {
    __shared__ uint32_t s_w[16];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int cells = int(a.dim * a.dim);
    uint32_t carry = 0;
    for (int b = 0; b < cells; b += 1024) {
        const int c = b + threadIdx.x;
        const uint32_t x = c < cells ? a.cnt[c] : 0u;
        const uint32_t inc = uint32_t(wave_incl_scan_i(int(x)));
        if (lane == 63) s_w[w] = inc;
        __syncthreads();
        uint32_t pre = 0, tot = 0;
        for (int k = 0; k < 16; k++) {
            const uint32_t v = s_w[k];
            if (k < w) pre += v;
            tot += v;
        }
        if (c < cells) {
            const uint32_t ex = carry + pre + inc - x;
            a.cnt[c] = ex;
            a.start[c] = a.base + ex;
        }
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.start[cells] = a.base + carry;
        *a.total = carry;
    }
}

// Every registered (cell, triangle) entry: the triangle's corners at the next
// position of the cell's run.  Order inside a cell is arbitrary: the sign
// test only counts hits.
__global__ __launch_bounds__(256) void k_par_fill(ParGridArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ntri) return;
    const uint32_t gt = uint32_t(a.tri0 + t);
    const ParCells c = par_cells(a, gt);
    for (uint32_t b = c.b0; b <= c.b1; b++)
        for (uint32_t x = c.a0; x <= c.a1; x++) {
            const uint32_t e = a.base + atomicAdd(a.cnt + size_t(b) * a.dim + x, 1u);
            for (int q = 0; q < 3; q++) {
                const float* p = a.verts + 3 * size_t(a.tris[3 * size_t(gt) + q]);
                a.soup[3 * size_t(e) + q] = make_float4(p[0], p[1], p[2], 0.f);
            }
        }
}

__global__ __launch_bounds__(256) void k_zero_u32(uint32_t* __restrict__ p, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

static inline unsigned nblocks(size_t n, int bs) { return unsigned((n + bs - 1) / bs); }

void launch_zero_u32(uint32_t* p, size_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_zero_u32, dim3(nblocks(n, 256)), dim3(256), 0, s, p, n);
}

void launch_refit_level(const MeshRefitArgs& a, const uint32_t* list, int count, hipStream_t s) {
    if (count > 0) hipLaunchKernelGGL(k_refit_level, dim3(nblocks(size_t(count), 256)), dim3(256), 0, s, a, list, count);
}
void launch_tri_soup(const MeshRefitArgs& a, uint32_t p0, int count, hipStream_t s) {
    if (count > 0) hipLaunchKernelGGL(k_tri_soup, dim3(nblocks(size_t(count), 256)), dim3(256), 0, s, a, p0, count);
}
void launch_dist_grid(const DistGridArgs& a, hipStream_t s) {
    const size_t cells = size_t(a.nx) * a.ny * a.nz;
    if (cells == 0) return;
    hipLaunchKernelGGL(k_dist_fill, dim3(nblocks(cells, 256)), dim3(256), 0, s, a);
    if (a.ntri > 0) hipLaunchKernelGGL(k_dist_scatter, dim3(nblocks(size_t(a.ntri), 4)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_dist_start, dim3(nblocks(cells, 256)), dim3(256), 0, s, a);
}
void launch_par_count_scan(const ParGridArgs& a, hipStream_t s) {
    if (a.dim == 0) return;
    if (a.ntri > 0) hipLaunchKernelGGL(k_par_count, dim3(nblocks(size_t(a.ntri), 256)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_par_scan, dim3(1), dim3(1024), 0, s, a);
}
void launch_par_fill(const ParGridArgs& a, hipStream_t s) {
    if (a.dim > 0 && a.ntri > 0) hipLaunchKernelGGL(k_par_fill, dim3(nblocks(size_t(a.ntri), 256)), dim3(256), 0, s, a);
}

}  // namespace xpbd
