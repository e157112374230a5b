// mesh_update.cpp -- animated colliders: XpbdSolver::update_mesh
// (scene.cpp:617-696) on the device.
//
// The reference patches the host vertex array, refits the BVH on the CPU
// (MeshBvh::refit, mesh_bvh.h:142-181) and re-uploads every vertex and the
// whole node arrays.  Here update_mesh uploads the mesh's vertices and
// velocities only, then refits the BVH on the device one height at a time,
// refreshes the copies of its boxes and triangle corners the narrowphase
// reads, and rebuilds the mesh's distance-bound and parity grids -- all into
// the buffers finalize_meshes allocated, with room to spare, so captured
// frames keep their pointers and stay valid.  When a grid outgrows its room
// (the mesh grew or deformed far), every mesh's derived data is rebuilt on the
// host as finalize_meshes does, and the frame graphs are dropped.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "devbuf.h"

namespace xpbd {

// Static tables for the device updates, built once the BVHs exist: each
// mesh's reachable nodes grouped by height (leaves first), their leaf-order
// ranges and stack words, each triangle's leaf position, each mesh's used
// vertices, and scratch sized to the largest grid.
void Solver::prepare_mesh_updates() {
    const size_t M = meshes_.size();
    size_t total_nodes = 0;
    for (const MeshBvhHost& b : bvh_) total_nodes += b.lo.size();
    std::vector<uint32_t> nfirst(std::max<size_t>(total_nodes, 1), 0), nlast(nfirst.size(), 0),
        nword(nfirst.size(), 0), list;
    refit_lvl_.assign(M, {});
    mesh_vref_.assign(M, {});
    for (size_t c = 0; c < M; c++) {
        const MeshBvhHost& b = bvh_[c];
        const uint32_t nb = uint32_t(node_off_[c]), pb = uint32_t(prim_off_[c]);
        if (!b.lo.empty()) {
            std::vector<uint32_t> order{b.root};
            for (size_t k = 0; k < order.size(); k++) {
                const uint32_t v = order[k];
                if (!((b.lo[v].packed >> 30) & 3u)) {
                    order.push_back(b.lo[v].packed & 0x3FFFFFFFu);
                    order.push_back(b.hi[v].packed & 0x3FFFFFFFu);
                }
            }
            std::vector<int> height(b.lo.size(), 0);
            int top = 0;
            for (size_t k = order.size(); k-- > 0;) {
                const uint32_t v = order[k], lp = b.lo[v].packed;
                if ((lp >> 30) & 3u) {
                    nfirst[nb + v] = (lp & 0x3FFFFFFFu) + pb;
                    nlast[nb + v] = (b.hi[v].packed & 0x3FFFFFFFu) - 1u + pb;
                } else {
                    const uint32_t l = lp & 0x3FFFFFFFu, r = b.hi[v].packed & 0x3FFFFFFFu;
                    height[v] = 1 + std::max(height[l], height[r]);
                    nfirst[nb + v] = std::min(nfirst[nb + l], nfirst[nb + r]);
                    nlast[nb + v] = std::max(nlast[nb + l], nlast[nb + r]);
                }
                nword[nb + v] = node_word(c, v);
                top = std::max(top, height[v]);
            }
            std::vector<std::vector<uint32_t>> by(size_t(top) + 1);
            for (const uint32_t v : order) by[size_t(height[v])].push_back(nb + v);
            for (const auto& lvl : by) {
                refit_lvl_[c].push_back(uint32_t(list.size()));
                list.insert(list.end(), lvl.begin(), lvl.end());
            }
            refit_lvl_[c].push_back(uint32_t(list.size()));
        }
        std::vector<uint32_t>& vr = mesh_vref_[c];
        const size_t t0 = tri_off_[c], nt = meshes_[c].tris.size() / 3;
        vr.assign(all_tris_.begin() + 3 * t0, all_tris_.begin() + 3 * (t0 + nt));
        std::sort(vr.begin(), vr.end());
        vr.erase(std::unique(vr.begin(), vr.end()), vr.end());
    }
    std::vector<uint32_t> tpos(std::max<size_t>(all_tris_.size() / 3, 1), 0);
    for (size_t k = 0; k < prim_h_.size(); k++) tpos[prim_h_[k]] = uint32_t(k);
    if (list.empty()) list.push_back(0);
    upload(refit_list_d_, list.data(), list.size());
    upload(nfirst_d_, nfirst.data(), nfirst.size());
    upload(nlast_d_, nlast.data(), nlast.size());
    upload(nword_d_, nword.data(), nword.size());
    upload(tpos_d_, tpos.data(), tpos.size());
    size_t cells = 1, pcells = 1;
    for (size_t c = 0; c < M; c++) {
        cells = std::max(cells, lb_room_[c]);
        pcells = std::max(pcells, size_t(pgrid_[c].dim) * pgrid_[c].dim);
    }
    dalloc(span_lo_d_, cells);
    dalloc(span_hi_d_, cells);
    dalloc(pcnt_d_, pcells);
    dalloc(ptotal_d_, 1);
}

bool Solver::update_mesh_on_device(size_t c) {
    MeshRefitArgs ra;
    ra.nodes = nodes_d_.p; ra.rec = cnodes_d_.p; ra.lbox = lbox_d_.p; ra.tsoup = tsoup_.p;
    ra.prim = prim_d_.p; ra.tris = tris_d_.p; ra.verts = verts_d_.p;
    ra.nprim = uint32_t(prim_h_.size());
    const std::vector<uint32_t>& lv = refit_lvl_[c];
    for (size_t l = 0; l + 1 < lv.size(); l++)
        launch_refit_level(ra, refit_list_d_.p + lv[l], int(lv[l + 1] - lv[l]), stream_);
    launch_tri_soup(ra, uint32_t(prim_off_[c]), int(bvh_[c].prim.size()), stream_);

    const std::vector<uint32_t>& vr = mesh_vref_[c];
    const int nt = int(meshes_[c].tris.size() / 3);
    DistGrid& d = dgrid_[c];
    if (d.nx > 0 && lb_room_[c] > 0) {
        // the grid follows the mesh's box grown by the cap, with the cell edge
        // it was built with (dist_grid_build)
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (const uint32_t v : vr)
            for (int k = 0; k < 3; k++) {
                const double x = all_verts_[3 * size_t(v) + k];
                lo[k] = std::min(lo[k], x);
                hi[k] = std::max(hi[k], x);
            }
        uint32_t n[3];
        for (int k = 0; k < 3; k++) {
            lo[k] -= d.cap;
            hi[k] += d.cap;
            n[k] = uint32_t(std::ceil((hi[k] - lo[k]) / d.h)) + 1;
        }
        if (size_t(n[0]) * n[1] * n[2] > lb_room_[c]) return false;
        d.ox = float(lo[0]); d.oy = float(lo[1]); d.oz = float(lo[2]);
        d.nx = n[0]; d.ny = n[1]; d.nz = n[2];
        d.lb.clear(); d.start.clear();   // the device holds the bounds now
        MeshInfo& mi = info_h_[c];
        mi.lb_ox = d.ox; mi.lb_oy = d.oy; mi.lb_oz = d.oz;
        mi.lb_nx = n[0]; mi.lb_ny = n[1]; mi.lb_nz = n[2];
        DistGridArgs da;
        da.tri0 = int(tri_off_[c]); da.ntri = nt;
        da.tris = tris_d_.p; da.verts = verts_d_.p; da.tpos = tpos_d_.p;
        da.ox = double(d.ox); da.oy = double(d.oy); da.oz = double(d.oz);
        da.hc = 1.0 / double(d.inv);
        da.nx = n[0]; da.ny = n[1]; da.nz = n[2];
        da.cap = d.cap;
        da.lb = lb_.p + mi.lb_off; da.slo = span_lo_d_.p; da.shi = span_hi_d_.p; da.lb_node = lb_node_.p + mi.lb_off;
        da.nodes = nodes_d_.p; da.nfirst = nfirst_d_.p; da.nlast = nlast_d_.p; da.nword = nword_d_.p;
        da.root = uint32_t(node_off_[c]) + bvh_[c].root;
        launch_dist_grid(da, stream_);
    }

    ParityGrid& g = pgrid_[c];
    if (meshes_[c].watertight && g.dim > 0) {
        // parity_grid_build's extent, pad and cell size over the moved corners
        const RayAxes r = parity_axes();
        const double Sx = double(r.Sx), Sy = double(r.Sy);
        double scale = 0.0, umin = 1e300, umax = -1e300, wmin = 1e300, wmax = -1e300;
        for (const uint32_t v : vr) {
            const float* p = all_verts_.data() + 3 * size_t(v);
            const double u = double(p[r.kx]) - Sx * double(p[r.kz]), w = double(p[r.ky]) - Sy * double(p[r.kz]);
            umin = std::min(umin, u); umax = std::max(umax, u);
            wmin = std::min(wmin, w); wmax = std::max(wmax, w);
            for (int k = 0; k < 3; k++) scale = std::max(scale, double(std::fabs(p[k])));
        }
        const double pad = 1e-4 * (1.0 + scale);
        umin -= pad; umax += pad; wmin -= pad; wmax += pad;
        const double ext = std::max(umax - umin, wmax - wmin);
        const float cell = float(ext / double(g.dim)) * 1.0001f;
        g.u0 = float(umin);
        g.w0 = float(wmin);
        g.inv = 1.0f / cell;
        g.start.clear(); g.tri.clear();   // the device holds the entries now
        MeshInfo& mi = info_h_[c];
        ParGridArgs pa;
        pa.tri0 = int(tri_off_[c]); pa.ntri = nt;
        pa.tris = tris_d_.p; pa.verts = verts_d_.p;
        pa.kx = r.kx; pa.ky = r.ky; pa.kz = r.kz;
        pa.Sx = Sx; pa.Sy = Sy; pa.pad = pad;
        pa.u0 = g.u0; pa.w0 = g.w0; pa.inv = g.inv; pa.dim = g.dim;
        pa.cnt = pcnt_d_.p; pa.start = pg_start_.p + mi.pg_cell_off; pa.base = uint32_t(pg_base_[c]);
        pa.total = ptotal_d_.p; pa.soup = pg_soup_.p;
        launch_zero_u32(pcnt_d_.p, size_t(g.dim) * g.dim, stream_);
        launch_par_count_scan(pa, stream_);
        uint32_t total = 0;
        HIPCHECK(hipMemcpyAsync(&total, ptotal_d_.p, sizeof(total), hipMemcpyDeviceToHost, stream_));
        HIPCHECK(hipStreamSynchronize(stream_));
        if (total > pg_room_[c]) return false;
        launch_par_fill(pa, stream_);
        mi.pg_u0 = g.u0; mi.pg_w0 = g.w0; mi.pg_inv = g.inv;
    }
    HIPCHECK(hipMemcpyAsync(info_d_.p, info_h_.data(), info_h_.size() * sizeof(MeshInfo), hipMemcpyHostToDevice, stream_));
    HIPCHECK(hipGetLastError());
    return true;
}

void Solver::rebuild_meshes_on_host() {
    HIPCHECK(hipStreamSynchronize(stream_));
    invalidate_graphs();
    for (size_t c = 0; c < meshes_.size(); c++) lbvh_refit(bvh_[c], all_verts_.data(), all_tris_.data() + 3 * tri_off_[c]);
    upload_nodes();
    upload_tri_soup();
    for (size_t c = 0; c < meshes_.size(); c++) build_query_grids(c);
    upload_query_grids();
    prepare_mesh_updates();
}

void Solver::update_mesh(size_t mi, const float* v, size_t nv, const float* vel) {
    if (!finalized_) throw std::runtime_error("update_mesh: call finalize_meshes() first");
    if (mi >= meshes_.size()) throw std::runtime_error("update_mesh: mesh_index out of range");
    if (nv != meshes_[mi].verts.size() / 3) throw std::runtime_error("update_mesh: vertex count mismatch");
    DeviceGuard g(device_);
    HIPCHECK(hipStreamSynchronize(stream_));
    const size_t vs = vert_off_[mi];
    std::memcpy(all_verts_.data() + 3 * vs, v, nv * 3 * sizeof(float));
    if (vel) std::memcpy(all_vels_.data() + 3 * vs, vel, nv * 3 * sizeof(float));
    else std::fill(all_vels_.begin() + 3 * vs, all_vels_.begin() + 3 * (vs + nv), 0.f);
    std::memcpy(meshes_[mi].verts.data(), v, nv * 3 * sizeof(float));
    if (nv) {
        HIPCHECK(hipMemcpyAsync(verts_d_.p + 3 * vs, all_verts_.data() + 3 * vs, nv * 3 * sizeof(float),
                                hipMemcpyHostToDevice, stream_));
        HIPCHECK(hipMemcpyAsync(vvel_d_.p + 3 * vs, all_vels_.data() + 3 * vs, nv * 3 * sizeof(float),
                                hipMemcpyHostToDevice, stream_));
    }
    if (!update_mesh_on_device(mi)) rebuild_meshes_on_host();
    HIPCHECK(hipStreamSynchronize(stream_));
}

}  // namespace xpbd
