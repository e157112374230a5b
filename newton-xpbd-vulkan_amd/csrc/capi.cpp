// capi.cpp -- extern "C" boundary over xpbd::Solver (declared in
// include/xpbd_capi.h).  No exception crosses the ABI: every entry point
// returns 0 or 1 and keeps the message for xpbd_last_error.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>
#include <stdexcept>
#include <string>

#include "../../include/xpbd_capi.h"
#include "solver.h"
#include "sort_entry.h"

struct xpbd_solver {
    xpbd::Solver s;
    std::string err;
};

namespace {
thread_local std::string g_err;

template <class F> int guarded(xpbd_solver* h, F&& fn) {
    if (!h) { g_err = "null solver handle"; return 1; }
    try {
        fn();
        h->err.clear();
        return 0;
    } catch (const std::exception& e) {
        h->err = e.what();
        return 1;
    }
}
}  // namespace

extern "C" {

void xpbd_config_default(xpbd_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->dt = 1.0f / 60.0f;
    c->substeps = 1;
    c->iterations = 2;
    c->particle_v_max = 1e5f;
    c->soft_contact_mu = 0.5f;
    c->shape_material_mu = 1.0f;
    c->soft_contact_relaxation = 0.9f;
    c->particle_adhesion = 0.0f;
    c->particle_mu = 0.5f;
    c->particle_cohesion = 0.0f;
    c->particle_max_radius = 0.02f;
    c->search_radius = 0.04f;
    c->soft_contact_margin = 0.01f;
    c->enable_restitution = 0;
    c->soft_contact_restitution = 0.0f;
    c->has_ground_plane = 1;
    c->ground_plane_altitude = 0.0f;
    c->gravity[0] = 0.0f; c->gravity[1] = 0.0f; c->gravity[2] = -9.81f;
    c->device_index = -1;
    c->compat_mode = 0;
    c->enable_validation = 0;
    c->enable_diagnostics = 0;
    c->enable_profile = 0;
    c->max_diag_pairs = 100000u;
    c->substep_batch_size = 15;
}

int xpbd_device_count(int* count) {
    if (!count) return 1;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return 0;
}

int xpbd_device_name(int index, char* buf, size_t len) {
    hipDeviceProp_t p;
    if (!buf || len == 0) { g_err = "null buffer"; return 1; }
    if (hipGetDeviceProperties(&p, index) != hipSuccess) { g_err = "invalid device index"; return 1; }
    std::strncpy(buf, p.name, len - 1);
    buf[len - 1] = '\0';
    return 0;
}

int xpbd_create(xpbd_solver** out) {
    if (!out) { g_err = "null output pointer"; return 1; }
    try {
        *out = new xpbd_solver();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        *out = nullptr;
        return 1;
    }
}

void xpbd_destroy(xpbd_solver* h) { delete h; }

const char* xpbd_last_error(const xpbd_solver* h) { return h ? h->err.c_str() : g_err.c_str(); }

int xpbd_init(xpbd_solver* h, const xpbd_config* c, size_t n, size_t mc) {
    return guarded(h, [&] {
        if (!c) throw std::runtime_error("null config");
        h->s.init(*c, n, mc);
    });
}
int xpbd_add_mesh(xpbd_solver* h, const float* v, size_t nv, const uint32_t* t, size_t nt, float margin,
                  const float* rot, int wt) {
    return guarded(h, [&] { h->s.add_mesh(v, nv, t, nt, margin, rot, wt); });
}
int xpbd_finalize_meshes(xpbd_solver* h) { return guarded(h, [&] { h->s.finalize_meshes(); }); }
int xpbd_set_particles(xpbd_solver* h, const float* q, const float* qd, const float* w, const float* r, const int32_t* f,
                       size_t n) {
    return guarded(h, [&] { h->s.set_particles(q, qd, w, r, f, n); });
}
int xpbd_update_mesh(xpbd_solver* h, size_t mi, const float* v, size_t nv, const float* vel) {
    return guarded(h, [&] { h->s.update_mesh(mi, v, nv, vel); });
}
int xpbd_step(xpbd_solver* h, float dt, int s, int it) { return guarded(h, [&] { h->s.step(dt, s, it); }); }
int xpbd_collide_only(xpbd_solver* h) { return guarded(h, [&] { h->s.collide_only(); }); }
int xpbd_step_solve_only(xpbd_solver* h, float dt, int s, int it) {
    return guarded(h, [&] { h->s.step_solve_only(dt, s, it); });
}
int xpbd_get_positions(xpbd_solver* h, float* out, size_t n) { return guarded(h, [&] { h->s.get_positions(out, n); }); }
int xpbd_get_velocities(xpbd_solver* h, float* out, size_t n) { return guarded(h, [&] { h->s.get_velocities(out, n); }); }
int xpbd_particle_count(const xpbd_solver* h, size_t* out) {
    if (!h || !out) return 1;
    *out = h->s.particle_count();
    return 0;
}
int xpbd_mesh_count(const xpbd_solver* h, size_t* out) {
    if (!h || !out) return 1;
    *out = h->s.mesh_count();
    return 0;
}
int xpbd_last_contact_count(const xpbd_solver* h, int* out) {
    if (!h || !out) return 1;
    *out = h->s.last_contact_count();
    return 0;
}
int xpbd_get_contact_particles(xpbd_solver* h, int32_t* out, size_t cap) {
    return guarded(h, [&] { h->s.get_contact_ints(false, out, cap); });
}
int xpbd_get_contact_shapes(xpbd_solver* h, int32_t* out, size_t cap) {
    return guarded(h, [&] { h->s.get_contact_ints(true, out, cap); });
}
int xpbd_get_contact_normals(xpbd_solver* h, float* out, size_t cap) {
    return guarded(h, [&] { h->s.get_contact_vec(0, out, cap); });
}
int xpbd_get_contact_body_positions(xpbd_solver* h, float* out, size_t cap) {
    return guarded(h, [&] { h->s.get_contact_vec(1, out, cap); });
}
int xpbd_get_contact_body_velocities(xpbd_solver* h, float* out, size_t cap) {
    return guarded(h, [&] { h->s.get_contact_vec(2, out, cap); });
}
int xpbd_set_contact_data(xpbd_solver* h, const int32_t* p, const int32_t* s, const float* bp, const float* nrm,
                          int count) {
    return guarded(h, [&] { h->s.set_contact_data(p, s, bp, nrm, count); });
}
int xpbd_get_sorted_cells(xpbd_solver* h, uint32_t* out, size_t n) { return guarded(h, [&] { h->s.get_sorted(false, out, n); }); }
int xpbd_get_sorted_point_ids(xpbd_solver* h, uint32_t* out, size_t n) {
    return guarded(h, [&] { h->s.get_sorted(true, out, n); });
}
int xpbd_get_cell_starts(xpbd_solver* h, int32_t* out, size_t n) { return guarded(h, [&] { h->s.get_cell_table(false, out, n); }); }
int xpbd_get_cell_ends(xpbd_solver* h, int32_t* out, size_t n) { return guarded(h, [&] { h->s.get_cell_table(true, out, n); }); }
int xpbd_get_deltas(xpbd_solver* h, float* out, size_t n) { return guarded(h, [&] { h->s.get_deltas(out, n); }); }
int xpbd_get_diag_pair_count(xpbd_solver* h, uint32_t* out) { return guarded(h, [&] { *out = h->s.diag_word(0); }); }
int xpbd_get_diag_overflow_count(xpbd_solver* h, uint32_t* out) { return guarded(h, [&] { *out = h->s.diag_word(1); }); }
int xpbd_get_diag_pairs(xpbd_solver* h, xpbd_pair_record* out, size_t cap, size_t* written) {
    return guarded(h, [&] {
        size_t w = h->s.get_diag_pairs(out, cap);
        if (written) *written = w;
    });
}
int xpbd_reset_diag_buffers(xpbd_solver* h) { return guarded(h, [&] { h->s.reset_diag(); }); }
// The reference binds these buffers but no shader writes them, so after the
// zero fill at init they read back as zeros (scene.cpp:116-128; the shape,
// apply and mesh shaders declare no binding for them).
int xpbd_get_p11_diag(xpbd_solver* h, uint32_t* out32) {
    return guarded(h, [&] { std::memset(out32, 0, XPBD_DIAG_WORDS * sizeof(uint32_t)); });
}
int xpbd_get_p11_apply_diag(xpbd_solver* h, uint32_t* out32) {
    return guarded(h, [&] { std::memset(out32, 0, XPBD_DIAG_WORDS * sizeof(uint32_t)); });
}
int xpbd_get_mesh_candidate_diag(xpbd_solver* h, uint32_t* out32) {
    return guarded(h, [&] { std::memset(out32, 0, XPBD_DIAG_WORDS * sizeof(uint32_t)); });
}
int xpbd_get_shape_contact_diag(xpbd_solver* h, float* out, size_t cap, size_t* len) {
    return guarded(h, [&] {
        size_t L = h->s.shape_contact_diag_len();
        if (len) *len = L;
        if (out) std::memset(out, 0, std::min(cap, L) * sizeof(float));
    });
}
int xpbd_get_mesh_bvh_node_count(xpbd_solver* h, size_t mi, size_t* nodes) {
    return guarded(h, [&] { *nodes = h->s.bvh_nodes(mi); });
}
int xpbd_get_mesh_bvh_lowers(xpbd_solver* h, size_t mi, float* out, size_t cap) {
    return guarded(h, [&] { h->s.bvh_half(mi, false, out, cap); });
}
int xpbd_get_mesh_bvh_uppers(xpbd_solver* h, size_t mi, float* out, size_t cap) {
    return guarded(h, [&] { h->s.bvh_half(mi, true, out, cap); });
}
int xpbd_print_profiler_summary(xpbd_solver* h) { return guarded(h, [&] { h->s.print_profiler_summary(); }); }
int xpbd_profile_stages(xpbd_solver* h, char* names, size_t names_len, double* ms, uint64_t* launches, size_t cap,
                        size_t* count) {
    return guarded(h, [&] {
        size_t c = h->s.profile_stages(names, names_len, ms, launches, cap);
        if (count) *count = c;
    });
}
int xpbd_profile_reset(xpbd_solver* h) { return guarded(h, [&] { h->s.profile_reset(); }); }

int xpbd_add_box_body(xpbd_solver* h, const float* pos, const float* quat, const float* half, float density, int flags) {
    return guarded(h, [&] { h->s.add_box_body(pos, quat, half, density, flags); });
}
int xpbd_set_body_params(xpbd_solver* h, float ad, float relax, float mt, float mr, int uw) {
    return guarded(h, [&] { h->s.set_body_params(ad, relax, mt, mr, uw); });
}
int xpbd_rollout_forward(xpbd_solver* h, const float* x0, const float* v0, float relax, float sdt, int substeps,
                         int every, float* x_out, float* v_out) {
    return guarded(h, [&] { h->s.rollout_forward(x0, v0, relax, sdt, substeps, every, x_out, v_out); });
}
int xpbd_rollout_backward(xpbd_solver* h, const float* gx, const float* gv, float* gx0, float* gv0, double* grelax) {
    return guarded(h, [&] { h->s.rollout_backward(gx, gv, gx0, gv0, grelax); });
}
int xpbd_mesh_query_stats(xpbd_solver* h, uint64_t* out6) { return guarded(h, [&] { h->s.mesh_query_stats(out6); }); }
int xpbd_init_batch(xpbd_solver* h, const xpbd_config* c, const uint64_t* counts, size_t ns, size_t mc) {
    return guarded(h, [&] {
        if (!c) throw std::runtime_error("null config");
        if (!counts && ns) throw std::runtime_error("null scene counts");
        std::vector<size_t> cnt(counts, counts + ns);
        h->s.init_batch(*c, cnt.data(), ns, mc);
    });
}
int xpbd_scene_count(const xpbd_solver* h, size_t* out) {
    if (!h || !out) return 1;
    *out = h->s.scene_count();
    return 0;
}
int xpbd_set_wrap_planes(xpbd_solver* h, int on) { return guarded(h, [&] { h->s.set_wrap_planes(on != 0); }); }
int xpbd_wrap_planes(const xpbd_solver* h, int* out) {
    if (!h || !out) return 1;
    *out = h->s.wrap_planes() ? 1 : 0;
    return 0;
}
int xpbd_set_graph_mode(xpbd_solver* h, int on) { return guarded(h, [&] { h->s.set_graph_mode(on != 0); }); }
int xpbd_radix_tile(const xpbd_solver* h, int* out) {
    if (!h || !out) { g_err = "null argument"; return 1; }
    *out = h->s.radix_tile();
    return 0;
}
int xpbd_pp_block(const xpbd_solver* h, int* out) {
    if (!h || !out) { g_err = "null argument"; return 1; }
    *out = h->s.pp_block();
    return 0;
}
int xpbd_set_tuning(xpbd_solver* h, int pp_block, int radix_tile) {
    return guarded(h, [&] { h->s.set_tuning(pp_block, radix_tile); });
}
int xpbd_cell_table_bytes(const xpbd_solver* h, size_t* out) {
    if (!h || !out) { g_err = "null argument"; return 1; }
    *out = h->s.cell_table_bytes();
    return 0;
}
int xpbd_set_deterministic(xpbd_solver* h, int on) { return guarded(h, [&] { h->s.set_deterministic(on != 0); }); }
int xpbd_deterministic(const xpbd_solver* h, int* out) {
    if (!h || !out) { g_err = "null argument"; return 1; }
    *out = h->s.deterministic() ? 1 : 0;
    return 0;
}
int xpbd_autotune(xpbd_solver* h, int frames, float dt, int substeps, int iterations, int* block) {
    return guarded(h, [&] {
        const int b = h->s.autotune(frames, dt, substeps, iterations);
        if (block) *block = b;
    });
}
int xpbd_body_count(const xpbd_solver* h, size_t* out) {
    if (!h || !out) return 1;
    *out = h->s.body_count();
    return 0;
}
int xpbd_get_bodies(xpbd_solver* h, float* q7, float* qd6, size_t cap) {
    return guarded(h, [&] { h->s.get_bodies(q7, qd6, cap); });
}
int xpbd_set_body_velocities(xpbd_solver* h, const float* qd6, size_t n) {
    return guarded(h, [&] { h->s.set_body_velocities(qd6, n); });
}

int xpbd_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                    size_t n, int key_bits, void* stream) {
    try {
        xpbd::sort_pairs_alloc(keys_in, vals_in, keys_out, vals_out, n, key_bits, static_cast<hipStream_t>(stream));
        g_err.clear();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

int xpbd_check_arith(uint64_t* out3) {
    try {
        if (!out3) throw std::runtime_error("xpbd_check_arith: null output");
        xpbd::check_arith(out3, nullptr);
        g_err.clear();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

int xpbd_sort_pairs_noalloc(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                            uint32_t* vals_out, size_t n, int key_bits, void* stream) {
    try {
        xpbd::sort_pairs_noalloc(keys_in, vals_in, keys_out, vals_out, n, key_bits, static_cast<hipStream_t>(stream));
        g_err.clear();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

int xpbd_lbvh_build(const float* verts, size_t nv, const uint32_t* tris, size_t nt, float* lowers, float* uppers,
                    uint32_t* prim, uint32_t* root) {
    (void)nv;
    try {
        xpbd::MeshBvhHost b;
        xpbd::lbvh_build(verts, tris, uint32_t(nt), b);
        std::memcpy(lowers, b.lo.data(), b.lo.size() * sizeof(xpbd::NodeHalf));
        std::memcpy(uppers, b.hi.data(), b.hi.size() * sizeof(xpbd::NodeHalf));
        std::memcpy(prim, b.prim.data(), b.prim.size() * sizeof(uint32_t));
        *root = b.root;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

int xpbd_lbvh_refit(const float* verts, const float* new_verts, size_t nv, const uint32_t* tris, size_t nt,
                    float* lowers, float* uppers) {
    try {
        xpbd::MeshBvhHost b;
        xpbd::lbvh_build(verts, tris, uint32_t(nt), b);
        std::vector<float> pts(new_verts, new_verts + 3 * nv);
        xpbd::lbvh_refit(b, pts.data(), tris);
        std::memcpy(lowers, b.lo.data(), b.lo.size() * sizeof(xpbd::NodeHalf));
        std::memcpy(uppers, b.hi.data(), b.hi.size() * sizeof(xpbd::NodeHalf));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

int xpbd_parity_grid_candidates(const float* verts, size_t nv, const uint32_t* tris, size_t nt, const float* points,
                                size_t np, uint32_t* counts, uint32_t* ids, size_t ids_cap, size_t* total) {
    try {
        for (size_t k = 0; k < 3 * nt; k++)
            if (tris[k] >= nv) throw std::runtime_error("parity grid: vertex index out of range");
        xpbd::ParityGrid g;
        xpbd::parity_grid_build(verts, tris, uint32_t(nt), 0u, g);
        size_t w = 0;
        for (size_t k = 0; k < np; k++) {
            const long c = g.dim ? xpbd::parity_grid_cell(g, points + 3 * k) : -1;
            uint32_t cnt = 0;
            if (c >= 0)
                for (uint32_t e = g.start[size_t(c)]; e < g.start[size_t(c) + 1]; e++, cnt++, w++)
                    if (w < ids_cap) ids[w] = g.tri[e];
            counts[k] = cnt;
        }
        if (total) *total = w;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

int xpbd_dist_grid_bounds(const float* verts, size_t nv, const uint32_t* tris, size_t nt, float cap, const float* points,
                          size_t np, float* bounds) {
    try {
        for (size_t k = 0; k < 3 * nt; k++)
            if (tris[k] >= nv) throw std::runtime_error("distance grid: vertex index out of range");
        xpbd::DistGrid g;
        xpbd::dist_grid_build(verts, tris, uint32_t(nt), cap, g);
        for (size_t k = 0; k < np; k++) bounds[k] = g.lb.empty() ? -1.0f : xpbd::dist_grid_bound(g, points + 3 * k);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

}  // extern "C"
