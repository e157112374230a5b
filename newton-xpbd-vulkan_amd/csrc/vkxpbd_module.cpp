// vkxpbd_module.cpp -- the `vkxpbd` Python module surface of
// src/bindings.cpp:8-301 (same classes, methods, argument names, defaults and
// array shapes), implemented on top of the C-ABI in include/xpbd_capi.h only.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/xpbd_capi.h"

namespace py = pybind11;

namespace {

struct CfgWrap {
    xpbd_config c;
    std::string shader_dir = "shaders";
    CfgWrap() { xpbd_config_default(&c); }
};

struct PySolver {
    xpbd_solver* h = nullptr;
    PySolver() {
        if (xpbd_create(&h) != 0) throw std::runtime_error(xpbd_last_error(nullptr));
    }
    ~PySolver() { xpbd_destroy(h); }
    void check(int rc) const {
        if (rc != 0) throw std::runtime_error(xpbd_last_error(h));
    }
    size_t n() const {
        size_t v = 0;
        xpbd_particle_count(h, &v);
        return v;
    }
    int count() const {
        int v = 0;
        xpbd_last_contact_count(h, &v);
        return v;
    }
};

using farr = py::array_t<float, py::array::c_style | py::array::forcecast>;
using iarr = py::array_t<int32_t, py::array::c_style | py::array::forcecast>;
using uarr = py::array_t<uint32_t, py::array::c_style | py::array::forcecast>;

py::array_t<float> vec3_array(size_t n) { return py::array_t<float>({n, size_t(3)}); }

}  // namespace

PYBIND11_MODULE(_vkxpbd, m) {
    m.doc() = "XPBD physics solver extension module (HIP, AMD Instinct MI355X)";

    auto cfg = py::class_<CfgWrap>(m, "SolverConfig");
    cfg.def(py::init<>());
#define CFG_F(name) cfg.def_property(#name, [](const CfgWrap& c) { return c.c.name; }, [](CfgWrap& c, decltype(xpbd_config::name) v) { c.c.name = v; })
#define CFG_B(name) cfg.def_property(#name, [](const CfgWrap& c) { return c.c.name != 0; }, [](CfgWrap& c, bool v) { c.c.name = v ? 1 : 0; })
    CFG_F(dt);
    CFG_F(substeps);
    CFG_F(iterations);
    CFG_F(particle_v_max);
    CFG_F(soft_contact_mu);
    CFG_F(shape_material_mu);
    CFG_F(soft_contact_relaxation);
    CFG_F(particle_adhesion);
    CFG_F(particle_mu);
    CFG_F(particle_cohesion);
    CFG_F(particle_max_radius);
    CFG_F(search_radius);
    CFG_F(soft_contact_margin);
    CFG_B(enable_restitution);
    CFG_F(soft_contact_restitution);
    CFG_B(has_ground_plane);
    CFG_F(ground_plane_altitude);
    cfg.def_readwrite("shader_dir", &CfgWrap::shader_dir);
    CFG_F(device_index);
    CFG_B(compat_mode);
    CFG_B(enable_validation);
    CFG_B(enable_diagnostics);
    CFG_B(enable_profile);
    CFG_F(max_diag_pairs);
    CFG_F(substep_batch_size);
#undef CFG_F
#undef CFG_B
    cfg.def_property(
        "gravity", [](const CfgWrap& c) { return std::vector<float>{c.c.gravity[0], c.c.gravity[1], c.c.gravity[2]}; },
        [](CfgWrap& c, const std::vector<float>& g) {
            if (g.size() >= 3) { c.c.gravity[0] = g[0]; c.c.gravity[1] = g[1]; c.c.gravity[2] = g[2]; }
        });

    m.def(
        "enumerate_devices",
        []() {
            int n = 0;
            xpbd_device_count(&n);
            std::vector<std::string> names;
            for (int i = 0; i < n; i++) {
                char buf[256];
                if (xpbd_device_name(i, buf, sizeof(buf)) == 0) names.emplace_back(buf);
            }
            return names;
        },
        "Return list of available GPU device names");

    py::class_<PySolver>(m, "XpbdSolver")
        .def(py::init<>())
        .def(
            "init",
            [](PySolver& s, const CfgWrap& c, size_t n, size_t mc) { s.check(xpbd_init(s.h, &c.c, n, mc)); },
            py::arg("config"), py::arg("num_particles"), py::arg("max_contacts") = 0)
        .def(
            "add_mesh",
            [](PySolver& s, farr vertices, uarr indices, float shape_margin, py::object shape_rot, py::object is_watertight) {
                const float* rot = nullptr;
                farr rot_arr;
                if (!shape_rot.is_none()) {
                    rot_arr = shape_rot.cast<farr>();
                    rot = rot_arr.data();
                }
                int wt = -1;
                if (!is_watertight.is_none()) wt = is_watertight.cast<bool>() ? 1 : 0;
                s.check(xpbd_add_mesh(s.h, vertices.data(), size_t(vertices.size()) / 3, indices.data(),
                                      size_t(indices.size()) / 3, shape_margin, rot, wt));
            },
            py::arg("vertices"), py::arg("indices"), py::arg("shape_margin") = 0.0f, py::arg("shape_rot") = py::none(),
            py::arg("is_watertight") = py::none())
        .def("finalize_meshes", [](PySolver& s) { s.check(xpbd_finalize_meshes(s.h)); })
        .def(
            "set_particles",
            [](PySolver& s, farr q, py::object qd, py::object inv_mass, py::object radii, py::object flags) {
                const size_t n = size_t(q.size()) / 3;
                farr qd_a, w_a, r_a;
                iarr f_a;
                const float* qd_p = nullptr;
                const float* w_p = nullptr;
                const float* r_p = nullptr;
                const int32_t* f_p = nullptr;
                if (!qd.is_none()) { qd_a = qd.cast<farr>(); qd_p = qd_a.data(); }
                if (!inv_mass.is_none()) { w_a = inv_mass.cast<farr>(); w_p = w_a.data(); }
                if (!radii.is_none()) { r_a = radii.cast<farr>(); r_p = r_a.data(); }
                if (!flags.is_none()) { f_a = flags.cast<iarr>(); f_p = f_a.data(); }
                s.check(xpbd_set_particles(s.h, q.data(), qd_p, w_p, r_p, f_p, n));
            },
            py::arg("q"), py::arg("qd") = py::none(), py::arg("inv_mass") = py::none(), py::arg("radii") = py::none(),
            py::arg("flags") = py::none())
        .def(
            "step",
            [](PySolver& s, float dt, int substeps, int iterations) {
                int rc;
                {
                    py::gil_scoped_release nogil;
                    rc = xpbd_step(s.h, dt, substeps, iterations);
                }
                s.check(rc);
            },
            py::arg("dt") = 1.0f / 60.0f, py::arg("substeps") = 1, py::arg("iterations") = 2)
        .def("get_positions",
             [](PySolver& s) {
                 size_t n = s.n();
                 auto r = vec3_array(n);
                 s.check(xpbd_get_positions(s.h, r.mutable_data(), n));
                 return r;
             })
        .def("get_velocities",
             [](PySolver& s) {
                 size_t n = s.n();
                 auto r = vec3_array(n);
                 s.check(xpbd_get_velocities(s.h, r.mutable_data(), n));
                 return r;
             })
        .def("collide_only",
             [](PySolver& s) {
                 int rc;
                 {
                     py::gil_scoped_release nogil;
                     rc = xpbd_collide_only(s.h);
                 }
                 s.check(rc);
             })
        .def("step_solve_only",
             [](PySolver& s, float dt, int substeps, int iterations) {
                 int rc;
                 {
                     py::gil_scoped_release nogil;
                     rc = xpbd_step_solve_only(s.h, dt, substeps, iterations);
                 }
                 s.check(rc);
             })
        .def(
            "update_mesh",
            [](PySolver& s, size_t mesh_index, farr verts, std::optional<farr> velocities) {
                const size_t nv = size_t(verts.size()) / 3;
                const float* vel = nullptr;
                if (velocities.has_value() && velocities->size() > 0) {
                    if (size_t(velocities->size()) / 3 != nv)
                        throw std::runtime_error("update_mesh: velocities size mismatch with vertices");
                    vel = velocities->data();
                }
                s.check(xpbd_update_mesh(s.h, mesh_index, verts.data(), nv, vel));
            },
            py::arg("mesh_index"), py::arg("verts"), py::arg("velocities") = py::none(),
            "Update vertex positions and optional velocities of an already-added mesh (for animated colliders). "
            "verts must be a flat float32 array of shape [n_verts*3] or [n_verts, 3]. "
            "Call this before step() each frame where the mesh has moved.")
        .def_property_readonly("particle_count", [](const PySolver& s) { return s.n(); })
        .def_property_readonly("mesh_count",
                               [](const PySolver& s) {
                                   size_t v = 0;
                                   xpbd_mesh_count(s.h, &v);
                                   return v;
                               })
        .def_property_readonly("last_contact_count", [](const PySolver& s) { return s.count(); })
        .def("get_contact_particles",
             [](PySolver& s) {
                 size_t c = size_t(std::max(s.count(), 0));
                 py::array_t<int32_t> r(c);
                 s.check(xpbd_get_contact_particles(s.h, r.mutable_data(), c));
                 return r;
             })
        .def("get_contact_shapes",
             [](PySolver& s) {
                 size_t c = size_t(std::max(s.count(), 0));
                 py::array_t<int32_t> r(c);
                 s.check(xpbd_get_contact_shapes(s.h, r.mutable_data(), c));
                 return r;
             })
        .def("get_contact_normals",
             [](PySolver& s) {
                 size_t c = size_t(std::max(s.count(), 0));
                 auto r = vec3_array(c);
                 s.check(xpbd_get_contact_normals(s.h, r.mutable_data(), c));
                 return r;
             })
        .def("get_contact_body_positions",
             [](PySolver& s) {
                 size_t c = size_t(std::max(s.count(), 0));
                 auto r = vec3_array(c);
                 s.check(xpbd_get_contact_body_positions(s.h, r.mutable_data(), c));
                 return r;
             })
        .def("get_contact_body_velocities",
             [](PySolver& s) {
                 size_t c = size_t(std::max(s.count(), 0));
                 auto r = vec3_array(c);
                 s.check(xpbd_get_contact_body_velocities(s.h, r.mutable_data(), c));
                 return r;
             })
        .def("set_contact_data",
             [](PySolver& s, py::array_t<int32_t> particles, py::array_t<int32_t> shapes,
                py::array_t<float> body_positions, py::array_t<float> normals, int contact_count) {
                 const size_t c = size_t(std::max(contact_count, 0));
                 if (size_t(particles.size()) < c || size_t(shapes.size()) < c || size_t(body_positions.size()) < 3 * c ||
                     size_t(normals.size()) < 3 * c)
                     throw std::runtime_error("set_contact_data: arrays shorter than contact_count");
                 auto p = py::array_t<int32_t, py::array::c_style | py::array::forcecast>::ensure(particles);
                 auto sh = py::array_t<int32_t, py::array::c_style | py::array::forcecast>::ensure(shapes);
                 auto bp = farr::ensure(body_positions);
                 auto nn = farr::ensure(normals);
                 s.check(xpbd_set_contact_data(s.h, p.data(), sh.data(), bp.data(), nn.data(), contact_count));
             })
        .def("get_sorted_cells",
             [](PySolver& s) {
                 size_t n = s.n();
                 py::array_t<uint32_t> r(n);
                 s.check(xpbd_get_sorted_cells(s.h, r.mutable_data(), n));
                 return r;
             })
        .def("get_sorted_point_ids",
             [](PySolver& s) {
                 size_t n = s.n();
                 py::array_t<uint32_t> r(n);
                 s.check(xpbd_get_sorted_point_ids(s.h, r.mutable_data(), n));
                 return r;
             })
        .def("get_cell_starts",
             [](PySolver& s) {
                 py::array_t<int32_t> r(XPBD_NUM_CELLS);
                 s.check(xpbd_get_cell_starts(s.h, r.mutable_data(), XPBD_NUM_CELLS));
                 return r;
             })
        .def("get_cell_ends",
             [](PySolver& s) {
                 py::array_t<int32_t> r(XPBD_NUM_CELLS);
                 s.check(xpbd_get_cell_ends(s.h, r.mutable_data(), XPBD_NUM_CELLS));
                 return r;
             })
        .def("get_deltas",
             [](PySolver& s) {
                 size_t n = s.n();
                 auto r = vec3_array(n);
                 s.check(xpbd_get_deltas(s.h, r.mutable_data(), n));
                 return r;
             })
        .def("get_diag_pair_count",
             [](PySolver& s) {
                 uint32_t v = 0;
                 s.check(xpbd_get_diag_pair_count(s.h, &v));
                 return v;
             })
        .def("get_diag_overflow_count",
             [](PySolver& s) {
                 uint32_t v = 0;
                 s.check(xpbd_get_diag_overflow_count(s.h, &v));
                 return v;
             })
        .def("reset_diag_buffers", [](PySolver& s) { s.check(xpbd_reset_diag_buffers(s.h)); })
        .def("get_diag_pairs",
             [](PySolver& s) {
                 uint32_t count = 0;
                 s.check(xpbd_get_diag_pair_count(s.h, &count));
                 std::vector<xpbd_pair_record> recs(count);
                 size_t w = 0;
                 if (count) s.check(xpbd_get_diag_pairs(s.h, recs.data(), recs.size(), &w));
                 py::list out;
                 for (size_t k = 0; k < w; k++) {
                     const xpbd_pair_record& r = recs[k];
                     py::dict d;
                     d["i"] = r.i;
                     d["j"] = r.j;
                     d["dist"] = r.dist;
                     d["err"] = r.err;
                     d["normal"] = py::make_tuple(r.normal[0], r.normal[1], r.normal[2]);
                     d["vrel"] = py::make_tuple(r.vrel[0], r.vrel[1], r.vrel[2]);
                     d["vn"] = r.vn;
                     d["vt_len"] = r.vt_len;
                     d["denom"] = r.denom;
                     d["lambda_n"] = r.lambda_n;
                     d["lambda_f"] = r.lambda_f;
                     d["delta_i"] = py::make_tuple(r.delta_i[0], r.delta_i[1], r.delta_i[2]);
                     d["iter_index"] = r.iter_index;
                     out.append(d);
                 }
                 return out;
             })
        .def("get_p11_diag",
             [](PySolver& s) {
                 py::array_t<uint32_t> r(XPBD_DIAG_WORDS);
                 s.check(xpbd_get_p11_diag(s.h, r.mutable_data()));
                 return r;
             })
        .def("get_p11_apply_diag",
             [](PySolver& s) {
                 py::array_t<uint32_t> r(XPBD_DIAG_WORDS);
                 s.check(xpbd_get_p11_apply_diag(s.h, r.mutable_data()));
                 return r;
             })
        .def("get_shape_contact_diag",
             [](PySolver& s) {
                 size_t len = 0;
                 s.check(xpbd_get_shape_contact_diag(s.h, nullptr, 0, &len));
                 py::array_t<float> r(len);
                 s.check(xpbd_get_shape_contact_diag(s.h, r.mutable_data(), len, &len));
                 return r;
             })
        .def("get_mesh_candidate_diag",
             [](PySolver& s) {
                 py::array_t<uint32_t> r(XPBD_DIAG_WORDS);
                 s.check(xpbd_get_mesh_candidate_diag(s.h, r.mutable_data()));
                 return r;
             })
        .def("get_mesh_bvh_lowers",
             [](PySolver& s, size_t mesh_index) {
                 size_t nodes = 0;
                 s.check(xpbd_get_mesh_bvh_node_count(s.h, mesh_index, &nodes));
                 py::array_t<float> r({nodes, size_t(4)});
                 s.check(xpbd_get_mesh_bvh_lowers(s.h, mesh_index, r.mutable_data(), nodes));
                 return r;
             })
        .def("get_mesh_bvh_uppers",
             [](PySolver& s, size_t mesh_index) {
                 size_t nodes = 0;
                 s.check(xpbd_get_mesh_bvh_node_count(s.h, mesh_index, &nodes));
                 py::array_t<float> r({nodes, size_t(4)});
                 s.check(xpbd_get_mesh_bvh_uppers(s.h, mesh_index, r.mutable_data(), nodes));
                 return r;
             })
        .def("print_profiler_summary", [](PySolver& s) { s.check(xpbd_print_profiler_summary(s.h)); })
        .def("profile_stages",
             [](PySolver& s) {
                 size_t count = 0;
                 std::vector<char> names(4096, 0);
                 std::vector<double> ms(64, 0.0);
                 std::vector<uint64_t> launches(64, 0);
                 s.check(xpbd_profile_stages(s.h, names.data(), names.size(), ms.data(), launches.data(), 64, &count));
                 py::dict d;
                 const char* p = names.data();
                 for (size_t k = 0; k < count && k < 64; k++) {
                     std::string nm(p);
                     p += nm.size() + 1;
                     d[py::str(nm)] = py::make_tuple(ms[k], launches[k]);
                 }
                 return d;
             },
             "Per-stage accumulated GPU milliseconds and launch counts (enable_profile).")
        .def("profile_reset", [](PySolver& s) { s.check(xpbd_profile_reset(s.h)); })
        .def(
            "init_batch",
            [](PySolver& s, const CfgWrap& c, py::array_t<int64_t, py::array::c_style | py::array::forcecast> counts,
               size_t mc) {
                std::vector<uint64_t> cnt(size_t(counts.size()));
                for (size_t k = 0; k < cnt.size(); k++) {
                    if (counts.data()[k] < 0) throw std::runtime_error("init_batch: negative scene size");
                    cnt[k] = uint64_t(counts.data()[k]);
                }
                s.check(xpbd_init_batch(s.h, &c.c, cnt.data(), cnt.size(), mc));
            },
            py::arg("config"), py::arg("scene_counts"), py::arg("max_contacts") = 0,
            "Batched scenes (extension): scene_counts[k] particles per scene, concatenated in scene order.")
        .def(
            "rollout_forward",
            [](PySolver& s, uintptr_t x0, uintptr_t v0, float relax, float sdt, int substeps, int every, uintptr_t xo,
               uintptr_t vo) {
                int rc;
                {
                    py::gil_scoped_release nogil;
                    rc = xpbd_rollout_forward(s.h, reinterpret_cast<const float*>(x0), reinterpret_cast<const float*>(v0),
                                              relax, sdt, substeps, every, reinterpret_cast<float*>(xo),
                                              reinterpret_cast<float*>(vo));
                }
                s.check(rc);
            },
            py::arg("x0_ptr"), py::arg("v0_ptr"), py::arg("relax"), py::arg("sdt"), py::arg("substeps"),
            py::arg("checkpoint_every"), py::arg("x_out_ptr"), py::arg("v_out_ptr"),
            "Differentiable rollout forward on device pointers (0 = null); see vkxpbd.autograd.")
        .def(
            "rollout_backward",
            [](PySolver& s, uintptr_t gx, uintptr_t gv, uintptr_t gx0, uintptr_t gv0) {
                double gr = 0.0;
                int rc;
                {
                    py::gil_scoped_release nogil;
                    rc = xpbd_rollout_backward(s.h, reinterpret_cast<const float*>(gx), reinterpret_cast<const float*>(gv),
                                               reinterpret_cast<float*>(gx0), reinterpret_cast<float*>(gv0), &gr);
                }
                s.check(rc);
                return gr;
            },
            py::arg("gx_ptr"), py::arg("gv_ptr"), py::arg("gx0_ptr"), py::arg("gv0_ptr"),
            "Reverse pass of the last rollout_forward; returns d/d relax.")
        .def("mesh_query_stats",
             [](PySolver& s) {
                 uint64_t v[6];
                 s.check(xpbd_mesh_query_stats(s.h, v));
                 py::dict d;
                 const char* k[6] = {"pairs", "root_passes", "node_pops", "closest_tests", "parity_tests", "contacts"};
                 for (int i = 0; i < 6; i++) d[k[i]] = v[i];
                 return d;
             })
        .def_property_readonly("scene_count",
                               [](const PySolver& s) {
                                   size_t v = 0;
                                   xpbd_scene_count(s.h, &v);
                                   return v;
                               })
        .def(
            "set_wrap_planes", [](PySolver& s, bool on) { s.check(xpbd_set_wrap_planes(s.h, on ? 1 : 0)); },
            py::arg("enabled"),
            "Sort by wrap plane and cell, so queries skip particles 128 cells away (default on from 262,144 "
            "particles; same results).")
        .def_property_readonly("wrap_planes",
                               [](const PySolver& s) {
                                   int v = 0;
                                   xpbd_wrap_planes(s.h, &v);
                                   return v != 0;
                               })
        .def(
            "set_graph_mode", [](PySolver& s, bool on) { s.check(xpbd_set_graph_mode(s.h, on ? 1 : 0)); },
            py::arg("enabled"), "Capture and replay whole frames as hipGraphs (default on).")
        .def(
            "set_deterministic", [](PySolver& s, bool on) { s.check(xpbd_set_deterministic(s.h, on ? 1 : 0)); },
            py::arg("enabled"),
            "Take every sum the reference takes with float atomics (shape deltas, body deltas, restitution) in a "
            "canonical order, so results do not depend on scheduling (default off).")
        .def_property_readonly("cell_table_bytes",
                               [](const PySolver& s) {
                                   size_t v = 0;
                                   xpbd_cell_table_bytes(s.h, &v);
                                   return v;
                               })
        .def_property_readonly("radix_tile",
                               [](const PySolver& s) {
                                   int v = 0;
                                   xpbd_radix_tile(s.h, &v);
                                   return v;
                               })
        .def_property_readonly("pp_block",
                               [](const PySolver& s) {
                                   int v = 0;
                                   xpbd_pp_block(s.h, &v);
                                   return v;
                               })
        .def(
            "set_tuning", [](PySolver& s, int pp_block, int radix_tile) { s.check(xpbd_set_tuning(s.h, pp_block, radix_tile)); },
            py::arg("pp_block"), py::arg("radix_tile"),
            "Set the launch shapes autotune picks (pp_block 64/128/256, radix_tile 1024/2048); results do not depend "
            "on them.")
        .def_property_readonly("deterministic",
                               [](const PySolver& s) {
                                   int v = 0;
                                   xpbd_deterministic(s.h, &v);
                                   return v != 0;
                               })
        .def(
            "autotune",
            [](PySolver& s, int frames, float dt, int substeps, int iterations) {
                int b = 0, rc;
                {
                    py::gil_scoped_release nogil;
                    rc = xpbd_autotune(s.h, frames, dt, substeps, iterations, &b);
                }
                s.check(rc);
                return b;
            },
            py::arg("frames") = 3, py::arg("dt") = 1.0f / 60.0f, py::arg("substeps") = 8, py::arg("iterations") = 1,
            "Time the particle-particle block sizes x radix tiles on the current state (state is restored), keep the "
            "fastest and return its block size.")
        .def(
            "add_box_body",
            [](PySolver& s, farr pos, farr quat, farr half_extents, float density, int flags) {
                if (pos.size() < 3 || quat.size() < 4 || half_extents.size() < 3)
                    throw std::runtime_error("add_box_body: pos[3], quat[4] (xyzw), half_extents[3] expected");
                s.check(xpbd_add_box_body(s.h, pos.data(), quat.data(), half_extents.data(), density, flags));
            },
            py::arg("pos"), py::arg("quat"), py::arg("half_extents"), py::arg("density") = 1000.0f, py::arg("flags") = 0,
            "Add a dynamic rigid box (extension: the reference wires no body path).")
        .def(
            "set_body_params",
            [](PySolver& s, float ad, float relax, float mt, float mr, bool uw) {
                s.check(xpbd_set_body_params(s.h, ad, relax, mt, mr, uw ? 1 : 0));
            },
            py::arg("angular_damping") = 0.05f, py::arg("rigid_contact_relaxation") = 0.8f,
            py::arg("mu_torsional") = 0.0f, py::arg("mu_rolling") = 0.0f, py::arg("use_constraint_weights") = true)
        .def_property_readonly("body_count",
                               [](const PySolver& s) {
                                   size_t v = 0;
                                   xpbd_body_count(s.h, &v);
                                   return v;
                               })
        .def("get_body_q",
             [](PySolver& s) {
                 size_t nb = 0;
                 xpbd_body_count(s.h, &nb);
                 py::array_t<float> r({nb, size_t(7)});
                 s.check(xpbd_get_bodies(s.h, r.mutable_data(), nullptr, nb));
                 return r;
             })
        .def("get_body_qd",
             [](PySolver& s) {
                 size_t nb = 0;
                 xpbd_body_count(s.h, &nb);
                 py::array_t<float> r({nb, size_t(6)});
                 s.check(xpbd_get_bodies(s.h, nullptr, r.mutable_data(), nb));
                 return r;
             })
        .def("set_body_velocities", [](PySolver& s, farr qd) {
            s.check(xpbd_set_body_velocities(s.h, qd.data(), size_t(qd.size()) / 6));
        });
}
