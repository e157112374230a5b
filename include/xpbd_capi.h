/*
 * xpbd_capi.h -- C-ABI of the MI355X XPBD particle solver (libvkxpbd.so).
 *
 * This is the drop-in boundary for the reference's GPU step: every entry point
 * replaces one method of vkx::XpbdSolver (src/scene.h:74-227) or one free
 * function of the vkxpbd pybind11 module (src/bindings.cpp:8-301).  The
 * Python module `vkxpbd` in this repository binds exactly these symbols.
 *
 * Conventions
 *   - Plain pointers and sizes; no C++ or torch types cross the ABI.
 *   - Every call returns int status: 0 = ok, nonzero = error.  The message is
 *     available from xpbd_last_error(h) and carries the reference's text where
 *     the reference throws (e.g. "Particle count mismatch").
 *   - Host arrays are copied in before the call returns (the reference's
 *     synchronous staging, src/buffer.h:80-108); getters copy out into
 *     caller-provided buffers whose capacity (in elements) is passed.
 *   - Vectors are packed xyz float32 (3 per element) on the host side.
 *   - Calls on one handle are not thread-safe (as the reference,
 *     src/buffer.h:110-139).  Distinct handles are independent.
 */
#ifndef XPBD_CAPI_H
#define XPBD_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XPBD_NUM_CELLS 2097152 /* 128^3 hash-grid cells, src/scene.cpp:17-19 */
#define XPBD_DIAG_WORDS 32     /* diagnostic word buffers, src/scene.cpp:120-128 */

/* Mirrors vkx::SolverConfig (src/scene.h:14-41) field for field, minus the
 * std::string shader_dir, which the Python layer keeps (no shaders here). */
typedef struct xpbd_config {
    float dt;
    int32_t substeps;
    int32_t iterations;
    float particle_v_max;
    float soft_contact_mu;
    float shape_material_mu;
    float soft_contact_relaxation;
    float particle_adhesion;
    float particle_mu;
    float particle_cohesion;
    float particle_max_radius;
    float search_radius;
    float soft_contact_margin;
    int32_t enable_restitution;
    float soft_contact_restitution;
    int32_t has_ground_plane;
    float ground_plane_altitude;
    float gravity[3];
    int32_t device_index;
    int32_t compat_mode;
    int32_t enable_validation;
    int32_t enable_diagnostics;
    int32_t enable_profile;
    uint32_t max_diag_pairs;
    int32_t substep_batch_size;
} xpbd_config;

/* One particle-particle diagnostic record, 80 bytes: vkx::GPUPairRecord
 * (src/scene.h:43-64). */
typedef struct xpbd_pair_record {
    uint32_t i, j;
    float dist, err;
    float normal[3];
    float vrel[3];
    float vn, vt_len, denom, lambda_n, lambda_f;
    float delta_i[3];
    uint32_t iter_index, pad0;
} xpbd_pair_record;

typedef struct xpbd_solver xpbd_solver;

/* Default values of vkx::SolverConfig (src/scene.h:14-41). */
void xpbd_config_default(xpbd_config* cfg);

/* enumerate_devices() (src/bindings.cpp:48; src/context.h:38-69): number of
 * HIP devices and the name of one of them. */
int xpbd_device_count(int* count);
int xpbd_device_name(int index, char* buf, size_t buflen);

/* vkx::XpbdSolver::XpbdSolver / ~XpbdSolver (src/scene.cpp:27-45). */
int xpbd_create(xpbd_solver** out);
void xpbd_destroy(xpbd_solver* h);
/* Message of the last failed call on h (empty if none).  h may be NULL, then
 * the message of the last failed xpbd_create/xpbd_device_* call is returned. */
const char* xpbd_last_error(const xpbd_solver* h);

/* XpbdSolver::init (src/scene.cpp:65-152).  max_contacts 0 means 2*n. */
int xpbd_init(xpbd_solver* h, const xpbd_config* cfg, size_t num_particles, size_t max_contacts);

/* XpbdSolver::add_mesh (src/scene.cpp:154-202).  rot_xyzw may be NULL
 * (identity); watertight_override -1 = auto-detect, 0 = false, 1 = true. */
int xpbd_add_mesh(xpbd_solver* h, const float* verts, size_t num_verts, const uint32_t* tris, size_t num_tris,
                  float shape_margin, const float* rot_xyzw, int watertight_override);

/* XpbdSolver::finalize_meshes (src/scene.cpp:204-390). */
int xpbd_finalize_meshes(xpbd_solver* h);

/* XpbdSolver::set_particles (src/scene.cpp:592-615).  qd NULL uploads zeros;
 * inv_mass / radii / flags NULL keep the previous device contents. */
int xpbd_set_particles(xpbd_solver* h, const float* q, const float* qd, const float* inv_mass, const float* radii,
                       const int32_t* flags, size_t n);

/* XpbdSolver::update_mesh (src/scene.cpp:617-696).  vels may be NULL (zero). */
int xpbd_update_mesh(xpbd_solver* h, size_t mesh_index, const float* verts, size_t num_verts, const float* vels);

/* XpbdSolver::step (src/scene.cpp:698-1092). */
int xpbd_step(xpbd_solver* h, float dt, int substeps, int iterations);
/* XpbdSolver::collide_only (src/scene.cpp:1229-1245). */
int xpbd_collide_only(xpbd_solver* h);
/* XpbdSolver::step_solve_only (src/scene.cpp:1094-1227). */
int xpbd_step_solve_only(xpbd_solver* h, float dt, int substeps, int iterations);

/* XpbdSolver::get_positions / get_velocities (src/scene.cpp:1247-1267):
 * out holds n*3 floats; n is clamped to the particle count. */
int xpbd_get_positions(xpbd_solver* h, float* out, size_t n);
int xpbd_get_velocities(xpbd_solver* h, float* out, size_t n);

/* particle_count / mesh_count / last_contact_count (src/scene.h:101-103). */
int xpbd_particle_count(const xpbd_solver* h, size_t* out);
int xpbd_mesh_count(const xpbd_solver* h, size_t* out);
int xpbd_last_contact_count(const xpbd_solver* h, int* out);

/* get_contact_* (src/scene.cpp:1269-1302): the first last_contact_count
 * entries; cap is the capacity of out in contacts. */
int xpbd_get_contact_particles(xpbd_solver* h, int32_t* out, size_t cap);
int xpbd_get_contact_shapes(xpbd_solver* h, int32_t* out, size_t cap);
int xpbd_get_contact_normals(xpbd_solver* h, float* out, size_t cap);
int xpbd_get_contact_body_positions(xpbd_solver* h, float* out, size_t cap);
int xpbd_get_contact_body_velocities(xpbd_solver* h, float* out, size_t cap);

/* XpbdSolver::set_contact_data (src/scene.cpp:1304-1315).  Body velocities
 * are not injected, as in the reference. */
int xpbd_set_contact_data(xpbd_solver* h, const int32_t* particles, const int32_t* shapes, const float* body_positions,
                          const float* normals, int contact_count);

/* Grid debug getters (src/scene.cpp:1317-1348).  sorted cells / ids / deltas
 * hold n (n*3 for deltas) elements; cell tables hold XPBD_NUM_CELLS. */
int xpbd_get_sorted_cells(xpbd_solver* h, uint32_t* out, size_t n);
int xpbd_get_sorted_point_ids(xpbd_solver* h, uint32_t* out, size_t n);
int xpbd_get_cell_starts(xpbd_solver* h, int32_t* out, size_t n);
int xpbd_get_cell_ends(xpbd_solver* h, int32_t* out, size_t n);
int xpbd_get_deltas(xpbd_solver* h, float* out, size_t n);

/* Diagnostics (src/scene.cpp:1350-1414). */
int xpbd_get_diag_pair_count(xpbd_solver* h, uint32_t* out);
int xpbd_get_diag_overflow_count(xpbd_solver* h, uint32_t* out);
/* Copies min(count, max_diag_pairs, cap) records; *written receives the number. */
int xpbd_get_diag_pairs(xpbd_solver* h, xpbd_pair_record* out, size_t cap, size_t* written);
int xpbd_reset_diag_buffers(xpbd_solver* h);
/* 32-word diagnostic buffers (get_p11_diag, get_p11_apply_diag,
 * get_mesh_candidate_diag). */
int xpbd_get_p11_diag(xpbd_solver* h, uint32_t* out32);
int xpbd_get_p11_apply_diag(xpbd_solver* h, uint32_t* out32);
int xpbd_get_mesh_candidate_diag(xpbd_solver* h, uint32_t* out32);
/* get_shape_contact_diag: *len receives the float count; out may be NULL to
 * query the length only. */
int xpbd_get_shape_contact_diag(xpbd_solver* h, float* out, size_t cap, size_t* len);

/* get_mesh_bvh_lowers / uppers (src/scene.cpp:1416-1436): nodes*4 floats,
 * w = packed u32 bits.  *nodes receives the node count (0 for a bad index). */
int xpbd_get_mesh_bvh_node_count(xpbd_solver* h, size_t mesh_index, size_t* nodes);
int xpbd_get_mesh_bvh_lowers(xpbd_solver* h, size_t mesh_index, float* out, size_t cap_nodes);
int xpbd_get_mesh_bvh_uppers(xpbd_solver* h, size_t mesh_index, float* out, size_t cap_nodes);

/* XpbdSolver::print_profiler_summary (src/scene.cpp:1438-1515). */
int xpbd_print_profiler_summary(xpbd_solver* h);

/* Per-stage GPU time accumulated by the profiler (HIP events on the solver's
 * stream; enable_profile or VKXPBD_PROFILE=1).  No reference counterpart as a
 * separate call: it exposes the numbers print_profiler_summary prints.
 * names receives up to cap NUL-separated stage names (may be NULL), ms and
 * launches receive per-stage totals; *count receives the stage count. */
int xpbd_profile_stages(xpbd_solver* h, char* names, size_t names_len, double* ms, uint64_t* launches, size_t cap,
                        size_t* count);
int xpbd_profile_reset(xpbd_solver* h);

/* Dynamic rigid boxes.  No counterpart in the reference's API: its body
 * shaders (integrate_bodies.comp, solve_body_contact_positions.comp,
 * apply_body_deltas.comp) are never dispatched and nothing creates dynamic
 * bodies.  Each box becomes a body and a shape (after the meshes); particles
 * collide with it (particle-shape contacts) and its corners with the ground
 * plane (rigid contacts).  pos xyz, quat xyzw, half extents xyz, density
 * kg/m^3, flags (2 = kinematic).  Call before finalize_meshes/step. */
int xpbd_add_box_body(xpbd_solver* h, const float* pos, const float* quat, const float* half_extents, float density,
                      int flags);
/* Body solver parameters (defaults 0.05, 0.8, 0, 0, 1): angular damping of
 * integrate_bodies, relaxation of the rigid contacts, torsional and rolling
 * friction, and whether apply_body_deltas divides by the rigid-contact count. */
int xpbd_set_body_params(xpbd_solver* h, float angular_damping, float rigid_relaxation, float mu_torsional,
                         float mu_rolling, int use_constraint_weights);
int xpbd_body_count(const xpbd_solver* h, size_t* out);
/* Body state: q7 = (px, py, pz, qx, qy, qz, qw), qd6 = (vx, vy, vz, wx, wy, wz)
 * per box in add order; either pointer may be NULL; cap in bodies. */
int xpbd_get_bodies(xpbd_solver* h, float* q7, float* qd6, size_t cap);
int xpbd_set_body_velocities(xpbd_solver* h, const float* qd6, size_t n);

/* Differentiable rollout (extension; the reference has no gradient path).
 * Runs `substeps` substeps of length sdt (one iteration each) from positions
 * x0 and velocities v0 with contact relaxation `relax`, keeping state
 * checkpoints every `checkpoint_every` substeps.  x0, v0, x_out, v_out are
 * DEVICE pointers to 3 floats per particle on the solver's device (x0/v0 NULL:
 * current state; outputs may be NULL).  The solver's state afterwards is the
 * final state.  Dynamic boxes and restitution are not supported. */
int xpbd_rollout_forward(xpbd_solver* h, const float* x0, const float* v0, float relax, float sdt, int substeps,
                         int checkpoint_every, float* x_out, float* v_out);
/* Reverse pass of the last rollout_forward: cotangents of the final positions
 * gx (required) and velocities gv (NULL = zero) -> cotangents of x0 and v0
 * (device pointers, NULL to skip) and of the relaxation (*grelax, host).  The
 * contact set of every substep is the forward's (frozen).  Restores the
 * forward's final state. */
int xpbd_rollout_backward(xpbd_solver* h, const float* gx, const float* gv, float* gx0, float* gv0, double* grelax);

/* Mesh narrowphase counters accumulated while the environment variable
 * VKXPBD_MESH_STATS=1 was set at init, since the last xpbd_profile_reset: out[0] active
 * particle x mesh pairs, [1] particles needing the tree search, [2] BVH node pops, [3]
 * closest-point triangle tests, [4] parity ray/triangle tests, [5] contacts.
 * Extension (the reference's mesh diagnostics buffer is never written). */
int xpbd_mesh_query_stats(xpbd_solver* h, uint64_t* out6);

/* Batched scenes: num_scenes independent particle scenes (counts[k] particles
 * each, concatenated in scene order in every particle array) advanced together
 * with one launch per stage.  Each scene has its own hash grid (the sort key
 * carries the scene index above the 21 cell bits) and sees the ground plane
 * and all meshes/boxes.  At most 2048 scenes.  Extension without a reference
 * counterpart: the reference steps one scene per XpbdSolver
 * (src/scene.cpp:65-152 is the single-scene init it generalises). */
int xpbd_init_batch(xpbd_solver* h, const xpbd_config* cfg, const uint64_t* counts, size_t num_scenes,
                    size_t max_contacts);
int xpbd_scene_count(const xpbd_solver* h, size_t* out);
/* Wrap planes: the hash grid wraps every 128 cells, so particles 128 cells
 * apart share a cell and every neighbour query scans them.  With wrap planes a
 * plain solver sorts by wrap plane << 21 | cell (two bits per axis of which
 * wrap a cell lies in) and queries read only their own wrap's run: fewer
 * candidates for scenes wider than 128 cells, one more radix pass (27-bit
 * keys) and a 1 GiB cell table.  xpbd_set_wrap_planes asks for the mode
 * (default: on from 262,144 particles; environment VKXPBD_WRAP_PLANES=1 or 0
 * forces it at init); xpbd_wrap_planes reports whether it is in effect: never
 * for a batched solver, nor while a pair could reach
 * 126 cells (3 r_max + particle_max_radius + 2|cohesion| >= 126 search_radius,
 * re-checked by set_particles).  Results are identical either way; the
 * sorted-cell and cell-table getters report the reference's order in both.
 * No reference counterpart (src/scene.cpp:885-914 sorts plain cell keys). */
int xpbd_set_wrap_planes(xpbd_solver* h, int on);
int xpbd_wrap_planes(const xpbd_solver* h, int* out);
/* Frame capture: when on (default; environment VKXPBD_GRAPHS=0 turns it off),
 * xpbd_step records each distinct frame once as a hipGraph and replays it.
 * Results are identical either way.  No reference counterpart (the reference
 * re-records its command buffer every frame, src/scene.cpp:805-992). */
int xpbd_set_graph_mode(xpbd_solver* h, int on);
/* Deterministic mode (no counterpart in the reference, whose float atomics
 * make such sums depend on scheduling): every sum the reference takes with
 * float atomics -- a particle's shape deltas and restitution updates, a
 * body's deltas -- is taken in a canonical order instead (a particle's
 * contacts by shape index; a box's particle contacts by particle id, then
 * its corner contacts by corner), the order of a sequential pass over the
 * contacts as the generators emit them.  Results then do not depend on
 * scheduling.  Off by default (VKXPBD_DETERMINISTIC=1 at init turns it on);
 * costs two radix sorts of the contact buffer per substep. */
int xpbd_set_deterministic(xpbd_solver* h, int on);
int xpbd_deterministic(const xpbd_solver* h, int* out);
/* Times the candidate launch shapes -- particle-particle block size (64,
 * 128, 256) x radix tile (1,024 or 2,048 keys per workgroup of the sort that
 * builds the cell tables) -- over `frames` frames of the current state,
 * restores the state, keeps the fastest and returns its block size in *block
 * (may be NULL).  Results do not depend on the choice. */
int xpbd_autotune(xpbd_solver* h, int frames, float dt, int substeps, int iterations, int* block);
/* Keys per workgroup of the radix sort currently in use (1,024 or 2,048). */
int xpbd_radix_tile(const xpbd_solver* h, int* out);
/* Particle-particle block size currently in use (64, 128 or 256). */
int xpbd_pp_block(const xpbd_solver* h, int* out);
/* Sets the launch shapes xpbd_autotune picks -- pp_block 64, 128 or 256,
 * radix_tile 1,024 or 2,048 -- e.g. to reuse one tuning for solvers of the
 * same workload (no counterpart in the reference).  Results do not depend
 * on the choice. */
int xpbd_set_tuning(xpbd_solver* h, int pp_block, int radix_tile);
/* Device bytes held by the cell-range tables: the dense 2^21-row table (2^24
 * or 2^27 rows with wrap planes) of a plain solver, or the row table of a
 * batched solver (8 bytes per grid row of each scene, 2^14 rows per scene,
 * plus 8 lower bounds per occupied group of 8 cells, at most 36 bytes per
 * particle). */
int xpbd_cell_table_bytes(const xpbd_solver* h, size_t* out);

/* Host-only LBVH build of one mesh, the builder finalize_meshes uses
 * (MeshBvh::build, src/mesh_bvh.h:40-69 -> src/mesh_lbvh.h:49-223).  Needs no
 * GPU.  lowers/uppers receive 2*num_tris packed nodes (4 floats each, w =
 * packed u32 bits), prim num_tris primitive ids, *root the root node. */
int xpbd_lbvh_build(const float* verts, size_t num_verts, const uint32_t* tris, size_t num_tris, float* lowers,
                    float* uppers, uint32_t* prim, uint32_t* root);
/* Build over verts, then refit to new_verts (MeshBvh::refit,
 * src/mesh_bvh.h:142-181).  Host only. */
int xpbd_lbvh_refit(const float* verts, const float* new_verts, size_t num_verts, const uint32_t* tris,
                    size_t num_tris, float* lowers, float* uppers);
/* Host only: builds the projected ray-parity grid the watertight sign test
 * uses and returns, per query point, the triangles registered in the point's
 * cell: counts[k] for point k, ids concatenated in point order up to
 * ids_cap entries (*total receives the full length).  Exists so the grid's
 * conservativeness (every triangle the ray test hits is a candidate) can be
 * checked without a GPU. */
int xpbd_parity_grid_candidates(const float* verts, size_t num_verts, const uint32_t* tris, size_t num_tris,
                                const float* points, size_t num_points, uint32_t* counts, uint32_t* ids,
                                size_t ids_cap, size_t* total);
/* Host only: builds the distance lower-bound grid of a mesh for the given cap
 * and returns, per point, the bound the mesh query uses for it (-1 for a
 * point outside the grid, which is farther than cap from every triangle).
 * Exists so the bounds can be checked against exact distances without a GPU. */
int xpbd_dist_grid_bounds(const float* verts, size_t num_verts, const uint32_t* tris, size_t num_tris, float cap,
                          const float* points, size_t num_points, float* bounds);

/* Stable LSD radix sort of n (key, value) pairs on caller-owned device
 * buffers of the current HIP device: (keys_out, vals_out) receive the pairs
 * ordered by key, equal keys in input order.  Keys must be below
 * 2^key_bits (key_bits in [1, 32]; the solver's cell keys take 21, or 24
 * with one wrap-plane bit per axis).  The inputs are not written and the
 * outputs must not overlap them.  Allocates its scratch (temporary pairs,
 * tile histograms, bin totals) for the call and returns once the sort has
 * completed on `stream` (a hipStream_t; NULL is the default stream).  Errors
 * are reported through xpbd_last_error(NULL).
 * Replaces GpuSort::create/record (src/gpu_sort.h:18-129), the reusable sort
 * the reference runs over (cell, particle id) pairs every substep. */
int xpbd_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                    size_t n, int key_bits, void* stream);

/* xpbd_sort_pairs without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as xpbd_sort_pairs, for at most 1,048,576 pairs
 * (larger n is rejected with an error); the temporaries live in a 9 MiB
 * static device array of the library, one per device it runs on, held until
 * the process ends.  Launches on `stream` and returns without waiting.  The
 * first call on a device looks up that array's address (a call with n = 0
 * does only this, e.g. before a capture).  Every call on a device shares
 * the array: calls that overlap on one device (other streams, other
 * threads) race. */
int xpbd_sort_pairs_noalloc(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                            uint32_t* vals_out, size_t n, int key_bits, void* stream);

/* Self-check of the device arithmetic the contact projection relies on; no
 * counterpart in the reference.  The solver takes correctly rounded
 * reciprocals and square roots through shorter gfx950 sequences (v_rcp_f32
 * with one Newton step; v_sqrt_f32 with a neighbour correction) inside the
 * ranges where they equal the compiler's correctly rounded 1.0f / d and
 * sqrtf(a).  This compares them there on every float, on the current HIP
 * device: out3[0] = reciprocal mismatches over |d| in [2^-126, 2^126),
 * out3[1] = square-root mismatches over a in [2^-96, +inf), out3[2] = at
 * a = +0.  All three are zero on a device where the solver's results hold. */
int xpbd_check_arith(uint64_t* out3);

#ifdef __cplusplus
}
#endif
#endif /* XPBD_CAPI_H */
