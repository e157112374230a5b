// solver.h -- host orchestrator of the MI355X XPBD step: owns device memory on
// one HIP device and one stream, and records the substep pipeline
// (src/scene.h:74-227 is the interface it reproduces).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/xpbd_capi.h"
#include "lbvh.h"
#include "xpbd_kernels.h"

namespace xpbd {

// Device allocation owned by the solver (freed in the destructor).
template <class T> struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
};

class Solver {
public:
    Solver();
    ~Solver();

    void init(const xpbd_config& cfg, size_t n, size_t max_contacts);
    // Batched mode: num_scenes independent particle scenes (counts[k]
    // particles each, concatenated in order), each with its own hash grid and
    // ground plane, advanced together (one launch per stage for all scenes).
    void init_batch(const xpbd_config& cfg, const size_t* counts, size_t num_scenes, size_t max_contacts);
    size_t scene_count() const { return num_scenes_; }
    bool wrap_planes() const { return planes_; }
    // hipGraph capture of whole frames in step() (default on; VKXPBD_GRAPHS=0 disables)
    void set_graph_mode(bool on) { use_graphs_ = on; invalidate_graphs(); }
    void set_wrap_planes(bool on) { planes_allowed_ = on; update_planes(); }
    // Deterministic mode: every float sum the reference takes with atomics
    // (shape deltas, body deltas, restitution) in a canonical order instead
    // (VKXPBD_DETERMINISTIC=1 at init also turns it on).
    void set_deterministic(bool on) { det_ = on; invalidate_graphs(); }
    bool deterministic() const { return det_; }
    // Times candidate launch shapes (particle-pair block size x radix tile)
    // on the current state (then restores it) and keeps the fastest; returns
    // the chosen particle-particle block size.
    int autotune(int frames, float dt, int substeps, int iterations);
    int radix_tile() const { return 256 * sort_items_; }
    int pp_block() const { return pp_block_; }
    // sets what autotune() picks: pp_block 64, 128 or 256; radix_tile 1,024 or 2,048
    void set_tuning(int pp_block, int radix_tile);
    // device bytes of the cell-range tables (dense rows, or the batched row
    // table, group bounds and tile scratch)
    size_t cell_table_bytes() const {
        return range_.n * sizeof(int32_t) + rkeys_.n * sizeof(uint32_t) + rows_.n * sizeof(uint2) +
               glb_.n * sizeof(uint32_t) + rslot_.n * sizeof(int32_t) + (gtcnt_.n + gtoff_.n) * sizeof(uint32_t);
    }
    void add_mesh(const float* v, size_t nv, const uint32_t* t, size_t nt, float margin, const float* rot, int wt);
    void finalize_meshes();
    void set_particles(const float* q, const float* qd, const float* w, const float* r, const int32_t* f, size_t n);
    void update_mesh(size_t mi, const float* v, size_t nv, const float* vel);
    void step(float dt, int substeps, int iterations);
    void collide_only();
    void step_solve_only(float dt, int substeps, int iterations);

    void get_positions(float* out, size_t n);
    void get_velocities(float* out, size_t n);
    size_t particle_count() const { return n_; }
    size_t mesh_count() const { return meshes_.size(); }
    int last_contact_count() const { return last_count_; }

    void get_contact_ints(bool shapes, int32_t* out, size_t cap);
    void get_contact_vec(int which, float* out, size_t cap);   // 0 normals, 1 body pos, 2 body vel
    void set_contact_data(const int32_t* p, const int32_t* s, const float* bp, const float* nrm, int count);
    void get_sorted(bool ids, uint32_t* out, size_t n);
    void get_cell_table(bool ends, int32_t* out, size_t n);
    void get_deltas(float* out, size_t n);

    uint32_t diag_word(int k);
    size_t get_diag_pairs(xpbd_pair_record* out, size_t cap);
    void reset_diag();
    size_t shape_contact_diag_len() const;

    size_t bvh_nodes(size_t mi) const { return mi < bvh_.size() ? bvh_[mi].lo.size() : 0; }
    void bvh_half(size_t mi, bool upper, float* out, size_t cap) const;

    // Dynamic box bodies (builder extension; the reference wires no body path).
    void add_box_body(const float* pos, const float* quat, const float* half, float density, int flags);
    void set_body_params(float angular_damping, float rigid_relax, float mu_t, float mu_r, int use_weights);
    size_t body_count() const { return boxes_.size(); }
    void get_bodies(float* q7, float* qd6, size_t cap);
    void set_body_velocities(const float* qd6, size_t n);

    // Differentiable rollout (one iteration per substep, contact set frozen
    // from the forward pass in the reverse pass).  Device pointers hold 3
    // floats per particle; x0/v0 may be null (current state).  The forward
    // keeps state checkpoints every `checkpoint_every` substeps; the backward
    // recomputes each segment from its checkpoint and runs the reverse
    // kernels, then restores the forward's final state.
    void rollout_forward(const float* x0, const float* v0, float relax, float sdt, int substeps, int checkpoint_every,
                         float* x_out, float* v_out);
    void rollout_backward(const float* gx, const float* gv, float* gx0, float* gv0, double* grelax);

    void print_profiler_summary();
    size_t profile_stages(char* names, size_t names_len, double* ms, uint64_t* launches, size_t cap);
    void profile_reset();
    // mesh query counters (environment VKXPBD_MESH_STATS=1): active particle x mesh
    // pairs, root-box passes, BVH node pops, closest-point tests, parity
    // triangle tests, contacts
    void mesh_query_stats(uint64_t out[6]);

private:
    struct MeshHost {
        std::vector<float> verts;
        std::vector<uint32_t> tris;
        float margin = 0.f;
        float rot[4] = {0.f, 0.f, 0.f, 1.f};
        bool watertight = false;
    };
    struct BoxHost {
        float pos[3], rot[4], half[3], density;
        int flags;
    };
    enum Stage {
        kMesh, kIntegrate, kSortUp, kSortScan, kSortScatter, kRanges, kShape, kPP, kRest, kCopyOut,
        kBodyContacts, kBodyIntegrate, kBodySolve, kBodyApply, kShapeApply, kNumStages
    };

    void require_init() const;
    void init_common(const xpbd_config& cfg, size_t n, size_t max_contacts);
    void invalidate_graphs();
    void record_frame(float dt, int substeps, int iterations);
    void alloc_all();
    void free_all();
    void upload_nodes();
    void free_tape();
    void swap_tape(int j);
    void ensure_tape();
    void build_query_grids(size_t mesh);
    void upload_tri_soup();
    uint32_t node_word(size_t mesh, uint32_t node) const;
    void upload_query_grids();
    // animated colliders (mesh_update.cpp): static tables for refitting and
    // rebuilding a mesh's derived data on the device; the in-place update
    // (false when a grid outgrew its room); the host rebuild of every mesh's
    // derived data (finalize's tail, after a refit of every BVH)
    void prepare_mesh_updates();
    bool update_mesh_on_device(size_t mesh);
    void rebuild_meshes_on_host();
    BodyArgs body_args() const;
    // fuse: kKeyed -- this substep's radix input was written by the previous
    // one (k_integrate skipped, its resets done by the first upsweep);
    // kKeyNext -- the last iteration writes the next substep's radix input
    static constexpr int kKeyed = 1, kKeyNext = 2;
    void substep(float sdt, int iterations, bool collide, int slot, bool zero_next, int fuse = 0);
    bool fuse_keys_ok(int iterations) const;
    int frame_fuse(int sub, int substeps, int iterations) const;
    BodyStepArgs body_step_args(float sdt);
    DetApplyArgs det_apply_args(const float4* Xin, const float4* Vin, int out, float sdt);
    // the state's move into the sorted order (integrated over sdt), taken
    // before the sort flips the layout
    SortGather sort_gather(float sdt);
    // last: the last pass also moves the state; rs: resets done by the first upsweep
    void run_sort(int tiles, const SortGather* last, const SortReset& rs = SortReset{});
    void update_planes();
    std::vector<uint64_t> reference_sorted();
    // cell ranges of the sorted keys; cleared = the previous rows are already
    // zero (k_integrate clears them when wrap planes are on)
    void build_ranges(const uint32_t* sorted_keys, bool cleared = false, const SortGather* gather = nullptr);
    void read_count(int slot);
    void ensure_det();   // deterministic mode's buffers (outside any graph capture)
    void stage_begin(Stage s);
    void stage_end(Stage s);
    void sync_profile();

    xpbd_config cfg_{};
    bool initialized_ = false, finalized_ = false, grid_valid_ = false;
    int device_ = 0;
    hipStream_t stream_ = nullptr;
    // mesh contacts of a substep run on side_ (default priority, like stream_)
    // alongside integrate, sort and cell ranges, which do not touch them; the
    // contact remap after the sort waits for them
    hipStream_t side_ = nullptr;
    hipEvent_t fork_ev_ = nullptr, join_ev_ = nullptr, sorted_ev_ = nullptr;
    size_t n_ = 0, cmax_ = 0;

    // batched scenes
    size_t num_scenes_ = 1;
    std::vector<size_t> scene_off_;
    DevBuf<int32_t> scene_d_;
    int sort_passes_ = 3, sort_bits_ = 7;
    void set_sort_shape(int key_bits);
    // plain solver: sort keys carry the wrap plane above the cell (see update_planes)
    static constexpr size_t kPlanesMinParticles = size_t(1) << 18;
    bool planes_ = false, planes_allowed_ = false;
    int plane_bits_ = 2;   // wrap-plane bits per axis (1 or 2)
    bool plane_bits_fixed_ = false;   // set by VKXPBD_PLANE_BITS; else chosen from the particles' extent
    int table_bits_ = -1;             // plane bits the cell table and sort were shaped for
    // particle-pair block size (64, 128 or 256; autotune() picks per scene):
    // 64 measured fastest on the 1M-particle pile (2.31 against 2.38 ms per frame)
    int pp_block_ = 64;
    int sort_items_ = 8;   // keys per radix thread: 4 (1,024-key tiles) or 8 (2,048); autotune() may pick
    // captured frames: rotation state and parameters -> executable graph
    struct FrameGraph {
        int s, p, o;
        float dt;
        int substeps, iterations;
        int lay;
        int s2, p2, o2, lay2;
        hipGraphExec_t exec;
    };
    std::vector<FrameGraph> graphs_;
    float relax_ = 0.f;   // contact relaxation of the solves (a rollout may override the config's)
    // rollout tape: checkpoints, per-substep records of the segment being
    // reversed, and the reverse-pass buffers
    struct Tape {
        bool valid = false;
        float sdt = 0.f, relax = 0.f;
        int substeps = 0, every = 0;
        std::vector<DevBuf<float4>> ckP, ckV;      // checkpoints: state, flags and ids (slot order)
        std::vector<DevBuf<int32_t>> ckF;
        std::vector<DevBuf<uint32_t>> ckI;
        std::vector<DevBuf<float4>> P0, V0, Xs, Vs, cbpos, cbvel, cnorm;
        std::vector<DevBuf<uint32_t>> keys, vals;
        std::vector<DevBuf<int32_t>> Fs, cpart, cshape;
        std::vector<DevBuf<float>> delta;
        DevBuf<int32_t> counts;
        DevBuf<float4> lx[2], lv[2], As, finP, finV;
        DevBuf<int32_t> finF;
        DevBuf<uint32_t> finI;
        DevBuf<float> shx, shv;
        DevBuf<double> relax_part;
    } tape_;
    bool use_graphs_ = true;
    // deterministic mode: canonical contact keys (by particle slot and shape;
    // by box and particle id), their sorted copies, the sort's scratch,
    // per-contact deltas and each box's range of the box-key order
    bool det_ = false;
    int det_sbits_ = 0, det_idbits_ = 0, det_bbits_ = 0, det_kb1_ = 0, det_kb2_ = 0;
    DevBuf<uint32_t> dk1_, dv1_, dk1s_, dv1s_, dk2_, dv2_, dk2s_, dv2s_;
    DevBuf<char> dscr_;
    DevBuf<float4> cfd_, cbd_, rcd_;
    DevBuf<int32_t> dseg_;
    // particle state: 3 rotating float4 buffers each (state / two iteration buffers)
    DevBuf<float4> P_[3], V_[3];
    int s_ = 0, p_ = 1, o_ = 2;
    // The particle state lives in slot order: after every substep's sort, the
    // slots follow that sort (neighbours adjacent, writes coalesced).  ids_
    // maps slot -> particle id; flags_ and ids_ are double-buffered (lay_ is
    // current) because the last radix pass writes the next order's copies.
    DevBuf<int32_t> flags_[2];
    DevBuf<uint32_t> ids_[2];
    int lay_ = 0;
    DevBuf<uint32_t> rank_;    // previous slot -> sorted slot, from the last sort
    DevBuf<uint2> kv_;         // radix input: (cell key, slot) at each particle id
    int32_t* flags_cur() { return flags_[lay_].p; }
    uint32_t* ids_cur() { return ids_[lay_].p; }
    std::vector<uint32_t> host_ids();
    // host I/O staging (persistent): a device buffer the getters fill in
    // particle order and a pinned host buffer it is copied to in one DMA
    DevBuf<float> io_dev_;
    float* io_host_ = nullptr;
    size_t io_cap_ = 0;
    void* stage_out(size_t words);   // ensures both hold `words` 4-byte words; returns io_dev_
    void copy_out(void* out, size_t bytes);   // io_dev_ -> pinned -> out
    std::vector<float> radius_h_, invm_h_;
    float rmax_ = 0.f;   // largest particle radius (host copy of the radii)
    // grid and sort
    DevBuf<uint32_t> keysA_, valsA_, keysB_, valsB_, hist_, bintot_;
    DevBuf<uint32_t> chunks_;          // chunk totals of the chunked radix passes, kSortMaxPasses arrays
    // batched solvers: radix tiles that never cross a scene (first key, count,
    // the scene's first tile, the next scene's first tile), for 1,024- and
    // 2,048-key tiles; the passes then sort each scene's key range by its cells
    DevBuf<int4> seg_tab_[2];
    int seg_tiles_[2] = {0, 0};
    bool seg_sort_ = true;             // VKXPBD_SEG_SORT=0: the scene bits as a radix digit instead
    void build_seg_tables();
    size_t sort_chunk_words_ = 0;      // words per pass array
    bool sort_chunked_ = false;        // chunked radix passes (VKXPBD_SORT_CHUNKED=1; else upsweep, binscan, scatter)
    bool batch_gather_rows_ = true;    // batched: state moved by the row-table build (VKXPBD_BATCH_GATHER_IN_ROWS=0: last radix pass)
    bool solve_in_walk_ = true;        // the walk solves its contacts (VKXPBD_SOLVE_IN_WALK=0: k_shape_solve)
    bool fuse_keys_ = true;            // substeps keyed by the previous one's apply (VKXPBD_FUSE_KEYS=0: off)
    DevBuf<float4> Xs_[1], Vs_[1];   // predicted state in sorted order (the first iteration's input)
    DevBuf<int32_t> range_;   // int2 per key (start, end); rows of the last build are reset before the next
    DevBuf<uint32_t> rkeys_;  // with wrap planes: the sorted keys whose rows range_ holds
    // batched solvers: the cell ranges as a row table (8 bytes per row of
    // every scene) over 8 lower bounds per occupied group of 8 cells, the
    // row each sorted position of the last build filled (-1: none), and the
    // per-tile scratch of the group numbering
    DevBuf<uint2> rows_;
    DevBuf<uint32_t> glb_;
    DevBuf<int32_t> rslot_;
    DevBuf<uint32_t> gtcnt_, gtoff_;
    RowTable rowtab() const { return RowTable{rows_.p, glb_.p}; }
    DevBuf<int32_t> tbl_;     // 2 x cells scratch for the dense getters
    // contacts: SoA of float4 payloads, counter slots
    DevBuf<int32_t> cpart_, cshape_, cc_;
    DevBuf<float4> cbpos_, cbvel_, cnorm_;
    int contact_slot_ = 0, last_count_ = 0;
    int32_t* count_host_ = nullptr;   // pinned
    DevBuf<float> dshape_, delta_;
    DevBuf<float> ppd_;   // 0 + pair sum of particles whose plane contact the pair solve added
    DevBuf<int> owner_;   // per sorted slot: contact that finishes its shape delta, else -1
    // bodies (one static body per mesh, scene.cpp:220-238) and shapes
    DevBuf<float4> body_q_, body_qd_, body_half_;
    DevBuf<float> body_com_, body_I_, body_invI_, body_invm_, body_delta_, body_invw_, mat_mu_, mat_mu_t_, mat_mu_r_;
    DevBuf<int32_t> body_flags_, shape_body_;
    std::vector<BoxHost> boxes_;
    float angular_damping_ = 0.05f, rigid_relax_ = 0.8f, mu_t_ = 0.f, mu_r_ = 0.f;
    int use_weights_ = 1;
    // rigid contacts (box corners vs the ground plane)
    DevBuf<int32_t> rc_s0_, rc_s1_;
    DevBuf<float4> rc_p0_, rc_p1_, rc_n_;
    DevBuf<float> rc_impulse_;
    int rigid_slot_ = 0;
    // meshes
    std::vector<MeshHost> meshes_;
    std::vector<MeshBvhHost> bvh_;
    std::vector<float> all_verts_, all_vels_;
    std::vector<uint32_t> all_tris_;
    std::vector<size_t> vert_off_, tri_off_, node_off_, prim_off_;
    DevBuf<float> verts_d_, vvel_d_;
    DevBuf<uint32_t> tris_d_, prim_d_, err_d_;
    DevBuf<float4> nodes_d_;   // interleaved (lower, upper) node boxes, re-based
    DevBuf<float4> cnodes_d_;  // per internal node: walk record (children and grandchildren boxes, stack words)
    DevBuf<float4> lbox_d_;    // leaf boxes by first primitive position
    std::vector<ParityGrid> pgrid_;        // per mesh (empty for non-watertight)
    DevBuf<uint32_t> pg_start_;            // all meshes' parity-grid cell starts, concatenated
    std::vector<DistGrid> dgrid_;          // per mesh distance lower bounds
    DevBuf<float> lb_;                     // all meshes' bounds, concatenated
    DevBuf<uint32_t> lb_node_;             // per bound cell: search start node
    std::vector<uint32_t> prim_h_;         // global triangle id per leaf position
    DevBuf<float4> tsoup_;                 // triangle corners in leaf order
    DevBuf<float4> pg_soup_;               // triangle corners of every parity-grid entry
    DevBuf<uint32_t> mcnt_;                // mesh queue counters: 2 substep parities x meshes x 2
    DevBuf<int2> mq_;                      // mesh candidate queue (particle, start node)
    std::vector<uint8_t> mesh_split_;      // per mesh: walk with four lanes per particle
    DevBuf<MeshInfo> info_d_;
    std::vector<MeshInfo> info_h_;
    float avg_edge_ = 0.f;
    // animated colliders: room per mesh in the bound grid (cells) and the
    // parity entries (first entry, capacity); refit order (global nodes by
    // height, per mesh the offsets of its heights); per global node its
    // leaf-order range and stack word; per global triangle its leaf position;
    // per mesh the vertices its triangles use; scratch for the rebuilds
    std::vector<size_t> lb_room_, pg_base_, pg_room_;
    DevBuf<uint32_t> refit_list_d_;
    std::vector<std::vector<uint32_t>> refit_lvl_;
    DevBuf<uint32_t> nfirst_d_, nlast_d_, nword_d_, tpos_d_;
    std::vector<std::vector<uint32_t>> mesh_vref_;
    DevBuf<uint32_t> span_lo_d_, span_hi_d_, pcnt_d_, ptotal_d_;
    // diagnostics
    DevBuf<uint32_t> diag_header_;
    DevBuf<uint64_t> mesh_stats_;
    bool mesh_stats_on_ = false;   // VKXPBD_MESH_STATS=1: count mesh query work (adds atomics)
    DevBuf<float> diag_pairs_;
    uint32_t max_pairs_ = 100000;
    // profiler
    bool profiling_ = false;
    std::vector<hipEvent_t> ev_pool_;
    struct Pending { int stage; hipEvent_t a, b; };
    std::vector<Pending> pending_;
    size_t ev_used_ = 0;
    double stage_ms_[kNumStages] = {};
    uint64_t stage_n_[kNumStages] = {};
    uint64_t prof_substeps_ = 0;
    double prof_wall_ms_ = 0.0;
    hipEvent_t cur_a_ = nullptr;
};

}  // namespace xpbd
