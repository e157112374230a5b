// solver.cpp -- host side of the MI355X XPBD step (see solver.h).
//
// Substep pipeline (K = iterations):
//   [k_mesh_*]                   particle/mesh narrowphase on x0      (if meshes;
//                                side stream, beside integrate and the sort)
//   k_integrate                  plane contacts on x0, cell keys of x_pred;
//                                resets the previous substep's cell rows
//   sort: (upsweep, binscan, scatter) per 7-bit digit, the last pass storing
//         the integrated state (x_pred, v_pred) in cell order
//   k_cell_ranges                (start, end) of every occupied cell
//   K x { k_shape_solve          (first one on the side stream, after the mesh)
//         k_pp_apply             particle-particle solve + delta apply
//         k_shape_apply }        shape deltas of the particles that have any
//   [k_restitution]
// Compared with src/scene.cpp:805-992 this drops the contact-counter fill,
// the 16 MiB cell-table fills, the delta fill and the odd-K copy-back, and
// replaces 15 sort dispatches by 9 (12 with wrap planes).
#include "solver.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "devbuf.h"
#include "sort_entry.h"

namespace xpbd {

namespace {

constexpr const char* kStageNames[] = {"mesh_contacts", "integrate",      "sort_upsweep",   "sort_binscan",
                                       "sort_scatter",  "cell_ranges",    "shape_solve",    "pp_solve_apply",
                                       "restitution",   "readback",       "body_contacts",  "integrate_bodies",
                                       "body_contact_solve", "apply_body_deltas", "shape_apply"};

}  // namespace

Solver::Solver() = default;

Solver::~Solver() {
    if (!initialized_) return;
    DeviceGuard g(device_);
    (void)hipStreamSynchronize(stream_);
    if (profiling_) print_profiler_summary();
    free_all();
}

void Solver::require_init() const {
    if (!initialized_) throw std::runtime_error("XpbdSolver not initialized");
}

void Solver::free_all() {
    for (int k = 0; k < 3; k++) { dfree(P_[k]); dfree(V_[k]); }
    for (int k = 0; k < 2; k++) { dfree(flags_[k]); dfree(ids_[k]); }
    dfree(rank_); dfree(kv_); dfree(keysA_); dfree(valsA_); dfree(keysB_); dfree(valsB_); dfree(hist_); dfree(bintot_); dfree(chunks_); dfree(seg_tab_[0]); dfree(seg_tab_[1]);
    dfree(Xs_[0]); dfree(Vs_[0]);
    dfree(body_half_); dfree(body_I_); dfree(body_invw_); dfree(mat_mu_t_); dfree(mat_mu_r_);
    dfree(rc_s0_); dfree(rc_s1_); dfree(rc_p0_); dfree(rc_p1_); dfree(rc_n_); dfree(rc_impulse_);
    dfree(range_); dfree(rkeys_); dfree(rows_); dfree(glb_); dfree(rslot_); dfree(gtcnt_); dfree(gtoff_); dfree(tbl_); dfree(cpart_); dfree(cshape_); dfree(cc_); dfree(cbpos_); dfree(cbvel_); dfree(cnorm_);
    dfree(dshape_); dfree(delta_); dfree(ppd_); dfree(owner_); dfree(body_q_); dfree(body_qd_); dfree(body_com_); dfree(body_invI_);
    dfree(body_invm_); dfree(body_delta_); dfree(mat_mu_); dfree(body_flags_); dfree(shape_body_);
    dfree(verts_d_); dfree(vvel_d_); dfree(tris_d_); dfree(prim_d_); dfree(err_d_); dfree(nodes_d_); dfree(cnodes_d_); dfree(lbox_d_); dfree(pg_start_); dfree(lb_); dfree(lb_node_); dfree(tsoup_); dfree(pg_soup_); dfree(mcnt_); dfree(mq_);
    dfree(info_d_); dfree(diag_header_); dfree(diag_pairs_); dfree(mesh_stats_);
    for (hipEvent_t e : ev_pool_) (void)hipEventDestroy(e);
    ev_pool_.clear();
    invalidate_graphs();
    free_tape();
    dfree(scene_d_);
    dfree(dk1_); dfree(dv1_); dfree(dk1s_); dfree(dv1s_); dfree(dk2_); dfree(dv2_); dfree(dk2s_); dfree(dv2s_);
    dfree(dscr_); dfree(cfd_); dfree(cbd_); dfree(rcd_); dfree(dseg_);
    if (count_host_) { (void)hipHostFree(count_host_); count_host_ = nullptr; }
    dfree(io_dev_);
    if (io_host_) { (void)hipHostFree(io_host_); io_host_ = nullptr; }
    io_cap_ = 0;
    if (fork_ev_) { (void)hipEventDestroy(fork_ev_); fork_ev_ = nullptr; }
    if (join_ev_) { (void)hipEventDestroy(join_ev_); join_ev_ = nullptr; }
    if (sorted_ev_) { (void)hipEventDestroy(sorted_ev_); sorted_ev_ = nullptr; }
    if (side_) { (void)hipStreamDestroy(side_); side_ = nullptr; }
    if (stream_) { (void)hipStreamDestroy(stream_); stream_ = nullptr; }
}

void Solver::alloc_all() {
    for (int k = 0; k < 3; k++) { dalloc(P_[k], n_); dalloc(V_[k], n_); }
    s_ = 0; p_ = 1; o_ = 2;
    for (int k = 0; k < 2; k++) { dalloc(flags_[k], n_); dalloc(ids_[k], n_); }
    lay_ = 0;
    dalloc(rank_, n_);
    dalloc(kv_, n_);
    radius_h_.assign(n_, 0.f);
    invm_h_.assign(n_, 0.f);
    rmax_ = 0.f;
    dalloc(keysA_, n_); dalloc(valsA_, n_); dalloc(keysB_, n_); dalloc(valsB_, n_);
    dalloc(Xs_[0], n_); dalloc(Vs_[0], n_);
    // radix tiles: 2,048 keys from 262,144 particles on, else 1,024 (autotune
    // may switch); the histograms are sized for the smaller tiles
    sort_items_ = sort_items_for((long long)n_);
    const size_t tile = size_t(256) * 4;
    const size_t tiles = (n_ + tile - 1) / tile;
    // (rows padded to whole chunks for the chunked passes)
    const size_t ctiles = (std::max<size_t>(tiles, 1) + kSortChunkTiles - 1) / kSortChunkTiles;
    dalloc(hist_, size_t(kSortMaxBins) * ctiles * kSortChunkTiles);
    dalloc(bintot_, kSortMaxBins);
    // chunk totals of the chunked passes: one array per pass index, each
    // 256 bins x the chunks of the smaller tiles; zero from here on between sorts
    sort_chunk_words_ = size_t(256) * ctiles;
    dalloc(chunks_, kSortMaxPasses * sort_chunk_words_);
    // the plain solver's dense 2^21-row cell table (the key is the row; the
    // keys of the rows written last, rkeys, let each build reset exactly those
    // rows first); a batched solver's row table instead: 8 bytes per row
    // (2^14 per scene) and 32 bytes per particle
    dalloc(range_, num_scenes_ > 1 ? size_t(4) : size_t(2) * kNumCells);
    dalloc(rkeys_, n_);
    if (num_scenes_ > 1) {
        dalloc(rows_, num_scenes_ << 14);           // every row empty
        dalloc(glb_, 8 * (n_ + 1));                 // at most one group per particle, and the sentinel
        dalloc(rslot_, n_);
        dfill(rslot_, 0xff);                        // no row filled
        dalloc(gtcnt_, size_t(group_tiles(int(n_))));
        dalloc(gtoff_, size_t(group_tiles(int(n_))));
    }
    dalloc(cpart_, cmax_); dalloc(cshape_, cmax_);
    dalloc(cbpos_, cmax_); dalloc(cbvel_, cmax_); dalloc(cnorm_, cmax_);
    dalloc(cc_, 8);   // [particle slot 0, slot 1, error flags, spare, rigid slot 0, slot 1, spare, spare]
    dalloc(dshape_, 3 * n_); dalloc(delta_, 3 * n_); dalloc(ppd_, 3 * n_);
    dalloc(owner_, n_);
    dfill(owner_, 0xff);   // -1: no shape delta pending
    dalloc(diag_header_, 2);
    dalloc(mesh_stats_, 8);
    dalloc(mq_, n_);   // mesh candidate queue (its counters are sized by finalize_meshes)
    max_pairs_ = cfg_.max_diag_pairs ? cfg_.max_diag_pairs : 100000u;
    dalloc(diag_pairs_, cfg_.enable_diagnostics ? size_t(20) * max_pairs_ : 20);
    HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&count_host_), 4 * sizeof(int32_t), hipHostMallocDefault));
    launch_iota(ids_[0].p, int(n_), stream_);   // slot order starts as particle order
    HIPCHECK(hipStreamSynchronize(stream_));
}

void Solver::init_common(const xpbd_config& cfg, size_t n, size_t max_contacts) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) throw std::runtime_error("No HIP devices found");
    if (n > (size_t(1) << 28)) throw std::runtime_error("num_particles exceeds the supported maximum of 2^28");
    if (initialized_) { DeviceGuard g(device_); (void)hipDeviceSynchronize(); free_all(); initialized_ = false; }
    cfg_ = cfg;
    relax_ = cfg.soft_contact_relaxation;
    n_ = n;
    cmax_ = max_contacts ? max_contacts : n * 2;
    if (cmax_ > size_t(0x7FFFFFFF)) throw std::runtime_error("max_contacts exceeds the int32 range");
    device_ = (cfg.device_index >= 0 && cfg.device_index < count) ? cfg.device_index : 0;
    DeviceGuard g(device_);
    HIPCHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    HIPCHECK(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&join_ev_, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&sorted_ev_, hipEventDisableTiming));
    alloc_all();
    const char* env = std::getenv("VKXPBD_PROFILE");
    bool env_on = env && (std::strcmp(env, "1") == 0 || std::strcmp(env, "true") == 0);
    profiling_ = cfg.enable_profile || env_on;
    if (profiling_) std::printf("[vkxpbd] GPU profiler enabled (HIP events)\n");
    const char* senv = std::getenv("VKXPBD_MESH_STATS");
    mesh_stats_on_ = senv && (std::strcmp(senv, "1") == 0 || std::strcmp(senv, "true") == 0);
    const char* denv = std::getenv("VKXPBD_DETERMINISTIC");
    det_ = denv && (std::strcmp(denv, "1") == 0 || std::strcmp(denv, "true") == 0);
    const char* genv = std::getenv("VKXPBD_GRAPHS");
    use_graphs_ = !(genv && std::strcmp(genv, "0") == 0);
    const char* cenv = std::getenv("VKXPBD_SORT_CHUNKED");
    sort_chunked_ = cenv && std::strcmp(cenv, "1") == 0;
    const char* genv2 = std::getenv("VKXPBD_BATCH_GATHER_IN_ROWS");
    batch_gather_rows_ = !(genv2 && std::strcmp(genv2, "0") == 0);
    const char* wenv = std::getenv("VKXPBD_SOLVE_IN_WALK");
    solve_in_walk_ = !(wenv && std::strcmp(wenv, "0") == 0);
    const char* fenv = std::getenv("VKXPBD_FUSE_KEYS");
    fuse_keys_ = !(fenv && std::strcmp(fenv, "0") == 0);
    // wrap planes pay for their 128 MiB or 1 GiB table and wider radix keys
    // on large scenes only; the environment can force them either way
    const char* penv = std::getenv("VKXPBD_WRAP_PLANES");
    planes_allowed_ = penv ? std::strcmp(penv, "1") == 0 : n_ >= kPlanesMinParticles;
    const char* benv = std::getenv("VKXPBD_PLANE_BITS");
    plane_bits_fixed_ = benv != nullptr;
    plane_bits_ = (benv && std::strcmp(benv, "1") == 0) ? 1 : 2;
    table_bits_ = -1;
    planes_ = false;
    update_planes();
    grid_valid_ = false;
    contact_slot_ = 0;
    rigid_slot_ = 0;
    last_count_ = 0;
    initialized_ = true;
    finalized_ = false;
}

void Solver::init(const xpbd_config& cfg, size_t n, size_t max_contacts) {
    num_scenes_ = 1;
    scene_off_.assign({0, n});
    set_sort_shape(kCellBits);
    init_common(cfg, n, max_contacts);
}

void Solver::init_batch(const xpbd_config& cfg, const size_t* counts, size_t num_scenes, size_t max_contacts) {
    if (num_scenes == 0) throw std::runtime_error("init_batch: num_scenes must be positive");
    // 11 key bits above the cell
    if (num_scenes > 2048) throw std::runtime_error("init_batch: at most 2048 scenes (11 key bits above the cell)");
    size_t total = 0;
    scene_off_.assign(1, 0);
    for (size_t k = 0; k < num_scenes; k++) { total += counts[k]; scene_off_.push_back(total); }
    num_scenes_ = num_scenes;
    int bits = kCellBits;
    while ((size_t(1) << (bits - kCellBits)) < num_scenes) bits++;
    set_sort_shape(bits);
    init_common(cfg, total, max_contacts);
    std::vector<int32_t> sc(total);
    for (size_t k = 0; k < num_scenes; k++)
        for (size_t i = scene_off_[k]; i < scene_off_[k + 1]; i++) sc[i] = int32_t(k);
    DeviceGuard g(device_);
    upload(scene_d_, sc.data(), sc.size());
    // a scene's particles keep their key range (ids and slots are
    // scene-major), so the passes may sort each scene by its cells alone
    const char* senv = std::getenv("VKXPBD_SEG_SORT");
    seg_sort_ = !(senv && std::strcmp(senv, "0") == 0);
    if (seg_sort_) {
        build_seg_tables();
        set_sort_shape(kCellBits);
    }
}

// Radix tiles of whole scenes for both tile sizes: each scene's keys
// [off, off + count) split into tiles of at most 1,024 or 2,048 keys.
void Solver::build_seg_tables() {
    for (int k = 0; k < 2; k++) {
        const int tile = 256 * (k ? 8 : 4);
        std::vector<int4> tab;
        for (size_t s = 0; s + 1 < scene_off_.size(); s++) {
            const int a = int(scene_off_[s]), b = int(scene_off_[s + 1]);
            const int t0 = int(tab.size());
            for (int x = a; x < b; x += tile) tab.push_back(make_int4(x, std::min(tile, b - x), t0, 0));
            for (size_t t = size_t(t0); t < tab.size(); t++) tab[t].w = int(tab.size());
        }
        seg_tiles_[k] = int(tab.size());
        if (tab.empty()) tab.push_back(make_int4(0, 0, 0, 0));
        upload(seg_tab_[k], tab.data(), tab.size());
    }
    // the tile histograms: room for the larger count of tiles
    const size_t need = size_t(kSortMaxBins) * size_t(std::max(seg_tiles_[0], 1));
    if (hist_.n < need) dalloc(hist_, need);
}

void Solver::add_mesh(const float* v, size_t nv, const uint32_t* t, size_t nt, float margin, const float* rot, int wt) {
    MeshHost m;
    m.verts.assign(v, v + nv * 3);
    m.tris.assign(t, t + nt * 3);
    m.margin = margin;
    if (rot) std::memcpy(m.rot, rot, sizeof(m.rot));
    if (wt >= 0) m.watertight = wt != 0;
    else if (nt == 0 || nv == 0) m.watertight = false;
    else {   // every undirected edge shared by exactly two triangles (scene.cpp:170-198)
        std::vector<uint64_t> e;
        e.reserve(nt * 3);
        for (size_t k = 0; k < nt; k++)
            for (int s = 0; s < 3; s++) {
                uint64_t a = t[k * 3 + s], b = t[k * 3 + (s + 1) % 3];
                e.push_back(a < b ? (a << 32) | b : (b << 32) | a);
            }
        std::sort(e.begin(), e.end());
        bool closed = true;
        for (size_t i = 0; i < e.size() && closed;) {
            size_t j = i + 1;
            while (j < e.size() && e[j] == e[i]) j++;
            closed = (j - i) == 2;
            i = j;
        }
        m.watertight = closed;
    }
    meshes_.push_back(std::move(m));
    finalized_ = false;
}

void Solver::upload_nodes() {
    size_t total = 0;
    for (auto& b : bvh_) total += b.lo.size();
    std::vector<float4> nl(total), nu(total);
    for (size_t c = 0; c < bvh_.size(); c++) {
        const uint32_t nb = uint32_t(node_off_[c]), pb = uint32_t(prim_off_[c]);
        for (size_t i = 0; i < bvh_[c].lo.size(); i++) {
            const NodeHalf L = bvh_[c].lo[i], U = bvh_[c].hi[i];
            uint32_t lw, uw;
            if ((L.packed >> 30) & 3u) {   // leaf: primitive range, re-based (scene.cpp:299-313)
                lw = ((L.packed & 0x3FFFFFFFu) + pb) | (1u << 30);
                uw = (U.packed & 0x3FFFFFFFu) + pb;
            } else {
                lw = (L.packed & 0x3FFFFFFFu) + nb;
                uw = (U.packed & 0x3FFFFFFFu) + nb;
            }
            float lf, uf;
            std::memcpy(&lf, &lw, 4);
            std::memcpy(&uf, &uw, 4);
            nl[nb + i] = make_float4(L.x, L.y, L.z, lf);
            nu[nb + i] = make_float4(U.x, U.y, U.z, uf);
        }
    }
    std::vector<float4> inter(2 * total);
    for (size_t i = 0; i < total; i++) { inter[2 * i] = nl[i]; inter[2 * i + 1] = nu[i]; }
    upload(nodes_d_, inter.data(), inter.size());
    // Walk records: per internal node its children's boxes and, for internal
    // children, their children's boxes (12 float4); each lower corner's w is
    // the child's stack word.  Leaves whose range is short carry it in the
    // word; their boxes are also stored by first primitive position.
    std::vector<float4> rec(12 * total, make_float4(0.f, 0.f, 0.f, 0.f));
    size_t nprim = 0;
    for (auto& b : bvh_) nprim += b.prim.size();
    std::vector<float4> lbox(2 * std::max<size_t>(nprim, 1), make_float4(0.f, 0.f, 0.f, 0.f));
    auto put = [](float4* dst, const NodeHalf& L, const NodeHalf& U, uint32_t w) {
        float fw;
        std::memcpy(&fw, &w, 4);
        dst[0] = make_float4(L.x, L.y, L.z, fw);
        dst[1] = make_float4(U.x, U.y, U.z, 0.f);
    };
    for (size_t c = 0; c < bvh_.size(); c++) {
        const uint32_t nb = uint32_t(node_off_[c]);
        const MeshBvhHost& b = bvh_[c];
        if (b.lo.empty()) continue;
        // only nodes reachable from the root: a collapsed leaf's former
        // descendants are leaves too and share its first primitive
        std::vector<uint32_t> reach{b.root};
        for (size_t k = 0; k < reach.size(); k++) {
            const uint32_t v = reach[k];
            if (!((b.lo[v].packed >> 30) & 3u)) {
                reach.push_back(b.lo[v].packed & 0x3FFFFFFFu);
                reach.push_back(b.hi[v].packed & 0x3FFFFFFFu);
            }
        }
        for (const uint32_t i : reach) {
            if ((b.lo[i].packed >> 30) & 3u) {   // leaf: box by first primitive
                const uint32_t li = (b.lo[i].packed & 0x3FFFFFFFu) + uint32_t(prim_off_[c]);
                if (li < nprim) put(&lbox[2 * size_t(li)], b.lo[i], b.hi[i], 0u);
                continue;
            }
            const uint32_t ch[2] = {b.lo[i].packed & 0x3FFFFFFFu, b.hi[i].packed & 0x3FFFFFFFu};
            for (int k = 0; k < 2; k++) {
                put(&rec[12 * (nb + i) + 2 * k], b.lo[ch[k]], b.hi[ch[k]], node_word(c, ch[k]));
                if ((b.lo[ch[k]].packed >> 30) & 3u) continue;
                const uint32_t gc[2] = {b.lo[ch[k]].packed & 0x3FFFFFFFu, b.hi[ch[k]].packed & 0x3FFFFFFFu};
                for (int m = 0; m < 2; m++)
                    put(&rec[12 * (nb + i) + 4 + 4 * k + 2 * m], b.lo[gc[m]], b.hi[gc[m]], node_word(c, gc[m]));
            }
        }
    }
    upload(cnodes_d_, rec.data(), rec.size());
    upload(lbox_d_, lbox.data(), lbox.size());
    for (size_t c = 0; c < bvh_.size() && c < info_h_.size(); c++)
        info_h_[c].root_word = bvh_[c].lo.empty() ? 0u : node_word(c, bvh_[c].root);
}

// Stack word of node v of mesh c for the tree walk: an internal node's global
// index; a leaf with 1..15 primitives as 1 << 31 | count << 27 | first
// primitive position; any other leaf as 1 << 31 | global node index.
uint32_t Solver::node_word(size_t c, uint32_t v) const {
    const MeshBvhHost& b = bvh_[c];
    const uint32_t g = uint32_t(node_off_[c]) + v;
    if (g >= (1u << 27)) throw std::runtime_error("mesh BVH exceeds 2^27 nodes");
    if (!((b.lo[v].packed >> 30) & 3u)) return g;
    const uint32_t li = (b.lo[v].packed & 0x3FFFFFFFu) + uint32_t(prim_off_[c]);
    const uint32_t ri = (b.hi[v].packed & 0x3FFFFFFFu) + uint32_t(prim_off_[c]);
    const uint32_t cnt = ri - li;
    if (cnt >= 1 && cnt <= 15 && li < (1u << 27)) return 0x80000000u | (cnt << 27) | li;
    return 0x80000000u | g;
}

// Host-side query accelerators of one mesh: the projected parity grid
// (watertight meshes) and the distance lower-bound grid.  The bound grid's cap
// covers the search distance of particles up to twice particle_max_radius;
// larger particles take the full tree walk.
void Solver::build_query_grids(size_t c) {
    const uint32_t nt = uint32_t(meshes_[c].tris.size() / 3);
    const uint32_t* tris = all_tris_.data() + 3 * tri_off_[c];
    pgrid_[c] = ParityGrid{};
    if (meshes_[c].watertight) parity_grid_build(all_verts_.data(), tris, nt, uint32_t(tri_off_[c]), pgrid_[c]);
    const float cap = 1.25f * ((cfg_.soft_contact_margin + meshes_[c].margin) + 2.0f * cfg_.particle_max_radius) +
                      avg_edge_ * 1e-3f;
    dist_grid_build(all_verts_.data(), tris, nt, cap, dgrid_[c], &bvh_[c]);
}

// Concatenates the per-mesh grids and refreshes the mesh descriptors.  Each
// mesh gets a block of room: parity cell starts (dim^2 + 1, global entry
// positions) and an entry segment of twice its entries plus 1,024; bounds and
// start nodes for twice its cells plus 64.  update_mesh rebuilds a mesh's
// grids inside its blocks on the device.
void Solver::upload_query_grids() {
    const size_t M = pgrid_.size();
    std::vector<uint32_t> start, lbn;
    std::vector<float> lb;
    std::vector<float4> psoup;
    lb_room_.assign(M, 0);
    pg_base_.assign(M, 0);
    pg_room_.assign(M, 0);
    for (size_t c = 0; c < M; c++) {
        const ParityGrid& g = pgrid_[c];
        info_h_[c].pg_u0 = g.u0;
        info_h_[c].pg_w0 = g.w0;
        info_h_[c].pg_inv = g.inv;
        info_h_[c].pg_dim = g.dim;
        info_h_[c].pg_cell_off = uint32_t(start.size());
        const uint32_t base = uint32_t(psoup.size() / 3);
        pg_base_[c] = base;
        for (const uint32_t v : g.start) start.push_back(base + v);
        for (uint32_t t : g.tri)   // the registered triangles' corners, in cell order
            for (int q = 0; q < 3; q++) {
                const uint32_t v = all_tris_[3 * size_t(t) + q];
                psoup.push_back(make_float4(all_verts_[3 * v], all_verts_[3 * v + 1], all_verts_[3 * v + 2], 0.f));
            }
        if (g.dim > 0) {
            pg_room_[c] = 2 * g.tri.size() + 1024;
            psoup.resize(3 * (base + pg_room_[c]), make_float4(0.f, 0.f, 0.f, 0.f));
        }
        const DistGrid& d = dgrid_[c];
        info_h_[c].lb_ox = d.ox; info_h_[c].lb_oy = d.oy; info_h_[c].lb_oz = d.oz; info_h_[c].lb_inv = d.inv;
        info_h_[c].lb_nx = d.nx; info_h_[c].lb_ny = d.ny; info_h_[c].lb_nz = d.nz;
        info_h_[c].lb_off = uint32_t(lb.size());
        info_h_[c].lb_cap = d.lb.empty() ? 0.0f : d.cap;   // cap 0: no early out
        lb.insert(lb.end(), d.lb.begin(), d.lb.end());
        for (size_t k = 0; k < d.lb.size(); k++)   // start nodes as stack words
            lbn.push_back(node_word(c, k < d.start.size() ? d.start[k] : bvh_[c].root));
        if (!d.lb.empty()) {
            lb_room_[c] = 2 * d.lb.size() + 64;
            lb.resize(info_h_[c].lb_off + lb_room_[c], 0.f);
            lbn.resize(info_h_[c].lb_off + lb_room_[c], 0u);
        }
    }
    start.push_back(uint32_t(psoup.size() / 3));
    upload(pg_start_, start.data(), start.size());
    upload(pg_soup_, psoup.data(), psoup.size());
    upload(lb_, lb.data(), lb.size());
    upload(lb_node_, lbn.data(), lbn.size());
    upload(info_d_, info_h_.data(), info_h_.size());
}

// Corners of every triangle in leaf order (the order the tree search reads
// them), the first corner's w carrying the global triangle id: one contiguous
// 48-byte read per triangle instead of an index-then-vertex gather.
void Solver::upload_tri_soup() {
    std::vector<float4> soup(3 * prim_h_.size());
    for (size_t k = 0; k < prim_h_.size(); k++) {
        const uint32_t t = prim_h_[k];
        for (int q = 0; q < 3; q++) {
            const uint32_t v = all_tris_[3 * size_t(t) + q];
            float w = 0.f;
            if (q == 0) std::memcpy(&w, &t, sizeof(w));
            soup[3 * k + q] = make_float4(all_verts_[3 * v], all_verts_[3 * v + 1], all_verts_[3 * v + 2], w);
        }
    }
    upload(tsoup_, soup.data(), soup.size());
}

void Solver::finalize_meshes() {
    if (!initialized_) throw std::runtime_error("XpbdSolver::init must be called before finalize_meshes");
    DeviceGuard g(device_);
    HIPCHECK(hipStreamSynchronize(stream_));
    invalidate_graphs();
    const size_t M = meshes_.size(), NB = boxes_.size();
    // shapes: 0 = ground plane (body -1), 1..M = meshes (static bodies 0..M-1),
    // M+1..M+NB = boxes (dynamic bodies M..M+NB-1)
    std::vector<int32_t> shape_body(M + 1 + NB, -1);
    for (size_t c = 0; c < M + NB; c++) shape_body[c + 1] = int32_t(c);
    std::vector<float> mat_mu(M + 1 + NB, cfg_.shape_material_mu), mat_mu_t(M + 1 + NB, mu_t_),
        mat_mu_r(M + 1 + NB, mu_r_);
    upload(shape_body_, shape_body.data(), shape_body.size());
    upload(mat_mu_, mat_mu.data(), mat_mu.size());
    upload(mat_mu_t_, mat_mu_t.data(), mat_mu_t.size());
    upload(mat_mu_r_, mat_mu_r.data(), mat_mu_r.size());
    // mesh bodies: identity pose, zero velocity / com / inverse mass and inertia;
    // boxes: solid boxes, m = rho * 8abc, I = m/3 (b^2 + c^2, a^2 + c^2, a^2 + b^2)
    const size_t B = std::max<size_t>(M + NB, 1);
    std::vector<float4> bq(2 * B, make_float4(0.f, 0.f, 0.f, 0.f)), bh(B, make_float4(0.f, 0.f, 0.f, 0.f));
    for (size_t b = 0; b < B; b++) bq[2 * b + 1] = make_float4(0.f, 0.f, 0.f, 1.f);
    std::vector<float> bI(9 * B, 0.f), binvI(9 * B, 0.f), binvm(B, 0.f);
    std::vector<int32_t> bflags(B, 0);
    for (size_t k = 0; k < NB; k++) {
        const BoxHost& x = boxes_[k];
        const size_t b = M + k;
        bq[2 * b] = make_float4(x.pos[0], x.pos[1], x.pos[2], 0.f);
        bq[2 * b + 1] = make_float4(x.rot[0], x.rot[1], x.rot[2], x.rot[3]);
        bh[b] = make_float4(x.half[0], x.half[1], x.half[2], 0.f);
        bflags[b] = x.flags;
        const float hx = x.half[0], hy = x.half[1], hz = x.half[2];
        const float m = x.density * 8.0f * hx * hy * hz;
        const float ix = m / 3.0f * (hy * hy + hz * hz), iy = m / 3.0f * (hx * hx + hz * hz),
                    iz = m / 3.0f * (hx * hx + hy * hy);
        binvm[b] = m > 0.0f ? 1.0f / m : 0.0f;
        bI[9 * b] = ix; bI[9 * b + 4] = iy; bI[9 * b + 8] = iz;
        binvI[9 * b] = ix > 0.f ? 1.0f / ix : 0.f;
        binvI[9 * b + 4] = iy > 0.f ? 1.0f / iy : 0.f;
        binvI[9 * b + 8] = iz > 0.f ? 1.0f / iz : 0.f;
    }
    upload(body_q_, bq.data(), bq.size());
    dalloc(body_qd_, 2 * B);
    upload(body_half_, bh.data(), bh.size());
    dalloc(body_com_, 3 * B);
    upload(body_I_, bI.data(), bI.size());
    upload(body_invI_, binvI.data(), binvI.size());
    upload(body_invm_, binvm.data(), binvm.size());
    upload(body_flags_, bflags.data(), bflags.size());
    dalloc(body_delta_, 8 * B);
    dalloc(body_invw_, B);
    const size_t rcmax = std::max<size_t>(8 * NB, 1);
    dalloc(rc_s0_, rcmax); dalloc(rc_s1_, rcmax);
    dalloc(rc_p0_, rcmax); dalloc(rc_p1_, rcmax); dalloc(rc_n_, rcmax);
    dalloc(rc_impulse_, 6 * rcmax);

    all_verts_.clear(); all_tris_.clear(); vert_off_.clear(); tri_off_.clear(); node_off_.clear(); prim_off_.clear();
    bvh_.assign(M, MeshBvhHost{});
    info_h_.assign(M, MeshInfo{});
    size_t nb = 0, pb = 0;
    std::vector<uint32_t> prim_global;
    for (size_t c = 0; c < M; c++) {
        const uint32_t vo = uint32_t(all_verts_.size() / 3), to = uint32_t(all_tris_.size() / 3);
        vert_off_.push_back(vo);
        tri_off_.push_back(to);
        all_verts_.insert(all_verts_.end(), meshes_[c].verts.begin(), meshes_[c].verts.end());
        for (uint32_t k : meshes_[c].tris) all_tris_.push_back(k + vo);
        const uint32_t nt = uint32_t(meshes_[c].tris.size() / 3);
        lbvh_build(all_verts_.data(), all_tris_.data() + size_t(to) * 3, nt, bvh_[c]);
        if (c == 0) {   // narrowphase epsilon from mesh 0's mean edge length (scene.cpp:274-281,325)
            double s = 0.0;
            const uint32_t* tt = all_tris_.data() + size_t(to) * 3;
            for (uint32_t k = 0; k < nt; k++)
                for (int e = 0; e < 3; e++) {
                    const uint32_t i0 = tt[3 * k + e], i1 = tt[3 * k + (e + 1) % 3];
                    const float dx = all_verts_[3 * i1] - all_verts_[3 * i0];
                    const float dy = all_verts_[3 * i1 + 1] - all_verts_[3 * i0 + 1];
                    const float dz = all_verts_[3 * i1 + 2] - all_verts_[3 * i0 + 2];
                    s += std::sqrt(dx * dx + dy * dy + dz * dz);
                }
            avg_edge_ = float(s / double(size_t(nt) * 3));
        }
        node_off_.push_back(nb);
        prim_off_.push_back(pb);
        info_h_[c].root = uint32_t(nb + bvh_[c].root);
        info_h_[c].margin = meshes_[c].margin;
        info_h_[c].watertight = meshes_[c].watertight ? 1u : 0u;
        info_h_[c].prim_off = uint32_t(pb);
        info_h_[c].prim_cnt = uint32_t(bvh_[c].prim.size());
        for (uint32_t p : bvh_[c].prim) prim_global.push_back(p + to);
        nb += bvh_[c].lo.size();
        pb += bvh_[c].prim.size();
    }
    if (M > 0) {
        std::printf("[VKXPBD DIAG] finalize_meshes: n_cubes=%zu, total_nodes=%zu\n", M, nb);
        for (size_t c = 0; c < M; c++)
            std::printf("  Mesh %zu: root=%u, nodes=%zu, prims=%zu\n", c, info_h_[c].root, bvh_[c].lo.size(),
                        bvh_[c].prim.size());
        std::fflush(stdout);
    }
    all_vels_.assign(all_verts_.size(), 0.f);
    upload(verts_d_, all_verts_.data(), all_verts_.size());
    upload(vvel_d_, all_vels_.data(), all_vels_.size());
    upload(tris_d_, all_tris_.data(), all_tris_.size());
    upload(prim_d_, prim_global.data(), prim_global.size());
    prim_h_ = prim_global;
    upload_nodes();
    upload_tri_soup();
    pgrid_.assign(M, ParityGrid{});
    dgrid_.assign(M, DistGrid{});
    for (size_t c = 0; c < M; c++) build_query_grids(c);
    upload_query_grids();
    prepare_mesh_updates();
    dalloc(mcnt_, 4 * std::max<size_t>(M, 1));
    // watertight meshes walk with four lanes per particle (VKXPBD_MESH_LANES=1: one)
    const char* lenv = std::getenv("VKXPBD_MESH_LANES");
    const bool split = !(lenv && std::strcmp(lenv, "1") == 0);
    mesh_split_.assign(M, 0);
    for (size_t c = 0; c < M; c++) mesh_split_[c] = (split && meshes_[c].watertight) ? 1 : 0;
    finalized_ = true;
}

void Solver::set_particles(const float* q, const float* qd, const float* w, const float* r, const int32_t* f, size_t n) {
    require_init();
    if (n != n_) throw std::runtime_error("Particle count mismatch");
    DeviceGuard g(device_);
    HIPCHECK(hipStreamSynchronize(stream_));
    if (w) std::memcpy(invm_h_.data(), w, n * sizeof(float));
    if (r) std::memcpy(radius_h_.data(), r, n * sizeof(float));
    const float old_rmax = rmax_;
    rmax_ = 0.0f;
    for (size_t i = 0; i < n; i++) rmax_ = std::max(rmax_, std::fabs(radius_h_[i]));
    // captured frames bake the particle-pair solve's reject bound (r_max) in
    if (rmax_ != old_rmax) invalidate_graphs();
    if (!plane_bits_fixed_ && q && n) {
        // one plane bit per axis (256 virtual cells, three 8-bit radix passes
        // instead of four 7-bit ones) while the particles span at most 192
        // cells on every axis; a scene that spreads further only aliases
        // more distant cells, which the neighbour test rejects
        float lo[3], hi[3];
        for (int k = 0; k < 3; k++) lo[k] = hi[k] = q[k];
        for (size_t i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) {
                lo[k] = std::min(lo[k], q[3 * i + k]);
                hi[k] = std::max(hi[k], q[3 * i + k]);
            }
        float span = 0.0f;
        for (int k = 0; k < 3; k++) span = std::max(span, (hi[k] - lo[k]) / cfg_.search_radius);
        plane_bits_ = span <= 192.0f ? 1 : 2;
    }
    update_planes();
    // written into the current slot order: slot t holds particle ids[t]
    const std::vector<uint32_t> ids = host_ids();
    std::vector<float4> P(n), V(n);
    std::vector<int32_t> F(f ? n : 0);
    for (size_t t = 0; t < n; t++) {
        const size_t i = ids[t];
        P[t] = make_float4(q[3 * i], q[3 * i + 1], q[3 * i + 2], radius_h_[i]);
        V[t] = qd ? make_float4(qd[3 * i], qd[3 * i + 1], qd[3 * i + 2], invm_h_[i]) : make_float4(0.f, 0.f, 0.f, invm_h_[i]);
        if (f) F[t] = f[i];
    }
    if (n) {
        HIPCHECK(hipMemcpy(P_[s_].p, P.data(), n * sizeof(float4), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(V_[s_].p, V.data(), n * sizeof(float4), hipMemcpyHostToDevice));
        if (f) HIPCHECK(hipMemcpy(flags_cur(), F.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
    }
}


BodyArgs Solver::body_args() const {
    BodyArgs b;
    b.q = body_q_.p; b.qd = body_qd_.p; b.com = body_com_.p; b.invI = body_invI_.p; b.invm = body_invm_.p;
    b.flags = body_flags_.p; b.shape_body = shape_body_.p; b.mat_mu = mat_mu_.p; b.delta = body_delta_.p;
    return b;
}

// ------------------------------------------------------------- profiler ----
void Solver::stage_begin(Stage) {
    if (!profiling_) return;
    if (ev_used_ + 2 > ev_pool_.size()) {
        for (int k = 0; k < 64; k++) {
            hipEvent_t e;
            // no system-scope fence: a default event record writes back the L2,
            // which would inflate the very stage times being measured
            HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
            ev_pool_.push_back(e);
        }
    }
    cur_a_ = ev_pool_[ev_used_++];
    HIPCHECK(hipEventRecord(cur_a_, stream_));
}
void Solver::stage_end(Stage s) {
    if (!profiling_) return;
    hipEvent_t b = ev_pool_[ev_used_++];
    HIPCHECK(hipEventRecord(b, stream_));
    pending_.push_back(Pending{int(s), cur_a_, b});
}
void Solver::sync_profile() {
    if (!profiling_) return;
    HIPCHECK(hipStreamSynchronize(stream_));
    for (const Pending& p : pending_) {
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, p.a, p.b));
        stage_ms_[p.stage] += ms;
        stage_n_[p.stage] += 1;
    }
    pending_.clear();
    ev_used_ = 0;
}

// ----------------------------------------------------------- wrap planes ----
// The hash grid wraps every 128 cells, so particles far apart can share a
// cell; every neighbour query then scans them and rejects them on distance.
// A plain solver sorts by wrap plane << 21 | cell instead (two bits per axis of
// which wrap a cell lies in, 2^27 rows): a query reads only its own wrap's
// run, which holds every particle the reference would accept, in the same
// order.  That holds while no accepted pair, seen from anywhere in its query
// box, can reach 126 cells: then particles of one physical cell that lie in
// different wraps never interact.  The check uses the largest radius set so
// far, so set_particles can switch the mode (and drops captured frames).
void Solver::update_planes() {
    const float rm = std::max(rmax_, cfg_.particle_max_radius);
    const float reach = 3.0f * rm + cfg_.particle_max_radius + 2.0f * std::fabs(cfg_.particle_cohesion);
    const bool want = planes_allowed_ && num_scenes_ == 1 && reach < 126.0f * cfg_.search_radius;
    if (want == planes_ && range_.p && (!planes_ || table_bits_ == plane_bits_)) return;
    DeviceGuard g(device_);
    HIPCHECK(hipStreamSynchronize(stream_));
    invalidate_graphs();
    planes_ = want;
    table_bits_ = plane_bits_;
    if (num_scenes_ == 1) set_sort_shape(planes_ ? kCellBits + 3 * plane_bits_ : kCellBits);
    if (num_scenes_ == 1) dalloc(range_, size_t(2) * (size_t(kNumCells) << (planes_ ? 3 * plane_bits_ : 0)));
    dalloc(rkeys_, n_);   // zero keys: the first reset touches row 0 only
    HIPCHECK(hipDeviceSynchronize());
    grid_valid_ = false;
}

// Cell ranges of the sorted keys.  Every table (2^21 rows, 2^24 or 2^27 with
// wrap planes, 2^21 per batched scene) has the previous build's rows reset
// first, from their keys (rkeys): a reset proportional to the particles, not
// to the table.
void Solver::build_ranges(const uint32_t* sorted_keys, bool cleared, const SortGather* gather) {
    const int n = int(n_);
    if (num_scenes_ > 1) {
        if (!cleared) launch_clear_rows(n, rslot_.p, rows_.p, stream_);
        launch_row_table(n, sorted_keys, gtcnt_.p, gtoff_.p, rows_.p, glb_.p, rslot_.p, valsA_.p, gather, stream_);
        return;
    }
    if (!cleared) launch_clear_ranges(n, rkeys_.p, range_.p, stream_);
    launch_cell_ranges(n, sorted_keys, range_.p, rkeys_.p, stream_);
}

// ----------------------------------------------------------------- step ----
// The fewest radix passes of at most kSortUseBits that cover the key, with
// the digit width spread evenly over them (never below 7 bits).
void Solver::set_sort_shape(int key_bits) {
    sort_passes_ = (key_bits + kSortUseBits - 1) / kSortUseBits;
    sort_bits_ = std::max(7, (key_bits + sort_passes_ - 1) / sort_passes_);
}


SortGather Solver::sort_gather(float sdt) {
    return SortGather{P_[s_].p, V_[s_].p, flags_cur(), ids_cur(), Xs_[0].p, Vs_[0].p,
                      flags_[lay_ ^ 1].p, ids_[lay_ ^ 1].p, rank_.p,
                      sdt, cfg_.particle_v_max, cfg_.gravity[0], cfg_.gravity[1], cfg_.gravity[2]};
}

void Solver::run_sort(int tiles, const SortGather* last, const SortReset& rs) {
    const int n = int(n_);
    uint32_t* kb[2] = {keysA_.p, keysB_.p};
    uint32_t* vb[2] = {valsA_.p, valsB_.p};
    // pass 0 reads k_integrate's (key, slot) pairs; the passes then alternate
    // between the A and B buffers and the last one must land in A (the
    // buffers the getters read)
    int dst = (sort_passes_ & 1) ? 0 : 1;
    // chunked passes (no bin-scan launch) for digits of at most 8 bits; pass
    // p's scatter zeroes the totals pass (p + 1) mod P adds into
    const int chunks = (tiles + kSortChunkTiles - 1) / kSortChunkTiles;
    // batched solvers: whole-scene tiles, each scene sorted in its own range
    const bool segmented = num_scenes_ > 1 && seg_sort_;
    if (segmented) tiles = seg_tiles_[sort_items_ == 8 ? 1 : 0];
    const bool chunked = !segmented && sort_chunked_ && sort_bits_ <= 8 && sort_passes_ >= 2 && sort_passes_ <= kSortMaxPasses &&
                         size_t(256) * size_t(chunks) <= sort_chunk_words_;
    for (int pass = 0; pass < sort_passes_; pass++) {
        const int shift = pass * sort_bits_;
        const uint32_t* kin = pass == 0 ? nullptr : kb[dst ^ 1];
        const uint32_t* vin = pass == 0 ? nullptr : vb[dst ^ 1];
        const uint2* pin = pass == 0 ? kv_.p : nullptr;
        SortPass ch;
        if (segmented) ch.seg = seg_tab_[sort_items_ == 8 ? 1 : 0].p;
        if (chunked) {
            ch.chunk = chunks_.p + size_t(pass) * sort_chunk_words_;
            ch.chunks = chunks;
            ch.tstride = chunks * kSortChunkTiles;
            ch.zero_next = chunks_.p + size_t((pass + 1) % sort_passes_) * sort_chunk_words_;
            ch.zero_count = chunks << sort_bits_;
        }
        stage_begin(kSortUp);
        launch_sort_upsweep(sort_bits_, sort_items_, n, kin, pin, shift, hist_.p, tiles, ch, pass == 0 ? rs : SortReset{},
                            stream_);
        stage_end(kSortUp);
        if (!chunked) {
            stage_begin(kSortScan);
            launch_sort_binscan(sort_bits_, hist_.p, bintot_.p, tiles, stream_);
            stage_end(kSortScan);
        }
        stage_begin(kSortScatter);
        launch_sort_scatter(sort_bits_, sort_items_, n, kin, vin, pin, kb[dst], vb[dst], shift, hist_.p, bintot_.p, tiles,
                            pass == sort_passes_ - 1 ? last : nullptr, ch, stream_);
        stage_end(kSortScatter);
        dst ^= 1;
    }
    lay_ ^= 1;   // flags and ids of the sorted order are current from here on
}

DetApplyArgs Solver::det_apply_args(const float4* Xin, const float4* Vin, int out, float sdt) {
    DetApplyArgs d;
    d.cmax = int(cmax_);
    d.key1s = dk1s_.p; d.val1s = dv1s_.p;
    d.pad1 = det_kb1_ >= 32 ? 0xFFFFFFFFu : (1u << det_kb1_) - 1u;
    d.sbits = det_sbits_;
    d.cfd = cfd_.p; d.delta = delta_.p;
    d.Ps = P_[s_].p; d.vals = valsA_.p; d.Xs = Xin; d.Vs = Vin;
    d.Po = P_[out].p; d.Vo = V_[out].p;
    d.dt = sdt; d.vmax = cfg_.particle_v_max;
    return d;
}

BodyStepArgs Solver::body_step_args(float sdt) {
    BodyStepArgs b;
    b.b0 = int(meshes_.size());
    b.b1 = int(meshes_.size() + boxes_.size());
    b.q = body_q_.p; b.qd = body_qd_.p; b.com = body_com_.p; b.I = body_I_.p; b.invI = body_invI_.p;
    b.invm = body_invm_.p; b.flags = body_flags_.p; b.delta = body_delta_.p; b.invw = body_invw_.p;
    b.dt = sdt; b.gx = cfg_.gravity[0]; b.gy = cfg_.gravity[1]; b.gz = cfg_.gravity[2];
    b.angular_damping = angular_damping_;
    b.use_weights = use_weights_;
    return b;
}

// Substeps inside a frame may key each other (the last iteration's apply
// writes the next substep's radix input, which k_integrate would compute from
// the same state by the same code) where k_integrate has no other work that
// is read: solvers without boxes, restitution or deterministic mode whose
// plane contacts come from a mesh's candidate pass, or that have no ground
// plane, or whose plane contacts are solved inside the particle-pair solve
// (then the contact buffer is read only after the frame, and k_integrate keys
// the last substep to emit them).  A frame's first substep is always keyed by
// k_integrate: the host may have changed the state before it.
bool Solver::fuse_keys_ok(int iterations) const {
    if (!fuse_keys_ || det_ || !boxes_.empty() || cfg_.enable_restitution || iterations < 1 || n_ == 0) return false;
    if (!meshes_.empty() || !cfg_.has_ground_plane) return true;
    // without a mesh k_integrate emits the plane contacts; with the plane
    // folded into the particle-pair solve (every plane contact fits) nothing
    // reads them before the frame ends, so only the last substep needs them
    return cmax_ >= n_;
}

int Solver::frame_fuse(int sub, int substeps, int iterations) const {
    if (!fuse_keys_ok(iterations)) return 0;
    const bool last_emits = meshes_.empty() && cfg_.has_ground_plane;   // k_integrate keys the last substep
    auto keyed = [&](int k) { return k > 0 && !(last_emits && k == substeps - 1); };
    return (keyed(sub) ? kKeyed : 0) | (sub + 1 < substeps && keyed(sub + 1) ? kKeyNext : 0);
}

void Solver::substep(float sdt, int iterations, bool collide, int slot, bool zero_next, int fuse) {
    const int n = int(n_);
    const int sort_tile = 256 * sort_items_;
    const int tiles = int((n_ + sort_tile - 1) / sort_tile);
    const int nbox = int(boxes_.size());
    const int M = int(meshes_.size());
    int* cc_gen = collide ? cc_.p + slot : nullptr;
    int* cc_solve = collide ? cc_.p + slot : cc_.p + contact_slot_;
    int* rcc_solve = collide ? cc_.p + 4 + slot : cc_.p + 4 + rigid_slot_;
    const float inv_cell = 1.0f / cfg_.search_radius;
    // the mesh walk is latency bound and fills few SIMDs: it overlaps the
    // bandwidth-bound integrate and sort (unless stage times are measured)
    const bool fork = collide && !meshes_.empty() && n > 0 && !profiling_;
    // with a mesh, its candidate pass emits the plane contacts too (on the
    // side stream), so the first shape solve needs nothing from the main one
    const bool plane_in_mesh = collide && !meshes_.empty() && cfg_.has_ground_plane;
    if (fork) {
        HIPCHECK(hipEventRecord(fork_ev_, stream_));
        HIPCHECK(hipStreamWaitEvent(side_, fork_ev_, 0));
    }
    const int* flags_prev = flags_cur();   // the previous slot order's flags (the sort flips them)
    // the ground plane's contacts solved inside the particle-pair solve (the
    // shape solve then has only the mesh contacts): every plane contact
    // emitted must be stored (no overflow possible) and no box shares the
    // shape dispatch
    const bool plane_fold = collide && cfg_.has_ground_plane && !det_ && nbox == 0 && cmax_ >= n_ * (1 + size_t(M));
    // without meshes and boxes every contact is the plane's: no shape solve
    const bool shape_work = !(plane_fold && M == 0);
    // the shape solve of iteration it: the first one in the previous slot
    // order (the contacts' order), integrating the state itself
    auto shape_args = [&](int it, const float4* Xin, const float4* Vin) {
        ShapeArgs sa{};
        sa.cmax = int(cmax_); sa.cc = cc_solve;
        const bool prev_order = it == 0 && n > 0 && !det_;
        sa.cfd = det_ ? cfd_.p : nullptr;
        sa.cbd = det_ ? cbd_.p : nullptr;
        sa.cpart = cpart_.p; sa.cshape = cshape_.p; sa.cbpos = cbpos_.p; sa.cbvel = cbvel_.p; sa.cnorm = cnorm_.p;
        sa.P = prev_order ? P_[s_].p : Xin; sa.V = prev_order ? V_[s_].p : Vin;
        sa.flags = prev_order ? flags_prev : flags_cur();
        sa.from_old = prev_order ? 1 : 0;
        sa.g_dt = sdt; sa.g_vmax = cfg_.particle_v_max;
        sa.gx = cfg_.gravity[0]; sa.gy = cfg_.gravity[1]; sa.gz = cfg_.gravity[2];
        sa.body = body_args();
        sa.pmu = cfg_.soft_contact_mu; sa.ka = cfg_.particle_adhesion; sa.dt = sdt; sa.relax = relax_;
        sa.dshape = dshape_.p; sa.owner = owner_.p;
        sa.skip_plane = plane_fold ? 1 : 0;
        return sa;
    };
    // on the side stream with only mesh contacts to solve (the plane's folded
    // into the particle-pair solve, no boxes): the walk solves each contact it
    // writes, so the first shape solve is not launched
    const bool solve_in_walk = fork && !det_ && nbox == 0 && plane_fold && iterations > 0 && solve_in_walk_;
    if (collide && !meshes_.empty()) {
        MeshArgs m{};
        m.n = n; m.P = P_[s_].p; m.flags = flags_cur();
        m.verts = verts_d_.p; m.tris = tris_d_.p; m.nodes = nodes_d_.p; m.cnodes = cnodes_d_.p; m.lbox = lbox_d_.p; m.prim = prim_d_.p;
        m.pg_start = pg_start_.p; m.pg_soup = pg_soup_.p; m.tsoup = tsoup_.p; m.lb = lb_.p; m.lb_node = lb_node_.p;
        m.mcnt = mcnt_.p; m.mq = mq_.p;
        m.slot = slot;
        m.vvel = vvel_d_.p; m.info = info_d_.p; m.nmesh = M;
        m.margin = cfg_.soft_contact_margin; m.eps = avg_edge_ * 1e-3f;
        m.cmax = int(cmax_); m.cc = cc_gen;
        m.cpart = cpart_.p; m.cshape = cshape_.p; m.cbpos = cbpos_.p; m.cbvel = cbvel_.p; m.cnorm = cnorm_.p;
        m.err = reinterpret_cast<uint32_t*>(cc_.p + 2);
        m.stats = mesh_stats_on_ ? mesh_stats_.p : nullptr;
        m.plane = plane_in_mesh ? 1 : 0; m.altitude = cfg_.ground_plane_altitude;
        m.solve_inline = solve_in_walk ? 1 : 0;
        if (solve_in_walk) m.solve = shape_args(0, nullptr, nullptr);
        stage_begin(kMesh);
        launch_mesh_contacts(m, mesh_split_.data(), fork, fork ? side_ : stream_);
        stage_end(kMesh);
    }
    if (collide && nbox > 0) {
        stage_begin(kBodyContacts);
        ParticleBoxArgs pb;
        pb.n = n; pb.P = P_[s_].p; pb.flags = flags_cur();
        pb.nbox = nbox; pb.b0 = M; pb.shape0 = M + 1;
        pb.q = body_q_.p; pb.half = body_half_.p; pb.margin = cfg_.soft_contact_margin;
        pb.cmax = int(cmax_); pb.cc = cc_gen;
        pb.cpart = cpart_.p; pb.cshape = cshape_.p; pb.cbpos = cbpos_.p; pb.cbvel = cbvel_.p; pb.cnorm = cnorm_.p;
        launch_particle_box_contacts(pb, stream_);
        if (cfg_.has_ground_plane) {
            BodyContactGenArgs bg;
            bg.nbox = nbox; bg.b0 = M; bg.shape0 = M + 1;
            bg.q = body_q_.p; bg.half = body_half_.p;
            bg.margin = cfg_.soft_contact_margin; bg.altitude = cfg_.ground_plane_altitude;
            bg.rcmax = 8 * nbox; bg.rcc = rcc_solve;
            bg.s0 = rc_s0_.p; bg.s1 = rc_s1_.p; bg.p0 = rc_p0_.p; bg.p1 = rc_p1_.p; bg.n = rc_n_.p;
            bg.fixed = det_ ? 1 : 0;
            launch_body_plane_contacts(bg, stream_);
        }
        launch_fill_zero(reinterpret_cast<int*>(rc_impulse_.p), rc_impulse_.n & ~size_t(3), stream_);
        stage_end(kBodyContacts);
    }
    const bool keyed = (fuse & kKeyed) != 0;
    if (n > 0 && !keyed) {
        IntegrateArgs a;
        a.n = n;
        a.Ps = P_[s_].p; a.Vs = V_[s_].p; a.flags = flags_cur();
        a.ids = ids_cur(); a.kv = kv_.p;
        a.dt = sdt; a.vmax = cfg_.particle_v_max;
        a.gx = cfg_.gravity[0]; a.gy = cfg_.gravity[1]; a.gz = cfg_.gravity[2];
        a.inv_cell = inv_cell;
        a.has_plane = (collide && cfg_.has_ground_plane && !plane_in_mesh) ? 1 : 0;
        a.margin = cfg_.soft_contact_margin; a.altitude = cfg_.ground_plane_altitude;
        a.cmax = int(cmax_); a.cc = plane_in_mesh ? nullptr : cc_gen;
        a.cc_next = zero_next ? cc_.p + (slot ^ 1) : nullptr;
        a.rcc_next = zero_next ? cc_.p + 4 + (slot ^ 1) : nullptr;
        a.mesh_next = zero_next ? cc_.p + 6 : nullptr;
        a.scene = num_scenes_ > 1 ? scene_d_.p : nullptr;
        a.planes = planes_ ? plane_bits_ : 0;
        a.cpart = cpart_.p; a.cshape = cshape_.p; a.cbpos = cbpos_.p; a.cbvel = cbvel_.p; a.cnorm = cnorm_.p;
        a.rkeys = num_scenes_ > 1 ? nullptr : rkeys_.p; a.range = range_.p;
        a.rslot = num_scenes_ > 1 ? rslot_.p : nullptr; a.rows = rows_.p;
        stage_begin(kIntegrate);
        launch_integrate(a, stream_);
        stage_end(kIntegrate);
    }
    const BodyStepArgs bs = body_step_args(sdt);
    if (nbox > 0) {
        stage_begin(kBodyIntegrate);
        launch_integrate_bodies(bs, stream_);
        stage_end(kBodyIntegrate);
        // the first shape solve (side stream) reads the box contacts and poses
        if (fork && iterations > 0 && !det_) HIPCHECK(hipEventRecord(sorted_ev_, stream_));
    }
    const int shape_grid = int(std::max<size_t>(1, std::min<size_t>((cmax_ + 255) / 256, 2048)));
    int cur = p_, nxt = o_;
    NextKey nk{};
    if (fuse & kKeyNext) {
        nk.kv = kv_.p;
        nk.gx = cfg_.gravity[0]; nk.gy = cfg_.gravity[1]; nk.gz = cfg_.gravity[2];
        nk.inv_cell = inv_cell;
        nk.planes = planes_ ? plane_bits_ : 0;
        nk.keys = num_scenes_ > 1 ? keysA_.p : nullptr;
    }
    if (n > 0) {
        // the state moves into the sorted order in the cell-range pass (one
        // thread per sorted slot): a plain solver's k_cell_ranges_gather, a
        // batched solver's row-table build (or its last radix pass)
        const SortGather g = sort_gather(sdt);
        const bool gather_in_ranges = num_scenes_ == 1 || batch_gather_rows_;
        SortReset rs;
        if (keyed) {   // k_integrate's resets (the rows: the last shape-delta apply's, if it ran)
            if (num_scenes_ > 1) {
                rs.rslot = rslot_.p;
                rs.rows = rows_.p;
            } else if (!shape_work) {
                rs.rkeys = rkeys_.p;
                rs.range = range_.p;
            }
            rs.cc_next = zero_next ? cc_.p + (slot ^ 1) : nullptr;
        }
        run_sort(tiles, gather_in_ranges ? nullptr : &g, rs);
        stage_begin(kRanges);
        if (num_scenes_ == 1) launch_cell_ranges_gather(n, keysA_.p, valsA_.p, range_.p, rkeys_.p, g, stream_);
        else build_ranges(keysA_.p, true, gather_in_ranges ? &g : nullptr);
        stage_end(kRanges);
        grid_valid_ = true;
        if (det_ && iterations > 0) {
            // deterministic mode: every contact is in (mesh stage joined), then
            // the canonical orders: by particle slot and shape, by box and id
            if (fork) {
                HIPCHECK(hipEventRecord(join_ev_, side_));
                HIPCHECK(hipStreamWaitEvent(stream_, join_ev_, 0));
            }
            const uint32_t pad1 = det_kb1_ >= 32 ? 0xFFFFFFFFu : (1u << det_kb1_) - 1u;
            const uint32_t pad2 = det_kb2_ >= 32 ? 0xFFFFFFFFu : (1u << det_kb2_) - 1u;
            DetKeyArgs ka;
            ka.cmax = int(cmax_); ka.cc = cc_solve;
            ka.cpart = cpart_.p; ka.cshape = cshape_.p; ka.remap = rank_.p; ka.ids = ids_cur();
            ka.sbits = det_sbits_; ka.idbits = det_idbits_; ka.box_shape0 = M + 1; ka.nbox = nbox;
            ka.pad1 = pad1; ka.pad2 = pad2;
            ka.key1 = dk1_.p; ka.val1 = dv1_.p; ka.key2 = nbox ? dk2_.p : nullptr; ka.val2 = nbox ? dv2_.p : nullptr;
            launch_det_keys(ka, stream_);
            sort_run(dk1_.p, dv1_.p, dk1s_.p, dv1s_.p, cmax_, det_kb1_, dscr_.p, stream_);
            if (nbox > 0) {
                sort_run(dk2_.p, dv2_.p, dk2s_.p, dv2s_.p, cmax_, det_kb2_, dscr_.p, stream_);
                launch_fill_zero(dseg_.p, dseg_.n, stream_);
                launch_det_box_ranges(int(cmax_), dk2s_.p, pad2, det_idbits_, dseg_.p, stream_);
            }
        }
        // contacts were generated in the previous slot order: the first shape
        // solve works in it and k_shape_apply maps them through rank (the key
        // kernel in deterministic mode); with no iteration a kernel does
        if (iterations == 0) {   // no solve: the predicted state still moves to the sorted order
            if (fork) {
                HIPCHECK(hipEventRecord(join_ev_, side_));
                HIPCHECK(hipStreamWaitEvent(stream_, join_ev_, 0));
            }
            launch_remap_contacts(cpart_.p, cc_solve, int(cmax_), rank_.p, shape_grid, stream_);
            HIPCHECK(hipMemcpyAsync(P_[cur].p, Xs_[0].p, n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_));
            HIPCHECK(hipMemcpyAsync(V_[cur].p, Vs_[0].p, n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_));
        }
    }
    for (int it = 0; it < iterations; it++) {
        // iteration input: the gathered predicted state, then the previous
        // iteration's output (both in sorted order)
        const float4* Xin = it == 0 ? Xs_[0].p : P_[cur].p;
        const float4* Vin = it == 0 ? Vs_[0].p : V_[cur].p;
        const ShapeArgs sa = shape_args(it, Xin, Vin);
        const bool prev_order = it == 0 && n > 0 && !det_;
        // the first shape solve follows the mesh contacts on the side stream,
        // from the previous slot order, so it waits for nothing on the main
        // stream but the boxes; the particle-particle solve (which does not
        // read the shape deltas) runs meanwhile and k_shape_apply joins the two
        const bool side = fork && it == 0 && !det_;
        if (side && nbox > 0) HIPCHECK(hipStreamWaitEvent(side_, sorted_ev_, 0));
        stage_begin(kShape);
        if (shape_work && !(it == 0 && solve_in_walk)) launch_shape_solve(sa, shape_grid, side ? side_ : stream_);
        stage_end(kShape);
        if (side) HIPCHECK(hipEventRecord(join_ev_, side_));
        PPArgs pa{};
        pa.n = n; pa.keys = keysA_.p; pa.vals = valsA_.p; pa.ids = ids_cur(); pa.range = range_.p;
        pa.Ps = P_[s_].p; pa.Xs = Xin; pa.Vs = Vin; pa.Fs = flags_cur();
        pa.delta = delta_.p; pa.Po = P_[nxt].p; pa.Vo = V_[nxt].p;
        pa.mu = cfg_.particle_mu; pa.coh = cfg_.particle_cohesion; pa.max_r = cfg_.particle_max_radius;
        pa.dt = sdt; pa.relax = relax_; pa.inv_cell = inv_cell; pa.vmax = cfg_.particle_v_max;
        pa.rmax = rmax_;
        pa.planes = planes_ ? plane_bits_ : 0;
        pa.rowtab = rowtab();
        pa.plane_fold = plane_fold ? 1 : 0;
        pa.altitude = cfg_.ground_plane_altitude; pa.margin = cfg_.soft_contact_margin; pa.ka = cfg_.particle_adhesion;
        pa.pmu = cfg_.soft_contact_mu; pa.mat_mu0 = cfg_.shape_material_mu; pa.ppd = ppd_.p;
        pa.iter_index = uint32_t(it); pa.max_diag_pairs = max_pairs_;
        pa.diag_header = diag_header_.p; pa.diag_pairs = diag_pairs_.p;
        const bool last_it = it + 1 == iterations;
        if (last_it && nk.kv != nullptr) nk.ids = ids_cur();
        pa.nk = last_it ? nk : NextKey{};
        stage_begin(kPP);
        launch_pp_apply(pa, cfg_.enable_diagnostics != 0, num_scenes_ > 1, pp_block_, stream_);
        stage_end(kPP);
        if (side) HIPCHECK(hipStreamWaitEvent(stream_, join_ev_, 0));
        stage_begin(kShapeApply);
        if (det_) {
            DetApplyArgs da = det_apply_args(Xin, Vin, nxt, sdt);
            launch_det_shape_apply(da, stream_);
        } else if (!shape_work) {   // nothing to apply; the contacts still move to the sorted slots
            if (prev_order) launch_remap_contacts(cpart_.p, cc_solve, int(cmax_), rank_.p, shape_grid, stream_);
        } else {
            ShapeApplyArgs aa{};
            aa.cmax = int(cmax_); aa.cc = cc_solve; aa.cpart = cpart_.p;
            aa.remap = prev_order ? rank_.p : nullptr;
            aa.owner = owner_.p; aa.dshape = dshape_.p; aa.delta = delta_.p;
            aa.Ps = P_[s_].p; aa.vals = valsA_.p; aa.Xs = Xin; aa.Vs = Vin;
            aa.Po = P_[nxt].p; aa.Vo = V_[nxt].p;
            aa.dt = sdt; aa.vmax = cfg_.particle_v_max;
            aa.plane_fold = plane_fold ? 1 : 0; aa.flags = flags_cur(); aa.ppd = ppd_.p;
            aa.altitude = cfg_.ground_plane_altitude; aa.margin = cfg_.soft_contact_margin; aa.ka = cfg_.particle_adhesion;
            aa.pmu = cfg_.soft_contact_mu; aa.mat_mu0 = cfg_.shape_material_mu; aa.relax = relax_;
            aa.nk = last_it ? nk : NextKey{};
            // keying the next substep: its rows' reset, here where nothing else runs
            aa.rk = last_it && nk.kv != nullptr ? rkeys_.p : nullptr;
            aa.range = range_.p;
            aa.n = n;
            launch_shape_apply(aa, shape_grid, stream_);
        }
        stage_end(kShapeApply);
        std::swap(cur, nxt);
        if (nbox > 0) {
            BodySolveArgs bv;
            bv.rcmax = 8 * nbox; bv.rcc = rcc_solve;
            bv.s0 = rc_s0_.p; bv.s1 = rc_s1_.p; bv.p0 = rc_p0_.p; bv.p1 = rc_p1_.p; bv.n = rc_n_.p;
            bv.shape_body = shape_body_.p; bv.mu = mat_mu_.p; bv.mu_t = mat_mu_t_.p; bv.mu_r = mat_mu_r_.p;
            bv.q = body_q_.p; bv.qd = body_qd_.p; bv.com = body_com_.p; bv.invI = body_invI_.p;
            bv.invm = body_invm_.p; bv.flags = body_flags_.p;
            bv.delta = body_delta_.p; bv.invw = body_invw_.p; bv.impulse = rc_impulse_.p;
            bv.relax = rigid_relax_; bv.dt = sdt;
            bv.rcd = det_ ? rcd_.p : nullptr;
            stage_begin(kBodySolve);
            launch_body_contact_solve(bv, stream_);
            if (det_) {
                DetBodyArgs db;
                db.nbox = nbox; db.b0 = M;
                db.seg = dseg_.p; db.val2s = dv2s_.p; db.cbd = cbd_.p;
                db.rcd = rcd_.p;
                db.delta = body_delta_.p;
                launch_det_body_reduce(db, stream_);
            }
            stage_end(kBodySolve);
            stage_begin(kBodyApply);
            launch_apply_body_deltas(bs, stream_);
            stage_end(kBodyApply);
        }
    }
    if (cfg_.enable_restitution) {
        RestArgs ra;
        ra.cmax = int(cmax_); ra.cc = cc_solve;
        ra.cpart = cpart_.p; ra.cshape = cshape_.p; ra.cbpos = cbpos_.p; ra.cbvel = cbvel_.p; ra.cnorm = cnorm_.p;
        ra.old = valsA_.p; ra.P0 = P_[s_].p; ra.V0 = V_[s_].p; ra.flags = flags_cur(); ra.V = V_[cur].p;
        ra.body = body_args();
        ra.ka = cfg_.particle_adhesion; ra.restitution = cfg_.soft_contact_restitution;
        stage_begin(kRest);
        if (det_ && n > 0 && iterations > 0) launch_det_restitution(ra, det_apply_args(nullptr, nullptr, cur, sdt), stream_);
        else launch_restitution(ra, shape_grid, stream_);
        stage_end(kRest);
    }
    const int old_s = s_;
    s_ = cur;
    p_ = old_s;
    o_ = nxt;
    HIPCHECK(hipGetLastError());
}

void Solver::read_count(int slot) {
    stage_begin(kCopyOut);
    HIPCHECK(hipMemcpyAsync(count_host_, cc_.p, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, stream_));
    stage_end(kCopyOut);
    HIPCHECK(hipStreamSynchronize(stream_));
    const int32_t c = count_host_[slot];
    last_count_ = c > int32_t(cmax_) ? int32_t(cmax_) : c;
    if (count_host_[2] != 0) {
        const int32_t e = count_host_[2];
        HIPCHECK(hipMemsetAsync(cc_.p + 2, 0, sizeof(int32_t), stream_));
        HIPCHECK(hipStreamSynchronize(stream_));
        throw std::runtime_error("mesh BVH traversal exceeded its 32-entry stack (flags " + std::to_string(e) + ")");
    }
}

// ---------------------------------------------------------- frame graphs ----
// A frame (counter reset + all substeps) is captured once per rotation state of
// the three particle buffers and of the flags/ids pair, and per (dt, substeps,
// iterations), then replayed
// with one hipGraphLaunch.  Every device pointer and scalar a frame bakes in is
// fixed between invalidations: the methods that reallocate buffers or change
// kernel arguments (finalize_meshes, update_mesh, body parameters, init) drop
// the cache.
void Solver::invalidate_graphs() {
    for (FrameGraph& f : graphs_)
        if (f.exec) (void)hipGraphExecDestroy(f.exec);
    graphs_.clear();
}

void Solver::record_frame(float dt, int substeps, int iterations) {
    const float sdt = dt / float(substeps);
    for (const FrameGraph& f : graphs_)
        if (f.s == s_ && f.p == p_ && f.o == o_ && f.lay == lay_ && f.dt == dt && f.substeps == substeps &&
            f.iterations == iterations) {
            HIPCHECK(hipGraphLaunch(f.exec, stream_));
            s_ = f.s2; p_ = f.p2; o_ = f.o2; lay_ = f.lay2;
            grid_valid_ = true;
            return;
        }
    if (graphs_.size() >= 8) invalidate_graphs();
    FrameGraph f{s_, p_, o_, dt, substeps, iterations, lay_, 0, 0, 0, 0, nullptr};
    HIPCHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    hipGraph_t graph = nullptr;
    try {
        launch_frame_begin(cc_.p, mcnt_.p, int(mcnt_.n), stream_);
        for (int sub = 0; sub < substeps; sub++)
            substep(sdt, iterations, true, sub & 1, true, frame_fuse(sub, substeps, iterations));
    } catch (...) {
        (void)hipStreamEndCapture(stream_, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        s_ = f.s; p_ = f.p; o_ = f.o; lay_ = f.lay;
        throw;
    }
    HIPCHECK(hipStreamEndCapture(stream_, &graph));
    const hipError_t e = hipGraphInstantiate(&f.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHECK(e);
    f.s2 = s_; f.p2 = p_; f.o2 = o_; f.lay2 = lay_;
    graphs_.push_back(f);
    HIPCHECK(hipGraphLaunch(f.exec, stream_));
}

int Solver::autotune(int frames, float dt, int substeps, int iterations) {
    require_init();
    if (!finalized_) finalize_meshes();
    DeviceGuard g(device_);
    if (n_ == 0 || substeps <= 0) return pp_block_;
    frames = std::max(frames, 1);
    HIPCHECK(hipStreamSynchronize(stream_));
    // Snapshot of everything a frame advances and everything the getters or a
    // later step_solve_only read: the state and its slot order, the bodies,
    // the sorted keys, the deltas and the last step's contact set (the trial
    // frames overwrite all of these).  The first four follow the rotating
    // slot buffers: they are restored into whichever buffers are current.
    struct Saved {
        void* live;
        DevBuf<char> copy;
        size_t bytes;
    };
    std::vector<Saved> saved;
    auto free_saved = [&] {
        for (Saved& x : saved) dfree(x.copy);
        saved.clear();
    };
    const size_t nc = std::min<size_t>(cmax_, size_t(std::max(last_count_, 0)));
    auto keep = [&](void* live, size_t bytes) {
        Saved x{live, {}, bytes};
        if (live && bytes) {
            dalloc(x.copy, bytes);
            HIPCHECK(hipMemcpyAsync(x.copy.p, live, bytes, hipMemcpyDeviceToDevice, stream_));
        }
        saved.push_back(x);
    };
    try {
        keep(P_[s_].p, n_ * sizeof(float4));
        keep(V_[s_].p, n_ * sizeof(float4));
        keep(flags_cur(), n_ * sizeof(int32_t));
        keep(ids_cur(), n_ * sizeof(uint32_t));
        keep(body_q_.p, body_q_.n * sizeof(float4));
        keep(body_qd_.p, body_qd_.n * sizeof(float4));
        keep(keysA_.p, n_ * sizeof(uint32_t));
        keep(delta_.p, 3 * n_ * sizeof(float));
        keep(cc_.p, cc_.n * sizeof(int32_t));
        keep(cpart_.p, nc * sizeof(int32_t));
        keep(cshape_.p, nc * sizeof(int32_t));
        keep(cbpos_.p, nc * sizeof(float4));
        keep(cbvel_.p, nc * sizeof(float4));
        keep(cnorm_.p, nc * sizeof(float4));
        keep(rc_s0_.p, rc_s0_.n * sizeof(int32_t));
        keep(rc_s1_.p, rc_s1_.n * sizeof(int32_t));
        keep(rc_p0_.p, rc_p0_.n * sizeof(float4));
        keep(rc_p1_.p, rc_p1_.n * sizeof(float4));
        keep(rc_n_.p, rc_n_.n * sizeof(float4));
        HIPCHECK(hipStreamSynchronize(stream_));
    } catch (...) {
        free_saved();
        throw;
    }
    const int keep_count = last_count_, keep_cslot = contact_slot_, keep_rslot = rigid_slot_;
    const bool keep_grid = grid_valid_;
    auto restore = [&] {
        void* now[4] = {P_[s_].p, V_[s_].p, flags_cur(), ids_cur()};
        for (size_t k = 0; k < saved.size(); k++) {
            if (!saved[k].copy.p) continue;
            HIPCHECK(hipMemcpyAsync(k < 4 ? now[k] : saved[k].live, saved[k].copy.p, saved[k].bytes,
                                    hipMemcpyDeviceToDevice, stream_));
        }
        HIPCHECK(hipStreamSynchronize(stream_));
        last_count_ = keep_count;
        contact_slot_ = keep_cslot;
        rigid_slot_ = keep_rslot;
        grid_valid_ = keep_grid;
    };
    // candidates: the particle-pair block size and the radix tile (keys per
    // workgroup of the sort that builds the cell tables)
    const int blocks[] = {64, 128, 256};
    const int items[] = {4, 8};
    int best = pp_block_, best_items = sort_items_;
    double best_ms = 1e300;
    try {
        for (int it : items) {
            for (int c : blocks) {
                pp_block_ = c;
                sort_items_ = it;
                invalidate_graphs();
                step(dt, substeps, iterations);   // warm-up (and capture)
                restore();
                const auto t0 = std::chrono::high_resolution_clock::now();
                for (int f = 0; f < frames; f++) step(dt, substeps, iterations);
                const double ms =
                    std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
                restore();
                if (ms < best_ms) { best_ms = ms; best = c; best_items = it; }
            }
        }
    } catch (...) {
        free_saved();
        throw;
    }
    free_saved();
    pp_block_ = best;
    sort_items_ = best_items;
    invalidate_graphs();
    return best;
}

void Solver::set_tuning(int pp_block, int radix_tile) {
    if (pp_block != 64 && pp_block != 128 && pp_block != 256)
        throw std::runtime_error("set_tuning: pp_block must be 64, 128 or 256");
    if (radix_tile != 1024 && radix_tile != 2048)
        throw std::runtime_error("set_tuning: radix_tile must be 1024 or 2048");
    pp_block_ = pp_block;
    sort_items_ = radix_tile / 256;
    invalidate_graphs();
}

// Bits of a count: the smallest b with v < 2^b (at least 1).
static int bits_for(size_t v) {
    int b = 1;
    while (b < 63 && (size_t(1) << b) <= v) b++;
    return b;
}

void Solver::ensure_det() {
    if (!det_) return;
    const size_t nshape = meshes_.size() + 1 + boxes_.size(), nbox = boxes_.size();
    // keys stay below the padding key (all ones): shape < 2^sbits - 1 and
    // box < 2^bbits - 1 leave the top value free
    const int sbits = bits_for(nshape), idbits = bits_for(n_ > 0 ? n_ - 1 : 0), bbits = bits_for(nbox);
    const int kb1 = bits_for(n_ > 0 ? n_ - 1 : 0) + sbits, kb2 = nbox ? bbits + idbits : 0;
    if (kb1 > 32 || kb2 > 32)
        throw std::runtime_error("deterministic mode: particles x shapes exceed the 32-bit contact keys");
    const bool same = det_sbits_ == sbits && det_idbits_ == idbits && det_bbits_ == bbits && dk1_.n == cmax_ &&
                      (nbox ? rcd_.n == 2 * 8 * nbox : rcd_.n == 0) && dscr_.p;
    if (same) return;
    DeviceGuard g(device_);
    HIPCHECK(hipStreamSynchronize(stream_));
    invalidate_graphs();
    det_sbits_ = sbits; det_idbits_ = idbits; det_bbits_ = bbits; det_kb1_ = kb1; det_kb2_ = kb2;
    dalloc(dk1_, cmax_); dalloc(dv1_, cmax_); dalloc(dk1s_, cmax_); dalloc(dv1s_, cmax_);
    dalloc(cfd_, cmax_); dalloc(cbd_, 2 * cmax_);
    if (nbox) {
        dalloc(dk2_, cmax_); dalloc(dv2_, cmax_); dalloc(dk2s_, cmax_); dalloc(dv2s_, cmax_);
        dalloc(rcd_, 2 * 8 * nbox);
        dalloc(dseg_, (2 * nbox + 3) & ~size_t(3));   // cleared in 16-byte steps
    } else {
        dfree(dk2_); dfree(dv2_); dfree(dk2s_); dfree(dv2s_); dfree(rcd_); dfree(dseg_);
    }
    // the layout depends on the digit width, which is not monotonic in the
    // key width: room for the larger of the two sorts
    dalloc(dscr_, std::max(sort_scratch_bytes(cmax_, kb1), nbox ? sort_scratch_bytes(cmax_, kb2) : size_t(0)));
}

void Solver::step(float dt, int substeps, int iterations) {
    require_init();
    if (!finalized_) finalize_meshes();
    ensure_det();
    DeviceGuard g(device_);
    const auto t0 = std::chrono::high_resolution_clock::now();
    const float sdt = dt / float(substeps);
    if (use_graphs_ && !profiling_ && substeps > 0) {
        record_frame(dt, substeps, iterations);
    } else {
        launch_frame_begin(cc_.p, mcnt_.p, int(mcnt_.n), stream_);
        for (int sub = 0; sub < substeps; sub++)
            substep(sdt, iterations, true, sub & 1, true, frame_fuse(sub, substeps, iterations));
    }
    if (substeps > 0) contact_slot_ = rigid_slot_ = (substeps - 1) & 1;
    read_count(substeps > 0 ? contact_slot_ : 0);
    if (profiling_) {
        sync_profile();
        prof_substeps_ += uint64_t(std::max(substeps, 0));
        prof_wall_ms_ += std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
    }
    const char* chk = std::getenv("VKXPBD_CHECK_SORT_ORDER");
    if (chk && (std::strcmp(chk, "1") == 0 || std::strcmp(chk, "true") == 0) && n_ > 0) {
        std::vector<uint32_t> k(n_);
        get_sorted(false, k.data(), n_);
        for (size_t i = 1; i < n_; i++)
            if (k[i] < k[i - 1]) {
                std::fprintf(stderr, "[VKXPBD SORT CHECK ERROR] keys[%zu] = %u < keys[%zu] = %u!\n", i, k[i], i - 1, k[i - 1]);
                return;
            }
        std::printf("[VKXPBD DEBUG] Sort check passed: %zu keys strictly non-decreasing.\n", n_);
    }
}

void Solver::step_solve_only(float dt, int substeps, int iterations) {
    require_init();
    if (!finalized_) finalize_meshes();
    ensure_det();
    DeviceGuard g(device_);
    const float sdt = dt / float(substeps);
    for (int sub = 0; sub < substeps; sub++) substep(sdt, iterations, false, contact_slot_, false);
    HIPCHECK(hipStreamSynchronize(stream_));
    sync_profile();
}

void Solver::collide_only() {
    require_init();
    if (!finalized_) finalize_meshes();
    DeviceGuard g(device_);
    launch_frame_begin(cc_.p, mcnt_.p, int(mcnt_.n), stream_);
    contact_slot_ = 0;
    PlaneArgs pl;
    pl.n = int(n_); pl.P = P_[s_].p; pl.flags = flags_cur();
    pl.margin = cfg_.soft_contact_margin; pl.altitude = cfg_.ground_plane_altitude;
    pl.cmax = int(cmax_); pl.cc = cc_.p;
    pl.cpart = cpart_.p; pl.cshape = cshape_.p; pl.cbpos = cbpos_.p; pl.cbvel = cbvel_.p; pl.cnorm = cnorm_.p;
    if (cfg_.has_ground_plane) launch_plane_contacts(pl, stream_);
    if (!boxes_.empty()) {
        ParticleBoxArgs pb;
        pb.n = int(n_); pb.P = P_[s_].p; pb.flags = flags_cur();
        pb.nbox = int(boxes_.size()); pb.b0 = int(meshes_.size()); pb.shape0 = int(meshes_.size()) + 1;
        pb.q = body_q_.p; pb.half = body_half_.p; pb.margin = cfg_.soft_contact_margin;
        pb.cmax = int(cmax_); pb.cc = cc_.p;
        pb.cpart = cpart_.p; pb.cshape = cshape_.p; pb.cbpos = cbpos_.p; pb.cbvel = cbvel_.p; pb.cnorm = cnorm_.p;
        launch_particle_box_contacts(pb, stream_);
    }
    if (!meshes_.empty()) {
        MeshArgs m{};
        m.n = int(n_); m.P = P_[s_].p; m.flags = flags_cur();
        m.verts = verts_d_.p; m.tris = tris_d_.p; m.nodes = nodes_d_.p; m.cnodes = cnodes_d_.p; m.lbox = lbox_d_.p; m.prim = prim_d_.p;
        m.pg_start = pg_start_.p; m.pg_soup = pg_soup_.p; m.tsoup = tsoup_.p; m.lb = lb_.p; m.lb_node = lb_node_.p;
        m.mcnt = mcnt_.p; m.mq = mq_.p;
        m.slot = 0;
        m.vvel = vvel_d_.p; m.info = info_d_.p; m.nmesh = int(meshes_.size());
        m.margin = cfg_.soft_contact_margin; m.eps = avg_edge_ * 1e-3f;
        m.cmax = int(cmax_); m.cc = cc_.p;
        m.cpart = cpart_.p; m.cshape = cshape_.p; m.cbpos = cbpos_.p; m.cbvel = cbvel_.p; m.cnorm = cnorm_.p;
        m.err = reinterpret_cast<uint32_t*>(cc_.p + 2);
        m.stats = mesh_stats_on_ ? mesh_stats_.p : nullptr;
        m.plane = 0;          // the plane's contacts come from k_plane_contacts above
        m.solve_inline = 0;   // contacts only, no solve
        launch_mesh_contacts(m, mesh_split_.data(), false, stream_);
    }
    HIPCHECK(hipGetLastError());
    read_count(0);
    sync_profile();
}

// -------------------------------------------------------------- getters ----
// Slot -> particle id of the current order (host copy).
std::vector<uint32_t> Solver::host_ids() {
    std::vector<uint32_t> ids(n_);
    if (n_) HIPCHECK(hipMemcpyAsync(ids.data(), ids_cur(), n_ * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
    return ids;
}
void* Solver::stage_out(size_t words) {
    if (words > io_cap_) {
        dfree(io_dev_);
        if (io_host_) { HIPCHECK(hipHostFree(io_host_)); io_host_ = nullptr; }
        io_cap_ = 0;
        dalloc(io_dev_, words);
        HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&io_host_), words * sizeof(float), hipHostMallocDefault));
        io_cap_ = words;
    }
    return io_dev_.p;
}
void Solver::copy_out(void* out, size_t bytes) {
    if (bytes) HIPCHECK(hipMemcpyAsync(io_host_, io_dev_.p, bytes, hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
    if (bytes) std::memcpy(out, io_host_, bytes);
}
// The state lives in slot order: the getters scatter it into particle order
// on the device (scene.cpp:1247-1267 reads it in particle order), then one
// DMA brings it to pinned host memory (the reference keeps host-cached
// staging for the same reason, buffer.h:164-219).
void Solver::get_positions(float* out, size_t n) {
    require_init();
    DeviceGuard g(device_);
    n = std::min(n, n_);
    if (!n_) return;
    float* d = static_cast<float*>(stage_out(3 * n_));
    launch_scatter_f3(d, P_[s_].p, ids_cur(), int(n_), stream_);
    HIPCHECK(hipGetLastError());
    copy_out(out, 3 * n * sizeof(float));
}
void Solver::get_velocities(float* out, size_t n) {
    require_init();
    DeviceGuard g(device_);
    n = std::min(n, n_);
    if (!n_) return;
    float* d = static_cast<float*>(stage_out(3 * n_));
    launch_scatter_f3(d, V_[s_].p, ids_cur(), int(n_), stream_);
    HIPCHECK(hipGetLastError());
    copy_out(out, 3 * n * sizeof(float));
}
void Solver::get_contact_ints(bool shapes, int32_t* out, size_t cap) {
    require_init();
    DeviceGuard g(device_);
    const size_t c = std::min<size_t>(size_t(std::max(last_count_, 0)), cap);
    if (!c) return;
    int* d = static_cast<int*>(stage_out(c));
    if (shapes) HIPCHECK(hipMemcpyAsync(d, cshape_.p, c * sizeof(int32_t), hipMemcpyDeviceToDevice, stream_));
    else launch_slots_to_ids(d, cpart_.p, ids_cur(), int(c), int(n_), stream_);   // contacts name slots
    HIPCHECK(hipGetLastError());
    copy_out(out, c * sizeof(int32_t));
}
void Solver::get_contact_vec(int which, float* out, size_t cap) {
    require_init();
    DeviceGuard g(device_);
    const size_t c = std::min<size_t>(size_t(std::max(last_count_, 0)), cap);
    if (!c) return;
    const float4* src = which == 0 ? cnorm_.p : (which == 1 ? cbpos_.p : cbvel_.p);
    float* d = static_cast<float*>(stage_out(3 * c));
    launch_pack3(d, src, int(c), stream_);
    HIPCHECK(hipGetLastError());
    copy_out(out, 3 * c * sizeof(float));
}
void Solver::set_contact_data(const int32_t* p, const int32_t* s, const float* bp, const float* nrm, int count) {
    require_init();
    DeviceGuard g(device_);
    HIPCHECK(hipStreamSynchronize(stream_));
    const int32_t dev_count = std::max(count, 0);
    const size_t c = std::min<size_t>(size_t(dev_count), cmax_);
    // the solve kernels index particles and shapes with these: reject ids a
    // kernel would read out of bounds
    const int32_t nshapes = int32_t(meshes_.size()) + 1;
    for (size_t i = 0; i < c; i++) {
        if (p[i] < 0 || size_t(p[i]) >= n_) throw std::runtime_error("set_contact_data: particle index out of range");
        if (s[i] < 0 || s[i] >= nshapes) throw std::runtime_error("set_contact_data: shape index out of range");
    }
    if (!finalized_) finalize_meshes();
    last_count_ = count;
    if (c) {
        std::vector<float4> b(c), nn(c);
        for (size_t i = 0; i < c; i++) {
            b[i] = make_float4(bp[3 * i], bp[3 * i + 1], bp[3 * i + 2], 0.f);
            nn[i] = make_float4(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2], 0.f);
        }
        // particle ids -> slots of the current order
        const std::vector<uint32_t> ids = host_ids();
        std::vector<int32_t> slot_of(n_), cp(c);
        for (size_t t = 0; t < n_; t++) slot_of[ids[t]] = int32_t(t);
        for (size_t i = 0; i < c; i++) cp[i] = slot_of[size_t(p[i])];
        HIPCHECK(hipMemcpy(cpart_.p, cp.data(), c * sizeof(int32_t), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(cshape_.p, s, c * sizeof(int32_t), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(cbpos_.p, b.data(), c * sizeof(float4), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(cnorm_.p, nn.data(), c * sizeof(float4), hipMemcpyHostToDevice));
    }
    HIPCHECK(hipMemcpy(cc_.p + contact_slot_, &dev_count, sizeof(int32_t), hipMemcpyHostToDevice));
}
// With wrap planes the device order is (plane, cell, id); the getters present
// the reference's order instead: cells ascending, particle ids ascending
// within a cell (what a stable sort of plain cell keys in id order gives).
std::vector<uint64_t> Solver::reference_sorted() {
    std::vector<uint32_t> k(n_), v(n_);
    if (n_) {
        HIPCHECK(hipMemcpyAsync(k.data(), keysA_.p, n_ * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
        HIPCHECK(hipMemcpyAsync(v.data(), ids_cur(), n_ * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
    }
    HIPCHECK(hipStreamSynchronize(stream_));
    std::vector<uint64_t> kv(n_);
    for (size_t t = 0; t < n_; t++) kv[t] = (uint64_t(k[t] & ((1u << kCellBits) - 1u)) << 32) | v[t];
    std::sort(kv.begin(), kv.end());
    return kv;
}

void Solver::get_sorted(bool ids, uint32_t* out, size_t n) {
    require_init();
    DeviceGuard g(device_);
    n = std::min(n, n_);
    if (planes_) {
        const std::vector<uint64_t> kv = reference_sorted();
        for (size_t t = 0; t < n; t++) out[t] = ids ? uint32_t(kv[t]) : uint32_t(kv[t] >> 32);
        return;
    }
    // after a sort the current slot order is the sorted order: its ids are the sorted ids
    if (n) HIPCHECK(hipMemcpyAsync(out, ids ? ids_cur() : keysA_.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
}
void Solver::get_cell_table(bool ends, int32_t* out, size_t n) {
    require_init();
    if (num_scenes_ > 1) throw std::runtime_error("cell tables are per scene in a batched solver; not available");
    DeviceGuard g(device_);
    n = std::min<size_t>(n, kNumCells);
    if (!grid_valid_ || n_ == 0) { std::memset(out, 0, n * sizeof(int32_t)); return; }
    if (planes_) {   // dense tables of the reference's order (k_cell_tables on the host)
        const std::vector<uint64_t> kv = reference_sorted();
        std::memset(out, 0, n * sizeof(int32_t));
        for (size_t t = 0; t < n_; t++) {
            const uint32_t c = uint32_t(kv[t] >> 32);
            if (c >= n) continue;
            if (!ends && (t == 0 || uint32_t(kv[t - 1] >> 32) != c)) out[c] = int32_t(t);
            if (ends && (t + 1 == n_ || uint32_t(kv[t + 1] >> 32) != c)) out[c] = int32_t(t + 1);
        }
        return;
    }
    if (!tbl_.p) dalloc(tbl_, size_t(2) * kNumCells);
    HIPCHECK(hipMemsetAsync(tbl_.p, 0, size_t(2) * kNumCells * sizeof(int32_t), stream_));
    launch_cell_tables(int(n_), keysA_.p, tbl_.p, tbl_.p + kNumCells, stream_);
    HIPCHECK(hipGetLastError());
    if (n) HIPCHECK(hipMemcpyAsync(out, tbl_.p + (ends ? kNumCells : 0), n * sizeof(int32_t), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
}
void Solver::get_deltas(float* out, size_t n) {
    require_init();
    DeviceGuard g(device_);
    n = std::min(n, n_);
    if (!n_) return;
    float* d = static_cast<float*>(stage_out(3 * n_));
    launch_scatter3(d, delta_.p, ids_cur(), int(n_), stream_);   // deltas are in sorted (current) order
    HIPCHECK(hipGetLastError());
    copy_out(out, 3 * n * sizeof(float));
}
uint32_t Solver::diag_word(int k) {
    if (!initialized_) return 0;
    DeviceGuard g(device_);
    uint32_t v = 0;
    HIPCHECK(hipMemcpyAsync(&v, diag_header_.p + k, sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
    return v;
}
size_t Solver::get_diag_pairs(xpbd_pair_record* out, size_t cap) {
    const uint32_t count = diag_word(0);
    size_t rd = std::min<size_t>(std::min<size_t>(count, max_pairs_), cap);
    if (!cfg_.enable_diagnostics) rd = std::min<size_t>(rd, 1);
    if (rd == 0) return 0;
    DeviceGuard g(device_);
    HIPCHECK(hipMemcpyAsync(out, diag_pairs_.p, rd * sizeof(xpbd_pair_record), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
    return rd;
}
void Solver::reset_diag() {
    if (!initialized_) return;
    DeviceGuard g(device_);
    HIPCHECK(hipMemsetAsync(diag_header_.p, 0, 2 * sizeof(uint32_t), stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
}
size_t Solver::shape_contact_diag_len() const {
    if (!initialized_) return 0;
    return cfg_.enable_diagnostics ? size_t(4) * cmax_ * 32 : 32;
}
// Boxes from the device (update_mesh refits there), packed words from the
// host topology (mesh-local, as the reference returns them).
void Solver::bvh_half(size_t mi, bool upper, float* out, size_t cap) const {
    if (mi >= bvh_.size()) return;
    const std::vector<NodeHalf>& v = upper ? bvh_[mi].hi : bvh_[mi].lo;
    const size_t n = std::min(cap, v.size());
    if (n == 0) return;
    std::vector<float4> nodes(2 * n);
    HIPCHECK(hipStreamSynchronize(stream_));
    HIPCHECK(hipMemcpy(nodes.data(), nodes_d_.p + 2 * node_off_[mi], 2 * n * sizeof(float4), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) {
        NodeHalf h = v[i];
        const float4 f = nodes[2 * i + (upper ? 1 : 0)];
        h.x = f.x; h.y = f.y; h.z = f.z;
        std::memcpy(out + 4 * i, &h, sizeof(NodeHalf));
    }
}

void Solver::add_box_body(const float* pos, const float* quat, const float* half, float density, int flags) {
    BoxHost b;
    std::memcpy(b.pos, pos, sizeof(b.pos));
    std::memcpy(b.rot, quat, sizeof(b.rot));
    std::memcpy(b.half, half, sizeof(b.half));
    b.density = density;
    b.flags = flags;
    boxes_.push_back(b);
    finalized_ = false;
}

void Solver::set_body_params(float angular_damping, float rigid_relax, float mu_t, float mu_r, int use_weights) {
    angular_damping_ = angular_damping;
    rigid_relax_ = rigid_relax;
    mu_t_ = mu_t;
    mu_r_ = mu_r;
    use_weights_ = use_weights;
    invalidate_graphs();
    if (finalized_) {   // update the per-shape coefficients in place; body state is kept
        DeviceGuard g(device_);
        HIPCHECK(hipStreamSynchronize(stream_));
        std::vector<float> t(mat_mu_t_.n, mu_t), r(mat_mu_r_.n, mu_r);
        if (!t.empty()) HIPCHECK(hipMemcpy(mat_mu_t_.p, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
        if (!r.empty()) HIPCHECK(hipMemcpy(mat_mu_r_.p, r.data(), r.size() * sizeof(float), hipMemcpyHostToDevice));
    }
}

void Solver::get_bodies(float* q7, float* qd6, size_t cap) {
    require_init();
    if (!finalized_) finalize_meshes();
    DeviceGuard g(device_);
    const size_t M = meshes_.size(), nb = std::min(cap, boxes_.size());
    std::vector<float4> q(2 * nb), qd(2 * nb);
    if (nb) {
        HIPCHECK(hipMemcpyAsync(q.data(), body_q_.p + 2 * M, 2 * nb * sizeof(float4), hipMemcpyDeviceToHost, stream_));
        HIPCHECK(hipMemcpyAsync(qd.data(), body_qd_.p + 2 * M, 2 * nb * sizeof(float4), hipMemcpyDeviceToHost, stream_));
    }
    HIPCHECK(hipStreamSynchronize(stream_));
    for (size_t k = 0; k < nb; k++) {
        if (q7) {
            const float v[7] = {q[2 * k].x, q[2 * k].y, q[2 * k].z, q[2 * k + 1].x, q[2 * k + 1].y, q[2 * k + 1].z, q[2 * k + 1].w};
            std::memcpy(q7 + 7 * k, v, sizeof(v));
        }
        if (qd6) {
            const float v[6] = {qd[2 * k].x, qd[2 * k].y, qd[2 * k].z, qd[2 * k + 1].x, qd[2 * k + 1].y, qd[2 * k + 1].z};
            std::memcpy(qd6 + 6 * k, v, sizeof(v));
        }
    }
}

void Solver::set_body_velocities(const float* qd6, size_t n) {
    require_init();
    if (!finalized_) finalize_meshes();
    if (n != boxes_.size()) throw std::runtime_error("set_body_velocities: body count mismatch");
    DeviceGuard g(device_);
    HIPCHECK(hipStreamSynchronize(stream_));
    std::vector<float4> qd(2 * n);
    for (size_t k = 0; k < n; k++) {
        qd[2 * k] = make_float4(qd6[6 * k], qd6[6 * k + 1], qd6[6 * k + 2], 0.f);
        qd[2 * k + 1] = make_float4(qd6[6 * k + 3], qd6[6 * k + 4], qd6[6 * k + 5], 0.f);
    }
    if (n) HIPCHECK(hipMemcpy(body_qd_.p + 2 * meshes_.size(), qd.data(), 2 * n * sizeof(float4), hipMemcpyHostToDevice));
}

void Solver::profile_reset() {
    for (int k = 0; k < kNumStages; k++) { stage_ms_[k] = 0.0; stage_n_[k] = 0; }
    prof_substeps_ = 0;
    prof_wall_ms_ = 0.0;
    if (initialized_) {
        DeviceGuard g(device_);
        HIPCHECK(hipStreamSynchronize(stream_));
        dfill(mesh_stats_, 0);
    }
}

void Solver::mesh_query_stats(uint64_t out[6]) {
    require_init();
    DeviceGuard g(device_);
    uint64_t v[8];
    HIPCHECK(hipStreamSynchronize(stream_));
    HIPCHECK(hipMemcpy(v, mesh_stats_.p, sizeof(v), hipMemcpyDeviceToHost));
    std::memcpy(out, v, 6 * sizeof(uint64_t));
}

size_t Solver::profile_stages(char* names, size_t names_len, double* ms, uint64_t* launches, size_t cap) {
    size_t off = 0;
    for (int k = 0; k < kNumStages; k++) {
        if (size_t(k) < cap) {
            if (ms) ms[k] = stage_ms_[k];
            if (launches) launches[k] = stage_n_[k];
        }
        if (names) {
            const size_t L = std::strlen(kStageNames[k]) + 1;
            if (off + L <= names_len) { std::memcpy(names + off, kStageNames[k], L); off += L; }
        }
    }
    return kNumStages;
}

void Solver::print_profiler_summary() {
    if (!profiling_ || prof_substeps_ == 0) return;
    double total = 0.0;
    for (int k = 0; k < kNumStages; k++) total += stage_ms_[k];
    const double ns = double(prof_substeps_);
    const double idle = prof_wall_ms_ > total ? prof_wall_ms_ - total : 0.0;
    std::printf("\n====================================================================================\n");
    std::printf("                      VKXPBD (HIP) GPU EVENT PROFILER SUMMARY\n");
    std::printf("====================================================================================\n");
    std::printf("Total Substeps Recorded: %llu\n", (unsigned long long)prof_substeps_);
    std::printf("Wall-Clock Step Time:    %10.2f ms (%8.4f ms/substep)\n", prof_wall_ms_, prof_wall_ms_ / ns);
    std::printf("Summed GPU Active Time:  %10.2f ms (%8.4f ms/substep)\n", total, total / ns);
    std::printf("Host / Sync Idle Gap:    %10.2f ms (%8.4f ms/substep, %5.1f%% of wall)\n", idle, idle / ns,
                100.0 * idle / (prof_wall_ms_ > 0 ? prof_wall_ms_ : 1.0));
    std::printf("------------------------------------------------------------------------------------\n");
    std::printf("%-28s | %10s | %14s | %9s | %8s\n", "Kernel / Stage", "Total (ms)", "Per-Substep (ms)", "% GPU", "Launches");
    for (int k = 0; k < kNumStages; k++)
        std::printf("%-28s | %10.3f | %14.4f | %8.2f%% | %8llu\n", kStageNames[k], stage_ms_[k], stage_ms_[k] / ns,
                    100.0 * stage_ms_[k] / (total > 0 ? total : 1.0), (unsigned long long)stage_n_[k]);
    std::printf("====================================================================================\n\n");
    std::fflush(stdout);
}

}  // namespace xpbd
