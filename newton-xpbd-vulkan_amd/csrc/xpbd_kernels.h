// xpbd_kernels.h -- launch interface between the host solver and the gfx950
// kernels (argument blocks passed by value, one per kernel).
#pragma once
#include "xpbd_device.h"

namespace xpbd {

// Radix sort geometry: LSD passes of 7- to 10-bit digits over tiles of 256
// threads x 4 keys.  The host uses the fewest passes of at most kSortUseBits
// that cover the key (three 7-bit passes for the 21-bit cell id): wider
// digits leave shorter same-bin runs per tile and measured slower (3 x 9 bits
// took longer than 4 x 7).
constexpr int kSortMaxBits = 10;
constexpr int kSortMaxBins = 1 << kSortMaxBits;
constexpr int kSortUseBits = 8;
constexpr int kCellBits = 21;       // batched solver: key = scene << 21 | cell
// Keys per radix thread: 8 (2,048-key tiles) from kSortLargeN keys on, where
// 1,024-key tiles measured 0.9% slower per frame (8-bit digits); 4 below it,
// where fewer, longer tiles leave most CUs idle.
constexpr int kSortLargeN = 1 << 18;
__host__ __device__ constexpr int sort_items_for(long long n) { return n >= kSortLargeN ? 8 : 4; }

// Work the first radix upsweep does for a substep keyed by the previous one
// (k_integrate skipped): the next substep's contact counter is zeroed and,
// unless the previous substep's k_shape_apply did it, the rows of the last
// cell-range build are reset from their keys (rkeys, as k_integrate resets
// them).  Null pointers: nothing.
struct SortReset {
    const uint32_t* rkeys = nullptr;
    int* range = nullptr;
    int* cc_next = nullptr;
    const int* rslot = nullptr;   // batched solvers: the row-table rows the last build filled (-1: none)
    uint2* rows = nullptr;
};

// Options of one radix pass.  Chunked passes (no bin-scan launch): tiles are
// grouped in chunks of kSortChunkTiles; chunk holds each chunk's per-bin
// totals, chunk[c][bin] (chunks x 2^bits words), summed by the upsweep with
// device atomics; the tile counts are hist[bin][tile] with rows of tstride
// words (a multiple of kSortChunkTiles).  zero_next / zero_count: the totals
// of the next pass, zeroed by this pass's scatter.  chunk == nullptr: the
// three-launch pass (upsweep, binscan, scatter).
// Segments (seg non-null, a batched solver's scenes): tile k holds keys
// [seg[k].x, seg[k].x + seg[k].y) of one segment whose tiles are
// [seg[k].z, seg[k].w), and every segment is sorted within its own key range:
// the scatter's base of a bin is the segment's first key, plus the bin's
// earlier bins inside the segment, plus the tile's count before it inside the
// segment, all differences of the global bin scan.
constexpr int kSortChunkTiles = 32;
constexpr int kSortMaxPasses = 4;   // 32-bit keys in 8-bit digits
struct SortPass {
    const int4* seg = nullptr;
    uint32_t* chunk = nullptr;
    int chunks = 0;
    int tstride = 0;
    uint32_t* zero_next = nullptr;
    int zero_count = 0;
};

// The particle state's move into the new sorted order (k_cell_ranges_gather
// for a plain solver, k_group_build or the last radix pass for a batched
// one): the sort's values are slots of the previous order; the particle at
// slot v (P, V, flags, id) is integrated (integrate_particle with dt, vmax,
// g) into Xs, Vs at its sorted slot, its flags and id follow it, and rank
// maps each previous slot to its new one.
struct SortGather {
    const float4* P; const float4* V; const int* flags; const uint32_t* ids;
    float4* Xs; float4* Vs; int* Fs; uint32_t* ids_out; uint32_t* rank;
    float dt, vmax, gx, gy, gz;
};

struct BodyArgs {
    const float4* q;      // 2 per body: (pos, 0), (quat xyzw)
    const float4* qd;     // 2 per body: (lin, 0), (ang, 0)
    const float* com;     // 3 per body
    const float* invI;    // 9 per body (row-major as the particle-shape solve reads it)
    const float* invm;
    const int* flags;
    const int* shape_body;
    const float* mat_mu;
    float* delta;         // 8 per body: lin at +0..2, ang at +4..6
};

// The next substep's radix input written where a substep finishes its
// particles (k_pp_apply, k_shape_apply): integrate_particle on the final
// state, its cell key (with the wrap plane or the scene above it), (key, slot)
// at the particle's id.  kv == nullptr: not written (k_integrate keys the
// substep).
struct NextKey {
    uint2* kv; const uint32_t* ids;   // sorted slot -> particle id
    float gx, gy, gz, inv_cell;
    int planes;
    const uint32_t* keys;             // batched solvers: the sorted keys, whose scene bits the key keeps; else null
};

struct IntegrateArgs {
    int n;
    const float4* Ps; const float4* Vs; const int* flags;   // slot order
    const uint32_t* ids;                // particle id of each slot
    uint2* kv;                          // radix input in particle-id order: (cell key, slot)
    float dt, vmax, gx, gy, gz, inv_cell;
    int has_plane; float margin, altitude;
    int cmax; int* cc; int* cc_next; int* rcc_next; int* mesh_next;
    const int* scene;   // batched mode: scene of each particle (key = scene << 21 | cell); else null
    int planes;         // key = wrap plane << 21 | cell (plain mode only)
    int* cpart; int* cshape; float4* cbpos; float4* cbvel; float4* cnorm;
    // non-null: zero the cell-table rows the last k_cell_ranges wrote (their
    // keys, in slot order), before this substep's ranges are built
    const uint32_t* rkeys; int* range;
    // batched solvers: the row-table rows the last build filled (per sorted
    // position, -1 for none) are emptied instead
    const int* rslot; uint2* rows;
};

struct PlaneArgs {
    int n; const float4* P; const int* flags;
    float margin, altitude;
    int cmax; int* cc;
    int* cpart; int* cshape; float4* cbpos; float4* cbvel; float4* cnorm;
};

// Rigid transform of a shape's body; identity and zeros for body < 0 (the
// ground plane), as solve_particle_shape_contacts.comp:123-131.
struct BodyFrame { f3 pos; float4 rot; f3 com; f3 v; f3 w; };
__device__ __forceinline__ BodyFrame load_body(const BodyArgs& b, int body) {
    BodyFrame F{{0.f, 0.f, 0.f}, make_float4(0.f, 0.f, 0.f, 1.f), {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    if (body >= 0) {
        F.pos = xyz(b.q[2 * body]);
        F.rot = b.q[2 * body + 1];
        F.com = {b.com[3 * body], b.com[3 * body + 1], b.com[3 * body + 2]};
        F.v = xyz(b.qd[2 * body]);
        F.w = xyz(b.qd[2 * body + 1]);
    }
    return F;
}


struct ShapeArgs {
    int cmax; const int* cc;
    int* cpart; const int* cshape; const float4* cbpos; const float4* cbvel; const float4* cnorm;
    // the particle state at the contacts' slots; from_old: the substep-start
    // state of the previous slot order (where the contacts were generated),
    // integrated here over g_dt exactly as the sorted copy is (so the solve
    // need not wait for the sort), and dshape/owner indexed by those slots
    const float4* P; const float4* V; const int* flags;
    int from_old; float g_dt, g_vmax, gx, gy, gz;
    int skip_plane;             // the ground plane's contacts are left to the particle-pair solve
    BodyArgs body;
    float pmu, ka, dt, relax;
    float* dshape;
    int* owner;                 // per sorted slot: largest contact index that added to dshape, else -1
    // deterministic mode (non-null): each contact's particle delta (w = 1
    // when it contributes) and body delta (2 float4: -dv with w flag, -dw)
    // are stored per contact instead of added with atomics
    float4* cfd; float4* cbd;
};

// One particle-shape contact of solve_particle_shape_contacts.comp:121-188:
// the particle's state (P, Vv), the shape's body (-1: none) and friction mu,
// the contact's body-local point and velocity and its normal.  Returns false
// where the shader adds nothing (farther than the adhesion distance, or no
// inverse mass on either side); else the particle's delta fd, the unweighted
// delta dtot and the lever arm r.  The shape solve runs every contact
// through it, and the particle-pair solve and the shape-delta apply run the
// ground plane's through it (body -1, the contact a pure function of the
// particle's substep-start position), so all three compute the same floats.
__device__ __forceinline__ bool shape_contact(float4 P, float4 Vv, int body, const BodyArgs& B, float mu, f3 cbpos,
                                              f3 cbvel, f3 n, float ka, float dt, float relax, f3& fd, f3& dtot, f3& r) {
    f3 px = xyz(P), pv = xyz(Vv);
    BodyFrame F = load_body(B, body);
    f3 bx = F.pos + quat_rotate_sc(F.rot, cbpos);
    r = bx - (F.pos + quat_rotate_sc(F.rot, F.com));
    f3 d = px - bx;
    float c = fma_(n.z, d.z, fma_(n.x, d.x, n.y * d.y)) - P.w;
    if (c > ka) return false;
    f3 bv = F.v + cross3(F.w, r) + quat_rotate_sc(F.rot, cbvel);
    f3 v = pv - bv;
    float lambda_n = c;
    f3 delta_n = n * lambda_n;
    float vn = n.y * v.y;
    vn = fma_(n.x, v.x, vn);
    vn = fma_(n.z, v.z, vn);
    f3 vt = {fma_(-n.x, vn, v.x), fma_(-n.y, vn, v.y), fma_(-n.z, vn, v.z)};
    float w1 = Vv.w;
    float w2 = 0.0f;
    if (body >= 0) {
        f3 ang = cross3(r, n);
        f3 ra = quat_rotate_inv_sc(F.rot, ang);
        const float* I = B.invI + 9 * body;   // row-major, as this shader reads it
        f3 Ira = {I[0] * ra.x + I[1] * ra.y + I[2] * ra.z, I[3] * ra.x + I[4] * ra.y + I[5] * ra.z,
                  I[6] * ra.x + I[7] * ra.y + I[8] * ra.z};
        w2 = B.invm[body] + dot3(ra, Ira);
    }
    float denom = w1 + w2;
    if (denom == 0.0f) return false;
    float l2 = vt.y * vt.y;
    l2 = fma_(vt.x, vt.x, l2);
    l2 = fma_(vt.z, vt.z, l2);
    const RootInv vl = exact_sqrt_inv(l2);
    const float vt_len = vl.d;
    float lambda_f = gmax(mu * lambda_n, -vt_len * dt);
    f3 vt_dir = vt_len > 0.0f ? exact_div_by(vt, vt_len, vl.inv) : f3{0.f, 0.f, 0.f};
    f3 diff = {fma_(vt_dir.x, lambda_f, -delta_n.x), fma_(vt_dir.y, lambda_f, -delta_n.y),
               fma_(vt_dir.z, lambda_f, -delta_n.z)};
    dtot = exact_div(diff, denom) * relax;
    fd = w1 * dtot;
    return true;
}

// One contact of solve_particle_shape_contacts.comp:104-200 (t: its index,
// pi: its particle's slot in a.P, s: its shape): the shader's early outs, the
// particle state (integrated over g_dt when it is the substep-start state of
// the previous slot order), shape_contact, then the float atomics into the
// particle's shape sum and the body's delta; the particle is marked with its
// largest contact index (owner), which k_shape_apply consumes and clears.
// k_shape_solve runs it over the contact buffer; the mesh walk may run it on
// each contact it writes, with the same values.
__device__ __forceinline__ void shape_solve_one(const ShapeArgs& a, int t, int pi, int s, f3 bpos, f3 bvel, f3 nrm) {
    const int body = a.body.shape_body[s];
    const int pf = a.flags[pi];
    if ((pf & 1) == 0) return;
    if (pf & 2) {
        if (body < 0) return;
        if (a.body.flags[body] & 4) return;
        if (a.body.invm[body] == 0.0f) return;
    }
    float4 P = a.P[pi], Vv = a.V[pi];
    if (a.from_old) integrate_particle(P, Vv, pf, a.g_dt, a.g_vmax, a.gx, a.gy, a.gz);
    const float mu = 0.5f * (a.pmu + a.body.mat_mu[s]);
    f3 fd, dtot, r;
    if (!shape_contact(P, Vv, body, a.body, mu, bpos, bvel, nrm, a.ka, a.dt, a.relax, fd, dtot, r)) return;
    atomicMax(&a.owner[pi], t);
    atomicAdd(&a.dshape[3 * pi + 0], fd.x);
    atomicAdd(&a.dshape[3 * pi + 1], fd.y);
    atomicAdd(&a.dshape[3 * pi + 2], fd.z);
    if (body >= 0 && a.body.invm[body] > 0.0f) {
        f3 dv = {dtot.x / a.dt, dtot.y / a.dt, dtot.z / a.dt};
        f3 dw = cross3(r, dv);
        float* bd = a.body.delta + 8 * body;
        atomicAdd(bd + 0, -dv.x); atomicAdd(bd + 1, -dv.y); atomicAdd(bd + 2, -dv.z);
        atomicAdd(bd + 4, -dw.x); atomicAdd(bd + 5, -dw.y); atomicAdd(bd + 6, -dw.z);
    }
}

// Adds the shape deltas to the particles that received any, after k_pp_apply
// wrote every particle with its particle-particle delta alone.
struct ShapeApplyArgs {
    int cmax; const int* cc;
    int* cpart;
    const uint32_t* remap;   // non-null: contacts, dshape and owner name previous slots; cpart is mapped through it
    int* owner; float* dshape; float* delta;        // sorted order; owner and dshape left cleared
    // plane_fold: the particle-pair solve added each particle's ground-plane
    // contact (and kept its own 0 + pair sum in ppd); a particle with shape
    // contacts as well gets (its shape sum + the plane's) + that sum here
    int plane_fold; const int* flags; const float* ppd;
    float altitude, margin, ka, pmu, mat_mu0, relax;
    const float4* Ps; const uint32_t* vals;         // substep-start positions, sorted slot -> their slot
    const float4* Xs; const float4* Vs;             // iteration input state, sorted order
    float4* Po; float4* Vo;                         // output state, sorted order
    float dt, vmax;
    NextKey nk;
    // non-null: reset the rows of this substep's cell-range build (their keys
    // rk, n of them) for the next substep's build, which k_integrate would do
    const uint32_t* rk; int* range; int n;
};

struct PPArgs {
    int n;
    const uint32_t* keys;   // sorted keys (batched mode: scene << 21 | cell)
    const uint32_t* vals;   // sorted slot -> slot of the substep-start order
    const uint32_t* ids;    // sorted slot -> particle id
    const int* range;
    const float4* Ps;                               // substep-start positions (substep-start order)
    const float4* Xs; const float4* Vs; const int* Fs;   // iteration input state, sorted order
    float* delta;                                   // sorted order
    float4* Po; float4* Vo;                         // output state, sorted order
    float mu, coh, max_r, dt, relax, inv_cell, vmax;
    float rmax;                                     // largest particle radius (bounds every accepted pair)
    int planes;                                     // keys carry the wrap plane above the cell
    RowTable rowtab;                                // batched solvers: the cell ranges
    // plane_fold: each particle's ground-plane contact (generated from its
    // substep-start position, solved as the shape solve would) is added here
    // ahead of the pair sum; ppd receives 0 + the pair sum of such particles
    int plane_fold; float altitude, margin, ka, pmu, mat_mu0; float* ppd;
    uint32_t iter_index, max_diag_pairs;
    uint32_t* diag_header; float* diag_pairs;
    NextKey nk;
};

struct RestArgs {
    int cmax; const int* cc;
    const int* cpart; const int* cshape; const float4* cbpos; const float4* cbvel; const float4* cnorm;
    const uint32_t* old;                    // sorted slot -> substep-start slot (P0, V0)
    const float4* P0; const float4* V0; const int* flags;
    float4* V;
    BodyArgs body;
    float ka, restitution;
};

struct MeshInfo {
    uint32_t root;
    float margin;
    uint32_t watertight;
    uint32_t prim_off;
    uint32_t prim_cnt;
    // projected parity grid (watertight meshes): origin, 1/cell, cells per
    // axis, first cell in the global cell-start table
    float pg_u0, pg_w0, pg_inv;
    uint32_t pg_dim, pg_cell_off;
    // distance lower-bound grid: origin, 1/cell, cells per axis, first cell in
    // the global table, and the distance cap the bounds are valid below
    float lb_ox, lb_oy, lb_oz, lb_inv;
    uint32_t lb_nx, lb_ny, lb_nz, lb_off;
    float lb_cap;
    uint32_t root_word;   // stack word of the root
    uint32_t pad[4];
};
static_assert(sizeof(MeshInfo) == 96, "MeshInfo layout");

struct MeshArgs {
    int n; const float4* P; const int* flags;
    const float* verts; const uint32_t* tris;
    const float4* nodes;   // interleaved (lower, upper) per node
    const float4* cnodes;  // per internal node, 12 float4: (lower, upper) of both children, then of the left
                           // child's children, then of the right child's; lower.w = the child's stack word
    const float4* lbox;    // (lower, upper) of the leaf starting at each leaf-order primitive position
    const uint32_t* prim;
    const float* vvel; const MeshInfo* info; int nmesh;
    float margin, eps;
    int cmax; int* cc;
    int* cpart; int* cshape; float4* cbpos; float4* cbvel; float4* cnorm;
    uint32_t* err;
    const uint32_t* pg_start;   // projected parity grids: CSR cell starts ...
    const float4* pg_soup;      // ... and each entry's triangle corners (3 float4)
    const float4* tsoup;        // triangle corners in leaf order (3 float4; first w = triangle id)
    const float* lb;            // distance lower bounds of all meshes' grids
    const uint32_t* lb_node;    // per bound cell: node the closest-face search starts at
    uint64_t* stats;            // optional query counters (6 words), null = off
    uint32_t* mcnt;             // per (substep parity, mesh): queue length and parity-queue length
    int slot;                   // substep parity: which counter pair this substep uses
    int2* mq;                   // candidate queue: (particle, search start node word)
    // solve_inline: the walk solves each contact it writes (shape_solve_one
    // with solve: the first shape solve, in the previous slot order)
    int solve_inline = 0;
    ShapeArgs solve{};
    // plane > 0: the first mesh's candidate pass also emits the ground-plane
    // contacts (generate_plane_contacts on the same substep-start positions)
    int plane = 0; float altitude = 0.0f;
};

// Rigid bodies (dynamic bodies occupy indices [b0, b1) of the body arrays).
struct BodyStepArgs {
    int b0, b1;
    float4* q; float4* qd;
    const float* com; const float* I; const float* invI; const float* invm; const int* flags;
    float* delta; float* invw;
    float dt, gx, gy, gz, angular_damping;
    int use_weights;
};

struct BodyContactGenArgs {
    int nbox, b0, shape0;
    const float4* q; const float4* half;
    float margin, altitude;
    int rcmax; int* rcc;
    int* s0; int* s1; float4* p0; float4* p1; float4* n;
    int fixed;   // deterministic mode: corner c of box k at slot 8k + c (inactive slots: s0 = s1 = 0)
};

// Deterministic mode (k_det.hip): canonical contact keys and the sums over them.
struct DetKeyArgs {
    int cmax; const int* cc;
    int* cpart; const int* cshape; const uint32_t* remap; const uint32_t* ids;
    int sbits, idbits, box_shape0, nbox;
    uint32_t pad1, pad2;
    uint32_t* key1; uint32_t* val1; uint32_t* key2; uint32_t* val2;   // key2 null without boxes
};
struct DetApplyArgs {
    int cmax;
    const uint32_t* key1s; const uint32_t* val1s;   // key1 order
    uint32_t pad1; int sbits;
    const float4* cfd;
    float* delta;
    const float4* Ps; const uint32_t* vals;
    const float4* Xs; const float4* Vs;
    float4* Po; float4* Vo;
    float dt, vmax;
};
struct DetBodyArgs {
    int nbox, b0;
    const int* seg; const uint32_t* val2s; const float4* cbd;   // seg null: no particle contacts
    const float4* rcd;                                           // null: no corner contacts
    float* delta;
};

struct ParticleBoxArgs {
    int n; const float4* P; const int* flags;
    int nbox, b0, shape0;
    const float4* q; const float4* half;
    float margin;
    int cmax; int* cc;
    int* cpart; int* cshape; float4* cbpos; float4* cbvel; float4* cnorm;
};

struct BodySolveArgs {
    int rcmax; const int* rcc;
    const int* s0; const int* s1; const float4* p0; const float4* p1; const float4* n;
    const int* shape_body; const float* mu; const float* mu_t; const float* mu_r;
    const float4* q; const float4* qd; const float* com; const float* invI; const float* invm; const int* flags;
    float* delta; float* invw; float* impulse;
    float relax, dt;
    float4* rcd;   // deterministic mode: body a's (linear with w flag, angular) delta per contact, no atomics
};

// Reverse pass of one substep (differentiable rollout).
struct AdjArgs {
    int n;
    const uint32_t* keys; const uint32_t* vals; const int* range; const int* Fs;
    const float4* Xs; const float4* Vs;     // sorted predicted state (w: radius, inverse mass)
    const float4* P0; const float4* V0;     // substep-start state, previous order (slot vals[t])
    const float* delta;                     // final deltas, sorted order
    const float4* lx_in; const float4* lv_in;   // end-state cotangents, sorted order
    float4* lx_out; float4* lv_out;             // start-state cotangents, previous order
    float4* As;                             // apply cotangent per sorted slot
    float* shx; float* shv;
    double* relax_part;
    float mu, coh, max_r, dt, relax, inv_cell, vmax, gx, gy, gz;
    float rmax;
    int planes;
    RowTable rowtab;   // batched solvers: the cell ranges
};
struct AdjShapeArgs {
    int cmax; const int* cc;
    const int* cpart; const int* cshape; const float4* cbpos; const float4* cbvel; const float4* cnorm;
    const int* flags; const float4* Xs; const float4* Vs; const float4* As;   // sorted order
    BodyArgs body;
    float pmu, ka, dt, relax;
    float* shx; float* shv; double* relax_part;
};

void launch_adj_prep(const AdjArgs& a, hipStream_t s);
void launch_adj_shape(const AdjShapeArgs& a, int grid, hipStream_t s);
void launch_adj_pp(const AdjArgs& a, bool batched, hipStream_t s);
void launch_xyz_in(float4* dst, const float* src, int n, hipStream_t s);
void launch_xyz_out(float* dst, const float4* src, int n, hipStream_t s);
void launch_fill_f4(float4* dst, int n, hipStream_t s);

void launch_integrate_bodies(const BodyStepArgs& a, hipStream_t s);
void launch_apply_body_deltas(const BodyStepArgs& a, hipStream_t s);
void launch_body_plane_contacts(const BodyContactGenArgs& a, hipStream_t s);
void launch_particle_box_contacts(const ParticleBoxArgs& a, hipStream_t s);
void launch_body_contact_solve(const BodySolveArgs& a, hipStream_t s);

void launch_integrate(const IntegrateArgs& a, hipStream_t s);
// Animated colliders (k_meshupd.hip): refit of one height's nodes (with the
// walk records and leaf boxes), triangle corners in leaf order, and the
// distance-bound and parity grids of one mesh rebuilt in place.
struct MeshRefitArgs {
    float4* nodes; float4* rec; float4* lbox; float4* tsoup;
    const uint32_t* prim; const uint32_t* tris; const float* verts;
    uint32_t nprim;
};
struct DistGridArgs {
    int tri0, ntri;                       // the mesh's global triangles
    const uint32_t* tris; const float* verts;
    const uint32_t* tpos;                 // global triangle -> global leaf position
    double ox, oy, oz, hc;                // origin and cell edge (as the bound lookups compute cells)
    uint32_t nx, ny, nz;
    float cap;
    float* lb; uint32_t* slo; uint32_t* shi; uint32_t* lb_node;   // the mesh's cells
    const float4* nodes; const uint32_t* nfirst; const uint32_t* nlast; const uint32_t* nword;
    uint32_t root;                        // global root node
};
struct ParGridArgs {
    int tri0, ntri;
    const uint32_t* tris; const float* verts;
    int kx, ky, kz;
    double Sx, Sy, pad;
    float u0, w0, inv;
    uint32_t dim;
    uint32_t* cnt;                        // dim^2 counters (zeroed before the count)
    uint32_t* start;                      // the mesh's dim^2 + 1 cell starts
    uint32_t base;                        // the mesh's first entry
    uint32_t* total;
    float4* soup;
};
void launch_refit_level(const MeshRefitArgs& a, const uint32_t* list, int count, hipStream_t s);
void launch_tri_soup(const MeshRefitArgs& a, uint32_t p0, int count, hipStream_t s);
void launch_dist_grid(const DistGridArgs& a, hipStream_t s);
void launch_par_count_scan(const ParGridArgs& a, hipStream_t s);
void launch_par_fill(const ParGridArgs& a, hipStream_t s);
void launch_zero_u32(uint32_t* p, size_t n, hipStream_t s);

void launch_plane_contacts(const PlaneArgs& a, hipStream_t s);
void launch_cell_ranges(int n, const uint32_t* keys, int* range, uint32_t* rkeys, hipStream_t s);
void launch_cell_ranges_gather(int n, const uint32_t* keys, const uint32_t* vals, int* range, uint32_t* rkeys,
                               const SortGather& g, hipStream_t s);
// batched solvers: the row table of the sorted keys (tile scratch of
// group_tiles(n) words each), recording per sorted position the row it
// filled; empty the last build's rows
int group_tiles(int n);
void launch_row_table(int n, const uint32_t* keys, uint32_t* tilecnt, uint32_t* tileoff, uint2* rows, uint32_t* glb,
                      int* rslot, const uint32_t* vals, const SortGather* gather, hipStream_t s);
void launch_clear_rows(int n, const int* rslot, uint2* rows, hipStream_t s);
void launch_clear_ranges(int n, const uint32_t* rkeys, int* range, hipStream_t s);
void launch_remap_contacts(int* cpart, const int* cc, int cmax, const uint32_t* map, int grid, hipStream_t s);
void launch_scatter_f3(float* dst, const float4* src, const uint32_t* ids, int n, hipStream_t s);
void launch_scatter3(float* dst, const float* src, const uint32_t* ids, int n, hipStream_t s);
void launch_slots_to_ids(int* dst, const int* src, const uint32_t* ids, int count, int n, hipStream_t s);
void launch_pack3(float* dst, const float4* src, int count, hipStream_t s);
void launch_iota(uint32_t* p, int n, hipStream_t s);
void launch_gather_f3(float4* dst, const float* src, const uint32_t* ids, int n, hipStream_t s);
void launch_cell_tables(int n, const uint32_t* keys, int* starts, int* ends, hipStream_t s);
void launch_shape_solve(const ShapeArgs& a, int grid, hipStream_t s);
void launch_shape_apply(const ShapeApplyArgs& a, int grid, hipStream_t s);
// batched: cell ranges come from the row table.  block: 64, 128 or 256.
void launch_pp_apply(const PPArgs& a, bool diag, bool batched, int block, hipStream_t s);
void launch_restitution(const RestArgs& a, int grid, hipStream_t s);
void launch_det_keys(const DetKeyArgs& a, hipStream_t s);
void launch_det_box_ranges(int cmax, const uint32_t* key2s, uint32_t pad, int idbits, int* seg, hipStream_t s);
void launch_det_shape_apply(const DetApplyArgs& a, hipStream_t s);
void launch_det_body_reduce(const DetBodyArgs& a, hipStream_t s);
void launch_det_restitution(const RestArgs& r, const DetApplyArgs& a, hipStream_t s);
void launch_zero_int(int* p, int count, hipStream_t s);
void launch_frame_begin(int* cc, uint32_t* mcnt, int nmcnt, hipStream_t s);
// watertight[ci] != 0: four lanes per walk (split subtrees) for mesh ci
// beside: the mesh stage runs on a second stream beside the radix passes
void launch_mesh_contacts(const MeshArgs& a, const uint8_t* watertight, bool beside, hipStream_t s);

// Radix sort: one pass = upsweep (tile histograms) -> bin-row scan -> scatter.
// items: keys per thread (4 or 8), a tile being 256 * items keys.
// Chunked passes (ch.chunk set, bits <= 8): upsweep -> scatter, hist
// tile-major and no bin totals.
void launch_sort_upsweep(int bits, int items, int n, const uint32_t* keys, const uint2* pairs, int shift, uint32_t* hist,
                         int tiles, const SortPass& ch, const SortReset& rs, hipStream_t s);
void launch_sort_binscan(int bits, uint32_t* hist, uint32_t* bintot, int tiles, hipStream_t s);
void launch_sort_scatter(int bits, int items, int n, const uint32_t* keys_in, const uint32_t* vals_in, const uint2* pairs_in,
                         uint32_t* keys_out, uint32_t* vals_out, int shift, const uint32_t* hist, const uint32_t* bintot,
                         int tiles, const SortGather* gather, const SortPass& ch, hipStream_t s);
void launch_fill_zero(int* p, size_t count, hipStream_t s);

}  // namespace xpbd
