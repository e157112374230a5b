// lbvh.h -- triangle-mesh LBVH in the packed node format the narrowphase
// reads (src/mesh_bvh.h:14-20): per node two 16-byte halves, lower =
// (min xyz, left child | leaf flag << 30), upper = (max xyz, right child);
// a leaf stores the primitive range [lo, hi) in the two packed words.
#pragma once
#include <cstdint>
#include <vector>

namespace xpbd {

struct NodeHalf {
    float x, y, z;
    uint32_t packed;
};
static_assert(sizeof(NodeHalf) == 16, "node half must be 16 bytes");

struct MeshBvhHost {
    std::vector<NodeHalf> lo, hi;   // 2 * num_tris nodes
    std::vector<uint32_t> prim;     // leaf-order triangle ids (mesh local)
    uint32_t root = 0;
};

// Builds the BVH of one mesh.  pts: vertex xyz; tris: 3 vertex indices per
// triangle (indices into pts).  Same construction as the reference
// (src/mesh_lbvh.h:49-223): 10-bit-per-axis Morton codes of triangle-box
// centroids, stable order, agglomerative bottom-up pairing on the
// common-prefix lengths, then subtrees spanning <= 4 primitives (or at depth
// >= 32) collapse to packed leaves.
void lbvh_build(const float* pts, const uint32_t* tris, uint32_t num_tris, MeshBvhHost& out);

// Recomputes node boxes from the current vertex positions over the nodes
// reachable from the root (src/mesh_bvh.h:142-181); topology is unchanged.
void lbvh_refit(MeshBvhHost& b, const float* pts, const uint32_t* tris);

// Grid of triangle projections along the fixed parity-ray direction of the
// watertight inside/outside test (generate_mesh_contacts.comp:375).  A ray
// with that direction can only hit a triangle whose projection contains the
// ray origin's projection; each triangle is registered in every cell its
// projected box, grown by `pad`, touches.  start has dim*dim + 1 entries.
struct ParityGrid {
    float u0 = 0.f, w0 = 0.f, inv = 0.f;
    uint32_t dim = 0;
    std::vector<uint32_t> start, tri;
};

// Projection of the parity ray: the axes (kz the dominant direction) and the
// shears of generate_mesh_contacts.comp's watertight ray test, in float.
struct RayAxes { int kx, ky, kz; float Sx, Sy; };
RayAxes parity_axes();

// pts/tris: vertex xyz and global vertex indices; tri_base: global id of the
// mesh's first triangle (stored in the grid).
void parity_grid_build(const float* pts, const uint32_t* tris, uint32_t num_tris, uint32_t tri_base, ParityGrid& g);

// Uniform grid over a mesh's box grown by `cap`; each cell holds a lower bound
// of the distance from any point of the cell to any triangle of the mesh
// (box-to-triangle-box distance, rounded down), capped at `cap`.  A point
// whose cell bound exceeds a query distance d < cap has no triangle within d,
// and a point outside the grid is farther than cap from every triangle.
//
// With a BVH, each cell also records the deepest node whose subtree holds
// every triangle within `cap` of the cell (the lowest common ancestor of
// their leaf positions; the root when none).  The depth-first closest-face
// search started there visits the same candidates in the same order as the
// search from the root, whose other branches hold no triangle within reach.
struct DistGrid {
    float ox = 0.f, oy = 0.f, oz = 0.f, inv = 0.f, cap = 0.f;
    double h = 0.0;                // cell edge the dimensions were derived with
    uint32_t nx = 0, ny = 0, nz = 0;
    std::vector<float> lb;
    std::vector<uint32_t> start;   // local node index per cell (with a BVH)
};
void dist_grid_build(const float* pts, const uint32_t* tris, uint32_t num_tris, float cap, DistGrid& g,
                     const MeshBvhHost* bvh = nullptr);
// Bound for a point in the device's float arithmetic; -1 outside the grid.
float dist_grid_bound(const DistGrid& g, const float p[3]);

// Cell of a ray origin, in the device's float arithmetic; -1 outside the grid.
long parity_grid_cell(const ParityGrid& g, const float p[3]);

}  // namespace xpbd
