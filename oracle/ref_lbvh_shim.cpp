// C entry points over the reference's own LBVH builder and CPU queries
// (src/mesh_bvh.h:24-322, src/mesh_lbvh.h:11-223), compiled from the
// reference sources where they lie.  Test infrastructure only: the tests use
// it to pin the oracle's LBVH restatement and the product's builder.
#include <cstring>
#include "mesh_bvh.h"

extern "C" {

// Builds the BVH of one mesh.  lowers/uppers receive 2*nt packed 16-byte
// node halves, prim receives nt primitive indices.
void ref_lbvh_build(const float* pts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                    float* lowers, float* uppers, uint32_t* prim, uint32_t* root) {
    vkx::MeshBvh b;
    std::vector<float> P(pts, pts + size_t(nv) * 3);
    std::vector<uint32_t> T(tris, tris + size_t(nt) * 3);
    b.build(P, T);
    std::memcpy(lowers, b.lowers.data(), 16 * b.lowers.size());
    std::memcpy(uppers, b.uppers.data(), 16 * b.uppers.size());
    std::memcpy(prim, b.primitive_indices.data(), 4 * b.primitive_indices.size());
    *root = b.root;
}

// Build, move the vertices to new_pts, refit (mesh_bvh.h:142-181).
void ref_lbvh_refit(const float* pts, const float* new_pts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                    float* lowers, float* uppers) {
    vkx::MeshBvh b;
    std::vector<float> P(pts, pts + size_t(nv) * 3);
    std::vector<uint32_t> T(tris, tris + size_t(nt) * 3);
    b.build(P, T);
    for (size_t i = 0; i < size_t(nv) * 3; i++) b.points[i] = new_pts[i];
    b.refit();
    std::memcpy(lowers, b.lowers.data(), 16 * b.lowers.size());
    std::memcpy(uppers, b.uppers.data(), 16 * b.uppers.size());
}

// Brute-force closest triangle (mesh_bvh.h:296-321).
int ref_query_closest(const float* pts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                      const float* p, float max_dist, uint32_t* face, float* point) {
    vkx::MeshBvh b;
    std::vector<float> P(pts, pts + size_t(nv) * 3);
    std::vector<uint32_t> T(tris, tris + size_t(nt) * 3);
    b.build(P, T);
    return b.query_closest(p, max_dist, *face, point) ? 1 : 0;
}

}
