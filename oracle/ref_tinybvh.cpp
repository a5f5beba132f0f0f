// ref_tinybvh.cpp — driver that builds the reference's own vendored tinybvh
// (packages/tinybvh/tiny_bvh.h, v1.6.7) into oracle/_ref, with the vector
// types the reference plugs into it (glm, gaussian_bvh.cuh:3-16).
//
// TEST INFRASTRUCTURE.  Compiled only where the reference tree is present;
// it pins orc_bvh.cpp.  Nothing here is product code.
#include <cstdint>
#include <vector>

#include <glm/glm.hpp>

namespace tinybvh {
using bvhint2 = glm::ivec2;
using bvhint3 = glm::ivec3;
using bvhuint2 = glm::uvec2;
using bvhuint3 = glm::uvec3;
using bvhuint4 = glm::uvec4;
using bvhvec2 = glm::vec2;
using bvhvec3 = glm::vec3;
using bvhvec4 = glm::vec4;
using bvhdbl3 = glm::dvec3;
using bvhmat4 = glm::mat4x4;
}  // namespace tinybvh
#define TINYBVH_USE_CUSTOM_VECTOR_TYPES
#define TINYBVH_IMPLEMENTATION
#include "tiny_bvh.h"

extern "C" {

// BVH::Build over n triangles (3 x vec4 each) and the leaf walk the reference
// makes of it (right child first, primIdx positions; gaussian_bvh.cu:14-34).
// leaf_order receives n entries.  Returns 0, or -1 if the walk does not
// cover n positions.
int ref_tinybvh_leaf_order(const float* verts, uint32_t n, uint32_t* leaf_order) {
    tinybvh::BVH bvh;
    bvh.Build(reinterpret_cast<const tinybvh::bvhvec4*>(verts), n);
    uint32_t k = 0;
    std::vector<uint32_t> st{0u};
    while (!st.empty()) {
        const uint32_t id = st.back();
        st.pop_back();
        if (id >= bvh.allocatedNodes) continue;
        const tinybvh::BVH::BVHNode& nd = bvh.bvhNode[id];
        if (nd.isLeaf()) {
            if (nd.leftFirst >= n) continue;
            for (uint32_t i = 0; i < nd.triCount && k < n; ++i) leaf_order[k++] = nd.leftFirst + i;
        } else {
            st.push_back(nd.leftFirst);
            st.push_back(nd.leftFirst + 1);
        }
    }
    return k == n ? 0 : -1;
}

}  // extern "C"
