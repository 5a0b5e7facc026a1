// bvh_build.h — host BVH construction for the GPU scene layout.
//
// Produces exactly the trees of the reference's BVHBuilder (12-bucket binned
// SAH, src/core/bvhBuilder.cpp:59-182) collapsed to 4-wide nodes
// (bvhBuilder.cpp:184-262), and the TrianglePack4 packing of Mesh::activate
// (src/shapes/mesh.cpp:57-106).  Large subtrees are built on worker threads;
// the result does not depend on the thread count.
#pragma once

#include <cstdint>
#include <vector>

#include "../../../include/gnd_hip.h"
#include "hmath.h"

namespace gnd {

struct PrimRef {
    Box box;
    Vec3 centroid;
    int id;
};

struct BinaryNode {
    Box box;
    uint32_t offset;      // leaf: first primitive; inner: right child
    uint16_t count;       // 0 for inner nodes (uint16 as in the reference BVHNode)
    uint8_t axis;
};

struct BinaryBVH {
    std::vector<BinaryNode> nodes;
    std::vector<int> order;  // leaf slot -> original primitive id
};

// Binned SAH build; `refs` is reordered in place.
BinaryBVH buildBinary(std::vector<PrimRef>& refs, int threads);

// Greedy largest-area collapse of a binary BVH into 4-wide nodes.
std::vector<gnd_bvh4_node> collapseWide(const std::vector<BinaryNode>& nodes);

struct MeshBuild {
    std::vector<gnd_bvh4_node> nodes;
    std::vector<gnd_tri_pack4> packs;
    std::vector<float> cdf;
    float totalArea = 0.0f;
    Box bounds;
};

// Mesh::activate: BLAS + packs + area CDF over de-indexed corner positions.
void buildMesh(const float* positions, uint32_t ntris, int threads, MeshBuild& out);

}  // namespace gnd
