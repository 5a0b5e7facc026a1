// bvh_build.h — host BVH construction for the GPU scene layout.
//
// Produces exactly the trees of the reference's BVHBuilder (12-bucket binned
// SAH, src/core/bvhBuilder.cpp:59-182) collapsed to 4-wide nodes
// (bvhBuilder.cpp:184-262), and the TrianglePack4 packing of Mesh::activate
// (src/shapes/mesh.cpp:57-106).  Large subtrees are built as tasks of a
// TaskPool (task_pool.h); the result does not depend on the thread count.
#pragma once

#include <cstdint>
#include <memory>
#include <utility>
#include <vector>

#include "../../../include/gnd_hip.h"
#include "hmath.h"

namespace gnd {

// Allocator that leaves trivially constructible elements uninitialised on
// resize, for arrays the build writes in full right after sizing them.
template <class T>
struct UninitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = UninitAlloc<U>;
    };
    UninitAlloc() = default;
    template <class U>
    UninitAlloc(const UninitAlloc<U>&) noexcept {}
    template <class U>
    void construct(U* p) noexcept {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... Args>
    void construct(U* p, Args&&... args) {
        ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...);
    }
};
template <class T>
using BuildVector = std::vector<T, UninitAlloc<T>>;

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

class TaskPool;

// Binned SAH build; `refs` is reordered in place.  pool may be null (one thread).
BinaryBVH buildBinary(std::vector<PrimRef>& refs, TaskPool* pool);

// Greedy largest-area collapse of a binary BVH into 4-wide nodes (pre-order
// layout); pool may be null.
std::vector<gnd_bvh4_node> collapseWide(const std::vector<BinaryNode>& nodes, TaskPool* pool = nullptr);

struct MeshBuild {
    std::vector<gnd_bvh4_node> nodes;
    BuildVector<gnd_tri_pack4> packs;
    std::vector<float> cdf;
    float totalArea = 0.0f;
    Box bounds;
};

// Mesh::activate: BLAS + packs + area CDF over de-indexed corner positions.
void buildMesh(const float* positions, uint32_t ntris, TaskPool* pool, MeshBuild& out);

}  // namespace gnd
