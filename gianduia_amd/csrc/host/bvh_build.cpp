// bvh_build.cpp — binned-SAH BVH build, 4-wide collapse and triangle packing
// on the host (see bvh_build.h for the reference functions it reproduces).
#include "bvh_build.h"
#include "task_pool.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <thread>

namespace gnd {

namespace {

constexpr int kBins = 12;          // nBuckets, bvhBuilder.cpp:79
constexpr int kMaxDepth = 63;      // leaf forced at depth 63, bvhBuilder.cpp:65
constexpr int kTaskMin = 1 << 12;  // subtrees at least this big become pool tasks

// every field is written when the node is made (slots are left uninitialised)
struct TreeNode {
    Box box;
    int kid[2];
    int axis;
    int start;
    int n;  // > 0 for leaves
    int subtreeSize;
};

// Binned SAH (bvhBuilder.cpp:59-182).  Independent subtrees are built as
// pool tasks.  Nodes live in one array sized for the worst case (2N - 1): the
// subtree over n primitives owns 2n - 1 consecutive slots (its root the
// first, then the left and the right subtree's slots), so threads never share
// a counter and a subtree's nodes stay together; the flattened tree is a
// depth-first walk, so the slot numbers do not matter.  Splitting the passes
// of the top nodes over threads as well (bounds, binning, a rank-based
// parallel form of the partition) measured slower than subtree tasks alone.
class SahBuilder {
public:
    SahBuilder(std::vector<PrimRef>& refs, TaskPool* pool)
        : refs_(refs), pool_(pool), nodes_(new TreeNode[refs.empty() ? 1 : 2 * refs.size() - 1]) {}
    int run() { return build(0, (int)refs_.size(), 0, 0); }
    const TreeNode& node(int i) const { return nodes_[i]; }
    TaskPool* pool() const { return pool_; }

private:
    std::vector<PrimRef>& refs_;
    TaskPool* pool_;
    std::unique_ptr<TreeNode[]> nodes_;

    // Bucket index of a centroid: int(12 * (c - lo) / extent), 12 -> 11
    static inline int binOf(const Vec3& c, int dim, float lo, float extent) {
        int b = (int)((float)kBins * (c[dim] - lo) / extent);
        return b == kBins ? kBins - 1 : b;
    }

    // std::partition on bidirectional iterators (the reference's partition)
    template <class G>
    int partition(int start, int end, G&& goesLeft) {
        int first = start, last = end;
        for (;;) {
            for (;;) {
                if (first == last) return first;
                if (goesLeft(refs_[first])) ++first;
                else break;
            }
            --last;
            for (;;) {
                if (first == last) return first;
                if (!goesLeft(refs_[last])) --last;
                else break;
            }
            std::swap(refs_[first], refs_[last]);
            ++first;
        }
    }

    int makeLeaf(int id, int start, int n, const Box& box) {
        TreeNode& t = nodes_[id];
        t.box = box;
        t.kid[0] = t.kid[1] = -1;
        t.axis = 0;
        t.start = start;
        t.n = n;
        t.subtreeSize = 1;
        return id;
    }

    int build(int start, int end, int depth, int id) {
        const int n = end - start;
        Box box;
        for (int i = start; i < end; ++i) box.grow(refs_[i].box);
        if (n == 1 || depth == kMaxDepth) return makeLeaf(id, start, n, box);
        Box cbox;
        for (int i = start; i < end; ++i) cbox.grow(refs_[i].centroid);
        const int dim = cbox.widestAxis();
        if (cbox.hi[dim] == cbox.lo[dim]) return makeLeaf(id, start, n, box);

        float extent = cbox.hi[dim] - cbox.lo[dim];
        if (extent == 0) extent = 1.0f;
        const float lo = cbox.lo[dim];

        int cnt[kBins] = {0};
        Box bb[kBins];
        for (int i = start; i < end; ++i) {
            const int b = binOf(refs_[i].centroid, dim, lo, extent);
            cnt[b]++;
            bb[b].grow(refs_[i].box);
        }
        // prefix / suffix sweeps (unions are exact min/max, so the sweep equals
        // the reference's per-split re-accumulation)
        Box left[kBins], right[kBins];
        int lc[kBins], rc[kBins];
        Box acc;
        int c = 0;
        for (int i = 0; i < kBins; ++i) {
            acc.grow(bb[i]);
            c += cnt[i];
            left[i] = acc;
            lc[i] = c;
        }
        acc = Box();
        c = 0;
        for (int i = kBins - 1; i >= 0; --i) {
            acc.grow(bb[i]);
            c += cnt[i];
            right[i] = acc;
            rc[i] = c;
        }
        const float parentArea = box.area();
        auto splitCost = [&](int i) {
            return 1 + ((float)lc[i] * left[i].area() + (float)rc[i + 1] * right[i + 1].area()) / parentArea;
        };
        float best = splitCost(0);
        int bestSplit = 0;
        for (int i = 1; i < kBins - 1; ++i) {
            float cost = splitCost(i);
            if (cost < best) { best = cost; bestSplit = i; }
        }
        if (!(n > 4 || best < (float)n)) return makeLeaf(id, start, n, box);

        const int mid = partition(start, end, [&](const PrimRef& r) {
            return binOf(r.centroid, dim, lo, extent) <= bestSplit;
        });

        int k0 = -1, k1 = -1;
        const int id0 = id + 1, id1 = id + 2 * (mid - start);
        if (pool_ && mid - start >= kTaskMin && end - mid >= kTaskMin) {
            std::atomic<int> pending{1};
            pool_->submit([&] {
                k0 = build(start, mid, depth + 1, id0);
                pending.fetch_sub(1, std::memory_order_release);
            });
            k1 = build(mid, end, depth + 1, id1);
            pool_->waitFor(pending);
        } else {
            k0 = build(start, mid, depth + 1, id0);
            k1 = build(mid, end, depth + 1, id1);
        }
        TreeNode& t = nodes_[id];
        t.kid[0] = k0;
        t.kid[1] = k1;
        t.box = nodes_[k0].box;
        t.box.grow(nodes_[k1].box);
        t.axis = dim;
        t.start = 0;
        t.n = 0;
        t.subtreeSize = 1 + nodes_[k0].subtreeSize + nodes_[k1].subtreeSize;
        return id;
    }
};

// depth-first flatten, left child at parent + 1 (bvhBuilder.cpp:29-44): the
// node of a subtree goes to `pos`, its left subtree from pos + 1, its right
// subtree after the left one; big subtrees are written as pool tasks
void flatten(const SahBuilder& b, int id, BinaryNode* out, uint32_t pos) {
    const TreeNode& t = b.node(id);
    BinaryNode& bn = out[pos];
    bn.box = t.box;
    if (t.n > 0) {
        bn.offset = (uint32_t)t.start;
        bn.count = (uint16_t)t.n;
        bn.axis = 0;
        return;
    }
    const uint32_t right = pos + 1 + (uint32_t)b.node(t.kid[0]).subtreeSize;
    bn.count = 0;
    bn.axis = (uint8_t)t.axis;
    bn.offset = right;
    TaskPool* pool = b.pool();
    if (pool && b.node(t.kid[0]).subtreeSize >= kTaskMin && b.node(t.kid[1]).subtreeSize >= kTaskMin) {
        std::atomic<int> pending{1};
        pool->submit([&] {
            flatten(b, t.kid[0], out, pos + 1);
            pending.fetch_sub(1, std::memory_order_release);
        });
        flatten(b, t.kid[1], out, right);
        pool->waitFor(pending);
    } else {
        flatten(b, t.kid[0], out, pos + 1);
        flatten(b, t.kid[1], out, right);
    }
}

}  // namespace

BinaryBVH buildBinary(std::vector<PrimRef>& refs, TaskPool* pool) {
    BinaryBVH out;
    if (refs.empty()) return out;
    SahBuilder b(refs, refs.size() >= (size_t)2 * kTaskMin ? pool : nullptr);
    const int root = b.run();
    out.nodes.resize(b.node(root).subtreeSize);
    flatten(b, root, out.nodes.data(), 0);
    out.order.resize(refs.size());
    // leaves own contiguous, depth-first ordered ranges of the partitioned
    // array, so slot i holds refs[i]
    for (size_t i = 0; i < refs.size(); ++i) out.order[i] = refs[i].id;
    return out;
}

namespace {
// Greedy largest-area collapse (bvhBuilder.cpp:184-262).  A wide node's
// children come from expanding its binary root: repeatedly replace the inner
// candidate with the largest box area by its two children, up to 4.  In the
// binary array (depth-first, left child at e + 1, right child at offset) the
// subtree of node e spans [e, end), so every candidate carries its end.
struct Cand {
    uint32_t node, end;
};
int expand(const std::vector<BinaryNode>& bin, Cand root, Cand* cand) {
    int nc = 1;
    cand[0] = root;
    while (nc < 4) {
        int best = -1;
        float bestArea = -1.0f;
        for (int i = 0; i < nc; ++i) {
            const BinaryNode& c = bin[cand[i].node];
            if (c.count == 0) {
                float a = c.box.area();
                if (a > bestArea) { bestArea = a; best = i; }
            }
        }
        if (best < 0) break;
        const Cand e = cand[best];
        for (int i = best; i + 1 < nc; ++i) cand[i] = cand[i + 1];
        cand[nc - 1] = {e.node + 1, bin[e.node].offset};
        cand[nc] = {bin[e.node].offset, e.end};
        ++nc;
    }
    return nc;
}

constexpr uint32_t kCollapseTask = 1 << 14;  // binary subtrees at least this big become tasks

// Wide nodes of the collapsed subtree of each wide root (indexed by binary
// node), so the write pass can place subtrees at their pre-order offsets.
uint32_t countWide(const std::vector<BinaryNode>& bin, Cand r, std::vector<uint32_t>& count, TaskPool* pool) {
    Cand cand[4];
    const int nc = expand(bin, r, cand);
    uint32_t total = 1;
    std::atomic<int> pending{0};
    uint32_t sub[4] = {0, 0, 0, 0};
    for (int i = 0; i < nc; ++i) {
        if (bin[cand[i].node].count > 0) continue;
        if (pool && cand[i].end - cand[i].node >= kCollapseTask) {
            pending.fetch_add(1, std::memory_order_relaxed);
            pool->submit([&, i] {
                sub[i] = countWide(bin, cand[i], count, pool);
                pending.fetch_sub(1, std::memory_order_release);
            });
        } else {
            sub[i] = countWide(bin, cand[i], count, nullptr);
        }
    }
    if (pool) pool->waitFor(pending);
    for (int i = 0; i < nc; ++i) total += sub[i];
    count[r.node] = total;
    return total;
}

void writeWide(const std::vector<BinaryNode>& bin, Cand r, uint32_t me, const std::vector<uint32_t>& count,
               gnd_bvh4_node* out, TaskPool* pool) {
    Cand cand[4];
    const int nc = expand(bin, r, cand);
    uint8_t type[4];
    uint32_t off[4], cnt[4];
    uint32_t next = me + 1;
    std::atomic<int> pending{0};
    for (int i = 0; i < nc; ++i) {
        const BinaryNode& c = bin[cand[i].node];
        if (c.count > 0) {
            type[i] = 2;
            off[i] = c.offset;
            cnt[i] = c.count;
            continue;
        }
        type[i] = 1;
        off[i] = next;
        cnt[i] = 0;
        const uint32_t at = next;
        next += count[cand[i].node];
        if (pool && cand[i].end - cand[i].node >= kCollapseTask) {
            pending.fetch_add(1, std::memory_order_relaxed);
            pool->submit([&, i, at] {
                writeWide(bin, cand[i], at, count, out, pool);
                pending.fetch_sub(1, std::memory_order_release);
            });
        } else {
            writeWide(bin, cand[i], at, count, out, nullptr);
        }
    }
    gnd_bvh4_node& n4 = out[me];
    std::memset(&n4, 0, sizeof(n4));
    const float inf = std::numeric_limits<float>::infinity();
    for (int i = 0; i < 4; ++i) {
        if (i < nc) {
            const Box& b = bin[cand[i].node].box;
            n4.child_type[i] = type[i];
            n4.offset[i] = off[i];
            n4.count[i] = cnt[i];
            n4.min_x[i] = b.lo.x; n4.min_y[i] = b.lo.y; n4.min_z[i] = b.lo.z;
            n4.max_x[i] = b.hi.x; n4.max_y[i] = b.hi.y; n4.max_z[i] = b.hi.z;
        } else {
            n4.child_type[i] = 0;
            n4.min_x[i] = n4.min_y[i] = n4.min_z[i] = inf;
            n4.max_x[i] = n4.max_y[i] = n4.max_z[i] = -inf;
        }
    }
    if (pool) pool->waitFor(pending);
}
}  // namespace

std::vector<gnd_bvh4_node> collapseWide(const std::vector<BinaryNode>& nodes, TaskPool* pool) {
    std::vector<gnd_bvh4_node> out;
    if (nodes.empty()) return out;
    const Cand root{0, (uint32_t)nodes.size()};
    std::vector<uint32_t> count(nodes.size(), 0);
    const uint32_t n = countWide(nodes, root, count, pool);
    out.resize(n);
    writeWide(nodes, root, 0, count, out.data(), pool);
    return out;
}

// runs fn(lo, hi) over [0, n) on the pool, or inline
template <class F>
void forRange(TaskPool* pool, int64_t n, F&& fn) {
    if (pool && n >= 4096) pool->parallelFor(n, fn);
    else fn(0, n);
}

void buildMesh(const float* P, uint32_t ntris, TaskPool* pool, MeshBuild& out) {
    std::vector<PrimRef> refs(ntris);
    forRange(pool, ntris, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            Box b;
            for (int k = 0; k < 3; ++k) b.grow(Vec3(P[9 * i + 3 * k], P[9 * i + 3 * k + 1], P[9 * i + 3 * k + 2]));
            refs[i] = {b, b.center(), (int)i};
        }
    });
    BinaryBVH bvh = buildBinary(refs, pool);
    if (!bvh.nodes.empty()) out.bounds = bvh.nodes[0].box;
    // packs of each leaf in node order: offsets by a prefix sum, then the
    // (gather-bound) packing in parallel
    std::vector<uint32_t> leaves, packBase;
    uint32_t npacks = 0;
    for (uint32_t i = 0; i < (uint32_t)bvh.nodes.size(); ++i) {
        if (bvh.nodes[i].count == 0) continue;
        leaves.push_back(i);
        packBase.push_back(npacks);
        npacks += (bvh.nodes[i].count + 3u) / 4u;
    }
    out.packs.clear();
    out.packs.resize(npacks);  // every lane is written below
    forRange(pool, (int64_t)leaves.size(), [&](int64_t lo, int64_t hi) {
        for (int64_t li = lo; li < hi; ++li) {
            BinaryNode& node = bvh.nodes[leaves[li]];
            const int first = (int)node.offset, count = node.count;
            const int np = (count + 3) / 4;
            for (int p = 0; p < np; ++p) {
                gnd_tri_pack4& pk = out.packs[packBase[li] + p];
                for (int l = 0; l < 4; ++l) {
                    int k = p * 4 + l;
                    if (k >= count) {  // padding lane: zero vertices, no primitive
                        pk.prim[l] = 0xFFFFFFFFu;
                        pk.v0x[l] = pk.v0y[l] = pk.v0z[l] = 0.0f;
                        pk.v1x[l] = pk.v1y[l] = pk.v1z[l] = 0.0f;
                        pk.v2x[l] = pk.v2y[l] = pk.v2z[l] = 0.0f;
                        continue;
                    }
                    uint32_t t = (uint32_t)bvh.order[first + k];
                    pk.prim[l] = t;
                    const float* v = P + 9 * t;
                    pk.v0x[l] = v[0]; pk.v0y[l] = v[1]; pk.v0z[l] = v[2];
                    pk.v1x[l] = v[3]; pk.v1y[l] = v[4]; pk.v1z[l] = v[5];
                    pk.v2x[l] = v[6]; pk.v2y[l] = v[7]; pk.v2z[l] = v[8];
                }
            }
            node.offset = packBase[li];
            node.count = (uint16_t)np;
        }
    });
    out.nodes = collapseWide(bvh.nodes, pool);
    out.cdf.assign(ntris + 1, 0.0f);
    float total = 0.0f;
    for (uint32_t i = 0; i < ntris; ++i) {
        const float* v = P + 9 * i;
        Vec3 p0(v[0], v[1], v[2]), p1(v[3], v[4], v[5]), p2(v[6], v[7], v[8]);
        float a = 0.5f * length(cross(p1 - p0, p2 - p0));
        total += a;
        out.cdf[i + 1] = total;
    }
    out.totalArea = total;
    if (total > 0.0f)
        for (uint32_t i = 1; i <= ntris; ++i) out.cdf[i] /= total;
}

}  // namespace gnd
