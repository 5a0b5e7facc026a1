// bvh_build.cpp — binned-SAH BVH build, 4-wide collapse and triangle packing
// on the host (see bvh_build.h for the reference functions it reproduces).
#include "bvh_build.h"

#include <atomic>
#include <future>
#include <memory>
#include <mutex>
#include <stdexcept>

namespace gnd {

namespace {

constexpr int kBins = 12;          // nBuckets, bvhBuilder.cpp:79
constexpr int kMaxDepth = 63;      // leaf forced at depth 63, bvhBuilder.cpp:65
constexpr int kParallelMin = 1 << 15;

struct TreeNode {
    Box box;
    TreeNode* kid[2] = {nullptr, nullptr};
    int axis = 0;
    int start = 0;
    int n = 0;  // > 0 for leaves
    int subtreeSize = 1;
};

class SahBuilder {
public:
    SahBuilder(std::vector<PrimRef>& refs, int threads) : refs_(refs), budget_(threads > 1 ? threads - 1 : 0) {}

    TreeNode* run() { return build(0, (int)refs_.size(), 0); }

    ~SahBuilder() {
        for (auto& p : pools_) delete[] p.first;
    }

private:
    std::vector<PrimRef>& refs_;
    std::atomic<int> budget_;
    std::mutex poolMutex_;
    std::vector<std::pair<TreeNode*, size_t>> pools_;

    TreeNode* newNode() {
        // small per-call allocations are fine at these node counts; keep
        // ownership in pools_ so the tree is freed in one place
        TreeNode* n = new TreeNode[1];
        std::lock_guard<std::mutex> g(poolMutex_);
        pools_.push_back({n, 1});
        return n;
    }

    // Bucket index of a centroid: int(12 * (c - lo) / extent), 12 -> 11
    static inline int binOf(const Vec3& c, int dim, float lo, float extent) {
        int b = (int)((float)kBins * (c[dim] - lo) / extent);
        return b == kBins ? kBins - 1 : b;
    }

    TreeNode* makeLeaf(TreeNode* node, int start, int n, const Box& box) {
        node->start = start;
        node->n = n;
        node->box = box;
        return node;
    }

    TreeNode* build(int start, int end, int depth) {
        TreeNode* node = newNode();
        const int n = end - start;
        Box box;
        for (int i = start; i < end; ++i) box.grow(refs_[i].box);
        if (n == 1 || depth == kMaxDepth) return makeLeaf(node, start, n, box);

        Box cbox;
        for (int i = start; i < end; ++i) cbox.grow(refs_[i].centroid);
        const int dim = cbox.widestAxis();
        if (cbox.hi[dim] == cbox.lo[dim]) return makeLeaf(node, start, n, box);

        float extent = cbox.hi[dim] - cbox.lo[dim];
        if (extent == 0) extent = 1.0f;
        const float lo = cbox.lo[dim];

        int cnt[kBins] = {0};
        Box bb[kBins];
        for (int i = start; i < end; ++i) {
            int b = binOf(refs_[i].centroid, dim, lo, extent);
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
        if (!(n > 4 || best < (float)n)) return makeLeaf(node, start, n, box);

        // Hoare partition (std::partition on bidirectional iterators)
        int first = start, last = end;
        auto goesLeft = [&](const PrimRef& r) { return binOf(r.centroid, dim, lo, extent) <= bestSplit; };
        for (;;) {
            for (;;) {
                if (first == last) goto partitioned;
                if (goesLeft(refs_[first])) ++first;
                else break;
            }
            --last;
            for (;;) {
                if (first == last) goto partitioned;
                if (!goesLeft(refs_[last])) --last;
                else break;
            }
            std::swap(refs_[first], refs_[last]);
            ++first;
        }
    partitioned:
        const int mid = first;

        TreeNode *k0 = nullptr, *k1 = nullptr;
        bool spawned = false;
        if (n > kParallelMin) {
            int b = budget_.load();
            while (b > 0 && !budget_.compare_exchange_weak(b, b - 1)) {
            }
            if (b > 0) {
                spawned = true;
                auto fut = std::async(std::launch::async, [&] { return build(start, mid, depth + 1); });
                k1 = build(mid, end, depth + 1);
                k0 = fut.get();
                budget_.fetch_add(1);
            }
        }
        if (!spawned) {
            k0 = build(start, mid, depth + 1);
            k1 = build(mid, end, depth + 1);
        }
        node->kid[0] = k0;
        node->kid[1] = k1;
        node->box = k0->box;
        node->box.grow(k1->box);
        node->axis = dim;
        node->n = 0;
        node->subtreeSize = 1 + k0->subtreeSize + k1->subtreeSize;
        return node;
    }
};

// depth-first flatten, left child at parent + 1 (bvhBuilder.cpp:29-44)
uint32_t flatten(const TreeNode* t, std::vector<BinaryNode>& out, uint32_t& next) {
    uint32_t me = next++;
    BinaryNode& bn = out[me];
    bn.box = t->box;
    if (t->n > 0) {
        bn.offset = (uint32_t)t->start;
        bn.count = (uint16_t)t->n;
        bn.axis = 0;
        return me;
    }
    bn.count = 0;
    bn.axis = (uint8_t)t->axis;
    flatten(t->kid[0], out, next);
    uint32_t r = flatten(t->kid[1], out, next);
    out[me].offset = r;
    return me;
}

}  // namespace

BinaryBVH buildBinary(std::vector<PrimRef>& refs, int threads) {
    BinaryBVH out;
    if (refs.empty()) return out;
    SahBuilder b(refs, threads);
    TreeNode* root = b.run();
    out.nodes.resize(root->subtreeSize);
    uint32_t next = 0;
    flatten(root, out.nodes, next);
    out.order.resize(refs.size());
    // leaves own contiguous, depth-first ordered ranges of the partitioned
    // array, so slot i holds refs[i]
    for (size_t i = 0; i < refs.size(); ++i) out.order[i] = refs[i].id;
    return out;
}

namespace {
uint32_t collapseNode(const std::vector<BinaryNode>& bin, std::vector<gnd_bvh4_node>& out, uint32_t root) {
    uint32_t me = (uint32_t)out.size();
    out.emplace_back();
    uint32_t cand[4];
    int nc = 1;
    cand[0] = root;
    while (nc < 4) {
        int best = -1;
        float bestArea = -1.0f;
        for (int i = 0; i < nc; ++i) {
            const BinaryNode& c = bin[cand[i]];
            if (c.count == 0) {
                float a = c.box.area();
                if (a > bestArea) { bestArea = a; best = i; }
            }
        }
        if (best < 0) break;
        uint32_t e = cand[best];
        for (int i = best; i + 1 < nc; ++i) cand[i] = cand[i + 1];
        cand[nc - 1] = e + 1;
        cand[nc] = bin[e].offset;
        ++nc;
    }
    uint8_t type[4];
    uint32_t off[4], cnt[4];
    for (int i = 0; i < nc; ++i) {
        const BinaryNode& c = bin[cand[i]];
        if (c.count > 0) { type[i] = 2; off[i] = c.offset; cnt[i] = c.count; }
        else { type[i] = 1; off[i] = collapseNode(bin, out, cand[i]); cnt[i] = 0; }
    }
    gnd_bvh4_node& n4 = out[me];
    std::memset(&n4, 0, sizeof(n4));
    const float inf = std::numeric_limits<float>::infinity();
    for (int i = 0; i < 4; ++i) {
        if (i < nc) {
            const Box& b = bin[cand[i]].box;
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
    return me;
}
}  // namespace

std::vector<gnd_bvh4_node> collapseWide(const std::vector<BinaryNode>& nodes) {
    std::vector<gnd_bvh4_node> out;
    if (nodes.empty()) return out;
    out.reserve(nodes.size() / 2 + 1);
    collapseNode(nodes, out, 0);
    return out;
}

void buildMesh(const float* P, uint32_t ntris, int threads, MeshBuild& out) {
    std::vector<PrimRef> refs(ntris);
    for (uint32_t i = 0; i < ntris; ++i) {
        Box b;
        for (int k = 0; k < 3; ++k) b.grow(Vec3(P[9 * i + 3 * k], P[9 * i + 3 * k + 1], P[9 * i + 3 * k + 2]));
        refs[i] = {b, b.center(), (int)i};
    }
    BinaryBVH bvh = buildBinary(refs, threads);
    if (!bvh.nodes.empty()) out.bounds = bvh.nodes[0].box;
    out.packs.clear();
    for (auto& node : bvh.nodes) {
        if (node.count == 0) continue;
        const int first = (int)node.offset, count = node.count;
        const uint32_t packBase = (uint32_t)out.packs.size();
        const int np = (count + 3) / 4;
        for (int p = 0; p < np; ++p) {
            gnd_tri_pack4 pk;
            std::memset(&pk, 0, sizeof(pk));
            for (int l = 0; l < 4; ++l) {
                int k = p * 4 + l;
                if (k >= count) { pk.prim[l] = 0xFFFFFFFFu; continue; }
                uint32_t t = (uint32_t)bvh.order[first + k];
                pk.prim[l] = t;
                const float* v = P + 9 * t;
                pk.v0x[l] = v[0]; pk.v0y[l] = v[1]; pk.v0z[l] = v[2];
                pk.v1x[l] = v[3]; pk.v1y[l] = v[4]; pk.v1z[l] = v[5];
                pk.v2x[l] = v[6]; pk.v2y[l] = v[7]; pk.v2z[l] = v[8];
            }
            out.packs.push_back(pk);
        }
        node.offset = packBase;
        node.count = (uint16_t)np;
    }
    out.nodes = collapseWide(bvh.nodes);
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
