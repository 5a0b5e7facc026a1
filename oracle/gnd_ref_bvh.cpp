// gnd_ref_bvh.cpp — oracle: BVH build/collapse, mesh packing, scene set-up
// and BVH4 traversal, restated from the reference.  Test infrastructure only.
#include "gnd_ref.h"

#include <climits>
#include <memory>
#include <stdexcept>

namespace gref {

// Bounds3f::rayIntersect (include/gianduia/math/bounds.h:108-128)
bool boundsRayIntersect(const Bounds& b, V3 o, V3 d, float tMax, float* t0o, float* t1o) {
    float t0 = 0.f, t1 = tMax;
    for (int i = 0; i < 3; ++i) {
        float inv = 1.f / d[i];
        float tNear = (b.pMin[i] - o[i]) * inv;
        float tFar = (b.pMax[i] - o[i]) * inv;
        if (tNear > tFar) std::swap(tNear, tFar);
        t0 = tNear > t0 ? tNear : t0;
        t1 = tFar < t1 ? tFar : t1;
        if (t0 > t1) return false;
    }
    if (t0o) *t0o = t0;
    if (t1o) *t1o = t1;
    return true;
}

// ------------------------------------------------------------ binary SAH build
// BVHBuilder::build (src/core/bvhBuilder.cpp:59-182).  The reference builds
// subtrees above 4096 primitives with tbb::parallel_invoke; subtree contents
// never depend on thread timing (each subtree owns its index range), only the
// leaf offsets do, and the sequential left-first order used here is one of the
// reference's possible outcomes.
namespace {
struct BuildNode {
    Bounds bounds;
    BuildNode* child[2] = {nullptr, nullptr};
    int axis = 0, first = 0, n = 0;
};

// std::partition on bidirectional iterators (the libstdc++/libc++ algorithm)
template <class Pred>
int partitionRange(std::vector<BuildInfo>& v, int first, int last, Pred pred) {
    while (true) {
        while (true) {
            if (first == last) return first;
            if (pred(v[first])) ++first;
            else break;
        }
        --last;
        while (true) {
            if (first == last) return first;
            if (!pred(v[last])) --last;
            else break;
        }
        std::swap(v[first], v[last]);
        ++first;
    }
}

struct Builder {
    std::vector<BuildInfo>& data;
    std::vector<BuildInfo> ordered;
    int orderedOffset = 0;
    int totalNodes = 0;
    std::vector<std::unique_ptr<BuildNode>> pool;

    explicit Builder(std::vector<BuildInfo>& d) : data(d), ordered(d.size()) {}

    BuildNode* build(int start, int end, int depth) {
        pool.emplace_back(new BuildNode());
        BuildNode* node = pool.back().get();
        ++totalNodes;
        Bounds bounds;
        for (int i = start; i < end; ++i) bounds = Union(bounds, data[i].bounds);
        int n = end - start;
        auto leaf = [&]() {
            int off = orderedOffset;
            orderedOffset += n;
            for (int i = 0; i < n; ++i) ordered[off + i] = data[start + i];
            node->first = off;
            node->n = n;
            node->bounds = bounds;
            return node;
        };
        if (n == 1 || depth == 63) return leaf();
        Bounds cb;
        for (int i = start; i < end; ++i) cb = Union(cb, data[i].centroid);
        int dim = cb.maximumExtent();
        if (cb.pMax[dim] == cb.pMin[dim]) return leaf();

        constexpr int nB = 12;
        int count[nB] = {0};
        Bounds bb[nB];
        float denom = cb.pMax[dim] - cb.pMin[dim];
        if (denom == 0) denom = 1.0f;
        auto bucketOf = [&](const BuildInfo& bi) {
            int b = (int)((float)nB * (bi.centroid - cb.pMin)[dim] / denom);
            if (b == nB) b = nB - 1;
            return b;
        };
        for (int i = start; i < end; ++i) {
            int b = bucketOf(data[i]);
            count[b]++;
            bb[b] = Union(bb[b], data[i].bounds);
        }
        float cost[nB - 1];
        float sa = bounds.surfaceArea();
        for (int i = 0; i < nB - 1; ++i) {
            Bounds b0, b1;
            int c0 = 0, c1 = 0;
            for (int j = 0; j <= i; ++j) { b0 = Union(b0, bb[j]); c0 += count[j]; }
            for (int j = i + 1; j < nB; ++j) { b1 = Union(b1, bb[j]); c1 += count[j]; }
            cost[i] = 1 + ((float)c0 * b0.surfaceArea() + (float)c1 * b1.surfaceArea()) / sa;
        }
        float minCost = cost[0];
        int minB = 0;
        for (int i = 1; i < nB - 1; ++i)
            if (cost[i] < minCost) { minCost = cost[i]; minB = i; }
        float leafCost = (float)n;
        if (n > 4 || minCost < leafCost) {
            int mid = partitionRange(data, start, end, [&](const BuildInfo& bi) { return bucketOf(bi) <= minB; });
            BuildNode* c0 = build(start, mid, depth + 1);
            BuildNode* c1 = build(mid, end, depth + 1);
            node->child[0] = c0;
            node->child[1] = c1;
            node->bounds = Union(c0->bounds, c1->bounds);
            node->axis = dim;
            node->n = 0;
            return node;
        }
        return leaf();
    }

    // flattenBVH (bvhBuilder.cpp:29-44): depth-first, left child = parent + 1
    int flatten(BuildNode* node, int* offset, std::vector<BVHNode>& out) {
        int cur = (*offset)++;
        out[cur].bounds = node->bounds;
        if (node->n > 0) {
            out[cur].offset = (uint32_t)node->first;
            out[cur].nPrimitives = (uint16_t)node->n;
        } else {
            out[cur].axis = (uint8_t)node->axis;
            out[cur].nPrimitives = 0;
            flatten(node->child[0], offset, out);
            out[cur].offset = (uint32_t)flatten(node->child[1], offset, out);
        }
        return cur;
    }
};
}  // namespace

BuildResult bvhBuild(std::vector<BuildInfo>& data) {
    BuildResult res;
    if (data.empty()) return res;
    Builder b(data);
    BuildNode* root = b.build(0, (int)data.size(), 0);
    res.nodes.resize(b.totalNodes);
    std::memset((void*)res.nodes.data(), 0, sizeof(BVHNode) * res.nodes.size());
    int off = 0;
    b.flatten(root, &off, res.nodes);
    res.orderedIndices.reserve(b.ordered.size());
    for (auto& bi : b.ordered) res.orderedIndices.push_back(bi.originalIndex);
    return res;
}

// BVHBuilder::buildWide (bvhBuilder.cpp:184-262): greedy expansion of the
// largest-area inner candidate until four children; children emitted in
// candidate order, subtrees allocated depth-first.
namespace {
int collapse(const std::vector<BVHNode>& bin, std::vector<BVHNode4>& out, int binIdx) {
    int idx = (int)out.size();
    out.push_back(BVHNode4());
    std::vector<int> cand = {binIdx};
    while (cand.size() < 4) {
        int best = -1;
        float bestArea = -1.0f;
        for (size_t i = 0; i < cand.size(); ++i) {
            const BVHNode& c = bin[cand[i]];
            if (c.nPrimitives == 0) {
                float a = c.bounds.surfaceArea();
                if (a > bestArea) { bestArea = a; best = (int)i; }
            }
        }
        if (best == -1) break;
        int e = cand[best];
        cand.erase(cand.begin() + best);
        cand.push_back(e + 1);
        cand.push_back((int)bin[e].offset);
    }
    struct Child { uint8_t type; uint32_t off, cnt; Bounds b; };
    std::vector<Child> ch;
    for (int ci : cand) {
        const BVHNode& c = bin[ci];
        if (c.nPrimitives > 0) ch.push_back({2, c.offset, (uint32_t)c.nPrimitives, c.bounds});
        else {
            int sub = collapse(bin, out, ci);
            ch.push_back({1, (uint32_t)sub, 0, c.bounds});
        }
    }
    BVHNode4& n4 = out[idx];
    for (int i = 0; i < 4; ++i) {
        if (i < (int)ch.size()) {
            n4.childType[i] = ch[i].type;
            n4.offset[i] = ch[i].off;
            n4.packCount[i] = ch[i].cnt;
            n4.minX[i] = ch[i].b.pMin.x; n4.minY[i] = ch[i].b.pMin.y; n4.minZ[i] = ch[i].b.pMin.z;
            n4.maxX[i] = ch[i].b.pMax.x; n4.maxY[i] = ch[i].b.pMax.y; n4.maxZ[i] = ch[i].b.pMax.z;
        } else {
            n4.childType[i] = 0;
            n4.offset[i] = 0;
            n4.packCount[i] = 0;
            n4.minX[i] = n4.minY[i] = n4.minZ[i] = kInf;
            n4.maxX[i] = n4.maxY[i] = n4.maxZ[i] = -kInf;
        }
    }
    return idx;
}
}  // namespace

std::vector<BVHNode4> bvhBuildWide(const std::vector<BVHNode>& bin) {
    std::vector<BVHNode4> out;
    if (bin.empty()) return out;
    out.reserve(bin.size() / 2 + 1);
    collapse(bin, out, 0);
    return out;
}

// ------------------------------------------------------------ mesh
// Mesh::activate (src/shapes/mesh.cpp:17-146)
void meshActivate(Mesh& m) {
    uint32_t nt = m.ntris();
    std::vector<BuildInfo> data;
    data.reserve(nt);
    for (uint32_t i = 0; i < nt; ++i) {
        Bounds b;
        b = Union(b, m.positions[3 * i + 0]);
        b = Union(b, m.positions[3 * i + 1]);
        b = Union(b, m.positions[3 * i + 2]);
        data.push_back({b, b.centroid(), (int)i});
    }
    BuildResult r = bvhBuild(data);
    if (!r.nodes.empty()) m.bounds = r.nodes[0].bounds;
    m.packs.clear();
    for (auto& node : r.nodes) {
        if (node.nPrimitives > 0) {
            int origOffset = (int)node.offset;
            int origCount = node.nPrimitives;
            int packOffset = (int)m.packs.size();
            int numPacks = (origCount + 3) / 4;
            for (int p = 0; p < numPacks; ++p) {
                TrianglePack4 pk;
                for (int l = 0; l < 4; ++l) {
                    pk.v0x[l] = pk.v0y[l] = pk.v0z[l] = 0.0f;
                    pk.v1x[l] = pk.v1y[l] = pk.v1z[l] = 0.0f;
                    pk.v2x[l] = pk.v2y[l] = pk.v2z[l] = 0.0f;
                    pk.primIndex[l] = 0xFFFFFFFFu;
                }
                for (int l = 0; l < 4; ++l) {
                    int tri = p * 4 + l;
                    if (tri < origCount) {
                        int orig = r.orderedIndices[origOffset + tri];
                        pk.primIndex[l] = (uint32_t)orig;
                        V3 p0 = m.positions[3 * orig], p1 = m.positions[3 * orig + 1], p2 = m.positions[3 * orig + 2];
                        pk.v0x[l] = p0.x; pk.v0y[l] = p0.y; pk.v0z[l] = p0.z;
                        pk.v1x[l] = p1.x; pk.v1y[l] = p1.y; pk.v1z[l] = p1.z;
                        pk.v2x[l] = p2.x; pk.v2y[l] = p2.y; pk.v2z[l] = p2.z;
                    }
                }
                m.packs.push_back(pk);
            }
            node.offset = (uint32_t)packOffset;
            node.nPrimitives = (uint16_t)numPacks;
        }
    }
    if (!r.nodes.empty()) m.nodes4 = bvhBuildWide(r.nodes);
    m.cdf.assign(nt + 1, 0.0f);
    m.totalArea = 0.0f;
    for (uint32_t i = 0; i < nt; ++i) {
        V3 e1 = m.positions[3 * i + 1] - m.positions[3 * i];
        V3 e2 = m.positions[3 * i + 2] - m.positions[3 * i];
        float area = 0.5f * Length(Cross(e1, e2));
        m.totalArea += area;
        m.cdf[i + 1] = m.totalArea;
    }
    if (m.totalArea > 0.0f)
        for (uint32_t i = 1; i <= nt; ++i) m.cdf[i] /= m.totalArea;
}

// Mesh::loadOBJ normal handling (mesh.cpp:163-228): file normals are
// normalized; missing normals are generated per shared vertex as the sum of
// unnormalized face normals, then divided by their length.
static void meshFromDesc(Mesh& m, const gnd_mesh_desc& d) {
    uint32_t n = d.ntris * 3;
    m.positions.resize(n);
    for (uint32_t i = 0; i < n; ++i)
        m.positions[i] = V3(d.positions[3 * i], d.positions[3 * i + 1], d.positions[3 * i + 2]);
    m.normals.resize(n);
    if (d.normals) {
        for (uint32_t i = 0; i < n; ++i)
            m.normals[i] = Normalize(V3(d.normals[3 * i], d.normals[3 * i + 1], d.normals[3 * i + 2]));
    } else {
        uint32_t nv = 0;
        for (uint32_t i = 0; i < n; ++i) nv = std::max(nv, (d.vertex_ids ? d.vertex_ids[i] : i) + 1);
        std::vector<V3> gen(nv, V3(0, 0, 0));
        for (uint32_t f = 0; f < d.ntris; ++f) {
            uint32_t i0 = d.vertex_ids ? d.vertex_ids[3 * f] : 3 * f;
            uint32_t i1 = d.vertex_ids ? d.vertex_ids[3 * f + 1] : 3 * f + 1;
            uint32_t i2 = d.vertex_ids ? d.vertex_ids[3 * f + 2] : 3 * f + 2;
            V3 p0 = m.positions[3 * f], p1 = m.positions[3 * f + 1], p2 = m.positions[3 * f + 2];
            V3 w = Cross(p1 - p0, p2 - p0);
            gen[i0] = gen[i0] + w;
            gen[i1] = gen[i1] + w;
            gen[i2] = gen[i2] + w;
        }
        for (auto& g : gen) {
            float len = Length(g);
            if (len > 0.0f) g = g / len;
        }
        for (uint32_t i = 0; i < n; ++i) m.normals[i] = gen[d.vertex_ids ? d.vertex_ids[i] : i];
    }
    if (d.uvs) m.uvs.assign(d.uvs, d.uvs + 2 * n);
}

static M4 toM4(const float* f) {
    M4 m;
    std::memcpy(m.m, f, sizeof(m.m));
    return m;
}

// Medium set-up: HomogeneousMedium ctor (homogeneous.cpp:11-17) and the
// HeterogeneousMedium macro-majorant grid (heterogeneous.cpp:20-104) over a
// dense grid whose active box is [0,n) in index space.
static void mediumFromDesc(Medium& md, const gnd_medium_desc& d) {
    md.d = d;
    md.sigma_a = C3(d.sigma_a[0], d.sigma_a[1], d.sigma_a[2]);
    md.sigma_s = C3(d.sigma_s[0], d.sigma_s[1], d.sigma_s[2]);
    md.sigma_t = md.sigma_a + md.sigma_s;
    md.toLocal = toM4(d.to_local);
    if (d.type != GND_MEDIUM_GRID) return;
    size_t nvox = (size_t)d.nx * d.ny * d.nz;
    md.density.assign(d.density, d.density + nvox);
    md.gridMin = V3(0, 0, 0);
    md.gridMax = V3((float)(d.nx - 1) + 1.0f, (float)(d.ny - 1) + 1.0f, (float)(d.nz - 1) + 1.0f);
    float ex = md.gridMax.x - md.gridMin.x, ey = md.gridMax.y - md.gridMin.y, ez = md.gridMax.z - md.gridMin.z;
    int target = d.macro_res > 0 ? d.macro_res : 32;
    float maxE = std::max({ex, ey, ez});
    md.resX = std::max(1, (int)std::round(target * (ex / maxE)));
    md.resY = std::max(1, (int)std::round(target * (ey / maxE)));
    md.resZ = std::max(1, (int)std::round(target * (ez / maxE)));
    md.cellX = ex / md.resX;
    md.cellY = ey / md.resY;
    md.cellZ = ez / md.resZ;
    md.majorants.assign((size_t)md.resX * md.resY * md.resZ, 0.0f);
    float maxT = std::max({md.sigma_t.r, md.sigma_t.g, md.sigma_t.b});
    for (int z = 0; z < d.nz; ++z)
        for (int y = 0; y < d.ny; ++y)
            for (int x = 0; x < d.nx; ++x) {
                float dens = md.density[((size_t)z * d.ny + y) * d.nx + x] * d.density_scale;
                if (dens <= 0.0f || !std::isfinite(dens)) continue;
                int cx = std::clamp((int)((x - md.gridMin.x) / md.cellX), 0, md.resX - 1);
                int cy = std::clamp((int)((y - md.gridMin.y) / md.cellY), 0, md.resY - 1);
                int cz = std::clamp((int)((z - md.gridMin.z) / md.cellZ), 0, md.resZ - 1);
                size_t fi = (size_t)cz * (md.resX * md.resY) + cy * md.resX + cx;
                float lm = dens * maxT;
                if (lm > md.majorants[fi]) md.majorants[fi] = lm;
            }
}

// Scene::activate + BVH::build (src/scene/scene.cpp:46-62, src/scene/bvh.cpp:10-36)
// Distribution1D (distribution.cpp:6-18)
void Distribution1D::build(const float* f, int n) {
    func.assign(f, f + n);
    cdf.assign(n + 1, 0.0f);
    cdf[0] = 0.0f;
    for (int i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + func[i - 1] / n;
    funcInt = cdf[n];
    if (funcInt == 0.0f) {
        for (int i = 1; i < n + 1; ++i) cdf[i] = float(i) / float(n);
    } else {
        for (int i = 1; i < n + 1; ++i) cdf[i] /= funcInt;
    }
}
// Distribution1D::SampleContinuous (distribution.cpp:20-34)
float Distribution1D::sampleContinuous(float u, float* pdf, int* off) const {
    auto it = std::upper_bound(cdf.begin(), cdf.end(), u);
    int offset = std::clamp<int>((int)std::distance(cdf.begin(), it) - 1, 0, (int)cdf.size() - 2);
    if (off) *off = offset;
    float du = u - cdf[offset];
    if ((cdf[offset + 1] - cdf[offset]) > 0.0f) du /= (cdf[offset + 1] - cdf[offset]);
    if (pdf) *pdf = (funcInt > 0.0f) ? func[offset] / funcInt : 0.0f;
    return (offset + du) / count();
}
// Distribution2D (distribution.cpp:53-80)
void Distribution2D::build(const float* data, int nu, int nv) {
    conditional.assign(nv, Distribution1D());
    std::vector<float> marginalFunc(nv);
    for (int v = 0; v < nv; ++v) {
        conditional[v].build(&data[(size_t)v * nu], nu);
        marginalFunc[v] = conditional[v].funcInt;
    }
    marginal.build(marginalFunc.data(), nv);
}
P2 Distribution2D::sampleContinuous(P2 u, float* pdf) const {
    float pdfs[2];
    int v;
    float d1 = marginal.sampleContinuous(u.y, &pdfs[1], &v);
    float d0 = conditional[v].sampleContinuous(u.x, &pdfs[0], nullptr);
    *pdf = pdfs[0] * pdfs[1];
    return {d0, d1};
}
float Distribution2D::pdf(P2 p) const {
    int iu = std::clamp<int>((int)(p.x * conditional[0].count()), 0, conditional[0].count() - 1);
    int iv = std::clamp<int>((int)(p.y * marginal.count()), 0, marginal.count() - 1);
    return conditional[iv].func[iu] / marginal.funcInt;
}

// Light constructors (pointlight.cpp:9-12, spotlight.cpp:9-22,
// directional.cpp:8-11, environment.cpp:14-27 with buildPDFs :106-120)
Light lightFromDesc(const gnd_light_desc& d) {
    Light L;
    L.d = d;
    L.d.env_rgb = nullptr;
    L.position = V3(d.position[0], d.position[1], d.position[2]);
    L.power = C3(d.power[0], d.power[1], d.power[2]);
    L.irradiance = C3(d.irradiance[0], d.irradiance[1], d.irradiance[2]);
    const V3 dir(d.direction[0], d.direction[1], d.direction[2]);
    if (d.type == GND_LIGHT_SPOT) {
        L.direction = Normalize(dir);
        float halfAngle = d.spot_size * 0.5f;
        float innerAngle = halfAngle * (1.0f - d.blend);
        L.cosTotalWidth = std::cos((kPi / 180.f) * halfAngle);
        L.cosFalloffStart = std::cos((kPi / 180.f) * innerAngle);
    } else if (d.type == GND_LIGHT_DIRECTIONAL) {
        L.direction = -Normalize(dir);
    } else if (d.type == GND_LIGHT_ENVIRONMENT) {
        if (!d.env_rgb || d.env_width <= 0 || d.env_height <= 0)
            throw std::runtime_error("environment light without an image");
        L.width = d.env_width;
        L.height = d.env_height;
        L.strength = d.strength;
        L.rgb.resize((size_t)L.width * L.height);
        for (size_t i = 0; i < L.rgb.size(); ++i)
            L.rgb[i] = C3(d.env_rgb[3 * i], d.env_rgb[3 * i + 1], d.env_rgb[3 * i + 2]);
        const int rows = L.height, cols = L.width;
        std::vector<float> img((size_t)rows * cols);
        for (int i = 0; i < rows; i++) {
            float sinTheta = std::sin(M_PI * (i + 0.5f) / rows);  // double sine, stored as float
            for (int j = 0; j < cols; j++) img[(size_t)i * cols + j] = L.rgb[(size_t)i * cols + j].luminance() * sinTheta;
        }
        L.dist.build(img.data(), cols, rows);
    }
    return L;
}

void sceneBuild(Scene& s, const gnd_scene_desc& d) {
    s.meshes.resize(d.nmeshes);
    for (uint32_t i = 0; i < d.nmeshes; ++i) {
        meshFromDesc(s.meshes[i], d.meshes[i]);
        meshActivate(s.meshes[i]);
    }
    s.materials.resize(d.nmaterials);
    for (uint32_t i = 0; i < d.nmaterials; ++i) s.materials[i].d = d.materials[i];
    s.media.resize(d.nmedia);
    for (uint32_t i = 0; i < d.nmedia; ++i) mediumFromDesc(s.media[i], d.media[i]);
    s.prims.resize(d.nprims);
    s.emitters.clear();
    for (uint32_t i = 0; i < d.nprims; ++i) {
        const gnd_primitive_desc& pd = d.prims[i];
        Prim& p = s.prims[i];
        p.shapeType = pd.shape_type;
        p.mesh = pd.mesh;
        p.radius = pd.radius;
        p.material = pd.material;
        p.toWorld = toM4(pd.to_world);
        p.toWorldInv = toM4(pd.to_world_inv);
        p.motion = motionFromMatrices(pd.to_world, pd.has_motion ? pd.to_world_end : pd.to_world);
        p.endInv = toM4(pd.has_motion ? pd.to_world_end_inv : pd.to_world_inv);
        p.radiance = C3(pd.radiance[0], pd.radiance[1], pd.radiance[2]);
        if (pd.has_emitter) p.emitter = 0;  // slot assigned below
        Bounds sb;
        if (p.shapeType == GND_SHAPE_MESH) sb = s.meshes[p.mesh].bounds;
        else { sb.pMin = V3(-p.radius, -p.radius, -p.radius); sb.pMax = V3(p.radius, p.radius, p.radius); }
        // Transform::operator()(Bounds3f) (transform.h:69-76)
        Bounds wb;
        wb.pMin = wb.pMax = xfPoint(p.toWorld, sb.corner(0));
        for (int c = 1; c < 8; ++c) wb = Union(wb, xfPoint(p.toWorld, sb.corner(c)));
        if (p.motion.animated) {
            // AnimatedTransform::motionBounds (animated.cpp:49-60): also the
            // box at time 1 and at the interpolated time 0.5
            for (const M4& E : {p.motion.end, motionInterpolate(p.motion, 0.5f)})
                for (int c = 0; c < 8; ++c) wb = Union(wb, xfPoint(E, sb.corner(c)));
        }
        p.worldBounds = wb;
    }
    // emitter list in document order: lights at their emitter_index, area
    // emitters of the primitives in the other slots in primitive order
    // (Scene::addChild, scene.cpp:9-40)
    s.lights.clear();
    s.envLight = -1;
    int nArea = 0;
    for (auto& p : s.prims) nArea += p.emitter >= 0 ? 1 : 0;
    const int nEm = nArea + (int)d.nlights;
    s.emitters.assign(nEm, INT32_MIN);
    for (uint32_t k = 0; k < d.nlights; ++k) {
        Light L = lightFromDesc(d.lights[k]);
        const int e = d.lights[k].emitter_index;
        if (e < 0 || e >= nEm || s.emitters[e] != INT32_MIN) throw std::runtime_error("bad light emitter_index");
        if (L.d.type == GND_LIGHT_ENVIRONMENT) {
            if (s.envLight >= 0) throw std::runtime_error("Scene: cannot add multiple infinite area lights!");
            s.envLight = (int)k;
        }
        s.emitters[e] = -(int)k - 1;
        s.lights.push_back(std::move(L));
    }
    {
        int slot = 0;
        for (uint32_t i = 0; i < d.nprims; ++i) {
            if (s.prims[i].emitter < 0) continue;
            while (s.emitters[slot] != INT32_MIN) ++slot;
            s.prims[i].emitter = slot;
            s.emitters[slot] = (int)i;
        }
    }
    std::vector<BuildInfo> data;
    for (uint32_t i = 0; i < d.nprims; ++i) {
        const Bounds& b = s.prims[i].worldBounds;
        data.push_back({b, b.pMin * 0.5f + b.pMax * 0.5f, (int)i});
    }
    BuildResult r = bvhBuild(data);
    s.tlasPrims = r.orderedIndices;
    s.tlas = r.nodes.empty() ? std::vector<BVHNode4>() : bvhBuildWide(r.nodes);

    s.camera.d = d.camera;
    s.camera.toWorld = toM4(d.camera.to_world);
    s.camera.toWorldInv = toM4(d.camera.to_world_inv);
    s.camera.motion = motionFromMatrices(d.camera.to_world, d.camera.has_motion ? d.camera.to_world_end
                                                                                : d.camera.to_world);
    // PerspectiveCamera / ThinLensCamera ctor: m_scale = tan(Radians(fov/2))
    s.camera.scale = std::tan((kPi / 180.f) * (d.camera.fov * 0.5f));
    s.camera.aspect = d.camera.width / (float)d.camera.height;
    s.integrator = d.integrator;
    s.maxDepth = d.max_depth;
    s.aoLength = d.ao_length > 0.0f ? d.ao_length : 5.0f;
    s.spp = d.spp;
}

// ------------------------------------------------------------ traversal
// traverseBVH4 (include/gianduia/core/bvhTraversal.h:19-98)
template <class Leaf>
static bool traverseBVH4(const std::vector<BVHNode4>& nodes, const Ray& ray, float& tMaxRef, bool shadow,
                         Leaf&& leaf, Counters* c) {
    if (nodes.empty()) return false;
    bool hitAny = false;
    int stack[256];
    int sp = 0;
    stack[sp++] = 0;
    const __m128 ox = _mm_set1_ps(ray.o.x), oy = _mm_set1_ps(ray.o.y), oz = _mm_set1_ps(ray.o.z);
    const float ix = 1.0f / ray.d.x, iy = 1.0f / ray.d.y, iz = 1.0f / ray.d.z;
    const __m128 idx = _mm_set1_ps(ix), idy = _mm_set1_ps(iy), idz = _mm_set1_ps(iz);
    const __m128 tMinV = _mm_set1_ps(ray.tMin);
    __m128 tMaxV = _mm_set1_ps(tMaxRef);
    while (sp > 0) {
        int ni = stack[--sp];
        const BVHNode4& n = nodes[ni];
        if (c) c->nodes++;
        __m128 t1x = _mm_mul_ps(_mm_sub_ps(_mm_load_ps(n.minX), ox), idx);
        __m128 t2x = _mm_mul_ps(_mm_sub_ps(_mm_load_ps(n.maxX), ox), idx);
        __m128 tminx = _mm_min_ps(t1x, t2x), tmaxx = _mm_max_ps(t1x, t2x);
        __m128 t1y = _mm_mul_ps(_mm_sub_ps(_mm_load_ps(n.minY), oy), idy);
        __m128 t2y = _mm_mul_ps(_mm_sub_ps(_mm_load_ps(n.maxY), oy), idy);
        __m128 tminy = _mm_min_ps(t1y, t2y), tmaxy = _mm_max_ps(t1y, t2y);
        __m128 t1z = _mm_mul_ps(_mm_sub_ps(_mm_load_ps(n.minZ), oz), idz);
        __m128 t2z = _mm_mul_ps(_mm_sub_ps(_mm_load_ps(n.maxZ), oz), idz);
        __m128 tminz = _mm_min_ps(t1z, t2z), tmaxz = _mm_max_ps(t1z, t2z);
        __m128 tmin = _mm_max_ps(_mm_max_ps(tminx, tminy), _mm_max_ps(tminz, tMinV));
        __m128 tmax = _mm_min_ps(_mm_min_ps(tmaxx, tmaxy), _mm_min_ps(tmaxz, tMaxV));
        int mask = _mm_movemask_ps(_mm_cmple_ps(tmin, tmax));
        if (mask == 0) continue;
        for (int i = 0; i < 4; ++i) {
            if ((mask & (1 << i)) == 0) continue;
            if (n.childType[i] == 1) {
                if (sp >= 256) throw std::runtime_error("oracle traversal stack overflow");
                stack[sp++] = (int)n.offset[i];
            } else if (n.childType[i] == 2) {
                if (leaf((int)n.offset[i], (int)n.packCount[i])) {
                    hitAny = true;
                    if (shadow) return true;
                    tMaxV = _mm_set1_ps(tMaxRef);
                }
            }
        }
    }
    return hitAny;
}

// Mesh::rayIntersect (src/shapes/mesh.cpp:231-326): 4-lane Moller-Trumbore,
// lanes committed in order with a strict t < tMax test.
static bool meshIntersect(const Mesh& m, Ray& ray, Hit& h, bool predicate, Counters* c) {
    const __m128 ox = _mm_set1_ps(ray.o.x), oy = _mm_set1_ps(ray.o.y), oz = _mm_set1_ps(ray.o.z);
    const __m128 dx = _mm_set1_ps(ray.d.x), dy = _mm_set1_ps(ray.d.y), dz = _mm_set1_ps(ray.d.z);
    const __m128 absMask = _mm_castsi128_ps(_mm_set1_epi32(0x7fffffff));
    auto leaf = [&](int offset, int count) -> bool {
        bool hitLocal = false;
        __m128 tMaxV = _mm_set1_ps(ray.tMax);
        for (int p = 0; p < count; ++p) {
            const TrianglePack4& pk = m.packs[offset + p];
            if (c) c->packs++;
            __m128 v0x = _mm_load_ps(pk.v0x), v0y = _mm_load_ps(pk.v0y), v0z = _mm_load_ps(pk.v0z);
            __m128 e1x = _mm_sub_ps(_mm_load_ps(pk.v1x), v0x), e1y = _mm_sub_ps(_mm_load_ps(pk.v1y), v0y),
                   e1z = _mm_sub_ps(_mm_load_ps(pk.v1z), v0z);
            __m128 e2x = _mm_sub_ps(_mm_load_ps(pk.v2x), v0x), e2y = _mm_sub_ps(_mm_load_ps(pk.v2y), v0y),
                   e2z = _mm_sub_ps(_mm_load_ps(pk.v2z), v0z);
            // pvec = d x e2
            __m128 px = _mm_sub_ps(_mm_mul_ps(dy, e2z), _mm_mul_ps(dz, e2y));
            __m128 py = _mm_sub_ps(_mm_mul_ps(dz, e2x), _mm_mul_ps(dx, e2z));
            __m128 pz = _mm_sub_ps(_mm_mul_ps(dx, e2y), _mm_mul_ps(dy, e2x));
            __m128 det = _mm_add_ps(_mm_add_ps(_mm_mul_ps(e1x, px), _mm_mul_ps(e1y, py)), _mm_mul_ps(e1z, pz));
            __m128 valid = _mm_cmpgt_ps(_mm_and_ps(det, absMask), _mm_set1_ps(1e-8f));
            __m128 invDet = _mm_div_ps(_mm_set1_ps(1.0f), det);
            __m128 tx = _mm_sub_ps(ox, v0x), ty = _mm_sub_ps(oy, v0y), tz = _mm_sub_ps(oz, v0z);
            __m128 u = _mm_mul_ps(
                _mm_add_ps(_mm_add_ps(_mm_mul_ps(tx, px), _mm_mul_ps(ty, py)), _mm_mul_ps(tz, pz)), invDet);
            valid = _mm_and_ps(valid, _mm_cmpge_ps(u, _mm_set1_ps(0.0f)));
            valid = _mm_and_ps(valid, _mm_cmple_ps(u, _mm_set1_ps(1.0f)));
            // qvec = tvec x e1
            __m128 qx = _mm_sub_ps(_mm_mul_ps(ty, e1z), _mm_mul_ps(tz, e1y));
            __m128 qy = _mm_sub_ps(_mm_mul_ps(tz, e1x), _mm_mul_ps(tx, e1z));
            __m128 qz = _mm_sub_ps(_mm_mul_ps(tx, e1y), _mm_mul_ps(ty, e1x));
            __m128 v = _mm_mul_ps(
                _mm_add_ps(_mm_add_ps(_mm_mul_ps(dx, qx), _mm_mul_ps(dy, qy)), _mm_mul_ps(dz, qz)), invDet);
            valid = _mm_and_ps(valid, _mm_cmpge_ps(v, _mm_set1_ps(0.0f)));
            valid = _mm_and_ps(valid, _mm_cmple_ps(_mm_add_ps(u, v), _mm_set1_ps(1.0f)));
            __m128 t = _mm_mul_ps(
                _mm_add_ps(_mm_add_ps(_mm_mul_ps(e2x, qx), _mm_mul_ps(e2y, qy)), _mm_mul_ps(e2z, qz)), invDet);
            valid = _mm_and_ps(valid, _mm_cmpgt_ps(t, _mm_set1_ps(ray.tMin)));
            valid = _mm_and_ps(valid, _mm_cmplt_ps(t, tMaxV));
            int mask = _mm_movemask_ps(valid);
            if (mask) {
                alignas(16) float ta[4], ua[4], va[4];
                _mm_store_ps(ta, t);
                _mm_store_ps(ua, u);
                _mm_store_ps(va, v);
                for (int l = 0; l < 4; ++l) {
                    if (mask & (1 << l)) {
                        if (ta[l] < ray.tMax) {
                            if (predicate) return true;
                            ray.tMax = ta[l];
                            tMaxV = _mm_set1_ps(ray.tMax);
                            h.t = ta[l];
                            h.u = ua[l];
                            h.v = va[l];
                            h.primIndex = pk.primIndex[l];
                            hitLocal = true;
                        }
                    }
                }
            }
        }
        return hitLocal;
    };
    return traverseBVH4(m.nodes4, ray, ray.tMax, predicate, leaf, c);
}

// Sphere::rayIntersect (src/shapes/sphere.cpp:15-39) with SolveQuadratic
// (include/gianduia/math/constants.h:46-54)
static bool sphereIntersect(float radius, const Ray& ray, Hit& h, bool predicate) {
    V3 ov = ray.o;
    float A = LengthSq(ray.d);
    float B = 2.0f * Dot(ov, ray.d);
    float C = LengthSq(ov) - (radius * radius);
    float discrim = B * B - 4.0f * A * C;
    if (discrim < 0) return false;
    float root = std::sqrt(discrim);
    float t0 = (-B - root) / (2.0f * A);
    float t1 = (-B + root) / (2.0f * A);
    float tHit;
    if (t0 > ray.tMin && t0 < ray.tMax) tHit = t0;
    else if (t1 > ray.tMin && t1 < ray.tMax) tHit = t1;
    else return false;
    if (predicate) return true;
    h.t = tHit;
    h.u = h.v = 0.0f;
    h.primIndex = 0;
    return true;
}

// Transform::operator()(Ray) (transform.h:63-67): tMin resets to Epsilon
static Ray toObject(const Prim& p, const Ray& rw) {
    M4 W, Wi;
    primAt(p, rw.time, W, Wi);  // interpolate(time).inverse() (primitive.cpp:48-50)
    Ray r;
    r.o = xfPoint(Wi, rw.o);
    r.d = xfVector(Wi, rw.d);
    r.tMax = rw.tMax;
    r.tMin = kEpsilon;
    return r;
}

// Primitive::rayIntersect + BVH::rayIntersect (primitive.cpp:48-59, bvh.cpp:38-66)
static bool sceneIntersect(const Scene& s, Ray& ray, Hit& hit, bool shadow, Counters* c) {
    hit.t = std::numeric_limits<float>::max();
    hit.prim = -1;
    if (c) c->rays++;
    auto leaf = [&](int offset, int count) -> bool {
        bool hitLocal = false;
        for (int p = 0; p < count; ++p) {
            int pi = s.tlasPrims[offset + p];
            const Prim& pr = s.prims[pi];
            if (c) c->insts++;
            Ray lr = toObject(pr, ray);
            Hit its;
            bool ok = pr.shapeType == GND_SHAPE_MESH ? meshIntersect(s.meshes[pr.mesh], lr, its, shadow, c)
                                                     : sphereIntersect(pr.radius, lr, its, shadow);
            if (ok) {
                if (shadow) return true;
                if (its.t < hit.t) {
                    ray.tMax = its.t;
                    hit = its;
                    hit.prim = pi;
                    hitLocal = true;
                }
            }
        }
        return hitLocal;
    };
    return traverseBVH4(s.tlas, ray, ray.tMax, shadow, leaf, c);
}

bool intersectClosest(const Scene& s, Ray& ray, Hit& hit, Counters* c) {
    return sceneIntersect(s, ray, hit, false, c);
}
bool intersectAny(const Scene& s, Ray& ray, Counters* c) {
    Hit h;
    return sceneIntersect(s, ray, h, true, c);
}

}  // namespace gref
