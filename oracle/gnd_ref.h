/*
 * gnd_ref.h — CPU restatement of the reference renderer's hot path.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  This library is the oracle that the
 * HIP renderer is checked against.  Only the test-suite, the smoke check and
 * the CPU-baseline leg of the benchmark may load it; the product library
 * never links or calls it.
 *
 * Every function restates the arithmetic of the reference (daverlez/gianduia)
 * in the reference's evaluation order and cites the file:line it follows.
 * The reference itself cannot be compiled here (glm, Highway, TBB, pugixml,
 * OpenVDB are absent), so the glm scalar forms and the Highway/SSE lane
 * semantics are restated from their published definitions:
 *   glm dot(a,b)      = (a.x*b.x + a.y*b.y) + a.z*b.z
 *   glm cross(x,y)    = (x.y*y.z - y.y*x.z, x.z*y.x - y.z*x.x, x.x*y.y - y.x*x.y)
 *   glm normalize(v)  = v * (1 / sqrt(dot(v,v)))
 *   glm mat4 * vec4   = (c0*x + c1*y) + (c2*z + c3*w)
 *   glm mat3 * vec3   = (c0*x + c1*y) + c2*z
 *   hn::Min / hn::Max = SSE minps / maxps   ((a<b)?a:b / (a>b)?a:b)
 * The BVH4 slab test and the 4-lane Moller-Trumbore use SSE intrinsics so the
 * lane semantics are the reference's own (x86-64 Highway target is SSE2+).
 *
 * Build flags: -O3 -ffp-contract=off -fno-fast-math (no FMA contraction).
 *
 * How the restatement is pinned (tests/test_oracle_known_answers.py,
 * tests/test_oracle_pins.py, tests/test_oracle_media_pins.py,
 * tests/test_golden_rays.py): the reference's own test values (Fresnel, GGX,
 * BSDF lobes, camera, bounds, frame, colour, sphere), published PCG32 and
 * SplitMix64 outputs, brute-force traversal (also of the committed ray-batch
 * fixtures), warp chi^2 / KS tests, BSDF energy and pdf integrals, HG
 * normalisation, homogeneous-media closed forms, the grid DDA over
 * multi-voxel macro cells (ratio tracking's mean and delta tracking's
 * escape probability and collision CDF against the closed-form optical
 * depth), and furnaces for the path, direct and volumetric path integrators
 * (the latter with a homogeneous and a grid medium).  Parity unpinned (no reference test or closed form
 * fixes them): the glm scalar evaluation orders above, glm::decompose /
 * slerp / mat4_cast / inverse of moving transforms (restated in
 * gnd_ref_motion.cpp, pinned only by closed forms), the Highway lane
 * semantics, OpenVDB BoxSampler / worldToIndex on the dense grid, and the
 * exact positions of area-light and polygonal-aperture samples (their
 * distributions, the thin-lens effects and the Gaussian filter's weights are
 * pinned by closed forms).
 */
#pragma once

#include <xmmintrin.h>
#include <emmintrin.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#include <algorithm>
#include <atomic>

#include "../include/gnd_scene.h"

namespace gref {

constexpr float kInf = std::numeric_limits<float>::infinity();
constexpr float kEpsilon = 0.0001f;                 // constants.h:13
constexpr float kPi = 3.14159265358979323846f;      // constants.h:15
constexpr float kInvPi = 0.31830988618379067154f;   // constants.h:17
constexpr float kInv2Pi = 0.15915494309189533577f;  // constants.h:18
constexpr float kInv4Pi = 0.07957747154594766788f;  // constants.h:19

// ---------------------------------------------------------------- vectors
struct V3 {
    float x = 0, y = 0, z = 0;
    V3() = default;
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
// glm vec3 / scalar divides every component (geometry.h:69)
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float Dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 Cross(V3 x, V3 y) {
    return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
inline float Length(V3 v) { return std::sqrt(Dot(v, v)); }
inline float LengthSq(V3 v) { return Dot(v, v); }
inline V3 Normalize(V3 v) { return v * (1.0f / std::sqrt(Dot(v, v))); }

struct P2 {
    float x = 0, y = 0;
};

// Color3f (include/gianduia/math/color.h:11-115); '/' multiplies by 1/s (:41-42)
struct C3 {
    float r = 0, g = 0, b = 0;
    C3() = default;
    explicit C3(float v) : r(v), g(v), b(v) {}
    C3(float a, float b_, float c) : r(a), g(b_), b(c) {}
    float operator[](int i) const { return i == 0 ? r : (i == 1 ? g : b); }
    float& operator[](int i) { return i == 0 ? r : (i == 1 ? g : b); }
    bool isBlack() const { return r <= 0.0f && g <= 0.0f && b <= 0.0f; }
    bool hasNaNs() const {
        return std::isnan(r) || std::isnan(g) || std::isnan(b) || std::isinf(r) ||
               std::isinf(g) || std::isinf(b);
    }
    float luminance() const { return r * 0.212671f + g * 0.715160f + b * 0.072169f; }
};
inline C3 operator+(C3 a, C3 b) { return {a.r + b.r, a.g + b.g, a.b + b.b}; }
inline C3 operator-(C3 a, C3 b) { return {a.r - b.r, a.g - b.g, a.b - b.b}; }
inline C3 operator*(C3 a, C3 b) { return {a.r * b.r, a.g * b.g, a.b * b.b}; }
inline C3 operator*(C3 a, float s) { return {a.r * s, a.g * s, a.b * s}; }
inline C3 operator/(C3 a, float s) {
    float inv = 1.0f / s;
    return {a.r * inv, a.g * inv, a.b * inv};
}

// ---------------------------------------------------------------- transforms
struct M4 {
    float m[16];  // column-major, m[c*4 + r]
};
// Transform::operator()(Point3f) (transform.h:43-50)
inline V3 xfPoint(const M4& M, V3 p) {
    const float* m = M.m;
    float o[4];
    for (int r = 0; r < 4; ++r)
        o[r] = (m[0 * 4 + r] * p.x + m[1 * 4 + r] * p.y) + (m[2 * 4 + r] * p.z + m[3 * 4 + r] * 1.0f);
    if (o[3] == 1.0f) return {o[0], o[1], o[2]};
    return V3(o[0], o[1], o[2]) / o[3];
}
// Transform::operator()(Vector3f) (transform.h:52-56)
inline V3 xfVector(const M4& M, V3 v) {
    const float* m = M.m;
    float o[3];
    for (int r = 0; r < 3; ++r)
        o[r] = (m[0 * 4 + r] * v.x + m[1 * 4 + r] * v.y) + (m[2 * 4 + r] * v.z + m[3 * 4 + r] * 0.0f);
    return {o[0], o[1], o[2]};
}
// Transform::operator()(Normal3f): transpose(mat3(mInv)) * n (transform.h:58-61)
inline V3 xfNormal(const M4& Minv, V3 n) {
    const float* m = Minv.m;
    float o[3];
    for (int r = 0; r < 3; ++r)
        o[r] = (m[r * 4 + 0] * n.x + m[r * 4 + 1] * n.y) + m[r * 4 + 2] * n.z;
    return {o[0], o[1], o[2]};
}

// ---------------------------------------------------------------- bounds
// Bounds3f (include/gianduia/math/bounds.h:11-106)
struct Bounds {
    V3 pMin{std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
            std::numeric_limits<float>::max()};
    V3 pMax{std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest(),
            std::numeric_limits<float>::lowest()};
    float surfaceArea() const {
        V3 d = pMax - pMin;
        return 2.f * (d.x * d.y + d.x * d.z + d.y * d.z);
    }
    int maximumExtent() const {
        V3 d = pMax - pMin;
        if (d.x > d.y && d.x > d.z) return 0;
        if (d.y > d.z) return 1;
        return 2;
    }
    V3 centroid() const { return pMin * 0.5f + pMax * 0.5f; }
    V3 corner(int c) const {
        return {(c & 1) ? pMax.x : pMin.x, (c & 2) ? pMax.y : pMin.y, (c & 4) ? pMax.z : pMin.z};
    }
};
inline float smin(float a, float b) { return (b < a) ? b : a; }  // std::min
inline float smax(float a, float b) { return (a < b) ? b : a; }  // std::max
inline Bounds Union(const Bounds& a, const Bounds& b) {
    Bounds r;
    r.pMin = {smin(a.pMin.x, b.pMin.x), smin(a.pMin.y, b.pMin.y), smin(a.pMin.z, b.pMin.z)};
    r.pMax = {smax(a.pMax.x, b.pMax.x), smax(a.pMax.y, b.pMax.y), smax(a.pMax.z, b.pMax.z)};
    return r;
}
inline Bounds Union(const Bounds& a, V3 p) {
    Bounds r;
    r.pMin = {smin(a.pMin.x, p.x), smin(a.pMin.y, p.y), smin(a.pMin.z, p.z)};
    r.pMax = {smax(a.pMax.x, p.x), smax(a.pMax.y, p.y), smax(a.pMax.z, p.z)};
    return r;
}
// Bounds3f::rayIntersect (bounds.h:108-128)
bool boundsRayIntersect(const Bounds& b, V3 o, V3 d, float tMax, float* t0o, float* t1o);

// ---------------------------------------------------------------- BVH
// BVHNode / BVHNode4 (include/gianduia/core/bvhBuilder.h:8-28)
struct BVHNode {
    Bounds bounds;
    uint32_t offset;   // leaf: primitivesOffset; inner: rightChildOffset
    uint16_t nPrimitives;
    uint8_t axis;
};
struct alignas(16) BVHNode4 {
    float minX[4], minY[4], minZ[4];
    float maxX[4], maxY[4], maxZ[4];
    uint32_t offset[4];
    uint32_t packCount[4];
    uint8_t childType[4];  // 0 empty, 1 inner, 2 leaf
};
struct BuildInfo {
    Bounds bounds;
    V3 centroid;
    int originalIndex;
};
struct BuildResult {
    std::vector<BVHNode> nodes;
    std::vector<int> orderedIndices;
};
BuildResult bvhBuild(std::vector<BuildInfo>& data);                  // bvhBuilder.cpp:59-182
std::vector<BVHNode4> bvhBuildWide(const std::vector<BVHNode>& bin);  // bvhBuilder.cpp:184-262

// Mesh::TrianglePack4 (include/gianduia/shapes/mesh.h:37-43)
struct alignas(16) TrianglePack4 {
    float v0x[4], v0y[4], v0z[4];
    float v1x[4], v1y[4], v1z[4];
    float v2x[4], v2y[4], v2z[4];
    uint32_t primIndex[4];
};

struct Counters {
    uint64_t nodes = 0, packs = 0, insts = 0, rays = 0;
};

struct Ray {
    V3 o, d;
    float tMin = kEpsilon;
    float tMax = kInf;
    int medium = -1;
    float time = 0.0f;  // ray.h:19, read by animated transforms only
};

struct Mesh {
    std::vector<V3> positions;  // de-indexed corners
    std::vector<V3> normals;    // per corner (always present after load, like loadOBJ)
    std::vector<float> uvs;     // 2 per corner, empty when absent
    std::vector<TrianglePack4> packs;
    std::vector<BVHNode4> nodes4;
    Bounds bounds;
    std::vector<float> cdf;
    float totalArea = 0.0f;
    uint32_t ntris() const { return (uint32_t)(positions.size() / 3); }
};
void meshActivate(Mesh& mesh);  // mesh.cpp:17-146

struct Hit {
    float t = 0.0f;
    float u = 0.0f, v = 0.0f;
    uint32_t primIndex = 0;
    int prim = -1;  // primitive index (scene order)
};

// ---------------------------------------------------------------- sampler
// PCG32 (include/gianduia/core/pcg32.h:5-32), IndependentSampler (src/samplers/independent.cpp)
struct PCG32 {
    uint64_t state = 0, inc = 1;
    void seed(uint64_t initstate, uint64_t initseq = 1) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        nextUInt();
        state += initstate;
        nextUInt();
    }
    uint32_t nextUInt() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xorshifted = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = uint32_t(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((~rot + 1u) & 31));
    }
    float nextFloat() { return std::ldexp((float)nextUInt(), -32); }
};
inline uint64_t splitMix64(uint64_t index) {  // independent.cpp:38-43
    uint64_t z = (index + 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
struct Sampler {
    PCG32 rng;
    uint64_t draws = 0;
    void seed(uint64_t idx) { rng.seed(splitMix64(idx)); }
    float next1D() { ++draws; return rng.nextFloat(); }
    // Point2f(nextFloat(), nextFloat()): left-to-right evaluation (Clang)
    P2 next2D() { P2 p; p.x = next1D(); p.y = next1D(); return p; }
};

// ---------------------------------------------------------------- scene
struct Material {
    gnd_material_desc d;
};
struct Medium {
    gnd_medium_desc d;
    C3 sigma_a, sigma_s, sigma_t;
    // grid data
    std::vector<float> density;
    V3 gridMin, gridMax;
    int resX = 1, resY = 1, resZ = 1;
    float cellX = 1, cellY = 1, cellZ = 1;
    std::vector<float> majorants;
    M4 toLocal;
};
// AnimatedTransform (src/math/animated.cpp, include/gianduia/math/animated.h):
// keyframes at times 0 and 1, decomposed into translation, rotation
// quaternion (x, y, z, w) and scale when the matrices differ
struct Motion {
    bool animated = false;
    M4 start, end;
    float T[2][3], R[2][4], S[2][3];
};
Motion motionFromMatrices(const float* start, const float* end);
M4 motionInterpolate(const Motion& mo, float time);
// glm::inverse (Transform(M) keeps it, transform.h:20)
M4 motionInverse(const M4& m);

struct Prim {
    int shapeType = GND_SHAPE_MESH;
    int mesh = -1;
    float radius = 1.0f;
    int material = 0;
    int emitter = -1;  // index into Scene::emitters
    C3 radiance;
    M4 toWorld, toWorldInv;
    Motion motion;  // Primitive ctor (primitive.cpp:6-11): toWorld -> toWorldEnd
    M4 endInv;      // toWorldEnd's inverse as given
    Bounds worldBounds;
};
// Primitive::getToWorld(time) and its inverse (primitive.cpp:48-51, 82-84)
void primAt(const Prim& p, float time, M4& W, M4& Wi);

struct Camera {
    gnd_camera_desc d;
    M4 toWorld, toWorldInv;
    Motion motion;
    float scale = 1.0f;
    float aspect = 1.0f;
};

// Distribution1D / Distribution2D (src/math/distribution.cpp:6-81)
struct Distribution1D {
    std::vector<float> func, cdf;
    float funcInt = 0.0f;
    void build(const float* f, int n);
    int count() const { return (int)func.size(); }
    float sampleContinuous(float u, float* pdf, int* off) const;
};
struct Distribution2D {
    std::vector<Distribution1D> conditional;  // one per row (v), over the columns (u)
    Distribution1D marginal;
    void build(const float* data, int nu, int nv);
    P2 sampleContinuous(P2 u, float* pdf) const;
    float pdf(P2 p) const;
};
// Point / spot / directional / environment emitters
// (src/emitters/{pointlight,spotlight,directional,environment}.cpp)
struct Light {
    gnd_light_desc d;
    V3 position, direction;  // spot: normalised cone axis; directional: towards the light
    C3 power, irradiance;
    float cosTotalWidth = 0.0f, cosFalloffStart = 0.0f;
    int width = 0, height = 0;  // environment image
    std::vector<C3> rgb;
    float strength = 1.0f;
    Distribution2D dist;
    bool isDelta() const { return d.type != GND_LIGHT_ENVIRONMENT; }
};

struct Scene {
    std::vector<Mesh> meshes;
    std::vector<Prim> prims;            // scene order
    std::vector<int> tlasPrims;         // TLAS leaf slot -> prim index (BVH::build reorder)
    std::vector<BVHNode4> tlas;
    std::vector<int> emitters;          // scene addChild order: prim index (area emitter)
                                        // or -(light index + 1)
    std::vector<Light> lights;
    int envLight = -1;                  // index into lights of the environment, if any
    std::vector<Material> materials;
    std::vector<Medium> media;
    Camera camera;
    int integrator = GND_INTEGRATOR_PATH;
    int maxDepth = 1000;
    uint32_t spp = 1;
    float aoLength = 5.0f;              // ao.cpp:9 ("length"; a desc value <= 0 selects it)
};

void sceneBuild(Scene& s, const gnd_scene_desc& d);
Light lightFromDesc(const gnd_light_desc& d);

bool intersectClosest(const Scene& s, Ray& ray, Hit& hit, Counters* c);
bool intersectAny(const Scene& s, Ray& ray, Counters* c);

// SurfaceInteraction (include/gianduia/shapes/shape.h:26-52), filled
struct Isect {
    float t = 0;
    V3 p;
    P2 uv;
    V3 n;
    V3 dpdu;
    int prim = -1;
    uint32_t primIndex = 0;
    int mediumInside = -1, mediumOutside = -1;
    bool hasInterface = false;
};
void fillInteraction(const Scene& s, const Ray& ray, const Hit& hit, Isect& it);

// ---------------------------------------------------------------- render
struct RenderOptions {
    int threads = 1;
    uint64_t seedOffset = 0;
    int sppBegin = 0, sppEnd = -1;   // pass range (default all)
    int rowBegin = 0, rowEnd = -1;   // row range (default all)
    bool accumulateSquares = false;
    bool aovs = false;               // also splat the per-sample AOV records
};
// AOVRecord (include/gianduia/core/film.h:13-19)
struct AOV {
    C3 albedo{0.0f};
    V3 normal{0.0f, 0.0f, 0.0f};
    float depth = -1.0f, roughness = 0.0f, metallic = 0.0f;
};
constexpr int kAovFloats = 9;  // albedo rgb, normal xyz, depth, roughness, metallic
struct Film {
    int width = 0, height = 0;
    std::vector<float> rgbw;     // 4 floats per pixel: r, g, b, weight
    std::vector<float> sq;       // optional per-pixel sum of squares (r,g,b)
    std::vector<float> aov;      // optional weighted AOV sums, kAovFloats per pixel
};
void render(const Scene& s, const RenderOptions& opt, Film& film, Counters* c);

}  // namespace gref
