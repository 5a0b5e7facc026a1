// dev_common.h — device-side types, arithmetic and the HBM scene layout.
//
// Arithmetic keeps the reference's evaluation order (glm scalar forms) and the
// library is compiled with -ffp-contract=off, correctly rounded fp32 divide and
// sqrt and fp32 denormals preserved, so the traversal reproduces the CPU bits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/gnd_hip.h"

namespace gdev {

#define GDEV_INLINE __device__ __forceinline__

#include "motion.h"

constexpr float kEps = 0.0001f;                     // constants.h:13
constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kInv2Pi = 0.15915494309189533577f;
constexpr float kInv4Pi = 0.07957747154594766788f;
constexpr float kFltMax = 3.402823466e+38f;
constexpr double kPiD = 3.14159265358979323846;
constexpr double kInvPiD = 0.31830988618379067154;  // M_1_PI
// buckets of the guide tables that narrow the CDF searches of environment
// sampling: entry j = the first CDF index whose value exceeds j / kEnvGuide
constexpr int kEnvGuide = 64;

GDEV_INLINE float finf() { return __builtin_inff(); }

struct V3 {
    float x, y, z;
};
GDEV_INLINE V3 mk(float x, float y, float z) { return V3{x, y, z}; }
GDEV_INLINE V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
GDEV_INLINE V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
GDEV_INLINE V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
GDEV_INLINE V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
GDEV_INLINE V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
GDEV_INLINE V3 divs(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }      // glm vec3 / s
GDEV_INLINE V3 cdiv(V3 a, float s) { float i = 1.0f / s; return a * i; }        // Color3f / s
GDEV_INLINE float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
GDEV_INLINE V3 cross(V3 a, V3 b) { return V3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
GDEV_INLINE float length(V3 v) { return __builtin_sqrtf(dot(v, v)); }
GDEV_INLINE float lengthSq(V3 v) { return dot(v, v); }
GDEV_INLINE V3 normalize(V3 v) { return v * (1.0f / __builtin_sqrtf(dot(v, v))); }
GDEV_INLINE bool isBlack(V3 c) { return c.x <= 0.0f && c.y <= 0.0f && c.z <= 0.0f; }
GDEV_INLINE float luminance(V3 c) { return c.x * 0.212671f + c.y * 0.715160f + c.z * 0.072169f; }
GDEV_INLINE bool hasNaN(V3 c) {
    return __builtin_isnan(c.x) || __builtin_isnan(c.y) || __builtin_isnan(c.z) || __builtin_isinf(c.x) ||
           __builtin_isinf(c.y) || __builtin_isinf(c.z);
}
// SSE minps / maxps select semantics: (a<b)?a:b and (a>b)?a:b
GDEV_INLINE float selMin(float a, float b) { return a < b ? a : b; }
GDEV_INLINE float selMax(float a, float b) { return a > b ? a : b; }
// std::min / std::max
GDEV_INLINE float stdMin(float a, float b) { return (b < a) ? b : a; }
GDEV_INLINE float stdMax(float a, float b) { return (a < b) ? b : a; }
GDEV_INLINE float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// 4x4 column-major matrix, m[c*4 + r]
struct M4 {
    float m[16];
};
// Transform::operator()(Point3f) (transform.h:43-50)
GDEV_INLINE V3 xfPoint(const M4& M, V3 p) {
    const float* m = M.m;
    float o0 = (m[0] * p.x + m[4] * p.y) + (m[8] * p.z + m[12] * 1.0f);
    float o1 = (m[1] * p.x + m[5] * p.y) + (m[9] * p.z + m[13] * 1.0f);
    float o2 = (m[2] * p.x + m[6] * p.y) + (m[10] * p.z + m[14] * 1.0f);
    float o3 = (m[3] * p.x + m[7] * p.y) + (m[11] * p.z + m[15] * 1.0f);
    if (o3 == 1.0f) return V3{o0, o1, o2};
    return divs(V3{o0, o1, o2}, o3);
}
GDEV_INLINE V3 xfVector(const M4& M, V3 v) {
    const float* m = M.m;
    return V3{(m[0] * v.x + m[4] * v.y) + (m[8] * v.z + m[12] * 0.0f),
              (m[1] * v.x + m[5] * v.y) + (m[9] * v.z + m[13] * 0.0f),
              (m[2] * v.x + m[6] * v.y) + (m[10] * v.z + m[14] * 0.0f)};
}
// AnimatedTransform at dt inside (0, 1) as an M4 (motion.h)
GDEV_INLINE M4 motionMatrix(const GndMotion& mo, float dt) {
    M4 M;
    gndMotionMatrix(mo, dt, M.m);
    return M;
}
// transpose(mat3(Minv)) * n (transform.h:58-61)
GDEV_INLINE V3 xfNormal(const M4& Minv, V3 n) {
    const float* m = Minv.m;
    return V3{(m[0] * n.x + m[1] * n.y) + m[2] * n.z, (m[4] * n.x + m[5] * n.y) + m[6] * n.z,
              (m[8] * n.x + m[9] * n.y) + m[10] * n.z};
}

// ------------------------------------------------------------------ HBM layout
// BVH4 node re-laid for the GPU: one 128-byte line.  Same float bounds as the
// reference BVHNode4; child[i] is an absolute node / pack / TLAS-slot index and
// meta[i] = count | (type << 30).
struct alignas(16) DNode {
    float4 xlo, xhi, ylo, yhi, zlo, zhi;
    uint4 child;
    uint4 meta;
};
static_assert(sizeof(DNode) == 128, "node must be one 128-byte line");

// TrianglePack4, 160 B (same SoA lanes as the reference)
struct alignas(16) DPack {
    float4 v0x, v0y, v0z, v1x, v1y, v1z, v2x, v2y, v2z;
    uint4 prim;
};

// Traversal view of a primitive: world->object matrix + BLAS root.  Kept per
// TLAS leaf slot (slot order, not primitive order), so an instance step reads
// one 80-byte record instead of a slot -> primitive index and then the record.
struct alignas(16) DInst {
    M4 worldToObject;
    int32_t root;   // absolute BLAS root node, -1 for an analytic sphere
    float radius;
    int32_t motion;  // index into DScene::motions, -1 when the primitive does not move
    int32_t prim;    // primitive index (BVH::rayIntersect's primitive)
};
static_assert(sizeof(DInst) == 80, "instance record is 5 x 16 bytes");

// Shading view of a primitive
struct alignas(16) DPrim {
    M4 toWorld;
    M4 toWorldInv;
    int32_t shape;      // GND_SHAPE_*
    int32_t mesh;
    int32_t material;
    int32_t emitter;    // index in emitter list, -1 when none
    uint32_t triBase, nTris, cdfBase, hasUV;
    float totalArea, radius;
    float radiance[3];
    int32_t motion;  // index into DScene::motions, -1 when the primitive does not move
};

// A moving primitive (Primitive ctor, primitive.cpp:6-11): its keyframes and
// their inverses as given, and the decomposition AnimatedTransform mixes
struct alignas(16) DPrimMotion {
    M4 start, startInv, end, endInv;
    GndMotion mo;
};

struct alignas(16) DMaterial {
    int32_t type, checker, mediumInside, mediumOutside;
    float albedo[3], roughness;
    float R[3], eta;
    float T[3], alpha;      // alpha = max(1e-4, roughnessToAlpha(roughness))
    float etaRGB[3], pad0;
    float k[3], pad1;
    float Kd[3], pad2;
    float Ks[3], pad3;
};

struct alignas(16) DMedium {
    int32_t type, nx, ny, nz;
    float sigma_a[3], g;
    float sigma_s[3], densityScale;
    float sigma_t[3], maxSigmaT;
    int32_t resX, resY, resZ, pad0;
    float cellX, cellY, cellZ, pad1;
    float gridMaxX, gridMaxY, gridMaxZ, pad2;
    M4 toLocal;
    const float* density;
    const float* majorants;
    // per voxel (i, j, k) of the grid: the 8 corners of its trilinear cell,
    // (i + di, j + dj, k + dk) at 4 di + 2 dj + dk (0 beyond the grid), as 2
    // float4 in one 32-byte sector; null for grids too large to expand
    const float4* corners;
};

// Point / spot / directional / environment emitter
struct DLight {
    int32_t type;
    float pos[3];
    float dir[3];      // spot: normalised cone axis; directional: towards the light
    float power[3];
    float irr[3];
    float cosTotal, cosFalloff;
};

struct DScene {
    const DNode* tlas;
    const DInst* tlasInst;    // per TLAS leaf slot: the instance record of the slot's primitive
    const DNode* blas;
    const DPack* packs;
    const DPrim* prims;
    const float* shading;     // 24 floats per triangle
    const float* cdf;
    const DMaterial* mats;
    const DMedium* media;
    const int32_t* emitters;  // emitter list: primitive index (area) or -(light + 1)
    int32_t nTlas, nEmitters, nPrims, hasIndex, hasGrid, hasMotion, nMedia;
    const DPrimMotion* motions;
    const DLight* lights;
    // environment map (envLight >= 0): RGB image, row-major, and its
    // Distribution2D: per-row conditional func / cdf / integral, marginal cdf
    const float* envRgb;
    const float* envCondFunc;  // envH * envW
    const float* envCondCdf;   // envH * (envW + 1)
    const float* envCondInt;   // envH
    const float* envMargCdf;   // envH + 1
    const int32_t* envCondGuide;  // envH * (kEnvGuide + 2): search ranges of the row CDFs
    const int32_t* envMargGuide;  // kEnvGuide + 2: search ranges of the marginal CDF
    float envMargInt, envStrength;
    int32_t nLights, envLight, envW, envH;
};

// Primitive::getToWorld(time) / its inverse (primitive.cpp:48-51, 82-84;
// AnimatedTransform::interpolate): the keyframes at or beyond the ends, else
// the interpolated matrix and its glm::inverse
GDEV_INLINE void motionAt(const DPrimMotion& m, float time, M4& M, M4& Minv) {
    if (time <= 0.0f) {
        M = m.start;
        Minv = m.startInv;
    } else if (time >= 1.0f) {
        M = m.end;
        Minv = m.endInv;
    } else {
        gndMotionMatrix(m.mo, (time - 0.0f) / (1.0f - 0.0f), M.m);
        gndInverse44(M.m, Minv.m);
    }
}

struct HitRec {
    float t, u, v;
    int32_t inst, prim;
};

}  // namespace gdev
