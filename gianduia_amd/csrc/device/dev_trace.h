// dev_trace.h — two-level BVH4 traversal on CDNA4, one ray per lane.
//
// Follows traverseBVH4 (include/gianduia/core/bvhTraversal.h:19-98),
// BVH::rayIntersect (src/scene/bvh.cpp:38-66), Primitive::rayIntersect
// (src/shapes/primitive.cpp:48-59), Mesh::rayIntersect (src/shapes/mesh.cpp:231-326)
// and Sphere::rayIntersect (src/shapes/sphere.cpp:15-39) bit for bit:
//  * slab test with SSE minps/maxps select semantics and the node's hit mask
//    computed once per pop with the tMax current at the pop;
//  * children visited 0..3: inner children pushed, leaves intersected inline,
//    so the LIFO pop order equals the reference's;
//  * Moller-Trumbore lanes evaluated in order with a strict t < tMax commit;
//  * object rays restart at tMin = 1e-4 (transform.h:63-67).
// The per-lane stack lives in LDS (short stack, `LDSD` entries per lane,
// stride = block size so lanes hit distinct banks) and spills to private
// scratch beyond that.  TLAS and BLAS entries share one stack.
#pragma once

#include "dev_common.h"

namespace gdev {

constexpr int kStackCap = 128;

template <int LDSD>
struct TravStack {
    int* lds;
    int stride;
    int sp;
    bool overflow;
    int spill[kStackCap - LDSD > 0 ? kStackCap - LDSD : 1];
    GDEV_INLINE void push(int v) {
        if (sp < LDSD) lds[sp * stride] = v;
        else if (sp < kStackCap) spill[sp - LDSD] = v;
        else overflow = true;
        ++sp;
    }
    GDEV_INLINE int pop() {
        --sp;
        if (sp < LDSD) return lds[sp * stride];
        if (sp < kStackCap) return spill[sp - LDSD];
        return 0;
    }
};

GDEV_INLINE float comp(const float4& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w)); }
GDEV_INLINE uint32_t compu(const uint4& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w)); }

// 4-wide slab test; bit i set when child i is hit (bvhTraversal.h:56-79)
GDEV_INLINE uint32_t slab4(const DNode& n, V3 o, V3 inv, float tmin, float tmax) {
    uint32_t mask = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float t1x = (comp(n.xlo, i) - o.x) * inv.x;
        float t2x = (comp(n.xhi, i) - o.x) * inv.x;
        float t1y = (comp(n.ylo, i) - o.y) * inv.y;
        float t2y = (comp(n.yhi, i) - o.y) * inv.y;
        float t1z = (comp(n.zlo, i) - o.z) * inv.z;
        float t2z = (comp(n.zhi, i) - o.z) * inv.z;
        float tn = selMax(selMax(selMin(t1x, t2x), selMin(t1y, t2y)), selMax(selMin(t1z, t2z), tmin));
        float tf = selMin(selMin(selMax(t1x, t2x), selMax(t1y, t2y)), selMin(selMax(t1z, t2z), tmax));
        mask |= (tn <= tf) ? (1u << i) : 0u;
    }
    return mask;
}

GDEV_INLINE DNode loadNode(const DNode* nodes, int idx) {
    const float4* p = reinterpret_cast<const float4*>(nodes + idx);
    DNode n;
    n.xlo = p[0]; n.xhi = p[1]; n.ylo = p[2]; n.yhi = p[3]; n.zlo = p[4]; n.zhi = p[5];
    n.child = reinterpret_cast<const uint4*>(p)[6];
    n.meta = reinterpret_cast<const uint4*>(p)[7];
    return n;
}

// Mesh::rayIntersect over one BLAS (object space).  lt is the object ray's
// mutable tMax; returns true when a hit was committed (or any hit for ANY).
template <bool ANY, bool COUNT, int LDSD>
GDEV_INLINE bool traceMesh(const DScene& S, int root, V3 o, V3 d, float& lt, float& hu, float& hv, int& hp,
                           TravStack<LDSD>& st, uint32_t* cnt) {
    const V3 inv = V3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    bool hitAny = false;
    const int base = st.sp;
    st.push(root);
    while (st.sp > base) {
        const int ni = st.pop();
        const DNode n = loadNode(S.blas, ni);
        if (COUNT) cnt[0]++;
        const uint32_t mask = slab4(n, o, inv, kEps, lt);
        if (mask == 0) continue;
        for (int i = 0; i < 4; ++i) {
            if (!((mask >> i) & 1u)) continue;
            const uint32_t meta = compu(n.meta, i);
            const uint32_t c = compu(n.child, i);
            if ((meta >> 30) == 1u) {
                st.push((int)c);
            } else if ((meta >> 30) == 2u) {
                const uint32_t np = meta & 0x3FFFFFFFu;
                for (uint32_t p = c; p < c + np; ++p) {
                    if (COUNT) cnt[1]++;
                    const float4* pk = reinterpret_cast<const float4*>(S.packs + p);
                    const float4 v0x = pk[0], v0y = pk[1], v0z = pk[2];
                    const float4 v1x = pk[3], v1y = pk[4], v1z = pk[5];
                    const float4 v2x = pk[6], v2y = pk[7], v2z = pk[8];
                    const uint4 prim = reinterpret_cast<const uint4*>(pk)[9];
#pragma unroll
                    for (int l = 0; l < 4; ++l) {
                        const float ax = comp(v0x, l), ay = comp(v0y, l), az = comp(v0z, l);
                        const float e1x = comp(v1x, l) - ax, e1y = comp(v1y, l) - ay, e1z = comp(v1z, l) - az;
                        const float e2x = comp(v2x, l) - ax, e2y = comp(v2y, l) - ay, e2z = comp(v2z, l) - az;
                        const float px = d.y * e2z - d.z * e2y;
                        const float py = d.z * e2x - d.x * e2z;
                        const float pz = d.x * e2y - d.y * e2x;
                        const float det = (e1x * px + e1y * py) + e1z * pz;
                        bool valid = __builtin_fabsf(det) > 1e-8f;
                        const float invDet = 1.0f / det;
                        const float tx = o.x - ax, ty = o.y - ay, tz = o.z - az;
                        const float u = ((tx * px + ty * py) + tz * pz) * invDet;
                        valid = valid && (u >= 0.0f) && (u <= 1.0f);
                        const float qx = ty * e1z - tz * e1y;
                        const float qy = tz * e1x - tx * e1z;
                        const float qz = tx * e1y - ty * e1x;
                        const float v = ((d.x * qx + d.y * qy) + d.z * qz) * invDet;
                        valid = valid && (v >= 0.0f) && ((u + v) <= 1.0f);
                        const float t = ((e2x * qx + e2y * qy) + e2z * qz) * invDet;
                        valid = valid && (t > kEps) && (t < lt);
                        if (valid) {
                            if (ANY) return true;
                            lt = t;
                            hu = u;
                            hv = v;
                            hp = (int)compu(prim, l);
                            hitAny = true;
                        }
                    }
                }
            }
        }
    }
    return hitAny;
}

// Sphere::rayIntersect with SolveQuadratic (object space, tMin = 1e-4)
GDEV_INLINE bool traceSphere(float radius, V3 o, V3 d, float& lt) {
    const float A = lengthSq(d);
    const float B = 2.0f * dot(o, d);
    const float C = lengthSq(o) - (radius * radius);
    const float disc = B * B - 4.0f * A * C;
    if (disc < 0) return false;
    const float root = __builtin_sqrtf(disc);
    const float t0 = (-B - root) / (2.0f * A);
    const float t1 = (-B + root) / (2.0f * A);
    if (t0 > kEps && t0 < lt) { lt = t0; return true; }
    if (t1 > kEps && t1 < lt) { lt = t1; return true; }
    return false;
}

// Scene::rayIntersect: closest hit (ANY = false) or any hit (ANY = true).
template <bool ANY, bool COUNT, int LDSD>
GDEV_INLINE bool traceRay(const DScene& S, V3 o, V3 d, float tmin, float tmax, HitRec& hit, TravStack<LDSD>& st,
                          uint32_t* cnt) {
    hit.inst = -1;
    hit.t = finf();
    if (S.nTlas == 0) return false;
    const V3 inv = V3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    float tcur = tmax;          // the world ray's mutable tMax
    float best = kFltMax;       // isect.t starts at FLT_MAX (bvh.cpp:39)
    bool hitAny = false;
    st.sp = 0;
    st.push(0);
    while (st.sp > 0) {
        const int ni = st.pop();
        const DNode n = loadNode(S.tlas, ni);
        if (COUNT) cnt[0]++;
        const uint32_t mask = slab4(n, o, inv, tmin, tcur);
        if (mask == 0) continue;
        for (int i = 0; i < 4; ++i) {
            if (!((mask >> i) & 1u)) continue;
            const uint32_t meta = compu(n.meta, i);
            const uint32_t c = compu(n.child, i);
            if ((meta >> 30) == 1u) {
                st.push((int)c);
            } else if ((meta >> 30) == 2u) {
                const uint32_t np = meta & 0x3FFFFFFFu;
                for (uint32_t p = c; p < c + np; ++p) {
                    const int pi = S.tlasPrims[p];
                    const DInst& I = S.inst[pi];
                    if (COUNT) cnt[2]++;
                    const V3 lo = xfPoint(I.worldToObject, o);
                    const V3 ld = xfVector(I.worldToObject, d);
                    float lt = tcur;
                    float hu = 0.0f, hv = 0.0f;
                    int hp = 0;
                    bool h;
                    if (I.root >= 0) h = traceMesh<ANY, COUNT, LDSD>(S, I.root, lo, ld, lt, hu, hv, hp, st, cnt);
                    else h = traceSphere(I.radius, lo, ld, lt);
                    if (h) {
                        if (ANY) return true;
                        if (lt < best) {
                            tcur = lt;
                            best = lt;
                            hit.t = lt;
                            hit.u = hu;
                            hit.v = hv;
                            hit.inst = pi;
                            hit.prim = hp;
                            hitAny = true;
                        }
                    }
                }
            }
        }
    }
    return hitAny;
}

}  // namespace gdev
