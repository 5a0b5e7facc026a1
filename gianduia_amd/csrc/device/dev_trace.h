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
//
// traceStream runs the traversal over a stream of rays so that persistent
// kernels can hand a lane a new ray as soon as its ray is done.
#pragma once

#include "dev_common.h"

namespace gdev {

constexpr int kStackCap = 128;

using LdsPtr = __attribute__((address_space(3))) int*;

template <int LDSD>
constexpr int spillSize() { return kStackCap - LDSD > 0 ? kStackCap - LDSD : 1; }

// Per-lane stack: `LDSD` entries in LDS (this lane's column, stride = block
// size), deeper entries in a private array owned by the kernel.  Only
// scalars and pointers live here, so the state stays in registers.
// `cap` (1..kStackCap, wave-uniform) bounds the depth: a push beyond it sets
// `overflow` and is dropped, so the walk skips that subtree and still ends
// (the caller reports the overflow as an error, the result is not used).
template <int LDSD>
struct TravStack {
    LdsPtr lds;
    int* spill;
    int stride;
    int sp;
    int cap;
    bool overflow;
    GDEV_INLINE void push(int v) {
        if (sp >= cap) {
            overflow = true;
            return;
        }
        if (sp < LDSD) lds[sp * stride] = v;
        else spill[sp - LDSD] = v;
        ++sp;
    }
    GDEV_INLINE int pop() {
        --sp;
        return sp < LDSD ? lds[sp * stride] : spill[sp - LDSD];
    }
    // Pushes c.x, c.y, c.z, c.w for the set bits 0..3 of `bits`, in that
    // order.  When four entries fit in LDS (the common case) this is branch
    // free: child i is written to the slot its rank among the set bits gives,
    // in increasing i, so a child whose bit is clear is overwritten by the
    // next pushed one or left above the new top (a dead slot).
    GDEV_INLINE void pushBits(const uint4& c, uint32_t bits) {
        const int np = __builtin_popcount(bits);
        if (sp + 4 <= LDSD && sp + np <= cap) {
            const int b0 = (int)(bits & 1u), b1 = (int)((bits >> 1) & 1u), b2 = (int)((bits >> 2) & 1u);
            lds[sp * stride] = (int)c.x;
            lds[(sp + b0) * stride] = (int)c.y;
            lds[(sp + b0 + b1) * stride] = (int)c.z;
            lds[(sp + b0 + b1 + b2) * stride] = (int)c.w;
            sp += np;
            return;
        }
        if (bits & 1u) push((int)c.x);
        if (bits & 2u) push((int)c.y);
        if (bits & 4u) push((int)c.z);
        if (bits & 8u) push((int)c.w);
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

// The whole 80-byte record in five 16-byte loads issued together (the BLAS
// root and the motion index arrive with the matrix).
GDEV_INLINE DInst loadInst(const DInst* inst, uint32_t slot) {
    const float4* p = reinterpret_cast<const float4*>(inst + slot);
    const float4 r0 = p[0], r1 = p[1], r2 = p[2], r3 = p[3];
    const int4 r4 = reinterpret_cast<const int4*>(p)[4];
    DInst I;
    I.worldToObject.m[0] = r0.x; I.worldToObject.m[1] = r0.y; I.worldToObject.m[2] = r0.z; I.worldToObject.m[3] = r0.w;
    I.worldToObject.m[4] = r1.x; I.worldToObject.m[5] = r1.y; I.worldToObject.m[6] = r1.z; I.worldToObject.m[7] = r1.w;
    I.worldToObject.m[8] = r2.x; I.worldToObject.m[9] = r2.y; I.worldToObject.m[10] = r2.z; I.worldToObject.m[11] = r2.w;
    I.worldToObject.m[12] = r3.x; I.worldToObject.m[13] = r3.y; I.worldToObject.m[14] = r3.z;
    I.worldToObject.m[15] = r3.w;
    I.root = r4.x;
    I.radius = __int_as_float(r4.y);
    I.motion = r4.z;
    I.prim = r4.w;
    return I;
}

// Mesh::rayIntersect's 4-lane Moller-Trumbore (mesh.cpp:262-319) on one pack:
// lanes evaluated in order, committed with a strict t < lt.  Returns true on a
// committed hit (for ANY: on the first valid lane).
template <bool ANY>
GDEV_INLINE bool testPackData(float4 v0x, float4 v0y, float4 v0z, float4 v1x, float4 v1y, float4 v1z, float4 v2x,
                              float4 v2y, float4 v2z, uint4 prim, V3 o, V3 d, float& lt, float& hu, float& hv,
                              int& hp) {
    bool hit = false;
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
            hit = true;
        }
    }
    return hit;
}

template <bool ANY>
GDEV_INLINE bool testPack(const DPack* packs, uint32_t p, V3 o, V3 d, float& lt, float& hu, float& hv, int& hp) {
    const float4* pk = reinterpret_cast<const float4*>(packs + p);
    return testPackData<ANY>(pk[0], pk[1], pk[2], pk[3], pk[4], pk[5], pk[6], pk[7], pk[8],
                             reinterpret_cast<const uint4*>(pk)[9], o, d, lt, hu, hv, hp);
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

// Slab test with IEEE min/max.  Equal to slab4 when no product is NaN: node
// bounds of non-empty children are finite, so that holds whenever the inverse
// direction is finite (empty children are ignored by their type).
GDEV_INLINE uint32_t slab4Fast(const DNode& n, V3 o, V3 inv, float tmin, float tmax) {
    uint32_t mask = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float t1x = (comp(n.xlo, i) - o.x) * inv.x;
        float t2x = (comp(n.xhi, i) - o.x) * inv.x;
        float t1y = (comp(n.ylo, i) - o.y) * inv.y;
        float t2y = (comp(n.yhi, i) - o.y) * inv.y;
        float t1z = (comp(n.zlo, i) - o.z) * inv.z;
        float t2z = (comp(n.zhi, i) - o.z) * inv.z;
        float tn = __builtin_fmaxf(__builtin_fmaxf(__builtin_fminf(t1x, t2x), __builtin_fminf(t1y, t2y)),
                                   __builtin_fmaxf(__builtin_fminf(t1z, t2z), tmin));
        float tf = __builtin_fminf(__builtin_fminf(__builtin_fmaxf(t1x, t2x), __builtin_fmaxf(t1y, t2y)),
                                   __builtin_fminf(__builtin_fmaxf(t1z, t2z), tmax));
        mask |= (tn <= tf) ? (1u << i) : 0u;
    }
    return mask;
}

GDEV_INLINE V3 selV3(bool c, V3 a, V3 b) { return V3{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z}; }

GDEV_INLINE bool finite3(V3 v) {
    return __builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z);
}

// The leaf children of one node, consumed in child order: each hit leaf
// child is a range of `count` consecutive entries (triangle packs of a BLAS
// leaf, primitive slots of a TLAS leaf) starting at its child index.  The
// node's child and meta words are copied whole and a mask marks the leaves
// still to consume, so recording a node's leaves is one unconditional copy
// (no per-child branches whose joins would copy the queue state).
struct LeafQueue {
    uint4 child, meta;  // the node's words (count in meta's low 30 bits)
    uint32_t mask;      // leaf children not consumed yet
    uint32_t next;      // position inside the lowest one's range
    GDEV_INLINE void clear() { mask = 0u; }
    GDEV_INLINE bool empty() const { return mask == 0u; }
    GDEV_INLINE void set(const uint4& c, const uint4& m, uint32_t leaves) {
        child = c;
        meta = m;
        mask = leaves;
        next = 0u;
    }
    GDEV_INLINE uint32_t take() {
        const int i = __builtin_ctz(mask);
        const uint32_t v = compu(child, i) + next;
        const uint32_t cnt = compu(meta, i) & 0x3FFFFFFFu;
        const bool last = next + 1u >= cnt;
        mask = last ? (mask & (mask - 1u)) : mask;
        next = last ? 0u : next + 1u;
        return v;
    }
};

// Inner children of a node: bit i when child i is a node (type 1).
GDEV_INLINE uint32_t innerBits(const uint4& meta) {
    uint32_t m = 0u;
#pragma unroll
    for (int i = 0; i < 4; ++i) m |= ((compu(meta, i) >> 30) == 1u) ? (1u << i) : 0u;
    return m;
}

// Children of a node with a non-empty leaf range: bit i when child i is a
// leaf (type 2) with a count above zero.
GDEV_INLINE uint32_t leafBits(const uint4& meta) {
    uint32_t m = 0u;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t w = compu(meta, i);
        m |= ((w >> 30) == 2u && (w & 0x3FFFFFFFu) != 0u) ? (1u << i) : 0u;
    }
    return m;
}

enum : int { ACT_NONE = 0, ACT_PACK = 1, ACT_NODE = 2, ACT_INST = 3 };

// The step kind run by the wave this iteration: the kind most lanes wait for,
// with lanes waiting for a pack test or an instance step counted twice (instance
// steps first on ties, then pack tests), so the expensive steps run with most
// of the wave active.  ACT_NONE when no lane has work.  Counting the rarer
// kinds double runs them before too many lanes idle behind node pops: 1080p
// scene traversal 42.6 -> 41.0 ms per frame; weights (2,5,5)/2 do no better.
#ifndef GND_PACK_WEIGHT
#define GND_PACK_WEIGHT 2
#endif
#ifndef GND_INST_WEIGHT
#define GND_INST_WEIGHT 2
#endif
GDEV_INLINE int pickAction(int act) {
    const uint64_t mP = __ballot(act == ACT_PACK);
    const uint64_t mN = __ballot(act == ACT_NODE);
    const uint64_t mI = __ballot(act == ACT_INST);
    const int nP = __popcll(mP) * GND_PACK_WEIGHT, nN = __popcll(mN), nI = __popcll(mI) * GND_INST_WEIGHT;
    if (nI > 0 && nI >= nN && nI >= nP) return ACT_INST;
    if (nP > 0 && nP >= nN) return ACT_PACK;
    if (nN > 0) return ACT_NODE;
    return ACT_NONE;
}

// Scene::rayIntersect over a stream of rays: closest hit (ANY = false) or any
// hit (ANY = true).  All lanes of the wave call this together.
//
// Each lane runs the reference traversal as a state machine whose steps are a
// BVH4 node pop, a TrianglePack4 test, or an instance start / finish.  Per
// lane the steps happen exactly in the reference order (a node's leaf
// children are tested, in child order, before the next pop; the instances of
// a TLAS leaf are traversed in order before the next TLAS pop), so results and
// visit counts are the reference's.  Across the wave, every loop iteration
// runs one kind of step (pickAction).
//
// Rays come from `fetch(idle, o, d, tmin, tmax, drained)`, called by the whole
// wave while `drained` is false: it returns true on the idle lanes it hands a
// ray and sets `drained` (wave-uniformly) once no ray is left.  A lane whose
// ray is done reports it through `finish(found, hit, nodePops)` and becomes
// idle, so a wave does not run at the pace of its slowest ray.  With COUNT,
// cnt[0..3] accumulate node pops, pack tests, instance tests and loop
// iterations, and cnt[4] the iterations in which the lane ran a step.
// MOTION: the scene has moving primitives (instances take the ray's time)
template <bool ANY, bool COUNT, int LDSD, bool MOTION = false, class Fetch, class Finish>
GDEV_INLINE void traceStream(const DScene& S, TravStack<LDSD>& st, uint32_t* cnt, Fetch&& fetch,
                             Finish&& finish) {
    V3 o = V3{0.0f, 0.0f, 0.0f}, d = o, winv = o;  // world ray
    float tmin = 0.0f, tcur = 0.0f;  // tcur: the world ray's mutable tMax
    float time = 0.0f;               // the ray's time (moving primitives)
    float best = kFltMax;            // isect.t, starts at FLT_MAX (bvh.cpp:39)
    V3 lo = o, ld = o, linv = o;     // object-space ray of the current instance
    float lt = 0.0f, hu = 0.0f, hv = 0.0f;
    int hp = 0, curInst = -1, level = 0, blasBase = 0;
    bool instHit = false, found = false, busy = false, done = true, drained = false;
    uint32_t nodes = 0;
    HitRec hit;
    hit.t = finf();
    hit.u = hit.v = 0.0f;
    hit.inst = hit.prim = -1;
    LeafQueue tq, bq;
    tq.clear();
    bq.clear();
    for (;;) {
        if (COUNT) cnt[3]++;
        int act = ACT_NONE;
        if (busy) {
            if (!done) {
                if (!bq.empty()) act = ACT_PACK;
                else if (level == 1) act = st.sp > blasBase ? ACT_NODE : ACT_INST;
                else if (!tq.empty()) act = ACT_INST;
                else if (st.sp > 0) act = ACT_NODE;
                else done = true;
            }
            if (done) {
                finish(found, hit, nodes);
                busy = false;
            }
        }
        if (!drained) {
            V3 ro, rd;
            float rtmin, rtmax, rtime = 0.0f;
            if (fetch(!busy, ro, rd, rtmin, rtmax, rtime, drained)) {
                busy = true;
                o = ro;
                d = rd;
                time = rtime;
                winv = V3{1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z};
                tmin = rtmin;
                tcur = rtmax;
                best = kFltMax;
                found = false;
                done = S.nTlas == 0;
                level = 0;
                blasBase = 0;
                instHit = false;
                nodes = 0;
                hit.t = finf();
                hit.u = hit.v = 0.0f;
                hit.inst = hit.prim = -1;
                tq.clear();
                bq.clear();
                st.sp = 0;
                if (!done) {
                    st.push(0);
                    act = ACT_NODE;
                }
            }
        }
        if (drained && __ballot(busy) == 0) break;
        const int pick = pickAction(act);
        if (pick == ACT_NONE || act != pick) continue;
        if (COUNT) cnt[4]++;
        // Primitive::rayIntersect returned: BVH::rayIntersect keeps the hit.
        // Run by the BLAS step that leaves the instance with nothing left to
        // visit, in the same iteration (no step of its own).
        auto leaveInstance = [&]() {
            if (instHit && lt < best) {
                tcur = lt;
                best = lt;
                hit.t = lt;
                hit.u = hu;
                hit.v = hv;
                hit.inst = curInst;
                hit.prim = hp;
                found = true;
            }
            level = 0;
        };
        if (act == ACT_PACK) {
            if (COUNT) cnt[1]++;
            const uint32_t p = bq.take();
            if (testPack<ANY>(S.packs, p, lo, ld, lt, hu, hv, hp)) {
                if (ANY) {
                    found = true;
                    done = true;
                } else {
                    instHit = true;
                }
            }
            if (bq.empty() && st.sp == blasBase) leaveInstance();
        }
        // the node this lane visits in this iteration: a popped one, or the
        // BLAS root of the instance it enters (entering an instance and
        // visiting its root are one step)
        int visit = -1;
        if (act == ACT_NODE) {
            visit = st.pop();
        } else if (act == ACT_INST) {
            if (level == 1) {
                leaveInstance();
            } else {
                const DInst I = loadInst(S.tlasInst, tq.take());
                const int pi = I.prim;
                if (COUNT) cnt[2]++;
                if (!MOTION || I.motion < 0) {
                    lo = xfPoint(I.worldToObject, o);
                    ld = xfVector(I.worldToObject, d);
                } else {
                    // Primitive::rayIntersect: interpolate(time).inverse() (primitive.cpp:48-50)
                    M4 W, Wi;
                    motionAt(S.motions[I.motion], time, W, Wi);
                    lo = xfPoint(Wi, o);
                    ld = xfVector(Wi, d);
                }
                lt = tcur;
                curInst = pi;
                instHit = false;
                if (I.root >= 0) {
                    linv = V3{1.0f / ld.x, 1.0f / ld.y, 1.0f / ld.z};
                    level = 1;
                    blasBase = st.sp;
                    visit = I.root;
                } else if (traceSphere(I.radius, lo, ld, lt)) {
                    if (ANY) {
                        found = true;
                        done = true;
                    } else if (lt < best) {
                        tcur = lt;
                        best = lt;
                        hit.t = lt;
                        hit.u = 0.0f;
                        hit.v = 0.0f;
                        hit.inst = pi;
                        hit.prim = 0;
                        found = true;
                    }
                }
            }
        }
        if (visit >= 0) {
            if (COUNT) {
                cnt[0]++;
                nodes++;
            }
            const bool blas = level == 1;
            const DNode n = loadNode(blas ? S.blas : S.tlas, visit);
            const V3 ro = selV3(blas, lo, o);
            const V3 ri = selV3(blas, linv, winv);
            const float rmin = blas ? kEps : tmin;
            const float rmax = blas ? lt : tcur;
            const bool fast = __all(finite3(ri));
            const uint32_t mask = fast ? slab4Fast(n, ro, ri, rmin, rmax) : slab4(n, ro, ri, rmin, rmax);
            // inner children are pushed in child order; the hit leaves are
            // then consumed in child order before the next pop (bvhTraversal.h:78-92)
            st.pushBits(n.child, mask & innerBits(n.meta));
            const uint32_t leaves = mask & leafBits(n.meta);
            if (blas) bq.set(n.child, n.meta, leaves);
            else tq.set(n.child, n.meta, leaves);
            if (blas && leaves == 0u && st.sp == blasBase) leaveInstance();
        }
    }
}

// One ray per lane (all lanes of the wave call together; inactive lanes pass
// active = false).  Returns whether a hit was found; `hit` holds the closest.
template <bool ANY, bool COUNT, int LDSD, bool MOTION = false>
GDEV_INLINE bool traceRay(const DScene& S, V3 o, V3 d, float tmin, float tmax, float time, HitRec& hit,
                          TravStack<LDSD>& st, uint32_t* cnt, bool active = true) {
    bool result = false;
    hit.t = finf();
    hit.u = hit.v = 0.0f;
    hit.inst = hit.prim = -1;
    traceStream<ANY, COUNT, LDSD, MOTION>(
        S, st, cnt,
        [&](bool idle, V3& ro, V3& rd, float& rtmin, float& rtmax, float& rtime, bool& drained) {
            drained = true;
            if (!idle || !active) return false;
            ro = o;
            rd = d;
            rtmin = tmin;
            rtmax = tmax;
            rtime = time;
            return true;
        },
        [&](bool found, const HitRec& h, uint32_t) {
            result = found;
            hit = h;
        });
    return result;
}

}  // namespace gdev
