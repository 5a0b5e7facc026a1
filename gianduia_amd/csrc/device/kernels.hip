// kernels.hip — CDNA4 kernels of the ray-tracing hot path and the C-ABI of
// libgnd_hip.so (declared in include/gnd_hip.h).
//
// Wavefront form: camera rays, closest-hit traversal, shading + next-event
// estimation, shadow traversal and the film gather run as separate kernels
// over struct-of-arrays path state in HBM.  Live paths are compacted between
// bounces with a wave-wide ballot + popcount and one atomic per wave.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "dev_common.h"
#include "dev_shade.h"
#include "dev_trace.h"
#include "denoise.h"

using namespace gdev;

// ============================================================================ device state

// The ray of each queued path, in queue order (the closest-hit kernel reads
// rays with coalesced loads and without a path-index indirection).
struct RayQ {
    float *ox, *oy, *oz, *dx, *dy, *dz;
};

struct PathSoA {
    float *tpr, *tpg, *tpb, *Lr, *Lg, *Lb;
    float *pdfPrev, *lx, *ly, *lz;
    float *weight, *pfx, *pfy;
    uint64_t* rng;
    uint32_t *c0, *c1, *flags;  // c0: bounces | nullBounces<<16; c1: surface | volume bounces<<16
    int32_t* job;
    int32_t* chunk;  // engine chunk of the slot (continuous batching)
    // event record between the NEE and scatter kernels (split mode)
    int32_t *evKind, *evPrim, *evMedium;
    float *evpx, *evpy, *evpz, *evnx, *evny, *evnz, *evtx, *evty, *evtz, *evu, *evv;
    float* time;  // the camera sample's ray time (moving primitives)
    // AOV record per slot (allocated when a job asks for AOVs, else null):
    // field k of slot p at aov[k * aovStride + p], fields as in kAovFloats
    float* aov;
    uint32_t aovStride;
};

// AOVRecord (include/gianduia/core/film.h:13-19): albedo rgb, shading normal
// xyz mapped to [0,1], depth along the camera forward axis, roughness, metallic
constexpr int kAovFloats = 9;

// Closest hits per queue entry, one 32-byte record each: (t, u, v, instance)
// then (primitive, unused x3).  A ray's hit is written once, when its
// traversal ends, by lanes that end at different times: one record keeps that
// write inside one 32-byte memory sector instead of dirtying a sector in each
// of five arrays.
struct HitSoA {
    uint4* rec;
};
GDEV_INLINE void storeHit(const HitSoA& H, int64_t i, bool found, const HitRec& h) {
    H.rec[2 * i] = make_uint4(__float_as_uint(found ? h.t : finf()), __float_as_uint(h.u), __float_as_uint(h.v),
                              (uint32_t)(found ? h.inst : -1));
    reinterpret_cast<int32_t*>(H.rec + 2 * i + 1)[0] = found ? h.prim : -1;
}
GDEV_INLINE HitRec loadHit(const HitSoA& H, int64_t i) {
    const uint4 a = H.rec[2 * i];
    HitRec h;
    h.t = __uint_as_float(a.x);
    h.u = __uint_as_float(a.y);
    h.v = __uint_as_float(a.z);
    h.inst = (int32_t)a.w;
    h.prim = reinterpret_cast<const int32_t*>(H.rec + 2 * i + 1)[0];
    return h;
}

struct ShadowSoA {
    int32_t *path, *medium;
    float *ox, *oy, *oz, *dx, *dy, *dz, *tmax;
    float *pr, *pg, *pb;   // pending contribution
    float *lpx, *lpy, *lpz;  // light point (transmittance segments)
    float* time;             // the shadow ray's time (moving primitives)
};

struct DJob {
    DCamera cam;
    float* film;
    float* filmSq;
    float* filmAov;  // weighted AOV sums, kAovFloats per pixel, or null
    uint64_t seedOffset;
    int32_t maxDepth, integrator, W, H;
    float aoLength;  // ao: occlusion ray tMax
};

struct DChunk {     // one pass of one job: npix path slots from slotBase
    uint32_t slotBase, npix;
    int32_t job, pass;
    uint32_t first;  // index of its first path in the camera launch that starts it
    int32_t id;      // engine chunk id (continuous batching), -1 otherwise
    uint32_t qBase;  // in-queue position of its first live path, after WaveArgs::camQueueBase
};

struct DSplatTask { // one job inside a wave: its chunks in pass order
    uint32_t pixBase;   // prefix over tasks of W*H
    int32_t job, chunkBegin, chunkEnd;
};

// flags layout
constexpr uint32_t F_SPECULAR = 1u;
constexpr uint32_t F_GBUF = 1u << 24;  // the path still needs its AOV albedo and normal
constexpr uint32_t F_TIME0 = 1u << 25;  // the path's ray has time 0 (direct's BSDF ray, direct.cpp:60)
GDEV_INLINE int flagChannel(uint32_t f) { return (int)((f >> 1) & 3u) - 1; }  // -1 none, 0..2
GDEV_INLINE int flagMedium(uint32_t f) { return (int)((f >> 8) & 0xFFFFu) - 1; }
GDEV_INLINE uint32_t setMedium(uint32_t f, int m) { return (f & ~(0xFFFFu << 8)) | ((uint32_t)(m + 1) << 8); }

struct Counters {
    unsigned long long nodes, packs, insts;      // closest-hit traversals
    unsigned long long snodes, spacks, sinsts;   // shadow / transmittance traversals
    unsigned long long nan, stackOverflow;
    unsigned long long maxRayNodes, badRays;     // worst closest-hit ray, rays with non-finite directions
    unsigned long long iters, siters;            // lane slots of traversal loop iterations (64 per wave step)
    unsigned long long steps, ssteps;            // of those, the lane slots that ran a step
    unsigned long long rays, srays, launches;    // closest-hit and shadow queries, bounces with work
    unsigned long long tailRays, tailSrays;      // the same queries made by the tail kernel
};

// Paths that end leave their engine chunk (continuous batching): one atomic
// per distinct chunk among the wave's ending lanes.  All lanes call this.
GDEV_INLINE void retirePaths(uint32_t* chunkLive, bool ending, int chunk) {
    if (!chunkLive) return;
    uint64_t m = __ballot(ending);
    while (m) {
        const int leader = __ffsll((unsigned long long)m) - 1;
        const int c = __shfl(chunk, leader);
        const uint64_t same = __ballot(ending && chunk == c);
        if ((int)__lane_id() == leader) atomicSub(&chunkLive[c], (uint32_t)__popcll(same));
        m &= ~same;
    }
}

// Block-aggregated append to three queues: one atomic per block and queue
// (a counter shared by the whole launch serialises its atomics, so fewer is
// faster).  Every thread of the block calls this; it returns this lane's
// slots in the queues.  c1 and c2 may be null.
// Threads per block of the shading kernels: at 2 waves per SIMD a CU holds
// two 256-thread blocks, so it starts the next block while the slowest wave
// of the other finishes (1080p shading 14.3 ms per frame with 512-thread
// blocks, 13.8 with 256, 14.0 with 128).
#ifndef GND_SHADE_BLOCK
#define GND_SHADE_BLOCK 256
#endif
constexpr int kShadeBlock = GND_SHADE_BLOCK;
// waves per SIMD the shading kernel and the transmittance shadow kernel are
// compiled for (register budget; on the volumetric scene the shading kernel
// gains nothing from 3; the persistent transmittance kernel takes 33.8 ms per
// frame at 3 waves (168 VGPRs or fewer), 43.4 ms at 2 and 52 ms at 4, which
// spills)
#ifndef GND_SHADE_WAVES
#define GND_SHADE_WAVES 2
#endif
#ifndef GND_SHADOWTR_WAVES
#define GND_SHADOWTR_WAVES 3
#endif
GDEV_INLINE void blockAppend3(uint32_t* c0, bool p0, uint32_t& pos0, uint32_t* c1, bool p1, uint32_t& pos1,
                              uint32_t* c2, bool p2, uint32_t& pos2) {
    __shared__ uint32_t sN[3][kShadeBlock / 64];
    __shared__ uint32_t sBase[3];
    const int w = (int)(threadIdx.x >> 6), nw = (int)(blockDim.x >> 6);
    const uint64_t m0 = __ballot(p0), m1 = __ballot(c1 && p1), m2 = __ballot(c2 && p2);
    if (__lane_id() == 0) {
        sN[0][w] = (uint32_t)__popcll(m0);
        sN[1][w] = (uint32_t)__popcll(m1);
        sN[2][w] = (uint32_t)__popcll(m2);
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        uint32_t total = 0;
        for (int i = 0; i < nw; ++i) total += sN[threadIdx.x][i];
        uint32_t* c = threadIdx.x == 0 ? c0 : (threadIdx.x == 1 ? c1 : c2);
        sBase[threadIdx.x] = (total && c) ? atomicAdd(c, total) : 0u;
    }
    __syncthreads();
    uint32_t pre0 = 0, pre1 = 0, pre2 = 0;
    for (int i = 0; i < w; ++i) {
        pre0 += sN[0][i];
        pre1 += sN[1][i];
        pre2 += sN[2][i];
    }
    const uint64_t lt = (1ull << __lane_id()) - 1ull;
    pos0 = sBase[0] + pre0 + (uint32_t)__popcll(m0 & lt);
    pos1 = sBase[1] + pre1 + (uint32_t)__popcll(m1 & lt);
    pos2 = sBase[2] + pre2 + (uint32_t)__popcll(m2 & lt);
    __syncthreads();  // sN / sBase are reused by the next call
}

// XCD-aware block order: the dispatcher deals workgroups round-robin over the
// 8 XCDs; remap so each XCD walks a contiguous range (coherent rays share L2).
GDEV_INLINE uint32_t xcdBlock() {
    const uint32_t g = gridDim.x, b = blockIdx.x;
    if (g % 8u != 0u) return b;
    return (b % 8u) * (g / 8u) + b / 8u;
}

template <int LDSD>
GDEV_INLINE TravStack<LDSD> makeStack(int* lds, int* spill, int cap) {
    TravStack<LDSD> st;
    st.lds = (LdsPtr)(lds + threadIdx.x);
    st.spill = spill;
    st.stride = (int)blockDim.x;
    st.sp = 0;
    st.cap = cap;
    st.overflow = false;
    return st;
}

GDEV_INLINE void flushCounters(Counters* c, const uint32_t* cnt, bool overflow, bool shadow = false,
                               uint32_t maxRay = 0, uint32_t bad = 0) {
    if (maxRay) atomicMax(&c->maxRayNodes, (unsigned long long)maxRay);
    if (bad) atomicAdd(&c->badRays, (unsigned long long)bad);
    unsigned long long a = cnt[0], b = cnt[1], d = cnt[2], it = cnt[3], w = cnt[4];
    // wave reduction then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off);
        b += __shfl_down(b, off);
        d += __shfl_down(d, off);
        it += __shfl_down(it, off);
        w += __shfl_down(w, off);
    }
    if (__lane_id() == 0) {
        atomicAdd(shadow ? &c->snodes : &c->nodes, a);
        atomicAdd(shadow ? &c->spacks : &c->packs, b);
        atomicAdd(shadow ? &c->sinsts : &c->insts, d);
        atomicAdd(shadow ? &c->siters : &c->iters, it);
        atomicAdd(shadow ? &c->ssteps : &c->steps, w);
    }
    if (overflow) atomicAdd(&c->stackOverflow, 1ull);
}

// ============================================================================ verify-mode traversal

// Every wave kernel runs with at most 256 threads per block (the tuning
// accepts 64, 128 or 256); the bound lets the compiler use more than 128
// VGPRs instead of spilling to scratch.
#define GND_WAVE_KERNEL __global__ __launch_bounds__(256)
// Traversal kernels are latency bound: their register budget is capped so
// that GND_TRACE_WAVES (closest hit) and GND_SHADOW_WAVES (any hit) waves fit
// on each SIMD.  Measured on the 1080p scene: 4 and 4 (no spills); 5 waves
// spill and lose more than the extra occupancy gains.
#ifndef GND_TRACE_WAVES
#define GND_TRACE_WAVES 4
#endif
#define GND_TRACE_KERNEL __global__ __launch_bounds__(256, GND_TRACE_WAVES)
#ifndef GND_SHADOW_WAVES
#define GND_SHADOW_WAVES 4
#endif

template <int LDSD, bool ANY, bool COUNT>
GND_WAVE_KERNEL void kTraceBatch(DScene S, gnd_ray_soa R, int64_t n, gnd_hit_soa H, int32_t* occ, Counters* ctr,
                                 int32_t stackCap) {
    extern __shared__ int sStack[];
    const int64_t i = (int64_t)xcdBlock() * blockDim.x + threadIdx.x;
    uint32_t cnt[5] = {0, 0, 0, 0, 0};
    int spill[spillSize<LDSD>()];
    TravStack<LDSD> st = makeStack<LDSD>(sStack, spill, stackCap);
    if (i < n) {
        HitRec h;
        bool r = traceRay<ANY, COUNT, LDSD>(S, V3{R.ox[i], R.oy[i], R.oz[i]}, V3{R.dx[i], R.dy[i], R.dz[i]},
                                           R.tmin[i], R.tmax[i], 0.0f, h, st, cnt);
        if (ANY) {
            occ[i] = r ? 1 : 0;
        } else {
            H.t[i] = r ? h.t : finf();
            H.inst[i] = r ? h.inst : -1;
            H.prim[i] = r ? h.prim : -1;
            H.u[i] = r ? h.u : 0.0f;
            H.v[i] = r ? h.v : 0.0f;
        }
    }
    if (COUNT) flushCounters(ctr, cnt, st.overflow);
    else if (st.overflow) atomicAdd(&ctr->stackOverflow, 1ull);
}

// ============================================================================ wavefront path tracer

struct WaveArgs {
    DScene S;
    const DJob* jobs;
    const DChunk* chunks;
    int32_t nChunks;
    uint32_t nSlots;
    PathSoA P;
    HitSoA Hq;            // hits indexed by queue slot
    ShadowSoA Sh;
    int32_t* qIn;
    int32_t* qOut;
    RayQ rIn, rOut;       // rays of the in- and out-queue entries
    int32_t* gIn;         // in-queue indices of the paths starting in a density
    int32_t* gOut;        // grid (split mode), and the same list being built
    uint32_t* counts;     // [0] in, [1] out, [2] shadow, [3] grid out, [4] grid in,
                          // [8..] ray feeds
    uint32_t* chunkLive;  // live paths per engine chunk, or null
    uint32_t camQueueBase;  // in-queue entries before the paths kCamera starts
    Counters* ctr;
    int32_t pathOnlyFast; // shadows may use any-hit
    int32_t split;        // NEE and scatter in separate kernels, density-grid
                          // delta tracking in kMedium ahead of them
    int32_t refill;       // idle lanes before a wave refills: closest-hit
    int32_t srefill;      // and any-hit traversal
    int32_t stackCap;     // traversal stack entries per lane (1..kStackCap)
};

GDEV_INLINE void storeRay(const RayQ& R, uint32_t i, V3 o, V3 d) {
    R.ox[i] = o.x; R.oy[i] = o.y; R.oz[i] = o.z;
    R.dx[i] = d.x; R.dy[i] = d.y; R.dz[i] = d.z;
}
GDEV_INLINE void loadRay(const RayQ& R, uint32_t i, V3& o, V3& d) {
    o = V3{R.ox[i], R.oy[i], R.oz[i]};
    d = V3{R.dx[i], R.dy[i], R.dz[i]};
}

// SamplerIntegrator::render per-sample set-up (integrator.h:88-125): seeding,
// film jitter, lens sample, chromatic-aberration channel, time, camera ray.
// One thread per new path: t in [0, nSlots) over the chunks of this launch
// (sorted by `first`); the path goes to slot slotBase + (t - first).
__global__ void kCamera(WaveArgs A) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nSlots) return;
    int lo = 0, hi = A.nChunks - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (A.chunks[mid].first <= t) lo = mid;
        else hi = mid - 1;
    }
    const DChunk ch = A.chunks[lo];
    const DJob& J = A.jobs[ch.job];
    const uint32_t pix = t - ch.first;
    const uint32_t slot = ch.slotBase + pix;
    const int W = J.W, H = J.H;
    const int x = (int)(pix % (uint32_t)W), y = (int)(pix / (uint32_t)W);
    const uint64_t pixelIdx = (uint64_t)(y * W + x);
    const uint64_t g = pixelIdx + (uint64_t)ch.pass * (uint64_t)(W * H) + J.seedOffset;
    Rng rng;
    rng.s = pcgSeed(splitMix64(g));
    const float jx = rng.next1D(), jy = rng.next1D();
    const float pfx = x + jx, pfy = y + jy;
    const float fx = pfx / W, fy = 1.0f - pfy / H;
    const float lx = rng.next1D(), ly = rng.next1D();
    const float chR = rng.next1D();
    int chan;
    float lambda;
    if (!J.cam.hasCA) { chan = -1; lambda = 0.0f; }
    else if (chR < 0.333333f) { chan = 0; lambda = 1.0f; }
    else if (chR < 0.666666f) { chan = 1; lambda = 0.0f; }
    else { chan = 2; lambda = -1.0f; }
    const float time = rng.next1D();
    V3 ro, rd;
    float rayTime;
    const float w = shootRay(J.cam, fx, fy, lx, ly, lambda, time, ro, rd, rayTime);
    PathSoA& P = A.P;
    P.time[slot] = rayTime;
    P.tpr[slot] = 1.0f; P.tpg[slot] = 1.0f; P.tpb[slot] = 1.0f;
    P.Lr[slot] = 0.0f; P.Lg[slot] = 0.0f; P.Lb[slot] = 0.0f;
    P.pdfPrev[slot] = 1.0f;
    P.lx[slot] = ro.x; P.ly[slot] = ro.y; P.lz[slot] = ro.z;
    P.weight[slot] = w;
    P.pfx[slot] = pfx;
    P.pfy[slot] = pfy;
    P.rng[slot] = rng.s;
    P.c0[slot] = 0u;
    P.c1[slot] = 0u;
    const bool aov = J.filmAov && P.aov;
    P.flags[slot] = F_SPECULAR | ((uint32_t)(chan + 1) << 1) | (aov ? F_GBUF : 0u);  // medium -1 -> field 0
    if (aov) {
        const float init[kAovFloats] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, -1.0f, 0.0f, 0.0f};
#pragma unroll
        for (int k = 0; k < kAovFloats; ++k) P.aov[(size_t)k * P.aovStride + slot] = init[k];
    }
    P.job[slot] = ch.job;
    P.chunk[slot] = ch.id;
    // the host laid out the in-queue: no atomics, and queue order = slot order
    if (J.maxDepth > 0) {
        const uint32_t q = A.camQueueBase + ch.qBase + pix;
        A.qIn[q] = (int32_t)slot;
        storeRay(A.rIn, q, ro, rd);
    }
}

// Persistent traversal kernels hand out rays with one atomic per wave on a
// counter zeroed before the launch.  A lane whose ray is done takes a new one
// as soon as `refill` lanes of its wave are idle (or the whole wave is), so a
// wave does not run at the pace of its slowest ray.  The best threshold
// depends on the scene: a BVH that stays in L2 makes traversal steps cheap
// next to a refill (fetching rays from HBM), so refills should gather more
// lanes; a large BVH makes idle lanes the bigger loss.  The traversal
// kernels take it from the tuning; the medium-walk kernels use kWalkRefill.
// Volumetric 720p frame: 24 idle lanes 72.4 ms, 40 lanes 70.0 ms, 56 lanes
// 70.1 ms (kMedium and kShadow together).
constexpr int kWalkRefill = 40;
// kShadow runs a traversal (all lanes waiting for one take part; a traversal
// lasts as long as its longest ray) only when more than kShadowTraceRatio
// times as many lanes wait for one as wait for a walk step (one step each):
// volumetric 720p transmittance shadows 14.6 ms per frame at a ratio of 1,
// 13.3 at 2, 12.6 at 4 and 8, 12.9 at 16, 16.8 at 1/2.
constexpr int kShadowTraceRatio = 4;
// Queue counters: [0] in-queue size, [1] out-queue, [2] shadow queue, then
// the ray-feed part counters of the two persistent traversals.
#ifndef GND_FEED_PARTS
#define GND_FEED_PARTS 1
#endif
constexpr int kParts = GND_FEED_PARTS;
constexpr uint32_t kFeedTrace = 8, kFeedShadow = kFeedTrace + kParts, kFeedMedium = kFeedShadow + kParts,
                   kFeedTail = kFeedMedium + kParts, kNumCounts = 64;

// The persistent traversals' ray feed: waves take rays in queue order with
// one atomic per refill.  The queue can be cut into kParts contiguous parts
// with a counter each (a wave starts on part blockIdx.x % kParts and moves on
// when its part is used up), which spreads the atomics over more addresses;
// measured on the 1080p scene, one part is fastest (1: 71 ms, 4: 74 ms,
// 16: 76 ms of traversal per frame): with a single counter every wave works
// on the same region of the queue, so the rays in flight are coherent and
// share BVH nodes in L2.
// With one counter (the default) a wave of the closest-hit traversal reserves
// kFeedBlock rays per atomic and hands them out over its next refills (block
// 0: one atomic per refill, for the k rays it hands out).  The counter is
// shared by every wave of the launch and its atomics serialise: with an
// atomic per refill, lowering the refill threshold from 40 to 16 lanes raised
// lane use from 0.48 to 0.55 yet took the 1080p traversal from 24.3 to 39.6 ms
// per frame.  Blocks of 256 rays: 1080p traversal 24.5 -> 23.2 ms, volumetric
// (a small BVH, so refills dominate) 13.4 -> 9.0 ms; 128 and 512 are slower.
// The any-hit traversal takes blocks of 256 too (1080p shadows 12.45 ms per
// frame at a 40-lane threshold without them, 12.18 at 16 lanes with them).
// The walk and tail kernels keep an atomic per refill: their waves' work is
// uneven, and reserved rays left in a slow wave delayed the launch's end
// (volumetric shadows 13.1 -> 14.5 ms with blocks of 64).
#ifndef GND_FEED_BLOCK
#define GND_FEED_BLOCK 256
#endif
constexpr uint32_t kFeedBlock = GND_FEED_BLOCK;
#ifndef GND_SHADOW_FEED_BLOCK
#define GND_SHADOW_FEED_BLOCK 256
#endif
#ifndef GND_WALK_FEED_BLOCK
#define GND_WALK_FEED_BLOCK 0
#endif
constexpr uint32_t kShadowFeedBlock = GND_SHADOW_FEED_BLOCK, kWalkFeedBlock = GND_WALK_FEED_BLOCK;
struct RayFeed {
    uint32_t* ctr;
    uint32_t n;
    int part, tried, refill;  // wave-uniform
    // one counter: the reserved rays not handed out yet, [resBase, resBase + resLeft)
    uint32_t resBase, block;
    int resLeft;
    bool exhausted;
    GDEV_INLINE void init(uint32_t* counters, uint32_t count, int refillLanes, uint32_t blockRays = kFeedBlock) {
        ctr = counters;
        n = count;
        refill = refillLanes;
        block = blockRays;
        part = (int)(blockIdx.x % (unsigned)kParts);
        tried = 0;
        resBase = 0;
        resLeft = 0;
        exhausted = false;
    }
    GDEV_INLINE uint32_t edge(int p) const { return (uint32_t)(((uint64_t)n * (uint64_t)p) / kParts); }
    // Idle lanes take the next rays (once `refill` lanes, or the whole wave,
    // are idle); returns this lane's queue index or -1.  Sets `drained`
    // (wave-uniformly) once every ray is handed out.
    GDEV_INLINE int64_t take(bool idle, uint64_t waveMask, bool& drained) {
        const uint64_t m = __ballot(idle);
        const int k = __popcll(m);
        if (k == 0 || (k < refill && m != waveMask)) return -1;
        const int leader = __ffsll((unsigned long long)m) - 1;
        if (block > 0 && kParts == 1) {
            // a new block when the reservation cannot serve every idle lane;
            // ranks below resLeft take the reservation, the next ones the block
            uint32_t bBase = 0;
            int bLeft = 0;
            if (resLeft < k && !exhausted) {
                uint32_t b = 0;
                if ((int)__lane_id() == leader) b = atomicAdd(ctr, block);
                bBase = __shfl(b, leader);
                bLeft = bBase < n ? (int)min(block, n - bBase) : 0;
                if (bBase >= n || n - bBase <= block) exhausted = true;
            }
            if (resLeft + bLeft == 0) {
                drained = true;
                return -1;
            }
            const int rank = __popcll(m & ((1ull << __lane_id()) - 1ull));
            int64_t r = -1;
            if (idle) {
                if (rank < resLeft) r = (int64_t)resBase + rank;
                else if (rank < resLeft + bLeft) r = (int64_t)bBase + (rank - resLeft);
            }
            const int given = min(k, resLeft + bLeft);
            if (bLeft > 0) {
                // the reservation is used up (a block is taken only when it
                // could not serve all k lanes)
                const int usedB = given - resLeft;
                resBase = bBase + (uint32_t)usedB;
                resLeft = bLeft - usedB;
            } else {
                resBase += (uint32_t)given;
                resLeft -= given;
            }
            return r;
        }
        while (tried < kParts) {
            const uint32_t lo = edge(part), len = edge(part + 1) - lo;
            uint32_t base = 0;
            if ((int)__lane_id() == leader) base = atomicAdd(&ctr[part], (uint32_t)k);
            base = __shfl(base, leader);
            if (base >= len || base + (uint32_t)k >= len) {
                part = (part + 1) % kParts;
                tried++;
            }
            if (base < len) {
                if (!idle) return -1;
                const uint32_t r = base + (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1ull));
                return r < len ? (int64_t)(lo + r) : -1;
            }
        }
        drained = true;
        return -1;
    }
};

// closest hit for the live path rays (Scene::rayIntersect at the top of Li)
template <int LDSD, bool COUNT, bool MOTION>
GND_TRACE_KERNEL void kTracePaths(WaveArgs A) {
    extern __shared__ int sStack[];
    const uint32_t n = A.counts[0];
    const uint64_t waveMask = __ballot(1);
    uint32_t cnt[5] = {0, 0, 0, 0, 0};
    uint32_t maxRay = 0, bad = 0;
    int spill[spillSize<LDSD>()];
    TravStack<LDSD> st = makeStack<LDSD>(sStack, spill, A.stackCap);
    RayFeed feed;
    feed.init(A.counts + kFeedTrace, n, A.refill);
    int64_t cur = -1;  // queue index of this lane's ray
    traceStream<false, COUNT, LDSD, MOTION>(
        A.S, st, cnt,
        [&](bool idle, V3& o, V3& d, float& tmin, float& tmax, float& time, bool& drained) {
            const int64_t i = feed.take(idle, waveMask, drained);
            if (i < 0) return false;
            cur = i;
            loadRay(A.rIn, (uint32_t)i, o, d);
            if (COUNT && !finite3(d)) bad++;
            tmin = kEps;
            tmax = finf();
            if (MOTION) {
                const int q = A.qIn[i];
                time = (A.P.flags[q] & F_TIME0) ? 0.0f : A.P.time[q];
            }
            return true;
        },
        [&](bool found, const HitRec& h, uint32_t nodes) {
            storeHit(A.Hq, cur, found, h);
            if (COUNT) maxRay = max(maxRay, nodes);
        });
    if (COUNT) flushCounters(A.ctr, cnt, st.overflow, false, maxRay, bad);
    else if (st.overflow) atomicAdd(&A.ctr->stackOverflow, 1ull);
}

struct PathLocal {
    V3 o, d, tp, L, last;
    float pdfPrev;
    float time;  // the camera sample's ray time
    Rng rng;
    uint32_t c0, c1, flags;
    int job;
};
// the time of the path's current ray
GDEV_INLINE float rayTime(const PathLocal& s) { return (s.flags & F_TIME0) ? 0.0f : s.time; }

GDEV_INLINE void loadPath(const PathSoA& P, int p, PathLocal& s) {
    s.tp = V3{P.tpr[p], P.tpg[p], P.tpb[p]};
    s.L = V3{P.Lr[p], P.Lg[p], P.Lb[p]};
    s.last = V3{P.lx[p], P.ly[p], P.lz[p]};
    s.pdfPrev = P.pdfPrev[p];
    s.time = P.time[p];
    s.rng.s = P.rng[p];
    s.c0 = P.c0[p];
    s.c1 = P.c1[p];
    s.flags = P.flags[p];
    s.job = P.job[p];
}
GDEV_INLINE void storePath(const PathSoA& P, int p, const PathLocal& s) {
    P.tpr[p] = s.tp.x; P.tpg[p] = s.tp.y; P.tpb[p] = s.tp.z;
    P.Lr[p] = s.L.x; P.Lg[p] = s.L.y; P.Lb[p] = s.L.z;
    P.lx[p] = s.last.x; P.ly[p] = s.last.y; P.lz[p] = s.last.z;
    P.pdfPrev[p] = s.pdfPrev;
    P.rng[p] = s.rng.s;
    P.c0[p] = s.c0;
    P.c1[p] = s.c1;
    P.flags[p] = s.flags;
}

// Event kinds between NEE and scatter
constexpr int EV_NONE = 0, EV_MEDIUM = 1, EV_SURFACE = 2;

struct Event {
    int kind;
    V3 p, n, dpdu;
    float u, v;
    int prim;
    int curMedium;
};

GDEV_INLINE int nextMedium(V3 wo, V3 wi, int cur, V3 n, int inside, int outside) {  // volPath.cpp:293-298
    if (dot(wi, n) > 0.0f) return dot(wo, n) > 0.0f ? cur : outside;
    return dot(wo, n) > 0.0f ? inside : cur;
}

// Scatter half of a bounce: phase or BSDF sampling, Russian roulette, next ray
// (path.cpp:159-177, volPath.cpp:88-111 and 222-248).  Returns true if alive.
// What a shading-kernel variant handles (its template argument FEAT):
//  kFeatAll  every integrator, AOVs, every emitter kind;
//  kFeatPath path-integrator jobs without AOVs on a scene whose emitters are
//            all area emitters (no environment, no other lights);
//  kFeatVol  the same with volpath jobs too (media).
// The leaner forms leave the other branches out of the code and of its
// register budget.
constexpr int kFeatAll = 0, kFeatPath = 1, kFeatVol = 2;

template <int FEAT = kFeatAll>
GDEV_INLINE bool scatterStep(const DScene& S, const DJob& J, PathLocal& s, const Event& ev) {
    const bool vol = FEAT != kFeatPath && J.integrator == GND_INTEGRATOR_VOLPATH;
    uint32_t bounces = s.c0 & 0xFFFFu, nullB = s.c0 >> 16;
    uint32_t surfB = s.c1 & 0xFFFFu, volB = s.c1 >> 16;
    if (FEAT != kFeatPath && ev.kind == EV_MEDIUM) {
        const DMedium& m = S.media[ev.curMedium];
        V3 wo = -s.d, wi;
        float ux = s.rng.next1D(), uy = s.rng.next1D();
        s.pdfPrev = hgSample(hgClamp(m.g), wo, wi, ux, uy);
        s.flags &= ~F_SPECULAR;
        s.last = ev.p;
        s.o = ev.p;
        s.d = wi;  // medium unchanged
        if (volB > 6) {
            float q = stdMax(0.05f, stdMin(luminance(s.tp), 0.99f));
            if (s.rng.next1D() > q) return false;
            s.tp = cdiv(s.tp, q);
        }
        bounces++;
        volB++;
        s.c0 = bounces | (nullB << 16);
        s.c1 = surfB | (volB << 16);
        return (int)bounces < J.maxDepth;
    }
    // surface
    Bsdf b;
    b.fr.x = ev.dpdu;
    b.fr.y = normalize(cross(ev.n, ev.dpdu));
    b.fr.z = ev.n;
    makeLobe(S.mats + S.prims[ev.prim].material, ev.u, ev.v, b.lobe);
    V3 wiW = V3{0, 0, 0};
    int st = 0;
    float sx = s.rng.next1D(), sy = s.rng.next1D();
    float uc = s.rng.next1D();
    V3 fc = b.sample(-s.d, wiW, sx, sy, uc, s.pdfPrev, st);
    if (isBlack(fc) || s.pdfPrev <= 1e-6f) return false;
    if (st & BSDF_SPECULAR) s.flags |= F_SPECULAR;
    else s.flags &= ~F_SPECULAR;
    s.tp = s.tp * fc;
    if (!vol) {
        if (bounces > 2) {
            float q = stdMax(0.05f, stdMin(luminance(s.tp), 0.99f));
            if (s.rng.next1D() > q) return false;
            s.tp = cdiv(s.tp, q);
        }
        s.last = ev.p;
        s.o = ev.p;
        s.d = wiW;
        bounces++;
        s.c0 = bounces | (nullB << 16);
        return (int)bounces < J.maxDepth;
    }
    V3 inc = -s.d;
    s.last = ev.p;
    int inside, outside;
    mediumInterface(S, ev.prim, inside, outside);
    int nm = nextMedium(inc, wiW, ev.curMedium, ev.n, inside, outside);
    s.o = ev.p;
    s.d = wiW;
    s.flags = setMedium(s.flags, nm);
    if (surfB > 2) {
        float q = stdMax(0.05f, stdMin(luminance(s.tp), 0.99f));
        if (s.rng.next1D() > q) return false;
        s.tp = cdiv(s.tp, q);
    }
    bounces++;
    surfB++;
    s.c0 = bounces | (nullB << 16);
    s.c1 = surfB | (volB << 16);
    return (int)bounces < J.maxDepth;
}

// The integrators' G-buffer writes (path.cpp:35-61 and 99-106, volPath.cpp:
// 117-146 and 184-191) into slot p's AOV record.  A miss: depth / roughness /
// metallic reset at bounce 0; albedo = the environment along the ray (black
// without one), normal 0.
GDEV_INLINE void aovPut(const PathSoA& P, int p, int k, float v) { P.aov[(size_t)k * P.aovStride + p] = v; }
GDEV_INLINE void aovMiss(const PathSoA& P, int p, PathLocal& s, uint32_t bounces, V3 envAlbedo) {
    if (bounces == 0) {
        aovPut(P, p, 6, -1.0f);
        aovPut(P, p, 7, 0.0f);
        aovPut(P, p, 8, 0.0f);
    }
    if (s.flags & F_GBUF) {
        aovPut(P, p, 0, envAlbedo.x);
        aovPut(P, p, 1, envAlbedo.y);
        aovPut(P, p, 2, envAlbedo.z);
#pragma unroll
        for (int k = 3; k < 6; ++k) aovPut(P, p, k, 0.0f);
        s.flags &= ~F_GBUF;
    }
}
// A surface with a BSDF: depth along Camera::getForward (camera.h:40-43),
// roughness and metallic (Material::getRoughness / getMetallic: conductor and
// glass roughness, conductor metallic 1) at bounce 0; albedo
// (Material::getAlbedo: matte albedo texture, conductor R) and the shading
// normal mapped to [0,1] at the first surface with a non-specular lobe, or at
// the last bounce.
GDEV_INLINE bool lobeNonSpecular(int kind) {
    return kind == L_LAMBERT || kind == L_MF_REFL || kind == L_MF_FRESNEL || kind == L_ROUGH_PLASTIC;
}
GDEV_INLINE void aovSurface(const WaveArgs& A, const DJob& J, int p, PathLocal& s, const Isect& it, const Bsdf& b,
                            uint32_t bounces) {
    const DMaterial& m = A.S.mats[A.S.prims[it.prim].material];
    if (bounces == 0) {
        const V3 fwd = xfVector(J.cam.toWorld, V3{0.0f, 0.0f, -1.0f});
        aovPut(A.P, p, 6, dot(it.p - s.last, fwd));
        aovPut(A.P, p, 7, (m.type == GND_MAT_CONDUCTOR || m.type == GND_MAT_GLASS) ? m.roughness : 0.0f);
        aovPut(A.P, p, 8, m.type == GND_MAT_CONDUCTOR ? 1.0f : 0.0f);
    }
    if ((s.flags & F_GBUF) && (lobeNonSpecular(b.lobe.kind) || (int)bounces == J.maxDepth - 1)) {
        const V3 alb = m.type == GND_MAT_MATTE ? b.lobe.albedo : (m.type == GND_MAT_CONDUCTOR ? c3(m.R) : sameV(0.0f));
        const V3 nn = V3{it.n.x / 2.0f, it.n.y / 2.0f, it.n.z / 2.0f} + sameV(0.5f);
        aovPut(A.P, p, 0, alb.x);
        aovPut(A.P, p, 1, alb.y);
        aovPut(A.P, p, 2, alb.z);
        aovPut(A.P, p, 3, nn.x);
        aovPut(A.P, p, 4, nn.y);
        aovPut(A.P, p, 5, nn.z);
        s.flags &= ~F_GBUF;
    }
}

// A shadow ray's time: the reference interaction's, except a point light's,
// whose shadow ray takes the time of the fresh light interaction, 0
// (pointlight.cpp:29)
GDEV_INLINE float shadowTime(const DScene& S, int e, float t) {
    return (e < 0 && S.lights[-e - 1].type == GND_LIGHT_POINT) ? 0.0f : t;
}

// First half of a bounce: medium sampling (volpath), emission with MIS, null
// BSDF pass-through and the NEE light sample with its shadow ray.
// Returns: 0 = dead, 1 = continues straight (already a new ray), 2 = needs scatter.
// pre: the medium was a density grid that kMedium already walked (throughput
// and sampler state updated); preHit / preP = its collision, if any.
template <int FEAT = kFeatAll>
GDEV_INLINE int neeStep(const WaveArgs& A, const DJob& J, int p, PathLocal& s, const HitRec& h, Event& ev,
                        bool& shadow, V3& so, V3& sd, float& stmax, int& smedium, V3& pending, V3& lightP,
                        float& stime, bool pre, bool preHit, V3 preP) {
    const DScene& S = A.S;
    const bool vol = FEAT != kFeatPath && J.integrator == GND_INTEGRATOR_VOLPATH;
    const bool hit = h.inst >= 0;
    const float invNE = 1.0f / (float)(uint32_t)S.nEmitters;
    shadow = false;
    uint32_t bounces = s.c0 & 0xFFFFu, nullB = s.c0 >> 16;
    const int medium = vol ? flagMedium(s.flags) : -1;
    if (medium >= 0) {
        MediumEvent me;
        if (pre) {
            me.valid = preHit;
            me.p = preP;
        } else {
            const float tm = hit ? h.t : finf();
            s.tp = s.tp * mediumSample(S.media[medium], s.o, s.d, kEps, tm, s.rng, me);
            if (me.valid) s.tp = s.tp * me.sigma_s;
        }
        if (me.valid) {
            if (isBlack(s.tp)) return 0;
            const float lsp = 1.0f / (float)(uint32_t)S.nEmitters;
            const int ep = pickEmitter(S, s.rng.next1D());
            const float ux = s.rng.next1D(), uy = s.rng.next1D();
            LightSample ls;
            bool delta;
            // the light is sampled from a fresh SurfaceInteraction holding
            // only the point (volPath.cpp:60-61): time 0
            emitterSample<FEAT != kFeatAll>(S, ep, me.p, ux, uy, ls, delta, 0.0f);
            const V3 Li = cdiv(ls.Li, lsp);
            if (ls.ok && ls.pdf > 1e-6f && !isBlack(Li)) {
                const V3 wi = normalize(ls.p - me.p);
                const float pp = hgPhase(dot(-s.d, wi), hgClamp(S.media[medium].g));
                pending = pp > 1e-6f ? s.tp * pp * Li * powerHeuristic(ls.pdf * lsp, delta ? 0.0f : pp) : sameV(0.0f);
                shadow = true;
                so = me.p; sd = ls.dir; stmax = ls.tmax; smedium = medium; lightP = ls.p;
                stime = 0.0f;
            }
            ev.kind = EV_MEDIUM;
            ev.p = me.p;
            ev.curMedium = medium;
            return 2;
        }
    }
    const bool aovOn = FEAT == kFeatAll && J.filmAov && A.P.aov;
    if (!hit) {
        // the environment along the escaping ray, MIS-weighted against NEE
        // unless the last bounce was specular (path.cpp:36-55, volPath.cpp:
        // 117-139); its pdf is taken from the origin, the reference point of
        // the default SurfaceInteraction there
        V3 env = sameV(0.0f);
        if (FEAT == kFeatAll && S.envLight >= 0) {
            env = envEval(S, s.d);
            if (!isBlack(env)) {
                float w = 1.0f;
                if (!(s.flags & F_SPECULAR)) {
                    const float pe = envPdf(S, sameV(0.0f), s.o + s.d * 1e5f) * invNE;
                    w = powerHeuristic(s.pdfPrev, pe);
                }
                s.L = s.L + s.tp * env * w;
            }
        }
        if (aovOn) aovMiss(A.P, p, s, bounces, env);
        return 0;
    }
    Isect it;
    fillInteraction(S, s.o, s.d, h, it, rayTime(s));
    const DPrim& pr = S.prims[it.prim];
    Bsdf b;
    const bool hasBsdf = makeBsdf(S, it, b);
    if (pr.emitter >= 0) {
        const bool skip = (bounces == 0 && !hasBsdf);
        if (!skip) {
            V3 Le = areaEval(pr, it.n, -s.d);
            if (!isBlack(Le)) {
                float w = 1.0f;
                if (!(s.flags & F_SPECULAR)) {
                    // dummyRef holds only the point (path.cpp:76, volPath.cpp:161): time 0
                    float lp = areaPdf(S, it.prim, s.last, it.p, it.n, 0.0f);
                    w = powerHeuristic(s.pdfPrev, lp * invNE);
                }
                s.L = s.L + s.tp * Le * w;
            }
        }
    }
    int inside = -1, outside = -1;
    if (vol) mediumInterface(S, it.prim, inside, outside);
    if (!hasBsdf) {
        if (nullB++ > 100) return 0;
        s.c0 = bounces | (nullB << 16);
        s.o = it.p;  // direction unchanged
        if (vol) s.flags = setMedium(s.flags, dot(s.d, it.n) > 0.0f ? outside : inside);
        return 1;
    }
    if (aovOn) aovSurface(A, J, p, s, it, b, bounces);
    const float lsp = 1.0f / (float)(uint32_t)S.nEmitters;
    const int ep = pickEmitter(S, s.rng.next1D());
    const float ux = s.rng.next1D(), uy = s.rng.next1D();
    LightSample ls;
    bool delta;
    emitterSample<FEAT != kFeatAll>(S, ep, it.p, ux, uy, ls, delta, rayTime(s));
    const V3 Li = cdiv(ls.Li, lsp);
    if (ls.ok && ls.pdf > 1e-6f && !isBlack(Li)) {
        const V3 wi = normalize(ls.p - it.p);
        const float bp = b.pdf(-s.d, wi);
        const V3 f = b.f(-s.d, wi);
        stime = shadowTime(S, ep, rayTime(s));
        if (!isBlack(f) && bp > 1e-6f)
            pending = s.tp * f * Li * __builtin_fabsf(dot(it.n, wi)) * powerHeuristic(ls.pdf * lsp, delta ? 0.0f : bp);
        else pending = sameV(0.0f);
        shadow = true;
        so = it.p; sd = ls.dir; stmax = ls.tmax; lightP = ls.p;
        smedium = vol ? nextMedium(-s.d, ls.dir, medium, it.n, inside, outside) : -1;
    }
    ev.kind = EV_SURFACE;
    ev.p = it.p;
    ev.n = it.n;
    ev.dpdu = it.dpdu;
    ev.u = it.u;
    ev.v = it.v;
    ev.prim = it.prim;
    ev.curMedium = medium;
    return 2;
}

// DirectLight::Li (src/integrators/direct.cpp:10-76) as two bounces of the
// wavefront loop.  Bounce 0, the camera hit: its emission, one light sample
// (balance-heuristic MIS, weight 1 for delta lights) whose shadow ray is a
// plain any-hit test, and one BSDF sample that becomes the path's next ray.
// Bounce 1, that ray's hit: its emission, MIS-weighted against the light
// pdf.  Between the two the sample's weight f*|cos|/pdf rides in the
// throughput, its pdf in pdfPrev, and F_SPECULAR holds "sampled type ==
// BSDF_SPECULAR" (direct.cpp:66: an equality test, false for every lobe
// here since each also carries a reflection or transmission bit).
// Returns 0 = done, 1 = continues with the BSDF ray.
GDEV_INLINE int directStep(const WaveArgs& A, PathLocal& s, const HitRec& h, bool& shadow, V3& so, V3& sd,
                           float& stmax, V3& pending, V3& lightP, float& stime) {
    const DScene& S = A.S;
    shadow = false;
    if (h.inst < 0) return 0;  // no environment term in this integrator
    Isect it;
    fillInteraction(S, s.o, s.d, h, it, rayTime(s));  // the BSDF ray has time 0 (direct.cpp:60)
    const DPrim& pr = S.prims[it.prim];
    const float lsp = 1.0f / (float)(uint32_t)S.nEmitters;
    if ((s.c0 & 0xFFFFu) > 0) {
        if (pr.emitter >= 0) {
            const V3 Lb = areaEval(pr, it.n, -s.d);
            if (!isBlack(Lb)) {
                float lpg = areaPdf(S, it.prim, s.last, it.p, it.n, s.time);  // pdf(isect, ...)
                if (__builtin_isinf(lpg)) lpg = 0.0f;
                const float pl = (s.flags & F_SPECULAR) ? 0.0f : lpg * lsp;
                const float den = s.pdfPrev + pl;
                if (den > 1e-6f) s.L = s.L + cdiv(s.tp * Lb * s.pdfPrev, den);
            }
        }
        return 0;
    }
    if (pr.emitter >= 0) s.L = s.L + areaEval(pr, it.n, -s.d);
    Bsdf b;
    if (!makeBsdf(S, it, b)) return 0;
    const int ep = pickEmitter(S, s.rng.next1D());
    const float ux = s.rng.next1D(), uy = s.rng.next1D();
    LightSample ls;
    bool delta;
    emitterSample(S, ep, it.p, ux, uy, ls, delta, s.time);
    const V3 Li = cdiv(ls.Li, lsp);
    if (ls.ok && ls.pdf > 1e-6f && !isBlack(Li)) {
        stime = shadowTime(S, ep, s.time);
        // the reference tests the shadow ray first and then weighs; the
        // weight draws no random numbers, so the ray is only cast when the
        // contribution is non-zero
        const V3 wi = normalize(ls.p - it.p);
        const float bp = b.pdf(-s.d, wi);
        const V3 f = b.f(-s.d, wi);
        if (!isBlack(f) && bp > 1e-6f) {
            const float pl = ls.pdf * lsp;
            const float pb = delta ? 0.0f : bp;
            const float den = pl + pb;
            if (den > 1e-6f) {
                pending = cdiv(f * Li * __builtin_fabsf(dot(it.n, wi)) * pl, den);
                shadow = true;
                so = it.p; sd = ls.dir; stmax = ls.tmax; lightP = ls.p;
            }
        }
    }
    V3 wiW = V3{0, 0, 0};
    int st = 0;
    const float sx = s.rng.next1D(), sy = s.rng.next1D();
    const float uc = s.rng.next1D();
    float bpdf = 0.0f;
    const V3 fc = b.sample(-s.d, wiW, sx, sy, uc, bpdf, st);
    if (isBlack(fc) || !(bpdf > 1e-6f)) return 0;
    s.tp = fc;
    s.pdfPrev = bpdf;
    if (st == BSDF_SPECULAR) s.flags |= F_SPECULAR;
    else s.flags &= ~F_SPECULAR;
    s.last = it.p;
    s.o = it.p;
    s.d = wiW;
    s.c0 = 1u;
    s.flags |= F_TIME0;
    return 1;
}

// AmbientOcclusion::Li (src/integrators/ao.cpp:14-40): a camera ray that
// misses is white; a hit casts one cosine-weighted ray of length aoLength
// about the shading normal (frame from a fixed helper axis), white when
// nothing is hit.  The occlusion ray goes through the shadow queue with a
// contribution of 1.
GDEV_INLINE void aoStep(const WaveArgs& A, const DJob& J, PathLocal& s, const HitRec& h, bool& shadow, V3& so,
                        V3& sd, float& stmax, V3& pending, V3& lightP, float& stime) {
    shadow = false;
    if (h.inst < 0) {
        s.L = sameV(1.0f);
        return;
    }
    Isect it;
    fillInteraction(A.S, s.o, s.d, h, it, rayTime(s));
    const float ux = s.rng.next1D(), uy = s.rng.next1D();
    const V3 ld = squareToCosineHemisphere(ux, uy);
    const V3 n = it.n;
    const V3 t1 = __builtin_fabsf(n.x) > 0.1f ? normalize(cross(V3{0.0f, 1.0f, 0.0f}, n))
                                              : normalize(cross(V3{1.0f, 0.0f, 0.0f}, n));
    const V3 t2 = cross(n, t1);
    sd = t1 * ld.x + t2 * ld.y + n * ld.z;
    so = it.p;
    stmax = J.aoLength;
    stime = 0.0f;  // Ray(isect.p, dir, length) (ao.cpp:34)
    lightP = it.p;
    pending = sameV(1.0f);
    shadow = true;
}

// Paths whose ray starts in a density grid, when kMedium runs (split mode).
GDEV_INLINE bool gridWalked(const WaveArgs& A, const DJob& J, uint32_t flags) {
    const int medium = J.integrator == GND_INTEGRATOR_VOLPATH ? flagMedium(flags) : -1;
    return A.split && medium >= 0 && A.S.media[medium].type == GND_MEDIUM_GRID;
}

// Whether this bounce's shadow ray of a volpath job may walk a density grid
// (ratio tracking in kShadow draws from the path's sampler between the NEE
// sample and the scatter).  When it cannot, nothing draws in between and the
// scatter runs in kShade, with the same draws in the same order.
GDEV_INLINE bool shadowMayWalk(const WaveArgs& A, const DJob& J, bool shadow, V3 so, V3 sd, float stmax) {
    if (!shadow || J.integrator != GND_INTEGRATOR_VOLPATH) return false;
    for (int m = 0; m < A.S.nMedia; ++m) {
        const DMedium& md = A.S.media[m];
        if (md.type == GND_MEDIUM_GRID && lineMayCrossGrid(md, so, sd, stmax)) return true;
    }
    return false;
}

// Delta tracking (Medium::sample, heterogeneous.cpp:251-376) for the paths of
// the in-queue whose ray starts in a density grid (the list gIn, built by the
// shading kernels of the bounce before), ahead of kShade: the first sampler
// use of a volpath bounce (volPath.cpp:40-48).  Persistent, one walk step per
// lane per iteration; a lane whose walk ends takes the next path, so a wave
// does not run at the pace of its longest walk.  Writes the throughput (times
// the 1/sigma_t weight and, at a collision, sigma_s) and the sampler state;
// evKind / evp* = whether and where the walk collided.
__global__ __launch_bounds__(256) void kMedium(WaveArgs A) {
    const uint32_t n = A.counts[4];
    const uint64_t waveMask = __ballot(1);
    const DScene& S = A.S;
    RayFeed feed;
    feed.init(A.counts + kFeedMedium, n, kWalkRefill, kWalkFeedBlock);
    bool drained = false;
    int64_t cur = -1;
    int p = -1, medium = 0;
    Rng rng;
    GridTrack<true> k;
    for (;;) {
        if (!drained) {
            const int64_t j = feed.take(cur < 0, waveMask, drained);
            if (j >= 0) {
                const int32_t i = A.gIn[j];
                const int q = A.qIn[i];
                medium = flagMedium(A.P.flags[q]);
                V3 o, d;
                loadRay(A.rIn, (uint32_t)i, o, d);
                const uint4 hr = A.Hq.rec[2 * (int64_t)i];
                const float tm = (int32_t)hr.w >= 0 ? __uint_as_float(hr.x) : finf();
                p = q;
                rng.s = A.P.rng[q];
                cur = i;
                // a ray that misses the grid box leaves the path as it is
                if (!k.start(S.media[medium], o, d, kEps, tm)) {
                    A.P.evKind[q] = EV_NONE;
                    cur = -1;
                }
            }
        }
        const bool busy = cur >= 0;
        if (__ballot(busy) == 0) {
            if (drained) break;
            continue;
        }
        if (busy && k.step(S.media[medium], rng)) {
            const DMedium& m = S.media[medium];
            V3 tp = V3{A.P.tpr[p], A.P.tpg[p], A.P.tpb[p]};
            if (k.hit) {
                V3 o, d;
                loadRay(A.rIn, (uint32_t)cur, o, d);
                MediumEvent me;
                gridEvent(m, o, d, k, me);
                tp = tp * gridWeight(m, k);
                tp = tp * me.sigma_s;
                A.P.evpx[p] = me.p.x; A.P.evpy[p] = me.p.y; A.P.evpz[p] = me.p.z;
            } else {
                tp = tp * sameV(1.0f);
            }
            A.P.tpr[p] = tp.x; A.P.tpg[p] = tp.y; A.P.tpb[p] = tp.z;
            A.P.rng[p] = rng.s;
            A.P.evKind[p] = k.hit ? EV_MEDIUM : EV_NONE;
            cur = -1;
        }
    }
}

// SPLIT = false: full bounce (NEE + scatter) in one kernel; shadow rays resolve
// afterwards (valid when shadow rays consume no random numbers).
// SPLIT = true: NEE only; the event is stored for kScatter.
// FEAT (see kFeatAll): the kFeatPath form's smaller register budget fits
// GND_SHADE_LEAN_WAVES waves per SIMD (1080p scene: 13.7 -> 12.7 ms of
// shading per frame at 3 waves, 13.2 at 2).
#ifndef GND_SHADE_LEAN_WAVES
#define GND_SHADE_LEAN_WAVES 3
#endif
#ifndef GND_SHADE_VOL_WAVES
#define GND_SHADE_VOL_WAVES 2
#endif
template <bool SPLIT, int FEAT = kFeatAll>
__global__ __launch_bounds__(kShadeBlock, FEAT == kFeatPath  ? GND_SHADE_LEAN_WAVES
                                          : FEAT == kFeatVol ? GND_SHADE_VOL_WAVES
                                                             : GND_SHADE_WAVES) void kShade(WaveArgs A) {
    static_assert(!(SPLIT && FEAT == kFeatPath), "the split kernels serve volumetric jobs");
    const uint32_t n = A.counts[0];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t start = blockIdx.x * blockDim.x;
    for (uint32_t base = start; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        const bool active = i < n;
        bool alive = false, shadow = false, ending = false, toGrid = false;
        int p = -1;
        PathLocal s;
        V3 so, sd, pending = sameV(0.0f), lightP;
        float stmax = 0.0f, stime = 0.0f;
        int smedium = -1;
        if (active) {
            p = A.qIn[i];
            loadPath(A.P, p, s);
            loadRay(A.rIn, i, s.o, s.d);
            const DJob& J = A.jobs[s.job];
            const HitRec h = loadHit(A.Hq, i);
            Event ev;
            ev.kind = EV_NONE;
            bool pre = false, preHit = false;
            V3 preP = sameV(0.0f);
            if (SPLIT && gridWalked(A, J, s.flags)) {
                pre = true;
                preHit = A.P.evKind[p] == EV_MEDIUM;
                if (preHit) preP = V3{A.P.evpx[p], A.P.evpy[p], A.P.evpz[p]};
            }
            int r;
            if (FEAT == kFeatAll && J.integrator == GND_INTEGRATOR_DIRECT) {
                r = directStep(A, s, h, shadow, so, sd, stmax, pending, lightP, stime);
            } else if (FEAT == kFeatAll && J.integrator == GND_INTEGRATOR_AO) {
                aoStep(A, J, s, h, shadow, so, sd, stmax, pending, lightP, stime);
                r = 0;
            } else {
                r = neeStep<FEAT>(A, J, p, s, h, ev, shadow, so, sd, stmax, smedium, pending, lightP, stime, pre,
                                  preHit, preP);
            }
            if (!A.pathOnlyFast && shadow) {
                // the general shadow kernel always needs the ray; the any-hit
                // path skips zero contributions
            } else if (shadow && isBlack(pending)) {
                shadow = false;
            }
            // split mode: the scatter waits for kScatter only when the shadow
            // ray may draw from the sampler first
            bool deferred = false;
            if (r == 1) alive = true;
            else if (r == 2) {
                if (SPLIT && shadowMayWalk(A, J, shadow, so, sd, stmax)) {
                    deferred = true;
                    alive = false;
                    A.P.evKind[p] = ev.kind;
                    A.P.evPrim[p] = ev.prim;
                    A.P.evMedium[p] = ev.curMedium;
                    A.P.evpx[p] = ev.p.x; A.P.evpy[p] = ev.p.y; A.P.evpz[p] = ev.p.z;
                    A.P.evnx[p] = ev.n.x; A.P.evny[p] = ev.n.y; A.P.evnz[p] = ev.n.z;
                    A.P.evtx[p] = ev.dpdu.x; A.P.evty[p] = ev.dpdu.y; A.P.evtz[p] = ev.dpdu.z;
                    A.P.evu[p] = ev.u; A.P.evv[p] = ev.v;
                } else {
                    alive = scatterStep<FEAT>(A.S, J, s, ev);
                }
            }
            if (SPLIT && !deferred) A.P.evKind[p] = EV_NONE;
            toGrid = SPLIT && alive && gridWalked(A, J, s.flags);
            storePath(A.P, p, s);
            ending = !alive && !deferred;
        }
        retirePaths(A.chunkLive, ending, ending && A.chunkLive ? A.P.chunk[p] : -1);
        uint32_t qs, ss, gs;
        blockAppend3(&A.counts[1], alive, qs, &A.counts[2], shadow, ss, SPLIT ? &A.counts[3] : nullptr, toGrid, gs);
        if (alive) {
            A.qOut[qs] = p;
            storeRay(A.rOut, qs, s.o, s.d);
        }
        if (toGrid) A.gOut[gs] = (int32_t)qs;
        if (shadow) {
            const ShadowSoA& Sh = A.Sh;
            Sh.path[ss] = p;
            Sh.medium[ss] = smedium;
            Sh.ox[ss] = so.x; Sh.oy[ss] = so.y; Sh.oz[ss] = so.z;
            Sh.dx[ss] = sd.x; Sh.dy[ss] = sd.y; Sh.dz[ss] = sd.z;
            Sh.tmax[ss] = stmax;
            Sh.time[ss] = stime;
            Sh.pr[ss] = pending.x; Sh.pg[ss] = pending.y; Sh.pb[ss] = pending.z;
            Sh.lpx[ss] = lightP.x; Sh.lpy[ss] = lightP.y; Sh.lpz[ss] = lightP.z;
        }
    }
}

// Split mode second half: the paths that recorded an event scatter now (after
// their shadow rays consumed transmittance samples).
template <int FEAT = kFeatAll>
__global__ __launch_bounds__(kShadeBlock) void kScatter(WaveArgs A) {
    const uint32_t nEv = A.counts[0];
    const int32_t* qEv = A.qIn;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < nEv; base += stride) {
        const uint32_t i = base + threadIdx.x;
        bool alive = false, ending = false, toGrid = false;
        int p = -1;
        PathLocal s;
        if (i < nEv) {
            p = qEv[i];
            const int kind = A.P.evKind[p];
            if (kind != EV_NONE) {
                loadPath(A.P, p, s);
                loadRay(A.rIn, i, s.o, s.d);
                Event ev;
                ev.kind = kind;
                ev.prim = A.P.evPrim[p];
                ev.curMedium = A.P.evMedium[p];
                ev.p = V3{A.P.evpx[p], A.P.evpy[p], A.P.evpz[p]};
                ev.n = V3{A.P.evnx[p], A.P.evny[p], A.P.evnz[p]};
                ev.dpdu = V3{A.P.evtx[p], A.P.evty[p], A.P.evtz[p]};
                ev.u = A.P.evu[p];
                ev.v = A.P.evv[p];
                alive = scatterStep<FEAT>(A.S, A.jobs[s.job], s, ev);
                toGrid = alive && gridWalked(A, A.jobs[s.job], s.flags);
                storePath(A.P, p, s);
                ending = !alive;
            }
        }
        retirePaths(A.chunkLive, ending, ending && A.chunkLive ? A.P.chunk[p] : -1);
        uint32_t qs, unused, gs;
        blockAppend3(&A.counts[1], alive, qs, nullptr, false, unused, &A.counts[3], toGrid, gs);
        if (alive) {
            A.qOut[qs] = p;
            storeRay(A.rOut, qs, s.o, s.d);
        }
        if (toGrid) A.gOut[gs] = (int32_t)qs;
    }
}

// Shadow rays when no surface is a null BSDF and no segment can cross a
// medium: one any-hit test each, persistent with refill like kTracePaths.
template <int LDSD, bool COUNT, bool MOTION>
__global__ __launch_bounds__(256, GND_SHADOW_WAVES) void kShadowAny(WaveArgs A) {
    extern __shared__ int sStack[];
    const uint32_t n = A.counts[2];
    uint32_t cnt[5] = {0, 0, 0, 0, 0};
    int spill[spillSize<LDSD>()];
    TravStack<LDSD> st = makeStack<LDSD>(sStack, spill, A.stackCap);
    const DScene& S = A.S;
    const ShadowSoA& Sh = A.Sh;
    const uint64_t waveMask = __ballot(1);
    RayFeed feed;
    feed.init(A.counts + kFeedShadow, n, A.srefill, kShadowFeedBlock);
    int64_t cur = -1;
    traceStream<true, COUNT, LDSD, MOTION>(
        S, st, cnt,
        [&](bool idle, V3& o, V3& d, float& tmin, float& tmax, float& time, bool& drained) {
            const int64_t i = feed.take(idle, waveMask, drained);
            if (i < 0) return false;
            cur = i;
            o = V3{Sh.ox[i], Sh.oy[i], Sh.oz[i]};
            d = V3{Sh.dx[i], Sh.dy[i], Sh.dz[i]};
            tmin = kEps;
            tmax = Sh.tmax[i];
            if (MOTION) time = Sh.time[i];
            return true;
        },
        [&](bool occluded, const HitRec&, uint32_t) {
            if (occluded) return;
            const int p = Sh.path[cur];
            A.P.Lr[p] = A.P.Lr[p] + Sh.pr[cur];
            A.P.Lg[p] = A.P.Lg[p] + Sh.pg[cur];
            A.P.Lb[p] = A.P.Lb[p] + Sh.pb[cur];
        });
    if (COUNT) flushCounters(A.ctr, cnt, st.overflow, true);
    else if (st.overflow) atomicAdd(&A.ctr->stackOverflow, 1ull);
}

// General shadow rays: closest-hit loop through null-BSDF surfaces with
// per-segment medium transmittance (path.cpp:122-140 / volPath.cpp:261-291).
// Persistent: each lane runs one shadow ray as a small state machine (trace
// the next segment, then walk a density grid one ratio-tracking step at a
// time) and takes the next ray when it ends.  An iteration either traverses
// or steps the walks, whichever more lanes wait for: a traversal costs many
// walk steps, so doing both every iteration would leave the walks crawling.
template <int LDSD, bool COUNT, bool MOTION>
__global__ __launch_bounds__(256, GND_SHADOWTR_WAVES) void kShadow(WaveArgs A) {
    extern __shared__ int sStack[];
    const uint32_t n = A.counts[2];
    uint32_t cnt[5] = {0, 0, 0, 0, 0};
    int spill[spillSize<LDSD>()];
    TravStack<LDSD> st = makeStack<LDSD>(sStack, spill, A.stackCap);
    const DScene& S = A.S;
    const ShadowSoA& Sh = A.Sh;
    const uint64_t waveMask = __ballot(1);
    RayFeed feed;
    feed.init(A.counts + kFeedShadow, n, kWalkRefill, kWalkFeedBlock);
    bool drained = false;
    constexpr int kIdle = 0, kTrace = 1, kWalk = 2;
    constexpr int kVisible = 0, kOccluded = 1, kOnward = 2;
    int64_t cur = -1;
    int phase = kIdle, medium = -1, iter = 0, p = -1;
    bool vol = false, opaque = false;  // opaque: direct / ao, any surface occludes
    float stime = 0.0f;                // the shadow ray's time (constant along its segments)
    V3 o = sameV(0.0f), d = sameV(0.0f), Tr = sameV(1.0f);
    float tmax = 0.0f;
    Rng rng;
    // what follows the segment in flight: visible, occluded, or onward from
    // the null surface at `no` into `nmedium` up to `ntmax`
    int next = kVisible, nmedium = -1;
    V3 no = sameV(0.0f);
    float ntmax = 0.0f;
    GridTrack<false> k;
    for (;;) {
        if (!drained) {
            const int64_t i = feed.take(phase == kIdle, waveMask, drained);
            if (i >= 0) {
                cur = i;
                p = Sh.path[i];
                o = V3{Sh.ox[i], Sh.oy[i], Sh.oz[i]};
                d = V3{Sh.dx[i], Sh.dy[i], Sh.dz[i]};
                tmax = Sh.tmax[i];
                stime = S.hasMotion ? Sh.time[i] : 0.0f;
                {
                    const int32_t integ = A.jobs[A.P.job[p]].integrator;
                    vol = integ == GND_INTEGRATOR_VOLPATH;
                    opaque = integ == GND_INTEGRATOR_DIRECT || integ == GND_INTEGRATOR_AO;
                }
                medium = Sh.medium[i];
                Tr = sameV(1.0f);
                rng.s = A.P.rng[p];
                iter = 0;
                phase = kTrace;
            }
        }
        const int nTrace = __popcll(__ballot(phase == kTrace)), nWalk = __popcll(__ballot(phase == kWalk));
        if (nTrace + nWalk == 0) {
            if (drained) break;
            continue;
        }
        // one traversal or one walk step per iteration (kShadowTraceRatio)
        const bool tracing = nTrace > nWalk * kShadowTraceRatio;
        bool segDone = false;
        if (tracing && phase == kTrace) {
            HitRec h;
            const bool hit = traceRay<false, COUNT, LDSD, MOTION>(S, o, d, kEps, tmax, stime, h, st, cnt);
            if (!hit) {
                next = kVisible;
            } else {
                Isect it;
                fillInteraction(S, o, d, h, it, stime);
                const DMaterial& m = S.mats[S.prims[it.prim].material];
                if (m.type != GND_MAT_INDEX || opaque) {
                    next = kOccluded;
                } else {
                    next = kOnward;
                    if (vol) {
                        int inside, outside;
                        mediumInterface(S, it.prim, inside, outside);
                        nmedium = dot(d, it.n) > 0.0f ? outside : inside;
                        const V3 lightP = V3{Sh.lpx[cur], Sh.lpy[cur], Sh.lpz[cur]};
                        ntmax = length(it.p - lightP) - kEps;
                    } else {
                        nmedium = medium;
                        ntmax = h.t - length(it.p - o);  // the hit distance minus the walked length
                    }
                    no = it.p;
                }
            }
            segDone = true;
            if (vol && medium >= 0) {
                const DMedium& md = S.media[medium];
                const float segT = hit ? h.t : tmax;
                if (md.type == GND_MEDIUM_GRID) {
                    if (k.start(md, o, d, kEps, segT)) {
                        phase = kWalk;
                        segDone = false;
                    } else {
                        Tr = Tr * sameV(1.0f);
                    }
                } else {
                    Tr = Tr * mediumTr(md, o, d, kEps, segT, rng);
                }
            }
        } else if (!tracing && phase == kWalk) {
            if (k.step(S.media[medium], rng)) {
                Tr = Tr * k.transmittance();
                segDone = true;
            }
        }
        if (segDone) {
            bool fin = true, visible = false;
            if (vol && isBlack(Tr)) {
            } else if (next == kVisible) {
                visible = true;
            } else if (next == kOnward && ++iter < 1024) {
                fin = false;
                o = no;
                medium = nmedium;
                tmax = ntmax;
                phase = kTrace;
            }
            if (fin) {
                A.P.rng[p] = rng.s;
                if (visible) {
                    const V3 pending = V3{Sh.pr[cur], Sh.pg[cur], Sh.pb[cur]};
                    const V3 c = vol ? pending * Tr : pending;
                    A.P.Lr[p] = A.P.Lr[p] + c.x;
                    A.P.Lg[p] = A.P.Lg[p] + c.y;
                    A.P.Lb[p] = A.P.Lb[p] + c.z;
                }
                cur = -1;
                phase = kIdle;
            }
        }
    }
    if (COUNT) flushCounters(A.ctr, cnt, st.overflow, true);
    else if (st.overflow) atomicAdd(&A.ctr->stackOverflow, 1ull);
}

// ============================================================================ tail kernel

// The last bounces of a wave, once few paths are left: a launch per bounce
// then lasts as long as its slowest ray, so most of the GPU idles.  kTail runs
// every path of the in-queue to its end in one launch.  Each lane takes a
// path (persistent, refilled from the queue as paths end) and repeats, per
// bounce: closest hit, medium sampling + emission + NEE (neeStep / directStep
// / aoStep), the shadow ray with its transmittance, then the scatter.  That is
// the order of random-number draws of the bounce kernels in split mode
// (kMedium, kShade<true>, kShadow, kScatter), and the radiance sums are formed
// in the same order, so a path gives the same bits whichever way it ran.
// Traversals are wave-collective (traceRay): lanes without a ray take part
// idle.  FAST: every shadow ray is one any-hit query (no null-BSDF surface,
// no medium on a segment).
// Waves per SIMD the tail kernel is compiled for: at 4 its registers are cut
// from about 240 to 128 and some spill, yet more resident walks and
// traversals hide more memory latency (best frame, over the tail threshold:
// 1080p 67.5 ms at 2 waves, 66.4 at 3, 65.9 at 4, 69.1 at 5; volumetric
// 97.6, 95.4, 91.7 at 4 with a 4 M threshold, 101.6 at 5).
#ifndef GND_TAIL_WAVES
#define GND_TAIL_WAVES 4
#endif
// a lane's work between refills is a whole bounce, so idle lanes refill early
#ifndef GND_TAIL_REFILL
#define GND_TAIL_REFILL 8
#endif
constexpr int kTailRefill = GND_TAIL_REFILL;
// paths reserved per atomic by a tail-kernel wave (0: one atomic per refill;
// 1080p frame 57.2 ms, blocks of 32 57.8, 64 57.9, 128 58.4)
#ifndef GND_TAIL_FEED_BLOCK
#define GND_TAIL_FEED_BLOCK 0
#endif
constexpr uint32_t kTailFeedBlock = GND_TAIL_FEED_BLOCK;
// traversal stack entries in LDS per lane of the tail kernel
constexpr int kTailLds = 16;
// FEAT: the jobs and scene the launch serves (see kFeatAll).
template <int LDSD, bool MOTION, bool FAST, int FEAT = kFeatAll>
__global__ __launch_bounds__(256, GND_TAIL_WAVES) void kTail(WaveArgs A) {
    extern __shared__ int sStack[];
    const uint32_t n = A.counts[0];
    const uint64_t waveMask = __ballot(1);
    uint32_t cnt[5] = {0, 0, 0, 0, 0};
    int spill[spillSize<LDSD>()];
    TravStack<LDSD> st = makeStack<LDSD>(sStack, spill, A.stackCap);
    const DScene& S = A.S;
    RayFeed feed;
    feed.init(A.counts + kFeedTail, n, kTailRefill, kTailFeedBlock);
    bool drained = false, busy = false;
    int p = -1;
    PathLocal s;
    uint32_t nRays = 0, nShadow = 0;
    for (;;) {
        if (!drained) {
            const int64_t i = feed.take(!busy, waveMask, drained);
            if (i >= 0) {
                p = A.qIn[i];
                loadPath(A.P, p, s);
                loadRay(A.rIn, (uint32_t)i, s.o, s.d);
                busy = true;
            }
        }
        if (__ballot(busy) == 0) {
            if (drained) break;
            continue;
        }
        // closest hit (Scene::rayIntersect at the top of Li)
        HitRec h;
        h.t = finf();
        h.u = h.v = 0.0f;
        h.inst = h.prim = -1;
        const bool found = traceRay<false, false, LDSD, MOTION>(S, s.o, s.d, kEps, finf(), rayTime(s), h, st, cnt,
                                                                busy);
        if (busy) {
            nRays++;
            if (!found) h.inst = h.prim = -1;
        }
        // emission, NEE and the event of this bounce
        bool shadow = false;
        int r = 0, smedium = -1;
        V3 so = sameV(0.0f), sd = sameV(0.0f), pending = sameV(0.0f), lightP = sameV(0.0f);
        float stmax = 0.0f, stime = 0.0f;
        Event ev;
        ev.kind = EV_NONE;
        const DJob* J = busy ? &A.jobs[s.job] : nullptr;
        if (busy) {
            if (FEAT == kFeatAll && J->integrator == GND_INTEGRATOR_DIRECT) {
                r = directStep(A, s, h, shadow, so, sd, stmax, pending, lightP, stime);
            } else if (FEAT == kFeatAll && J->integrator == GND_INTEGRATOR_AO) {
                aoStep(A, *J, s, h, shadow, so, sd, stmax, pending, lightP, stime);
                r = 0;
            } else {
                r = neeStep<FEAT>(A, *J, p, s, h, ev, shadow, so, sd, stmax, smedium, pending, lightP, stime, false,
                                  false, sameV(0.0f));
            }
            if (FAST && shadow && isBlack(pending)) shadow = false;
            if (shadow) nShadow++;
        }
        // the shadow ray, segment by segment through null-BSDF surfaces
        // (path.cpp:122-140, volPath.cpp:261-291)
        {
            const bool vol = FEAT != kFeatPath && busy && J->integrator == GND_INTEGRATOR_VOLPATH;
            const bool opaque = FEAT == kFeatAll && busy &&
                                (J->integrator == GND_INTEGRATOR_DIRECT || J->integrator == GND_INTEGRATOR_AO);
            bool open = shadow, visible = false;
            V3 Tr = sameV(1.0f), o = so;
            float tmax = stmax;
            int medium = smedium, iter = 0;
            while (__ballot(open)) {
                HitRec sh;
                const bool hit = traceRay<FAST, false, LDSD, MOTION>(S, o, sd, kEps, tmax, stime, sh, st, cnt, open);
                if (!open) continue;
                if (FAST) {
                    visible = !hit;
                    open = false;
                    continue;
                }
                int next = 0, nmedium = -1;  // 0 visible, 1 occluded, 2 onward
                V3 no = sameV(0.0f);
                float ntmax = 0.0f;
                if (hit) {
                    Isect it;
                    fillInteraction(S, o, sd, sh, it, stime);
                    const DMaterial& m = S.mats[S.prims[it.prim].material];
                    if (m.type != GND_MAT_INDEX || opaque) {
                        next = 1;
                    } else {
                        next = 2;
                        if (vol) {
                            int inside, outside;
                            mediumInterface(S, it.prim, inside, outside);
                            nmedium = dot(sd, it.n) > 0.0f ? outside : inside;
                            ntmax = length(it.p - lightP) - kEps;
                        } else {
                            nmedium = medium;
                            ntmax = sh.t - length(it.p - o);
                        }
                        no = it.p;
                    }
                }
                if (vol && medium >= 0) Tr = Tr * mediumTr(S.media[medium], o, sd, kEps, hit ? sh.t : tmax, s.rng);
                if (vol && isBlack(Tr)) {
                    open = false;
                } else if (next == 0) {
                    visible = true;
                    open = false;
                } else if (next == 2 && ++iter < 1024) {
                    o = no;
                    medium = nmedium;
                    tmax = ntmax;
                } else {
                    open = false;
                }
            }
            if (visible) s.L = s.L + (vol ? pending * Tr : pending);
        }
        // scatter, or the end of the path
        bool alive = false;
        if (busy) {
            if (r == 1) alive = true;
            else if (r == 2) alive = scatterStep<FEAT>(S, *J, s, ev);
        }
        const bool ending = busy && !alive;
        if (ending) {
            storePath(A.P, p, s);
            busy = false;
        }
        retirePaths(A.chunkLive, ending, ending && A.chunkLive ? A.P.chunk[p] : -1);
    }
    // queries for the ray statistics (closest-hit and shadow), one atomic per wave
    unsigned long long a = nRays, b = nShadow;
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off);
        b += __shfl_down(b, off);
    }
    if (__lane_id() == 0) {
        if (a) atomicAdd(&A.ctr->tailRays, a);
        if (b) atomicAdd(&A.ctr->tailSrays, b);
    }
    if (st.overflow) atomicAdd(&A.ctr->stackOverflow, 1ull);
}

// Film gather (Film::addSample, film.h:26-69): one thread per (job, pixel)
// sums the filtered contributions of the neighbouring samples of each pass in
// pass order, so results do not depend on scheduling.
__global__ void kSplat(WaveArgs A, const DSplatTask* tasks, int nTasks, uint32_t totalPix) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= totalPix) return;
    int lo = 0, hi = nTasks - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (tasks[mid].pixBase <= t) lo = mid;
        else hi = mid - 1;
    }
    const DSplatTask tk = tasks[lo];
    const DJob& J = A.jobs[tk.job];
    const uint32_t pix = t - tk.pixBase;
    const int W = J.W, H = J.H;
    const int x = (int)(pix % (uint32_t)W), y = (int)(pix / (uint32_t)W);
    const DCamera& c = J.cam;
    const float rx = c.filter == GND_FILTER_BOX ? 0.5f : c.radiusX;
    const float ry = c.filter == GND_FILTER_BOX ? 0.5f : c.radiusY;
    // A sample of pixel column qx has x + jitter in [qx, qx + 1] (the float sum
    // may round up to qx + 1), so its support [px - r - 0.5, px + r - 0.5]
    // holds no pixel centre further than ceil(r) columns away: a neighbour
    // beyond that fails the x0..x1 / y0..y1 test below, so skipping it leaves
    // the sums and their order unchanged.
    const int reach = max(0, (int)ceilf(stdMax(rx, ry)));
    float4 acc = reinterpret_cast<float4*>(J.film)[pix];
    float3 accSq = J.filmSq ? reinterpret_cast<float3*>(J.filmSq)[pix] : float3{0, 0, 0};
    const bool aov = J.filmAov && A.P.aov;
    float accA[kAovFloats];
#pragma unroll
    for (int k = 0; k < kAovFloats; ++k) accA[k] = aov ? J.filmAov[(size_t)pix * kAovFloats + k] : 0.0f;
    unsigned long long nan = 0;
    for (int ci = tk.chunkBegin; ci < tk.chunkEnd; ++ci) {
        const DChunk ch = A.chunks[ci];
        float sr = 0.0f, sg = 0.0f, sb = 0.0f, sw = 0.0f;
        float sa[kAovFloats];
#pragma unroll
        for (int k = 0; k < kAovFloats; ++k) sa[k] = 0.0f;
        for (int dy = -reach; dy <= reach; ++dy) {
            const int qy = y + dy;
            if (qy < 0 || qy >= H) continue;
            for (int dx = -reach; dx <= reach; ++dx) {
                const int qx = x + dx;
                if (qx < 0 || qx >= W) continue;
                const uint32_t q = ch.slotBase + (uint32_t)(qy * W + qx);
                const float px = A.P.pfx[q], py = A.P.pfy[q];
                const int x0 = max(0, (int)ceilf(px - rx - 0.5f));
                const int x1 = min(W - 1, (int)floorf(px + rx - 0.5f));
                const int y0 = max(0, (int)ceilf(py - ry - 0.5f));
                const int y1 = min(H - 1, (int)floorf(py + ry - 0.5f));
                if (x < x0 || x > x1 || y < y0 || y > y1) continue;
                const float w = filterEval(c, px - (x + 0.5f), py - (y + 0.5f));
                if (!(w > 0.0f)) continue;
                // final sample value (integrator.h:128-141)
                V3 L = V3{A.P.Lr[q], A.P.Lg[q], A.P.Lb[q]} * A.P.weight[q];
                const int chan = flagChannel(A.P.flags[q]);
                if (chan == 0) L = V3{L.x * 3.0f, 0.0f, 0.0f};
                else if (chan == 1) L = V3{0.0f, L.y * 3.0f, 0.0f};
                else if (chan == 2) L = V3{0.0f, 0.0f, L.z * 3.0f};
                if (hasNaN(L)) {
                    L = sameV(0.0f);
                    if (dx == 0 && dy == 0) nan++;
                }
                sr += L.x * w;
                sg += L.y * w;
                sb += L.z * w;
                sw += w;
                if (aov) {
#pragma unroll
                    for (int k = 0; k < kAovFloats; ++k) sa[k] += A.P.aov[(size_t)k * A.P.aovStride + q] * w;
                }
            }
        }
        acc.x += sr;
        acc.y += sg;
        acc.z += sb;
        acc.w += sw;
        if (aov) {
#pragma unroll
            for (int k = 0; k < kAovFloats; ++k) accA[k] += sa[k];
        }
        if (J.filmSq && sw > 0.0f) {
            float vr = sr / sw, vg = sg / sw, vb = sb / sw;
            accSq.x += vr * vr;
            accSq.y += vg * vg;
            accSq.z += vb * vb;
        }
    }
    reinterpret_cast<float4*>(J.film)[pix] = acc;
    if (J.filmSq) reinterpret_cast<float3*>(J.filmSq)[pix] = accSq;
    if (aov) {
#pragma unroll
        for (int k = 0; k < kAovFloats; ++k) J.filmAov[(size_t)pix * kAovFloats + k] = accA[k];
    }
    if (nan) atomicAdd(&A.ctr->nan, nan);
}

// plain streaming copy (bandwidth probe and HBM-counter calibration)
__global__ void kCopyProbe(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

__global__ void kFill(float* p, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void kFillU32(uint32_t* p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// End of a bounce: the out-queue (and its density-grid list) becomes the next
// in-queue; out, shadow and ray-fetch counters restart at zero (a kernel, so the step can be captured
// in a graph).
__global__ void kRotate(uint32_t* counts, Counters* c) {
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        c->rays += counts[0];
        c->srays += counts[2];
        if (counts[0]) c->launches += 1ull;
        counts[0] = counts[1];
        counts[1] = 0u;
        counts[2] = 0u;
        counts[4] = counts[3];
        counts[3] = 0u;
    } else if (t >= kFeedTrace && t < kNumCounts) {
        counts[t] = 0u;
    }
}

// ============================================================================ host side

struct gnd_hip_scene {
    DScene view{};
    std::vector<void*> allocations;
    size_t bytes = 0;
    int hasIndex = 0, hasGrid = 0, nMedia = 0, hasMotion = 0;
    uint32_t nPrims = 0;
    size_t accelBytes = 0;  // TLAS, instances, BLAS and triangle packs
};

// Queues of one bounce loop for `capacity` paths in flight: hit records and
// path indices by queue position, the rays of the two path queues, the
// shadow queue, the two density-grid lists and the queue counters (160 bytes
// per entry: the 32-byte hit record and 32 four-byte SoA arrays, each 256-B
// aligned).  A bounce loop reads and
// writes only its own queues, so loops with their own queue sets can run
// concurrently over one pool of path slots.
struct QueueSet {
    size_t capacity = 0;
    void* buf = nullptr;
    HitSoA Hq{};
    ShadowSoA Sh{};
    int32_t* q[2] = {nullptr, nullptr};
    RayQ rq[2] = {};
    int32_t* gl[2] = {nullptr, nullptr};  // density-grid lists of the two queues
    uint32_t* counts = nullptr;  // kNumCounts queue and ray-feed counters
};

// Path state per slot (140 bytes in 34 SoA arrays, each 256-B aligned) and
// the queues of gnd_hip_render's bounce loop: 288 bytes per path slot.
struct WaveBuffers {
    size_t capacity = 0;
    void* pathBuf = nullptr;
    void* aovBuf = nullptr;  // per-slot AOV records, allocated on first use
    PathSoA P{};
    QueueSet Q;
};

// Scratch of gnd_hip_denoise_noalloc: the denoiser's four float4 records per
// pixel (two radiance buffers, two guide buffers) for up to kDnStaticPixels
// pixels, 128 MiB on every device the library runs on, held until the
// process ends.  Its name is a host-side stand-in: host code uses the device
// address that gnd_hip_init looks up (gnd_hip_ctx::dnStatic).  Every call on
// one device shares it, so calls that overlap on one device race (calls on
// one context are serialised by its lock).  Each call writes every record it
// reads, so nothing needs clearing between calls.
constexpr size_t kDnStaticPixels = (size_t)1 << 21;  // 2,097,152 pixels: 1920 x 1080 fits
static __device__ __align__(256) float4 scratch[4 * kDnStaticPixels];

struct gnd_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::mutex mu;
    // Up to 256M paths in flight per wave (300 B each with one queue set,
    // 77 GB of the 288 GB; a render allocates only what its jobs' samples
    // need): a 1080p x 16 spp frame is one wave, and a batch of many jobs
    // takes few waves, so few bounces run below the GPU's width (64-job
    // batch over the 1080p scene, 527M samples: 40M slots 347, 120M 421,
    // 264M 434, 540M 437 Msamples/s).
    // 32 LDS stack entries per lane: as fast as 16 on the 1080p scene, and
    // the deeper stack rarely spills to scratch (16 traversal waves per CU,
    // 64 lanes x 32 entries x 4 B each, take 128 KiB of its 160 KiB of LDS)
    gnd_hip_tuning tuning{128, 32, 256000000, 0, 0, 0, 0, 0};
    bool counters = false;
    int numCUs = 0;
    size_t hbm = 0;
    std::string name;
    Counters* dCounters = nullptr;
    WaveBuffers wb;  // path state of gnd_hip_render (grown on demand)
    DJob* dJobs = nullptr;
    size_t jobCap = 0;
    DChunk* dChunks = nullptr;
    size_t chunkCap = 0;
    DSplatTask* dTasks = nullptr;
    size_t taskCap = 0;
    std::atomic<int> cancel{0};  // set from any thread by gnd_hip_cancel
    hipEvent_t ev[2] = {nullptr, nullptr};
    float4* dnScratch = nullptr;  // denoiser buffers (grown on demand)
    size_t dnCap = 0;             // pixels dnScratch holds
    float4* dnStatic = nullptr;   // device address of `scratch` on this device
    // instantiated bounce chains (use_graph), keyed by their exact launches
    struct CachedGraph {
        std::string key;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        uint64_t lastUse = 0;
    };
    std::vector<CachedGraph> graphs;
    uint64_t graphClock = 0;
    std::map<std::string, gnd_hip_tuning> tuned;  // gnd_hip_autotune results per scene and job shape
    std::map<std::string, int> chainDepth;         // bounces before the tail kernel, per wave shape
};

// gnd_hip_autotune's probe: at most this many samples (one default wave)
constexpr uint64_t kAutotuneSamples = 40000000;

namespace {

thread_local std::string g_lastError;

struct HipError : std::runtime_error {
    int code;
    HipError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define GND_CHECK(expr)                                                                                     \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess)                                                                               \
            throw HipError(e_ == hipErrorOutOfMemory ? GND_E_NOMEM : GND_E_DEVICE,                         \
                           std::string(#expr) + ": " + hipGetErrorString(e_));                              \
    } while (0)

template <class F>
int guarded(gnd_hip_ctx* ctx, F&& f) {
    try {
        f();
        return GND_OK;
    } catch (const HipError& e) {
        if (ctx) ctx->err = e.what();
        g_lastError = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = e.what();
        g_lastError = e.what();
        return GND_E_INVALID;
    }
}

// Corner records of a density grid (the layout DMedium::corners holds):
// thread v writes voxel v's 8 trilinear corners, 0 beyond the grid.
__global__ __launch_bounds__(256) void kGridCorners(const float* __restrict__ d, float4* __restrict__ out, int nx,
                                                    int ny, int nz) {
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nxy = (size_t)nx * ny;
    if (v >= nxy * (size_t)nz) return;
    const int x = (int)(v % (size_t)nx), y = (int)((v / (size_t)nx) % (size_t)ny), z = (int)(v / nxy);
    auto at = [&](int i, int j, int k) { return (i < nx && j < ny && k < nz) ? d[(size_t)k * nxy + (size_t)j * nx + i] : 0.0f; };
    out[2 * v] = make_float4(at(x, y, z), at(x, y, z + 1), at(x, y + 1, z), at(x, y + 1, z + 1));
    out[2 * v + 1] = make_float4(at(x + 1, y, z), at(x + 1, y, z + 1), at(x + 1, y + 1, z), at(x + 1, y + 1, z + 1));
}
// grids whose corner records would take more than this keep the one-by-one lookups
constexpr size_t kGridCornerBytes = size_t(4) << 30;

template <class T>
T* devAlloc(gnd_hip_scene* s, size_t count, const T* src) {
    if (count == 0) count = 1;
    void* p = nullptr;
    GND_CHECK(hipMalloc(&p, count * sizeof(T)));
    s->allocations.push_back(p);
    s->bytes += count * sizeof(T);
    if (src) GND_CHECK(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice));
    return static_cast<T*>(p);
}

DNode toDNode(const gnd_bvh4_node& n, uint32_t innerBase, uint32_t leafBase) {
    DNode d;
    float* lo[3] = {&d.xlo.x, &d.ylo.x, &d.zlo.x};
    float* hi[3] = {&d.xhi.x, &d.yhi.x, &d.zhi.x};
    const float* slo[3] = {n.min_x, n.min_y, n.min_z};
    const float* shi[3] = {n.max_x, n.max_y, n.max_z};
    for (int a = 0; a < 3; ++a)
        for (int i = 0; i < 4; ++i) {
            lo[a][i] = slo[a][i];
            hi[a][i] = shi[a][i];
        }
    uint32_t* ch = &d.child.x;
    uint32_t* me = &d.meta.x;
    for (int i = 0; i < 4; ++i) {
        uint32_t type = n.child_type[i];
        if (type == 1) ch[i] = n.offset[i] + innerBase;
        else if (type == 2) ch[i] = n.offset[i] + leafBase;
        else ch[i] = 0;
        if (n.count[i] > 0x3FFFFFFFu) throw std::runtime_error("BVH leaf count too large");
        me[i] = n.count[i] | (type << 30);
    }
    return d;
}

// Light constructors (pointlight.cpp:9-12, spotlight.cpp:9-22, directional.cpp:8-11)
DLight makeLight(const gnd_light_desc& l) {
    DLight L;
    std::memset(&L, 0, sizeof(L));
    L.type = l.type;
    for (int c = 0; c < 3; ++c) {
        L.pos[c] = l.position[c];
        L.power[c] = l.power[c];
        L.irr[c] = l.irradiance[c];
    }
    const float dx = l.direction[0], dy = l.direction[1], dz = l.direction[2];
    if (l.type == GND_LIGHT_SPOT || l.type == GND_LIGHT_DIRECTIONAL) {
        // Normalize(v) = v * (1 / sqrt(dot)); the directional light points back
        const float inv = 1.0f / std::sqrt((dx * dx + dy * dy) + dz * dz);
        const float sgn = l.type == GND_LIGHT_DIRECTIONAL ? -1.0f : 1.0f;
        L.dir[0] = sgn * (dx * inv);
        L.dir[1] = sgn * (dy * inv);
        L.dir[2] = sgn * (dz * inv);
    }
    if (l.type == GND_LIGHT_SPOT) {
        float halfAngle = l.spot_size * 0.5f;
        float innerAngle = halfAngle * (1.0f - l.blend);
        L.cosTotal = std::cos((kPi / 180.f) * halfAngle);
        L.cosFalloff = std::cos((kPi / 180.f) * innerAngle);
    }
    if (l.type == GND_LIGHT_ENVIRONMENT && (!l.env_rgb || l.env_width <= 0 || l.env_height <= 0))
        throw std::runtime_error("environment light without an image");
    return L;
}

// EnvironmentMap::buildPDFs (environment.cpp:106-120) with Distribution1D /
// Distribution2D (distribution.cpp:6-18, 53-64): per row the luminance times
// sin(theta of the row centre), a piecewise-constant CDF per row and over the
// rows' integrals.
struct EnvTables {
    std::vector<float> rgb, condFunc, condCdf, condInt, margCdf;
    std::vector<int32_t> condGuide, margGuide;
    float margInt = 0.0f;
};
// guide[j] = std::upper_bound(cdf, cdf + n + 1, j / kEnvGuide) - cdf for
// j = 0..kEnvGuide, guide[kEnvGuide + 1] = n + 1 (u = 1 exactly, which the
// sampler can draw)
static void guide1D(const float* cdf, int n, int32_t* g) {
    for (int j = 0; j <= kEnvGuide; ++j) {
        const float u = (float)j / (float)kEnvGuide;
        g[j] = (int32_t)(std::upper_bound(cdf, cdf + n + 1, u) - cdf);
    }
    g[kEnvGuide + 1] = n + 1;
}
static void dist1D(const float* f, int n, float* cdf, float& funcInt) {
    cdf[0] = 0.0f;
    for (int i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + f[i - 1] / n;
    funcInt = cdf[n];
    if (funcInt == 0.0f) {
        for (int i = 1; i < n + 1; ++i) cdf[i] = float(i) / float(n);
    } else {
        for (int i = 1; i < n + 1; ++i) cdf[i] /= funcInt;
    }
}
EnvTables envTables(const gnd_light_desc& l) {
    EnvTables t;
    const int W = l.env_width, H = l.env_height;
    t.rgb.assign(l.env_rgb, l.env_rgb + (size_t)W * H * 3);
    t.condFunc.resize((size_t)W * H);
    for (int i = 0; i < H; i++) {
        float sinTheta = std::sin(M_PI * (i + 0.5f) / H);  // double sine, stored as float
        for (int j = 0; j < W; j++) {
            const float* c = &t.rgb[3 * ((size_t)i * W + j)];
            const float lum = c[0] * 0.212671f + c[1] * 0.715160f + c[2] * 0.072169f;  // Color3f::luminance
            t.condFunc[(size_t)i * W + j] = lum * sinTheta;
        }
    }
    t.condCdf.resize((size_t)H * (W + 1));
    t.condInt.resize(H);
    for (int v = 0; v < H; ++v) dist1D(&t.condFunc[(size_t)v * W], W, &t.condCdf[(size_t)v * (W + 1)], t.condInt[v]);
    t.margCdf.resize(H + 1);
    dist1D(t.condInt.data(), H, t.margCdf.data(), t.margInt);
    t.condGuide.resize((size_t)H * (kEnvGuide + 2));
    for (int v = 0; v < H; ++v)
        guide1D(&t.condCdf[(size_t)v * (W + 1)], W, &t.condGuide[(size_t)v * (kEnvGuide + 2)]);
    t.margGuide.resize(kEnvGuide + 2);
    guide1D(t.margCdf.data(), H, t.margGuide.data());
    return t;
}

DCamera makeCamera(const gnd_camera_desc& c) {
    DCamera d{};
    std::memcpy(d.toWorld.m, c.to_world, sizeof(d.toWorld.m));
    d.type = c.type;
    d.width = c.width;
    d.height = c.height;
    d.blades = c.blades;
    // PerspectiveCamera / ThinLensCamera ctor: m_scale = tan(Radians(fov / 2))
    d.scale = std::tan((kPi / 180.f) * (c.fov * 0.5f));
    d.aspect = c.width / (float)c.height;
    d.lensRadius = c.lens_radius;
    d.focalDistance = c.focal_distance;
    d.k1 = c.k1;
    d.caAxial = c.ca_axial;
    d.caLateral = c.ca_lateral;
    d.vignetting = c.vignetting;
    d.filter = c.filter;
    d.radiusX = c.filter_radius[0];
    d.radiusY = c.filter_radius[1];
    d.alpha = c.filter_alpha;
    d.expX = std::exp(-c.filter_alpha * c.filter_radius[0] * c.filter_radius[0]);
    d.expY = std::exp(-c.filter_alpha * c.filter_radius[1] * c.filter_radius[1]);
    d.hasCA = (c.type == GND_CAMERA_THINLENS && (c.ca_lateral != 0.0f || c.ca_axial != 0.0f)) ? 1 : 0;
    // AnimatedTransform(tStart, 0, tEnd, 1): animated when the matrices differ
    std::memcpy(d.toWorldEnd.m, c.has_motion ? c.to_world_end : c.to_world, sizeof(d.toWorldEnd.m));
    d.animated = c.has_motion && std::memcmp(c.to_world, c.to_world_end, sizeof(c.to_world)) != 0 ? 1 : 0;
    if (d.animated) {
        gndMotionFrom(c.to_world, c.to_world_end, d.motion);
    }
    d.rolling = c.type == GND_CAMERA_THINLENS && c.rolling_shutter ? 1 : 0;
    d.scanDuration = c.scan_duration;
    d.exposureDuration = c.exposure_duration;
    return d;
}

template <class T>
void ensureDev(T*& p, size_t& cap, size_t need) {
    if (need <= cap) return;
    if (p) GND_CHECK(hipFree(p));
    size_t n = std::max(need, cap * 2);
    GND_CHECK(hipMalloc((void**)&p, n * sizeof(T)));
    cap = n;
}

// four-byte fields per slot / per queue entry beside the 32-byte hit record
constexpr size_t kPathFields4 = 33, kQueueFields4 = 32;

size_t alignedField(size_t cap, size_t elemBytes) { return ((cap * elemBytes + 255) / 256) * 256; }
size_t pathStateBytes(size_t cap) { return kPathFields4 * alignedField(cap, 4) + alignedField(cap, 8); }
size_t queueSetBytes(size_t cap) { return kQueueFields4 * alignedField(cap, 4) + alignedField(cap, 32); }
size_t waveBufferBytes(size_t cap) { return pathStateBytes(cap) + queueSetBytes(cap); }

// Hands out 256-B aligned SoA fields of `cap` entries from one allocation.
struct FieldCarver {
    char* base;
    size_t cap, off = 0;
    void* take(size_t elemBytes) {
        void* p = base + off;
        off += alignedField(cap, elemBytes);
        return p;
    }
};

void freeQueueSet(QueueSet& qs) {
    if (qs.buf) (void)hipFree(qs.buf);
    if (qs.counts) (void)hipFree(qs.counts);
    qs = QueueSet{};
}

void ensureQueueSet(QueueSet& qs, size_t cap) {
    if (!qs.counts) GND_CHECK(hipMalloc(&qs.counts, kNumCounts * sizeof(uint32_t)));
    if (cap <= qs.capacity) return;
    if (qs.buf) GND_CHECK(hipFree(qs.buf));
    qs.buf = nullptr;
    qs.capacity = 0;
    const size_t bytes = queueSetBytes(cap);
    GND_CHECK(hipMalloc(&qs.buf, bytes));
    FieldCarver c{static_cast<char*>(qs.buf), cap};
    qs.Hq.rec = (uint4*)c.take(32);
    ShadowSoA& S = qs.Sh;
    S.path = (int32_t*)c.take(4);
    S.medium = (int32_t*)c.take(4);
    float** g[] = {&S.ox, &S.oy, &S.oz, &S.dx, &S.dy, &S.dz, &S.tmax, &S.pr, &S.pg, &S.pb, &S.lpx, &S.lpy, &S.lpz,
                   &S.time};
    for (auto pp : g) *pp = (float*)c.take(4);
    qs.q[0] = (int32_t*)c.take(4);
    qs.q[1] = (int32_t*)c.take(4);
    qs.gl[0] = (int32_t*)c.take(4);
    qs.gl[1] = (int32_t*)c.take(4);
    for (RayQ& r : qs.rq)
        for (float** pp : {&r.ox, &r.oy, &r.oz, &r.dx, &r.dy, &r.dz}) *pp = (float*)c.take(4);
    if (c.off > bytes) throw std::runtime_error("queue layout overflow");
    qs.capacity = cap;
}

void freeWaveBuffers(WaveBuffers& wb) {
    if (wb.pathBuf) (void)hipFree(wb.pathBuf);
    if (wb.aovBuf) (void)hipFree(wb.aovBuf);
    freeQueueSet(wb.Q);
    wb = WaveBuffers{};
}

// Path state for `cap` slots; with `queues`, also the queue set of one bounce
// loop over them.
void ensureWaveBuffers(WaveBuffers& wb, size_t cap, bool queues = true) {
    if (queues) ensureQueueSet(wb.Q, cap);
    if (cap <= wb.capacity) return;
    if (wb.pathBuf) GND_CHECK(hipFree(wb.pathBuf));
    wb.pathBuf = nullptr;
    if (wb.aovBuf) GND_CHECK(hipFree(wb.aovBuf));
    wb.aovBuf = nullptr;
    wb.P.aov = nullptr;
    wb.P.aovStride = 0;
    wb.capacity = 0;
    // 33 four-byte fields and one eight-byte field per slot
    const size_t bytes = pathStateBytes(cap);
    GND_CHECK(hipMalloc(&wb.pathBuf, bytes));
    FieldCarver c{static_cast<char*>(wb.pathBuf), cap};
    PathSoA& P = wb.P;
    P.rng = (uint64_t*)c.take(8);
    float** f[] = {&P.tpr, &P.tpg, &P.tpb, &P.Lr, &P.Lg, &P.Lb,
                   &P.pdfPrev, &P.lx, &P.ly, &P.lz, &P.weight, &P.pfx, &P.pfy, &P.evpx, &P.evpy, &P.evpz,
                   &P.evnx, &P.evny, &P.evnz, &P.evtx, &P.evty, &P.evtz, &P.evu, &P.evv, &P.time};
    for (auto pp : f) *pp = (float*)c.take(4);
    P.c0 = (uint32_t*)c.take(4);
    P.c1 = (uint32_t*)c.take(4);
    P.flags = (uint32_t*)c.take(4);
    P.job = (int32_t*)c.take(4);
    P.chunk = (int32_t*)c.take(4);
    P.evKind = (int32_t*)c.take(4);
    P.evPrim = (int32_t*)c.take(4);
    P.evMedium = (int32_t*)c.take(4);
    if (c.off > bytes) throw std::runtime_error("path state layout overflow");
    wb.capacity = cap;
}

// AOV records for every slot of the wave buffers (36 bytes per slot)
void ensureAovBuffers(WaveBuffers& wb) {
    if (wb.aovBuf || wb.capacity == 0) return;
    GND_CHECK(hipMalloc(&wb.aovBuf, wb.capacity * kAovFloats * sizeof(float)));
    wb.P.aov = static_cast<float*>(wb.aovBuf);
    wb.P.aovStride = (uint32_t)wb.capacity;
}

using TraceKernel = void (*)(DScene, gnd_ray_soa, int64_t, gnd_hit_soa, int32_t*, Counters*, int32_t);

template <bool ANY, bool COUNT>
TraceKernel pickTrace(int lds) {
    switch (lds) {
        case 8: return kTraceBatch<8, ANY, COUNT>;
        case 24: return kTraceBatch<24, ANY, COUNT>;
        case 32: return kTraceBatch<32, ANY, COUNT>;
        default: return kTraceBatch<16, ANY, COUNT>;
    }
}
// LDS stack depths compiled in: 8, 16, 24, 32 entries per lane
int normLds(int lds) { return lds <= 8 ? 8 : (lds <= 16 ? 16 : (lds <= 24 ? 24 : 32)); }
// traversal stack entries per lane: the tuning's limit, or the whole stack
int stackCapOf(const gnd_hip_tuning& t) { return t.stack_limit > 0 ? std::min(t.stack_limit, kStackCap) : kStackCap; }

using WaveKernel = void (*)(WaveArgs);
template <bool COUNT, bool MOTION>
WaveKernel pickTracePaths(int lds) {
    switch (lds) {
        case 8: return kTracePaths<8, COUNT, MOTION>;
        case 24: return kTracePaths<24, COUNT, MOTION>;
        case 32: return kTracePaths<32, COUNT, MOTION>;
        default: return kTracePaths<16, COUNT, MOTION>;
    }
}
template <bool FAST, bool COUNT, bool MOTION>
WaveKernel pickShadow(int lds) {
    switch (lds) {
        case 8: return FAST ? kShadowAny<8, COUNT, MOTION> : kShadow<8, COUNT, MOTION>;
        case 24: return FAST ? kShadowAny<24, COUNT, MOTION> : kShadow<24, COUNT, MOTION>;
        case 32: return FAST ? kShadowAny<32, COUNT, MOTION> : kShadow<32, COUNT, MOTION>;
        default: return FAST ? kShadowAny<16, COUNT, MOTION> : kShadow<16, COUNT, MOTION>;
    }
}
template <bool MOTION>
void pickBounceKernels(bool fast, bool cnt, int lds, WaveKernel& trace, WaveKernel& shadow) {
    trace = cnt ? pickTracePaths<true, MOTION>(lds) : pickTracePaths<false, MOTION>(lds);
    shadow = fast ? (cnt ? pickShadow<true, true, MOTION>(lds) : pickShadow<true, false, MOTION>(lds))
                  : (cnt ? pickShadow<false, true, MOTION>(lds) : pickShadow<false, false, MOTION>(lds));
}

float elapsedMs(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    GND_CHECK(hipEventElapsedTime(&ms, a, b));
    return ms;
}

}  // namespace

// ============================================================================ C-ABI

// Defined with the host scene shell: frees the device copies that host scenes
// made on `ctx` (called at shutdown while the context is still valid).
void gndReleaseContextUploads(gnd_hip_ctx* ctx);
static int sceneUpload(gnd_hip_ctx* ctx, const gnd_hip_scene_arrays* a, gnd_hip_scene** out);

extern "C" {

const char* gnd_hip_last_error(const gnd_hip_ctx* ctx) { return ctx ? ctx->err.c_str() : g_lastError.c_str(); }

int gnd_hip_init(int device, gnd_hip_ctx** out) {
    if (!out) return GND_E_INVALID;
    gnd_hip_ctx* ctx = new gnd_hip_ctx();
    int rc = guarded(ctx, [&] {
        int n = 0;
        GND_CHECK(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) throw std::runtime_error("no such HIP device");
        ctx->device = device;
        GND_CHECK(hipSetDevice(device));
        hipDeviceProp_t prop;
        GND_CHECK(hipGetDeviceProperties(&prop, device));
        ctx->numCUs = prop.multiProcessorCount;
        ctx->hbm = prop.totalGlobalMem;
        ctx->name = prop.gcnArchName;
        GND_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        GND_CHECK(hipMalloc(&ctx->dCounters, sizeof(Counters)));
        GND_CHECK(hipMemset(ctx->dCounters, 0, sizeof(Counters)));
        GND_CHECK(hipEventCreate(&ctx->ev[0]));
        GND_CHECK(hipEventCreate(&ctx->ev[1]));
        // the no-allocation denoiser's scratch: this device's copy of the array
        GND_CHECK(hipGetSymbolAddress((void**)&ctx->dnStatic, HIP_SYMBOL(scratch)));
    });
    if (rc != GND_OK) {
        g_lastError = ctx->err;
        delete ctx;
        *out = nullptr;
        return rc;
    }
    *out = ctx;
    return GND_OK;
}

void gnd_hip_shutdown(gnd_hip_ctx* ctx) {
    if (!ctx) return;
    gndReleaseContextUploads(ctx);
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    freeWaveBuffers(ctx->wb);
    if (ctx->dCounters) (void)hipFree(ctx->dCounters);
    if (ctx->dJobs) (void)hipFree(ctx->dJobs);
    if (ctx->dChunks) (void)hipFree(ctx->dChunks);
    if (ctx->dTasks) (void)hipFree(ctx->dTasks);
    if (ctx->dnScratch) (void)hipFree(ctx->dnScratch);
    for (auto e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& g : ctx->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int gnd_hip_device_info(gnd_hip_ctx* ctx, char* name, int name_len, int* num_cus, size_t* hbm_bytes) {
    if (!ctx) return GND_E_INVALID;
    if (name && name_len > 0) {
        std::strncpy(name, ctx->name.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    if (num_cus) *num_cus = ctx->numCUs;
    if (hbm_bytes) *hbm_bytes = ctx->hbm;
    return GND_OK;
}

int gnd_hip_set_tuning(gnd_hip_ctx* ctx, const gnd_hip_tuning* t) {
    if (!ctx || !t) return GND_E_INVALID;
    if (t->block_size != 64 && t->block_size != 128 && t->block_size != 256) {
        ctx->err = "block_size must be 64, 128 or 256";
        return GND_E_INVALID;
    }
    if (t->queue_capacity < 1024) {
        ctx->err = "queue_capacity too small";
        return GND_E_INVALID;
    }
    if (t->refill_lanes < 0 || t->refill_lanes > 64 || t->shadow_refill_lanes < 0 || t->shadow_refill_lanes > 64) {
        ctx->err = "refill lanes must be 0 (per scene) or 1..64";
        return GND_E_INVALID;
    }
    if (t->tail_paths < -1) {
        ctx->err = "tail_paths must be -1 (no tail kernel), 0 (per device) or a path count";
        return GND_E_INVALID;
    }
    if (t->stack_limit < 0 || t->stack_limit > kStackCap) {
        ctx->err = "stack_limit must be 0 (the full stack) or 1.." + std::to_string(kStackCap);
        return GND_E_INVALID;
    }
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->tuning = *t;
    ctx->tuning.stack_lds = normLds(t->stack_lds);
    return GND_OK;
}

int gnd_hip_get_tuning(gnd_hip_ctx* ctx, gnd_hip_tuning* t) {
    if (!ctx || !t) return GND_E_INVALID;
    *t = ctx->tuning;
    return GND_OK;
}

int gnd_hip_set_counters(gnd_hip_ctx* ctx, int enabled) {
    if (!ctx) return GND_E_INVALID;
    ctx->counters = enabled != 0;
    return GND_OK;
}

int gnd_hip_malloc(gnd_hip_ctx* ctx, size_t bytes, void** dptr) {
    return guarded(ctx, [&] {
        GND_CHECK(hipSetDevice(ctx->device));
        GND_CHECK(hipMalloc(dptr, bytes ? bytes : 1));
    });
}
int gnd_hip_free(gnd_hip_ctx* ctx, void* dptr) {
    return guarded(ctx, [&] { GND_CHECK(hipFree(dptr)); });
}
int gnd_hip_memcpy_h2d(gnd_hip_ctx* ctx, void* dst, const void* src, size_t bytes) {
    return guarded(ctx, [&] {
        GND_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        GND_CHECK(hipStreamSynchronize(ctx->stream));
    });
}
int gnd_hip_memcpy_d2h(gnd_hip_ctx* ctx, void* dst, const void* src, size_t bytes) {
    return guarded(ctx, [&] {
        GND_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        GND_CHECK(hipStreamSynchronize(ctx->stream));
    });
}
int gnd_hip_memset(gnd_hip_ctx* ctx, void* dst, int value, size_t bytes) {
    return guarded(ctx, [&] {
        GND_CHECK(hipMemsetAsync(dst, value, bytes, ctx->stream));
        GND_CHECK(hipStreamSynchronize(ctx->stream));
    });
}
int gnd_hip_synchronize(gnd_hip_ctx* ctx) {
    return guarded(ctx, [&] { GND_CHECK(hipStreamSynchronize(ctx->stream)); });
}

int gnd_hip_scene_upload(gnd_hip_ctx* ctx, const gnd_hip_scene_arrays* a, gnd_hip_scene** out) {
    if (!ctx || !a || !out || !a->desc) return GND_E_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    return sceneUpload(ctx, a, out);
}

}  // extern "C"

// gnd_hip_scene_upload with the context lock held by the caller
static int sceneUpload(gnd_hip_ctx* ctx, const gnd_hip_scene_arrays* a, gnd_hip_scene** out) {
    gnd_hip_scene* s = new gnd_hip_scene();
    int rc = guarded(ctx, [&] {
        GND_CHECK(hipSetDevice(ctx->device));
        const gnd_scene_desc& d = *a->desc;
        // stack entries are non-negative int node indices
        if (a->n_blas_nodes > 0x7FFFFFFFu || a->n_tlas_nodes > 0x7FFFFFFFu)
            throw std::runtime_error("scene too large for the traversal stack encoding");
        // TLAS
        std::vector<DNode> tl(a->n_tlas_nodes);
        for (uint32_t i = 0; i < a->n_tlas_nodes; ++i) tl[i] = toDNode(a->tlas_nodes[i], 0, 0);
        // BLAS: children made absolute per mesh
        std::vector<DNode> bl(a->n_blas_nodes);
        for (uint32_t m = 0; m < a->n_meshes; ++m) {
            const gnd_mesh_info& mi = a->meshes[m];
            if (mi.node_base + mi.n_nodes > a->n_blas_nodes || mi.pack_base + mi.n_packs > a->n_packs)
                throw std::runtime_error("mesh slice out of range");
            for (uint32_t i = 0; i < mi.n_nodes; ++i)
                bl[mi.node_base + i] = toDNode(a->blas_nodes[mi.node_base + i], mi.node_base, mi.pack_base);
        }
        if (d.nprims == 0) throw std::runtime_error("scene has no primitives");
        std::vector<DInst> inst(d.nprims);
        std::vector<DPrim> prims(d.nprims);
        std::vector<DPrimMotion> motions;
        std::vector<int32_t> emitters;
        for (uint32_t i = 0; i < d.nprims; ++i) {
            const gnd_primitive_desc& p = d.prims[i];
            DInst& I = inst[i];
            std::memcpy(I.worldToObject.m, p.to_world_inv, sizeof(I.worldToObject.m));
            I.radius = p.radius;
            I.motion = -1;
            I.prim = (int32_t)i;
            // a moving primitive (AnimatedTransform: animated when the matrices differ)
            int motion = -1;
            if (p.has_motion && std::memcmp(p.to_world, p.to_world_end, sizeof(p.to_world)) != 0) {
                DPrimMotion m;
                std::memcpy(m.start.m, p.to_world, sizeof(m.start.m));
                std::memcpy(m.startInv.m, p.to_world_inv, sizeof(m.startInv.m));
                std::memcpy(m.end.m, p.to_world_end, sizeof(m.end.m));
                std::memcpy(m.endInv.m, p.to_world_end_inv, sizeof(m.endInv.m));
                gndMotionFrom(p.to_world, p.to_world_end, m.mo);
                motion = (int)motions.size();
                motions.push_back(m);
            }
            I.motion = motion;
            DPrim& P = prims[i];
            std::memset(&P, 0, sizeof(P));
            std::memcpy(P.toWorld.m, p.to_world, sizeof(P.toWorld.m));
            std::memcpy(P.toWorldInv.m, p.to_world_inv, sizeof(P.toWorldInv.m));
            P.shape = p.shape_type;
            P.mesh = p.mesh;
            P.material = p.material;
            if (p.material < 0 || (uint32_t)p.material >= d.nmaterials) throw std::runtime_error("bad material index");
            P.radius = p.radius;
            P.radiance[0] = p.radiance[0]; P.radiance[1] = p.radiance[1]; P.radiance[2] = p.radiance[2];
            P.emitter = p.has_emitter ? 0 : -1;  // slot assigned below
            P.motion = motion;
            if (p.shape_type == GND_SHAPE_MESH) {
                if (p.mesh < 0 || (uint32_t)p.mesh >= a->n_meshes) throw std::runtime_error("bad mesh index");
                const gnd_mesh_info& mi = a->meshes[p.mesh];
                if (mi.n_nodes == 0) throw std::runtime_error("mesh without geometry");
                I.root = (int32_t)mi.node_base;
                P.triBase = mi.tri_base;
                P.nTris = mi.n_tris;
                P.cdfBase = mi.cdf_base;
                P.hasUV = mi.has_uv;
                P.totalArea = mi.total_area;
            } else {
                I.root = -1;
            }
        }
        // emitter list in document order (Scene::addChild, scene.cpp:9-40):
        // the lights at their emitter_index, the area emitters of the
        // primitives in the other slots in primitive order
        int nArea = 0;
        for (uint32_t i = 0; i < d.nprims; ++i) nArea += prims[i].emitter >= 0 ? 1 : 0;
        const int nEm = nArea + (int)d.nlights;
        emitters.assign(nEm, INT32_MIN);
        std::vector<DLight> lights(std::max<uint32_t>(1, d.nlights));
        int envLight = -1;
        for (uint32_t k = 0; k < d.nlights; ++k) {
            const gnd_light_desc& ld = d.lights[k];
            const int e = ld.emitter_index;
            if (e < 0 || e >= nEm || emitters[e] != INT32_MIN) throw std::runtime_error("bad light emitter_index");
            emitters[e] = -(int32_t)k - 1;
            lights[k] = makeLight(ld);
            if (ld.type == GND_LIGHT_ENVIRONMENT) {
                if (envLight >= 0) throw std::runtime_error("more than one environment light");
                envLight = (int)k;
            }
        }
        {
            int slot = 0;
            for (uint32_t i = 0; i < d.nprims; ++i) {
                if (prims[i].emitter < 0) continue;
                while (emitters[slot] != INT32_MIN) ++slot;
                prims[i].emitter = slot;
                emitters[slot] = (int32_t)i;
            }
        }
        if (emitters.empty()) throw std::runtime_error("scene has no emitter");
        std::vector<DMaterial> mats(d.nmaterials);
        for (uint32_t i = 0; i < d.nmaterials; ++i) {
            const gnd_material_desc& m = d.materials[i];
            DMaterial& M = mats[i];
            std::memset(&M, 0, sizeof(M));
            M.type = m.type;
            M.checker = m.albedo_is_checker;
            M.mediumInside = m.medium_inside;
            M.mediumOutside = m.medium_outside;
            for (int c = 0; c < 3; ++c) {
                M.albedo[c] = m.albedo[c];
                M.R[c] = m.R[c];
                M.T[c] = m.T[c];
                M.etaRGB[c] = m.eta_rgb[c];
                M.k[c] = m.k_rgb[c];
                M.Kd[c] = m.Kd[c];
                M.Ks[c] = m.Ks[c];
            }
            M.eta = m.eta;
            M.roughness = m.roughness;
            // roughnessToAlpha then the distribution's clamp (microfacet.h:26-39)
            float r = std::max(m.roughness, 1e-3f);
            M.alpha = std::max(1e-4f, r * r);
            if ((m.medium_inside >= (int)d.nmedia) || (m.medium_outside >= (int)d.nmedia))
                throw std::runtime_error("bad medium index");
        }
        std::vector<DMedium> media(d.nmedia);
        for (uint32_t i = 0; i < d.nmedia; ++i) {
            const gnd_medium_desc& m = d.media[i];
            DMedium& M = media[i];
            std::memset(&M, 0, sizeof(M));
            M.type = m.type;
            for (int c = 0; c < 3; ++c) {
                M.sigma_a[c] = m.sigma_a[c];
                M.sigma_s[c] = m.sigma_s[c];
                M.sigma_t[c] = m.sigma_a[c] + m.sigma_s[c];
            }
            M.g = m.g;
            M.maxSigmaT = std::max({M.sigma_t[0], M.sigma_t[1], M.sigma_t[2]});
            std::memcpy(M.toLocal.m, m.to_local, sizeof(M.toLocal.m));
            if (m.type == GND_MEDIUM_GRID) {
                if (!m.density || m.nx <= 0 || m.ny <= 0 || m.nz <= 0) throw std::runtime_error("bad grid medium");
                s->hasGrid = 1;
                M.nx = m.nx; M.ny = m.ny; M.nz = m.nz;
                M.densityScale = m.density_scale;
                // macro majorant grid (heterogeneous.cpp:63-103) over the dense [0, n) box
                float ex = (float)(m.nx - 1) + 1.0f, ey = (float)(m.ny - 1) + 1.0f, ez = (float)(m.nz - 1) + 1.0f;
                M.gridMaxX = ex; M.gridMaxY = ey; M.gridMaxZ = ez;
                int target = m.macro_res > 0 ? m.macro_res : 32;
                float maxE = std::max({ex, ey, ez});
                M.resX = std::max(1, (int)std::round(target * (ex / maxE)));
                M.resY = std::max(1, (int)std::round(target * (ey / maxE)));
                M.resZ = std::max(1, (int)std::round(target * (ez / maxE)));
                M.cellX = ex / M.resX;
                M.cellY = ey / M.resY;
                M.cellZ = ez / M.resZ;
                std::vector<float> maj((size_t)M.resX * M.resY * M.resZ, 0.0f);
                for (int z = 0; z < m.nz; ++z)
                    for (int y = 0; y < m.ny; ++y)
                        for (int x = 0; x < m.nx; ++x) {
                            float dens = m.density[((size_t)z * m.ny + y) * m.nx + x] * m.density_scale;
                            if (dens <= 0.0f || !std::isfinite(dens)) continue;
                            int cx = std::clamp((int)((x - 0.0f) / M.cellX), 0, M.resX - 1);
                            int cy = std::clamp((int)((y - 0.0f) / M.cellY), 0, M.resY - 1);
                            int cz = std::clamp((int)((z - 0.0f) / M.cellZ), 0, M.resZ - 1);
                            size_t fi = (size_t)cz * (M.resX * M.resY) + cy * M.resX + cx;
                            float lm = dens * M.maxSigmaT;
                            if (lm > maj[fi]) maj[fi] = lm;
                        }
                M.density = devAlloc<float>(s, (size_t)m.nx * m.ny * m.nz, m.density);
                M.majorants = devAlloc<float>(s, maj.size(), maj.data());
                // corner records of the voxels, expanded on the device
                const size_t nv = (size_t)m.nx * m.ny * m.nz;
                if (nv * 32 <= kGridCornerBytes) {
                    float4* cr = devAlloc<float4>(s, 2 * nv, nullptr);
                    hipLaunchKernelGGL(kGridCorners, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, nullptr,
                                       M.density, cr, m.nx, m.ny, m.nz);
                    GND_CHECK(hipGetLastError());
                    GND_CHECK(hipStreamSynchronize(nullptr));
                    M.corners = cr;
                }
            }
        }
        for (uint32_t i = 0; i < d.nmaterials; ++i)
            if (d.materials[i].type == GND_MAT_INDEX) s->hasIndex = 1;
        DScene& v = s->view;
        v.tlas = devAlloc<DNode>(s, tl.size(), tl.data());
        // instance records in TLAS slot order
        std::vector<DInst> slotInst(a->n_tlas_prims);
        for (uint32_t k = 0; k < a->n_tlas_prims; ++k) {
            const int32_t pi = a->tlas_prims[k];
            if (pi < 0 || (uint32_t)pi >= d.nprims) throw std::runtime_error("TLAS primitive index out of range");
            slotInst[k] = inst[pi];
        }
        v.tlasInst = devAlloc<DInst>(s, slotInst.size(), slotInst.data());
        v.blas = devAlloc<DNode>(s, bl.size(), bl.data());
        v.packs = reinterpret_cast<const DPack*>(devAlloc<gnd_tri_pack4>(s, a->n_packs, a->packs));
        s->accelBytes = tl.size() * sizeof(DNode) + slotInst.size() * sizeof(DInst) + bl.size() * sizeof(DNode) +
                        (size_t)a->n_packs * sizeof(gnd_tri_pack4);
        v.prims = devAlloc<DPrim>(s, prims.size(), prims.data());
        v.shading = devAlloc<float>(s, (size_t)a->n_tris * 24, a->shading);
        v.cdf = devAlloc<float>(s, a->n_cdf, a->cdf);
        v.mats = devAlloc<DMaterial>(s, mats.size(), mats.data());
        v.media = devAlloc<DMedium>(s, media.size(), media.data());
        v.emitters = devAlloc<int32_t>(s, emitters.size(), emitters.data());
        v.lights = devAlloc<DLight>(s, lights.size(), lights.data());
        v.motions = motions.empty() ? nullptr : devAlloc<DPrimMotion>(s, motions.size(), motions.data());
        v.hasMotion = motions.empty() ? 0 : 1;
        s->hasMotion = v.hasMotion;
        v.nLights = (int32_t)d.nlights;
        v.envLight = envLight;
        if (envLight >= 0) {
            const EnvTables t = envTables(d.lights[envLight]);
            v.envRgb = devAlloc<float>(s, t.rgb.size(), t.rgb.data());
            v.envCondFunc = devAlloc<float>(s, t.condFunc.size(), t.condFunc.data());
            v.envCondCdf = devAlloc<float>(s, t.condCdf.size(), t.condCdf.data());
            v.envCondInt = devAlloc<float>(s, t.condInt.size(), t.condInt.data());
            v.envMargCdf = devAlloc<float>(s, t.margCdf.size(), t.margCdf.data());
            v.envCondGuide = devAlloc<int32_t>(s, t.condGuide.size(), t.condGuide.data());
            v.envMargGuide = devAlloc<int32_t>(s, t.margGuide.size(), t.margGuide.data());
            v.envMargInt = t.margInt;
            v.envStrength = d.lights[envLight].strength;
            v.envW = d.lights[envLight].env_width;
            v.envH = d.lights[envLight].env_height;
        }
        v.nTlas = (int32_t)tl.size();
        v.nEmitters = (int32_t)emitters.size();
        v.nPrims = (int32_t)d.nprims;
        v.hasIndex = s->hasIndex;
        v.hasGrid = s->hasGrid;
        v.nMedia = (int32_t)d.nmedia;
        s->nPrims = d.nprims;
        s->nMedia = (int)d.nmedia;
    });
    if (rc != GND_OK) {
        for (void* p : s->allocations) (void)hipFree(p);
        delete s;
        *out = nullptr;
        return rc;
    }
    *out = s;
    return GND_OK;
}

extern "C" {

void gnd_hip_scene_free(gnd_hip_ctx* ctx, gnd_hip_scene* s) {
    if (!s) return;
    if (ctx) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
    }
    for (void* p : s->allocations) (void)hipFree(p);
    delete s;
}

size_t gnd_hip_scene_bytes(const gnd_hip_scene* s) { return s ? s->bytes : 0; }

static int traceBatch(gnd_hip_ctx* ctx, gnd_hip_scene* scene, const gnd_ray_soa* rays, int64_t n,
                      const gnd_hit_soa* hits, int32_t* occ, bool any, gnd_hip_stats* stats) {
    if (!ctx || !scene || !rays || n < 0 || (!any && !hits) || (any && !occ)) return GND_E_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        GND_CHECK(hipSetDevice(ctx->device));
        if (n == 0) return;
        GND_CHECK(hipMemsetAsync(ctx->dCounters, 0, sizeof(Counters), ctx->stream));
        const int B = ctx->tuning.block_size, lds = normLds(ctx->tuning.stack_lds);
        int64_t blocks = (n + B - 1) / B;
        blocks = (blocks + 7) / 8 * 8;
        if (blocks > INT32_MAX) throw std::runtime_error("ray batch too large");
        gnd_hit_soa h{};
        if (hits) h = *hits;
        TraceKernel k = any ? (ctx->counters ? pickTrace<true, true>(lds) : pickTrace<true, false>(lds))
                            : (ctx->counters ? pickTrace<false, true>(lds) : pickTrace<false, false>(lds));
        GND_CHECK(hipEventRecord(ctx->ev[0], ctx->stream));
        hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(B), (size_t)B * lds * sizeof(int), ctx->stream,
                           scene->view, *rays, n, h, occ, ctx->dCounters, stackCapOf(ctx->tuning));
        GND_CHECK(hipGetLastError());
        GND_CHECK(hipEventRecord(ctx->ev[1], ctx->stream));
        Counters c;
        GND_CHECK(hipMemcpyAsync(&c, ctx->dCounters, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
        GND_CHECK(hipStreamSynchronize(ctx->stream));
        if (c.stackOverflow) throw HipError(GND_E_STACK, "traversal stack overflow");
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->ms_total = stats->ms_trace = elapsedMs(ctx->ev[0], ctx->ev[1]);
            stats->rays_closest = any ? 0 : (uint64_t)n;
            stats->rays_shadow = any ? (uint64_t)n : 0;
            stats->node_visits = c.nodes;
            stats->pack_tests = c.packs;
            stats->inst_tests = c.insts;
            stats->trace_launches = 1;
        }
    });
}

int gnd_hip_trace_closest(gnd_hip_ctx* ctx, gnd_hip_scene* scene, const gnd_ray_soa* rays, int64_t n,
                          const gnd_hit_soa* hits, gnd_hip_stats* stats) {
    return traceBatch(ctx, scene, rays, n, hits, nullptr, false, stats);
}

int gnd_hip_trace_any(gnd_hip_ctx* ctx, gnd_hip_scene* scene, const gnd_ray_soa* rays, int64_t n,
                      int32_t* occluded, gnd_hip_stats* stats) {
    return traceBatch(ctx, scene, rays, n, nullptr, occluded, true, stats);
}

int gnd_hip_cancel(gnd_hip_ctx* ctx) {
    if (!ctx) return GND_E_INVALID;
    ctx->cancel = 1;
    return GND_OK;
}

// ---------------------------------------------------------------- render

struct WavePlan {
    std::vector<DChunk> chunks;
    std::vector<DSplatTask> tasks;
    uint32_t slots = 0, pix = 0;
    uint32_t live = 0;  // paths with max depth > 0 (the first in-queue)
    int32_t maxDepth = 0;
};

// Bounces a job's paths may take: max_depth for path / volpath; the camera
// hit and one BSDF ray for direct; the camera hit for ao (their max_depth is
// not used, as in the reference).
static int jobDepth(const gnd_hip_job& j) {
    if (j.integrator == GND_INTEGRATOR_DIRECT) return 2;
    if (j.integrator == GND_INTEGRATOR_AO) return 1;
    return j.max_depth;
}

// Packs passes of the jobs into waves of at most `cap` slots.  A job's passes
// enter waves in pass order; every pass is whole (the film gather needs the
// neighbouring samples of the same pass).
static std::vector<WavePlan> planWaves(const std::vector<gnd_hip_job>& jobs, size_t cap) {
    std::vector<WavePlan> waves;
    std::vector<uint32_t> next(jobs.size(), 0);
    for (;;) {
        WavePlan w;
        bool any = false;
        for (size_t j = 0; j < jobs.size(); ++j) {
            const uint32_t npix = (uint32_t)jobs[j].camera.width * (uint32_t)jobs[j].camera.height;
            if (next[j] >= jobs[j].spp) continue;
            int first = (int)w.chunks.size();
            while (next[j] < jobs[j].spp && (w.slots + npix <= cap || (w.slots == 0))) {
                w.chunks.push_back({w.slots, npix, (int32_t)j, (int32_t)next[j], w.slots, -1, 0u});
                w.slots += npix;
                next[j]++;
                any = true;
                if (w.slots >= cap) break;
            }
            if ((int)w.chunks.size() > first) {
                w.tasks.push_back({w.pix, (int32_t)j, first, (int32_t)w.chunks.size()});
                w.pix += npix;
                w.maxDepth = std::max(w.maxDepth, jobDepth(jobs[j]));
            }
            if (w.slots >= cap) break;
        }
        if (!any) break;
        // in-queue layout of the camera paths: live chunks back to back
        for (DChunk& c : w.chunks) {
            c.qBase = w.live;
            if (jobDepth(jobs[c.job]) > 0) w.live += c.npix;
        }
        waves.push_back(std::move(w));
    }
    return waves;
}

// Kernels and launch shape of a bounce for one scene and job mix.
struct BouncePlan {
    WaveKernel trace = nullptr, shadow = nullptr, shade = nullptr, scatter = nullptr;
    bool split = false, fastShadow = false;
    int B = 128;
    int refill = 32, srefill = 40;
    int stackCap = kStackCap;
    WaveKernel tail = nullptr;  // kTail for this scene
    uint32_t tailPaths = 0;     // in-queue sizes at or below this run kTail
    unsigned tailGridCap = 8;
    size_t shm = 0;
    unsigned gridCap = 8;
    // blocks for n queue entries: one entry per thread up to the grid cap
    // (the traversal and shading kernels loop over their queue)
    unsigned grid(uint32_t n) const {
        unsigned b = (unsigned)((n + (uint32_t)B - 1) / (uint32_t)B);
        b = std::min(b, gridCap);
        return std::max(8u, (b + 7) / 8 * 8);
    }
    // kernels with `block` threads per block, the same thread cap
    unsigned gridFor(uint32_t n, unsigned block) const {
        const unsigned cap = std::max(8u, gridCap * (unsigned)B / block);
        unsigned b = (unsigned)((n + block - 1) / block);
        b = std::min(b, cap);
        return std::max(8u, (b + 7) / 8 * 8);
    }
    unsigned shadeGrid(uint32_t n) const { return gridFor(n, (unsigned)kShadeBlock); }
};

// What the jobs of a launch ask of the shading kernels.
struct JobMix {
    bool vol = false;     // a volpath job
    bool other = false;   // a direct or AO job
    bool aov = false;     // a job with film AOVs
    void add(const gnd_hip_job& j) {
        vol = vol || j.integrator == GND_INTEGRATOR_VOLPATH;
        other = other || j.integrator == GND_INTEGRATOR_DIRECT || j.integrator == GND_INTEGRATOR_AO;
        aov = aov || j.film_aov != nullptr;
    }
};

static BouncePlan makeBouncePlan(const gnd_hip_ctx* ctx, const gnd_hip_scene* scene, const JobMix& mix) {
    BouncePlan K;
    const bool anyVol = mix.vol;
    // the split NEE / scatter kernels keep transmittance sampling in the
    // reference order when shadow rays cross a density grid
    K.split = anyVol && scene->hasGrid;
    // path jobs without AOVs on a scene lit by area emitters only take the
    // lean shading kernels: path / volpath jobs without AOVs on a scene lit
    // by area emitters only
    const bool areaOnly = scene->view.nLights == 0 && scene->view.envLight < 0;
    const int feat = (!mix.other && !mix.aov && areaOnly) ? (mix.vol ? kFeatVol : kFeatPath) : kFeatAll;
    if (K.split) K.shade = feat == kFeatVol ? kShade<true, kFeatVol> : kShade<true, kFeatAll>;
    else if (feat == kFeatPath) K.shade = kShade<false, kFeatPath>;
    else K.shade = feat == kFeatVol ? kShade<false, kFeatVol> : kShade<false, kFeatAll>;
    K.scatter = feat == kFeatVol ? kScatter<kFeatVol> : kScatter<kFeatAll>;
    // any-hit shadows are exact when no surface is a null BSDF and no
    // shadow segment can cross a medium (the path integrator ignores media)
    K.fastShadow = !scene->hasIndex && (!anyVol || scene->nMedia == 0);
    K.B = ctx->tuning.block_size;
    const int lds = normLds(ctx->tuning.stack_lds);
    K.shm = (size_t)K.B * lds * sizeof(int);
    const bool cnt = ctx->counters;
    // moving primitives: the kernels that give each ray its time
    if (scene->hasMotion) pickBounceKernels<true>(K.fastShadow, cnt, lds, K.trace, K.shadow);
    else pickBounceKernels<false>(K.fastShadow, cnt, lds, K.trace, K.shadow);
#ifndef GND_GRID_WAVES_PER_CU
#define GND_GRID_WAVES_PER_CU 64
#endif
    // up to 64 waves per CU in flight: the traversal is latency bound
    K.gridCap = (unsigned)std::max(8, ctx->numCUs * GND_GRID_WAVES_PER_CU * 64 / K.B);
    // refill thresholds: measured best on a 36-triangle Cornell box (40 / 64;
    // closest hit 823 Msamples/s at 40, 811 at 16, 799 at 56) and the 1.95M-
    // triangle scene (16 / 16; 599 at 16, 592 at 24, 583 at 32; any hit 605 at
    // 16, 602 at 24, 600 at 40), with the block-reserving ray feed; a BVH
    // within one XCD's 4 MiB L2 counts as small
    const bool small = scene->accelBytes <= (size_t(4) << 20);
    K.refill = ctx->tuning.refill_lanes > 0 ? ctx->tuning.refill_lanes : (small ? 40 : 16);
    K.srefill = ctx->tuning.shadow_refill_lanes > 0 ? ctx->tuning.shadow_refill_lanes : (small ? 64 : 16);
    K.stackCap = stackCapOf(ctx->tuning);
    if (scene->hasMotion) K.tail = K.fastShadow ? kTail<kTailLds, true, true> : kTail<kTailLds, true, false>;
    else if (feat == kFeatPath)
        K.tail = K.fastShadow ? kTail<kTailLds, false, true, kFeatPath> : kTail<kTailLds, false, false, kFeatPath>;
    else if (feat == kFeatVol)
        K.tail = K.fastShadow ? kTail<kTailLds, false, true, kFeatVol> : kTail<kTailLds, false, false, kFeatVol>;
    else K.tail = K.fastShadow ? kTail<kTailLds, false, true> : kTail<kTailLds, false, false>;
    // resident lanes of the tail kernel: GND_TAIL_WAVES waves on each of 4 SIMDs per CU
    const uint32_t tailLanes = (uint32_t)ctx->numCUs * 4u * GND_TAIL_WAVES * 64u;
    K.tailGridCap = std::max(8u, tailLanes / (unsigned)K.B);
    const int32_t tp = ctx->tuning.tail_paths;
    // default: eight paths per resident lane (2M on 256 CUs at 4 waves per
    // SIMD), measured near the best on the 1080p scaled, volumetric and
    // Cornell-box frames
    K.tailPaths = tp < 0 ? 0u : (tp == 0 ? 8u * tailLanes : (uint32_t)tp);
    return K;
}

// Device-timed spans of the render stream (kind 0 closest hit, 1 shadow,
// 2 camera / shading / film, 3 tail kernel).
struct Timeline {
    hipStream_t stream;
    std::vector<hipEvent_t> evs;
    struct Span { size_t a, b; int kind; };
    std::vector<Span> spans;
    explicit Timeline(hipStream_t s) : stream(s) {}
    ~Timeline() {
        for (auto e : evs) (void)hipEventDestroy(e);
    }
    size_t mark() {
        hipEvent_t e;
        GND_CHECK(hipEventCreate(&e));
        GND_CHECK(hipEventRecord(e, stream));
        evs.push_back(e);
        return evs.size() - 1;
    }
    void lap(size_t& since, int kind) {
        const size_t now = mark();
        spans.push_back({since, now, kind});
        since = now;
    }
    void sum(double& trace, double& shadow, double& other, double& tail) const {
        for (const Span& sp : spans) {
            const double ms = elapsedMs(evs[sp.a], evs[sp.b]);
            (sp.kind == 0 ? trace : sp.kind == 1 ? shadow : sp.kind == 3 ? tail : other) += ms;
        }
    }
};

// in-queue `cur` of the two (with its rays and density-grid list), out-queue the other
static void bindQueues(WaveArgs& A, const QueueSet& qs, int cur) {
    A.qIn = qs.q[cur];
    A.qOut = qs.q[1 - cur];
    A.rIn = qs.rq[cur];
    A.rOut = qs.rq[1 - cur];
    A.gIn = qs.gl[cur];
    A.gOut = qs.gl[1 - cur];
}

// One bounce over the in-queue A.qIn (count in counts[0], at most `live`):
// closest hit, shading and NEE, shadow rays, the split scatter, then the queue
// rotation.  No host synchronisation, so the chain can be captured.
static void launchBounce(hipStream_t stream, const WaveArgs& A, const BouncePlan& K, uint32_t live, Timeline* tl) {
    const unsigned g = K.grid(live);
    size_t t = tl ? tl->mark() : 0;
    hipLaunchKernelGGL(K.trace, dim3(g), dim3(K.B), K.shm, stream, A);
    GND_CHECK(hipGetLastError());
    if (tl) tl->lap(t, 0);
    const unsigned gs = K.shadeGrid(live);
    if (K.split) {
        // density-grid delta tracking for the paths that start in a grid
        hipLaunchKernelGGL(kMedium, dim3(K.gridFor(live, 256)), dim3(256), 0, stream, A);
        GND_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(K.shade, dim3(gs), dim3(kShadeBlock), 0, stream, A);
    GND_CHECK(hipGetLastError());
    if (tl) tl->lap(t, 2);
    // NEE emits at most one shadow ray per path
    hipLaunchKernelGGL(K.shadow, dim3(g), dim3(K.B), K.shm, stream, A);
    GND_CHECK(hipGetLastError());
    if (tl) tl->lap(t, 1);
    if (K.split) {
        // the paths of this bounce with an event scatter after their shadow
        // rays consumed the transmittance samples
        hipLaunchKernelGGL(K.scatter, dim3(gs), dim3(kShadeBlock), 0, stream, A);
        GND_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(kRotate, dim3(1), dim3(64), 0, stream, A.counts, A.ctr);
    GND_CHECK(hipGetLastError());
    if (tl) tl->lap(t, 2);
}

// The rest of the in-queue's paths in one kTail launch, then the queue
// counters restart at zero (no path is left).
static void launchTail(hipStream_t stream, const WaveArgs& A, const BouncePlan& K, uint32_t live, Timeline* tl) {
    size_t t = tl ? tl->mark() : 0;
    const unsigned g = std::max(8u, std::min(K.tailGridCap, (live + (unsigned)K.B - 1) / (unsigned)K.B) / 8 * 8);
    hipLaunchKernelGGL(K.tail, dim3(g), dim3(K.B), (size_t)K.B * kTailLds * sizeof(int), stream, A);
    GND_CHECK(hipGetLastError());
    hipLaunchKernelGGL(kFillU32, dim3(1), dim3(64), 0, stream, A.counts, (size_t)kNumCounts, 0u);
    GND_CHECK(hipGetLastError());
    if (tl) tl->lap(t, 3);
}

// The instantiated chain of `depth` bounces starting from the queues bound in
// A.  A chain's nodes hold their kernel arguments and launch shapes by value,
// so a cached chain whose arguments and shapes are byte-identical is the same
// work: it is replayed instead of captured again (the key is exactly those
// bytes, so buffers reallocated elsewhere never match a stale chain).  Up to
// kGraphCache chains are kept, least recently used replaced first.
constexpr size_t kGraphCache = 32;

// The bytes a chain's launches are made of: the wave's kernel arguments, the
// first bounce's launch shapes and kernels.
static std::string chainKey(const WaveArgs& A0, const BouncePlan& K, uint32_t live) {
    std::string key;
    auto put = [&](const void* p, size_t n) { key.append(static_cast<const char*>(p), n); };
    const unsigned shape[6] = {K.grid(live), K.shadeGrid(live), K.gridFor(live, 256), (unsigned)K.B,
                               K.tailPaths, K.tailGridCap};
    const void* fns[5] = {(const void*)K.trace, (const void*)K.shadow, (const void*)K.tail, (const void*)K.shade,
                          (const void*)K.scatter};
    const int flags[2] = {K.split ? 1 : 0, (int)K.shm};
    put(&A0, sizeof(A0));
    put(shape, sizeof(shape));
    put(fns, sizeof(fns));
    put(flags, sizeof(flags));
    return key;
}

// `depth` bounces, then (with `tail`) the tail kernel over what is left.
static hipGraphExec_t bounceGraph(gnd_hip_ctx* ctx, const WaveArgs& A0, const BouncePlan& K, uint32_t live,
                                  int depth, bool tail) {
    std::string key = chainKey(A0, K, live);
    const int tag[2] = {depth, tail ? 1 : 0};
    key.append(reinterpret_cast<const char*>(tag), sizeof(tag));
    for (auto& c : ctx->graphs)
        if (c.key == key) {
            c.lastUse = ++ctx->graphClock;
            return c.exec;
        }
    hipGraph_t gr = nullptr;
    hipGraphExec_t ge = nullptr;
    WaveArgs A = A0;
    GND_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    try {
        for (int d = 0; d < depth; ++d) {  // bounce d + 1 reads the queues bounce d wrote
            launchBounce(ctx->stream, A, K, live, nullptr);
            std::swap(A.qIn, A.qOut);
            std::swap(A.rIn, A.rOut);
            std::swap(A.gIn, A.gOut);
        }
        // the tail kernel is persistent: its grid bounds the workers, not the paths
        if (tail) launchTail(ctx->stream, A, K, K.tailPaths, nullptr);
    } catch (...) {
        if (hipStreamEndCapture(ctx->stream, &gr) == hipSuccess && gr) (void)hipGraphDestroy(gr);
        throw;
    }
    GND_CHECK(hipStreamEndCapture(ctx->stream, &gr));
    const hipError_t ie = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(gr);
        GND_CHECK(ie);
    }
    if (ctx->graphs.size() >= kGraphCache) {
        auto victim = std::min_element(ctx->graphs.begin(), ctx->graphs.end(),
                                       [](const auto& a, const auto& b) { return a.lastUse < b.lastUse; });
        (void)hipGraphExecDestroy(victim->exec);
        (void)hipGraphDestroy(victim->graph);
        ctx->graphs.erase(victim);
    }
    ctx->graphs.push_back({key, gr, ge, ++ctx->graphClock});
    return ge;
}

static uint32_t readLive(gnd_hip_ctx* ctx, const uint32_t* counts) {
    uint32_t live = 0;
    GND_CHECK(hipMemcpyAsync(&live, counts, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    GND_CHECK(hipStreamSynchronize(ctx->stream));
    return live;
}

static DJob makeDJob(const gnd_hip_job& j) {
    DJob J;
    J.cam = makeCamera(j.camera);
    J.film = j.film;
    J.filmSq = j.film_sq;
    J.filmAov = j.film_aov;
    J.seedOffset = j.seed_offset;
    J.maxDepth = jobDepth(j);
    J.integrator = j.integrator;
    J.aoLength = j.ao_length > 0.0f ? j.ao_length : 5.0f;
    J.W = j.camera.width;
    J.H = j.camera.height;
    return J;
}

static void clearFilm(hipStream_t stream, const DJob& J) {
    const size_t np = (size_t)J.W * J.H;
    hipLaunchKernelGGL(kFill, dim3((unsigned)((np * 4 + 255) / 256)), dim3(256), 0, stream, J.film, np * 4, 0.0f);
    if (J.filmSq)
        hipLaunchKernelGGL(kFill, dim3((unsigned)((np * 3 + 255) / 256)), dim3(256), 0, stream, J.filmSq, np * 3,
                           0.0f);
    if (J.filmAov)
        hipLaunchKernelGGL(kFill, dim3((unsigned)((np * kAovFloats + 255) / 256)), dim3(256), 0, stream, J.filmAov,
                           np * kAovFloats, 0.0f);
}

static void fillStats(gnd_hip_stats* stats, const Counters& c, uint64_t samples) {
    stats->samples = samples;
    stats->rays_closest = c.rays + c.tailRays;
    stats->rays_shadow = c.srays + c.tailSrays;
    stats->rays_tail = c.tailRays;
    stats->trace_launches = c.launches;
    stats->node_visits = c.nodes;
    stats->pack_tests = c.packs;
    stats->inst_tests = c.insts;
    stats->shadow_node_visits = c.snodes;
    stats->shadow_pack_tests = c.spacks;
    stats->shadow_inst_tests = c.sinsts;
    stats->max_ray_nodes = c.maxRayNodes;
    stats->bad_rays = c.badRays;
    stats->lane_iterations = c.iters;
    stats->shadow_lane_iterations = c.siters;
    stats->lane_steps = c.steps;
    stats->shadow_lane_steps = c.ssteps;
    stats->nan_samples = c.nan;
}

int gnd_hip_render(gnd_hip_ctx* ctx, gnd_hip_scene* scene, const gnd_hip_job* jobsIn, int njobs,
                   gnd_hip_stats* stats) {
    if (!ctx || !scene || !jobsIn || njobs <= 0) return GND_E_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->cancel = 0;
    return guarded(ctx, [&] {
        GND_CHECK(hipSetDevice(ctx->device));
        std::vector<gnd_hip_job> jobs(jobsIn, jobsIn + njobs);
        size_t maxPix = 0;
        JobMix mix;
        uint64_t samples = 0;
        for (auto& j : jobs) {
            if (j.camera.width <= 0 || j.camera.height <= 0 || !j.film)
                throw std::runtime_error("job needs a positive resolution and a film");
            if (j.integrator < GND_INTEGRATOR_PATH || j.integrator > GND_INTEGRATOR_AO)
                throw std::runtime_error("unknown integrator");
            maxPix = std::max(maxPix, (size_t)j.camera.width * j.camera.height);
            mix.add(j);
            samples += (uint64_t)j.camera.width * j.camera.height * j.spp;
        }
        // slots for one wave: the capacity, or all the samples when fewer
        const size_t cap = std::max<size_t>(std::min<uint64_t>((uint64_t)ctx->tuning.queue_capacity, samples), maxPix);
        if (cap > 0x7FFFFFFFull) throw std::runtime_error("wave capacity too large");
        ensureWaveBuffers(ctx->wb, cap);
        for (int j = 0; j < njobs; ++j)
            if (jobs[j].film_aov) ensureAovBuffers(ctx->wb);
        std::vector<DJob> dj(jobs.size());
        for (size_t j = 0; j < jobs.size(); ++j) {
            dj[j] = makeDJob(jobs[j]);
            if (!jobs[j].accumulate) clearFilm(ctx->stream, dj[j]);
        }
        ensureDev(ctx->dJobs, ctx->jobCap, dj.size());
        GND_CHECK(hipMemcpyAsync(ctx->dJobs, dj.data(), dj.size() * sizeof(DJob), hipMemcpyHostToDevice,
                                 ctx->stream));
        GND_CHECK(hipMemsetAsync(ctx->dCounters, 0, sizeof(Counters), ctx->stream));
        const std::vector<WavePlan> waves = planWaves(jobs, cap);
        const BouncePlan K = makeBouncePlan(ctx, scene, mix);
        const bool graph = ctx->tuning.use_graph != 0;
        Timeline tl(ctx->stream);
        Timeline* tlp = graph ? nullptr : &tl;
        const size_t t0 = tl.mark();
        for (const WavePlan& w : waves) {
            if (ctx->cancel) throw HipError(GND_E_CANCELLED, "render cancelled");
            ensureDev(ctx->dChunks, ctx->chunkCap, w.chunks.size());
            ensureDev(ctx->dTasks, ctx->taskCap, w.tasks.size());
            GND_CHECK(hipMemcpyAsync(ctx->dChunks, w.chunks.data(), w.chunks.size() * sizeof(DChunk),
                                     hipMemcpyHostToDevice, ctx->stream));
            GND_CHECK(hipMemcpyAsync(ctx->dTasks, w.tasks.data(), w.tasks.size() * sizeof(DSplatTask),
                                     hipMemcpyHostToDevice, ctx->stream));
            hipLaunchKernelGGL(kFillU32, dim3(1), dim3(64), 0, ctx->stream, ctx->wb.Q.counts, (size_t)kNumCounts, 0u);
            WaveArgs A{};
            A.S = scene->view;
            A.jobs = ctx->dJobs;
            A.chunks = ctx->dChunks;
            A.nChunks = (int32_t)w.chunks.size();
            A.nSlots = w.slots;
            A.P = ctx->wb.P;
            A.Hq = ctx->wb.Q.Hq;
            A.Sh = ctx->wb.Q.Sh;
            A.counts = ctx->wb.Q.counts;
            A.chunkLive = nullptr;
            A.ctr = ctx->dCounters;
            A.pathOnlyFast = K.fastShadow ? 1 : 0;
            A.split = K.split ? 1 : 0;
            A.refill = K.refill;
            A.srefill = K.srefill;
            A.stackCap = K.stackCap;
            bindQueues(A, ctx->wb.Q, 0);
            size_t tc = tl.mark();
            A.camQueueBase = 0;
            hipLaunchKernelGGL(kCamera, dim3((w.slots + 255) / 256), dim3(256), 0, ctx->stream, A);
            GND_CHECK(hipGetLastError());
            hipLaunchKernelGGL(kFillU32, dim3(1), dim3(64), 0, ctx->stream, ctx->wb.Q.counts, (size_t)1, w.live);
            if (tlp) tl.lap(tc, 2);
            int cur = 0;
            uint32_t live = w.live;
            // with the hipGraph chain: the first wave of a shape runs the host
            // loop below and records how many bounces come before the tail
            // kernel; later waves of that shape launch those bounces and the
            // tail as one graph, with no host synchronisation.  Without a
            // tail kernel the chain is a fixed max depth + 1 bounces (kernels
            // exit at once on an empty queue) and paths still alive after it
            // continue in the loop.
            std::string learnKey;
            if (graph && K.tailPaths > 0) {
                bindQueues(A, ctx->wb.Q, cur);
                std::string key = chainKey(A, K, live);
                auto it = ctx->chainDepth.find(key);
                if (it != ctx->chainDepth.end()) {
                    GND_CHECK(hipGraphLaunch(bounceGraph(ctx, A, K, live, it->second, true), ctx->stream));
                    live = 0;
                } else {
                    learnKey.swap(key);
                }
            } else if (graph) {
                const int depth = std::max(1, w.maxDepth + 1);
                bindQueues(A, ctx->wb.Q, cur);
                GND_CHECK(hipGraphLaunch(bounceGraph(ctx, A, K, live, depth, false), ctx->stream));
                if (depth % 2) cur = 1 - cur;
                live = readLive(ctx, ctx->wb.Q.counts);
            }
            int bounces = 0;
            while (live > 0) {
                bindQueues(A, ctx->wb.Q, cur);
                if (live <= K.tailPaths) {
                    launchTail(ctx->stream, A, K, live, tlp);
                    live = 0;
                    break;
                }
                launchBounce(ctx->stream, A, K, live, tlp);
                cur = 1 - cur;
                bounces++;
                live = readLive(ctx, ctx->wb.Q.counts);
                if (ctx->cancel.load()) throw HipError(GND_E_CANCELLED, "render cancelled");
            }
            if (!learnKey.empty()) {
                if (ctx->chainDepth.size() >= 4096) ctx->chainDepth.clear();
                ctx->chainDepth[learnKey] = bounces;
            }
            size_t ts = tl.mark();
            hipLaunchKernelGGL(kSplat, dim3((w.pix + 255) / 256), dim3(256), 0, ctx->stream, A,
                               (const DSplatTask*)ctx->dTasks, (int)w.tasks.size(), w.pix);
            GND_CHECK(hipGetLastError());
            if (tlp) tl.lap(ts, 2);
        }
        const size_t t1 = tl.mark();
        Counters c;
        GND_CHECK(hipMemcpyAsync(&c, ctx->dCounters, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
        GND_CHECK(hipStreamSynchronize(ctx->stream));
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->ms_total = elapsedMs(tl.evs[t0], tl.evs[t1]);
            tl.sum(stats->ms_trace, stats->ms_shadow, stats->ms_shade, stats->ms_tail);
            fillStats(stats, c, samples);
        }
        if (c.stackOverflow) throw HipError(GND_E_STACK, "traversal stack overflow");
    });
}

int gnd_hip_denoise(gnd_hip_ctx* ctx, const float* color, const float* albedo, const float* normal, int width,
                    int height, float* output, double* ms) {
    if (!ctx || !color || !output || width <= 0 || height <= 0) return GND_E_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        GND_CHECK(hipSetDevice(ctx->device));
        const size_t np = (size_t)width * (size_t)height;
        if (np > ((size_t)1 << 28)) throw std::runtime_error("image too large to denoise");
        if (np > ctx->dnCap) {
            if (ctx->dnScratch) GND_CHECK(hipFree(ctx->dnScratch));
            ctx->dnScratch = nullptr;
            ctx->dnCap = 0;
            GND_CHECK(hipMalloc((void**)&ctx->dnScratch, 4 * np * sizeof(float4)));
            ctx->dnCap = np;
        }
        GND_CHECK(hipEventRecord(ctx->ev[0], ctx->stream));
        GND_CHECK(gndDenoiseLaunch(ctx->stream, color, albedo, normal, width, height, output, ctx->dnScratch, nullptr,
                                   nullptr));
        GND_CHECK(hipEventRecord(ctx->ev[1], ctx->stream));
        GND_CHECK(hipEventSynchronize(ctx->ev[1]));
        if (ms) *ms = elapsedMs(ctx->ev[0], ctx->ev[1]);
    });
}

int gnd_hip_denoise_noalloc(gnd_hip_ctx* ctx, const float* color, const float* albedo, const float* normal,
                            int width, int height, float* output, double* ms) {
    if (!ctx || !color || !output || width <= 0 || height <= 0) return GND_E_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    const size_t np = (size_t)width * (size_t)height;
    if (np > kDnStaticPixels) {
        ctx->err = "gnd_hip_denoise_noalloc: " + std::to_string(width) + "x" + std::to_string(height) +
                   " exceeds its static scratch of " + std::to_string(kDnStaticPixels) + " pixels";
        return GND_E_INVALID;
    }
    if (!ctx->dnStatic) {
        ctx->err = "gnd_hip_denoise_noalloc: the static scratch has no device address";
        return GND_E_DEVICE;
    }
    return guarded(ctx, [&] {
        GND_CHECK(hipSetDevice(ctx->device));
        GND_CHECK(hipEventRecord(ctx->ev[0], ctx->stream));
        GND_CHECK(gndDenoiseLaunch(ctx->stream, color, albedo, normal, width, height, output, ctx->dnStatic, nullptr,
                                   nullptr));
        GND_CHECK(hipEventRecord(ctx->ev[1], ctx->stream));
        GND_CHECK(hipEventSynchronize(ctx->ev[1]));
        if (ms) *ms = elapsedMs(ctx->ev[0], ctx->ev[1]);
    });
}

int gnd_hip_copy_probe(gnd_hip_ctx* ctx, size_t bytes, int repeats, double* ms_per_copy) {
    if (!ctx || bytes < 16 || repeats <= 0) return GND_E_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    void *a = nullptr, *b = nullptr;
    int rc = guarded(ctx, [&] {
        GND_CHECK(hipSetDevice(ctx->device));
        const size_t n = bytes / 16;
        GND_CHECK(hipMalloc(&a, n * 16));
        GND_CHECK(hipMalloc(&b, n * 16));
        hipLaunchKernelGGL(kFill, dim3((unsigned)((n * 4 + 255) / 256)), dim3(256), 0, ctx->stream, (float*)a,
                           n * 4, 1.0f);
        const unsigned grid = (unsigned)std::max(1, ctx->numCUs * 16);
        hipLaunchKernelGGL(kCopyProbe, dim3(grid), dim3(256), 0, ctx->stream, (const float4*)a, (float4*)b, n);
        GND_CHECK(hipEventRecord(ctx->ev[0], ctx->stream));
        for (int r = 0; r < repeats; ++r)
            hipLaunchKernelGGL(kCopyProbe, dim3(grid), dim3(256), 0, ctx->stream, (const float4*)a, (float4*)b, n);
        GND_CHECK(hipGetLastError());
        GND_CHECK(hipEventRecord(ctx->ev[1], ctx->stream));
        GND_CHECK(hipEventSynchronize(ctx->ev[1]));
        if (ms_per_copy) *ms_per_copy = elapsedMs(ctx->ev[0], ctx->ev[1]) / repeats;
    });
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int gnd_hip_autotune(gnd_hip_ctx* ctx, gnd_hip_scene* scene, const gnd_hip_job* job, gnd_hip_tuning* out) {
    if (!ctx || !scene || !job || !out || job->spp == 0 || job->camera.width <= 0 || job->camera.height <= 0)
        return GND_E_INVALID;
    // probe: the job itself, cut to at most one default wave of samples (its
    // film is overwritten), timed on the device (best of two) for every
    // candidate.  Coordinate search over block size x LDS stack depth, then
    // the closest-hit and the any-hit refill thresholds, then the wave
    // capacity (whole passes per wave), keeping the fastest setting of each
    // in turn.  The result is cached per scene and job shape.
    const uint64_t npix = (uint64_t)job->camera.width * (uint64_t)job->camera.height;
    gnd_hip_job probe = *job;
    probe.film_sq = nullptr;
    probe.spp = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(job->spp, kAutotuneSamples / npix));
    char keyBuf[160];
    std::snprintf(keyBuf, sizeof(keyBuf), "%p %d %d %u %d %d %d %d", (void*)scene, job->camera.width,
                  job->camera.height, probe.spp, job->integrator, job->max_depth, job->camera.type,
                  job->film_aov ? 1 : 0);
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        auto it = ctx->tuned.find(keyBuf);
        if (it != ctx->tuned.end()) {
            ctx->tuning = it->second;
            *out = it->second;
            return GND_OK;
        }
    }
    const gnd_hip_tuning start = ctx->tuning;
    gnd_hip_tuning best = start;
    double bestMs = 1e300;
    gnd_hip_stats st;
    int rc = gnd_hip_render(ctx, scene, &probe, 1, &st);  // warm-up: buffers, code objects
    auto tryTuning = [&](const gnd_hip_tuning& t) {
        if (rc != GND_OK) return;
        rc = gnd_hip_set_tuning(ctx, &t);
        double ms = 1e300;
        for (int r = 0; r < 2 && rc == GND_OK; ++r) {
            rc = gnd_hip_render(ctx, scene, &probe, 1, &st);
            if (rc == GND_OK) ms = std::min(ms, st.ms_total);
        }
        if (rc == GND_OK && ms < bestMs) {
            bestMs = ms;
            best = t;
        }
    };
    for (int b : {64, 128, 256})
        for (int l : {8, 16, 24, 32}) {
            gnd_hip_tuning t = start;
            t.block_size = b;
            t.stack_lds = l;
            tryTuning(t);
        }
    const gnd_hip_tuning shape = best;
    for (int r : {12, 16, 24, 32, 40, 48, 56}) {
        gnd_hip_tuning t = shape;
        t.refill_lanes = r;
        tryTuning(t);
    }
    const gnd_hip_tuning closest = best;
    for (int r : {16, 24, 32, 40, 48, 56, 64}) {
        gnd_hip_tuning t = closest;
        t.shadow_refill_lanes = r;
        tryTuning(t);
    }
    // the path count below which the tail kernel finishes a wave
    const gnd_hip_tuning shadows = best;
    for (int tp : {-1, 131072, 524288, 1048576, 2097152, 4194304, 8388608}) {
        gnd_hip_tuning t = shadows;
        t.tail_paths = tp;
        tryTuning(t);
    }
    // wave capacity: all of the probe's passes in one wave, or 1/2, 1/4, 1/8
    // of them per wave (at least one pass)
    const gnd_hip_tuning refills = best;
    for (int div : {1, 2, 4, 8}) {
        const uint64_t passes = std::max<uint64_t>(1, (probe.spp + div - 1) / div);
        gnd_hip_tuning t = refills;
        t.queue_capacity = (int64_t)std::max<uint64_t>(1024, passes * npix);
        if (t.queue_capacity == refills.queue_capacity) continue;
        tryTuning(t);
    }
    if (rc != GND_OK) {
        (void)gnd_hip_set_tuning(ctx, &start);
        return rc;
    }
    rc = gnd_hip_set_tuning(ctx, &best);
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        ctx->tuned[keyBuf] = ctx->tuning;
        *out = ctx->tuning;
    }
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------- serving engine
//
// Continuous batching.  A server owns one pool of path slots and several
// bounce loops ("lanes"), each with its own queues, counters and stream.  A
// lane renders the jobs of one resident scene, so lanes of different scenes
// run at the same time on the device.  Requests go into an inbox under a
// short lock and never wait for device work.  The worker thread admits
// passes of waiting jobs into free slots of their scene's lane whenever that
// lane is between bounces, launches one bounce per lane on the lane's stream
// without waiting for the others, and, as each lane's bounce completes,
// retires passes whose paths have all ended (per-chunk live counters) and
// gathers finished passes into their job's film in pass order, so a film is
// bit-identical whatever else shares the GPU.  Scenes stay resident in HBM
// until the budget needs their space (least recently used scene without a
// lane first).

using SrvClock = std::chrono::steady_clock;

struct SrvScene {
    const gnd_hip_scene_arrays* arrays = nullptr;  // borrowed until the server is destroyed
    gnd_hip_scene* dev = nullptr;                  // worker-owned
    size_t bytes = 0;
    uint64_t lastUse = 0;
    int lane = -1;  // lane rendering it, -1 none
};

struct SrvJob {
    uint64_t ticket = 0;
    int scene = -1;
    gnd_hip_job job{};
    // published under the server lock
    int32_t state = 0;  // 0 queued, 1 running, 2 done, < 0 failed (GND_E_*)
    uint32_t passesDone = 0;
    SrvClock::time_point tSubmit, tAdmit, tDone;
    bool cancelRequested = false;
    // owned by the worker
    int lane = -1;
    int slot = -1;                    // entry in the device job table
    uint32_t nextPass = 0;            // next pass to admit
    uint32_t nextGather = 0;          // next pass to gather into the film
    uint32_t chunksOut = 0;           // passes admitted and not yet gathered or dropped
    int32_t failCode = 0;             // != 0: stop, drop what is in flight, then fail with it
    SrvClock::time_point blockedSince{};  // first failed attempt to find slots for its next pass
    std::map<uint32_t, int> finished; // pass -> chunk whose paths have all ended
};

struct SrvChunk {
    SrvJob* job = nullptr;
    uint32_t pass = 0, slotBase = 0, npix = 0;
};

// D2H words a lane reads back after each bounce (pinned host memory):
// [0] live paths, [2..3] stack overflows (u64), [4 + c] live paths of chunk c
constexpr int kLaneReadHead = 4;

struct SrvLane {
    int scene = -1;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    QueueSet Q;
    Counters* dCtr = nullptr;
    BouncePlan K{};
    uint32_t live = 0;  // paths in the in-queue
    int cur = 0;
    bool inFlight = false;
    uint32_t readChunks = 0;  // chunk ids [0, readChunks) read back by the bounce in flight
    unsigned long long overflowSeen = 0;
    std::vector<int> running;        // chunks with paths in flight
    std::vector<SrvJob*> jobs;       // admitted jobs not yet finished
    std::vector<int> fresh;          // chunks admitted for the next launch
    std::vector<SrvJob*> freshJobs;  // jobs admitted for the next launch
    DChunk* dCam = nullptr;
    size_t camCap = 0;
    DChunk* dGather = nullptr;
    size_t gatherCap = 0;
    DSplatTask* dTasks = nullptr;
    size_t taskCap = 0;
    char* hStage = nullptr;  // pinned H2D staging, reused once the stream is idle
    size_t stageCap = 0;
    uint32_t* hRead = nullptr;  // pinned D2H words (kLaneReadHead + kSrvMaxChunks)
};

struct gnd_hip_server {
    gnd_hip_ctx* ctx = nullptr;
    gnd_hip_server_config cfg{};
    size_t pathBytes = 0;  // path state + the lanes' queues
    SrvClock::time_point t0;
    // ---- shared with callers, under mu
    std::mutex mu;
    std::condition_variable cv, done;
    bool stop = false;
    std::vector<SrvScene> scenes;  // appended by add_scene under mu; other fields worker-owned
    std::unordered_map<uint64_t, std::unique_ptr<SrvJob>> jobs;
    std::vector<SrvJob*> inbox;
    std::vector<uint64_t> cancels;
    uint64_t nextTicket = 1;
    uint64_t samplesDone = 0;
    int residentScenes = 0;
    size_t residentBytes = 0;
    // ---- owned by the worker
    std::thread worker;
    WaveBuffers wb;  // path state of the slot pool (queues live in the lanes)
    std::vector<SrvLane> lanes;
    std::deque<SrvJob*> pending;  // submitted, passes left to admit (submission order)
    uint64_t useClock = 0;
    DJob* dJobs = nullptr;
    std::vector<int> freeJobSlots;
    uint32_t* dLive = nullptr;  // live paths per chunk
    std::vector<SrvChunk> chunks;
    std::vector<int> freeChunks;
    std::map<uint32_t, uint32_t> freeSlots;  // begin -> end of free path-slot ranges
};

namespace {

constexpr int kSrvMaxChunks = 1 << 16;
constexpr int kSrvDefaultLanes = 4;
constexpr int kSrvMaxLanes = 16;

double msSince(const gnd_hip_server* s, SrvClock::time_point t) {
    return std::chrono::duration<double, std::milli>(t - s->t0).count();
}

bool takeSlots(gnd_hip_server* s, uint32_t n, uint32_t& base) {
    for (auto it = s->freeSlots.begin(); it != s->freeSlots.end(); ++it) {
        const uint32_t b = it->first, e = it->second;
        if (e - b < n) continue;
        s->freeSlots.erase(it);
        if (e - b > n) s->freeSlots[b + n] = e;
        base = b;
        return true;
    }
    return false;
}

void giveSlots(gnd_hip_server* s, uint32_t b, uint32_t n) {
    uint32_t e = b + n;
    auto next = s->freeSlots.lower_bound(b);
    if (next != s->freeSlots.end() && next->first == e) {
        e = next->second;
        next = s->freeSlots.erase(next);
    }
    if (next != s->freeSlots.begin()) {
        auto prev = std::prev(next);
        if (prev->second == b) {
            prev->second = e;
            return;
        }
    }
    s->freeSlots[b] = e;
}

size_t sceneEstimate(const gnd_hip_scene_arrays* a) {
    size_t b = (size_t)a->n_tlas_nodes * sizeof(DNode) + (size_t)a->n_blas_nodes * sizeof(DNode) +
               (size_t)a->n_packs * sizeof(DPack) + (size_t)a->n_tris * 24 * sizeof(float) +
               (size_t)a->n_cdf * sizeof(float) + (size_t)a->n_tlas_prims * sizeof(DInst);
    const gnd_scene_desc& d = *a->desc;
    b += (size_t)d.nprims * sizeof(DPrim) + (size_t)d.nmaterials * sizeof(DMaterial);
    for (uint32_t m = 0; m < d.nmedia; ++m)
        if (d.media[m].type == GND_MEDIUM_GRID) {
            // the dense grid with its majorants, and the corner records
            const size_t nv = (size_t)d.media[m].nx * d.media[m].ny * d.media[m].nz;
            b += nv * 4 * 2 + (nv * 32 <= kGridCornerBytes ? nv * 32 : 0);
        }
    return b;
}

size_t residentBytesOf(const gnd_hip_server* s) {
    size_t b = s->pathBytes + (s->wb.aovBuf ? s->wb.capacity * kAovFloats * sizeof(float) : 0);
    for (const SrvScene& sc : s->scenes)
        if (sc.dev) b += sc.bytes;
    return b;
}

void publishResidency(gnd_hip_server* s) {
    int n = 0;
    for (const SrvScene& sc : s->scenes) n += sc.dev ? 1 : 0;
    const size_t b = residentBytesOf(s);
    std::lock_guard<std::mutex> lk(s->mu);
    s->residentScenes = n;
    s->residentBytes = b;
}

// Makes scene `i` resident, evicting scenes no lane renders (least recently
// used first) while the budget would be exceeded.
void makeResident(gnd_hip_server* s, int i) {
    SrvScene& sc = s->scenes[i];
    sc.lastUse = ++s->useClock;
    if (sc.dev) return;
    const size_t need = sceneEstimate(sc.arrays);
    for (;;) {
        if (residentBytesOf(s) + need <= s->cfg.hbm_budget) break;
        int victim = -1;
        for (int k = 0; k < (int)s->scenes.size(); ++k) {
            const SrvScene& o = s->scenes[k];
            if (k == i || !o.dev || o.lane >= 0) continue;
            if (victim < 0 || o.lastUse < s->scenes[victim].lastUse) victim = k;
        }
        if (victim < 0) throw HipError(GND_E_NOMEM, "scene does not fit the server's HBM budget");
        gnd_hip_scene_free(s->ctx, s->scenes[victim].dev);
        s->scenes[victim].dev = nullptr;
    }
    gnd_hip_scene* dev = nullptr;
    const int rc = sceneUpload(s->ctx, sc.arrays, &dev);  // the worker holds the context lock
    if (rc != GND_OK) throw HipError(rc, std::string("scene upload: ") + s->ctx->err);
    sc.dev = dev;
    sc.bytes = gnd_hip_scene_bytes(dev);
    publishResidency(s);
}

// A state change to publish under the server lock (kSrvProgress: passes
// gathered so far, the state unchanged).
struct SrvEvent {
    SrvJob* job;
    int32_t state;
    SrvClock::time_point t;
};
constexpr int32_t kSrvProgress = 100;

void publish(gnd_hip_server* s, const std::vector<SrvEvent>& evs) {
    if (evs.empty()) return;
    std::lock_guard<std::mutex> lk(s->mu);
    for (const SrvEvent& e : evs) {
        SrvJob* j = e.job;
        if (e.state == kSrvProgress) {
            j->passesDone = j->nextGather;
            continue;
        }
        if (e.state == 1) {
            j->tAdmit = e.t;
        } else {
            if (j->state == 0) j->tAdmit = e.t;
            j->tDone = e.t;
            if (e.state == 2)
                s->samplesDone += (uint64_t)j->job.camera.width * (uint64_t)j->job.camera.height * j->job.spp;
        }
        j->state = e.state;
        j->passesDone = j->nextGather;
    }
    s->done.notify_all();
}

// Removes a job from its lane's lists and the job table, as it ends.
void endJob(gnd_hip_server* s, SrvJob* j) {
    if (j->lane >= 0) {
        auto& v = s->lanes[j->lane].jobs;
        v.erase(std::remove(v.begin(), v.end(), j), v.end());
        auto& f = s->lanes[j->lane].freshJobs;
        f.erase(std::remove(f.begin(), f.end(), j), f.end());
    }
    if (j->slot >= 0) s->freeJobSlots.push_back(j->slot);
    j->slot = -1;
}

void releaseChunk(gnd_hip_server* s, int c) {
    SrvChunk& ch = s->chunks[c];
    giveSlots(s, ch.slotBase, ch.npix);
    if (ch.job) ch.job->chunksOut--;
    ch = SrvChunk{};
    s->freeChunks.push_back(c);
}

bool laneIdle(const SrvLane& L) {
    return !L.inFlight && L.live == 0 && L.running.empty() && L.jobs.empty() && L.fresh.empty();
}

// The lane of scene `sc`: the one rendering it, else an idle lane (the scene
// is made resident first), else -1.
int laneFor(gnd_hip_server* s, int sc) {
    if (s->scenes[sc].lane >= 0) return s->scenes[sc].lane;
    for (int l = 0; l < (int)s->lanes.size(); ++l) {
        SrvLane& L = s->lanes[l];
        if (!laneIdle(L)) continue;
        if (L.scene >= 0) s->scenes[L.scene].lane = -1;
        L.scene = -1;
        makeResident(s, sc);
        L.scene = sc;
        s->scenes[sc].lane = l;
        return l;
    }
    return -1;
}

// Admission between bounces.  Passes go to waiting jobs one per job per
// round, in submission order, until no more fit, so a job that arrives while
// a long one runs gets slots as soon as passes retire.  A job waits (without
// holding back the others) while its scene has no lane, its lane is
// mid-bounce or its pass does not fit; once the oldest waiting job has found
// no room for longer than kSrvReserve, later jobs stop taking slots until it
// fits (a large pass is not starved by small ones).
constexpr auto kSrvReserve = std::chrono::milliseconds(20);

void admit(gnd_hip_server* s, std::vector<SrvEvent>& evs) {
    const auto now = SrvClock::now();
    auto drop = [&](SrvJob* j) { s->pending.erase(std::remove(s->pending.begin(), s->pending.end(), j), s->pending.end()); };
    std::vector<SrvJob*> gone;
    for (SrvJob* j : s->pending) {
        if (!j->failCode) continue;
        // cancelled or failed: no more passes; a job with a lane ends there
        // once its passes in flight are dropped
        if (j->lane < 0) evs.push_back({j, j->failCode, now});
        gone.push_back(j);
    }
    for (SrvJob* j : gone) drop(j);
    bool reserve = false;
    for (bool any = true; any && !reserve;) {
        any = false;
        gone.clear();
        bool oldest = true;
        for (SrvJob* j : s->pending) {
            const bool first = oldest;
            oldest = false;
            int l = j->lane;
            if (l < 0) {
                try {
                    l = laneFor(s, j->scene);
                } catch (const HipError& e) {
                    s->ctx->err = e.what();
                    evs.push_back({j, e.code, now});
                    gone.push_back(j);
                    continue;
                }
                if (l < 0 || s->freeJobSlots.empty()) continue;
            }
            SrvLane& L = s->lanes[l];
            if (L.inFlight) continue;
            const uint32_t npix = (uint32_t)j->job.camera.width * (uint32_t)j->job.camera.height;
            uint32_t base = 0;
            if (s->freeChunks.empty() || !takeSlots(s, npix, base)) {
                if (j->blockedSince == SrvClock::time_point{}) j->blockedSince = now;
                else if (first && now - j->blockedSince > kSrvReserve) {
                    reserve = true;
                    break;
                }
                continue;
            }
            j->blockedSince = SrvClock::time_point{};
            if (j->lane < 0) {
                if (j->job.film_aov && !s->wb.aovBuf) {
                    ensureAovBuffers(s->wb);
                    publishResidency(s);
                }
                j->lane = l;
                j->slot = s->freeJobSlots.back();
                s->freeJobSlots.pop_back();
                L.jobs.push_back(j);
                L.freshJobs.push_back(j);
                evs.push_back({j, 1, now});
            }
            const int c = s->freeChunks.back();
            s->freeChunks.pop_back();
            s->chunks[c] = SrvChunk{j, j->nextPass, base, npix};
            j->chunksOut++;
            L.fresh.push_back(c);
            j->nextPass++;
            any = true;
            if (j->nextPass == j->job.spp) gone.push_back(j);
        }
        for (SrvJob* j : gone) drop(j);
    }
}

WaveArgs laneArgs(gnd_hip_server* s, const SrvLane& L) {
    WaveArgs A{};
    A.S = s->scenes[L.scene].dev->view;
    A.jobs = s->dJobs;
    A.P = s->wb.P;
    A.Hq = L.Q.Hq;
    A.Sh = L.Q.Sh;
    A.counts = L.Q.counts;
    A.chunkLive = s->dLive;
    A.ctr = L.dCtr;
    A.pathOnlyFast = L.K.fastShadow ? 1 : 0;
    A.split = L.K.split ? 1 : 0;
    A.refill = L.K.refill;
    A.srefill = L.K.srefill;
    A.stackCap = L.K.stackCap;
    return A;
}

void ensureStage(SrvLane& L, size_t bytes) {
    if (bytes <= L.stageCap) return;
    if (L.hStage) GND_CHECK(hipHostFree(L.hStage));
    L.hStage = nullptr;
    L.stageCap = 0;
    const size_t n = std::max(bytes, L.stageCap * 2);
    GND_CHECK(hipHostMalloc((void**)&L.hStage, n, hipHostMallocDefault));
    L.stageCap = n;
}

// Launches one step of an idle lane on its stream: records and film clears of
// newly admitted jobs, camera rays of newly admitted passes, one bounce over
// every live path, and the read-back of the counts; no host wait.
void launchLane(gnd_hip_server* s, SrvLane& L) {
    gnd_hip_ctx* ctx = s->ctx;
    JobMix mix;
    for (SrvJob* j : L.jobs) mix.add(j->job);
    const gnd_hip_scene* scene = s->scenes[L.scene].dev;
    L.K = makeBouncePlan(ctx, scene, mix);
    WaveArgs A = laneArgs(s, L);
    if (!L.fresh.empty()) {
        const size_t nj = L.freshJobs.size(), nc = L.fresh.size();
        ensureStage(L, nj * sizeof(DJob) + nc * (sizeof(DChunk) + sizeof(uint32_t)) + 64);
        DJob* sj = reinterpret_cast<DJob*>(L.hStage);
        uint32_t* sl = reinterpret_cast<uint32_t*>(sj + nj);
        DChunk* sc = reinterpret_cast<DChunk*>(sl + nc + (nc & 1));
        for (size_t k = 0; k < nj; ++k) {
            SrvJob* j = L.freshJobs[k];
            sj[k] = makeDJob(j->job);
            GND_CHECK(hipMemcpyAsync(s->dJobs + j->slot, &sj[k], sizeof(DJob), hipMemcpyHostToDevice, L.stream));
            if (!j->job.accumulate) clearFilm(L.stream, sj[k]);
        }
        uint32_t n = 0, nLive = 0;
        for (size_t k = 0; k < nc; ++k) {
            const int c = L.fresh[k];
            const SrvChunk& ch = s->chunks[c];
            const bool lives = jobDepth(ch.job->job) > 0;
            sc[k] = DChunk{ch.slotBase, ch.npix, ch.job->slot, (int32_t)ch.pass, n, c, nLive};
            n += ch.npix;
            if (lives) nLive += ch.npix;
            sl[k] = lives ? ch.npix : 0u;
            GND_CHECK(hipMemcpyAsync(s->dLive + c, &sl[k], sizeof(uint32_t), hipMemcpyHostToDevice, L.stream));
            L.running.push_back(c);
        }
        ensureDev(L.dCam, L.camCap, nc);
        GND_CHECK(hipMemcpyAsync(L.dCam, sc, nc * sizeof(DChunk), hipMemcpyHostToDevice, L.stream));
        A.chunks = L.dCam;
        A.nChunks = (int32_t)nc;
        A.nSlots = n;
        bindQueues(A, L.Q, L.cur);
        A.camQueueBase = L.live;  // after the survivors of the last bounce
        hipLaunchKernelGGL(kCamera, dim3((n + 255) / 256), dim3(256), 0, L.stream, A);
        GND_CHECK(hipGetLastError());
        L.live += nLive;
        hipLaunchKernelGGL(kFillU32, dim3(1), dim3(64), 0, L.stream, L.Q.counts, (size_t)1, L.live);
        L.fresh.clear();
        L.freshJobs.clear();
    }
    if (L.live > 0) {
        bindQueues(A, L.Q, L.cur);
        if (L.live <= L.K.tailPaths) {
            launchTail(L.stream, A, L.K, L.live, nullptr);
        } else {
            launchBounce(L.stream, A, L.K, L.live, nullptr);
            L.cur = 1 - L.cur;
        }
    }
    uint32_t used = 0;  // chunk ids are handed out lowest first
    for (int c : L.running) used = std::max(used, (uint32_t)c + 1u);
    L.readChunks = used;
    GND_CHECK(hipMemcpyAsync(L.hRead, L.Q.counts, sizeof(uint32_t), hipMemcpyDeviceToHost, L.stream));
    GND_CHECK(hipMemcpyAsync(L.hRead + 2, &L.dCtr->stackOverflow, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             L.stream));
    if (used)
        GND_CHECK(hipMemcpyAsync(L.hRead + kLaneReadHead, s->dLive, (size_t)used * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, L.stream));
    GND_CHECK(hipEventRecord(L.done, L.stream));
    L.inFlight = true;
}

// After a lane's step completed: stack-overflow check, retirement, the film
// gather of finished passes in pass order, and the end of finished jobs.
void finishLane(gnd_hip_server* s, SrvLane& L, std::vector<SrvEvent>& evs) {
    L.inFlight = false;
    L.live = L.hRead[0];
    unsigned long long overflow = 0;
    std::memcpy(&overflow, L.hRead + 2, sizeof(overflow));
    if (overflow > L.overflowSeen) {
        // as gnd_hip_render: the images of the jobs this bounce served are not valid
        L.overflowSeen = overflow;
        for (SrvJob* j : L.jobs)
            if (!j->failCode) j->failCode = GND_E_STACK;
        s->ctx->err = "traversal stack overflow";
    }
    std::vector<int> still;
    for (int c : L.running) {
        if ((uint32_t)c < L.readChunks && L.hRead[kLaneReadHead + c] == 0) {
            SrvChunk& ch = s->chunks[c];
            ch.job->finished[ch.pass] = c;
        } else {
            still.push_back(c);
        }
    }
    L.running.swap(still);
    // gather finished passes in pass order; a failing job's passes are dropped
    std::vector<DChunk> gc;
    std::vector<DSplatTask> tasks;
    std::vector<int> released;
    uint32_t pix = 0;
    for (SrvJob* j : L.jobs) {
        if (j->failCode) {
            for (auto& kv : j->finished) released.push_back(kv.second);
            j->finished.clear();
            continue;
        }
        const int first = (int)gc.size();
        for (auto it = j->finished.find(j->nextGather); it != j->finished.end();
             it = j->finished.find(j->nextGather)) {
            const SrvChunk& ch = s->chunks[it->second];
            gc.push_back({ch.slotBase, ch.npix, j->slot, (int32_t)ch.pass, 0u, it->second, 0u});
            released.push_back(it->second);
            j->finished.erase(it);
            j->nextGather++;
        }
        if ((int)gc.size() > first) {
            tasks.push_back({pix, j->slot, first, (int32_t)gc.size()});
            pix += (uint32_t)j->job.camera.width * (uint32_t)j->job.camera.height;
            if (j->nextGather < j->job.spp) evs.push_back({j, kSrvProgress, SrvClock::now()});
        }
    }
    if (!tasks.empty()) {
        const size_t bytes = gc.size() * sizeof(DChunk) + tasks.size() * sizeof(DSplatTask);
        ensureStage(L, bytes);
        std::memcpy(L.hStage, gc.data(), gc.size() * sizeof(DChunk));
        std::memcpy(L.hStage + gc.size() * sizeof(DChunk), tasks.data(), tasks.size() * sizeof(DSplatTask));
        ensureDev(L.dGather, L.gatherCap, gc.size());
        ensureDev(L.dTasks, L.taskCap, tasks.size());
        GND_CHECK(hipMemcpyAsync(L.dGather, L.hStage, gc.size() * sizeof(DChunk), hipMemcpyHostToDevice, L.stream));
        GND_CHECK(hipMemcpyAsync(L.dTasks, L.hStage + gc.size() * sizeof(DChunk), tasks.size() * sizeof(DSplatTask),
                                 hipMemcpyHostToDevice, L.stream));
        WaveArgs A = laneArgs(s, L);
        A.chunks = L.dGather;
        A.nChunks = (int32_t)gc.size();
        hipLaunchKernelGGL(kSplat, dim3((pix + 255) / 256), dim3(256), 0, L.stream, A, (const DSplatTask*)L.dTasks,
                           (int)tasks.size(), pix);
        GND_CHECK(hipGetLastError());
        GND_CHECK(hipStreamSynchronize(L.stream));
    }
    for (int c : released) releaseChunk(s, c);
    const auto now = SrvClock::now();
    std::vector<SrvJob*> ended;
    for (SrvJob* j : L.jobs) {
        if (j->failCode ? j->chunksOut == 0 : j->nextGather == j->job.spp) ended.push_back(j);
    }
    for (SrvJob* j : ended) {
        if (j->failCode) {
            // passes not admitted yet are dropped with the job
            s->pending.erase(std::remove(s->pending.begin(), s->pending.end(), j), s->pending.end());
        }
        endJob(s, j);
        evs.push_back({j, j->failCode ? j->failCode : 2, now});
    }
    if (laneIdle(L) && L.scene >= 0) {
        s->scenes[L.scene].lane = -1;
        L.scene = -1;
    }
}

// A device error: every job with passes in flight fails, the lanes are reset
// and their slots returned; queued jobs stay queued.
void failInFlight(gnd_hip_server* s, int code, std::vector<SrvEvent>& evs) {
    const auto now = SrvClock::now();
    for (SrvLane& L : s->lanes) {
        (void)hipStreamSynchronize(L.stream);
        std::vector<SrvJob*> js = L.jobs;
        for (SrvJob* j : js) {
            j->failCode = code;
            s->pending.erase(std::remove(s->pending.begin(), s->pending.end(), j), s->pending.end());
            endJob(s, j);
            evs.push_back({j, code, now});
        }
        L.jobs.clear();
        L.running.clear();
        L.fresh.clear();
        L.freshJobs.clear();
        L.live = 0;
        L.cur = 0;
        L.inFlight = false;
        hipLaunchKernelGGL(kFillU32, dim3(1), dim3(64), 0, L.stream, L.Q.counts, (size_t)kNumCounts, 0u);
        (void)hipStreamSynchronize(L.stream);
        if (L.scene >= 0) s->scenes[L.scene].lane = -1;
        L.scene = -1;
    }
    for (int c = 0; c < (int)s->chunks.size(); ++c)
        if (s->chunks[c].job) releaseChunk(s, c);
}

bool workerBusy(const gnd_hip_server* s) {
    if (!s->pending.empty()) return true;
    for (const SrvLane& L : s->lanes)
        if (!laneIdle(L)) return true;
    return false;
}

void serverLoop(gnd_hip_server* s) {
    gnd_hip_ctx* ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    for (;;) {
        std::vector<SrvEvent> evs;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (!s->stop && s->inbox.empty() && s->cancels.empty() && !workerBusy(s))
                s->cv.wait(lk, [&] { return s->stop || !s->inbox.empty() || !s->cancels.empty(); });
            if (s->stop) break;
            for (SrvJob* j : s->inbox) s->pending.push_back(j);
            s->inbox.clear();
            for (uint64_t t : s->cancels) {
                auto it = s->jobs.find(t);
                if (it == s->jobs.end()) continue;
                SrvJob* j = it->second.get();
                if ((j->state == 0 || j->state == 1) && !j->failCode) j->failCode = GND_E_CANCELLED;
            }
            s->cancels.clear();
        }
        try {
            {
                std::lock_guard<std::mutex> g(ctx->mu);
                admit(s, evs);
                // cancelled jobs of a lane between bounces that hold no chunks end here
                for (SrvLane& L : s->lanes) {
                    if (L.inFlight) continue;
                    std::vector<SrvJob*> ended;
                    for (SrvJob* j : L.jobs)
                        if (j->failCode && j->chunksOut == 0) ended.push_back(j);
                    for (SrvJob* j : ended) {
                        endJob(s, j);
                        evs.push_back({j, j->failCode, SrvClock::now()});
                    }
                }
                for (SrvLane& L : s->lanes) {
                    if (L.scene >= 0 && !L.inFlight && (!L.fresh.empty() || L.live > 0)) launchLane(s, L);
                    if (laneIdle(L) && L.scene >= 0) {
                        s->scenes[L.scene].lane = -1;
                        L.scene = -1;
                    }
                }
            }
            publish(s, evs);
            evs.clear();
            // wait for the first lane to finish its step
            bool any = false, finished = false;
            for (SrvLane& L : s->lanes) {
                if (!L.inFlight) continue;
                any = true;
                const hipError_t q = hipEventQuery(L.done);
                if (q == hipErrorNotReady) continue;
                GND_CHECK(q);
                std::lock_guard<std::mutex> g(ctx->mu);
                finishLane(s, L, evs);
                finished = true;
            }
            publish(s, evs);
            if (any && !finished) std::this_thread::sleep_for(std::chrono::microseconds(20));
        } catch (const std::exception& e) {
            const int code = dynamic_cast<const HipError*>(&e) ? static_cast<const HipError&>(e).code : GND_E_DEVICE;
            std::vector<SrvEvent> fail;
            {
                std::lock_guard<std::mutex> g(ctx->mu);
                ctx->err = e.what();
                failInFlight(s, code, fail);
            }
            publish(s, evs);
            publish(s, fail);
        }
    }
    for (SrvLane& L : s->lanes) (void)hipStreamSynchronize(L.stream);
}

void statusOf(const gnd_hip_server* s, const SrvJob& j, gnd_hip_job_status* st) {
    using ms = std::chrono::duration<double, std::milli>;
    const bool admitted = j.state != 0, ended = j.state == 2 || j.state < 0;
    st->state = j.state;
    st->passes_done = j.passesDone;
    st->queue_ms = admitted ? ms(j.tAdmit - j.tSubmit).count() : 0.0;
    st->run_ms = j.state == 2 ? ms(j.tDone - j.tAdmit).count() : 0.0;
    st->latency_ms = ended ? ms(j.tDone - j.tSubmit).count() : 0.0;
    st->t_submit_ms = msSince(s, j.tSubmit);
    st->t_admit_ms = admitted ? msSince(s, j.tAdmit) : -1.0;
    st->t_done_ms = ended ? msSince(s, j.tDone) : -1.0;
}

void destroyLanes(gnd_hip_server* s) {
    for (SrvLane& L : s->lanes) {
        if (L.stream) (void)hipStreamSynchronize(L.stream);
        freeQueueSet(L.Q);
        for (void* p : {(void*)L.dCtr, (void*)L.dCam, (void*)L.dGather, (void*)L.dTasks})
            if (p) (void)hipFree(p);
        if (L.hStage) (void)hipHostFree(L.hStage);
        if (L.hRead) (void)hipHostFree(L.hRead);
        if (L.done) (void)hipEventDestroy(L.done);
        if (L.stream) (void)hipStreamDestroy(L.stream);
    }
    s->lanes.clear();
}

}  // namespace

extern "C" {

int gnd_hip_server_create(gnd_hip_ctx* ctx, const gnd_hip_server_config* cfg, gnd_hip_server** out) {
    if (!ctx || !cfg || !out || cfg->path_slots < 1024 || cfg->max_jobs < 1 || cfg->lanes > kSrvMaxLanes)
        return GND_E_INVALID;
    gnd_hip_server* s = new gnd_hip_server();
    s->ctx = ctx;
    s->cfg = *cfg;
    s->t0 = SrvClock::now();
    if (s->cfg.hbm_budget == 0) s->cfg.hbm_budget = ctx->hbm / 10 * 9;
    if (s->cfg.lanes == 0) s->cfg.lanes = kSrvDefaultLanes;
    const size_t slots = cfg->path_slots;
    s->pathBytes = pathStateBytes(slots) + s->cfg.lanes * queueSetBytes(slots);
    int rc = guarded(ctx, [&] {
        std::lock_guard<std::mutex> g(ctx->mu);
        GND_CHECK(hipSetDevice(ctx->device));
        if (s->pathBytes > s->cfg.hbm_budget) throw HipError(GND_E_NOMEM, "path slots exceed the HBM budget");
        ensureWaveBuffers(s->wb, slots, false);
        GND_CHECK(hipMalloc(&s->dJobs, (size_t)cfg->max_jobs * sizeof(DJob)));
        GND_CHECK(hipMalloc(&s->dLive, (size_t)kSrvMaxChunks * sizeof(uint32_t)));
        s->lanes.resize(s->cfg.lanes);
        for (SrvLane& L : s->lanes) {
            GND_CHECK(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
            GND_CHECK(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
            ensureQueueSet(L.Q, slots);
            GND_CHECK(hipMalloc(&L.dCtr, sizeof(Counters)));
            GND_CHECK(hipMemsetAsync(L.dCtr, 0, sizeof(Counters), L.stream));
            GND_CHECK(hipHostMalloc((void**)&L.hRead, (kLaneReadHead + (size_t)kSrvMaxChunks) * sizeof(uint32_t),
                                    hipHostMallocDefault));
            hipLaunchKernelGGL(kFillU32, dim3(1), dim3(64), 0, L.stream, L.Q.counts, (size_t)kNumCounts, 0u);
            GND_CHECK(hipStreamSynchronize(L.stream));
        }
    });
    if (rc != GND_OK) {
        destroyLanes(s);
        freeWaveBuffers(s->wb);
        if (s->dJobs) (void)hipFree(s->dJobs);
        if (s->dLive) (void)hipFree(s->dLive);
        delete s;
        return rc;
    }
    for (int i = (int)cfg->max_jobs - 1; i >= 0; --i) s->freeJobSlots.push_back(i);
    s->chunks.resize(kSrvMaxChunks);
    for (int c = kSrvMaxChunks - 1; c >= 0; --c) s->freeChunks.push_back(c);
    s->freeSlots[0] = cfg->path_slots;
    s->residentBytes = residentBytesOf(s);
    s->worker = std::thread(serverLoop, s);
    *out = s;
    return GND_OK;
}

void gnd_hip_server_destroy(gnd_hip_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    gnd_hip_ctx* ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    destroyLanes(s);
    for (SrvScene& sc : s->scenes)
        if (sc.dev) gnd_hip_scene_free(ctx, sc.dev);
    freeWaveBuffers(s->wb);
    for (void* p : {(void*)s->dJobs, (void*)s->dLive})
        if (p) (void)hipFree(p);
    delete s;
}

int gnd_hip_server_add_scene(gnd_hip_server* s, const gnd_hip_scene_arrays* arrays, int32_t* scene_id) {
    if (!s || !arrays || !arrays->desc || !scene_id) return GND_E_INVALID;
    // the worker reads the scene list without the lock: pause it while the
    // list may reallocate (it holds the context lock across each step)
    std::lock_guard<std::mutex> g(s->ctx->mu);
    std::lock_guard<std::mutex> lk(s->mu);
    SrvScene sc;
    sc.arrays = arrays;
    s->scenes.push_back(sc);
    *scene_id = (int32_t)s->scenes.size() - 1;
    return GND_OK;
}

int gnd_hip_server_submit(gnd_hip_server* s, int32_t scene_id, const gnd_hip_job* job, uint64_t* ticket) {
    const auto arrival = SrvClock::now();  // before any lock: latency runs from here
    if (!s || !job || !ticket || job->spp == 0 || !job->film || job->camera.width <= 0 || job->camera.height <= 0 ||
        job->integrator < GND_INTEGRATOR_PATH || job->integrator > GND_INTEGRATOR_AO)
        return GND_E_INVALID;
    if ((uint64_t)job->camera.width * (uint64_t)job->camera.height > s->cfg.path_slots) {
        std::lock_guard<std::mutex> g(s->ctx->mu);
        s->ctx->err = "one pass of the job needs more path slots than the server has";
        return GND_E_NOMEM;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (scene_id < 0 || scene_id >= (int32_t)s->scenes.size()) return GND_E_INVALID;
    const uint64_t t = s->nextTicket++;
    auto j = std::make_unique<SrvJob>();
    j->ticket = t;
    j->scene = scene_id;
    j->job = *job;
    j->tSubmit = arrival;
    s->inbox.push_back(j.get());
    s->jobs.emplace(t, std::move(j));
    *ticket = t;
    s->cv.notify_all();
    return GND_OK;
}

int gnd_hip_server_poll(gnd_hip_server* s, uint64_t ticket, gnd_hip_job_status* st) {
    if (!s || !st) return GND_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return GND_E_INVALID;
    statusOf(s, *it->second, st);
    return GND_OK;
}

int gnd_hip_server_wait(gnd_hip_server* s, uint64_t ticket, double timeout_s, gnd_hip_job_status* st) {
    if (!s || !st) return GND_E_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return GND_E_INVALID;
    SrvJob* j = it->second.get();
    const double secs = std::min(std::max(0.0, timeout_s), 1e7);  // NaN -> 0
    const auto until = SrvClock::now() + std::chrono::duration_cast<SrvClock::duration>(
                                             std::chrono::duration<double>(secs));
    const bool ended = s->done.wait_until(lk, until, [&] { return j->state == 2 || j->state < 0; });
    statusOf(s, *j, st);
    return ended ? GND_OK : GND_E_TIMEOUT;
}

int gnd_hip_server_cancel(gnd_hip_server* s, uint64_t ticket) {
    if (!s) return GND_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return GND_E_INVALID;
    SrvJob* j = it->second.get();
    if (j->state == 0 || j->state == 1) {
        j->cancelRequested = true;
        s->cancels.push_back(ticket);
        s->cv.notify_all();
    }
    return GND_OK;
}

int gnd_hip_server_release(gnd_hip_server* s, uint64_t ticket) {
    if (!s) return GND_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return GND_E_INVALID;
    if (it->second->state != 2 && it->second->state >= 0) return GND_E_BUSY;
    s->jobs.erase(it);
    return GND_OK;
}

int gnd_hip_server_stats(gnd_hip_server* s, gnd_hip_stats* stats, int32_t* resident_scenes, size_t* resident_bytes) {
    if (!s) return GND_E_INVALID;
    uint64_t samples = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        samples = s->samplesDone;
        if (resident_scenes) *resident_scenes = s->residentScenes;
        if (resident_bytes) *resident_bytes = s->residentBytes;
    }
    if (!stats) return GND_OK;
    std::lock_guard<std::mutex> g(s->ctx->mu);
    return guarded(s->ctx, [&] {
        GND_CHECK(hipSetDevice(s->ctx->device));
        // the lanes' counters summed (kernels in flight may still add to them)
        Counters sum{};
        for (SrvLane& L : s->lanes) {
            Counters c;
            GND_CHECK(hipMemcpyAsync(&c, L.dCtr, sizeof(c), hipMemcpyDeviceToHost, s->ctx->stream));
            GND_CHECK(hipStreamSynchronize(s->ctx->stream));
            sum.nodes += c.nodes;
            sum.packs += c.packs;
            sum.insts += c.insts;
            sum.snodes += c.snodes;
            sum.spacks += c.spacks;
            sum.sinsts += c.sinsts;
            sum.nan += c.nan;
            sum.stackOverflow += c.stackOverflow;
            sum.maxRayNodes = std::max(sum.maxRayNodes, c.maxRayNodes);
            sum.badRays += c.badRays;
            sum.iters += c.iters;
            sum.siters += c.siters;
            sum.steps += c.steps;
            sum.ssteps += c.ssteps;
            sum.rays += c.rays;
            sum.srays += c.srays;
            sum.tailRays += c.tailRays;
            sum.tailSrays += c.tailSrays;
            sum.launches += c.launches;
        }
        std::memset(stats, 0, sizeof(*stats));
        fillStats(stats, sum, samples);
    });
}

}  // extern "C"
