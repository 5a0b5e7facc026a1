// gnd_ref_capi.cpp — extern "C" entry points of the oracle for the test-suite
// and the benchmark's CPU-baseline leg (loaded through ctypes).  Test
// infrastructure only (see gnd_ref.h).
#include "gnd_ref.h"

#include <chrono>
#include <stdexcept>
#include <string>

using namespace gref;

namespace {
thread_local std::string g_err;
}

extern "C" {

const char* gnd_ref_last_error() { return g_err.c_str(); }

void* gnd_ref_scene_create(const gnd_scene_desc* d) {
    try {
        Scene* s = new Scene();
        sceneBuild(*s, *d);
        return s;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void gnd_ref_scene_free(void* h) { delete static_cast<Scene*>(h); }

// counts[0]=tlas nodes, [1]=sum of mesh BVH4 nodes, [2]=sum of packs, [3]=emitters
void gnd_ref_scene_counts(void* h, int64_t* counts) {
    Scene* s = static_cast<Scene*>(h);
    int64_t nb = 0, np = 0;
    for (auto& m : s->meshes) {
        nb += (int64_t)m.nodes4.size();
        np += (int64_t)m.packs.size();
    }
    counts[0] = (int64_t)s->tlas.size();
    counts[1] = nb;
    counts[2] = np;
    counts[3] = (int64_t)s->emitters.size();
}

// Raw reference-layout arrays: BVHNode4 is 144 B, TrianglePack4 is 160 B.
int64_t gnd_ref_mesh_nodes(void* h, int mesh, void* out) {
    Scene* s = static_cast<Scene*>(h);
    const auto& v = s->meshes[mesh].nodes4;
    if (out) std::memcpy(out, v.data(), v.size() * sizeof(BVHNode4));
    return (int64_t)v.size();
}
int64_t gnd_ref_mesh_packs(void* h, int mesh, void* out) {
    Scene* s = static_cast<Scene*>(h);
    const auto& v = s->meshes[mesh].packs;
    if (out) std::memcpy(out, v.data(), v.size() * sizeof(TrianglePack4));
    return (int64_t)v.size();
}
int64_t gnd_ref_mesh_cdf(void* h, int mesh, float* out, float* total_area) {
    Scene* s = static_cast<Scene*>(h);
    const auto& v = s->meshes[mesh].cdf;
    if (out) std::memcpy(out, v.data(), v.size() * sizeof(float));
    if (total_area) *total_area = s->meshes[mesh].totalArea;
    return (int64_t)v.size();
}
int64_t gnd_ref_tlas_nodes(void* h, void* out) {
    Scene* s = static_cast<Scene*>(h);
    if (out) std::memcpy(out, s->tlas.data(), s->tlas.size() * sizeof(BVHNode4));
    return (int64_t)s->tlas.size();
}
int64_t gnd_ref_tlas_prims(void* h, int32_t* out) {
    Scene* s = static_cast<Scene*>(h);
    if (out)
        for (size_t i = 0; i < s->tlasPrims.size(); ++i) out[i] = s->tlasPrims[i];
    return (int64_t)s->tlasPrims.size();
}
int gnd_ref_sizeof_node4() { return (int)sizeof(BVHNode4); }
int gnd_ref_sizeof_pack() { return (int)sizeof(TrianglePack4); }

// Closest-hit (any_hit = 0) or any-hit (any_hit = 1) queries over a ray batch.
// Per-ray visit counters (nodes, packs, instance tests) go to counts[3*i..]
// when counts is non-NULL.
int gnd_ref_trace(void* h, const gnd_ray_soa* rays, int64_t n, const gnd_hit_soa* hits, int any_hit,
                  int64_t* counts) {
    Scene* s = static_cast<Scene*>(h);
    for (int64_t i = 0; i < n; ++i) {
        Ray r;
        r.o = V3(rays->ox[i], rays->oy[i], rays->oz[i]);
        r.d = V3(rays->dx[i], rays->dy[i], rays->dz[i]);
        r.tMin = rays->tmin[i];
        r.tMax = rays->tmax[i];
        Counters c;
        if (any_hit) {
            bool occ = intersectAny(*s, r, &c);
            hits->inst[i] = occ ? 1 : 0;
        } else {
            Hit hh;
            bool ok = intersectClosest(*s, r, hh, &c);
            if (ok) {
                hits->t[i] = hh.t;
                hits->inst[i] = hh.prim;
                hits->prim[i] = (int32_t)hh.primIndex;
                hits->u[i] = hh.u;
                hits->v[i] = hh.v;
            } else {
                hits->t[i] = kInf;
                hits->inst[i] = -1;
                hits->prim[i] = -1;
                hits->u[i] = 0.0f;
                hits->v[i] = 0.0f;
            }
        }
        if (counts) {
            counts[3 * i + 0] = (int64_t)c.nodes;
            counts[3 * i + 1] = (int64_t)c.packs;
            counts[3 * i + 2] = (int64_t)c.insts;
        }
    }
    return 0;
}

// Progressive render of passes [spp_begin, spp_end) into rgbw (4 floats per
// pixel: weighted r,g,b sums and weight sum), optional sq (per-pixel sums of
// squared per-pass values) and optional aov (9 floats per pixel: weighted
// albedo rgb, normal xyz, depth, roughness, metallic sums).  stats:
// [0]=seconds, [1]=ray queries, [2]=node visits, [3]=pack tests,
// [4]=instance tests.
int gnd_ref_render(void* h, int threads, int spp_begin, int spp_end, int row_begin, int row_end,
                   uint64_t seed_offset, float* rgbw, float* sq, double* stats, float* aov) {
    try {
        Scene* s = static_cast<Scene*>(h);
        RenderOptions o;
        o.threads = threads;
        o.sppBegin = spp_begin;
        o.sppEnd = spp_end;
        o.rowBegin = row_begin;
        o.rowEnd = row_end;
        o.seedOffset = seed_offset;
        o.accumulateSquares = sq != nullptr;
        o.aovs = aov != nullptr;
        Film f;
        Counters c;
        auto t0 = std::chrono::steady_clock::now();
        render(*s, o, f, stats ? &c : nullptr);
        auto t1 = std::chrono::steady_clock::now();
        std::memcpy(rgbw, f.rgbw.data(), f.rgbw.size() * sizeof(float));
        if (sq) std::memcpy(sq, f.sq.data(), f.sq.size() * sizeof(float));
        if (aov) std::memcpy(aov, f.aov.data(), f.aov.size() * sizeof(float));
        if (stats) {
            stats[0] = std::chrono::duration<double>(t1 - t0).count();
            stats[1] = (double)c.rays;
            stats[2] = (double)c.nodes;
            stats[3] = (double)c.packs;
            stats[4] = (double)c.insts;
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// PCG32 / SplitMix64 probes for known-answer checks
void gnd_ref_pcg32(uint64_t initstate, uint64_t initseq, int n, uint32_t* out) {
    PCG32 r;
    r.seed(initstate, initseq);
    for (int i = 0; i < n; ++i) out[i] = r.nextUInt();
}
uint64_t gnd_ref_splitmix64(uint64_t idx) { return splitMix64(idx); }
void gnd_ref_sampler_floats(uint64_t idx, int n, float* out) {
    Sampler s;
    s.seed(idx);
    for (int i = 0; i < n; ++i) out[i] = s.next1D();
}

}  // extern "C"
