/*
 * gnd_hip.h — C-ABI of the MI355X ray-tracing library (libgnd_hip.so).
 *
 * The reference renderer (daverlez/gianduia) has no process or device
 * boundary: SamplerIntegrator::render (include/gianduia/core/integrator.h:49-155)
 * calls Scene::rayIntersect (src/scene/scene.cpp:64-71) and the virtual
 * Li() of the path / volpath integrators (src/integrators/path.cpp:31-178,
 * src/integrators/volPath.cpp:37-249) on CPU threads.  This header is the one
 * boundary the GPU build adds: the host shell keeps the scene graph, the XML
 * loader and the BVH4 build, flattens them into the arrays below and drives the
 * device through these entry points.  Plain pointers and sizes only; every
 * function returns 0 on success or a negative GND_E_* status, with text from
 * gnd_hip_last_error().  No C++ exception crosses this boundary.
 *
 * Thread-safety: calls on one context are serialized internally; separate
 * contexts are independent.
 */
#ifndef GND_HIP_H
#define GND_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "gnd_scene.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GND_OK = 0,
    GND_E_INVALID = -1,   /* bad argument / malformed scene (reference: std::runtime_error) */
    GND_E_DEVICE = -2,    /* HIP runtime error                                  */
    GND_E_NOMEM = -3,     /* device allocation failed / over the HBM budget     */
    GND_E_CANCELLED = -4, /* render stopped by gnd_hip_cancel at a pass boundary */
    GND_E_STACK = -5,     /* traversal stack overflow detected on the device   */
    GND_E_TIMEOUT = -6,   /* gnd_hip_server_wait: the job is still queued or running */
    GND_E_BUSY = -7       /* gnd_hip_server_release: the job has not ended yet  */
};

typedef struct gnd_hip_ctx gnd_hip_ctx;
typedef struct gnd_hip_scene gnd_hip_scene;

/* BVHNode4, reference layout (include/gianduia/core/bvhBuilder.h:21-28): 144 B.
 * Offsets are relative to the owning mesh (BLAS) or to tlas_prims (TLAS). */
typedef struct gnd_bvh4_node {
    float min_x[4], min_y[4], min_z[4];
    float max_x[4], max_y[4], max_z[4];
    uint32_t offset[4];
    uint32_t count[4];
    uint8_t child_type[4]; /* 0 empty, 1 inner, 2 leaf */
    uint8_t pad[12];
} gnd_bvh4_node;

/* Mesh::TrianglePack4 (include/gianduia/shapes/mesh.h:37-43): 160 B. */
typedef struct gnd_tri_pack4 {
    float v0x[4], v0y[4], v0z[4];
    float v1x[4], v1y[4], v1z[4];
    float v2x[4], v2y[4], v2z[4];
    uint32_t prim[4]; /* triangle index in the mesh, 0xFFFFFFFF for padding */
} gnd_tri_pack4;

typedef struct gnd_mesh_info {
    uint32_t node_base, n_nodes; /* slice of blas_nodes                        */
    uint32_t pack_base, n_packs; /* slice of packs                             */
    uint32_t tri_base, n_tris;   /* slice of shading (and cdf at tri_base+mesh) */
    uint32_t cdf_base;           /* slice of cdf, n_tris + 1 entries            */
    uint32_t has_uv;
    float total_area;            /* Mesh::m_totalArea (mesh.cpp:121-145)        */
    float bounds_min[3], bounds_max[3];
} gnd_mesh_info;

/* Host-built, flattened scene.  All arrays are borrowed for the duration of
 * gnd_hip_scene_upload and copied to HBM. */
typedef struct gnd_hip_scene_arrays {
    const gnd_scene_desc* desc;      /* primitives, materials, media, camera   */
    uint32_t n_tlas_nodes;
    const gnd_bvh4_node* tlas_nodes; /* BVH::build + buildWide (src/scene/bvh.cpp:10-36) */
    uint32_t n_tlas_prims;
    const int32_t* tlas_prims;       /* TLAS leaf slot -> primitive (scene order)  */
    uint32_t n_meshes;
    const gnd_mesh_info* meshes;
    uint32_t n_blas_nodes;
    const gnd_bvh4_node* blas_nodes; /* all meshes, each relative to its mesh  */
    uint32_t n_packs;
    const gnd_tri_pack4* packs;
    uint32_t n_tris;
    const float* shading;            /* 24 floats per triangle: p0 p1 p2 n0 n1 n2 uv0 uv1 uv2 */
    uint32_t n_cdf;
    const float* cdf;                /* per-mesh triangle-area CDFs            */
} gnd_hip_scene_arrays;

/* One render job: a camera over a resident scene, rendered progressively
 * pass by pass like SamplerIntegrator::render.  film points to device memory
 * of width*height*4 floats (weighted r,g,b sums and weight sum, the Film
 * accumulator of include/gianduia/core/film.h:110-117); the caller resolves
 * rgb = sum / weight.  film_sq (optional) receives per-pixel sums of squared
 * per-pass values for variance estimates. */
typedef struct gnd_hip_job {
    gnd_camera_desc camera;
    uint32_t spp;
    int32_t max_depth;
    int32_t integrator;     /* GND_INTEGRATOR_PATH / _VOLPATH / _DIRECT / _AO  */
    uint64_t seed_offset;   /* added to pixelIdx + pass*W*H before SplitMix64 */
    float* film;            /* device pointer                                 */
    float* film_sq;         /* device pointer or NULL                         */
    float* film_aov;        /* device pointer or NULL: per pixel the filter-weighted
                               sums of the AOVRecord (film.h:13-19) of each sample:
                               albedo rgb, normal xyz (mapped to [0,1]), depth,
                               roughness, metallic (9 floats); divide by the film's
                               weight to resolve (film.cpp:50-95)                 */
    float ao_length;        /* GND_INTEGRATOR_AO: occlusion ray tMax (> 0)     */
    int32_t accumulate;     /* 0: the films are cleared first (Film::clear); 1: the
                               passes are added onto the sums already in them, so a
                               render split into calls (seed_offset advancing by
                               spp*W*H) gives the bits of one call               */
} gnd_hip_job;

typedef struct gnd_hip_stats {
    double ms_total;        /* wall time of the call on the device stream     */
    double ms_trace;        /* sum of closest-hit traversal kernel times      */
    double ms_shadow;       /* sum of shadow traversal kernel times           */
    double ms_shade;        /* sum of shading / camera / film kernel times    */
    uint64_t samples;       /* camera samples (W*H*spp summed over jobs)      */
    uint64_t rays_closest;  /* closest-hit queries                            */
    uint64_t rays_shadow;   /* shadow / transmittance queries                 */
    uint64_t node_visits;   /* BVH4 node pops (filled when counters are on)   */
    uint64_t pack_tests;    /* TrianglePack4 tests (counters on)              */
    uint64_t inst_tests;    /* instance (primitive) tests (counters on)       */
    uint64_t trace_launches;
    uint64_t nan_samples;   /* samples zeroed by the NaN/Inf guard            */
    uint64_t shadow_node_visits;  /* the same counters for shadow rays      */
    uint64_t shadow_pack_tests;
    uint64_t shadow_inst_tests;
    uint64_t max_ray_nodes; /* node pops of the worst closest-hit ray (counters on) */
    uint64_t bad_rays;      /* rays with a non-finite direction (counters on)       */
    uint64_t lane_iterations;        /* traversal loop iterations x 64 lanes, closest hit (counters on) */
    uint64_t shadow_lane_iterations; /* the same for shadow rays (counters on)                        */
    double ms_tail;         /* time in the tail kernel (the last bounces of each wave) */
    uint64_t rays_tail;     /* closest-hit queries made by the tail kernel (in rays_closest) */
    uint64_t lane_steps;    /* of lane_iterations, the lane slots that ran a traversal step
                               (a node visit, a pack test, an instance entry with its root
                               visit, or an instance exit) (counters on)              */
    uint64_t shadow_lane_steps;  /* the same for shadow rays (counters on)             */
} gnd_hip_stats;

/* Tunables picked by gnd_hip_autotune or set by the caller. */
typedef struct gnd_hip_tuning {
    int32_t block_size;       /* threads per block of the traversal kernels (64..256) */
    int32_t stack_lds;        /* per-lane traversal stack entries held in LDS (8/16/24/32) */
    int64_t queue_capacity;   /* paths in flight per wave (default 256M; a render
                                 allocates min(this, its samples) path slots)         */
    int32_t use_graph;        /* capture the bounce loop in a hipGraph                */
    int32_t refill_lanes;     /* idle lanes (1..64) before a closest-hit traversal wave
                                 takes new rays; 0 = per scene (40 when the BVH fits in
                                 4 MiB, else 16)                                        */
    int32_t shadow_refill_lanes;  /* the same for any-hit shadow traversal; 0 = per
                                 scene (64, else 16)                                    */
    int32_t stack_limit;      /* traversal stack entries per lane: 0 = all 128 (the
                                 reference's int stack[64] per level, TLAS + BLAS);
                                 1..128 caps it, so the GND_E_STACK path can be
                                 exercised on small scenes                              */
    int32_t tail_paths;       /* once a wave's in-queue holds at most this many paths,
                                 one persistent tail kernel runs each of them to its
                                 end (a launch per bounce would last as long as its
                                 slowest ray); 0 = eight per resident lane of that
                                 kernel (2M on 256 CUs), -1 = never                    */
} gnd_hip_tuning;

/* ---- context ------------------------------------------------------------ */
int gnd_hip_init(int device, gnd_hip_ctx** out);
void gnd_hip_shutdown(gnd_hip_ctx* ctx);
const char* gnd_hip_last_error(const gnd_hip_ctx* ctx);
int gnd_hip_device_info(gnd_hip_ctx* ctx, char* name, int name_len, int* num_cus, size_t* hbm_bytes);
int gnd_hip_set_tuning(gnd_hip_ctx* ctx, const gnd_hip_tuning* t);
int gnd_hip_get_tuning(gnd_hip_ctx* ctx, gnd_hip_tuning* t);
/* Traversal counters (node pops, pack tests, instance tests) on/off. */
int gnd_hip_set_counters(gnd_hip_ctx* ctx, int enabled);

/* ---- device memory helpers (for bindings without their own allocator) ---- */
int gnd_hip_malloc(gnd_hip_ctx* ctx, size_t bytes, void** dptr);
int gnd_hip_free(gnd_hip_ctx* ctx, void* dptr);
int gnd_hip_memcpy_h2d(gnd_hip_ctx* ctx, void* dst, const void* src, size_t bytes);
int gnd_hip_memcpy_d2h(gnd_hip_ctx* ctx, void* dst, const void* src, size_t bytes);
int gnd_hip_memset(gnd_hip_ctx* ctx, void* dst, int value, size_t bytes);
int gnd_hip_synchronize(gnd_hip_ctx* ctx);

/* ---- scene ---------------------------------------------------------------- */
/* Scene::activate's device half (src/scene/scene.cpp:46-62): uploads the
 * host-built TLAS/BLAS/packs and material/light/medium tables to HBM. */
int gnd_hip_scene_upload(gnd_hip_ctx* ctx, const gnd_hip_scene_arrays* arrays, gnd_hip_scene** out);
void gnd_hip_scene_free(gnd_hip_ctx* ctx, gnd_hip_scene* scene);
size_t gnd_hip_scene_bytes(const gnd_hip_scene* scene);

/* ---- traversal (verify-mode entry points, bit-exact with the reference) ---- */
/* Scene::rayIntersect(ray, isect) closest hit (src/scene/scene.cpp:64-66 ->
 * src/scene/bvh.cpp:38-66).  rays/hits hold DEVICE pointers; n rays.  On a
 * miss hit.t = +inf and hit.inst = hit.prim = -1. */
int gnd_hip_trace_closest(gnd_hip_ctx* ctx, gnd_hip_scene* scene, const gnd_ray_soa* rays, int64_t n,
                          const gnd_hit_soa* hits, gnd_hip_stats* stats);
/* Scene::rayIntersect(ray) any hit (src/scene/scene.cpp:68-71):
 * occluded[i] = 1 if any surface lies in (tmin, tmax), else 0. */
int gnd_hip_trace_any(gnd_hip_ctx* ctx, gnd_hip_scene* scene, const gnd_ray_soa* rays, int64_t n,
                      int32_t* occluded, gnd_hip_stats* stats);

/* ---- rendering ------------------------------------------------------------ */
/* SamplerIntegrator::render for one or more jobs over one resident scene.
 * Jobs are rendered in one wavefront stream: paths of every job share the
 * same bounce launches.  Films are cleared first (Film::clear). */
int gnd_hip_render(gnd_hip_ctx* ctx, gnd_hip_scene* scene, const gnd_hip_job* jobs, int njobs,
                   gnd_hip_stats* stats);
/* Integrator::cancel (include/gianduia/core/integrator.h:22): honoured at the
 * next wave boundary of a running gnd_hip_render on this context. */
int gnd_hip_cancel(gnd_hip_ctx* ctx);

/* Bandwidth probe: `repeats` launches of a plain streaming copy of `bytes`
 * between two device buffers; ms_per_copy receives the average kernel time.
 * Used to calibrate HBM counters (a known-size copy reads and writes bytes). */
int gnd_hip_copy_probe(gnd_hip_ctx* ctx, size_t bytes, int repeats, double* ms_per_copy);

/* Denoises a resolved film image in place of Denoiser::execute
 * (src/core/denoiser.cpp:6-55: OIDN filter "RT", hdr, albedo and normal
 * guides).  OIDN's trained network is not available here, so the filter is
 * an edge-avoiding a-trous wavelet filter (five passes) on albedo-demodulated
 * radiance guided by the albedo and normal images; see denoise.hip.  All
 * pointers are device memory of width*height*3 floats: color (radiance),
 * albedo (or NULL), normal (the AOV's n*0.5+0.5 with 0 = no surface, or
 * NULL), output (may be color).  ms (optional): device time of the call. */
int gnd_hip_denoise(gnd_hip_ctx* ctx, const float* color, const float* albedo, const float* normal, int width,
                    int height, float* output, double* ms);

/* gnd_hip_denoise without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * kernels and results as gnd_hip_denoise; its scratch is a static device array
 * of 2,097,152 pixels (128 MiB per device, held until the process ends), so
 * width*height above that is refused with GND_E_INVALID.  Calls that overlap
 * on one device share that array and race. */
int gnd_hip_denoise_noalloc(gnd_hip_ctx* ctx, const float* color, const float* albedo, const float* normal,
                            int width, int height, float* output, double* ms);

/* Picks the block size (64/128/256) and LDS stack depth (8/16/24/32), then
 * the closest-hit and the any-hit refill thresholds (24..64), then the
 * tail-kernel threshold (off, 128K..8M paths), then the wave capacity (all of
 * the probe's passes in one wave, or a half, a quarter or an eighth of them
 * per wave), by timing `job` cut to at most 40M samples on
 * `scene` (best of two) per candidate, keeping the fastest of each in turn;
 * sets the result as the context's tuning and returns it.  The result is
 * remembered per scene and job shape, so a second call returns at once.
 * job->film is overwritten.  No counterpart in the reference (its CPU render
 * has no launch shape). */
int gnd_hip_autotune(gnd_hip_ctx* ctx, gnd_hip_scene* scene, const gnd_hip_job* job, gnd_hip_tuning* out);

/* ---- serving ---------------------------------------------------------------
 * No counterpart in the reference (it renders one scene per process,
 * include/gianduia/core/integrator.h:49-155).  A server owns a pool of path
 * slots and renders jobs by continuous batching.  A request goes into an
 * inbox and never waits for device work.  The server runs up to `lanes`
 * bounce loops at once, each on its own stream and each rendering the jobs of
 * one resident scene, so jobs on different scenes run concurrently; between
 * two bounces of a lane, passes of waiting jobs of its scene join it (path
 * slots go to jobs in submission order).  Each job's finished passes are
 * gathered into its film in pass order, so a film is bit-identical to
 * gnd_hip_render of the same job alone.  Scenes stay resident in HBM until
 * the budget needs their space (least recently used scene without a lane
 * first).  A worker thread drives the engine; every call is thread-safe. */
typedef struct gnd_hip_server gnd_hip_server;

typedef struct gnd_hip_server_config {
    size_t hbm_budget;     /* bytes for path slots + lane queues + resident scenes;
                              0 = 90% of HBM                                      */
    uint32_t path_slots;   /* paths in flight; a pass needs W*H of them. Each slot
                              costs 140 B of path state plus 160 B of queues per lane */
    uint32_t max_jobs;     /* jobs admitted at once                               */
    uint32_t lanes;        /* concurrent bounce loops (scenes rendered at once),
                              1..16; 0 = 4                                        */
} gnd_hip_server_config;

typedef struct gnd_hip_job_status {
    int32_t state;         /* 0 queued, 1 running, 2 done, < 0 failed (GND_E_*)   */
    uint32_t passes_done;  /* passes gathered into the film                      */
    double queue_ms;       /* submission -> first admission                      */
    double run_ms;         /* first admission -> film complete                   */
    double latency_ms;     /* submission -> film complete (or failure)           */
    double t_submit_ms;    /* arrival (entry of gnd_hip_server_submit), ms since the
                              server was created                                  */
    double t_admit_ms;     /* first admission, same clock; -1 while queued       */
    double t_done_ms;      /* film complete or failure, same clock; -1 before    */
} gnd_hip_job_status;

int gnd_hip_server_create(gnd_hip_ctx* ctx, const gnd_hip_server_config* cfg, gnd_hip_server** out);
/* Stops the worker (jobs in flight are abandoned) and frees the server. */
void gnd_hip_server_destroy(gnd_hip_server* srv);
/* Adds a host-built scene to the pool; `arrays` stays borrowed until the
 * server is destroyed (scenes are uploaded on first use and re-uploaded after
 * eviction). */
int gnd_hip_server_add_scene(gnd_hip_server* srv, const gnd_hip_scene_arrays* arrays, int32_t* scene_id);
/* Queues a job (job->film: device memory the caller keeps until it is done). */
int gnd_hip_server_submit(gnd_hip_server* srv, int32_t scene_id, const gnd_hip_job* job, uint64_t* ticket);
int gnd_hip_server_poll(gnd_hip_server* srv, uint64_t ticket, gnd_hip_job_status* status);
/* Waits until the job is done or failed (GND_OK, status filled), or until
 * timeout_s passes (GND_E_TIMEOUT, status filled: the job still owns its film). */
int gnd_hip_server_wait(gnd_hip_server* srv, uint64_t ticket, double timeout_s, gnd_hip_job_status* status);
/* Asks the server to stop a queued or running job (Integrator::cancel for one
 * request): it admits no more of its passes, drops those in flight and then
 * fails it with GND_E_CANCELLED; once a wait reports that, no kernel touches
 * its film any more.  No effect on a job that has ended. */
int gnd_hip_server_cancel(gnd_hip_server* srv, uint64_t ticket);
/* Forgets an ended job (its status can no longer be read); GND_E_BUSY while
 * it is queued or running.  A long-running server releases what it has read. */
int gnd_hip_server_release(gnd_hip_server* srv, uint64_t ticket);
/* Totals since creation (samples of finished jobs, rays, counters). */
int gnd_hip_server_stats(gnd_hip_server* srv, gnd_hip_stats* stats, int32_t* resident_scenes,
                         size_t* resident_bytes);

#ifdef __cplusplus
}
#endif

#endif
