/*
 * gnd_scene.h — flat, pointer-and-size description of a Gianduia scene.
 *
 * This is the data format that crosses every C boundary of the build: the
 * host shell (XML loader, procedural generators) produces it, the HIP library
 * consumes it in gnd_hip_scene_upload, and the CPU restatement of the
 * reference renderer consumes the same bytes.  It carries exactly what the
 * reference's parsed object graph carries for the hot path:
 *
 *   mesh        — de-indexed corners, as Mesh::loadOBJ leaves them
 *                 (src/shapes/mesh.cpp:197-228): 3 positions per triangle,
 *                 optional per-corner normals and uvs.
 *   primitive   — Primitive = shape + static toWorld transform (matrix and
 *                 the analytic inverse composed at parse time,
 *                 src/core/parser.cpp:112-165, include/gianduia/math/transform.h:37-39)
 *                 + material + optional area emitter (src/shapes/primitive.cpp:10-37).
 *   material    — matte / conductor / glass / plastic / index with constant
 *                 textures (src/materials/{matte,conductor,glass,plastic,index}.cpp).
 *   medium      — homogeneous or dense-grid heterogeneous
 *                 (src/materials/{homogeneous,heterogeneous}.cpp).
 *   camera      — perspective / thinlens + film filter
 *                 (src/cameras/{perspective,thinLens}.cpp, include/gianduia/core/filter.h).
 *   light       — point / spot / directional / environment emitters
 *                 (src/emitters/{pointlight,spotlight,directional,environment}.cpp).
 *
 * Matrices are column-major float[16] (m[c*4 + r]), the memory order of the
 * reference's glm::mat4.
 */
#ifndef GND_SCENE_H
#define GND_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { GND_SHAPE_MESH = 0, GND_SHAPE_SPHERE = 1 };

enum {
    GND_MAT_MATTE = 0,      /* src/materials/matte.cpp:46-53      */
    GND_MAT_CONDUCTOR = 1,  /* src/materials/conductor.cpp:101-120 */
    GND_MAT_GLASS = 2,      /* src/materials/glass.cpp:98-115      */
    GND_MAT_PLASTIC = 3,    /* src/materials/plastic.cpp:74-91     */
    GND_MAT_INDEX = 4       /* src/materials/index.cpp:10-16 (null BSDF) */
};

enum { GND_MEDIUM_HOMOGENEOUS = 0, GND_MEDIUM_GRID = 1 };
enum { GND_CAMERA_PERSPECTIVE = 0, GND_CAMERA_THINLENS = 1 };
enum { GND_FILTER_GAUSSIAN = 0, GND_FILTER_BOX = 1 };
/* path: src/integrators/path.cpp; volpath: volPath.cpp; direct: direct.cpp
 * (emission + one light sample + one BSDF sample, any-hit shadows); ao:
 * ao.cpp (one cosine-weighted occlusion ray of length ao_length)          */
enum { GND_INTEGRATOR_PATH = 0, GND_INTEGRATOR_VOLPATH = 1, GND_INTEGRATOR_DIRECT = 2, GND_INTEGRATOR_AO = 3 };
enum { GND_LIGHT_POINT = 0, GND_LIGHT_SPOT = 1, GND_LIGHT_DIRECTIONAL = 2, GND_LIGHT_ENVIRONMENT = 3 };

typedef struct gnd_mesh_desc {
    uint32_t ntris;
    const float* positions;  /* 9*ntris floats: p0,p1,p2 of each triangle   */
    const float* normals;    /* 9*ntris floats or NULL (then smooth normals  */
                             /* are generated like loadOBJ does)            */
    const float* uvs;        /* 6*ntris floats or NULL                      */
    const uint32_t* vertex_ids; /* 3*ntris shared-vertex ids used only when */
                             /* normals == NULL to build smooth normals;    */
                             /* NULL means every corner is its own vertex   */
} gnd_mesh_desc;

typedef struct gnd_material_desc {
    int32_t type;
    int32_t albedo_is_checker;  /* matte with no albedo: checkerboard default */
    float albedo[3];            /* matte                                       */
    float R[3];                 /* conductor, glass                            */
    float T[3];                 /* glass                                       */
    float eta_rgb[3];           /* conductor eta                               */
    float k_rgb[3];             /* conductor k                                 */
    float Kd[3], Ks[3];         /* plastic                                     */
    float eta;                  /* glass, plastic IOR                          */
    float roughness;
    int32_t medium_inside;      /* -1 = none                                   */
    int32_t medium_outside;     /* -1 = none                                   */
} gnd_material_desc;

typedef struct gnd_medium_desc {
    int32_t type;
    float sigma_a[3];
    float sigma_s[3];
    float g;
    /* dense grid (heterogeneous): values at integer index coordinates,       */
    /* x fastest; the whole [0,n) box is the active region                    */
    int32_t nx, ny, nz;
    const float* density;
    float density_scale;
    int32_t macro_res;          /* majorant grid target resolution (ref: 32)  */
    float to_local[16];         /* world -> grid index space                  */
    float to_local_inv[16];
} gnd_medium_desc;

typedef struct gnd_primitive_desc {
    int32_t shape_type;
    int32_t mesh;               /* index into meshes when shape_type == MESH   */
    float radius;               /* sphere radius                               */
    int32_t material;           /* index into materials                        */
    int32_t has_emitter;        /* area emitter attached                       */
    float radiance[3];
    float to_world[16];
    float to_world_inv[16];
    /* motion blur (Primitive ctor, primitive.cpp:6-11): with has_motion the
     * primitive moves from to_world at time 0 to to_world_end at time 1
     * (AnimatedTransform; rays carry the camera sample's time)              */
    int32_t has_motion;
    float to_world_end[16];
    float to_world_end_inv[16];
} gnd_primitive_desc;

typedef struct gnd_camera_desc {
    int32_t type;
    int32_t width, height;
    float fov;
    float to_world[16];
    float to_world_inv[16];
    float lens_radius, focal_distance;
    float k1, ca_axial, ca_lateral;
    int32_t vignetting;
    int32_t blades;
    int32_t filter;
    float filter_radius[2];
    float filter_alpha;
    /* motion blur (Camera ctor, camera.h:20-25, AnimatedTransform): with
     * has_motion the camera moves from to_world at time 0 to to_world_end at
     * time 1 (decomposed, mixed / slerped per ray); rolling shutter
     * (thin lens only, thinLens.cpp:29-31, 81-87): time = (1 - film y) *
     * scan_duration + sample time * exposure_duration                      */
    int32_t has_motion;
    float to_world_end[16];
    int32_t rolling_shutter;
    float scan_duration, exposure_duration;
} gnd_camera_desc;

/* An emitter that is not a shape (Scene::addChild's EEmitter case,
 * src/scene/scene.cpp:33-40).  The scene's emitter list, from which NEE picks
 * uniformly (scene.h:28-34), holds the area emitters of the primitives and
 * these lights in document order: `emitter_index` is this light's position in
 * it; the area emitters fill the other positions in primitive order.  At most
 * one environment light. */
typedef struct gnd_light_desc {
    int32_t type;            /* GND_LIGHT_*                                            */
    int32_t emitter_index;
    float position[3];       /* point, spot                                            */
    float direction[3];      /* spot: cone axis; directional: direction the light
                                travels ("direction", default (0,-1,0))                 */
    float power[3];          /* point, spot (default 10)                               */
    float irradiance[3];     /* directional (default 1)                                */
    float spot_size;         /* spot: full cone angle in degrees (default 45)          */
    float blend;             /* spot: soft-edge fraction (default 0.15)                */
    float strength;          /* environment (default 1)                                */
    int32_t env_width, env_height;
    const float* env_rgb;    /* environment: width*height RGB floats, row 0 = theta 0
                                (+z), column 0 = phi 0 (the EXR's pixel layout)         */
} gnd_light_desc;

typedef struct gnd_scene_desc {
    uint32_t nmeshes;
    const gnd_mesh_desc* meshes;
    uint32_t nprims;
    const gnd_primitive_desc* prims;
    uint32_t nmaterials;
    const gnd_material_desc* materials;
    uint32_t nmedia;
    const gnd_medium_desc* media;
    gnd_camera_desc camera;
    int32_t integrator;
    int32_t max_depth;
    uint32_t spp;
    uint32_t nlights;
    const gnd_light_desc* lights;
    float ao_length;         /* ao integrator: occlusion ray length ("length",
                                default 5, ao.cpp:9)                           */
} gnd_scene_desc;

/* Struct-of-arrays ray batch (32 B per ray) and hit batch (20 B per ray).   */
typedef struct gnd_ray_soa {
    float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
} gnd_ray_soa;

typedef struct gnd_hit_soa {
    float* t;          /* hit distance, +inf (ray tmax untouched) when missed */
    int32_t* inst;     /* primitive index in scene order, -1 when missed      */
    int32_t* prim;     /* triangle index inside the mesh (0 for spheres)      */
    float *u, *v;      /* barycentrics of the hit                             */
} gnd_hit_soa;

#ifdef __cplusplus
}
#endif

#endif
