/*
 * gnd_host.h — C API of the host scene shell in libgnd_hip.so.
 *
 * The host shell keeps what the reference keeps on the CPU: the scene graph
 * parsed from the reference's XML schema (src/core/parser.cpp:41-235), the
 * meshes with their BLAS build (src/shapes/mesh.cpp:17-146) and the TLAS build
 * (src/scene/bvh.cpp:10-36).  It flattens a scene into gnd_hip_scene_arrays
 * for gnd_hip_scene_upload, and offers the procedural scenes of the benchmark
 * configurations.  A render through gnd_host_render behaves like
 * SamplerIntegrator::render (include/gianduia/core/integrator.h:49-155) with
 * the film resolved on the host.
 */
#ifndef GND_HOST_H
#define GND_HOST_H

#include "gnd_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gnd_host_scene gnd_host_scene;

const char* gnd_host_last_error(void);

/* Copies a caller-built description (all arrays are copied). */
int gnd_host_scene_create(const gnd_scene_desc* desc, gnd_host_scene** out);
/* Reference XML schema: scene/camera/sampler/integrator/primitive/shape(mesh
 * with an OBJ file, sphere)/material/emitter(area)/medium/filter. */
int gnd_host_scene_load_xml(const char* path, gnd_host_scene** out);
/* Procedural scenes: "cbox" (512x512 Cornell box), "scaled" (1M-triangle
 * displaced icosphere + 64 instances, 1920x1080 thin lens), "volumetric"
 * (Cornell box + glass sphere with a homogeneous medium + 256^3 density grid),
 * "procedural" (the scaled generator at a target triangle count).  params is
 * "key=value;key=value" (seed, tris, width, height, spp, depth, grid). */
int gnd_host_scene_generate(const char* name, const char* params, gnd_host_scene** out);
void gnd_host_scene_free(gnd_host_scene* s);

/* ---- the reference's object graph, object by object ----
 * A scene made from the objects the reference's parser creates instead of
 * from a file: each object with its class tag, registered type name, name
 * attribute and PropertyList (GndFactory::createInstance(type, props),
 * src/core/parser.cpp:171-193), then each parent -> child link in document
 * order (GndObject::addChild, parser.cpp:200-211).  gnd_host_builder_finish
 * reads the graph exactly as gnd_host_scene_load_xml reads a file holding
 * the same objects, so a binding can hand over the Scene* the reference has
 * built.  Relative file names (OBJ meshes, EXR maps) resolve against
 * base_dir (FileResolver's base path). */
enum {
    GND_PROP_BOOLEAN = 0, /* PropertyList::setBoolean: ivalue 0 / 1           */
    GND_PROP_INTEGER = 1, /* setInteger: ivalue                               */
    GND_PROP_FLOAT = 2,   /* setFloat: v[0]                                   */
    GND_PROP_STRING = 3,  /* setString: svalue                                */
    GND_PROP_POINT = 4,   /* setPoint3: v[0..2]                               */
    GND_PROP_VECTOR2 = 5, /* setVector2 / setPoint2: v[0..1]                  */
    GND_PROP_VECTOR = 6,  /* setVector3: v[0..2]                              */
    GND_PROP_COLOR = 7,   /* setColor: v[0..2]                                */
    GND_PROP_TRANSFORM = 8 /* setTransform: matrix = 32 floats, the matrix and
                              then its inverse (Transform::getMatrix /
                              getInverse), each column-major                 */
};
typedef struct gnd_prop {
    const char* name;
    int32_t kind;        /* GND_PROP_* */
    int32_t ivalue;
    float v[3];
    const char* svalue;
    const float* matrix;
} gnd_prop;
typedef struct gnd_host_builder gnd_host_builder;
int gnd_host_builder_create(const char* base_dir, gnd_host_builder** out);
/* tag: "scene", "primitive", "shape", "camera", "integrator", "sampler",
 * "texture", "material", "emitter", "filter" or "medium"
 * (GndObject::classTypeToString, include/gianduia/core/object.h); type: the
 * name the object was created by ("mesh", "matte", "thinlens", ...; the tag
 * for scene and primitive); name: the object's name or NULL.  The
 * properties are copied.  *id receives the object's id. */
int gnd_host_builder_object(gnd_host_builder* b, const char* tag, const char* type, const char* name,
                            const gnd_prop* props, int nprops, int32_t* id);
/* parent->addChild(child); an object added to two parents is shared (a
 * <ref id>). */
int gnd_host_builder_add_child(gnd_host_builder* b, int32_t parent, int32_t child);
/* The scene rooted at `root` (a "scene" object), with the XML loader's
 * checks and errors. */
int gnd_host_builder_finish(gnd_host_builder* b, int32_t root, gnd_host_scene** out);
void gnd_host_builder_free(gnd_host_builder* b);

/* Description of the scene (valid until the scene is freed). */
const gnd_scene_desc* gnd_host_scene_desc(gnd_host_scene* s);
/* Builds (once) and returns the flattened arrays (valid until freed).
 * threads <= 0 uses every hardware thread.  build_ms receives the time. */
int gnd_host_scene_build(gnd_host_scene* s, int threads, const gnd_hip_scene_arrays** out, double* build_ms);

/* Camera overrides used by batch jobs. */
int gnd_host_camera_lookat(gnd_camera_desc* cam, const float* origin, const float* target, const float* up);

/* Integrator::render for the scene's own camera/sampler/integrator on a
 * context: uploads (once per context), renders every pass and writes the
 * resolved RGB image (width*height*3 floats, Film::resolve) to rgb.  spp <= 0
 * keeps the scene's sample count. */
int gnd_host_render(gnd_hip_ctx* ctx, gnd_host_scene* s, int spp, uint64_t seed_offset, float* rgb,
                    gnd_hip_stats* stats);

/* ---- Integrator / Sampler plugin interface ----------------------------------
 * Integrator (include/gianduia/core/integrator.h:18-40): render(scene),
 * cancel(), setCallback(sampleIndex, film).  Integrators are created by their
 * XML name through the factory (include/gianduia/core/factory.h:26-36):
 * "path" and "volpath" (the GPU integrators under the reference's names),
 * "path_hip", "volpath_hip".  render() renders the scene's camera at its
 * sampler's spp into the integrator's film, in batches of passes_per_callback
 * passes; the callback receives the index of the last pass done and the
 * resolved film (width*height*3 floats, Film::resolve).  cancel() stops at the
 * next batch (from any thread, also from inside the callback). */
typedef struct gnd_host_integrator gnd_host_integrator;
typedef void (*gnd_host_progress_fn)(int sample_index, const float* rgb, int width, int height, void* user);

/* max_depth < 0 keeps the integrator default (1000, path.cpp:9). */
int gnd_host_integrator_create(gnd_hip_ctx* ctx, const char* name, int max_depth, int passes_per_callback,
                               gnd_host_integrator** out);
/* The scene's own <integrator type=... maxDepth=...>. */
int gnd_host_integrator_from_scene(gnd_hip_ctx* ctx, gnd_host_scene* s, int passes_per_callback,
                                   gnd_host_integrator** out);
int gnd_host_integrator_set_callback(gnd_host_integrator* it, gnd_host_progress_fn fn, void* user);
int gnd_host_integrator_render(gnd_host_integrator* it, gnd_host_scene* s);
int gnd_host_integrator_cancel(gnd_host_integrator* it);
/* Resolved film (rgb may be NULL to query the size) and the passes rendered. */
int gnd_host_integrator_film(gnd_host_integrator* it, float* rgb, int* width, int* height, int* passes);
/* Film AOVs (include/gianduia/core/film.h:13-107): set_aovs before render()
 * makes the film also hold the resolved AOV images; aovs() copies them as
 * width*height*9 floats per pixel: albedo rgb, normal xyz (mapped to [0,1]),
 * depth, roughness, metallic (where no sample landed: 0, 0, -1, 0, 0).
 * GND_E_INVALID if the last render had no AOVs. */
int gnd_host_integrator_set_aovs(gnd_host_integrator* it, int enabled);
int gnd_host_integrator_aovs(gnd_host_integrator* it, float* out, int* width, int* height);
/* Film::savePNG / Film::saveEXR (integrator.h:152-153, film.cpp:153-194):
 * ".png" (8-bit sRGB), ".exr" (OpenEXR scanline, float R G B, uncompressed)
 * or ".pfm" (32-bit float RGB). */
int gnd_host_integrator_save(gnd_host_integrator* it, const char* path);
/* The end of SamplerIntegrator::render (integrator.h:152-153): with a path
 * stem set (FileResolver's output path / output name, fileResolver.cpp:44-59),
 * every render, also a cancelled one, ends by writing <stem>.exr and
 * <stem>.png of the film; NULL or "" turns it off (the default). */
int gnd_host_integrator_set_output(gnd_host_integrator* it, const char* path_stem);
/* One resolved AOV image (Image2D::saveEXR of the film's AOV images,
 * image2d.h:62-97): aov = "albedo" | "normal" (R G B) or "depth" |
 * "roughness" | "metallic" (Y), path ".exr" or ".pfm". */
int gnd_host_integrator_save_aov(gnd_host_integrator* it, const char* aov, const char* path);
/* Denoiser::execute on the integrator's film (src/core/denoiser.cpp:6-55,
 * src/app/application.cpp:299-300): the resolved radiance is replaced by its
 * denoised version (gnd_hip_denoise), guided by the film's albedo and normal
 * images when AOVs are on; save() afterwards writes the denoised image. */
int gnd_host_integrator_denoise(gnd_host_integrator* it);
void gnd_host_integrator_free(gnd_host_integrator* it);
/* Registered integrator names, comma separated; returns the full length. */
int gnd_host_integrator_names(char* buf, int len);
/* Writes a width*height*3 float RGB image as ".png", ".exr" or ".pfm" (the film writers). */
int gnd_host_image_save(const char* path, const float* rgb, int width, int height);
/* Reads a scanline OpenEXR file the way Film(const std::string&) does
 * (src/core/film.cpp:32-48 via Image2D::loadEXR, include/gianduia/core/
 * image2d.h:100-134): R, G, B as floats over the data window, a missing
 * channel read as 0.  NONE / RLE / ZIPS / ZIP compression, UINT / HALF /
 * FLOAT channels.  Writes the size; with rgb non-NULL also the pixels
 * (width*height*3 floats, capacity = floats available). */
int gnd_host_image_load(const char* path, int* width, int* height, float* rgb, size_t capacity);
/* IndependentSampler (src/samplers/independent.cpp:7-47): the first n floats
 * of sample index `sample_index`. */
int gnd_host_sampler_floats(uint64_t sample_index, int n, float* out);

#ifdef __cplusplus
}
#endif

#endif
