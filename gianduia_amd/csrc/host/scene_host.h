// scene_host.h — owned scene storage of the host shell and its flattening
// into the arrays the device library uploads.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../../include/gnd_hip.h"

struct gnd_prop;
#include "bvh_build.h"
#include "hmath.h"

namespace gnd {


struct MeshData {
    std::vector<float> positions;    // 9 per triangle
    std::vector<float> normals;      // 9 per triangle or empty
    std::vector<float> uvs;          // 6 per triangle or empty
    std::vector<uint32_t> vertexIds; // 3 per triangle or empty
    uint32_t ntris() const { return (uint32_t)(positions.size() / 9); }
};

struct MediumData {
    gnd_medium_desc desc{};
    std::vector<float> density;
};

struct LightData {
    gnd_light_desc desc{};
    std::vector<float> env;  // environment: width*height*3 RGB
};

// Light defaults of the reference constructors (pointlight.cpp:9-12,
// spotlight.cpp:9-22, directional.cpp:8-11, environment.cpp:14-27).
LightData lightDefaults(int type);

class HostScene {
public:
    std::vector<MeshData> meshes;
    std::vector<gnd_primitive_desc> prims;
    std::vector<gnd_material_desc> materials;
    std::vector<MediumData> media;
    std::vector<LightData> lights;
    gnd_camera_desc camera{};
    int integrator = GND_INTEGRATOR_PATH;
    int maxDepth = 1000;
    float aoLength = 5.0f;  // ao integrator "length" (ao.cpp:9)
    uint32_t spp = 1;

    // helpers for generators and the XML loader
    int addMesh(MeshData m);
    int addMaterial(const gnd_material_desc& m);
    int addPrim(int mesh, const Xform& toWorld, int material, const float* radiance = nullptr);
    // the primitive moves to `end` at time 1 (toWorldEnd, primitive.cpp:6-11)
    void setMotion(int prim, const Xform& end);
    int addSphere(float radius, const Xform& toWorld, int material, const float* radiance = nullptr);
    // appends a light after every emitter declared so far (document order)
    void addLight(LightData l);
    int emitterCount() const;
    void setCamera(int type, int w, int h, float fov, const Xform& toWorld);

    void copyFrom(const gnd_scene_desc& d);
    const gnd_scene_desc& desc();  // rebuilt on each call
    const gnd_hip_scene_arrays& build(int threads, double* ms);

private:
    gnd_scene_desc desc_{};
    std::vector<gnd_mesh_desc> meshDescs_;
    std::vector<gnd_medium_desc> mediumDescs_;
    std::vector<gnd_light_desc> lightDescs_;
    bool built_ = false;
    // flattened storage
    std::vector<gnd_bvh4_node> tlas_;
    std::vector<int32_t> tlasPrims_;
    std::vector<gnd_mesh_info> meshInfo_;
    BuildVector<gnd_bvh4_node> blas_;
    BuildVector<gnd_tri_pack4> packs_;
    BuildVector<float> shading_;
    BuildVector<float> cdf_;
    gnd_hip_scene_arrays arrays_{};
};

gnd_material_desc materialMatte(float r, float g, float b);
gnd_material_desc materialConductor(const float* R, const float* eta, const float* k, float roughness);
gnd_material_desc materialGlass(float eta, float roughness, int inside = -1, int outside = -1);
gnd_material_desc materialPlastic(const float* Kd, const float* Ks, float eta, float roughness);
gnd_material_desc materialIndex(int inside, int outside);

// The objects the reference's parser creates (parser.cpp:41-235), handed
// over one by one: class tag, registered type name, name attribute and
// PropertyList, then the parent -> child links in document order; build()
// reads them into a scene exactly as the XML loader reads the same file
// (xml_loader.cpp).
class ObjectGraph {
public:
    explicit ObjectGraph(const std::string& baseDir);
    ~ObjectGraph();
    int add(const std::string& tag, const std::string& type, const std::string& name, const gnd_prop* props,
            int nprops);
    void addChild(int parent, int child);
    std::unique_ptr<HostScene> build(int root) const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

// procedural scenes (generators.cpp)
std::unique_ptr<HostScene> generateScene(const std::string& name, const std::string& params);
// reference XML schema loader (xml_loader.cpp)
std::unique_ptr<HostScene> loadXmlScene(const std::string& path);

}  // namespace gnd
