// scene_host.cpp — owned scene storage and flattening (Scene::activate,
// Mesh::activate and BVH::build of the reference, host half).
#include "scene_host.h"

#include <atomic>
#include <chrono>
#include <stdexcept>
#include <thread>

namespace gnd {

gnd_material_desc materialMatte(float r, float g, float b) {
    gnd_material_desc m{};
    m.type = GND_MAT_MATTE;
    m.albedo[0] = r; m.albedo[1] = g; m.albedo[2] = b;
    m.medium_inside = m.medium_outside = -1;
    return m;
}

// Conductor defaults R = eta = k = 1, roughness 0 (conductor.cpp:66-99)
gnd_material_desc materialConductor(const float* R, const float* eta, const float* k, float roughness) {
    gnd_material_desc m{};
    m.type = GND_MAT_CONDUCTOR;
    for (int c = 0; c < 3; ++c) {
        m.R[c] = R ? R[c] : 1.0f;
        m.eta_rgb[c] = eta ? eta[c] : 1.0f;
        m.k_rgb[c] = k ? k[c] : 1.0f;
    }
    m.roughness = roughness;
    m.medium_inside = m.medium_outside = -1;
    return m;
}

// Glass defaults R = T = 1, eta 1.5, roughness 0 (glass.cpp:63-96)
gnd_material_desc materialGlass(float eta, float roughness, int inside, int outside) {
    gnd_material_desc m{};
    m.type = GND_MAT_GLASS;
    for (int c = 0; c < 3; ++c) m.R[c] = m.T[c] = 1.0f;
    m.eta = eta;
    m.roughness = roughness;
    m.medium_inside = inside;
    m.medium_outside = outside;
    return m;
}

// Plastic defaults Kd = Ks = 0.25, eta 1.5, roughness 0.1 (plastic.cpp:51-72)
gnd_material_desc materialPlastic(const float* Kd, const float* Ks, float eta, float roughness) {
    gnd_material_desc m{};
    m.type = GND_MAT_PLASTIC;
    for (int c = 0; c < 3; ++c) {
        m.Kd[c] = Kd ? Kd[c] : 0.25f;
        m.Ks[c] = Ks ? Ks[c] : 0.25f;
    }
    m.eta = eta;
    m.roughness = roughness;
    m.medium_inside = m.medium_outside = -1;
    return m;
}

gnd_material_desc materialIndex(int inside, int outside) {
    gnd_material_desc m{};
    m.type = GND_MAT_INDEX;
    m.medium_inside = inside;
    m.medium_outside = outside;
    return m;
}

int HostScene::addMesh(MeshData m) {
    meshes.push_back(std::move(m));
    built_ = false;
    return (int)meshes.size() - 1;
}

int HostScene::addMaterial(const gnd_material_desc& m) {
    materials.push_back(m);
    return (int)materials.size() - 1;
}

static void setXform(gnd_primitive_desc& p, const Xform& x) {
    std::memcpy(p.to_world, x.m.m, sizeof(p.to_world));
    std::memcpy(p.to_world_inv, x.inv.m, sizeof(p.to_world_inv));
}

int HostScene::addPrim(int mesh, const Xform& toWorld, int material, const float* radiance) {
    gnd_primitive_desc p{};
    p.shape_type = GND_SHAPE_MESH;
    p.mesh = mesh;
    p.radius = 1.0f;
    p.material = material;
    if (radiance) {
        p.has_emitter = 1;
        p.radiance[0] = radiance[0]; p.radiance[1] = radiance[1]; p.radiance[2] = radiance[2];
    }
    setXform(p, toWorld);
    prims.push_back(p);
    built_ = false;
    return (int)prims.size() - 1;
}

int HostScene::addSphere(float radius, const Xform& toWorld, int material, const float* radiance) {
    int i = addPrim(-1, toWorld, material, radiance);
    prims[i].shape_type = GND_SHAPE_SPHERE;
    prims[i].radius = radius;
    return i;
}

void HostScene::setCamera(int type, int w, int h, float fov, const Xform& toWorld) {
    camera = gnd_camera_desc{};
    camera.type = type;
    camera.width = w;
    camera.height = h;
    camera.fov = fov;
    std::memcpy(camera.to_world, toWorld.m.m, sizeof(camera.to_world));
    std::memcpy(camera.to_world_inv, toWorld.inv.m, sizeof(camera.to_world_inv));
    camera.focal_distance = 10.0f;  // thinLens.cpp:18
    camera.filter = GND_FILTER_GAUSSIAN;  // film.cpp:29 default filter
    camera.filter_radius[0] = camera.filter_radius[1] = 2.0f;
    camera.filter_alpha = 2.0f;
}

void HostScene::copyFrom(const gnd_scene_desc& d) {
    meshes.clear();
    for (uint32_t i = 0; i < d.nmeshes; ++i) {
        const gnd_mesh_desc& md = d.meshes[i];
        MeshData m;
        m.positions.assign(md.positions, md.positions + 9 * (size_t)md.ntris);
        if (md.normals) m.normals.assign(md.normals, md.normals + 9 * (size_t)md.ntris);
        if (md.uvs) m.uvs.assign(md.uvs, md.uvs + 6 * (size_t)md.ntris);
        if (md.vertex_ids) m.vertexIds.assign(md.vertex_ids, md.vertex_ids + 3 * (size_t)md.ntris);
        meshes.push_back(std::move(m));
    }
    prims.assign(d.prims, d.prims + d.nprims);
    materials.assign(d.materials, d.materials + d.nmaterials);
    media.clear();
    for (uint32_t i = 0; i < d.nmedia; ++i) {
        MediumData md;
        md.desc = d.media[i];
        if (md.desc.type == GND_MEDIUM_GRID && md.desc.density)
            md.density.assign(md.desc.density, md.desc.density + (size_t)md.desc.nx * md.desc.ny * md.desc.nz);
        media.push_back(std::move(md));
    }
    camera = d.camera;
    integrator = d.integrator;
    maxDepth = d.max_depth;
    spp = d.spp;
    built_ = false;
}

const gnd_scene_desc& HostScene::desc() {
    meshDescs_.resize(meshes.size());
    for (size_t i = 0; i < meshes.size(); ++i) {
        gnd_mesh_desc& md = meshDescs_[i];
        md.ntris = meshes[i].ntris();
        md.positions = meshes[i].positions.data();
        md.normals = meshes[i].normals.empty() ? nullptr : meshes[i].normals.data();
        md.uvs = meshes[i].uvs.empty() ? nullptr : meshes[i].uvs.data();
        md.vertex_ids = meshes[i].vertexIds.empty() ? nullptr : meshes[i].vertexIds.data();
    }
    mediumDescs_.resize(media.size());
    for (size_t i = 0; i < media.size(); ++i) {
        mediumDescs_[i] = media[i].desc;
        mediumDescs_[i].density = media[i].density.empty() ? nullptr : media[i].density.data();
    }
    desc_.nmeshes = (uint32_t)meshes.size();
    desc_.meshes = meshDescs_.data();
    desc_.nprims = (uint32_t)prims.size();
    desc_.prims = prims.data();
    desc_.nmaterials = (uint32_t)materials.size();
    desc_.materials = materials.data();
    desc_.nmedia = (uint32_t)media.size();
    desc_.media = mediumDescs_.data();
    desc_.camera = camera;
    desc_.integrator = integrator;
    desc_.max_depth = maxDepth;
    desc_.spp = spp;
    return desc_;
}

// Corner normals as Mesh::loadOBJ leaves them (mesh.cpp:163-228): given
// normals are normalized, missing ones are generated per shared vertex.
static void cornerNormals(const MeshData& m, float* out /* 9 per tri, stride 24 */) {
    const uint32_t nt = m.ntris();
    if (!m.normals.empty()) {
        for (uint32_t c = 0; c < 3 * nt; ++c) {
            Vec3 n = normalize(Vec3(m.normals[3 * c], m.normals[3 * c + 1], m.normals[3 * c + 2]));
            float* o = out + 24 * (c / 3) + 9 + 3 * (c % 3);
            o[0] = n.x; o[1] = n.y; o[2] = n.z;
        }
        return;
    }
    uint32_t nv = 0;
    for (uint32_t c = 0; c < 3 * nt; ++c) nv = std::max(nv, (m.vertexIds.empty() ? c : m.vertexIds[c]) + 1);
    std::vector<Vec3> acc(nv, Vec3(0, 0, 0));
    for (uint32_t f = 0; f < nt; ++f) {
        const float* p = &m.positions[9 * f];
        Vec3 p0(p[0], p[1], p[2]), p1(p[3], p[4], p[5]), p2(p[6], p[7], p[8]);
        Vec3 w = cross(p1 - p0, p2 - p0);
        for (int k = 0; k < 3; ++k) {
            uint32_t id = m.vertexIds.empty() ? 3 * f + k : m.vertexIds[3 * f + k];
            acc[id] = acc[id] + w;
        }
    }
    for (auto& a : acc) {
        float len = length(a);
        if (len > 0.0f) a = a / len;
    }
    for (uint32_t c = 0; c < 3 * nt; ++c) {
        const Vec3& n = acc[m.vertexIds.empty() ? c : m.vertexIds[c]];
        float* o = out + 24 * (c / 3) + 9 + 3 * (c % 3);
        o[0] = n.x; o[1] = n.y; o[2] = n.z;
    }
}

const gnd_hip_scene_arrays& HostScene::build(int threads, double* ms) {
    auto t0 = std::chrono::steady_clock::now();
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    if (!built_) {
        for (auto& p : prims) {
            if (p.shape_type == GND_SHAPE_MESH && (p.mesh < 0 || p.mesh >= (int)meshes.size()))
                throw std::runtime_error("primitive references a missing mesh");
            if (p.material < 0 || p.material >= (int)materials.size())
                throw std::runtime_error("primitive references a missing material");
        }
        bool anyEmitter = false;
        for (auto& p : prims) anyEmitter = anyEmitter || p.has_emitter;
        if (!anyEmitter) throw std::runtime_error("scene has no emitter");
        const size_t nm = meshes.size();
        std::vector<MeshBuild> mb(nm);
        // BLAS builds: big meshes use the thread budget internally, small ones run side by side
        {
            std::vector<std::thread> pool;
            std::atomic<size_t> next{0};
            int workers = std::min<int>(threads, (int)std::max<size_t>(nm, 1));
            auto work = [&]() {
                for (;;) {
                    size_t i = next.fetch_add(1);
                    if (i >= nm) break;
                    int inner = meshes[i].ntris() > 100000 ? threads : 1;
                    buildMesh(meshes[i].positions.data(), meshes[i].ntris(), inner, mb[i]);
                }
            };
            for (int w = 1; w < workers; ++w) pool.emplace_back(work);
            work();
            for (auto& t : pool) t.join();
        }
        meshInfo_.assign(nm, gnd_mesh_info{});
        blas_.clear();
        packs_.clear();
        cdf_.clear();
        uint32_t ntot = 0;
        for (size_t i = 0; i < nm; ++i) ntot += meshes[i].ntris();
        shading_.assign((size_t)ntot * 24, 0.0f);
        uint32_t triBase = 0;
        for (size_t i = 0; i < nm; ++i) {
            gnd_mesh_info& mi = meshInfo_[i];
            mi.node_base = (uint32_t)blas_.size();
            mi.n_nodes = (uint32_t)mb[i].nodes.size();
            mi.pack_base = (uint32_t)packs_.size();
            mi.n_packs = (uint32_t)mb[i].packs.size();
            mi.tri_base = triBase;
            mi.n_tris = meshes[i].ntris();
            mi.cdf_base = (uint32_t)cdf_.size();
            mi.has_uv = meshes[i].uvs.empty() ? 0u : 1u;
            mi.total_area = mb[i].totalArea;
            mi.bounds_min[0] = mb[i].bounds.lo.x; mi.bounds_min[1] = mb[i].bounds.lo.y; mi.bounds_min[2] = mb[i].bounds.lo.z;
            mi.bounds_max[0] = mb[i].bounds.hi.x; mi.bounds_max[1] = mb[i].bounds.hi.y; mi.bounds_max[2] = mb[i].bounds.hi.z;
            blas_.insert(blas_.end(), mb[i].nodes.begin(), mb[i].nodes.end());
            packs_.insert(packs_.end(), mb[i].packs.begin(), mb[i].packs.end());
            cdf_.insert(cdf_.end(), mb[i].cdf.begin(), mb[i].cdf.end());
            float* sh = &shading_[(size_t)triBase * 24];
            const MeshData& m = meshes[i];
            for (uint32_t t = 0; t < m.ntris(); ++t) {
                std::memcpy(sh + 24 * t, &m.positions[9 * t], 9 * sizeof(float));
                if (!m.uvs.empty()) std::memcpy(sh + 24 * t + 18, &m.uvs[6 * t], 6 * sizeof(float));
            }
            cornerNormals(m, sh);
            triBase += m.ntris();
        }
        // TLAS over primitive world bounds (bvh.cpp:10-36, primitive.cpp:74-76)
        std::vector<PrimRef> refs(prims.size());
        for (size_t i = 0; i < prims.size(); ++i) {
            const gnd_primitive_desc& p = prims[i];
            Box sb;
            if (p.shape_type == GND_SHAPE_MESH) sb = mb[p.mesh].bounds;
            else {
                sb.lo = Vec3(-p.radius, -p.radius, -p.radius);
                sb.hi = Vec3(p.radius, p.radius, p.radius);
            }
            Mat4 M;
            std::memcpy(M.m, p.to_world, sizeof(M.m));
            Box wb;
            Vec3 c0 = xformPoint(M, sb.corner(0));
            wb.lo = wb.hi = c0;
            for (int c = 1; c < 8; ++c) wb.grow(xformPoint(M, sb.corner(c)));
            refs[i] = {wb, wb.lo * 0.5f + wb.hi * 0.5f, (int)i};
        }
        BinaryBVH tl = buildBinary(refs, 1);
        tlasPrims_.assign(tl.order.begin(), tl.order.end());
        tlas_ = collapseWide(tl.nodes);
        built_ = true;
    }
    const gnd_scene_desc& d = desc();
    arrays_.desc = &d;
    arrays_.n_tlas_nodes = (uint32_t)tlas_.size();
    arrays_.tlas_nodes = tlas_.data();
    arrays_.n_tlas_prims = (uint32_t)tlasPrims_.size();
    arrays_.tlas_prims = tlasPrims_.data();
    arrays_.n_meshes = (uint32_t)meshInfo_.size();
    arrays_.meshes = meshInfo_.data();
    arrays_.n_blas_nodes = (uint32_t)blas_.size();
    arrays_.blas_nodes = blas_.data();
    arrays_.n_packs = (uint32_t)packs_.size();
    arrays_.packs = packs_.data();
    arrays_.n_tris = (uint32_t)(shading_.size() / 24);
    arrays_.shading = shading_.data();
    arrays_.n_cdf = (uint32_t)cdf_.size();
    arrays_.cdf = cdf_.data();
    if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return arrays_;
}

}  // namespace gnd
