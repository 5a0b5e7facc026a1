// scene_host.cpp — owned scene storage and flattening (Scene::activate,
// Mesh::activate and BVH::build of the reference, host half).
#include "scene_host.h"

#include "../device/motion.h"

#include <atomic>
#include <chrono>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <thread>

#include "task_pool.h"

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
    std::memcpy(p.to_world_end, x.m.m, sizeof(p.to_world_end));
    std::memcpy(p.to_world_end_inv, x.inv.m, sizeof(p.to_world_end_inv));
}

void HostScene::setMotion(int prim, const Xform& end) {
    gnd_primitive_desc& p = prims.at((size_t)prim);
    p.has_motion = 1;
    std::memcpy(p.to_world_end, end.m.m, sizeof(p.to_world_end));
    std::memcpy(p.to_world_end_inv, end.inv.m, sizeof(p.to_world_end_inv));
    built_ = false;
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
    std::memcpy(camera.to_world_end, toWorld.m.m, sizeof(camera.to_world_end));
    camera.scan_duration = 0.9f;      // thinLens.cpp:30
    camera.exposure_duration = 0.1f;  // thinLens.cpp:31
}

LightData lightDefaults(int type) {
    LightData l;
    gnd_light_desc& d = l.desc;
    d.type = type;
    d.emitter_index = -1;
    d.direction[0] = 0.0f;
    d.direction[1] = type == GND_LIGHT_SPOT ? 0.0f : -1.0f;
    d.direction[2] = type == GND_LIGHT_SPOT ? -1.0f : 0.0f;
    for (int c = 0; c < 3; ++c) {
        d.power[c] = 10.0f;
        d.irradiance[c] = 1.0f;
    }
    d.spot_size = 45.0f;
    d.blend = 0.15f;
    d.strength = 1.0f;
    return l;
}

int HostScene::emitterCount() const {
    int n = (int)lights.size();
    for (auto& p : prims) n += p.has_emitter ? 1 : 0;
    return n;
}

void HostScene::addLight(LightData l) {
    if (l.desc.type == GND_LIGHT_ENVIRONMENT) {
        for (auto& o : lights)
            if (o.desc.type == GND_LIGHT_ENVIRONMENT)
                throw std::runtime_error("Scene: cannot add multiple infinite area lights!");
        if (l.desc.env_width <= 0 || l.desc.env_height <= 0 ||
            l.env.size() != (size_t)l.desc.env_width * l.desc.env_height * 3)
            throw std::runtime_error("environment light needs a width*height RGB image");
    }
    l.desc.emitter_index = emitterCount();
    lights.push_back(std::move(l));
    built_ = false;
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
    lights.clear();
    for (uint32_t i = 0; i < d.nlights; ++i) {
        LightData l;
        l.desc = d.lights[i];
        if (l.desc.type == GND_LIGHT_ENVIRONMENT && l.desc.env_rgb)
            l.env.assign(l.desc.env_rgb, l.desc.env_rgb + (size_t)l.desc.env_width * l.desc.env_height * 3);
        lights.push_back(std::move(l));
    }
    camera = d.camera;
    integrator = d.integrator;
    maxDepth = d.max_depth;
    aoLength = d.ao_length > 0.0f ? d.ao_length : 5.0f;
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
    lightDescs_.resize(lights.size());
    for (size_t i = 0; i < lights.size(); ++i) {
        lightDescs_[i] = lights[i].desc;
        lightDescs_[i].env_rgb = lights[i].env.empty() ? nullptr : lights[i].env.data();
    }
    desc_.camera = camera;
    desc_.integrator = integrator;
    desc_.max_depth = maxDepth;
    desc_.ao_length = aoLength;
    desc_.spp = spp;
    desc_.nlights = (uint32_t)lights.size();
    desc_.lights = lightDescs_.empty() ? nullptr : lightDescs_.data();
    return desc_;
}

// Corner normals as Mesh::loadOBJ leaves them (mesh.cpp:163-228): given
// normals are normalized, missing ones are generated per shared vertex.
// runs fn(lo, hi) over [0, n) on the pool, or inline
template <class F>
static void forRange(TaskPool* pool, int64_t n, F&& fn) {
    if (pool && n >= 4096) pool->parallelFor(n, fn);
    else fn(0, n);
}

// The smooth normals are sums of the area-weighted face normals of each
// vertex in face order (as a sequential scatter would add them); a
// vertex -> incident-corner list kept in face order lets every vertex sum its
// own faces in that order, in parallel, with the same bits.
static void cornerNormals(const MeshData& m, float* out /* 9 per tri, stride 24 */, TaskPool* pool) {
    const uint32_t nt = m.ntris();
    if (!m.normals.empty()) {
        forRange(pool, 3 * (int64_t)nt, [&](int64_t lo, int64_t hi) {
            for (int64_t c = lo; c < hi; ++c) {
                Vec3 n = normalize(Vec3(m.normals[3 * c], m.normals[3 * c + 1], m.normals[3 * c + 2]));
                float* o = out + 24 * (c / 3) + 9 + 3 * (c % 3);
                o[0] = n.x; o[1] = n.y; o[2] = n.z;
            }
        });
        return;
    }
    auto vid = [&](uint32_t c) { return m.vertexIds.empty() ? c : m.vertexIds[c]; };
    uint32_t nv = 0;
    for (uint32_t c = 0; c < 3 * nt; ++c) nv = std::max(nv, vid(c) + 1);
    std::vector<Vec3> w(nt);
    forRange(pool, nt, [&](int64_t lo, int64_t hi) {
        for (int64_t f = lo; f < hi; ++f) {
            const float* p = &m.positions[9 * f];
            Vec3 p0(p[0], p[1], p[2]), p1(p[3], p[4], p[5]), p2(p[6], p[7], p[8]);
            w[f] = cross(p1 - p0, p2 - p0);
        }
    });
    std::vector<uint32_t> first(nv + 1, 0), faces(3 * (size_t)nt);
    for (uint32_t c = 0; c < 3 * nt; ++c) first[vid(c) + 1]++;
    for (uint32_t v = 0; v < nv; ++v) first[v + 1] += first[v];
    {
        std::vector<uint32_t> fill(first.begin(), first.end() - 1);
        for (uint32_t c = 0; c < 3 * nt; ++c) faces[fill[vid(c)]++] = c / 3;
    }
    std::vector<Vec3> acc(nv);
    forRange(pool, nv, [&](int64_t lo, int64_t hi) {
        for (int64_t v = lo; v < hi; ++v) {
            Vec3 a(0, 0, 0);
            for (uint32_t i = first[v]; i < first[v + 1]; ++i) a = a + w[faces[i]];
            float len = length(a);
            if (len > 0.0f) a = a / len;
            acc[v] = a;
        }
    });
    forRange(pool, 3 * (int64_t)nt, [&](int64_t lo, int64_t hi) {
        for (int64_t c = lo; c < hi; ++c) {
            const Vec3& n = acc[vid((uint32_t)c)];
            float* o = out + 24 * (c / 3) + 9 + 3 * (c % 3);
            o[0] = n.x; o[1] = n.y; o[2] = n.z;
        }
    });
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
        if (emitterCount() == 0) throw std::runtime_error("scene has no emitter");
        {
            std::vector<char> taken((size_t)emitterCount(), 0);
            for (auto& l : lights) {
                const int e = l.desc.emitter_index;
                if (e < 0 || e >= emitterCount() || taken[e]) throw std::runtime_error("bad light emitter_index");
                taken[e] = 1;
            }
        }
        const size_t nm = meshes.size();
        std::vector<MeshBuild> mb(nm);
        // one pool for the whole build: the BLAS builds are tasks, and big
        // meshes split their own work into tasks of the same pool
        std::unique_ptr<TaskPool> pool;
        if (threads > 1) pool = std::make_unique<TaskPool>(threads - 1);
        if (pool) {
            std::atomic<int> pending{(int)nm};
            for (size_t i = 0; i < nm; ++i)
                pool->submit([&, i] {
                    buildMesh(meshes[i].positions.data(), meshes[i].ntris(), pool.get(), mb[i]);
                    pending.fetch_sub(1, std::memory_order_release);
                });
            pool->waitFor(pending);
        } else {
            for (size_t i = 0; i < nm; ++i) buildMesh(meshes[i].positions.data(), meshes[i].ntris(), nullptr, mb[i]);
        }
        // the scene-wide arrays are sized once (no zero fill: every element is
        // written below) and each mesh's part is copied in parallel
        meshInfo_.assign(nm, gnd_mesh_info{});
        size_t nNodes = 0, nPacks = 0, nCdf = 0, nTris = 0;
        for (size_t i = 0; i < nm; ++i) {
            gnd_mesh_info& mi = meshInfo_[i];
            mi.node_base = (uint32_t)nNodes;
            mi.pack_base = (uint32_t)nPacks;
            mi.cdf_base = (uint32_t)nCdf;
            mi.tri_base = (uint32_t)nTris;
            nNodes += mb[i].nodes.size();
            nPacks += mb[i].packs.size();
            nCdf += mb[i].cdf.size();
            nTris += meshes[i].ntris();
        }
        blas_.clear();
        packs_.clear();
        cdf_.clear();
        shading_.clear();
        blas_.resize(nNodes);
        packs_.resize(nPacks);
        cdf_.resize(nCdf);
        shading_.resize(nTris * 24);
        for (size_t i = 0; i < nm; ++i) {
            gnd_mesh_info& mi = meshInfo_[i];
            mi.n_nodes = (uint32_t)mb[i].nodes.size();
            mi.n_packs = (uint32_t)mb[i].packs.size();
            mi.n_tris = meshes[i].ntris();
            mi.has_uv = meshes[i].uvs.empty() ? 0u : 1u;
            mi.total_area = mb[i].totalArea;
            mi.bounds_min[0] = mb[i].bounds.lo.x; mi.bounds_min[1] = mb[i].bounds.lo.y; mi.bounds_min[2] = mb[i].bounds.lo.z;
            mi.bounds_max[0] = mb[i].bounds.hi.x; mi.bounds_max[1] = mb[i].bounds.hi.y; mi.bounds_max[2] = mb[i].bounds.hi.z;
            forRange(pool.get(), (int64_t)mb[i].nodes.size(), [&](int64_t lo, int64_t hi) {
                std::copy(mb[i].nodes.begin() + lo, mb[i].nodes.begin() + hi, blas_.begin() + mi.node_base + lo);
            });
            forRange(pool.get(), (int64_t)mb[i].packs.size(), [&](int64_t lo, int64_t hi) {
                std::copy(mb[i].packs.begin() + lo, mb[i].packs.begin() + hi, packs_.begin() + mi.pack_base + lo);
            });
            std::copy(mb[i].cdf.begin(), mb[i].cdf.end(), cdf_.begin() + mi.cdf_base);
            float* sh = &shading_[(size_t)mi.tri_base * 24];
            const MeshData& m = meshes[i];
            forRange(pool.get(), m.ntris(), [&](int64_t lo, int64_t hi) {
                for (int64_t t = lo; t < hi; ++t) {
                    std::memcpy(sh + 24 * t, &m.positions[9 * t], 9 * sizeof(float));
                    if (!m.uvs.empty()) std::memcpy(sh + 24 * t + 18, &m.uvs[6 * t], 6 * sizeof(float));
                    else std::memset(sh + 24 * t + 18, 0, 6 * sizeof(float));
                }
            });
            cornerNormals(m, sh, pool.get());
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
            if (p.has_motion && std::memcmp(p.to_world, p.to_world_end, sizeof(p.to_world)) != 0) {
                // AnimatedTransform::motionBounds (animated.cpp:49-60): the
                // box at time 0, at time 1 and at the interpolated time 0.5
                GndMotion mo;
                gndMotionFrom(p.to_world, p.to_world_end, mo);
                Mat4 ends[2];
                std::memcpy(ends[0].m, p.to_world_end, sizeof(ends[0].m));
                gndMotionMatrix(mo, 0.5f, ends[1].m);
                for (const Mat4& E : ends)
                    for (int c = 0; c < 8; ++c) wb.grow(xformPoint(E, sb.corner(c)));
            }
            refs[i] = {wb, wb.lo * 0.5f + wb.hi * 0.5f, (int)i};
        }
        BinaryBVH tl = buildBinary(refs, nullptr);
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
