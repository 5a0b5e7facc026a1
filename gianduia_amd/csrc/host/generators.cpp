// generators.cpp — procedural scenes of the benchmark workloads, built in code
// with the reference's object model (primitives = mesh/sphere + transform +
// material + optional area emitter).  Sizes and seeds are fixed here so the
// CPU restatement and the GPU render the same bytes.
#include <cstdio>
#include <map>
#include <random>
#include <sstream>
#include <stdexcept>

#include "scene_host.h"

namespace gnd {

namespace {

using Params = std::map<std::string, std::string>;

Params parseParams(const std::string& s) {
    Params p;
    std::stringstream ss(s);
    std::string kv;
    while (std::getline(ss, kv, ';')) {
        auto e = kv.find('=');
        if (e == std::string::npos) continue;
        p[kv.substr(0, e)] = kv.substr(e + 1);
    }
    return p;
}
double num(const Params& p, const char* k, double dflt) {
    auto it = p.find(k);
    return it == p.end() ? dflt : std::stod(it->second);
}

void pushTri(MeshData& m, Vec3 a, Vec3 b, Vec3 c, Vec3 n) {
    for (Vec3 v : {a, b, c}) {
        m.positions.push_back(v.x); m.positions.push_back(v.y); m.positions.push_back(v.z);
        m.normals.push_back(n.x); m.normals.push_back(n.y); m.normals.push_back(n.z);
    }
}

// filter=gaussian keeps the film default (Gaussian, radius 2, alpha 2);
// anything else, or no key, selects the box filter the workloads use.
void applyFilter(HostScene& s, const Params& p) {
    auto it = p.find("filter");
    if (it == p.end() || it->second != "gaussian") s.camera.filter = GND_FILTER_BOX;
}

// Quad a-b-c-d (counter-clockwise seen from the side the normal points to)
void pushQuad(MeshData& m, Vec3 a, Vec3 b, Vec3 c, Vec3 d, Vec3 n) {
    pushTri(m, a, b, c, n);
    pushTri(m, a, c, d, n);
}

MeshData quadMesh(Vec3 a, Vec3 b, Vec3 c, Vec3 d, Vec3 n) {
    MeshData m;
    pushQuad(m, a, b, c, d, n);
    return m;
}

// Unit cube [-0.5, 0.5]^3 with outward face normals (12 triangles)
MeshData cubeMesh() {
    MeshData m;
    const float h = 0.5f;
    Vec3 p[8];
    for (int i = 0; i < 8; ++i) p[i] = Vec3((i & 1) ? h : -h, (i & 2) ? h : -h, (i & 4) ? h : -h);
    pushQuad(m, p[0], p[4], p[6], p[2], Vec3(-1, 0, 0));
    pushQuad(m, p[1], p[3], p[7], p[5], Vec3(1, 0, 0));
    pushQuad(m, p[0], p[1], p[5], p[4], Vec3(0, -1, 0));
    pushQuad(m, p[2], p[6], p[7], p[3], Vec3(0, 1, 0));
    pushQuad(m, p[0], p[2], p[3], p[1], Vec3(0, 0, -1));
    pushQuad(m, p[4], p[5], p[7], p[6], Vec3(0, 0, 1));
    return m;
}

// Seeded 3D value noise (hashed lattice values, smoothstep trilinear blend)
struct ValueNoise {
    uint32_t seed;
    explicit ValueNoise(uint32_t s) : seed(s) {}
    float lattice(int x, int y, int z) const {
        uint32_t h = seed * 0x9E3779B1u;
        h ^= (uint32_t)x * 0x85EBCA77u;
        h = (h << 13) | (h >> 19);
        h ^= (uint32_t)y * 0xC2B2AE3Du;
        h = (h << 13) | (h >> 19);
        h ^= (uint32_t)z * 0x27D4EB2Fu;
        h ^= h >> 15;
        h *= 0x2C1B3C6Du;
        h ^= h >> 12;
        h *= 0x297A2D39u;
        h ^= h >> 15;
        return (h & 0xFFFFFF) / float(0xFFFFFF);  // [0,1]
    }
    float eval(float x, float y, float z) const {
        int ix = (int)std::floor(x), iy = (int)std::floor(y), iz = (int)std::floor(z);
        float fx = x - ix, fy = y - iy, fz = z - iz;
        auto s = [](float t) { return t * t * (3.0f - 2.0f * t); };
        float sx = s(fx), sy = s(fy), sz = s(fz);
        float r = 0.0f;
        for (int k = 0; k < 8; ++k) {
            int dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
            float w = (dx ? sx : 1 - sx) * (dy ? sy : 1 - sy) * (dz ? sz : 1 - sz);
            r += w * lattice(ix + dx, iy + dy, iz + dz);
        }
        return r;
    }
    // octaves of value noise in [-1, 1] (amplitude halves, frequency doubles)
    float fbm(Vec3 p, int octaves, float baseFreq) const {
        float sum = 0, amp = 1, norm = 0, f = baseFreq;
        for (int o = 0; o < octaves; ++o) {
            sum += amp * (2.0f * eval(p.x * f + 17.0f * o, p.y * f + 31.0f * o, p.z * f + 47.0f * o) - 1.0f);
            norm += amp;
            amp *= 0.5f;
            f *= 2.0f;
        }
        return sum / norm;
    }
};

// Icosphere: icosahedron subdivided `level` times (20 * 4^level triangles),
// vertices displaced along the radius by amplitude * fbm.  Vertices are shared
// (vertexIds) so shading normals are generated smooth, as for an OBJ without
// normals.
MeshData icosphere(int level, float amplitude, uint32_t seed, int octaves) {
    const float t = (1.0f + std::sqrt(5.0f)) / 2.0f;
    std::vector<Vec3> v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                           {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& p : v) p = normalize(p);
    std::vector<uint32_t> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2,
                               10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5,
                               2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (int l = 0; l < level; ++l) {
        std::map<uint64_t, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            v.push_back(normalize((v[a] + v[b]) * 0.5f));
            uint32_t id = (uint32_t)v.size() - 1;
            mid[key] = id;
            return id;
        };
        std::vector<uint32_t> nf;
        nf.reserve(f.size() * 4);
        for (size_t i = 0; i < f.size(); i += 3) {
            uint32_t a = f[i], b = f[i + 1], c = f[i + 2];
            uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            uint32_t tri[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), tri, tri + 12);
        }
        f.swap(nf);
    }
    ValueNoise noise(seed);
    for (auto& p : v) p = p * (1.0f + amplitude * noise.fbm(p, octaves, 2.0f));
    MeshData m;
    m.positions.reserve(f.size() * 3);
    m.vertexIds.assign(f.begin(), f.end());
    for (uint32_t id : f) {
        m.positions.push_back(v[id].x); m.positions.push_back(v[id].y); m.positions.push_back(v[id].z);
    }
    return m;
}

// Latitude-longitude sphere with `seg` segments and `rings` rings:
// 2*seg*(rings-1) triangles (seg=100, rings=51 -> 10,000), displaced radially.
MeshData latlongSphere(int seg, int rings, float amplitude, uint32_t seed) {
    std::vector<Vec3> v;
    ValueNoise noise(seed);
    for (int r = 0; r <= rings; ++r) {
        float th = 3.14159265358979f * r / rings;
        for (int s = 0; s < seg; ++s) {
            float ph = 2.0f * 3.14159265358979f * s / seg;
            Vec3 p(std::sin(th) * std::cos(ph), std::cos(th), std::sin(th) * std::sin(ph));
            if (r == 0 || r == rings) p = Vec3(0, r == 0 ? 1.0f : -1.0f, 0);
            v.push_back(p * (1.0f + amplitude * noise.fbm(p, 4, 3.0f)));
        }
    }
    MeshData m;
    auto id = [&](int r, int s) { return (uint32_t)(r * seg + (s % seg)); };
    auto tri = [&](uint32_t a, uint32_t b, uint32_t c) {
        for (uint32_t k : {a, b, c}) {
            m.positions.push_back(v[k].x); m.positions.push_back(v[k].y); m.positions.push_back(v[k].z);
            m.vertexIds.push_back(k);
        }
    };
    for (int r = 0; r < rings; ++r)
        for (int s = 0; s < seg; ++s) {
            if (r != 0) tri(id(r, s), id(r, s + 1), id(r + 1, s));
            if (r != rings - 1) tri(id(r, s + 1), id(r + 1, s + 1), id(r + 1, s));
        }
    return m;
}

// ---------------------------------------------------------------- scenes

// Cornell box: walls = five faces of [-1,1]^3 (front open), short box 0.6^3
// rotated +15 deg, tall box 0.6x1.2x0.6 rotated -18 deg, 0.5x0.5 ceiling light
// at y = 0.999 facing -y with radiance (17,12,4), pinhole fov 40 at (0,0,3.9).
std::unique_ptr<HostScene> cornellBox(const Params& p) {
    auto s = std::make_unique<HostScene>();
    const float R[3] = {0.63f, 0.065f, 0.05f}, G[3] = {0.14f, 0.45f, 0.09f};
    int white = s->addMaterial(materialMatte(0.73f, 0.73f, 0.73f));
    int red = s->addMaterial(materialMatte(R[0], R[1], R[2]));
    int green = s->addMaterial(materialMatte(G[0], G[1], G[2]));
    int black = s->addMaterial(materialMatte(0.0f, 0.0f, 0.0f));
    Xform I;
    Vec3 c[8];
    for (int i = 0; i < 8; ++i) c[i] = Vec3((i & 1) ? 1.f : -1.f, (i & 2) ? 1.f : -1.f, (i & 4) ? 1.f : -1.f);
    // floor (normal +y), ceiling (-y), back (+z), left (+x), right (-x)
    s->addPrim(s->addMesh(quadMesh(c[0], c[4], c[5], c[1], Vec3(0, 1, 0))), I, white);
    s->addPrim(s->addMesh(quadMesh(c[2], c[3], c[7], c[6], Vec3(0, -1, 0))), I, white);
    s->addPrim(s->addMesh(quadMesh(c[0], c[1], c[3], c[2], Vec3(0, 0, 1))), I, white);
    s->addPrim(s->addMesh(quadMesh(c[0], c[2], c[6], c[4], Vec3(1, 0, 0))), I, red);
    s->addPrim(s->addMesh(quadMesh(c[1], c[5], c[7], c[3], Vec3(-1, 0, 0))), I, green);
    int cube = s->addMesh(cubeMesh());
    Xform shortBox = xformScale(0.6f, 0.6f, 0.6f).then(xformRotate(15.0f, 0, 1, 0)).then(xformTranslate(0.33f, -0.7f, 0.35f));
    Xform tallBox = xformScale(0.6f, 1.2f, 0.6f).then(xformRotate(-18.0f, 0, 1, 0)).then(xformTranslate(-0.35f, -0.4f, -0.3f));
    const int sb = s->addPrim(cube, shortBox, white);
    s->addPrim(cube, tallBox, white);
    // movebox=d: the short box slides d along x and turns 20 degrees about y
    // over the shutter interval (motion blur)
    if (const double mv = num(p, "movebox", 0.0); mv != 0.0)
        s->setMotion(sb, xformScale(0.6f, 0.6f, 0.6f).then(xformRotate(35.0f, 0, 1, 0))
                             .then(xformTranslate(0.33f + (float)mv, -0.7f, 0.35f)));
    const float Le[3] = {17.0f, 12.0f, 4.0f};
    const float h = 0.25f, y = 0.999f;
    s->addPrim(s->addMesh(quadMesh(Vec3(-h, y, -h), Vec3(h, y, -h), Vec3(h, y, h), Vec3(-h, y, h), Vec3(0, -1, 0))), I,
               black, Le);
    int W = (int)num(p, "width", 512), H = (int)num(p, "height", 512);
    s->setCamera(GND_CAMERA_PERSPECTIVE, W, H, 40.0f, xformLookAt(Vec3(0, 0, 3.9f), Vec3(0, 0, 0), Vec3(0, 1, 0)));
    applyFilter(*s, p);
    s->spp = (uint32_t)num(p, "spp", 64);
    s->maxDepth = (int)num(p, "depth", 8);
    s->integrator = GND_INTEGRATOR_PATH;
    return s;
}

// Scaled geometry: a displaced icosphere (level 8 = 1,310,720 triangles by
// default; `level` overrides) and `instances` (64) randomly transformed copies of
// a 10,000-triangle displaced sphere, each object drawn from conductor / glass
// (IOR 1.5) / plastic with GGX roughness U[0.05, 0.5]; a floor quad and one
// quad light; thin-lens camera, aperture radius 0.05.
std::unique_ptr<HostScene> scaledScene(const Params& p) {
    auto s = std::make_unique<HostScene>();
    uint32_t seed = (uint32_t)num(p, "seed", 0);
    int level = (int)num(p, "level", 8);
    int ninst = (int)num(p, "instances", 64);
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> U(0.0f, 1.0f);
    const float goldEta[3] = {0.143f, 0.375f, 1.442f}, goldK[3] = {3.983f, 2.386f, 1.603f};
    auto randomMaterial = [&]() {
        int kind = std::min(2, (int)(U(rng) * 3.0f));
        float rough = 0.05f + 0.45f * U(rng);
        if (kind == 0) {
            float Rc[3] = {0.6f + 0.4f * U(rng), 0.6f + 0.4f * U(rng), 0.6f + 0.4f * U(rng)};
            return materialConductor(Rc, goldEta, goldK, rough);
        }
        if (kind == 1) return materialGlass(1.5f, rough);
        float Kd[3] = {0.1f + 0.7f * U(rng), 0.1f + 0.7f * U(rng), 0.1f + 0.7f * U(rng)};
        float Ks[3] = {0.25f, 0.25f, 0.25f};
        return materialPlastic(Kd, Ks, 1.5f, rough);
    };
    int big = s->addMesh(icosphere(level, 0.1f, seed + 1, 4));
    s->addPrim(big, Xform(), s->addMaterial(randomMaterial()));
    int small = s->addMesh(latlongSphere(100, 51, 0.1f, seed + 2));
    for (int i = 0; i < ninst; ++i) {
        float sc = 0.08f + 0.14f * U(rng);
        float ang = 360.0f * U(rng);
        Vec3 axis(U(rng) - 0.5f, U(rng) - 0.5f, U(rng) - 0.5f);
        if (length(axis) < 1e-3f) axis = Vec3(0, 1, 0);
        float phi = 6.2831853f * U(rng), rad = 1.35f + 1.6f * U(rng), yy = -0.9f + 1.9f * U(rng);
        Xform x = xformScale(sc, sc, sc)
                      .then(xformRotate(ang, axis.x, axis.y, axis.z))
                      .then(xformTranslate(rad * std::cos(phi), yy, rad * std::sin(phi)));
        const int id = s->addPrim(small, x, s->addMaterial(randomMaterial()));
        // instmove=d: every instance moves d along a random direction (motion blur)
        if (const double mv = num(p, "instmove", 0.0); mv != 0.0) {
            std::uniform_real_distribution<float> Um(-1.0f, 1.0f);
            std::mt19937 mrng(seed * 7919u + (uint32_t)i);
            Vec3 dir(Um(mrng), Um(mrng), Um(mrng));
            if (length(dir) < 1e-3f) dir = Vec3(1, 0, 0);
            dir = normalize(dir) * (float)mv;
            s->setMotion(id, x.then(xformTranslate(dir.x, dir.y, dir.z)));
        }
    }
    int floorMat = s->addMaterial(materialMatte(0.5f, 0.5f, 0.5f));
    const float F = 6.0f, fy = -1.25f;
    s->addPrim(s->addMesh(quadMesh(Vec3(-F, fy, -F), Vec3(-F, fy, F), Vec3(F, fy, F), Vec3(F, fy, -F), Vec3(0, 1, 0))),
               Xform(), floorMat);
    const float Le[3] = {12.0f, 11.0f, 10.0f};
    const float L = 1.0f, ly = 3.5f;
    s->addPrim(s->addMesh(quadMesh(Vec3(-L, ly, -L), Vec3(L, ly, -L), Vec3(L, ly, L), Vec3(-L, ly, L), Vec3(0, -1, 0))),
               Xform(), s->addMaterial(materialMatte(0, 0, 0)), Le);
    int W = (int)num(p, "width", 1920), H = (int)num(p, "height", 1080);
    s->setCamera(GND_CAMERA_THINLENS, W, H, 40.0f, xformLookAt(Vec3(0, 1.2f, 5.5f), Vec3(0, 0, 0), Vec3(0, 1, 0)));
    s->camera.lens_radius = 0.05f;
    s->camera.focal_distance = 5.6f;
    applyFilter(*s, p);
    s->spp = (uint32_t)num(p, "spp", 16);
    s->maxDepth = (int)num(p, "depth", 16);
    s->integrator = GND_INTEGRATOR_PATH;
    return s;
}

// Volumetric: the Cornell box plus a smooth glass sphere (IOR 1.5) holding a
// homogeneous medium (sigma_t = 2, albedo 0.8, HG g = 0.3) and an index-material
// box holding an n^3 (256^3) density grid of seeded 5-octave fBm clamped to
// [0,1] with sigma_t max 8 (albedo 0.8); volpath, 1280x720, 32 spp, depth 32.
std::unique_ptr<HostScene> volumetricScene(const Params& p) {
    Params q = p;
    if (!q.count("width")) q["width"] = "1280";
    if (!q.count("height")) q["height"] = "720";
    auto s = cornellBox(q);
    uint32_t seed = (uint32_t)num(p, "seed", 0);
    int n = (int)num(p, "grid", 256);
    MediumData hom;
    hom.desc.type = GND_MEDIUM_HOMOGENEOUS;
    for (int c = 0; c < 3; ++c) { hom.desc.sigma_s[c] = 1.6f; hom.desc.sigma_a[c] = 0.4f; }
    hom.desc.g = 0.3f;
    Mat4 I = Mat4::identity();
    std::memcpy(hom.desc.to_local, I.m, sizeof(I.m));
    std::memcpy(hom.desc.to_local_inv, I.m, sizeof(I.m));
    s->media.push_back(std::move(hom));
    int glass = s->addMaterial(materialGlass(1.5f, 0.0f, 0, -1));
    s->addSphere(0.3f, xformTranslate(0.35f, -0.1f, 0.35f), glass);
    // density grid inside the box [-0.85,-0.05] x [-1,-0.2] x [-0.2,0.6]
    // (grid=0 leaves it out: homogeneous medium only)
    if (n > 0) {
        MediumData grid;
        grid.desc.type = GND_MEDIUM_GRID;
        for (int c = 0; c < 3; ++c) { grid.desc.sigma_s[c] = 6.4f; grid.desc.sigma_a[c] = 1.6f; }
        grid.desc.g = 0.0f;
        grid.desc.nx = grid.desc.ny = grid.desc.nz = n;
        grid.desc.density_scale = 1.0f;
        grid.desc.macro_res = 32;
        grid.density.resize((size_t)n * n * n);
        ValueNoise noise(seed + 7);
        for (int z = 0; z < n; ++z)
            for (int y = 0; y < n; ++y)
                for (int x = 0; x < n; ++x) {
                    Vec3 u((x + 0.5f) / n, (y + 0.5f) / n, (z + 0.5f) / n);
                    float v = 0.5f + 0.5f * noise.fbm(u, 5, 4.0f);
                    v = (v - 0.35f) * 2.0f;
                    grid.density[((size_t)z * n + y) * n + x] = std::min(1.0f, std::max(0.0f, v));
                }
        const float bx0 = -0.85f, by0 = -1.0f, bz0 = -0.2f, ext = 0.8f;
        // world -> index: (p - b0) * n / ext
        Xform toLocal = xformTranslate(-bx0, -by0, -bz0).then(xformScale(n / ext, n / ext, n / ext));
        std::memcpy(grid.desc.to_local, toLocal.m.m, sizeof(toLocal.m.m));
        std::memcpy(grid.desc.to_local_inv, toLocal.inv.m, sizeof(toLocal.inv.m));
        s->media.push_back(std::move(grid));
        int index = s->addMaterial(materialIndex(1, -1));
        Xform boxX = xformScale(ext, ext, ext)
                         .then(xformTranslate(bx0 + ext * 0.5f, by0 + ext * 0.5f, bz0 + ext * 0.5f));
        s->addPrim(s->addMesh(cubeMesh()), boxX, index);
    }
    s->spp = (uint32_t)num(p, "spp", 32);
    s->maxDepth = (int)num(p, "depth", 32);
    s->integrator = GND_INTEGRATOR_VOLPATH;
    return s;
}

// Scene-pool generator: the scaled generator at a target triangle count
// (icosphere level and instance count chosen to land near `tris`).
std::unique_ptr<HostScene> proceduralScene(const Params& p) {
    double tris = num(p, "tris", 1e5);
    Params q = p;
    int level = 2;
    while (level < 9 && 20.0 * std::pow(4.0, level + 1) <= tris * 0.8) ++level;
    double rest = std::max(0.0, tris - 20.0 * std::pow(4.0, level));
    q["level"] = std::to_string(level);
    q["instances"] = std::to_string((int)std::min(1024.0, std::round(rest / 10000.0)));
    return scaledScene(q);
}

// Procedural sky for the environment light: a width 2n x height n
// latitude-longitude image (row 0 = +z, the reference's spherical
// coordinates), y up: a horizon-to-zenith gradient above the horizon, a dim
// ground below it and a small bright sun disk, so that the map's importance
// sampling has a peak to find.
std::vector<float> skyImage(int n) {
    const int W = 2 * n, H = n;
    std::vector<float> img((size_t)W * H * 3);
    const Vec3 sun = normalize(Vec3(0.4f, 0.8f, 0.3f));
    const float cosSun = std::cos(3.0f * 3.14159265f / 180.0f);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const float th = 3.14159265f * (y + 0.5f) / H, ph = 2.0f * 3.14159265f * (x + 0.5f) / W;
            const Vec3 d(std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th));
            Vec3 c;
            if (d.y >= 0.0f) {
                const float t = d.y;
                c = Vec3(0.9f, 0.9f, 1.0f) * (1.0f - t) + Vec3(0.3f, 0.5f, 0.9f) * t;
            } else {
                c = Vec3(0.25f, 0.22f, 0.2f);
            }
            if (dot(d, sun) > cosSun) c = c + Vec3(50.0f, 45.0f, 40.0f);
            float* o = &img[3 * ((size_t)y * W + x)];
            o[0] = c.x; o[1] = c.y; o[2] = c.z;
        }
    return img;
}

// lights=<comma list> appends non-area emitters to any generated scene,
// after its area emitters, in list order: point, spot, sun (directional),
// env (the procedural sky, envres=n rows, default 64; envstrength).
void addLights(HostScene& s, const Params& p) {
    auto it = p.find("lights");
    if (it == p.end() || it->second.empty()) return;
    std::stringstream ss(it->second);
    std::string k;
    while (std::getline(ss, k, ',')) {
        LightData l;
        if (k == "point") {
            l = lightDefaults(GND_LIGHT_POINT);
            const float pos[3] = {0.3f, 0.7f, 0.6f}, pw[3] = {6.0f, 5.5f, 5.0f};
            for (int c = 0; c < 3; ++c) { l.desc.position[c] = pos[c]; l.desc.power[c] = pw[c]; }
        } else if (k == "spot") {
            l = lightDefaults(GND_LIGHT_SPOT);
            const float pos[3] = {-0.4f, 0.8f, 1.6f}, dir[3] = {0.3f, -0.7f, -1.0f}, pw[3] = {30.0f, 28.0f, 24.0f};
            for (int c = 0; c < 3; ++c) {
                l.desc.position[c] = pos[c];
                l.desc.direction[c] = dir[c];
                l.desc.power[c] = pw[c];
            }
            l.desc.spot_size = 50.0f;
            l.desc.blend = 0.3f;
        } else if (k == "sun") {
            l = lightDefaults(GND_LIGHT_DIRECTIONAL);
            const float dir[3] = {-0.3f, -1.0f, -0.4f}, irr[3] = {2.0f, 1.9f, 1.7f};
            for (int c = 0; c < 3; ++c) { l.desc.direction[c] = dir[c]; l.desc.irradiance[c] = irr[c]; }
        } else if (k == "env") {
            l = lightDefaults(GND_LIGHT_ENVIRONMENT);
            const int n = (int)num(p, "envres", 64);
            if (n < 1 || n > 8192) throw std::runtime_error("envres out of range");
            l.env = skyImage(n);
            l.desc.env_width = 2 * n;
            l.desc.env_height = n;
            l.desc.strength = (float)num(p, "envstrength", 1.0);
        } else {
            throw std::runtime_error("unknown light kind: " + k);
        }
        s.addLight(std::move(l));
    }
}

}  // namespace

std::unique_ptr<HostScene> generateScene(const std::string& name, const std::string& params) {
    Params p = parseParams(params);
    std::unique_ptr<HostScene> s;
    if (name == "cbox") s = cornellBox(p);
    else if (name == "scaled") s = scaledScene(p);
    else if (name == "volumetric") s = volumetricScene(p);
    else if (name == "procedural") s = proceduralScene(p);
    else throw std::runtime_error("unknown scene generator: " + name);
    addLights(*s, p);
    // integrator=path|volpath|direct|ao (aolength=) swaps the scene's integrator
    auto it = p.find("integrator");
    if (it != p.end()) {
        const std::string& t = it->second;
        if (t == "path") s->integrator = GND_INTEGRATOR_PATH;
        else if (t == "volpath") s->integrator = GND_INTEGRATOR_VOLPATH;
        else if (t == "direct") s->integrator = GND_INTEGRATOR_DIRECT;
        else if (t == "ao") s->integrator = GND_INTEGRATOR_AO;
        else throw std::runtime_error("unknown integrator: " + t);
    }
    s->aoLength = (float)num(p, "aolength", 5.0);
    // cammove=<distance>: the camera slides that far along its right axis and
    // turns by camturn degrees about world y over the shutter interval
    // (motion blur); rolling=1 turns on the thin lens's rolling shutter
    const double move = num(p, "cammove", 0.0), turn = num(p, "camturn", 0.0);
    if (move != 0.0 || turn != 0.0) {
        Xform start;  // only the forward matrix matters here
        std::memcpy(start.m.m, s->camera.to_world, sizeof(start.m.m));
        const Xform end = start.then(xformTranslate((float)move, 0.0f, 0.0f))
                              .then(xformRotate((float)turn, 0.0f, 1.0f, 0.0f));
        s->camera.has_motion = 1;
        std::memcpy(s->camera.to_world_end, end.m.m, sizeof(s->camera.to_world_end));
    }
    if (num(p, "rolling", 0.0) != 0.0) s->camera.rolling_shutter = 1;
    return s;
}

}  // namespace gnd
