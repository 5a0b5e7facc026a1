// xml_loader.cpp — the reference's scene XML schema (src/core/parser.cpp:41-235)
// read with a small in-tree XML reader (pugixml is not part of this build).
//
// Supported: <scene>, <camera type="perspective|thinlens"> with <filter
// type="box|gaussian">, <sampler type="independent">, <integrator
// type="path|volpath|direct|ao" (or with a _hip suffix)>, <primitive> with <shape
// type="mesh"> (OBJ file) or <shape type="sphere">, <material
// type="matte|conductor|glass|plastic|index"> with nested <medium
// name="inside|outside" type="homogeneous">, <emitter type="area"> inside a
// primitive, scene-level <emitter type="point|spot|directional|environment">
// (the environment image an OpenEXR file), <ref id>
// sharing of shapes, and <transform> with translate/scale/rotate/lookat steps
// composed as t = step * t (parser.cpp:112-165).
#include <algorithm>
#include <cctype>
#include <cstring>
#include <deque>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

#include "../../../include/gnd_host.h"
#include "integrator.h"
#include "scene_host.h"

namespace gnd {

namespace {

struct XNode {
    std::string tag;
    std::map<std::string, std::string> attr;
    std::vector<XNode> kids;
    std::string attrOr(const char* k, const std::string& d = "") const {
        auto it = attr.find(k);
        return it == attr.end() ? d : it->second;
    }
};

class XmlReader {
public:
    explicit XmlReader(const std::string& text) : s_(text) {}
    XNode parseDocument() {
        skipMisc();
        XNode root = parseElement();
        return root;
    }

private:
    const std::string& s_;
    size_t i_ = 0;

    [[noreturn]] void fail(const std::string& m) {
        throw std::runtime_error("XML parsing failed: " + m + " at offset " + std::to_string(i_));
    }
    void skipWs() {
        while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
    }
    bool startsWith(const char* p) const { return s_.compare(i_, std::strlen(p), p) == 0; }
    void skipMisc() {
        for (;;) {
            skipWs();
            if (startsWith("<?")) {
                size_t e = s_.find("?>", i_);
                if (e == std::string::npos) fail("unterminated declaration");
                i_ = e + 2;
            } else if (startsWith("<!--")) {
                size_t e = s_.find("-->", i_);
                if (e == std::string::npos) fail("unterminated comment");
                i_ = e + 3;
            } else if (startsWith("<!")) {
                size_t e = s_.find('>', i_);
                if (e == std::string::npos) fail("unterminated directive");
                i_ = e + 1;
            } else {
                return;
            }
        }
    }
    std::string name() {
        size_t b = i_;
        while (i_ < s_.size() && (std::isalnum((unsigned char)s_[i_]) || s_[i_] == '_' || s_[i_] == '-' ||
                                  s_[i_] == ':' || s_[i_] == '.'))
            ++i_;
        if (b == i_) fail("expected a name");
        return s_.substr(b, i_ - b);
    }
    static std::string unescape(const std::string& v) {
        std::string o;
        for (size_t k = 0; k < v.size(); ++k) {
            if (v[k] == '&') {
                size_t e = v.find(';', k);
                std::string ent = v.substr(k + 1, e - k - 1);
                if (ent == "lt") o += '<';
                else if (ent == "gt") o += '>';
                else if (ent == "amp") o += '&';
                else if (ent == "quot") o += '"';
                else if (ent == "apos") o += '\'';
                k = e;
            } else o += v[k];
        }
        return o;
    }
    XNode parseElement() {
        if (i_ >= s_.size() || s_[i_] != '<') fail("expected '<'");
        ++i_;
        XNode n;
        n.tag = name();
        for (;;) {
            skipWs();
            if (startsWith("/>")) { i_ += 2; return n; }
            if (i_ < s_.size() && s_[i_] == '>') { ++i_; break; }
            std::string k = name();
            skipWs();
            if (i_ >= s_.size() || s_[i_] != '=') fail("expected '='");
            ++i_;
            skipWs();
            char q = s_[i_];
            if (q != '"' && q != '\'') fail("expected a quoted value");
            size_t e = s_.find(q, i_ + 1);
            if (e == std::string::npos) fail("unterminated attribute");
            n.attr[k] = unescape(s_.substr(i_ + 1, e - i_ - 1));
            i_ = e + 1;
        }
        for (;;) {
            skipMisc();
            if (startsWith("</")) {
                i_ += 2;
                std::string c = name();
                if (c != n.tag) fail("mismatched </" + c + ">");
                skipWs();
                if (i_ >= s_.size() || s_[i_] != '>') fail("expected '>'");
                ++i_;
                return n;
            }
            if (i_ < s_.size() && s_[i_] == '<') {
                n.kids.push_back(parseElement());
            } else {
                size_t e = s_.find('<', i_);  // text content is ignored by the schema
                if (e == std::string::npos) fail("unterminated element");
                i_ = e;
            }
        }
    }
};

// parseVector3 (parser.cpp:15-26): "x, y, z" or "x y z"
Vec3 parseVec3(const std::string& v) {
    std::stringstream ss(v);
    float x, y, z;
    char c1, c2;
    if ((ss >> x >> c1 >> y >> c2 >> z) && c1 == ',' && c2 == ',') return Vec3(x, y, z);
    std::stringstream s2(v);
    if (!(s2 >> x >> y >> z)) throw std::runtime_error("Could not parse vector from string: " + v);
    return Vec3(x, y, z);
}
float asFloat(const std::string& v, float d = 0.0f) { return v.empty() ? d : std::stof(v); }

struct Props {
    std::map<std::string, float> f;
    std::map<std::string, int> i;
    std::map<std::string, bool> b;
    std::map<std::string, std::string> s;
    std::map<std::string, Vec3> v;
    std::map<std::string, Xform> t;
    float getF(const char* k, float d) const { auto it = f.find(k); return it == f.end() ? d : it->second; }
    int getI(const char* k, int d) const { auto it = i.find(k); return it == i.end() ? d : it->second; }
    bool getB(const char* k, bool d) const { auto it = b.find(k); return it == b.end() ? d : it->second; }
    bool hasV(const char* k) const { return v.count(k) > 0; }
    Vec3 getV(const char* k, Vec3 d) const { auto it = v.find(k); return it == v.end() ? d : it->second; }
    std::string getS(const char* k, const std::string& d) const { auto it = s.find(k); return it == s.end() ? d : it->second; }
};

Props readProps(const XNode& n, std::vector<const XNode*>& children) {
    Props p;
    for (const XNode& c : n.kids) {
        const std::string nm = c.attrOr("name");
        if (c.tag == "boolean") p.b[nm] = c.attrOr("value") == "true" || c.attrOr("value") == "1";
        else if (c.tag == "integer") p.i[nm] = std::stoi(c.attrOr("value", "0"));
        else if (c.tag == "float") p.f[nm] = asFloat(c.attrOr("value"));
        else if (c.tag == "string") p.s[nm] = c.attrOr("value");
        else if (c.tag == "point" || c.tag == "vector" || c.tag == "vector3" || c.tag == "color") {
            if (c.attr.count("value")) p.v[nm] = parseVec3(c.attrOr("value"));
            else if (c.tag == "color")
                p.v[nm] = Vec3(asFloat(c.attrOr("r")), asFloat(c.attrOr("g")), asFloat(c.attrOr("b")));
            else p.v[nm] = Vec3(asFloat(c.attrOr("x")), asFloat(c.attrOr("y")), asFloat(c.attrOr("z")));
        } else if (c.tag == "vector2") {
            std::string val = c.attrOr("value");
            p.v[nm] = Vec3(asFloat(c.attrOr("x")), asFloat(c.attrOr("y")), 0.0f);
            if (!val.empty()) {
                std::stringstream ss(val);
                float x, y;
                char cm;
                if (!((ss >> x >> cm >> y) && cm == ',')) {
                    std::stringstream s2(val);
                    s2 >> x >> y;
                }
                p.v[nm] = Vec3(x, y, 0.0f);
            }
        } else if (c.tag == "transform") {
            Xform t;
            for (const XNode& st : c.kids) {
                Xform step;
                if (st.tag == "translate")
                    step = xformTranslate(asFloat(st.attrOr("x")), asFloat(st.attrOr("y")), asFloat(st.attrOr("z")));
                else if (st.tag == "scale") {
                    float v = asFloat(st.attrOr("value"));
                    if (v != 0.0f) step = xformScale(v, v, v);
                    else step = xformScale(asFloat(st.attrOr("x"), 1.f), asFloat(st.attrOr("y"), 1.f),
                                           asFloat(st.attrOr("z"), 1.f));
                } else if (st.tag == "rotate") {
                    float a = asFloat(st.attrOr("angle"));
                    Vec3 ax = st.attr.count("axis") ? parseVec3(st.attrOr("axis"))
                                                    : Vec3(asFloat(st.attrOr("x")), asFloat(st.attrOr("y")),
                                                           asFloat(st.attrOr("z")));
                    step = xformRotate(a, ax.x, ax.y, ax.z);
                } else if (st.tag == "lookat") {
                    step = xformLookAt(parseVec3(st.attrOr("origin")), parseVec3(st.attrOr("target")),
                                       parseVec3(st.attrOr("up")));
                }
                t = t.then(step);
            }
            p.t[nm] = t;
        } else {
            children.push_back(&c);
        }
    }
    return p;
}

// An object of the scene description as the reference's parser builds it
// (parser.cpp:41-235): its class tag ("scene", "primitive", "shape", ...),
// its registered type name, its name attribute, its properties and its child
// objects in document order.  A child shared through <ref id> (or given to
// two parents) is the same node under each parent.
struct ONode {
    std::string tag, type, name;
    Props p;
    std::vector<const ONode*> kids;
};
struct OTree {
    std::deque<ONode> nodes;  // stable addresses
    ONode& add() {
        nodes.emplace_back();
        return nodes.back();
    }
};

// XML element -> object node, registering its id before its children are
// read, as parser.cpp:188-211 does
const ONode* fromXml(const XNode& n, OTree& t, std::map<std::string, const ONode*>& ids) {
    ONode& o = t.add();
    o.tag = n.tag;
    o.type = n.attrOr("type");
    o.name = n.attrOr("name");
    std::vector<const XNode*> kids;
    o.p = readProps(n, kids);
    const std::string id = n.attrOr("id");
    if (!id.empty()) {
        if (ids.count(id)) throw std::runtime_error("Parsing error: id <" + id + "> already exists.");
        ids[id] = &o;
    }
    for (const XNode* k : kids) {
        if (k->tag == "ref") {
            const std::string rid = k->attrOr("id");
            auto it = ids.find(rid);
            if (it == ids.end()) throw std::runtime_error("Parsing error: no object referencing " + rid + " found.");
            o.kids.push_back(it->second);
        } else {
            o.kids.push_back(fromXml(*k, t, ids));
        }
    }
    return &o;
}

// Minimal OBJ reader (v / vn / vt / f with polygon fans), de-indexed like
// Mesh::loadOBJ (mesh.cpp:148-229)
MeshData loadObj(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Failed to load OBJ: " + path);
    std::vector<Vec3> v, vn;
    std::vector<std::pair<float, float>> vt;
    struct Idx { int v, t, n; };
    std::vector<Idx> corners;
    std::string line;
    auto fix = [](int k, size_t n) { return k < 0 ? (int)n + k : k - 1; };
    while (std::getline(in, line)) {
        std::stringstream ss(line);
        std::string tag;
        ss >> tag;
        if (tag == "v") { float x, y, z; ss >> x >> y >> z; v.emplace_back(x, y, z); }
        else if (tag == "vn") { float x, y, z; ss >> x >> y >> z; vn.emplace_back(x, y, z); }
        else if (tag == "vt") { float x, y; ss >> x >> y; vt.emplace_back(x, y); }
        else if (tag == "f") {
            std::vector<Idx> poly;
            std::string tok;
            while (ss >> tok) {
                Idx id{-1, -1, -1};
                int slot = 0;
                std::string cur;
                for (size_t k = 0; k <= tok.size(); ++k) {
                    if (k == tok.size() || tok[k] == '/') {
                        if (!cur.empty()) {
                            int val = std::stoi(cur);
                            if (slot == 0) id.v = fix(val, v.size());
                            else if (slot == 1) id.t = fix(val, vt.size());
                            else id.n = fix(val, vn.size());
                        }
                        cur.clear();
                        ++slot;
                    } else cur += tok[k];
                }
                poly.push_back(id);
            }
            for (size_t k = 1; k + 1 < poly.size(); ++k) {
                corners.push_back(poly[0]);
                corners.push_back(poly[k]);
                corners.push_back(poly[k + 1]);
            }
        }
    }
    if (corners.empty()) throw std::runtime_error("Mesh: no geometry loaded!");
    MeshData m;
    bool hasN = !vn.empty(), hasT = !vt.empty();
    for (const Idx& c : corners) {
        const Vec3& p = v.at(c.v);
        m.positions.insert(m.positions.end(), {p.x, p.y, p.z});
        m.vertexIds.push_back((uint32_t)c.v);
        if (hasN) {
            Vec3 n = c.n >= 0 ? vn.at(c.n) : Vec3(0, 0, 0);
            m.normals.insert(m.normals.end(), {n.x, n.y, n.z});
        }
        if (hasT) {
            auto t = c.t >= 0 ? vt.at(c.t) : std::make_pair(0.0f, 0.0f);
            m.uvs.insert(m.uvs.end(), {t.first, t.second});
        }
    }
    return m;
}

struct Loader {
    HostScene& sc;
    std::string baseDir;
    std::map<const ONode*, int> meshIds;  // shapes shared by several primitives
    std::map<const ONode*, float> sphereRadius;

    int readMedium(const ONode& n) {
        const Props& p = n.p;
        MediumData md;
        const std::string& type = n.type;
        if (type != "homogeneous") throw std::runtime_error("unsupported medium type: " + type);
        md.desc.type = GND_MEDIUM_HOMOGENEOUS;
        Vec3 sa = p.getV("sigma_a", Vec3(0.1f, 0.1f, 0.1f)), ss = p.getV("sigma_s", Vec3(0.1f, 0.1f, 0.1f));
        for (int c = 0; c < 3; ++c) { md.desc.sigma_a[c] = sa[c]; md.desc.sigma_s[c] = ss[c]; }
        md.desc.g = p.getF("g", 0.0f);
        Mat4 I = Mat4::identity();
        std::memcpy(md.desc.to_local, I.m, sizeof(I.m));
        std::memcpy(md.desc.to_local_inv, I.m, sizeof(I.m));
        sc.media.push_back(std::move(md));
        return (int)sc.media.size() - 1;
    }

    int readMaterial(const ONode& n) {
        const Props& p = n.p;
        const std::string& type = n.type;
        int inside = -1, outside = -1;
        for (const ONode* k : n.kids) {
            if (k->tag == "medium") {
                int id = readMedium(*k);
                if (k->name == "outside") outside = id;
                else inside = id;
            }
        }
        gnd_material_desc m{};
        if (type == "matte") {
            m = materialMatte(0, 0, 0);
            if (p.hasV("albedo")) { Vec3 a = p.getV("albedo", Vec3()); m.albedo[0] = a.x; m.albedo[1] = a.y; m.albedo[2] = a.z; }
            else m.albedo_is_checker = 1;
        } else if (type == "conductor") {
            float R[3], E[3], K[3];
            Vec3 r = p.getV("R", Vec3(1, 1, 1)), e = p.getV("eta", Vec3(1, 1, 1)), k = p.getV("k", Vec3(1, 1, 1));
            for (int c = 0; c < 3; ++c) { R[c] = r[c]; E[c] = e[c]; K[c] = k[c]; }
            m = materialConductor(R, E, K, p.getF("roughness", 0.0f));
        } else if (type == "glass") {
            m = materialGlass(p.getF("eta", 1.5f), p.getF("roughness", 0.0f), inside, outside);
            Vec3 r = p.getV("R", Vec3(1, 1, 1)), t = p.getV("T", Vec3(1, 1, 1));
            for (int c = 0; c < 3; ++c) { m.R[c] = r[c]; m.T[c] = t[c]; }
        } else if (type == "plastic") {
            float Kd[3], Ks[3];
            Vec3 kd = p.getV("Kd", Vec3(0.25f, 0.25f, 0.25f)), ks = p.getV("Ks", Vec3(0.25f, 0.25f, 0.25f));
            for (int c = 0; c < 3; ++c) { Kd[c] = kd[c]; Ks[c] = ks[c]; }
            m = materialPlastic(Kd, Ks, p.getF("eta", 1.5f), p.getF("roughness", 0.1f));
        } else if (type == "index") {
            m = materialIndex(inside, outside);
        } else {
            throw std::runtime_error("unsupported material type: " + type);
        }
        return sc.addMaterial(m);
    }

    void readPrimitive(const ONode& n) {
        const Props& p = n.p;
        Xform x = p.t.count("toWorld") ? p.t.at("toWorld") : Xform();

        int mesh = -1, material = -1;
        bool sphere = false, emitter = false;
        float radius = 1.0f, Le[3] = {0, 0, 0};
        for (const ONode* k : n.kids) {
            if (k->tag == "shape") {
                auto mi = meshIds.find(k);
                if (mi != meshIds.end()) {
                    mesh = mi->second;
                    continue;
                }
                auto si = sphereRadius.find(k);
                if (si != sphereRadius.end()) {
                    sphere = true;
                    radius = si->second;
                    continue;
                }
                const Props& sp = k->p;
                const std::string& type = k->type;
                if (type == "mesh") {
                    std::string f = sp.getS("filename", "");
                    if (f.empty()) throw std::runtime_error("Mesh: filename missing!");
                    std::string path = (!f.empty() && f[0] == '/') ? f : baseDir + f;
                    mesh = sc.addMesh(loadObj(path));
                    meshIds[k] = mesh;
                } else if (type == "sphere") {
                    sphere = true;
                    radius = sp.getF("radius", 1.0f);
                    sphereRadius[k] = radius;
                } else {
                    throw std::runtime_error("unsupported shape type: " + type);
                }
            } else if (k->tag == "material") {
                material = readMaterial(*k);
            } else if (k->tag == "emitter") {
                const Props& ep = k->p;
                if (k->type != "area") throw std::runtime_error("only area emitters are supported");
                emitter = true;
                Vec3 r = ep.getV("radiance", Vec3(0, 0, 0));
                Le[0] = r.x; Le[1] = r.y; Le[2] = r.z;
            }
        }
        if (mesh < 0 && !sphere) throw std::runtime_error("Primitive: defined without a Shape!");
        if (material < 0) {
            gnd_material_desc m = materialMatte(0, 0, 0);
            m.albedo_is_checker = 1;  // missing texture matte fallback (primitive.cpp:39-49)
            material = sc.addMaterial(m);
        }
        const int id = sphere ? sc.addSphere(radius, x, material, emitter ? Le : nullptr)
                              : sc.addPrim(mesh, x, material, emitter ? Le : nullptr);
        if (p.t.count("toWorldEnd")) sc.setMotion(id, p.t.at("toWorldEnd"));  // primitive.cpp:6-11
    }

    // A scene-level emitter (Scene::addChild, scene.cpp:33-40): point,
    // spot, directional or an environment map read from an EXR file.  Its
    // position in the emitter list is its place in the document.
    void readEmitter(const ONode& n) {
        const Props& p = n.p;
        const std::string& type = n.type;
        int t;
        if (type == "point") t = GND_LIGHT_POINT;
        else if (type == "spot") t = GND_LIGHT_SPOT;
        else if (type == "directional") t = GND_LIGHT_DIRECTIONAL;
        else if (type == "environment") t = GND_LIGHT_ENVIRONMENT;
        else if (type == "area") throw std::runtime_error("an area emitter belongs inside a <primitive>");
        else throw std::runtime_error("unsupported emitter type: " + type);
        LightData l = lightDefaults(t);
        gnd_light_desc& d = l.desc;
        auto put = [](float* dst, Vec3 v) { dst[0] = v.x; dst[1] = v.y; dst[2] = v.z; };
        if (t == GND_LIGHT_POINT || t == GND_LIGHT_SPOT) {
            if (!p.hasV("position")) throw std::runtime_error("Property 'position' is missing");
            put(d.position, p.getV("position", Vec3()));
            put(d.power, p.getV("power", Vec3(d.power[0], d.power[1], d.power[2])));
        }
        if (t == GND_LIGHT_SPOT || t == GND_LIGHT_DIRECTIONAL)
            put(d.direction, p.getV("direction", Vec3(d.direction[0], d.direction[1], d.direction[2])));
        if (t == GND_LIGHT_SPOT) {
            d.spot_size = p.getF("spot_size", d.spot_size);
            d.blend = p.getF("blend", d.blend);
        }
        if (t == GND_LIGHT_DIRECTIONAL) put(d.irradiance, p.getV("irradiance", Vec3(d.irradiance[0], d.irradiance[1], d.irradiance[2])));
        if (t == GND_LIGHT_ENVIRONMENT) {
            const std::string f = p.getS("filename", "");
            const std::string path = (!f.empty() && f[0] == '/') ? f : baseDir + f;
            std::ifstream probe(path, std::ios::binary);
            if (f.empty() || !probe) throw std::runtime_error("EnvMap: Invalid environment map path.");
            const size_t dot = path.find_last_of('.');
            if (dot == std::string::npos || path.substr(dot) != ".exr")
                throw std::runtime_error("EnvMap: Invalid environment map extension.");
            readEXR(path, d.env_width, d.env_height, l.env);
            d.strength = p.getF("strength", 1.0f);
        }
        sc.addLight(std::move(l));
    }

    void readCamera(const ONode& n) {
        const Props& p = n.p;
        const std::string& type = n.type;
        Xform x = p.t.count("toWorld") ? p.t.at("toWorld") : Xform();
        sc.setCamera(type == "thinlens" ? GND_CAMERA_THINLENS : GND_CAMERA_PERSPECTIVE, p.getI("width", 1280),
                     p.getI("height", 720), p.getF("fov", 45.0f), x);
        gnd_camera_desc& c = sc.camera;
        c.lens_radius = p.getF("lens_radius", 0.0f);
        c.focal_distance = p.getF("focal_distance", 10.0f);
        c.k1 = p.getF("k1", 0.0f);
        c.ca_axial = p.getF("ca_axial", 0.0f);
        c.ca_lateral = p.getF("ca_lateral", 0.0f);
        c.vignetting = p.getB("vignetting", false) ? 1 : 0;
        c.blades = p.getI("blades", 0);
        if (p.t.count("toWorldEnd")) {  // camera.h:22-25
            c.has_motion = 1;
            std::memcpy(c.to_world_end, p.t.at("toWorldEnd").m.m, sizeof(c.to_world_end));
        }
        if (type == "thinlens") {  // thinLens.cpp:29-31
            c.rolling_shutter = p.getB("rolling_shutter", false) ? 1 : 0;
            c.scan_duration = p.getF("scan_duration", 0.9f);
            c.exposure_duration = p.getF("exposure_duration", 0.1f);
        }
        for (const ONode* k : n.kids) {
            if (k->tag == "filter") {
                const Props& fp = k->p;
                if (k->type == "box") c.filter = GND_FILTER_BOX;
                else {
                    c.filter = GND_FILTER_GAUSSIAN;
                    Vec3 r = fp.getV("radius", Vec3(2.0f, 2.0f, 0.0f));
                    c.filter_radius[0] = r.x;
                    c.filter_radius[1] = r.y;
                    c.filter_alpha = fp.getF("alpha", 2.0f);
                }
            }
        }
    }

    void readScene(const ONode& root) {
        if (root.tag != "scene") throw std::runtime_error("root element must be <scene>");
        bool haveCam = false, haveInt = false, haveSampler = false;
        for (const ONode* k : root.kids) {
            if (k->tag == "camera") {
                if (haveCam) throw std::runtime_error("Scene: cannot add multiple cameras!");
                haveCam = true;
                readCamera(*k);
            } else if (k->tag == "sampler") {
                if (haveSampler) throw std::runtime_error("Scene: cannot add multiple samplers!");
                haveSampler = true;
                sc.spp = (uint32_t)k->p.getI("spp", 1);
            } else if (k->tag == "integrator") {
                if (haveInt) throw std::runtime_error("Scene: cannot add multiple integrators!");
                haveInt = true;
                const Props& ip = k->p;
                const std::string& t = k->type;
                if (t == "path" || t == "path_hip") sc.integrator = GND_INTEGRATOR_PATH;
                else if (t == "volpath" || t == "volpath_hip") sc.integrator = GND_INTEGRATOR_VOLPATH;
                else if (t == "direct" || t == "direct_hip") sc.integrator = GND_INTEGRATOR_DIRECT;
                else if (t == "ao" || t == "ao_hip") sc.integrator = GND_INTEGRATOR_AO;
                else throw std::runtime_error("unsupported integrator: " + t);
                // ao.cpp:9; a length <= 0 (no occluder can be nearer than
                // tMin) is kept as a tiny positive length with the same result
                const float len = ip.getF("length", 5.0f);
                sc.aoLength = len > 0.0f ? len : 1e-30f;
                sc.maxDepth = ip.getI("maxDepth", 1000);
            } else if (k->tag == "primitive") {
                readPrimitive(*k);
            } else if (k->tag == "emitter") {
                readEmitter(*k);
            } else {
                throw std::runtime_error("Scene: cannot add specified child!");
            }
        }
        if (!haveCam) throw std::runtime_error("Scene: no camera specified!");
        if (!haveInt) throw std::runtime_error("Scene: no integrator specified!");
        if (!haveSampler) throw std::runtime_error("Scene: no sampler specified!");
    }
};

}  // namespace

std::unique_ptr<HostScene> loadXmlScene(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("XML parsing failed: cannot open " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    std::string text = ss.str();
    XmlReader r(text);
    XNode root = r.parseDocument();
    OTree tree;
    std::map<std::string, const ONode*> ids;
    const ONode* top = fromXml(root, tree, ids);
    auto sc = std::make_unique<HostScene>();
    Loader L{*sc};
    size_t slash = path.find_last_of('/');
    L.baseDir = slash == std::string::npos ? "" : path.substr(0, slash + 1);
    L.readScene(*top);
    return sc;
}

struct ObjectGraph::Impl {
    OTree tree;
    std::vector<ONode*> nodes;
    std::string baseDir;
};

ObjectGraph::ObjectGraph(const std::string& baseDir) : impl_(std::make_unique<Impl>()) {
    impl_->baseDir = baseDir;
    if (!baseDir.empty() && baseDir.back() != '/') impl_->baseDir += '/';
}

ObjectGraph::~ObjectGraph() = default;

int ObjectGraph::add(const std::string& tag, const std::string& type, const std::string& name, const gnd_prop* props,
                     int nprops) {
    // the reference's class tags (GndObject::classTypeToString, object.h)
    static const char* const kTags[] = {"scene",   "primitive", "shape",    "camera",  "integrator", "sampler",
                                        "texture", "material",  "emitter", "filter", "medium"};
    if (std::find_if(std::begin(kTags), std::end(kTags), [&](const char* t) { return tag == t; }) == std::end(kTags))
        throw std::runtime_error("unknown object class: " + tag);
    if (nprops < 0 || (nprops > 0 && !props)) throw std::runtime_error("invalid property list");
    ONode& o = impl_->tree.add();
    o.tag = tag;
    o.type = type;
    o.name = name;
    for (int i = 0; i < nprops; ++i) {
        const gnd_prop& q = props[i];
        if (!q.name) throw std::runtime_error("a property of " + tag + " has no name");
        const std::string nm = q.name;
        switch (q.kind) {
            case GND_PROP_BOOLEAN: o.p.b[nm] = q.ivalue != 0; break;
            case GND_PROP_INTEGER: o.p.i[nm] = q.ivalue; break;
            case GND_PROP_FLOAT: o.p.f[nm] = q.v[0]; break;
            case GND_PROP_STRING: o.p.s[nm] = q.svalue ? q.svalue : ""; break;
            case GND_PROP_POINT:
            case GND_PROP_VECTOR:
            case GND_PROP_COLOR: o.p.v[nm] = Vec3(q.v[0], q.v[1], q.v[2]); break;
            case GND_PROP_VECTOR2: o.p.v[nm] = Vec3(q.v[0], q.v[1], 0.0f); break;
            case GND_PROP_TRANSFORM: {
                if (!q.matrix) throw std::runtime_error("transform " + nm + " has no matrix");
                Xform x;
                std::memcpy(x.m.m, q.matrix, sizeof(x.m.m));
                std::memcpy(x.inv.m, q.matrix + 16, sizeof(x.inv.m));
                o.p.t[nm] = x;
                break;
            }
            default: throw std::runtime_error("property " + nm + " has an unknown kind");
        }
    }
    impl_->nodes.push_back(&o);
    return (int)impl_->nodes.size() - 1;
}

void ObjectGraph::addChild(int parent, int child) {
    const int n = (int)impl_->nodes.size();
    if (parent < 0 || parent >= n || child < 0 || child >= n || parent == child)
        throw std::runtime_error("invalid object ids");
    impl_->nodes[parent]->kids.push_back(impl_->nodes[child]);
}

std::unique_ptr<HostScene> ObjectGraph::build(int root) const {
    if (root < 0 || root >= (int)impl_->nodes.size()) throw std::runtime_error("invalid root object");
    auto sc = std::make_unique<HostScene>();
    Loader L{*sc};
    L.baseDir = impl_->baseDir;
    L.readScene(*impl_->nodes[root]);
    return sc;
}

}  // namespace gnd
