// gnd_ref_render.cpp — oracle: shading, lights, media, camera, integrators and
// the progressive render driver, restated from the reference.  Test
// infrastructure only (see gnd_ref.h).
#include "gnd_ref.h"

#include <cmath>
#include <thread>
#include <stdexcept>

namespace gref {

static inline float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ------------------------------------------------------------ interaction
static Ray toObjectRay(const Prim& p, const Ray& rw) {
    M4 W, Wi;
    primAt(p, rw.time, W, Wi);
    Ray r;
    r.o = xfPoint(Wi, rw.o);
    r.d = xfVector(Wi, rw.d);
    r.tMax = rw.tMax;
    return r;
}

static V3 dpduFallback(V3 n) {  // mesh.cpp:370-373
    if (std::abs(n.x) > std::abs(n.y)) return V3(-n.z, 0.0f, n.x) / std::sqrt(n.x * n.x + n.z * n.z);
    return V3(0.0f, n.z, -n.y) / std::sqrt(n.y * n.y + n.z * n.z);
}

// Mesh::fillInteraction (mesh.cpp:328-390), Sphere::fillInteraction
// (sphere.cpp:41-55), Primitive::fillInteraction (primitive.cpp:61-72)
void fillInteraction(const Scene& s, const Ray& rw, const Hit& h, Isect& it) {
    const Prim& pr = s.prims[h.prim];
    it.t = h.t;
    it.prim = h.prim;
    it.primIndex = h.primIndex;
    V3 P, N, dpdu;
    P2 uv;
    if (pr.shapeType == GND_SHAPE_MESH) {
        const Mesh& m = s.meshes[pr.mesh];
        uint32_t i0 = 3 * h.primIndex, i1 = i0 + 1, i2 = i0 + 2;
        V3 p0 = m.positions[i0], p1 = m.positions[i1], p2 = m.positions[i2];
        float u = h.u, v = h.v, w = 1.0f - u - v;
        P = p0 * w + p1 * u + p2 * v;
        N = Normalize(m.normals[i0] * w + m.normals[i1] * u + m.normals[i2] * v);
        uv.x = u;
        uv.y = v;
        if (!m.uvs.empty()) {
            float u0 = m.uvs[2 * i0], v0 = m.uvs[2 * i0 + 1], u1 = m.uvs[2 * i1], v1 = m.uvs[2 * i1 + 1],
                  u2 = m.uvs[2 * i2], v2 = m.uvs[2 * i2 + 1];
            uv.x = u0 * w + u1 * u + u2 * v;
            uv.y = v0 * w + v1 * u + v2 * v;
            V3 dp1 = p1 - p0, dp2 = p2 - p0;
            float du1 = u1 - u0, dv1 = v1 - v0, du2 = u2 - u0, dv2 = v2 - v0;
            float det = du1 * dv2 - dv1 * du2;
            if (std::abs(det) < 1e-8f) dpdu = dpduFallback(N);
            else {
                float invDet = 1.0f / det;
                dpdu = (dp1 * dv2 - dp2 * dv1) * invDet;
            }
            dpdu = Normalize(dpdu - N * Dot(N, dpdu));
        } else {
            dpdu = Normalize(dpduFallback(N));
        }
    } else {
        Ray rl = toObjectRay(pr, rw);
        P = rl.o + rl.d * h.t;
        N = Normalize(P - V3(0, 0, 0));
        float phi = std::atan2(N.x, N.z);
        if (phi < 0.0f) phi = (float)((double)phi + 2.0f * M_PI);
        uv.x = phi * kInv2Pi;
        float theta = std::acos(clampf(N.y, -1.0f, 1.0f));
        uv.y = theta * kInvPi;
        dpdu = V3((float)((double)(P.z * 2.0f) * M_PI), 0.0f, (float)((double)(-P.x * 2.0f) * M_PI));
    }
    it.uv = uv;
    M4 W, Wi;
    primAt(pr, rw.time, W, Wi);  // interpolate(rWorld.time) (primitive.cpp:61-72)
    it.p = xfPoint(W, P);
    it.n = Normalize(xfNormal(Wi, N));
    it.dpdu = Normalize(xfVector(W, dpdu));
    const gnd_material_desc& md = s.materials[pr.material].d;
    it.hasInterface = false;
    it.mediumInside = it.mediumOutside = -1;
    // Material::applyMediums (material.h:119-126), glass and index only
    if ((md.type == GND_MAT_GLASS || md.type == GND_MAT_INDEX) && (md.medium_inside >= 0 || md.medium_outside >= 0)) {
        it.hasInterface = true;
        it.mediumInside = md.medium_inside;
        it.mediumOutside = md.medium_outside;
    }
}

// ------------------------------------------------------------ warps (src/math/warp.cpp)
static P2 squareToUniformDisk(P2 s) {  // warp.cpp:7-25 (M_PI is double)
    float px = 2.0f * s.x - 1.0f, py = 2.0f * s.y - 1.0f;
    if (std::abs(px) < kEpsilon && std::abs(py) < kEpsilon) return {0, 0};
    float r, theta;
    if (std::abs(px) > std::abs(py)) {
        r = px;
        theta = (float)((M_PI / 4.0f) * (double)(py / px));
    } else {
        r = py;
        theta = (float)((M_PI / 2.0f) - (M_PI / 4.0f) * (double)(px / py));
    }
    return {r * std::cos(theta), r * std::sin(theta)};
}
static V3 squareToCosineHemisphere(P2 s) {  // warp.cpp:68-77
    P2 d = squareToUniformDisk(s);
    float r = d.x * d.x + d.y * d.y;
    return V3(d.x, d.y, std::sqrt(1 - r));
}
static float squareToCosineHemispherePdf(V3 v) {  // warp.cpp:79-84
    if (v.z <= 0.0f || std::abs(1 - LengthSq(v)) > kEpsilon) return 0.0f;
    return kInvPi * v.z;
}
static V3 squareToUniformSphere(P2 s) {  // warp.cpp:34-43 with cosThetaMax = -1
    float theta = std::acos(1 - (1 - (-1.0f)) * s.x);
    float phi = (float)(2.0f * M_PI * (double)s.y);
    return V3(std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta));
}
static P2 squareToUniformPolygon(P2 s, int blades) {  // warp.cpp:86-113
    if (blades < 3) return squareToUniformDisk(s);
    float bf = s.x * blades;
    int bi = std::min((int)bf, blades - 1);
    float u0 = bf - bi, u1 = s.y;
    float dt = 2.0f * kPi / blades;
    float th1 = bi * dt + kPi / 2.0f, th2 = (bi + 1) * dt + kPi / 2.0f;
    float r = std::sqrt(u0);
    float v1x = std::cos(th1), v1y = std::sin(th1), v2x = std::cos(th2), v2y = std::sin(th2);
    return {r * ((1.0f - u1) * v1x + u1 * v2x), r * ((1.0f - u1) * v1y + u1 * v2y)};
}

// ------------------------------------------------------------ frame (math/frame.h)
struct Frame {
    V3 x{1, 0, 0}, y{0, 1, 0}, z{0, 0, 1};
    Frame() = default;
    Frame(V3 a, V3 b, V3 c) : x(a), y(b), z(c) {}
    static Frame fromNormal(V3 n) {  // frame.h:17-25 (Duff et al.)
        Frame f;
        f.z = n;
        float sign = std::copysign(1.0f, n.z);
        const float a = -1.0f / (sign + n.z);
        const float b = n.x * n.y * a;
        f.x = V3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
        f.y = V3(b, sign + n.y * n.y * a, -n.y);
        return f;
    }
    V3 toLocal(V3 v) const { return V3(Dot(v, x), Dot(v, y), Dot(v, z)); }
    V3 toWorld(V3 v) const { return x * v.x + y * v.y + z * v.z; }
};

// ------------------------------------------------------------ fresnel (materials/fresnel.h)
static float FrDielectric(float cosI, float etaExt, float etaInt) {  // fresnel.h:13-40
    if (etaExt == etaInt) return 0.0f;
    cosI = clampf(cosI, -1.f, 1.f);
    if (cosI < 0.f) { std::swap(etaExt, etaInt); cosI = -cosI; }
    float sinI = std::sqrt(std::max(0.f, 1.f - cosI * cosI));
    float sinT = etaExt / etaInt * sinI;
    if (sinT >= 1.f) return 1.f;
    float cosT = std::sqrt(std::max(0.f, 1.f - sinT * sinT));
    float rParl = ((etaInt * cosI) - (etaExt * cosT)) / ((etaInt * cosI) + (etaExt * cosT));
    float rPerp = ((etaExt * cosI) - (etaInt * cosT)) / ((etaExt * cosI) + (etaInt * cosT));
    return (rParl * rParl + rPerp * rPerp) / 2.f;
}
static bool Refract(V3 wi, V3 n, float eta, V3* wt) {  // fresnel.h:47-58
    float cosI = Dot(n, wi);
    float sin2I = std::max(0.f, 1.f - cosI * cosI);
    float sin2T = eta * eta * sin2I;
    if (sin2T >= 1.f) return false;
    float cosT = std::sqrt(1.f - sin2T);
    *wt = -wi * eta + n * (eta * cosI - cosT);
    return true;
}
static float conductorChannel(float cosI, float etaI, float etaT, float k) {  // fresnel.h:111-133
    float cosI2 = cosI * cosI;
    float sinI2 = std::max(0.0f, 1.f - cosI2);
    float eta = etaT / etaI, etak = k / etaI;
    float eta2 = eta * eta, etak2 = etak * etak;
    float t0 = eta2 - etak2 - sinI2;
    float a2b2 = std::sqrt(t0 * t0 + 4.f * eta2 * etak2);
    float t1 = a2b2 + cosI2;
    float a = std::sqrt(std::max(0.0f, 0.5f * (a2b2 + t0)));
    float t2 = 2.f * std::abs(cosI) * a;
    float Rs = (t1 - t2) / (t1 + t2);
    float t3 = cosI2 * a2b2 + sinI2 * sinI2;
    float t4 = t2 * sinI2;
    float Rp = Rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (Rp + Rs);
}

// ------------------------------------------------------------ GGX (src/math/microfacet.cpp)
struct GGX {
    float ax, ay;
    explicit GGX(float alpha) : ax(std::max(1e-4f, alpha)), ay(std::max(1e-4f, alpha)) {}
    static float roughnessToAlpha(float r) { r = std::max(r, 1e-3f); return r * r; }
    float D(V3 wh) const {
        if (std::abs(wh.z) < kEpsilon) return 0.0f;
        float d = (wh.x * wh.x) / (ax * ax) + (wh.y * wh.y) / (ay * ay) + (wh.z * wh.z);
        return 1.0f / (kPi * ax * ay * d * d);
    }
    float G1(V3 w) const {
        if (std::abs(w.z) < kEpsilon) return 0.0f;
        float a2tan2 = (w.x * w.x * ax * ax + w.y * w.y * ay * ay) / (w.z * w.z);
        float lambda = 0.5f * (-1.0f + std::sqrt(1.0f + a2tan2));
        return 1.0f / (1.0f + lambda);
    }
    float G(V3 wo, V3 wi) const { return G1(wo) * G1(wi); }
    V3 sample_wh(V3 wo, P2 s) const {  // microfacet.cpp:30-56 (Heitz 2018)
        bool flip = wo.z < 0.0f;
        V3 woh = flip ? -wo : wo;
        V3 Vh = Normalize(V3(ax * woh.x, ay * woh.y, woh.z));
        float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
        V3 T1 = lensq > 0.0f ? V3(-Vh.y, Vh.x, 0.0f) / std::sqrt(lensq) : V3(1.0f, 0.0f, 0.0f);
        V3 T2 = Cross(Vh, T1);
        float r = std::sqrt(s.x);
        float phi = 2.0f * kPi * s.y;
        float t1 = r * std::cos(phi);
        float t2 = r * std::sin(phi);
        float sv = 0.5f * (1.0f + Vh.z);
        t2 = (1.0f - sv) * std::sqrt(std::max(0.0f, 1.0f - t1 * t1)) + sv * t2;
        V3 Nh = t1 * T1 + t2 * T2 + std::sqrt(std::max(0.0f, 1.0f - t1 * t1 - t2 * t2)) * Vh;
        V3 wh = Normalize(V3(ax * Nh.x, ay * Nh.y, std::max(0.0f, Nh.z)));
        if (flip) wh = -wh;
        return wh;
    }
    float pdf(V3 wo, V3 wh) const {
        if (wo.z * wh.z < 0.0f) return 0.0f;
        float cosO = std::abs(wo.z);
        if (cosO == 0.0f) return 0.0f;
        float dotOH = std::abs(Dot(wo, wh));
        return D(wh) * G1(wo) * dotOH / cosO;
    }
};

// ------------------------------------------------------------ BxDFs (src/materials/reflection.cpp)
enum { BSDF_REFLECTION = 1, BSDF_TRANSMISSION = 2, BSDF_DIFFUSE = 4, BSDF_GLOSSY = 8, BSDF_SPECULAR = 16 };
enum LobeKind { LAMBERT, SPEC_REFL, FRESNEL_SPEC, MF_REFL, MF_FRESNEL, SMOOTH_PLASTIC, ROUGH_PLASTIC };

struct Lobe {
    LobeKind kind;
    C3 R, T, Kd, Ks, eta3, k3;
    float etaExt = 1.0f, etaInt = 1.0f, eta = 1.5f;
    GGX ggx{1.0f};
    C3 conductorF(float cosI) const {
        return C3(conductorChannel(cosI, 1.0f, eta3.r, k3.r), conductorChannel(cosI, 1.0f, eta3.g, k3.g),
                  conductorChannel(cosI, 1.0f, eta3.b, k3.b));
    }
    int type() const {
        switch (kind) {
            case LAMBERT: return BSDF_REFLECTION | BSDF_DIFFUSE;
            case SPEC_REFL: return BSDF_REFLECTION | BSDF_SPECULAR;
            case FRESNEL_SPEC: return BSDF_REFLECTION | BSDF_TRANSMISSION | BSDF_SPECULAR;
            case MF_REFL: return BSDF_REFLECTION | BSDF_GLOSSY;
            case MF_FRESNEL: return BSDF_REFLECTION | BSDF_TRANSMISSION | BSDF_GLOSSY;
            case SMOOTH_PLASTIC: return BSDF_REFLECTION | BSDF_DIFFUSE | BSDF_SPECULAR;
            default: return BSDF_REFLECTION | BSDF_DIFFUSE | BSDF_GLOSSY;
        }
    }
    C3 f(V3 wo, V3 wi) const {
        switch (kind) {
            case LAMBERT: return R * kInvPi;  // reflection.cpp:8-10
            case SPEC_REFL: case FRESNEL_SPEC: return C3(0.0f);
            case MF_REFL: {  // reflection.cpp:103-121
                if (wo.z * wi.z <= 0.0f) return C3(0.0f);
                float cO = std::abs(wo.z), cI = std::abs(wi.z);
                if (cO == 0.0f || cI == 0.0f) return C3(0.0f);
                V3 wh = wo + wi;
                if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) return C3(0.0f);
                wh = Normalize(wh);
                if (wh.z < 0.0f) wh = -wh;
                if (Dot(wo, wh) * wo.z < 0.0f || Dot(wi, wh) * wi.z < 0.0f) return C3(0.0f);
                C3 F = conductorF(Dot(wi, wh));
                float D = ggx.D(wh), G = ggx.G(wo, wi);
                return R * D * G * F / (4.0f * cI * cO);
            }
            case MF_FRESNEL: {  // reflection.cpp:151-191
                bool reflect = wo.z * wi.z > 0.0f;
                bool entering = wo.z > 0.0f;
                float etaI = entering ? etaExt : etaInt, etaT = entering ? etaInt : etaExt;
                V3 wh = reflect ? Normalize(wo + wi) : Normalize(wo * etaI + wi * etaT);
                if (wh.z * wo.z < 0.0f) wh = -wh;
                float dO = Dot(wo, wh), dI = Dot(wi, wh);
                if (reflect) { if (dO <= 0.0f || dI <= 0.0f) return C3(0.0f); }
                else { if (dO <= 0.0f || dI >= 0.0f) return C3(0.0f); }
                float F = FrDielectric(dO, etaI, etaT);
                float D = ggx.D(wh), G = ggx.G(wo, wi);
                if (reflect) return R * D * G * F / (4.0f * std::abs(wo.z * wi.z));
                if (dO * dI > 0.0f) return C3(0.0f);
                float denom = (etaI * dO + etaT * dI);
                denom *= denom;
                float factor = std::abs(dO * dI) / std::abs(wo.z * wi.z);
                return T * factor * (etaI * etaI) * D * G * (1.0f - F) / denom;
            }
            case SMOOTH_PLASTIC: {  // reflection.cpp:258-264
                if (wo.z < kEpsilon || wi.z < kEpsilon) return C3(0.0f);
                float Fo = FrDielectric(wo.z, 1.0f, eta), Fi = FrDielectric(wi.z, 1.0f, eta);
                return Kd * (float)M_1_PI * (1.0f - Fo) * (1.0f - Fi);
            }
            default: {  // ROUGH_PLASTIC reflection.cpp:295-313
                if (wo.z < kEpsilon || wi.z < kEpsilon) return C3(0.0f);
                float Fo = FrDielectric(wo.z, 1.0f, eta), Fi = FrDielectric(wi.z, 1.0f, eta);
                C3 diffuse = Kd * (float)M_1_PI * (1.0f - Fo) * (1.0f - Fi);
                V3 wh = Normalize(wo + wi);
                if (wh.z < 0.0f) wh = -wh;
                if (Dot(wo, wh) * wo.z < 0.0f || Dot(wi, wh) * wi.z < 0.0f) return C3(0.0f);
                float Fh = FrDielectric(Dot(wo, wh), 1.0f, eta);
                float D = ggx.D(wh), G = ggx.G(wo, wi);
                C3 spec = Ks * D * G * Fh / (4.0f * wo.z * wi.z);
                return diffuse + spec;
            }
        }
    }
    float pdf(V3 wo, V3 wi) const {
        switch (kind) {
            case LAMBERT: return squareToCosineHemispherePdf(wi);
            case SPEC_REFL: case FRESNEL_SPEC: return 0.0f;
            case MF_REFL: {  // reflection.cpp:140-146
                if (wo.z * wi.z <= 0.0f) return 0.0f;
                V3 wh = Normalize(wo + wi);
                if (wh.z < 0.0f) wh = -wh;
                return ggx.pdf(wo, wh) / (4.0f * Dot(wo, wh));
            }
            case MF_FRESNEL: {  // reflection.cpp:227-253
                bool reflect = wo.z * wi.z > 0.0f;
                bool entering = wo.z > 0.0f;
                float etaI = entering ? etaExt : etaInt, etaT = entering ? etaInt : etaExt;
                V3 wh = reflect ? Normalize(wo + wi) : Normalize(wo * etaI + wi * etaT);
                if (wh.z * wo.z < 0.0f) wh = -wh;
                float dO = Dot(wo, wh), dI = Dot(wi, wh);
                float F = FrDielectric(dO, etaI, etaT);
                float pwh = ggx.pdf(wo, wh);
                if (reflect) return F * pwh / (4.0f * std::abs(dO));
                if (dO * dI > 0.0f) return 0.0f;
                float denom = (etaI * dO + etaT * dI);
                float dwh = (etaT * etaT * std::abs(dI)) / (denom * denom);
                return (1.0f - F) * pwh * dwh;
            }
            case SMOOTH_PLASTIC: {
                if (wo.z < kEpsilon || wi.z < kEpsilon) return 0.0f;
                float Fo = FrDielectric(wo.z, 1.0f, eta);
                return (1.0f - Fo) * squareToCosineHemispherePdf(wi);
            }
            default: {  // reflection.cpp:337-348
                if (wo.z < kEpsilon || wi.z < kEpsilon) return 0.0f;
                V3 wh = Normalize(wo + wi);
                if (wh.z < 0.0f) wh = -wh;
                float ps = FrDielectric(wo.z, 1.0f, eta);
                float pdfSpec = ggx.pdf(wo, wh) / (4.0f * Dot(wo, wh));
                float pdfDiff = squareToCosineHemispherePdf(wi);
                return ps * pdfSpec + (1.0f - ps) * pdfDiff;
            }
        }
    }
    // returns f; sets wi, pdf, type (type untouched when the reference leaves it unset)
    C3 sample(V3 wo, V3& wi, P2 s, float uc, float& pdf, int& stype) const {
        switch (kind) {
            case LAMBERT:  // reflection.cpp:12-27
                if (wo.z <= 0.0f) { pdf = 0.0f; return C3(0.0f); }
                wi = squareToCosineHemisphere(s);
                pdf = squareToCosineHemispherePdf(wi);
                stype = type();
                return R * kInvPi;
            case SPEC_REFL:  // reflection.cpp:40-48
                wi = V3(-wo.x, -wo.y, wo.z);
                pdf = 1.0f;
                stype = type();
                return R * conductorF(wo.z);
            case FRESNEL_SPEC: {  // reflection.cpp:61-94
                float F = FrDielectric(wo.z, etaExt, etaInt);
                if (s.x < F) {
                    wi = V3(-wo.x, -wo.y, wo.z);
                    stype = BSDF_SPECULAR | BSDF_REFLECTION;
                    pdf = F;
                    return R * F;
                }
                bool entering = wo.z > 0;
                float etaI = entering ? etaExt : etaInt, etaT = entering ? etaInt : etaExt;
                V3 n(0, 0, 1);
                if (!entering) n = V3(0, 0, -1);
                V3 wr;
                if (!Refract(wo, n, etaI / etaT, &wr)) return C3(0.f);
                wi = wr;
                stype = BSDF_SPECULAR | BSDF_TRANSMISSION;
                pdf = 1.f - F;
                float er = etaI / etaT;
                return T * er * er * (1.0f - F);
            }
            case MF_REFL: {  // reflection.cpp:123-138
                if (std::abs(wo.z) < kEpsilon) { pdf = 0.0f; return C3(0.0f); }
                V3 wh = ggx.sample_wh(wo, s);
                float dO = Dot(wo, wh);
                if (dO <= 0.0f) { pdf = 0.0f; return C3(0.0f); }
                wi = -wo + wh * 2.0f * dO;
                if (wo.z * wi.z <= 0.0f) { pdf = 0.0f; return C3(0.0f); }
                pdf = this->pdf(wo, wi);
                stype = type();
                return f(wo, wi);
            }
            case MF_FRESNEL: {  // reflection.cpp:193-225
                if (std::abs(wo.z) < kEpsilon) return C3(0.0f);
                bool entering = wo.z > 0.0f;
                float etaI = entering ? etaExt : etaInt, etaT = entering ? etaInt : etaExt;
                V3 wh = ggx.sample_wh(wo, s);
                float dO = Dot(wo, wh);
                if (dO <= 0.0f) { pdf = 0.0f; return C3(0.0f); }
                float F = FrDielectric(dO, etaI, etaT);
                if (uc < F) {
                    wi = -wo + wh * 2.0f * dO;
                    if (wo.z * wi.z <= 0.0f) { pdf = 0.0f; return C3(0.0f); }
                    stype = BSDF_REFLECTION | BSDF_GLOSSY;
                } else {
                    if (!Refract(wo, wh, etaI / etaT, &wi)) { pdf = 0.0f; return C3(0.0f); }
                    if (wo.z * wi.z > 0.0f) { pdf = 0.0f; return C3(0.0f); }
                    stype = BSDF_TRANSMISSION | BSDF_GLOSSY;
                }
                pdf = this->pdf(wo, wi);
                return f(wo, wi);
            }
            case SMOOTH_PLASTIC: {  // reflection.cpp:266-284
                if (wo.z < kEpsilon) { pdf = 0.0f; return C3(0.0f); }
                float ps = FrDielectric(wo.z, 1.0f, eta);
                if (uc < ps) {
                    wi = V3(-wo.x, -wo.y, wo.z);
                    pdf = ps;
                    stype = BSDF_REFLECTION | BSDF_SPECULAR;
                    return Ks;
                }
                wi = squareToCosineHemisphere(s);
                if (wi.z < kEpsilon) { pdf = 0.0f; return C3(0.0f); }
                pdf = (1.0f - ps) * squareToCosineHemispherePdf(wi);
                stype = BSDF_REFLECTION | BSDF_DIFFUSE;
                return f(wo, wi);
            }
            default: {  // ROUGH_PLASTIC reflection.cpp:315-335
                if (wo.z < kEpsilon) { pdf = 0.0f; return C3(0.0f); }
                float ps = FrDielectric(wo.z, 1.0f, eta);
                if (uc < ps) {
                    V3 wh = ggx.sample_wh(wo, s);
                    float dO = Dot(wo, wh);
                    if (dO <= 0.0f) { pdf = 0.0f; return C3(0.0f); }
                    wi = -wo + wh * 2.0f * dO;
                    if (wi.z < kEpsilon) { pdf = 0.0f; return C3(0.0f); }
                    stype = BSDF_REFLECTION | BSDF_GLOSSY;
                } else {
                    wi = squareToCosineHemisphere(s);
                    if (wi.z < kEpsilon) { pdf = 0.0f; return C3(0.0f); }
                    stype = BSDF_REFLECTION | BSDF_DIFFUSE;
                }
                pdf = this->pdf(wo, wi);
                return f(wo, wi);
            }
        }
    }
};

// BSDF with one lobe (src/materials/bsdf.cpp:6-136); every material on the
// path adds exactly one lobe with weight 1, so selection probability is 1.
struct BSDF {
    Frame frame;
    Lobe lobe;
    bool nonSpecular() const { return (lobe.type() & BSDF_SPECULAR) == 0; }
    C3 f(V3 woW, V3 wiW) const {
        V3 wo = frame.toLocal(woW), wi = frame.toLocal(wiW);
        if (wo.z == 0.0f) return C3(0.0f);
        return C3(0.0f) + lobe.f(wo, wi);
    }
    float pdf(V3 woW, V3 wiW) const {
        V3 wo = frame.toLocal(woW), wi = frame.toLocal(wiW);
        if (wo.z == 0.0f) return 0.0f;
        return 0.0f + 1.0f * lobe.pdf(wo, wi);
    }
    C3 sample(V3 woW, V3* wiW, P2 s, float uc, float* pdf, int* stype) const {
        V3 wo = frame.toLocal(woW);
        if (wo.z == 0.0f) return C3(0.0f);
        V3 wi(0, 0, 0);
        float lobePdf = 0.0f;
        int st = 0;
        C3 fv = lobe.sample(wo, wi, s, uc, lobePdf, st);
        *wiW = frame.toWorld(wi);
        *stype = st;
        if (lobePdf == 0.0f || fv.isBlack()) { *pdf = 0.0f; return C3(0.0f); }
        if (st & BSDF_SPECULAR) {
            *pdf = lobePdf * 1.0f;
            return fv / *pdf;
        }
        float tp = lobePdf * 1.0f;
        *pdf = tp;
        if (tp < kEpsilon) return C3(0.0f);
        return fv * std::abs(wi.z) / tp;
    }
};

// Material::computeScatteringFunctions (matte.cpp:46-53, conductor.cpp:101-120,
// glass.cpp:98-115, plastic.cpp:74-91, index.cpp:10-16).  Returns false for the
// null BSDF.
static bool makeBSDF(const Scene& s, const Isect& it, BSDF& b) {
    const gnd_material_desc& m = s.materials[s.prims[it.prim].material].d;
    if (m.type == GND_MAT_INDEX) return false;
    V3 t = it.dpdu;
    V3 bt = Normalize(Cross(it.n, it.dpdu));
    b.frame = Frame(t, bt, it.n);
    Lobe& L = b.lobe;
    switch (m.type) {
        case GND_MAT_MATTE:
            L.kind = LAMBERT;
            if (m.albedo_is_checker) {  // checkerboard_color default (checkerboard.h), scale 0.1
                float ux = it.uv.x / 0.1f, uy = it.uv.y / 0.1f;
                int si = (int)std::floor(ux), ti = (int)std::floor(uy);
                L.R = ((si + ti) % 2 == 0) ? C3(0.0f) : C3(1.0f, 0.0f, 1.0f);
            } else L.R = C3(m.albedo[0], m.albedo[1], m.albedo[2]);
            break;
        case GND_MAT_CONDUCTOR:
            L.R = C3(m.R[0], m.R[1], m.R[2]);
            L.eta3 = C3(m.eta_rgb[0], m.eta_rgb[1], m.eta_rgb[2]);
            L.k3 = C3(m.k_rgb[0], m.k_rgb[1], m.k_rgb[2]);
            if (m.roughness == 0.0f) L.kind = SPEC_REFL;
            else { L.kind = MF_REFL; L.ggx = GGX(GGX::roughnessToAlpha(m.roughness)); }
            break;
        case GND_MAT_GLASS:
            L.R = C3(m.R[0], m.R[1], m.R[2]);
            L.T = C3(m.T[0], m.T[1], m.T[2]);
            L.etaExt = 1.0f;
            L.etaInt = m.eta;
            if (m.roughness == 0.0f) L.kind = FRESNEL_SPEC;
            else { L.kind = MF_FRESNEL; L.ggx = GGX(GGX::roughnessToAlpha(m.roughness)); }
            break;
        case GND_MAT_PLASTIC:
            L.Kd = C3(m.Kd[0], m.Kd[1], m.Kd[2]);
            L.Ks = C3(m.Ks[0], m.Ks[1], m.Ks[2]);
            L.eta = m.eta;
            if (m.roughness == 0.0f) L.kind = SMOOTH_PLASTIC;
            else { L.kind = ROUGH_PLASTIC; L.ggx = GGX(GGX::roughnessToAlpha(m.roughness)); }
            break;
        default: return false;
    }
    return true;
}

// ------------------------------------------------------------ area light (src/emitters/area.cpp)
struct LightSample {
    V3 p, n;
    float pdf = 0.0f;
    Ray shadow;
    C3 Li;
};
static C3 areaEval(const Prim& pr, V3 n, V3 w) {  // area.cpp:45-49
    if (Dot(n, w) > 0.0f) return pr.radiance;
    return C3(0.0f);
}
// area.cpp:51-88 with Mesh::sampleSurface (mesh.cpp:396-434) / Sphere::sampleSurface (sphere.cpp:64-81)
static C3 areaSample(const Scene& s, const Prim& pr, V3 refP, P2 u, LightSample& ls, bool& pdfSet, float time) {
    M4 W, Wi;
    primAt(pr, time, W, Wi);  // getToWorld(ref.time) (area.cpp:54)
    pdfSet = false;
    V3 lp, ln;
    float areaPdf;
    if (pr.shapeType == GND_SHAPE_MESH) {
        const Mesh& m = s.meshes[pr.mesh];
        if (m.totalArea == 0.0f) areaPdf = 0.0f, lp = V3(0, 0, 0), ln = V3(0, 0, 0);
        else {
            auto it = std::lower_bound(m.cdf.begin(), m.cdf.end(), u.x);
            size_t ti = (size_t)std::distance(m.cdf.begin(), it);
            ti = std::clamp(ti, (size_t)1, m.cdf.size() - 1) - 1;
            float pdfTri = m.cdf[ti + 1] - m.cdf[ti];
            float rx = (u.x - m.cdf[ti]) / pdfTri;
            rx = clampf(rx, 0.0f, 1.0f);
            float sq = std::sqrt(rx);
            float bu = 1.0f - sq, bv = u.y * sq, bw = 1.0f - bu - bv;
            uint32_t i0 = 3 * (uint32_t)ti;
            lp = m.positions[i0] * bw + m.positions[i0 + 1] * bu + m.positions[i0 + 2] * bv;
            ln = Normalize(m.normals[i0] * bw + m.normals[i0 + 1] * bu + m.normals[i0 + 2] * bv);
            areaPdf = 1.0f / m.totalArea;
        }
    } else {
        V3 r = squareToUniformSphere(u);
        lp = r * pr.radius;
        ln = Normalize(lp - V3(0, 0, 0));
        areaPdf = kInv4Pi / (pr.radius * pr.radius);
    }
    ls.p = xfPoint(W, lp);
    ls.n = Normalize(xfNormal(Wi, ln));
    Frame fr = Frame::fromNormal(ln);
    V3 t1 = xfVector(W, fr.x), t2 = xfVector(W, fr.y);
    float areaScale = Length(Cross(t1, t2));
    if (areaScale > 0.0f) areaPdf /= areaScale;
    ls.shadow.o = refP;
    ls.shadow.d = Normalize(ls.p - refP);
    ls.shadow.tMin = kEpsilon;
    ls.shadow.tMax = Length(ls.p - refP) - kEpsilon;
    ls.shadow.time = time;  // area.cpp:75
    float dist = LengthSq(refP - ls.p);
    V3 wi = ls.shadow.d;
    float cosT = Dot(ls.n, -wi);
    if (areaPdf <= kEpsilon || cosT <= kEpsilon) return C3(0.0f);
    float ap = areaPdf * dist / cosT;
    ls.pdf = ap;
    pdfSet = true;
    return areaEval(pr, ls.n, Normalize(refP - ls.p)) / ls.pdf;
}
// area.cpp:90-119
static float areaPdfAt(const Scene& s, const Prim& pr, V3 refP, V3 infoP, V3 infoN, float time) {
    M4 W, Wi;
    primAt(pr, time, W, Wi);  // getToWorld(ref.time) (area.cpp:92)
    V3 ln = xfNormal(W, infoN);  // toLocal(n): transpose(mat3(toWorld.m)) * n
    float areaPdf = pr.shapeType == GND_SHAPE_MESH
                        ? (s.meshes[pr.mesh].totalArea > 0.0f ? 1.0f / s.meshes[pr.mesh].totalArea : 0.0f)
                        : kInv4Pi / (pr.radius * pr.radius);
    Frame fr = Frame::fromNormal(ln);
    V3 t1 = xfVector(W, fr.x), t2 = xfVector(W, fr.y);
    float areaScale = Length(Cross(t1, t2));
    if (areaScale > 0.0f) areaPdf /= areaScale;
    float dist = LengthSq(refP - infoP);
    V3 wi = Normalize(infoP - refP);
    float cosT = Dot(infoN, -wi);
    if (areaPdf <= kEpsilon || cosT <= kEpsilon) return 0.0f;
    return areaPdf * dist / cosT;
}

static inline float powerHeuristic(float f, float g) {  // path.cpp:13-17
    float d = (f * f) + (g * g);
    return d > 0.0f ? (f * f) / d : 0.0f;
}
static int pickEmitter(const Scene& s, float rnd) {  // scene.h:28-34
    size_t n = s.emitters.size();
    size_t idx = std::min((size_t)std::floor((float)n * rnd), n - 1);
    return s.emitters[idx];
}

// ------------------------------------------------------------ lights (src/emitters/*.cpp)
// SphericalCoordinates / SphericalDirection (constants.h:80-100)
static P2 sphericalCoordinates(V3 w) {
    P2 r{std::acos(std::clamp(w.z, -1.0f, 1.0f)), std::atan2(w.y, w.x)};
    if (r.y < 0.0f) r.y += 2.0f * kPi;
    return r;
}
static V3 sphericalDirection(float theta, float phi) {
    float cosTheta = std::cos(theta), sinTheta = std::sin(theta);
    float cosPhi = std::cos(phi), sinPhi = std::sin(phi);
    return V3(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
}
// Film::getPixelBilinear (film.cpp:125-151): u wraps, v clamps
static C3 envBilinear(const Light& L, float u, float v) {
    const int W = L.width, H = L.height;
    float x = u * W - 0.5f;
    float y = v * H - 0.5f;
    int x0 = (int)std::floor(x);
    int y0 = (int)std::floor(y);
    float dx = x - x0;
    float dy = y - y0;
    int px0 = (x0 % W + W) % W;
    int px1 = (px0 + 1) % W;
    int py0 = std::clamp(y0, 0, H - 1);
    int py1 = std::clamp(y0 + 1, 0, H - 1);
    C3 c00 = L.rgb[(size_t)py0 * W + px0], c10 = L.rgb[(size_t)py0 * W + px1];
    C3 c01 = L.rgb[(size_t)py1 * W + px0], c11 = L.rgb[(size_t)py1 * W + px1];
    C3 c0 = c00 * (1.0f - dx) + c10 * dx;
    C3 c1 = c01 * (1.0f - dx) + c11 * dx;
    return c0 * (1.0f - dy) + c1 * dy;
}
// EnvironmentMap::eval (environment.cpp:31-40)
static C3 envEval(const Light& L, V3 w) {
    P2 sph = sphericalCoordinates(w);
    float u = sph.y * M_1_PI * 0.5f;
    float v = sph.x * M_1_PI;
    return envBilinear(L, u, v) * L.strength;  // float * Color3f, one product per channel
}
// EnvironmentMap::pdf (environment.cpp:75-91)
static float envPdf(const Light& L, V3 refP, V3 infoP) {
    V3 wi = Normalize(infoP - refP);
    P2 sph = sphericalCoordinates(wi);
    float sinTheta = std::sin(sph.x);
    if (sinTheta <= 0.0f) return 0.0f;
    float u = sph.y * M_1_PI * 0.5f;
    float v = sph.x * M_1_PI;
    float mapPdf = L.dist.pdf({u, v});
    return mapPdf / (2.0f * M_PI * M_PI * sinTheta);
}
// Emitter::sample of the non-area emitters (pointlight.cpp:18-33,
// spotlight.cpp:28-60, directional.cpp:17-33, environment.cpp:42-73)
static C3 lightSample(const Light& L, V3 refP, P2 u, LightSample& ls, float time) {
    switch (L.d.type) {
        case GND_LIGHT_POINT: {
            C3 I = L.power * kInv4Pi;
            float squaredDist = LengthSq(L.position - refP);
            ls.p = L.position;
            ls.pdf = 1.0f;
            ls.shadow.o = refP;
            ls.shadow.d = Normalize(ls.p - refP);
            ls.shadow.tMin = kEpsilon;
            ls.shadow.tMax = Length(ls.p - refP) - kEpsilon;
            ls.shadow.time = 0.0f;  // info.time of a fresh interaction (pointlight.cpp:29)
            return I / squaredDist;
        }
        case GND_LIGHT_SPOT: {
            V3 wi = Normalize(L.position - refP);
            float dist = Length(L.position - refP);
            float squaredDist = dist * dist;
            ls.p = L.position;
            ls.pdf = 1.0f;
            ls.shadow.o = refP;
            ls.shadow.d = wi;
            ls.shadow.tMin = kEpsilon;
            ls.shadow.tMax = dist - kEpsilon;
            ls.shadow.time = time;  // spotlight.cpp:42
            float cosTheta = Dot(-wi, L.direction);
            if (cosTheta <= L.cosTotalWidth) return C3(0.0f);
            float falloff = 1.0f;
            if (cosTheta < L.cosFalloffStart) {
                float t = (cosTheta - L.cosTotalWidth) / (L.cosFalloffStart - L.cosTotalWidth);
                falloff = (t * t) * (3.0f - 2.0f * t);
            }
            C3 I = L.power * kInv4Pi;
            return (I * falloff) / squaredDist;
        }
        case GND_LIGHT_DIRECTIONAL: {
            V3 wi = L.direction;
            ls.p = refP + wi * 1e5f;
            ls.pdf = 1.0f;
            ls.shadow.o = refP;
            ls.shadow.d = wi;
            ls.shadow.tMin = kEpsilon;
            ls.shadow.tMax = std::numeric_limits<float>::max();
            ls.shadow.time = time;  // directional.cpp:31
            return L.irradiance;
        }
        default: {  // environment
            float mapPdf;
            P2 uv = L.dist.sampleContinuous(u, &mapPdf);
            if (mapPdf == 0.0f) {
                ls.pdf = 0.0f;
                return C3(0.0f);
            }
            float theta = uv.y * M_PI;
            float phi = uv.x * 2.0f * M_PI;
            V3 wi = sphericalDirection(theta, phi);
            ls.p = refP + wi * 1e5f;
            ls.n = -wi;
            ls.shadow = Ray();
            ls.shadow.o = refP;
            ls.shadow.d = wi;
            ls.shadow.time = time;  // environment.cpp:62
            float sinTheta = std::sin(theta);
            if (sinTheta <= 0.0f) {
                ls.pdf = 0.0f;
                return C3(0.0f);
            }
            ls.pdf = mapPdf / (2.0f * M_PI * M_PI * sinTheta);
            return envEval(L, wi) / ls.pdf;
        }
    }
}
// one NEE light sample from emitter-list entry `e` (area or other light);
// `delta` is Emitter::isDelta
static C3 emitterSample(const Scene& s, int e, V3 refP, P2 u, LightSample& ls, bool& pdfSet, bool& delta,
                        float time) {
    if (e >= 0) {
        delta = false;
        return areaSample(s, s.prims[e], refP, u, ls, pdfSet, time);
    }
    const Light& L = s.lights[-e - 1];
    delta = L.isDelta();
    pdfSet = true;
    return lightSample(L, refP, u, ls, time);
}
// the environment seen by a ray that left the scene, with the MIS weight
// against NEE (path.cpp:36-55, volPath.cpp:117-139); pdf against the origin
// as the reference's default SurfaceInteraction reference point
static bool envMiss(const Scene& s, const Ray& r, bool specular, float pdfPrev, C3& Le, float& w) {
    if (s.envLight < 0) return false;
    const Light& L = s.lights[s.envLight];
    Le = envEval(L, r.d);
    if (Le.isBlack()) return false;
    w = 1.0f;
    if (!specular) {
        float pe = envPdf(L, V3(0, 0, 0), r.o + r.d * 1e5f) * (1.0f / s.emitters.size());
        w = powerHeuristic(pdfPrev, pe);
    }
    return true;
}

// ------------------------------------------------------------ media
struct MediumEvent {
    bool valid = false;
    V3 p, wo;
    C3 sigma_a, sigma_s;
    float g = 0.0f;
};
static float hgPhase(float cosT, float g) {  // phase.h:86-89
    float denom = 1.0f + g * g + 2.0f * g * cosT;
    return kInv4Pi * (1.0f - g * g) / (denom * std::sqrt(denom));
}
static float hgClamp(float g) { return clampf(g, -0.999f, 0.999f); }
static float hgSample(float g, V3 wo, V3* wi, P2 u) {  // phase.h:49-73 (sampled density is HG(-g); kept as is)
    float cosT;
    if (std::abs(g) < 1e-3f) cosT = 1.0f - 2.0f * u.x;
    else {
        float sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * u.x);
        cosT = (1.0f + g * g - sq * sq) / (2.0f * g);
    }
    float sinT = std::sqrt(std::max(0.0f, 1.0f - cosT * cosT));
    float phi = (float)(2.0f * M_PI * (double)u.y);
    V3 v1, v2;
    if (std::abs(wo.x) > std::abs(wo.y)) {
        float inv = 1.0f / std::sqrt(wo.x * wo.x + wo.z * wo.z);
        v1 = V3(-wo.z * inv, 0.0f, wo.x * inv);
    } else {
        float inv = 1.0f / std::sqrt(wo.y * wo.y + wo.z * wo.z);
        v1 = V3(0.0f, wo.z * inv, -wo.y * inv);
    }
    v2 = Cross(wo, v1);
    *wi = v1 * (sinT * std::cos(phi)) + v2 * (sinT * std::sin(phi)) + wo * cosT;
    return hgPhase(cosT, g);
}

// OpenVDB BoxSampler::isSample on a dense grid: trilinear, lerp z, then y, then x
static float gridSample(const Medium& m, double x, double y, double z) {
    int ix = (int)std::floor(x), iy = (int)std::floor(y), iz = (int)std::floor(z);
    double ux = x - ix, uy = y - iy, uz = z - iz;
    auto at = [&](int i, int j, int k) -> float {
        if (i < 0 || j < 0 || k < 0 || i >= m.d.nx || j >= m.d.ny || k >= m.d.nz) return 0.0f;
        return m.density[((size_t)k * m.d.ny + j) * m.d.nx + i];
    };
    float d000 = at(ix, iy, iz), d001 = at(ix, iy, iz + 1), d010 = at(ix, iy + 1, iz), d011 = at(ix, iy + 1, iz + 1);
    float d100 = at(ix + 1, iy, iz), d101 = at(ix + 1, iy, iz + 1), d110 = at(ix + 1, iy + 1, iz),
          d111 = at(ix + 1, iy + 1, iz + 1);
    float a = d000 + (float)((d001 - d000) * uz);
    float b = d010 + (float)((d011 - d010) * uz);
    float r1 = a + (float)((b - a) * uy);
    a = d100 + (float)((d101 - d100) * uz);
    b = d110 + (float)((d111 - d110) * uz);
    float r2 = a + (float)((b - a) * uy);
    return r1 + (float)(ux * (r2 - r1));
}

// Heterogeneous set-up shared by Tr and sample (heterogeneous.cpp:106-162, 251-305)
struct GridWalk {
    Ray ir;
    float t, tMax;
    int cx, cy, cz, stepX, stepY, stepZ;
    float tDX, tDY, tDZ, tMX, tMY, tMZ;
};
static bool gridSetup(const Medium& m, const Ray& ray, GridWalk& g) {
    V3 lo = xfPoint(m.toLocal, ray.o);
    V3 lt = xfPoint(m.toLocal, ray.o + ray.d);
    double iox = lo.x, ioy = lo.y, ioz = lo.z;
    double idx = (double)lt.x - iox, idy = (double)lt.y - ioy, idz = (double)lt.z - ioz;
    g.ir.o = V3((float)iox, (float)ioy, (float)ioz);
    g.ir.d = V3((float)idx, (float)idy, (float)idz);
    g.ir.tMin = ray.tMin;
    g.ir.tMax = ray.tMax;
    g.t = ray.tMin;
    g.tMax = ray.tMax;
    Bounds gb;
    gb.pMin = m.gridMin;
    gb.pMax = m.gridMax;
    float h0, h1;
    if (!boundsRayIntersect(gb, g.ir.o, g.ir.d, g.ir.tMax, &h0, &h1)) return false;
    g.t = std::max(g.t, h0);
    g.tMax = std::min(g.tMax, h1);
    V3 pe = g.ir.o + g.ir.d * g.t;
    g.cx = std::clamp((int)std::floor((pe.x - m.gridMin.x) / m.cellX), 0, m.resX - 1);
    g.cy = std::clamp((int)std::floor((pe.y - m.gridMin.y) / m.cellY), 0, m.resY - 1);
    g.cz = std::clamp((int)std::floor((pe.z - m.gridMin.z) / m.cellZ), 0, m.resZ - 1);
    g.stepX = g.ir.d.x > 0.0f ? 1 : -1;
    g.stepY = g.ir.d.y > 0.0f ? 1 : -1;
    g.stepZ = g.ir.d.z > 0.0f ? 1 : -1;
    const float FM = std::numeric_limits<float>::max();
    g.tDX = g.ir.d.x != 0.0f ? std::abs(m.cellX / g.ir.d.x) : FM;
    g.tDY = g.ir.d.y != 0.0f ? std::abs(m.cellY / g.ir.d.y) : FM;
    g.tDZ = g.ir.d.z != 0.0f ? std::abs(m.cellZ / g.ir.d.z) : FM;
    float nbx = m.gridMin.x + (g.cx + (g.stepX > 0 ? 1 : 0)) * m.cellX;
    float nby = m.gridMin.y + (g.cy + (g.stepY > 0 ? 1 : 0)) * m.cellY;
    float nbz = m.gridMin.z + (g.cz + (g.stepZ > 0 ? 1 : 0)) * m.cellZ;
    g.tMX = g.ir.d.x != 0.0f ? (nbx - g.ir.o.x) / g.ir.d.x : FM;
    g.tMY = g.ir.d.y != 0.0f ? (nby - g.ir.o.y) / g.ir.d.y : FM;
    g.tMZ = g.ir.d.z != 0.0f ? (nbz - g.ir.o.z) / g.ir.d.z : FM;
    if (g.tMX <= g.t) g.tMX += g.tDX;
    if (g.tMY <= g.t) g.tMY += g.tDY;
    if (g.tMZ <= g.t) g.tMZ += g.tDZ;
    return true;
}
static void gridAdvance(GridWalk& g) {
    if (g.tMX < g.tMY) {
        if (g.tMX < g.tMZ) { g.tMX += g.tDX; g.cx += g.stepX; }
        else { g.tMZ += g.tDZ; g.cz += g.stepZ; }
    } else {
        if (g.tMY < g.tMZ) { g.tMY += g.tDY; g.cy += g.stepY; }
        else { g.tMZ += g.tDZ; g.cz += g.stepZ; }
    }
}

// Medium::Tr (homogeneous.cpp:22-30, heterogeneous.cpp:106-229)
static C3 mediumTr(const Medium& m, const Ray& ray, Sampler& smp) {
    if (m.d.type == GND_MEDIUM_HOMOGENEOUS) {
        if (std::isinf(ray.tMax)) return C3(0.0f);
        return C3(std::exp(-m.sigma_t.r * ray.tMax), std::exp(-m.sigma_t.g * ray.tMax),
                  std::exp(-m.sigma_t.b * ray.tMax));
    }
    C3 tr(1.0f);
    GridWalk g;
    if (!gridSetup(m, ray, g)) return tr;
    float t = g.t;
    while (t < g.tMax) {
        if (g.cx < 0 || g.cx >= m.resX || g.cy < 0 || g.cy >= m.resY || g.cz < 0 || g.cz >= m.resZ) break;
        float maj = m.majorants[(size_t)g.cz * (m.resX * m.resY) + g.cy * m.resX + g.cx];
        float tExit = std::min({g.tMX, g.tMY, g.tMZ, g.tMax});
        if (maj > 0.0f) {
            float invMaj = 1.0f / maj;
            while (t < tExit) {
                float step = -std::log(1.0f - smp.next1D()) * invMaj;
                if (step <= 0.0f) step = 1e-5f;
                if (t + step >= tExit) break;
                t += step;
                double px = g.ir.o.x + g.ir.d.x * t, py = g.ir.o.y + g.ir.d.y * t, pz = g.ir.o.z + g.ir.d.z * t;
                float dens = gridSample(m, px, py, pz) * m.d.density_scale;
                if (dens < 1e-5f) dens = 0.0f;
                if (dens > 0.0f) {
                    C3 lst = m.sigma_t * dens;
                    tr = tr * (C3(1.0f) - lst * invMaj);
                    float maxTr = std::max({tr.r, tr.g, tr.b});
                    if (maxTr < 1e-4f) return C3(0.0f);
                    if (maxTr < 0.1f) {
                        float q = std::max(0.05f, 1.0f - maxTr);
                        if (smp.next1D() < q) return C3(0.0f);
                        tr = tr / (1.0f - q);
                    }
                }
            }
        }
        t = tExit;
        if (t >= g.tMax) break;
        gridAdvance(g);
    }
    return tr;
}

// Medium::sample (homogeneous.cpp:32-63, heterogeneous.cpp:251-376)
static C3 mediumSample(const Medium& m, const Ray& ray, Sampler& smp, MediumEvent& mi) {
    if (m.d.type == GND_MEDIUM_HOMOGENEOUS) {
        int ch = std::min((int)(smp.next1D() * 3.0f), 2);
        float dist = -std::log(1.0f - smp.next1D()) / m.sigma_t[ch];
        float t = std::min(dist, ray.tMax);
        bool sampled = t < ray.tMax;
        C3 tr(std::exp(-m.sigma_t.r * t), std::exp(-m.sigma_t.g * t), std::exp(-m.sigma_t.b * t));
        C3 pc = sampled ? (m.sigma_t * tr) : tr;
        float pdf = (pc.r + pc.g + pc.b) / 3.0f;
        if (pdf == 0.0f) return C3(0.0f);
        if (sampled) {
            mi.valid = true;
            mi.p = ray.o + ray.d * t;
            mi.wo = -ray.d;
            mi.g = hgClamp(m.d.g);
            mi.sigma_a = m.sigma_a;
            mi.sigma_s = m.sigma_s;
        }
        return tr / pdf;
    }
    GridWalk g;
    if (!gridSetup(m, ray, g)) return C3(1.0f);
    float t = g.t;
    while (t < g.tMax) {
        if (g.cx < 0 || g.cx >= m.resX || g.cy < 0 || g.cy >= m.resY || g.cz < 0 || g.cz >= m.resZ) break;
        float maj = m.majorants[(size_t)g.cz * (m.resX * m.resY) + g.cy * m.resX + g.cx];
        float tExit = std::min({g.tMX, g.tMY, g.tMZ, g.tMax});
        if (maj > 0.0f) {
            float invMaj = 1.0f / maj;
            while (t < tExit) {
                float step = -std::log(1.0f - smp.next1D()) * invMaj;
                if (step <= 0.0f) step = 1e-5f;
                if (t + step >= tExit) break;
                t += step;
                double px = g.ir.o.x + g.ir.d.x * t, py = g.ir.o.y + g.ir.d.y * t, pz = g.ir.o.z + g.ir.d.z * t;
                float dens = gridSample(m, px, py, pz) * m.d.density_scale;
                if (dens < 1e-5f) dens = 0.0f;
                if (dens > 0.0f) {
                    int ch = std::min((int)(smp.next1D() * 3.0f), 2);
                    float prob = dens * m.sigma_t[ch] * invMaj;
                    if (smp.next1D() < prob) {
                        mi.valid = true;
                        mi.p = ray.o + ray.d * t;
                        mi.wo = -ray.d;
                        mi.g = hgClamp(m.d.g);
                        C3 lst = m.sigma_t * dens;
                        mi.sigma_a = m.sigma_a * dens;
                        mi.sigma_s = m.sigma_s * dens;
                        return C3(1.0f) / lst[ch];
                    }
                }
            }
        }
        t = tExit;
        if (t >= g.tMax) break;
        gridAdvance(g);
    }
    return C3(1.0f);
}

// ------------------------------------------------------------ camera
// PerspectiveCamera::shootRay (perspective.cpp:15-32), ThinLensCamera::shootRay (thinLens.cpp:34-95)
struct CamSample {
    P2 pFilm, pLens;
    float lambdaOffset = 0.0f;
    float time = 0.0f;
};
static float shootRay(const Camera& c, const CamSample& cs, Ray& ray) {
    V3 dir, org(0, 0, 0);
    float weight = 1.0f;
    if (c.d.type == GND_CAMERA_PERSPECTIVE) {
        float x = (2.0f * cs.pFilm.x - 1.0f) * c.aspect * c.scale;
        float y = (2.0f * cs.pFilm.y - 1.0f) * c.scale;
        dir = Normalize(V3(x, y, -1.0f));
    } else {
        float ndcX = 2.0f * cs.pFilm.x - 1.0f, ndcY = 2.0f * cs.pFilm.y - 1.0f;
        float r2 = ndcX * ndcX + ndcY * ndcY;
        float dist = 1.0f + c.d.k1 * r2;
        float distX = ndcX * dist, distY = ndcY * dist;
        float scaleMod = c.scale * (1.0f + c.d.ca_lateral * cs.lambdaOffset);
        float x = distX * c.aspect * scaleMod, y = distY * scaleMod;
        dir = V3(x, y, -1.0f);
        if (c.d.lens_radius > 0.0f) {
            P2 pl = c.d.blades >= 3 ? squareToUniformPolygon(cs.pLens, c.d.blades) : squareToUniformDisk(cs.pLens);
            pl.x = pl.x * c.d.lens_radius;
            pl.y = pl.y * c.d.lens_radius;
            float fdm = c.d.focal_distance * (1.0f + c.d.ca_axial * cs.lambdaOffset);
            float ft = fdm / std::abs(dir.z);
            V3 pFocus = org + dir * ft;
            org = V3(pl.x, pl.y, 0.0f);
            dir = pFocus - org;
        }
        dir = Normalize(dir);
        if (c.d.vignetting) {
            float ct = std::abs(dir.z);
            float c2 = ct * ct;
            weight = c2 * c2;
        }
    }
    // rolling shutter (thinLens.cpp:81-87), then the camera transform at that
    // time (AnimatedTransform::interpolate)
    float finalTime = cs.time;
    if (c.d.type != GND_CAMERA_PERSPECTIVE && c.d.rolling_shutter) {
        float scanline = 1.0f - cs.pFilm.y;
        finalTime = scanline * c.d.scan_duration + (cs.time * c.d.exposure_duration);
    }
    const M4 toWorld = motionInterpolate(c.motion, finalTime);
    ray = Ray();
    ray.o = xfPoint(toWorld, org);
    ray.d = xfVector(toWorld, dir);
    ray.time = finalTime;
    return weight;
}

// ------------------------------------------------------------ AOVs
// Material::getAlbedo / getRoughness / getMetallic (material.h:24-34;
// matte.cpp:55-57 albedo texture, conductor.cpp:122-132 R / roughness / 1,
// glass.cpp:117-119 roughness; plastic and index keep the defaults)
static C3 aovAlbedo(const gnd_material_desc& m, const BSDF& b) {
    if (m.type == GND_MAT_MATTE) return b.lobe.R;  // the albedo texture at the hit
    if (m.type == GND_MAT_CONDUCTOR) return C3(m.R[0], m.R[1], m.R[2]);
    return C3(0.0f);
}
static float aovRoughness(const gnd_material_desc& m) {
    return (m.type == GND_MAT_CONDUCTOR || m.type == GND_MAT_GLASS) ? m.roughness : 0.0f;
}
static float aovMetallic(const gnd_material_desc& m) { return m.type == GND_MAT_CONDUCTOR ? 1.0f : 0.0f; }
// Camera::getForward (camera.h:40-43): the camera-space -z axis in world space
static V3 cameraForward(const Camera& c) { return xfVector(c.toWorld, V3(0.0f, 0.0f, -1.0f)); }

// The G-buffer writes of both integrators (path.cpp:35-61 and 99-106,
// volPath.cpp:117-146 and 184-191).  A miss: depth / metallic / roughness
// reset at bounce 0, albedo = the environment along the ray (black without
// one), normal 0.
static void aovMiss(AOV* a, const Scene& s, V3 dir, int bounces, bool& needs) {
    if (!a) return;
    if (bounces == 0) {
        a->depth = -1.0f;
        a->metallic = 0.0f;
        a->roughness = 0.0f;
    }
    if (needs) {
        a->albedo = s.envLight >= 0 ? envEval(s.lights[s.envLight], dir) : C3(0.0f);
        a->normal = V3(0.0f, 0.0f, 0.0f);
        needs = false;
    }
}
// A surface with a BSDF: depth / metallic / roughness at bounce 0; albedo and
// shading normal (mapped to [0,1]) at the first surface with a non-specular
// lobe, or at the last bounce.
static void aovSurface(AOV* a, const Scene& s, const Isect& it, const BSDF& b, V3 primaryO, int bounces,
                       bool& needs) {
    if (!a) return;
    const gnd_material_desc& m = s.materials[s.prims[it.prim].material].d;
    if (bounces == 0) {
        a->depth = Dot(it.p - primaryO, cameraForward(s.camera));
        a->metallic = aovMetallic(m);
        a->roughness = aovRoughness(m);
    }
    if (needs && (b.nonSpecular() || bounces == s.maxDepth - 1)) {
        a->albedo = aovAlbedo(m, b);
        a->normal = it.n / 2.0f + V3(0.5f, 0.5f, 0.5f);
        needs = false;
    }
}

// ------------------------------------------------------------ integrators
// PathIntegrator::Li (src/integrators/path.cpp:31-178)
static C3 LiPath(const Scene& s, Ray r, Sampler& smp, Counters* cnt, AOV* aov) {
    C3 L(0.0f), tp(1.0f);
    int bounces = 0, nullBounces = 0;
    bool specular = true;
    float pdfPrev = 1.0f;
    V3 lastScatter = r.o;
    const V3 primaryO = r.o;
    bool needsG = aov != nullptr;
    float invNE = 1.0f / s.emitters.size();
    while (bounces < s.maxDepth) {
        Hit h;
        if (!intersectClosest(s, r, h, cnt)) {
            C3 Le;
            float we;
            if (envMiss(s, r, specular, pdfPrev, Le, we)) L = L + tp * Le * we;
            aovMiss(aov, s, r.d, bounces, needsG);
            break;
        }
        Isect it;
        fillInteraction(s, r, h, it);
        const Prim& pr = s.prims[it.prim];
        BSDF bsdf;
        bool hasBsdf = makeBSDF(s, it, bsdf);
        if (pr.emitter >= 0) {
            bool skip = (bounces == 0 && !hasBsdf);
            if (!skip) {
                C3 Le = areaEval(pr, it.n, -r.d);
                if (!Le.isBlack()) {
                    float w = 1.0f;
                    if (!specular) {
                        float lp = areaPdfAt(s, pr, lastScatter, it.p, it.n, 0.0f);  // dummyRef: time 0
                        w = powerHeuristic(pdfPrev, lp * invNE);
                    }
                    L = L + tp * Le * w;
                }
            }
        }
        if (!hasBsdf) {
            if (nullBounces++ > 100) break;
            Ray nr;
            nr.time = r.time;  // the ray's time carries on (path.cpp:90, 139, 176)
            nr.o = it.p;
            nr.d = r.d;
            r = nr;
            continue;
        }
        aovSurface(aov, s, it, bsdf, primaryO, bounces, needsG);
        // next event estimation
        float lsp = 1.0f / s.emitters.size();
        int ep = pickEmitter(s, smp.next1D());
        LightSample ls;
        bool pdfSet;
        bool delta;
        C3 Li = emitterSample(s, ep, it.p, smp.next2D(), ls, pdfSet, delta, r.time);
        Li = Li / lsp;
        if (pdfSet && ls.pdf > 1e-6f && !Li.isBlack()) {
            bool occluded = false;
            Ray tr = ls.shadow;
            while (true) {
                Hit sh;
                if (!intersectClosest(s, tr, sh, cnt)) break;
                Isect sit;
                fillInteraction(s, tr, sh, sit);
                BSDF sb;
                if (makeBSDF(s, sit, sb)) { occluded = true; break; }
                float nextT = tr.tMax - Length(sit.p - tr.o);
                Ray nr;
                nr.time = tr.time;  // the ray's time carries on (path.cpp:90, 139, 176)
                nr.o = sit.p;
                nr.d = tr.d;
                nr.tMax = nextT;
                tr = nr;
            }
            if (!occluded) {
                V3 wi = Normalize(ls.p - it.p);
                float bp = bsdf.pdf(-r.d, wi);
                C3 f = bsdf.f(-r.d, wi);
                if (!f.isBlack() && bp > 1e-6f) {
                    float wl = powerHeuristic(ls.pdf * lsp, delta ? 0.0f : bp);
                    L = L + tp * f * Li * std::abs(Dot(it.n, wi)) * wl;
                }
            }
        }
        // BSDF sampling
        V3 wi;
        int st = 0;
        P2 u2 = smp.next2D();
        float uc = smp.next1D();
        C3 fc = bsdf.sample(-r.d, &wi, u2, uc, &pdfPrev, &st);
        if (fc.isBlack() || pdfPrev <= 1e-6f) break;
        specular = (st & BSDF_SPECULAR) != 0;
        tp = tp * fc;
        if (bounces > 2) {
            float q = std::max(0.05f, std::min(tp.luminance(), 0.99f));
            if (smp.next1D() > q) break;
            tp = tp / q;
        }
        lastScatter = it.p;
        Ray nr;
        nr.time = r.time;  // the ray's time carries on (path.cpp:90, 139, 176)
        nr.o = it.p;
        nr.d = wi;
        r = nr;
        bounces++;
    }
    return L;
}

// VolumetricPathIntegrator::evaluateTr (volPath.cpp:261-291)
static C3 evaluateTr(const Scene& s, Ray r, Sampler& smp, V3 lightP, Counters* cnt) {
    C3 Tr(1.0f);
    while (true) {
        Hit h;
        bool hit = intersectClosest(s, r, h, cnt);
        if (r.medium >= 0) {
            Ray trr = r;
            trr.tMax = hit ? h.t : r.tMax;
            Tr = Tr * mediumTr(s.media[r.medium], trr, smp);
        }
        if (Tr.isBlack()) return Tr;
        if (!hit) break;
        Isect it;
        fillInteraction(s, r, h, it);
        BSDF b;
        if (makeBSDF(s, it, b)) return C3(0.0f);
        Ray nr;
        nr.time = r.time;  // the ray's time carries on (path.cpp:90, 139, 176)
        nr.o = it.p;
        nr.d = r.d;
        nr.medium = Dot(r.d, it.n) > 0.0f ? it.mediumOutside : it.mediumInside;  // getMedium
        nr.tMin = kEpsilon;
        nr.tMax = Length(it.p - lightP) - kEpsilon;
        r = nr;
    }
    return Tr;
}
static int nextMedium(V3 wo, V3 wi, int cur, const Isect& it) {  // volPath.cpp:293-298
    if (Dot(wi, it.n) > 0.0f) return Dot(wo, it.n) > 0.0f ? cur : it.mediumOutside;
    return Dot(wo, it.n) > 0.0f ? it.mediumInside : cur;
}

// VolumetricPathIntegrator::Li (volPath.cpp:37-249)
// DirectLight::Li (src/integrators/direct.cpp:10-76): emission of the
// camera hit, one light sample (balance heuristic, p_bsdf = 0 for delta
// lights; the shadow ray a plain any-hit test) and one BSDF sample whose hit
// adds its emission; "sampledType == BSDF_SPECULAR" (direct.cpp:66) is an
// equality test.  No environment term and no AOV writes.
static C3 LiDirect(const Scene& s, Ray r, Sampler& smp, Counters* cnt) {
    Hit h;
    if (!intersectClosest(s, r, h, cnt)) return C3(0.0f);
    C3 L(0.0f);
    Isect it;
    fillInteraction(s, r, h, it);
    const Prim& pr = s.prims[it.prim];
    if (pr.emitter >= 0) L = L + areaEval(pr, it.n, -r.d);
    BSDF bsdf;
    if (!makeBSDF(s, it, bsdf)) return L;
    float lightSelectPdf = 1.0f / s.emitters.size();
    int ep = pickEmitter(s, smp.next1D());
    LightSample ls;
    bool pdfSet, delta;
    C3 Li = emitterSample(s, ep, it.p, smp.next2D(), ls, pdfSet, delta, r.time);
    Li = Li / lightSelectPdf;
    if (pdfSet && ls.pdf > 1e-6f && !Li.isBlack()) {
        Ray sr = ls.shadow;
        if (!intersectAny(s, sr, cnt)) {
            V3 wi = Normalize(ls.p - it.p);
            float bsdfPdf = bsdf.pdf(-r.d, wi);
            C3 f = bsdf.f(-r.d, wi);
            if (!f.isBlack() && bsdfPdf > 1e-6f) {
                float pLight = ls.pdf * lightSelectPdf;
                float pBsdf = delta ? 0.0f : bsdfPdf;
                float misDenom = pLight + pBsdf;
                if (misDenom > 1e-6f) L = L + f * Li * std::abs(Dot(it.n, wi)) * pLight / misDenom;
            }
        }
    }
    V3 wi;
    float bsdfPdf = 0.0f;
    int st = 0;
    P2 u2 = smp.next2D();
    float uc = smp.next1D();
    C3 f = bsdf.sample(-r.d, &wi, u2, uc, &bsdfPdf, &st);
    if (!f.isBlack() && bsdfPdf > 1e-6f) {
        Ray sec;
        sec.o = it.p;
        sec.d = wi;
        Hit h2;
        if (intersectClosest(s, sec, h2, cnt)) {
            Isect it2;
            fillInteraction(s, sec, h2, it2);
            const Prim& p2 = s.prims[it2.prim];
            if (p2.emitter >= 0) {
                C3 Lb = areaEval(p2, it2.n, -wi);
                if (!Lb.isBlack()) {
                    float lightPdfGeo = areaPdfAt(s, p2, it.p, it2.p, it2.n, r.time);  // pdf(isect, ...)
                    if (std::isinf(lightPdfGeo)) lightPdfGeo = 0.0f;
                    float pLight = st == BSDF_SPECULAR ? 0.0f : lightPdfGeo * lightSelectPdf;
                    float misDenom = bsdfPdf + pLight;
                    if (misDenom > 1e-6f) L = L + f * Lb * bsdfPdf / misDenom;
                }
            }
        }
    }
    return L;
}

// AmbientOcclusion::Li (src/integrators/ao.cpp:14-40)
static C3 LiAO(const Scene& s, Ray r, Sampler& smp, Counters* cnt) {
    Hit h;
    if (!intersectClosest(s, r, h, cnt)) return C3(1.0f);
    Isect it;
    fillInteraction(s, r, h, it);
    V3 local = squareToCosineHemisphere(smp.next2D());
    V3 n = it.n;
    V3 t1 = std::abs(n.x) > 0.1f ? Normalize(Cross(V3(0.0f, 1.0f, 0.0f), n)) : Normalize(Cross(V3(1.0f, 0.0f, 0.0f), n));
    V3 t2 = Cross(n, t1);
    Ray sec;
    sec.o = it.p;
    sec.d = t1 * local.x + t2 * local.y + n * local.z;
    sec.tMax = s.aoLength;
    if (intersectAny(s, sec, cnt)) return C3(0.0f);
    return C3(1.0f);
}

static C3 LiVolPath(const Scene& s, Ray r, Sampler& smp, Counters* cnt, AOV* aov) {
    C3 L(0.0f), tp(1.0f);
    int bounces = 0, surfB = 0, volB = 0, nullB = 0;
    float pdfPrev = 1.0f;
    bool specular = true;
    V3 lastScatter = r.o;
    const V3 primaryO = r.o;
    bool needsG = aov != nullptr;
    float invNE = 1.0f / s.emitters.size();
    while (bounces < s.maxDepth) {
        Hit h;
        bool hitSurface = intersectClosest(s, r, h, cnt);
        MediumEvent mi;
        if (r.medium >= 0) tp = tp * mediumSample(s.media[r.medium], r, smp, mi);
        if (mi.valid) {
            // no volumetric emission on the path (Medium::Le is zero without a temperature grid)
            tp = tp * mi.sigma_s;
            if (tp.isBlack()) break;
            float lsp = 1.0f / s.emitters.size();
            int ep = pickEmitter(s, smp.next1D());
            LightSample ls;
            bool pdfSet;
            bool delta;
            // mi_ref holds only the point (volPath.cpp:60-61): time 0
            C3 Li = emitterSample(s, ep, mi.p, smp.next2D(), ls, pdfSet, delta, 0.0f);
            Li = Li / lsp;
            if (pdfSet && ls.pdf > 1e-6f && !Li.isBlack()) {
                Ray sr = ls.shadow;
                sr.medium = r.medium;
                C3 Tr = evaluateTr(s, sr, smp, ls.p, cnt);
                if (!Tr.isBlack()) {
                    V3 wi = Normalize(ls.p - mi.p);
                    float pp = hgPhase(Dot(mi.wo, wi), mi.g);
                    if (pp > 1e-6f) {
                        float wl = powerHeuristic(ls.pdf * lsp, delta ? 0.0f : pp);
                        L = L + tp * pp * Li * wl * Tr;
                    }
                }
            }
            V3 wi;
            pdfPrev = hgSample(mi.g, mi.wo, &wi, smp.next2D());
            specular = false;
            int prevMedium = r.medium;
            lastScatter = mi.p;
            Ray nr;
            nr.time = r.time;  // the ray's time carries on (path.cpp:90, 139, 176)
            nr.o = mi.p;
            nr.d = wi;
            nr.medium = prevMedium;
            r = nr;
            if (volB > 6) {
                float q = std::max(0.05f, std::min(tp.luminance(), 0.99f));
                if (smp.next1D() > q) break;
                tp = tp / q;
            }
            bounces++;
            volB++;
            continue;
        }
        int curMedium = r.medium;
        if (!hitSurface) {
            C3 Le;
            float we;
            if (envMiss(s, r, specular, pdfPrev, Le, we)) L = L + tp * Le * we;
            aovMiss(aov, s, r.d, bounces, needsG);
            break;
        }
        Isect it;
        fillInteraction(s, r, h, it);
        const Prim& pr = s.prims[it.prim];
        BSDF bsdf;
        bool hasBsdf = makeBSDF(s, it, bsdf);
        if (pr.emitter >= 0) {
            bool skip = (bounces == 0 && !hasBsdf);
            if (!skip) {
                C3 Le = areaEval(pr, it.n, -r.d);
                if (!Le.isBlack()) {
                    float w = 1.0f;
                    if (!specular) {
                        float lp = areaPdfAt(s, pr, lastScatter, it.p, it.n, 0.0f);  // dummyRef: time 0
                        w = powerHeuristic(pdfPrev, lp * invNE);
                    }
                    L = L + tp * Le * w;
                }
            }
        }
        if (!hasBsdf) {
            if (nullB++ > 100) break;
            Ray nr;
            nr.time = r.time;  // the ray's time carries on (path.cpp:90, 139, 176)
            nr.o = it.p;
            nr.d = r.d;
            nr.medium = Dot(r.d, it.n) > 0.0f ? it.mediumOutside : it.mediumInside;
            r = nr;
            continue;
        }
        aovSurface(aov, s, it, bsdf, primaryO, bounces, needsG);
        float lsp = 1.0f / s.emitters.size();
        int ep = pickEmitter(s, smp.next1D());
        LightSample ls;
        bool pdfSet;
        bool delta;
        C3 Li = emitterSample(s, ep, it.p, smp.next2D(), ls, pdfSet, delta, r.time);
        Li = Li / lsp;
        if (pdfSet && ls.pdf > 1e-6f && !Li.isBlack()) {
            Ray sr = ls.shadow;
            sr.medium = nextMedium(-r.d, sr.d, curMedium, it);
            C3 Tr = evaluateTr(s, sr, smp, ls.p, cnt);
            if (!Tr.isBlack()) {
                V3 wi = Normalize(ls.p - it.p);
                float bp = bsdf.pdf(-r.d, wi);
                C3 f = bsdf.f(-r.d, wi);
                if (!f.isBlack() && bp > 1e-6f) {
                    float wl = powerHeuristic(ls.pdf * lsp, delta ? 0.0f : bp);
                    L = L + tp * f * Li * std::abs(Dot(it.n, wi)) * wl * Tr;
                }
            }
        }
        int st = 0;
        V3 wi;
        P2 u2 = smp.next2D();
        float uc = smp.next1D();
        C3 fc = bsdf.sample(-r.d, &wi, u2, uc, &pdfPrev, &st);
        if (fc.isBlack() || pdfPrev <= 1e-6f) break;
        specular = (st & BSDF_SPECULAR) != 0;
        tp = tp * fc;
        V3 inc = -r.d;
        lastScatter = it.p;
        Ray nr;
        nr.time = r.time;  // the ray's time carries on (path.cpp:90, 139, 176)
        nr.o = it.p;
        nr.d = wi;
        nr.medium = nextMedium(inc, wi, curMedium, it);
        r = nr;
        if (surfB > 2) {
            float q = std::max(0.05f, std::min(tp.luminance(), 0.99f));
            if (smp.next1D() > q) break;
            tp = tp / q;
        }
        bounces++;
        surfB++;
    }
    return L;
}

// SamplerIntegrator::Li of the scene's integrator
static C3 integrate(const Scene& s, const Ray& ray, Sampler& smp, Counters* c, AOV* ap) {
    switch (s.integrator) {
        case GND_INTEGRATOR_VOLPATH: return LiVolPath(s, ray, smp, c, ap);
        case GND_INTEGRATOR_DIRECT: return LiDirect(s, ray, smp, c);
        case GND_INTEGRATOR_AO: return LiAO(s, ray, smp, c);
        default: return LiPath(s, ray, smp, c, ap);
    }
}

// ------------------------------------------------------------ film + driver
// Film::addSample filter weights (film.h:26-69, filter.h:20-54)
static float filterEval(const Camera& c, float ox, float oy) {
    if (c.d.filter == GND_FILTER_BOX) return (std::abs(ox) <= 0.5f && std::abs(oy) <= 0.5f) ? 1.0f : 0.0f;
    float a = c.d.filter_alpha;
    float ex = std::exp(-a * c.d.filter_radius[0] * c.d.filter_radius[0]);
    float ey = std::exp(-a * c.d.filter_radius[1] * c.d.filter_radius[1]);
    float gx = std::max(0.0f, std::exp(-a * ox * ox) - ex);
    float gy = std::max(0.0f, std::exp(-a * oy * oy) - ey);
    return gx * gy;
}

struct SampleRec {
    float px, py;  // pFilm in pixel units
    C3 L;
    AOV aov;
};

// SamplerIntegrator::render (include/gianduia/core/integrator.h:49-150).  Each
// pass computes its samples row-parallel, then splats them with a per-pixel
// gather in a fixed neighbour order (the reference's CAS adds have no fixed
// order; the gather gives every run the same sums).
void render(const Scene& s, const RenderOptions& opt, Film& film, Counters* cnt) {
    const Camera& cam = s.camera;
    int W = cam.d.width, H = cam.d.height;
    film.width = W;
    film.height = H;
    film.rgbw.assign((size_t)W * H * 4, 0.0f);
    if (opt.accumulateSquares) film.sq.assign((size_t)W * H * 3, 0.0f);
    if (opt.aovs) film.aov.assign((size_t)W * H * kAovFloats, 0.0f);
    if (s.emitters.empty()) throw std::runtime_error("oracle render: scene has no emitter");
    int rb = opt.rowBegin, re = opt.rowEnd < 0 ? H : opt.rowEnd;
    int sb = opt.sppBegin, se = opt.sppEnd < 0 ? (int)s.spp : opt.sppEnd;
    float rx = cam.d.filter == GND_FILTER_BOX ? 0.5f : cam.d.filter_radius[0];
    float ry = cam.d.filter == GND_FILTER_BOX ? 0.5f : cam.d.filter_radius[1];
    int reach = (int)std::ceil(std::max(rx, ry)) + 1;
    std::vector<SampleRec> samples((size_t)W * H);
    int nt = std::max(1, opt.threads);
    std::vector<Counters> tc(nt);
    bool hasCA = cam.d.type == GND_CAMERA_THINLENS && (cam.d.ca_lateral != 0.0f || cam.d.ca_axial != 0.0f);
    for (int sIdx = sb; sIdx < se; ++sIdx) {
        std::atomic<int> nextRow{rb};
        auto work = [&](int tid) {
            Sampler smp;
            Counters* c = cnt ? &tc[tid] : nullptr;
            for (;;) {
                int y = nextRow.fetch_add(1);
                if (y >= re) break;
                for (int x = 0; x < W; ++x) {
                    uint64_t pixelIdx = (uint64_t)(y * W + x);
                    uint64_t g = pixelIdx + (uint64_t)sIdx * (uint64_t)(W * H) + opt.seedOffset;
                    smp.seed(g);
                    CamSample cs;
                    P2 ps = smp.next2D();
                    float pfx = x + ps.x, pfy = y + ps.y;
                    cs.pFilm.x = pfx / W;
                    cs.pFilm.y = 1.0f - pfy / H;
                    cs.pLens = smp.next2D();
                    float chR = smp.next1D();
                    int ch;
                    if (!hasCA) { ch = -1; cs.lambdaOffset = 0.0f; }
                    else if (chR < 0.333333f) { ch = 0; cs.lambdaOffset = 1.0f; }
                    else if (chR < 0.666666f) { ch = 1; cs.lambdaOffset = 0.0f; }
                    else { ch = 2; cs.lambdaOffset = -1.0f; }
                    cs.time = smp.next1D();
                    Ray ray;
                    float w = shootRay(cam, cs, ray);
                    AOV aov;
                    AOV* ap = opt.aovs ? &aov : nullptr;
                    C3 raw = integrate(s, ray, smp, c, ap);
                    raw = raw * w;
                    C3 nc(0.0f);
                    if (ch == -1) nc = raw;
                    else if (ch == 0) nc.r = raw.r * 3.0f;
                    else if (ch == 1) nc.g = raw.g * 3.0f;
                    else nc.b = raw.b * 3.0f;
                    if (nc.hasNaNs()) nc = C3(0.0f);
                    samples[(size_t)y * W + x] = {pfx, pfy, nc, aov};
                }
            }
        };
        std::vector<std::thread> th;
        for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
        work(0);
        for (auto& t : th) t.join();
        // gather splat
        std::atomic<int> nextRow2{rb};
        auto splat = [&]() {
            for (;;) {
                int y = nextRow2.fetch_add(1);
                if (y >= re) break;
                for (int x = 0; x < W; ++x) {
                    float sr = 0, sg = 0, sbb = 0, sw = 0;
                    float sa[kAovFloats] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                    for (int dy = -reach; dy <= reach; ++dy) {
                        int qy = y + dy;
                        if (qy < rb || qy >= re) continue;
                        for (int dx = -reach; dx <= reach; ++dx) {
                            int qx = x + dx;
                            if (qx < 0 || qx >= W) continue;
                            const SampleRec& q = samples[(size_t)qy * W + qx];
                            int x0 = std::max(0, (int)std::ceil(q.px - rx - 0.5f));
                            int x1 = std::min(W - 1, (int)std::floor(q.px + rx - 0.5f));
                            int y0 = std::max(0, (int)std::ceil(q.py - ry - 0.5f));
                            int y1 = std::min(H - 1, (int)std::floor(q.py + ry - 0.5f));
                            if (x < x0 || x > x1 || y < y0 || y > y1) continue;
                            float wgt = filterEval(cam, q.px - (x + 0.5f), q.py - (y + 0.5f));
                            if (wgt > 0.0f) {
                                sr += q.L.r * wgt;
                                sg += q.L.g * wgt;
                                sbb += q.L.b * wgt;
                                sw += wgt;
                                if (opt.aovs) {
                                    const float v[kAovFloats] = {q.aov.albedo.r, q.aov.albedo.g, q.aov.albedo.b,
                                                                 q.aov.normal.x, q.aov.normal.y, q.aov.normal.z,
                                                                 q.aov.depth,    q.aov.roughness, q.aov.metallic};
                                    for (int k = 0; k < kAovFloats; ++k) sa[k] += v[k] * wgt;
                                }
                            }
                        }
                    }
                    float* acc = &film.rgbw[((size_t)y * W + x) * 4];
                    acc[0] += sr;
                    acc[1] += sg;
                    acc[2] += sbb;
                    acc[3] += sw;
                    if (opt.aovs) {
                        float* av = &film.aov[((size_t)y * W + x) * kAovFloats];
                        for (int k = 0; k < kAovFloats; ++k) av[k] += sa[k];
                    }
                    if (opt.accumulateSquares && sw > 0.0f) {
                        float* q2 = &film.sq[((size_t)y * W + x) * 3];
                        float vr = sr / sw, vg = sg / sw, vb = sbb / sw;
                        q2[0] += vr * vr;
                        q2[1] += vg * vg;
                        q2[2] += vb * vb;
                    }
                }
            }
        };
        std::vector<std::thread> th2;
        for (int t = 1; t < nt; ++t) th2.emplace_back(splat);
        splat();
        for (auto& t : th2) t.join();
    }
    if (cnt)
        for (auto& c : tc) {
            cnt->rays += c.rays;
            cnt->nodes += c.nodes;
            cnt->packs += c.packs;
            cnt->insts += c.insts;
        }
}

}  // namespace gref

// ------------------------------------------------------------ probes for known-answer tests
using namespace gref;
extern "C" {

float gnd_ref_fr_dielectric(float cosI, float etaExt, float etaInt) { return FrDielectric(cosI, etaExt, etaInt); }

int gnd_ref_refract(const float* wi, const float* n, float eta, float* wt) {
    V3 o;
    bool ok = Refract(V3(wi[0], wi[1], wi[2]), V3(n[0], n[1], n[2]), eta, &o);
    if (ok) { wt[0] = o.x; wt[1] = o.y; wt[2] = o.z; }
    return ok ? 1 : 0;
}

void gnd_ref_fr_conductor(float cosI, const float* etaI, const float* etaT, const float* k, float* out) {
    for (int c = 0; c < 3; ++c) out[c] = conductorChannel(cosI, etaI[c], etaT[c], k[c]);
}

float gnd_ref_roughness_to_alpha(float r) { return GGX::roughnessToAlpha(r); }

// which: 0 D(w1), 1 G1(w1), 2 G(w1,w2), 3 pdf(w1=wo, w2=wh)
float gnd_ref_ggx(float alpha, int which, const float* w1, const float* w2) {
    GGX g(alpha);
    V3 a(w1[0], w1[1], w1[2]), b(w2 ? w2[0] : 0, w2 ? w2[1] : 0, w2 ? w2[2] : 0);
    switch (which) {
        case 0: return g.D(a);
        case 1: return g.G1(a);
        case 2: return g.G(a, b);
        default: return g.pdf(a, b);
    }
}
void gnd_ref_ggx_sample(float alpha, const float* wo, const float* u, float* wh) {
    GGX g(alpha);
    V3 r = g.sample_wh(V3(wo[0], wo[1], wo[2]), P2{u[0], u[1]});
    wh[0] = r.x; wh[1] = r.y; wh[2] = r.z;
}

static Lobe lobeFromDesc(const gnd_material_desc* m) {
    Lobe L;
    switch (m->type) {
        case GND_MAT_MATTE: L.kind = LAMBERT; L.R = C3(m->albedo[0], m->albedo[1], m->albedo[2]); break;
        case GND_MAT_CONDUCTOR:
            L.R = C3(m->R[0], m->R[1], m->R[2]);
            L.eta3 = C3(m->eta_rgb[0], m->eta_rgb[1], m->eta_rgb[2]);
            L.k3 = C3(m->k_rgb[0], m->k_rgb[1], m->k_rgb[2]);
            if (m->roughness == 0.0f) L.kind = SPEC_REFL;
            else { L.kind = MF_REFL; L.ggx = GGX(GGX::roughnessToAlpha(m->roughness)); }
            break;
        case GND_MAT_GLASS:
            L.R = C3(m->R[0], m->R[1], m->R[2]);
            L.T = C3(m->T[0], m->T[1], m->T[2]);
            L.etaExt = 1.0f;
            L.etaInt = m->eta;
            if (m->roughness == 0.0f) L.kind = FRESNEL_SPEC;
            else { L.kind = MF_FRESNEL; L.ggx = GGX(GGX::roughnessToAlpha(m->roughness)); }
            break;
        default:
            L.Kd = C3(m->Kd[0], m->Kd[1], m->Kd[2]);
            L.Ks = C3(m->Ks[0], m->Ks[1], m->Ks[2]);
            L.eta = m->eta;
            if (m->roughness == 0.0f) L.kind = SMOOTH_PLASTIC;
            else { L.kind = ROUGH_PLASTIC; L.ggx = GGX(GGX::roughnessToAlpha(m->roughness)); }
    }
    return L;
}
// Local-frame lobe probes: f (out rgb), pdf, sample (out rgb f, wi, pdf, type)
void gnd_ref_lobe_f(const gnd_material_desc* m, const float* wo, const float* wi, float* out) {
    C3 f = lobeFromDesc(m).f(V3(wo[0], wo[1], wo[2]), V3(wi[0], wi[1], wi[2]));
    out[0] = f.r; out[1] = f.g; out[2] = f.b;
}
float gnd_ref_lobe_pdf(const gnd_material_desc* m, const float* wo, const float* wi) {
    return lobeFromDesc(m).pdf(V3(wo[0], wo[1], wo[2]), V3(wi[0], wi[1], wi[2]));
}
int gnd_ref_lobe_sample(const gnd_material_desc* m, const float* wo, const float* u, float uc, float* f,
                        float* wi, float* pdf) {
    V3 w(0, 0, 0);
    float p = 0.0f;
    int st = 0;
    C3 r = lobeFromDesc(m).sample(V3(wo[0], wo[1], wo[2]), w, P2{u[0], u[1]}, uc, p, st);
    f[0] = r.r; f[1] = r.g; f[2] = r.b;
    wi[0] = w.x; wi[1] = w.y; wi[2] = w.z;
    *pdf = p;
    return st;
}
// BSDF::sample in a local frame (identity frame): returns f*|cos|/pdf weight
int gnd_ref_bsdf_sample_weight(const gnd_material_desc* m, const float* wo, const float* u, float uc,
                               float* weight, float* wi, float* pdf) {
    BSDF b;
    b.lobe = lobeFromDesc(m);
    V3 w;
    int st = 0;
    float p = 0.0f;
    C3 r = b.sample(V3(wo[0], wo[1], wo[2]), &w, P2{u[0], u[1]}, u[2], &p, &st);
    (void)uc;
    weight[0] = r.r; weight[1] = r.g; weight[2] = r.b;
    wi[0] = w.x; wi[1] = w.y; wi[2] = w.z;
    *pdf = p;
    return st;
}

void gnd_ref_warp(int which, const float* u, float* out) {
    P2 s{u[0], u[1]};
    switch (which) {
        case 0: { P2 d = squareToUniformDisk(s); out[0] = d.x; out[1] = d.y; out[2] = 0; break; }
        case 1: { V3 v = squareToCosineHemisphere(s); out[0] = v.x; out[1] = v.y; out[2] = v.z; break; }
        case 2: { V3 v = squareToUniformSphere(s); out[0] = v.x; out[1] = v.y; out[2] = v.z; break; }
        default: { P2 d = squareToUniformPolygon(s, which); out[0] = d.x; out[1] = d.y; out[2] = 0; }
    }
}
float gnd_ref_cosine_hemisphere_pdf(const float* v) { return squareToCosineHemispherePdf(V3(v[0], v[1], v[2])); }

void gnd_ref_frame(const float* n, float* xyz) {
    Frame f = Frame::fromNormal(V3(n[0], n[1], n[2]));
    xyz[0] = f.x.x; xyz[1] = f.x.y; xyz[2] = f.x.z;
    xyz[3] = f.y.x; xyz[4] = f.y.y; xyz[5] = f.y.z;
    xyz[6] = f.z.x; xyz[7] = f.z.y; xyz[8] = f.z.z;
}

float gnd_ref_hg_pdf(float g, const float* wo, const float* wi) {
    return hgPhase(Dot(V3(wo[0], wo[1], wo[2]), V3(wi[0], wi[1], wi[2])), hgClamp(g));
}
float gnd_ref_hg_sample(float g, const float* wo, const float* u, float* wi) {
    V3 w;
    float p = hgSample(hgClamp(g), V3(wo[0], wo[1], wo[2]), &w, P2{u[0], u[1]});
    wi[0] = w.x; wi[1] = w.y; wi[2] = w.z;
    return p;
}

float gnd_ref_luminance(float r, float g, float b) { return C3(r, g, b).luminance(); }

int gnd_ref_bounds_intersect(const float* bmin, const float* bmax, const float* o, const float* d, float tmax,
                             float* t01) {
    Bounds b;
    b.pMin = V3(bmin[0], bmin[1], bmin[2]);
    b.pMax = V3(bmax[0], bmax[1], bmax[2]);
    return boundsRayIntersect(b, V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]), tmax, &t01[0], &t01[1]) ? 1 : 0;
}

// Camera ray for a given camera sample (pFilm in [0,1]^2 with y up, pLens,
// lambdaOffset).  out: o[3], d[3], weight.
void gnd_ref_camera_ray(void* h, const float* pfilm, const float* plens, float lambda_offset, float time,
                        float* out) {
    Scene* s = static_cast<Scene*>(h);
    CamSample cs;
    cs.pFilm = P2{pfilm[0], pfilm[1]};
    cs.pLens = P2{plens[0], plens[1]};
    cs.lambdaOffset = lambda_offset;
    cs.time = time;
    Ray r;
    float w = shootRay(s->camera, cs, r);
    out[0] = r.o.x; out[1] = r.o.y; out[2] = r.o.z;
    out[3] = r.d.x; out[4] = r.d.y; out[5] = r.d.z;
    out[6] = w;
    out[7] = r.time;
}

// Surface interaction of a closest hit: out p[3], n[3], dpdu[3], uv[2]
int gnd_ref_surface(void* h, const float* o, const float* d, float* out) {
    Scene* s = static_cast<Scene*>(h);
    Ray r;
    r.o = V3(o[0], o[1], o[2]);
    r.d = V3(d[0], d[1], d[2]);
    Hit hh;
    if (!intersectClosest(*s, r, hh, nullptr)) return -1;
    Isect it;
    fillInteraction(*s, r, hh, it);
    out[0] = it.p.x; out[1] = it.p.y; out[2] = it.p.z;
    out[3] = it.n.x; out[4] = it.n.y; out[5] = it.n.z;
    out[6] = it.dpdu.x; out[7] = it.dpdu.y; out[8] = it.dpdu.z;
    out[9] = it.uv.x; out[10] = it.uv.y;
    return it.prim;
}

// Medium probes on a world-space segment: Tr (which=0) or sample (which=1).
// out: rgb result, then for sample: valid flag, p[3]; draws = sampler draws.
int gnd_ref_medium(void* h, int medium, int which, const float* o, const float* d, float tmax, uint64_t seed,
                   float* out) {
    Scene* s = static_cast<Scene*>(h);
    Ray r;
    r.o = V3(o[0], o[1], o[2]);
    r.d = V3(d[0], d[1], d[2]);
    r.tMax = tmax;
    Sampler smp;
    smp.seed(seed);
    if (which == 0) {
        C3 t = mediumTr(s->media[medium], r, smp);
        out[0] = t.r; out[1] = t.g; out[2] = t.b;
    } else {
        MediumEvent mi;
        C3 t = mediumSample(s->media[medium], r, smp, mi);
        out[0] = t.r; out[1] = t.g; out[2] = t.b;
        out[3] = mi.valid ? 1.0f : 0.0f;
        out[4] = mi.p.x; out[5] = mi.p.y; out[6] = mi.p.z;
    }
    return (int)smp.draws;
}

float gnd_ref_grid_sample(void* h, int medium, double x, double y, double z) {
    Scene* s = static_cast<Scene*>(h);
    return gridSample(s->media[medium], x, y, z);
}

// One NEE sample of light `k` from refP with the 2D sample u: out = sampled
// point (3), pdf, returned radiance Li (3), shadow direction (3), tMax.
void gnd_ref_light_sample(void* h, int k, const float* refP, const float* u, float* out) {
    Scene* s = static_cast<Scene*>(h);
    LightSample ls;
    C3 Li = lightSample(s->lights[k], V3(refP[0], refP[1], refP[2]), P2{u[0], u[1]}, ls, 0.0f);
    const float v[11] = {ls.p.x, ls.p.y, ls.p.z, ls.pdf, Li.r, Li.g, Li.b,
                         ls.shadow.d.x, ls.shadow.d.y, ls.shadow.d.z, ls.shadow.tMax};
    for (int i = 0; i < 11; ++i) out[i] = v[i];
}
// AreaLight::sample of primitive `prim` from refP at time 0: out p[3], n[3], pdf
int gnd_ref_area_sample(void* h, int prim, const float* refP, const float* u, float* out) {
    Scene* s = static_cast<Scene*>(h);
    LightSample ls;
    bool pdfSet = false;
    areaSample(*s, s->prims[prim], V3(refP[0], refP[1], refP[2]), P2{u[0], u[1]}, ls, pdfSet, 0.0f);
    const float v[7] = {ls.p.x, ls.p.y, ls.p.z, ls.n.x, ls.n.y, ls.n.z, pdfSet ? ls.pdf : 0.0f};
    for (int i = 0; i < 7; ++i) out[i] = v[i];
    return pdfSet ? 1 : 0;
}
// EnvironmentMap::pdf towards direction dir from the origin, and eval along it
float gnd_ref_env_pdf(void* h, const float* dir, float* Le) {
    Scene* s = static_cast<Scene*>(h);
    const Light& L = s->lights[s->envLight];
    V3 w(dir[0], dir[1], dir[2]);
    C3 e = envEval(L, w);
    if (Le) { Le[0] = e.r; Le[1] = e.g; Le[2] = e.b; }
    return envPdf(L, V3(0, 0, 0), w * 1e5f);
}

// Single-path radiance for (pixel, pass): out rgb, then (if aov is not null)
// the path's AOV record: albedo rgb, normal xyz, depth, roughness, metallic
void gnd_ref_path_sample(void* h, int x, int y, int pass, uint64_t seed_offset, float* out, float* aov) {
    Scene* s = static_cast<Scene*>(h);
    int W = s->camera.d.width, H = s->camera.d.height;
    Sampler smp;
    smp.seed((uint64_t)(y * W + x) + (uint64_t)pass * (uint64_t)(W * H) + seed_offset);
    CamSample cs;
    P2 ps = smp.next2D();
    cs.pFilm.x = (x + ps.x) / W;
    cs.pFilm.y = 1.0f - (y + ps.y) / H;
    cs.pLens = smp.next2D();
    smp.next1D();
    cs.time = smp.next1D();
    Ray ray;
    float w = shootRay(s->camera, cs, ray);
    AOV a;
    AOV* ap = aov ? &a : nullptr;
    C3 L = integrate(*s, ray, smp, nullptr, ap);
    L = L * w;
    if (L.hasNaNs()) L = C3(0.0f);
    out[0] = L.r; out[1] = L.g; out[2] = L.b;
    if (aov) {
        const float v[kAovFloats] = {a.albedo.r, a.albedo.g, a.albedo.b, a.normal.x, a.normal.y,
                                     a.normal.z, a.depth,    a.roughness, a.metallic};
        for (int k = 0; k < kAovFloats; ++k) aov[k] = v[k];
    }
}

}  // extern "C"
