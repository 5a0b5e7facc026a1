// dev_shade.h — surface interaction, BSDFs, area-light sampling, media and
// camera rays on the device.  Each function follows the reference routine it
// names; transcendentals come from the ROCm device math library.
#pragma once

#include "dev_common.h"

namespace gdev {

// ------------------------------------------------------------------ PCG32 / sampler
// PCG32 (include/gianduia/core/pcg32.h:5-32) with initseq = 1 (inc = 3);
// IndependentSampler::seed(idx) = PCG32.seed(splitMix64(idx)) (src/samplers/independent.cpp:22-44)
GDEV_INLINE uint64_t splitMix64(uint64_t z) {
    z = z + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
constexpr uint64_t kPcgMul = 6364136223846793005ULL;
constexpr uint64_t kPcgInc = 3ULL;
GDEV_INLINE uint64_t pcgSeed(uint64_t initstate) {
    uint64_t s = 0u;
    s = s * kPcgMul + kPcgInc;
    s += initstate;
    s = s * kPcgMul + kPcgInc;
    return s;
}
struct Rng {
    uint64_t s;
    GDEV_INLINE uint32_t nextUInt() {
        uint64_t old = s;
        s = old * kPcgMul + kPcgInc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31));
    }
    // ldexp((float)u, -32): round to float, then an exact power-of-two scale
    GDEV_INLINE float next1D() { return (float)nextUInt() * 2.3283064365386963e-10f; }
};

// ------------------------------------------------------------------ warps (src/math/warp.cpp)
GDEV_INLINE void squareToUniformDisk(float sx, float sy, float& ox, float& oy) {  // warp.cpp:7-25
    float px = 2.0f * sx - 1.0f, py = 2.0f * sy - 1.0f;
    if (__builtin_fabsf(px) < kEps && __builtin_fabsf(py) < kEps) { ox = 0.0f; oy = 0.0f; return; }
    float r, theta;
    if (__builtin_fabsf(px) > __builtin_fabsf(py)) {
        r = px;
        theta = (float)((kPiD / 4.0f) * (double)(py / px));
    } else {
        r = py;
        theta = (float)((kPiD / 2.0f) - (kPiD / 4.0f) * (double)(px / py));
    }
    ox = r * cosf(theta);
    oy = r * sinf(theta);
}
GDEV_INLINE V3 squareToCosineHemisphere(float sx, float sy) {  // warp.cpp:68-77
    float dx, dy;
    squareToUniformDisk(sx, sy, dx, dy);
    float r = dx * dx + dy * dy;
    return V3{dx, dy, __builtin_sqrtf(1 - r)};
}
GDEV_INLINE float cosineHemispherePdf(V3 v) {  // warp.cpp:79-84
    if (v.z <= 0.0f || __builtin_fabsf(1 - lengthSq(v)) > kEps) return 0.0f;
    return kInvPi * v.z;
}
GDEV_INLINE V3 squareToUniformSphere(float sx, float sy) {  // warp.cpp:34-43, cosThetaMax = -1
    float theta = acosf(1 - (1 - (-1.0f)) * sx);
    float phi = (float)(2.0f * kPiD * (double)sy);
    float st = sinf(theta);
    return V3{st * cosf(phi), st * sinf(phi), cosf(theta)};
}
GDEV_INLINE void squareToUniformPolygon(float sx, float sy, int blades, float& ox, float& oy) {  // warp.cpp:86-113
    if (blades < 3) { squareToUniformDisk(sx, sy, ox, oy); return; }
    float bf = sx * blades;
    int bi = (int)bf < blades - 1 ? (int)bf : blades - 1;
    float u0 = bf - bi, u1 = sy;
    float dt = 2.0f * kPi / blades;
    float th1 = bi * dt + kPi / 2.0f, th2 = (bi + 1) * dt + kPi / 2.0f;
    float r = __builtin_sqrtf(u0);
    ox = r * ((1.0f - u1) * cosf(th1) + u1 * cosf(th2));
    oy = r * ((1.0f - u1) * sinf(th1) + u1 * sinf(th2));
}

// ------------------------------------------------------------------ frame (math/frame.h)
struct Frame {
    V3 x, y, z;
    GDEV_INLINE V3 toLocal(V3 v) const { return V3{dot(v, x), dot(v, y), dot(v, z)}; }
    GDEV_INLINE V3 toWorld(V3 v) const { return x * v.x + y * v.y + z * v.z; }
};
GDEV_INLINE Frame frameFromNormal(V3 n) {  // frame.h:17-25
    Frame f;
    f.z = n;
    float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    f.x = V3{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    f.y = V3{b, sign + n.y * n.y * a, -n.y};
    return f;
}

// ------------------------------------------------------------------ fresnel (materials/fresnel.h)
GDEV_INLINE float frDielectric(float cosI, float etaExt, float etaInt) {  // fresnel.h:13-40
    if (etaExt == etaInt) return 0.0f;
    cosI = clampf(cosI, -1.f, 1.f);
    if (cosI < 0.f) { float t = etaExt; etaExt = etaInt; etaInt = t; cosI = -cosI; }
    float sinI = __builtin_sqrtf(stdMax(0.f, 1.f - cosI * cosI));
    float sinT = etaExt / etaInt * sinI;
    if (sinT >= 1.f) return 1.f;
    float cosT = __builtin_sqrtf(stdMax(0.f, 1.f - sinT * sinT));
    float rParl = ((etaInt * cosI) - (etaExt * cosT)) / ((etaInt * cosI) + (etaExt * cosT));
    float rPerp = ((etaExt * cosI) - (etaInt * cosT)) / ((etaExt * cosI) + (etaInt * cosT));
    return (rParl * rParl + rPerp * rPerp) / 2.f;
}
GDEV_INLINE bool refract(V3 wi, V3 n, float eta, V3& wt) {  // fresnel.h:47-58
    float cosI = dot(n, wi);
    float sin2I = stdMax(0.f, 1.f - cosI * cosI);
    float sin2T = eta * eta * sin2I;
    if (sin2T >= 1.f) return false;
    float cosT = __builtin_sqrtf(1.f - sin2T);
    wt = (-wi) * eta + n * (eta * cosI - cosT);
    return true;
}
GDEV_INLINE float conductorChannel(float cosI, float eta, float k) {  // fresnel.h:111-133, etaI = 1
    float cosI2 = cosI * cosI;
    float sinI2 = stdMax(0.0f, 1.f - cosI2);
    float e = eta / 1.0f, ek = k / 1.0f;
    float eta2 = e * e, etak2 = ek * ek;
    float t0 = eta2 - etak2 - sinI2;
    float a2b2 = __builtin_sqrtf(t0 * t0 + 4.f * eta2 * etak2);
    float t1 = a2b2 + cosI2;
    float a = __builtin_sqrtf(stdMax(0.0f, 0.5f * (a2b2 + t0)));
    float t2 = 2.f * __builtin_fabsf(cosI) * a;
    float Rs = (t1 - t2) / (t1 + t2);
    float t3 = cosI2 * a2b2 + sinI2 * sinI2;
    float t4 = t2 * sinI2;
    float Rp = Rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (Rp + Rs);
}

// ------------------------------------------------------------------ GGX (src/math/microfacet.cpp)
GDEV_INLINE float ggxD(float a, V3 wh) {
    if (__builtin_fabsf(wh.z) < kEps) return 0.0f;
    float d = (wh.x * wh.x) / (a * a) + (wh.y * wh.y) / (a * a) + (wh.z * wh.z);
    return 1.0f / (kPi * a * a * d * d);
}
GDEV_INLINE float ggxG1(float a, V3 w) {
    if (__builtin_fabsf(w.z) < kEps) return 0.0f;
    float a2tan2 = (w.x * w.x * a * a + w.y * w.y * a * a) / (w.z * w.z);
    float lambda = 0.5f * (-1.0f + __builtin_sqrtf(1.0f + a2tan2));
    return 1.0f / (1.0f + lambda);
}
GDEV_INLINE float ggxG(float a, V3 wo, V3 wi) { return ggxG1(a, wo) * ggxG1(a, wi); }
GDEV_INLINE V3 ggxSample(float a, V3 wo, float sx, float sy) {  // microfacet.cpp:30-56
    bool flip = wo.z < 0.0f;
    V3 woh = flip ? -wo : wo;
    V3 Vh = normalize(V3{a * woh.x, a * woh.y, woh.z});
    float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    V3 T1 = lensq > 0.0f ? divs(V3{-Vh.y, Vh.x, 0.0f}, __builtin_sqrtf(lensq)) : V3{1.0f, 0.0f, 0.0f};
    V3 T2 = cross(Vh, T1);
    float r = __builtin_sqrtf(sx);
    float phi = 2.0f * kPi * sy;
    float t1 = r * cosf(phi);
    float t2 = r * sinf(phi);
    float s = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - s) * __builtin_sqrtf(stdMax(0.0f, 1.0f - t1 * t1)) + s * t2;
    V3 Nh = T1 * t1 + T2 * t2 + Vh * __builtin_sqrtf(stdMax(0.0f, 1.0f - t1 * t1 - t2 * t2));
    V3 wh = normalize(V3{a * Nh.x, a * Nh.y, stdMax(0.0f, Nh.z)});
    return flip ? -wh : wh;
}
GDEV_INLINE float ggxPdf(float a, V3 wo, V3 wh) {  // microfacet.cpp:58-66
    if (wo.z * wh.z < 0.0f) return 0.0f;
    float cosO = __builtin_fabsf(wo.z);
    if (cosO == 0.0f) return 0.0f;
    return ggxD(a, wh) * ggxG1(a, wo) * __builtin_fabsf(dot(wo, wh)) / cosO;
}

// ------------------------------------------------------------------ BxDFs (src/materials/reflection.cpp)
enum : int { BSDF_REFLECTION = 1, BSDF_TRANSMISSION = 2, BSDF_DIFFUSE = 4, BSDF_GLOSSY = 8, BSDF_SPECULAR = 16 };
enum : int { L_LAMBERT, L_SPEC_REFL, L_FRESNEL_SPEC, L_MF_REFL, L_MF_FRESNEL, L_SMOOTH_PLASTIC, L_ROUGH_PLASTIC };

struct Lobe {
    int kind;
    const DMaterial* m;
    V3 albedo;  // matte albedo (constant or checkerboard)
};

GDEV_INLINE V3 c3(const float* f) { return V3{f[0], f[1], f[2]}; }
GDEV_INLINE V3 conductorF(const DMaterial& m, float cosI) {
    return V3{conductorChannel(cosI, m.etaRGB[0], m.k[0]), conductorChannel(cosI, m.etaRGB[1], m.k[1]),
              conductorChannel(cosI, m.etaRGB[2], m.k[2])};
}
GDEV_INLINE V3 sameV(float v) { return V3{v, v, v}; }

GDEV_INLINE V3 lobeF(const Lobe& L, V3 wo, V3 wi) {
    const DMaterial& m = *L.m;
    switch (L.kind) {
        case L_LAMBERT: return L.albedo * kInvPi;
        case L_SPEC_REFL:
        case L_FRESNEL_SPEC: return sameV(0.0f);
        case L_MF_REFL: {  // reflection.cpp:103-121
            if (wo.z * wi.z <= 0.0f) return sameV(0.0f);
            float cO = __builtin_fabsf(wo.z), cI = __builtin_fabsf(wi.z);
            if (cO == 0.0f || cI == 0.0f) return sameV(0.0f);
            V3 wh = wo + wi;
            if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) return sameV(0.0f);
            wh = normalize(wh);
            if (wh.z < 0.0f) wh = -wh;
            if (dot(wo, wh) * wo.z < 0.0f || dot(wi, wh) * wi.z < 0.0f) return sameV(0.0f);
            V3 F = conductorF(m, dot(wi, wh));
            float D = ggxD(m.alpha, wh), G = ggxG(m.alpha, wo, wi);
            return cdiv(c3(m.R) * D * G * F, 4.0f * cI * cO);
        }
        case L_MF_FRESNEL: {  // reflection.cpp:151-191
            bool reflect = wo.z * wi.z > 0.0f;
            bool entering = wo.z > 0.0f;
            float etaI = entering ? 1.0f : m.eta, etaT = entering ? m.eta : 1.0f;
            V3 wh = reflect ? normalize(wo + wi) : normalize(wo * etaI + wi * etaT);
            if (wh.z * wo.z < 0.0f) wh = -wh;
            float dO = dot(wo, wh), dI = dot(wi, wh);
            if (reflect) { if (dO <= 0.0f || dI <= 0.0f) return sameV(0.0f); }
            else { if (dO <= 0.0f || dI >= 0.0f) return sameV(0.0f); }
            float F = frDielectric(dO, etaI, etaT);
            float D = ggxD(m.alpha, wh), G = ggxG(m.alpha, wo, wi);
            if (reflect) return cdiv(c3(m.R) * D * G * F, 4.0f * __builtin_fabsf(wo.z * wi.z));
            if (dO * dI > 0.0f) return sameV(0.0f);
            float denom = (etaI * dO + etaT * dI);
            denom *= denom;
            float factor = __builtin_fabsf(dO * dI) / __builtin_fabsf(wo.z * wi.z);
            return cdiv(c3(m.T) * factor * (etaI * etaI) * D * G * (1.0f - F), denom);
        }
        case L_SMOOTH_PLASTIC: {  // reflection.cpp:258-264
            if (wo.z < kEps || wi.z < kEps) return sameV(0.0f);
            float Fo = frDielectric(wo.z, 1.0f, m.eta), Fi = frDielectric(wi.z, 1.0f, m.eta);
            return c3(m.Kd) * (float)0.318309886183790671538 * (1.0f - Fo) * (1.0f - Fi);
        }
        default: {  // L_ROUGH_PLASTIC, reflection.cpp:295-313
            if (wo.z < kEps || wi.z < kEps) return sameV(0.0f);
            float Fo = frDielectric(wo.z, 1.0f, m.eta), Fi = frDielectric(wi.z, 1.0f, m.eta);
            V3 diffuse = c3(m.Kd) * (float)0.318309886183790671538 * (1.0f - Fo) * (1.0f - Fi);
            V3 wh = normalize(wo + wi);
            if (wh.z < 0.0f) wh = -wh;
            if (dot(wo, wh) * wo.z < 0.0f || dot(wi, wh) * wi.z < 0.0f) return sameV(0.0f);
            float Fh = frDielectric(dot(wo, wh), 1.0f, m.eta);
            float D = ggxD(m.alpha, wh), G = ggxG(m.alpha, wo, wi);
            V3 spec = cdiv(c3(m.Ks) * D * G * Fh, 4.0f * wo.z * wi.z);
            return diffuse + spec;
        }
    }
}

GDEV_INLINE float lobePdf(const Lobe& L, V3 wo, V3 wi) {
    const DMaterial& m = *L.m;
    switch (L.kind) {
        case L_LAMBERT: return cosineHemispherePdf(wi);
        case L_SPEC_REFL:
        case L_FRESNEL_SPEC: return 0.0f;
        case L_MF_REFL: {  // reflection.cpp:140-146
            if (wo.z * wi.z <= 0.0f) return 0.0f;
            V3 wh = normalize(wo + wi);
            if (wh.z < 0.0f) wh = -wh;
            return ggxPdf(m.alpha, wo, wh) / (4.0f * dot(wo, wh));
        }
        case L_MF_FRESNEL: {  // reflection.cpp:227-253
            bool reflect = wo.z * wi.z > 0.0f;
            bool entering = wo.z > 0.0f;
            float etaI = entering ? 1.0f : m.eta, etaT = entering ? m.eta : 1.0f;
            V3 wh = reflect ? normalize(wo + wi) : normalize(wo * etaI + wi * etaT);
            if (wh.z * wo.z < 0.0f) wh = -wh;
            float dO = dot(wo, wh), dI = dot(wi, wh);
            float F = frDielectric(dO, etaI, etaT);
            float pwh = ggxPdf(m.alpha, wo, wh);
            if (reflect) return F * pwh / (4.0f * __builtin_fabsf(dO));
            if (dO * dI > 0.0f) return 0.0f;
            float denom = (etaI * dO + etaT * dI);
            float dwh = (etaT * etaT * __builtin_fabsf(dI)) / (denom * denom);
            return (1.0f - F) * pwh * dwh;
        }
        case L_SMOOTH_PLASTIC: {
            if (wo.z < kEps || wi.z < kEps) return 0.0f;
            return (1.0f - frDielectric(wo.z, 1.0f, m.eta)) * cosineHemispherePdf(wi);
        }
        default: {  // reflection.cpp:337-348
            if (wo.z < kEps || wi.z < kEps) return 0.0f;
            V3 wh = normalize(wo + wi);
            if (wh.z < 0.0f) wh = -wh;
            float ps = frDielectric(wo.z, 1.0f, m.eta);
            float pdfSpec = ggxPdf(m.alpha, wo, wh) / (4.0f * dot(wo, wh));
            return ps * pdfSpec + (1.0f - ps) * cosineHemispherePdf(wi);
        }
    }
}

GDEV_INLINE int lobeType(int kind) {
    switch (kind) {
        case L_LAMBERT: return BSDF_REFLECTION | BSDF_DIFFUSE;
        case L_SPEC_REFL: return BSDF_REFLECTION | BSDF_SPECULAR;
        case L_FRESNEL_SPEC: return BSDF_REFLECTION | BSDF_TRANSMISSION | BSDF_SPECULAR;
        case L_MF_REFL: return BSDF_REFLECTION | BSDF_GLOSSY;
        case L_MF_FRESNEL: return BSDF_REFLECTION | BSDF_TRANSMISSION | BSDF_GLOSSY;
        case L_SMOOTH_PLASTIC: return BSDF_REFLECTION | BSDF_DIFFUSE | BSDF_SPECULAR;
        default: return BSDF_REFLECTION | BSDF_DIFFUSE | BSDF_GLOSSY;
    }
}

// BxDF::sample; leaves wi/pdf/type untouched on the paths where the reference does
GDEV_INLINE V3 lobeSample(const Lobe& L, V3 wo, V3& wi, float sx, float sy, float uc, float& pdf, int& st) {
    const DMaterial& m = *L.m;
    switch (L.kind) {
        case L_LAMBERT:  // reflection.cpp:12-27
            if (wo.z <= 0.0f) { pdf = 0.0f; return sameV(0.0f); }
            wi = squareToCosineHemisphere(sx, sy);
            pdf = cosineHemispherePdf(wi);
            st = lobeType(L_LAMBERT);
            return L.albedo * kInvPi;
        case L_SPEC_REFL:  // reflection.cpp:40-48
            wi = V3{-wo.x, -wo.y, wo.z};
            pdf = 1.0f;
            st = lobeType(L_SPEC_REFL);
            return c3(m.R) * conductorF(m, wo.z);
        case L_FRESNEL_SPEC: {  // reflection.cpp:61-94
            float F = frDielectric(wo.z, 1.0f, m.eta);
            if (sx < F) {
                wi = V3{-wo.x, -wo.y, wo.z};
                st = BSDF_SPECULAR | BSDF_REFLECTION;
                pdf = F;
                return c3(m.R) * F;
            }
            bool entering = wo.z > 0;
            float etaI = entering ? 1.0f : m.eta, etaT = entering ? m.eta : 1.0f;
            V3 n = entering ? V3{0, 0, 1} : V3{0, 0, -1};
            V3 wr;
            if (!refract(wo, n, etaI / etaT, wr)) return sameV(0.0f);
            wi = wr;
            st = BSDF_SPECULAR | BSDF_TRANSMISSION;
            pdf = 1.f - F;
            float er = etaI / etaT;
            return c3(m.T) * er * er * (1.0f - F);
        }
        case L_MF_REFL: {  // reflection.cpp:123-138
            if (__builtin_fabsf(wo.z) < kEps) { pdf = 0.0f; return sameV(0.0f); }
            V3 wh = ggxSample(m.alpha, wo, sx, sy);
            float dO = dot(wo, wh);
            if (dO <= 0.0f) { pdf = 0.0f; return sameV(0.0f); }
            wi = -wo + wh * 2.0f * dO;
            if (wo.z * wi.z <= 0.0f) { pdf = 0.0f; return sameV(0.0f); }
            pdf = lobePdf(L, wo, wi);
            st = lobeType(L_MF_REFL);
            return lobeF(L, wo, wi);
        }
        case L_MF_FRESNEL: {  // reflection.cpp:193-225
            if (__builtin_fabsf(wo.z) < kEps) return sameV(0.0f);
            bool entering = wo.z > 0.0f;
            float etaI = entering ? 1.0f : m.eta, etaT = entering ? m.eta : 1.0f;
            V3 wh = ggxSample(m.alpha, wo, sx, sy);
            float dO = dot(wo, wh);
            if (dO <= 0.0f) { pdf = 0.0f; return sameV(0.0f); }
            float F = frDielectric(dO, etaI, etaT);
            if (uc < F) {
                wi = -wo + wh * 2.0f * dO;
                if (wo.z * wi.z <= 0.0f) { pdf = 0.0f; return sameV(0.0f); }
                st = BSDF_REFLECTION | BSDF_GLOSSY;
            } else {
                if (!refract(wo, wh, etaI / etaT, wi)) { pdf = 0.0f; return sameV(0.0f); }
                if (wo.z * wi.z > 0.0f) { pdf = 0.0f; return sameV(0.0f); }
                st = BSDF_TRANSMISSION | BSDF_GLOSSY;
            }
            pdf = lobePdf(L, wo, wi);
            return lobeF(L, wo, wi);
        }
        case L_SMOOTH_PLASTIC: {  // reflection.cpp:266-284
            if (wo.z < kEps) { pdf = 0.0f; return sameV(0.0f); }
            float ps = frDielectric(wo.z, 1.0f, m.eta);
            if (uc < ps) {
                wi = V3{-wo.x, -wo.y, wo.z};
                pdf = ps;
                st = BSDF_REFLECTION | BSDF_SPECULAR;
                return c3(m.Ks);
            }
            wi = squareToCosineHemisphere(sx, sy);
            if (wi.z < kEps) { pdf = 0.0f; return sameV(0.0f); }
            pdf = (1.0f - ps) * cosineHemispherePdf(wi);
            st = BSDF_REFLECTION | BSDF_DIFFUSE;
            return lobeF(L, wo, wi);
        }
        default: {  // L_ROUGH_PLASTIC, reflection.cpp:315-335
            if (wo.z < kEps) { pdf = 0.0f; return sameV(0.0f); }
            float ps = frDielectric(wo.z, 1.0f, m.eta);
            if (uc < ps) {
                V3 wh = ggxSample(m.alpha, wo, sx, sy);
                float dO = dot(wo, wh);
                if (dO <= 0.0f) { pdf = 0.0f; return sameV(0.0f); }
                wi = -wo + wh * 2.0f * dO;
                if (wi.z < kEps) { pdf = 0.0f; return sameV(0.0f); }
                st = BSDF_REFLECTION | BSDF_GLOSSY;
            } else {
                wi = squareToCosineHemisphere(sx, sy);
                if (wi.z < kEps) { pdf = 0.0f; return sameV(0.0f); }
                st = BSDF_REFLECTION | BSDF_DIFFUSE;
            }
            pdf = lobePdf(L, wo, wi);
            return lobeF(L, wo, wi);
        }
    }
}

// Material -> lobe (matte.cpp:46-53, conductor.cpp:101-120, glass.cpp:98-115,
// plastic.cpp:74-91); returns false for the index (null) material.
GDEV_INLINE bool makeLobe(const DMaterial* m, float u, float v, Lobe& L) {
    L.m = m;
    switch (m->type) {
        case GND_MAT_MATTE:
            L.kind = L_LAMBERT;
            if (m->checker) {  // checkerboard_color fallback (checkerboard.h), scale 0.1
                int si = (int)floorf(u / 0.1f), ti = (int)floorf(v / 0.1f);
                L.albedo = ((si + ti) % 2 == 0) ? sameV(0.0f) : V3{1.0f, 0.0f, 1.0f};
            } else L.albedo = c3(m->albedo);
            return true;
        case GND_MAT_CONDUCTOR: L.kind = m->roughness == 0.0f ? L_SPEC_REFL : L_MF_REFL; return true;
        case GND_MAT_GLASS: L.kind = m->roughness == 0.0f ? L_FRESNEL_SPEC : L_MF_FRESNEL; return true;
        case GND_MAT_PLASTIC: L.kind = m->roughness == 0.0f ? L_SMOOTH_PLASTIC : L_ROUGH_PLASTIC; return true;
        default: return false;
    }
}

// BSDF with its single lobe (src/materials/bsdf.cpp:6-136)
struct Bsdf {
    Frame fr;
    Lobe lobe;
    GDEV_INLINE V3 f(V3 woW, V3 wiW) const {
        V3 wo = fr.toLocal(woW), wi = fr.toLocal(wiW);
        if (wo.z == 0.0f) return sameV(0.0f);
        return sameV(0.0f) + lobeF(lobe, wo, wi);
    }
    GDEV_INLINE float pdf(V3 woW, V3 wiW) const {
        V3 wo = fr.toLocal(woW), wi = fr.toLocal(wiW);
        if (wo.z == 0.0f) return 0.0f;
        return 0.0f + 1.0f * lobePdf(lobe, wo, wi);
    }
    // returns the f*|cos|/pdf weight; pdfOut untouched when wo.z == 0 (bsdf.cpp:71-72)
    GDEV_INLINE V3 sample(V3 woW, V3& wiW, float sx, float sy, float uc, float& pdfOut, int& st) const {
        V3 wo = fr.toLocal(woW);
        if (wo.z == 0.0f) return sameV(0.0f);
        V3 wi = V3{0, 0, 0};
        float lp = 0.0f;
        int t = 0;
        V3 fv = lobeSample(lobe, wo, wi, sx, sy, uc, lp, t);
        wiW = fr.toWorld(wi);
        st = t;
        if (lp == 0.0f || isBlack(fv)) { pdfOut = 0.0f; return sameV(0.0f); }
        if (t & BSDF_SPECULAR) {
            pdfOut = lp * 1.0f;
            return cdiv(fv, pdfOut);
        }
        float tp = lp * 1.0f;
        pdfOut = tp;
        if (tp < kEps) return sameV(0.0f);
        return cdiv(fv * __builtin_fabsf(wi.z), tp);
    }
};

// ------------------------------------------------------------------ surface interaction
struct Isect {
    V3 p, n, dpdu;
    float u, v;
    int prim;
};

GDEV_INLINE V3 dpduFallback(V3 n) {  // mesh.cpp:370-373
    if (__builtin_fabsf(n.x) > __builtin_fabsf(n.y))
        return divs(V3{-n.z, 0.0f, n.x}, __builtin_sqrtf(n.x * n.x + n.z * n.z));
    return divs(V3{0.0f, n.z, -n.y}, __builtin_sqrtf(n.y * n.y + n.z * n.z));
}

// The primitive's object-to-world matrix and inverse at a ray time
// (Primitive::getToWorld, primitive.cpp:82-84)
GDEV_INLINE void primXf(const DScene& S, const DPrim& pr, float time, M4& W, M4& Wi) {
    W = pr.toWorld;
    Wi = pr.toWorldInv;
    if (pr.motion >= 0) motionAt(S.motions[pr.motion], time, W, Wi);
}

// Mesh::fillInteraction / Sphere::fillInteraction + Primitive::fillInteraction
// (mesh.cpp:328-390, sphere.cpp:41-55, primitive.cpp:61-72); time: the ray's
GDEV_INLINE void fillInteraction(const DScene& S, V3 ro, V3 rd, const HitRec& h, Isect& it, float time) {
    const DPrim& pr = S.prims[h.inst];
    M4 W, Wi;
    primXf(S, pr, time, W, Wi);
    V3 P, N, dpdu;
    float uu = h.u, vv = h.v;
    if (pr.shape == GND_SHAPE_MESH) {
        const float* sh = S.shading + (size_t)(pr.triBase + (uint32_t)h.prim) * 24;
        V3 p0 = V3{sh[0], sh[1], sh[2]}, p1 = V3{sh[3], sh[4], sh[5]}, p2 = V3{sh[6], sh[7], sh[8]};
        V3 n0 = V3{sh[9], sh[10], sh[11]}, n1 = V3{sh[12], sh[13], sh[14]}, n2 = V3{sh[15], sh[16], sh[17]};
        float u = h.u, v = h.v, w = 1.0f - u - v;
        P = p0 * w + p1 * u + p2 * v;
        N = normalize(n0 * w + n1 * u + n2 * v);
        if (pr.hasUV) {
            float u0 = sh[18], v0 = sh[19], u1 = sh[20], v1 = sh[21], u2 = sh[22], v2 = sh[23];
            uu = u0 * w + u1 * u + u2 * v;
            vv = v0 * w + v1 * u + v2 * v;
            V3 dp1 = p1 - p0, dp2 = p2 - p0;
            float du1 = u1 - u0, dv1 = v1 - v0, du2 = u2 - u0, dv2 = v2 - v0;
            float det = du1 * dv2 - dv1 * du2;
            if (__builtin_fabsf(det) < 1e-8f) dpdu = dpduFallback(N);
            else {
                float invDet = 1.0f / det;
                dpdu = (dp1 * dv2 - dp2 * dv1) * invDet;
            }
            dpdu = normalize(dpdu - N * dot(N, dpdu));
        } else {
            dpdu = normalize(dpduFallback(N));
        }
    } else {
        V3 lo = xfPoint(Wi, ro);
        V3 ld = xfVector(Wi, rd);
        P = lo + ld * h.t;
        N = normalize(P - V3{0, 0, 0});
        float phi = atan2f(N.x, N.z);
        if (phi < 0.0f) phi = (float)((double)phi + 2.0f * kPiD);
        uu = phi * kInv2Pi;
        float theta = acosf(clampf(N.y, -1.0f, 1.0f));
        vv = theta * kInvPi;
        dpdu = V3{(float)((double)(P.z * 2.0f) * kPiD), 0.0f, (float)((double)(-P.x * 2.0f) * kPiD)};
    }
    it.u = uu;
    it.v = vv;
    it.prim = h.inst;
    it.p = xfPoint(W, P);
    it.n = normalize(xfNormal(Wi, N));
    it.dpdu = normalize(xfVector(W, dpdu));
}

GDEV_INLINE bool makeBsdf(const DScene& S, const Isect& it, Bsdf& b) {
    const DMaterial* m = S.mats + S.prims[it.prim].material;
    if (!makeLobe(m, it.u, it.v, b.lobe)) return false;
    b.fr.x = it.dpdu;
    b.fr.y = normalize(cross(it.n, it.dpdu));
    b.fr.z = it.n;
    return true;
}

// medium interface of a surface (Material::applyMediums, material.h:119-126)
GDEV_INLINE void mediumInterface(const DScene& S, int prim, int& inside, int& outside) {
    const DMaterial& m = S.mats[S.prims[prim].material];
    inside = outside = -1;
    if ((m.type == GND_MAT_GLASS || m.type == GND_MAT_INDEX) && (m.mediumInside >= 0 || m.mediumOutside >= 0)) {
        inside = m.mediumInside;
        outside = m.mediumOutside;
    }
}

// ------------------------------------------------------------------ area light (src/emitters/area.cpp)
GDEV_INLINE V3 areaEval(const DPrim& pr, V3 n, V3 w) {  // area.cpp:45-49
    if (dot(n, w) > 0.0f) return V3{pr.radiance[0], pr.radiance[1], pr.radiance[2]};
    return sameV(0.0f);
}

struct LightSample {
    V3 p, n;
    float pdf;
    V3 dir;
    float tmax;
    V3 Li;
    bool ok;  // pdf written (else the reference leaves it unset and Li is black)
};

// AreaLight::sample with Mesh::sampleSurface / Sphere::sampleSurface; time:
// the reference interaction's (area.cpp:54)
GDEV_INLINE void areaSample(const DScene& S, int primIdx, V3 refP, float sx, float sy, LightSample& ls, float time) {
    const DPrim& pr = S.prims[primIdx];
    M4 W, Wi;
    primXf(S, pr, time, W, Wi);
    V3 lp = V3{0, 0, 0}, ln = V3{0, 0, 0};
    float areaPdf = 0.0f;
    if (pr.shape == GND_SHAPE_MESH) {
        if (pr.totalArea != 0.0f) {
            const float* cdf = S.cdf + pr.cdfBase;
            const uint32_t n = pr.nTris + 1;
            // std::lower_bound: first element not less than sx
            uint32_t lo = 0, cnt = n;
            while (cnt > 0) {
                uint32_t step = cnt / 2, mid = lo + step;
                if (cdf[mid] < sx) { lo = mid + 1; cnt -= step + 1; }
                else cnt = step;
            }
            uint32_t ti = lo < 1u ? 1u : (lo > n - 1 ? n - 1 : lo);
            ti -= 1;
            float pdfTri = cdf[ti + 1] - cdf[ti];
            float rx = clampf((sx - cdf[ti]) / pdfTri, 0.0f, 1.0f);
            float sq = __builtin_sqrtf(rx);
            float bu = 1.0f - sq, bv = sy * sq, bw = 1.0f - bu - bv;
            const float* sh = S.shading + (size_t)(pr.triBase + ti) * 24;
            lp = V3{sh[0], sh[1], sh[2]} * bw + V3{sh[3], sh[4], sh[5]} * bu + V3{sh[6], sh[7], sh[8]} * bv;
            ln = normalize(V3{sh[9], sh[10], sh[11]} * bw + V3{sh[12], sh[13], sh[14]} * bu +
                           V3{sh[15], sh[16], sh[17]} * bv);
            areaPdf = 1.0f / pr.totalArea;
        }
    } else {
        V3 r = squareToUniformSphere(sx, sy);
        lp = r * pr.radius;
        ln = normalize(lp - V3{0, 0, 0});
        areaPdf = kInv4Pi / (pr.radius * pr.radius);
    }
    ls.p = xfPoint(W, lp);
    ls.n = normalize(xfNormal(Wi, ln));
    Frame fr = frameFromNormal(ln);
    V3 t1 = xfVector(W, fr.x), t2 = xfVector(W, fr.y);
    float areaScale = length(cross(t1, t2));
    if (areaScale > 0.0f) areaPdf /= areaScale;
    ls.dir = normalize(ls.p - refP);
    ls.tmax = length(ls.p - refP) - kEps;
    float dist = lengthSq(refP - ls.p);
    float cosT = dot(ls.n, -ls.dir);
    ls.ok = false;
    ls.Li = sameV(0.0f);
    if (areaPdf <= kEps || cosT <= kEps) return;
    ls.pdf = areaPdf * dist / cosT;
    ls.ok = true;
    ls.Li = cdiv(areaEval(pr, ls.n, normalize(refP - ls.p)), ls.pdf);
}

// AreaLight::pdf (area.cpp:90-119); time: the reference interaction's (area.cpp:92)
GDEV_INLINE float areaPdf(const DScene& S, int primIdx, V3 refP, V3 p, V3 n, float time) {
    const DPrim& pr = S.prims[primIdx];
    M4 W, Wi;
    primXf(S, pr, time, W, Wi);
    V3 ln = xfNormal(W, n);
    float ap = pr.shape == GND_SHAPE_MESH ? (pr.totalArea > 0.0f ? 1.0f / pr.totalArea : 0.0f)
                                          : kInv4Pi / (pr.radius * pr.radius);
    Frame fr = frameFromNormal(ln);
    V3 t1 = xfVector(W, fr.x), t2 = xfVector(W, fr.y);
    float areaScale = length(cross(t1, t2));
    if (areaScale > 0.0f) ap /= areaScale;
    float dist = lengthSq(refP - p);
    V3 wi = normalize(p - refP);
    float cosT = dot(n, -wi);
    if (ap <= kEps || cosT <= kEps) return 0.0f;
    return ap * dist / cosT;
}

GDEV_INLINE float powerHeuristic(float f, float g) {  // path.cpp:13-17
    float d = (f * f) + (g * g);
    return d > 0.0f ? (f * f) / d : 0.0f;
}
GDEV_INLINE int pickEmitter(const DScene& S, float rnd) {  // scene.h:28-34
    uint32_t n = (uint32_t)S.nEmitters;
    float f = floorf((float)n * rnd);
    uint32_t idx = (uint32_t)f;
    if (idx > n - 1) idx = n - 1;
    return S.emitters[idx];
}

// ------------------------------------------------------------------ non-area lights
// SphericalCoordinates / SphericalDirection (constants.h:80-100)
GDEV_INLINE void sphericalCoordinates(V3 w, float& theta, float& phi) {
    theta = acosf(clampf(w.z, -1.0f, 1.0f));
    phi = atan2f(w.y, w.x);
    if (phi < 0.0f) phi += 2.0f * kPi;
}
// Film::getPixelBilinear (film.cpp:125-151): u wraps, v clamps
GDEV_INLINE V3 envBilinear(const DScene& S, float u, float v) {
    const int W = S.envW, H = S.envH;
    float x = u * W - 0.5f;
    float y = v * H - 0.5f;
    int x0 = (int)floorf(x);
    int y0 = (int)floorf(y);
    float dx = x - x0;
    float dy = y - y0;
    int px0 = (x0 % W + W) % W;
    int px1 = (px0 + 1) % W;
    int py0 = y0 < 0 ? 0 : (y0 > H - 1 ? H - 1 : y0);
    int py1 = y0 + 1 < 0 ? 0 : (y0 + 1 > H - 1 ? H - 1 : y0 + 1);
    auto px = [&](int xx, int yy) {
        const float* c = S.envRgb + 3 * ((size_t)yy * W + xx);
        return V3{c[0], c[1], c[2]};
    };
    V3 c0 = px(px0, py0) * (1.0f - dx) + px(px1, py0) * dx;
    V3 c1 = px(px0, py1) * (1.0f - dx) + px(px1, py1) * dx;
    return c0 * (1.0f - dy) + c1 * dy;
}
// EnvironmentMap::eval (environment.cpp:31-40); u, v in double as phi * M_1_PI
GDEV_INLINE V3 envEval(const DScene& S, V3 w) {
    float theta, phi;
    sphericalCoordinates(w, theta, phi);
    float u = (float)((double)phi * kInvPiD * 0.5);
    float v = (float)((double)theta * kInvPiD);
    return envBilinear(S, u, v) * S.envStrength;
}
// Distribution2D::Pdf (distribution.cpp:75-79)
GDEV_INLINE float envMapPdf(const DScene& S, float u, float v) {
    int iu = (int)(u * S.envW), iv = (int)(v * S.envH);
    iu = iu < 0 ? 0 : (iu > S.envW - 1 ? S.envW - 1 : iu);
    iv = iv < 0 ? 0 : (iv > S.envH - 1 ? S.envH - 1 : iv);
    return S.envCondFunc[(size_t)iv * S.envW + iu] / S.envMargInt;
}
// EnvironmentMap::pdf (environment.cpp:75-91)
GDEV_INLINE float envPdf(const DScene& S, V3 refP, V3 infoP) {
    V3 wi = normalize(infoP - refP);
    float theta, phi;
    sphericalCoordinates(wi, theta, phi);
    float sinTheta = sinf(theta);
    if (sinTheta <= 0.0f) return 0.0f;
    float u = (float)((double)phi * kInvPiD * 0.5);
    float v = (float)((double)theta * kInvPiD);
    float mapPdf = envMapPdf(S, u, v);
    return (float)((double)mapPdf / (2.0 * kPiD * kPiD * (double)sinTheta));
}
// Distribution1D::SampleContinuous (distribution.cpp:20-34) over func / cdf
// of n entries: std::upper_bound, clamp, linear remap within the bucket.
// The search runs over [guide[k], guide[k+1]] for u's bucket k = floor(u *
// kEnvGuide) (exact: a power-of-two product): upper_bound is monotone in u and
// guide[j] is its value at j / kEnvGuide, so the first entry greater than u
// lies in that range, and the result is upper_bound's over the whole table.
GDEV_INLINE float sample1D(const float* func, const float* cdf, const int32_t* guide, int n, float funcInt, float u,
                           float& pdf, int& off) {
    int k = (int)(u * (float)kEnvGuide);
    k = k < 0 ? 0 : (k > kEnvGuide ? kEnvGuide : k);
    int lo = guide[k], cnt = guide[k + 1] - lo;  // first cdf entry greater than u
    while (cnt > 0) {
        int step = cnt / 2, mid = lo + step;
        if (!(u < cdf[mid])) { lo = mid + 1; cnt -= step + 1; }
        else cnt = step;
    }
    int offset = lo - 1;
    offset = offset < 0 ? 0 : (offset > n - 1 ? n - 1 : offset);
    off = offset;
    float du = u - cdf[offset];
    if ((cdf[offset + 1] - cdf[offset]) > 0.0f) du /= (cdf[offset + 1] - cdf[offset]);
    pdf = (funcInt > 0.0f) ? func[offset] / funcInt : 0.0f;
    return (offset + du) / n;
}
// Emitter::sample of the non-area emitters (pointlight.cpp:18-33,
// spotlight.cpp:28-60, directional.cpp:17-33, environment.cpp:42-73); the
// pdf is always written, so ls.ok
GDEV_INLINE void lightSample(const DScene& S, int k, V3 refP, float ux, float uy, LightSample& ls) {
    const DLight& L = S.lights[k];
    ls.ok = true;
    ls.Li = sameV(0.0f);
    if (L.type == GND_LIGHT_POINT || L.type == GND_LIGHT_SPOT) {
        const V3 lp = V3{L.pos[0], L.pos[1], L.pos[2]};
        const V3 I = V3{L.power[0], L.power[1], L.power[2]} * kInv4Pi;
        ls.p = lp;
        ls.pdf = 1.0f;
        if (L.type == GND_LIGHT_POINT) {
            float squaredDist = lengthSq(lp - refP);
            ls.dir = normalize(ls.p - refP);
            ls.tmax = length(ls.p - refP) - kEps;
            ls.Li = cdiv(I, squaredDist);
        } else {
            V3 wi = normalize(lp - refP);
            float dist = length(lp - refP);
            float squaredDist = dist * dist;
            ls.dir = wi;
            ls.tmax = dist - kEps;
            float cosTheta = dot(-wi, V3{L.dir[0], L.dir[1], L.dir[2]});
            if (cosTheta <= L.cosTotal) return;
            float falloff = 1.0f;
            if (cosTheta < L.cosFalloff) {
                float t = (cosTheta - L.cosTotal) / (L.cosFalloff - L.cosTotal);
                falloff = (t * t) * (3.0f - 2.0f * t);
            }
            ls.Li = cdiv(I * falloff, squaredDist);
        }
        return;
    }
    if (L.type == GND_LIGHT_DIRECTIONAL) {
        const V3 wi = V3{L.dir[0], L.dir[1], L.dir[2]};
        ls.p = refP + wi * 1e5f;
        ls.pdf = 1.0f;
        ls.dir = wi;
        ls.tmax = kFltMax;
        ls.Li = V3{L.irr[0], L.irr[1], L.irr[2]};
        return;
    }
    // environment: Distribution2D::SampleContinuous (distribution.cpp:65-73)
    float pdf1, pdf0;
    int v, unused;
    const float d1 = sample1D(S.envCondInt, S.envMargCdf, S.envMargGuide, S.envH, S.envMargInt, uy, pdf1, v);
    const float d0 = sample1D(S.envCondFunc + (size_t)v * S.envW, S.envCondCdf + (size_t)v * (S.envW + 1),
                              S.envCondGuide + (size_t)v * (kEnvGuide + 2), S.envW, S.envCondInt[v], ux, pdf0, unused);
    const float mapPdf = pdf0 * pdf1;
    if (mapPdf == 0.0f) {
        ls.pdf = 0.0f;
        return;
    }
    float theta = (float)((double)d1 * kPiD);
    float phi = (float)((double)(d0 * 2.0f) * kPiD);
    float cosT = cosf(theta), sinT = sinf(theta), cosP = cosf(phi), sinP = sinf(phi);
    V3 wi = V3{sinT * cosP, sinT * sinP, cosT};
    ls.p = refP + wi * 1e5f;
    ls.dir = wi;
    ls.tmax = finf();
    float sinTheta = sinf(theta);
    if (sinTheta <= 0.0f) {
        ls.pdf = 0.0f;
        return;
    }
    ls.pdf = (float)((double)mapPdf / (2.0 * kPiD * kPiD * (double)sinTheta));
    ls.Li = cdiv(envEval(S, wi), ls.pdf);
}
// the NEE sample of emitter-list entry e (area emitter primitive, or
// -(light + 1)); delta = Emitter::isDelta.  AREA_ONLY: the scene's emitter
// list holds area emitters only (no light entries to branch to).
template <bool AREA_ONLY = false>
GDEV_INLINE void emitterSample(const DScene& S, int e, V3 refP, float ux, float uy, LightSample& ls, bool& delta,
                               float time) {
    if (AREA_ONLY || e >= 0) {
        delta = false;
        areaSample(S, e, refP, ux, uy, ls, time);
        return;
    }
    const int k = -e - 1;
    delta = S.lights[k].type != GND_LIGHT_ENVIRONMENT;
    lightSample(S, k, refP, ux, uy, ls);
}

// ------------------------------------------------------------------ media
GDEV_INLINE float hgPhase(float cosT, float g) {  // phase.h:86-89
    float denom = 1.0f + g * g + 2.0f * g * cosT;
    return kInv4Pi * (1.0f - g * g) / (denom * __builtin_sqrtf(denom));
}
GDEV_INLINE float hgSample(float g, V3 wo, V3& wi, float ux, float uy) {  // phase.h:49-73 (as written)
    float cosT;
    if (__builtin_fabsf(g) < 1e-3f) cosT = 1.0f - 2.0f * ux;
    else {
        float sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * ux);
        cosT = (1.0f + g * g - sq * sq) / (2.0f * g);
    }
    float sinT = __builtin_sqrtf(stdMax(0.0f, 1.0f - cosT * cosT));
    float phi = (float)(2.0f * kPiD * (double)uy);
    V3 v1;
    if (__builtin_fabsf(wo.x) > __builtin_fabsf(wo.y)) {
        float inv = 1.0f / __builtin_sqrtf(wo.x * wo.x + wo.z * wo.z);
        v1 = V3{-wo.z * inv, 0.0f, wo.x * inv};
    } else {
        float inv = 1.0f / __builtin_sqrtf(wo.y * wo.y + wo.z * wo.z);
        v1 = V3{0.0f, wo.z * inv, -wo.y * inv};
    }
    V3 v2 = cross(wo, v1);
    wi = v1 * (sinT * cosf(phi)) + v2 * (sinT * sinf(phi)) + wo * cosT;
    return hgPhase(cosT, g);
}
GDEV_INLINE float hgClamp(float g) { return clampf(g, -0.999f, 0.999f); }

// OpenVDB BoxSampler::isSample on the dense grid: lerp z, then y, then x.
// The 8 corner values come from the voxel's corner record (one sector)
// inside the grid, else one by one.
GDEV_INLINE float gridFetch(const DMedium& m, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= m.nx || j >= m.ny || k >= m.nz) return 0.0f;
    return m.density[((size_t)k * m.ny + j) * m.nx + i];
}
GDEV_INLINE float gridSample(const DMedium& m, double x, double y, double z) {
    int ix = (int)floor(x), iy = (int)floor(y), iz = (int)floor(z);
    double ux = x - ix, uy = y - iy, uz = z - iz;
    float d000, d001, d010, d011, d100, d101, d110, d111;
    if (m.corners && (unsigned)ix < (unsigned)m.nx && (unsigned)iy < (unsigned)m.ny && (unsigned)iz < (unsigned)m.nz) {
        const float4* r = m.corners + 2 * (((size_t)iz * m.ny + iy) * m.nx + ix);
        const float4 c0 = r[0], c1 = r[1];
        d000 = c0.x; d001 = c0.y; d010 = c0.z; d011 = c0.w;
        d100 = c1.x; d101 = c1.y; d110 = c1.z; d111 = c1.w;
    } else {
        d000 = gridFetch(m, ix, iy, iz); d001 = gridFetch(m, ix, iy, iz + 1);
        d010 = gridFetch(m, ix, iy + 1, iz); d011 = gridFetch(m, ix, iy + 1, iz + 1);
        d100 = gridFetch(m, ix + 1, iy, iz); d101 = gridFetch(m, ix + 1, iy, iz + 1);
        d110 = gridFetch(m, ix + 1, iy + 1, iz); d111 = gridFetch(m, ix + 1, iy + 1, iz + 1);
    }
    float a = d000 + (float)((d001 - d000) * uz);
    float b = d010 + (float)((d011 - d010) * uz);
    float r1 = a + (float)((b - a) * uy);
    a = d100 + (float)((d101 - d100) * uz);
    b = d110 + (float)((d111 - d110) * uz);
    float r2 = a + (float)((b - a) * uy);
    return r1 + (float)(ux * (r2 - r1));
}

// Macro-cell DDA state; the step along an axis is the sign of the index-space
// direction there (+1 if positive, else -1).
struct GridWalk {
    V3 io, id;
    float t, tMax;
    int cx, cy, cz;
    float tdx, tdy, tdz, tmx, tmy, tmz;
};

GDEV_INLINE int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
GDEV_INLINE float stdMin4(float a, float b, float c, float d) {  // std::min({a,b,c,d})
    float r = a;
    if (b < r) r = b;
    if (c < r) r = c;
    if (d < r) r = d;
    return r;
}

// heterogeneous.cpp:106-162 / 251-305: index-space ray, clip to the grid box,
// macro-cell DDA set-up
GDEV_INLINE bool gridSetup(const DMedium& m, V3 o, V3 d, float tmin, float tmax, GridWalk& g) {
    V3 lo = xfPoint(m.toLocal, o);
    V3 lt = xfPoint(m.toLocal, o + d);
    g.io = lo;
    g.id = V3{(float)((double)lt.x - (double)lo.x), (float)((double)lt.y - (double)lo.y),
              (float)((double)lt.z - (double)lo.z)};
    g.t = tmin;
    g.tMax = tmax;
    // Bounds3f::rayIntersect (bounds.h:108-128) over [0, n]^3, one axis at a
    // time (no arrays, so nothing is indexed through private memory)
    float t0 = 0.f, t1 = tmax;
    auto slab = [&](float oo, float dd, float bmax) {
        const float inv = 1.f / dd;
        float tn = (0.0f - oo) * inv, tf = (bmax - oo) * inv;
        if (tn > tf) { const float t = tn; tn = tf; tf = t; }
        t0 = tn > t0 ? tn : t0;
        t1 = tf < t1 ? tf : t1;
        return !(t0 > t1);
    };
    if (!slab(g.io.x, g.id.x, m.gridMaxX) || !slab(g.io.y, g.id.y, m.gridMaxY) || !slab(g.io.z, g.id.z, m.gridMaxZ))
        return false;
    g.t = stdMax(g.t, t0);
    g.tMax = stdMin(g.tMax, t1);
    V3 pe = g.io + g.id * g.t;
    g.cx = clampi((int)floorf((pe.x - 0.0f) / m.cellX), 0, m.resX - 1);
    g.cy = clampi((int)floorf((pe.y - 0.0f) / m.cellY), 0, m.resY - 1);
    g.cz = clampi((int)floorf((pe.z - 0.0f) / m.cellZ), 0, m.resZ - 1);
    g.tdx = g.id.x != 0.0f ? __builtin_fabsf(m.cellX / g.id.x) : kFltMax;
    g.tdy = g.id.y != 0.0f ? __builtin_fabsf(m.cellY / g.id.y) : kFltMax;
    g.tdz = g.id.z != 0.0f ? __builtin_fabsf(m.cellZ / g.id.z) : kFltMax;
    float nbx = 0.0f + (g.cx + (g.id.x > 0.0f ? 1 : 0)) * m.cellX;
    float nby = 0.0f + (g.cy + (g.id.y > 0.0f ? 1 : 0)) * m.cellY;
    float nbz = 0.0f + (g.cz + (g.id.z > 0.0f ? 1 : 0)) * m.cellZ;
    g.tmx = g.id.x != 0.0f ? (nbx - g.io.x) / g.id.x : kFltMax;
    g.tmy = g.id.y != 0.0f ? (nby - g.io.y) / g.id.y : kFltMax;
    g.tmz = g.id.z != 0.0f ? (nbz - g.io.z) / g.id.z : kFltMax;
    if (g.tmx <= g.t) g.tmx += g.tdx;
    if (g.tmy <= g.t) g.tmy += g.tdy;
    if (g.tmz <= g.t) g.tmz += g.tdz;
    return true;
}
// One DDA step (heterogeneous.cpp:206-224): x if tmx < tmy and tmx < tmz, y if
// tmy <= tmx and tmy < tmz, else z.  Written as selects of values: branches
// that each update a different field let the compiler merge them into one
// update through a pointer to the chosen field, which puts the whole walk
// state in private memory (volumetric 720p frame 91.0 -> 88.7 ms).
GDEV_INLINE void gridAdvance(GridWalk& g) {
    const bool xl = g.tmx < g.tmy;
    const bool ax = xl && g.tmx < g.tmz;
    const bool ay = !xl && g.tmy < g.tmz;
    const bool az = !ax && !ay;
    g.tmx = ax ? g.tmx + g.tdx : g.tmx;
    g.tmy = ay ? g.tmy + g.tdy : g.tmy;
    g.tmz = az ? g.tmz + g.tdz : g.tmz;
    g.cx += ax ? (g.id.x > 0.0f ? 1 : -1) : 0;
    g.cy += ay ? (g.id.y > 0.0f ? 1 : -1) : 0;
    g.cz += az ? (g.id.z > 0.0f ? 1 : -1) : 0;
}

// The macro-cell walk of ratio tracking (Medium::Tr, heterogeneous.cpp:106-229,
// SAMPLE = false) and delta tracking (Medium::sample, heterogeneous.cpp:251-376,
// SAMPLE = true) in resumable form: step() enters one cell or makes one
// tentative collision and returns true once the walk has ended.  The scalar
// functions below loop on it; the wavefront kernels interleave the steps of
// many walks so that a lane whose walk ends takes new work.
//   ratio tracking: `tr` is the transmittance, or zero if `hit` (the walk
//                   ended below the throughput floor or by roulette)
//   delta tracking: `hit` = a real collision at `t`, density `dens`, channel `ch`
template <bool SAMPLE>
struct GridTrack {
    GridWalk g;
    float t, tExit, invMaj;
    V3 tr;
    float dens;
    int ch;
    bool inCell, hit;
    GDEV_INLINE bool start(const DMedium& m, V3 o, V3 d, float tmin, float tmax) {
        tr = sameV(1.0f);
        inCell = false;
        hit = false;
        if (!gridSetup(m, o, d, tmin, tmax, g)) return false;
        t = g.t;
        return true;
    }
    // t = the cell's exit; the walk ends there or moves to the next cell
    GDEV_INLINE bool leaveCell() {
        inCell = false;
        t = tExit;
        if (t >= g.tMax) return true;
        gridAdvance(g);
        return false;
    }
    GDEV_INLINE bool step(const DMedium& m, Rng& rng) {
        if (!inCell) {
            if (!(t < g.tMax)) return true;
            if (g.cx < 0 || g.cx >= m.resX || g.cy < 0 || g.cy >= m.resY || g.cz < 0 || g.cz >= m.resZ) return true;
            const float maj = m.majorants[(size_t)g.cz * (m.resX * m.resY) + g.cy * m.resX + g.cx];
            tExit = stdMin4(g.tmx, g.tmy, g.tmz, g.tMax);
            if (!(maj > 0.0f)) return leaveCell();
            invMaj = 1.0f / maj;
            inCell = true;
        }
        if (!(t < tExit)) return leaveCell();
        float stp = -logf(1.0f - rng.next1D()) * invMaj;
        if (stp <= 0.0f) stp = 1e-5f;
        if (t + stp >= tExit) return leaveCell();
        t += stp;
        double px = g.io.x + g.id.x * t, py = g.io.y + g.id.y * t, pz = g.io.z + g.id.z * t;
        float dn = gridSample(m, px, py, pz) * m.densityScale;
        if (dn < 1e-5f) dn = 0.0f;
        if (!(dn > 0.0f)) return false;
        if (SAMPLE) {
            float r = rng.next1D() * 3.0f;
            int c = (int)r < 2 ? (int)r : 2;
            float prob = dn * m.sigma_t[c] * invMaj;
            if (rng.next1D() < prob) {
                hit = true;
                dens = dn;
                ch = c;
                return true;
            }
            return false;
        } else {
            V3 lst = V3{m.sigma_t[0], m.sigma_t[1], m.sigma_t[2]} * dn;
            tr = tr * (sameV(1.0f) - lst * invMaj);
            float maxTr = stdMax(stdMax(tr.x, tr.y), tr.z);
            if (maxTr < 1e-4f) {
                hit = true;
                return true;
            }
            if (maxTr < 0.1f) {
                float q = stdMax(0.05f, 1.0f - maxTr);
                if (rng.next1D() < q) {
                    hit = true;
                    return true;
                }
                tr = cdiv(tr, 1.0f - q);
            }
            return false;
        }
    }
    // ratio tracking: the segment's transmittance
    GDEV_INLINE V3 transmittance() const { return hit ? sameV(0.0f) : tr; }
};

// Whether the line o + t d, t in [0, tmax], may pass through density grid m:
// the line in the grid's index space against the grid's box grown by a
// margin, over a slightly longer t range (conservative: a ray that ratio
// tracking would walk in the grid always passes, including the segments a
// shadow ray continues along from null-surface crossings near the line).
// Comparisons with NaN keep the answer "may".
GDEV_INLINE bool lineMayCrossGrid(const DMedium& m, V3 o, V3 d, float tmax) {
    const V3 lo = xfPoint(m.toLocal, o);
    const V3 lt = xfPoint(m.toLocal, o + d);
    const float dl[3] = {lt.x - lo.x, lt.y - lo.y, lt.z - lo.z};
    const float ol[3] = {lo.x, lo.y, lo.z};
    const float bmax[3] = {m.gridMaxX, m.gridMaxY, m.gridMaxZ};
    const float mx = 1e-3f * stdMax(stdMax(m.gridMaxX, m.gridMaxY), m.gridMaxZ) + 1e-3f;
    float t0 = -1e-3f, t1 = __builtin_isinf(tmax) ? finf() : tmax * 1.001f + 1e-3f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (__builtin_fabsf(dl[a]) < 1e-30f) {
            if (ol[a] < -mx || ol[a] > bmax[a] + mx) return false;
            continue;
        }
        const float inv = 1.0f / dl[a];
        float tn = (-mx - ol[a]) * inv, tf = (bmax[a] + mx - ol[a]) * inv;
        if (tn > tf) {
            const float t = tn;
            tn = tf;
            tf = t;
        }
        if (tn > t0) t0 = tn;
        if (tf < t1) t1 = tf;
        if (t0 > t1) return false;
    }
    return true;
}

// Medium::Tr (homogeneous.cpp:22-30; heterogeneous.cpp:106-229 ratio tracking)
GDEV_INLINE V3 mediumTr(const DMedium& m, V3 o, V3 d, float tmin, float tmax, Rng& rng) {
    if (m.type == GND_MEDIUM_HOMOGENEOUS) {
        if (__builtin_isinf(tmax)) return sameV(0.0f);
        return V3{expf(-m.sigma_t[0] * tmax), expf(-m.sigma_t[1] * tmax), expf(-m.sigma_t[2] * tmax)};
    }
    GridTrack<false> k;
    if (!k.start(m, o, d, tmin, tmax)) return sameV(1.0f);
    while (!k.step(m, rng)) {
    }
    return k.transmittance();
}

struct MediumEvent {
    bool valid;
    V3 p;
    V3 sigma_s;
};

// a delta-tracking collision as a medium event, and the throughput factor
// 1 / sigma_t(channel) it returns (heterogeneous.cpp:251-376)
GDEV_INLINE void gridEvent(const DMedium& m, V3 o, V3 d, const GridTrack<true>& k, MediumEvent& ev) {
    ev.valid = true;
    ev.p = o + d * k.t;
    ev.sigma_s = V3{m.sigma_s[0], m.sigma_s[1], m.sigma_s[2]} * k.dens;
}
GDEV_INLINE V3 gridWeight(const DMedium& m, const GridTrack<true>& k) {
    V3 lst = V3{m.sigma_t[0], m.sigma_t[1], m.sigma_t[2]} * k.dens;
    float sc = k.ch == 0 ? lst.x : (k.ch == 1 ? lst.y : lst.z);
    return cdiv(sameV(1.0f), sc);
}

// Medium::sample (homogeneous.cpp:32-63; heterogeneous.cpp:251-376 delta tracking)
GDEV_INLINE V3 mediumSample(const DMedium& m, V3 o, V3 d, float tmin, float tmax, Rng& rng, MediumEvent& ev) {
    ev.valid = false;
    const V3 st = V3{m.sigma_t[0], m.sigma_t[1], m.sigma_t[2]};
    if (m.type == GND_MEDIUM_HOMOGENEOUS) {
        float r = rng.next1D() * 3.0f;
        int ch = (int)r < 2 ? (int)r : 2;
        float dist = -logf(1.0f - rng.next1D()) / m.sigma_t[ch];
        float t = stdMin(dist, tmax);
        bool sampled = t < tmax;
        V3 tr = V3{expf(-m.sigma_t[0] * t), expf(-m.sigma_t[1] * t), expf(-m.sigma_t[2] * t)};
        V3 pc = sampled ? (st * tr) : tr;
        float pdf = (pc.x + pc.y + pc.z) / 3.0f;
        if (pdf == 0.0f) return sameV(0.0f);
        if (sampled) {
            ev.valid = true;
            ev.p = o + d * t;
            ev.sigma_s = V3{m.sigma_s[0], m.sigma_s[1], m.sigma_s[2]};
        }
        return cdiv(tr, pdf);
    }
    GridTrack<true> k;
    if (!k.start(m, o, d, tmin, tmax)) return sameV(1.0f);
    while (!k.step(m, rng)) {
    }
    if (!k.hit) return sameV(1.0f);
    gridEvent(m, o, d, k, ev);
    return gridWeight(m, k);
}

// ------------------------------------------------------------------ camera
// PerspectiveCamera::shootRay (perspective.cpp:15-32), ThinLensCamera::shootRay (thinLens.cpp:34-95)
struct DCamera {
    M4 toWorld;
    int type, width, height, blades;
    float scale, aspect, lensRadius, focalDistance;
    float k1, caAxial, caLateral;
    int vignetting;
    int filter;
    float radiusX, radiusY, alpha;
    float expX, expY;
    int hasCA, animated;      // animated: the camera moves (motion blur)
    int rolling;              // thin lens rolling shutter
    float scanDuration, exposureDuration;
    M4 toWorldEnd;
    GndMotion motion;
};

// time: the sample's time draw (rolling shutter remaps it on the thin lens);
// out: the ray's time, which only animated geometry reads
GDEV_INLINE float shootRay(const DCamera& c, float fx, float fy, float lx, float ly, float lambdaOffset, float time,
                           V3& ro, V3& rd, float& rayTime) {
    V3 dir, org = V3{0, 0, 0};
    float weight = 1.0f;
    if (c.type == GND_CAMERA_PERSPECTIVE) {
        float x = (2.0f * fx - 1.0f) * c.aspect * c.scale;
        float y = (2.0f * fy - 1.0f) * c.scale;
        dir = normalize(V3{x, y, -1.0f});
    } else {
        float ndcX = 2.0f * fx - 1.0f, ndcY = 2.0f * fy - 1.0f;
        float r2 = ndcX * ndcX + ndcY * ndcY;
        float dist = 1.0f + c.k1 * r2;
        float distX = ndcX * dist, distY = ndcY * dist;
        float scaleMod = c.scale * (1.0f + c.caLateral * lambdaOffset);
        float x = distX * c.aspect * scaleMod, y = distY * scaleMod;
        dir = V3{x, y, -1.0f};
        if (c.lensRadius > 0.0f) {
            float px, py;
            if (c.blades >= 3) squareToUniformPolygon(lx, ly, c.blades, px, py);
            else squareToUniformDisk(lx, ly, px, py);
            px = px * c.lensRadius;
            py = py * c.lensRadius;
            float fdm = c.focalDistance * (1.0f + c.caAxial * lambdaOffset);
            float ft = fdm / __builtin_fabsf(dir.z);
            V3 pFocus = org + dir * ft;
            org = V3{px, py, 0.0f};
            dir = pFocus - org;
        }
        dir = normalize(dir);
        if (c.vignetting) {
            float ct = __builtin_fabsf(dir.z);
            float c2 = ct * ct;
            weight = c2 * c2;
        }
    }
    // thinLens.cpp:81-87: rolling shutter; perspective.cpp:27-29 uses the draw
    float t = time;
    if (c.type != GND_CAMERA_PERSPECTIVE && c.rolling) {
        const float scanline = 1.0f - fy;
        t = scanline * c.scanDuration + (time * c.exposureDuration);
    }
    rayTime = t;
    // AnimatedTransform::interpolate: the keyframes at or beyond the ends
    if (!c.animated || t <= 0.0f) {
        ro = xfPoint(c.toWorld, org);
        rd = xfVector(c.toWorld, dir);
    } else if (t >= 1.0f) {
        ro = xfPoint(c.toWorldEnd, org);
        rd = xfVector(c.toWorldEnd, dir);
    } else {
        const M4 M = motionMatrix(c.motion, (t - 0.0f) / (1.0f - 0.0f));
        ro = xfPoint(M, org);
        rd = xfVector(M, dir);
    }
    return weight;
}

// Film filters (include/gianduia/core/filter.h:20-54)
GDEV_INLINE float filterEval(const DCamera& c, float ox, float oy) {
    if (c.filter == GND_FILTER_BOX) return (__builtin_fabsf(ox) <= 0.5f && __builtin_fabsf(oy) <= 0.5f) ? 1.0f : 0.0f;
    float gx = stdMax(0.0f, expf(-c.alpha * ox * ox) - c.expX);
    float gy = stdMax(0.0f, expf(-c.alpha * oy * oy) - c.expY);
    return gx * gy;
}

}  // namespace gdev
