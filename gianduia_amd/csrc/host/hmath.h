// hmath.h — host-side vector / bounds / matrix arithmetic of the scene shell.
//
// The BVH build and the scene flattening must produce the same floats as the
// reference host code, so every operation here keeps the reference's
// evaluation order (glm scalar forms; include/gianduia/math/geometry.h:243-257,
// include/gianduia/math/bounds.h:11-106, include/gianduia/math/transform.h:37-76).
// Compiled with -ffp-contract=off.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

namespace gnd {

struct Vec3 {
    float x = 0, y = 0, z = 0;
    Vec3() = default;
    Vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator-(Vec3 a) { return {-a.x, -a.y, -a.z}; }
inline Vec3 operator*(Vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline Vec3 operator/(Vec3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float dot(Vec3 a, Vec3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline Vec3 cross(Vec3 a, Vec3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
inline float length(Vec3 v) { return std::sqrt(dot(v, v)); }
inline Vec3 normalize(Vec3 v) { return v * (1.0f / std::sqrt(dot(v, v))); }

// std::min / std::max on floats: (b < a) ? b : a and (a < b) ? b : a
inline float fmin_std(float a, float b) { return (b < a) ? b : a; }
inline float fmax_std(float a, float b) { return (a < b) ? b : a; }

struct Box {
    Vec3 lo{std::numeric_limits<float>::max(), std::numeric_limits<float>::max(), std::numeric_limits<float>::max()};
    Vec3 hi{std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest(),
            std::numeric_limits<float>::lowest()};
    void grow(Vec3 p) {
        lo = {fmin_std(lo.x, p.x), fmin_std(lo.y, p.y), fmin_std(lo.z, p.z)};
        hi = {fmax_std(hi.x, p.x), fmax_std(hi.y, p.y), fmax_std(hi.z, p.z)};
    }
    void grow(const Box& b) {
        lo = {fmin_std(lo.x, b.lo.x), fmin_std(lo.y, b.lo.y), fmin_std(lo.z, b.lo.z)};
        hi = {fmax_std(hi.x, b.hi.x), fmax_std(hi.y, b.hi.y), fmax_std(hi.z, b.hi.z)};
    }
    float area() const {
        Vec3 d = hi - lo;
        return 2.f * (d.x * d.y + d.x * d.z + d.y * d.z);
    }
    int widestAxis() const {
        Vec3 d = hi - lo;
        if (d.x > d.y && d.x > d.z) return 0;
        if (d.y > d.z) return 1;
        return 2;
    }
    Vec3 center() const { return lo * 0.5f + hi * 0.5f; }
    Vec3 corner(int c) const { return {(c & 1) ? hi.x : lo.x, (c & 2) ? hi.y : lo.y, (c & 4) ? hi.z : lo.z}; }
};

// 4x4 column-major (m[c*4 + r]), the memory order of glm::mat4
struct Mat4 {
    float m[16];
    static Mat4 identity() {
        Mat4 r;
        std::memset(r.m, 0, sizeof(r.m));
        r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
        return r;
    }
    float& at(int c, int r) { return m[c * 4 + r]; }
    float at(int c, int r) const { return m[c * 4 + r]; }
};
// glm mat4 * mat4: column i = ((A0*B[i][0] + A1*B[i][1]) + A2*B[i][2]) + A3*B[i][3]
inline Mat4 mul(const Mat4& a, const Mat4& b) {
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int row = 0; row < 4; ++row)
            r.at(i, row) = ((a.at(0, row) * b.at(i, 0) + a.at(1, row) * b.at(i, 1)) + a.at(2, row) * b.at(i, 2)) +
                           a.at(3, row) * b.at(i, 3);
    return r;
}
// Transform::operator()(Point3f) (transform.h:43-50)
inline Vec3 xformPoint(const Mat4& M, Vec3 p) {
    float o[4];
    for (int r = 0; r < 4; ++r)
        o[r] = (M.at(0, r) * p.x + M.at(1, r) * p.y) + (M.at(2, r) * p.z + M.at(3, r) * 1.0f);
    if (o[3] == 1.0f) return {o[0], o[1], o[2]};
    return Vec3(o[0], o[1], o[2]) / o[3];
}
inline Vec3 xformVector(const Mat4& M, Vec3 v) {
    float o[3];
    for (int r = 0; r < 3; ++r)
        o[r] = (M.at(0, r) * v.x + M.at(1, r) * v.y) + (M.at(2, r) * v.z + M.at(3, r) * 0.0f);
    return {o[0], o[1], o[2]};
}

// A transform with its analytic inverse (Transform in transform.h:12-40)
struct Xform {
    Mat4 m = Mat4::identity();
    Mat4 inv = Mat4::identity();
    // Transform::operator*: (m1*m2, inv2*inv1)
    Xform then(const Xform& step) const {  // step * this (parser.cpp:162: t = step * t)
        Xform r;
        r.m = mul(step.m, m);
        r.inv = mul(inv, step.inv);
        return r;
    }
};
Xform xformTranslate(float x, float y, float z);
Xform xformScale(float x, float y, float z);
Xform xformRotate(float degrees, float ax, float ay, float az);
Xform xformLookAt(Vec3 origin, Vec3 target, Vec3 up);
Mat4 invert(const Mat4& m);

}  // namespace gnd
