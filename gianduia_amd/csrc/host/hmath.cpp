// hmath.cpp — transform factories of the scene shell (transform.h:101-144).
#include "hmath.h"

namespace gnd {

Xform xformTranslate(float x, float y, float z) {
    Xform t;
    t.m.at(3, 0) = x; t.m.at(3, 1) = y; t.m.at(3, 2) = z;
    t.inv.at(3, 0) = -x; t.inv.at(3, 1) = -y; t.inv.at(3, 2) = -z;
    return t;
}

Xform xformScale(float x, float y, float z) {
    Xform t;
    t.m.at(0, 0) = x; t.m.at(1, 1) = y; t.m.at(2, 2) = z;
    t.inv.at(0, 0) = 1.f / x; t.inv.at(1, 1) = 1.f / y; t.inv.at(2, 2) = 1.f / z;
    return t;
}

// glm::rotate(mat4(1), radians(deg), axis); the inverse is the transpose
Xform xformRotate(float degrees, float ax, float ay, float az) {
    const float a = degrees * 0.01745329251994329576923690768489f;
    const float c = std::cos(a), s = std::sin(a);
    Vec3 axis = normalize(Vec3(ax, ay, az));
    Vec3 tmp = axis * (1.0f - c);
    Xform t;
    Mat4& R = t.m;
    R.at(0, 0) = c + tmp.x * axis.x;
    R.at(0, 1) = tmp.x * axis.y + s * axis.z;
    R.at(0, 2) = tmp.x * axis.z - s * axis.y;
    R.at(1, 0) = tmp.y * axis.x - s * axis.z;
    R.at(1, 1) = c + tmp.y * axis.y;
    R.at(1, 2) = tmp.y * axis.z + s * axis.x;
    R.at(2, 0) = tmp.z * axis.x + s * axis.y;
    R.at(2, 1) = tmp.z * axis.y - s * axis.x;
    R.at(2, 2) = c + tmp.z * axis.z;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) t.inv.at(i, j) = R.at(j, i);
    return t;
}

// General 4x4 inverse (cofactor expansion in double, rounded to float)
Mat4 invert(const Mat4& M) {
    double a[16], inv[16];
    for (int i = 0; i < 16; ++i) a[i] = M.m[i];
    inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] +
             a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
    inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] -
             a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
    inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] +
             a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
    inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] -
              a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
    inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] -
             a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
    inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] +
             a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
    inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] -
             a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
    inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] +
              a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
    inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] +
             a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
    inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] -
             a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
    inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] +
              a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
    inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] -
              a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
    inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] -
             a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
    inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] +
             a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
    inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] -
              a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
    inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] +
              a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
    double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
    Mat4 r;
    for (int i = 0; i < 16; ++i) r.m[i] = (float)(inv[i] / det);
    return r;
}

// glm::lookAt (right-handed); Transform::LookAt keeps m = inverse(view), inv = view
Xform xformLookAt(Vec3 eye, Vec3 center, Vec3 up) {
    Vec3 f = normalize(center - eye);
    Vec3 s = normalize(cross(f, up));
    Vec3 u = cross(s, f);
    Mat4 v = Mat4::identity();
    v.at(0, 0) = s.x; v.at(1, 0) = s.y; v.at(2, 0) = s.z;
    v.at(0, 1) = u.x; v.at(1, 1) = u.y; v.at(2, 1) = u.z;
    v.at(0, 2) = -f.x; v.at(1, 2) = -f.y; v.at(2, 2) = -f.z;
    v.at(3, 0) = -dot(s, eye);
    v.at(3, 1) = -dot(u, eye);
    v.at(3, 2) = dot(f, eye);
    Xform t;
    t.m = invert(v);
    t.inv = v;
    return t;
}

}  // namespace gnd
