// motion.h — AnimatedTransform (src/math/animated.cpp) for the host scene
// build and the device kernels: keyframes at times 0 and 1 decomposed into
// translation, rotation and scale, interpolated per ray time.  The glm
// routines the reference calls are restated from glm's published sources:
//   decompose (gtx/matrix_decompose.inl): translation = column 3; scale and
//     shear by Gram-Schmidt over columns 0..2 (length sqrt((x x + y y) + z z),
//     normalise v * 1 / length); a handedness flip negates scale and columns;
//     rotation quaternion (x, y, z, w) by the trace method
//   mix(x, y, a) = x * (1 - a) + y * a
//   slerp (gtc/quaternion.inl): dot = (w w' + x x') + (y y' + z z'); negate
//     the second when dot < 0; linear mix when dot > 1 - FLT_EPSILON, else
//     (sin((1 - a) t) x + sin(a t) y) / sin(t), t = acos(dot)
//   translate(I, T) * mat4_cast(R) * scale(I, S), mat4 products with columns
//     ((A0 b0 + A1 b1) + A2 b2) + A3 b3
//   inverse (Transform(M) keeps glm::inverse(M); detail::compute_inverse):
//     cofactors, columns scaled by the alternating signs, then * 1 / det
// Matrices are column-major float[16] (m[c * 4 + r]).
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__) || defined(__HIP__)
#define GND_HD __host__ __device__ inline
#else
#define GND_HD inline
#endif

struct GndMotion {
    float t0[3], t1[3], s0[3], s1[3], r0[4], r1[4];
};

GND_HD void gndDecompose(const float* m, float T[3], float Q[4], float S[3]) {
    const float w = m[15];
    float L[16];
    for (int k = 0; k < 16; ++k) L[k] = m[k] / w;
    for (int k = 0; k < 3; ++k) T[k] = L[12 + k];
    float c[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) c[i][j] = L[i * 4 + j];
    float shear[3];
    auto len = [](const float* v) { return sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]); };
    auto dot = [](const float* a, const float* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; };
    S[0] = len(c[0]);
    for (int k = 0; k < 3; ++k) c[0][k] = c[0][k] * 1.0f / S[0];
    shear[2] = dot(c[0], c[1]);
    for (int k = 0; k < 3; ++k) c[1][k] = c[1][k] * 1.0f + c[0][k] * -shear[2];
    S[1] = len(c[1]);
    for (int k = 0; k < 3; ++k) c[1][k] = c[1][k] * 1.0f / S[1];
    shear[1] = dot(c[0], c[2]);
    for (int k = 0; k < 3; ++k) c[2][k] = c[2][k] * 1.0f + c[0][k] * -shear[1];
    shear[0] = dot(c[1], c[2]);
    for (int k = 0; k < 3; ++k) c[2][k] = c[2][k] * 1.0f + c[1][k] * -shear[0];
    S[2] = len(c[2]);
    for (int k = 0; k < 3; ++k) c[2][k] = c[2][k] * 1.0f / S[2];
    const float x[3] = {c[1][1] * c[2][2] - c[2][1] * c[1][2], c[1][2] * c[2][0] - c[2][2] * c[1][0],
                        c[1][0] * c[2][1] - c[2][0] * c[1][1]};
    if (dot(c[0], x) < 0.0f)
        for (int i = 0; i < 3; ++i) {
            S[i] *= -1.0f;
            for (int k = 0; k < 3; ++k) c[i][k] *= -1.0f;
        }
    const float tr = c[0][0] + c[1][1] + c[2][2];
    if (tr > 0.0f) {
        float root = sqrtf(tr + 1.0f);
        Q[3] = 0.5f * root;
        root = 0.5f / root;
        Q[0] = root * (c[1][2] - c[2][1]);
        Q[1] = root * (c[2][0] - c[0][2]);
        Q[2] = root * (c[0][1] - c[1][0]);
    } else {
        int i = 0;
        if (c[1][1] > c[0][0]) i = 1;
        if (c[2][2] > c[i][i]) i = 2;
        const int j = (i + 1) % 3, k = (j + 1) % 3;
        float root = sqrtf(c[i][i] - c[j][j] - c[k][k] + 1.0f);
        Q[i] = 0.5f * root;
        root = 0.5f / root;
        Q[j] = root * (c[i][j] + c[j][i]);
        Q[k] = root * (c[i][k] + c[k][i]);
        Q[3] = root * (c[j][k] - c[k][j]);
    }
}

GND_HD void gndMotionFrom(const float* start, const float* end, GndMotion& mo) {
    gndDecompose(start, mo.t0, mo.r0, mo.s0);
    gndDecompose(end, mo.t1, mo.r1, mo.s1);
}

GND_HD float gndMix(float x, float y, float a) { return x * (1.0f - a) + y * a; }

GND_HD void gndMul44(const float* A, const float* B, float* R) {
    for (int i = 0; i < 4; ++i)
        for (int r = 0; r < 4; ++r)
            R[i * 4 + r] = ((A[0 * 4 + r] * B[i * 4 + 0] + A[1 * 4 + r] * B[i * 4 + 1]) + A[2 * 4 + r] * B[i * 4 + 2]) +
                           A[3 * 4 + r] * B[i * 4 + 3];
}

// AnimatedTransform::interpolate (animated.cpp:30-47) at dt inside (0, 1)
GND_HD void gndMotionMatrix(const GndMotion& mo, float dt, float* out) {
    float T[3], S[3], q[4];
    for (int k = 0; k < 3; ++k) {
        T[k] = gndMix(mo.t0[k], mo.t1[k], dt);
        S[k] = gndMix(mo.s0[k], mo.s1[k], dt);
    }
    const float* x = mo.r0;
    float z[4] = {mo.r1[0], mo.r1[1], mo.r1[2], mo.r1[3]};
    float cosT = (x[3] * z[3] + x[0] * z[0]) + (x[1] * z[1] + x[2] * z[2]);
    if (cosT < 0.0f) {
        for (int k = 0; k < 4; ++k) z[k] = -z[k];
        cosT = -cosT;
    }
    if (cosT > 1.0f - 1.1920929e-7f) {
        for (int k = 0; k < 4; ++k) q[k] = gndMix(x[k], z[k], dt);
    } else {
        const float angle = acosf(cosT);
        const float a = sinf((1.0f - dt) * angle), b = sinf(dt * angle), sn = sinf(angle);
        for (int k = 0; k < 4; ++k) q[k] = (a * x[k] + b * z[k]) / sn;
    }
    float Tm[16], Rm[16], Sm[16], TR[16];
    for (int k = 0; k < 16; ++k) Tm[k] = Rm[k] = Sm[k] = (k % 5 == 0) ? 1.0f : 0.0f;
    for (int r = 0; r < 4; ++r) {
        const float i0 = r == 0 ? 1.0f : 0.0f, i1 = r == 1 ? 1.0f : 0.0f, i2 = r == 2 ? 1.0f : 0.0f,
                    i3 = r == 3 ? 1.0f : 0.0f;
        Tm[12 + r] = ((i0 * T[0] + i1 * T[1]) + i2 * T[2]) + i3;
    }
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 4; ++r) Sm[c * 4 + r] = ((c == r) ? 1.0f : 0.0f) * S[c];
    const float qxx = q[0] * q[0], qyy = q[1] * q[1], qzz = q[2] * q[2];
    const float qxz = q[0] * q[2], qxy = q[0] * q[1], qyz = q[1] * q[2];
    const float qwx = q[3] * q[0], qwy = q[3] * q[1], qwz = q[3] * q[2];
    Rm[0] = 1.0f - 2.0f * (qyy + qzz);
    Rm[1] = 2.0f * (qxy + qwz);
    Rm[2] = 2.0f * (qxz - qwy);
    Rm[4] = 2.0f * (qxy - qwz);
    Rm[5] = 1.0f - 2.0f * (qxx + qzz);
    Rm[6] = 2.0f * (qyz + qwx);
    Rm[8] = 2.0f * (qxz + qwy);
    Rm[9] = 2.0f * (qyz - qwx);
    Rm[10] = 1.0f - 2.0f * (qxx + qyy);
    gndMul44(Tm, Rm, TR);
    gndMul44(TR, Sm, out);
}

// glm::inverse of a 4x4 matrix
GND_HD void gndInverse44(const float* M, float* out) {
#define GM(c, r) M[(c) * 4 + (r)]
    const float c00 = GM(2, 2) * GM(3, 3) - GM(3, 2) * GM(2, 3), c02 = GM(1, 2) * GM(3, 3) - GM(3, 2) * GM(1, 3);
    const float c03 = GM(1, 2) * GM(2, 3) - GM(2, 2) * GM(1, 3), c04 = GM(2, 1) * GM(3, 3) - GM(3, 1) * GM(2, 3);
    const float c06 = GM(1, 1) * GM(3, 3) - GM(3, 1) * GM(1, 3), c07 = GM(1, 1) * GM(2, 3) - GM(2, 1) * GM(1, 3);
    const float c08 = GM(2, 1) * GM(3, 2) - GM(3, 1) * GM(2, 2), c10 = GM(1, 1) * GM(3, 2) - GM(3, 1) * GM(1, 2);
    const float c11 = GM(1, 1) * GM(2, 2) - GM(2, 1) * GM(1, 2), c12 = GM(2, 0) * GM(3, 3) - GM(3, 0) * GM(2, 3);
    const float c14 = GM(1, 0) * GM(3, 3) - GM(3, 0) * GM(1, 3), c15 = GM(1, 0) * GM(2, 3) - GM(2, 0) * GM(1, 3);
    const float c16 = GM(2, 0) * GM(3, 2) - GM(3, 0) * GM(2, 2), c18 = GM(1, 0) * GM(3, 2) - GM(3, 0) * GM(1, 2);
    const float c19 = GM(1, 0) * GM(2, 2) - GM(2, 0) * GM(1, 2), c20 = GM(2, 0) * GM(3, 1) - GM(3, 0) * GM(2, 1);
    const float c22 = GM(1, 0) * GM(3, 1) - GM(3, 0) * GM(1, 1), c23 = GM(1, 0) * GM(2, 1) - GM(2, 0) * GM(1, 1);
    const float f0[4] = {c00, c00, c02, c03}, f1[4] = {c04, c04, c06, c07}, f2[4] = {c08, c08, c10, c11};
    const float f3[4] = {c12, c12, c14, c15}, f4[4] = {c16, c16, c18, c19}, f5[4] = {c20, c20, c22, c23};
    const float v0[4] = {GM(1, 0), GM(0, 0), GM(0, 0), GM(0, 0)}, v1[4] = {GM(1, 1), GM(0, 1), GM(0, 1), GM(0, 1)};
    const float v2[4] = {GM(1, 2), GM(0, 2), GM(0, 2), GM(0, 2)}, v3[4] = {GM(1, 3), GM(0, 3), GM(0, 3), GM(0, 3)};
    float inv[16];
    for (int k = 0; k < 4; ++k) {
        const float sa = (k % 2 == 0) ? 1.0f : -1.0f;
        inv[0 * 4 + k] = ((v1[k] * f0[k] - v2[k] * f1[k]) + v3[k] * f2[k]) * sa;
        inv[1 * 4 + k] = ((v0[k] * f0[k] - v2[k] * f3[k]) + v3[k] * f4[k]) * -sa;
        inv[2 * 4 + k] = ((v0[k] * f1[k] - v1[k] * f3[k]) + v3[k] * f5[k]) * sa;
        inv[3 * 4 + k] = ((v0[k] * f2[k] - v1[k] * f4[k]) + v2[k] * f5[k]) * -sa;
    }
    const float d0 = GM(0, 0) * inv[0], d1 = GM(0, 1) * inv[4], d2 = GM(0, 2) * inv[8], d3 = GM(0, 3) * inv[12];
    const float one = 1.0f / ((d0 + d1) + (d2 + d3));
    for (int k = 0; k < 16; ++k) out[k] = inv[k] * one;
#undef GM
}
