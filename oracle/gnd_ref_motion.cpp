// gnd_ref_motion.cpp — AnimatedTransform (src/math/animated.cpp) restated.
//
// TEST INFRASTRUCTURE, NOT PRODUCT CODE (see gnd_ref.h).  The reference
// decomposes both keyframes with glm::decompose and interpolates with
// glm::mix / glm::slerp / glm::mat4_cast; glm is not available here, so those
// are restated from glm's published sources (gtx/matrix_decompose.inl,
// gtc/quaternion.inl, ext/matrix_transform.inl, the mat4 product):
//   decompose:  translation = column 3; scale / shear by Gram-Schmidt over
//               columns 0..2 (length = sqrt((x x + y y) + z z), normalise =
//               v * 1 / length); a handedness flip negates scale and columns;
//               quaternion from the orthonormal columns by the trace method
//   mix(x,y,a) = x * (1 - a) + y * a
//   slerp:      dot = (w w' + x x') + (y y' + z z'); negate y when dot < 0;
//               linear mix when dot > 1 - FLT_EPSILON, else
//               (sin((1-a)t) x + sin(a t) y) / sin(t), t = acos(dot)
//   M = translate(I, T) * mat4_cast(R) * scale(I, S), each product with
//       columns ((A0 b0 + A1 b1) + A2 b2) + A3 b3
// Parity with glm's own arithmetic is unpinned (glm absent); the closed forms
// in tests/test_oracle_pins.py (decompose of known T R S, keyframes at the
// ends, half-way rotation angle) pin what they can.
#include <cmath>
#include <cstring>

#include "gnd_ref.h"

namespace gref {
namespace {

void decompose(const float* m, float T[3], float Q[4], float S[3]) {
    const float w = m[15];
    float L[16];
    for (int k = 0; k < 16; ++k) L[k] = m[k] / w;
    for (int k = 0; k < 3; ++k) T[k] = L[12 + k];
    float col[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) col[i][j] = L[i * 4 + j];
    auto len = [](const float* v) { return std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]); };
    auto dot = [](const float* a, const float* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; };
    auto normalise = [&](float* v) {
        const float l = len(v);
        for (int k = 0; k < 3; ++k) v[k] = v[k] * 1.0f / l;
    };
    auto subtract = [](float* a, const float* b, float s) {
        for (int k = 0; k < 3; ++k) a[k] = a[k] * 1.0f + b[k] * s;
    };
    float shear[3];
    S[0] = len(col[0]);
    normalise(col[0]);
    shear[2] = dot(col[0], col[1]);
    subtract(col[1], col[0], -shear[2]);
    S[1] = len(col[1]);
    normalise(col[1]);
    shear[1] = dot(col[0], col[2]);
    subtract(col[2], col[0], -shear[1]);
    shear[0] = dot(col[1], col[2]);
    subtract(col[2], col[1], -shear[0]);
    S[2] = len(col[2]);
    normalise(col[2]);
    const float x[3] = {col[1][1] * col[2][2] - col[2][1] * col[1][2], col[1][2] * col[2][0] - col[2][2] * col[1][0],
                        col[1][0] * col[2][1] - col[2][0] * col[1][1]};
    if (dot(col[0], x) < 0.0f) {
        for (int i = 0; i < 3; ++i) {
            S[i] *= -1.0f;
            for (int k = 0; k < 3; ++k) col[i][k] *= -1.0f;
        }
    }
    const float tr = col[0][0] + col[1][1] + col[2][2];
    if (tr > 0.0f) {
        float root = std::sqrt(tr + 1.0f);
        Q[3] = 0.5f * root;
        root = 0.5f / root;
        Q[0] = root * (col[1][2] - col[2][1]);
        Q[1] = root * (col[2][0] - col[0][2]);
        Q[2] = root * (col[0][1] - col[1][0]);
    } else {
        const int next[3] = {1, 2, 0};
        int i = 0;
        if (col[1][1] > col[0][0]) i = 1;
        if (col[2][2] > col[i][i]) i = 2;
        const int j = next[i], k = next[j];
        float root = std::sqrt(col[i][i] - col[j][j] - col[k][k] + 1.0f);
        Q[i] = 0.5f * root;
        root = 0.5f / root;
        Q[j] = root * (col[i][j] + col[j][i]);
        Q[k] = root * (col[i][k] + col[k][i]);
        Q[3] = root * (col[j][k] - col[k][j]);
    }
}

float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }

M4 product(const M4& A, const M4& B) {
    M4 R;
    for (int i = 0; i < 4; ++i)
        for (int r = 0; r < 4; ++r)
            R.m[i * 4 + r] = ((A.m[0 * 4 + r] * B.m[i * 4 + 0] + A.m[1 * 4 + r] * B.m[i * 4 + 1]) +
                              A.m[2 * 4 + r] * B.m[i * 4 + 2]) +
                             A.m[3 * 4 + r] * B.m[i * 4 + 3];
    return R;
}

M4 identity() {
    M4 I;
    for (int k = 0; k < 16; ++k) I.m[k] = (k % 5 == 0) ? 1.0f : 0.0f;
    return I;
}

}  // namespace

Motion motionFromMatrices(const float* start, const float* end) {
    Motion mo;
    std::memcpy(mo.start.m, start, sizeof(mo.start.m));
    std::memcpy(mo.end.m, end, sizeof(mo.end.m));
    mo.animated = std::memcmp(start, end, sizeof(mo.start.m)) != 0;  // Transform != compares m
    if (mo.animated) {
        decompose(start, mo.T[0], mo.R[0], mo.S[0]);
        decompose(end, mo.T[1], mo.R[1], mo.S[1]);
    }
    return mo;
}

// AnimatedTransform::interpolate (animated.cpp:30-47), keyframes at 0 and 1
M4 motionInterpolate(const Motion& mo, float time) {
    if (!mo.animated || time <= 0.0f) return mo.start;
    if (time >= 1.0f) return mo.end;
    const float dt = (time - 0.0f) / (1.0f - 0.0f);
    float T[3], S[3], q[4];
    for (int k = 0; k < 3; ++k) {
        T[k] = mixf(mo.T[0][k], mo.T[1][k], dt);
        S[k] = mixf(mo.S[0][k], mo.S[1][k], dt);
    }
    const float* x = mo.R[0];
    float z[4] = {mo.R[1][0], mo.R[1][1], mo.R[1][2], mo.R[1][3]};
    float cosT = (x[3] * z[3] + x[0] * z[0]) + (x[1] * z[1] + x[2] * z[2]);
    if (cosT < 0.0f) {
        for (float& c : z) c = -c;
        cosT = -cosT;
    }
    if (cosT > 1.0f - 1.1920929e-7f) {
        for (int k = 0; k < 4; ++k) q[k] = mixf(x[k], z[k], dt);
    } else {
        const float angle = std::acos(cosT);
        const float a = std::sin((1.0f - dt) * angle), b = std::sin(dt * angle), sn = std::sin(angle);
        for (int k = 0; k < 4; ++k) q[k] = (a * x[k] + b * z[k]) / sn;
    }
    M4 Tm = identity(), Sm = identity(), Rm = identity();
    for (int r = 0; r < 4; ++r) {
        const float i0 = r == 0 ? 1.0f : 0.0f, i1 = r == 1 ? 1.0f : 0.0f, i2 = r == 2 ? 1.0f : 0.0f,
                    i3 = r == 3 ? 1.0f : 0.0f;
        Tm.m[12 + r] = ((i0 * T[0] + i1 * T[1]) + i2 * T[2]) + i3;
    }
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 4; ++r) Sm.m[c * 4 + r] = ((c == r) ? 1.0f : 0.0f) * S[c];
    const float qxx = q[0] * q[0], qyy = q[1] * q[1], qzz = q[2] * q[2];
    const float qxz = q[0] * q[2], qxy = q[0] * q[1], qyz = q[1] * q[2];
    const float qwx = q[3] * q[0], qwy = q[3] * q[1], qwz = q[3] * q[2];
    Rm.m[0] = 1.0f - 2.0f * (qyy + qzz);
    Rm.m[1] = 2.0f * (qxy + qwz);
    Rm.m[2] = 2.0f * (qxz - qwy);
    Rm.m[4] = 2.0f * (qxy - qwz);
    Rm.m[5] = 1.0f - 2.0f * (qxx + qzz);
    Rm.m[6] = 2.0f * (qyz + qwx);
    Rm.m[8] = 2.0f * (qxz + qwy);
    Rm.m[9] = 2.0f * (qyz - qwx);
    Rm.m[10] = 1.0f - 2.0f * (qxx + qyy);
    return product(product(Tm, Rm), Sm);
}

// glm::inverse (detail::compute_inverse<4, 4>): the cofactor form, columns
// scaled by alternating signs, then times 1 / determinant
M4 motionInverse(const M4& mat) {
    const float* m = mat.m;
    auto e = [&](int c, int r) { return m[c * 4 + r]; };
    const float c00 = e(2, 2) * e(3, 3) - e(3, 2) * e(2, 3), c02 = e(1, 2) * e(3, 3) - e(3, 2) * e(1, 3);
    const float c03 = e(1, 2) * e(2, 3) - e(2, 2) * e(1, 3), c04 = e(2, 1) * e(3, 3) - e(3, 1) * e(2, 3);
    const float c06 = e(1, 1) * e(3, 3) - e(3, 1) * e(1, 3), c07 = e(1, 1) * e(2, 3) - e(2, 1) * e(1, 3);
    const float c08 = e(2, 1) * e(3, 2) - e(3, 1) * e(2, 2), c10 = e(1, 1) * e(3, 2) - e(3, 1) * e(1, 2);
    const float c11 = e(1, 1) * e(2, 2) - e(2, 1) * e(1, 2), c12 = e(2, 0) * e(3, 3) - e(3, 0) * e(2, 3);
    const float c14 = e(1, 0) * e(3, 3) - e(3, 0) * e(1, 3), c15 = e(1, 0) * e(2, 3) - e(2, 0) * e(1, 3);
    const float c16 = e(2, 0) * e(3, 2) - e(3, 0) * e(2, 2), c18 = e(1, 0) * e(3, 2) - e(3, 0) * e(1, 2);
    const float c19 = e(1, 0) * e(2, 2) - e(2, 0) * e(1, 2), c20 = e(2, 0) * e(3, 1) - e(3, 0) * e(2, 1);
    const float c22 = e(1, 0) * e(3, 1) - e(3, 0) * e(1, 1), c23 = e(1, 0) * e(2, 1) - e(2, 0) * e(1, 1);
    const float fac[6][4] = {{c00, c00, c02, c03}, {c04, c04, c06, c07}, {c08, c08, c10, c11},
                             {c12, c12, c14, c15}, {c16, c16, c18, c19}, {c20, c20, c22, c23}};
    float vec[4][4];
    for (int r = 0; r < 4; ++r) {
        vec[r][0] = e(1, r);
        vec[r][1] = vec[r][2] = vec[r][3] = e(0, r);
    }
    M4 inv;
    for (int k = 0; k < 4; ++k) {
        const float sign = (k % 2 == 0) ? 1.0f : -1.0f;
        inv.m[0 * 4 + k] = ((vec[1][k] * fac[0][k] - vec[2][k] * fac[1][k]) + vec[3][k] * fac[2][k]) * sign;
        inv.m[1 * 4 + k] = ((vec[0][k] * fac[0][k] - vec[2][k] * fac[3][k]) + vec[3][k] * fac[4][k]) * -sign;
        inv.m[2 * 4 + k] = ((vec[0][k] * fac[1][k] - vec[1][k] * fac[3][k]) + vec[3][k] * fac[5][k]) * sign;
        inv.m[3 * 4 + k] = ((vec[0][k] * fac[2][k] - vec[1][k] * fac[4][k]) + vec[2][k] * fac[5][k]) * -sign;
    }
    const float det = (e(0, 0) * inv.m[0] + e(0, 1) * inv.m[4]) + (e(0, 2) * inv.m[8] + e(0, 3) * inv.m[12]);
    const float one = 1.0f / det;
    for (float& v : inv.m) v = v * one;
    return inv;
}

void primAt(const Prim& p, float time, M4& W, M4& Wi) {
    if (!p.motion.animated || time <= 0.0f) {
        W = p.toWorld;
        Wi = p.toWorldInv;
    } else if (time >= 1.0f) {
        W = p.motion.end;
        Wi = p.endInv;
    } else {
        W = motionInterpolate(p.motion, time);
        Wi = motionInverse(W);
    }
}

}  // namespace gref

using namespace gref;

extern "C" {
void gnd_ref_motion_matrix(const float* start, const float* end, float time, float* out) {
    const M4 m = motionInterpolate(motionFromMatrices(start, end), time);
    std::memcpy(out, m.m, sizeof(m.m));
}
void gnd_ref_decompose(const float* m, float* T, float* Q, float* S) { decompose(m, T, Q, S); }
}
