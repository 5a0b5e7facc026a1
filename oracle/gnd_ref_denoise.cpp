// gnd_ref_denoise.cpp — CPU restatement of the film denoiser filter.
//
// TEST INFRASTRUCTURE, NOT PRODUCT CODE (see gnd_ref.h).  The reference
// denoises with Intel Open Image Denoise (src/core/denoiser.cpp:6-55), a
// trained network that is not available here; the GPU library replaces it
// with an edge-avoiding a-trous wavelet filter (Dammertz et al., HPG 2010)
// on albedo-demodulated radiance.  This file restates that filter, step for
// step and in the same float evaluation order, so the HIP kernels can be
// checked against it.  It is a restatement of this build's filter, not of
// the reference: parity with OIDN is unpinned.
#include <cmath>
#include <cstddef>
#include <vector>

namespace {

struct F4 {
    float x, y, z, w;
};

float lum3(float r, float g, float b) { return r * 0.212671f + g * 0.715160f + b * 0.072169f; }

constexpr float kDemodEps = 1e-3f;
constexpr float kInvSigmaA2 = 100.0f;
constexpr float kSigmaL = 4.0f;
constexpr int kPasses = 5;

}  // namespace

extern "C" void gnd_ref_denoise(const float* color, const float* albedo, const float* normal, int W, int H,
                                float* out) {
    const size_t np = (size_t)W * H;
    std::vector<F4> E(np), E2(np), G1(np), G2(np);
    // demodulate, pack the guides
    for (size_t p = 0; p < np; ++p) {
        float ar = 1.0f, ag = 1.0f, ab = 1.0f;
        float er = color[3 * p], eg = color[3 * p + 1], eb = color[3 * p + 2];
        if (albedo) {
            ar = albedo[3 * p];
            ag = albedo[3 * p + 1];
            ab = albedo[3 * p + 2];
            er = color[3 * p] / (ar + kDemodEps);
            eg = color[3 * p + 1] / (ag + kDemodEps);
            eb = color[3 * p + 2] / (ab + kDemodEps);
        }
        float nx = 0.0f, ny = 0.0f, nz = 0.0f, valid = 0.0f;
        if (normal) {
            const float vx = normal[3 * p], vy = normal[3 * p + 1], vz = normal[3 * p + 2];
            if (vx != 0.0f || vy != 0.0f || vz != 0.0f) {
                nx = 2.0f * vx - 1.0f;
                ny = 2.0f * vy - 1.0f;
                nz = 2.0f * vz - 1.0f;
                const float l2 = (nx * nx + ny * ny) + nz * nz;
                if (l2 > 0.0f) {
                    const float inv = 1.0f / std::sqrt(l2);
                    nx *= inv;
                    ny *= inv;
                    nz *= inv;
                    valid = 1.0f;
                } else {
                    nx = ny = nz = 0.0f;
                }
            }
        }
        E[p] = {er, eg, eb, 0.0f};
        G1[p] = {ar, ag, ab, valid};
        G2[p] = {nx, ny, nz, 0.0f};
    }
    // luminance variance over the 5x5 window
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float s = 0.0f, s2 = 0.0f;
            int n = 0;
            for (int dy = -2; dy <= 2; ++dy) {
                const int qy = y + dy;
                if (qy < 0 || qy >= H) continue;
                for (int dx = -2; dx <= 2; ++dx) {
                    const int qx = x + dx;
                    if (qx < 0 || qx >= W) continue;
                    const F4& e = E[(size_t)qy * W + qx];
                    const float l = lum3(e.x, e.y, e.z);
                    s += l;
                    s2 += l * l;
                    ++n;
                }
            }
            const float inv = 1.0f / (float)n;
            const float m = s * inv;
            float var = s2 * inv - m * m;
            var = var > 0.0f ? var : 0.0f;
            const F4& c = E[(size_t)y * W + x];
            E2[(size_t)y * W + x] = {c.x, c.y, c.z, var};
        }
    const float h[5] = {1.0f / 16.0f, 1.0f / 4.0f, 3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};
    std::vector<F4>* in = &E2;
    std::vector<F4>* o = &E;
    for (int it = 0; it < kPasses; ++it) {
        const int step = 1 << it;
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const size_t p = (size_t)y * W + x;
                const F4 ep = (*in)[p], gp = G1[p], nrm = G2[p];
                const float lp = lum3(ep.x, ep.y, ep.z);
                const float sigma = kSigmaL * std::sqrt(ep.w);
                float sr = 0.0f, sg = 0.0f, sb = 0.0f, sv = 0.0f, sw = 0.0f;
                for (int j = 0; j < 5; ++j) {
                    const int qy = y + (j - 2) * step;
                    if (qy < 0 || qy >= H) continue;
                    for (int i = 0; i < 5; ++i) {
                        const int qx = x + (i - 2) * step;
                        if (qx < 0 || qx >= W) continue;
                        const size_t q = (size_t)qy * W + qx;
                        const F4 eq = (*in)[q], gq = G1[q], nq = G2[q];
                        float wn;
                        if (gp.w > 0.0f && gq.w > 0.0f) {
                            float c = (nrm.x * nq.x + nrm.y * nq.y) + nrm.z * nq.z;
                            c = c > 0.0f ? c : 0.0f;
                            for (int k = 0; k < 7; ++k) c = c * c;
                            wn = c;
                        } else {
                            wn = gp.w == gq.w ? 1.0f : 0.0f;
                        }
                        float wa = 1.0f;
                        if (albedo) {
                            const float dr = gp.x - gq.x, dg = gp.y - gq.y, db = gp.z - gq.z;
                            wa = std::exp(-((dr * dr + dg * dg) + db * db) * kInvSigmaA2);
                        }
                        const float lq = lum3(eq.x, eq.y, eq.z);
                        const float wl = sigma > 0.0f ? std::exp(-std::fabs(lp - lq) / sigma) : (lp == lq ? 1.0f : 0.0f);
                        const float w = ((h[i] * h[j]) * wn) * wa * wl;
                        sr += w * eq.x;
                        sg += w * eq.y;
                        sb += w * eq.z;
                        sv += (w * w) * eq.w;
                        sw += w;
                    }
                }
                const float inv = 1.0f / sw;
                (*o)[p] = {sr * inv, sg * inv, sb * inv, (sv * inv) * inv};
            }
        std::swap(in, o);
    }
    for (size_t p = 0; p < np; ++p) {
        const F4& e = (*in)[p];
        if (albedo) {
            out[3 * p] = e.x * (albedo[3 * p] + kDemodEps);
            out[3 * p + 1] = e.y * (albedo[3 * p + 1] + kDemodEps);
            out[3 * p + 2] = e.z * (albedo[3 * p + 2] + kDemodEps);
        } else {
            out[3 * p] = e.x;
            out[3 * p + 1] = e.y;
            out[3 * p + 2] = e.z;
        }
    }
}
