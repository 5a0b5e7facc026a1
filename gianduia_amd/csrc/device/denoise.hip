// denoise.hip — film denoiser behind gnd_hip_denoise.
//
// The reference denoises the resolved film with Intel Open Image Denoise
// (Denoiser::execute, src/core/denoiser.cpp:6-55): filter "RT" on the
// radiance with the albedo and normal AOVs as guides, hdr = true, the result
// written back over the radiance.  OIDN is a trained network whose weights
// ship only inside its binary library, which is not part of this build, so
// this is a different filter with the same interface and role: the
// edge-avoiding a-trous wavelet filter of Dammertz, Sewtz, Hanika and Lensch
// ("Edge-Avoiding A-Trous Wavelet Transform for fast Global Illumination
// Filtering", HPG 2010) on albedo-demodulated radiance, with the luminance
// edge-stopping scaled by a per-pixel variance estimate that is propagated
// through the passes (as in SVGF, Schied et al., HPG 2017).
//
//   prep:   E = radiance / (albedo + 1e-3) per channel (radiance itself
//           without an albedo guide); guides packed as float4 (albedo, normal
//           validity) and (unit normal).
//   var:    luminance variance of E over the 5x5 window (clamped at the
//           borders) into E.w.
//   a-trous: five passes with steps 1, 2, 4, 8, 16: 5x5 B3-spline taps
//           (1/16, 1/4, 3/8, 1/4, 1/16) times the edge-stopping weights
//             normal:    max(0, n_p . n_q)^128 (both valid; else 1 when both
//                        are invalid, 0 when one is)
//             albedo:    exp(-|a_p - a_q|^2 / 0.1^2)
//             luminance: exp(-|l_p - l_q| / (4 sqrt(var_p))), or [l_p == l_q]
//                        when var_p = 0;
//           the variance is filtered with the squared weights.
//   finish: radiance = E * (albedo + 1e-3).
// Every weight depends on ratios of radiance values only, so scaling the
// input by a power of two scales the output by exactly the same factor (an
// HDR filter, as hdr = true asks of OIDN).
//
// Memory: one thread per pixel, float4 records.  An a-trous pass stages each
// 16 x 16 tile of a sub-lattice (pixels `step` apart) with its halo in LDS, so
// the 25 taps of a pixel come from LDS (the whole 1080p working set, 130 MB,
// stays in the Infinity Cache between passes).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "denoise.h"

namespace {

constexpr float kDemodEps = 1e-3f;
constexpr float kInvSigmaA2 = 100.0f;  // 1 / 0.1^2
constexpr float kSigmaL = 4.0f;

__device__ __forceinline__ float lum3(float r, float g, float b) { return r * 0.212671f + g * 0.715160f + b * 0.072169f; }

__global__ __launch_bounds__(256) void kDnPrep(const float* __restrict__ color, const float* __restrict__ albedo,
                                               const float* __restrict__ normal, int W, int H,
                                               float4* __restrict__ E, float4* __restrict__ G1,
                                               float4* __restrict__ G2) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    const size_t p = (size_t)y * W + x;
    const float cr = color[3 * p], cg = color[3 * p + 1], cb = color[3 * p + 2];
    float ar = 1.0f, ag = 1.0f, ab = 1.0f;
    float er = cr, eg = cg, eb = cb;
    if (albedo) {
        ar = albedo[3 * p];
        ag = albedo[3 * p + 1];
        ab = albedo[3 * p + 2];
        er = cr / (ar + kDemodEps);
        eg = cg / (ag + kDemodEps);
        eb = cb / (ab + kDemodEps);
    }
    float nx = 0.0f, ny = 0.0f, nz = 0.0f, valid = 0.0f;
    if (normal) {
        const float vx = normal[3 * p], vy = normal[3 * p + 1], vz = normal[3 * p + 2];
        if (vx != 0.0f || vy != 0.0f || vz != 0.0f) {  // the AOV's miss value is 0
            nx = 2.0f * vx - 1.0f;
            ny = 2.0f * vy - 1.0f;
            nz = 2.0f * vz - 1.0f;
            const float l2 = (nx * nx + ny * ny) + nz * nz;
            if (l2 > 0.0f) {
                const float inv = 1.0f / __builtin_sqrtf(l2);
                nx *= inv;
                ny *= inv;
                nz *= inv;
                valid = 1.0f;
            } else {
                nx = ny = nz = 0.0f;
            }
        }
    }
    E[p] = float4{er, eg, eb, 0.0f};
    G1[p] = float4{ar, ag, ab, valid};
    G2[p] = float4{nx, ny, nz, 0.0f};
}

__global__ __launch_bounds__(256) void kDnVariance(const float4* __restrict__ Ein, int W, int H,
                                                   float4* __restrict__ Eout) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    float s = 0.0f, s2 = 0.0f;
    int n = 0;
    for (int dy = -2; dy <= 2; ++dy) {
        const int qy = y + dy;
        if (qy < 0 || qy >= H) continue;
        for (int dx = -2; dx <= 2; ++dx) {
            const int qx = x + dx;
            if (qx < 0 || qx >= W) continue;
            const float4 e = Ein[(size_t)qy * W + qx];
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
    const float4 c = Ein[(size_t)y * W + x];
    Eout[(size_t)y * W + x] = float4{c.x, c.y, c.z, var};
}

// One a-trous pass over the sub-lattice of pixels congruent to (ox, oy)
// modulo `step`: a block takes a 16 x 16 tile of that lattice (pixels step
// apart), stages the tile and its 2-point halo (20 x 20 records of radiance
// and both guides, 19 KB) in LDS once, and every pixel then reads its 25 taps
// from LDS instead of from L2.  The taps, their order and the arithmetic are
// those of the per-pixel gather.
constexpr int kDnTile = 16, kDnSpan = kDnTile + 4;

__global__ __launch_bounds__(256) void kDnAtrous(const float4* __restrict__ Ein, const float4* __restrict__ G1,
                                                 const float4* __restrict__ G2, int W, int H, int step,
                                                 int useAlbedo, float4* __restrict__ Eout) {
    __shared__ float4 sE[kDnSpan * kDnSpan], sG[kDnSpan * kDnSpan], sN[kDnSpan * kDnSpan];
    const int ox = (int)(blockIdx.x % (unsigned)step), tx = (int)(blockIdx.x / (unsigned)step);
    const int oy = (int)(blockIdx.y % (unsigned)step), ty = (int)(blockIdx.y / (unsigned)step);
    for (int k = threadIdx.x; k < kDnSpan * kDnSpan; k += 256) {
        const int px = ox + (tx * kDnTile - 2 + k % kDnSpan) * step;
        const int py = oy + (ty * kDnTile - 2 + k / kDnSpan) * step;
        if (px >= 0 && px < W && py >= 0 && py < H) {
            const size_t q = (size_t)py * W + px;
            sE[k] = Ein[q];
            sG[k] = G1[q];
            sN[k] = G2[q];
        }
    }
    __syncthreads();
    const int li = (int)(threadIdx.x & 15), lj = (int)(threadIdx.x >> 4);
    const int x = ox + (tx * kDnTile + li) * step, y = oy + (ty * kDnTile + lj) * step;
    if (x >= W || y >= H) return;
    const float h[5] = {1.0f / 16.0f, 1.0f / 4.0f, 3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};
    const int c = (lj + 2) * kDnSpan + (li + 2);
    const float4 ep = sE[c], gp = sG[c], np = sN[c];
    const float lp = lum3(ep.x, ep.y, ep.z);
    const float sigma = kSigmaL * __builtin_sqrtf(ep.w);
    float sr = 0.0f, sg = 0.0f, sb = 0.0f, sv = 0.0f, sw = 0.0f;
    for (int j = 0; j < 5; ++j) {
        const int qy = y + (j - 2) * step;
        if (qy < 0 || qy >= H) continue;
        for (int i = 0; i < 5; ++i) {
            const int qx = x + (i - 2) * step;
            if (qx < 0 || qx >= W) continue;
            const int k = (lj + j) * kDnSpan + (li + i);
            const float4 eq = sE[k], gq = sG[k], nq = sN[k];
            float wn;
            if (gp.w > 0.0f && gq.w > 0.0f) {
                float cs = (np.x * nq.x + np.y * nq.y) + np.z * nq.z;
                cs = cs > 0.0f ? cs : 0.0f;
#pragma unroll
                for (int r = 0; r < 7; ++r) cs = cs * cs;  // ^128
                wn = cs;
            } else {
                wn = gp.w == gq.w ? 1.0f : 0.0f;
            }
            float wa = 1.0f;
            if (useAlbedo) {
                const float dr = gp.x - gq.x, dg = gp.y - gq.y, db = gp.z - gq.z;
                wa = expf(-((dr * dr + dg * dg) + db * db) * kInvSigmaA2);
            }
            const float lq = lum3(eq.x, eq.y, eq.z);
            const float wl = sigma > 0.0f ? expf(-__builtin_fabsf(lp - lq) / sigma) : (lp == lq ? 1.0f : 0.0f);
            const float w = ((h[i] * h[j]) * wn) * wa * wl;
            sr += w * eq.x;
            sg += w * eq.y;
            sb += w * eq.z;
            sv += (w * w) * eq.w;
            sw += w;
        }
    }
    const float inv = 1.0f / sw;  // the centre tap alone is > 0
    Eout[(size_t)y * W + x] = float4{sr * inv, sg * inv, sb * inv, (sv * inv) * inv};
}

__global__ __launch_bounds__(256) void kDnFinish(const float4* __restrict__ E, const float* __restrict__ albedo,
                                                 int W, int H, float* __restrict__ out) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    const size_t p = (size_t)y * W + x;
    const float4 e = E[p];
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

}  // namespace

hipError_t gndDenoiseLaunch(hipStream_t stream, const float* color, const float* albedo, const float* normal, int W,
                            int H, float* out, float4* scratch, hipEvent_t passBegin, hipEvent_t passEnd) {
    const size_t np = (size_t)W * H;
    float4* E0 = scratch;
    float4* E1 = scratch + np;
    float4* G1 = scratch + 2 * np;
    float4* G2 = scratch + 3 * np;
    const dim3 grid((unsigned)((W + 15) / 16), (unsigned)((H + 15) / 16)), block(256);
    hipLaunchKernelGGL(kDnPrep, grid, block, 0, stream, color, albedo, normal, W, H, E0, G1, G2);
    hipLaunchKernelGGL(kDnVariance, grid, block, 0, stream, E0, W, H, E1);
    float4* in = E1;
    float4* o = E0;
    for (int it = 0; it < kDenoisePasses; ++it) {
        if (it == 0 && passBegin) (void)hipEventRecord(passBegin, stream);
        const int step = 1 << it;
        // step x step sub-lattices, each tiled 16 x 16 (blocks of a sub-lattice interleave)
        const unsigned lw = (unsigned)((W + step - 1) / step), lh = (unsigned)((H + step - 1) / step);
        const dim3 tgrid(((lw + kDnTile - 1) / kDnTile) * (unsigned)step, ((lh + kDnTile - 1) / kDnTile) * (unsigned)step);
        hipLaunchKernelGGL(kDnAtrous, tgrid, block, 0, stream, in, G1, G2, W, H, step, albedo ? 1 : 0, o);
        if (it == 0 && passEnd) (void)hipEventRecord(passEnd, stream);
        float4* t = in;
        in = o;
        o = t;
    }
    hipLaunchKernelGGL(kDnFinish, grid, block, 0, stream, in, albedo, W, H, out);
    return hipGetLastError();
}
