// integrator.h — the reference's Integrator / Sampler / Film plugin
// interfaces on the host, with the GPU integrators behind them.
//
//   Integrator      include/gianduia/core/integrator.h:18-40 (render, cancel,
//                   preprocess, setCallback with (sampleIndex, film))
//   Sampler         include/gianduia/core/sampler.h:8-27 (clone, seed,
//                   next1D, next2D, getSampleCount)
//   IndependentSampler  src/samplers/independent.cpp:7-47
//   Film            include/gianduia/core/film.h:26-69 (accumulate, resolve)
//   registration    include/gianduia/core/factory.h:26-36 (by XML name;
//                   a duplicate name overrides with a warning)
//
// The GPU integrators render whole passes on the device through the C ABI
// (gnd_hip_render), accumulating into one device film, and copy its sums into
// the host Film between callbacks.
#pragma once

#include <atomic>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/gnd_hip.h"

namespace gnd {

class Sampler {
public:
    virtual ~Sampler() = default;
    virtual std::unique_ptr<Sampler> clone() const = 0;
    virtual void seed(uint64_t sampleIndex) = 0;
    virtual float next1D() = 0;
    virtual std::pair<float, float> next2D() = 0;
    virtual size_t getSampleCount() const = 0;
};

// PCG32 seeded with splitMix64(index) and stream 1; floats are u32 * 2^-32
// (so 1.0f is possible, pcg32.h:29-31).  The device sampler draws the same
// sequence per (pixel, pass).
class IndependentSampler : public Sampler {
public:
    explicit IndependentSampler(size_t spp) : spp_(spp) {}
    std::unique_ptr<Sampler> clone() const override;
    void seed(uint64_t sampleIndex) override;
    float next1D() override;
    std::pair<float, float> next2D() override;
    size_t getSampleCount() const override { return spp_; }

private:
    uint32_t nextU32();
    size_t spp_;
    uint64_t state_ = 0, inc_ = 3;
};

// Host film: weighted RGB sums and weights per pixel (what the device
// gathers per pass), resolved to RGB = sum / weight; optionally the AOV sums
// (AOVRecord, film.h:13-19), resolved the same way (film.cpp:50-95).
class Film {
public:
    static constexpr int kAovs = 9;  // albedo rgb, normal xyz, depth, roughness, metallic
    Film() = default;
    Film(int w, int h) { reset(w, h); }
    void reset(int w, int h, bool aovs = false);
    void clear();
    // adds device-format accumulators (w*h*4 floats: r, g, b, weight)
    void accumulate(const float* rgbw);
    // takes a device film's sums as they are (w*h*4 floats), and its AOV sums (w*h*9)
    void setSums(const float* rgbw);
    void setAovSums(const float* sums);
    void resolve();
    int width() const { return w_; }
    int height() const { return h_; }
    bool hasAovs() const { return !aovAcc_.empty(); }
    const std::vector<float>& rgb() const { return rgb_; }   // resolved, w*h*3
    const std::vector<float>& aovs() const { return aov_; }  // resolved, w*h*9 (where the
                                                              // weight is 0: 0,0,0, 0,0,0, -1, 0, 0)
    void savePFM(const std::string& path) const;
    void savePNG(const std::string& path) const;  // 8-bit sRGB, values clamped to [0, 1]
    void saveEXR(const std::string& path) const;  // float R, G, B (Image2D::saveEXR, image2d.h:62-97)
    // one resolved AOV image: "albedo" / "normal" (R, G, B) or "depth" /
    // "roughness" / "metallic" (Y), as ".exr" or ".pfm"
    void saveAov(const std::string& kind, const std::string& path) const;
    // overwrites the resolved radiance (w*h*3) until the next resolve(), as
    // Denoiser::execute writes into the film's radiance image
    void replaceRgb(const float* rgb) { rgb_.assign(rgb, rgb + (size_t)w_ * h_ * 3); }

private:
    int w_ = 0, h_ = 0;
    std::vector<float> acc_;  // r, g, b, weight sums per pixel (the device film)
    std::vector<float> rgb_;
    std::vector<float> aovAcc_;
    std::vector<float> aov_;
};

// Scanline OpenEXR file, uncompressed, 32-bit float channels (channels = 3:
// interleaved R, G, B input; 1: Y), rows top to bottom.
void writeEXR(const std::string& path, const float* px, int w, int h, int channels);
// PFM (little-endian, rows bottom to top), 3 channels ("PF") or 1 ("Pf")
void writePFM(const std::string& path, const float* px, int w, int h, int channels);
// Scanline OpenEXR file -> interleaved R, G, B floats over its data window,
// rows top to bottom (channels it lacks read as 0); NONE / RLE / ZIPS / ZIP
// compression, UINT / HALF / FLOAT channels.  Throws on anything else.
void readEXR(const std::string& path, int& width, int& height, std::vector<float>& rgb);

// Denoiser (include/gianduia/core/denoiser.h, src/core/denoiser.cpp:6-55):
// execute(film) denoises the film's resolved radiance in place on the GPU
// (gnd_hip_denoise), guided by the film's albedo and normal images when it
// has AOVs.
class Denoiser {
public:
    void execute(gnd_hip_ctx* ctx, Film* film);
};

// What an integrator renders: the scene description and its device copy.
struct RenderScene {
    const gnd_scene_desc* desc = nullptr;
    gnd_hip_scene* device = nullptr;
};

class Integrator {
public:
    using RenderCallback = std::function<void(int sampleIndex, const Film& film)>;
    virtual ~Integrator() = default;
    virtual void render(const RenderScene& scene) = 0;
    virtual void cancel() { stop_ = true; }
    virtual void preprocess(const RenderScene&) {}
    void setCallback(RenderCallback cb) { callback_ = std::move(cb); }
    // also fill the film's AOV images (Integrator::Li's AOVRecord argument)
    void setAovs(bool on) { aovs_ = on; }
    Film& film() { return film_; }
    int passesDone() const { return passes_; }

protected:
    void notifyUpdate(int sampleIndex) {
        if (callback_) callback_(sampleIndex, film_);
    }
    std::atomic<bool> stop_{false};
    RenderCallback callback_;
    Film film_;
    int passes_ = 0;
    bool aovs_ = false;
};

// path / volpath on the GPU: passes are rendered in batches of
// `passesPerCallback` (one gnd_hip_render each, seed offset = first pass *
// W*H so pass s draws the reference's stream pixelIdx + s*W*H, every batch
// after the first adding onto the device film), copied into the film,
// resolved, and reported to the callback.  cancel() stops at
// the next batch and also interrupts the batch in flight.
class HipIntegrator : public Integrator {
public:
    HipIntegrator(gnd_hip_ctx* ctx, int integrator, int maxDepth, int passesPerCallback)
        : ctx_(ctx), integrator_(integrator), maxDepth_(maxDepth), batch_(passesPerCallback) {}
    void render(const RenderScene& scene) override;
    void cancel() override;

private:
    gnd_hip_ctx* ctx_;
    int integrator_, maxDepth_, batch_;
};

struct IntegratorParams {
    gnd_hip_ctx* ctx = nullptr;
    int maxDepth = 1000;  // path.cpp:9, volPath.cpp:12
    int passesPerCallback = 1;
};

class IntegratorFactory {
public:
    using Creator = std::function<std::unique_ptr<Integrator>(const IntegratorParams&)>;
    static void registerClass(const std::string& name, Creator c);
    static std::unique_ptr<Integrator> create(const std::string& name, const IntegratorParams& p);
    static std::vector<std::string> names();
};

// Registers "path", "volpath", "direct", "ao" and their "_hip" aliases (idempotent;
// called by the factory before its first use, so static-library linking
// cannot drop the registrations).
void registerBuiltinIntegrators();

}  // namespace gnd
