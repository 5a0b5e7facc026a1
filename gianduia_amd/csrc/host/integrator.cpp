// integrator.cpp — host Integrator / Sampler / Film plugin interfaces and the
// GPU path / volpath integrators behind them.
#include "integrator.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <stdexcept>

namespace gnd {

// ---------------------------------------------------------------- sampler
// SplitMix64 (independent.cpp:7-20) and PCG32 (pcg32.h:5-32)
namespace {
uint64_t splitMix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
constexpr uint64_t kMul = 6364136223846793005ULL;
}  // namespace

std::unique_ptr<Sampler> IndependentSampler::clone() const { return std::make_unique<IndependentSampler>(*this); }

void IndependentSampler::seed(uint64_t sampleIndex) {
    inc_ = (1ULL << 1u) | 1u;  // initseq 1
    state_ = 0;
    nextU32();
    state_ += splitMix64(sampleIndex);
    nextU32();
}

uint32_t IndependentSampler::nextU32() {
    const uint64_t old = state_;
    state_ = old * kMul + inc_;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
}

float IndependentSampler::next1D() { return std::ldexp((float)nextU32(), -32); }

std::pair<float, float> IndependentSampler::next2D() {
    // Point2f(nextFloat(), nextFloat()), evaluated left to right
    const float a = next1D();
    const float b = next1D();
    return {a, b};
}

// ---------------------------------------------------------------- film
void Film::reset(int w, int h, bool aovs) {
    w_ = w;
    h_ = h;
    acc_.assign((size_t)w * h * 4, 0.0);
    rgb_.assign((size_t)w * h * 3, 0.0f);
    aovAcc_.assign(aovs ? (size_t)w * h * kAovs : 0, 0.0);
    aov_.assign(aovs ? (size_t)w * h * kAovs : 0, 0.0f);
}

void Film::clear() {
    std::fill(acc_.begin(), acc_.end(), 0.0);
    std::fill(rgb_.begin(), rgb_.end(), 0.0f);
    std::fill(aovAcc_.begin(), aovAcc_.end(), 0.0);
    std::fill(aov_.begin(), aov_.end(), 0.0f);
}

void Film::accumulate(const float* rgbw) {
    for (size_t i = 0; i < acc_.size(); ++i) acc_[i] += rgbw[i];
}

void Film::setSums(const float* rgbw) { std::copy(rgbw, rgbw + acc_.size(), acc_.begin()); }

void Film::setAovSums(const float* sums) { std::copy(sums, sums + aovAcc_.size(), aovAcc_.begin()); }

// Film::resolve (src/core/film.cpp:50-94): value = sum * (1 / weight); where
// the weight is 0, black and the AOV defaults (depth -1)
void Film::resolve() {
    const size_t n = (size_t)w_ * h_;
    for (size_t i = 0; i < n; ++i) {
        const float wsum = acc_[4 * i + 3];
        const float inv = wsum > 0.0f ? 1.0f / wsum : 0.0f;
        for (int c = 0; c < 3; ++c) rgb_[3 * i + c] = acc_[4 * i + c] * inv;
        if (!aovAcc_.empty()) {
            for (int k = 0; k < kAovs; ++k) aov_[kAovs * i + k] = aovAcc_[kAovs * i + k] * inv;
            if (!(wsum > 0.0f)) aov_[kAovs * i + 6] = -1.0f;
        }
    }
}

void writePFM(const std::string& path, const float* px, int w, int h, int channels) {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    f << (channels == 3 ? "PF" : "Pf") << "\n" << w << " " << h << "\n-1.0\n";  // little-endian, bottom row first
    for (int y = h - 1; y >= 0; --y)
        f.write(reinterpret_cast<const char*>(px + (size_t)y * w * channels),
                (std::streamsize)w * channels * sizeof(float));
    if (!f) throw std::runtime_error("write failed: " + path);
}

void Film::savePFM(const std::string& path) const { writePFM(path, rgb_.data(), w_, h_, 3); }

namespace {
void putBytes(std::vector<unsigned char>& v, const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    v.insert(v.end(), b, b + n);
}
template <class T>
void putLE(std::vector<unsigned char>& v, T x) {  // the host is little-endian, as EXR is
    putBytes(v, &x, sizeof(T));
}
void putStr(std::vector<unsigned char>& v, const char* s) { putBytes(v, s, std::strlen(s) + 1); }
void attribute(std::vector<unsigned char>& v, const char* name, const char* type,
               const std::vector<unsigned char>& value) {
    putStr(v, name);
    putStr(v, type);
    putLE<int32_t>(v, (int32_t)value.size());
    v.insert(v.end(), value.begin(), value.end());
}
}  // namespace

// OpenEXR single-part scanline file: magic 20000630, version 2, the required
// header attributes, a table of scanline-block offsets, then one block per
// row (y, byte count, each channel's row in channel-name order).  Written
// uncompressed; OpenEXR readers take any compression the header names.
void writeEXR(const std::string& path, const float* px, int w, int h, int channels) {
    if (channels != 1 && channels != 3) throw std::runtime_error("EXR: 1 or 3 channels");
    std::vector<unsigned char> out;
    putLE<int32_t>(out, 20000630);
    putLE<int32_t>(out, 2);
    // channels, sorted by name (B, G, R), 32-bit float, sampling 1
    static const char* rgbNames[3] = {"B", "G", "R"};
    static const int rgbIndex[3] = {2, 1, 0};  // position of B, G, R in the input
    std::vector<unsigned char> ch;
    for (int c = 0; c < channels; ++c) {
        putStr(ch, channels == 3 ? rgbNames[c] : "Y");
        putLE<int32_t>(ch, 2);            // pixel type FLOAT
        putLE<uint8_t>(ch, 0);            // pLinear
        putLE<uint8_t>(ch, 0);
        putLE<uint8_t>(ch, 0);
        putLE<uint8_t>(ch, 0);            // reserved
        putLE<int32_t>(ch, 1);            // xSampling
        putLE<int32_t>(ch, 1);            // ySampling
    }
    ch.push_back(0);
    attribute(out, "channels", "chlist", ch);
    attribute(out, "compression", "compression", {0});  // NO_COMPRESSION
    std::vector<unsigned char> box;
    for (int32_t v : {0, 0, w - 1, h - 1}) putLE<int32_t>(box, v);
    attribute(out, "dataWindow", "box2i", box);
    attribute(out, "displayWindow", "box2i", box);
    attribute(out, "lineOrder", "lineOrder", {0});  // INCREASING_Y
    std::vector<unsigned char> f1, v2;
    putLE<float>(f1, 1.0f);
    attribute(out, "pixelAspectRatio", "float", f1);
    putLE<float>(v2, 0.0f);
    putLE<float>(v2, 0.0f);
    attribute(out, "screenWindowCenter", "v2f", v2);
    attribute(out, "screenWindowWidth", "float", f1);
    out.push_back(0);  // end of header
    const size_t rowBytes = (size_t)w * channels * sizeof(float);
    const size_t table = out.size();
    const size_t first = table + (size_t)h * sizeof(uint64_t);
    for (int y = 0; y < h; ++y) putLE<uint64_t>(out, (uint64_t)(first + (size_t)y * (8 + rowBytes)));
    std::vector<float> row((size_t)w);
    for (int y = 0; y < h; ++y) {
        putLE<int32_t>(out, y);
        putLE<int32_t>(out, (int32_t)rowBytes);
        for (int c = 0; c < channels; ++c) {
            const int src = channels == 3 ? rgbIndex[c] : 0;
            for (int x = 0; x < w; ++x) row[x] = px[((size_t)y * w + x) * channels + src];
            putBytes(out, row.data(), row.size() * sizeof(float));
        }
    }
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    f.write(reinterpret_cast<const char*>(out.data()), (std::streamsize)out.size());
    if (!f) throw std::runtime_error("write failed: " + path);
}

void Film::saveEXR(const std::string& path) const { writeEXR(path, rgb_.data(), w_, h_, 3); }

void Film::saveAov(const std::string& kind, const std::string& path) const {
    if (aov_.empty()) throw std::runtime_error("the film has no AOVs");
    int first, n;
    if (kind == "albedo") { first = 0; n = 3; }
    else if (kind == "normal") { first = 3; n = 3; }
    else if (kind == "depth") { first = 6; n = 1; }
    else if (kind == "roughness") { first = 7; n = 1; }
    else if (kind == "metallic") { first = 8; n = 1; }
    else throw std::runtime_error("unknown AOV: " + kind);
    const size_t np = (size_t)w_ * h_;
    std::vector<float> img(np * n);
    for (size_t i = 0; i < np; ++i)
        for (int c = 0; c < n; ++c) img[i * n + c] = aov_[i * kAovs + first + c];
    const auto dot = path.find_last_of('.');
    const std::string ext = dot == std::string::npos ? "" : path.substr(dot + 1);
    if (ext == "exr") writeEXR(path, img.data(), w_, h_, n);
    else if (ext == "pfm") writePFM(path, img.data(), w_, h_, n);
    else throw std::runtime_error("unsupported image format: " + path);
}

namespace {
uint32_t crc32(const unsigned char* p, size_t n, uint32_t c = 0xFFFFFFFFu) {
    for (size_t i = 0; i < n; ++i) {
        c ^= p[i];
        for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (0xEDB88320u & (0u - (c & 1u)));
    }
    return c;
}
void put32(std::vector<unsigned char>& v, uint32_t x) {
    for (int s = 24; s >= 0; s -= 8) v.push_back((unsigned char)(x >> s));
}
void chunk(std::ofstream& f, const char* type, const std::vector<unsigned char>& data) {
    std::vector<unsigned char> b;
    put32(b, (uint32_t)data.size());
    b.insert(b.end(), type, type + 4);
    b.insert(b.end(), data.begin(), data.end());
    std::vector<unsigned char> tail;
    put32(tail, crc32(b.data() + 4, b.size() - 4) ^ 0xFFFFFFFFu);
    f.write(reinterpret_cast<const char*>(b.data()), (std::streamsize)b.size());
    f.write(reinterpret_cast<const char*>(tail.data()), 4);
}
unsigned char toSRGB8(float v) {
    v = std::min(1.0f, std::max(0.0f, v));
    const float s = v <= 0.0031308f ? 12.92f * v : 1.055f * std::pow(v, 1.0f / 2.4f) - 0.055f;
    return (unsigned char)std::lround(s * 255.0f);
}
}  // namespace

// PNG, 8-bit RGB, zlib stream of stored (uncompressed) deflate blocks
void Film::savePNG(const std::string& path) const {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    f.write(reinterpret_cast<const char*>(sig), 8);
    std::vector<unsigned char> ihdr;
    put32(ihdr, (uint32_t)w_);
    put32(ihdr, (uint32_t)h_);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8 bit, RGB, deflate, no filter set, no interlace
    chunk(f, "IHDR", ihdr);
    std::vector<unsigned char> raw;
    raw.reserve((size_t)h_ * (1 + 3 * (size_t)w_));
    for (int y = 0; y < h_; ++y) {
        raw.push_back(0);  // filter type none
        for (int x = 0; x < w_; ++x)
            for (int c = 0; c < 3; ++c) raw.push_back(toSRGB8(rgb_[3 * ((size_t)y * w_ + x) + c]));
    }
    std::vector<unsigned char> z = {0x78, 0x01};
    uint32_t a = 1, b = 0;  // Adler-32
    for (unsigned char ch : raw) {
        a = (a + ch) % 65521u;
        b = (b + a) % 65521u;
    }
    for (size_t off = 0; off < raw.size() || off == 0; off += 65535) {
        const size_t len = std::min<size_t>(65535, raw.size() - off);
        z.push_back(off + len >= raw.size() ? 1 : 0);
        z.push_back((unsigned char)(len & 0xFF));
        z.push_back((unsigned char)(len >> 8));
        z.push_back((unsigned char)(~len & 0xFF));
        z.push_back((unsigned char)((~len >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + (std::ptrdiff_t)off, raw.begin() + (std::ptrdiff_t)(off + len));
        if (raw.empty()) break;
    }
    put32(z, (b << 16) | a);
    chunk(f, "IDAT", z);
    chunk(f, "IEND", {});
}

// ---------------------------------------------------------------- GPU integrators
namespace {
void check(int rc, gnd_hip_ctx* ctx) {
    if (rc != GND_OK) throw std::runtime_error(gnd_hip_last_error(ctx));
}
}  // namespace

void HipIntegrator::render(const RenderScene& scene) {
    stop_ = false;
    passes_ = 0;
    if (!ctx_ || !scene.desc || !scene.device) throw std::runtime_error("integrator needs a context and a scene");
    preprocess(scene);
    const gnd_scene_desc& d = *scene.desc;
    const int W = d.camera.width, H = d.camera.height;
    const uint32_t spp = d.spp;
    film_.reset(W, H, aovs_);
    float* dev = nullptr;
    float* devAov = nullptr;
    check(gnd_hip_malloc(ctx_, (size_t)W * H * 4 * sizeof(float), (void**)&dev), ctx_);
    std::vector<float> acc((size_t)W * H * 4), aovSums;
    try {
        if (aovs_) {
            check(gnd_hip_malloc(ctx_, (size_t)W * H * Film::kAovs * sizeof(float), (void**)&devAov), ctx_);
            aovSums.resize((size_t)W * H * Film::kAovs);
        }
        for (uint32_t s = 0; s < spp && !stop_; s += (uint32_t)batch_) {
            gnd_hip_job job{};
            job.camera = d.camera;
            job.spp = std::min<uint32_t>((uint32_t)batch_, spp - s);
            job.max_depth = maxDepth_;
            job.integrator = integrator_;
            job.ao_length = d.ao_length;
            job.seed_offset = (uint64_t)s * (uint64_t)W * (uint64_t)H;
            job.film = dev;
            job.film_aov = devAov;
            // the passes of every batch after the first add onto the device
            // film in pass order, so the film holds the bits of one render of
            // all passes so far whatever the batch size
            job.accumulate = s > 0 ? 1 : 0;
            const int rc = gnd_hip_render(ctx_, scene.device, &job, 1, nullptr);
            if (rc == GND_E_CANCELLED) break;
            check(rc, ctx_);
            check(gnd_hip_memcpy_d2h(ctx_, acc.data(), dev, acc.size() * sizeof(float)), ctx_);
            film_.setSums(acc.data());
            if (devAov) {
                check(gnd_hip_memcpy_d2h(ctx_, aovSums.data(), devAov, aovSums.size() * sizeof(float)), ctx_);
                film_.setAovSums(aovSums.data());
            }
            film_.resolve();
            passes_ += (int)job.spp;
            notifyUpdate((int)(s + job.spp - 1));  // integrator.h:147
        }
    } catch (...) {
        gnd_hip_free(ctx_, dev);
        if (devAov) gnd_hip_free(ctx_, devAov);
        throw;
    }
    gnd_hip_free(ctx_, dev);
    if (devAov) gnd_hip_free(ctx_, devAov);
}

void Denoiser::execute(gnd_hip_ctx* ctx, Film* film) {
    if (!film || film->width() <= 0 || film->rgb().empty()) return;
    const int W = film->width(), H = film->height();
    const size_t n = (size_t)W * H * 3;
    std::vector<float> albedo, normal;
    if (film->hasAovs()) {
        albedo.resize(n);
        normal.resize(n);
        const std::vector<float>& a = film->aovs();
        for (size_t i = 0; i < (size_t)W * H; ++i)
            for (int c = 0; c < 3; ++c) {
                albedo[3 * i + c] = a[i * Film::kAovs + c];
                normal[3 * i + c] = a[i * Film::kAovs + 3 + c];
            }
    }
    float* d[3] = {nullptr, nullptr, nullptr};  // color (also the output), albedo, normal
    const int nbuf = albedo.empty() ? 1 : 3;
    try {
        for (int k = 0; k < nbuf; ++k) check(gnd_hip_malloc(ctx, n * sizeof(float), (void**)&d[k]), ctx);
        check(gnd_hip_memcpy_h2d(ctx, d[0], film->rgb().data(), n * sizeof(float)), ctx);
        if (nbuf == 3) {
            check(gnd_hip_memcpy_h2d(ctx, d[1], albedo.data(), n * sizeof(float)), ctx);
            check(gnd_hip_memcpy_h2d(ctx, d[2], normal.data(), n * sizeof(float)), ctx);
        }
        check(gnd_hip_denoise(ctx, d[0], d[1], d[2], W, H, d[0], nullptr), ctx);
        std::vector<float> out(n);
        check(gnd_hip_memcpy_d2h(ctx, out.data(), d[0], n * sizeof(float)), ctx);
        film->replaceRgb(out.data());
    } catch (...) {
        for (float* p : d)
            if (p) gnd_hip_free(ctx, p);
        throw;
    }
    for (float* p : d)
        if (p) gnd_hip_free(ctx, p);
}

void HipIntegrator::cancel() {
    stop_ = true;
    if (ctx_) gnd_hip_cancel(ctx_);
}

// ---------------------------------------------------------------- factory
namespace {
std::mutex g_factoryMu;
std::map<std::string, IntegratorFactory::Creator>& registry() {
    static std::map<std::string, IntegratorFactory::Creator> r;
    return r;
}
}  // namespace

void IntegratorFactory::registerClass(const std::string& name, Creator c) {
    std::lock_guard<std::mutex> g(g_factoryMu);
    auto& r = registry();
    if (r.count(name)) std::fprintf(stderr, "Warning: integrator \"%s\" registered twice, overriding\n", name.c_str());
    r[name] = std::move(c);
}

std::unique_ptr<Integrator> IntegratorFactory::create(const std::string& name, const IntegratorParams& p) {
    registerBuiltinIntegrators();
    Creator c;
    {
        std::lock_guard<std::mutex> g(g_factoryMu);
        auto it = registry().find(name);
        if (it == registry().end()) throw std::runtime_error("unknown integrator: " + name);
        c = it->second;
    }
    return c(p);
}

std::vector<std::string> IntegratorFactory::names() {
    registerBuiltinIntegrators();
    std::lock_guard<std::mutex> g(g_factoryMu);
    std::vector<std::string> out;
    for (auto& kv : registry()) out.push_back(kv.first);
    return out;
}

void registerBuiltinIntegrators() {
    static std::once_flag once;
    std::call_once(once, [] {
        auto make = [](int type) {
            return [type](const IntegratorParams& p) -> std::unique_ptr<Integrator> {
                if (!p.ctx) throw std::runtime_error("GPU integrators need a device context");
                return std::make_unique<HipIntegrator>(p.ctx, type, p.maxDepth, std::max(1, p.passesPerCallback));
            };
        };
        // the CPU names select the GPU integrators too: a scene file keeps
        // <integrator type="path"> and renders on the device
        IntegratorFactory::registerClass("path", make(GND_INTEGRATOR_PATH));
        IntegratorFactory::registerClass("volpath", make(GND_INTEGRATOR_VOLPATH));
        IntegratorFactory::registerClass("path_hip", make(GND_INTEGRATOR_PATH));
        IntegratorFactory::registerClass("volpath_hip", make(GND_INTEGRATOR_VOLPATH));
        IntegratorFactory::registerClass("direct", make(GND_INTEGRATOR_DIRECT));
        IntegratorFactory::registerClass("direct_hip", make(GND_INTEGRATOR_DIRECT));
        IntegratorFactory::registerClass("ao", make(GND_INTEGRATOR_AO));
        IntegratorFactory::registerClass("ao_hip", make(GND_INTEGRATOR_AO));
    });
}

}  // namespace gnd
