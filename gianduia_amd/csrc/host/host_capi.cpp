// host_capi.cpp — extern "C" surface of the host scene shell (include/gnd_host.h).
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/gnd_host.h"
#include "integrator.h"
#include "scene_host.h"

struct gnd_host_scene {
    std::unique_ptr<gnd::HostScene> scene;
    // device copies per context (uploaded on first render)
    std::vector<std::pair<gnd_hip_ctx*, gnd_hip_scene*>> uploads;
    std::mutex mu;
};

struct gnd_host_builder {
    std::unique_ptr<gnd::ObjectGraph> graph;
};

struct gnd_host_integrator {
    std::unique_ptr<gnd::Integrator> impl;
    gnd_hip_ctx* ctx = nullptr;
    std::string output;  // path stem saved as .exr and .png after each render, or empty
};

namespace {
thread_local std::string g_hostErr;

// Host scenes holding device copies: a context being shut down frees the
// copies made on it first (gndReleaseContextUploads), so a host scene freed
// later never touches a dead context.  Lock order: g_uploadMu, then a
// scene's mu.
std::mutex g_uploadMu;
std::set<gnd_host_scene*> g_withUploads;

// The scene's device copy on `ctx` (built and uploaded on first use).
gnd_hip_scene* deviceScene(gnd_hip_ctx* ctx, gnd_host_scene* s) {
    std::lock_guard<std::mutex> r(g_uploadMu);
    std::lock_guard<std::mutex> g(s->mu);
    for (auto& u : s->uploads)
        if (u.first == ctx) return u.second;
    const gnd_hip_scene_arrays& arr = s->scene->build(0, nullptr);
    gnd_hip_scene* dev = nullptr;
    if (gnd_hip_scene_upload(ctx, &arr, &dev) != GND_OK) throw std::runtime_error(gnd_hip_last_error(ctx));
    s->uploads.push_back({ctx, dev});
    g_withUploads.insert(s);
    return dev;
}

template <class F>
int hostGuard(F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        g_hostErr = e.what();
        return GND_E_INVALID;
    }
}
}  // namespace

extern "C" {

const char* gnd_host_last_error(void) { return g_hostErr.c_str(); }

int gnd_host_scene_create(const gnd_scene_desc* desc, gnd_host_scene** out) {
    if (!desc || !out) return GND_E_INVALID;
    return hostGuard([&] {
        auto* h = new gnd_host_scene();
        h->scene = std::make_unique<gnd::HostScene>();
        h->scene->copyFrom(*desc);
        *out = h;
        return GND_OK;
    });
}

int gnd_host_scene_load_xml(const char* path, gnd_host_scene** out) {
    if (!path || !out) return GND_E_INVALID;
    return hostGuard([&] {
        auto* h = new gnd_host_scene();
        h->scene = gnd::loadXmlScene(path);
        *out = h;
        return GND_OK;
    });
}

int gnd_host_builder_create(const char* base_dir, gnd_host_builder** out) {
    if (!out) return GND_E_INVALID;
    return hostGuard([&] {
        auto* b = new gnd_host_builder();
        b->graph = std::make_unique<gnd::ObjectGraph>(base_dir ? base_dir : "");
        *out = b;
        return GND_OK;
    });
}

int gnd_host_builder_object(gnd_host_builder* b, const char* tag, const char* type, const char* name,
                            const gnd_prop* props, int nprops, int32_t* id) {
    if (!b || !tag || !id) return GND_E_INVALID;
    return hostGuard([&] {
        *id = b->graph->add(tag, type ? type : tag, name ? name : "", props, nprops);
        return GND_OK;
    });
}

int gnd_host_builder_add_child(gnd_host_builder* b, int32_t parent, int32_t child) {
    if (!b) return GND_E_INVALID;
    return hostGuard([&] {
        b->graph->addChild(parent, child);
        return GND_OK;
    });
}

int gnd_host_builder_finish(gnd_host_builder* b, int32_t root, gnd_host_scene** out) {
    if (!b || !out) return GND_E_INVALID;
    return hostGuard([&] {
        auto h = std::make_unique<gnd_host_scene>();
        h->scene = b->graph->build(root);
        *out = h.release();
        return GND_OK;
    });
}

void gnd_host_builder_free(gnd_host_builder* b) { delete b; }

int gnd_host_scene_generate(const char* name, const char* params, gnd_host_scene** out) {
    if (!name || !out) return GND_E_INVALID;
    return hostGuard([&] {
        auto* h = new gnd_host_scene();
        h->scene = gnd::generateScene(name, params ? params : "");
        *out = h;
        return GND_OK;
    });
}

void gnd_host_scene_free(gnd_host_scene* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> r(g_uploadMu);
        g_withUploads.erase(s);
        std::lock_guard<std::mutex> g(s->mu);
        for (auto& u : s->uploads) gnd_hip_scene_free(u.first, u.second);
        s->uploads.clear();
    }
    delete s;
}

const gnd_scene_desc* gnd_host_scene_desc(gnd_host_scene* s) { return s ? &s->scene->desc() : nullptr; }

int gnd_host_scene_build(gnd_host_scene* s, int threads, const gnd_hip_scene_arrays** out, double* build_ms) {
    if (!s || !out) return GND_E_INVALID;
    return hostGuard([&] {
        std::lock_guard<std::mutex> g(s->mu);
        *out = &s->scene->build(threads, build_ms);
        return GND_OK;
    });
}

int gnd_host_camera_lookat(gnd_camera_desc* cam, const float* o, const float* t, const float* up) {
    if (!cam || !o || !t || !up) return GND_E_INVALID;
    gnd::Xform x = gnd::xformLookAt(gnd::Vec3(o[0], o[1], o[2]), gnd::Vec3(t[0], t[1], t[2]),
                                    gnd::Vec3(up[0], up[1], up[2]));
    std::memcpy(cam->to_world, x.m.m, sizeof(cam->to_world));
    std::memcpy(cam->to_world_inv, x.inv.m, sizeof(cam->to_world_inv));
    return GND_OK;
}

// Integrator::render with the film resolved on the host (Film::resolve,
// src/core/film.cpp:50-94): rgb = sum * (1 / weight), 0 where weight is 0.
int gnd_host_render(gnd_hip_ctx* ctx, gnd_host_scene* s, int spp, uint64_t seed_offset, float* rgb,
                    gnd_hip_stats* stats) {
    if (!ctx || !s || !rgb) return GND_E_INVALID;
    gnd_hip_scene* dev = nullptr;
    int rc = hostGuard([&] {
        dev = deviceScene(ctx, s);
        return GND_OK;
    });
    if (rc != GND_OK) return rc;
    const gnd_scene_desc& d = s->scene->desc();
    const size_t np = (size_t)d.camera.width * d.camera.height;
    void* film = nullptr;
    rc = gnd_hip_malloc(ctx, np * 4 * sizeof(float), &film);
    if (rc != GND_OK) return rc;
    gnd_hip_job job{};
    job.camera = d.camera;
    job.spp = spp > 0 ? (uint32_t)spp : d.spp;
    job.max_depth = d.max_depth;
    job.integrator = d.integrator;
    job.ao_length = d.ao_length;
    job.seed_offset = seed_offset;
    job.film = static_cast<float*>(film);
    rc = gnd_hip_render(ctx, dev, &job, 1, stats);
    std::vector<float> acc(np * 4);
    if (rc == GND_OK) rc = gnd_hip_memcpy_d2h(ctx, acc.data(), film, acc.size() * sizeof(float));
    gnd_hip_free(ctx, film);
    if (rc != GND_OK) {
        g_hostErr = gnd_hip_last_error(ctx);
        return rc;
    }
    for (size_t i = 0; i < np; ++i) {
        float w = acc[4 * i + 3];
        if (w > 0.0f) {
            float inv = 1.0f / w;
            rgb[3 * i + 0] = acc[4 * i + 0] * inv;
            rgb[3 * i + 1] = acc[4 * i + 1] * inv;
            rgb[3 * i + 2] = acc[4 * i + 2] * inv;
        } else {
            rgb[3 * i + 0] = rgb[3 * i + 1] = rgb[3 * i + 2] = 0.0f;
        }
    }
    return GND_OK;
}

// ---- Integrator plugin interface (include/gianduia/core/integrator.h:18-40)

int gnd_host_integrator_create(gnd_hip_ctx* ctx, const char* name, int max_depth, int passes_per_callback,
                               gnd_host_integrator** out) {
    if (!ctx || !name || !out) return GND_E_INVALID;
    return hostGuard([&] {
        gnd::IntegratorParams p;
        p.ctx = ctx;
        if (max_depth >= 0) p.maxDepth = max_depth;
        p.passesPerCallback = passes_per_callback > 0 ? passes_per_callback : 1;
        auto* h = new gnd_host_integrator();
        h->impl = gnd::IntegratorFactory::create(name, p);
        h->ctx = ctx;
        *out = h;
        return GND_OK;
    });
}

int gnd_host_integrator_from_scene(gnd_hip_ctx* ctx, gnd_host_scene* s, int passes_per_callback,
                                   gnd_host_integrator** out) {
    if (!ctx || !s || !out) return GND_E_INVALID;
    const gnd_scene_desc& d = s->scene->desc();
    static const char* names[] = {"path", "volpath", "direct", "ao"};
    if (d.integrator < GND_INTEGRATOR_PATH || d.integrator > GND_INTEGRATOR_AO) return GND_E_INVALID;
    return gnd_host_integrator_create(ctx, names[d.integrator], d.max_depth,
                                      passes_per_callback, out);
}

int gnd_host_integrator_set_callback(gnd_host_integrator* it, gnd_host_progress_fn fn, void* user) {
    if (!it) return GND_E_INVALID;
    if (!fn) {
        it->impl->setCallback(nullptr);
        return GND_OK;
    }
    it->impl->setCallback([fn, user](int sampleIndex, const gnd::Film& film) {
        fn(sampleIndex, film.rgb().data(), film.width(), film.height(), user);
    });
    return GND_OK;
}

int gnd_host_integrator_render(gnd_host_integrator* it, gnd_host_scene* s) {
    if (!it || !s) return GND_E_INVALID;
    return hostGuard([&] {
        gnd::RenderScene rs;
        rs.device = deviceScene(it->ctx, s);
        rs.desc = &s->scene->desc();
        it->impl->render(rs);
        // SamplerIntegrator::render ends with film->saveEXR(); film->savePNG()
        // (integrator.h:152-153), also after a cancel
        if (!it->output.empty()) {
            it->impl->film().saveEXR(it->output + ".exr");
            it->impl->film().savePNG(it->output + ".png");
        }
        return GND_OK;
    });
}

int gnd_host_integrator_set_output(gnd_host_integrator* it, const char* path_stem) {
    if (!it) return GND_E_INVALID;
    it->output = path_stem ? path_stem : "";
    return GND_OK;
}

int gnd_host_integrator_cancel(gnd_host_integrator* it) {
    if (!it) return GND_E_INVALID;
    it->impl->cancel();
    return GND_OK;
}

int gnd_host_integrator_film(gnd_host_integrator* it, float* rgb, int* width, int* height, int* passes) {
    if (!it) return GND_E_INVALID;
    const gnd::Film& f = it->impl->film();
    if (width) *width = f.width();
    if (height) *height = f.height();
    if (passes) *passes = it->impl->passesDone();
    if (rgb) std::memcpy(rgb, f.rgb().data(), f.rgb().size() * sizeof(float));
    return GND_OK;
}

int gnd_host_integrator_set_aovs(gnd_host_integrator* it, int enabled) {
    if (!it) return GND_E_INVALID;
    it->impl->setAovs(enabled != 0);
    return GND_OK;
}

int gnd_host_integrator_aovs(gnd_host_integrator* it, float* out, int* width, int* height) {
    if (!it) return GND_E_INVALID;
    const gnd::Film& f = it->impl->film();
    if (width) *width = f.width();
    if (height) *height = f.height();
    if (!f.hasAovs()) {
        g_hostErr = "the film has no AOVs (enable them before render)";
        return GND_E_INVALID;
    }
    if (out) std::memcpy(out, f.aovs().data(), f.aovs().size() * sizeof(float));
    return GND_OK;
}

int gnd_host_integrator_save(gnd_host_integrator* it, const char* path) {
    if (!it || !path) return GND_E_INVALID;
    return hostGuard([&] {
        const std::string p(path);
        const auto dot = p.find_last_of('.');
        const std::string ext = dot == std::string::npos ? "" : p.substr(dot + 1);
        if (ext == "pfm") it->impl->film().savePFM(p);
        else if (ext == "png") it->impl->film().savePNG(p);
        else if (ext == "exr") it->impl->film().saveEXR(p);
        else throw std::runtime_error("unsupported image format: " + p);
        return GND_OK;
    });
}

int gnd_host_integrator_save_aov(gnd_host_integrator* it, const char* aov, const char* path) {
    if (!it || !aov || !path) return GND_E_INVALID;
    return hostGuard([&] {
        it->impl->film().saveAov(aov, path);
        return GND_OK;
    });
}

int gnd_host_integrator_denoise(gnd_host_integrator* it) {
    if (!it) return GND_E_INVALID;
    return hostGuard([&] {
        gnd::Denoiser().execute(it->ctx, &it->impl->film());
        return GND_OK;
    });
}

void gnd_host_integrator_free(gnd_host_integrator* it) { delete it; }

int gnd_host_integrator_names(char* buf, int len) {
    return hostGuard([&] {
        std::string all;
        for (const std::string& n : gnd::IntegratorFactory::names()) all += (all.empty() ? "" : ",") + n;
        if (buf && len > 0) {
            std::strncpy(buf, all.c_str(), (size_t)len - 1);
            buf[len - 1] = '\0';
        }
        return (int)all.size();
    });
}

int gnd_host_image_save(const char* path, const float* rgb, int width, int height) {
    if (!path || !rgb || width <= 0 || height <= 0) return GND_E_INVALID;
    return hostGuard([&] {
        gnd::Film f(width, height);
        std::vector<float> rgbw((size_t)width * height * 4);
        for (size_t i = 0; i < (size_t)width * height; ++i) {
            for (int c = 0; c < 3; ++c) rgbw[4 * i + c] = rgb[3 * i + c];
            rgbw[4 * i + 3] = 1.0f;
        }
        f.accumulate(rgbw.data());
        f.resolve();
        const std::string p(path);
        if (p.size() > 4 && p.compare(p.size() - 4, 4, ".pfm") == 0) f.savePFM(p);
        else if (p.size() > 4 && p.compare(p.size() - 4, 4, ".png") == 0) f.savePNG(p);
        else if (p.size() > 4 && p.compare(p.size() - 4, 4, ".exr") == 0) f.saveEXR(p);
        else throw std::runtime_error("unsupported image format: " + p);
        return GND_OK;
    });
}

int gnd_host_image_load(const char* path, int* width, int* height, float* rgb, size_t capacity) {
    if (!path || !width || !height) return GND_E_INVALID;
    return hostGuard([&] {
        int w = 0, h = 0;
        std::vector<float> px;
        gnd::readEXR(path, w, h, px);
        *width = w;
        *height = h;
        if (rgb) {
            if (capacity < px.size()) throw std::runtime_error("image buffer too small");
            std::memcpy(rgb, px.data(), px.size() * sizeof(float));
        }
        return GND_OK;
    });
}

int gnd_host_sampler_floats(uint64_t sample_index, int n, float* out) {
    if (n < 0 || (n > 0 && !out)) return GND_E_INVALID;
    gnd::IndependentSampler smp(1);
    smp.seed(sample_index);
    for (int i = 0; i < n; ++i) out[i] = smp.next1D();
    return GND_OK;
}

}  // extern "C"

// Called by gnd_hip_shutdown while `ctx` is still valid: frees the device
// copies host scenes made on it and forgets them.
void gndReleaseContextUploads(gnd_hip_ctx* ctx) {
    std::lock_guard<std::mutex> r(g_uploadMu);
    for (auto it = g_withUploads.begin(); it != g_withUploads.end();) {
        gnd_host_scene* s = *it;
        std::lock_guard<std::mutex> g(s->mu);
        for (auto u = s->uploads.begin(); u != s->uploads.end();) {
            if (u->first == ctx) {
                gnd_hip_scene_free(ctx, u->second);
                u = s->uploads.erase(u);
            } else {
                ++u;
            }
        }
        it = s->uploads.empty() ? g_withUploads.erase(it) : std::next(it);
    }
}
