// host_capi.cpp — extern "C" surface of the host scene shell (include/gnd_host.h).
#include <cmath>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/gnd_host.h"
#include "scene_host.h"

struct gnd_host_scene {
    std::unique_ptr<gnd::HostScene> scene;
    // device copies per context (uploaded on first render)
    std::vector<std::pair<gnd_hip_ctx*, gnd_hip_scene*>> uploads;
    std::mutex mu;
};

namespace {
thread_local std::string g_hostErr;

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
    for (auto& u : s->uploads) gnd_hip_scene_free(u.first, u.second);
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
    {
        std::lock_guard<std::mutex> g(s->mu);
        for (auto& u : s->uploads)
            if (u.first == ctx) dev = u.second;
        if (!dev) {
            const gnd_hip_scene_arrays* arr = nullptr;
            int rc = hostGuard([&] {
                arr = &s->scene->build(0, nullptr);
                return GND_OK;
            });
            if (rc != GND_OK) return rc;
            rc = gnd_hip_scene_upload(ctx, arr, &dev);
            if (rc != GND_OK) {
                g_hostErr = gnd_hip_last_error(ctx);
                return rc;
            }
            s->uploads.push_back({ctx, dev});
        }
    }
    const gnd_scene_desc& d = s->scene->desc();
    const size_t np = (size_t)d.camera.width * d.camera.height;
    void* film = nullptr;
    int rc = gnd_hip_malloc(ctx, np * 4 * sizeof(float), &film);
    if (rc != GND_OK) return rc;
    gnd_hip_job job{};
    job.camera = d.camera;
    job.spp = spp > 0 ? (uint32_t)spp : d.spp;
    job.max_depth = d.max_depth;
    job.integrator = d.integrator;
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

}  // extern "C"
