// Sanitizer driver for the host shell and the oracle (test infrastructure,
// built with -fsanitize=address,undefined by tests/test_sanitizers.py).
//
// For an XML scene file and each procedural scene: load or generate it, build
// its BVHs on four threads, flatten it, build the oracle's scene from the
// flattened description and render two rows of one pass with the oracle on
// two threads.  Also reads an EXR file.  Exits non-zero on any error; the
// sanitizers abort on the first memory or undefined-behaviour error.
#include <cmath>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "../gianduia_amd/csrc/host/scene_host.h"

namespace gnd {
void readEXR(const std::string& path, int& width, int& height, std::vector<float>& rgb);
}

extern "C" {
void* gnd_ref_scene_create(const gnd_scene_desc* d);
void gnd_ref_scene_free(void* h);
const char* gnd_ref_last_error();
int gnd_ref_render(void* h, int threads, int spp_begin, int spp_end, int row_begin, int row_end,
                   uint64_t seed_offset, float* rgbw, float* sq, double* stats, float* aov);
}

static int check(const char* name, gnd::HostScene& s) {
    double ms = 0.0;
    s.build(4, &ms);
    const gnd_scene_desc& d = s.desc();
    void* ref = gnd_ref_scene_create(&d);
    if (!ref) {
        std::fprintf(stderr, "%s: oracle scene failed: %s\n", name, gnd_ref_last_error());
        return 1;
    }
    const int W = d.camera.width, H = d.camera.height;
    std::vector<float> rgbw((size_t)W * H * 4, 0.0f);
    double stats[5] = {0, 0, 0, 0, 0};
    const int r0 = H / 2, r1 = H / 2 + 2 < H ? H / 2 + 2 : H;
    const int rc = gnd_ref_render(ref, 2, 0, 1, r0, r1, 0, rgbw.data(), nullptr, stats, nullptr);
    gnd_ref_scene_free(ref);
    if (rc != 0) {
        std::fprintf(stderr, "%s: oracle render failed: %s\n", name, gnd_ref_last_error());
        return 1;
    }
    double wsum = 0.0;
    for (int y = r0; y < r1; ++y)
        for (int x = 0; x < W; ++x) {
            const float* p = &rgbw[((size_t)y * W + x) * 4];
            for (int c = 0; c < 4; ++c)
                if (!std::isfinite(p[c])) {
                    std::fprintf(stderr, "%s: non-finite film value\n", name);
                    return 1;
                }
            wsum += p[3];
        }
    std::printf("%s: %dx%d, build %.1f ms, film weight %.1f over rows %d-%d\n", name, W, H, ms, wsum, r0, r1);
    return wsum > 0.0 ? 0 : 1;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s scene.xml image.exr\n", argv[0]);
        return 2;
    }
    int bad = 0;
    try {
        auto xml = gnd::loadXmlScene(argv[1]);
        bad |= check("xml", *xml);
        const char* gens[][2] = {{"cbox", "width=48;height=32"},
                                 {"scaled", "level=3;instances=4;width=48;height=27"},
                                 {"scaled", "level=3;instances=4;width=48;height=27;lights=point,spot,sun,env;envres=32"},
                                 {"volumetric", "grid=32;width=40;height=24"},
                                 {"cbox", "width=32;height=32;integrator=direct"},
                                 {"cbox", "width=32;height=32;integrator=ao"}};
        for (auto& g : gens) {
            auto s = gnd::generateScene(g[0], g[1]);
            bad |= check((std::string(g[0]) + " " + g[1]).c_str(), *s);
        }
        int w = 0, h = 0;
        std::vector<float> rgb;
        gnd::readEXR(argv[2], w, h, rgb);
        if (rgb.size() != (size_t)w * h * 3) {
            std::fprintf(stderr, "exr: %zu values for %dx%d\n", rgb.size(), w, h);
            bad = 1;
        }
        std::printf("exr: %dx%d\n", w, h);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return bad;
}
