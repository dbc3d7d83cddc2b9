// The C++ facade used the way a caller of the reference uses its API
// (src/main.cpp run*): configuration 1 built through the tracer:: factories,
// create_pt_renderer(params)->render(scene, Film(res, create_gaussin_filter(0.5, 0.6))).
// Writes each render's float RGB image to <prefix>_<k>.f32:
//   a: the scene as built;
//   b: the same renderer after scene->set_camera(second camera);
//   c: a fresh renderer on the scene with the second camera;
//   d: the first renderer on another scene object (configuration 1, seed 5);
//   e: create_bdpt_renderer(BDPTRendererParams{.., 5, 5, spp 16})->render on the
//      first scene with its second camera.
#include <cstdio>
#include <stdexcept>
#include <string>

#include "tracer/scenes.h"

using namespace tracer;

static void save(const RenderTarget& t, const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    const size_t n = size_t(t.color.width()) * size_t(t.color.height()) * 3;
    const size_t w = std::fwrite(t.color.get_raw_data(), sizeof(float), n, f);
    std::fclose(f);
    if (w != n) throw std::runtime_error("short write " + path);
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <output prefix>\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    try {
        auto b = scenes::build_config(1, 0, 1.0);
        PTRendererParams params;
        params.spp = 16;
        params.min_depth = b.min_depth;
        params.max_depth = b.max_depth;
        params.direct_light_sample_num = 1;
        GpuRenderOptions opt;
        opt.seed = 7;
        auto r = create_pt_renderer(params, opt);
        Film film(Point2i(64, 48), create_gaussin_filter(0.5f, 0.6f));
        save(r->render(*b.scene, film), out + "_a.f32");
        b.scene->set_camera(create_thin_lens_camera(64.0f / 48.0f, Point3f(0.3f, 0.2f, 3.0f), Point3f(0.0f, -0.2f, 0.0f),
                                                     Vector3f(0.0f, 1.0f, 0.0f), 0.8f, 0.0f, 1.0f));
        save(r->render(*b.scene, film), out + "_b.f32");
        auto r2 = create_pt_renderer(params, opt);
        save(r2->render(*b.scene, film), out + "_c.f32");
        auto b2 = scenes::build_config(1, 5, 1.0);
        save(r->render(*b2.scene, film), out + "_d.f32");
        BDPTRendererParams bp;
        bp.max_camera_vertex_count = 5;
        bp.max_light_vertex_count = 5;
        bp.spp = 16;
        auto rb = create_bdpt_renderer(bp, opt);
        save(rb->render(*b.scene, film), out + "_e.f32");
    } catch (const std::exception& e) {
        std::fprintf(stderr, "facade_harness: %s\n", e.what());
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
