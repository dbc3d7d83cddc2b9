/*
 * scenes.h — synthetic benchmark scenes (configurations 1-5) and OBJ scenes
 * built through the tracer:: factory API, and a C ABI to fetch them as
 * tr_scene_desc.
 *
 * There is no reference counterpart: the reference hard-codes its scenes in
 * src/main.cpp run*() (src/main.cpp:4-313) and loads external OBJ assets.
 * These generators play that role for the benchmark configurations.  All
 * randomness is PCG32 with a fixed draw order, so a (config, seed) pair always
 * gives the same scene.
 */
#ifndef TRACER_SCENES_H
#define TRACER_SCENES_H

#include <stdint.h>
#include "tracer_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct trs_scene trs_scene;

/* config: 1..5.  seed: PCG32 seed (the configurations use seed = config - 1).
 * detail: 1.0 = full size; < 1 shrinks the heavy parts (sphere count,
 * terrain resolution) for quick tests.  Returns NULL on failure and sets
 * *err (static storage) when err is not NULL. */
trs_scene* trs_create(int32_t config, uint64_t seed, double detail, const char** err);
/* A scene read from a Wavefront OBJ file (and its MTL libraries) through
 * load_model_from_file and create_texture_from_file, wired as the
 * reference's run() does (src/main.cpp:212-282): Phong materials, emissive
 * materials as area lights, MediumInterface{vacuum inside, fog outside} (the
 * fog of src/main.cpp:242 when outside_fog != 0, else vacuum), Gaussian
 * filter (0.5, 0.6), 3 primitives per BVH leaf; a material with Ni > 1 is a
 * glass.  Returns NULL and sets *err on a read or parse error, a triangle
 * without a material or an image texture. */
trs_scene* trs_load_obj(const char* obj_path, const tr_camera* camera, int32_t outside_fog, const char** err);
/* trs_load_obj with a heterogeneous medium outside every surface:
 * create_heterogeneous_medium(local_to_world (row-major 3x4),
 * create_image_texture3d(nx x ny x nz floats, x fastest), albedo, g,
 * max_scattering_count, false), src/medium/heterogeneous.cpp:101-108: the
 * albedo is the constant RGB albedo[3] or, when albedo_grid is not NULL, an
 * image of nx x ny x nz grey values; g the constant or, when g_grid is not
 * NULL, an image of nx x ny x nz values. */
trs_scene* trs_load_obj_grid(const char* obj_path, const tr_camera* camera, const float* density, int32_t nx, int32_t ny,
                             int32_t nz, const float* local_to_world, const float* albedo, float g,
                             const float* albedo_grid, const float* g_grid, int32_t max_scattering_count,
                             const char** err);
/* Reads a Radiance .hdr image (the loader behind create_texture2d_from_file):
 * *width, *height, and when rgb is not NULL its height rows of width linear
 * RGB texels (cap floats at least).  Returns 0, or 1 and *err on failure. */
int32_t trs_read_hdr(const char* path, int32_t* width, int32_t* height, float* rgb, int64_t cap, const char** err);
/* Reads an 8-bit PNG (the loader behind create_texture2d_from_file): rows
 * bottom to top, as the reference flips 8-bit images on load; grey and
 * palette images expand to RGB, alpha is dropped. */
int32_t trs_read_png(const char* path, int32_t* width, int32_t* height, uint8_t* rgb, int64_t cap, const char** err);
/* write_image_to_hdr / write_image_to_png of height rows of width RGB texels
 * (floats / bytes).  Return 0, or 1 and *err on failure. */
int32_t trs_write_hdr(const char* path, int32_t width, int32_t height, const float* rgb, const char** err);
int32_t trs_write_png(const char* path, int32_t width, int32_t height, const uint8_t* rgb, const char** err);
/* In place: create_aces_tone_mapper(aces_exposure) when aces_exposure >= 0,
 * then create_gamma_corrector(gamma) when gamma > 0.  Returns 0, or 1 on bad
 * arguments. */
int32_t trs_postprocess(float* rgb, int32_t width, int32_t height, float aces_exposure, float gamma);
const tr_scene_desc* trs_desc(const trs_scene* s);
/* Film, spp and depths of the configuration (PTRendererParams + film). */
void trs_render_defaults(const trs_scene* s, tr_render_params* out);
int32_t trs_max_leaf(const trs_scene* s);
void trs_destroy(trs_scene* s);

#ifdef __cplusplus
}

#include "tracer/tracer.hpp"

namespace tracer {
namespace scenes {
/* Configuration 1..5 as the factory API builds it: the scene objects (with
 * camera, lights and environment set) that trs_create flattens, the film
 * filter and the render defaults. */
struct BuiltScene {
    RC<Scene> scene;
    RC<Filter> filter;
    int width = 0, height = 0, spp = 0, min_depth = 0, max_depth = 0;
};
BuiltScene build_config(int config, uint64_t seed, double detail);
}  // namespace scenes
}  // namespace tracer
#endif
#endif /* TRACER_SCENES_H */
