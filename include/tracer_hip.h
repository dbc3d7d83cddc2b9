/*
 * tracer_hip.h — C ABI of libtracer_hip.so, the MI355X (gfx950) wavefront path
 * tracer.  Plain C types, plain pointers and sizes; no exceptions cross it.
 *
 * The reference (wyzwzz/Tracer) has no FFI: its public surface is the C++
 * factory API in namespace tracer (the headers under src/factory/) and the abstract
 * Renderer/Scene classes (src/core/renderer.hpp:12-18, src/core/scene.hpp:11-33).
 * This ABI is what that API flattens to; the C++ facade in include/tracer/
 * keeps the reference's names and signatures and calls into it.  Each entry
 * point below names the reference interface it replaces.
 *
 * Conventions
 *   - Every call returns tr_status (0 = TR_OK).  tr_last_error(ctx) returns a
 *     message for the last failure on that context.
 *   - One HIP stream per context; calls on one context are serialised, and
 *     separate contexts are independent.  A context is bound to one device.
 *   - Host pointers are caller-owned and only read (or written) during the
 *     call.  Functions whose name ends in _device take device pointers and
 *     enqueue on the context's stream without synchronising.
 */
#ifndef TRACER_HIP_H
#define TRACER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t tr_status;
enum {
    TR_OK = 0,
    TR_ERR_INVALID = 1,   /* bad argument or inconsistent scene */
    TR_ERR_HIP = 2,       /* a HIP runtime call failed */
    TR_ERR_OOM = 3,       /* device allocation failed */
    TR_ERR_STATE = 4      /* call order violated (e.g. render before upload) */
};

typedef struct tr_ctx tr_ctx;

/* Material kinds.  TR_MAT_LAMBERT is the reference's Phong material with
 * ks = 0 (src/material/phong.cpp:29-70): base_color is kd, already divided by
 * max(1, max_i kd_i).  TR_MAT_DISNEY is create_disney
 * (src/factory/material.hpp:37-52, src/material/disney.cpp) with every
 * texture constant.  TR_MAT_PHONG is create_phong_material(ka, kd, ks, ns)
 * (src/factory/material.hpp:19-24) with constant textures: base_color = kd,
 * specular_scale = ks, shininess = ns as given; the library applies the
 * energy normalisation and lobe weights of Phong::shading
 * (src/material/phong.cpp:36-53).  TR_MAT_GLASS is create_glass(kr, kt, ior)
 * (src/factory/material.hpp:13-17, src/material/glass.cpp) with constant
 * textures: base_color = kr, specular_scale = kt, ior > 0; the refracted
 * colour is sqrt(kt) as Glass::shading computes it (pass kt = kr for the
 * reference's map_kr == map_kt case). */
enum { TR_MAT_LAMBERT = 0, TR_MAT_DISNEY = 1, TR_MAT_PHONG = 2, TR_MAT_GLASS = 3 };

typedef struct tr_material {
    int32_t type;
    float base_color[3];
    float metallic, roughness, transmission, transmission_roughness, ior;
    float specular_scale[3];
    float specular_tint, anisotropic, sheen, sheen_tint, clearcoat, clearcoat_gloss;
    float shininess;   /* Phong exponent ns (TR_MAT_PHONG) */
    /* 0: none; k > 0: textures[k - 1] of the scene gives base_color per hit
     * (kd of TR_MAT_PHONG, base_color of TR_MAT_DISNEY) at the hit's
     * interpolated uv (Texture2D::evaluate, src/core/texture.hpp:57-83). */
    int32_t base_color_texture;
    /* TR_MAT_DISNEY only: 0: the constant above; k > 0: textures[k - 1] gives
     * the parameter per hit at the hit's uv (Disney::shading evaluates every
     * texture, src/material/disney.cpp:604-621; a scalar parameter takes the
     * red channel, Texture2D::evaluate_s, src/core/texture.hpp:73-82), in the
     * order metallic, roughness, transmission, transmission_roughness, ior,
     * specular_scale (RGB), specular_tint, anisotropic, sheen, sheen_tint,
     * clearcoat, clearcoat_gloss (TR_PARAM_*). */
    int32_t param_texture[12];
    /* TR_MAT_DISNEY only: 0: none; k > 0: textures[k - 1] is the normal map of
     * NormalMapper::reorient (src/core/material.hpp:37-48): the texel c gives
     * the tangent-space normal 2c - 1, and the shading frame is turned so
     * that its z axis is that normal. */
    int32_t normal_texture;
} tr_material;

enum {
    TR_PARAM_METALLIC = 0, TR_PARAM_ROUGHNESS, TR_PARAM_TRANSMISSION, TR_PARAM_TRANSMISSION_ROUGHNESS, TR_PARAM_IOR,
    TR_PARAM_SPECULAR_SCALE, TR_PARAM_SPECULAR_TINT, TR_PARAM_ANISOTROPIC, TR_PARAM_SHEEN, TR_PARAM_SHEEN_TINT,
    TR_PARAM_CLEARCOAT, TR_PARAM_CLEARCOAT_GLOSS, TR_PARAM_COUNT
};

/* An image texture (create_texture2d_from_file, an MTL map): height rows of
 * width RGB texels, 8-bit (rgb8, read as value / 255) or float (rgbf); the
 * other pointer is NULL.  Sampled bilinearly with wrap_repeat. */
typedef struct tr_texture {
    int32_t width, height;
    const uint8_t* rgb8;
    const float* rgbf;
} tr_texture;

/* create_homogeneous_medium(sigma_a, sigma_s, g, max_scattering_count),
 * src/factory/medium.hpp:21-26, src/medium/homogeneous.cpp:12-18.
 *
 * With density != NULL the medium is create_heterogeneous_medium(
 * local_to_world, density, albedo, g, max_scattering_count,
 * white_for_indirect), src/factory/medium.hpp:13-19,
 * src/medium/heterogeneous.cpp:10-108: a density grid of nx * ny * nz floats
 * (x fastest, then y, then z; sigma_t = density) looked up trilinearly at the
 * local position world_to_local(p) with every coordinate wrapped into [0, 1]
 * (Texture3D's wrap_repeat).  sigma_s then holds the constant albedo, g the
 * constant asymmetry, and sigma_a is unused.  local_to_world is a row-major
 * 3x4 affine matrix; its inverse maps world positions to grid coordinates.
 * albedo_grid (anx * any * anz RGB triples) and g_grid (gnx * gny * gnz
 * values), when not NULL, replace the constant albedo and g: both are looked
 * up trilinearly at the collision's local position, as HeterogeneousMedium::
 * sample reads albedo->evaluate and g->evaluate_s there
 * (src/medium/heterogeneous.cpp:66-78); g is kept inside (-1, 1) as the
 * constant is. */
typedef struct tr_medium {
    float sigma_a[3];
    float sigma_s[3];
    float g;
    int32_t max_scattering_count;  /* scatters in a run before transmittance only; 0: none */
    const float* density;          /* NULL: homogeneous */
    int32_t nx, ny, nz;
    int32_t white_for_indirect;    /* albedo 1 after the first scatter of a run */
    float local_to_world[12];
    const float* albedo_grid;      /* NULL: sigma_s is the albedo */
    int32_t anx, any, anz;
    const float* g_grid;           /* NULL: g is the asymmetry */
    int32_t gnx, gny, gnz;
} tr_medium;

/* create_thin_lens_camera(aspect, pos, target, up, fov, lens_radius,
 * focal_dist), src/factory/camera.hpp:9-11; fov is the vertical field of view
 * in radians (src/camera/thin_lens_camera.cpp:31). */
typedef struct tr_camera {
    float pos[3];
    float target[3];
    float up[3];
    float fov_y;
    float lens_radius;
    float focal_distance;
    float aspect;
} tr_camera;

/* A flattened scene: what create_triangle_mesh + create_geometric_primitive +
 * create_bvh_accel::build + create_general_scene (+ lights, environment
 * light, camera) describe in the reference.  Triangle i is one
 * GeometricPrimitive (src/shape/geometry.cpp:10-75): its material, its
 * MediumInterface (-1 = vacuum) and, if emission is non-zero, a one-sided
 * DiffuseLight (src/light/diffuse_light.cpp) on the side of its geometric
 * normal cross(B-A, C-A). */
typedef struct tr_scene_desc {
    int64_t num_vertices;
    const float* positions;        /* 3 * num_vertices, world space */
    const float* normals;          /* 3 * num_vertices, or NULL (flat shading) */
    const float* uvs;              /* 2 * num_vertices, or NULL */
    int64_t num_triangles;
    const int32_t* indices;        /* 3 * num_triangles */
    const int32_t* material_ids;   /* num_triangles */
    const float* emission;         /* 3 * num_triangles, or NULL */
    const int32_t* medium_inside;  /* num_triangles, or NULL (all vacuum) */
    const int32_t* medium_outside; /* num_triangles, or NULL (all vacuum) */
    int32_t num_materials;
    int32_t num_media;
    const tr_material* materials;
    const tr_medium* media;
    /* Environment light.  has_env with env_width == 0: create_ibl_light(
     * create_constant_texture2d(env_radiance)); with env_width > 0: an image
     * environment (create_ibl_light over an HDR texture,
     * src/light/environment_light.cpp:20-196), see env_pixels below. */
    int32_t has_env;
    float env_radiance[3];
    /* Directional light (no reference counterpart): direction points from the
     * scene towards the sun; irradiance is measured normal to it. */
    int32_t has_sun;
    float sun_direction[3];
    float sun_irradiance[3];
    tr_camera camera;
    /* Image environment: env_height rows of env_width linear RGB texels
     * (3 floats each), row 0 at v = 0 (theta = 0), sampled bilinearly
     * (src/utility/sampler.hpp:12-25).  A world direction d maps to light
     * space as env_world_to_light * d (row-major 3x3 rotation), then to
     * u = phi / 2pi, v = theta / pi of the light-space direction. */
    int32_t env_width, env_height;
    const float* env_pixels;
    float env_world_to_light[9];
    /* textures that materials name by base_color_texture (uvs required) */
    int32_t num_textures;
    const tr_texture* textures;
} tr_scene_desc;

/* BVH options.  max_leaf_primitives mirrors create_bvh_accel(int)
 * (src/factory/accelerator.hpp:12); the device layout is always the 4-wide
 * node format described in the library. */
typedef struct tr_bvh_opts {
    int32_t max_leaf_primitives;  /* leaf size bound for the binary SAH pass (reference default 3) */
    int32_t num_threads;          /* host build threads, 0 = all */
} tr_bvh_opts;

enum { TR_FILTER_BOX = 0, TR_FILTER_GAUSSIAN = 1 };
enum { TR_TR_ANALYTIC = 0, TR_TR_RATIO = 1 };

/* PTRendererParams (src/factory/renderer.hpp:12-19) plus the film and the
 * wavefront/guide controls that have no reference counterpart. */
typedef struct tr_render_params {
    int32_t width, height;
    int32_t spp;
    int32_t min_depth;                 /* Russian roulette (survival 0.9) when depth > min_depth */
    int32_t max_depth;
    int32_t direct_light_sample_num;   /* light samples per vertex, averaged (reference default 1; 0 = 1;
                                          > 1 with guide = 0 only, at most 64) */
    int32_t filter_type;               /* TR_FILTER_* (reference: Gaussian only) */
    float filter_radius;               /* pixels, in (0, 16] */
    float filter_alpha;
    int32_t shadow_transmittance;      /* TR_TR_* */
    uint64_t seed;
    int32_t sample_offset;             /* index of the first sample, for progressive rendering */
    int32_t samples_per_pass;          /* samples per pixel in flight at once, 0 = auto */
    int32_t guide;                     /* 0 = plain path tracer, 1 = learned path guiding (the network, and
                                          the learned light sampling of guide_light_grid when > 0),
                                          2 = the learned light sampling alone (needs guide_light_grid > 0) */
    int32_t guide_train_passes;        /* passes that train the guide before it is used (guide=1) */
    float guide_fraction;              /* one-sample MIS probability of the guide (0.5) */
    float guide_lr;                    /* Adam learning rate (1e-3) */
    int32_t profile;                   /* 1 = time kernels with HIP events; 2 = also count BVH
                                          node visits / triangle tests (slower) */
    int32_t guide_max_depth;           /* guide only vertices of depth < this (0 = every depth) */
    float guide_train_fraction;        /* MIS probability of the guide during training passes (0: they
                                          sample the BSDF alone and only record, so their image samples
                                          have the plain estimator's variance) */
    int32_t tail_paths;                /* plain passes: once at most this many paths are alive, one
                                          kernel carries them to their ends (0 = 131072, < 0 = never) */
    int32_t guide_light_grid;          /* guide = 1, 2: > 0 learns where light comes from: the scene box
                                          is cut into guide_light_grid^3 cells (at most 64), training passes
                                          learn per cell which lights and which parts of them deliver
                                          unoccluded light, and later passes draw next-event light samples
                                          from that (mixed with the power-based choice); 0 = off.  What is
                                          learned carries over to later calls on the context (progressive
                                          rendering) and starts again with a call that trains from sample 0
                                          (sample_offset = 0), a new scene or tr_guide_init */
} tr_render_params;

typedef struct tr_render_stats {
    uint64_t camera_paths;
    uint64_t extension_rays;      /* closest-hit rays traced */
    uint64_t shadow_rays;         /* any-hit rays traced */
    uint64_t nonfinite_dropped;   /* samples dropped for NaN/Inf (pixel_sampler_renderer.cpp:83-87) */
    uint64_t guide_fallbacks;     /* guided vertices that fell back to BSDF-only sampling */
    uint64_t guide_train_samples;
    double ms_total;              /* wall time of the render on the device (HIP events) */
    double ms_extend;             /* k_extend total (profile = 1) */
    double ms_shade;
    double ms_shadow;
    double ms_other;
    double ms_guide;              /* guide inference (profile = 1) */
    uint64_t extend_launches;
    uint64_t node_visits;         /* profile = 2: BVH nodes visited by extension rays */
    uint64_t tri_tests;           /* profile = 2: triangle tests by extension rays */
    uint64_t guide_vertices;      /* surface vertices sampled with the guide mixture */
    uint64_t guide_inferences;    /* network evaluations (surface hits of guided passes) */
    double ms_guide_train;        /* guide training between passes (profile = 1) */
    uint64_t tail_extension_rays; /* extension rays traced by the tail kernel (included above) */
    double ms_tail;               /* tail kernel (profile = 1) */
} tr_render_stats;

/* ---- context ---------------------------------------------------------- */
tr_status tr_create(tr_ctx** out, int32_t device);
void tr_destroy(tr_ctx* ctx);
const char* tr_last_error(const tr_ctx* ctx);
/* Returns a static string describing the library build (arch, version). */
const char* tr_version(void);
/* Page-locked host memory, e.g. for film buffers that tr_render fills every
 * frame: device-to-host copies into it run by DMA at full PCIe bandwidth
 * instead of through a staging buffer.  No counterpart in the reference (its
 * film lives in host memory).  Free with tr_host_free, whose ctx may be NULL
 * (for buffers that outlive their context). */
tr_status tr_host_alloc(tr_ctx* ctx, int64_t bytes, void** ptr);
tr_status tr_host_free(tr_ctx* ctx, void* ptr);

/* ---- scene (replaces Aggregate::build + create_general_scene,
 *      src/core/aggregate.hpp:19, src/factory/scene.hpp:14) ---------------- */
tr_status tr_scene_upload(tr_ctx* ctx, const tr_scene_desc* scene);
/* BVHAccel::build (src/accelerator/bvh.cpp:78-105): host SAH build, 4-wide
 * collapse, upload.  Requires tr_scene_upload first. */
tr_status tr_bvh_build(tr_ctx* ctx, const tr_bvh_opts* opts);
/* Scene::set_camera (src/core/scene.hpp:15-31) for an uploaded scene: the
 * camera of later renders and camera batches, without rebuilding the BVH.
 * TR_ERR_INVALID for a degenerate camera (position = target, up parallel to
 * the view direction, fov outside (0, pi), non-finite values). */
tr_status tr_set_camera(tr_ctx* ctx, const tr_camera* camera);
/* Number of nodes / bytes of the device BVH and triangle arrays. */
tr_status tr_bvh_info(tr_ctx* ctx, int64_t* num_nodes, int64_t* node_bytes, int64_t* tri_bytes);

/* Host-only build of the device BVH (no GPU needed): 4-wide nodes of 128 bytes
 * and 48-byte leaf-ordered triangles, as uploaded by tr_bvh_build.  Call with
 * nodes_out = tris_out = NULL to get *num_nodes first.  For inspection and
 * tests. */
tr_status tr_bvh4_export(const tr_scene_desc* scene, const tr_bvh_opts* opts, void* nodes_out, int64_t nodes_cap,
                         void* tris_out, int64_t* num_nodes, int32_t* depth);
/* The first n outputs of the PCG32 stream the renderer draws for camera
 * sample `sample` of pixel `pixel` (host evaluation of the device code). */
void tr_rng_sequence(uint64_t seed, uint32_t pixel, uint32_t sample, int64_t n, uint32_t* out);

/* ---- render (replaces Renderer::render, src/core/renderer.hpp:16, for the
 *      PathTraceRenderer of src/renderer/path_trace_renderer.cpp) ---------- */
/* film_sums: W*H*4 doubles (R, G, B, filter weight), accumulated onto what the
 * caller passes in when accumulate != 0 (progressive), else overwritten.
 * image: W*H*3 floats, sums / weight (Film::write_render_target,
 * src/core/render.hpp:105-115).  Either may be NULL. */
tr_status tr_render(tr_ctx* ctx, const tr_render_params* params, double* film_sums,
                    int32_t accumulate, float* image, tr_render_stats* stats);

/* ---- bidirectional path tracing (replaces BDPTRenderer::render,
 *      src/renderer/bdpt_renderer.cpp:1087-1185, the renderer behind
 *      create_bdpt_renderer, src/factory/renderer.hpp:39-48) ------------- */
typedef struct tr_bdpt_params {
    int32_t width, height;
    int32_t spp;
    int32_t max_camera_vertex_count;   /* vertices of a camera subpath, the camera's included (reference
                                          default 10), in [1, 32] */
    int32_t max_light_vertex_count;    /* vertices of a light subpath, the light's included (reference
                                          default 10), in [0, 32] */
    int32_t filter_type;               /* TR_FILTER_*: the filter of the camera-side estimate; the light-
                                          tracing (t = 1) splats go to the pixel they land in, weighted
                                          1 / spp, as the reference adds its splat image */
    float filter_radius;
    float filter_alpha;
    uint64_t seed;
    int32_t sample_offset;             /* index of the first sample, for progressive rendering */
    int32_t samples_per_pass;          /* samples per pixel in flight at once, 0 = auto */
    int32_t only_s;                    /* testing: 0 = every strategy, MIS-weighted; k > 0 keeps only the
                                          strategies with s = k - 1 light vertices, each with weight 1 */
    int32_t only_t;                    /* testing: 0 = every strategy; k > 0 keeps only the strategies with
                                          t = k camera vertices, each with weight 1 */
    int32_t profile;                   /* 1 = time the kernels with HIP events */
    int32_t reserved;
} tr_bdpt_params;

/* Renders with every connection strategy (s light vertices, t camera
 * vertices) of the reference, combined by the balance heuristic; every light
 * path of at most max_camera_vertex_count + max_light_vertex_count vertices
 * gets weight 1 in total.  film_sums, accumulate, image and stats as for
 * tr_render: the light-tracing splats of the call are folded into the
 * call's film sums so that sums / weight resolves to the camera-side estimate
 * plus splats / spp.  stats->extension_rays counts subpath rays,
 * shadow_rays connection rays; with profile = 1, ms_extend is subpath
 * traversal, ms_shade the subpath vertices, ms_tail the connection
 * strategies, ms_shadow their visibility, ms_other subpath starts, gather
 * and film.  The directional sun light has no
 * counterpart in the reference's BDPT: scenes with one are rejected
 * (TR_ERR_INVALID). */
tr_status tr_render_bdpt(tr_ctx* ctx, const tr_bdpt_params* params, double* film_sums, int32_t accumulate,
                         float* image, tr_render_stats* stats);

/* ---- ray queries (replace Aggregate::intersect_p / intersect,
 *      src/accelerator/bvh.cpp:112-185, with the watertight test) --------- */
/* rays: n * 8 floats (ox, oy, oz, tmin, dx, dy, dz, tmax); directions need not
 * be normalised.  hits: n * 4 floats (t, prim as int32 bits, b1, b2); prim =
 * -1 on a miss; b1/b2 weight the triangle's 2nd/3rd vertex.  Ties in t go to
 * the lowest primitive index. */
tr_status tr_trace_closest(tr_ctx* ctx, const float* rays, float* hits, int64_t n);
tr_status tr_trace_any(tr_ctx* ctx, const float* rays, int32_t* occluded, int64_t n);
/* Device-pointer forms, enqueued on the context stream. */
tr_status tr_trace_closest_device(tr_ctx* ctx, const float* d_rays, float* d_hits, int64_t n);
tr_status tr_trace_any_device(tr_ctx* ctx, const float* d_rays, int32_t* d_occluded, int64_t n);
/* tr_trace_closest_device without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Same rays, hits
 * and results as tr_trace_closest_device; the traversal stack spill area is a
 * static 12.6 MB array per device instead of the context's allocation, so the
 * scene's BVH may need at most 56 stack entries (4-wide depth 17) — deeper
 * trees get TR_ERR_INVALID — and at most 32768 threads traverse at once.  Calls
 * that overlap on one GPU (any contexts) share that array and race. */
tr_status tr_trace_closest_device_noalloc(tr_ctx* ctx, const float* d_rays, float* d_hits, int64_t n);
/* Closest-hit with per-ray visit counters (nodes, triangles) for the
 * algorithmic-bytes model; slower, never used by tr_render. */
tr_status tr_trace_closest_counted(tr_ctx* ctx, const float* rays, float* hits,
                                   uint32_t* node_visits, uint32_t* tri_tests, int64_t n);
tr_status tr_stream(tr_ctx* ctx, void** hip_stream);
tr_status tr_synchronize(tr_ctx* ctx);

/* ---- BSDF batches (replace BSDF::eval / pdf / sample of DisneyBSDF and the
 *      Phong/Lambert AggregateBSDF, src/core/bsdf.hpp:95-112) ------------- */
/* All vectors in the shading frame (shading normal = +z).  ng: geometric
 * normal per entry, or NULL (ng = +z).  f: n*3, pdf: n. */
tr_status tr_bsdf_eval_batch(tr_ctx* ctx, const tr_material* mats, int32_t num_mats,
                             const int32_t* mat_ids, const float* wi, const float* wo,
                             const float* ng, int64_t n, float* f, float* pdf);
/* u: n*3 uniforms (lobe selector, u, v).  wi: n*3, f: n*3, pdf: n
 * (pdf = 0 for an invalid sample). */
tr_status tr_bsdf_sample_batch(tr_ctx* ctx, const tr_material* mats, int32_t num_mats,
                               const int32_t* mat_ids, const float* wo, const float* ng,
                               const float* u, int64_t n, float* wi, float* f, float* pdf);

/* The film of the last tr_render (kept on the device) as 8-bit RGB, rows of
 * width pixels (cap >= width * height * 3 bytes): colour = sums / weight,
 * then create_aces_tone_mapper(aces_exposure) when aces_exposure >= 0
 * (src/post_processor/aces.cpp), colour^inv_gamma when inv_gamma > 0
 * (src/post_processor/gamma.cpp), then clamp(int(255 colour), 0, 255)
 * (src/main.cpp:290-299). */
tr_status tr_film_to_ldr(tr_ctx* ctx, float aces_exposure, float inv_gamma, uint8_t* rgb, int64_t cap);

/* Environment light of the uploaded scene (after tr_bvh_build), for tests:
 * IBL::sample_li with uniforms u (n*2: u picks phi, v theta) -> unit
 * directions wi (n*3), solid-angle pdf (0 = no sample) and radiance L (n*3)
 * (src/light/environment_light.cpp:92-119); and IBL::light_emit / pdf at
 * unit directions (n*3) (:179-196).  TR_ERR_STATE without an environment. */
tr_status tr_env_sample_batch(tr_ctx* ctx, const float* u, int64_t n, float* wi, float* pdf, float* L);
tr_status tr_env_eval_batch(tr_ctx* ctx, const float* dirs, int64_t n, float* L, float* pdf);
/* Lookups in the uploaded scene (after tr_bvh_build), for tests: the density
 * of grid medium `medium` at world positions (n*3)
 * (HeterogeneousMedium's density->evaluate_s(world_to_local(p)),
 * src/medium/heterogeneous.cpp:45-47 with src/utility/sampler.hpp:47-67), and
 * texture `texture` (1-based, as materials name it) at uv (n*2) -> rgb (n*3)
 * (Texture2D::evaluate, src/core/texture.hpp:57-83, texels interpolated as
 * floats).  TR_ERR_INVALID for a medium without a grid or a texture id out of
 * range; TR_ERR_STATE when no material names a texture. */
tr_status tr_medium_density_batch(tr_ctx* ctx, int32_t medium, const float* pos, int64_t n, float* density);
/* The albedo (n*3) and asymmetry (n) that delta tracking in grid medium
 * `medium` uses at world points pos (n*3): albedo->evaluate and
 * g->evaluate_s at the local position (src/medium/heterogeneous.cpp:66-78)
 * from the medium's albedo and g grids, or its constants. */
tr_status tr_medium_albedo_g_batch(tr_ctx* ctx, int32_t medium, const float* pos, int64_t n, float* albedo, float* g);
tr_status tr_texture_eval_batch(tr_ctx* ctx, int32_t texture, const float* uv, int64_t n, float* rgb);

/* Test batches over the uploaded scene (after tr_bvh_build) through the
 * renderer's own device code.
 *
 * tr_camera_ray_batch: ThinLensCamera::generate_ray
 * (src/camera/thin_lens_camera.cpp:40-59) at film positions film_uv (n*2, in
 * [0, 1]: pixel position / film size) with lens samples lens_uv (n*2) ->
 * rays (n*8: origin, 0, unit direction, 3e38).
 *
 * tr_light_pmf: the light-selection pmf of next-event estimation
 * (compute_light_power_distribution, src/renderer/direct_illumination.cpp:257-264,
 * over DiffuseLight::power = pi L A and IBL::power = radiance integral * R^2,
 * src/light/diffuse_light.cpp:17-20, src/light/environment_light.cpp:88-90):
 * entries 0..num_area_lights-1 are the emissive triangles in scene order, then
 * the environment, then the sun (0 when absent).  Call with pmf = NULL to get
 * *n = num_area_lights + 2.
 *
 * tr_light_pick_batch: Distribution1D::sample_discrete
 * (src/utility/distribution.hpp:49-67) of that pmf at uniforms u (n, in
 * [0, 1)) -> light index (same numbering) and its pmf.
 *
 * tr_hg_sample_batch: HenyeyGreensteinPhaseFunction::sample
 * (src/medium/phase_function.hpp:23-48) around wo (n*3) with uniforms u (n*2)
 * -> wi (n*3) and pdf (n).  Needs no scene.
 *
 * tr_film_accumulate_batch: the film's Tile::add_sample + merge
 * (src/core/render.hpp:37-57, 90-100) as tr_render applies it: radiance holds
 * width * height * spp RGB samples, sample s of pixel (x, y) at index
 * (s * height + y) * width + x, placed at the film position tr_render draws for
 * sample sample_offset + s of that pixel (seed of params); params give the
 * filter.  film_sums (W*H*4 doubles) is overwritten.  At most 2^26 samples.
 *
 * tr_shading_frame_batch: the shading frame of SurfaceIntersection
 * (src/shape/triangle.cpp:178-211: shading_coord from the normal derivatives
 * of compute_ss_ts, src/utility/coordinate.hpp:54-68, or coordinate(sn),
 * src/utility/geometry.hpp:887-894) as the renderer builds it at a hit of
 * primitive prims[i] (scene order) with barycentrics bary (n*2: weights of
 * vertex 1 and vertex 2) -> frames (n*9: x, y, z axes).  At most 2^26 hits. */
tr_status tr_camera_ray_batch(tr_ctx* ctx, const float* film_uv, const float* lens_uv, int64_t n, float* rays);
tr_status tr_light_pmf(tr_ctx* ctx, float* pmf, int64_t cap, int64_t* n);
tr_status tr_light_pick_batch(tr_ctx* ctx, const float* u, int64_t n, int32_t* idx, float* pmf);
tr_status tr_hg_sample_batch(tr_ctx* ctx, float g, const float* wo, const float* u, int64_t n, float* wi, float* pdf);
tr_status tr_shading_frame_batch(tr_ctx* ctx, const int32_t* prims, const float* bary, int64_t n, float* frames);
tr_status tr_film_accumulate_batch(tr_ctx* ctx, const tr_render_params* params, const float* radiance,
                                   double* film_sums);

/* ---- learned path guiding (no reference counterpart) ---------------------
 * MLP: 96 input features (81 used: position in the scene box, its sin/cos at
 * 12 octaves per axis, shading normal, outgoing direction) -> 4 hidden ReLU
 * layers of `hidden` units -> 48 outputs (40 used: per lobe k of 8, logit,
 * mean (3, normalised), log concentration clamped to [-4, 10]).  bf16 weights
 * and activations on MFMA with FP32 accumulation; FP32 master weights and
 * Adam.  Parameter layout (floats): for each layer l = 0..4 its weights
 * W_l[in][out] row-major (in = 96 for l = 0, else hidden; out = 48 for l = 4,
 * else hidden), then the biases of layers 0..4. */
typedef struct tr_guide_config {
    int32_t hidden;            /* 64 or 128 */
    int32_t batch;             /* training samples per Adam step (1 << 18) */
    int32_t steps_per_pass;    /* Adam steps after each training pass (16) */
    int32_t record_capacity;   /* records aimed at per training pass (1 << 20); buffers grow so none is dropped */
    float lr, beta1, beta2, eps;   /* Adam (1e-3, 0.9, 0.99, 1e-8) */
    float weight_clamp;        /* training weights are clamped at this multiple of their mean (1000) */
    int32_t reserved0;
    uint64_t seed;             /* parameter initialisation */
} tr_guide_config;

/* Fills *cfg with the defaults above. */
void tr_guide_default_config(tr_guide_config* cfg);
/* (Re)initialises the network and its Adam state; cfg NULL = defaults.
 * tr_render with guide = 1 calls this with defaults if it was never called. */
tr_status tr_guide_init(tr_ctx* ctx, const tr_guide_config* cfg);
tr_status tr_guide_num_params(tr_ctx* ctx, int64_t* n);
tr_status tr_guide_get_params(tr_ctx* ctx, float* params, int64_t n);
/* Replaces the master parameters (and their bf16 copies); Adam state is reset. */
tr_status tr_guide_set_params(tr_ctx* ctx, const float* params, int64_t n);
/* Box of the position encoding; by default (and after tr_guide_init) the
 * bounds of the uploaded scene. */
tr_status tr_guide_set_box(tr_ctx* ctx, const float lo[3], const float hi[3]);
/* points: n * 9 floats (x, n, wo); raw: n * 48 head outputs. */
tr_status tr_guide_forward(tr_ctx* ctx, const float* points, int64_t n, float* raw);
/* One training batch from host records, n * 16 floats each (x, n, wo, wi,
 * weight, 3 unused).  flags bit 0: take an Adam step; bit 1: normalise the
 * weights (w * n / sum(w), clamped at weight_clamp, / n) as training during
 * rendering does, else use them as given.  loss: sum of w * -log(q(wi) + 1e-3)
 * over the batch; grad: num_params floats, its gradient.  Either may be NULL. */
tr_status tr_guide_train_batch(tr_ctx* ctx, const float* records, int64_t n, int32_t flags, float* loss,
                               float* grad);
/* The mixture head as the renderer uses it, on given head outputs raw (n*48):
 * samples (n*3) drawn with uniforms u (n*3), the density at them (pdf_samples)
 * and at the unit directions dirs (n*3) (pdf_dirs); both -1 when raw is not
 * finite (the renderer's BSDF-only fallback). */
tr_status tr_guide_mixture_eval(tr_ctx* ctx, const float* raw, const float* dirs, const float* u, int64_t n,
                                float* samples, float* pdf_samples, float* pdf_dirs);
/* Weighted training samples of the most recent training pass of tr_render
 * (the layout tr_guide_train_batch takes, with the sampling pdf in float 14):
 * copies min(cap, count) of them to out and the count to *n. */
tr_status tr_guide_last_samples(tr_ctx* ctx, float* out, int64_t cap, int64_t* n);
/* MFMA layout check: D (16x16) = bf16(A) (16x32) * bf16(B) (32x16), row-major. */
tr_status tr_guide_mfma_selftest(tr_ctx* ctx, const float* A, const float* B, float* D);

#ifdef __cplusplus
}
#endif
#endif /* TRACER_HIP_H */
