// C ABI of libtracer_hip.so (include/tracer_hip.h): context, scene upload,
// BVH build, render orchestration and the batch query entry points.
#include "tracer_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "bvh_build.h"
#include "capi_internal.h"

using namespace trhip;
using namespace trcapi;

// The ctypes bindings (tracer_amd/abi.py) assume these layouts.
static_assert(sizeof(tr_material) == 132, "tr_material layout");
static_assert(sizeof(tr_medium) == 152, "tr_medium layout");
static_assert(sizeof(tr_camera) == 52, "tr_camera layout");
static_assert(sizeof(tr_render_params) == 96, "tr_render_params layout");
static_assert(sizeof(tr_render_stats) == 160, "tr_render_stats layout");
static_assert(sizeof(tr_bdpt_params) == 64, "tr_bdpt_params layout");
static_assert(sizeof(tr_guide_config) == 48, "tr_guide_config layout");

namespace {

// Eighths of a wave's traversing lanes that must hold a leaf before the wave
// tests triangles, for closest-hit and any-hit rays (TraceScene::leaf_share,
// leaf_share_any).  TRACER_HIP_LEAF_SHARE / TRACER_HIP_LEAF_SHARE_ANY (0-8)
// override them for tuning.
int env_share(const char* name, int dflt) {
    const char* e = std::getenv(name);
    if (e && e[0] >= '0' && e[0] <= '8' && e[1] == '\0') return e[0] - '0';
    return dflt;
}
int trace_leaf_share() {
    static const int v = env_share("TRACER_HIP_LEAF_SHARE", 4);
    return v;
}
int trace_leaf_share_any() {
    static const int v = env_share("TRACER_HIP_LEAF_SHARE_ANY", 4);
    return v;
}
// Rounds a shade window must span before the plain shade kernels order it by
// shading class (RenderConsts::shade_sort_rounds); TRACER_HIP_SHADE_SORT (0-8)
// overrides it for tuning, 5 and above never sort.
int shade_sort_rounds() {
    static const int v = env_share("TRACER_HIP_SHADE_SORT", 2);
    return v;
}

// DisneyBSDF constructor (src/material/disney.cpp:347-399) evaluated once per material.
DevMaterial make_material_untextured(const tr_material& m);

void check_material(const tr_material& m) {
    if (m.type != TR_MAT_LAMBERT && m.type != TR_MAT_DISNEY && m.type != TR_MAT_PHONG && m.type != TR_MAT_GLASS)
        throw InvalidArg("unknown material type");
    if (m.type == TR_MAT_GLASS) {
        if (!(m.ior > 0.0f && m.ior < 1e30f)) throw InvalidArg("glass ior must be finite and > 0");
        for (int k = 0; k < 3; ++k)
            if (!(m.base_color[k] >= 0.0f && m.specular_scale[k] >= 0.0f)) throw InvalidArg("glass kr and kt must be >= 0");
    }
    if (m.type == TR_MAT_PHONG) {
        if (!(m.shininess >= 0.0f && m.shininess < 1e30f)) throw InvalidArg("Phong shininess must be finite and >= 0");
        for (int k = 0; k < 3; ++k)
            if (!(m.base_color[k] >= 0.0f && m.specular_scale[k] >= 0.0f)) throw InvalidArg("Phong kd and ks must be >= 0");
    }
}

// Phong::shading (src/material/phong.cpp:29-60) with constant textures: kd
// and ks divided by max(1, max_i(kd_i + ks_i)), lobe weights by luminance
// (0.5 each when both are black).  The aggregate's albedo kd + ks is what the
// black-fringe fallback uses (src/material/aggregate_bsdf.hpp:19-23).
static DevMaterial make_phong(const tr_material& m) {
    DevMaterial d{};
    d.type = MAT_PHONG;
    f3 kd = mk(m.base_color[0], m.base_color[1], m.base_color[2]);
    f3 ks = mk(m.specular_scale[0], m.specular_scale[1], m.specular_scale[2]);
    const float dem = std::max(1.0f, max3(kd + ks));
    kd = kd * (1.0f / dem);
    ks = ks * (1.0f / dem);
    const float ld = lum(kd), ls = lum(ks);
    d.w_diffuse = ld + ls > kEps ? ld / (ld + ls) : 0.5f;
    d.w_spec = 1.0f - d.w_diffuse;
    d.C = kd + ks;
    d.Ctint = kd;
    d.spec_scale = ks;
    d.ns = m.shininess;
    d.guide_share = d.w_diffuse;   // the guide shares the broad (diffuse) lobe's selection weight
    return d;
}

// Glass::shading (src/material/glass.cpp:96-112): reflection colour kr,
// refraction colour sqrt(kt), the ior as given.  Never guided.
static DevMaterial make_glass(const tr_material& m) {
    DevMaterial d{};
    d.type = MAT_GLASS;
    d.C = mk(m.base_color[0], m.base_color[1], m.base_color[2]);
    d.spec_scale = mk(std::sqrt(m.specular_scale[0]), std::sqrt(m.specular_scale[1]), std::sqrt(m.specular_scale[2]));
    d.ior = m.ior;
    d.guide_share = 0.0f;
    return d;
}

DevMaterial make_material(const tr_material& m) {
    DevMaterial d = make_material_untextured(m);
    d.tex = m.base_color_texture;
    if (d.tex > 0 && d.type == MAT_PHONG)   // kept raw: normalised with kd per hit
        d.spec_scale = mk(m.specular_scale[0], m.specular_scale[1], m.specular_scale[2]);
    return d;
}

// Whether a material reads a texture per hit (base colour / kd, a Disney
// parameter or a normal map); has_param_textures: the last two.
bool has_param_textures(const tr_material& m) {
    bool any = m.normal_texture > 0;
    for (int k = 0; k < TR_PARAM_COUNT; ++k) any = any || m.param_texture[k] > 0;
    return any;
}
bool uses_texture(const tr_material& m) { return m.base_color_texture > 0 || has_param_textures(m); }

// The constant parameters of a Disney (or Lambert) material.
DisneyIn disney_in(const tr_material& m) {
    DisneyIn p;
    p.base_color = mk(m.base_color[0], m.base_color[1], m.base_color[2]);
    p.metallic = m.metallic;
    p.roughness = m.roughness;
    p.transmission = m.transmission;
    p.transmission_roughness = m.transmission_roughness;
    p.ior = m.ior;
    p.specular_scale = mk(m.specular_scale[0], m.specular_scale[1], m.specular_scale[2]);
    p.specular_tint = m.specular_tint;
    p.anisotropic = m.anisotropic;
    p.sheen = m.sheen;
    p.sheen_tint = m.sheen_tint;
    p.clearcoat = m.clearcoat;
    p.clearcoat_gloss = m.clearcoat_gloss;
    return p;
}

DevMaterial make_material_untextured(const tr_material& m) {
    if (m.type == TR_MAT_PHONG) return make_phong(m);
    if (m.type == TR_MAT_GLASS) return make_glass(m);
    DevMaterial d{};
    d.type = m.type == TR_MAT_DISNEY ? MAT_DISNEY : MAT_LAMBERT;
    disney_setup(disney_in(m), d);
    if (d.type == MAT_LAMBERT) {
        d.metallic = 0.0f;
        d.transmission = 0.0f;
    }
    // Guided sampling only where the BSDF is broad: the guide network does not
    // see material parameters, so at transmissive or glossy metal vertices its
    // samples would mostly land where f vanishes.  Elsewhere the guide takes
    // the broad part of the BSDF's lobe-selection weight (diffuse, and
    // specular once alpha_x >= 0.25) and leaves the sharp lobes (smooth
    // specular, clearcoat) to the BSDF sampler: the guide fraction at a vertex
    // is guide_fraction * guide_share.
    const bool guided = d.type == MAT_LAMBERT || (m.transmission <= 0.0f && (m.metallic < 0.5f || d.ax >= 0.25f));
    float share = 1.0f;
    if (d.type != MAT_LAMBERT) {
        const float tot = d.w_diffuse + d.w_trans + d.w_spec + d.w_cc;
        const float broad = d.w_diffuse + (d.ax >= 0.25f ? d.w_spec : 0.0f);
        share = tot > 0.0f ? broad / tot : 0.0f;
    }
    d.guide_share = guided ? share : 0.0f;
    return d;
}

// Inverse of a row-major 3x4 affine matrix (cofactors, double precision).
bool invert_affine(const float* a, double* out) {
    const double m00 = a[0], m01 = a[1], m02 = a[2], m10 = a[4], m11 = a[5], m12 = a[6];
    const double m20 = a[8], m21 = a[9], m22 = a[10];
    const double c00 = m11 * m22 - m12 * m21, c01 = m02 * m21 - m01 * m22, c02 = m01 * m12 - m02 * m11;
    const double c10 = m12 * m20 - m10 * m22, c11 = m00 * m22 - m02 * m20, c12 = m02 * m10 - m00 * m12;
    const double c20 = m10 * m21 - m11 * m20, c21 = m01 * m20 - m00 * m21, c22 = m00 * m11 - m01 * m10;
    const double det = m00 * c00 + m01 * c10 + m02 * c20;
    if (!(std::abs(det) > 1e-30) || !std::isfinite(det)) return false;
    const double r[9] = {c00 / det, c01 / det, c02 / det, c10 / det, c11 / det, c12 / det, c20 / det, c21 / det, c22 / det};
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) out[4 * i + j] = r[3 * i + j];
        out[4 * i + 3] = -(r[3 * i] * a[3] + r[3 * i + 1] * a[7] + r[3 * i + 2] * a[11]);
    }
    return true;
}

// A grid medium (HeterogeneousMedium, src/medium/heterogeneous.cpp:12-29):
// the majorant is the largest voxel, at least eps = 3e-4 (src/common.hpp:47);
// sigma_s carries the albedo.
DevMedium make_grid_medium(const tr_medium& m, const std::vector<float>& grid, int offset) {
    DevMedium d{};
    d.sigma_s = mk(m.sigma_s[0], m.sigma_s[1], m.sigma_s[2]);
    d.g = std::min(std::max(m.g, -1.0f + kEps), 1.0f - kEps);
    float mx = 0.0f;
    for (float v : grid) mx = std::max(mx, v);
    d.majorant = std::max(mx, kEps);
    d.sigma_t = f3s(d.majorant);
    d.max_scatter = m.max_scattering_count;   // 0: no scattering, transmittance only (as the reference)
    d.grid = offset;
    d.nx = m.nx;
    d.ny = m.ny;
    d.nz = m.nz;
    double inv[12];
    invert_affine(m.local_to_world, inv);
    for (int k = 0; k < 12; ++k) d.w2l[k] = float(inv[k]);
    d.white_indirect = m.white_for_indirect ? 1 : 0;
    d.albedo_grid = -1;
    d.g_grid = -1;
    return d;
}

DevMedium make_medium(const tr_medium& m) {
    DevMedium d{};
    d.grid = -1;
    d.albedo_grid = -1;
    d.g_grid = -1;
    d.sigma_s = mk(m.sigma_s[0], m.sigma_s[1], m.sigma_s[2]);
    const f3 sa = mk(m.sigma_a[0], m.sigma_a[1], m.sigma_a[2]);
    d.sigma_t = sa + d.sigma_s;
    // HomogeneousMedium clamps g into (-1, 1) (src/medium/homogeneous.cpp:16-17)
    d.g = std::min(std::max(m.g, -1.0f + kEps), 1.0f - kEps);
    d.majorant = max3(d.sigma_t);
    d.max_scatter = m.max_scattering_count;   // 0: no scattering, transmittance only (as the reference)
    if (d.majorant > 0.0f) {
        d.p_absorb = (sa.x + sa.y + sa.z) / (3.0f * d.majorant);
        d.p_scatter = (d.sigma_s.x + d.sigma_s.y + d.sigma_s.z) / (3.0f * d.majorant);
        d.p_null = 1.0f - d.p_absorb - d.p_scatter;
        if (d.p_null < 1e-6f) {
            d.p_null = 0.0f;
            d.p_scatter = 1.0f - d.p_absorb;
        }
    } else {
        d.majorant = 1e-30f;
        d.p_absorb = 0.0f;
        d.p_scatter = 0.0f;
        d.p_null = 1.0f;
    }
    return d;
}

void validate_scene(const tr_scene_desc* d) {
    if (!d) throw InvalidArg("scene is NULL");
    if (d->num_vertices <= 0 || !d->positions) throw InvalidArg("scene has no vertices");
    if (d->num_triangles <= 0 || !d->indices || !d->material_ids) throw InvalidArg("scene has no triangles");
    if (d->num_materials <= 0 || !d->materials) throw InvalidArg("scene has no materials");
    if (d->num_media < 0 || (d->num_media > 0 && !d->media)) throw InvalidArg("media array missing");
    if (d->has_env && d->env_width != 0) {
        if (d->env_width < 0 || d->env_height <= 0 || !d->env_pixels) throw InvalidArg("environment image: bad size or NULL texels");
        if (int64_t(d->env_width) * d->env_height > (int64_t(1) << 28)) throw InvalidArg("environment image too large");
        const size_t n = size_t(d->env_width) * size_t(d->env_height) * 3;
        for (size_t i = 0; i < n; ++i)
            if (!(d->env_pixels[i] >= 0.0f && d->env_pixels[i] < 3.0e38f))
                throw InvalidArg("environment texels must be finite and >= 0");
        // a rotation: R R^T = I (the light -> world map is its transpose)
        const float* r = d->env_world_to_light;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const float dotij = r[3 * i] * r[3 * j] + r[3 * i + 1] * r[3 * j + 1] + r[3 * i + 2] * r[3 * j + 2];
                if (!(std::fabs(dotij - (i == j ? 1.0f : 0.0f)) < 1e-4f))
                    throw InvalidArg("env_world_to_light must be a rotation");
            }
    }
    // non-finite vertex data would reach the BVH build (bin index of a NaN
    // centroid) and the shading frames
    for (int64_t i = 0; i < 3 * d->num_vertices; ++i)
        if (!std::isfinite(d->positions[i])) throw InvalidArg("vertex positions must be finite");
    if (d->normals)
        for (int64_t i = 0; i < 3 * d->num_vertices; ++i)
            if (!std::isfinite(d->normals[i])) throw InvalidArg("vertex normals must be finite");
    if (d->uvs)
        for (int64_t i = 0; i < 2 * d->num_vertices; ++i)
            if (!std::isfinite(d->uvs[i])) throw InvalidArg("vertex uvs must be finite");
    if (d->emission)
        for (int64_t i = 0; i < 3 * d->num_triangles; ++i)
            if (!(d->emission[i] >= 0.0f && d->emission[i] < 3.0e38f)) throw InvalidArg("emission must be finite and >= 0");
    for (int64_t t = 0; t < d->num_triangles; ++t) {
        for (int k = 0; k < 3; ++k) {
            const int32_t v = d->indices[3 * t + k];
            if (v < 0 || v >= d->num_vertices) throw InvalidArg("triangle " + std::to_string(t) + " has a vertex index out of range");
        }
        const int32_t m = d->material_ids[t];
        if (m < 0 || m >= d->num_materials) throw InvalidArg("triangle " + std::to_string(t) + " has an invalid material id");
        if (d->medium_inside) {
            const int32_t mi = d->medium_inside[t];
            if (mi < -1 || mi >= d->num_media) throw InvalidArg("invalid medium_inside id");
        }
        if (d->medium_outside) {
            const int32_t mo = d->medium_outside[t];
            if (mo < -1 || mo >= d->num_media) throw InvalidArg("invalid medium_outside id");
        }
    }
    for (int i = 0; i < d->num_materials; ++i) check_material(d->materials[i]);
    if (d->num_textures < 0 || (d->num_textures > 0 && !d->textures)) throw InvalidArg("texture array missing");
    for (int i = 0; i < d->num_textures; ++i) {
        const tr_texture& t = d->textures[i];
        if (t.width <= 0 || t.height <= 0 || int64_t(t.width) * t.height > (int64_t(1) << 28))
            throw InvalidArg("texture " + std::to_string(i) + ": bad size");
        if ((t.rgb8 == nullptr) == (t.rgbf == nullptr))
            throw InvalidArg("texture " + std::to_string(i) + ": give exactly one of rgb8 and rgbf");
        if (t.rgbf)
            for (size_t k = 0; k < size_t(t.width) * size_t(t.height) * 3; ++k)
                if (!(t.rgbf[k] >= 0.0f && t.rgbf[k] < 3.0e38f))
                    throw InvalidArg("texture " + std::to_string(i) + ": texels must be finite and >= 0");
    }
    for (int i = 0; i < d->num_materials; ++i) {
        const tr_material& m = d->materials[i];
        int ids[TR_PARAM_COUNT + 2];
        ids[0] = m.base_color_texture;
        for (int k = 0; k < TR_PARAM_COUNT; ++k) ids[1 + k] = m.param_texture[k];
        ids[TR_PARAM_COUNT + 1] = m.normal_texture;
        for (int k = 0; k < TR_PARAM_COUNT + 2; ++k) {
            const int t = ids[k];
            if (t < 0 || t > d->num_textures) throw InvalidArg("material " + std::to_string(i) + ": texture id out of range");
            if (t > 0 && !d->uvs) throw InvalidArg("a textured material needs vertex uvs");
            if (t > 0 && m.type == TR_MAT_GLASS) throw InvalidArg("glass takes no texture");
            if (t > 0 && k > 0 && m.type != TR_MAT_DISNEY)
                throw InvalidArg("only Disney materials take parameter textures and normal maps");
        }
    }
    for (int i = 0; i < d->num_media; ++i) {
        const tr_medium& m = d->media[i];
        if (!(m.g > -1.0f && m.g < 1.0f)) throw InvalidArg("medium g must lie in (-1, 1)");  // homogeneous.cpp:16
        if (m.density) {
            if (m.nx < 1 || m.ny < 1 || m.nz < 1 || int64_t(m.nx) * m.ny * m.nz > (int64_t(1) << 31) - 1)
                throw InvalidArg("density grid dimensions must be positive and hold fewer than 2^31 voxels");
            const size_t n = size_t(m.nx) * size_t(m.ny) * size_t(m.nz);
            for (size_t k = 0; k < n; ++k)
                if (!(m.density[k] >= 0.0f && std::isfinite(m.density[k])))
                    throw InvalidArg("density grid values must be finite and >= 0");
            double inv[12];
            if (!invert_affine(m.local_to_world, inv)) throw InvalidArg("medium local_to_world is singular");
            auto check_dims = [](int x, int y, int z, int ch, const char* what) {
                if (x < 1 || y < 1 || z < 1 || int64_t(x) * y * z * ch > (int64_t(1) << 31) - 1)
                    throw InvalidArg(std::string(what) + " grid dimensions must be positive and hold fewer than 2^31 values");
                return size_t(x) * size_t(y) * size_t(z) * size_t(ch);
            };
            if (m.albedo_grid) {
                const size_t n = check_dims(m.anx, m.any, m.anz, 3, "albedo");
                for (size_t k = 0; k < n; ++k)
                    if (!(m.albedo_grid[k] >= 0.0f && std::isfinite(m.albedo_grid[k])))
                        throw InvalidArg("albedo grid values must be finite and >= 0");
            }
            if (m.g_grid) {
                const size_t n = check_dims(m.gnx, m.gny, m.gnz, 1, "g");
                for (size_t k = 0; k < n; ++k)
                    if (!(m.g_grid[k] >= -1.0f && m.g_grid[k] <= 1.0f))
                        throw InvalidArg("g grid values must lie in [-1, 1]");
            }
        } else if (m.albedo_grid || m.g_grid) {
            throw InvalidArg("albedo and g grids need a density grid");
        }
        for (int k = 0; k < 3; ++k)
            if (m.sigma_a[k] < 0 || m.sigma_s[k] < 0) throw InvalidArg("medium coefficients must be >= 0");
    }
}

// Cutpoint table over the light CDF (Chen and Asau 1974): entry j holds the
// first index whose CDF value exceeds j / K.  K is a power of two, so u * K and
// j / K are exact and a scan from entry floor(u * K) ends at the index the
// binary search of Distribution1D::sample_discrete finds
// (src/utility/distribution.hpp:49-67), after one or two CDF reads.
std::vector<int> light_cutpoints(const std::vector<float>& cdf) {
    const int n = int(cdf.size());
    int K = 1;
    while (K < n && K < (1 << 20)) K <<= 1;
    std::vector<int> g(static_cast<size_t>(K));
    int i = 0;
    for (int j = 0; j < K; ++j) {
        const float x = float(j) / float(K);
        while (i < n - 1 && !(cdf[size_t(i)] > x)) ++i;
        g[size_t(j)] = i;
    }
    return g;
}

// ThinLensCamera (thin_lens_camera.cpp:23-36) with look_at's basis
// (transform.cpp:96-124).
DevCamera device_camera(const tr_camera& cm) {
    DevCamera d{};
    const f3 pos = mk(cm.pos[0], cm.pos[1], cm.pos[2]);
    const f3 dir = normalize(mk(cm.target[0], cm.target[1], cm.target[2]) - pos);
    const f3 right = normalize(cross(normalize(mk(cm.up[0], cm.up[1], cm.up[2])), dir));
    d.pos = pos;
    d.dir = dir;
    d.right = right;
    d.up = cross(dir, right);
    d.focal = cm.focal_distance > 0.0f ? cm.focal_distance : 1.0f;
    d.film_h = 2.0f * d.focal * std::tan(cm.fov_y * 0.5f);
    d.film_w = (cm.aspect > 0.0f ? cm.aspect : 1.0f) * d.film_h;
    d.lens_radius = cm.lens_radius;
    return d;
}

void check_camera(const tr_camera& cm) {
    const float* v[3] = {cm.pos, cm.target, cm.up};
    for (const float* a : v)
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(a[k])) throw InvalidArg("camera vectors must be finite");
    const double dx = double(cm.target[0]) - cm.pos[0], dy = double(cm.target[1]) - cm.pos[1],
                 dz = double(cm.target[2]) - cm.pos[2];
    const double dl = std::sqrt(dx * dx + dy * dy + dz * dz);
    const double ul = std::sqrt(double(cm.up[0]) * cm.up[0] + double(cm.up[1]) * cm.up[1] + double(cm.up[2]) * cm.up[2]);
    if (!(dl > 0.0) || !(ul > 0.0)) throw InvalidArg("camera target must differ from its position and up must not be zero");
    const double cx = cm.up[1] * dz - cm.up[2] * dy, cy = cm.up[2] * dx - cm.up[0] * dz, cz = cm.up[0] * dy - cm.up[1] * dx;
    if (!(std::sqrt(cx * cx + cy * cy + cz * cz) > 1e-6 * dl * ul)) throw InvalidArg("camera up is parallel to its view direction");
    if (!(cm.fov_y > 0.0f && cm.fov_y < 3.14159f)) throw InvalidArg("camera fov_y must lie in (0, pi)");
    if (!(cm.lens_radius >= 0.0f && std::isfinite(cm.lens_radius))) throw InvalidArg("camera lens radius must be >= 0");
    if (!(cm.focal_distance >= 0.0f && std::isfinite(cm.focal_distance)) || !(cm.aspect >= 0.0f && std::isfinite(cm.aspect)))
        throw InvalidArg("camera focal distance and aspect must be finite and >= 0");
}

// compute_ss_ts (src/utility/coordinate.hpp:54-68) over the vertex-normal
// differences AB, AC and the uv differences (ab, ac): the derivative along v,
// ts = (m10 AB - m00 AC) / det, with w = 1 where SurfaceIntersection keeps it
// (det != 0 and both derivatives at least eps long, triangle.cpp:205-207);
// w = 0 leaves the frame to coordinate(sn) per hit.
float4 normal_dv(f3 AB, f3 AC, float abx, float aby, float acx, float acy) {
    const float m00 = abx, m01 = aby, m10 = acx, m11 = acy;
    const float det = m00 * m11 - m01 * m10;
    if (det == 0.0f) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float inv_det = 1.0f / det;
    const f3 ss = AB * (m11 * inv_det) - AC * (m01 * inv_det);
    const f3 ts = AB * (m10 * inv_det) - AC * (m00 * inv_det);
    if (dot(ss, ss) < kEps || dot(ts, ts) < kEps) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return f4(ts, 1.0f);
}

void build_device_scene(tr_ctx* c, const tr_bvh_opts* opts) {
    HostScene& h = c->hs;
    const int64_t nt = int64_t(h.mat.size());
    const int64_t nv = int64_t(h.P.size() / 3);
    BvhBuildOptions bo;
    if (opts) {
        bo.max_leaf = opts->max_leaf_primitives > 0 ? opts->max_leaf_primitives : 4;
        bo.threads = opts->num_threads;
    }
    BvhBuildResult r = build_bvh4(h.P.data(), nv, h.idx.data(), nt, bo);
    c->num_nodes = int64_t(r.nodes.size());
    c->bvh_depth = r.depth;
    c->lcache.grid = 0;   // a new scene: the light cache starts again
    c->lcache.built = false;
    c->max_stack = 3 * r.depth + 4;

    // leaf-ordered shading records; the light table lists the emissive
    // triangles in scene order, as the reference's scene.lights does, so that
    // the power distribution over it samples the same light for the same u
    std::vector<float4> shade(size_t(nt) * kShadeRecord);   // the fourth float4 (geometric normal) is filled on the device
    std::vector<int> tri_to_light(size_t(nt), -1);
    int nA = 0;
    {
        std::vector<int>& leaf_of = c->leaf_of;
        leaf_of.assign(size_t(nt), 0);
        for (int64_t i = 0; i < nt; ++i) leaf_of[size_t(r.tris[size_t(i)].prim)] = int(i);
        for (int64_t prim = 0; prim < nt && !h.emission.empty(); ++prim)
            if (h.emission[3 * prim] != 0.0f || h.emission[3 * prim + 1] != 0.0f || h.emission[3 * prim + 2] != 0.0f)
                tri_to_light[size_t(leaf_of[size_t(prim)])] = nA++;
    }
    std::vector<float4> lights(size_t(nA) * 4);
    std::vector<double> power(size_t(nA) + 2, 0.0);
    // world bounding sphere (Bounds3::bounding_sphere, geometry.hpp:752-755)
    const f3 wc = mk(0.5f * (h.lo[0] + h.hi[0]), 0.5f * (h.lo[1] + h.hi[1]), 0.5f * (h.lo[2] + h.hi[2]));
    const float wr = length(mk(h.hi[0], h.hi[1], h.hi[2]) - wc);
    for (int64_t i = 0; i < nt; ++i) {
        const TriPacked& tp = r.tris[size_t(i)];
        const int prim = tp.prim;
        f3 n[3];
        const f3 A = mk(tp.v0[0], tp.v0[1], tp.v0[2]), B = mk(tp.v1[0], tp.v1[1], tp.v1[2]), C = mk(tp.v2[0], tp.v2[1], tp.v2[2]);
        const f3 cr = cross(B - A, C - A);
        const f3 gn = length(cr) > 0.0f ? normalize(cr) : mk(0.0f, 0.0f, 1.0f);
        for (int k = 0; k < 3; ++k) {
            if (!h.N.empty()) {
                const int v = h.idx[3 * size_t(prim) + k];
                f3 nn = mk(h.N[3 * size_t(v)], h.N[3 * size_t(v) + 1], h.N[3 * size_t(v) + 2]);
                const float ln = length(nn);
                n[k] = ln > 0.0f ? nn * (1.0f / ln) : gn;
            } else {
                n[k] = gn;
            }
        }
        const int mi = h.med_in.empty() ? -1 : h.med_in[size_t(prim)];
        const int mo = h.med_out.empty() ? -1 : h.med_out[size_t(prim)];
        const int med = (mi & 0xffff) | (mo << 16);
        float4* rec = shade.data() + kShadeRecord * size_t(i);
        rec[0] = f4(n[0], __builtin_bit_cast(float, int32_t(h.mat[size_t(prim)])));
        rec[1] = f4(n[1], __builtin_bit_cast(float, int32_t(tri_to_light[size_t(i)])));
        rec[2] = f4(n[2], __builtin_bit_cast(float, int32_t(med)));
        rec[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const int li = tri_to_light[size_t(i)];
        if (li >= 0) {
            const f3 L = mk(h.emission[3 * prim], h.emission[3 * prim + 1], h.emission[3 * prim + 2]);
            const double area = 0.5 * std::sqrt(double(cr.x) * cr.x + double(cr.y) * cr.y + double(cr.z) * cr.z);
            // (L, area), (A, pmf), (B, triangle), (C, 0): the vertices travel with
            // the light so that sampling it needs no second dependent load
            lights[4 * size_t(li)] = f4(L, float(area));
            lights[4 * size_t(li) + 1] = f4(A, 0.0f);
            lights[4 * size_t(li) + 2] = f4(B, __builtin_bit_cast(float, int32_t(i)));
            lights[4 * size_t(li) + 3] = f4(C, 0.0f);
            power[size_t(li)] = double(kPi) * lum(L) * area;   // DiffuseLight::power, diffuse_light.cpp:17-20
        }
    }
    // IBL::power (environment_light.cpp:81-83): the radiance integral times R^2
    if (h.has_env && h.envmap.w > 0) {
        const double* ri = h.envmap.radiance_integral;
        power[size_t(nA)] = (0.2126 * ri[0] + 0.7152 * ri[1] + 0.0722 * ri[2]) * double(wr) * double(wr);
    } else if (h.has_env) {
        power[size_t(nA)] = double(lum(mk(h.env[0], h.env[1], h.env[2]))) * 4.0 * kPi * wr * wr;
    }
    if (h.has_sun) power[size_t(nA) + 1] = double(lum(mk(h.sun_E[0], h.sun_E[1], h.sun_E[2]))) * kPi * wr * wr;
    double tot = 0.0;
    for (double p : power) tot += p;
    std::vector<float> pmf(power.size(), 0.0f), cdf(power.size(), 1.0f);
    double acc = 0.0;
    for (size_t i = 0; i < power.size(); ++i) {
        pmf[i] = tot > 0.0 ? float(power[i] / tot) : 0.0f;
        acc += tot > 0.0 ? power[i] / tot : 0.0;
        cdf[i] = float(acc);
    }
    for (int i = 0; i < nA; ++i) lights[4 * size_t(i) + 1].w = pmf[size_t(i)];
    const std::vector<int> guide = light_cutpoints(cdf);

    std::vector<DevMaterial> dm(h.mats.size());
    c->has_delta = false;
    // Disney materials with parameter textures or a normal map: one row each
    // of texture ids (base colour, the twelve parameters, the normal map) and
    // constant parameters, referenced by DevMaterial::texset
    std::vector<int> mat_tex;
    std::vector<float> mat_raw;
    for (size_t i = 0; i < h.mats.size(); ++i) {
        const tr_material& m = h.mats[i];
        dm[i] = make_material(m);
        c->has_delta = c->has_delta || dm[i].type == MAT_GLASS;
        if (m.type == TR_MAT_DISNEY && has_param_textures(m)) {
            dm[i].texset = int(mat_tex.size() / kMatTexRow) + 1;
            mat_tex.push_back(m.base_color_texture);
            for (int k = 0; k < TR_PARAM_COUNT; ++k) mat_tex.push_back(m.param_texture[k]);
            mat_tex.push_back(m.normal_texture);
            mat_tex.resize(size_t(dm[i].texset) * kMatTexRow, 0);
            const DisneyIn p = disney_in(m);
            const float raw[kMatRawRow] = {p.base_color.x, p.base_color.y, p.base_color.z, p.metallic, p.roughness,
                                           p.transmission, p.transmission_roughness, p.ior, p.specular_scale.x,
                                           p.specular_scale.y, p.specular_scale.z, p.specular_tint, p.anisotropic,
                                           p.sheen, p.sheen_tint, p.clearcoat, p.clearcoat_gloss, 0.0f, 0.0f, 0.0f};
            mat_raw.insert(mat_raw.end(), raw, raw + kMatRawRow);
        }
    }
    std::vector<DevMedium> dmed(h.media.size());
    std::vector<float> voxels;
    for (size_t i = 0; i < h.media.size(); ++i) {
        if (h.grids[i].empty()) {
            dmed[i] = make_medium(h.media[i]);
        } else {
            dmed[i] = make_grid_medium(h.media[i], h.grids[i], int(voxels.size()));
            voxels.insert(voxels.end(), h.grids[i].begin(), h.grids[i].end());
            // albedo and g grids follow in the same pool
            if (!h.albedo_grids[i].empty()) {
                dmed[i].albedo_grid = int(voxels.size());
                dmed[i].anx = h.media[i].anx; dmed[i].any = h.media[i].any; dmed[i].anz = h.media[i].anz;
                voxels.insert(voxels.end(), h.albedo_grids[i].begin(), h.albedo_grids[i].end());
            }
            if (!h.g_grids[i].empty()) {
                dmed[i].g_grid = int(voxels.size());
                dmed[i].gnx = h.media[i].gnx; dmed[i].gny = h.media[i].gny; dmed[i].gnz = h.media[i].gnz;
                voxels.insert(voxels.end(), h.g_grids[i].begin(), h.g_grids[i].end());
            }
        }
    }
    if (voxels.size() > size_t(INT32_MAX)) throw InvalidArg("medium grids hold more than 2^31 values in total");

    hipStream_t st = c->stream;
    upload(c->nodes, r.nodes.data(), r.nodes.size() * sizeof(Node4), st);
    upload(c->tris, r.tris.data(), r.tris.size() * sizeof(TriPacked), st);
    upload(c->tri_shade, shade.data(), shade.size() * sizeof(float4), st);
    launch_geometric_normals(c->tris.as<float4>(), c->tri_shade.as<float4>(), nt, st);
    // shading class per leaf-ordered triangle (kept until the stream synchronises)
    std::vector<unsigned char> cls(static_cast<size_t>(nt));
    for (int64_t i = 0; i < nt; ++i) {
        const int type = h.mats[size_t(h.mat[size_t(r.tris[size_t(i)].prim)])].type;
        cls[size_t(i)] = (unsigned char)(type == TR_MAT_LAMBERT ? kShadeClassLambert
                                         : type == TR_MAT_DISNEY ? kShadeClassDisney : kShadeClassOther);
    }
    upload(c->tri_cls, cls.data(), cls.size(), st);
    upload(c->mats, dm.data(), dm.size() * sizeof(DevMaterial), st);
    upload(c->media, dmed.data(), dmed.size() * sizeof(DevMedium), st);
    if (!voxels.empty()) upload(c->density, voxels.data(), voxels.size() * sizeof(float), st);
    upload(c->lights, lights.data(), lights.size() * sizeof(float4), st);
    upload(c->light_cdf, cdf.data(), cdf.size() * sizeof(float), st);
    upload(c->light_pmf, pmf.data(), pmf.size() * sizeof(float), st);
    upload(c->light_guide, guide.data(), guide.size() * sizeof(int), st);
    // material textures: per-triangle uvs in leaf order, texel pools, a table
    bool any_tex = false;
    for (const tr_material& m : h.mats) any_tex = any_tex || uses_texture(m);
    if (!mat_tex.empty()) {
        upload(c->mat_tex, mat_tex.data(), mat_tex.size() * sizeof(int), st);
        upload(c->mat_raw, mat_raw.data(), mat_raw.size() * sizeof(float), st);
    }
    if (any_tex) {
        std::vector<float4> uv(size_t(nt) * 2);
        for (int64_t i = 0; i < nt; ++i) {
            const int prim = r.tris[size_t(i)].prim;
            float q[6];
            for (int k = 0; k < 3; ++k) {
                const int v = h.idx[3 * size_t(prim) + k];
                q[2 * k] = h.UV[2 * size_t(v)];
                q[2 * k + 1] = h.UV[2 * size_t(v) + 1];
            }
            uv[2 * size_t(i)] = make_float4(q[0], q[1], q[2], q[3]);
            uv[2 * size_t(i) + 1] = make_float4(q[4], q[5], 0.0f, 0.0f);
        }
        std::vector<int4> info;
        std::vector<uchar4> p8;
        std::vector<float4> pf;
        for (const HostTexture& t : h.tex) {
            const size_t n = size_t(t.w) * size_t(t.h);
            const bool is8 = !t.rgb8.empty();
            if ((is8 ? p8.size() : pf.size()) + n > size_t(INT32_MAX)) throw InvalidArg("textures too large");
            info.push_back(make_int4(t.w, t.h, is8 ? 1 : 0, int(is8 ? p8.size() : pf.size())));
            for (size_t k = 0; k < n; ++k) {
                if (is8) p8.push_back(make_uchar4(t.rgb8[3 * k], t.rgb8[3 * k + 1], t.rgb8[3 * k + 2], 0));
                else pf.push_back(make_float4(t.rgbf[3 * k], t.rgbf[3 * k + 1], t.rgbf[3 * k + 2], 0.0f));
            }
        }
        if (p8.empty()) p8.push_back(make_uchar4(0, 0, 0, 0));
        if (pf.empty()) pf.push_back(make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        upload(c->tri_uv, uv.data(), uv.size() * sizeof(float4), st);
        upload(c->tex_info, info.data(), info.size() * sizeof(int4), st);
        upload(c->tex8, p8.data(), p8.size() * sizeof(uchar4), st);
        upload(c->texf, pf.data(), pf.size() * sizeof(float4), st);
    }
    // normal derivatives for the tangent frames of anisotropic materials and
    // normal maps (the only users of the frame's orientation about the normal)
    bool any_oriented = false;
    for (const tr_material& m : h.mats)
        any_oriented = any_oriented || (m.type == TR_MAT_DISNEY &&
                                        (m.anisotropic != 0.0f || m.param_texture[TR_PARAM_ANISOTROPIC] > 0 ||
                                         m.normal_texture > 0));
    const bool want_tan = any_oriented && !h.UV.empty() && !h.N.empty();
    if (want_tan) {
        std::vector<float4> dv(static_cast<size_t>(nt));
        for (int64_t i = 0; i < nt; ++i) {
            const int prim = r.tris[size_t(i)].prim;
            // the shading record's vertex normals: normalised as Triangle::intersect does
            f3 n[3];
            float u[3], v[3];
            for (int k = 0; k < 3; ++k) {
                const int vi = h.idx[3 * size_t(prim) + k];
                n[k] = xyz(shade[kShadeRecord * size_t(i) + k]);
                u[k] = h.UV[2 * size_t(vi)];
                v[k] = h.UV[2 * size_t(vi) + 1];
            }
            dv[size_t(i)] = normal_dv(n[1] - n[0], n[2] - n[0], u[1] - u[0], v[1] - v[0], u[2] - u[0], v[2] - v[0]);
        }
        upload(c->tri_tan, dv.data(), dv.size() * sizeof(float4), st);
    }
    const EnvMapHost& em = h.envmap;
    DevEnvMap env{};
    if (h.has_env && em.w > 0) {
        std::vector<float4> tex(size_t(em.w) * size_t(em.h));
        for (size_t i = 0; i < tex.size(); ++i) tex[i] = make_float4(em.rgb[3 * i], em.rgb[3 * i + 1], em.rgb[3 * i + 2], 0.0f);
        upload(c->env_tex, tex.data(), tex.size() * sizeof(float4), st);
        upload(c->env_cond_cdf, em.cond_cdf.data(), em.cond_cdf.size() * sizeof(float), st);
        upload(c->env_cond_func, em.cond_func.data(), em.cond_func.size() * sizeof(float), st);
        upload(c->env_marg_cdf, em.marg_cdf.data(), em.marg_cdf.size() * sizeof(float), st);
        upload(c->env_marg_func, em.marg_func.data(), em.marg_func.size() * sizeof(float), st);
        env.tex = c->env_tex.as<float4>();
        env.cond_cdf = c->env_cond_cdf.as<float>();
        env.cond_func = c->env_cond_func.as<float>();
        env.marg_cdf = c->env_marg_cdf.as<float>();
        env.marg_func = c->env_marg_func.as<float>();
        env.w = em.w;
        env.h = em.h;
        env.nu = em.nu;
        env.nv = em.nv;
        env.marg_int = em.marg_int;
        env.r0 = mk(em.rot[0], em.rot[1], em.rot[2]);
        env.r1 = mk(em.rot[3], em.rot[4], em.rot[5]);
        env.r2 = mk(em.rot[6], em.rot[7], em.rot[8]);
    }
    hip_check(hipStreamSynchronize(st), "scene upload");

    // traversal spill area: (max_stack - LDS entries) per thread of the largest grid
    const int threads = trace_grid_threads();
    const int extra = std::max(1, c->max_stack - kLdsStack);
    c->ovf_node.reserve(size_t(extra) * threads * sizeof(int));
    c->ovf_t.reserve(size_t(extra) * threads * sizeof(float));
    c->ovf_threads = threads;

    DevScene& s = c->ds;
    s.tr.nodes = c->nodes.as<float4>();
    s.tr.tris = c->tris.as<float4>();
    s.tr.max_stack = c->max_stack;
    s.tr.leaf_share = trace_leaf_share();
    s.tr.leaf_share_any = trace_leaf_share_any();
    s.tri_shade = c->tri_shade.as<float4>();
    s.tri_cls = c->tri_cls.as<unsigned char>();
    s.mats = c->mats.as<DevMaterial>();
    s.media = c->media.as<DevMedium>();
    s.lights = c->lights.as<float4>();
    s.light_cdf = c->light_cdf.as<float>();
    s.light_pmf = c->light_pmf.as<float>();
    s.light_guide = c->light_guide.as<int>();
    s.light_guide_size = int(guide.size());
    s.num_area_lights = nA;
    s.num_light_entries = tot > 0.0 ? nA + 2 : 0;
    s.has_env = h.has_env;
    s.has_sun = h.has_sun;
    s.env_L = mk(h.env[0], h.env[1], h.env[2]);
    s.env = env;
    s.tri_uv = any_tex ? c->tri_uv.as<float4>() : nullptr;
    s.mat_tex = mat_tex.empty() ? nullptr : c->mat_tex.as<int>();
    s.mat_raw = mat_raw.empty() ? nullptr : c->mat_raw.as<float>();
    s.tri_tan = want_tan ? c->tri_tan.as<float4>() : nullptr;
    s.tex_info = any_tex ? c->tex_info.as<int4>() : nullptr;
    s.tex8 = any_tex ? c->tex8.as<uchar4>() : nullptr;
    s.texf = any_tex ? c->texf.as<float4>() : nullptr;
    s.has_tex = any_tex ? 1 : 0;
    s.density = voxels.empty() ? nullptr : c->density.as<float>();
    s.has_grid = voxels.empty() ? 0 : 1;
    const f3 sd = mk(h.sun_dir[0], h.sun_dir[1], h.sun_dir[2]);
    s.sun_dir = length(sd) > 0.0f ? normalize(sd) : mk(0.0f, 0.0f, 1.0f);
    s.sun_E = mk(h.sun_E[0], h.sun_E[1], h.sun_E[2]);
    s.wcenter = wc;
    s.wradius = wr;
    s.cam = device_camera(h.cam);
    c->have_bvh = true;
}

// Path slots for P paths and shadow slots for nls light samples each.
void ensure_paths(tr_ctx* c, int64_t P, int nls = 1) {
    const size_t ns = size_t(P) * size_t(nls);
    c->sh_o.reserve(ns * 16); c->sh_d.reserve(ns * 16); c->sh_c.reserve(ns * 16);
    if (P <= c->path_capacity) return;
    const size_t n = size_t(P);
    c->ray_o.reserve(n * 16); c->ray_d.reserve(n * 16); c->hit.reserve(n * 16);
    c->thr.reserve(n * 16); c->rad.reserve(n * 16); c->rng.reserve(n * 8); c->info.reserve(n * 8);
    c->q0.reserve(n * 4); c->q1.reserve(n * 4);
    c->path_capacity = P;
}

// Samples per pixel in flight when the caller leaves it to the library. The
// late bounces of a pass trace only the few paths still alive, and a launch of
// a few thousand rays costs about as much as one of a hundred thousand, so the
// pass holds as many paths as memory allows up to 2^26 (a 1024x1024 frame of
// 64 spp in one pass). Path slots may use a quarter of the free device memory
// plus what this context already holds.
int auto_samples_per_pass(tr_ctx* c, int64_t npix, bool guided, int nls = 1) {
    const int64_t slot_bytes = 5 * 16 + 2 * 8 + 2 * 4 + 3 * 16 * nls + (guided ? kGuideOut * 4 + 2 * 4 : 0);
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const int64_t held = guided ? std::min(c->path_capacity, c->guide.path_capacity) : c->path_capacity;
    const int64_t fit = held + int64_t(free_b / 4) / slot_bytes;
    const int64_t target = std::min<int64_t>(int64_t(1) << 26, fit);
    return int(std::max<int64_t>(1, target / npix));
}

DevPaths make_paths(tr_ctx* c) {
    DevPaths p{};
    p.ray_o = c->ray_o.as<float4>(); p.ray_d = c->ray_d.as<float4>(); p.hit = c->hit.as<float4>();
    p.thr = c->thr.as<float4>(); p.rad = c->rad.as<float4>(); p.rng = c->rng.as<uint2>(); p.info = c->info.as<uint2>();
    p.queue[0] = c->q0.as<int>(); p.queue[1] = c->q1.as<int>();
    p.sh_o = c->sh_o.as<float4>(); p.sh_d = c->sh_d.as<float4>(); p.sh_c = c->sh_c.as<float4>();
    p.counters = c->counters.as<unsigned int>();
    p.stats = c->stats.as<unsigned long long>();
    p.ovf_node = c->ovf_node.as<int>();
    p.ovf_t = c->ovf_t.as<float>();
    p.ovf_threads = c->ovf_threads;
    return p;
}

// Film filter of a render: box or GaussianFilter (src/filter/gaussian_filter.cpp)
// of any radius up to kMaxFilterRadius pixels (the gather of k_accumulate
// visits (2 floor(r + 1/2) + 1)^2 neighbours per sample).
constexpr float kMaxFilterRadius = 16.0f;
void check_filter(int type, float radius, float alpha) {
    if (type != TR_FILTER_BOX && type != TR_FILTER_GAUSSIAN) throw InvalidArg("unknown filter");
    if (!(radius > 0.0f && radius <= kMaxFilterRadius)) throw InvalidArg("filter radius must lie in (0, 16] pixels");
    if (type == TR_FILTER_GAUSSIAN && !(alpha >= 0.0f && alpha < 1e30f))
        throw InvalidArg("Gaussian filter alpha must be finite and >= 0");
}
void check_filter(const tr_render_params& P) { check_filter(P.filter_type, P.filter_radius, P.filter_alpha); }

// Learned light-sample distribution (guide_light_grid): bins per cell aim
// at about kLcBinsTarget (S^2 strata per emissive triangle, S in [1, 16], plus
// the environment and the sun); eps is the share of the power-based choice.
// On configuration 4 (1024^2, 256 spp, 16^3 cells) S = 16 beat 24 and 32
// (1.91 and 1.84 against 2.22 equal-time relMSE gain: finer strata get too
// few training samples each), and eps 0.1 edged out 0.2 and 0.05.
constexpr int kLcBinsTarget = 1024;
constexpr float kLcEps = 0.1f;

// restart: a render that trains from sample 0 begins a new image, and the
// distribution starts again from the power-based choice.
void lc_prepare(tr_ctx* c, int G, int64_t max_paths, bool training, bool restart) {
    LightCacheState& L = c->lcache;
    const int nA = c->ds.num_area_lights;
    if (L.grid != G || L.num_area != nA || restart) {
        int S = nA > 0 ? int(std::floor(std::sqrt(double(kLcBinsTarget) / double(nA)))) : 1;
        S = std::min(std::max(S, 1), 16);
        L.grid = 0;
        L.built = false;
        L.strata = S;
        L.num_area = nA;
        L.bins = nA * S * S + 2;
        const size_t n = size_t(G) * G * G * size_t(L.bins);
        L.sum.reserve(n * 8);
        L.hits.reserve(size_t(G) * G * G * 4);
        hip_check(hipMemsetAsync(L.hits.ptr, 0, size_t(G) * G * G * 4, c->stream), "light cache clear");
        L.pmf.reserve(n * 4);
        L.table.reserve(n * 8);
        L.work.reserve(n * 4);
        hip_check(hipMemsetAsync(L.sum.ptr, 0, n * 8, c->stream), "light cache clear");
        for (int k = 0; k < 3; ++k) {
            const float ext = c->hs.hi[k] - c->hs.lo[k];
            L.lo[k] = c->hs.lo[k];
            L.scale[k] = ext > 0.0f ? float(G) / ext : 0.0f;
        }
        L.grid = G;
    }
    if (training && max_paths > L.sh_cap) {
        L.sh_g.reserve(size_t(max_paths) * 16);   // one shadow ray per path and bounce at most
        L.sh_cap = max_paths;
    }
}

DevLightCache lc_device(tr_ctx* c, bool train, bool use) {
    const LightCacheState& L = c->lcache;
    DevLightCache d{};
    d.grid = L.grid;
    d.strata = L.strata;
    d.bins = L.bins;
    d.num_area = L.num_area;
    d.lo = mk(L.lo[0], L.lo[1], L.lo[2]);
    d.scale = mk(L.scale[0], L.scale[1], L.scale[2]);
    d.pmf = L.pmf.as<float>();
    d.table = L.table.as<float2>();
    d.sum = L.sum.as<unsigned long long>();
    d.hits = L.hits.as<unsigned int>();
    d.train = train ? 1 : 0;
    d.use = use ? 1 : 0;
    return d;
}

void lc_build(tr_ctx* c) {
    LightCacheState& L = c->lcache;
    const int cells = L.grid * L.grid * L.grid;
    launch_lc_build(lc_device(c, false, false), c->light_pmf.as<float>(), kLcEps, cells, L.pmf.as<float>(),
                    L.table.as<float2>(), L.work.as<int>(), c->stream);
    hip_check(hipGetLastError(), "k_lc_build launch");
    L.built = true;
}

// Bidirectional path tracing: per camera sample 80-byte vertices of both
// subpaths, two subpath rays with their hits, weights and RNG states, and per
// strategy a contribution and (at most) one connection ray.
constexpr int kBdMaxVertices = 32;
int64_t bdpt_slot_bytes(int nc, int nl) {
    const int64_t K = int64_t(nc) * (nl + 1);
    return int64_t(nc + nl) * int64_t(sizeof(BdVertex)) + 2 * 4 + 2 * (32 + 16 + 16 + 8 + 2 * 4) + 16 + K * (16 + 32 + 4);
}
void bdpt_ensure(tr_ctx* c, int64_t slots, int nc, int nl, int64_t npix) {
    BdptState& B = c->bdpt;
    const int64_t K = int64_t(nc) * (nl + 1);
    const int64_t verts = slots * (nc + nl), entries = slots * K;
    if (slots > B.slots) {
        const size_t n = size_t(slots);
        B.ncv.reserve(n * 4); B.nlv.reserve(n * 4); B.rad.reserve(n * 16);
        B.gray.reserve(2 * n * 32); B.ghit.reserve(2 * n * 16); B.gbeta.reserve(2 * n * 16); B.grng.reserve(2 * n * 8);
        B.gq0.reserve(2 * n * 4); B.gq1.reserve(2 * n * 4);
        B.slots = slots;
    }
    B.cv.reserve(size_t(slots) * nc * sizeof(BdVertex));
    B.lv.reserve(size_t(slots) * std::max(nl, 1) * sizeof(BdVertex));
    B.vertices = std::max(B.vertices, verts);
    if (entries > B.entries) {
        B.con.reserve(size_t(entries) * 16);
        B.sray.reserve(size_t(entries) * 32);
        B.sidx.reserve(size_t(entries) * 4);
        B.entries = entries;
    }
    if (npix > B.pixels) {
        B.splat.reserve(size_t(npix) * 24);
        B.w0.reserve(size_t(npix) * 8);
        B.pixels = npix;
    }
}

hipEvent_t ev(tr_ctx* c, size_t i) {
    while (c->events.size() <= i) {
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        c->events.push_back(e);
    }
    return c->events[i];
}

}  // namespace

extern "C" {

const char* tr_version(void) { return "tracer_hip 0.2 (gfx950, wavefront path tracer, 4-wide BVH)"; }

tr_status tr_bvh4_export(const tr_scene_desc* d, const tr_bvh_opts* opts, void* nodes_out, int64_t nodes_cap,
                         void* tris_out, int64_t* num_nodes, int32_t* depth) {
    try {
        validate_scene(d);
        BvhBuildOptions bo;
        if (opts) {
            bo.max_leaf = opts->max_leaf_primitives > 0 ? opts->max_leaf_primitives : 4;
            bo.threads = opts->num_threads;
        }
        BvhBuildResult r = build_bvh4(d->positions, d->num_vertices, d->indices, d->num_triangles, bo);
        if (num_nodes) *num_nodes = int64_t(r.nodes.size());
        if (depth) *depth = r.depth;
        if (nodes_out) {
            if (nodes_cap < int64_t(r.nodes.size())) return TR_ERR_INVALID;
            std::memcpy(nodes_out, r.nodes.data(), r.nodes.size() * sizeof(Node4));
        }
        if (tris_out) std::memcpy(tris_out, r.tris.data(), r.tris.size() * sizeof(TriPacked));
        return TR_OK;
    } catch (const std::exception&) {
        return TR_ERR_INVALID;
    }
}

void tr_rng_sequence(uint64_t seed, uint32_t pixel, uint32_t sample, int64_t n, uint32_t* out) {
    Pcg r = pcg_for_sample(seed, pixel, sample);
    for (int64_t i = 0; i < n; ++i) out[i] = r.next();
}

tr_status tr_create(tr_ctx** out, int32_t device) {
    if (!out) return TR_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return TR_ERR_HIP;
    if (device < 0 || device >= count) return TR_ERR_INVALID;
    tr_ctx* c = new tr_ctx();
    c->device = device;
    tr_status s = guarded(c, [&] {
        hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(noalloc_scratch_init(), "hipGetSymbolAddress (allocation-free scratch)");
        c->counters.reserve(size_t(kNumCounters) * kCounterStride * sizeof(unsigned int));
        c->stats.reserve(128);
    });
    if (s != TR_OK) {
        delete c;
        return s;
    }
    *out = c;
    return TR_OK;
}

void tr_destroy(tr_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* tr_last_error(const tr_ctx* c) { return c ? c->err.c_str() : "null context"; }

tr_status tr_host_alloc(tr_ctx* c, int64_t bytes, void** ptr) {
    return guarded(c, [&] {
        if (!ptr) throw InvalidArg("ptr is NULL");
        if (bytes <= 0) throw InvalidArg("bytes must be positive");
        *ptr = nullptr;
        hip_check(hipHostMalloc(ptr, size_t(bytes), hipHostMallocDefault), "hipHostMalloc");
    });
}

tr_status tr_host_free(tr_ctx* c, void* ptr) {
    if (!c) return !ptr || hipHostFree(ptr) == hipSuccess ? TR_OK : TR_ERR_HIP;
    return guarded(c, [&] {
        if (ptr) hip_check(hipHostFree(ptr), "hipHostFree");
    });
}

tr_status tr_scene_upload(tr_ctx* c, const tr_scene_desc* d) {
    return guarded(c, [&] {
        validate_scene(d);
        HostScene& h = c->hs;
        const size_t nv = size_t(d->num_vertices), nt = size_t(d->num_triangles);
        h.P.assign(d->positions, d->positions + 3 * nv);
        if (d->normals) h.N.assign(d->normals, d->normals + 3 * nv); else h.N.clear();
        if (d->uvs) h.UV.assign(d->uvs, d->uvs + 2 * nv); else h.UV.clear();
        h.tex.clear();
        for (int i = 0; i < d->num_textures; ++i) {
            const tr_texture& t = d->textures[i];
            HostTexture ht;
            ht.w = t.width;
            ht.h = t.height;
            const size_t n = size_t(t.width) * size_t(t.height) * 3;
            if (t.rgb8) ht.rgb8.assign(t.rgb8, t.rgb8 + n);
            else ht.rgbf.assign(t.rgbf, t.rgbf + n);
            h.tex.push_back(std::move(ht));
        }
        h.idx.assign(d->indices, d->indices + 3 * nt);
        h.mat.assign(d->material_ids, d->material_ids + nt);
        if (d->emission) h.emission.assign(d->emission, d->emission + 3 * nt); else h.emission.clear();
        if (d->medium_inside) h.med_in.assign(d->medium_inside, d->medium_inside + nt); else h.med_in.clear();
        if (d->medium_outside) h.med_out.assign(d->medium_outside, d->medium_outside + nt); else h.med_out.clear();
        h.mats.assign(d->materials, d->materials + d->num_materials);
        if (d->num_media) h.media.assign(d->media, d->media + d->num_media); else h.media.clear();
        h.grids.assign(h.media.size(), {});
        h.albedo_grids.assign(h.media.size(), {});
        h.g_grids.assign(h.media.size(), {});
        for (size_t i = 0; i < h.media.size(); ++i) {
            tr_medium& m = h.media[i];
            if (m.density) {
                h.grids[i].assign(m.density, m.density + size_t(m.nx) * size_t(m.ny) * size_t(m.nz));
                m.density = nullptr;   // the copies here are the grids from here on
            }
            if (m.albedo_grid) {
                h.albedo_grids[i].assign(m.albedo_grid, m.albedo_grid + 3 * size_t(m.anx) * size_t(m.any) * size_t(m.anz));
                m.albedo_grid = nullptr;
            }
            if (m.g_grid) {
                h.g_grids[i].assign(m.g_grid, m.g_grid + size_t(m.gnx) * size_t(m.gny) * size_t(m.gnz));
                m.g_grid = nullptr;
            }
        }
        h.has_env = d->has_env;
        h.has_sun = d->has_sun;
        std::memcpy(h.env, d->env_radiance, sizeof h.env);
        h.envmap = EnvMapHost{};
        if (d->has_env && d->env_width > 0) {
            EnvMapHost& e = h.envmap;
            e.w = d->env_width;
            e.h = d->env_height;
            e.rgb.assign(d->env_pixels, d->env_pixels + size_t(e.w) * size_t(e.h) * 3);
            std::memcpy(e.rot, d->env_world_to_light, sizeof e.rot);
            build_envmap(e);
        }
        std::memcpy(h.sun_dir, d->sun_direction, sizeof h.sun_dir);
        std::memcpy(h.sun_E, d->sun_irradiance, sizeof h.sun_E);
        h.cam = d->camera;
        for (int k = 0; k < 3; ++k) { h.lo[k] = INFINITY; h.hi[k] = -INFINITY; }
        for (size_t t = 0; t < nt; ++t)
            for (int k = 0; k < 3; ++k) {
                const float* v = &h.P[3 * size_t(h.idx[3 * t + k])];
                for (int a = 0; a < 3; ++a) { h.lo[a] = std::min(h.lo[a], v[a]); h.hi[a] = std::max(h.hi[a], v[a]); }
            }
        c->have_scene = true;
        c->have_bvh = false;
    });
}

tr_status tr_set_camera(tr_ctx* c, const tr_camera* camera) {
    return guarded(c, [&] {
        if (!camera) throw InvalidArg("camera is NULL");
        if (!c->have_scene) throw StateError("tr_set_camera before tr_scene_upload");
        check_camera(*camera);
        c->hs.cam = *camera;
        if (c->have_bvh) c->ds.cam = device_camera(*camera);
    });
}

tr_status tr_bvh_build(tr_ctx* c, const tr_bvh_opts* opts) {
    return guarded(c, [&] {
        if (!c->have_scene) throw StateError("tr_bvh_build before tr_scene_upload");
        build_device_scene(c, opts);
    });
}

tr_status tr_bvh_info(tr_ctx* c, int64_t* num_nodes, int64_t* node_bytes, int64_t* tri_bytes) {
    return guarded(c, [&] {
        if (!c->have_bvh) throw StateError("no BVH");
        if (num_nodes) *num_nodes = c->num_nodes;
        if (node_bytes) *node_bytes = c->num_nodes * int64_t(sizeof(Node4));
        if (tri_bytes) *tri_bytes = int64_t(c->hs.mat.size()) * int64_t(sizeof(TriPacked));
    });
}

tr_status tr_stream(tr_ctx* c, void** s) {
    return guarded(c, [&] {
        if (!s) throw InvalidArg("null out pointer");
        *s = c->stream;
    });
}

tr_status tr_synchronize(tr_ctx* c) {
    return guarded(c, [&] { hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize"); });
}

tr_status tr_trace_closest_device(tr_ctx* c, const float* d_rays, float* d_hits, int64_t n) {
    return guarded(c, [&] {
        if (!c->have_bvh) throw StateError("tr_trace_closest before tr_bvh_build");
        if (n < 0 || (n > 0 && (!d_rays || !d_hits))) throw InvalidArg("bad ray batch");
        launch_trace_closest(c->ds.tr, reinterpret_cast<const float4*>(d_rays), reinterpret_cast<float4*>(d_hits), n,
                             c->ovf_node.as<int>(), c->ovf_t.as<float>(), c->ovf_threads, c->stream);
        hip_check(hipGetLastError(), "k_trace_closest launch");
    });
}

tr_status tr_trace_closest_device_noalloc(tr_ctx* c, const float* d_rays, float* d_hits, int64_t n) {
    return guarded(c, [&] {
        if (!c->have_bvh) throw StateError("tr_trace_closest_device_noalloc before tr_bvh_build");
        if (n < 0 || (n > 0 && (!d_rays || !d_hits))) throw InvalidArg("bad ray batch");
        if (c->max_stack > noalloc_max_stack())
            throw InvalidArg("the BVH needs " + std::to_string(c->max_stack) + " traversal stack entries; "
                             "tr_trace_closest_device_noalloc holds at most " + std::to_string(noalloc_max_stack()));
        try {
            launch_trace_closest_noalloc(c->ds.tr, reinterpret_cast<const float4*>(d_rays),
                                         reinterpret_cast<float4*>(d_hits), n, c->stream);
        } catch (const std::invalid_argument& e) {
            throw InvalidArg(e.what());
        }
        hip_check(hipGetLastError(), "k_trace_closest launch");
    });
}

tr_status tr_trace_any_device(tr_ctx* c, const float* d_rays, int32_t* d_occ, int64_t n) {
    return guarded(c, [&] {
        if (!c->have_bvh) throw StateError("tr_trace_any before tr_bvh_build");
        if (n < 0 || (n > 0 && (!d_rays || !d_occ))) throw InvalidArg("bad ray batch");
        launch_trace_any(c->ds.tr, reinterpret_cast<const float4*>(d_rays), d_occ, n, c->ovf_node.as<int>(),
                         c->ovf_t.as<float>(), c->ovf_threads, c->stream);
        hip_check(hipGetLastError(), "k_trace_any launch");
    });
}

tr_status tr_trace_closest(tr_ctx* c, const float* rays, float* hits, int64_t n) {
    return guarded(c, [&] {
        if (!c->have_bvh) throw StateError("tr_trace_closest before tr_bvh_build");
        if (n < 0 || (n > 0 && (!rays || !hits))) throw InvalidArg("bad ray batch");
        if (n == 0) return;
        upload(c->s_in, rays, size_t(n) * 32, c->stream);
        c->s_out.reserve(size_t(n) * 16);
        launch_trace_closest(c->ds.tr, c->s_in.as<float4>(), c->s_out.as<float4>(), n, c->ovf_node.as<int>(),
                             c->ovf_t.as<float>(), c->ovf_threads, c->stream);
        hip_check(hipGetLastError(), "k_trace_closest launch");
        hip_check(hipMemcpyAsync(hits, c->s_out.ptr, size_t(n) * 16, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

tr_status tr_trace_any(tr_ctx* c, const float* rays, int32_t* occ, int64_t n) {
    return guarded(c, [&] {
        if (!c->have_bvh) throw StateError("tr_trace_any before tr_bvh_build");
        if (n < 0 || (n > 0 && (!rays || !occ))) throw InvalidArg("bad ray batch");
        if (n == 0) return;
        upload(c->s_in, rays, size_t(n) * 32, c->stream);
        c->s_out.reserve(size_t(n) * 4);
        launch_trace_any(c->ds.tr, c->s_in.as<float4>(), c->s_out.as<int>(), n, c->ovf_node.as<int>(), c->ovf_t.as<float>(),
                         c->ovf_threads, c->stream);
        hip_check(hipGetLastError(), "k_trace_any launch");
        hip_check(hipMemcpyAsync(occ, c->s_out.ptr, size_t(n) * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

tr_status tr_trace_closest_counted(tr_ctx* c, const float* rays, float* hits, uint32_t* nvis, uint32_t* ntri, int64_t n) {
    return guarded(c, [&] {
        if (!c->have_bvh) throw StateError("tr_trace_closest_counted before tr_bvh_build");
        if (n < 0 || (n > 0 && (!rays || !hits || !nvis || !ntri))) throw InvalidArg("bad ray batch");
        if (n == 0) return;
        upload(c->s_in, rays, size_t(n) * 32, c->stream);
        c->s_out.reserve(size_t(n) * 16);
        c->s_aux.reserve(size_t(n) * 4);
        c->s_aux2.reserve(size_t(n) * 4);
        launch_trace_counted(c->ds.tr, c->s_in.as<float4>(), c->s_out.as<float4>(), c->s_aux.as<uint32_t>(),
                             c->s_aux2.as<uint32_t>(), n, c->ovf_node.as<int>(), c->ovf_t.as<float>(), c->ovf_threads,
                             c->stream);
        hip_check(hipGetLastError(), "k_trace_counted launch");
        hip_check(hipMemcpyAsync(hits, c->s_out.ptr, size_t(n) * 16, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipMemcpyAsync(nvis, c->s_aux.ptr, size_t(n) * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipMemcpyAsync(ntri, c->s_aux2.ptr, size_t(n) * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

static void bsdf_common(tr_ctx* c, const tr_material* mats, int32_t num_mats, const int32_t* ids, int64_t n) {
    if (!mats || num_mats <= 0 || !ids || n < 0) throw InvalidArg("bad BSDF batch");
    std::vector<DevMaterial> dm(static_cast<size_t>(num_mats));
    for (int i = 0; i < num_mats; ++i) {
        check_material(mats[i]);
        dm[size_t(i)] = make_material(mats[i]);
    }
    for (int64_t i = 0; i < n; ++i)
        if (ids[i] < 0 || ids[i] >= num_mats) throw InvalidArg("material id out of range");
    upload(c->s_aux3, dm.data(), dm.size() * sizeof(DevMaterial), c->stream);
    upload(c->s_aux4, ids, size_t(n) * 4, c->stream);
}

tr_status tr_bsdf_eval_batch(tr_ctx* c, const tr_material* mats, int32_t num_mats, const int32_t* ids, const float* wi,
                             const float* wo, const float* ng, int64_t n, float* f, float* pdf) {
    return guarded(c, [&] {
        if (!wi || !wo || !f || !pdf) throw InvalidArg("bad BSDF batch");
        bsdf_common(c, mats, num_mats, ids, n);
        if (n == 0) return;
        const size_t v3 = size_t(n) * 12;
        c->s_in.reserve(3 * v3);
        char* in = c->s_in.as<char>();
        hip_check(hipMemcpyAsync(in, wi, v3, hipMemcpyHostToDevice, c->stream), "H2D");
        hip_check(hipMemcpyAsync(in + v3, wo, v3, hipMemcpyHostToDevice, c->stream), "H2D");
        if (ng) hip_check(hipMemcpyAsync(in + 2 * v3, ng, v3, hipMemcpyHostToDevice, c->stream), "H2D");
        c->s_out.reserve(v3 + size_t(n) * 4);
        char* out = c->s_out.as<char>();
        launch_bsdf_eval(c->s_aux3.as<DevMaterial>(), c->s_aux4.as<int>(), reinterpret_cast<float*>(in),
                         reinterpret_cast<float*>(in + v3), ng ? reinterpret_cast<float*>(in + 2 * v3) : nullptr, n,
                         reinterpret_cast<float*>(out), reinterpret_cast<float*>(out + v3), c->stream);
        hip_check(hipGetLastError(), "k_bsdf_eval launch");
        hip_check(hipMemcpyAsync(f, out, v3, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipMemcpyAsync(pdf, out + v3, size_t(n) * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

tr_status tr_film_to_ldr(tr_ctx* c, float aces_exposure, float inv_gamma, uint8_t* rgb, int64_t cap) {
    return guarded(c, [&] {
        if (c->film_w <= 0) throw StateError("tr_film_to_ldr before tr_render");
        const int64_t npix = int64_t(c->film_w) * c->film_h;
        if (!rgb || cap < npix * 3) throw InvalidArg("rgb is NULL or smaller than width * height * 3 bytes");
        c->s_out.reserve(size_t(npix) * 3);
        launch_film_ldr(c->film.as<double4>(), c->s_out.as<unsigned char>(), int(npix), aces_exposure, inv_gamma,
                        c->stream);
        hip_check(hipGetLastError(), "k_film_ldr launch");
        hip_check(hipMemcpyAsync(rgb, c->s_out.ptr, size_t(npix) * 3, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

static void env_common(tr_ctx* c, const void* a, const void* b, const void* d, int64_t n) {
    if (!c->have_bvh) throw StateError("environment batch before tr_bvh_build");
    if (!c->hs.has_env) throw StateError("the scene has no environment light");
    if (n < 0 || n > (int64_t(1) << 26)) throw InvalidArg("n must lie in [0, 2^26]");
    if (!a || !b || !d) throw InvalidArg("NULL array");
}

tr_status tr_env_sample_batch(tr_ctx* c, const float* u, int64_t n, float* wi, float* pdf, float* L) {
    return guarded(c, [&] {
        env_common(c, u, wi, pdf, n);
        if (!L) throw InvalidArg("NULL array");
        if (n == 0) return;
        const size_t v3 = size_t(n) * 12;
        upload(c->s_in, u, size_t(n) * 8, c->stream);
        c->s_out.reserve(2 * v3 + size_t(n) * 4);
        char* out = c->s_out.as<char>();
        launch_env_sample(c->ds, c->s_in.as<float>(), n, reinterpret_cast<float*>(out),
                          reinterpret_cast<float*>(out + 2 * v3), reinterpret_cast<float*>(out + v3), c->stream);
        hip_check(hipGetLastError(), "k_env_sample launch");
        hip_check(hipMemcpyAsync(wi, out, v3, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipMemcpyAsync(L, out + v3, v3, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipMemcpyAsync(pdf, out + 2 * v3, size_t(n) * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

tr_status tr_env_eval_batch(tr_ctx* c, const float* dirs, int64_t n, float* L, float* pdf) {
    return guarded(c, [&] {
        env_common(c, dirs, L, pdf, n);
        if (n == 0) return;
        const size_t v3 = size_t(n) * 12;
        upload(c->s_in, dirs, v3, c->stream);
        c->s_out.reserve(v3 + size_t(n) * 4);
        char* out = c->s_out.as<char>();
        launch_env_eval(c->ds, c->s_in.as<float>(), n, reinterpret_cast<float*>(out), reinterpret_cast<float*>(out + v3),
                        c->stream);
        hip_check(hipGetLastError(), "k_env_eval launch");
        hip_check(hipMemcpyAsync(L, out, v3, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipMemcpyAsync(pdf, out + v3, size_t(n) * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

tr_status tr_medium_density_batch(tr_ctx* c, int32_t medium, const float* pos, int64_t n, float* density) {
    return guarded(c, [&] {
        if (!c->have_bvh) throw StateError("medium batch before tr_bvh_build");
        if (medium < 0 || size_t(medium) >= c->hs.media.size() || c->hs.grids[size_t(medium)].empty())
            throw InvalidArg("medium is not a density grid of the scene");
        if (n < 0 || n > (int64_t(1) << 26)) throw InvalidArg("n must lie in [0, 2^26]");
        if (!pos || !density) throw InvalidArg("NULL array");
        if (n == 0) return;
        upload(c->s_in, pos, size_t(n) * 12, c->stream);
        c->s_out.reserve(size_t(n) * 4);
        launch_grid_density(c->ds, medium, c->s_in.as<float>(), n, c->s_out.as<float>(), c->stream);
        hip_check(hipGetLastError(), "k_grid_density launch");
        hip_check(hipMemcpyAsync(density, c->s_out.as<float>(), size_t(n) * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

tr_status tr_medium_albedo_g_batch(tr_ctx* c, int32_t medium, const float* pos, int64_t n, float* albedo, float* g) {
    return guarded(c, [&] {
        if (!c->have_bvh) throw StateError("medium batch before tr_bvh_build");
        if (medium < 0 || size_t(medium) >= c->hs.media.size() || c->hs.grids[size_t(medium)].empty())
            throw InvalidArg("medium is not a density grid of the scene");
        if (n < 0 || n > (int64_t(1) << 26)) throw InvalidArg("n must lie in [0, 2^26]");
        if (!pos || !albedo || !g) throw InvalidArg("NULL array");
        if (n == 0) return;
        upload(c->s_in, pos, size_t(n) * 12, c->stream);
        c->s_out.reserve(size_t(n) * 16);
        float* out = c->s_out.as<float>();
        launch_grid_albedo_g(c->ds, medium, c->s_in.as<float>(), n, out, out + 3 * n, c->stream);
        hip_check(hipGetLastError(), "k_grid_albedo_g launch");
        hip_check(hipMemcpyAsync(albedo, out, size_t(n) * 12, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipMemcpyAsync(g, out + 3 * n, size_t(n) * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

tr_status tr_texture_eval_batch(tr_ctx* c, int32_t texture, const float* uv, int64_t n, float* rgb) {
    return guarded(c, [&] {
        if (!c->have_bvh) throw StateError("texture batch before tr_bvh_build");
        if (texture < 1 || size_t(texture) > c->hs.tex.size()) throw InvalidArg("texture id out of range");
        bool used = false;   // texture tables are uploaded only for scenes whose materials name one
        for (const tr_material& m : c->hs.mats) used = used || uses_texture(m);
        if (!used) throw StateError("no material of the scene names a texture");
        if (n < 0 || n > (int64_t(1) << 26)) throw InvalidArg("n must lie in [0, 2^26]");
        if (!uv || !rgb) throw InvalidArg("NULL array");
        if (n == 0) return;
        upload(c->s_in, uv, size_t(n) * 8, c->stream);
        c->s_out.reserve(size_t(n) * 12);
        launch_tex_eval(c->ds, texture, c->s_in.as<float>(), n, c->s_out.as<float>(), c->stream);
        hip_check(hipGetLastError(), "k_tex_eval launch");
        hip_check(hipMemcpyAsync(rgb, c->s_out.as<float>(), size_t(n) * 12, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

tr_status tr_bsdf_sample_batch(tr_ctx* c, const tr_material* mats, int32_t num_mats, const int32_t* ids, const float* wo,
                               const float* ng, const float* u, int64_t n, float* wi, float* f, float* pdf) {
    return guarded(c, [&] {
        if (!wo || !u || !wi || !f || !pdf) throw InvalidArg("bad BSDF batch");
        bsdf_common(c, mats, num_mats, ids, n);
        if (n == 0) return;
        const size_t v3 = size_t(n) * 12;
        c->s_in.reserve(3 * v3);
        char* in = c->s_in.as<char>();
        hip_check(hipMemcpyAsync(in, wo, v3, hipMemcpyHostToDevice, c->stream), "H2D");
        hip_check(hipMemcpyAsync(in + v3, u, v3, hipMemcpyHostToDevice, c->stream), "H2D");
        if (ng) hip_check(hipMemcpyAsync(in + 2 * v3, ng, v3, hipMemcpyHostToDevice, c->stream), "H2D");
        c->s_out.reserve(2 * v3 + size_t(n) * 4);
        char* out = c->s_out.as<char>();
        launch_bsdf_sample(c->s_aux3.as<DevMaterial>(), c->s_aux4.as<int>(), reinterpret_cast<float*>(in),
                           ng ? reinterpret_cast<float*>(in + 2 * v3) : nullptr, reinterpret_cast<float*>(in + v3), n,
                           reinterpret_cast<float*>(out), reinterpret_cast<float*>(out + v3),
                           reinterpret_cast<float*>(out + 2 * v3), c->stream);
        hip_check(hipGetLastError(), "k_bsdf_sample launch");
        hip_check(hipMemcpyAsync(wi, out, v3, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipMemcpyAsync(f, out + v3, v3, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipMemcpyAsync(pdf, out + 2 * v3, size_t(n) * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

tr_status tr_camera_ray_batch(tr_ctx* c, const float* film_uv, const float* lens_uv, int64_t n, float* rays) {
    return guarded(c, [&] {
        if (!c->have_bvh) throw StateError("camera batch before tr_bvh_build");
        if (n < 0 || n > (int64_t(1) << 26)) throw InvalidArg("n must lie in [0, 2^26]");
        if (!film_uv || !lens_uv || !rays) throw InvalidArg("NULL array");
        if (n == 0) return;
        c->s_in.reserve(size_t(n) * 16);
        float* in = c->s_in.as<float>();
        hip_check(hipMemcpyAsync(in, film_uv, size_t(n) * 8, hipMemcpyHostToDevice, c->stream), "H2D");
        hip_check(hipMemcpyAsync(in + 2 * n, lens_uv, size_t(n) * 8, hipMemcpyHostToDevice, c->stream), "H2D");
        c->s_out.reserve(size_t(n) * 32);
        launch_camera_rays(c->ds, in, in + 2 * n, n, c->s_out.as<float4>(), c->stream);
        hip_check(hipGetLastError(), "k_camera_rays launch");
        hip_check(hipMemcpyAsync(rays, c->s_out.ptr, size_t(n) * 32, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

tr_status tr_shading_frame_batch(tr_ctx* c, const int32_t* prims, const float* bary, int64_t n, float* frames) {
    return guarded(c, [&] {
        if (!c->have_bvh) throw StateError("shading-frame batch before tr_bvh_build");
        if (n < 0 || n > (int64_t(1) << 26)) throw InvalidArg("n must lie in [0, 2^26]");
        if (!prims || !bary || !frames) throw InvalidArg("NULL array");
        if (n == 0) return;
        std::vector<int32_t> leaf(static_cast<size_t>(n));
        const int64_t nt = int64_t(c->leaf_of.size());
        for (int64_t i = 0; i < n; ++i) {
            if (prims[i] < 0 || prims[i] >= nt) throw InvalidArg("primitive index out of range");
            leaf[size_t(i)] = c->leaf_of[size_t(prims[i])];
        }
        c->s_in.reserve(size_t(n) * 12);
        int* dtri = c->s_in.as<int>();
        float* db = reinterpret_cast<float*>(dtri + n);
        hip_check(hipMemcpyAsync(dtri, leaf.data(), size_t(n) * 4, hipMemcpyHostToDevice, c->stream), "H2D");
        hip_check(hipMemcpyAsync(db, bary, size_t(n) * 8, hipMemcpyHostToDevice, c->stream), "H2D");
        c->s_out.reserve(size_t(n) * 36);
        launch_shading_frames(c->ds, dtri, db, n, c->s_out.as<float>(), c->stream);
        hip_check(hipGetLastError(), "k_shading_frames launch");
        hip_check(hipMemcpyAsync(frames, c->s_out.ptr, size_t(n) * 36, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

tr_status tr_light_pmf(tr_ctx* c, float* pmf, int64_t cap, int64_t* n) {
    return guarded(c, [&] {
        if (!c->have_bvh) throw StateError("tr_light_pmf before tr_bvh_build");
        if (!n) throw InvalidArg("n is NULL");
        const int nA = c->ds.num_area_lights;
        *n = nA + 2;
        if (!pmf) return;
        if (cap < nA + 2) throw InvalidArg("pmf holds fewer than num_area_lights + 2 entries");
        hip_check(hipMemcpy(pmf, c->light_pmf.ptr, size_t(nA + 2) * sizeof(float), hipMemcpyDeviceToHost), "D2H");
    });
}

tr_status tr_light_pick_batch(tr_ctx* c, const float* u, int64_t n, int32_t* idx, float* pmf) {
    return guarded(c, [&] {
        if (!c->have_bvh) throw StateError("light batch before tr_bvh_build");
        if (n < 0 || n > (int64_t(1) << 26)) throw InvalidArg("n must lie in [0, 2^26]");
        if (!u || !idx || !pmf) throw InvalidArg("NULL array");
        if (c->ds.num_light_entries == 0) throw StateError("the scene has no light");
        if (n == 0) return;
        for (int64_t i = 0; i < n; ++i)
            if (!(u[i] >= 0.0f && u[i] < 1.0f)) throw InvalidArg("uniforms must lie in [0, 1)");
        upload(c->s_in, u, size_t(n) * 4, c->stream);
        c->s_out.reserve(size_t(n) * 8);
        int* di = c->s_out.as<int>();
        float* dp = reinterpret_cast<float*>(di + n);
        launch_light_pick(c->ds, c->s_in.as<float>(), n, di, dp, c->stream);
        hip_check(hipGetLastError(), "k_light_pick launch");
        hip_check(hipMemcpyAsync(idx, di, size_t(n) * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipMemcpyAsync(pmf, dp, size_t(n) * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

tr_status tr_hg_sample_batch(tr_ctx* c, float g, const float* wo, const float* u, int64_t n, float* wi, float* pdf) {
    return guarded(c, [&] {
        if (!(g > -1.0f && g < 1.0f)) throw InvalidArg("g must lie in (-1, 1)");
        if (n < 0 || n > (int64_t(1) << 26)) throw InvalidArg("n must lie in [0, 2^26]");
        if (!wo || !u || !wi || !pdf) throw InvalidArg("NULL array");
        if (n == 0) return;
        c->s_in.reserve(size_t(n) * 20);
        float* in = c->s_in.as<float>();
        hip_check(hipMemcpyAsync(in, wo, size_t(n) * 12, hipMemcpyHostToDevice, c->stream), "H2D");
        hip_check(hipMemcpyAsync(in + 3 * n, u, size_t(n) * 8, hipMemcpyHostToDevice, c->stream), "H2D");
        c->s_out.reserve(size_t(n) * 16);
        float* out = c->s_out.as<float>();
        launch_hg_sample(g, in, in + 3 * n, n, out, out + 3 * n, c->stream);
        hip_check(hipGetLastError(), "k_hg_sample launch");
        hip_check(hipMemcpyAsync(wi, out, size_t(n) * 12, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipMemcpyAsync(pdf, out + 3 * n, size_t(n) * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

tr_status tr_film_accumulate_batch(tr_ctx* c, const tr_render_params* prm, const float* radiance, double* film_sums) {
    return guarded(c, [&] {
        if (!prm || !radiance || !film_sums) throw InvalidArg("NULL argument");
        const tr_render_params& P = *prm;
        if (P.width <= 0 || P.height <= 0 || P.spp <= 0) throw InvalidArg("width, height and spp must be positive");
        check_filter(P);
        const int64_t npix = int64_t(P.width) * P.height;
        const int64_t n = npix * P.spp;
        if (n > (int64_t(1) << 26)) throw InvalidArg("width * height * spp must be at most 2^26");
        ensure_paths(c, n);
        upload(c->s_in, radiance, size_t(n) * 12, c->stream);
        DevPaths p = make_paths(c);
        launch_unpack_radiance(c->s_in.as<float>(), n, p.rad, c->stream);
        c->film.reserve(size_t(npix) * sizeof(double4));
        hip_check(hipMemsetAsync(c->film.ptr, 0, size_t(npix) * 32, c->stream), "film clear");
        hip_check(hipMemsetAsync(c->stats.ptr, 0, 128, c->stream), "stats clear");
        RenderConsts rc{};
        rc.width = P.width; rc.height = P.height;
        rc.filter = P.filter_type; rc.filter_radius = P.filter_radius; rc.filter_alpha = P.filter_alpha;
        rc.seed = P.seed;
        rc.spp_in_pass = P.spp;
        rc.sample_base = P.sample_offset;
        launch_accumulate(p, rc, c->film.as<double4>(), c->stream);
        hip_check(hipGetLastError(), "k_accumulate launch");
        hip_check(hipMemcpyAsync(film_sums, c->film.ptr, size_t(npix) * 32, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        c->film_w = P.width;
        c->film_h = P.height;
    });
}

tr_status tr_render(tr_ctx* c, const tr_render_params* prm, double* film_sums, int32_t accumulate, float* image,
                    tr_render_stats* stats) {
    return guarded(c, [&] {
        if (!c->have_bvh) throw StateError("tr_render before tr_bvh_build");
        if (!prm) throw InvalidArg("params is NULL");
        const tr_render_params& P = *prm;
        if (P.width <= 0 || P.height <= 0 || P.spp <= 0) throw InvalidArg("width, height and spp must be positive");
        if (P.max_depth < 0 || P.min_depth < 0) throw InvalidArg("depths must be >= 0");
        if (P.direct_light_sample_num < 0 || P.direct_light_sample_num > 64)
            throw InvalidArg("direct_light_sample_num must lie in [0, 64] (0 = 1)");
        const int nls = std::max(1, P.direct_light_sample_num);
        if (nls > 1 && P.guide != 0)
            throw InvalidArg("direct_light_sample_num > 1 is supported by the plain path tracer (guide = 0)");
        check_filter(P);
        if (P.guide < 0 || P.guide > 2) throw InvalidArg("guide must be 0, 1 or 2");
        if (P.guide == 2 && P.guide_light_grid <= 0)
            throw InvalidArg("guide = 2 (learned light sampling alone) needs guide_light_grid > 0");
        if (P.guide && !(P.guide_fraction >= 0.0f && P.guide_fraction < 1.0f))
            throw InvalidArg("guide_fraction must lie in [0, 1) (0 = the default 0.5)");
        if (P.guide && P.guide_train_passes < 0) throw InvalidArg("guide_train_passes must be >= 0");
        if (P.guide_light_grid < 0 || P.guide_light_grid > 64) throw InvalidArg("guide_light_grid must lie in [0, 64]");
        const int64_t npix = int64_t(P.width) * P.height;
        if (npix > (int64_t(1) << 28)) throw InvalidArg("image too large");
        int spp_pp = P.samples_per_pass;
        if (spp_pp <= 0) spp_pp = auto_samples_per_pass(c, npix, P.guide == 1, nls);
        spp_pp = std::min(spp_pp, P.spp);
        const int64_t maxP = npix * spp_pp;
        if (maxP > (int64_t(1) << 30)) throw InvalidArg("too many paths per pass");
        ensure_paths(c, maxP, nls);
        c->film.reserve(size_t(npix) * sizeof(double4));
        c->film_w = P.width;
        c->film_h = P.height;
        hipStream_t st = c->stream;
        if (accumulate && film_sums)
            hip_check(hipMemcpyAsync(c->film.ptr, film_sums, size_t(npix) * 32, hipMemcpyHostToDevice, st), "H2D film");
        else
            hip_check(hipMemsetAsync(c->film.ptr, 0, size_t(npix) * 32, st), "film clear");
        hip_check(hipMemsetAsync(c->stats.ptr, 0, 128, st), "stats clear");
        // guide 1: the network (and the light cache when guide_light_grid > 0);
        // guide 2: the light cache alone, on the plain kernels
        const bool guided = P.guide == 1;
        const bool any_training = P.guide != 0 && P.guide_train_passes > 0;
        if (guided) guide_prepare(c, maxP, any_training);
        const int lgrid = P.guide != 0 ? P.guide_light_grid : 0;
        if (lgrid > 0) lc_prepare(c, lgrid, maxP, any_training, any_training && P.sample_offset == 0);
        DevPaths p = make_paths(c);
        if (guided) guide_fill_paths(c, p, any_training);
        RenderConsts rc{};
        rc.width = P.width; rc.height = P.height;
        rc.min_depth = P.min_depth; rc.max_depth = P.max_depth;
        rc.filter = P.filter_type; rc.filter_radius = P.filter_radius; rc.filter_alpha = P.filter_alpha;
        rc.shadow_tr = P.shadow_transmittance;
        rc.nls = nls;
        rc.shade_sort_rounds = shade_sort_rounds();
        rc.seed = P.seed;
        const float guide_frac = guided ? (P.guide_fraction > 0.0f ? P.guide_fraction : 0.5f) : 0.0f;
        if (guided && !(P.guide_train_fraction >= 0.0f && P.guide_train_fraction < 1.0f))
            throw InvalidArg("guide_train_fraction must lie in [0, 1)");
        rc.guide_max_depth = P.guide_max_depth;
        const float guide_lr = P.guide_lr > 0.0f ? P.guide_lr : c->guide.cfg.lr;
        uint64_t trained = 0;
        size_t evi = 0;
        const bool prof = P.profile != 0;
        double ms_ext = 0, ms_sh = 0, ms_shd = 0, ms_oth = 0, ms_gd = 0, ms_gt = 0, ms_tl = 0;
        // plain passes finish their last bounces in one kernel once at most
        // this many paths are alive
        const int tail = P.tail_paths == 0 ? 131072 : P.tail_paths;
        const bool use_tail = tail > 0 && !guided && P.profile != 2 && nls == 1;
        std::vector<std::pair<size_t, int>> marks;   // (event index, kind)
        hipEvent_t e0 = ev(c, evi++), e1 = ev(c, evi++);
        hip_check(hipEventRecord(e0, st), "event");
        auto timed = [&](int kind, auto&& launch) {
            if (!prof) { launch(); return; }
            const size_t a = evi++, b = evi++;
            hip_check(hipEventRecord(ev(c, a), st), "event");
            launch();
            hip_check(hipEventRecord(ev(c, b), st), "event");
            marks.push_back({a, kind});
        };
        uint64_t launches = 0;
        int pass = 0;
        for (int done = 0; done < P.spp; done += spp_pp, ++pass) {
            const int k = std::min(spp_pp, P.spp - done);
            rc.spp_in_pass = k;
            rc.sample_base = P.sample_offset + done;
            const int num_paths = int(npix * k);
            const bool training = P.guide != 0 && pass < P.guide_train_passes;
            rc.guide_record = training && guided ? 1 : 0;
            rc.guide_frac = training ? P.guide_train_fraction : guide_frac;
            if (training && guided) {
                rc.rec_threshold = guide_record_threshold(c, num_paths, P.max_depth);
                guide_fill_paths(c, p, true);   // the record buffers may have grown
                rc.rec_seed = uint32_t(mix64(P.seed ^ (0x7261696eull + uint64_t(rc.sample_base))));
                guide_begin_pass(c);
            }
            if (lgrid > 0) {
                // training passes after the first sample the cache they refine
                p.lc = lc_device(c, training, c->lcache.built);
                p.sh_g = c->lcache.sh_g.as<float4>();
            }
            int cur = 0;
            timed(3, [&] { launch_generate(c->ds, p, rc, num_paths, cur, st); });
            // specular bounces do not count towards the depth: a scene with
            // glass needs up to kMaxSpecularDepth more rounds
            const int rounds = P.max_depth + 1 + (c->has_delta ? kMaxSpecularDepth : 0);
            for (int b = 0; b < rounds; ++b) {
                rc.first_bounce = b == 0 ? 1 : 0;
                if (use_tail) timed(6, [&] { launch_tail(c->ds, p, rc, cur, std::min(tail, num_paths), st); });
                timed(0, [&] { launch_extend(c->ds.tr, p, cur, num_paths, P.profile == 2, st); });
                if (guided && rc.guide_frac > 0.0f) timed(4, [&] { guide_infer(c, p, cur, num_paths); });
                timed(1, [&] { launch_shade(c->ds, p, rc, cur, num_paths, guided, st); });
                if (nls > 1) timed(2, [&] { launch_shadow_multi(c->ds.tr, p, cur, num_paths, nls, st); });
                else timed(2, [&] { launch_shadow(c->ds.tr, p, num_paths, 0, rc.guide_record, st); });
                ++launches;
                cur = 1 - cur;
            }
            timed(3, [&] { launch_accumulate(p, rc, c->film.as<double4>(), st); });
            hip_check(hipGetLastError(), "render launch");
            if (training)
                timed(5, [&] {
                    if (lgrid > 0) lc_build(c);
                    if (guided) trained += uint64_t(guide_train_pass(c, p, P.seed + uint64_t(pass), guide_lr));
                });
        }
        hip_check(hipEventRecord(e1, st), "event");
        if (film_sums)
            hip_check(hipMemcpyAsync(film_sums, c->film.ptr, size_t(npix) * 32, hipMemcpyDeviceToHost, st), "D2H film");
        if (image) {
            c->s_out.reserve(size_t(npix) * 12);
            launch_resolve(c->film.as<double4>(), c->s_out.as<float>(), int(npix), st);
            hip_check(hipMemcpyAsync(image, c->s_out.ptr, size_t(npix) * 12, hipMemcpyDeviceToHost, st), "D2H image");
        }
        unsigned long long hs[16] = {0};
        hip_check(hipMemcpyAsync(hs, c->stats.ptr, sizeof hs, hipMemcpyDeviceToHost, st), "D2H stats");
        hip_check(hipStreamSynchronize(st), "render sync");
        if (stats) {
            std::memset(stats, 0, sizeof *stats);
            stats->camera_paths = uint64_t(npix) * uint64_t(P.spp);
            stats->extension_rays = hs[0];
            stats->shadow_rays = hs[1];
            stats->nonfinite_dropped = hs[2];
            float ms = 0.0f;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            stats->ms_total = ms;
            for (auto& m : marks) {
                float t = 0.0f;
                hip_check(hipEventElapsedTime(&t, ev(c, m.first), ev(c, m.first + 1)), "elapsed");
                if (m.second == 0) ms_ext += t;
                else if (m.second == 1) ms_sh += t;
                else if (m.second == 2) ms_shd += t;
                else if (m.second == 4) ms_gd += t;
                else if (m.second == 5) ms_gt += t;
                else if (m.second == 6) ms_tl += t;
                else ms_oth += t;
            }
            stats->ms_extend = ms_ext;
            stats->ms_shade = ms_sh;
            stats->ms_shadow = ms_shd;
            stats->ms_other = ms_oth;
            stats->ms_guide = ms_gd;
            stats->extend_launches = launches;
            stats->node_visits = hs[3];
            stats->tri_tests = hs[4];
            stats->guide_fallbacks = hs[5];
            stats->guide_vertices = hs[6];
            stats->guide_train_samples = trained;
            stats->guide_inferences = hs[7];
            stats->ms_guide_train = ms_gt;
            stats->tail_extension_rays = hs[8];
            stats->ms_tail = ms_tl;
        }
    });
}

tr_status tr_render_bdpt(tr_ctx* c, const tr_bdpt_params* prm, double* film_sums, int32_t accumulate, float* image,
                         tr_render_stats* stats) {
    return guarded(c, [&] {
        if (!c->have_bvh) throw StateError("tr_render_bdpt before tr_bvh_build");
        if (!prm) throw InvalidArg("params is NULL");
        const tr_bdpt_params& P = *prm;
        if (P.width <= 0 || P.height <= 0 || P.spp <= 0) throw InvalidArg("width, height and spp must be positive");
        const int nc = P.max_camera_vertex_count, nl = P.max_light_vertex_count;
        if (nc < 1 || nc > kBdMaxVertices) throw InvalidArg("max_camera_vertex_count must lie in [1, 32]");
        if (nl < 0 || nl > kBdMaxVertices) throw InvalidArg("max_light_vertex_count must lie in [0, 32]");
        check_filter(P.filter_type, P.filter_radius, P.filter_alpha);
        if (P.only_s < 0 || P.only_s > nl + 1 || P.only_t < 0 || P.only_t > nc)
            throw InvalidArg("only_s must lie in [0, max_light_vertex_count + 1] and only_t in [0, max_camera_vertex_count]");
        if (P.samples_per_pass < 0 || P.sample_offset < 0) throw InvalidArg("samples_per_pass and sample_offset must be >= 0");
        if (c->ds.has_sun) throw InvalidArg("bidirectional path tracing does not support the directional sun light");
        const int64_t npix = int64_t(P.width) * P.height;
        if (npix > (int64_t(1) << 28)) throw InvalidArg("image too large");
        const int64_t K = int64_t(nc) * (nl + 1);
        int spp_pp = P.samples_per_pass;
        if (spp_pp <= 0) {
            // samples in flight: a quarter of the free device memory plus what the
            // context holds, at most 2^26 strategy entries
            size_t free_b = 0, total_b = 0;
            hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
            const int64_t fit = c->bdpt.slots + int64_t(free_b / 4) / bdpt_slot_bytes(nc, nl);
            const int64_t target = std::min<int64_t>(fit, (int64_t(1) << 26) / K);
            spp_pp = int(std::max<int64_t>(1, target / npix));
        }
        spp_pp = std::min(spp_pp, P.spp);
        const int64_t slots = npix * spp_pp;
        if (slots * K >= (int64_t(1) << 31) || slots > (int64_t(1) << 29))
            throw InvalidArg("too many samples per pass for the strategy table (lower samples_per_pass)");
        bdpt_ensure(c, slots, nc, nl, npix);
        c->film.reserve(size_t(npix) * sizeof(double4));
        c->film_w = P.width;
        c->film_h = P.height;
        hipStream_t st = c->stream;
        BdptState& S = c->bdpt;
        if (accumulate && film_sums)
            hip_check(hipMemcpyAsync(c->film.ptr, film_sums, size_t(npix) * 32, hipMemcpyHostToDevice, st), "H2D film");
        else
            hip_check(hipMemsetAsync(c->film.ptr, 0, size_t(npix) * 32, st), "film clear");
        hip_check(hipMemsetAsync(c->stats.ptr, 0, 128, st), "stats clear");
        hip_check(hipMemsetAsync(S.splat.ptr, 0, size_t(npix) * 24, st), "splat clear");
        launch_bd_snapshot(c->film.as<double4>(), S.w0.as<double>(), int(npix), st);

        BdPaths b{};
        b.cv = S.cv.as<BdVertex>(); b.lv = S.lv.as<BdVertex>();
        b.ncv = S.ncv.as<int>(); b.nlv = S.nlv.as<int>();
        b.gray = S.gray.as<float4>(); b.ghit = S.ghit.as<float4>(); b.gbeta = S.gbeta.as<float4>();
        b.grng = S.grng.as<uint2>();
        b.gq[0] = S.gq0.as<int>(); b.gq[1] = S.gq1.as<int>();
        b.counters = c->counters.as<unsigned int>();
        b.con = S.con.as<float4>(); b.sray = S.sray.as<float4>(); b.sidx = S.sidx.as<int>();
        b.splat = S.splat.as<unsigned long long>();
        b.rad = S.rad.as<float4>();
        b.stats = c->stats.as<unsigned long long>();
        b.ovf_node = c->ovf_node.as<int>(); b.ovf_t = c->ovf_t.as<float>(); b.ovf_threads = c->ovf_threads;
        BdConsts k{};
        k.width = P.width; k.height = P.height; k.npix = int(npix);
        k.nc = nc; k.nl = nl; k.K = int(K);
        k.only_s = P.only_s - 1;
        k.only_t = P.only_t;
        k.seed = P.seed;
        const DevCamera& cam = c->ds.cam;
        k.film_area = cam.film_w * cam.film_h;
        k.lens_area = cam.lens_radius > 0.0f ? float(kPi) * cam.lens_radius * cam.lens_radius : 1.0f;
        k.env_pdf_pos = 1.0f / (float(kPi) * c->ds.wradius * c->ds.wradius);
        // k_accumulate reads the camera-side radiance of each slot
        DevPaths p{};
        p.rad = b.rad;
        p.stats = b.stats;
        RenderConsts rc{};
        rc.width = P.width; rc.height = P.height;
        rc.filter = P.filter_type; rc.filter_radius = P.filter_radius; rc.filter_alpha = P.filter_alpha;
        rc.seed = P.seed;

        size_t evi = 0;
        const bool prof = P.profile != 0;
        std::vector<std::pair<size_t, int>> marks;
        hipEvent_t e0 = ev(c, evi++), e1 = ev(c, evi++);
        hip_check(hipEventRecord(e0, st), "event");
        auto timed = [&](int kind, auto&& launch) {
            if (!prof) { launch(); return; }
            const size_t a = evi++, bb = evi++;
            hip_check(hipEventRecord(ev(c, a), st), "event");
            launch();
            hip_check(hipEventRecord(ev(c, bb), st), "event");
            marks.push_back({a, kind});
        };
        const int steps = std::max(nc, nl) - 1;
        for (int done = 0; done < P.spp; done += spp_pp) {
            const int kk = std::min(spp_pp, P.spp - done);
            k.slots = int(npix * kk);
            k.sample_base = P.sample_offset + done;
            rc.spp_in_pass = kk;
            rc.sample_base = k.sample_base;
            hip_check(hipMemsetAsync(c->counters.ptr, 0, size_t(3) * kCounterStride * sizeof(unsigned int), st), "counters");
            timed(3, [&] { launch_bd_start(c->ds, b, k, st); });
            int cur = 0;
            for (int step = 0; step < steps; ++step) {
                timed(0, [&] { launch_bd_trace(c->ds.tr, b, cur, 2 * k.slots, st); });
                timed(1, [&] { launch_bd_extend(c->ds, b, k, cur, st); });
                cur = 1 - cur;
            }
            timed(4, [&] { launch_bd_connect(c->ds, b, k, st); });
            timed(2, [&] { launch_bd_occlude(c->ds.tr, b, int64_t(k.slots) * K, st); });
            timed(3, [&] {
                launch_bd_gather(b, k, st);
                launch_accumulate(p, rc, c->film.as<double4>(), st);
            });
            hip_check(hipGetLastError(), "bdpt launch");
        }
        launch_bd_fold(b, c->film.as<double4>(), S.w0.as<double>(), int(npix), P.spp, st);
        hip_check(hipEventRecord(e1, st), "event");
        if (film_sums)
            hip_check(hipMemcpyAsync(film_sums, c->film.ptr, size_t(npix) * 32, hipMemcpyDeviceToHost, st), "D2H film");
        if (image) {
            c->s_out.reserve(size_t(npix) * 12);
            launch_resolve(c->film.as<double4>(), c->s_out.as<float>(), int(npix), st);
            hip_check(hipMemcpyAsync(image, c->s_out.ptr, size_t(npix) * 12, hipMemcpyDeviceToHost, st), "D2H image");
        }
        unsigned long long hs[16] = {0};
        hip_check(hipMemcpyAsync(hs, c->stats.ptr, sizeof hs, hipMemcpyDeviceToHost, st), "D2H stats");
        hip_check(hipStreamSynchronize(st), "bdpt sync");
        if (stats) {
            std::memset(stats, 0, sizeof *stats);
            stats->camera_paths = uint64_t(npix) * uint64_t(P.spp);
            stats->extension_rays = hs[0];
            stats->shadow_rays = hs[1];
            stats->nonfinite_dropped = hs[2] + hs[9];
            float ms = 0.0f;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            stats->ms_total = ms;
            double acc[5] = {0, 0, 0, 0, 0};
            for (auto& m : marks) {
                float t = 0.0f;
                hip_check(hipEventElapsedTime(&t, ev(c, m.first), ev(c, m.first + 1)), "elapsed");
                acc[m.second] += t;
            }
            stats->ms_extend = acc[0];
            stats->ms_shade = acc[1];
            stats->ms_shadow = acc[2];
            stats->ms_other = acc[3];
            stats->ms_tail = acc[4];
        }
    });
}

}  // extern "C"
