// Device-resident scene: geometry for traversal, per-triangle shading
// records, materials, media, lights and camera.  One struct, passed to the
// kernels by value.
#pragma once
#include "tr_bsdf.h"
#include "tr_trace.h"

namespace trhip {

// Homogeneous medium (src/medium/homogeneous.cpp:12-18) prepared for delta
// tracking: majorant = max channel sigma_t, collision-type probabilities from
// the channel-averaged coefficients (spectral tracking; for RGB-equal media
// this is plain analogue free-flight sampling).
// A homogeneous medium, or (grid >= 0) a heterogeneous one whose sigma_t is a
// density grid: sigma_s then holds the albedo, majorant the grid maximum.
struct DevMedium {
    f3 sigma_s;
    float g;
    f3 sigma_t;
    float majorant;
    float p_scatter, p_absorb, p_null;
    int max_scatter;
    int grid;              // first voxel in DevScene::density, -1 if homogeneous
    int nx, ny, nz;
    float w2l[12];         // world to grid coordinates, row-major 3x4
    int white_indirect;    // albedo 1 after the first scatter of a run
    int albedo_grid;       // first value in DevScene::density of an RGB albedo grid (anx * any * anz triples), -1: sigma_s
    int g_grid;            // first value of an asymmetry grid (gnx * gny * gnz), -1: g
    int pad0;
    int anx, any, anz, gnx, gny, gnz;
    int pad1[2];
};

// ThinLensCamera (src/camera/thin_lens_camera.cpp:23-36) with look_at's
// left-handed basis (src/utility/transform.cpp:96-124).
struct DevCamera {
    f3 pos, right, up, dir;
    float film_w, film_h, focal, lens_radius;
};

// Image environment (IBL, src/light/environment_light.cpp:20-196): texels
// and the Distribution2D built on the host (distribution.hpp:80-110).  w = 0
// means a constant environment of radiance env_L.
struct DevEnvMap {
    const float4* tex;          // h rows of w texels (RGB), row 0 at v = 0
    const float* cond_cdf;      // nv rows of nu + 1
    const float* cond_func;     // nv rows of nu
    const float* marg_cdf;      // nv + 1
    const float* marg_func;     // nv
    int w, h, nu, nv;
    float marg_int;
    f3 r0, r1, r2;              // rows of the world -> light rotation
};

constexpr int kMatTexRow = 16;
// Shading classes, in the order the shade kernels visit them: paths whose ray
// left the scene (and unused window entries), Lambert surfaces, Disney
// surfaces, every other material.
enum : int { kShadeClassMiss = 0, kShadeClassLambert = 1, kShadeClassDisney = 2, kShadeClassOther = 3, kShadeClasses = 4 };
constexpr int kMatRawRow = 20;

struct DevScene {
    TraceScene tr;
    // 4 float4 (one 64-byte record) per leaf-ordered triangle: (n0, material),
    // (n1, light), (n2, media), (geometric normal, 0); a surface hit reads its
    // record and not the triangle's vertices
    const float4* tri_shade;
    // shading class per leaf-ordered triangle (kShadeClassLambert, _Disney,
    // _Other): the shade kernels group the paths of a window by it
    const unsigned char* tri_cls;
    const DevMaterial* mats;
    const DevMedium* media;
    const float4* lights;      // 4 float4 per area light: (L, area), (A, pmf), (B, tri), (C, 0)
    const float* light_cdf;    // over [area lights..., env, sun]
    const float* light_pmf;
    const int* light_guide;    // cutpoints: first CDF index above j / light_guide_size
    int light_guide_size;      // a power of two
    int num_area_lights;
    int num_light_entries;
    int has_env, has_sun;
    f3 env_L;
    DevEnvMap env;
    // material textures (has_tex): per-triangle uvs (2 float4 in leaf order),
    // per texture (w, h, 8-bit, texel offset), 8-bit and float texel pools
    const float4* tri_uv;
    // per leaf-ordered triangle: the normal derivative along v of
    // compute_ss_ts (xyz) and 1 in w where it is used (scenes with texture
    // coordinates and an anisotropic material or a normal map; NULL otherwise)
    const float4* tri_tan;
    // Disney materials with per-hit parameters (DevMaterial::texset = 1 + row):
    // kMatTexRow texture ids (base colour, the twelve parameters in
    // tr_material::param_texture order, the normal map; 0 = constant) and
    // kMatRawRow constants in DisneyIn order
    const int* mat_tex;
    const float* mat_raw;
    const int4* tex_info;
    const uchar4* tex8;
    const float4* texf;
    int has_tex;
    // heterogeneous media: voxel pool of every density grid
    const float* density;
    int has_grid;
    f3 sun_dir;
    f3 sun_E;
    f3 wcenter;
    float wradius;
    DevCamera cam;
};

constexpr int kShadeRecord = 4;   // float4 per shading record
struct TriShade {
    f3 n0, n1, n2;
    f3 ng;   // normalize(cross(v1 - v0, v2 - v0)), written by k_geometric_normals
    int material, light, med_in, med_out;
};
TR_D TriShade load_shade(const DevScene& s, int tri) {
    const float4* p = s.tri_shade + size_t(tri) * kShadeRecord;
    const float4 a = p[0], b = p[1], c = p[2], e = p[3];
    TriShade r;
    r.n0 = xyz(a); r.n1 = xyz(b); r.n2 = xyz(c);
    r.ng = xyz(e);
    r.material = __float_as_int(a.w);
    r.light = __float_as_int(b.w);
    const int med = __float_as_int(c.w);
    r.med_in = (med << 16) >> 16;   // sign-extended low half
    r.med_out = med >> 16;
    return r;
}

// Tangent axes of a surface hit as SurfaceIntersection::shading_coord
// (src/shape/triangle.cpp:200-211) builds them around the interpolated
// normal ns: x = cross(dndv, ns), y = cross(ns, x), normalised.  dndv is the
// triangle's normal derivative along v (compute_ss_ts,
// src/utility/coordinate.hpp:54-68) where the table holds one, else the
// second axis of coordinate(ns) (src/utility/geometry.hpp:887-894), for which
// x is coordinate's first axis.  Only anisotropic BSDFs and normal maps (whose
// texels are tangent-space normals) see the difference from any other frame
// around ns.
TR_D void shading_frame(const DevScene& s, int tri, f3 ns, f3& tx, f3& ty) {
    const float4 tv = s.tri_tan ? s.tri_tan[tri] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (tv.w != 0.0f) {
        tx = normalize(cross(xyz(tv), ns));
    } else if (fabsf(ns.x) > fabsf(ns.y)) {
        tx = mk(-ns.z, 0.0f, ns.x) * (1.0f / sqrtf(ns.x * ns.x + ns.z * ns.z));
    } else {
        tx = mk(0.0f, ns.z, -ns.y) * (1.0f / sqrtf(ns.y * ns.y + ns.z * ns.z));
    }
    ty = cross(ns, tx);
}

// Light selection by power (compute_light_power_distribution,
// src/renderer/direct_illumination.cpp:257-264; Distribution1D::sample_discrete,
// src/utility/distribution.hpp:49-67): the first CDF entry above u, or the
// last entry.  The scan starts at the cutpoint of u's bucket, which lies at or
// before that entry, so it returns what a binary search returns.
TR_D int pick_light(const DevScene& s, float u, float& pmf) {
    const int last = s.num_light_entries - 1;
    const int K = s.light_guide_size;
    const int j = min(int(u * float(K)), K - 1);
    int lo = s.light_guide[max(j, 0)];
    while (lo < last && !(s.light_cdf[lo] > u)) ++lo;
    while (lo > 0 && s.light_pmf[lo] == 0.0f) --lo;
    pmf = s.light_pmf[lo];
    return lo;
}

// Texture2D::evaluate (core/texture.hpp:57-83: wrap_repeat) and
// LinearSampler::Sample2D (utility/sampler.hpp:12-25) on the environment.
TR_D f3 env_texel(const DevEnvMap& e, float u, float v) {
    u = fminf(fmaxf(u - floorf(u), 0.0f), 1.0f) * float(e.w - 1);
    v = fminf(fmaxf(v - floorf(v), 0.0f), 1.0f) * float(e.h - 1);
    const int u0 = min(int(u), e.w - 1), u1 = min(u0 + 1, e.w - 1);
    const int v0 = min(int(v), e.h - 1), v1 = min(v0 + 1, e.h - 1);
    const float du = u - float(u0), dv = v - float(v0);
    const f3 a = xyz(e.tex[size_t(v0) * e.w + u0]), b = xyz(e.tex[size_t(v0) * e.w + u1]);
    const f3 c = xyz(e.tex[size_t(v1) * e.w + u0]), d = xyz(e.tex[size_t(v1) * e.w + u1]);
    return (a * (1.0f - du) + b * du) * (1.0f - dv) + (c * (1.0f - du) + d * du) * dv;
}

// Distribution1D::sample_continuous (distribution.hpp:33-47): the last cdf
// entry not above u, the offset within it, the cell's pdf.
TR_D float sample_cont(const float* cdf, const float* func, float func_int, int n, float u, float& pdf, int& off) {
    int lo = 0, hi = n - 1;   // largest i in [0, n - 1] with cdf[i] <= u
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cdf[mid] <= u) lo = mid;
        else hi = mid - 1;
    }
    off = lo;
    float du = u - cdf[lo];
    const float w = cdf[lo + 1] - cdf[lo];
    if (w > 0.0f) du /= w;
    pdf = func_int > 0.0f ? func[lo] / func_int : 0.0f;
    return (float(lo) + du) / float(n);
}

// Light-space direction of a world direction and its (u, v) = (phi / 2pi, theta / pi).
TR_D void env_uv(const DevEnvMap& e, f3 d, float& u, float& v, float& sin_theta) {
    const f3 l = normalize(mk(dot(e.r0, d), dot(e.r1, d), dot(e.r2, d)));
    float phi = (l.x == 0.0f && l.y == 0.0f) ? 0.0f : atan2f(l.y, l.x);
    if (phi < 0.0f) phi += 2.0f * kPi;
    const float theta = acosf(fminf(fmaxf(l.z, -1.0f), 1.0f));
    u = fminf(fmaxf(phi * (0.5f / kPi), 0.0f), 1.0f);
    v = fminf(fmaxf(theta * (1.0f / kPi), 0.0f), 1.0f);
    sin_theta = sinf(theta);
}

// IBL::light_emit and IBL::pdf (environment_light.cpp:179-196) for a world
// direction.  The pdf is the density sample_env draws:
// conditional func / marginal integral / (2 pi^2 sin theta) (the reference's
// pdf() multiplies by the marginal integral instead, defect Q28).
// IMG = false compiles the constant environment only (kernels for scenes
// without an image environment keep their register budget).
template <bool IMG>
TR_D f3 env_eval(const DevScene& s, f3 d, float& pdf) {
    const DevEnvMap& e = s.env;
    if (!IMG || e.w == 0) {
        pdf = 1.0f / (4.0f * kPi);
        return s.env_L;
    }
    float u, v, st;
    env_uv(e, d, u, v, st);
    if (st == 0.0f || !(e.marg_int > 0.0f)) {
        pdf = 0.0f;
    } else {
        const int iu = min(max(int(u * float(e.nu)), 0), e.nu - 1);
        const int iv = min(max(int(v * float(e.nv)), 0), e.nv - 1);
        pdf = e.cond_func[size_t(iv) * e.nu + iu] / e.marg_int / (2.0f * kPi * kPi * st);
    }
    return env_texel(e, u, v);
}

// Texture2D::evaluate (wrap_repeat) with LinearSampler::Sample2D in floating
// point; 8-bit texels read as value / 255.
// Density of a grid medium at a world position: world to grid coordinates,
// each wrapped into [0, 1] (Texture3D::wrap_repeat, src/core/texture.hpp:114-116),
// then trilinear over the voxels (LinearSampler::Sample3D,
// src/utility/sampler.hpp:47-67).
// Grid coordinates of world point p in medium m: world_to_local, every
// coordinate wrapped into [0, 1] (Texture3D's wrap_repeat).
TR_D f3 grid_local(const DevMedium& m, f3 p) {
    const float* w = m.w2l;
    const float lx = w[0] * p.x + w[1] * p.y + w[2] * p.z + w[3];
    const float ly = w[4] * p.x + w[5] * p.y + w[6] * p.z + w[7];
    const float lz = w[8] * p.x + w[9] * p.y + w[10] * p.z + w[11];
    auto wrap = [](float x) { return fminf(fmaxf(x - floorf(x), 0.0f), 1.0f); };
    return mk(wrap(lx), wrap(ly), wrap(lz));
}

// Trilinear lookup (LinearSampler::Sample3D, src/utility/sampler.hpp:47-67)
// of an nx * ny * nz grid of CH floats per voxel (x fastest) at wrapped local
// coordinates l: CH values into out.
template <int CH>
TR_D void grid_sample(const float* g, int nx, int ny, int nz, f3 l, float* out) {
    const float u = l.x * float(nx - 1), v = l.y * float(ny - 1), k = l.z * float(nz - 1);
    const int u0 = min(int(u), nx - 1), v0 = min(int(v), ny - 1), k0 = min(int(k), nz - 1);
    const int u1 = min(u0 + 1, nx - 1), v1 = min(v0 + 1, ny - 1), k1 = min(k0 + 1, nz - 1);
    const float du = u - float(u0), dv = v - float(v0), dk = k - float(k0);
    const size_t sy = size_t(nx), sz = size_t(nx) * size_t(ny);
    auto at = [&](int x, int y, int z, int c) {
        return __ldg(g + (size_t(z) * sz + size_t(y) * sy + size_t(x)) * CH + c);
    };
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        const float c00 = at(u0, v0, k0, c) * (1.0f - du) + at(u1, v0, k0, c) * du;
        const float c10 = at(u0, v1, k0, c) * (1.0f - du) + at(u1, v1, k0, c) * du;
        const float c01 = at(u0, v0, k1, c) * (1.0f - du) + at(u1, v0, k1, c) * du;
        const float c11 = at(u0, v1, k1, c) * (1.0f - du) + at(u1, v1, k1, c) * du;
        out[c] = (c00 * (1.0f - dv) + c10 * dv) * (1.0f - dk) + (c01 * (1.0f - dv) + c11 * dv) * dk;
    }
}

TR_D float grid_density(const DevScene& s, const DevMedium& m, f3 p) {
    float d;
    grid_sample<1>(s.density + m.grid, m.nx, m.ny, m.nz, grid_local(m, p), &d);
    return d;
}

// Albedo and asymmetry of a grid medium at p (HeterogeneousMedium::sample
// reads albedo->evaluate and g->evaluate_s at the collision,
// src/medium/heterogeneous.cpp:66-78): their grids, or the constants.  g is
// kept inside (-1, 1) as the constant is.
TR_D f3 medium_albedo(const DevScene& s, const DevMedium& m, f3 p) {
    if (m.albedo_grid < 0) return m.sigma_s;
    float a[3];
    grid_sample<3>(s.density + m.albedo_grid, m.anx, m.any, m.anz, grid_local(m, p), a);
    return mk(a[0], a[1], a[2]);
}
TR_D float medium_g(const DevScene& s, const DevMedium& m, f3 p) {
    if (m.g_grid < 0) return m.g;
    float g;
    grid_sample<1>(s.density + m.g_grid, m.gnx, m.gny, m.gnz, grid_local(m, p), &g);
    return fminf(fmaxf(g, -1.0f + kEps), 1.0f - kEps);
}

// Ratio-tracking transmittance of a grid medium over [0, dist) along d
// (HeterogeneousMedium::tr, src/medium/heterogeneous.cpp:35-52).
TR_D float grid_tr(const DevScene& s, const DevMedium& m, f3 o, f3 d, float dist, Pcg& rng) {
    const float inv = 1.0f / m.majorant;
    float T = 1.0f, t = 0.0f;
    for (int k = 0; k < (1 << 16); ++k) {
        t -= logf(1.0f - rng.uniform()) * inv;
        if (t >= dist) break;
        T *= 1.0f - grid_density(s, m, o + d * t) * inv;
        if (T <= 0.0f) break;
    }
    return T;
}

TR_D f3 tex_lookup(const DevScene& s, int t, float u, float v) {
    const int4 ti = s.tex_info[t];
    const int w = ti.x, h = ti.y;
    u = fminf(fmaxf(u - floorf(u), 0.0f), 1.0f) * float(w - 1);
    v = fminf(fmaxf(v - floorf(v), 0.0f), 1.0f) * float(h - 1);
    const int u0 = min(int(u), w - 1), u1 = min(u0 + 1, w - 1);
    const int v0 = min(int(v), h - 1), v1 = min(v0 + 1, h - 1);
    const float du = u - float(u0), dv = v - float(v0);
    const size_t o = size_t(ti.w);
    f3 a, b, c, d;
    if (ti.z) {
        auto ld = [&](int x, int y) {
            const uchar4 q = s.tex8[o + size_t(y) * w + x];
            return mk(float(q.x), float(q.y), float(q.z)) * (1.0f / 255.0f);
        };
        a = ld(u0, v0); b = ld(u1, v0); c = ld(u0, v1); d = ld(u1, v1);
    } else {
        a = xyz(s.texf[o + size_t(v0) * w + u0]); b = xyz(s.texf[o + size_t(v0) * w + u1]);
        c = xyz(s.texf[o + size_t(v1) * w + u0]); d = xyz(s.texf[o + size_t(v1) * w + u1]);
    }
    return (a * (1.0f - du) + b * du) * (1.0f - dv) + (c * (1.0f - du) + d * du) * dv;
}

// The material at a textured hit: kd (Phong) or base_color (Disney) from the
// texture at the interpolated uv (alpha weights B, beta C, triangle.cpp:178-190),
// and the constants that depend on it recomputed as Phong::shading and the
// DisneyBSDF constructor compute them (phong.cpp:36-53, disney.cpp:347-399).
TR_D void apply_texture(const DevScene& s, DevMaterial& m, int tri, float b1, float b2) {
    const float4 p = s.tri_uv[2 * size_t(tri)], q = s.tri_uv[2 * size_t(tri) + 1];
    const float b0 = 1.0f - b1 - b2;
    const float u = p.x * b0 + p.z * b1 + q.x * b2, v = p.y * b0 + p.w * b1 + q.y * b2;
    const f3 c = tex_lookup(s, m.tex - 1, u, v);
    if (m.type == MAT_PHONG) {
        const f3 ks_raw = m.spec_scale;
        const float dem = fmaxf(1.0f, max3(c + ks_raw));
        const f3 kd = c * (1.0f / dem), ks = ks_raw * (1.0f / dem);
        const float ld = lum(kd), ls = lum(ks);
        m.w_diffuse = ld + ls > kEps ? ld / (ld + ls) : 0.5f;
        m.w_spec = 1.0f - m.w_diffuse;
        m.C = kd + ks;
        m.Ctint = kd;
        m.spec_scale = ks;
        if (m.guide_share > 0.0f) m.guide_share = m.w_diffuse;
    } else if (m.type == MAT_LAMBERT) {
        m.C = c * (1.0f / fmaxf(1.0f, max3(c)));
    } else if (m.type == MAT_DISNEY) {
        m.C = c;
        const float l = lum(c);
        m.Ctint = l > 0.0f ? c * (1.0f / l) : f3s(1.0f);
        const float A = fminf(fmaxf(l * (1.0f - m.metallic), 0.3f), 0.7f);
        const float B = 1.0f - A;
        m.w_diffuse = A * (1.0f - m.transmission);
        m.w_trans = A * m.transmission;
        m.w_spec = B * 2.0f / (2.0f + m.clearcoat);
        m.w_cc = B * m.clearcoat / (2.0f + m.clearcoat);
        if (m.guide_share > 0.0f) {   // the broad share of the lobe weights, as make_material sets it
            const float tot = m.w_diffuse + m.w_trans + m.w_spec + m.w_cc;
            const float broad = m.w_diffuse + (m.ax >= 0.25f ? m.w_spec : 0.0f);
            m.guide_share = tot > 0.0f ? broad / tot : 0.0f;
        }
    }
}

// Texture coordinates of a hit (alpha weights vertex 1, beta vertex 2).
TR_D void hit_uv(const DevScene& s, int tri, float b1, float b2, float& u, float& v) {
    const float4 p = s.tri_uv[2 * size_t(tri)], q = s.tri_uv[2 * size_t(tri) + 1];
    const float b0 = 1.0f - b1 - b2;
    u = p.x * b0 + p.z * b1 + q.x * b2;
    v = p.y * b0 + p.w * b1 + q.y * b2;
}

// A Disney material with textured parameters (DevMaterial::texset > 0) at
// uv: every parameter read from its texture (a scalar takes the red channel,
// Texture2D::evaluate_s) or its constant, and the DisneyBSDF constructor rerun
// (Disney::shading, src/material/disney.cpp:604-621).
TR_D void disney_params_at(const DevScene& s, DevMaterial& m, float u, float v) {
    const int* ids = s.mat_tex + size_t(m.texset - 1) * kMatTexRow;
    const float* r = s.mat_raw + size_t(m.texset - 1) * kMatRawRow;
    auto col = [&](int id, float a, float b, float c) { return id > 0 ? tex_lookup(s, id - 1, u, v) : mk(a, b, c); };
    auto scal = [&](int k, float c) { return ids[1 + k] > 0 ? tex_lookup(s, ids[1 + k] - 1, u, v).x : c; };
    DisneyIn in;
    in.base_color = col(ids[0], r[0], r[1], r[2]);
    in.metallic = scal(0, r[3]);
    in.roughness = scal(1, r[4]);
    in.transmission = scal(2, r[5]);
    in.transmission_roughness = scal(3, r[6]);
    in.ior = scal(4, r[7]);
    in.specular_scale = col(ids[6], r[8], r[9], r[10]);
    in.specular_tint = scal(6, r[11]);
    in.anisotropic = scal(7, r[12]);
    in.sheen = scal(8, r[13]);
    in.sheen_tint = scal(9, r[14]);
    in.clearcoat = scal(10, r[15]);
    in.clearcoat_gloss = scal(11, r[16]);
    disney_setup(in, m);
}

// NormalMapper::reorient (src/core/material.hpp:37-48) of the shading frame
// (tx, ty, ns) at uv when the material has a normal map: z = the frame's
// image of normalize(2c - 1) for the texel c, x = cross(y, z), y = cross(z, x)
// (Coord::rotate_to_new_z, src/utility/coordinate.hpp:107-113), normalised.
TR_D void normal_map_frame(const DevScene& s, const DevMaterial& m, float u, float v, f3& tx, f3& ty, f3& ns) {
    const int id = s.mat_tex[size_t(m.texset - 1) * kMatTexRow + 13];
    if (id <= 0) return;
    const f3 c = tex_lookup(s, id - 1, u, v);
    const f3 ln = normalize(mk(c.x * 2.0f - 1.0f, c.y * 2.0f - 1.0f, c.z * 2.0f - 1.0f));
    const f3 z = normalize(tx * ln.x + ty * ln.y + ns * ln.z);
    const f3 x = normalize(cross(ty, z));
    ty = normalize(cross(z, x));
    tx = x;
    ns = z;
}

// Everything a material reads per hit: textured Disney parameters and the
// normal map, or the kd / base-colour texture.
TR_D void apply_material_textures(const DevScene& s, DevMaterial& m, int tri, float b1, float b2, f3& tx, f3& ty,
                                  f3& ns) {
    if (m.texset > 0) {
        float u, v;
        hit_uv(s, tri, b1, b2, u, v);
        disney_params_at(s, m, u, v);
        normal_map_frame(s, m, u, v, tx, ty, ns);
    } else if (m.tex > 0) {
        apply_texture(s, m, tri, b1, b2);
    }
}

struct LightSampleD {
    f3 pos;     // point on the light (area) or a far point (env / sun)
    f3 wi;      // unit direction from the reference point
    float dist; // distance to pos
    f3 L;
    float pdf;  // solid-angle pdf without selection; 0 = invalid; < 0 = delta light
    bool area;
};

// DiffuseLight::sample_li (src/light/diffuse_light.cpp:44-60) over a triangle
// (Triangle::sample, src/shape/triangle.cpp:61-97: b.x weights A, b.y weights B);
// a constant environment is sampled uniformly over the sphere, an image one
// by its 2D distribution, and the sun is a delta direction.
template <bool IMG>
TR_D LightSampleD sample_light(const DevScene& s, int entry, f3 ref, float u, float v) {
    LightSampleD r;
    r.pdf = 0.0f;
    r.area = false;
    if (entry < s.num_area_lights) {
        const float4* lr = s.lights + 4 * size_t(entry);
        const float4 l0 = lr[0];
        const f3 A = xyz(lr[1]), B = xyz(lr[2]), C = xyz(lr[3]);
        const float su = sqrtf(u);
        const float bx = 1.0f - su, by = v * su;
        const f3 pos = A * bx + B * by + C * (1.0f - bx - by);
        const f3 n = normalize(cross(B - A, C - A));
        const f3 to_ref = ref - pos;
        const float d2 = dot(to_ref, to_ref);
        const float d = sqrtf(d2);
        const float cosl = dot(n, to_ref) / d;
        if (!(cosl > 0.0f)) return r;
        r.pos = pos;
        r.wi = to_ref * (-1.0f / d);
        r.dist = d;
        r.L = xyz(l0);
        r.pdf = d2 / (l0.w * cosl);
        r.area = true;
        return r;
    }
    if (IMG && entry == s.num_area_lights && s.env.w > 0) {
        // IBL::sample_li (environment_light.cpp:92-119): v picks theta, u phi
        const DevEnvMap& e = s.env;
        float pv, pu;
        int iv, iu;
        const float d1 = sample_cont(e.marg_cdf, e.marg_func, e.marg_int, e.nv, v, pv, iv);
        const float* ccdf = e.cond_cdf + size_t(iv) * (e.nu + 1);
        const float* cfun = e.cond_func + size_t(iv) * e.nu;
        const float d0 = sample_cont(ccdf, cfun, e.marg_func[iv], e.nu, u, pu, iu);
        const float map_pdf = pu * pv;
        if (!(map_pdf > 0.0f)) return r;
        const float theta = d1 * kPi, phi = d0 * 2.0f * kPi;
        float st, ct, sp, cp;
        sincosf(theta, &st, &ct);
        sincosf(phi, &sp, &cp);
        if (st == 0.0f) return r;
        const f3 l = mk(st * cp, st * sp, ct);
        // light -> world is the transpose of the world -> light rotation
        r.wi = normalize(mk(e.r0.x * l.x + e.r1.x * l.y + e.r2.x * l.z, e.r0.y * l.x + e.r1.y * l.y + e.r2.y * l.z,
                            e.r0.z * l.x + e.r1.z * l.y + e.r2.z * l.z));
        r.dist = 2.0f * s.wradius;
        r.pos = ref + r.wi * r.dist;
        r.L = env_texel(e, d0, d1);
        r.pdf = map_pdf / (2.0f * kPi * kPi * st);
        return r;
    }
    if (entry == s.num_area_lights) {
        const float z = 1.0f - 2.0f * u;
        const float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        float sp, cp;
        sincosf(2.0f * kPi * v, &sp, &cp);
        r.wi = mk(rr * cp, rr * sp, z);
        r.dist = 2.0f * s.wradius;
        r.pos = ref + r.wi * r.dist;
        r.L = s.env_L;
        r.pdf = 1.0f / (4.0f * kPi);
        return r;
    }
    r.wi = s.sun_dir;
    r.dist = 2.0f * s.wradius;
    r.pos = ref + r.wi * r.dist;
    r.L = s.sun_E;
    r.pdf = -1.0f;
    return r;
}

// Henyey-Greenstein phase function (src/medium/phase_function.hpp:12-77);
// u = -cos(wi, wo) as in the reference.
TR_D float hg_phase(float g, float u) {
    const float g2 = g * g;
    const float dem = 1.0f + g2 - 2.0f * g * u;
    return (1.0f - g2) / (4.0f * kPi * dem * sqrtf(dem));
}
TR_D f3 hg_sample(float g, f3 wo, float u1, float u2, float& pdf) {
    const float s = 2.0f * u1 - 1.0f;
    float u;
    if (fabsf(g) < 0.001f) u = s;
    else {
        const float g2 = g * g;
        const float q = (1.0f - g2) / (1.0f + g * s);
        u = (1.0f + g2 - q * q) / (2.0f * g);
    }
    const float ct = -u;
    const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float sp, cp;
    sincosf(2.0f * kPi * u2, &sp, &cp);
    f3 t, b;
    const f3 n = normalize(wo);
    onb(n, t, b);
    pdf = hg_phase(g, u);
    return normalize(t * (st * sp) + b * (st * cp) + n * ct);
}

TR_D void camera_ray(const DevCamera& c, float fx, float fy, float lu, float lv, f3& o, f3& d) {
    const f3 film = mk((0.5f - fx) * c.film_w, (0.5f - fy) * c.film_h, c.focal);
    float dx = 0.0f, dy = 0.0f;
    if (c.lens_radius > 0.0f) concentric_disk(lu, lv, dx, dy);
    const f3 lens = mk(c.lens_radius * dx, c.lens_radius * dy, 0.0f);
    const f3 ld = normalize(film - lens);
    o = c.pos + c.right * lens.x + c.up * lens.y;
    d = normalize(c.right * ld.x + c.up * ld.y + c.dir * ld.z);
}

}  // namespace trhip
