// refcpu — TEST INFRASTRUCTURE ONLY (see ref_math.hpp header).
//
// Scene, lights, medium, camera, film and the path-tracing loop of the
// reference, plus the extern "C" entry points the tests and the CPU-baseline
// leg of bench.py call through ctypes.
//
//   PixelSamplerRenderer::render   src/renderer/pixel_sampler_renderer.cpp:24-109
//   PathTraceRenderer::eval_pixel_li  src/renderer/path_trace_renderer.cpp:27-210
//   NEE / MIS helpers             src/renderer/direct_illumination.cpp:11-264
//   Triangle surface frames       src/shape/triangle.cpp:178-213, :61-97
//   DiffuseLight                  src/light/diffuse_light.cpp:17-60
//   IBL (constant or image)        src/light/environment_light.cpp:20-196 (ref_env.hpp)
//   HomogeneousMedium / vacuum    src/medium/homogeneous.cpp:12-66, src/medium/vacuum.cpp:8-29
//   ThinLensCamera::generate_ray  src/camera/thin_lens_camera.cpp:23-59, src/utility/transform.cpp:96-124
//   GaussianFilter / Film tiles   src/filter/gaussian_filter.cpp:9-35, src/core/render.hpp:37-115
//   parallel_for_2d               src/utility/parallel.hpp:65-107
//
// Modes
//   quirks = 1 (Reference): the reference's estimator as written, including
//     Q1 (medium segment sampled from the camera-ray origin), Q4 (scatter
//     weight 1), Q5 (MIS-BSDF light hit attenuated by Tr(ray.o, isect.pos)),
//     Q6 (no environment NEE at medium vertices), Q23 (emission of the
//     surface a segment reaches added before the segment's medium sample),
//     Q27 (specular flag kept across a medium scatter), Q26 (no direct light at a
//     vertex whose continuation sample fails), the Disney Q2/Q3/Q21 and the
//     Phong/Lambert Q22/Q24 and the 8-bit texture interpolation Q29.
//   quirks = 0 (Fixed): the same loop with those defects corrected.
//   nee = 0: every light sampled at every vertex (reference); nee = 1: one
//     light picked by power (compute_light_power_distribution,
//     direct_illumination.cpp:257-264).
// The RNG is PCG32 seeded per (pixel, sample) so renders are reproducible
// (the reference seeds one stream per thread, sampler.hpp:145-147).
#include "ref_math.hpp"
#include "ref_rng.hpp"
#include "ref_triangle.hpp"
#include "ref_bvh.hpp"
#include "ref_bsdf.hpp"
#include "ref_env.hpp"
#include "tracer_hip.h"

#include <atomic>
#include <chrono>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace refcpu {

struct Sampler {
    Pcg32 rng;
    float get() { return rng.next_float(); }
};

struct Isect {
    V3f pos, wo;
    Coord<float> geo, shd;
    int prim = -1;
    float t = 0;
    float u = 0, v = 0;   // interpolated uv (alpha weights B, beta C)
    V3f eps_offset(const V3f& dir) const {   // intersection.hpp:19-23
        if (dot(geo.z, dir) > 0) return pos + geo.z * kEps;
        return pos - geo.z * kEps;
    }
};

// An image texture of the scene: 8-bit or float RGB texels.
struct RefTex {
    int w = 0, h = 0;
    std::vector<uint8_t> rgb8;
    std::vector<float> rgbf;
    // Texture2D::evaluate (texture.hpp:57-83, wrap_repeat) with
    // LinearSampler::Sample2D (sampler.hpp:12-25).
    //   Q29 on 8-bit texels the reference interpolates Point3<uint8_t>
    //       values: every weighted term and sum is truncated to an integer
    //       before the final / 255 (image_texture2d.cpp:24-28).  Fixed mode
    //       interpolates texel / 255 in floating point.
    RGB<float> eval(double u, double v, bool quirks) const {
        u = std::clamp(u - std::floor(u), 0.0, 1.0) * (w - 1);
        v = std::clamp(v - std::floor(v), 0.0, 1.0) * (h - 1);
        const int u0 = std::clamp(int(u), 0, w - 1), u1 = std::clamp(u0 + 1, 0, w - 1);
        const int v0 = std::clamp(int(v), 0, h - 1), v1 = std::clamp(v0 + 1, 0, h - 1);
        const double du = u - u0, dv = v - v0;
        RGB<float> out;
        for (int k = 0; k < 3; ++k) {
            auto at = [&](int x, int y) { return (size_t(y) * size_t(w) + size_t(x)) * 3 + size_t(k); };
            if (!rgb8.empty() && quirks) {
                auto q = [](double x) { return int(uint8_t(x)); };
                const int top = q(q(rgb8[at(u0, v0)] * (1.0 - du)) + q(rgb8[at(u1, v0)] * du));
                const int bot = q(q(rgb8[at(u0, v1)] * (1.0 - du)) + q(rgb8[at(u1, v1)] * du));
                out[k] = float(q(q(top * (1.0 - dv)) + q(bot * dv))) / 255.0f;
            } else {
                auto tx = [&](int x, int y) { return rgb8.empty() ? double(rgbf[at(x, y)]) : double(rgb8[at(x, y)]) / 255.0; };
                out[k] = float((tx(u0, v0) * (1.0 - du) + tx(u1, v0) * du) * (1.0 - dv) +
                               (tx(u0, v1) * (1.0 - du) + tx(u1, v1) * du) * dv);
            }
        }
        return out;
    }
};

struct MediumSample { bool scattered = false; V3f pos; RGB<float> w{1, 1, 1}; float g = 0; };

// Heterogeneous medium (src/medium/heterogeneous.cpp:10-108) over a float
// density grid, x fastest: world_to_local = inverse(local_to_world), every
// coordinate wrapped into [0, 1] (texture.hpp:114-116), trilinear in double
// (LinearSampler::Sample3D, sampler.hpp:47-67); majorant = the largest voxel,
// at least eps (common.hpp:47).
struct RefGrid {
    std::vector<float> d;   // empty: homogeneous medium
    int nx = 0, ny = 0, nz = 0;
    float w2l[12] = {};
    float inv_max = 0;
    RGB<float> albedo;
    float g = 0;
    // albedo->evaluate and g->evaluate_s at the collision (heterogeneous.cpp:66-78)
    // from their own grids when given: RGB triples, asymmetries
    std::vector<float> ad, gd;
    int anx = 0, any = 0, anz = 0, gnx = 0, gny = 0, gnz = 0;
    bool white = false;
    void init(const tr_medium& m) {
        d.assign(m.density, m.density + size_t(m.nx) * size_t(m.ny) * size_t(m.nz));
        nx = m.nx; ny = m.ny; nz = m.nz;
        g = m.g;
        if (m.albedo_grid) {
            ad.assign(m.albedo_grid, m.albedo_grid + 3 * size_t(m.anx) * size_t(m.any) * size_t(m.anz));
            anx = m.anx; any = m.any; anz = m.anz;
        }
        if (m.g_grid) {
            gd.assign(m.g_grid, m.g_grid + size_t(m.gnx) * size_t(m.gny) * size_t(m.gnz));
            gnx = m.gnx; gny = m.gny; gnz = m.gnz;
        }
        float mx = 0;
        for (float v : d) mx = std::max(mx, v);
        inv_max = 1 / std::max(mx, kEps);
        albedo = RGB<float>(m.sigma_s[0], m.sigma_s[1], m.sigma_s[2]);
        white = m.white_for_indirect != 0;
        // inverse of the 3x4 affine local_to_world by cofactors
        const float* a = m.local_to_world;
        const double c00 = double(a[5]) * a[10] - double(a[6]) * a[9], c01 = double(a[2]) * a[9] - double(a[1]) * a[10];
        const double c02 = double(a[1]) * a[6] - double(a[2]) * a[5], c10 = double(a[6]) * a[8] - double(a[4]) * a[10];
        const double c11 = double(a[0]) * a[10] - double(a[2]) * a[8], c12 = double(a[2]) * a[4] - double(a[0]) * a[6];
        const double c20 = double(a[4]) * a[9] - double(a[5]) * a[8], c21 = double(a[1]) * a[8] - double(a[0]) * a[9];
        const double c22 = double(a[0]) * a[5] - double(a[1]) * a[4];
        const double det = a[0] * c00 + a[1] * c10 + a[2] * c20;
        const double r[9] = {c00 / det, c01 / det, c02 / det, c10 / det, c11 / det, c12 / det, c20 / det, c21 / det, c22 / det};
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) w2l[4 * i + j] = float(r[3 * i + j]);
            w2l[4 * i + 3] = float(-(r[3 * i] * a[3] + r[3 * i + 1] * a[7] + r[3 * i + 2] * a[11]));
        }
    }
    // channel c of a grid of ch values per voxel (x fastest) at world point p
    float lookup(const std::vector<float>& grid, int gx, int gy, int gz, int ch, int c, const V3f& p) const {
        auto wrap = [](double x) { return std::clamp(x - std::floor(x), 0.0, 1.0); };
        const float lx = w2l[0] * p.x + w2l[1] * p.y + w2l[2] * p.z + w2l[3];
        const float ly = w2l[4] * p.x + w2l[5] * p.y + w2l[6] * p.z + w2l[7];
        const float lz = w2l[8] * p.x + w2l[9] * p.y + w2l[10] * p.z + w2l[11];
        const double u = wrap(lx) * (gx - 1), v = wrap(ly) * (gy - 1), k = wrap(lz) * (gz - 1);
        const int u0 = std::clamp(int(u), 0, gx - 1), u1 = std::clamp(u0 + 1, 0, gx - 1);
        const int v0 = std::clamp(int(v), 0, gy - 1), v1 = std::clamp(v0 + 1, 0, gy - 1);
        const int k0 = std::clamp(int(k), 0, gz - 1), k1 = std::clamp(k0 + 1, 0, gz - 1);
        const double du = u - u0, dv = v - v0, dk = k - k0;
        auto at = [&](int x, int y, int z) {
            return double(grid[((size_t(z) * size_t(gy) + size_t(y)) * size_t(gx) + size_t(x)) * size_t(ch) + size_t(c)]);
        };
        return float(((at(u0, v0, k0) * (1.0 - du) + at(u1, v0, k0) * du) * (1.0 - dv) +
                      (at(u0, v1, k0) * (1.0 - du) + at(u1, v1, k0) * du) * dv) * (1.0 - dk) +
                     ((at(u0, v0, k1) * (1.0 - du) + at(u1, v0, k1) * du) * (1.0 - dv) +
                      (at(u0, v1, k1) * (1.0 - du) + at(u1, v1, k1) * du) * dv) * dk);
    }
    float density(const V3f& p) const { return lookup(d, nx, ny, nz, 1, 0, p); }
    RGB<float> albedo_at(const V3f& p) const {
        if (ad.empty()) return albedo;
        return RGB<float>(lookup(ad, anx, any, anz, 3, 0, p), lookup(ad, anx, any, anz, 3, 1, p),
                          lookup(ad, anx, any, anz, 3, 2, p));
    }
    float g_at(const V3f& p) const { return gd.empty() ? g : lookup(gd, gnx, gny, gnz, 1, 0, p); }
    // HeterogeneousMedium::tr (heterogeneous.cpp:35-52): ratio tracking
    float tr(const V3f& a, const V3f& b, Sampler& sp) const {
        float res = 1, t = 0;
        const float tmax = float((b - a).length());
        const V3f dir = (b - a).normalize();
        while (true) {
            t += -std::log(1 - sp.get()) * inv_max;
            if (t >= tmax) break;
            res *= 1 - density(a + dir * t) * inv_max;
        }
        return res;
    }
    // HeterogeneousMedium::sample (heterogeneous.cpp:54-82): delta tracking,
    // weight = albedo at a scatter (1 after the first scatter of a run when
    // white_for_indirect), 1 when the segment is crossed
    MediumSample sample(const V3f& a, const V3f& b, Sampler& sp, bool indirect) const {
        MediumSample r;
        const float tmax = float((a - b).length());
        const V3f dir = (b - a).normalize();
        float t = 0;
        while (true) {
            t += -std::log(1 - sp.get()) * inv_max;
            if (t >= tmax) break;
            const V3f pos = a + dir * t;
            if (sp.get() < density(pos) * inv_max) {
                r.scattered = true;
                r.pos = pos;
                r.w = indirect && white ? RGB<float>(1) : albedo_at(pos);
                r.g = g_at(pos);
                return r;
            }
        }
        return r;
    }
};

struct Scene {
    std::vector<V3f> P, N;
    std::vector<RefTex> tex;
    std::vector<float> UV;
    std::vector<int> idx, mat, med_in, med_out;
    std::vector<float> emis;
    std::vector<tr_material> mats;
    std::vector<tr_medium> media;
    std::vector<RefGrid> grids;       // per medium; empty d: homogeneous
    RefBVH bvh;
    std::vector<int> lights;          // emissive triangles, one DiffuseLight each
    bool has_env = false;
    RGB<float> env;
    std::shared_ptr<RefEnvMap> envmap;   // image environment, else the constant env
    RGB<float> env_radiance(const V3f& dir) const { return envmap ? envmap->light_emit(dir) : env; }
    bool has_sun = false;
    V3f sun_dir;
    RGB<float> sun_E;
    tr_camera cam;
    Bounds3 world;
    V3f wcenter;
    float wradius = 0;
    // power distribution over [area lights..., env, sun]
    std::vector<float> lpmf, lcdf;
    int n_light_entries = 0;

    V3f vtx(int t, int k) const { return P[idx[3 * t + k]]; }
    float tri_area(int t) const {   // triangle.cpp:54-59
        V3f A = vtx(t, 0), B = vtx(t, 1), C = vtx(t, 2);
        return float(0.5 * cross(B - A, C - A).length());
    }
    RGB<float> tri_emission(int t) const {
        if (emis.empty()) return RGB<float>(0);
        return RGB<float>(emis[3 * t], emis[3 * t + 1], emis[3 * t + 2]);
    }
};

// triangle.cpp:178-213 — surface interaction for a hit (alpha -> B, beta -> C)
static void fill_isect(const Scene& s, int prim, const V3f& o, const V3f& d, float t, float alpha, float beta, Isect& is) {
    const int i0 = s.idx[3 * prim], i1 = s.idx[3 * prim + 1], i2 = s.idx[3 * prim + 2];
    const V3f A = s.P[i0], B = s.P[i1], C = s.P[i2];
    const V3f AB = B - A, AC = C - A;
    float ua = 0, va = 0, ub = 0, vb = 0, uc = 0, vc = 0;
    if (!s.UV.empty()) {
        ua = s.UV[2 * i0]; va = s.UV[2 * i0 + 1];
        ub = s.UV[2 * i1]; vb = s.UV[2 * i1 + 1];
        uc = s.UV[2 * i2]; vc = s.UV[2 * i2 + 1];
    }
    V3f gn = cross(AB, AC).normalize();
    V3f nA = gn, nB = gn, nC = gn;
    if (!s.N.empty()) { nA = s.N[i0].normalize(); nB = s.N[i1].normalize(); nC = s.N[i2].normalize(); }
    is.wo = -d;
    is.pos = o + d * t;
    is.prim = prim;
    is.t = t;
    is.u = ua + (ub - ua) * alpha + (uc - ua) * beta;
    is.v = va + (vb - va) * alpha + (vc - va) * beta;
    V3f dpdu, dpdv;
    compute_ss_ts(AB, AC, ub - ua, vb - va, uc - ua, vc - va, gn, dpdu, dpdv);
    dpdu = cross(dpdv, gn);
    dpdv = cross(gn, dpdu);
    is.geo = Coord<float>(dpdu, dpdv, gn);
    V3f sn = nA + (nB - nA) * alpha + (nC - nA) * beta;
    V3f dndu, dndv;
    compute_ss_ts(nB - nA, nC - nA, ub - ua, vb - va, uc - ua, vc - va, sn, dndu, dndv);
    if (dndu.length_squared() < kEps || dndv.length_squared() < kEps) coordinate_system(sn, dndu, dndv);
    dndu = cross(dndv, sn);
    dndv = cross(sn, dndu);
    is.shd = Coord<float>(dndu, dndv, sn);
}

// Reference closest hit: BVHAccel::intersect_p with Moller-Trumbore; the
// surface interaction is formed for every accepted candidate as in
// geometry.cpp:49-60 / triangle.cpp:175-213.
static bool intersect_ref(const Scene& s, const V3f& o, const V3f& dn, float tmin, float tmax, Isect* is) {
    float tm = tmax;
    return s.bvh.traverse(o, dn, tm, [&](int p, float& tcur) {
        MTHit h;
        if (!mt_intersect(o, dn, tmin, tcur, s.vtx(p, 0), s.vtx(p, 1), s.vtx(p, 2), &h)) return false;
        tcur = h.t;
        if (is) fill_isect(s, p, o, dn, h.t, h.alpha, h.beta, *is);
        return true;
    }, false);
}
static bool occluded_ref(const Scene& s, const V3f& o, const V3f& dn, float tmin, float tmax) {
    float tm = tmax;
    return s.bvh.traverse(o, dn, tm, [&](int p, float& tcur) {
        return mt_intersect(o, dn, tmin, tcur, s.vtx(p, 0), s.vtx(p, 1), s.vtx(p, 2), nullptr);
    }, true);
}

// ----------------------------------------------------------------------------
// Lights
struct LightSample { V3f pos; RGB<float> L; float pdf = 0; bool valid = false; };

// diffuse_light.cpp:44-60 with triangle.cpp:61-97 (b.x weights A, b.y weights B)
static LightSample sample_area(const Scene& s, int tri, const V3f& ref, float u, float v) {
    LightSample ls;
    float su0 = std::sqrt(u);
    float bx = 1 - su0, by = v * su0;
    V3f A = s.vtx(tri, 0), B = s.vtx(tri, 1), C = s.vtx(tri, 2);
    V3f pos = A * bx + B * by + C * (1 - bx - by);
    V3f gn = cross(B - A, C - A).normalize();
    if (dot(gn, ref - pos) <= 0) return ls;
    V3f sp2ref = ref - pos;
    float dist2 = float(sp2ref.length_squared());
    float area_pdf = 1 / s.tri_area(tri);
    ls.pdf = area_pdf * dist2 / std::abs(float(dot(gn, sp2ref) / sp2ref.length()));
    ls.pos = pos;
    ls.L = s.tri_emission(tri);
    ls.valid = true;
    return ls;
}
// diffuse_light.cpp:31-42
static float pdf_area(const Scene& s, int tri, const V3f& ref, const V3f& pos) {
    V3f A = s.vtx(tri, 0), B = s.vtx(tri, 1), C = s.vtx(tri, 2);
    V3f n = cross(B - A, C - A).normalize();
    if (dot(n, ref - pos) <= 0) return 0;
    float area = s.tri_area(tri);
    V3f p2r = ref - pos;
    float d2 = float(p2r.length_squared());
    float ac = std::abs(float(dot(n, p2r) / (n.length() * p2r.length())));
    return d2 / (area * ac);
}
// environment_light.cpp:92-119,179-196 for a 1x1 constant texture: the 2D
// distribution is uniform, theta = pi v, phi = 2 pi u, pdf = 1/(2 pi^2 sin).
static LightSample sample_env(const Scene& s, const V3f& ref, float u, float v) {
    LightSample ls;
    if (s.envmap) {
        V3f wi;
        if (!s.envmap->sample(u, v, wi, ls.pdf, ls.L)) return ls;
        ls.pos = ref + wi * (2 * s.wradius);
        ls.valid = true;
        return ls;
    }
    float theta = v * kPi<float>, phi = u * 2 * kPi<float>;
    float st = std::sin(theta), ct = std::cos(theta);
    V3f wi = V3f(st * std::cos(phi), st * std::sin(phi), ct).normalize();
    if (st == 0) return ls;
    ls.pdf = 1 / (2 * kPi<float> * kPi<float> * st);
    ls.L = s.env;
    ls.pos = ref + wi * (2 * s.wradius);
    ls.valid = true;
    return ls;
}
static float pdf_env(const Scene& s, const V3f& dir, bool quirks) {
    if (s.envmap) return s.envmap->pdf(dir, quirks);
    float theta = std::acos(std::clamp<float>(dir.z / float(dir.length()), -1, 1));
    float st = std::sin(theta);
    if (st == 0) return 0;
    return 1 / (2 * kPi<float> * kPi<float> * st);
}

// ----------------------------------------------------------------------------
// Medium: homogeneous.cpp:26-66 (quirks) / spectral-MIS free flight (fixed)
static RGB<float> medium_tr(const tr_medium* m, const V3f& a, const V3f& b) {
    if (!m) return RGB<float>(1);
    float d = float((a - b).length());
    return RGB<float>(std::exp(-(m->sigma_a[0] + m->sigma_s[0]) * d), std::exp(-(m->sigma_a[1] + m->sigma_s[1]) * d),
                      std::exp(-(m->sigma_a[2] + m->sigma_s[2]) * d));
}
static MediumSample medium_sample(const tr_medium* m, const V3f& a, const V3f& b, Sampler& sp, bool quirks) {
    MediumSample r;
    if (!m) return r;
    RGB<float> ss(m->sigma_s[0], m->sigma_s[1], m->sigma_s[2]);
    RGB<float> st(m->sigma_a[0] + ss.r, m->sigma_a[1] + ss.g, m->sigma_a[2] + ss.b);
    if (ss.is_black()) { r.w = medium_tr(m, a, b); return r; }
    float u = sp.get(), v = sp.get();
    int ch = std::min<int>(int(u * 3), 2);
    float dist = -std::log(1 - v) / st[ch];
    float tmax = float((b - a).length());
    V3f dir = (b - a).normalize();
    bool scat = dist < tmax;
    RGB<float> tr;
    for (int i = 0; i < 3; ++i) tr[i] = std::exp(-st[i] * std::min(dist, tmax));
    RGB<float> density = scat ? (quirks ? ss * tr : st * tr) : tr;
    float pdf = (density.r + density.g + density.b) / 3;
    if (quirks) pdf = std::max(pdf, kEps);
    r.w = tr / pdf;
    if (scat) r.w *= ss;
    if (scat) { r.scattered = true; r.pos = a + dir * dist; }
    return r;
}

// ----------------------------------------------------------------------------
// BSDF dispatch (float, reference precision)
struct SurfBSDF {
    int type = 0;
    std::unique_ptr<DisneyRef<float>> dis;
    std::unique_ptr<LambertRef<float>> lam;
    std::unique_ptr<PhongRef<float>> pho;
    std::unique_ptr<GlassRef<float>> gla;
    RGB<float> eval(const V3f& wi, const V3f& wo) const {
        return type == TR_MAT_DISNEY ? dis->eval(wi, wo)
             : type == TR_MAT_PHONG  ? pho->eval(wi, wo)
             : type == TR_MAT_GLASS  ? gla->eval(wi, wo)
                                     : lam->eval(wi, wo);
    }
    float pdf(const V3f& wi, const V3f& wo) const {
        return type == TR_MAT_DISNEY ? dis->pdf(wi, wo)
             : type == TR_MAT_PHONG  ? pho->pdf(wi, wo)
             : type == TR_MAT_GLASS  ? gla->pdf(wi, wo)
                                     : lam->pdf(wi, wo);
    }
    BSDFSample<float> sample(const V3f& wo, float a, float b, float c) const {
        return type == TR_MAT_DISNEY ? dis->sample(wo, a, b, c)
             : type == TR_MAT_PHONG  ? pho->sample(wo, a, b, c)
             : type == TR_MAT_GLASS  ? gla->sample(wo, a, b, c)
                                     : lam->sample(wo, a, b, c);
    }
};
static DisneyParams to_params(const tr_material& m) {
    DisneyParams p;
    for (int i = 0; i < 3; ++i) { p.base_color[i] = m.base_color[i]; p.specular_scale[i] = m.specular_scale[i]; }
    p.metallic = m.metallic; p.roughness = m.roughness; p.transmission = m.transmission;
    p.transmission_roughness = m.transmission_roughness; p.ior = m.ior; p.specular_tint = m.specular_tint;
    p.anisotropic = m.anisotropic; p.sheen = m.sheen; p.sheen_tint = m.sheen_tint; p.clearcoat = m.clearcoat;
    p.clearcoat_gloss = m.clearcoat_gloss;
    return p;
}
template <class T> static RGB<T> phong_kd(const tr_material& m) { return RGB<T>(m.base_color[0], m.base_color[1], m.base_color[2]); }
template <class T> static RGB<T> phong_ks(const tr_material& m) { return RGB<T>(m.specular_scale[0], m.specular_scale[1], m.specular_scale[2]); }
static void make_bsdf(const Scene& s, const Isect& is, bool quirks, SurfBSDF& b) {
    tr_material m = s.mats[s.mat[is.prim]];
    if (m.base_color_texture > 0) {   // kd / base_color from the texture at the hit's uv
        const RGB<float> c = s.tex[size_t(m.base_color_texture) - 1].eval(is.u, is.v, quirks);
        m.base_color[0] = c.r; m.base_color[1] = c.g; m.base_color[2] = c.b;
        if (m.type == TR_MAT_LAMBERT) {   // Phong(ks = 0): kd / max(1, max_i kd_i)
            const float dem = std::max(1.0f, std::max(c.r, std::max(c.g, c.b)));
            for (int k = 0; k < 3; ++k) m.base_color[k] /= dem;
        }
    }
    b.type = m.type;
    if (m.type == TR_MAT_DISNEY) {
        // Disney::shading (disney.cpp:604-621): every parameter texture at the
        // hit's uv, evaluate_s taking the red channel (texture.hpp:73-82), and
        // the shading frame turned by NormalMapper::reorient (material.hpp:37-48)
        DisneyParams p = to_params(m);
        auto scal = [&](int k, float c) {
            return m.param_texture[k] > 0 ? s.tex[size_t(m.param_texture[k]) - 1].eval(is.u, is.v, quirks).r : c;
        };
        p.metallic = scal(TR_PARAM_METALLIC, p.metallic);
        p.roughness = scal(TR_PARAM_ROUGHNESS, p.roughness);
        p.transmission = scal(TR_PARAM_TRANSMISSION, p.transmission);
        p.transmission_roughness = scal(TR_PARAM_TRANSMISSION_ROUGHNESS, p.transmission_roughness);
        p.ior = scal(TR_PARAM_IOR, p.ior);
        if (m.param_texture[TR_PARAM_SPECULAR_SCALE] > 0) {
            const RGB<float> c = s.tex[size_t(m.param_texture[TR_PARAM_SPECULAR_SCALE]) - 1].eval(is.u, is.v, quirks);
            p.specular_scale[0] = c.r; p.specular_scale[1] = c.g; p.specular_scale[2] = c.b;
        }
        p.specular_tint = scal(TR_PARAM_SPECULAR_TINT, p.specular_tint);
        p.anisotropic = scal(TR_PARAM_ANISOTROPIC, p.anisotropic);
        p.sheen = scal(TR_PARAM_SHEEN, p.sheen);
        p.sheen_tint = scal(TR_PARAM_SHEEN_TINT, p.sheen_tint);
        p.clearcoat = scal(TR_PARAM_CLEARCOAT, p.clearcoat);
        p.clearcoat_gloss = scal(TR_PARAM_CLEARCOAT_GLOSS, p.clearcoat_gloss);
        Coord<float> shd = is.shd;
        if (m.normal_texture > 0) {
            const RGB<float> c = s.tex[size_t(m.normal_texture) - 1].eval(is.u, is.v, quirks);
            const V3f ln = V3f(c.r * 2 - 1, c.g * 2 - 1, c.b * 2 - 1).normalize();
            const V3f wn = shd.to_global(ln);
            const V3f nx = cross(shd.y, wn);   // coordinate.hpp:107-113
            const V3f ny = cross(wn, nx);
            shd = Coord<float>(nx, ny, wn);
        }
        b.dis.reset(new DisneyRef<float>(is.geo, shd, p, !quirks));
    }
    else if (m.type == TR_MAT_PHONG) b.pho.reset(new PhongRef<float>(is.geo, is.shd, phong_kd<float>(m), phong_ks<float>(m), m.shininess, !quirks));
    else if (m.type == TR_MAT_GLASS) b.gla.reset(new GlassRef<float>(is.geo, is.shd, phong_kd<float>(m), phong_ks<float>(m), m.ior));
    else b.lam.reset(new LambertRef<float>(is.geo, is.shd, RGB<float>(m.base_color[0], m.base_color[1], m.base_color[2]), !quirks));
}

// ----------------------------------------------------------------------------
// Film filter and Tile::add_sample (src/core/render.hpp:37-57): a sample at
// film position (x, y) in pixel units goes to every pixel (qx, qy) of the tile
// with qx in [ceil(x - 0.5 - r), floor(x - 0.5 + r)] (same in y), weighted by
// the filter at the offset from the pixel centre.  GaussianFilter::eval
// (src/filter/gaussian_filter.cpp:20-30): prod over x and y of
// max(0, exp(-alpha d^2) - exp(-alpha r^2)).  The box filter (no reference
// counterpart) weights 1 inside |d| <= r; at r <= 0.5 it keeps a sample in its
// own pixel.
struct FilmFilter {
    int type;
    float r, alpha;
    float eval(float dx, float dy) const {
        if (type == TR_FILTER_BOX) return (std::abs(dx) <= r && std::abs(dy) <= r) ? 1.0f : 0.0f;
        const float e = std::exp(-alpha * r * r);
        auto g = [&](float x) { return std::max(0.0f, std::exp(-alpha * x * x) - e); };
        return g(dx) * g(dy);
    }
    // Film::get_film_tile (render.hpp:79-88): pixel bounds [x0, x1) x [y0, y1)
    // padded by the radius and clipped to the film.
    void tile_bounds(int x0, int y0, int x1, int y1, int W, int H, int& px0, int& py0, int& px1, int& py1) const {
        px0 = std::max(0, int(std::ceil(float(x0) - 0.5f - r)));
        py0 = std::max(0, int(std::ceil(float(y0) - 0.5f - r)));
        px1 = std::min(W, int(std::floor(float(x1) - 0.5f + r)) + 1);
        py1 = std::min(H, int(std::floor(float(y1) - 0.5f + r)) + 1);
    }
    // Tile::add_sample into a tile buffer of (R, G, B, weight) doubles over
    // [px0, px1) x [py0, py1).
    void add_sample(float x, float y, float Lr, float Lg, float Lb, int px0, int py0, int px1, int py1,
                    double* tile) const {
        const int tw = px1 - px0;
        if (type == TR_FILTER_BOX && r <= 0.5f) {
            const int qx = int(std::floor(x)), qy = int(std::floor(y));
            if (qx < px0 || qy < py0 || qx >= px1 || qy >= py1) return;
            double* tp = &tile[(size_t(qy - py0) * tw + (qx - px0)) * 4];
            tp[0] += Lr; tp[1] += Lg; tp[2] += Lb; tp[3] += 1.0;
            return;
        }
        const float dxp = x - 0.5f, dyp = y - 0.5f;
        const int qx0 = std::max(px0, int(std::ceil(dxp - r))), qx1 = std::min(px1, int(std::floor(dxp + r)) + 1);
        const int qy0 = std::max(py0, int(std::ceil(dyp - r))), qy1 = std::min(py1, int(std::floor(dyp + r)) + 1);
        for (int qy = qy0; qy < qy1; ++qy)
            for (int qx = qx0; qx < qx1; ++qx) {
                const float w = eval(x - (float(qx) + 0.5f), y - (float(qy) + 0.5f));
                double* tp = &tile[(size_t(qy - py0) * tw + (qx - px0)) * 4];
                tp[0] += double(Lr) * w; tp[1] += double(Lg) * w; tp[2] += double(Lb) * w; tp[3] += w;
            }
    }
};

struct RenderCfg {
    int W, H, spp, min_depth, max_depth, nls;
    int filter; float fr, fa;
    uint64_t seed;
    int sample_offset;
    bool quirks;
    int nee;  // 0 all lights, 1 one by power
};

struct Renderer {
    const Scene& s;
    const RenderCfg& c;
    std::atomic<uint64_t> ext_rays{0}, shadow_rays{0};
    Renderer(const Scene& sc, const RenderCfg& cf) : s(sc), c(cf) {}

    const tr_medium* med(int id) const { return id < 0 ? nullptr : &s.media[id]; }
    const RefGrid* grid_of(const tr_medium* m) const {
        if (!m) return nullptr;
        const RefGrid& g = s.grids[size_t(m - s.media.data())];
        return g.d.empty() ? nullptr : &g;
    }
    // transmittance of a segment: closed form (homogeneous) or ratio tracking (grid)
    RGB<float> seg_tr(const tr_medium* m, const V3f& a, const V3f& b, Sampler& sp) const {
        const RefGrid* g = grid_of(m);
        return g ? RGB<float>(g->tr(a, b, sp)) : medium_tr(m, a, b);
    }
    int wo_medium_id(const Isect& is) const {   // intersection.hpp:36-41
        return dot(is.wo, is.geo.z) >= 0 ? s.med_out[is.prim] : s.med_in[is.prim];
    }
    int medium_id(const Isect& is, const V3f& d) const { return dot(d, is.geo.z) >= 0 ? s.med_out[is.prim] : s.med_in[is.prim]; }

    bool intersect(const V3f& o, const V3f& d, float tmin, float tmax, Isect* is) {
        ext_rays++;
        return intersect_ref(s, o, d, tmin, tmax, is);
    }
    bool occluded(const V3f& o, const V3f& d, float tmin, float tmax) {
        shadow_rays++;
        return occluded_ref(s, o, d, tmin, tmax);
    }

    // One light strategy evaluated at a shading point.  `li` indexes the light
    // list: [0, nA) area lights, nA = env, nA+1 = sun.  Returns radiance
    // contribution without the light-selection probability.
    template <class EvalF, class PdfF>
    RGB<float> sample_one_light(int li, const V3f& p, const V3f& ng, bool surface, int medium_towards_light_id_surface,
                                const tr_medium* cur_med, EvalF&& evalf, PdfF&& pdff, float sel_pdf, Sampler& sp) {
        const int nA = int(s.lights.size());
        float u0 = sp.get(), u1 = sp.get(); sp.get(); sp.get(); sp.get();   // Sample5
        if (li < nA) {
            const int tri = s.lights[li];
            LightSample ls = sample_area(s, tri, p, u0, u1);
            if (!ls.valid || ls.L.is_black() || !ls.pdf) return RGB<float>(0);
            ls.pdf *= sel_pdf;
            V3f to = ls.pos - p;
            float dist = float(to.length()) - kEps;
            if (dist <= kEps) return RGB<float>(0);
            V3f dir = to.normalize();
            if (surface) {
                if (occluded(p, dir, kEps, dist)) return RGB<float>(0);
            } else {
                // scene.visible(light_pos, p): Ray(p, dir, eps, |p-q| - eps)
                float dd = float((ls.pos - p).length());
                if (occluded(p, dir, kEps, dd - kEps)) return RGB<float>(0);
            }
            RGB<float> f = evalf(dir);
            if (f.is_black()) return RGB<float>(0);
            float bp = pdff(dir);
            const tr_medium* m = surface ? med(medium_towards_light_id_surface) : cur_med;
            RGB<float> tr = seg_tr(m, ls.pos, p, sp);
            RGB<float> val = tr * ls.L * f;
            if (surface) val *= std::abs(cosv(dir, ng));
            float w = power_heuristic(ls.pdf, bp);
            return val * (w / ls.pdf);
        }
        if (li == nA) {   // environment
            if (!surface && c.quirks) return RGB<float>(0);   // Q6
            LightSample ls = sample_env(s, p, u0, u1);
            if (!ls.valid || ls.L.is_black()) return RGB<float>(0);
            ls.pdf *= sel_pdf;
            V3f to = ls.pos - p;
            float dist = float(to.length()) - kEps;
            if (dist <= kEps) return RGB<float>(0);
            V3f dir = to.normalize();
            if (occluded(p, dir, kEps, dist)) return RGB<float>(0);
            RGB<float> f = evalf(dir);
            if (f.is_black()) return RGB<float>(0);
            float bp = pdff(dir);
            RGB<float> val = ls.L * f;
            if (surface) val *= std::abs(cosv(dir, ng));
            return val * (power_heuristic(ls.pdf, bp) / ls.pdf);
        }
        // sun (delta light, no MIS)
        V3f dir = s.sun_dir.normalize();
        if (occluded(p, dir, kEps, 2 * s.wradius)) return RGB<float>(0);
        RGB<float> f = evalf(dir);
        if (f.is_black()) return RGB<float>(0);
        RGB<float> val = s.sun_E * f;
        if (surface) val *= std::abs(cosv(dir, ng));
        return val / sel_pdf;
    }

    float light_select_pdf(int li) const { return c.nee == 1 ? s.lpmf[li] : 1.0f; }

    // Direct lighting at a vertex: all lights, or one by power.
    template <class EvalF, class PdfF>
    RGB<float> direct(const V3f& p, const V3f& ng, bool surface, const Isect* is, const tr_medium* cur_med,
                      EvalF&& evalf, PdfF&& pdff, Sampler& sp) {
        RGB<float> acc;
        if (c.nee == 0) {
            const int nA = int(s.lights.size());
            for (int li = 0; li < nA; ++li)
                acc += sample_light_generic(li, p, ng, surface, is, cur_med, evalf, pdff, 1.0f, sp);
            if (s.has_env) acc += sample_light_generic(nA, p, ng, surface, is, cur_med, evalf, pdff, 1.0f, sp);
            if (s.has_sun) acc += sample_light_generic(nA + 1, p, ng, surface, is, cur_med, evalf, pdff, 1.0f, sp);
        } else if (s.n_light_entries > 0) {
            float u = sp.get();
            int li = int(std::upper_bound(s.lcdf.begin(), s.lcdf.end(), u) - s.lcdf.begin());
            li = std::min(li, s.n_light_entries - 1);
            while (li > 0 && s.lpmf[li] == 0) --li;
            acc += sample_light_generic(li, p, ng, surface, is, cur_med, evalf, pdff, s.lpmf[li], sp);
        }
        return acc;
    }
    template <class EvalF, class PdfF>
    RGB<float> sample_light_generic(int li, const V3f& p, const V3f& ng, bool surface, const Isect* is, const tr_medium* cur_med,
                                    EvalF&& evalf, PdfF&& pdff, float sel, Sampler& sp) {
        // Surface + area light: the medium between the point and the light is
        // the one on the light's side of the surface (direct_illumination.cpp:55),
        // known only once the light point is sampled.
        if (surface && li < int(s.lights.size())) {
            float u0 = sp.get(), u1 = sp.get(); sp.get(); sp.get(); sp.get();
            const int tri = s.lights[li];
            LightSample ls = sample_area(s, tri, p, u0, u1);
            if (!ls.valid || ls.L.is_black() || !ls.pdf) return RGB<float>(0);
            ls.pdf *= sel;
            V3f to = ls.pos - p;
            float dist = float(to.length()) - kEps;
            if (dist <= kEps) return RGB<float>(0);
            V3f dir = to.normalize();
            if (occluded(p, dir, kEps, dist)) return RGB<float>(0);
            RGB<float> f = evalf(dir);
            if (f.is_black()) return RGB<float>(0);
            float bp = pdff(dir);
            RGB<float> tr = seg_tr(med(medium_id(*is, dir)), ls.pos, p, sp);
            RGB<float> val = tr * ls.L * f * std::abs(cosv(dir, ng));
            return val * (power_heuristic(ls.pdf, bp) / ls.pdf);
        }
        return sample_one_light(li, p, ng, surface, -1, cur_med, evalf, pdff, sel, sp);
    }

    // light pdf of the strategy that would have sampled `dir` towards a hit on
    // area light triangle `tri` / the environment (including selection).
    float area_light_pdf(int tri, const V3f& ref, const V3f& pos) const {
        float p = pdf_area(s, tri, ref, pos);
        if (c.nee == 1) {
            int li = light_index_of_tri(tri);
            p *= li >= 0 ? s.lpmf[li] : 0.0f;
        }
        return p;
    }
    float env_light_pdf(const V3f& dir) const {
        float p = pdf_env(s, dir, c.quirks);
        if (c.nee == 1) p *= s.lpmf[s.lights.size()];
        return p;
    }
    std::vector<int> tri_to_light;
    int light_index_of_tri(int tri) const { return tri_to_light[tri]; }

    // direct_illumination.cpp:144-198 (surface) / :200-255 (medium)
    template <class SampleF>
    RGB<float> sample_bsdf_mis(const V3f& origin, const V3f& ng, bool surface, const Isect* is, const tr_medium* cur_med,
                               SampleF&& samplef, Sampler& sp) {
        float a = sp.get(), b = sp.get(), cc = sp.get();
        BSDFSample<float> bs = samplef(a, b, cc);
        if (!bs.valid()) return RGB<float>(0);
        V3f wi = bs.wi.normalize();
        V3f o = surface ? is->eps_offset(wi) : origin;
        Isect t;
        bool hit = intersect(o, wi, 0, std::numeric_limits<float>::max(), &t);
        if (!hit) {
            if (!s.has_env) return RGB<float>(0);
            RGB<float> Le = s.env_radiance(wi);
            RGB<float> f = Le * bs.f;
            if (surface) f *= std::abs(cosv(ng, wi));
            float lp = env_light_pdf(wi);
            return f * (power_heuristic(bs.pdf, lp) / bs.pdf);
        }
        RGB<float> Le = s.tri_emission(t.prim);
        if (Le.is_black()) return RGB<float>(0);
        if (!(dot(t.geo.z, t.wo) > 0)) return RGB<float>(0);   // one-sided light_emit
        RGB<float> tr;
        if (surface) {
            const tr_medium* m = med(medium_id(*is, bs.wi));
            tr = c.quirks ? seg_tr(m, o, is->pos, sp) : seg_tr(m, o, t.pos, sp);   // Q5
        } else {
            tr = seg_tr(cur_med, o, t.pos, sp);
        }
        RGB<float> f = tr * Le * bs.f;
        if (surface) f *= std::abs(cosv(ng, wi));
        float lp = area_light_pdf(t.prim, o, t.pos);
        return f * (power_heuristic(bs.pdf, lp) / bs.pdf);
    }

    // path_trace_renderer.cpp:27-210
    RGB<float> Li(V3f ro, V3f rd, Sampler& sp) {
        RGB<float> coef(1), L(0);
        const V3f cam_o = ro;
        V3f o = ro, d = rd.normalize();
        bool specular = false;
        int scattering = 0, s_depth = 0;
        for (int depth = 0; depth < c.max_depth; ++depth) {
            if (depth > c.min_depth) {
                if (sp.get() > 0.9f) break;
                coef = coef / 0.9f;
            }
            if (!coef.is_finite()) break;
            Isect is;
            bool found = intersect(o, d, 0, std::numeric_limits<float>::max(), &is);
            // Q23: the reference adds the emission of the surface it reached
            // before it samples the medium of the segment (path_trace_renderer.cpp:66-80
            // before :84-122), so an emitter seen through a medium is not
            // attenuated.  Fixed: the emission is added after the medium
            // decision, only when the segment reaches the surface.
            RGB<float> Le_hit(0);
            if (depth == 0 || specular) {
                if (found) {
                    RGB<float> Le = s.tri_emission(is.prim);
                    if (!Le.is_black() && dot(is.geo.z, -d) > 0) Le_hit = Le;
                } else if (s.has_env) {
                    L += coef * s.env_radiance(d);
                }
            }
            if (c.quirks) L += coef * Le_hit;
            if (!found) break;
            const int mid = wo_medium_id(is);
            const tr_medium* m = med(mid);
            const int max_sc = m ? m->max_scattering_count : std::numeric_limits<int>::max();
            if (scattering < max_sc) {
                MediumSample ms = grid_of(m) ? grid_of(m)->sample(c.quirks ? cam_o : o, is.pos, sp, scattering > 0)
                                              : medium_sample(m, c.quirks ? cam_o : o, is.pos, sp, c.quirks);   // Q1
                coef *= ms.w;
                if (ms.scattered) {
                    scattering++;
                    HGRef<float> hg{std::clamp<float>(grid_of(m) ? ms.g : m->g, -1 + kEps, 1 - kEps)};
                    const V3f wo = -((is.pos - (c.quirks ? cam_o : o)).normalize());
                    auto evalf = [&](const V3f& wi) { return RGB<float>(hg.eval(wi, wo)); };
                    auto pdff = [&](const V3f& wi) { return hg.eval(wi, wo); };
                    auto samplef = [&](float a, float b, float) { return hg.sample(wo, a, b); };
                    RGB<float> di;
                    for (int i = 0; i < c.nls; ++i) {
                        di += coef * direct(ms.pos, V3f(0, 0, 1), false, nullptr, m, evalf, pdff, sp);
                        di += coef * sample_bsdf_mis(ms.pos, V3f(0, 0, 1), false, nullptr, m, samplef, sp);
                    }
                    L += di * (1.0f / c.nls);
                    float a = sp.get(), b = sp.get(); sp.get();
                    BSDFSample<float> ps = hg.sample(wo, a, b);
                    if (!ps.valid()) break;
                    o = ms.pos;
                    d = ps.wi.normalize();
                    coef *= ps.f / ps.pdf;
                    // Q27: the reference keeps the specular flag of the last surface
                    // vertex across a medium scatter (path_trace_renderer.cpp:94-117
                    // continue without resetting it), so after glass -> scatter the
                    // next hit's emission is added unweighted on top of the scatter
                    // vertex's light sample and MIS ray.  Fixed: a scatter vertex is
                    // not specular.
                    if (!c.quirks) specular = false;
                    continue;
                }
            } else {
                coef *= seg_tr(m, o, is.pos, sp);
                scattering = 0;
            }
            if (!c.quirks) L += coef * Le_hit;
            SurfBSDF bsdf;
            make_bsdf(s, is, c.quirks, bsdf);
            float a = sp.get(), b = sp.get(), cc = sp.get();
            BSDFSample<float> bs = bsdf.sample(is.wo, a, b, cc);
            // Q26: the reference ends the path before its direct-light estimate
            // when the continuation sample fails (path_trace_renderer.cpp:128-131
            // before :141-153), losing that vertex's direct light.  Fixed: the
            // direct light is estimated first and the path ends after it.
            const bool cont_ok = !(bs.f.is_black() || bs.pdf < kEps);
            if (!cont_ok && c.quirks) break;
            // path_trace_renderer.cpp:133-138: a specular (glass) bounce does not count
            // towards the depth, up to max_specular_depth = 20 of them
            specular = cont_ok && bs.delta;
            if (specular && s_depth < 20) {
                ++s_depth;
                --depth;
            }
            auto evalf = [&](const V3f& wi) { return bsdf.eval(wi, is.wo); };
            auto pdff = [&](const V3f& wi) { return bsdf.pdf(wi, is.wo); };
            auto samplef = [&](float x, float y, float z) { return bsdf.sample(is.wo, x, y, z); };
            RGB<float> di;
            if (!bs.delta) {   // path_trace_renderer.cpp:141: no light sample at a specular vertex
                for (int i = 0; i < c.nls; ++i) {
                    di += coef * direct(is.pos, is.geo.z, true, &is, nullptr, evalf, pdff, sp);
                    di += coef * sample_bsdf_mis(is.pos, is.geo.z, true, &is, nullptr, samplef, sp);
                }
            }
            L += di * (1.0f / c.nls);
            if (!cont_ok) break;
            coef *= bs.f * (std::abs(cosv(is.geo.z, bs.wi)) / bs.pdf);
            if (!coef.is_finite()) break;
            V3f nd = bs.wi.normalize();
            o = is.eps_offset(bs.wi);
            d = nd;
        }
        if (!L.is_finite()) return RGB<float>(0);
        return L;
    }

    // thin_lens_camera.cpp:40-59 with look_at (transform.cpp:96-124)
    void camera_ray(float fx, float fy, float lu, float lv, V3f& o, V3f& d) const {
        const tr_camera& cm = s.cam;
        V3f pos(cm.pos[0], cm.pos[1], cm.pos[2]), tgt(cm.target[0], cm.target[1], cm.target[2]), up(cm.up[0], cm.up[1], cm.up[2]);
        V3f dir = (tgt - pos).normalize();
        V3f right = cross(up.normalize(), dir).normalize();
        V3f nup = cross(dir, right);
        float Hf = 2 * cm.focal_distance * std::tan(cm.fov_y * 0.5f);
        float Wf = cm.aspect * Hf;
        V3f film((0.5f - fx) * Wf, (0.5f - fy) * Hf, cm.focal_distance);
        float dx, dy;
        concentric_sample_disk(lu, lv, dx, dy);
        V3f lens(cm.lens_radius * dx, cm.lens_radius * dy, 0);
        V3f ld = (film - lens).normalize();
        o = pos + right * lens.x + nup * lens.y;
        d = right * ld.x + nup * ld.y + dir * ld.z;
    }

    void render(double* sums, int threads) {
        const int W = c.W, H = c.H;
        const int TS = 16;
        const int tx = (W + TS - 1) / TS, ty = (H + TS - 1) / TS;
        std::atomic<int> next{0};
        std::mutex mtx;
        auto worker = [&]() {
            std::vector<double> tile;
            for (;;) {
                int t = next++;
                if (t >= tx * ty) break;
                int x0 = (t % tx) * TS, y0 = (t / tx) * TS;
                int x1 = std::min(W, x0 + TS), y1 = std::min(H, y0 + TS);
                // tile padded by the filter radius (Film::get_film_tile, render.hpp:79-88)
                const FilmFilter flt{c.filter, c.fr, c.fa};
                int px0 = 0, py0 = 0, px1 = 0, py1 = 0;
                flt.tile_bounds(x0, y0, x1, y1, W, H, px0, py0, px1, py1);
                int tw = px1 - px0;
                tile.assign(size_t(tw) * (py1 - py0) * 4, 0.0);
                for (int y = y0; y < y1; ++y)
                    for (int x = x0; x < x1; ++x) {
                        const uint64_t pix = uint64_t(y) * W + x;
                        for (int i = 0; i < c.spp; ++i) {
                            const uint64_t sidx = uint64_t(c.sample_offset) + i;
                            Sampler sp;
                            sp.rng.seed(mix64(c.seed ^ mix64(pix * 0x9E3779B97F4A7C15ULL + sidx)), pix);
                            float fu = sp.get(), fv = sp.get(), lu = sp.get(), lv = sp.get();
                            float pxf = x + fu, pyf = y + fv;
                            V3f o, d;
                            camera_ray(pxf / W, pyf / H, lu, lv, o, d);
                            RGB<float> L = Li(o, d, sp);
                            if (!L.is_finite()) continue;
                            flt.add_sample(pxf, pyf, L.r, L.g, L.b, px0, py0, px1, py1, tile.data());
                        }
                    }
                std::lock_guard<std::mutex> lk(mtx);   // render.hpp:90-100
                for (int y = py0; y < py1; ++y)
                    for (int x = px0; x < px1; ++x) {
                        const double* tp = &tile[(size_t(y - py0) * tw + (x - px0)) * 4];
                        double* fp = &sums[(size_t(y) * W + x) * 4];
                        for (int k = 0; k < 4; ++k) fp[k] += tp[k];
                    }
            }
        };
        std::vector<std::thread> pool;
        for (int i = 0; i < threads; ++i) pool.emplace_back(worker);
        for (auto& t : pool) t.join();
    }
};

static Scene* build_scene(const tr_scene_desc* d, int max_leaf) {
    Scene* s = new Scene();
    const int64_t nv = d->num_vertices, nt = d->num_triangles;
    s->P.resize(nv);
    for (int64_t i = 0; i < nv; ++i) s->P[i] = V3f(d->positions[3 * i], d->positions[3 * i + 1], d->positions[3 * i + 2]);
    if (d->normals) {
        s->N.resize(nv);
        for (int64_t i = 0; i < nv; ++i) s->N[i] = V3f(d->normals[3 * i], d->normals[3 * i + 1], d->normals[3 * i + 2]);
    }
    if (d->uvs) s->UV.assign(d->uvs, d->uvs + 2 * nv);
    for (int i = 0; i < d->num_textures; ++i) {
        const tr_texture& t = d->textures[i];
        RefTex rt;
        rt.w = t.width;
        rt.h = t.height;
        const size_t n = size_t(t.width) * size_t(t.height) * 3;
        if (t.rgb8) rt.rgb8.assign(t.rgb8, t.rgb8 + n);
        else rt.rgbf.assign(t.rgbf, t.rgbf + n);
        s->tex.push_back(std::move(rt));
    }
    s->idx.assign(d->indices, d->indices + 3 * nt);
    s->mat.assign(d->material_ids, d->material_ids + nt);
    if (d->emission) s->emis.assign(d->emission, d->emission + 3 * nt);
    s->med_in.assign(nt, -1);
    s->med_out.assign(nt, -1);
    if (d->medium_inside) s->med_in.assign(d->medium_inside, d->medium_inside + nt);
    if (d->medium_outside) s->med_out.assign(d->medium_outside, d->medium_outside + nt);
    s->mats.assign(d->materials, d->materials + d->num_materials);
    if (d->num_media) s->media.assign(d->media, d->media + d->num_media);
    s->grids.assign(s->media.size(), RefGrid{});
    for (size_t i = 0; i < s->media.size(); ++i)
        if (s->media[i].density) {
            s->grids[i].init(s->media[i]);
            s->media[i].density = nullptr;   // the grid copies above are used from here on
            s->media[i].albedo_grid = nullptr;
            s->media[i].g_grid = nullptr;
        }
    std::vector<Bounds3> pb(nt);
    for (int64_t t = 0; t < nt; ++t) {
        Bounds3 b;
        b.grow(s->vtx(int(t), 0)); b.grow(s->vtx(int(t), 1)); b.grow(s->vtx(int(t), 2));
        pb[t] = b;
        s->world.grow(b);
        if (!s->tri_emission(int(t)).is_black()) s->lights.push_back(int(t));
    }
    s->bvh.build(pb, max_leaf);
    s->has_env = d->has_env != 0;
    s->env = RGB<float>(d->env_radiance[0], d->env_radiance[1], d->env_radiance[2]);
    if (d->has_env && d->env_width > 0)
        s->envmap = std::make_shared<RefEnvMap>(d->env_width, d->env_height, d->env_pixels, d->env_world_to_light);
    s->has_sun = d->has_sun != 0;
    s->sun_dir = V3f(d->sun_direction[0], d->sun_direction[1], d->sun_direction[2]);
    s->sun_E = RGB<float>(d->sun_irradiance[0], d->sun_irradiance[1], d->sun_irradiance[2]);
    s->cam = d->camera;
    // Bounds3::bounding_sphere (geometry.hpp:752-755)
    s->wcenter = (s->world.lo + s->world.hi) * 0.5f;
    s->wradius = float((s->world.hi - s->wcenter).length());
    // power distribution: DiffuseLight pi*L*A, IBL (L*4pi)*R^2, sun E*pi*R^2
    const int nA = int(s->lights.size());
    s->lpmf.assign(nA + 2, 0.0f);
    for (int i = 0; i < nA; ++i) s->lpmf[i] = kPi<float> * s->tri_emission(s->lights[i]).lum() * s->tri_area(s->lights[i]);
    if (s->has_env && s->envmap) s->lpmf[nA] = s->envmap->avg_radiance.lum() * s->wradius * s->wradius;   // IBL::power
    else if (s->has_env) s->lpmf[nA] = s->env.lum() * 4 * kPi<float> * s->wradius * s->wradius;
    if (s->has_sun) s->lpmf[nA + 1] = s->sun_E.lum() * kPi<float> * s->wradius * s->wradius;
    double tot = 0;
    for (float p : s->lpmf) tot += p;
    s->lcdf.assign(nA + 2, 1.0f);
    double acc = 0;
    for (int i = 0; i < nA + 2; ++i) {
        s->lpmf[i] = tot > 0 ? float(s->lpmf[i] / tot) : 0.0f;
        acc += s->lpmf[i];
        s->lcdf[i] = float(acc);
    }
    s->n_light_entries = nA + 2;
    return s;
}

}  // namespace refcpu

using namespace refcpu;

extern "C" {

// ---- scene -------------------------------------------------------------
void* ref_scene_create(const tr_scene_desc* d, int max_leaf) { return build_scene(d, max_leaf); }
void ref_scene_destroy(void* h) { delete static_cast<Scene*>(h); }
int64_t ref_bvh_node_count(void* h) { return int64_t(static_cast<Scene*>(h)->bvh.nodes.size()); }
// Copies the flattened 32-byte nodes and the reordered primitive indices.
void ref_bvh_export(void* h, void* nodes_out, int32_t* prims_out) {
    Scene* s = static_cast<Scene*>(h);
    std::memcpy(nodes_out, s->bvh.nodes.data(), s->bvh.nodes.size() * sizeof(RefNode));
    std::memcpy(prims_out, s->bvh.prims.data(), s->bvh.prims.size() * sizeof(int));
}
void ref_light_pmf(void* h, float* pmf_out) {
    Scene* s = static_cast<Scene*>(h);
    std::memcpy(pmf_out, s->lpmf.data(), s->lpmf.size() * sizeof(float));
}
// Distribution1D (src/utility/distribution.hpp:12-29) over the light powers
// and its sample_discrete (:49-67): the last entry of the running-sum table
// cdf[0..n] (cdf[0] = 0, cdf[n] = 1) that is <= u, clamped to [0, n - 1];
// pmf = func[i] / (func_int * n).  Light i < num_area is the scene's i-th
// emissive triangle in triangle order, then the environment, then the sun.
void ref_light_pick(void* h, const float* u, int64_t n, int32_t* idx, float* pmf) {
    Scene* s = static_cast<Scene*>(h);
    const int m = int(s->lpmf.size());
    std::vector<float> func(s->lpmf), cdf(size_t(m) + 1, 0.0f);
    for (int i = 0; i < m; ++i) cdf[size_t(i) + 1] = cdf[size_t(i)] + func[size_t(i)] / float(m);
    const float func_int = cdf[size_t(m)];
    for (int i = 1; i <= m; ++i) cdf[size_t(i)] = func_int > 0 ? cdf[size_t(i)] / func_int : float(i) / float(m);
    for (int64_t k = 0; k < n; ++k) {
        int off = m;   // last cdf entry <= u
        while (off > 0 && !(cdf[size_t(off)] <= u[k])) --off;
        off = std::min(std::max(off, 0), m - 1);
        idx[k] = off;
        pmf[k] = func_int > 0 ? func[size_t(off)] / (func_int * float(m)) : 0.0f;
    }
}

// Tile::add_sample over a whole W x H film (one tile covering it, so only the
// film bounds clip): samples at film positions pos (n*2, pixel units) with
// radiance L (n*3), in order, accumulated onto sums (W*H*4 doubles).
void ref_film_splat(int W, int H, int filter, float radius, float alpha, const float* pos, const float* L, int64_t n,
                    double* sums) {
    const FilmFilter f{filter, radius, alpha};
    for (int64_t k = 0; k < n; ++k)
        f.add_sample(pos[2 * k], pos[2 * k + 1], L[3 * k], L[3 * k + 1], L[3 * k + 2], 0, 0, W, H, sums);
}

// ---- ray queries --------------------------------------------------------
// rays: n*8 (ox,oy,oz,tmin,dx,dy,dz,tmax); hits: n*4 (t, prim bits, b1, b2)
static void par_for(int64_t n, int threads, const std::function<void(int64_t, int64_t)>& fn) {
    if (threads <= 1 || n < 1024) { fn(0, n); return; }
    std::vector<std::thread> pool;
    int64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        int64_t a = t * chunk, b = std::min(n, a + chunk);
        if (a < b) pool.emplace_back(fn, a, b);
    }
    for (auto& th : pool) th.join();
}
static inline void put_hit(float* h, float t, int prim, float b1, float b2) {
    h[0] = t; std::memcpy(&h[1], &prim, 4); h[2] = b1; h[3] = b2;
}

// Reference semantics: BVHAccel::intersect_p + Moller-Trumbore (normalised d).
void ref_trace_closest_mt(void* hs, const float* rays, float* hits, int64_t n, int threads) {
    Scene* s = static_cast<Scene*>(hs);
    par_for(n, threads, [&](int64_t a, int64_t b) {
        for (int64_t i = a; i < b; ++i) {
            const float* r = rays + 8 * i;
            V3f o(r[0], r[1], r[2]), d = V3f(r[4], r[5], r[6]).normalize();
            float tm = r[7];
            int best = -1; MTHit bh{0, 0, 0};
            s->bvh.traverse(o, d, tm, [&](int p, float& tcur) {
                MTHit h;
                if (!mt_intersect(o, d, r[3], tcur, s->vtx(p, 0), s->vtx(p, 1), s->vtx(p, 2), &h)) return false;
                tcur = h.t; best = p; bh = h;
                return true;
            }, false);
            put_hit(hits + 4 * i, best >= 0 ? bh.t : 0.0f, best, bh.alpha, bh.beta);
        }
    });
}

// Brute force over all triangles with the watertight test; lowest id wins ties.
void ref_trace_closest_wt_brute(void* hs, const float* rays, float* hits, int64_t n, int threads) {
    Scene* s = static_cast<Scene*>(hs);
    const int nt = int(s->mat.size());
    par_for(n, threads, [&](int64_t a, int64_t b) {
        for (int64_t i = a; i < b; ++i) {
            const float* r = rays + 8 * i;
            V3f o(r[0], r[1], r[2]), d(r[4], r[5], r[6]);
            WTRay wr = wt_setup(o, d);
            float best_t = r[7]; int best = -1; WTHit bh{0, 0, 0};
            for (int p = 0; p < nt; ++p) {
                WTHit h;
                if (!wt_intersect(wr, r[3], best_t, s->vtx(p, 0), s->vtx(p, 1), s->vtx(p, 2), &h)) continue;
                if (h.t < best_t || best < 0 || p < best) { best_t = h.t; best = p; bh = h; }
            }
            put_hit(hits + 4 * i, best >= 0 ? bh.t : 0.0f, best, bh.b1, bh.b2);
        }
    });
}

// Same result as the brute force, through the reference BVH with a
// conservative slab test (Ize, "Robust BVH Ray Traversal", JCGT 2013:
// interval widened by 1 +/- 2*gamma(3)) so no box holding a hit is culled.
static inline bool slab_conservative(const Bounds3& b, const V3f& o, const V3f& inv, float tmin, float tmax, float& tnear) {
    const float g = 2.0f * (3 * 5.96046448e-08f) / (1 - 3 * 5.96046448e-08f);
    float t0 = tmin, t1 = tmax;
    for (int k = 0; k < 3; ++k) {
        float tn = (b.lo[k] - o[k]) * inv[k];
        float tf = (b.hi[k] - o[k]) * inv[k];
        if (tn > tf) std::swap(tn, tf);
        tn *= 1 - g;
        tf *= 1 + g;
        t0 = tn > t0 ? tn : t0;
        t1 = tf < t1 ? tf : t1;
        if (t0 > t1) return false;
    }
    tnear = t0;
    return true;
}
static inline V3f safe_inv(const V3f& d) {
    auto f = [](float x) { float ax = std::fabs(x) < 1e-30f ? std::copysign(1e-30f, x) : x; return 1.0f / ax; };
    return V3f(f(d.x), f(d.y), f(d.z));
}
void ref_trace_closest_wt_bvh(void* hs, const float* rays, float* hits, int64_t n, int threads) {
    Scene* s = static_cast<Scene*>(hs);
    par_for(n, threads, [&](int64_t a, int64_t b) {
        std::vector<int> stack;
        for (int64_t i = a; i < b; ++i) {
            const float* r = rays + 8 * i;
            V3f o(r[0], r[1], r[2]), d(r[4], r[5], r[6]);
            WTRay wr = wt_setup(o, d);
            V3f inv = safe_inv(d);
            float best_t = r[7]; int best = -1; WTHit bh{0, 0, 0};
            stack.clear();
            if (!s->bvh.nodes.empty()) stack.push_back(0);
            while (!stack.empty()) {
                int ni = stack.back(); stack.pop_back();
                const RefNode& nd = s->bvh.nodes[ni];
                float tn;
                if (!slab_conservative(nd.b, o, inv, r[3], best_t, tn)) continue;
                if (nd.count > 0) {
                    for (int k = 0; k < nd.count; ++k) {
                        int p = s->bvh.prims[nd.offset + k];
                        WTHit h;
                        if (!wt_intersect(wr, r[3], best_t, s->vtx(p, 0), s->vtx(p, 1), s->vtx(p, 2), &h)) continue;
                        if (h.t < best_t || best < 0 || p < best) { best_t = h.t; best = p; bh = h; }
                    }
                } else {
                    stack.push_back(nd.offset);
                    stack.push_back(ni + 1);
                }
            }
            put_hit(hits + 4 * i, best >= 0 ? bh.t : 0.0f, best, bh.b1, bh.b2);
        }
    });
}
void ref_trace_any_wt(void* hs, const float* rays, int32_t* occ, int64_t n, int threads) {
    Scene* s = static_cast<Scene*>(hs);
    const int nt = int(s->mat.size());
    par_for(n, threads, [&](int64_t a, int64_t b) {
        std::vector<int> stack;
        for (int64_t i = a; i < b; ++i) {
            const float* r = rays + 8 * i;
            V3f o(r[0], r[1], r[2]), d(r[4], r[5], r[6]);
            WTRay wr = wt_setup(o, d);
            V3f inv = safe_inv(d);
            int hit = 0;
            stack.clear();
            if (!s->bvh.nodes.empty()) stack.push_back(0);
            while (!stack.empty() && !hit) {
                int ni = stack.back(); stack.pop_back();
                const RefNode& nd = s->bvh.nodes[ni];
                float tn;
                if (!slab_conservative(nd.b, o, inv, r[3], r[7], tn)) continue;
                if (nd.count > 0) {
                    for (int k = 0; k < nd.count && !hit; ++k) {
                        int p = s->bvh.prims[nd.offset + k];
                        if (wt_intersect(wr, r[3], r[7], s->vtx(p, 0), s->vtx(p, 1), s->vtx(p, 2), nullptr)) hit = 1;
                    }
                } else {
                    stack.push_back(nd.offset);
                    stack.push_back(ni + 1);
                }
            }
            occ[i] = hit;
            (void)nt;
        }
    });
}
// Single-triangle tests, for known-answer fixtures.
int ref_mt_single(const float* ray, const float* tri, float* out) {
    MTHit h;
    V3f o(ray[0], ray[1], ray[2]), d = V3f(ray[4], ray[5], ray[6]).normalize();
    if (!mt_intersect(o, d, ray[3], ray[7], V3f(tri[0], tri[1], tri[2]), V3f(tri[3], tri[4], tri[5]), V3f(tri[6], tri[7], tri[8]), &h)) return 0;
    out[0] = h.t; out[1] = h.alpha; out[2] = h.beta;
    return 1;
}
int ref_wt_single(const float* ray, const float* tri, float* out) {
    WTHit h;
    WTRay wr = wt_setup(V3f(ray[0], ray[1], ray[2]), V3f(ray[4], ray[5], ray[6]));
    if (!wt_intersect(wr, ray[3], ray[7], V3f(tri[0], tri[1], tri[2]), V3f(tri[3], tri[4], tri[5]), V3f(tri[6], tri[7], tri[8]), &h)) return 0;
    out[0] = h.t; out[1] = h.b1; out[2] = h.b2;
    return 1;
}

// ---- BSDF batches ---------------------------------------------------------
// Vectors in the shading frame (shading normal +z); ng per entry or NULL.
// precision: 0 float (reference), 1 double (FP64 twin).  fixed: quirk switch.
void ref_bsdf_eval(const tr_material* mats, const int32_t* ids, const float* wi, const float* wo, const float* ng,
                   int64_t n, int precision, int fixed, double* f, double* pdf) {
    for (int64_t i = 0; i < n; ++i) {
        const tr_material& m = mats[ids[i]];
        auto run = [&](auto zero) {
            using T = decltype(zero);
            V3<T> vi(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]), vo(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
            V3<T> g = ng ? V3<T>(ng[3 * i], ng[3 * i + 1], ng[3 * i + 2]) : V3<T>(0, 0, 1);
            Coord<T> shd;
            Coord<T> geo = Coord<T>::from_z(g);
            RGB<T> fv; T pv;
            if (m.type == TR_MAT_DISNEY) {
                DisneyRef<T> b(geo, shd, to_params(m), fixed != 0);
                fv = b.eval(vi, vo); pv = b.pdf(vi, vo);
            } else if (m.type == TR_MAT_PHONG) {
                PhongRef<T> b(geo, shd, phong_kd<T>(m), phong_ks<T>(m), T(m.shininess), fixed != 0);
                fv = b.eval(vi, vo); pv = b.pdf(vi, vo);
            } else if (m.type == TR_MAT_GLASS) {
                fv = RGB<T>(0); pv = 0;
            } else {
                LambertRef<T> b(geo, shd, RGB<T>(m.base_color[0], m.base_color[1], m.base_color[2]), fixed != 0);
                fv = b.eval(vi, vo); pv = b.pdf(vi, vo);
            }
            f[3 * i] = fv.r; f[3 * i + 1] = fv.g; f[3 * i + 2] = fv.b; pdf[i] = pv;
        };
        if (precision) run(double(0)); else run(float(0));
    }
}
void ref_bsdf_sample(const tr_material* mats, const int32_t* ids, const float* wo, const float* ng, const float* u,
                     int64_t n, int precision, int fixed, double* wi, double* f, double* pdf) {
    for (int64_t i = 0; i < n; ++i) {
        const tr_material& m = mats[ids[i]];
        auto run = [&](auto zero) {
            using T = decltype(zero);
            V3<T> vo(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
            V3<T> g = ng ? V3<T>(ng[3 * i], ng[3 * i + 1], ng[3 * i + 2]) : V3<T>(0, 0, 1);
            Coord<T> shd;
            Coord<T> geo = Coord<T>::from_z(g);
            BSDFSample<T> s;
            if (m.type == TR_MAT_DISNEY) {
                DisneyRef<T> b(geo, shd, to_params(m), fixed != 0);
                s = b.sample(vo, T(u[3 * i]), T(u[3 * i + 1]), T(u[3 * i + 2]));
            } else if (m.type == TR_MAT_PHONG) {
                PhongRef<T> b(geo, shd, phong_kd<T>(m), phong_ks<T>(m), T(m.shininess), fixed != 0);
                s = b.sample(vo, T(u[3 * i]), T(u[3 * i + 1]), T(u[3 * i + 2]));
            } else if (m.type == TR_MAT_GLASS) {
                GlassRef<T> b(geo, shd, phong_kd<T>(m), phong_ks<T>(m), T(m.ior));
                s = b.sample(vo, T(u[3 * i]), T(u[3 * i + 1]), T(u[3 * i + 2]));
            } else {
                LambertRef<T> b(geo, shd, RGB<T>(m.base_color[0], m.base_color[1], m.base_color[2]), fixed != 0);
                s = b.sample(vo, T(u[3 * i]), T(u[3 * i + 1]), T(u[3 * i + 2]));
            }
            wi[3 * i] = s.wi.x; wi[3 * i + 1] = s.wi.y; wi[3 * i + 2] = s.wi.z;
            f[3 * i] = s.f.r; f[3 * i + 1] = s.f.g; f[3 * i + 2] = s.f.b;
            pdf[i] = s.valid() ? s.pdf : 0;
        };
        if (precision) run(double(0)); else run(float(0));
    }
}
// HG phase: value for cos between wi and wo, and sampling.
// One 8-bit texture on its own: lookups at (u, v) pairs (RGB out).
void ref_tex_eval8(int w, int h, const uint8_t* rgb8, const float* uv, int64_t n, int quirks, float* out) {
    RefTex t;
    t.w = w;
    t.h = h;
    t.rgb8.assign(rgb8, rgb8 + size_t(w) * size_t(h) * 3);
    for (int64_t i = 0; i < n; ++i) {
        const RGB<float> c = t.eval(uv[2 * i], uv[2 * i + 1], quirks != 0);
        out[3 * i] = c.r; out[3 * i + 1] = c.g; out[3 * i + 2] = c.b;
    }
}

// A heterogeneous medium on its own: densities at world points, and over
// `trials` segments a -> b the mean ratio-tracking transmittance, the fraction
// of delta-tracking samples that scatter and their mean weight (red channel).
void ref_grid_density(const tr_medium* m, const float* pts, int64_t n, float* out) {
    RefGrid g;
    g.init(*m);
    for (int64_t i = 0; i < n; ++i) out[i] = g.density(V3f(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
}
// Albedo (RGB) and asymmetry of the medium at world points: its grids or constants.
void ref_grid_albedo_g(const tr_medium* m, const float* pts, int64_t n, float* albedo, float* g) {
    RefGrid r;
    r.init(*m);
    for (int64_t i = 0; i < n; ++i) {
        const V3f p(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        const RGB<float> a = r.albedo_at(p);
        albedo[3 * i] = a.r; albedo[3 * i + 1] = a.g; albedo[3 * i + 2] = a.b;
        g[i] = r.g_at(p);
    }
}
void ref_grid_stats(const tr_medium* m, const float* a, const float* b, int64_t trials, uint64_t seed, int indirect,
                    double* out) {
    RefGrid g;
    g.init(*m);
    Sampler sp;
    sp.rng.seed(seed, 7);
    const V3f pa(a[0], a[1], a[2]), pb(b[0], b[1], b[2]);
    double tr = 0, sc = 0, w = 0;
    for (int64_t i = 0; i < trials; ++i) {
        tr += g.tr(pa, pb, sp);
        const MediumSample ms = g.sample(pa, pb, sp, indirect != 0);
        if (ms.scattered) { sc += 1; w += ms.w.r; }
    }
    out[0] = tr / double(trials);
    out[1] = sc / double(trials);
    out[2] = sc > 0 ? w / sc : 0.0;
}

// Image environment on its own: IBL sample_li / light_emit / pdf.
void* ref_env_create(int w, int h, const float* texels, const float* world_to_light) {
    return new RefEnvMap(w, h, texels, world_to_light);
}
void ref_env_destroy(void* e) { delete static_cast<RefEnvMap*>(e); }
void ref_env_sample(void* e, const float* u, int64_t n, float* wi, float* pdf, float* L) {
    const RefEnvMap* m = static_cast<RefEnvMap*>(e);
    for (int64_t i = 0; i < n; ++i) {
        V3f d;
        float p = 0;
        RGB<float> l;
        if (!m->sample(u[2 * i], u[2 * i + 1], d, p, l)) { p = 0; d = V3f(0, 0, 0); l = RGB<float>(0); }
        wi[3 * i] = d.x; wi[3 * i + 1] = d.y; wi[3 * i + 2] = d.z;
        L[3 * i] = l.r; L[3 * i + 1] = l.g; L[3 * i + 2] = l.b;
        pdf[i] = p;
    }
}
void ref_env_eval(void* e, const float* dirs, int64_t n, int quirks, float* L, float* pdf) {
    const RefEnvMap* m = static_cast<RefEnvMap*>(e);
    for (int64_t i = 0; i < n; ++i) {
        const V3f d(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
        const RGB<float> l = m->light_emit(d);
        L[3 * i] = l.r; L[3 * i + 1] = l.g; L[3 * i + 2] = l.b;
        pdf[i] = m->pdf(d, quirks != 0);
    }
}
void ref_env_radiance_integral(void* e, float* out3) {
    const RefEnvMap* m = static_cast<RefEnvMap*>(e);
    out3[0] = m->avg_radiance.r; out3[1] = m->avg_radiance.g; out3[2] = m->avg_radiance.b;
}
double ref_hg_eval(double g, double cos_wi_wo) { HGRef<double> h{g}; return h.phase(-cos_wi_wo); }
void ref_hg_sample(float g, const float* wo, const float* u, int64_t n, float* wi, float* pdf) {
    HGRef<float> h{g};
    for (int64_t i = 0; i < n; ++i) {
        BSDFSample<float> s = h.sample(V3f(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), u[2 * i], u[2 * i + 1]);
        wi[3 * i] = s.wi.x; wi[3 * i + 1] = s.wi.y; wi[3 * i + 2] = s.wi.z; pdf[i] = s.pdf;
    }
}
// Shading frames of SurfaceIntersection (triangle.cpp:178-211, through
// fill_isect) at hits of prims[i] with barycentrics (alpha, beta) = bary[2i..]:
// x, y, z axes as nine floats per hit.
void ref_shading_frames(void* hs, const int32_t* prims, const float* bary, int64_t n, float* out) {
    Scene* s = static_cast<Scene*>(hs);
    for (int64_t i = 0; i < n; ++i) {
        Isect is;
        fill_isect(*s, prims[i], V3f(0, 0, 0), V3f(0, 0, 1), 1.0f, bary[2 * i], bary[2 * i + 1], is);
        const V3f ax[3] = {is.shd.x, is.shd.y, is.shd.z};
        for (int k = 0; k < 3; ++k) {
            out[9 * i + 3 * k] = ax[k].x;
            out[9 * i + 3 * k + 1] = ax[k].y;
            out[9 * i + 3 * k + 2] = ax[k].z;
        }
    }
}

void ref_camera_rays(void* hs, const float* film_uv, const float* lens_uv, int64_t n, float* rays) {
    Scene* s = static_cast<Scene*>(hs);
    RenderCfg c{};
    Renderer r(*s, c);
    for (int64_t i = 0; i < n; ++i) {
        V3f o, d;
        r.camera_ray(film_uv[2 * i], film_uv[2 * i + 1], lens_uv[2 * i], lens_uv[2 * i + 1], o, d);
        d = d.normalize();
        float* q = rays + 8 * i;
        q[0] = o.x; q[1] = o.y; q[2] = o.z; q[3] = 0; q[4] = d.x; q[5] = d.y; q[6] = d.z; q[7] = std::numeric_limits<float>::max();
    }
}

// ---- render ----------------------------------------------------------------
// sums: W*H*4 doubles, accumulated.  Returns milliseconds of render() only;
// rays_out (optional) gets extension and shadow ray counts.
double ref_render(void* hs, const tr_render_params* p, int quirks, int nee, int threads, double* sums, uint64_t* rays_out) {
    Scene* s = static_cast<Scene*>(hs);
    RenderCfg c;
    c.W = p->width; c.H = p->height; c.spp = p->spp; c.min_depth = p->min_depth; c.max_depth = p->max_depth;
    c.nls = std::max(1, p->direct_light_sample_num);
    c.filter = p->filter_type; c.fr = p->filter_radius; c.fa = p->filter_alpha;
    c.seed = p->seed; c.sample_offset = p->sample_offset; c.quirks = quirks != 0; c.nee = nee;
    Renderer r(*s, c);
    r.tri_to_light.assign(s->mat.size(), -1);
    for (size_t i = 0; i < s->lights.size(); ++i) r.tri_to_light[s->lights[i]] = int(i);
    if (threads <= 0) threads = int(std::thread::hardware_concurrency());
    auto t0 = std::chrono::steady_clock::now();
    r.render(sums, threads);
    auto t1 = std::chrono::steady_clock::now();
    if (rays_out) { rays_out[0] = r.ext_rays.load(); rays_out[1] = r.shadow_rays.load(); }
    return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

}  // extern "C"

// ---- closed-form helpers exposed for known-answer tests --------------------
extern "C" {
void ref_pcg32(uint64_t state, uint64_t stream, int default_stream, int64_t n, uint32_t* out) {
    Pcg32 g;
    if (default_stream) g.seed_default_stream(state);
    else g.seed(state, stream);
    for (int64_t i = 0; i < n; ++i) out[i] = g.next_u32();
}
double ref_fresnel(double eta_i, double eta_o, double cos_i) { return dielectric_fresnel<double>(eta_i, eta_o, cos_i); }
// anisotropic GTR2 and GTR1 distributions of a unit half vector, double precision
double ref_D_gtr2(double hx, double hy, double hz, double ax, double ay) {
    V3d h(hx, hy, hz);
    double sphi, cphi;
    double phi = local_phi(h);
    sphi = std::sin(phi); cphi = std::cos(phi);
    return mf::anisotropic_gtr2(sphi, cphi, local_cos_to_sin(hz), hz, ax, ay);
}
double ref_D_gtr1(double hz, double alpha) { return mf::gtr1(local_cos_to_sin(hz), hz, alpha); }
double ref_G1_gtr2(double wx, double wy, double wz, double ax, double ay) {
    V3d w(wx, wy, wz);
    double phi = local_phi(w);
    return mf::smith_anisotropic_gtr2(std::cos(phi), std::sin(phi), ax, ay, local_tan_theta(w));
}
}
