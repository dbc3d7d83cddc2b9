// Bidirectional path tracing kernels: camera and light subpaths, the
// connection strategies with their multiple-importance weights, the gather of
// each sample's camera-side radiance and the light-tracing splats.
//
// Replaces BDPTRenderer::render and its helpers
// (src/renderer/bdpt_renderer.cpp:289-1185).  The strategies, the vertex
// densities and their reassignment per strategy follow the reference; three
// of its weights are corrected, each of which leaves the reference's
// combination of strategies biased (and so are two more of its choices, the
// environment's emission disk and the shading-normal adjoint, below):
//   * compute_mis_weight (bdpt_renderer.cpp:741-762) never counts the
//     strategy with no light vertex: its light loop adds a ratio only for
//     i >= 1.  A path that the camera subpath reaches on its own is then
//     weighted by every other strategy as if that strategy did not exist,
//     while its own weight counts the others, and the weights of the path sum
//     to more than 1.  Here the i = 0 ratio is added (no light here is a
//     delta light).
//   * The weights sum over every split of the path, including splits whose
//     camera or light part is longer than max_camera_vertex_count or
//     max_light_vertex_count allow.  Those strategies never run, so paths of
//     more than min(nc, nl + 1) vertices lose energy.  Here a split counts
//     only if its strategy is sampled, and every path of at most nc + nl
//     vertices is weighted to exactly 1.
//   * mis_weight_tx_s0 (:816-830) evaluates the environment's emission
//     density at light_to_ref, opposite to the direction it samples; here at
//     the sampled direction (the same for a constant environment).
//   * IBL::sample_le (environment_light.cpp:142-145) starts a light subpath
//     on a disk at world_center - wi R and sends it along -wi, away from the
//     scene: no such subpath ever hits it, while the weights still count the
//     strategies that would use one.  Here the disk lies at world_center +
//     wi R, on the side the light comes from.
//   * Light subpaths evaluate the BSDF in its radiance form, whose normal
//     correction |wi.ns| / |wi.ng| is not the adjoint's where shading and
//     geometric normals differ (only refraction's eta^2 changes with the
//     transport mode, disney.cpp:98, glass.cpp:65); here importance transport
//     takes |wo.ns| / |wo.ng| (Veach 1997, sec. 5.3), so camera and light
//     subpaths agree over interpolated normals and normal maps.
//   * At a specular vertex the reference divides by 1 instead of the vertex's
//     density (:748, :755) and keeps the sampled lobe weight as the density
//     of the vertex after it (:369-378).  Two strategies then disagree on
//     their density ratio for a path through glass, and its weights do not
//     sum to 1.  Here the vertex after a specular one gets density 0, as the
//     vertex before it already does, both count as 1, and the specular
//     vertex's own ratio is kept: every strategy computes the same ratios.
// Path-tracer conventions carry over: triangle emission is one-sided about
// the geometric normal, media are sampled by delta tracking, connections
// through a medium take its transmittance (ratio tracking in a density grid).
#include <algorithm>

#include "tr_bdpt.h"

namespace trhip {

namespace {

constexpr int kBdBlock = 256;
constexpr uint64_t kLightStream = 0x6c69676874707468ull;
constexpr uint64_t kConnStream = 0x636f6e6e65637473ull;

TR_D int vtype(const BdVertex& v) { return __float_as_int(v.beta.w) & 7; }
TR_D bool vdelta(const BdVertex& v) { return (__float_as_int(v.beta.w) & kBdDelta) != 0; }
TR_D int vaux(const BdVertex& v) { return __float_as_int(v.wo.w); }

TR_D Pcg load_pcg(uint2 s, uint64_t inc) {
    Pcg r;
    r.state = (uint64_t(s.y) << 32) | s.x;
    r.inc = inc;
    return r;
}
TR_D uint2 save_pcg(const Pcg& r) { return make_uint2(uint32_t(r.state), uint32_t(r.state >> 32)); }

// Density `pdf` of a direction sampled at `from` as a density per unit area
// (or per solid angle for the environment) at vertex v
// (bdpt::pdf_solid_angle_to_area, bdpt_renderer.cpp:228-258).
TR_D float pdf_to_area(float pdf, f3 from, const BdVertex& v) {
    const int ty = vtype(v);
    if (ty == BD_ENV) return pdf;
    const f3 w = xyz(v.p) - from;
    const float d2 = dot(w, w);
    if (ty == BD_MEDIUM) return pdf / d2;
    return pdf * fabsf(dot(xyz(v.n), w)) / (sqrtf(d2) * d2);
}

// A scattering vertex with its shading frame; its material is read from the
// scene's table where it is evaluated (a copy is made only for a textured one).
struct BdScatter {
    bool medium;
    f3 pos, wo, ng, ns, tx, ty;
    int mat, tri;
    bool glass, textured;
    float b1, b2, g;
    int med_in, med_out;
};

TR_D void load_scatter(const DevScene& s, const BdVertex& v, BdScatter& c) {
    c.pos = xyz(v.p);
    c.wo = xyz(v.wo);
    const int aux = vaux(v);
    c.b1 = v.x.x;
    c.b2 = v.x.y;
    c.tri = aux;
    if (vtype(v) == BD_MEDIUM) {
        c.medium = true;
        c.g = medium_g(s, s.media[aux], c.pos);
        c.med_in = c.med_out = aux;
        c.ng = c.ns = c.tx = c.ty = f3s(0.0f);
        c.mat = 0;
        c.glass = c.textured = false;
        return;
    }
    c.medium = false;
    c.g = 0.0f;
    c.ng = xyz(v.n);
    const TriShade sr = load_shade(s, aux);
    c.ns = normalize(sr.n0 * (1.0f - c.b1 - c.b2) + sr.n1 * c.b1 + sr.n2 * c.b2);
    shading_frame(s, aux, c.ns, c.tx, c.ty);
    c.mat = sr.material;
    const DevMaterial& m = s.mats[sr.material];
    c.glass = m.type == MAT_GLASS;
    c.textured = m.tex > 0 || m.texset > 0;
    if (m.texset > 0) {   // the normal map turns the frame once; sc_bsdf reads the parameters
        float u, v;
        hit_uv(s, aux, c.b1, c.b2, u, v);
        normal_map_frame(s, m, u, v, c.tx, c.ty, c.ns);
    }
    c.med_in = sr.med_in;
    c.med_out = sr.med_out;
}

// The vertex's material (textured: kd / base colour or the Disney
// parameters at its uv).
TR_D DevMaterial sc_material(const DevScene& s, const BdScatter& c) {
    DevMaterial m = s.mats[c.mat];
    if (c.textured) {
        if (m.texset > 0) {
            float u, v;
            hit_uv(s, c.tri, c.b1, c.b2, u, v);
            disney_params_at(s, m, u, v);
        } else {
            apply_texture(s, m, c.tri, c.b1, c.b2);
        }
    }
    return m;
}
TR_D BsdfEval sc_bsdf(const DevScene& s, const BdScatter& c, f3 lwi, f3 lwo, f3 lng) {
    if (c.textured) return bsdf_eval(sc_material(s, c), lwi, lwo, lng);
    return bsdf_eval(s.mats[c.mat], lwi, lwo, lng);
}

TR_D f3 sc_local(const BdScatter& c, f3 w) { return mk(dot(w, c.tx), dot(w, c.ty), dot(w, c.ns)); }

// Light-subpath (importance) transport differs from radiance transport only
// in refraction: glass and Disney transmission leave out the radiance
// scaling by (eta_i / eta_t)^2 (glass.cpp:65, disney.cpp:98).
TR_D float adjoint_scale(const DevMaterial& m, f3 lwi, f3 lwo) {
    const bool refracts = m.type == MAT_GLASS || (m.type == MAT_DISNEY && m.transmission > 0.0f);
    if (!refracts || !(lwi.z * lwo.z < 0.0f)) return 1.0f;
    return lwo.z > 0.0f ? m.ior * m.ior : 1.0f / (m.ior * m.ior);
}

// Importance transport through a vertex whose shading normal differs from
// its geometric one (Veach 1997, sec. 5.3): the BSDF's value carries
// |wi.ns| / |wi.ng| for radiance (normal_correct_factor, coordinate.hpp:43-47),
// and a light subpath needs |wo.ns| / |wo.ng| instead; the reference's BDPT
// keeps the radiance form, so its light subpaths over shading normals
// disagree with its camera subpaths.  Black-fringe directions (a Lambert lobe
// about the geometric normal) and specular vertices keep their value.
TR_D float shading_adjoint(const DevMaterial& m, f3 lwi, f3 lwo, f3 lng) {
    if (m.type == MAT_GLASS || fringe(lwi, lng) || fringe(lwo, lng)) return 1.0f;
    const float num = fabsf(lwo.z) * fabsf(dot(lwi, lng)), den = fabsf(dot(lwo, lng)) * fabsf(lwi.z);
    return den > 0.0f ? num / den : 0.0f;
}

// f(wi, wo) at the vertex (BSDF::eval / the phase function); importance:
// light-subpath transport.
TR_D f3 sc_f(const DevScene& s, const BdScatter& c, f3 wi, bool importance) {
    if (c.medium) return f3s(hg_phase(c.g, -dot(wi, c.wo)));
    const f3 lwi = sc_local(c, wi), lwo = sc_local(c, c.wo), lng = sc_local(c, c.ng);
    const DevMaterial m = sc_material(s, c);
    f3 f = bsdf_eval(m, lwi, lwo, lng).f;
    if (importance) f *= adjoint_scale(m, lwi, lwo) * shading_adjoint(m, lwi, lwo, lng);
    return f;
}

// Density of sampling w1 when w2 is the outgoing direction (BSDF::pdf(wi = w1,
// wo = w2), src/core/bsdf.hpp:35); 0 at a specular vertex.
TR_D float sc_pdf(const DevScene& s, const BdScatter& c, f3 w1, f3 w2) {
    if (c.medium) return hg_phase(c.g, -dot(w1, w2));
    if (c.glass) return 0.0f;
    return sc_bsdf(s, c, sc_local(c, w1), sc_local(c, w2), sc_local(c, c.ng)).pdf;
}

// Medium on the side of direction w (SurfaceIntersection::medium, src/core/intersection.hpp:36-41).
TR_D int sc_medium(const BdScatter& c, f3 w) {
    if (c.medium) return c.med_in;
    return dot(w, c.ng) > 0.0f ? c.med_out : c.med_in;
}

// Transmittance over [0, dist) from o along w (Medium::tr): closed form in a
// homogeneous medium, ratio tracking in a density grid.
TR_D f3 bd_tr(const DevScene& s, int mid, f3 o, f3 w, float dist, Pcg& rng) {
    if (mid < 0) return f3s(1.0f);
    const DevMedium& m = s.media[mid];
    if (m.grid >= 0) return f3s(grid_tr(s, m, o, w, dist, rng));
    return mk(expf(-m.sigma_t.x * dist), expf(-m.sigma_t.y * dist), expf(-m.sigma_t.z * dist));
}

// Free-flight sampling of a segment [0, t) in medium m, as the path tracer
// samples it (delta tracking in a density grid, spectral tracking in a
// homogeneous medium), without a limit on scattering events: the subpaths end
// at their vertex counts.  Returns 0: the segment was crossed, 1: scattered at
// x, 2: absorbed; T takes the tracking weights.
TR_D int track_segment(const DevScene& s, const DevMedium& m, f3 o, f3 d, float t, Pcg& rng, f3& T, f3& x) {
    if (m.grid >= 0) {
        const float inv = 1.0f / m.majorant;
        float tt = 0.0f;
        for (int k = 0; k < (1 << 20); ++k) {
            tt -= logf(1.0f - rng.uniform()) * inv;
            if (tt >= t) return 0;
            const f3 p = o + d * tt;
            if (rng.uniform() < grid_density(s, m, p) * inv) {
                T *= medium_albedo(s, m, p);   // the albedo of a grid medium at the collision
                x = p;
                return 1;
            }
        }
        return 0;
    }
    float tt = 0.0f;
    for (int k = 0; k < 100000; ++k) {
        tt -= logf(1.0f - rng.uniform()) / m.majorant;
        if (tt >= t) return 0;
        const float xi = rng.uniform();
        if (xi < m.p_absorb) return 2;
        if (xi < m.p_absorb + m.p_scatter) {
            T *= m.sigma_s * (1.0f / (m.majorant * m.p_scatter));
            x = o + d * tt;
            return 1;
        }
        T *= (f3s(m.majorant) - m.sigma_t) * (1.0f / (m.majorant * m.p_null));
    }
    return 0;
}

TR_D void put_vertex(BdVertex& v, f3 pos, float pdf_fwd, f3 n, float pdf_bwd, f3 beta, int type, f3 wo, int aux, float b1,
                     float b2) {
    v.p = f4(pos, pdf_fwd);
    v.n = f4(n, pdf_bwd);
    v.beta = f4(beta, __int_as_float(type));
    v.wo = f4(wo, __int_as_float(aux));
    v.x = make_float4(b1, b2, 0.0f, 0.0f);
}

}  // namespace

// ---- subpath starts -------------------------------------------------------
// Camera vertex and ray (generate_camera_subpath, :289-313; the film and lens
// samples are the path tracer's, so k_accumulate finds the same film
// position); light vertex and ray (generate_light_subpath, :386-430): a light
// chosen by power, a point on it, a cosine-distributed direction about its
// normal, or for the environment a direction from its distribution and an
// origin on the disk of the scene's bounding sphere facing it
// (IBL::sample_le, environment_light.cpp:123-157).
__global__ __launch_bounds__(kBdBlock) void k_bd_start(DevScene s, BdPaths b, BdConsts k) {
    __shared__ BlockAppendLds<kBdBlock> ap;
    const int i = blockIdx.x * kBdBlock + threadIdx.x;
    bool cam_go = false, light_go = false;
    if (i < k.slots) {
        const uint32_t pixel = uint32_t(i % k.npix);
        const uint32_t sample = uint32_t(k.sample_base + i / k.npix);
        Pcg rng = pcg_for_sample(k.seed, pixel, sample);
        const float fu = rng.uniform(), fv = rng.uniform(), lu = rng.uniform(), lv = rng.uniform();
        const int px = int(pixel % uint32_t(k.width)), py = int(pixel / uint32_t(k.width));
        const float fx = (float(px) + fu) / float(k.width), fy = (float(py) + fv) / float(k.height);
        f3 o, d;
        camera_ray(s.cam, fx, fy, lu, lv, o, d);
        // ThinLensCamera::pdf_we (thin_lens_camera.cpp:85-95); We cos / (pdf_pos pdf_dir) = 1
        const float cz = dot(d, s.cam.dir);
        const float pdf_dir = s.cam.focal * s.cam.focal / (k.film_area * cz * cz * cz);
        put_vertex(b.cv[size_t(i) * k.nc], o, 1.0f / k.lens_area, s.cam.dir, 0.0f, f3s(k.lens_area), BD_CAMERA,
                   f3s(0.0f), -1, 0.0f, 0.0f);
        b.ncv[i] = 1;
        b.gray[2 * size_t(i)] = f4(o, 0.0f);
        b.gray[2 * size_t(i) + 1] = f4(d, 3.0e38f);
        b.gbeta[i] = make_float4(1.0f, 1.0f, 1.0f, pdf_dir);
        b.grng[i] = save_pcg(rng);
        cam_go = k.nc > 1;

        Pcg lr = pcg_for_sample(k.seed ^ kLightStream, pixel, sample);
        int nl = 0;
        const size_t r = size_t(k.slots) + size_t(i);
        if (k.nl > 0 && s.num_light_entries > 0) {
            float pmf;
            const int e = pick_light(s, lr.uniform(), pmf);
            const float u1 = lr.uniform(), u2 = lr.uniform(), u3 = lr.uniform(), u4 = lr.uniform();
            BdVertex& l0 = b.lv[size_t(i) * k.nl];
            if (pmf > 0.0f && e < s.num_area_lights) {
                // DiffuseLight::sample_le (diffuse_light.cpp:62-80) on a triangle
                const float4* lrec = s.lights + 4 * size_t(e);
                const float4 q = lrec[0];
                const f3 A = xyz(lrec[1]), B = xyz(lrec[2]), C = xyz(lrec[3]);
                const float su = sqrtf(u1);
                const float bx = 1.0f - su, by = u2 * su;
                const f3 pos = A * bx + B * by + C * (1.0f - bx - by);
                const f3 n = normalize(cross(B - A, C - A));
                f3 tx, ty;
                onb(n, tx, ty);
                const f3 lc = cosine_hemisphere(u3, u4);
                const float pdir = lc.z * kInvPi;
                const float init_pdf = pmf / q.w;   // selection x 1 / area
                if (pdir > 0.0f && init_pdf > 0.0f) {
                    const f3 dir = normalize(tx * lc.x + ty * lc.y + n * lc.z);
                    put_vertex(l0, pos, 0.0f, n, init_pdf, f3s(1.0f / init_pdf), BD_AREA, f3s(0.0f), e, 0.0f, 0.0f);
                    nl = 1;
                    b.gray[2 * r] = f4(pos + n * kEps, 0.0f);
                    b.gray[2 * r + 1] = f4(dir, 3.0e38f);
                    b.gbeta[r] = f4(xyz(q) * (lc.z / (init_pdf * pdir)), pdir);
                    light_go = k.nl > 1;
                }
            } else if (pmf > 0.0f && e == s.num_area_lights) {
                const LightSampleD ls = sample_light<true>(s, e, s.wcenter, u1, u2);
                if (ls.pdf > 0.0f) {
                    // the disk faces the scene from the light's side: the
                    // reference places it at center - wi R and shoots along
                    // -wi (environment_light.cpp:142-145), away from the
                    // scene, so its light subpaths never hit anything
                    f3 tx, ty;
                    onb(ls.wi, tx, ty);
                    float dx, dy;
                    concentric_disk(u3, u4, dx, dy);
                    const f3 pos = s.wcenter + (tx * dx + ty * dy + ls.wi) * s.wradius;
                    const f3 dir = -ls.wi;
                    const float init_pdf = ls.pdf * pmf;
                    put_vertex(l0, pos, 0.0f, dir, init_pdf, f3s(1.0f / init_pdf), BD_ENV, f3s(0.0f), e, 0.0f, 0.0f);
                    nl = 1;
                    b.gray[2 * r] = f4(pos, kEps);
                    b.gray[2 * r + 1] = f4(dir, 3.0e38f);
                    b.gbeta[r] = f4(ls.L * (1.0f / (init_pdf * k.env_pdf_pos)), ls.pdf);
                    light_go = k.nl > 1;
                }
            }
        }
        b.nlv[i] = nl;
        b.grng[r] = save_pcg(lr);
    }
    unsigned int ic, il;
    block_append2(ap, bd_counter(b, 0), cam_go, bd_counter(b, 0), light_go, ic, il);
    if (cam_go) b.gq[0][ic] = i;
    if (light_go) b.gq[0][il] = k.slots + i;
}

// ---- subpath extension ----------------------------------------------------
// One hit of a subpath ray: the vertex it makes (an escaping camera ray makes
// an environment vertex), its density from the previous vertex, the reverse
// density of the previous vertex once the next direction is sampled, and the
// next ray (generate_camera_subpath :314-381, generate_light_subpath :431-496).
// Returns whether the subpath goes on.
TR_D bool extend_one(const DevScene& s, const BdPaths& b, const BdConsts& k, int ray) {
    const bool light = ray >= k.slots;
    const int slot = light ? ray - k.slots : ray;
    const int maxc = light ? k.nl : k.nc;
    BdVertex* V = light ? b.lv + size_t(slot) * k.nl : b.cv + size_t(slot) * k.nc;
    int* cntp = light ? b.nlv + slot : b.ncv + slot;
    int cnt = *cntp;
    const float4 ro = b.gray[2 * size_t(ray)], rd = b.gray[2 * size_t(ray) + 1];
    const float4 h = b.ghit[ray];
    const float4 gb = b.gbeta[ray];
    f3 beta = xyz(gb);
    const float pdf_sa = gb.w;
    const f3 o = xyz(ro), d = xyz(rd);
    const int tri = __float_as_int(h.y);
    BdVertex& prev = V[cnt - 1];
    // the first vertex after an environment light: the density of the ray's
    // origin on the disk (generate_light_subpath, :497-499)
    const bool after_env = light && cnt == 1 && vtype(prev) == BD_ENV;
    if (tri < 0) {
        if (!light && s.has_env) {
            put_vertex(V[cnt], o + d * (2.0f * s.wradius), pdf_sa, -d, 0.0f, beta, BD_ENV, f3s(0.0f), -1, 0.0f, 0.0f);
            *cntp = cnt + 1;
        }
        return false;
    }
    const uint32_t pixel = uint32_t(slot % k.npix);
    Pcg rng = load_pcg(b.grng[ray], (uint64_t(pixel) << 1) | 1u);
    const float t = h.x;
    const TriShade sr = load_shade(s, tri);
    const f3 ng = sr.ng;
    const int mid = dot(d, ng) <= 0.0f ? sr.med_out : sr.med_in;
    int ev = 0;
    f3 x;
    if (mid >= 0) ev = track_segment(s, s.media[mid], o, d, t, rng, beta, x);
    if (ev == 2) return false;
    const f3 wo = -d;
    if (ev == 1) {
        const f3 dx = x - o;
        const float pa = pdf_sa / dot(dx, dx);
        put_vertex(V[cnt], x, light ? 0.0f : pa, f3s(0.0f), light ? (after_env ? k.env_pdf_pos : pa) : 0.0f, beta,
                   BD_MEDIUM, wo, mid, 0.0f, 0.0f);
        *cntp = ++cnt;
        if (cnt >= maxc) return false;
        float ppdf;
        const f3 wi = hg_sample(medium_g(s, s.media[mid], x), wo, rng.uniform(), rng.uniform(), ppdf);
        // the phase function is symmetric: the reverse density is ppdf
        const float rev = pdf_to_area(ppdf, x, prev);
        if (light) prev.p.w = rev;
        else prev.n.w = rev;
        b.gray[2 * size_t(ray)] = f4(x, 0.0f);
        b.gray[2 * size_t(ray) + 1] = f4(wi, 3.0e38f);
        b.gbeta[ray] = f4(beta, ppdf);
        b.grng[ray] = save_pcg(rng);
        return true;
    }
    const f3 pos = o + d * t;
    const float pa = pdf_sa * fabsf(dot(ng, wo)) / (t * t);
    const float b1 = h.z, b2 = h.w;
    DevMaterial mat = s.mats[sr.material];
    f3 ns = normalize(sr.n0 * (1.0f - b1 - b2) + sr.n1 * b1 + sr.n2 * b2);
    f3 tx, ty;
    shading_frame(s, tri, ns, tx, ty);
    apply_material_textures(s, mat, tri, b1, b2, tx, ty, ns);
    const bool delta = mat.type == MAT_GLASS;
    put_vertex(V[cnt], pos, light ? 0.0f : pa, ng, light ? (after_env ? k.env_pdf_pos : pa) : 0.0f, beta,
               BD_SURFACE | (delta ? kBdDelta : 0), wo, tri, b1, b2);
    *cntp = ++cnt;
    if (cnt >= maxc) return false;
    const f3 lwo = mk(dot(wo, tx), dot(wo, ty), dot(wo, ns)), lng = mk(dot(ng, tx), dot(ng, ty), dot(ng, ns));
    const float su = rng.uniform(), u1 = rng.uniform(), u2 = rng.uniform();
    const BsdfSample bs = bsdf_sample(mat, lwo, lng, su, u1, u2);
    if (!(bs.pdf > 0.0f) || is_black(bs.f)) return false;
    const f3 wi = normalize(tx * bs.wi.x + ty * bs.wi.y + ns * bs.wi.z);
    f3 f = bs.f;
    if (light) f *= adjoint_scale(mat, bs.wi, lwo) * shading_adjoint(mat, bs.wi, lwo, lng);
    const float rpdf = delta ? 0.0f : bsdf_eval(mat, lwo, bs.wi, lng).pdf;
    const float rev = pdf_to_area(rpdf, pos, prev);
    if (light) prev.p.w = rev;
    else prev.n.w = rev;
    beta *= f * (fabsf(dot(ng, wi)) / bs.pdf);
    if (!finite3(beta)) return false;
    const f3 org = dot(ng, wi) > 0.0f ? pos + ng * kEps : pos - ng * kEps;
    b.gray[2 * size_t(ray)] = f4(org, 0.0f);
    b.gray[2 * size_t(ray) + 1] = f4(wi, 3.0e38f);
    // after a specular vertex the next vertex's density is 0 (counted as 1 by
    // the weights), as the vertex before it gets reverse density 0
    b.gbeta[ray] = f4(beta, delta ? 0.0f : bs.pdf);
    b.grng[ray] = save_pcg(rng);
    return true;
}

__global__ __launch_bounds__(kBdBlock) void k_bd_extend(DevScene s, BdPaths b, BdConsts k, int cur) {
    __shared__ BlockAppendLds<kBdBlock> ap;
    const unsigned int n = *bd_counter(b, cur);
    for (unsigned int base = blockIdx.x * kBdBlock; base < n; base += gridDim.x * kBdBlock) {
        const unsigned int idx = base + threadIdx.x;
        int ray = -1;
        bool go = false;
        if (idx < n) {
            ray = b.gq[cur][idx];
            go = extend_one(s, b, k, ray);
        }
        unsigned int q, unused;
        block_append2(ap, bd_counter(b, 1 - cur), go, bd_counter(b, 1 - cur), false, q, unused);
        if (go) b.gq[1 - cur][q] = ray;
    }
}

// ---- connections ----------------------------------------------------------
namespace {

// Whether strategy (s, t) is ever sampled.
TR_D bool bd_sampled(int s, int t, const BdConsts& k) {
    return t >= 1 && t <= k.nc && s <= k.nl && !(t == 1 && s <= 1);
}

// The balance-heuristic weight of strategy (s, t) (compute_mis_weight,
// bdpt_renderer.cpp:741-762) with the densities the strategy reassigns:
// camera[t-1].pdf_bwd = b_bwd, camera[t-2].pdf_bwd = a_bwd,
// light[s-1].pdf_fwd = c_fwd, light[s-2].pdf_fwd = d_fwd.  A ratio is
// added only for a strategy that is sampled, the strategy with no light
// vertex is counted (the lights are not delta lights), and a specular vertex
// divides by its density like any other (its neighbours carry density 0).
TR_D float bd_remap(float f) { return (isfinite(f) && f > 0.0f) ? f : 1.0f; }
TR_D float mis_weight(const BdVertex* C, int t, const BdVertex* L, int s, float a_bwd, float b_bwd, float c_fwd,
                      float d_fwd, const BdConsts& k) {
    const int n = s + t;
    float sum = 1.0f, cur = 1.0f;
    for (int i = s - 1; i >= 0; --i) {
        const float pf = i == s - 1 ? c_fwd : (i == s - 2 ? d_fwd : L[i].p.w);
        const bool dl = vdelta(L[i]);
        cur *= bd_remap(pf) / bd_remap(L[i].n.w);
        const bool dprev = i >= 1 ? vdelta(L[i - 1]) : false;
        if (!dl && !dprev && bd_sampled(i, n - i, k)) sum += cur;
    }
    cur = 1.0f;
    for (int i = t - 1; i >= 1; --i) {
        const float pb = i == t - 1 ? b_bwd : (i == t - 2 ? a_bwd : C[i].n.w);
        const bool dc = vdelta(C[i]);
        cur *= bd_remap(pb) / bd_remap(C[i].p.w);
        if (!dc && !vdelta(C[i - 1]) && bd_sampled(n - i, i, k)) sum += cur;
    }
    return 1.0f / sum;
}

struct BdConn {
    f3 c;          // weighted contribution, unoccluded
    int splat;     // film pixel of a t = 1 strategy, -1 = the sample's own pixel
    bool ray;      // needs a visibility test
    f3 o, d;
    float tmax;
};

// Strategy j = (t - 1) * (nl + 1) + s of slot `slot`.
TR_D void connect_one(const DevScene& s, const BdPaths& b, const BdConsts& k, int slot, int j, BdConn& cn) {
    const int t = j / (k.nl + 1) + 1, sl = j % (k.nl + 1);
    const int nct = b.ncv[slot], nlt = b.nlv[slot];
    if (t > nct || sl > nlt || !bd_sampled(sl, t, k)) return;
    if (k.only_s >= 0 && sl != k.only_s) return;
    if (k.only_t >= 1 && t != k.only_t) return;
    const bool mis = k.only_s < 0 && k.only_t < 1;
    const BdVertex* C = b.cv + size_t(slot) * k.nc;
    const BdVertex* L = b.lv + size_t(slot) * k.nl;
    // an environment vertex ends the camera subpath: only s = 0 uses it (:991-993)
    if (t > 1 && sl != 0 && vtype(C[t - 1]) == BD_ENV) return;
    const uint32_t pixel = uint32_t(slot % k.npix);
    const uint32_t sample = uint32_t(k.sample_base + slot / k.npix);
    Pcg rng = pcg_seed(mix64(k.seed ^ kConnStream) ^ ((uint64_t(sample) << 32) | pixel), uint64_t(j));

    if (sl == 0) {
        // the camera subpath ends on a light (t2_s0_path_contrib :502-525,
        // tx_s0_path_contrib :527-548, mis_weight_tx_s0 :764-831)
        const BdVertex& bv = C[t - 1];
        const BdVertex& av = C[t - 2];
        const int ty = vtype(bv);
        f3 Le;
        float sel, a_bwd;
        if (ty == BD_ENV) {
            const f3 wi = -xyz(bv.n);
            float epdf;
            Le = env_eval<true>(s, wi, epdf);
            sel = s.light_pmf[s.num_area_lights] * epdf;
            a_bwd = k.env_pdf_pos;
        } else if (ty == BD_SURFACE) {
            const int li = load_shade(s, vaux(bv)).light;
            if (li < 0) return;
            const float4 q = s.lights[4 * li];
            const f3 to_a = xyz(av.p) - xyz(bv.p);
            const float da = length(to_a);
            const float cosl = dot(xyz(bv.n), to_a) / da;
            if (!(cosl > 0.0f)) return;   // one-sided emission (diffuse_light.cpp:26-29)
            Le = xyz(q);
            sel = s.lights[4 * li + 1].w / q.w;
            a_bwd = pdf_to_area(cosl * kInvPi, xyz(bv.p), av);
        } else {
            return;
        }
        cn.c = xyz(bv.beta) * Le;
        if (mis) cn.c *= mis_weight(C, t, L, 0, a_bwd, sel, 0.0f, 0.0f, k);
        return;
    }

    if (t == 1) {
        // light tracing: the light subpath's end seen by the camera
        // (t1_sx_path_contrib :616-676, mis_weight_t1_sx :879-912)
        const BdVertex& bv = L[sl - 1];
        BdScatter B;
        load_scatter(s, bv, B);
        const f3 cpos = xyz(C[0].p);
        const f3 dv = B.pos - cpos;
        const float d2 = dot(dv, dv), dist = sqrtf(d2);
        const f3 w = dv * (1.0f / dist);
        const DevCamera& cam = s.cam;
        // ThinLensCamera::eval_we (thin_lens_camera.cpp:61-83)
        const float cz = dot(w, cam.dir);
        if (!(cz > 0.0f)) return;
        const f3 lens = cpos - cam.pos;
        const float fxp = dot(lens, cam.right) + cam.focal / cz * dot(w, cam.right);
        const float fyp = dot(lens, cam.up) + cam.focal / cz * dot(w, cam.up);
        const float px = (0.5f - fxp / cam.film_w) * float(k.width);
        const float py = (0.5f - fyp / cam.film_h) * float(k.height);
        if (!(px >= 0.0f && px <= float(k.width) && py >= 0.0f && py <= float(k.height))) return;
        const float cz2 = cz * cz;
        const float we = cam.focal * cam.focal / (k.film_area * k.lens_area * cz2 * cz2);
        const f3 l2c = -w;
        const f3 f = sc_f(s, B, l2c, true);
        if (is_black(f)) return;
        float G = cz / d2;
        if (!B.medium) G *= fabsf(dot(B.ng, l2c));
        const f3 tr = bd_tr(s, sc_medium(B, l2c), B.pos, l2c, dist, rng);
        cn.c = f * xyz(bv.beta) * tr * (we * G / C[0].p.w);
        if (mis) {
            const float pdf_dir = cam.focal * cam.focal / (k.film_area * cz2 * cz);
            const float c_fwd = pdf_to_area(pdf_dir, cpos, bv);
            const float d_fwd = pdf_to_area(sc_pdf(s, B, B.wo, l2c), B.pos, L[sl - 2]);
            cn.c *= mis_weight(C, 1, L, sl, 0.0f, 0.0f, c_fwd, d_fwd, k);
        }
        cn.splat = min(int(px), k.width - 1) + k.width * min(int(py), k.height - 1);
        cn.ray = true;
        cn.o = B.pos;
        cn.d = l2c;
        cn.tmax = dist - kEps;
        return;
    }

    const BdVertex& bv = C[t - 1];
    const BdVertex& av = C[t - 2];
    BdScatter B;
    load_scatter(s, bv, B);
    if (sl == 1) {
        // one light vertex (tx_s1_path_contrib :550-614, mis_weight_tx_s1 :833-877)
        const BdVertex& cv = L[0];
        if (vtype(cv) == BD_AREA) {
            const f3 cpos = xyz(cv.p);
            const f3 dv = cpos - B.pos;
            const float d2 = dot(dv, dv), dist = sqrtf(d2);
            const f3 w = dv * (1.0f / dist);
            const float cosl = -dot(xyz(cv.n), w);
            if (!(cosl > 0.0f)) return;
            f3 f = sc_f(s, B, w, false);
            if (!B.medium) f *= fabsf(dot(B.ng, w));
            if (is_black(f)) return;
            const f3 Le = xyz(s.lights[4 * vaux(cv)]);
            const f3 tr = bd_tr(s, sc_medium(B, w), B.pos, w, dist, rng);
            cn.c = xyz(bv.beta) * f * Le * tr * (cosl / (d2 * cv.n.w));
            if (mis) {
                const float b_bwd = pdf_to_area(cosl * kInvPi, cpos, bv);
                const float a_bwd = pdf_to_area(sc_pdf(s, B, B.wo, w), B.pos, av);
                const float c_fwd = pdf_to_area(sc_pdf(s, B, w, B.wo), B.pos, cv);
                cn.c *= mis_weight(C, t, L, 1, a_bwd, b_bwd, c_fwd, 0.0f, k);
            }
            cn.ray = true;
            cn.o = B.pos;
            cn.d = w;
            cn.tmax = dist - kEps;
        } else {
            const f3 w = -xyz(cv.n);
            f3 f = sc_f(s, B, w, false);
            if (!B.medium) f *= fabsf(dot(B.ng, w));
            if (is_black(f)) return;
            float epdf;
            const f3 Le = env_eval<true>(s, w, epdf);
            cn.c = xyz(bv.beta) * f * Le * (1.0f / cv.n.w);
            if (mis) {
                const float a_bwd = pdf_to_area(sc_pdf(s, B, B.wo, w), B.pos, av);
                const float c_fwd = sc_pdf(s, B, w, B.wo);
                cn.c *= mis_weight(C, t, L, 1, a_bwd, k.env_pdf_pos, c_fwd, 0.0f, k);
            }
            cn.ray = true;
            cn.o = B.pos;
            cn.d = w;
            cn.tmax = 2.0f * s.wradius;
        }
        return;
    }

    // two subpaths joined (tx_sx_path_contrib :678-737, mis_weight_tx_sx :914-956);
    // the camera-side vertex is finished before the light-side one is loaded,
    // so only one shading frame and material are live at a time
    const BdVertex& cv = L[sl - 1];
    const BdVertex& dvx = L[sl - 2];
    const f3 cpos = xyz(cv.p);
    const f3 dv = cpos - B.pos;
    const float d2 = dot(dv, dv), dist = sqrtf(d2);
    const f3 w = dv * (1.0f / dist);
    f3 c = sc_f(s, B, w, false);
    if (is_black(c)) return;
    float G = 1.0f / d2;
    if (!B.medium) G *= fabsf(dot(B.ng, w));
    c *= bd_tr(s, sc_medium(B, w), B.pos, w, dist, rng) * xyz(bv.beta);
    float a_bwd = 0.0f, c_fwd = 0.0f;
    if (mis) {
        a_bwd = pdf_to_area(sc_pdf(s, B, B.wo, w), B.pos, av);
        c_fwd = pdf_to_area(sc_pdf(s, B, w, B.wo), B.pos, cv);
    }
    const f3 bpos = B.pos;
    BdScatter Cc;
    load_scatter(s, cv, Cc);
    const f3 fc = sc_f(s, Cc, -w, true);
    if (is_black(fc)) return;
    if (!Cc.medium) G *= fabsf(dot(Cc.ng, w));
    cn.c = c * xyz(cv.beta) * fc * G;
    if (mis) {
        const float b_bwd = pdf_to_area(sc_pdf(s, Cc, -w, Cc.wo), cpos, bv);
        const float d_fwd = pdf_to_area(sc_pdf(s, Cc, Cc.wo, -w), cpos, dvx);
        cn.c *= mis_weight(C, t, L, sl, a_bwd, b_bwd, c_fwd, d_fwd, k);
    }
    cn.ray = true;
    cn.o = bpos;
    cn.d = w;
    cn.tmax = dist - kEps;
}

}  // namespace

// One thread per (slot, strategy): the weighted contribution of the strategy
// as if unoccluded, and its visibility ray appended to the connection queue.
// Held to three waves per SIMD (168 registers, 120 bytes of spills) instead of
// two (175 registers): configuration 1 connections 181 -> 137 ms per 8-spp
// frame, the same film.
__global__ __launch_bounds__(kBdBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_bd_connect(DevScene s, BdPaths b,
                                                                                                 BdConsts k) {
    __shared__ BlockAppendLds<kBdBlock> ap;
    const int64_t total = int64_t(k.slots) * k.K;
    for (int64_t base = int64_t(blockIdx.x) * kBdBlock; base < total; base += int64_t(gridDim.x) * kBdBlock) {
        const int64_t e = base + threadIdx.x;
        BdConn cn;
        cn.c = f3s(0.0f);
        cn.splat = -1;
        cn.ray = false;
        if (e < total) connect_one(s, b, k, int(e / k.K), int(e % k.K), cn);
        if (cn.ray && (max3(cn.c) <= 0.0f || !(cn.tmax > kEps))) {
            cn.ray = false;
            cn.c = f3s(0.0f);
        }
        unsigned int q, unused;
        block_append2(ap, bd_counter(b, 2), cn.ray, bd_counter(b, 2), false, q, unused);
        if (e < total) {
            b.con[e] = f4(cn.c, __int_as_float(cn.splat));
            if (cn.ray) {
                b.sray[2 * size_t(q)] = f4(cn.o, kEps);
                b.sray[2 * size_t(q) + 1] = f4(cn.d, cn.tmax);
                b.sidx[q] = int(e);
            }
        }
    }
}

// Per slot, in strategy order: the camera-side radiance for the film, and the
// light-tracing splats into the fixed-point splat image.
__global__ __launch_bounds__(kBdBlock) void k_bd_gather(BdPaths b, BdConsts k) {
    const int slot = blockIdx.x * kBdBlock + threadIdx.x;
    if (slot >= k.slots) return;
    f3 L = f3s(0.0f);
    unsigned long long bad = 0;
    const float4* c = b.con + size_t(slot) * k.K;
    for (int j = 0; j < k.K; ++j) {
        const float4 v = c[j];
        if (v.x == 0.0f && v.y == 0.0f && v.z == 0.0f) continue;
        const int px = __float_as_int(v.w);
        if (px < 0) {
            L += xyz(v);
            continue;
        }
        if (!finite3(xyz(v))) {
            ++bad;
            continue;
        }
        const float ch[3] = {v.x, v.y, v.z};
        for (int q = 0; q < 3; ++q) {
            if (!(ch[q] > 0.0f)) continue;
            const double fx = double(fminf(ch[q], kBdSplatMax)) * kBdSplatScale + 0.5;
            atomicAdd(&b.splat[3 * size_t(px) + q], (unsigned long long)fx);
        }
    }
    b.rad[slot] = f4(L, 0.0f);
    if (bad) atomicAdd(&b.stats[9], bad);
}

__global__ void k_bd_snapshot(const double4* film, double* w0, int npix) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < npix) w0[p] = film[p].w;
}

__global__ void k_bd_fold(unsigned long long* splat, double4* film, const double* w0, int npix, double inv_spp) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    double4 f = film[p];
    const double s = (f.w - w0[p]) * inv_spp / kBdSplatScale;
    f.x += double(splat[3 * size_t(p)]) * s;
    f.y += double(splat[3 * size_t(p) + 1]) * s;
    f.z += double(splat[3 * size_t(p) + 2]) * s;
    film[p] = f;
    splat[3 * size_t(p)] = 0ull;
    splat[3 * size_t(p) + 1] = 0ull;
    splat[3 * size_t(p) + 2] = 0ull;
}

void launch_bd_start(const DevScene& s, const BdPaths& b, const BdConsts& k, hipStream_t st) {
    hipLaunchKernelGGL(k_bd_start, dim3((k.slots + kBdBlock - 1) / kBdBlock), dim3(kBdBlock), 0, st, s, b, k);
}
void launch_bd_extend(const DevScene& s, const BdPaths& b, const BdConsts& k, int cur, hipStream_t st) {
    const int64_t rays = 2 * int64_t(k.slots);
    const int blocks = int(std::min<int64_t>((rays + kBdBlock - 1) / kBdBlock, 16384));
    hipLaunchKernelGGL(k_bd_extend, dim3(blocks), dim3(kBdBlock), 0, st, s, b, k, cur);
}
void launch_bd_connect(const DevScene& s, const BdPaths& b, const BdConsts& k, hipStream_t st) {
    const int64_t total = int64_t(k.slots) * k.K;
    const int blocks = int(std::min<int64_t>((total + kBdBlock - 1) / kBdBlock, 65536));
    hipLaunchKernelGGL(k_bd_connect, dim3(blocks), dim3(kBdBlock), 0, st, s, b, k);
}
void launch_bd_gather(const BdPaths& b, const BdConsts& k, hipStream_t st) {
    hipLaunchKernelGGL(k_bd_gather, dim3((k.slots + kBdBlock - 1) / kBdBlock), dim3(kBdBlock), 0, st, b, k);
}
void launch_bd_snapshot(const double4* film, double* w0, int npix, hipStream_t st) {
    hipLaunchKernelGGL(k_bd_snapshot, dim3((npix + 255) / 256), dim3(256), 0, st, film, w0, npix);
}
void launch_bd_fold(const BdPaths& b, double4* film, const double* w0, int npix, int spp, hipStream_t st) {
    hipLaunchKernelGGL(k_bd_fold, dim3((npix + 255) / 256), dim3(256), 0, st, b.splat, film, w0, npix, 1.0 / double(spp));
}

}  // namespace trhip
