// Wavefront path-tracing kernels: camera ray generation, shading (surface and
// medium vertices, next-event estimation, MIS, Russian roulette), film
// accumulation in FP64 and resolve.
//
// Replaces PixelSamplerRenderer::render (src/renderer/pixel_sampler_renderer.cpp:24-109)
// and PathTraceRenderer::eval_pixel_li (src/renderer/path_trace_renderer.cpp:27-210)
// with the NEE/MIS helpers of src/renderer/direct_illumination.cpp.  The
// recursion of eval_pixel_li is unrolled into one kernel launch per bounce
// over a compacted queue of live paths:
//   * the reference traces, per vertex, a continuation ray and a separate
//     BSDF-sampled MIS ray; here the continuation ray is the MIS ray (its hit
//     on an emitter is weighted against the light-sampling pdf), which has the
//     same expectation with one closest-hit ray per vertex;
//   * NEE samples one light chosen by power instead of every light
//     (path_trace_renderer.cpp:141-153 loops over scene.lights);
//   * homogeneous media use delta tracking (spectral tracking for
//     chromatic coefficients) with the medium of a segment decided at its end
//     as in the reference (SurfaceIntersection::wo_medium,
//     src/core/intersection.hpp:36-41);
//   * the reference defects Q1, Q4, Q5, Q6 and Q23 (direct emission added
//     before the medium sample) are not reproduced.
#include <algorithm>

#include "tr_guide.h"
#include "tr_kernels.h"

namespace trhip {

// info.y layout: depth (8 bits) | specular bounces not counted in depth (8 bits) << 8 |
// medium scatter count (16 bits) << 16.  Each scatter event takes one depth, so the
// count stays below 256.
TR_D uint32_t pack_info(int depth, int sdepth, int scatter) {
    return uint32_t(depth) | (uint32_t(sdepth) << 8) | (uint32_t(scatter) << 16);
}


TR_D Pcg load_rng(const DevPaths& p, int slot, uint32_t pixel) {
    const uint2 s = p.rng[slot];
    Pcg r;
    r.state = (uint64_t(s.y) << 32) | s.x;
    r.inc = (uint64_t(pixel) << 1) | 1u;
    return r;
}
TR_D void store_rng(const DevPaths& p, int slot, const Pcg& r) {
    p.rng[slot] = make_uint2(uint32_t(r.state), uint32_t(r.state >> 32));
}

__global__ __launch_bounds__(256) void k_generate(DevScene s, DevPaths p, RenderConsts c, int num_paths, int cur) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        *queue_counter(p, cur) = (unsigned int)num_paths;
        *queue_counter(p, 1 - cur) = 0u;
        *queue_counter(p, kShadowCounter) = 0u;
        *queue_counter(p, kTailBlocksCounter) = 0u;
    }
    if (i >= num_paths) return;
    const int npix = c.width * c.height;
    const uint32_t pixel = uint32_t(i % npix);
    const uint32_t sample = uint32_t(c.sample_base + i / npix);
    Pcg rng = pcg_for_sample(c.seed, pixel, sample);
    const float fu = rng.uniform(), fv = rng.uniform(), lu = rng.uniform(), lv = rng.uniform();
    const int px = int(pixel % uint32_t(c.width)), py = int(pixel / uint32_t(c.width));
    // pixel_sampler_renderer.cpp:62-71: film position in [0,1)^2
    const float fx = (float(px) + fu) / float(c.width), fy = (float(py) + fv) / float(c.height);
    f3 o, d;
    camera_ray(s.cam, fx, fy, lu, lv, o, d);
    p.ray_o[i] = f4(o, 0.0f);
    p.ray_d[i] = f4(d, 3.0e38f);
    // throughput (1, 1, 1, pdf 0), radiance (0, 0, 0, no cell) and (pixel,
    // depth 0) are not stored: the first shading of path i starts from them
    // (first_path_state)
    store_rng(p, i, rng);
    p.queue[cur][i] = i;
    if (p.grec) p.grec[i] = -1;
}

// Transmittance of a homogeneous medium over a shadow segment: closed form,
// or a ratio-tracking estimate (Novak et al. 2014) with the medium majorant.
TR_D f3 shadow_tr(const DevMedium& m, float dist, int mode, Pcg& rng) {
    if (mode == 0) return mk(expf(-m.sigma_t.x * dist), expf(-m.sigma_t.y * dist), expf(-m.sigma_t.z * dist));
    f3 T = f3s(1.0f);
    float t = 0.0f;
    const f3 ratio = mk(1.0f - m.sigma_t.x / m.majorant, 1.0f - m.sigma_t.y / m.majorant, 1.0f - m.sigma_t.z / m.majorant);
    for (int k = 0; k < 4096; ++k) {
        t -= logf(1.0f - rng.uniform()) / m.majorant;
        if (t >= dist) break;
        T *= ratio;
        if (max3(T) <= 0.0f) break;
    }
    return T;
}

// Transmittance of a shadow segment from o along d: ratio tracking through a
// grid medium (in the kernels specialised for extended scenes), else the
// homogeneous estimate above.
template <bool EXT>
TR_D f3 segment_tr(const DevScene& s, const DevMedium& m, f3 o, f3 d, float dist, int mode, Pcg& rng) {
    if (EXT && m.grid >= 0) return f3s(grid_tr(s, m, o, d, dist, rng));
    return shadow_tr(m, dist, mode, rng);
}

struct ShadowOut {
    bool emit;
    f3 o;
    float tmin;
    f3 d;
    float tmax;
    f3 contrib;
};

// Training-record selection: a path records its surface vertices when this
// hash of its slot falls below the pass threshold (murmur3 finaliser).
TR_D unsigned int record_hash(unsigned int slot, unsigned int seed) {
    unsigned int h = slot * 0x9E3779B1u ^ seed;
    h ^= h >> 16; h *= 0x85EBCA6Bu;
    h ^= h >> 13; h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

// Guide mixture of a path slot from the inference kernel's head outputs.
TR_D bool load_mixture(const DevPaths& p, int slot, Mixture& mix) {
    float raw[5 * kGuideLobes];
    const float4* g = reinterpret_cast<const float4*>(p.graw + size_t(slot) * kGuideOut);
#pragma unroll
    for (int k = 0; k < 10; ++k) {
        const float4 v = g[k];
        raw[4 * k] = v.x; raw[4 * k + 1] = v.y; raw[4 * k + 2] = v.z; raw[4 * k + 3] = v.w;
    }
    if (!mixture_raw_valid(raw)) return false;
    mix = mixture_from_raw(raw);
    return true;
}

// Outputs of shade_path for the appends of the wavefront kernel.
struct ShadeOut {
    bool cont;                   // the path continues: its next ray, throughput and RNG are stored
    ShadowOut sh;                // light sample to test with a shadow ray
    bool g_used, g_fallback;     // guided vertex / fallback to the BSDF (non-finite head outputs)
    bool want_rec, want_nee_rec; // training records of this vertex
    float4 rec0, rec1, rec2, rec3, aux0;
    float4 nrec0, nrec1, nrec2, nrec3;
    float nee_weight;
    int lc_cell, lc_bin;         // light-cache training: cell and bin of the light sample (-1: none)
    float lc_val;                // its unoccluded contribution / sampling density (luminance)
};

// Light-cache mode of a shading kernel: 0 off, 1 sampling (the cache is
// trained), 2 training (light samples by power note their cell and bin),
// 3 decided per pass by DevLightCache::use / train (the guide kernels).
template <int LCM> TR_D bool lc_use(const DevPaths& p) { return LCM == 1 || (LCM == 3 && p.lc.use); }
template <int LCM> TR_D bool lc_train(const DevPaths& p) { return LCM == 2 || (LCM == 3 && p.lc.train); }

// Next-event light choice: by power, or (guided passes with a trained light
// cache) from the cache of the vertex's cell.  Draws the same three uniforms
// either way.  Training passes note the cell and bin of the power-based choice.
// here_cell: the cell of x when the cache is sampled.
template <int LCM>
TR_D int nee_pick(const DevScene& s, const DevPaths& p, f3 x, int here_cell, Pcg& rng, float& lu, float& lv, float& pmf,
                  int& cell, int& bin) {
    cell = -1;
    bin = -1;
    if (lc_use<LCM>(p)) {
        const int b = lc_sample_bin(p.lc, here_cell, rng.uniform(), pmf);
        lu = rng.uniform();
        lv = rng.uniform();
        if (lc_train<LCM>(p)) {
            cell = here_cell;
            bin = b;
        }
        return lc_bin_light(p.lc, b, lu, lv, pmf);
    }
    const int e = pick_light(s, rng.uniform(), pmf);
    lu = rng.uniform();
    lv = rng.uniform();
    if (lc_train<LCM>(p)) {
        cell = lc_cell(p.lc, x);
        bin = lc_bin_of_sample(p.lc, e, lu, lv);
    }
    return e;
}

// Selection probability of light `light` hit at barycentrics (b1, b2) for the
// vertex whose cell is prev_cell (-1: the power-based choice), per unit of the
// light's area parameters; light == num_area_lights: the environment.
template <int LCM>
TR_D float nee_select_pdf(const DevScene& s, const DevPaths& p, int prev_cell, int light, float b1, float b2) {
    if (lc_use<LCM>(p) && prev_cell >= 0) {
        if (light >= s.num_area_lights) return lc_pmf(p.lc, prev_cell, lc_env_bin(p.lc));
        return lc_pmf(p.lc, prev_cell, lc_bin_of_hit(p.lc, light, b1, b2)) * float(p.lc.strata * p.lc.strata);
    }
    return light >= s.num_area_lights ? s.light_pmf[s.num_area_lights] : s.lights[4 * light + 1].w;
}

// A light sample of a vertex with several (direct_light_sample_num > 1),
// in its fixed shadow slot e; w of the contribution = 0 until the shadow
// kernel finds the segment unoccluded.
TR_D void put_shadow(const DevPaths& p, size_t e, const ShadowOut& sh) {
    p.sh_o[e] = f4(sh.o, sh.tmin);
    p.sh_d[e] = f4(sh.d, sh.tmax);
    p.sh_c[e] = f4(sh.contrib, 0.0f);
}

// The state k_generate leaves implicit for path `slot` of a pass (its slot is
// its generation index): throughput (1, 1, 1) with no BSDF pdf, no radiance
// and no light-cache cell yet, its pixel, depth 0.
TR_D void first_path_state(const RenderConsts& c, int slot, float4& tp, float4& rp, uint2& info) {
    tp = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    rp = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
    info = make_uint2(uint32_t(slot) % uint32_t(c.width * c.height), pack_info(0, 0, 0));
}

// One vertex of the path in `slot`: reads its state and hit, adds emission,
// samples the continuation and a light, and stores the updated state.
// first: the path's first vertex, whose state is implicit (first_path_state).
// EXT: the scene has an image environment or textured materials (their
// code is compiled in only then).
template <bool GUIDE, bool EXT, int LCM, bool MULTI = false>
TR_D void shade_path(const DevScene& s, const DevPaths& p, const RenderConsts& c, int slot, bool first, ShadeOut& out) {
    bool& cont = out.cont;
    ShadowOut& sh = out.sh;
    bool &g_used = out.g_used, &g_fallback = out.g_fallback, &want_rec = out.want_rec, &want_nee_rec = out.want_nee_rec;
    float4 &rec0 = out.rec0, &rec1 = out.rec1, &rec2 = out.rec2, &rec3 = out.rec3, &aux0 = out.aux0;
    float4 &nrec0 = out.nrec0, &nrec1 = out.nrec1, &nrec2 = out.nrec2, &nrec3 = out.nrec3;
    float& nee_weight = out.nee_weight;
    cont = false;
    sh.emit = false;
    g_used = g_fallback = want_rec = want_nee_rec = false;
    rec0 = rec1 = rec2 = rec3 = aux0 = nrec0 = nrec1 = nrec2 = nrec3 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    nee_weight = 0.0f;
    out.lc_cell = out.lc_bin = -1;
    out.lc_val = 0.0f;
    int next_cell = -1;   // cell of this vertex for the next hit's MIS weight (guided passes with a light cache)
    // direct_light_sample_num light samples per vertex (MULTI): sample k goes to
    // shadow slot slot * nls + k, marked unused (tmax < 0) unless it carries light
    const int nls = MULTI ? c.nls : 1;
    const float fnls = float(nls);
    if (MULTI)
        for (int k = 0; k < nls; ++k) p.sh_d[size_t(slot) * nls + k] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    float4 next_o = make_float4(0, 0, 0, 0), next_d = make_float4(0, 0, 0, 0);
    const float4 ro = p.ray_o[slot], rd = p.ray_d[slot], h = p.hit[slot];
    float4 tp, rp;
    uint2 info;
    if (first) first_path_state(c, slot, tp, rp, info);
    else {
        tp = p.thr[slot];
        rp = p.rad[slot];
        info = p.info[slot];
    }
    const int depth = int(info.y & 0xffu);
    int sdepth = int((info.y >> 8) & 0xffu);
    int scatter = int(info.y >> 16);
    int next_depth = depth + 1;
    Pcg rng = load_rng(p, slot, info.x);
    f3 T = xyz(tp);
    const float prev_pdf = tp.w;
    f3 L = xyz(rp);
    const f3 o = xyz(ro), d = xyz(rd);
    const int prev_cell = __float_as_int(rp.w);
    const int tri = __float_as_int(h.y);
    bool alive = true;
    if (tri < 0) {
        // miss: the environment (environment_light.cpp:179-196)
        if (s.has_env) {
            float epdf;
            const f3 Le = env_eval<EXT>(s, d, epdf);
            float w = 1.0f;
            if (prev_pdf > 0.0f)
                w = power_heuristic(prev_pdf, nee_select_pdf<LCM>(s, p, prev_cell, s.num_area_lights, 0.0f, 0.0f) * epdf * fnls);
            L += T * Le * w;
        }
        alive = false;
    } else {
        const float t = h.x;
        const TriShade sr = load_shade(s, tri);
        const f3 ng = sr.ng;
        const float dn = dot(d, ng);
        // segment medium: the side of the hit surface the ray arrives from
        const int mid = dn <= 0.0f ? sr.med_out : sr.med_in;
        bool scattered = false;
        f3 sp_pos;
        if (EXT && mid >= 0 && s.media[mid].grid >= 0) {
            // heterogeneous medium: delta tracking against the grid maximum
            // (HeterogeneousMedium::sample, src/medium/heterogeneous.cpp:54-82),
            // ratio tracking when the run of scatters is at its limit
            const DevMedium m = s.media[mid];
            if (depth >= c.max_depth || scatter >= m.max_scatter) {
                T *= grid_tr(s, m, o, d, t, rng);
                if (scatter >= m.max_scatter) scatter = 0;
            } else {
                const float inv = 1.0f / m.majorant;
                float tt = 0.0f;
                for (int k = 0; k < (1 << 20); ++k) {
                    tt -= logf(1.0f - rng.uniform()) * inv;
                    if (tt >= t) break;
                    const f3 x = o + d * tt;
                    const float dens = grid_density(s, m, x);
                    if (rng.uniform() < dens * inv) {
                        if (!(m.white_indirect && scatter > 0)) T *= medium_albedo(s, m, x);
                        scattered = true;
                        sp_pos = x;
                        break;
                    }
                }
            }
        } else if (mid >= 0) {
            const DevMedium m = s.media[mid];
            if (depth >= c.max_depth || scatter >= m.max_scatter) {
                T *= mk(expf(-m.sigma_t.x * t), expf(-m.sigma_t.y * t), expf(-m.sigma_t.z * t));
                if (scatter >= m.max_scatter) scatter = 0;
            } else {
                // delta / spectral tracking with the majorant
                float tt = 0.0f;
                for (int k = 0; k < 100000; ++k) {
                    tt -= logf(1.0f - rng.uniform()) / m.majorant;
                    if (tt >= t) break;
                    const float xi = rng.uniform();
                    if (xi < m.p_absorb) { alive = false; break; }
                    if (xi < m.p_absorb + m.p_scatter) {
                        T *= m.sigma_s * (1.0f / (m.majorant * m.p_scatter));
                        scattered = true;
                        sp_pos = o + d * tt;
                        break;
                    }
                    T *= (f3s(m.majorant) - m.sigma_t) * (1.0f / (m.majorant * m.p_null));
                }
            }
        }
        if (!alive) {
        } else if (scattered) {
            // ---- medium vertex (path_trace_renderer.cpp:94-115) ----
            const DevMedium m = s.media[mid];
            const float mg = EXT ? medium_g(s, m, sp_pos) : m.g;   // the asymmetry at the collision
            ++scatter;
            const f3 wo = -d;
            if (lc_use<LCM>(p)) next_cell = lc_cell(p.lc, sp_pos);
            auto nee_medium = [&](int k) {
                float pmf, lu, lv;
                int lcell, lbin;
                const int e = nee_pick<LCM>(s, p, sp_pos, next_cell, rng, lu, lv, pmf, lcell, lbin);
                const LightSampleD ls = sample_light<EXT>(s, e, sp_pos, lu, lv);
                if (ls.pdf != 0.0f && pmf > 0.0f && !is_black(ls.L)) {
                    const float ph = hg_phase(mg, -dot(ls.wi, wo));
                    float w = 1.0f;
                    float lp = pmf;
                    if (ls.pdf > 0.0f) { lp *= ls.pdf; w = power_heuristic(lp * fnls, ph); }
                    f3 trs = f3s(1.0f);
                    if (ls.area) trs = segment_tr<EXT>(s, m, sp_pos, ls.wi, ls.dist, c.shadow_tr, rng);
                    const f3 contrib = T * ls.L * trs * (ph * w / (lp * fnls));
                    if (max3(contrib) > 0.0f && finite3(contrib)) {
                        if (LCM >= 2 && lbin >= 0) {
                            out.lc_cell = lcell;
                            out.lc_bin = lbin;
                            out.lc_val = fminf(lum(ls.L * trs) / lp, kLcMaxValue);
                        }
                        sh.emit = !MULTI;
                        sh.o = sp_pos;
                        sh.tmin = kEps;
                        sh.d = ls.wi;
                        sh.tmax = ls.area ? ls.dist - kEps : ls.dist;
                        sh.contrib = contrib;
                        if (MULTI) put_shadow(p, size_t(slot) * nls + k, sh);
                    }
                }
            };
            if (s.num_light_entries > 0) {
                if constexpr (MULTI) {
                    for (int k = 0; k < nls; ++k) nee_medium(k);
                } else {
                    nee_medium(0);
                }
            }
            float ppdf;
            const f3 wi = hg_sample(mg, wo, rng.uniform(), rng.uniform(), ppdf);
            // f / pdf = 1 for the phase function
            next_o = f4(sp_pos, 0.0f);
            next_d = f4(wi, 3.0e38f);
            tp = f4(T, ppdf);
            cont = true;
        } else {
            // ---- surface vertex ----
            // At the maximum depth the hit only completes the previous vertex's BSDF
            // (MIS) sample; after a specular vertex the reference has no such sample.
            if (sr.light >= 0 && dn < 0.0f && (depth < c.max_depth || prev_pdf > 0.0f)) {
                // emission with MIS against light sampling (direct_illumination.cpp:176-197)
                const float4 l0 = s.lights[4 * sr.light];
                float w = 1.0f;
                if (prev_pdf > 0.0f) {
                    const float lp = nee_select_pdf<LCM>(s, p, prev_cell, sr.light, h.z, h.w) * t * t / (l0.w * (-dn)) * fnls;
                    w = power_heuristic(prev_pdf, lp);
                }
                L += T * xyz(l0) * w;
            }
            if (depth >= c.max_depth) {
                alive = false;
            } else {
                const float b1 = h.z, b2 = h.w;
                const f3 pos = o + d * t;
                f3 ns = normalize(sr.n0 * (1.0f - b1 - b2) + sr.n1 * b1 + sr.n2 * b2);
                f3 tx, ty;
                shading_frame(s, tri, ns, tx, ty);
                DevMaterial mat = s.mats[sr.material];
                if (EXT) apply_material_textures(s, mat, tri, b1, b2, tx, ty, ns);
                auto to_local = [&](f3 v) { return mk(dot(v, tx), dot(v, ty), dot(v, ns)); };
                const f3 wo = -d;
                const f3 lwo = to_local(wo), lng = to_local(ng);
                Mixture mix;
                bool use_guide = false;
                const bool guide_here = GUIDE && mat.guide_share > 0.0f;
                if (guide_here && c.guide_frac > 0.0f && (c.guide_max_depth <= 0 || depth < c.guide_max_depth)) {
                    use_guide = load_mixture(p, slot, mix);
                    g_fallback = !use_guide;
                    g_used = use_guide;
                }
                const float frac = c.guide_frac * mat.guide_share;
                const bool rec_path = guide_here && c.guide_record &&
                                      record_hash(uint32_t(slot), c.rec_seed) < c.rec_threshold;
                const float su = rng.uniform(), u1 = rng.uniform(), u2 = rng.uniform();
                f3 wi, fs;
                float pdf;
                bool delta = false;
                // guide samples are folded into the hemisphere of wo about the shading normal
                const f3 nsd = dot(wo, ns) >= 0.0f ? ns : -ns;
                if (use_guide) {
                    float pdf_b;
                    if (rng.uniform() < frac) {
                        wi = fold_to_hemisphere(mixture_sample(mix, su, u1, u2), nsd);
                        const BsdfEval be = bsdf_eval(mat, to_local(wi), lwo, lng);
                        fs = be.f;
                        pdf_b = be.pdf;
                    } else {
                        const BsdfSample bs = bsdf_sample(mat, lwo, lng, su, u1, u2);
                        wi = normalize(tx * bs.wi.x + ty * bs.wi.y + ns * bs.wi.z);
                        fs = bs.f;
                        pdf_b = bs.pdf;
                    }
                    pdf = is_black(fs) ? 0.0f : frac * mixture_pdf_folded(mix, wi, nsd) + (1.0f - frac) * pdf_b;
                } else {
                    const BsdfSample bs = bsdf_sample(mat, lwo, lng, su, u1, u2);
                    wi = normalize(tx * bs.wi.x + ty * bs.wi.y + ns * bs.wi.z);
                    fs = bs.f;
                    pdf = bs.pdf;
                    delta = bs.delta;
                }
                // Next-event estimation does not depend on the continuation sample.  The
                // reference ends the path before estimating direct light when the
                // continuation sample fails (path_trace_renderer.cpp:128-131 before
                // :141-153), which loses that vertex's direct light; here the direct
                // light is estimated at every vertex and the path ends afterwards.
                {
                    // next-event estimation (direct_illumination.cpp:35-109); none at a
                    // specular vertex (path_trace_renderer.cpp:141)
                    if (lc_use<LCM>(p)) next_cell = lc_cell(p.lc, pos);
                    auto nee_surface = [&](int k) {
                        float pmf, lu, lv;
                        int lcell, lbin;
                        const int e = nee_pick<LCM>(s, p, pos, next_cell, rng, lu, lv, pmf, lcell, lbin);
                        const LightSampleD ls = sample_light<EXT>(s, e, pos, lu, lv);
                        if (ls.pdf != 0.0f && pmf > 0.0f && !is_black(ls.L)) {
                            const float dist = ls.area ? ls.dist - kEps : ls.dist;
                            if (dist > kEps) {
                                const BsdfEval be = bsdf_eval(mat, to_local(ls.wi), lwo, lng);
                                if (!is_black(be.f)) {
                                    float w = 1.0f;
                                    float lp = pmf;
                                    if (ls.pdf > 0.0f) {
                                        lp *= ls.pdf;
                                        const float sp = use_guide ? frac * mixture_pdf_folded(mix, ls.wi, nsd) +
                                                                         (1.0f - frac) * be.pdf
                                                                   : be.pdf;
                                        w = power_heuristic(lp * fnls, sp);
                                    }
                                    f3 trs = f3s(1.0f);
                                    if (ls.area) {
                                        const int lm = dot(ls.wi, ng) >= 0.0f ? sr.med_out : sr.med_in;
                                        if (lm >= 0) trs = segment_tr<EXT>(s, s.media[lm], pos, ls.wi, ls.dist, c.shadow_tr, rng);
                                    }
                                    const float cosg = fabsf(dot(ls.wi, ng));
                                    const f3 contrib = T * ls.L * be.f * trs * (cosg * w / (lp * fnls));
                                    if (max3(contrib) > 0.0f && finite3(contrib)) {
                                        if (LCM >= 2 && lbin >= 0) {
                                            out.lc_cell = lcell;
                                            out.lc_bin = lbin;
                                            out.lc_val = fminf(lum(ls.L * trs) / lp, kLcMaxValue);
                                        }
                                        sh.emit = !MULTI;
                                        sh.o = pos;
                                        sh.tmin = kEps;
                                        sh.d = ls.wi;
                                        sh.tmax = dist;
                                        sh.contrib = contrib;
                                        if (MULTI) put_shadow(p, size_t(slot) * nls + k, sh);
                                        if (rec_path) {
                                            // light-sample record: MIS-weighted f |cos| L / p_light,
                                            // kept if the shadow ray reaches the light
                                            nee_weight = lum(ls.L * be.f * trs) * (cosg * w / lp);
                                            want_nee_rec = nee_weight > 0.0f && isfinite(nee_weight);
                                            nrec0 = make_float4(pos.x, pos.y, pos.z, ns.x);
                                            nrec1 = make_float4(ns.y, ns.z, wo.x, wo.y);
                                            nrec2 = make_float4(wo.z, ls.wi.x, ls.wi.y, ls.wi.z);
                                            nrec3 = make_float4(float(depth), __int_as_float(slot), lp, 0.0f);
                                        }
                                    }
                                }
                            }
                        }
                    };
                    if (!delta && s.num_light_entries > 0) {
                        if constexpr (MULTI) {
                            for (int k = 0; k < nls; ++k) nee_surface(k);
                        } else {
                            nee_surface(0);
                        }
                    }
                }
                if (is_black(fs) || !(pdf >= kEps)) {
                    alive = false;
                } else {
                    // continuation (path_trace_renderer.cpp:155-163)
                    const f3 Tv = T;
                    T *= fs * (fabsf(dot(ng, wi)) / pdf);
                    const f3 org = dot(ng, wi) > 0.0f ? pos + ng * kEps : pos - ng * kEps;
                    next_o = f4(org, 0.0f);
                    next_d = f4(wi, 3.0e38f);
                    // a specular vertex passes pdf 0: the next hit's emission is not
                    // weighted against light sampling
                    tp = f4(T, delta ? 0.0f : pdf);
                    if (delta && sdepth < kMaxSpecularDepth) {
                        ++sdepth;
                        next_depth = depth;
                    }
                    cont = finite3(T);
                    if (rec_path && cont) {
                        // record (x, n, wo, wi); its weight lum((L_end - L_next) * rr / Tv)
                        // estimates f |cos| L_i / pdf, with emission MIS-weighted, once the
                        // path has ended
                        want_rec = true;
                        rec0 = make_float4(pos.x, pos.y, pos.z, ns.x);
                        rec1 = make_float4(ns.y, ns.z, wo.x, wo.y);
                        rec2 = make_float4(wo.z, wi.x, wi.y, wi.z);
                        rec3 = make_float4(float(depth), __int_as_float(slot), pdf, 0.0f);
                        aux0 = make_float4(Tv.x > 0.0f ? 1.0f / Tv.x : 0.0f, Tv.y > 0.0f ? 1.0f / Tv.y : 0.0f,
                                           Tv.z > 0.0f ? 1.0f / Tv.z : 0.0f, 0.0f);
                    }
                }
            }
        }
    }
    if (cont) {
        // Russian roulette before the next segment (path_trace_renderer.cpp:54-58)
        if (next_depth > c.min_depth) {
            if (rng.uniform() > 0.9f) cont = false;
            else {
                tp.x /= 0.9f; tp.y /= 0.9f; tp.z /= 0.9f;
                aux0.x *= 0.9f; aux0.y *= 0.9f; aux0.z *= 0.9f;
            }
        }
    }
    want_rec = want_rec && cont;
    if (GUIDE && c.guide_record) {
        // radiance gathered before this vertex closes the previous vertex's record
        const int pr = p.grec[slot];
        if (pr >= 0) p.rec_aux[2 * pr + 1] = make_float4(rp.x, rp.y, rp.z, 1.0f);
    }
    // rad.w: the cell of this vertex for the next hit's MIS weight (read only
    // by the kernels that sample the light cache)
    p.rad[slot] = f4(L, __int_as_float(next_cell));
    if (cont) {
        p.ray_o[slot] = next_o;
        p.ray_d[slot] = next_d;
        p.thr[slot] = tp;
        p.info[slot] = make_uint2(info.x, pack_info(next_depth, sdepth, scatter));
        store_rng(p, slot, rng);
    }
}

// GUIDE: surface vertices sample the continuation direction by one-sample MIS
// between the guide mixture (probability c.guide_frac times the material's
// guide_share) and the BSDF, and weight it, light samples and emitter hits
// with the combined pdf.  A vertex
// whose head outputs are not finite uses the plain BSDF estimator, drawing
// the same random numbers as the plain kernel.
// A workgroup takes a window of up to kShadeSortRounds x kShadeBlock queue
// entries, orders them by the shading class of their hit (kShadeClass*, a
// counting sort in LDS that keeps queue order within a class) and shades the
// window in that order, kShadeBlock paths per round.  Waves then hold paths
// that take the same branches: rays that left the scene finish in a few
// instructions instead of idling beside a Disney evaluation.  A path's result
// does not depend on which thread shades it, so the film is unchanged.
constexpr int kShadeSortRounds = 4;
template <int BLOCK>
struct ShadeSortLds {
    int slot[kShadeSortRounds * BLOCK];
    unsigned int cnt[kShadeClasses * kShadeSortRounds * (BLOCK / 64)];   // (class, round, wave)
};
static_assert(kShadeClasses * kShadeSortRounds * (kShadeBlock / 64) == 64, "one wave scans the class counts");

// Fills sort.slot with the window [base, base + rounds * BLOCK) of queue q
// (entries at or past n hold -1), ordered by class.  Every thread of the
// block calls it; it ends with a barrier.
template <int BLOCK>
TR_D void shade_sort_window(ShadeSortLds<BLOCK>& sort, const DevScene& s, const DevPaths& p, const int* q,
                            unsigned int n, unsigned int base, int rounds) {
    constexpr int WAVES = BLOCK / 64;
    const int lane = __lane_id(), wave = int(threadIdx.x) >> 6;
    const unsigned long long below = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
    int slot[kShadeSortRounds], key[kShadeSortRounds];
    unsigned int rank[kShadeSortRounds];
#pragma unroll
    for (int r = 0; r < kShadeSortRounds; ++r) {
        slot[r] = -1;
        key[r] = r < rounds ? int(kShadeClassMiss) : -1;   // -1: outside the window
        const unsigned int i = base + unsigned(r * BLOCK) + threadIdx.x;
        if (r < rounds && i < n) {
            slot[r] = q[i];
            const int tri = __float_as_int(p.hit[slot[r]].y);
            if (tri >= 0) key[r] = int(s.tri_cls[tri]);
        }
        rank[r] = 0;
#pragma unroll
        for (int k = 0; k < kShadeClasses; ++k) {
            const unsigned long long m = __ballot(key[r] == k);
            if (key[r] == k) rank[r] = (unsigned int)__popcll(m & below);
            if (lane == 0) sort.cnt[(k * kShadeSortRounds + r) * WAVES + wave] = (unsigned int)__popcll(m);
        }
    }
    __syncthreads();
    if (wave == 0) {
        // exclusive prefix over (class, round, wave): where each wave's run of a class starts
        const unsigned int own = sort.cnt[lane];
        unsigned int v = own;
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned int t = __shfl_up(v, off);
            if (lane >= off) v += t;
        }
        sort.cnt[lane] = v - own;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kShadeSortRounds; ++r)
        if (key[r] >= 0) sort.slot[sort.cnt[(key[r] * kShadeSortRounds + r) * WAVES + wave] + rank[r]] = slot[r];
    __syncthreads();
}

// The guided kernels shade in queue order: the window sort's registers would
// halve their occupancy.
template <bool GUIDE, bool EXT, int LCM, bool MULTI = false>
TR_D void shade_queue(const DevScene& s, const DevPaths& p, const RenderConsts& c, int cur) {
    constexpr bool SORT = !GUIDE;
    __shared__ BlockAppendLds<kShadeBlock> app;
    __shared__ ShadeSortLds<kShadeBlock> sort;
    const unsigned int n = *queue_counter(p, cur);
    const int* q = p.queue[cur];
    int* qn = p.queue[1 - cur];
    unsigned int n_guided = 0, n_fallback = 0;
    // Rounds per window: as many as keep every workgroup of the grid busy.
    const unsigned int per_block = n / (gridDim.x * unsigned(kShadeBlock));
    const int rounds = per_block < 1u ? 1 : (per_block > unsigned(kShadeSortRounds) ? kShadeSortRounds : int(per_block));
    const unsigned int window = unsigned(rounds) * kShadeBlock;
    const unsigned int stride = gridDim.x * window;
    const bool sorted = SORT && rounds >= c.shade_sort_rounds;
    // The loop bounds are uniform across the block: every thread reaches the
    // block-wide sort and appends.
    for (unsigned int base = blockIdx.x * window; base < n; base += stride) {
        if (sorted) shade_sort_window<kShadeBlock>(sort, s, p, q, n, base, rounds);
        for (int r = 0; r < rounds; ++r) {
            int slot_in;
            if (sorted) {
                // the round's waves take its 64-entry runs in a rotated order, so
                // that the costly classes (last in the window) do not gather on one SIMD
                const int rot = int(blockIdx.x + unsigned(r)) & (kShadeBlock / 64 - 1);
                slot_in = sort.slot[r * kShadeBlock + ((int(threadIdx.x) + 64 * rot) & (kShadeBlock - 1))];
            } else {
                const unsigned int i = base + unsigned(r * kShadeBlock) + threadIdx.x;
                slot_in = i < n ? q[i] : -1;
            }
            const bool active = slot_in >= 0;
            ShadeOut o;
            o.cont = false;
            o.sh.emit = false;
            o.g_used = o.g_fallback = o.want_rec = o.want_nee_rec = false;
            int slot = 0;
            if (active) {
                slot = slot_in;
                shade_path<GUIDE, EXT, LCM, MULTI>(s, p, c, slot, c.first_bounce != 0, o);
            }
            const bool cont = o.cont, g_used = o.g_used, g_fallback = o.g_fallback;
            const bool want_rec = o.want_rec, want_nee_rec = o.want_nee_rec;
            const ShadowOut& sh = o.sh;
            const float4 &rec0 = o.rec0, &rec1 = o.rec1, &rec2 = o.rec2, &rec3 = o.rec3, &aux0 = o.aux0;
            const float4 &nrec0 = o.nrec0, &nrec1 = o.nrec1, &nrec2 = o.nrec2, &nrec3 = o.nrec3;
            const float nee_weight = o.nee_weight;
            n_guided += g_used ? 1u : 0u;
            n_fallback += g_fallback ? 1u : 0u;
            int nee_ri = -1;
            if (GUIDE && c.guide_record) {
                unsigned int ri, nri;
                block_append2(app, p.rec_count, want_rec, p.rec_count, want_nee_rec, ri, nri);
                {
                    const bool stored = want_rec && ri < (unsigned int)p.rec_cap;
                    if (stored) {
                        p.rec[4 * ri] = rec0;
                        p.rec[4 * ri + 1] = rec1;
                        p.rec[4 * ri + 2] = rec2;
                        p.rec[4 * ri + 3] = rec3;
                        p.rec_aux[2 * ri] = aux0;
                        p.rec_aux[2 * ri + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    }
                    if (cont) p.grec[slot] = stored ? int(ri) : -1;
                }
                if (want_nee_rec && nri < (unsigned int)p.rec_cap) {
                    nee_ri = int(nri);
                    p.rec[4 * nri] = nrec0;
                    p.rec[4 * nri + 1] = nrec1;
                    p.rec[4 * nri + 2] = nrec2;
                    p.rec[4 * nri + 3] = nrec3;
                    p.rec_aux[2 * nri] = make_float4(nee_weight, 0.0f, 0.0f, 0.0f);
                    p.rec_aux[2 * nri + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                }
            }
            unsigned int qi, si;
            block_append2(app, queue_counter(p, 1 - cur), cont, queue_counter(p, kShadowCounter), sh.emit, qi, si);
            if (cont) qn[qi] = slot;
            if (sh.emit) {
                p.sh_o[si] = f4(sh.o, sh.tmin);
                p.sh_d[si] = f4(sh.d, sh.tmax);
                p.sh_c[si] = f4(sh.contrib, __int_as_float(slot));
                if (GUIDE && c.guide_record) p.sh_rec[si] = nee_ri;
                if (lc_train<LCM>(p))
                    p.sh_g[si] = make_float4(__int_as_float(o.lc_cell), __int_as_float(o.lc_bin), o.lc_val, 0.0f);
            }
        }
    }
    if (GUIDE) {
        // guided and fallback vertex counts: one atomic per wave and launch
        for (int off = 32; off > 0; off >>= 1) {
            n_guided += __shfl_xor(n_guided, off);
            n_fallback += __shfl_xor(n_fallback, off);
        }
        if (__lane_id() == 0) {
            if (n_guided) atomicAdd(&p.stats[6], (unsigned long long)n_guided);
            if (n_fallback) atomicAdd(&p.stats[5], (unsigned long long)n_fallback);
        }
    }
}

template <bool GUIDE, bool EXT, int LCM>
__global__ __launch_bounds__(kShadeBlock) void k_shade(DevScene s, DevPaths p, RenderConsts c, int cur) {
    shade_queue<GUIDE, EXT, LCM>(s, p, c, cur);
}
// The plain renderer with direct_light_sample_num > 1 light samples per vertex.
template <bool EXT>
__global__ __launch_bounds__(kShadeBlock) void k_shade_multi(DevScene s, DevPaths p, RenderConsts c, int cur) {
    shade_queue<false, EXT, 0, true>(s, p, c, cur);
}

// Plain scenes (no image environment, textures or grids) keep three waves
// per SIMD, also when the light cache adds its few registers.
template <int LCM>
__global__ __launch_bounds__(kShadeBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_shade3(DevScene s, DevPaths p,
                                                                                                 RenderConsts c, int cur) {
    shade_queue<false, false, LCM>(s, p, c, cur);
}

// The plain kernel of scenes with an image environment, textures or density
// grids, held to three waves per SIMD like the plain one.
template <int LCM>
__global__ __launch_bounds__(kShadeBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_shade3x(DevScene s, DevPaths p,
                                                                                                  RenderConsts c, int cur) {
    shade_queue<false, true, LCM>(s, p, c, cur);
}

// The last bounces of a plain pass in one launch: once the queue holds at
// most max_tail paths, each thread takes one and carries it to its end
// (closest hit, shading, shadow ray, next bounce).  Late bounces hold few
// paths, and launched one by one each costs as much as its slowest ray; here
// a path's bounces follow each other directly.  A path draws the same random
// numbers, makes the same decisions and adds the same contributions in the
// same order as in the per-bounce kernels, so the film is identical.  The
// last workgroup to finish empties the queue, so the per-bounce kernels that
// follow find no work; with more than max_tail paths nothing happens.
template <bool EXT, int LCM>
__global__ __launch_bounds__(kShadeBlock) void k_tail(DevScene s, DevPaths p, RenderConsts c, int cur,
                                                      unsigned int max_tail) {
    __shared__ TraceStack<kShadeBlock, kLdsStack> st;
    const int tid = threadIdx.x;
    const int gt = blockIdx.x * blockDim.x + tid;
    const int stride = gridDim.x * blockDim.x;
    const unsigned int queued = *queue_counter(p, cur);
    const bool run = queued <= max_tail;
    const unsigned int n = run ? queued : 0u;
    const int* q = p.queue[cur];
    unsigned long long n_ext = 0, n_sh = 0;
    for (unsigned int i = gt; i < n; i += stride) {
        const int slot = q[i];
        bool cont = true;
        bool first = c.first_bounce != 0;
        while (cont) {
            const float4 a = p.ray_o[slot], b = p.ray_d[slot];
            TraceHit h;
            trace<false, false, kShadeBlock, kLdsStack>(s.tr, xyz(a), a.w, xyz(b), b.w, st, tid, p.ovf_node + gt,
                                                        p.ovf_t + gt, stride, h);
            p.hit[slot] = make_float4(h.t, __int_as_float(h.tri), h.b1, h.b2);
            ++n_ext;
            ShadeOut o;
            shade_path<false, EXT, LCM>(s, p, c, slot, first, o);
            first = false;
            if (o.sh.emit) {
                ++n_sh;
                TraceHit hs;
                if (!trace<true, false, kShadeBlock, kLdsStack>(s.tr, o.sh.o, o.sh.tmin, o.sh.d, o.sh.tmax, st, tid,
                                                                p.ovf_node + gt, p.ovf_t + gt, stride, hs)) {
                    float4 r = p.rad[slot];
                    r.x += o.sh.contrib.x; r.y += o.sh.contrib.y; r.z += o.sh.contrib.z;
                    p.rad[slot] = r;
                    if (lc_train<LCM>(p) && o.lc_cell >= 0 && o.lc_val > 0.0f)
                        lc_record(p.lc, o.lc_cell, o.lc_bin, o.lc_val);
                }
            }
            cont = o.cont;
        }
    }
    // ray counts: one atomic per wave and counter
    for (int off = 32; off > 0; off >>= 1) {
        n_ext += __shfl_xor(n_ext, off);
        n_sh += __shfl_xor(n_sh, off);
    }
    if (__lane_id() == 0) {
        if (n_ext) {
            atomicAdd(&p.stats[0], n_ext);
            atomicAdd(&p.stats[8], n_ext);
        }
        if (n_sh) atomicAdd(&p.stats[1], n_sh);
    }
    // every workgroup has read the queue length before it counts itself done
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int* done = queue_counter(p, kTailBlocksCounter);
        if (atomicAdd(done, 1u) == gridDim.x - 1) {
            *done = 0u;
            if (run) *queue_counter(p, cur) = 0u;
        }
    }
}

// Film accumulation for one pass (Tile::add_sample + Film::merge_film_tile,
// src/core/render.hpp:37-57, 90-100): every pixel gathers the samples of the
// pass that its filter footprint covers, in a fixed order, into FP64 sums —
// deterministic without atomics.  A sample at film position (qx + fu, qy + fv)
// reaches pixel px when |qx + fu - px - 0.5| <= r, so only pixels with
// |qx - px| <= floor(r + 0.5) can contribute.  Sample positions are re-derived
// from the sample's RNG stream (its first two numbers).  A sample whose
// radiance is not finite counts as black with its filter weight, as
// eval_pixel_li returns black for an invalid L
// (src/renderer/path_trace_renderer.cpp:199-203) and the film adds it.
__global__ __launch_bounds__(256) void k_accumulate(DevPaths p, RenderConsts c, double4* film) {
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    const int npix = c.width * c.height;
    if (pix >= npix) return;
    const int px = pix % c.width, py = pix / c.width;
    double4 acc = film[pix];
    unsigned long long bad = 0;
    // a box of radius <= 1/2 keeps every sample in its own pixel
    const bool own_only = c.filter == 0 && c.filter_radius <= 0.5f;
    const int R = own_only ? 0 : int(floorf(c.filter_radius + 0.5f));
    const float e = expf(-c.filter_alpha * c.filter_radius * c.filter_radius);
    for (int sidx = 0; sidx < c.spp_in_pass; ++sidx) {
        for (int dy = -R; dy <= R; ++dy) {
            const int qy = py + dy;
            if (qy < 0 || qy >= c.height) continue;
            for (int dx = -R; dx <= R; ++dx) {
                const int qx = px + dx;
                if (qx < 0 || qx >= c.width) continue;
                const int qpix = qy * c.width + qx;
                const int slot = sidx * npix + qpix;
                float4 L = p.rad[slot];
                const bool fin = isfinite(L.x) && isfinite(L.y) && isfinite(L.z);
                if (!fin) {
                    if (dx == 0 && dy == 0) ++bad;
                    L = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                }
                float w = 1.0f;
                if (!own_only) {
                    Pcg rng = pcg_for_sample(c.seed, uint32_t(qpix), uint32_t(c.sample_base + sidx));
                    const float fu = rng.uniform(), fv = rng.uniform();
                    const float ox = float(qx) + fu - (float(px) + 0.5f), oy = float(qy) + fv - (float(py) + 0.5f);
                    if (!(fabsf(ox) <= c.filter_radius && fabsf(oy) <= c.filter_radius)) continue;
                    if (c.filter != 0) {
                        // GaussianFilter::eval (src/filter/gaussian_filter.cpp:20-30)
                        const float gx = fmaxf(0.0f, expf(-c.filter_alpha * ox * ox) - e);
                        const float gy = fmaxf(0.0f, expf(-c.filter_alpha * oy * oy) - e);
                        w = gx * gy;
                    }
                }
                if (w == 0.0f) continue;
                acc.x += double(L.x) * w;
                acc.y += double(L.y) * w;
                acc.z += double(L.z) * w;
                acc.w += double(w);
            }
        }
    }
    film[pix] = acc;
    if (bad) atomicAdd(&p.stats[2], bad);
}

__global__ void k_resolve(const double4* film, float* image, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 f = film[i];
    float r = 0.0f, g = 0.0f, b = 0.0f;
    if (f.w > 0.0) { r = float(f.x / f.w); g = float(f.y / f.w); b = float(f.z / f.w); }
    image[3 * i] = r; image[3 * i + 1] = g; image[3 * i + 2] = b;
}

// The film as 8-bit RGB in one pass: resolve (sums / weight), the ACES
// curve of exposure * colour (src/post_processor/aces.cpp) when exposure >= 0,
// colour^inv_gamma (src/post_processor/gamma.cpp) when inv_gamma > 0, then
// clamp(int(255 c), 0, 255) as src/main.cpp:290-299 quantises; one pixel
// (32 bytes in, 3 out) per thread.
TR_D float aces_curve(float x) {
    return fminf(fmaxf((x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f), 0.0f), 1.0f);
}
TR_D unsigned char to_byte(float c, float exposure, float inv_gamma) {
    if (exposure >= 0.0f) c = aces_curve(c * exposure);
    if (inv_gamma > 0.0f) c = powf(c, inv_gamma);
    const float v = c * 255.0f;
    if (!(v > 0.0f)) return 0;            // also NaN
    return v >= 255.0f ? 255 : (unsigned char)int(v);
}
__global__ void k_film_ldr(const double4* film, unsigned char* rgb, int n, float exposure, float inv_gamma) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 f = film[i];
    float r = 0.0f, g = 0.0f, b = 0.0f;
    if (f.w > 0.0) { r = float(f.x / f.w); g = float(f.y / f.w); b = float(f.z / f.w); }
    rgb[3 * size_t(i)] = to_byte(r, exposure, inv_gamma);
    rgb[3 * size_t(i) + 1] = to_byte(g, exposure, inv_gamma);
    rgb[3 * size_t(i) + 2] = to_byte(b, exposure, inv_gamma);
}

__global__ void k_bsdf_eval(const DevMaterial* mats, const int* ids, const float* wi, const float* wo, const float* ng,
                            int64_t n, float* f, float* pdf) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DevMaterial m = mats[ids[i]];
    const f3 g = ng ? mk(ng[3 * i], ng[3 * i + 1], ng[3 * i + 2]) : mk(0.0f, 0.0f, 1.0f);
    const BsdfEval e = bsdf_eval(m, mk(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]), mk(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), g);
    f[3 * i] = e.f.x; f[3 * i + 1] = e.f.y; f[3 * i + 2] = e.f.z;
    pdf[i] = e.pdf;
}

__global__ void k_bsdf_sample(const DevMaterial* mats, const int* ids, const float* wo, const float* ng, const float* u,
                              int64_t n, float* wi, float* f, float* pdf) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DevMaterial m = mats[ids[i]];
    const f3 g = ng ? mk(ng[3 * i], ng[3 * i + 1], ng[3 * i + 2]) : mk(0.0f, 0.0f, 1.0f);
    const BsdfSample s = bsdf_sample(m, mk(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), g, u[3 * i], u[3 * i + 1], u[3 * i + 2]);
    const bool ok = s.pdf > 0.0f && !is_black(s.f);
    wi[3 * i] = s.wi.x; wi[3 * i + 1] = s.wi.y; wi[3 * i + 2] = s.wi.z;
    f[3 * i] = s.f.x; f[3 * i + 1] = s.f.y; f[3 * i + 2] = s.f.z;
    pdf[i] = ok ? s.pdf : 0.0f;
}

static int shade_blocks(int max_paths) {
    int b = (max_paths + kShadeBlock - 1) / kShadeBlock;
    if (b < 1) b = 1;
    if (b > 4096) b = 4096;
    return b;
}

void launch_generate(const DevScene& s, const DevPaths& p, const RenderConsts& c, int num_paths, int cur, hipStream_t st) {
    const int blocks = (num_paths + 255) / 256;
    hipLaunchKernelGGL(k_generate, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, st, s, p, c, num_paths, cur);
}
// Light-cache mode of the plain kernels for this pass (see lc_use / lc_train).
static int plain_lc_mode(const DevPaths& p) {
    if (p.lc.grid <= 0) return 0;
    if (p.lc.train && p.lc.use) return 3;
    return p.lc.train ? 2 : (p.lc.use ? 1 : 0);
}
template <bool EXT, int LCM>
static void launch_shade_plain(const DevScene& s, const DevPaths& p, const RenderConsts& c, int cur, dim3 g, hipStream_t st) {
    if constexpr (EXT) hipLaunchKernelGGL((k_shade3x<LCM>), g, dim3(kShadeBlock), 0, st, s, p, c, cur);
    else hipLaunchKernelGGL((k_shade3<LCM>), g, dim3(kShadeBlock), 0, st, s, p, c, cur);
}
void launch_shade(const DevScene& s, const DevPaths& p, const RenderConsts& c, int cur, int max_paths, bool guide,
                  hipStream_t st) {
    const dim3 g(shade_blocks(max_paths)), b(kShadeBlock);
    const bool img = s.env.w > 0 || s.has_tex || s.has_grid;
    if (c.nls > 1) {
        if (img) hipLaunchKernelGGL((k_shade_multi<true>), g, b, 0, st, s, p, c, cur);
        else hipLaunchKernelGGL((k_shade_multi<false>), g, b, 0, st, s, p, c, cur);
        return;
    }
    if (guide) {
        if (img) hipLaunchKernelGGL((k_shade<true, true, 3>), g, b, 0, st, s, p, c, cur);
        else hipLaunchKernelGGL((k_shade<true, false, 3>), g, b, 0, st, s, p, c, cur);
        return;
    }
    switch (plain_lc_mode(p) + (img ? 4 : 0)) {
        case 0: launch_shade_plain<false, 0>(s, p, c, cur, g, st); break;
        case 1: launch_shade_plain<false, 1>(s, p, c, cur, g, st); break;
        case 2: launch_shade_plain<false, 2>(s, p, c, cur, g, st); break;
        case 3: launch_shade_plain<false, 3>(s, p, c, cur, g, st); break;
        case 4: launch_shade_plain<true, 0>(s, p, c, cur, g, st); break;
        case 5: launch_shade_plain<true, 1>(s, p, c, cur, g, st); break;
        case 6: launch_shade_plain<true, 2>(s, p, c, cur, g, st); break;
        default: launch_shade_plain<true, 3>(s, p, c, cur, g, st); break;
    }
}
void launch_tail(const DevScene& s, const DevPaths& p, const RenderConsts& c, int cur, int max_tail, hipStream_t st) {
    // one thread per path, and no more threads than the traversal spill area has columns
    int blocks = std::min((max_tail + kShadeBlock - 1) / kShadeBlock, p.ovf_threads / kShadeBlock);
    if (blocks < 1) blocks = 1;
    const bool img = s.env.w > 0 || s.has_tex || s.has_grid;
    const dim3 g(blocks), b(kShadeBlock);
    const unsigned int mt = (unsigned int)max_tail;
    switch (plain_lc_mode(p) + (img ? 4 : 0)) {
        case 0: hipLaunchKernelGGL((k_tail<false, 0>), g, b, 0, st, s, p, c, cur, mt); break;
        case 1: hipLaunchKernelGGL((k_tail<false, 1>), g, b, 0, st, s, p, c, cur, mt); break;
        case 2: hipLaunchKernelGGL((k_tail<false, 2>), g, b, 0, st, s, p, c, cur, mt); break;
        case 3: hipLaunchKernelGGL((k_tail<false, 3>), g, b, 0, st, s, p, c, cur, mt); break;
        case 4: hipLaunchKernelGGL((k_tail<true, 0>), g, b, 0, st, s, p, c, cur, mt); break;
        case 5: hipLaunchKernelGGL((k_tail<true, 1>), g, b, 0, st, s, p, c, cur, mt); break;
        case 6: hipLaunchKernelGGL((k_tail<true, 2>), g, b, 0, st, s, p, c, cur, mt); break;
        default: hipLaunchKernelGGL((k_tail<true, 3>), g, b, 0, st, s, p, c, cur, mt); break;
    }
}
void launch_accumulate(const DevPaths& p, const RenderConsts& c, double4* film, hipStream_t st) {
    const int n = c.width * c.height;
    hipLaunchKernelGGL(k_accumulate, dim3((n + 255) / 256), dim3(256), 0, st, p, c, film);
}
void launch_resolve(const double4* film, float* image, int num_pixels, hipStream_t st) {
    hipLaunchKernelGGL(k_resolve, dim3((num_pixels + 255) / 256), dim3(256), 0, st, film, image, num_pixels);
}
void launch_film_ldr(const double4* film, unsigned char* rgb, int num_pixels, float exposure, float inv_gamma,
                     hipStream_t st) {
    hipLaunchKernelGGL(k_film_ldr, dim3((num_pixels + 255) / 256), dim3(256), 0, st, film, rgb, num_pixels, exposure,
                       inv_gamma);
}
void launch_bsdf_eval(const DevMaterial* mats, const int* ids, const float* wi, const float* wo, const float* ng, int64_t n,
                      float* f, float* pdf, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_bsdf_eval, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, mats, ids, wi, wo, ng, n, f, pdf);
}
void launch_bsdf_sample(const DevMaterial* mats, const int* ids, const float* wo, const float* ng, const float* u, int64_t n,
                        float* wi, float* f, float* pdf, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_bsdf_sample, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, mats, ids, wo, ng, u, n, wi, f, pdf);
}

// Environment light of the uploaded scene: samples (u: n*2) and lookups
// (dirs: n*3) through the renderer's own functions.
__global__ void k_env_sample(DevScene s, const float* u, int64_t n, float* wi, float* pdf, float* L) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const LightSampleD r = sample_light<true>(s, s.num_area_lights, mk(0.0f, 0.0f, 0.0f), u[2 * i], u[2 * i + 1]);
    wi[3 * i] = r.wi.x; wi[3 * i + 1] = r.wi.y; wi[3 * i + 2] = r.wi.z;
    L[3 * i] = r.L.x; L[3 * i + 1] = r.L.y; L[3 * i + 2] = r.L.z;
    pdf[i] = r.pdf;
}
__global__ void k_env_eval(DevScene s, const float* d, int64_t n, float* L, float* pdf) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float p;
    const f3 e = env_eval<true>(s, normalize(mk(d[3 * i], d[3 * i + 1], d[3 * i + 2])), p);
    L[3 * i] = e.x; L[3 * i + 1] = e.y; L[3 * i + 2] = e.z;
    pdf[i] = p;
}
__global__ void k_grid_albedo_g(DevScene s, int medium, const float* x, int64_t n, float* albedo, float* g) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DevMedium& m = s.media[medium];
    const f3 p = mk(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    const f3 a = medium_albedo(s, m, p);
    albedo[3 * i] = a.x; albedo[3 * i + 1] = a.y; albedo[3 * i + 2] = a.z;
    g[i] = medium_g(s, m, p);
}
__global__ void k_grid_density(DevScene s, int medium, const float* x, int64_t n, float* out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = grid_density(s, s.media[medium], mk(x[3 * i], x[3 * i + 1], x[3 * i + 2]));
}
__global__ void k_tex_eval(DevScene s, int texture, const float* uv, int64_t n, float* out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f3 c = tex_lookup(s, texture - 1, uv[2 * i], uv[2 * i + 1]);
    out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
}
// Light-cache distributions from the training sums, one thread per cell:
// p(bin) = a sum(bin) / sum(cell) + (1 - a) default(bin), a = (1 - eps) h /
// (h + h0) with h the cell's unoccluded training samples, where the default
// is the power-based light choice spread evenly over a light's strata (cells
// without an unoccluded sample keep the default alone), then Vose's alias
// table over the bins; `work` holds one cell's small and large lists (small
// from the front, large from the back).
__global__ __launch_bounds__(256) void k_lc_build(DevLightCache lc, const float* light_pmf, float eps, int cells,
                                                  float* pmf, float2* table, int* work) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= cells) return;
    const int B = lc.bins, S2 = lc.strata * lc.strata, env = lc.num_area * S2;
    const size_t base = size_t(cell) * B;
    unsigned long long tot = 0;
    for (int b = 0; b < B; ++b) tot += lc.sum[base + b];
    const double inv = tot ? 1.0 / double(tot) : 0.0;
    const float h = float(lc.hits[cell]);
    const float a = tot ? (1.0f - eps) * h / (h + kLcHits0) : 0.0f;
    for (int b = 0; b < B; ++b) {
        const float def = b < env ? light_pmf[b / S2] / float(S2) : light_pmf[lc.num_area + (b - env)];
        const float q = a * float(double(lc.sum[base + b]) * inv) + (1.0f - a) * def;
        pmf[base + b] = q;
    }
    int* w = work + base;
    int ns = 0, nl = 0;
    for (int b = 0; b < B; ++b) {
        const float q = pmf[base + b] * float(B);
        table[base + b] = make_float2(q, __int_as_float(b));
        if (q < 1.0f) w[ns++] = b;
        else w[B - 1 - nl++] = b;
    }
    while (ns > 0 && nl > 0) {
        const int sm = w[--ns];
        const int lg = w[B - nl];
        const float qs = table[base + sm].x;
        table[base + sm] = make_float2(qs, __int_as_float(lg));
        const float ql = (table[base + lg].x + qs) - 1.0f;
        table[base + lg].x = ql;
        if (ql < 1.0f) {
            --nl;
            w[ns++] = lg;
        }
    }
    while (nl > 0) { const int lg = w[B - nl--]; table[base + lg] = make_float2(1.0f, __int_as_float(lg)); }
    while (ns > 0) { const int sm = w[--ns]; table[base + sm] = make_float2(1.0f, __int_as_float(sm)); }
}
void launch_lc_build(const DevLightCache& lc, const float* light_pmf, float eps, int cells, float* pmf, float2* table,
                     int* work, hipStream_t st) {
    hipLaunchKernelGGL(k_lc_build, dim3((cells + 255) / 256), dim3(256), 0, st, lc, light_pmf, eps, cells, pmf, table,
                       work);
}

// Geometric normal of every leaf-ordered triangle into its shading record
// (the fourth float4), computed on the device as the shading code computed it
// from the vertices.
__global__ void k_geometric_normals(const float4* tris, float4* tri_shade, int64_t nt) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const float4* tv = tris + size_t(i) * 3;
    const f3 v0 = xyz(tv[0]), v1 = xyz(tv[1]), v2 = xyz(tv[2]);
    tri_shade[size_t(i) * kShadeRecord + 3] = f4(normalize(cross(v1 - v0, v2 - v0)), 0.0f);
}
void launch_geometric_normals(const float4* tris, float4* tri_shade, int64_t nt, hipStream_t st) {
    if (nt <= 0) return;
    hipLaunchKernelGGL(k_geometric_normals, dim3(unsigned((nt + 255) / 256)), dim3(256), 0, st, tris, tri_shade, nt);
}

// Test batches through the renderer's own device functions: camera rays
// (ThinLensCamera::generate_ray), light selection (pick_light), HG sampling,
// and film samples with given radiance for k_accumulate.
__global__ void k_shading_frames(DevScene s, const int* tri, const float* bary, int64_t n, float* out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float b1 = bary[2 * i], b2 = bary[2 * i + 1];
    const TriShade sr = load_shade(s, tri[i]);
    const f3 ns = normalize(sr.n0 * (1.0f - b1 - b2) + sr.n1 * b1 + sr.n2 * b2);
    f3 tx, ty;
    shading_frame(s, tri[i], ns, tx, ty);
    const float v[9] = {tx.x, tx.y, tx.z, ty.x, ty.y, ty.z, ns.x, ns.y, ns.z};
    for (int k = 0; k < 9; ++k) out[9 * i + k] = v[k];
}
__global__ void k_camera_rays(DevScene s, const float* film_uv, const float* lens_uv, int64_t n, float4* rays) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 o, d;
    camera_ray(s.cam, film_uv[2 * i], film_uv[2 * i + 1], lens_uv[2 * i], lens_uv[2 * i + 1], o, d);
    rays[2 * i] = f4(o, 0.0f);
    rays[2 * i + 1] = f4(d, 3.0e38f);
}
__global__ void k_light_pick(DevScene s, const float* u, int64_t n, int* idx, float* pmf) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float pm;
    idx[i] = pick_light(s, u[i], pm);
    pmf[i] = pm;
}
__global__ void k_hg_sample(float g, const float* wo, const float* u, int64_t n, float* wi, float* pdf) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float pp;
    const f3 w = hg_sample(g, mk(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), u[2 * i], u[2 * i + 1], pp);
    wi[3 * i] = w.x; wi[3 * i + 1] = w.y; wi[3 * i + 2] = w.z;
    pdf[i] = pp;
}
__global__ void k_unpack_radiance(const float* L, int64_t n, float4* rad) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    rad[i] = make_float4(L[3 * i], L[3 * i + 1], L[3 * i + 2], 0.0f);
}
void launch_shading_frames(const DevScene& s, const int* tri, const float* bary, int64_t n, float* frames,
                           hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_shading_frames, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, s, tri, bary, n, frames);
}
void launch_camera_rays(const DevScene& s, const float* film_uv, const float* lens_uv, int64_t n, float4* rays,
                        hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_camera_rays, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, s, film_uv, lens_uv, n, rays);
}
void launch_light_pick(const DevScene& s, const float* u, int64_t n, int* idx, float* pmf, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_light_pick, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, s, u, n, idx, pmf);
}
void launch_hg_sample(float g, const float* wo, const float* u, int64_t n, float* wi, float* pdf, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_hg_sample, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, g, wo, u, n, wi, pdf);
}
void launch_unpack_radiance(const float* L, int64_t n, float4* rad, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_unpack_radiance, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, L, n, rad);
}
void launch_grid_albedo_g(const DevScene& s, int medium, const float* x, int64_t n, float* albedo, float* g,
                          hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_grid_albedo_g, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, s, medium, x, n, albedo, g);
}
void launch_grid_density(const DevScene& s, int medium, const float* x, int64_t n, float* out, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_grid_density, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, s, medium, x, n, out);
}
void launch_tex_eval(const DevScene& s, int texture, const float* uv, int64_t n, float* out, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_tex_eval, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, s, texture, uv, n, out);
}
void launch_env_sample(const DevScene& s, const float* u, int64_t n, float* wi, float* pdf, float* L, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_env_sample, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, s, u, n, wi, pdf, L);
}
void launch_env_eval(const DevScene& s, const float* d, int64_t n, float* L, float* pdf, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_env_eval, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, s, d, n, L, pdf);
}

}  // namespace trhip
