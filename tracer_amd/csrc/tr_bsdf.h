// Materials on the device: Disney BSDF, Lambert and Phong, evaluated in the
// shading frame (shading normal = +z) with the geometric normal given in that
// frame.
//
// Replaces DisneyBSDF (src/material/disney.cpp:14-548) with its helpers
// (src/material/utility/microfacet.cpp, src/utility/reflection.hpp), the
// Phong AggregateBSDF (src/material/phong.cpp:29-70,
// src/material/aggregate_bsdf.hpp, src/material/bxdf/diffuse_bxdf.cpp,
// src/material/bxdf/phong_specular_bxdf.cpp) — Lambert is its ks = 0 case —
// plus the black-fringe handling of LocalBSDF (src/core/bsdf.hpp:114-165).
//
// The formulas are the reference's, rewritten for FP32 accuracy and without
// transcendental round trips:
//   * sin/cos of a direction's azimuth come from x/r and y/r instead of
//     sin(atan2(y, x)) and cos(atan2(y, x)) (disney.cpp:74-89, algebraically
//     identical);
//   * sin^2(theta) comes from x^2 + y^2 instead of 1 - cos^2 (no
//     cancellation near the pole), so D and G keep their relative accuracy
//     at roughness 0.05 (alpha = 0.0025).
// Sampling is exclusive lobe selection with the macro-Fresnel split for
// lwo.z < 0, transmission pdf uses cos(theta_h), refraction is about the
// sampled microfacet normal, and Lambert eval carries the shading-normal
// correction that its sample() applies — the corrections that make the
// returned pdf the sampling density (reference defects Q2, Q3, Q21, Q22).
#pragma once
#include "tr_device.h"

namespace trhip {

enum : int { MAT_LAMBERT = 0, MAT_DISNEY = 1, MAT_PHONG = 2, MAT_GLASS = 3 };

// Per-material constants, precomputed on the host from the create_disney
// parameters exactly as the DisneyBSDF constructor does (disney.cpp:347-399).
// Phong uses C = kd + ks (the aggregate albedo of the black-fringe fallback),
// Ctint = kd, spec_scale = ks, ns, w_diffuse and w_spec, all after the
// energy normalisation of Phong::shading.  Glass uses C = the reflection
// colour, spec_scale = the refraction colour and ior.
struct DevMaterial {
    int type;
    float metallic, roughness, transmission;
    f3 C; float ior;
    f3 Ctint; float sheen;
    f3 spec_scale; float sheen_tint;
    float spec_tint, clearcoat, ax, ay;
    float tax, tay, cc_rough, w_diffuse;
    float w_trans, w_spec, w_cc;
    float guide_share;   // guide mixture probability at this material, as a share of the guide fraction (host, make_material)
    float ns;            // Phong exponent
    int tex;             // 0 or 1 + index of the texture giving kd / base_color per hit
    int texset;          // 0 or 1 + row of DevScene::mat_tex / mat_raw: Disney parameters or a normal map per hit
};

// Parameters of create_disney (src/factory/material.hpp:37-52) at one surface
// point: the values of its thirteen textures there.
struct DisneyIn {
    f3 base_color;
    float metallic, roughness, transmission, transmission_roughness, ior;
    f3 specular_scale;
    float specular_tint, anisotropic, sheen, sheen_tint, clearcoat, clearcoat_gloss;
};

// DisneyBSDF constructor (src/material/disney.cpp:347-399): lobe constants and
// lobe-selection weights from the parameters.  type, guide_share, ns, tex and
// texset are left as they are.
TR_HD void disney_setup(const DisneyIn& p, DevMaterial& d) {
    d.C = p.base_color;
    const float l = lum(d.C);
    d.Ctint = l > 0.0f ? d.C * (1.0f / l) : f3s(1.0f);
    d.metallic = p.metallic;
    d.roughness = p.roughness;
    d.transmission = p.transmission;
    d.ior = fmaxf(1.01f, p.ior);
    d.spec_scale = p.specular_scale;
    d.spec_tint = p.specular_tint;
    d.sheen = p.sheen;
    d.sheen_tint = p.sheen_tint;
    d.clearcoat = p.clearcoat;
    const float aspect = p.anisotropic > 0.0f ? sqrtf(1.0f - 0.9f * p.anisotropic) : 1.0f;
    d.ax = fmaxf(0.001f, p.roughness * p.roughness / aspect);
    d.ay = fmaxf(0.001f, p.roughness * p.roughness * aspect);
    d.tax = fmaxf(0.001f, p.transmission_roughness * p.transmission_roughness / aspect);
    d.tay = fmaxf(0.001f, p.transmission_roughness * p.transmission_roughness * aspect);
    float cr = (1.0f - p.clearcoat_gloss) * 0.1f + p.clearcoat_gloss * 0.0f;
    cr *= cr;
    d.cc_rough = fmaxf(cr, 0.0001f);
    const float A = fminf(fmaxf(l * (1.0f - p.metallic), 0.3f), 0.7f);
    const float B = 1.0f - A;
    d.w_diffuse = A * (1.0f - p.transmission);
    d.w_trans = A * p.transmission;
    d.w_spec = B * 2.0f / (2.0f + p.clearcoat);
    d.w_cc = B * p.clearcoat / (2.0f + p.clearcoat);
}

struct BsdfSample {
    f3 wi;
    f3 f;
    float pdf;
    bool delta;   // a specular (Dirac) sample: f and pdf are the weights of a delta lobe
};

TR_HD float pow5(float x) { const float t2 = x * x; return t2 * t2 * x; }
TR_HD float one_minus_5(float c) { return pow5(1.0f - c); }

// reflection.hpp:16-40
TR_HD float fresnel_dielectric(float eta_i, float eta_o, float cos_i) {
    if (cos_i < 0.0f) { const float t = eta_i; eta_i = eta_o; eta_o = t; cos_i = -cos_i; }
    const float sin_i = sqrtf(fmaxf(0.0f, 1.0f - cos_i * cos_i));
    const float sin_t = eta_o / eta_i * sin_i;
    if (sin_t >= 1.0f) return 1.0f;
    const float cos_t = sqrtf(fmaxf(0.0f, 1.0f - sin_t * sin_t));
    const float para = (eta_i * cos_i - eta_o * cos_t) / (eta_i * cos_i + eta_o * cos_t);
    const float perp = (eta_o * cos_i - eta_i * cos_t) / (eta_o * cos_i + eta_i * cos_t);
    return 0.5f * (para * para + perp * perp);
}

// 1 - fresnel_dielectric(eta_i, eta_o, cos_i) when the transmitted cosine
// cos_t is known, formed without cancellation: 1 - r^2 = (1 - r)(1 + r) for
// the s and p amplitudes, i.e.
// T = 2 ni no ci ct (1/(ni ci + no ct)^2 + 1/(no ci + ni ct)^2).
TR_HD float fresnel_transmittance(float eta_i, float eta_o, float cos_i, float cos_t) {
    if (cos_i < 0.0f) { const float t = eta_i; eta_i = eta_o; eta_o = t; cos_i = -cos_i; }
    const float a = eta_i * cos_i + eta_o * cos_t, b = eta_o * cos_i + eta_i * cos_t;
    return 2.0f * eta_i * eta_o * cos_i * cos_t * (1.0f / (a * a) + 1.0f / (b * b));
}

// Anisotropic GTR2 (GGX) normal distribution, microfacet.cpp:15-21, for a
// unit half vector h: 1 / (pi ax ay (hx^2/ax^2 + hy^2/ay^2 + hz^2)^2).
TR_HD float D_gtr2(f3 h, float ax, float ay) {
    const float x = h.x / ax, y = h.y / ay;
    const float s = x * x + y * y + h.z * h.z;
    return 1.0f / (kPi * ax * ay * s * s);
}
// Smith G1 for anisotropic GTR2, microfacet.cpp:38-43:
// 2 / (1 + sqrt(1 + (ax^2 x^2 + ay^2 y^2) / z^2)).
TR_HD float G1_gtr2(f3 w, float ax, float ay) {
    const float r2 = w.x * w.x + w.y * w.y;
    if (r2 <= 0.0f) return 1.0f;
    const float a = ax * w.x, b = ay * w.y;
    const float t = (a * a + b * b) / (w.z * w.z);
    return 2.0f / (1.0f + sqrtf(1.0f + t));
}
// GTR1 (clearcoat) distribution, microfacet.cpp:29-35.
TR_HD float D_gtr1(f3 h, float alpha) {
    if (alpha >= 1.0f) return kInvPi;
    const float a2 = alpha * alpha;
    const float sin2 = h.x * h.x + h.y * h.y;
    return (a2 - 1.0f) / (2.0f * kPi * logf(alpha) * (a2 * h.z * h.z + sin2));
}
// Smith G for GTR2 with fixed alpha, microfacet.cpp:46-51.
TR_HD float G1_smith(f3 w, float alpha) {
    const float r2 = w.x * w.x + w.y * w.y;
    if (r2 <= 0.0f) return 1.0f;
    const float t2 = r2 / (w.z * w.z);
    return 2.0f / (1.0f + sqrtf(1.0f + alpha * alpha * t2));
}

// microfacet.cpp:58-80 (Heitz 2014 visible-normal sampling)
TR_HD f3 sample_vndf(f3 ve, float ax, float ay, float u, float v) {
    const f3 vh = normalize(mk(ax * ve.x, ay * ve.y, ve.z));
    const float lensq = vh.x * vh.x + vh.y * vh.y;
    const f3 t1 = lensq > kEps ? mk(-vh.y, vh.x, 0.0f) * (1.0f / sqrtf(lensq)) : mk(1.0f, 0.0f, 0.0f);
    const f3 t2 = cross(vh, t1);
    const float r = sqrtf(u);
    float sp, cp;
    sincosf(2.0f * kPi * v, &sp, &cp);
    const float p1 = r * cp;
    const float p2r = r * sp;
    const float s = 0.5f * (1.0f + vh.z);
    const float p2 = (1.0f - s) * sqrtf(1.0f - p1 * p1) + s * p2r;
    const f3 nh = t1 * p1 + t2 * p2 + vh * sqrtf(fmaxf(0.0f, 1.0f - p1 * p1 - p2 * p2));
    return normalize(mk(ax * nh.x, ay * nh.y, fmaxf(0.0f, nh.z)));
}
// microfacet.cpp:92-109 (D * cos sampling of anisotropic GTR2).  With
// A = cos^2(phi)/ax^2 + sin^2(phi)/ay^2 and D = (1 - A) v + A:
// cos^2(theta) = A (1 - v) / D and sin^2(theta) = v / D, both formed
// directly (sqrt(1 - cos^2) loses half the digits near the pole).
TR_HD f3 sample_gtr2_dcos(float ax, float ay, float u, float v) {
    float s, c;
    sincosf(2.0f * kPi * u, &s, &c);
    float sp = ay * s, cp = ax * c;
    const float nor = 1.0f / sqrtf(sp * sp + cp * cp);
    sp *= nor;
    cp *= nor;
    const float A = (cp / ax) * (cp / ax) + (sp / ay) * (sp / ay);
    const float D = (1.0f - A) * v + A;
    const float ct = sqrtf(A * (1.0f - v) / D);
    const float st = sqrtf(v / D);
    return normalize(mk(st * cp, st * sp, ct));
}
// microfacet.cpp:125-134.  cos^2(theta) = (alpha^(2-2v) - 1) / (alpha^2 - 1);
// both it and sin^2(theta) = alpha^2 (alpha^(-2v) - 1) / (1 - alpha^2) are
// formed with expm1 so the sharp lobe (alpha = 1e-4) keeps its digits.
TR_HD f3 sample_gtr1(float alpha, float u, float v) {
    float s, c;
    sincosf(2.0f * kPi * u, &s, &c);
    const float la = logf(alpha);
    const float a2 = alpha * alpha;
    const float one_m_a2 = 1.0f - a2;
    const float ct2 = -expm1f((2.0f - 2.0f * v) * la) / one_m_a2;
    const float st2 = a2 * expm1f(-2.0f * v * la) / one_m_a2;
    const float ct = sqrtf(fmaxf(0.0f, ct2)), st = sqrtf(fmaxf(0.0f, st2));
    return normalize(mk(st * c, st * s, ct));
}

// ---- Disney lobes (local frame) -------------------------------------------
// disney.cpp:60-109
TR_HD f3 dis_f_transmission(const DevMaterial& m, f3 wi, f3 wo) {
    const float ci = wi.z, co = wo.z;
    const float eta = co > 0.0f ? m.ior : 1.0f / m.ior;
    const f3 v = wo + wi * eta;
    const float lv = length(v);
    f3 h = v * (1.0f / lv);
    // sdem = dot(wo, h) + eta dot(wi, h) = dot(v, h) = +-|v| (no cancellation)
    float sdem = lv;
    if (h.z < 0.0f) { h = -h; sdem = -lv; }
    const float dwo = dot(wo, h), dwi = dot(wi, h);
    // With h along wo + eta wi, Snell's law holds about h for any unit pair,
    // so the transmitted cosine of fresnel_dielectric(ior, 1, dwo) is |dwi|
    // exactly when wo lies on h's side in the orientation the Fresnel term
    // assumes; computing it as sqrt(1 - sin_t^2) loses ~1/cos_t^2 digits
    // near the critical angle.
    const bool snell = (dwo > 0.0f) == (co > 0.0f);
    const float T = snell ? fresnel_transmittance(m.ior, 1.0f, dwo, fabsf(dwi))
                          : 1.0f - fresnel_dielectric(m.ior, 1.0f, dwo);
    const float D = D_gtr2(h, m.tax, m.tay);
    const float G = G1_gtr2(wi, m.tax, m.tay) * G1_gtr2(wo, m.tax, m.tay);
    // eta^2 * (1/eta)^2 = 1 for radiance transport (disney.cpp:98,102-105)
    const float val = T * D * G * dwi * dwo / (ci * co * sdem * sdem);
    const float tf = co > 0.0f ? m.transmission : 1.0f;
    const float k = (1.0f - m.metallic) * tf * fabsf(val);
    return mk(sqrtf(m.C.x), sqrtf(m.C.y), sqrtf(m.C.z)) * k;
}
// disney.cpp:110-145
TR_HD f3 dis_f_inner(const DevMaterial& m, f3 wi, f3 wo) {
    const f3 h = -normalize(wi + wo);
    const float F = fresnel_dielectric(m.ior, 1.0f, dot(wo, h));
    const float D = D_gtr2(h, m.tax, m.tay);
    const float G = G1_gtr2(wi, m.tax, m.tay) * G1_gtr2(wo, m.tax, m.tay);
    return m.C * (m.transmission * fabsf(F * D * G / (4.0f * wi.z * wo.z)));
}
// disney.cpp:146-209 (outside reflection: diffuse + retro, sheen, specular, clearcoat)
TR_HD f3 dis_f_reflect(const DevMaterial& m, f3 wi, f3 wo) {
    const float ci = wi.z, co = wo.z;
    const f3 h = normalize(wi + wo);
    const float cd = dot(wi, h);
    const float Fcd = one_minus_5(cd);
    f3 diffuse = f3s(0.0f), sheen = f3s(0.0f);
    if (m.metallic < 1.0f) {
        const float FL = one_minus_5(ci), FV = one_minus_5(co);
        const float RR = 2.0f * m.roughness * cd * cd;
        const float k = kInvPi * ((1.0f - 0.5f * FL) * (1.0f - 0.5f * FV) + RR * (FL + FV + FL * FV * (RR - 1.0f)));
        diffuse = m.C * k;
        if (m.sheen > 0.0f) sheen = lerp3(f3s(1.0f), m.Ctint, m.sheen_tint) * (4.0f * m.sheen * Fcd);
    }
    const f3 Cspec = lerp3(lerp3(f3s(1.0f), m.Ctint, m.spec_tint), m.C, m.metallic);
    const float Fd = fresnel_dielectric(m.ior, 1.0f, cd);
    const f3 cond = Cspec + (f3s(1.0f) - Cspec) * Fcd;
    const f3 F = lerp3(m.spec_scale * Cspec * Fd, cond, m.metallic);
    const float D = D_gtr2(h, m.ax, m.ay);
    const float G = G1_gtr2(wi, m.ax, m.ay) * G1_gtr2(wo, m.ax, m.ay);
    f3 r = (diffuse + sheen) * ((1.0f - m.metallic) * (1.0f - m.transmission)) + F * (D * G / fabsf(4.0f * ci * co));
    if (m.clearcoat > 0.0f) {
        const float Dc = D_gtr1(h, m.cc_rough);
        const float Fc = 0.04f + 0.96f * Fcd;
        const float Gc = G1_smith(wi, 0.25f) * G1_smith(wo, 0.25f);
        r += f3s(m.clearcoat * Dc * Fc * Gc / fabsf(4.0f * ci * co));
    }
    return r;
}
// disney.cpp:210-235, with cos(theta_h) as the D * cos sampling density
TR_HD float dis_pdf_transmission(const DevMaterial& m, f3 wi, f3 wo) {
    const float eta = wo.z > 0.0f ? m.ior : 1.0f / m.ior;
    const f3 v = wo + wi * eta;
    const float lv = length(v);
    f3 h = v * (1.0f / lv);
    float sdem = lv;
    if (h.z < 0.0f) { h = -h; sdem = -lv; }
    const float dwo = dot(wo, h), dwi = dot(wi, h);
    if (((wo.z > 0.0f) != (dwo > 0.0f)) || ((wi.z > 0.0f) != (dwi > 0.0f))) return 0.0f;
    const float jac = eta * eta * dwi / (sdem * sdem);
    return fabsf(h.z * D_gtr2(h, m.tax, m.tay) * jac);
}
// disney.cpp:236-251
TR_HD float dis_pdf_inner(const DevMaterial& m, f3 wi, f3 wo) {
    const f3 h = -normalize(wi + wo);
    return fabsf(h.z * D_gtr2(h, m.tax, m.tay) / (4.0f * dot(wi, h)));
}
// disney.cpp:499-534 (local, |z| >= eps already checked)
TR_HD float dis_pdf_local(const DevMaterial& m, f3 wi, f3 wo) {
    if (wo.z < 0.0f) {
        if (m.transmission == 0.0f) return 0.0f;
        const float F = fminf(fmaxf(fresnel_dielectric(m.ior, 1.0f, wo.z), 0.1f), 0.9f);
        if (wi.z > 0.0f) return (1.0f - F) * dis_pdf_transmission(m, wi, wo);
        return F * dis_pdf_inner(m, wi, wo);
    }
    if (wi.z < 0.0f) return m.w_trans * dis_pdf_transmission(m, wi, wo);
    const f3 h = normalize(wi + wo);
    const float pd = wi.z * kInvPi;
    const float ps = G1_gtr2(wo, m.ax, m.ay) * D_gtr2(h, m.ax, m.ay) / (4.0f * wo.z);
    const float pc = h.z * D_gtr1(h, m.cc_rough) / (4.0f * dot(wi, h));
    return m.w_diffuse * pd + m.w_spec * ps + m.w_cc * pc;
}
TR_HD f3 dis_f_local(const DevMaterial& m, f3 wi, f3 wo) {
    if (wi.z * wo.z < 0.0f) {
        if (m.transmission == 0.0f) return f3s(0.0f);
        return dis_f_transmission(m, wi, wo);
    }
    if (wi.z < 0.0f && wo.z < 0.0f) {
        if (m.transmission == 0.0f) return f3s(0.0f);
        return dis_f_inner(m, wi, wo);
    }
    return dis_f_reflect(m, wi, wo);
}
// disney.cpp:285-305 (refraction about the sampled normal)
TR_HD bool dis_sample_transmission(const DevMaterial& m, f3 wo, float u, float v, f3& wi) {
    const f3 h = sample_gtr2_dcos(m.tax, m.tay, u, v);
    if (h.z <= 0.0f) return false;
    const float dho = dot(h, wo);
    if ((wo.z > 0.0f) != (dho > 0.0f)) return false;
    const float eta = wo.z > 0.0f ? 1.0f / m.ior : m.ior;
    const f3 n = dho > 0.0f ? h : -h;
    const float ci = fabsf(dho);
    const float st2 = eta * eta * fmaxf(0.0f, 1.0f - ci * ci);
    if (st2 >= 1.0f) return false;
    const float ct = sqrtf(1.0f - st2);
    wi = normalize(n * (eta * ci - ct) - wo * eta);
    if (wi.z * wo.z > 0.0f || ((wi.z > 0.0f) != (dot(h, wi) > 0.0f))) return false;
    return true;
}
TR_HD bool dis_sample_local(const DevMaterial& m, f3 wo, float su, float u, float v, f3& wi) {
    if (wo.z < 0.0f) {
        if (m.transmission == 0.0f) return false;
        const float F = fminf(fmaxf(fresnel_dielectric(m.ior, 1.0f, wo.z), 0.1f), 0.9f);
        if (su < F) {
            const f3 h = sample_gtr2_dcos(m.tax, m.tay, u, v);
            if (h.z <= 0.0f) return false;
            const f3 w = h * (2.0f * dot(wo, h)) - wo;
            if (w.z > 0.0f) return false;
            wi = normalize(w);
            return true;
        }
        return dis_sample_transmission(m, wo, u, v, wi);
    }
    float sel = su;
    if (sel < m.w_diffuse) { wi = cosine_hemisphere(u, v); return true; }
    sel -= m.w_diffuse;
    if (sel < m.w_trans) return dis_sample_transmission(m, wo, u, v, wi);
    sel -= m.w_trans;
    f3 h;
    if (sel < m.w_spec) h = normalize(sample_vndf(wo, m.ax, m.ay, u, v));
    else h = sample_gtr1(m.cc_rough, u, v);
    if (h.z <= 0.0f) return false;
    wi = normalize(h * (2.0f * dot(wo, h)) - wo);
    return wi.z > 0.0f;
}

// ---- full BSDF: fringe handling, local normalisation, normal correction ----
TR_HD bool fringe(f3 w, f3 ng) { return (dot(w, ng) > 0.0f) != (w.z > 0.0f); }
// coordinate.hpp:43-47 with the shading normal = +z
TR_HD float normal_correct(f3 ng, f3 wi) {
    const float lw = length(wi);
    const float dem = fabsf(dot(ng, wi) / (length(ng) * lw));
    return dem < kEps ? 1.0f : fabsf(wi.z / lw) / dem;
}

struct BsdfEval {
    f3 f;
    float pdf;
};

// PhongSpecularBXDF (phong_specular_bxdf.cpp:7-40): (ns + 1) / (2 pi) cos^ns(a)
// with a the angle between wi and the mirror direction of wo about the
// shading normal, reflect(wo, z) = (-wo.x, -wo.y, wo.z); zero when
// cos(a) <= 0.  Both the value (times ks) and the pdf of the lobe.
TR_HD float phong_lobe(f3 lwi, f3 lwo, float ns) {
    const float c = lwi.z * lwo.z - lwi.x * lwo.x - lwi.y * lwo.y;
    return c > 0.0f ? (ns + 1.0f) * (0.5f * kInvPi) * powf(c, ns) : 0.0f;
}

// eval + pdf for directions wi, wo (shading frame), geometric normal ng.
// Glass is a pair of delta lobes: zero for every pair of directions.
TR_HD BsdfEval bsdf_eval(const DevMaterial& m, f3 wi, f3 wo, f3 ng) {
    BsdfEval r{f3s(0.0f), 0.0f};
    if (m.type == MAT_GLASS) return r;
    if (fringe(wi, ng) || fringe(wo, ng)) {
        const bool gi = dot(wi, ng) > 0.0f, go = dot(wo, ng) > 0.0f;
        if (gi && go) r.f = m.C * kInvPi;
        if (gi == go) r.pdf = dot(wi, ng) / (length(wi) * length(ng)) * kInvPi;
        return r;
    }
    const f3 lwi = normalize(wi), lwo = normalize(wo);
    if (m.type == MAT_LAMBERT) {
        if (lwi.z <= 0.0f || lwo.z <= 0.0f) return r;
        r.f = m.C * (kInvPi * normal_correct(ng, wi));
        r.pdf = lwi.z * kInvPi;
        return r;
    }
    if (m.type == MAT_PHONG) {
        // both lobes vanish unless wi and wo lie above the shading surface
        if (lwi.z <= 0.0f || lwo.z <= 0.0f) return r;
        const float ps = phong_lobe(lwi, lwo, m.ns);
        r.f = (m.Ctint * kInvPi + m.spec_scale * ps) * normal_correct(ng, wi);
        r.pdf = m.w_diffuse * lwi.z * kInvPi + m.w_spec * ps;
        return r;
    }
    if (fabsf(lwi.z) < kEps || fabsf(lwo.z) < kEps) return r;
    r.f = dis_f_local(m, lwi, lwo) * normal_correct(ng, wi);
    r.pdf = dis_pdf_local(m, lwi, lwo);
    return r;
}

// Sample wi given wo; returns pdf = 0 when no valid direction was produced.
// (gt, gb): a tangent frame around ng for the fringe fallback.
// GlassBSDF::sample (src/material/glass.cpp:24-67): reflection with
// probability F = dielectric Fresnel at lwo.z (ior inside, 1 outside), else
// refraction about the shading normal on wo's side; f = colour * weight /
// |lwi.z| times the shading-normal correction, radiance transport scales the
// refracted weight by (eta_i / eta_t)^2.  Total internal reflection has F = 1.
TR_HD BsdfSample glass_sample(const DevMaterial& m, f3 wo, f3 ng, float su) {
    BsdfSample s{f3s(0.0f), f3s(0.0f), 0.0f, true};
    const f3 lwo = normalize(wo);
    const float F = fresnel_dielectric(m.ior, 1.0f, lwo.z);
    f3 wi;
    f3 col;
    float w;
    if (su < F) {
        wi = mk(-lwo.x, -lwo.y, lwo.z);
        col = m.C;
        w = F;
    } else {
        const bool outside = lwo.z > 0.0f;
        const float eta = outside ? 1.0f / m.ior : m.ior;   // eta_i / eta_t
        const float ci = fabsf(lwo.z);
        const float st2 = eta * eta * fmaxf(0.0f, 1.0f - ci * ci);
        if (st2 >= 1.0f) return s;                          // refract_dir, reflection.hpp:43-51
        const float ct = sqrtf(1.0f - st2);
        const float nz = outside ? 1.0f : -1.0f;
        wi = normalize(mk(-eta * lwo.x, -eta * lwo.y, (eta * ci - ct) * nz - eta * lwo.z));
        col = m.spec_scale * (eta * eta);
        w = fresnel_transmittance(m.ior, 1.0f, lwo.z, ct);   // 1 - F without cancellation
    }
    s.wi = wi;
    s.f = col * (w / fabsf(wi.z) * normal_correct(ng, wi));
    s.pdf = w;
    return s;
}

TR_HD BsdfSample bsdf_sample(const DevMaterial& m, f3 wo, f3 ng, float su, float u, float v) {
    if (m.type == MAT_GLASS) return glass_sample(m, wo, ng, su);
    BsdfSample s{f3s(0.0f), f3s(0.0f), 0.0f, false};
    if (fringe(wo, ng)) {
        // bsdf.hpp:141-153: cosine sampling about the geometric normal
        if (!(dot(wo, ng) > 0.0f)) return s;
        const f3 l = cosine_hemisphere(u, v);
        const float pd = l.z * kInvPi;
        if (pd < kEps) return s;
        const f3 n = normalize(ng);
        f3 t, b;
        onb(n, t, b);
        const f3 wi = normalize(t * l.x + b * l.y + n * l.z);
        s.wi = wi;
        s.f = m.C * (kInvPi * normal_correct(ng, wi));
        s.pdf = pd;
        return s;
    }
    const f3 lwo = normalize(wo);
    f3 wi;
    if (m.type == MAT_LAMBERT) {
        if (lwo.z <= 0.0f) return s;
        wi = cosine_hemisphere(u, v);
        const float pd = wi.z * kInvPi;
        if (pd < 0.001f) return s;   // diffuse_bxdf.cpp:25-26
        wi = normalize(wi);
        if (fringe(wi, ng)) {
            // both hemispheres disagree for wi: evaluate through the fringe path
            const BsdfEval e = bsdf_eval(m, wi, wo, ng);
            s.wi = wi; s.f = e.f; s.pdf = e.pdf;
            return s;
        }
        s.wi = wi;
        s.f = m.C * (kInvPi * normal_correct(ng, wi));
        s.pdf = pd;
        return s;
    }
    if (m.type == MAT_PHONG) {
        if (lwo.z <= 0.0f) return s;
        if (su < m.w_diffuse) {
            // DiffuseBXDF::sample (diffuse_bxdf.cpp:20-31)
            wi = cosine_hemisphere(u, v);
            if (wi.z * kInvPi < 0.001f) return s;
            wi = normalize(wi);
        } else {
            // PowCosineSampleHemisphere about the mirror direction
            // (phong_specular_bxdf.cpp:7-17, :55-69) in the frame that
            // Coord::create_from_z builds (coordinate.hpp:84-93), so a
            // uniform triple maps to the reference's direction; below the
            // surface: no sample
            const f3 z = normalize(mk(-lwo.x, -lwo.y, lwo.z));
            const f3 y = fabsf(z.x) > fabsf(z.y) ? normalize(mk(-z.z, 0.0f, z.x)) : normalize(mk(0.0f, z.z, -z.y));
            const f3 x = cross(y, z);
            const float ct = powf(u, 1.0f / (m.ns + 1.0f));
            const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
            float sp, cp;
            sincosf(2.0f * kPi * v, &sp, &cp);
            wi = normalize(x * (st * cp) + y * (st * sp) + z * ct);
            if (wi.z <= 0.0f) return s;
        }
        // the aggregate's value and pdf at wi: both lobes, weighted
        // (aggregate_bsdf.hpp:84-98); a wi below the geometric surface takes
        // the black-fringe path of eval
        const BsdfEval e = bsdf_eval(m, wi, wo, ng);
        s.wi = wi;
        s.f = e.f;
        s.pdf = e.pdf;
        return s;
    }
    if (fabsf(lwo.z) < kEps) return s;
    if (!dis_sample_local(m, lwo, su, u, v, wi)) return s;
    const BsdfEval e = bsdf_eval(m, wi, wo, ng);
    s.wi = wi;
    s.f = e.f;
    s.pdf = e.pdf;
    return s;
}

}  // namespace trhip
