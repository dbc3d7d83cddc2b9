// refcpu — TEST INFRASTRUCTURE ONLY (see ref_math.hpp header).
//
// BSDFs on the path, templated on the arithmetic type T so one restatement
// yields both the reference-precision (float) BSDF and its FP64 twin.
//
//  * DisneyRef<T>: src/material/disney.cpp:14-548 (the "2015 bsdf (old)"
//    DisneyBSDF) with its helpers src/material/utility/microfacet.cpp:7-134,
//    src/material/utility/microfacet.hpp:10-14 (one_minus_5),
//    src/utility/reflection.hpp:16-51 and the black-fringe handling of
//    LocalBSDF, src/core/bsdf.hpp:114-165.
//    `fixed` switches off three reference defects that make the sampling
//    density differ from the returned pdf:
//      Q2  sample(): no `else` after the diffuse branch, and the lobe choice
//          for lwo.z < 0 ignores the macro-Fresnel split used by pdf()
//          (disney.cpp:465-488 vs :510-519).  Fixed: exclusive lobe
//          selection; for lwo.z < 0 inner reflection with probability F.
//      Q3  pdf_transmission() multiplies by dot(lwi,lwh) where the D*cos
//          sampling density needs cos(theta_h) (disney.cpp:234).
//      Q21 refract_dir() takes cos(theta_i) = |lwo.z| instead of
//          dot(lwo, wh) (reflection.hpp:44), so sample_transmission() does
//          not refract about the sampled microfacet normal.
//  * LambertRef<T>: Phong with ks = 0 (src/material/phong.cpp:29-70) is an
//    AggregateBSDF (src/material/aggregate_bsdf.hpp:36-118) holding one
//    DiffuseBXDF (src/material/bxdf/diffuse_bxdf.cpp:14-38) of weight 1.
//  * PhongRef<T>: Phong with any kd, ks, ns (src/material/phong.cpp:29-70):
//    the same AggregateBSDF with a DiffuseBXDF and a PhongSpecularBXDF
//    (src/material/bxdf/phong_specular_bxdf.cpp:7-79), weighted by luminance.
//      Q24 sample() picks the lobe with (u - 0.001) * sum(w) <= w_d
//          (aggregate_bsdf.hpp:62-76), so the diffuse lobe is chosen with
//          probability w_d + 0.001 while the returned pdf weighs it by w_d.
//          Fixed: u * sum(w) <= w_d.  Q22 applies as for LambertRef.
//  * GlassRef<T>: Glass / GlassBSDF (src/material/glass.cpp:7-146), a pair
//    of delta lobes: reflection with probability F, refraction otherwise.
//  * HGRef<T>: Henyey-Greenstein phase function,
//    src/medium/phase_function.hpp:12-77.
#pragma once
#include "ref_math.hpp"

namespace refcpu {

template <class T>
struct BSDFSample {
    V3<T> wi;
    RGB<T> f;
    T pdf = 0;
    bool delta = false;   // BSDFSampleResult::is_delta (bsdf.hpp:17-25)
    bool valid() const { return pdf > 0 && !f.is_black(); }  // bsdf.hpp:90-92
};

// Fields of create_disney (src/factory/material.hpp:37-52), already
// evaluated from constant textures.
struct DisneyParams {
    float base_color[3];
    float metallic, roughness, transmission, transmission_roughness, ior;
    float specular_scale[3];
    float specular_tint, anisotropic, sheen, sheen_tint, clearcoat, clearcoat_gloss;
};

namespace mf {
template <class T> inline T sqr(T x) { return x * x; }
template <class T> inline T one_minus_5(T x) { T t = 1 - x; T t2 = t * t; return t2 * t2 * t; }
template <class T> inline T compute_alpha(T sphi, T cphi, T ax, T ay) { return sqr(cphi / ax) + sqr(sphi / ay); }
template <class T> inline T anisotropic_gtr2(T sphi, T cphi, T sth, T cth, T ax, T ay) {
    const T A = compute_alpha(sphi, cphi, ax, ay);
    const T RD = sqr(sth) * A + sqr(cth);
    return 1 / (kPi<T> * ax * ay * sqr(RD));
}
template <class T> inline T gtr1(T sth, T cth, T alpha) {
    if (alpha >= 1) return kInvPi<T>;
    const T U = sqr(alpha) - 1;
    const T LD = 2 * kPi<T> * std::log(alpha);
    const T RD = sqr(alpha * cth) + sqr(sth);
    return U / (LD * RD);
}
template <class T> inline T smith_anisotropic_gtr2(T cphi, T sphi, T ax, T ay, T tth) {
    const T t = sqr(ax * cphi) + sqr(ay * sphi);
    const T sv = 1 + t * sqr(tth);
    const T lambda = -T(0.5) + T(0.5) * std::sqrt(sv);
    return 1 / (1 + lambda);
}
template <class T> inline T smith_gtr2(T tth, T alpha) {
    if (!tth) return 1;
    const T root = alpha * tth;
    return 2 / (1 + std::sqrt(1 + root * root));
}
// microfacet.cpp:58-80 (Heitz 2014 visible-normal sampling)
template <class T> inline V3<T> sample_vndf(const V3<T>& ve, T ax, T ay, T u, T v) {
    const V3<T> vh = V3<T>(ax * ve.x, ay * ve.y, ve.z).normalize();
    const T lensq = vh.x * vh.x + vh.y * vh.y;
    const V3<T> t1 = lensq > T(kEps) ? V3<T>(-vh.y, vh.x, 0).div(std::sqrt(lensq)) : V3<T>(1, 0, 0);
    const V3<T> t2 = cross(vh, t1);
    const T r = std::sqrt(u);
    const T phi = 2 * kPi<T> * v;
    const T t_1 = r * std::cos(phi);
    const T _t_2 = r * std::sin(phi);
    const T s = T(0.5) * (1 + vh.z);
    const T t_2 = (1 - s) * std::sqrt(1 - t_1 * t_1) + s * _t_2;
    const V3<T> nh = t1 * t_1 + t2 * t_2 + vh * std::sqrt(std::max<T>(0, 1 - t_1 * t_1 - t_2 * t_2));
    return V3<T>(ax * nh.x, ay * nh.y, std::max<T>(0, nh.z)).normalize();
}
// microfacet.cpp:92-109 (D*cos sampling)
template <class T> inline V3<T> sample_anisotropic_gtr2(T ax, T ay, T u, T v) {
    T sphi = ay * std::sin(2 * kPi<T> * u);
    T cphi = ax * std::cos(2 * kPi<T> * u);
    const T nor = 1 / std::sqrt(sqr(sphi) + sqr(cphi));
    sphi *= nor;
    cphi *= nor;
    const T A = compute_alpha(sphi, cphi, ax, ay);
    const T cth = std::sqrt(A * (1 - v) / ((1 - A) * v + A));
    const T sth = std::sqrt(std::max<T>(0, 1 - sqr(cth)));
    return V3<T>(sth * cphi, sth * sphi, cth).normalize();
}
// microfacet.cpp:125-134
template <class T> inline V3<T> sample_gtr1(T alpha, T u, T v) {
    const T phi = 2 * kPi<T> * u;
    const T cth = std::sqrt((std::pow(alpha, 2 - 2 * v) - 1) / (sqr(alpha) - 1));
    const T sth = local_cos_to_sin(cth);
    return V3<T>(sth * std::cos(phi), sth * std::sin(phi), cth).normalize();
}
}  // namespace mf

// reflection.hpp:16-40
template <class T> inline T dielectric_fresnel(T eta_i, T eta_o, T cos_i) {
    if (cos_i < 0) { std::swap(eta_i, eta_o); cos_i = -cos_i; }
    const T sin_i = std::sqrt(std::max<T>(0, 1 - cos_i * cos_i));
    const T sin_t = eta_o / eta_i * sin_i;
    if (sin_t >= 1) return 1;
    const T cos_t = std::sqrt(std::max<T>(0, 1 - sin_t * sin_t));
    const T para = (eta_i * cos_i - eta_o * cos_t) / (eta_i * cos_i + eta_o * cos_t);
    const T perp = (eta_o * cos_i - eta_i * cos_t) / (eta_o * cos_i + eta_i * cos_t);
    return T(0.5) * (para * para + perp * perp);
}
// reflection.hpp:43-51 (Q21: reference takes cos_i = |lwo.z|)
template <class T> inline bool refract_dir(const V3<T>& lwo, const V3<T>& n, T eta, bool fixed, V3<T>& out) {
    const T cos_i = fixed ? dot(lwo, n) : std::abs(lwo.z);
    const T sin_i2 = std::max<T>(0, 1 - cos_i * cos_i);
    const T sin_t2 = eta * eta * sin_i2;
    if (sin_t2 >= 1) return false;
    const T cos_t = std::sqrt(1 - sin_t2);
    out = n * (eta * cos_i - cos_t) - lwo * eta;
    return true;
}

template <class T>
class DisneyRef {
public:
    Coord<T> geo, shd;
    bool fixed = false;
    RGB<T> C, Ctint, specular_scale;
    T metallic, roughness, specular_tint, anisotropic, sheen, sheen_tint, clearcoat;
    T transmission, IOR, transmission_roughness, trans_ax, trans_ay, ax, ay, cc_rough;
    T w_diffuse, w_specular, w_clearcoat, w_transmission;

    // disney.cpp:347-399
    DisneyRef(const Coord<T>& g, const Coord<T>& s, const DisneyParams& p, bool fixed_mode) : geo(g), shd(s), fixed(fixed_mode) {
        C = RGB<T>(p.base_color[0], p.base_color[1], p.base_color[2]);
        const T lum = C.lum();
        Ctint = lum > 0 ? C / lum : RGB<T>(1);
        metallic = p.metallic;
        roughness = p.roughness;
        specular_scale = RGB<T>(p.specular_scale[0], p.specular_scale[1], p.specular_scale[2]);
        specular_tint = p.specular_tint;
        anisotropic = p.anisotropic;
        sheen = p.sheen;
        sheen_tint = p.sheen_tint;
        transmission = p.transmission;
        transmission_roughness = p.transmission_roughness;
        IOR = std::max<T>(T(1.01f), T(p.ior));
        const T aspect = anisotropic > 0 ? std::sqrt(1 - T(0.9f) * anisotropic) : T(1);
        ax = std::max<T>(T(0.001f), mf::sqr(roughness) / aspect);
        ay = std::max<T>(T(0.001f), mf::sqr(roughness) * aspect);
        trans_ax = std::max<T>(T(0.001f), mf::sqr(transmission_roughness) / aspect);
        trans_ay = std::max<T>(T(0.001f), mf::sqr(transmission_roughness) * aspect);
        clearcoat = p.clearcoat;
        cc_rough = mix(T(0.1f), T(0), T(p.clearcoat_gloss));
        cc_rough *= cc_rough;
        cc_rough = std::max<T>(cc_rough, T(0.0001f));
        const T A = std::clamp<T>(C.lum() * (1 - metallic), T(0.3f), T(0.7f));
        const T B = 1 - A;
        w_diffuse = A * (1 - transmission);
        w_transmission = A * transmission;
        w_specular = B * 2 / (2 + clearcoat);
        w_clearcoat = B * clearcoat / (2 + clearcoat);
    }

    // bsdf.hpp:124-160
    bool fringe(const V3<T>& w) const { return geo.positive_z(w) ^ shd.positive_z(w); }
    RGB<T> eval_fringe(const V3<T>& wi, const V3<T>& wo) const {
        if (geo.positive_z(wi) & geo.positive_z(wo)) return C * kInvPi<T>;
        return RGB<T>(0);
    }
    T pdf_fringe(const V3<T>& wi, const V3<T>& wo) const {
        if (geo.positive_z(wi) ^ geo.positive_z(wo)) return 0;
        return geo.to_local(wi).normalize().z * kInvPi<T>;
    }

    void angles(const V3<T>& w, T& sphi, T& cphi) const {
        T phi = local_phi(w);
        sphi = std::sin(phi);
        cphi = std::cos(phi);
    }
    T D_aniso(const V3<T>& h, T axx, T ayy) const {
        T sp, cp;
        angles(h, sp, cp);
        return mf::anisotropic_gtr2(sp, cp, local_cos_to_sin(h.z), h.z, axx, ayy);
    }
    T G_aniso(const V3<T>& wi, const V3<T>& wo, T axx, T ayy) const {
        T spi, cpi, spo, cpo;
        angles(wi, spi, cpi);
        angles(wo, spo, cpo);
        return mf::smith_anisotropic_gtr2(cpi, spi, axx, ayy, local_tan_theta(wi)) *
               mf::smith_anisotropic_gtr2(cpo, spo, axx, ayy, local_tan_theta(wo));
    }

    // disney.cpp:60-109
    RGB<T> f_transmission(const V3<T>& lwi, const V3<T>& lwo) const {
        const T ci = lwi.z, co = lwo.z;
        const T eta = co > 0 ? IOR : 1 / IOR;
        V3<T> h = (lwo + lwi * eta).normalize();
        if (h.z < 0) h = -h;
        const T cd = dot(lwo, h);
        const T F = dielectric_fresnel(IOR, T(1), cd);
        const T D = D_aniso(h, trans_ax, trans_ay);
        const T G = G_aniso(lwi, lwo, trans_ax, trans_ay);
        const T sdem = cd + eta * dot(lwi, h);
        const T corr = 1 / eta;  // TransportMode::Radiance
        const RGB<T> sqrtC = sqrt(C);
        const T val = (1 - F) * D * G * eta * eta * dot(lwi, h) * dot(lwo, h) * corr * corr / (ci * co * sdem * sdem);
        const T trans_factor = co > 0 ? transmission : T(1);
        return sqrtC * ((1 - metallic) * trans_factor * std::abs(val));
    }
    // disney.cpp:110-145
    RGB<T> f_inner_reflection(const V3<T>& lwi, const V3<T>& lwo) const {
        const V3<T> h = -(lwi + lwo).normalize();
        const T cd = dot(lwo, h);
        const T F = dielectric_fresnel(IOR, T(1), cd);
        const T D = D_aniso(h, trans_ax, trans_ay);
        const T G = G_aniso(lwi, lwo, trans_ax, trans_ay);
        return C * (transmission * std::abs(F * D * G / (4 * lwi.z * lwo.z)));
    }
    // disney.cpp:146-155
    RGB<T> f_diffuse(T ci, T co, T cd) const {
        const RGB<T> lam = C / kPi<T>;
        const T FL = mf::one_minus_5(ci), FV = mf::one_minus_5(co);
        const T RR = 2 * roughness * cd * cd;
        const RGB<T> retro = C / kPi<T> * (RR * (FL + FV + FL * FV * (RR - 1)));
        return lam * ((1 - T(0.5) * FL) * (1 - T(0.5) * FV)) + retro;
    }
    // disney.cpp:156-159
    RGB<T> f_sheen(T cd) const { return mixc(RGB<T>(1), Ctint, sheen_tint) * (4 * sheen * mf::one_minus_5(cd)); }
    // disney.cpp:160-197
    RGB<T> f_specular(const V3<T>& lwi, const V3<T>& lwo) const {
        const T ci = lwi.z, co = lwo.z;
        const V3<T> h = (lwi + lwo).normalize();
        const T cd = dot(lwi, h);
        const RGB<T> Cspec = mixc(mixc(RGB<T>(1), Ctint, specular_tint), C, metallic);
        const RGB<T> dfres = Cspec * dielectric_fresnel(IOR, T(1), cd);
        const RGB<T> cfres = Cspec + (RGB<T>(1) - Cspec) * mf::one_minus_5(cd);
        const RGB<T> F = mixc(specular_scale * dfres, cfres, metallic);
        const T D = D_aniso(h, ax, ay);
        const T G = G_aniso(lwi, lwo, ax, ay);
        return F * (D * G / std::abs(4 * ci * co));
    }
    // disney.cpp:198-209
    RGB<T> f_clearcoat(T ci, T co, T ti, T to, T sh, T ch, T cd) const {
        const T D = mf::gtr1(sh, ch, cc_rough);
        const T F = T(0.04f) + (1 - T(0.04f)) * mf::one_minus_5(cd);
        const T G = mf::smith_gtr2(ti, T(0.25f)) * mf::smith_gtr2(to, T(0.25f));
        return RGB<T>(clearcoat * D * F * G / std::abs(4 * ci * co));
    }
    // disney.cpp:210-235 (Q3 in reference mode)
    T pdf_transmission(const V3<T>& lwi, const V3<T>& lwo) const {
        const T eta = lwo.z > 0 ? IOR : 1 / IOR;
        V3<T> h = (lwo + lwi * eta).normalize();
        if (h.z < 0) h = -h;
        if (((lwo.z > 0) != (dot(h, lwo) > 0)) || ((lwi.z > 0) != (dot(h, lwi) > 0))) return 0;
        const T sdem = dot(lwo, h) + eta * dot(lwi, h);
        const T dwh_dwi = eta * eta * dot(lwi, h) / (sdem * sdem);
        const T D = D_aniso(h, trans_ax, trans_ay);
        const T lead = fixed ? h.z : dot(lwi, h);
        return std::abs(lead * D * dwh_dwi);
    }
    // disney.cpp:236-251
    T pdf_inner_reflection(const V3<T>& lwi, const V3<T>& lwo) const {
        const V3<T> h = -(lwi + lwo).normalize();
        const T cd = dot(lwi, h);
        const T D = D_aniso(h, trans_ax, trans_ay);
        return std::abs(h.z * D / (4 * cd));
    }
    // disney.cpp:256-284
    void pdf_spec_cc(const V3<T>& lwi, const V3<T>& lwo, T& ps, T& pc) const {
        const V3<T> h = (lwi + lwo).normalize();
        const T ch = h.z, sh = local_cos_to_sin(ch), cd = dot(lwi, h);
        const T cpo = std::cos(local_phi(lwo));
        const T spo = local_cos_to_sin(cpo);
        const T to = local_tan_theta(lwo);
        const T Ds = D_aniso(h, ax, ay);
        ps = mf::smith_anisotropic_gtr2(cpo, spo, ax, ay, to) * Ds / (4 * lwo.z);
        const T Dc = mf::gtr1(sh, ch, cc_rough);
        pc = ch * Dc / (4 * cd);
    }
    // disney.cpp:285-305
    bool sample_transmission(const V3<T>& lwo, T u, T v, V3<T>& lwi) const {
        const V3<T> h = mf::sample_anisotropic_gtr2(trans_ax, trans_ay, u, v);
        if (h.z <= 0) return false;
        if ((lwo.z > 0) != (dot(h, lwo) > 0)) return false;
        const T eta = lwo.z > 0 ? 1 / IOR : IOR;
        const V3<T> owh = dot(h, lwo) > 0 ? h : -h;
        V3<T> r;
        if (!refract_dir(lwo, owh, eta, fixed, r)) return false;
        lwi = r.normalize();
        if (lwi.z * lwo.z > 0 || ((lwi.z > 0) != (dot(h, lwi) > 0))) return false;
        return true;
    }
    // disney.cpp:306-318
    bool sample_inner_reflection(const V3<T>& lwo, T u, T v, V3<T>& lwi) const {
        const V3<T> h = mf::sample_anisotropic_gtr2(trans_ax, trans_ay, u, v);
        if (h.z <= 0) return false;
        const V3<T> w = h * (2 * dot(lwo, h)) - lwo;
        if (w.z > 0) return false;
        lwi = w.normalize();
        return true;
    }
    // disney.cpp:322-344
    bool sample_specular(const V3<T>& lwo, T u, T v, V3<T>& lwi) const {
        const V3<T> h = mf::sample_vndf(lwo, ax, ay, u, v).normalize();
        if (h.z <= 0) return false;
        lwi = (h * (2 * dot(lwo, h)) - lwo).normalize();
        return lwi.z > 0;
    }
    bool sample_clearcoat(const V3<T>& lwo, T u, T v, V3<T>& lwi) const {
        const V3<T> h = mf::sample_gtr1(cc_rough, u, v);
        if (h.z <= 0) return false;
        lwi = (h * (2 * dot(lwo, h)) - lwo).normalize();
        return lwi.z > 0;
    }

    // disney.cpp:401-455
    RGB<T> eval(const V3<T>& wi, const V3<T>& wo) const {
        if (fringe(wi) | fringe(wo)) return eval_fringe(wi, wo);
        const V3<T> lwi = shd.to_local(wi).normalize();
        const V3<T> lwo = shd.to_local(wo).normalize();
        if (std::abs(lwi.z) < T(kEps) || std::abs(lwo.z) < T(kEps)) return RGB<T>(0);
        const T ncf = normal_correct_factor(geo.z, shd.z, wi);
        if (lwi.z * lwo.z < 0) {
            if (transmission == 0) return RGB<T>(0);
            return f_transmission(lwi, lwo) * ncf;
        } else if (lwi.z < 0 && lwo.z < 0) {
            if (transmission == 0) return RGB<T>(0);
            return f_inner_reflection(lwi, lwo) * ncf;
        }
        const T ci = lwi.z, co = lwo.z;
        const V3<T> h = (lwi + lwo).normalize();
        const T cd = dot(lwi, h);
        RGB<T> diffuse, sh;
        if (metallic < 1) {
            diffuse = f_diffuse(ci, co, cd);
            if (sheen > 0) sh = f_sheen(cd);
        }
        RGB<T> spec = f_specular(lwi, lwo);
        RGB<T> cc;
        if (clearcoat > 0) {
            const T ti = local_tan_theta(lwi), to = local_tan_theta(lwo);
            const T ch = h.z, shh = local_cos_to_sin(ch);
            cc = f_clearcoat(ci, co, ti, to, shh, ch, cd);
        }
        RGB<T> ret = (diffuse + sh) * ((1 - metallic) * (1 - transmission)) + spec + cc;
        return ret * ncf;
    }

    // disney.cpp:499-534
    T pdf(const V3<T>& wi, const V3<T>& wo) const {
        if (fringe(wi) | fringe(wo)) return pdf_fringe(wi, wo);
        const V3<T> lwi = shd.to_local(wi).normalize();
        const V3<T> lwo = shd.to_local(wo).normalize();
        if (std::abs(lwi.z) < T(kEps) || std::abs(lwo.z) < T(kEps)) return 0;
        if (lwo.z < 0) {
            if (transmission == 0) return 0;
            T F = dielectric_fresnel(IOR, T(1), lwo.z);
            F = std::clamp<T>(F, T(0.1f), T(0.9f));
            if (lwi.z > 0) return (1 - F) * pdf_transmission(lwi, lwo);
            return F * pdf_inner_reflection(lwi, lwo);
        }
        if (lwi.z < 0) return w_transmission * pdf_transmission(lwi, lwo);
        const T pd = lwi.z * kInvPi<T>;
        T ps, pc;
        pdf_spec_cc(lwi, lwo, ps, pc);
        return w_diffuse * pd + w_specular * ps + w_clearcoat * pc;
    }

    // disney.cpp:457-497 (Q2 in reference mode); u3 = (selector, u, v)
    BSDFSample<T> sample(const V3<T>& wo, T su, T u, T v) const {
        BSDFSample<T> r;
        if (fringe(wo)) {
            // bsdf.hpp:141-153
            if (!geo.positive_z(wo)) return r;
            V3<T> l = cosine_sample_hemisphere(u, v);
            T pd = l.z * kInvPi<T>;
            if (pd < T(kEps)) return r;
            V3<T> wi = geo.to_global(l).normalize();
            r.wi = wi;
            r.f = C * (kInvPi<T> * normal_correct_factor(geo.z, shd.z, wi));
            r.pdf = pd;
            return r;
        }
        const V3<T> lwo = shd.to_local(wo).normalize();
        if (std::abs(lwo.z) < T(kEps)) return r;
        if (lwo.z < 0 && transmission == 0) return r;
        V3<T> lwi;
        bool ok = false;
        if (!fixed) {
            T sel = su;
            bool have = false;
            if (sel < w_diffuse) { lwi = cosine_sample_hemisphere(u, v); have = true; }
            else sel -= w_diffuse;
            // the transmission/specular/clearcoat chain always runs and overwrites
            if (sel < w_transmission) have = sample_transmission(lwo, u, v, lwi);
            else if (sel -= w_transmission, sel < w_specular) have = sample_specular(lwo, u, v, lwi);
            else have = sample_clearcoat(lwo, u, v, lwi);
            ok = have && !lwi.is_zero();
            if (!have) return r;
        } else if (lwo.z < 0) {
            T F = std::clamp<T>(dielectric_fresnel(IOR, T(1), lwo.z), T(0.1f), T(0.9f));
            if (su < F) ok = sample_inner_reflection(lwo, u, v, lwi);
            else ok = sample_transmission(lwo, u, v, lwi);
        } else {
            T sel = su;
            if (sel < w_diffuse) { lwi = cosine_sample_hemisphere(u, v); ok = true; }
            else if ((sel -= w_diffuse) < w_transmission) ok = sample_transmission(lwo, u, v, lwi);
            else if ((sel -= w_transmission) < w_specular) ok = sample_specular(lwo, u, v, lwi);
            else ok = sample_clearcoat(lwo, u, v, lwi);
        }
        if (!ok || lwi.is_zero()) return r;
        const V3<T> wi = shd.to_global(lwi);
        r.wi = wi;
        r.f = eval(wi, wo);
        r.pdf = pdf(wi, wo);
        return r;
    }
};

// Lambert through Phong(kd, ks = 0): aggregate_bsdf.hpp:36-118 +
// diffuse_bxdf.cpp:14-38.
//   Q22 eval() carries no normal-correction factor while sample() multiplies
//       by it (aggregate_bsdf.hpp:96-98), and sample() keeps a direction that
//       lies in the shading hemisphere but below the geometric surface.
//       Fixed: eval() includes the factor and such a sample is evaluated
//       through eval() (it is black there), so sample() and eval() agree.
template <class T>
class LambertRef {
public:
    Coord<T> geo, shd;
    RGB<T> albedo;
    bool fixed = false;
    LambertRef(const Coord<T>& g, const Coord<T>& s, const RGB<T>& a, bool fixed_mode = false)
        : geo(g), shd(s), albedo(a), fixed(fixed_mode) {}
    bool fringe(const V3<T>& w) const { return geo.positive_z(w) ^ shd.positive_z(w); }
    RGB<T> eval(const V3<T>& wi, const V3<T>& wo) const {
        if (fringe(wi) | fringe(wo)) {
            if (geo.positive_z(wi) & geo.positive_z(wo)) return albedo * kInvPi<T>;
            return RGB<T>(0);
        }
        const V3<T> lwi = shd.to_local(wi).normalize(), lwo = shd.to_local(wo).normalize();
        if (!lwi.z || !lwo.z) return RGB<T>(0);
        if (lwi.z <= 0 || lwo.z <= 0) return RGB<T>(0);
        if (fixed) return (albedo / kPi<T>) * normal_correct_factor(geo.z, shd.z, wi);
        return albedo / kPi<T>;
    }
    T pdf(const V3<T>& wi, const V3<T>& wo) const {
        if (fringe(wi) | fringe(wo)) {
            if (geo.positive_z(wi) ^ geo.positive_z(wo)) return 0;
            return geo.to_local(wi).normalize().z * kInvPi<T>;
        }
        const V3<T> lwi = shd.to_local(wi).normalize(), lwo = shd.to_local(wo).normalize();
        if (!lwi.z || !lwo.z) return 0;
        if (lwi.z <= 0 || lwo.z <= 0) return 0;
        return lwi.z * kInvPi<T>;
    }
    BSDFSample<T> sample(const V3<T>& wo, T su, T u, T v) const {
        BSDFSample<T> r;
        if (fringe(wo)) {
            if (!geo.positive_z(wo)) return r;
            V3<T> l = cosine_sample_hemisphere(u, v);
            T pd = l.z * kInvPi<T>;
            if (pd < T(kEps)) return r;
            V3<T> wi = geo.to_global(l).normalize();
            r.wi = wi;
            r.f = albedo * (kInvPi<T> * normal_correct_factor(geo.z, shd.z, wi));
            r.pdf = pd;
            return r;
        }
        const V3<T> lwo = shd.to_local(wo).normalize();
        if (!lwo.z) return r;
        (void)su;  // selector (su - 0.001) * 1 always picks the single lobe
        if (lwo.z <= 0) return r;
        V3<T> lwi = cosine_sample_hemisphere(u, v);
        T pd = lwi.z * kInvPi<T>;
        if (pd < T(0.001f)) return r;
        lwi = lwi.normalize();
        const V3<T> wi = shd.to_global(lwi);
        r.wi = wi;
        if (fixed && fringe(wi)) {
            r.f = eval(wi, wo);
            r.pdf = pdf(wi, wo);
            return r;
        }
        r.f = (albedo / kPi<T>) * normal_correct_factor(geo.z, shd.z, wi);
        r.pdf = pd;
        return r;
    }
};

// Phong(kd, ks, ns) through AggregateBSDF<2> (aggregate_bsdf.hpp:36-118):
// the constructor restates Phong::shading (phong.cpp:29-60).
template <class T>
class PhongRef {
public:
    Coord<T> geo, shd;
    RGB<T> kd, ks, albedo;
    T ns = 1, wd = T(0.5), ws = T(0.5);
    bool fixed = false;
    PhongRef(const Coord<T>& g, const Coord<T>& s, const RGB<T>& kd_in, const RGB<T>& ks_in, T ns_in,
             bool fixed_mode = false)
        : geo(g), shd(s), kd(kd_in), ks(ks_in), ns(ns_in), fixed(fixed_mode) {
        T dem = 1;   // phong.cpp:36-41
        for (int i = 0; i < 3; ++i) dem = std::max(dem, kd[i] + ks[i]);
        kd = kd / dem;
        ks = ks / dem;
        const T ld = kd.lum(), ls = ks.lum();   // phong.cpp:43-53
        wd = ld + ls > T(kEps) ? ld / (ld + ls) : T(0.5);
        ws = 1 - wd;
        albedo = kd + ks;   // phong.cpp:55-56
    }
    bool fringe(const V3<T>& w) const { return geo.positive_z(w) ^ shd.positive_z(w); }
    // phong_specular_bxdf.cpp:19-21: (e + 1) / (2 pi) cos^e
    static T pow_cos_pdf(T e, T c) { return (e + 1) * (T(0.5) * kInvPi<T>) * std::pow(c, e); }
    // phong_specular_bxdf.cpp:29-51 (value / ks and pdf are the same lobe)
    T spec_lobe(const V3<T>& lwi, const V3<T>& lwo) const {
        if (lwi.z <= 0 || lwo.z <= 0) return 0;
        const V3<T> ideal(-lwo.x, -lwo.y, lwo.z);   // reflect(lwo, z), reflection.hpp:12-14
        const T c = dot(ideal, lwi) / (lwi.length() * ideal.length());
        return c <= 0 ? T(0) : pow_cos_pdf(ns, c);
    }
    RGB<T> f_local(const V3<T>& lwi, const V3<T>& lwo) const {   // sum of the two evaluates
        if (lwi.z <= 0 || lwo.z <= 0) return RGB<T>(0);
        return kd * kInvPi<T> + ks * spec_lobe(lwi, lwo);
    }
    T pdf_local(const V3<T>& lwi, const V3<T>& lwo) const {   // weighted, before / sum(w)
        if (lwi.z <= 0 || lwo.z <= 0) return 0;
        return wd * lwi.z * kInvPi<T> + ws * spec_lobe(lwi, lwo);
    }
    RGB<T> eval(const V3<T>& wi, const V3<T>& wo) const {
        if (fringe(wi) | fringe(wo)) {   // bsdf.hpp:114-127
            if (geo.positive_z(wi) & geo.positive_z(wo)) return albedo * kInvPi<T>;
            return RGB<T>(0);
        }
        const V3<T> lwi = shd.to_local(wi).normalize(), lwo = shd.to_local(wo).normalize();
        if (!lwi.z || !lwo.z) return RGB<T>(0);
        if (fixed) return f_local(lwi, lwo) * normal_correct_factor(geo.z, shd.z, wi);
        return f_local(lwi, lwo);
    }
    T pdf(const V3<T>& wi, const V3<T>& wo) const {
        if (fringe(wi) | fringe(wo)) {
            if (geo.positive_z(wi) ^ geo.positive_z(wo)) return 0;
            return geo.to_local(wi).normalize().z * kInvPi<T>;
        }
        const V3<T> lwi = shd.to_local(wi).normalize(), lwo = shd.to_local(wo).normalize();
        if (!lwi.z || !lwo.z) return 0;
        const T sw = wd + ws;
        return sw > 0 ? pdf_local(lwi, lwo) / sw : T(0);
    }
    BSDFSample<T> sample(const V3<T>& wo, T su, T u, T v) const {
        BSDFSample<T> r;
        if (fringe(wo)) {   // bsdf.hpp:141-153
            if (!geo.positive_z(wo)) return r;
            V3<T> l = cosine_sample_hemisphere(u, v);
            T pd = l.z * kInvPi<T>;
            if (pd < T(kEps)) return r;
            V3<T> wi = geo.to_global(l).normalize();
            r.wi = wi;
            r.f = albedo * (kInvPi<T> * normal_correct_factor(geo.z, shd.z, wi));
            r.pdf = pd;
            return r;
        }
        const V3<T> lwo = shd.to_local(wo).normalize();
        if (!lwo.z) return r;
        const T sum = wd + ws;
        const bool diffuse = fixed ? su * sum <= wd : (su - T(0.001)) * sum <= wd;
        if (lwo.z <= 0) return r;   // both BXDF samplers reject wo below the surface
        V3<T> lwi;
        RGB<T> f_own;
        T pdf_own;
        if (diffuse) {   // diffuse_bxdf.cpp:20-31
            lwi = cosine_sample_hemisphere(u, v);
            pdf_own = lwi.z * kInvPi<T>;
            if (pdf_own < T(0.001f)) return r;
            f_own = kd * kInvPi<T>;
        } else {         // phong_specular_bxdf.cpp:7-17, :55-69
            const V3<T> ideal = V3<T>(-lwo.x, -lwo.y, lwo.z).normalize();
            const T ct = std::pow(u, 1 / (ns + 1));
            const T st = std::sqrt(std::max(T(0), 1 - ct * ct));
            const T phi = 2 * kPi<T> * v;
            const V3<T> l = V3<T>(st * std::cos(phi), st * std::sin(phi), ct).normalize();
            lwi = Coord<T>::from_z(ideal).to_global(l).normalize();
            if (lwi.z <= 0) return r;
            pdf_own = pow_cos_pdf(ns, l.z);   // the sampler's own pdf, phong_specular_bxdf.cpp:67
            f_own = ks * spec_lobe(lwi, lwo);
        }
        if (!(pdf_own > 0)) return r;   // BXDFSampleResult::is_valid, bxdf.hpp:16-18
        lwi = lwi.normalize();
        const V3<T> wi = shd.to_global(lwi);
        r.wi = wi;
        if (fixed) {
            r.f = eval(wi, wo);
            r.pdf = pdf(wi, wo);
            return r;
        }
        // aggregate_bsdf.hpp:84-98: the chosen lobe's pdf times its weight
        // plus the other lobe's pdf times its weight (not divided by sum(w)),
        // f of both lobes, times the normal correction
        if (diffuse) {
            r.f = f_own + ks * spec_lobe(lwi, lwo);
            r.pdf = pdf_own * wd + spec_lobe(lwi, lwo) * ws;
        } else {
            r.f = f_own + ((lwi.z <= 0 || lwo.z <= 0) ? RGB<T>(0) : kd * kInvPi<T>);
            r.pdf = pdf_own * ws + ((lwi.z <= 0 || lwo.z <= 0) ? T(0) : lwi.z * kInvPi<T>) * wd;
        }
        r.f = r.f * normal_correct_factor(geo.z, shd.z, wi);
        return r;
    }
};

// Glass(kr, kt, ior): GlassBSDF with color_reflect = kr and color_refract =
// sqrt(kt) (glass.cpp:96-112; kt = kr when map_kr == map_kt), Fresnel of
// DielectricFresnelPoint(ior, 1) (fresnel_point.hpp).  eval and pdf are 0.
template <class T>
class GlassRef {
public:
    Coord<T> geo, shd;
    RGB<T> kr, kt;
    T ior;
    GlassRef(const Coord<T>& g, const Coord<T>& s, const RGB<T>& kr_in, const RGB<T>& kt_in, T ior_in)
        : geo(g), shd(s), kr(kr_in), kt(sqrt(kt_in)), ior(ior_in) {}
    RGB<T> eval(const V3<T>&, const V3<T>&) const { return RGB<T>(0); }
    T pdf(const V3<T>&, const V3<T>&) const { return 0; }
    // glass.cpp:24-67
    BSDFSample<T> sample(const V3<T>& wo, T su, T, T) const {
        BSDFSample<T> r;
        r.delta = true;
        const V3<T> lwo = shd.to_local(wo).normalize();
        const V3<T> nor = lwo.z > 0 ? V3<T>(0, 0, 1) : V3<T>(0, 0, -1);
        const T fr = dielectric_fresnel(ior, T(1), lwo.z);
        V3<T> lwi;
        RGB<T> col;
        T w;
        if (su < fr) {
            lwi = V3<T>(-lwo.x, -lwo.y, lwo.z);
            col = kr;
            w = fr;
        } else {
            const T eta_i = lwo.z > 0 ? T(1) : ior;
            const T eta_t = lwo.z > 0 ? ior : T(1);
            const T eta = eta_i / eta_t;
            V3<T> t;
            if (!refract_dir(lwo, nor, eta, false, t)) return r;
            lwi = t.normalize();
            col = kt * (eta * eta);   // radiance transport
            w = 1 - fr;
        }
        const V3<T> wi = shd.to_global(lwi);
        r.wi = wi;
        r.f = col * (w / std::abs(lwi.z)) * normal_correct_factor(geo.z, shd.z, wi);
        r.pdf = w;
        return r;
    }
};

// phase_function.hpp:12-77
template <class T>
struct HGRef {
    T g;
    T phase(T u) const {
        T g2 = g * g;
        T dem = 1 + g2 - 2 * g * u;
        return (1 - g2) / (4 * kPi<T> * dem * std::sqrt(dem));
    }
    T eval(const V3<T>& wi, const V3<T>& wo) const { return phase(-cosv(wi, wo)); }
    BSDFSample<T> sample(const V3<T>& wo, T u, T v) const {
        T s = 2 * u - 1, uu;
        if (std::abs(g) < T(0.001f)) uu = s;
        else { T g2 = g * g; uu = (1 + g2 - mf::sqr((1 - g2) / (1 + g * s))) / (2 * g); }
        T ct = -uu;
        T st = std::sqrt(std::max<T>(0, 1 - ct * ct));
        T phi = 2 * kPi<T> * v;
        V3<T> l(st * std::sin(phi), st * std::cos(phi), ct);
        BSDFSample<T> r;
        r.wi = Coord<T>::from_z(wo).to_global(l);
        T pv = phase(uu);
        r.f = RGB<T>(pv);
        r.pdf = pv;
        return r;
    }
};

}  // namespace refcpu
