// refcpu — TEST INFRASTRUCTURE ONLY (see ref_math.hpp header).
//
// Image environment light: IBL over an HDR texture
// (src/light/environment_light.cpp:20-196) with Distribution1D / 2D
// (src/utility/distribution.hpp:12-110), the texture lookup of
// Texture2D::evaluate (src/core/texture.hpp:57-83, wrap_repeat) and
// LinearSampler::Sample2D (src/utility/sampler.hpp:12-25).  Arithmetic is
// float where the reference's `real` is float.
//   Q28 IBL::pdf() returns conditional func * marginal func_int where the
//       density sample_li draws is conditional func / marginal func_int
//       (distribution.hpp:102-106 vs :93-100).  Fixed mode divides.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

#include "ref_math.hpp"

namespace refcpu {

struct RefDist1D {   // distribution.hpp:12-78
    std::vector<float> func, cdf;
    float func_int = 0;
    RefDist1D(const float* f, int n) : func(f, f + n), cdf(size_t(n) + 1, 0.0f) {
        for (int i = 0; i < n; ++i) cdf[size_t(i) + 1] = cdf[size_t(i)] + func[size_t(i)] / float(n);
        func_int = cdf[size_t(n)];
        if (!func_int) {
            for (int i = 1; i < n + 1; ++i) cdf[size_t(i)] = float(i) / float(n);
        } else {
            for (int i = 1; i < n + 1; ++i) cdf[size_t(i)] /= func_int;
        }
    }
    int count() const { return int(func.size()); }
    float sample_continuous(float u, float* pdf, int* off) const {
        // last cdf entry <= u (lower_bound over the reversed cdf with greater<>)
        auto p = std::lower_bound(cdf.rbegin(), cdf.rend(), u, std::greater<float>());
        int offset = int(cdf.size()) - 1 - int(p - cdf.rbegin());
        offset = std::clamp<int>(offset, 0, int(cdf.size()) - 2);
        *off = offset;
        float du = u - cdf[size_t(offset)];
        if (cdf[size_t(offset) + 1] - cdf[size_t(offset)] > 0) du /= cdf[size_t(offset) + 1] - cdf[size_t(offset)];
        *pdf = func_int > 0 ? func[size_t(offset)] / func_int : 0;
        return (float(offset) + du) / float(count());
    }
};

class RefEnvMap {
public:
    int w = 0, h = 0;
    std::vector<float> rgb;
    float rot[9];
    std::vector<RefDist1D> cond;
    std::vector<RefDist1D> marg;   // one element
    RGB<float> avg_radiance;       // IBL::avg_radiance: sum of L * texel solid angle

    RefEnvMap(int width, int height, const float* texels, const float* world_to_light)
        : w(width), h(height), rgb(texels, texels + size_t(width) * size_t(height) * 3) {
        std::copy(world_to_light, world_to_light + 9, rot);
        const float pi = kPi<float>;
        for (int y = 0; y < h; ++y) {   // environment_light.cpp:24-35
            const float v0 = float(y) / float(h), v1 = float(y + 1) / float(h);
            for (int x = 0; x < w; ++x) {
                const float u0 = float(x) / float(w), u1 = float(x + 1) / float(w);
                const float da = std::abs(2 * pi * (u1 - u0) * (std::cos(pi * v1) - std::cos(pi * v0)));
                avg_radiance += evaluate((u0 + u1) / 2, (v0 + v1) / 2) * da;
            }
        }
        const int nw = std::min(w, 360), nh = std::min(h, 180);   // environment_light.cpp:36-79
        std::vector<float> img(size_t(nw) * size_t(nh));
        for (int y = 0; y < nh; ++y) {
            const float y0 = float(y) / float(nh), y1 = float(y + 1) / float(nh);
            const int sy0 = std::max<int>(0, int(std::floor(y0 * float(h))) - 1);
            const int sy1 = std::min<int>(h - 1, int(std::floor(y1 * float(h))) + 1);
            for (int x = 0; x < nw; ++x) {
                const float x0 = float(x) / float(nw), x1 = float(x + 1) / float(nw);
                const int sx0 = std::max<int>(0, int(std::floor(x0 * float(w))) - 1);
                const int sx1 = std::min<int>(w - 1, int(std::floor(x1 * float(w))) + 1);
                float lum = 0;
                for (int sy = sy0; sy <= sy1; ++sy)
                    for (int sx = sx0; sx <= sx1; ++sx)
                        lum += evaluate(float(sx + 0.5) / float(w), float(sy + 0.5) / float(h)).lum();
                const float da = std::abs(2 * pi * (x1 - x0) * pi * (y1 - y0) * std::sin(pi * 0.5f * (y1 + y0)));
                img[size_t(y) * size_t(nw) + size_t(x)] = lum * da;
            }
        }
        cond.reserve(size_t(nh));
        std::vector<float> mf;
        for (int y = 0; y < nh; ++y) {   // Distribution2D, distribution.hpp:80-91
            cond.emplace_back(&img[size_t(y) * size_t(nw)], nw);
            mf.push_back(cond.back().func_int);
        }
        marg.emplace_back(mf.data(), nh);
    }

    RGB<float> evaluate(double u, double v) const {
        u = std::clamp(u - std::floor(u), 0.0, 1.0) * (w - 1);
        v = std::clamp(v - std::floor(v), 0.0, 1.0) * (h - 1);
        const int u0 = std::clamp(int(u), 0, w - 1), u1 = std::clamp(u0 + 1, 0, w - 1);
        const int v0 = std::clamp(int(v), 0, h - 1), v1 = std::clamp(v0 + 1, 0, h - 1);
        const double du = u - u0, dv = v - v0;
        RGB<float> out;
        for (int k = 0; k < 3; ++k) {
            auto at = [&](int x, int y) { return double(rgb[(size_t(y) * size_t(w) + size_t(x)) * 3 + size_t(k)]); };
            out[k] = float((at(u0, v0) * (1.0 - du) + at(u1, v0) * du) * (1.0 - dv) +
                           (at(u0, v1) * (1.0 - du) + at(u1, v1) * du) * dv);
        }
        return out;
    }
    V3<float> to_light(const V3<float>& d) const {
        return V3<float>(rot[0] * d.x + rot[1] * d.y + rot[2] * d.z, rot[3] * d.x + rot[4] * d.y + rot[5] * d.z,
                         rot[6] * d.x + rot[7] * d.y + rot[8] * d.z);
    }
    V3<float> to_world(const V3<float>& l) const {   // inverse of a rotation: its transpose
        return V3<float>(rot[0] * l.x + rot[3] * l.y + rot[6] * l.z, rot[1] * l.x + rot[4] * l.y + rot[7] * l.z,
                         rot[2] * l.x + rot[5] * l.y + rot[8] * l.z);
    }
    static void uv_of(const V3<float>& l, float& u, float& v, float& sin_theta) {   // geometry.hpp:912-921
        float phi = (!l.y && !l.x) ? 0.0f : std::atan2(l.y, l.x);
        if (phi < 0) phi += 2 * kPi<float>;
        const float theta = std::acos(std::clamp<float>(l.z, -1, 1));
        u = std::clamp<float>(phi / (2 * kPi<float>), 0, 1);
        v = std::clamp<float>(theta / kPi<float>, 0, 1);
        sin_theta = std::sin(theta);
    }
    // light_emit (environment_light.cpp:179-188)
    RGB<float> light_emit(const V3<float>& dir) const {
        float u, v, st;
        uv_of(to_light(dir.normalize()).normalize(), u, v, st);
        return evaluate(u, v);
    }
    // pdf (environment_light.cpp:189-196), Q28 in reference mode
    float pdf(const V3<float>& dir, bool quirks) const {
        float u, v, st;
        uv_of(to_light(dir.normalize()).normalize(), u, v, st);
        if (st == 0) return 0;
        const int nu = cond[0].count(), nv = marg[0].count();
        const int iu = std::clamp<int>(int(u * float(nu)), 0, nu - 1);
        const int iv = std::clamp<int>(int(v * float(nv)), 0, nv - 1);
        const float f = cond[size_t(iv)].func[size_t(iu)];
        const float p = quirks ? f * marg[0].func_int : (marg[0].func_int > 0 ? f / marg[0].func_int : 0.0f);
        return p / (2 * kPi<float> * kPi<float> * st);
    }
    // sample_li (environment_light.cpp:92-119): false when no direction
    bool sample(float su, float sv, V3<float>& wi, float& pdf_out, RGB<float>& L) const {
        float pm, pc;
        int iv, iu;
        const float d1 = marg[0].sample_continuous(sv, &pm, &iv);
        const float d0 = cond[size_t(iv)].sample_continuous(su, &pc, &iu);
        const float map_pdf = pc * pm;
        if (map_pdf == 0) return false;
        const float theta = d1 * kPi<float>, phi = d0 * 2 * kPi<float>;
        const float st = std::sin(theta), ct = std::cos(theta);
        if (st == 0) return false;
        wi = to_world(V3<float>(st * std::cos(phi), st * std::sin(phi), ct)).normalize();
        pdf_out = map_pdf / (2 * kPi<float> * kPi<float> * st);
        L = light_emit(wi);
        return true;
    }
};

}  // namespace refcpu
