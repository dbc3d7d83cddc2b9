// IBL preprocessing (src/light/environment_light.cpp:20-80) and the
// piecewise-constant distributions of src/utility/distribution.hpp:12-110.
#include "envmap.h"

#include <algorithm>
#include <cmath>

namespace trcapi {

namespace {
constexpr double kPiD = 3.14159265358979323846;

// Distribution1D (distribution.hpp:12-31): cdf[i + 1] = cdf[i] + f[i] / n,
// normalised by its last entry (func_int); a zero function gets a uniform cdf.
float build_1d(const float* f, int n, float* func, float* cdf) {
    cdf[0] = 0.0f;
    for (int i = 0; i < n; ++i) {
        func[i] = f[i];
        cdf[i + 1] = cdf[i] + f[i] / float(n);
    }
    const float func_int = cdf[n];
    if (func_int == 0.0f) {
        for (int i = 1; i <= n; ++i) cdf[i] = float(i) / float(n);
    } else {
        for (int i = 1; i <= n; ++i) cdf[i] /= func_int;
    }
    return func_int;
}
}  // namespace

void env_lookup(const EnvMapHost& e, double u, double v, double out[3]) {
    u = std::clamp(u - std::floor(u), 0.0, 1.0);
    v = std::clamp(v - std::floor(v), 0.0, 1.0);
    u *= double(e.w - 1);
    v *= double(e.h - 1);
    const int u0 = std::clamp(int(u), 0, e.w - 1), u1 = std::clamp(u0 + 1, 0, e.w - 1);
    const int v0 = std::clamp(int(v), 0, e.h - 1), v1 = std::clamp(v0 + 1, 0, e.h - 1);
    const double du = u - u0, dv = v - v0;
    auto at = [&](int x, int y, int k) { return double(e.rgb[(size_t(y) * size_t(e.w) + size_t(x)) * 3 + size_t(k)]); };
    for (int k = 0; k < 3; ++k)
        out[k] = (at(u0, v0, k) * (1.0 - du) + at(u1, v0, k) * du) * (1.0 - dv) +
                 (at(u0, v1, k) * (1.0 - du) + at(u1, v1, k) * du) * dv;
}

void build_envmap(EnvMapHost& e) {
    // The reference computes in float (real); so does this restatement, so that
    // cell borders land on the same texels.
    const int W = e.w, H = e.h;
    const float pi = float(kPiD);
    // avg_radiance: radiance at each texel centre times the texel's solid angle
    float avg[3] = {0.0f, 0.0f, 0.0f};
    for (int y = 0; y < H; ++y) {
        const float v0 = float(y) / float(H), v1 = float(y + 1) / float(H);
        for (int x = 0; x < W; ++x) {
            const float u0 = float(x) / float(W), u1 = float(x + 1) / float(W);
            const float da = std::abs(2.0f * pi * (u1 - u0) * (std::cos(pi * v1) - std::cos(pi * v0)));
            double c[3];
            env_lookup(e, 0.5f * (u0 + u1), 0.5f * (v0 + v1), c);
            for (int k = 0; k < 3; ++k) avg[k] += float(c[k]) * da;
        }
    }
    for (int k = 0; k < 3; ++k) e.radiance_integral[k] = avg[k];
    // luminance of at most 360 x 180 cells, each summing the texels it covers
    // plus a one-texel border, times the cell's solid angle
    const int nw = std::min(W, 360), nh = std::min(H, 180);
    std::vector<float> img(size_t(nw) * size_t(nh));
    for (int y = 0; y < nh; ++y) {
        const float y0 = float(y) / float(nh), y1 = float(y + 1) / float(nh);
        const int sy0 = std::max(0, int(std::floor(y0 * float(H))) - 1);
        const int sy1 = std::min(H - 1, int(std::floor(y1 * float(H))) + 1);
        for (int x = 0; x < nw; ++x) {
            const float x0 = float(x) / float(nw), x1 = float(x + 1) / float(nw);
            const int sx0 = std::max(0, int(std::floor(x0 * float(W))) - 1);
            const int sx1 = std::min(W - 1, int(std::floor(x1 * float(W))) + 1);
            float lum = 0.0f;
            for (int sy = sy0; sy <= sy1; ++sy)
                for (int sx = sx0; sx <= sx1; ++sx) {
                    double c[3];
                    env_lookup(e, double(float(sx + 0.5) / float(W)), double(float(sy + 0.5) / float(H)), c);
                    const float cf[3] = {float(c[0]), float(c[1]), float(c[2])};
                    lum += 0.2126f * cf[0] + 0.7152f * cf[1] + 0.0722f * cf[2];
                }
            const float da = std::abs(2.0f * pi * (x1 - x0) * pi * (y1 - y0) * std::sin(pi * 0.5f * (y1 + y0)));
            img[size_t(y) * size_t(nw) + size_t(x)] = lum * da;
        }
    }
    e.nu = nw;
    e.nv = nh;
    e.cond_func.assign(size_t(nw) * size_t(nh), 0.0f);
    e.cond_cdf.assign(size_t(nw + 1) * size_t(nh), 0.0f);
    std::vector<float> row_int(static_cast<size_t>(nh));
    for (int y = 0; y < nh; ++y)
        row_int[size_t(y)] = build_1d(&img[size_t(y) * size_t(nw)], nw, &e.cond_func[size_t(y) * size_t(nw)],
                                      &e.cond_cdf[size_t(y) * size_t(nw + 1)]);
    e.marg_func.assign(static_cast<size_t>(nh), 0.0f);
    e.marg_cdf.assign(static_cast<size_t>(nh) + 1, 0.0f);
    e.marg_int = build_1d(row_int.data(), nh, e.marg_func.data(), e.marg_cdf.data());
}

}  // namespace trcapi
