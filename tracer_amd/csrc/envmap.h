// Image environment light on the host: the texture lookup and the 2D sampling
// distribution of IBL (src/light/environment_light.cpp:20-80), built once per
// scene upload and copied to the device as plain arrays.
#pragma once
#include <cstdint>
#include <vector>

namespace trcapi {

struct EnvMapHost {
    int w = 0, h = 0;
    std::vector<float> rgb;     // h rows of w texels, 3 floats each, row 0 at v = 0
    float rot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};   // world -> light, row-major
    // Distribution2D over an nu x nv grid (distribution.hpp:80-110): per row v
    // the conditional func (nu) and cdf (nu + 1), then the marginal over rows.
    int nu = 0, nv = 0;
    std::vector<float> cond_func, cond_cdf, marg_func, marg_cdf;
    float marg_int = 0.0f;
    double radiance_integral[3] = {0, 0, 0};   // sum of L(texel centre) * solid angle of the texel
};

// Texture2D::evaluate with wrap_repeat (core/texture.hpp:57-83) and
// LinearSampler::Sample2D (utility/sampler.hpp:12-25).
void env_lookup(const EnvMapHost& e, double u, double v, double out[3]);

// Fills the distribution and the radiance integral from w, h, rgb.
void build_envmap(EnvMapHost& e);

}  // namespace trcapi
