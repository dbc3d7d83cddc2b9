// Device math, random numbers and shared constants for the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define TR_HD __host__ __device__ __forceinline__
#define TR_D __device__ __forceinline__

namespace trhip {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;
// Offset / guard epsilon of the reference (src/common.hpp:47).
constexpr float kEps = 3e-4f;

struct f3 {
    float x, y, z;
};
TR_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
TR_HD f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
TR_HD f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
TR_HD f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
TR_HD f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
TR_HD f3 operator*(float s, f3 a) { return {a.x * s, a.y * s, a.z * s}; }
TR_HD f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
TR_HD f3 operator/(f3 a, float s) { float r = 1.0f / s; return {a.x * r, a.y * r, a.z * r}; }
TR_HD f3& operator+=(f3& a, f3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
TR_HD f3& operator*=(f3& a, f3 b) { a.x *= b.x; a.y *= b.y; a.z *= b.z; return a; }
TR_HD f3& operator*=(f3& a, float s) { a.x *= s; a.y *= s; a.z *= s; return a; }
TR_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
TR_HD f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
TR_HD float length(f3 a) { return sqrtf(dot(a, a)); }
TR_HD f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
TR_HD float lum(f3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
TR_HD bool is_black(f3 c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }
TR_HD float max3(f3 c) { return fmaxf(c.x, fmaxf(c.y, c.z)); }
TR_HD f3 f3s(float s) { return {s, s, s}; }
TR_HD f3 lerp3(f3 a, f3 b, float t) { return a * (1.0f - t) + b * t; }
TR_HD f3 xyz(float4 v) { return {v.x, v.y, v.z}; }
TR_HD float4 f4(f3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
TR_HD bool finite3(f3 c) { return isfinite(c.x) && isfinite(c.y) && isfinite(c.z); }

// Orthonormal basis from a unit normal (Duff et al., "Building an
// Orthonormal Basis, Revisited", JCGT 6(1), 2017).
TR_HD void onb(f3 n, f3& t, f3& b) {
    const float s = copysignf(1.0f, n.z);
    const float a = -1.0f / (s + n.z);
    const float bb = n.x * n.y * a;
    t = {1.0f + s * n.x * n.x * a, s * bb, -s * n.x};
    b = {bb, s + n.y * n.y * a, -n.y};
}

// PCG32 (O'Neill 2014): 64-bit LCG, XSH-RR output of the previous state.
struct Pcg {
    uint64_t state;
    uint64_t inc;
    TR_HD uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    // top 24 bits * 2^-24, in [0, 1)
    TR_HD float uniform() { return float(next() >> 8) * (1.0f / 16777216.0f); }
};
TR_HD uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
// pcg32_srandom_r(initstate, initseq)
TR_HD Pcg pcg_seed(uint64_t initstate, uint64_t initseq) {
    Pcg p;
    p.state = 0;
    p.inc = (initseq << 1u) | 1u;
    p.next();
    p.state += initstate;
    p.next();
    return p;
}
// The stream of one camera sample: keyed by the render seed, the pixel and
// the sample index; the pixel also selects the PCG stream.
TR_HD Pcg pcg_for_sample(uint64_t seed, uint32_t pixel, uint32_t sample) {
    return pcg_seed(mix64(seed ^ mix64(uint64_t(pixel) * 0x9E3779B97F4A7C15ULL + sample)), pixel);
}

// Power heuristic with beta = 2 (src/core/sampling.hpp:43-46).
TR_HD float power_heuristic(float f, float g) {
    const float ff = f * f, gg = g * g;
    return ff / (ff + gg);
}

// Concentric disk mapping (src/core/sampling.cpp:8-22).
TR_HD void concentric_disk(float u, float v, float& ox, float& oy) {
    const float x = 2.0f * u - 1.0f, y = 2.0f * v - 1.0f;
    if (x == 0.0f && y == 0.0f) { ox = oy = 0.0f; return; }
    float r, th;
    if (fabsf(x) > fabsf(y)) { r = x; th = (kPi * 0.25f) * (y / x); }
    else { r = y; th = kPi * 0.5f - (kPi * 0.25f) * (x / y); }
    float s, c;
    sincosf(th, &s, &c);
    ox = r * c;
    oy = r * s;
}
TR_HD f3 cosine_hemisphere(float u, float v) {
    float dx, dy;
    concentric_disk(u, v, dx, dy);
    return {dx, dy, sqrtf(fmaxf(0.0f, 1.0f - dx * dx - dy * dy))};
}

}  // namespace trhip
