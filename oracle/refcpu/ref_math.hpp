// refcpu — CPU restatement of wyzwzz/Tracer's unidirectional path tracer.
//
// TEST INFRASTRUCTURE ONLY.  This directory is the oracle: tests, the smoke
// check and the CPU-baseline leg of bench.py may call it, as the checker.  The
// product (tracer_amd/) never links, loads or calls anything in oracle/.
//
// ref_math.hpp — vector math with the reference's rounding behaviour.
//   * Vector3<T>                     src/utility/geometry.hpp (class Vector3)
//   * cross() evaluated in double     src/utility/geometry.hpp:583-588
//   * length()/normalize() via double src/utility/geometry.hpp:443-445, 564-575
//   * operator/(U) multiplies by a double reciprocal  geometry.hpp:430-433
// Every template is instantiated for float (reference semantics) and double
// (the FP64 twin used to pin the GPU's FP32 BSDF).
#pragma once
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <limits>

namespace refcpu {

template <class T> constexpr T kPi = T(3.141592653589793238462643383);
template <class T> constexpr T kInvPi = T(1) / kPi<T>;
// src/common.hpp:47 — eps = 3e-4 (used for ray offsets and many guards)
constexpr float kEps = 3e-4f;

template <class T>
struct V3 {
    T x = 0, y = 0, z = 0;
    V3() = default;
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    template <class U> explicit V3(const V3<U>& o) : x(T(o.x)), y(T(o.y)), z(T(o.z)) {}
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    V3 operator+(const V3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    V3 operator-(const V3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    V3 operator-() const { return {-x, -y, -z}; }
    V3 operator*(T s) const { return {x * s, y * s, z * s}; }
    V3& operator+=(const V3& o) { x += o.x; y += o.y; z += o.z; return *this; }
    bool is_zero() const { return x == 0 && y == 0 && z == 0; }
    // geometry.hpp:443 — the sum of squares is formed in T, then widened.
    double length_squared() const { return x * x + y * y + z * z; }
    double length() const { return std::sqrt(length_squared()); }
    // geometry.hpp:430-433 — division multiplies by a double reciprocal.
    V3 div(double f) const {
        double inv = 1.0 / f;
        return V3(T(x * inv), T(y * inv), T(z * inv));
    }
    V3 normalize() const { return div(length()); }
};
template <class T> inline V3<T> operator*(T s, const V3<T>& v) { return v * s; }
template <class T> inline T dot(const V3<T>& a, const V3<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// geometry.hpp:583-588 — cross product formed in double, rounded to T.
template <class T> inline V3<T> cross(const V3<T>& a, const V3<T>& b) {
    double ax = a.x, ay = a.y, az = a.z, bx = b.x, by = b.y, bz = b.z;
    return V3<T>(T(ay * bz - az * by), T(az * bx - ax * bz), T(ax * by - ay * bx));
}
template <class T> inline V3<T> normalize(const V3<T>& v) { return v.normalize(); }
template <class T> inline V3<T> vmin(const V3<T>& a, const V3<T>& b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
template <class T> inline V3<T> vmax(const V3<T>& a, const V3<T>& b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }
// geometry.hpp:943-945 — cos(w1,w2) = dot / (|w1||w2|)
template <class T> inline T cosv(const V3<T>& a, const V3<T>& b) { return T(dot(a, b) / (a.length() * b.length())); }
template <class T> inline T abs_cosv(const V3<T>& a, const V3<T>& b) { return std::abs(cosv(a, b)); }

using V3f = V3<float>;
using V3d = V3<double>;

// RGB spectrum, src/core/spectrum.hpp:12-162
template <class T>
struct RGB {
    T r = 0, g = 0, b = 0;
    RGB() = default;
    explicit RGB(T v) : r(v), g(v), b(v) {}
    RGB(T a, T c, T d) : r(a), g(c), b(d) {}
    template <class U> explicit RGB(const RGB<U>& o) : r(T(o.r)), g(T(o.g)), b(T(o.b)) {}
    T operator[](int i) const { return i == 0 ? r : (i == 1 ? g : b); }
    T& operator[](int i) { return i == 0 ? r : (i == 1 ? g : b); }
    RGB operator+(const RGB& o) const { return {r + o.r, g + o.g, b + o.b}; }
    RGB operator-(const RGB& o) const { return {r - o.r, g - o.g, b - o.b}; }
    RGB operator*(const RGB& o) const { return {r * o.r, g * o.g, b * o.b}; }
    RGB operator*(T s) const { return {r * s, g * s, b * s}; }
    RGB operator/(T s) const { return {r / s, g / s, b / s}; }
    RGB& operator+=(const RGB& o) { r += o.r; g += o.g; b += o.b; return *this; }
    RGB& operator*=(const RGB& o) { r *= o.r; g *= o.g; b *= o.b; return *this; }
    RGB& operator*=(T s) { r *= s; g *= s; b *= s; return *this; }
    // spectrum.hpp:94-96
    T lum() const { return T(0.2126) * r + T(0.7152) * g + T(0.0722) * b; }
    // "is_back" = all components zero (spectrum.hpp)
    bool is_black() const { return r == 0 && g == 0 && b == 0; }
    bool is_finite() const { return std::isfinite(r) && std::isfinite(g) && std::isfinite(b); }
    T max_comp() const { return std::max(r, std::max(g, b)); }
};
template <class T> inline RGB<T> operator*(T s, const RGB<T>& c) { return c * s; }
template <class T> inline RGB<T> sqrt(const RGB<T>& c) { return {std::sqrt(c.r), std::sqrt(c.g), std::sqrt(c.b)}; }
template <class T, class F> inline T mix(const T& a, const T& b, F f) { return (1 - f) * a + f * b; }
template <class T> inline RGB<T> mixc(const RGB<T>& a, const RGB<T>& b, T f) { return a * (1 - f) + b * f; }

// Orthonormal frame, src/utility/coordinate.hpp:15-116.  The constructor
// normalises all three axes (coordinate.hpp:78-82).
template <class T>
struct Coord {
    V3<T> x{1, 0, 0}, y{0, 1, 0}, z{0, 0, 1};
    Coord() = default;
    Coord(const V3<T>& a, const V3<T>& b, const V3<T>& c) : x(a.normalize()), y(b.normalize()), z(c.normalize()) {}
    V3<T> to_local(const V3<T>& v) const { return {dot(x, v), dot(y, v), dot(z, v)}; }
    V3<T> to_global(const V3<T>& v) const { return x * v.x + y * v.y + z * v.z; }
    bool positive_z(const V3<T>& v) const { return dot(v, z) > 0; }
    // coordinate.hpp:84-93
    static Coord from_z(const V3<T>& nz) {
        V3<T> z = nz.normalize();
        V3<T> y;
        if (std::abs(z.x) > std::abs(z.y)) y = V3<T>(-z.z, 0, z.x).normalize();
        else y = V3<T>(0, z.z, -z.y).normalize();
        return Coord(cross(y, z), y, z);
    }
};

// geometry.hpp:887-894 — build two tangents from one vector.
template <class T> inline void coordinate_system(const V3<T>& v1, V3<T>& v2, V3<T>& v3) {
    if (std::abs(v1.x) > std::abs(v1.y)) v2 = V3<T>(-v1.z, 0, v1.x).div(std::sqrt(v1.x * v1.x + v1.z * v1.z));
    else v2 = V3<T>(0, v1.z, -v1.y).div(std::sqrt(v1.y * v1.y + v1.z * v1.z));
    v3 = cross(v1, v2);
}

// coordinate.hpp:43-47
template <class T> inline T normal_correct_factor(const V3<T>& g, const V3<T>& s, const V3<T>& wi) {
    T dem = std::abs(cosv(g, wi));
    return dem < T(kEps) ? T(1) : std::abs(cosv(s, wi)) / dem;
}

// coordinate.hpp:54-68 — tangents from positional/UV differences.
template <class T>
inline void compute_ss_ts(const V3<T>& AB, const V3<T>& AC, T abx, T aby, T acx, T acy, const V3<T>& n, V3<T>& ss, V3<T>& ts) {
    const T m00 = abx, m01 = aby, m10 = acx, m11 = acy;
    const T det = m00 * m11 - m01 * m10;
    if (det != 0) {
        const T inv_det = 1 / det;
        ss = AB * (m11 * inv_det) - AC * (m01 * inv_det);
        ts = AB * (m10 * inv_det) - AC * (m00 * inv_det);
    } else {
        coordinate_system(n, ss, ts);
    }
}

// geometry.hpp:912-933 — local spherical helpers.
template <class T> inline T local_phi(const V3<T>& w) {
    if (w.y == 0 && w.x == 0) return 0;
    T r = std::atan2(w.y, w.x);
    return r < 0 ? r + 2 * kPi<T> : r;
}
template <class T> inline T local_tan_theta(const V3<T>& w) {
    const T t = 1 - w.z * w.z;
    if (t <= 0) return 0;
    return std::sqrt(t) / w.z;
}
template <class T> inline T local_cos_to_sin(T c) { return std::sqrt(std::max<T>(0, 1 - c * c)); }

// sampling.hpp:43-46
template <class T> inline T power_heuristic(T f, T g) { return (f * f) / (f * f + g * g); }

// sampling.cpp:8-22
template <class T> inline void concentric_sample_disk(T u, T v, T& ox, T& oy) {
    T x = u * 2 - 1, y = v * 2 - 1;
    if (x == 0 && y == 0) { ox = oy = 0; return; }
    T theta, r;
    if (std::abs(x) > std::abs(y)) { r = x; theta = kPi<T> / 4 * (y / x); }
    else { r = y; theta = kPi<T> / 2 - kPi<T> / 4 * (x / y); }
    ox = r * std::cos(theta);
    oy = r * std::sin(theta);
}
// sampling.cpp:24-28
template <class T> inline V3<T> cosine_sample_hemisphere(T u, T v) {
    T dx, dy;
    concentric_sample_disk(u, v, dx, dy);
    T z = std::sqrt(std::max<T>(0, 1 - dx * dx - dy * dy));
    return {dx, dy, z};
}

}  // namespace refcpu
