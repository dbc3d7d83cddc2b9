// refcpu — TEST INFRASTRUCTURE ONLY (see ref_math.hpp header).
//
// Ray/triangle tests.
//  * mt_intersect(): the reference's Moller-Trumbore test,
//    src/shape/triangle.cpp:117-143 (any-hit) and :145-175 (closest-hit core):
//    cross products in double rounded to float, exact-zero determinant test,
//    inclusive edges, t accepted in [t_min, t_max].  alpha weights vertex B,
//    beta weights vertex C.
//  * wt_intersect(): the watertight test of Woop, Benthin and Wald, "Watertight
//    Ray/Triangle Intersection", JCGT 2(1), 2013 — the algorithm the GPU
//    kernels implement.  Restated here in plain FP32 with the same operation
//    order (and explicit fmaf where the GPU uses fma) so that a brute-force
//    loop over all triangles is a bit-exact oracle for the GPU's BVH
//    traversal.  Built with -ffp-contract=off.
//  * Ties in t are broken by the lowest primitive index, for both the
//    brute-force oracle and the GPU.  (The reference lets the last-tested
//    primitive win, triangle.cpp:174 + shape/geometry.cpp:52, which depends
//    on traversal order.)
#pragma once
#include "ref_math.hpp"
#include <cfloat>

namespace refcpu {

struct MTHit { float t, alpha, beta; };

// triangle.cpp:151-174
inline bool mt_intersect(const V3f& o, const V3f& d, float tmin, float tmax,
                         const V3f& A, const V3f& B, const V3f& C, MTHit* out) {
    const V3f AB = B - A;
    const V3f AC = C - A;
    V3f s1 = cross(d, AC);
    float div = dot(s1, AB);
    if (!div) return false;
    float inv_div = 1 / div;
    const V3f AO = o - A;
    float alpha = dot(AO, s1) * inv_div;
    if (alpha < 0) return false;
    V3f s2 = cross(AO, AB);
    float beta = dot(d, s2) * inv_div;
    if (beta < 0 || alpha + beta > 1) return false;
    float t = dot(AC, s2) * inv_div;
    if (t < tmin || t > tmax) return false;
    if (out) *out = {t, alpha, beta};
    return true;
}

// Per-ray constants of the watertight test (Woop et al. 2013, section 3).
struct WTRay {
    V3f o;
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

inline WTRay wt_setup(const V3f& o, const V3f& d) {
    WTRay r;
    r.o = o;
    float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    int kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (d[kz] < 0.0f) std::swap(kx, ky);
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = d[kx] / d[kz];
    r.Sy = d[ky] / d[kz];
    r.Sz = 1.0f / d[kz];
    return r;
}

struct WTHit { float t, b1, b2; };  // b1 weights vertex 1 (B), b2 weights vertex 2 (C)

// Returns true for a hit with tmin <= t <= tmax.
inline bool wt_intersect(const WTRay& r, float tmin, float tmax,
                         const V3f& v0, const V3f& v1, const V3f& v2, WTHit* out) {
    const V3f A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
    const float Ax = std::fmaf(-r.Sx, A[r.kz], A[r.kx]);
    const float Ay = std::fmaf(-r.Sy, A[r.kz], A[r.ky]);
    const float Bx = std::fmaf(-r.Sx, B[r.kz], B[r.kx]);
    const float By = std::fmaf(-r.Sy, B[r.kz], B[r.ky]);
    const float Cx = std::fmaf(-r.Sx, C[r.kz], C[r.kx]);
    const float Cy = std::fmaf(-r.Sy, C[r.kz], C[r.ky]);
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        // Exact fallback on edges (the products of two floats are exact in double).
        U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
        V = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
        W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.Sz * A[r.kz];
    const float Bz = r.Sz * B[r.kz];
    const float Cz = r.Sz * C[r.kz];
    const float T = U * Az + V * Bz + W * Cz;
    const float rcp = 1.0f / det;
    const float t = T * rcp;
    if (!(t >= tmin && t <= tmax)) return false;
    if (out) *out = {t, V * rcp, W * rcp};
    return true;
}

}  // namespace refcpu
