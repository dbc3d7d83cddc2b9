// Learned light-sample distribution for next-event estimation (the light
// part of path guiding; no reference counterpart).
//
// The reference picks every light at every vertex and samples a point
// uniformly over its area (path_trace_renderer.cpp:141-153,
// diffuse_light.cpp:44-60).  When most of a light is hidden from most of the
// scene (a small opening between the light and the rest), nearly every such
// sample is occluded.  Here the scene box is cut into G^3 cells, and each
// emissive triangle into S x S strata of equal area (the strata of its
// uniform-area parameterisation bx = 1 - sqrt(u), by = v sqrt(u),
// Triangle::sample, triangle.cpp:61-97).  Bins are (triangle, stratum) pairs
// plus the environment and the sun.  Training passes sum, per cell and bin,
// the luminance of the unoccluded (transmitted) light samples divided by
// their sampling density: a Monte Carlo estimate of each bin's share of the
// light arriving in the cell.  The receiving vertex's BSDF and cosine are left
// out: they differ between the vertices that share a cell and only add noise
// (on configuration 4 leaving out the BSDF lowered the guided relMSE by 9%).  Later passes draw a bin from
// a * learned + (1 - a) * (power-based default), a = (1 - eps) h / (h + h0)
// for a cell with h unoccluded training samples (a cell that saw little keeps
// mostly the default), by an alias table, then a point uniformly within the
// bin.  The default part keeps every bin that can
// contribute at a positive density, so the estimator stays unbiased; its MIS
// weights use the same density at both ends of a connection (the cell of the
// vertex that sampled the light travels with the path).
#pragma once
#include "tr_device.h"

namespace trhip {

struct DevLightCache {
    int grid;             // cells per axis; 0 = the cache is off
    int strata;           // S: strata per side of each emissive triangle
    int bins;             // B = num_area_lights * S^2 + 2
    int num_area;         // emissive triangles
    f3 lo, scale;         // cell of x: floor((x - lo) * scale), clamped
    const float* pmf;     // [cell][bin] sampling probability of each bin
    const float2* table;  // [cell][bin] alias table: (probability of keeping the bin, alias as int bits)
    unsigned long long* sum;   // [cell][bin] training sums, fixed point (kLcFixed)
    unsigned int* hits;   // [cell] unoccluded training samples
    int train;            // shadow rays of this pass add their contributions to sum
    int use;              // next-event estimation of this pass samples the cache
};

// Fixed-point scale of the training sums: integer atomics make them (and the
// distribution built from them) independent of the order of the additions.
constexpr float kLcFixed = 1048576.0f;     // 2^20
constexpr float kLcMaxValue = 1.0e6f;      // contributions above this are clamped
constexpr float kLcHits0 = 64.0f;          // h0: unoccluded samples at which a cell's learned part has half its weight

// Adds one unoccluded training sample to its cell and bin.
TR_D void lc_record(const DevLightCache& lc, int cell, int bin, float value) {
    atomicAdd(lc.sum + size_t(cell) * lc.bins + bin, (unsigned long long)(value * kLcFixed));
    atomicAdd(lc.hits + cell, 1u);
}

TR_D int lc_cell(const DevLightCache& lc, f3 x) {
    const int g = lc.grid;
    const int ix = min(max(int((x.x - lc.lo.x) * lc.scale.x), 0), g - 1);
    const int iy = min(max(int((x.y - lc.lo.y) * lc.scale.y), 0), g - 1);
    const int iz = min(max(int((x.z - lc.lo.z) * lc.scale.z), 0), g - 1);
    return (iz * g + iy) * g + ix;
}

TR_D int lc_env_bin(const DevLightCache& lc) { return lc.num_area * lc.strata * lc.strata; }

// Stratum index (i + S j) of the point with area parameters (u, v).
TR_D int lc_stratum(const DevLightCache& lc, float u, float v) {
    const int S = lc.strata;
    const int i = min(int(u * float(S)), S - 1), j = min(int(v * float(S)), S - 1);
    return j * S + i;
}

// Bin of a point on area light `light` hit with barycentric weights b1, b2 of
// its second and third vertex: u = (1 - bx)^2 = (b1 + b2)^2, v = by / sqrt(u).
TR_D int lc_bin_of_hit(const DevLightCache& lc, int light, float b1, float b2) {
    const float su = b1 + b2;
    const float u = su * su;
    const float v = su > 0.0f ? b1 / su : 0.0f;
    return light * lc.strata * lc.strata + lc_stratum(lc, u, v);
}

TR_D float lc_pmf(const DevLightCache& lc, int cell, int bin) { return lc.pmf[size_t(cell) * lc.bins + bin]; }

// Draws a bin of cell with uniform u by the alias table; pmf is its probability.
TR_D int lc_sample_bin(const DevLightCache& lc, int cell, float u, float& pmf) {
    const float x = u * float(lc.bins);
    const int k = min(int(x), lc.bins - 1);
    const size_t base = size_t(cell) * lc.bins;
    const float2 t = lc.table[base + k];
    const int bin = (x - float(k)) < t.x ? k : __float_as_int(t.y);
    pmf = lc.pmf[base + bin];
    return bin;
}

// The light entry of a bin (area light index, num_area = environment,
// num_area + 1 = sun).  For an area light, (lu, lv) are moved into the bin's
// stratum and pmf becomes a density per unit of the light's area parameters
// (times S^2).
TR_D int lc_bin_light(const DevLightCache& lc, int bin, float& lu, float& lv, float& pmf) {
    const int S = lc.strata, S2 = S * S;
    const int env = lc.num_area * S2;
    if (bin >= env) return lc.num_area + (bin - env);
    const int light = bin / S2, st = bin - light * S2;
    lu = (float(st % S) + lu) / float(S);
    lv = (float(st / S) + lv) / float(S);
    pmf *= float(S2);
    return light;
}

// Bin of a light sample drawn without the cache (training passes): area light
// e at area parameters (lu, lv), or the environment / sun entry.
TR_D int lc_bin_of_sample(const DevLightCache& lc, int e, float lu, float lv) {
    if (e >= lc.num_area) return lc_env_bin(lc) + (e - lc.num_area);
    return e * lc.strata * lc.strata + lc_stratum(lc, lu, lv);
}

}  // namespace trhip
