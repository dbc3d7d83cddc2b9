// Bidirectional path tracing on the device: subpath vertices, the pass
// state and the launchers of k_bdpt.hip / k_trace.hip.
//
// Replaces BDPTRenderer::render (src/renderer/bdpt_renderer.cpp:1087-1185)
// with its helpers bdpt::generate_camera_subpath (:289-384),
// generate_light_subpath (:386-500), the four connection cases (:502-737),
// compute_mis_weight (:741-762) and the mis_weight_* reassignments
// (:764-956).  One pass holds one camera sample per slot for a whole number of
// frames; its camera and light subpaths are traced together as a wavefront,
// then every (slot, strategy) pair is evaluated by one thread, the
// connections that need a visibility test are traced as one compacted any-hit
// queue, and each slot's camera-side radiance is summed in strategy order.
#pragma once
#include "tr_kernels.h"

namespace trhip {

enum : int { BD_CAMERA = 0, BD_AREA = 1, BD_ENV = 2, BD_SURFACE = 3, BD_MEDIUM = 4 };
constexpr int kBdDelta = 8;   // flag bit of the type word: a specular (Dirac) scattering vertex

// One subpath vertex (bdpt::Vertex, bdpt_renderer.cpp:31-117), 80 bytes.
//   p:    position, pdf_fwd (area density of reaching it from the camera side)
//   n:    surface: geometric normal; camera: view direction; area light: the
//         emitter's normal; environment: direction from the light into the
//         scene (light_to_ref); pdf_bwd (area density from the light side)
//   beta: accumulated weight (accu_coef), type word (type | kBdDelta) as int bits
//   wo:   surface / medium: direction towards the previous vertex of its
//         subpath; w: surface: leaf-ordered triangle; area light: light index;
//         medium: medium index (int bits)
//   x:    surface: barycentrics b1, b2 of the hit
struct BdVertex {
    float4 p, n, beta, wo, x;
};

// Fixed-point scale of the light-tracing splats (t = 1 strategies): integer
// sums do not depend on the order of the atomic adds.
constexpr double kBdSplatScale = 268435456.0;   // 2^28
constexpr float kBdSplatMax = 1.0e9f;           // larger splat values are clamped

struct BdPaths {
    BdVertex* cv;             // camera subpaths [slot][nc]
    BdVertex* lv;             // light subpaths [slot][nl]
    int* ncv;                 // vertices of each camera subpath
    int* nlv;                 // vertices of each light subpath
    float4* gray;             // subpath rays [2 * slots][2]: (origin, tmin), (direction, tmax);
                              // ray r < slots extends camera subpath r, r >= slots light subpath r - slots
    float4* ghit;             // (t, triangle, b1, b2) per subpath ray
    float4* gbeta;            // (weight rgb, solid-angle pdf of the ray's direction)
    uint2* grng;              // PCG32 state per subpath
    int* gq[2];               // subpath rays of the current / next step
    unsigned int* counters;   // [0], [1] the two queues, [2] the shadow queue (kCounterStride apart)
    float4* con;              // [slot][K]: (MIS-weighted unoccluded contribution rgb, splat pixel or -1)
    float4* sray;             // compacted connection rays [cap][2]
    int* sidx;                // the con entry of each connection ray
    unsigned long long* splat;   // [pixel][3] light-tracing sums, kBdSplatScale fixed point
    float4* rad;              // [slot] camera-side radiance (input of k_accumulate)
    unsigned long long* stats;   // [0] subpath rays, [1] connection rays, [2] non-finite samples,
                                 // [9] non-finite splats
    int* ovf_node;            // traversal stack spill area
    float* ovf_t;
    int ovf_threads;
};

struct BdConsts {
    int width, height, npix;
    int slots;                // camera samples in the pass (spp_in_pass whole frames)
    int nc, nl;               // max_camera_vertex_count, max_light_vertex_count
    int K;                    // strategies per slot: nc * (nl + 1), index (t - 1) * (nl + 1) + s
    int sample_base;
    int only_s, only_t;       // >= 0 / >= 1: keep only strategies with this s / t, unweighted
    unsigned long long seed;
    float film_area, lens_area;   // focal-plane film area, lens area (1 for a pinhole)
    float env_pdf_pos;        // 1 / (pi R^2): density of an environment light ray's origin
};

TR_D unsigned int* bd_counter(const BdPaths& b, int k) { return b.counters + k * kCounterStride; }

// k_bdpt.hip
void launch_bd_start(const DevScene& s, const BdPaths& b, const BdConsts& k, hipStream_t st);
void launch_bd_extend(const DevScene& s, const BdPaths& b, const BdConsts& k, int cur, hipStream_t st);
void launch_bd_connect(const DevScene& s, const BdPaths& b, const BdConsts& k, hipStream_t st);
void launch_bd_gather(const BdPaths& b, const BdConsts& k, hipStream_t st);
// film[p].xyz += splat[p] / 2^28 * (film[p].w - w0[p]) / spp: the light-tracing
// image added to the resolved pixel with weight 1 / spp (bdpt_renderer.cpp:1175-1182);
// then the splat sums are cleared.
void launch_bd_fold(const BdPaths& b, double4* film, const double* w0, int npix, int spp, hipStream_t st);
void launch_bd_snapshot(const double4* film, double* w0, int npix, hipStream_t st);
// k_trace.hip: closest hits of the subpath rays in queue cur (clears queue 1 - cur's counter);
// any-hit tests of the connection rays (occluded entries of con are zeroed).
void launch_bd_trace(const TraceScene& sc, const BdPaths& b, int cur, int max_rays, hipStream_t st);
void launch_bd_occlude(const TraceScene& sc, const BdPaths& b, int64_t max_rays, hipStream_t st);

}  // namespace trhip
