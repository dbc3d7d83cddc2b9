// Host-side launchers for the kernels in k_trace.hip and k_render.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "tr_lightcache.h"
#include "tr_scene.h"

namespace trhip {

constexpr int kTraceBlock = 256;
constexpr int kLdsStack = 8;
constexpr int kShadeBlock = 256;
// Specular bounces that do not count towards the depth (path_trace_renderer.cpp:20, :133-138).
constexpr int kMaxSpecularDepth = 20;

// Path state of the wavefront (one slot per camera sample in flight).
struct DevPaths {
    float4* ray_o;    // (origin, tmin)
    float4* ray_d;    // (direction, tmax)
    float4* hit;      // (t, leaf-ordered triangle, b1, b2)
    float4* thr;      // (throughput rgb, pdf of the sampled direction; 0 = camera ray)
    float4* rad;      // (radiance rgb, cell of the vertex that sampled the ray as int bits: -1 = none)
    uint2* rng;       // PCG32 state
    uint2* info;      // (pixel, depth | scatter count << 8)
    int* queue[2];    // active path slots for the current / next bounce
    float4* sh_o;     // shadow rays: (origin, tmin)
    float4* sh_d;     //              (direction, tmax)
    float4* sh_c;     //              (contribution rgb, path slot)
    unsigned int* counters;   // queue lengths 0, 1 and the shadow queue length 2, kCounterStride apart
    unsigned long long* stats;  // [0] extension rays, [1] shadow rays, [2] non-finite samples,
                                // [3] node visits, [4] triangle tests
    int* ovf_node;    // traversal stack spill area
    float* ovf_t;
    int ovf_threads;  // threads the spill area is sized for
    // path guiding (null without the guide); stats[5] vertices that fell back
    // to BSDF sampling, stats[6] vertices sampled with the guide mixture,
    // stats[7] network evaluations
    const float* graw;        // [slot][48] head outputs of the guide MLP
    int* grec;                // training record waiting for the radiance of its continuation, -1 = none
    float4* rec;              // training records [cap][4]: (x, n.x), (n.yz, wo.xy), (wo.z, wi), (-, slot, pdf, -)
    float4* rec_aux;          // [cap][2]: continuation records (rr / throughput before the vertex, -),
                              //   (radiance before the next vertex, 1); light-sample records
                              //   (weight, -, -, -), (-, -, -, 2 once the shadow ray found the light)
    unsigned int* rec_count;
    int rec_cap;
    int* sh_rec;              // light-sample record of each shadow ray, -1 = none
    // learned light-sample distribution (guided passes; lc.grid == 0: off)
    DevLightCache lc;
    float4* sh_g;             // training passes: (cell, bin as int bits, contribution / density, -) per shadow ray
};

// The queue counters sit 1 KB apart: appends to different queues then go to
// different cache lines and do not serialise behind each other.
constexpr int kCounterStride = 256;
constexpr int kNumCounters = 4;
constexpr int kShadowCounter = 2;
constexpr int kTailBlocksCounter = 3;   // workgroups of the current k_tail launch that have finished
TR_D unsigned int* queue_counter(const DevPaths& p, int k) { return p.counters + k * kCounterStride; }

// Block-aggregated append to two counters: one atomic per counter and block
// (the counters are device-wide and every atomic on them serialises, so a
// wave-level append made the shade kernel wait on them).  Every thread of the
// block must call it: it synchronises the block.  Slots are consecutive in
// thread order.  c0 and c1 may be the same counter.
template <int BLOCK>
struct BlockAppendLds {
    unsigned int cnt[2][BLOCK / 64];
    unsigned int base[2];
};

template <int BLOCK>
TR_D void block_append2(BlockAppendLds<BLOCK>& sh, unsigned int* c0, bool p0, unsigned int* c1, bool p1,
                        unsigned int& i0, unsigned int& i1) {
    const int lane = __lane_id(), wave = int(threadIdx.x) >> 6;
    const unsigned long long m0 = __ballot(p0), m1 = __ballot(p1);
    if (lane == 0) {
        sh.cnt[0][wave] = (unsigned int)__popcll(m0);
        sh.cnt[1][wave] = (unsigned int)__popcll(m1);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        unsigned int tot = 0;
        for (int w = 0; w < BLOCK / 64; ++w) tot += sh.cnt[threadIdx.x][w];
        sh.base[threadIdx.x] = tot ? atomicAdd(threadIdx.x == 0 ? c0 : c1, tot) : 0u;
    }
    __syncthreads();
    unsigned int b0 = sh.base[0], b1 = sh.base[1];
    for (int w = 0; w < wave; ++w) {
        b0 += sh.cnt[0][w];
        b1 += sh.cnt[1][w];
    }
    const unsigned long long below = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
    i0 = b0 + (unsigned int)__popcll(m0 & below);
    i1 = b1 + (unsigned int)__popcll(m1 & below);
    __syncthreads();   // sh is reused by the next call
}

struct RenderConsts {
    int width, height;
    int min_depth, max_depth;
    int spp_in_pass;     // samples per pixel in this pass
    int sample_base;     // sample index of the pass's first sample
    int filter;          // TR_FILTER_*
    float filter_radius, filter_alpha;
    int shadow_tr;       // TR_TR_*
    int nls;             // light samples per vertex
    int shade_sort_rounds;   // shade kernels order a window by class when it spans at least this many rounds
    int first_bounce;        // the queue holds the camera paths of k_generate, whose throughput, radiance
                             // and depth are implicit (1, 0, 0) rather than stored
    unsigned long long seed;
    float guide_frac;    // probability of sampling the guide mixture at a surface vertex
    int guide_max_depth; // guide only vertices of depth < this (0 = every depth)
    int guide_record;    // collect training records in this pass
    unsigned int rec_threshold;   // a path records when hash(slot, rec_seed) < rec_threshold
    unsigned int rec_seed;
};

// k_trace.hip
void launch_trace_closest(const TraceScene& sc, const float4* rays, float4* hits, int64_t n, int* ovf_node,
                          float* ovf_t, int ovf_threads, hipStream_t st);
void launch_trace_any(const TraceScene& sc, const float4* rays, int* occ, int64_t n, int* ovf_node, float* ovf_t,
                      int ovf_threads, hipStream_t st);
void launch_trace_counted(const TraceScene& sc, const float4* rays, float4* hits, uint32_t* nvis, uint32_t* ntri,
                          int64_t n, int* ovf_node, float* ovf_t, int ovf_threads, hipStream_t st);
// Closest hits with the traversal spill area in a file-scope device array
// (no device allocation).  noalloc_scratch_init looks up its address on the
// current device (once per device); noalloc_max_stack is the deepest traversal
// stack it holds.
hipError_t noalloc_scratch_init();
int noalloc_max_stack();
void launch_trace_closest_noalloc(const TraceScene& sc, const float4* rays, float4* hits, int64_t n, hipStream_t st);
// wavefront: extend the current queue, trace the shadow queue
void launch_extend(const TraceScene& sc, const DevPaths& p, int cur, int max_paths, bool count, hipStream_t st);
// record != 0: validate the light-sample training records of unoccluded shadow rays
void launch_shadow(const TraceScene& sc, const DevPaths& p, int max_shadow, int accumulate_atomic, int record,
                   hipStream_t st);
// direct_light_sample_num > 1: the light samples of queue cur's paths
// (nls per path, at slot * nls + k), then their unoccluded contributions added
// to the paths' radiance in sample order.
void launch_shadow_multi(const TraceScene& sc, const DevPaths& p, int cur, int max_paths, int nls, hipStream_t st);
// Threads of the largest trace-kernel grid on the current device: the spill
// area holds one column per thread.
int trace_grid_threads();

// k_render.hip
void launch_generate(const DevScene& s, const DevPaths& p, const RenderConsts& c, int num_paths, int cur, hipStream_t st);
void launch_shade(const DevScene& s, const DevPaths& p, const RenderConsts& c, int cur, int max_paths, bool guide,
                  hipStream_t st);
// Plain passes: when queue cur holds at most max_tail paths, every one is
// carried to its end and the queue emptied; otherwise nothing happens.
void launch_tail(const DevScene& s, const DevPaths& p, const RenderConsts& c, int cur, int max_tail, hipStream_t st);
void launch_accumulate(const DevPaths& p, const RenderConsts& c, double4* film, hipStream_t st);
void launch_resolve(const double4* film, float* image, int num_pixels, hipStream_t st);
void launch_bsdf_eval(const DevMaterial* mats, const int* ids, const float* wi, const float* wo, const float* ng, int64_t n,
                      float* f, float* pdf, hipStream_t st);
void launch_film_ldr(const double4* film, unsigned char* rgb, int num_pixels, float exposure, float inv_gamma,
                     hipStream_t st);
void launch_env_sample(const DevScene& s, const float* u, int64_t n, float* wi, float* pdf, float* L, hipStream_t st);
void launch_env_eval(const DevScene& s, const float* d, int64_t n, float* L, float* pdf, hipStream_t st);
void launch_grid_albedo_g(const DevScene& s, int medium, const float* x, int64_t n, float* albedo, float* g,
                          hipStream_t st);
void launch_grid_density(const DevScene& s, int medium, const float* x, int64_t n, float* out, hipStream_t st);
void launch_tex_eval(const DevScene& s, int texture, const float* uv, int64_t n, float* out, hipStream_t st);
// Light-cache sampling tables (pmf, alias table) of every cell from the training sums.
void launch_lc_build(const DevLightCache& lc, const float* light_pmf, float eps, int cells, float* pmf, float2* table,
                     int* work, hipStream_t st);
void launch_geometric_normals(const float4* tris, float4* tri_shade, int64_t nt, hipStream_t st);
void launch_shading_frames(const DevScene& s, const int* tri, const float* bary, int64_t n, float* frames,
                           hipStream_t st);
void launch_camera_rays(const DevScene& s, const float* film_uv, const float* lens_uv, int64_t n, float4* rays,
                        hipStream_t st);
void launch_light_pick(const DevScene& s, const float* u, int64_t n, int* idx, float* pmf, hipStream_t st);
void launch_hg_sample(float g, const float* wo, const float* u, int64_t n, float* wi, float* pdf, hipStream_t st);
void launch_unpack_radiance(const float* L, int64_t n, float4* rad, hipStream_t st);
void launch_bsdf_sample(const DevMaterial* mats, const int* ids, const float* wo, const float* ng, const float* u,
                        int64_t n, float* wi, float* f, float* pdf, hipStream_t st);

}  // namespace trhip
