// Traversal kernels.  Built with -ffp-contract=off: the watertight triangle
// test must round exactly like the CPU restatement it is checked against.
#include <algorithm>
#include <atomic>
#include <mutex>
#include <stdexcept>
#include <string>

#include "tr_bdpt.h"
#include "tr_kernels.h"

namespace trhip {

namespace {

enum TraceKernel { TK_CLOSEST, TK_ANY, TK_COUNTED, TK_EXTEND, TK_EXTEND_COUNT, TK_SHADOW, TK_SHADOW_MULTI, TK_BD_TRACE,
                   TK_BD_OCCLUDE, TK_NUM };

// The trace kernels loop over their rays with a grid stride, so every wave
// gets an equal share.  A grid that is not a whole number of the kernel's
// resident waves leaves its last round part-empty: on configuration 2
// (64 spp per pass) a 2048-block k_extend (1.6 rounds at 5 waves per SIMD)
// took 48.5 ms per frame against 36.6 ms at three full rounds.
constexpr int kTraceRounds = 3;
constexpr int kMaxDevices = 64;
std::atomic<int> g_grid[kMaxDevices][TK_NUM];

template <typename K>
int resident_rounds(K kernel) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kTraceBlock, 0) != hipSuccess)
        throw std::runtime_error("trace kernel occupancy query failed");
    return std::max(1, per_cu) * std::max(1, cus) * kTraceRounds;
}

}  // namespace

__global__ __launch_bounds__(kTraceBlock) void k_trace_closest(TraceScene sc, const float4* __restrict__ rays,
                                                                float4* __restrict__ hits, int64_t n, int* ovf_node,
                                                                float* ovf_t) {
    __shared__ TraceStack<kTraceBlock, kLdsStack> st;
    const int tid = threadIdx.x;
    const int gt = blockIdx.x * kTraceBlock + tid;
    const int stride = gridDim.x * kTraceBlock;
    for (int64_t i = gt; i < n; i += stride) {
        const float4 a = rays[2 * i], b = rays[2 * i + 1];
        TraceHit h;
        trace<false, false, kTraceBlock, kLdsStack>(sc, xyz(a), a.w, xyz(b), b.w, st, tid, ovf_node + gt, ovf_t + gt,
                                                    stride, h);
        hits[i] = make_float4(h.t, __int_as_float(h.prim), h.b1, h.b2);
    }
}

__global__ __launch_bounds__(kTraceBlock) void k_trace_any(TraceScene sc, const float4* __restrict__ rays,
                                                            int* __restrict__ occ, int64_t n, int* ovf_node, float* ovf_t) {
    __shared__ TraceStack<kTraceBlock, kLdsStack> st;
    const int tid = threadIdx.x;
    const int gt = blockIdx.x * kTraceBlock + tid;
    const int stride = gridDim.x * kTraceBlock;
    for (int64_t i = gt; i < n; i += stride) {
        const float4 a = rays[2 * i], b = rays[2 * i + 1];
        TraceHit h;
        occ[i] = trace<true, false, kTraceBlock, kLdsStack>(sc, xyz(a), a.w, xyz(b), b.w, st, tid, ovf_node + gt,
                                                            ovf_t + gt, stride, h)
                     ? 1
                     : 0;
    }
}

__global__ __launch_bounds__(kTraceBlock) void k_trace_counted(TraceScene sc, const float4* __restrict__ rays,
                                                                float4* __restrict__ hits, uint32_t* nvis, uint32_t* ntri,
                                                                int64_t n, int* ovf_node, float* ovf_t) {
    __shared__ TraceStack<kTraceBlock, kLdsStack> st;
    const int tid = threadIdx.x;
    const int gt = blockIdx.x * kTraceBlock + tid;
    const int stride = gridDim.x * kTraceBlock;
    for (int64_t i = gt; i < n; i += stride) {
        const float4 a = rays[2 * i], b = rays[2 * i + 1];
        TraceHit h;
        uint32_t cv = 0, ct = 0;
        trace<false, true, kTraceBlock, kLdsStack>(sc, xyz(a), a.w, xyz(b), b.w, st, tid, ovf_node + gt, ovf_t + gt,
                                                   stride, h, &cv, &ct);
        hits[i] = make_float4(h.t, __int_as_float(h.prim), h.b1, h.b2);
        nvis[i] = cv;
        ntri[i] = ct;
    }
}

// Work distribution of the wavefront trace kernels.  Each wave owns a
// contiguous share of the queue (neighbouring entries: nearby pixels) and its
// lanes traverse their own rays.  When fewer than kRefillLanes of the 64 lanes
// are still traversing, the idle lanes take the next entries of the share.
// Measured on configuration 2 (per frame): waiting for every lane (one ray
// per lane at a time) 25.6 ms of k_extend and 8.9 ms of k_shadow; refilling
// below 48 busy lanes 22.2 and 7.7 ms; below 32, 22.1 and 8.0 ms; refilling
// any idle lane at once, 24.6 and 7.8 ms (the wave's rays drift apart).
// Re-swept with leaf postponement and class-ordered shading (40, 48 or 56
// lanes, separately for closest-hit and any-hit rays): configuration 2 and 5
// frames all within 1% of each other.
// Fetching single rays from a global work counter was much slower (the
// atomics contend), and so was dealing the queue out in chunks of 256 entries
// from a device counter once a wave's chunk ran out (configuration 2:
// k_extend 15.4 -> 39.2 ms, k_shadow 5.7 -> 27.2 ms per frame).
constexpr int kRefillLanes = 48;

struct WaveQueue {
    unsigned int next, end;   // wave-uniform

    TR_D explicit WaveQueue(unsigned int n) {
        const unsigned int waves = gridDim.x * (kTraceBlock / 64);
        const unsigned int w = blockIdx.x * (kTraceBlock / 64) + (threadIdx.x >> 6);
        const unsigned int per = (n + waves - 1) / waves;
        next = __builtin_amdgcn_readfirstlane(min(n, w * per));
        end = __builtin_amdgcn_readfirstlane(min(n, next + per));
    }

    // Called by every lane of the wave: idle lanes that receive a queue entry
    // call fetch(entry) and become busy.
    template <typename F>
    TR_D void refill(bool& busy, F&& fetch) {
        if (__popcll(__ballot(busy)) < kRefillLanes && next < end) {
            const unsigned long long idle = __ballot(!busy);
            const unsigned int idx = next + (unsigned int)__popcll(idle & ((1ull << __lane_id()) - 1ull));
            if (!busy && idx < end) {
                fetch(idx);
                busy = true;
            }
            next = __builtin_amdgcn_readfirstlane(min(end, next + (unsigned int)__popcll(idle)));
        }
    }
};

// Wavefront extension: closest hit for every path slot in the current queue.
// COUNT: also accumulate node visits and triangle tests into stats[3..4].
template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void k_extend(TraceScene sc, DevPaths p, int cur) {
    __shared__ TraceStack<kTraceBlock, kLdsStack> st;
    const int tid = threadIdx.x;
    const int gt = blockIdx.x * kTraceBlock + tid;
    const int stride = gridDim.x * kTraceBlock;
    const unsigned int n = *queue_counter(p, cur);
    if (gt == 0) {
        *queue_counter(p, 1 - cur) = 0u;   // next queue and shadow queue are filled by the shade kernel
        *queue_counter(p, kShadowCounter) = 0u;
        atomicAdd(&p.stats[0], (unsigned long long)n);
    }
    const int* q = p.queue[cur];
    uint32_t cv = 0, ct = 0;
    WaveQueue wq(n);
    TraceRay r;
    bool busy = false;
    int slot = 0;
    while (true) {
        wq.refill(busy, [&](unsigned int idx) {
            slot = q[idx];
            const float4 a = p.ray_o[slot], b = p.ray_d[slot];
            r.init(xyz(a), a.w, xyz(b), b.w);
        });
        if (__ballot(busy) == 0ull) break;
        if (busy && !r.step<false, COUNT, kTraceBlock, kLdsStack>(sc, st, tid, p.ovf_node + gt, p.ovf_t + gt, stride,
                                                                  &cv, &ct)) {
            TraceHit h;
            r.result(h);
            p.hit[slot] = make_float4(h.t, __int_as_float(h.tri), h.b1, h.b2);
            busy = false;
        }
    }
    if (COUNT) {
        atomicAdd(&p.stats[3], (unsigned long long)cv);
        atomicAdd(&p.stats[4], (unsigned long long)ct);
    }
}

// Wavefront shadow rays: any hit; unoccluded contributions are added to the
// radiance of their path.  record != 0: unoccluded rays validate their
// light-sample training record.
__global__ __launch_bounds__(kTraceBlock) void k_shadow(TraceScene sc, DevPaths p, int atomic_add, int record) {
    __shared__ TraceStack<kTraceBlock, kLdsStack> st;
    const int tid = threadIdx.x;
    const int gt = blockIdx.x * kTraceBlock + tid;
    const int stride = gridDim.x * kTraceBlock;
    const unsigned int n = *queue_counter(p, kShadowCounter);
    if (gt == 0) atomicAdd(&p.stats[1], (unsigned long long)n);
    WaveQueue wq(n);
    TraceRay r;
    bool busy = false;
    unsigned int i = 0;
    while (true) {
        wq.refill(busy, [&](unsigned int idx) {
            i = idx;
            const float4 a = p.sh_o[i], b = p.sh_d[i];
            r.init(xyz(a), a.w, xyz(b), b.w);
        });
        if (__ballot(busy) == 0ull) break;
        if (!busy || r.step<true, false, kTraceBlock, kLdsStack>(sc, st, tid, p.ovf_node + gt, p.ovf_t + gt, stride,
                                                                 nullptr, nullptr))
            continue;
        busy = false;
        TraceHit h;
        if (!r.result(h)) {
            if (record) {
                const int ri = p.sh_rec[i];
                if (ri >= 0) p.rec_aux[2 * ri + 1] = make_float4(0.0f, 0.0f, 0.0f, 2.0f);
            }
            if (p.lc.train) {
                // light-cache training: the unoccluded sample's contribution / density
                const float4 g = p.sh_g[i];
                const int cell = __float_as_int(g.x);
                if (cell >= 0 && g.z > 0.0f) lc_record(p.lc, cell, __float_as_int(g.y), g.z);
            }
            const float4 c = p.sh_c[i];
            const int slot = __float_as_int(c.w);
            if (atomic_add) {
                atomicAdd(&p.rad[slot].x, c.x);
                atomicAdd(&p.rad[slot].y, c.y);
                atomicAdd(&p.rad[slot].z, c.z);
            } else {
                float4 rad = p.rad[slot];
                rad.x += c.x; rad.y += c.y; rad.z += c.z;
                p.rad[slot] = rad;
            }
        }
    }
}

// Shadow rays of vertices with several light samples (direct_light_sample_num
// > 1): sample k of the path in queue entry i sits at slot * nls + k; unused
// samples have tmax < 0.  An unoccluded sample sets w of its contribution to 1;
// k_gather_multi then adds the path's samples in order, so the film stays
// deterministic.
__global__ __launch_bounds__(kTraceBlock) void k_shadow_multi(TraceScene sc, DevPaths p, int cur, int nls) {
    __shared__ TraceStack<kTraceBlock, kLdsStack> st;
    const int tid = threadIdx.x;
    const int gt = blockIdx.x * kTraceBlock + tid;
    const int stride = gridDim.x * kTraceBlock;
    const unsigned int n = *queue_counter(p, cur) * (unsigned int)nls;
    const int* q = p.queue[cur];
    WaveQueue wq(n);
    TraceRay r;
    bool busy = false, valid = false;
    size_t e = 0;
    while (true) {
        wq.refill(busy, [&](unsigned int idx) {
            e = size_t(q[idx / (unsigned int)nls]) * nls + idx % (unsigned int)nls;
            const float4 a = p.sh_o[e], b = p.sh_d[e];
            valid = b.w >= 0.0f;
            r.init(xyz(a), a.w, xyz(b), valid ? b.w : -1.0f);
        });
        if (__ballot(busy) == 0ull) break;
        if (!busy || r.step<true, false, kTraceBlock, kLdsStack>(sc, st, tid, p.ovf_node + gt, p.ovf_t + gt, stride,
                                                                 nullptr, nullptr))
            continue;
        busy = false;
        TraceHit h;
        if (valid && !r.result(h)) p.sh_c[e].w = 1.0f;
    }
}

__global__ __launch_bounds__(256) void k_gather_multi(DevPaths p, int cur, int nls) {
    const unsigned int n = *queue_counter(p, cur);
    const int* q = p.queue[cur];
    unsigned long long rays = 0;
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int slot = q[i];
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        for (int k = 0; k < nls; ++k) {
            const size_t e = size_t(slot) * nls + k;
            if (!(p.sh_d[e].w >= 0.0f)) continue;
            ++rays;
            const float4 c = p.sh_c[e];
            if (c.w > 0.5f) { ax += c.x; ay += c.y; az += c.z; }
        }
        float4 rad = p.rad[slot];
        rad.x += ax; rad.y += ay; rad.z += az;
        p.rad[slot] = rad;
    }
    for (int off = 32; off > 0; off >>= 1) rays += __shfl_xor(rays, off);
    if (__lane_id() == 0 && rays) atomicAdd(&p.stats[1], rays);
}

// Bidirectional path tracing: closest hits of the subpath rays in queue cur.
__global__ __launch_bounds__(kTraceBlock) void k_bd_trace(TraceScene sc, BdPaths b, int cur) {
    __shared__ TraceStack<kTraceBlock, kLdsStack> st;
    const int tid = threadIdx.x;
    const int gt = blockIdx.x * kTraceBlock + tid;
    const int stride = gridDim.x * kTraceBlock;
    const unsigned int n = *bd_counter(b, cur);
    if (gt == 0) {
        *bd_counter(b, 1 - cur) = 0u;   // filled by k_bd_extend
        atomicAdd(&b.stats[0], (unsigned long long)n);
    }
    const int* q = b.gq[cur];
    WaveQueue wq(n);
    TraceRay r;
    bool busy = false;
    int ray = 0;
    while (true) {
        wq.refill(busy, [&](unsigned int idx) {
            ray = q[idx];
            const float4 a = b.gray[2 * size_t(ray)], d = b.gray[2 * size_t(ray) + 1];
            r.init(xyz(a), a.w, xyz(d), d.w);
        });
        if (__ballot(busy) == 0ull) break;
        if (busy && !r.step<false, false, kTraceBlock, kLdsStack>(sc, st, tid, b.ovf_node + gt, b.ovf_t + gt, stride,
                                                                  nullptr, nullptr)) {
            TraceHit h;
            r.result(h);
            b.ghit[ray] = make_float4(h.t, __int_as_float(h.tri), h.b1, h.b2);
            busy = false;
        }
    }
}

// Bidirectional path tracing: visibility of the compacted connection rays;
// an occluded connection contributes nothing.
__global__ __launch_bounds__(kTraceBlock) void k_bd_occlude(TraceScene sc, BdPaths b) {
    __shared__ TraceStack<kTraceBlock, kLdsStack> st;
    const int tid = threadIdx.x;
    const int gt = blockIdx.x * kTraceBlock + tid;
    const int stride = gridDim.x * kTraceBlock;
    const unsigned int n = *bd_counter(b, 2);
    if (gt == 0) atomicAdd(&b.stats[1], (unsigned long long)n);
    WaveQueue wq(n);
    TraceRay r;
    bool busy = false;
    unsigned int i = 0;
    while (true) {
        wq.refill(busy, [&](unsigned int idx) {
            i = idx;
            const float4 a = b.sray[2 * size_t(i)], d = b.sray[2 * size_t(i) + 1];
            r.init(xyz(a), a.w, xyz(d), d.w);
        });
        if (__ballot(busy) == 0ull) break;
        if (!busy || r.step<true, false, kTraceBlock, kLdsStack>(sc, st, tid, b.ovf_node + gt, b.ovf_t + gt, stride,
                                                                 nullptr, nullptr))
            continue;
        busy = false;
        TraceHit h;
        if (r.result(h)) {
            float4* c = b.con + b.sidx[i];
            c->x = 0.0f;
            c->y = 0.0f;
            c->z = 0.0f;
        }
    }
}

namespace {

// Blocks of a full-size launch of kernel k on the current device.
int grid_blocks(TraceKernel k) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) throw std::runtime_error("no current device");
    int b = g_grid[dev][k].load(std::memory_order_relaxed);
    if (b > 0) return b;
    switch (k) {
        case TK_CLOSEST: b = resident_rounds(k_trace_closest); break;
        case TK_ANY: b = resident_rounds(k_trace_any); break;
        case TK_COUNTED: b = resident_rounds(k_trace_counted); break;
        case TK_EXTEND: b = resident_rounds(k_extend<false>); break;
        case TK_EXTEND_COUNT: b = resident_rounds(k_extend<true>); break;
        case TK_SHADOW_MULTI: b = resident_rounds(k_shadow_multi); break;
        case TK_BD_TRACE: b = resident_rounds(k_bd_trace); break;
        case TK_BD_OCCLUDE: b = resident_rounds(k_bd_occlude); break;
        default: b = resident_rounds(k_shadow); break;
    }
    g_grid[dev][k].store(b, std::memory_order_relaxed);
    return b;
}

// Blocks for n rays, at most the full-size grid and at most what the spill
// area (one column per thread) covers.
int trace_blocks(TraceKernel k, int64_t n, int ovf_threads) {
    int64_t blocks = (n + kTraceBlock - 1) / kTraceBlock;
    blocks = std::min<int64_t>(blocks, grid_blocks(k));
    blocks = std::min<int64_t>(blocks, ovf_threads / kTraceBlock);
    if (blocks < 1) throw std::runtime_error("traversal spill area smaller than one block");
    return int(blocks);
}

}  // namespace

int trace_grid_threads() {
    int b = 0;
    for (int k = 0; k < TK_NUM; ++k) b = std::max(b, grid_blocks(TraceKernel(k)));
    return b * kTraceBlock;
}

// ---- allocation-free closest-hit queries ---------------------------------
// tr_trace_closest_device_noalloc traverses with its spill area in this
// file-scope array instead of the context's allocation: kNoAllocEntries stack
// entries (node id, then entry distance as float bits) for each of
// kNoAllocThreads threads, 12.6 MB per device, held until the process ends.
// Every call on a device shares that device's copy: calls that overlap on one
// GPU race.  The array's name is only a host-side stand-in; its device
// address is looked up once per device (noalloc_scratch_init) and used from
// then on.
constexpr int kNoAllocBlocks = 128;
constexpr int kNoAllocThreads = kNoAllocBlocks * kTraceBlock;
constexpr int kNoAllocEntries = 48;
static __device__ __align__(256) int scratch[2 * kNoAllocEntries * kNoAllocThreads];

namespace {
std::mutex g_scratch_mtx;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch_init() {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mtx);
    if (g_scratch_ptr[dev]) return hipSuccess;
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = ptr;
    return hipSuccess;
}

int noalloc_max_stack() { return kNoAllocEntries + kLdsStack; }

void launch_trace_closest_noalloc(const TraceScene& sc, const float4* rays, float4* hits, int64_t n, hipStream_t st) {
    if (n <= 0) return;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) throw std::runtime_error("no current device");
    void* ptr = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mtx);
        ptr = g_scratch_ptr[dev];
    }
    if (!ptr) throw std::runtime_error("the allocation-free scratch of this device is not set up");
    if (sc.max_stack - kLdsStack > kNoAllocEntries)
        throw std::invalid_argument("the BVH is too deep for the allocation-free spill area (at most " +
                                    std::to_string(kNoAllocEntries + kLdsStack) + " stack entries)");
    const int64_t want = (n + kTraceBlock - 1) / kTraceBlock;
    const int blocks = int(std::min<int64_t>(want, kNoAllocBlocks));
    int* node = static_cast<int*>(ptr);
    float* t = reinterpret_cast<float*>(node + size_t(kNoAllocEntries) * kNoAllocThreads);
    hipLaunchKernelGGL(k_trace_closest, dim3(blocks), dim3(kTraceBlock), 0, st, sc, rays, hits, n, node, t);
}

void launch_trace_closest(const TraceScene& sc, const float4* rays, float4* hits, int64_t n, int* ovf_node, float* ovf_t,
                          int ovf_threads, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_trace_closest, dim3(trace_blocks(TK_CLOSEST, n, ovf_threads)), dim3(kTraceBlock), 0, st, sc, rays, hits, n, ovf_node, ovf_t);
}
void launch_trace_any(const TraceScene& sc, const float4* rays, int* occ, int64_t n, int* ovf_node, float* ovf_t,
                      int ovf_threads, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_trace_any, dim3(trace_blocks(TK_ANY, n, ovf_threads)), dim3(kTraceBlock), 0, st, sc, rays, occ, n, ovf_node, ovf_t);
}
void launch_trace_counted(const TraceScene& sc, const float4* rays, float4* hits, uint32_t* nvis, uint32_t* ntri, int64_t n,
                          int* ovf_node, float* ovf_t, int ovf_threads, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_trace_counted, dim3(trace_blocks(TK_COUNTED, n, ovf_threads)), dim3(kTraceBlock), 0, st, sc, rays, hits, nvis, ntri, n,
                       ovf_node, ovf_t);
}
void launch_extend(const TraceScene& sc, const DevPaths& p, int cur, int max_paths, bool count, hipStream_t st) {
    if (count)
        hipLaunchKernelGGL(k_extend<true>, dim3(trace_blocks(TK_EXTEND_COUNT, max_paths, p.ovf_threads)), dim3(kTraceBlock), 0, st, sc, p, cur);
    else
        hipLaunchKernelGGL(k_extend<false>, dim3(trace_blocks(TK_EXTEND, max_paths, p.ovf_threads)), dim3(kTraceBlock), 0, st, sc, p, cur);
}
void launch_shadow_multi(const TraceScene& sc, const DevPaths& p, int cur, int max_paths, int nls, hipStream_t st) {
    const int64_t n = int64_t(max_paths) * nls;
    hipLaunchKernelGGL(k_shadow_multi, dim3(trace_blocks(TK_SHADOW_MULTI, n, p.ovf_threads)), dim3(kTraceBlock), 0, st, sc, p,
                       cur, nls);
    const int blocks = int(std::min<int64_t>((max_paths + 255) / 256, 4096));
    hipLaunchKernelGGL(k_gather_multi, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, st, p, cur, nls);
}
void launch_bd_trace(const TraceScene& sc, const BdPaths& b, int cur, int max_rays, hipStream_t st) {
    hipLaunchKernelGGL(k_bd_trace, dim3(trace_blocks(TK_BD_TRACE, max_rays, b.ovf_threads)), dim3(kTraceBlock), 0, st, sc, b,
                       cur);
}
void launch_bd_occlude(const TraceScene& sc, const BdPaths& b, int64_t max_rays, hipStream_t st) {
    hipLaunchKernelGGL(k_bd_occlude, dim3(trace_blocks(TK_BD_OCCLUDE, max_rays, b.ovf_threads)), dim3(kTraceBlock), 0, st,
                       sc, b);
}
void launch_shadow(const TraceScene& sc, const DevPaths& p, int max_shadow, int accumulate_atomic, int record,
                   hipStream_t st) {
    hipLaunchKernelGGL(k_shadow, dim3(trace_blocks(TK_SHADOW, max_shadow, p.ovf_threads)), dim3(kTraceBlock), 0, st, sc, p, accumulate_atomic,
                       record);
}

}  // namespace trhip
