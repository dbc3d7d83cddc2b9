// Traversal kernels.  Built with -ffp-contract=off: the watertight triangle
// test must round exactly like the CPU restatement it is checked against.
#include "tr_kernels.h"

namespace trhip {

constexpr int kMaxTraceBlocks = 2048;

int trace_grid_threads(int64_t n) {
    int64_t blocks = (n + kTraceBlock - 1) / kTraceBlock;
    if (blocks < 1) blocks = 1;
    if (blocks > kMaxTraceBlocks) blocks = kMaxTraceBlocks;
    return int(blocks) * kTraceBlock;
}

static int trace_blocks(int64_t n) { return trace_grid_threads(n) / kTraceBlock; }

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

// Wavefront extension: closest hit for every path slot in the current queue,
// one ray per lane at a time in queue order: a wave's lanes traverse rays of
// neighbouring queue entries (nearby pixels) together.  Letting each lane run
// ahead through its own sequence of rays measured slower on the MI355X
// (0.32 vs 0.30 ms per launch on configuration 2), and fetching rays from a
// global work counter much slower (0.49 ms: the atomics contend).
// COUNT: also accumulate node visits and triangle tests into stats[3..4].
template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void k_extend(TraceScene sc, DevPaths p, int cur) {
    __shared__ TraceStack<kTraceBlock, kLdsStack> st;
    const int tid = threadIdx.x;
    const int gt = blockIdx.x * kTraceBlock + tid;
    const int stride = gridDim.x * kTraceBlock;
    const unsigned int n = p.counters[cur];
    if (gt == 0) {
        p.counters[1 - cur] = 0u;   // next queue and shadow queue are filled by the shade kernel
        p.counters[2] = 0u;
        atomicAdd(&p.stats[0], (unsigned long long)n);
    }
    const int* q = p.queue[cur];
    uint32_t cv = 0, ct = 0;
    for (unsigned int i = gt; i < n; i += stride) {
        const int slot = q[i];
        const float4 a = p.ray_o[slot], b = p.ray_d[slot];
        TraceHit h;
        trace<false, COUNT, kTraceBlock, kLdsStack>(sc, xyz(a), a.w, xyz(b), b.w, st, tid, p.ovf_node + gt,
                                                    p.ovf_t + gt, stride, h, &cv, &ct);
        p.hit[slot] = make_float4(h.t, __int_as_float(h.tri), h.b1, h.b2);
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
    const unsigned int n = p.counters[2];
    if (gt == 0) atomicAdd(&p.stats[1], (unsigned long long)n);
    for (unsigned int i = gt; i < n; i += stride) {
        const float4 a = p.sh_o[i], b = p.sh_d[i];
        TraceHit h;
        const bool occ = trace<true, false, kTraceBlock, kLdsStack>(sc, xyz(a), a.w, xyz(b), b.w, st, tid,
                                                                    p.ovf_node + gt, p.ovf_t + gt, stride, h);
        if (!occ) {
            if (record) {
                const int ri = p.sh_rec[i];
                if (ri >= 0) p.rec_aux[2 * ri + 1] = make_float4(0.0f, 0.0f, 0.0f, 2.0f);
            }
            const float4 c = p.sh_c[i];
            const int slot = __float_as_int(c.w);
            if (atomic_add) {
                atomicAdd(&p.rad[slot].x, c.x);
                atomicAdd(&p.rad[slot].y, c.y);
                atomicAdd(&p.rad[slot].z, c.z);
            } else {
                float4 r = p.rad[slot];
                r.x += c.x; r.y += c.y; r.z += c.z;
                p.rad[slot] = r;
            }
        }
    }
}

void launch_trace_closest(const TraceScene& sc, const float4* rays, float4* hits, int64_t n, int* ovf_node, float* ovf_t,
                          int ovf_threads, hipStream_t st) {
    if (n <= 0) return;
    const int blocks = trace_blocks(n);
    (void)ovf_threads;
    hipLaunchKernelGGL(k_trace_closest, dim3(blocks), dim3(kTraceBlock), 0, st, sc, rays, hits, n, ovf_node, ovf_t);
}
void launch_trace_any(const TraceScene& sc, const float4* rays, int* occ, int64_t n, int* ovf_node, float* ovf_t,
                      int ovf_threads, hipStream_t st) {
    if (n <= 0) return;
    (void)ovf_threads;
    hipLaunchKernelGGL(k_trace_any, dim3(trace_blocks(n)), dim3(kTraceBlock), 0, st, sc, rays, occ, n, ovf_node, ovf_t);
}
void launch_trace_counted(const TraceScene& sc, const float4* rays, float4* hits, uint32_t* nvis, uint32_t* ntri, int64_t n,
                          int* ovf_node, float* ovf_t, int ovf_threads, hipStream_t st) {
    if (n <= 0) return;
    (void)ovf_threads;
    hipLaunchKernelGGL(k_trace_counted, dim3(trace_blocks(n)), dim3(kTraceBlock), 0, st, sc, rays, hits, nvis, ntri, n,
                       ovf_node, ovf_t);
}
void launch_extend(const TraceScene& sc, const DevPaths& p, int cur, int max_paths, bool count, hipStream_t st) {
    if (count)
        hipLaunchKernelGGL(k_extend<true>, dim3(trace_blocks(max_paths)), dim3(kTraceBlock), 0, st, sc, p, cur);
    else
        hipLaunchKernelGGL(k_extend<false>, dim3(trace_blocks(max_paths)), dim3(kTraceBlock), 0, st, sc, p, cur);
}
void launch_shadow(const TraceScene& sc, const DevPaths& p, int max_shadow, int accumulate_atomic, int record,
                   hipStream_t st) {
    hipLaunchKernelGGL(k_shadow, dim3(trace_blocks(max_shadow)), dim3(kTraceBlock), 0, st, sc, p, accumulate_atomic,
                       record);
}

}  // namespace trhip
