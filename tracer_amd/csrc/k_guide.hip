// Guide MLP kernels on gfx950 bf16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32
// accumulation: inference over the wavefront queue, fused forward + mixture
// loss + backward for training, weight / bias gradient reductions, Adam, and
// the conversion of a training pass's path records into weighted samples.
// Fragment layouts and the transposed-layer scheme are described in tr_guide.h.
#include <hipcub/hipcub.hpp>

#include "tr_guide_kernels.h"

namespace trhip {

typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16, round to nearest even (v_cvt_pk_bf16_f32 on gfx950)
TR_D unsigned short f2bf(float x) { return __builtin_bit_cast(unsigned short, static_cast<__bf16>(x)); }
TR_D float bf2f(unsigned short h) { return __uint_as_float(uint32_t(h) << 16); }
TR_D f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
TR_D bf16x8 as_bf(u16x8 v) { return __builtin_bit_cast(bf16x8, v); }
TR_D u16x8 as_u16(bf16x8 v) { return __builtin_bit_cast(u16x8, v); }

// Compiler-only barrier at the top of each sample-tile iteration: without it
// the weight-fragment loads, loop-invariant and (by type) unaliased with the
// stores, are hoisted out of the loop into hundreds of registers and spilled.
TR_D void keep_loads_in_loop() { asm volatile("" ::: "memory"); }

// Weight fragments and biases of the MLP.  Global data is read with buffer
// loads: one scalar resource plus a per-lane offset, so no per-load 64-bit
// addresses occupy vector registers.  LDS: the forward fragments staged by
// the workgroup.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
template <bool LDSW>
struct WeightSrc {
    __amdgpu_buffer_rsrc_t frag;
    __amdgpu_buffer_rsrc_t bias;
    const bf16x8* lds;
    // fragment at byte offset `byte` of the fragment buffer (64 lanes x 16 bytes)
    TR_D bf16x8 frag_at(int byte, int lane) const {
        if (LDSW) return lds[byte / 16 + lane];
        return __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(frag, lane * 16, byte, 0));
    }
    // bias[index + 4g .. index + 4g + 3]
    TR_D f32x4 bias4(int index, int g) const {
        return __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(bias, g * 16, index * 4, 0));
    }
};

TR_D __amdgpu_buffer_rsrc_t make_rsrc(const void* p, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, bytes, 0x00020000);
}
// 16-bit store at byte offset voff + soff (< 2^31, checked by the host) of a
// buffer: voff per lane, soff uniform (a scalar register).
TR_D void store_u16(__amdgpu_buffer_rsrc_t r, unsigned short v, int voff, int soff) {
    __builtin_amdgcn_raw_buffer_store_b16(v, r, voff, soff, 0);
}
// perm_feature(g, j) - 4 g: the lane-independent part of a slot's row
constexpr int slot_row(int j) { return j < 4 ? j : 12 + j; }
template <bool LDSW>
TR_D WeightSrc<LDSW> forward_src(const GuideNet& net, int H, const bf16x8* lds) {
    const GuideDims d{H};
    WeightSrc<LDSW> w;
    w.frag = make_rsrc(net.fwd, d.fwd_size() * 2);
    w.bias = make_rsrc(net.bias, (4 * H + kGuideOut) * 4);
    w.lds = lds;
    return w;
}
TR_D WeightSrc<false> backward_src(const GuideNet& net, int H) {
    const GuideDims d{H};
    WeightSrc<false> w;
    w.frag = make_rsrc(net.bwd, d.bwd_size() * 2);
    w.bias = make_rsrc(net.bias, (4 * H + kGuideOut) * 4);
    w.lds = nullptr;
    return w;
}

// acc[MT][NT] = A(layer) * b; fragment (mt, s) of the layer at byte
// base + (mt * KS + s) * 1024.
template <int KS, int MT, int NT, class Src>
TR_D void mlp_layer(const Src& src, int base, const bf16x8 (&b)[KS][NT], f32x4 (&acc)[MT][NT], int lane) {
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const bf16x8 a = src.frag_at(base + (mt * KS + s) * 1024, lane);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = mfma(a, b[s][nt], acc[mt][nt]);
        }
}

// D tiles + bias, ReLU -> bf16 B fragments of the next layer: slot j of lane
// group g in K step s is row perm(g, j) of the step, i.e. tile 2s + j / 4, row 4g + j % 4.
template <int MT, int NT, class Src>
TR_D void pack_relu(const f32x4 (&acc)[MT][NT], const Src& src, int bias_base, bf16x8 (&b)[MT / 2][NT], int lane) {
    const int g = lane >> 4;
#pragma unroll
    for (int s = 0; s < MT / 2; ++s) {
        const f32x4 b0 = src.bias4(bias_base + 32 * s, g), b1 = src.bias4(bias_base + 32 * s + 16, g);
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            u16x8 v;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int t = 2 * s + (j >> 2), i = j & 3;
                v[j] = f2bf(fmaxf(acc[t][nt][i] + (j < 4 ? b0[i] : b1[i]), 0.0f));
            }
            b[s][nt] = as_bf(v);
        }
    }
}

// Input features of the NT sample columns of this lane as B fragments: K step
// s is axis s; lane groups 0..2 hold octaves 4g .. 4g + 3 as (sin, cos) pairs,
// lane group 3 the position, normal and direction components.  v_sin_f32 /
// v_cos_f32 take their argument in revolutions, so sin(2 pi frac(2^k p))
// needs no range reduction; frac(2^k p) is exact in floating point.
template <int NT>
TR_D void encode(const GuideBox& box, const f3 (&x)[NT], const f3 (&n)[NT], const f3 (&wo)[NT], bf16x8 (&b)[3][NT],
                 int lane) {
    const int g = lane >> 4;
    const bool trig = g < 3;
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const float p = fminf(fmaxf((comp(x[nt], s) - comp(box.lo, s)) * comp(box.inv_ext, s), 0.0f), 1.0f);
            const float other[8] = {p, comp(n[nt], s), comp(wo[nt], s), 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            u16x8 v;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const float y = ldexpf(p, 4 * g + m);
                const float fr = y - floorf(y);
                v[2 * m] = f2bf(trig ? __builtin_amdgcn_sinf(fr) : other[2 * m]);
                v[2 * m + 1] = f2bf(trig ? __builtin_amdgcn_cosf(fr) : other[2 * m + 1]);
            }
            b[s][nt] = as_bf(v);
        }
}

// Hidden layers 0..3; h[l] holds the (post-ReLU) input fragments of layer l + 1.
template <int H, int NT, class Src>
TR_D void forward_hidden(const Src& src, const bf16x8 (&xin)[3][NT], bf16x8 (&h)[4][H / 32][NT], int lane) {
    const GuideDims d{H};
    f32x4 acc[H / 16][NT];
    mlp_layer<3, H / 16, NT>(src, d.fwd_offset(0) * 2, xin, acc, lane);
    pack_relu<H / 16, NT>(acc, src, 0, h[0], lane);
#pragma unroll
    for (int l = 1; l < 4; ++l) {
        mlp_layer<H / 32, H / 16, NT>(src, d.fwd_offset(l) * 2, h[l - 1], acc, lane);
        pack_relu<H / 16, NT>(acc, src, l * H, h[l], lane);
    }
}

// Whole forward pass; out holds the 48 head outputs including the output bias.
template <int H, int NT, class Src>
TR_D void forward_all(const GuideBox& box, const Src& src, const f3 (&x)[NT], const f3 (&nn)[NT], const f3 (&wo)[NT],
                      bf16x8 (&xin)[3][NT], bf16x8 (&hh)[4][H / 32][NT], f32x4 (&out)[3][NT], int lane) {
    const GuideDims d{H};
    encode<NT>(box, x, nn, wo, xin, lane);
    forward_hidden<H, NT>(src, xin, hh, lane);
    mlp_layer<H / 32, 3, NT>(src, d.fwd_offset(4) * 2, hh[3], out, lane);
    const int g = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 3; ++mt) {
        const f32x4 bo = src.bias4(4 * H + mt * 16, g);
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) out[mt][nt] += bo;
    }
}

// ---- inference over the wavefront queue --------------------------------------
// Head outputs for every queued path whose extension ray hit a surface.  With
// WLDS the forward fragments are staged in LDS once per workgroup.
// With WLDS: the workgroup copies the forward fragments (fwd_size * 2 bytes,
// the dynamic LDS its launcher passes) into LDS.
template <int H, bool WLDS>
TR_D void stage_forward(const GuideNet& net, bf16x8* lds_w) {
    if (WLDS) {
        const int nf = GuideDims{H}.fwd_size() / 8;
        for (int i = threadIdx.x; i < nf; i += blockDim.x) lds_w[i] = net.fwd[i];
        __syncthreads();
    }
}

template <int H, int NT, bool WLDS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_guide_infer(GuideNet net, DevScene s, DevPaths p, int cur, float* raw_out) {
    extern __shared__ bf16x8 lds_w[];
    stage_forward<H, WLDS>(net, lds_w);
    const WeightSrc<WLDS> src = forward_src<WLDS>(net, H, lds_w);
    const unsigned int n = *queue_counter(p, cur);
    const int* q = p.queue[cur];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4;
    const unsigned int per_block = (BLOCK / 64) * 16 * NT;
    unsigned long long evals = 0;   // network evaluations of this wave, added to stats[7] once
    for (unsigned int bbase = blockIdx.x * per_block; bbase < n; bbase += gridDim.x * per_block) {
        keep_loads_in_loop();
        const unsigned int base = bbase + wave * 16 * NT;
        f3 x[NT], nn[NT], wo[NT];
        int slot[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const unsigned int idx = base + nt * 16 + (lane & 15);
            slot[nt] = -1;
            x[nt] = nn[nt] = wo[nt] = f3s(0.0f);
            if (idx < n) {
                const int sl = q[idx];
                const float4 h = p.hit[sl];
                const int tri = __float_as_int(h.y);
                if (tri >= 0) {
                    slot[nt] = sl;
                    const float4 ro = p.ray_o[sl], rd = p.ray_d[sl];
                    const f3 dir = xyz(rd);
                    x[nt] = xyz(ro) + dir * h.x;
                    const TriShade sr = load_shade(s, tri);
                    nn[nt] = normalize(sr.n0 * (1.0f - h.z - h.w) + sr.n1 * h.z + sr.n2 * h.w);
                    wo[nt] = -dir;
                }
            }
        }
        bf16x8 xin[3][NT];
        bf16x8 hh[4][H / 32][NT];
        f32x4 out[3][NT];
        forward_all<H, NT>(net.box, src, x, nn, wo, xin, hh, out, lane);
        // evaluations of this wave's tile, counted once: lanes 0..15 hold its sample columns
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) evals += __popcll(__ballot(lane < 16 && slot[nt] >= 0));
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            if (slot[nt] < 0) continue;
            float* o = raw_out + size_t(slot[nt]) * kGuideOut;
#pragma unroll
            for (int mt = 0; mt < 3; ++mt)
                *reinterpret_cast<float4*>(o + mt * 16 + 4 * g) =
                    make_float4(out[mt][nt][0], out[mt][nt][1], out[mt][nt][2], out[mt][nt][3]);
        }
    }
    if (lane == 0 && evals > 0) atomicAdd(&p.stats[7], evals);
}

// ---- inference on explicit points (x, n, wo: 9 floats each) -------------------
// The same forward pass and weight staging as k_guide_infer, so that its
// checks against the oracle cover the kernels the renderer runs.
template <int H, int NT, bool WLDS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_guide_forward_points(GuideNet net, const float* pts, int n, float* raw_out) {
    extern __shared__ bf16x8 lds_w[];
    stage_forward<H, WLDS>(net, lds_w);
    const WeightSrc<WLDS> src = forward_src<WLDS>(net, H, lds_w);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4;
    const int per_block = (BLOCK / 64) * 16 * NT;
    for (int bbase = blockIdx.x * per_block; bbase < n; bbase += gridDim.x * per_block) {
        keep_loads_in_loop();
        const int base = bbase + wave * 16 * NT;
        f3 x[NT], nn[NT], wo[NT];
        int idxs[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int idx = base + nt * 16 + (lane & 15);
            idxs[nt] = idx;
            x[nt] = nn[nt] = wo[nt] = f3s(0.0f);
            if (idx < n) {
                const float* r = pts + size_t(idx) * 9;
                x[nt] = mk(r[0], r[1], r[2]);
                nn[nt] = mk(r[3], r[4], r[5]);
                wo[nt] = mk(r[6], r[7], r[8]);
            }
        }
        bf16x8 xin[3][NT];
        bf16x8 hh[4][H / 32][NT];
        f32x4 out[3][NT];
        forward_all<H, NT>(net.box, src, x, nn, wo, xin, hh, out, lane);
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            if (idxs[nt] >= n) continue;
#pragma unroll
            for (int mt = 0; mt < 3; ++mt)
                *reinterpret_cast<float4*>(raw_out + size_t(idxs[nt]) * kGuideOut + mt * 16 + 4 * g) =
                    make_float4(out[mt][nt][0], out[mt][nt][1], out[mt][nt][2], out[mt][nt][3]);
        }
    }
}

// ---- training: forward, mixture loss gradient, backward --------------------------
// Sample j < n_pad of the batch; j >= n contributes no gradient.  Layer inputs
// (bf16, as the forward pass consumed them) and output deltas (bf16, as the
// backward pass consumed them) are stored feature-major for the weight
// gradient GEMMs: acts[row][j], deltas[row][j].
constexpr int kLossStride = kGuideOut + 1;   // LDS row stride, odd to spread banks

template <int H, int NT>
TR_D void guide_train_tiles(const GuideNet& net, const GuideBatch& b, const GuideScratch& sc) {
    __shared__ float lds_o[4][16 * NT][kLossStride];
    const GuideDims d{H};
    const WeightSrc<false> fsrc = forward_src<false>(net, H, nullptr);
    const WeightSrc<false> bsrc = backward_src(net, H);
    const __amdgpu_buffer_rsrc_t acts = make_rsrc(sc.acts, d.act_rows() * b.n_pad * 2);
    const __amdgpu_buffer_rsrc_t deltas = make_rsrc(sc.deltas, d.delta_rows() * b.n_pad * 2);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4;
    const int per_block = 4 * 16 * NT;
    const int n_pad = b.n_pad;
    float wnorm = 1.0f;
    if (b.normalize) {
        const float ws = *b.wsum;
        wnorm = ws > 0.0f ? float(b.n_total) / ws : 0.0f;
    }
    float loss_acc = 0.0f;
    for (int bbase = blockIdx.x * per_block; bbase < n_pad; bbase += gridDim.x * per_block) {
        keep_loads_in_loop();
        const int base = bbase + wave * 16 * NT;
        f3 x[NT], nn[NT], wo[NT];
        int idxs[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int idx = base + nt * 16 + (lane & 15);
            idxs[nt] = idx;
            x[nt] = nn[nt] = wo[nt] = f3s(0.0f);
            if (idx < b.n) {
                const size_t r = size_t((unsigned long long)idx * b.mul + b.add) % (unsigned long long)b.n_total;
                const float4 r0 = b.rec[4 * r], r1 = b.rec[4 * r + 1], r2 = b.rec[4 * r + 2];
                x[nt] = mk(r0.x, r0.y, r0.z);
                nn[nt] = mk(r0.w, r1.x, r1.y);
                wo[nt] = mk(r1.z, r1.w, r2.x);
            }
        }
        bf16x8 xin[3][NT];
        bf16x8 hh[4][H / 32][NT];
        f32x4 out[3][NT];
        forward_all<H, NT>(net.box, fsrc, x, nn, wo, xin, hh, out, lane);
        // layer inputs -> acts; row of slot (s, j) = row base + 32 s + 4 g + slot_row(j)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int idx = idxs[nt];
            if (idx >= n_pad) continue;
            const int voff = (4 * g * n_pad + idx) * 2;
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                const u16x8 v = as_u16(xin[s][nt]);
#pragma unroll
                for (int j = 0; j < 8; ++j) store_u16(acts, v[j], voff, (32 * s + slot_row(j)) * n_pad * 2);
            }
#pragma unroll
            for (int l = 0; l < 4; ++l)
#pragma unroll
                for (int s = 0; s < H / 32; ++s) {
                    const u16x8 v = as_u16(hh[l][s][nt]);
#pragma unroll
                    for (int j = 0; j < 8; ++j)
                        store_u16(acts, v[j], voff, (d.act_row(l + 1) + 32 * s + slot_row(j)) * n_pad * 2);
                }
        }
        // head outputs -> LDS rows, one row per sample
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int mt = 0; mt < 3; ++mt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int f = mt * 16 + 4 * g + i;
                    lds_o[wave][nt * 16 + (lane & 15)][f] = out[mt][nt][i];
                }
        __syncthreads();
        if (lane < 16 * NT) {
            const int idx = base + lane;
            float* row = lds_o[wave][lane];
            float w = 0.0f;
            f3 wi = f3s(0.0f);
            if (idx < b.n) {
                const size_t r = size_t((unsigned long long)idx * b.mul + b.add) % (unsigned long long)b.n_total;
                const float4 r2 = b.rec[4 * r + 2], r3 = b.rec[4 * r + 3];
                w = r3.x;
                if (b.normalize) w = fminf(w * wnorm, b.clamp) / float(b.n);
                wi = mk(r2.y, r2.z, r2.w);
            }
            if (w > 0.0f) {
                loss_acc += mixture_loss_grad_row(row, wi, w);
            } else {
                for (int f = 0; f < kGuideOut; ++f) row[f] = 0.0f;
            }
        }
        __syncthreads();
        // output deltas in the D layout
        f32x4 dout[3][NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int mt = 0; mt < 3; ++mt)
#pragma unroll
                for (int i = 0; i < 4; ++i) dout[mt][nt][i] = lds_o[wave][nt * 16 + (lane & 15)][mt * 16 + 4 * g + i];
        __syncthreads();
        // bf16 deltas: stored, and the B operand (K = 48 outputs padded to 64) of the layer-4 backward
        bf16x8 bd[2][NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            u16x8 v0, v1;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int i = j & 3;
                v0[j] = f2bf(dout[j >> 2][nt][i]);
                v1[j] = j < 4 ? f2bf(dout[2][nt][i]) : (unsigned short)0;
            }
            bd[0][nt] = as_bf(v0);
            bd[1][nt] = as_bf(v1);
            const int idx = idxs[nt];
            if (idx < n_pad) {
                const int voff = (4 * g * n_pad + idx) * 2;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    store_u16(deltas, v0[j], voff, (d.delta_row(4) + slot_row(j)) * n_pad * 2);
                    if (j < 4) store_u16(deltas, v1[j], voff, (d.delta_row(4) + 32 + j) * n_pad * 2);
                }
            }
        }
        // dL/d(input of layer 4) = W4 delta4
        f32x4 gacc[H / 16][NT];
        mlp_layer<2, H / 16, NT>(bsrc, d.bwd_offset(4) * 2, bd, gacc, lane);
        bf16x8 bh[H / 32][NT];
#pragma unroll
        for (int l = 3; l >= 0; --l) {
            // delta_l = dL/d(output of layer l) * ReLU'(pre-activation): the stored
            // post-ReLU activation is positive exactly where the derivative is 1
#pragma unroll
            for (int s = 0; s < H / 32; ++s)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    const u16x8 a = as_u16(hh[l][s][nt]);
                    u16x8 v;
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const int t = 2 * s + (j >> 2), i = j & 3;
                        const bool pos = a[j] != 0 && !(a[j] & 0x8000);
                        v[j] = pos ? f2bf(gacc[t][nt][i]) : (unsigned short)0;
                    }
                    bh[s][nt] = as_bf(v);
                    const int idx = idxs[nt];
                    if (idx < n_pad) {
                        const int voff = (4 * g * n_pad + idx) * 2;
#pragma unroll
                        for (int j = 0; j < 8; ++j)
                            store_u16(deltas, v[j], voff, (d.delta_row(l) + 32 * s + slot_row(j)) * n_pad * 2);
                    }
                }
            if (l > 0) mlp_layer<H / 32, H / 16, NT>(bsrc, d.bwd_offset(l) * 2, bh, gacc, lane);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) loss_acc += __shfl_down(loss_acc, off);
    if (lane == 0 && loss_acc != 0.0f) atomicAdd(sc.loss, loss_acc);
}

// One wave per 16 samples.  At width 64 the kernel is held to 168 registers
// (28 bytes of spills), three waves per SIMD instead of one for 32 samples per
// wave at 295 registers: 81.7 -> 65.4 us per 2^16-sample step, with the same
// activations, deltas and gradients.  Width 128 would spill 356 bytes at that
// limit and keeps one wave per SIMD.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_guide_train64(GuideNet net,
                                                                                               GuideBatch b,
                                                                                               GuideScratch sc) {
    guide_train_tiles<64, 1>(net, b, sc);
}
__global__ __launch_bounds__(256) void k_guide_train128(GuideNet net, GuideBatch b, GuideScratch sc) {
    guide_train_tiles<128, 1>(net, b, sc);
}

// ---- weight gradients: dW_l[in][out] = sum_j act_l[in][j] delta_l[out][j] -----------
// One wavefront per (chunk, layer, in tile, out tile): a 16 x 16 tile with K =
// the chunk's samples; partial sums per chunk.
__global__ __launch_bounds__(256) void k_guide_dw(GuideDims d, const unsigned short* acts, const unsigned short* deltas,
                                                  int n_pad, int num_tasks, float* partial) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const int task = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (task >= num_tasks) return;
    int tiles_per_chunk = 0;
    for (int l = 0; l < kGuideLayers; ++l) tiles_per_chunk += (d.in(l) / 16) * (d.out(l) / 16);
    const int c = task / tiles_per_chunk;
    int t = task % tiles_per_chunk, l = 0;
    while (t >= (d.in(l) / 16) * (d.out(l) / 16)) {
        t -= (d.in(l) / 16) * (d.out(l) / 16);
        ++l;
    }
    const int nts = d.out(l) / 16;
    const int mt = t / nts, nt = t % nts;
    const unsigned short* arow = acts + size_t(d.act_row(l) + mt * 16 + (lane & 15)) * n_pad;
    const unsigned short* drow = deltas + size_t(d.delta_row(l) + nt * 16 + (lane & 15)) * n_pad;
    const int s0 = c * kGuideChunk, s1 = min(n_pad, s0 + kGuideChunk);
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k = s0; k < s1; k += 32) {
        const u16x4 a0 = *reinterpret_cast<const u16x4*>(arow + k + 4 * g);
        const u16x4 a1 = *reinterpret_cast<const u16x4*>(arow + k + 16 + 4 * g);
        const u16x4 b0 = *reinterpret_cast<const u16x4*>(drow + k + 4 * g);
        const u16x4 b1 = *reinterpret_cast<const u16x4*>(drow + k + 16 + 4 * g);
        const u16x8 av = {a0[0], a0[1], a0[2], a0[3], a1[0], a1[1], a1[2], a1[3]};
        const u16x8 bv = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
        acc = mfma(as_bf(av), as_bf(bv), acc);
    }
    float* dst = partial + size_t(c) * d.num_weights() + d.w_offset(l);
#pragma unroll
    for (int i = 0; i < 4; ++i) dst[(mt * 16 + 4 * g + i) * d.out(l) + nt * 16 + (lane & 15)] = acc[i];
}

// Bias gradients: one wavefront per (chunk, delta row) sums the row over the chunk.
__global__ __launch_bounds__(256) void k_guide_db(GuideDims d, const unsigned short* deltas, int n_pad, int nchunks,
                                                  float* partial) {
    const int lane = threadIdx.x & 63;
    const int task = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int rows = d.delta_rows();
    if (task >= rows * nchunks) return;
    const int row = task % rows, c = task / rows;
    const int s0 = c * kGuideChunk, s1 = min(n_pad, s0 + kGuideChunk);
    const unsigned short* src = deltas + size_t(row) * n_pad;
    float sum = 0.0f;
    for (int k = s0 + 4 * lane; k < s1; k += 256) {
        const u16x4 v = *reinterpret_cast<const u16x4*>(src + k);
        sum += (bf2f(v[0]) + bf2f(v[1])) + (bf2f(v[2]) + bf2f(v[3]));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off);
    if (lane == 0) partial[size_t(c) * rows + row] = sum;
}

// Layer and (in, out) of weight p in the master layout.
TR_HD void weight_coords(const GuideDims& d, int p, int& l, int& in, int& out) {
    l = 0;
    while (l + 1 < kGuideLayers && p >= d.w_offset(l + 1)) ++l;
    const int r = p - d.w_offset(l);
    in = r / d.out(l);
    out = r % d.out(l);
}

// Sum of the gradient partials; with update, one Adam step on the FP32 master
// parameters and a refresh of their bf16 fragment copies.
__global__ void k_guide_adam(GuideDims d, const float* pw, const float* pb, int nchunks, float* master, float* m, float* v,
                             unsigned short* fwd, unsigned short* bwd, float lr, float b1, float b2, float eps, float bc1,
                             float bc2, int update, float* grad_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int nw = d.num_weights(), np = d.num_params();
    if (p >= np) return;
    float gsum = 0.0f;
    if (p < nw) {
        for (int c = 0; c < nchunks; ++c) gsum += pw[size_t(c) * nw + p];
    } else {
        const int rows = d.delta_rows();
        for (int c = 0; c < nchunks; ++c) gsum += pb[size_t(c) * rows + (p - nw)];
    }
    if (grad_out) grad_out[p] = gsum;
    if (!update) return;
    const float mm = b1 * m[p] + (1.0f - b1) * gsum;
    const float vv = b2 * v[p] + (1.0f - b2) * gsum * gsum;
    m[p] = mm;
    v[p] = vv;
    const float w = master[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    master[p] = w;
    if (p < nw) {
        int l, in, out;
        weight_coords(d, p, l, in, out);
        const unsigned short h = f2bf(w);
        fwd[fwd_frag_index(d, l, in, out)] = h;
        bwd[bwd_frag_index(d, l, in, out)] = h;
    }
}

// bf16 fragments from the master weights.
__global__ void k_guide_pack(GuideDims d, const float* master, unsigned short* fwd, unsigned short* bwd) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.num_weights()) return;
    int l, in, out;
    weight_coords(d, p, l, in, out);
    const unsigned short h = f2bf(master[p]);
    fwd[fwd_frag_index(d, l, in, out)] = h;
    bwd[bwd_frag_index(d, l, in, out)] = h;
}

// Training pass records -> weighted samples.  Two kinds of record share the
// buffer: a continuation record (flag 1) weighs its sampled direction with
// lum((L_end - L_next) * rr / T_v), the luminance of f |cos| L_i / pdf with the
// emission at the next vertex MIS-weighted (L_next: the path radiance when the
// next vertex was shaded, L_end: when the path ended); a light-sample record
// (flag 2, set by the shadow kernel when the light was visible) carries the
// MIS-weighted f |cos| L / p_light of its light direction.  By the partition
// of unity of the MIS weights the two kinds together estimate the gradient of
// the full product target.  Records with a non-positive or non-finite weight
// carry no gradient and are dropped.  Kept samples are appended in arrival
// order with a key (path slot, vertex depth, kind) that identifies them; the
// order kernels below sort them by it, so that training does not depend on
// the order in which workgroups ran.
constexpr int kFinalizeBlock = 1024;

__global__ __launch_bounds__(kFinalizeBlock) void k_guide_finalize(DevPaths p, int num_records, float4* train,
                                                                   unsigned long long* keys, unsigned int* vals,
                                                                   unsigned int* train_count) {
    __shared__ BlockAppendLds<kFinalizeBlock> app;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool keep = false;
    float w = 0.0f;
    float4 r0{}, r1{}, r2{}, r3{};
    if (i < num_records) {
        const float4 a1 = p.rec_aux[2 * i + 1];
        if (a1.w == 1.0f || a1.w == 2.0f) {
            const float4 a0 = p.rec_aux[2 * i];
            r3 = p.rec[4 * i + 3];
            if (a1.w == 1.0f) {
                const float4 le = p.rad[__float_as_int(r3.y)];
                w = lum(mk((le.x - a1.x) * a0.x, (le.y - a1.y) * a0.y, (le.z - a1.z) * a0.z));
            } else {
                w = a0.x;   // light-sample record whose shadow ray reached the light
            }
            keep = isfinite(w) && w > 0.0f;
            if (keep) {
                r0 = p.rec[4 * i];
                r1 = p.rec[4 * i + 1];
                r2 = p.rec[4 * i + 2];
            }
        }
    }
    // one atomic per block for the sample count
    unsigned int ti, unused;
    block_append2(app, train_count, keep, train_count, false, ti, unused);
    if (keep) {
        train[4 * ti] = r0;
        train[4 * ti + 1] = r1;
        train[4 * ti + 2] = r2;
        train[4 * ti + 3] = make_float4(w, 0.0f, r3.z, 0.0f);
        const unsigned int slot = unsigned(__float_as_int(r3.y)), depth = unsigned(r3.x);
        keys[ti] = ((unsigned long long)slot << 9) | (depth << 1) | (p.rec_aux[2 * i + 1].w == 2.0f ? 1u : 0u);
        vals[ti] = ti;
    }
}

// Samples in key order, and per-workgroup sums of their weights: one wave
// shuffle tree and the waves in order, so the sums do not depend on timing.
constexpr int kGatherBlock = 256;

__global__ __launch_bounds__(kGatherBlock) void k_guide_gather(const float4* raw, const unsigned int* order, int n,
                                                               float4* train, float* partial) {
    __shared__ float wpart[kGatherBlock / 64];
    const int i = blockIdx.x * kGatherBlock + threadIdx.x;
    float w = 0.0f;
    if (i < n) {
        const unsigned int j = order[i];
#pragma unroll
        for (int k = 0; k < 4; ++k) train[4 * i + k] = raw[4 * size_t(j) + k];
        w = raw[4 * size_t(j) + 3].x;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) w += __shfl_down(w, off);
    if ((threadIdx.x & 63) == 0) wpart[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (int k = 0; k < kGatherBlock / 64; ++k) t += wpart[k];
        partial[blockIdx.x] = t;
    }
}

// Sum of the workgroup partials in a fixed order (one workgroup).
__global__ __launch_bounds__(256) void k_guide_sum(const float* partial, int nb, float* out) {
    __shared__ float s[256];
    float t = 0.0f;
    for (int k = threadIdx.x; k < nb; k += 256) t += partial[k];
    s[threadIdx.x] = t;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

// MFMA lane-layout check: D = A B for one 16 x 32 by 32 x 16 product, A and B
// row-major fp32 (rounded to bf16 here), D row-major 16 x 16.
__global__ void k_mfma_selftest(const float* A, const float* B, float* D) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
    u16x8 av, bv;
    for (int j = 0; j < 8; ++j) {
        const int k = perm_feature(g, j);
        av[j] = f2bf(A[(lane & 15) * 32 + k]);
        bv[j] = f2bf(B[k * 16 + (lane & 15)]);
    }
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = mfma(as_bf(av), as_bf(bv), acc);
    for (int i = 0; i < 4; ++i) D[(4 * g + i) * 16 + (lane & 15)] = acc[i];
}

// Mixture head on given outputs: a direction sampled with uniforms u, the
// mixture density at it and at given directions (-1 when the outputs are not
// finite).  Exposes the device sampler and density used by the shade kernel.
__global__ void k_guide_mixture(const float* raw, const float* dirs, const float* u, int n, float* samples,
                                float* pdf_samples, float* pdf_dirs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float r[5 * kGuideLobes];
    for (int f = 0; f < 5 * kGuideLobes; ++f) r[f] = raw[size_t(i) * kGuideOut + f];
    if (!mixture_raw_valid(r)) {
        samples[3 * i] = samples[3 * i + 1] = samples[3 * i + 2] = 0.0f;
        pdf_samples[i] = pdf_dirs[i] = -1.0f;
        return;
    }
    const Mixture m = mixture_from_raw(r);
    const f3 d = mixture_sample(m, u[3 * i], u[3 * i + 1], u[3 * i + 2]);
    samples[3 * i] = d.x;
    samples[3 * i + 1] = d.y;
    samples[3 * i + 2] = d.z;
    pdf_samples[i] = mixture_pdf(m, d);
    pdf_dirs[i] = mixture_pdf(m, mk(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
}

// ---- launchers ----------------------------------------------------------------
static int guide_blocks(int64_t n, int per_block) {
    int64_t b = (n + per_block - 1) / per_block;
    return int(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

// Compute units and the LDS a workgroup may allocate on the current device.
static void device_limits(int& cus, int& lds_per_block) {
    int dev = 0;
    cus = 0;
    lds_per_block = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 0;
    if (hipDeviceGetAttribute(&lds_per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        lds_per_block = 0;
}

// The forward fragments are staged in LDS once per workgroup, and the
// workgroups stay resident and stride over the queue, so that the staging is
// paid once per workgroup rather than once per tile: width 64 (43 KB) with
// three 256-thread workgroups per CU, width 128 (135 KB) with one 1024-thread
// workgroup per CU where the device lets a workgroup hold that much LDS.
void launch_guide_infer(const GuideNet& net, int H, const DevScene& s, const DevPaths& p, int cur, int max_paths,
                        float* raw_out, hipStream_t st) {
    int cus = 0, lds_max = 0;
    device_limits(cus, lds_max);
    const size_t lds = size_t(GuideDims{H}.fwd_size()) * 2;
    if (H == 64) {
        int blocks = guide_blocks(max_paths, 128);
        if (cus > 0) blocks = std::min(blocks, 3 * cus);
        hipLaunchKernelGGL((k_guide_infer<64, 2, true, 256>), dim3(blocks), dim3(256), lds, st, net, s, p, cur, raw_out);
    } else if (lds_max > 0 && size_t(lds_max) >= lds && cus > 0) {
        const int blocks = std::min(guide_blocks(max_paths, 256), cus);
        hipLaunchKernelGGL((k_guide_infer<128, 1, true, 1024>), dim3(blocks), dim3(1024), lds, st, net, s, p, cur,
                           raw_out);
    } else {
        hipLaunchKernelGGL((k_guide_infer<128, 1, false, 256>), dim3(guide_blocks(max_paths, 64)), dim3(256), 0, st, net,
                           s, p, cur, raw_out);
    }
}
void launch_guide_forward_points(const GuideNet& net, int H, const float* pts, int n, float* raw_out, hipStream_t st) {
    if (n <= 0) return;
    int cus = 0, lds_max = 0;
    device_limits(cus, lds_max);
    const size_t lds = size_t(GuideDims{H}.fwd_size()) * 2;
    if (H == 64) {
        int blocks = guide_blocks(n, 128);
        if (cus > 0) blocks = std::min(blocks, 3 * cus);
        hipLaunchKernelGGL((k_guide_forward_points<64, 2, true, 256>), dim3(blocks), dim3(256), lds, st, net, pts, n,
                           raw_out);
    } else if (lds_max > 0 && size_t(lds_max) >= lds && cus > 0) {
        const int blocks = std::min(guide_blocks(n, 256), cus);
        hipLaunchKernelGGL((k_guide_forward_points<128, 1, true, 1024>), dim3(blocks), dim3(1024), lds, st, net, pts, n,
                           raw_out);
    } else {
        hipLaunchKernelGGL((k_guide_forward_points<128, 1, false, 256>), dim3(guide_blocks(n, 64)), dim3(256), 0, st, net,
                           pts, n, raw_out);
    }
}
void launch_guide_train(const GuideNet& net, int H, const GuideBatch& b, const GuideScratch& s, hipStream_t st) {
    if (H == 64)
        hipLaunchKernelGGL(k_guide_train64, dim3(guide_blocks(b.n_pad, 64)), dim3(256), 0, st, net, b, s);
    else
        hipLaunchKernelGGL(k_guide_train128, dim3(guide_blocks(b.n_pad, 64)), dim3(256), 0, st, net, b, s);
}
void launch_guide_grads(const GuideDims& d, const GuideScratch& s, int n_pad, hipStream_t st) {
    const int nchunks = (n_pad + kGuideChunk - 1) / kGuideChunk;
    int tiles = 0;
    for (int l = 0; l < kGuideLayers; ++l) tiles += (d.in(l) / 16) * (d.out(l) / 16);
    const int tasks = tiles * nchunks;
    hipLaunchKernelGGL(k_guide_dw, dim3((tasks + 3) / 4), dim3(256), 0, st, d, s.acts, s.deltas, n_pad, tasks, s.partial_w);
    const int btasks = d.delta_rows() * nchunks;
    hipLaunchKernelGGL(k_guide_db, dim3((btasks + 3) / 4), dim3(256), 0, st, d, s.deltas, n_pad, nchunks, s.partial_b);
}
void launch_guide_adam(const GuideDims& d, const GuideScratch& s, int n_pad, float* master, float* m, float* v,
                       unsigned short* fwd, unsigned short* bwd, float lr, float b1, float b2, float eps, int step,
                       int update, float* grad_out, hipStream_t st) {
    const int nchunks = (n_pad + kGuideChunk - 1) / kGuideChunk;
    const float bc1 = 1.0f - powf(b1, float(step)), bc2 = 1.0f - powf(b2, float(step));
    const int np = d.num_params();
    hipLaunchKernelGGL(k_guide_adam, dim3((np + 255) / 256), dim3(256), 0, st, d, s.partial_w, s.partial_b, nchunks, master,
                       m, v, fwd, bwd, lr, b1, b2, eps, bc1, bc2, update, grad_out);
}
void launch_guide_pack(const GuideDims& d, const float* master, unsigned short* fwd, unsigned short* bwd, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_pack, dim3((d.num_weights() + 255) / 256), dim3(256), 0, st, d, master, fwd, bwd);
}
void launch_guide_finalize(const DevPaths& p, int num_records, float4* raw, unsigned long long* keys, unsigned int* vals,
                           unsigned int* train_count, hipStream_t st) {
    if (num_records <= 0) return;
    hipLaunchKernelGGL(k_guide_finalize, dim3((num_records + kFinalizeBlock - 1) / kFinalizeBlock), dim3(kFinalizeBlock), 0,
                       st, p, num_records, raw, keys, vals, train_count);
}
// Keys are (slot << 9) | (depth << 1) | kind with slot < 2^30 and depth < 256.
constexpr int kSampleKeyBits = 39;

size_t guide_order_temp_bytes(int n) {
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, static_cast<const unsigned long long*>(nullptr),
                                           static_cast<unsigned long long*>(nullptr), static_cast<const unsigned int*>(nullptr),
                                           static_cast<unsigned int*>(nullptr), n, 0, kSampleKeyBits) != hipSuccess)
        return 0;
    return bytes;
}
hipError_t launch_guide_order(int n, unsigned long long* keys, unsigned long long* keys_sorted, unsigned int* vals,
                              unsigned int* vals_sorted, void* temp, size_t temp_bytes, const float4* raw, float4* train,
                              float* partial, float* wsum, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    size_t bytes = temp_bytes;
    const hipError_t e = hipcub::DeviceRadixSort::SortPairs(temp, bytes, keys, keys_sorted, vals, vals_sorted, n, 0,
                                                            kSampleKeyBits, st);
    if (e != hipSuccess) return e;
    const int nb = (n + kGatherBlock - 1) / kGatherBlock;
    hipLaunchKernelGGL(k_guide_gather, dim3(nb), dim3(kGatherBlock), 0, st, raw, vals_sorted, n, train, partial);
    hipLaunchKernelGGL(k_guide_sum, dim3(1), dim3(256), 0, st, partial, nb, wsum);
    return hipGetLastError();
}
void launch_guide_mixture(const float* raw, const float* dirs, const float* u, int n, float* samples,
                          float* pdf_samples, float* pdf_dirs, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_guide_mixture, dim3((n + 255) / 256), dim3(256), 0, st, raw, dirs, u, n, samples, pdf_samples,
                       pdf_dirs);
}
void launch_mfma_selftest(const float* A, const float* B, float* D, hipStream_t st) {
    hipLaunchKernelGGL(k_mfma_selftest, dim3(1), dim3(64), 0, st, A, B, D);
}

}  // namespace trhip
