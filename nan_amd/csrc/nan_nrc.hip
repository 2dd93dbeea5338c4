// Learned radiance cache for gfx950: multiresolution hash-grid encoding plus
// a fully fused ReLU MLP on MFMA (FP16 operands, FP32 accumulation), trained
// online with Adam on a relative L2 loss.
//
// Encoding (Mueller et al., "Instant Neural Graphics Primitives with a
// Multiresolution Hash Encoding", 2022): level l has resolution
// floor(base * b^l), b = exp((ln max - ln base) / (levels - 1)); a level whose
// (res+1)^3 lattice fits its table is indexed densely, the others through the
// spatial hash x ^ y*2654435761 ^ z*805459861 (mod table size); each level
// contributes 2 trilinearly interpolated features.  Network input: 32 hash
// features, normal, albedo, zeros to 48.  Four hidden ReLU layers of 64 and a
// linear output of 16 (RGB in the first 3); no biases.
//
// MLP layout.  Every layer is computed transposed, Y^T = W^T X^T, with
// v_mfma_f32_16x16x16_f16 tiles: A = a 16x16 block of W^T (lane%16 = output
// unit, 4*(lane/16)+i = input unit), B = a 16x16 block of X^T (lane%16 =
// sample, 4*(lane/16)+i = input unit), D = Y^T (lane%16 = sample,
// 4*(lane/16)+i = output unit).  D is already the B operand of the next layer,
// so activations stay in registers from input to output.  One wave computes
// a group of 64 samples (4 tiles of 16).  Activations in memory use the same
// "fragment order": for group g, tile t, 16-unit block kb, lane, item:
//   ((g * 4 + t) * KB + kb) * 256 + lane * 4 + item,
// so a fragment is one coalesced 8-byte load per lane.
//
// Backward (per wave, layer by layer over its groups): dX^T = W dY^T (A = W
// block, B = dY^T in registers), masked by the saved activations; dW^T =
// dY^T X summed over samples, with X and dY^T transposed through LDS so that
// samples become the MFMA reduction dimension; per-wave dW partials are summed
// in a fixed order by the Adam kernel, so MLP gradients are deterministic.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "nan_nrc_internal.h"

using namespace nanx;

namespace {

constexpr uint64_t kFusedStepRecords = 32768;  // largest minibatch the fused step takes

constexpr int kWtPad = 8;           // LDS row padding (halves) of the transposed weights
constexpr int kTPad = 8;            // LDS row padding (halves) of the backward transposes
constexpr int kFwdBlock = 256;      // 4 waves
constexpr int kEncGroup = 64 * 48;  // halves of encoded input per group
constexpr int kActGroup = 64 * 64;  // halves of a 64-wide activation per group

__host__ __device__ constexpr int wt_lds_offset(int l) {
    int o = 0;
    for (int k = 0; k < l; ++k) o += nrc_layer_out(k) * (nrc_layer_in(k) + kWtPad);
    return o;
}
constexpr int kWtLds = wt_lds_offset(kNrcLayers);

__device__ __forceinline__ size_t frag_index(uint32_t s, uint32_t unit, uint32_t kb_count) {
    uint32_t g = s >> 6, t = (s >> 4) & 3u, r = s & 15u;
    uint32_t kb = unit >> 4, u = unit & 15u;
    uint32_t lane = r + 16u * (u >> 2);
    return ((size_t)(g * 4u + t) * kb_count + kb) * 256u + lane * 4u + (u & 3u);
}

// ---------------------------------------------------------------------------
// hash-grid encoding
// ---------------------------------------------------------------------------
struct Corners {
    uint32_t idx[8];
    float w[8];
};

__device__ __forceinline__ Corners level_corners(const NrcGrid &E, uint32_t level, const float *pos) {
    Corners c;
    uint32_t base[3];
    float f[3];
    level_cell(E, level, pos, base, f);
    for (uint32_t k = 0; k < 8; ++k) cell_corner(E, level, base, f, k, c.idx[k], c.w[k]);
    return c;
}

// One thread per (sample, pair of levels) over np (a multiple of 64) samples,
// level-major (blockIdx.y = the pair, so a wave gathers 64 neighbouring
// samples from one level's table): 16 corner gathers of 4 bytes (both
// features of an entry) and one 8-byte store of the four features in
// fragment order.  Pair 'pairs' writes normal, albedo and zero padding.
// Samples >= n get zeros.
// Hash levels per thread (even).  Four levels per thread (32 gathers in
// flight, half the threads) measured no faster: 13.2M queries of the
// one-bounce form took 2.88 ms per frame against 2.63 ms with two.
constexpr uint32_t kEncLevels = 2;

__global__ void nrc_encode_fwd(NrcGrid E, const float *pos, const float *nrm, const float *alb, uint32_t n,
                               uint32_t np, _Float16 *enc, float *enc_f32 /* nullable [n][48] */) {
    const uint32_t groups = (E.levels + kEncLevels - 1u) / kEncLevels;
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x, p = blockIdx.y;
    if (s >= np) return;
    if (p == groups) {
        float v[kNrcIn - 32];
        for (int k = 0; k < kNrcIn - 32; ++k) v[k] = 0.0f;
        if (s < n)
            for (int a = 0; a < 3; ++a) {
                v[a] = nrm[3 * (size_t)s + a];
                v[3 + a] = alb[3 * (size_t)s + a];
            }
        for (uint32_t u = 2 * kEncLevels * groups; u < 32u; u += 4) {  // hash features of absent levels
            *(half4_t *)&enc[frag_index(s, u, 3)] = half4_t{(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f,
                                                            (_Float16)0.0f};
            if (enc_f32 && s < n)
                for (int i = 0; i < 4; ++i) enc_f32[(size_t)s * kNrcIn + u + i] = 0.0f;
        }
        for (int k = 0; k < kNrcIn - 32; k += 4) {
            half4_t h = {(_Float16)v[k], (_Float16)v[k + 1], (_Float16)v[k + 2], (_Float16)v[k + 3]};
            *(half4_t *)&enc[frag_index(s, 32 + k, 3)] = h;
            if (enc_f32 && s < n)
                for (int i = 0; i < 4; ++i) enc_f32[(size_t)s * kNrcIn + 32 + k + i] = (float)h[i];
        }
        return;
    }
    float f[2 * kEncLevels];
    for (uint32_t k = 0; k < 2 * kEncLevels; ++k) f[k] = 0.0f;
    if (s < n) {
        for (uint32_t j = 0; j < kEncLevels; ++j) {
            uint32_t level = kEncLevels * p + j;
            if (level >= E.levels) break;
            Corners c = level_corners(E, level, pos + 3 * (size_t)s);
            const half2_t *tab = reinterpret_cast<const half2_t *>(E.table) + (size_t)level * E.entries;
            for (int k = 0; k < 8; ++k) {
                half2_t v = tab[c.idx[k]];
                f[2 * j] += c.w[k] * (float)v[0];
                f[2 * j + 1] += c.w[k] * (float)v[1];
            }
        }
    }
    for (uint32_t q = 0; q < kEncLevels / 2; ++q) {
        uint32_t u = 2 * kEncLevels * p + 4 * q;
        half4_t h = {(_Float16)f[4 * q], (_Float16)f[4 * q + 1], (_Float16)f[4 * q + 2], (_Float16)f[4 * q + 3]};
        *(half4_t *)&enc[frag_index(s, u, 3)] = h;
        if (enc_f32 && s < n)
            for (int i = 0; i < 4; ++i) enc_f32[(size_t)s * kNrcIn + u + i] = (float)h[i];
    }
}


// One thread per (sample, corner) of one level (blockIdx.y = level): eight
// consecutive lanes take the eight corners of a sample, one packed atomic
// each.  The atomics are the cost; eight times the threads of one per
// (sample, level) keep more of them in flight (73 -> 43 us for 8,192
// samples).
__global__ void nrc_encode_bwd(NrcGrid E, const float *pos, uint32_t n, const float *denc, __half2 *grad_table) {
    uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t level = blockIdx.y, s = (uint32_t)(gid >> 3), k = (uint32_t)(gid & 7u);
    if (s >= n) return;
    float g0 = denc[(size_t)s * kNrcIn + 2 * level], g1 = denc[(size_t)s * kNrcIn + 2 * level + 1];
    if (g0 == 0.0f && g1 == 0.0f) return;
    g0 *= kTableGradScale;
    g1 *= kTableGradScale;
    uint32_t base[3], idx;
    float f[3], w;
    level_cell(E, level, pos + 3 * (size_t)s, base, f);
    cell_corner(E, level, base, f, k, idx, w);
    __half2 *g = grad_table + (size_t)level * E.entries;
    unsafeAtomicAdd(&g[idx], __halves2half2(__float2half(w * g0), __float2half(w * g1)));
}

// ---------------------------------------------------------------------------
// fused MLP forward
// ---------------------------------------------------------------------------
// acc[t][ot] = W^T block (ot, .) x X^T tile t, over IN/16 input blocks taken
// in pairs on the K=32 MFMA (an odd last block on K=16).
template <int IN, int OUT>
__device__ __forceinline__ void layer_fwd(const _Float16 *wl, const half4_t (&x)[4][4], floatx4_t (&acc)[4][4]) {
    const int lane = __lane_id(), r = lane & 15, q = lane >> 4;
#pragma unroll
    for (int ot = 0; ot < OUT / 16; ++ot) {
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t][ot] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
        const _Float16 *row = &wl[(ot * 16 + r) * (IN + kWtPad) + 4 * q];
#pragma unroll
        for (int kb = 0; kb + 1 < IN / 16; kb += 2) {
            half4_t a0 = *(const half4_t *)&row[kb * 16], a1 = *(const half4_t *)&row[kb * 16 + 16];
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t][ot] = mfma32(a0, a1, x[t][kb], x[t][kb + 1], acc[t][ot]);
        }
        if ((IN / 16) % 2) {
            half4_t a = *(const half4_t *)&row[(IN / 16 - 1) * 16];
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t][ot] = mfma16(a, x[t][IN / 16 - 1], acc[t][ot]);
        }
    }
}

// ReLU after the rounding to FP16, on packed halves (rounding is monotone and
// keeps zero, so max(round(a), 0) = round(max(a, 0)) up to the sign of zero).
__device__ __forceinline__ void relu_to_half(const floatx4_t (&acc)[4][4], half4_t (&x)[4][4]) {
    const half4_t zero = {(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f};
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
            x[t][kb] = __builtin_elementwise_max(__builtin_convertvector(acc[t][kb], half4_t), zero);
}

// Copies the per-layer [out][in] FP16 weights into LDS rows padded by kWtPad,
// 16 bytes at a time (every layer width and offset is a multiple of 8 halves).
__device__ __forceinline__ void stage_weights(const _Float16 *wt, _Float16 *wl, uint32_t tid, uint32_t threads) {
    static_assert(kNrcIn % 8 == 0 && kNrcWidth % 8 == 0 && kWtPad % 8 == 0, "16-byte weight rows");
    for (int l = 0; l < kNrcLayers; ++l) {
        const int in = nrc_layer_in(l), chunks = in * nrc_layer_out(l) / 8;
        const uint4 *src = reinterpret_cast<const uint4 *>(wt + nrc_layer_offset(l));
        _Float16 *dst = wl + wt_lds_offset(l);
        for (int i = (int)tid; i < chunks; i += (int)threads) {
            const int e = 8 * i;
            *reinterpret_cast<uint4 *>(&dst[(e / in) * (in + kWtPad) + e % in]) = src[i];
        }
    }
}

// wt: fp16 weights, per layer [out][in].  enc: encoded input (fragment order,
// KB = 3).  pred[s] = first four outputs of sample s < n.  save (SAVE): the
// inputs of layers 1..4, [layer-1][group][4096] fragment order.
template <bool SAVE>
__global__ __launch_bounds__(kFwdBlock) __attribute__((amdgpu_waves_per_eu(SAVE ? 3 : 4, 4))) void nrc_mlp_fwd(const _Float16 *wt, const _Float16 *enc, uint32_t groups,
                                                          uint32_t n, float4 *pred, _Float16 *save,
                                                          uint32_t cap_groups, float *stats) {
    __shared__ __align__(16) _Float16 wl[kWtLds];
    if (SAVE && blockIdx.x == 0 && threadIdx.x == 0) {  // training statistics of this step
        stats[0] = 0.0f;
        stats[1] = (float)n;
        stats[2] = 0.0f;
    }
    stage_weights(wt, wl, threadIdx.x, kFwdBlock);
    __syncthreads();
    const int lane = __lane_id(), r = lane & 15, q = lane >> 4;
    const uint32_t waves = gridDim.x * (kFwdBlock / 64);
    for (uint32_t g = blockIdx.x * (kFwdBlock / 64) + threadIdx.x / 64; g < groups; g += waves) {
        // weights are read from LDS per group: kept in registers across groups
        // they would leave room for two waves per SIMD
        __asm__ volatile("" ::: "memory");
        half4_t x[4][4];
        floatx4_t acc[4][4];
        const _Float16 *e = enc + (size_t)g * kEncGroup;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int kb = 0; kb < 3; ++kb) x[t][kb] = *(const half4_t *)&e[(t * 3 + kb) * 256 + lane * 4];
        layer_fwd<kNrcIn, kNrcWidth>(wl + wt_lds_offset(0), x, acc);
#pragma unroll
        for (int l = 1; l < kNrcLayers; ++l) {
            relu_to_half(acc, x);
            if (SAVE) {
                _Float16 *sv = save + ((size_t)(l - 1) * cap_groups + g) * kActGroup;
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int kb = 0; kb < 4; ++kb) *(half4_t *)&sv[(t * 4 + kb) * 256 + lane * 4] = x[t][kb];
            }
            if (l < kNrcLayers - 1) layer_fwd<kNrcWidth, kNrcWidth>(wl + wt_lds_offset(l), x, acc);
            else layer_fwd<kNrcWidth, kNrcOut>(wl + wt_lds_offset(l), x, acc);
        }
        if (q == 0) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                uint32_t s = g * 64 + t * 16 + r;
                if (s < n) pred[s] = make_float4(acc[t][0][0], acc[t][0][1], acc[t][0][2], acc[t][0][3]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// loss
// ---------------------------------------------------------------------------
// Per-sample gradient of sum_c (p-t)^2 / (p^2 + 0.01) (denominator held
// constant), clamped to +-kGradClamp, written as the output-layer dY
// (fragment order, KB = 1) for every sample < np; adds the loss sum to
// stats[0].  Invalid and padding samples get zero gradient.
constexpr int kLossBlock = 1024;

__global__ __launch_bounds__(kLossBlock) void nrc_loss(const float4 *pred, const float *target, const uint32_t *valid, uint32_t n, uint32_t np,
                         _Float16 *dy, float *stats) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    float loss = 0.0f;
    bool is_valid = false;
    if (s < np) {
        float gr[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (s < n && (!valid || valid[s])) {
            is_valid = true;
            float4 p4 = pred[s];
            float p[3] = {p4.x, p4.y, p4.z};
            for (int c = 0; c < 3; ++c) {
                float tc = target[3 * (size_t)s + c];
                float d = p[c] * p[c] + 0.01f;
                float diff = p[c] - tc;
                loss += diff * diff / d;
                gr[c] = fminf(fmaxf(2.0f * diff / d, -kGradClamp), kGradClamp);
            }
        }
        uint32_t g = s >> 6, t = (s >> 4) & 3u, r = s & 15u;
        _Float16 *d = dy + (size_t)g * kActGroup + t * 256;
        half4_t h0 = {(_Float16)gr[0], (_Float16)gr[1], (_Float16)gr[2], (_Float16)gr[3]};
        half4_t z = {(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f};
        *(half4_t *)&d[r * 4] = h0;
        *(half4_t *)&d[(r + 16) * 4] = z;
        *(half4_t *)&d[(r + 32) * 4] = z;
        *(half4_t *)&d[(r + 48) * 4] = z;
    }
    // block sums (fixed order), one atomic per block and statistic
    __shared__ float part[2][kLossBlock / 64];
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off);
    uint64_t ballot = __ballot(is_valid);
    if (__lane_id() == 0) {
        part[0][threadIdx.x / 64] = loss;
        part[1][threadIdx.x / 64] = (float)__popcll(ballot);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float l = 0.0f, v = 0.0f;
        for (int k = 0; k < kLossBlock / 64; ++k) {
            l += part[0][k];
            v += part[1][k];
        }
        if (l != 0.0f) atomicAdd(&stats[0], l);
        if (v != 0.0f) atomicAdd(&stats[2], v);
    }
}

// ---------------------------------------------------------------------------
// fused MLP backward
// ---------------------------------------------------------------------------
template <int IN, int OUT, int L>
__device__ __forceinline__ void layer_bwd(const _Float16 *wn, const _Float16 *xin, uint32_t x_stride, _Float16 *dyb,
                                          uint32_t groups, float *part, float *denc, _Float16 *tx, _Float16 *ty) {
    constexpr int KI = IN / 16, KO = OUT / 16;
    constexpr int TS = 64 + kTPad;
    const int lane = __lane_id(), r = lane & 15, q = lane >> 4;
    floatx4_t dw[KO][KI];
#pragma unroll
    for (int ot = 0; ot < KO; ++ot)
#pragma unroll
        for (int it = 0; it < KI; ++it) dw[ot][it] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
    const _Float16 *w = wn + nrc_layer_offset(L);  // [in][out]
    for (uint32_t g = blockIdx.x; g < groups; g += gridDim.x) {
        _Float16 *dyg = dyb + (size_t)g * kActGroup;
        const _Float16 *xg = xin + (size_t)g * x_stride;
        half4_t dy[4][KO], x[4][KI];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
#pragma unroll
            for (int kb = 0; kb < KO; ++kb) dy[t][kb] = *(const half4_t *)&dyg[(t * KO + kb) * 256 + lane * 4];
#pragma unroll
            for (int kb = 0; kb < KI; ++kb) x[t][kb] = *(const half4_t *)&xg[(t * KI + kb) * 256 + lane * 4];
        }
        // transposes: ty[out unit][sample], tx[in unit][sample]
#pragma unroll
        for (int t = 0; t < 4; ++t) {
#pragma unroll
            for (int kb = 0; kb < KO; ++kb)
#pragma unroll
                for (int i = 0; i < 4; ++i) ty[(kb * 16 + 4 * q + i) * TS + t * 16 + r] = dy[t][kb][i];
#pragma unroll
            for (int kb = 0; kb < KI; ++kb)
#pragma unroll
                for (int i = 0; i < 4; ++i) tx[(kb * 16 + 4 * q + i) * TS + t * 16 + r] = x[t][kb][i];
        }
        __syncthreads();
        // dW^T[o][i] += sum_s dY[s][o] X[s][i]
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            half4_t b[KI];
#pragma unroll
            for (int it = 0; it < KI; ++it) b[it] = *(const half4_t *)&tx[(it * 16 + r) * TS + ks * 16 + 4 * q];
#pragma unroll
            for (int ot = 0; ot < KO; ++ot) {
                half4_t a = *(const half4_t *)&ty[(ot * 16 + r) * TS + ks * 16 + 4 * q];
#pragma unroll
                for (int it = 0; it < KI; ++it) dw[ot][it] = mfma16(a, b[it], dw[ot][it]);
            }
        }
        __syncthreads();
        // dX^T[i][s] = sum_o W[i][o] dY^T[o][s]
#pragma unroll
        for (int it = 0; it < KI; ++it) {
            floatx4_t acc[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ko = 0; ko < KO; ++ko) {
                half4_t a = *(const half4_t *)&w[(it * 16 + r) * OUT + ko * 16 + 4 * q];
#pragma unroll
                for (int t = 0; t < 4; ++t) acc[t] = mfma16(a, dy[t][ko], acc[t]);
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                if (L > 0) {
                    half4_t h;
#pragma unroll
                    for (int i = 0; i < 4; ++i) h[i] = (_Float16)((float)x[t][it][i] > 0.0f ? acc[t][i] : 0.0f);
                    *(half4_t *)&dyg[(t * KI + it) * 256 + lane * 4] = h;  // dY of the layer below (KB = 4)
                } else {
                    uint32_t s = g * 64 + t * 16 + r;
                    *(float4 *)&denc[(size_t)s * kNrcIn + it * 16 + 4 * q] =
                        make_float4(acc[t][0], acc[t][1], acc[t][2], acc[t][3]);
                }
            }
        }
    }
    // partial dW in weight order [out][in]
    float *p = part + nrc_layer_offset(L);
#pragma unroll
    for (int ot = 0; ot < KO; ++ot)
#pragma unroll
        for (int it = 0; it < KI; ++it)
#pragma unroll
            for (int i = 0; i < 4; ++i) p[(ot * 16 + 4 * q + i) * IN + it * 16 + r] = dw[ot][it][i];
}

// One wave per block; block b owns groups b, b + grid, ...  dyb holds the
// output-layer gradient on entry (KB = 1 per group) and is overwritten layer
// by layer.  partial: [grid][kNrcWeights].  denc: [groups * 64][48].
__global__ __launch_bounds__(64) void nrc_mlp_bwd(const _Float16 *wn, const _Float16 *enc, const _Float16 *save,
                                                   uint32_t cap_groups, _Float16 *dyb, uint32_t groups,
                                                   float *partial, float *denc) {
    __shared__ _Float16 tx[64 * (64 + kTPad)], ty[64 * (64 + kTPad)];
    float *part = partial + (size_t)blockIdx.x * kNrcWeights;
    const size_t layer = (size_t)cap_groups * kActGroup;
    layer_bwd<kNrcWidth, kNrcOut, 4>(wn, save + 3 * layer, kActGroup, dyb, groups, part, denc, tx, ty);
    layer_bwd<kNrcWidth, kNrcWidth, 3>(wn, save + 2 * layer, kActGroup, dyb, groups, part, denc, tx, ty);
    layer_bwd<kNrcWidth, kNrcWidth, 2>(wn, save + 1 * layer, kActGroup, dyb, groups, part, denc, tx, ty);
    layer_bwd<kNrcWidth, kNrcWidth, 1>(wn, save, kActGroup, dyb, groups, part, denc, tx, ty);
    layer_bwd<kNrcIn, kNrcWidth, 0>(wn, enc, kEncGroup, dyb, groups, part, denc, tx, ty);
}

// ---------------------------------------------------------------------------
// optimiser
// ---------------------------------------------------------------------------
// Fixed-order sum of the per-wave partials for weights [k0, k0 + 32):
// thread (k, q) sums partials q, q + 8, ...; the 8 sums combine in order.
constexpr int kReduceWeights = 32, kReduceSplit = 8;

__device__ __forceinline__ float reduce_partials(const float *partial, uint32_t parts, uint32_t k, float *lds) {
    const uint32_t kl = threadIdx.x % kReduceWeights, q = threadIdx.x / kReduceWeights;
    float g = 0.0f;
    if (k < (uint32_t)kNrcWeights)
        for (uint32_t p = q; p < parts; p += kReduceSplit) g += partial[(size_t)p * kNrcWeights + k];
    lds[q * kReduceWeights + kl] = g;
    __syncthreads();
    float sum = 0.0f;
    for (int j = 0; j < kReduceSplit; ++j) sum += lds[j * kReduceWeights + kl];
    return sum;
}

// MLP weights: Adam on the reduced gradient (block of 256 = 32 weights x 8).
__global__ __launch_bounds__(256) void nrc_adam_mlp(float *w, float *m, float *v, const float *partial,
                                                     uint32_t parts, _Float16 *w16t, _Float16 *w16n,
                                                     const float *stats, AdamArgs a, AdamState *as) {
    __shared__ float lds[kReduceWeights * kReduceSplit];
    __shared__ float corr[2];
    uint32_t k = blockIdx.x * kReduceWeights + threadIdx.x % kReduceWeights;
    if (threadIdx.x == 0) {
        uint32_t step = as->step + 1u;
        corr[0] = bias_correction(a.b1, step);
        corr[1] = bias_correction(a.b2, step);
        if (blockIdx.x == 0) {
            as->next_step = step;
            as->c1 = corr[0];
            as->c2 = corr[1];
        }
    }
    float g = reduce_partials(partial, parts, k, lds);  // synchronises the block
    if (threadIdx.x >= kReduceWeights || k >= (uint32_t)kNrcWeights) return;
    const float c1 = corr[0], c2 = corr[1];
    g *= grad_scale(stats);
    float mi = a.b1 * m[k] + (1.0f - a.b1) * g;
    float vi = a.b2 * v[k] + (1.0f - a.b2) * g * g;
    m[k] = mi;
    v[k] = vi;
    float wi = w[k] - a.lr * (mi / c1) / (sqrtf(vi / c2) + a.eps);
    w[k] = wi;
    w16t[k] = (_Float16)wi;
    // plain [in][out] copy for the backward kernel
    int l = 0;
    while (l + 1 < kNrcLayers && (int)k >= nrc_layer_offset(l + 1)) ++l;
    int in = nrc_layer_in(l), out = nrc_layer_out(l), rel = (int)k - nrc_layer_offset(l);
    int o = rel / in, i = rel % in;
    w16n[nrc_layer_offset(l) + i * out + o] = (_Float16)wi;
}

// Hash table, one entry (2 features) per thread: dense, or lazy (entries with
// zero gradient keep their state).  Non-finite FP16 gradient sums are dropped.
// An entry's state is one 32-byte record, so a touched entry costs one read
// and one write of a single sector wherever it lies in the table.
__global__ void nrc_adam_table(TableState *st, __half2 *grad, __half2 *w16, uint64_t n, const float *stats, AdamArgs a,
                               AdamState *as, int dense) {
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const float c1 = as->c1, c2 = as->c2;
    if (k == 0) as->step = as->next_step;
    if (k >= n) return;
    __half2 g2 = grad[k];
    float g[2] = {__low2float(g2), __high2float(g2)};
    if (g[0] == 0.0f && g[1] == 0.0f && !dense) return;
    grad[k] = __halves2half2(__float2half(0.0f), __float2half(0.0f));
    float sc = grad_scale(stats) / kTableGradScale;
    float4 *rec = reinterpret_cast<float4 *>(st + k);
    float4 wm = rec[0], vp = rec[1];
    float wv[2] = {wm.x, wm.y}, mv[2] = {wm.z, wm.w}, vv[2] = {vp.x, vp.y};
    for (int c = 0; c < 2; ++c) {
        float gc = isfinite(g[c]) ? g[c] * sc : 0.0f;
        mv[c] = a.b1 * mv[c] + (1.0f - a.b1) * gc;
        vv[c] = a.b2 * vv[c] + (1.0f - a.b2) * gc * gc;
        wv[c] = wv[c] - a.lr * (mv[c] / c1) / (sqrtf(vv[c] / c2) + a.eps);
    }
    rec[0] = make_float4(wv[0], wv[1], mv[0], mv[1]);
    rec[1] = make_float4(vv[0], vv[1], 0.0f, 0.0f);
    w16[k] = __halves2half2(__float2half(wv[0]), __float2half(wv[1]));
}

// FP32 master weights of the table in and out of the state records (one
// float2 per entry), for the contiguous host copies of create/get/set.
__global__ void nrc_table_scatter(TableState *st, const float2 *w, uint64_t n) {
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    *reinterpret_cast<float2 *>(st[k].w) = w[k];
}

__global__ void nrc_table_gather(const TableState *st, float2 *w, uint64_t n) {
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    w[k] = *reinterpret_cast<const float2 *>(st[k].w);
}

__global__ __launch_bounds__(256) void nrc_sum_partials(const float *partial, uint32_t parts, const float *stats,
                                                         float *out) {
    __shared__ float lds[kReduceWeights * kReduceSplit];
    uint32_t k = blockIdx.x * kReduceWeights + threadIdx.x % kReduceWeights;
    float g = reduce_partials(partial, parts, k, lds);
    if (threadIdx.x < kReduceWeights && k < (uint32_t)kNrcWeights) out[k] = g * grad_scale(stats);
}

// The accumulated hash-table gradient as a float pair per entry, on the
// scale of the mean loss (what the table's Adam step uses), clearing it.
__global__ void nrc_table_grad_out(__half2 *grad, uint64_t n, const float *stats, float2 *out) {
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    __half2 g = grad[k];
    grad[k] = __halves2half2(__float2half(0.0f), __float2half(0.0f));
    float sc = grad_scale(stats) / kTableGradScale;
    out[k] = make_float2(__low2float(g) * sc, __high2float(g) * sc);
}

__global__ void nrc_scale_rows(const float *in, uint32_t n, const float *stats, float *out) {
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (uint64_t)n * kNrcIn) return;
    out[k] = in[k] * grad_scale(stats);
}

__global__ void nrc_rgb(const float4 *pred, uint32_t n, float *out, int raw) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float4 p = pred[s];
    if (raw) {
        reinterpret_cast<float4 *>(out)[s] = p;
    } else {
        out[3 * (size_t)s + 0] = fmaxf(p.x, 0.0f);
        out[3 * (size_t)s + 1] = fmaxf(p.y, 0.0f);
        out[3 * (size_t)s + 2] = fmaxf(p.z, 0.0f);
    }
}

__global__ void nrc_add_loss(const float *stats, float *loss_out) {
    if (threadIdx.x == 0) *loss_out += stats[0] * grad_scale(stats);
}

// Level l's entries in 32-bit words (one FP16 pair each) that a query can
// read: the (res+1)^3 lattice of a densely indexed level, the whole table of
// a hashed one.
__host__ __device__ inline uint32_t snapshot_words(const NrcGrid &E, uint32_t l) {
    uint32_t r1 = E.res[l] + 1u;
    return E.dense[l] ? r1 * r1 * r1 : E.entries;
}

// The snapshot copy: blockIdx.y < levels copies that level's used entries,
// blockIdx.y == levels the FP16 weights; snap holds the weights, then the
// table in its own layout.
__global__ void nrc_snapshot_copy(NrcGrid E, const uint32_t *w16t, uint32_t *snap) {
    const uint32_t l = blockIdx.y;
    const uint32_t *src = w16t;
    uint32_t *dst = snap;
    uint32_t count = kNrcWeights / 2;
    if (l < E.levels) {
        const size_t at = (size_t)l * E.entries;
        src = reinterpret_cast<const uint32_t *>(E.table) + at;
        dst = snap + kNrcWeights / 2 + at;
        count = snapshot_words(E, l);
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
}  // namespace

namespace nanx {
NrcGrid nrc_grid(const NanNrc *c) {
    NrcGrid E;
    E.table = c->t16;
    E.levels = c->cfg.levels;
    E.entries = c->table_entries;
    for (int l = 0; l < kNrcMaxLevels; ++l) {
        E.res[l] = c->level_res[l];
        E.dense[l] = c->level_dense[l];
    }
    for (int a = 0; a < 3; ++a) {
        E.bmin[a] = c->cfg.bbox_min[a];
        E.bscale[a] = 1.0f / (c->cfg.bbox_max[a] - c->cfg.bbox_min[a]);
    }
    return E;
}
}  // namespace nanx

namespace {

NrcGrid grid_of(const NanNrc *c) { return nrc_grid(c); }

uint32_t groups_for(uint32_t n) { return (n + 63u) / 64u; }

// Four resident blocks of four waves per CU: 1024 blocks fill the chip once,
// and each block stages the weights once for all the groups it takes.
uint32_t fwd_blocks(uint32_t groups) { return std::max<uint32_t>(1, std::min<uint32_t>((groups + 3) / 4, 1024)); }

uint32_t bwd_waves(uint32_t groups) { return std::max<uint32_t>(1, std::min<uint32_t>(groups, kNrcBwdWaves)); }

// Encoded input of n samples into enc (every entry of the groups used is
// written, padding rows included).
void encode_to(NanNrc *c, const float *pos, const float *nrm, const float *alb, uint32_t n, _Float16 *enc,
               float *enc_f32, hipStream_t s) {
    uint32_t np = groups_for(n) * 64;
    dim3 grid(grid_for(np, 256), (c->cfg.levels + kEncLevels - 1) / kEncLevels + 1);
    hipLaunchKernelGGL(nrc_encode_fwd, grid, dim3(256), 0, s, grid_of(c), pos, nrm, alb, n, np, enc, enc_f32);
}

void encode(NanNrc *c, const float *pos, const float *nrm, const float *alb, uint32_t n, float *enc_f32,
            hipStream_t s) {
    encode_to(c, pos, nrm, alb, n, c->enc, enc_f32, s);
}

// Forward (saving activations), loss and backward into c->wpart / c->denc.
void forward_backward(NanNrc *c, const float *pos, const float *nrm, const float *alb, const float *target,
                      const uint32_t *valid, uint32_t n, hipStream_t s) {
    uint32_t groups = groups_for(n), np = groups * 64;
    encode_to(c, pos, nrm, alb, n, c->tenc, nullptr, s);
    hipLaunchKernelGGL((nrc_mlp_fwd<true>), dim3(fwd_blocks(groups)), dim3(kFwdBlock), 0, s, c->w16t, c->tenc,
                       groups, n, c->tpred, c->save, c->cap_train_groups, c->stats);
    hipLaunchKernelGGL(nrc_loss, dim3(grid_for(np, kLossBlock)), dim3(kLossBlock), 0, s, c->tpred, target, valid, n,
                       np, c->dy, c->stats);
    hipLaunchKernelGGL(nrc_mlp_bwd, dim3(bwd_waves(groups)), dim3(64), 0, s, c->w16n, c->tenc, c->save,
                       c->cap_train_groups, c->dy, groups, c->wpart, c->denc);
}

}  // namespace

NanNrc::~NanNrc() {
    for (void *p : {(void *)w32, (void *)wm, (void *)wv, (void *)wpart, (void *)w16t, (void *)w16n, (void *)ts,
                    (void *)tg, (void *)t16, (void *)enc, (void *)save, (void *)dy, (void *)denc, (void *)pred,
                    (void *)tenc, (void *)tpred, (void *)snap, (void *)stats, (void *)lossp, adam})
        hipFree(p);
}

namespace nanx {

int nrc_train_one(NanNrc *c, const float *pos, const float *nrm, const float *alb, const float *target,
                  const uint32_t *valid, uint32_t n, hipStream_t s);

// Prediction scratch (encoded input, outputs) for n samples.  Growing it
// allocates, so it is refused while `s` is being captured into a graph.
int nrc_reserve(NanNrc *c, uint32_t n, hipStream_t s) {
    uint32_t groups = groups_for(n);
    if (groups <= c->cap_groups) return 0;
    if (stream_capturing(s))
        return fail(nullptr, 0, "nan_nrc: %u samples exceed the %u reserved; reserve before graph capture", n,
                    c->cap_groups * 64);
    HIP_TRY(hipDeviceSynchronize(), nullptr, 0);  // earlier launches may still use the old buffers
    hipFree(c->enc);
    hipFree(c->pred);
    c->enc = nullptr;
    c->pred = nullptr;
    c->cap_groups = 0;
    uint32_t cap = std::max<uint32_t>(groups, 1024);
    HIP_TRY(hipMalloc((void **)&c->enc, (size_t)cap * kEncGroup * 2), nullptr, 0);
    HIP_TRY(hipMalloc((void **)&c->pred, (size_t)cap * 64 * sizeof(float4)), nullptr, 0);
    c->cap_groups = cap;
    return 0;
}

// Training scratch (encoded input, outputs, saved activations, gradients)
// for n samples.
int nrc_reserve_train(NanNrc *c, uint32_t n, hipStream_t s) {
    uint32_t groups = groups_for(n);
    if (groups <= c->cap_train_groups) return 0;
    if (stream_capturing(s))
        return fail(nullptr, 0, "nan_nrc: %u training records exceed the %u reserved; reserve before graph capture",
                    n, c->cap_train_groups * 64);
    HIP_TRY(hipDeviceSynchronize(), nullptr, 0);
    for (void *p : {(void *)c->save, (void *)c->dy, (void *)c->denc, (void *)c->tenc, (void *)c->tpred}) hipFree(p);
    c->save = c->dy = c->tenc = nullptr;
    c->denc = nullptr;
    c->tpred = nullptr;
    c->cap_train_groups = 0;
    uint32_t cap = std::max<uint32_t>(groups, 1024);
    HIP_TRY(hipMalloc((void **)&c->tenc, (size_t)cap * kEncGroup * 2), nullptr, 0);
    HIP_TRY(hipMalloc((void **)&c->tpred, (size_t)cap * 64 * sizeof(float4)), nullptr, 0);
    HIP_TRY(hipMalloc((void **)&c->save, (size_t)(kNrcLayers - 1) * cap * kActGroup * 2), nullptr, 0);
    HIP_TRY(hipMalloc((void **)&c->dy, (size_t)cap * kActGroup * 2), nullptr, 0);
    HIP_TRY(hipMalloc((void **)&c->denc, (size_t)cap * 64 * kNrcIn * 4), nullptr, 0);
    c->cap_train_groups = cap;
    return 0;
}

int nrc_predict(NanNrc *c, const float *pos, const float *nrm, const float *alb, uint32_t n, float4 *pred,
                hipStream_t s, bool from_snapshot) {
    if (n == 0) return 0;
    if (from_snapshot && !c->snap) return fail(nullptr, 0, "nan_nrc: no parameter snapshot reserved");
    if (nrc_reserve(c, n, s)) return -1;
    uint32_t groups = groups_for(n);
    NrcGrid E = grid_of(c);
    const _Float16 *wt = c->w16t;
    if (from_snapshot) {
        wt = c->snap;
        E.table = c->snap + kNrcWeights;
    }
    uint32_t np = groups * 64;
    dim3 grid(grid_for(np, 256), (c->cfg.levels + kEncLevels - 1) / kEncLevels + 1);
    hipLaunchKernelGGL(nrc_encode_fwd, grid, dim3(256), 0, s, E, pos, nrm, alb, n, np, c->enc, (float *)nullptr);
    hipLaunchKernelGGL((nrc_mlp_fwd<false>), dim3(fwd_blocks(groups)), dim3(kFwdBlock), 0, s, wt, c->enc, groups, n,
                       pred, (_Float16 *)nullptr, c->cap_groups, (float *)nullptr);
    return check_launch("nrc predict");
}

int nrc_reserve_snapshot(NanNrc *c, hipStream_t s) {
    if (c->snap) return 0;
    if (stream_capturing(s))
        return fail(nullptr, 0, "nan_nrc: parameter snapshot not reserved; reserve before graph capture");
    HIP_TRY(hipMalloc((void **)&c->snap, ((size_t)kNrcWeights + c->table_params) * 2), nullptr, 0);
    return 0;
}

int nrc_snapshot(NanNrc *c, hipStream_t s) {
    if (nrc_reserve_snapshot(c, s)) return -1;
    uint32_t most = kNrcWeights / 2;
    for (uint32_t l = 0; l < c->cfg.levels; ++l) most = std::max(most, snapshot_words(grid_of(c), l));
    dim3 grid(std::min<uint32_t>(grid_for(most, 256), 64), c->cfg.levels + 1);
    hipLaunchKernelGGL(nrc_snapshot_copy, grid, dim3(256), 0, s, grid_of(c), (const uint32_t *)c->w16t,
                       (uint32_t *)c->snap);
    return check_launch("nrc snapshot");
}

int nrc_infer_into(NanNrc *c, const float *pos, const float *nrm, const float *alb, uint32_t n, _Float16 *enc,
                   uint32_t enc_groups, float4 *pred, float *out, hipStream_t s) {
    if (n == 0) return 0;
    uint32_t groups = groups_for(n);
    if (groups > enc_groups) return fail(nullptr, 0, "nan_nrc: %u samples exceed the %u scratch groups", n, enc_groups);
    encode_to(c, pos, nrm, alb, n, enc, nullptr, s);
    hipLaunchKernelGGL((nrc_mlp_fwd<false>), dim3(fwd_blocks(groups)), dim3(kFwdBlock), 0, s, c->w16t, enc, groups,
                       n, pred, (_Float16 *)nullptr, enc_groups, (float *)nullptr);
    hipLaunchKernelGGL(nrc_rgb, dim3(grid_for(n, 256)), dim3(256), 0, s, pred, n, out, 0);
    return 0;
}

int nrc_train_step(NanNrc *c, const float *pos, const float *nrm, const float *alb, const float *target,
                   const uint32_t *valid, uint32_t n, hipStream_t s) {
    if (n == 0) return 0;
    uint32_t parts = std::max<uint32_t>(1, std::min<uint32_t>(c->cfg.minibatches, n));
    // consecutive slices of the records, one Adam step each; counts are the slice sizes
    for (uint32_t k = 0; k < parts; ++k) {
        uint64_t a = (uint64_t)n * k / parts, b = (uint64_t)n * (k + 1) / parts;
        const uint32_t *v = valid ? valid + a : nullptr;
        // the fused step (measured against the seven-launch step: 38 vs 54 us at
        // 4,096 records, 50 vs 64 at 8,192, 67 vs 78 at 16,384, 105 vs 111 at 32,768)
        const bool fused = c->fused_train && b - a <= kFusedStepRecords && nrc_fused_eligible(c, (uint32_t)(b - a));
        if (fused ? nrc_train_fused(c, pos + 3 * a, nrm + 3 * a, alb + 3 * a, target + 3 * a, v, (uint32_t)(b - a), s)
                  : nrc_train_one(c, pos + 3 * a, nrm + 3 * a, alb + 3 * a, target + 3 * a, v, (uint32_t)(b - a), s))
            return -1;
    }
    return 0;
}

int nrc_train_one(NanNrc *c, const float *pos, const float *nrm, const float *alb, const float *target,
                  const uint32_t *valid, uint32_t n, hipStream_t s) {
    if (nrc_reserve_train(c, n, s)) return -1;
    forward_backward(c, pos, nrm, alb, target, valid, n, s);
    hipLaunchKernelGGL(nrc_encode_bwd, dim3(grid_for((uint64_t)n * 8, 256), c->cfg.levels), dim3(256), 0, s,
                       grid_of(c), pos, n, c->denc, c->tg);
    uint32_t reduce_blocks = grid_for(kNrcWeights, kReduceWeights);
    const NanNrcConfig &cf = c->cfg;
    AdamArgs a;
    a.lr = cf.learning_rate;
    a.b1 = cf.beta1;
    a.b2 = cf.beta2;
    a.eps = cf.epsilon;
    hipLaunchKernelGGL(nrc_adam_mlp, dim3(reduce_blocks), dim3(256), 0, s, c->w32, c->wm, c->wv, c->wpart,
                       bwd_waves(groups_for(n)), c->w16t, c->w16n, c->stats, a, (AdamState *)c->adam);
    uint64_t entries = c->table_params / 2;
    hipLaunchKernelGGL(nrc_adam_table, dim3(grid_for(entries, 256)), dim3(256), 0, s, c->ts, c->tg,
                       (__half2 *)c->t16, entries, c->stats, a, (AdamState *)c->adam,
                       (int)cf.dense_table_adam);
    return check_launch("nrc train");
}

}  // namespace nanx

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace {

void to_half(const float *src, size_t n, std::vector<_Float16> &dst) {
    dst.resize(n);
    for (size_t i = 0; i < n; ++i) dst[i] = (_Float16)src[i];
}

// fp16 copies (transposed [out][in] and plain [in][out]) of fp32 weights
// stored per layer as [out][in]
void weight_copies(const float *w, std::vector<_Float16> &wt, std::vector<_Float16> &wn) {
    to_half(w, kNrcWeights, wt);
    wn.assign(kNrcWeights, (_Float16)0.0f);
    for (int l = 0; l < kNrcLayers; ++l) {
        int in = nrc_layer_in(l), out = nrc_layer_out(l), off = nrc_layer_offset(l);
        for (int o = 0; o < out; ++o)
            for (int i = 0; i < in; ++i) wn[off + i * out + o] = wt[off + o * in + i];
    }
}


// Host <-> table master weights (table_params floats, entry-major pairs)
// through a temporary device buffer: contiguous copies, then a scatter into
// or a gather out of the state records.
int table_weights_in(NanNrc *c, const float *host) {
    uint64_t entries = c->table_params / 2;
    float2 *tmp = nullptr;
    HIP_TRY(hipMalloc((void **)&tmp, entries * sizeof(float2)), nullptr, 0);
    hipError_t e = hipMemcpy(tmp, host, entries * sizeof(float2), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(nrc_table_scatter, dim3(grid_for(entries, 256)), dim3(256), 0, 0, c->ts, tmp, entries);
        e = hipDeviceSynchronize();
    }
    hipFree(tmp);
    HIP_TRY(e, nullptr, 0);
    return 0;
}

int table_weights_out(const NanNrc *c, float *host) {
    uint64_t entries = c->table_params / 2;
    float2 *tmp = nullptr;
    HIP_TRY(hipMalloc((void **)&tmp, entries * sizeof(float2)), nullptr, 0);
    hipLaunchKernelGGL(nrc_table_gather, dim3(grid_for(entries, 256)), dim3(256), 0, 0, c->ts, tmp, entries);
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(host, tmp, entries * sizeof(float2), hipMemcpyDeviceToHost);
    hipFree(tmp);
    HIP_TRY(e, nullptr, 0);
    return 0;
}

}  // namespace

extern "C" {

NAN_API int nan_nrc_create(const NanNrcConfig *cfg, NanNrc **out, char *err, size_t cap) {
    if (!cfg || !out) return fail(err, cap, "nan_nrc_create: null argument");
    *out = nullptr;
    if (cfg->levels == 0 || cfg->levels > (uint32_t)kNrcMaxLevels || cfg->features_per_level != 2)
        return fail(err, cap, "nan_nrc_create: levels must be 1..16 with 2 features per level");
    if (cfg->log2_table_size < 8 || cfg->log2_table_size > 24)
        return fail(err, cap, "nan_nrc_create: log2_table_size must be 8..24");
    if (cfg->width != (uint32_t)kNrcWidth || cfg->hidden_layers != (uint32_t)kNrcHidden)
        return fail(err, cap, "nan_nrc_create: the fused MLP has %d hidden layers of width %d", kNrcHidden,
                    kNrcWidth);
    if (!(cfg->base_resolution >= 1.0f) || !(cfg->max_resolution >= cfg->base_resolution) ||
        !(cfg->max_resolution <= 65536.0f))
        return fail(err, cap, "nan_nrc_create: need 1 <= base_resolution <= max_resolution <= 65536");
    for (int a = 0; a < 3; ++a)
        if (!(cfg->bbox_max[a] > cfg->bbox_min[a])) return fail(err, cap, "nan_nrc_create: empty bounding box");
    if (cfg->minibatches == 0 || cfg->minibatches > 64)
        return fail(err, cap, "nan_nrc_create: minibatches must be 1..64");
    if (!(cfg->learning_rate > 0.0f) || !(cfg->beta1 >= 0.0f && cfg->beta1 < 1.0f) ||
        !(cfg->beta2 >= 0.0f && cfg->beta2 < 1.0f) || !(cfg->epsilon > 0.0f))
        return fail(err, cap, "nan_nrc_create: invalid Adam parameters");
    NanNrc *c = new NanNrc();
    c->cfg = *cfg;
    c->table_entries = 1u << cfg->log2_table_size;
    c->table_params = (uint64_t)cfg->levels * c->table_entries * 2;
    double b = cfg->levels > 1 ? std::exp((std::log((double)cfg->max_resolution) -
                                           std::log((double)cfg->base_resolution)) / (double)(cfg->levels - 1))
                               : 1.0;
    for (uint32_t l = 0; l < (uint32_t)kNrcMaxLevels; ++l) {
        double r = l < cfg->levels ? std::floor((double)cfg->base_resolution * std::pow(b, (double)l)) : 1.0;
        c->level_res[l] = (uint32_t)std::max(1.0, r);
        double lattice = std::pow((double)c->level_res[l] + 1.0, 3.0);
        c->level_dense[l] = lattice <= (double)c->table_entries ? 1u : 0u;
    }
    // initialisation: table U(-1e-4, 1e-4); weights U(+-sqrt(6 / (fan_in + fan_out)))
    std::mt19937 gen(cfg->seed);
    std::vector<float> w(kNrcWeights), t(c->table_params);
    for (int l = 0; l < kNrcLayers; ++l) {
        float lim = std::sqrt(6.0f / (float)(nrc_layer_in(l) + nrc_layer_out(l)));
        std::uniform_real_distribution<float> u(-lim, lim);
        for (int i = 0; i < nrc_layer_in(l) * nrc_layer_out(l); ++i) w[nrc_layer_offset(l) + i] = u(gen);
    }
    std::uniform_real_distribution<float> ut(-1e-4f, 1e-4f);
    for (auto &x : t) x = ut(gen);
    std::vector<_Float16> wt16, wn16, t16;
    weight_copies(w.data(), wt16, wn16);
    to_half(t.data(), t.size(), t16);
    std::vector<float> zw(kNrcWeights, 0.0f), stats(4, 0.0f);
    bool ok = upload(&c->w32, w.data(), w.size() * 4, err, cap) == 0 &&
              upload(&c->w16t, wt16.data(), wt16.size() * 2, err, cap) == 0 &&
              upload(&c->w16n, wn16.data(), wn16.size() * 2, err, cap) == 0 &&
              upload(&c->wm, zw.data(), zw.size() * 4, err, cap) == 0 &&
              upload(&c->wv, zw.data(), zw.size() * 4, err, cap) == 0 &&
              upload(&c->wpart, nullptr, (size_t)kNrcBwdWaves * kNrcWeights * 4, err, cap) == 0 &&
              upload(&c->ts, nullptr, t.size() / 2 * sizeof(TableState), err, cap) == 0 &&
              hipMemset(c->ts, 0, t.size() / 2 * sizeof(TableState)) == hipSuccess &&
              table_weights_in(c, t.data()) == 0 &&
              upload(&c->t16, t16.data(), t16.size() * 2, err, cap) == 0 &&
              upload(&c->tg, nullptr, t.size() * 2, err, cap) == 0 &&
              upload(&c->stats, stats.data(), stats.size() * 4, err, cap) == 0 &&
              upload(&c->lossp, nullptr, (size_t)kNrcBwdWaves * sizeof(float2), err, cap) == 0 &&
              upload(&c->adam, zw.data(), 16, err, cap) == 0 &&
              hipMemset(c->tg, 0, t.size() * 2) == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
              nrc_reserve_train(c, 65536, nullptr) == 0;
    if (!ok) {
        if (err && cap && !err[0]) fail(err, cap, "nan_nrc_create: device allocation failed");
        delete c;
        return -1;
    }
    *out = c;
    return 0;
}

NAN_API void nan_nrc_destroy(NanNrc *nrc) { delete nrc; }

NAN_API int nan_nrc_param_counts(const NanNrc *nrc, uint64_t *mlp_weights, uint64_t *table_params) {
    if (!nrc) return fail(nullptr, 0, "nan_nrc_param_counts: null cache");
    if (mlp_weights) *mlp_weights = kNrcWeights;
    if (table_params) *table_params = nrc->table_params;
    return 0;
}

NAN_API int nan_nrc_get_params(const NanNrc *nrc, float *weights, float *table) {
    if (!nrc) return fail(nullptr, 0, "nan_nrc_get_params: null cache");
    HIP_TRY(hipDeviceSynchronize(), nullptr, 0);
    if (weights) HIP_TRY(hipMemcpy(weights, nrc->w32, (size_t)kNrcWeights * 4, hipMemcpyDeviceToHost), nullptr, 0);
    if (table && table_weights_out(nrc, table)) return -1;
    return 0;
}

NAN_API int nan_nrc_set_params(NanNrc *nrc, const float *weights, const float *table) {
    if (!nrc) return fail(nullptr, 0, "nan_nrc_set_params: null cache");
    HIP_TRY(hipDeviceSynchronize(), nullptr, 0);
    if (weights) {
        std::vector<_Float16> wt, wn;
        weight_copies(weights, wt, wn);
        HIP_TRY(hipMemcpy(nrc->w32, weights, (size_t)kNrcWeights * 4, hipMemcpyHostToDevice), nullptr, 0);
        HIP_TRY(hipMemcpy(nrc->w16t, wt.data(), wt.size() * 2, hipMemcpyHostToDevice), nullptr, 0);
        HIP_TRY(hipMemcpy(nrc->w16n, wn.data(), wn.size() * 2, hipMemcpyHostToDevice), nullptr, 0);
    }
    if (table) {
        std::vector<_Float16> h;
        to_half(table, nrc->table_params, h);
        if (table_weights_in(nrc, table)) return -1;
        HIP_TRY(hipMemcpy(nrc->t16, h.data(), h.size() * 2, hipMemcpyHostToDevice), nullptr, 0);
    }
    return 0;
}

NAN_API int nan_nrc_infer(NanNrc *nrc, const float *pos, const float *nrm, const float *alb, float *out, uint32_t n,
                          void *stream) {
    if (!nrc || (n && (!pos || !nrm || !alb || !out))) return fail(nullptr, 0, "nan_nrc_infer: null argument");
    if (n == 0) return 0;
    hipStream_t s = as_stream(stream);
    if (nrc_reserve(nrc, n, s) || nrc_infer_into(nrc, pos, nrm, alb, n, nrc->enc, nrc->cap_groups, nrc->pred, out, s))
        return -1;
    return check_launch("nrc infer");
}

NAN_API int nan_nrc_forward_raw(NanNrc *nrc, const float *pos, const float *nrm, const float *alb, float *out,
                                uint32_t n, void *stream) {
    if (!nrc || (n && (!pos || !nrm || !alb || !out))) return fail(nullptr, 0, "nan_nrc_forward_raw: null argument");
    if (n == 0) return 0;
    hipStream_t s = as_stream(stream);
    if (nrc_predict(nrc, pos, nrm, alb, n, nrc->pred, s)) return -1;
    hipLaunchKernelGGL(nrc_rgb, dim3(grid_for(n, 256)), dim3(256), 0, s, nrc->pred, n, out, 1);
    return check_launch("nrc forward");
}

NAN_API int nan_nrc_set_fused_training(NanNrc *nrc, int fused) {
    if (!nrc) return fail(nullptr, 0, "nan_nrc_set_fused_training: null cache");
    nrc->fused_train = fused != 0;
    return 0;
}

NAN_API int nan_nrc_stats(NanNrc *nrc, float *out) {
    if (!nrc || !out) return fail(nullptr, 0, "nan_nrc_stats: null argument");
    float st[4];
    HIP_TRY(hipDeviceSynchronize(), nullptr, 0);
    HIP_TRY(hipMemcpy(st, nrc->stats, sizeof(st), hipMemcpyDeviceToHost), nullptr, 0);
    out[0] = st[0] / (3.0f * std::max(st[1], 1.0f));
    out[1] = st[2] > 0.0f ? st[2] : st[1];
    return 0;
}

NAN_API int nan_nrc_time_forward(NanNrc *nrc, const float *pos, const float *nrm, const float *alb, uint32_t n,
                                 uint32_t repeats, float *ms_out, void *stream) {
    if (!nrc || !ms_out || (n && (!pos || !nrm || !alb))) return fail(nullptr, 0, "nan_nrc_time_forward: null argument");
    *ms_out = 0.0f;
    if (n == 0 || repeats == 0) return 0;
    hipStream_t s = as_stream(stream);
    if (nrc_reserve(nrc, n, s)) return -1;
    uint32_t groups = groups_for(n);
    encode(nrc, pos, nrm, alb, n, nullptr, s);
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a), nullptr, 0);
    if (hipEventCreate(&b) != hipSuccess) {
        hipEventDestroy(a);
        return fail(nullptr, 0, "nan_nrc_time_forward: hipEventCreate failed");
    }
    hipEventRecord(a, s);
    for (uint32_t r = 0; r < repeats; ++r)
        hipLaunchKernelGGL((nrc_mlp_fwd<false>), dim3(fwd_blocks(groups)), dim3(kFwdBlock), 0, s, nrc->w16t,
                           nrc->enc, groups, n, nrc->pred, (_Float16 *)nullptr, nrc->cap_groups, (float *)nullptr);
    hipEventRecord(b, s);
    hipError_t e = hipEventSynchronize(b);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (e != hipSuccess) return fail(nullptr, 0, "nan_nrc_time_forward: %s", hipGetErrorString(e));
    *ms_out = ms / (float)repeats;
    return check_launch("nrc forward timing");
}

NAN_API int nan_nrc_encode(NanNrc *nrc, const float *pos, const float *nrm, const float *alb, float *out, uint32_t n,
                           void *stream) {
    if (!nrc || (n && (!pos || !nrm || !alb || !out))) return fail(nullptr, 0, "nan_nrc_encode: null argument");
    if (n == 0) return 0;
    if (nrc_reserve(nrc, n, as_stream(stream))) return -1;
    encode(nrc, pos, nrm, alb, n, out, as_stream(stream));
    return check_launch("nrc encode");
}

NAN_API int nan_nrc_train(NanNrc *nrc, const float *pos, const float *nrm, const float *alb, const float *target,
                          uint32_t n, float *loss_out, void *stream) {
    if (!nrc || (n && (!pos || !nrm || !alb || !target))) return fail(nullptr, 0, "nan_nrc_train: null argument");
    if (n == 0) return 0;
    hipStream_t s = as_stream(stream);
    if (nrc_train_step(nrc, pos, nrm, alb, target, nullptr, n, s)) return -1;
    if (loss_out) hipLaunchKernelGGL(nrc_add_loss, dim3(1), dim3(64), 0, s, nrc->stats, loss_out);
    return check_launch("nrc train");
}

NAN_API int nan_nrc_table_grad(NanNrc *nrc, const float *pos, const float *nrm, const float *alb,
                               const float *target, uint32_t n, float *table_grad, void *stream) {
    if (!nrc || (n && (!pos || !nrm || !alb || !target || !table_grad)))
        return fail(nullptr, 0, "nan_nrc_table_grad: null argument");
    if (n == 0) return 0;
    hipStream_t s = as_stream(stream);
    if (nrc_reserve_train(nrc, n, s)) return -1;
    forward_backward(nrc, pos, nrm, alb, target, nullptr, n, s);
    hipLaunchKernelGGL(nrc_encode_bwd, dim3(grid_for((uint64_t)n * 8, 256), nrc->cfg.levels), dim3(256), 0, s,
                       grid_of(nrc), pos, n, nrc->denc, nrc->tg);
    uint64_t entries = nrc->table_params / 2;
    hipLaunchKernelGGL(nrc_table_grad_out, dim3(grid_for(entries, 256)), dim3(256), 0, s, nrc->tg, entries,
                       nrc->stats, (float2 *)table_grad);
    return check_launch("nrc table grad");
}

NAN_API int nan_nrc_backward_raw(NanNrc *nrc, const float *pos, const float *nrm, const float *alb,
                                 const float *target, uint32_t n, float *weight_grad, float *input_grad,
                                 void *stream) {
    if (!nrc || (n && (!pos || !nrm || !alb || !target || !weight_grad || !input_grad)))
        return fail(nullptr, 0, "nan_nrc_backward_raw: null argument");
    if (n == 0) return 0;
    hipStream_t s = as_stream(stream);
    if (nrc->fused_train && nrc_fused_eligible(nrc, n))
        return nrc_fused_grads(nrc, pos, nrm, alb, target, n, weight_grad, input_grad, s);
    if (nrc_reserve_train(nrc, n, s)) return -1;
    forward_backward(nrc, pos, nrm, alb, target, nullptr, n, s);
    hipLaunchKernelGGL(nrc_sum_partials, dim3(grid_for(kNrcWeights, kReduceWeights)), dim3(256), 0, s, nrc->wpart,
                       bwd_waves(groups_for(n)), nrc->stats, weight_grad);
    hipLaunchKernelGGL(nrc_scale_rows, dim3(grid_for((uint64_t)n * kNrcIn, 256)), dim3(256), 0, s, nrc->denc, n,
                       nrc->stats, input_grad);
    return check_launch("nrc backward");
}

}  // extern "C"
