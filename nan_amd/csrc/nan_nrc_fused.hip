// Fused Adam step of the learned radiance cache, for minibatches of up to
// kNrcBwdWaves * 64 records, on dense and hashed grid levels alike.  Three
// launches per step instead of seven:
//
//   nrc_fused_fb          four waves per 64-record group (four 16-record
//                         tiles), each owning one 16-unit block of every
//                         layer: hash-grid encoding, the 48 -> 64x4 -> 16
//                         ReLU MLP forward (the layers' inputs kept in LDS for
//                         the backward), the relative L2 loss, and the
//                         backward.  The group's weight gradients over its 64
//                         records go to its own partial in MFMA fragment order
//                         (no cross-group reduction); each record's gradient
//                         with respect to its hash features goes to a
//                         per-level array.
//   nrc_fused_update      the weights' Adam on the fixed-order sum of the
//                         partials (sixteen threads per element) and the
//                         step's loss statistics, beside the table-gradient
//                         scatter: one thread per (record, level, corner), one
//                         packed FP16 atomic each.
//   nrc_fused_table_adam  the table's lazy Adam over each dense level's
//                         lattice and each hashed level's whole table.
//
// The step computes what nrc_train_one does: the same loss, gradient clamp,
// rounding points (activations and dY in FP16, FP32 accumulation), FP16
// table-gradient sums and Adam with its step count on the device; the MLP
// gradients are summed in a different fixed order.
//
// Every kernel keeps its LDS within 64 KB per workgroup: a first form of the
// forward-backward kernel with 93 KB (four waves' gradients staged at once)
// computed wrong forward values at -O3 on gfx950 although the device reports
// 160 KB per block.  LDS float atomics were measured too: about 200 cycles
// per ds_add_f32 wave instruction, which made summing table gradients per
// z-slab in LDS slower than the global packed atomics used here.  The
// atomics run from their own grid of one thread per (record, corner): issued
// from the forward-backward kernel's few waves they reached about 13 G/s,
// against about 30 G/s spread over many threads.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>

#include "nan_nrc_internal.h"

using namespace nanx;

namespace {

constexpr int kX0Stride = kNrcIn + 8;  // LDS row (halves) of the encoded input
constexpr int kTS = 64 + 8;            // LDS row (halves) of the backward transposes
constexpr int kUpdBlock = 1024;
constexpr int kMlpSplit = 16;  // threads summing the partials of one weight
constexpr int kMlpPerBlock = kUpdBlock / kMlpSplit;

struct FbArgs {
    NrcGrid E;
    const _Float16 *w16t, *w16n;  // per layer [out][in] and [in][out]
    const float *pos, *nrm, *alb, *target;
    const uint32_t *valid;  // nullable
    uint32_t n, groups, cap;
    float *partial;    // [gridDim.x][kNrcWeights], MFMA fragment order
    float2 *gtab;      // [level][cap]: each record's gradient wrt its two features of the level; null: none
    float2 *lossp;     // [gridDim.x]: loss sum, valid records
    AdamState *as;     // null: no Adam bookkeeping
    float b1, b2;
    float *input_grad;  // nullable: every record's gradient wrt its 48 inputs, times input_scale, [n][48]
    float input_scale;
};

struct UpdArgs {
    NrcGrid E;
    const float *pos;
    const float2 *gtab;
    uint32_t cap, scatter_blocks;
    uint32_t n, parts;
    const float *partial;
    float *w32, *wm, *wv;
    _Float16 *w16t, *w16n;
    TableState *ts;
    __half2 *tgrad, *t16;
    uint32_t entries, levels, lattice_total;
    uint32_t lattice_begin[kNrcMaxLevels + 1];  // prefix of the levels' spans: (res+1)^3 dense, entries hashed
    const float2 *lossp;
    float *stats;
    AdamState *as;
    float lr, b1, b2, eps;
    uint32_t dense_adam, mlp_blocks;
};

constexpr int kFbWaves = 4;  // waves per 64-record group of the forward-backward kernel

// ReLU after the rounding to FP16 (rounding is monotone and keeps zero).
__device__ __forceinline__ half4_t relu_h(const floatx4_t &acc) {
    const half4_t zero = {(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f};
    return __builtin_elementwise_max(__builtin_convertvector(acc, half4_t), zero);
}

// dY of the layer below from W dY^T: kept where the layer's input (a ReLU
// output) is positive, rounded to FP16.
__device__ __forceinline__ half4_t relu_mask(const half4_t &xin, const floatx4_t &dx) {
    half4_t o;
#pragma unroll
    for (int i = 0; i < 4; ++i) o[i] = (_Float16)((float)xin[i] > 0.0f ? dx[i] : 0.0f);
    return o;
}

// One 16x16 block of a weight gradient over the group's 64 records, dW^T =
// dY^T X with the records as the MFMA reduction dimension: rows ot of the
// transposed dY (ty) against rows it of the transposed X (tx).
__device__ __forceinline__ floatx4_t dw_block(const _Float16 *ty, const _Float16 *tx, int ot, int it, int r, int q) {
    floatx4_t d = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
        d = mfma16(*(const half4_t *)&ty[(ot * 16 + r) * kTS + ks * 16 + 4 * q],
                   *(const half4_t *)&tx[(it * 16 + r) * kTS + ks * 16 + 4 * q], d);
    return d;
}

// The 16 rows (units 16b..16b+15) of a transposed operand, written from the
// four tiles' fragments of block b (lane%16 = record of the tile, 4*(lane/16)+i
// = unit).
__device__ __forceinline__ void put_t(_Float16 *t, int b, const half4_t (&v)[4], int r, int q) {
#pragma unroll
    for (int tt = 0; tt < 4; ++tt)
#pragma unroll
        for (int i = 0; i < 4; ++i) t[(b * 16 + 4 * q + i) * kTS + tt * 16 + r] = v[tt][i];
}

// FULL: also the gradient with respect to all 48 inputs (a test hook).
//
// Four waves share a 64-record group, each owning one 16-unit block: in the
// forward, wave w computes units 16w..16w+15 of every hidden layer for all
// four 16-record tiles and the output layer for tile w; the layers'
// activations are exchanged through LDS (the same copy the backward reads).
// In the backward, wave w computes the weight-gradient rows of output block w
// and the input gradient of input block w; the dY blocks are exchanged
// through LDS in fragment order, and transposed through LDS for the weight
// gradients.  A single wave per group spent most of its time issuing one
// serial instruction stream while three of the CU's four SIMDs idled.  The
// partials (64 KB per group, read once by the update launch on any XCD) are
// stored non-temporally, so they do not sit dirty in L2 for the
// end-of-kernel write-back (16.3 against 18.4 us per 16,384-record step).
template <bool FULL>
__global__ __launch_bounds__(64 * kFbWaves) void nrc_fused_fb(FbArgs A) {
    // u: the encoded input rows (record-major) until every wave holds its
    // fragments, then the backward's dY exchange [tile][block][lane]
    __shared__ __align__(16) half4_t u[4][4][64];
    __shared__ __align__(16) _Float16 tx[64 * kTS], ty[64 * kTS];
    __shared__ half4_t sv[kNrcHidden][4][4][64];  // inputs of layers 1..4, fragment order
    __shared__ float2 lred[kFbWaves];
    static_assert(64 * kX0Stride * sizeof(_Float16) <= sizeof(u), "encoded rows must fit the exchange buffer");
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r = lane & 15, q = lane >> 4;
    _Float16 *x0s = reinterpret_cast<_Float16 *>(&u[0][0][0]);
    const uint32_t g = blockIdx.x;  // one 64-record group per block
    if (A.as && g == 0 && threadIdx.x == 0) {  // this step's bias corrections
        const uint32_t step = A.as->step + 1u;
        A.as->next_step = step;
        A.as->c1 = bias_correction(A.b1, step);
        A.as->c2 = bias_correction(A.b2, step);
    }
    float *part = A.partial + (size_t)g * kNrcWeights;
    constexpr int NI0 = FULL ? 3 : 2;  // input blocks of layer 0 whose gradient is needed
    // This wave's weight fragments, requested before the encoding's gathers:
    // forward, output block w of each hidden layer ([out][in]) and all of the
    // output layer; backward, input block w of each layer ([in][out]).
    half4_t wf0[3], wf1[4], wf2[4], wf3[4], wf4[4], wb1[4], wb2[4], wb3[4], wb0[4], wb4;
    {
        const _Float16 *wt = A.w16t, *wn = A.w16n;
#pragma unroll
        for (int kb = 0; kb < 3; ++kb) wf0[kb] = *(const half4_t *)&wt[(w * 16 + r) * kNrcIn + kb * 16 + 4 * q];
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            wf1[kb] = *(const half4_t *)&wt[nrc_layer_offset(1) + (w * 16 + r) * kNrcWidth + kb * 16 + 4 * q];
            wf2[kb] = *(const half4_t *)&wt[nrc_layer_offset(2) + (w * 16 + r) * kNrcWidth + kb * 16 + 4 * q];
            wf3[kb] = *(const half4_t *)&wt[nrc_layer_offset(3) + (w * 16 + r) * kNrcWidth + kb * 16 + 4 * q];
            wf4[kb] = *(const half4_t *)&wt[nrc_layer_offset(4) + r * kNrcWidth + kb * 16 + 4 * q];
            wb1[kb] = *(const half4_t *)&wn[nrc_layer_offset(1) + (w * 16 + r) * kNrcWidth + kb * 16 + 4 * q];
            wb2[kb] = *(const half4_t *)&wn[nrc_layer_offset(2) + (w * 16 + r) * kNrcWidth + kb * 16 + 4 * q];
            wb3[kb] = *(const half4_t *)&wn[nrc_layer_offset(3) + (w * 16 + r) * kNrcWidth + kb * 16 + 4 * q];
            // layer 0 has 48 inputs: blocks 0..2 only
            wb0[kb] = w < NI0 ? *(const half4_t *)&wn[(w * 16 + r) * kNrcWidth + kb * 16 + 4 * q] : half4_t{};
        }
        wb4 = *(const half4_t *)&wn[nrc_layer_offset(4) + (w * 16 + r) * kNrcOut + 4 * q];
    }
    // the targets of tile w, read by the loss
    const uint32_t sw = g * 64u + (uint32_t)w * 16u + (uint32_t)r;
    const bool counted = q == 0 && sw < A.n && (!A.valid || A.valid[sw]);
    float tgt[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) tgt[c] = counted ? A.target[3 * (size_t)sw + c] : 0.0f;
    {  // encoding: lane = record, wave w the levels w, w + 4, ... (dense or
       // hashed corner indexing per level); lanes past n read the cell at the
       // box corner and store zeros.
        const uint32_t s = g * 64u + (uint32_t)lane;
        const bool live = s < A.n;
        _Float16 *row = x0s + lane * kX0Stride;
        float p[3] = {0.0f, 0.0f, 0.0f};
        if (live)
            for (int a = 0; a < 3; ++a) p[a] = A.pos[3 * (size_t)s + a];
        const uint32_t levels = A.E.levels;
        for (uint32_t l = (uint32_t)w; l < (uint32_t)kNrcMaxLevels; l += kFbWaves) {
            half2_t val = {(_Float16)0.0f, (_Float16)0.0f};
            if (l < levels) {
                uint32_t base[3];
                float f[3];
                level_cell(A.E, l, p, base, f);
                const half2_t *tab = reinterpret_cast<const half2_t *>(A.E.table) + (size_t)l * A.E.entries;
                half2_t v[8];
                float wt[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    uint32_t idx;
                    cell_corner(A.E, l, base, f, (uint32_t)k, idx, wt[k]);
                    v[k] = tab[idx];
                }
                float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    f0 += wt[k] * (float)v[k][0];
                    f1 += wt[k] * (float)v[k][1];
                }
                if (live) val = half2_t{(_Float16)f0, (_Float16)f1};
            }
            *reinterpret_cast<half2_t *>(&row[2 * l]) = val;
        }
        if (w == kFbWaves - 1) {
            for (int a = 0; a < 3; ++a) {
                row[32 + a] = (_Float16)(live ? A.nrm[3 * (size_t)s + a] : 0.0f);
                row[35 + a] = (_Float16)(live ? A.alb[3 * (size_t)s + a] : 0.0f);
            }
            for (int k = 38; k < kNrcIn; ++k) row[k] = (_Float16)0.0f;
        }
    }
    __syncthreads();
    half4_t x0[4][3];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int kb = 0; kb < 3; ++kb) x0[t][kb] = *(const half4_t *)&x0s[(t * 16 + r) * kX0Stride + kb * 16 + 4 * q];
    // forward: wave w's block of each hidden layer for the four tiles
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        floatx4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mfma32(wf0[0], wf0[1], x0[t][0], x0[t][1], acc);
        acc = mfma16(wf0[2], x0[t][2], acc);
        sv[0][t][w][lane] = relu_h(acc);
    }
    __syncthreads();  // also: every wave holds its x0 fragments, so u is free
#pragma unroll
    for (int l = 1; l < kNrcHidden; ++l) {
        const half4_t(&wl)[4] = l == 1 ? wf1 : (l == 2 ? wf2 : wf3);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            half4_t x[4];
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) x[kb] = sv[l - 1][t][kb][lane];
            floatx4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mfma32(wl[0], wl[1], x[0], x[1], acc);
            acc = mfma32(wl[2], wl[3], x[2], x[3], acc);
            sv[l][t][w][lane] = relu_h(acc);
        }
        __syncthreads();
    }
    // output layer and loss for tile w: lanes of quarter 0 hold outputs 0..3
    // of record r
    float loss = 0.0f, nvalid = 0.0f;
    half4_t dyw[4];  // this wave's dY block of the layer being differentiated, per tile
    half4_t dy4 = {(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f};
    {
        half4_t x[4];
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) x[kb] = sv[3][w][kb][lane];
        floatx4_t out = {0.0f, 0.0f, 0.0f, 0.0f};
        out = mfma32(wf4[0], wf4[1], x[0], x[1], out);
        out = mfma32(wf4[2], wf4[3], x[2], x[3], out);
        if (counted) {
            float gr[3];
            for (int c = 0; c < 3; ++c) {
                const float pc = out[c], tc = tgt[c];
                const float d = pc * pc + 0.01f, diff = pc - tc;
                loss += diff * diff / d;
                gr[c] = fminf(fmaxf(2.0f * diff / d, -kGradClamp), kGradClamp);
            }
            dy4 = half4_t{(_Float16)gr[0], (_Float16)gr[1], (_Float16)gr[2], (_Float16)0.0f};
            nvalid = 1.0f;
        }
    }
    // backward of the output layer (16 outputs): dY of tile w transposed into
    // ty rows 0..15, block w of its input (sv[3]) transposed into tx, dY
    // fragments exchanged through u
    half4_t xin[4];
    {
#pragma unroll
        for (int i = 0; i < 4; ++i) ty[(4 * q + i) * kTS + w * 16 + r] = dy4[i];
#pragma unroll
        for (int t = 0; t < 4; ++t) xin[t] = sv[3][t][w][lane];
        put_t(tx, w, xin, r, q);
        u[w][0][lane] = dy4;
        __syncthreads();
        float4 *p = reinterpret_cast<float4 *>(part + nrc_layer_offset(4)) + lane;
        const floatx4_t d = dw_block(ty, tx, 0, w, r, q);
        __builtin_nontemporal_store(d, reinterpret_cast<floatx4_t *>(&p[w * 64]));
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            floatx4_t dx = {0.0f, 0.0f, 0.0f, 0.0f};
            dx = mfma16(wb4, u[t][0][lane], dx);
            dyw[t] = relu_mask(xin[t], dx);
        }
        __syncthreads();
    }
    // hidden layers 3, 2, 1
#pragma unroll
    for (int l = kNrcHidden - 1; l >= 1; --l) {
        const half4_t(&wl)[4] = l == 1 ? wb1 : (l == 2 ? wb2 : wb3);
        put_t(ty, w, dyw, r, q);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            u[t][w][lane] = dyw[t];
            xin[t] = sv[l - 1][t][w][lane];
        }
        put_t(tx, w, xin, r, q);
        __syncthreads();
        float4 *p = reinterpret_cast<float4 *>(part + nrc_layer_offset(l)) + lane;
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const floatx4_t d = dw_block(ty, tx, w, it, r, q);
            __builtin_nontemporal_store(d, reinterpret_cast<floatx4_t *>(&p[(w * 4 + it) * 64]));
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            floatx4_t dx = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ko = 0; ko < 4; ++ko) dx = mfma16(wl[ko], u[t][ko][lane], dx);
            dyw[t] = relu_mask(xin[t], dx);
        }
        __syncthreads();
    }
    // layer 0 (48 inputs): weight-gradient rows of output block w, and the
    // gradient of input block w < NI0 (the hash features; FULL: all inputs)
    put_t(ty, w, dyw, r, q);
#pragma unroll
    for (int t = 0; t < 4; ++t) u[t][w][lane] = dyw[t];
    if (w < 3) {
#pragma unroll
        for (int t = 0; t < 4; ++t) xin[t] = w == 0 ? x0[t][0] : (w == 1 ? x0[t][1] : x0[t][2]);
        put_t(tx, w, xin, r, q);
    }
    __syncthreads();
    {
        float4 *p = reinterpret_cast<float4 *>(part) + lane;
#pragma unroll
        for (int it = 0; it < 3; ++it) {
            const floatx4_t d = dw_block(ty, tx, w, it, r, q);
            __builtin_nontemporal_store(d, reinterpret_cast<floatx4_t *>(&p[(w * 3 + it) * 64]));
        }
    }
    if (w < NI0) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            floatx4_t dx = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ko = 0; ko < 4; ++ko) dx = mfma16(wb0[ko], u[t][ko][lane], dx);
            const uint32_t s = g * 64u + (uint32_t)t * 16u + (uint32_t)r;
            if (s >= A.n) continue;
            if (FULL && A.input_grad)
                *reinterpret_cast<float4 *>(&A.input_grad[(size_t)s * kNrcIn + w * 16 + 4 * q]) =
                    make_float4(dx[0] * A.input_scale, dx[1] * A.input_scale, dx[2] * A.input_scale,
                                dx[3] * A.input_scale);
            if (!A.gtab || w >= 2) continue;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t lv = (uint32_t)(w * 8 + 2 * q + h);
                if (lv < A.E.levels) A.gtab[(size_t)lv * A.cap + s] = make_float2(dx[2 * h], dx[2 * h + 1]);
            }
        }
    }
    // the group's loss: each wave's lanes, then the four waves in order
    for (int off = 32; off > 0; off >>= 1) {
        loss += __shfl_down(loss, off);
        nvalid += __shfl_down(nvalid, off);
    }
    if (lane == 0) lred[w] = make_float2(loss, nvalid);
    __syncthreads();
    if (threadIdx.x == 0)
        A.lossp[g] = make_float2(((lred[0].x + lred[1].x) + lred[2].x) + lred[3].x,
                                 ((lred[0].y + lred[1].y) + lred[2].y) + lred[3].y);
}

// Weight (index into [layer][out][in] order) of element f of a partial in
// fragment order: per layer, blocks (ot, it) of 64 lanes x 4, lane r + 16q,
// item j holding dW^T[ot*16 + 4q + j][it*16 + r].
__device__ __forceinline__ uint32_t frag_weight(uint32_t f) {
    int l = 0;
    while (l + 1 < kNrcLayers && (int)f >= nrc_layer_offset(l + 1)) ++l;
    const uint32_t in = (uint32_t)nrc_layer_in(l), ki = in / 16u, rel = f - (uint32_t)nrc_layer_offset(l);
    const uint32_t blk = rel / 256u, lane = (rel / 4u) % 64u, j = rel % 4u;
    const uint32_t ot = blk / ki, it = blk % ki, r = lane % 16u, q = lane / 16u;
    return (uint32_t)nrc_layer_offset(l) + (ot * 16u + 4u * q + j) * in + it * 16u + r;
}

__device__ __forceinline__ void adam2(float (&w)[2], float (&m)[2], float (&v)[2], const float (&g)[2], float lr,
                                      float b1, float b2, float eps, float c1, float c2) {
    for (int c = 0; c < 2; ++c) {
        m[c] = b1 * m[c] + (1.0f - b1) * g[c];
        v[c] = b2 * v[c] + (1.0f - b2) * g[c] * g[c];
        w[c] = w[c] - lr * (m[c] / c1) / (sqrtf(v[c] / c2) + eps);
    }
}

// Hash-table gradients: one thread per (record, corner) of one level
// (scatter_blocks blocks per level), eight neighbouring lanes the eight corners of a record, one
// packed FP16 atomic each (many threads keep the device's atomic rate busy).
__device__ __forceinline__ void scatter_corner(const UpdArgs &U, uint32_t gid, uint32_t level) {
    const uint32_t s = gid >> 3, k = gid & 7u;
    if (s >= U.n) return;
    const float2 g = U.gtab[(size_t)level * U.cap + s];
    if (g.x == 0.0f && g.y == 0.0f) return;
    uint32_t base[3], idx;
    float f[3], w;
    level_cell(U.E, level, U.pos + 3 * (size_t)s, base, f);
    cell_corner(U.E, level, base, f, k, idx, w);
    unsafeAtomicAdd(&U.tgrad[(size_t)level * U.entries + idx],
                    __halves2half2(__float2half(w * g.x * kTableGradScale), __float2half(w * g.y * kTableGradScale)));
}

// Blocks [0, mlp_blocks): the weights' Adam and the step's loss statistics;
// then scatter_blocks blocks per level: the table-gradient atomics.
__global__ __launch_bounds__(kUpdBlock) void nrc_fused_update(UpdArgs U) {
    __shared__ float red[kUpdBlock];
    const uint32_t tid = threadIdx.x;
    if (blockIdx.x >= U.mlp_blocks) {
        const uint32_t b = blockIdx.x - U.mlp_blocks;
        scatter_corner(U, (b % U.scatter_blocks) * kUpdBlock + tid, b / U.scatter_blocks);
        return;
    }
    const float scale = 1.0f / (3.0f * fmaxf((float)U.n, 1.0f));
    const float c1 = U.as->c1, c2 = U.as->c2;
    if (blockIdx.x == 0) {  // the step's loss statistics: a fixed-order tree over the per-wave sums
        __shared__ float2 lred[kNrcBwdWaves];
        for (uint32_t p = tid; p < (uint32_t)kNrcBwdWaves; p += kUpdBlock)
            lred[p] = p < U.parts ? U.lossp[p] : make_float2(0.0f, 0.0f);
        __syncthreads();
        for (uint32_t w = kNrcBwdWaves / 2; w > 0; w >>= 1) {
            if (tid < w) lred[tid] = make_float2(lred[tid].x + lred[tid + w].x, lred[tid].y + lred[tid + w].y);
            __syncthreads();
        }
        if (tid == 0) {
            U.stats[0] = lred[0].x;
            U.stats[1] = (float)U.n;
            U.stats[2] = lred[0].y;
        }
    }
    // partial element f (fragment order): thread (sub, fl) sums partials sub,
    // sub + 16, ... eight loads at a time, then four, then one; the sixteen
    // sums combine in a fixed order
    const uint32_t fl = tid % kMlpPerBlock, sub = tid / kMlpPerBlock;
    const uint32_t f = blockIdx.x * kMlpPerBlock + fl;
    float g = 0.0f;
    if (f < (uint32_t)kNrcWeights) {
        uint32_t p = sub;
        for (; p + 7 * kMlpSplit < U.parts; p += 8 * kMlpSplit) {  // eight loads in flight
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = U.partial[(size_t)(p + j * kMlpSplit) * kNrcWeights + f];
            g += ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
        }
        for (; p + 3 * kMlpSplit < U.parts; p += 4 * kMlpSplit) {
            const float a = U.partial[(size_t)p * kNrcWeights + f];
            const float b = U.partial[(size_t)(p + kMlpSplit) * kNrcWeights + f];
            const float c = U.partial[(size_t)(p + 2 * kMlpSplit) * kNrcWeights + f];
            const float d = U.partial[(size_t)(p + 3 * kMlpSplit) * kNrcWeights + f];
            g += ((a + b) + c) + d;
        }
        for (; p < U.parts; p += kMlpSplit) g += U.partial[(size_t)p * kNrcWeights + f];
    }
    red[tid] = g;
    __syncthreads();
    if (sub != 0 || f >= (uint32_t)kNrcWeights) return;
    g = 0.0f;
#pragma unroll
    for (int k = 0; k < kMlpSplit; ++k) g += red[k * kMlpPerBlock + fl];
    g *= scale;
    const uint32_t k = frag_weight(f);
    const float mi = U.b1 * U.wm[k] + (1.0f - U.b1) * g;
    const float vi = U.b2 * U.wv[k] + (1.0f - U.b2) * g * g;
    U.wm[k] = mi;
    U.wv[k] = vi;
    const float wi = U.w32[k] - U.lr * (mi / c1) / (sqrtf(vi / c2) + U.eps);
    U.w32[k] = wi;
    U.w16t[k] = (_Float16)wi;
    int l = 0;
    while (l + 1 < kNrcLayers && (int)k >= nrc_layer_offset(l + 1)) ++l;
    const int in = nrc_layer_in(l), out = nrc_layer_out(l), rel = (int)k - nrc_layer_offset(l);
    U.w16n[nrc_layer_offset(l) + (rel % in) * out + rel / in] = (_Float16)wi;
}

// The table's lazy Adam, one entry per thread: a dense level's lattice, a
// hashed level's whole table; the step's last kernel stores Adam's step
// count.
__global__ __launch_bounds__(kUpdBlock) void nrc_fused_table_adam(UpdArgs U) {
    const float scale = 1.0f / (3.0f * fmaxf((float)U.n, 1.0f));
    const float c1 = U.as->c1, c2 = U.as->c2;
    const uint32_t i = blockIdx.x * kUpdBlock + threadIdx.x;
    if (i == 0) U.as->step = U.as->next_step;
    if (i >= U.lattice_total) return;
    uint32_t l = 0;
    while (i >= U.lattice_begin[l + 1]) ++l;
    const size_t k = (size_t)l * U.entries + (i - U.lattice_begin[l]);
    const __half2 g2 = U.tgrad[k];
    float g[2] = {__low2float(g2), __high2float(g2)};
    if (g[0] == 0.0f && g[1] == 0.0f && !U.dense_adam) return;
    U.tgrad[k] = __halves2half2(__float2half(0.0f), __float2half(0.0f));
    const float sc = scale / kTableGradScale;
    for (int c = 0; c < 2; ++c) g[c] = isfinite(g[c]) ? g[c] * sc : 0.0f;
    float4 *rec = reinterpret_cast<float4 *>(U.ts + k);
    const float4 wm = rec[0], vp = rec[1];
    float w[2] = {wm.x, wm.y}, m[2] = {wm.z, wm.w}, v[2] = {vp.x, vp.y};
    adam2(w, m, v, g, U.lr, U.b1, U.b2, U.eps, c1, c2);
    rec[0] = make_float4(w[0], w[1], m[0], m[1]);
    rec[1] = make_float4(v[0], v[1], 0.0f, 0.0f);
    U.t16[k] = __halves2half2(__float2half(w[0]), __float2half(w[1]));
}

// Fixed-order sum of the per-wave weight-gradient partials, times scale.
__global__ void nrc_fused_sum(const float *partial, uint32_t parts, float scale, float *out) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= (uint32_t)kNrcWeights) return;
    float g = 0.0f;
    for (uint32_t p = 0; p < parts; ++p) g += partial[(size_t)p * kNrcWeights + f];
    out[frag_weight(f)] = g * scale;
}

FbArgs fb_args(NanNrc *c, const float *pos, const float *nrm, const float *alb, const float *target,
               const uint32_t *valid, uint32_t n) {
    FbArgs A;
    A.E = nrc_grid(c);
    A.w16t = c->w16t;
    A.w16n = c->w16n;
    A.pos = pos;
    A.nrm = nrm;
    A.alb = alb;
    A.target = target;
    A.valid = valid;
    A.n = n;
    A.groups = (n + 63u) / 64u;
    A.cap = c->cap_train_groups * 64u;
    A.partial = c->wpart;
    A.gtab = reinterpret_cast<float2 *>(c->denc);
    A.lossp = c->lossp;
    A.as = reinterpret_cast<AdamState *>(c->adam);
    A.b1 = c->cfg.beta1;
    A.b2 = c->cfg.beta2;
    A.input_grad = nullptr;
    A.input_scale = 0.0f;
    return A;
}

}  // namespace

namespace nanx {

bool nrc_fused_eligible(const NanNrc *c, uint32_t n) {
    (void)c;
    return n <= (uint32_t)kNrcBwdWaves * 64u;  // one 64-record group per block and partial
}

int nrc_train_fused(NanNrc *c, const float *pos, const float *nrm, const float *alb, const float *target,
                    const uint32_t *valid, uint32_t n, hipStream_t s) {
    if (n == 0) return 0;
    if (!nrc_fused_eligible(c, n))
        return fail(nullptr, 0, "nan_nrc: the fused training step takes at most %u records",
                    (uint32_t)kNrcBwdWaves * 64u);
    if (nrc_reserve_train(c, n, s)) return -1;
    FbArgs A = fb_args(c, pos, nrm, alb, target, valid, n);
    const uint32_t parts = A.groups;
    hipLaunchKernelGGL(nrc_fused_fb<false>, dim3(parts), dim3(64 * kFbWaves), 0, s, A);
    const NanNrcConfig &cf = c->cfg;
    UpdArgs U;
    U.E = A.E;
    U.pos = pos;
    U.gtab = A.gtab;
    U.cap = A.cap;
    U.n = n;
    U.parts = parts;
    U.partial = c->wpart;
    U.w32 = c->w32;
    U.wm = c->wm;
    U.wv = c->wv;
    U.w16t = c->w16t;
    U.w16n = c->w16n;
    U.ts = c->ts;
    U.tgrad = c->tg;
    U.t16 = reinterpret_cast<__half2 *>(c->t16);
    U.entries = c->table_entries;
    U.levels = cf.levels;
    U.lattice_begin[0] = 0;
    for (uint32_t l = 0; l < (uint32_t)kNrcMaxLevels; ++l) {
        const uint32_t r1 = c->level_res[l] + 1u;
        const uint32_t span = c->level_dense[l] ? r1 * r1 * r1 : c->table_entries;
        U.lattice_begin[l + 1] = U.lattice_begin[l] + (l < cf.levels ? span : 0u);
    }
    U.lattice_total = U.lattice_begin[cf.levels];
    U.lossp = c->lossp;
    U.stats = c->stats;
    U.as = A.as;
    U.lr = cf.learning_rate;
    U.b1 = cf.beta1;
    U.b2 = cf.beta2;
    U.eps = cf.epsilon;
    U.dense_adam = cf.dense_table_adam ? 1u : 0u;
    U.mlp_blocks = ((uint32_t)kNrcWeights + kMlpPerBlock - 1) / kMlpPerBlock;
    U.scatter_blocks = (uint32_t)(((uint64_t)n * 8u + kUpdBlock - 1) / kUpdBlock);
    hipLaunchKernelGGL(nrc_fused_update, dim3(U.mlp_blocks + U.scatter_blocks * cf.levels), dim3(kUpdBlock), 0, s, U);
    hipLaunchKernelGGL(nrc_fused_table_adam, dim3((U.lattice_total + kUpdBlock - 1) / kUpdBlock), dim3(kUpdBlock), 0,
                       s, U);
    return check_launch("nrc fused training step");
}

int nrc_fused_grads(NanNrc *c, const float *pos, const float *nrm, const float *alb, const float *target, uint32_t n,
                    float *weight_grad, float *input_grad, hipStream_t s) {
    if (n == 0) return 0;
    if (!nrc_fused_eligible(c, n))
        return fail(nullptr, 0, "nan_nrc: the fused gradients take at most %u records",
                    (uint32_t)kNrcBwdWaves * 64u);
    FbArgs A = fb_args(c, pos, nrm, alb, target, nullptr, n);
    const uint32_t parts = A.groups;
    const float scale = 1.0f / (3.0f * (float)n);
    A.gtab = nullptr;  // parameters and their gradient buffers stay untouched
    A.as = nullptr;
    A.input_grad = input_grad;
    A.input_scale = scale;
    hipLaunchKernelGGL(nrc_fused_fb<true>, dim3(parts), dim3(64 * kFbWaves), 0, s, A);
    hipLaunchKernelGGL(nrc_fused_sum, dim3((kNrcWeights + 255) / 256), dim3(256), 0, s, c->wpart, parts, scale,
                       weight_grad);
    return check_launch("nrc fused gradients");
}

}  // namespace nanx
