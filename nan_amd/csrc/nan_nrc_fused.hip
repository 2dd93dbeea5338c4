// Fused Adam step of the learned radiance cache, for caches whose hash-grid
// levels are all densely indexed ((res+1)^3 entries fit the table: the grids
// of small scenes).  Three launches per step instead of seven:
//
//   nrc_fused_fb          one wave per 64-record group (four 16-record
//                         tiles): hash-grid encoding, the 48 -> 64x4 -> 16
//                         ReLU MLP forward (the layers' inputs kept in LDS for
//                         the backward), the relative L2 loss, and the
//                         backward.  The wave's weight gradients over its 64
//                         records go to its own partial in MFMA fragment order
//                         (no cross-wave reduction); each record's gradient
//                         with respect to its hash features goes to a
//                         per-level array.
//   nrc_fused_update      the weights' Adam on the fixed-order sum of the
//                         partials (sixteen threads per element) and the
//                         step's loss statistics, beside the table-gradient
//                         scatter: one thread per (record, level, corner), one
//                         packed FP16 atomic each.
//   nrc_fused_table_adam  the table's lazy Adam over the dense lattices.
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
    uint32_t lattice_begin[kNrcMaxLevels + 1];  // prefix of the levels' (res+1)^3
    const float2 *lossp;
    float *stats;
    AdamState *as;
    float lr, b1, b2, eps;
    uint32_t dense_adam, mlp_blocks;
};

// acc[t] = W^T x[t] for the four tiles (B/D fragment layout: lane%16 = record
// of the tile, 4*(lane/16)+i = unit); each A fragment of W ([out][in],
// global) feeds four MFMAs.  The layer's fragments are loaded together
// before the products: one round trip to L2 or memory per layer (the update
// kernel has just rewritten the weights, so they are not in this XCD's L2).
template <int IN, int OUT>
__device__ __forceinline__ void fwd4(const _Float16 *w, const half4_t (&x)[4][IN / 16],
                                     floatx4_t (&acc)[4][OUT / 16]) {
    constexpr int KB = IN / 16, KO = OUT / 16;
    const int lane = __lane_id(), r = lane & 15, q = lane >> 4;
    half4_t a[KO][KB];
#pragma unroll
    for (int ot = 0; ot < KO; ++ot)
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) a[ot][kb] = *(const half4_t *)&w[(ot * 16 + r) * IN + kb * 16 + 4 * q];
#pragma unroll
    for (int ot = 0; ot < KO; ++ot) {
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t][ot] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb + 1 < KB; kb += 2)
#pragma unroll
            for (int t = 0; t < 4; ++t)
                acc[t][ot] = mfma32(a[ot][kb], a[ot][kb + 1], x[t][kb], x[t][kb + 1], acc[t][ot]);
        if (KB % 2)
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t][ot] = mfma16(a[ot][KB - 1], x[t][KB - 1], acc[t][ot]);
    }
}

// ReLU after the rounding to FP16 (rounding is monotone and keeps zero).
template <int KB>
__device__ __forceinline__ void relu4(const floatx4_t (&acc)[4][KB], half4_t (&x)[4][KB]) {
    const half4_t zero = {(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f};
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
            x[t][kb] = __builtin_elementwise_max(__builtin_convertvector(acc[t][kb], half4_t), zero);
}

// Weight gradient of layer L over the wave's 64 records: dW^T = dY^T X with
// the records as the MFMA reduction dimension (both operands transposed
// through LDS), one 16x16 block at a time, stored to the wave's partial in
// fragment order (see frag_weight).
template <int L>
__device__ __forceinline__ void bwd_dw(const half4_t (&xin)[4][nrc_layer_in(L) / 16],
                                       const half4_t (&dy)[4][nrc_layer_out(L) / 16], _Float16 *tx, _Float16 *ty,
                                       float *part) {
    constexpr int IN = nrc_layer_in(L), OUT = nrc_layer_out(L), KI = IN / 16, KO = OUT / 16;
    const int lane = __lane_id(), r = lane & 15, q = lane >> 4;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int kb = 0; kb < KO; ++kb)
#pragma unroll
            for (int i = 0; i < 4; ++i) ty[(kb * 16 + 4 * q + i) * kTS + t * 16 + r] = dy[t][kb][i];
#pragma unroll
        for (int kb = 0; kb < KI; ++kb)
#pragma unroll
            for (int i = 0; i < 4; ++i) tx[(kb * 16 + 4 * q + i) * kTS + t * 16 + r] = xin[t][kb][i];
    }
    __syncthreads();  // one wave per block: orders the transposes' writes and reads
    float4 *p = reinterpret_cast<float4 *>(part + nrc_layer_offset(L)) + lane;
#pragma unroll
    for (int ot = 0; ot < KO; ++ot) {
        half4_t a[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) a[ks] = *(const half4_t *)&ty[(ot * 16 + r) * kTS + ks * 16 + 4 * q];
#pragma unroll
        for (int it = 0; it < KI; ++it) {
            floatx4_t d = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
                d = mfma16(a[ks], *(const half4_t *)&tx[(it * 16 + r) * kTS + ks * 16 + 4 * q], d);
            p[(ot * KI + it) * 64] = make_float4(d[0], d[1], d[2], d[3]);
        }
    }
    __syncthreads();  // the next layer rewrites the transposes
}

// The A fragments of W dY^T for input blocks [0, NI) of layer L (w16n is
// [in][out]), loaded together.
template <int L, int NI>
__device__ __forceinline__ void load_wn(const _Float16 *w16n, half4_t (&a)[NI][nrc_layer_out(L) / 16]) {
    constexpr int OUT = nrc_layer_out(L), KO = OUT / 16;
    const int lane = __lane_id(), r = lane & 15, q = lane >> 4;
    const _Float16 *wn = w16n + nrc_layer_offset(L);
#pragma unroll
    for (int it = 0; it < NI; ++it)
#pragma unroll
        for (int ko = 0; ko < KO; ++ko) a[it][ko] = *(const half4_t *)&wn[(it * 16 + r) * OUT + ko * 16 + 4 * q];
}

// W dY^T for input block it of layer L over the four tiles.
template <int KO>
__device__ __forceinline__ void bwd_dx(const half4_t (&a)[KO], const half4_t (&dy)[4][KO], floatx4_t (&dx)[4]) {
#pragma unroll
    for (int t = 0; t < 4; ++t) dx[t] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ko = 0; ko < KO; ++ko)
#pragma unroll
        for (int t = 0; t < 4; ++t) dx[t] = mfma16(a[ko], dy[t][ko], dx[t]);
}

// Backward of layer L >= 1: its weight gradient, and the dY of the layer
// below: W dY^T where the layer's input (a ReLU output) is positive, rounded
// to FP16.  The weights' loads are in flight during the weight gradient.
template <int L>
__device__ __forceinline__ void bwd4(const _Float16 *w16n, const half4_t (&xin)[4][4],
                                     const half4_t (&dy)[4][nrc_layer_out(L) / 16], _Float16 *tx, _Float16 *ty,
                                     float *part, half4_t (&dyn)[4][4]) {
    constexpr int KO = nrc_layer_out(L) / 16;
    half4_t a[4][KO];
    load_wn<L, 4>(w16n, a);
    bwd_dw<L>(xin, dy, tx, ty, part);
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        floatx4_t dx[4];
        bwd_dx<KO>(a[it], dy, dx);
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                dyn[t][it][i] = (_Float16)((float)xin[t][it][i] > 0.0f ? dx[t][i] : 0.0f);
    }
}

__device__ __forceinline__ void save_act(half4_t (&sv)[4][4][64], const half4_t (&x)[4][4]) {
    const int lane = __lane_id();
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) sv[t][kb][lane] = x[t][kb];
}

__device__ __forceinline__ void load_act(const half4_t (&sv)[4][4][64], half4_t (&x)[4][4]) {
    const int lane = __lane_id();
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) x[t][kb] = sv[t][kb][lane];
}

// FULL: also the gradient with respect to all 48 inputs (a test hook).
template <bool FULL>
__global__ __launch_bounds__(64) void nrc_fused_fb(FbArgs A) {
    __shared__ __align__(16) _Float16 x0s[64 * kX0Stride];
    __shared__ __align__(16) _Float16 tx[64 * kTS], ty[64 * kTS];
    __shared__ half4_t sv[kNrcHidden][4][4][64];  // inputs of layers 1..4, fragment order
    const int lane = __lane_id(), r = lane & 15, q = lane >> 4;
    if (A.as && blockIdx.x == 0 && lane == 0) {  // this step's bias corrections
        const uint32_t step = A.as->step + 1u;
        A.as->next_step = step;
        A.as->c1 = bias_correction(A.b1, step);
        A.as->c2 = bias_correction(A.b2, step);
    }
    float *part = A.partial + (size_t)blockIdx.x * kNrcWeights;
    float loss = 0.0f, nvalid = 0.0f;
    const half4_t zero4 = {(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f};
    {
        const uint32_t g = blockIdx.x;  // one 64-record group per wave
        {  // encoding: lane = record.  Every level is densely indexed here; the
           // 32 gathers of four levels are issued together (lanes past n read
           // the cell at the box corner and store zeros).
            const uint32_t s = g * 64u + (uint32_t)lane;
            const bool live = s < A.n;
            _Float16 *row = x0s + lane * kX0Stride;
            float p[3] = {0.0f, 0.0f, 0.0f};
            if (live)
                for (int a = 0; a < 3; ++a) p[a] = A.pos[3 * (size_t)s + a];
            const uint32_t levels = A.E.levels;
            for (uint32_t l0 = 0; l0 < levels; l0 += 4) {
                half2_t v[4][8];
                float w[4][8];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t l = min(l0 + (uint32_t)j, levels - 1u);
                    uint32_t base[3];
                    float f[3];
                    level_cell(A.E, l, p, base, f);
                    const half2_t *tab = reinterpret_cast<const half2_t *>(A.E.table) + (size_t)l * A.E.entries;
                    const uint32_t r1 = A.E.res[l] + 1u;
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
                        v[j][k] = tab[(base[0] + dx) + r1 * ((base[1] + dy) + r1 * (base[2] + dz))];
                        w[j][k] = (dx ? f[0] : 1.0f - f[0]) * (dy ? f[1] : 1.0f - f[1]) * (dz ? f[2] : 1.0f - f[2]);
                    }
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        f0 += w[j][k] * (float)v[j][k][0];
                        f1 += w[j][k] * (float)v[j][k][1];
                    }
                    if (l0 + (uint32_t)j < levels)
                        *reinterpret_cast<half2_t *>(&row[2 * (l0 + j)]) =
                            live ? half2_t{(_Float16)f0, (_Float16)f1} : half2_t{(_Float16)0.0f, (_Float16)0.0f};
                }
            }
            for (uint32_t l = levels; l < (uint32_t)kNrcMaxLevels; ++l)
                *reinterpret_cast<half2_t *>(&row[2 * l]) = half2_t{(_Float16)0.0f, (_Float16)0.0f};
            for (int a = 0; a < 3; ++a) {
                row[32 + a] = (_Float16)(s < A.n ? A.nrm[3 * (size_t)s + a] : 0.0f);
                row[35 + a] = (_Float16)(s < A.n ? A.alb[3 * (size_t)s + a] : 0.0f);
            }
            for (int k = 38; k < kNrcIn; ++k) row[k] = (_Float16)0.0f;
        }
        __syncthreads();
        half4_t x0[4][3];
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int kb = 0; kb < 3; ++kb)
                x0[t][kb] = *(const half4_t *)&x0s[(t * 16 + r) * kX0Stride + kb * 16 + 4 * q];
        // forward; the inputs of layers 1..4 are kept in LDS for the backward
        half4_t dy4[4][1];
        {
            half4_t x[4][4];
            floatx4_t acc[4][4];
            fwd4<kNrcIn, kNrcWidth>(A.w16t + nrc_layer_offset(0), x0, acc);
            relu4(acc, x);
            save_act(sv[0], x);
#pragma unroll
            for (int l = 1; l < kNrcHidden; ++l) {
                fwd4<kNrcWidth, kNrcWidth>(A.w16t + nrc_layer_offset(l), x, acc);
                relu4(acc, x);
                save_act(sv[l], x);
            }
            floatx4_t out[4][1];
            fwd4<kNrcWidth, kNrcOut>(A.w16t + nrc_layer_offset(kNrcLayers - 1), x, out);
            // loss: lanes of quarter 0 hold outputs 0..3 of record r of each tile
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                dy4[t][0] = zero4;
                const uint32_t s = g * 64u + (uint32_t)t * 16u + (uint32_t)r;
                if (q == 0 && s < A.n && (!A.valid || A.valid[s])) {
                    float gr[3];
                    for (int c = 0; c < 3; ++c) {
                        const float pc = out[t][0][c], tc = A.target[3 * (size_t)s + c];
                        const float d = pc * pc + 0.01f, diff = pc - tc;
                        loss += diff * diff / d;
                        gr[c] = fminf(fmaxf(2.0f * diff / d, -kGradClamp), kGradClamp);
                    }
                    dy4[t][0] = half4_t{(_Float16)gr[0], (_Float16)gr[1], (_Float16)gr[2], (_Float16)0.0f};
                    nvalid += 1.0f;
                }
            }
        }
        // backward, layer by layer
        floatx4_t dh[4][FULL ? 3 : 2];  // gradients of the 32 hash features (FULL: all 48 inputs)
        {
            half4_t dy[4][4], dyb[4][4], xin[4][4];
            load_act(sv[3], xin);
            bwd4<4>(A.w16n, xin, dy4, tx, ty, part, dy);
            load_act(sv[2], xin);
            bwd4<3>(A.w16n, xin, dy, tx, ty, part, dyb);
            load_act(sv[1], xin);
            bwd4<2>(A.w16n, xin, dyb, tx, ty, part, dy);
            load_act(sv[0], xin);
            bwd4<1>(A.w16n, xin, dy, tx, ty, part, dyb);
            half4_t a0[FULL ? 3 : 2][4];
            load_wn<0, FULL ? 3 : 2>(A.w16n, a0);
            bwd_dw<0>(x0, dyb, tx, ty, part);
#pragma unroll
            for (int it = 0; it < (FULL ? 3 : 2); ++it) {
                floatx4_t dx[4];
                bwd_dx<4>(a0[it], dyb, dx);
#pragma unroll
                for (int t = 0; t < 4; ++t) dh[t][it] = dx[t];
            }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t s = g * 64u + (uint32_t)t * 16u + (uint32_t)r;
            if (s >= A.n) continue;
            if (FULL && A.input_grad) {
#pragma unroll
                for (int it = 0; it < (FULL ? 3 : 2); ++it)
                    *reinterpret_cast<float4 *>(&A.input_grad[(size_t)s * kNrcIn + it * 16 + 4 * q]) =
                        make_float4(dh[t][it][0] * A.input_scale, dh[t][it][1] * A.input_scale,
                                    dh[t][it][2] * A.input_scale, dh[t][it][3] * A.input_scale);
            }
            if (!A.gtab) continue;
#pragma unroll
            for (int it = 0; it < 2; ++it)
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const uint32_t lv = (uint32_t)(it * 8 + 2 * q + h);
                    if (lv < A.E.levels) A.gtab[(size_t)lv * A.cap + s] = make_float2(dh[t][it][2 * h], dh[t][it][2 * h + 1]);
                }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        loss += __shfl_down(loss, off);
        nvalid += __shfl_down(nvalid, off);
    }
    if (lane == 0) A.lossp[blockIdx.x] = make_float2(loss, nvalid);
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
// (blockIdx.y), eight neighbouring lanes the eight corners of a record, one
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
// blocks after them (blockIdx.y = level): the table-gradient atomics.
__global__ __launch_bounds__(kUpdBlock) void nrc_fused_update(UpdArgs U) {
    __shared__ float red[kUpdBlock];
    const uint32_t tid = threadIdx.x;
    if (blockIdx.x >= U.mlp_blocks) {
        scatter_corner(U, (blockIdx.x - U.mlp_blocks) * kUpdBlock + tid, blockIdx.y);
        return;
    }
    if (blockIdx.y != 0) return;
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
    // sub + 16, ... four loads at a time; the sixteen sums combine in a fixed
    // order
    const uint32_t fl = tid % kMlpPerBlock, sub = tid / kMlpPerBlock;
    const uint32_t f = blockIdx.x * kMlpPerBlock + fl;
    float g = 0.0f;
    if (f < (uint32_t)kNrcWeights) {
        uint32_t p = sub;
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

// The table's lazy Adam over the dense lattices, one entry per thread; the
// step's last kernel stores Adam's step count.
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
    if (n > (uint32_t)kNrcBwdWaves * 64u) return false;  // one 64-record group per wave and partial
    for (uint32_t l = 0; l < c->cfg.levels; ++l)
        if (!c->level_dense[l]) return false;
    return true;
}

int nrc_train_fused(NanNrc *c, const float *pos, const float *nrm, const float *alb, const float *target,
                    const uint32_t *valid, uint32_t n, hipStream_t s) {
    if (n == 0) return 0;
    if (!nrc_fused_eligible(c, n))
        return fail(nullptr, 0, "nan_nrc: the fused training step needs densely indexed levels and at most %u records",
                    (uint32_t)kNrcBwdWaves * 64u);
    if (nrc_reserve_train(c, n, s)) return -1;
    FbArgs A = fb_args(c, pos, nrm, alb, target, valid, n);
    const uint32_t parts = A.groups;
    hipLaunchKernelGGL(nrc_fused_fb<false>, dim3(parts), dim3(64), 0, s, A);
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
        U.lattice_begin[l + 1] = U.lattice_begin[l] + (l < cf.levels ? r1 * r1 * r1 : 0u);
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
    hipLaunchKernelGGL(nrc_fused_update, dim3(U.mlp_blocks + U.scatter_blocks, cf.levels), dim3(kUpdBlock), 0, s, U);
    hipLaunchKernelGGL(nrc_fused_table_adam, dim3((U.lattice_total + kUpdBlock - 1) / kUpdBlock), dim3(kUpdBlock), 0,
                       s, U);
    return check_launch("nrc fused training step");
}

int nrc_fused_grads(NanNrc *c, const float *pos, const float *nrm, const float *alb, const float *target, uint32_t n,
                    float *weight_grad, float *input_grad, hipStream_t s) {
    if (n == 0) return 0;
    if (!nrc_fused_eligible(c, n))
        return fail(nullptr, 0, "nan_nrc: the fused gradients need densely indexed levels and at most %u records",
                    (uint32_t)kNrcBwdWaves * 64u);
    FbArgs A = fb_args(c, pos, nrm, alb, target, nullptr, n);
    const uint32_t parts = A.groups;
    const float scale = 1.0f / (3.0f * (float)n);
    A.gtab = nullptr;  // parameters and their gradient buffers stay untouched
    A.as = nullptr;
    A.input_grad = input_grad;
    A.input_scale = scale;
    hipLaunchKernelGGL(nrc_fused_fb<true>, dim3(parts), dim3(64), 0, s, A);
    hipLaunchKernelGGL(nrc_fused_sum, dim3((kNrcWeights + 255) / 256), dim3(256), 0, s, c->wpart, parts, scale,
                       weight_grad);
    return check_launch("nrc fused gradients");
}

}  // namespace nanx
