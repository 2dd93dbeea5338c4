// Learned radiance cache: state shared by the network kernels (nan_nrc.hip)
// and the learned texel tracer (nan_learned.hip).  Not part of the C ABI.
#pragma once

#include <hip/hip_fp16.h>

#include "nan_internal.h"

namespace nanx {

constexpr int kNrcIn = 48;      // padded network input: 32 hash features, normal, albedo, zeros
constexpr int kNrcWidth = 64;   // hidden width
constexpr int kNrcOut = 16;     // padded output: RGB in the first 3
constexpr int kNrcHidden = 4;   // hidden layers
constexpr int kNrcLayers = kNrcHidden + 1;
constexpr int kNrcMaxLevels = 16;
constexpr int kNrcBwdWaves = 512;  // waves (one per block) of the backward kernel
constexpr int kNrcEncGroup = 64 * kNrcIn;  // halves of encoded input per 64-sample group

__host__ __device__ constexpr int nrc_layer_in(int l) { return l == 0 ? kNrcIn : kNrcWidth; }
__host__ __device__ constexpr int nrc_layer_out(int l) { return l == kNrcLayers - 1 ? kNrcOut : kNrcWidth; }
__host__ __device__ constexpr int nrc_layer_offset(int l) {
    int o = 0;
    for (int k = 0; k < l; ++k) o += nrc_layer_in(k) * nrc_layer_out(k);
    return o;
}
constexpr int kNrcWeights = nrc_layer_offset(kNrcLayers);  // 16,384

// Encoding as seen by the kernels.
struct NrcGrid {
    const _Float16 *table;  // [levels][entries][2]
    uint32_t levels, entries;
    uint32_t res[kNrcMaxLevels];
    uint32_t dense[kNrcMaxLevels];
    float bmin[3], bscale[3];  // unit-cube coordinate = (p - bmin) * bscale
};


// ---------------------------------------------------------------------------
// device helpers shared by the network kernels
// ---------------------------------------------------------------------------
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));

// A K=32 product over two K=16 fragments, issued as two K=16 MFMAs.  The
// single K=32 instruction (v_mfma_f32_16x16x32_f16) was measured wrong when
// the compiler rewrote the upper half of its A or B registers a few
// instructions after issue (the products of every tile but the first came
// out wrong in a four-wave kernel; the K=16 form is exact there).
__device__ __forceinline__ floatx4_t mfma16(half4_t a, half4_t b, floatx4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ floatx4_t mfma32(half4_t a0, half4_t a1, half4_t b0, half4_t b1, floatx4_t c) {
    return mfma16(a1, b1, mfma16(a0, b0, c));
}

// Per-record gradient bound of the relative L2 loss (before the 1/3n scale).
constexpr float kGradClamp = 16384.0f;

// Table gradients accumulate as packed FP16 pairs (one global_atomic_pk_add_f16
// per corner carries both features), pre-scaled by kTableGradScale to keep
// sums of many contributions inside FP16 range.
constexpr float kTableGradScale = 1.0f / 16.0f;

// The cell of a position in one level: lower corner and fractions.
__device__ __forceinline__ void level_cell(const NrcGrid &E, uint32_t level, const float *pos, uint32_t (&base)[3],
                                           float (&f)[3]) {
    float fr = (float)E.res[level];
    for (int a = 0; a < 3; ++a) {
        float u = fminf(fmaxf((pos[a] - E.bmin[a]) * E.bscale[a], 0.0f), 1.0f);
        float x = u * fr;
        float fl = fminf(floorf(x), fr - 1.0f);
        base[a] = (uint32_t)fl;
        f[a] = x - fl;
    }
}

// Corner k (bit 0: +x, bit 1: +y, bit 2: +z) of a cell: table entry (dense
// indexing where (res+1)^3 fits, spatial hash otherwise) and trilinear weight.
__device__ __forceinline__ void cell_corner(const NrcGrid &E, uint32_t level, const uint32_t (&base)[3],
                                            const float (&f)[3], uint32_t k, uint32_t &idx, float &w) {
    uint32_t r1 = E.res[level] + 1u;
    uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
    uint32_t x = base[0] + dx, y = base[1] + dy, z = base[2] + dz;
    idx = E.dense[level] ? x + r1 * (y + r1 * z)
                         : ((x * 1u) ^ (y * 2654435761u) ^ (z * 805459861u)) & (E.entries - 1u);
    w = (dx ? f[0] : 1.0f - f[0]) * (dy ? f[1] : 1.0f - f[1]) * (dz ? f[2] : 1.0f - f[2]);
}

struct AdamArgs {
    float lr, b1, b2, eps;
};

// Adam's step count lives on the device, so a captured training step that is
// replayed takes the next step's bias corrections 1 - beta^step each time.
// The first kernel of a step reads `step` and publishes step + 1 and its
// corrections; a later kernel of the step reads those and stores `step` (no
// kernel writes a field that it reads).
struct AdamState {
    uint32_t step, next_step;
    float c1, c2;
};

// Adam's bias correction 1 - beta^step as -expm1(step * ln beta): FP32, no
// cancellation for beta^step near 1, and cheap enough for one thread of a
// training kernel (beta = 0 gives 1 for every step >= 1).
__device__ __forceinline__ float bias_correction(float beta, uint32_t step) {
    return -expm1f((float)step * logf(beta));
}

__device__ __forceinline__ float grad_scale(const float *stats) {
    return 1.0f / (3.0f * fmaxf(stats[1], 1.0f));
}

}  // namespace nanx

// Optimiser state of one hash-table entry (both features): FP32 master
// weights, Adam first and second moments, padding to 32 bytes.
struct alignas(32) TableState {
    float w[2], m[2], v[2], pad[2];
};

struct NanNrc {
    NanNrcConfig cfg{};
    uint32_t table_entries = 0;
    uint64_t table_params = 0;
    uint32_t level_res[nanx::kNrcMaxLevels] = {};
    uint32_t level_dense[nanx::kNrcMaxLevels] = {};
    // MLP: fp32 master [layer][out][in], fp16 copies transposed ([out][in]) and
    // plain ([in][out]), Adam moments, per-wave gradient partials
    float *w32 = nullptr, *wm = nullptr, *wv = nullptr, *wpart = nullptr;
    _Float16 *w16t = nullptr, *w16n = nullptr;
    // hash table: optimiser state records, packed fp16 gradient pairs, fp16
    // copy read by the encoding
    TableState *ts = nullptr;
    __half2 *tg = nullptr;
    _Float16 *t16 = nullptr;
    // batch scratch (capacity in 64-sample groups): prediction (enc, pred)
    // and training (save, dy, denc)
    uint32_t cap_groups = 0, cap_train_groups = 0;
    _Float16 *enc = nullptr;   // [groups][3072] encoded input, MFMA fragment order
    _Float16 *save = nullptr;  // [4][groups][4096] inputs of layers 1..4
    _Float16 *dy = nullptr;    // [groups][4096] gradient wrt the current layer's output
    float *denc = nullptr;     // [groups * 64][48]
    float4 *pred = nullptr;    // [groups * 64]
    // the training step's own encoded input and outputs (cap_train_groups), so
    // that queries on another stream may run beside a step
    _Float16 *tenc = nullptr;
    float4 *tpred = nullptr;
    // copy of what the queries read (FP16 weights [kNrcWeights], then the FP16
    // table) for queries that run beside training (nrc_snapshot)
    _Float16 *snap = nullptr;
    float *stats = nullptr;    // [0] loss sum, [1] normalising count, [2] valid records (floats), [3] spare
    float2 *lossp = nullptr;   // [kNrcBwdWaves] per-block loss sums of the fused step
    bool fused_train = true;   // three-launch step for minibatches of at most kNrcBwdWaves * 64 records
    void *adam = nullptr;      // Adam step count and bias corrections (16 bytes, device)
    ~NanNrc();
};

namespace nanx {

// The encoding's view of a cache (table, level resolutions, box).
NrcGrid nrc_grid(const NanNrc *c);

// Grow the prediction / training scratch to hold n samples (refused while
// `stream` is being captured into a graph: growing allocates).
int nrc_reserve(NanNrc *c, uint32_t n, hipStream_t stream);
int nrc_reserve_train(NanNrc *c, uint32_t n, hipStream_t stream);
// Network prediction for n samples into pred[n] (RGB + one spare output),
// from the live parameters or from the last nrc_snapshot.
int nrc_predict(NanNrc *c, const float *pos, const float *nrm, const float *alb, uint32_t n, float4 *pred,
                hipStream_t stream, bool from_snapshot = false);
// Copy the parameters the queries read into c->snap (allocated by
// nrc_reserve_snapshot, refused while `stream` is being captured), so that
// queries can run while a training step on another stream changes them.
int nrc_reserve_snapshot(NanNrc *c, hipStream_t stream);
int nrc_snapshot(NanNrc *c, hipStream_t stream);
// The cache query of nan_nrc_infer with its scratch passed in: encoded input
// enc (enc_groups groups of 64 samples) and raw outputs pred (enc_groups * 64
// float4), then the clamped RGB into out[n][3].  Allocates nothing.
int nrc_infer_into(NanNrc *c, const float *pos, const float *nrm, const float *alb, uint32_t n, _Float16 *enc,
                   uint32_t enc_groups, float4 *pred, float *out, hipStream_t stream);
// Adam steps (cfg.minibatches of them, over consecutive slices) on the
// relative L2 loss over n records.  valid (nullable): per-record flag;
// invalid records contribute nothing, and the loss is normalised by the
// slice size.
int nrc_train_step(NanNrc *c, const float *pos, const float *nrm, const float *alb, const float *target,
                   const uint32_t *valid, uint32_t n, hipStream_t stream);
// One Adam step in three launches (nan_nrc_fused.hip), for steps of at most
// kNrcBwdWaves * 64 records on dense and hashed levels; the same loss and
// update as the seven-launch step.
bool nrc_fused_eligible(const NanNrc *c, uint32_t n);
int nrc_train_fused(NanNrc *c, const float *pos, const float *nrm, const float *alb, const float *target,
                    const uint32_t *valid, uint32_t n, hipStream_t stream);
// The fused forward and backward alone: gradients of the mean loss with
// respect to the weights and to the 48 inputs (test hook).
int nrc_fused_grads(NanNrc *c, const float *pos, const float *nrm, const float *alb, const float *target, uint32_t n,
                    float *weight_grad, float *input_grad, hipStream_t stream);

}  // namespace nanx
