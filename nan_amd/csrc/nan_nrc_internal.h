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
    float *stats = nullptr;    // [0] loss sum, [1] normalising count, [2] valid records (floats), [3] spare
    void *adam = nullptr;      // Adam step count and bias corrections (16 bytes, device)
    ~NanNrc();
};

namespace nanx {

// Grow the prediction / training scratch to hold n samples (refused while
// `stream` is being captured into a graph: growing allocates).
int nrc_reserve(NanNrc *c, uint32_t n, hipStream_t stream);
int nrc_reserve_train(NanNrc *c, uint32_t n, hipStream_t stream);
// Network prediction for n samples into pred[n] (RGB + one spare output).
int nrc_predict(NanNrc *c, const float *pos, const float *nrm, const float *alb, uint32_t n, float4 *pred,
                hipStream_t stream);
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

}  // namespace nanx
