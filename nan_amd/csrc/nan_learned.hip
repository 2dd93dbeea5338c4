// Learned form of the texel tracer: plain path suffixes train the radiance
// cache, then every texel sample takes the sun term and one cosine-sampled
// bounce as the plain tracer does (texture_space_path_tracer.slang:86-117)
// and ends in the cache at the first hit y1:
//   plain   E = sun(x0) + (az/pdf) * (Le(y1) + suffix(y1))
//   learned E = sun(x0) + (az/pdf) * (Le(y1) + max(0, cache(y1)))
// where suffix(y1) is the radiance trace_radiance (slang:57-84) gathers after
// y1's own emission, i.e. the cache's training target.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "nan_nrc_internal.h"
#include "nan_tracer.h"

using namespace nanx;

// Device working set (plain pointers, passed to kernels by value).
struct LearnedDev {
    float *tr_pos, *tr_nrm, *tr_alb, *tr_target;  // training records [n][3]
    uint32_t *tr_valid;
    float *pv_rec;  // NAN_LEARNED_OVERLAP: the previous call's records, trained on while this call
                    // traces: pos, nrm, alb, target [cap][3] and valid u32 [cap], one after another
    float *q_pos, *q_nrm, *q_alb, *q_w;  // cache queries [slots * spp][3], weight (0: none)
    float4 *q_pred;
    float4 *partial;  // per payload slot: sum over samples of the terms outside the cache
    const uint32_t *frame;  // device [iteration, reset] in device-frame mode, else null (host values)
};

struct LearnedBuffers {
    LearnedDev d{};
    uint32_t cap_train = 0;
    uint64_t cap_query = 0;
    uint32_t cap_prev = 0;     // records pv_rec holds
    // NAN_LEARNED_OVERLAP: the parameter snapshot and the training of the
    // previous call's records run on side_stream, forked from and joined into
    // the caller's stream
    hipStream_t side_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_snap = nullptr, ev_traced = nullptr, ev_join = nullptr;
    ~LearnedBuffers() {
        for (void *p : {(void *)d.tr_pos, (void *)d.tr_nrm, (void *)d.tr_alb, (void *)d.tr_target,
                        (void *)d.tr_valid, (void *)d.q_pos, (void *)d.q_nrm, (void *)d.q_alb, (void *)d.q_w,
                        (void *)d.q_pred, (void *)d.partial, (void *)d.pv_rec})
            hipFree(p);
        for (hipEvent_t e : {ev_fork, ev_snap, ev_traced, ev_join})
            if (e) hipEventDestroy(e);
        if (side_stream) hipStreamDestroy(side_stream);
    }
};

namespace {

// ---------------------------------------------------------------------------
// light sampling of emissive triangles (NEE), combined with the cosine-sampled
// bounce by the power heuristic (Veach, "Robust Monte Carlo Methods for Light
// Transport Simulation", 1997, ch. 9).  Emission is two-sided, as shading
// reports it for either face, so pdfs use |cos| at the emitter.
// ---------------------------------------------------------------------------
struct EmitterSample {
    f3 dir, le;
    float dist, pdf;  // pdf in solid angle
};

__device__ __forceinline__ float power_weight(float a, float b) {
    float a2 = a * a, b2 = b * b;
    return a2 / (a2 + b2);
}

// Uniform point over the emitters' total area: triangle by area CDF, point by
// the square-root parametrisation.
template <bool MAT>
__device__ __forceinline__ bool sample_emitter(const DevScene &S, f3 o, float u0, float u1, float u2,
                                               EmitterSample &es) {
    uint32_t lo = 0, hi = S.emitter_count - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (u0 < S.emitter_cdf[mid]) hi = mid;
        else lo = mid + 1;
    }
    const float4 *e = S.emitters + 5 * (size_t)lo;
    float4 r0 = e[0], r1 = e[1], r2 = e[2], r3 = e[3], r4 = e[4];
    float su = sqrtf(u1);
    float a = su * (1.0f - u2), b = su * u2;
    f3 p = mk(r0.x + a * r1.x + b * r2.x, r0.y + a * r1.y + b * r2.y, r0.z + a * r1.z + b * r2.z);
    f3 d = p - o;
    float d2 = dot(d, d);
    if (!(d2 > 1e-12f)) return false;
    es.dist = sqrtf(d2);
    es.dir = d * (1.0f / es.dist);
    float cy = fabsf(dot(mk(r3.x, r3.y, r3.z), es.dir));
    if (!(cy > 1e-6f)) return false;
    es.pdf = d2 / (cy * S.emitter_area);
    es.le = mk(r4.x, r4.y, r4.z);
    if (MAT && __float_as_uint(r4.w) != 0u) {
        // emission at the sampled point (barycentrics 1 - a - b, a, b): the
        // emissive texture times the factor, zero where the alpha test cuts
        // the surface away
        uint32_t tri = __float_as_uint(r1.w);
        if (!alpha_accept(S, tri, a, b)) es.le = mk(0.0f, 0.0f, 0.0f);
        const uint32_t mat = __float_as_uint(S.tri_shading[5 * (size_t)tri + 4].w) & 0x7FFFFFFFu;
        const uint32_t etex = __float_as_uint(S.materials[3 * (size_t)mat + 2].y);
        if (etex != kInvalid) {
            const float4 *uvp = S.tri_uv + 2 * (size_t)tri;
            const float4 u01 = uvp[0], u2 = uvp[1];
            const float w0 = 1.0f - a - b;
            float4 tx = sample_texture(S, etex, w0 * u01.x + a * u01.z + b * u2.x, w0 * u01.y + a * u01.w + b * u2.y);
            es.le = mk(es.le.x * tx.x, es.le.y * tx.y, es.le.z * tx.z);
        }
    }
    return true;
}

// Solid-angle pdf with which light sampling picks the point hit at distance t
// along dir on triangle tri.
__device__ __forceinline__ float emitter_pdf(const DevScene &S, uint32_t tri, f3 dir, float t) {
    float cy = fabsf(dot(world_flat_normal(S, tri), dir));
    return cy > 1e-6f ? t * t / (cy * S.emitter_area) : INFINITY;
}

template <int STACK, int BLOCK, bool COUNT, bool MAT>
__device__ __forceinline__ bool visible_to(const DevScene &S, f3 o, f3 dir, float tmax, uint32_t *col,
                                           PathCounters &pc) {
    Ray r;
    r.o = o;
    r.d = dir;
    r.tmin = 0.0f;
    r.tmax = tmax;
    ray_setup(r);
    Hit h{r.tmax, 0.0f, 0.0f, kInvalid};
    TraversalStats st{0, 0};
    bool occluded = traverse<STACK, BLOCK, true, COUNT, MAT>(S, r, h, col, st);
    if (COUNT) {
        pc.shadow++;
        pc.shadow_nodes += st.nodes;
        pc.shadow_tris += st.tris;
    }
    return !occluded;
}

// Light-sampled emission reaching the surface sd, MIS-weighted against the
// cosine-sampled bounce; returns Le * cos / pdf * w (cosine-weighted, no BSDF).
template <int STACK, int BLOCK, bool COUNT, bool MAT>
__device__ __forceinline__ f3 emitter_nee(const DevScene &S, const Shading &sd, Rng &rng, uint32_t *col,
                                          PathCounters &pc) {
    float u0 = rng.next_1d(), u1 = rng.next_1d(), u2 = rng.next_1d();
    EmitterSample es;
    if (!sample_emitter<MAT>(S, sd.origin_front, u0, u1, u2, es)) return mk(0.0f, 0.0f, 0.0f);
    float cx = dot(sd.frame.n, es.dir);
    if (!(cx > 0.0f)) return mk(0.0f, 0.0f, 0.0f);
    if (!visible_to<STACK, BLOCK, COUNT, MAT>(S, sd.origin_front, es.dir, es.dist * (1.0f - 1e-4f), col, pc))
        return mk(0.0f, 0.0f, 0.0f);
    float w = power_weight(es.pdf, cx * kInvPi);
    return es.le * (cx / es.pdf * w);
}

// Radiance leaving y1 (shading sd, reached by the first bounce) without its
// emission: trace_radiance (texture_space_path_tracer.slang:57-84) from its
// sun term at y1 on, for the remaining max_bounces - 1 bounces.  With nee,
// emission is also light-sampled at every vertex that has a successor, and
// bounce-sampled emission carries the matching MIS weight.
template <int STACK, int BLOCK, bool COUNT, bool MAT>
__device__ f3 suffix_radiance(const DevScene &S, Shading sd, uint32_t max_bounces, bool nee, Rng &rng,
                              uint32_t *col, PathCounters &pc) {
    const Light &L = S.light;
    f3 T = mk(1.0f, 1.0f, 1.0f), rad = mk(0.0f, 0.0f, 0.0f);
    for (uint32_t bounce = 0;; ++bounce) {
        float ux = rng.next_1d(), uy = rng.next_1d();
        f3 ld = cone_sample(L.sun_dir, L.cos_half, ux, uy);
        float ndl = fmaxf(0.0f, dot(sd.frame.n, ld));
        f3 f = sd.diffuse * kInvPi;
        if (L.sun_on && ndl > 0.0f && shadow_visible<STACK, BLOCK, COUNT, MAT>(S, sd.origin_front, ld, col, pc))
            rad = rad + T * ((f * L.sun_illum) * ndl);
        if (bounce + 1 >= max_bounces) break;
        if (nee) rad = rad + T * (f * emitter_nee<STACK, BLOCK, COUNT, MAT>(S, sd, rng, col, pc));
        float ex = rng.next_1d(), ey = rng.next_1d();
        f3 sdir;
        float pdf;
        cosine_sample(ex, ey, sdir, pdf);
        if (pdf <= 1e-6f) break;
        float az = fabsf(sdir.z);
        T = T * mk(f.x * az / pdf, f.y * az / pdf, f.z * az / pdf);
        Ray r;
        r.o = sd.origin_front;
        r.d = to_global(sd.frame, sdir);
        r.tmin = 0.0f;
        r.tmax = 1000.0f;
        ray_setup(r);
        Hit h{r.tmax, 0.0f, 0.0f, kInvalid};
        TraversalStats st{0, 0};
        bool hit = traverse<STACK, BLOCK, false, COUNT, MAT>(S, r, h, col, st);
        if (COUNT) {
            pc.extend++;
            pc.nodes += st.nodes;
            pc.tris += st.tris;
        }
        if (!hit) {
            rad = rad + T * miss_radiance(L, r.d);
            break;
        }
        f3 dir = r.d;
        shading_data<MAT>(S, h.id, h.b1, h.b2, -dir, sd);
        if (sd.emissive.x > 0.0f || sd.emissive.y > 0.0f || sd.emissive.z > 0.0f) {
            float w = nee ? power_weight(pdf, emitter_pdf(S, h.id, dir, h.t)) : 1.0f;
            rad = rad + T * sd.emissive * w;
        }
    }
    return rad;
}

struct FirstHit {
    float weight;  // az / pdf of the cosine-sampled direction (0: no direction)
    float pdf;
    float t;
    uint32_t tri;
    f3 dir;
};

// One cosine-sampled bounce from sd; on a hit fills y1's shading.
template <int STACK, int BLOCK, bool COUNT, bool MAT>
__device__ __forceinline__ bool first_bounce(const DevScene &S, const Shading &sd, float ex, float ey, FirstHit &fh,
                                             Shading &y1, uint32_t *col, PathCounters &pc) {
    f3 sdir;
    float pdf;
    cosine_sample(ex, ey, sdir, pdf);
    fh.weight = 0.0f;
    fh.pdf = pdf;
    if (pdf <= 1e-6f) return false;
    fh.weight = fabsf(sdir.z) / pdf;
    Ray r;
    r.o = sd.origin_front;
    r.d = to_global(sd.frame, sdir);
    r.tmin = 0.0f;
    r.tmax = 1000.0f;
    ray_setup(r);
    fh.dir = r.d;
    Hit h{r.tmax, 0.0f, 0.0f, kInvalid};
    TraversalStats st{0, 0};
    bool hit = traverse<STACK, BLOCK, false, COUNT, MAT>(S, r, h, col, st);
    if (COUNT) {
        pc.extend++;
        pc.nodes += st.nodes;
        pc.tris += st.tris;
    }
    fh.t = h.t;
    fh.tri = h.id;
    if (hit) shading_data<MAT>(S, h.id, h.b1, h.b2, -fh.dir, y1);
    return hit;
}

// Training records: record j draws a front texel uniformly (and, on a
// double-sided instance, a side), follows one plain path from it and records
// y1 (position, normal, albedo) with the suffix radiance as target.  In the
// texel-query form it records x0 instead, with the indirect irradiance of
// the path as target: (az/pdf) times the sky on a miss, else times the
// suffix radiance of y1 plus y1's emission (MIS-weighted against the light
// sampling x0 does when emitters are light-sampled).
template <int STACK, int BLOCK, bool COUNT, bool MAT>
__device__ __forceinline__ void train_path(const DevScene &S, const DevLayout &Ly, uint32_t iteration, uint32_t spp,
                                           uint32_t max_bounces, uint32_t n, uint32_t flags, const LearnedDev &B,
                                           uint32_t j, uint32_t *col, PathCounters &pc) {
    if (j < n) {
        bool nee = (flags & NAN_LEARNED_EMITTER_NEE) && S.emitter_count > 0;
        bool at_x0 = (flags & NAN_LEARNED_TEXEL_QUERY) != 0;
        TrainStart st = train_start(Ly, j, iteration);
        Rng rng;
        TexelSurface ts = texel_surface(Ly, st.w, iteration, spp, rng);
        bool back = (ts.side.y & 1u) && st.back;
        rng.state = st.rng;
        f3 nf = world_flat_normal(S, ts.tri);
        Shading sd0, y1;
        shading_data<MAT>(S, ts.tri, ts.b1, ts.b2, back ? -nf : nf, sd0);
        float ex = rng.next_1d(), ey = rng.next_1d();
        FirstHit fh;
        bool valid = first_bounce<STACK, BLOCK, COUNT, MAT>(S, sd0, ex, ey, fh, y1, col, pc);
        f3 target = mk(0.0f, 0.0f, 0.0f);
        if (at_x0) {
            if (valid) {
                target = suffix_radiance<STACK, BLOCK, COUNT, MAT>(S, y1, max_bounces, nee, rng, col, pc);
                float wm = nee ? power_weight(fh.pdf, emitter_pdf(S, fh.tri, fh.dir, fh.t)) : 1.0f;
                target = (target + y1.emissive * wm) * fh.weight;
            } else if (fh.weight > 0.0f) {
                target = miss_radiance(S.light, fh.dir) * fh.weight;
            }
            valid = true;
            put3(B.tr_pos, j, sd0.origin_front);
            put3(B.tr_nrm, j, sd0.frame.n);
            put3(B.tr_alb, j, sd0.diffuse);
        } else if (valid) {
            target = suffix_radiance<STACK, BLOCK, COUNT, MAT>(S, y1, max_bounces, nee, rng, col, pc);
            put3(B.tr_pos, j, y1.origin_front);
            put3(B.tr_nrm, j, y1.frame.n);
            put3(B.tr_alb, j, y1.diffuse);
        } else {
            put3(B.tr_pos, j, mk(0.0f, 0.0f, 0.0f));
            put3(B.tr_nrm, j, mk(0.0f, 0.0f, 0.0f));
            put3(B.tr_alb, j, mk(0.0f, 0.0f, 0.0f));
        }
        put3(B.tr_target, j, target);
        B.tr_valid[j] = valid ? 1u : 0u;
    }
}

template <int STACK, int BLOCK, bool COUNT, bool MAT>
__global__ __launch_bounds__(BLOCK) void learned_train_paths(DevScene S, DevLayout Ly, uint32_t iteration,
                                                              uint32_t spp, uint32_t max_bounces, uint32_t n,
                                                              uint32_t flags, LearnedDev B,
                                                              unsigned long long *counters) {
    if (B.frame) iteration = B.frame[0];
    __shared__ uint32_t stack[STACK * BLOCK];
    PathCounters pc{0, 0, 0, 0, 0, 0, 0};
    train_path<STACK, BLOCK, COUNT, MAT>(S, Ly, iteration, spp, max_bounces, n, flags, B,
                                         blockIdx.x * BLOCK + threadIdx.x, stack + threadIdx.x, pc);
    if (COUNT) flush_counters(pc, counters);
}

__device__ __forceinline__ void texel_slots(const DevLayout &Ly, uint64_t w, uint32_t &front, uint32_t &back,
                                            bool &two_sided) {
    uint32_t f = Ly.texel_face[w];
    uint32_t inst = Ly.face_inst[f];
    uint4 ii = Ly.instances[inst];
    uint2 side = Ly.sides[inst];
    uint32_t local = (uint32_t)(w - Ly.front_prefix[inst]);
    front = ii.z + local;
    back = side.x + local;
    two_sided = (side.y & 1u) != 0;
}

// Every texel sample: sun term at x0, (with nee) light-sampled emission, and
// the first bounce; queries for the cache go to q = slot * spp + k (weight 0:
// no query).
template <int STACK, int BLOCK, bool COUNT, bool MAT>
__device__ __forceinline__ void texel_first_bounce(const DevScene &S, const DevLayout &Ly, uint32_t iteration,
                                                   uint32_t spp, uint32_t flags, const LearnedDev &B, uint64_t w,
                                                   uint32_t *col, PathCounters &pc) {
    if (w < Ly.total_front) {
        const Light &L = S.light;
        bool nee = (flags & NAN_LEARNED_EMITTER_NEE) && S.emitter_count > 0;
        Rng rng;
        TexelSurface ts = texel_surface(Ly, w, iteration, spp, rng);
        uint32_t nspp = spp > 1u ? spp : 1u;
        f3 nf = world_flat_normal(S, ts.tri);
        for (int side = 0; side < 2; ++side) {
            if (side == 1 && !(ts.side.y & 1u)) break;
            uint32_t slot = side ? ts.side.x + ts.local : ts.global_texel;
            Shading sd, y1;
            shading_data<MAT>(S, ts.tri, ts.b1, ts.b2, side ? -nf : nf, sd);
            f3 acc = mk(0.0f, 0.0f, 0.0f);
            for (uint32_t k = 0; k < nspp; ++k) {
                float ux = rng.next_1d(), uy = rng.next_1d();
                f3 ld = cone_sample(L.sun_dir, L.cos_half, ux, uy);
                float ndl = fmaxf(0.0f, dot(sd.frame.n, ld));
                if (L.sun_on && ndl > 0.0f && shadow_visible<STACK, BLOCK, COUNT, MAT>(S, sd.origin_front, ld, col, pc))
                    acc = acc + L.sun_illum * ndl;
                if (nee) acc = acc + emitter_nee<STACK, BLOCK, COUNT, MAT>(S, sd, rng, col, pc);
                float ex = rng.next_1d(), ey = rng.next_1d();
                FirstHit fh;
                uint64_t q = (uint64_t)slot * nspp + k;
                bool hit = first_bounce<STACK, BLOCK, COUNT, MAT>(S, sd, ex, ey, fh, y1, col, pc);
                if (hit) {
                    if (y1.emissive.x > 0.0f || y1.emissive.y > 0.0f || y1.emissive.z > 0.0f) {
                        float wm = nee ? power_weight(fh.pdf, emitter_pdf(S, fh.tri, fh.dir, fh.t)) : 1.0f;
                        acc = acc + y1.emissive * (fh.weight * wm);
                    }
                    put3(B.q_pos, q, y1.origin_front);
                    put3(B.q_nrm, q, y1.frame.n);
                    put3(B.q_alb, q, y1.diffuse);
                    B.q_w[q] = fh.weight;
                } else {
                    if (fh.weight > 0.0f) acc = acc + miss_radiance(L, fh.dir) * fh.weight;
                    put3(B.q_pos, q, mk(0.0f, 0.0f, 0.0f));
                    put3(B.q_nrm, q, mk(0.0f, 0.0f, 0.0f));
                    put3(B.q_alb, q, mk(0.0f, 0.0f, 0.0f));
                    B.q_w[q] = 0.0f;
                }
            }
            B.partial[slot] = make_float4(acc.x, acc.y, acc.z, 0.0f);
        }
    }
}

template <int STACK, int BLOCK, bool COUNT, bool MAT>
__global__ __launch_bounds__(BLOCK) void learned_first_bounce(DevScene S, DevLayout Ly, uint32_t iteration,
                                                               uint32_t spp, uint32_t flags, LearnedDev B,
                                                               unsigned long long *counters) {
    if (B.frame) iteration = B.frame[0];
    __shared__ uint32_t stack[STACK * BLOCK];
    PathCounters pc{0, 0, 0, 0, 0, 0, 0};
    texel_first_bounce<STACK, BLOCK, COUNT, MAT>(S, Ly, iteration, spp, flags, B,
                                                 (uint64_t)blockIdx.x * BLOCK + threadIdx.x, stack + threadIdx.x, pc);
    if (COUNT) flush_counters(pc, counters);
}

// One launch for a call's texel samples (blocks [0, texel_blocks)) and its
// training paths (the blocks after them): both are latency bound, and one
// launch costs the longer of the two instead of their sum.
template <int STACK, int BLOCK, bool MAT>
__global__ __launch_bounds__(BLOCK) void learned_trace(DevScene S, DevLayout Ly, uint32_t iteration, uint32_t spp,
                                                        uint32_t max_bounces, uint32_t n_train, uint32_t flags,
                                                        uint32_t texel_blocks, LearnedDev B) {
    if (B.frame) iteration = B.frame[0];
    __shared__ uint32_t stack[STACK * BLOCK];
    PathCounters pc{0, 0, 0, 0, 0, 0, 0};
    if (blockIdx.x < texel_blocks)
        texel_first_bounce<STACK, BLOCK, false, MAT>(S, Ly, iteration, spp, flags, B,
                                                     (uint64_t)blockIdx.x * BLOCK + threadIdx.x, stack + threadIdx.x,
                                                     pc);
    else
        train_path<STACK, BLOCK, false, MAT>(S, Ly, iteration, spp, max_bounces, n_train, flags, B,
                                             (blockIdx.x - texel_blocks) * BLOCK + threadIdx.x, stack + threadIdx.x,
                                             pc);
}

// The call's training records into the previous-frame buffer (same layout,
// cap records per array) for the next call's training.
__global__ void learned_keep_records(LearnedDev B, uint32_t n, uint32_t cap) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float *o = B.pv_rec;
    for (int a = 0; a < 3; ++a) {
        o[3 * (size_t)j + a] = B.tr_pos[3 * (size_t)j + a];
        o[3 * ((size_t)cap + j) + a] = B.tr_nrm[3 * (size_t)j + a];
        o[3 * (2 * (size_t)cap + j) + a] = B.tr_alb[3 * (size_t)j + a];
        o[3 * (3 * (size_t)cap + j) + a] = B.tr_target[3 * (size_t)j + a];
    }
    reinterpret_cast<uint32_t *>(o + 12 * (size_t)cap)[j] = B.tr_valid[j];
}

// Texel-query form: every texel sample takes the sun term at x0 and (with
// nee) MIS-weighted light-sampled emission, and queries the cache at x0
// itself for the rest of its irradiance; no bounce ray is traced.
template <int STACK, int BLOCK, bool COUNT, bool MAT>
__global__ __launch_bounds__(BLOCK) void learned_texel_direct(DevScene S, DevLayout Ly, uint32_t iteration,
                                                               uint32_t spp, uint32_t flags, LearnedDev B,
                                                               unsigned long long *counters) {
    if (B.frame) iteration = B.frame[0];
    __shared__ uint32_t stack[STACK * BLOCK];
    uint32_t *col = stack + threadIdx.x;
    PathCounters pc{0, 0, 0, 0, 0, 0, 0};
    uint64_t w = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (w < Ly.total_front) {
        const Light &L = S.light;
        bool nee = (flags & NAN_LEARNED_EMITTER_NEE) && S.emitter_count > 0;
        Rng rng;
        TexelSurface ts = texel_surface(Ly, w, iteration, spp, rng);
        uint32_t nspp = spp > 1u ? spp : 1u;
        f3 nf = world_flat_normal(S, ts.tri);
        for (int side = 0; side < 2; ++side) {
            if (side == 1 && !(ts.side.y & 1u)) break;
            uint32_t slot = side ? ts.side.x + ts.local : ts.global_texel;
            Shading sd;
            shading_data<MAT>(S, ts.tri, ts.b1, ts.b2, side ? -nf : nf, sd);
            f3 acc = mk(0.0f, 0.0f, 0.0f);
            for (uint32_t k = 0; k < nspp; ++k) {
                float ux = rng.next_1d(), uy = rng.next_1d();
                f3 ld = cone_sample(L.sun_dir, L.cos_half, ux, uy);
                float ndl = fmaxf(0.0f, dot(sd.frame.n, ld));
                if (L.sun_on && ndl > 0.0f && shadow_visible<STACK, BLOCK, COUNT, MAT>(S, sd.origin_front, ld, col, pc))
                    acc = acc + L.sun_illum * ndl;
                if (nee) acc = acc + emitter_nee<STACK, BLOCK, COUNT, MAT>(S, sd, rng, col, pc);
                uint64_t q = (uint64_t)slot * nspp + k;
                put3(B.q_pos, q, sd.origin_front);
                put3(B.q_nrm, q, sd.frame.n);
                put3(B.q_alb, q, sd.diffuse);
                B.q_w[q] = 1.0f;
            }
            B.partial[slot] = make_float4(acc.x, acc.y, acc.z, 0.0f);
        }
    }
    if (COUNT) flush_counters(pc, counters);
}

template <bool COUNT>
__global__ void learned_finish(DevLayout Ly, LearnedDev B, float4 *payload, uint32_t spp, uint32_t reset,
                               uint32_t flags, unsigned long long *counters) {
    if (B.frame) reset = B.frame[1];
    PathCounters pc{0, 0, 0, 0, 0, 0, 0};
    uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < Ly.total_front) {
        uint32_t slots[2];
        bool two;
        texel_slots(Ly, w, slots[0], slots[1], two);
        uint32_t nspp = spp > 1u ? spp : 1u;
        for (int side = 0; side < (two ? 2 : 1); ++side) {
            uint32_t slot = slots[side];
            float4 p = B.partial[slot];
            f3 e = mk(p.x, p.y, p.z);
            for (uint32_t k = 0; k < nspp; ++k) {
                uint64_t q = (uint64_t)slot * nspp + k;
                float wq = B.q_w[q];
                if (wq > 0.0f) {
                    float4 c = B.q_pred[q];
                    e = e + mk(fmaxf(c.x, 0.0f), fmaxf(c.y, 0.0f), fmaxf(c.z, 0.0f)) * wq;
                }
            }
            float fs = (float)nspp;
            if (flags & NAN_LEARNED_RAMPED_MEAN) store_ramped(payload, slot, mk(e.x / fs, e.y / fs, e.z / fs), reset);
            else store_accumulated(payload, slot, mk(e.x / fs, e.y / fs, e.z / fs), reset);
            if (COUNT) pc.slots++;
        }
    }
    if (COUNT) flush_counters(pc, counters);
}

template <typename T>
int grow(T **p, size_t bytes) {
    hipFree(*p);
    *p = nullptr;
    HIP_TRY(hipMalloc((void **)p, bytes), nullptr, 0);
    return 0;
}

// Grows the learned forms' record and query buffers; refused while `s` is
// being captured into a graph (growing allocates and synchronises).
int reserve(NanTracer *t, uint32_t n_train, uint64_t n_query, hipStream_t s) {
    if (!t->learned) {
        if (stream_capturing(s))
            return fail(nullptr, 0, "learned forms: buffers not reserved; call nan_tst_learned_reserve before "
                                    "graph capture");
        t->learned = new LearnedBuffers();
    }
    LearnedBuffers &B = *t->learned;
    LearnedDev &d = B.d;
    if (n_train > B.cap_train || n_query > B.cap_query) {
        if (stream_capturing(s))
            return fail(nullptr, 0, "learned forms: %u training paths / %llu queries exceed the reserved %u / %llu; "
                                    "call nan_tst_learned_reserve before graph capture", n_train,
                        (unsigned long long)n_query, B.cap_train, (unsigned long long)B.cap_query);
        HIP_TRY(hipDeviceSynchronize(), nullptr, 0);
    }
    if (n_train > B.cap_train) {
        size_t n = n_train;
        B.cap_train = 0;
        if (grow(&d.tr_pos, n * 12) || grow(&d.tr_nrm, n * 12) || grow(&d.tr_alb, n * 12) ||
            grow(&d.tr_target, n * 12) || grow(&d.tr_valid, n * 4))
            return -1;
        B.cap_train = n_train;
    }
    if (n_query > B.cap_query) {
        size_t n = n_query;
        B.cap_query = 0;
        if (grow(&d.q_pos, n * 12) || grow(&d.q_nrm, n * 12) || grow(&d.q_alb, n * 12) || grow(&d.q_w, n * 4) ||
            grow(&d.q_pred, n * 16))
            return -1;
        B.cap_query = n_query;
    }
    if (!d.partial) {
        if (stream_capturing(s)) return fail(nullptr, 0, "learned forms: buffers not reserved before graph capture");
        if (grow(&d.partial, std::max<uint64_t>(t->total_payload, 1) * 16)) return -1;
    }
    return 0;
}

// NAN_LEARNED_OVERLAP's working set: the previous call's records (zeroed, so
// the first call trains on nothing), the cache's parameter snapshot and the
// side stream with its events.
int reserve_overlap(NanTracer *t, NanNrc *c, uint32_t n_train, hipStream_t s) {
    if (nrc_reserve_snapshot(c, s)) return -1;
    LearnedBuffers &B = *t->learned;
    if (n_train > B.cap_prev) {
        if (stream_capturing(s))
            return fail(nullptr, 0, "learned forms: %u training paths exceed the %u reserved for overlapped "
                                    "training; call nan_tst_learned_reserve before graph capture", n_train, B.cap_prev);
        HIP_TRY(hipDeviceSynchronize(), nullptr, 0);
        B.cap_prev = 0;
        if (grow(&B.d.pv_rec, (size_t)n_train * 13 * 4)) return -1;
        HIP_TRY(hipMemset(B.d.pv_rec, 0, (size_t)n_train * 13 * 4), nullptr, 0);
        HIP_TRY(hipDeviceSynchronize(), nullptr, 0);
        B.cap_prev = n_train;
    }
    if (!B.side_stream) {
        if (stream_capturing(s))
            return fail(nullptr, 0, "learned forms: overlap stream not created; call nan_tst_learned_reserve before "
                                    "graph capture");
        HIP_TRY(hipStreamCreateWithFlags(&B.side_stream, hipStreamNonBlocking), nullptr, 0);
        for (hipEvent_t *e : {&B.ev_fork, &B.ev_snap, &B.ev_traced, &B.ev_join})
            HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming), nullptr, 0);
    }
    return 0;
}

LearnedIO learned_io(const LearnedDev &B) {
    return LearnedIO{B.tr_pos, B.tr_nrm, B.tr_alb, B.tr_target, B.tr_valid, B.q_pos, B.q_nrm, B.q_alb, B.q_w, B.q_pred,
                     B.frame};
}

// The learned forms on the wavefront queues (same results as launch() below).
// Training batches up to this many paths trace faster as one thread per path
// (one launch) than through the queues (a dozen launches per bounce whose
// fixed cost dominates small batches); both write the same records.
constexpr uint32_t kQueuedTrainPaths = 524288;

int launch_wavefront(NanTracer *t, NanNrc *c, float4 *payload, uint32_t iteration, uint32_t bounces, uint32_t reset,
                     uint32_t n_train, uint32_t flags, hipStream_t s) {
    LearnedIO io = learned_io(t->learned->d);
    bool at_x0 = (flags & NAN_LEARNED_TEXEL_QUERY) != 0;
    if (n_train) {
        if (n_train < kQueuedTrainPaths && t->scene->stack_depth != 0) {
            StageTimer tm(t, s, kStageLearnedPaths);
            const DevScene &S = t->scene->dev;
            DevLayout Ly = dev_layout(t);
            const LearnedDev &B = t->learned->d;
#define NAN_TRAIN_PATHS(D, M)                                                                                   \
    hipLaunchKernelGGL((learned_train_paths<D, kTraceBlock, false, M>), dim3(grid_for(n_train, kTraceBlock)),     \
                       dim3(kTraceBlock), 0, s, S, Ly, iteration, 1u, bounces, n_train, flags, B,                \
                       (unsigned long long *)nullptr)
            const bool mat = S.mat_features != 0u;
            if (t->scene->stack_depth <= 32) {
                if (mat) NAN_TRAIN_PATHS(32, true);
                else NAN_TRAIN_PATHS(32, false);
            } else {
                if (mat) NAN_TRAIN_PATHS(64, true);
                else NAN_TRAIN_PATHS(64, false);
            }
#undef NAN_TRAIN_PATHS
            if (check_launch("learned training paths")) return -1;
        } else if (wavefront_learned_train_paths(t, io, iteration, n_train, bounces, at_x0, s)) {
            return -1;
        }
        StageTimer tm(t, s, kStageCacheTrain);
        if (nrc_train_step(c, io.tr_pos, io.tr_nrm, io.tr_alb, io.tr_target, io.tr_valid, n_train, s)) return -1;
    }
    for (uint32_t side = 0; side < 2; ++side) {
        uint32_t count = 0;
        if (wavefront_learned_side(t, io, iteration, side, at_x0, &count, s)) return -1;
        if (count == 0) continue;
        {
            StageTimer tm(t, s, kStageCacheQuery);
            if (nrc_predict(c, io.q_pos, io.q_nrm, io.q_alb, count, io.q_pred, s)) return -1;
        }
        if (wavefront_learned_finish(t, io, payload, side, reset, at_x0, (flags & NAN_LEARNED_RAMPED_MEAN) != 0, s))
            return -1;
    }
    return 0;
}

// NAN_LEARNED_OVERLAP: the cache trains on the previous call's records on
// the side stream while this call traces its texel samples and training
// paths in one launch; the cache queries wait for the training.
template <int STACK, bool MAT>
int launch_overlap(NanTracer *t, NanNrc *c, float4 *payload, uint32_t iteration, uint32_t nspp, uint32_t bounces,
                   uint32_t reset, uint32_t n_train, uint32_t flags, hipStream_t s) {
    const DevScene &S = t->scene->dev;
    DevLayout Ly = dev_layout(t);
    LearnedBuffers &LB = *t->learned;
    const LearnedDev &B = LB.d;
    hipStream_t side = LB.side_stream;
    HIP_TRY(hipEventRecord(LB.ev_fork, s), nullptr, 0);
    HIP_TRY(hipStreamWaitEvent(side, LB.ev_fork, 0), nullptr, 0);
    // the queries read the parameters as the previous call left them, copied
    // before this call's steps change them
    if (nrc_snapshot(c, side)) return -1;
    HIP_TRY(hipEventRecord(LB.ev_snap, side), nullptr, 0);
    // the previous call's records; zeros (no valid record, so a step that
    // changes no parameter) before the first overlapped call
    const uint32_t cap = LB.cap_prev;
    if (n_train) {
        StageTimer tm(t, side, kStageCacheTrain);
        const float *pv = B.pv_rec;
        if (nrc_train_step(c, pv, pv + 3 * (size_t)cap, pv + 6 * (size_t)cap, pv + 9 * (size_t)cap,
                           reinterpret_cast<const uint32_t *>(pv + 12 * (size_t)cap), n_train, side))
            return -1;
    }
    {
        StageTimer tm(t, s, kStageFirstBounce);
        const uint32_t texel_blocks = grid_for(t->total_front, kTraceBlock);
        const uint32_t train_blocks = grid_for(n_train, kTraceBlock);
        hipLaunchKernelGGL((learned_trace<STACK, kTraceBlock, MAT>), dim3(texel_blocks + train_blocks),
                           dim3(kTraceBlock), 0, s, S, Ly, iteration, nspp, bounces, n_train, flags, texel_blocks, B);
    }
    HIP_TRY(hipEventRecord(LB.ev_traced, s), nullptr, 0);
    HIP_TRY(hipStreamWaitEvent(side, LB.ev_traced, 0), nullptr, 0);
    if (n_train)
        hipLaunchKernelGGL(learned_keep_records, dim3(grid_for(n_train, 256)), dim3(256), 0, side, B, n_train, cap);
    HIP_TRY(hipEventRecord(LB.ev_join, side), nullptr, 0);
    HIP_TRY(hipStreamWaitEvent(s, LB.ev_snap, 0), nullptr, 0);
    {
        StageTimer tm(t, s, kStageCacheQuery);
        uint64_t nq = t->total_payload * nspp;
        if (nrc_predict(c, B.q_pos, B.q_nrm, B.q_alb, (uint32_t)nq, B.q_pred, s, true)) return -1;
    }
    {
        StageTimer tm(t, s, kStageFinish);
        hipLaunchKernelGGL((learned_finish<false>), dim3(grid_for(t->total_front, kFlatBlock)), dim3(kFlatBlock), 0,
                           s, Ly, B, payload, nspp, reset, flags, (unsigned long long *)nullptr);
    }
    HIP_TRY(hipStreamWaitEvent(s, LB.ev_join, 0), nullptr, 0);
    return check_launch("learned tracer (overlapped training)");
}

template <int STACK, bool MAT>
int launch_mat(NanTracer *t, NanNrc *c, float4 *payload, uint32_t iteration, uint32_t spp, uint32_t bounces,
               uint32_t reset, uint32_t n_train, uint32_t flags, unsigned long long *counters, hipStream_t s) {
    const DevScene &S = t->scene->dev;
    DevLayout Ly = dev_layout(t);
    const LearnedDev &B = t->learned->d;
    uint32_t nspp = spp > 1u ? spp : 1u;
    if ((flags & NAN_LEARNED_OVERLAP) && !counters && !(flags & NAN_LEARNED_TEXEL_QUERY))
        return launch_overlap<STACK, MAT>(t, c, payload, iteration, nspp, bounces, reset, n_train, flags, s);
    {
        StageTimer tm(t, s, kStageLearnedPaths);
        if (n_train) {
            if (counters)
                hipLaunchKernelGGL((learned_train_paths<STACK, kTraceBlock, true, MAT>), dim3(grid_for(n_train, kTraceBlock)),
                                   dim3(kTraceBlock), 0, s, S, Ly, iteration, nspp, bounces, n_train, flags, B, counters);
            else
                hipLaunchKernelGGL((learned_train_paths<STACK, kTraceBlock, false, MAT>),
                                   dim3(grid_for(n_train, kTraceBlock)), dim3(kTraceBlock), 0, s, S, Ly, iteration,
                                   nspp, bounces, n_train, flags, B, counters);
        }
    }
    if (n_train) {
        StageTimer tm(t, s, kStageCacheTrain);
        if (nrc_train_step(c, B.tr_pos, B.tr_nrm, B.tr_alb, B.tr_target, B.tr_valid, n_train, s)) return -1;
    }
    {
        StageTimer tm(t, s, kStageFirstBounce);
        uint32_t grid = grid_for(t->total_front, kTraceBlock);
        bool at_x0 = (flags & NAN_LEARNED_TEXEL_QUERY) != 0;
        if (counters) {
            if (at_x0)
                hipLaunchKernelGGL((learned_texel_direct<STACK, kTraceBlock, true, MAT>), dim3(grid), dim3(kTraceBlock), 0,
                                   s, S, Ly, iteration, nspp, flags, B, counters);
            else
                hipLaunchKernelGGL((learned_first_bounce<STACK, kTraceBlock, true, MAT>), dim3(grid), dim3(kTraceBlock), 0,
                                   s, S, Ly, iteration, nspp, flags, B, counters);
        } else {
            if (at_x0)
                hipLaunchKernelGGL((learned_texel_direct<STACK, kTraceBlock, false, MAT>), dim3(grid), dim3(kTraceBlock),
                                   0, s, S, Ly, iteration, nspp, flags, B, counters);
            else
                hipLaunchKernelGGL((learned_first_bounce<STACK, kTraceBlock, false, MAT>), dim3(grid), dim3(kTraceBlock),
                                   0, s, S, Ly, iteration, nspp, flags, B, counters);
        }
    }
    {
        StageTimer tm(t, s, kStageCacheQuery);
        uint64_t nq = t->total_payload * nspp;
        if (nrc_predict(c, B.q_pos, B.q_nrm, B.q_alb, (uint32_t)nq, B.q_pred, s)) return -1;
    }
    {
        StageTimer tm(t, s, kStageFinish);
        uint32_t grid = grid_for(t->total_front, kFlatBlock);
        if (counters)
            hipLaunchKernelGGL((learned_finish<true>), dim3(grid), dim3(kFlatBlock), 0, s, Ly, B, payload, nspp,
                               reset, flags, counters);
        else
            hipLaunchKernelGGL((learned_finish<false>), dim3(grid), dim3(kFlatBlock), 0, s, Ly, B, payload, nspp,
                               reset, flags, counters);
    }
    return check_launch("learned tracer");
}

template <int STACK>
int launch(NanTracer *t, NanNrc *c, float4 *payload, uint32_t iteration, uint32_t spp, uint32_t bounces,
           uint32_t reset, uint32_t n_train, uint32_t flags, unsigned long long *counters, hipStream_t s) {
    if (t->scene->dev.mat_features)
        return launch_mat<STACK, true>(t, c, payload, iteration, spp, bounces, reset, n_train, flags, counters, s);
    return launch_mat<STACK, false>(t, c, payload, iteration, spp, bounces, reset, n_train, flags, counters, s);
}

}  // namespace

namespace nanx {
void learned_release(NanTracer *t) {
    delete t->learned;
    t->learned = nullptr;
}
}  // namespace nanx

extern "C" {

NAN_API int nan_tst_execute_learned(NanTracer *tracer, NanNrc *nrc, void *payload, uint32_t iteration,
                                    uint32_t samples_per_texel, uint32_t max_bounces, uint32_t reset,
                                    uint32_t train_samples, uint32_t flags, uint64_t *counters, void *stream) {
    if (!tracer || !nrc || !payload) return fail(nullptr, 0, "nan_tst_execute_learned: null argument");
    if (tracer->scene->batch || tracer->scene_count != 1)
        return fail(nullptr, 0, "nan_tst_execute_learned: scene batches are not supported");
    if (tracer->scene->roulette)
        return fail(nullptr, 0, "nan_tst_execute_learned: Russian roulette is not supported by the learned forms");
    if (tracer->scene->stack_depth == 0)
        return fail(nullptr, 0, "nan_tst_execute_learned: BVH too deep for the per-texel kernels");
    if (max_bounces == 0) return fail(nullptr, 0, "nan_tst_execute_learned: max_bounces must be at least 1");
    if (flags & ~(uint32_t)(NAN_LEARNED_EMITTER_NEE | NAN_LEARNED_RAMPED_MEAN | NAN_LEARNED_TEXEL_QUERY |
                            NAN_LEARNED_RNG_PCG | NAN_LEARNED_OVERLAP))
        return fail(nullptr, 0, "nan_tst_execute_learned: unknown flags");
    if ((flags & NAN_LEARNED_OVERLAP) && (flags & NAN_LEARNED_TEXEL_QUERY))
        return fail(nullptr, 0, "nan_tst_execute_learned: overlapped training is a one-bounce-form option");
    tracer->rng_mode = (flags & NAN_LEARNED_RNG_PCG) ? NAN_RNG_PCG : NAN_RNG_REFERENCE;
    if (tracer->total_front == 0) return 0;
    uint32_t nspp = samples_per_texel > 1u ? samples_per_texel : 1u;
    uint64_t nq = tracer->total_payload * nspp;
    if (nq >= (1ull << 31)) return fail(nullptr, 0, "nan_tst_execute_learned: too many cache queries per frame");
    hipStream_t s = as_stream(stream);
    if (reserve(tracer, train_samples, nq, s) || nrc_reserve(nrc, (uint32_t)nq, s) ||
        nrc_reserve_train(nrc, train_samples, s) ||
        ((flags & NAN_LEARNED_OVERLAP) && reserve_overlap(tracer, nrc, train_samples, s)))
        return -1;
    unsigned long long *c = reinterpret_cast<unsigned long long *>(counters);
    // device-frame mode: iteration and reset come from the tracer's device
    // frame, which the call advances at its end (a captured call replays
    // successive iterations)
    tracer->learned->d.frame = wavefront_frame(tracer);
    // wavefront mode runs the forms on the tracer's queues when they need no
    // per-texel extras: 1 sample per texel, no emitter light sampling, no
    // counters, training paths within the tracer's path slots
    bool nee = (flags & NAN_LEARNED_EMITTER_NEE) && tracer->scene->dev.emitter_count > 0;
    int rc;
    if (tracer->mode == NAN_TRACER_WAVEFRONT && nspp == 1 && !nee && !c && !(flags & NAN_LEARNED_OVERLAP) &&
        train_samples <= tracer->total_front)
        rc = launch_wavefront(tracer, nrc, (float4 *)payload, iteration, max_bounces, reset, train_samples, flags, s);
    else if (tracer->scene->stack_depth <= 32)
        rc = launch<32>(tracer, nrc, (float4 *)payload, iteration, nspp, max_bounces, reset, train_samples, flags, c,
                        s);
    else
        rc = launch<64>(tracer, nrc, (float4 *)payload, iteration, nspp, max_bounces, reset, train_samples, flags, c,
                        s);
    if (rc == 0) wavefront_advance_frame(tracer, s);
    return rc;
}

NAN_API int nan_tst_learned_reserve(NanTracer *tracer, NanNrc *nrc, uint32_t samples_per_texel,
                                    uint32_t train_samples) {
    if (!tracer || !nrc) return fail(nullptr, 0, "nan_tst_learned_reserve: null argument");
    uint32_t nspp = samples_per_texel > 1u ? samples_per_texel : 1u;
    uint64_t nq = tracer->total_payload * nspp;
    if (nq >= (1ull << 31)) return fail(nullptr, 0, "nan_tst_learned_reserve: too many cache queries per frame");
    if (stream_capturing(nullptr))
        return fail(nullptr, 0, "nan_tst_learned_reserve: called during graph capture");
    if (reserve(tracer, train_samples, nq, nullptr) || nrc_reserve(nrc, (uint32_t)std::max<uint64_t>(nq, 1), nullptr) ||
        nrc_reserve_train(nrc, std::max(train_samples, 1u), nullptr) ||
        reserve_overlap(tracer, nrc, std::max(train_samples, 1u), nullptr))
        return -1;
    return 0;
}

NAN_API int nan_tst_learned_records(NanTracer *tracer, uint32_t iteration, uint32_t n, uint32_t max_bounces,
                                    uint32_t flags, float *pos, float *nrm, float *alb, float *target,
                                    uint32_t *valid, void *stream) {
    if (!tracer || (n && (!pos || !nrm || !alb || !target || !valid)))
        return fail(nullptr, 0, "nan_tst_learned_records: null argument");
    if (tracer->scene->batch || tracer->scene_count != 1)
        return fail(nullptr, 0, "nan_tst_learned_records: scene batches are not supported");
    if (tracer->scene->stack_depth == 0)
        return fail(nullptr, 0, "nan_tst_learned_records: BVH too deep for the per-texel kernels");
    if (tracer->scene->roulette)
        return fail(nullptr, 0, "nan_tst_learned_records: Russian roulette is not supported by the learned forms");
    if (max_bounces == 0) return fail(nullptr, 0, "nan_tst_learned_records: max_bounces must be at least 1");
    if (flags & ~(uint32_t)(NAN_LEARNED_EMITTER_NEE | NAN_LEARNED_RAMPED_MEAN | NAN_LEARNED_TEXEL_QUERY |
                            NAN_LEARNED_RNG_PCG))
        return fail(nullptr, 0, "nan_tst_learned_records: unknown flags");
    tracer->rng_mode = (flags & NAN_LEARNED_RNG_PCG) ? NAN_RNG_PCG : NAN_RNG_REFERENCE;
    if (n == 0 || tracer->total_front == 0) return 0;
    hipStream_t s = as_stream(stream);
    if (reserve(tracer, n, 0, s)) return -1;
    tracer->learned->d.frame = nullptr;  // records of the given iteration
    const LearnedDev &B = tracer->learned->d;
    bool nee = (flags & NAN_LEARNED_EMITTER_NEE) && tracer->scene->dev.emitter_count > 0;
    if (tracer->mode == NAN_TRACER_WAVEFRONT && !nee && n <= tracer->total_front) {
        if (wavefront_learned_train_paths(tracer, learned_io(B), iteration, n, max_bounces,
                                          (flags & NAN_LEARNED_TEXEL_QUERY) != 0, s))
            return -1;
    } else {
        const DevScene &S = tracer->scene->dev;
        DevLayout Ly = dev_layout(tracer);
#define NAN_TRAIN_PATHS(D, M)                                                                                \
    hipLaunchKernelGGL((learned_train_paths<D, kTraceBlock, false, M>), dim3(grid_for(n, kTraceBlock)),        \
                       dim3(kTraceBlock), 0, s, S, Ly, iteration, 1u, max_bounces, n, flags, B, nullptr)
        const bool mat = S.mat_features != 0u;
        if (tracer->scene->stack_depth <= 32) {
            if (mat) NAN_TRAIN_PATHS(32, true);
            else NAN_TRAIN_PATHS(32, false);
        } else {
            if (mat) NAN_TRAIN_PATHS(64, true);
            else NAN_TRAIN_PATHS(64, false);
        }
#undef NAN_TRAIN_PATHS
        if (check_launch("learned training paths")) return -1;
    }
    size_t b3 = (size_t)n * 12;
    HIP_TRY(hipMemcpyAsync(pos, B.tr_pos, b3, hipMemcpyDeviceToDevice, s), nullptr, 0);
    HIP_TRY(hipMemcpyAsync(nrm, B.tr_nrm, b3, hipMemcpyDeviceToDevice, s), nullptr, 0);
    HIP_TRY(hipMemcpyAsync(alb, B.tr_alb, b3, hipMemcpyDeviceToDevice, s), nullptr, 0);
    HIP_TRY(hipMemcpyAsync(target, B.tr_target, b3, hipMemcpyDeviceToDevice, s), nullptr, 0);
    HIP_TRY(hipMemcpyAsync(valid, B.tr_valid, (size_t)n * 4, hipMemcpyDeviceToDevice, s), nullptr, 0);
    return 0;
}

}  // extern "C"
