// Device-side per-texel path functions shared by the megakernel and the
// wavefront pipeline (texture_space_path_tracer.slang:57-216).
#pragma once

#include "nan_internal.h"

// ---------------------------------------------------------------------------
// path tracing device functions
// ---------------------------------------------------------------------------
struct PathCounters {
    uint32_t extend, shadow, nodes, tris, slots, shadow_nodes, shadow_tris;
};

template <int STACK, int BLOCK, bool COUNT, bool MAT>
__device__ __forceinline__ bool shadow_visible(const DevScene &S, f3 origin, f3 dir, uint32_t *col,
                                               PathCounters &pc) {
    Ray r;
    r.o = origin;
    r.d = dir;
    r.tmin = 0.0f;
    r.tmax = 1000.0f;
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

// Russian roulette (an extension; the reference's paths always run to
// max_bounces): a path that would take another bounce survives with
// probability p = min(1, max component of its throughput) and carries T / p,
// so the estimate stays unbiased.  One draw, taken only when the path would
// continue, in the oracle's order.
__device__ __forceinline__ bool roulette_survives(f3 &T, Rng &rng) {
    float p = fminf(1.0f, fmaxf(fmaxf(T.x, T.y), T.z));
    float u = rng.next_1d();
    if (!(u < p)) return false;
    T = mk(T.x / p, T.y / p, T.z / p);
    return true;
}

// trace_radiance (texture_space_path_tracer.slang:57-84)
template <int STACK, int BLOCK, bool COUNT, bool MAT>
__device__ f3 trace_radiance(const DevScene &S, f3 origin, f3 dir, uint32_t max_bounces, Rng &rng, uint32_t *col,
                             PathCounters &pc, bool roulette = false) {
    const Light &L = S.light;
    f3 T = mk(1.0f, 1.0f, 1.0f), rad = mk(0.0f, 0.0f, 0.0f);
    for (uint32_t bounce = 0; bounce < max_bounces; ++bounce) {
        Ray r;
        r.o = origin;
        r.d = dir;
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
        if (!hit) return rad + T * miss_radiance(L, dir);
        Shading sd;
        shading_data<MAT>(S, h.id, h.b1, h.b2, -dir, sd);
        rad = rad + T * sd.emissive;
        float ux = rng.next_1d(), uy = rng.next_1d();
        f3 ld = cone_sample(L.sun_dir, L.cos_half, ux, uy);
        float ndl = fmaxf(0.0f, dot(sd.frame.n, ld));
        if (L.sun_on && ndl > 0.0f && shadow_visible<STACK, BLOCK, COUNT, MAT>(S, sd.origin_front, ld, col, pc)) {
            f3 f = sd.diffuse * kInvPi;
            f3 li = (f * L.sun_illum) * ndl;
            rad = rad + T * li;
        }
        float ex = rng.next_1d(), ey = rng.next_1d();
        f3 sdir;
        float pdf;
        cosine_sample(ex, ey, sdir, pdf);
        if (pdf <= 1e-6f) return rad;
        f3 f = sd.diffuse * kInvPi;
        float az = fabsf(sdir.z);
        T = T * mk(f.x * az / pdf, f.y * az / pdf, f.z * az / pdf);
        if (roulette && bounce + 1 < max_bounces && !roulette_survives(T, rng)) return rad;
        origin = sd.origin_front;
        dir = to_global(sd.frame, sdir);
    }
    return rad;
}

// estimate_surface_irradiance (texture_space_path_tracer.slang:86-117)
template <int STACK, int BLOCK, bool COUNT, bool MAT>
__device__ f3 estimate_irradiance(const DevScene &S, uint32_t tri, float b1, float b2, f3 view,
                                  uint32_t spp, uint32_t max_bounces, Rng &rng, uint32_t *col, PathCounters &pc,
                                  bool roulette = false) {
    const Light &L = S.light;
    Shading sd;
    shading_data<MAT>(S, tri, b1, b2, view, sd);
    f3 acc = mk(0.0f, 0.0f, 0.0f);
    for (uint32_t k = 0; k < spp; ++k) {
        f3 sample = mk(0.0f, 0.0f, 0.0f);
        float ux = rng.next_1d(), uy = rng.next_1d();
        f3 ld = cone_sample(L.sun_dir, L.cos_half, ux, uy);
        float ndl = fmaxf(0.0f, dot(sd.frame.n, ld));
        if (L.sun_on && ndl > 0.0f && shadow_visible<STACK, BLOCK, COUNT, MAT>(S, sd.origin_front, ld, col, pc))
            sample = L.sun_illum * ndl;
        float ex = rng.next_1d(), ey = rng.next_1d();
        f3 sdir;
        float pdf;
        cosine_sample(ex, ey, sdir, pdf);
        if (pdf > 1e-6f) {
            f3 rad = trace_radiance<STACK, BLOCK, COUNT, MAT>(S, sd.origin_front, to_global(sd.frame, sdir), max_bounces,
                                                         rng, col, pc, roulette);
            float az = fabsf(sdir.z);
            sample = sample + mk(rad.x * az / pdf, rad.y * az / pdf, rad.z * az / pdf);
        }
        acc = acc + sample;
    }
    float fs = (float)spp;
    return mk(acc.x / fs, acc.y / fs, acc.z / fs);
}

// store_accumulated_irradiance (texture_space_path_tracer.slang:119-138)
__device__ __forceinline__ void store_accumulated(float4 *payload, uint32_t slot, f3 e, uint32_t reset) {
    float4 p = payload[slot];
    uint32_t prev_count = reset ? 0u : __float_as_uint(p.w);
    f3 prev = prev_count == 0 ? mk(0.0f, 0.0f, 0.0f) : mk(p.x, p.y, p.z);
    uint32_t next = prev_count + 1 < kMaxTexelSamples ? prev_count + 1 : kMaxTexelSamples;
    float w = 1.0f / (float)(prev_count + 1);
    float4 o;
    o.x = prev.x + w * (e.x - prev.x);
    o.y = prev.y + w * (e.y - prev.y);
    o.z = prev.z + w * (e.z - prev.z);
    o.w = __uint_as_float(next);
    payload[slot] = o;
}

// Progressive mean with frame k (1-based since the reset) weighted by k:
// m_k = m_{k-1} + 2 / (k + 1) * (e_k - m_{k-1}) (learned forms: the cache
// improves over the frames, so later estimates are the better ones).
__device__ __forceinline__ void store_ramped(float4 *payload, uint32_t slot, f3 e, uint32_t reset) {
    float4 p = payload[slot];
    uint32_t prev_count = reset ? 0u : __float_as_uint(p.w);
    f3 prev = prev_count == 0 ? mk(0.0f, 0.0f, 0.0f) : mk(p.x, p.y, p.z);
    uint32_t next = prev_count + 1 < kMaxTexelSamples ? prev_count + 1 : kMaxTexelSamples;
    float w = 2.0f / (float)(prev_count + 2);
    payload[slot] = make_float4(prev.x + w * (e.x - prev.x), prev.y + w * (e.y - prev.y), prev.z + w * (e.z - prev.z),
                                __uint_as_float(next));
}

// 32-bit integer hash ("lowbias32", C. Wellons, hash-prospector, 2018).
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ void put3(float *a, uint64_t i, f3 v) {
    a[3 * i + 0] = v.x;
    a[3 * i + 1] = v.y;
    a[3 * i + 2] = v.z;
}

// Training record j of the learned forms: a uniformly drawn front texel and,
// on a double-sided instance, a side; its RNG stream is its own.
struct TrainStart {
    uint64_t w;
    bool back;
    uint32_t rng;
};
__device__ __forceinline__ TrainStart train_start(const DevLayout &Ly, uint32_t j, uint32_t iteration) {
    TrainStart ts;
    uint32_t h = mix32(j ^ mix32(iteration * 0x9E3779B9u + 0x632BE5ABu));
    ts.w = ((uint64_t)h * Ly.total_front) >> 32;
    ts.back = (mix32(h ^ 0x5bd1e995u) & 1u) != 0;  // applies on double-sided instances only
    ts.rng = mix32(h + 0x68E31DA4u);
    return ts;
}

struct TexelSurface {
    uint32_t instance, prim, tri, global_texel, local, scene;
    float b1, b2;
    uint2 side;
};

// compute_main prologue (texture_space_path_tracer.slang:142-199) for flat
// front texel w; leaves rng after the jitter draws.
__device__ __forceinline__ TexelSurface texel_surface(const DevLayout &Ly, uint64_t w, uint32_t iteration,
                                                      uint32_t spp, Rng &rng) {
    TexelSurface ts;
    uint32_t f = Ly.texel_face[w];
    uint32_t inst = Ly.face_inst[f];
    uint4 ii = Ly.instances[inst];
    uint2 face = Ly.faces[f];
    ts.side = Ly.sides[inst];
    ts.scene = Ly.face_scene ? Ly.face_scene[f] : 0u;
    // the RNG and the reported instance are scene-local (batched scenes keep
    // their single-scene streams)
    uint32_t inst_local = Ly.scene_inst_base ? inst - Ly.scene_inst_base[ts.scene] : inst;
    ts.instance = inst_local;
    ts.prim = f - ii.x;
    ts.tri = f;  // layout faces follow the scene's global triangle order
    ts.local = (uint32_t)(w - Ly.front_prefix[inst]);
    ts.global_texel = ii.z + ts.local;
    uint32_t i, j;
    address_to_ij(ts.local - face.x, face.y, i, j);
    float fr = (float)face.y;
    float base_x = (float)i / fr, base_y = (float)j / fr;
    if (Ly.rng_pcg) rng.seed_pcg(ts.global_texel, inst_local, iteration * (spp > 1u ? spp : 1u));
    else rng.seed(ts.global_texel, inst_local, iteration * (spp > 1u ? spp : 1u));
    float bx = base_x, by = base_y;
    for (int attempt = 0; attempt < 32; ++attempt) {
        float ux = rng.next_1d();
        float uy = rng.next_1d();
        float cx = base_x + (ux - 0.5f) / fr;
        float cy = base_y + (uy - 0.5f) / fr;
        if (cx < 0.0f || cy < 0.0f || cx + cy > 1.0f) continue;
        f3 cb = mk(cx, cy, 1.0f - cx - cy);
        if (nearest_texel(cb, ii.z, face) == ts.global_texel) {
            bx = cx;
            by = cy;
            break;
        }
    }
    ts.b1 = by;
    ts.b2 = fmaxf(0.0f, 1.0f - bx - by);
    return ts;
}

__device__ __forceinline__ void flush_counters(const PathCounters &pc, unsigned long long *counters) {
    __shared__ unsigned int block_sum[7];
    if (threadIdx.x < 7) block_sum[threadIdx.x] = 0;
    __syncthreads();
    atomicAdd(&block_sum[0], pc.extend);
    atomicAdd(&block_sum[1], pc.shadow);
    atomicAdd(&block_sum[2], pc.nodes);
    atomicAdd(&block_sum[3], pc.tris);
    atomicAdd(&block_sum[4], pc.slots);
    atomicAdd(&block_sum[5], pc.shadow_nodes);
    atomicAdd(&block_sum[6], pc.shadow_tris);
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&counters[NAN_COUNTER_EXTEND_RAYS], (unsigned long long)block_sum[0]);
        atomicAdd(&counters[NAN_COUNTER_SHADOW_RAYS], (unsigned long long)block_sum[1]);
        atomicAdd(&counters[NAN_COUNTER_NODE_VISITS], (unsigned long long)block_sum[2]);
        atomicAdd(&counters[NAN_COUNTER_TRI_TESTS], (unsigned long long)block_sum[3]);
        atomicAdd(&counters[NAN_COUNTER_SLOT_UPDATES], (unsigned long long)block_sum[4]);
        atomicAdd(&counters[NAN_COUNTER_SHADOW_NODE_VISITS], (unsigned long long)block_sum[5]);
        atomicAdd(&counters[NAN_COUNTER_SHADOW_TRI_TESTS], (unsigned long long)block_sum[6]);
    }
}

