// Wavefront texture-space path tracer for gfx950.
//
// The per-texel work of texture_space_path_tracer.slang:140-216 split into
// stages so that the traversal kernels run alone, with few registers and
// many waves per SIMD:
//
//   generate  ray-gen over the Mesh Colors lattice, surface shading, the
//             surface sun shadow ray and the first bounce ray
//   extend    closest hit for the live bounce rays          (traversal)
//   shade     emission, sun next-event ray, Lambert bounce, compaction of
//             the surviving paths into the next ray queue by wave ballot
//   shadow    any-hit for the queued sun rays             (traversal)
//   finish    running-mean store into the front/back 16-byte slots
//
// repeated `max_bounces` times per side.  Back sides of double-sided
// instances run after every front side has finished, continuing each
// texel's RNG where its front path left it, so the draw order and every
// floating-point operation equal the single-kernel form (and the CPU
// oracle): the payload is bit-identical.
//
// Traversal kernels are persistent: a fixed grid sized from the occupancy
// calculator; a lane whose ray is done takes the next queued ray (refills of a
// wave share one atomic).
//
// Queues are split into kParts partitions with one counter each, on its own
// 64-byte line.  A device-scope atomic on one address is serialized at the
// memory side (about 11 ns on MI355X whatever the caller); on 8 addresses they
// proceed in parallel.  Producers append per block (one atomic per block and
// queue) into partition blockIdx % kParts; consumers walk the partitions in
// order, traversal waves starting at their own and moving on when it is empty.
// Which lane traces which ray changes nothing in the results.
// Their stack is a 16-entry ring per lane in LDS that spills its oldest
// entries to a per-lane column in HBM when a traversal goes deeper.
#include "nan_tracer.h"

using namespace nanx;

namespace {

constexpr int kWfBlock = 256;       // stage kernels
constexpr int kTravBlock = 256;     // traversal kernels
constexpr int kLdsStackMax = 32;    // largest LDS ring (entries per lane)
constexpr int kMaxStack = 128;      // deepest traversal the spill area holds
constexpr uint32_t kTargetSun = 0u;
constexpr uint32_t kTargetRad = 1u;
constexpr uint32_t kCounterSlots = 1024;
constexpr uint32_t kParts = 8;       // queue partitions, one counter each
constexpr uint32_t kCtrStride = 16;  // uint32 words between counters (one 64-byte line each)
constexpr uint32_t kCounterWords = kCounterSlots * kParts * kCtrStride;
// The double-sided texels' sweeps run on a second stream with grids of
// 1/kSideDiv of the main ones, so their short, latency-bound launches hold
// few CUs while the single-sided sweep fills the rest (0: one stream).
#ifndef NAN_WF_SIDE_DIV
#define NAN_WF_SIDE_DIV 4
#endif
constexpr uint32_t kSideDiv = NAN_WF_SIDE_DIV;
constexpr uint64_t kSideLimit = 8;  // split when at most 1/kSideLimit of front texels are double-sided

}  // namespace

struct WfBuffers {
    uint32_t capacity = 0;      // paths (front texels)
    uint32_t *rng = nullptr;    // [cap] random stream state of each path
    uint32_t rng_pcg = 0;       // 1: PCG streams (NAN_RNG_PCG); set per call
    float4 *T = nullptr;        // [cap] throughput.xyz, |cos| of the first bounce
    float4 *rad = nullptr;      // [cap] path radiance.xyz, pdf of the first bounce
    float4 *sun = nullptr;      // [cap] surface sun irradiance of the current sample
    float4 *acc = nullptr;      // [cap] sum of samples over spp (spp > 1 only)
    // queues: kParts partitions of qcap (ray) or 2 qcap (shadow) entries
    float4 *ray_o[2] = {nullptr, nullptr};  // origin.xyz, path id bits
    float4 *ray_d[2] = {nullptr, nullptr};  // direction.xyz, scene bits
    float4 *hits = nullptr;     // t, b1, b2, triangle id bits (indexed like the ray queue)
    float4 *sh_o = nullptr;     // origin.xyz, path id bits
    float4 *sh_d = nullptr;     // direction.xyz, target bits
    float4 *sh_c = nullptr;     // contribution.xyz, scene bits
    uint32_t qcap = 0;          // entries per partition of a ray queue
    uint32_t *qidx = nullptr;   // [cap] learned forms: cache query index of each path
    uint32_t *counters = nullptr;   // [kCounterWords]: slot s, partition j at (s * kParts + j) * kCtrStride
    uint32_t *frame = nullptr;      // [iteration, reset] in device-frame mode
    uint32_t *spill = nullptr;      // [(kMaxStack - ring) * spill_lanes]
    uint32_t spill_lanes = 0;
    uint32_t trav_blocks = 0;       // persistent grid of the traversal kernels
    uint32_t ring = 16;             // LDS ring entries per lane
    uint32_t blocks_per_cu_req = 0; // requested persistent blocks per CU (0: occupancy limit)
    uint32_t gen_blocks = 0;        // grids of the stage kernels: resident capacity (CUs x occupancy),
    uint32_t shade_blocks = 0;      // so a grid-stride loop never waits for a second round of blocks
    uint32_t finish_blocks = 0;
    // Host only.  The double-sided texels' paths (front sweep, then back
    // sweep) run on a second stream beside the single-sided texels' front
    // sweep: `ds` is their queue set (own queues and spill area, this set's
    // per-path arrays and counters).  Null when the tracer has no such split.
    WfBuffers *ds = nullptr;
    hipStream_t side_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t *queue_counter(uint32_t *counters, uint32_t slot, uint32_t part) {
    return counters + ((size_t)slot * kParts + part) * kCtrStride;
}

// Entry counts of a queue's partitions, each clamped to the partition size.
__device__ __forceinline__ void queue_counts(const uint32_t *counters, uint32_t slot, uint32_t part_cap,
                                             uint32_t (&c)[kParts]) {
#pragma unroll
    for (uint32_t j = 0; j < kParts; ++j) c[j] = min(counters[((size_t)slot * kParts + j) * kCtrStride], part_cap);
}

// Storage index of the k-th entry of a queue (partitions in order).
__device__ __forceinline__ uint32_t queue_entry(const uint32_t (&c)[kParts], uint32_t k, uint32_t part_cap) {
    uint32_t j = 0, off = k;
#pragma unroll
    for (uint32_t t = 0; t + 1 < kParts; ++t) {
        bool past = j == t && off >= c[t];
        off -= past ? c[t] : 0u;
        j += past ? 1u : 0u;
    }
    return j * part_cap + off;
}

// Appends the wanting threads of the block to two queues (partition
// blockIdx % kParts of each) with one atomic per queue; every thread of the
// block must call it.  Returns storage indexes (valid where wanted); an
// entry that would not fit its partition gets 0xFFFFFFFF and is dropped (the
// partition sizes make that impossible for grids that are multiples of
// kParts).
__device__ __forceinline__ void block_append2(bool want_a, bool want_b, uint32_t *ctr_a, uint32_t *ctr_b,
                                              uint32_t cap_a, uint32_t cap_b, uint32_t &pos_a, uint32_t &pos_b) {
    constexpr uint32_t kWaves = kWfBlock / 64;
    __shared__ uint32_t s_a[kWaves + 1], s_b[kWaves + 1];
    unsigned long long ma = __ballot(want_a), mb = __ballot(want_b);
    uint32_t wave = threadIdx.x >> 6;
    if (lane_id() == 0) {
        s_a[wave] = (uint32_t)__popcll(ma);
        s_b[wave] = (uint32_t)__popcll(mb);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ta = 0, tb = 0;
        for (uint32_t w = 0; w < kWaves; ++w) {
            uint32_t ca = s_a[w], cb = s_b[w];
            s_a[w] = ta;
            s_b[w] = tb;
            ta += ca;
            tb += cb;
        }
        s_a[kWaves] = ta ? atomicAdd(ctr_a, ta) : 0u;
        s_b[kWaves] = tb ? atomicAdd(ctr_b, tb) : 0u;
    }
    __syncthreads();
    unsigned long long below = (1ull << lane_id()) - 1ull;
    uint32_t part = blockIdx.x % kParts;
    uint32_t ia = s_a[kWaves] + s_a[wave] + (uint32_t)__popcll(ma & below);
    uint32_t ib = s_b[kWaves] + s_b[wave] + (uint32_t)__popcll(mb & below);
    pos_a = ia < cap_a ? part * cap_a + ia : 0xFFFFFFFFu;
    pos_b = ib < cap_b ? part * cap_b + ib : 0xFFFFFFFFu;
    __syncthreads();  // s_a / s_b are reused by the next call
}

// Resumable traversal with a RING-entry LDS ring per lane that spills its
// oldest entries to HBM, in two phases a wave runs in step (after Aila and
// Laine, "Understanding the Efficiency of Ray Traversal on GPUs", 2009):
// in the node phase each lane visits one inner node, or parks the leaf it
// reached and moves on to its next stacked entry; in the leaf phase every lane
// with a parked leaf tests that leaf's triangles.  A wave thus never runs the
// node and the triangle code in the same step for different lanes.  The
// result equals traverse() in nan_device.h (closest hit = smallest
// (t, triangle id); any-hit stops at the first hit in [tmin, tmax]), in
// whatever order nodes and leaves are visited.
constexpr uint32_t kNone = 0xFFFFFFFFu;  // not a node or leaf entry (leaves hold at most 8 triangles)

struct TravState {
    uint32_t cur;   // next entry to visit, kNone once the stack is exhausted (then sp == 0)
    uint32_t sp;    // unsigned, so the ring slot sp % RING of a power-of-two ring is one AND
    uint32_t pend;  // leaf parked for the next leaf phase, or kNone
};

template <int RING>
struct LaneStack {
    uint32_t *lds_col, *spill_col;
    uint32_t spill_stride;
    __device__ __forceinline__ void push(TravState &ts, uint32_t e) const {
        uint32_t *slot = lds_col + (ts.sp % RING) * kTravBlock;
        if (ts.sp >= RING) spill_col[(size_t)(ts.sp - RING) * spill_stride] = *slot;
        *slot = e;
        ++ts.sp;
    }
    __device__ __forceinline__ uint32_t pop(TravState &ts) const {
        --ts.sp;
        uint32_t *slot = lds_col + (ts.sp % RING) * kTravBlock;
        uint32_t e = *slot;
        if (ts.sp >= RING) *slot = spill_col[(size_t)(ts.sp - RING) * spill_stride];
        return e;
    }
};

// Node phase step; returns true when the ray has nothing left to visit.
template <int RING, bool ANY_HIT, bool COUNT>
__device__ __forceinline__ bool node_step(const DevScene &S, const Ray &r, float slack, const Hit &hit,
                                          TravState &ts, const LaneStack<RING> &stk, TraversalStats &st) {
    auto push = [&](uint32_t e) { stk.push(ts, e); };
    auto pop = [&]() -> uint32_t { return stk.pop(ts); };
    if (ts.cur == kNone) return ts.pend == kNone;
    if (ts.cur & kLeafBit) {
        if (ts.pend != kNone) return false;  // waits for the leaf phase
        ts.pend = ts.cur;
        ts.cur = ts.sp ? pop() : kNone;
        return false;
    }
    {
        const uint4 *np = S.nodes + 4 * (size_t)ts.cur;
        uint4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3];
        if (COUNT) st.nodes++;
        uint32_t e0, e1, e2, e3;
        int nhit = node_children<ANY_HIT>(w0, w1, w2, w3, r, __builtin_fmaf(ANY_HIT ? r.tmax : hit.t, kPad, slack), slack, e0,
                                          e1, e2, e3);
        if (nhit == 0) {
            ts.cur = ts.sp ? pop() : kNone;
            return ts.cur == kNone && ts.pend == kNone;
        }
        if (nhit > 3) push(e3);
        if (nhit > 2) push(e2);
        if (nhit > 1) push(e1);
        ts.cur = e0;
        return false;
    }
}

// Leaf phase step: tests the parked leaf; returns true when the ray is done.
template <bool ANY_HIT, bool COUNT, bool MAT>
__device__ __forceinline__ bool leaf_step(const DevScene &S, const Ray &r, Hit &hit, TravState &ts,
                                          TraversalStats &st) {
    if (ts.pend != kNone) {
        uint32_t first = ts.pend & kLeafFirstMask;
        uint32_t count = ((ts.pend >> kLeafCountShift) & 0xFu) + 1u;
        ts.pend = kNone;
        for (uint32_t i = 0; i < count; ++i) {
            const float4 *tp = S.tris + 3 * (size_t)(first + i);
            float4 a = tp[0], b = tp[1], c = tp[2];
            if (COUNT) st.tris++;
            float t, b1, b2;
            if (tri_intersect(r, mk(a.x, a.y, a.z), mk(a.w, b.x, b.y), mk(b.z, b.w, c.x), t, b1, b2)) {
                uint32_t id = __float_as_uint(c.y);
                bool tested = MAT && __float_as_uint(c.z) != 0u;  // non-opaque material
                if (ANY_HIT) {
                    if (tested && !alpha_accept(S, id, b1, b2)) continue;
                    hit.t = t;
                    hit.id = id;
                    return true;
                }
                if ((t < hit.t || (t == hit.t && id < hit.id)) && (!tested || alpha_accept(S, id, b1, b2))) {
                    hit.t = t;
                    hit.b1 = b1;
                    hit.b2 = b2;
                    hit.id = id;
                }
            }
        }
    }
    return ts.cur == kNone;
}

// Persistent traversal over a ray queue with per-lane refill: a lane whose
// ray has finished takes the next queued ray (one atomic per wave for all
// refilling lanes), so a wave keeps tracing while any rays remain instead
// of waiting for its slowest lane.
template <int RING, bool ANY_HIT, bool COUNT, bool MAT, typename Load, typename Store>
__device__ __forceinline__ void trace_queue(const DevScene &S, uint32_t *counters, uint32_t count_slot,
                                            uint32_t fetch_slot, uint32_t part_cap, uint32_t *lds_col,
                                            uint32_t *spill_col, uint32_t spill_stride, uint32_t steps, Load load,
                                            Store store, uint32_t &n_rays, uint32_t &n_nodes, uint32_t &n_tris) {
    // partition sizes in LDS (registers are the traversal's)
    __shared__ uint32_t s_cnt[kParts];
    if (threadIdx.x < kParts)
        s_cnt[threadIdx.x] = min(counters[((size_t)count_slot * kParts + threadIdx.x) * kCtrStride], part_cap);
    __syncthreads();
    const LaneStack<RING> stk{lds_col, spill_col, spill_stride};
    Ray r;
    Hit h;
    float slack = 0.0f;
    TravState ts{0, 0, kNone};
    TraversalStats st{0, 0};
    uint32_t k = 0;
    bool has = false;
    bool drained = false;
    uint32_t part = blockIdx.x % kParts, tried = 0;  // wave-uniform
    while (true) {
        // refill idle lanes from the current partition, moving on while it is empty
        while (!drained) {
            unsigned long long mask = __ballot(!has);
            uint32_t n = (uint32_t)__popcll(mask);
            if (!n) break;
            uint32_t c = s_cnt[part];
            uint32_t base = 0;
            int leader = __ffsll((long long)mask) - 1;
            if ((int)lane_id() == leader) base = atomicAdd(queue_counter(counters, fetch_slot, part), n);
            base = __shfl(base, leader);
            uint32_t avail = base < c ? min(n, c - base) : 0u;
            if (!has) {
                uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane_id()) - 1ull));
                if (rank < avail) {
                    k = part * part_cap + base + rank;
                    uint32_t root;
                    load(k, r, root, slack);
                    ray_setup(r);
                    h = Hit{r.tmax, 0.0f, 0.0f, kInvalid};
                    ts = TravState{root, 0, kNone};
                    st = TraversalStats{0, 0};
                    has = true;
                }
            }
            if (base + n >= c) {
                part = part + 1 == kParts ? 0u : part + 1;
                if (++tried == kParts) drained = true;
            }
        }
        if (__ballot(has) == 0ull) break;
        for (uint32_t s = 0; s < steps; ++s) {
            // leaf phase once the lanes waiting on a parked leaf are at least
            // half as many as those that can still visit nodes (measured
            // against 1/8, 1/4, 1, 2 and "all waiting": this is the fastest)
            bool blocked = has && ts.pend != kNone && (ts.cur == kNone || (ts.cur & kLeafBit));
            bool movable = has && ts.cur != kNone && (!(ts.cur & kLeafBit) || ts.pend == kNone);
            unsigned long long mb = __ballot(blocked), mm = __ballot(movable);
            bool leaf_phase = mb && (!mm || 2 * __popcll(mb) >= __popcll(mm));
            if (has) {
                bool done = leaf_phase ? leaf_step<ANY_HIT, COUNT, MAT>(S, r, h, ts, st)
                                       : node_step<RING, ANY_HIT, COUNT>(S, r, slack, h, ts, stk, st);
                if (done) {
                    store(k, h);
                    has = false;
                    if (COUNT) {
                        n_rays++;
                        n_nodes += st.nodes;
                        n_tris += st.tris;
                    }
                }
            }
            if (__ballot(has) == 0ull) break;
        }
    }
}

__device__ __forceinline__ void add_counter(unsigned long long *counters, int slot, uint32_t v) {
    // wave-reduce, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (lane_id() == 0 && v) atomicAdd(&counters[slot], (unsigned long long)v);
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void wf_zero_counters(uint32_t *c, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) c[i] = 0u;
}

// End of a frame in device-frame mode: next iteration, no reset.
__global__ void wf_advance_frame(uint32_t *frame) {
    if (threadIdx.x == 0) {
        frame[0] += 1u;
        frame[1] = 0u;
    }
}

struct GenParams {
    const uint32_t *frame; // device [iteration, reset] (null: use iteration below)
    const uint32_t *list;  // null: path i is front texel i
    uint32_t count;
    uint32_t side;         // 0 front, 1 back
    uint32_t sample;       // 0 .. spp-1
    uint32_t iteration, spp;
    uint32_t ext_slot, sh_slot;  // counter slots of the queues written
};

template <bool MAT>
__global__ __launch_bounds__(kWfBlock, 6) void wf_generate(DevScene S, DevLayout Ly, WfBuffers W, GenParams P) {
    uint32_t stride = gridDim.x * kWfBlock;
    uint32_t *ext_count = queue_counter(W.counters, P.ext_slot, blockIdx.x % kParts);
    uint32_t *sh_count = queue_counter(W.counters, P.sh_slot, blockIdx.x % kParts);
    for (uint32_t base = blockIdx.x * kWfBlock; base < P.count; base += stride) {
        uint32_t i = base + threadIdx.x;
        bool active = i < P.count;
        bool want_shadow = false, want_ray = false;
        f3 origin = mk(0.0f, 0.0f, 0.0f), ld = origin, dir = origin, contrib = origin;
        uint32_t w = 0, scene = 0;
        if (active) {
            w = P.list ? P.list[i] : i;
            Rng rng;
            TexelSurface ts = texel_surface(Ly, w, P.frame ? P.frame[0] : P.iteration, P.spp, rng);
            scene = ts.scene;
            const Light &L = S.scenes[scene].light;
            if (P.side != 0 || P.sample != 0) rng.state = W.rng[w];
            f3 n = world_flat_normal(S, ts.tri);
            Shading sd;
            shading_data<MAT>(S, ts.tri, ts.b1, ts.b2, P.side ? -n : n, sd);
            float ux = rng.next_1d(), uy = rng.next_1d();
            ld = cone_sample(L.sun_dir, L.cos_half, ux, uy);
            float ndl = fmaxf(0.0f, dot(sd.frame.n, ld));
            want_shadow = L.sun_on && ndl > 0.0f;
            contrib = L.sun_illum * ndl;
            float ex = rng.next_1d(), ey = rng.next_1d();
            f3 sdir;
            float pdf;
            cosine_sample(ex, ey, sdir, pdf);
            float az = fabsf(sdir.z);
            want_ray = pdf > 1e-6f;
            dir = to_global(sd.frame, sdir);
            origin = sd.origin_front;
            W.T[w] = make_float4(1.0f, 1.0f, 1.0f, az);
            W.rad[w] = make_float4(0.0f, 0.0f, 0.0f, pdf);
            W.sun[w] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            W.rng[w] = rng.state;
        }
        uint32_t qs, qr;
        block_append2(active && want_shadow, active && want_ray, sh_count, ext_count, 2 * W.qcap, W.qcap, qs, qr);
        if (active && want_shadow && qs != 0xFFFFFFFFu) {
            W.sh_o[qs] = make_float4(origin.x, origin.y, origin.z, __uint_as_float(w));
            W.sh_d[qs] = make_float4(ld.x, ld.y, ld.z, __uint_as_float(kTargetSun));
            W.sh_c[qs] = make_float4(contrib.x, contrib.y, contrib.z, __uint_as_float(scene));
        }
        if (active && want_ray && qr != 0xFFFFFFFFu) {
            W.ray_o[0][qr] = make_float4(origin.x, origin.y, origin.z, __uint_as_float(w));
            W.ray_d[0][qr] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(scene));
        }
    }
}

template <int RING, bool COUNT, bool MAT>
__global__ __launch_bounds__(kTravBlock, 8) void wf_extend(DevScene S, WfBuffers W, uint32_t q, uint32_t count_slot,
                                                          uint32_t fetch_slot, uint32_t steps,
                                                          unsigned long long *counters) {
    __shared__ uint32_t stack[RING * kTravBlock];
    uint32_t gid = blockIdx.x * kTravBlock + threadIdx.x;
    uint32_t n_rays = 0, n_nodes = 0, n_tris = 0;
    const float4 *ro = W.ray_o[q], *rd = W.ray_d[q];
    float4 *hits = W.hits;
    trace_queue<RING, false, COUNT, MAT>(
        S, W.counters, count_slot, fetch_slot, W.qcap, stack + threadIdx.x, W.spill + gid, W.spill_lanes, steps,
        [&](uint32_t k, Ray &r, uint32_t &root, float &slack) {
            float4 o = ro[k], d = rd[k];
            r.o = mk(o.x, o.y, o.z);
            r.d = mk(d.x, d.y, d.z);
            r.tmin = 0.0f;
            r.tmax = 1000.0f;
            const SceneRec &sc = S.scenes[__float_as_uint(d.w)];
            root = sc.root;
            slack = sc.slack;
        },
        [&](uint32_t k, const Hit &h) { hits[k] = make_float4(h.t, h.b1, h.b2, __uint_as_float(h.id)); },
        n_rays, n_nodes, n_tris);
    if (COUNT) {
        add_counter(counters, NAN_COUNTER_EXTEND_RAYS, n_rays);
        add_counter(counters, NAN_COUNTER_NODE_VISITS, n_nodes);
        add_counter(counters, NAN_COUNTER_TRI_TESTS, n_tris);
    }
}

template <int RING, bool COUNT, bool MAT>
__global__ __launch_bounds__(kTravBlock, 8) void wf_shadow(DevScene S, WfBuffers W, uint32_t count_slot,
                                                          uint32_t fetch_slot, uint32_t steps,
                                                          unsigned long long *counters) {
    __shared__ uint32_t stack[RING * kTravBlock];
    uint32_t gid = blockIdx.x * kTravBlock + threadIdx.x;
    uint32_t n_rays = 0, n_nodes = 0, n_tris = 0;
    trace_queue<RING, true, COUNT, MAT>(
        S, W.counters, count_slot, fetch_slot, 2 * W.qcap, stack + threadIdx.x, W.spill + gid, W.spill_lanes, steps,
        [&](uint32_t k, Ray &r, uint32_t &root, float &slack) {
            float4 o = W.sh_o[k], d = W.sh_d[k];
            r.o = mk(o.x, o.y, o.z);
            r.d = mk(d.x, d.y, d.z);
            r.tmin = 0.0f;
            r.tmax = 1000.0f;
            const SceneRec &sc = S.scenes[__float_as_uint(W.sh_c[k].w)];
            root = sc.root;
            slack = sc.slack;
        },
        [&](uint32_t k, const Hit &h) {
            if (h.id != kInvalid) return;  // occluded
            float4 o = W.sh_o[k], d = W.sh_d[k], c = W.sh_c[k];
            uint32_t p = __float_as_uint(o.w);
            if (__float_as_uint(d.w) == kTargetSun) {
                W.sun[p] = make_float4(c.x, c.y, c.z, 0.0f);
            } else {
                float4 v = W.rad[p];
                v.x = v.x + c.x;
                v.y = v.y + c.y;
                v.z = v.z + c.z;
                W.rad[p] = v;
            }
        },
        n_rays, n_nodes, n_tris);
    if (COUNT) {
        add_counter(counters, NAN_COUNTER_SHADOW_RAYS, n_rays);
        add_counter(counters, NAN_COUNTER_SHADOW_NODE_VISITS, n_nodes);
        add_counter(counters, NAN_COUNTER_SHADOW_TRI_TESTS, n_tris);
    }
}

struct ShadeParams {
    uint32_t q_in, q_out;
    uint32_t in_slot, out_slot, sh_slot;
    uint32_t bounce, max_bounces;
    // learned-form training paths (0: none): at the first hit y1, record 1 its
    // position, normal and albedo as the record's features, record 2 its
    // emission aside in W.sun; either way y1's emission stays out of rad
    uint32_t train;
    LearnedIO io;
};

// one iteration of trace_radiance's loop body (texture_space_path_tracer.slang:60-83)
template <bool MAT>
__global__ __launch_bounds__(kWfBlock, 5) void wf_shade(DevScene S, WfBuffers W, ShadeParams P) {
    uint32_t cnt[kParts];
    queue_counts(W.counters, P.in_slot, W.qcap, cnt);
    uint32_t count = 0;
#pragma unroll
    for (uint32_t j = 0; j < kParts; ++j) count += cnt[j];
    uint32_t *sh_count = queue_counter(W.counters, P.sh_slot, blockIdx.x % kParts);
    uint32_t *out_count = queue_counter(W.counters, P.out_slot, blockIdx.x % kParts);
    uint32_t stride = gridDim.x * kWfBlock;
    for (uint32_t base = blockIdx.x * kWfBlock; base < count; base += stride) {
        uint32_t kq = base + threadIdx.x;
        bool active = kq < count;
        uint32_t k = active ? queue_entry(cnt, kq, W.qcap) : 0u;
        bool want_shadow = false, want_ray = false;
        f3 origin = mk(0.0f, 0.0f, 0.0f), ld = origin, dir = origin, contrib = origin;
        uint32_t p = 0, scene = 0;
        if (active) {
            float4 ro = W.ray_o[P.q_in][k], rd = W.ray_d[P.q_in][k];
            float4 hv = W.hits[k];
            p = __float_as_uint(ro.w);
            scene = __float_as_uint(rd.w);
            const SceneRec &sc = S.scenes[scene];
            const Light &L = sc.light;
            uint32_t max_bounces = sc.max_bounces ? sc.max_bounces : P.max_bounces;
            float4 Tv = W.T[p], Rv = W.rad[p];
            f3 T = mk(Tv.x, Tv.y, Tv.z), rad = mk(Rv.x, Rv.y, Rv.z);
            uint32_t id = __float_as_uint(hv.w);
            bool first_train = P.train != 0 && P.bounce == 1;
            if (id == kInvalid) {
                rad = rad + T * miss_radiance(L, mk(rd.x, rd.y, rd.z));
            } else {
                Rng rng;
                rng.state = W.rng[p];
                rng.pcg = W.rng_pcg;
                f3 d = mk(rd.x, rd.y, rd.z);
                Shading sd;
                shading_data<MAT>(S, id, hv.y, hv.z, -d, sd);
                if (!first_train) {
                    rad = rad + T * sd.emissive;
                } else if (P.train == 1) {
                    put3(P.io.tr_pos, p, sd.origin_front);
                    put3(P.io.tr_nrm, p, sd.frame.n);
                    put3(P.io.tr_alb, p, sd.diffuse);
                    P.io.tr_valid[p] = 1u;
                } else {
                    W.sun[p] = make_float4(sd.emissive.x, sd.emissive.y, sd.emissive.z, 0.0f);
                }
                float ux = rng.next_1d(), uy = rng.next_1d();
                ld = cone_sample(L.sun_dir, L.cos_half, ux, uy);
                float ndl = fmaxf(0.0f, dot(sd.frame.n, ld));
                if (L.sun_on && ndl > 0.0f) {
                    f3 f = sd.diffuse * kInvPi;
                    f3 li = (f * L.sun_illum) * ndl;
                    contrib = T * li;
                    want_shadow = true;
                }
                float ex = rng.next_1d(), ey = rng.next_1d();
                f3 sdir;
                float pdf;
                cosine_sample(ex, ey, sdir, pdf);
                if (pdf > 1e-6f) {
                    f3 f = sd.diffuse * kInvPi;
                    float az = fabsf(sdir.z);
                    T = T * mk(f.x * az / pdf, f.y * az / pdf, f.z * az / pdf);
                    origin = sd.origin_front;
                    dir = to_global(sd.frame, sdir);
                    want_ray = P.bounce < max_bounces;
                    if (want_ray && sc.roulette) want_ray = roulette_survives(T, rng);
                }
                W.rng[p] = rng.state;
                W.T[p] = make_float4(T.x, T.y, T.z, Tv.w);
            }
            W.rad[p] = make_float4(rad.x, rad.y, rad.z, Rv.w);
        }
        uint32_t qs, qr;
        block_append2(active && want_shadow, active && want_ray, sh_count, out_count, 2 * W.qcap, W.qcap, qs, qr);
        if (active && want_shadow && qs != 0xFFFFFFFFu) {
            W.sh_o[qs] = make_float4(origin.x, origin.y, origin.z, __uint_as_float(p));
            W.sh_d[qs] = make_float4(ld.x, ld.y, ld.z, __uint_as_float(kTargetRad));
            W.sh_c[qs] = make_float4(contrib.x, contrib.y, contrib.z, __uint_as_float(scene));
        }
        if (active && want_ray && qr != 0xFFFFFFFFu) {
            W.ray_o[P.q_out][qr] = make_float4(origin.x, origin.y, origin.z, __uint_as_float(p));
            W.ray_d[P.q_out][qr] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(scene));
        }
    }
}

struct FinishParams {
    const uint32_t *frame; // device [iteration, reset] (null: use reset below)
    const uint32_t *list;
    uint32_t count, side, sample, spp, reset;
    float4 *const *payloads;  // per-scene payload (null: the launch's payload)
};

template <bool COUNT>
__global__ __launch_bounds__(kWfBlock) void wf_finish(DevLayout Ly, WfBuffers W, float4 *payload, FinishParams P,
                                                     unsigned long long *counters) {
    uint32_t stride = gridDim.x * kWfBlock;
    for (uint32_t base = blockIdx.x * kWfBlock; base < P.count; base += stride) {
        uint32_t i = base + threadIdx.x;
        uint32_t slots = 0;
        if (i < P.count) {
            uint32_t w = P.list ? P.list[i] : i;
            float4 sv = W.sun[w], Rv = W.rad[w], Tv = W.T[w];
            f3 sample = mk(sv.x, sv.y, sv.z);
            float az = Tv.w, pdf = Rv.w;
            if (pdf > 1e-6f) sample = sample + mk(Rv.x * az / pdf, Rv.y * az / pdf, Rv.z * az / pdf);
            f3 acc;
            if (P.spp > 1) {
                float4 a = P.sample == 0 ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : W.acc[w];
                acc = mk(a.x, a.y, a.z) + sample;
                W.acc[w] = make_float4(acc.x, acc.y, acc.z, 0.0f);
            } else {
                acc = mk(0.0f, 0.0f, 0.0f) + sample;
            }
            if (P.sample + 1 == P.spp) {
                float fs = (float)P.spp;
                f3 e = mk(acc.x / fs, acc.y / fs, acc.z / fs);
                uint32_t f = Ly.texel_face[w];
                uint32_t inst = Ly.face_inst[f];
                uint32_t local = w - Ly.front_prefix[inst];
                uint32_t slot = P.side ? Ly.sides[inst].x + local : Ly.instances[inst].z + local;
                float4 *dst = P.payloads ? P.payloads[Ly.face_scene ? Ly.face_scene[f] : 0u] : payload;
                store_accumulated(dst, slot, e, P.frame ? P.frame[1] : P.reset);
                slots = 1;
            }
        }
        if (COUNT) add_counter(counters, NAN_COUNTER_SLOT_UPDATES, slots);
    }
}

// ---------------------------------------------------------------------------
// learned forms (1 sample per texel, no emitter light sampling): the kernels of
// nan_learned.hip restaged on the queues, with the same random draws and the
// same floating-point operations in the same order
// ---------------------------------------------------------------------------
struct LearnedGenParams {
    const uint32_t *list;  // null: path i is front texel i
    uint32_t count, side, iteration;
    uint32_t ext_slot, sh_slot;
    uint32_t at_x0;        // 1: query the cache at the texel itself (no bounce ray)
};

// Texel samples of one side: the sun shadow ray, then the first bounce ray
// (its query is filled at the hit) or, at_x0, the cache query at the texel.
template <bool MAT>
__global__ __launch_bounds__(kWfBlock, 6) void wf_learned_generate(DevScene S, DevLayout Ly, WfBuffers W,
                                                                   LearnedIO io, LearnedGenParams P) {
    uint32_t stride = gridDim.x * kWfBlock;
    uint32_t *ext_count = queue_counter(W.counters, P.ext_slot, blockIdx.x % kParts);
    uint32_t *sh_count = queue_counter(W.counters, P.sh_slot, blockIdx.x % kParts);
    for (uint32_t base = blockIdx.x * kWfBlock; base < P.count; base += stride) {
        uint32_t i = base + threadIdx.x;
        bool active = i < P.count;
        bool want_shadow = false, want_ray = false;
        f3 origin = mk(0.0f, 0.0f, 0.0f), ld = origin, dir = origin, contrib = origin;
        uint32_t w = 0, scene = 0;
        if (active) {
            w = P.list ? P.list[i] : i;
            Rng rng;
            TexelSurface ts = texel_surface(Ly, w, io.frame ? io.frame[0] : P.iteration, 1u, rng);
            scene = ts.scene;
            const Light &L = S.scenes[scene].light;
            if (P.side != 0) rng.state = W.rng[w];
            f3 n = world_flat_normal(S, ts.tri);
            Shading sd;
            shading_data<MAT>(S, ts.tri, ts.b1, ts.b2, P.side ? -n : n, sd);
            float ux = rng.next_1d(), uy = rng.next_1d();
            ld = cone_sample(L.sun_dir, L.cos_half, ux, uy);
            float ndl = fmaxf(0.0f, dot(sd.frame.n, ld));
            want_shadow = L.sun_on && ndl > 0.0f;
            contrib = L.sun_illum * ndl;
            origin = sd.origin_front;
            W.sun[w] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (P.at_x0) {
                put3(io.q_pos, i, sd.origin_front);
                put3(io.q_nrm, i, sd.frame.n);
                put3(io.q_alb, i, sd.diffuse);
                io.q_w[i] = 1.0f;
            } else {
                float ex = rng.next_1d(), ey = rng.next_1d();
                f3 sdir;
                float pdf;
                cosine_sample(ex, ey, sdir, pdf);
                want_ray = pdf > 1e-6f;
                dir = to_global(sd.frame, sdir);
                W.T[w] = make_float4(1.0f, 1.0f, 1.0f, fabsf(sdir.z));
                W.rad[w] = make_float4(0.0f, 0.0f, 0.0f, pdf);
                W.qidx[w] = i;
                put3(io.q_pos, i, mk(0.0f, 0.0f, 0.0f));
                put3(io.q_nrm, i, mk(0.0f, 0.0f, 0.0f));
                put3(io.q_alb, i, mk(0.0f, 0.0f, 0.0f));
                io.q_w[i] = 0.0f;
            }
            W.rng[w] = rng.state;
        }
        uint32_t qs, qr;
        block_append2(active && want_shadow, active && want_ray, sh_count, ext_count, 2 * W.qcap, W.qcap, qs, qr);
        if (active && want_shadow && qs != 0xFFFFFFFFu) {
            W.sh_o[qs] = make_float4(origin.x, origin.y, origin.z, __uint_as_float(w));
            W.sh_d[qs] = make_float4(ld.x, ld.y, ld.z, __uint_as_float(kTargetSun));
            W.sh_c[qs] = make_float4(contrib.x, contrib.y, contrib.z, __uint_as_float(scene));
        }
        if (active && want_ray && qr != 0xFFFFFFFFu) {
            W.ray_o[0][qr] = make_float4(origin.x, origin.y, origin.z, __uint_as_float(w));
            W.ray_d[0][qr] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(scene));
        }
    }
}

// First hits of the learned bounce: the hit's emission (or the sky on a
// miss) aside in rad.xyz, the hit's cache query at the path's query index.
template <bool MAT>
__global__ __launch_bounds__(kWfBlock, 5) void wf_learned_hit(DevScene S, WfBuffers W, LearnedIO io,
                                                              uint32_t in_slot) {
    uint32_t cnt[kParts];
    queue_counts(W.counters, in_slot, W.qcap, cnt);
    uint32_t count = 0;
#pragma unroll
    for (uint32_t j = 0; j < kParts; ++j) count += cnt[j];
    for (uint32_t kq = blockIdx.x * kWfBlock + threadIdx.x; kq < count; kq += gridDim.x * kWfBlock) {
        uint32_t k = queue_entry(cnt, kq, W.qcap);
        float4 ro = W.ray_o[0][k], rd = W.ray_d[0][k];
        float4 hv = W.hits[k];
        uint32_t p = __float_as_uint(ro.w);
        const Light &L = S.scenes[__float_as_uint(rd.w)].light;
        uint32_t i = W.qidx[p];
        f3 d = mk(rd.x, rd.y, rd.z);
        float4 Rv = W.rad[p];
        uint32_t id = __float_as_uint(hv.w);
        f3 term;
        if (id == kInvalid) {
            term = miss_radiance(L, d);
        } else {
            Shading y1;
            shading_data<MAT>(S, id, hv.y, hv.z, -d, y1);
            term = y1.emissive;
            put3(io.q_pos, i, y1.origin_front);
            put3(io.q_nrm, i, y1.frame.n);
            put3(io.q_alb, i, y1.diffuse);
            io.q_w[i] = 1.0f;
        }
        W.rad[p] = make_float4(term.x, term.y, term.z, Rv.w);
    }
}

struct LearnedFinishParams {
    const uint32_t *list;
    uint32_t count, side, reset, at_x0, ramped;
};

// E = sun + (az/pdf) * (Le(y1) or sky) + (az/pdf) * max(0, cache(y1)), or at
// the texel E = sun + max(0, cache(x0)); into the side's payload slot.
__global__ __launch_bounds__(kWfBlock) void wf_learned_finish(DevLayout Ly, WfBuffers W, LearnedIO io,
                                                              float4 *payload, LearnedFinishParams P) {
    for (uint32_t i = blockIdx.x * kWfBlock + threadIdx.x; i < P.count; i += gridDim.x * kWfBlock) {
        uint32_t w = P.list ? P.list[i] : i;
        float4 sv = W.sun[w];
        f3 acc = mk(sv.x, sv.y, sv.z);
        float4 c = io.q_pred[i];
        f3 pred = mk(fmaxf(c.x, 0.0f), fmaxf(c.y, 0.0f), fmaxf(c.z, 0.0f));
        if (P.at_x0) {
            acc = acc + pred * 1.0f;
        } else {
            float4 Rv = W.rad[w];
            float pdf = Rv.w;
            if (pdf > 1e-6f) {
                float weight = W.T[w].w / pdf;
                acc = acc + mk(Rv.x, Rv.y, Rv.z) * weight;
                if (io.q_w[i] > 0.0f) acc = acc + pred * weight;
            }
        }
        uint32_t f = Ly.texel_face[w];
        uint32_t inst = Ly.face_inst[f];
        uint32_t local = w - Ly.front_prefix[inst];
        uint32_t slot = P.side ? Ly.sides[inst].x + local : Ly.instances[inst].z + local;
        uint32_t reset = io.frame ? io.frame[1] : P.reset;
        if (P.ramped) store_ramped(payload, slot, acc, reset);
        else store_accumulated(payload, slot, acc, reset);
    }
}

// Training paths: record j starts like learned_train_paths in nan_learned.hip
// (texel, side, own RNG stream, cosine bounce, no sun term at x0).
template <bool MAT>
__global__ __launch_bounds__(kWfBlock, 6) void wf_train_generate(DevScene S, DevLayout Ly, WfBuffers W,
                                                                 LearnedIO io, uint32_t n, uint32_t iteration,
                                                                 uint32_t at_x0, uint32_t ext_slot) {
    if (io.frame) iteration = io.frame[0];
    uint32_t stride = gridDim.x * kWfBlock;
    uint32_t *ext_count = queue_counter(W.counters, ext_slot, blockIdx.x % kParts);
    uint32_t *unused = queue_counter(W.counters, ext_slot + 1, blockIdx.x % kParts);
    for (uint32_t base = blockIdx.x * kWfBlock; base < n; base += stride) {
        uint32_t j = base + threadIdx.x;
        bool active = j < n;
        bool want_ray = false;
        f3 origin = mk(0.0f, 0.0f, 0.0f), dir = origin;
        uint32_t scene = 0;
        if (active) {
            TrainStart st = train_start(Ly, j, iteration);
            Rng rng;
            TexelSurface ts = texel_surface(Ly, st.w, iteration, 1u, rng);
            bool back = (ts.side.y & 1u) && st.back;
            rng.state = st.rng;
            scene = ts.scene;
            f3 nf = world_flat_normal(S, ts.tri);
            Shading sd0;
            shading_data<MAT>(S, ts.tri, ts.b1, ts.b2, back ? -nf : nf, sd0);
            float ex = rng.next_1d(), ey = rng.next_1d();
            f3 sdir;
            float pdf;
            cosine_sample(ex, ey, sdir, pdf);
            want_ray = pdf > 1e-6f;
            origin = sd0.origin_front;
            dir = to_global(sd0.frame, sdir);
            W.T[j] = make_float4(1.0f, 1.0f, 1.0f, fabsf(sdir.z));
            W.rad[j] = make_float4(0.0f, 0.0f, 0.0f, pdf);
            W.sun[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            W.rng[j] = rng.state;
            if (at_x0) {
                put3(io.tr_pos, j, sd0.origin_front);
                put3(io.tr_nrm, j, sd0.frame.n);
                put3(io.tr_alb, j, sd0.diffuse);
                io.tr_valid[j] = 1u;
            } else {
                put3(io.tr_pos, j, mk(0.0f, 0.0f, 0.0f));
                put3(io.tr_nrm, j, mk(0.0f, 0.0f, 0.0f));
                put3(io.tr_alb, j, mk(0.0f, 0.0f, 0.0f));
                io.tr_valid[j] = 0u;
            }
        }
        uint32_t qs, qr;
        block_append2(false, active && want_ray, unused, ext_count, 2 * W.qcap, W.qcap, qs, qr);
        if (active && want_ray && qr != 0xFFFFFFFFu) {
            W.ray_o[0][qr] = make_float4(origin.x, origin.y, origin.z, __uint_as_float(j));
            W.ray_d[0][qr] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(scene));
        }
    }
}

// Training targets: the suffix radiance of y1, or at the texel
// (suffix + Le(y1)) * az/pdf (the sky * az/pdf on a miss).
__global__ __launch_bounds__(kWfBlock) void wf_train_finish(WfBuffers W, LearnedIO io, uint32_t n, uint32_t at_x0) {
    for (uint32_t j = blockIdx.x * kWfBlock + threadIdx.x; j < n; j += gridDim.x * kWfBlock) {
        float4 Rv = W.rad[j];
        f3 rad = mk(Rv.x, Rv.y, Rv.z);
        f3 target = mk(0.0f, 0.0f, 0.0f);
        if (at_x0) {
            float pdf = Rv.w;
            if (pdf > 1e-6f) {
                float4 le = W.sun[j];
                target = (rad + mk(le.x, le.y, le.z) * 1.0f) * (W.T[j].w / pdf);
            }
        } else if (io.tr_valid[j]) {
            target = rad;
        }
        put3(io.tr_target, j, target);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace nanx {

void wavefront_release(NanTracer *t) {
    WfBuffers *W = t->wf;
    if (!W) return;
    if (WfBuffers *D = W->ds) {
        hipFree(D->ray_o[0]);
        hipFree(D->ray_o[1]);
        hipFree(D->ray_d[0]);
        hipFree(D->ray_d[1]);
        hipFree(D->hits);
        hipFree(D->sh_o);
        hipFree(D->sh_d);
        hipFree(D->sh_c);
        hipFree(D->spill);
        delete D;
    }
    if (W->ev_fork) hipEventDestroy(W->ev_fork);
    if (W->ev_join) hipEventDestroy(W->ev_join);
    if (W->side_stream) hipStreamDestroy(W->side_stream);
    hipFree(W->rng);
    hipFree(W->T);
    hipFree(W->rad);
    hipFree(W->sun);
    hipFree(W->acc);
    hipFree(W->ray_o[0]);
    hipFree(W->ray_o[1]);
    hipFree(W->ray_d[0]);
    hipFree(W->ray_d[1]);
    hipFree(W->hits);
    hipFree(W->sh_o);
    hipFree(W->sh_d);
    hipFree(W->sh_c);
    hipFree(W->counters);
    hipFree(W->frame);
    hipFree(W->spill);
    hipFree(W->qidx);
    delete W;
    t->wf = nullptr;
}

template <int RING, bool MAT>
static int occupancy_blocks(int *per_cu) {
    int a = 0, b = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, wf_extend<RING, false, MAT>, kTravBlock, 0), nullptr, 0);
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, wf_shadow<RING, false, MAT>, kTravBlock, 0), nullptr, 0);
    *per_cu = std::max(1, std::min(a, b));
    return 0;
}

// The whole working set of the wavefront forms (per-path state, queues,
// counters, spill area, the second stream and its events), sized from the
// layout and the launch shape.  Allocated when the tracer is created and
// again when nan_tracer_tune changes the launch shape, never on the execute
// path, so the first execute may be captured into a graph.
static int wavefront_alloc(NanTracer *t) {
    const WfTuning &tu = t->tuning;
    WfBuffers *W = new WfBuffers();
    t->wf = W;
    W->ring = tu.ring;
    W->blocks_per_cu_req = tu.blocks_per_cu;
    uint32_t cap = (uint32_t)std::max<uint64_t>(t->total_front, 1);
    W->capacity = cap;
    size_t c4 = (size_t)cap * 16;
    HIP_TRY(hipMalloc((void **)&W->rng, (size_t)cap * 4), nullptr, 0);
    HIP_TRY(hipMalloc((void **)&W->T, c4), nullptr, 0);
    HIP_TRY(hipMalloc((void **)&W->rad, c4), nullptr, 0);
    HIP_TRY(hipMalloc((void **)&W->sun, c4), nullptr, 0);
    HIP_TRY(hipMalloc((void **)&W->acc, c4), nullptr, 0);
    HIP_TRY(hipMalloc((void **)&W->qidx, (size_t)cap * 4), nullptr, 0);
    // A stage grid that is a multiple of kParts hands partition j the
    // 256-entry chunks c = j (mod kParts) of its input, so no partition
    // receives more than ceil(n / 256 / kParts) * 256 <= n / kParts + 256
    // entries per producer launch (shadow queues take two producers).
    W->qcap = (uint32_t)(((cap + kParts - 1) / kParts + 256 + 255) / 256 * 256);
    size_t q4 = (size_t)W->qcap * kParts * 16;
    for (int k = 0; k < 2; ++k) {
        HIP_TRY(hipMalloc((void **)&W->ray_o[k], q4), nullptr, 0);
        HIP_TRY(hipMalloc((void **)&W->ray_d[k], q4), nullptr, 0);
    }
    HIP_TRY(hipMalloc((void **)&W->hits, q4), nullptr, 0);
    HIP_TRY(hipMalloc((void **)&W->sh_o, 2 * q4), nullptr, 0);
    HIP_TRY(hipMalloc((void **)&W->sh_d, 2 * q4), nullptr, 0);
    HIP_TRY(hipMalloc((void **)&W->sh_c, 2 * q4), nullptr, 0);
    HIP_TRY(hipMalloc((void **)&W->counters, (size_t)kCounterWords * 4), nullptr, 0);
    HIP_TRY(hipMalloc((void **)&W->frame, 2 * 4), nullptr, 0);
    int dev = 0, cus = 0, per_cu = 1;
    HIP_TRY(hipGetDevice(&dev), nullptr, 0);
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), nullptr, 0);
    // the kernel variants the scene runs (MAT: textured or alpha-tested materials)
    const bool mat = t->scene->dev.mat_features != 0u;
    int rc = W->ring == 16 ? (mat ? occupancy_blocks<16, true>(&per_cu) : occupancy_blocks<16, false>(&per_cu))
             : W->ring == 24 ? (mat ? occupancy_blocks<24, true>(&per_cu) : occupancy_blocks<24, false>(&per_cu))
                             : (mat ? occupancy_blocks<32, true>(&per_cu) : occupancy_blocks<32, false>(&per_cu));
    if (rc) return rc;
    if (tu.blocks_per_cu) per_cu = std::min(per_cu, (int)tu.blocks_per_cu);
    W->trav_blocks = (uint32_t)(cus * per_cu);
    int og = 1, os = 1, of = 1;
    if (mat) {
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&og, wf_generate<true>, kWfBlock, 0), nullptr, 0);
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&os, wf_shade<true>, kWfBlock, 0), nullptr, 0);
    } else {
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&og, wf_generate<false>, kWfBlock, 0), nullptr, 0);
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&os, wf_shade<false>, kWfBlock, 0), nullptr, 0);
    }
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&of, wf_finish<false>, kWfBlock, 0), nullptr, 0);
    auto parts_multiple = [](int blocks) { return (uint32_t)std::max<int>(kParts, blocks / kParts * kParts); };
    W->gen_blocks = parts_multiple(cus * std::max(1, og));
    W->shade_blocks = parts_multiple(cus * std::max(1, os));
    W->finish_blocks = (uint32_t)(cus * std::max(1, of));
    W->spill_lanes = W->trav_blocks * kTravBlock;
    HIP_TRY(hipMalloc((void **)&W->spill, (size_t)(kMaxStack - W->ring) * W->spill_lanes * 4), nullptr, 0);
    // The side stream carries two sweeps of the double-sided texels on a
    // quarter of the GPU: with at most 1/8 of the front texels double-sided
    // its work stays within its share.  More runs in one-stream order: with
    // 26% (the 4.0M-triangle combined scene) a split frame took 30.2 ms with
    // quarter-width, 27.2 ms with third-width and 24.9 ms with half-width
    // side grids, against 23.8 ms in one-stream order.
    if (kSideDiv != 0 && t->ds_count != 0 && t->ss_count != 0 && (uint64_t)t->ds_count * kSideLimit <= t->total_front) {
        WfBuffers *D = new WfBuffers(*W);
        W->ds = D;
        D->ds = nullptr;
        D->side_stream = nullptr;
        D->ev_fork = D->ev_join = nullptr;
        D->ray_o[0] = D->ray_o[1] = D->ray_d[0] = D->ray_d[1] = nullptr;
        D->hits = D->sh_o = D->sh_d = D->sh_c = nullptr;
        D->spill = nullptr;
        D->qidx = nullptr;
        auto share = [](uint32_t blocks) { return std::max<uint32_t>(kParts, blocks / kSideDiv / kParts * kParts); };
        D->trav_blocks = share(W->trav_blocks);
        D->gen_blocks = share(W->gen_blocks);
        D->shade_blocks = share(W->shade_blocks);
        D->finish_blocks = std::max<uint32_t>(1u, W->finish_blocks / kSideDiv);
        D->spill_lanes = D->trav_blocks * kTravBlock;
        D->qcap = (uint32_t)(((t->ds_count + kParts - 1) / kParts + 256 + 255) / 256 * 256);
        size_t d4 = (size_t)D->qcap * kParts * 16;
        for (int k = 0; k < 2; ++k) {
            HIP_TRY(hipMalloc((void **)&D->ray_o[k], d4), nullptr, 0);
            HIP_TRY(hipMalloc((void **)&D->ray_d[k], d4), nullptr, 0);
        }
        HIP_TRY(hipMalloc((void **)&D->hits, d4), nullptr, 0);
        HIP_TRY(hipMalloc((void **)&D->sh_o, 2 * d4), nullptr, 0);
        HIP_TRY(hipMalloc((void **)&D->sh_d, 2 * d4), nullptr, 0);
        HIP_TRY(hipMalloc((void **)&D->sh_c, 2 * d4), nullptr, 0);
        HIP_TRY(hipMalloc((void **)&D->spill, (size_t)(kMaxStack - D->ring) * D->spill_lanes * 4), nullptr, 0);
        HIP_TRY(hipStreamCreateWithFlags(&W->side_stream, hipStreamNonBlocking), nullptr, 0);
        HIP_TRY(hipEventCreateWithFlags(&W->ev_fork, hipEventDisableTiming), nullptr, 0);
        HIP_TRY(hipEventCreateWithFlags(&W->ev_join, hipEventDisableTiming), nullptr, 0);
    }
    return 0;
}

int wavefront_prepare(NanTracer *t) {
    const WfTuning &tu = t->tuning;
    if (t->wf && t->wf->ring == tu.ring && t->wf->blocks_per_cu_req == tu.blocks_per_cu) return 0;
    wavefront_release(t);
    if (wavefront_alloc(t)) {
        std::string msg = g_last_error;
        wavefront_release(t);
        return fail(nullptr, 0, "wavefront buffers: %s", msg.c_str());
    }
    return 0;
}

// Persistent grid of a traversal launch: the occupancy-sized grid, or fewer
// blocks (a multiple of kParts) when the queue holds at most max_rays rays.
static uint32_t trav_grid(const WfBuffers &W, uint64_t max_rays) {
    uint64_t need = (grid_for(max_rays, kTravBlock) + kParts - 1) / kParts * kParts;
    return (uint32_t)std::min<uint64_t>(W.trav_blocks, std::max<uint64_t>(need, kParts));
}

// Launches kernel<MAT> with MAT = the scene's material features.
#define NAN_LAUNCH_MAT(kernel, grid, block, stream, ...)                                 \
    do {                                                                                 \
        if (S.mat_features) hipLaunchKernelGGL(kernel<true>, grid, block, 0, stream, __VA_ARGS__);  \
        else hipLaunchKernelGGL(kernel<false>, grid, block, 0, stream, __VA_ARGS__);     \
    } while (0)

template <int RING>
static void launch_extend(const WfBuffers &W, const DevScene &S, uint32_t q, uint32_t count_slot, uint32_t fetch,
                          uint32_t steps, unsigned long long *counters, hipStream_t stream,
                          uint64_t max_rays = ~0ull) {
    uint32_t blocks = trav_grid(W, max_rays);
#define NAN_EXTEND(C, M)                                                                                    \
    hipLaunchKernelGGL((wf_extend<RING, C, M>), dim3(blocks), dim3(kTravBlock), 0, stream, S, W, q, count_slot, \
                       fetch, steps, counters)
    if (counters) {
        if (S.mat_features) NAN_EXTEND(true, true);
        else NAN_EXTEND(true, false);
    } else {
        if (S.mat_features) NAN_EXTEND(false, true);
        else NAN_EXTEND(false, false);
    }
#undef NAN_EXTEND
}

template <int RING>
static void launch_shadow(const WfBuffers &W, const DevScene &S, uint32_t count_slot, uint32_t fetch, uint32_t steps,
                          unsigned long long *counters, hipStream_t stream, uint64_t max_rays = ~0ull) {
    uint32_t blocks = trav_grid(W, max_rays);
#define NAN_SHADOW(C, M)                                                                                    \
    hipLaunchKernelGGL((wf_shadow<RING, C, M>), dim3(blocks), dim3(kTravBlock), 0, stream, S, W, count_slot, \
                       fetch, steps, counters)
    if (counters) {
        if (S.mat_features) NAN_SHADOW(true, true);
        else NAN_SHADOW(true, false);
    } else {
        if (S.mat_features) NAN_SHADOW(false, true);
        else NAN_SHADOW(false, false);
    }
#undef NAN_SHADOW
}

// Queue counter slots one sample of one sweep takes: the first ray and
// shadow queues, then per bounce a refill counter for extend and for shadow
// and the next ray and shadow queues.
static inline uint32_t sample_slots(uint32_t bounces) { return 4 * bounces + 2; }

// The sweeps of one set of paths (`list` null: front texel i is path i) on
// one queue set and stream: per sample, generate, `bounces` rounds of extend /
// shade / shadow, finish.  Counter slots come from [slot, slot_end); when a
// sample would not fit, the range [slot_begin, slot_end) is zeroed again on
// the stream and reused, so any number of samples and sweeps fits a range
// that holds one sample (checked by the caller before the first launch).
static int run_sweeps(NanTracer *t, WfBuffers &W, float4 *payload, const uint32_t *list, uint32_t count,
                      uint32_t side, uint32_t iteration, uint32_t spp, uint32_t bounces, uint32_t reset,
                      unsigned long long *counters, hipStream_t stream, uint32_t &slot, uint32_t slot_begin,
                      uint32_t slot_end) {
    if (count == 0) return 0;
    const DevScene &S = t->scene->dev;
    DevLayout Ly = dev_layout(t);
    uint32_t nspp = spp > 1 ? spp : 1;
    uint32_t steps = std::max(1u, t->tuning.refill_steps);
    uint32_t round8 = (grid_for(count, kWfBlock) + kParts - 1) / kParts * kParts;
    uint32_t gen_blocks = std::min<uint32_t>(W.gen_blocks, round8);
    uint32_t shade_blocks = std::min<uint32_t>(W.shade_blocks, round8);
    uint32_t finish_blocks = std::min<uint32_t>(W.finish_blocks, grid_for(count, kWfBlock));
    for (uint32_t k = 0; k < nspp; ++k) {
        if (slot + sample_slots(bounces) > slot_end) {
            // every earlier launch on this stream is done with the range by
            // the time the zeroing runs (stream order): start it over
            hipLaunchKernelGGL(wf_zero_counters, dim3(64), dim3(1024), 0, stream,
                               W.counters + (size_t)slot_begin * kParts * kCtrStride,
                               (slot_end - slot_begin) * kParts * kCtrStride);
            slot = slot_begin;
        }
        uint32_t ext_slot = slot++, sh_slot = slot++;
        GenParams gp{t->device_frame ? W.frame : nullptr, list, count, side, k, iteration, spp, ext_slot, sh_slot};
        {
            StageTimer tm(t, stream, kStageGenerate);
            NAN_LAUNCH_MAT(wf_generate, dim3(gen_blocks), dim3(kWfBlock), stream, S, Ly, W, gp);
        }
        uint32_t q = 0;
        for (uint32_t b = 1; b <= bounces; ++b) {
            uint32_t fetch = slot++;
            {
                StageTimer tm(t, stream, kStageExtend);
                if (W.ring == 16) launch_extend<16>(W, S, q, ext_slot, fetch, steps, counters, stream, count);
                else if (W.ring == 24) launch_extend<24>(W, S, q, ext_slot, fetch, steps, counters, stream, count);
                else launch_extend<32>(W, S, q, ext_slot, fetch, steps, counters, stream, count);
            }
            uint32_t next_ext = slot++;
            ShadeParams sp{q, 1u - q, ext_slot, next_ext, sh_slot, b, bounces};
            {
                StageTimer tm(t, stream, kStageShade);
                NAN_LAUNCH_MAT(wf_shade, dim3(shade_blocks), dim3(kWfBlock), stream, S, W, sp);
            }
            uint32_t sfetch = slot++;
            {
                StageTimer tm(t, stream, kStageShadow);
                // at most one sun ray per path from generate (bounce 1) and one from shade
                uint64_t max_sh = 2ull * count;
                if (W.ring == 16) launch_shadow<16>(W, S, sh_slot, sfetch, steps, counters, stream, max_sh);
                else if (W.ring == 24) launch_shadow<24>(W, S, sh_slot, sfetch, steps, counters, stream, max_sh);
                else launch_shadow<32>(W, S, sh_slot, sfetch, steps, counters, stream, max_sh);
            }
            sh_slot = slot++;
            ext_slot = next_ext;
            q = 1u - q;
        }
        FinishParams fp{t->device_frame ? W.frame : nullptr, list, count, side, k, nspp, reset,
                        t->scene->batch ? t->payload_table : nullptr};
        StageTimer tm(t, stream, kStageFinish);
        if (counters)
            hipLaunchKernelGGL(wf_finish<true>, dim3(finish_blocks), dim3(kWfBlock), 0, stream, Ly, W, payload, fp,
                               counters);
        else
            hipLaunchKernelGGL(wf_finish<false>, dim3(finish_blocks), dim3(kWfBlock), 0, stream, Ly, W, payload, fp,
                               counters);
    }
    return 0;
}

// One frame.  A double-sided texel's back path starts from the random state
// its front path ended with (texture_space_path_tracer.slang:204-214), so its
// back sweep waits for its front sweep; the back sweep is small and made of
// short, latency-bound launches.  With both kinds of texels present the
// double-sided texels' two sweeps therefore run on a second stream, forked
// from and joined back into the caller's stream, beside the single-sided
// texels' front sweep, which fills the GPU meanwhile.  Every path touches
// only its own per-path entries and payload slots, and each stream has its
// own queues, spill area and counter slots, so the results are those of the
// one-stream order.  Stage timing (profiling) uses the one-stream order.
// The random streams of the current call, on every queue set.
static void set_rng_mode(NanTracer *t) {
    uint32_t pcg = t->rng_mode == NAN_RNG_PCG ? 1u : 0u;
    t->wf->rng_pcg = pcg;
    if (t->wf->ds) t->wf->ds->rng_pcg = pcg;
}

int wavefront_execute(NanTracer *t, float4 *payload, uint32_t iteration, uint32_t spp, uint32_t bounces,
                      uint32_t reset, unsigned long long *counters, hipStream_t stream) {
    if (t->scene->info.max_stack > (uint32_t)kMaxStack)
        return fail(nullptr, 0, "wavefront: BVH needs a %u-entry stack (max %d)", t->scene->info.max_stack,
                    kMaxStack);
    if (t->scene->batch && !t->payload_table_set)
        return fail(nullptr, 0, "wavefront: batch payloads are not set (nan_batch_set_payloads)");
    if (!t->wf) return fail(nullptr, 0, "wavefront: working set missing (a failed nan_tracer_tune?)");
    // each stream's counter range must hold one sample's queues; checked
    // before any launch, so a refused call leaves the payload untouched
    if ((uint64_t)sample_slots(bounces) > kCounterSlots / 2)
        return fail(nullptr, 0, "wavefront: max_bounces %u exceeds the %u the queue counters hold", bounces,
                    (kCounterSlots / 2 - 2) / 4);
    WfBuffers &W = *t->wf;
    set_rng_mode(t);
    bool split = W.ds != nullptr && !t->profile;
    hipLaunchKernelGGL(wf_zero_counters, dim3(64), dim3(1024), 0, stream, W.counters, kCounterWords);
    if (!split) {
        uint32_t slot = 0;
        if (run_sweeps(t, W, payload, nullptr, (uint32_t)t->total_front, 0, iteration, spp, bounces, reset,
                       counters, stream, slot, 0, kCounterSlots) ||
            run_sweeps(t, W, payload, t->ds_list, t->ds_count, 1, iteration, spp, bounces, reset, counters, stream,
                       slot, 0, kCounterSlots))
            return -1;
    } else {
        WfBuffers &D = *W.ds;
        hipStream_t side_stream = W.side_stream;
        HIP_TRY(hipEventRecord(W.ev_fork, stream), nullptr, 0);
        HIP_TRY(hipStreamWaitEvent(side_stream, W.ev_fork, 0), nullptr, 0);
        uint32_t dslot = kCounterSlots / 2, sslot = 0;
        int rc = run_sweeps(t, D, payload, t->ds_list, t->ds_count, 0, iteration, spp, bounces, reset, counters,
                            side_stream, dslot, kCounterSlots / 2, kCounterSlots);
        if (!rc)
            rc = run_sweeps(t, D, payload, t->ds_list, t->ds_count, 1, iteration, spp, bounces, reset, counters,
                            side_stream, dslot, kCounterSlots / 2, kCounterSlots);
        if (!rc)
            rc = run_sweeps(t, W, payload, t->ss_list, t->ss_count, 0, iteration, spp, bounces, reset, counters,
                            stream, sslot, 0, kCounterSlots / 2);
        // the join is recorded whatever happened, so a capture never ends with the side stream forked
        hipError_t e1 = hipEventRecord(W.ev_join, side_stream);
        hipError_t e2 = hipStreamWaitEvent(stream, W.ev_join, 0);
        if (rc) return -1;
        HIP_TRY(e1, nullptr, 0);
        HIP_TRY(e2, nullptr, 0);
    }
    if (t->device_frame) hipLaunchKernelGGL(wf_advance_frame, dim3(1), dim3(64), 0, stream, W.frame);
    return check_launch("wavefront");
}

const uint32_t *wavefront_frame(const NanTracer *t) { return t->device_frame && t->wf ? t->wf->frame : nullptr; }

void wavefront_advance_frame(NanTracer *t, hipStream_t stream) {
    if (t->device_frame && t->wf) hipLaunchKernelGGL(wf_advance_frame, dim3(1), dim3(64), 0, stream, t->wf->frame);
}

int wavefront_set_frame(NanTracer *t, uint32_t iteration, uint32_t reset) {
    if (!t->wf) return fail(nullptr, 0, "wavefront: working set missing (a failed nan_tracer_tune?)");
    uint32_t v[2] = {iteration, reset ? 1u : 0u};
    HIP_TRY(hipMemcpy(t->wf->frame, v, sizeof(v), hipMemcpyHostToDevice), nullptr, 0);
    return 0;
}


static int learned_checks(NanTracer *t) {
    if (t->scene->info.max_stack > (uint32_t)kMaxStack)
        return fail(nullptr, 0, "wavefront: BVH needs a %u-entry stack (max %d)", t->scene->info.max_stack,
                    kMaxStack);
    if (!t->wf) return fail(nullptr, 0, "wavefront: working set missing (a failed nan_tracer_tune?)");
    set_rng_mode(t);
    return 0;
}

static void launch_extend_any(const WfBuffers &W, const DevScene &S, uint32_t q, uint32_t count_slot, uint32_t fetch,
                              uint32_t steps, hipStream_t stream, uint64_t max_rays) {
    if (W.ring == 16) launch_extend<16>(W, S, q, count_slot, fetch, steps, nullptr, stream, max_rays);
    else if (W.ring == 24) launch_extend<24>(W, S, q, count_slot, fetch, steps, nullptr, stream, max_rays);
    else launch_extend<32>(W, S, q, count_slot, fetch, steps, nullptr, stream, max_rays);
}

static void launch_shadow_any(const WfBuffers &W, const DevScene &S, uint32_t count_slot, uint32_t fetch,
                              uint32_t steps, hipStream_t stream, uint64_t max_rays) {
    if (W.ring == 16) launch_shadow<16>(W, S, count_slot, fetch, steps, nullptr, stream, max_rays);
    else if (W.ring == 24) launch_shadow<24>(W, S, count_slot, fetch, steps, nullptr, stream, max_rays);
    else launch_shadow<32>(W, S, count_slot, fetch, steps, nullptr, stream, max_rays);
}

int wavefront_learned_train_paths(NanTracer *t, const LearnedIO &io, uint32_t iteration, uint32_t n,
                                  uint32_t bounces, bool at_x0, hipStream_t stream) {
    if (learned_checks(t)) return -1;
    WfBuffers &W = *t->wf;
    if (n == 0) return 0;
    if (n > W.capacity) return fail(nullptr, 0, "wavefront: %u training paths exceed the %u path slots", n, W.capacity);
    if (sample_slots(bounces) + 2 > kCounterSlots) return fail(nullptr, 0, "wavefront: too many stages");
    const DevScene &S = t->scene->dev;
    DevLayout Ly = dev_layout(t);
    uint32_t steps = std::max(1u, t->tuning.refill_steps);
    uint32_t round8 = (grid_for(n, kWfBlock) + kParts - 1) / kParts * kParts;
    uint32_t gen_blocks = std::min<uint32_t>(W.gen_blocks, round8);
    uint32_t shade_blocks = std::min<uint32_t>(W.shade_blocks, round8);
    StageTimer tm(t, stream, kStageLearnedPaths);
    hipLaunchKernelGGL(wf_zero_counters, dim3(64), dim3(1024), 0, stream, W.counters, kCounterWords);
    uint32_t slot = 0;
    uint32_t ext_slot = slot;
    slot += 2;  // the generator's ray queue and an unused second queue
    NAN_LAUNCH_MAT(wf_train_generate, dim3(gen_blocks), dim3(kWfBlock), stream, S, Ly, W, io, n, iteration,
                   at_x0 ? 1u : 0u, ext_slot);
    uint32_t sh_slot = slot++;
    uint32_t q = 0;
    for (uint32_t b = 1; b <= bounces; ++b) {
        uint32_t fetch = slot++;
        launch_extend_any(W, S, q, ext_slot, fetch, steps, stream, n);
        uint32_t next_ext = slot++;
        ShadeParams sp{q, 1u - q, ext_slot, next_ext, sh_slot, b, bounces, at_x0 ? 2u : 1u, io};
        NAN_LAUNCH_MAT(wf_shade, dim3(shade_blocks), dim3(kWfBlock), stream, S, W, sp);
        uint32_t sfetch = slot++;
        launch_shadow_any(W, S, sh_slot, sfetch, steps, stream, n);
        sh_slot = slot++;
        ext_slot = next_ext;
        q = 1u - q;
    }
    hipLaunchKernelGGL(wf_train_finish, dim3(std::min<uint32_t>(W.finish_blocks, grid_for(n, kWfBlock))),
                       dim3(kWfBlock), 0, stream, W, io, n, at_x0 ? 1u : 0u);
    return check_launch("wavefront training paths");
}

int wavefront_learned_side(NanTracer *t, const LearnedIO &io, uint32_t iteration, uint32_t side, bool at_x0,
                           uint32_t *count_out, hipStream_t stream) {
    if (learned_checks(t)) return -1;
    WfBuffers &W = *t->wf;
    uint32_t count = side ? t->ds_count : (uint32_t)t->total_front;
    *count_out = count;
    if (count == 0) return 0;
    const DevScene &S = t->scene->dev;
    DevLayout Ly = dev_layout(t);
    uint32_t steps = std::max(1u, t->tuning.refill_steps);
    uint32_t round8 = (grid_for(count, kWfBlock) + kParts - 1) / kParts * kParts;
    StageTimer tm(t, stream, kStageFirstBounce);
    hipLaunchKernelGGL(wf_zero_counters, dim3(64), dim3(1024), 0, stream, W.counters, kCounterWords);
    LearnedGenParams gp{side ? t->ds_list : nullptr, count, side, iteration, 0u, 1u, at_x0 ? 1u : 0u};
    NAN_LAUNCH_MAT(wf_learned_generate, dim3(std::min<uint32_t>(W.gen_blocks, round8)), dim3(kWfBlock), stream, S, Ly,
                   W, io, gp);
    if (!at_x0) {
        launch_extend_any(W, S, 0u, 0u, 2u, steps, stream, count);
        NAN_LAUNCH_MAT(wf_learned_hit, dim3(std::min<uint32_t>(W.shade_blocks, round8)), dim3(kWfBlock), stream, S,
                       W, io, 0u);
    }
    launch_shadow_any(W, S, 1u, 3u, steps, stream, count);
    return check_launch("wavefront learned samples");
}

int wavefront_learned_finish(NanTracer *t, const LearnedIO &io, float4 *payload, uint32_t side, uint32_t reset,
                             bool at_x0, bool ramped, hipStream_t stream) {
    WfBuffers &W = *t->wf;
    uint32_t count = side ? t->ds_count : (uint32_t)t->total_front;
    if (count == 0) return 0;
    StageTimer tm(t, stream, kStageFinish);
    LearnedFinishParams fp{side ? t->ds_list : nullptr, count, side, reset, at_x0 ? 1u : 0u, ramped ? 1u : 0u};
    hipLaunchKernelGGL(wf_learned_finish, dim3(std::min<uint32_t>(W.finish_blocks, grid_for(count, kWfBlock))),
                       dim3(kWfBlock), 0, stream, dev_layout(t), W, io, payload, fp);
    return check_launch("wavefront learned finish");
}

}  // namespace nanx
