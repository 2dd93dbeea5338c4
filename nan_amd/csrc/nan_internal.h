// Internal host/device declarations shared by the translation units of
// libnan_hip (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "nan_device.h"
#include "../../include/nan_hip.h"

using namespace nan;

namespace nanx {

inline thread_local std::string g_last_error;

inline int fail(char *err, size_t cap, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    if (err && cap) {
        strncpy(err, buf, cap - 1);
        err[cap - 1] = '\0';
    }
    return -1;
}

#define HIP_TRY(expr, err, cap)                                                                  \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess) return fail(err, cap, "%s: %s", #expr, hipGetErrorString(_e));     \
    } while (0)

template <typename T>
inline int upload(T **dst, const void *src, size_t bytes, char *err, size_t cap) {
    *dst = nullptr;
    if (bytes == 0) bytes = 16;
    HIP_TRY(hipMalloc((void **)dst, bytes), err, cap);
    if (src) HIP_TRY(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), err, cap);
    return 0;
}

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// True while `s` is being captured into a graph, when nothing may allocate
// or synchronise the device.  A failed query counts as capturing.
inline bool stream_capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess) {
        (void)hipGetLastError();
        return true;
    }
    return st != hipStreamCaptureStatusNone;
}

constexpr int kTraceBlock = 128;
constexpr int kFlatBlock = 256;
constexpr uint32_t kMaterialWords = 12;  // material record, 3 float4 (nan_hip.h)

int wavefront_set_frame(NanTracer *t, uint32_t iteration, uint32_t reset);

// Launch shape of the wavefront traversal kernels (nan_tracer_tune).
struct WfTuning {
    uint32_t blocks_per_cu = 0;  // persistent grid: 0 = occupancy limit
    uint32_t ring = 16;          // LDS stack ring entries per lane: 16, 24 or 32
    uint32_t refill_steps = 16;  // traversal steps between lane refills
};

}  // namespace nanx

struct NanScene {
    DevScene dev{};
    std::vector<void *> allocations;
    NanSceneInfo info{};
    uint32_t stack_depth = 0;
    std::vector<uint32_t> instance_triangles;  // triangles of each instance's mesh
    // host copies used to assemble batches of scenes
    std::vector<uint32_t> h_nodes;
    std::vector<float> h_tris, h_shade, h_xf, h_itf;
    std::vector<uint32_t> h_mats;
    std::vector<float> h_tri_uv;                // 8 floats per triangle; empty without textures
    std::vector<uint32_t> h_textures, h_texels;  // texture table (4 words each) and RGBA8 texels
    bool batch = false;  // assembled from several scenes (wavefront only)
    float lighting[16] = {};            // the lighting record last set
    uint32_t roulette = 0;              // Russian roulette on this scene's paths (1: on)
    const float4 *sky_lut = nullptr;    // sky-view table (caller-owned device memory) or null
    ~NanScene() {
        for (void *p : allocations) hipFree(p);
    }
};

struct NanTracer {
    NanScene *scene = nullptr;
    uint32_t rng_mode = NAN_RNG_REFERENCE;  // random streams of the current call
    uint2 *faces = nullptr;
    uint4 *instances = nullptr;
    uint2 *sides = nullptr;
    uint32_t *texel_face = nullptr;
    uint32_t *face_inst = nullptr;
    uint32_t *front_prefix = nullptr;
    uint32_t *ds_list = nullptr;      // front texels of double-sided instances
    uint32_t ds_count = 0;
    uint32_t *ss_list = nullptr;      // the other front texels (ascending)
    uint32_t ss_count = 0;
    // batches of scenes (nan_batch_create)
    uint32_t scene_count = 1;
    uint32_t *face_scene = nullptr;
    uint32_t *scene_inst_base = nullptr;
    bool device_frame = false;         // iteration/reset come from a device counter (graph replay)
    float4 **payload_table = nullptr;  // device array of per-scene payload pointers
    bool payload_table_set = false;
    NanScene *owned_scene = nullptr;   // the batch's concatenated scene
    // host copies of the layout tables (to assemble batches)
    std::vector<uint32_t> h_faces, h_instances, h_sides, h_texel_face, h_face_inst, h_prefix, h_ds;
    uint32_t face_count = 0, instance_count = 0;
    uint64_t total_front = 0, total_payload = 0;
    int mode = 1;                     // 0 megakernel, 1 wavefront
    struct WfBuffers *wf = nullptr;   // wavefront working set (wavefront_prepare)
    struct LearnedBuffers *learned = nullptr;  // learned-cache working set (nan_tst_learned_reserve)
    nanx::WfTuning tuning;
    // stage timing: (stage, start, end) event triples recorded while profiling
    bool profile = false;
    std::vector<hipEvent_t> event_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> stage_events;
    size_t events_used = 0;
    ~NanTracer();
};

namespace nanx {
void wavefront_release(NanTracer *t);
void learned_release(NanTracer *t);
}

inline NanTracer::~NanTracer() {
    nanx::wavefront_release(this);
    nanx::learned_release(this);
    for (hipEvent_t e : event_pool) hipEventDestroy(e);
    hipFree(face_scene);
    hipFree(scene_inst_base);
    hipFree(payload_table);
    delete owned_scene;
    {
        hipFree(ds_list);
        hipFree(ss_list);
        hipFree(faces);
        hipFree(instances);
        hipFree(sides);
        hipFree(texel_face);
        hipFree(face_inst);
        hipFree(front_prefix);
    }
}

struct DevLayout {
    const uint2 *faces;
    const uint4 *instances;        // face_offset global; texel_offset scene-local
    const uint2 *sides;            // back_texel_offset scene-local
    const uint32_t *texel_face;    // global face of each flat front texel
    const uint32_t *face_inst;     // global instance of each face
    const uint32_t *front_prefix;  // flat index of each instance's first front texel
    uint64_t total_front;
    const uint32_t *face_scene;       // scene of each face (null: one scene)
    const uint32_t *scene_inst_base;  // first global instance of each scene (null: one scene)
    uint32_t rng_pcg;                 // 1: per-texel PCG streams, 0: the reference LCG
};


namespace nanx {

inline DevLayout dev_layout(const NanTracer *t) {
    DevLayout L;
    L.faces = t->faces;
    L.instances = t->instances;
    L.sides = t->sides;
    L.texel_face = t->texel_face;
    L.face_inst = t->face_inst;
    L.front_prefix = t->front_prefix;
    L.total_front = t->total_front;
    L.face_scene = t->face_scene;
    L.scene_inst_base = t->scene_inst_base;
    L.rng_pcg = t->rng_mode == NAN_RNG_PCG ? 1u : 0u;
    return L;
}

// Stage ids of nan_tracer_stage_times.
enum Stage {
    kStageGenerate = 0,
    kStageExtend,
    kStageShade,
    kStageShadow,
    kStageFinish,
    kStageMegakernel,
    kStageLearnedPaths,
    kStageCacheTrain,
    kStageFirstBounce,
    kStageCacheQuery,
    kStageCount
};

// Brackets one kernel launch with events when the tracer is profiling.
struct StageTimer {
    NanTracer *t;
    hipStream_t stream;
    int stage;
    hipEvent_t end = nullptr;
    StageTimer(NanTracer *t_, hipStream_t s, int st) : t(t_), stream(s), stage(st) {
        if (!t->profile) return;
        while (t->event_pool.size() < t->events_used + 2) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) {
                t->profile = false;
                return;
            }
            t->event_pool.push_back(e);
        }
        hipEvent_t start = t->event_pool[t->events_used++];
        end = t->event_pool[t->events_used++];
        hipEventRecord(start, stream);
        t->stage_events.push_back({stage, {start, end}});
    }
    ~StageTimer() {
        if (end) hipEventRecord(end, stream);
    }
};

inline uint32_t grid_for(uint64_t n, int block) { return (uint32_t)((n + block - 1) / block); }

// Lighting record (float[16]) plus the scene's sky-view table -> kernel Light.
Light make_light(const float *lighting16, const float4 *sky_lut);
float atmosphere_horizon();

inline int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(nullptr, 0, "%s launch failed: %s", what, hipGetErrorString(e));
    return 0;
}

// Wavefront tracer (nan_wavefront.hip).  wavefront_prepare allocates the
// working set for the tracer's layout and launch shape (again after a launch
// shape change); the execute paths only use it.
int wavefront_prepare(NanTracer *t);
int wavefront_execute(NanTracer *t, float4 *payload, uint32_t iteration, uint32_t spp, uint32_t bounces,
                      uint32_t reset, unsigned long long *counters, hipStream_t stream);

// Device buffers of the learned forms (plain pointers, owned by nan_learned.hip).
struct LearnedIO {
    float *tr_pos, *tr_nrm, *tr_alb, *tr_target;  // training records [n][3]
    uint32_t *tr_valid;
    float *q_pos, *q_nrm, *q_alb, *q_w;  // cache queries [n][3], weight (0: none)
    float4 *q_pred;
    const uint32_t *frame;  // device [iteration, reset] in device-frame mode, else null (host values)
};

// Learned forms on the wavefront queues (nan_wavefront.hip), 1 sample per
// texel, no emitter light sampling; same records, queries and payloads as the
// per-texel kernels of nan_learned.hip bit for bit.
//   train_paths: n training records (one plain path each) into io.tr_*.
//   side: one side's texel samples: sun term and first bounce (or, at_x0, the
//     texel itself) with their cache queries io.q_* at the side's path index;
//     *count = paths of the side.
//   finish: that side's estimates (plus max(0, io.q_pred)) into the payload.
int wavefront_learned_train_paths(NanTracer *t, const LearnedIO &io, uint32_t iteration, uint32_t n,
                                  uint32_t bounces, bool at_x0, hipStream_t stream);
int wavefront_learned_side(NanTracer *t, const LearnedIO &io, uint32_t iteration, uint32_t side, bool at_x0,
                           uint32_t *count, hipStream_t stream);
int wavefront_learned_finish(NanTracer *t, const LearnedIO &io, float4 *payload, uint32_t side, uint32_t reset,
                             bool at_x0, bool ramped, hipStream_t stream);
// Device-frame mode: the tracer's device [iteration, reset] (null when off)
// and the end-of-frame step to the next iteration.
const uint32_t *wavefront_frame(const NanTracer *t);
void wavefront_advance_frame(NanTracer *t, hipStream_t stream);


}  // namespace nanx
