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

constexpr int kTraceBlock = 128;
constexpr int kFlatBlock = 256;

}  // namespace nanx

struct NanScene {
    DevScene dev{};
    std::vector<void *> allocations;
    NanSceneInfo info{};
    uint32_t stack_depth = 0;
    std::vector<uint32_t> instance_triangles;  // triangles of each instance's mesh
    ~NanScene() {
        for (void *p : allocations) hipFree(p);
    }
};

struct NanTracer {
    NanScene *scene = nullptr;
    uint2 *faces = nullptr;
    uint4 *instances = nullptr;
    uint2 *sides = nullptr;
    uint32_t *texel_face = nullptr;
    uint32_t *face_inst = nullptr;
    uint32_t *front_prefix = nullptr;
    uint32_t *ds_list = nullptr;      // front texels of double-sided instances
    uint32_t ds_count = 0;
    uint32_t face_count = 0, instance_count = 0;
    uint64_t total_front = 0, total_payload = 0;
    int mode = 1;                     // 0 megakernel, 1 wavefront
    struct WfBuffers *wf = nullptr;   // wavefront working set, allocated on first use
    // stage timing: (stage, start, end) event triples recorded while profiling
    bool profile = false;
    std::vector<hipEvent_t> event_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> stage_events;
    size_t events_used = 0;
    ~NanTracer();
};

namespace nanx {
void wavefront_release(NanTracer *t);
}

inline NanTracer::~NanTracer() {
    nanx::wavefront_release(this);
    for (hipEvent_t e : event_pool) hipEventDestroy(e);
    {
        hipFree(ds_list);
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
    const uint4 *instances;
    const uint2 *sides;
    const uint32_t *texel_face;
    const uint32_t *face_inst;
    const uint32_t *front_prefix;
    uint64_t total_front;
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
    return L;
}

// Stage ids of nan_tracer_stage_times.
enum Stage { kStageGenerate = 0, kStageExtend, kStageShade, kStageShadow, kStageFinish, kStageMegakernel, kStageCount };

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

inline int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(nullptr, 0, "%s launch failed: %s", what, hipGetErrorString(e));
    return 0;
}

// Wavefront tracer (nan_wavefront.hip).
int wavefront_execute(NanTracer *t, float4 *payload, uint32_t iteration, uint32_t spp, uint32_t bounces,
                      uint32_t reset, unsigned long long *counters, hipStream_t stream);

}  // namespace nanx
