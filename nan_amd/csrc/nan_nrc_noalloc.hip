// Cache queries that allocate nothing: nan_nrc_infer with its scratch (the
// encoded network input and the raw network outputs) in a file-scope device
// array instead of buffers the library grows on demand.  For callers that
// cannot allocate device memory: code under graph capture, tight serving
// loops, harnesses that forbid it.
//
// Every device the library runs on holds one copy of `scratch` (7 MiB) until
// the process ends.  Its device address is looked up once per device, with
// that device current, on the entry point's first call there; host code uses
// only that address (the array's name is a host-side stand-in).  All calls on
// one device share the array: calls that overlap on one device (on different
// streams, or from different threads without ordering) race.  Every entry of
// the groups a call uses is written before it is read (padding rows
// included), so nothing has to start at zero.
#include <mutex>

#include "nan_nrc_internal.h"

using namespace nanx;

namespace {

constexpr uint32_t kNoAllocMaxSamples = 65536;
constexpr uint32_t kNoAllocGroups = kNoAllocMaxSamples / 64;
constexpr size_t kEncBytes = (size_t)kNoAllocGroups * nanx::kNrcEncGroup * sizeof(_Float16);  // 6 MiB
constexpr size_t kPredBytes = (size_t)kNoAllocMaxSamples * sizeof(float4);                      // 1 MiB
static_assert(kEncBytes % 256 == 0, "prediction block must stay 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kEncBytes + kPredBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void *g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per device).
int device_scratch(void **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), nullptr, 0);
    if (dev < 0 || dev >= kMaxDevices) return fail(nullptr, 0, "nan_nrc_infer_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(nullptr, 0, "nan_nrc_infer_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
        g_scratch[dev] = ptr;
    }
    *out = g_scratch[dev];
    return 0;
}

}  // namespace

extern "C" {

NAN_API int nan_nrc_infer_noalloc(NanNrc *nrc, const float *pos, const float *nrm, const float *alb, float *out,
                                  uint32_t n, void *stream) {
    if (!nrc || (n && (!pos || !nrm || !alb || !out))) return fail(nullptr, 0, "nan_nrc_infer_noalloc: null argument");
    if (n > kNoAllocMaxSamples)
        return fail(nullptr, 0, "nan_nrc_infer_noalloc: %u samples exceed the maximum of %u per call", n,
                    kNoAllocMaxSamples);
    if (n == 0) return 0;
    void *base = nullptr;
    if (device_scratch(&base)) return -1;
    _Float16 *enc = static_cast<_Float16 *>(base);
    float4 *pred = reinterpret_cast<float4 *>(static_cast<unsigned char *>(base) + kEncBytes);
    if (nrc_infer_into(nrc, pos, nrm, alb, n, enc, kNoAllocGroups, pred, out, as_stream(stream))) return -1;
    return check_launch("nrc infer (no allocation)");
}

}  // extern "C"
