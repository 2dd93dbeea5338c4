// libnan_hip: kernels and C ABI of the MI355X texture-space path tracer.
//
// Kernels (one HIP thread per unit of work, wave64, LDS traversal stacks):
//   tst_megakernel  one progressive iteration over all front texels
//                   (texture_space_path_tracer.slang:140-216)
//   tst_raygen      texel ray generation only (slang:142-199)
//   trace_rays      closest hit + occlusion for a ray batch (scene.slang:1017-1172)
//   rgb9e5_pack     float3 slot -> uint32 (mesh_colors_rgb9e5_pack.slang:7-16)
//   resolve         primary rays into the camera (mesh_colors_resolve.slang:29-73)
//   accumulate / tonemap (accumulator.slang, tone_mapper.slang)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "nan_bvh.h"
#include "nan_internal.h"
#include "nan_tracer.h"

using namespace nanx;

template <int STACK, int BLOCK, bool COUNT, bool MAT>
__global__ __launch_bounds__(BLOCK) void tst_megakernel(DevScene S, DevLayout Ly, float4 *payload,
                                                        uint32_t iteration, uint32_t spp, uint32_t max_bounces,
                                                        uint32_t reset, unsigned long long *counters) {
    __shared__ uint32_t stack[STACK * BLOCK];
    uint32_t *col = stack + threadIdx.x;
    PathCounters pc{0, 0, 0, 0, 0, 0, 0};
    uint64_t w = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (w < Ly.total_front) {
        Rng rng;
        TexelSurface ts = texel_surface(Ly, w, iteration, spp, rng);
        uint32_t nspp = spp > 1u ? spp : 1u;
        f3 n = world_flat_normal(S, ts.tri);
        const bool roulette = S.scenes[ts.scene].roulette != 0u;
        f3 e = estimate_irradiance<STACK, BLOCK, COUNT, MAT>(S, ts.tri, ts.b1, ts.b2, n, nspp, max_bounces, rng, col, pc,
                                                        roulette);
        store_accumulated(payload, ts.global_texel, e, reset);
        if (COUNT) pc.slots++;
        if (ts.side.y & 1u) {
            f3 eb = estimate_irradiance<STACK, BLOCK, COUNT, MAT>(S, ts.tri, ts.b1, ts.b2, -n, nspp, max_bounces, rng, col,
                                                             pc, roulette);
            store_accumulated(payload, ts.side.x + ts.local, eb, reset);
            if (COUNT) pc.slots++;
        }
    }
    if (COUNT) flush_counters(pc, counters);
}

__global__ void tst_raygen_kernel(DevLayout Ly, uint32_t iteration, uint32_t spp, uint4 *out, uint32_t *rng_out) {
    uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= Ly.total_front) return;
    Rng rng;
    TexelSurface ts = texel_surface(Ly, w, iteration, spp, rng);
    out[w] = make_uint4(ts.instance, ts.prim, __float_as_uint(ts.b1), __float_as_uint(ts.b2));
    if (rng_out) rng_out[w] = rng.state;
}

template <int STACK, int BLOCK, bool MAT>
__global__ __launch_bounds__(BLOCK) void trace_rays_kernel(DevScene S, const float4 *rays, uint64_t n, float4 *hits,
                                                           uint2 *ids, uint32_t *occluded) {
    __shared__ uint32_t stack[STACK * BLOCK];
    uint32_t *col = stack + threadIdx.x;
    uint64_t k = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (k >= n) return;
    float4 a = rays[2 * k], b = rays[2 * k + 1];
    Ray r;
    r.o = mk(a.x, a.y, a.z);
    r.tmin = a.w;
    r.d = mk(b.x, b.y, b.z);
    r.tmax = b.w;
    ray_setup(r);
    Hit h{r.tmax, 0.0f, 0.0f, kInvalid};
    TraversalStats st{0, 0};
    bool hit = traverse<STACK, BLOCK, false, false, MAT>(S, r, h, col, st);
    if (hit) {
        hits[k] = make_float4(h.t, h.b1, h.b2, 0.0f);
        ids[k] = S.tri_ids[h.id];
    } else {
        hits[k] = make_float4(-1.0f, 0.0f, 0.0f, 0.0f);
        ids[k] = make_uint2(kInvalid, kInvalid);
    }
    if (occluded) {
        Hit o{r.tmax, 0.0f, 0.0f, kInvalid};
        occluded[k] = traverse<STACK, BLOCK, true, false, MAT>(S, r, o, col, st) ? 1u : 0u;
    }
}

__global__ void rgb9e5_pack_kernel(const float4 *payload, uint32_t *out, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = payload[i];
    out[i] = pack_rgb9e5(p.x, p.y, p.z);
}

__global__ void rgb9e5_unpack_kernel(const uint32_t *in, float4 *out, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 v = unpack_rgb9e5(in[i]);
    out[i] = make_float4(v.x, v.y, v.z, 0.0f);
}

// ---------------------------------------------------------------------------
// resolve
// ---------------------------------------------------------------------------
struct DevCamera {
    f3 pos, u, v, w;
    float jx, jy, near, far;
};

__device__ f3 sky_with_sun_disk(const Light &L, f3 dir) {
    f3 out = miss_radiance(L, dir);
    const float ground = 6.360f, vy = 6.360f + 0.0002f;
    f3 oc = mk(0.0f, vy, 0.0f);
    float a = dot(dir, dir), bq = 2.0f * dot(oc, dir), c = dot(oc, oc) - ground * ground;
    float disc = bq * bq - 4.0f * a * c;
    bool hits_ground = false;
    if (disc >= 0.0f) {
        float sq = sqrtf(disc);
        float t0 = (-bq - sq) / (2.0f * a), t1 = (-bq + sq) / (2.0f * a);
        hits_ground = t0 > 0.0f || t1 > 0.0f;
    }
    if (!hits_ground && L.sun_on) {
        float vdl = dot(dir, L.sun_dir);
        if (vdl > L.cos_half) {
            float solid = 2.0f * kPi * (1.0f - L.cos_half);
            float soft = saturate(2.0f * (vdl - L.cos_half) / (1.0f - L.cos_half));
            f3 disk = (L.transmittance * mk(L.outer.x / solid, L.outer.y / solid, L.outer.z / solid)) * soft;
            out = out + disk;
        }
    }
    return out;
}

template <int STACK, int BLOCK, bool FROZEN, bool MAT>
__global__ __launch_bounds__(BLOCK) void resolve_kernel(DevScene S, DevLayout Ly, const void *cache, DevCamera cam,
                                                        float4 *out, uint32_t width, uint32_t height) {
    __shared__ uint32_t stack[STACK * BLOCK];
    uint32_t *col = stack + threadIdx.x;
    uint64_t k = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (k >= (uint64_t)width * height) return;
    uint32_t x = (uint32_t)(k % width), y = (uint32_t)(k / width);
    float px = ((float)x + 0.5f) / (float)width;
    float py = ((float)y + 0.5f) / (float)height;
    px = px + -cam.jx;
    py = py + cam.jy;
    float nx = 2.0f * px + -1.0f, ny = -2.0f * py + 1.0f;
    f3 d = normalize(cam.u * nx + cam.v * ny + cam.w);
    float inv_cos = 1.0f / dot(normalize(cam.w), d);
    Ray r;
    r.o = cam.pos;
    r.d = d;
    r.tmin = cam.near * inv_cos;
    r.tmax = cam.far * inv_cos;
    ray_setup(r);
    Hit h{r.tmax, 0.0f, 0.0f, kInvalid};
    TraversalStats st{0, 0};
    f3 color;
    if (!traverse<STACK, BLOCK, false, false, MAT>(S, r, h, col, st)) {
        color = sky_with_sun_disk(S.light, d);
    } else {
        uint2 ids = S.tri_ids[h.id];
        f3 fn = world_flat_normal(S, h.id);
        bool back = dot(fn, -d) <= 0.0f;
        uint4 ii = Ly.instances[ids.x];
        uint2 side = Ly.sides[ids.x];
        if (back && !(side.y & 1u)) {
            color = mk(0.0f, 0.0f, 0.0f);
        } else {
            uint2 face = Ly.faces[ii.x + ids.y];
            f3 bc = mk(1.0f - h.b1 - h.b2, h.b1, h.b2);
            f3 E = sample_irradiance<FROZEN>(cache, bc, ii.z, face, side, back);
            Shading sd;
            shading_data<MAT>(S, h.id, h.b1, h.b2, -d, sd);
            color = sd.emissive + (sd.diffuse * kInvPi) * E;
        }
    }
    out[k] = make_float4(color.x, color.y, color.z, 1.0f);
}

__global__ void accumulate_kernel(const float4 *in, float4 *accum, float4 *out, uint64_t n, int reset) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 v = reset ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : accum[i];
    float4 x = in[i];
    v.x += x.x;
    v.y += x.y;
    v.z += x.z;
    v.w += 1.0f;
    accum[i] = v;
    out[i] = make_float4(v.x / v.w, v.y / v.w, v.z / v.w, 1.0f);
}

// LpmToneMap + LinearToRec709 (tone_mapper.slang:8-81)
__device__ f3 lpm_tonemap(f3 color, float hdr_max, float exposure, float contrast) {
    const f3 luma709 = mk(0.2126f, 0.7152f, 0.0722f);
    const f3 crosstalk = mk(1.0f, 0.5f, 0.0625f);
    float mid_in = hdr_max * 0.18f * exp2f(-exposure);
    float mid_out = 0.18f;
    float cs = contrast;
    float z0 = -powf(mid_in, contrast);
    float z1 = powf(hdr_max, cs) * powf(mid_in, contrast);
    float z2 = powf(hdr_max, contrast) * powf(mid_in, cs) * mid_out;
    float z3 = powf(hdr_max, cs) * mid_out;
    float z4 = powf(mid_in, cs) * mid_out;
    float bias_x = -((z0 + (mid_out * (z1 - z2)) / (z3 - z4)) / z4);
    float w0 = powf(hdr_max, cs) * powf(mid_in, contrast);
    float w1 = powf(hdr_max, contrast) * powf(mid_in, cs) * mid_out;
    float w2 = powf(hdr_max, cs) * mid_out;
    float w3 = powf(mid_in, cs) * mid_out;
    float bias_y = (w0 - w1) / (w2 - w3);
    float maxc = fmaxf(fmaxf(color.x, color.y), color.z);
    if (maxc <= 0.0f) return mk(0.0f, 0.0f, 0.0f);
    float rcp = 1.0f / maxc;
    f3 ratio = color * rcp;
    float luma = dot(color, luma709);
    luma = powf(fmaxf(luma, 0.0f), contrast);
    luma = luma / (luma * bias_x + bias_y);
    float luma_ratio = dot(ratio, luma709);
    float ratio_scale = saturate(luma / fmaxf(luma_ratio, 0.00001f));
    f3 o = mk(saturate(ratio.x * ratio_scale), saturate(ratio.y * ratio_scale), saturate(ratio.z * ratio_scale));
    f3 cap = crosstalk * mk(1.0f - o.x, 1.0f - o.y, 1.0f - o.z);
    float luma_add = saturate(luma - dot(o, luma709));
    float cap_luma = dot(cap, luma709);
    if (cap_luma > 0.00001f) {
        float t = luma_add / cap_luma;
        o = mk(saturate(o.x + t * cap.x), saturate(o.y + t * cap.y), saturate(o.z + t * cap.z));
    }
    luma_add = saturate(luma - dot(o, luma709));
    o = mk(saturate(o.x + luma_add * (1.0f / luma709.x)), saturate(o.y + luma_add * (1.0f / luma709.y)),
           saturate(o.z + luma_add * (1.0f / luma709.z)));
    return o;
}

__device__ __forceinline__ float rec709(float x) { return x < 0.018f ? 4.5f * x : 1.099f * powf(x, 0.45f) - 0.099f; }

__global__ void tonemap_kernel(const float4 *in, float4 *out, uint64_t n, float exposure) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 x = in[i];
    f3 t = lpm_tonemap(mk(x.x, x.y, x.z), 12.0f, exposure, 1.0f);
    out[i] = make_float4(rec709(t.x), rec709(t.y), rec709(t.z), 1.0f);
}

// Issue-rate calibration: eight independent FP32 FMA chains per thread, four
// rounds per loop trip (32 v_fma_f32 per trip), no memory traffic but one
// store; with every SIMD full the instruction rate is the VALU issue ceiling
// of plain (non-packed) FP32 code such as the traversal kernels.
constexpr int kValuChains = 8;
constexpr int kValuUnroll = 4;

__global__ __launch_bounds__(256) void valu_calibration_kernel(float *out, uint32_t trips, float a, float b) {
    float v[kValuChains];
#pragma unroll
    for (int k = 0; k < kValuChains; ++k) v[k] = (float)(threadIdx.x + k);
    for (uint32_t t = 0; t < trips; ++t) {
#pragma unroll
        for (int u = 0; u < kValuUnroll; ++u)
#pragma unroll
            for (int k = 0; k < kValuChains; ++k) v[k] = __builtin_fmaf(v[k], a, b);
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < kValuChains; ++k) s += v[k];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
namespace nanx {

// Horizon angle seen from the view point 200 m above the ground
// (scene.slang:474): FP32 up to the arccosine, which is taken in double and
// rounded, so every build of this library (and the oracle) holds one value.
float atmosphere_horizon() {
    const float ground = 6.360f, vy = 6.360f + 0.0002f;
    float h = sqrtf(vy * vy);
    float q = sqrtf(h * h - ground * ground) / h;
    return (float)std::acos((double)std::fmin(std::fmax(q, -1.0f), 1.0f)) - 0.5f * kPi;
}

Light make_light(const float *L, const float4 *sky_lut) {
    Light l;
    l.sun_dir = mk(L[0], L[1], L[2]);
    l.cos_half = L[3];
    l.outer = mk(L[4], L[5], L[6]);
    l.transmittance = mk(L[8], L[9], L[10]);
    l.sun_illum = mk(L[4] * L[8], L[5] * L[9], L[6] * L[10]);
    l.sky = mk(L[12], L[13], L[14]);
    l.sun_on = l.sun_illum.x > 0.0f || l.sun_illum.y > 0.0f || l.sun_illum.z > 0.0f;
    l.sky_lut = sky_lut;
    l.horizon = atmosphere_horizon();
    return l;
}

}  // namespace nanx

namespace {



template <int STACK>
void launch_megakernel(const NanTracer *t, float4 *payload, uint32_t iteration, uint32_t spp, uint32_t bounces,
                       uint32_t reset, unsigned long long *counters, hipStream_t stream) {
    DevLayout L = dev_layout(t);
    uint32_t grid = grid_for(t->total_front, kTraceBlock);
    const bool mat = t->scene->dev.mat_features != 0u;
#define NAN_MEGAKERNEL(C, M)                                                                                \
    hipLaunchKernelGGL((tst_megakernel<STACK, kTraceBlock, C, M>), dim3(grid), dim3(kTraceBlock), 0, stream, \
                       t->scene->dev, L, payload, iteration, spp, bounces, reset, counters)
    if (counters) {
        if (mat) NAN_MEGAKERNEL(true, true);
        else NAN_MEGAKERNEL(true, false);
    } else {
        if (mat) NAN_MEGAKERNEL(false, true);
        else NAN_MEGAKERNEL(false, false);
    }
#undef NAN_MEGAKERNEL
}

template <int STACK>
void launch_trace_rays(const NanScene *s, const float4 *rays, uint64_t n, float4 *hits, uint2 *ids, uint32_t *occ,
                       hipStream_t stream) {
    if (s->dev.mat_features)
        hipLaunchKernelGGL((trace_rays_kernel<STACK, kTraceBlock, true>), dim3(grid_for(n, kTraceBlock)),
                           dim3(kTraceBlock), 0, stream, s->dev, rays, n, hits, ids, occ);
    else
        hipLaunchKernelGGL((trace_rays_kernel<STACK, kTraceBlock, false>), dim3(grid_for(n, kTraceBlock)),
                           dim3(kTraceBlock), 0, stream, s->dev, rays, n, hits, ids, occ);
}

template <int STACK>
void launch_resolve(const NanTracer *t, const void *cache, int frozen, const DevCamera &cam, float4 *out,
                    uint32_t w, uint32_t h, hipStream_t stream) {
    DevLayout L = dev_layout(t);
    uint64_t n = (uint64_t)w * h;
    const bool mat = t->scene->dev.mat_features != 0u;
#define NAN_RESOLVE(F, M)                                                                                 \
    hipLaunchKernelGGL((resolve_kernel<STACK, kTraceBlock, F, M>), dim3(grid_for(n, kTraceBlock)),         \
                       dim3(kTraceBlock), 0, stream, t->scene->dev, L, cache, cam, out, w, h)
    if (frozen) {
        if (mat) NAN_RESOLVE(true, true);
        else NAN_RESOLVE(true, false);
    } else {
        if (mat) NAN_RESOLVE(false, true);
        else NAN_RESOLVE(false, false);
    }
#undef NAN_RESOLVE
}


// Kernel stack depths compiled in; a scene whose worst-case traversal needs
// more than the largest is rejected at creation.
constexpr uint32_t kStackDepths[] = {32, 64};

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

NAN_API int nan_version(void) { return NAN_HIP_ABI_VERSION; }

NAN_API const char *nan_last_error(void) { return g_last_error.c_str(); }

// Texel decode table of every texture: c / 255, then the inverse sRGB
// transfer function (IEC 61966-2-1) of c / 255, each in double, rounded once.
static void make_texel_lut(float *lut) {
    for (int c = 0; c < 256; ++c) {
        double x = c / 255.0;
        lut[c] = (float)x;
        lut[256 + c] = (float)(x <= 0.04045 ? x / 12.92 : std::pow((x + 0.055) / 1.055, 2.4));
    }
}

// 1 when some material is textured or alpha tested (DevScene::mat_features)
static uint32_t material_features(const uint32_t *materials, size_t count) {
    for (size_t m = 0; m < count; ++m) {
        const uint32_t *w = materials + kMaterialWords * m;
        if (w[7] != 0u || w[8] != 0xFFFFFFFFu || w[9] != 0xFFFFFFFFu || w[10] != 0xFFFFFFFFu) return 1u;
    }
    return 0u;
}

static int check_materials_and_textures(const NanSceneDesc *d, char *err, size_t cap) {
    if (d->texture_count && (!d->textures || !d->texels))
        return fail(err, cap, "nan_scene_create: null texture buffer");
    for (uint32_t k = 0; k < d->texture_count; ++k) {
        const uint32_t *t = d->textures + 4 * (size_t)k;
        if (t[1] == 0 || t[2] == 0 || t[1] > (1u << 24) || t[2] > (1u << 24) ||
            (uint64_t)t[0] + (uint64_t)t[1] * t[2] > d->texel_count)
            return fail(err, cap, "nan_scene_create: texture %u (%ux%u at texel %u) out of range", k, t[1], t[2], t[0]);
    }
    for (uint32_t m = 0; m < d->material_count; ++m) {
        const uint32_t *w = d->materials + kMaterialWords * (size_t)m;
        if (w[7] > 2u) return fail(err, cap, "nan_scene_create: material %u has unknown alpha_mode %u", m, w[7]);
        for (int j = 8; j < 11; ++j)
            if (w[j] != 0xFFFFFFFFu && w[j] >= d->texture_count)
                return fail(err, cap, "nan_scene_create: material %u references missing texture %u", m, w[j]);
    }
    return 0;
}

// Textures, the texel decode table and the per-triangle uv records on the
// device (nothing without textures); the frame index of the alpha test always.
static int upload_textures(NanScene *s, const uint32_t *materials, size_t material_count,
                           const std::vector<uint32_t> &textures, const std::vector<uint32_t> &texels,
                           const std::vector<float> &tri_uv, char *err, size_t cap) {
    auto up = [&](auto **dst, const void *src, size_t bytes) -> int {
        int rc = upload(dst, src, bytes, err, cap);
        if (rc == 0) s->allocations.push_back((void *)*dst);
        return rc;
    };
    const uint32_t zero = 0;
    uint32_t *frame;
    if (up(&frame, &zero, 4)) return -1;
    s->dev.frame_index = frame;
    s->dev.mat_features = material_features(materials, material_count);
    if (textures.empty()) return 0;
    float lut[512];
    make_texel_lut(lut);
    uint4 *dtex;
    uint32_t *dtexels;
    float *dlut;
    float4 *duv;
    int rc = 0;
    rc |= up(&dtex, textures.data(), textures.size() * 4);
    rc |= up(&dtexels, texels.data(), texels.size() * 4);
    rc |= up(&dlut, lut, sizeof(lut));
    rc |= up(&duv, tri_uv.data(), tri_uv.size() * 4);
    if (rc) return -1;
    s->dev.textures = dtex;
    s->dev.texels = dtexels;
    s->dev.texel_lut = dlut;
    s->dev.tri_uv = duv;
    return 0;
}

NAN_API int nan_scene_create(const NanSceneDesc *d, const NanBvhOptions *options, NanScene **out, char *err,
                             size_t cap) {
    if (!d || !out) return fail(err, cap, "nan_scene_create: null argument");
    *out = nullptr;
    if (!d->vertices || !d->indices || !d->mesh_descs || !d->instance_descs || !d->transforms ||
        !d->inv_transposes || !d->materials || !d->lighting)
        return fail(err, cap, "nan_scene_create: null scene buffer");
    if (d->instance_count == 0) return fail(err, cap, "nan_scene_create: scene has no instances");
    if (check_materials_and_textures(d, err, cap)) return -1;
    auto t0 = std::chrono::steady_clock::now();
    // validate descriptors and count triangles
    uint64_t tri_total = 0;
    for (uint32_t m = 0; m < d->mesh_count; ++m) {
        const uint32_t *md = d->mesh_descs + 4 * (size_t)m;
        if ((uint64_t)md[2] + md[0] > d->vertex_count || (uint64_t)md[3] + md[1] > d->index_count || md[1] % 3)
            return fail(err, cap, "nan_scene_create: mesh %u descriptor out of range", m);
    }
    for (uint32_t i = 0; i < d->instance_count; ++i) {
        const uint32_t *id = d->instance_descs + 4 * (size_t)i;
        if (id[0] >= d->mesh_count || id[1] >= d->material_count || id[2] >= d->transform_count)
            return fail(err, cap, "nan_scene_create: instance %u references a missing mesh/material/transform", i);
        tri_total += d->mesh_descs[4 * (size_t)id[0] + 1] / 3;
    }
    for (uint32_t m = 0; m < d->mesh_count; ++m) {
        const uint32_t *md = d->mesh_descs + 4 * (size_t)m;
        for (uint32_t k = 0; k < md[1]; ++k)
            if (d->indices[md[3] + k] >= md[0])
                return fail(err, cap, "nan_scene_create: mesh %u index out of range", m);
    }
    if (tri_total == 0) return fail(err, cap, "nan_scene_create: scene has no triangles");
    if (tri_total > nanbvh::kLeafFirstMask) return fail(err, cap, "nan_scene_create: too many triangles");
    // world-space triangles: M (x, y, z, 1) in float64, rounded once
    std::vector<float> world((size_t)tri_total * 9);
    std::vector<uint32_t> tri_ids((size_t)tri_total * 2);
    uint64_t g = 0;
    for (uint32_t i = 0; i < d->instance_count; ++i) {
        const uint32_t *id = d->instance_descs + 4 * (size_t)i;
        const uint32_t *md = d->mesh_descs + 4 * (size_t)id[0];
        const float *m = d->transforms + 16 * (size_t)id[2];
        uint32_t ntri = md[1] / 3;
        for (uint32_t p = 0; p < ntri; ++p, ++g) {
            tri_ids[2 * g] = i;
            tri_ids[2 * g + 1] = p;
            for (int k = 0; k < 3; ++k) {
                const float *v = d->vertices + 8 * (size_t)(md[2] + d->indices[md[3] + 3 * p + k]);
                for (int r = 0; r < 3; ++r) {
                    double w = (double)m[4 * r + 0] * v[0] + (double)m[4 * r + 1] * v[1];
                    w = w + (double)m[4 * r + 2] * v[2];
                    w = w + (double)m[4 * r + 3];
                    world[9 * g + 3 * k + r] = (float)w;
                }
            }
        }
    }
    // per-triangle shading records (load_triangle in nan_device.h)
    std::vector<float> shade_rec((size_t)tri_total * 20);
    g = 0;
    for (uint32_t i = 0; i < d->instance_count; ++i) {
        const uint32_t *id = d->instance_descs + 4 * (size_t)i;
        const uint32_t *md = d->mesh_descs + 4 * (size_t)id[0];
        uint32_t ntri = md[1] / 3;
        uint32_t matw = id[1] | ((id[3] & 1u) << 31);
        for (uint32_t p = 0; p < ntri; ++p, ++g) {
            const float *v[3];
            for (int k = 0; k < 3; ++k) v[k] = d->vertices + 8 * (size_t)(md[2] + d->indices[md[3] + 3 * p + k]);
            float *r = &shade_rec[20 * g];
            for (int k = 0; k < 3; ++k) {
                r[4 * k + 0] = v[k][0];
                r[4 * k + 1] = v[k][1];
                r[4 * k + 2] = v[k][2];
                r[4 * k + 3] = v[k][3];
            }
            r[12] = v[0][4];
            r[13] = v[0][5];
            r[14] = v[1][4];
            r[15] = v[1][5];
            r[16] = v[2][4];
            r[17] = v[2][5];
            std::memcpy(&r[18], &id[2], 4);
            std::memcpy(&r[19], &matw, 4);
        }
    }
    nanbvh::BuildOptions bo;
    if (options) {
        if (options->sah_bins) bo.bins = options->sah_bins;
        if (options->max_leaf_size) bo.max_leaf = options->max_leaf_size;
        bo.threads = options->build_threads;
    }
    nanbvh::Bvh4 bvh;
    const char *berr = nullptr;
    if (!nanbvh::build_bvh4(world.data(), (uint32_t)tri_total, bo, &bvh, &berr))
        return fail(err, cap, "nan_scene_create: %s", berr ? berr : "BVH build failed");
    uint32_t depth = 0;
    for (uint32_t s : kStackDepths)
        if (bvh.max_stack <= s) { depth = s; break; }
    if (!depth) return fail(err, cap, "nan_scene_create: BVH needs a %u-entry stack (max 64)", bvh.max_stack);
    // triangles of non-opaque materials carry the alpha-test flag in their record
    for (size_t r = 0; r < bvh.tris.size(); r += nanbvh::kTriWords) {
        uint32_t id;
        std::memcpy(&id, &bvh.tris[r + 9], 4);
        const uint32_t *inst = d->instance_descs + 4 * (size_t)tri_ids[2 * (size_t)id];
        uint32_t flag = d->materials[kMaterialWords * (size_t)inst[1] + 7] != 0u ? 1u : 0u;
        std::memcpy(&bvh.tris[r + 10], &flag, 4);
    }
    // per-triangle uv records (with textures): uv0 uv1 | uv2 0 0
    std::vector<float> tri_uv;
    if (d->texture_count) {
        tri_uv.assign((size_t)tri_total * 8, 0.0f);
        g = 0;
        for (uint32_t i = 0; i < d->instance_count; ++i) {
            const uint32_t *id = d->instance_descs + 4 * (size_t)i;
            const uint32_t *md = d->mesh_descs + 4 * (size_t)id[0];
            uint32_t ntri = md[1] / 3;
            for (uint32_t p = 0; p < ntri; ++p, ++g)
                for (int k = 0; k < 3; ++k) {
                    const float *v = d->vertices + 8 * (size_t)(md[2] + d->indices[md[3] + 3 * p + k]);
                    tri_uv[8 * g + 2 * k] = v[6];
                    tri_uv[8 * g + 2 * k + 1] = v[7];
                }
        }
    }
    auto t1 = std::chrono::steady_clock::now();

    NanScene *s = new NanScene();
    auto up = [&](auto **dst, const void *src, size_t bytes) -> int {
        int rc = upload(dst, src, bytes, err, cap);
        if (rc == 0) s->allocations.push_back((void *)*dst);
        return rc;
    };
    // vertices: 32-byte stride copied as-is (pos.xyz n.x | n.yz uv)
    uint4 *nodes;
    float4 *tris, *verts, *xf, *itf, *mats, *shrec;
    uint2 *ids;
    uint32_t *indices;
    uint4 *mesh_descs, *inst_descs;
    size_t bytes_total = 0;
    int rc = 0;
    rc |= up(&nodes, bvh.nodes.data(), bvh.nodes.size() * 4);
    rc |= up(&tris, bvh.tris.data(), bvh.tris.size() * 4);
    rc |= up(&ids, tri_ids.data(), tri_ids.size() * 4);
    rc |= up(&verts, d->vertices, (size_t)d->vertex_count * 32);
    rc |= up(&indices, d->indices, (size_t)d->index_count * 4);
    rc |= up(&mesh_descs, d->mesh_descs, (size_t)d->mesh_count * 16);
    rc |= up(&inst_descs, d->instance_descs, (size_t)d->instance_count * 16);
    rc |= up(&xf, d->transforms, (size_t)d->transform_count * 64);
    rc |= up(&itf, d->inv_transposes, (size_t)d->transform_count * 64);
    rc |= up(&mats, d->materials, (size_t)d->material_count * kMaterialWords * 4);
    rc |= up(&shrec, shade_rec.data(), shade_rec.size() * 4);
    std::vector<uint32_t> h_textures(d->textures, d->textures + (size_t)d->texture_count * 4);
    std::vector<uint32_t> h_texels;
    if (d->texture_count) h_texels.assign(d->texels, d->texels + d->texel_count);
    rc |= upload_textures(s, d->materials, d->material_count, h_textures, h_texels, tri_uv, err, cap);
    if (rc) {
        delete s;
        return -1;
    }
    bytes_total = bvh.nodes.size() * 4 + bvh.tris.size() * 4 + tri_ids.size() * 4 + (size_t)d->vertex_count * 32 +
                  (size_t)d->index_count * 4 + (size_t)d->mesh_count * 16 + (size_t)d->instance_count * 16 +
                  (size_t)d->transform_count * 128 + (size_t)d->material_count * kMaterialWords * 4 +
                  shade_rec.size() * 4 + h_textures.size() * 4 + h_texels.size() * 4 + tri_uv.size() * 4;
    s->dev.nodes = nodes;
    s->dev.tris = tris;
    s->dev.tri_ids = ids;
    s->dev.vertices = verts;
    s->dev.indices = indices;
    s->dev.mesh_descs = mesh_descs;
    s->dev.instance_descs = inst_descs;
    s->dev.transforms = xf;
    s->dev.inv_transposes = itf;
    s->dev.materials = mats;
    s->dev.tri_shading = shrec;
    s->h_nodes = std::move(bvh.nodes);
    s->h_tris = std::move(bvh.tris);
    s->h_shade = std::move(shade_rec);
    s->h_xf.assign(d->transforms, d->transforms + (size_t)d->transform_count * 16);
    s->h_itf.assign(d->inv_transposes, d->inv_transposes + (size_t)d->transform_count * 16);
    s->h_mats.assign(d->materials, d->materials + (size_t)d->material_count * kMaterialWords);
    s->h_tri_uv = std::move(tri_uv);
    s->h_textures = std::move(h_textures);
    s->h_texels = std::move(h_texels);
    s->dev.root = bvh.root_entry;
    {
        // emissive triangles (material emission > 0), areas in float64
        std::vector<float> em;
        std::vector<double> areas;
        for (uint64_t t = 0; t < tri_total; ++t) {
            const uint32_t *id = d->instance_descs + 4 * (size_t)tri_ids[2 * t];
            const uint32_t *mw = d->materials + kMaterialWords * (size_t)id[1];
            float le[3];
            std::memcpy(le, mw + 4, 12);
            if (!(le[0] > 0.0f || le[1] > 0.0f || le[2] > 0.0f)) continue;
            const float *w = world.data() + 9 * t;
            double e1[3], e2[3], c[3];
            for (int r = 0; r < 3; ++r) {
                e1[r] = (double)w[3 + r] - w[r];
                e2[r] = (double)w[6 + r] - w[r];
            }
            c[0] = e1[1] * e2[2] - e1[2] * e2[1];
            c[1] = e1[2] * e2[0] - e1[0] * e2[2];
            c[2] = e1[0] * e2[1] - e1[1] * e2[0];
            double len = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            if (!(len > 0.0)) continue;
            // textured or alpha-tested emitters evaluate the point they sample
            uint32_t tri_bits = (uint32_t)t, varying = (mw[9] != 0xFFFFFFFFu || mw[7] != 0u) ? 1u : 0u;
            float tri_f, varying_f;
            std::memcpy(&tri_f, &tri_bits, 4);
            std::memcpy(&varying_f, &varying, 4);
            float rec[20] = {w[0], w[1], w[2], (float)(0.5 * len),
                             (float)e1[0], (float)e1[1], (float)e1[2], tri_f,
                             (float)e2[0], (float)e2[1], (float)e2[2], 0.0f,
                             (float)(c[0] / len), (float)(c[1] / len), (float)(c[2] / len), 0.0f,
                             le[0], le[1], le[2], varying_f};
            em.insert(em.end(), rec, rec + 20);
            areas.push_back(0.5 * len);
        }
        if (!areas.empty()) {
            double total = 0.0;
            for (double a : areas) total += a;
            std::vector<float> cdf(areas.size());
            double run = 0.0;
            for (size_t k = 0; k < areas.size(); ++k) {
                run += areas[k];
                cdf[k] = (float)(run / total);
            }
            cdf.back() = 1.0f;
            float4 *demit;
            float *dcdf;
            if (up(&demit, em.data(), em.size() * 4) || up(&dcdf, cdf.data(), cdf.size() * 4)) {
                delete s;
                return -1;
            }
            s->dev.emitters = demit;
            s->dev.emitter_cdf = dcdf;
            s->dev.emitter_count = (uint32_t)areas.size();
            s->dev.emitter_area = (float)total;
        }
    }
    float max_coord = 1.0f;
    for (float v : world) max_coord = std::max(max_coord, std::fabs(v));
    s->dev.slack = max_coord * 0x1p-18f;
    std::memcpy(s->lighting, d->lighting, sizeof(s->lighting));
    s->dev.light = make_light(s->lighting, nullptr);
    {
        SceneRec rec{};
        rec.root = s->dev.root;
        rec.slack = s->dev.slack;
        rec.max_bounces = 0;
        rec.roulette = 0;
        rec.light = s->dev.light;
        SceneRec *drec;
        if (up(&drec, &rec, sizeof(rec))) {
            delete s;
            return -1;
        }
        s->dev.scenes = drec;
    }
    s->stack_depth = depth;
    s->instance_triangles.resize(d->instance_count);
    for (uint32_t i = 0; i < d->instance_count; ++i)
        s->instance_triangles[i] = d->mesh_descs[4 * (size_t)d->instance_descs[4 * (size_t)i] + 1] / 3;
    s->info.triangle_count = tri_total;
    s->info.node_count = bvh.node_count;
    s->info.leaf_count = bvh.leaf_count;
    s->info.device_bytes = bytes_total;
    s->info.max_stack = bvh.max_stack;
    s->info.stack_depth = depth;
    s->info.build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    s->info.sah_cost = bvh.sah_cost;
    *out = s;
    return 0;
}

NAN_API void nan_scene_destroy(NanScene *scene) { delete scene; }

NAN_API int nan_scene_info(const NanScene *scene, NanSceneInfo *out) {
    if (!scene || !out) return fail(nullptr, 0, "nan_scene_info: null argument");
    *out = scene->info;
    return 0;
}

// The front texels in [0, total_front) that `ds` (ascending) does not hold:
// the paths of the single-sided instances, which the wavefront traces apart
// from the double-sided ones.
static int upload_single_sided(NanTracer *t, const std::vector<uint32_t> &ds, char *err, size_t cap) {
    std::vector<uint32_t> ss;
    ss.reserve((size_t)t->total_front - ds.size());
    size_t j = 0;
    for (uint64_t w = 0; w < t->total_front; ++w) {
        if (j < ds.size() && ds[j] == w) {
            ++j;
            continue;
        }
        ss.push_back((uint32_t)w);
    }
    if (j != ds.size()) return fail(err, cap, "double-sided texel list is not ascending");
    t->ss_count = (uint32_t)ss.size();
    return upload(&t->ss_list, ss.data(), ss.size() * 4, err, cap);
}

static int refresh_light(NanScene *scene) {
    scene->dev.light = make_light(scene->lighting, scene->sky_lut);
    SceneRec rec{};
    rec.root = scene->dev.root;
    rec.slack = scene->dev.slack;
    rec.roulette = scene->roulette;
    rec.light = scene->dev.light;
    HIP_TRY(hipMemcpy((void *)scene->dev.scenes, &rec, sizeof(rec), hipMemcpyHostToDevice), nullptr, 0);
    return 0;
}

NAN_API int nan_scene_set_lighting(NanScene *scene, const float *lighting16) {
    if (!scene || !lighting16) return fail(nullptr, 0, "nan_scene_set_lighting: null argument");
    std::memcpy(scene->lighting, lighting16, sizeof(scene->lighting));
    return refresh_light(scene);
}

NAN_API int nan_scene_set_russian_roulette(NanScene *scene, uint32_t enable) {
    if (!scene) return fail(nullptr, 0, "nan_scene_set_russian_roulette: null scene");
    if (scene->batch) return fail(nullptr, 0, "nan_scene_set_russian_roulette: set it on each scene of a batch");
    scene->roulette = enable ? 1u : 0u;
    return refresh_light(scene);
}

__global__ void set_frame_index_kernel(uint32_t *dst, uint32_t value) { *dst = value; }

static int set_frame_index(NanScene *scene, uint32_t frame_index, void *stream) {
    hipLaunchKernelGGL(set_frame_index_kernel, dim3(1), dim3(1), 0, as_stream(stream),
                       const_cast<uint32_t *>(scene->dev.frame_index), frame_index);
    HIP_TRY(hipGetLastError(), nullptr, 0);
    return 0;
}

NAN_API int nan_scene_set_frame_index(NanScene *scene, uint32_t frame_index, void *stream) {
    if (!scene) return fail(nullptr, 0, "nan_scene_set_frame_index: null scene");
    if (scene->batch) return fail(nullptr, 0, "nan_scene_set_frame_index: use nan_batch_set_frame_index on a batch");
    return set_frame_index(scene, frame_index, stream);
}

NAN_API int nan_scene_set_sky_view_lut(NanScene *scene, const float *sky_view_lut) {
    if (!scene) return fail(nullptr, 0, "nan_scene_set_sky_view_lut: null scene");
    if (scene->batch) return fail(nullptr, 0, "nan_scene_set_sky_view_lut: set the table on each scene of a batch");
    scene->sky_lut = reinterpret_cast<const float4 *>(sky_view_lut);
    return refresh_light(scene);
}

NAN_API int nan_trace_rays(NanScene *scene, const float *rays, uint64_t n, float *hits, uint32_t *ids,
                           uint32_t *occluded, void *stream) {
    if (!scene || (n && (!rays || !hits || !ids))) return fail(nullptr, 0, "nan_trace_rays: null argument");
    if (n == 0) return 0;
    const float4 *r = reinterpret_cast<const float4 *>(rays);
    if (scene->stack_depth <= 32)
        launch_trace_rays<32>(scene, r, n, (float4 *)hits, (uint2 *)ids, occluded, as_stream(stream));
    else
        launch_trace_rays<64>(scene, r, n, (float4 *)hits, (uint2 *)ids, occluded, as_stream(stream));
    return check_launch("trace_rays");
}

NAN_API int nan_tracer_create(NanScene *scene, const uint32_t *face_table, uint32_t face_count,
                              const uint32_t *instance_table, const uint32_t *side_table, uint32_t instance_count,
                              uint64_t total_surface_texels, uint64_t total_payload_count, NanTracer **out,
                              char *err, size_t cap) {
    if (!scene || !face_table || !instance_table || !side_table || !out)
        return fail(err, cap, "nan_tracer_create: null argument");
    *out = nullptr;
    if (instance_count != scene->instance_triangles.size())
        return fail(err, cap, "nan_tracer_create: layout has %u instances, scene has %zu", instance_count,
                    scene->instance_triangles.size());
    if (total_payload_count >= (1ull << 32)) return fail(err, cap, "nan_tracer_create: payload exceeds 2^32 slots");
    // validate the tables against each other: faces tile each instance's range
    std::vector<uint32_t> texel_face;
    texel_face.reserve(total_surface_texels);
    std::vector<uint32_t> face_inst(face_count);
    std::vector<uint32_t> prefix(instance_count);
    uint64_t front = 0, payload_end = 0;
    uint64_t faces_before = 0;
    for (uint32_t i = 0; i < instance_count; ++i) {
        const uint32_t *ii = instance_table + 4 * (size_t)i;
        const uint32_t *sd = side_table + 2 * (size_t)i;
        if ((uint64_t)ii[0] + ii[1] > face_count)
            return fail(err, cap, "nan_tracer_create: instance %u face range out of bounds", i);
        if (ii[1] != scene->instance_triangles[i])
            return fail(err, cap, "nan_tracer_create: instance %u face_count differs from its mesh", i);
        if (ii[0] != faces_before)
            return fail(err, cap, "nan_tracer_create: instance %u faces are not in instance order", i);
        faces_before += ii[1];
        prefix[i] = (uint32_t)front;
        uint64_t local = 0;
        for (uint32_t f = ii[0]; f < ii[0] + ii[1]; ++f) {
            const uint32_t *fi = face_table + 2 * (size_t)f;
            if (fi[1] == 0 || fi[1] > 4096) return fail(err, cap, "nan_tracer_create: face %u bad resolution", f);
            if (fi[0] != local) return fail(err, cap, "nan_tracer_create: face %u address is not contiguous", f);
            uint64_t pc = (uint64_t)(fi[1] + 1) * (fi[1] + 2) / 2;
            face_inst[f] = i;
            for (uint64_t k = 0; k < pc; ++k) texel_face.push_back(f);
            local += pc;
        }
        if (local != ii[3]) return fail(err, cap, "nan_tracer_create: instance %u texel_count mismatch", i);
        if ((uint64_t)ii[2] + ii[3] > total_payload_count)
            return fail(err, cap, "nan_tracer_create: instance %u slots out of bounds", i);
        if (sd[1] & 1u) {
            if ((uint64_t)sd[0] + ii[3] > total_payload_count)
                return fail(err, cap, "nan_tracer_create: instance %u back slots out of bounds", i);
            payload_end = std::max<uint64_t>(payload_end, (uint64_t)sd[0] + ii[3]);
        }
        payload_end = std::max<uint64_t>(payload_end, (uint64_t)ii[2] + ii[3]);
        front += ii[3];
    }
    if (front != total_surface_texels) return fail(err, cap, "nan_tracer_create: total_surface_texels mismatch");
    if (front >= (1ull << 32)) return fail(err, cap, "nan_tracer_create: too many texels");
    // every instance must exist in the scene
    NanTracer *t = new NanTracer();
    t->scene = scene;
    t->face_count = face_count;
    t->instance_count = instance_count;
    t->total_front = front;
    t->total_payload = total_payload_count;
    int rc = 0;
    rc |= upload(&t->faces, face_table, (size_t)face_count * 8, err, cap);
    rc |= upload(&t->instances, instance_table, (size_t)instance_count * 16, err, cap);
    rc |= upload(&t->sides, side_table, (size_t)instance_count * 8, err, cap);
    rc |= upload(&t->texel_face, texel_face.data(), texel_face.size() * 4, err, cap);
    rc |= upload(&t->face_inst, face_inst.data(), face_inst.size() * 4, err, cap);
    rc |= upload(&t->front_prefix, prefix.data(), prefix.size() * 4, err, cap);
    std::vector<uint32_t> ds;
    for (uint32_t i = 0; i < instance_count; ++i)
        if (side_table[2 * (size_t)i + 1] & 1u)
            for (uint32_t k = 0; k < instance_table[4 * (size_t)i + 3]; ++k) ds.push_back(prefix[i] + k);
    t->ds_count = (uint32_t)ds.size();
    rc |= upload(&t->ds_list, ds.data(), ds.size() * 4, err, cap);
    rc |= upload_single_sided(t, ds, err, cap);
    t->h_faces.assign(face_table, face_table + (size_t)face_count * 2);
    t->h_instances.assign(instance_table, instance_table + (size_t)instance_count * 4);
    t->h_sides.assign(side_table, side_table + (size_t)instance_count * 2);
    t->h_texel_face = std::move(texel_face);
    t->h_face_inst = std::move(face_inst);
    t->h_prefix = std::move(prefix);
    t->h_ds = std::move(ds);
    if (!rc && wavefront_prepare(t)) rc = fail(err, cap, "nan_tracer_create: %s", g_last_error.c_str());
    if (rc) {
        delete t;
        return -1;
    }
    *out = t;
    return 0;
}

NAN_API void nan_tracer_destroy(NanTracer *tracer) { delete tracer; }

NAN_API int nan_tracer_tables(const NanTracer *tracer, const void **face_table, const void **instance_table,
                              const void **side_table) {
    if (!tracer) return fail(nullptr, 0, "nan_tracer_tables: null tracer");
    if (face_table) *face_table = tracer->faces;
    if (instance_table) *instance_table = tracer->instances;
    if (side_table) *side_table = tracer->sides;
    return 0;
}

static int tst_execute(NanTracer *t, void *payload, uint32_t iteration, uint32_t spp, uint32_t bounces,
                       uint32_t reset, uint32_t rng_mode, uint64_t *counters, void *stream) {
    if (!t || (!payload && !t->scene->batch)) return fail(nullptr, 0, "nan_tst_execute: null argument");
    if (rng_mode != NAN_RNG_REFERENCE && rng_mode != NAN_RNG_PCG)
        return fail(nullptr, 0, "nan_tst_execute: unknown rng_mode %u", rng_mode);
    t->rng_mode = rng_mode;
    if (t->total_front == 0) return 0;
    if (t->scene->batch && t->mode != NAN_TRACER_WAVEFRONT)
        return fail(nullptr, 0, "nan_tst_execute: scene batches run in wavefront mode only");
    unsigned long long *c = reinterpret_cast<unsigned long long *>(counters);
    if (t->mode == NAN_TRACER_WAVEFRONT)
        return wavefront_execute(t, (float4 *)payload, iteration, spp, bounces, reset, c, as_stream(stream));
    StageTimer tm(t, as_stream(stream), kStageMegakernel);
    if (t->scene->stack_depth <= 32)
        launch_megakernel<32>(t, (float4 *)payload, iteration, spp, bounces, reset, c, as_stream(stream));
    else
        launch_megakernel<64>(t, (float4 *)payload, iteration, spp, bounces, reset, c, as_stream(stream));
    return check_launch("tst_megakernel");
}

NAN_API int nan_tracer_profile(NanTracer *tracer, int enable) {
    if (!tracer) return fail(nullptr, 0, "nan_tracer_profile: null tracer");
    tracer->profile = enable != 0;
    tracer->stage_events.clear();
    tracer->events_used = 0;
    return 0;
}

NAN_API int nan_tracer_stage_times(NanTracer *tracer, double *ms, uint32_t *launches) {
    if (!tracer || !ms) return fail(nullptr, 0, "nan_tracer_stage_times: null argument");
    for (int k = 0; k < kStageCount; ++k) {
        ms[k] = 0.0;
        if (launches) launches[k] = 0;
    }
    for (auto &se : tracer->stage_events) {
        HIP_TRY(hipEventSynchronize(se.second.second), nullptr, 0);
        float dt = 0.0f;
        HIP_TRY(hipEventElapsedTime(&dt, se.second.first, se.second.second), nullptr, 0);
        ms[se.first] += dt;
        if (launches) launches[se.first]++;
    }
    tracer->stage_events.clear();
    tracer->events_used = 0;
    return 0;
}

NAN_API int nan_tracer_set_mode(NanTracer *tracer, int mode) {
    if (!tracer) return fail(nullptr, 0, "nan_tracer_set_mode: null tracer");
    if (mode != NAN_TRACER_MEGAKERNEL && mode != NAN_TRACER_WAVEFRONT)
        return fail(nullptr, 0, "nan_tracer_set_mode: unknown mode %d", mode);
    if (mode == NAN_TRACER_MEGAKERNEL && tracer->scene->stack_depth == 0)
        return fail(nullptr, 0, "nan_tracer_set_mode: BVH too deep for the megakernel");
    tracer->mode = mode;
    return 0;
}

NAN_API int nan_tst_execute(NanTracer *tracer, void *payload, uint32_t iteration, uint32_t samples_per_texel,
                            uint32_t max_bounces, uint32_t reset, uint32_t rng_mode, void *stream) {
    return tst_execute(tracer, payload, iteration, samples_per_texel, max_bounces, reset, rng_mode, nullptr,
                       stream);
}

NAN_API int nan_tst_execute_counted(NanTracer *tracer, void *payload, uint32_t iteration, uint32_t samples_per_texel,
                                    uint32_t max_bounces, uint32_t reset, uint32_t rng_mode, uint64_t *counters,
                                    void *stream) {
    if (!counters) return fail(nullptr, 0, "nan_tst_execute_counted: null counters");
    return tst_execute(tracer, payload, iteration, samples_per_texel, max_bounces, reset, rng_mode, counters,
                       stream);
}

NAN_API int nan_tst_raygen(NanTracer *tracer, uint32_t iteration, uint32_t spp, uint32_t *out, uint32_t *rng_state,
                           void *stream) {
    if (!tracer || !out) return fail(nullptr, 0, "nan_tst_raygen: null argument");
    if (tracer->total_front == 0) return 0;
    DevLayout Ly = dev_layout(tracer);
    Ly.rng_pcg = 0u;  // the reference's stream
    hipLaunchKernelGGL(tst_raygen_kernel, dim3(grid_for(tracer->total_front, kFlatBlock)), dim3(kFlatBlock), 0,
                       as_stream(stream), Ly, iteration, spp, (uint4 *)out, rng_state);
    return check_launch("tst_raygen");
}

NAN_API int nan_rgb9e5_pack(const void *payload, uint32_t *packed, uint64_t n, void *stream) {
    if (n == 0) return 0;
    if (!payload || !packed) return fail(nullptr, 0, "nan_rgb9e5_pack: null argument");
    hipLaunchKernelGGL(rgb9e5_pack_kernel, dim3(grid_for(n, kFlatBlock)), dim3(kFlatBlock), 0, as_stream(stream),
                       (const float4 *)payload, packed, n);
    return check_launch("rgb9e5_pack");
}

NAN_API int nan_rgb9e5_unpack(const uint32_t *packed, float *out, uint64_t n, void *stream) {
    if (n == 0) return 0;
    if (!packed || !out) return fail(nullptr, 0, "nan_rgb9e5_unpack: null argument");
    hipLaunchKernelGGL(rgb9e5_unpack_kernel, dim3(grid_for(n, kFlatBlock)), dim3(kFlatBlock), 0, as_stream(stream),
                       packed, (float4 *)out, n);
    return check_launch("rgb9e5_unpack");
}

NAN_API int nan_resolve(NanTracer *tracer, const void *cache, int frozen, const float *c, float *out,
                        uint32_t width, uint32_t height, void *stream) {
    if (!tracer || !cache || !c || !out) return fail(nullptr, 0, "nan_resolve: null argument");
    if (width == 0 || height == 0) return 0;
    DevCamera cam;
    cam.pos = mk(c[0], c[1], c[2]);
    cam.jx = c[3];
    cam.u = mk(c[4], c[5], c[6]);
    cam.jy = c[7];
    cam.v = mk(c[8], c[9], c[10]);
    cam.near = c[11];
    cam.w = mk(c[12], c[13], c[14]);
    cam.far = c[15];
    if (tracer->scene->stack_depth <= 32)
        launch_resolve<32>(tracer, cache, frozen, cam, (float4 *)out, width, height, as_stream(stream));
    else
        launch_resolve<64>(tracer, cache, frozen, cam, (float4 *)out, width, height, as_stream(stream));
    return check_launch("resolve");
}

NAN_API int nan_accumulate(const float *input, float *accum, float *output, uint64_t n, int reset, void *stream) {
    if (n == 0) return 0;
    if (!input || !accum || !output) return fail(nullptr, 0, "nan_accumulate: null argument");
    hipLaunchKernelGGL(accumulate_kernel, dim3(grid_for(n, kFlatBlock)), dim3(kFlatBlock), 0, as_stream(stream),
                       (const float4 *)input, (float4 *)accum, (float4 *)output, n, reset);
    return check_launch("accumulate");
}

NAN_API int nan_tonemap(const float *input, float *output, uint64_t n, float exposure, void *stream) {
    if (n == 0) return 0;
    if (!input || !output) return fail(nullptr, 0, "nan_tonemap: null argument");
    hipLaunchKernelGGL(tonemap_kernel, dim3(grid_for(n, kFlatBlock)), dim3(kFlatBlock), 0, as_stream(stream),
                       (const float4 *)input, (float4 *)output, n, exposure);
    return check_launch("tonemap");
}


NAN_API int nan_calibrate_valu(uint32_t trips, double *ms_out, double *fma_wave_insts_out) {
    if (!ms_out || !fma_wave_insts_out || trips == 0) return fail(nullptr, 0, "nan_calibrate_valu: bad argument");
    int dev = 0, cus = 0, per_cu = 0;
    HIP_TRY(hipGetDevice(&dev), nullptr, 0);
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), nullptr, 0);
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, valu_calibration_kernel, 256, 0), nullptr, 0);
    uint32_t blocks = (uint32_t)(cus * std::max(per_cu, 1));
    float *out = nullptr;
    HIP_TRY(hipMalloc((void **)&out, (size_t)blocks * 256 * sizeof(float)), nullptr, 0);
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(valu_calibration_kernel, dim3(blocks), dim3(256), 0, nullptr, out, trips, 0.999f, 1e-3f);
        hipEventRecord(a, nullptr);
        hipLaunchKernelGGL(valu_calibration_kernel, dim3(blocks), dim3(256), 0, nullptr, out, trips, 0.999f, 1e-3f);
        hipEventRecord(b, nullptr);
        e = hipEventSynchronize(b);
    }
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    hipFree(out);
    HIP_TRY(e, nullptr, 0);
    *ms_out = ms;
    *fma_wave_insts_out = (double)blocks * (256 / 64) * (double)trips * kValuUnroll * kValuChains;
    return check_launch("valu calibration");
}

NAN_API int nan_tracer_tune(NanTracer *tracer, uint32_t blocks_per_cu, uint32_t lds_stack, uint32_t refill_steps) {
    if (!tracer) return fail(nullptr, 0, "nan_tracer_tune: null tracer");
    if (lds_stack != 16 && lds_stack != 24 && lds_stack != 32)
        return fail(nullptr, 0, "nan_tracer_tune: lds_stack must be 16, 24 or 32");
    if (refill_steps == 0 || refill_steps > 4096) return fail(nullptr, 0, "nan_tracer_tune: refill_steps out of range");
    tracer->tuning.blocks_per_cu = blocks_per_cu;
    tracer->tuning.ring = lds_stack;
    tracer->tuning.refill_steps = refill_steps;
    // a new launch shape resizes the working set here, not in the next execute
    return wavefront_prepare(tracer);
}

NAN_API int nan_batch_create(NanScene *const *scenes, NanTracer *const *tracers, uint32_t count, NanTracer **out,
                             char *err, size_t cap) {
    if (!scenes || !tracers || !out || count == 0) return fail(err, cap, "nan_batch_create: null argument");
    *out = nullptr;
    uint64_t nodes = 0, tris = 0, xfs = 0, mats = 0, faces = 0, insts = 0, front = 0, ds = 0, texel_total = 0;
    uint32_t max_stack = 0;
    bool any_textures = false;
    for (uint32_t k = 0; k < count; ++k) {
        const NanScene *sc = scenes[k];
        const NanTracer *tr = tracers[k];
        if (!sc || !tr || tr->scene != sc) return fail(err, cap, "nan_batch_create: tracer %u is not over scene %u", k, k);
        if (sc->batch || tr->scene_count != 1) return fail(err, cap, "nan_batch_create: batches cannot be nested");
        nodes += sc->h_nodes.size() / nanbvh::kNodeWords;
        tris += sc->h_tris.size() / nanbvh::kTriWords;
        xfs += sc->h_xf.size() / 16;
        mats += sc->h_mats.size() / kMaterialWords;
        texel_total += sc->h_texels.size();
        any_textures = any_textures || !sc->h_textures.empty();
        faces += tr->face_count;
        insts += tr->instance_count;
        front += tr->total_front;
        ds += tr->ds_count;
        max_stack = std::max(max_stack, sc->info.max_stack);
    }
    if (tris > nanbvh::kLeafFirstMask || nodes >= (1ull << 31) || front >= (1ull << 32) || faces >= (1ull << 32) ||
        texel_total >= (1ull << 32))
        return fail(err, cap, "nan_batch_create: batch too large for 32-bit indices");
    std::vector<uint32_t> h_nodes;
    std::vector<float> h_tris, h_shade, h_xf, h_itf;
    std::vector<uint32_t> h_mats, h_faces, h_inst, h_sides, h_tf, h_fi, h_prefix, h_ds, h_face_scene, h_inst_base;
    std::vector<uint32_t> h_textures, h_texels;
    std::vector<float> h_tri_uv;
    if (any_textures) h_tri_uv.reserve(tris * 8);
    std::vector<SceneRec> recs(count);
    h_nodes.reserve(nodes * nanbvh::kNodeWords);
    h_tris.reserve(tris * nanbvh::kTriWords);
    h_shade.reserve(tris * 20);
    uint32_t node_off = 0, tri_off = 0, xf_off = 0, mat_off = 0, face_off = 0, inst_off = 0, front_off = 0;
    auto rebase_entry = [&](uint32_t e) -> uint32_t {
        if (e == 0xFFFFFFFFu) return e;
        if (e & nanbvh::kLeafBit) return (e & ~nanbvh::kLeafFirstMask) | ((e & nanbvh::kLeafFirstMask) + tri_off);
        return e + node_off;
    };
    for (uint32_t k = 0; k < count; ++k) {
        const NanScene *sc = scenes[k];
        const NanTracer *tr = tracers[k];
        size_t n0 = h_nodes.size();
        h_nodes.insert(h_nodes.end(), sc->h_nodes.begin(), sc->h_nodes.end());
        for (size_t i = n0; i < h_nodes.size(); i += nanbvh::kNodeWords)
            for (int c = 0; c < 4; ++c) h_nodes[i + 10 + c] = rebase_entry(h_nodes[i + 10 + c]);
        size_t t0 = h_tris.size();
        h_tris.insert(h_tris.end(), sc->h_tris.begin(), sc->h_tris.end());
        for (size_t i = t0; i < h_tris.size(); i += nanbvh::kTriWords) {
            uint32_t id;
            std::memcpy(&id, &h_tris[i + 9], 4);
            id += tri_off;
            std::memcpy(&h_tris[i + 9], &id, 4);
        }
        size_t s0 = h_shade.size();
        h_shade.insert(h_shade.end(), sc->h_shade.begin(), sc->h_shade.end());
        for (size_t i = s0; i < h_shade.size(); i += 20) {
            uint32_t xf, mw;
            std::memcpy(&xf, &h_shade[i + 18], 4);
            std::memcpy(&mw, &h_shade[i + 19], 4);
            xf += xf_off;
            mw = (mw & 0x80000000u) | ((mw & 0x7FFFFFFFu) + mat_off);
            std::memcpy(&h_shade[i + 18], &xf, 4);
            std::memcpy(&h_shade[i + 19], &mw, 4);
        }
        h_xf.insert(h_xf.end(), sc->h_xf.begin(), sc->h_xf.end());
        h_itf.insert(h_itf.end(), sc->h_itf.begin(), sc->h_itf.end());
        size_t m0 = h_mats.size();
        h_mats.insert(h_mats.end(), sc->h_mats.begin(), sc->h_mats.end());
        const uint32_t tex_off = (uint32_t)(h_textures.size() / 4), texel_off = (uint32_t)h_texels.size();
        for (size_t i = m0; i < h_mats.size(); i += kMaterialWords)
            for (int j = 8; j < 11; ++j)
                if (h_mats[i + j] != 0xFFFFFFFFu) h_mats[i + j] += tex_off;
        for (size_t i = 0; i < sc->h_textures.size(); i += 4) {
            h_textures.push_back(sc->h_textures[i] + texel_off);
            h_textures.insert(h_textures.end(), sc->h_textures.begin() + i + 1, sc->h_textures.begin() + i + 4);
        }
        h_texels.insert(h_texels.end(), sc->h_texels.begin(), sc->h_texels.end());
        if (any_textures) {
            if (sc->h_tri_uv.empty())
                h_tri_uv.resize(h_tri_uv.size() + sc->h_tris.size() / nanbvh::kTriWords * 8, 0.0f);
            else
                h_tri_uv.insert(h_tri_uv.end(), sc->h_tri_uv.begin(), sc->h_tri_uv.end());
        }
        recs[k] = SceneRec{};
        recs[k].root = rebase_entry(sc->dev.root);
        recs[k].slack = sc->dev.slack;
        recs[k].max_bounces = 0;
        recs[k].roulette = sc->roulette;
        recs[k].light = sc->dev.light;
        // layout
        h_faces.insert(h_faces.end(), tr->h_faces.begin(), tr->h_faces.end());
        for (uint32_t i = 0; i < tr->instance_count; ++i) {
            const uint32_t *ii = &tr->h_instances[4 * (size_t)i];
            h_inst.push_back(ii[0] + face_off);
            h_inst.push_back(ii[1]);
            h_inst.push_back(ii[2]);
            h_inst.push_back(ii[3]);
            h_sides.push_back(tr->h_sides[2 * (size_t)i]);
            h_sides.push_back(tr->h_sides[2 * (size_t)i + 1]);
            h_prefix.push_back(tr->h_prefix[i] + front_off);
        }
        for (uint32_t f : tr->h_texel_face) h_tf.push_back(f + face_off);
        for (uint32_t i : tr->h_face_inst) h_fi.push_back(i + inst_off);
        for (uint32_t w : tr->h_ds) h_ds.push_back(w + front_off);
        h_face_scene.insert(h_face_scene.end(), tr->face_count, k);
        h_inst_base.push_back(inst_off);
        node_off += (uint32_t)(sc->h_nodes.size() / nanbvh::kNodeWords);
        tri_off += (uint32_t)(sc->h_tris.size() / nanbvh::kTriWords);
        xf_off += (uint32_t)(sc->h_xf.size() / 16);
        mat_off += (uint32_t)(sc->h_mats.size() / kMaterialWords);
        face_off += tr->face_count;
        inst_off += tr->instance_count;
        front_off += (uint32_t)tr->total_front;
    }
    NanScene *bs = new NanScene();
    bs->batch = true;
    auto up = [&](auto **dst, const void *src, size_t bytes) -> int {
        int rc = upload(dst, src, bytes, err, cap);
        if (rc == 0) bs->allocations.push_back((void *)*dst);
        return rc;
    };
    uint4 *d_nodes;
    float4 *d_tris, *d_shade, *d_xf, *d_itf, *d_mats;
    SceneRec *d_recs;
    int rc = 0;
    rc |= up(&d_nodes, h_nodes.data(), h_nodes.size() * 4);
    rc |= up(&d_tris, h_tris.data(), h_tris.size() * 4);
    rc |= up(&d_shade, h_shade.data(), h_shade.size() * 4);
    rc |= up(&d_xf, h_xf.data(), h_xf.size() * 4);
    rc |= up(&d_itf, h_itf.data(), h_itf.size() * 4);
    rc |= up(&d_mats, h_mats.data(), h_mats.size() * 4);
    rc |= up(&d_recs, recs.data(), recs.size() * sizeof(SceneRec));
    rc |= upload_textures(bs, h_mats.data(), h_mats.size() / kMaterialWords, h_textures, h_texels, h_tri_uv, err, cap);
    if (rc) {
        delete bs;
        return -1;
    }
    bs->dev.nodes = d_nodes;
    bs->dev.tris = d_tris;
    bs->dev.tri_shading = d_shade;
    bs->dev.transforms = d_xf;
    bs->dev.inv_transposes = d_itf;
    bs->dev.materials = d_mats;
    bs->dev.scenes = d_recs;
    bs->dev.root = recs[0].root;
    bs->dev.slack = recs[0].slack;
    bs->dev.light = recs[0].light;
    bs->info.triangle_count = tris;
    bs->info.node_count = nodes;
    bs->info.max_stack = max_stack;
    bs->stack_depth = 0;

    NanTracer *t = new NanTracer();
    t->scene = bs;
    t->owned_scene = bs;
    t->scene_count = count;
    t->face_count = (uint32_t)faces;
    t->instance_count = (uint32_t)insts;
    t->total_front = front;
    t->ds_count = (uint32_t)ds;
    t->mode = NAN_TRACER_WAVEFRONT;
    rc = 0;
    rc |= upload(&t->faces, h_faces.data(), h_faces.size() * 4, err, cap);
    rc |= upload(&t->instances, h_inst.data(), h_inst.size() * 4, err, cap);
    rc |= upload(&t->sides, h_sides.data(), h_sides.size() * 4, err, cap);
    rc |= upload(&t->texel_face, h_tf.data(), h_tf.size() * 4, err, cap);
    rc |= upload(&t->face_inst, h_fi.data(), h_fi.size() * 4, err, cap);
    rc |= upload(&t->front_prefix, h_prefix.data(), h_prefix.size() * 4, err, cap);
    rc |= upload(&t->ds_list, h_ds.data(), h_ds.size() * 4, err, cap);
    rc |= upload_single_sided(t, h_ds, err, cap);
    rc |= upload(&t->face_scene, h_face_scene.data(), h_face_scene.size() * 4, err, cap);
    rc |= upload(&t->scene_inst_base, h_inst_base.data(), h_inst_base.size() * 4, err, cap);
    rc |= upload(&t->payload_table, nullptr, (size_t)count * sizeof(float4 *), err, cap);
    if (!rc && wavefront_prepare(t)) rc = fail(err, cap, "nan_batch_create: %s", g_last_error.c_str());
    if (rc) {
        delete t;
        return -1;
    }
    t->payload_table_set = false;
    *out = t;
    return 0;
}

NAN_API int nan_batch_set_payloads(NanTracer *batch, void *const *payloads) {
    if (!batch || !payloads || !batch->scene->batch || !batch->payload_table)
        return fail(nullptr, 0, "nan_batch_set_payloads: not a batch or null payloads");
    for (uint32_t k = 0; k < batch->scene_count; ++k)
        if (!payloads[k]) return fail(nullptr, 0, "nan_batch_set_payloads: payload %u is null", k);
    HIP_TRY(hipMemcpy(batch->payload_table, payloads, (size_t)batch->scene_count * sizeof(void *),
                      hipMemcpyHostToDevice), nullptr, 0);
    batch->payload_table_set = true;
    return 0;
}

NAN_API int nan_batch_set_bounces(NanTracer *batch, const uint32_t *max_bounces) {
    if (!batch || !max_bounces || !batch->scene->batch) return fail(nullptr, 0, "nan_batch_set_bounces: not a batch");
    std::vector<SceneRec> recs(batch->scene_count);
    HIP_TRY(hipMemcpy(recs.data(), batch->scene->dev.scenes, recs.size() * sizeof(SceneRec), hipMemcpyDeviceToHost),
            nullptr, 0);
    for (uint32_t k = 0; k < batch->scene_count; ++k) recs[k].max_bounces = max_bounces[k];
    HIP_TRY(hipMemcpy((void *)batch->scene->dev.scenes, recs.data(), recs.size() * sizeof(SceneRec),
                      hipMemcpyHostToDevice), nullptr, 0);
    return 0;
}

NAN_API int nan_batch_set_frame_index(NanTracer *batch, uint32_t frame_index, void *stream) {
    if (!batch || !batch->scene->batch) return fail(nullptr, 0, "nan_batch_set_frame_index: not a batch");
    return set_frame_index(batch->scene, frame_index, stream);
}

NAN_API int nan_tracer_device_frame(NanTracer *tracer, int enable, uint32_t iteration, uint32_t reset) {
    if (!tracer) return fail(nullptr, 0, "nan_tracer_device_frame: null tracer");
    if (enable && tracer->mode != NAN_TRACER_WAVEFRONT)
        return fail(nullptr, 0, "nan_tracer_device_frame: wavefront mode only");
    tracer->device_frame = enable != 0;
    if (enable) return wavefront_set_frame(tracer, iteration, reset);
    return 0;
}

}  // extern "C"
