// Mesh Colors surface filter: one-ring, normal-aware diffusion of the texel
// cache (MeshColorsSurfaceFilter, mesh_colors_surface_filter.py:12-128;
// kernel mesh_colors_surface_filter.slang:17-253).  One launch covers every
// front texel of every instance through the tracer's flat texel->face map
// (the reference dispatches per instance and binary-searches the face); each
// texel filters its front slot and, on double-sided instances, its back slot.
#include <hip/hip_runtime.h>

#include "nan_internal.h"

using namespace nanx;

namespace {

constexpr uint32_t kInvalidAdjacency = 0xFFFFFFFFu;

// mesh_colors_filter_offset (mesh_colors_surface_filter.slang:17-29)
__device__ __forceinline__ int2 filter_offset(uint32_t k) {
    const int dx[6] = {1, -1, 0, 0, 1, -1}, dy[6] = {0, 0, 1, -1, -1, 1};
    return make_int2(dx[k], dy[k]);
}

// mesh_colors_map_one_edge_tap (mesh_colors.slang:152-200): barycentrics on
// the neighbour of a tap one lattice step across `crossed`.
__device__ __forceinline__ bool map_one_edge_tap(int ci, int cj, uint32_t resolution, uint32_t adj_edge, bool flip,
                                                 uint32_t crossed, float nb[3]) {
    float inv = 1.0f / (float)resolution;
    float src[3] = {(float)ci * inv, (float)cj * inv, 1.0f - (float)(ci + cj) * inv};
    uint32_t s0 = crossed, s1 = (crossed + 1) % 3, sop = (crossed + 2) % 3;
    uint32_t n0 = adj_edge, n1 = (adj_edge + 1) % 3, nop = (adj_edge + 2) % 3;
    float outside = src[sop];
    uint32_t m0 = flip ? n1 : n0, m1 = flip ? n0 : n1;
    nb[0] = nb[1] = nb[2] = 0.0f;
    nb[m0] = src[s0] + outside;
    nb[m1] = src[s1] + outside;
    nb[nop] = -outside;
    for (int k = 0; k < 3; ++k)
        if (nb[k] < -1e-5f || nb[k] > 1.0f + 1e-5f) return false;
    for (int k = 0; k < 3; ++k) nb[k] = fmaxf(nb[k], 0.0f);
    float total = fmaxf((nb[0] + nb[1]) + nb[2], 1e-6f);
    for (int k = 0; k < 3; ++k) nb[k] = nb[k] / total;
    return true;
}

// mesh_colors_normal_filter_weight (mesh_colors.slang:202-214)
__device__ __forceinline__ float normal_filter_weight(f3 c, f3 s, float sigma) {
    float cosine = fminf(fmaxf(dot(c, s), 0.0f), 1.0f);
    if (sigma <= 1e-6f) return cosine >= 1.0f - 1e-6f ? 1.0f : 0.0f;
    float a = acosf(cosine) / sigma;
    return expf(-0.5f * a * a);
}

__global__ void surface_filter_kernel(DevLayout Ly, const uint4 *adjacency, const float4 *normals,
                                      const float4 *src, float4 *dst, float neighbor_weight, float sigma) {
    uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= Ly.total_front) return;
    uint32_t f = Ly.texel_face[w];
    uint32_t inst = Ly.face_inst[f];
    uint4 ii = Ly.instances[inst];
    uint2 face = Ly.faces[f];
    uint2 side = Ly.sides[inst];
    uint32_t local = (uint32_t)(w - Ly.front_prefix[inst]);
    uint32_t i, j;
    address_to_ij(local - face.x, face.y, i, j);
    uint32_t front = ii.z + local;
    float4 cn = normals[f];
    f3 center_normal = mk(cn.x, cn.y, cn.z);
    uint4 adj = adjacency[f];
    const int R = (int)face.y;
    for (int sd = 0; sd < 2; ++sd) {
        bool back = sd == 1;
        if (back && !(side.y & 1u)) break;
        uint32_t center = side_texel(front, ii.z, side, back);
        float4 c = src[center];
        f3 acc = mk(c.x, c.y, c.z);
        float wsum = 1.0f;
        for (uint32_t k = 0; k < 6; ++k) {
            int2 off = filter_offset(k);
            int ci = (int)i + off.x, cj = (int)j + off.y;
            if (ci >= 0 && cj >= 0 && ci + cj <= R) {
                uint32_t t = lattice_address((uint32_t)ci, (uint32_t)cj, ii.z, face);
                float4 v = src[side_texel(t, ii.z, side, back)];
                acc = acc + mk(v.x, v.y, v.z) * neighbor_weight;
                wsum = wsum + neighbor_weight;
                continue;
            }
            uint32_t crossed = ci < 0 ? 1u : (cj < 0 ? 2u : 0u);
            uint32_t packed = crossed == 0 ? adj.x : (crossed == 1 ? adj.y : adj.z);
            if (packed == kInvalidAdjacency) continue;
            uint32_t adj_face = ii.x + (packed >> 3);
            float nb[3];
            if (!map_one_edge_tap(ci, cj, face.y, packed & 3u, (packed & 4u) != 0, crossed, nb)) continue;
            float4 an = normals[adj_face];
            float tw = neighbor_weight * normal_filter_weight(center_normal, mk(an.x, an.y, an.z), sigma);
            if (!(tw > 0.0f)) continue;
            f3 v = sample_irradiance<false>(src, mk(nb[0], nb[1], nb[2]), ii.z, Ly.faces[adj_face], side, back);
            acc = acc + v * tw;
            wsum = wsum + tw;
        }
        float inv = fmaxf(wsum, 1e-6f);
        dst[center] = make_float4(acc.x / inv, acc.y / inv, acc.z / inv, c.w);
    }
}

}  // namespace

extern "C" {

NAN_API int nan_surface_filter(NanTracer *tracer, const uint32_t *adjacency, const float *face_normals,
                               const void *src, void *dst, float neighbor_weight, float normal_sigma_radians,
                               void *stream) {
    if (!tracer || !adjacency || !face_normals || !src || !dst)
        return fail(nullptr, 0, "nan_surface_filter: null argument");
    if (src == dst) return fail(nullptr, 0, "nan_surface_filter: requires distinct source and destination");
    if (tracer->scene->batch) return fail(nullptr, 0, "nan_surface_filter: scene batches are not supported");
    if (tracer->total_front == 0) return 0;
    hipLaunchKernelGGL(surface_filter_kernel, dim3(grid_for(tracer->total_front, kFlatBlock)), dim3(kFlatBlock), 0,
                       as_stream(stream), dev_layout(tracer), reinterpret_cast<const uint4 *>(adjacency),
                       reinterpret_cast<const float4 *>(face_normals), static_cast<const float4 *>(src),
                       static_cast<float4 *>(dst), neighbor_weight, normal_sigma_radians);
    return check_launch("surface_filter");
}

}  // extern "C"
