// Atmosphere lookup tables (atmosphere.py / atmosphere.slang of the
// reference; Hillaire, "A Scalable and Production Ready Sky and Atmosphere
// Rendering Technique", EGSR 2020): transmittance 256x64, multiple
// scattering 32x32 and sky-view 200x200, each a float4 image stored row by
// row ([y][x], x along u).  Units are megameters; the FP32 steps follow the
// reference.  Texture reads are bilinear with clamp-to-edge addressing
// written out (the reference's linear sampler), one thread per texel.
#include <hip/hip_runtime.h>

#include "nan_internal.h"

using namespace nanx;

namespace {

constexpr float kPi = 3.14159265358f;
constexpr float kGround = 6.360f, kTop = 6.460f;
constexpr int kTW = 256, kTH = 64, kMW = 32, kMH = 32, kSW = 200, kSH = 200;

struct Atmo {
    f3 rs;  // Rayleigh scattering per megameter
    f3 oz;  // ozone absorption
};

__device__ __forceinline__ float length(f3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 divs(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float safeacos(float x) { return acosf(fminf(fmaxf(x, -1.0f), 1.0f)); }

__device__ __forceinline__ float mie_phase(float c) {
    const float g = 0.8f, scale = 3.0f / (8.0f * kPi);
    float num = (1.0f - g * g) * (1.0f + c * c);
    float denom = (2.0f + g * g) * powf(1.0f + g * g - 2.0f * g * c, 1.5f);
    return scale * num / denom;
}

__device__ __forceinline__ float rayleigh_phase(float c) { return (3.0f / (16.0f * kPi)) * (1.0f + c * c); }

// getScatteringValues (atmosphere.slang:45-62)
__device__ __forceinline__ void scattering(f3 pos, f3 &rayleigh, float &mie, f3 &extinction) {
    const f3 rs_base = mk(5.802f, 13.558f, 33.1f), oz_base = mk(0.650f, 1.881f, 0.085f);
    float alt = (length(pos) - kGround) * 1000.0f;
    float rd = expf(-alt / 8.0f), md = expf(-alt / 1.2f);
    rayleigh = rs_base * rd;
    float ra = 0.0f * rd;
    mie = 3.996f * md;
    float ma = 4.4f * md;
    f3 oz = oz_base * fmaxf(0.0f, 1.0f - fabsf(alt - 25.0f) / 15.0f);
    extinction = mk(rayleigh.x + ra + mie + ma, rayleigh.y + ra + mie + ma, rayleigh.z + ra + mie + ma) + oz;
}

// rayIntersectSphere (atmosphere.slang:68-77)
__device__ __forceinline__ float sphere_hit(f3 ro, f3 rd, float rad) {
    float b = dot(ro, rd);
    float c = dot(ro, ro) - rad * rad;
    if (c > 0.0f && b > 0.0f) return -1.0f;
    float discr = b * b - c;
    if (discr < 0.0f) return -1.0f;
    if (discr > b * b) return -b + sqrtf(discr);
    return -b - sqrtf(discr);
}

// bilinear, clamp to edge
__device__ __forceinline__ f3 sample_clamp(const float4 *lut, int w, int h, float u, float v) {
    float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float tx = x - fx, ty = y - fy;
    int x0 = (int)fx, y0 = (int)fy;
    int xa = min(max(x0, 0), w - 1), xb = min(max(x0 + 1, 0), w - 1);
    int ya = min(max(y0, 0), h - 1), yb = min(max(y0 + 1, 0), h - 1);
    float4 a = lut[ya * w + xa], b = lut[ya * w + xb], c = lut[yb * w + xa], d = lut[yb * w + xb];
    f3 r0 = mk(a.x, a.y, a.z) * (1.0f - tx) + mk(b.x, b.y, b.z) * tx;
    f3 r1 = mk(c.x, c.y, c.z) * (1.0f - tx) + mk(d.x, d.y, d.z) * tx;
    return r0 * (1.0f - ty) + r1 * ty;
}

// getValFromTLUT / getValFromMultiScattLUT (atmosphere.slang:83-113)
__device__ __forceinline__ f3 lut_at(const float4 *lut, int w, int h, f3 pos, f3 sun) {
    float height = length(pos);
    f3 up = divs(pos, height);
    float c = dot(sun, up);
    float u = fminf(fmaxf(0.5f + 0.5f * c, 0.0f), 1.0f);
    float v = fminf(fmaxf((height - kGround) / (kTop - kGround), 0.0f), 1.0f);
    return sample_clamp(lut, w, h, u, v);
}

__device__ __forceinline__ float lerpf(float a, float b, float t) { return a + t * (b - a); }

// getSunTransmittance (atmosphere.slang:117-140)
__device__ f3 sun_transmittance(f3 pos, f3 sun) {
    if (sphere_hit(pos, sun, kGround) > 0.0f) return mk(0.0f, 0.0f, 0.0f);
    float atmo = sphere_hit(pos, sun, kTop);
    float t = 0.0f;
    f3 tr = mk(1.0f, 1.0f, 1.0f);
    for (float i = 0.0f; i < 40.0f; i += 1.0f) {
        float nt = ((i + 0.3f) / 40.0f) * atmo;
        float dt = nt - t;
        t = nt;
        f3 rs, ext;
        float ms;
        scattering(pos + sun * t, rs, ms, ext);
        tr = tr * mk(expf(-dt * ext.x), expf(-dt * ext.y), expf(-dt * ext.z));
    }
    return tr;
}

__device__ __forceinline__ void lut_coords(int px, int py, int w, int h, f3 &pos, f3 &sun) {
    float u = ((float)px + 0.5f) / (float)w, v = ((float)py + 0.5f) / (float)h;
    float c = 2.0f * u - 1.0f;
    float theta = safeacos(c);
    pos = mk(0.0f, lerpf(kGround, kTop, v), 0.0f);
    sun = normalize(mk(0.0f, c, -sinf(theta)));
}

__global__ void transmittance_lut_kernel(float4 *out) {
    int px = blockIdx.x * blockDim.x + threadIdx.x, py = blockIdx.y * blockDim.y + threadIdx.y;
    if (px >= kTW || py >= kTH) return;
    f3 pos, sun;
    lut_coords(px, py, kTW, kTH, pos, sun);
    f3 t = sun_transmittance(pos, sun);
    out[py * kTW + px] = make_float4(t.x, t.y, t.z, 1.0f);
}

// getMulScattValues (atmosphere.slang:170-240) and Eq. 10
__global__ void multiscatt_lut_kernel(const float4 *tlut, float4 *out) {
    int px = blockIdx.x * blockDim.x + threadIdx.x, py = blockIdx.y * blockDim.y + threadIdx.y;
    if (px >= kMW || py >= kMH) return;
    f3 pos, sun;
    lut_coords(px, py, kMW, kMH, pos, sun);
    const f3 ground_albedo = mk(0.3f, 0.3f, 0.3f);
    const int sq = 8;
    float inv = 1.0f / (float)(sq * sq);
    f3 lum_total = mk(0.0f, 0.0f, 0.0f), fms = mk(0.0f, 0.0f, 0.0f);
    for (int i = 0; i < sq; ++i) {
        for (int j = 0; j < sq; ++j) {
            float theta = kPi * ((float)i + 0.5f) / (float)sq;
            float phi = safeacos(1.0f - 2.0f * ((float)j + 0.5f) / (float)sq);
            float cp = cosf(phi), sp = sinf(phi), ct = cosf(theta), st = sinf(theta);
            f3 dir = mk(sp * st, cp, sp * ct);
            float atmo = sphere_hit(pos, dir, kTop), grd = sphere_hit(pos, dir, kGround);
            float tmax = grd > 0.0f ? grd : atmo;
            float c = dot(dir, sun);
            float mp = mie_phase(c), rp = rayleigh_phase(-c);
            f3 lum = mk(0.0f, 0.0f, 0.0f), lum_factor = mk(0.0f, 0.0f, 0.0f), tr = mk(1.0f, 1.0f, 1.0f);
            float t = 0.0f;
            for (float s = 0.0f; s < 20.0f; s += 1.0f) {
                float nt = ((s + 0.3f) / 20.0f) * tmax;
                float dt = nt - t;
                t = nt;
                f3 np = pos + dir * t;
                f3 rs, ext;
                float ms;
                scattering(np, rs, ms, ext);
                f3 st_ = mk(expf(-dt * ext.x), expf(-dt * ext.y), expf(-dt * ext.z));
                f3 snp = mk(rs.x + ms, rs.y + ms, rs.z + ms);
                f3 sf = mk((snp.x - snp.x * st_.x) / ext.x, (snp.y - snp.y * st_.y) / ext.y,
                           (snp.z - snp.z * st_.z) / ext.z);
                lum_factor = lum_factor + tr * sf;
                f3 sunt = lut_at(tlut, kTW, kTH, np, sun);
                f3 ris = rs * rp;
                float mis = ms * mp;
                f3 ins = mk(ris.x + mis, ris.y + mis, ris.z + mis) * sunt;
                f3 si = mk((ins.x - ins.x * st_.x) / ext.x, (ins.y - ins.y * st_.y) / ext.y,
                           (ins.z - ins.z * st_.z) / ext.z);
                lum = lum + si * tr;
                tr = tr * st_;
            }
            if (grd > 0.0f) {
                f3 hit = pos + dir * grd;
                if (dot(pos, sun) > 0.0f) {
                    hit = normalize(hit) * kGround;
                    lum = lum + tr * ground_albedo * lut_at(tlut, kTW, kTH, hit, sun);
                }
            }
            fms = fms + lum_factor * inv;
            lum_total = lum_total + lum * inv;
        }
    }
    f3 psi = mk(lum_total.x / (1.0f - fms.x), lum_total.y / (1.0f - fms.y), lum_total.z / (1.0f - fms.z));
    out[py * kMW + px] = make_float4(psi.x, psi.y, psi.z, 1.0f);
}

// raymarchScattering + sky_view_lut_compute (atmosphere.slang:267-397)
__global__ void sky_view_lut_kernel(const float4 *tlut, const float4 *mslut, f3 sun_direction, float4 *out) {
    int px = blockIdx.x * blockDim.x + threadIdx.x, py = blockIdx.y * blockDim.y + threadIdx.y;
    if (px >= kSW || py >= kSH) return;
    float u = ((float)px + 0.5f) / (float)kSW, v = ((float)py + 0.5f) / (float)kSH;
    float az = (u - 0.5f) * 2.0f * kPi;
    float adj;
    if (v < 0.5f) {
        float c = 1.0f - 2.0f * v;
        adj = -c * c;
    } else {
        float c = v * 2.0f - 1.0f;
        adj = c * c;
    }
    f3 view = mk(0.0f, kGround + 0.0002f, 0.0f);
    float height = length(view);
    f3 up = divs(view, height);
    float horizon = safeacos(sqrtf(height * height - kGround * kGround) / height) - 0.5f * kPi;
    float alt = adj * 0.5f * kPi - horizon;
    float ca = cosf(alt);
    f3 dir = mk(ca * sinf(az), sinf(alt), -ca * cosf(az));
    float sun_alt = (0.5f * kPi) - acosf(dot(sun_direction, up));
    f3 sun = mk(0.0f, sinf(sun_alt), -cosf(sun_alt));
    float atmo = sphere_hit(view, dir, kTop), grd = sphere_hit(view, dir, kGround);
    float tmax = grd < 0.0f ? atmo : grd;
    float c = dot(dir, sun);
    float mp = mie_phase(c), rp = rayleigh_phase(-c);
    f3 lum = mk(0.0f, 0.0f, 0.0f), tr = mk(1.0f, 1.0f, 1.0f);
    float t = 0.0f;
    const float steps = 32.0f;
    for (float i = 0.0f; i < steps; i += 1.0f) {
        float nt = ((i + 0.3f) / steps) * tmax;
        float dt = nt - t;
        t = nt;
        f3 np = view + dir * t;
        f3 rs, ext;
        float ms;
        scattering(np, rs, ms, ext);
        f3 st_ = mk(expf(-dt * ext.x), expf(-dt * ext.y), expf(-dt * ext.z));
        f3 sunt = lut_at(tlut, kTW, kTH, np, sun);
        f3 psi = lut_at(mslut, kMW, kMH, np, sun);
        f3 ris = rs * (sunt * rp + psi);
        f3 mis = (sunt * mp + psi) * ms;
        f3 ins = ris + mis;
        f3 si = mk((ins.x - ins.x * st_.x) / ext.x, (ins.y - ins.y * st_.y) / ext.y, (ins.z - ins.z * st_.z) / ext.z);
        lum = lum + si * tr;
        tr = tr * st_;
    }
    out[py * kSW + px] = make_float4(lum.x, lum.y, lum.z, 1.0f);
}

// get_sun_transmittance_at_view (scene.slang:997-1009)
__global__ void view_transmittance_kernel(const float4 *tlut, f3 sun_direction, float *out) {
    if (threadIdx.x != 0) return;
    f3 view = mk(0.0f, kGround + 0.0002f, 0.0f);
    f3 t = lut_at(tlut, kTW, kTH, view, normalize(sun_direction));
    out[0] = t.x;
    out[1] = t.y;
    out[2] = t.z;
    out[3] = 1.0f;
}

// Scene.sample_sky_lut (scene.slang:469-518) for a list of directions
__global__ void sky_lookup_kernel(Light L, const float4 *dirs, uint64_t n, float4 *out) {
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float4 d = dirs[k];
    f3 r = sky_view_radiance(L, mk(d.x, d.y, d.z));
    out[k] = make_float4(r.x, r.y, r.z, 0.0f);
}

dim3 grid2(int w, int h) { return dim3((w + 7) / 8, (h + 7) / 8); }

}  // namespace

extern "C" {

NAN_API int nan_atmosphere_transmittance_lut(float *out, void *stream) {
    if (!out) return fail(nullptr, 0, "nan_atmosphere_transmittance_lut: null output");
    hipLaunchKernelGGL(transmittance_lut_kernel, grid2(kTW, kTH), dim3(8, 8), 0, as_stream(stream),
                       reinterpret_cast<float4 *>(out));
    return check_launch("transmittance_lut");
}

NAN_API int nan_atmosphere_multiscattering_lut(const float *transmittance, float *out, void *stream) {
    if (!transmittance || !out) return fail(nullptr, 0, "nan_atmosphere_multiscattering_lut: null argument");
    hipLaunchKernelGGL(multiscatt_lut_kernel, grid2(kMW, kMH), dim3(8, 8), 0, as_stream(stream),
                       reinterpret_cast<const float4 *>(transmittance), reinterpret_cast<float4 *>(out));
    return check_launch("multiscattering_lut");
}

NAN_API int nan_atmosphere_sky_view_lut(const float *transmittance, const float *multiscattering,
                                        const float *sun_direction3, float *out, void *stream) {
    if (!transmittance || !multiscattering || !sun_direction3 || !out)
        return fail(nullptr, 0, "nan_atmosphere_sky_view_lut: null argument");
    f3 sun = mk(sun_direction3[0], sun_direction3[1], sun_direction3[2]);
    hipLaunchKernelGGL(sky_view_lut_kernel, grid2(kSW, kSH), dim3(8, 8), 0, as_stream(stream),
                       reinterpret_cast<const float4 *>(transmittance),
                       reinterpret_cast<const float4 *>(multiscattering), sun, reinterpret_cast<float4 *>(out));
    return check_launch("sky_view_lut");
}

NAN_API int nan_atmosphere_view_transmittance(const float *transmittance, const float *sun_direction3, float *out4,
                                              void *stream) {
    if (!transmittance || !sun_direction3 || !out4)
        return fail(nullptr, 0, "nan_atmosphere_view_transmittance: null argument");
    f3 sun = mk(sun_direction3[0], sun_direction3[1], sun_direction3[2]);
    if (!(dot(sun, sun) > 0.0f)) return fail(nullptr, 0, "nan_atmosphere_view_transmittance: zero sun direction");
    hipLaunchKernelGGL(view_transmittance_kernel, dim3(1), dim3(64), 0, as_stream(stream),
                       reinterpret_cast<const float4 *>(transmittance), sun, out4);
    return check_launch("view_transmittance");
}

NAN_API int nan_sky_view_lookup(const float *sky_view_lut, const float *lighting16, const float *dirs, uint64_t n,
                                float *out, void *stream) {
    if (!sky_view_lut || !lighting16 || (n && (!dirs || !out)))
        return fail(nullptr, 0, "nan_sky_view_lookup: null argument");
    if (n == 0) return 0;
    Light L = make_light(lighting16, reinterpret_cast<const float4 *>(sky_view_lut));
    hipLaunchKernelGGL(sky_lookup_kernel, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream), L,
                       reinterpret_cast<const float4 *>(dirs), n, reinterpret_cast<float4 *>(out));
    return check_launch("sky_view_lookup");
}

}  // extern "C"
