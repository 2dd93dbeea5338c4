// Device-side building blocks of the texture-space tracer for gfx950.
//
// Arithmetic contract (shared with the CPU oracle): FP32 with no implicit
// contraction (this translation unit is compiled with -ffp-contract=off),
// __builtin_fmaf exactly where the reference writes mad(), correctly
// rounded division and sqrt (hipcc's default for FP32), and sin/cos of
// 2*pi*u from sincos_2pi().  Under this contract every payload slot equals
// the oracle's bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nan {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr uint32_t kMaxTexelSamples = 65535u;
constexpr float kPi = 0x1.921fb6p+1f;
constexpr float kTwoPi = 0x1.921fb6p+2f;
constexpr float kInvPi = 0x1.45f306p-2f;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kLeafCountShift = 27;
constexpr uint32_t kLeafFirstMask = (1u << kLeafCountShift) - 1u;
// Ray-box slack.  A child box is visited when its entry distance is within
// far * (1 + 2^-10) + slack and within best_t * (1 + 2^-10) + slack (each one
// FMA), where
// slack = 2^-18 * max(1, largest world coordinate) (DevScene::slack).  The
// relative term covers the slab test's rounding (2*gamma_3 of Ize, "Robust
// BVH Ray Traversal", JCGT 2013); the absolute term covers the FP32 triangle
// test reporting a distance below its box's true entry (or a hit just behind
// the origin) by a few ulps of the coordinates, so culling never changes the
// closest hit.
constexpr float kPad = 0x1.004p+0f;

struct f3 {
    float x, y, z;
};
__host__ __device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ __forceinline__ f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__host__ __device__ __forceinline__ f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__host__ __device__ __forceinline__ f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
__host__ __device__ __forceinline__ f3 fabs3(f3 a) { return mk(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
__host__ __device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ f3 normalize(f3 a) {
    float inv = 1.0f / sqrtf(dot(a, a));
    return a * inv;
}
__device__ __forceinline__ float comp(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
__device__ __forceinline__ float saturate(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
__device__ __forceinline__ float fracf(float x) { return x - floorf(x); }
__device__ __forceinline__ float exp2i(int e) { return __int_as_float((e + 127) << 23); }

// atan2 and asin from correctly rounded operations only, so that the sky
// lookup below is bit-identical to the oracle's: arctan on [0, 1] by the
// polynomial of Abramowitz & Stegun 4.4.49 (|error| <= 2e-8), extended to
// all octants by atan(a) = pi/2 - atan(1/a) and the signs of y and x.
__device__ __forceinline__ float nan_atan2(float y, float x) {
    const float c[9] = {1.0f,           -0.3333314528f, 0.1999355085f, -0.1420889944f, 0.1065626393f,
                        -0.0752896400f, 0.0429096138f,  -0.0161657367f, 0.0028662257f};
    float ax = fabsf(x), ay = fabsf(y);
    float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    if (!(mx > 0.0f)) return 0.0f;
    float a = mn / mx, s = a * a;
    float p = c[8];
    for (int k = 7; k >= 0; --k) p = fmaf(p, s, c[k]);
    float r = a * p;
    if (ay > ax) r = 0.5f * kPi - r;
    if (x < 0.0f) r = kPi - r;
    return y < 0.0f ? -r : r;
}
__device__ __forceinline__ float nan_asin(float x) {
    x = fminf(fmaxf(x, -1.0f), 1.0f);
    return nan_atan2(x, sqrtf((1.0f - x) * (1.0f + x)));
}


// sin(2*pi*u), cos(2*pi*u) for u in [0, 1]: exact quarter-turn reduction on
// u, then degree-9/10 Taylor polynomials on [-pi/4, pi/4].
__device__ __forceinline__ void sincos_2pi(float u, float *s_out, float *c_out) {
    float k = rintf(u * 4.0f);
    float v = u - k * 0.25f;
    float a = v * kTwoPi;
    float a2 = a * a;
    float p = __builtin_fmaf(a2, 0x1.71de3ap-19f, -0x1.a01a02p-13f);
    p = __builtin_fmaf(a2, p, 0x1.111112p-7f);
    p = __builtin_fmaf(a2, p, -0x1.555556p-3f);
    float s = __builtin_fmaf(a * a2, p, a);
    float q = __builtin_fmaf(a2, -0x1.27e4fcp-22f, 0x1.a01a02p-16f);
    q = __builtin_fmaf(a2, q, -0x1.6c16c2p-10f);
    q = __builtin_fmaf(a2, q, 0x1.555556p-5f);
    q = __builtin_fmaf(a2, q, -0.5f);
    float c = __builtin_fmaf(a2, q, 1.0f);
    int quadrant = ((int)k) & 3;
    float rs = quadrant == 0 ? s : (quadrant == 1 ? c : (quadrant == 2 ? -s : -c));
    float rc = quadrant == 0 ? c : (quadrant == 1 ? -s : (quadrant == 2 ? -c : s));
    *s_out = rs;
    *c_out = rc;
}

// ---------------------------------------------------------------------------
// RNG: 32-bit LCG (common.slang:164-210)
// ---------------------------------------------------------------------------
// MurmurHash3's block mix and finaliser over the key (h, k) (Appleby,
// MurmurHash3, public domain; common.slang:520-545).
__device__ __forceinline__ uint32_t murmur_key_mix(uint32_t h, uint32_t k) {
    k *= 0xcc9e2d51u;
    k = (k << 15) | (k >> 17);
    k *= 0x1b873593u;
    h ^= k;
    h = (h << 13) | (h >> 19);
    h = h * 5u + 0xe6546b64u;
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

// Per-texel random streams.  pcg == 0: the reference's LCG (common.slang:
// 164-210), seeded from (texel, instance, sample).  pcg == 1: PCG
// (O'Neill, "PCG: A Family of Simple Fast Space-Efficient Statistically Good
// Algorithms for Random Number Generation", 2014) with the RXS-M-XS output
// on a 32-bit LCG state, started at a hash of (texel, instance, sample).
// Both map a 32-bit draw to [0, 1) as the reference does, (r >> 8) * 2^-24.
// pcg comes from a kernel argument, so it stays in a scalar register.
struct Rng {
    uint32_t state;
    uint32_t pcg;
    __device__ __forceinline__ void seed(uint32_t x, uint32_t y, uint32_t s) {
        state = (x * 100591u) ^ (y * 102191u) ^ (s * 103643u);
        pcg = 0u;
    }
    __device__ __forceinline__ void seed_pcg(uint32_t x, uint32_t y, uint32_t s) {
        state = murmur_key_mix(murmur_key_mix(x, y), s);
        pcg = 1u;
    }
    __device__ __forceinline__ float next_1d() {
        uint32_t r;
        if (pcg == 0u) {
            state = 1664525u * state + 1013904223u;
            r = state;
        } else {
            uint32_t old = state;
            state = old * 747796405u + 2891336453u;
            uint32_t word = ((old >> ((old >> 28u) + 4u)) ^ old) * 277803737u;
            r = (word >> 22u) ^ word;
        }
        return (float)(r >> 8) * 0x1p-24f;
    }
};

// ---------------------------------------------------------------------------
// scene as seen by the kernels
// ---------------------------------------------------------------------------
struct Light {
    f3 sun_dir;
    float cos_half;
    f3 sun_illum;     // color * intensity * transmittance
    f3 outer;         // color * intensity
    f3 transmittance;
    f3 sky;                 // constant sky radiance (used when sky_lut is null)
    int sun_on;
    const float4 *sky_lut;  // sky-view table float4[200][200] (scene.slang:469-518), or null
    float horizon;          // horizon angle of the view point, radians (sky_lut only)
};

// Sky-view table lookup (Scene.sample_sky_lut, scene.slang:469-518): the
// inverse of the table's altitude/azimuth mapping around the sun, bilinear
// with wrap-around addressing (the reference's default sampler), times the
// sun's outer-space illuminance.  A direction along the vertical has no
// azimuth; it reads azimuth 0.
__device__ __forceinline__ f3 sky_view_radiance(const Light &L, f3 dir) {
    const f3 view = mk(0.0f, 6.360f + 0.0002f, 0.0f);
    float height = sqrtf(dot(view, view));
    f3 up = mk(view.x / height, view.y / height, view.z / height);
    float alt = fminf(fmaxf(nan_asin(dot(dir, up)), -0.5f * kPi + 1e-2f), 0.5f * kPi - 1e-2f);
    float az = 0.0f;
    if (!(fabsf(alt) > 0.5f * kPi - 0.0001f)) {
        f3 right = cross(L.sun_dir, up);
        f3 forward = cross(up, right);
        f3 p = dir - up * dot(dir, up);
        if (dot(p, p) > 0.0f) {
            p = normalize(p);
            az = nan_atan2(dot(p, right), dot(p, forward));
        }
    }
    float adj = (alt + L.horizon) / (0.5f * kPi);
    float v = adj < 0.0f ? 0.5f * (1.0f - sqrtf(fminf(fmaxf(-adj, 0.0f), 1.0f)))
                         : 0.5f * (1.0f + sqrtf(fminf(fmaxf(adj, 0.0f), 1.0f)));
    float u = az / (2.0f * kPi) + 0.5f;
    if (!(u >= 0.0f && u <= 1.0f)) u = 0.5f;  // a NaN direction reads a valid texel
    if (!(v >= 0.0f && v <= 1.0f)) v = 0.5f;
    float x = u * 200.0f - 0.5f, y = v * 200.0f - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float tx = x - fx, ty = y - fy;
    int x0 = (int)fx, y0 = (int)fy;  // u, v lie in [0, 1]: x0, y0 in [-1, 199]
    int xa = (x0 + 200) % 200, xb = (x0 + 201) % 200, ya = (y0 + 200) % 200, yb = (y0 + 201) % 200;
    float4 a = L.sky_lut[ya * 200 + xa], b = L.sky_lut[ya * 200 + xb];
    float4 c = L.sky_lut[yb * 200 + xa], d = L.sky_lut[yb * 200 + xb];
    f3 r0 = mk(a.x, a.y, a.z) * (1.0f - tx) + mk(b.x, b.y, b.z) * tx;
    f3 r1 = mk(c.x, c.y, c.z) * (1.0f - tx) + mk(d.x, d.y, d.z) * tx;
    return (r0 * (1.0f - ty) + r1 * ty) * L.outer;
}

// radiance of a ray that leaves the scene: the sky-view table when the scene
// has one, else the constant sky
__device__ __forceinline__ f3 miss_radiance(const Light &L, f3 dir) {
    return L.sky_lut ? sky_view_radiance(L, dir) : L.sky;
}

// Per-scene record for kernels that trace several scenes in one launch.
struct SceneRec {
    uint32_t root;         // BVH root entry (rebased into the shared node array)
    float slack;           // absolute box-test slack of the scene
    uint32_t max_bounces;  // 0: use the launch's value
    uint32_t roulette;     // 1: Russian roulette on every bounce that would continue
    Light light;
};

struct DevScene {
    const uint4 *nodes;           // 64-byte BVH4 nodes
    const float4 *tris;           // 48-byte triangle records, 3 float4 each
    const uint2 *tri_ids;         // global triangle id -> (instance, primitive)
    const float4 *vertices;       // 2 float4 per vertex: pos.xyz n.x | n.yz uv
    const uint32_t *indices;
    const uint4 *mesh_descs;
    const uint4 *instance_descs;
    const float4 *transforms;     // 4 float4 rows
    const float4 *inv_transposes; // 4 float4 rows
    // 3 float4: base.rgb flags | emissive.rgb alpha_mode | base_tex emissive_tex normal_tex cutoff
    const float4 *materials;
    const float4 *tri_shading;    // 5 float4 per global triangle (load_triangle)
    uint32_t root;
    float slack;                  // absolute box-test slack (see kPad)
    Light light;
    const SceneRec *scenes;       // one record per scene traced together (>= 1)
    // emissive triangles for light sampling (learned form only; null in batches):
    // 5 float4 each: v0 area | e1 | e2 | unit normal | emission
    const float4 *emitters;
    const float *emitter_cdf;     // cumulative area fraction, last entry 1
    uint32_t emitter_count;
    float emitter_area;           // total emissive area
    // textures (null without; see sample_texture)
    const float4 *tri_uv;         // 2 float4 per global triangle: uv0 uv1 | uv2 0 0
    const uint4 *textures;        // first texel, width, height, flags (bit 0: sRGB colour)
    const uint32_t *texels;       // RGBA8, R in the low byte
    const float *texel_lut;       // [0, 256): c / 255; [256, 512): sRGB decode of c / 255
    const uint32_t *frame_index;  // frame index of the stochastic alpha test (device memory)
    // 1: some material is textured or alpha tested.  Selects the kernels'
    // MAT variants; the others never read a texture or test an alpha, so a
    // scene without such materials keeps their registers and occupancy.
    uint32_t mat_features;
};

struct Ray {
    f3 o, d;
    float tmin, tmax;
    int kx, ky, kz;
    float Sx, Sy, Sz;
    f3 inv;
};

__device__ __forceinline__ void ray_setup(Ray &r) {
    float ax = fabsf(r.d.x), ay = fabsf(r.d.y), az = fabsf(r.d.z);
    int kz = 0;
    float am = ax;
    if (ay > am) { kz = 1; am = ay; }
    if (az > am) { kz = 2; }
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dz = comp(r.d, kz);
    if (dz < 0.0f) { int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = comp(r.d, kx) / dz;
    r.Sy = comp(r.d, ky) / dz;
    r.Sz = 1.0f / dz;
    float dx = fabsf(r.d.x) < 1e-30f ? copysignf(1e-30f, r.d.x) : r.d.x;
    float dy = fabsf(r.d.y) < 1e-30f ? copysignf(1e-30f, r.d.y) : r.d.y;
    float dzz = fabsf(r.d.z) < 1e-30f ? copysignf(1e-30f, r.d.z) : r.d.z;
    r.inv = mk(1.0f / dx, 1.0f / dy, 1.0f / dzz);
}

struct Hit {
    float t, b1, b2;
    uint32_t id;  // global triangle id, kInvalid on a miss
};

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013).
__device__ __forceinline__ bool tri_intersect(const Ray &r, f3 v0, f3 v1, f3 v2, float &t_out, float &b1_out,
                                              float &b2_out) {
    f3 A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
    float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
    float Ax = comp(A, r.kx) - r.Sx * Akz;
    float Ay = comp(A, r.ky) - r.Sy * Akz;
    float Bx = comp(B, r.kx) - r.Sx * Bkz;
    float By = comp(B, r.ky) - r.Sy * Bkz;
    float Cx = comp(C, r.kx) - r.Sx * Ckz;
    float Cy = comp(C, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= r.tmin && t <= r.tmax)) return false;
    t_out = t;
    b1_out = V * rcp;
    b2_out = W * rcp;
    return true;
}

struct TraversalStats {
    uint32_t nodes, tris;
};

// ---------------------------------------------------------------------------
// textures and the alpha test (scene.slang:124-149, 313-368; common.slang:516-577)
// ---------------------------------------------------------------------------
// One axis of a bilinear footprint under wrap addressing: the two texels it
// straddles and the weight of the second.  u is wrapped to [0, 1] first
// (u - floor(u)), so x = u * n - 0.5 lies in [-0.5, n - 0.5] and no index
// leaves the texture; a non-finite u reads as 0.
__device__ __forceinline__ void wrap_axis(float u, uint32_t n, uint32_t &i0, uint32_t &i1, float &a) {
    if (!(fabsf(u) <= 0x1.fffffep+127f)) u = 0.0f;
    float f = u - floorf(u);
    float x = f * (float)n - 0.5f;
    float fl = floorf(x);
    a = x - fl;
    int k = (int)fl;
    i0 = k < 0 ? n - 1u : (uint32_t)k;
    i1 = i0 + 1u == n ? 0u : i0 + 1u;
}

__device__ __forceinline__ float4 texel_value(const float *lut, const float *colour_lut, uint32_t w) {
    return make_float4(colour_lut[w & 0xFFu], colour_lut[(w >> 8) & 0xFFu], colour_lut[(w >> 16) & 0xFFu],
                       lut[w >> 24]);
}

// SampleLevel(linear_sampler, uv, 0) of an RGBA8 texture: the four texels
// decoded (c / 255, RGB through the sRGB curve for colour textures) and
// blended along x, then y, one FMA per lerp.
__device__ __forceinline__ float4 sample_texture(const DevScene &S, uint32_t tex, float u, float v) {
    const uint4 d = S.textures[tex];
    const float *clut = S.texel_lut + ((d.w & 1u) ? 256 : 0);
    uint32_t x0, x1, y0, y1;
    float ax, ay;
    wrap_axis(u, d.y, x0, x1, ax);
    wrap_axis(v, d.z, y0, y1, ay);
    const uint32_t *row0 = S.texels + d.x + (size_t)y0 * d.y;
    const uint32_t *row1 = S.texels + d.x + (size_t)y1 * d.y;
    const float4 t00 = texel_value(S.texel_lut, clut, row0[x0]), t10 = texel_value(S.texel_lut, clut, row0[x1]);
    const float4 t01 = texel_value(S.texel_lut, clut, row1[x0]), t11 = texel_value(S.texel_lut, clut, row1[x1]);
    auto lerp2 = [&](float p00, float p10, float p01, float p11) {
        float r0 = __builtin_fmaf(ax, p10 - p00, p00);
        float r1 = __builtin_fmaf(ax, p11 - p01, p01);
        return __builtin_fmaf(ay, r1 - r0, r0);
    };
    return make_float4(lerp2(t00.x, t10.x, t01.x, t11.x), lerp2(t00.y, t10.y, t01.y, t11.y),
                       lerp2(t00.z, t10.z, t01.z, t11.z), lerp2(t00.w, t10.w, t01.w, t11.w));
}

// float -> int32 by truncation, saturating, NaN -> 0 (the GPU's conversion,
// spelled out so that no case depends on it)
__device__ __forceinline__ int32_t to_i32_sat(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return -2147483647 - 1;
    return (int32_t)x;
}

// Hash of the 1e-4-unit voxel holding p (position_hashing(p, 10000),
// common.slang:547-564).
__device__ __forceinline__ uint32_t position_hash(f3 p) {
    uint32_t vx = ((uint32_t)to_i32_sat(p.x * 10000.0f) + 0x3ffffu) & 0x7ffffu;
    uint32_t vy = ((uint32_t)to_i32_sat(p.y * 10000.0f) + 0x3ffffu) & 0x7ffffu;
    uint32_t vz = ((uint32_t)to_i32_sat(p.z * 10000.0f) + 0x3ffffu) & 0x7ffffu;
    uint32_t k0 = ((vx & 0x7fffu) << 7) | (vy & 0x7fu);
    uint32_t k1 = ((vy & 0xfffu) << 20) | ((vz & 0x7fffu) << 1);
    return murmur_key_mix(k0, k1);
}

// Alpha test of a candidate hit on a non-opaque triangle (eval_alpha_test,
// scene.slang:124-149, as Scene.intersect applies it, 1025-1060): mask
// keeps opacity >= cutoff; blend keeps opacity >= u, u from the hash of the
// world position's voxel and the frame index.  Opacity is the base-colour
// texture's alpha at the hit's uv, 1 without one.
__device__ __forceinline__ bool alpha_accept(const DevScene &S, uint32_t tri, float b1, float b2) {
    const float4 *rec = S.tri_shading + 5 * (size_t)tri;
    const float4 r4 = rec[4];
    const uint32_t mat = __float_as_uint(r4.w) & 0x7FFFFFFFu;
    const float4 m1 = S.materials[3 * (size_t)mat + 1], m2 = S.materials[3 * (size_t)mat + 2];
    const uint32_t mode = __float_as_uint(m1.w);
    if (mode == 0u) return true;
    const float w0 = 1.0f - b1 - b2;
    float opacity = 1.0f;
    const uint32_t base_tex = __float_as_uint(m2.x);
    if (base_tex != kInvalid) {
        const float4 *uvp = S.tri_uv + 2 * (size_t)tri;
        const float4 u01 = uvp[0], u2 = uvp[1];
        float u = w0 * u01.x + b1 * u01.z + b2 * u2.x;
        float v = w0 * u01.y + b1 * u01.w + b2 * u2.y;
        opacity = sample_texture(S, base_tex, u, v).w;
    }
    if (mode == 1u) return opacity >= m2.w;
    const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2];
    const f3 p = mk(w0 * r0.x + b1 * r1.x + b2 * r2.x, w0 * r0.y + b1 * r1.y + b2 * r2.y,
                    w0 * r0.z + b1 * r1.z + b2 * r2.z);
    const float4 *M = S.transforms + 4 * (size_t)__float_as_uint(r4.z);
    const float4 x0 = M[0], x1 = M[1], x2 = M[2];
    const f3 wp = mk(x0.x * p.x + x0.y * p.y + x0.z * p.z + x0.w, x1.x * p.x + x1.y * p.y + x1.z * p.z + x1.w,
                     x2.x * p.x + x2.y * p.y + x2.z * p.z + x2.w);
    const uint32_t seed = murmur_key_mix(position_hash(wp), *S.frame_index);
    return opacity >= (float)seed * 0x1p-32f;
}

// The children of a 4-wide node a ray enters: conservative slab tests on the
// quantized boxes (decode origin + q * 2^e, one rounding), culled beyond
// `limit`; returns how many survive, in e0..e3 nearest first (also for any
// hit: a sun ray's occluders tend to be near, measured 1% faster than node
// order).  The visiting order never changes a result.
template <bool ANY_HIT>
__device__ __forceinline__ int node_children(const uint4 &w0, const uint4 &w1, const uint4 &w2, const uint4 &w3,
                                             const Ray &r, float limit, float slack, uint32_t &e0, uint32_t &e1,
                                             uint32_t &e2, uint32_t &e3) {
    f3 origin = mk(__uint_as_float(w0.x), __uint_as_float(w0.y), __uint_as_float(w0.z));
    uint32_t valid = w0.w >> 24;
    uint32_t qlx = w1.x, qly = w1.y, qlz = w1.z, qhx = w1.w, qhy = w2.x, qhz = w2.y;
    // slab distance of plane q: (origin + q*s - o) * inv = q * (s*inv) + (origin - o) * inv,
    // one FMA per plane.  s = 2^e (e: the signed exponent byte), so s*inv =
    // ldexp(inv, e) is exact; the rest rounds (origin - o), the product b and
    // the FMA: |error| <= 2eps|b| + eps|t|.  The eps|t| part is relative and
    // inside kPad.  The 2eps|b| part is not: |b| = |(origin - o) * inv| grows as
    // 1/|d| along an axis the ray nearly follows (a ray straight down from just
    // above a vertex has |b| ~ 1e9 and t errors of tens of units on that axis,
    // where the true slab interval is all or nothing).  So each axis's entry
    // planes start from b - 2^-21|b| and its exit planes from b + 2^-21|b|
    // (one FMA each, |b| a free source modifier): 8eps|b| covers the 2eps|b|
    // and the FMA's own rounding, so the entry never exceeds and the exit
    // never falls short of the exact slab distance by more than eps|t|.
    const float ax = __builtin_ldexpf(r.inv.x, (int)(int8_t)(w0.w & 0xFFu)),
                ay = __builtin_ldexpf(r.inv.y, (int)(int8_t)((w0.w >> 8) & 0xFFu)),
                az = __builtin_ldexpf(r.inv.z, (int)(int8_t)((w0.w >> 16) & 0xFFu));
    const float bx = (origin.x - r.o.x) * r.inv.x, by = (origin.y - r.o.y) * r.inv.y,
                bz = (origin.z - r.o.z) * r.inv.z;
    const float bnx = __builtin_fmaf(fabsf(bx), -0x1p-21f, bx), bfx = __builtin_fmaf(fabsf(bx), 0x1p-21f, bx);
    const float bny = __builtin_fmaf(fabsf(by), -0x1p-21f, by), bfy = __builtin_fmaf(fabsf(by), 0x1p-21f, by);
    const float bnz = __builtin_fmaf(fabsf(bz), -0x1p-21f, bz), bfz = __builtin_fmaf(fabsf(bz), 0x1p-21f, bz);
    // the entry plane of each axis follows the sign of the direction (s*inv has
    // the sign of inv and the FMA is monotone in q), so the near and far
    // planes of all four children are picked once per node instead of
    // ordering every pair: the same values as min/max of the two planes
    const uint32_t nqx = r.inv.x >= 0.0f ? qlx : qhx, fqx = r.inv.x >= 0.0f ? qhx : qlx;
    const uint32_t nqy = r.inv.y >= 0.0f ? qly : qhy, fqy = r.inv.y >= 0.0f ? qhy : qly;
    const uint32_t nqz = r.inv.z >= 0.0f ? qlz : qhz, fqz = r.inv.z >= 0.0f ? qhz : qlz;
    float tn[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float tnx = __builtin_fmaf((float)((nqx >> (8 * c)) & 0xFFu), ax, bnx);
        float tfx = __builtin_fmaf((float)((fqx >> (8 * c)) & 0xFFu), ax, bfx);
        float tny = __builtin_fmaf((float)((nqy >> (8 * c)) & 0xFFu), ay, bny);
        float tfy = __builtin_fmaf((float)((fqy >> (8 * c)) & 0xFFu), ay, bfy);
        float tnz = __builtin_fmaf((float)((nqz >> (8 * c)) & 0xFFu), az, bnz);
        float tfz = __builtin_fmaf((float)((fqz >> (8 * c)) & 0xFFu), az, bfz);
        float near = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, r.tmin));
        float far = __builtin_fmaf(fminf(fminf(tfx, tfy), tfz), kPad, slack);
        far = fminf(far, limit);
        tn[c] = (((valid >> c) & 1u) && near <= far) ? near : INFINITY;  // +inf: culled
    }
    e0 = w2.z;
    e1 = w2.w;
    e2 = w3.x;
    e3 = w3.y;
    float k0 = tn[0], k1 = tn[1], k2 = tn[2], k3 = tn[3];
    int n = (k0 < INFINITY) + (k1 < INFINITY) + (k2 < INFINITY) + (k3 < INFINITY);
    // a 4-input sorting network on the entry distance; culled children go last
    auto cswap = [](float &ka, uint32_t &ea, float &kb, uint32_t &eb) {
        bool sw = kb < ka;
        float k = sw ? kb : ka;
        uint32_t e = sw ? eb : ea;
        kb = sw ? ka : kb;
        eb = sw ? ea : eb;
        ka = k;
        ea = e;
    };
    cswap(k0, e0, k1, e1);
    cswap(k2, e2, k3, e3);
    cswap(k0, e0, k2, e2);
    cswap(k1, e1, k3, e3);
    cswap(k1, e1, k2, e2);
    return n;
}

// Stack of child entries in LDS: STACK rows of BLOCK lanes, lane-major so
// the 64 lanes of a wave hit 64 different banks.
template <int STACK, int BLOCK, bool ANY_HIT, bool COUNT, bool MAT>
__device__ __forceinline__ bool traverse(const DevScene &S, const Ray &r, Hit &hit, uint32_t *lds_col,
                                         TraversalStats &st) {
    uint32_t cur = S.root;
    int sp = 0;
    while (true) {
        if (!(cur & kLeafBit)) {
            const uint4 *np = S.nodes + 4 * (size_t)cur;
            uint4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3];
            if (COUNT) st.nodes++;
            uint32_t e0, e1, e2, e3;
            int nhit = node_children<ANY_HIT>(w0, w1, w2, w3, r, __builtin_fmaf(ANY_HIT ? r.tmax : hit.t, kPad, S.slack), S.slack,
                                              e0, e1, e2, e3);
            if (nhit == 0) {
                if (sp == 0) break;
                cur = lds_col[(--sp) * BLOCK];
                continue;
            }
            // push the far children (farthest first), continue with the nearest
            if (nhit > 3) lds_col[(sp++) * BLOCK] = e3;
            if (nhit > 2) lds_col[(sp++) * BLOCK] = e2;
            if (nhit > 1) lds_col[(sp++) * BLOCK] = e1;
            cur = e0;
            continue;
        }
        // leaf
        uint32_t first = cur & kLeafFirstMask;
        uint32_t count = ((cur >> kLeafCountShift) & 0xFu) + 1u;
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
        if (sp == 0) break;
        cur = lds_col[(--sp) * BLOCK];
    }
    return hit.id != kInvalid;
}

// ---------------------------------------------------------------------------
// shading (scene.slang:178-406, common.slang:54-162, 452-490)
// ---------------------------------------------------------------------------
struct Frame {
    f3 n, t, b;
};
__device__ __forceinline__ Frame make_frame(f3 n) {
    Frame f;
    f.n = n;
    f.t = normalize(cross(n, mk(-n.z, n.x, -n.y)));
    f.b = cross(n, f.t);
    return f;
}
__device__ __forceinline__ f3 to_global(const Frame &f, f3 w) { return f.t * w.x + f.b * w.y + f.n * w.z; }

struct Shading {
    f3 origin_front;
    Frame frame;
    f3 diffuse, emissive;
};

__device__ __forceinline__ f3 row3(float4 r) { return mk(r.x, r.y, r.z); }
__device__ __forceinline__ f3 mul33(const float4 *m, f3 v) {
    return mk(dot(row3(m[0]), v), dot(row3(m[1]), v), dot(row3(m[2]), v));
}

// Per-triangle shading record (5 float4 = 80 B, global triangle order):
//   v0.xyz n0.x | v1.xyz n1.x | v2.xyz n2.x | n0.yz n1.yz | n2.yz transform_id material_id|odd<<31
// holding the object-space vertices and normals the reference gathers through
// instance -> mesh -> index -> vertex (scene.slang:271-300), copied bit for bit.
__device__ __forceinline__ void load_triangle(const DevScene &S, uint32_t tri, f3 &v0, f3 &v1, f3 &v2, f3 (&n)[3],
                                              uint32_t &xf, uint32_t &mat, bool &odd) {
    const float4 *rec = S.tri_shading + 5 * (size_t)tri;
    float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3], r4 = rec[4];
    v0 = mk(r0.x, r0.y, r0.z);
    v1 = mk(r1.x, r1.y, r1.z);
    v2 = mk(r2.x, r2.y, r2.z);
    n[0] = mk(r0.w, r3.x, r3.y);
    n[1] = mk(r1.w, r3.z, r3.w);
    n[2] = mk(r2.w, r4.x, r4.y);
    xf = __float_as_uint(r4.z);
    uint32_t m = __float_as_uint(r4.w);
    mat = m & 0x7FFFFFFFu;
    odd = (m >> 31) != 0u;
}

template <bool MAT>
__device__ __forceinline__ void shading_data(const DevScene &S, uint32_t tri, float b1, float b2, f3 view,
                                             Shading &sd) {
    f3 v0, v1, v2, nrm[3];
    uint32_t xf, mat;
    bool odd;
    load_triangle(S, tri, v0, v1, v2, nrm, xf, mat, odd);
    const float4 *M = S.transforms + 4 * (size_t)xf;
    const float4 *IT = S.inv_transposes + 4 * (size_t)xf;
    float4 m0 = M[0], m1 = M[1], m2 = M[2];
    float4 it0 = IT[0], it1 = IT[1], it2 = IT[2], it3 = IT[3];
    f3 bc = mk(1.0f - b1 - b2, b1, b2);

    const float c0 = 5.9604644775390625e-8f, c1 = 1.7881397695873602e-7f, c2 = 1.1920931797249068e-7f;
    f3 e1 = v1 - v0, e2 = v2 - v0;
    f3 op = mk(v0.x + __builtin_fmaf(b1, e1.x, b2 * e2.x), v0.y + __builtin_fmaf(b1, e1.y, b2 * e2.y),
               v0.z + __builtin_fmaf(b1, e1.z, b2 * e2.z));
    f3 on = cross(e1, e2);
    f3 wp = mk(m0.w + __builtin_fmaf(m0.x, op.x, __builtin_fmaf(m0.y, op.y, m0.z * op.z)),
               m1.w + __builtin_fmaf(m1.x, op.x, __builtin_fmaf(m1.y, op.y, m1.z * op.z)),
               m2.w + __builtin_fmaf(m2.x, op.x, __builtin_fmaf(m2.y, op.y, m2.z * op.z)));
    f3 wnu = mk(dot(row3(it0), on), dot(row3(it1), on), dot(row3(it2), on));
    float world_scale = 1.0f / sqrtf(fmaxf(dot(wnu, wnu), 1e-30f));
    f3 world_normal = wnu * world_scale;
    f3 ae1 = fabs3(e1), ae2 = fabs3(e2);
    f3 ext3 = (ae1 + ae2) + fabs3(ae1 - ae2);
    float extent = fmaxf(fmaxf(ext3.x, ext3.y), ext3.z);
    f3 object_error = fabs3(v0) * c0 + mk(c1 * extent, c1 * extent, c1 * extent);
    f3 aop = fabs3(op);
    f3 world_error = mk(c1 * dot(fabs3(row3(m0)), aop) + c2 * fabsf(m0.w),
                        c1 * dot(fabs3(row3(m1)), aop) + c2 * fabsf(m1.w),
                        c1 * dot(fabs3(row3(m2)), aop) + c2 * fabsf(m2.w));
    f3 awp = fabs3(wp);
    f3 te = mk(fabsf(it0.x) * awp.x + fabsf(it1.x) * awp.y + fabsf(it2.x) * awp.z + fabsf(it3.x) * 1.0f,
               fabsf(it0.y) * awp.x + fabsf(it1.y) * awp.y + fabsf(it2.y) * awp.z + fabsf(it3.y) * 1.0f,
               fabsf(it0.z) * awp.x + fabsf(it1.z) * awp.y + fabsf(it2.z) * awp.z + fabsf(it3.z) * 1.0f);
    object_error = object_error + te * c2;
    float object_offset = dot(object_error, fabs3(on));
    float world_offset = dot(world_error, fabs3(world_normal));
    world_offset = __builtin_fmaf(world_scale, object_offset, world_offset);

    f3 ni = nrm[0] * bc.x + nrm[1] * bc.y + nrm[2] * bc.z;
    f3 geometry_normal = normalize(mk(dot(row3(it0), ni), dot(row3(it1), ni), dot(row3(it2), ni)));
    f3 face_normal = world_normal;
    const float4 *mp = S.materials + 3 * (size_t)mat;
    float4 ma = mp[0], mb = mp[1], mc = mp[2];
    f3 diffuse = mk(ma.x, ma.y, ma.z), emissive = mk(mb.x, mb.y, mb.z);
    f3 shading_normal = geometry_normal;
    const uint32_t base_tex = __float_as_uint(mc.x), emissive_tex = __float_as_uint(mc.y),
                   normal_tex = __float_as_uint(mc.z);
    if (MAT && (base_tex & emissive_tex & normal_tex) != kInvalid) {
        // textures at the hit's uv (scene.slang:297, 313-368): base colour
        // replaces the factor, emission scales it, the normal map tilts the
        // interpolated normal in a tangent frame from object-space dp/duv
        const float4 *uvp = S.tri_uv + 2 * (size_t)tri;
        const float4 u01 = uvp[0], u2 = uvp[1];
        float u = bc.x * u01.x + bc.y * u01.z + bc.z * u2.x;
        float v = bc.x * u01.y + bc.y * u01.w + bc.z * u2.y;
        if (base_tex != kInvalid) {
            float4 t = sample_texture(S, base_tex, u, v);
            diffuse = mk(t.x, t.y, t.z);
        }
        if (emissive_tex != kInvalid) {
            float4 t = sample_texture(S, emissive_tex, u, v);
            emissive = mk(t.x * mb.x, t.y * mb.y, t.z * mb.z);
        }
        if (normal_tex != kInvalid) {
            float4 t = sample_texture(S, normal_tex, u, v);
            f3 tn = normalize(mk(t.x * 2.0f - 1.0f, t.y * 2.0f - 1.0f, t.z * 2.0f - 1.0f));
            float d1u = u01.z - u01.x, d1v = u01.w - u01.y, d2u = u2.x - u01.x, d2v = u2.y - u01.y;
            float det = d1u * d2v - d1v * d2u;
            if (fabsf(det) > 1e-6f) {
                f3 tg = e1 * d2v - e2 * d1v;
                tg = normalize(mk(tg.x / det, tg.y / det, tg.z / det));
                tg = mk(dot(row3(it0), tg), dot(row3(it1), tg), dot(row3(it2), tg));
                tg = normalize(tg - geometry_normal * dot(tg, geometry_normal));
                f3 bt = cross(geometry_normal, tg);
                shading_normal = normalize(tg * tn.x + bt * tn.y + geometry_normal * tn.z);
            }
        }
    }
    bool double_sided = (__float_as_uint(ma.w) & 1u) != 0;
    bool front = dot(face_normal, view) > 0.0f;
    float sign = (double_sided && !front) ? -1.0f : 1.0f;
    face_normal = face_normal * sign;
    shading_normal = shading_normal * sign;
    f3 D = -view;
    float dn = 2.0f * dot(shading_normal, D);
    f3 R = D - shading_normal * dn;
    float k = dot(R, face_normal);
    if (k < 0.0f) shading_normal = normalize(normalize(R - face_normal * k) - D);
    sd.origin_front = wp + face_normal * world_offset;
    sd.frame = make_frame(shading_normal);
    sd.diffuse = diffuse;
    sd.emissive = emissive;
}

// texture_space_path_tracer.slang:39-55
__device__ __forceinline__ f3 world_flat_normal(const DevScene &S, uint32_t tri) {
    f3 p0, p1, p2, nrm[3];
    uint32_t xf, mat;
    bool odd;
    load_triangle(S, tri, p0, p1, p2, nrm, xf, mat, odd);
    f3 n = normalize(cross(p1 - p0, p2 - p0));
    if (odd) n = -n;
    const float4 *IT = S.inv_transposes + 4 * (size_t)xf;
    float4 it0 = IT[0], it1 = IT[1], it2 = IT[2];
    return normalize(mk(dot(row3(it0), n), dot(row3(it1), n), dot(row3(it2), n)));
}

__device__ __forceinline__ void cosine_sample(float ex, float ey, f3 &dir, float &pdf) {
    float s, c;
    sincos_2pi(ex, &s, &c);
    float ct = sqrtf(ey);
    float st = sqrtf(1.0f - ct * ct);
    dir = mk(st * c, st * s, ct);
    pdf = ct * kInvPi;
}

__device__ __forceinline__ f3 cone_sample(f3 axis, float cos_half, float ux, float uy) {
    float ct = 1.0f - ux * (1.0f - cos_half);
    float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float s, c;
    sincos_2pi(uy, &s, &c);
    f3 local = mk(c * st, s * st, ct);
    Frame f = make_frame(axis);
    return to_global(f, local);
}

// ---------------------------------------------------------------------------
// Mesh Colors lattice (mesh_colors.slang:37-65, 255-363)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t patch_count(uint32_t r) { return (r + 1) * (r + 2) / 2; }

__device__ __forceinline__ uint32_t lattice_address(uint32_t i, uint32_t j, uint32_t texel_offset, uint2 face) {
    uint32_t row = i * (face.y + 1) - i * (i - 1) / 2;
    return texel_offset + face.x + row + j;
}

__device__ __forceinline__ void address_to_ij(uint32_t address, uint32_t resolution, uint32_t &i, uint32_t &j) {
    uint32_t left = 0, right = resolution;
    while (left < right) {
        uint32_t middle = (left + right + 1) / 2;
        uint32_t ma = middle * (resolution + 1) - middle * (middle - 1) / 2;
        if (ma <= address) left = middle; else right = middle - 1;
    }
    i = left;
    j = address - (left * (resolution + 1) - left * (left - 1) / 2);
}

__device__ __forceinline__ uint32_t nearest_texel(f3 b, uint32_t texel_offset, uint2 face) {
    uint32_t R = face.y;
    float fr = (float)R;
    f3 sc = mk(saturate(b.x) * fr, saturate(b.y) * fr, saturate(b.z) * fr);
    uint32_t i = (uint32_t)sc.x, j = (uint32_t)sc.y;
    i = i > R ? R : i;
    j = j > R - i ? R - i : j;
    f3 w = mk(fracf(sc.x), fracf(sc.y), fracf(sc.z));
    bool upper = fabsf(w.x + w.y + w.z - 2.0f) <= 1e-3f;
    if (upper) w = mk(1.0f - w.x, 1.0f - w.y, 1.0f - w.z);
    if (w.x + w.y + w.z <= 1e-6f || i >= R || j >= R - i) return lattice_address(i, j, texel_offset, face);
    float largest = fmaxf(w.x, fmaxf(w.y, w.z));
    if (upper) {
        if (fabsf(largest - w.x) < 1e-3f) return lattice_address(i, j + 1, texel_offset, face);
        if (fabsf(largest - w.y) < 1e-3f) return lattice_address(i + 1, j, texel_offset, face);
        return lattice_address(i + 1, j + 1, texel_offset, face);
    }
    if (fabsf(largest - w.x) < 1e-3f) return lattice_address(i + 1, j, texel_offset, face);
    if (fabsf(largest - w.y) < 1e-3f) return lattice_address(i, j + 1, texel_offset, face);
    return lattice_address(i, j, texel_offset, face);
}

// ---------------------------------------------------------------------------
// RGB9E5 (mesh_colors.slang:218-253), exact to the float64 restatement of
// test_mesh_colors_rgb9e5.py:15-50 for every float32 input.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t round_half_up(float x) {
    float f = floorf(x);
    return (uint32_t)f + ((x - f) >= 0.5f ? 1u : 0u);
}

__device__ __forceinline__ uint32_t pack_rgb9e5(float r, float g, float b) {
    r = r != r ? 0.0f : fminf(fmaxf(r, 0.0f), 65408.0f);
    g = g != g ? 0.0f : fminf(fmaxf(g, 0.0f), 65408.0f);
    b = b != b ? 0.0f : fminf(fmaxf(b, 0.0f), 65408.0f);
    float m = fmaxf(r, fmaxf(g, b));
    if (m < 0x1p-25f) return 0u;
    int e = (int)((__float_as_uint(m) >> 23) & 0xFFu) - 127 + 1;
    e = e < -15 ? -15 : e;
    uint32_t shared = (uint32_t)(e + 15);
    float scale = exp2i((int)shared - 24);
    uint32_t mr = round_half_up(r / scale), mg = round_half_up(g / scale), mbv = round_half_up(b / scale);
    uint32_t mm = max(mr, max(mg, mbv));
    if (mm == 512u && shared < 31u) {
        shared += 1u;
        scale *= 2.0f;
        mr = round_half_up(r / scale);
        mg = round_half_up(g / scale);
        mbv = round_half_up(b / scale);
    }
    mr = min(mr, 511u);
    mg = min(mg, 511u);
    mbv = min(mbv, 511u);
    return mr | (mg << 9) | (mbv << 18) | (shared << 27);
}

__device__ __forceinline__ f3 unpack_rgb9e5(uint32_t p) {
    float scale = exp2i((int)(p >> 27) - 24);
    return mk((float)(p & 0x1ffu) * scale, (float)((p >> 9) & 0x1ffu) * scale, (float)((p >> 18) & 0x1ffu) * scale);
}

// Slot sampling shared by the resolve and surface-filter kernels.
__device__ __forceinline__ uint32_t side_texel(uint32_t front, uint32_t texel_offset, uint2 side, bool back) {
    return back ? side.x + (front - texel_offset) : front;
}

template <bool FROZEN>
__device__ __forceinline__ f3 load_cache(const void *cache, uint32_t slot) {
    if (FROZEN) return unpack_rgb9e5(static_cast<const uint32_t *>(cache)[slot]);
    float4 p = static_cast<const float4 *>(cache)[slot];
    return mk(p.x, p.y, p.z);
}

// mesh_colors_sample_irradiance (mesh_colors.slang:365-426)
template <bool FROZEN>
__device__ f3 sample_irradiance(const void *cache, f3 b, uint32_t texel_offset, uint2 face, uint2 side, bool back) {
    uint32_t R = face.y;
    float fr = (float)R;
    f3 sc = mk(saturate(b.x) * fr, saturate(b.y) * fr, saturate(b.z) * fr);
    uint32_t i = (uint32_t)sc.x, j = (uint32_t)sc.y;
    i = i > R ? R : i;
    j = j > R - i ? R - i : j;
    f3 w = mk(fracf(sc.x), fracf(sc.y), fracf(sc.z));
    bool upper = fabsf(w.x + w.y + w.z - 2.0f) <= 1e-3f;
    if (upper) w = mk(1.0f - w.x, 1.0f - w.y, 1.0f - w.z);
    if (w.x + w.y + w.z <= 1e-6f || i >= R || j >= R - i)
        return load_cache<FROZEN>(cache, side_texel(lattice_address(i, j, texel_offset, face), texel_offset, side, back));
    uint32_t ta, tb, tc;
    if (upper) {
        ta = lattice_address(i, j + 1, texel_offset, face);
        tb = lattice_address(i + 1, j, texel_offset, face);
        tc = lattice_address(i + 1, j + 1, texel_offset, face);
    } else {
        ta = lattice_address(i + 1, j, texel_offset, face);
        tb = lattice_address(i, j + 1, texel_offset, face);
        tc = lattice_address(i, j, texel_offset, face);
    }
    f3 A = load_cache<FROZEN>(cache, side_texel(ta, texel_offset, side, back));
    f3 B = load_cache<FROZEN>(cache, side_texel(tb, texel_offset, side, back));
    f3 C = load_cache<FROZEN>(cache, side_texel(tc, texel_offset, side, back));
    return A * w.x + B * w.y + C * w.z;
}


}  // namespace nan
