/*
 * CPU oracle for the texture-space (Mesh Colors) path tracer.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates the reference algorithm in
 * plain C so that the HIP product can be checked against it.  Only tests,
 * the smoke check and the CPU-baseline leg of the benchmark load it; the
 * product never links, calls or falls back to it.
 *
 * Every function cites the reference source it follows (paths relative to
 * the reference root).  The arithmetic follows one written-down FP32
 * specification shared with the product: no implicit contraction (compile
 * with -ffp-contract=off), explicit fmaf() exactly where the reference
 * writes mad(), correctly rounded sqrtf and division, and sin/cos of
 * 2*pi*u from the range reduction and Taylor polynomials in sincos_2pi()
 * (the reference leaves sin/cos to the shader compiler, so its own values
 * are implementation defined).  Under that specification the oracle and the
 * product give bit-identical payloads.
 *
 * What pins each part (beyond agreement with the product):
 *   - layout, lattice, find_face, nearest_texel, RNG: closed forms and the
 *     reference's known answers (tests/test_oracle_*.py);
 *   - RGB9E5: the reference's float64 restatement
 *     (test_mesh_colors_rgb9e5.py:15-50) over exhaustive channel sweeps;
 *   - ray/triangle: an FP64 brute-force nearest hit (orc_intersect_fp64)
 *     and tinybvh's occlusion query built from the reference tree;
 *   - shading/paths: white-furnace and closed-form emitter invariants;
 *   - Russian roulette (an extension): the emitting closed box, whose
 *     samples are pi * Le * (integer hit count) with a geometric tail and
 *     the no-roulette mean (tests/test_roulette_oracle.py).
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* constants                                                           */
/* ------------------------------------------------------------------ */
#define ORC_INVALID 0xFFFFFFFFu
#define ORC_MAX_SAMPLES 65535u
static const float ORC_PI = 0x1.921fb6p+1f;
static const float ORC_TWO_PI = 0x1.921fb6p+2f;
static const float ORC_INV_PI = 0x1.45f306p-2f;

/* ------------------------------------------------------------------ */
/* small float3 helpers, no contraction                                */
/* ------------------------------------------------------------------ */
typedef struct { float x, y, z; } f3;
static inline f3 mk(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 mul(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline f3 muls(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline f3 neg(f3 a) { return mk(-a.x, -a.y, -a.z); }
static inline f3 fabs3(f3 a) { return mk(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
static inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline f3 normalize(f3 a) {
    float inv = 1.0f / sqrtf(dot(a, a));
    return muls(a, inv);
}
static inline float comp(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

/* sin(2*pi*u), cos(2*pi*u) for u in [0, 1]: exact quarter-turn reduction
 * on u, then degree-9/10 Taylor polynomials on [-pi/4, pi/4]. */
static void sincos_2pi(float u, float *s_out, float *c_out) {
    float k = rintf(u * 4.0f);
    float v = u - k * 0.25f;
    float a = v * ORC_TWO_PI;
    float a2 = a * a;
    float p = fmaf(a2, 0x1.71de3ap-19f, -0x1.a01a02p-13f);
    p = fmaf(a2, p, 0x1.111112p-7f);
    p = fmaf(a2, p, -0x1.555556p-3f);
    float s = fmaf(a * a2, p, a);
    float q = fmaf(a2, -0x1.27e4fcp-22f, 0x1.a01a02p-16f);
    q = fmaf(a2, q, -0x1.6c16c2p-10f);
    q = fmaf(a2, q, 0x1.555556p-5f);
    q = fmaf(a2, q, -0.5f);
    float c = fmaf(a2, q, 1.0f);
    int quadrant = ((int)k) & 3;
    switch (quadrant) {
    case 0: *s_out = s; *c_out = c; break;
    case 1: *s_out = c; *c_out = -s; break;
    case 2: *s_out = -s; *c_out = -c; break;
    default: *s_out = -c; *c_out = s; break;
    }
}

/* ------------------------------------------------------------------ */
/* RNG (common.slang:164-210)                                          */
/* ------------------------------------------------------------------ */
/* pcg == 0: the reference's LCG.  pcg == 1: the production streams, PCG
 * (O'Neill 2014) with the RXS-M-XS output on a 32-bit LCG state, started at
 * murmur_key_mix(murmur_key_mix(x, y), sample_index); no reference
 * counterpart.  Both map a draw r to (r >> 8) * 2^-24. */
typedef struct { uint32_t state, pcg; } orc_rng;
static uint32_t murmur_key_mix(uint32_t h, uint32_t k);
static inline void rng_seed_mode(orc_rng *r, uint32_t x, uint32_t y, uint32_t sample_index, uint32_t pcg) {
    r->pcg = pcg;
    if (pcg) r->state = murmur_key_mix(murmur_key_mix(x, y), sample_index);
    else r->state = (x * 100591u) ^ (y * 102191u) ^ (sample_index * 103643u);
}
static inline uint32_t rng_next(orc_rng *r) {
    if (!r->pcg) {
        r->state = 1664525u * r->state + 1013904223u;
        return r->state;
    }
    uint32_t old = r->state;
    r->state = old * 747796405u + 2891336453u;
    uint32_t word = ((old >> ((old >> 28u) + 4u)) ^ old) * 277803737u;
    return (word >> 22u) ^ word;
}
static inline float rng_1d(orc_rng *r) { return (float)(rng_next(r) >> 8) * 0x1p-24f; }

ORC_API uint32_t orc_rng_stream(uint32_t x, uint32_t y, uint32_t s, uint32_t count, float *out) {
    orc_rng r;
    rng_seed_mode(&r, x, y, s, 0);
    for (uint32_t i = 0; i < count; ++i) out[i] = rng_1d(&r);
    return r.state;
}

/* count raw 32-bit draws of the PCG stream of (x, y, s); returns the state after them */
ORC_API uint32_t orc_pcg_stream(uint32_t x, uint32_t y, uint32_t s, uint32_t count, uint32_t *out) {
    orc_rng r;
    rng_seed_mode(&r, x, y, s, 1);
    for (uint32_t i = 0; i < count; ++i) out[i] = rng_next(&r);
    return r.state;
}

/* ------------------------------------------------------------------ */
/* Mesh Colors lattice (mesh_colors.slang:37-65, 255-363)              */
/* ------------------------------------------------------------------ */
static inline uint32_t patch_count(uint32_t r) { return (r + 1) * (r + 2) / 2; }

static inline uint32_t lattice_address(uint32_t i, uint32_t j, const uint32_t *inst, const uint32_t *face) {
    uint32_t row = i * (face[1] + 1) - i * (i - 1) / 2;
    return inst[2] + face[0] + row + j;
}

ORC_API void orc_address_to_ij(uint32_t address, uint32_t resolution, uint32_t *ij) {
    uint32_t left = 0, right = resolution;
    while (left < right) {
        uint32_t middle = (left + right + 1) / 2;
        uint32_t ma = middle * (resolution + 1) - middle * (middle - 1) / 2;
        if (ma <= address) left = middle; else right = middle - 1;
    }
    uint32_t cumulative = left * (resolution + 1) - left * (left - 1) / 2;
    ij[0] = left;
    ij[1] = address - cumulative;
}

/* binary search over one instance's faces (mesh_colors.slang:272-305) */
static int find_face(const uint32_t *faces, const uint32_t *inst, uint32_t local_texel,
                     const uint32_t **face_out, uint32_t *prim_out) {
    int left = (int)inst[0];
    int right = (int)(inst[0] + inst[1]) - 1;
    while (left <= right) {
        int middle = (left + right) / 2;
        const uint32_t *f = faces + 2 * (size_t)middle;
        int local = (int)local_texel - (int)f[0];
        int count = (int)patch_count(f[1]);
        if (local < 0) right = middle - 1;
        else if (local >= count) left = middle + 1;
        else { *face_out = f; *prim_out = (uint32_t)middle - inst[0]; return 1; }
    }
    return 0;
}

static inline float saturate(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
static inline float fracf(float x) { return x - floorf(x); }

/* mesh_colors_nearest_texel (mesh_colors.slang:325-363) */
static uint32_t nearest_texel(f3 b, const uint32_t *inst, const uint32_t *face) {
    uint32_t R = face[1];
    float fr = (float)R;
    f3 sc = mk(saturate(b.x) * fr, saturate(b.y) * fr, saturate(b.z) * fr);
    uint32_t i = (uint32_t)sc.x, j = (uint32_t)sc.y;
    if (i > R) i = R;
    if (j > R - i) j = R - i;
    f3 w = mk(fracf(sc.x), fracf(sc.y), fracf(sc.z));
    int upper = fabsf(w.x + w.y + w.z - 2.0f) <= 1e-3f;
    if (upper) w = mk(1.0f - w.x, 1.0f - w.y, 1.0f - w.z);
    if (w.x + w.y + w.z <= 1e-6f || i >= R || j >= R - i) return lattice_address(i, j, inst, face);
    float largest = fmaxf(w.x, fmaxf(w.y, w.z));
    if (upper) {
        if (fabsf(largest - w.x) < 1e-3f) return lattice_address(i, j + 1, inst, face);
        if (fabsf(largest - w.y) < 1e-3f) return lattice_address(i + 1, j, inst, face);
        return lattice_address(i + 1, j + 1, inst, face);
    }
    if (fabsf(largest - w.x) < 1e-3f) return lattice_address(i + 1, j, inst, face);
    if (fabsf(largest - w.y) < 1e-3f) return lattice_address(i, j + 1, inst, face);
    return lattice_address(i, j, inst, face);
}

ORC_API uint32_t orc_nearest_texel(const float *bary, const uint32_t *inst, const uint32_t *face) {
    return nearest_texel(mk(bary[0], bary[1], bary[2]), inst, face);
}

/* Texel ray-gen: compute_main prologue (texture_space_path_tracer.slang:142-199).
 * Writes the surface hit (instance, prim, bary.x bits, bary.y bits) and leaves
 * the RNG positioned after the jitter draws. */
typedef struct { uint32_t instance, prim; float b1, b2; } orc_hit_info;

static int texel_raygen(const uint32_t *faces, const uint32_t *instances, uint32_t instance_index,
                        uint32_t local_texel, uint32_t iteration, uint32_t spp, uint32_t pcg,
                        orc_hit_info *hit, orc_rng *rng, const uint32_t **face_out) {
    const uint32_t *inst = instances + 4 * (size_t)instance_index;
    const uint32_t *face;
    uint32_t prim;
    if (!find_face(faces, inst, local_texel, &face, &prim)) return 0;
    uint32_t global_texel = inst[2] + local_texel;
    uint32_t ij[2];
    orc_address_to_ij(local_texel - face[0], face[1], ij);
    float fr = (float)face[1];
    float base_x = (float)ij[0] / fr, base_y = (float)ij[1] / fr;
    rng_seed_mode(rng, global_texel, instance_index, iteration * (spp > 1 ? spp : 1), pcg);
    float bx = base_x, by = base_y;
    for (int attempt = 0; attempt < 32; ++attempt) {
        float ux = rng_1d(rng);
        float uy = rng_1d(rng);
        float cx = base_x + (ux - 0.5f) / fr;
        float cy = base_y + (uy - 0.5f) / fr;
        if (cx < 0.0f || cy < 0.0f || cx + cy > 1.0f) continue;
        f3 cb = mk(cx, cy, 1.0f - cx - cy);
        if (nearest_texel(cb, inst, face) == global_texel) { bx = cx; by = cy; break; }
    }
    float bz = fmaxf(0.0f, 1.0f - bx - by);
    hit->instance = instance_index;
    hit->prim = prim;
    hit->b1 = by;
    hit->b2 = bz;
    if (face_out) *face_out = face;
    return 1;
}

/* Flat front-texel order: instances in order, local texels in order. */
ORC_API int orc_texel_raygen(const uint32_t *faces, const uint32_t *instances, uint32_t instance_count,
                             uint32_t iteration, uint32_t spp, uint32_t *out, uint32_t *out_rng) {
    size_t w = 0;
    for (uint32_t i = 0; i < instance_count; ++i) {
        uint32_t count = instances[4 * (size_t)i + 3];
        for (uint32_t t = 0; t < count; ++t, ++w) {
            orc_hit_info h;
            orc_rng rng;
            if (!texel_raygen(faces, instances, i, t, iteration, spp, 0, &h, &rng, NULL)) return -1;
            out[4 * w + 0] = h.instance;
            out[4 * w + 1] = h.prim;
            memcpy(&out[4 * w + 2], &h.b1, 4);
            memcpy(&out[4 * w + 3], &h.b2, 4);
            if (out_rng) out_rng[w] = rng.state;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* RGB9E5 (mesh_colors.slang:218-253; float64 restatement in          */
/* test_mesh_colors_rgb9e5.py:15-50)                                   */
/* ------------------------------------------------------------------ */
static inline uint32_t round_half_up(float x) {
    /* floor(x + 0.5) without the rounding of the addition */
    float f = floorf(x);
    return (uint32_t)f + ((x - f) >= 0.5f ? 1u : 0u);
}

ORC_API uint32_t orc_pack_rgb9e5(const float *rgb) {
    float v[3];
    for (int k = 0; k < 3; ++k) {
        float x = rgb[k];
        if (isnan(x)) x = 0.0f;
        x = fminf(fmaxf(x, 0.0f), 65408.0f);
        v[k] = x;
    }
    float m = fmaxf(v[0], fmaxf(v[1], v[2]));
    if (m < 0x1p-25f) return 0;
    int e = ilogbf(m) + 1;
    if (e < -15) e = -15;
    uint32_t shared = (uint32_t)(e + 15);
    float scale = ldexpf(1.0f, (int)shared - 24);
    uint32_t mant[3];
    for (int k = 0; k < 3; ++k) mant[k] = round_half_up(v[k] / scale);
    uint32_t mm = mant[0] > mant[1] ? mant[0] : mant[1];
    mm = mm > mant[2] ? mm : mant[2];
    if (mm == 512 && shared < 31) {
        shared += 1;
        scale *= 2.0f;
        for (int k = 0; k < 3; ++k) mant[k] = round_half_up(v[k] / scale);
    }
    for (int k = 0; k < 3; ++k) if (mant[k] > 511) mant[k] = 511;
    return mant[0] | (mant[1] << 9) | (mant[2] << 18) | (shared << 27);
}

ORC_API void orc_unpack_rgb9e5(uint32_t packed, float *rgb) {
    float scale = ldexpf(1.0f, (int)(packed >> 27) - 24);
    rgb[0] = (float)(packed & 0x1ff) * scale;
    rgb[1] = (float)((packed >> 9) & 0x1ff) * scale;
    rgb[2] = (float)((packed >> 18) & 0x1ff) * scale;
}

/* words for (x, 0, 0) where x runs over the float32 bit patterns start..start+count-1 */
ORC_API void orc_pack_rgb9e5_red_range(uint32_t start, uint32_t count, uint32_t *out) {
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t bits = start + i;
        float rgb[3] = {0.0f, 0.0f, 0.0f};
        memcpy(&rgb[0], &bits, 4);
        out[i] = orc_pack_rgb9e5(rgb);
    }
}

ORC_API void orc_pack_rgb9e5_array(const float *payload, uint32_t stride_floats, uint64_t n, uint32_t *out) {
    for (uint64_t i = 0; i < n; ++i) out[i] = orc_pack_rgb9e5(payload + i * stride_floats);
}

/* ------------------------------------------------------------------ */
/* scene                                                               */
/* ------------------------------------------------------------------ */
typedef struct {
    const float *vertices;        /* [V][8] object-space pos, normal, uv */
    uint32_t vertex_count;
    const uint32_t *indices;      /* [T*3] mesh-local indices, meshes concatenated */
    uint32_t index_count;
    const uint32_t *mesh_descs;   /* [M][4] vertex_count, index_count, vertex_offset, index_offset */
    uint32_t mesh_count;
    const uint32_t *instance_descs; /* [I][4] mesh, material, transform, flags */
    uint32_t instance_count;
    const float *transforms;      /* [X][16] row-major object-to-world */
    const float *inv_transposes;  /* [X][16] row-major transpose(inverse(M)) */
    uint32_t transform_count;
    const uint32_t *materials;    /* [Mat][12] base.rgb f32, flags, emissive.rgb f32, alpha_mode,
                                     base/emissive/normal texture (0xFFFFFFFF none), alpha_cutoff f32 */
    uint32_t material_count;
    const float *lighting;        /* [16] see Lighting.pack */
    const uint32_t *textures;     /* [Tx][4] first texel, width, height, flags (bit 0 sRGB) */
    uint32_t texture_count;
    const uint32_t *texels;       /* RGBA8, R in the low byte */
    uint64_t texel_count;
} orc_scene_desc;
#define ORC_MATERIAL_WORDS 12u

typedef struct { float lo[3], hi[3]; uint32_t a, b; /* internal: left,right ; leaf: first,count|LEAF */ } orc_bnode;
#define ORC_LEAF 0x80000000u

typedef struct {
    orc_scene_desc d;
    uint32_t tri_count;
    float *tri;          /* [T][9] world-space vertices in global triangle order */
    uint32_t *tri_inst;  /* [T] */
    uint32_t *tri_prim;  /* [T] */
    uint32_t *inst_tri_offset; /* [I] */
    orc_bnode *nodes;
    uint32_t node_count;
    uint32_t *order;     /* leaf triangle order */
    int use_bvh;
    const float *sky_lut; /* sky-view table [200][200][4] (caller-owned) or NULL */
    float slack;         /* absolute box-test slack, see box_hit */
    int roulette;        /* Russian roulette on paths (extension), see trace_radiance */
    uint32_t frame_index; /* stochastic alpha test's frame index (scene.slang:146-148) */
    uint32_t rng_pcg;    /* 1: PCG streams (rng_1d), 0: the reference's LCG */
    float lut[512];      /* texel decode: c / 255, then the sRGB decode of c / 255 */
} orc_scene;

static void build_world_triangles(orc_scene *s) {
    const orc_scene_desc *d = &s->d;
    uint32_t total = 0;
    s->inst_tri_offset = (uint32_t *)malloc(sizeof(uint32_t) * (d->instance_count ? d->instance_count : 1));
    for (uint32_t i = 0; i < d->instance_count; ++i) {
        s->inst_tri_offset[i] = total;
        total += d->mesh_descs[4 * (size_t)d->instance_descs[4 * (size_t)i] + 1] / 3;
    }
    s->tri_count = total;
    s->tri = (float *)malloc(sizeof(float) * 9 * (size_t)(total ? total : 1));
    s->tri_inst = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(total ? total : 1));
    s->tri_prim = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(total ? total : 1));
    for (uint32_t i = 0; i < d->instance_count; ++i) {
        const uint32_t *inst = d->instance_descs + 4 * (size_t)i;
        const uint32_t *mesh = d->mesh_descs + 4 * (size_t)inst[0];
        const float *m = d->transforms + 16 * (size_t)inst[2];
        uint32_t ntri = mesh[1] / 3;
        for (uint32_t p = 0; p < ntri; ++p) {
            size_t g = (size_t)s->inst_tri_offset[i] + p;
            s->tri_inst[g] = i;
            s->tri_prim[g] = p;
            for (int k = 0; k < 3; ++k) {
                uint32_t vi = mesh[2] + d->indices[mesh[3] + 3 * p + k];
                const float *v = d->vertices + 8 * (size_t)vi;
                for (int r = 0; r < 3; ++r) {
                    /* world = M (x, y, z, 1), evaluated in float64, rounded once */
                    double w = (double)m[4 * r + 0] * v[0] + (double)m[4 * r + 1] * v[1];
                    w = w + (double)m[4 * r + 2] * v[2];
                    w = w + (double)m[4 * r + 3];
                    s->tri[9 * g + 3 * k + r] = (float)w;
                }
            }
        }
    }
}


/* ------------------------------------------------------------------ */
/* textures and the alpha test                                          */
/* ------------------------------------------------------------------ */
/* Texel decode table: c / 255, then the inverse sRGB transfer function
 * (IEC 61966-2-1) of c / 255, each in double, rounded once.  The reference
 * loads base-colour and emissive textures as sRGB, the rest as UNORM
 * (texture_manager.py:93-94, 120-127); the GPU decodes per texel. */
static void make_texel_lut(float *lut) {
    for (int c = 0; c < 256; ++c) {
        double x = c / 255.0;
        lut[c] = (float)x;
        lut[256 + c] = (float)(x <= 0.04045 ? x / 12.92 : pow((x + 0.055) / 1.055, 2.4));
    }
}

ORC_API void orc_texel_lut(float *out512) { make_texel_lut(out512); }

/* Wrap addressing on one axis (the reference's default sampler,
 * scene.py:116): u wrapped into [0, 1], texel centres at (i + 0.5) / n, the
 * two texels a bilinear footprint straddles and the second one's weight. */
static void wrap_axis(float u, uint32_t n, uint32_t *i0, uint32_t *i1, float *a) {
    if (!(fabsf(u) <= FLT_MAX)) u = 0.0f;
    float f = u - floorf(u);
    float x = f * (float)n - 0.5f;
    float fl = floorf(x);
    *a = x - fl;
    int k = (int)fl;
    *i0 = k < 0 ? n - 1u : (uint32_t)k;
    *i1 = *i0 + 1u == n ? 0u : *i0 + 1u;
}

typedef struct { float r, g, b, a; } orc_rgba;

static orc_rgba texel_value(const float *lut, const float *clut, uint32_t w) {
    orc_rgba t = {clut[w & 0xFFu], clut[(w >> 8) & 0xFFu], clut[(w >> 16) & 0xFFu], lut[w >> 24]};
    return t;
}

static float lerp2(float ax, float ay, float p00, float p10, float p01, float p11) {
    float r0 = fmaf(ax, p10 - p00, p00);
    float r1 = fmaf(ax, p11 - p01, p01);
    return fmaf(ay, r1 - r0, r0);
}

/* SampleLevel(linear_sampler, uv, 0) on mip 0 of an RGBA8 texture
 * (scene.slang:124-131, 313-368): bilinear over the decoded texels, along
 * x then y, one fmaf per lerp. */
static orc_rgba sample_texture(const orc_scene *s, uint32_t tex, float u, float v) {
    const uint32_t *d = s->d.textures + 4 * (size_t)tex;
    const float *clut = s->lut + ((d[3] & 1u) ? 256 : 0);
    uint32_t x0, x1, y0, y1;
    float ax, ay;
    wrap_axis(u, d[1], &x0, &x1, &ax);
    wrap_axis(v, d[2], &y0, &y1, &ay);
    const uint32_t *row0 = s->d.texels + d[0] + (size_t)y0 * d[1];
    const uint32_t *row1 = s->d.texels + d[0] + (size_t)y1 * d[1];
    orc_rgba t00 = texel_value(s->lut, clut, row0[x0]), t10 = texel_value(s->lut, clut, row0[x1]);
    orc_rgba t01 = texel_value(s->lut, clut, row1[x0]), t11 = texel_value(s->lut, clut, row1[x1]);
    orc_rgba out = {lerp2(ax, ay, t00.r, t10.r, t01.r, t11.r), lerp2(ax, ay, t00.g, t10.g, t01.g, t11.g),
                    lerp2(ax, ay, t00.b, t10.b, t01.b, t11.b), lerp2(ax, ay, t00.a, t10.a, t01.a, t11.a)};
    return out;
}

ORC_API void orc_sample_texture(const orc_scene *s, uint32_t tex, const float *uv, uint64_t n, float *out4) {
    for (uint64_t i = 0; i < n; ++i) {
        orc_rgba t = sample_texture(s, tex, uv[2 * i], uv[2 * i + 1]);
        out4[4 * i] = t.r, out4[4 * i + 1] = t.g, out4[4 * i + 2] = t.b, out4[4 * i + 3] = t.a;
    }
}

/* murmur_add + murmur_mix (common.slang:520-545): MurmurHash3's block mix
 * of k into h and its finaliser. */
static uint32_t murmur_key_mix(uint32_t h, uint32_t k) {
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

/* int3(pos * scale) of the reference: truncation; out of range saturates and
 * NaN gives 0 (what the GPU's conversion does) */
static int32_t to_i32_sat(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}

/* position_hashing(pos, 10000) (common.slang:547-564) */
static uint32_t position_hash(f3 p) {
    uint32_t vx = ((uint32_t)to_i32_sat(p.x * 10000.0f) + 0x3ffffu) & 0x7ffffu;
    uint32_t vy = ((uint32_t)to_i32_sat(p.y * 10000.0f) + 0x3ffffu) & 0x7ffffu;
    uint32_t vz = ((uint32_t)to_i32_sat(p.z * 10000.0f) + 0x3ffffu) & 0x7ffffu;
    uint32_t k0 = ((vx & 0x7fffu) << 7) | (vy & 0x7fu);
    uint32_t k1 = ((vy & 0xfffu) << 20) | ((vz & 0x7fffu) << 1);
    return murmur_key_mix(k0, k1);
}

ORC_API uint32_t orc_murmur_key_mix(uint32_t h, uint32_t k) { return murmur_key_mix(h, k); }

ORC_API void orc_alpha_hash(const float *pos, uint64_t n, uint32_t frame_index, uint32_t *out) {
    for (uint64_t i = 0; i < n; ++i)
        out[i] = murmur_key_mix(position_hash(mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2])), frame_index);
}

static inline float u32_bits_to_f(uint32_t w) { float f; memcpy(&f, &w, 4); return f; }

/* eval_alpha_test on a candidate hit (scene.slang:124-149) as
 * Scene.intersect / get_visibility apply it (scene.slang:1017-1172):
 * opaque materials always keep the hit; mask keeps opacity >= cutoff;
 * blend keeps opacity >= float(murmur_key_mix(position_hashing(world
 * position, 10000), frame_index)) / 2^32.  Opacity is the base-colour
 * texture's alpha at the interpolated uv, 1 without a texture.  World
 * position and uv interpolate as get_vertex does (scene.slang:157-177). */
static int alpha_accept(const orc_scene *s, uint32_t g, float b1, float b2) {
    const orc_scene_desc *d = &s->d;
    const uint32_t *inst = d->instance_descs + 4 * (size_t)s->tri_inst[g];
    const uint32_t *mat = d->materials + ORC_MATERIAL_WORDS * (size_t)inst[1];
    if (mat[7] == 0u) return 1;
    const uint32_t *mesh = d->mesh_descs + 4 * (size_t)inst[0];
    const uint32_t *idx = d->indices + mesh[3] + 3 * (size_t)s->tri_prim[g];
    const float *V0 = d->vertices + 8 * (size_t)(mesh[2] + idx[0]);
    const float *V1 = d->vertices + 8 * (size_t)(mesh[2] + idx[1]);
    const float *V2 = d->vertices + 8 * (size_t)(mesh[2] + idx[2]);
    float w0 = 1.0f - b1 - b2;
    float opacity = 1.0f;
    if (mat[8] != ORC_INVALID) {
        float u = w0 * V0[6] + b1 * V1[6] + b2 * V2[6];
        float v = w0 * V0[7] + b1 * V1[7] + b2 * V2[7];
        opacity = sample_texture(s, mat[8], u, v).a;
    }
    if (mat[7] == 1u) return opacity >= u32_bits_to_f(mat[11]);
    float p[3], w[3];
    for (int k = 0; k < 3; ++k) p[k] = w0 * V0[k] + b1 * V1[k] + b2 * V2[k];
    const float *M = d->transforms + 16 * (size_t)inst[2];
    for (int r = 0; r < 3; ++r) w[r] = M[4 * r] * p[0] + M[4 * r + 1] * p[1] + M[4 * r + 2] * p[2] + M[4 * r + 3];
    uint32_t seed = murmur_key_mix(position_hash(mk(w[0], w[1], w[2])), s->frame_index);
    return opacity >= (float)seed * 0x1p-32f;
}

static inline int alpha_tested(const orc_scene *s, uint32_t g) {
    const orc_scene_desc *d = &s->d;
    return d->materials[ORC_MATERIAL_WORDS * (size_t)d->instance_descs[4 * (size_t)s->tri_inst[g] + 1] + 7] != 0u;
}

/* --- oracle BVH: binary, object-median split on the longest centroid axis --- */
typedef struct { float c[3]; uint32_t id; } orc_ref;

static void bounds_of(const orc_scene *s, const uint32_t *ids, uint32_t n, float *lo, float *hi) {
    for (int k = 0; k < 3; ++k) { lo[k] = FLT_MAX; hi[k] = -FLT_MAX; }
    for (uint32_t i = 0; i < n; ++i) {
        const float *t = s->tri + 9 * (size_t)ids[i];
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k) {
                float x = t[3 * v + k];
                if (x < lo[k]) lo[k] = x;
                if (x > hi[k]) hi[k] = x;
            }
    }
}

static void select_median(orc_ref *refs, uint32_t n, uint32_t kth, int axis) {
    /* quickselect on refs[].c[axis], ties broken by id */
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        orc_ref pivot = refs[(lo + hi) / 2];
        uint32_t i = lo, j = hi;
        while (i <= j) {
            while (refs[i].c[axis] < pivot.c[axis] ||
                   (refs[i].c[axis] == pivot.c[axis] && refs[i].id < pivot.id)) ++i;
            while (refs[j].c[axis] > pivot.c[axis] ||
                   (refs[j].c[axis] == pivot.c[axis] && refs[j].id > pivot.id)) --j;
            if (i <= j) {
                orc_ref t = refs[i]; refs[i] = refs[j]; refs[j] = t;
                ++i;
                if (j == 0) break;
                --j;
            }
        }
        if (kth <= j) hi = j;
        else if (kth >= i) lo = i;
        else break;
    }
}

typedef struct { uint32_t node, begin, end; } orc_task;

static void build_bvh(orc_scene *s) {
    uint32_t n = s->tri_count;
    orc_ref *refs = (orc_ref *)malloc(sizeof(orc_ref) * (size_t)(n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) {
        const float *t = s->tri + 9 * (size_t)i;
        for (int k = 0; k < 3; ++k) refs[i].c[k] = (t[k] + t[3 + k] + t[6 + k]) * (1.0f / 3.0f);
        refs[i].id = i;
    }
    size_t cap = 2 * (size_t)(n ? n : 1);
    s->nodes = (orc_bnode *)malloc(sizeof(orc_bnode) * cap);
    s->order = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    s->node_count = 1;
    orc_task *stack = (orc_task *)malloc(sizeof(orc_task) * 256);
    int sp = 0;
    stack[sp++] = (orc_task){0, 0, n};
    uint32_t *ids = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    while (sp > 0) {
        orc_task t = stack[--sp];
        uint32_t cnt = t.end - t.begin;
        for (uint32_t i = 0; i < cnt; ++i) ids[i] = refs[t.begin + i].id;
        orc_bnode *nd = &s->nodes[t.node];
        bounds_of(s, ids, cnt, nd->lo, nd->hi);
        if (cnt <= 4) {
            for (uint32_t i = 0; i < cnt; ++i) s->order[t.begin + i] = ids[i];
            nd->a = t.begin;
            nd->b = cnt | ORC_LEAF;
            continue;
        }
        float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (uint32_t i = t.begin; i < t.end; ++i)
            for (int k = 0; k < 3; ++k) {
                if (refs[i].c[k] < clo[k]) clo[k] = refs[i].c[k];
                if (refs[i].c[k] > chi[k]) chi[k] = refs[i].c[k];
            }
        int axis = 0;
        float ext = chi[0] - clo[0];
        for (int k = 1; k < 3; ++k) if (chi[k] - clo[k] > ext) { ext = chi[k] - clo[k]; axis = k; }
        uint32_t mid = cnt / 2;
        select_median(refs + t.begin, cnt, mid, axis);
        uint32_t left = s->node_count, right = s->node_count + 1;
        s->node_count += 2;
        nd = &s->nodes[t.node];
        nd->a = left;
        nd->b = right;
        stack[sp++] = (orc_task){right, t.begin + mid, t.end};
        stack[sp++] = (orc_task){left, t.begin, t.begin + mid};
    }
    free(ids);
    free(stack);
    free(refs);
}

ORC_API orc_scene *orc_scene_create(const orc_scene_desc *desc, int use_bvh) {
    orc_scene *s = (orc_scene *)calloc(1, sizeof(orc_scene));
    s->d = *desc;
    s->use_bvh = use_bvh;
    make_texel_lut(s->lut);
    build_world_triangles(s);
    float m = 1.0f;
    for (size_t i = 0; i < 9 * (size_t)s->tri_count; ++i) m = fmaxf(m, fabsf(s->tri[i]));
    s->slack = m * 0x1p-18f;
    if (use_bvh) build_bvh(s);
    return s;
}

ORC_API void orc_scene_set_sky_view_lut(orc_scene *s, const float *sky_lut) { s->sky_lut = sky_lut; }
ORC_API void orc_scene_set_russian_roulette(orc_scene *s, int enable) { s->roulette = enable != 0; }
ORC_API void orc_scene_set_frame_index(orc_scene *s, uint32_t frame_index) { s->frame_index = frame_index; }
ORC_API void orc_scene_set_rng_mode(orc_scene *s, uint32_t pcg) { s->rng_pcg = pcg != 0; }

ORC_API void orc_scene_destroy(orc_scene *s) {
    if (!s) return;
    free(s->tri); free(s->tri_inst); free(s->tri_prim); free(s->inst_tri_offset);
    free(s->nodes); free(s->order);
    free(s);
}

ORC_API uint32_t orc_scene_triangle_count(const orc_scene *s) { return s->tri_count; }
ORC_API void orc_scene_world_triangles(const orc_scene *s, float *out) {
    memcpy(out, s->tri, sizeof(float) * 9 * (size_t)s->tri_count);
}

/* ------------------------------------------------------------------ */
/* ray / triangle: watertight test (Woop, Benthin, Wald, JCGT 2013)    */
/* ------------------------------------------------------------------ */
typedef struct {
    f3 o, d;
    float tmin, tmax;
    int kx, ky, kz;
    float Sx, Sy, Sz;
    f3 inv;   /* slab reciprocals with zero components replaced */
    int flat[3];  /* direction component below 1e-30: the ray keeps that coordinate */
} orc_ray;

static void ray_setup(orc_ray *r) {
    float a[3] = {fabsf(r->d.x), fabsf(r->d.y), fabsf(r->d.z)};
    int kz = 0;
    if (a[1] > a[kz]) kz = 1;
    if (a[2] > a[kz]) kz = 2;
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (comp(r->d, kz) < 0.0f) { int t = kx; kx = ky; ky = t; }
    r->kx = kx; r->ky = ky; r->kz = kz;
    float dz = comp(r->d, kz);
    r->Sx = comp(r->d, kx) / dz;
    r->Sy = comp(r->d, ky) / dz;
    r->Sz = 1.0f / dz;
    float dd[3] = {r->d.x, r->d.y, r->d.z}, iv[3];
    for (int k = 0; k < 3; ++k) {
        float x = dd[k];
        r->flat[k] = fabsf(x) < 1e-30f;
        if (r->flat[k]) x = copysignf(1e-30f, x);
        iv[k] = 1.0f / x;
    }
    r->inv = mk(iv[0], iv[1], iv[2]);
}

/* returns 1 on a hit inside [tmin, tmax] and writes t, b1 (weight of v1), b2 (v2) */
static int tri_intersect(const orc_ray *r, const float *tri, float *t_out, float *b1_out, float *b2_out) {
    f3 A = sub(mk(tri[0], tri[1], tri[2]), r->o);
    f3 B = sub(mk(tri[3], tri[4], tri[5]), r->o);
    f3 C = sub(mk(tri[6], tri[7], tri[8]), r->o);
    float Akz = comp(A, r->kz), Bkz = comp(B, r->kz), Ckz = comp(C, r->kz);
    float Ax = comp(A, r->kx) - r->Sx * Akz;
    float Ay = comp(A, r->ky) - r->Sy * Akz;
    float Bx = comp(B, r->kx) - r->Sx * Bkz;
    float By = comp(B, r->ky) - r->Sy * Bkz;
    float Cx = comp(C, r->kx) - r->Sx * Ckz;
    float Cy = comp(C, r->ky) - r->Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return 0;
    float det = U + V + W;
    if (det == 0.0f) return 0;
    float Az = r->Sz * Akz, Bz = r->Sz * Bkz, Cz = r->Sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= r->tmin && t <= r->tmax)) return 0;
    *t_out = t;
    *b1_out = V * rcp;
    *b2_out = W * rcp;
    return 1;
}

/* Box-test slack.  A box is visited when its entry distance is within
 * far * (1 + 2^-10) + slack and within best_t * (1 + 2^-10) + slack, where
 * slack = 2^-18 * max(1, largest world coordinate).  The relative term covers
 * the slab test's rounding (2 gamma_3, Ize 2013); the absolute term covers the
 * FP32 triangle test reporting a distance below its box's true entry (or a
 * hit just behind the origin) by a few ulps of the coordinates. */
#define ORC_PAD 0x1.004p+0f  /* 1 + 2^-10 */

/* An axis the ray does not move along (|d| < 1e-30: under 1e-27 over the
 * longest ray) meets the slab iff the origin lies in it, faces included: with
 * the clamped reciprocal a ray lying in a face plane would get the slab
 * interval [-huge, 0] and lose the box, and an edge hit on that face with it. */
static int box_hit(const orc_ray *r, const float *lo, const float *hi, float tfar_limit, float slack) {
    const float o3[3] = {r->o.x, r->o.y, r->o.z};
    for (int k = 0; k < 3; ++k)
        if (r->flat[k] && !(o3[k] >= lo[k] && o3[k] <= hi[k])) return 0;
    float t0x = r->flat[0] ? -INFINITY : (lo[0] - r->o.x) * r->inv.x;
    float t1x = r->flat[0] ? INFINITY : (hi[0] - r->o.x) * r->inv.x;
    float t0y = r->flat[1] ? -INFINITY : (lo[1] - r->o.y) * r->inv.y;
    float t1y = r->flat[1] ? INFINITY : (hi[1] - r->o.y) * r->inv.y;
    float t0z = r->flat[2] ? -INFINITY : (lo[2] - r->o.z) * r->inv.z;
    float t1z = r->flat[2] ? INFINITY : (hi[2] - r->o.z) * r->inv.z;
    float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), r->tmin));
    float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z)) * ORC_PAD + slack;
    tf = fminf(tf, tfar_limit);
    return tn <= tf;
}

typedef struct { int valid; float t, b1, b2; uint32_t id; } orc_hit;

/* closest hit with the tie rule "smallest global triangle id" */
static orc_hit closest_hit(const orc_scene *s, const orc_ray *r, uint64_t *tri_tests) {
    orc_hit h = {0, r->tmax, 0.0f, 0.0f, ORC_INVALID};
    uint64_t tests = 0;
    if (!s->use_bvh) {
        for (uint32_t g = 0; g < s->tri_count; ++g) {
            float t, b1, b2;
            ++tests;
            if (tri_intersect(r, s->tri + 9 * (size_t)g, &t, &b1, &b2)) {
                if ((t < h.t || (t == h.t && g < h.id)) && (!alpha_tested(s, g) || alpha_accept(s, g, b1, b2))) {
                    h.valid = 1; h.t = t; h.b1 = b1; h.b2 = b2; h.id = g;
                }
            }
        }
    } else if (s->tri_count) {
        uint32_t stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const orc_bnode *nd = &s->nodes[stack[--sp]];
            float limit = h.t * ORC_PAD + s->slack;
            if (!box_hit(r, nd->lo, nd->hi, limit, s->slack)) continue;
            if (nd->b & ORC_LEAF) {
                uint32_t cnt = nd->b & ~ORC_LEAF;
                for (uint32_t i = 0; i < cnt; ++i) {
                    uint32_t g = s->order[nd->a + i];
                    float t, b1, b2;
                    ++tests;
                    if (tri_intersect(r, s->tri + 9 * (size_t)g, &t, &b1, &b2)) {
                        if ((t < h.t || (t == h.t && g < h.id)) &&
                            (!alpha_tested(s, g) || alpha_accept(s, g, b1, b2))) {
                            h.valid = 1; h.t = t; h.b1 = b1; h.b2 = b2; h.id = g;
                        }
                    }
                }
            } else {
                stack[sp++] = nd->b;
                stack[sp++] = nd->a;
            }
        }
    }
    if (tri_tests) *tri_tests += tests;
    return h;
}

static int any_hit(const orc_scene *s, const orc_ray *r) {
    if (!s->use_bvh) {
        for (uint32_t g = 0; g < s->tri_count; ++g) {
            float t, b1, b2;
            if (tri_intersect(r, s->tri + 9 * (size_t)g, &t, &b1, &b2) &&
                (!alpha_tested(s, g) || alpha_accept(s, g, b1, b2)))
                return 1;
        }
        return 0;
    }
    if (!s->tri_count) return 0;
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const orc_bnode *nd = &s->nodes[stack[--sp]];
        if (!box_hit(r, nd->lo, nd->hi, r->tmax * ORC_PAD + s->slack, s->slack)) continue;
        if (nd->b & ORC_LEAF) {
            uint32_t cnt = nd->b & ~ORC_LEAF;
            for (uint32_t i = 0; i < cnt; ++i) {
                float t, b1, b2;
                uint32_t g = s->order[nd->a + i];
                if (tri_intersect(r, s->tri + 9 * (size_t)g, &t, &b1, &b2) &&
                    (!alpha_tested(s, g) || alpha_accept(s, g, b1, b2)))
                    return 1;
            }
        } else {
            stack[sp++] = nd->b;
            stack[sp++] = nd->a;
        }
    }
    return 0;
}

/* Debug cross-check: with orc_set_crosscheck(1), every BVH query is repeated
 * by brute force and the first disagreeing ray is kept. */
static int g_crosscheck = 0;
static float g_bad_ray[8];
static int g_bad_kind = 0;
static pthread_mutex_t g_bad_lock = PTHREAD_MUTEX_INITIALIZER;

ORC_API void orc_set_crosscheck(int on) { g_crosscheck = on; g_bad_kind = 0; }
ORC_API int orc_crosscheck_result(float *ray8) { memcpy(ray8, g_bad_ray, sizeof(g_bad_ray)); return g_bad_kind; }

static void record_bad(const orc_ray *r, int kind) {
    pthread_mutex_lock(&g_bad_lock);
    if (!g_bad_kind) {
        g_bad_kind = kind;
        float q[8] = {r->o.x, r->o.y, r->o.z, r->tmin, r->d.x, r->d.y, r->d.z, r->tmax};
        memcpy(g_bad_ray, q, sizeof(q));
    }
    pthread_mutex_unlock(&g_bad_lock);
}

/* rays: [n][8] = origin.xyz, tmin, dir.xyz, tmax.  Outputs t, (b1,b2), instance, prim;
 * misses report instance = prim = 0xFFFFFFFF. */
ORC_API void orc_intersect(const orc_scene *s, const float *rays, uint64_t n, float *out_t, float *out_b,
                           uint32_t *out_inst, uint32_t *out_prim, uint32_t *out_occluded) {
    for (uint64_t i = 0; i < n; ++i) {
        const float *q = rays + 8 * i;
        orc_ray r;
        r.o = mk(q[0], q[1], q[2]);
        r.tmin = q[3];
        r.d = mk(q[4], q[5], q[6]);
        r.tmax = q[7];
        ray_setup(&r);
        orc_hit h = closest_hit(s, &r, NULL);
        out_t[i] = h.valid ? h.t : -1.0f;
        out_b[2 * i] = h.b1;
        out_b[2 * i + 1] = h.b2;
        out_inst[i] = h.valid ? s->tri_inst[h.id] : ORC_INVALID;
        out_prim[i] = h.valid ? s->tri_prim[h.id] : ORC_INVALID;
        if (out_occluded) out_occluded[i] = (uint32_t)any_hit(s, &r);
    }
}

/* FP64 brute force nearest hit (Moller-Trumbore in double over the float32
 * world-space vertices).  Also reports the smallest |t - t_other| over all
 * other hit triangles and the smallest |barycentric| of the nearest hit, so a
 * test can exclude rays whose hit is within a tie margin. */
ORC_API void orc_intersect_fp64(const orc_scene *s, const double *rays, uint64_t n, double *out_t,
                                double *out_b, uint32_t *out_inst, uint32_t *out_prim,
                                double *out_t_gap, double *out_edge) {
    for (uint64_t i = 0; i < n; ++i) {
        const double *q = rays + 8 * i;
        double best = q[7], second = INFINITY, bb1 = 0, bb2 = 0, edge = 0;
        uint32_t best_id = ORC_INVALID;
        int any = 0;
        for (uint32_t g = 0; g < s->tri_count; ++g) {
            const float *t = s->tri + 9 * (size_t)g;
            double v0[3] = {t[0], t[1], t[2]}, v1[3] = {t[3], t[4], t[5]}, v2[3] = {t[6], t[7], t[8]};
            double e1[3], e2[3], pv[3], tv[3], qv[3];
            for (int k = 0; k < 3; ++k) { e1[k] = v1[k] - v0[k]; e2[k] = v2[k] - v0[k]; }
            const double *d = q + 4;
            pv[0] = d[1] * e2[2] - d[2] * e2[1];
            pv[1] = d[2] * e2[0] - d[0] * e2[2];
            pv[2] = d[0] * e2[1] - d[1] * e2[0];
            double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
            if (det == 0.0) continue;
            double inv = 1.0 / det;
            for (int k = 0; k < 3; ++k) tv[k] = q[k] - v0[k];
            double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
            if (u < 0.0 || u > 1.0) continue;
            qv[0] = tv[1] * e1[2] - tv[2] * e1[1];
            qv[1] = tv[2] * e1[0] - tv[0] * e1[2];
            qv[2] = tv[0] * e1[1] - tv[1] * e1[0];
            double v = (d[0] * qv[0] + d[1] * qv[1] + d[2] * qv[2]) * inv;
            if (v < 0.0 || u + v > 1.0) continue;
            double tt = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
            if (tt < q[3] || tt > q[7]) continue;
            if (!any || tt < best) {
                if (any) second = best < second ? best : second;
                best = tt; bb1 = u; bb2 = v; best_id = g; any = 1;
                double w = 1.0 - u - v;
                edge = fmin(fmin(u, v), w);
            } else if (tt < second) {
                second = tt;
            }
        }
        out_t[i] = any ? best : -1.0;
        out_b[2 * i] = bb1;
        out_b[2 * i + 1] = bb2;
        out_inst[i] = any ? s->tri_inst[best_id] : ORC_INVALID;
        out_prim[i] = any ? s->tri_prim[best_id] : ORC_INVALID;
        out_t_gap[i] = any ? second - best : INFINITY;
        out_edge[i] = edge;
    }
}

/* ------------------------------------------------------------------ */
/* shading (scene.slang:178-406, common.slang:54-162, 452-490)        */
/* ------------------------------------------------------------------ */
typedef struct { f3 n, t, b; } orc_frame;

static orc_frame make_frame(f3 n) {
    orc_frame f;
    f.n = n;
    f.t = normalize(cross(n, mk(-n.z, n.x, -n.y)));
    f.b = cross(n, f.t);
    return f;
}
static inline f3 to_global(const orc_frame *f, f3 w) {
    return add(add(muls(f->t, w.x), muls(f->b, w.y)), muls(f->n, w.z));
}

typedef struct {
    f3 position, origin_front, face_normal, view;
    orc_frame frame;
    f3 diffuse, emissive;
} orc_shading;

static inline f3 row3(const float *m, int r) { return mk(m[4 * r], m[4 * r + 1], m[4 * r + 2]); }
static inline f3 mul33(const float *m, f3 v) { return mk(dot(row3(m, 0), v), dot(row3(m, 1), v), dot(row3(m, 2), v)); }

static void shading_data(const orc_scene *s, uint32_t instance, uint32_t prim, float b1, float b2, f3 view,
                         orc_shading *sd) {
    const orc_scene_desc *d = &s->d;
    const uint32_t *inst = d->instance_descs + 4 * (size_t)instance;
    const uint32_t *mesh = d->mesh_descs + 4 * (size_t)inst[0];
    const float *M = d->transforms + 16 * (size_t)inst[2];
    const float *IT = d->inv_transposes + 16 * (size_t)inst[2];
    const uint32_t *idx = d->indices + mesh[3] + 3 * (size_t)prim;
    const float *V0 = d->vertices + 8 * (size_t)(mesh[2] + idx[0]);
    const float *V1 = d->vertices + 8 * (size_t)(mesh[2] + idx[1]);
    const float *V2 = d->vertices + 8 * (size_t)(mesh[2] + idx[2]);
    f3 v0 = mk(V0[0], V0[1], V0[2]), v1 = mk(V1[0], V1[1], V1[2]), v2 = mk(V2[0], V2[1], V2[2]);
    f3 bc = mk(1.0f - b1 - b2, b1, b2);

    /* get_triangle_spawn_data (scene.slang:178-269) */
    const float c0 = 5.9604644775390625e-8f, c1 = 1.7881397695873602e-7f, c2 = 1.1920931797249068e-7f;
    f3 e1 = sub(v1, v0), e2 = sub(v2, v0);
    f3 op = mk(v0.x + fmaf(b1, e1.x, b2 * e2.x), v0.y + fmaf(b1, e1.y, b2 * e2.y),
               v0.z + fmaf(b1, e1.z, b2 * e2.z));
    f3 on = cross(e1, e2);
    float wp[3];
    for (int r = 0; r < 3; ++r)
        wp[r] = M[4 * r + 3] + fmaf(M[4 * r + 0], op.x, fmaf(M[4 * r + 1], op.y, M[4 * r + 2] * op.z));
    f3 world_position = mk(wp[0], wp[1], wp[2]);
    f3 wnu = mul33(IT, on);
    float world_scale = 1.0f / sqrtf(fmaxf(dot(wnu, wnu), 1e-30f));
    f3 world_normal = muls(wnu, world_scale);
    f3 ae1 = fabs3(e1), ae2 = fabs3(e2);
    f3 ext3 = add(add(ae1, ae2), fabs3(sub(ae1, ae2)));
    float extent = fmaxf(fmaxf(ext3.x, ext3.y), ext3.z);
    f3 object_error = add(muls(fabs3(v0), c0), mk(c1 * extent, c1 * extent, c1 * extent));
    f3 aop = fabs3(op);
    float we[3];
    for (int r = 0; r < 3; ++r) we[r] = c1 * dot(fabs3(row3(M, r)), aop) + c2 * fabsf(M[4 * r + 3]);
    f3 world_error = mk(we[0], we[1], we[2]);
    f3 awp = fabs3(world_position);
    float te[3];
    for (int c = 0; c < 3; ++c)
        te[c] = fabsf(IT[0 * 4 + c]) * awp.x + fabsf(IT[1 * 4 + c]) * awp.y + fabsf(IT[2 * 4 + c]) * awp.z +
                fabsf(IT[3 * 4 + c]) * 1.0f;
    object_error = add(object_error, muls(mk(te[0], te[1], te[2]), c2));
    float object_offset = dot(object_error, fabs3(on));
    float world_offset = dot(world_error, fabs3(world_normal));
    world_offset = fmaf(world_scale, object_offset, world_offset);

    /* get_shading_data (scene.slang:271-406) */
    f3 n0 = mk(V0[3], V0[4], V0[5]), n1 = mk(V1[3], V1[4], V1[5]), n2 = mk(V2[3], V2[4], V2[5]);
    f3 ni = add(add(muls(n0, bc.x), muls(n1, bc.y)), muls(n2, bc.z));
    f3 geometry_normal = normalize(mul33(IT, ni));
    f3 face_normal = world_normal;
    const uint32_t *mat = d->materials + ORC_MATERIAL_WORDS * (size_t)inst[1];
    float mf[8];
    memcpy(mf, mat, sizeof(mf));
    f3 diffuse = mk(mf[0], mf[1], mf[2]), emissive = mk(mf[4], mf[5], mf[6]);
    f3 shading_normal = geometry_normal;
    if (mat[8] != ORC_INVALID || mat[9] != ORC_INVALID || mat[10] != ORC_INVALID) {
        /* textures at the interpolated uv (scene.slang:297, 313-368): the
         * base-colour texture replaces the factor, the emissive texture
         * scales it, the normal map tilts the interpolated normal in a
         * tangent frame from object-space dp/duv (before the two-sided flip) */
        float u = bc.x * V0[6] + bc.y * V1[6] + bc.z * V2[6];
        float v = bc.x * V0[7] + bc.y * V1[7] + bc.z * V2[7];
        if (mat[8] != ORC_INVALID) {
            orc_rgba t = sample_texture(s, mat[8], u, v);
            diffuse = mk(t.r, t.g, t.b);
        }
        if (mat[9] != ORC_INVALID) {
            orc_rgba t = sample_texture(s, mat[9], u, v);
            emissive = mk(t.r * mf[4], t.g * mf[5], t.b * mf[6]);
        }
        if (mat[10] != ORC_INVALID) {
            orc_rgba t = sample_texture(s, mat[10], u, v);
            f3 tn = normalize(mk(t.r * 2.0f - 1.0f, t.g * 2.0f - 1.0f, t.b * 2.0f - 1.0f));
            float d1u = V1[6] - V0[6], d1v = V1[7] - V0[7], d2u = V2[6] - V0[6], d2v = V2[7] - V0[7];
            float det = d1u * d2v - d1v * d2u;
            if (fabsf(det) > 1e-6f) {
                f3 tg = sub(muls(e1, d2v), muls(e2, d1v));
                tg = normalize(mk(tg.x / det, tg.y / det, tg.z / det));
                tg = mul33(IT, tg);
                tg = normalize(sub(tg, muls(geometry_normal, dot(tg, geometry_normal))));
                f3 bt = cross(geometry_normal, tg);
                shading_normal = normalize(add(add(muls(tg, tn.x), muls(bt, tn.y)), muls(geometry_normal, tn.z)));
            }
        }
    }
    int double_sided = (mat[3] & 1u) != 0;
    int front = dot(face_normal, view) > 0.0f;
    float sign = (double_sided && !front) ? -1.0f : 1.0f;
    face_normal = muls(face_normal, sign);
    shading_normal = muls(shading_normal, sign);
    /* adjust_shading_normal (common.slang:475-490) with D = -view */
    f3 D = neg(view);
    float dn = 2.0f * dot(shading_normal, D);
    f3 R = sub(D, muls(shading_normal, dn));
    float k = dot(R, face_normal);
    if (k < 0.0f) shading_normal = normalize(sub(normalize(sub(R, muls(face_normal, k))), D));

    sd->position = world_position;
    sd->origin_front = add(world_position, muls(face_normal, world_offset));
    sd->face_normal = face_normal;
    sd->view = view;
    sd->frame = make_frame(shading_normal);
    sd->diffuse = diffuse;
    sd->emissive = emissive;
}

/* world flat normal (texture_space_path_tracer.slang:39-55) */
static f3 world_flat_normal(const orc_scene *s, uint32_t instance, uint32_t prim) {
    const orc_scene_desc *d = &s->d;
    const uint32_t *inst = d->instance_descs + 4 * (size_t)instance;
    const uint32_t *mesh = d->mesh_descs + 4 * (size_t)inst[0];
    const float *IT = d->inv_transposes + 16 * (size_t)inst[2];
    const uint32_t *idx = d->indices + mesh[3] + 3 * (size_t)prim;
    const float *P0 = d->vertices + 8 * (size_t)(mesh[2] + idx[0]);
    const float *P1 = d->vertices + 8 * (size_t)(mesh[2] + idx[1]);
    const float *P2 = d->vertices + 8 * (size_t)(mesh[2] + idx[2]);
    f3 p0 = mk(P0[0], P0[1], P0[2]), p1 = mk(P1[0], P1[1], P1[2]), p2 = mk(P2[0], P2[1], P2[2]);
    f3 n = normalize(cross(sub(p1, p0), sub(p2, p0)));
    if (inst[3] & 1u) n = neg(n);
    return normalize(mul33(IT, n));
}

/* cosine_sample_hemisphere (common.slang:54-68): xyz, pdf */
static void cosine_sample(float ex, float ey, f3 *dir, float *pdf) {
    float s, c;
    sincos_2pi(ex, &s, &c);
    float ct = sqrtf(ey);
    float st = sqrtf(1.0f - ct * ct);
    *dir = mk(st * c, st * s, ct);
    *pdf = ct * ORC_INV_PI;
}

/* sample_cone_direction (common.slang:452-470), direction only */
static f3 cone_sample(f3 axis, float cos_half, float ux, float uy) {
    float ct = 1.0f - ux * (1.0f - cos_half);
    float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float s, c;
    sincos_2pi(uy, &s, &c);
    f3 local = mk(c * st, s * st, ct);
    orc_frame f = make_frame(axis);
    return to_global(&f, local);
}

typedef struct {
    uint64_t rays_extend, rays_shadow, tri_tests;
} orc_counters;

typedef struct {
    f3 sun_dir; float cos_half;
    f3 sun_illum;  /* color * intensity * transmittance */
    f3 outer;      /* color * intensity */
    f3 sky;
    int sun_on;
    const float *sky_lut;  /* sky-view table [200][200][4] or NULL */
    float horizon;
} orc_light;

/* horizon angle of the view point (scene.slang:474): FP32 up to the
 * arccosine, taken in double and rounded */
static float atmosphere_horizon(void) {
    const float ground = 6.360f, vy = 6.360f + 0.0002f;
    float h = sqrtf(vy * vy);
    float q = sqrtf(h * h - ground * ground) / h;
    return (float)acos((double)fminf(fmaxf(q, -1.0f), 1.0f)) - 0.5f * ORC_PI;
}

static orc_light load_light_lut(const float *L, const float *sky_lut) {
    orc_light l;
    l.sun_dir = mk(L[0], L[1], L[2]);
    l.cos_half = L[3];
    l.sun_illum = mk(L[4] * L[8], L[5] * L[9], L[6] * L[10]);
    l.outer = mk(L[4], L[5], L[6]);
    l.sky = mk(L[12], L[13], L[14]);
    l.sun_on = l.sun_illum.x > 0.0f || l.sun_illum.y > 0.0f || l.sun_illum.z > 0.0f;
    l.sky_lut = sky_lut;
    l.horizon = atmosphere_horizon();
    return l;
}

/* atan2 from correctly rounded operations: Abramowitz & Stegun 4.4.49 on
 * [0, 1] (|error| <= 2e-8) with octant reduction; asin through it. */
static float nan_atan2(float y, float x) {
    static const float c[9] = {1.0f,           -0.3333314528f, 0.1999355085f, -0.1420889944f, 0.1065626393f,
                               -0.0752896400f, 0.0429096138f,  -0.0161657367f, 0.0028662257f};
    float ax = fabsf(x), ay = fabsf(y);
    float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    if (!(mx > 0.0f)) return 0.0f;
    float a = mn / mx, s2 = a * a;
    float p = c[8];
    for (int k = 7; k >= 0; --k) p = fmaf(p, s2, c[k]);
    float r = a * p;
    if (ay > ax) r = 0.5f * ORC_PI - r;
    if (x < 0.0f) r = ORC_PI - r;
    return y < 0.0f ? -r : r;
}
static float nan_asin(float x) {
    x = fminf(fmaxf(x, -1.0f), 1.0f);
    return nan_atan2(x, sqrtf((1.0f - x) * (1.0f + x)));
}

/* Scene.sample_sky_lut (scene.slang:469-518), bilinear with wrap addressing */
static f3 sky_view_radiance(const orc_light *L, f3 dir) {
    const f3 view = mk(0.0f, 6.360f + 0.0002f, 0.0f);
    float height = sqrtf(dot(view, view));
    f3 up = mk(view.x / height, view.y / height, view.z / height);
    float alt = fminf(fmaxf(nan_asin(dot(dir, up)), -0.5f * ORC_PI + 1e-2f), 0.5f * ORC_PI - 1e-2f);
    float az = 0.0f;
    if (!(fabsf(alt) > 0.5f * ORC_PI - 0.0001f)) {
        f3 right = cross(L->sun_dir, up);
        f3 forward = cross(up, right);
        f3 p = sub(dir, muls(up, dot(dir, up)));
        if (dot(p, p) > 0.0f) {
            p = normalize(p);
            az = nan_atan2(dot(p, right), dot(p, forward));
        }
    }
    float adj = (alt + L->horizon) / (0.5f * ORC_PI);
    float v = adj < 0.0f ? 0.5f * (1.0f - sqrtf(fminf(fmaxf(-adj, 0.0f), 1.0f)))
                         : 0.5f * (1.0f + sqrtf(fminf(fmaxf(adj, 0.0f), 1.0f)));
    float u = az / (2.0f * ORC_PI) + 0.5f;
    if (!(u >= 0.0f && u <= 1.0f)) u = 0.5f;
    if (!(v >= 0.0f && v <= 1.0f)) v = 0.5f;
    float x = u * 200.0f - 0.5f, y = v * 200.0f - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float tx = x - fx, ty = y - fy;
    int x0 = (int)fx, y0 = (int)fy;
    int xa = (x0 + 200) % 200, xb = (x0 + 201) % 200, ya = (y0 + 200) % 200, yb = (y0 + 201) % 200;
    const float *a = L->sky_lut + 4 * (ya * 200 + xa), *b = L->sky_lut + 4 * (ya * 200 + xb);
    const float *c = L->sky_lut + 4 * (yb * 200 + xa), *d = L->sky_lut + 4 * (yb * 200 + xb);
    f3 r0 = add(muls(mk(a[0], a[1], a[2]), 1.0f - tx), muls(mk(b[0], b[1], b[2]), tx));
    f3 r1 = add(muls(mk(c[0], c[1], c[2]), 1.0f - tx), muls(mk(d[0], d[1], d[2]), tx));
    return mul(add(muls(r0, 1.0f - ty), muls(r1, ty)), L->outer);
}

static f3 miss_radiance(const orc_light *L, f3 dir) { return L->sky_lut ? sky_view_radiance(L, dir) : L->sky; }

ORC_API void orc_sky_view_lookup(const float *sky_lut, const float *lighting16, const float *dirs, uint64_t n,
                                 float *out) {
    orc_light L = load_light_lut(lighting16, sky_lut);
    for (uint64_t k = 0; k < n; ++k) {
        f3 r = sky_view_radiance(&L, mk(dirs[4 * k], dirs[4 * k + 1], dirs[4 * k + 2]));
        out[4 * k] = r.x; out[4 * k + 1] = r.y; out[4 * k + 2] = r.z; out[4 * k + 3] = 0.0f;
    }
}

static int shadow_visible(const orc_scene *s, f3 origin, f3 dir, orc_counters *cnt) {
    orc_ray r;
    r.o = origin; r.d = dir; r.tmin = 0.0f; r.tmax = 1000.0f;
    ray_setup(&r);
    cnt->rays_shadow++;
    int occ = any_hit(s, &r);
    if (g_crosscheck && s->use_bvh) {
        orc_scene tmp = *s;
        tmp.use_bvh = 0;
        if (any_hit(&tmp, &r) != occ) record_bad(&r, 2);
    }
    return !occ;
}

/* trace_radiance (texture_space_path_tracer.slang:57-84) */
static f3 trace_radiance(const orc_scene *s, const orc_light *L, f3 origin, f3 dir, uint32_t max_bounces,
                         orc_rng *rng, orc_counters *cnt) {
    f3 T = mk(1.0f, 1.0f, 1.0f), rad = mk(0.0f, 0.0f, 0.0f);
    for (uint32_t bounce = 0; bounce < max_bounces; ++bounce) {
        orc_ray r;
        r.o = origin; r.d = dir; r.tmin = 0.0f; r.tmax = 1000.0f;
        ray_setup(&r);
        cnt->rays_extend++;
        orc_hit h = closest_hit(s, &r, &cnt->tri_tests);
        if (g_crosscheck && s->use_bvh) {
            orc_scene tmp = *s;
            tmp.use_bvh = 0;
            orc_hit b = closest_hit(&tmp, &r, NULL);
            if (b.valid != h.valid || b.id != h.id || b.t != h.t) record_bad(&r, 1);
        }
        if (!h.valid) {
            rad = add(rad, mul(T, miss_radiance(L, dir)));
            return rad;
        }
        orc_shading sd;
        shading_data(s, s->tri_inst[h.id], s->tri_prim[h.id], h.b1, h.b2, neg(dir), &sd);
        rad = add(rad, mul(T, sd.emissive));
        /* sample_directional_light (scene.slang:954-995) */
        float ux = rng_1d(rng), uy = rng_1d(rng);
        f3 ld = cone_sample(L->sun_dir, L->cos_half, ux, uy);
        float ndl = fmaxf(0.0f, dot(sd.frame.n, ld));
        if (L->sun_on && ndl > 0.0f && shadow_visible(s, sd.origin_front, ld, cnt)) {
            f3 f = muls(sd.diffuse, ORC_INV_PI);
            f3 li = muls(mul(f, L->sun_illum), ndl);
            rad = add(rad, mul(T, li));
        }
        float ex = rng_1d(rng), ey = rng_1d(rng);
        f3 sdir;
        float pdf;
        cosine_sample(ex, ey, &sdir, &pdf);
        if (pdf <= 1e-6f) return rad;
        f3 f = muls(sd.diffuse, ORC_INV_PI);
        float az = fabsf(sdir.z);
        T = mul(T, mk(f.x * az / pdf, f.y * az / pdf, f.z * az / pdf));
        /* Russian roulette (no counterpart in the reference, whose paths run
         * to max_bounces): only a path that would take another bounce draws;
         * it survives with probability p = min(1, max component of T) and
         * carries T / p on, which keeps the estimate unbiased. */
        if (s->roulette && bounce + 1 < max_bounces) {
            float p = fminf(1.0f, fmaxf(fmaxf(T.x, T.y), T.z));
            float u = rng_1d(rng);
            if (!(u < p)) return rad;
            T = mk(T.x / p, T.y / p, T.z / p);
        }
        origin = sd.origin_front;
        dir = to_global(&sd.frame, sdir);
    }
    return rad;
}

/* estimate_surface_irradiance (texture_space_path_tracer.slang:86-117) */
static f3 estimate_irradiance(const orc_scene *s, const orc_light *L, uint32_t instance, uint32_t prim, float b1,
                              float b2, f3 view, uint32_t spp, uint32_t max_bounces, orc_rng *rng,
                              orc_counters *cnt) {
    orc_shading sd;
    shading_data(s, instance, prim, b1, b2, view, &sd);
    f3 acc = mk(0.0f, 0.0f, 0.0f);
    for (uint32_t k = 0; k < spp; ++k) {
        f3 sample = mk(0.0f, 0.0f, 0.0f);
        /* sample_directional_irradiance (scene.slang:940-952) */
        float ux = rng_1d(rng), uy = rng_1d(rng);
        f3 ld = cone_sample(L->sun_dir, L->cos_half, ux, uy);
        float ndl = fmaxf(0.0f, dot(sd.frame.n, ld));
        if (L->sun_on && ndl > 0.0f && shadow_visible(s, sd.origin_front, ld, cnt))
            sample = muls(L->sun_illum, ndl);
        float ex = rng_1d(rng), ey = rng_1d(rng);
        f3 sdir;
        float pdf;
        cosine_sample(ex, ey, &sdir, &pdf);
        if (pdf > 1e-6f) {
            f3 rad = trace_radiance(s, L, sd.origin_front, to_global(&sd.frame, sdir), max_bounces, rng, cnt);
            float az = fabsf(sdir.z);
            sample = add(sample, mk(rad.x * az / pdf, rad.y * az / pdf, rad.z * az / pdf));
        }
        acc = add(acc, sample);
    }
    float fs = (float)spp;
    return mk(acc.x / fs, acc.y / fs, acc.z / fs);
}

/* store_accumulated_irradiance (texture_space_path_tracer.slang:119-138) */
static void store_accumulated(float *payload, uint32_t slot, f3 e, uint32_t reset) {
    float *p = payload + 4 * (size_t)slot;
    uint32_t prev_count;
    memcpy(&prev_count, &p[3], 4);
    if (reset) prev_count = 0;
    f3 prev = prev_count == 0 ? mk(0.0f, 0.0f, 0.0f) : mk(p[0], p[1], p[2]);
    uint32_t next = prev_count + 1 < ORC_MAX_SAMPLES ? prev_count + 1 : ORC_MAX_SAMPLES;
    float w = 1.0f / (float)(prev_count + 1);
    p[0] = prev.x + w * (e.x - prev.x);
    p[1] = prev.y + w * (e.y - prev.y);
    p[2] = prev.z + w * (e.z - prev.z);
    memcpy(&p[3], &next, 4);
}

typedef struct {
    const orc_scene *s;
    const uint32_t *faces, *instances, *sides;
    uint32_t instance_count;
    float *payload;
    uint32_t iteration, spp, max_bounces, reset;
    const uint64_t *work;   /* flat front texel indices to process */
    uint64_t work_count;
    const uint32_t *front_prefix;
    orc_counters counters;
    uint64_t samples;
} orc_job;

static uint32_t instance_of(const uint32_t *front_prefix, uint32_t n, uint64_t w) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) / 2;
        if ((uint64_t)front_prefix[mid] <= w) lo = mid; else hi = mid - 1;
    }
    return lo;
}

/* compute_main (texture_space_path_tracer.slang:140-216) for one front texel */
static void trace_texel(orc_job *j, uint32_t instance, uint32_t local) {
    const uint32_t *inst = j->instances + 4 * (size_t)instance;
    const uint32_t *side = j->sides + 2 * (size_t)instance;
    orc_hit_info h;
    orc_rng rng;
    const uint32_t *face;
    if (!texel_raygen(j->faces, j->instances, instance, local, j->iteration, j->spp, j->s->rng_pcg, &h, &rng,
                      &face))
        return;
    uint32_t spp = j->spp > 1 ? j->spp : 1;
    orc_light L = load_light_lut(j->s->d.lighting, j->s->sky_lut);
    f3 n = world_flat_normal(j->s, instance, h.prim);
    f3 e = estimate_irradiance(j->s, &L, instance, h.prim, h.b1, h.b2, n, spp, j->max_bounces, &rng, &j->counters);
    uint32_t global_texel = inst[2] + local;
    store_accumulated(j->payload, global_texel, e, j->reset);
    j->samples++;
    if (side[1] & 1u) {
        f3 eb = estimate_irradiance(j->s, &L, instance, h.prim, h.b1, h.b2, neg(n), spp, j->max_bounces, &rng,
                                    &j->counters);
        store_accumulated(j->payload, side[0] + local, eb, j->reset);
        j->samples++;
    }
}

static void *job_main(void *arg) {
    orc_job *j = (orc_job *)arg;
    for (uint64_t k = 0; k < j->work_count; ++k) {
        uint64_t w = j->work[k];
        uint32_t i = instance_of(j->front_prefix, j->instance_count, w);
        trace_texel(j, i, (uint32_t)(w - j->front_prefix[i]));
    }
    return NULL;
}

/* Trace the listed flat front texels (all of them when work == NULL).
 * counters_out: [rays_extend, rays_shadow, tri_tests, slot_updates]. */
ORC_API int orc_trace_texels(const orc_scene *s, const uint32_t *faces, const uint32_t *instances,
                             const uint32_t *sides, uint32_t instance_count, float *payload, uint32_t iteration,
                             uint32_t spp, uint32_t max_bounces, uint32_t reset, const uint64_t *work,
                             uint64_t work_count, int threads, uint64_t *counters_out) {
    if (instance_count == 0) return -1;
    uint32_t *prefix = (uint32_t *)malloc(sizeof(uint32_t) * instance_count);
    uint64_t total = 0;
    for (uint32_t i = 0; i < instance_count; ++i) {
        prefix[i] = (uint32_t)total;
        total += instances[4 * (size_t)i + 3];
    }
    uint64_t *all = NULL;
    if (!work) {
        all = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(total ? total : 1));
        for (uint64_t w = 0; w < total; ++w) all[w] = w;
        work = all;
        work_count = total;
    }
    if (threads < 1) threads = 1;
    orc_job *jobs = (orc_job *)calloc((size_t)threads, sizeof(orc_job));
    pthread_t *tid = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    /* interleaved chunks of 64 texels balance uneven path lengths */
    uint64_t *lists = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(work_count ? work_count : 1));
    uint64_t *starts = (uint64_t *)calloc((size_t)threads + 1, sizeof(uint64_t));
    {
        uint64_t pos = 0;
        for (int t = 0; t < threads; ++t) {
            starts[t] = pos;
            for (uint64_t c = (uint64_t)t * 64; c < work_count; c += (uint64_t)threads * 64)
                for (uint64_t k = c; k < c + 64 && k < work_count; ++k) lists[pos++] = work[k];
        }
        starts[threads] = pos;
    }
    for (int t = 0; t < threads; ++t) {
        orc_job *j = &jobs[t];
        j->s = s; j->faces = faces; j->instances = instances; j->sides = sides;
        j->instance_count = instance_count; j->payload = payload; j->iteration = iteration;
        j->spp = spp; j->max_bounces = max_bounces; j->reset = reset;
        j->work = lists + starts[t]; j->work_count = starts[t + 1] - starts[t];
        j->front_prefix = prefix;
        if (threads == 1) job_main(j);
        else pthread_create(&tid[t], NULL, job_main, j);
    }
    if (threads > 1) for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
    if (counters_out) {
        memset(counters_out, 0, sizeof(uint64_t) * 4);
        for (int t = 0; t < threads; ++t) {
            counters_out[0] += jobs[t].counters.rays_extend;
            counters_out[1] += jobs[t].counters.rays_shadow;
            counters_out[2] += jobs[t].counters.tri_tests;
            counters_out[3] += jobs[t].samples;
        }
    }
    free(lists); free(starts); free(jobs); free(tid); free(prefix); free(all);
    return 0;
}

/* ------------------------------------------------------------------ */
/* primary-ray resolve (mesh_colors_resolve.slang:29-73,               */
/* mesh_colors_frozen_resolve.slang:16-135, camera.slang:31-66)        */
/* ------------------------------------------------------------------ */
static f3 load_slot(const float *payload, const uint32_t *packed, uint32_t slot) {
    if (packed) {
        float rgb[3];
        orc_unpack_rgb9e5(packed[slot], rgb);
        return mk(rgb[0], rgb[1], rgb[2]);
    }
    const float *p = payload + 4 * (size_t)slot;
    return mk(p[0], p[1], p[2]);
}

static inline uint32_t side_texel(uint32_t front, const uint32_t *inst, const uint32_t *side, int back) {
    return back ? side[0] + (front - inst[2]) : front;
}

/* mesh_colors_sample_irradiance (mesh_colors.slang:365-426) */
static f3 sample_irradiance(const float *payload, const uint32_t *packed, f3 b, const uint32_t *inst,
                            const uint32_t *face, const uint32_t *side, int back) {
    uint32_t R = face[1];
    float fr = (float)R;
    f3 sc = mk(saturate(b.x) * fr, saturate(b.y) * fr, saturate(b.z) * fr);
    uint32_t i = (uint32_t)sc.x, j = (uint32_t)sc.y;
    if (i > R) i = R;
    if (j > R - i) j = R - i;
    f3 w = mk(fracf(sc.x), fracf(sc.y), fracf(sc.z));
    int upper = fabsf(w.x + w.y + w.z - 2.0f) <= 1e-3f;
    if (upper) w = mk(1.0f - w.x, 1.0f - w.y, 1.0f - w.z);
    if (w.x + w.y + w.z <= 1e-6f || i >= R || j >= R - i)
        return load_slot(payload, packed, side_texel(lattice_address(i, j, inst, face), inst, side, back));
    uint32_t ta, tb, tc;
    if (upper) {
        ta = lattice_address(i, j + 1, inst, face);
        tb = lattice_address(i + 1, j, inst, face);
        tc = lattice_address(i + 1, j + 1, inst, face);
    } else {
        ta = lattice_address(i + 1, j, inst, face);
        tb = lattice_address(i, j + 1, inst, face);
        tc = lattice_address(i, j, inst, face);
    }
    f3 A = load_slot(payload, packed, side_texel(ta, inst, side, back));
    f3 B = load_slot(payload, packed, side_texel(tb, inst, side, back));
    f3 C = load_slot(payload, packed, side_texel(tc, inst, side, back));
    return add(add(muls(A, w.x), muls(B, w.y)), muls(C, w.z));
}

/* The lookup alone, for the checks that pin it (a linear field resolves to
 * the query's barycentrics, a lattice point to its own texel): bary float[3],
 * inst uint32[4], face uint32[2], side uint32[2] rows of the layout tables;
 * payload float[][4] or packed RGB9E5 words (payload NULL); out float[3]. */
ORC_API void orc_sample_irradiance(const float *payload, const uint32_t *packed, const float *bary,
                                   const uint32_t *inst, const uint32_t *face, const uint32_t *side, int back,
                                   float *out) {
    f3 e = sample_irradiance(payload, packed, mk(bary[0], bary[1], bary[2]), inst, face, side, back);
    out[0] = e.x;
    out[1] = e.y;
    out[2] = e.z;
}

/* sky plus sun disk for a primary miss (scene.slang:431-527) */
static f3 sky_with_sun_disk(const orc_light *L, const float *lighting, f3 dir) {
    f3 out = miss_radiance(L, dir);
    /* ground test: ray from (0, 6.3602, 0) against the planet sphere of radius 6.36 */
    const float ground = 6.360f, vy = 6.360f + 0.0002f;
    f3 oc = mk(0.0f, vy, 0.0f);
    float a = dot(dir, dir), bq = 2.0f * dot(oc, dir), c = dot(oc, oc) - ground * ground;
    float disc = bq * bq - 4.0f * a * c;
    int hits_ground = 0;
    if (disc >= 0.0f) {
        float sq = sqrtf(disc);
        float t0 = (-bq - sq) / (2.0f * a), t1 = (-bq + sq) / (2.0f * a);
        hits_ground = t0 > 0.0f || t1 > 0.0f;
    }
    if (!hits_ground && L->sun_on) {
        float vdl = dot(dir, L->sun_dir);
        if (vdl > L->cos_half) {
            float solid = 2.0f * ORC_PI * (1.0f - L->cos_half);
            f3 outer = mk(lighting[4], lighting[5], lighting[6]);
            f3 trans = mk(lighting[8], lighting[9], lighting[10]);
            float soft = saturate(2.0f * (vdl - L->cos_half) / (1.0f - L->cos_half));
            f3 disk = muls(mul(trans, mk(outer.x / solid, outer.y / solid, outer.z / solid)), soft);
            out = add(out, disk);
        }
    }
    return out;
}

/* camera: float[16] as Camera.pack(); writes out[H][W][4] */
ORC_API void orc_resolve(const orc_scene *s, const uint32_t *faces, const uint32_t *instances,
                         const uint32_t *sides, const float *payload, const uint32_t *packed, const float *camera,
                         uint32_t width, uint32_t height, uint32_t row_begin, uint32_t row_end, float *out) {
    orc_light L = load_light_lut(s->d.lighting, s->sky_lut);
    f3 pos = mk(camera[0], camera[1], camera[2]);
    f3 U = mk(camera[4], camera[5], camera[6]), Vv = mk(camera[8], camera[9], camera[10]);
    f3 W = mk(camera[12], camera[13], camera[14]);
    float jx = camera[3], jy = camera[7], near = camera[11], far = camera[15];
    f3 wn = normalize(W);
    for (uint32_t y = row_begin; y < row_end && y < height; ++y)
        for (uint32_t x = 0; x < width; ++x) {
            float px = ((float)x + 0.5f) / (float)width;
            float py = ((float)y + 0.5f) / (float)height;
            px = px + -jx;
            py = py + jy;
            float nx = 2.0f * px + -1.0f, ny = -2.0f * py + 1.0f;
            f3 d = normalize(add(add(muls(U, nx), muls(Vv, ny)), W));
            float inv_cos = 1.0f / dot(wn, d);
            orc_ray r;
            r.o = pos; r.d = d; r.tmin = near * inv_cos; r.tmax = far * inv_cos;
            ray_setup(&r);
            orc_hit h = closest_hit(s, &r, NULL);
            f3 color;
            if (!h.valid) {
                color = sky_with_sun_disk(&L, s->d.lighting, d);
            } else {
                uint32_t instance = s->tri_inst[h.id], prim = s->tri_prim[h.id];
                f3 fn = world_flat_normal(s, instance, prim);
                int back = dot(fn, neg(d)) <= 0.0f;
                const uint32_t *inst = instances + 4 * (size_t)instance;
                const uint32_t *side = sides + 2 * (size_t)instance;
                if (back && !(side[1] & 1u)) {
                    color = mk(0.0f, 0.0f, 0.0f);
                } else {
                    const uint32_t *face = faces + 2 * (size_t)(inst[0] + prim);
                    f3 bc = mk(1.0f - h.b1 - h.b2, h.b1, h.b2);
                    f3 E = sample_irradiance(payload, packed, bc, inst, face, side, back);
                    orc_shading sd;
                    shading_data(s, instance, prim, h.b1, h.b2, neg(d), &sd);
                    color = add(sd.emissive, mul(muls(sd.diffuse, ORC_INV_PI), E));
                }
            }
            float *o = out + 4 * ((size_t)y * width + x);
            o[0] = color.x; o[1] = color.y; o[2] = color.z; o[3] = 1.0f;
        }
}

/* ------------------------------------------------------------------ */
/* accumulator and tone map (accumulator.slang:3-29, tone_mapper.slang) */
/* ------------------------------------------------------------------ */
ORC_API void orc_accumulate(const float *input, float *accum, float *output, uint64_t n, int reset) {
    for (uint64_t i = 0; i < n; ++i) {
        const float *in = input + 4 * i;
        float *a = accum + 4 * i;
        float v[4] = {0, 0, 0, 0};
        if (!reset) memcpy(v, a, sizeof(v));
        v[0] += in[0]; v[1] += in[1]; v[2] += in[2]; v[3] += 1.0f;
        memcpy(a, v, sizeof(v));
        float *o = output + 4 * i;
        o[0] = v[0] / v[3]; o[1] = v[1] / v[3]; o[2] = v[2] / v[3]; o[3] = 1.0f;
    }
}

static f3 lpm_tonemap(f3 color, float hdr_max, float exposure, float contrast) {
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
    f3 ratio = muls(color, rcp);
    float luma = dot(color, luma709);
    luma = powf(fmaxf(luma, 0.0f), contrast);
    luma = luma / (luma * bias_x + bias_y);
    float luma_ratio = dot(ratio, luma709);
    float ratio_scale = saturate(luma / fmaxf(luma_ratio, 0.00001f));
    f3 out = mk(saturate(ratio.x * ratio_scale), saturate(ratio.y * ratio_scale), saturate(ratio.z * ratio_scale));
    f3 cap = mul(crosstalk, mk(1.0f - out.x, 1.0f - out.y, 1.0f - out.z));
    float luma_add = saturate(luma - dot(out, luma709));
    float cap_luma = dot(cap, luma709);
    if (cap_luma > 0.00001f) {
        float t = luma_add / cap_luma;
        out = mk(saturate(out.x + t * cap.x), saturate(out.y + t * cap.y), saturate(out.z + t * cap.z));
    }
    luma_add = saturate(luma - dot(out, luma709));
    f3 rl = mk(1.0f / luma709.x, 1.0f / luma709.y, 1.0f / luma709.z);
    out = mk(saturate(out.x + luma_add * rl.x), saturate(out.y + luma_add * rl.y), saturate(out.z + luma_add * rl.z));
    return out;
}

static float rec709(float x) { return x < 0.018f ? 4.5f * x : 1.099f * powf(x, 0.45f) - 0.099f; }

ORC_API void orc_tonemap(const float *input, float *output, uint64_t n, float exposure) {
    for (uint64_t i = 0; i < n; ++i) {
        const float *in = input + 4 * i;
        f3 t = lpm_tonemap(mk(in[0], in[1], in[2]), 12.0f, exposure, 1.0f);
        float *o = output + 4 * i;
        o[0] = rec709(t.x); o[1] = rec709(t.y); o[2] = rec709(t.z); o[3] = 1.0f;
    }
}

ORC_API void orc_sincos_2pi(const float *u, uint64_t n, float *s, float *c) {
    for (uint64_t i = 0; i < n; ++i) sincos_2pi(u[i], &s[i], &c[i]);
}
