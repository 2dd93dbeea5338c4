// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only;
// test infrastructure): the BVH builder of libnan_hip (quantization, leaf
// packing, stack-need computation) and the oracle (scene setup, BVH and
// brute-force hits, texel tracing, resolve, RGB9E5, atmosphere tables) on
// seeded random scenes, including degenerate triangles, mirrored and scaled
// instances and axis-parallel rays.  Besides the sanitizers' own checks it
// verifies the builder's invariants (every triangle in exactly one leaf,
// decoded child boxes contain their triangles, the stack bound holds for a
// full traversal) and that the oracle's BVH answers equal its brute force
// bit for bit.  Exit status 0 when every check passes.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../include/nan_hip.h"
#include "../nan_amd/csrc/nan_bvh.h"

extern "C" {
struct orc_scene;
orc_scene *orc_scene_create(const NanSceneDesc *desc, int use_bvh);
void orc_scene_destroy(orc_scene *s);
uint32_t orc_scene_triangle_count(const orc_scene *s);
void orc_scene_world_triangles(const orc_scene *s, float *out);
void orc_intersect(const orc_scene *s, const float *rays, uint64_t n, float *out_t, float *out_b, uint32_t *out_inst,
                   uint32_t *out_prim, uint32_t *out_occluded);
int orc_trace_texels(const orc_scene *s, const uint32_t *faces, const uint32_t *instances, const uint32_t *sides,
                     uint32_t instance_count, float *payload, uint32_t iteration, uint32_t spp, uint32_t max_bounces,
                     uint32_t reset, const uint64_t *work, uint64_t work_count, int threads, uint64_t *counters_out);
void orc_resolve(const orc_scene *s, const uint32_t *faces, const uint32_t *instances, const uint32_t *sides,
                 const float *payload, const uint32_t *packed, const float *camera, uint32_t width, uint32_t height,
                 uint32_t row_begin, uint32_t row_end, float *out);
void orc_pack_rgb9e5_array(const float *payload, uint32_t stride_floats, uint64_t n, uint32_t *out);
void orc_unpack_rgb9e5(uint32_t packed, float *rgb);
void orc_atmo_transmittance_lut(float *out);
void orc_atmo_multiscattering_lut(const float *tlut, float *out);
void orc_atmo_sky_view_lut(const float *tlut, const float *mslut, const float *sun_direction3, float *out);
}

static int g_failures = 0;

#define CHECK(cond, ...)                                \
    do {                                                \
        if (!(cond)) {                                  \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);          \
            std::fprintf(stderr, "\n");                 \
            ++g_failures;                               \
        }                                               \
    } while (0)

struct RandomScene {
    std::vector<float> vertices;  // [V][8]
    std::vector<uint32_t> indices, mesh_descs, instance_descs, materials, textures, texels;
    std::vector<float> transforms, inv_transposes, lighting;
    uint32_t triangles = 0;
};

static void identity(float *m) {
    std::memset(m, 0, 16 * sizeof(float));
    m[0] = m[5] = m[10] = m[15] = 1.0f;
}

static RandomScene make_scene(uint32_t seed) {
    std::mt19937 gen(seed);
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    RandomScene s;
    uint32_t meshes = 1 + seed % 3;
    for (uint32_t m = 0; m < meshes; ++m) {
        uint32_t v0 = (uint32_t)(s.vertices.size() / 8), i0 = (uint32_t)s.indices.size();
        uint32_t tris = 20 + (uint32_t)(gen() % 400);
        float scale = m == 2 ? 1e3f : (m == 1 ? 1e-2f : 3.0f);
        for (uint32_t t = 0; t < tris; ++t) {
            float cx = u(gen) * scale, cy = u(gen) * scale, cz = u(gen) * scale;
            for (int k = 0; k < 3; ++k) {
                float p[3] = {cx + 0.1f * scale * u(gen), cy + 0.1f * scale * u(gen), cz + 0.1f * scale * u(gen)};
                if (t % 17 == 0 && k == 2) {  // degenerate: repeat the first vertex
                    const float *q = &s.vertices[s.vertices.size() - 16];
                    p[0] = q[0], p[1] = q[1], p[2] = q[2];
                }
                if (t % 23 == 0) p[1] = cy;  // axis-aligned triangle
                float n[3] = {0.0f, 1.0f, 0.0f};
                float uv[2] = {3.0f * u(gen), 3.0f * u(gen)};  // wraps, negative too
                s.vertices.insert(s.vertices.end(), {p[0], p[1], p[2], n[0], n[1], n[2], uv[0], uv[1]});
            }
            for (uint32_t k = 0; k < 3; ++k) s.indices.push_back(3 * t + k);
        }
        s.mesh_descs.insert(s.mesh_descs.end(), {3 * tris, 3 * tris, v0, i0});
    }
    uint32_t instances = 1 + seed % 4;
    for (uint32_t i = 0; i < instances; ++i) {
        float m[16], it[16];
        identity(m);
        identity(it);
        float sx = (i % 2) ? -1.5f : 0.5f;  // odd instances mirrored
        m[0] = sx, m[5] = 2.0f, m[10] = 1.0f;
        m[3] = u(gen) * 5.0f, m[7] = u(gen) * 5.0f, m[11] = u(gen) * 5.0f;
        it[0] = 1.0f / sx, it[5] = 0.5f, it[10] = 1.0f;
        s.transforms.insert(s.transforms.end(), m, m + 16);
        s.inv_transposes.insert(s.inv_transposes.end(), it, it + 16);
        uint32_t mesh = i % meshes;
        s.instance_descs.insert(s.instance_descs.end(), {mesh, i % 2, i, sx < 0.0f ? 1u : 0u});
        s.triangles += s.mesh_descs[4 * mesh + 1] / 3;
    }
    auto fbits = [](float f) { uint32_t b; std::memcpy(&b, &f, 4); return b; };
    // material 0 single-sided, opaque, base-colour and normal textures; material 1
    // double-sided, emissive, blend alpha from the base-colour texture
    const uint32_t none = 0xFFFFFFFFu;
    s.materials = {fbits(0.7f), fbits(0.7f), fbits(0.7f), 0u, fbits(0.0f), fbits(0.0f), fbits(0.0f), 0u,
                   0u, none, 1u, fbits(0.5f),
                   fbits(0.5f), fbits(0.2f), fbits(0.1f), 1u, fbits(2.0f), fbits(1.0f), fbits(0.5f), 2u,
                   0u, none, none, fbits(0.5f)};
    // texture 0: 5x3 sRGB colour, texture 1: 2x7 normals
    s.textures = {0u, 5u, 3u, 1u, 15u, 2u, 7u, 0u};
    for (int k = 0; k < 29; ++k) s.texels.push_back((uint32_t)gen());
    float sun[16] = {0.3f, 0.8f, 0.2f, 0.9999f, 3.0f, 3.0f, 3.0f, 0.0f, 1.0f, 1.0f, 1.0f, 0.0f, 0.2f, 0.3f, 0.4f, 0.0f};
    float len = std::sqrt(sun[0] * sun[0] + sun[1] * sun[1] + sun[2] * sun[2]);
    sun[0] /= len, sun[1] /= len, sun[2] /= len;
    s.lighting.assign(sun, sun + 16);
    return s;
}

static NanSceneDesc desc_of(const RandomScene &s) {
    NanSceneDesc d{};
    d.vertices = s.vertices.data();
    d.vertex_count = (uint32_t)(s.vertices.size() / 8);
    d.indices = s.indices.data();
    d.index_count = (uint32_t)s.indices.size();
    d.mesh_descs = s.mesh_descs.data();
    d.mesh_count = (uint32_t)(s.mesh_descs.size() / 4);
    d.instance_descs = s.instance_descs.data();
    d.instance_count = (uint32_t)(s.instance_descs.size() / 4);
    d.transforms = s.transforms.data();
    d.inv_transposes = s.inv_transposes.data();
    d.transform_count = (uint32_t)(s.transforms.size() / 16);
    d.materials = s.materials.data();
    d.material_count = (uint32_t)(s.materials.size() / 12);
    d.lighting = s.lighting.data();
    d.textures = s.textures.data();
    d.texture_count = (uint32_t)(s.textures.size() / 4);
    d.texels = s.texels.data();
    d.texel_count = s.texels.size();
    return d;
}

// Builder invariants, and a full traversal (every child of every node) to
// check the stack bound.
static void check_bvh(const std::vector<float> &world, uint32_t tris, uint32_t max_leaf) {
    nanbvh::BuildOptions opt;
    opt.max_leaf = max_leaf;
    opt.threads = 2;
    nanbvh::Bvh4 b;
    const char *err = nullptr;
    bool ok = nanbvh::build_bvh4(world.data(), tris, opt, &b, &err);
    CHECK(ok, "build_bvh4: %s", err ? err : "?");
    if (!ok) return;
    CHECK(b.tris.size() == (size_t)tris * nanbvh::kTriWords, "triangle records %zu", b.tris.size());
    std::vector<int> seen(tris, 0);
    for (uint32_t k = 0; k < tris; ++k) {
        uint32_t id;
        std::memcpy(&id, &b.tris[(size_t)k * nanbvh::kTriWords + 9], 4);
        CHECK(id < tris, "triangle id %u", id);
        if (id < tris) seen[id]++;
    }
    for (uint32_t k = 0; k < tris; ++k) CHECK(seen[k] == 1, "triangle %u in %d leaves", k, seen[k]);
    if (b.root_entry & nanbvh::kLeafBit) return;
    // depth-first over every child, the order the kernels push entries
    struct Item { uint32_t entry; int depth; };
    std::vector<Item> stack{{0u, 0}};
    uint32_t deepest = 1;
    while (!stack.empty()) {
        deepest = std::max<uint32_t>(deepest, (uint32_t)stack.size());
        Item it = stack.back();
        stack.pop_back();
        CHECK(it.entry < b.node_count, "node %u of %u", it.entry, b.node_count);
        if (it.entry >= b.node_count) continue;
        const uint32_t *nd = &b.nodes[(size_t)it.entry * nanbvh::kNodeWords];
        uint32_t valid = nd[3] >> 24;
        for (int c = 0; c < 4; ++c) {
            if (!((valid >> c) & 1u)) continue;
            float lo[3], hi[3];
            nanbvh::decode_child(nd, c, lo, hi);
            uint32_t e = nd[10 + c];
            if (e & nanbvh::kLeafBit) {
                uint32_t first = e & nanbvh::kLeafFirstMask, count = ((e >> nanbvh::kLeafCountShift) & 7u) + 1u;
                CHECK(first + count <= tris, "leaf %u+%u", first, count);
                for (uint32_t k = first; k < first + count && k < tris; ++k)
                    for (int v = 0; v < 3; ++v)
                        for (int a = 0; a < 3; ++a) {
                            float x = b.tris[(size_t)k * nanbvh::kTriWords + 3 * v + a];
                            CHECK(x >= lo[a] && x <= hi[a], "vertex outside its decoded box");
                        }
            } else {
                stack.push_back({e, it.depth + 1});
            }
        }
    }
    CHECK(deepest <= b.max_stack + 1, "traversal needed %u stack entries, bound %u", deepest, b.max_stack);
}

int main() {
    std::mt19937 gen(7);
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    for (uint32_t seed = 0; seed < 12; ++seed) {
        RandomScene rs = make_scene(seed);
        NanSceneDesc d = desc_of(rs);
        orc_scene *bvh = orc_scene_create(&d, 1), *brute = orc_scene_create(&d, 0);
        uint32_t T = orc_scene_triangle_count(bvh);
        CHECK(T == rs.triangles, "oracle has %u triangles, scene %u", T, rs.triangles);
        std::vector<float> world((size_t)T * 9);
        orc_scene_world_triangles(bvh, world.data());
        for (uint32_t leaf : {1u, 4u, 8u}) check_bvh(world, T, leaf);
        // random rays, a quarter of them axis-parallel, some from vertices
        const uint32_t n = 3000;
        std::vector<float> rays((size_t)n * 8, 0.0f);
        for (uint32_t r = 0; r < n; ++r) {
            float *q = &rays[(size_t)r * 8];
            if (r % 3 == 0) {
                const float *v = &world[(size_t)(gen() % T) * 9];
                q[0] = v[0], q[1] = v[1] + 1e-3f, q[2] = v[2];
            } else {
                q[0] = 8.0f * u(gen), q[1] = 8.0f * u(gen), q[2] = 8.0f * u(gen);
            }
            float dx = u(gen), dy = u(gen), dz = u(gen);
            if (r % 4 == 0) {
                int axis = (int)(gen() % 3);
                dx = axis == 0 ? 1.0f : 0.0f, dy = axis == 1 ? -1.0f : 0.0f, dz = axis == 2 ? 1.0f : 0.0f;
            }
            float l = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (l == 0.0f) dx = 1.0f, l = 1.0f;
            q[4] = dx / l, q[5] = dy / l, q[6] = dz / l, q[7] = 1000.0f;
        }
        std::vector<float> t0(n), t1(n), b0(2 * n), b1(2 * n);
        std::vector<uint32_t> i0(n), i1(n), p0(n), p1(n), o0(n), o1(n);
        orc_intersect(bvh, rays.data(), n, t0.data(), b0.data(), i0.data(), p0.data(), o0.data());
        orc_intersect(brute, rays.data(), n, t1.data(), b1.data(), i1.data(), p1.data(), o1.data());
        int diff = 0;
        for (uint32_t r = 0; r < n; ++r)
            diff += std::memcmp(&t0[r], &t1[r], 4) != 0 || i0[r] != i1[r] || p0[r] != p1[r] || o0[r] != o1[r];
        CHECK(diff == 0, "scene %u: BVH and brute force differ on %d of %u rays", seed, diff, n);
        // a layout with resolution 2 on every face (6 texels), double-sided material 1
        uint32_t I = d.instance_count;
        std::vector<uint32_t> faces, inst, sides;
        uint32_t slots = 0, face_off = 0;
        for (uint32_t i = 0; i < I; ++i) {
            uint32_t ft = rs.mesh_descs[4 * rs.instance_descs[4 * i] + 1] / 3;
            for (uint32_t f = 0; f < ft; ++f) faces.insert(faces.end(), {6u * f, 2u});
            inst.insert(inst.end(), {face_off, ft, slots, 6u * ft});
            slots += 6u * ft;
            face_off += ft;
            bool two = rs.instance_descs[4 * i + 1] == 1;
            sides.insert(sides.end(), {two ? slots : 0xFFFFFFFFu, two ? 1u : 0u});
            if (two) slots += 6u * ft;
        }
        std::vector<float> payload((size_t)slots * 4, 0.0f);
        uint64_t counters[4] = {0, 0, 0, 0};
        for (uint32_t it = 0; it < 2; ++it) {
            int rc = orc_trace_texels(bvh, faces.data(), inst.data(), sides.data(), I, payload.data(), it, 2, 3,
                                      it == 0, nullptr, 0, 2, counters);
            CHECK(rc == 0, "trace_texels rc %d", rc);
        }
        bool finite = true;
        for (float x : payload) finite = finite && std::isfinite(x);
        CHECK(finite, "scene %u: non-finite payload", seed);
        std::vector<uint32_t> words(slots);
        orc_pack_rgb9e5_array(payload.data(), 4, slots, words.data());
        float cam[16] = {0.0f, 0.0f, -20.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f,
                         0.0f, 1.0f, 0.0f, 0.1f, 0.0f, 0.0f, 1.0f, 1000.0f};
        std::vector<float> img(24 * 16 * 4);
        orc_resolve(bvh, faces.data(), inst.data(), sides.data(), payload.data(), nullptr, cam, 24, 16, 0, 16,
                    img.data());
        orc_resolve(bvh, faces.data(), inst.data(), sides.data(), nullptr, words.data(), cam, 24, 16, 3, 9,
                    img.data());
        orc_scene_destroy(bvh);
        orc_scene_destroy(brute);
    }
    std::vector<float> tl(256 * 64 * 4), ms(32 * 32 * 4), sky(200 * 200 * 4);
    orc_atmo_transmittance_lut(tl.data());
    orc_atmo_multiscattering_lut(tl.data(), ms.data());
    float sun[3] = {0.0f, 0.5f, 0.8660254f};
    orc_atmo_sky_view_lut(tl.data(), ms.data(), sun, sky.data());
    float rgb[3];
    orc_unpack_rgb9e5(0xFFFFFFFFu, rgb);
    if (g_failures) {
        std::fprintf(stderr, "%d checks failed\n", g_failures);
        return 1;
    }
    std::printf("sanitize_check ok: 12 scenes (BVH invariants, 3000 rays each, texels, resolve), atmosphere tables\n");
    return 0;
}
