// CPU baseline / cross-check harness around the reference's vendored tinybvh
// (vertex_baker/cpp/third_party/tinybvh/tiny_bvh.h, v1.8.0 @ 0e45842, MIT),
// compiled where it lies in the reference tree by `make -C oracle ref`.
// TEST INFRASTRUCTURE ONLY: occlusion cross-checks and the CPU baseline.
//
// Workload shape follows the reference's visibility tracer
// (vertex_baker/cpp/tinybvh_visibility.cpp:145-230): a binned-SAH BVH::Build
// over the world-space triangles, IsOccluded per ray, one worker per core.
#define TINYBVH_IMPLEMENTATION
#include "tiny_bvh.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <thread>
#include <vector>

struct RefBvh {
    std::vector<tinybvh::bvhvec4> verts;
    tinybvh::BVH bvh;
    double build_ms = 0.0;
};

extern "C" {

// world: [T][9] floats
__attribute__((visibility("default"))) void *ref_bvh_create(const float *world, uint32_t tris) {
    RefBvh *r = new RefBvh();
    r->verts.resize((size_t)tris * 3);
    for (size_t i = 0; i < (size_t)tris * 3; ++i)
        r->verts[i] = tinybvh::bvhvec4(world[3 * i], world[3 * i + 1], world[3 * i + 2], 0.0f);
    auto t0 = std::chrono::steady_clock::now();
    r->bvh.Build(r->verts.data(), tris);
    r->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return r;
}

__attribute__((visibility("default"))) double ref_bvh_build_ms(void *h) { return ((RefBvh *)h)->build_ms; }

__attribute__((visibility("default"))) void ref_bvh_destroy(void *h) { delete (RefBvh *)h; }

// rays: [n][8] origin, tmin (ignored: tinybvh starts at 0), dir, tmax.
// Writes occluded[n]; returns trace milliseconds.
__attribute__((visibility("default"))) double ref_bvh_occluded(void *h, const float *rays, uint64_t n,
                                                                uint32_t *occluded, int threads) {
    RefBvh *r = (RefBvh *)h;
    if (threads < 1) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    auto work = [&](uint64_t begin, uint64_t end) {
        for (uint64_t i = begin; i < end; ++i) {
            const float *q = rays + 8 * i;
            tinybvh::Ray ray(tinybvh::bvhvec3(q[0], q[1], q[2]), tinybvh::bvhvec3(q[4], q[5], q[6]), q[7]);
            occluded[i] = r->bvh.IsOccluded(ray) ? 1u : 0u;
        }
    };
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work, n * t / threads, n * (t + 1) / threads);
    work(0, n / threads);
    for (auto &th : pool) th.join();
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // extern "C"
