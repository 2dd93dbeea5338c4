// Host-side BVH construction for the MI355X traversal kernels.
//
// A binned-SAH binary BVH is built over world-space triangles and collapsed
// into a 4-wide tree whose nodes are 64-byte records with 8-bit quantized
// child boxes (the compressed-wide-BVH idea of Ylitie, Karras and Laine,
// HPG 2017, reduced to 4 children so that one node is exactly one 64-byte
// line).  The reference has no software BVH: it builds a BLAS per mesh and a
// TLAS through the vendor API (scene.py:440-533) and traverses them in
// hardware.
#pragma once

#include <cstdint>
#include <vector>

namespace nanbvh {

// Node word layout (16 x uint32 = 64 bytes):
//   w0..w2  origin.xyz (float)
//   w3      exp_x | exp_y << 8 | exp_z << 16 (int8 each) | valid_mask << 24
//   w4..w6  qlo_x, qlo_y, qlo_z: one byte per child
//   w7..w9  qhi_x, qhi_y, qhi_z: one byte per child
//   w10..13 child entries
//   w14,15  reserved (0)
// Child decode: lo = origin + qlo * 2^exp (float add, exact product); the
// builder checks that every decoded box contains the child box under that
// exact float arithmetic.
// Child entry: internal -> node index (< 2^31);
//              leaf -> LEAF_BIT | (count - 1) << LEAF_COUNT_SHIFT | first triangle.
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kLeafCountShift = 27;
constexpr uint32_t kLeafFirstMask = (1u << kLeafCountShift) - 1u;
constexpr uint32_t kMaxLeafSize = 8;
constexpr uint32_t kNodeWords = 16;
// Triangle record: 12 floats = 48 bytes: v0.xyz v1.x | v1.yz v2.xy | v2.z id flags pad
// (flags: written 0 here; the scene sets 1 for a triangle whose material is alpha tested)
constexpr uint32_t kTriWords = 12;

struct BuildOptions {
    uint32_t bins = 16;
    uint32_t max_leaf = 4;
    uint32_t threads = 0;
};

struct Bvh4 {
    std::vector<uint32_t> nodes;      // node_count * kNodeWords
    std::vector<float> tris;          // triangle_count * kTriWords, leaf order
    uint32_t root_entry = 0;          // node 0 or a leaf entry for tiny scenes
    uint32_t node_count = 0;
    uint32_t leaf_count = 0;
    uint32_t max_stack = 0;           // worst-case traversal stack entries
    float sah_cost = 0.0f;
};

// world: triangle_count * 9 floats (v0, v1, v2), triangle id = index.
// Returns false and sets *error on failure.
bool build_bvh4(const float *world, uint32_t triangle_count, const BuildOptions &opt, Bvh4 *out,
                const char **error);

// Decoded child box as the kernels compute it (for tests and checks).
void decode_child(const uint32_t *node, int child, float lo[3], float hi[3]);

}  // namespace nanbvh
