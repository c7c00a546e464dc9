// Host-side wide-BVH builder for the GPU traversal kernels.
//
// Replaces the reference's two CPU builders:
//   QBVHScene::Build (src/liblightmetrica/scene.qbvh.cpp:453-591): 12-bin SAH on the
//     longest centroid axis, two binary levels per 4-wide node, <= 64-triangle leaves;
//   BVHScene::Build (src/liblightmetrica/scene.bvh.cpp:202-388): pbrt-style 12-bucket SAH.
// This builder bins all three axes (16 bins by default) with a traversal-cost term,
// then collapses the binary tree greedily (largest surface area first) into 4-wide
// nodes, quantizes each node's child boxes to 8 bits per plane relative to the node
// (64-byte nodes, two per 128-byte line), and stores leaves of at most 16 triangles as
// contiguous 48-byte vertex records.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace lmgpu {

// A 4-wide node with full-precision child boxes (the builder's intermediate form).
struct BoxNode4 {
    float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
    uint32_t child[4];
};

// The device node, 64 bytes, read as four 16-byte loads:
//   origin.xyz + packed exponents | lo_x hi_x lo_y hi_y | lo_z hi_z child0 child1 | child2 child3 pad
// Child i's box plane on axis a is origin[a] + q * 2^(e[a] - 127), q the byte i of the plane's
// word; planes are rounded outward and widened by one more quantum (the grid starts one
// quantum below the node box and spans it with one quantum to spare above).
struct alignas(64) Node4 {
    float origin[3];
    uint32_t exps;  // e_x | e_y << 8 | e_z << 16 (biased 8-bit exponents of the quanta)
    uint32_t q_lo_x, q_hi_x, q_lo_y, q_hi_y;
    uint32_t q_lo_z, q_hi_z;
    uint32_t child[4];
    uint32_t pad[2];
};
static_assert(sizeof(Node4) == 64, "Node4 is 64 bytes");

// Child reference encoding.
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmptyChild = 0xffffffffu;
constexpr uint32_t kMaxLeafCount = 16;          // 4-bit count field
constexpr uint32_t kMaxLeafFirst = 0x07ffffffu;  // 27-bit first-record field
inline uint32_t make_leaf_ref(uint32_t first, uint32_t count) {
    return kLeafBit | ((count - 1u) << 27) | first;
}

// Triangle record in leaf order: v0.w carries the scene-order triangle index.
struct alignas(16) TriRecord {
    float v0[4];
    float v1[4];
    float v2[4];
};
static_assert(sizeof(TriRecord) == 48, "TriRecord is 48 bytes");

struct BuildParams {
    uint32_t max_leaf_size = 4;
    uint32_t bins = 16;
    float traversal_cost = 1.2f;
    int num_threads = 0;
};

struct WideBVH {
    std::vector<Node4> nodes;
    std::vector<BoxNode4> box_nodes;  // the same tree with unquantized boxes
    std::vector<TriRecord> tris;
    uint32_t num_leaves = 0;
    uint32_t max_depth = 0;  // inner-node levels on the deepest root-to-leaf path
    double sah_cost = 0.0;
    double build_seconds = 0.0;
};

// positions: 3 floats per vertex; faces: 3 indices per triangle.
// Returns false with `err` set when the input cannot be built.
bool build_wide_bvh(const float* positions, uint32_t num_vertices, const uint32_t* faces, uint32_t num_triangles,
                    const BuildParams& params, WideBVH& out, std::string& err);

}  // namespace lmgpu
