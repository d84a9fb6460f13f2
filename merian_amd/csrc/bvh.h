// 4-wide BVH for the gfx950 traversal kernels: layout and host builder.
//
// Replaces the driver acceleration-structure build of the reference
// (src/merian/vk/raytrace/as_builder_blas.cpp:137, as_builder_tlas.cpp:131). Hit results do not
// depend on the tree shape: traversal is conservative and ties at equal t go to the lowest
// triangle id, so any correct tree returns the brute-force closest hit.
#pragma once

#include <stdint.h>

#include <vector>

namespace mrt {

// One node = one 128-byte line: child boxes as structure-of-arrays so a lane reads them with
// six 16-byte loads (lo/hi of one axis 16 bytes apart, so the ray's direction octant picks
// the near and far plane by load address), then the four children as ready traversal entries.
struct alignas(128) BVH4Node {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    uint32_t entry[4];  // inner child: node index; leaf child: leaf entry (below); empty: 0xFFFFFFFF
    uint32_t _pad[4];
};
static_assert(sizeof(BVH4Node) == 128, "BVH4Node must be one 128-byte line");

// Triangle record in leaf order: the Moller-Trumbore operands, three 16-byte loads.
//   a = (v0.xyz, id bits)   b = (e1.xyz, flags bits)   c = (e2.xyz, 0)
// flags: bit 0 = back-face culling on, bits 8..15 = 8-bit instance mask.
struct alignas(16) TriRecord {
    float a[4], b[4], c[4];
};

// Leaf stack entries: bit 31 set, first record in bits 3..30, (count - 1) in bits 0..2.
constexpr uint32_t LEAF_FLAG = 0x80000000u;
constexpr uint32_t EMPTY_ENTRY = 0xFFFFFFFFu;
inline uint32_t leaf_entry(uint32_t first, uint32_t count) { return LEAF_FLAG | (first << 3) | (count - 1u); }
constexpr int MAX_LEAF = 8; // most triangles per leaf the encoding allows (the build default is 4)

struct BVH4 {
    std::vector<BVH4Node> nodes;   // nodes[0] is the root
    std::vector<TriRecord> tris;   // leaf order
    uint32_t depth = 0;            // longest root-to-leaf path in nodes
    uint32_t root_leaf = 0;        // nonzero: the whole scene is one leaf entry (tiny scenes)
};

// v0/e1/e2: per-triangle float[3] operands exactly as the intersection test uses them
// (e1 = v1 - v0, e2 = v2 - v0 rounded in float). flags as in TriRecord.
// 8-wide node with child boxes quantised to 8 bits on a per-node power-of-two grid (80 bytes,
// five 16-byte loads): child box = fma(q, 2^e, origin) per plane, every decoded box containing
// the padded child box exactly as the BVH4 stores it. Inner children are stored consecutively
// from child_base in slot order; the triangles of the leaf children consecutively from tri_base.
// One code byte per slot makes the slot's traversal entry a select and an add:
//   leaf:  (first triangle - tri_base) << 3 | (count - 1)   (<= BVH8_LEAF_CODE_MAX)
//   empty: BVH8_EMPTY
//   inner: BVH8_INNER + rank among the inner children
// so entry = (inner ? child_base - BVH8_INNER : LEAF | tri_base << 3) + code.
constexpr uint8_t BVH8_EMPTY = 0xF0, BVH8_INNER = 0xF8;
constexpr int BVH8_MAX_LEAF = 4; // 7 leaves of 4 before a slot: offset 28, code 227
constexpr uint32_t BVH8_LEAF_CODE_MAX = 28u << 3 | 3u;
struct alignas(16) BVH8Node {
    float px, py, pz;       // grid origin
    int8_t ex, ey, ez;      // grid step 2^e per axis
    uint8_t inner_mask;     // bit k: slot k is an inner child
    uint32_t child_base;    // node index of the first inner child
    uint32_t tri_base;      // first triangle record of the leaf children
    uint8_t code[8];        // per slot, as above
    uint8_t qlox[8], qhix[8], qloy[8], qhiy[8], qloz[8], qhiz[8]; // empty slot: lo 255, hi 0
};
static_assert(sizeof(BVH8Node) == 80, "BVH8Node must be 80 bytes");
static_assert(BVH8_LEAF_CODE_MAX < BVH8_EMPTY, "leaf codes below the sentinels");

struct BVH8 {
    std::vector<BVH8Node> nodes; // nodes[0] is the root
    std::vector<TriRecord> tris; // per node, its leaf children's triangles in slot order
    uint32_t depth = 0;
};

// max_leaf: most triangles per leaf (1..MAX_LEAF); traversal_cost: SAH cost of one node
// visit relative to one triangle test.
BVH8 build_bvh8(const float* v0, const float* e1, const float* e2, const uint32_t* flags, uint32_t n,
                int threads, int max_leaf = 4, float traversal_cost = 1.0f);

BVH4 build_bvh4(const float* v0, const float* e1, const float* e2, const uint32_t* flags, uint32_t n,
                int threads, int max_leaf = 4, float traversal_cost = 1.0f);

} // namespace mrt
