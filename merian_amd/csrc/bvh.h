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
// six 16-byte loads, then the four child references and leaf sizes.
struct alignas(128) BVH4Node {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    int32_t child[4];   // inner child: node index; leaf child: first triangle record; empty: -1
    uint32_t counts;    // 8 bits per child: triangles in a leaf child, 0 for inner / empty
    uint32_t _pad[3];
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
constexpr int MAX_LEAF = 8; // most triangles per leaf the encoding allows (the build default is 4)

struct BVH4 {
    std::vector<BVH4Node> nodes;   // nodes[0] is the root
    std::vector<TriRecord> tris;   // leaf order
    uint32_t depth = 0;            // longest root-to-leaf path in nodes
    uint32_t root_leaf = 0;        // nonzero: the whole scene is one leaf entry (tiny scenes)
};

// v0/e1/e2: per-triangle float[3] operands exactly as the intersection test uses them
// (e1 = v1 - v0, e2 = v2 - v0 rounded in float). flags as in TriRecord.
// max_leaf: most triangles per leaf (1..MAX_LEAF); traversal_cost: SAH cost of one node
// visit relative to one triangle test.
BVH4 build_bvh4(const float* v0, const float* e1, const float* e2, const uint32_t* flags, uint32_t n,
                int threads, int max_leaf = 4, float traversal_cost = 1.0f);

} // namespace mrt
