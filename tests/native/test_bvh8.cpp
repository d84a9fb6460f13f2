// CPU check of the 8-wide quantised BVH (merian_amd/csrc/bvh.cpp): every decoded child box (the
// kernel's fma(q, 2^e, origin)) contains every triangle below it, every triangle appears exactly
// once, inner children are consecutive and leaf triangles follow tri_base in slot order, and the
// slot code bytes give exactly those entries (inner rank, leaf offset and count, empty).
#include "../../merian_amd/csrc/bvh.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

using namespace mrt;

static int failures = 0;
#define CHECK(c)                                                       \
    do {                                                               \
        if (!(c)) {                                                    \
            if (failures < 10) std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            failures++;                                                \
        }                                                              \
    } while (0)

struct Soup {
    std::vector<float> v0, e1, e2;
    std::vector<uint32_t> flags;
};

static float decode(uint8_t q, int e, float p) { return std::fmaf(float(q), std::ldexp(1.0f, e), p); }

// Walks the tree as the traversal kernel indexes it; every node and triangle index it forms
// must lie inside the arrays, and no path may be deeper than b.depth (the stack bound).
static void walk(const BVH8& b, const Soup& s, uint32_t ni, std::vector<int>& seen, uint32_t depth) {
    CHECK(ni < b.nodes.size());
    CHECK(depth <= b.depth);
    if (ni >= b.nodes.size()) return;
    const BVH8Node& n = b.nodes[ni];
    const uint8_t* qlo[3] = {n.qlox, n.qloy, n.qloz};
    const uint8_t* qhi[3] = {n.qhix, n.qhiy, n.qhiz};
    const int e[3] = {n.ex, n.ey, n.ez};
    const float p[3] = {n.px, n.py, n.pz};
    uint32_t inner = 0, tri = n.tri_base;
    for (int k = 0; k < 8; k++) {
        float lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = decode(qlo[a][k], e[a], p[a]);
            hi[a] = decode(qhi[a][k], e[a], p[a]);
        }
        const uint32_t code = n.code[k];
        const bool is_inner = code >= BVH8_INNER, is_empty = code == BVH8_EMPTY;
        CHECK(is_inner == bool(n.inner_mask & (1u << k)));
        CHECK(is_inner || is_empty || code <= BVH8_LEAF_CODE_MAX);
        // the kernel's entry: (inner ? child_base - BVH8_INNER : LEAF | tri_base << 3) + code
        const uint32_t entry = (is_inner ? n.child_base - uint32_t(BVH8_INNER) : (0x80000000u | n.tri_base << 3)) + code;
        const uint32_t cnt = is_inner || is_empty ? 0u : (code & 7u) + 1u;
        if (is_inner) {
            CHECK(entry == n.child_base + inner);
            walk(b, s, n.child_base + inner++, seen, depth + 1);
        } else if (!is_empty) {
            CHECK(cnt <= uint32_t(BVH8_MAX_LEAF));
            CHECK(entry == (0x80000000u | tri << 3 | (cnt - 1u)));
            CHECK(size_t(tri) + cnt <= b.tris.size());
            if (size_t(tri) + cnt > b.tris.size()) return;
            for (uint32_t t = 0; t < cnt; t++, tri++) {
                const TriRecord& r = b.tris[tri];
                uint32_t id;
                std::memcpy(&id, &r.a[3], 4);
                CHECK(id < seen.size());
                if (id < seen.size()) seen[id]++;
                for (int a = 0; a < 3; a++) {
                    const float x0 = r.a[a], x1 = r.a[a] + r.b[a], x2 = r.a[a] + r.c[a];
                    CHECK(lo[a] <= std::fmin(x0, std::fmin(x1, x2)) && hi[a] >= std::fmax(x0, std::fmax(x1, x2)));
                }
            }
        } else {
            for (int a = 0; a < 3; a++) CHECK(lo[a] > hi[a]); // empty slot never hits
        }
    }
}

int main() {
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> U(0.f, 1.f);
    for (int trial = 0; trial < 4; trial++) {
        Soup s;
        const uint32_t n = trial == 0 ? 1u : (trial == 1 ? 7u : (trial == 2 ? 1000u : 60000u));
        const float scale = trial == 3 ? 1000.f : 1.f;
        for (uint32_t i = 0; i < n; i++) {
            const float c[3] = {U(rng) * scale, U(rng) * scale, U(rng) * scale};
            const float l = 0.02f * scale * U(rng);
            for (int k = 0; k < 3; k++) {
                s.v0.push_back(c[k]);
                s.e1.push_back(l * (U(rng) - 0.5f));
                s.e2.push_back(l * (U(rng) - 0.5f));
            }
            s.flags.push_back(0xFF00u);
        }
        for (int leaf : {1, 2, 4, 8}) {
            const BVH8 b = build_bvh8(s.v0.data(), s.e1.data(), s.e2.data(), s.flags.data(), n, 2, leaf, 1.0f);
            std::vector<int> seen(n, 0);
            walk(b, s, 0, seen, 1);
            for (uint32_t i = 0; i < n; i++) CHECK(seen[i] == 1);
            CHECK(b.tris.size() == n);
        }
    }
    {
        // empty scene: one node whose eight slots are empty boxes
        const BVH8 b = build_bvh8(nullptr, nullptr, nullptr, nullptr, 0, 1, 4, 1.0f);
        CHECK(b.nodes.size() == 1 && b.tris.empty() && b.depth == 1);
        CHECK(b.nodes[0].inner_mask == 0);
        for (int k = 0; k < 8; k++) CHECK(b.nodes[0].code[k] == BVH8_EMPTY);
        std::vector<int> seen;
        Soup s;
        walk(b, s, 0, seen, 1);
    }
    if (failures) {
        std::printf("%d failures\n", failures);
        return 1;
    }
    std::printf("bvh8: all checks passed\n");
    return 0;
}
