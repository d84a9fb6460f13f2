// Binned-SAH BVH2 build, collapsed into the 128-byte BVH4 layout of bvh.h.
//
// The reference leaves BVH construction to the Vulkan driver (PreferFastTrace for static mesh
// groups, src/merian-shaders/scene/scene.cpp:1116-1123). This builder is the project's own:
// 32-bin SAH per axis over centroid bounds, leaves of at most max_leaf (<= MAX_LEAF) triangles, top levels
// built in parallel threads. Child boxes are rounded outward and padded so that the slab test
// (one FMA per plane) can never reject a box containing a hit the triangle test reports.
#include "bvh.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>

namespace mrt {
namespace {

struct AABB {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const AABB& b) {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void grow(const float* p) {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    float area() const {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0.f) return 0.f;
        return 2.f * (dx * dy + dy * dz + dz * dx);
    }
};

struct Prim {
    AABB box;
    float c[3];
};

struct Node2 {
    AABB box;
    int32_t left = -1, right = -1;
    uint32_t first = 0, count = 0;
};

constexpr int BINS = 32;

struct Builder {
    const std::vector<Prim>& prims;
    std::vector<uint32_t> idx;
    std::vector<Node2> nodes;
    std::atomic<uint32_t> next{1};
    int par_depth;
    uint32_t max_leaf = 4;
    float traversal_cost = 1.0f;

    Builder(const std::vector<Prim>& p, std::vector<uint32_t> order, int threads)
        : prims(p), idx(std::move(order)), nodes(std::max<size_t>(1, 2 * idx.size())) {
        par_depth = 0;
        while ((1 << par_depth) < threads && par_depth < 6) par_depth++;
    }

    void make_leaf(Node2& n, uint32_t b, uint32_t e) {
        n.left = n.right = -1;
        n.first = b;
        n.count = e - b;
    }

    void build(uint32_t ni, uint32_t b, uint32_t e, int depth) {
        Node2& node = nodes[ni];
        AABB box, cbox;
        for (uint32_t i = b; i < e; i++) {
            box.grow(prims[idx[i]].box);
            cbox.grow(prims[idx[i]].c);
        }
        node.box = box;
        const uint32_t n = e - b;
        if (n <= 1) {
            make_leaf(node, b, e);
            return;
        }
        // binned SAH over centroid bounds
        float best_cost = FLT_MAX;
        int best_axis = -1, best_split = 0;
        for (int ax = 0; ax < 3; ax++) {
            const float ext = cbox.hi[ax] - cbox.lo[ax];
            if (!(ext > 0.f)) continue;
            const float scale = float(BINS) / ext;
            AABB bb[BINS];
            uint32_t cnt[BINS] = {};
            for (uint32_t i = b; i < e; i++) {
                const Prim& p = prims[idx[i]];
                int k = int((p.c[ax] - cbox.lo[ax]) * scale);
                k = std::min(std::max(k, 0), BINS - 1);
                cnt[k]++;
                bb[k].grow(p.box);
            }
            float right_area[BINS];
            uint32_t right_cnt[BINS];
            AABB acc;
            uint32_t c = 0;
            for (int k = BINS - 1; k > 0; k--) {
                acc.grow(bb[k]);
                c += cnt[k];
                right_area[k] = acc.area();
                right_cnt[k] = c;
            }
            AABB lacc;
            uint32_t lc = 0;
            for (int k = 1; k < BINS; k++) {
                lacc.grow(bb[k - 1]);
                lc += cnt[k - 1];
                if (lc == 0 || right_cnt[k] == 0) continue;
                const float cost = lacc.area() * float(lc) + right_area[k] * float(right_cnt[k]);
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = ax;
                    best_split = k;
                }
            }
        }
        const float leaf_cost = box.area() * float(n);
        uint32_t mid;
        if (best_axis < 0) {
            if (n <= max_leaf) {
                make_leaf(node, b, e);
                return;
            }
            mid = b + n / 2; // coincident centroids: split by index
        } else {
            if (n <= max_leaf && best_cost + traversal_cost * box.area() >= leaf_cost) {
                make_leaf(node, b, e);
                return;
            }
            const float ext = cbox.hi[best_axis] - cbox.lo[best_axis];
            const float scale = float(BINS) / ext;
            const float lo = cbox.lo[best_axis];
            auto it = std::partition(idx.begin() + b, idx.begin() + e, [&](uint32_t pi) {
                int k = int((prims[pi].c[best_axis] - lo) * scale);
                k = std::min(std::max(k, 0), BINS - 1);
                return k < best_split;
            });
            mid = uint32_t(it - idx.begin());
            if (mid == b || mid == e) mid = b + n / 2;
        }
        const uint32_t l = next.fetch_add(2);
        const uint32_t r = l + 1;
        node.left = int32_t(l);
        node.right = int32_t(r);
        node.count = 0;
        if (depth < par_depth && n > 65536) {
            std::thread t([&, l, b, mid, depth] { build(l, b, mid, depth + 1); });
            build(r, mid, e, depth + 1);
            t.join();
        } else {
            build(l, b, mid, depth + 1);
            build(r, mid, e, depth + 1);
        }
    }
};

// pad a child box outward: rounding of the FMA slab test and of the triangle test
void padded(const AABB& b, float mag, float lo[3], float hi[3]) {
    float ext = 0.f;
    for (int k = 0; k < 3; k++) ext = std::max(ext, b.hi[k] - b.lo[k]);
    const float pad = ext * (1.0f / 1024.0f) + mag * (1.0f / 65536.0f) + 1e-30f;
    for (int k = 0; k < 3; k++) {
        lo[k] = b.lo[k] - pad;
        hi[k] = b.hi[k] + pad;
    }
}

struct Collapser {
    const std::vector<Node2>& n2;
    BVH4& out;
    float mag;

    uint32_t collapse(int32_t ni, uint32_t depth) {
        int32_t kids[4];
        int nk = 0;
        const Node2& root = n2[ni];
        kids[nk++] = root.left;
        kids[nk++] = root.right;
        while (nk < 4) {
            int best = -1;
            float best_area = -1.f;
            for (int i = 0; i < nk; i++) {
                const Node2& c = n2[kids[i]];
                if (c.left >= 0 && c.box.area() > best_area) {
                    best_area = c.box.area();
                    best = i;
                }
            }
            if (best < 0) break;
            const Node2& c = n2[kids[best]];
            kids[best] = c.left;
            kids[nk++] = c.right;
        }
        const uint32_t me = uint32_t(out.nodes.size());
        out.nodes.emplace_back();
        out.depth = std::max(out.depth, depth);
        BVH4Node node;
        std::memset(&node, 0, sizeof(node));
        for (int i = 0; i < 4; i++) {
            if (i >= nk) {
                node.lox[i] = node.loy[i] = node.loz[i] = 1e30f;
                node.hix[i] = node.hiy[i] = node.hiz[i] = -1e30f;
                node.entry[i] = EMPTY_ENTRY;
                continue;
            }
            const Node2& c = n2[kids[i]];
            float lo[3], hi[3];
            padded(c.box, mag, lo, hi);
            node.lox[i] = lo[0];
            node.loy[i] = lo[1];
            node.loz[i] = lo[2];
            node.hix[i] = hi[0];
            node.hiy[i] = hi[1];
            node.hiz[i] = hi[2];
            if (c.left < 0) {
                node.entry[i] = leaf_entry(c.first, c.count);
            } else {
                node.entry[i] = collapse(kids[i], depth + 1);
            }
        }
        out.nodes[me] = node;
        return me;
    }
};

float down(double v) {
    float f = float(v);
    if (double(f) > v) f = std::nextafter(f, -INFINITY);
    return f;
}
float up(double v) {
    float f = float(v);
    if (double(f) < v) f = std::nextafter(f, INFINITY);
    return f;
}

} // namespace

namespace {

// Primitive boxes (as the intersection test sees the vertices) and the binned-SAH BVH2.
struct Prepared {
    std::vector<Prim> prims;
    std::vector<uint32_t> order;
    float mag = 0.f;
};

Prepared prepare(const float* v0, const float* e1, const float* e2, uint32_t n) {
    Prepared pr;
    pr.prims.resize(n);
    pr.order.reserve(n);
    for (uint32_t i = 0; i < n; i++) {
        Prim& p = pr.prims[i];
        bool ok = true;
        for (int k = 0; k < 3; k++) {
            // vertices as the intersection test sees them: v0, v0 + e1, v0 + e2 (exact in double)
            const double a = v0[3 * i + k], b = a + double(e1[3 * i + k]), c = a + double(e2[3 * i + k]);
            if (!std::isfinite(a) || !std::isfinite(b) || !std::isfinite(c)) ok = false;
            p.box.lo[k] = down(std::min(a, std::min(b, c)));
            p.box.hi[k] = up(std::max(a, std::max(b, c)));
            p.c[k] = float((std::min(a, std::min(b, c)) + std::max(a, std::max(b, c))) * 0.5);
        }
        if (!ok) continue; // a non-finite triangle can never be hit
        for (int k = 0; k < 3; k++)
            pr.mag = std::max(pr.mag, std::max(std::fabs(p.box.lo[k]), std::fabs(p.box.hi[k])));
        pr.order.push_back(i);
    }
    return pr;
}

void set_record(TriRecord& t, uint32_t i, const float* v0, const float* e1, const float* e2, const uint32_t* flags) {
    for (int c = 0; c < 3; c++) {
        t.a[c] = v0[3 * i + c];
        t.b[c] = e1[3 * i + c];
        t.c[c] = e2[3 * i + c];
    }
    std::memcpy(&t.a[3], &i, 4);
    std::memcpy(&t.b[3], &flags[i], 4);
    t.c[3] = 0.f;
}

// Children of a BVH2 node for a k-wide node: repeatedly open the largest inner child.
int gather_children(const std::vector<Node2>& n2, int32_t ni, int width, int32_t* kids) {
    int nk = 0;
    kids[nk++] = n2[ni].left;
    kids[nk++] = n2[ni].right;
    while (nk < width) {
        int best = -1;
        float best_area = -1.f;
        for (int i = 0; i < nk; i++) {
            const Node2& c = n2[kids[i]];
            if (c.left >= 0 && c.box.area() > best_area) {
                best_area = c.box.area();
                best = i;
            }
        }
        if (best < 0) break;
        const Node2 c = n2[kids[best]];
        kids[best] = c.left;
        kids[nk++] = c.right;
    }
    return nk;
}

// 8-bit grid of one node: origin, exponent per axis, quantised child planes containing `lo`/`hi`.
struct Quant {
    float p[3];
    int e[3];
};

Quant node_grid(const float (*lo)[3], const float (*hi)[3], int nk) {
    Quant q;
    for (int a = 0; a < 3; a++) {
        float mn = FLT_MAX, mx = -FLT_MAX;
        for (int k = 0; k < nk; k++) {
            mn = std::min(mn, lo[k][a]);
            mx = std::max(mx, hi[k][a]);
        }
        q.p[a] = mn;
        const double ext = double(mx) - double(mn);
        int e = -126;
        if (ext > 0) {
            e = int(std::ceil(std::log2(ext / 250.0)));
            while (std::ldexp(250.0, e) < ext) e++;
        }
        q.e[a] = std::max(-126, std::min(127, e));
    }
    return q;
}

uint8_t quant_lo(float v, float p, float s) {
    double q = std::floor((double(v) - double(p)) / double(s));
    int qi = int(std::max(0.0, std::min(255.0, q)));
    while (qi > 0 && std::fmaf(float(qi), s, p) > v) qi--;
    return uint8_t(qi);
}

uint8_t quant_hi(float v, float p, float s) {
    double q = std::ceil((double(v) - double(p)) / double(s));
    int qi = int(std::max(0.0, std::min(255.0, q)));
    while (qi < 255 && std::fmaf(float(qi), s, p) < v) qi++;
    return uint8_t(qi);
}

struct Collapser8 {
    const std::vector<Node2>& n2;
    const std::vector<uint32_t>& idx;
    const float *v0, *e1, *e2;
    const uint32_t* flags;
    BVH8& out;
    float mag;

    // Places the node's children in slots by corner: slot bit a set = the child lies toward the
    // high end of axis a (children scored by their centre's offset from the node centre,
    // relative to the node's extent, greedily best pair first). A ray whose direction signs are
    // the bits o (bit a: negative along a) then meets the children roughly in slot order i ^ o.
    void assign_slots(const int32_t* kids, int nk, int32_t* slot_kid) const {
        for (int s = 0; s < 8; s++) slot_kid[s] = -1;
        float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        float ctr[8][3];
        for (int k = 0; k < nk; k++) {
            const AABB& b = n2[kids[k]].box;
            for (int a = 0; a < 3; a++) {
                ctr[k][a] = 0.5f * (b.lo[a] + b.hi[a]);
                mn[a] = std::min(mn[a], b.lo[a]);
                mx[a] = std::max(mx[a], b.hi[a]);
            }
        }
        float score[8][8];
        for (int k = 0; k < nk; k++)
            for (int s = 0; s < 8; s++) {
                float sc = 0.f;
                for (int a = 0; a < 3; a++) {
                    const float ext = mx[a] - mn[a];
                    const float off = ext > 0.f ? (ctr[k][a] - 0.5f * (mn[a] + mx[a])) / ext : 0.f;
                    sc += ((s >> a) & 1) ? off : -off;
                }
                score[k][s] = sc;
            }
        bool kid_done[8] = {}, slot_done[8] = {};
        for (int n = 0; n < nk; n++) {
            int bk = -1, bs = -1;
            float best = -FLT_MAX;
            for (int k = 0; k < nk; k++)
                for (int s = 0; s < 8; s++)
                    if (!kid_done[k] && !slot_done[s] && score[k][s] > best) {
                        best = score[k][s];
                        bk = k;
                        bs = s;
                    }
            kid_done[bk] = slot_done[bs] = true;
            slot_kid[bs] = kids[bk];
        }
    }

    // fills out.nodes[me] from BVH2 node ni (or a lone leaf when ni is a leaf)
    void emit(int32_t ni, uint32_t me, uint32_t depth) {
        out.depth = std::max(out.depth, depth);
        int32_t kids[8];
        int nk;
        if (n2[ni].left < 0) {
            kids[0] = ni;
            nk = 1;
        } else {
            nk = gather_children(n2, ni, 8, kids);
        }
        int32_t sk[8];
        assign_slots(kids, nk, sk);
        float lo[8][3], hi[8][3];
        for (int k = 0; k < nk; k++) padded(n2[kids[k]].box, mag, lo[k], hi[k]);
        const Quant g = node_grid(lo, hi, nk);
        BVH8Node node;
        std::memset(&node, 0, sizeof(node));
        node.px = g.p[0];
        node.py = g.p[1];
        node.pz = g.p[2];
        node.ex = int8_t(g.e[0]);
        node.ey = int8_t(g.e[1]);
        node.ez = int8_t(g.e[2]);
        uint8_t* qlo[3] = {node.qlox, node.qloy, node.qloz};
        uint8_t* qhi[3] = {node.qhix, node.qhiy, node.qhiz};
        for (int k = 0; k < 8; k++)
            for (int a = 0; a < 3; a++) {
                qlo[a][k] = 255;
                qhi[a][k] = 0;
            }
        int n_inner = 0;
        for (int k = 0; k < nk; k++)
            if (n2[kids[k]].left >= 0) n_inner++;
        node.child_base = uint32_t(out.nodes.size());
        out.nodes.resize(out.nodes.size() + size_t(n_inner));
        node.tri_base = uint32_t(out.tris.size());
        std::memset(node.code, BVH8_EMPTY, 8);
        uint32_t rank = 0;
        for (int s = 0; s < 8; s++) {
            if (sk[s] < 0) continue;
            const Node2& c = n2[sk[s]];
            float clo[3], chi[3];
            padded(c.box, mag, clo, chi);
            for (int a = 0; a < 3; a++) {
                const float sc = std::ldexp(1.0f, g.e[a]);
                qlo[a][s] = quant_lo(clo[a], g.p[a], sc);
                qhi[a][s] = quant_hi(chi[a], g.p[a], sc);
            }
            if (c.left >= 0) {
                node.inner_mask |= uint8_t(1u << s);
                node.code[s] = uint8_t(BVH8_INNER + rank++);
            } else {
                const uint32_t off = uint32_t(out.tris.size()) - node.tri_base;
                node.code[s] = uint8_t(off << 3 | (c.count - 1u)); // count in [1, BVH8_MAX_LEAF]
                for (uint32_t t = 0; t < c.count; t++) {
                    out.tris.emplace_back();
                    set_record(out.tris.back(), idx[c.first + t], v0, e1, e2, flags);
                }
            }
        }
        out.nodes[me] = node;
        uint32_t slot = node.child_base;
        for (int s = 0; s < 8; s++)
            if (sk[s] >= 0 && n2[sk[s]].left >= 0) emit(sk[s], slot++, depth + 1);
    }
};

} // namespace

BVH4 build_bvh4(const float* v0, const float* e1, const float* e2, const uint32_t* flags, uint32_t n,
                int threads, int max_leaf, float traversal_cost) {
    if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
    Prepared pr = prepare(v0, e1, e2, n);
    const std::vector<Prim>& prims = pr.prims;
    std::vector<uint32_t>& order = pr.order;
    const float mag = pr.mag;
    BVH4 out;
    const uint32_t m = uint32_t(order.size());
    Builder bld(prims, std::move(order), threads);
    bld.max_leaf = uint32_t(std::min(std::max(max_leaf, 1), MAX_LEAF));
    bld.traversal_cost = traversal_cost > 0.f ? traversal_cost : 1.0f;
    if (m > 0) bld.build(0, 0, m, 0);
    // triangle records in leaf order
    out.tris.resize(m);
    for (uint32_t k = 0; k < m; k++) {
        const uint32_t i = bld.idx[k];
        TriRecord& t = out.tris[k];
        for (int c = 0; c < 3; c++) {
            t.a[c] = v0[3 * i + c];
            t.b[c] = e1[3 * i + c];
            t.c[c] = e2[3 * i + c];
        }
        std::memcpy(&t.a[3], &i, 4);
        std::memcpy(&t.b[3], &flags[i], 4);
        t.c[3] = 0.f;
    }
    if (m == 0) {
        // empty scene: one node with four empty children
        BVH4Node node;
        std::memset(&node, 0, sizeof(node));
        for (int i = 0; i < 4; i++) {
            node.lox[i] = node.loy[i] = node.loz[i] = 1e30f;
            node.hix[i] = node.hiy[i] = node.hiz[i] = -1e30f;
            node.entry[i] = EMPTY_ENTRY;
        }
        out.nodes.push_back(node);
        out.depth = 1;
        return out;
    }
    const Node2& root = bld.nodes[0];
    if (root.left < 0) {
        // single leaf: wrap it in a root node with one leaf child
        BVH4Node node;
        std::memset(&node, 0, sizeof(node));
        for (int i = 0; i < 4; i++) {
            node.lox[i] = node.loy[i] = node.loz[i] = 1e30f;
            node.hix[i] = node.hiy[i] = node.hiz[i] = -1e30f;
            node.entry[i] = EMPTY_ENTRY;
        }
        float lo[3], hi[3];
        padded(root.box, mag, lo, hi);
        node.lox[0] = lo[0];
        node.loy[0] = lo[1];
        node.loz[0] = lo[2];
        node.hix[0] = hi[0];
        node.hiy[0] = hi[1];
        node.hiz[0] = hi[2];
        node.entry[0] = leaf_entry(root.first, root.count);
        out.nodes.push_back(node);
        out.depth = 1;
        return out;
    }
    out.nodes.reserve(bld.next.load() / 2 + 2);
    Collapser col{bld.nodes, out, mag};
    col.collapse(0, 1);
    return out;
}

BVH8 build_bvh8(const float* v0, const float* e1, const float* e2, const uint32_t* flags, uint32_t n,
                int threads, int max_leaf, float traversal_cost) {
    if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
    Prepared pr = prepare(v0, e1, e2, n);
    const float mag = pr.mag;
    const uint32_t m = uint32_t(pr.order.size());
    Builder bld(pr.prims, std::move(pr.order), threads);
    bld.max_leaf = uint32_t(std::min(std::max(max_leaf, 1), std::min(MAX_LEAF, BVH8_MAX_LEAF)));
    bld.traversal_cost = traversal_cost > 0.f ? traversal_cost : 1.0f;
    BVH8 out;
    if (m == 0) {
        BVH8Node node;
        std::memset(&node, 0, sizeof(node));
        std::memset(node.code, BVH8_EMPTY, 8);
        std::memset(node.qlox, 255, 8);
        std::memset(node.qloy, 255, 8);
        std::memset(node.qloz, 255, 8);
        node.ex = node.ey = node.ez = 0;
        out.nodes.push_back(node);
        out.depth = 1;
        return out;
    }
    bld.build(0, 0, m, 0);
    out.nodes.reserve(bld.next.load() / 4 + 2);
    out.tris.reserve(m);
    out.nodes.emplace_back();
    Collapser8 col{bld.nodes, bld.idx, v0, e1, e2, flags, out, mag};
    col.emit(0, 0, 1);
    return out;
}

} // namespace mrt
