// Closest-hit traversal of the BVH4 (bvh.h) on gfx950: one ray per lane, the traversal stack
// in LDS (one column per lane, so the 64 lanes of a wave hit 64 distinct banks), child order
// near-to-far by a 4-key sorting network, octant-selected slab planes (one FMA per plane).
//
// Result = Scene.trace_ray (include/merian-shaders/scene/scene.slang:62-76) with the ray query
// of acceleration-structure.slang:59-72: back-face culling per triangle, 8-bit instance mask,
// t in [tmin, tmax], ties at equal t to the lowest triangle id.
#pragma once

#include "bvh.h"
#include "mrt_device.h"
#include "mrt_kernels.h"

namespace mrt {

struct HitRec {
    uint32_t id; // global triangle id, 0xFFFFFFFF = miss
    float t, u, v;
};

struct DeviceBVH {
    const BVH4Node* nodes;
    const float4* tris; // TriRecord as 3 x float4
};

// Textbook Moller-Trumbore (1997) in a fixed operation order; no contraction (the file is
// compiled with -ffp-contract=off), so CPU and GPU round identically.
__device__ __forceinline__ void test_triangle(const float4 A, const float4 B, const float4 Cc, f3 o, f3 d, float tmin,
                                              float tmax, uint32_t ray_mask, HitRec& best) {
    const uint32_t flags = __float_as_uint(B.w);
    if (((flags >> 8) & ray_mask) == 0u) return;
    const f3 e1{B.x, B.y, B.z}, e2{Cc.x, Cc.y, Cc.z};
    const f3 p = cross(d, e2);
    const float det = dot(e1, p);
    if (flags & 1u) {
        if (!(det > 0.0f)) return;
    } else {
        if (!(det != 0.0f)) return;
    }
    const float inv = 1.0f / det;
    const f3 s = o - f3{A.x, A.y, A.z};
    const float u = dot(s, p) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) return;
    const f3 q = cross(s, e1);
    const float v = dot(d, q) * inv;
    if (!(v >= 0.0f && u + v <= 1.0f)) return;
    const float t = dot(e2, q) * inv;
    if (!(t >= tmin && t <= tmax)) return;
    const uint32_t id = __float_as_uint(A.w);
    if (t < best.t || (t == best.t && id < best.id)) best = HitRec{id, t, u, v};
}

__device__ __forceinline__ void cswap(float& ka, uint32_t& ea, float& kb, uint32_t& eb) {
    const bool s = kb < ka;
    const float tk = s ? kb : ka;
    const uint32_t te = s ? eb : ea;
    kb = s ? ka : kb;
    eb = s ? ea : eb;
    ka = tk;
    ea = te;
}

constexpr uint32_t TRAVERSAL_DONE = 0xFFFFFFFFu;

// Work counters of one traversal (COUNT variant only): node visits and triangle tests.
struct TraversalCount {
    uint32_t nodes = 0, tris = 0;
};

// Traversal stack of one lane: the first SL entries in an LDS column (stride = threads per
// block, so a wave's 64 lanes touch 64 distinct banks), deeper entries in a per-lane column of a
// global overflow area (stride = threads of the grid). Near-first traversal rarely goes past SL.
// The LDS column stride is the block size, a compile-time constant (a shift, not a multiply).
struct LaneStack {
    uint32_t* lds; // this lane's column
    uint32_t* ovf;
    uint32_t ovf_stride;
};
template <int SL> __device__ __forceinline__ void stack_push(const LaneStack& s, int& sp, uint32_t e) {
    if (sp < SL) {
        s.lds[sp * TRACE_BLOCK] = e;
    } else {
        s.ovf[size_t(sp - SL) * s.ovf_stride] = e;
    }
    sp++;
}
template <int SL> __device__ __forceinline__ uint32_t stack_pop(const LaneStack& s, int& sp) {
    --sp;
    // explicit address spaces keep the two loads apart (a ds_read and a global load rather
    // than one flat load through a selected pointer)
    typedef const __attribute__((address_space(3))) uint32_t* lds_ptr;
    typedef const __attribute__((address_space(1))) uint32_t* global_ptr;
    if (sp < SL) return ((lds_ptr)s.lds)[sp * TRACE_BLOCK];
    return ((global_ptr)s.ovf)[size_t(sp - SL) * s.ovf_stride];
}
template <int SL> __device__ __forceinline__ uint32_t stack_peek(const LaneStack& s, int i) {
    typedef const __attribute__((address_space(3))) uint32_t* lds_ptr;
    typedef const __attribute__((address_space(1))) uint32_t* global_ptr;
    if (i < SL) return ((lds_ptr)s.lds)[i * TRACE_BLOCK];
    return ((global_ptr)s.ovf)[size_t(i - SL) * s.ovf_stride];
}
template <int SL> __device__ __forceinline__ void stack_poke(const LaneStack& s, int i, uint32_t e) {
    if (i < SL) {
        s.lds[i * TRACE_BLOCK] = e;
    } else {
        s.ovf[size_t(i - SL) * s.ovf_stride] = e;
    }
}

// Per-lane traversal state between steps.
struct RayState {
    f3 o, d;
    float ix, iy, iz, ox, oy, oz;
    float t0, tmin, tmax;
    float tlim;   // far limit of the box tests: min(tmax, best.t) widened by 2 ulp
    uint32_t neg; // bit k: direction component k is negative
    uint32_t pn, pf; // BVH4: byte offsets of the near / far plane row of axis k in byte k
    uint32_t psel[3]; // BVH8: v_perm selector per axis picking the near plane word of a node
                      // (lo for a positive direction, hi for a negative one); swapped operands
                      // pick the far word
    HitRec best;
    int sp;
    uint32_t cur;  // next entry to process; TRAVERSAL_DONE when the stack is exhausted
    uint32_t leaf; // postponed leaf entry awaiting its triangle tests; TRAVERSAL_DONE when none
};

__device__ __forceinline__ bool is_inner(uint32_t e) { return !(e & LEAF_FLAG); }
__device__ __forceinline__ bool is_leaf(uint32_t e) { return e != TRAVERSAL_DONE && (e & LEAF_FLAG); }

__device__ __forceinline__ void ray_begin(RayState& r, f3 o, f3 d, float tmin, float tmax) {
    r.o = o;
    r.d = d;
    const float dx = fabsf(d.x) < 1e-20f ? copysignf(1e-20f, d.x) : d.x;
    const float dy = fabsf(d.y) < 1e-20f ? copysignf(1e-20f, d.y) : d.y;
    const float dz = fabsf(d.z) < 1e-20f ? copysignf(1e-20f, d.z) : d.z;
    // the hardware reciprocal (1 ulp) instead of a correctly rounded division: 1/d only feeds
    // the box tests, whose boxes the builder pads by extent/1024 + magnitude/65536, far above
    // an ulp of the slab distances; the triangle test uses d itself
    r.ix = __builtin_amdgcn_rcpf(dx);
    r.iy = __builtin_amdgcn_rcpf(dy);
    r.iz = __builtin_amdgcn_rcpf(dz);
    r.ox = o.x * r.ix;
    r.oy = o.y * r.iy;
    r.oz = o.z * r.iz;
    r.neg = (r.ix < 0.f ? 1u : 0u) | (r.iy < 0.f ? 2u : 0u) | (r.iz < 0.f ? 4u : 0u);
    r.tmin = tmin;
    r.tmax = tmax;
    r.t0 = tmin - fabsf(tmin) * 2.4e-7f;
    r.best = HitRec{TRAVERSAL_DONE, __builtin_inff(), 0.f, 0.f};
    r.tlim = fminf(tmax, r.best.t) * 1.00000024f;
    // BVH4Node rows: lox hix loy hiy loz hiz, 16 bytes each
    const uint32_t nx = r.neg & 1u, ny = (r.neg >> 1) & 1u, nz = (r.neg >> 2) & 1u;
    r.pn = (nx << 4) | ((32u + (ny << 4)) << 8) | ((64u + (nz << 4)) << 16);
    r.pf = ((16u - (nx << 4))) | ((48u - (ny << 4)) << 8) | ((80u - (nz << 4)) << 16);
    r.psel[0] = nx ? 0x07060504u : 0x03020100u;
    r.psel[1] = ny ? 0x07060504u : 0x03020100u;
    r.psel[2] = nz ? 0x07060504u : 0x03020100u;
    r.sp = 0;
    r.cur = 0u; // root node (always an inner node)
    r.leaf = TRAVERSAL_DONE;
}

// Tests the triangles of the postponed leaf (ANY: the first hit ends the traversal).
template <int SL, bool COUNT, bool ANY>
__device__ __forceinline__ void leaf_test(const float4* __restrict__ tris, RayState& r, uint32_t ray_mask,
                                          TraversalCount& count) {
    const uint32_t first = (r.leaf & 0x7FFFFFFFu) >> 3;
    const uint32_t cnt = (r.leaf & 7u) + 1u;
    // two records per round, both fetched before either is tested (one memory round trip)
    for (uint32_t k = 0; k < cnt; k += 2) {
        const uint32_t a = first + k, b = first + min(k + 1u, cnt - 1u);
        const float4 A0 = tris[3 * a], B0 = tris[3 * a + 1], C0 = tris[3 * a + 2];
        const float4 A1 = tris[3 * b], B1 = tris[3 * b + 1], C1 = tris[3 * b + 2];
        test_triangle(A0, B0, C0, r.o, r.d, r.tmin, r.tmax, ray_mask, r.best);
        if (k + 1u < cnt) test_triangle(A1, B1, C1, r.o, r.d, r.tmin, r.tmax, ray_mask, r.best);
    }
    if constexpr (COUNT) count.tris += cnt;
    r.leaf = TRAVERSAL_DONE;
    r.tlim = fminf(r.tmax, r.best.t) * 1.00000024f;
    if constexpr (ANY) {
        if (r.best.id != TRAVERSAL_DONE) {
            r.cur = TRAVERSAL_DONE;
            r.sp = 0;
        }
    }
}

// One inner-node step (r.cur is an inner node): test its four children, push the hit ones far
// to near, continue with the nearest.
template <int SL, bool COUNT>
__device__ __forceinline__ void node_step(const DeviceBVH bvh, RayState& r, const LaneStack& stk,
                                          TraversalCount& count) {
    if constexpr (COUNT) count.nodes++;
    // near and far plane rows of each axis chosen by address: one FMA per plane, no selects
    const char* nb = reinterpret_cast<const char*>(bvh.nodes);
    const uint32_t off = r.cur << 7;
    auto row = [&](uint32_t o) { return *reinterpret_cast<const float4*>(nb + (off + o)); };
    const float4 nxr = row(r.pn & 0xFFu), fxr = row(r.pf & 0xFFu);
    const float4 nyr = row((r.pn >> 8) & 0xFFu), fyr = row((r.pf >> 8) & 0xFFu);
    const float4 nzr = row(r.pn >> 16), fzr = row(r.pf >> 16);
    const uint4 ent4 = *reinterpret_cast<const uint4*>(nb + (off + 96u));
    const float tlim = r.tlim;
    float key[4];
    uint32_t ent[4];
#define MRT_CHILD(c, NX, FX, NY, FY, NZ, FZ, E)                                                              \
    {                                                                                                        \
        const float tn = fmaxf(fmaxf(__builtin_fmaf(NX, r.ix, -r.ox), __builtin_fmaf(NY, r.iy, -r.oy)),      \
                               fmaxf(__builtin_fmaf(NZ, r.iz, -r.oz), r.t0));                               \
        const float tf = fminf(fminf(__builtin_fmaf(FX, r.ix, -r.ox), __builtin_fmaf(FY, r.iy, -r.oy)),      \
                               fminf(__builtin_fmaf(FZ, r.iz, -r.oz), tlim));                               \
        key[c] = (tn <= tf) ? tn : __builtin_inff();                                                         \
        ent[c] = E;                                                                                          \
    }
    MRT_CHILD(0, nxr.x, fxr.x, nyr.x, fyr.x, nzr.x, fzr.x, ent4.x)
    MRT_CHILD(1, nxr.y, fxr.y, nyr.y, fyr.y, nzr.y, fzr.y, ent4.y)
    MRT_CHILD(2, nxr.z, fxr.z, nyr.z, fyr.z, nzr.z, fzr.z, ent4.z)
    MRT_CHILD(3, nxr.w, fxr.w, nyr.w, fyr.w, nzr.w, fzr.w, ent4.w)
#undef MRT_CHILD
    cswap(key[0], ent[0], key[1], ent[1]);
    cswap(key[2], ent[2], key[3], ent[3]);
    cswap(key[0], ent[0], key[2], ent[2]);
    cswap(key[1], ent[1], key[3], ent[3]);
    cswap(key[1], ent[1], key[2], ent[2]);
    const float inf = __builtin_inff();
    if (key[3] != inf) stack_push<SL>(stk, r.sp, ent[3]);
    if (key[2] != inf) stack_push<SL>(stk, r.sp, ent[2]);
    if (key[1] != inf) stack_push<SL>(stk, r.sp, ent[1]);
    if (key[0] != inf) {
        r.cur = ent[0];
    } else {
        r.cur = r.sp > 0 ? stack_pop<SL>(stk, r.sp) : TRAVERSAL_DONE;
    }
}

// Entry of BVH8 slot c from the node's second 16 bytes (child base, triangle base, code words):
// the slot's code byte picked by a byte permute (selector bytes 0x0C read as zero), then
// base + code (bvh.h).
__device__ __forceinline__ uint32_t bvh8_entry(const uint4 h1, uint32_t c) {
    const uint32_t code = __builtin_amdgcn_perm(h1.w, h1.z, c | 0x0C0C0C00u);
    return (code >= BVH8_INNER ? h1.x - uint32_t(BVH8_INNER) : LEAF_FLAG | (h1.y << 3)) + code;
}

// Grouped BVH8 stack entries: a node index (< 2^24) and, in the top byte, the node's hit
// children not yet visited as a mask over visiting positions (position i = slot i ^ r.neg).
constexpr uint32_t BVH8_GROUP_NODE = 0xFFFFFFu;

struct DeviceBVH8 { // near-first order by sorted distances, one stack entry per pushed child
    const BVH8Node* nodes;
    const float4* tris; // TriRecord as 3 x float4
};
struct DeviceBVH8G { // octant visiting order, one stack group per node
    const BVH8Node* nodes;
    const float4* tris;
};

// Next entry to visit: BVH4 and sorted BVH8 pop a ready entry; grouped BVH8 takes the first
// remaining child of the top group (re-reading its node's 16-byte header) and drops the group
// once it is empty.
template <int SL>
__device__ __forceinline__ uint32_t pop_next(const DeviceBVH&, RayState& r, const LaneStack& stk) {
    return r.sp > 0 ? stack_pop<SL>(stk, r.sp) : TRAVERSAL_DONE;
}
template <int SL>
__device__ __forceinline__ uint32_t pop_next(const DeviceBVH8&, RayState& r, const LaneStack& stk) {
    return r.sp > 0 ? stack_pop<SL>(stk, r.sp) : TRAVERSAL_DONE;
}
template <int SL>
__device__ __forceinline__ uint32_t pop_next(const DeviceBVH8G bvh, RayState& r, const LaneStack& stk) {
    if (r.sp == 0) return TRAVERSAL_DONE;
    const uint32_t g = stack_peek<SL>(stk, r.sp - 1);
    const uint32_t node = g & BVH8_GROUP_NODE;
    const uint32_t m = g >> 24;
    const uint4 h1 = reinterpret_cast<const uint4*>(bvh.nodes + node)[1];
    const uint32_t rest = m & (m - 1u);
    if (rest == 0u) {
        r.sp--;
    } else {
        stack_poke<SL>(stk, r.sp - 1, node | (rest << 24));
    }
    return bvh8_entry(h1, uint32_t(__builtin_ctz(m)) ^ r.neg);
}

// One inner-node step over the 8-wide quantised BVH, near-first by sorted entry distance: slab-test the eight child boxes, sort the
// hit children near to far, push all but the nearest, continue with it.
//
// Slab distances are taken in the node's grid: a plane at grid coordinate q lies at q 2^e + p,
// so its distance is q (2^e / d) + (p / d - o / d) = fma(q, a, b) with a = 2^e ix (exact: a
// power-of-two scale) and b = fma(p, ix, -ox) once per node and axis, one FMA per plane after
// the byte conversion. The sort key of a hit child is its entry distance's bit pattern with the
// child slot in the low three bits (misses: all ones), so a 19-comparator network of integer
// min/max sorts keys alone; a pushed child's entry is rebuilt from its slot's code byte.
// Only the visiting order depends on the keys; which children are visited does not.
template <int SL, bool COUNT>
__device__ __forceinline__ void node_step(const DeviceBVH8 bvh, RayState& r, const LaneStack& stk,
                                          TraversalCount& count) {
    if constexpr (COUNT) count.nodes++;
    const uint4* np = reinterpret_cast<const uint4*>(bvh.nodes + r.cur);
    const uint4 h0 = np[0], h1 = np[1], qx = np[2], qy = np[3], qz = np[4];
    const int ex = int(int8_t(h0.w & 0xFFu)), ey = int(int8_t((h0.w >> 8) & 0xFFu)), ez = int(int8_t((h0.w >> 16) & 0xFFu));
    const float ax = ldexpf(r.ix, ex), ay = ldexpf(r.iy, ey), az = ldexpf(r.iz, ez); // 2^e ix, exact
    const float bx = __builtin_fmaf(__uint_as_float(h0.x), r.ix, -r.ox);
    const float by = __builtin_fmaf(__uint_as_float(h0.y), r.iy, -r.oy);
    const float bz = __builtin_fmaf(__uint_as_float(h0.z), r.iz, -r.oz);
    const float tlim = r.tlim;
    // near / far plane bytes per axis, picked by the ray's octant with one byte permute per word
    // (children 0-3 in the low word, 4-7 in the high word)
    const uint32_t sx = r.psel[0], sy = r.psel[1], sz = r.psel[2];
    const uint32_t nx0 = __builtin_amdgcn_perm(qx.z, qx.x, sx), nx1 = __builtin_amdgcn_perm(qx.w, qx.y, sx);
    const uint32_t fx0 = __builtin_amdgcn_perm(qx.x, qx.z, sx), fx1 = __builtin_amdgcn_perm(qx.y, qx.w, sx);
    const uint32_t ny0 = __builtin_amdgcn_perm(qy.z, qy.x, sy), ny1 = __builtin_amdgcn_perm(qy.w, qy.y, sy);
    const uint32_t fy0 = __builtin_amdgcn_perm(qy.x, qy.z, sy), fy1 = __builtin_amdgcn_perm(qy.y, qy.w, sy);
    const uint32_t nz0 = __builtin_amdgcn_perm(qz.z, qz.x, sz), nz1 = __builtin_amdgcn_perm(qz.w, qz.y, sz);
    const uint32_t fz0 = __builtin_amdgcn_perm(qz.x, qz.z, sz), fz1 = __builtin_amdgcn_perm(qz.y, qz.w, sz);
    constexpr uint32_t MISS = 0xFFFFFFFFu;
    uint32_t key[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const uint32_t sh = 8u * uint32_t(c & 3);
        const float NX = __builtin_fmaf(float(((c < 4 ? nx0 : nx1) >> sh) & 0xFFu), ax, bx);
        const float FX = __builtin_fmaf(float(((c < 4 ? fx0 : fx1) >> sh) & 0xFFu), ax, bx);
        const float NY = __builtin_fmaf(float(((c < 4 ? ny0 : ny1) >> sh) & 0xFFu), ay, by);
        const float FY = __builtin_fmaf(float(((c < 4 ? fy0 : fy1) >> sh) & 0xFFu), ay, by);
        const float NZ = __builtin_fmaf(float(((c < 4 ? nz0 : nz1) >> sh) & 0xFFu), az, bz);
        const float FZ = __builtin_fmaf(float(((c < 4 ? fz0 : fz1) >> sh) & 0xFFu), az, bz);
        const float tn = fmaxf(fmaxf(NX, NY), fmaxf(NZ, r.t0));
        const float tf = fminf(fminf(FX, FY), fminf(FZ, tlim));
        const uint32_t code = ((c < 4 ? h1.z : h1.w) >> sh) & 0xFFu;
        key[c] = (tn <= tf && code != BVH8_EMPTY) ? ((__float_as_uint(tn) & ~7u) | uint32_t(c)) : MISS;
    }
    // entry of the slot in a sorted key's low bits
    auto entry_of = [&](uint32_t k) { return bvh8_entry(h1, k & 7u); };
#define MRT_CS(a, b)                                                                                           \
    {                                                                                                          \
        const uint32_t lo_ = min(key[a], key[b]);                                                              \
        key[b] = max(key[a], key[b]);                                                                          \
        key[a] = lo_;                                                                                          \
    }
    MRT_CS(0, 1) MRT_CS(2, 3) MRT_CS(4, 5) MRT_CS(6, 7)
    MRT_CS(0, 2) MRT_CS(1, 3) MRT_CS(4, 6) MRT_CS(5, 7)
    MRT_CS(1, 2) MRT_CS(5, 6)
    MRT_CS(0, 4) MRT_CS(1, 5) MRT_CS(2, 6) MRT_CS(3, 7)
    MRT_CS(2, 4) MRT_CS(3, 5)
    MRT_CS(1, 2) MRT_CS(3, 4) MRT_CS(5, 6)
#undef MRT_CS
#pragma unroll
    for (int c = 7; c >= 1; c--)
        if (key[c] != MISS) stack_push<SL>(stk, r.sp, entry_of(key[c]));
    if (key[0] != MISS) {
        r.cur = entry_of(key[0]);
    } else {
        r.cur = r.sp > 0 ? stack_pop<SL>(stk, r.sp) : TRAVERSAL_DONE;
    }
}


// One inner-node step over the 8-wide quantised BVH in octant order: slab-test the eight child
// boxes, continue with the first hit child in the ray's visiting order and push the others as
// one group. Fewer instructions per node than the sorted step, but a pop re-reads a node
// header: it wins on long incoherent traversals and loses where pops are frequent.
//
// Slab distances are taken in the node's grid: a plane at grid coordinate q lies at q 2^e + p,
// so its distance is q (2^e / d) + (p / d - o / d) = fma(q, a, b) with a = 2^e ix (exact: a
// power-of-two scale) and b = fma(p, ix, -ox) once per node and axis, one FMA per plane after
// the byte conversion. Visiting order: the builder puts a child in the slot of the node corner
// it lies toward (bit a: high end of axis a), so a ray whose direction signs are r.neg meets
// them roughly in the order slot = i ^ r.neg, i = 0..7 (the ordering of compressed wide BVHs,
// Ylitie, Karras and Laine 2017); the hit mask is permuted into that order by three
// conditional bit-group swaps, and no distances are sorted. Only the visiting order depends on
// it; which children are visited does not.
template <int SL, bool COUNT>
__device__ __forceinline__ void node_step(const DeviceBVH8G bvh, RayState& r, const LaneStack& stk,
                                          TraversalCount& count) {
    if constexpr (COUNT) count.nodes++;
    const uint4* np = reinterpret_cast<const uint4*>(bvh.nodes + r.cur);
    const uint4 h0 = np[0], h1 = np[1], qx = np[2], qy = np[3], qz = np[4];
    const int ex = int(int8_t(h0.w & 0xFFu)), ey = int(int8_t((h0.w >> 8) & 0xFFu)), ez = int(int8_t((h0.w >> 16) & 0xFFu));
    const float ax = ldexpf(r.ix, ex), ay = ldexpf(r.iy, ey), az = ldexpf(r.iz, ez); // 2^e ix, exact
    const float bx = __builtin_fmaf(__uint_as_float(h0.x), r.ix, -r.ox);
    const float by = __builtin_fmaf(__uint_as_float(h0.y), r.iy, -r.oy);
    const float bz = __builtin_fmaf(__uint_as_float(h0.z), r.iz, -r.oz);
    const float tlim = r.tlim;
    // near / far plane bytes per axis, picked by the ray's octant with one byte permute per word
    // (children 0-3 in the low word, 4-7 in the high word)
    const uint32_t sx = r.psel[0], sy = r.psel[1], sz = r.psel[2];
    const uint32_t nx0 = __builtin_amdgcn_perm(qx.z, qx.x, sx), nx1 = __builtin_amdgcn_perm(qx.w, qx.y, sx);
    const uint32_t fx0 = __builtin_amdgcn_perm(qx.x, qx.z, sx), fx1 = __builtin_amdgcn_perm(qx.y, qx.w, sx);
    const uint32_t ny0 = __builtin_amdgcn_perm(qy.z, qy.x, sy), ny1 = __builtin_amdgcn_perm(qy.w, qy.y, sy);
    const uint32_t fy0 = __builtin_amdgcn_perm(qy.x, qy.z, sy), fy1 = __builtin_amdgcn_perm(qy.y, qy.w, sy);
    const uint32_t nz0 = __builtin_amdgcn_perm(qz.z, qz.x, sz), nz1 = __builtin_amdgcn_perm(qz.w, qz.y, sz);
    const uint32_t fz0 = __builtin_amdgcn_perm(qz.x, qz.z, sz), fz1 = __builtin_amdgcn_perm(qz.y, qz.w, sz);
    uint32_t hm = 0u; // bit c: slot c is hit
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const uint32_t sh = 8u * uint32_t(c & 3);
        const float NX = __builtin_fmaf(float(((c < 4 ? nx0 : nx1) >> sh) & 0xFFu), ax, bx);
        const float FX = __builtin_fmaf(float(((c < 4 ? fx0 : fx1) >> sh) & 0xFFu), ax, bx);
        const float NY = __builtin_fmaf(float(((c < 4 ? ny0 : ny1) >> sh) & 0xFFu), ay, by);
        const float FY = __builtin_fmaf(float(((c < 4 ? fy0 : fy1) >> sh) & 0xFFu), ay, by);
        const float NZ = __builtin_fmaf(float(((c < 4 ? nz0 : nz1) >> sh) & 0xFFu), az, bz);
        const float FZ = __builtin_fmaf(float(((c < 4 ? fz0 : fz1) >> sh) & 0xFFu), az, bz);
        const float tn = fmaxf(fmaxf(NX, NY), fmaxf(NZ, r.t0));
        const float tf = fminf(fminf(FX, FY), fminf(FZ, tlim));
        const uint32_t code = ((c < 4 ? h1.z : h1.w) >> sh) & 0xFFu;
        hm |= (tn <= tf && code != BVH8_EMPTY) ? (1u << c) : 0u;
    }
    // position i of the visiting order is slot i ^ neg: swap bit pairs, pairs of pairs, nibbles
    uint32_t pm = hm;
    if (r.neg & 1u) pm = ((pm & 0x55u) << 1) | ((pm >> 1) & 0x55u);
    if (r.neg & 2u) pm = ((pm & 0x33u) << 2) | ((pm >> 2) & 0x33u);
    if (r.neg & 4u) pm = ((pm & 0x0Fu) << 4) | (pm >> 4);
    if (pm == 0u) {
        r.cur = pop_next<SL>(bvh, r, stk);
        return;
    }
    const uint32_t rest = pm & (pm - 1u);
    if (rest != 0u) stack_push<SL>(stk, r.sp, r.cur | (rest << 24));
    r.cur = bvh8_entry(h1, uint32_t(__builtin_ctz(pm)) ^ r.neg);
}

// Moves a leaf entry in `cur` into the (empty) postponed slot and takes the next entry, so the
// lane can keep stepping through inner nodes while its leaf waits for the wave's next leaf
// phase (speculative traversal, Aila & Laine 2009).
template <int SL, class BVH>
__device__ __forceinline__ void postpone_leaf(const BVH& bvh, RayState& r, const LaneStack& stk) {
    if (r.leaf == TRAVERSAL_DONE && is_leaf(r.cur)) {
        r.leaf = r.cur;
        r.cur = pop_next<SL>(bvh, r, stk);
    }
}

// Persistent closest-hit loop of one wave: every lane runs one traversal step per iteration and,
// when its ray is done, hands the hit to `sink` and later takes a fresh ray from `src`, so lanes
// with short rays do not idle while others finish long ones.
//
// Ray indices are handed out in wave-private chunks, so the single device-wide cursor `work`
// (zero at launch) sees a few atomics per wave instead of one per refill: every wave first owns
// a static chunk of n / (2 * waves) rays (rounded down to whole waves of 64), then claims
// guided chunks of (unclaimed rays) / (2 * waves), 64..1024 rays, from `work`.
// Refilled lanes issue their ray loads before the still-active lanes' traversal step, so both
// wait on one memory round trip.
template <int SL, bool COUNT, class Src, class Sink, bool ANY = false, class BVH = DeviceBVH>
__device__ __forceinline__ void trace_persistent(const BVH bvh, const Src& src, const Sink& sink, uint32_t n,
                                                 uint32_t* work, uint32_t ray_mask, const LaneStack& stk,
                                                 TraversalCount& count, uint32_t refill, uint32_t leaf_lanes) {
    const uint32_t lane = __lane_id();
    const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t c0 = (n / (2u * waves)) & ~63u;
    const uint32_t dyn0 = c0 * waves; // <= n / 2: dynamic claims start here
    uint32_t pool = wave * c0, pool_end = pool + c0;
    uint32_t seen = dyn0;
    bool claims_done = dyn0 >= n;
    RayState r;
    r.cur = TRAVERSAL_DONE;
    bool active = false;
    uint32_t idx = 0;
    while (true) {
        const uint64_t idle = __ballot(!active);
        const uint32_t nidle = uint32_t(__popcll(idle));
        bool fresh = false;
        if (nidle >= refill && !(claims_done && pool == pool_end)) {
            const uint32_t avail = pool_end - pool;
            uint32_t lo = 0, hi = 0;
            if (avail < nidle && !claims_done) {
                const uint32_t left = n - (seen < n ? seen : n);
                const uint32_t chunk = min(max((left / (2u * waves)) & ~63u, 64u), 1024u);
                const uint32_t leader = uint32_t(__ffsll((unsigned long long)idle) - 1);
                uint32_t got = 0;
                if (lane == leader) got = atomicAdd(work, chunk);
                const uint64_t base = uint64_t(__shfl(got, int(leader))) + dyn0;
                seen = base < n ? uint32_t(base) : n;
                if (base < n) {
                    lo = uint32_t(base);
                    hi = uint32_t(base + chunk < n ? base + chunk : n);
                }
                if (base + chunk >= n) claims_done = true;
            }
            if (!active) {
                const uint32_t k = uint32_t(__popcll(idle & ((1ull << lane) - 1ull)));
                uint32_t my = TRAVERSAL_DONE;
                if (k < avail) {
                    my = pool + k;
                } else if (k - avail < hi - lo) {
                    my = lo + (k - avail);
                }
                if (my != TRAVERSAL_DONE) {
                    idx = my;
                    src.load(idx, r.o, r.d, r.tmin, r.tmax); // this lane's state is dead until ray_begin
                    fresh = true;
                }
            }
            if (avail >= nidle) {
                pool += nidle;
            } else {
                pool = lo + min(nidle - avail, hi - lo);
                pool_end = hi;
            }
        }
        // One wave-uniform phase per iteration: an inner-node step for every lane whose current
        // entry is an inner node, or a leaf test for every lane with a postponed leaf. Leaf
        // phases run once `leaf_lanes` lanes can do nothing but their leaf (or no lane has an
        // inner node left), so both code paths run with most lanes active.
        const bool inner_ok = active && is_inner(r.cur);
        const bool has_leaf = active && r.leaf != TRAVERSAL_DONE;
        const uint64_t mi = __ballot(inner_ok);
        const uint64_t ml = __ballot(has_leaf);
        if ((mi | ml) == 0 && __ballot(fresh) == 0) break;
        const bool leaf_phase = mi == 0 || uint32_t(__popcll(ml & ~mi)) >= leaf_lanes;
        if (leaf_phase) {
            if (has_leaf) leaf_test<SL, COUNT, ANY>(bvh.tris, r, ray_mask, count);
        } else {
            if (inner_ok) node_step<SL, COUNT>(bvh, r, stk, count);
        }
        if (active) {
            postpone_leaf<SL>(bvh, r, stk);
            if (r.cur == TRAVERSAL_DONE && r.leaf == TRAVERSAL_DONE) {
                sink.store(idx, r.best);
                active = false;
            }
        }
        if (fresh) {
            ray_begin(r, r.o, r.d, r.tmin, r.tmax);
            active = true;
        }
    }
}

} // namespace mrt
