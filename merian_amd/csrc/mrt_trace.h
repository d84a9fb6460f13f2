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
__device__ __forceinline__ void test_triangle(const float4* __restrict__ tris, uint32_t rec, f3 o, f3 d, float tmin,
                                              float tmax, uint32_t ray_mask, HitRec& best) {
    const float4 A = tris[3 * rec + 0];
    const float4 B = tris[3 * rec + 1];
    const float4 Cc = tris[3 * rec + 2];
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

// stack: this lane's column base in LDS; stride = threads per block.
template <int STACK, bool COUNT = false>
__device__ __forceinline__ HitRec traverse(const DeviceBVH bvh, f3 o, f3 d, float tmin, float tmax, uint32_t ray_mask,
                                           uint32_t* stack, int stride, TraversalCount* count = nullptr) {
    HitRec best{TRAVERSAL_DONE, __builtin_inff(), 0.f, 0.f};
    const float dx = fabsf(d.x) < 1e-20f ? copysignf(1e-20f, d.x) : d.x;
    const float dy = fabsf(d.y) < 1e-20f ? copysignf(1e-20f, d.y) : d.y;
    const float dz = fabsf(d.z) < 1e-20f ? copysignf(1e-20f, d.z) : d.z;
    const float ix = 1.0f / dx, iy = 1.0f / dy, iz = 1.0f / dz;
    const float ox = o.x * ix, oy = o.y * iy, oz = o.z * iz;
    const bool nx = ix < 0.f, ny = iy < 0.f, nz = iz < 0.f;
    const float t0 = tmin - fabsf(tmin) * 2.4e-7f;
    int sp = 0;
    uint32_t cur = 0u; // root node
    while (true) {
        if (cur & LEAF_FLAG) {
            const uint32_t first = (cur & 0x7FFFFFFFu) >> 3;
            const uint32_t cnt = (cur & 7u) + 1u;
            for (uint32_t k = 0; k < cnt; k++) test_triangle(bvh.tris, first + k, o, d, tmin, tmax, ray_mask, best);
            if constexpr (COUNT) count->tris += cnt;
            cur = sp > 0 ? stack[(--sp) * stride] : TRAVERSAL_DONE;
        } else {
            if constexpr (COUNT) count->nodes++;
            const float4* np = reinterpret_cast<const float4*>(bvh.nodes + cur);
            const float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
            const int4 ch = reinterpret_cast<const int4*>(np)[6];
            const uint32_t counts = reinterpret_cast<const uint32_t*>(np)[28];
            const float tlim = fminf(tmax, best.t) * 1.00000024f;
            float key[4];
            uint32_t ent[4];
#define MRT_CHILD(c, LX, HX, LY, HY, LZ, HZ, CH)                                                        \
    {                                                                                                   \
        const float tnx = __builtin_fmaf(nx ? HX : LX, ix, -ox), tfx = __builtin_fmaf(nx ? LX : HX, ix, -ox); \
        const float tny = __builtin_fmaf(ny ? HY : LY, iy, -oy), tfy = __builtin_fmaf(ny ? LY : HY, iy, -oy); \
        const float tnz = __builtin_fmaf(nz ? HZ : LZ, iz, -oz), tfz = __builtin_fmaf(nz ? LZ : HZ, iz, -oz); \
        const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, t0));                                       \
        const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tlim));                                     \
        const uint32_t cc = (counts >> (8 * c)) & 0xFFu;                                                \
        key[c] = (tn <= tf) ? tn : __builtin_inff();                                                    \
        ent[c] = cc ? (LEAF_FLAG | (uint32_t(CH) << 3) | (cc - 1u)) : uint32_t(CH);                     \
    }
            MRT_CHILD(0, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, ch.x)
            MRT_CHILD(1, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, ch.y)
            MRT_CHILD(2, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, ch.z)
            MRT_CHILD(3, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, ch.w)
#undef MRT_CHILD
            cswap(key[0], ent[0], key[1], ent[1]);
            cswap(key[2], ent[2], key[3], ent[3]);
            cswap(key[0], ent[0], key[2], ent[2]);
            cswap(key[1], ent[1], key[3], ent[3]);
            cswap(key[1], ent[1], key[2], ent[2]);
            const float inf = __builtin_inff();
            if (key[3] != inf) stack[(sp++) * stride] = ent[3];
            if (key[2] != inf) stack[(sp++) * stride] = ent[2];
            if (key[1] != inf) stack[(sp++) * stride] = ent[1];
            if (key[0] != inf) {
                cur = ent[0];
            } else {
                cur = sp > 0 ? stack[(--sp) * stride] : TRAVERSAL_DONE;
            }
        }
        if (cur == TRAVERSAL_DONE) break;
    }
    return best;
}

} // namespace mrt
