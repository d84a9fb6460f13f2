// gfx950 kernels of the path tracer: ray queries, camera rays, the GBuffer primary pass and
// the wavefront pt loop (init -> [trace -> shade] x bounces -> accumulate).
//
// The reference runs pt as one ray-generation megakernel per pixel
// (include/merian-graph/nodes/render_pt/render_pt.slang:92-130) over driver ray queries. Here
// each bounce is two kernels over a device queue of live paths: a lean traversal kernel (LDS
// stack, low VGPR count, high occupancy) and a shading kernel that samples the material and
// appends the continuation ray with one wave-aggregated atomic per wavefront.
#include "mrt_guide.h"
#include "mrt_kernels.h"
#include "mrt_mcpg.h"
#include "mrt_trace.h"

#include <algorithm>
#include <vector>

namespace mrt {

namespace {

constexpr int SHADE_BLOCK = 256;
constexpr int SHADE_ROUNDS = 2;                          // paths per thread per tile
constexpr int SHADE_TILE = SHADE_BLOCK * SHADE_ROUNDS;   // paths a block groups by class at once
// Passes of at least SHADE_WIDE_PATHS paths shade tiles of SHADE_WIDE_ROUNDS paths per thread:
// wider class groups and a more coherent continuation queue, while every resident block still
// gets many tiles (at 1920x1080, 64 spp: 1.0 % shorter frames, 1.8 % with NEE; a 512x512
// 4-spp pass, 1 M paths, runs 2.3 % slower with them).
constexpr int SHADE_WIDE_ROUNDS = 3;
constexpr uint32_t SHADE_WIDE_PATHS = 1u << 23;
constexpr float SCENE_RAY_TMAX = 1e7f;     // scene/hit.slang:8
constexpr float ENV_MAP_DISTANCE = 1e7f;   // scene/environment-map.slang:11
constexpr float THROUGHPUT_THRESHOLD = 1e-5f; // scene/scene.slang:24

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Streamed path data (queues, hits, path state) is touched once per bounce; non-temporal
// accesses keep it from evicting BVH nodes and triangles from L2 and the Infinity Cache.
typedef float v4f __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_stream(const float4* p) {
    const v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st_stream(float4* p, float4 v) {
    const v4f t = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(t, reinterpret_cast<v4f*>(p));
}

// -----------------------------------------------------------------------------------------
// camera (camera.slang:45-56)

struct RayF {
    f3 o, d;
    float tmin, tmax;
};

__device__ __forceinline__ RayF camera_ray(const mrt_camera& c, uint32_t px, uint32_t py, uint32_t w, uint32_t h) {
    const float p_x = ((float(px) + 0.5f) + c.jitter[0]) / float(w);
    const float p_y = ((float(py) + 0.5f) + c.jitter[1]) / float(h);
    const float ndc_x = 2.0f * p_x + -1.0f;
    const float ndc_y = -2.0f * p_y + 1.0f;
    const f3 U = ld3(c.U), V = ld3(c.V), W = ld3(c.W);
    RayF r;
    r.o = ld3(c.position);
    r.d = normalize(ndc_x * U + ndc_y * V + W);
    const float inv_cos = 1.f / dot(normalize(W), r.d);
    r.tmin = c.near_plane * inv_cos;
    r.tmax = c.far_plane * inv_cos;
    return r;
}

// Ray sources and hit sinks of the persistent traversal loop.
struct SrcRays { // mrt_ray array: (origin, tmin), (direction, tmax)
    const mrt_ray* rays;
    __device__ __forceinline__ void load(uint32_t i, f3& o, f3& d, float& tmin, float& tmax) const {
        const float4* rp = reinterpret_cast<const float4*>(rays + i);
        const float4 a = rp[0], b = rp[1];
        o = f3{a.x, a.y, a.z};
        d = f3{b.x, b.y, b.z};
        tmin = a.w;
        tmax = b.w;
    }
};
struct SrcQueue { // path queue: (origin, path id), (direction, -); continuation rays [0, 1e7]
    const float4* qo;
    const float4* qd;
    __device__ __forceinline__ void load(uint32_t i, f3& o, f3& d, float& tmin, float& tmax) const {
        const float4 a = ld_stream(qo + i), b = ld_stream(qd + i);
        o = f3{a.x, a.y, a.z};
        d = f3{b.x, b.y, b.z};
        tmin = 0.f;
        tmax = SCENE_RAY_TMAX;
    }
};
struct SrcCamera { // primary rays of a w x h image
    mrt_camera cam;
    uint32_t w, h;
    __device__ __forceinline__ void load(uint32_t i, f3& o, f3& d, float& tmin, float& tmax) const {
        const RayF r = camera_ray(cam, i % w, i / w, w, h);
        o = r.o;
        d = r.d;
        tmin = r.tmin;
        tmax = r.tmax;
    }
};
struct SinkHits { // mrt_hit array
    mrt_hit* hits;
    __device__ __forceinline__ void store(uint32_t i, const HitRec& h) const {
        const bool miss = h.id == TRAVERSAL_DONE;
        reinterpret_cast<float4*>(hits)[i] = make_float4(__uint_as_float(h.id), miss ? __builtin_inff() : h.t,
                                                         miss ? 0.f : h.u, miss ? 0.f : h.v);
    }
};
struct SinkQueue { // hits of the path queue: (t, u, v, id)
    float4* hits;
    __device__ __forceinline__ void store(uint32_t i, const HitRec& h) const {
        st_stream(hits + i, make_float4(h.t, h.u, h.v, __uint_as_float(h.id)));
    }
};
struct SinkGBuffer { // PackedHitInfo (gbuffer.slang:175-196); one pretransformed group
    const uint4* tri_verts;
    const uint32_t* geom_first;
    mrt_packed_hit* out;
    __device__ __forceinline__ void store(uint32_t i, const HitRec& h) const {
        uint4 p;
        if (h.id == TRAVERSAL_DONE) {
            p = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u);
        } else {
            const uint32_t g = tri_verts[h.id].w;
            p.x = 0u;                           // instance_id
            p.y = h.id - geom_first[g];         // primitive_id within the geometry
            p.z = (g & 0xFFFFu);                // geometry_index | instance_index << 16
            p.w = uint32_t(float_to_half(h.u)) | (uint32_t(float_to_half(h.v)) << 16);
        }
        reinterpret_cast<uint4*>(out)[i] = p;
    }
};

struct SrcShadow { // NEE shadow queue: (origin, radiance target), (direction, -); [0, 1e7] like a continuation ray
    const float4* so;
    const float4* sd;
    __device__ __forceinline__ void load(uint32_t i, f3& o, f3& d, float& tmin, float& tmax) const {
        const float4 a = ld_stream(so + i), b = ld_stream(sd + i);
        o = f3{a.x, a.y, a.z};
        d = f3{b.x, b.y, b.z};
        tmin = 0.f;
        tmax = SCENE_RAY_TMAX;
    }
};
// Where a shadow ray's contribution goes (the w of its origin): the radiance of the path's
// continuation ray in the next queue (its slot), or, for a path that ended at the vertex, the
// path's own record (SHADOW_TO_PATH | path index).
constexpr uint32_t SHADOW_TO_PATH = 0x80000000u;
struct SinkShadow { // unoccluded: add the queued contribution to the path's radiance
    const float4* so;
    const float4* sc;
    float4* state; // PathBuffers.state
    float4* qr;    // radiance column of the queue the shaded vertices' continuation rays went to
    __device__ __forceinline__ void store(uint32_t i, const HitRec& h) const {
        if (h.id != TRAVERSAL_DONE) return;
        const uint32_t to = __float_as_uint(so[i].w);
        const float4 c = ld_stream(sc + i);
        float4* const p = (to & SHADOW_TO_PATH) ? state + 2 * size_t(to & ~SHADOW_TO_PATH) + 1 : qr + to;
        float4 b = *p;
        b.x = b.x + c.x;
        b.y = b.y + c.y;
        b.z = b.z + c.z;
        *p = b;
    }
};

// Adds this lane's traversal counts to counters[0] (node visits) and counters[1] (triangle
// tests) with one atomic per wave.
__device__ __forceinline__ void flush_counts(const TraversalCount& c, unsigned long long* counters) {
    unsigned long long n = c.nodes, t = c.tris;
    for (int off = 32; off > 0; off >>= 1) {
        n += __shfl_xor(n, off);
        t += __shfl_xor(t, off);
    }
    if (__lane_id() == 0) {
        atomicAdd(counters + 0, n);
        atomicAdd(counters + 1, t);
    }
}

// -----------------------------------------------------------------------------------------
// ray queries

// Closest-hit kernel for every ray stream (ray queries, the GBuffer's primary rays, the path
// queue): a persistent grid, one ray per lane with refill, LDS stack + global overflow.
// Ray count: *count_ptr when given (device-side queue size), else n_fixed.
// ANY: visibility queries (first hit ends the traversal).
// W: traversal of the scene's tree (4: BVH4Node; 8: quantised BVH8Node, sorted entries; 9: the
// same nodes in octant order with grouped stack entries).
template <bool COUNT, class Src, class Sink, bool ANY = false, int W = 4>
__global__ __launch_bounds__(TRACE_BLOCK) void k_trace(DeviceScene S, Src src, Sink sink,
                                                       const uint32_t* __restrict__ count_ptr, uint32_t n_fixed,
                                                       uint32_t* __restrict__ work, uint32_t mask, TraceScratch ts,
                                                       unsigned long long* counters) {
    __shared__ uint32_t stack[TRACE_LDS_STACK * TRACE_BLOCK];
    const uint32_t n = count_ptr ? *count_ptr : n_fixed;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const LaneStack stk{stack + threadIdx.x, ts.ovf + gtid, gridDim.x * blockDim.x};
    TraversalCount cnt;
    if constexpr (W == 9) {
        trace_persistent<TRACE_LDS_STACK, COUNT, Src, Sink, ANY>(DeviceBVH8G{S.nodes8, S.tris}, src, sink, n, work,
                                                                 mask, stk, cnt, ts.refill, ts.leaf_lanes);
    } else if constexpr (W == 8) {
        trace_persistent<TRACE_LDS_STACK, COUNT, Src, Sink, ANY>(DeviceBVH8{S.nodes8, S.tris}, src, sink, n, work,
                                                                 mask, stk, cnt, ts.refill, ts.leaf_lanes);
    } else {
        trace_persistent<TRACE_LDS_STACK, COUNT, Src, Sink, ANY>(DeviceBVH{S.nodes, S.tris}, src, sink, n, work, mask,
                                                                 stk, cnt, ts.refill, ts.leaf_lanes);
    }
    if constexpr (COUNT) flush_counts(cnt, counters);
}

__global__ void k_camera_rays(mrt_camera cam, uint32_t w, uint32_t h, mrt_ray* __restrict__ rays) {
    const uint64_t n = uint64_t(w) * h;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const RayF r = camera_ray(cam, uint32_t(i % w), uint32_t(i / w), w, h);
        float4* o = reinterpret_cast<float4*>(rays + i);
        o[0] = make_float4(r.o.x, r.o.y, r.o.z, r.tmin);
        o[1] = make_float4(r.d.x, r.d.y, r.d.z, r.tmax);
    }
}

// -----------------------------------------------------------------------------------------
// environment (environment-map.slang:13-68, utils/math.slang:243-249)

__device__ __forceinline__ f3 env_texel(const DeviceScene& S, int x, int y) {
    const int w = int(S.env_w), h = int(S.env_h);
    if (S.env_wrap == MRT_WRAP_REPEAT) {
        x = ((x % w) + w) % w;
        y = ((y % h) + h) % h;
    } else {
        x = min(max(x, 0), w - 1);
        y = min(max(y, 0), h - 1);
    }
    const float4 t = S.env[size_t(y) * w + x];
    return {t.x, t.y, t.z};
}
__device__ __forceinline__ float tlerp(float x, float y, float t) { return x + t * (y - x); }
__device__ __forceinline__ f3 tlerp(f3 x, f3 y, float t) {
    return {tlerp(x.x, y.x, t), tlerp(x.y, y.y, t), tlerp(x.z, y.z, t)};
}
__device__ __forceinline__ f3 env_eval(const DeviceScene& S, f3 dir) {
    if (S.env_type == MRT_ENV_NONE) return {0, 0, 0};
    const float* m = S.env_to_local;
    const f3 d{m[0] * dir.x + m[1] * dir.y + m[2] * dir.z, m[3] * dir.x + m[4] * dir.y + m[5] * dir.z,
               m[6] * dir.x + m[7] * dir.y + m[8] * dir.z};
    f2 uv;
    if (S.env_type == MRT_ENV_LATLONG) {
        const f3 p = normalize(d);
        uv.x = matan2(p.x, -p.z) * INV_2PI + 0.5f;
        uv.y = macos(p.y) * M_1_PI_F;
    } else {
        const f3 a{fabsf(d.x), fabsf(d.y), fabsf(d.z)};
        const float r = sqrtf(fmaxf(0.f, 1.f - a.z));
        const float mm = fmaxf(a.x, a.y);
        const float b = mm == 0.f ? 0.f : fminf(a.x, a.y) / mm;
        float phi = matan(b) * (2.f / PI);
        if (a.x < a.y) phi = 1.f - phi;
        float v = phi * r;
        float u = r - v;
        if (d.z < 0.f) {
            const float mu = u;
            u = 1.f - v;
            v = 1.f - mu;
        }
        u = d.x < 0.f ? -u : u;
        v = d.y < 0.f ? -v : v;
        uv = {0.5f * (u + 1.f), 0.5f * (v + 1.f)};
    }
    const float fx = uv.x * float(S.env_w) - 0.5f;
    const float fy = uv.y * float(S.env_h) - 0.5f;
    const float x0f = floorf(fx), y0f = floorf(fy);
    const float ax = fx - x0f, ay = fy - y0f;
    const int x0 = int(x0f), y0 = int(y0f);
    const f3 c = tlerp(tlerp(env_texel(S, x0, y0), env_texel(S, x0 + 1, y0), ax),
                       tlerp(env_texel(S, x0, y0 + 1), env_texel(S, x0 + 1, y0 + 1), ax), ay);
    return S.env_intensity * c;
}

// -----------------------------------------------------------------------------------------
// environment importance sampling for NEE (no counterpart in the reference's pt): texel by the
// alias table, uniform position inside the texel, mapped to a direction. Solid-angle pdf:
//   lat-long   p_texel * W * H / (2 pi^2 sin(theta))   (d omega = 2 pi^2 sin(theta) du dv)
//   equal-area p_texel * W * H / (4 pi)                (area preserving)

// Clarberg's equal-area square -> sphere map (pbrt-v4 EqualAreaSquareToSphere); inverse of the
// lookup in env_eval.
__device__ __forceinline__ f3 equal_area_square_to_sphere(float pu, float pv) {
    const float u = 2.f * pu - 1.f, v = 2.f * pv - 1.f;
    const float up = fabsf(u), vp = fabsf(v);
    const float sd = 1.f - (up + vp);
    const float r = 1.f - fabsf(sd);
    const float phi = (r == 0.f ? 1.f : (vp - up) / r + 1.f) * (PI / 4.f);
    const float z = copysignf(1.f - r * r, sd);
    float sn, cs;
    msincos(phi, sn, cs);
    const float k = r * sqrtf(fmaxf(0.f, 2.f - r * r));
    return {copysignf(cs, u) * k, copysignf(sn, v) * k, z};
}

__device__ __forceinline__ f3 env_to_world(const DeviceScene& S, f3 l) { // to_local^T (a rotation)
    const float* m = S.env_to_local;
    return {m[0] * l.x + m[3] * l.y + m[6] * l.z, m[1] * l.x + m[4] * l.y + m[7] * l.z,
            m[2] * l.x + m[5] * l.y + m[8] * l.z};
}

// Solid-angle density with which env_sample produces world direction `dir`.
__device__ __forceinline__ float env_pdf(const DeviceScene& S, f3 dir) {
    if (S.env_type == MRT_ENV_NONE) return 0.f;
    const float* m = S.env_to_local;
    const f3 d = normalize(f3{m[0] * dir.x + m[1] * dir.y + m[2] * dir.z, m[3] * dir.x + m[4] * dir.y + m[5] * dir.z,
                              m[6] * dir.x + m[7] * dir.y + m[8] * dir.z});
    const float n = float(S.env_w) * float(S.env_h);
    if (S.env_type == MRT_ENV_LATLONG) {
        const float u = matan2(d.x, -d.z) * INV_2PI + 0.5f;
        const float v = macos(d.y) * M_1_PI_F;
        const int x = min(max(int(u * float(S.env_w)), 0), int(S.env_w) - 1);
        const int y = min(max(int(v * float(S.env_h)), 0), int(S.env_h) - 1);
        const float sin_t = sqrtf(fmaxf(0.f, 1.f - d.y * d.y));
        if (sin_t <= 0.f) return 0.f;
        return S.env_samp[size_t(y) * S.env_w + x].z * n / (2.f * PI * PI * sin_t);
    }
    const f3 a{fabsf(d.x), fabsf(d.y), fabsf(d.z)};
    const float r = sqrtf(fmaxf(0.f, 1.f - a.z));
    const float mm = fmaxf(a.x, a.y);
    const float b = mm == 0.f ? 0.f : fminf(a.x, a.y) / mm;
    float phi = matan(b) * (2.f / PI);
    if (a.x < a.y) phi = 1.f - phi;
    float v = phi * r;
    float u = r - v;
    if (d.z < 0.f) {
        const float mu = u;
        u = 1.f - v;
        v = 1.f - mu;
    }
    u = 0.5f * ((d.x < 0.f ? -u : u) + 1.f);
    v = 0.5f * ((d.y < 0.f ? -v : v) + 1.f);
    const int x = min(max(int(u * float(S.env_w)), 0), int(S.env_w) - 1);
    const int y = min(max(int(v * float(S.env_h)), 0), int(S.env_h) - 1);
    return S.env_samp[size_t(y) * S.env_w + x].z * n / (4.f * PI);
}

// Samples a world direction toward the environment; returns false for an empty environment.
__device__ __forceinline__ bool env_sample(const DeviceScene& S, PCG& rng, f3& dir, float& pdf) {
    if (S.env_type == MRT_ENV_NONE) return false;
    const uint32_t n = S.env_w * S.env_h;
    const float r0 = rng.next_float(), r1 = rng.next_float();
    const float r2 = rng.next_float(), r3 = rng.next_float();
    uint32_t i = min(uint32_t(r0 * float(n)), n - 1u);
    const float4 e = S.env_samp[i];
    if (r1 >= e.x) i = __float_as_uint(e.y);
    const float p_texel = S.env_samp[i].z;
    if (!(p_texel > 0.f)) return false;
    const uint32_t x = i % S.env_w, y = i / S.env_w;
    const float u = (float(x) + r2) / float(S.env_w);
    const float v = (float(y) + r3) / float(S.env_h);
    f3 l;
    if (S.env_type == MRT_ENV_LATLONG) {
        float st, ct, sp, cp;
        msincos(v * PI, st, ct);
        msincos((u - 0.5f) * TWO_PI, sp, cp);
        if (st <= 0.f) return false;
        l = f3{st * sp, ct, -(st * cp)};
        pdf = p_texel * float(n) / (2.f * PI * PI * st);
    } else {
        l = equal_area_square_to_sphere(u, v);
        pdf = p_texel * float(n) / (4.f * PI);
    }
    dir = normalize(env_to_world(S, l));
    return pdf > 0.f;
}

__device__ __forceinline__ float power_heuristic(float a, float b) {
    const float a2 = a * a;
    return a2 / (a2 + b * b);
}

// -----------------------------------------------------------------------------------------
// shading points (scene.slang:312-498, hit.slang:106-147)

struct ShadingPoint {
    f3 pos;
    uint32_t material; // material entry of a surface hit
    uint32_t enc_face_normal;
    bool env; // the ray left the scene (or hit a UseEnvMap mesh)
    f3 emission;
    Frame frame;
    bool has_material;
    MaterialSample mat;
};

__device__ __forceinline__ f3 xyz(float4 v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ f2 uv_of(uint32_t bits) {
    return {half_to_float(uint16_t(bits & 0xFFFFu)), half_to_float(uint16_t(bits >> 16))};
}

// Vertex data of a hit triangle from its 64-byte shading record (DeviceScene.tri_shade: one
// gather per hit instead of the vertex ids and three vertex records); tangents, which only
// geometries with HasTangents use, come from the vertex records.
__device__ __forceinline__ void shading_point(const DeviceScene& S, f3 ro, f3 rd, uint32_t gid, float bu, float bv, float ior,
                              bool with_material, ShadingPoint& sp) {
    float4 r0{}, r1{}, r2{}, r3{};
    uint32_t geom = 0;
    bool surface = gid != TRAVERSAL_DONE;
    if (surface) {
        const float4* t = S.tri_shade + 4 * size_t(gid);
        r0 = t[0];
        r1 = t[1];
        r2 = t[2];
        r3 = t[3];
        geom = __float_as_uint(r3.w) & TRI_SHADE_GEOM_MASK;
        surface = (S.geoms[geom].flags & MRT_GEOM_USE_ENV_MAP) == 0;
    }
    if (!surface) {
        sp.pos = ro + rd * ENV_MAP_DISTANCE;
        sp.enc_face_normal = encode_normal(-rd);
        sp.emission = env_eval(S, -(-rd));
        sp.frame = frame_from_normal(-rd);
        sp.has_material = false;
        sp.env = true;
        return;
    }
    sp.env = false;
    const mrt_geometry G = S.geoms[geom];
    const f3 b{(1.0f - bu) - bv, bu, bv};
    const f3 p0 = xyz(r0), p1 = xyz(r1), p2 = xyz(r2);
    const f3 pos = b.x * p0 + b.y * p1 + b.z * p2;
    const f3 e01 = p1 - p0, e02 = p2 - p0;
    f3 face_normal = normalize(cross(e01, e02));
    if (G.flags & MRT_GEOM_FLIP_FACING) face_normal = -face_normal;
    f3 vn;
    if (G.flags & MRT_GEOM_FLAT_SHADING) {
        vn = face_normal;
    } else {
        vn = normalize(b.x * decode_normal(__float_as_uint(r0.w)) + b.y * decode_normal(__float_as_uint(r1.w)) +
                       b.z * decode_normal(__float_as_uint(r2.w)));
    }
    const f2 uv0 = uv_of(__float_as_uint(r3.x)), uv1 = uv_of(__float_as_uint(r3.y)), uv2 = uv_of(__float_as_uint(r3.z));
    Frame frame;
    if (G.flags & MRT_GEOM_HAS_TANGENTS) {
        const uint4 tv = S.tri_verts[gid];
        const mrt_vertex v0 = S.verts[tv.x], v1 = S.verts[tv.y], v2 = S.verts[tv.z];
        const float w0 = (v0.encoded_tangent & 1u) ? -1.f : 1.f;
        const float w1 = (v1.encoded_tangent & 1u) ? -1.f : 1.f;
        const float w2 = (v2.encoded_tangent & 1u) ? -1.f : 1.f;
        const f3 txyz = b.x * decode_normal(v0.encoded_tangent & ~1u) + b.y * decode_normal(v1.encoded_tangent & ~1u) +
                        b.z * decode_normal(v2.encoded_tangent & ~1u);
        const float tw = b.x * w0 + b.y * w1 + b.z * w2;
        frame = frame_from_nt(vn, normalize(txyz), signf(tw));
    } else {
        const f2 d1{uv1.x - uv0.x, uv1.y - uv0.y};
        const f2 d2{uv2.x - uv0.x, uv2.y - uv0.y};
        const float det = d1.x * d2.y - d2.x * d1.y;
        if (fabsf(det) > 1e-8f) {
            const float ws = (G.flags & MRT_GEOM_FLIP_FACING) ? 1.f : -1.f;
            frame = frame_from_nt(vn, (d2.y * e01 - d1.y * e02) / det, signf(det) * ws);
        } else {
            frame = frame_from_normal(vn);
        }
    }
    bool front = true;
    if (G.flags & MRT_GEOM_TWO_SIDED) {
        front = dot(-rd, face_normal) >= 0.f;
        if (!front) frame.r2 = -frame.r2;
    }
    sp.pos = pos;
    sp.material = G.material_id;
    sp.enc_face_normal = encode_normal(face_normal);
    sp.has_material = with_material;
    const mrt_material& m = S.mats[G.material_id];
    if (with_material) {
        ShadingData sd{frame, ior, front};
        sp.mat = get_material_sample(m, sd, sp.emission, sp.frame);
    } else {
        sp.emission = material_emission(m, front);
        sp.frame = frame;
    }
}

// -----------------------------------------------------------------------------------------
// wavefront pt

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

// Global triangle id and barycentrics of a gbuffer texel (PackedHitInfo, gbuffer.slang:40-61);
// misses, and hits of a stale or foreign gbuffer that name no triangle of this scene, give
// TRAVERSAL_DONE (shaded as misses).
__device__ __forceinline__ uint32_t gbuffer_hit(const DeviceScene& S, const mrt_packed_hit* gbuffer, uint32_t p,
                                                float& bu, float& bv) {
    const uint4 gb = reinterpret_cast<const uint4*>(gbuffer)[p];
    uint32_t gid = TRAVERSAL_DONE;
    bu = bv = 0.f;
    if (gb.x != 0xFFFFFFFFu) {
        const uint32_t geometry = gb.x + (gb.z & 0xFFFFu); // GeometryID = instance_id + geometry_index
        if (geometry < S.n_geoms && gb.y < S.geoms[geometry].primitive_count) gid = S.geom_first[geometry] + gb.y;
        bu = half_to_float(uint16_t(gb.w & 0xFFFFu));
        bv = half_to_float(uint16_t(gb.w >> 16));
    }
    return gid;
}

// Seeds one path per (pixel, sample) of the chunk with the primary ray and gbuffer hit.
__global__ void k_pt_init(DeviceScene S, PtArgs A, const mrt_packed_hit* __restrict__ gbuffer, PathBuffers P,
                          uint32_t* count0) {
    const uint32_t stride = gridDim.x * blockDim.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) *count0 = A.n_paths;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n_paths; i += stride) {
        const uint32_t p = i % A.npix, sl = i / A.npix;
        const uint32_t x = p % A.width, y = p / A.width;
        const uint32_t s = A.s0 + sl;
        const RayF r = camera_ray(A.cam, x, y, A.width, A.height);
        float bu, bv;
        const uint32_t gid = gbuffer_hit(S, gbuffer, p, bu, bv);
        st_stream(P.q_o[0] + i, make_float4(r.o.x, r.o.y, r.o.z, __uint_as_float(i)));
        st_stream(P.q_d[0] + i, make_float4(r.d.x, r.d.y, r.d.z, 0.f));
        st_stream(P.hits + i, make_float4(0.f, bu, bv, __uint_as_float(gid)));
        // queue slot i holds path i: the state goes to either place at the same index
        float4* const st = state_in_queue(A.mode) ? P.q_t[0] : P.state;
        float4* const sr = state_in_queue(A.mode) ? P.q_r[0] : P.state + 1;
        const size_t k = state_in_queue(A.mode) ? size_t(i) : 2 * size_t(i);
        st_stream(st + k, make_float4(1.f, 1.f, 1.f, 1.f));
        st_stream(sr + k, make_float4(0.f, 0.f, 0.f, __uint_as_float(pcg4d16(x, y, A.frame, s))));
        if (A.record_signature) P.sig[i] = sig_step(0x9E3779B9u * (s + 1u), gid);
    }
}

// Block-level append of continuation rays: lanes stage their ray in LDS (one LDS atomic per
// wave), then the block reserves its range of the queue with one global atomic and copies the
// staged rays out contiguously. One global atomic per SHADE_TILE rays instead of one per wave
// keeps the queue counter (a single address every block of the grid appends to) off the
// critical path.
// STATE: the path's state (throughput and ior, radiance and RNG) is staged with the ray and
// lands in the same queue slot.
struct RayStage {
    float4* o;      // one entry per path of the kernel's tile each
    float4* d;
    float4* t;      // path state, STATE stages only
    float4* r;
    uint32_t* meta; // [0] staged count, [1] reserved queue base, [2] count being flushed
};

// Returns the ray's index among the staged rays (the flush writes it to the reserved base plus
// that index); meaningless for a lane that does not push.
template <bool STATE = false>
__device__ __forceinline__ uint32_t stage_push(const RayStage& st, bool push, float4 o, float4 d, float4 t = float4{},
                                               float4 r = float4{}) {
    const uint64_t b = __ballot(push);
    if (!b) return 0u;
    const uint32_t lane = lane_id();
    const uint32_t leader = uint32_t(__ffsll((unsigned long long)b) - 1);
    uint32_t slot = 0;
    if (lane == leader) slot = atomicAdd(st.meta, uint32_t(__popcll(b)));
    slot = __shfl(slot, int(leader)) + uint32_t(__popcll(b & ((1ull << lane) - 1ull)));
    if (push) {
        st.o[slot] = o;
        st.d[slot] = d;
        if constexpr (STATE) {
            st.t[slot] = t;
            st.r[slot] = r;
        }
    }
    return slot;
}

// Block-wide append without staging: slots within the block from an LDS counter, the block's
// range from one global atomic. Every thread of the block calls this (it synchronises the
// block twice); `scratch` is two LDS words.
__device__ __forceinline__ uint32_t block_append(bool push, uint32_t* __restrict__ counter, uint32_t* scratch) {
    const uint64_t b = __ballot(push);
    const uint32_t lane = lane_id();
    const uint32_t leader = b ? uint32_t(__ffsll((unsigned long long)b) - 1) : 0u;
    uint32_t off = 0;
    if (b && lane == leader) off = atomicAdd(scratch, uint32_t(__popcll(b)));
    off = __shfl(off, int(leader)) + uint32_t(__popcll(b & ((1ull << lane) - 1ull)));
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t n = scratch[0];
        scratch[1] = n ? atomicAdd(counter, n) : 0u;
        scratch[0] = 0u;
    }
    __syncthreads();
    return scratch[1] + off;
}

// Every thread of the block calls this; at most the stage's capacity (its kernel's tile) may be
// staged between calls.
// nt, nr: the queue's path-state arrays (STATE stages only). Returns the queue index of the first
// flushed ray.
template <bool STATE = false>
__device__ __forceinline__ uint32_t stage_flush(const RayStage& st, float4* __restrict__ no, float4* __restrict__ nd,
                                            uint32_t* __restrict__ count_out, float4* __restrict__ nt = nullptr,
                                            float4* __restrict__ nr = nullptr) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t n = st.meta[0];
        st.meta[2] = n;
        st.meta[1] = n ? atomicAdd(count_out, n) : 0u;
        st.meta[0] = 0u;
    }
    __syncthreads();
    const uint32_t n = st.meta[2], base = st.meta[1];
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
        st_stream(no + base + k, st.o[k]);
        st_stream(nd + base + k, st.d[k]);
        if constexpr (STATE) {
            st_stream(nt + base + k, st.t[k]);
            st_stream(nr + base + k, st.r[k]);
        }
    }
    __syncthreads();
    return base;
}

// Groups the tile's paths [tile, tile + ROUNDS * SHADE_BLOCK) by shading class (counting sort in LDS,
// stable within a class): order[k] = queue slot of the k-th path in class order; slots past
// the queue end sort last. Ends with the block synchronised.
template <int ROUNDS>
__device__ __forceinline__ void shade_sort_tile(const DeviceScene& S, const float4* __restrict__ hits, uint32_t tile, uint32_t n,
                                uint32_t* order, uint32_t* cnt, uint32_t* total) {
    constexpr uint32_t K = SHADE_CLASSES + 1;
    constexpr uint32_t WAVES = SHADE_BLOCK / 64;
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    uint32_t cls[ROUNDS], rank[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const uint32_t j = tile + uint32_t(r) * SHADE_BLOCK + threadIdx.x;
        uint32_t c = SHADE_CLASSES;
        if (j < n) {
            const uint32_t gid = reinterpret_cast<const uint32_t*>(hits + j)[3];
            // the shading record's last 16 bytes: also brings the line in for the shading pass
            c = gid == TRAVERSAL_DONE ? uint32_t(SHADE_CLASS_ENV)
                                      : __float_as_uint(S.tri_shade[4 * size_t(gid) + 3].w) >> TRI_SHADE_CLASS_SHIFT;
        }
        cls[r] = c;
        rank[r] = 0;
        for (uint32_t k = 0; k < K; k++) {
            const uint64_t b = __ballot(c == k);
            if (c == k) rank[r] = uint32_t(__popcll(b & ((1ull << lane) - 1ull)));
            if (lane == 0) cnt[(uint32_t(r) * WAVES + wave) * K + k] = uint32_t(__popcll(b));
        }
    }
    __syncthreads();
    if (threadIdx.x < K) {
        // exclusive prefix of class k over the (round, wave) groups, and the class total
        const uint32_t k = threadIdx.x;
        uint32_t run = 0;
        for (uint32_t g = 0; g < ROUNDS * WAVES; g++) {
            const uint32_t v = cnt[g * K + k];
            cnt[g * K + k] = run;
            run += v;
        }
        total[k] = run;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        uint32_t off = 0;
        for (uint32_t k = 0; k < cls[r]; k++) off += total[k];
        order[off + cnt[(uint32_t(r) * WAVES + wave) * K + cls[r]] + rank[r]] =
            tile + uint32_t(r) * SHADE_BLOCK + threadIdx.x;
    }
    __syncthreads();
}

// Continues a shaded surface vertex with a BSDF sample (trace_path, render_pt.slang:44-75):
// throughput *= weight, ior follows transmission, and ShadingPoint.make_ray (hit.slang:134-146)
// gives the next ray unless the geometric side disagrees with the transmission flag.
// Returns whether the path continues; pdf = the sample's density.
__device__ __forceinline__ bool bsdf_continue(const ShadingPoint& sp, f3 wi_local, f3 ng, bool wi_below, PCG& rng,
                                              f3& thr, float& ior, f3& next_o, f3& next_d, float& pdf) {
    const Sample s = sp.mat.sample_eval(wi_local, rng);
    if (!s.valid) return false;
    const f3 wo = sp.frame.to_world(s.wo);
    thr = thr * s.weight;
    const bool transmission = wi_local.z * s.wo.z < 0.f;
    if (transmission) ior /= sp.mat.eta;
    const float wo_ng = dot(wo, ng);
    if ((wi_below != (wo_ng < 0.f)) != transmission) return false;
    next_o = offset_ray_origin(sp.pos, wo_ng < 0.f ? -ng : ng);
    next_d = wo;
    pdf = s.pdf;
    return true;
}

// Vertices where the guide is sampled: opaque surfaces that are not near-specular (the
// mixture cannot follow a near-delta lobe; the BSDF alone samples those).
__device__ __forceinline__ bool guide_eligible(const mrt_material& m) {
    if (m.model == MRT_MAT_DIFFUSE) return true;
    if (m.transmission_weight > 0.f) return false;
    return !(m.metalness > 0.f && m.specular_alpha[0] * m.specular_alpha[1] < 1e-4f);
}

// Shades vertex `bounce` of every queued path and appends its continuation ray
// (trace_path, render_pt.slang:24-90). In MRT_MODE_NEE it also samples the environment
// (one shadow ray per vertex, power-heuristic MIS with the BSDF sample; an escaping BSDF-sampled
// ray then carries the complementary weight). In MRT_MODE_GUIDED an eligible vertex samples
// the guide's lobes with probability A.guide_fraction, else the BSDF, and weighs the direction
// by f / (a q + (1 - a) p_bsdf) (one-sample MIS); sample-0 paths of the first chunk append a
// training record per guided vertex.
// The path's state is read from the input queue's slot (qt, qr) and written with the
// continuation ray to the output queue's (nt, nr); a path that ends writes its radiance to
// P.state once (every mode this kernel runs keeps its state in the queue). A NEE shadow ray names
// where its contribution goes: its path's slot in the output queue (each round's continuation
// rays are flushed before the round's shadow rays are appended) or its path's record.
// BSDF mode fits 168 VGPRs (one spilled, no measurable cost): 3 waves per SIMD instead of 2.
template <int MODE, int ROUNDS>
__global__ __launch_bounds__(SHADE_BLOCK) __attribute__((amdgpu_waves_per_eu(MODE == MRT_MODE_BSDF ? 3 : 1))) void k_shade(DeviceScene S, PtArgs A, int bounce, PathBuffers P,
                                                       const float4* __restrict__ qo, const float4* __restrict__ qd,
                                                       const float4* __restrict__ qt, const float4* __restrict__ qr,
                                                       const uint32_t* __restrict__ count_in, float4* __restrict__ no,
                                                       float4* __restrict__ nd, float4* __restrict__ nt,
                                                       float4* __restrict__ nr, uint32_t* __restrict__ count_out,
                                                       uint32_t* __restrict__ shadow_count) {
    constexpr bool queued = state_in_queue(MODE);
    constexpr int TILE = ROUNDS * SHADE_BLOCK;
    __shared__ uint32_t s_order[TILE];
    __shared__ uint32_t s_cnt[ROUNDS * (SHADE_BLOCK / 64) * (SHADE_CLASSES + 1)];
    __shared__ uint32_t s_total[SHADE_CLASSES + 1];
    __shared__ float4 s_stage_o[TILE], s_stage_d[TILE];
    __shared__ float4 s_stage_t[queued ? TILE : 1], s_stage_r[queued ? TILE : 1];
    __shared__ uint32_t s_meta[3];
    __shared__ uint32_t s_app[2]; // block_append scratch (records, shadow rays)
    const RayStage stage{s_stage_o, s_stage_d, s_stage_t, s_stage_r, s_meta};
    if (threadIdx.x == 0) {
        s_meta[0] = 0u;
        s_app[0] = 0u;
    }
    __syncthreads();
    const uint32_t n = *count_in;
    const bool last = bounce + 1 >= A.max_path_length;
    constexpr bool nee = MODE == MRT_MODE_NEE;
    constexpr bool guided = MODE == MRT_MODE_GUIDED;
    // Continuation hits arrive in arbitrary order: group each tile by shading class so a wave
    // runs one material's code. Primary hits arrive in pixel order and are coherent already;
    // the last vertex only adds emission.
    const bool sort = bounce > 0 && !last;
    for (uint32_t tile = blockIdx.x * TILE; tile < n; tile += gridDim.x * TILE) {
        if (sort) shade_sort_tile<ROUNDS>(S, P.hits, tile, n, s_order, s_cnt, s_total);
        for (int round = 0; round < ROUNDS; round++) {
            const uint32_t slot = uint32_t(round) * SHADE_BLOCK + threadIdx.x;
            const uint32_t i = sort ? s_order[slot] : tile + slot;
            bool push = false, shadow = false, record = false;
            f3 next_o{0, 0, 0}, next_d{0, 0, 0};
            float next_pdf = 0.f; // NEE: rides in the continuation ray's direction.w
            f3 rec_pos{0, 0, 0}, rec_rad{0, 0, 0}, rd_keep{0, 0, 0}, thr_keep{0, 0, 0};
            float rec_pdf = 0.f;
            f3 sh_o{0, 0, 0}, sh_d{0, 0, 0}, sh_c{0, 0, 0};
            float4 out_t{}, out_r{};
            uint32_t pid = 0;
            if (i < n) {
                const float4 o4 = ld_stream(qo + i), d4 = ld_stream(qd + i), h4 = ld_stream(P.hits + i);
                pid = __float_as_uint(o4.w);
                const f3 ro{o4.x, o4.y, o4.z}, rd{d4.x, d4.y, d4.z};
                const uint32_t gid = __float_as_uint(h4.w);
                const float4 sa = queued ? ld_stream(qt + i) : ld_stream(P.state + 2 * size_t(pid));
                const float4 sb = queued ? ld_stream(qr + i) : ld_stream(P.state + 2 * size_t(pid) + 1);
                f3 thr{sa.x, sa.y, sa.z};
                float ior = sa.w;
                f3 rad{sb.x, sb.y, sb.z};
                PCG rng{__float_as_uint(sb.w)};
                ShadingPoint sp;
                shading_point(S, ro, rd, gid, h4.y, h4.z, ior, bounce == 0 || !last, sp);
                f3 emission = sp.emission;
                if (nee && bounce > 0 && sp.env) {
                    // the environment was also reachable by NEE from the previous vertex
                    emission = emission * power_heuristic(d4.w, env_pdf(S, rd));
                }
                if (bounce == 0) {
                    rad = A.emission_on_primary ? thr * emission : f3{0.f, 0.f, 0.f};
                } else {
                    rad = rad + thr * emission;
                    if (A.record_signature) P.sig[pid] = sig_step(P.sig[pid], gid);
                }
                if (!last && sp.has_material &&
                    !(thr.x < THROUGHPUT_THRESHOLD && thr.y < THROUGHPUT_THRESHOLD && thr.z < THROUGHPUT_THRESHOLD)) {
                    const f3 wi_local = sp.frame.to_local(-rd);
                    const f3 ng = decode_normal(sp.enc_face_normal);
                    const bool wi_below = dot(-rd, ng) < 0.f;
                    if (nee) {
                        f3 wl;
                        float pdf_l;
                        if (env_sample(S, rng, wl, pdf_l)) {
                            const f3 wl_local = sp.frame.to_local(wl);
                            const f3 f = sp.mat.eval(wi_local, wl_local);
                            const float wl_ng = dot(wl, ng);
                            const bool trans = wi_local.z * wl_local.z < 0.f;
                            // Only directions the BSDF sampler can produce (pdf > 0): the BSDF-only
                            // estimator integrates f over that support (the GGX BTDF's eval is
                            // non-zero outside it), and NEE must estimate the same integral.
                            const float pb = sp.mat.pdf(wi_local, wl_local);
                            if (pb > 0.f && (f.x > 0.f || f.y > 0.f || f.z > 0.f) &&
                                ((wi_below != (wl_ng < 0.f)) == trans)) {
                                const float w = power_heuristic(env_pdf(S, wl), pb);
                                sh_c = ((thr * f) * env_eval(S, wl)) * (w / pdf_l);
                                sh_o = offset_ray_origin(sp.pos, wl_ng < 0.f ? -ng : ng);
                                sh_d = wl;
                                shadow = finite3(sh_c) && (sh_c.x > 0.f || sh_c.y > 0.f || sh_c.z > 0.f);
                            }
                        }
                    }
                    float alpha = 0.f;
                    GuideLobes L;
                    if (guided && uint32_t(bounce) < A.guided_bounces && guide_eligible(S.mats[sp.material])) {
                        // bounce 0: queue slot i = sample * npix + pixel, lobes computed per pixel
                    unpack_lobes(P.lobes + 4 * size_t(bounce == 0 ? i % A.npix : i), L);
                        if (L.pi_sum > 0.f) alpha = A.guide_fraction;
                    }
                    if (guided && alpha > 0.f) {
                        // one-sample MIS of the guide's mixture and the BSDF
                        bool ok = true;
                        f3 wo_l;
                        if (rng.next_float() < alpha) {
                            const float u0 = rng.next_float();
                            const f2 u12 = rng.next_float2();
                            wo_l = sp.frame.to_local(lobes_sample(L, u0, u12.x, u12.y));
                        } else {
                            const Sample s = sp.mat.sample_eval(wi_local, rng);
                            ok = s.valid;
                            wo_l = s.wo;
                        }
                        if (ok) {
                            const f3 wo = sp.frame.to_world(wo_l);
                            const f3 f = sp.mat.eval(wi_local, wo_l);
                            const float pb = sp.mat.pdf(wi_local, wo_l);
                            const float p = alpha * lobes_pdf(L, wo) + (1.f - alpha) * pb;
                            const float wo_ng = dot(wo, ng);
                            const bool transmission = wi_local.z * wo_l.z < 0.f;
                            // as NEE: only directions inside the BSDF sampler's support (pb > 0),
                            // the domain the BSDF-only estimator integrates over; this also bounds
                            // the weight by f / ((1 - alpha) pb)
                            if (pb > 0.f && p > 0.f && (f.x > 0.f || f.y > 0.f || f.z > 0.f) &&
                                (wi_below != (wo_ng < 0.f)) == transmission) {
                                thr = thr * (f / p);
                                next_o = offset_ray_origin(sp.pos, wo_ng < 0.f ? -ng : ng);
                                next_d = wo;
                                push = finite3(thr);
                                record = push && A.s0 == 0 && pid < A.npix &&
                                         (A.guide_record_shift == 0u ||
                                          ((pid * 2654435761u) >> (32u - A.guide_record_shift)) == 0u);
                                rec_pos = sp.pos;
                                rec_rad = rad;
                                rec_pdf = p;
                                rd_keep = rd;
                                thr_keep = thr;
                            } else {
                                thr = f3{0.f, 0.f, 0.f};
                            }
                        } else {
                            thr = f3{0.f, 0.f, 0.f};
                        }
                    } else {
                        float pdf = 0.f;
                        push = bsdf_continue(sp, wi_local, ng, wi_below, rng, thr, ior, next_o, next_d, pdf);
                        if (nee && push) next_pdf = pdf;
                    }
                }
                out_t = make_float4(thr.x, thr.y, thr.z, ior);
                out_r = make_float4(rad.x, rad.y, rad.z, __uint_as_float(rng.state));
                if (!queued) st_stream(P.state + 2 * size_t(pid), out_t);
                if (!queued || !push) st_stream(P.state + 2 * size_t(pid) + 1, out_r);
            }
            const uint32_t staged =
                stage_push<queued>(stage, push, make_float4(next_o.x, next_o.y, next_o.z, __uint_as_float(pid)),
                                   make_float4(next_d.x, next_d.y, next_d.z, next_pdf), out_t, out_r);
            if (guided) {
                const uint32_t r = block_append(record, P.grec_count, s_app);
                if (record && r < P.grec_cap) {
                    float4* g = P.grec + 5 * size_t(r);
                    g[0] = make_float4(rec_pos.x, rec_pos.y, rec_pos.z, __uint_as_float(pid));
                    g[1] = make_float4(-rd_keep.x, -rd_keep.y, -rd_keep.z, rec_pdf);
                    g[2] = make_float4(next_d.x, next_d.y, next_d.z, __uint_as_float(uint32_t(bounce)));
                    g[3] = make_float4(rec_rad.x, rec_rad.y, rec_rad.z, 0.f);
                    g[4] = make_float4(thr_keep.x, thr_keep.y, thr_keep.z, 0.f);
                }
            }
            if (nee) {
                // this round's continuation rays are flushed first, so that a shadow ray can name
                // the queue slot its path went to
                const uint32_t base = stage_flush<queued>(stage, no, nd, count_out, nt, nr);
                const uint32_t to = push ? base + staged : (SHADOW_TO_PATH | pid);
                const uint32_t q = block_append(shadow, shadow_count, s_app);
                if (shadow) {
                    st_stream(P.sq_o + q, make_float4(sh_o.x, sh_o.y, sh_o.z, __uint_as_float(to)));
                    st_stream(P.sq_d + q, make_float4(sh_d.x, sh_d.y, sh_d.z, 0.f));
                    st_stream(P.sq_c + q, make_float4(sh_c.x, sh_c.y, sh_c.z, 0.f));
                }
            }
        }
        // (NEE flushed every round) the flush also keeps s_order intact until every round read it
        if (!nee) stage_flush<queued>(stage, no, nd, count_out, nt, nr);
    }
}

// Primary vertices of a chunk in MRT_MODE_BSDF, in place of k_pt_init and the bounce-0 shade:
// the samples of a pixel share its primary ray and gbuffer hit (render_pt.slang:99-103), so a
// thread shades the pixel's hit once and continues every sample of the chunk from it.
// Results per path are those of k_pt_init + k_shade<MRT_MODE_BSDF> at bounce 0.
__global__ __launch_bounds__(SHADE_BLOCK) __attribute__((amdgpu_waves_per_eu(2))) void k_shade_primary(
    DeviceScene S, PtArgs A, const mrt_packed_hit* __restrict__ gbuffer, PathBuffers P, float4* __restrict__ no,
    float4* __restrict__ nd, float4* __restrict__ nt, float4* __restrict__ nr, uint32_t* __restrict__ count_out) {
    __shared__ float4 s_stage_o[SHADE_TILE], s_stage_d[SHADE_TILE], s_stage_t[SHADE_TILE], s_stage_r[SHADE_TILE];
    __shared__ uint32_t s_meta[3];
    const RayStage stage{s_stage_o, s_stage_d, s_stage_t, s_stage_r, s_meta};
    if (threadIdx.x == 0) s_meta[0] = 0u;
    __syncthreads();
    const uint32_t stride = gridDim.x * blockDim.x;
    const bool last = A.max_path_length <= 1;
    for (uint32_t base = blockIdx.x * blockDim.x; base < A.npix; base += stride) {
        const uint32_t p = base + threadIdx.x;
        const bool valid = p < A.npix;
        uint32_t x = 0, y = 0, gid = TRAVERSAL_DONE;
        ShadingPoint sp;
        sp.has_material = false;
        sp.emission = f3{0.f, 0.f, 0.f};
        f3 wi_local{0.f, 0.f, 1.f}, ng{0.f, 0.f, 1.f};
        bool wi_below = false;
        if (valid) {
            x = p % A.width;
            y = p / A.width;
            const RayF r = camera_ray(A.cam, x, y, A.width, A.height);
            float bu, bv;
            gid = gbuffer_hit(S, gbuffer, p, bu, bv);
            shading_point(S, r.o, r.d, gid, bu, bv, 1.f, true, sp);
            wi_local = sp.frame.to_local(-r.d);
            ng = decode_normal(sp.enc_face_normal);
            wi_below = dot(-r.d, ng) < 0.f;
        }
        const f3 one{1.f, 1.f, 1.f};
        const f3 rad = A.emission_on_primary ? one * sp.emission : f3{0.f, 0.f, 0.f};
        const bool cont = valid && !last && sp.has_material;
        for (uint32_t sl = 0; sl < A.chunk; sl++) {
            const uint32_t pid = sl * A.npix + p;
            bool push = false;
            f3 next_o{0.f, 0.f, 0.f}, next_d{0.f, 0.f, 0.f};
            float4 out_t{}, out_r{};
            if (valid) {
                const uint32_t s = A.s0 + sl;
                PCG rng{pcg4d16(x, y, A.frame, s)};
                f3 thr = one;
                float ior = 1.f, pdf = 0.f;
                if (cont) push = bsdf_continue(sp, wi_local, ng, wi_below, rng, thr, ior, next_o, next_d, pdf);
                // the state travels with the ray; a path that ends here leaves its radiance
                out_t = make_float4(thr.x, thr.y, thr.z, ior);
                out_r = make_float4(rad.x, rad.y, rad.z, __uint_as_float(rng.state));
                if (!push) st_stream(P.state + 2 * size_t(pid) + 1, out_r);
                if (A.record_signature) P.sig[pid] = sig_step(0x9E3779B9u * (s + 1u), gid);
            }
            stage_push<true>(stage, push, make_float4(next_o.x, next_o.y, next_o.z, __uint_as_float(pid)),
                             make_float4(next_d.x, next_d.y, next_d.z, 0.f), out_t, out_r);
            // the block stages up to SHADE_BLOCK rays per sample: flush every SHADE_ROUNDS samples
            if ((sl + 1) % SHADE_ROUNDS == 0 || sl + 1 == A.chunk) stage_flush<true>(stage, no, nd, count_out, nt, nr);
        }
    }
}

// GBuffer albedo per pixel: the BSDF's directional albedo towards the camera in the material's
// shading frame for a surface, the environment's emission for a miss (gbuffer_rt.slang:76-77,92);
// floored at 1e-2 per channel for demodulation (gbuffer.slang:135-140).
__global__ void k_gbuffer_albedo(DeviceScene S, mrt_camera cam, uint32_t w, uint32_t h,
                                 const mrt_packed_hit* __restrict__ gbuffer, float4* __restrict__ out, int floor) {
    const uint32_t n = w * h;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        const uint32_t x = p % w, y = p / w;
        const RayF r = camera_ray(cam, x, y, w, h);
        float bu, bv;
        const uint32_t gid = gbuffer_hit(S, gbuffer, p, bu, bv);
        ShadingPoint sp;
        shading_point(S, r.o, r.d, gid, bu, bv, 1.f, true, sp);
        const f3 a = sp.has_material ? sp.mat.get_albedo(sp.frame.to_local(-r.d)) : sp.emission;
        out[p] = floor ? make_float4(fmaxf(a.x, 1e-2f), fmaxf(a.y, 1e-2f), fmaxf(a.z, 1e-2f), 1.f)
                       : make_float4(a.x, a.y, a.z, 1.f);
    }
}

// -----------------------------------------------------------------------------------------
// Markov-chain path guiding (render_pt_mcpg.slang:39-244)

// Seeds pass A.s0 of an MCPG frame, one path per pixel: the samples of a pixel run one after
// another on one RNG stream (render_pt_mcpg.slang:224-244), so the first pass seeds it with
// pcg3d16(pixel, frame) and later passes continue from where the previous sample's path left
// it. Path state as k_pt_init.
__global__ void k_mcpg_init(DeviceScene S, PtArgs A, const mrt_packed_hit* __restrict__ gbuffer, PathBuffers P,
                            uint32_t* count0) {
    const uint32_t stride = gridDim.x * blockDim.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) *count0 = A.npix;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < A.npix; p += stride) {
        const uint32_t x = p % A.width, y = p / A.width;
        const RayF r = camera_ray(A.cam, x, y, A.width, A.height);
        float bu, bv;
        const uint32_t gid = gbuffer_hit(S, gbuffer, p, bu, bv);
        const uint32_t rng = A.s0 == 0 ? pcg3d16(x, y, A.frame) : __float_as_uint(P.state[2 * size_t(p) + 1].w);
        st_stream(P.q_o[0] + p, make_float4(r.o.x, r.o.y, r.o.z, __uint_as_float(p)));
        st_stream(P.q_d[0] + p, make_float4(r.d.x, r.d.y, r.d.z, 0.f));
        st_stream(P.hits + p, make_float4(0.f, bu, bv, __uint_as_float(gid)));
        st_stream(P.state + 2 * size_t(p), make_float4(1.f, 1.f, 1.f, 1.f));
        st_stream(P.state + 2 * size_t(p) + 1, make_float4(0.f, 0.f, 0.f, __uint_as_float(rng)));
    }
}

// What a scatter vertex's chain update needs once its continuation ray has been traced: the
// lobe's chain state (the RIS-selected one, or a fresh state when the BSDF was sampled) and
// the slot it came from, the RIS score sum, the vertex position and normal, the light-cache
// cell its own lookup found, the BSDF value and the sampling pdf.
struct McpgCarry {
    McState sel;
    uint32_t sel_slot;
    float score_sum;
    f3 pos;
    float wo_p;
    f3 normal;
    GridPos lc;
    f3 bsdf;
};
__device__ __forceinline__ void carry_store(float4* __restrict__ c, uint32_t pid, const McpgCarry& v) {
    float4* p = c + 5 * size_t(pid);
    p[0] = make_float4(v.sel.w_tgt.x, v.sel.w_tgt.y, v.sel.w_tgt.z, v.sel.sum_w);
    p[1] = make_float4(v.sel.w_cos, __uint_as_float(v.sel.mv_z_n), __uint_as_float(v.sel_slot), v.score_sum);
    p[2] = make_float4(v.pos.x, v.pos.y, v.pos.z, v.wo_p);
    p[3] = make_float4(v.normal.x, v.normal.y, v.normal.z, __uint_as_float(v.lc.slot));
    p[4] = make_float4(v.bsdf.x, v.bsdf.y, v.bsdf.z, __uint_as_float(v.lc.key));
}
__device__ __forceinline__ McpgCarry carry_load(const float4* __restrict__ c, uint32_t pid) {
    const float4* p = c + 5 * size_t(pid);
    const float4 a = p[0], b = p[1], d = p[2], e = p[3], f = p[4];
    McpgCarry v;
    // mv.xy and T are not carried: the update overwrites them before a store
    v.sel = McState{{a.x, a.y, a.z}, a.w, b.x, 0u, __float_as_uint(b.y), 0.f};
    v.sel_slot = __float_as_uint(b.z);
    v.score_sum = b.w;
    v.pos = {d.x, d.y, d.z};
    v.wo_p = d.w;
    v.normal = {e.x, e.y, e.z};
    v.lc = GridPos{__float_as_uint(e.w), __float_as_uint(f.w)};
    v.bsdf = {f.x, f.y, f.z};
    return v;
}

// The end of a scatter vertex's loop iteration (render_pt_mcpg.slang:161-178): the light
// cache learns the incident radiance over the sampling pdf (floored at lc_min_pdf); with
// probability mc_f * mc_samples / score_sum the chain takes the sample and is stored at the
// vertex, else the missing-light heuristic may kill the cell the lobe came from.
__device__ __forceinline__ void mcpg_complete(const McpgArgs& M, McpgCarry& c, f3 lc_incident, f3 next_pos, f3 wo,
                                              f3 cam, f3 prev_cam, uint32_t frame, PCG& rng) {
    const f3 lb = lc_incident * c.bsdf;
    const float mc_f = yuv_luminance(lb / c.wo_p);
    if (!isfinite(mc_f)) return;
    lc_update(M.lc, c.pos, c.normal, lb / fmaxf(c.wo_p, M.lc_min_pdf), cam, prev_cam, frame, rng, c.lc);
    if (rng.next_float() * c.score_sum < mc_f * float(M.samples)) {
        c.sel.add_sample(c.pos, mc_f, next_pos, 0u, 0u, M.time); // static scene: zero motion
        mc_store(M.mc, c.sel, c.pos, c.normal, cam, frame, rng);
    } else if (M.missing_light && c.sel_slot != GRID_INVALID_SLOT && c.sel.light_missing(mc_f, wo, c.pos, M.prior)) {
        mc_kill(M.mc, c.sel_slot);
    }
}

// Shades vertex `bounce` of every queued MCPG path (trace_path, render_pt_mcpg.slang:39-182).
// A continuation hit first completes the vertex it was scattered from: incident radiance
// (the hit's emission, else the light cache there), throughput and radiance, then that
// vertex's cache and chain updates. A vertex that scatters picks a lobe by RIS over
// M.samples chain lookups, samples it with probability p_guide (else the BSDF), weighs the
// direction by the stochastic-MIS pdf and carries its update inputs to the next bounce.
// Two waves per SIMD, without spills: with the path-state columns in the kernel's arguments, three
// waves spill 15 registers and the 1080p MCPG frame takes 533 ms instead of 529 (an earlier
// version, with fewer spills at three waves, measured the opposite: 11 % off the kernel).
__global__ __launch_bounds__(SHADE_BLOCK) __attribute__((amdgpu_waves_per_eu(2))) void k_shade_mcpg(
                                                            DeviceScene S, PtArgs A, McpgArgs M, int bounce,
                                                            PathBuffers P, const float4* __restrict__ qo,
                                                            const float4* __restrict__ qd,
                                                            const uint32_t* __restrict__ count_in,
                                                            float4* __restrict__ no, float4* __restrict__ nd,
                                                            uint32_t* __restrict__ count_out) {
    __shared__ float4 s_stage_o[SHADE_TILE], s_stage_d[SHADE_TILE];
    __shared__ uint32_t s_meta[3];
    const RayStage stage{s_stage_o, s_stage_d, nullptr, nullptr, s_meta};
    if (threadIdx.x == 0) s_meta[0] = 0u;
    __syncthreads();
    const uint32_t n = *count_in;
    const bool last = bounce + 1 >= A.max_path_length;
    const f3 cam{M.cam[0], M.cam[1], M.cam[2]}, prev_cam{M.prev_cam[0], M.prev_cam[1], M.prev_cam[2]};
    for (uint32_t tile = blockIdx.x * SHADE_TILE; tile < n; tile += gridDim.x * SHADE_TILE) {
        for (int round = 0; round < SHADE_ROUNDS; round++) {
            const uint32_t i = tile + uint32_t(round) * SHADE_BLOCK + threadIdx.x;
            bool push = false;
            f3 next_o{0, 0, 0}, next_d{0, 0, 0};
            uint32_t pid = 0;
            if (i < n) {
                const float4 o4 = ld_stream(qo + i), d4 = ld_stream(qd + i), h4 = ld_stream(P.hits + i);
                pid = __float_as_uint(o4.w);
                const f3 ro{o4.x, o4.y, o4.z}, rd{d4.x, d4.y, d4.z};
                const uint32_t gid = __float_as_uint(h4.w);
                const float4 sa = ld_stream(P.state + 2 * size_t(pid)), sb = ld_stream(P.state + 2 * size_t(pid) + 1);
                f3 thr{sa.x, sa.y, sa.z};
                float ior = sa.w;
                f3 rad{sb.x, sb.y, sb.z};
                PCG rng{__float_as_uint(sb.w)};
                ShadingPoint sp;
                shading_point(S, ro, rd, gid, h4.y, h4.z, ior, bounce == 0 || !last, sp);
                GridPos lc_here = invalid_grid_pos();
                if (bounce == 0) {
                    rad = A.emission_on_primary ? thr * sp.emission : f3{0.f, 0.f, 0.f};
                } else {
                    McpgCarry c = carry_load(M.carry, pid);
                    const f3 incident = sp.emission;
                    // evaluated whether or not its value is used, as the reference's select()
                    const f3 cached = lc_get(M.lc, sp.pos, sp.frame.r2, cam, prev_cam, rng, lc_here);
                    const bool direct = incident.x > 0.f || incident.y > 0.f || incident.z > 0.f ||
                                        (!M.lc_tail && A.max_path_length == 2);
                    const f3 lc_incident = direct ? incident : cached;
                    thr = thr * (c.bsdf / c.wo_p);
                    rad = rad + thr * ((M.lc_tail && bounce == A.max_path_length - 1) ? lc_incident : incident);
                    mcpg_complete(M, c, lc_incident, sp.pos, rd, cam, prev_cam, A.frame, rng);
                }
                if (!last && sp.has_material &&
                    !(thr.x < THROUGHPUT_THRESHOLD && thr.y < THROUGHPUT_THRESHOLD && thr.z < THROUGHPUT_THRESHOLD)) {
                    const f3 cur = sp.pos, normal = sp.frame.r2;
                    const f3 wi_local = sp.frame.to_local(-rd);
                    // resample one lobe from the chains (RIS); slot 0 holds the selected one
                    McState sel = mc_zero();
                    uint32_t sel_slot = GRID_INVALID_SLOT;
                    float score_sum = 0.f;
                    float scores[MCPG_MAX_SAMPLES];
                    Vmf vmfs[MCPG_MAX_SAMPLES];
#pragma unroll
                    for (int k = 0; k < MCPG_MAX_SAMPLES; k++) {
                        scores[k] = 0.f;
                        vmfs[k] = Vmf{0.f, {0.f, 0.f, 0.f}};
                    }
#pragma unroll
                    for (int k = 0; k < MCPG_MAX_SAMPLES; k++) {
                        if (k < M.samples) {
                            GridPos gp;
                            const McState s = mc_load(M.mc, cur, normal, cam, M.time, rng, gp);
                            score_sum += s.sum_w;
                            if (rng.next_float() < s.sum_w / score_sum) {
                                sel = s;
                                sel_slot = gp.slot;
                                vmfs[k] = vmfs[0];
                                scores[k] = scores[0];
                                scores[0] = s.sum_w;
                                vmfs[0] = mc_vmf(s, cur, M.prior);
                            } else {
                                scores[k] = s.sum_w;
                                vmfs[k] = mc_vmf(s, cur, M.prior);
                            }
                        }
                    }
                    const float bsdf_alpha = sp.mat.alpha();
                    const bool can_guide = score_sum > 0.f && bsdf_alpha >= M.alpha_threshold;
                    const float p_guide = M.roughness ? M.p_guiding * bsdf_alpha : M.p_guiding;
                    const bool guide = can_guide && rng.next_float() < p_guide;
                    f3 wo_world{0, 0, 0}, wo_local{0, 0, 0};
                    bool ok = true;
                    if (guide) {
                        wo_world = vmfs[0].sample(rng);
                        wo_local = sp.frame.to_local(wo_world);
                    } else {
                        const Sample s = sp.mat.sample_eval(wi_local, rng); // BSDF.sample: same draws
                        ok = s.valid;
                        if (ok) {
                            wo_local = s.wo;
                            wo_world = sp.frame.to_world(s.wo);
                            sel = mc_zero();
                            sel_slot = GRID_INVALID_SLOT;
                            (void)grid_target_level(M.mc, cur, cam, rng); // the reference's unused level draw
                        }
                    }
                    if (ok) {
                        float mix = 0.f;
                        if (can_guide) {
#pragma unroll
                            for (int k = 0; k < MCPG_MAX_SAMPLES; k++)
                                if (k < M.samples) mix += scores[k] * vmfs[k].pdf(wo_world);
                            mix /= score_sum;
                        }
                        const float bsdf_w = can_guide ? (1.0f - p_guide) : 1.0f;
                        const float pb = sp.mat.pdf(wi_local, wo_local);
                        const float wo_p = bsdf_w * pb + (can_guide ? p_guide : 0.f) * mix;
                        const f3 bsdf = M.bsdf_support_only && !(pb > 0.f) ? f3{0.f, 0.f, 0.f}
                                                                           : sp.mat.eval(wi_local, wo_local);
                        const bool transmission = wi_local.z * wo_local.z < 0.f;
                        bool has_next = false;
                        if (bsdf.x > 0.f || bsdf.y > 0.f || bsdf.z > 0.f) { // ShadingPoint.make_ray
                            const f3 ng = decode_normal(sp.enc_face_normal);
                            const float wo_ng = dot(wo_world, ng);
                            if (((dot(-rd, ng) < 0.f) != (wo_ng < 0.f)) == transmission) {
                                has_next = true;
                                next_o = offset_ray_origin(sp.pos, wo_ng < 0.f ? -ng : ng);
                                next_d = wo_world;
                            }
                        }
                        McpgCarry c{sel, sel_slot, score_sum, cur, wo_p, normal, lc_here, bsdf};
                        if (has_next) {
                            if (transmission) ior /= sp.mat.eta;
                            carry_store(M.carry, pid, c);
                            push = true;
                        } else {
                            mcpg_complete(M, c, f3{0.f, 0.f, 0.f}, cur, wo_world, cam, prev_cam, A.frame, rng);
                        }
                    }
                }
                st_stream(P.state + 2 * size_t(pid), make_float4(thr.x, thr.y, thr.z, ior));
                st_stream(P.state + 2 * size_t(pid) + 1, make_float4(rad.x, rad.y, rad.z, __uint_as_float(rng.state)));
            }
            stage_push(stage, push, make_float4(next_o.x, next_o.y, next_o.z, __uint_as_float(pid)),
                       make_float4(next_d.x, next_d.y, next_d.z, 0.f));
        }
        stage_flush(stage, no, nd, count_out);
    }
}

// Training records of the guide from the first chunk's sample-0 paths, once their radiance
// is final: incident radiance along the sampled direction = (final radiance - radiance before
// the vertex's continuation) / throughput after it; weight = its luminance / sampling pdf.
// Record r keeps its arrival slot; training reorders the records by key first.
__global__ void k_guide_targets(PathBuffers P) {
    const uint32_t n = min(*P.grec_count, P.grec_cap);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        const float4* g = P.grec + 5 * size_t(r);
        const float4 a = g[0], b = g[1], c = g[2], rb = g[3], th = g[4];
        const float4 rf = P.state[2 * size_t(__float_as_uint(a.w)) + 1];
        const f3 L{th.x > 0.f ? fmaxf(rf.x - rb.x, 0.f) / th.x : 0.f, th.y > 0.f ? fmaxf(rf.y - rb.y, 0.f) / th.y : 0.f,
                   th.z > 0.f ? fmaxf(rf.z - rb.z, 0.f) / th.z : 0.f};
        float w = b.w > 0.f ? yuv_luminance(L) / b.w : 0.f;
        if (!(w >= 0.f && w < 3.0e38f)) w = 0.f;
        // in arrival order; the key (path, bounce) is unique per record (paths below 2^26, at most
        // 64 vertices) and fixes the training order (launch_guide_order_records)
        float4* t = P.gtrain + 3 * size_t(r);
        t[0] = make_float4(a.x, a.y, a.z, w);
        t[1] = make_float4(b.x, b.y, b.z, 0.f);
        t[2] = make_float4(c.x, c.y, c.z, __uint_as_float(__float_as_uint(a.w) * 64u + __float_as_uint(c.w)));
    }
}

// Sums the chunk's samples per pixel in sample order with the NaN/Inf rejection of
// render_pt.slang:116-127; the last chunk divides by spp and writes the output.
__global__ void k_accumulate(PtArgs A, PathBuffers P, int first_chunk, int last_chunk, float4* __restrict__ out,
                             uint32_t* __restrict__ sig_out) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < A.npix; p += stride) {
        float4 acc = first_chunk ? make_float4(0.f, 0.f, 0.f, 0.f) : P.accum[p];
        uint32_t sa = first_chunk ? 0u : P.sig_accum[p];
        f3 irr{acc.x, acc.y, acc.z};
        float m2 = acc.w;
        // render_pt.slang:106-108,115: every sample divided by the floored GBuffer albedo
        const float4 dm = A.demodulate ? P.demod[p] : make_float4(1.f, 1.f, 1.f, 1.f);
        for (uint32_t sl = 0; sl < A.chunk; sl++) {
            const uint32_t i = sl * A.npix + p;
            const float4 b = ld_stream(P.state + 2 * size_t(i) + 1);
            const f3 c = A.demodulate ? f3{b.x / dm.x, b.y / dm.y, b.z / dm.z} : f3{b.x, b.y, b.z};
            if (A.record_signature) sa += P.sig[i];
            if (finite3(c)) {
                irr = irr + c;
                const float l = yuv_luminance(c);
                m2 += l * l;
            }
        }
        if (last_chunk) {
            if (A.spp > 0) {
                irr = irr / float(A.spp);
                m2 /= float(A.spp);
            }
            out[p] = make_float4(irr.x, irr.y, irr.z, m2);
            if (sig_out) sig_out[p] = sa;
        } else {
            P.accum[p] = make_float4(irr.x, irr.y, irr.z, m2);
            P.sig_accum[p] = sa;
        }
    }
}

// Running mean over frames: acc = count == 0 ? src : acc + (src - acc) / (count + 1).
__global__ void k_running_mean(float4* __restrict__ acc, const float4* __restrict__ src, uint32_t n, uint32_t count) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const float w = 1.f / float(count + 1u);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float4 s = src[i];
        if (count == 0u) {
            acc[i] = s;
        } else {
            const float4 a = acc[i];
            acc[i] = make_float4(a.x + (s.x - a.x) * w, a.y + (s.y - a.y) * w, a.z + (s.z - a.z) * w,
                                 a.w + (s.w - a.w) * w);
        }
    }
}

int grid_for(uint64_t n, int block, int max_blocks) {
    const uint64_t b = (n + block - 1) / block;
    return int(b < uint64_t(max_blocks) ? (b > 0 ? b : 1) : max_blocks);
}

constexpr int MAX_BLOCKS = 256 * 16; // 256 CUs

// Blocks of k_trace<...> that fit on the device at once (the persistent grid).
template <bool COUNT, class Src, class Sink, bool ANY, int W> int trace_resident_blocks() {
    // computed once per process (thread-safe static initialisation)
    static const int blocks = [] {
        int per_cu = 0, dev = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<COUNT, Src, Sink, ANY, W>, TRACE_BLOCK, 0) !=
                hipSuccess ||
            hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 256; // query failed: one block per CU of an MI355X still drains every queue
        return per_cu * cus > 0 ? per_cu * cus : 1;
    }();
    return blocks;
}

// Launches the persistent trace kernel over at most n_max rays (the actual count is *count_ptr
// when given). `work` must be zero on entry.
template <bool COUNT, class Src, class Sink, bool ANY = false>
hipError_t launch_trace(const DeviceScene& s, const TraceScratch& ts, const Src& src, const Sink& sink,
                        const uint32_t* count_ptr, uint32_t n_max, uint32_t* work, uint32_t mask,
                        unsigned long long* counters, hipStream_t st) {
    if (n_max == 0) return hipSuccess;
    if (s.stack_need > TRACE_LDS_STACK + ts.entries) return hipErrorInvalidValue; // overflow area too small
    if (s.width == 8 ? !s.nodes8 : !s.nodes) return hipErrorInvalidValue;
    const int w = s.width == 8 ? (s.bvh8_order ? 9 : 8) : 4;
    const int resident = w == 9   ? trace_resident_blocks<COUNT, Src, Sink, ANY, 9>()
                         : w == 8 ? trace_resident_blocks<COUNT, Src, Sink, ANY, 8>()
                                  : trace_resident_blocks<COUNT, Src, Sink, ANY, 4>();
    const int cap = std::max(1, int(int64_t(resident) * int64_t(ts.grid_pct) / 100));
    int g = grid_for(n_max, TRACE_BLOCK, cap);
    if (uint64_t(g) * TRACE_BLOCK > ts.threads) g = int(ts.threads / TRACE_BLOCK);
    if (g <= 0) return hipErrorInvalidValue;
    if (w == 9) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_trace<COUNT, Src, Sink, ANY, 9>), dim3(g), dim3(TRACE_BLOCK), 0, st, s,
                           src, sink, count_ptr, n_max, work, mask, ts, counters);
    } else if (w == 8) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_trace<COUNT, Src, Sink, ANY, 8>), dim3(g), dim3(TRACE_BLOCK), 0, st, s,
                           src, sink, count_ptr, n_max, work, mask, ts, counters);
    } else {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_trace<COUNT, Src, Sink, ANY, 4>), dim3(g), dim3(TRACE_BLOCK), 0, st, s,
                           src, sink, count_ptr, n_max, work, mask, ts, counters);
    }
    return hipGetLastError();
}

} // namespace

// ---------------------------------------------------------------------------------------------

void trace_prepare() {
    // evaluates the cached occupancy queries of every trace-kernel variant the frame launches,
    // so that no device query runs while a stream is being captured
    (void)trace_resident_blocks<false, SrcCamera, SinkGBuffer, false, 4>();
    (void)trace_resident_blocks<false, SrcQueue, SinkQueue, false, 4>();
    (void)trace_resident_blocks<false, SrcShadow, SinkShadow, true, 4>();
    (void)trace_resident_blocks<false, SrcCamera, SinkGBuffer, false, 8>();
    (void)trace_resident_blocks<false, SrcQueue, SinkQueue, false, 8>();
    (void)trace_resident_blocks<false, SrcShadow, SinkShadow, true, 8>();
    (void)trace_resident_blocks<false, SrcCamera, SinkGBuffer, false, 9>();
    (void)trace_resident_blocks<false, SrcQueue, SinkQueue, false, 9>();
    (void)trace_resident_blocks<false, SrcShadow, SinkShadow, true, 9>();
}

uint32_t trace_max_threads() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    // at most 32 waves of 64 lanes per CU
    return uint32_t(cus > 0 ? cus : 256) * 32u * 64u;
}

hipError_t launch_trace_rays(const DeviceScene& s, const TraceScratch& ts, const mrt_ray* rays, uint64_t n,
                             mrt_hit* hits, hipStream_t st, unsigned long long* counters) {
    if (n == 0) return hipSuccess;
    if (n > 0xFFFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, ts.work, 1u);
    if (counters)
        return launch_trace<true>(s, ts, SrcRays{rays}, SinkHits{hits}, nullptr, uint32_t(n), ts.work, 0xFFu, counters,
                                  st);
    return launch_trace<false>(s, ts, SrcRays{rays}, SinkHits{hits}, nullptr, uint32_t(n), ts.work, 0xFFu, nullptr, st);
}

hipError_t launch_running_mean(float4* acc, const float4* src, uint32_t n, uint32_t count, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_running_mean, dim3(grid_for(n, 256, MAX_BLOCKS)), dim3(256), 0, st, acc, src, n, count);
    return hipGetLastError();
}

hipError_t launch_camera_rays(const mrt_camera& cam, uint32_t w, uint32_t h, mrt_ray* rays, hipStream_t st) {
    const uint64_t n = uint64_t(w) * h;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_camera_rays, dim3(grid_for(n, 256, MAX_BLOCKS)), dim3(256), 0, st, cam, w, h, rays);
    return hipGetLastError();
}

hipError_t launch_gbuffer(const DeviceScene& s, const TraceScratch& ts, const mrt_camera& cam, uint32_t w, uint32_t h,
                          uint32_t mask, mrt_packed_hit* out, hipStream_t st) {
    const uint64_t n = uint64_t(w) * h;
    if (n == 0) return hipSuccess;
    if (n > 0xFFFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, ts.work, 1u);
    return launch_trace<false>(s, ts, SrcCamera{cam, w, h}, SinkGBuffer{s.tri_verts, s.geom_first, out}, nullptr,
                               uint32_t(n), ts.work, mask, nullptr, st);
}

hipError_t launch_gbuffer_albedo(const DeviceScene& s, const mrt_camera& cam, uint32_t w, uint32_t h,
                                 const mrt_packed_hit* gbuffer, float4* out, int floor, hipStream_t st) {
    const uint64_t n = uint64_t(w) * h;
    if (n == 0) return hipSuccess;
    if (n > 0xFFFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gbuffer_albedo, dim3(grid_for(n, 256, MAX_BLOCKS)), dim3(256), 0, st, s, cam, w, h, gbuffer,
                       out, floor);
    return hipGetLastError();
}

hipError_t launch_pt_frame(const DeviceScene& s, const TraceScratch& ts, PtArgs a, const mrt_packed_hit* gbuffer,
                           PathBuffers& pb, float4* out, uint32_t* sig_out, hipStream_t st, FrameTiming* timing,
                           uint64_t* continuation_segments, unsigned long long* trav_counters, const GuideNet* guide,
                           const McpgArgs* mcpg) {
    const int L = a.max_path_length;
    const bool guided = a.mode == MRT_MODE_GUIDED;
    if (guided && (!guide || !pb.lobes || !pb.grec || !pb.gtrain || !pb.grec_count)) return hipErrorInvalidValue;
    if (state_in_queue(a.mode) && (!pb.q_t[0] || !pb.q_t[1] || !pb.q_r[0] || !pb.q_r[1])) return hipErrorInvalidValue;
    const bool mc = a.mode == MRT_MODE_MCPG;
    if (mc && (!mcpg || !mcpg->carry || mcpg->samples < 0 || mcpg->samples > MCPG_MAX_SAMPLES))
        return hipErrorInvalidValue;
    if (guided) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, pb.grec_count, 1u);
    const uint32_t spp = uint32_t(a.spp > 0 ? a.spp : 0);
    if (spp == 0 || a.npix == 0) {
        // render_pt.slang:124-129 with spp = 0: irradiance stays 0 (fill kernels, no memset
        // nodes, so the frame can be captured in a graph)
        if (a.npix) {
            hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(4ull * a.npix, 256, MAX_BLOCKS)), dim3(256), 0, st,
                               reinterpret_cast<uint32_t*>(out), 4u * a.npix);
            if (sig_out)
                hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(a.npix, 256, MAX_BLOCKS)), dim3(256), 0, st, sig_out,
                                   a.npix);
        }
        return hipGetLastError();
    }
    // MCPG: the samples of a pixel follow each other on one RNG stream, one pass per sample
    const uint32_t chunk = mc ? 1u : (uint32_t(pb.capacity / a.npix) < spp ? uint32_t(pb.capacity / a.npix) : spp);
    if (mc && pb.capacity < a.npix) return hipErrorInvalidValue;
    if (chunk == 0) return hipErrorInvalidValue;
    if (a.mode == MRT_MODE_NEE && uint64_t(a.npix) * chunk > NEE_MAX_PATHS) return hipErrorInvalidValue;
    const uint32_t n_chunks = (spp + chunk - 1) / chunk;
    const uint32_t per_chunk = uint32_t(L + 1);
    // timing: one event before and after every launch, read once the frame has completed
    std::vector<hipEvent_t> ev;
    std::vector<int> kind; // per event pair: 0 trace, 1 shade
    hipError_t terr = hipSuccess; // first timing failure: timing stops, the frame still renders
    auto mark = [&](int k) {
        if (!timing || terr != hipSuccess) return;
        hipEvent_t e;
        terr = hipEventCreate(&e);
        if (terr != hipSuccess) return;
        ev.push_back(e);
        terr = hipEventRecord(e, st);
        if (k >= 0) kind.push_back(k);
    };
    mark(-1);
    if (a.demodulate) {
        if (!pb.demod) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_gbuffer_albedo, dim3(grid_for(a.npix, 256, MAX_BLOCKS)), dim3(256), 0, st, s, a.cam,
                           a.width, a.height, gbuffer, pb.demod, 1);
    }
    // [chunk][bounce] arrays: queue sizes, trace cursors, shadow-queue sizes, shadow cursors
    const uint32_t nz = n_chunks * per_chunk;
    hipLaunchKernelGGL(k_zero_u32, dim3((4 * nz + 255) / 256), dim3(256), 0, st, pb.counters, 4 * nz);
    const bool nee = a.mode == MRT_MODE_NEE;
    hipError_t err = hipSuccess;
    for (uint32_t c = 0; c < n_chunks && err == hipSuccess; c++) {
        a.s0 = c * chunk;
        a.chunk = (a.s0 + chunk <= spp) ? chunk : spp - a.s0;
        a.n_paths = a.npix * a.chunk;
        uint32_t* cnt = pb.counters + c * per_chunk;
        uint32_t* work = pb.counters + nz + c * per_chunk;
        uint32_t* scnt = pb.counters + 2 * nz + c * per_chunk;
        uint32_t* swork = pb.counters + 3 * nz + c * per_chunk;
        // BSDF mode shades the primary vertices per pixel (k_shade_primary); the other modes
        // seed the path queue and shade bounce 0 like every other bounce
        const bool fused_primary = a.mode == MRT_MODE_BSDF;
        if (mc) {
            hipLaunchKernelGGL(k_mcpg_init, dim3(grid_for(a.npix, 256, MAX_BLOCKS)), dim3(256), 0, st, s, a, gbuffer, pb,
                               cnt);
        } else if (fused_primary) {
            mark(-1);
            hipLaunchKernelGGL(k_shade_primary, dim3(grid_for(a.npix, SHADE_BLOCK, MAX_BLOCKS)), dim3(SHADE_BLOCK), 0,
                               st, s, a, gbuffer, pb, pb.q_o[1], pb.q_d[1], pb.q_t[1], pb.q_r[1], cnt + 1);
            mark(1);
        } else {
            hipLaunchKernelGGL(k_pt_init, dim3(grid_for(a.n_paths, 256, MAX_BLOCKS)), dim3(256), 0, st, s, a, gbuffer,
                               pb, cnt);
        }
        for (int b = fused_primary ? 1 : 0; b < L && err == hipSuccess; b++) {
            const int qi = b & 1;
            if (b > 0) {
                mark(-1);
                const SrcQueue src{pb.q_o[qi], pb.q_d[qi]};
                const SinkQueue sink{pb.hits};
                err = trav_counters ? launch_trace<true>(s, ts, src, sink, cnt + b, a.n_paths, work + b,
                                                         a.instance_mask, trav_counters, st)
                                    : launch_trace<false>(s, ts, src, sink, cnt + b, a.n_paths, work + b,
                                                          a.instance_mask, nullptr, st);
                mark(0);
            }
            if (guided && b + 1 < L && uint32_t(b) < a.guided_bounces && err == hipSuccess) {
                mark(-1);
                // the samples of a pixel share the primary ray and hit, so bounce 0 queries once
                // per pixel (the pass's sample-0 entries, queue slots [0, npix))
                err = b == 0 ? launch_guide_infer_queue(*guide, s, pb.q_o[qi], pb.q_d[qi], pb.hits, nullptr, a.npix,
                                                        pb.lobes, st)
                             : launch_guide_infer_queue(*guide, s, pb.q_o[qi], pb.q_d[qi], pb.hits, cnt + b,
                                                        a.n_paths, pb.lobes, st);
                mark(1);
            }
            mark(-1);
            const bool wide = a.n_paths >= SHADE_WIDE_PATHS;
            const dim3 sg(grid_for(a.n_paths, SHADE_TILE, MAX_BLOCKS));
            const dim3 sgw(grid_for(a.n_paths, SHADE_WIDE_ROUNDS * SHADE_BLOCK, MAX_BLOCKS));
            if (mc) {
                hipLaunchKernelGGL(k_shade_mcpg, sg, dim3(SHADE_BLOCK), 0, st, s, a, *mcpg, b, pb, pb.q_o[qi],
                                   pb.q_d[qi], cnt + b, pb.q_o[qi ^ 1], pb.q_d[qi ^ 1], cnt + b + 1);
            } else if (nee) {
                if (wide)
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_shade<MRT_MODE_NEE, SHADE_WIDE_ROUNDS>), sgw, dim3(SHADE_BLOCK), 0, st, s, a,
                                       b, pb, pb.q_o[qi], pb.q_d[qi], pb.q_t[qi], pb.q_r[qi], cnt + b, pb.q_o[qi ^ 1],
                                       pb.q_d[qi ^ 1], pb.q_t[qi ^ 1], pb.q_r[qi ^ 1], cnt + b + 1, scnt + b);
                else
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_shade<MRT_MODE_NEE, SHADE_ROUNDS>), sg, dim3(SHADE_BLOCK), 0, st, s, a, b,
                                       pb, pb.q_o[qi], pb.q_d[qi], pb.q_t[qi], pb.q_r[qi], cnt + b, pb.q_o[qi ^ 1],
                                       pb.q_d[qi ^ 1], pb.q_t[qi ^ 1], pb.q_r[qi ^ 1], cnt + b + 1, scnt + b);
            } else if (guided) {
                if (wide)
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_shade<MRT_MODE_GUIDED, SHADE_WIDE_ROUNDS>), sgw, dim3(SHADE_BLOCK), 0, st, s, a,
                                       b, pb, pb.q_o[qi], pb.q_d[qi], pb.q_t[qi], pb.q_r[qi], cnt + b, pb.q_o[qi ^ 1],
                                       pb.q_d[qi ^ 1], pb.q_t[qi ^ 1], pb.q_r[qi ^ 1], cnt + b + 1, scnt + b);
                else
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_shade<MRT_MODE_GUIDED, SHADE_ROUNDS>), sg, dim3(SHADE_BLOCK), 0, st, s, a, b,
                                       pb, pb.q_o[qi], pb.q_d[qi], pb.q_t[qi], pb.q_r[qi], cnt + b, pb.q_o[qi ^ 1],
                                       pb.q_d[qi ^ 1], pb.q_t[qi ^ 1], pb.q_r[qi ^ 1], cnt + b + 1, scnt + b);
            } else {
                if (wide)
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_shade<MRT_MODE_BSDF, SHADE_WIDE_ROUNDS>), sgw, dim3(SHADE_BLOCK), 0, st, s, a,
                                       b, pb, pb.q_o[qi], pb.q_d[qi], pb.q_t[qi], pb.q_r[qi], cnt + b, pb.q_o[qi ^ 1],
                                       pb.q_d[qi ^ 1], pb.q_t[qi ^ 1], pb.q_r[qi ^ 1], cnt + b + 1, scnt + b);
                else
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_shade<MRT_MODE_BSDF, SHADE_ROUNDS>), sg, dim3(SHADE_BLOCK), 0, st, s, a, b,
                                       pb, pb.q_o[qi], pb.q_d[qi], pb.q_t[qi], pb.q_r[qi], cnt + b, pb.q_o[qi ^ 1],
                                       pb.q_d[qi ^ 1], pb.q_t[qi ^ 1], pb.q_r[qi ^ 1], cnt + b + 1, scnt + b);
            }
            mark(1);
            if (nee && b + 1 < L && err == hipSuccess) {
                mark(-1);
                err = launch_trace<false, SrcShadow, SinkShadow, true>(
                    s, ts, SrcShadow{pb.sq_o, pb.sq_d}, SinkShadow{pb.sq_o, pb.sq_c, pb.state, pb.q_r[qi ^ 1]}, scnt + b, a.n_paths,
                    swork + b, a.instance_mask, nullptr, st);
                mark(0);
            }
        }
        if (guided && c == 0)
            hipLaunchKernelGGL(k_guide_targets, dim3(grid_for(pb.grec_cap, 256, MAX_BLOCKS)), dim3(256), 0, st, pb);
        hipLaunchKernelGGL(k_accumulate, dim3(grid_for(a.npix, 256, MAX_BLOCKS)), dim3(256), 0, st, a, pb,
                           int(c == 0), int(c + 1 == n_chunks), out, sig_out);
    }
    mark(-1);
    if (timing) {
        float ms_trace = 0.f, ms_shade = 0.f;
        if (terr == hipSuccess) terr = hipEventSynchronize(ev.back());
        // events: [frame start] then per launch (start, end) pairs, then [frame end]
        for (size_t k = 0; k < kind.size() && terr == hipSuccess; k++) {
            float ms = 0.f;
            terr = hipEventElapsedTime(&ms, ev[1 + 2 * k], ev[2 + 2 * k]);
            (kind[k] == 0 ? ms_trace : ms_shade) += ms;
        }
        if (terr == hipSuccess) terr = hipEventElapsedTime(&timing->ms_total, ev.front(), ev.back());
        timing->ms_trace = ms_trace;
        timing->ms_shade = ms_shade;
        timing->trace_launches = 0;
        for (int k : kind) timing->trace_launches += (k == 0);
        for (auto& e : ev) (void)hipEventDestroy(e); // release; nothing to recover on failure
    }
    if (err != hipSuccess) return err;
    if (terr != hipSuccess) return terr;
    if (continuation_segments) {
        std::vector<uint32_t> h(nz);
        err = hipMemcpyAsync(h.data(), pb.counters, sizeof(uint32_t) * nz, hipMemcpyDeviceToHost, st);
        if (err == hipSuccess) err = hipStreamSynchronize(st);
        if (err != hipSuccess) return err;
        uint64_t segs = 0;
        for (uint32_t c = 0; c < n_chunks; c++)
            for (int b = 1; b < L; b++) segs += h[c * per_chunk + b];
        *continuation_segments = segs;
        if (timing && nee) {
            std::vector<uint32_t> hs(nz);
            err = hipMemcpy(hs.data(), pb.counters + 2 * nz, sizeof(uint32_t) * nz, hipMemcpyDeviceToHost);
            if (err != hipSuccess) return err;
            for (uint32_t v : hs) timing->shadow_rays += v;
        }
    }
    return hipGetLastError();
}

// Clears n words with a fill kernel (no memset node, so callers may capture it in a graph).
hipError_t launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(n, 256, MAX_BLOCKS)), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

} // namespace mrt
