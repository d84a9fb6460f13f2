// CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// A plain C++ restatement of the reference's `pt` hot path (render_pt.slang and the Slang
// library it imports), used to check the HIP product path bit-for-bit (hit ids, t) and to
// within tolerance (radiance). Built by oracle/Makefile with -O2 -ffp-contract=off into
// oracle/liboracle.so; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// load it. Every function cites the reference file:line it restates (paths relative to the
// reference repository root).
//
// Pinned by (tests/test_oracle_*.py): brute-force Moller-Trumbore on small inputs, closed
// forms (camera basis, RNG integer identities, oct encoding round trips), the reference's BSDF
// statistical invariants (tests/test_bsdf.cpp:31-58,285-334) and its Lambert known answer
// (tests/test_material_system.cpp:165-215), and end-to-end known answers (convex Lambert
// object under a constant environment: pixel = albedo * L exactly).

#include "oracle_bsdf.h"
#include "../include/merian_amd.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <thread>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------------------
// Oct normal encoding (include/merian-shaders/utils/encoding.slang:6-55)

inline float copysignf_(float a, float b) { return std::copysign(a, b); }

inline int16_t pack_snorm(float v) {
    const float c = clampf(v, -1.f, 1.f) * 32767.0f;
    return int16_t(std::nearbyint(c)); // round half to even
}
inline uint32_t encode_normal(f3 n) {
    const float inv_l1 = 1.0f / ((std::fabs(n.x) + std::fabs(n.y)) + std::fabs(n.z));
    f2 p;
    if (n.z < 0.0f) {
        p.x = copysignf_(1.0f - std::fabs(n.y * inv_l1), n.x);
        p.y = copysignf_(1.0f - std::fabs(n.x * inv_l1), n.y);
    } else {
        p.x = n.x * inv_l1;
        p.y = n.y * inv_l1;
    }
    return uint32_t(uint16_t(pack_snorm(p.x))) | (uint32_t(uint16_t(pack_snorm(p.y))) << 16);
}
inline f3 decode_normal(uint32_t enc) {
    const float px = clampf(float(int16_t(enc & 0xFFFFu)) / 32767.0f, -1.f, 1.f);
    const float py = clampf(float(int16_t(enc >> 16)) / 32767.0f, -1.f, 1.f);
    f3 v{px, py, (1.0f - std::fabs(px)) - std::fabs(py)};
    if (v.z < 0.0f) {
        const float x = v.x, y = v.y;
        v.x = copysignf_(1.0f - std::fabs(y), x);
        v.y = copysignf_(1.0f - std::fabs(x), y);
    }
    return normalize(v);
}

// scene/hit.slang:14-23 (Wachter & Binder 2019)
inline float offset_component(float p, float n) {
    const int32_t io = int32_t(n * 256.f);
    int32_t bits;
    std::memcpy(&bits, &p, 4);
    bits += (p < 0.f) ? -io : io;
    float ip;
    std::memcpy(&ip, &bits, 4);
    return std::fabs(p) < (1.f / 16.f) ? p + n * (1.f / 65536.f) : ip;
}
inline f3 offset_ray_origin(f3 p, f3 n) {
    return {offset_component(p.x, n.x), offset_component(p.y, n.y), offset_component(p.z, n.z)};
}

// ---------------------------------------------------------------------------------------
// Frames (utils/frame.slang:7-135)

struct Frame {
    f3 r0, r1, r2; // tangent, bitangent, normal (rows)
    f3 to_local(f3 v) const { return {dot(r0, v), dot(r1, v), dot(r2, v)}; }
    // mul(v, basis) = v.x*r0 + v.y*r1 + v.z*r2
    f3 to_world(f3 v) const {
        return {v.x * r0.x + v.y * r1.x + v.z * r2.x, v.x * r0.y + v.y * r1.y + v.z * r2.y,
                v.x * r0.z + v.y * r1.z + v.z * r2.z};
    }
};
// make_frame: Duff et al. 2017 (frame.slang:7-13)
inline Frame frame_from_normal(f3 z) {
    const float s = (z.z >= 0) ? 1.f : -1.f;
    const float a = -1.0f / (s + z.z);
    const float b = z.x * z.y * a;
    Frame f;
    f.r0 = {1.0f + s * z.x * z.x * a, s * b, -s * z.x};
    f.r1 = {b, s + z.y * z.y * a, -z.y};
    f.r2 = z;
    return f;
}
// OrthonormalFrame(n, t, handedness) (frame.slang:90-96)
inline Frame frame_from_nt(f3 n, f3 t, float hs) {
    const f3 N = normalize(n);
    const f3 T = normalize(t - N * dot(t, N));
    const f3 B = cross(N, T) * hs;
    return Frame{T, B, N};
}
// apply_normal_map (frame.slang:123-132)
inline Frame apply_normal_map(const Frame& f, f3 n_local) {
    const f3 new_n = f.to_world(n_local);
    const f3 old_t = f.r0;
    const f3 new_t = normalize(old_t - new_n * dot(old_t, new_n));
    const f3 new_b = cross(new_n, new_t);
    return Frame{new_t, new_b, new_n};
}

// ---------------------------------------------------------------------------------------
// Camera (camera.slang:29-56; camera.cpp:263-266,318-362)

struct Ray {
    f3 o, d;
    float tmin, tmax;
};

inline f3 ld3(const float* p) { return {p[0], p[1], p[2]}; }

inline Ray camera_ray(const mrt_camera& c, uint32_t px, uint32_t py, uint32_t w, uint32_t h) {
    const float p_x = ((float(px) + 0.5f) + c.jitter[0]) / float(w);
    const float p_y = ((float(py) + 0.5f) + c.jitter[1]) / float(h);
    const float ndc_x = 2.0f * p_x + -1.0f;
    const float ndc_y = -2.0f * p_y + 1.0f;
    const f3 U = ld3(c.U), V = ld3(c.V), W = ld3(c.W);
    Ray r;
    r.o = ld3(c.position);
    r.d = normalize(ndc_x * U + ndc_y * V + W);
    const float inv_cos = 1.f / dot(normalize(W), r.d);
    r.tmin = c.near_plane * inv_cos;
    r.tmax = c.far_plane * inv_cos;
    return r;
}

// ---------------------------------------------------------------------------------------
// Ray-triangle: textbook Moller-Trumbore (1997), fixed operation order, no FMA contraction.
// Culling: the front side is the one cross(e01, e02) points to (scene-data.slangh:20,
// scene.slang:331-333); a back-face-culled triangle needs det > 0.

struct TriRec {
    f3 v0, e1, e2;
    uint32_t id;
    uint8_t cull;
    uint8_t mask;
};

inline bool intersect_mt(const Ray& r, const TriRec& t, float& tt, float& u, float& v) {
    const f3 p = cross(r.d, t.e2);
    const float det = dot(t.e1, p);
    if (t.cull) {
        if (!(det > 0.0f)) return false;
    } else {
        if (!(det != 0.0f)) return false;
    }
    const float inv = 1.0f / det;
    const f3 s = r.o - t.v0;
    u = dot(s, p) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    const f3 q = cross(s, t.e1);
    v = dot(r.d, q) * inv;
    if (!(v >= 0.0f && u + v <= 1.0f)) return false;
    tt = dot(t.e2, q) * inv;
    return tt >= r.tmin && tt <= r.tmax;
}

struct HitRec {
    uint32_t id;
    float t, u, v;
};

inline void consider(const Ray& r, const TriRec& tr, uint32_t ray_mask, HitRec& best) {
    if ((tr.mask & ray_mask) == 0u) return;
    float t, u, v;
    if (!intersect_mt(r, tr, t, u, v)) return;
    if (t < best.t || (t == best.t && tr.id < best.id)) best = HitRec{tr.id, t, u, v};
}

// Conservative BVH2 (median split), boxes in double and inflated, so that every triangle the
// Moller-Trumbore test could report is visited; the result equals brute force.
struct BVH {
    struct Node {
        double lo[3], hi[3];
        int32_t left, right; // children, or -1
        uint32_t first, count;
    };
    std::vector<Node> nodes;
    std::vector<TriRec> tris; // reordered

    void build(std::vector<TriRec> in) {
        tris = std::move(in);
        nodes.clear();
        nodes.reserve(2 * tris.size() / 2 + 2);
        std::vector<double> cen(3 * tris.size());
        for (size_t i = 0; i < tris.size(); i++) {
            const TriRec& t = tris[i];
            const double a[3] = {t.v0.x, t.v0.y, t.v0.z};
            const double b[3] = {(double)(t.v0.x + t.e1.x), (double)(t.v0.y + t.e1.y), (double)(t.v0.z + t.e1.z)};
            const double c[3] = {(double)(t.v0.x + t.e2.x), (double)(t.v0.y + t.e2.y), (double)(t.v0.z + t.e2.z)};
            for (int k = 0; k < 3; k++) cen[3 * i + k] = (a[k] + b[k] + c[k]) / 3.0;
        }
        std::vector<uint32_t> idx(tris.size());
        for (size_t i = 0; i < idx.size(); i++) idx[i] = uint32_t(i);
        if (!tris.empty()) build_rec(idx, cen, 0, uint32_t(idx.size()));
        std::vector<TriRec> re(tris.size());
        for (size_t i = 0; i < idx.size(); i++) re[i] = tris[idx[i]];
        tris.swap(re);
    }
    void tri_bounds(const TriRec& t, double lo[3], double hi[3]) const {
        const float pa[3] = {t.v0.x, t.v0.y, t.v0.z};
        const float e1[3] = {t.e1.x, t.e1.y, t.e1.z};
        const float e2[3] = {t.e2.x, t.e2.y, t.e2.z};
        for (int k = 0; k < 3; k++) {
            // v1 = v0 + e1 is not exact in float; bound with the exact double sums
            const double a = pa[k], b = (double)pa[k] + e1[k], c = (double)pa[k] + e2[k];
            lo[k] = std::min(lo[k], std::min(a, std::min(b, c)));
            hi[k] = std::max(hi[k], std::max(a, std::max(b, c)));
        }
    }
    int32_t build_rec(std::vector<uint32_t>& idx, const std::vector<double>& cen, uint32_t b, uint32_t e) {
        const int32_t ni = int32_t(nodes.size());
        nodes.push_back(Node{});
        Node nd;
        for (int k = 0; k < 3; k++) {
            nd.lo[k] = DBL_MAX;
            nd.hi[k] = -DBL_MAX;
        }
        for (uint32_t i = b; i < e; i++) tri_bounds(tris[idx[i]], nd.lo, nd.hi);
        double mag = 0;
        for (int k = 0; k < 3; k++) mag = std::max(mag, std::max(std::fabs(nd.lo[k]), std::fabs(nd.hi[k])));
        const double pad = 1e-5 * mag + 1e-30;
        for (int k = 0; k < 3; k++) {
            nd.lo[k] -= pad;
            nd.hi[k] += pad;
        }
        nd.left = nd.right = -1;
        nd.first = b;
        nd.count = e - b;
        if (e - b > 4) {
            double clo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, chi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
            for (uint32_t i = b; i < e; i++)
                for (int k = 0; k < 3; k++) {
                    clo[k] = std::min(clo[k], cen[3 * idx[i] + k]);
                    chi[k] = std::max(chi[k], cen[3 * idx[i] + k]);
                }
            int ax = 0;
            for (int k = 1; k < 3; k++)
                if (chi[k] - clo[k] > chi[ax] - clo[ax]) ax = k;
            const uint32_t mid = (b + e) / 2;
            std::nth_element(idx.begin() + b, idx.begin() + mid, idx.begin() + e,
                             [&](uint32_t x, uint32_t y) { return cen[3 * x + ax] < cen[3 * y + ax]; });
            nd.count = 0;
            nodes[ni] = nd;
            const int32_t l = build_rec(idx, cen, b, mid);
            const int32_t r = build_rec(idx, cen, mid, e);
            nodes[ni].left = l;
            nodes[ni].right = r;
            return ni;
        }
        nodes[ni] = nd;
        return ni;
    }
    static bool box_hit(const Node& n, const Ray& r, double best_t) {
        const double o[3] = {r.o.x, r.o.y, r.o.z};
        const double d[3] = {r.d.x, r.d.y, r.d.z};
        double t0 = r.tmin, t1 = std::min((double)r.tmax, best_t);
        // widen the interval so round-off in Moller-Trumbore's t can never hide a hit
        t0 = t0 - 1e-6 * std::fabs(t0) - 1e-30;
        t1 = t1 + 1e-6 * std::fabs(t1) + 1e-30;
        for (int k = 0; k < 3; k++) {
            if (d[k] == 0.0) {
                if (o[k] < n.lo[k] || o[k] > n.hi[k]) return false;
                continue;
            }
            const double inv = 1.0 / d[k];
            double a = (n.lo[k] - o[k]) * inv, bb = (n.hi[k] - o[k]) * inv;
            if (a > bb) std::swap(a, bb);
            t0 = std::max(t0, a);
            t1 = std::min(t1, bb);
            if (t0 > t1) return false;
        }
        return true;
    }
    HitRec closest(const Ray& r, uint32_t ray_mask) const {
        HitRec best{0xFFFFFFFFu, INFINITY, 0.f, 0.f};
        if (nodes.empty()) return best;
        int32_t stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp) {
            const Node& n = nodes[stack[--sp]];
            if (!box_hit(n, r, best.t)) continue;
            if (n.left < 0) {
                for (uint32_t i = 0; i < n.count; i++) consider(r, tris[n.first + i], ray_mask, best);
            } else {
                stack[sp++] = n.right;
                stack[sp++] = n.left;
            }
        }
        return best;
    }
};

// ---------------------------------------------------------------------------------------
// Scene (host mirror of the Scene parameter block, scene.slang:30-54)

struct Scene {
    std::vector<mrt_vertex> vertices;
    std::vector<uint32_t> indices;
    std::vector<mrt_geometry> geoms;
    std::vector<uint32_t> geom_first; // global id of each geometry's first triangle
    std::vector<mrt_material> materials;
    mrt_env env{};
    std::vector<float> env_rgba;
    std::vector<TriRec> tri_list; // input order, global ids
    BVH bvh;
    bool use_bvh = true;

    uint32_t vertex_index(const mrt_geometry& g, uint32_t prim, int k) const {
        if (g.flags & MRT_GEOM_INDEX_UINT32) return indices[g.index_offset + 3 * prim + k];
        return 3 * prim + uint32_t(k);
    }
    f3 vpos(const mrt_geometry& g, uint32_t local) const {
        const mrt_vertex& v = vertices[g.vertex_offset + local];
        return {v.position[0], v.position[1], v.position[2]};
    }
    void build() {
        geom_first.resize(geoms.size());
        uint32_t gid = 0;
        tri_list.clear();
        for (size_t g = 0; g < geoms.size(); g++) {
            geom_first[g] = gid;
            const mrt_geometry& G = geoms[g];
            for (uint32_t p = 0; p < G.primitive_count; p++) {
                const f3 a = vpos(G, vertex_index(G, p, 0));
                const f3 b = vpos(G, vertex_index(G, p, 1));
                const f3 c = vpos(G, vertex_index(G, p, 2));
                TriRec t;
                t.v0 = a;
                t.e1 = b - a;
                t.e2 = c - a;
                t.id = gid + p;
                t.cull = (G.flags & MRT_GEOM_TWO_SIDED) ? 0 : 1;
                t.mask = uint8_t(G.instance_mask & 0xFFu);
                tri_list.push_back(t);
            }
            gid += G.primitive_count;
        }
        if (use_bvh) bvh.build(tri_list);
    }
    HitRec trace(const Ray& r, uint32_t mask) const {
        if (use_bvh) return bvh.closest(r, mask);
        HitRec best{0xFFFFFFFFu, INFINITY, 0.f, 0.f};
        for (const TriRec& t : tri_list) consider(r, t, mask, best);
        return best;
    }
    // global triangle id -> (geometry, primitive)
    void locate(uint32_t gid, uint32_t& g, uint32_t& prim) const {
        const auto it = std::upper_bound(geom_first.begin(), geom_first.end(), gid);
        g = uint32_t(it - geom_first.begin()) - 1;
        prim = gid - geom_first[g];
    }
};

// Hit (scene/hit.slang:25-65); one pretransformed group: instance_id = 0, instance_index = 0.
struct Hit {
    uint32_t instance_id, instance_index, primitive_id, geometry_index;
    f2 bary;
    bool valid() const { return instance_id != 0xFFFFFFFFu; }
};
inline Hit invalid_hit() { return Hit{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, {0.f, 0.f}}; }

inline Hit to_hit(const Scene& sc, const HitRec& h) {
    if (h.id == 0xFFFFFFFFu) return invalid_hit();
    uint32_t g, p;
    sc.locate(h.id, g, p);
    return Hit{0u, 0u, p, g, {h.u, h.v}};
}

// PackedHitInfo (gbuffer.slang:175-196)
inline mrt_packed_hit pack_hit(const Hit& h) {
    mrt_packed_hit p;
    p.instance_id = h.instance_id;
    p.primitive_id = h.primitive_id;
    p.geometry_index = uint16_t(h.geometry_index);
    p.instance_index = uint16_t(h.instance_index);
    p.barycentrics[0] = f32_to_f16(h.bary.x);
    p.barycentrics[1] = f32_to_f16(h.bary.y);
    return p;
}
inline Hit unpack_hit(const mrt_packed_hit& p) {
    return Hit{p.instance_id, uint32_t(p.instance_index), p.primitive_id, uint32_t(p.geometry_index),
               {f16_to_f32(p.barycentrics[0]), f16_to_f32(p.barycentrics[1])}};
}

// ---------------------------------------------------------------------------------------
// Environment (environment-map.slang:13-68, utils/math.slang:243-251)

inline f3 texel(const Scene& sc, int x, int y) {
    const mrt_env& e = sc.env;
    const int w = int(e.width), h = int(e.height);
    if (e.wrap == MRT_WRAP_REPEAT) {
        x = ((x % w) + w) % w;
        y = ((y % h) + h) % h;
    } else {
        x = std::min(std::max(x, 0), w - 1);
        y = std::min(std::max(y, 0), h - 1);
    }
    const float* p = &sc.env_rgba[4 * (size_t(y) * w + x)];
    return {p[0], p[1], p[2]};
}
// bilinear LOD0 fetch, texel centres at (i + 0.5) / size; weights blended as x + t * (y - x)
inline float tlerp(float x, float y, float t) { return x + t * (y - x); }
inline f3 tlerp(f3 x, f3 y, float t) { return {tlerp(x.x, y.x, t), tlerp(x.y, y.y, t), tlerp(x.z, y.z, t)}; }
inline f3 sample_bilinear(const Scene& sc, f2 uv) {
    const float fx = uv.x * float(sc.env.width) - 0.5f;
    const float fy = uv.y * float(sc.env.height) - 0.5f;
    const float x0f = std::floor(fx), y0f = std::floor(fy);
    const float ax = fx - x0f, ay = fy - y0f;
    const int x0 = int(x0f), y0 = int(y0f);
    const f3 t00 = texel(sc, x0, y0), t10 = texel(sc, x0 + 1, y0);
    const f3 t01 = texel(sc, x0, y0 + 1), t11 = texel(sc, x0 + 1, y0 + 1);
    return tlerp(tlerp(t00, t10, ax), tlerp(t01, t11, ax), ay);
}
inline f3 mul3(const float* m, f3 v) {
    return {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
            m[6] * v.x + m[7] * v.y + m[8] * v.z};
}
// EqualAreaOctMap::dir_to_uv (environment-map.slang:44-63)
inline f2 equal_area_dir_to_uv(f3 d) {
    const f3 a{std::fabs(d.x), std::fabs(d.y), std::fabs(d.z)};
    const float r = std::sqrt(std::fmax(0.f, 1.f - a.z));
    const float m = std::fmax(a.x, a.y);
    const float b = m == 0.f ? 0.f : std::fmin(a.x, a.y) / m;
    float phi = o_atan(b) * (2.f / PI);
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
    return {0.5f * (u + 1.f), 0.5f * (v + 1.f)};
}
// world_to_latlong_map (utils/math.slang:243-249)
inline f2 latlong_uv(f3 dir) {
    const f3 p = normalize(dir);
    return {o_atan2(p.x, -p.z) * INV_2PI + 0.5f, o_acos(p.y) * M_1_PI_F};
}
// env.eval(sd) with sd.wi = -ray.d, so the looked-up direction is -sd.wi.
inline f3 env_eval(const Scene& sc, f3 minus_wi) {
    const mrt_env& e = sc.env;
    if (e.type == MRT_ENV_NONE) return {0, 0, 0};
    const f3 d = mul3(e.to_local, minus_wi);
    const f2 uv = e.type == MRT_ENV_LATLONG ? latlong_uv(d) : equal_area_dir_to_uv(d);
    return e.intensity * sample_bilinear(sc, uv);
}

// ---------------------------------------------------------------------------------------
// Materials (material-system.slang:51-72, diffuse-material.slang, openpbr-material.slang)

struct ShadingData {
    f3 wi, pos;
    f2 uv;
    Frame frame;
    f3 face_normal;
    uint32_t bounce;
    float ior;
    bool front_facing;
};

// One material sample: which tree it is, the tree, and the interior eta.
struct MaterialSample {
    int model; // MRT_MAT_*
    Lambert diffuse;
    OpenPBRBSDF pbr;
    float eta;
    Sample sample_eval(f3 wi, PCG& rng) const {
        return model == MRT_MAT_DIFFUSE ? diffuse.sample_eval(wi, rng) : pbr.sample_eval(wi, rng);
    }
    f3 eval(f3 wi, f3 wo) const { return model == MRT_MAT_DIFFUSE ? diffuse.eval(wi, wo) : pbr.eval(wi, wo); }
    float pdf(f3 wi, f3 wo) const { return model == MRT_MAT_DIFFUSE ? diffuse.pdf(wi, wo) : pbr.pdf(wi, wo); }
    float alpha() const { return model == MRT_MAT_DIFFUSE ? diffuse.alpha() : pbr.alpha(); }
    f3 get_albedo(f3 wi) const { return model == MRT_MAT_DIFFUSE ? diffuse.get_albedo(wi) : pbr.get_albedo(wi); }
};

constexpr float MIN_ROUGHNESS = 0.0316f; // material.slang:23

inline f3 material_emission(const mrt_material& m, const ShadingData& sd) {
    if (m.model == MRT_MAT_DIFFUSE) return {0, 0, 0};
    if (!sd.front_facing) return {0, 0, 0};
    return ld3(m.emission);
}

inline MaterialSample get_material_sample(const mrt_material& m, const ShadingData& sd, f3& emission,
                                          Frame& shading_frame) {
    MaterialSample s{};
    s.model = int(m.model);
    if (m.model == MRT_MAT_DIFFUSE) {
        emission = {0, 0, 0};
        shading_frame = apply_normal_map(sd.frame, f3{0, 0, 1});
        s.diffuse = Lambert{ld3(m.base_color)};
        s.eta = 1.0f;
        return s;
    }
    const bool has_tr = (m.features & MRT_PBR_TRANSMISSION) != 0;
    const bool has_coat = (m.features & MRT_PBR_CLEARCOAT) != 0;
    const bool has_sheen = (m.features & MRT_PBR_SHEEN) != 0;
    const f3 base = ld3(m.base_color);
    const float metal = m.metalness;
    const float min_alpha = pow2(MIN_ROUGHNESS);
    f2 alpha{std::fmax(m.specular_alpha[0], min_alpha), std::fmax(m.specular_alpha[1], min_alpha)};
    shading_frame = sd.frame;
    const float ior_t = (m.specular_ior - 1.0f) / (m.specular_ior + 1.0f);
    const float dielectric_f0 = clampf(m.specular_weight * ior_t * ior_t, 0.0f, 1.0f);
    float eta = 1.0f, tw = 0.0f;
    f3 tcol{1.0f, 1.0f, 1.0f};
    if (has_tr) {
        eta = sd.ior / (sd.front_facing ? m.specular_ior : 1.0f);
        tw = m.transmission_weight;
        tcol = ld3(m.transmission_color);
    }
    const float alpha_iso = has_tr ? std::sqrt(alpha.x * alpha.y) : alpha.x;
    const Dielectric diel{OpaqueDielectric{GGX{alpha}, Lambert{base}, dielectric_f0},
                          RoughDielectric{alpha_iso, eta, tcol}, tw};
    const Substrate substrate{diel, Metal{GGX{alpha}, base}, metal};
    float cw = 0.0f, coat_alpha = min_alpha, coat_f0 = 0.0f;
    if (has_coat) {
        cw = m.coat_weight;
        coat_alpha = std::fmax(m.coat_alpha, min_alpha);
        const float ct = (m.coat_ior - 1.0f) / (m.coat_ior + 1.0f);
        coat_f0 = ct * ct;
    }
    const WithCoat with_coat{substrate, Coated{GGX{{coat_alpha, coat_alpha}}, substrate, coat_f0}, cw};
    float sw = 0.0f, sheen_alpha = 0.0f;
    f3 sheen_col{0.0f, 0.0f, 0.0f};
    if (has_sheen) {
        sw = m.sheen_weight;
        sheen_col = ld3(m.sheen_color);
        sheen_alpha = m.sheen_alpha;
    }
    s.pbr = OpenPBRBSDF{with_coat, Sheen::make(sheen_col, sheen_alpha), sw};
    emission = material_emission(m, sd);
    s.eta = eta;
    return s;
}

// ---------------------------------------------------------------------------------------
// Shading points (scene.slang:312-498, hit.slang:106-147)

constexpr float ENV_MAP_DISTANCE = 1e7f;
constexpr float SCENE_RAY_TMAX = 1e7f;
constexpr float SCENE_THROUGHPUT_THRESHOLD = 1e-5f;

struct ShadingPoint {
    f3 pos;
    uint32_t enc_face_normal;
    f3 emission;
    Frame shading_frame;
    bool has_material;
    MaterialSample mat;
};

inline ShadingPoint env_shading_point(const Scene& sc, const Ray& ray) {
    ShadingPoint sp{};
    sp.pos = ray.o + ray.d * ENV_MAP_DISTANCE;
    sp.enc_face_normal = encode_normal(-ray.d);
    sp.emission = env_eval(sc, -(-ray.d));
    sp.shading_frame = frame_from_normal(-ray.d);
    sp.has_material = false;
    return sp;
}

// The geometric part of Scene.get_surface (scene.slang:322-382): face normal (FlipFacing),
// vertex normal (FlatShading), and the shading frame from per-vertex tangents or the Lengyel
// UV-derived fallback (scene.slang:357-371).
struct SurfaceGeom {
    f3 pos, face_normal, vertex_normal;
    f2 uv0, uv1, uv2;
    Frame frame;
};
inline SurfaceGeom surface_geometry(const mrt_vertex& v0, const mrt_vertex& v1, const mrt_vertex& v2, uint32_t flags,
                                    f3 b) {
    const f3 p0 = ld3(v0.position), p1 = ld3(v1.position), p2 = ld3(v2.position);
    SurfaceGeom g;
    g.pos = b.x * p0 + b.y * p1 + b.z * p2;
    const f3 e01 = p1 - p0, e02 = p2 - p0;
    f3 face_normal = normalize(cross(e01, e02));
    if (flags & MRT_GEOM_FLIP_FACING) face_normal = -face_normal;
    f3 vertex_normal;
    if (flags & MRT_GEOM_FLAT_SHADING) {
        vertex_normal = face_normal;
    } else {
        vertex_normal = normalize(b.x * decode_normal(v0.encoded_normal) + b.y * decode_normal(v1.encoded_normal) +
                                  b.z * decode_normal(v2.encoded_normal));
    }
    const f2 uv0{f16_to_f32(v0.uv[0]), f16_to_f32(v0.uv[1])};
    const f2 uv1{f16_to_f32(v1.uv[0]), f16_to_f32(v1.uv[1])};
    const f2 uv2{f16_to_f32(v2.uv[0]), f16_to_f32(v2.uv[1])};
    Frame frame;
    if (flags & MRT_GEOM_HAS_TANGENTS) {
        const auto dec_t = [](uint32_t enc, float& w) {
            w = (enc & 1u) ? -1.f : 1.f;
            return decode_normal(enc & ~1u);
        };
        float w0, w1, w2;
        const f3 t0 = dec_t(v0.encoded_tangent, w0), t1 = dec_t(v1.encoded_tangent, w1),
                 t2 = dec_t(v2.encoded_tangent, w2);
        const f3 txyz = b.x * t0 + b.y * t1 + b.z * t2;
        const float tw = b.x * w0 + b.y * w1 + b.z * w2;
        frame = frame_from_nt(vertex_normal, normalize(txyz), signf(tw));
    } else {
        const f2 d1{uv1.x - uv0.x, uv1.y - uv0.y};
        const f2 d2{uv2.x - uv0.x, uv2.y - uv0.y};
        const float det = d1.x * d2.y - d2.x * d1.y;
        if (std::fabs(det) > 1e-8f) {
            const float ws = (flags & MRT_GEOM_FLIP_FACING) ? 1.f : -1.f;
            const f3 t = (d2.y * e01 - d1.y * e02) / det;
            frame = frame_from_nt(vertex_normal, t, signf(det) * ws);
        } else {
            frame = frame_from_normal(vertex_normal);
        }
    }
    g.face_normal = face_normal;
    g.vertex_normal = vertex_normal;
    g.uv0 = uv0;
    g.uv1 = uv1;
    g.uv2 = uv2;
    g.frame = frame;
    return g;
}

inline ShadingPoint surface_shading_point(const Scene& sc, const Ray& ray, const Hit& hit, uint32_t bounce,
                                          float ior, bool with_material) {
    const mrt_geometry& G = sc.geoms[hit.instance_id + hit.geometry_index];
    const mrt_vertex& v0 = sc.vertices[G.vertex_offset + sc.vertex_index(G, hit.primitive_id, 0)];
    const mrt_vertex& v1 = sc.vertices[G.vertex_offset + sc.vertex_index(G, hit.primitive_id, 1)];
    const mrt_vertex& v2 = sc.vertices[G.vertex_offset + sc.vertex_index(G, hit.primitive_id, 2)];
    const f3 b{(1.0f - hit.bary.x) - hit.bary.y, hit.bary.x, hit.bary.y};
    const SurfaceGeom g = surface_geometry(v0, v1, v2, G.flags, b);
    const f3 pos = g.pos, face_normal = g.face_normal;
    const f2 uv0 = g.uv0, uv1 = g.uv1, uv2 = g.uv2;
    Frame frame = g.frame;
    f3 fn = face_normal;
    bool front = true;
    if (G.flags & MRT_GEOM_TWO_SIDED) {
        front = dot(-ray.d, face_normal) >= 0.f;
        if (!front) {
            frame.r2 = -frame.r2;
            fn = -fn;
        }
    }
    ShadingData sd;
    sd.wi = -ray.d;
    sd.pos = pos;
    sd.uv = {b.x * uv0.x + b.y * uv1.x + b.z * uv2.x, b.x * uv0.y + b.y * uv1.y + b.z * uv2.y};
    sd.frame = frame;
    sd.face_normal = fn;
    sd.bounce = bounce;
    sd.ior = ior;
    sd.front_facing = front;

    ShadingPoint sp{};
    sp.pos = pos;
    sp.enc_face_normal = encode_normal(face_normal);
    sp.has_material = with_material;
    const mrt_material& m = sc.materials[G.material_id];
    if (with_material) {
        sp.mat = get_material_sample(m, sd, sp.emission, sp.shading_frame);
    } else {
        sp.emission = material_emission(m, sd);
        sp.shading_frame = frame;
    }
    return sp;
}

inline ShadingPoint get_shading_point(const Scene& sc, const Ray& ray, const Hit& hit, uint32_t bounce, float ior,
                                      bool with_material) {
    if (hit.valid()) {
        const mrt_geometry& G = sc.geoms[hit.instance_id + hit.geometry_index];
        if ((G.flags & MRT_GEOM_USE_ENV_MAP) == 0)
            return surface_shading_point(sc, ray, hit, bounce, ior, with_material);
    }
    return env_shading_point(sc, ray);
}

// ShadingPoint.make_ray (hit.slang:134-146)
inline bool make_ray(const ShadingPoint& sp, f3 wi, f3 wo, bool transmission, Ray& out) {
    const f3 ng = decode_normal(sp.enc_face_normal);
    const float wo_ng = dot(wo, ng);
    if (((dot(wi, ng) < 0.f) != (wo_ng < 0.f)) != transmission) return false;
    out.o = offset_ray_origin(sp.pos, wo_ng < 0.f ? -ng : ng);
    out.d = wo;
    out.tmin = 0.f;
    out.tmax = SCENE_RAY_TMAX;
    return true;
}

// The GBuffer albedo of a primary shading point (gbuffer_rt.slang:76-77 for a surface: the BSDF's
// directional albedo towards the camera in the material's shading frame; :92 for a miss: the
// environment's emission).
inline f3 gbuffer_albedo(const Ray& ray, const ShadingPoint& sp) {
    return sp.has_material ? sp.mat.get_albedo(sp.shading_frame.to_local(-ray.d)) : sp.emission;
}
// get_demodulation_albedo (gbuffer.slang:135-140): floored at 1e-2 per channel.
inline f3 demodulation_albedo(const Ray& ray, const ShadingPoint& sp) {
    const f3 a = gbuffer_albedo(ray, sp);
    return {std::fmax(a.x, 1e-2f), std::fmax(a.y, 1e-2f), std::fmax(a.z, 1e-2f)};
}

// yuv_luminance (colors/colorspace_yuv.slang:5-7)
inline float yuv_luminance(f3 c) { return dot(c, f3{0.299f, 0.587f, 0.114f}); }

inline uint32_t sig_step(uint32_t h, uint32_t id) { return (h ^ (id + 0x7F4A7C15u)) * 0x01000193u; }

// trace_path (render_pt.slang:24-90)
// pcg3d16 (utils/hash.slang:83-90): the MCPG renderer's per-pixel seed
inline uint32_t pcg3d16_(uint32_t x, uint32_t y, uint32_t z) {
    const uint32_t a = 1664525u, c = 1013904223u;
    uint32_t v[3] = {x * a + c, y * a + c, z * a + c};
    v[0] += v[1] * v[2];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    v[0] += v[1] * v[2];
    return v[0];
}

inline f3 trace_path(const Scene& sc, const mrt_pt_params& prm, Ray ray, ShadingPoint sp, PCG& rng,
                     uint32_t& sig, uint64_t& segs) {
    float current_ior = 1.0f;
    const f3 throughput0{1.f, 1.f, 1.f};
    f3 throughput = throughput0;
    f3 radiance = prm.emission_on_primary ? throughput * sp.emission : f3{0, 0, 0};
    for (int bounce = 1; bounce < prm.max_path_length; bounce++) {
        if (!sp.has_material || (throughput.x < SCENE_THROUGHPUT_THRESHOLD && throughput.y < SCENE_THROUGHPUT_THRESHOLD &&
                                 throughput.z < SCENE_THROUGHPUT_THRESHOLD))
            break;
        const Frame frame = sp.shading_frame;
        const f3 wi_local = frame.to_local(-ray.d);
        const Sample s = sp.mat.sample_eval(wi_local, rng);
        if (!s.valid) break;
        const f3 wo = frame.to_world(s.wo);
        throughput = throughput * s.weight;
        const bool transmission = wi_local.z * s.wo.z < 0.f;
        if (transmission) current_ior /= sp.mat.eta;
        Ray next;
        if (!make_ray(sp, -ray.d, wo, transmission, next)) break;
        ray = next;
        const HitRec hr = sc.trace(ray, prm.instance_mask);
        segs++;
        sig = sig_step(sig, hr.id);
        sp = get_shading_point(sc, ray, to_hit(sc, hr), uint32_t(bounce), current_ior, bounce + 1 < prm.max_path_length);
        radiance = radiance + throughput * sp.emission;
    }
    return radiance;
}

} // namespace orc

using namespace orc;

// =======================================================================================
// C API (ctypes)
// =======================================================================================

extern "C" {

__attribute__((visibility("default"))) uint32_t orc_pcg4d16(uint32_t x, uint32_t y, uint32_t z, uint32_t w) {
    return pcg4d16(x, y, z, w);
}
__attribute__((visibility("default"))) void orc_pcg_stream(uint32_t seed, uint32_t n, uint32_t* out_u32,
                                                           float* out_f) {
    PCG r{seed};
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t u = r.next();
        if (out_u32) out_u32[i] = u;
        if (out_f) out_f[i] = u32_to_unit_float(u);
    }
}
__attribute__((visibility("default"))) uint32_t orc_encode_normal(const float* n) { return encode_normal(ld3(n)); }
__attribute__((visibility("default"))) void orc_decode_normal(uint32_t e, float* out) {
    const f3 v = decode_normal(e);
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
}
__attribute__((visibility("default"))) void orc_offset_ray_origin(const float* p, const float* n, float* out) {
    const f3 v = offset_ray_origin(ld3(p), ld3(n));
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
}
// elementary functions: op 0 sin 1 cos 2 atan 3 atan2(x, y) 4 acos 5 log2 6 exp2 7 exp 8 pow(x, y)
__attribute__((visibility("default"))) void orc_math(int op, const float* x, const float* y, uint64_t n, float* out) {
    for (uint64_t i = 0; i < n; i++) {
        const float a = x[i], b = y ? y[i] : 0.f;
        float r = 0.f;
        switch (op) {
        case 0: r = o_sin(a); break;
        case 1: r = o_cos(a); break;
        case 2: r = o_atan(a); break;
        case 3: r = o_atan2(a, b); break;
        case 4: r = o_acos(a); break;
        case 5: r = o_log2(a); break;
        case 6: r = o_exp2(a); break;
        case 7: r = o_exp(a); break;
        default: r = o_pow(a, b); break;
        }
        out[i] = r;
    }
}
// Frames (frame.slang:7-13,90-96,123-132): op 0 make_frame(a); op 1 OrthonormalFrame(a, b, hs);
// op 2 OrthonormalFrame(a, b, hs).apply_normal_map(c). out = rows T, B, N (9 floats).
__attribute__((visibility("default"))) void orc_frame(int32_t op, const float* a, const float* b, float hs,
                                                      const float* c, float* out) {
    Frame f = op == 0 ? frame_from_normal(ld3(a)) : frame_from_nt(ld3(a), ld3(b), hs);
    if (op == 2) f = apply_normal_map(f, ld3(c));
    const f3 rows[3] = {f.r0, f.r1, f.r2};
    for (int r = 0; r < 3; r++) {
        out[3 * r] = rows[r].x;
        out[3 * r + 1] = rows[r].y;
        out[3 * r + 2] = rows[r].z;
    }
}
// EqualAreaOctMap::dir_to_uv (environment-map.slang:44-63) over n directions (float[3n] -> float[2n]).
__attribute__((visibility("default"))) void orc_equal_area_uv(const float* d, uint64_t n, float* uv) {
    for (uint64_t i = 0; i < n; i++) {
        const f2 r = equal_area_dir_to_uv(ld3(d + 3 * i));
        uv[2 * i] = r.x;
        uv[2 * i + 1] = r.y;
    }
}
// Geometric part of Scene.get_surface (scene.slang:322-382) for one triangle and barycentrics
// (b1, b2): out = frame rows T, B, N, face normal, vertex normal (15 floats).
__attribute__((visibility("default"))) void orc_surface_frame(const mrt_vertex* v, uint32_t flags, const float* bary,
                                                              float* out) {
    const f3 b{(1.0f - bary[0]) - bary[1], bary[0], bary[1]};
    const SurfaceGeom g = surface_geometry(v[0], v[1], v[2], flags, b);
    const f3 rows[5] = {g.frame.r0, g.frame.r1, g.frame.r2, g.face_normal, g.vertex_normal};
    for (int r = 0; r < 5; r++) {
        out[3 * r] = rows[r].x;
        out[3 * r + 1] = rows[r].y;
        out[3 * r + 2] = rows[r].z;
    }
}
__attribute__((visibility("default"))) uint16_t orc_f32_to_f16(float f) { return f32_to_f16(f); }
__attribute__((visibility("default"))) float orc_f16_to_f32(uint16_t h) { return f16_to_f32(h); }

// Camera(...) + write_to + advance_jitter(frame) (camera.cpp:10-40,263-266,318-362)
__attribute__((visibility("default"))) void orc_camera_look_at(const float* eye, const float* target, const float* up_in,
                                                               float vfov, float aspect, float near_p, float far_p,
                                                               int32_t jitter, uint32_t frame, mrt_camera* out) {
    const f3 e = ld3(eye), t = ld3(target);
    const f3 up = normalize(ld3(up_in));
    const f3 fwd = normalize(t - e);
    const f3 right = normalize(cross(fwd, up));
    const float focus = length(t - e);
    const float h = std::tan(vfov * 0.5f);
    const f3 U = right * (h * aspect * focus);
    const f3 V = cross(right, fwd) * (h * focus);
    const f3 W = fwd * focus;
    for (int k = 0; k < 3; k++) out->position[k] = eye[k];
    out->U[0] = U.x; out->U[1] = U.y; out->U[2] = U.z;
    out->V[0] = V.x; out->V[1] = V.y; out->V[2] = V.z;
    out->W[0] = W.x; out->W[1] = W.y; out->W[2] = W.z;
    out->near_plane = near_p;
    out->far_plane = far_p;
    out->jitter[0] = out->jitter[1] = 0.f;
    if (jitter == MRT_JITTER_R2) {
        const float n = float(frame % 16u);
        out->jitter[0] = std::fmod(0.5f + (0.7548776662466927f * n), 1.0f) - 0.5f;
        out->jitter[1] = std::fmod(0.5f + (0.5698402909980532f * n), 1.0f) - 0.5f;
    }
}

__attribute__((visibility("default"))) void orc_camera_rays(const mrt_camera* cam, uint32_t w, uint32_t h,
                                                            mrt_ray* out) {
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) {
            const Ray r = camera_ray(*cam, x, y, w, h);
            mrt_ray& o = out[size_t(y) * w + x];
            o.origin[0] = r.o.x; o.origin[1] = r.o.y; o.origin[2] = r.o.z;
            o.direction[0] = r.d.x; o.direction[1] = r.d.y; o.direction[2] = r.d.z;
            o.tmin = r.tmin;
            o.tmax = r.tmax;
        }
}

static Scene* make_scene_from_desc(const mrt_scene_desc* d, int use_bvh) {
    Scene* s = new Scene();
    s->vertices.assign(d->vertices, d->vertices + d->vertex_count);
    s->indices.assign(d->indices, d->indices + d->index_count);
    s->geoms.assign(d->geometries, d->geometries + d->geometry_count);
    s->materials.assign(d->materials, d->materials + d->material_count);
    s->env = d->env;
    if (d->env.type != MRT_ENV_NONE) {
        s->env_rgba.assign(d->env.rgba, d->env.rgba + size_t(d->env.width) * d->env.height * 4);
        s->env.rgba = s->env_rgba.data();
    }
    s->use_bvh = use_bvh != 0;
    s->build();
    return s;
}

__attribute__((visibility("default"))) void* orc_scene_create(const mrt_scene_desc* d, int use_bvh) {
    return make_scene_from_desc(d, use_bvh);
}
// Geometry-only scene (one geometry, all triangles), as mrt_scene_create_triangles.
__attribute__((visibility("default"))) void* orc_scene_create_triangles(const float* pos, uint32_t nv,
                                                                        const uint32_t* idx, uint32_t nt,
                                                                        int two_sided, int use_bvh) {
    std::vector<mrt_vertex> verts(nv);
    for (uint32_t i = 0; i < nv; i++) {
        std::memset(&verts[i], 0, sizeof(mrt_vertex));
        verts[i].position[0] = pos[3 * i];
        verts[i].position[1] = pos[3 * i + 1];
        verts[i].position[2] = pos[3 * i + 2];
    }
    mrt_geometry g{};
    g.vertex_offset = 0;
    g.index_offset = 0;
    g.primitive_count = nt;
    g.material_id = 0;
    g.flags = uint16_t(MRT_GEOM_PRETRANSFORMED | MRT_GEOM_INDEX_UINT32 | (two_sided ? MRT_GEOM_TWO_SIDED : 0));
    g.instance_mask = 0xFF;
    mrt_material m{};
    mrt_scene_desc d{};
    d.vertices = verts.data();
    d.vertex_count = nv;
    d.indices = idx;
    d.index_count = 3 * nt;
    d.geometries = &g;
    d.geometry_count = 1;
    d.materials = &m;
    d.material_count = 1;
    return make_scene_from_desc(&d, use_bvh);
}
__attribute__((visibility("default"))) void orc_scene_destroy(void* s) { delete static_cast<Scene*>(s); }

extern "C++" template <class F> void parallel_for(uint64_t n, int threads, F f) {
    if (threads <= 0) threads = int(std::thread::hardware_concurrency());
    if (threads < 1) threads = 1;
    std::atomic<uint64_t> next{0};
    const uint64_t chunk = 256;
    auto work = [&]() {
        for (;;) {
            const uint64_t b = next.fetch_add(chunk);
            if (b >= n) break;
            const uint64_t e = std::min(n, b + chunk);
            for (uint64_t i = b; i < e; i++) f(i);
        }
    };
    std::vector<std::thread> ts;
    for (int t = 1; t < threads; t++) ts.emplace_back(work);
    work();
    for (auto& t : ts) t.join();
}

__attribute__((visibility("default"))) void orc_trace_closest(void* sp, const mrt_ray* rays, uint64_t n, mrt_hit* hits,
                                                              int threads) {
    const Scene& sc = *static_cast<Scene*>(sp);
    parallel_for(n, threads, [&](uint64_t i) {
        const mrt_ray& r = rays[i];
        const Ray ray{ld3(r.origin), ld3(r.direction), r.tmin, r.tmax};
        const HitRec h = sc.trace(ray, 0xFFu);
        hits[i].triangle = h.id;
        hits[i].t = h.id == 0xFFFFFFFFu ? INFINITY : h.t;
        hits[i].u = h.id == 0xFFFFFFFFu ? 0.f : h.u;
        hits[i].v = h.id == 0xFFFFFFFFu ? 0.f : h.v;
    });
}

// GBuffer (Raytraced): gbuffer_rt.slang:18-94 (hit part only)
__attribute__((visibility("default"))) void orc_gbuffer(void* sp, const mrt_camera* cam, uint32_t w, uint32_t h,
                                                        uint32_t mask, mrt_packed_hit* out, int threads) {
    const Scene& sc = *static_cast<Scene*>(sp);
    parallel_for(uint64_t(w) * h, threads, [&](uint64_t i) {
        const uint32_t x = uint32_t(i % w), y = uint32_t(i / w);
        const Ray r = camera_ray(*cam, x, y, w, h);
        out[i] = pack_hit(to_hit(sc, sc.trace(r, mask)));
    });
}

// The GBuffer's albedo texture (gbuffer.slang:160-162 stores float4(albedo, 1)) for a hit texture.
__attribute__((visibility("default"))) void orc_gbuffer_albedo(void* sp, const mrt_camera* cam, uint32_t w, uint32_t h,
                                                               const mrt_packed_hit* gbuffer, float* out, int threads) {
    const Scene& sc = *static_cast<Scene*>(sp);
    parallel_for(uint64_t(w) * h, threads, [&](uint64_t i) {
        const uint32_t x = uint32_t(i % w), y = uint32_t(i / w);
        const Ray r = camera_ray(*cam, x, y, w, h);
        const ShadingPoint s = get_shading_point(sc, r, unpack_hit(gbuffer[i]), 0u, 1.0f, true);
        const f3 a = gbuffer_albedo(r, s);
        out[4 * i + 0] = a.x;
        out[4 * i + 1] = a.y;
        out[4 * i + 2] = a.z;
        out[4 * i + 3] = 1.f;
    });
}

// Render (Path-traced) main (render_pt.slang:92-130) over the pixel rectangle
// [x0, x0+cw) x [y0, y0+ch) of a w x h image. out: float4 per rectangle pixel.
// sig (optional): per-pixel sum over samples of a hash of each path's hit id sequence.
// Returns the number of continuation segments traced.
__attribute__((visibility("default"))) uint64_t orc_pt_render(void* scp, const mrt_pt_params* prm,
                                                              const mrt_camera* cam, uint32_t frame, uint32_t w,
                                                              uint32_t h, const mrt_packed_hit* gbuffer, uint32_t x0,
                                                              uint32_t y0, uint32_t cw, uint32_t ch, float* out,
                                                              uint32_t* sig, int threads) {
    const Scene& sc = *static_cast<Scene*>(scp);
    std::atomic<uint64_t> total{0};
    parallel_for(uint64_t(cw) * ch, threads, [&](uint64_t i) {
        const uint32_t x = x0 + uint32_t(i % cw), y = y0 + uint32_t(i / cw);
        const Hit primary_hit = unpack_hit(gbuffer[size_t(y) * w + x]);
        const Ray primary_ray = camera_ray(*cam, x, y, w, h);
        const ShadingPoint primary_sp = get_shading_point(sc, primary_ray, primary_hit, 0u, 1.0f, true);
        const f3 demod = prm->demodulate_albedo ? demodulation_albedo(primary_ray, primary_sp) : f3{1.f, 1.f, 1.f};
        f3 irr{0, 0, 0};
        float m2 = 0.f;
        uint32_t psig = 0;
        uint64_t segs = 0;
        const uint32_t pid = primary_hit.valid() ? primary_hit.primitive_id + sc.geom_first[primary_hit.geometry_index]
                                                 : 0xFFFFFFFFu;
        for (int s = 0; s < prm->spp; s++) {
            PCG rng{pcg4d16(x, y, frame, uint32_t(s))};
            uint32_t ps = sig_step(0x9E3779B9u * uint32_t(s + 1), pid);
            const f3 contrib = trace_path(sc, *prm, primary_ray, primary_sp, rng, ps, segs) / demod;
            psig += ps;
            if (finite3(contrib)) {
                irr = irr + contrib;
                const float l = yuv_luminance(contrib);
                m2 += l * l;
            }
        }
        if (prm->spp > 0) {
            irr = irr / float(prm->spp);
            m2 /= float(prm->spp);
        }
        out[4 * i + 0] = irr.x;
        out[4 * i + 1] = irr.y;
        out[4 * i + 2] = irr.z;
        out[4 * i + 3] = m2;
        if (sig) sig[i] = psig;
        total += segs;
    });
    return total.load();
}

// ---- BSDF probes for the statistical invariant tests ----------------------------------
// op: 0 sample_eval, 1 eval, 2 pdf, 3 get_albedo (directional albedo towards wi)
// kind: 0 Lambert(albedo) 1 GGX(alpha.xy) 2 FresnelMix(GGX(alpha.x), Lambert(albedo), f0)
//       3 Conductor(GGX(alpha.x), f0rgb=albedo) 4 RoughDielectric(alpha.x, eta, tint=albedo)
//       5 Sheen(albedo, alpha.x) 6 GGXBTDF(alpha.x, eta, albedo)
//       100 material entry (`mat`), with front_facing / ior of the shading data.
struct orc_bsdf_cfg {
    int32_t kind;
    float albedo[3];
    float alpha[2];
    float f0;
    float eta;
    int32_t front_facing;
    float ior;
    mrt_material mat;
};

extern "C++" template <class B> void bsdf_batch(const B& b, int op, const float* wi, const float* wo, uint32_t seed,
                                          uint64_t n, float* o_wo, float* o_pdf, float* o_w, int32_t* o_valid) {
    for (uint64_t i = 0; i < n; i++) {
        const f3 WI = ld3(wi + 3 * i);
        if (op == 0) {
            PCG rng{pcg4d16(seed, uint32_t(i), uint32_t(i >> 32), 0x5bd1e995u)};
            const Sample s = b.sample_eval(WI, rng);
            o_valid[i] = s.valid;
            o_wo[3 * i] = s.wo.x; o_wo[3 * i + 1] = s.wo.y; o_wo[3 * i + 2] = s.wo.z;
            o_pdf[i] = s.pdf;
            o_w[3 * i] = s.weight.x; o_w[3 * i + 1] = s.weight.y; o_w[3 * i + 2] = s.weight.z;
        } else if (op == 1) {
            const f3 f = b.eval(WI, ld3(wo + 3 * i));
            o_w[3 * i] = f.x; o_w[3 * i + 1] = f.y; o_w[3 * i + 2] = f.z;
        } else if (op == 3) {
            const f3 a = b.get_albedo(WI);
            o_w[3 * i] = a.x; o_w[3 * i + 1] = a.y; o_w[3 * i + 2] = a.z;
        } else {
            o_pdf[i] = b.pdf(WI, ld3(wo + 3 * i));
        }
    }
}

// op: 0 sample_eval(wi) with stream i of `seed`; 1 eval(wi, wo); 2 pdf(wi, wo). wi/wo per item.
__attribute__((visibility("default"))) void orc_bsdf(const orc_bsdf_cfg* c, int op, const float* wi, const float* wo,
                                                     uint32_t seed, uint64_t n, float* o_wo, float* o_pdf, float* o_w,
                                                     int32_t* o_valid) {
    const f3 alb = ld3(c->albedo);
    switch (c->kind) {
    case 0: bsdf_batch(Lambert{alb}, op, wi, wo, seed, n, o_wo, o_pdf, o_w, o_valid); break;
    case 1: bsdf_batch(GGX{{c->alpha[0], c->alpha[1]}}, op, wi, wo, seed, n, o_wo, o_pdf, o_w, o_valid); break;
    case 2:
        bsdf_batch(OpaqueDielectric{GGX{{c->alpha[0], c->alpha[0]}}, Lambert{alb}, c->f0}, op, wi, wo, seed, n, o_wo,
                   o_pdf, o_w, o_valid);
        break;
    case 3: bsdf_batch(Metal{GGX{{c->alpha[0], c->alpha[0]}}, alb}, op, wi, wo, seed, n, o_wo, o_pdf, o_w, o_valid); break;
    case 4: bsdf_batch(RoughDielectric{c->alpha[0], c->eta, alb}, op, wi, wo, seed, n, o_wo, o_pdf, o_w, o_valid); break;
    case 5: bsdf_batch(Sheen::make(alb, c->alpha[0]), op, wi, wo, seed, n, o_wo, o_pdf, o_w, o_valid); break;
    case 6: bsdf_batch(GGXBTDF{c->alpha[0], c->eta, alb}, op, wi, wo, seed, n, o_wo, o_pdf, o_w, o_valid); break;
    default: {
        ShadingData sd{};
        sd.frame = frame_from_normal(f3{0, 0, 1});
        sd.front_facing = c->front_facing != 0;
        sd.ior = c->ior;
        f3 em;
        Frame sf;
        const MaterialSample ms = get_material_sample(c->mat, sd, em, sf);
        bsdf_batch(ms, op, wi, wo, seed, n, o_wo, o_pdf, o_w, o_valid);
    }
    }
}

} // extern "C"

#include "oracle_mcpg.h"
