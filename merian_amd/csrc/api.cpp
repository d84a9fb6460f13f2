// C ABI of include/merian_amd.h: contexts with an HBM budget, device-resident scenes (BVH build
// + upload), ray queries, the GBuffer pass and one frame of pt.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/merian_amd.h"
#include "bvh.h"
#include "mrt_kernels.h"
#include "mrt_api_internal.h"

using namespace mrt;

// ------------------------------------------------------------------------------------------
// context

namespace mrt {

mrt_status fail(mrt_ctx* c, mrt_status s, const std::string& msg) {
    if (c) c->err = msg;
    return s;
}

mrt_status hip_fail(mrt_ctx* c, hipError_t e, const char* what) {
    return fail(c, MRT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

void* dev_alloc(mrt_ctx* c, size_t bytes, mrt_status* st) {
    if (bytes == 0) bytes = 16;
    if (c->in_use + bytes > c->budget) {
        *st = fail(c, MRT_ERR_BUDGET,
                   "HBM budget exceeded: " + std::to_string(c->in_use + bytes) + " > " + std::to_string(c->budget));
        return nullptr;
    }
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        *st = fail(c, MRT_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
        return nullptr;
    }
    c->in_use += bytes;
    c->sizes[p] = bytes;
    *st = MRT_OK;
    return p;
}

void dev_free(mrt_ctx* c, void* p) {
    if (!p) return;
    auto it = c->sizes.find(p);
    if (it != c->sizes.end()) {
        c->in_use -= it->second;
        c->sizes.erase(it);
    }
    const hipError_t e = hipFree(p);
    if (e != hipSuccess) c->err = std::string("hipFree: ") + hipGetErrorString(e);
}

// NULL selects the context's own stream; hipStreamLegacy selects the HIP null stream (the
// default stream of most callers, torch's included, whose handle is 0 and so cannot be passed
// as itself).
hipStream_t pick_stream(mrt_ctx* c, void* stream) {
    if (!stream) return c->stream;
    if (static_cast<hipStream_t>(stream) == hipStreamLegacy) return hipStream_t(nullptr);
    return static_cast<hipStream_t>(stream);
}

void free_path_buffers(mrt_ctx* c) {
    PathBuffers& p = c->paths;
    for (void* q : {(void*)p.state, (void*)p.sig, (void*)p.q_o[0], (void*)p.q_o[1], (void*)p.q_d[0],
                    (void*)p.q_d[1], (void*)p.q_t[0], (void*)p.q_t[1], (void*)p.q_r[0], (void*)p.q_r[1], (void*)p.hits, (void*)p.sq_o, (void*)p.sq_d, (void*)p.sq_c,
                    (void*)p.counters, (void*)p.accum, (void*)p.demod, (void*)p.sig_accum})
        dev_free(c, q);
    p = PathBuffers{};
    c->paths_npix = 0;
    c->paths_counters = 0;
}

// Path buffers for `paths` paths, `npix` pixels and `counters` queue counters.
mrt_status ensure_path_buffers(mrt_ctx* c, uint64_t paths, uint64_t npix, uint64_t counters) {
    PathBuffers& p = c->paths;
    if (p.capacity >= paths && c->paths_npix >= npix && c->paths_counters >= counters) return MRT_OK;
    paths = std::max<uint64_t>(paths, p.capacity);
    npix = std::max<uint64_t>(npix, c->paths_npix);
    counters = std::max<uint64_t>(counters, c->paths_counters);
    free_path_buffers(c);
    mrt_status st = MRT_OK;
    auto A = [&](size_t bytes) { return st == MRT_OK ? dev_alloc(c, bytes, &st) : nullptr; };
    p.state = static_cast<float4*>(A(32 * paths));
    p.sig = static_cast<uint32_t*>(A(4 * paths));
    for (int k = 0; k < 2; k++) {
        p.q_o[k] = static_cast<float4*>(A(16 * paths));
        p.q_d[k] = static_cast<float4*>(A(16 * paths));
        p.q_t[k] = static_cast<float4*>(A(16 * paths));
        p.q_r[k] = static_cast<float4*>(A(16 * paths));
    }
    p.hits = static_cast<float4*>(A(16 * paths));
    p.sq_o = static_cast<float4*>(A(16 * paths));
    p.sq_d = static_cast<float4*>(A(16 * paths));
    p.sq_c = static_cast<float4*>(A(16 * paths));
    // per chunk and bounce: queue sizes, trace cursors, shadow-queue sizes, shadow cursors
    p.counters = static_cast<uint32_t*>(A(4 * 4 * counters));
    p.accum = static_cast<float4*>(A(16 * npix));
    p.demod = static_cast<float4*>(A(16 * npix));
    p.sig_accum = static_cast<uint32_t*>(A(4 * npix));
    if (st != MRT_OK) {
        const std::string keep = c->err;
        free_path_buffers(c);
        c->err = keep;
        return st;
    }
    p.capacity = paths;
    c->paths_npix = npix;
    c->paths_counters = counters;
    return MRT_OK;
}

// Global part of the traversal stacks, sized for the deepest scene traced so far.
mrt_status ensure_trace_scratch(mrt_ctx* c, const mrt_scene* s) {
    const uint32_t need = s->ds.stack_need > uint32_t(TRACE_LDS_STACK) ? s->ds.stack_need - TRACE_LDS_STACK : 1u;
    TraceScratch& t = c->trace;
    if (t.work && t.entries >= need) return MRT_OK;
    if (t.ovf) dev_free(c, t.ovf);
    if (t.work) dev_free(c, t.work);
    t = TraceScratch{};
    const uint32_t threads = trace_max_threads();
    mrt_status st = MRT_OK;
    t.ovf = static_cast<uint32_t*>(dev_alloc(c, size_t(threads) * need * 4, &st));
    if (st != MRT_OK) return st;
    t.work = static_cast<uint32_t*>(dev_alloc(c, 64, &st));
    if (st != MRT_OK) {
        dev_free(c, t.ovf);
        t = TraceScratch{};
        return st;
    }
    t.threads = threads;
    t.entries = need;
    return MRT_OK;
}

TraceScratch scene_scratch(const mrt_ctx* c, const mrt_scene* s) {
    TraceScratch t = c ? c->trace : TraceScratch{};
    t.refill = s->tuning.refill_lanes >= 1 && s->tuning.refill_lanes <= 64 ? s->tuning.refill_lanes : 16u;
    t.grid_pct = s->tuning.grid_percent >= 1 && s->tuning.grid_percent <= 100 ? s->tuning.grid_percent : 100u;
    t.leaf_lanes = s->tuning.leaf_lanes >= 1 && s->tuning.leaf_lanes <= 64 ? s->tuning.leaf_lanes : 16u;
    return t;
}

static bool capturing(hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

hipError_t scratch_begin(mrt_ctx* c, hipStream_t st) {
    if (!c->scratch_event || c->scratch_stream == st || capturing(st)) return hipSuccess;
    return hipStreamWaitEvent(st, c->scratch_event, 0);
}

hipError_t scratch_end(mrt_ctx* c, hipStream_t st) {
    if (capturing(st)) return hipSuccess;
    if (!c->scratch_event) {
        const hipError_t e = hipEventCreateWithFlags(&c->scratch_event, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    c->scratch_stream = st;
    return hipEventRecord(c->scratch_event, st);
}

} // namespace mrt

extern "C" {

#ifndef MRT_SOURCE_HASH
#define MRT_SOURCE_HASH "unknown"
#endif
// The build script passes a hash of the library's sources; loaders compare it with the sources
// beside the library to detect a stale build.
MRT_API const char* mrt_version(void) { return "merian_amd 0.2 (gfx950, HIP) sources " MRT_SOURCE_HASH; }

MRT_API uint32_t merian_node_plugin_abi_version(void) { return 1u; }

MRT_API mrt_status mrt_create(int32_t device, uint64_t hbm_budget, mrt_ctx** out) {
    if (!out) return MRT_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return MRT_ERR_DEVICE;
    if (device < 0 || device >= n) return MRT_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return MRT_ERR_DEVICE;
    mrt_ctx* c = new (std::nothrow) mrt_ctx();
    if (!c) return MRT_ERR_ALLOC;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return MRT_ERR_DEVICE;
    }
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return MRT_ERR_DEVICE;
    }
    c->budget = hbm_budget ? hbm_budget : uint64_t(double(free_b) * 0.9);
    if (noalloc_setup(c) != MRT_OK) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return MRT_ERR_DEVICE;
    }
    const char* mp = std::getenv("MRT_MAX_PATHS");
    c->max_paths = mp ? std::strtoull(mp, nullptr, 10) : (uint64_t(1) << 24);
    *out = c;
    return MRT_OK;
}

MRT_API mrt_status mrt_destroy(mrt_ctx* c) {
    if (!c) return MRT_ERR_INVALID;
    // teardown is best effort: every resource is released even if one step fails
    const bool ok = hipSetDevice(c->device) == hipSuccess && hipStreamSynchronize(c->stream) == hipSuccess;
    free_path_buffers(c);
    for (auto& kv : c->sizes) (void)hipFree(kv.first);
    c->sizes.clear();
    if (c->scratch_event) (void)hipEventDestroy(c->scratch_event);
    const bool destroyed = hipStreamDestroy(c->stream) == hipSuccess;
    delete c;
    return ok && destroyed ? MRT_OK : MRT_ERR_DEVICE;
}

MRT_API const char* mrt_last_error(mrt_ctx* c) { return c ? c->err.c_str() : "null context"; }

MRT_API uint64_t mrt_memory_in_use(mrt_ctx* c) { return c ? c->in_use : 0; }

MRT_API mrt_status mrt_alloc_dev(mrt_ctx* c, uint64_t bytes, void** d_ptr) {
    if (!c || !d_ptr) return MRT_ERR_INVALID;
    const hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(c, e, "mrt_alloc_dev");
    mrt_status st = MRT_OK;
    *d_ptr = dev_alloc(c, bytes, &st);
    return st;
}

MRT_API mrt_status mrt_free_dev(mrt_ctx* c, void* d_ptr) {
    if (!c) return MRT_ERR_INVALID;
    if (d_ptr && !c->sizes.count(d_ptr)) return fail(c, MRT_ERR_INVALID, "mrt_free_dev: not a context allocation");
    const hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "mrt_free_dev");
    dev_free(c, d_ptr);
    return MRT_OK;
}

MRT_API mrt_status mrt_download(mrt_ctx* c, void* host, const void* d_src, uint64_t bytes) {
    if (!c || (bytes && (!host || !d_src))) return MRT_ERR_INVALID;
    hipError_t e = hipMemcpyAsync(host, d_src, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, "mrt_download");
}

MRT_API mrt_status mrt_upload(mrt_ctx* c, void* d_dst, const void* host, uint64_t bytes) {
    if (!c || (bytes && (!host || !d_dst))) return MRT_ERR_INVALID;
    hipError_t e = hipMemcpyAsync(d_dst, host, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, "mrt_upload");
}

MRT_API mrt_status mrt_accumulate_dev(mrt_ctx* c, float* d_accum, const float* d_src, uint32_t pixels, uint32_t count,
                                      void* stream) {
    if (!c || (pixels && (!d_accum || !d_src))) return fail(c, MRT_ERR_INVALID, "mrt_accumulate_dev: null argument");
    const hipError_t e = launch_running_mean(reinterpret_cast<float4*>(d_accum), reinterpret_cast<const float4*>(d_src),
                                             pixels, count, pick_stream(c, stream));
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, "accumulate kernel");
}

MRT_API mrt_status mrt_synchronize(mrt_ctx* c) {
    if (!c) return MRT_ERR_INVALID;
    const hipError_t e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, "hipStreamSynchronize");
}

// ------------------------------------------------------------------------------------------
// scenes

static mrt_status upload(mrt_ctx* c, mrt_scene* s, const void* src, size_t bytes, void** dst) {
    mrt_status st = MRT_OK;
    *dst = dev_alloc(c, bytes, &st);
    if (st != MRT_OK) return st;
    s->allocs.push_back(*dst);
    s->bytes += bytes;
    if (bytes && src) {
        const hipError_t e = hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy (scene upload)");
    }
    return MRT_OK;
}

// Importance-sampling table of the environment for NEE: per texel (prob, alias, p, 0) of a
// Vose alias table over weights = mean(rgb) x solid-angle factor (sin(theta) at the texel centre
// for the lat-long map; constant for the equal-area map), p = normalised weight.
static std::vector<float> build_env_sampling(const mrt_env& e) {
    const size_t n = size_t(e.width) * e.height;
    std::vector<double> w(n);
    double total = 0.0;
    for (size_t i = 0; i < n; i++) {
        const float* t = e.rgba + 4 * i;
        double v = std::max(0.0, (double(t[0]) + double(t[1]) + double(t[2])) / 3.0);
        if (!std::isfinite(v)) v = 0.0;
        if (e.type == MRT_ENV_LATLONG) v *= std::sin((double(i / e.width) + 0.5) / double(e.height) * M_PI);
        w[i] = v;
        total += v;
    }
    std::vector<float> out(4 * n, 0.f);
    if (!(total > 0.0)) {
        for (size_t i = 0; i < n; i++) {
            out[4 * i] = 1.f;
            uint32_t a = uint32_t(i);
            std::memcpy(&out[4 * i + 1], &a, 4);
        }
        return out;
    }
    std::vector<double> q(n);
    std::vector<uint32_t> small, large, alias(n);
    std::vector<double> prob(n, 1.0);
    for (size_t i = 0; i < n; i++) {
        q[i] = w[i] / total * double(n);
        (q[i] < 1.0 ? small : large).push_back(uint32_t(i));
        alias[i] = uint32_t(i);
    }
    while (!small.empty() && !large.empty()) {
        const uint32_t l = small.back(), g = large.back();
        small.pop_back();
        prob[l] = q[l];
        alias[l] = g;
        q[g] = (q[g] + q[l]) - 1.0;
        if (q[g] < 1.0) {
            large.pop_back();
            small.push_back(g);
        }
    }
    for (size_t i = 0; i < n; i++) {
        out[4 * i] = float(prob[i]);
        std::memcpy(&out[4 * i + 1], &alias[i], 4);
        out[4 * i + 2] = float(w[i] / total);
    }
    return out;
}

static void scene_free(mrt_ctx* c, mrt_scene* s) {
    if (!s) return;
    for (void* p : s->allocs) dev_free(c, p);
    delete s;
}

MRT_API mrt_status mrt_scene_create(mrt_ctx* c, const mrt_scene_desc* d, mrt_scene** out) {
    if (!c || !d || !out) return fail(c, MRT_ERR_INVALID, "mrt_scene_create: null argument");
    *out = nullptr;
    if (const hipError_t e = hipSetDevice(c->device); e != hipSuccess) return hip_fail(c, e, "mrt_scene_create");
    if (d->geometry_count > 65535u) return fail(c, MRT_ERR_UNSUPPORTED, "more than 65535 geometries (u16 ids)");
    if (d->material_count == 0 && d->geometry_count) return fail(c, MRT_ERR_INVALID, "no materials");
    // flatten triangles in geometry order: global id = geometry prefix + primitive
    uint64_t n_tris = 0;
    for (uint32_t g = 0; g < d->geometry_count; g++) n_tris += d->geometries[g].primitive_count;
    if (n_tris >= (uint64_t(1) << 28)) return fail(c, MRT_ERR_UNSUPPORTED, "more than 2^28 triangles");
    std::vector<float> v0(3 * n_tris), e1(3 * n_tris), e2(3 * n_tris);
    std::vector<uint32_t> flags(n_tris), geom_first(std::max<uint32_t>(1, d->geometry_count));
    std::vector<uint32_t> tri_verts(4 * n_tris);
    std::vector<float> tri_shade(16 * std::max<uint64_t>(1, n_tris), 0.f);
    uint64_t gid = 0;
    for (uint32_t g = 0; g < d->geometry_count; g++) {
        const mrt_geometry& G = d->geometries[g];
        if (G.material_id >= d->material_count) return fail(c, MRT_ERR_INVALID, "material id out of range");
        geom_first[g] = uint32_t(gid);
        const mrt_material& M = d->materials[G.material_id];
        uint8_t cls = SHADE_CLASS_DIFFUSE;
        if (G.flags & MRT_GEOM_USE_ENV_MAP) {
            cls = SHADE_CLASS_ENV;
        } else if (M.model != MRT_MAT_DIFFUSE) {
            cls = (M.features & MRT_PBR_TRANSMISSION) ? SHADE_CLASS_TRANSMISSIVE
                  : M.metalness >= 0.5f              ? SHADE_CLASS_METAL
                                                     : SHADE_CLASS_DIELECTRIC;
        }
        const uint32_t geom_word = g | (uint32_t(cls) << TRI_SHADE_CLASS_SHIFT);
        for (uint32_t p = 0; p < G.primitive_count; p++, gid++) {
            uint32_t vi[3];
            for (int k = 0; k < 3; k++) {
                uint32_t local;
                if (G.flags & MRT_GEOM_INDEX_UINT32) {
                    const uint64_t ii = uint64_t(G.index_offset) + 3ull * p + k;
                    if (ii >= d->index_count) return fail(c, MRT_ERR_INVALID, "index buffer overrun");
                    local = d->indices[ii];
                } else {
                    local = 3u * p + uint32_t(k);
                }
                const uint64_t v = uint64_t(G.vertex_offset) + local;
                if (v >= d->vertex_count) return fail(c, MRT_ERR_INVALID, "vertex index out of range");
                vi[k] = uint32_t(v);
            }
            const float* a = d->vertices[vi[0]].position;
            const float* b = d->vertices[vi[1]].position;
            const float* cc = d->vertices[vi[2]].position;
            for (int k = 0; k < 3; k++) {
                v0[3 * gid + k] = a[k];
                e1[3 * gid + k] = b[k] - a[k];
                e2[3 * gid + k] = cc[k] - a[k];
            }
            flags[gid] = ((G.flags & MRT_GEOM_TWO_SIDED) ? 0u : 1u) | ((G.instance_mask & 0xFFu) << 8);
            tri_verts[4 * gid + 0] = vi[0];
            tri_verts[4 * gid + 1] = vi[1];
            tri_verts[4 * gid + 2] = vi[2];
            tri_verts[4 * gid + 3] = g;
            float* ts = &tri_shade[16 * gid];
            for (int k = 0; k < 3; k++) {
                const mrt_vertex& v = d->vertices[vi[k]];
                std::memcpy(ts + 4 * k, v.position, 12);
                std::memcpy(ts + 4 * k + 3, &v.encoded_normal, 4);
                const uint32_t uv = uint32_t(v.uv[0]) | (uint32_t(v.uv[1]) << 16);
                std::memcpy(ts + 12 + k, &uv, 4);
            }
            std::memcpy(ts + 15, &geom_word, 4);
        }
    }
    const mrt_tuning defaults = mrt_scene{}.tuning;
    const BVH4 bvh = build_bvh4(v0.data(), e1.data(), e2.data(), flags.data(), uint32_t(n_tris), 0,
                                int(defaults.bvh_max_leaf), defaults.bvh_traversal_cost);
    mrt_scene* s = new mrt_scene();
    if (n_tris) {
        for (int k = 0; k < 3; k++) {
            s->bbox_min[k] = v0[k];
            s->bbox_max[k] = v0[k];
        }
        for (uint64_t t = 0; t < n_tris; t++)
            for (int k = 0; k < 3; k++) {
                const float a = v0[3 * t + k], b = a + e1[3 * t + k], cc = a + e2[3 * t + k];
                s->bbox_min[k] = std::min(s->bbox_min[k], std::min(a, std::min(b, cc)));
                s->bbox_max[k] = std::max(s->bbox_max[k], std::max(a, std::max(b, cc)));
            }
    }
    s->ds.stack_need = 3u * bvh.depth + 4u;
    s->ds.width = 4;
    s->ds.nodes8 = nullptr;
    s->n_tris = n_tris;
    s->n_nodes = bvh.nodes.size();
    s->depth = bvh.depth;
    mrt_status st = MRT_OK;
    void* p = nullptr;
#define UP(SRC, BYTES, FIELD, T)                                     \
    if (st == MRT_OK) {                                              \
        st = upload(c, s, SRC, BYTES, &p);                           \
        s->ds.FIELD = static_cast<T>(p);                             \
    }
    UP(bvh.nodes.data(), bvh.nodes.size() * sizeof(BVH4Node), nodes, const BVH4Node*)
    UP(bvh.tris.data(), bvh.tris.size() * sizeof(TriRecord), tris, const float4*)
    UP(tri_verts.data(), tri_verts.size() * 4, tri_verts, const uint4*)
    UP(tri_shade.data(), tri_shade.size() * 4, tri_shade, const float4*)
    UP(d->vertices, size_t(d->vertex_count) * sizeof(mrt_vertex), verts, const mrt_vertex*)
    UP(d->geometries, size_t(d->geometry_count) * sizeof(mrt_geometry), geoms, const mrt_geometry*)
    UP(geom_first.data(), geom_first.size() * 4, geom_first, const uint32_t*)
    UP(d->materials, size_t(d->material_count) * sizeof(mrt_material), mats, const mrt_material*)
    const mrt_env& e = d->env;
    s->ds.env_type = e.type;
    s->ds.env_wrap = e.wrap;
    s->ds.env_w = e.width;
    s->ds.env_h = e.height;
    s->ds.env_intensity = e.intensity;
    std::memcpy(s->ds.env_to_local, e.to_local, sizeof(e.to_local));
    if (e.type != MRT_ENV_NONE) {
        if (!e.rgba || e.width == 0 || e.height == 0) st = fail(c, MRT_ERR_INVALID, "environment without texels");
        UP(e.rgba, size_t(e.width) * e.height * 16, env, const float4*)
        if (st == MRT_OK) {
            const std::vector<float> samp = build_env_sampling(e);
            UP(samp.data(), samp.size() * 4, env_samp, const float4*)
        }
    } else {
        s->ds.env = nullptr;
        s->ds.env_samp = nullptr;
    }
#undef UP
    s->ds.n_tris = uint32_t(n_tris);
    s->ds.n_geoms = d->geometry_count;
    if (st == MRT_OK) {
        const hipError_t he = hipStreamSynchronize(c->stream);
        if (he != hipSuccess) st = hip_fail(c, he, "scene upload");
    }
    if (st != MRT_OK) {
        scene_free(c, s);
        return st;
    }
    s->h_v0 = std::move(v0);
    s->h_e1 = std::move(e1);
    s->h_e2 = std::move(e2);
    s->h_flags = std::move(flags);
    *out = s;
    return MRT_OK;
}

MRT_API mrt_status mrt_scene_create_triangles(mrt_ctx* c, const float* positions, uint32_t vertex_count,
                                              const uint32_t* indices, uint32_t triangle_count, int32_t two_sided,
                                              mrt_scene** out) {
    if (!c || !positions || !indices || !out) return fail(c, MRT_ERR_INVALID, "null argument");
    std::vector<mrt_vertex> verts(vertex_count);
    for (uint32_t i = 0; i < vertex_count; i++) {
        std::memset(&verts[i], 0, sizeof(mrt_vertex));
        std::memcpy(verts[i].position, positions + 3ull * i, 12);
    }
    mrt_geometry g{};
    g.primitive_count = triangle_count;
    g.flags = uint16_t(MRT_GEOM_PRETRANSFORMED | MRT_GEOM_INDEX_UINT32 | (two_sided ? MRT_GEOM_TWO_SIDED : 0));
    g.instance_mask = 0xFFu;
    mrt_material m{};
    mrt_scene_desc d{};
    d.vertices = verts.data();
    d.vertex_count = vertex_count;
    d.indices = indices;
    d.index_count = 3u * triangle_count;
    d.geometries = &g;
    d.geometry_count = 1;
    d.materials = &m;
    d.material_count = 1;
    d.env.type = MRT_ENV_NONE;
    return mrt_scene_create(c, &d, out);
}

MRT_API mrt_status mrt_scene_destroy(mrt_ctx* c, mrt_scene* s) {
    if (!c || !s) return MRT_ERR_INVALID;
    const hipError_t e = hipStreamSynchronize(c->stream); // kernels may still read the scene
    scene_free(c, s);
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, "mrt_scene_destroy");
}

MRT_API mrt_status mrt_scene_info(mrt_scene* s, uint64_t* tris, uint64_t* nodes, uint64_t* bytes) {
    if (!s) return MRT_ERR_INVALID;
    if (tris) *tris = s->n_tris;
    if (nodes) *nodes = s->n_nodes;
    if (bytes) *bytes = s->bytes;
    return MRT_OK;
}

// Camera(...), recompute_ray_basis, write_to and advance_jitter
// (src/merian/utils/camera/camera.cpp:10-40, 263-266, 318-362).
MRT_API mrt_status mrt_camera_look_at(const float eye[3], const float target[3], const float up_in[3], float vfov,
                                      float aspect, float near_plane, float far_plane, int32_t jitter, uint32_t frame,
                                      mrt_camera* out) {
    if (!eye || !target || !up_in || !out) return MRT_ERR_INVALID;
    auto norm = [](const float v[3], float o[3]) {
        const float inv = 1.0f / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int k = 0; k < 3; k++) o[k] = v[k] * inv;
    };
    auto cross = [](const float a[3], const float b[3], float o[3]) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    float up[3], diff[3], fwd[3], rc[3], right[3], vdir[3];
    norm(up_in, up);
    for (int k = 0; k < 3; k++) diff[k] = target[k] - eye[k];
    norm(diff, fwd);
    cross(fwd, up, rc);
    norm(rc, right);
    const float focus = std::sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
    const float h = std::tan(vfov * 0.5f);
    cross(right, fwd, vdir);
    const float su = h * aspect * focus, sv = h * focus;
    for (int k = 0; k < 3; k++) {
        out->position[k] = eye[k];
        out->U[k] = right[k] * su;
        out->V[k] = vdir[k] * sv;
        out->W[k] = fwd[k] * focus;
    }
    out->near_plane = near_plane;
    out->far_plane = far_plane;
    out->jitter[0] = out->jitter[1] = 0.f;
    if (jitter == MRT_JITTER_R2) {
        const float n = float(frame % 16u);
        out->jitter[0] = std::fmod(0.5f + (0.7548776662466927f * n), 1.0f) - 0.5f;
        out->jitter[1] = std::fmod(0.5f + (0.5698402909980532f * n), 1.0f) - 0.5f;
    } else if (jitter != MRT_JITTER_NONE) {
        return MRT_ERR_INVALID;
    }
    return MRT_OK;
}

// ------------------------------------------------------------------------------------------
// ray queries, gbuffer, pt

MRT_API mrt_status mrt_trace_closest_dev(mrt_ctx* c, mrt_scene* s, const mrt_ray* rays, uint64_t n, mrt_hit* hits,
                                         void* stream) {
    if (!c || !s || (n && (!rays || !hits))) return fail(c, MRT_ERR_INVALID, "mrt_trace_closest_dev: bad argument");
    mrt_status st = ensure_trace_scratch(c, s);
    if (st != MRT_OK) return st;
    const hipStream_t strm = pick_stream(c, stream);
    hipError_t e = scratch_begin(c, strm);
    if (e == hipSuccess) e = launch_trace_rays(s->ds, scene_scratch(c, s), rays, n, hits, strm);
    if (e == hipSuccess) e = scratch_end(c, strm);
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, "trace kernel");
}

MRT_API mrt_status mrt_trace_count_dev(mrt_ctx* c, mrt_scene* s, const mrt_ray* rays, uint64_t n, mrt_hit* hits,
                                       uint64_t counts[2]) {
    if (!c || !s || !counts || (n && (!rays || !hits))) return fail(c, MRT_ERR_INVALID, "mrt_trace_count_dev: bad argument");
    mrt_status st = ensure_trace_scratch(c, s);
    if (st != MRT_OK) return st;
    void* dc = dev_alloc(c, 16, &st);
    if (st != MRT_OK) return st;
    hipError_t e = scratch_begin(c, c->stream);
    if (e == hipSuccess) e = launch_zero_u32(static_cast<uint32_t*>(dc), 4, c->stream);
    if (e == hipSuccess)
        e = launch_trace_rays(s->ds, scene_scratch(c, s), rays, n, hits, c->stream, static_cast<unsigned long long*>(dc));
    if (e == hipSuccess) e = scratch_end(c, c->stream);
    unsigned long long h[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(h, dc, 16, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dev_free(c, dc);
    if (e != hipSuccess) return hip_fail(c, e, "mrt_trace_count_dev");
    counts[0] = h[0];
    counts[1] = h[1];
    return MRT_OK;
}

MRT_API mrt_status mrt_trace_closest(mrt_ctx* c, mrt_scene* s, const mrt_ray* rays, uint64_t n, mrt_hit* hits) {
    if (!c || !s || (n && (!rays || !hits))) return fail(c, MRT_ERR_INVALID, "mrt_trace_closest: bad argument");
    if (n == 0) return MRT_OK;
    mrt_status st = ensure_trace_scratch(c, s);
    if (st != MRT_OK) return st;
    void* dr = dev_alloc(c, n * sizeof(mrt_ray), &st);
    if (st != MRT_OK) return st;
    void* dh = dev_alloc(c, n * sizeof(mrt_hit), &st);
    if (st != MRT_OK) {
        dev_free(c, dr);
        return st;
    }
    hipError_t e = scratch_begin(c, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dr, rays, n * sizeof(mrt_ray), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = launch_trace_rays(s->ds, scene_scratch(c, s), static_cast<mrt_ray*>(dr), n, static_cast<mrt_hit*>(dh),
                              c->stream);
    if (e == hipSuccess) e = scratch_end(c, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hits, dh, n * sizeof(mrt_hit), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dev_free(c, dr);
    dev_free(c, dh);
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, "mrt_trace_closest");
}

MRT_API mrt_status mrt_camera_rays_dev(mrt_ctx* c, const mrt_camera* cam, uint32_t w, uint32_t h, mrt_ray* rays,
                                       void* stream) {
    if (!c || !cam || (uint64_t(w) * h && !rays)) return fail(c, MRT_ERR_INVALID, "mrt_camera_rays_dev: bad argument");
    const hipError_t e = launch_camera_rays(*cam, w, h, rays, pick_stream(c, stream));
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, "camera ray kernel");
}

MRT_API mrt_status mrt_gbuffer_dev(mrt_ctx* c, mrt_scene* s, const mrt_camera* cam, uint32_t w, uint32_t h,
                                   uint32_t mask, mrt_packed_hit* out, void* stream) {
    if (!c || !s || !cam || (uint64_t(w) * h && !out)) return fail(c, MRT_ERR_INVALID, "mrt_gbuffer_dev: bad argument");
    const mrt_status st = ensure_trace_scratch(c, s);
    if (st != MRT_OK) return st;
    const hipStream_t strm = pick_stream(c, stream);
    hipError_t e = scratch_begin(c, strm);
    if (e == hipSuccess) e = launch_gbuffer(s->ds, scene_scratch(c, s), *cam, w, h, mask & 0xFFu, out, strm);
    if (e == hipSuccess) e = scratch_end(c, strm);
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, "gbuffer kernel");
}

MRT_API mrt_status mrt_gbuffer_albedo_dev(mrt_ctx* c, mrt_scene* s, const mrt_camera* cam, uint32_t w, uint32_t h,
                                          const mrt_packed_hit* gbuffer, float* albedo, void* stream) {
    if (!c || !s || !cam || (uint64_t(w) * h && (!gbuffer || !albedo)))
        return fail(c, MRT_ERR_INVALID, "mrt_gbuffer_albedo_dev: bad argument");
    const hipError_t e = launch_gbuffer_albedo(s->ds, *cam, w, h, gbuffer, reinterpret_cast<float4*>(albedo), 0,
                                               pick_stream(c, stream));
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, "gbuffer albedo kernel");
}

} // extern "C"

namespace mrt {
mrt_status check_params(mrt_ctx* c, const mrt_pt_params* p) {
    if (!p) return fail(c, MRT_ERR_INVALID, "null params");
    if (p->spp < 0) return fail(c, MRT_ERR_INVALID, "samples per pixel < 0");
    if (p->max_path_length < 1 || p->max_path_length > 64)
        return fail(c, MRT_ERR_INVALID, "max path length outside [1, 64]");
    if (p->mode != MRT_MODE_BSDF && p->mode != MRT_MODE_NEE)
        return fail(c, MRT_ERR_UNSUPPORTED, "render mode not supported by this entry point");
    return MRT_OK;
}
} // namespace mrt

namespace mrt {
mrt_status render_frame_dev(mrt_ctx* c, mrt_scene* s, const mrt_pt_params* prm, const mrt_camera* cam, uint32_t frame,
                            uint32_t w, uint32_t h, const mrt_packed_hit* gbuffer, float* irradiance,
                            uint32_t* signature, mrt_pt_stats* stats, void* stream, mrt_guide* guide,
                            mrt_mcpg* mcpg) {
    if (!c || !s || !cam) return fail(c, MRT_ERR_INVALID, "pt frame: null argument");
    mrt_status st = MRT_OK;
    const uint64_t npix = uint64_t(w) * h;
    if (npix == 0) return MRT_OK;
    if (!gbuffer || !irradiance) return fail(c, MRT_ERR_INVALID, "mrt_pt_render_dev: null buffer");
    if (npix > 0xFFFFFFFFull) return fail(c, MRT_ERR_UNSUPPORTED, "image too large");
    if (const hipError_t e = hipSetDevice(c->device); e != hipSuccess) return hip_fail(c, e, "pt frame");
    const uint64_t spp = uint64_t(prm->spp);
    const uint64_t max_paths = s->tuning.max_paths ? s->tuning.max_paths : c->max_paths;
    uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(spp ? spp : 1, max_paths / npix));
    const uint64_t path_limit = prm->mode == MRT_MODE_NEE && !mcpg ? NEE_MAX_PATHS : 0xFFFFFFFFull;
    if (npix * chunk > path_limit) chunk = path_limit / npix;
    if (chunk == 0) return fail(c, MRT_ERR_UNSUPPORTED, "image too large for one sample per pass");
    if (mcpg) chunk = 1; // the samples of a pixel run one after another
    const uint64_t n_chunks = spp ? (spp + chunk - 1) / chunk : 1;
    st = ensure_path_buffers(c, npix * chunk, npix, n_chunks * uint64_t(prm->max_path_length + 1));
    if (st != MRT_OK) return st;
    st = ensure_trace_scratch(c, s);
    if (st != MRT_OK) return st;
    // capacity may exceed this frame's need; the frame uses exactly `chunk` samples per pass
    PathBuffers pb = c->paths;
    pb.capacity = npix * chunk;
    PtArgs a{};
    a.cam = *cam;
    a.width = w;
    a.height = h;
    a.npix = uint32_t(npix);
    a.frame = frame;
    a.spp = prm->spp;
    a.max_path_length = prm->max_path_length;
    a.emission_on_primary = prm->emission_on_primary;
    a.demodulate = prm->demodulate_albedo ? 1 : 0;
    a.record_signature = prm->record_signature && signature ? 1 : 0;
    a.instance_mask = prm->instance_mask & 0xFFu;
    a.mode = mcpg ? int32_t(MRT_MODE_MCPG) : prm->mode;
    McpgArgs margs{};
    if (mcpg) {
        st = mcpg_prepare_frame(c, mcpg, npix, *cam, margs);
        if (st != MRT_OK) return st;
    }
    const GuideNet* net = nullptr;
    if (guide) {
        st = guide_prepare_frame(c, guide, npix * chunk, pb, a.guide_fraction, a.guided_bounces, &net);
        if (st != MRT_OK) return st;
        // one record per guided vertex of the first pass's sample-0 paths: when more would
        // arrive than the buffer holds, keep a pseudo-random 2^-shift of the pixels rather than
        // the first pixels of the queue (which would be the top of the image)
        const uint64_t verts = std::min<uint64_t>(a.guided_bounces, uint64_t(prm->max_path_length - 1));
        uint32_t shift = 0;
        while (shift < 31 && ((npix * verts) >> shift) > uint64_t(pb.grec_cap)) shift++;
        a.guide_record_shift = shift;
    }
    FrameTiming t;
    uint64_t segs = 0;
    const hipStream_t strm = pick_stream(c, stream);
    unsigned long long* tc = nullptr;
    if (prm->count_traversal && stats) {
        tc = static_cast<unsigned long long*>(dev_alloc(c, 16, &st));
        if (st != MRT_OK) return st;
        if (const hipError_t e = launch_zero_u32(reinterpret_cast<uint32_t*>(tc), 4, strm); e != hipSuccess) {
            dev_free(c, tc);
            return hip_fail(c, e, "traversal counters");
        }
    }
    hipError_t e = scratch_begin(c, strm);
    if (e == hipSuccess)
        e = launch_pt_frame(s->ds, scene_scratch(c, s), a, gbuffer, pb, reinterpret_cast<float4*>(irradiance),
                            a.record_signature ? signature : nullptr, strm, stats ? &t : nullptr, stats ? &segs : nullptr,
                            tc, net, mcpg ? &margs : nullptr);
    unsigned long long hc[2] = {0, 0};
    hipError_t ec = hipSuccess;
    if (tc) {
        ec = hipMemcpyAsync(hc, tc, 16, hipMemcpyDeviceToHost, strm);
        if (ec == hipSuccess) ec = hipStreamSynchronize(strm);
        dev_free(c, tc);
    }
    if (e != hipSuccess) return hip_fail(c, e, "pt frame");
    if (ec != hipSuccess) return hip_fail(c, ec, "traversal counters");
    if (mcpg) mcpg_end_frame(mcpg, *cam);
    if (guide) {
        st = guide_train_frame(c, guide, strm); // reads the frame's records from the path buffers
        if (st != MRT_OK) return st;
    }
    if (const hipError_t ee = scratch_end(c, strm); ee != hipSuccess) return hip_fail(c, ee, "pt frame");
    if (stats) {
        stats->primary_segments = npix;
        stats->continuation_segments = segs;
        stats->shadow_rays = t.shadow_rays;
        stats->ms_trace = t.ms_trace;
        stats->ms_shade = t.ms_shade;
        stats->ms_total = t.ms_total;
        stats->trace_launches = t.trace_launches;
        stats->node_visits = hc[0];
        stats->triangle_tests = hc[1];
    }
    return MRT_OK;
}
} // namespace mrt

extern "C" {

MRT_API mrt_status mrt_pt_render_dev(mrt_ctx* c, mrt_scene* s, const mrt_pt_params* prm, const mrt_camera* cam,
                                     uint32_t frame, uint32_t w, uint32_t h, const mrt_packed_hit* gbuffer,
                                     float* irradiance, uint32_t* signature, mrt_pt_stats* stats, void* stream) {
    const mrt_status st = check_params(c, prm);
    if (st != MRT_OK) return st;
    return render_frame_dev(c, s, prm, cam, frame, w, h, gbuffer, irradiance, signature, stats, stream, nullptr);
}

MRT_API mrt_status mrt_render_frame(mrt_ctx* c, mrt_scene* s, const mrt_pt_params* prm, const mrt_camera* cam,
                                    uint32_t frame, uint32_t w, uint32_t h, float* irradiance, uint32_t* signature,
                                    mrt_packed_hit* gbuffer, mrt_pt_stats* stats) {
    if (!c || !s || !cam || !irradiance) return fail(c, MRT_ERR_INVALID, "mrt_render_frame: null argument");
    mrt_status st = check_params(c, prm);
    if (st != MRT_OK) return st;
    const uint64_t npix = uint64_t(w) * h;
    if (npix == 0) return MRT_OK;
    void* dg = dev_alloc(c, npix * sizeof(mrt_packed_hit), &st);
    if (st != MRT_OK) return st;
    void* dout = dev_alloc(c, npix * 16, &st);
    void* dsig = st == MRT_OK ? dev_alloc(c, npix * 4, &st) : nullptr;
    if (st == MRT_OK) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipError_t e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
        if (e != hipSuccess) st = hip_fail(c, e, "mrt_render_frame events");
        if (st == MRT_OK)
            st = mrt_gbuffer_dev(c, s, cam, w, h, prm->instance_mask, static_cast<mrt_packed_hit*>(dg), nullptr);
        if (st == MRT_OK && (e = hipEventRecord(e1, c->stream)) != hipSuccess) st = hip_fail(c, e, "mrt_render_frame events");
        if (st == MRT_OK)
            st = mrt_pt_render_dev(c, s, prm, cam, frame, w, h, static_cast<mrt_packed_hit*>(dg),
                                   static_cast<float*>(dout), signature ? static_cast<uint32_t*>(dsig) : nullptr, stats,
                                   nullptr);
        if (st == MRT_OK && stats) {
            e = hipEventSynchronize(e1);
            if (e == hipSuccess) e = hipEventElapsedTime(&stats->ms_gbuffer, e0, e1);
            if (e != hipSuccess) st = hip_fail(c, e, "mrt_render_frame gbuffer time");
        }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        e = hipSuccess;
        if (st == MRT_OK) e = hipMemcpyAsync(irradiance, dout, npix * 16, hipMemcpyDeviceToHost, c->stream);
        if (st == MRT_OK && e == hipSuccess && signature)
            e = hipMemcpyAsync(signature, dsig, npix * 4, hipMemcpyDeviceToHost, c->stream);
        if (st == MRT_OK && e == hipSuccess && gbuffer)
            e = hipMemcpyAsync(gbuffer, dg, npix * sizeof(mrt_packed_hit), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (st == MRT_OK && e != hipSuccess) st = hip_fail(c, e, "mrt_render_frame copy");
    }
    dev_free(c, dg);
    dev_free(c, dout);
    dev_free(c, dsig);
    return st;
}

} // extern "C"
