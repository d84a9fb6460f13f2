// Autotuning of include/merian_amd.h (mrt_tune, mrt_scene_get/set_tuning): per-scene BVH build
// settings (leaf size, SAH traversal cost) and launch settings (paths in flight, refill
// threshold, persistent grid size), chosen by timing trial frames of the caller's view.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/merian_amd.h"
#include "bvh.h"
#include "mrt_api_internal.h"

using namespace mrt;

namespace mrt {

mrt_status rebuild_bvh(mrt_ctx* c, mrt_scene* s, uint32_t width, uint32_t max_leaf, float traversal_cost) {
    if (s->h_flags.size() != s->n_tris) return fail(c, MRT_ERR_UNSUPPORTED, "scene keeps no triangle copy");
    if (width != 4 && width != 8) return fail(c, MRT_ERR_INVALID, "BVH width must be 4 or 8");
    const float *v0 = s->h_v0.data(), *e1 = s->h_e1.data(), *e2 = s->h_e2.data();
    const uint32_t* fl = s->h_flags.data();
    const uint32_t n = uint32_t(s->n_tris);
    // build on the host, then swap the device tree
    const void* node_src;
    const void* tri_src;
    size_t node_bytes, tri_bytes, n_nodes;
    uint32_t depth;
    BVH4 b4;
    BVH8 b8;
    if (width == 8) {
        b8 = build_bvh8(v0, e1, e2, fl, n, 0, int(max_leaf), traversal_cost);
        node_src = b8.nodes.data();
        node_bytes = b8.nodes.size() * sizeof(BVH8Node);
        n_nodes = b8.nodes.size();
        tri_src = b8.tris.data();
        tri_bytes = b8.tris.size() * sizeof(TriRecord);
        depth = b8.depth;
    } else {
        b4 = build_bvh4(v0, e1, e2, fl, n, 0, int(max_leaf), traversal_cost);
        node_src = b4.nodes.data();
        node_bytes = b4.nodes.size() * sizeof(BVH4Node);
        n_nodes = b4.nodes.size();
        tri_src = b4.tris.data();
        tri_bytes = b4.tris.size() * sizeof(TriRecord);
        depth = b4.depth;
    }
    // kernels in flight may still read the old tree
    if (const hipError_t se = hipStreamSynchronize(c->stream); se != hipSuccess) return hip_fail(c, se, "bvh rebuild");
    mrt_status st = MRT_OK;
    void* nodes = dev_alloc(c, node_bytes, &st);
    if (st != MRT_OK) return st;
    void* tris = dev_alloc(c, tri_bytes, &st);
    if (st != MRT_OK) {
        dev_free(c, nodes);
        return st;
    }
    hipError_t e = hipMemcpyAsync(nodes, node_src, node_bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(tris, tri_src, tri_bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        dev_free(c, nodes);
        dev_free(c, tris);
        return hip_fail(c, e, "BVH upload");
    }
    void* olds[3] = {(void*)s->ds.nodes, (void*)s->ds.nodes8, (void*)s->ds.tris};
    for (void* old : olds) {
        if (!old) continue;
        auto it = std::find(s->allocs.begin(), s->allocs.end(), old);
        if (it != s->allocs.end()) s->allocs.erase(it);
        s->bytes -= c->sizes.count(old) ? c->sizes[old] : 0;
        dev_free(c, old);
    }
    s->ds.nodes = width == 4 ? static_cast<const BVH4Node*>(nodes) : nullptr;
    s->ds.nodes8 = width == 8 ? static_cast<const BVH8Node*>(nodes) : nullptr;
    s->ds.width = width;
    s->ds.tris = static_cast<const float4*>(tris);
    s->allocs.push_back(nodes);
    s->allocs.push_back(tris);
    s->bytes += node_bytes + tri_bytes;
    // at most width - 1 entries per level (a grouped BVH8 traversal needs one)
    s->ds.stack_need = (width - 1u) * depth + 4u;
    s->n_nodes = n_nodes;
    s->depth = depth;
    s->tuning.bvh_width = width;
    s->tuning.bvh_max_leaf = max_leaf;
    s->tuning.bvh_traversal_cost = traversal_cost;
    return MRT_OK;
}

} // namespace mrt

namespace {

bool valid(const mrt_tuning& t) {
    return t.refill_lanes >= 1 && t.refill_lanes <= 64 && t.grid_percent >= 1 && t.grid_percent <= 100 &&
           t.bvh_max_leaf >= 1 && t.bvh_max_leaf <= uint32_t(MAX_LEAF) && t.bvh_traversal_cost > 0.f &&
           (t.bvh_width == 0 || t.bvh_width == 4 || t.bvh_width == 8) && t.leaf_lanes <= 64 && t.bvh8_order <= 1;
}

// Device time of one trial frame (gbuffer + pt) in ms, after one untimed frame.
mrt_status trial(mrt_ctx* c, mrt_scene* s, const mrt_tune_request& r, float& ms) {
    mrt_pt_params p = r.params;
    p.spp = std::min(std::max(p.spp, 1), 64); // whole passes of the real frame up to 64 spp
    p.record_signature = 0;
    p.count_traversal = 0;
    const uint64_t npix = uint64_t(r.width) * r.height;
    mrt_status st = MRT_OK;
    void* gb = dev_alloc(c, npix * sizeof(mrt_packed_hit), &st);
    if (st != MRT_OK) return st;
    void* out = dev_alloc(c, npix * 16, &st);
    if (st != MRT_OK) {
        dev_free(c, gb);
        return st;
    }
    float best = 1e30f;
    for (int k = 0; k < 3 && st == MRT_OK; k++) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipError_t e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
        if (e != hipSuccess) st = hip_fail(c, e, "tune trial events");
        if (st == MRT_OK)
            st = mrt_gbuffer_dev(c, s, &r.camera, r.width, r.height, p.instance_mask,
                                 static_cast<mrt_packed_hit*>(gb), nullptr);
        if (st == MRT_OK)
            st = render_frame_dev(c, s, &p, &r.camera, uint32_t(k), r.width, r.height,
                                  static_cast<mrt_packed_hit*>(gb), static_cast<float*>(out), nullptr, nullptr,
                                  nullptr, nullptr);
        float t = 0.f;
        if (st == MRT_OK) {
            e = hipEventRecord(e1, c->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
            if (e != hipSuccess) st = hip_fail(c, e, "tune trial timing");
        }
        if (st == MRT_OK && k > 0) best = std::min(best, t); // the first frame may (re)allocate buffers
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
    dev_free(c, gb);
    dev_free(c, out);
    ms = best;
    return st;
}

} // namespace

extern "C" {

MRT_API mrt_status mrt_scene_get_tuning(mrt_scene* s, mrt_tuning* out) {
    if (!s || !out) return MRT_ERR_INVALID;
    *out = s->tuning;
    return MRT_OK;
}

MRT_API mrt_status mrt_scene_set_tuning(mrt_ctx* c, mrt_scene* s, const mrt_tuning* t) {
    if (!c || !s || !t) return MRT_ERR_INVALID;
    if (!valid(*t)) return fail(c, MRT_ERR_INVALID, "tuning values out of range");
    if (const hipError_t e = hipSetDevice(c->device); e != hipSuccess) return hip_fail(c, e, "mrt_scene_set_tuning");
    const uint32_t width = t->bvh_width ? t->bvh_width : 4u;
    // a grouped stack entry packs the node index into 24 bits (mrt_trace.h); a tree has fewer
    // nodes than triangles, so the check precedes any rebuild
    if (width == 8 && t->bvh8_order == 1 && s->n_tris > (uint64_t(1) << 24))
        return fail(c, MRT_ERR_UNSUPPORTED, "BVH8 octant order needs at most 2^24 triangles");
    if (t->bvh_max_leaf != s->tuning.bvh_max_leaf || t->bvh_traversal_cost != s->tuning.bvh_traversal_cost ||
        width != s->tuning.bvh_width) {
        const mrt_status st = rebuild_bvh(c, s, width, t->bvh_max_leaf, t->bvh_traversal_cost);
        if (st != MRT_OK) return st;
    }
    s->tuning = *t;
    s->tuning.bvh_width = width;
    if (s->tuning.leaf_lanes == 0) s->tuning.leaf_lanes = 16;
    s->ds.bvh8_order = width == 8 ? t->bvh8_order : 0u;
    return MRT_OK;
}

MRT_API mrt_status mrt_tune(mrt_ctx* c, mrt_scene* s, const mrt_tune_request* r, mrt_tuning* best_out) {
    if (!c || !s || !r) return MRT_ERR_INVALID;
    if (r->width == 0 || r->height == 0) return fail(c, MRT_ERR_INVALID, "mrt_tune: empty image");
    const mrt_status pc = check_params(c, &r->params);
    if (pc != MRT_OK) return pc;
    if (const hipError_t e = hipSetDevice(c->device); e != hipSuccess) return hip_fail(c, e, "mrt_tune");
    const auto t0 = std::chrono::steady_clock::now();
    const double budget = r->time_budget_ms > 0.f ? double(r->time_budget_ms) : 20000.0;
    auto elapsed = [&] { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); };

    // MRT_TUNE_VERBOSE=1 prints every trial to stderr
    const char* ev = std::getenv("MRT_TUNE_VERBOSE");
    const bool verbose = ev && ev[0] == '1';
    mrt_tuning best = s->tuning;
    if (best.max_paths == 0) best.max_paths = c->max_paths;
    mrt_status st = mrt_scene_set_tuning(c, s, &best);
    if (st != MRT_OK) return st;
    st = trial(c, s, *r, best.ms_per_frame);
    if (st != MRT_OK) return st;

    auto try_one = [&](mrt_tuning cand) -> mrt_status {
        if (elapsed() > budget) return MRT_OK;
        mrt_status e = mrt_scene_set_tuning(c, s, &cand);
        if (e != MRT_OK) return e;
        float ms = 0.f;
        e = trial(c, s, *r, ms);
        if (e == MRT_ERR_BUDGET || e == MRT_ERR_ALLOC) {
            c->err.clear(); // a candidate that does not fit is skipped
            return mrt_scene_set_tuning(c, s, &best);
        }
        if (e != MRT_OK) return e;
        cand.ms_per_frame = ms;
        if (verbose)
            std::fprintf(stderr,
                         "mrt_tune: width %u order %u leaf %u cost %.2f paths %llu refill %u leaf-phase %u grid %u%% -> "
                         "%.3f ms (best %.3f)\n",
                         cand.bvh_width, cand.bvh8_order, cand.bvh_max_leaf, double(cand.bvh_traversal_cost),
                         (unsigned long long)cand.max_paths, cand.refill_lanes, cand.leaf_lanes, cand.grid_percent,
                         double(ms), double(best.ms_per_frame));
        if (ms < best.ms_per_frame) best = cand;
        return mrt_scene_set_tuning(c, s, &best);
    };
    const uint64_t npix = uint64_t(r->width) * r->height;
    for (int pass = 0; pass < 2 && elapsed() < budget; pass++) {
        const mrt_tuning start = best;
        // BVH build settings
        struct B {
            uint32_t width, leaf;
            float cost;
        };
        const B bvh[] = {{4, 4, 1.0f}, {4, 2, 1.0f}, {4, 8, 1.0f}, {8, 2, 1.0f}, {8, 4, 1.0f},
                         {8, 1, 1.0f}, {4, 4, 0.5f}, {4, 4, 2.0f}};
        for (const B& b : bvh) {
            if (b.width == best.bvh_width && b.leaf == best.bvh_max_leaf && b.cost == best.bvh_traversal_cost) continue;
            mrt_tuning cand = best;
            cand.bvh_width = b.width;
            cand.bvh_max_leaf = b.leaf;
            cand.bvh_traversal_cost = b.cost;
            if ((st = try_one(cand)) != MRT_OK) return st;
        }
        // paths in flight (whole samples of the image per pass): up to 2^28 (about 48 GB of path
        // state), so that a 1080p frame of 64 spp can run as one pass; a candidate beyond the
        // context's HBM budget is skipped
        for (uint64_t mp : {uint64_t(1) << 21, uint64_t(1) << 22, uint64_t(1) << 23, uint64_t(1) << 24,
                            uint64_t(1) << 25, uint64_t(1) << 26, uint64_t(1) << 27, uint64_t(1) << 28}) {
            if (mp == best.max_paths || mp < npix) continue;
            mrt_tuning cand = best;
            cand.max_paths = mp;
            if ((st = try_one(cand)) != MRT_OK) return st;
        }
        for (uint32_t rf : {4u, 8u, 16u, 24u, 32u, 48u, 64u}) {
            if (rf == best.refill_lanes) continue;
            mrt_tuning cand = best;
            cand.refill_lanes = rf;
            if ((st = try_one(cand)) != MRT_OK) return st;
        }
        for (uint32_t ll : {4u, 8u, 16u, 24u, 32u, 48u}) {
            if (ll == best.leaf_lanes) continue;
            mrt_tuning cand = best;
            cand.leaf_lanes = ll;
            if ((st = try_one(cand)) != MRT_OK) return st;
        }
        for (uint32_t gp : {50u, 75u, 100u}) {
            if (gp == best.grid_percent) continue;
            mrt_tuning cand = best;
            cand.grid_percent = gp;
            if ((st = try_one(cand)) != MRT_OK) return st;
        }
        if (best.bvh_width == 8) {
            mrt_tuning cand = best;
            cand.bvh8_order ^= 1u;
            if (cand.bvh8_order == 0u || s->n_tris <= (uint64_t(1) << 24))
                if ((st = try_one(cand)) != MRT_OK) return st;
        }
        if (std::memcmp(&start, &best, sizeof(best)) == 0) break;
    }
    st = mrt_scene_set_tuning(c, s, &best);
    if (st == MRT_OK && best_out) *best_out = best;
    return st;
}

} // extern "C"
