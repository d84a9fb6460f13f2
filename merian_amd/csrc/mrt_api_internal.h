// Internal state behind the opaque handles of include/merian_amd.h.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/merian_amd.h"
#include "mrt_kernels.h"

struct mrt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;  // bytes the context may hold
    uint64_t in_use = 0;  // bytes it holds
    uint64_t max_paths = 1ull << 24; // paths in flight per wavefront pass
    std::unordered_map<void*, size_t> sizes;
    std::string err;
    mrt::PathBuffers paths{};
    uint64_t paths_npix = 0, paths_counters = 0;
    mrt::TraceScratch trace{};
    // last launch that used the context's shared scratch (trace stacks and cursor, path buffers)
    hipEvent_t scratch_event = nullptr;
    hipStream_t scratch_stream = nullptr;
};

struct mrt_scene {
    mrt::DeviceScene ds{};
    std::vector<void*> allocs;
    uint64_t bytes = 0, n_tris = 0, n_nodes = 0;
    uint32_t depth = 0;
    float bbox_min[3] = {0.f, 0.f, 0.f}, bbox_max[3] = {0.f, 0.f, 0.f}; // bounds of the triangles
    // triangle operands kept on the host for BVH rebuilds (tuning)
    std::vector<float> h_v0, h_e1, h_e2;
    std::vector<uint32_t> h_flags;
    // defaults: what mrt_tune found best on the 1080p 2M-triangle workload, grid and paths aside
    mrt_tuning tuning{0, 24, 100, 2, 1.0f, 4, 0.f, 16, 0};
};

namespace mrt {
mrt_status fail(mrt_ctx* c, mrt_status s, const std::string& msg);
mrt_status hip_fail(mrt_ctx* c, hipError_t e, const char* what);
void* dev_alloc(mrt_ctx* c, size_t bytes, mrt_status* st);
void dev_free(mrt_ctx* c, void* p);
hipStream_t pick_stream(mrt_ctx* c, void* stream);
mrt_status ensure_path_buffers(mrt_ctx* c, uint64_t paths, uint64_t npix, uint64_t counters);
mrt_status check_params(mrt_ctx* c, const mrt_pt_params* p);
mrt_status ensure_trace_scratch(mrt_ctx* c, const mrt_scene* s);
// The context's traversal scratch with the scene's launch settings.
TraceScratch scene_scratch(const mrt_ctx* c, const mrt_scene* s);
// Every launch that uses the context's scratch (traversal stacks and work cursor, path buffers)
// is bracketed by these two calls: a launch on a stream other than the previous user's first
// waits for that user's work, then records itself as the last user. Launches on one stream are
// ordered by the stream already. A stream being captured into a graph is skipped (an event
// recorded outside a capture cannot be waited on inside it): captured launches are ordered by
// the caller.
hipError_t scratch_begin(mrt_ctx* c, hipStream_t st);
// Looks up the device address of the allocation-free entry point's fixed scratch on the
// context's device (noalloc.cpp); mrt_create calls it with that device current.
mrt_status noalloc_setup(mrt_ctx* c);
hipError_t scratch_end(mrt_ctx* c, hipStream_t st);
// Rebuilds and uploads the scene's BVH with the given node width (4 or 8), leaf size and SAH
// traversal cost.
mrt_status rebuild_bvh(mrt_ctx* c, mrt_scene* s, uint32_t width, uint32_t max_leaf, float traversal_cost);
// One pt frame on the device (mrt_pt_render_dev; with a guide, mrt_guided_render_dev; with
// MCPG state, mrt_mcpg_render_dev).
mrt_status render_frame_dev(mrt_ctx* c, mrt_scene* s, const mrt_pt_params* prm, const mrt_camera* cam, uint32_t frame,
                            uint32_t w, uint32_t h, const mrt_packed_hit* gbuffer, float* irradiance,
                            uint32_t* signature, mrt_pt_stats* stats, void* stream, mrt_guide* guide,
                            mrt_mcpg* mcpg = nullptr);
// MCPG hooks of render_frame_dev (mcpg_api.cpp): the carry buffer for `paths` paths and the
// frame's kernel arguments; after the frame, its camera becomes the previous one.
mrt_status mcpg_prepare_frame(mrt_ctx* c, mrt_mcpg* m, uint64_t paths, const mrt_camera& cam, McpgArgs& a);
void mcpg_end_frame(mrt_mcpg* m, const mrt_camera& cam);
// Guide hooks of render_frame_dev (guide_api.cpp): buffers for `paths` paths in flight, and the
// training of the frame's records after it was enqueued.
mrt_status guide_prepare_frame(mrt_ctx* c, mrt_guide* g, uint64_t paths, PathBuffers& pb, float& fraction,
                               uint32_t& guided_bounces, const GuideNet** net);
mrt_status guide_train_frame(mrt_ctx* c, mrt_guide* g, hipStream_t st);
} // namespace mrt
