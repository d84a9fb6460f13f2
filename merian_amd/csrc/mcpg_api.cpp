// C ABI of the MCPG renderer (include/merian_amd.h, mrt_mcpg_*): the two hash grids it keeps
// across frames, the per-path carry buffer and the frame entry point. The frame itself is
// launch_pt_frame in MRT_MODE_MCPG (kernels.hip: k_mcpg_init, k_shade_mcpg).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "mrt_api_internal.h"

struct mrt_mcpg {
    mrt_mcpg_config cfg{};
    uint2* mc_keys = nullptr;
    float4* mc_data = nullptr; // 2 float4 (32 B) per slot
    uint2* lc_keys = nullptr;
    uint2* lc_data = nullptr;  // 8 B per slot
    float4* carry = nullptr;   // 5 float4 per path in flight
    uint64_t carry_paths = 0;
    bool has_prev = false;     // prev_cam is valid (a frame ran since the last reset)
    float prev_cam[3] = {0.f, 0.f, 0.f};
};

using namespace mrt;

namespace {

// Slots per grid: at most 2^28 (the node's own limits are 1e8 for the light cache), so every
// grid's word count fits the 32-bit fill kernel.
constexpr uint32_t MAX_GRID_SLOTS = 1u << 28;

bool grid_ok(const mrt_mcpg_grid_config& g) {
    return g.buffer_size >= 1u && g.buffer_size <= MAX_GRID_SLOTS && g.probe_count >= 1u && g.probe_count <= 32u &&
           g.locality_bits <= 5u && g.stochastic_interpolation <= 1u && g.tan_alpha_half > 0.f && std::isfinite(g.tan_alpha_half) &&
           std::isfinite(g.level_bias) && g.distribution_dimension > 0.f && std::isfinite(g.distribution_dimension);
}

bool config_ok(const mrt_mcpg_config& c) {
    return grid_ok(c.mc) && grid_ok(c.lc) && c.mc_samples >= 0 && c.mc_samples <= MRT_MCPG_MAX_SAMPLES &&
           c.p_guiding >= 0.f && c.p_guiding <= 1.f && (c.sampling_type == MRT_MCPG_MIS ||
                                                         c.sampling_type == MRT_MCPG_MIS_ROUGHNESS) &&
           c.guiding_alpha_threshold >= 0.f && std::isfinite(c.guiding_alpha_threshold) && c.dir_guide_prior >= 0.f &&
           std::isfinite(c.dir_guide_prior) && c.lc_min_pdf >= 0.f && std::isfinite(c.lc_min_pdf) &&
           (c.bsdf_support_only == 0 || c.bsdf_support_only == 1);
}

// Zero-fills both grids with a fill kernel on the context stream and waits for it. Frames may
// have run on a caller's stream, so the whole device is drained first.
mrt_status clear_grids(mrt_ctx* c, mrt_mcpg* m) {
    const uint64_t mc = m->cfg.mc.buffer_size, lc = m->cfg.lc.buffer_size;
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = launch_zero_u32(reinterpret_cast<uint32_t*>(m->mc_keys), uint32_t(2 * mc), c->stream);
    if (e == hipSuccess) e = launch_zero_u32(reinterpret_cast<uint32_t*>(m->mc_data), uint32_t(8 * mc), c->stream);
    if (e == hipSuccess) e = launch_zero_u32(reinterpret_cast<uint32_t*>(m->lc_keys), uint32_t(2 * lc), c->stream);
    if (e == hipSuccess) e = launch_zero_u32(reinterpret_cast<uint32_t*>(m->lc_data), uint32_t(2 * lc), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    m->has_prev = false;
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, "mcpg grid reset");
}

void free_grids(mrt_ctx* c, mrt_mcpg* m) {
    for (void* p : {(void*)m->mc_keys, (void*)m->mc_data, (void*)m->lc_keys, (void*)m->lc_data, (void*)m->carry})
        if (p) dev_free(c, p);
    m->mc_keys = nullptr;
    m->mc_data = nullptr;
    m->lc_keys = nullptr;
    m->lc_data = nullptr;
    m->carry = nullptr;
    m->carry_paths = 0;
}

HashGridDev grid_dev(const mrt_mcpg_grid_config& g, uint2* keys, void* data, bool mc) {
    HashGridDev d{};
    d.keys = keys;
    d.data = data;
    d.size = g.buffer_size;
    d.probes = g.probe_count;
    d.locality = g.locality_bits;
    // the MC grid's instantiation fixes jittered cells and levels and lock-free writers
    // (mcpg_common.slang:29-30, mc.slang:135-137); the light cache's takes the property
    // (hashed-irradiance-cache.slang:43-45)
    d.stochastic = mc ? 1u : g.stochastic_interpolation;
    d.level_stochastic = mc ? 1u : 0u;
    d.lock_free = mc ? 1u : 0u;
    d.tan_alpha_half = g.tan_alpha_half;
    d.level_bias = g.level_bias;
    d.dist_dim = g.distribution_dimension;
    return d;
}

} // namespace

namespace mrt {
// Carry buffer for `paths` paths and the frame's kernel arguments (render_frame_dev).
mrt_status mcpg_prepare_frame(mrt_ctx* c, mrt_mcpg* m, uint64_t paths, const mrt_camera& cam, McpgArgs& a) {
    if (m->carry_paths < paths) {
        if (m->carry) dev_free(c, m->carry);
        m->carry = nullptr;
        m->carry_paths = 0;
        mrt_status st = MRT_OK;
        m->carry = static_cast<float4*>(dev_alloc(c, 80 * paths, &st));
        if (st != MRT_OK) return st;
        m->carry_paths = paths;
    }
    const mrt_mcpg_config& k = m->cfg;
    a = McpgArgs{};
    a.mc = grid_dev(k.mc, m->mc_keys, m->mc_data, true);
    a.lc = grid_dev(k.lc, m->lc_keys, m->lc_data, false);
    a.carry = m->carry;
    a.samples = k.mc_samples;
    a.p_guiding = k.p_guiding;
    a.roughness = k.sampling_type == MRT_MCPG_MIS_ROUGHNESS ? 1 : 0;
    a.alpha_threshold = k.guiding_alpha_threshold;
    a.prior = k.dir_guide_prior;
    a.missing_light = k.missing_light_heuristic ? 1 : 0;
    a.lc_tail = k.use_light_cache_tail ? 1 : 0;
    a.lc_min_pdf = k.lc_min_pdf;
    a.bsdf_support_only = k.bsdf_support_only;
    for (int i = 0; i < 3; i++) {
        a.cam[i] = cam.position[i];
        a.prev_cam[i] = m->has_prev ? m->prev_cam[i] : cam.position[i];
    }
    a.time = 0.f; // static scenes: zero motion, so the state's time stamp never moves a target
    return MRT_OK;
}
void mcpg_end_frame(mrt_mcpg* m, const mrt_camera& cam) {
    for (int i = 0; i < 3; i++) m->prev_cam[i] = cam.position[i];
    m->has_prev = true;
}
} // namespace mrt

extern "C" {

MRT_API void mrt_mcpg_default_config(mrt_mcpg_config* out) {
    if (!out) return;
    mrt_mcpg_config k{};
    // render_pt_mcpg.hpp:125-152 and the HashGrid defaults (hash_grid.hpp:42-44)
    k.mc = mrt_mcpg_grid_config{32777259u, 2u, 3u, 1u, 0.006f, 0.0f, 2.0f, 0u};
    k.lc = mrt_mcpg_grid_config{4000037u, 4u, 3u, 0u, 0.006f, 0.0f, 2.0f, 0u};
    k.mc_samples = 5;
    k.p_guiding = 0.5f;
    k.sampling_type = MRT_MCPG_MIS_ROUGHNESS;
    k.guiding_alpha_threshold = 0.05f;
    k.dir_guide_prior = 0.2f;
    k.missing_light_heuristic = 1;
    k.use_light_cache_tail = 0;
    k.lc_min_pdf = 1.0f;
    *out = k;
}

MRT_API mrt_status mrt_mcpg_create(mrt_ctx* c, const mrt_mcpg_config* config, mrt_mcpg** out) {
    if (!c || !config || !out) return fail(c, MRT_ERR_INVALID, "mrt_mcpg_create: null argument");
    *out = nullptr;
    if (!config_ok(*config)) return fail(c, MRT_ERR_INVALID, "mrt_mcpg_create: configuration out of range");
    if (const hipError_t e = hipSetDevice(c->device); e != hipSuccess) return hip_fail(c, e, "mrt_mcpg_create");
    mrt_mcpg* m = new mrt_mcpg();
    m->cfg = *config;
    const uint64_t mc = config->mc.buffer_size, lc = config->lc.buffer_size;
    mrt_status st = MRT_OK;
    m->mc_keys = static_cast<uint2*>(dev_alloc(c, 8 * mc, &st));
    if (st == MRT_OK) m->mc_data = static_cast<float4*>(dev_alloc(c, 32 * mc, &st));
    if (st == MRT_OK) m->lc_keys = static_cast<uint2*>(dev_alloc(c, 8 * lc, &st));
    if (st == MRT_OK) m->lc_data = static_cast<uint2*>(dev_alloc(c, 8 * lc, &st));
    if (st == MRT_OK) st = clear_grids(c, m);
    if (st != MRT_OK) {
        free_grids(c, m);
        delete m;
        return st;
    }
    *out = m;
    return MRT_OK;
}

MRT_API mrt_status mrt_mcpg_destroy(mrt_ctx* c, mrt_mcpg* m) {
    if (!c || !m) return MRT_ERR_INVALID;
    hipError_t e = hipSetDevice(c->device);
    if (e == hipSuccess) e = hipDeviceSynchronize(); // kernels on any stream may still use the grids
    free_grids(c, m);
    delete m;
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, "mrt_mcpg_destroy");
}

MRT_API mrt_status mrt_mcpg_reset(mrt_ctx* c, mrt_mcpg* m) {
    if (!c || !m) return MRT_ERR_INVALID;
    if (const hipError_t e = hipSetDevice(c->device); e != hipSuccess) return hip_fail(c, e, "mrt_mcpg_reset");
    return clear_grids(c, m);
}

MRT_API mrt_status mrt_mcpg_render_dev(mrt_ctx* c, mrt_scene* s, mrt_mcpg* m, const mrt_pt_params* prm,
                                       const mrt_camera* cam, uint32_t frame, uint32_t w, uint32_t h,
                                       const mrt_packed_hit* gbuffer, float* irradiance, mrt_pt_stats* stats,
                                       void* stream) {
    if (!c || !s || !m || !prm || !cam) return fail(c, MRT_ERR_INVALID, "mrt_mcpg_render_dev: null argument");
    mrt_pt_params p = *prm;
    p.mode = MRT_MODE_BSDF; // validated as a plain pt frame, then run as MCPG
    p.record_signature = 0;
    const mrt_status st = check_params(c, &p);
    if (st != MRT_OK) return st;
    return render_frame_dev(c, s, &p, cam, frame, w, h, gbuffer, irradiance, nullptr, stats, stream, nullptr, m);
}

MRT_API mrt_status mrt_mcpg_read_grids(mrt_ctx* c, mrt_mcpg* m, uint32_t* mc_keys, void* mc_states, uint32_t* lc_keys,
                                       void* lc_entries) {
    if (!c || !m) return MRT_ERR_INVALID;
    // frames may have run on a caller's stream: drain the device before the copies
    if (const hipError_t e = hipSetDevice(c->device); e != hipSuccess) return hip_fail(c, e, "mrt_mcpg_read_grids");
    if (const hipError_t e = hipDeviceSynchronize(); e != hipSuccess) return hip_fail(c, e, "mrt_mcpg_read_grids");
    const uint64_t mc = m->cfg.mc.buffer_size, lc = m->cfg.lc.buffer_size;
    hipError_t e = hipSuccess;
    if (mc_keys) e = hipMemcpy(mc_keys, m->mc_keys, 8 * mc, hipMemcpyDeviceToHost);
    if (e == hipSuccess && mc_states) e = hipMemcpy(mc_states, m->mc_data, 32 * mc, hipMemcpyDeviceToHost);
    if (e == hipSuccess && lc_keys) e = hipMemcpy(lc_keys, m->lc_keys, 8 * lc, hipMemcpyDeviceToHost);
    if (e == hipSuccess && lc_entries) e = hipMemcpy(lc_entries, m->lc_data, 8 * lc, hipMemcpyDeviceToHost);
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, "mrt_mcpg_read_grids");
}

} // extern "C"
