// Host <-> device interface of the gfx950 kernels (kernels.hip): the device-resident scene
// view, the per-frame arguments of the wavefront path tracer and the launch entry points
// called by the C ABI (api.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "bvh.h"
#include "../../include/merian_amd.h"

namespace mrt {

// Scene as the kernels see it (the reference's Scene parameter block, scene.slang:30-54).
struct DeviceScene {
    const BVH4Node* nodes;       // width 4
    const BVH8Node* nodes8;      // width 8 (quantised)
    uint32_t width;              // BVH node width in use: 4 or 8
    uint32_t bvh8_order;         // BVH8 visiting order: 0 sorted entries, 1 octant groups (mrt_tuning)
    const float4* tris;          // TriRecord stream, leaf order
    const uint4* tri_verts;      // per global triangle id: 3 global vertex indices, geometry index
    const float4* tri_shade;     // per global triangle id, 64 B: (p0, n0), (p1, n1), (p2, n2),
                                 // (uv0, uv1, uv2, geometry | shading class << 24); n = oct normal
                                 // bits, uv = 2 x half bits
    const mrt_vertex* verts;     // PackedVertexData
    const mrt_geometry* geoms;   // GeometryData
    const uint32_t* geom_first;  // global id of each geometry's first triangle
    const mrt_material* mats;    // material entries
    const float4* env;           // environment texels (RGBA32F), row-major
    const float4* env_samp;      // per texel: alias-table probability, alias index bits, pmf, 0
    int32_t env_type, env_wrap;
    uint32_t env_w, env_h;
    float env_intensity;
    float env_to_local[9];
    uint32_t n_tris;
    uint32_t n_geoms;
    uint32_t stack_need;         // traversal stack entries the tree can need (width - 1 per level + slack)
};

// Shading classes: the shade kernel groups a block's paths by class so each wave runs one
// material code path. Hit triangles carry theirs (DeviceScene.tri_shade); misses are ENV.
enum : uint32_t {
    SHADE_CLASS_ENV = 0,         // environment (a miss, or a UseEnvMap geometry)
    SHADE_CLASS_DIFFUSE = 1,     // Lambert
    SHADE_CLASS_METAL = 2,       // OpenPBR, metalness >= 0.5, no transmission
    SHADE_CLASS_DIELECTRIC = 3,  // OpenPBR, metalness < 0.5, no transmission
    SHADE_CLASS_TRANSMISSIVE = 4, // OpenPBR with transmission
    SHADE_CLASSES = 5
};
constexpr uint32_t TRI_SHADE_GEOM_MASK = 0xFFFFFFu;
constexpr uint32_t TRI_SHADE_CLASS_SHIFT = 24;

// Entries of the per-lane traversal stack kept in LDS; deeper entries go to TraceScratch.ovf.
constexpr int TRACE_LDS_STACK = 16;
// Threads per block of the trace kernels (also the LDS stack's column stride).
constexpr int TRACE_BLOCK = 128;

// Per-context scratch of the trace kernels: the global part of the traversal stacks (`entries`
// per thread, `threads` columns) and a ray cursor for launches that do not bring their own.
struct TraceScratch {
    uint32_t* ovf;
    uint32_t threads;
    uint32_t entries;
    uint32_t* work;
    uint32_t refill = 24;   // idle lanes before a wave refetches rays (1..64)
    uint32_t grid_pct = 100; // persistent grid size, percent of the resident blocks
    uint32_t leaf_lanes = 16; // leaf-only lanes before a wave runs a leaf-test phase (1..64)
};

// Clears n words with a fill kernel (graph-capturable, unlike a memset).
hipError_t launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t st);

// Most trace-kernel threads resident on the current device (the persistent grid's upper bound).
uint32_t trace_max_threads();
// Runs the one-time occupancy queries of the frame's trace kernels (call before capturing).
void trace_prepare();

// Frame-constant arguments of the wavefront pt kernels.
struct PtArgs {
    mrt_camera cam;
    uint32_t width, height, npix, frame;
    int32_t spp, max_path_length, emission_on_primary, record_signature;
    uint32_t instance_mask;
    int32_t mode;     // MRT_MODE_*
    float guide_fraction; // MRT_MODE_GUIDED: probability of sampling the guide at an eligible vertex
    uint32_t guided_bounces; // MRT_MODE_GUIDED: vertices [0, guided_bounces) use the guide
    // MRT_MODE_GUIDED: training records come from the pixels p with
    // (p * 2654435761) >> (32 - shift) == 0 (all pixels when 0), a pseudo-random 2^-shift of the
    // image, so that a frame whose records outnumber the record buffer still trains on all of it
    uint32_t guide_record_shift;
    int32_t demodulate; // divide every sample by the pixel's floored GBuffer albedo
    uint32_t s0;      // first sample index of this chunk
    uint32_t chunk;   // samples in this chunk
    uint32_t n_paths; // npix * chunk
};

// Device buffers of one chunk of paths (structure of arrays).
struct PathBuffers {
    float4* state;     // per path 32 B: [2 pid] = throughput.xyz, current ior; [2 pid + 1] = radiance.xyz,
                       // PCG state bits. Modes whose state travels with the queue (state_in_queue) keep
                       // only [2 pid + 1] here: the radiance of a path that has ended.
    uint32_t* sig;     // hit-sequence hash per path
    float4* q_o[2];    // ray queue: origin.xyz, path index bits
    float4* q_d[2];    // ray queue: direction.xyz, NEE: pdf of the sampled direction (MIS weight of an
                       // escaping ray)
    float4* q_t[2];    // state of the queued path (state_in_queue modes): throughput.xyz, current ior
    float4* q_r[2];    //   radiance.xyz, PCG state bits
    float4* hits;      // t, u, v, triangle id bits; indexed like the queue being shaded
    uint32_t* counters;// per chunk and bounce: queue sizes, then the trace kernels' ray cursors
    float4* sq_o;      // shadow queue: origin.xyz, path index bits
    float4* sq_d;      // shadow queue: direction.xyz
    float4* sq_c;      // shadow queue: unoccluded contribution.rgb
    float4* accum;     // per pixel running sum: radiance.xyz, luminance^2
    float4* demod;     // per pixel demodulation albedo (floored GBuffer albedo), "demodulate albedo" only
    uint32_t* sig_accum;
    // MRT_MODE_GUIDED only
    uint4* lobes;      // packed guide lobes per queue entry (4 x uint4)
    float4* grec;      // raw training records: (pos, path), (wo, pdf), (dir, bounce), (radiance before, -), (throughput after, -)
    float4* gtrain;    // training records in arrival order: (pos, weight), (wo, -), (dir, key = path * 64 + bounce)
    uint32_t* grec_count;
    uint32_t grec_cap;
    uint64_t capacity; // paths
};

// Modes whose per-path state rides in the ray queue beside the ray (q_t, q_r), so the shading
// kernels read and write it in queue order instead of gathering it by path index. MCPG keeps it
// per path: its passes continue a pixel's RNG stream from one sample to the next.
__host__ __device__ constexpr bool state_in_queue(int mode) { return mode != MRT_MODE_MCPG; }
// Paths per pass in NEE mode: a shadow ray names its path's queue slot or path index in 31 bits.
constexpr uint64_t NEE_MAX_PATHS = 0x7FFFFFFFull;

struct GuideNet;

// One hash grid of the MCPG renderer (utils/hash-grid.slang:44-100) in device memory.
struct HashGridDev {
    uint2* keys;               // per slot: (cell key, frame stamp | lock bit); zero after a reset
    void* data;                // per slot payload (chain state 32 B, light-cache entry 8 B)
    uint32_t size;             // slots
    uint32_t probes;           // PROBE_COUNT
    uint32_t locality;         // LOCALITY_BITS
    uint32_t stochastic;       // STOCHASTIC_INTERPOLATION (jittered cell)
    uint32_t level_stochastic; // LEVEL_STOCHASTIC_INTERPOLATION (jittered level)
    uint32_t lock_free;        // LOCK_FREE writers
    float tan_alpha_half;      // grid_tan_alpha_half
    float level_bias;          // grid_level_bias
    float dist_dim;            // grid_distribution_dimension
};

// Frame arguments of the MCPG kernels (render_pt_mcpg.slang:12-30 constants, the two grids,
// the cameras) and the per-path record that carries a scatter vertex's chain update to the
// next bounce: 5 float4 per path (selected state, score sum, position, pdf, normal,
// light-cache cell, BSDF value).
struct McpgArgs {
    HashGridDev mc, lc;
    float4* carry;
    int32_t samples;         // mc_samples (RIS candidates, <= 8)
    float p_guiding;
    int32_t roughness;       // GUIDING_DIRECTIONAL_SAMPLING_MIS_ROUGHNESS
    float alpha_threshold;   // guiding_alpha_threshold
    float prior;             // dir_guide_prior
    int32_t missing_light;   // missing_light_heuristic
    int32_t lc_tail;         // use_light_cache_tail
    float lc_min_pdf;
    int32_t bsdf_support_only; // a direction with BSDF sampling pdf 0 contributes nothing
    float cam[3], prev_cam[3];
    float time;
};

// Launch entry points (kernels.hip). All asynchronous on `stream`.
// counters (optional, device): node visits and triangle tests are added to counters[0..1]
hipError_t launch_trace_rays(const DeviceScene& s, const TraceScratch& ts, const mrt_ray* rays, uint64_t n,
                             mrt_hit* hits, hipStream_t st, unsigned long long* counters = nullptr);
hipError_t launch_camera_rays(const mrt_camera& cam, uint32_t w, uint32_t h, mrt_ray* rays, hipStream_t st);
hipError_t launch_running_mean(float4* acc, const float4* src, uint32_t n, uint32_t count, hipStream_t st);
// The GBuffer albedo of every pixel of a hit texture (gbuffer_rt.slang:76-77,92): with floor != 0
// the demodulation albedo max(albedo, 1e-2) (gbuffer.slang:138-140), else float4(albedo, 1).
hipError_t launch_gbuffer_albedo(const DeviceScene& s, const mrt_camera& cam, uint32_t w, uint32_t h,
                                 const mrt_packed_hit* gbuffer, float4* out, int floor, hipStream_t st);
hipError_t launch_gbuffer(const DeviceScene& s, const TraceScratch& ts, const mrt_camera& cam, uint32_t w, uint32_t h,
                          uint32_t mask, mrt_packed_hit* out, hipStream_t st);

struct FrameTiming {
    float ms_trace = 0.f, ms_shade = 0.f, ms_total = 0.f;
    uint32_t trace_launches = 0;
    uint64_t shadow_rays = 0;
};
// One frame of pt over `gbuffer`; out = float4 irradiance per pixel; sig_out optional.
// counters_host (optional) receives every chunk's queue sizes after the frame completes.
// guide (MRT_MODE_GUIDED): the network whose lobes are inferred for every bounce's queue; the
// first chunk's sample-0 paths write training records (pb.grec / pb.gtrain).
// mcpg (MRT_MODE_MCPG): grids and carry buffer; the frame runs one pass per sample.
hipError_t launch_pt_frame(const DeviceScene& s, const TraceScratch& ts, PtArgs a, const mrt_packed_hit* gbuffer,
                           PathBuffers& pb, float4* out, uint32_t* sig_out, hipStream_t st, FrameTiming* timing,
                           uint64_t* continuation_segments, unsigned long long* trav_counters = nullptr,
                           const GuideNet* guide = nullptr, const McpgArgs* mcpg = nullptr);

} // namespace mrt
