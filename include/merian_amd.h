/*
 * merian_amd.h — C ABI of the MI355X (gfx950) path-tracing library behind the
 * "Render (Path-traced)" node.
 *
 * Conventions
 *   - Every entry point returns mrt_status (0 = ok, < 0 = error). No C++ exception
 *     crosses this boundary; the reference raises node_error / compilation_failed
 *     inside the graph (src/merian-graph/graph/graph.cpp:385-395) and the node host
 *     maps a negative status onto that.
 *   - Host buffers are owned by the caller; device memory behind the opaque handles is
 *     owned by the library. Entry points suffixed _dev take device pointers that the
 *     caller owns and an optional hipStream_t: NULL = the context's own stream (created
 *     non-blocking, so it does not wait for the HIP null stream); hipStreamLegacy = the HIP
 *     null stream; anything else = that stream.
 *   - One host thread per context (the reference graph is single-threaded,
 *     src/merian-graph/graph/graph.cpp:68-253). Work is asynchronous on the context's
 *     stream; host-buffer entry points synchronise before returning.
 *   - The context's device scratch (traversal stacks, path buffers) is shared by its
 *     trace, gbuffer and pt launches. When a _dev call passes a stream other than the
 *     previous such call's, the library orders the new launch after the previous one with
 *     an event, so calls on different streams serialise rather than race. Inside a stream
 *     capture (hipStreamBeginCapture) that ordering is the caller's: capture every call of
 *     one context on one stream.
 *   - mrt_last_error(ctx) describes the most recent failure on that context.
 *
 * Layouts mirror the reference's packed device structs where a layout is fixed there
 * (PackedVertexData 24 B, PackedHitInfo 16 B); everything else is this library's own.
 */
#ifndef MERIAN_AMD_H
#define MERIAN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define MRT_API __attribute__((visibility("default")))
#else
#define MRT_API
#endif

typedef int32_t mrt_status;
enum {
    MRT_OK = 0,
    MRT_ERR_INVALID = -1,     /* bad argument (null, size, unknown enum) */
    MRT_ERR_DEVICE = -2,      /* HIP runtime error */
    MRT_ERR_BUDGET = -3,      /* HBM budget of the context would be exceeded */
    MRT_ERR_UNSUPPORTED = -4, /* a feature the reference has but this build does not */
    MRT_ERR_NOT_READY = -5,   /* job not finished (mrt_job_poll) */
    MRT_ERR_CANCELLED = -6,   /* job was cancelled */
    MRT_ERR_ALLOC = -7        /* device allocation failed */
};

typedef struct mrt_ctx mrt_ctx;
typedef struct mrt_scene mrt_scene;

/* ------------------------------------------------------------------------------------
 * Scene description (host side, copied by mrt_scene_create)
 * ---------------------------------------------------------------------------------- */

/* = PackedVertexData (include/merian-shaders/scene/scene-data.slangh:75-101): 24 bytes. */
typedef struct mrt_vertex {
    float position[3];
    uint32_t encoded_normal;  /* packSnorm2x16(octahedral_project(n)) */
    uint16_t uv[2];           /* IEEE half bits */
    uint32_t encoded_tangent; /* LSB = handedness, upper bits = oct direction */
} mrt_vertex;

/* GeometryDataFlags (scene-data.slangh:13-41). */
enum {
    MRT_GEOM_PRETRANSFORMED = 0x1,
    MRT_GEOM_FLIP_FACING = 0x8,
    MRT_GEOM_INDEX_UINT32 = 0x40, /* else sequential indices 3i, 3i+1, 3i+2 */
    MRT_GEOM_FLAT_SHADING = 0x80,
    MRT_GEOM_HAS_TANGENTS = 0x100,
    MRT_GEOM_USE_ENV_MAP = 0x200,
    MRT_GEOM_TWO_SIDED = 0x400
};

/* GeometryData (scene-data.slangh:119-129) with buffer addresses replaced by offsets
 * into the scene's shared vertex / index arrays (the reference suballocates shared
 * buffers the same way, src/merian-shaders/scene/scene.cpp:1942-1972). All meshes are
 * world-space (pretransformed); instance transforms are baked at scene build. */
typedef struct mrt_geometry {
    uint32_t vertex_offset;   /* first vertex in mrt_scene_desc.vertices */
    uint32_t index_offset;    /* first uint32 in mrt_scene_desc.indices (3 per triangle) */
    uint32_t primitive_count; /* triangles */
    uint16_t material_id;
    uint16_t flags;           /* MRT_GEOM_* */
    uint32_t instance_mask;   /* 8-bit TLAS instance mask (scene.cpp:1825) */
} mrt_geometry;

/* Material models (material_system.cpp:124-195 registers type ids; here fixed). */
enum { MRT_MAT_DIFFUSE = 0, MRT_MAT_OPENPBR = 1 };
/* OpenPBRMaterial<TRANSMISSION, CLEARCOAT, SHEEN, VOLUME> feature bits
 * (include/merian-shaders/shading/materials/openpbr-material.slang:51-76). */
enum {
    MRT_PBR_TRANSMISSION = 0x1,
    MRT_PBR_CLEARCOAT = 0x2,
    MRT_PBR_SHEEN = 0x4,
    MRT_PBR_VOLUME = 0x8
};

/* One material entry, 128 bytes. DiffuseMaterial uses base_color only
 * (diffuse-material.slang:7-43); OpenPBR uses every field its features enable. */
typedef struct mrt_material {
    uint32_t model;    /* MRT_MAT_* */
    uint32_t features; /* MRT_PBR_* */
    float base_color[3];
    float opacity;
    float metalness;
    float specular_alpha[2]; /* GGX alpha per axis (tangent, bitangent) */
    float specular_weight;
    float specular_ior;
    float emission[3];
    float transmission_weight;
    float transmission_color[3];
    float coat_weight;
    float coat_alpha;
    float coat_ior;
    float sheen_weight;
    float sheen_color[3];
    float sheen_alpha;
    float absorption[3];
    float _pad[3];
} mrt_material;

/* Environment maps (environment-map.slang:13-68). */
enum { MRT_ENV_NONE = 0, MRT_ENV_LATLONG = 1, MRT_ENV_EQUAL_AREA_OCT = 2 };
enum { MRT_WRAP_REPEAT = 0, MRT_WRAP_CLAMP = 1 };
typedef struct mrt_env {
    int32_t type;      /* MRT_ENV_* */
    int32_t wrap;      /* MRT_WRAP_*: lat-long uses repeat (scene.cpp:785-795), pbrt's
                          equal-area map clamp-to-edge (pbrt_scene.cpp:924-969) */
    uint32_t width;
    uint32_t height;
    const float* rgba; /* width*height*4 floats, row-major, row 0 = v 0 */
    float intensity;
    float to_local[9]; /* row-major 3x3, applied to the world direction */
} mrt_env;

typedef struct mrt_scene_desc {
    const mrt_vertex* vertices;
    uint32_t vertex_count;
    const uint32_t* indices;
    uint32_t index_count;
    const mrt_geometry* geometries;
    uint32_t geometry_count;
    const mrt_material* materials;
    uint32_t material_count;
    mrt_env env;
} mrt_scene_desc;

/* Camera as the shader sees it (camera.slang:6-93; written by Camera::write_to,
 * src/merian/utils/camera/camera.cpp:349-362). */
typedef struct mrt_camera {
    float position[3];
    float U[3];
    float V[3];
    float W[3];
    float near_plane;
    float far_plane;
    float jitter[2];
} mrt_camera;

/* Jitter sequences (camera.cpp:318-339). */
enum { MRT_JITTER_NONE = 0, MRT_JITTER_R2 = 1 };

/* ------------------------------------------------------------------------------------
 * Rays and hits
 * ---------------------------------------------------------------------------------- */

typedef struct mrt_ray {
    float origin[3];
    float tmin;
    float direction[3];
    float tmax;
} mrt_ray;

/* Closest hit: triangle id (global, in input order; 0xFFFFFFFF = miss), t, barycentrics
 * of v1 and v2 (the ray-query convention, acceleration-structure.slang:30-35). */
typedef struct mrt_hit {
    uint32_t triangle;
    float t;
    float u;
    float v;
} mrt_hit;

/* = PackedHitInfo (include/merian-shaders/gbuffer.slang:40-61): the "gbuffer" hit
 * texel, 16 bytes. Miss: instance_id = primitive_id = 0xFFFFFFFF. */
typedef struct mrt_packed_hit {
    uint32_t instance_id;
    uint32_t primitive_id;
    uint16_t geometry_index;
    uint16_t instance_index;
    uint16_t barycentrics[2]; /* IEEE half bits, round-to-nearest-even */
} mrt_packed_hit;

/* ------------------------------------------------------------------------------------
 * Render parameters: the "Render (Path-traced)" properties
 * (src/merian-graph/nodes/render_pt/render_pt.cpp:147-187) plus build-side modes.
 * ---------------------------------------------------------------------------------- */
/* A frame runs in passes of whole samples that fit the paths in flight (mrt_tuning.max_paths);
 * an MRT_MODE_NEE pass holds at most 2^31 - 1 paths (its shadow rays name their path in 31
 * bits), so an NEE image of 2^31 pixels or more is refused with MRT_ERR_UNSUPPORTED. */
enum { MRT_MODE_BSDF = 0, MRT_MODE_NEE = 1, MRT_MODE_GUIDED = 2, MRT_MODE_MCPG = 3 };

typedef struct mrt_pt_params {
    int32_t spp;                 /* "samples per pixel" */
    int32_t max_path_length;     /* "max path length" (segments incl. the primary hit) */
    int32_t emission_on_primary; /* "emission on primary" */
    int32_t demodulate_albedo;   /* "demodulate albedo": every sample is divided by the pixel's GBuffer
                                    albedo floored at 1e-2 (render_pt.slang:104-116) */
    uint32_t instance_mask;      /* bits "0".."7" */
    int32_t mode;                /* MRT_MODE_*: BSDF = the reference estimator */
    int32_t record_signature;    /* write a per-pixel hash of the hit ids (parity aid) */
    int32_t count_traversal;     /* fill node_visits / triangle_tests (slower counting kernels) */
} mrt_pt_params;

/* Per-frame statistics (counters kept by the device queues). */
typedef struct mrt_pt_stats {
    uint64_t primary_segments;      /* closest-hit queries of the gbuffer pass */
    uint64_t continuation_segments; /* closest-hit queries of the bounce loop */
    uint64_t shadow_rays;           /* NEE visibility queries (MRT_MODE_NEE/GUIDED) */
    float ms_gbuffer;               /* device time of the gbuffer pass */
    float ms_trace;                 /* device time spent in the intersect kernel */
    float ms_shade;                 /* device time in shading / sampling kernels */
    float ms_total;                 /* device time of the whole frame */
    uint32_t trace_launches;        /* intersect-kernel launches of the frame */
    uint32_t _pad;
    uint64_t node_visits;           /* BVH4 node visits (count_traversal only) */
    uint64_t triangle_tests;        /* ray-triangle tests (count_traversal only) */
} mrt_pt_stats;

/* ------------------------------------------------------------------------------------
 * Context, scene
 * ---------------------------------------------------------------------------------- */

/* Library version / build identity ("gfx950"). */
MRT_API const char* mrt_version(void);

/* Create a context on HIP device `device`. hbm_budget bytes (0 = 90% of free memory)
 * bound every allocation the context makes, as the serving loop's admission limit. */
MRT_API mrt_status mrt_create(int32_t device, uint64_t hbm_budget, mrt_ctx** out);
MRT_API mrt_status mrt_destroy(mrt_ctx* ctx);
MRT_API const char* mrt_last_error(mrt_ctx* ctx);
/* Bytes of device memory the context currently holds. */
MRT_API uint64_t mrt_memory_in_use(mrt_ctx* ctx);
/* Waits for all work on the context's stream. */
MRT_API mrt_status mrt_synchronize(mrt_ctx* ctx);

/* Builds the BVH (binned SAH, 4-wide) and uploads geometry, materials and the
 * environment map; replaces the driver BLAS/TLAS build (scene.cpp:1672-1892). */
MRT_API mrt_status mrt_scene_create(mrt_ctx* ctx, const mrt_scene_desc* desc, mrt_scene** out);
MRT_API mrt_status mrt_scene_destroy(mrt_ctx* ctx, mrt_scene* scene);
/* Scene statistics: triangles, BVH nodes, device bytes. */
MRT_API mrt_status mrt_scene_info(mrt_scene* scene, uint64_t* triangles, uint64_t* nodes,
                                  uint64_t* device_bytes);
/* Geometry-only scene for ray queries: vertex positions float[3*n_vertices] and
 * uint32 triangle indices; two_sided != 0 disables back-face culling. */
MRT_API mrt_status mrt_scene_create_triangles(mrt_ctx* ctx, const float* positions,
                                              uint32_t vertex_count, const uint32_t* indices,
                                              uint32_t triangle_count, int32_t two_sided,
                                              mrt_scene** out);

/* Camera from look-at parameters, as Camera(...) + write_to + advance_jitter
 * (camera.cpp:10-40, 263-266, 318-362). vfov in radians. */
MRT_API mrt_status mrt_camera_look_at(const float eye[3], const float target[3],
                                      const float up[3], float vfov, float aspect,
                                      float near_plane, float far_plane, int32_t jitter,
                                      uint32_t frame, mrt_camera* out);

/* ------------------------------------------------------------------------------------
 * Ray queries: Scene.trace_ray (include/merian-shaders/scene/scene.slang:62-76) with
 * back-face culling and ties at equal t broken by the lowest triangle id.
 * ---------------------------------------------------------------------------------- */
MRT_API mrt_status mrt_trace_closest(mrt_ctx* ctx, mrt_scene* scene, const mrt_ray* rays,
                                     uint64_t n, mrt_hit* hits);
MRT_API mrt_status mrt_trace_closest_dev(mrt_ctx* ctx, mrt_scene* scene, const mrt_ray* d_rays,
                                         uint64_t n, mrt_hit* d_hits, void* stream);
/* As mrt_trace_closest_dev, with the counting kernel: counts[0] = node visits of the scene's
 * tree (4- or 8-wide), counts[1] = ray-triangle tests (host array). Synchronous; for the
 * bandwidth roofline. */
MRT_API mrt_status mrt_trace_count_dev(mrt_ctx* ctx, mrt_scene* scene, const mrt_ray* d_rays, uint64_t n,
                                       mrt_hit* d_hits, uint64_t counts[2]);
/* Allocation-free closest-hit queries. No counterpart in the reference: this entry point is
 * for callers that cannot allocate device memory (code under hipGraph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same results as mrt_trace_closest_dev, bit for bit.
 * Its traversal scratch (a work cursor and the traversal stacks' overflow columns) is a fixed
 * array of MRT_NOALLOC_SCRATCH_BYTES in device memory, held by every GPU the library runs on
 * until the process ends; its address is looked up per device by mrt_create. The persistent grid
 * is capped at what that array holds for the scene's tree depth; a tree too deep for even one
 * block returns MRT_ERR_UNSUPPORTED. Every call on one GPU shares the array: calls that overlap
 * on one GPU (other streams, other contexts, other threads) race and must be ordered by the
 * caller. Calls no hipMalloc, hipFree or allocating library routine. */
#define MRT_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
MRT_API mrt_status mrt_trace_closest_noalloc_dev(mrt_ctx* ctx, mrt_scene* scene, const mrt_ray* d_rays,
                                                 uint64_t n, mrt_hit* d_hits, void* stream);
/* Pinhole rays of a w x h image on the device (camera.slang:45-56), for coherent
 * ray-query workloads. */
MRT_API mrt_status mrt_camera_rays_dev(mrt_ctx* ctx, const mrt_camera* cam, uint32_t width,
                                       uint32_t height, mrt_ray* d_rays, void* stream);

/* ------------------------------------------------------------------------------------
 * GBuffer (Raytraced) + Render (Path-traced)
 * ---------------------------------------------------------------------------------- */

/* Primary rays + closest hit into PackedHitInfo (gbuffer_rt.slang:18-94). */
MRT_API mrt_status mrt_gbuffer_dev(mrt_ctx* ctx, mrt_scene* scene, const mrt_camera* cam,
                                   uint32_t width, uint32_t height, uint32_t instance_mask,
                                   mrt_packed_hit* d_gbuffer, void* stream);

/* The GBuffer's albedo texture for a hit texture (gbuffer_rt.slang:76-77,92; stored as
 * float4(albedo, 1), gbuffer.slang:160-162): per pixel the BSDF's directional albedo towards the
 * camera for a surface hit, the environment's emission for a miss. d_albedo: float4 per pixel. */
MRT_API mrt_status mrt_gbuffer_albedo_dev(mrt_ctx* ctx, mrt_scene* scene, const mrt_camera* cam, uint32_t width,
                                          uint32_t height, const mrt_packed_hit* d_gbuffer, float* d_albedo,
                                          void* stream);

/* One frame of pt (render_pt.slang:92-130): reads the gbuffer, writes irradiance
 * float4(mean rgb, mean yuv-luminance^2) per pixel, row-major. d_signature (optional,
 * uint32 per pixel) receives a hash of every path's hit sequence when
 * params->record_signature is set. stats may be NULL. */
MRT_API mrt_status mrt_pt_render_dev(mrt_ctx* ctx, mrt_scene* scene, const mrt_pt_params* params,
                                     const mrt_camera* cam, uint32_t frame, uint32_t width,
                                     uint32_t height, const mrt_packed_hit* d_gbuffer,
                                     float* d_irradiance, uint32_t* d_signature,
                                     mrt_pt_stats* stats, void* stream);

/* Host-buffer convenience: gbuffer + pt for one frame, results copied to the host. */
MRT_API mrt_status mrt_render_frame(mrt_ctx* ctx, mrt_scene* scene, const mrt_pt_params* params,
                                    const mrt_camera* cam, uint32_t frame, uint32_t width,
                                    uint32_t height, float* irradiance, uint32_t* signature,
                                    mrt_packed_hit* gbuffer, mrt_pt_stats* stats);

/* ------------------------------------------------------------------------------------
 * Autotuning (no counterpart in the reference, whose driver picks its own acceleration
 * structure and dispatch shapes). A scene carries the launch and BVH settings used when
 * it is rendered; mrt_tune measures candidate settings on the caller's view and keeps
 * the fastest. No setting changes results: traversal is exact whatever the tree and the
 * estimator accumulates in sample order whatever the pass size.
 * ---------------------------------------------------------------------------------- */
typedef struct mrt_tuning {
    uint64_t max_paths;       /* paths in flight per wavefront pass (0 = the context's) */
    uint32_t refill_lanes;    /* idle lanes of 64 before a traversal wave fetches new rays */
    uint32_t grid_percent;    /* persistent traversal grid, percent of the resident blocks */
    uint32_t bvh_max_leaf;    /* triangles per BVH leaf, 1..8 */
    float bvh_traversal_cost; /* SAH cost of a node visit relative to a triangle test */
    uint32_t bvh_width;       /* BVH node width: 4 (full-precision boxes, 128-byte nodes) or 8
                                 (8-bit quantised boxes, 80-byte nodes) */
    float ms_per_frame;       /* frame time measured for these settings (mrt_tune) */
    uint32_t leaf_lanes;      /* lanes of 64 left with nothing but a postponed leaf before a
                                 traversal wave runs a leaf-test phase, 1..64 (0 = 16) */
    uint32_t bvh8_order;      /* BVH8 visiting order: 0 = near-first by sorted entry distance,
                                 one stack entry per child; 1 = octant order, one stack group
                                 per node (fewer instructions per node, a node header re-read
                                 per pop: faster on long incoherent traversals; at most 2^24
                                 triangles). Results are the same either way. */
} mrt_tuning;

typedef struct mrt_tune_request {
    mrt_camera camera;        /* the view to tune on */
    uint32_t width, height;
    mrt_pt_params params;     /* spp is capped at 64 for the trial frames */
    float time_budget_ms;     /* wall time the search may take (0 = 20000) */
    uint32_t _pad;
} mrt_tune_request;

MRT_API mrt_status mrt_scene_get_tuning(mrt_scene* scene, mrt_tuning* out);
/* Applies settings; the BVH is rebuilt when its leaf size or traversal cost changes. */
MRT_API mrt_status mrt_scene_set_tuning(mrt_ctx* ctx, mrt_scene* scene, const mrt_tuning* tuning);
/* Coordinate search over BVH (width, leaf size, traversal cost), paths in flight, refill
 * threshold and grid size; the fastest settings stay applied and are returned. */
MRT_API mrt_status mrt_tune(mrt_ctx* ctx, mrt_scene* scene, const mrt_tune_request* request,
                            mrt_tuning* best);

/* ------------------------------------------------------------------------------------
 * Device buffers and the accumulation pass, for node hosts that link only this ABI.
 * ---------------------------------------------------------------------------------- */
/* Device allocation charged against the context's budget. */
MRT_API mrt_status mrt_alloc_dev(mrt_ctx* ctx, uint64_t bytes, void** d_ptr);
MRT_API mrt_status mrt_free_dev(mrt_ctx* ctx, void* d_ptr);
/* Synchronous copies on the context's stream. */
MRT_API mrt_status mrt_download(mrt_ctx* ctx, void* host, const void* d_src, uint64_t bytes);
MRT_API mrt_status mrt_upload(mrt_ctx* ctx, void* d_dst, const void* host, uint64_t bytes);
/* Running mean of float4 images (Accumulate, accumulate.slang:101 with a plain history):
 * d_accum = count == 0 ? d_src : d_accum + (d_src - d_accum) / (count + 1), per component. */
MRT_API mrt_status mrt_accumulate_dev(mrt_ctx* ctx, float* d_accum, const float* d_src, uint32_t pixels,
                                      uint32_t count, void* stream);

/* ------------------------------------------------------------------------------------
 * Serving loop: continuous batching of render jobs (gbuffer + pt of one frame each).
 *
 * No counterpart in the reference, which renders one frame per Graph::run
 * (src/merian-graph/graph/graph.cpp:68-253). A server owns a scheduler thread, up to
 * max_slots jobs in flight (one HIP stream each, so their kernels overlap on the
 * device) and an HBM budget. Jobs are admitted in submission order as soon as a slot
 * is free and their buffers fit the budget; jobs that arrive while others run join
 * without waiting for them. Each job's launches are captured in a hipGraph, replayed
 * through the slot's executable graph (updated in place when the shape repeats).
 * Per-job RNG depends on (pixel, frame, sample) only, so a job's result is the same
 * alone and in any mix, bit for bit.
 * ---------------------------------------------------------------------------------- */
typedef struct mrt_server mrt_server;

typedef struct mrt_server_config {
    uint64_t hbm_budget;         /* bytes for the buffers of jobs in flight; reserved from the
                                    context's budget at creation (0 = half of what is left) */
    uint32_t max_slots;          /* jobs in flight at once (0 = 4, at most 16) */
    uint32_t max_paths_per_job;  /* paths in flight per job (0 = 2^24); more spp -> chunks */
    int32_t use_graphs;          /* 1: capture each job in a hipGraph; 0: launch directly */
    uint32_t _pad;
} mrt_server_config;

typedef struct mrt_job_desc {
    mrt_scene* scene;            /* resident scene (must outlive the job) */
    mrt_camera camera;
    mrt_pt_params params;        /* mode BSDF or NEE; record_signature/count_traversal ignored */
    uint32_t width, height;
    uint32_t frame;              /* RNG frame index of this job */
    uint32_t _pad;
    float* irradiance;           /* host float4[width*height], written when the job is done */
} mrt_job_desc;

enum {
    MRT_JOB_QUEUED = 0,
    MRT_JOB_RUNNING = 1,
    MRT_JOB_DONE = 2,
    MRT_JOB_CANCELLED = 3,
    MRT_JOB_FAILED = 4
};

typedef struct mrt_job_status {
    int32_t state;               /* MRT_JOB_* */
    mrt_status error;            /* status of a failed job */
    uint64_t segments;           /* closest-hit segments traced (primary + continuation) */
    uint64_t device_bytes;       /* buffers the job held while in flight */
    double queued_ms;            /* submit -> admitted */
    double latency_ms;           /* submit -> result in the caller's buffer */
} mrt_job_status;

typedef struct mrt_server_stats {
    uint64_t submitted, completed, cancelled, failed;
    uint64_t segments;           /* over completed jobs */
    double p50_ms, p99_ms, max_ms, mean_ms; /* latency of completed jobs (nearest rank) */
    double wall_ms;              /* first submit -> last completion */
    uint64_t peak_bytes;         /* most buffer bytes held at once */
    uint32_t peak_in_flight;     /* most jobs in flight at once */
    uint32_t graph_instantiations, graph_updates, _pad;
} mrt_server_stats;

MRT_API mrt_status mrt_server_create(mrt_ctx* ctx, const mrt_server_config* config,
                                     mrt_server** out);
/* Cancels queued jobs, waits for jobs in flight, releases the reservation. */
MRT_API mrt_status mrt_server_destroy(mrt_server* server);
MRT_API const char* mrt_server_last_error(mrt_server* server);
/* Device bytes a job needs while in flight (its admission charge). */
MRT_API mrt_status mrt_job_bytes(mrt_server* server, const mrt_job_desc* job, uint64_t* bytes);
/* Queues a job; thread-safe. MRT_ERR_BUDGET if it could never fit the server's budget. */
MRT_API mrt_status mrt_job_submit(mrt_server* server, const mrt_job_desc* job, uint64_t* job_id);
/* MRT_OK with the status filled in (any state); MRT_ERR_INVALID for an unknown id. */
MRT_API mrt_status mrt_job_poll(mrt_server* server, uint64_t job_id, mrt_job_status* status);
/* Waits until the job is done, cancelled or failed, or timeout_ms passes (< 0: no limit;
 * MRT_ERR_NOT_READY on time-out). */
MRT_API mrt_status mrt_job_wait(mrt_server* server, uint64_t job_id, int32_t timeout_ms,
                                mrt_job_status* status);
/* A queued job is dropped; a job in flight finishes on the device but its result is not
 * written. MRT_ERR_INVALID if the job is unknown or already finished. */
MRT_API mrt_status mrt_job_cancel(mrt_server* server, uint64_t job_id);
MRT_API mrt_status mrt_server_stats_get(mrt_server* server, mrt_server_stats* stats);
/* Clears the latency history and counters (jobs in flight keep running). */
MRT_API mrt_status mrt_server_stats_reset(mrt_server* server);

/* ------------------------------------------------------------------------------------
 * Learned path guiding (MRT_MODE_GUIDED).
 *
 * No counterpart in the reference's pt; the closest analogues are the MCPG renderer's
 * guided one-sample MIS (render_pt_mcpg.slang:104-136,181) and its von Mises-Fisher lobes
 * (von-mises-fisher.slang:29-45). A network maps a shading position and the outgoing
 * direction to an 8-lobe vMF mixture over incident directions:
 *   input   64 = multiresolution hash grid (8 levels x 4 features, fp16 tables; a cell's
 *                two x-neighbour entries are adjacent, one 16-byte load per cell edge)
 *                + 16 real spherical harmonics (degree <= 3) of the direction to the viewer
 *                + 16 zeros
 *   layers  4 hidden layers of width 64 (ReLU), linear output of 40 = 8 x (logit, log kappa,
 *           mu.xyz); no biases; fp16 weights and activations, fp32 accumulation (MFMA)
 * It is trained online on the frame's own radiance samples (loss: -sum w log q(dir) /
 * sum w, w = incident luminance / sampling pdf) with Adam. At an eligible vertex the
 * renderer samples the mixture with probability guide_fraction, else the BSDF, and weighs
 * by the one-sample MIS pdf, so the estimator stays unbiased.
 * ---------------------------------------------------------------------------------- */
typedef struct mrt_guide mrt_guide;

enum { MRT_GUIDE_LEVELS = 8, MRT_GUIDE_FEATURES = 4, MRT_GUIDE_WIDTH = 64, MRT_GUIDE_LAYERS = 5,
       MRT_GUIDE_OUTPUTS = 40, MRT_GUIDE_LOBES = 8 };

typedef struct mrt_guide_config {
    float bbox_min[3], bbox_max[3]; /* position normalisation box (all zero: scene bounds) */
    uint32_t seed;                  /* weight initialisation */
    uint32_t log2_table;            /* hash-table entries per level (0 = 14: the eight levels'
                                       tables, 1 MB, stay in L2 while inference gathers from
                                       them; at the 1080p workload 2^14 was both faster and,
                                       with short online training, more accurate than larger
                                       tables); must hold max_resolution + 1 entries */
    uint32_t base_resolution;       /* coarsest grid resolution (0 = 16) */
    uint32_t max_resolution;        /* finest grid resolution (0 = 2048) */
    uint32_t train_batch;           /* records per training step (0 = 65536) */
    uint32_t train_steps;           /* training steps per rendered frame (0 = 4) */
    uint32_t max_records;           /* training records collected per frame (0 = 2^20) */
    float learning_rate;            /* Adam step size (0 = 5e-3) */
    float guide_fraction;           /* probability of sampling the guide (0 = 0.5) */
    uint32_t guided_bounces;        /* vertices 0 .. guided_bounces-1 may sample the guide; later
                                       ones sample the BSDF only and skip inference (0 = all) */
} mrt_guide_config;

/* Per-level grid resolutions and table size the guide uses (host-side computation). */
MRT_API mrt_status mrt_guide_layout(const mrt_guide_config* config, uint32_t resolutions[MRT_GUIDE_LEVELS],
                                    uint32_t* table_entries);
MRT_API mrt_status mrt_guide_create(mrt_ctx* ctx, mrt_scene* scene, const mrt_guide_config* config,
                                    mrt_guide** out);
MRT_API mrt_status mrt_guide_destroy(mrt_ctx* ctx, mrt_guide* guide);
/* Position normalisation in use: p' = clamp((p - offset) * scale, 0, 1) per axis. */
MRT_API mrt_status mrt_guide_normalisation(mrt_guide* guide, float offset[3], float scale[3]);
/* Parameters as fp32 (host arrays): weights [5][64][64] (row = output; layer 5 rows >= 40 are
 * zero), grid [8][table][4]. Either pointer may be NULL. */
MRT_API mrt_status mrt_guide_get_params(mrt_ctx* ctx, mrt_guide* guide, float* weights, float* grid);
/* Replaces the parameters (rounded to fp16 for the kernels) and clears the Adam state. */
MRT_API mrt_status mrt_guide_set_params(mrt_ctx* ctx, mrt_guide* guide, const float* weights,
                                        const float* grid);
/* Network outputs (40 fp32 per query) for host positions/directions (float[3*n] each). */
MRT_API mrt_status mrt_guide_forward(mrt_ctx* ctx, mrt_guide* guide, const float* positions,
                                     const float* directions, uint32_t n, float* outputs);
/* Training records: float[12*n] = (position.xyz, weight, wo.xyz, 0, dir.xyz, 0); weight =
 * incident luminance / sampling pdf. Gradients of the loss (sum over records, scaled by
 * n / sum of weights) without an update: weights [5][64][64], grid [8][table][4] (host
 * fp32, either may be NULL); *loss = mean weighted -log q. */
MRT_API mrt_status mrt_guide_gradient(mrt_ctx* ctx, mrt_guide* guide, const float* records,
                                      uint32_t n, float* d_weights, float* d_grid, float* loss);
/* `steps` Adam steps over consecutive batches of the host records; *loss = last step's loss. */
MRT_API mrt_status mrt_guide_train(mrt_ctx* ctx, mrt_guide* guide, const float* records,
                                   uint32_t n, uint32_t steps, float* loss);
/* One guided frame (MRT_MODE_GUIDED): as mrt_pt_render_dev, sampling the guide at eligible
 * vertices; with train != 0 the frame's records then train the guide (train_steps steps), in
 * an order fixed by (pixel, vertex), so training is deterministic. At most 2^26 pixels
 * (MRT_ERR_UNSUPPORTED beyond). */
MRT_API mrt_status mrt_guided_render_dev(mrt_ctx* ctx, mrt_scene* scene, mrt_guide* guide,
                                         const mrt_pt_params* params, const mrt_camera* cam,
                                         uint32_t frame, uint32_t width, uint32_t height,
                                         const mrt_packed_hit* d_gbuffer, float* d_irradiance,
                                         int32_t train, mrt_pt_stats* stats, void* stream);

/* ------------------------------------------------------------------------------------
 * Markov-chain path guiding: the surface transport of the "Render (Path-traced, MCPG)" node
 * (src/merian-graph/nodes/render_pt_mcpg/render_pt_mcpg.cpp:204-258,
 * render_pt_mcpg.slang:39-244). Guiding state persists across frames in two device hash
 * grids (utils/hash-grid.slang):
 *   - the MC grid: per cell a 32-byte Markov-chain state (mc.slang:15-128: weighted mean of
 *     the hit points a vertex's good samples reached, mean cosine, count) from which a von
 *     Mises-Fisher lobe is derived; a scattering vertex resamples one of `mc_samples`
 *     stochastic lookups (RIS) and samples its lobe with probability p_guide, else the BSDF,
 *     weighing by the mixture of all looked-up lobes and the BSDF (stochastic MIS);
 *   - the light cache: per cell a running mean of incident radiance (half3 + count,
 *     hashed-irradiance-cache.slang) that stands in for the incident radiance of a
 *     non-emitting hit when a chain is updated.
 * The samples of a pixel run one after another on one RNG stream per pixel and frame
 * (pcg3d16(pixel, frame)), as the node's per-pixel loop does, so a frame runs spp passes of
 * one path per pixel. Paths of one pass update the grids concurrently, as the reference's
 * threads do; with one path per pass (a 1x1 image) the sequence is the oracle's exactly.
 * Scenes are static here: motion vectors are zero and prev_pos == pos. The volume pass and
 * the debug outputs of the node are not part of this path. mrt_pt_params.mode is ignored.
 * ---------------------------------------------------------------------------------- */
typedef struct mrt_mcpg mrt_mcpg;

enum { MRT_MCPG_MIS = 0, MRT_MCPG_MIS_ROUGHNESS = 1, MRT_MCPG_MAX_SAMPLES = 8 };

typedef struct mrt_mcpg_grid_config {
    uint32_t buffer_size;            /* slots ("adaptive grid buf size" / "LC buffer size"), <= 2^28 */
    uint32_t probe_count;            /* slots probed before evicting (1..32) */
    uint32_t locality_bits;          /* 2^n-wide cell tiles in Morton order (0..5; 0 = scatter) */
    uint32_t stochastic_interpolation; /* light cache: jitter the cell per lookup (the MC grid
                                          always jitters cell and level) */
    float tan_alpha_half;            /* "grid tan(alpha/2)": cell size per unit camera distance */
    float level_bias;                /* "grid level bias" */
    float distribution_dimension;    /* "grid distribution dimension" */
    uint32_t _reserved;
} mrt_mcpg_grid_config;

typedef struct mrt_mcpg_config {
    mrt_mcpg_grid_config mc;         /* the "mc" child properties */
    mrt_mcpg_grid_config lc;         /* the "lc" child properties */
    int32_t mc_samples;              /* "MC samples" (0..MRT_MCPG_MAX_SAMPLES) */
    float p_guiding;                 /* "guiding prob" */
    int32_t sampling_type;           /* "directional sampling type": MRT_MCPG_MIS(_ROUGHNESS) */
    float guiding_alpha_threshold;   /* "roughness threshold" squared */
    float dir_guide_prior;           /* "ML prior" */
    int32_t missing_light_heuristic; /* "missing light heuristic" */
    int32_t use_light_cache_tail;    /* "surf: use LC" */
    float lc_min_pdf;                /* "LC min pdf" */
    int32_t bsdf_support_only;       /* no counterpart in the node (0 there): 1 counts a guided
                                        direction only where the BSDF's own sampling pdf is > 0,
                                        the domain the plain pt integrates over (the GGX BTDF
                                        evaluates to non-zero outside it) */
} mrt_mcpg_config;

/* The node's defaults (render_pt_mcpg.hpp:125-152, hash_grid.hpp:42-44). */
MRT_API void mrt_mcpg_default_config(mrt_mcpg_config* out);
/* Allocates both grids (zeroed, as the node's reset on its first frame). */
MRT_API mrt_status mrt_mcpg_create(mrt_ctx* ctx, const mrt_mcpg_config* config, mrt_mcpg** out);
/* Waits for the device (frames on any stream may still use the grids), then frees them. */
MRT_API mrt_status mrt_mcpg_destroy(mrt_ctx* ctx, mrt_mcpg* mcpg);
/* Clears both grids and forgets the previous camera (RenderMCPG::process at iteration 0). */
MRT_API mrt_status mrt_mcpg_reset(mrt_ctx* ctx, mrt_mcpg* mcpg);
/* One MCPG frame over a gbuffer (as mrt_pt_render_dev): d_irradiance = per pixel mean
 * radiance.rgb, mean luminance^2. `frame` seeds the RNG and stamps grid slots; the
 * previous camera is the one of the last frame rendered with this state (the current one
 * after a reset). Frames of one mrt_mcpg are stream-ordered by the caller: two frames of the
 * same state in flight on different streams race on its grids and carry buffer. */
MRT_API mrt_status mrt_mcpg_render_dev(mrt_ctx* ctx, mrt_scene* scene, mrt_mcpg* mcpg,
                                       const mrt_pt_params* params, const mrt_camera* cam,
                                       uint32_t frame, uint32_t width, uint32_t height,
                                       const mrt_packed_hit* d_gbuffer, float* d_irradiance,
                                       mrt_pt_stats* stats, void* stream);
/* Host copies of the grids after the device has drained (any pointer may be NULL): keys
 * uint32[2*size] (key, stamp), chain states 32 bytes per slot (w_tgt.xyz, sum_w, w_cos,
 * mv half3, N u16, T), light-cache entries 8 bytes per slot (irr half3, N u16). */
MRT_API mrt_status mrt_mcpg_read_grids(mrt_ctx* ctx, mrt_mcpg* mcpg, uint32_t* mc_keys, void* mc_states,
                                       uint32_t* lc_keys, void* lc_entries);

/* ------------------------------------------------------------------------------------
 * Graph descriptions (host only, no device work): merian's graph JSON (schema v4, v1-v3
 * read), fragment merge, includes and command-line overrides, as GraphDescription does
 * (src/merian-graph/graph/graph_description.cpp:119-857). Strings are UTF-8 JSON; every
 * returned string is allocated by the library and released with mrt_string_free.
 * search_dirs is a ':'-separated list (may be NULL). errors (may be NULL) receives the
 * messages of a failed call, one per line.
 * ---------------------------------------------------------------------------------- */
/* GraphDescription::apply_cli: MRT_OK when every override applied, MRT_ERR_INVALID with
 * the (partially) modified config otherwise. */
MRT_API mrt_status mrt_graph_apply_cli(const char* config_json, int32_t argc, const char* const* argv,
                                       const char* search_dirs, char** out_json, char** errors);
/* GraphDescription::resolve_includes. */
MRT_API mrt_status mrt_graph_resolve_includes(const char* config_json, const char* search_dirs,
                                              char** out_json, char** errors);
/* GraphDescription::merge_into(base, overwrite). */
MRT_API mrt_status mrt_graph_merge(const char* base_json, const char* overwrite_json, char** out_json);
/* GraphDescription::from_json + to_json: any schema version in, v4 out. */
MRT_API mrt_status mrt_graph_store(const char* graph_json, char** out_json, char** errors);
/* GraphDescription::cli_help. */
MRT_API mrt_status mrt_graph_cli_help(const char* config_json, const char* search_dirs, char** out_text);
MRT_API void mrt_string_free(char* s);

/* pbrt-v4 scenes as the "PBRT Scene" node loads them (src/merian-shaders/scene/pbrt_scene.cpp).
 * mrt_pbrt_inspect (host only): JSON with the geometry, vertex and triangle counts, every
 * material entry, the environment, the camera (eye, forward, up, vfov radians, width,
 * height) and the loader's warnings. */
MRT_API mrt_status mrt_pbrt_inspect(const char* path, char** out_json, char** errors);
/* The loaded scene as a host description (no device work), owned by the library until
 * mrt_pbrt_desc_free: what mrt_pbrt_scene_create uploads, for callers that inspect or edit a
 * scene before uploading it. The camera is the file's (camera_from_file) or the default camera
 * framing the scene's bounds (pbrt_scene.cpp:994-1003), whose image size is the caller's and whose
 * aspect ratio is fixed. */
typedef struct mrt_pbrt_desc {
    mrt_scene_desc scene;
    int32_t has_camera;
    int32_t camera_from_file;
    float eye[3], forward[3], up[3];
    float vfov;             /* radians, vertical */
    float aspect;           /* default camera: fixed aspect ratio; 0 = the image's */
    uint32_t width, height; /* the file's film resolution */
    uint32_t warning_count; /* loader warnings (listed by mrt_pbrt_inspect) */
} mrt_pbrt_desc;
MRT_API mrt_status mrt_pbrt_load_desc(const char* path, mrt_pbrt_desc** out, char** errors);
MRT_API void mrt_pbrt_desc_free(mrt_pbrt_desc* desc);
/* Loads and uploads the scene (as mrt_scene_create). */
MRT_API mrt_status mrt_pbrt_scene_create(mrt_ctx* ctx, const char* path, mrt_scene** out);
/* The graph runner (merian-graph-run's command line over this library's node host);
 * returns the process exit code. */
MRT_API int32_t mrt_graph_run_main(int32_t argc, const char* const* argv);

/* ------------------------------------------------------------------------------------
 * Node plugin ABI (include/merian-graph/plugin/node_plugin.hpp:17-29). The registry
 * argument is this library's node registry (merian_amd/host), not the Vulkan one.
 * ---------------------------------------------------------------------------------- */
MRT_API uint32_t merian_node_plugin_abi_version(void);
/* Registers "PBRT Scene", "GBuffer (Raytraced)", "Render (Path-traced)", "Render (Path-traced,
 * MCPG)", "Accumulate" and "Image Write" into a node registry of this library's graph host (mrt::graph::NodeRegistry;
 * the reference passes merian::NodeRegistry& with the same calling convention). */
typedef struct mrt_node_registry mrt_node_registry;
MRT_API void merian_register_nodes(mrt_node_registry* registry);

#ifdef __cplusplus
}
#endif

#endif /* MERIAN_AMD_H */
