// mrt_trace_closest_noalloc_dev: closest-hit ray queries that allocate no device memory.
//
// The trace kernel needs scratch beyond its rays and hits: a work cursor (one word, zero at the
// start of a launch) and the global part of every lane's traversal stack (the entries past the
// LDS part, one column per lane of the persistent grid). mrt_trace_closest_dev takes both from
// the context, which allocates them on first use. This variant takes them from `scratch`, a
// fixed array in device memory that every GPU the library runs on holds until the process ends.
//
// Layout of `scratch` (words): [0, 64) the work cursor and padding, [64, SCRATCH_WORDS) the stack
// overflow columns: entry k of lane t at 64 + k * threads + t, for threads lanes and `entries`
// entries per lane, threads * entries <= SCRATCH_WORDS - 64. The grid is capped at `threads`
// lanes, so a deep tree runs a smaller persistent grid rather than needing more memory.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/merian_amd.h"
#include "mrt_api_internal.h"
#include "mrt_kernels.h"

using namespace mrt;

namespace {

constexpr uint64_t SCRATCH_BYTES = uint64_t(MRT_NOALLOC_SCRATCH_BYTES);
constexpr uint64_t SCRATCH_WORDS = SCRATCH_BYTES / 4;
constexpr uint64_t CURSOR_WORDS = 64;
constexpr int MAX_DEVICES = 64;

// In host code `scratch` names a host-side stand-in, not the device array: device addresses
// come from hipGetSymbolAddress (scratch_ptr[device]).
static __device__ __align__(256) uint32_t scratch[SCRATCH_WORDS];

std::mutex g_mutex;
uint32_t* g_scratch_ptr[MAX_DEVICES] = {};

} // namespace

namespace mrt {

// Looks up the device address of `scratch` on the context's device (the current device), once
// per device. Called by mrt_create.
mrt_status noalloc_setup(mrt_ctx* c) {
    if (c->device < 0 || c->device >= MAX_DEVICES) return fail(c, MRT_ERR_UNSUPPORTED, "device index above 63");
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!g_scratch_ptr[c->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(c, e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch symbol");
        g_scratch_ptr[c->device] = static_cast<uint32_t*>(p);
    }
    return MRT_OK;
}

} // namespace mrt

extern "C" {

MRT_API mrt_status mrt_trace_closest_noalloc_dev(mrt_ctx* c, mrt_scene* s, const mrt_ray* rays, uint64_t n,
                                                 mrt_hit* hits, void* stream) {
    if (!c || !s || (n && (!rays || !hits)))
        return fail(c, MRT_ERR_INVALID, "mrt_trace_closest_noalloc_dev: bad argument");
    if (n > 0xFFFFFFFFull) return fail(c, MRT_ERR_INVALID, "mrt_trace_closest_noalloc_dev: more than 2^32-1 rays");
    if (n == 0) return MRT_OK;
    uint32_t* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mutex);
        if (c->device >= 0 && c->device < MAX_DEVICES) base = g_scratch_ptr[c->device];
    }
    if (!base) return fail(c, MRT_ERR_DEVICE, "mrt_trace_closest_noalloc_dev: scratch not set up on this device");
    const uint32_t need = s->ds.stack_need;
    const uint64_t entries = need > uint32_t(TRACE_LDS_STACK) ? need - TRACE_LDS_STACK : 1u;
    uint64_t threads = (SCRATCH_WORDS - CURSOR_WORDS) / entries / TRACE_BLOCK * TRACE_BLOCK;
    if (threads > trace_max_threads()) threads = trace_max_threads();
    if (threads < uint64_t(TRACE_BLOCK))
        return fail(c, MRT_ERR_UNSUPPORTED,
                    "mrt_trace_closest_noalloc_dev: the scene's BVH needs " + std::to_string(entries) +
                        " overflow stack entries per lane; the fixed scratch of " + std::to_string(SCRATCH_BYTES) +
                        " bytes holds at most " + std::to_string((SCRATCH_WORDS - CURSOR_WORDS) / TRACE_BLOCK) +
                        " for one block");
    TraceScratch ts = scene_scratch(nullptr, s);
    ts.work = base;
    ts.ovf = base + CURSOR_WORDS;
    ts.threads = uint32_t(threads);
    ts.entries = uint32_t(entries);
    // the cursor is cleared by a fill kernel inside launch_trace_rays (no memset node)
    const hipError_t e = launch_trace_rays(s->ds, ts, rays, n, hits, pick_stream(c, stream));
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, "trace kernel (no-alloc)");
}

} // extern "C"
