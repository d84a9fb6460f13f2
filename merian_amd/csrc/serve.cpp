// Serving loop of include/merian_amd.h (mrt_server_*, mrt_job_*): continuous batching of
// render jobs. A scheduler thread admits queued jobs in submission order into free slots
// while their buffers fit the HBM budget (serve_policy.h), captures each admitted job's
// launches (gbuffer + pt frame) into a hipGraph replayed on the slot's own stream, and
// polls the slots' completion events. Jobs in different slots run concurrently on the
// device; a job that arrives while others run starts as soon as a slot and memory free up.
//
// The reference has no serving loop: merian renders one frame per Graph::run
// (src/merian-graph/graph/graph.cpp:68-253) on one queue.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/merian_amd.h"
#include "mrt_api_internal.h"
#include "mrt_kernels.h"
#include "serve_policy.h"

using namespace mrt;

namespace {

constexpr uint32_t MAX_SLOTS = 16;
constexpr size_t KEEP_FINISHED = 1u << 16; // finished job records kept for poll/wait

struct Job {
    uint64_t id = 0;
    mrt_job_desc d{};
    JobLayout layout{};
    int32_t state = MRT_JOB_QUEUED;
    mrt_status error = MRT_OK;
    bool cancel = false;
    uint64_t segments = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
};

// Device buffers of one job in flight (stream-ordered allocations on the slot's stream).
struct JobBuffers {
    mrt_packed_hit* gbuf = nullptr;
    float4* out = nullptr;
    PathBuffers pb{};
    TraceScratch ts{};
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t* h_counters = nullptr; // pinned
    uint64_t h_counters_cap = 0;
    void* h_result = nullptr; // pinned staging of the job's irradiance
    uint64_t h_result_cap = 0;
    char* block = nullptr;   // device block the slot's jobs run in, kept from job to job
    uint64_t block_bytes = 0;
    double t_alloc_ms = 0; // host time of the last job's block resize (MRT_SERVE_TRACE)
    Job* job = nullptr;
    JobBuffers buf;
};

double now_ms(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

} // namespace

struct mrt_server {
    mrt_ctx* ctx = nullptr;
    int device = 0;
    mrt_server_config cfg{};
    uint64_t reserved = 0;
    uint32_t trace_threads = 0;
    std::chrono::steady_clock::time_point t0;
    // the server's own stream-ordered pool: job buffers come from it and return to it, it keeps
    // at most the server's budget of freed memory, and it is trimmed and destroyed with the server
    hipMemPool_t pool = nullptr;

    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::thread th;
    bool stop = false;

    Admission adm{0, 1};
    std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
    std::deque<uint64_t> finished;
    uint64_t next_id = 1;
    std::vector<Slot> slots;

    // statistics (guarded by m)
    uint64_t submitted = 0, completed = 0, cancelled = 0, failed = 0, segments = 0;
    std::vector<double> latencies;
    double first_submit = -1, last_done = 0;
    uint64_t peak_bytes = 0;
    uint64_t held = 0; // bytes of the slots' blocks (written by the scheduler thread)
    uint32_t peak_in_flight = 0, graph_inst = 0, graph_upd = 0;
    std::string err;
};

namespace {

JobShape shape_of(const mrt_server* s, const mrt_job_desc& d) {
    JobShape sh{};
    sh.npix = uint64_t(d.width) * d.height;
    sh.spp = uint32_t(d.params.spp > 0 ? d.params.spp : 0);
    sh.L = uint32_t(d.params.max_path_length);
    sh.nee = d.params.mode == MRT_MODE_NEE;
    sh.demod = d.params.demodulate_albedo != 0;
    sh.max_paths = sh.nee ? std::min<uint64_t>(s->cfg.max_paths_per_job, NEE_MAX_PATHS) : s->cfg.max_paths_per_job;
    sh.trace_threads = s->trace_threads;
    const uint32_t need = d.scene->ds.stack_need;
    sh.stack_entries = need > uint32_t(TRACE_LDS_STACK) ? need - TRACE_LDS_STACK : 1u;
    return sh;
}

mrt_status validate(const mrt_job_desc* d, std::string* why) {
    if (!d || !d->scene || !d->irradiance) {
        *why = "job: null scene or irradiance buffer";
        return MRT_ERR_INVALID;
    }
    const mrt_pt_params& p = d->params;
    if (p.spp < 0 || p.max_path_length < 1 || p.max_path_length > 64) {
        *why = "job: samples per pixel < 0 or max path length outside [1, 64]";
        return MRT_ERR_INVALID;
    }
    if (p.mode != MRT_MODE_BSDF && p.mode != MRT_MODE_NEE) {
        *why = "job: render mode not supported by the server";
        return MRT_ERR_UNSUPPORTED;
    }
    const uint64_t npix = uint64_t(d->width) * d->height;
    if (npix == 0 || npix > (1ull << 26)) {
        *why = "job: image size must be in [1, 2^26] pixels";
        return MRT_ERR_INVALID;
    }
    return MRT_OK;
}

// The job's buffers carved from the slot's block in the order job_layout counts them (each
// 256-byte aligned); the block holds at least the job's size class.
hipError_t carve_job(Slot& sl, const Job& j, const JobShape& sh) {
    JobBuffers& b = sl.buf;
    b = JobBuffers{};
    const JobLayout& l = j.layout;
    if (!sl.block || sl.block_bytes < l.bytes) return hipErrorInvalidValue;
    char* base = sl.block;
    uint64_t off = 0;
    auto A = [&](uint64_t bytes) -> void* {
        void* p = base + off;
        off += round256(bytes);
        return p;
    };
    const uint64_t P = l.paths, N = sh.npix;
    b.gbuf = static_cast<mrt_packed_hit*>(A(16 * N));
    b.out = static_cast<float4*>(A(16 * N));
    b.pb.state = static_cast<float4*>(A(32 * P));
    b.pb.sig = static_cast<uint32_t*>(A(4 * P));
    for (int k = 0; k < 2; k++) {
        b.pb.q_o[k] = static_cast<float4*>(A(16 * P));
        b.pb.q_d[k] = static_cast<float4*>(A(16 * P));
        b.pb.q_t[k] = static_cast<float4*>(A(16 * P));
        b.pb.q_r[k] = static_cast<float4*>(A(16 * P));
    }
    b.pb.hits = static_cast<float4*>(A(16 * P));
    if (sh.nee) {
        b.pb.sq_o = static_cast<float4*>(A(16 * P));
        b.pb.sq_d = static_cast<float4*>(A(16 * P));
        b.pb.sq_c = static_cast<float4*>(A(16 * P));
    }
    b.pb.counters = static_cast<uint32_t*>(A(4 * l.counters));
    b.pb.accum = static_cast<float4*>(A(16 * N));
    b.pb.sig_accum = static_cast<uint32_t*>(A(4 * N));
    if (sh.demod) b.pb.demod = static_cast<float4*>(A(16 * N));
    b.pb.capacity = P;
    b.ts.ovf = static_cast<uint32_t*>(A(uint64_t(l.threads) * std::max<uint32_t>(1, sh.stack_entries) * 4));
    b.ts.work = static_cast<uint32_t*>(A(64));
    b.ts.threads = l.threads;
    b.ts.entries = sh.stack_entries;
    const TraceScratch tuned = scene_scratch(nullptr, j.d.scene);
    b.ts.refill = tuned.refill;
    b.ts.grid_pct = tuned.grid_pct;
    b.ts.leaf_lanes = tuned.leaf_lanes;
    // job_layout counts the same buffers: the carve never passes the block
    return off <= l.bytes ? hipSuccess : hipErrorInvalidValue;
}

// Returns the slot's block to the server's pool (stream-ordered after the slot's work).
hipError_t release_block(Slot& sl) {
    hipError_t e = hipSuccess;
    if (sl.block) e = hipFreeAsync(sl.block, sl.stream);
    sl.block = nullptr;
    sl.block_bytes = 0;
    return e;
}

// Makes the slot's block exactly `bytes` (a size class) unless it is that already.
hipError_t resize_block(hipMemPool_t pool, Slot& sl, uint64_t bytes) {
    if (sl.block && sl.block_bytes == bytes) return hipSuccess;
    hipError_t e = release_block(sl);
    if (e != hipSuccess) return e;
    e = hipMallocFromPoolAsync(reinterpret_cast<void**>(&sl.block), bytes, pool, sl.stream);
    if (e != hipSuccess) {
        sl.block = nullptr;
        return e;
    }
    sl.block_bytes = bytes;
    return hipSuccess;
}

// Enqueues one job's frame on the slot's stream (captured in a graph when enabled).
hipError_t enqueue_frame(mrt_server* s, Slot& sl, const Job& j) {
    const mrt_job_desc& d = j.d;
    PtArgs a{};
    a.cam = d.camera;
    a.width = d.width;
    a.height = d.height;
    a.npix = d.width * d.height;
    a.frame = d.frame;
    a.spp = d.params.spp;
    a.max_path_length = d.params.max_path_length;
    a.emission_on_primary = d.params.emission_on_primary;
    a.demodulate = d.params.demodulate_albedo ? 1 : 0;
    a.record_signature = 0;
    a.instance_mask = d.params.instance_mask & 0xFFu;
    a.mode = d.params.mode;
    const DeviceScene& ds = d.scene->ds;
    auto launches = [&]() -> hipError_t {
        hipError_t e = launch_gbuffer(ds, sl.buf.ts, d.camera, d.width, d.height, a.instance_mask, sl.buf.gbuf,
                                      sl.stream);
        if (e != hipSuccess) return e;
        return launch_pt_frame(ds, sl.buf.ts, a, sl.buf.gbuf, sl.buf.pb, sl.buf.out, nullptr, sl.stream, nullptr,
                               nullptr, nullptr);
    };
    if (!s->cfg.use_graphs) return launches();
    hipError_t e = hipStreamBeginCapture(sl.stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    const hipError_t le = launches();
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(sl.stream, &g);
    if (le != hipSuccess) {
        if (g) (void)hipGraphDestroy(g); // the launch error is the one reported
        return le;
    }
    if (e != hipSuccess) return e;
    bool updated = false;
    if (sl.exec) {
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult res;
        updated = hipGraphExecUpdate(sl.exec, g, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
        if (!updated) {
            (void)hipGetLastError(); // clear the failed update's error
            e = hipGraphExecDestroy(sl.exec);
            sl.exec = nullptr;
            if (e != hipSuccess) {
                (void)hipGraphDestroy(g);
                return e;
            }
        }
    }
    if (!updated) e = hipGraphInstantiate(&sl.exec, g, nullptr, nullptr, 0);
    const hipError_t ge = hipGraphDestroy(g);
    if (e == hipSuccess) e = ge;
    if (e != hipSuccess) {
        if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
        sl.exec = nullptr;
        return e;
    }
    {
        std::lock_guard<std::mutex> lk(s->m);
        (updated ? s->graph_upd : s->graph_inst)++;
    }
    return hipGraphLaunch(sl.exec, sl.stream);
}

hipError_t start_job(mrt_server* s, Slot& sl, Job& j) {
    const JobShape sh = shape_of(s, j.d);
    const uint64_t nz = j.layout.n_chunks * uint64_t(sh.L + 1);
    if (nz > sl.h_counters_cap) {
        hipError_t e = sl.h_counters ? hipHostFree(sl.h_counters) : hipSuccess;
        sl.h_counters = nullptr;
        sl.h_counters_cap = 0;
        if (e != hipSuccess) return e;
        e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_counters), nz * 4, hipHostMallocDefault);
        if (e != hipSuccess) return e;
        sl.h_counters_cap = nz;
    }
    // the result goes device -> pinned staging on the slot stream, so the completion event
    // covers it and the scheduler thread only copies host memory (no pageable DMA)
    const uint64_t out_bytes = 16ull * j.d.width * j.d.height;
    if (out_bytes > sl.h_result_cap) {
        hipError_t e = sl.h_result ? hipHostFree(sl.h_result) : hipSuccess;
        sl.h_result = nullptr;
        sl.h_result_cap = 0;
        if (e != hipSuccess) return e;
        e = hipHostMalloc(&sl.h_result, out_bytes, hipHostMallocDefault);
        if (e != hipSuccess) return e;
        sl.h_result_cap = out_bytes;
    }
    hipError_t e = carve_job(sl, j, sh);
    if (e == hipSuccess) e = enqueue_frame(s, sl, j);
    // queue sizes and the result -> pinned host memory, then the completion event
    if (e == hipSuccess)
        e = hipMemcpyAsync(sl.h_counters, sl.buf.pb.counters, nz * 4, hipMemcpyDeviceToHost, sl.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(sl.h_result, sl.buf.out, out_bytes, hipMemcpyDeviceToHost, sl.stream);
    if (e == hipSuccess) e = hipEventRecord(sl.done, sl.stream);
    return e;
}

void finish_job(mrt_server* s, Slot& sl, hipError_t status) {
    Job& j = *sl.job;
    mrt_status result = MRT_OK;
    if (status == hipSuccess) {
        const uint32_t L = uint32_t(j.d.params.max_path_length), per = L + 1;
        uint64_t segs = uint64_t(j.d.width) * j.d.height;
        if (j.d.params.spp > 0) // spp = 0 traces the gbuffer only and leaves the counters unset
            for (uint64_t c = 0; c < j.layout.n_chunks; c++)
                for (uint32_t b = 1; b < L; b++) segs += sl.h_counters[c * per + b];
        j.segments = segs;
        bool cancel;
        {
            std::lock_guard<std::mutex> lk(s->m);
            cancel = j.cancel;
        }
        if (!cancel) std::memcpy(j.d.irradiance, sl.h_result, 16ull * j.d.width * j.d.height);
    } else {
        result = MRT_ERR_DEVICE;
    }
    sl.buf = JobBuffers{}; // the block stays with the slot for its next job
    std::lock_guard<std::mutex> lk(s->m);
    j.t_done = now_ms(s->t0);
    if (result != MRT_OK) {
        j.state = MRT_JOB_FAILED;
        j.error = result;
        s->failed++;
        s->err = std::string("job ") + std::to_string(j.id) + ": " +
                 (status != hipSuccess ? hipGetErrorString(status) : "result copy failed");
    } else if (j.cancel) {
        j.state = MRT_JOB_CANCELLED;
        j.error = MRT_ERR_CANCELLED;
        s->cancelled++;
    } else {
        j.state = MRT_JOB_DONE;
        s->completed++;
        s->segments += j.segments;
        s->latencies.push_back(j.t_done - j.t_submit);
        s->last_done = j.t_done;
    }
    s->adm.release(j.layout.block);
    s->finished.push_back(j.id);
    while (s->finished.size() > KEEP_FINISHED) {
        s->jobs.erase(s->finished.front());
        s->finished.pop_front();
    }
    sl.job = nullptr;
    s->cv_done.notify_all();
}

void scheduler(mrt_server* s) {
    // a thread that cannot select the device fails every job it admits with that error
    const hipError_t dev = hipSetDevice(s->device);
    // MRT_SERVE_TRACE=1 prints, per job, the host time spent starting it (allocation, capture,
    // launch) and finishing it, and the pool's reserved bytes, to stderr
    const char* tr = std::getenv("MRT_SERVE_TRACE");
    const bool trace = tr && tr[0] == '1';
    std::unique_lock<std::mutex> lk(s->m);
    while (true) {
        // admit while the head of the queue fits: beside the jobs in flight (budget), in a free
        // slot whose block fits or can grow within the budget (plan_slot)
        while (true) {
            uint64_t id = 0, bytes = 0;
            if (!s->adm.peek(&id, &bytes)) break;
            std::vector<uint64_t> held(s->slots.size());
            std::vector<bool> busy(s->slots.size());
            for (size_t k = 0; k < s->slots.size(); k++) {
                held[k] = s->slots[k].block_bytes;
                busy[k] = s->slots[k].job != nullptr;
            }
            SlotPlan plan;
            if (!plan_slot(held, busy, s->adm.budget(), bytes, &plan)) break;
            (void)s->adm.admit(&id, &bytes);
            Slot& sl = s->slots[size_t(plan.slot)];
            Job& j = *s->jobs.at(id);
            j.state = MRT_JOB_RUNNING;
            j.t_admit = now_ms(s->t0);
            sl.job = &j;
            s->peak_in_flight = std::max(s->peak_in_flight, s->adm.busy());
            lk.unlock();
            const auto t_start = std::chrono::steady_clock::now();
            hipError_t e = dev;
            for (int r : plan.release)
                if (e == hipSuccess) e = release_block(s->slots[size_t(r)]);
            if (e == hipSuccess) e = resize_block(s->pool, sl, plan.block);
            sl.t_alloc_ms = now_ms(t_start);
            uint64_t held_now = 0;
            for (const Slot& t : s->slots) held_now += t.block_bytes;
            if (e == hipSuccess) e = start_job(s, sl, j);
            if (trace) {
                uint64_t reserved = 0;
                (void)hipMemPoolGetAttribute(s->pool, hipMemPoolAttrReservedMemCurrent, &reserved);
                std::fprintf(stderr, "serve: job %llu %ux%u spp %d start %.3f ms (alloc %.3f ms), pool reserved %.2f GB\n",
                             (unsigned long long)j.id, j.d.width, j.d.height, j.d.params.spp, now_ms(t_start),
                             sl.t_alloc_ms, double(reserved) / 1e9);
            }
            if (e != hipSuccess) {
                (void)hipStreamSynchronize(sl.stream); // drain what was queued; the job fails with e
                finish_job(s, sl, e);
            }
            lk.lock();
            s->held = held_now;
            s->peak_bytes = std::max(s->peak_bytes, held_now);
        }
        const bool in_flight = s->adm.busy() > 0;
        if (!in_flight) {
            if (s->stop) break;
            s->cv_work.wait(lk, [&] { return s->stop || s->adm.queued() > 0; });
            continue;
        }
        lk.unlock();
        bool progressed = false;
        for (Slot& sl : s->slots) {
            if (!sl.job) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) continue;
            const auto t_fin = std::chrono::steady_clock::now();
            const uint64_t fin_id = sl.job->id;
            finish_job(s, sl, q);
            if (trace) std::fprintf(stderr, "serve: job %llu finish %.3f ms\n", (unsigned long long)fin_id, now_ms(t_fin));
            progressed = true;
        }
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(50));
        lk.lock();
    }
}

void fill_status(const Job& j, mrt_job_status* st) {
    st->state = j.state;
    st->error = j.error;
    st->segments = j.segments;
    st->device_bytes = j.layout.block;
    st->queued_ms = j.state == MRT_JOB_QUEUED ? 0.0 : j.t_admit - j.t_submit;
    st->latency_ms = (j.state == MRT_JOB_DONE || j.state == MRT_JOB_FAILED || j.state == MRT_JOB_CANCELLED)
                         ? j.t_done - j.t_submit
                         : 0.0;
}

} // namespace

extern "C" {

MRT_API mrt_status mrt_server_create(mrt_ctx* c, const mrt_server_config* cfg, mrt_server** out) {
    if (!c || !out) return MRT_ERR_INVALID;
    *out = nullptr;
    mrt_server_config k = cfg ? *cfg : mrt_server_config{};
    if (k.max_slots == 0) k.max_slots = 4;
    if (k.max_slots > MAX_SLOTS) return fail(c, MRT_ERR_INVALID, "max_slots above 16");
    if (k.max_paths_per_job == 0) k.max_paths_per_job = 1u << 24;
    const uint64_t left = c->budget > c->in_use ? c->budget - c->in_use : 0;
    if (k.hbm_budget == 0) k.hbm_budget = left / 2;
    if (k.hbm_budget > left) return fail(c, MRT_ERR_BUDGET, "server budget exceeds what the context has left");
    if (const hipError_t e = hipSetDevice(c->device); e != hipSuccess) return hip_fail(c, e, "mrt_server_create");
    std::unique_ptr<mrt_server> s(new mrt_server());
    s->ctx = c;
    s->device = c->device;
    s->cfg = k;
    s->trace_threads = trace_max_threads();
    trace_prepare();
    s->t0 = std::chrono::steady_clock::now();
    s->adm = Admission(k.hbm_budget, k.max_slots);
    // freed job buffers stay in the server's pool, up to its budget, instead of going back to
    // the driver on every synchronisation
    hipMemPoolProps props{};
    props.allocType = hipMemAllocationTypePinned;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = c->device;
    if (const hipError_t e = hipMemPoolCreate(&s->pool, &props); e != hipSuccess) {
        s->pool = nullptr;
        return hip_fail(c, e, "server memory pool");
    }
    uint64_t keep = k.hbm_budget;
    if (const hipError_t e = hipMemPoolSetAttribute(s->pool, hipMemPoolAttrReleaseThreshold, &keep); e != hipSuccess) {
        (void)hipMemPoolDestroy(s->pool);
        return hip_fail(c, e, "server memory pool");
    }
    s->slots.resize(k.max_slots);
    for (Slot& sl : s->slots) {
        hipError_t e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e != hipSuccess) {
            for (Slot& t : s->slots) { // best effort: the creation error is the one reported
                if (t.done) (void)hipEventDestroy(t.done);
                if (t.stream) (void)hipStreamDestroy(t.stream);
            }
            (void)hipMemPoolDestroy(s->pool);
            return hip_fail(c, e, "server streams");
        }
    }
    c->in_use += k.hbm_budget;
    s->reserved = k.hbm_budget;
    mrt_server* raw = s.release();
    raw->th = std::thread(scheduler, raw);
    *out = raw;
    return MRT_OK;
}

MRT_API mrt_status mrt_server_destroy(mrt_server* s) {
    if (!s) return MRT_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->m);
        // drop queued jobs; jobs in flight finish
        for (auto& kv : s->jobs) {
            Job& j = *kv.second;
            if (j.state == MRT_JOB_QUEUED && s->adm.remove(j.id)) {
                j.state = MRT_JOB_CANCELLED;
                j.error = MRT_ERR_CANCELLED;
                s->cancelled++;
            }
        }
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    if (s->th.joinable()) s->th.join();
    // teardown is best effort: every slot is released even if one step fails
    bool ok = hipSetDevice(s->device) == hipSuccess;
    for (Slot& sl : s->slots) {
        ok &= hipStreamSynchronize(sl.stream) == hipSuccess;
        if (sl.exec) ok &= hipGraphExecDestroy(sl.exec) == hipSuccess;
        if (sl.h_counters) ok &= hipHostFree(sl.h_counters) == hipSuccess;
        if (sl.h_result) ok &= hipHostFree(sl.h_result) == hipSuccess;
        ok &= release_block(sl) == hipSuccess;
        ok &= hipStreamSynchronize(sl.stream) == hipSuccess;
        ok &= hipEventDestroy(sl.done) == hipSuccess;
        ok &= hipStreamDestroy(sl.stream) == hipSuccess;
    }
    if (s->pool) {
        // every job buffer was freed on its slot stream, and every slot stream has drained
        ok &= hipMemPoolTrimTo(s->pool, 0) == hipSuccess;
        ok &= hipMemPoolDestroy(s->pool) == hipSuccess;
    }
    s->ctx->in_use -= s->reserved;
    delete s;
    return ok ? MRT_OK : MRT_ERR_DEVICE;
}

MRT_API const char* mrt_server_last_error(mrt_server* s) {
    if (!s) return "null server";
    std::lock_guard<std::mutex> lk(s->m);
    thread_local std::string copy;
    copy = s->err;
    return copy.c_str();
}

MRT_API mrt_status mrt_job_bytes(mrt_server* s, const mrt_job_desc* d, uint64_t* bytes) {
    if (!s || !bytes) return MRT_ERR_INVALID;
    std::string why;
    const mrt_status v = validate(d, &why);
    if (v != MRT_OK) {
        std::lock_guard<std::mutex> lk(s->m);
        s->err = why;
        return v;
    }
    *bytes = job_layout(shape_of(s, *d)).block;
    return MRT_OK;
}

MRT_API mrt_status mrt_job_submit(mrt_server* s, const mrt_job_desc* d, uint64_t* id) {
    if (!s || !id) return MRT_ERR_INVALID;
    std::string why;
    const mrt_status v = validate(d, &why);
    std::lock_guard<std::mutex> lk(s->m);
    if (v != MRT_OK) {
        s->err = why;
        return v;
    }
    if (s->stop) {
        s->err = "server is shutting down";
        return MRT_ERR_INVALID;
    }
    std::unique_ptr<Job> j(new Job());
    j->d = *d;
    j->layout = job_layout(shape_of(s, *d));
    if (!s->adm.can_ever_fit(j->layout.block)) {
        s->err = "job needs " + std::to_string(j->layout.block) + " bytes, more than the server budget of " +
                 std::to_string(s->adm.budget());
        return MRT_ERR_BUDGET;
    }
    j->id = s->next_id++;
    j->t_submit = now_ms(s->t0);
    if (s->first_submit < 0) s->first_submit = j->t_submit;
    s->submitted++;
    *id = j->id;
    s->adm.push(j->id, j->layout.block);
    s->jobs.emplace(j->id, std::move(j));
    s->cv_work.notify_one();
    return MRT_OK;
}

MRT_API mrt_status mrt_job_poll(mrt_server* s, uint64_t id, mrt_job_status* st) {
    if (!s || !st) return MRT_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return MRT_ERR_INVALID;
    fill_status(*it->second, st);
    return MRT_OK;
}

MRT_API mrt_status mrt_job_wait(mrt_server* s, uint64_t id, int32_t timeout_ms, mrt_job_status* st) {
    if (!s || !st) return MRT_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return MRT_ERR_INVALID;
    Job* j = it->second.get();
    auto terminal = [&] {
        return j->state == MRT_JOB_DONE || j->state == MRT_JOB_CANCELLED || j->state == MRT_JOB_FAILED;
    };
    if (timeout_ms < 0) {
        s->cv_done.wait(lk, terminal);
    } else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), terminal)) {
        fill_status(*j, st);
        return MRT_ERR_NOT_READY;
    }
    fill_status(*j, st);
    return MRT_OK;
}

MRT_API mrt_status mrt_job_cancel(mrt_server* s, uint64_t id) {
    if (!s) return MRT_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return MRT_ERR_INVALID;
    Job& j = *it->second;
    if (j.state == MRT_JOB_QUEUED) {
        s->adm.remove(id);
        j.state = MRT_JOB_CANCELLED;
        j.error = MRT_ERR_CANCELLED;
        j.t_done = now_ms(s->t0);
        s->cancelled++;
        s->finished.push_back(id);
        s->cv_done.notify_all();
        return MRT_OK;
    }
    if (j.state == MRT_JOB_RUNNING && !j.cancel) {
        j.cancel = true;
        return MRT_OK;
    }
    return MRT_ERR_INVALID;
}

MRT_API mrt_status mrt_server_stats_get(mrt_server* s, mrt_server_stats* o) {
    if (!s || !o) return MRT_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->m);
    std::memset(o, 0, sizeof(*o));
    o->submitted = s->submitted;
    o->completed = s->completed;
    o->cancelled = s->cancelled;
    o->failed = s->failed;
    o->segments = s->segments;
    o->p50_ms = percentile(s->latencies, 50.0);
    o->p99_ms = percentile(s->latencies, 99.0);
    o->max_ms = percentile(s->latencies, 100.0);
    double sum = 0;
    for (double v : s->latencies) sum += v;
    o->mean_ms = s->latencies.empty() ? 0.0 : sum / double(s->latencies.size());
    o->wall_ms = s->first_submit >= 0 && s->last_done > s->first_submit ? s->last_done - s->first_submit : 0.0;
    o->peak_bytes = s->peak_bytes;
    o->peak_in_flight = s->peak_in_flight;
    o->graph_instantiations = s->graph_inst;
    o->graph_updates = s->graph_upd;
    return MRT_OK;
}

MRT_API mrt_status mrt_server_stats_reset(mrt_server* s) {
    if (!s) return MRT_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->m);
    s->submitted = s->completed = s->cancelled = s->failed = s->segments = 0;
    s->latencies.clear();
    s->first_submit = -1;
    s->last_done = 0;
    s->peak_bytes = s->held;
    s->peak_in_flight = s->adm.busy();
    s->graph_inst = s->graph_upd = 0;
    return MRT_OK;
}

} // extern "C"
