// C ABI of the learned guide (include/merian_amd.h, mrt_guide_*): parameter storage, Adam
// state and training buffers on the device, host entry points for tests and tools, and the
// per-frame hooks of the guided pt (render_frame_dev in api.cpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/merian_amd.h"
#include "mrt_api_internal.h"
#include "mrt_guide.h"

using namespace mrt;

struct mrt_guide {
    mrt_guide_config cfg{};
    GuideNet net{};
    GuideParams P{};
    GuideTrainBuffers T{};    // kblocks = capacity
    float* dW = nullptr;      // reduced weight gradient [5][64][64]
    float4* host_records = nullptr; // records uploaded by mrt_guide_gradient / mrt_guide_train
    uint32_t host_records_cap = 0;
    // frame buffers
    uint4* lobes = nullptr;
    uint64_t lobes_cap = 0;   // queue entries
    float4* grec = nullptr;
    float4* gtrain = nullptr;
    float4* gtrain_ordered = nullptr; // the frame's records in their fixed training order
    uint32_t* order_keys = nullptr;   // 2 x max_records (sort input, output)
    uint32_t* order_vals = nullptr;
    void* order_temp = nullptr;
    size_t order_temp_bytes = 0;
    uint32_t* grec_count = nullptr;
    uint32_t* h_count = nullptr; // pinned
    uint64_t steps = 0;       // Adam steps taken
    std::vector<void*> allocs;
};

namespace {

mrt_guide_config with_defaults(const mrt_guide_config* in) {
    mrt_guide_config c = in ? *in : mrt_guide_config{};
    if (c.log2_table == 0) c.log2_table = 14;
    if (c.base_resolution == 0) c.base_resolution = 16;
    if (c.max_resolution == 0) c.max_resolution = 2048;
    if (c.train_batch == 0) c.train_batch = 1u << 16;
    if (c.train_steps == 0) c.train_steps = 4;
    if (c.max_records == 0) c.max_records = 1u << 20;
    if (c.learning_rate == 0.f) c.learning_rate = 5e-3f;
    if (c.guide_fraction == 0.f) c.guide_fraction = 0.5f;
    if (c.guided_bounces == 0) c.guided_bounces = 64;
    return c;
}

mrt_status check_config(const mrt_guide_config& c) {
    if (c.log2_table < 8 || c.log2_table > 24) return MRT_ERR_INVALID;
    if (c.base_resolution < 1 || c.max_resolution < c.base_resolution || c.max_resolution > (1u << 20))
        return MRT_ERR_INVALID;
    // a hashed level stores each grid row along x whole: the finest row must fit one table
    if (uint64_t(c.max_resolution) + 1 > (uint64_t(1) << c.log2_table)) return MRT_ERR_INVALID;
    if (c.train_batch > (1u << 22) || c.max_records > (1u << 26)) return MRT_ERR_INVALID;
    if (!(c.guide_fraction > 0.f && c.guide_fraction < 1.f)) return MRT_ERR_INVALID;
    return MRT_OK;
}

void layout(const mrt_guide_config& c, uint32_t res[MRT_GUIDE_LEVELS]) {
    const double b = std::log(double(c.max_resolution) / double(c.base_resolution)) / double(MRT_GUIDE_LEVELS - 1);
    for (int l = 0; l < MRT_GUIDE_LEVELS; l++)
        res[l] = uint32_t(std::floor(double(c.base_resolution) * std::exp(b * double(l)) + 0.5));
}

// Deterministic host generator for the initial weights (xorshift64*).
struct InitRng {
    uint64_t s;
    explicit InitRng(uint32_t seed) : s(0x9E3779B97F4A7C15ull ^ (uint64_t(seed) * 0xD1B54A32D192ED03ull)) {
        if (!s) s = 1;
    }
    float uniform() { // [-1, 1)
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        const uint64_t r = s * 0x2545F4914F6CDD1Dull;
        return float(double(r >> 11) * (1.0 / 9007199254740992.0)) * 2.f - 1.f;
    }
};

template <class T> T* galloc(mrt_ctx* c, mrt_guide* g, size_t count, mrt_status* st) {
    if (*st != MRT_OK) return nullptr;
    void* p = dev_alloc(c, count * sizeof(T), st);
    if (*st == MRT_OK) g->allocs.push_back(p);
    return static_cast<T*>(p);
}

// Releases every buffer of the guide; false if freeing the pinned counter failed.
bool guide_free(mrt_ctx* c, mrt_guide* g) {
    if (!g) return true;
    for (void* p : g->allocs) dev_free(c, p);
    const bool ok = !g->h_count || hipHostFree(g->h_count) == hipSuccess;
    delete g;
    return ok;
}

// Training buffers for batches of up to `cap` records.
mrt_status ensure_train(mrt_ctx* c, mrt_guide* g, uint32_t cap) {
    if (g->T.H && g->T.cap >= cap) return MRT_OK;
    for (void* p : {(void*)g->T.H, (void*)g->T.D, (void*)g->T.partials}) {
        if (!p) continue;
        dev_free(c, p);
        g->allocs.erase(std::find(g->allocs.begin(), g->allocs.end(), p));
    }
    mrt_status st = MRT_OK;
    const uint32_t kb = (cap + G_WGRAD_RECORDS - 1) / G_WGRAD_RECORDS;
    g->T.H = galloc<_Float16>(c, g, size_t(G_LAYERS) * cap * G_WIDTH, &st);
    g->T.D = galloc<_Float16>(c, g, size_t(G_LAYERS) * cap * G_WIDTH, &st);
    g->T.partials = galloc<float>(c, g, size_t(G_LAYERS) * kb * G_WIDTH * G_WIDTH, &st);
    if (st != MRT_OK) {
        g->T.H = g->T.D = nullptr;
        g->T.partials = nullptr;
        g->T.cap = 0;
        return st;
    }
    g->T.cap = cap;
    g->T.kblocks = kb;
    return MRT_OK;
}

mrt_status upload_records(mrt_ctx* c, mrt_guide* g, const float* rec, uint32_t n) {
    if (n > g->host_records_cap) {
        if (g->host_records) {
            dev_free(c, g->host_records);
            g->allocs.erase(std::find(g->allocs.begin(), g->allocs.end(), (void*)g->host_records));
        }
        mrt_status st = MRT_OK;
        g->host_records = galloc<float4>(c, g, size_t(3) * n, &st);
        if (st != MRT_OK) {
            g->host_records = nullptr;
            g->host_records_cap = 0;
            return st;
        }
        g->host_records_cap = n;
    }
    const hipError_t e = hipMemcpyAsync(g->host_records, rec, size_t(n) * 48, hipMemcpyHostToDevice, c->stream);
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, "guide records upload");
}

// One Adam step over n device records.
mrt_status train_step(mrt_ctx* c, mrt_guide* g, const float4* rec, uint32_t n, hipStream_t st) {
    GuideTrainBuffers tb = g->T;
    hipError_t e = launch_guide_gradient(g->net, rec, n, tb, st);
    if (e != hipSuccess) return hip_fail(c, e, "guide gradient");
    g->steps++;
    const float bc1 = float(1.0 - std::pow(0.9, double(g->steps)));
    const float bc2 = float(1.0 - std::pow(0.99, double(g->steps)));
    e = launch_guide_adam(g->P, tb, g->cfg.learning_rate, n ? 1.f / float(n) : 0.f, bc1, bc2, st);
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, "guide adam");
}

// Guided frames may run (and train) on a caller's stream: the calls that read, write or free
// the guide's buffers on the context stream drain the device first.
mrt_status drain(mrt_ctx* c, const char* what) {
    hipError_t e = hipSetDevice(c->device);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, what);
}

} // namespace

namespace mrt {

mrt_status guide_prepare_frame(mrt_ctx* c, mrt_guide* g, uint64_t paths, PathBuffers& pb, float& fraction,
                               uint32_t& guided_bounces, const GuideNet** net) {
    if (paths > g->lobes_cap) {
        if (g->lobes) {
            dev_free(c, g->lobes);
            g->allocs.erase(std::find(g->allocs.begin(), g->allocs.end(), (void*)g->lobes));
        }
        mrt_status st = MRT_OK;
        g->lobes = galloc<uint4>(c, g, size_t(4) * paths, &st);
        if (st != MRT_OK) {
            g->lobes = nullptr;
            g->lobes_cap = 0;
            return st;
        }
        g->lobes_cap = paths;
    }
    pb.lobes = g->lobes;
    pb.grec = g->grec;
    pb.gtrain = g->gtrain;
    pb.grec_count = g->grec_count;
    pb.grec_cap = g->cfg.max_records;
    fraction = g->cfg.guide_fraction;
    guided_bounces = g->cfg.guided_bounces;
    *net = &g->net;
    return MRT_OK;
}

mrt_status guide_train_frame(mrt_ctx* c, mrt_guide* g, hipStream_t st) {
    if (g->cfg.train_steps == 0) return MRT_OK;
    hipError_t e = hipMemcpyAsync(g->h_count, g->grec_count, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(c, e, "guide record count");
    const uint32_t n = std::min(*g->h_count, g->cfg.max_records);
    if (n == 0) return MRT_OK;
    // records arrive in the order blocks finish; sorted by (path, bounce) and permuted, the
    // training order, hence the trained guide, is the same on every run
    e = launch_guide_order_records(g->gtrain, n, g->order_keys, g->order_vals, g->order_temp, g->order_temp_bytes,
                                   g->gtrain_ordered, st);
    if (e != hipSuccess) return hip_fail(c, e, "guide record order");
    // the steps take disjoint slices of the (permuted) records while the frame has enough of
    // them: no record drives more than one step of a frame
    const uint32_t steps = g->cfg.train_steps;
    const uint32_t batch = std::min(g->cfg.train_batch, std::max(1u, (n + steps - 1) / steps));
    mrt_status s = ensure_train(c, g, batch);
    if (s != MRT_OK) return s;
    for (uint32_t k = 0; k < steps; k++) {
        const uint32_t off = uint32_t((uint64_t(k) * batch) % n);
        const uint32_t cnt = std::min(batch, n - off);
        s = train_step(c, g, g->gtrain_ordered + 3 * size_t(off), cnt, st);
        if (s != MRT_OK) return s;
    }
    return MRT_OK;
}

} // namespace mrt

extern "C" {

MRT_API mrt_status mrt_guide_layout(const mrt_guide_config* config, uint32_t resolutions[MRT_GUIDE_LEVELS], uint32_t* table) {
    const mrt_guide_config c = with_defaults(config);
    if (check_config(c) != MRT_OK) return MRT_ERR_INVALID;
    if (resolutions) layout(c, resolutions);
    if (table) *table = 1u << c.log2_table;
    return MRT_OK;
}

MRT_API mrt_status mrt_guide_create(mrt_ctx* c, mrt_scene* s, const mrt_guide_config* config, mrt_guide** out) {
    if (!c || !s || !out) return fail(c, MRT_ERR_INVALID, "mrt_guide_create: null argument");
    *out = nullptr;
    const mrt_guide_config cfg = with_defaults(config);
    if (check_config(cfg) != MRT_OK) return fail(c, MRT_ERR_INVALID, "mrt_guide_create: configuration out of range");
    if (const hipError_t e = hipSetDevice(c->device); e != hipSuccess) return hip_fail(c, e, "mrt_guide_create");
    mrt_guide* g = new mrt_guide();
    g->cfg = cfg;
    layout(cfg, g->net.res);
    g->net.table = 1u << cfg.log2_table;
    float lo[3], hi[3];
    bool box = false;
    for (int k = 0; k < 3; k++) box = box || cfg.bbox_min[k] != 0.f || cfg.bbox_max[k] != 0.f;
    for (int k = 0; k < 3; k++) {
        lo[k] = box ? cfg.bbox_min[k] : s->bbox_min[k];
        hi[k] = box ? cfg.bbox_max[k] : s->bbox_max[k];
        const float ext = std::max(hi[k] - lo[k], 1e-6f);
        g->net.offset[k] = lo[k] - 1e-3f * ext;
        g->net.scale[k] = 1.f / (ext * 1.002f);
    }
    mrt_status st = MRT_OK;
    const size_t nw = size_t(G_LAYERS) * G_WIDTH * G_WIDTH;
    const uint64_t ng = uint64_t(G_LEVELS) * g->net.table * G_FEATS;
    g->P.w = galloc<float>(c, g, nw, &st);
    g->P.w_m = galloc<float>(c, g, nw, &st);
    g->P.w_v = galloc<float>(c, g, nw, &st);
    g->P.frag_fwd = galloc<half8_t>(c, g, size_t(G_FRAGS) * 64, &st);
    g->P.frag_bwd = galloc<half8_t>(c, g, size_t(G_FRAGS) * 64, &st);
    g->P.g = galloc<float>(c, g, ng, &st);
    g->P.g_m = galloc<float>(c, g, ng, &st);
    g->P.g_v = galloc<float>(c, g, ng, &st);
    g->P.grid = galloc<_Float16>(c, g, ng, &st);
    g->P.grid_count = ng;
    g->T.grid_grad = galloc<unsigned long long>(c, g, ng, &st);
    g->T.scalars = galloc<float>(c, g, 4, &st);
    g->dW = galloc<float>(c, g, nw, &st);
    g->grec = galloc<float4>(c, g, size_t(5) * cfg.max_records, &st);
    g->gtrain = galloc<float4>(c, g, size_t(3) * cfg.max_records, &st);
    g->gtrain_ordered = galloc<float4>(c, g, size_t(3) * cfg.max_records, &st);
    g->order_keys = galloc<uint32_t>(c, g, size_t(2) * cfg.max_records, &st);
    g->order_vals = galloc<uint32_t>(c, g, size_t(2) * cfg.max_records, &st);
    g->order_temp_bytes = launch_guide_order_temp_bytes(cfg.max_records);
    if (st == MRT_OK && g->order_temp_bytes == 0) st = fail(c, MRT_ERR_DEVICE, "guide record sort: no temporary size");
    g->order_temp = galloc<char>(c, g, std::max<size_t>(g->order_temp_bytes, 1), &st);
    g->grec_count = galloc<uint32_t>(c, g, 1, &st);
    if (st == MRT_OK && hipHostMalloc(reinterpret_cast<void**>(&g->h_count), 4, hipHostMallocDefault) != hipSuccess) {
        g->h_count = nullptr;
        st = fail(c, MRT_ERR_ALLOC, "pinned host memory");
    }
    if (st != MRT_OK) {
        guide_free(c, g);
        return st;
    }
    g->net.frag_fwd = g->P.frag_fwd;
    g->net.frag_bwd = g->P.frag_bwd;
    g->net.grid = g->P.grid;
    // initial parameters: uniform weights scaled by sqrt(6 / (fan_in + fan_out)), output layer
    // halved (a broad, near-uniform mixture to start from), grid entries in [-1e-4, 1e-4]
    InitRng rng(cfg.seed);
    std::vector<float> w(nw, 0.f);
    for (int l = 0; l < G_LAYERS; l++) {
        const int n_out = l + 1 < G_LAYERS ? G_WIDTH : G_OUT;
        const float a = std::sqrt(6.f / float(G_WIDTH + n_out)) * (l + 1 < G_LAYERS ? 1.f : 0.5f);
        for (int o = 0; o < n_out; o++)
            for (int i = 0; i < G_WIDTH; i++) w[(size_t(l) * G_WIDTH + o) * G_WIDTH + i] = a * rng.uniform();
    }
    std::vector<float> gr(ng);
    for (uint64_t k = 0; k < ng; k++) gr[k] = 1e-4f * rng.uniform();
    const mrt_status ps = mrt_guide_set_params(c, g, w.data(), gr.data());
    if (ps != MRT_OK) {
        guide_free(c, g);
        return ps;
    }
    *out = g;
    return MRT_OK;
}

MRT_API mrt_status mrt_guide_destroy(mrt_ctx* c, mrt_guide* g) {
    if (!c || !g) return MRT_ERR_INVALID;
    const mrt_status st = drain(c, "mrt_guide_destroy"); // kernels may still use the guide's buffers
    const bool freed = guide_free(c, g);
    if (st != MRT_OK) return st;
    return freed ? MRT_OK : fail(c, MRT_ERR_DEVICE, "mrt_guide_destroy: hipHostFree failed");
}

MRT_API mrt_status mrt_guide_normalisation(mrt_guide* g, float offset[3], float scale[3]) {
    if (!g || !offset || !scale) return MRT_ERR_INVALID;
    for (int k = 0; k < 3; k++) {
        offset[k] = g->net.offset[k];
        scale[k] = g->net.scale[k];
    }
    return MRT_OK;
}

MRT_API mrt_status mrt_guide_get_params(mrt_ctx* c, mrt_guide* g, float* weights, float* grid) {
    if (!c || !g) return MRT_ERR_INVALID;
    if (const mrt_status st = drain(c, "mrt_guide_get_params"); st != MRT_OK) return st;
    hipError_t e = hipSuccess;
    if (weights)
        e = hipMemcpyAsync(weights, g->P.w, size_t(G_LAYERS) * G_WIDTH * G_WIDTH * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && grid) e = hipMemcpyAsync(grid, g->P.g, g->P.grid_count * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, "mrt_guide_get_params");
}

MRT_API mrt_status mrt_guide_set_params(mrt_ctx* c, mrt_guide* g, const float* weights, const float* grid) {
    if (!c || !g) return MRT_ERR_INVALID;
    if (const mrt_status st = drain(c, "mrt_guide_set_params"); st != MRT_OK) return st;
    const size_t nw = size_t(G_LAYERS) * G_WIDTH * G_WIDTH;
    std::vector<float> w;
    if (weights) {
        // rows >= 40 of the output layer stay zero
        w.assign(weights, weights + nw);
        for (size_t o = G_OUT; o < size_t(G_WIDTH); o++)
            for (size_t i = 0; i < size_t(G_WIDTH); i++) w[(size_t(G_LAYERS - 1) * G_WIDTH + o) * G_WIDTH + i] = 0.f;
    }
    hipError_t e = hipSuccess;
    if (weights) e = hipMemcpyAsync(g->P.w, w.data(), nw * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && grid) e = hipMemcpyAsync(g->P.g, grid, g->P.grid_count * 4, hipMemcpyHostToDevice, c->stream);
    // Adam state and gradients restart from zero
    if (e == hipSuccess) e = hipMemsetAsync(g->P.w_m, 0, nw * 4, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(g->P.w_v, 0, nw * 4, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(g->P.g_m, 0, g->P.grid_count * 4, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(g->P.g_v, 0, g->P.grid_count * 4, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(g->T.grid_grad, 0, g->P.grid_count * sizeof(unsigned long long), c->stream);
    if (e == hipSuccess) e = launch_guide_pack(g->P, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    g->steps = 0;
    return e == hipSuccess ? MRT_OK : hip_fail(c, e, "mrt_guide_set_params");
}

MRT_API mrt_status mrt_guide_forward(mrt_ctx* c, mrt_guide* g, const float* positions, const float* directions,
                                     uint32_t n, float* outputs) {
    if (!c || !g || (n && (!positions || !directions || !outputs)))
        return fail(c, MRT_ERR_INVALID, "mrt_guide_forward: null argument");
    if (n == 0) return MRT_OK;
    if (const mrt_status st = drain(c, "mrt_guide_forward"); st != MRT_OK) return st;
    std::vector<float4> hp(n), hd(n);
    for (uint32_t i = 0; i < n; i++) {
        hp[i] = make_float4(positions[3 * i], positions[3 * i + 1], positions[3 * i + 2], 0.f);
        hd[i] = make_float4(directions[3 * i], directions[3 * i + 1], directions[3 * i + 2], 0.f);
    }
    mrt_status st = MRT_OK;
    void* dp = dev_alloc(c, size_t(n) * 16, &st);
    void* dd = st == MRT_OK ? dev_alloc(c, size_t(n) * 16, &st) : nullptr;
    void* dout = st == MRT_OK ? dev_alloc(c, size_t(n) * G_OUT * 4, &st) : nullptr;
    hipError_t e = hipSuccess;
    if (st == MRT_OK) {
        e = hipMemcpyAsync(dp, hp.data(), size_t(n) * 16, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dd, hd.data(), size_t(n) * 16, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = launch_guide_forward(g->net, static_cast<float4*>(dp), static_cast<float4*>(dd), n,
                                     static_cast<float*>(dout), c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(outputs, dout, size_t(n) * G_OUT * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) st = hip_fail(c, e, "mrt_guide_forward");
    }
    dev_free(c, dp);
    dev_free(c, dd);
    dev_free(c, dout);
    return st;
}

MRT_API mrt_status mrt_guide_gradient(mrt_ctx* c, mrt_guide* g, const float* records, uint32_t n, float* d_weights,
                                      float* d_grid, float* loss) {
    if (!c || !g || (n && !records)) return fail(c, MRT_ERR_INVALID, "mrt_guide_gradient: null argument");
    if (n == 0) return fail(c, MRT_ERR_INVALID, "mrt_guide_gradient: no records");
    mrt_status st = drain(c, "mrt_guide_gradient");
    if (st == MRT_OK) st = ensure_train(c, g, n);
    if (st == MRT_OK) st = upload_records(c, g, records, n);
    if (st != MRT_OK) return st;
    GuideTrainBuffers tb = g->T;
    const size_t gbytes = g->P.grid_count * sizeof(unsigned long long);
    hipError_t e = hipMemsetAsync(g->T.grid_grad, 0, gbytes, c->stream);
    if (e == hipSuccess) e = launch_guide_gradient(g->net, g->host_records, n, tb, c->stream);
    if (e == hipSuccess) e = launch_guide_reduce_wgrad(tb, g->dW, c->stream);
    float sc[2] = {0.f, 0.f};
    std::vector<unsigned long long> fx(d_grid ? g->P.grid_count : 0);
    if (e == hipSuccess && d_weights)
        e = hipMemcpyAsync(d_weights, g->dW, size_t(G_LAYERS) * G_WIDTH * G_WIDTH * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && d_grid) e = hipMemcpyAsync(fx.data(), g->T.grid_grad, gbytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(sc, g->T.scalars, 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    // leave the gradient accumulator cleared for training
    if (e == hipSuccess) e = hipMemsetAsync(g->T.grid_grad, 0, gbytes, c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "mrt_guide_gradient");
    for (size_t k = 0; k < fx.size(); k++) d_grid[k] = float(double((long long)fx[k]) / double(G_GRID_FX));
    if (loss) *loss = sc[1] / float(n);
    return MRT_OK;
}

MRT_API mrt_status mrt_guide_train(mrt_ctx* c, mrt_guide* g, const float* records, uint32_t n, uint32_t steps,
                                   float* loss) {
    if (!c || !g || (n && !records)) return fail(c, MRT_ERR_INVALID, "mrt_guide_train: null argument");
    if (n == 0 || steps == 0) return MRT_OK;
    const uint32_t batch = std::min(n, g->cfg.train_batch);
    mrt_status st = drain(c, "mrt_guide_train");
    if (st == MRT_OK) st = ensure_train(c, g, batch);
    if (st == MRT_OK) st = upload_records(c, g, records, n);
    if (st != MRT_OK) return st;
    uint32_t last = batch;
    for (uint32_t k = 0; k < steps && st == MRT_OK; k++) {
        const uint32_t off = uint32_t((uint64_t(k) * batch) % n);
        last = std::min(batch, n - off);
        st = train_step(c, g, g->host_records + 3 * size_t(off), last, c->stream);
    }
    if (st != MRT_OK) return st;
    float sc[2] = {0.f, 0.f};
    hipError_t e = hipMemcpyAsync(sc, g->T.scalars, 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "mrt_guide_train");
    if (loss) *loss = sc[1] / float(last);
    return MRT_OK;
}

MRT_API mrt_status mrt_guided_render_dev(mrt_ctx* c, mrt_scene* s, mrt_guide* g, const mrt_pt_params* prm,
                                         const mrt_camera* cam, uint32_t frame, uint32_t w, uint32_t h,
                                         const mrt_packed_hit* gbuffer, float* irradiance, int32_t train,
                                         mrt_pt_stats* stats, void* stream) {
    if (!c || !s || !g || !prm) return fail(c, MRT_ERR_INVALID, "mrt_guided_render_dev: null argument");
    if (prm->mode != MRT_MODE_GUIDED) return fail(c, MRT_ERR_INVALID, "mrt_guided_render_dev: mode must be GUIDED");
    if (prm->spp < 0 || prm->max_path_length < 1 || prm->max_path_length > 64)
        return fail(c, MRT_ERR_INVALID, "samples per pixel < 0 or max path length outside [1, 64]");
    // training records are keyed pixel * 64 + vertex in 32 bits
    if (uint64_t(w) * h > (uint64_t(1) << 26))
        return fail(c, MRT_ERR_UNSUPPORTED, "mrt_guided_render_dev: more than 2^26 pixels");
    // without training the records are still written but no step runs
    const uint32_t steps = g->cfg.train_steps;
    if (!train) g->cfg.train_steps = 0;
    const mrt_status st =
        render_frame_dev(c, s, prm, cam, frame, w, h, gbuffer, irradiance, nullptr, stats, stream, g);
    g->cfg.train_steps = steps;
    return st;
}

} // extern "C"
