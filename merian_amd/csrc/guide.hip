// gfx950 kernels of the learned guide (include/merian_amd.h, MRT_MODE_GUIDED): hash-grid + SH
// encoding, the 64-wide MLP on MFMA (fp16 operands, fp32 accumulation), lobe packing for the
// shading kernel, and training (fused forward/backward per record tile, a weight-gradient GEMM
// over the batch, Adam). Layout conventions: mrt_guide.h.
#include "mrt_guide.h"

#include <rocprim/device/device_radix_sort.hpp>

namespace mrt {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int GB = 256;        // threads per block (4 waves)
constexpr int NT = 2;          // 16-query tiles per wave iteration
constexpr int WQ = 16 * NT;    // queries per wave iteration
constexpr int BQ = 4 * WQ;     // queries per block iteration
constexpr int INT_ = 1;        // inference: 16-query tiles per wave iteration
constexpr int IWQ = 16 * INT_;
constexpr int IBQ = 4 * IWQ;

__device__ __forceinline__ floatx4 mfma(const half8_t a, const half8_t b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Real spherical harmonics of degrees 0..3 (orthonormal on the sphere).
__device__ __forceinline__ float sh_coef(int k, float x, float y, float z) {
    switch (k) {
    case 0: return 0.28209479177387814f;
    case 1: return -0.48860251190291987f * y;
    case 2: return 0.48860251190291987f * z;
    case 3: return -0.48860251190291987f * x;
    case 4: return 1.0925484305920792f * (x * y);
    case 5: return -1.0925484305920792f * (y * z);
    case 6: return 0.94617469575755997f * (z * z) - 0.31539156525251999f;
    case 7: return -1.0925484305920792f * (x * z);
    case 8: return 0.54627421529603959f * (x * x - y * y);
    case 9: return 0.59004358992664352f * y * (-3.f * (x * x) + y * y);
    case 10: return 2.8906114426405538f * (x * y) * z;
    case 11: return 0.45704579946446572f * y * (1.f - 5.f * (z * z));
    case 12: return 0.3731763325901154f * z * (5.f * (z * z) - 3.f);
    case 13: return 0.45704579946446572f * x * (1.f - 5.f * (z * z));
    case 14: return 1.4453057213202769f * z * (x * x - y * y);
    default: return 0.59004358992664352f * x * (3.f * (y * y) - x * x);
    }
}

// Corners of grid level l around normalised position pn: table indices and trilinear weights
// (corner bit 0 = x, 1 = y, 2 = z). Both layouts keep a grid row along x in consecutive
// entries: dense levels are x-major, and a hashed level places row (y, z) at an offset in
// [0, table - side] drawn from a hash of (y, z) (multiply-high range reduction), so
// idx[2r + 1] == idx[2r] + 1 < table and one 8-byte load fetches each x-edge of the cell.
__device__ __forceinline__ void grid_corners(const GuideNet& net, int l, const float pn[3], uint32_t idx[8], float w[8]) {
    const uint32_t res = net.res[l];
    int ci[3];
    float fr[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float x = pn[a] * float(res);
        int c = int(floorf(x));
        c = c >= int(res) ? int(res) - 1 : (c < 0 ? 0 : c);
        ci[a] = c;
        fr[a] = x - float(c);
    }
    const uint64_t side = uint64_t(res) + 1;
    const bool dense = side * side * side <= uint64_t(net.table);
    const uint32_t span = net.table - uint32_t(side) + 1u; // row offsets; side <= table (checked at create)
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const uint32_t cx = uint32_t(ci[0] + (c & 1)), cy = uint32_t(ci[1] + ((c >> 1) & 1)),
                       cz = uint32_t(ci[2] + ((c >> 2) & 1));
        idx[c] = dense ? uint32_t(cx + side * (cy + side * cz))
                       : uint32_t((uint64_t((cy * 2654435761u) ^ (cz * 805459861u)) * span) >> 32) + cx;
        const float wx = (c & 1) ? fr[0] : 1.f - fr[0];
        const float wy = (c & 2) ? fr[1] : 1.f - fr[1];
        const float wz = (c & 4) ? fr[2] : 1.f - fr[2];
        w[c] = (wx * wy) * wz;
    }
}

__device__ __forceinline__ void normalise(const GuideNet& net, f3 p, float pn[3]) {
    pn[0] = fminf(1.f, fmaxf(0.f, (p.x - net.offset[0]) * net.scale[0]));
    pn[1] = fminf(1.f, fmaxf(0.f, (p.y - net.offset[1]) * net.scale[1]));
    pn[2] = fminf(1.f, fmaxf(0.f, (p.z - net.offset[2]) * net.scale[2]));
}

// Entries i and i + 1 of one level's table (the two x-corners of a cell edge, G_FEATS fp16
// features each) in one 16-byte load (8-byte aligned).
struct __attribute__((packed, aligned(8))) EntryPair {
    uint32_t a0, a1, b0, b1;
};
static_assert(G_FEATS == 4, "one entry = two 32-bit words");

// Grid levels of lane group g: chunk-0 slots s < 4 hold features 4g + s (level g), slots s >= 4
// hold features 16 + 4g + s - 4 (level 4 + g), as feat_of(0, g, s) = G_FEATS * level + feature.
__device__ __forceinline__ int slot_level(int g, int h) { return 4 * h + g; }

__device__ __forceinline__ void add_entry(float f[4], float w, uint32_t lo, uint32_t hi) {
    f[0] += w * half_to_float(uint16_t(lo & 0xFFFFu));
    f[1] += w * half_to_float(uint16_t(lo >> 16));
    f[2] += w * half_to_float(uint16_t(hi & 0xFFFFu));
    f[3] += w * half_to_float(uint16_t(hi >> 16));
}

// Encoding of one query by lane group g: chunk 0 = grid levels g and 4 + g (four features
// each), chunk 1 = SH coefficients 4g..4g+3, then zeros.
__device__ __forceinline__ void encode(const GuideNet& net, f3 p, f3 wo, int g, half8_t x[2]) {
    float pn[3];
    normalise(net, p, pn);
    const uint32_t* tab = reinterpret_cast<const uint32_t*>(net.grid);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int l = slot_level(g, h);
        uint32_t idx[8];
        float w[8];
        grid_corners(net, l, pn, idx, w);
        const uint32_t* lt = tab + size_t(l) * net.table * 2;
        float f[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const EntryPair v = *reinterpret_cast<const EntryPair*>(lt + 2 * size_t(idx[2 * e]));
            add_entry(f, w[2 * e], v.a0, v.a1);
            add_entry(f, w[2 * e + 1], v.b0, v.b1);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) x[0][4 * h + k] = _Float16(f[k]);
    }
#pragma unroll
    for (int s = 0; s < 4; s++) x[1][s] = _Float16(sh_coef(4 * g + s, wo.x, wo.y, wo.z));
#pragma unroll
    for (int s = 4; s < 8; s++) x[1][s] = _Float16(0.f);
}

// acc[t][ot] = W_layer (rows 16 ot..) * x[t] for ot < OT and each of the T tiles.
template <int OT, int T = NT>
__device__ __forceinline__ void layer_mfma(const half8_t* frag, int layer, const half8_t (&x)[T][2],
                                           floatx4 (&acc)[T][4], int lane) {
#pragma unroll
    for (int t = 0; t < T; t++)
#pragma unroll
        for (int ot = 0; ot < 4; ot++) acc[t][ot] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < 2; c++) {
#pragma unroll
        for (int ot = 0; ot < OT; ot++) {
            const half8_t a = frag[((layer * 4 + ot) * 2 + c) * 64 + lane];
#pragma unroll
            for (int t = 0; t < T; t++) acc[t][ot] = mfma(a, x[t][c], acc[t][ot]);
        }
    }
}

// D layout -> B operand of the next layer (ReLU, fp16).
template <int T = NT>
__device__ __forceinline__ void relu_to_operand(const floatx4 (&acc)[T][4], half8_t (&x)[T][2]) {
#pragma unroll
    for (int t = 0; t < T; t++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int s = 0; s < 8; s++) x[t][c][s] = _Float16(fmaxf(acc[t][2 * c + (s >> 2)][s & 3], 0.f));
}

// ---------------------------------------------------------------------------------------------
// inference

// Query sources: position and direction to the viewer of query i (false: no query).
struct QueueQueries {
    DeviceScene S;
    const float4 *qo, *qd, *hits;
    const uint32_t* count; // queue size on the device; null: n_fixed
    uint32_t n_fixed;
    __device__ __forceinline__ uint32_t n() const { return count ? *count : n_fixed; }
    __device__ __forceinline__ bool load(uint32_t i, f3& p, f3& wo) const {
        const float4 h = hits[i];
        const uint32_t gid = __float_as_uint(h.w);
        const float4 d = qd[i];
        wo = f3{-d.x, -d.y, -d.z};
        if (gid == 0xFFFFFFFFu) {
            p = f3{0.f, 0.f, 0.f};
            return false;
        }
        const float4* t = S.tri_shade + 4 * size_t(gid); // (p0, n0), (p1, n1), (p2, n2), ...
        const float4 a = t[0], b = t[1], c = t[2];
        const f3 p0{a.x, a.y, a.z};
        const f3 p1{b.x, b.y, b.z};
        const f3 p2{c.x, c.y, c.z};
        p = p0 + h.y * (p1 - p0) + h.z * (p2 - p0);
        return true;
    }
};
struct ArrayQueries {
    const float4 *pos, *wo;
    uint32_t count;
    __device__ __forceinline__ uint32_t n() const { return count; }
    __device__ __forceinline__ bool load(uint32_t i, f3& p, f3& w) const {
        const float4 a = pos[i], b = wo[i];
        p = f3{a.x, a.y, a.z};
        w = f3{b.x, b.y, b.z};
        return true;
    }
};

// Output rows of the inference kernels' last layer: lane group g (lanes 16g..16g+15) holds D rows
// 16 ot + 4 g + r (ot < 3, r < 4) of its query, so the last layer's rows are permuted when its
// weights are staged: slot k = 4 ot + r of group g computes output 10 g + k for k < 10 (lobes 2g
// and 2g+1, five outputs each), a zero row for k = 10, 11. Each lane then holds two whole
// lobes and no output has to move between lanes through LDS.
__device__ __forceinline__ int out_row_of(int row) {
    const int ot = row >> 4, g = (row & 15) >> 2, k = 4 * ot + (row & 3);
    return k < 10 ? 10 * g + k : -1;
}

// Sinks of one query's ten outputs held by lane group g: raw outputs or packed lobes. Every lane
// of the wave calls store (the lobe softmax reduces across the query's four lane groups).
struct RawSink {
    float* out;
    __device__ __forceinline__ void store(uint32_t i, bool valid, const float (&o)[10], int g) const {
        if (!valid) return;
#pragma unroll
        for (int k = 0; k < 10; k++) out[size_t(i) * G_OUT + 10 * g + k] = o[k];
    }
};
struct LobeSink {
    uint4* lobes;
    __device__ __forceinline__ void store(uint32_t i, bool valid, const float (&o)[10], int g) const {
        // softmax over the 8 lobe logits: two per lane, the other six in lanes +-16, +-32
        float zmax = fmaxf(o[0], o[5]);
        zmax = fmaxf(zmax, __shfl_xor(zmax, 16));
        zmax = fmaxf(zmax, __shfl_xor(zmax, 32));
        const float e0 = expf(o[0] - zmax), e1 = expf(o[5] - zmax);
        float sum = e0 + e1;
        sum += __shfl_xor(sum, 16);
        sum += __shfl_xor(sum, 32);
        if (!valid) return;
        uint32_t w[4];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const float pi = (h ? e1 : e0) / sum;
            const float kappa = expf(fminf(G_KAPPA_LOG_MAX, fmaxf(G_KAPPA_LOG_MIN, o[5 * h + 1])));
            f3 v{o[5 * h + 2], o[5 * h + 3], o[5 * h + 4]};
            const float len = sqrtf(dot(v, v));
            const f3 mu = len > 1e-6f ? v * (1.f / len) : f3{0.f, 0.f, 1.f};
            w[2 * h] = uint32_t(float_to_half(pi)) | (uint32_t(float_to_half(kappa)) << 16);
            w[2 * h + 1] = encode_normal(mu);
        }
        lobes[size_t(i) * 4 + g] = make_uint4(w[0], w[1], w[2], w[3]);
    }
};

// Inference: one 16-query tile per wave iteration, the forward weights in LDS (the last layer's
// rows permuted, out_row_of), 4 waves per block. LDS holds nothing else (40 KB: four blocks per
// CU), a query's outputs stay in the registers of the lanes that computed them, and one tile
// keeps the kernel within 128 registers without spills: four waves per SIMD hide the latency of
// the hash-grid gathers. The grid is capped at the resident blocks (1024), so the weights are
// staged once per block.
template <class Q, class Sink>
__global__ __launch_bounds__(GB) __attribute__((amdgpu_waves_per_eu(4))) void k_guide_infer(GuideNet net, Q q, Sink sink) {
    __shared__ half8_t frag[G_FRAGS * 64]; // 40 KB: forward fragments of all layers
    for (int k = threadIdx.x; k < G_FRAGS * 64; k += GB) {
        const int layer = k / (8 * 64);
        if (layer < G_LAYERS - 1) {
            frag[k] = net.frag_fwd[k];
        } else {
            // [layer][ot][c][lane']: lane' = 16 kg + rr holds row 16 ot + rr of k-group kg
            const int ot = (k >> 7) & 3, c = (k >> 6) & 1, l = k & 63;
            const int o = out_row_of(16 * ot + (l & 15));
            frag[k] = o < 0 ? half8_t{} : net.frag_fwd[((layer * 4 + (o >> 4)) * 2 + c) * 64 + (l & ~15) + (o & 15)];
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4;
    const uint32_t n = q.n();
    const uint32_t stride = gridDim.x * IBQ;
    for (uint32_t base = blockIdx.x * IBQ + wave * IWQ; base < n; base += stride) { // wave-uniform
        half8_t x[INT_][2];
        floatx4 acc[INT_][4];
        bool valid[INT_];
#pragma unroll
        for (int t = 0; t < INT_; t++) {
            const uint32_t i = base + t * 16 + (lane & 15);
            f3 p{0.f, 0.f, 0.f}, wo{0.f, 0.f, 1.f};
            valid[t] = i < n && q.load(i, p, wo);
            encode(net, p, wo, g, x[t]);
        }
#pragma unroll
        for (int layer = 0; layer < 4; layer++) {
            layer_mfma<4, INT_>(frag, layer, x, acc, lane);
            relu_to_operand<INT_>(acc, x);
        }
        layer_mfma<3, INT_>(frag, 4, x, acc, lane);
#pragma unroll
        for (int t = 0; t < INT_; t++) {
            float o[10];
#pragma unroll
            for (int k = 0; k < 10; k++) o[k] = acc[t][k >> 2][k & 3];
            sink.store(base + t * 16 + (lane & 15), valid[t], o, g);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// training

__global__ void k_guide_sum_weights(const float4* __restrict__ rec, uint32_t n, float* __restrict__ scalars) {
    __shared__ float part[1024];
    float s = 0.f;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s += rec[3 * size_t(i)].w;
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        scalars[0] = part[0];
        scalars[1] = 0.f;
    }
}

// Stores the 8 fp16 values of a B operand (features feat_of(c, g, s)) of record r.
__device__ __forceinline__ void store_operand(_Float16* base, uint32_t r, int g, const half8_t x[2]) {
    _Float16* row = base + size_t(r) * G_WIDTH;
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int h = 0; h < 2; h++) {
            typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
            half4_t v{x[c][4 * h], x[c][4 * h + 1], x[c][4 * h + 2], x[c][4 * h + 3]};
            *reinterpret_cast<half4_t*>(row + 16 * (2 * c + h) + 4 * g) = v;
        }
}

// d(-log q)/d(outputs) of one record for lobes 4*half_ .. 4*half_+3, scaled by s; returns -log q.
__device__ __forceinline__ float lobe_gradient(const float* o, f3 d, float s, int half_, float* g) {
    float zmax = o[0];
    for (int j = 1; j < G_LOBES; j++) zmax = fmaxf(zmax, o[5 * j]);
    float e[G_LOBES], esum = 0.f;
    for (int j = 0; j < G_LOBES; j++) {
        e[j] = expf(o[5 * j] - zmax);
        esum += e[j];
    }
    float pi[G_LOBES], kap[G_LOBES], f[G_LOBES], cm[G_LOBES], q = 0.f;
    f3 mu[G_LOBES];
    float len[G_LOBES];
    bool clamped[G_LOBES];
    for (int j = 0; j < G_LOBES; j++) {
        pi[j] = e[j] / esum;
        const float r = o[5 * j + 1];
        clamped[j] = r < G_KAPPA_LOG_MIN || r > G_KAPPA_LOG_MAX;
        kap[j] = expf(fminf(G_KAPPA_LOG_MAX, fmaxf(G_KAPPA_LOG_MIN, r)));
        const f3 v{o[5 * j + 2], o[5 * j + 3], o[5 * j + 4]};
        len[j] = sqrtf(dot(v, v));
        mu[j] = len[j] > 1e-6f ? v * (1.f / len[j]) : f3{0.f, 0.f, 0.f};
        cm[j] = dot(mu[j], d);
        f[j] = kap[j] / (6.2831853071795864f * -expm1f(-2.f * kap[j])) * expf(kap[j] * (cm[j] - 1.f));
        q += pi[j] * f[j];
    }
    for (int k = 0; k < 20; k++) g[k] = 0.f;
    if (!(q > 0.f)) return 0.f;
    for (int k = 0; k < 4; k++) {
        const int j = 4 * half_ + k;
        const float r = pi[j] * f[j] / q;
        const float kk = kap[j];
        g[5 * k] = s * (pi[j] - r);
        const float dlogf = 1.f / kk - 2.f / expm1f(2.f * kk) + (cm[j] - 1.f);
        g[5 * k + 1] = clamped[j] ? 0.f : s * (-r * dlogf * kk);
        if (len[j] > 1e-6f) {
            g[5 * k + 2] = s * (-r * kk * (d.x - cm[j] * mu[j].x) / len[j]);
            g[5 * k + 3] = s * (-r * kk * (d.y - cm[j] * mu[j].y) / len[j]);
            g[5 * k + 4] = s * (-r * kk * (d.z - cm[j] * mu[j].z) / len[j]);
        }
    }
    for (int k = 0; k < 20; k++) g[k] = fminf(G_DOUT_CLAMP, fmaxf(-G_DOUT_CLAMP, g[k]));
    return -logf(q);
}

// Forward and backward of BQ records per block iteration: stores every layer's input (H) and
// output gradient (D) for the weight-gradient GEMM and adds the grid gradient.
__global__ __launch_bounds__(GB) void k_guide_fwd_bwd(GuideNet net, const float4* __restrict__ rec, uint32_t n,
                                                      GuideTrainBuffers B) {
    __shared__ float outs[4][WQ][48];
    __shared__ float douts[4][WQ][48];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4;
    const float mean_w = B.scalars[0] / float(n > 0 ? n : 1);
    for (uint32_t base = blockIdx.x * BQ; base < n; base += gridDim.x * BQ) {
        half8_t x[5][NT][2];
        floatx4 acc[NT][4];
        uint32_t rid[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            rid[t] = base + wave * WQ + t * 16 + (lane & 15);
            f3 p{0.f, 0.f, 0.f}, wo{0.f, 0.f, 1.f};
            if (rid[t] < n) {
                const float4 a = rec[3 * size_t(rid[t])], b = rec[3 * size_t(rid[t]) + 1];
                p = f3{a.x, a.y, a.z};
                wo = f3{b.x, b.y, b.z};
            }
            encode(net, p, wo, g, x[0][t]);
            if (rid[t] < n) store_operand(B.H, rid[t], g, x[0][t]);
        }
#pragma unroll
        for (int layer = 0; layer < 4; layer++) {
            half8_t (&xi)[NT][2] = x[layer];
            layer_mfma<4>(net.frag_fwd, layer, xi, acc, lane);
            relu_to_operand(acc, x[layer + 1]);
#pragma unroll
            for (int t = 0; t < NT; t++)
                if (rid[t] < n) store_operand(B.H + size_t(layer + 1) * B.cap * G_WIDTH, rid[t], g, x[layer + 1][t]);
        }
        layer_mfma<3>(net.frag_fwd, 4, x[4], acc, lane);
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int ot = 0; ot < 3; ot++)
#pragma unroll
                for (int r = 0; r < 4; r++) outs[wave][t * 16 + (lane & 15)][16 * ot + 4 * g + r] = acc[t][ot][r];
        __syncthreads();
        {
            const int qi = lane >> 1, half_ = lane & 1;
            const uint32_t i = base + wave * WQ + qi;
            float gr[20];
            float nlq = 0.f, s = 0.f;
            if (i < n) {
                const float4 a = rec[3 * size_t(i)], c = rec[3 * size_t(i) + 2];
                s = mean_w > 0.f ? fminf(G_WEIGHT_SCALE_MAX, a.w / mean_w) : 0.f;
                nlq = lobe_gradient(outs[wave][qi], f3{c.x, c.y, c.z}, s, half_, gr);
            } else {
                for (int k = 0; k < 20; k++) gr[k] = 0.f;
            }
            for (int k = 0; k < 20; k++) douts[wave][qi][20 * half_ + k] = gr[k];
            if (half_ == 1)
                for (int k = 40; k < 48; k++) douts[wave][qi][k] = 0.f;
            // loss: one lane per record, one atomic per wave
            float l = (half_ == 0 && i < n) ? s * nlq : 0.f;
            for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
            if (lane == 0) atomicAdd(B.scalars + 1, l);
        }
        __syncthreads();
        // gradient at the outputs as B operand (features >= 48 are zero)
        half8_t d[NT][2];
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int s = 0; s < 8; s++) {
                    const int fo = feat_of(c, g, s);
                    d[t][c][s] = _Float16(fo < 48 ? douts[wave][t * 16 + (lane & 15)][fo] : 0.f);
                }
#pragma unroll
        for (int layer = 4; layer >= 0; layer--) {
#pragma unroll
            for (int t = 0; t < NT; t++)
                if (rid[t] < n) store_operand(B.D + size_t(layer) * B.cap * G_WIDTH, rid[t], g, d[t]);
            layer_mfma<4>(net.frag_bwd, layer, d, acc, lane); // acc = W^T d: gradient at the layer's input
            if (layer > 0) {
#pragma unroll
                for (int t = 0; t < NT; t++)
#pragma unroll
                    for (int c = 0; c < 2; c++)
#pragma unroll
                        for (int s = 0; s < 8; s++)
                            d[t][c][s] = _Float16(float(x[layer][t][c][s]) > 0.f ? acc[t][2 * c + (s >> 2)][s & 3] : 0.f);
            }
        }
        // acc = gradient at the encoding: in-tile 0 holds features 4g+r (level g), in-tile 1
        // holds 16+4g+r (level 4+g)
#pragma unroll
        for (int t = 0; t < NT; t++) {
            if (rid[t] >= n) continue;
            const float4 a = rec[3 * size_t(rid[t])];
            float pn[3];
            normalise(net, f3{a.x, a.y, a.z}, pn);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int l = slot_level(g, h);
                uint32_t idx[8];
                float w[8];
                grid_corners(net, l, pn, idx, w);
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    unsigned long long* gg = B.grid_grad + (size_t(l) * net.table + idx[c]) * G_FEATS;
#pragma unroll
                    for (int k = 0; k < G_FEATS; k++) {
                        const float v = w[c] * acc[t][h][k];
                        // non-finite contributions are dropped (they would poison Adam's moments)
                        if (v != 0.f && fabsf(v) < __builtin_inff()) {
                            const float cv = fminf(fmaxf(v, -G_GRID_FX_CLAMP), G_GRID_FX_CLAMP);
                            atomicAdd(gg + k, (unsigned long long)__float2ll_rn(cv * G_GRID_FX));
                        }
                    }
                }
            }
        }
    }
}

// dW_layer = sum over records of D_layer (outputs) x H_layer (inputs): per block a range of
// G_WGRAD_RECORDS records, staged 32 at a time through LDS as [feature][record]; wave w
// computes output rows 16w..16w+15 for all 64 inputs.
__global__ __launch_bounds__(GB) void k_guide_wgrad(GuideTrainBuffers B, uint32_t n) {
    __shared__ __attribute__((aligned(16))) _Float16 Dt[64][40]; // rows of 80 bytes: 16-byte aligned reads
    __shared__ __attribute__((aligned(16))) _Float16 Ht[64][40];
    const int layer = blockIdx.y, kb = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4;
    const _Float16* D = B.D + size_t(layer) * B.cap * G_WIDTH;
    const _Float16* H = B.H + size_t(layer) * B.cap * G_WIDTH;
    const uint32_t r0 = kb * G_WGRAD_RECORDS;
    const uint32_t r1 = min(n, r0 + G_WGRAD_RECORDS);
    floatx4 acc[4];
    for (int it = 0; it < 4; it++) acc[it] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (uint32_t rb = r0; rb < r1; rb += 32) {
        {
            const int r = threadIdx.x >> 3, f0 = (threadIdx.x & 7) * 8;
            const uint32_t rr = rb + r;
            half8_t dv = {}, hv = {};
            if (rr < r1) {
                dv = *reinterpret_cast<const half8_t*>(D + size_t(rr) * G_WIDTH + f0);
                hv = *reinterpret_cast<const half8_t*>(H + size_t(rr) * G_WIDTH + f0);
            }
#pragma unroll
            for (int j = 0; j < 8; j++) {
                Dt[f0 + j][r] = dv[j];
                Ht[f0 + j][r] = hv[j];
            }
        }
        __syncthreads();
        const half8_t a = *reinterpret_cast<const half8_t*>(&Dt[16 * wave + (lane & 15)][8 * g]);
#pragma unroll
        for (int it = 0; it < 4; it++) {
            const half8_t b = *reinterpret_cast<const half8_t*>(&Ht[16 * it + (lane & 15)][8 * g]);
            acc[it] = mfma(a, b, acc[it]);
        }
        __syncthreads();
    }
    float* P = B.partials + (size_t(layer) * B.kblocks + kb) * G_WIDTH * G_WIDTH;
#pragma unroll
    for (int it = 0; it < 4; it++)
#pragma unroll
        for (int r = 0; r < 4; r++) P[(16 * wave + 4 * g + r) * G_WIDTH + 16 * it + (lane & 15)] = acc[it][r];
}

__global__ void k_guide_reduce_wgrad(GuideTrainBuffers B, float* __restrict__ dW) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= uint32_t(G_LAYERS * G_WIDTH * G_WIDTH)) return;
    const uint32_t layer = e / (G_WIDTH * G_WIDTH), k = e % (G_WIDTH * G_WIDTH);
    float s = 0.f;
    for (uint32_t kb = 0; kb < B.kblocks; kb++) s += B.partials[(size_t(layer) * B.kblocks + kb) * G_WIDTH * G_WIDTH + k];
    dW[e] = s;
}

// Fragment positions of weight W_layer[o][i]: forward (A = W) and backward (A = W^T).
__device__ __forceinline__ void frag_positions(int layer, int o, int i, uint32_t& fwd, int& fs, uint32_t& bwd, int& bs) {
    {
        const int ti = i >> 4, c = ti >> 1, g = (i & 15) >> 2;
        fs = 4 * (ti & 1) + (i & 3);
        fwd = uint32_t(((layer * 4 + (o >> 4)) * 2 + c) * 64 + 16 * g + (o & 15));
    }
    {
        const int to = o >> 4, c = to >> 1, g = (o & 15) >> 2;
        bs = 4 * (to & 1) + (o & 3);
        bwd = uint32_t(((layer * 4 + (i >> 4)) * 2 + c) * 64 + 16 * g + (i & 15));
    }
}

__global__ void k_guide_adam_w(GuideParams p, GuideTrainBuffers B, float lr, float scale, float bc1, float bc2,
                               int update) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= uint32_t(G_LAYERS * G_WIDTH * G_WIDTH)) return;
    const int layer = int(e / (G_WIDTH * G_WIDTH)), o = int((e / G_WIDTH) % G_WIDTH), i = int(e % G_WIDTH);
    float w = p.w[e];
    if (update && !(layer == G_LAYERS - 1 && o >= G_OUT)) {
        float gsum = 0.f;
        for (uint32_t kb = 0; kb < B.kblocks; kb++)
            gsum += B.partials[(size_t(layer) * B.kblocks + kb) * G_WIDTH * G_WIDTH + uint32_t(o) * G_WIDTH + i];
        const float gr = gsum * scale;
        const float m = 0.9f * p.w_m[e] + 0.1f * gr;
        const float v = 0.99f * p.w_v[e] + 0.01f * gr * gr;
        p.w_m[e] = m;
        p.w_v[e] = v;
        w -= lr * (m / bc1) / (sqrtf(v / bc2) + 1e-10f);
        p.w[e] = w;
    }
    uint32_t fwd, bwd;
    int fs, bs;
    frag_positions(layer, o, i, fwd, fs, bwd, bs);
    reinterpret_cast<_Float16*>(p.frag_fwd + fwd)[fs] = _Float16(w);
    reinterpret_cast<_Float16*>(p.frag_bwd + bwd)[bs] = _Float16(w);
}

__global__ void k_guide_adam_grid(GuideParams p, unsigned long long* __restrict__ grad, float lr, float scale,
                                  float bc1, float bc2, int update) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < p.grid_count; e += stride) {
        float w = p.g[e];
        if (update) {
            const float gr = float((long long)grad[e]) * (1.f / G_GRID_FX) * scale;
            grad[e] = 0ull;
            float m = p.g_m[e], v = p.g_v[e];
            if (gr != 0.f || m != 0.f) {
                m = 0.9f * m + 0.1f * gr;
                v = 0.99f * v + 0.01f * gr * gr;
                p.g_m[e] = m;
                p.g_v[e] = v;
                w -= lr * (m / bc1) / (sqrtf(v / bc2) + 1e-10f);
                p.g[e] = w;
            }
        }
        p.grid[e] = _Float16(w);
    }
}

// Sort keys (the third float4's w of each record) and their record indices.
__global__ void k_guide_record_keys(const float4* __restrict__ in, uint32_t n, uint32_t* __restrict__ keys,
                                    uint32_t* __restrict__ vals) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        keys[r] = __float_as_uint(in[3 * size_t(r) + 2].w);
        vals[r] = r;
    }
}

// Record k of the sorted order -> slot (k * 2654435761) mod n (2654435761 is prime and n < 2^32,
// so the map is a permutation); the key is cleared.
__global__ void k_guide_record_place(const float4* __restrict__ in, const uint32_t* __restrict__ order, uint32_t n,
                                     float4* __restrict__ out) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const float4* a = in + 3 * size_t(order[k]);
        float4* b = out + 3 * ((uint64_t(k) * 2654435761ull) % n);
        const float4 c = a[2];
        b[0] = a[0];
        b[1] = a[1];
        b[2] = make_float4(c.x, c.y, c.z, 0.f);
    }
}

int blocks_for(uint64_t n, int per_block, int cap) {
    const uint64_t b = (n + per_block - 1) / per_block;
    return int(b < uint64_t(cap) ? (b > 0 ? b : 1) : cap);
}

} // namespace

hipError_t launch_guide_infer_queue(const GuideNet& net, const DeviceScene& s, const float4* qo, const float4* qd,
                                    const float4* hits, const uint32_t* count, uint32_t n_max, uint4* lobes,
                                    hipStream_t st) {
    if (n_max == 0) return hipSuccess;
    const QueueQueries q{s, qo, qd, hits, count, n_max};
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_guide_infer<QueueQueries, LobeSink>), dim3(blocks_for(n_max, IBQ, 1024)),
                       dim3(GB), 0, st, net, q, LobeSink{lobes});
    return hipGetLastError();
}

hipError_t launch_guide_forward(const GuideNet& net, const float4* pos, const float4* wo, uint32_t n, float* out,
                                hipStream_t st) {
    if (n == 0) return hipSuccess;
    const ArrayQueries q{pos, wo, n};
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_guide_infer<ArrayQueries, RawSink>), dim3(blocks_for(n, IBQ, 1024)), dim3(GB),
                       0, st, net, q, RawSink{out});
    return hipGetLastError();
}

hipError_t launch_guide_gradient(const GuideNet& net, const float4* records, uint32_t n, GuideTrainBuffers& b,
                                 hipStream_t st) {
    if (n > b.cap) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_guide_sum_weights, dim3(1), dim3(1024), 0, st, records, n, b.scalars);
    if (n == 0) return hipGetLastError();
    hipLaunchKernelGGL(k_guide_fwd_bwd, dim3(blocks_for(n, BQ, 1024)), dim3(GB), 0, st, net, records, n, b);
    const uint32_t kb = (n + G_WGRAD_RECORDS - 1) / G_WGRAD_RECORDS;
    if (kb > b.kblocks) return hipErrorInvalidValue;
    // the partials are laid out for exactly kb blocks; the reductions read the same layout
    b.kblocks = kb;
    hipLaunchKernelGGL(k_guide_wgrad, dim3(kb, G_LAYERS), dim3(GB), 0, st, b, n);
    return hipGetLastError();
}

hipError_t launch_guide_reduce_wgrad(const GuideTrainBuffers& b, float* dW, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_reduce_wgrad, dim3((G_LAYERS * G_WIDTH * G_WIDTH + 255) / 256), dim3(256), 0, st, b, dW);
    return hipGetLastError();
}

hipError_t launch_guide_adam(const GuideParams& p, const GuideTrainBuffers& b, float lr, float scale, float bc1,
                             float bc2, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_adam_w, dim3((G_LAYERS * G_WIDTH * G_WIDTH + 255) / 256), dim3(256), 0, st, p, b, lr,
                       scale, bc1, bc2, 1);
    hipLaunchKernelGGL(k_guide_adam_grid, dim3(blocks_for(p.grid_count, 256, 8192)), dim3(256), 0, st, p, b.grid_grad,
                       lr, scale, bc1, bc2, 1);
    return hipGetLastError();
}

hipError_t launch_guide_pack(const GuideParams& p, hipStream_t st) {
    GuideTrainBuffers none{};
    hipLaunchKernelGGL(k_guide_adam_w, dim3((G_LAYERS * G_WIDTH * G_WIDTH + 255) / 256), dim3(256), 0, st, p, none,
                       0.f, 0.f, 1.f, 1.f, 0);
    hipLaunchKernelGGL(k_guide_adam_grid, dim3(blocks_for(p.grid_count, 256, 8192)), dim3(256), 0, st, p, nullptr,
                       0.f, 0.f, 1.f, 1.f, 0);
    return hipGetLastError();
}

size_t launch_guide_order_temp_bytes(uint32_t cap) {
    size_t bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, bytes, static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                  static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr), cap, 0u, 32u) !=
        hipSuccess)
        return 0;
    return bytes;
}

hipError_t launch_guide_order_records(const float4* in, uint32_t n, uint32_t* keys, uint32_t* vals, void* temp,
                                      size_t temp_bytes, float4* out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const int g = blocks_for(n, 256, 4096);
    hipLaunchKernelGGL(k_guide_record_keys, dim3(g), dim3(256), 0, st, in, n, keys, vals);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t bytes = temp_bytes;
    e = rocprim::radix_sort_pairs(temp, bytes, keys, keys + n, vals, vals + n, n, 0u, 32u, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_guide_record_place, dim3(g), dim3(256), 0, st, in, vals + n, n, out);
    return hipGetLastError();
}

} // namespace mrt
