// Learned guide: device view of the network, the packed lobe format the shading kernel samples
// from, and the launchers of guide.hip.
//
// MFMA layout (v_mfma_f32_16x16x32_f16, one wave): A is 16x32 (lane l holds row l%16,
// k = 8*(l/16)+s), B is 32x16 (lane l holds column l%16, same k), D is 16x16 (lane l holds
// column l%16, rows 4*(l/16)+i). The network runs transposed, D = W * X^T: rows are features,
// columns are queries. A lane of D holds features 16*t + 4*g + i (t = row tile, g = l/16) of
// query l%16, which is exactly what the same lane needs as B operand of the next layer if the
// k-slots of chunk c are assigned to features feat_of(c, g, s) = 16*(2c + s/4) + 4g + s%4.
// The weight fragments are stored with the same assignment, so activations never move
// between lanes from layer to layer.
#pragma once

#include <hip/hip_runtime.h>

#include "mrt_device.h"
#include "mrt_kernels.h"

namespace mrt {

constexpr int G_LEVELS = 8, G_FEATS = 4, G_WIDTH = 64, G_LAYERS = 5, G_OUT = 40, G_LOBES = 8; // G_FEATS per entry
constexpr float G_KAPPA_LOG_MIN = -2.302585093f; // kappa in [0.1, 5000]
constexpr float G_KAPPA_LOG_MAX = 8.517193191f;
constexpr float G_DOUT_CLAMP = 4096.f;
constexpr float G_WEIGHT_SCALE_MAX = 64.f;
constexpr int G_FRAGS = G_LAYERS * 4 * 2; // [layer][row tile][k chunk], 64 lanes x 8 halves each

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));

__host__ __device__ inline int feat_of(int c, int g, int s) { return 16 * (2 * c + (s >> 2)) + 4 * g + (s & 3); }

// Network view of the kernels.
struct GuideNet {
    const half8_t* frag_fwd; // W_k fragments: [layer][out tile][chunk][lane]
    const half8_t* frag_bwd; // W_k^T fragments: [layer][in tile][chunk][lane]
    const _Float16* grid;    // [level][table][G_FEATS]
    uint32_t res[G_LEVELS];
    uint32_t table;
    float offset[3], scale[3];
};

// One packed lobe: weight (fp16), kappa (fp16), mean direction (oct, 2 x snorm16).
// A query's 8 lobes are 64 bytes (4 x uint4).
struct GuideLobes {
    float pi[G_LOBES], kappa[G_LOBES];
    f3 mu[G_LOBES];
    float pi_sum;
};

__device__ __forceinline__ void unpack_lobes(const uint4* p, GuideLobes& L) {
    L.pi_sum = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint4 v = p[k];
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int j = 2 * k + h;
            L.pi[j] = half_to_float(uint16_t(w[2 * h] & 0xFFFFu));
            L.kappa[j] = half_to_float(uint16_t(w[2 * h] >> 16));
            L.mu[j] = decode_normal(w[2 * h + 1]);
            L.pi_sum += L.pi[j];
        }
    }
}

__device__ __forceinline__ float vmf_pdf(float kappa, float cos_mu) {
    return kappa / (6.2831853071795864f * -expm1f(-2.f * kappa)) * expf(kappa * (cos_mu - 1.f));
}

// Density of the mixture (weights normalised by their sum).
__device__ __forceinline__ float lobes_pdf(const GuideLobes& L, f3 d) {
    if (!(L.pi_sum > 0.f)) return 0.f;
    float q = 0.f;
#pragma unroll
    for (int j = 0; j < G_LOBES; j++) q += L.pi[j] * vmf_pdf(L.kappa[j], dot(L.mu[j], d));
    return q / L.pi_sum;
}

// Samples the mixture: lobe by weight (u0), then the vMF lobe (u1, u2):
// cos = 1 + log(u1 + (1 - u1) exp(-2 kappa)) / kappa, azimuth 2 pi u2 around mu.
__device__ __forceinline__ f3 lobes_sample(const GuideLobes& L, float u0, float u1, float u2) {
    // The chosen lobe's kappa and mu ride along in selects while the weights are walked, so no
    // lobe index exists that the compiler could turn into an indexed (scratch-resident) array.
    float t = u0 * L.pi_sum;
    float kappa = L.kappa[G_LOBES - 1];
    f3 mu = L.mu[G_LOBES - 1];
    bool found = false;
#pragma unroll
    for (int k = 0; k < G_LOBES - 1; k++) {
        const bool take = !found && t < L.pi[k];
        kappa = take ? L.kappa[k] : kappa;
        mu.x = take ? L.mu[k].x : mu.x;
        mu.y = take ? L.mu[k].y : mu.y;
        mu.z = take ? L.mu[k].z : mu.z;
        if (!found && !take) t -= L.pi[k];
        found = found || take;
    }
    float c = 1.f + logf(u1 + (1.f - u1) * expf(-2.f * kappa)) / kappa;
    c = fminf(1.f, fmaxf(-1.f, c));
    const float s = sqrtf(fmaxf(0.f, 1.f - c * c));
    const Frame fr = frame_from_normal(mu);
    const float phi = 6.2831853071795864f * u2;
    return fr.to_world(f3{s * cosf(phi), s * sinf(phi), c});
}

// Launchers (guide.hip).
// Lobes of the queued hits [0, *count) of one bounce (count null: [0, n_max)): 64 bytes per
// queue entry.
hipError_t launch_guide_infer_queue(const GuideNet& net, const DeviceScene& s, const float4* qo, const float4* qd,
                                    const float4* hits, const uint32_t* count, uint32_t n_max, uint4* lobes,
                                    hipStream_t st);
// Network outputs for n queries (pos, wo as float4 xyz): 40 floats per query.
hipError_t launch_guide_forward(const GuideNet& net, const float4* pos, const float4* wo, uint32_t n, float* out,
                                hipStream_t st);

// Training buffers of one step over up to `cap` records.
struct GuideTrainBuffers {
    _Float16* H;       // [layer][cap][64] inputs of each layer
    _Float16* D;       // [layer][cap][64] gradients at each layer's outputs
    float* partials;   // [layer][kblocks][64][64]
    // [level][table][G_FEATS] grid gradient as 64-bit fixed point (G_GRID_FX units): integer
    // atomics sum it exactly, so the sum does not depend on the order the records arrive in
    unsigned long long* grid_grad;
    float* scalars;    // [0] weight sum, [1] loss sum
    uint32_t cap, kblocks;
};
constexpr uint32_t G_WGRAD_RECORDS = 2048; // records per weight-gradient block
constexpr float G_GRID_FX = 16777216.f;      // grid-gradient fixed point: 2^24 units per 1
constexpr float G_GRID_FX_CLAMP = 4096.f;    // one corner's contribution, as the output gradients

// A frame's training records in a fixed order: the records (3 float4 each; the third's w holds
// the unique key pid * 64 + bounce) sorted by key, then record k of the sorted order written to
// slot (k * 2654435761) mod n of `out`, so consecutive training steps mix pixels and bounces.
// `temp` holds temp_bytes (launch_guide_order_temp_bytes(cap)); keys/vals hold 2 * cap entries.
size_t launch_guide_order_temp_bytes(uint32_t cap);
hipError_t launch_guide_order_records(const float4* in, uint32_t n, uint32_t* keys, uint32_t* vals, void* temp,
                                      size_t temp_bytes, float4* out, hipStream_t st);

// Records: 3 float4 each, (pos, weight), (wo, 0), (dir, 0). Computes the weight sum, the
// forward and backward passes, the grid gradient (added to grid_grad) and the weight-gradient
// partials of n <= cap records.
hipError_t launch_guide_gradient(const GuideNet& net, const float4* records, uint32_t n, GuideTrainBuffers& b,
                                 hipStream_t st);
// Sums the weight-gradient partials into dW [5][64][64] (sum over records).
hipError_t launch_guide_reduce_wgrad(const GuideTrainBuffers& b, float* dW, hipStream_t st);

struct GuideParams {
    float* w;      // master weights [5][64][64]
    float* w_m;    // Adam moments
    float* w_v;
    half8_t* frag_fwd;
    half8_t* frag_bwd;
    float* g;      // master grid [G_LEVELS][table][G_FEATS]
    float* g_m;
    float* g_v;
    _Float16* grid;
    uint64_t grid_count;
};
// Adam step (scale multiplies the summed gradients; bc1/bc2 = bias corrections), then repacks.
hipError_t launch_guide_adam(const GuideParams& p, const GuideTrainBuffers& b, float lr, float scale, float bc1,
                             float bc2, hipStream_t st);
// Rebuilds the fp16 fragments and grid from the master parameters.
hipError_t launch_guide_pack(const GuideParams& p, hipStream_t st);

} // namespace mrt
