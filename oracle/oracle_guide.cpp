// CPU oracle of the learned guide — TEST INFRASTRUCTURE ONLY (tests/, never the product).
//
// Restates the guide network of include/merian_amd.h (hash-grid + SH encoding, 4 hidden ReLU
// layers of width 64, 40 outputs decoded into 8 von Mises-Fisher lobes) and the gradient of
// its training loss, in two modes:
//   mirror: the rounding points of the GPU kernels (fp16 weights, grid features, activations
//           and output gradients; fp32 accumulation) — compared with the GPU within tolerance;
//   exact:  double precision without rounding — pinned by finite differences (tests).
// The reference has no learned guide; its vMF lobe (von-mises-fisher.slang:29-45) uses the
// same density kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.w - 1)).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle_math.h"

namespace orc_guide {

constexpr int LEVELS = 8, FEATS = 4, WIDTH = 64, LAYERS = 5, OUTPUTS = 40, LOBES = 8; // FEATS per grid entry
constexpr float KAPPA_LOG_MIN = -2.302585093f; // kappa in [0.1, 5000]
constexpr float KAPPA_LOG_MAX = 8.517193191f;
constexpr float DOUT_CLAMP = 4096.f;           // output gradients are clamped to +-4096 (fp16 range)
constexpr float WEIGHT_SCALE_MAX = 64.f;       // per-record weight relative to the batch mean

struct Spec {
    uint32_t res[LEVELS];
    uint32_t table;
    float offset[3], scale[3];
};

inline float r16(float x) { return orc::f16_to_f32(orc::f32_to_f16(x)); }

template <bool MIRROR> struct Num;
template <> struct Num<true> {
    using T = float;
    static float round(float x) { return r16(x); }
};
template <> struct Num<false> {
    using T = double;
    static double round(double x) { return x; }
};

// Real spherical harmonics of degrees 0..3 (orthonormal on the sphere).
template <class T> void sh16(T x, T y, T z, T* o) {
    o[0] = T(0.28209479177387814);
    o[1] = T(-0.48860251190291987) * y;
    o[2] = T(0.48860251190291987) * z;
    o[3] = T(-0.48860251190291987) * x;
    o[4] = T(1.0925484305920792) * (x * y);
    o[5] = T(-1.0925484305920792) * (y * z);
    o[6] = T(0.94617469575755997) * (z * z) - T(0.31539156525251999);
    o[7] = T(-1.0925484305920792) * (x * z);
    o[8] = T(0.54627421529603959) * (x * x - y * y);
    o[9] = T(0.59004358992664352) * y * (T(-3) * (x * x) + y * y);
    o[10] = T(2.8906114426405538) * (x * y) * z;
    o[11] = T(0.45704579946446572) * y * (T(1) - T(5) * (z * z));
    o[12] = T(0.3731763325901154) * z * (T(5) * (z * z) - T(3));
    o[13] = T(0.45704579946446572) * x * (T(1) - T(5) * (z * z));
    o[14] = T(1.4453057213202769) * z * (x * x - y * y);
    o[15] = T(0.59004358992664352) * x * (T(3) * (y * y) - x * x);
}

// Grid corner of one level: 8 table indices and trilinear weights (corner bit 0 = x, 1 = y, 2 = z).
// Dense levels are x-major; a hashed level stores grid row (y, z) as a run of side entries at
// offset floor(hash(y, z) * (table - side + 1) / 2^32), so a cell's two x-corners are
// consecutive entries.
template <class T> void corners(const Spec& s, int l, const float* pn, uint32_t* idx, T* w) {
    const uint32_t res = s.res[l];
    int ci[3];
    T fr[3];
    for (int a = 0; a < 3; a++) {
        const float x = pn[a] * float(res);
        int c = int(std::floor(x));
        if (c >= int(res)) c = int(res) - 1;
        if (c < 0) c = 0;
        ci[a] = c;
        fr[a] = T(x) - T(c);
    }
    const uint64_t side = uint64_t(res) + 1;
    const bool dense = side * side * side <= s.table;
    const uint32_t span = s.table - uint32_t(side) + 1u;
    for (int c = 0; c < 8; c++) {
        const uint32_t cx = uint32_t(ci[0] + (c & 1)), cy = uint32_t(ci[1] + ((c >> 1) & 1)),
                       cz = uint32_t(ci[2] + ((c >> 2) & 1));
        idx[c] = dense ? uint32_t(cx + side * (cy + side * cz))
                       : uint32_t((uint64_t((cy * 2654435761u) ^ (cz * 805459861u)) * span) >> 32) + cx;
        const T wx = (c & 1) ? fr[0] : T(1) - fr[0];
        const T wy = (c & 2) ? fr[1] : T(1) - fr[1];
        const T wz = (c & 4) ? fr[2] : T(1) - fr[2];
        w[c] = (wx * wy) * wz;
    }
}

inline void normalise(const Spec& s, const float* p, float* pn) {
    for (int a = 0; a < 3; a++) pn[a] = std::min(1.f, std::max(0.f, (p[a] - s.offset[a]) * s.scale[a]));
}

// Network state of one query (inputs of every layer, outputs).
template <class T> struct Act {
    T h[LAYERS][WIDTH]; // h[k] = input of layer k (h[0] = encoding)
    T out[WIDTH];
    uint32_t idx[LEVELS][8];
    T cw[LEVELS][8];
};

template <bool MIRROR>
void forward(const Spec& s, const std::vector<typename Num<MIRROR>::T>& W, const std::vector<typename Num<MIRROR>::T>& G,
             const float* p, const float* wo, Act<typename Num<MIRROR>::T>& a) {
    using T = typename Num<MIRROR>::T;
    float pn[3];
    normalise(s, p, pn);
    for (int l = 0; l < LEVELS; l++) {
        corners<T>(s, l, pn, a.idx[l], a.cw[l]);
        for (int f = 0; f < FEATS; f++) {
            T acc = 0;
            for (int c = 0; c < 8; c++) acc += a.cw[l][c] * G[(size_t(l) * s.table + a.idx[l][c]) * FEATS + f];
            a.h[0][FEATS * l + f] = Num<MIRROR>::round(acc);
        }
    }
    T sh[16];
    sh16<T>(T(wo[0]), T(wo[1]), T(wo[2]), sh);
    for (int k = 0; k < 16; k++) a.h[0][32 + k] = Num<MIRROR>::round(sh[k]);
    for (int k = 48; k < 64; k++) a.h[0][k] = 0;
    for (int k = 0; k < LAYERS; k++) {
        const int n_out = k + 1 < LAYERS ? WIDTH : OUTPUTS;
        for (int o = 0; o < WIDTH; o++) {
            T acc = 0;
            if (o < n_out)
                for (int i = 0; i < WIDTH; i++) acc += W[(size_t(k) * WIDTH + o) * WIDTH + i] * a.h[k][i];
            if (k + 1 < LAYERS) {
                a.h[k + 1][o] = Num<MIRROR>::round(acc > T(0) ? acc : T(0));
            } else {
                a.out[o] = acc;
            }
        }
    }
}

// Lobes of the output vector: weights (softmax), kappa, unit mean direction, and the raw
// quantities the gradient needs.
template <class T> struct Lobes {
    T pi[LOBES], kappa[LOBES], mu[LOBES][3], vlen[LOBES];
    bool kappa_clamped[LOBES];
};

template <class T> void decode(const T* out, Lobes<T>& L) {
    T zmax = out[0];
    for (int j = 1; j < LOBES; j++) zmax = std::max(zmax, out[5 * j]);
    T sum = 0;
    for (int j = 0; j < LOBES; j++) {
        L.pi[j] = std::exp(out[5 * j] - zmax);
        sum += L.pi[j];
    }
    for (int j = 0; j < LOBES; j++) {
        L.pi[j] /= sum;
        const T r = out[5 * j + 1];
        L.kappa_clamped[j] = r < T(KAPPA_LOG_MIN) || r > T(KAPPA_LOG_MAX);
        L.kappa[j] = std::exp(std::min(T(KAPPA_LOG_MAX), std::max(T(KAPPA_LOG_MIN), r)));
        const T vx = out[5 * j + 2], vy = out[5 * j + 3], vz = out[5 * j + 4];
        const T len = std::sqrt(vx * vx + vy * vy + vz * vz);
        L.vlen[j] = len;
        const T inv = len > T(1e-6) ? T(1) / len : T(0);
        L.mu[j][0] = vx * inv;
        L.mu[j][1] = vy * inv;
        L.mu[j][2] = vz * inv;
    }
}

template <class T> T vmf(T kappa, T cos_mu) {
    return kappa / (T(2.0 * M_PI) * -std::expm1(T(-2) * kappa)) * std::exp(kappa * (cos_mu - T(1)));
}

template <class T> T mixture_pdf(const Lobes<T>& L, const T* d) {
    T q = 0;
    for (int j = 0; j < LOBES; j++) {
        const T c = L.mu[j][0] * d[0] + L.mu[j][1] * d[1] + L.mu[j][2] * d[2];
        q += L.pi[j] * vmf(L.kappa[j], c);
    }
    return q;
}

// d(-log q(d)) / d(out), scaled by s.
template <class T> void output_gradient(const T* out, const T* d, T s, T* g) {
    Lobes<T> L;
    decode(out, L);
    T f[LOBES], q = 0;
    for (int j = 0; j < LOBES; j++) {
        const T c = L.mu[j][0] * d[0] + L.mu[j][1] * d[1] + L.mu[j][2] * d[2];
        f[j] = vmf(L.kappa[j], c);
        q += L.pi[j] * f[j];
    }
    for (int o = 0; o < WIDTH; o++) g[o] = 0;
    if (!(q > T(0))) return;
    for (int j = 0; j < LOBES; j++) {
        const T r = L.pi[j] * f[j] / q;
        const T k = L.kappa[j];
        const T c = L.mu[j][0] * d[0] + L.mu[j][1] * d[1] + L.mu[j][2] * d[2];
        g[5 * j] = s * (L.pi[j] - r);
        const T dlogf_dk = T(1) / k - T(2) / std::expm1(T(2) * k) + (c - T(1));
        g[5 * j + 1] = L.kappa_clamped[j] ? T(0) : s * (-r * dlogf_dk * k);
        if (L.vlen[j] > T(1e-6)) {
            for (int a = 0; a < 3; a++) g[5 * j + 2 + a] = s * (-r * k * (d[a] - c * L.mu[j][a]) / L.vlen[j]);
        }
    }
    for (int o = 0; o < OUTPUTS; o++) g[o] = std::min(T(DOUT_CLAMP), std::max(T(-DOUT_CLAMP), g[o]));
}

template <bool MIRROR>
float gradient(const Spec& s, const float* Wf, const float* Gf, const float* rec, uint32_t n, float* dW, float* dG) {
    using T = typename Num<MIRROR>::T;
    std::vector<T> W(size_t(LAYERS) * WIDTH * WIDTH), G(size_t(LEVELS) * s.table * FEATS);
    for (size_t i = 0; i < W.size(); i++) W[i] = Num<MIRROR>::round(T(Wf[i]));
    for (size_t i = 0; i < G.size(); i++) G[i] = Num<MIRROR>::round(T(Gf[i]));
    std::vector<T> gW(W.size(), T(0)), gG(G.size(), T(0));
    T wsum = 0;
    for (uint32_t r = 0; r < n; r++) wsum += T(rec[12 * r + 3]);
    const T mean_w = n ? wsum / T(n) : T(0);
    T loss = 0;
    Act<T> a;
    for (uint32_t r = 0; r < n; r++) {
        const float* R = rec + 12 * r;
        const T s_r = mean_w > T(0) ? std::min(T(WEIGHT_SCALE_MAX), T(R[3]) / mean_w) : T(0);
        forward<MIRROR>(s, W, G, R, R + 4, a);
        const T d[3] = {T(R[8]), T(R[9]), T(R[10])};
        Lobes<T> L;
        decode(a.out, L);
        const T q = mixture_pdf(L, d);
        if (q > T(0)) loss += s_r * -std::log(q);
        T g[WIDTH], dh[WIDTH], dp[WIDTH];
        output_gradient<T>(a.out, d, s_r, g);
        for (int o = 0; o < WIDTH; o++) dp[o] = Num<MIRROR>::round(g[o]);
        for (int k = LAYERS - 1; k >= 0; k--) {
            // dp = gradient at the pre-activation outputs of layer k (as used by the kernels)
            for (int o = 0; o < WIDTH; o++)
                for (int i = 0; i < WIDTH; i++) gW[(size_t(k) * WIDTH + o) * WIDTH + i] += dp[o] * a.h[k][i];
            for (int i = 0; i < WIDTH; i++) {
                T acc = 0;
                for (int o = 0; o < WIDTH; o++) acc += W[(size_t(k) * WIDTH + o) * WIDTH + i] * dp[o];
                dh[i] = acc;
            }
            if (k > 0)
                for (int i = 0; i < WIDTH; i++) dp[i] = Num<MIRROR>::round(a.h[k][i] > T(0) ? dh[i] : T(0));
        }
        // dh = gradient at the encoding; scatter the grid part to the 8 corners of each level
        for (int l = 0; l < LEVELS; l++)
            for (int c = 0; c < 8; c++)
                for (int f = 0; f < FEATS; f++)
                    gG[(size_t(l) * s.table + a.idx[l][c]) * FEATS + f] += a.cw[l][c] * dh[FEATS * l + f];
    }
    if (dW)
        for (size_t i = 0; i < gW.size(); i++) dW[i] = float(gW[i]);
    if (dG)
        for (size_t i = 0; i < gG.size(); i++) dG[i] = float(gG[i]);
    return n ? float(loss / T(n)) : 0.f;
}

} // namespace orc_guide

using namespace orc_guide;

extern "C" {

// Spec layout: res[8], table, offset[3], scale[3] (host struct of 15 words).
void orc_guide_forward(const Spec* s, const float* Wf, const float* Gf, const float* pos, const float* wo, uint32_t n,
                       float* out, float* encoding, int exact) {
    if (exact) {
        std::vector<double> W(Wf, Wf + size_t(LAYERS) * WIDTH * WIDTH), G(Gf, Gf + size_t(LEVELS) * s->table * FEATS);
        Act<double> a;
        for (uint32_t q = 0; q < n; q++) {
            forward<false>(*s, W, G, pos + 3 * q, wo + 3 * q, a);
            for (int o = 0; o < OUTPUTS; o++) out[OUTPUTS * q + o] = float(a.out[o]);
            if (encoding)
                for (int i = 0; i < WIDTH; i++) encoding[WIDTH * q + i] = float(a.h[0][i]);
        }
        return;
    }
    std::vector<float> W(size_t(LAYERS) * WIDTH * WIDTH), G(size_t(LEVELS) * s->table * FEATS);
    for (size_t i = 0; i < W.size(); i++) W[i] = r16(Wf[i]);
    for (size_t i = 0; i < G.size(); i++) G[i] = r16(Gf[i]);
    Act<float> a;
    for (uint32_t q = 0; q < n; q++) {
        forward<true>(*s, W, G, pos + 3 * q, wo + 3 * q, a);
        for (int o = 0; o < OUTPUTS; o++) out[OUTPUTS * q + o] = a.out[o];
        if (encoding)
            for (int i = 0; i < WIDTH; i++) encoding[WIDTH * q + i] = a.h[0][i];
    }
}

float orc_guide_gradient(const Spec* s, const float* Wf, const float* Gf, const float* records, uint32_t n,
                         float* dW, float* dG, int exact) {
    return exact ? gradient<false>(*s, Wf, Gf, records, n, dW, dG) : gradient<true>(*s, Wf, Gf, records, n, dW, dG);
}

void orc_guide_sh(const float* d, float* out) {
    double o[16];
    sh16<double>(d[0], d[1], d[2], o);
    for (int k = 0; k < 16; k++) out[k] = float(o[k]);
}

// Mixture density of decoded outputs (40 per query) at a direction; pi/kappa/mu optional outputs.
double orc_guide_mixture_pdf(const float* out40, const float* dir, float* pi, float* kappa, float* mu) {
    double o[WIDTH] = {0};
    for (int i = 0; i < OUTPUTS; i++) o[i] = out40[i];
    Lobes<double> L;
    decode(o, L);
    if (pi)
        for (int j = 0; j < LOBES; j++) {
            pi[j] = float(L.pi[j]);
            kappa[j] = float(L.kappa[j]);
            for (int a = 0; a < 3; a++) mu[3 * j + a] = float(L.mu[j][a]);
        }
    const double d[3] = {dir[0], dir[1], dir[2]};
    return mixture_pdf(L, d);
}

double orc_vmf_pdf(double kappa, double cos_mu) { return vmf<double>(kappa, cos_mu); }

} // extern "C"
