// CPU oracle — TEST INFRASTRUCTURE ONLY. Included at the end of oracle.cpp.
//
// Restatement of the MCPG renderer's surface transport (render_pt_mcpg.slang:39-244) with
// the Slang library it imports: the hash grid (utils/hash-grid.slang), the Markov-chain state
// (render_pt_mcpg/mc.slang), von Mises-Fisher lobes (distributions/von-mises-fisher.slang),
// the light cache (light-cache/hashed-irradiance-cache.slang), grid helpers
// (utils/grid.slang:37-72, utils/envmap.slang:7-31, utils/morton.slang:24-32) and
// log1p / expm1 / x_over_expm1 (utils/math.slang:31-70). Pixels and samples run one after
// another in pixel order, as one GPU thread per pixel would if the threads ran in turn.
//
// Conventions where the reference leaves a choice (stated so the product can follow them):
// natural log = log2(x) ln 2 of the oracle's log2; pow(2, x) = exp2(x); round() ties to even;
// float -> int saturates (NaN -> 0); mad() is unfused; static scenes (zero motion vectors,
// prev_pos == pos, scene time 0).

namespace orc {

constexpr uint32_t HG_LOCK = 0x80000000u;
constexpr uint32_t HG_INVALID = 0xFFFFFFFFu;
constexpr float HG_LEVEL_OFFSET = 32.0f;

inline float o_ln(float x) { return o_log2(x) * 0.6931471805599453f; }
inline float o_log1p(float x) {
    const float u = x + 1.0f;
    if (u == 1.0f) return x;
    const float y = o_ln(u);
    if (x < 1.0f) return (x * y) / (u - 1.0f);
    return y;
}
inline float o_expm1(float x) {
    const float u = o_exp(x);
    if (u == 1.0f) return x;
    const float y = u - 1.0f;
    if (std::fabs(x) < 1.0f) return (x * y) / o_ln(u);
    return y;
}
inline float o_x_over_expm1(float x) {
    const float u = o_exp(x);
    if (u == 1.0f) return 1.0f;
    const float y = u - 1.0f;
    if (std::fabs(x) < 1.0f) return o_ln(u) / y;
    return x / y;
}
inline int32_t o_f2i(float f) {
    if (std::isnan(f)) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f < -2147483648.0f) return INT32_MIN;
    return int32_t(f);
}

struct GridCfg {
    uint32_t size, probes, locality;
    bool stochastic, level_stochastic, lock_free;
    float tan_half, bias, dim;
};

struct GPos {
    uint32_t slot = HG_INVALID, key = 0;
    bool valid() const { return slot != HG_INVALID; }
};

struct I3v {
    int32_t x, y, z;
};

// HashGrid<T, ...> (hash-grid.slang:44-532): split storage (keys, payloads), GRID_POWER 2,
// one step per unit size.
template <class T> struct HashGrid {
    GridCfg c{};
    std::vector<uint32_t> keys; // 2 per slot: key, stamp | lock
    std::vector<T> data;

    void reset() {
        keys.assign(2 * size_t(c.size), 0u);
        data.assign(c.size, T{});
    }
    uint32_t target_level(f3 pos, f3 cam, PCG& rng) const {
        const f3 d = cam - pos;
        const float dist2 = std::fmax(dot(d, d), 1e-12f);
        const float level = std::fmax(1.0f * ((0.5f * o_ln(dist2)) / o_ln(2.0f) + HG_LEVEL_OFFSET) + c.bias, 1.0f);
        if (c.level_stochastic) {
            const float fl = std::floor(level);
            const float frac = level - fl;
            return uint32_t(fl + (rng.next_float() < frac ? 1.0f : 0.0f));
        }
        return uint32_t(std::floor(level));
    }
    uint32_t distributed_level(f3 pos, f3 cam, PCG& rng) const {
        const float t = float(target_level(pos, cam, rng));
        const float u = rng.next_float();
        const float level = t + ((-1.0f / c.dim) * o_ln(1.0f - u)) / o_ln(2.0f);
        return uint32_t(std::nearbyint(level)); // default rounding mode: ties to even
    }
    float inv_width(uint32_t level) const {
        return (0.5f / c.tan_half) * o_exp2(HG_LEVEL_OFFSET - (float(level) - c.bias) / 1.0f);
    }
    static I3v interpolate(f3 g, float r) {
        const I3v lo{o_f2i(std::floor(g.x)), o_f2i(std::floor(g.y)), o_f2i(std::floor(g.z))};
        const float fx = g.x - float(lo.x), fy = g.y - float(lo.y), fz = g.z - float(lo.z);
        // the eight cell vertices in the reference's order, each with its trilinear weight
        const int off[8][3] = {{1, 1, 1}, {1, 1, 0}, {1, 0, 1}, {1, 0, 0}, {0, 1, 1}, {0, 1, 0}, {0, 0, 1}, {0, 0, 0}};
        float acc = 0.0f;
        for (int k = 0; k < 7; k++) {
            const float wx = off[k][0] ? fx : 1.0f - fx;
            const float wy = off[k][1] ? fy : 1.0f - fy;
            const float wz = off[k][2] ? fz : 1.0f - fz;
            acc += (wx * wy) * wz;
            if (r <= acc) return {lo.x + off[k][0], lo.y + off[k][1], lo.z + off[k][2]};
        }
        return lo;
    }
    I3v grid_idx(f3 pos, float iw, PCG& rng) const {
        const f3 g{pos.x * iw + 0.5f, pos.y * iw + 0.5f, pos.z * iw + 0.5f};
        if (c.stochastic) return interpolate(g, rng.next_float());
        return {o_f2i(std::floor(g.x)), o_f2i(std::floor(g.y)), o_f2i(std::floor(g.z))};
    }
    static uint32_t cube_face(f3 w) {
        const float ax = std::fabs(w.x), ay = std::fabs(w.y), az = std::fabs(w.z);
        if (ax > ay && ax > az) return w.x < 0.0f ? 1u : 0u;
        if (ay > az) return w.y < 0.0f ? 3u : 2u;
        return w.z < 0.0f ? 5u : 4u;
    }
    static uint32_t normal_bin(f3 n, uint32_t level) {
        if ((level & 1u) == 0u) return cube_face(n);
        const float r2 = 1.41421356237309504880f;
        const float s = n.x + n.y, t = r2 * n.z;
        return cube_face(f3{r2 * (n.x - n.y), s - t, s + t});
    }
    static uint32_t spread_bits(uint32_t v) { // morton_encode3d per axis
        v &= 0x3FFu;
        v = (v ^ (v << 16)) & 0x030000FFu;
        v = (v ^ (v << 8)) & 0x0300F00Fu;
        v = (v ^ (v << 4)) & 0x030C30C3u;
        v = (v ^ (v << 2)) & 0x09249249u;
        return v;
    }
    GPos cell(I3v gi, uint32_t level, f3 n) const {
        const uint32_t word = (level << 3) | normal_bin(n, level);
        const uint32_t tiles = c.size >> (3u * c.locality);
        GPos p;
        if (c.locality == 0u || tiles == 0u) {
            p.key = pcg4d16(uint32_t(gi.x), uint32_t(gi.y), uint32_t(gi.z), word);
            p.slot = p.key % c.size;
        } else {
            const int L = int(c.locality);
            const uint32_t m = (1u << c.locality) - 1u;
            const uint32_t th = pcg4d16(uint32_t(gi.x >> L), uint32_t(gi.y >> L), uint32_t(gi.z >> L), word);
            const uint32_t local = spread_bits(uint32_t(gi.x) & m) | (spread_bits(uint32_t(gi.y) & m) << 1) |
                                   (spread_bits(uint32_t(gi.z) & m) << 2);
            p.key = th ^ (local * 0x9E3779B9u);
            p.slot = ((th % tiles) << (3u * c.locality)) + local;
        }
        if ((c.probes & (c.probes - 1u)) == 0u) p.slot &= ~(c.probes - 1u);
        return p;
    }
    uint32_t wrap(uint32_t base, uint32_t i) const {
        uint32_t s = base + i;
        if (s >= c.size) s -= c.size;
        return s;
    }
    GPos find(GPos p) const {
        for (uint32_t i = 0; i < c.probes; i++) {
            const uint32_t s = wrap(p.slot, i);
            if (keys[2 * size_t(s)] == p.key) return GPos{s, p.key};
        }
        return GPos{};
    }
    GPos find_level(uint32_t level, f3 pos, f3 n, PCG& rng) const {
        return find(cell(grid_idx(pos, inv_width(level), rng), level, n));
    }
    // slot a writer takes: the key's slot, else the oldest stamp (ties -> later slot)
    uint32_t victim(GPos p, uint32_t frame, bool& matched) const {
        uint32_t target = p.slot, max_age = 0;
        matched = false;
        for (uint32_t i = 0; i < c.probes; i++) {
            const uint32_t s = wrap(p.slot, i);
            if (keys[2 * size_t(s)] == p.key) {
                matched = true;
                return s;
            }
            const uint32_t age = frame - (keys[2 * size_t(s) + 1] & ~HG_LOCK);
            if (age >= max_age) {
                max_age = age;
                target = s;
            }
        }
        return target;
    }
    bool lock(uint32_t s, uint32_t frame) {
        if (c.lock_free) return true;
        const uint32_t locked = frame | HG_LOCK;
        const uint32_t old = keys[2 * size_t(s) + 1];
        keys[2 * size_t(s) + 1] = locked;
        return old != locked;
    }
    void store(uint32_t s, uint32_t key, uint32_t frame, const T& t) {
        data[s] = t;
        keys[2 * size_t(s)] = key;
        keys[2 * size_t(s) + 1] = frame & ~HG_LOCK;
    }
};

// MCState (mc.slang:15-128)
struct MCState {
    f3 w_tgt{0, 0, 0};
    float sum_w = 0.f, w_cos = 0.f;
    uint16_t mv[3] = {0, 0, 0};
    uint16_t N = 0;
    float T = 0.f;

    f3 target() const { return w_tgt / sum_w; }
    f3 dir(f3 from) const { return normalize(target() - from); }
    float prior(f3 pos, float dir_prior) const {
        const f3 d = target() - pos;
        return std::fmax(0.0001f, dir_prior / dot(d, d));
    }
    float mean_cos(f3 pos, float dir_prior) const {
        const uint32_t n = N;
        const float n2 = float(n * n);
        const float r = ((n2 * w_cos) / sum_w) / (n2 + prior(pos, dir_prior));
        return std::fmin(std::fmax(r, 0.0f), 0.9999999f);
    }
    float kappa(f3 pos, float dir_prior) const {
        const float r = mean_cos(pos, dir_prior);
        return (3.0f * r - r * r * r) / (1.0f - r * r);
    }
    void add_sample(f3 pos, float w, f3 tgt, float time) {
        N = uint16_t(std::min<uint32_t>(uint32_t(N) + 1u, 1024u));
        const float a = std::fmax(1.0f / float(N), 0.01f);
        sum_w = lerp(sum_w, w, a);
        w_tgt = lerp(w_tgt, w * tgt, a);
        const float c = w * std::fmax(0.0f, dot(normalize(tgt - pos), dir(pos)));
        w_cos = std::fmin(std::fmax(lerp(w_cos, c, a), 0.0f), sum_w);
        mv[0] = mv[1] = mv[2] = 0; // zero motion vector (static scene)
        T = time;
    }
};

struct Vmf {
    float kappa = 0.f;
    f3 mu{0, 0, 0};
    // vmf_pdf_tokuyoshi
    float pdf(f3 w) const {
        const f3 d = w - mu;
        return (o_exp((-0.5f * kappa) * dot(d, d)) * o_x_over_expm1(-2.0f * kappa)) / 12.566370614359172f;
    }
    // vmf_sample_tokuyosh(mu, kappa, next_float2)
    f3 sample(PCG& rng) const {
        const f2 u = rng.next_float2();
        const float phi = 6.283185307179586f * u.x;
        const float r =
            kappa > (1.192092896e-07f / 4.0f) ? o_log1p(u.y * o_expm1(-2.0f * kappa)) / kappa : -2.0f * u.y;
        const float st = std::sqrt(-(r * r + 2.0f * r));
        float s, c;
        o_sincos(phi, s, c);
        return frame_from_normal(mu).to_world(f3{c * st, s * st, 1.0f + r});
    }
};

inline Vmf state_vmf(const MCState& s, f3 pos, float dir_prior) {
    const f3 d = s.dir(pos);
    const float k = s.kappa(pos, dir_prior);
    if (std::isnan(k) || std::isnan(d.x) || std::isnan(d.y) || std::isnan(d.z)) return Vmf{};
    return Vmf{k, d};
}

struct LcData {
    uint16_t irr[3] = {0, 0, 0};
    uint16_t N = 0;
    f3 value() const { return f3{f16_to_f32(irr[0]), f16_to_f32(irr[1]), f16_to_f32(irr[2])}; }
};

inline void lc_accumulate(LcData& d, f3 irr) {
    if (std::isnan(irr.x) || std::isnan(irr.y) || std::isnan(irr.z)) return;
    const f3 smp{clampf(irr.x, 0.0f, 65504.0f), clampf(irr.y, 0.0f, 65504.0f), clampf(irr.z, 0.0f, 65504.0f)};
    f3 mean = d.value();
    if (!finite3(mean)) {
        mean = smp;
        d.N = 0;
    }
    d.N = uint16_t(std::min<uint32_t>(uint32_t(d.N) + 1u, 256u));
    const f3 v = lerp(mean, smp, 1.0f / float(d.N));
    d.irr[0] = f32_to_f16(v.x);
    d.irr[1] = f32_to_f16(v.y);
    d.irr[2] = f32_to_f16(v.z);
}

struct McpgCfg {
    int samples;
    float p_guiding;
    bool roughness;
    float alpha_threshold, prior;
    bool missing_light, lc_tail;
    float lc_min_pdf;
    bool bsdf_support_only;
};

struct McpgState {
    HashGrid<MCState> mc;
    HashGrid<LcData> lc;
    McpgCfg cfg{};
    bool has_prev = false;
    f3 prev_cam{0, 0, 0};

    // MCPG::load with its grid position (mc.slang:156-173)
    MCState mc_load(f3 pos, f3 n, f3 cam, float time, PCG& rng, GPos& at) const {
        const uint32_t level = mc.distributed_level(pos, cam, rng);
        at = mc.find_level(level, pos, n, rng);
        MCState s = at.valid() ? mc.data[at.slot] : MCState{};
        if (s.sum_w > 0.0f) {
            const f3 mvv{f16_to_f32(s.mv[0]), f16_to_f32(s.mv[1]), f16_to_f32(s.mv[2])};
            s.w_tgt = s.w_tgt + (s.sum_w * (time - s.T)) * mvv;
        } else {
            s.sum_w = 0.0f;
        }
        return s;
    }
    // MCPG::store = distributed_set (hash-grid.slang:497-505, 438-471)
    void mc_store(const MCState& s, f3 pos, f3 n, f3 cam, uint32_t frame, PCG& rng) {
        const uint32_t level = mc.distributed_level(pos, cam, rng);
        const GPos p = mc.cell(mc.grid_idx(pos, mc.inv_width(level), rng), level, n);
        bool matched;
        const uint32_t s_ = mc.victim(p, frame, matched);
        if (!mc.lock(s_, frame)) return;
        mc.store(s_, p.key, frame, s);
    }
    // HashedIrradianceCache::get with prev camera (hash-grid.slang:246-270)
    f3 lc_get(f3 pos, f3 n, f3 cam, f3 pcam, PCG& rng, GPos& found) const {
        const uint32_t level = lc.target_level(pos, cam, rng);
        found = lc.find_level(level, pos, n, rng);
        auto sanitized = [&](uint32_t s) {
            const f3 v = lc.data[s].value();
            return finite3(v) ? v : f3{0, 0, 0};
        };
        if (found.valid()) return sanitized(found.slot);
        if (cam.x != pcam.x || cam.y != pcam.y || cam.z != pcam.z) {
            const uint32_t pl = lc.target_level(pos, pcam, rng);
            if (level != pl) {
                const GPos q = lc.find_level(pl, pos, n, rng);
                if (q.valid()) return sanitized(q.slot);
            }
        }
        return f3{0, 0, 0};
    }
    // HashedIrradianceCache::update (hashed-irradiance-cache.slang:64-80)
    void lc_update(f3 pos, f3 n, f3 irr, f3 cam, f3 pcam, uint32_t frame, PCG& rng, GPos at) {
        if (at.valid()) {
            if (!lc.lock(at.slot, frame)) return;
            LcData d = lc.data[at.slot];
            lc_accumulate(d, irr);
            lc.store(at.slot, at.key, frame, d);
            return;
        }
        const uint32_t level = lc.target_level(pos, cam, rng);
        const GPos p = lc.cell(lc.grid_idx(pos, lc.inv_width(level), rng), level, n);
        bool matched;
        const uint32_t s = lc.victim(p, frame, matched);
        if (!lc.lock(s, frame)) return;
        LcData d = lc.data[s];
        if (!matched) {
            const uint32_t pl = lc.target_level(pos, pcam, rng);
            if (level != pl) {
                const GPos q = lc.find_level(pl, pos, n, rng);
                if (q.valid()) d = lc.data[q.slot];
            }
        }
        lc_accumulate(d, irr);
        lc.store(s, p.key, frame, d);
    }
};

// trace_path (render_pt_mcpg.slang:39-182), vacuum everywhere (transmittance 1)
inline f3 mcpg_trace_path(const Scene& sc, McpgState& M, const mrt_pt_params& prm, Ray ray, ShadingPoint sp, f3 cam,
                          f3 pcam, uint32_t frame, PCG& rng, uint64_t& segs) {
    const McpgCfg& k = M.cfg;
    const int L = prm.max_path_length;
    const float time = 0.0f;
    float current_ior = 1.0f;
    f3 throughput{1.0f, 1.0f, 1.0f};
    f3 radiance = prm.emission_on_primary ? throughput * sp.emission : f3{0, 0, 0};
    GPos lc_grid_pos;
    for (int bounce = 1; bounce < L; bounce++) {
        if (!sp.has_material || (throughput.x < SCENE_THROUGHPUT_THRESHOLD && throughput.y < SCENE_THROUGHPUT_THRESHOLD &&
                                 throughput.z < SCENE_THROUGHPUT_THRESHOLD))
            break;
        const GPos cur_lc_grid_pos = lc_grid_pos;
        lc_grid_pos = GPos{};
        const f3 cur_pos = sp.pos;
        const Frame frame_ = sp.shading_frame;
        const f3 normal = frame_.r2;
        const MaterialSample& bsdf_m = sp.mat;
        const f3 wi_local = frame_.to_local(-ray.d);

        MCState selected;
        GPos selected_pos;
        float score_sum = 0.0f;
        std::vector<float> scores(size_t(std::max(k.samples, 1)), 0.0f);
        std::vector<Vmf> vmfs(size_t(std::max(k.samples, 1)));
        for (int i = 0; i < k.samples; i++) {
            GPos gp;
            const MCState s = M.mc_load(cur_pos, normal, cam, time, rng, gp);
            score_sum += s.sum_w;
            if (rng.next_float() < s.sum_w / score_sum) {
                selected = s;
                selected_pos = gp;
                vmfs[size_t(i)] = vmfs[0];
                scores[size_t(i)] = scores[0];
                scores[0] = s.sum_w;
                vmfs[0] = state_vmf(s, cur_pos, k.prior);
            } else {
                scores[size_t(i)] = s.sum_w;
                vmfs[size_t(i)] = state_vmf(s, cur_pos, k.prior);
            }
        }
        const float bsdf_alpha = bsdf_m.alpha();
        const bool can_guide = score_sum > 0.0f && bsdf_alpha >= k.alpha_threshold;
        const float p_guide = k.roughness ? k.p_guiding * bsdf_alpha : k.p_guiding;
        const bool guide = can_guide && rng.next_float() < p_guide;
        f3 wo_world, wo_local;
        if (guide) {
            wo_world = vmfs[0].sample(rng);
            wo_local = frame_.to_local(wo_world);
        } else {
            const Sample s = bsdf_m.sample_eval(wi_local, rng);
            if (!s.valid) break;
            wo_local = s.wo;
            wo_world = frame_.to_world(s.wo);
            selected = MCState{};
            selected_pos = GPos{};
            (void)M.mc.target_level(cur_pos, cam, rng);
        }
        float wo_p = 0.0f;
        if (can_guide) {
            for (int i = 0; i < k.samples; i++) wo_p += scores[size_t(i)] * vmfs[size_t(i)].pdf(wo_world);
            wo_p /= score_sum;
        }
        const float bsdf_w = can_guide ? (1.0f - p_guide) : 1.0f;
        const float pb = bsdf_m.pdf(wi_local, wo_local);
        wo_p = bsdf_w * pb + (can_guide ? p_guide : 0.0f) * wo_p;
        // bsdf_support_only (no counterpart in the reference): directions outside the BSDF
        // sampler's support contribute nothing
        const f3 bsdf = k.bsdf_support_only && !(pb > 0.0f) ? f3{0, 0, 0} : bsdf_m.eval(wi_local, wo_local);
        const bool transmission = wi_local.z * wo_local.z < 0.f;
        Ray next;
        bool has_next = false;
        if (bsdf.x > 0.0f || bsdf.y > 0.0f || bsdf.z > 0.0f) has_next = make_ray(sp, -ray.d, wo_world, transmission, next);

        f3 incident{0, 0, 0}, lc_incident{0, 0, 0};
        f3 next_pos = cur_pos;
        if (has_next) {
            if (transmission) current_ior /= bsdf_m.eta;
            ray = next;
            const HitRec hr = sc.trace(ray, prm.instance_mask);
            segs++;
            sp = get_shading_point(sc, ray, to_hit(sc, hr), uint32_t(bounce), current_ior, bounce + 1 < L);
            next_pos = sp.pos;
            incident = sp.emission;
            const f3 cached = M.lc_get(next_pos, sp.shading_frame.r2, cam, pcam, rng, lc_grid_pos);
            lc_incident = (incident.x > 0.0f || incident.y > 0.0f || incident.z > 0.0f || (!k.lc_tail && L == 2))
                              ? incident
                              : cached;
            throughput = throughput * (bsdf / wo_p);
            radiance = radiance + throughput * ((k.lc_tail && bounce == L - 1) ? lc_incident : incident);
        }
        const float mc_f = yuv_luminance((lc_incident * bsdf) / wo_p);
        if (!std::isinf(mc_f) && !std::isnan(mc_f)) {
            M.lc_update(cur_pos, normal, (lc_incident * bsdf) / std::fmax(wo_p, k.lc_min_pdf), cam, pcam, frame, rng,
                        cur_lc_grid_pos);
            if (rng.next_float() * score_sum < mc_f * float(k.samples)) {
                selected.add_sample(cur_pos, mc_f, next_pos, time);
                M.mc_store(selected, cur_pos, normal, cam, frame, rng);
            } else if (k.missing_light && selected_pos.valid()) {
                // light_missing_heuristic (mc.slang:53-70)
                bool kill = !(mc_f > 1e-3f * selected.sum_w);
                if (kill) {
                    const float c = dot(wo_world, selected.dir(cur_pos));
                    kill = !(c < 0.9f + 0.1f * selected.mean_cos(cur_pos, k.prior));
                }
                if (kill) M.mc.data[selected_pos.slot].sum_w = 0.0f;
            }
        }
        if (!has_next) break;
    }
    return radiance;
}

} // namespace orc

extern "C" {

// grid: size, probes, locality, stochastic (light cache only), tan(alpha/2), bias, dimension
struct orc_mcpg_grid {
    uint32_t size, probes, locality, stochastic;
    float tan_half, bias, dim;
    uint32_t _pad;
};
struct orc_mcpg_cfg {
    orc_mcpg_grid mc, lc;
    int32_t samples;
    float p_guiding;
    int32_t roughness;
    float alpha_threshold, prior;
    int32_t missing_light, lc_tail;
    float lc_min_pdf;
    int32_t bsdf_support_only;
};

__attribute__((visibility("default"))) void* orc_mcpg_create(const orc_mcpg_cfg* k) {
    McpgState* m = new McpgState();
    m->mc.c = GridCfg{k->mc.size, k->mc.probes, k->mc.locality, true, true, true, k->mc.tan_half, k->mc.bias, k->mc.dim};
    m->lc.c = GridCfg{k->lc.size,       k->lc.probes,   k->lc.locality, k->lc.stochastic != 0, false,
                      false,            k->lc.tan_half, k->lc.bias,     k->lc.dim};
    m->cfg = McpgCfg{k->samples,       k->p_guiding, k->roughness != 0, k->alpha_threshold,
                     k->prior,         k->missing_light != 0, k->lc_tail != 0, k->lc_min_pdf,
                     k->bsdf_support_only != 0};
    m->mc.reset();
    m->lc.reset();
    return m;
}
__attribute__((visibility("default"))) void orc_mcpg_destroy(void* m) { delete static_cast<McpgState*>(m); }

// Grid contents in the product's byte layout: keys (2 u32 per slot), chain states 32 B,
// light-cache entries 8 B.
__attribute__((visibility("default"))) void orc_mcpg_grids(void* mp, uint32_t* mc_keys, float* mc_states,
                                                           uint32_t* lc_keys, uint16_t* lc_entries) {
    const McpgState& m = *static_cast<McpgState*>(mp);
    for (size_t s = 0; s < m.mc.data.size(); s++) {
        if (mc_keys) {
            mc_keys[2 * s] = m.mc.keys[2 * s];
            mc_keys[2 * s + 1] = m.mc.keys[2 * s + 1];
        }
        if (mc_states) {
            const MCState& x = m.mc.data[s];
            float* o = mc_states + 8 * s;
            o[0] = x.w_tgt.x;
            o[1] = x.w_tgt.y;
            o[2] = x.w_tgt.z;
            o[3] = x.sum_w;
            o[4] = x.w_cos;
            const uint32_t w5 = uint32_t(x.mv[0]) | (uint32_t(x.mv[1]) << 16);
            const uint32_t w6 = uint32_t(x.mv[2]) | (uint32_t(x.N) << 16);
            std::memcpy(o + 5, &w5, 4);
            std::memcpy(o + 6, &w6, 4);
            o[7] = x.T;
        }
    }
    for (size_t s = 0; s < m.lc.data.size(); s++) {
        if (lc_keys) {
            lc_keys[2 * s] = m.lc.keys[2 * s];
            lc_keys[2 * s + 1] = m.lc.keys[2 * s + 1];
        }
        if (lc_entries) {
            for (int j = 0; j < 3; j++) lc_entries[4 * s + j] = m.lc.data[s].irr[j];
            lc_entries[4 * s + 3] = m.lc.data[s].N;
        }
    }
}

// One MCPG frame (render_pixel, render_pt_mcpg.slang:219-248) over the whole w x h image, one
// pixel after another; out: float4 per pixel. Returns the continuation segments traced.
__attribute__((visibility("default"))) uint64_t orc_mcpg_render(void* scp, void* mp, const mrt_pt_params* prm,
                                                                const mrt_camera* cam, uint32_t frame, uint32_t w,
                                                                uint32_t h, const mrt_packed_hit* gbuffer, float* out) {
    const Scene& sc = *static_cast<Scene*>(scp);
    McpgState& M = *static_cast<McpgState*>(mp);
    const f3 cpos = ld3(cam->position);
    const f3 pcam = M.has_prev ? M.prev_cam : cpos;
    uint64_t segs = 0;
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) {
            const size_t i = size_t(y) * w + x;
            PCG rng{pcg3d16_(x, y, frame)};
            const Ray primary_ray = camera_ray(*cam, x, y, w, h);
            const ShadingPoint primary_sp = get_shading_point(sc, primary_ray, unpack_hit(gbuffer[i]), 0u, 1.0f, true);
            const f3 demod =
                prm->demodulate_albedo ? demodulation_albedo(primary_ray, primary_sp) : f3{1.f, 1.f, 1.f};
            f3 irr{0, 0, 0};
            float m2 = 0.f;
            for (int s = 0; s < prm->spp; s++) {
                const f3 contrib =
                    mcpg_trace_path(sc, M, *prm, primary_ray, primary_sp, cpos, pcam, frame, rng, segs) / demod;
                if (finite3(contrib)) {
                    irr = irr + contrib;
                    const float l = yuv_luminance(contrib);
                    m2 += l * l;
                }
            }
            if (prm->spp > 0) {
                irr = irr / float(prm->spp);
                m2 /= float(prm->spp);
            }
            out[4 * i + 0] = irr.x;
            out[4 * i + 1] = irr.y;
            out[4 * i + 2] = irr.z;
            out[4 * i + 3] = m2;
        }
    M.prev_cam = cpos;
    M.has_prev = true;
    return segs;
}

// Probes for the invariant tests. op 0: vMF pdf(w) for (kappa, mu); op 1: vMF sample from
// PCG(seed) per row; op 2: log1p; op 3: expm1; op 4: x / expm1(x).
__attribute__((visibility("default"))) void orc_mcpg_probe(int op, const float* a, uint64_t n, float* out,
                                                           uint32_t seed) {
    for (uint64_t i = 0; i < n; i++) {
        switch (op) {
        case 0: {
            const Vmf v{a[7 * i], f3{a[7 * i + 1], a[7 * i + 2], a[7 * i + 3]}};
            out[i] = v.pdf(f3{a[7 * i + 4], a[7 * i + 5], a[7 * i + 6]});
            break;
        }
        case 1: {
            const Vmf v{a[4 * i], f3{a[4 * i + 1], a[4 * i + 2], a[4 * i + 3]}};
            PCG rng{pcg4d16(uint32_t(i), seed, 7u, 11u)};
            const f3 d = v.sample(rng);
            out[3 * i] = d.x;
            out[3 * i + 1] = d.y;
            out[3 * i + 2] = d.z;
            break;
        }
        case 2: out[i] = o_log1p(a[i]); break;
        case 3: out[i] = o_expm1(a[i]); break;
        default: out[i] = o_x_over_expm1(a[i]); break;
        }
    }
}

// Hash-grid addressing probe: for each position (3 floats), the light-cache-style cell at a
// given level and normal: out = (slot, key) of get_hash_grid_pos.
__attribute__((visibility("default"))) void orc_mcpg_cells(const orc_mcpg_grid* g, const float* pos, const float* nrm,
                                                           const uint32_t* levels, uint64_t n, uint32_t* out) {
    HashGrid<LcData> G;
    G.c = GridCfg{g->size, g->probes, g->locality, false, false, false, g->tan_half, g->bias, g->dim};
    PCG rng{1u};
    for (uint64_t i = 0; i < n; i++) {
        const f3 p{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}, nn{nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]};
        const GPos c = G.cell(G.grid_idx(p, G.inv_width(levels[i]), rng), levels[i], nn);
        out[2 * i] = c.slot;
        out[2 * i + 1] = c.key;
    }
}

// MCState add_sample / vMF derivation probe (mc.slang:73-117): state[8] in/out as the
// product's 32-byte layout, one sample (pos, w, target) applied; kv = (kappa, mu).
__attribute__((visibility("default"))) void orc_mcpg_add_sample(float* state, const float* pos, float w,
                                                                const float* tgt, float prior, float* kv) {
    MCState s;
    s.w_tgt = f3{state[0], state[1], state[2]};
    s.sum_w = state[3];
    s.w_cos = state[4];
    uint32_t w6;
    std::memcpy(&w6, state + 6, 4);
    s.N = uint16_t(w6 >> 16);
    s.add_sample(ld3(pos), w, ld3(tgt), 0.0f);
    state[0] = s.w_tgt.x;
    state[1] = s.w_tgt.y;
    state[2] = s.w_tgt.z;
    state[3] = s.sum_w;
    state[4] = s.w_cos;
    const uint32_t n6 = uint32_t(s.N) << 16, z = 0;
    std::memcpy(state + 5, &z, 4);
    std::memcpy(state + 6, &n6, 4);
    state[7] = s.T;
    const Vmf v = state_vmf(s, ld3(pos), prior);
    kv[0] = v.kappa;
    kv[1] = v.mu.x;
    kv[2] = v.mu.y;
    kv[3] = v.mu.z;
}

} // extern "C"
