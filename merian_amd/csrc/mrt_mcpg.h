// Markov-chain path guiding (MCPG) on the device: the hash grid both caches live in
// (utils/hash-grid.slang), the per-cell chain state (render_pt_mcpg/mc.slang:15-128), its
// von Mises-Fisher lobe (distributions/von-mises-fisher.slang:29-45) and the light cache
// (light-cache/hashed-irradiance-cache.slang), as the MCPG renderer's surface transport uses
// them (render_pt_mcpg.slang:39-213).
//
// Grids live in HBM as two arrays per grid: 8-byte keys (cell hash, frame stamp | lock bit)
// and payloads (chain state 32 B, light-cache entry 8 B), so a probe reads only keys. Every
// function takes the path's PCG by reference and draws in the reference's order. The
// elementary functions are the deterministic ones of mrt_math.h; natural log is
// log2(x) ln 2, pow(2, x) is exp2(x), and log1p / expm1 / x / expm1(x) follow the published
// methods the reference cites (Goldberg 1991; Higham 2002).
#pragma once

#include "mrt_device.h"
#include "mrt_kernels.h"
#include "mrt_math.h"

namespace mrt {

#define MCPG_D __device__ __forceinline__

constexpr uint32_t GRID_LOCK_BIT = 0x80000000u;     // hash-grid.slang:10
constexpr float GRID_LEVEL_OFFSET = 32.0f;          // hash-grid.slang:12
constexpr uint32_t GRID_INVALID_SLOT = 0xFFFFFFFFu; // hash-grid.slang:13
constexpr uint32_t MC_MAX_N = 1024;                 // mc.slang:11
constexpr float MC_MIN_ALPHA = 0.01f;               // mc.slang:12
constexpr uint32_t LC_MAX_N = 256;                  // hashed-irradiance-cache.slang:6
constexpr float LC_MAX_VALUE = 65504.0f;            // float16_t.maxValue
constexpr float FOUR_PI_F = 12.566370614359172f;
constexpr float FLT_EPS_F = 1.192092896e-07f;
constexpr float SQRT2_F = 1.41421356237309504880f;
constexpr int MCPG_MAX_SAMPLES = 8; // RIS candidates per vertex the kernels hold in registers

// ---------------------------------------------------------------------------------------
// scalar helpers

MCPG_D float mlog(float x) { return mlog2(x) * LN2_F; }
// log(1 + x) (utils/math.slang:31-42, Goldberg 1991)
MCPG_D float mlog1p(float x) {
    const float u = x + 1.0f;
    if (u == 1.0f) return x;
    const float y = mlog(u);
    if (x < 1.0f) return (x * y) / (u - 1.0f);
    return y;
}
// exp(x) - 1 (utils/math.slang:44-56, Higham 2002)
MCPG_D float mexpm1(float x) {
    const float u = mexp(x);
    if (u == 1.0f) return x;
    const float y = u - 1.0f;
    if (fabsf(x) < 1.0f) return (x * y) / mlog(u);
    return y;
}
// x / (exp(x) - 1) (utils/math.slang:60-70)
MCPG_D float x_over_expm1(float x) {
    const float u = mexp(x);
    if (u == 1.0f) return 1.0f;
    const float y = u - 1.0f;
    if (fabsf(x) < 1.0f) return mlog(u) / y;
    return x / y;
}
// float -> int32 with a defined result out of range (NaN -> 0, saturating), identical on the
// host and the device.
MCPG_D int32_t to_i32(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f < -2147483648.0f) return -2147483647 - 1;
    return int32_t(f);
}

// ---------------------------------------------------------------------------------------
// hash grid (utils/hash-grid.slang:44-532), GRID_POWER 2, one step per unit size


struct GridPos {
    uint32_t slot, key;
    MCPG_D bool valid() const { return slot != GRID_INVALID_SLOT; }
};
MCPG_D GridPos invalid_grid_pos() { return GridPos{GRID_INVALID_SLOT, 0u}; }

struct I3 {
    int32_t x, y, z;
};

// target_level_for_pos (hash-grid.slang:103-114)
MCPG_D uint32_t grid_target_level(const HashGridDev& g, f3 pos, f3 cam, PCG& rng) {
    const f3 d = cam - pos;
    const float dist2 = fmaxf(dot(d, d), 1e-12f);
    float level = (0.5f * mlog(dist2)) / mlog(2.0f) + GRID_LEVEL_OFFSET;
    level = fmaxf(level + g.level_bias, 1.0f);
    const float fl = floorf(level);
    if (g.level_stochastic) return uint32_t(fl + ((rng.next_float() < level - fl) ? 1.0f : 0.0f));
    return uint32_t(fl);
}
// distributed_level_for_pos (hash-grid.slang:117-123): the target level plus an exponential
// spread toward coarser levels; round() resolves ties to even.
MCPG_D uint32_t grid_distributed_level(const HashGridDev& g, f3 pos, f3 cam, PCG& rng) {
    const uint32_t t = grid_target_level(g, pos, cam, rng);
    const float u = rng.next_float();
    const float spread = ((-1.0f / g.dist_dim) * mlog(1.0f - u)) / mlog(2.0f);
    return uint32_t(rintf(float(t) + spread));
}
// grid_inv_width_for_level (hash-grid.slang:126-131)
MCPG_D float grid_inv_width(const HashGridDev& g, uint32_t level) {
    return (0.5f / g.tan_alpha_half) * mexp2(GRID_LEVEL_OFFSET - (float(level) - g.level_bias));
}
// grid_idx_interpolate_scaled (utils/grid.slang:37-72): a vertex of the cell with the
// trilinear weights as probabilities
MCPG_D I3 grid_idx_interpolate(f3 p, float r) {
    const I3 lo{to_i32(floorf(p.x)), to_i32(floorf(p.y)), to_i32(floorf(p.z))};
    const f3 f{p.x - float(lo.x), p.y - float(lo.y), p.z - float(lo.z)};
    float sum = (f.x * f.y) * f.z;
    if (r <= sum) return {lo.x + 1, lo.y + 1, lo.z + 1};
    sum += (f.x * f.y) * (1.0f - f.z);
    if (r <= sum) return {lo.x + 1, lo.y + 1, lo.z};
    sum += (f.x * (1.0f - f.y)) * f.z;
    if (r <= sum) return {lo.x + 1, lo.y, lo.z + 1};
    sum += (f.x * (1.0f - f.y)) * (1.0f - f.z);
    if (r <= sum) return {lo.x + 1, lo.y, lo.z};
    sum += ((1.0f - f.x) * f.y) * f.z;
    if (r <= sum) return {lo.x, lo.y + 1, lo.z + 1};
    sum += ((1.0f - f.x) * f.y) * (1.0f - f.z);
    if (r <= sum) return {lo.x, lo.y + 1, lo.z};
    sum += ((1.0f - f.x) * (1.0f - f.y)) * f.z;
    if (r <= sum) return {lo.x, lo.y, lo.z + 1};
    return lo;
}
// grid_idx (hash-grid.slang:133-140): lattice staggered by half a cell
MCPG_D I3 grid_idx(const HashGridDev& g, f3 pos, float inv_width, PCG& rng) {
    const f3 s{pos.x * inv_width + 0.5f, pos.y * inv_width + 0.5f, pos.z * inv_width + 0.5f};
    if (g.stochastic) return grid_idx_interpolate(s, rng.next_float());
    return {to_i32(floorf(s.x)), to_i32(floorf(s.y)), to_i32(floorf(s.z))};
}
// cubemap_side / cubemap_side_45deg (utils/envmap.slang:7-31)
MCPG_D uint32_t cubemap_side(f3 w) {
    const f3 a{fabsf(w.x), fabsf(w.y), fabsf(w.z)};
    const bool xd = a.x > a.y && a.x > a.z, yd = a.y > a.z;
    return (xd ? 0u : yd ? 2u : 4u) + uint32_t((xd ? w.x : yd ? w.y : w.z) < 0.0f);
}
MCPG_D uint32_t cubemap_side_45deg(f3 w) {
    const float s = w.x + w.y;
    const float t = SQRT2_F * w.z;
    const f3 d{SQRT2_F * (w.x - w.y), s - t, s + t};
    return cubemap_side(d);
}
// level_normal_word (hash-grid.slang:143-151): level << 3 | normal bin (axis-aligned cube
// faces on even levels, 45-degree turned faces on odd levels)
MCPG_D uint32_t level_normal_word(uint32_t level, f3 n) {
    return (level << 3) | ((level & 1u) == 0u ? cubemap_side(n) : cubemap_side_45deg(n));
}
// morton_encode3d (utils/morton.slang:24-32)
MCPG_D uint32_t morton_part(uint32_t p) {
    p &= 0x000003FFu;
    p = (p ^ (p << 16)) & 0x030000FFu;
    p = (p ^ (p << 8)) & 0x0300F00Fu;
    p = (p ^ (p << 4)) & 0x030C30C3u;
    p = (p ^ (p << 2)) & 0x09249249u;
    return p;
}
MCPG_D uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return morton_part(x) | (morton_part(y) << 1) | (morton_part(z) << 2);
}
// get_hash_grid_pos (hash-grid.slang:154-180): the slot a cell starts probing at and its key
MCPG_D GridPos grid_cell(const HashGridDev& g, I3 gi, uint32_t ln) {
    const uint32_t tiles = g.locality >= 11u ? 0u : (g.size >> (3u * g.locality));
    uint32_t slot, key;
    if (g.locality == 0u || tiles == 0u) {
        key = pcg4d16(uint32_t(gi.x), uint32_t(gi.y), uint32_t(gi.z), ln);
        slot = key % g.size;
    } else {
        const uint32_t mask = (1u << g.locality) - 1u;
        const int32_t L = int32_t(g.locality);
        const uint32_t th = pcg4d16(uint32_t(gi.x >> L), uint32_t(gi.y >> L), uint32_t(gi.z >> L), ln);
        const uint32_t cell = morton3(uint32_t(gi.x) & mask, uint32_t(gi.y) & mask, uint32_t(gi.z) & mask);
        key = th ^ (cell * 0x9E3779B9u);
        slot = ((th % tiles) << (3u * g.locality)) + cell;
    }
    if ((g.probes & (g.probes - 1u)) == 0u) slot &= ~(g.probes - 1u);
    return GridPos{slot, key};
}
MCPG_D uint32_t next_slot(const HashGridDev& g, uint32_t s) {
    s++;
    return s >= g.size ? s - g.size : s;
}
// find (hash-grid.slang:191-204): linear probe for the key
MCPG_D GridPos grid_find(const HashGridDev& g, GridPos p) {
    uint32_t slot = p.slot;
    for (uint32_t i = 0; i < g.probes; i++) {
        if (g.keys[slot].x == p.key) return GridPos{slot, p.key};
        slot = next_slot(g, slot);
    }
    return invalid_grid_pos();
}
// find_level (hash-grid.slang:224-228)
MCPG_D GridPos grid_find_level(const HashGridDev& g, uint32_t level, f3 pos, f3 normal, PCG& rng) {
    const I3 gi = grid_idx(g, pos, grid_inv_width(g, level), rng);
    return grid_find(g, grid_cell(g, gi, level_normal_word(level, normal)));
}
// The slot a writer of `cell` takes (hash-grid.slang:326-345): the matching slot, else the
// least recently stamped of the probed slots (ties: the later one).
MCPG_D uint32_t grid_write_slot(const HashGridDev& g, GridPos cell, uint32_t frame, bool& matched) {
    uint32_t target = cell.slot, max_age = 0u, idx = cell.slot;
    matched = false;
    for (uint32_t i = 0; i < g.probes; i++) {
        if (i > 0) idx = next_slot(g, idx);
        const uint2 probe = g.keys[idx];
        if (probe.x == cell.key) {
            target = idx;
            matched = true;
            break;
        }
        const uint32_t age = frame - (probe.y & ~GRID_LOCK_BIT);
        if (age >= max_age) {
            max_age = age;
            target = idx;
        }
    }
    return target;
}

// ---------------------------------------------------------------------------------------
// Markov-chain state (mc.slang:15-128): 32 bytes per cell, laid out as the reference's
// scalar layout: w_tgt.xyz, sum_w | w_cos, mv.xy (half), mv.z (half) | N (u16), T

struct McState {
    f3 w_tgt;
    float sum_w;
    float w_cos;
    uint32_t mv_xy;   // half bits
    uint32_t mv_z_n;  // mv.z half bits | N << 16
    float T;
    MCPG_D uint32_t N() const { return mv_z_n >> 16; }
    MCPG_D f3 mv() const {
        return {half_to_float(uint16_t(mv_xy & 0xFFFFu)), half_to_float(uint16_t(mv_xy >> 16)),
                half_to_float(uint16_t(mv_z_n & 0xFFFFu))};
    }
    MCPG_D f3 target() const { return w_tgt / sum_w; }
    MCPG_D f3 dir(f3 from) const { return normalize(target() - from); }
    MCPG_D float prior(f3 pos, float dir_guide_prior) const {
        const f3 d = target() - pos;
        return fmaxf(0.0001f, dir_guide_prior / dot(d, d));
    }
    MCPG_D float mean_cos(f3 pos, float dir_guide_prior) const {
        const float nn = float(N() * N());
        return clampf(((nn * w_cos) / sum_w) / (nn + prior(pos, dir_guide_prior)), 0.0f, 0.9999999f);
    }
    MCPG_D float kappa(f3 pos, float dir_guide_prior) const {
        const float r = mean_cos(pos, dir_guide_prior);
        return (3.0f * r - (r * r) * r) / (1.0f - r * r);
    }
    MCPG_D bool light_missing(float mc_f, f3 wo, f3 pos, float dir_guide_prior) const {
        if (mc_f > 1e-3f * sum_w) return false;
        const float c = dot(wo, dir(pos));
        if (c < 0.9f + 0.1f * mean_cos(pos, dir_guide_prior)) return false;
        return true;
    }
    // add_sample (mc.slang:96-117): maximum-likelihood update with an exponentially
    // weighted average; the motion vector is stored as given
    MCPG_D void add_sample(f3 pos, float w, f3 tgt, uint32_t mv_xy_bits, uint16_t mv_z_bits, float time) {
        const uint32_t n = N() + 1u < MC_MAX_N ? N() + 1u : MC_MAX_N;
        const float a = fmaxf(1.0f / float(n), MC_MIN_ALPHA);
        sum_w = lerp(sum_w, w, a);
        w_tgt = lerp(w_tgt, w * tgt, a);
        w_cos = clampf(lerp(w_cos, w * fmaxf(0.0f, dot(normalize(tgt - pos), dir(pos))), a), 0.0f, sum_w);
        mv_xy = mv_xy_bits;
        mv_z_n = uint32_t(mv_z_bits) | (n << 16);
        T = time;
    }
};

MCPG_D McState mc_zero() { return McState{{0.f, 0.f, 0.f}, 0.f, 0.f, 0u, 0u, 0.f}; }

// von Mises-Fisher lobe (von-mises-fisher.slang:29-45, Tokuyoshi's stable forms)
struct Vmf {
    float kappa;
    f3 mu;
    MCPG_D float pdf(f3 w) const {
        const f3 d = w - mu;
        return (mexp((-0.5f * kappa) * dot(d, d)) * x_over_expm1(-2.0f * kappa)) / FOUR_PI_F;
    }
    MCPG_D f3 sample(PCG& rng) const {
        const f2 u = rng.next_float2();
        const float phi = TWO_PI * u.x;
        const float r = kappa > (FLT_EPS_F / 4.0f) ? mlog1p(u.y * mexpm1(-2.0f * kappa)) / kappa : -2.0f * u.y;
        const float st = sqrtf(-(r * r + 2.0f * r));
        float sp, cp;
        msincos(phi, sp, cp);
        return frame_from_normal(mu).to_world(f3{cp * st, sp * st, 1.0f + r});
    }
};
// get_vmf (mc.slang:73-82)
MCPG_D Vmf mc_vmf(const McState& s, f3 pos, float dir_guide_prior) {
    const f3 d = s.dir(pos);
    const float k = s.kappa(pos, dir_guide_prior);
    if (k != k || d.x != d.x || d.y != d.y || d.z != d.z) return Vmf{0.0f, {0.f, 0.f, 0.f}};
    return Vmf{k, d};
}

// ---------------------------------------------------------------------------------------
// MC grid (MCPG<...>, mc.slang:130-200): jittered cells and levels, lock-free writers

MCPG_D McState mc_slot_load(const HashGridDev& g, uint32_t slot) {
    const float4* d = static_cast<const float4*>(g.data) + 2 * size_t(slot);
    const float4 a = d[0], b = d[1];
    return McState{{a.x, a.y, a.z}, a.w, b.x, __float_as_uint(b.y), __float_as_uint(b.z), b.w};
}
MCPG_D void mc_slot_store(const HashGridDev& g, uint32_t slot, const McState& s) {
    float4* d = static_cast<float4*>(g.data) + 2 * size_t(slot);
    d[0] = make_float4(s.w_tgt.x, s.w_tgt.y, s.w_tgt.z, s.sum_w);
    d[1] = make_float4(s.w_cos, __uint_as_float(s.mv_xy), __uint_as_float(s.mv_z_n), s.T);
}
// load (mc.slang:156-173): a distributed lookup; an empty or killed cell reads sum_w = 0
MCPG_D McState mc_load(const HashGridDev& g, f3 pos, f3 normal, f3 cam, float time, PCG& rng, GridPos& pos_out) {
    const uint32_t level = grid_distributed_level(g, pos, cam, rng);
    pos_out = grid_find_level(g, level, pos, normal, rng);
    McState s = pos_out.valid() ? mc_slot_load(g, pos_out.slot) : mc_zero();
    if (s.sum_w > 0.0f) {
        s.w_tgt = s.w_tgt + (s.sum_w * (time - s.T)) * s.mv();
    } else {
        s.sum_w = 0.0f;
    }
    return s;
}
// store (mc.slang:175-182) = distributed_set (hash-grid.slang:497-505): replaces the cell
MCPG_D void mc_store(const HashGridDev& g, const McState& s, f3 pos, f3 normal, f3 cam, uint32_t frame, PCG& rng) {
    const uint32_t level = grid_distributed_level(g, pos, cam, rng);
    const I3 gi = grid_idx(g, pos, grid_inv_width(g, level), rng);
    const GridPos cell = grid_cell(g, gi, level_normal_word(level, normal));
    bool matched;
    const uint32_t target = grid_write_slot(g, cell, frame, matched);
    mc_slot_store(g, target, s); // payload before key (no fence, as the reference's shader)
    g.keys[target] = make_uint2(cell.key, frame & ~GRID_LOCK_BIT);
}
// kill (mc.slang:194-197): invalidates the cell the lobe was read from
MCPG_D void mc_kill(const HashGridDev& g, uint32_t slot) {
    static_cast<float*>(g.data)[8 * size_t(slot) + 3] = 0.0f;
}

// ---------------------------------------------------------------------------------------
// light cache (HashedIrradianceCache, hashed-irradiance-cache.slang:8-85): half3 mean +
// u16 count per cell, exact level, locked writers

struct LcEntry {
    uint32_t xy; // half bits
    uint32_t zn; // half bits | N << 16
};
MCPG_D f3 lc_irr(LcEntry e) {
    return {half_to_float(uint16_t(e.xy & 0xFFFFu)), half_to_float(uint16_t(e.xy >> 16)),
            half_to_float(uint16_t(e.zn & 0xFFFFu))};
}
MCPG_D LcEntry lc_slot_load(const HashGridDev& g, uint32_t slot) {
    const uint2 v = static_cast<const uint2*>(g.data)[slot];
    return LcEntry{v.x, v.y};
}
MCPG_D void lc_slot_store(const HashGridDev& g, uint32_t slot, LcEntry e) {
    static_cast<uint2*>(g.data)[slot] = make_uint2(e.xy, e.zn);
}
// hashed_irradiance_cache_sanitize: inf / NaN read as zero
MCPG_D f3 lc_sanitized(LcEntry e) {
    const f3 v = lc_irr(e);
    return finite3(v) ? v : f3{0.f, 0.f, 0.f};
}
// HashedIrradianceCacheUpdate::update (hashed-irradiance-cache.slang:19-36)
MCPG_D void lc_apply(LcEntry& e, f3 irr) {
    if (irr.x != irr.x || irr.y != irr.y || irr.z != irr.z) return;
    const f3 smp{clampf(irr.x, 0.0f, LC_MAX_VALUE), clampf(irr.y, 0.0f, LC_MAX_VALUE), clampf(irr.z, 0.0f, LC_MAX_VALUE)};
    f3 mean = lc_irr(e);
    uint32_t n = e.zn >> 16;
    if (!finite3(mean)) {
        mean = smp;
        n = 0u;
    }
    n = n + 1u < LC_MAX_N ? n + 1u : LC_MAX_N;
    const f3 v = lerp(mean, smp, 1.0f / float(n));
    e.xy = uint32_t(float_to_half(v.x)) | (uint32_t(float_to_half(v.y)) << 16);
    e.zn = uint32_t(float_to_half(v.z)) | (n << 16);
}
// get (hash-grid.slang:246-270 via hashed-irradiance-cache.slang:52-62): the cell at the
// camera's level; where the camera moved and that level changed, the previous camera's level
// is read instead (found then stays invalid, so an update does not write there)
MCPG_D f3 lc_get(const HashGridDev& g, f3 pos, f3 normal, f3 cam, f3 prev_cam, PCG& rng, GridPos& found) {
    const uint32_t level = grid_target_level(g, pos, cam, rng);
    found = grid_find_level(g, level, pos, normal, rng);
    if (found.valid()) return lc_sanitized(lc_slot_load(g, found.slot));
    if (cam.x != prev_cam.x || cam.y != prev_cam.y || cam.z != prev_cam.z) {
        const uint32_t prev_level = grid_target_level(g, pos, prev_cam, rng);
        if (level != prev_level) {
            const GridPos p = grid_find_level(g, prev_level, pos, normal, rng);
            if (p.valid()) return lc_sanitized(lc_slot_load(g, p.slot));
        }
    }
    return {0.f, 0.f, 0.f};
}
// A writer's lock (hash-grid.slang:348-356): the slot's stamp swapped to frame | lock bit;
// a writer that finds the slot locked in this frame skips its update.
MCPG_D bool lc_lock(const HashGridDev& g, uint32_t slot, uint32_t frame) {
    if (g.lock_free) return true;
    const uint32_t locked = frame | GRID_LOCK_BIT;
    const uint32_t old = atomicExch(reinterpret_cast<uint32_t*>(g.keys) + 2 * size_t(slot) + 1, locked);
    return old != locked;
}
// update (hashed-irradiance-cache.slang:64-80): at the cell the lookup found, else at the
// camera's level, seeding a new cell from the previous camera's level when that differs
MCPG_D void lc_update(const HashGridDev& g, f3 pos, f3 normal, f3 irr, f3 cam, f3 prev_cam, uint32_t frame, PCG& rng,
                      GridPos at) {
    if (at.valid()) { // update_at (hash-grid.slang:373-390)
        if (!lc_lock(g, at.slot, frame)) return;
        LcEntry e = lc_slot_load(g, at.slot);
        lc_apply(e, irr);
        lc_slot_store(g, at.slot, e);
        g.keys[at.slot] = make_uint2(at.key, frame & ~GRID_LOCK_BIT);
        return;
    }
    // _update_at_level<true> (hash-grid.slang:310-370)
    const uint32_t level = grid_target_level(g, pos, cam, rng);
    const I3 gi = grid_idx(g, pos, grid_inv_width(g, level), rng);
    const GridPos cell = grid_cell(g, gi, level_normal_word(level, normal));
    bool matched;
    const uint32_t target = grid_write_slot(g, cell, frame, matched);
    if (!lc_lock(g, target, frame)) return;
    LcEntry e = lc_slot_load(g, target);
    if (!matched) {
        const uint32_t prev_level = grid_target_level(g, pos, prev_cam, rng);
        if (level != prev_level) {
            const GridPos p = grid_find_level(g, prev_level, pos, normal, rng);
            if (p.valid()) e = lc_slot_load(g, p.slot);
        }
    }
    lc_apply(e, irr);
    lc_slot_store(g, target, e);
    g.keys[target] = make_uint2(cell.key, frame & ~GRID_LOCK_BIT);
}

#undef MCPG_D

} // namespace mrt
