// Node types of the graph host: "PBRT Scene", "GBuffer (Raytraced)", "Render (Path-traced)",
// "Render (Path-traced, MCPG)", "Accumulate" and "Image Write" (names and connectors of
// merian's node_registry.cpp:63-96).
#include <cmath>
#include <cstring>
#include <sys/stat.h>

#include "graph.h"

namespace mrt::graph {

namespace {

using json::Value;

// Properties with defaults: read a key (keeping the default when absent), record the result.
class Props {
  public:
    explicit Props(const Value& in) : in_(in.is_object() ? in : Value::object()), out_(Value::object()) {
        // keys this node does not interpret are kept as given (e.g. UI-only settings)
        for (const auto& [k, v] : in_.as_object()) out_[k] = v;
    }
    bool get_bool(const std::string& k, bool def) {
        const Value* v = in_.find(k);
        const bool r = v && v->is_bool() ? v->as_bool() : def;
        out_[k] = r;
        return r;
    }
    int64_t get_int(const std::string& k, int64_t def) {
        const Value* v = in_.find(k);
        const int64_t r = v && v->is_number() ? v->as_int() : def;
        out_[k] = Value(r);
        return r;
    }
    double get_float(const std::string& k, double def) {
        const Value* v = in_.find(k);
        const double r = v && v->is_number() ? v->as_double() : def;
        out_[k] = Value(r);
        return r;
    }
    std::string get_string(const std::string& k, const std::string& def) {
        const Value* v = in_.find(k);
        const std::string r = v && v->is_string() ? v->as_string() : def;
        out_[k] = r;
        return r;
    }
    Value get_object(const std::string& k, const Value& def) {
        const Value* v = in_.find(k);
        Value r = def;
        if (v && v->is_object()) GraphDescription::merge_into(r, *v);
        out_[k] = r;
        return r;
    }
    const Value& result() const { return out_; }

  private:
    Value in_, out_;
};

void check(mrt_ctx* c, mrt_status s, const char* what) {
    if (s != MRT_OK) throw node_error(std::string(what) + ": " + mrt_last_error(c));
}

// A context-owned device buffer that grows on demand.
struct DeviceBuffer {
    void* d = nullptr;
    uint64_t bytes = 0;
    void ensure(mrt_ctx* c, uint64_t need) {
        if (need <= bytes && d) return;
        release(c);
        check(c, mrt_alloc_dev(c, need, &d), "device allocation");
        bytes = need;
    }
    void release(mrt_ctx* c) {
        if (d) mrt_free_dev(c, d);
        d = nullptr;
        bytes = 0;
    }
};

uint32_t mask_bits(Props& p) {
    uint32_t m = 0;
    for (int k = 0; k < 8; k++)
        if (p.get_bool(std::to_string(k), true)) m |= 1u << k;
    return m;
}

// ------------------------------------------------------------------------------------------

// Loads a pbrt-v4 file (pbrt.h) into a device scene; "env" with Type "LatLong" and a PFM Path
// replaces the environment by a lat-long map (repeat), rotated by Yaw / Pitch in degrees.
class PbrtSceneNode : public Node {
  public:
    std::vector<std::string> inputs() const override { return {}; }
    std::vector<std::string> outputs() const override { return {"scene"}; }
    Value properties(const Value& in) override {
        Props p(in);
        file_ = p.get_string("file", "");
        Value env_def = Value::object();
        env_def["Type"] = "Empty";
        env_def["Path"] = "";
        env_def["Yaw"] = 0.0;
        env_def["Pitch"] = 0.0;
        env_ = p.get_object("env", env_def);
        p.get_object("settings", Value::object());
        dirty_ = true;
        return p.result();
    }
    NodeStatusFlags process(RunInfo& info, const std::map<std::string, const Resource*>&,
                            std::map<std::string, Resource>& out) override {
        if (dirty_) load(info);
        out["scene"] = res_;
        return 0;
    }
    void release(mrt_ctx* c) override {
        if (res_.scene) mrt_scene_destroy(c, res_.scene);
        res_.scene = nullptr;
    }

  private:
    void load(RunInfo& info) {
        if (file_.empty()) throw node_error("no scene file (property \"file\")");
        pbrt::Scene sc = pbrt::load(file_);
        for (const std::string& w : sc.warnings)
            if (info.log) info.log->push_back("PBRT Scene: " + w);
        mrt_scene_desc d = sc.desc();
        std::vector<float> latlong;
        const std::string type = env_.value("Type", "Empty").as_string();
        const std::string path = env_.value("Path", "").as_string();
        if (type == "LatLong" && !path.empty()) {
            uint32_t w = 0, h = 0;
            if (!pbrt::read_pfm(path, latlong, w, h)) throw node_error("cannot read environment map '" + path + "' (PFM)");
            d.env.type = MRT_ENV_LATLONG;
            d.env.wrap = MRT_WRAP_REPEAT;
            d.env.width = w;
            d.env.height = h;
            d.env.rgba = latlong.data();
            d.env.intensity = 1.f;
            const double yaw = env_.value("Yaw", 0.0).as_double() * M_PI / 180.0;
            const double pitch = env_.value("Pitch", 0.0).as_double() * M_PI / 180.0;
            const double cy = std::cos(yaw), sy = std::sin(yaw), cp = std::cos(pitch), sp = std::sin(pitch);
            // world -> map: yaw about +y, then pitch about +x
            const double R[9] = {cy, 0, -sy, sy * sp, cp, cy * sp, sy * cp, -sp, cy * cp};
            for (int k = 0; k < 9; k++) d.env.to_local[k] = float(R[k]);
        }
        if (res_.scene) mrt_scene_destroy(info.ctx, res_.scene);
        res_.scene = nullptr;
        check(info.ctx, mrt_scene_create(info.ctx, &d, &res_.scene), "scene upload");
        res_.camera = sc.camera;
        res_.version++;
        dirty_ = false;
    }
    std::string file_;
    Value env_;
    bool dirty_ = true;
    SceneResource res_;
};

// Primary rays + closest hit into the 16-byte GBuffer (gbuffer.cpp:103-146) and the albedo
// texture. The extent is the scene camera's resolution when the file has a camera
// (gbuffer.cpp:104-112), else width x height; pbrt cameras use the R2 jitter sequence.
class GBufferNode : public Node {
  public:
    std::vector<std::string> inputs() const override { return {"scene"}; }
    std::vector<std::string> outputs() const override { return {"gbuffer"}; }
    Value properties(const Value& in) override {
        Props p(in);
        mask_ = mask_bits(p);
        width_ = uint32_t(p.get_int("width", 1920));
        height_ = uint32_t(p.get_int("height", 1080));
        return p.result();
    }
    NodeStatusFlags process(RunInfo& info, const std::map<std::string, const Resource*>& in,
                            std::map<std::string, Resource>& out) override {
        auto it = in.find("scene");
        if (it == in.end() || !it->second || !std::holds_alternative<SceneResource>(*it->second))
            throw node_error("input \"scene\" is not connected");
        const SceneResource& s = std::get<SceneResource>(*it->second);
        GBufferResource g;
        g.scene = s.scene;
        const bool file_extent = s.camera.present && s.camera.has_resolution;
        g.width = file_extent ? s.camera.width : width_;
        g.height = file_extent ? s.camera.height : height_;
        float eye[3] = {3, 3, 3}, target[3] = {0, 0, 0}, up[3] = {0, 1, 0};
        float vfov = float(60.0 * M_PI / 180.0);
        if (s.camera.present) {
            for (int k = 0; k < 3; k++) {
                eye[k] = s.camera.eye[k];
                target[k] = s.camera.eye[k] + s.camera.forward[k];
                up[k] = s.camera.up[k];
            }
            vfov = s.camera.vfov;
        }
        const float aspect = s.camera.present && s.camera.aspect > 0.f ? s.camera.aspect : float(g.width) / float(g.height);
        check(info.ctx,
              mrt_camera_look_at(eye, target, up, vfov, aspect, 0.01f, 1000.f, MRT_JITTER_R2, uint32_t(info.iteration),
                                 &g.camera),
              "camera");
        buf_.ensure(info.ctx, uint64_t(g.width) * g.height * sizeof(mrt_packed_hit));
        albedo_.ensure(info.ctx, uint64_t(g.width) * g.height * 16);
        g.d = static_cast<mrt_packed_hit*>(buf_.d);
        g.albedo = static_cast<float*>(albedo_.d);
        check(info.ctx, mrt_gbuffer_dev(info.ctx, s.scene, &g.camera, g.width, g.height, mask_, g.d, nullptr), "gbuffer");
        check(info.ctx, mrt_gbuffer_albedo_dev(info.ctx, s.scene, &g.camera, g.width, g.height, g.d, g.albedo, nullptr),
              "gbuffer albedo");
        out["gbuffer"] = g;
        return 0;
    }
    void release(mrt_ctx* c) override {
        buf_.release(c);
        albedo_.release(c);
    }

  private:
    uint32_t mask_ = 0xFF, width_ = 1920, height_ = 1080;
    DeviceBuffer buf_, albedo_;
};

// "Render (Path-traced)" (render_pt.cpp:26-34,151-181): scene + gbuffer in, irradiance out.
class RenderPTNode : public Node {
  public:
    std::vector<std::string> inputs() const override { return {"scene", "gbuffer"}; }
    std::vector<std::string> outputs() const override { return {"irradiance"}; }
    Value properties(const Value& in) override {
        Props p(in);
        spp_ = int32_t(p.get_int("samples per pixel", 1));
        max_path_length_ = int32_t(p.get_int("max path length", 5));
        emission_on_primary_ = p.get_bool("emission on primary", true);
        p.get_bool("shader execution reordering", false); // accepted; wavefront queues already reorder
        demodulate_ = p.get_bool("demodulate albedo", false);
        mask_ = mask_bits(p);
        const std::string fmt = p.get_string("irradiance format", "R32G32B32A32Sfloat");
        if (fmt != "R32G32B32A32Sfloat")
            throw node_error("irradiance format '" + fmt + "' is not supported (R32G32B32A32Sfloat only)");
        return p.result();
    }
    NodeStatusFlags process(RunInfo& info, const std::map<std::string, const Resource*>& in,
                            std::map<std::string, Resource>& out) override {
        auto gi = in.find("gbuffer");
        if (gi == in.end() || !gi->second || !std::holds_alternative<GBufferResource>(*gi->second))
            throw node_error("input \"gbuffer\" is not connected");
        auto si = in.find("scene");
        if (si == in.end() || !si->second || !std::holds_alternative<SceneResource>(*si->second))
            throw node_error("input \"scene\" is not connected");
        const GBufferResource& g = std::get<GBufferResource>(*gi->second);
        const SceneResource& s = std::get<SceneResource>(*si->second);
        mrt_pt_params p{};
        p.spp = spp_;
        p.max_path_length = max_path_length_;
        p.emission_on_primary = emission_on_primary_;
        p.demodulate_albedo = demodulate_;
        p.instance_mask = mask_;
        p.mode = MRT_MODE_BSDF;
        // a changed path length invalidates accumulated history (render_pt.cpp:102-105)
        if (max_path_length_ != emitted_max_path_length_) {
            emitted_max_path_length_ = max_path_length_;
            if (info.send_event) info.send_event("bounces_changed");
        }
        ImageResource img;
        img.width = g.width;
        img.height = g.height;
        buf_.ensure(info.ctx, uint64_t(g.width) * g.height * 16);
        img.d = static_cast<float*>(buf_.d);
        check(info.ctx,
              mrt_pt_render_dev(info.ctx, s.scene, &p, &g.camera, uint32_t(info.iteration), g.width, g.height, g.d,
                                img.d, nullptr, nullptr, nullptr),
              "pt");
        out["irradiance"] = img;
        return 0;
    }
    void release(mrt_ctx* c) override { buf_.release(c); }

  private:
    int32_t spp_ = 1, max_path_length_ = 5;
    int32_t emitted_max_path_length_ = 5; // the default: a loaded length other than 5 is announced once
    bool emission_on_primary_ = true, demodulate_ = false;
    uint32_t mask_ = 0xFF;
    DeviceBuffer buf_;
};

// "Render (Path-traced, MCPG)" (render_pt_mcpg.cpp:204-258,480-650): scene + gbuffer in,
// irradiance out; the Markov-chain and light-cache grids persist across iterations and are
// cleared on iteration 0. The surface transport only: the volume pass, distance chains and
// debug outputs are accepted as properties and not run.
class RenderMCPGNode : public Node {
  public:
    std::vector<std::string> inputs() const override { return {"scene", "gbuffer"}; }
    std::vector<std::string> outputs() const override { return {"irradiance"}; }
    Value properties(const Value& in) override {
        Props p(in);
        p.get_bool("raygen executor", false);
        const bool reference_mode = p.get_bool("reference mode", false);
        emission_on_primary_ = p.get_bool("emission on primary", true);
        demodulate_ = p.get_bool("demodulate albedo", false);
        spp_ = int32_t(p.get_int("samples per pixel", 1));
        max_path_length_ = int32_t(p.get_int("max path length", 5));
        mrt_mcpg_config k;
        mrt_mcpg_default_config(&k);
        const float p_guiding = float(p.get_float("guiding prob", 0.5));
        k.p_guiding = reference_mode ? 0.0f : p_guiding;
        const Value* st = in.is_object() ? in.find("directional sampling type") : nullptr;
        if (st && st->is_string()) {
            if (st->as_string() == "MIS") k.sampling_type = MRT_MCPG_MIS;
            else if (st->as_string() == "MIS+Roughness") k.sampling_type = MRT_MCPG_MIS_ROUGHNESS;
            else throw node_error("unknown directional sampling type '" + st->as_string() + "'");
        } else if (st && st->is_number()) {
            k.sampling_type = int32_t(st->as_int());
        }
        const float rough = float(p.get_float("roughness threshold", std::sqrt(0.05)));
        k.guiding_alpha_threshold = rough * rough;
        for (const char* v : {"volume samples per pixel", "distance MC samples"}) p.get_int(v, 1);
        for (const char* v : {"volume guiding prob", "distance guiding prob", "distance MC base width",
                              "distance MC max width", "distance MC level spread", "volume forward project min z"})
            p.get_float(v, 0.0);
        for (const char* v : {"volume: use LC", "volume forward project"}) p.get_bool(v, true);
        const auto grid = [&](const Value& child, mrt_mcpg_grid_config& g) {
            if (const Value* v = child.find("grid tan(alpha/2)"); v && v->is_number()) g.tan_alpha_half = float(v->as_double());
            if (const Value* v = child.find("grid level bias"); v && v->is_number()) g.level_bias = float(v->as_double());
            if (const Value* v = child.find("grid distribution dimension"); v && v->is_number())
                g.distribution_dimension = float(v->as_double());
        };
        Value mc = p.get_object("mc", Value::object());
        if (const Value* v = mc.find("ML prior"); v && v->is_number()) k.dir_guide_prior = float(v->as_double());
        if (const Value* v = mc.find("MC samples"); v && v->is_number()) k.mc_samples = int32_t(v->as_int());
        if (const Value* v = mc.find("missing light heuristic"); v && v->is_bool()) k.missing_light_heuristic = v->as_bool();
        if (const Value* v = mc.find("adaptive grid buf size"); v && v->is_number()) k.mc.buffer_size = uint32_t(v->as_int());
        if (const Value* v = mc.find("MC probe count"); v && v->is_number()) k.mc.probe_count = uint32_t(v->as_int());
        if (const Value* v = mc.find("locality bits"); v && v->is_number()) k.mc.locality_bits = uint32_t(v->as_int());
        grid(mc, k.mc);
        Value lc = p.get_object("lc", Value::object());
        if (const Value* v = lc.find("surf: use LC"); v && v->is_bool()) k.use_light_cache_tail = v->as_bool();
        if (const Value* v = lc.find("LC buffer size"); v && v->is_number()) k.lc.buffer_size = uint32_t(v->as_int());
        if (const Value* v = lc.find("LC probe count"); v && v->is_number()) k.lc.probe_count = uint32_t(v->as_int());
        if (const Value* v = lc.find("LC stochastic interpolation"); v && v->is_bool())
            k.lc.stochastic_interpolation = v->as_bool();
        if (const Value* v = lc.find("LC locality bits"); v && v->is_number()) k.lc.locality_bits = uint32_t(v->as_int());
        if (const Value* v = lc.find("LC min pdf"); v && v->is_number()) k.lc_min_pdf = float(v->as_double());
        grid(lc, k.lc);
        mask_ = mask_bits(p);
        p.get_int("debug output selector", 0);
        const std::string fmt = p.get_string("irradiance format", "R32G32B32A32Sfloat");
        if (fmt != "R32G32B32A32Sfloat")
            throw node_error("irradiance format '" + fmt + "' is not supported (R32G32B32A32Sfloat only)");
        if (std::memcmp(&k, &cfg_, sizeof(k)) != 0) {
            cfg_ = k;
            recreate_ = true; // grid sizes or constants changed: the node rebuilds its state
        }
        return p.result();
    }
    NodeStatusFlags process(RunInfo& info, const std::map<std::string, const Resource*>& in,
                            std::map<std::string, Resource>& out) override {
        auto gi = in.find("gbuffer");
        if (gi == in.end() || !gi->second || !std::holds_alternative<GBufferResource>(*gi->second))
            throw node_error("input \"gbuffer\" is not connected");
        auto si = in.find("scene");
        if (si == in.end() || !si->second || !std::holds_alternative<SceneResource>(*si->second))
            throw node_error("input \"scene\" is not connected");
        const GBufferResource& g = std::get<GBufferResource>(*gi->second);
        const SceneResource& s = std::get<SceneResource>(*si->second);
        if (recreate_ || !state_) {
            if (state_) mrt_mcpg_destroy(info.ctx, state_);
            state_ = nullptr;
            check(info.ctx, mrt_mcpg_create(info.ctx, &cfg_, &state_), "mcpg state");
            recreate_ = false;
        } else if (info.iteration == 0) {
            check(info.ctx, mrt_mcpg_reset(info.ctx, state_), "mcpg reset");
        }
        mrt_pt_params p{};
        p.spp = spp_;
        p.max_path_length = max_path_length_;
        p.emission_on_primary = emission_on_primary_;
        p.demodulate_albedo = demodulate_;
        p.instance_mask = mask_;
        if (max_path_length_ != emitted_max_path_length_) { // render_pt_mcpg.cpp:212-215
            emitted_max_path_length_ = max_path_length_;
            if (info.send_event) info.send_event("bounces_changed");
        }
        ImageResource img;
        img.width = g.width;
        img.height = g.height;
        buf_.ensure(info.ctx, uint64_t(g.width) * g.height * 16);
        img.d = static_cast<float*>(buf_.d);
        check(info.ctx,
              mrt_mcpg_render_dev(info.ctx, s.scene, state_, &p, &g.camera, uint32_t(info.iteration), g.width,
                                  g.height, g.d, img.d, nullptr, nullptr),
              "mcpg");
        out["irradiance"] = img;
        return 0;
    }
    void release(mrt_ctx* c) override {
        buf_.release(c);
        if (state_) mrt_mcpg_destroy(c, state_);
        state_ = nullptr;
    }

  private:
    int32_t spp_ = 1, max_path_length_ = 5, emitted_max_path_length_ = 5;
    bool emission_on_primary_ = true, demodulate_ = false;
    uint32_t mask_ = 0xFF;
    mrt_mcpg_config cfg_{};
    bool recreate_ = true;
    mrt_mcpg* state_ = nullptr;
    DeviceBuffer buf_;
};

// "Accumulate": running mean of its source over iterations (accumulate.cpp), restarted on the
// first iteration, when the source's extent changes, when "clear" is set, and when an event
// matching "clear event pattern" arrives (accumulate.cpp:60-64,154-158; default "/user/clear").
// A restart sends the node's own "clear" event. "max history" caps the count (a number, or
// "inf"). The reference's SVGF-style reprojection, firefly filter and adaptivity settings are
// accepted and kept but not applied (a static camera needs none of them to converge).
class AccumulateNode : public Node {
  public:
    std::vector<std::string> inputs() const override { return {"src", "gbuffer", "prev_history", "prev_out"}; }
    std::vector<std::string> outputs() const override { return {"out", "history"}; }
    Value properties(const Value& in) override {
        Props p(in);
        const Value* mh = in.find("max history");
        max_history_ = mh && mh->is_number() ? uint64_t(std::max<int64_t>(1, mh->as_int())) : UINT64_MAX;
        if (!in.contains("max history")) p.get_string("max history", "inf");
        const std::string pattern = p.get_string("clear event pattern", "/user/clear");
        if (pattern != pattern_) {
            pattern_ = pattern;
            reconnect_ = true; // listeners are registered in on_connected
        }
        clear_ = clear_ || p.get_bool("clear", false);
        return p.result();
    }
    NodeStatusFlags on_connected(ConnectInfo& ci) override {
        ci.listen(pattern_, [this](const GraphEvent&) {
            clear_ = true;
            return true;
        });
        reconnect_ = false;
        return 0;
    }
    NodeStatusFlags process(RunInfo& info, const std::map<std::string, const Resource*>& in,
                            std::map<std::string, Resource>& out) override {
        auto si = in.find("src");
        if (si == in.end() || !si->second || !std::holds_alternative<ImageResource>(*si->second))
            throw node_error("input \"src\" is not connected");
        const ImageResource& src = std::get<ImageResource>(*si->second);
        if (src.width != w_ || src.height != h_ || info.iteration == 0 || clear_) {
            count_ = 0;
            w_ = src.width;
            h_ = src.height;
            clear_ = false;
            cleared_++;
            if (info.send_event) info.send_event("clear");
        }
        buf_.ensure(info.ctx, uint64_t(w_) * h_ * 16);
        const uint32_t n = uint32_t(std::min<uint64_t>(count_, max_history_ - 1));
        check(info.ctx, mrt_accumulate_dev(info.ctx, static_cast<float*>(buf_.d), src.d, w_ * h_, n, nullptr),
              "accumulate");
        count_++;
        ImageResource img{static_cast<float*>(buf_.d), w_, h_};
        out["out"] = img;
        out["history"] = img;
        return reconnect_ ? NEEDS_RECONNECT : 0;
    }
    void release(mrt_ctx* c) override { buf_.release(c); }

  private:
    uint64_t count_ = 0, max_history_ = UINT64_MAX, cleared_ = 0;
    uint32_t w_ = 0, h_ = 0;
    std::string pattern_ = "/user/clear";
    bool clear_ = false, reconnect_ = true;
    DeviceBuffer buf_;
};

// "Image Write" (image_write.cpp:23-30,155-330), iteration trigger: captures its source as PFM
// at iteration "iteration" (1-indexed) when "enable" is set, or once when "record_next" is set;
// after a capture the target iteration becomes iteration * "iteration power" + "iteration
// offset". "exit at iteration" ends the run. The filename is a format string over
// {iteration}, {record_iteration}, {width}, {height} (optionally {name:06}); ".pfm" is appended.
class ImageWriteNode : public Node {
  public:
    std::vector<std::string> inputs() const override { return {"src"}; }
    std::vector<std::string> outputs() const override { return {}; }
    Value properties(const Value& in) override {
        Props p(in);
        const std::string f = p.get_string("format", "PFM");
        if (f != "PFM") throw node_error("format '" + f + "' is not supported (PFM only)");
        filename_ = p.get_string("filename", "image_{iteration:06}");
        enable_ = p.get_bool("enable", false);
        record_next_ = p.get_bool("record_next", false);
        const std::string trig = p.get_string("trigger", "iteration");
        if (trig != "iteration") throw node_error("trigger '" + trig + "' is not supported (iteration only)");
        record_iteration_ = p.get_int("iteration", 1);
        power_ = p.get_int("iteration power", 1);
        offset_ = p.get_int("iteration offset", 0);
        exit_at_ = p.get_int("exit at iteration", -1);
        if (p.get_float("scale", 1.0) != 1.0) throw node_error("scale != 1 is not supported");
        return p.result();
    }
    NodeStatusFlags process(RunInfo& info, const std::map<std::string, const Resource*>& in,
                            std::map<std::string, Resource>&) override {
        auto si = in.find("src");
        if (si == in.end() || !si->second || !std::holds_alternative<ImageResource>(*si->second))
            throw node_error("input \"src\" is not connected");
        const ImageResource& src = std::get<ImageResource>(*si->second);
        if (record_next_ || (enable_ && iteration_ == record_iteration_)) {
            std::vector<float> host(size_t(src.width) * src.height * 4);
            check(info.ctx, mrt_download(info.ctx, host.data(), src.d, host.size() * 4), "download");
            const std::string path = format(src) + ".pfm";
            if (!pbrt::write_pfm(path, host.data(), src.width, src.height))
                throw node_error("cannot write '" + path + "'");
            if (info.log) info.log->push_back("Image Write: wrote " + path);
            record_next_ = false;
            if (enable_) record_iteration_ = record_iteration_ * power_ + offset_;
        }
        if (exit_at_ == iteration_) info.stop = true;
        iteration_++;
        return 0;
    }

  private:
    std::string format(const ImageResource& src) const {
        std::string out;
        const std::string& f = filename_;
        for (size_t k = 0; k < f.size(); k++) {
            if (f[k] != '{') {
                out += f[k];
                continue;
            }
            const size_t e = f.find('}', k);
            if (e == std::string::npos) throw node_error("bad filename format '" + f + "'");
            std::string field = f.substr(k + 1, e - k - 1), spec;
            const size_t colon = field.find(':');
            if (colon != std::string::npos) {
                spec = field.substr(colon + 1);
                field = field.substr(0, colon);
            }
            int64_t v;
            if (field == "iteration") v = iteration_;
            else if (field == "record_iteration") v = record_iteration_;
            else if (field == "width") v = src.width;
            else if (field == "height") v = src.height;
            else throw node_error("unknown filename variable '" + field + "'");
            std::string num = std::to_string(v);
            if (!spec.empty() && spec[0] == '0') {
                const size_t w = size_t(std::stoul(spec.substr(1)));
                if (num.size() < w) num = std::string(w - num.size(), '0') + num;
            }
            out += num;
            k = e;
        }
        return out;
    }
    std::string filename_;
    bool enable_ = false, record_next_ = false;
    int64_t record_iteration_ = 1, power_ = 1, offset_ = 0, exit_at_ = -1, iteration_ = 1;
};

} // namespace

void register_nodes(NodeRegistry& r) {
    r.register_node_type("PBRT Scene", "Loads a pbrt-v4 scene file.",
                         [] { return std::unique_ptr<Node>(new PbrtSceneNode()); });
    r.register_node_type("GBuffer (Raytraced)", "Primary rays and closest hits into the GBuffer (gfx950).",
                         [] { return std::unique_ptr<Node>(new GBufferNode()); });
    r.register_node_type("Render (Path-traced)", "Wavefront path tracer on gfx950.",
                         [] { return std::unique_ptr<Node>(new RenderPTNode()); });
    r.register_node_type("Render (Path-traced, MCPG)", "Markov-chain path guiding on gfx950 (surface transport).",
                         [] { return std::unique_ptr<Node>(new RenderMCPGNode()); });
    r.register_node_type("Accumulate", "Running mean of the source over iterations.",
                         [] { return std::unique_ptr<Node>(new AccumulateNode()); });
    r.register_node_type("Image Write", "Stores a graph output as PFM image.",
                         [] { return std::unique_ptr<Node>(new ImageWriteNode()); });
}

} // namespace mrt::graph

// Plugin entry point (include/merian-graph/plugin/node_plugin.hpp:17-29, plugin_export.hpp:10-18):
// registers this library's node types into a registry of this host.
extern "C" MRT_API void merian_register_nodes(mrt_node_registry* registry) {
    if (registry) mrt::graph::register_nodes(*reinterpret_cast<mrt::graph::NodeRegistry*>(registry));
}
