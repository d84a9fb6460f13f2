// C ABI of the graph description engine (include/merian_amd.h, mrt_graph_*): host-only entry
// points over GraphDescription for runners and bindings that do not link C++.
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/merian_amd.h"
#include "graph_description.h"
#include "pbrt.h"

using namespace mrt;
using mrt::graph::GraphDescription;

namespace {

char* dup(const std::string& s) {
    char* p = static_cast<char*>(std::malloc(s.size() + 1));
    if (p) std::memcpy(p, s.c_str(), s.size() + 1);
    return p;
}

std::vector<std::string> split_dirs(const char* dirs) {
    std::vector<std::string> out;
    if (!dirs) return out;
    std::string s = dirs;
    size_t p = 0;
    while (true) {
        const size_t q = s.find(':', p);
        const std::string d = s.substr(p, q == std::string::npos ? std::string::npos : q - p);
        if (!d.empty()) out.push_back(d);
        if (q == std::string::npos) break;
        p = q + 1;
    }
    return out;
}

void put_errors(char** errors, const std::vector<std::string>& msgs) {
    if (!errors) return;
    std::string all;
    for (const std::string& m : msgs) all += m + "\n";
    *errors = dup(all);
}

bool parse(const char* text, json::Value& v, std::vector<std::string>& errs) {
    if (!text) {
        errs.push_back("null JSON text");
        return false;
    }
    try {
        v = json::parse(text);
        return true;
    } catch (const std::exception& e) {
        errs.push_back(e.what());
        return false;
    }
}

} // namespace

extern "C" {

MRT_API mrt_status mrt_graph_apply_cli(const char* config_json, int32_t argc, const char* const* argv,
                                       const char* search_dirs, char** out_json, char** errors) {
    std::vector<std::string> errs;
    json::Value cfg;
    if (!parse(config_json, cfg, errs) || argc < 0 || (argc > 0 && !argv)) {
        put_errors(errors, errs);
        return MRT_ERR_INVALID;
    }
    std::vector<std::string> args;
    for (int32_t k = 0; k < argc; k++) args.emplace_back(argv[k] ? argv[k] : "");
    bool ok;
    try {
        ok = GraphDescription::apply_cli(cfg, args, split_dirs(search_dirs), &errs);
    } catch (const std::exception& e) {
        errs.push_back(e.what());
        ok = false;
    }
    if (out_json) *out_json = dup(json::dump(cfg, 4));
    put_errors(errors, errs);
    return ok ? MRT_OK : MRT_ERR_INVALID;
}

MRT_API mrt_status mrt_graph_resolve_includes(const char* config_json, const char* search_dirs, char** out_json,
                                              char** errors) {
    std::vector<std::string> errs;
    json::Value cfg;
    if (!parse(config_json, cfg, errs)) {
        put_errors(errors, errs);
        return MRT_ERR_INVALID;
    }
    bool ok;
    try {
        ok = GraphDescription::resolve_includes(cfg, split_dirs(search_dirs), &errs);
    } catch (const std::exception& e) {
        errs.push_back(e.what());
        ok = false;
    }
    if (out_json) *out_json = dup(json::dump(cfg, 4));
    put_errors(errors, errs);
    return ok ? MRT_OK : MRT_ERR_INVALID;
}

MRT_API mrt_status mrt_graph_merge(const char* base_json, const char* overwrite_json, char** out_json) {
    std::vector<std::string> errs;
    json::Value base, over;
    if (!parse(base_json, base, errs) || !parse(overwrite_json, over, errs)) return MRT_ERR_INVALID;
    GraphDescription::merge_into(base, over);
    if (out_json) *out_json = dup(json::dump(base, 4));
    return MRT_OK;
}

MRT_API mrt_status mrt_graph_store(const char* graph_json, char** out_json, char** errors) {
    std::vector<std::string> errs;
    json::Value g;
    if (!parse(graph_json, g, errs)) {
        put_errors(errors, errs);
        return MRT_ERR_INVALID;
    }
    try {
        const GraphDescription d = GraphDescription::from_json(g);
        if (out_json) *out_json = dup(json::dump(d.to_json(), 4));
        return MRT_OK;
    } catch (const std::exception& e) {
        errs.push_back(e.what());
        put_errors(errors, errs);
        return MRT_ERR_INVALID;
    }
}

MRT_API mrt_status mrt_graph_cli_help(const char* config_json, const char* search_dirs, char** out_text) {
    std::vector<std::string> errs;
    json::Value cfg;
    if (!parse(config_json, cfg, errs)) return MRT_ERR_INVALID;
    try {
        if (out_text) *out_text = dup(GraphDescription::cli_help(cfg, split_dirs(search_dirs)));
    } catch (const std::exception&) {
        return MRT_ERR_INVALID;
    }
    return MRT_OK;
}

MRT_API void mrt_string_free(char* s) { std::free(s); }

MRT_API mrt_status mrt_pbrt_inspect(const char* path, char** out_json, char** errors) {
    if (!path) return MRT_ERR_INVALID;
    try {
        const pbrt::Scene sc = pbrt::load(path);
        json::Value j = json::Value::object();
        j["geometries"] = int64_t(sc.geometries.size());
        j["vertices"] = int64_t(sc.vertices.size());
        j["triangles"] = int64_t(sc.indices.size() / 3);
        json::Value geoms = json::Value::array();
        for (const mrt_geometry& g : sc.geometries) {
            json::Value e = json::Value::object();
            e["material_id"] = int64_t(g.material_id);
            e["flags"] = int64_t(g.flags);
            e["primitive_count"] = int64_t(g.primitive_count);
            geoms.push_back(e);
        }
        j["geometry_list"] = geoms;
        json::Value mats = json::Value::array();
        for (const mrt_material& m : sc.materials) {
            json::Value e = json::Value::object();
            auto v3 = [](const float* f) { return json::Value(json::Array{double(f[0]), double(f[1]), double(f[2])}); };
            e["model"] = int64_t(m.model);
            e["features"] = int64_t(m.features);
            e["base_color"] = v3(m.base_color);
            e["metalness"] = double(m.metalness);
            e["specular_alpha"] = json::Value(json::Array{double(m.specular_alpha[0]), double(m.specular_alpha[1])});
            e["specular_weight"] = double(m.specular_weight);
            e["specular_ior"] = double(m.specular_ior);
            e["transmission_weight"] = double(m.transmission_weight);
            e["emission"] = v3(m.emission);
            mats.push_back(e);
        }
        j["materials"] = mats;
        json::Value env = json::Value::object();
        env["width"] = int64_t(sc.env_width);
        env["height"] = int64_t(sc.env_height);
        env["scale"] = double(sc.env_scale);
        if (sc.env_width == 1 && sc.env_height == 1)
            env["constant"] = json::Value(json::Array{double(sc.env_rgba[0]), double(sc.env_rgba[1]), double(sc.env_rgba[2])});
        j["env"] = env;
        json::Value cam = json::Value::object();
        cam["present"] = sc.camera.present;
        auto a3 = [](const float* f) { return json::Value(json::Array{double(f[0]), double(f[1]), double(f[2])}); };
        if (sc.camera.present) {
            cam["eye"] = a3(sc.camera.eye);
            cam["forward"] = a3(sc.camera.forward);
            cam["up"] = a3(sc.camera.up);
            cam["vfov"] = double(sc.camera.vfov);
            cam["from_file"] = sc.camera.has_resolution;
            if (sc.camera.aspect > 0.f) cam["aspect"] = double(sc.camera.aspect);
        }
        cam["width"] = int64_t(sc.camera.width);
        cam["height"] = int64_t(sc.camera.height);
        j["camera"] = cam;
        json::Value w = json::Value::array();
        for (const std::string& m : sc.warnings) w.push_back(m);
        j["warnings"] = w;
        if (out_json) *out_json = dup(json::dump(j, 4));
        return MRT_OK;
    } catch (const std::exception& e) {
        put_errors(errors, {e.what()});
        return MRT_ERR_INVALID;
    }
}

namespace {
// mrt_pbrt_desc followed by the loaded scene its pointers refer to
struct PbrtDescHolder {
    mrt_pbrt_desc pub;
    pbrt::Scene scene;
};
} // namespace

MRT_API mrt_status mrt_pbrt_load_desc(const char* path, mrt_pbrt_desc** out, char** errors) {
    if (!path || !out) return MRT_ERR_INVALID;
    *out = nullptr;
    try {
        auto* h = new PbrtDescHolder();
        h->scene = pbrt::load(path);
        const pbrt::Camera& c = h->scene.camera;
        h->pub.scene = h->scene.desc();
        h->pub.has_camera = c.present ? 1 : 0;
        h->pub.camera_from_file = c.present && c.has_resolution ? 1 : 0;
        for (int k = 0; k < 3; k++) {
            h->pub.eye[k] = c.present ? c.eye[k] : 0.f;
            h->pub.forward[k] = c.present ? c.forward[k] : 0.f;
            h->pub.up[k] = c.present ? c.up[k] : 0.f;
        }
        h->pub.vfov = c.vfov;
        h->pub.aspect = c.aspect;
        h->pub.width = c.width;
        h->pub.height = c.height;
        h->pub.warning_count = uint32_t(h->scene.warnings.size());
        *out = &h->pub;
        return MRT_OK;
    } catch (const std::exception& e) {
        put_errors(errors, {e.what()});
        return MRT_ERR_INVALID;
    }
}

MRT_API void mrt_pbrt_desc_free(mrt_pbrt_desc* desc) { delete reinterpret_cast<PbrtDescHolder*>(desc); }

MRT_API mrt_status mrt_pbrt_scene_create(mrt_ctx* ctx, const char* path, mrt_scene** out) {
    if (!ctx || !path || !out) return MRT_ERR_INVALID;
    try {
        const pbrt::Scene sc = pbrt::load(path);
        const mrt_scene_desc d = sc.desc();
        return mrt_scene_create(ctx, &d, out);
    } catch (const std::exception&) {
        return MRT_ERR_INVALID;
    }
}

} // extern "C"
