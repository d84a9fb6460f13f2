// C ABI of the graph description engine (include/merian_amd.h, mrt_graph_*): host-only entry
// points over GraphDescription for runners and bindings that do not link C++.
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/merian_amd.h"
#include "graph_description.h"

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

} // extern "C"
