// Graph descriptions as merian stores them (schema v4; v1-v3 read), the fragment merge, include
// resolution and the command-line override engine of the graph runner.
//
// The behaviour is merian's GraphDescription (src/merian-graph/graph/graph_description.cpp):
// from_json/to_json (:119-238; the v3 node array becomes the v4 map, :246-256), connection
// strings (:296-323), merge_into with "$+$key" appends (:706-729), resolve_includes (:659-705)
// and apply_cli (:731-857): value, flag and variant overrides, fragment merges before pointer
// sets, variant defaults and persisted selections, positional arguments for the "--" entry.
// Errors are collected as messages instead of logged.
#pragma once

#include <map>
#include <set>
#include <string>
#include <vector>

#include "json.h"

namespace mrt::graph {

struct NodeEntry {
    std::string type;
    json::Value properties = json::Value::object();
    bool enabled = true;
    uint32_t linearization_order = 0;
    json::Value metadata = json::Value::object();
    // output name -> destination node -> destination inputs
    std::map<std::string, std::map<std::string, std::set<std::string>>> outputs;
};

class GraphDescription {
  public:
    std::map<std::string, NodeEntry> nodes;
    json::Value graph_properties = json::Value::object();
    json::Value profiler_properties = json::Value::object();
    json::Value cli = json::Value::object();

    // Throws std::runtime_error on an invalid description.
    static GraphDescription from_json(const json::Value& v);
    json::Value to_json() const;

    void add_node(const std::string& type, const std::string& id, const json::Value& properties);
    void add_connection(const std::string& src, const std::string& dst, const std::string& output,
                        const std::string& input);

    // Recursive object merge; scalars and plain arrays replace; "$+$key": [...] appends the
    // elements missing from base["key"].
    static void merge_into(json::Value& base, const json::Value& overwrite);

    // Replaces config by the merge of its "include" files (in order, resolved against
    // search_dirs, nested includes relative to their own file) with config on top.
    static bool resolve_includes(json::Value& config, const std::vector<std::string>& search_dirs,
                                 std::vector<std::string>* errors = nullptr);

    // Applies the config's "cli" overrides to the arguments. Returns false on any error;
    // config may be partially modified then.
    static bool apply_cli(json::Value& config, const std::vector<std::string>& args,
                          const std::vector<std::string>& search_dirs = {},
                          std::vector<std::string>* errors = nullptr);

    // Usage text of the overrides a config declares (after its default variants are applied).
    static std::string cli_help(const json::Value& config, const std::vector<std::string>& search_dirs = {});
};

// Quotes one argument for a space-joined positional string, as merian's quote_arg
// (include/merian/utils/string.hpp:114-126): arguments without blanks or quotes stay as they
// are; others are wrapped in double quotes with inner quotes doubled.
std::string quote_arg(const std::string& value);

} // namespace mrt::graph
