// Graph descriptions as merian stores them (schema v4; v1-v3 read), the fragment merge, include
// resolution and the command-line override engine of the graph runner.
//
// Semantics follow merian's GraphDescription (src/merian-graph/graph/graph_description.cpp):
// from_json/to_json (:119-238, v3 array -> v4 map :246-256), parse of connections (:296-323),
// merge_into (:706-729, "$+$key" appends), resolve_includes (:659-705), apply_cli (:731-857)
// with value / flag / variant overrides (:443-557), patches that merge before they set (:593-610),
// variant defaults and persisted selections (:612-643). Errors are collected as messages
// instead of logged.
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

    // Applies the config's "cli" overrides to the arguments (see the file comment). Returns
    // false on any error; config may be partially modified then.
    static bool apply_cli(json::Value& config, const std::vector<std::string>& args,
                          const std::vector<std::string>& search_dirs = {},
                          std::vector<std::string>* errors = nullptr);

    static std::string cli_help(const json::Value& config, const std::vector<std::string>& search_dirs = {});

  private:
    static void parse_v1(const json::Value& v, GraphDescription& d);
    static void parse_v4(const json::Value& v, GraphDescription& d);
};

// Quotes one argument for a space-joined positional string (merian's quote_arg,
// include/merian/utils/string.hpp:114-126).
std::string quote_arg(const std::string& value);

} // namespace mrt::graph
