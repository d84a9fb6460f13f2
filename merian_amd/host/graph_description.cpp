// Graph description engine of the graph host (graph_description.h).
//
// The file has four parts:
//   1. files: locating JSON files against a list of search directories and reading them;
//   2. the description schema: names, connection strings and the readers of each version;
//   3. fragment merge and include expansion;
//   4. command-line overrides. The "cli" block of a config declares named overrides. Each is
//      read into a Declaration; applying one runs an Action (merge fragment files, then assign
//      values at JSON pointers). An OverrideRun applies one command line to a config: variant
//      defaults first, then the "--name value" pairs in argument order (a pair whose name is
//      not declared yet waits until a merged fragment declares it), then the positional rest.
#include "graph_description.h"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <optional>
#include <stdexcept>
#include <sys/stat.h>

namespace mrt::graph {

using json::Value;

namespace {

// Error messages of one engine call, and whether there were any.
struct Log {
    std::vector<std::string>* sink = nullptr;
    bool ok = true;
    void error(const std::string& message) {
        ok = false;
        if (sink) sink->push_back(message);
    }
};

// ------------------------------------------------------------------------------------------
// 1. files

bool is_regular_file(const std::string& path) {
    struct stat st;
    return ::stat(path.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

// A name is tried as given, then (if relative) inside each search directory in order.
std::optional<std::string> locate(const std::string& name, const std::vector<std::string>& dirs) {
    if (is_regular_file(name)) return name;
    if (name.empty() || name.front() == '/') return std::nullopt;
    for (const std::string& dir : dirs) {
        if (dir.empty()) continue;
        std::string candidate = dir + "/" + name;
        if (is_regular_file(candidate)) return candidate;
    }
    return std::nullopt;
}

std::string directory_of(const std::string& path) {
    const size_t slash = path.rfind('/');
    if (slash == std::string::npos) return ".";
    return slash == 0 ? std::string("/") : path.substr(0, slash);
}

std::optional<Value> load_json(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return std::nullopt;
    const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    Value v;
    if (!json::try_parse(text, v)) return std::nullopt;
    return v;
}

// "include" and "merge" accept one file name or a list of them.
std::vector<std::string> file_names(const Value& v) {
    std::vector<std::string> names;
    if (v.is_string()) names.push_back(v.as_string());
    if (v.is_array())
        for (const Value& item : v.as_array())
            if (item.is_string()) names.push_back(item.as_string());
    return names;
}

// ------------------------------------------------------------------------------------------
// 2. schema

// Node ids and connector names are joined with '.', '->' and '/' in connection strings and JSON
// pointers, so none of them may appear inside a name.
void require_plain_name(const std::string& name, const char* what) {
    for (const char* reserved : {".", "->", "/"})
        if (name.find(reserved) != std::string::npos)
            throw std::invalid_argument(std::string(what) + " \"" + name + "\" may not contain '.', '->' or '/'");
}

struct Endpoint {
    std::string node, input;
};

// "node.input", split at the first '.'
std::optional<Endpoint> parse_endpoint(const std::string& text) {
    const size_t dot = text.find('.');
    if (dot == std::string::npos || dot == 0 || dot + 1 == text.size()) return std::nullopt;
    return Endpoint{text.substr(0, dot), text.substr(dot + 1)};
}

// "output->node.input"; an output name has no '-'
std::optional<std::pair<std::string, Endpoint>> parse_link(const std::string& text) {
    const size_t arrow = text.find("->");
    if (arrow == std::string::npos || arrow == 0) return std::nullopt;
    std::string output = text.substr(0, arrow);
    if (output.find('-') != std::string::npos) return std::nullopt;
    std::optional<Endpoint> to = parse_endpoint(text.substr(arrow + 2));
    if (!to) return std::nullopt;
    return std::make_pair(std::move(output), std::move(*to));
}

const std::string& string_field(const Value& obj, const char* key, const std::string& where) {
    const Value* f = obj.find(key);
    if (!f || !f->is_string())
        throw std::runtime_error(where + ": \"" + key + "\" must be a string");
    return f->as_string();
}

void read_sections(const Value& v, GraphDescription& d) {
    if (const Value* gp = v.find("graph_properties")) d.graph_properties = *gp;
    if (const Value* pp = v.find("profiler")) d.profiler_properties = *pp;
}

// Schema v1 and v2: nodes keyed by id with a "disable" flag, connections listed separately.
void read_v1(const Value& v, GraphDescription& d) {
    read_sections(v, d);
    if (const Value* nodes = v.find("nodes"); nodes && nodes->is_object()) {
        for (const auto& [id, n] : nodes->as_object()) {
            d.add_node(string_field(n, "type", "node \"" + id + "\""), id, n.value("properties", Value()));
            d.nodes[id].enabled = !n.value("disable", false).as_bool();
        }
    }
    if (const Value* links = v.find("connections"); links && links->is_array()) {
        for (const Value& c : links->as_array()) {
            const std::string where = "connection";
            d.add_connection(string_field(c, "src", where), string_field(c, "dst", where),
                             string_field(c, "src_output", where), string_field(c, "dst_input", where));
        }
    }
}

// Schema v4: nodes keyed by id, each listing its outgoing connections.
void read_v4(const Value& v, GraphDescription& d) {
    read_sections(v, d);
    if (const Value* c = v.find("cli")) d.cli = *c;
    const Value* nodes = v.find("nodes");
    if (!nodes || !nodes->is_object()) return;
    for (const auto& [id, n] : nodes->as_object()) {
        const Value* type = n.find("type");
        if (!type || !type->is_string())
            throw std::runtime_error("node \"" + id + "\" has no \"type\" (a cli pointer into an id that names no " +
                                     "node creates an entry like this)");
        d.add_node(type->as_string(), id, n.value("properties", Value()));
        NodeEntry& e = d.nodes[id];
        e.enabled = n.value("enabled", true).as_bool();
        if (const Value* lo = n.find("linearization order")) e.linearization_order = uint32_t(lo->as_int());
        if (const Value* md = n.find("metadata")) e.metadata = *md;
        const Value* outs = n.find("outputs");
        if (!outs) continue;
        if (outs->is_array()) {
            for (const Value& link : outs->as_array()) {
                const std::string text = link.is_string() ? link.as_string() : json::dump(link);
                const auto parsed = parse_link(text);
                if (!parsed)
                    throw std::runtime_error("node \"" + id + "\": output \"" + text +
                                             "\" is not of the form output->node.input");
                d.add_connection(id, parsed->second.node, parsed->first, parsed->second.input);
            }
        } else if (outs->is_object()) {
            for (const auto& [output, targets] : outs->as_object()) {
                if (!targets.is_array()) throw std::runtime_error("node \"" + id + "\": targets of \"" + output +
                                                                  "\" must be a list");
                for (const Value& target : targets.as_array()) {
                    const std::string text = target.is_string() ? target.as_string() : json::dump(target);
                    const auto to = parse_endpoint(text);
                    if (!to)
                        throw std::runtime_error("node \"" + id + "\": target \"" + text + "\" of output \"" + output +
                                                 "\" is not of the form node.input");
                    d.add_connection(id, to->node, output, to->input);
                }
            }
        }
    }
}

// Schema v3 lists nodes as an array of objects with an "id"; v4 keys them by it.
Value v3_as_v4(const Value& v) {
    Value out = v;
    const Value* nodes = v.find("nodes");
    if (!nodes || !nodes->is_array()) return out;
    Value keyed = Value::object();
    for (const Value& n : nodes->as_array()) {
        Value entry = n;
        const std::string id = string_field(n, "id", "v3 node");
        entry.erase("id");
        keyed[id] = std::move(entry);
    }
    out["nodes"] = std::move(keyed);
    return out;
}

// ------------------------------------------------------------------------------------------
// 3. includes

// Depth-first expansion: the included files (each expanded first, relative to its own
// directory) are merged in order, then the including config on top. `open` holds the canonical
// paths of the files being expanded; meeting one of them again is a cycle.
void expand_includes(Value& config, const std::vector<std::string>& dirs, std::set<std::string>& open, Log& log) {
    const Value* include = config.find("include");
    if (!include) return;
    Value merged = Value::object();
    for (const std::string& name : file_names(*include)) {
        const std::optional<std::string> path = locate(name, dirs);
        if (!path) {
            log.error("include \"" + name + "\": file not found in the search path");
            continue;
        }
        char* real = ::realpath(path->c_str(), nullptr);
        const std::string key = real ? std::string(real) : *path;
        std::free(real);
        if (open.count(key)) {
            log.error("include \"" + *path + "\" includes itself (cycle)");
            continue;
        }
        std::optional<Value> part = load_json(*path);
        if (!part) {
            log.error("include \"" + *path + "\": not a JSON file");
            continue;
        }
        open.insert(key);
        std::vector<std::string> inner{directory_of(*path)};
        inner.insert(inner.end(), dirs.begin(), dirs.end());
        expand_includes(*part, inner, open, log);
        open.erase(key);
        GraphDescription::merge_into(merged, *part);
    }
    GraphDescription::merge_into(merged, config);
    config = std::move(merged);
}

// ------------------------------------------------------------------------------------------
// 4. command-line overrides

// What applying an override does: merge the fragment files in order, then assign values at
// JSON pointers (so a set always wins over a fragment of the same action).
struct Action {
    std::vector<std::string> fragments;
    std::vector<std::pair<std::string, Value>> assignments;

    // {"merge": file or [files], "set": {pointer: value, ...}}
    static Action from_spec(const Value& spec) {
        Action a;
        if (!spec.is_object()) return a;
        if (const Value* m = spec.find("merge")) a.fragments = file_names(*m);
        if (const Value* s = spec.find("set"); s && s->is_object())
            for (const auto& [pointer, value] : s->as_object()) a.assignments.emplace_back(pointer, value);
        return a;
    }
    size_t size() const { return fragments.size() + assignments.size(); }
};

// One entry of the "cli" block.
struct Declaration {
    enum class Form { Value, Flag, Variant };
    Form form = Form::Value;
    std::string name;         // given as --name; "--" receives the positional arguments
    bool required = false;
    std::string pointer;      // Value: where the argument is stored
    Action with;              // Value: applied with the argument; Flag: applied for "on"
    Action off;               // Flag: applied for "off"
    std::map<std::string, Action> variants;
    std::optional<std::string> fallback; // Variant: applied when the command line does not choose
    std::optional<std::string> selected; // Variant: a persisted choice, not applied again
};

Declaration::Form form_of(const Value& spec) {
    if (const Value* t = spec.find("type"); t && t->is_string()) {
        if (t->as_string() == "variant") return Declaration::Form::Variant;
        if (t->as_string() == "flag") return Declaration::Form::Flag;
        return Declaration::Form::Value;
    }
    if (spec.contains("variants")) return Declaration::Form::Variant;
    return spec.contains("pointer") ? Declaration::Form::Value : Declaration::Form::Flag;
}

// The declarations of a config. Entries set to null (removed by a merged fragment) are skipped;
// malformed ones are reported and skipped.
std::map<std::string, Declaration> declarations_of(const Value& config, Log* log) {
    std::map<std::string, Declaration> out;
    const Value* cli = config.find("cli");
    if (!cli || !cli->is_object()) return out;
    for (const auto& [name, spec] : cli->as_object()) {
        if (spec.is_null()) continue;
        if (!spec.is_object()) {
            if (log) log->error("cli entry \"" + name + "\" is not an object");
            continue;
        }
        Declaration d;
        d.name = name;
        d.form = form_of(spec);
        d.required = spec.value("required", false).as_bool();
        if (d.form == Declaration::Form::Value) {
            const Value* p = spec.find("pointer");
            if (!p || !p->is_string()) {
                if (log) log->error("cli entry \"" + name + "\" has no \"pointer\"");
                continue;
            }
            d.pointer = p->as_string();
            d.with = Action::from_spec(spec);
        } else if (d.form == Declaration::Form::Flag) {
            d.with = Action::from_spec(spec);
            if (const Value* off = spec.find("off")) d.off = Action::from_spec(*off);
        } else {
            if (const Value* vs = spec.find("variants"); vs && vs->is_object())
                for (const auto& [variant, action] : vs->as_object()) d.variants.emplace(variant, Action::from_spec(action));
            if (const Value* f = spec.find("default"); f && f->is_string()) d.fallback = f->as_string();
            if (const Value* s = spec.find("selected"); s && s->is_string()) d.selected = s->as_string();
        }
        out.emplace(name, std::move(d));
    }
    return out;
}

std::string lowercase(std::string s) {
    for (char& c : s) c = char(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

// State named by a flag's argument.
std::optional<bool> switch_state(const std::string& text) {
    const std::string t = lowercase(text);
    for (const char* on : {"", "on", "true", "yes", "1"})
        if (t == on) return true;
    for (const char* off : {"off", "false", "no", "0"})
        if (t == off) return false;
    return std::nullopt;
}

// A command-line value becomes JSON (a number stays a number), unless what it replaces is a
// string: file names such as "123" stay strings. Text that is not JSON stays text.
Value argument_value(const Value* current, const std::string& text) {
    if (current && current->is_string()) return Value(text);
    Value parsed;
    if (json::try_parse(text, parsed)) return parsed;
    return Value(text);
}

bool names_fragment_file(const std::string& value) {
    const std::string ext = ".json";
    const bool json_ext = value.size() >= ext.size() && value.compare(value.size() - ext.size(), ext.size(), ext) == 0;
    return json_ext || value.find('/') != std::string::npos;
}

std::string joined_names(const std::map<std::string, Action>& variants, const char* separator) {
    std::string s;
    for (const auto& kv : variants) {
        if (!s.empty()) s += separator;
        s += kv.first;
    }
    return s;
}

bool is_option(const std::string& arg) { return arg.size() > 2 && arg[0] == '-' && arg[1] == '-'; }

// One application of a command line to a config.
class OverrideRun {
  public:
    OverrideRun(Value& config, const std::vector<std::string>& args, const std::vector<std::string>& dirs, Log& log)
        : cfg_(config), args_(args), dirs_(dirs), log_(log), used_(args.size(), false) {
        for (const std::string& a : args_)
            if (is_option(a)) named_.insert(a.substr(2)); // an explicit choice suppresses the default
    }

    void run() {
        apply_defaults();
        if (!take_options()) return;
        take_positionals();
        check_required();
        prune_removed();
    }

    // Applies the defaults of the variants that neither have a selection nor are named on the
    // command line, re-reading the declarations after each (a fragment may declare more).
    void apply_defaults() {
        while (true) {
            const auto decls = declarations_of(cfg_, nullptr);
            const Declaration* next = nullptr;
            for (const auto& kv : decls) {
                const Declaration& d = kv.second;
                if (d.form == Declaration::Form::Variant && d.fallback && !d.selected && !named_.count(d.name) &&
                    !failed_defaults_.count(d.name)) {
                    next = &d;
                    break;
                }
            }
            if (!next) return;
            const std::string name = next->name, choice = *next->fallback;
            const auto v = next->variants.find(choice);
            if (v == next->variants.end()) {
                log_.error("--" + name + ": its default \"" + choice + "\" names no declared variant");
                failed_defaults_.insert(name);
                continue;
            }
            const Action action = v->second;
            perform(action);
            cfg_["cli"][name]["selected"] = choice;
        }
    }

  private:
    Value& cfg_;
    const std::vector<std::string>& args_;
    const std::vector<std::string>& dirs_;
    Log& log_;
    std::vector<bool> used_;
    std::set<std::string> named_;           // every --name on the command line
    std::set<std::string> given_;           // overrides that received a value
    std::set<std::string> failed_defaults_; // defaults reported as undeclared

    bool perform(const Action& a) {
        for (const std::string& name : a.fragments) {
            const std::optional<std::string> path = locate(name, dirs_);
            if (!path) {
                log_.error("fragment \"" + name + "\": file not found in the search path");
                return false;
            }
            const std::optional<Value> fragment = load_json(*path);
            if (!fragment) {
                log_.error("fragment \"" + *path + "\": not a JSON file");
                return false;
            }
            GraphDescription::merge_into(cfg_, *fragment);
        }
        for (const auto& [pointer, value] : a.assignments) json::pointer_ref(cfg_, pointer) = value;
        return true;
    }

    std::optional<Action> action_for(const Declaration& d, const std::string& value) {
        switch (d.form) {
        case Declaration::Form::Value: {
            Action a = d.with;
            a.assignments.insert(a.assignments.begin(),
                                 {d.pointer, argument_value(json::pointer_find(cfg_, d.pointer), value)});
            return a;
        }
        case Declaration::Form::Flag: {
            const std::optional<bool> on = switch_state(value);
            if (!on) {
                log_.error("--" + d.name + " takes on or off, not \"" + value + "\"");
                return std::nullopt;
            }
            return *on ? d.with : d.off;
        }
        case Declaration::Form::Variant: {
            const auto v = d.variants.find(value);
            if (v != d.variants.end()) return v->second;
            if (names_fragment_file(value)) return Action{{value}, {}};
            log_.error("--" + d.name + " " + value + ": neither a declared variant (" +
                       joined_names(d.variants, ", ") + ") nor a .json fragment file");
            return std::nullopt;
        }
        }
        return std::nullopt;
    }

    // The first unused "--name" whose name is declared now (or is the built-in "merge").
    std::optional<size_t> next_option(const std::map<std::string, Declaration>& decls) const {
        for (size_t i = 0; i < args_.size(); i++) {
            if (used_[i] || !is_option(args_[i])) continue;
            const std::string name = args_[i].substr(2);
            if (name == "merge" || decls.count(name)) return i;
        }
        return std::nullopt;
    }

    // Applies "--name value" pairs until none is applicable. Returns false if a name lacks its
    // value (the run stops there).
    bool take_options() {
        while (true) {
            const auto decls = declarations_of(cfg_, &log_);
            const std::optional<size_t> at = next_option(decls);
            if (!at) return true;
            const size_t i = *at;
            if (i + 1 >= args_.size() || used_[i + 1]) {
                log_.error("missing value after " + args_[i]);
                return false;
            }
            used_[i] = used_[i + 1] = true;
            const std::string name = args_[i].substr(2), value = args_[i + 1];
            given_.insert(name);
            if (name == "merge") {
                perform(Action{{value}, {}});
            } else {
                const Declaration& d = decls.at(name);
                const std::optional<Action> a = action_for(d, value);
                if (a && perform(*a) && d.form == Declaration::Form::Variant) cfg_["cli"][name]["selected"] = value;
            }
            apply_defaults();
        }
    }

    // Tokens no option used are positional: one is stored as it is, several are quoted and
    // joined with spaces, into the "--" override.
    void take_positionals() {
        std::vector<std::string> rest;
        for (size_t i = 0; i < args_.size(); i++)
            if (!used_[i]) rest.push_back(args_[i]);
        if (rest.empty()) return;
        std::string text;
        if (rest.size() == 1) {
            text = rest.front();
        } else {
            for (const std::string& a : rest) text += (text.empty() ? "" : " ") + quote_arg(a);
        }
        given_.insert("--");
        const auto decls = declarations_of(cfg_, nullptr);
        const auto it = decls.find("--");
        if (it == decls.end()) {
            log_.error("positional arguments (" + text + ") but the graph declares no \"--\" entry to receive them");
            return;
        }
        if (const std::optional<Action> a = action_for(it->second, text)) perform(*a);
    }

    void check_required() {
        for (const auto& [name, d] : declarations_of(cfg_, nullptr)) {
            if (!d.required || given_.count(name) || d.fallback || d.selected) continue;
            if (d.form == Declaration::Form::Value) {
                // a value the config already holds (a stored session) satisfies the requirement
                const Value* v = json::pointer_find(cfg_, d.pointer);
                if (v && (v->is_string() ? !v->as_string().empty() : !v->is_null())) continue;
            }
            log_.error(name == "--" ? std::string("the graph requires positional arguments")
                                    : "the graph requires --" + name);
        }
    }

    void prune_removed() {
        Value* cli = const_cast<Value*>(cfg_.find("cli"));
        if (!cli || !cli->is_object()) return;
        json::Object& entries = cli->as_object();
        for (auto it = entries.begin(); it != entries.end();) it = it->second.is_null() ? entries.erase(it) : ++it;
    }
};

} // namespace

std::string quote_arg(const std::string& value) {
    if (!value.empty() && value.find_first_of(" \t\n\v\f\r\"") == std::string::npos) return value;
    std::string out = "\"";
    for (const char c : value) {
        out.push_back(c);
        if (c == '"') out.push_back('"');
    }
    out.push_back('"');
    return out;
}

// ------------------------------------------------------------------------------------------
// description

void GraphDescription::add_node(const std::string& type, const std::string& id, const Value& properties) {
    require_plain_name(id, "node id");
    if (nodes.count(id)) throw std::runtime_error("node id \"" + id + "\" appears twice");
    NodeEntry e;
    e.type = type;
    e.properties = properties.is_null() ? Value::object() : properties;
    nodes.emplace(id, std::move(e));
}

void GraphDescription::add_connection(const std::string& src, const std::string& dst, const std::string& output,
                                      const std::string& input) {
    const auto it = nodes.find(src);
    if (it == nodes.end()) throw std::runtime_error("connection from \"" + src + "\", which is not a node");
    require_plain_name(output, "output");
    require_plain_name(input, "input");
    it->second.outputs[output][dst].insert(input);
}

GraphDescription GraphDescription::from_json(const Value& v) {
    GraphDescription d;
    const Value* version = v.find("version");
    const int64_t ver = version ? version->as_int() : 1;
    switch (ver) {
    case 1:
    case 2: read_v1(v, d); break;
    case 3: read_v4(v3_as_v4(v), d); break;
    case 4: read_v4(v, d); break;
    default: throw std::runtime_error("graph schema version " + std::to_string(ver) + " is not supported (1-4)");
    }
    return d;
}

Value GraphDescription::to_json() const {
    Value out = Value::object();
    out["version"] = 4;
    if (!graph_properties.empty()) out["graph_properties"] = graph_properties;
    if (!profiler_properties.empty()) out["profiler"] = profiler_properties;
    if (!cli.empty()) out["cli"] = cli;
    Value all = Value::object();
    for (const auto& [id, e] : nodes) {
        Value n = Value::object();
        n["type"] = e.type;
        n["enabled"] = e.enabled;
        if (e.linearization_order) n["linearization order"] = e.linearization_order;
        if (!e.properties.empty()) n["properties"] = e.properties;
        if (!e.metadata.empty()) n["metadata"] = e.metadata;
        Value links = Value::array();
        for (const auto& [output, targets] : e.outputs)
            for (const auto& [dst, inputs] : targets)
                for (const std::string& input : inputs) links.push_back(output + "->" + dst + "." + input);
        if (!links.empty()) n["outputs"] = std::move(links);
        all[id] = std::move(n);
    }
    if (!all.empty()) out["nodes"] = std::move(all);
    return out;
}

void GraphDescription::merge_into(Value& base, const Value& over) {
    if (!base.is_object() || !over.is_object()) {
        base = over;
        return;
    }
    static const std::string append_prefix = "$+$";
    for (const auto& [key, value] : over.as_object()) {
        if (!value.is_array() || key.compare(0, append_prefix.size(), append_prefix) != 0) {
            merge_into(base[key], value);
            continue;
        }
        Value& list = base[key.substr(append_prefix.size())];
        if (!list.is_array()) list = Value::array();
        for (const Value& item : value.as_array()) {
            const json::Array& present = list.as_array();
            if (std::find(present.begin(), present.end(), item) == present.end()) list.push_back(item);
        }
    }
}

bool GraphDescription::resolve_includes(Value& config, const std::vector<std::string>& dirs,
                                        std::vector<std::string>* errors) {
    Log log{errors};
    std::set<std::string> open;
    expand_includes(config, dirs, open, log);
    return log.ok;
}

bool GraphDescription::apply_cli(Value& config, const std::vector<std::string>& args,
                                 const std::vector<std::string>& dirs, std::vector<std::string>* errors) {
    Log log{errors};
    OverrideRun(config, args, dirs, log).run();
    return log.ok;
}

std::string GraphDescription::cli_help(const Value& config, const std::vector<std::string>& dirs) {
    Value settled = config;
    Log quiet;
    OverrideRun(settled, {}, dirs, quiet).apply_defaults();
    const auto decls = declarations_of(settled, nullptr);
    if (decls.empty()) return "this graph declares no options\n";
    std::string text = "options of this graph:\n";
    for (const auto& [name, d] : decls) {
        std::string usage, effect;
        switch (d.form) {
        case Declaration::Form::Value:
            usage = name == "--" ? "ARGS... (after the graph file)" : "--" + name + " VALUE";
            effect = "sets " + d.pointer;
            if (d.with.size()) effect += ", and applies " + std::to_string(d.with.size()) + " more change(s)";
            break;
        case Declaration::Form::Flag:
            usage = "--" + name + " on|off";
            effect = "switch";
            break;
        case Declaration::Form::Variant:
            usage = "--" + name + " " + joined_names(d.variants, "|") + "|FILE.json";
            effect = "variant";
            if (d.selected) effect += ", now " + *d.selected;
            else if (d.fallback) effect += ", by default " + *d.fallback;
            break;
        }
        text += "  " + usage + (d.required ? "   (required)" : "") + "\n      " + effect + "\n";
    }
    return text;
}

} // namespace mrt::graph
