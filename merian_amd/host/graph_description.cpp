// Graph description engine of the graph host (graph_description.h).
#include "graph_description.h"

#include <algorithm>
#include <cctype>
#include <fstream>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <sys/stat.h>

namespace mrt::graph {

using json::Array;
using json::Object;
using json::Value;

namespace {

bool file_exists(const std::string& p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

std::string parent_dir(const std::string& p) {
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string(".") : (k == 0 ? std::string("/") : p.substr(0, k));
}

std::optional<std::string> resolve_file(const std::string& file, const std::vector<std::string>& dirs) {
    if (file_exists(file)) return file;
    if (!file.empty() && file[0] == '/') return std::nullopt;
    for (const std::string& d : dirs) {
        const std::string cand = d.empty() ? file : d + "/" + file;
        if (file_exists(cand)) return cand;
    }
    return std::nullopt;
}

bool read_json_file(const std::string& path, Value& out) {
    std::ifstream in(path);
    if (!in) return false;
    std::stringstream ss;
    ss << in.rdbuf();
    return json::try_parse(ss.str(), out);
}

void note(std::vector<std::string>* errors, const std::string& msg) {
    if (errors) errors->push_back(msg);
}

void validate_identifier(const std::string& id, const char* what) {
    if (id.find('.') != std::string::npos || id.find("->") != std::string::npos || id.find('/') != std::string::npos)
        throw std::invalid_argument(std::string(what) + " '" + id + "' contains a reserved character ('.', '->' or '/')");
}

// "output->dst.input": output has no '-', dst has no '.'
bool split_arrow(const std::string& s, std::string& out, std::string& dst, std::string& in) {
    const size_t a = s.find("->");
    if (a == std::string::npos || a == 0) return false;
    out = s.substr(0, a);
    if (out.find('-') != std::string::npos) return false;
    const std::string rest = s.substr(a + 2);
    const size_t d = rest.find('.');
    if (d == std::string::npos || d == 0 || d + 1 >= rest.size()) return false;
    dst = rest.substr(0, d);
    in = rest.substr(d + 1);
    return true;
}

bool split_dot(const std::string& s, std::string& dst, std::string& in) {
    const size_t d = s.find('.');
    if (d == std::string::npos || d == 0 || d + 1 >= s.size()) return false;
    dst = s.substr(0, d);
    in = s.substr(d + 1);
    return true;
}

// ------------------------------------------------------------------------------------------
// command-line overrides

struct Patch {
    std::vector<std::pair<std::string, Value>> set; // JSON pointer -> value
    std::vector<std::string> merges;                // fragment files, merged before the sets
};

Patch parse_patch(const Value& spec) {
    Patch p;
    const Value set = spec.value("set", Value::object());
    if (set.is_object())
        for (const auto& [ptr, v] : set.as_object()) p.set.emplace_back(ptr, v);
    if (const Value* m = spec.find("merge")) {
        if (m->is_array()) {
            for (const Value& f : m->as_array()) p.merges.push_back(f.as_string());
        } else {
            p.merges.push_back(m->as_string());
        }
    }
    return p;
}

enum class Kind { Value_, Flag, Variant };

struct Override {
    std::string name;
    Kind kind = Kind::Value_;
    bool required = false;
    std::string pointer;
    Patch extra, on, off;
    std::map<std::string, Patch> variants;
    std::optional<std::string> default_variant, selected;
};

Kind classify(const Value& spec) {
    if (const Value* t = spec.find("type")) {
        const std::string& s = t->as_string();
        if (s == "variant") return Kind::Variant;
        if (s == "flag") return Kind::Flag;
        return Kind::Value_;
    }
    if (spec.contains("variants")) return Kind::Variant;
    return spec.contains("pointer") ? Kind::Value_ : Kind::Flag;
}

std::vector<Override> parse_overrides(const Value& config) {
    std::vector<Override> out;
    const Value cli = config.value("cli", Value::object());
    if (!cli.is_object()) return out;
    for (const auto& [name, spec] : cli.as_object()) {
        if (spec.is_null()) continue; // removed by a merged fragment
        Override o;
        o.name = name;
        o.kind = classify(spec);
        o.required = spec.value("required", false).as_bool();
        switch (o.kind) {
        case Kind::Value_:
            o.pointer = spec.find("pointer")->as_string();
            o.extra = parse_patch(spec);
            break;
        case Kind::Flag:
            o.on = parse_patch(spec);
            if (const Value* off = spec.find("off")) o.off = parse_patch(*off);
            break;
        case Kind::Variant:
            if (const Value* vs = spec.find("variants"))
                for (const auto& [k, p] : vs->as_object()) o.variants.emplace(k, parse_patch(p));
            if (const Value* d = spec.find("default")) o.default_variant = d->as_string();
            if (const Value* s = spec.find("selected")) o.selected = s->as_string();
            break;
        }
        out.push_back(std::move(o));
    }
    return out;
}

std::optional<bool> parse_bool(std::string t) {
    for (char& c : t) c = char(std::tolower(static_cast<unsigned char>(c)));
    if (t.empty() || t == "on" || t == "true" || t == "1" || t == "yes") return true;
    if (t == "off" || t == "false" || t == "0" || t == "no") return false;
    return std::nullopt;
}

// A command-line value keeps the type of what it replaces: text where the target is a
// string, JSON where it parses, text otherwise.
Value typed_value(const Value& config, const std::string& ptr, const std::string& text) {
    const Value* cur = json::pointer_find(config, ptr);
    if (cur && cur->is_string()) return Value(text);
    Value parsed;
    return json::try_parse(text, parsed) ? parsed : Value(text);
}

std::string variant_names(const Override& o, const char* sep) {
    std::string s;
    for (const auto& [k, _] : o.variants) s += (s.empty() ? "" : sep) + k;
    return s;
}

std::optional<Patch> resolve(const Override& o, const std::string& value, const Value& config,
                             std::vector<std::string>* errors) {
    switch (o.kind) {
    case Kind::Value_: {
        Patch p = o.extra;
        p.set.insert(p.set.begin(), {o.pointer, typed_value(config, o.pointer, value)});
        return p;
    }
    case Kind::Flag:
        if (const auto b = parse_bool(value)) return *b ? o.on : o.off;
        note(errors, "--" + o.name + " expects on|off, got '" + value + "'");
        return std::nullopt;
    case Kind::Variant: {
        auto it = o.variants.find(value);
        if (it != o.variants.end()) return it->second;
        const bool is_json = value.size() >= 5 && value.compare(value.size() - 5, 5, ".json") == 0;
        if (is_json || value.find('/') != std::string::npos) {
            Patch p;
            p.merges.push_back(value); // a custom fragment file
            return p;
        }
        note(errors, "--" + o.name + ": unknown '" + value + "', expected one of " + variant_names(o, ", ") +
                         " or a .json fragment");
        return std::nullopt;
    }
    }
    return std::nullopt;
}

bool apply_patch(Value& config, const Patch& p, const std::vector<std::string>& dirs, std::vector<std::string>* errors) {
    for (const std::string& f : p.merges) {
        const auto path = resolve_file(f, dirs);
        if (!path) {
            note(errors, "merge file '" + f + "' not found");
            return false;
        }
        Value frag;
        if (!read_json_file(*path, frag)) {
            note(errors, "merge file '" + *path + "' is not valid JSON");
            return false;
        }
        GraphDescription::merge_into(config, frag);
    }
    for (const auto& [ptr, v] : p.set) json::pointer_ref(config, ptr) = v;
    return true;
}

// Applies variant defaults that are neither selected already nor suppressed by an explicit
// argument, re-reading the overrides after each (a merged fragment may declare more).
bool settle_variants(Value& config, const std::set<std::string>& suppressed, const std::vector<std::string>& dirs,
                     std::vector<std::string>* errors) {
    bool ok = true;
    std::set<std::string> settled;
    while (true) {
        const std::vector<Override> ovs = parse_overrides(config);
        const Override* pending = nullptr;
        for (const Override& o : ovs) {
            if (o.kind == Kind::Variant && (o.default_variant || o.selected) && !settled.count(o.name) &&
                !suppressed.count(o.name)) {
                pending = &o;
                break;
            }
        }
        if (!pending) return ok;
        settled.insert(pending->name);
        if (pending->selected) continue;
        auto v = pending->variants.find(*pending->default_variant);
        if (v == pending->variants.end()) {
            note(errors, "--" + pending->name + ": default '" + *pending->default_variant + "' is not declared");
            ok = false;
            continue;
        }
        const std::string name = pending->name, def = *pending->default_variant;
        ok = apply_patch(config, v->second, dirs, errors) && ok;
        config["cli"][name]["selected"] = def;
    }
}

std::vector<std::string> file_list(const Value& v) {
    std::vector<std::string> out;
    if (v.is_array()) {
        for (const Value& f : v.as_array()) out.push_back(f.as_string());
    } else {
        out.push_back(v.as_string());
    }
    return out;
}

bool includes_rec(Value& config, const std::vector<std::string>& dirs, std::set<std::string>& active,
                  std::vector<std::string>* errors) {
    const Value* inc = config.find("include");
    if (!inc) return true;
    bool ok = true;
    Value base = Value::object();
    for (const std::string& f : file_list(*inc)) {
        const auto path = resolve_file(f, dirs);
        if (!path) {
            note(errors, "include file '" + f + "' not found");
            ok = false;
            continue;
        }
        char* rp = ::realpath(path->c_str(), nullptr);
        const std::string key = rp ? std::string(rp) : *path;
        std::free(rp);
        if (!active.insert(key).second) {
            note(errors, "include cycle at '" + *path + "'");
            ok = false;
            continue;
        }
        Value included;
        if (!read_json_file(*path, included)) {
            note(errors, "include file '" + *path + "' is not valid JSON");
            ok = false;
        } else {
            std::vector<std::string> nested{parent_dir(*path)};
            nested.insert(nested.end(), dirs.begin(), dirs.end());
            ok = includes_rec(included, nested, active, errors) && ok;
            GraphDescription::merge_into(base, included);
        }
        active.erase(key);
    }
    GraphDescription::merge_into(base, config);
    config = std::move(base);
    return ok;
}

} // namespace

std::string quote_arg(const std::string& value) {
    if (!value.empty() && value.find_first_of(" \t\n\v\f\r\"") == std::string::npos) return value;
    std::string q = "\"";
    for (const char c : value) {
        if (c == '"') q += '"';
        q += c;
    }
    return q + "\"";
}

// ------------------------------------------------------------------------------------------
// description

void GraphDescription::add_node(const std::string& type, const std::string& id, const Value& properties) {
    validate_identifier(id, "node identifier");
    if (nodes.count(id)) throw std::runtime_error("node identifier '" + id + "' is not unique");
    NodeEntry e;
    e.type = type;
    e.properties = properties.is_null() ? Value::object() : properties;
    nodes.emplace(id, std::move(e));
}

void GraphDescription::add_connection(const std::string& src, const std::string& dst, const std::string& output,
                                      const std::string& input) {
    auto it = nodes.find(src);
    if (it == nodes.end()) throw std::runtime_error("connection from unknown node '" + src + "'");
    validate_identifier(output, "output name");
    validate_identifier(input, "input name");
    it->second.outputs[output][dst].insert(input);
}

void GraphDescription::parse_v1(const Value& v, GraphDescription& d) {
    if (const Value* gp = v.find("graph_properties")) d.graph_properties = *gp;
    if (const Value* pp = v.find("profiler")) d.profiler_properties = *pp;
    if (const Value* ns = v.find("nodes")) {
        for (const auto& [id, n] : ns->as_object()) {
            d.add_node(n.find("type")->as_string(), id, n.value("properties", Value()));
            d.nodes[id].enabled = !n.value("disable", false).as_bool();
        }
    }
    if (const Value* cs = v.find("connections")) {
        for (const Value& c : cs->as_array())
            d.add_connection(c.find("src")->as_string(), c.find("dst")->as_string(), c.find("src_output")->as_string(),
                             c.find("dst_input")->as_string());
    }
}

void GraphDescription::parse_v4(const Value& v, GraphDescription& d) {
    if (const Value* gp = v.find("graph_properties")) d.graph_properties = *gp;
    if (const Value* pp = v.find("profiler")) d.profiler_properties = *pp;
    if (const Value* c = v.find("cli")) d.cli = *c;
    const Value* ns = v.find("nodes");
    if (!ns) return;
    for (const auto& [id, n] : ns->as_object()) {
        const Value* t = n.find("type");
        if (!t)
            throw std::runtime_error("node '" + id +
                                     "' has no type (a cli pointer to a nonexistent node id creates such an entry)");
        d.add_node(t->as_string(), id, n.value("properties", Value()));
        NodeEntry& e = d.nodes[id];
        e.enabled = n.value("enabled", true).as_bool();
        if (const Value* lo = n.find("linearization order")) e.linearization_order = uint32_t(lo->as_int());
        if (const Value* md = n.find("metadata")) e.metadata = *md;
    }
    for (const auto& [id, n] : ns->as_object()) {
        const Value* outs = n.find("outputs");
        if (!outs) continue;
        if (outs->is_object()) {
            for (const auto& [out, targets] : outs->as_object())
                for (const Value& t : targets.as_array()) {
                    std::string dst, in;
                    if (!split_dot(t.as_string(), dst, in))
                        throw std::runtime_error("invalid connection target '" + t.as_string() + "' for output '" + out +
                                                 "' in node '" + id + "': expected 'dst.input'");
                    d.add_connection(id, dst, out, in);
                }
        } else if (outs->is_array()) {
            for (const Value& c : outs->as_array()) {
                std::string out, dst, in;
                if (!split_arrow(c.as_string(), out, dst, in))
                    throw std::runtime_error("invalid connection '" + c.as_string() + "' in node '" + id +
                                             "': expected 'output->dst.input'");
                d.add_connection(id, dst, out, in);
            }
        }
    }
}

GraphDescription GraphDescription::from_json(const Value& v) {
    GraphDescription d;
    const Value* ver = v.find("version");
    if (!ver) {
        parse_v1(v, d);
        return d;
    }
    const int64_t version = ver->as_int();
    if (version == 2) {
        parse_v1(v, d);
    } else if (version == 3) {
        Value v4 = v;
        if (const Value* ns = v.find("nodes")) {
            Value keyed = Value::object();
            for (const Value& n : ns->as_array()) {
                Value e = n;
                const std::string id = e.find("id")->as_string();
                e.erase("id");
                keyed[id] = e;
            }
            v4["nodes"] = keyed;
        }
        parse_v4(v4, d);
    } else if (version == 4) {
        parse_v4(v, d);
    } else {
        throw std::runtime_error("schema version " + std::to_string(version) + " unsupported");
    }
    return d;
}

Value GraphDescription::to_json() const {
    Value j = Value::object();
    j["version"] = 4;
    if (!graph_properties.empty()) j["graph_properties"] = graph_properties;
    if (!profiler_properties.empty()) j["profiler"] = profiler_properties;
    if (!cli.empty()) j["cli"] = cli;
    Value ns = Value::object();
    for (const auto& [id, e] : nodes) {
        Value n = Value::object();
        n["type"] = e.type;
        n["enabled"] = e.enabled;
        if (e.linearization_order) n["linearization order"] = e.linearization_order;
        if (!e.properties.empty()) n["properties"] = e.properties;
        if (!e.metadata.empty()) n["metadata"] = e.metadata;
        if (!e.outputs.empty()) {
            Value outs = Value::array();
            for (const auto& [out, dsts] : e.outputs)
                for (const auto& [dst, ins] : dsts)
                    for (const std::string& in : ins) outs.push_back(out + "->" + dst + "." + in);
            n["outputs"] = outs;
        }
        ns[id] = n;
    }
    if (!ns.empty()) j["nodes"] = ns;
    return j;
}

void GraphDescription::merge_into(Value& base, const Value& over) {
    if (base.is_object() && over.is_object()) {
        static const std::string append = "$+$";
        for (const auto& [k, v] : over.as_object()) {
            if (v.is_array() && k.compare(0, append.size(), append) == 0) {
                Value& target = base[k.substr(append.size())];
                if (!target.is_array()) target = Value::array();
                for (const Value& e : v.as_array()) {
                    const Array& ta = target.as_array();
                    if (std::find(ta.begin(), ta.end(), e) == ta.end()) target.push_back(e);
                }
            } else {
                merge_into(base[k], v);
            }
        }
    } else {
        base = over;
    }
}

bool GraphDescription::resolve_includes(Value& config, const std::vector<std::string>& dirs,
                                        std::vector<std::string>* errors) {
    std::set<std::string> active;
    return includes_rec(config, dirs, active, errors);
}

bool GraphDescription::apply_cli(Value& config, const std::vector<std::string>& args,
                                 const std::vector<std::string>& dirs, std::vector<std::string>* errors) {
    // an explicit --name suppresses that variant's default, even if its value then fails
    std::set<std::string> raw;
    for (const std::string& a : args)
        if (a.size() > 2 && a.compare(0, 2, "--") == 0) raw.insert(a.substr(2));
    bool ok = settle_variants(config, raw, dirs, errors);

    // one recognised "--name value" at a time, re-reading the overrides after each, so that
    // overrides declared by merged fragments are recognised in the same invocation
    std::vector<bool> used(args.size(), false);
    std::set<std::string> given;
    while (true) {
        const std::vector<Override> ovs = parse_overrides(config);
        bool progressed = false;
        for (size_t i = 0; i < args.size(); i++) {
            const std::string& a = args[i];
            if (used[i] || a.size() <= 2 || a.compare(0, 2, "--") != 0) continue;
            const std::string name = a.substr(2);
            const Override* o = nullptr;
            for (const Override& c : ovs)
                if (c.name == name) o = &c;
            if (name != "merge" && !o) continue;
            if (i + 1 >= args.size() || used[i + 1]) {
                note(errors, "missing value for override '" + a + "'");
                return false;
            }
            used[i] = used[i + 1] = true;
            given.insert(name);
            const std::string value = args[i + 1];
            if (name == "merge") {
                Patch p;
                p.merges.push_back(value);
                ok = apply_patch(config, p, dirs, errors) && ok;
            } else {
                const bool variant = o->kind == Kind::Variant;
                const std::optional<Patch> p = resolve(*o, value, config, errors);
                const bool done = p && apply_patch(config, *p, dirs, errors);
                ok = done && ok;
                if (done && variant) config["cli"][name]["selected"] = value;
            }
            ok = settle_variants(config, raw, dirs, errors) && ok;
            progressed = true;
            break;
        }
        if (!progressed) break;
    }

    // the remaining tokens are positional and feed the "--" override
    std::vector<std::string> pos;
    for (size_t i = 0; i < args.size(); i++)
        if (!used[i]) pos.push_back(args[i]);
    std::string joined;
    if (pos.size() == 1) {
        joined = pos[0];
    } else {
        for (const std::string& a : pos) joined += (joined.empty() ? "" : " ") + quote_arg(a);
    }
    const std::vector<Override> ovs = parse_overrides(config);
    if (!pos.empty()) {
        given.insert("--");
        const Override* o = nullptr;
        for (const Override& c : ovs)
            if (c.name == "--") o = &c;
        if (!o) {
            note(errors, "positional arguments given but '--' is not declared in the cli block");
            ok = false;
        } else {
            const std::optional<Patch> p = resolve(*o, joined, config, errors);
            ok = p && apply_patch(config, *p, dirs, errors) && ok;
        }
    }
    for (const Override& o : ovs) {
        if (!o.required || given.count(o.name) || o.default_variant || o.selected) continue;
        if (o.kind == Kind::Value_) {
            // a value already materialised in the config (a stored session) satisfies it
            const Value* v = json::pointer_find(config, o.pointer);
            if (v && (v->is_string() ? !v->as_string().empty() : !v->is_null())) continue;
        }
        note(errors, o.name == "--" ? std::string("missing required positional args (see --help)")
                                    : "missing required --" + o.name);
        ok = false;
    }
    if (Value* cli = const_cast<Value*>(config.find("cli")); cli && cli->is_object()) {
        Object& co = cli->as_object();
        for (auto it = co.begin(); it != co.end();) it = it->second.is_null() ? co.erase(it) : std::next(it);
    }
    return ok;
}

std::string GraphDescription::cli_help(const Value& config, const std::vector<std::string>& dirs) {
    Value settled = config;
    settle_variants(settled, {}, dirs, nullptr);
    const std::vector<Override> ovs = parse_overrides(settled);
    if (ovs.empty()) return "this graph declares no cli overrides\n";
    std::string out = "cli overrides:\n";
    for (const Override& o : ovs) {
        std::string usage, target;
        switch (o.kind) {
        case Kind::Value_:
            usage = o.name == "--" ? "args... (after graph.json)" : "--" + o.name + " <value>";
            target = o.pointer;
            if (size_t n = o.extra.set.size() + o.extra.merges.size()) target += " (+" + std::to_string(n) + " more)";
            break;
        case Kind::Flag:
            usage = "--" + o.name + " <on|off>";
            target = "on/off";
            break;
        case Kind::Variant:
            usage = "--" + o.name + " <" + variant_names(o, "|") + "|file.json>";
            target = o.selected ? "selected: " + *o.selected : (o.default_variant ? "default: " + *o.default_variant : "");
            break;
        }
        usage.resize(std::max<size_t>(usage.size(), 28), ' ');
        out += "  " + usage + " -> " + target + (o.required ? " (required)" : "") + "\n";
    }
    return out;
}

} // namespace mrt::graph
