// Node registry and graph execution of the graph host (graph.h).
#include "graph.h"

#include <algorithm>
#include <cstdio>
#include <deque>
#include <set>

namespace mrt::graph {

void NodeRegistry::register_node_type(const std::string& name, const std::string& description,
                                      std::function<std::unique_ptr<Node>()> create) {
    types_[name] = NodeType{name, description, std::move(create)};
}

const NodeType* NodeRegistry::find(const std::string& name) const {
    auto it = types_.find(name);
    return it == types_.end() ? nullptr : &it->second;
}

std::vector<std::string> NodeRegistry::names() const {
    std::vector<std::string> out;
    for (const auto& [k, _] : types_) out.push_back(k);
    return out;
}

NodeRegistry& NodeRegistry::builtin() {
    static NodeRegistry r = [] {
        NodeRegistry reg;
        register_nodes(reg);
        return reg;
    }();
    return r;
}

Graph::Graph(mrt_ctx* ctx, NodeRegistry& registry) : ctx_(ctx), registry_(registry) {}

Graph::~Graph() {
    for (auto& [_, inst] : nodes_) inst.node->release(ctx_);
}

void Graph::load(const GraphDescription& d) {
    for (auto& [_, inst] : nodes_) inst.node->release(ctx_);
    nodes_.clear();
    order_.clear();
    warnings_.clear();
    desc_ = d;
    for (const auto& [id, e] : d.nodes) {
        if (!e.enabled) continue;
        const NodeType* t = registry_.find(e.type);
        if (!t) {
            warnings_.push_back("node '" + id + "' of type '" + e.type + "' is not provided by this build; skipped");
            continue;
        }
        Instance inst;
        inst.id = id;
        inst.type = e.type;
        inst.node = t->create();
        try {
            inst.props = inst.node->properties(e.properties);
        } catch (const std::exception& ex) {
            throw node_error("node '" + id + "' (" + e.type + "): " + ex.what());
        }
        nodes_.emplace(id, std::move(inst));
    }
    for (const auto& [src, e] : d.nodes) {
        auto s = nodes_.find(src);
        for (const auto& [out, dsts] : e.outputs)
            for (const auto& [dst, ins] : dsts) {
                auto t = nodes_.find(dst);
                if (s == nodes_.end() || t == nodes_.end()) continue; // an endpoint was skipped
                if (src == dst) continue; // feedback of a node's own previous output: kept internally
                const auto outs = s->second.node->outputs();
                if (std::find(outs.begin(), outs.end(), out) == outs.end())
                    throw node_error("node '" + src + "' has no output '" + out + "'");
                const auto ins_decl = t->second.node->inputs();
                for (const std::string& in : ins) {
                    if (std::find(ins_decl.begin(), ins_decl.end(), in) == ins_decl.end())
                        throw node_error("node '" + dst + "' has no input '" + in + "'");
                    t->second.sources[in] = {src, out};
                }
            }
    }
    order_nodes();
    connected_ = false;
}

// Topological order (Kahn), ties by identifier.
void Graph::order_nodes() {
    order_.clear();
    std::map<std::string, int> indeg;
    std::map<std::string, std::set<std::string>> succ;
    for (const auto& [id, inst] : nodes_) {
        indeg[id];
        for (const auto& [in, srcp] : inst.sources)
            if (succ[srcp.first].insert(id).second) indeg[id]++;
    }
    std::set<std::string> ready;
    for (const auto& [id, k] : indeg)
        if (k == 0) ready.insert(id);
    while (!ready.empty()) {
        const std::string id = *ready.begin();
        ready.erase(ready.begin());
        order_.push_back(id);
        for (const std::string& n : succ[id])
            if (--indeg[n] == 0) ready.insert(n);
    }
    if (order_.size() != nodes_.size()) throw node_error("the graph has a cycle");
}

GraphDescription Graph::store() const {
    GraphDescription d = desc_;
    for (const auto& [id, inst] : nodes_) d.nodes[id].properties = inst.props;
    return d;
}

const Resource* Graph::output(const std::string& node, const std::string& name) const {
    auto n = nodes_.find(node);
    if (n == nodes_.end()) return nullptr;
    auto o = n->second.outputs.find(name);
    return o == n->second.outputs.end() ? nullptr : &o->second;
}

std::vector<EventPattern> parse_event_patterns(const std::string& list, std::vector<std::string>* invalid) {
    std::vector<EventPattern> out;
    size_t begin = 0;
    while (begin <= list.size()) {
        size_t end = list.find(',', begin);
        if (end == std::string::npos) end = list.size();
        const std::string item = list.substr(begin, end - begin);
        begin = end + 1;
        if (item.empty()) continue;
        const size_t a = item.find('/');
        const size_t b = a == std::string::npos ? a : item.find('/', a + 1);
        if (b == std::string::npos || item.find('/', b + 1) != std::string::npos || b + 1 == item.size()) {
            if (invalid) invalid->push_back(item);
            continue;
        }
        out.push_back(EventPattern{item.substr(0, a), item.substr(a + 1, b - a - 1), item.substr(b + 1)});
    }
    return out;
}

void Graph::dispatch(const GraphEvent& e) {
    // a listener may send further events; iterate over a copy of the list
    const auto listeners = listeners_;
    for (const auto& [pattern, listener] : listeners)
        if (pattern.matches(e)) listener(e);
}

void Graph::send_event(const std::string& name) { dispatch(GraphEvent{"", "user", name}); }

// Every node's on_connected, in run order, with a fresh set of listeners (the reference
// re-registers listeners on each connect, graph.cpp:268-290).
void Graph::connect() {
    listeners_.clear();
    for (const std::string& id : order_) {
        Instance& inst = nodes_.at(id);
        ConnectInfo ci;
        ci.ctx = ctx_;
        ci.identifier = id;
        ci.listen = [this, id](const std::string& patterns, EventListener listener) {
            std::vector<std::string> bad;
            for (const EventPattern& p : parse_event_patterns(patterns, &bad)) listeners_.emplace_back(p, listener);
            for (const std::string& b : bad)
                warnings_.push_back("node '" + id + "': invalid event pattern '" + b + "' (expected type/identifier/event)");
        };
        try {
            inst.node->on_connected(ci);
        } catch (const std::exception& ex) {
            throw node_error("node '" + id + "' (" + inst.type + "): " + ex.what());
        }
    }
    connected_ = true;
}

void Graph::remove_node(const std::string& id) {
    auto it = nodes_.find(id);
    if (it == nodes_.end()) return;
    it->second.node->release(ctx_);
    nodes_.erase(it);
    for (auto& [_, inst] : nodes_)
        for (auto s = inst.sources.begin(); s != inst.sources.end();)
            s = s->second.first == id ? inst.sources.erase(s) : std::next(s);
    desc_.nodes[id].enabled = false;
    order_nodes();
    connected_ = false;
}

void Graph::set_properties(const std::string& node, const json::Value& properties) {
    auto it = nodes_.find(node);
    if (it == nodes_.end()) throw node_error("no node '" + node + "'");
    json::Value merged = it->second.props;
    GraphDescription::merge_into(merged, properties);
    try {
        it->second.props = it->second.node->properties(merged);
    } catch (const std::exception& ex) {
        throw node_error("node '" + node + "' (" + it->second.type + "): " + ex.what());
    }
}

void Graph::run() {
    if (!connected_) connect();
    RunInfo info;
    info.ctx = ctx_;
    info.iteration = iteration_;
    info.log = &warnings_;
    std::map<std::string, NodeStatusFlags> flags;
    const auto inputs_of = [this](const Instance& inst) {
        std::map<std::string, const Resource*> in;
        for (const auto& [name, srcp] : inst.sources) in[name] = output(srcp.first, srcp.second);
        return in;
    };
    const auto sender = [this, &info](const Instance& inst) {
        info.send_event = [this, type = inst.type, id = inst.id](const std::string& name) {
            dispatch(GraphEvent{type, id, name});
        };
    };
    for (const std::string& id : order_) {
        Instance& inst = nodes_.at(id);
        sender(inst);
        try {
            flags[id] |= inst.node->pre_process(info, inputs_of(inst));
        } catch (const std::exception& ex) {
            throw node_error("node '" + id + "' (" + inst.type + "): " + ex.what());
        }
    }
    for (const std::string& id : order_) {
        Instance& inst = nodes_.at(id);
        sender(inst);
        const auto t0 = std::chrono::steady_clock::now();
        try {
            flags[id] |= inst.node->process(info, inputs_of(inst), inst.outputs);
        } catch (const std::exception& ex) {
            throw node_error("node '" + id + "' (" + inst.type + "): " + ex.what());
        }
        if (timing_) {
            mrt_synchronize(ctx_);
            inst.total_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
    }
    info.send_event = nullptr;
    iteration_++;
    stop_ = stop_ || info.stop;
    // status flags: removed nodes leave the graph, and any of them or a reconnect request
    // makes the next run call on_connected again. One iteration is in flight at a time here,
    // so RESET_IN_FLIGHT_DATA has nothing to drop.
    for (const auto& [id, f] : flags) {
        if (f & REMOVE_NODE) remove_node(id);
        if (f & NEEDS_RECONNECT) connected_ = false;
    }
}

std::string Graph::report() const {
    std::string out = "node times (average over " + std::to_string(iteration_) + " iterations):\n";
    for (const std::string& id : order_) {
        const Instance& inst = nodes_.at(id);
        char buf[256];
        std::snprintf(buf, sizeof buf, "  %-24s %-22s %10.3f ms\n", id.c_str(), inst.type.c_str(),
                      iteration_ ? inst.total_ms / double(iteration_) : 0.0);
        out += buf;
    }
    return out;
}

} // namespace mrt::graph
