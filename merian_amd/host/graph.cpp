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
    // topological order (Kahn), ties by identifier
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

void Graph::run() {
    RunInfo info;
    info.ctx = ctx_;
    info.iteration = iteration_;
    info.log = &warnings_;
    for (const std::string& id : order_) {
        Instance& inst = nodes_.at(id);
        std::map<std::string, const Resource*> in;
        for (const auto& [name, srcp] : inst.sources) in[name] = output(srcp.first, srcp.second);
        const auto t0 = std::chrono::steady_clock::now();
        try {
            inst.node->process(info, in, inst.outputs);
        } catch (const std::exception& ex) {
            throw node_error("node '" + id + "' (" + inst.type + "): " + ex.what());
        }
        if (timing_) {
            mrt_synchronize(ctx_);
            inst.total_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
    }
    iteration_++;
    stop_ = stop_ || info.stop;
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
