// Node host of the graph runner: nodes with merian's lifecycle reduced to the pt path
// (src/merian-graph/node/node.hpp:53-230: inputs, outputs, properties as JSON, process once per
// run in topological order), the registry of node types (node_registry.cpp:63-96) and a graph
// built from a GraphDescription. Nodes reach the device only through the C ABI
// (include/merian_amd.h).
#pragma once

#include <chrono>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

#include "../../include/merian_amd.h"
#include "graph_description.h"
#include "pbrt.h"

namespace mrt::graph {

// merian's node_error: the node cannot run with its current configuration.
struct node_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct SceneResource {
    mrt_scene* scene = nullptr;
    pbrt::Camera camera;  // camera of the scene file (camera.present = false: none)
    uint64_t version = 0; // increments when the scene is reloaded
};
struct GBufferResource {
    mrt_packed_hit* d = nullptr;
    uint32_t width = 0, height = 0;
    mrt_camera camera{};
    mrt_scene* scene = nullptr;
};
struct ImageResource { // float4 per pixel, row-major, device memory
    float* d = nullptr;
    uint32_t width = 0, height = 0;
};
using Resource = std::variant<std::monostate, SceneResource, GBufferResource, ImageResource>;

struct RunInfo {
    mrt_ctx* ctx = nullptr;
    uint64_t iteration = 0; // total graph iterations so far (0-based), the RNG frame index
    bool stop = false;      // set by a node to end the run loop (Image Write "exit at iteration")
    std::vector<std::string>* log = nullptr;
};

class Node {
  public:
    virtual ~Node() = default;
    virtual std::vector<std::string> inputs() const = 0;
    virtual std::vector<std::string> outputs() const = 0;
    // Loads properties (missing keys keep their defaults; throws node_error on a value the
    // node cannot honour) and returns them with every key filled in.
    virtual json::Value properties(const json::Value& in) = 0;
    virtual void process(RunInfo& info, const std::map<std::string, const Resource*>& in,
                         std::map<std::string, Resource>& out) = 0;
    virtual void release(mrt_ctx*) {}
};

struct NodeType {
    std::string name, description;
    std::function<std::unique_ptr<Node>()> create;
};

class NodeRegistry {
  public:
    void register_node_type(const std::string& name, const std::string& description,
                            std::function<std::unique_ptr<Node>()> create);
    const NodeType* find(const std::string& name) const;
    std::vector<std::string> names() const;
    // "PBRT Scene", "GBuffer (Raytraced)", "Render (Path-traced)", "Accumulate", "Image Write"
    static NodeRegistry& builtin();

  private:
    std::map<std::string, NodeType> types_;
};

// Registers this library's node types into a registry (the plugin entry point's work).
void register_nodes(NodeRegistry& registry);

class Graph {
  public:
    explicit Graph(mrt_ctx* ctx, NodeRegistry& registry = NodeRegistry::builtin());
    ~Graph();
    // Instantiates the enabled nodes. Node types this build does not provide (windows, UI,
    // post-processing) are skipped with a warning, together with their connections.
    void load(const GraphDescription& d);
    // The description with every node's current properties.
    GraphDescription store() const;
    // One iteration: every node in topological order.
    void run();
    bool stop_requested() const { return stop_; }
    uint64_t iterations() const { return iteration_; }
    // Per-node average time over the run (device work included), one line per node.
    std::string report() const;
    const std::vector<std::string>& warnings() const { return warnings_; }
    const Resource* output(const std::string& node, const std::string& name) const;
    void set_timing(bool on) { timing_ = on; }

  private:
    struct Instance {
        std::string id, type;
        std::unique_ptr<Node> node;
        json::Value props;
        // input name -> (source node, output name)
        std::map<std::string, std::pair<std::string, std::string>> sources;
        std::map<std::string, Resource> outputs;
        double total_ms = 0.0;
    };
    mrt_ctx* ctx_;
    NodeRegistry& registry_;
    GraphDescription desc_;
    std::map<std::string, Instance> nodes_;
    std::vector<std::string> order_;
    std::vector<std::string> warnings_;
    uint64_t iteration_ = 0;
    bool stop_ = false;
    bool timing_ = false;
};

} // namespace mrt::graph
