// Node host of the graph runner: nodes with merian's lifecycle (src/merian-graph/graph/node.hpp:
// 53-230): inputs and outputs, properties as JSON, on_connected after the graph is wired,
// pre_process and process once per run in topological order, each returning status flags, and
// graph events ("<node type>/<identifier>/<event>", graph_events.cpp:7-63) that nodes send and
// listen to. Also the registry of node types (node_registry.cpp:63-96) and a graph built from a
// GraphDescription. Nodes reach the device only through the C ABI (include/merian_amd.h).
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
    mrt_packed_hit* d = nullptr; // hit texture (PackedHitInfo per pixel)
    float* albedo = nullptr;     // albedo texture, float4(albedo, 1) per pixel
    uint32_t width = 0, height = 0;
    mrt_camera camera{};
    mrt_scene* scene = nullptr;
};
struct ImageResource { // float4 per pixel, row-major, device memory
    float* d = nullptr;
    uint32_t width = 0, height = 0;
};
using Resource = std::variant<std::monostate, SceneResource, GBufferResource, ImageResource>;

// Status flags of a node's hooks (node.hpp:57-65).
using NodeStatusFlags = uint32_t;
enum NodeStatusFlagBits : uint32_t {
    NEEDS_RECONNECT = 0b1,      // the graph calls on_connected of every node before the next run
    RESET_IN_FLIGHT_DATA = 0b10, // data kept for iterations in flight is dropped
    REMOVE_NODE = 0b100          // the node is taken out of the graph after this run
};

// One event. Nodes send events under their type and identifier; the host sends user events
// (identifier "user") and graph events (identifier "graph") with an empty node type.
struct GraphEvent {
    std::string node_type, identifier, name;
};
// Returns true if the event was handled.
using EventListener = std::function<bool(const GraphEvent&)>;

// A listener pattern "<node type>/<identifier>/<event>": an empty type or identifier matches
// any; several patterns are separated by commas.
struct EventPattern {
    std::string node_type, identifier, name;
    bool matches(const GraphEvent& e) const {
        return e.name == name && (node_type.empty() || node_type == e.node_type) &&
               (identifier.empty() || identifier == e.identifier);
    }
};
// Parses a comma-separated pattern list; malformed entries are returned in `invalid`.
std::vector<EventPattern> parse_event_patterns(const std::string& list, std::vector<std::string>* invalid = nullptr);

struct RunInfo {
    mrt_ctx* ctx = nullptr;
    uint64_t iteration = 0; // total graph iterations so far (0-based), the RNG frame index
    bool stop = false;      // set by a node to end the run loop (Image Write "exit at iteration")
    std::vector<std::string>* log = nullptr;
    // Sends an event from the node being run (its type and identifier are filled in); listeners
    // are called before send_event returns.
    std::function<void(const std::string& name)> send_event;
};

// What a node may do in on_connected.
struct ConnectInfo {
    mrt_ctx* ctx = nullptr;
    std::string identifier;
    // Registers a listener for a comma-separated pattern list (NodeIOLayout::register_event_listener).
    std::function<void(const std::string& patterns, EventListener listener)> listen;
};

class Node {
  public:
    virtual ~Node() = default;
    virtual std::vector<std::string> inputs() const = 0;
    virtual std::vector<std::string> outputs() const = 0;
    // Loads properties (missing keys keep their defaults; throws node_error on a value the
    // node cannot honour) and returns them with every key filled in.
    virtual json::Value properties(const json::Value& in) = 0;
    // After the graph was (re)wired, before the next run.
    virtual NodeStatusFlags on_connected(ConnectInfo&) { return 0; }
    // Every run, for every node in topological order, before any node's process.
    virtual NodeStatusFlags pre_process(RunInfo&, const std::map<std::string, const Resource*>&) { return 0; }
    virtual NodeStatusFlags process(RunInfo& info, const std::map<std::string, const Resource*>& in,
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
    // One iteration: on_connected of every node first if the graph was (re)wired, then
    // pre_process and process of every node in topological order.
    void run();
    // A user event ("/user/<name>"), delivered to the listeners at once.
    void send_event(const std::string& name);
    // Loads new properties into one node, as a configuration change between iterations.
    void set_properties(const std::string& node, const json::Value& properties);
    // Identifiers of the nodes in the graph, in run order.
    const std::vector<std::string>& order() const { return order_; }
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
    bool connected_ = false;
    std::vector<std::pair<EventPattern, EventListener>> listeners_;

    void connect();
    void dispatch(const GraphEvent& e);
    void remove_node(const std::string& id);
    void order_nodes();
};

} // namespace mrt::graph
