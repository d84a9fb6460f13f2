// CPU unit test of the graph host's node lifecycle (merian_amd/host/graph.cpp): on_connected
// before the first run and after NEEDS_RECONNECT, pre_process of every node before any process,
// REMOVE_NODE, and graph events matched by "<node type>/<identifier>/<event>" patterns
// (graph_events.cpp:13-63 semantics). Fake node types, no device: the C ABI calls the host makes
// outside the nodes are stubbed. Built and run by tests/test_native.py with g++.
#include "../../merian_amd/host/graph.h"

#include <cstdio>
#include <string>
#include <vector>

static int failures = 0;
#define CHECK(c)                                                       \
    do {                                                               \
        if (!(c)) {                                                    \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            failures++;                                                \
        }                                                              \
    } while (0)

extern "C" mrt_status mrt_synchronize(mrt_ctx*) { return MRT_OK; }
namespace mrt::graph {
void register_nodes(NodeRegistry&) {}
} // namespace mrt::graph

using namespace mrt::graph;
using mrt::json::Value;

static std::vector<std::string> trace; // "<id>:<hook>" in call order

// "Emitter": sends the event named by its "event" property once per run when "fire" is set;
// returns the status flags named by "flags" from process.
struct Emitter : Node {
    std::string id, event;
    bool fire = false;
    uint32_t flags = 0;
    std::vector<std::string> inputs() const override { return {"in"}; }
    std::vector<std::string> outputs() const override { return {"out"}; }
    Value properties(const Value& in) override {
        event = in.value("event", Value("changed")).as_string();
        fire = in.value("fire", false).as_bool();
        flags = uint32_t(in.value("flags", Value(int64_t(0))).as_int());
        return in;
    }
    NodeStatusFlags on_connected(ConnectInfo& ci) override {
        id = ci.identifier;
        trace.push_back(id + ":connect");
        return 0;
    }
    NodeStatusFlags pre_process(RunInfo&, const std::map<std::string, const Resource*>&) override {
        trace.push_back(id + ":pre");
        return 0;
    }
    NodeStatusFlags process(RunInfo& info, const std::map<std::string, const Resource*>&,
                            std::map<std::string, Resource>& out) override {
        trace.push_back(id + ":process");
        if (fire) info.send_event(event);
        out["out"] = ImageResource{};
        const uint32_t f = flags;
        flags = 0;
        return f;
    }
};

// "Listener": counts the events matching its "pattern" property.
struct Listener : Node {
    std::string id, pattern;
    int* hits = nullptr;
    std::vector<GraphEvent>* seen = nullptr;
    std::vector<std::string> inputs() const override { return {"in"}; }
    std::vector<std::string> outputs() const override { return {"out"}; }
    Value properties(const Value& in) override {
        pattern = in.value("pattern", Value("")).as_string();
        return in;
    }
    NodeStatusFlags on_connected(ConnectInfo& ci) override {
        id = ci.identifier;
        trace.push_back(id + ":connect");
        ci.listen(pattern, [this](const GraphEvent& e) {
            if (seen) seen->push_back(e);
            return true;
        });
        return 0;
    }
    NodeStatusFlags process(RunInfo&, const std::map<std::string, const Resource*>&,
                            std::map<std::string, Resource>& out) override {
        trace.push_back(id + ":process");
        out["out"] = ImageResource{};
        return 0;
    }
};

static std::vector<GraphEvent> seen_a, seen_b;

int main() {
    {   // pattern parsing
        std::vector<std::string> bad;
        const auto ps = parse_event_patterns("/user/clear,//bounces_changed,Render (Path-traced)/render/x,a/b,,x/y/", &bad);
        CHECK(ps.size() == 3);
        CHECK(bad.size() == 2);
        CHECK(ps[0].node_type.empty() && ps[0].identifier == "user" && ps[0].name == "clear");
        CHECK(ps[1].node_type.empty() && ps[1].identifier.empty() && ps[1].name == "bounces_changed");
        CHECK(ps[2].node_type == "Render (Path-traced)" && ps[2].identifier == "render");
        const GraphEvent user{"", "user", "clear"}, node{"Render (Path-traced)", "render", "bounces_changed"};
        CHECK(ps[0].matches(user) && !ps[0].matches(node));
        CHECK(ps[1].matches(node) && !ps[1].matches(user));
        CHECK(!ps[2].matches(node)); // event name differs
        CHECK((EventPattern{"", "", "clear"}.matches(user)));
        CHECK(!(EventPattern{"Accumulate", "", "clear"}.matches(user)));
    }
    NodeRegistry reg;
    reg.register_node_type("Emitter", "", [] { return std::unique_ptr<Node>(new Emitter()); });
    reg.register_node_type("Listener", "", [] {
        static int k = 0;
        auto* l = new Listener();
        l->seen = (k++ % 2 == 0) ? &seen_a : &seen_b;
        return std::unique_ptr<Node>(l);
    });
    GraphDescription d;
    Value pe = Value::object();
    pe["event"] = "bounces_changed";
    d.add_node("Emitter", "e", pe);
    Value pa = Value::object();
    pa["pattern"] = "//bounces_changed";
    d.add_node("Listener", "la", pa); // created first (map order): seen_a
    Value pb = Value::object();
    pb["pattern"] = "/user/clear,Emitter//other";
    d.add_node("Listener", "lb", pb); // seen_b
    d.add_connection("e", "la", "out", "in");
    d.add_connection("la", "lb", "out", "in");
    Graph g(nullptr, reg);
    g.load(d);
    CHECK((g.order() == std::vector<std::string>{"e", "la", "lb"}));

    // first run: on_connected of every node, then every pre_process, then every process
    g.run();
    CHECK((trace == std::vector<std::string>{"e:connect", "la:connect", "lb:connect", "e:pre", "e:process",
                                             "la:process", "lb:process"}));
    CHECK(seen_a.empty() && seen_b.empty());

    // a node event reaches the listener whose pattern matches, with the sender's type and id
    trace.clear();
    Value fire = Value::object();
    fire["fire"] = true;
    g.set_properties("e", fire);
    g.run();
    CHECK(seen_a.size() == 1 && seen_b.empty());
    CHECK(seen_a.size() == 1 && seen_a[0].node_type == "Emitter" && seen_a[0].identifier == "e" &&
          seen_a[0].name == "bounces_changed");
    CHECK((trace == std::vector<std::string>{"e:pre", "e:process", "la:process", "lb:process"})); // no reconnect

    // a user event reaches "/user/clear"; a node event named "other" reaches "Emitter//other"
    g.send_event("clear");
    CHECK(seen_b.size() == 1 && seen_b[0].identifier == "user");
    Value other = Value::object();
    other["event"] = "other";
    g.set_properties("e", other);
    g.run();
    CHECK(seen_b.size() == 2 && seen_b[1].name == "other" && seen_a.size() == 1);

    // NEEDS_RECONNECT: on_connected of every node before the next run, listeners registered once
    Value rc = Value::object();
    rc["fire"] = false;
    rc["flags"] = int64_t(NEEDS_RECONNECT);
    g.set_properties("e", rc);
    trace.clear();
    g.run();
    g.run();
    CHECK((trace == std::vector<std::string>{"e:pre", "e:process", "la:process", "lb:process", "e:connect",
                                             "la:connect", "lb:connect", "e:pre", "e:process", "la:process",
                                             "lb:process"}));
    Value fire2 = Value::object();
    fire2["fire"] = true;
    fire2["event"] = "bounces_changed";
    g.set_properties("e", fire2);
    g.run();
    CHECK(seen_a.size() == 2); // one listener after the reconnect, not two

    // REMOVE_NODE: the node and its connections leave the graph
    Value rm = Value::object();
    rm["fire"] = false;
    rm["flags"] = int64_t(REMOVE_NODE);
    g.set_properties("lb", rm); // Listener ignores flags: remove the emitter instead
    g.set_properties("e", rm);
    g.run();
    CHECK((g.order() == std::vector<std::string>{"la", "lb"}));
    CHECK(!g.store().nodes.at("e").enabled);
    trace.clear();
    g.run();
    CHECK((trace == std::vector<std::string>{"la:connect", "lb:connect", "la:process", "lb:process"}));
    CHECK(g.iterations() == 8);

    if (failures) std::printf("%d failures\n", failures);
    return failures ? 1 : 0;
}
