// Graph runner of the graph host: merian-graph-run's command line (src/merian-graph-run/main.cpp:
// runner flags :31-51,81-119, include resolution :185, cli overrides :198, load :223, run loop
// :243-247, --print-times :258-260) over this library's node host.
#include <cstdio>
#include <fstream>
#include <optional>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/merian_amd.h"
#include "graph.h"

using namespace mrt;
using namespace mrt::graph;

namespace {

void usage() {
    std::printf(
        "usage: merian-amd-graph-run [options] [graph.json [args...]]\n"
        "  --max-iterations=<N>          quit after N graph iterations\n"
        "  --print-times                 print per-node times averaged over the run\n"
        "  --merge <file.json>           deep-merge a JSON file into the config (repeatable)\n"
        "  --<name> <value>              set an override declared in the graph's \"cli\" block\n"
        "  args...                       positional args feed the cli \"--\" override\n"
        "  --validation=<..>, --loglevel=<..>, --plugin-path=<..>, --time-delta=<ms>: accepted, no effect\n"
        "  --help                        with a graph.json: also lists its cli overrides\n");
}

struct Options {
    std::optional<std::string> config;
    bool help = false, print_times = false;
    std::optional<uint64_t> max_iterations;
    std::vector<std::string> graph_args;
};

std::optional<Options> parse(const std::vector<std::string>& args) {
    Options o;
    bool found = false;
    for (const std::string& a : args) o.help = o.help || a == "--help" || a == "-h";
    for (size_t i = 0; i < args.size(); i++) {
        const std::string& a = args[i];
        auto starts = [&](const char* p) { return a.compare(0, std::char_traits<char>::length(p), p) == 0; };
        if (a == "--help" || a == "-h") continue;
        if (a == "--validation" || starts("--validation=") || starts("--loglevel=") || starts("--plugin-path=") ||
            starts("--time-delta=")) {
            continue;
        } else if (starts("--max-iterations=")) {
            o.max_iterations = std::stoull(a.substr(a.find('=') + 1));
        } else if (a == "--print-times") {
            o.print_times = true;
        } else if (found) {
            o.graph_args.push_back(a);
        } else if (starts("--")) {
            o.graph_args.push_back(a); // an override before the config keeps its value adjacent
            if (i + 1 < args.size()) o.graph_args.push_back(args[++i]);
        } else if (starts("-")) {
            std::fprintf(stderr, "unknown option '%s'\n", a.c_str());
            usage();
            return std::nullopt;
        } else {
            o.config = a;
            found = true;
        }
    }
    return o;
}

std::string dir_of(const std::string& p) {
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? "." : p.substr(0, k);
}

} // namespace

extern "C" MRT_API int32_t mrt_graph_run_main(int32_t argc, const char* const* argv) {
    std::vector<std::string> args;
    for (int32_t k = 1; k < argc; k++) args.emplace_back(argv[k]);
    const auto opt = parse(args);
    if (!opt) return 1;
    if (!opt->config) {
        if (opt->help) {
            usage();
            return 0;
        }
        std::fprintf(stderr, "no graph config given\n");
        return 1;
    }
    std::ifstream in(*opt->config);
    if (!in) {
        std::fprintf(stderr, "could not open config '%s'\n", opt->config->c_str());
        return 1;
    }
    std::stringstream ss;
    ss << in.rdbuf();
    json::Value config;
    try {
        config = json::parse(ss.str());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "could not parse config '%s': %s\n", opt->config->c_str(), e.what());
        return 1;
    }
    const std::vector<std::string> dirs{dir_of(*opt->config)};
    std::vector<std::string> errors;
    auto flush = [&] {
        for (const std::string& e : errors) std::fprintf(stderr, "error: %s\n", e.c_str());
        errors.clear();
    };
    if (!GraphDescription::resolve_includes(config, dirs, &errors)) {
        flush();
        return 1;
    }
    if (opt->help) {
        usage();
        std::printf("%s", GraphDescription::cli_help(config, dirs).c_str());
        return 0;
    }
    if (!GraphDescription::apply_cli(config, opt->graph_args, dirs, &errors)) {
        flush();
        std::printf("%s", GraphDescription::cli_help(config, dirs).c_str());
        return 1;
    }
    mrt_ctx* ctx = nullptr;
    if (mrt_create(0, 0, &ctx) != MRT_OK) {
        std::fprintf(stderr, "no usable HIP device\n");
        return 1;
    }
    int32_t rc = 0;
    try {
        const GraphDescription desc = GraphDescription::from_json(config);
        Graph g(ctx);
        g.set_timing(opt->print_times);
        g.load(desc);
        for (const std::string& w : g.warnings()) std::fprintf(stderr, "warning: %s\n", w.c_str());
        size_t shown = g.warnings().size();
        for (uint64_t it = 0; !g.stop_requested() && (!opt->max_iterations || it < *opt->max_iterations); it++) {
            g.run();
            for (; shown < g.warnings().size(); shown++) std::fprintf(stderr, "%s\n", g.warnings()[shown].c_str());
        }
        mrt_synchronize(ctx);
        if (opt->print_times) std::printf("%s", g.report().c_str());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        rc = 1;
    }
    mrt_destroy(ctx);
    return rc;
}
