// merian-amd-graph-run: the graph runner executable (all logic in mrt_graph_run_main).
#include "../../include/merian_amd.h"

extern "C" int32_t mrt_graph_run_main(int32_t argc, const char* const* argv);

int main(int argc, char** argv) { return mrt_graph_run_main(argc, argv); }
