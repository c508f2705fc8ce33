// bfs-directionopt: the lonestar direction-optimising BFS app entry point
// (lonestar/analytics/cpu/bfs/bfsDirectionOpt.cpp: flags :53-111, main
// :405-520) on the GPU direction-optimising BFS of the C ABI (gx_bfs_do:
// push while the frontier is small, pull over in-edges with a frontier
// bitmap otherwise, switching by -alpha and -beta as :259-300).  As in the
// reference, node data holds the BFS parent, not the level: the source keeps
// 0 and unreached nodes keep DIST_INFINITY (2147483646).  Parents are not
// unique; the levels they imply are.  -algo=Async runs the same kernel
// (push-only levels are the same), Auto picks by the degree test of
// Utils.h:37-57 as the reference does.
#include "lonestar.hpp"

using namespace lonestar;

int main(int argc, char** argv) {
  Cli cli;
  cli.value("startNode", "0", "Node to start search from (default value 0)");
  cli.value("reportNode", "1", "Node to report distance to (default value 1)");
  cli.value("numRuns", "1", "Number of runs (default value 1)");
  cli.value("alpha", "15", "alpha value to change direction in direction-optimization (default value 15)");
  cli.value("beta", "18", "beta value to change direction in direction-optimization (default value 18)");
  cli.value("preAlloc", "400", "Number of pages to preAlloc (default value 400)");
  cli.value("numPrint", "10",
            "Print parents for the numPrint number of nodes for verification if verification is on (default value "
            "10)");
  cli.value("exec", "PARALLEL", "Choose SERIAL or PARALLEL execution (default value PARALLEL):");
  cli.value("algo", "Auto", "Choose an algorithm (default value Auto): SyncDO, Async, Auto");
  cli.parse(argc, argv, "bfs-directionopt");
  Stats stats;
  start(cli, stats, "Breadth-first Search",
        "Computes the shortest path from a source node to all nodes in a directed graph using a modified "
        "Bellman-Ford algorithm");
  Timer total(stats, "TimerTotal");
  total.start();
  HostGraph hg;
  read_graph(cli.input(), hg);
  std::cout << "Read " << hg.n() << " nodes, " << hg.m() << " edges\n";
  const uint64_t src = cli.u64("startNode"), report = cli.u64("reportNode");
  if (static_cast<int64_t>(src) >= hg.n() || static_cast<int64_t>(report) >= hg.n()) {
    std::cerr << "failed to set report: " << report << " or failed to set source: " << src << "\n";
    std::abort();
  }
  std::string algo = cli.str("algo");
  if (algo != "SyncDO" && algo != "Async" && algo != "Auto") die("unknown -algo " + algo);
  const int alpha = static_cast<int>(cli.u64("alpha")), beta = static_cast<int>(cli.u64("beta"));
  if (alpha <= 0 || beta <= 0) die("-alpha and -beta must be positive");
  gx_graph* g = upload(hg, cli.on("symmetricGraph"), false);
  DeviceArray<uint32_t> level(hg.n()), parent(hg.n());
  std::cout << "Running " << algo << " algorithm with " << (cli.str("exec") == "SERIAL" ? "SERIAL" : "PARALLEL")
            << " execution\n";
  std::cout << "WARNING: This bfs version uses bi-directional CSR graph "
            << "and assigns parent instead of the shortest distance from source\n";
  std::cout << " Execution started\n";
  Timer exec(stats, "Timer_0");
  exec.start();
  if (algo == "Auto") {
    Timer autot(stats, "AutoAlgo_0");
    autot.start();
    algo = approximately_power_law(hg) ? "SyncDO" : "Async";
    autot.stop();
    autot.report();
    info("Choosing ", algo, " algorithm");
  }
  const uint64_t runs = std::max<uint64_t>(1, cli.u64("numRuns"));
  for (uint64_t run = 0; run < runs; ++run) {
    std::cout << "BFS::go run " << run << " called\n";
    Timer trun(stats, ("Timer_Run" + std::to_string(run)).c_str(), "BFS");
    trun.start();
    check(gx_bfs_do(g, static_cast<uint32_t>(src), static_cast<uint32_t>(alpha), static_cast<uint32_t>(beta),
                    level.p, parent.p, nullptr),
          "gx_bfs_do");
    hip_check(hipDeviceSynchronize(), "bfs");
    trun.stop();
    trun.report();
  }
  exec.stop();
  std::vector<uint32_t> par = parent.host();
  for (int64_t v = 0; v < hg.n(); ++v) {
    if (static_cast<uint64_t>(v) == src)
      par[v] = 0;  // the reference's source keeps the 0 it was initialised with
    else if (par[v] == UINT32_MAX)
      par[v] = GX_DIST_INFINITY;
  }
  std::cout << "Node " << report << " has parent " << par[report] << "\n";
  if (!cli.on("noverify")) {
    const uint64_t np = std::min<uint64_t>(cli.u64("numPrint"), static_cast<uint64_t>(hg.n()));
    for (uint64_t v = 0; v < np; ++v) std::cout << "parent[" << v << "] : " << par[v] << "\n";
  }
  write_node_values(cli.str("o"), par, [](FILE* f, uint32_t v) { return std::fprintf(f, "%u\n", v) > 0; });
  check(gx_graph_free(g), "gx_graph_free");
  total.stop();
  exec.report();
  total.report();
  stats.print(cli.str("statFile"));
  return 0;
}
