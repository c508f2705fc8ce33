// sssp: the lonestar SSSP app entry point (lonestar/analytics/cpu/sssp/
// SSSP.cpp: flags :43-88, main :350-500) on the GPU delta-stepping kernel of
// the C ABI.  -delta is the bucket shift, as in the reference.  -algo names
// are accepted; AutoAlgo reports its choice with the reference's power-law
// test (Utils.h:37-57), and every choice runs the same kernel, whose
// distances are the unique shortest-path distances.
#include "lonestar.hpp"

using namespace lonestar;

namespace {

const char* kName = "Single Source Shortest Path";
const char* kDesc =
    "Computes the shortest path from a source node to all nodes in a directed "
    "graph using a modified chaotic iteration algorithm";

bool verify(const HostGraph& hg, const std::vector<uint32_t>& d, uint32_t src) {
  if (d[src] != 0) {
    std::cerr << "ERROR: source has non-zero dist value == " << d[src] << std::endl;
    return false;
  }
  size_t not_visited = 0;
  for (int64_t v = 0; v < hg.n(); ++v)
    if (d[v] >= GX_DIST_INFINITY) ++not_visited;
  if (not_visited)
    std::cerr << not_visited << " unvisited nodes; this is an error if the graph is strongly connected\n";
  const uint32_t* w = static_cast<const uint32_t*>(hg.csr.edge_data);
  bool bad = false;
  uint32_t maxd = 0;
  for (int64_t u = 0; u < hg.n(); ++u) {
    const uint32_t sd = d[u];
    if (sd == GX_DIST_INFINITY) continue;
    maxd = std::max(maxd, sd);
    for (int32_t e = hg.csr.row[u]; e < hg.csr.row[u + 1]; ++e) {
      const uint32_t v = static_cast<uint32_t>(hg.csr.col[e]);
      if (d[v] > sd + w[e]) {
        std::cout << "Wrong label: " << d[v] << ", on node: " << v << ", correct label from src node " << u
                  << " is " << sd + w[e] << "\n";
        bad = true;
      }
    }
  }
  if (bad) {
    std::cerr << "node found with incorrect distance\n";
    return false;
  }
  std::cout << "max dist: " << maxd << "\n";
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  Cli cli;
  cli.value("startNode", "0", "Node to start search from (default value 0)");
  cli.value("reportNode", "1", "Node to report distance to(default value 1)");
  cli.value("delta", "13", "Shift value for the deltastep (default value 13)");
  cli.value("algo", "Auto",
            "Choose an algorithm (default value auto): deltaTile, deltaStep, deltaStepBarrier, serDeltaTile, "
            "serDelta, dijkstraTile, dijkstra, topo, topoTile, Auto");
  cli.parse(argc, argv, "sssp");
  Stats stats;
  start(cli, stats, kName, kDesc);
  Timer total(stats, "TimerTotal");
  total.start();

  HostGraph hg;
  std::cout << "Reading from file: " << cli.input() << "\n";
  read_graph(cli.input(), hg);
  std::cout << "Read " << hg.n() << " nodes, " << hg.m() << " edges\n";
  const uint64_t src = cli.u64("startNode"), report = cli.u64("reportNode");
  if (static_cast<int64_t>(src) >= hg.n() || static_cast<int64_t>(report) >= hg.n()) {
    std::cerr << "failed to set report: " << report << " or failed to set source: " << src << "\n";
    std::abort();
  }
  const uint32_t shift = static_cast<uint32_t>(cli.u64("delta"));
  std::string algo = cli.str("algo");
  if (algo == "deltaStep" || algo == "deltaTile" || algo == "serDelta" || algo == "serDeltaTile") {
    std::cout << "INFO: Using delta-step of " << (1u << shift) << "\n";
    std::cout << "WARNING: Performance varies considerably due to delta parameter.\n";
    std::cout << "WARNING: Do not expect the default to be good for your graph.\n";
  }
  gx_graph* g = upload(hg, cli.on("symmetricGraph"), true);
  DeviceArray<uint32_t> dist(hg.n());
  std::cout << "Running " << algo << " algorithm\n";
  Timer exec(stats, "Timer_0");
  exec.start();
  if (algo == "Auto") {
    Timer autot(stats, "AutoAlgo_0");
    autot.start();
    algo = approximately_power_law(hg) ? "deltaStep" : "deltaStepBarrier";
    autot.stop();
    autot.report();
    info("Choosing ", algo, " algorithm");
  }
  check(gx_sssp_delta(g, static_cast<uint32_t>(src), shift, dist.p, nullptr), "gx_sssp_delta");
  hip_check(hipDeviceSynchronize(), "sssp");
  exec.stop();
  const std::vector<uint32_t> d = dist.host();
  std::cout << "Node " << report << " has distance " << d[report] << "\n";
  write_distances(cli.str("o"), d);
  uint64_t maxd = 0, sum = 0, visited = 0;
  for (uint32_t x : d)
    if (x != GX_DIST_INFINITY) {
      maxd = std::max<uint64_t>(maxd, x);
      sum += x;
      ++visited;
    }
  info("# visited nodes is ", visited);
  info("Max distance is ", maxd);
  info("Sum of visited distances is ", sum);
  if (!cli.on("noverify")) {
    if (verify(hg, d, static_cast<uint32_t>(src)))
      std::cout << "Verification successful.\n";
    else
      die("verification failed");
  }
  check(gx_graph_free(g), "gx_graph_free");
  total.stop();
  exec.report();
  total.report();
  stats.print(cli.str("statFile"));
  return 0;
}
