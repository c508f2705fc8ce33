// bfs: the lonestar BFS app entry point (lonestar/analytics/cpu/bfs/bfs.cpp:
// flags :45-76, main :318-415) on the GPU top-down BFS of the C ABI.
// -algo and -exec are accepted for script compatibility; every choice runs
// the same level-synchronous GPU kernel, whose levels are unique.
#include "lonestar.hpp"

using namespace lonestar;

namespace {

const char* kName = "Breadth-first Search";
const char* kDesc =
    "Computes the shortest path from a source node to all nodes in a directed "
    "graph using a modified Bellman-Ford algorithm";

// BFS_SSSP::verify (lonestar/liblonestar/include/Lonestar/BFS_SSSP.h:180-263):
// source at 0 and no edge can shorten a reached vertex.
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
  bool bad = false;
  uint32_t maxd = 0;
  for (int64_t u = 0; u < hg.n(); ++u) {
    const uint32_t sd = d[u];
    if (sd == GX_DIST_INFINITY) continue;
    maxd = std::max(maxd, sd);
    for (int32_t e = hg.csr.row[u]; e < hg.csr.row[u + 1]; ++e) {
      const uint32_t v = static_cast<uint32_t>(hg.csr.col[e]);
      if (d[v] > sd + 1) {
        std::cout << "Wrong label: " << d[v] << ", on node: " << v << ", correct label from src node " << u
                  << " is " << sd + 1 << "\n";
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
  cli.value("reportNode", "1", "Node to report distance to (default value 1)");
  cli.value("exec", "PARALLEL", "Choose SERIAL or PARALLEL execution (default value PARALLEL)");
  cli.value("algo", "SyncTile", "Choose an algorithm (default value SyncTile): AsyncTile, Async, SyncTile, Sync");
  cli.parse(argc, argv, "bfs");
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
  gx_graph* g = upload(hg, cli.on("symmetricGraph"), false);
  DeviceArray<uint32_t> level(hg.n());
  std::cout << "Running " << cli.str("algo") << " algorithm with "
            << (cli.str("exec") == "SERIAL" ? "SERIAL" : "PARALLEL") << " execution\n";
  Timer exec(stats, "Timer_0");
  exec.start();
  check(gx_bfs(g, static_cast<uint32_t>(src), level.p, nullptr), "gx_bfs");
  hip_check(hipDeviceSynchronize(), "bfs");
  exec.stop();
  const std::vector<uint32_t> d = level.host();
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
