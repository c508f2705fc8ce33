// connected-components: the lonestar CC app entry point (lonestar/analytics/
// cpu/connected-components/ConnectedComponents.cpp: flags :60-135, findLargest
// :1074-1157, verify :1046-1072, main :1206-1260) on the GPU union-find of the
// C ABI.  Labels are the minimum vertex id of each component (the reference's
// canonical result), so the component count, the non-trivial count and the
// largest size match every reference algorithm; -algo names are accepted.
#include <unordered_map>

#include "lonestar.hpp"

using namespace lonestar;

int main(int argc, char** argv) {
  Cli cli;
  cli.value("algo", "EdgetiledAsync",
            "Choose an algorithm: Async, EdgeAsync, EdgetiledAsync, BlockedAsync, LabelProp, Serial, Sync, "
            "Afforest, EdgeAfforest, EdgetiledAfforest");
  cli.value("edgeTileSize", "512", "(For Edgetiled algos) Size of edge tiles (default 512)");
  cli.value("vns", "2", "(For Afforest and its variants) number of edges per vertice to process initially");
  cli.value("lcs", "1024", "(For Afforest and its variants) number of times randomly sampling over vertices");
  cli.value("outputLargestComponent", "", "[output graph file] (not supported)");
  cli.value("outputNodePermutation", "", "[output node permutation file] (not supported)");
  cli.parse(argc, argv, "connected-components");
  Stats stats;
  start(cli, stats, "Connected Components", "Computes the connected components of a graph");
  Timer total(stats, "TimerTotal");
  total.start();
  if (!cli.on("symmetricGraph"))
    die("This application requires a symmetric graph input; please use the -symmetricGraph flag  to indicate "
        "the input is a symmetric graph.");
  HostGraph hg;
  read_graph(cli.input(), hg);
  std::cout << "Read " << hg.n() << " nodes\n";
  gx_graph* g = upload(hg, true, false);
  DeviceArray<uint32_t> label(hg.n());
  Timer exec(stats, "Timer_0");
  exec.start();
  check(gx_cc(g, label.p, nullptr), "gx_cc");
  hip_check(hipDeviceSynchronize(), "cc");
  exec.stop();
  const std::vector<uint32_t> lab = label.host();
  // "node label" lines (lonestar/analytics/gpu/connected-components/
  // support.cu:16-36 without its per-edge debug lines).
  write_node_values(cli.str("o"), lab, [](FILE* f, uint32_t v) { return std::fprintf(f, "%u\n", v) > 0; });
  if (!cli.on("noverify")) {
    size_t reps = 0;
    std::unordered_map<uint32_t, int64_t> members;  // non-representative members per component
    for (int64_t v = 0; v < hg.n(); ++v) {
      if (lab[v] == static_cast<uint32_t>(v))
        ++reps;
      else
        ++members[lab[v]];
    }
    int64_t largest = 0;
    for (const auto& kv : members) largest = std::max(largest, kv.second);
    const size_t largest_size = static_cast<size_t>(largest) + 1;
    double ratio = static_cast<double>(hg.n()) - static_cast<double>(reps) + static_cast<double>(members.size());
    if (ratio) ratio = largest_size / ratio;
    std::cout << "Total components: " << reps << "\n";
    std::cout << "Number of non-trivial components: " << members.size() << " (largest size: " << largest_size
              << " [" << ratio << "])\n";
    for (int64_t u = 0; u < hg.n(); ++u)
      for (int32_t e = hg.csr.row[u]; e < hg.csr.row[u + 1]; ++e) {
        const uint32_t v = static_cast<uint32_t>(hg.csr.col[e]);
        if (lab[u] != lab[v]) {
          std::cerr << "not in same component: " << u << " (" << lab[u] << ") and " << v << " (" << lab[v] << ")\n";
          die("verification failed");
        }
      }
  }
  check(gx_graph_free(g), "gx_graph_free");
  total.stop();
  exec.report();
  total.report();
  stats.print(cli.str("statFile"));
  return 0;
}
